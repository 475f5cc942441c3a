// generate.cc — deterministic synthetic lattices, pair queries and multi-pin
// nets.  The stream is xoshiro256** seeded through splitmix64, so every
// workload is reproducible from its seed on any host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "apgx.h"
#include "common.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (uint64_t& word : s) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      word = x ^ (x >> 31);
    }
  }
  static uint64_t Rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t Next() {
    const uint64_t result = Rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = Rotl(s[3], 45);
    return result;
  }
  // Uniform integer in [lo, hi] by 32x32->64 multiply-shift of the top word.
  uint32_t Range(uint32_t lo, uint32_t hi) {
    const uint64_t span = static_cast<uint64_t>(hi) - lo + 1;
    return lo + static_cast<uint32_t>(((Next() >> 32) * span) >> 32);
  }
  // Uniform index in [0, n) for n up to 2^64 (64x64 -> high 64).
  uint64_t Index(uint64_t n) {
    return static_cast<uint64_t>((static_cast<unsigned __int128>(Next()) * n) >> 64);
  }
  bool Ppm(uint32_t ppm) { return Range(0, 999999) < ppm; }
  double Unit() { return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" int apgx_generate_lattice(const apgx_lattice_gen_params* p, uint32_t* cost) {
  if (p == nullptr || cost == nullptr || p->width == 0 || p->height == 0 || p->layers == 0 ||
      p->cost_min == 0 || p->cost_min > p->cost_max || p->cost_max > APGX_MAX_CELL_COST ||
      p->blocked_ppm > 1000000 || p->keepout_ppm > 900000 ||
      (p->keepout_ppm != 0 &&
       (p->keepout_min_side == 0 || p->keepout_min_side > p->keepout_max_side ||
        p->keepout_max_side > p->width || p->keepout_max_side > p->height))) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_lattice: invalid parameters");
  }
  Rng rng(p->seed);
  const uint64_t plane = static_cast<uint64_t>(p->width) * p->height;
  const uint64_t total = plane * p->layers;
  // Cell costs and i.i.d. blocking, one cost draw and one blocking draw per cell
  // in cell-index order.
  for (uint64_t c = 0; c < total; ++c) {
    const uint32_t value = rng.Range(p->cost_min, p->cost_max);
    const bool blocked = p->blocked_ppm != 0 && rng.Ppm(p->blocked_ppm);
    cost[c] = blocked ? 0 : value;
  }
  // Rectangular keepouts per layer: sides uniform in [min, max], positions
  // uniform with the rectangle inside the lattice, added until the layer's
  // blocked fraction reaches the target.
  if (p->keepout_ppm != 0) {
    const uint64_t target = (plane * p->keepout_ppm + 999999) / 1000000;
    for (uint32_t l = 0; l < p->layers; ++l) {
      uint32_t* layer = cost + l * plane;
      uint64_t blocked = 0;
      for (uint64_t i = 0; i < plane; ++i) blocked += layer[i] == 0;
      while (blocked < target) {
        const uint32_t w = rng.Range(p->keepout_min_side, p->keepout_max_side);
        const uint32_t h = rng.Range(p->keepout_min_side, p->keepout_max_side);
        const uint32_t x0 = rng.Range(0, p->width - w);
        const uint32_t y0 = rng.Range(0, p->height - h);
        for (uint32_t y = y0; y < y0 + h; ++y) {
          for (uint32_t x = x0; x < x0 + w; ++x) {
            uint32_t& cell = layer[static_cast<uint64_t>(y) * p->width + x];
            if (cell != 0) {
              cell = 0;
              ++blocked;
            }
          }
        }
      }
    }
  }
  apgx::ClearError();
  return APGX_OK;
}

extern "C" int apgx_generate_pairs(const uint32_t* cost, uint32_t width, uint32_t height,
                                   uint32_t layers, uint32_t count, uint64_t seed,
                                   uint32_t* pins_out) {
  if (cost == nullptr || pins_out == nullptr || width == 0 || height == 0 || layers == 0) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_pairs: invalid parameters");
  }
  const uint64_t total = static_cast<uint64_t>(width) * height * layers;
  uint64_t free_cells = 0;
  for (uint64_t c = 0; c < total; ++c) free_cells += cost[c] != 0;
  if (free_cells < 2) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_pairs: fewer than two free cells");
  }
  Rng rng(seed);
  for (uint32_t q = 0; q < count; ++q) {
    uint64_t s, t;
    do {
      s = rng.Index(total);
    } while (cost[s] == 0);
    do {
      t = rng.Index(total);
    } while (cost[t] == 0 || t == s);
    pins_out[2 * q] = static_cast<uint32_t>(s);
    pins_out[2 * q + 1] = static_cast<uint32_t>(t);
  }
  apgx::ClearError();
  return APGX_OK;
}

extern "C" int apgx_generate_nets(const uint32_t* cost, uint32_t width, uint32_t height,
                                  uint32_t layers, uint32_t count, uint32_t min_pins,
                                  uint32_t max_pins, uint32_t min_side, uint32_t max_side,
                                  uint64_t seed, apgx_net_query* nets_out, uint32_t* pins_out,
                                  uint64_t* pin_total) {
  if (cost == nullptr || nets_out == nullptr || pins_out == nullptr || pin_total == nullptr ||
      min_pins < 2 || min_pins > max_pins || max_pins > APGX_MAX_NET_PINS || min_side == 0 ||
      min_side > max_side || max_side > width || max_side > height) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_nets: invalid parameters");
  }
  Rng rng(seed);
  const uint64_t plane = static_cast<uint64_t>(width) * height;
  // Summed-area table of free cells over all layers: window free counts in O(1).
  const uint64_t sw = static_cast<uint64_t>(width) + 1;
  std::vector<uint64_t> sat(sw * (static_cast<uint64_t>(height) + 1), 0);
  for (uint32_t y = 0; y < height; ++y) {
    uint64_t row = 0;
    for (uint32_t x = 0; x < width; ++x) {
      for (uint32_t l = 0; l < layers; ++l) row += cost[l * plane + static_cast<uint64_t>(y) * width + x] != 0;
      sat[(y + 1) * sw + x + 1] = sat[y * sw + x + 1] + row;
    }
  }
  uint64_t used = 0;
  const double ratio = std::log2(static_cast<double>(max_side) / min_side);
  for (uint32_t n = 0; n < count; ++n) {
    // Square window, side log-uniform in [min_side, max_side].
    uint32_t side = static_cast<uint32_t>(std::floor(min_side * std::exp2(ratio * rng.Unit())));
    side = std::clamp(side, min_side, max_side);
    const uint32_t x0 = rng.Range(0, width - side);
    const uint32_t y0 = rng.Range(0, height - side);
    const uint32_t want = rng.Range(min_pins, max_pins);
    const uint64_t x1 = x0 + side, y1 = y0 + side;
    const uint64_t window_free = sat[y1 * sw + x1] - sat[y0 * sw + x1] - sat[y1 * sw + x0] + sat[y0 * sw + x0];
    const uint32_t pins = static_cast<uint32_t>(std::min<uint64_t>(want, window_free));
    nets_out[n].net_id = n + 1;
    nets_out[n].pin_offset = static_cast<uint32_t>(used);
    nets_out[n].pin_count = pins;
    nets_out[n].window = apgx_window{x0, y0, side, side};
    uint32_t placed = 0;
    while (placed < pins) {
      const uint32_t l = rng.Range(0, layers - 1);
      const uint32_t x = rng.Range(x0, x0 + side - 1);
      const uint32_t y = rng.Range(y0, y0 + side - 1);
      const uint64_t c = l * plane + static_cast<uint64_t>(y) * width + x;
      if (cost[c] == 0) continue;
      bool duplicate = false;
      for (uint32_t i = 0; i < placed; ++i) duplicate |= pins_out[used + i] == c;
      if (duplicate) continue;
      pins_out[used + placed++] = static_cast<uint32_t>(c);
    }
    used += pins;
  }
  *pin_total = used;
  apgx::ClearError();
  return APGX_OK;
}

extern "C" int apgx_generate_board(const apgx_board_gen_params* p, uint32_t* cost,
                                   apgx_net_query* nets_out, uint32_t* pins_out, uint64_t* pin_total) {
  if (p == nullptr || cost == nullptr || nets_out == nullptr || pins_out == nullptr ||
      pin_total == nullptr || p->bga_fields == 0 || p->bga_pads == 0 || p->bga_pitch == 0 ||
      p->min_pins < 2 || p->min_pins > p->max_pins || p->max_pins > APGX_MAX_NET_PINS ||
      static_cast<uint64_t>(p->bga_pads) * p->bga_pitch > std::min(p->width, p->height) ||
      p->pad_pin_ppm > 1000000) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_board: invalid parameters");
  }
  apgx_lattice_gen_params lp{p->width, p->height, p->layers, p->cost_min, p->cost_max, 0,
                             p->keepout_ppm, p->keepout_min_side, p->keepout_max_side, 0, p->seed};
  if (int rc = apgx_generate_lattice(&lp, cost); rc != APGX_OK) return rc;
  Rng rng(p->seed ^ 0xB6A0B6A0B6A0B6A0ull);
  const uint64_t plane = static_cast<uint64_t>(p->width) * p->height;
  const uint32_t side = p->bga_pads * p->bga_pitch;
  // Field placement: uniform positions, rejecting overlaps (bounded attempts).
  std::vector<uint32_t> fx, fy;
  for (uint32_t f = 0, tries = 0; f < p->bga_fields; ++tries) {
    if (tries > 100000) return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_board: fields do not fit");
    const uint32_t x = rng.Range(0, p->width - side), y = rng.Range(0, p->height - side);
    bool overlap = false;
    for (uint32_t g = 0; g < f; ++g)
      overlap |= x < fx[g] + side && fx[g] < x + side && y < fy[g] + side && fy[g] < y + side;
    if (overlap) continue;
    fx.push_back(x);
    fy.push_back(y);
    ++f;
  }
  const uint32_t pads = p->bga_pads * p->bga_pads;
  auto pad_cell = [&](uint32_t f, uint32_t i) {
    const uint32_t x = fx[f] + (i % p->bga_pads) * p->bga_pitch + p->bga_pitch / 2;
    const uint32_t y = fy[f] + (i / p->bga_pads) * p->bga_pitch + p->bga_pitch / 2;
    return static_cast<uint64_t>(y) * p->width + x;  // layer 0
  };
  for (uint32_t f = 0; f < p->bga_fields; ++f)
    for (uint32_t i = 0; i < pads; ++i) {
      uint32_t& c = cost[pad_cell(f, i)];
      if (c == 0) c = rng.Range(p->cost_min, p->cost_max);
    }
  // Pad cells are reserved for pad pins: scattered pins never land on one.
  std::vector<uint8_t> used(plane * p->layers, 0);
  for (uint32_t f = 0; f < p->bga_fields; ++f)
    for (uint32_t i = 0; i < pads; ++i) used[pad_cell(f, i)] = 1;
  // Pin access (pin_clearance r > 0): cells within Chebyshev distance r of a
  // pin on its layer, and the cell above each pad, are kept out of reach of
  // scattered pins and forced free at the end, so no pin is walled in by
  // keepouts or by another net's pin.
  const uint32_t clr = p->pin_clearance;
  std::vector<uint8_t> near(clr ? plane * p->layers : 0, 0);
  auto mark_near = [&](uint64_t c) {
    const uint64_t l = c / plane, y = (c % plane) / p->width, x = c % p->width;
    for (int64_t dy = -static_cast<int64_t>(clr); dy <= static_cast<int64_t>(clr); ++dy)
      for (int64_t dx = -static_cast<int64_t>(clr); dx <= static_cast<int64_t>(clr); ++dx) {
        const int64_t yy = static_cast<int64_t>(y) + dy, xx = static_cast<int64_t>(x) + dx;
        if (yy >= 0 && yy < p->height && xx >= 0 && xx < p->width)
          near[l * plane + static_cast<uint64_t>(yy) * p->width + static_cast<uint64_t>(xx)] = 1;
      }
  };
  if (clr)
    for (uint32_t f = 0; f < p->bga_fields; ++f)
      for (uint32_t i = 0; i < pads; ++i) {
        mark_near(pad_cell(f, i));
        if (p->layers > 1) near[plane + pad_cell(f, i)] = 1;
      }
  std::vector<uint32_t> free_pads(static_cast<size_t>(p->bga_fields) * pads);
  std::vector<uint32_t> free_count(p->bga_fields, pads);
  for (uint32_t f = 0; f < p->bga_fields; ++f)
    for (uint32_t i = 0; i < pads; ++i) free_pads[static_cast<size_t>(f) * pads + i] = i;
  // Takes a random unused pad of field f (or of the next field with one left).
  auto take_pad = [&](uint32_t f, uint64_t* out) {
    for (uint32_t k = 0; k < p->bga_fields; ++k) {
      const uint32_t g = (f + k) % p->bga_fields;
      if (free_count[g] == 0) continue;
      uint32_t* list = &free_pads[static_cast<size_t>(g) * pads];
      const uint32_t j = rng.Range(0, free_count[g] - 1);
      const uint32_t pad = list[j];
      list[j] = list[--free_count[g]];
      *out = pad_cell(g, pad);
      return true;
    }
    return false;
  };
  uint64_t w = 0;
  for (uint32_t n = 0; n < p->nets; ++n) {
    const uint32_t k = rng.Range(p->min_pins, p->max_pins);
    const uint32_t f = rng.Range(0, p->bga_fields - 1);
    const uint64_t base = w;
    uint64_t c = 0;
    if (!take_pad(f, &c)) return apgx::Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_generate_board: pads exhausted");
    used[c] = 1;
    pins_out[w++] = static_cast<uint32_t>(c);
    const uint32_t cx = fx[f] + side / 2, cy = fy[f] + side / 2;
    const uint32_t r = p->scatter_radius ? p->scatter_radius : std::max(p->width, p->height);
    const uint32_t x0 = cx > r ? cx - r : 0, x1 = std::min(p->width - 1, cx + r);
    const uint32_t y0 = cy > r ? cy - r : 0, y1 = std::min(p->height - 1, cy + r);
    for (uint32_t i = 1; i < k; ++i) {
      if (rng.Ppm(p->pad_pin_ppm)) {
        const uint32_t g = rng.Range(0, 1) ? f : rng.Range(0, p->bga_fields - 1);
        if (!take_pad(g, &c)) return apgx::Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_generate_board: pads exhausted");
      } else {
        uint32_t tries = 0;
        do {
          if (++tries > 1000000) return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_generate_board: no free cell");
          const uint32_t l = rng.Range(0, p->layers - 1);
          c = l * plane + static_cast<uint64_t>(rng.Range(y0, y1)) * p->width + rng.Range(x0, x1);
        } while (cost[c] == 0 || used[c] || (clr && near[c]));
        if (clr) mark_near(c);
      }
      used[c] = 1;
      pins_out[w++] = static_cast<uint32_t>(c);
    }
    uint32_t bx0 = p->width, by0 = p->height, bx1 = 0, by1 = 0;
    for (uint64_t j = base; j < w; ++j) {
      const uint32_t x = static_cast<uint32_t>(pins_out[j] % p->width);
      const uint32_t y = static_cast<uint32_t>((pins_out[j] / p->width) % p->height);
      bx0 = std::min(bx0, x); bx1 = std::max(bx1, x);
      by0 = std::min(by0, y); by1 = std::max(by1, y);
    }
    const uint32_t m = p->window_margin;
    const uint32_t wx0 = bx0 > m ? bx0 - m : 0, wy0 = by0 > m ? by0 - m : 0;
    const uint32_t wx1 = std::min(p->width - 1, bx1 + m), wy1 = std::min(p->height - 1, by1 + m);
    nets_out[n].net_id = n + 1;
    nets_out[n].pin_offset = static_cast<uint32_t>(base);
    nets_out[n].pin_count = k;
    nets_out[n].window = apgx_window{wx0, wy0, wx1 - wx0 + 1, wy1 - wy0 + 1};
  }
  if (clr) {
    for (uint64_t c = 0; c < plane * p->layers; ++c)
      if (near[c] && cost[c] == 0) cost[c] = rng.Range(p->cost_min, p->cost_max);
  }
  *pin_total = w;
  apgx::ClearError();
  return APGX_OK;
}
