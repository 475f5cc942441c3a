// prices.hip — candidate-allocator price epoch on cell resources.
//
// Restates the price update of PlanNegotiatedRegeneration
// (src/allocator/negotiated_regeneration_plan.cc:697-790) with
// NextPresentFactor (:430-442) for the dense lattice, one resource per cell in
// cell order (the reference's canonical resource order becomes the cell
// index):
//   factor   = initial + epoch * increment                      (checked)
//   entries  = cells with prior history != 0, plus overused cells
//   overused cell: present = factor * overuse, history += inc * overuse,
//                  total = present + history                    (checked)
//   other entries: present = 0, total = history
// followed by the reference's bounds in its order: hot-resource count, then
// per overused cell (cell order) the entry bound and the three overflow
// checks, then per entry (cell order) the value bound and the running
// aggregate (overflow, then bound).  The first failing check is the outcome
// and leaves the planes untouched, as the reference returns an error instead
// of a snapshot.
//
// Kernels: one HBM pass computes per-cell values and per-block partials
// (counts, clamped sums, first failing cells); one workgroup then walks the
// block partials in the reference's check order (totals, block-wide scans to
// find where a running count or sum first crosses its bound, the exact cell
// inside that block) and records the outcome; a second pass writes the planes
// when every check passed.  The three launches run back to back and the host
// reads one 64-byte outcome record at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"

namespace apgx {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kBlockCells = kThreads * kPerThread;
constexpr uint64_t kNone = ~0ull;
constexpr uint64_t kSumClamp = 1ull << 33;  // per-entry clamp of block sums (values above fail the value bound)

struct Params {
  uint64_t factor, inc, max_value;
  uint64_t cells;
};

struct CellEval {
  bool over_cell;     // usage > capacity
  bool entry;         // price entry after the update
  bool new_entry;     // overused, no prior history
  uint8_t arith;      // 0 ok, else APGX_PRICE_* overflow outcome of this overused cell
  bool value_fail;
  uint64_t present, history, total;
};

__host__ __device__ inline CellEval Evaluate(const Params& p, uint32_t usage, uint32_t cap, uint32_t hist) {
  CellEval e{};
  const uint64_t over = usage > cap ? usage - cap : 0u;
  e.over_cell = over > 0;
  e.history = hist;
  e.entry = hist != 0 || e.over_cell;
  e.new_entry = e.over_cell && hist == 0;
  if (e.over_cell) {
    uint64_t present = 0, add = 0, h = 0, t = 0;
    if (__builtin_mul_overflow(p.factor, over, &present) || __builtin_mul_overflow(p.inc, over, &add)) {
      e.arith = APGX_PRICE_MULTIPLY_OVERFLOW;
    } else if (__builtin_add_overflow(static_cast<uint64_t>(hist), add, &h)) {
      e.arith = APGX_PRICE_HISTORY_OVERFLOW;
    } else if (__builtin_add_overflow(present, h, &t)) {
      e.arith = APGX_PRICE_TOTAL_OVERFLOW;
    }
    e.present = present;
    e.history = h;
    e.total = t;
  } else {
    e.total = hist;
  }
  e.value_fail = e.entry && e.arith == 0 &&
                 (e.present > p.max_value || e.history > p.max_value || e.total > p.max_value);
  return e;
}

struct BlockPartial {
  unsigned long long over_count, new_count, entry_count, sum_total;
  unsigned long long first_arith, first_value;  // cell index or kNone
  unsigned long long arith_kind, pad;
};

__device__ inline uint32_t CapAt(const uint8_t* cap, uint64_t c) { return cap ? cap[c] : 1u; }

__global__ __launch_bounds__(kThreads) void PriceScan(Params p, const uint32_t* __restrict__ usage,
                                                      const uint8_t* __restrict__ cap,
                                                      const uint32_t* __restrict__ hist,
                                                      BlockPartial* __restrict__ out) {
  __shared__ unsigned long long s_over, s_new, s_entry, s_sum, s_arith, s_value;
  if (threadIdx.x == 0) {
    s_over = s_new = s_entry = s_sum = 0;
    s_arith = s_value = kNone;
  }
  __syncthreads();
  unsigned long long over = 0, nw = 0, ent = 0, sum = 0, fa = kNone, fv = kNone;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlockCells;
  for (int k = 0; k < kPerThread; ++k) {
    const uint64_t c = base + static_cast<uint64_t>(k) * kThreads + threadIdx.x;  // coalesced
    if (c >= p.cells) break;
    const CellEval e = Evaluate(p, usage[c], CapAt(cap, c), hist[c]);
    over += e.over_cell;
    nw += e.new_entry;
    ent += e.entry;
    if (e.entry) sum += min(e.total, kSumClamp);
    if (e.arith && c < fa) fa = c;
    if (e.value_fail && c < fv) fv = c;
  }
  atomicAdd(&s_over, over);
  atomicAdd(&s_new, nw);
  atomicAdd(&s_entry, ent);
  atomicAdd(&s_sum, sum);
  atomicMin(&s_arith, fa);
  atomicMin(&s_value, fv);
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockPartial b{};
    b.over_count = s_over;
    b.new_count = s_new;
    b.entry_count = s_entry;
    b.sum_total = s_sum;
    b.first_arith = s_arith;
    b.first_value = s_value;
    b.arith_kind = 0;
    out[blockIdx.x] = b;
  }
}

struct Limits {
  uint64_t max_entries, max_hot, max_agg;
};

// Outcome of an epoch, written by PriceDecide (64 bytes).
struct Decision {
  unsigned long long outcome;    // APGX_PRICE_*
  unsigned long long cell;       // failing cell or kNone
  unsigned long long hot;        // overused cells
  unsigned long long entries;    // price entries after the update
  unsigned long long aggregate;  // sum of totals (passing epochs)
  unsigned long long pass;       // 1: every check passed and PriceWrite writes the planes
  unsigned long long pad[2];
};
static_assert(sizeof(Decision) == 64, "outcome record");

// The cell of the `target`-th (0-based) new entry of one block, in cell order.
__device__ uint64_t LocateNewEntry(const Params& p, const uint32_t* usage, const uint8_t* cap, const uint32_t* hist,
                                   uint64_t block, uint64_t target) {
  uint64_t seen = 0;
  const uint64_t begin = block * kBlockCells, end = min(p.cells, begin + kBlockCells);
  for (uint64_t c = begin; c < end; ++c) {
    if (!Evaluate(p, usage[c], CapAt(cap, c), hist[c]).new_entry) continue;
    if (seen == target) return c;
    ++seen;
  }
  return kNone;
}

// The entry of one block, in cell order, at which `agg` plus the running
// total first overflows (*overflow) or exceeds `bound`.
__device__ uint64_t LocateAggregate(const Params& p, const uint32_t* usage, const uint8_t* cap, const uint32_t* hist,
                                    uint64_t block, uint64_t agg, uint64_t bound, bool* overflow) {
  const uint64_t begin = block * kBlockCells, end = min(p.cells, begin + kBlockCells);
  for (uint64_t c = begin; c < end; ++c) {
    const CellEval e = Evaluate(p, usage[c], CapAt(cap, c), hist[c]);
    if (!e.entry) continue;
    uint64_t next = 0;
    if (__builtin_add_overflow(agg, e.total, &next)) {
      *overflow = true;
      return c;
    }
    agg = next;
    if (agg > bound) return c;
  }
  return kNone;
}

constexpr int kDecideThreads = 1024;

// Exclusive prefix sums of s[0 .. kDecideThreads) in place; returns the
// total (thread 0; the callers synchronise around it).
__device__ unsigned __int128 ExclusiveScan(unsigned __int128* s) {
  unsigned __int128 run = 0;
  for (int i = 0; i < kDecideThreads; ++i) {
    const unsigned __int128 v = s[i];
    s[i] = run;
    run += v;
  }
  return run;
}

// The epoch's checks in the reference's order (negotiated_regeneration_plan.cc
// :690-790), over the block partials of PriceScan: the hot-resource bound;
// then, over the overused cells in cell order, the entry bound (the cell where
// prior entries + new entries so far reach the bound) against the first
// arithmetic overflow; then, over the entries in cell order, the first value
// bound against the first crossing of the aggregate bound.  Each thread owns
// a contiguous run of blocks; a running count or sum crosses its bound in
// exactly one run, found from a scan of the per-thread sums, whose owner finds
// the block, and thread 0 walks that block for the cell.
__global__ __launch_bounds__(kDecideThreads) void PriceDecide(Params p, Limits lim, const uint32_t* __restrict__ usage,
                                                              const uint8_t* __restrict__ cap,
                                                              const uint32_t* __restrict__ hist,
                                                              const BlockPartial* __restrict__ part, uint64_t blocks,
                                                              Decision* __restrict__ out) {
  __shared__ unsigned long long s_over, s_entries, s_prior, s_fa, s_fv;
  __shared__ unsigned __int128 s_scan[kDecideThreads];
  __shared__ unsigned long long s_hit, s_hit_before, s_entry_cell, s_done;
  const int t = threadIdx.x;
  const uint64_t chunk = (blocks + kDecideThreads - 1) / kDecideThreads;
  const uint64_t b0 = min(blocks, static_cast<uint64_t>(t) * chunk), b1 = min(blocks, b0 + chunk);
  if (t == 0) {
    s_over = s_entries = s_prior = 0;
    s_fa = s_fv = kNone;
    s_entry_cell = kNone;
    s_done = 0;
  }
  __syncthreads();
  unsigned long long over = 0, ent = 0, prior = 0, fa = kNone, fv = kNone;
  for (uint64_t i = b0; i < b1; ++i) {
    const BlockPartial q = part[i];
    over += q.over_count;
    ent += q.entry_count;
    prior += q.entry_count - q.new_count;
    fa = min(fa, q.first_arith);
    fv = min(fv, q.first_value);
  }
  atomicAdd(&s_over, over);
  atomicAdd(&s_entries, ent);
  atomicAdd(&s_prior, prior);
  atomicMin(&s_fa, fa);
  atomicMin(&s_fv, fv);
  __syncthreads();
  const uint64_t total_over = s_over, entries = s_entries, prior_entries = s_prior;
  const uint64_t first_arith = s_fa, first_value = s_fv;
  Decision d{};
  d.cell = kNone;
  d.hot = total_over;
  d.entries = entries;
  // 1. Hot-resource bound, before any price.
  if (total_over > lim.max_hot) {
    if (t == 0) {
      d.outcome = APGX_PRICE_HOT_RESOURCE_BOUND;
      *out = d;
    }
    return;
  }
  // 2. Entry bound: the k-th new entry with prior + k reaching the bound.
  const uint64_t new_entries = entries - prior_entries;
  if (new_entries > 0 && prior_entries + new_entries > lim.max_entries) {
    const uint64_t k = lim.max_entries > prior_entries ? lim.max_entries - prior_entries : 0;
    unsigned long long mine = 0;
    for (uint64_t i = b0; i < b1; ++i) mine += part[i].new_count;
    s_scan[t] = mine;
    if (t == 0) s_hit = kNone;
    __syncthreads();
    if (t == 0) ExclusiveScan(s_scan);
    __syncthreads();
    const unsigned __int128 before = s_scan[t];
    if (before <= k && k < before + mine) {
      uint64_t seen = static_cast<uint64_t>(before);
      for (uint64_t i = b0; i < b1; ++i) {
        if (seen + part[i].new_count > k) {
          s_hit = i;
          s_hit_before = seen;
          break;
        }
        seen += part[i].new_count;
      }
    }
    __syncthreads();
    if (t == 0 && s_hit != kNone) s_entry_cell = LocateNewEntry(p, usage, cap, hist, s_hit, k - s_hit_before);
    __syncthreads();
  }
  // 3. The entry bound or the first overflowing overused cell, whichever comes first.
  if (t == 0) {
    const uint64_t entry_cell = s_entry_cell;
    if (entry_cell != kNone && entry_cell <= first_arith) {
      d.outcome = APGX_PRICE_ENTRY_BOUND;
      d.cell = entry_cell;
      s_done = 1;
    } else if (first_arith != kNone) {
      d.outcome = Evaluate(p, usage[first_arith], CapAt(cap, first_arith), hist[first_arith]).arith;
      d.cell = first_arith;
      s_done = 1;
    }
  }
  __syncthreads();
  if (s_done) {
    if (t == 0) *out = d;
    return;
  }
  // 4. Entries in cell order: the running aggregate's first crossing.
  unsigned __int128 mine = 0;
  for (uint64_t i = b0; i < b1; ++i) mine += part[i].sum_total;
  s_scan[t] = mine;
  if (t == 0) s_hit = kNone;
  __syncthreads();
  unsigned __int128 sum = 0;  // the sum of every entry's total (thread 0)
  if (t == 0) sum = ExclusiveScan(s_scan);
  __syncthreads();
  const unsigned __int128 before = s_scan[t];
  if (before <= lim.max_agg && before + mine > lim.max_agg) {
    unsigned __int128 seen = before;
    for (uint64_t i = b0; i < b1; ++i) {
      if (seen + part[i].sum_total > lim.max_agg) {
        s_hit = i;
        s_hit_before = static_cast<uint64_t>(seen);
        break;
      }
      seen += part[i].sum_total;
    }
  }
  __syncthreads();
  if (t != 0) return;
  uint64_t agg_cell = kNone;
  bool overflow = false;
  if (s_hit != kNone) agg_cell = LocateAggregate(p, usage, cap, hist, s_hit, s_hit_before, lim.max_agg, &overflow);
  if (first_value != kNone && first_value <= agg_cell) {
    d.outcome = APGX_PRICE_VALUE_BOUND;
    d.cell = first_value;
  } else if (agg_cell != kNone) {
    d.outcome = overflow ? APGX_PRICE_AGGREGATE_OVERFLOW : APGX_PRICE_AGGREGATE_BOUND;
    d.cell = agg_cell;
  } else {
    d.outcome = APGX_PRICE_OK;
    d.aggregate = static_cast<uint64_t>(sum);
    d.pass = 1;
  }
  *out = d;
}

// Writes the planes of an epoch whose checks passed (a failing epoch leaves
// them untouched).
__global__ __launch_bounds__(kThreads) void PriceWrite(Params p, const uint32_t* __restrict__ usage,
                                                       const uint8_t* __restrict__ cap,
                                                       uint32_t* __restrict__ hist, uint32_t* __restrict__ present,
                                                       uint32_t* __restrict__ total, const Decision* __restrict__ d) {
  const uint64_t c = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x;
  if (c >= p.cells || d->pass == 0) return;
  const CellEval e = Evaluate(p, usage[c], CapAt(cap, c), hist[c]);
  hist[c] = static_cast<uint32_t>(e.history);
  if (present) present[c] = static_cast<uint32_t>(e.present);
  if (total) total[c] = static_cast<uint32_t>(e.total);
}

struct Buffers {
  uint32_t* usage = nullptr;
  uint8_t* cap = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* present = nullptr;
  uint32_t* total = nullptr;
  BlockPartial* part = nullptr;
  Decision* decision = nullptr;
  ~Buffers() {
    for (void* q : {static_cast<void*>(usage), static_cast<void*>(cap), static_cast<void*>(hist),
                    static_cast<void*>(present), static_cast<void*>(total), static_cast<void*>(part),
                    static_cast<void*>(decision)})
      if (q) (void)hipFree(q);
  }
};

// Policy scalars of an epoch: the present factor (checked) and the bounds.
// Returns false (with out->outcome set) when the factor overflows.
bool EpochParams(uint64_t epoch_index, const apgx_price_policy* policy, const apgx_price_limits* limits,
                 uint64_t cells, Params* p, Limits* l, apgx_price_epoch* out) {
  apgx_price_policy pol;
  apgx_price_limits lim;
  apgx_price_defaults(&pol, &lim);
  if (policy) pol = *policy;
  if (limits) lim = *limits;
  std::memset(out, 0, sizeof(*out));
  out->epoch_index = epoch_index;
  out->outcome_cell = kNone;
  *p = Params{};
  if (__builtin_mul_overflow(epoch_index, pol.present_factor_increment, &p->factor) ||
      __builtin_add_overflow(p->factor, pol.initial_present_factor, &p->factor)) {
    out->outcome = APGX_PRICE_PRESENT_FACTOR_OVERFLOW;
    return false;
  }
  out->present_factor = p->factor;
  p->inc = pol.historical_price_increment;
  p->max_value = lim.maximum_price_value;
  p->cells = cells;
  *l = Limits{lim.maximum_price_entries, lim.maximum_hot_resources, lim.maximum_aggregate_price};
  return true;
}

// One epoch on device planes with the given partial and outcome scratch:
// scan, decide, write (only when the checks pass), then the outcome record.
// `history` is updated in place; present and total may be null.
int RunEpoch(const Params& p, const Limits& l, const uint32_t* usage, const uint8_t* cap, uint32_t* history,
             uint32_t* present, uint32_t* total, BlockPartial* part, Decision* decision, apgx_price_epoch* out) {
  const uint64_t blocks = (p.cells + kBlockCells - 1) / kBlockCells;
  hipLaunchKernelGGL(PriceScan, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, nullptr, p, usage, cap,
                     history, part);
  APGX_HIP(hipGetLastError(), "price scan launch");
  hipLaunchKernelGGL(PriceDecide, dim3(1), dim3(kDecideThreads), 0, nullptr, p, l, usage, cap, history, part, blocks,
                     decision);
  APGX_HIP(hipGetLastError(), "price decision launch");
  hipLaunchKernelGGL(PriceWrite, dim3(static_cast<unsigned>((p.cells + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     nullptr, p, usage, cap, history, present, total, decision);
  APGX_HIP(hipGetLastError(), "price write launch");
  Decision d{};
  APGX_HIP(hipMemcpy(&d, decision, sizeof(d), hipMemcpyDeviceToHost), "price outcome readback");
  out->outcome = static_cast<uint32_t>(d.outcome);
  out->outcome_cell = d.cell;
  out->hot_resources = d.hot;
  out->price_entries = d.entries;
  out->aggregate_price = d.pass ? d.aggregate : 0;
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" void apgx_price_defaults(apgx_price_policy* policy, apgx_price_limits* limits) {
  if (policy) {
    policy->initial_present_factor = 100;
    policy->present_factor_increment = 100;
    policy->historical_price_increment = 10;
  }
  if (limits) {
    limits->maximum_price_entries = 1000000;
    limits->maximum_hot_resources = 1000000;
    limits->maximum_price_value = 10000000;
    limits->maximum_aggregate_price = 1000000000000ull;
  }
}

extern "C" int apgx_price_update(uint64_t epoch_index, const apgx_price_policy* policy,
                                 const apgx_price_limits* limits, uint64_t cells, const uint32_t* usage,
                                 const uint8_t* capacity, uint32_t* history, uint32_t* present, uint32_t* total,
                                 apgx_price_epoch* out) {
  using namespace apgx;
  if (out == nullptr || (cells && (usage == nullptr || history == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_update: null argument");
  if (limits && limits->maximum_price_value > 0xFFFFFFFFull)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_update: maximum_price_value must fit the u32 price planes");
  Params p;
  Limits l;
  if (!EpochParams(epoch_index, policy, limits, cells, &p, &l, out) || cells == 0) {
    ClearError();
    return APGX_OK;
  }
  const uint64_t blocks = (cells + kBlockCells - 1) / kBlockCells;
  Buffers b;
  APGX_HIP(hipMalloc(&b.usage, 4 * cells), "hipMalloc");
  APGX_HIP(hipMalloc(&b.hist, 4 * cells), "hipMalloc");
  APGX_HIP(hipMalloc(&b.part, sizeof(BlockPartial) * blocks), "hipMalloc");
  APGX_HIP(hipMalloc(&b.decision, sizeof(Decision)), "hipMalloc");
  if (present) APGX_HIP(hipMalloc(&b.present, 4 * cells), "hipMalloc");
  if (total) APGX_HIP(hipMalloc(&b.total, 4 * cells), "hipMalloc");
  APGX_HIP(hipMemcpy(b.usage, usage, 4 * cells, hipMemcpyHostToDevice), "usage upload");
  APGX_HIP(hipMemcpy(b.hist, history, 4 * cells, hipMemcpyHostToDevice), "history upload");
  if (capacity) {
    APGX_HIP(hipMalloc(&b.cap, cells), "hipMalloc");
    APGX_HIP(hipMemcpy(b.cap, capacity, cells, hipMemcpyHostToDevice), "capacity upload");
  }
  if (int rc = RunEpoch(p, l, b.usage, b.cap, b.hist, b.present, b.total, b.part, b.decision, out)) return rc;
  if (out->outcome == APGX_PRICE_OK) {
    APGX_HIP(hipMemcpy(history, b.hist, 4 * cells, hipMemcpyDeviceToHost), "history readback");
    if (present) APGX_HIP(hipMemcpy(present, b.present, 4 * cells, hipMemcpyDeviceToHost), "present readback");
    if (total) APGX_HIP(hipMemcpy(total, b.total, 4 * cells, hipMemcpyDeviceToHost), "total readback");
  }
  ClearError();
  return APGX_OK;
}

// ---------------------------------------------------------------------------
// No-allocation price epoch.  Same kernels as apgx_price_update, on device
// planes, with the block partials and the outcome record in `scratch`, a
// fixed array in this library's device image (one copy per GPU, held until
// the process ends): 65,536 partials of 64 bytes (2^26 cells at most) and the
// 64-byte record, 4 MiB + 64 B.  Its device address is looked up once per
// device, with that device current, on the entry point's first call there.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.  Nothing in it needs clearing: each call writes every partial it
// reads and the whole outcome record.
// ---------------------------------------------------------------------------
namespace apgx {
namespace {

constexpr uint64_t kNoAllocBlocks = 1ull << 16;
constexpr uint64_t kNoAllocCells = kNoAllocBlocks * kBlockCells;
constexpr size_t kScratchBytes = kNoAllocBlocks * sizeof(BlockPartial) + sizeof(Decision);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxScratchDevices] = {};

// Device address of this device's `scratch` (the array's name is a host
// stand-in, never a device address).
int ScratchAddress(void** out) {
  int device = -1;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice");
  if (device < 0 || device >= kMaxScratchDevices)
    return Fail(APGX_E_UNSUPPORTED, "apgx_price_update_no_alloc: device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_address[device] == nullptr) {
    void* ptr = nullptr;
    APGX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_address[device] = ptr;
  }
  *out = g_scratch_address[device];
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_price_update_no_alloc(uint64_t epoch_index, const apgx_price_policy* policy,
                                          const apgx_price_limits* limits, uint64_t cells, const uint32_t* usage,
                                          const uint8_t* capacity, uint32_t* history, uint32_t* present,
                                          uint32_t* total, apgx_price_epoch* out) {
  using namespace apgx;
  if (out == nullptr || (cells && (usage == nullptr || history == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_update_no_alloc: null argument");
  if (limits && limits->maximum_price_value > 0xFFFFFFFFull)
    return Fail(APGX_E_INTERNAL_INVARIANT,
                "apgx_price_update_no_alloc: maximum_price_value must fit the u32 price planes");
  if (cells > kNoAllocCells)
    return Fail(APGX_E_RESOURCE_EXHAUSTED,
                "apgx_price_update_no_alloc: at most 67108864 cells (its fixed scratch holds 65536 block partials)");
  void* base = nullptr;
  if (int rc = ScratchAddress(&base)) return rc;
  Params p;
  Limits l;
  if (!EpochParams(epoch_index, policy, limits, cells, &p, &l, out) || cells == 0) {
    ClearError();
    return APGX_OK;
  }
  auto* part = static_cast<BlockPartial*>(base);
  auto* decision = reinterpret_cast<Decision*>(static_cast<unsigned char*>(base) + kNoAllocBlocks * sizeof(BlockPartial));
  if (int rc = RunEpoch(p, l, usage, capacity, history, present, total, part, decision, out)) return rc;
  ClearError();
  return APGX_OK;
}

// ---------------------------------------------------------------------------
// Price snapshots (NegotiatedPriceSnapshot, include/apgar/allocator/
// negotiated_regeneration_plan.h:49-60) on cell resources.  Capture compacts
// the nonzero price cells of the present and history planes, in cell order, on
// the device: a counting pass per 4096-cell chunk, an exclusive scan of the
// chunk counts, and a scatter pass whose wave ballots and prefix counts give
// every entry its slot.  The identity is the reference's SnapshotIdentity
// field order (negotiated_regeneration_plan.cc:108-127) with a cell resource
// (layer, x, y) in place of an edge resource.  Validation restates
// ValidatePriorSnapshot (:211-396) check for check; restore scatters a
// validated snapshot back into dense planes so the next epoch can resume.
// ---------------------------------------------------------------------------
namespace apgx {
namespace {

constexpr int kSnapPerThread = 16;
constexpr int kSnapChunk = kThreads * kSnapPerThread;  // cells per block

__global__ __launch_bounds__(kThreads) void SnapshotCount(const uint32_t* __restrict__ present,
                                                          const uint32_t* __restrict__ history, uint64_t cells,
                                                          uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSnapChunk;
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < kSnapPerThread; ++k) {
    const uint64_t c = base + static_cast<uint64_t>(k) * kThreads + threadIdx.x;
    n += (c < cells && (present[c] | history[c]) != 0) ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off, 64);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += s_wave[w];
    counts[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kThreads) void SnapshotScatter(const uint32_t* __restrict__ present,
                                                            const uint32_t* __restrict__ history, uint64_t cells,
                                                            const uint64_t* __restrict__ offsets,
                                                            apgx_price_entry* __restrict__ out) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSnapChunk;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t slot = offsets[blockIdx.x];
  for (int k = 0; k < kSnapPerThread; ++k) {
    const uint64_t c = base + static_cast<uint64_t>(k) * kThreads + threadIdx.x;
    uint32_t p = 0, h = 0;
    if (c < cells) {
      p = present[c];
      h = history[c];
    }
    const bool keep = (p | h) != 0;
    const uint64_t mask = __ballot(keep);
    const uint32_t below = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    if (lane == 0) s_wave[wave] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      before += w < wave ? s_wave[w] : 0u;
      total += s_wave[w];
    }
    if (keep) {
      apgx_price_entry e;
      e.cell = c;
      e.present_price = p;
      e.historical_price = h;
      e.total_price = static_cast<uint64_t>(p) + h;
      out[slot + before + below] = e;
    }
    slot += total;
    __syncthreads();  // s_wave is rewritten by the next sub-chunk
  }
}

__global__ void Fill32(uint32_t* __restrict__ p, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void SnapshotRestore(const apgx_price_entry* __restrict__ entries, uint64_t count,
                                uint32_t* __restrict__ present, uint32_t* __restrict__ history) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const apgx_price_entry e = entries[i];
  present[e.cell] = static_cast<uint32_t>(e.present_price);
  history[e.cell] = static_cast<uint32_t>(e.historical_price);
}

uint64_t NonzeroFinish(const StableHash& h) {
  const uint64_t v = h.Finish();
  return v == 0 ? 1 : v;
}

}  // namespace
}  // namespace apgx

extern "C" uint64_t apgx_price_policy_identity(const apgx_price_policy* policy) {
  // PricePolicyIdentity (src/allocator/negotiated_regeneration_plan.cc:98-106).
  if (policy == nullptr) return 0;
  apgx::StableHash h;
  h.String("APGAR-P4R06-NEGOTIATED-PRICE-POLICY-V1");
  h.U64(policy->initial_present_factor);
  h.U64(policy->present_factor_increment);
  h.U64(policy->historical_price_increment);
  return apgx::NonzeroFinish(h);
}

extern "C" uint64_t apgx_price_snapshot_identity(const apgx_price_snapshot* s, const apgx_price_entry* entries) {
  if (s == nullptr || (s->price_count && entries == nullptr)) return 0;
  apgx::StableHash h;
  h.String("APGX-CELL-PRICE-SNAPSHOT-V1");
  h.U64(s->associations.board_content_hash);
  h.U64(s->associations.compiler_profile_fingerprint);
  h.U32(s->associations.geometry_compiler_version);
  h.U32(s->associations.width);
  h.U32(s->associations.height);
  h.U32(s->associations.layers);
  h.U32(s->capacity_units);
  h.U64(s->policy_identity);
  h.U64(s->prior_snapshot_identity);
  h.U64(s->epoch_index);
  h.U64(s->present_factor);
  h.U64(s->price_count);
  const uint64_t plane = static_cast<uint64_t>(s->associations.width) * s->associations.height;
  for (uint64_t i = 0; i < s->price_count; ++i) {
    const apgx_price_entry& e = entries[i];
    const uint64_t w = s->associations.width ? s->associations.width : 1;
    h.U32(plane ? static_cast<uint32_t>(e.cell / plane) : 0u);          // layer
    h.I64(static_cast<int64_t>(e.cell % w));                            // x
    h.I64(static_cast<int64_t>(plane ? (e.cell % plane) / w : 0u));     // y
    h.U64(e.present_price);
    h.U64(e.historical_price);
    h.U64(e.total_price);
  }
  return apgx::NonzeroFinish(h);
}

extern "C" int apgx_price_snapshot_capture(const apgx_price_associations* associations, uint32_t capacity_units,
                                           const apgx_price_policy* policy, uint64_t epoch_index,
                                           uint64_t prior_snapshot_identity, const uint32_t* present,
                                           const uint32_t* history, apgx_price_entry* entries,
                                           uint64_t entry_capacity, apgx_price_snapshot* out) {
  using namespace apgx;
  if (associations == nullptr || policy == nullptr || out == nullptr || present == nullptr || history == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_snapshot_capture: null argument");
  const uint64_t cells = static_cast<uint64_t>(associations->width) * associations->height * associations->layers;
  uint64_t factor = 0;
  if (__builtin_mul_overflow(epoch_index, policy->present_factor_increment, &factor) ||
      __builtin_add_overflow(factor, policy->initial_present_factor, &factor))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_snapshot_capture: present factor overflows");
  std::memset(out, 0, sizeof(*out));
  out->associations = *associations;
  out->capacity_units = capacity_units;
  out->policy_identity = apgx_price_policy_identity(policy);
  out->prior_snapshot_identity = prior_snapshot_identity;
  out->epoch_index = epoch_index;
  out->present_factor = factor;
  uint64_t total = 0;
  if (cells) {
    const uint64_t blocks = (cells + kSnapChunk - 1) / kSnapChunk;
    uint32_t *d_present = nullptr, *d_history = nullptr, *d_counts = nullptr;
    uint64_t* d_offsets = nullptr;
    apgx_price_entry* d_entries = nullptr;
    struct Guard {
      std::vector<void*> p;
      ~Guard() {
        for (void* q : p)
          if (q) (void)hipFree(q);
      }
    } guard;
    APGX_HIP(hipMalloc(&d_present, 4 * cells), "hipMalloc");
    guard.p.push_back(d_present);
    APGX_HIP(hipMalloc(&d_history, 4 * cells), "hipMalloc");
    guard.p.push_back(d_history);
    APGX_HIP(hipMalloc(&d_counts, 4 * blocks), "hipMalloc");
    guard.p.push_back(d_counts);
    APGX_HIP(hipMalloc(&d_offsets, 8 * blocks), "hipMalloc");
    guard.p.push_back(d_offsets);
    APGX_HIP(hipMemcpy(d_present, present, 4 * cells, hipMemcpyHostToDevice), "present upload");
    APGX_HIP(hipMemcpy(d_history, history, 4 * cells, hipMemcpyHostToDevice), "history upload");
    hipLaunchKernelGGL(SnapshotCount, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, nullptr, d_present,
                       d_history, cells, d_counts);
    APGX_HIP(hipGetLastError(), "snapshot count launch");
    std::vector<uint32_t> counts(blocks);
    APGX_HIP(hipMemcpy(counts.data(), d_counts, 4 * blocks, hipMemcpyDeviceToHost), "count readback");
    std::vector<uint64_t> offsets(blocks);
    for (uint64_t b = 0; b < blocks; ++b) {
      offsets[b] = total;
      total += counts[b];
    }
    out->price_count = total;
    if (entries != nullptr && total > entry_capacity)
      return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_price_snapshot_capture: entry capacity too small");
    if (entries != nullptr && total) {
      APGX_HIP(hipMalloc(&d_entries, sizeof(apgx_price_entry) * total), "hipMalloc");
      guard.p.push_back(d_entries);
      APGX_HIP(hipMemcpy(d_offsets, offsets.data(), 8 * blocks, hipMemcpyHostToDevice), "offset upload");
      hipLaunchKernelGGL(SnapshotScatter, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, nullptr, d_present,
                         d_history, cells, d_offsets, d_entries);
      APGX_HIP(hipGetLastError(), "snapshot scatter launch");
      APGX_HIP(hipMemcpy(entries, d_entries, sizeof(apgx_price_entry) * total, hipMemcpyDeviceToHost),
               "entry readback");
    }
  }
  if (entries == nullptr && total) {  // size query only
    ClearError();
    return APGX_OK;
  }
  out->snapshot_identity = apgx_price_snapshot_identity(out, entries);
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_price_snapshot_validate(const apgx_price_snapshot* prior, const apgx_price_entry* entries,
                                            const apgx_price_associations* expected, uint32_t capacity_units,
                                            const apgx_price_policy* policy, const apgx_price_limits* limits,
                                            uint64_t maximum_epoch_index, uint64_t maximum_overuse_per_epoch,
                                            apgx_price_check* out) {
  using namespace apgx;
  if (prior == nullptr || expected == nullptr || policy == nullptr || limits == nullptr || out == nullptr ||
      (prior->price_count && entries == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_snapshot_validate: null argument");
  std::memset(out, 0, sizeof(*out));
  out->entry_index = kNone;
  auto reject = [&](const char* id, uint64_t index = kNone) {
    out->valid = 0;
    out->entry_index = index;
    std::strncpy(out->invariant, id, sizeof(out->invariant) - 1);
    ClearError();
    return APGX_OK;
  };
  const apgx_price_associations& a = prior->associations;
  if (a.board_content_hash != expected->board_content_hash ||
      a.compiler_profile_fingerprint != expected->compiler_profile_fingerprint ||
      a.geometry_compiler_version != expected->geometry_compiler_version || a.width != expected->width ||
      a.height != expected->height || a.layers != expected->layers || prior->capacity_units != capacity_units)
    return reject("allocator.negotiated_plan.prior_association.v1");
  if (prior->policy_identity != apgx_price_policy_identity(policy))
    return reject("allocator.negotiated_plan.prior_policy.v1");
  if ((prior->epoch_index == 0) != (prior->prior_snapshot_identity == 0))
    return reject("allocator.negotiated_plan.prior_chain.v1");
  uint64_t factor = 0;
  if (__builtin_mul_overflow(prior->epoch_index, policy->present_factor_increment, &factor) ||
      __builtin_add_overflow(factor, policy->initial_present_factor, &factor))
    return reject("allocator.negotiated_plan.prior_schedule_overflow.v1");
  if (prior->present_factor != factor) return reject("allocator.negotiated_plan.prior_present_factor.v1");
  if (prior->price_count > limits->maximum_price_entries)
    return reject("allocator.negotiated_plan.prior_price_count.v1");
  const uint64_t cells = static_cast<uint64_t>(a.width) * a.height * a.layers;
  uint64_t aggregate = 0;
  for (uint64_t i = 0; i < prior->price_count; ++i) {
    const apgx_price_entry& e = entries[i];
    if (e.cell >= cells || (i > 0 && entries[i - 1].cell >= e.cell) ||
        (e.present_price == 0 && e.historical_price == 0))
      return reject("allocator.negotiated_plan.prior_price_order.v1", i);
    uint64_t total = 0;
    if (__builtin_add_overflow(e.present_price, e.historical_price, &total))
      return reject("allocator.negotiated_plan.prior_price_sum_overflow.v1", i);
    if (total != e.total_price) return reject("allocator.negotiated_plan.prior_price_total.v1", i);
    if (policy->historical_price_increment == 0 || e.historical_price % policy->historical_price_increment != 0)
      return reject("allocator.negotiated_plan.prior_history_schedule.v1", i);
    const uint64_t cumulative = e.historical_price / policy->historical_price_increment;
    uint64_t current = cumulative;
    if (prior->present_factor == 0) {
      if (e.present_price != 0) return reject("allocator.negotiated_plan.prior_present_schedule.v1", i);
    } else {
      if (e.present_price % prior->present_factor != 0)
        return reject("allocator.negotiated_plan.prior_present_schedule.v1", i);
      current = e.present_price / prior->present_factor;
    }
    if (current > maximum_overuse_per_epoch)
      return reject("allocator.negotiated_plan.prior_current_overuse_bound.v1", i);
    if (cumulative < current) return reject("allocator.negotiated_plan.prior_history_contribution.v1", i);
    uint64_t max_prior = 0;
    if (__builtin_mul_overflow(prior->epoch_index, maximum_overuse_per_epoch, &max_prior))
      return reject("allocator.negotiated_plan.prior_history_bound_overflow.v1", i);
    if (cumulative - current > max_prior)
      return reject("allocator.negotiated_plan.prior_cumulative_overuse_bound.v1", i);
    if (e.present_price > limits->maximum_price_value || e.historical_price > limits->maximum_price_value ||
        e.total_price > limits->maximum_price_value)
      return reject("allocator.negotiated_plan.prior_price_bound.v1", i);
    if (__builtin_add_overflow(aggregate, e.total_price, &aggregate))
      return reject("allocator.negotiated_plan.prior_aggregate_price_overflow.v1", i);
    if (aggregate > limits->maximum_aggregate_price)
      return reject("allocator.negotiated_plan.prior_aggregate_price_bound.v1", i);
  }
  if (prior->snapshot_identity == 0 || prior->snapshot_identity != apgx_price_snapshot_identity(prior, entries))
    return reject("allocator.negotiated_plan.prior_identity.v1");
  if (prior->epoch_index >= maximum_epoch_index) return reject("allocator.negotiated_plan.epoch_bound.v1");
  out->valid = 1;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_price_snapshot_restore(const apgx_price_snapshot* snapshot, const apgx_price_entry* entries,
                                           uint32_t* present, uint32_t* history) {
  using namespace apgx;
  if (snapshot == nullptr || present == nullptr || history == nullptr || (snapshot->price_count && !entries))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_snapshot_restore: null argument");
  const apgx_price_associations& a = snapshot->associations;
  const uint64_t cells = static_cast<uint64_t>(a.width) * a.height * a.layers;
  const uint64_t n = snapshot->price_count;
  for (uint64_t i = 0; i < n; ++i)
    if (entries[i].cell >= cells || entries[i].present_price > 0xFFFFFFFFull ||
        entries[i].historical_price > 0xFFFFFFFFull)
      return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_price_snapshot_restore: entry outside the u32 planes");
  if (cells == 0) {
    ClearError();
    return APGX_OK;
  }
  uint32_t *d_present = nullptr, *d_history = nullptr;
  apgx_price_entry* d_entries = nullptr;
  struct Guard {
    std::vector<void*> p;
    ~Guard() {
      for (void* q : p)
        if (q) (void)hipFree(q);
    }
  } guard;
  APGX_HIP(hipMalloc(&d_present, 4 * cells), "hipMalloc");
  guard.p.push_back(d_present);
  APGX_HIP(hipMalloc(&d_history, 4 * cells), "hipMalloc");
  guard.p.push_back(d_history);
  hipLaunchKernelGGL(Fill32, dim3(static_cast<unsigned>((cells + 255) / 256)), dim3(256), 0, nullptr, d_present, cells);
  hipLaunchKernelGGL(Fill32, dim3(static_cast<unsigned>((cells + 255) / 256)), dim3(256), 0, nullptr, d_history, cells);
  APGX_HIP(hipGetLastError(), "plane clear launch");
  if (n) {
    APGX_HIP(hipMalloc(&d_entries, sizeof(apgx_price_entry) * n), "hipMalloc");
    guard.p.push_back(d_entries);
    APGX_HIP(hipMemcpy(d_entries, entries, sizeof(apgx_price_entry) * n, hipMemcpyHostToDevice), "entry upload");
    hipLaunchKernelGGL(SnapshotRestore, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, nullptr, d_entries,
                       n, d_present, d_history);
    APGX_HIP(hipGetLastError(), "snapshot restore launch");
  }
  APGX_HIP(hipMemcpy(present, d_present, 4 * cells, hipMemcpyDeviceToHost), "present readback");
  APGX_HIP(hipMemcpy(history, d_history, 4 * cells, hipMemcpyDeviceToHost), "history readback");
  ClearError();
  return APGX_OK;
}
