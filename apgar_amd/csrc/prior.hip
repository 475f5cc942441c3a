// prior.hip — learned history-cost prior: a small convolutional
// encoder-decoder, trained and run on the GPU by hand-written kernels.
//
// The reference has no learned component; this prior only seeds the history
// plane of the negotiated allocator (apgx_negotiate), whose exact searches and
// overuse accounting still judge every route.
//
// Network (per sample, a grid of gh x gw lattice tiles, NHWC activations):
//   input   2L planes (per layer: pin density, blocked fraction), padded to 16
//   level 0 conv 16->32, conv 32->32          (gh x gw)        skip e0
//   level 1 avgpool2, conv, conv              (gh/2 x gw/2)    skip e1
//   level 2 avgpool2, conv, conv              (gh/4 x gw/4)    skip e2
//   level 3 avgpool2, conv, conv              (gh/8 x gw/8)    bottleneck
//   decoder up2(x) + skip, conv, conv   at levels 2, 1, 0
//   head    conv 32->16 (the first L outputs are the per-layer predictions)
// Every conv is 3x3, stride 1, zero padding 1, with bias; all but the head
// apply ReLU.  Convolutions run on bf16 MFMA (v_mfma_f32_16x16x32_bf16):
// weights and activations enter the matrix cores as bf16 (round to nearest
// even), products accumulate in fp32.  Master weights, biases, activations,
// gradients and Adam moments are fp32.
//
// Kernels: ConvFwd (implicit GEMM, pixels x output channels, K = 9 x Cin),
// the input gradient as ConvFwd over flipped / transposed weights, ConvWgrad
// (MFMA over pixel chunks into per-split partial sums, reduced in a fixed
// order, so gradients are deterministic), pooling / upsample-add and their
// adjoints, masked MSE loss, Adam.  A training step is captured once as a
// hipGraph and replayed with new batch indices in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "lattice_internal.h"

namespace apgx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kC = 32;        // channels of every hidden tensor
constexpr int kCin = 16;      // padded input channels
constexpr int kCout = 16;     // padded head outputs
constexpr int kConvs = 15;
constexpr int kSplitPixels = 256;  // pixels per weight-gradient partial sum

__host__ __device__ inline uint16_t ToBf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return static_cast<uint16_t>(u >> 16);
}

__device__ inline bf16x8 AsBf16(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

template <int CIN>
constexpr int PaddedK() { return (9 * CIN + 31) / 32 * 32; }

// y[p][co] = act(b[co] + sum_{tap, ci} x[shift(p, tap)][ci] * w[co][tap][ci])
// wt: bf16 [COUT][PaddedK<CIN>()] with k = tap * CIN + ci (zero beyond 9*CIN).
// One wave computes 16 pixels x COUT; four waves per block.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void ConvFwd(const float* __restrict__ x, const uint16_t* __restrict__ wt,
                                               const float* __restrict__ bias, float* __restrict__ y,
                                               int n, int h, int w, int relu) {
  constexpr int KP = PaddedK<CIN>();
  constexpr int NB = COUT / 16;
  const int lane = threadIdx.x & 63;
  const int P = n * h * w;
  const int p0 = (blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6)) * 16;
  if (p0 >= P) return;
  const int r16 = lane & 15, grp = lane >> 4;
  const int p = p0 + r16;
  const bool pv = p < P;
  const int pn = p / (h * w), py = (p / w) % h, px = p % w;
  f32x4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < KP / 32; ++c) {
    const int k0 = c * 32 + grp * 8;
    const int tap = k0 / CIN, ci0 = k0 % CIN;
    u16x8 av = {0, 0, 0, 0, 0, 0, 0, 0};
    if (tap < 9 && pv) {
      const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
      if (yy >= 0 && yy < h && xx >= 0 && xx < w) {
        const float4* src = reinterpret_cast<const float4*>(x + ((static_cast<size_t>(pn) * h + yy) * w + xx) * CIN + ci0);
        const float4 a0 = src[0], a1 = src[1];
        av = u16x8{ToBf16(a0.x), ToBf16(a0.y), ToBf16(a0.z), ToBf16(a0.w),
                   ToBf16(a1.x), ToBf16(a1.y), ToBf16(a1.z), ToBf16(a1.w)};
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const u16x8 bv = *reinterpret_cast<const u16x8*>(wt + static_cast<size_t>(b * 16 + r16) * KP + k0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AsBf16(av), AsBf16(bv), acc[b], 0, 0, 0);
    }
  }
  // Lane holds D[row = 4 * grp + i][col = r16] of each 16 x 16 block.
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int co = b * 16 + r16;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int pp = p0 + 4 * grp + i;
      if (pp < P) {
        float v = acc[b][i] + bb;
        if (relu) v = fmaxf(v, 0.f);
        y[static_cast<size_t>(pp) * COUT + co] = v;
      }
    }
  }
}

// Weight gradient partial sums:
//   part[s][co][k] = sum_{p in split s} x[shift(p, tap(k))][ci(k)] * dy[p][co]
// as an MFMA GEMM (rows k, columns co, reduction over 32-pixel chunks).
// Grid: (ceil(9*CIN/16), COUT/16, splits), one wave per block.
template <int CIN, int COUT>
__global__ __launch_bounds__(64) void ConvWgrad(const float* __restrict__ x, const float* __restrict__ dy,
                                                float* __restrict__ part, int n, int h, int w) {
  constexpr int KT = 9 * CIN;
  const int lane = threadIdx.x;
  const int r16 = lane & 15, grp = lane >> 4;
  const int kk = blockIdx.x * 16 + r16;
  const int co = blockIdx.y * 16 + r16;
  const int s = blockIdx.z;
  const int P = n * h * w;
  const int tap = kk / CIN, ci = kk % CIN;
  const int dyy = tap / 3 - 1, dxx = tap % 3 - 1;
  const bool kv = kk < KT;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int pbeg = s * kSplitPixels;
  const int pend = min(P, pbeg + kSplitPixels);
  for (int q0 = pbeg; q0 < pend; q0 += 32) {
    u16x8 av, bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int q = q0 + grp * 8 + j;
      float a = 0.f, b = 0.f;
      if (q < pend) {
        b = dy[static_cast<size_t>(q) * COUT + co];
        if (kv) {
          const int qn = q / (h * w), qy = (q / w) % h + dyy, qx = q % w + dxx;
          if (qy >= 0 && qy < h && qx >= 0 && qx < w)
            a = x[((static_cast<size_t>(qn) * h + qy) * w + qx) * CIN + ci];
        }
      }
      av[j] = ToBf16(a);
      bv[j] = ToBf16(b);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AsBf16(av), AsBf16(bv), acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = blockIdx.x * 16 + 4 * grp + i;
    if (row < KT) part[(static_cast<size_t>(s) * COUT + co) * KT + row] = acc[i];
  }
}

// dw[i] = sum over splits in order; db[co] = sum_p dy[p][co] (one block per co).
__global__ __launch_bounds__(256) void ReduceSplits(const float* __restrict__ part, float* __restrict__ dw,
                                                    int count, int splits) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int k = 0; k < splits; ++k) s += part[static_cast<size_t>(k) * count + i];
  dw[i] = s;
}

__global__ __launch_bounds__(256) void BiasGrad(const float* __restrict__ dy, float* __restrict__ db, int P,
                                                int cout) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (int p = threadIdx.x; p < P; p += 256) s += dy[static_cast<size_t>(p) * cout + co];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

// bf16 packing of one conv's fp32 master weights w[co][tap][ci]:
//   fwd[co][k = tap*cin + ci]              (padded to kp_f)
//   bwd[ci][k = (8 - tap)*cout + co]       (flipped, transposed; padded to kp_b)
// so that the input gradient is ConvFwd<cout, cin>(dy_pre, bwd).
__global__ __launch_bounds__(256) void PackWeights(const float* __restrict__ w, uint16_t* __restrict__ fwd,
                                                   uint16_t* __restrict__ bwd, int cin, int cout, int kp_f,
                                                   int kp_b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nf = cout * kp_f, nb = cin * kp_b;
  if (i < nf) {
    const int co = i / kp_f, k = i % kp_f;
    fwd[i] = k < 9 * cin ? ToBf16(w[co * 9 * cin + k]) : 0;
  }
  if (i < nb) {
    const int ci = i / kp_b, k = i % kp_b;
    uint16_t v = 0;
    if (k < 9 * cout) {
      const int tap = 8 - k / cout, co = k % cout;
      v = ToBf16(w[(co * 9 + tap) * cin + ci]);
    }
    bwd[i] = v;
  }
}

__global__ __launch_bounds__(256) void ReluMask(const float* __restrict__ post, const float* __restrict__ g,
                                                float* __restrict__ out, int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) out[i] = post[i] > 0.f ? g[i] : 0.f;
}

// out[n][y][x][c] = mean of in over the 2x2 block (in: 2h x 2w).
__global__ __launch_bounds__(256) void AvgPool2(const float* __restrict__ in, float* __restrict__ out, int n,
                                                int h, int w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n * h * w * kC) return;
  const int c = i % kC, x = (i / kC) % w, y = (i / (kC * w)) % h, b = i / (kC * w * h);
  const int W2 = 2 * w;
  const float* s = in + ((static_cast<size_t>(b) * 2 * h + 2 * y) * W2 + 2 * x) * kC + c;
  out[i] = 0.25f * (s[0] + s[kC] + s[W2 * kC] + s[W2 * kC + kC]);
}

// Adjoint of AvgPool2, accumulated: gin += gout / 4 at each covered cell.
__global__ __launch_bounds__(256) void AvgPool2Bwd(const float* __restrict__ gout, float* __restrict__ gin,
                                                   int n, int h, int w) {
  const int i = blockIdx.x * 256 + threadIdx.x;  // over the fine grid (2h x 2w)
  const int H2 = 2 * h, W2 = 2 * w;
  if (i >= n * H2 * W2 * kC) return;
  const int c = i % kC, x = (i / kC) % W2, y = (i / (kC * W2)) % H2, b = i / (kC * W2 * H2);
  gin[i] += 0.25f * gout[((static_cast<size_t>(b) * h + y / 2) * w + x / 2) * kC + c];
}

// out = up2(coarse) + skip on the fine grid (h x w = 2x coarse).
__global__ __launch_bounds__(256) void UpAdd(const float* __restrict__ coarse, const float* __restrict__ skip,
                                             float* __restrict__ out, int n, int h, int w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n * h * w * kC) return;
  const int c = i % kC, x = (i / kC) % w, y = (i / (kC * w)) % h, b = i / (kC * w * h);
  out[i] = coarse[((static_cast<size_t>(b) * (h / 2) + y / 2) * (w / 2) + x / 2) * kC + c] + skip[i];
}

// Adjoint of UpAdd: gskip = gout; gcoarse = sum of gout over each 2x2 block.
__global__ __launch_bounds__(256) void UpAddBwd(const float* __restrict__ gout, float* __restrict__ gcoarse,
                                                float* __restrict__ gskip, int n, int h, int w) {
  const int i = blockIdx.x * 256 + threadIdx.x;  // over the coarse grid
  const int hc = h / 2, wc = w / 2;
  if (i < n * h * w * kC) gskip[i] = gout[i];
  if (i >= n * hc * wc * kC) return;
  const int c = i % kC, x = (i / kC) % wc, y = (i / (kC * wc)) % hc, b = i / (kC * wc * hc);
  const float* s = gout + ((static_cast<size_t>(b) * h + 2 * y) * w + 2 * x) * kC + c;
  gcoarse[i] = s[0] + s[kC] + s[w * kC] + s[w * kC + kC];
}

// Batch assembly from the device-resident data set: sample idx[b] under the
// dihedral transform aug[b] (bit 0 flip x, bit 1 flip y, bit 2 transpose;
// transposes need a square grid).  Inputs are planar [sample][c][y][x];
// the batch is NHWC with channels padded with zeros.
__global__ __launch_bounds__(256) void Gather(const float* __restrict__ data, int channels, int cpad,
                                              const uint32_t* __restrict__ idx, const uint32_t* __restrict__ aug,
                                              float* __restrict__ out, int n, int h, int w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n * h * w * cpad) return;
  const int c = i % cpad, x = (i / cpad) % w, y = (i / (cpad * w)) % h, b = i / (cpad * w * h);
  if (c >= channels) {
    out[i] = 0.f;
    return;
  }
  const uint32_t t = aug ? aug[b] : 0u;
  int sx = x, sy = y;
  if (t & 4u) { const int tmp = sx; sx = sy; sy = tmp; }
  if (t & 1u) sx = w - 1 - sx;
  if (t & 2u) sy = h - 1 - sy;
  out[i] = data[((static_cast<size_t>(idx[b]) * channels + c) * h + sy) * w + sx];
}

// Masked MSE over the first `used` of kCout channels: one block, fixed order.
// loss[0] = mean squared error; g = d loss / d out.
__global__ __launch_bounds__(256) void MseLoss(const float* __restrict__ out, const float* __restrict__ target,
                                               float* __restrict__ g, float* __restrict__ loss, int P, int used) {
  __shared__ float red[256];
  const float inv = 1.f / static_cast<float>(P * used);
  float s = 0.f;
  for (int i = threadIdx.x; i < P * kCout; i += 256) {
    const int c = i % kCout;
    float d = 0.f;
    if (c < used) d = out[i] - target[i];
    s += d * d;
    g[i] = 2.f * d * inv;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * inv;
}

// Adam on the flat fp32 parameter vector; step t (1-based) in device memory.
__global__ __launch_bounds__(256) void AdamStep(float* __restrict__ w, const float* __restrict__ g,
                                                float* __restrict__ m, float* __restrict__ v,
                                                const uint32_t* __restrict__ step, float lr, float b1, float b2,
                                                float eps, int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const float t = static_cast<float>(*step);
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, t));
  const float vh = vi / (1.f - powf(b2, t));
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void BumpStep(uint32_t* step) { step[0] += 1u; }

__global__ __launch_bounds__(256) void FillZero(float* p, int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = 0.f;
}

inline unsigned Blocks(long long n, int per = 256) { return static_cast<unsigned>((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Host-side network description.
// ---------------------------------------------------------------------------
enum Tensor : int {
  kX, kE0a, kE0, kP1, kE1a, kE1, kP2, kE2a, kE2, kP3, kE3a, kE3,
  kU2, kD2a, kD2, kU1, kD1a, kD1, kU0, kD0a, kD0, kOut, kTensors
};
struct ConvSpec { int in, out, cin, cout, level, relu; };
// Forward order of the ops; backward is the reverse.
enum OpKind { kConv, kPool, kUp };
struct Op { OpKind kind; int a, b, c; };  // conv: index; pool: in,out; up: coarse,skip,out

constexpr ConvSpec kConvSpec[kConvs] = {
    {kX, kE0a, kCin, kC, 0, 1},   {kE0a, kE0, kC, kC, 0, 1}, {kP1, kE1a, kC, kC, 1, 1},
    {kE1a, kE1, kC, kC, 1, 1},    {kP2, kE2a, kC, kC, 2, 1}, {kE2a, kE2, kC, kC, 2, 1},
    {kP3, kE3a, kC, kC, 3, 1},    {kE3a, kE3, kC, kC, 3, 1}, {kU2, kD2a, kC, kC, 2, 1},
    {kD2a, kD2, kC, kC, 2, 1},    {kU1, kD1a, kC, kC, 1, 1}, {kD1a, kD1, kC, kC, 1, 1},
    {kU0, kD0a, kC, kC, 0, 1},    {kD0a, kD0, kC, kC, 0, 1}, {kD0, kOut, kC, kCout, 0, 0},
};
const Op kOps[] = {
    {kConv, 0, 0, 0},         {kConv, 1, 0, 0},        {kPool, kE0, kP1, 0},   {kConv, 2, 0, 0},
    {kConv, 3, 0, 0},         {kPool, kE1, kP2, 0},    {kConv, 4, 0, 0},       {kConv, 5, 0, 0},
    {kPool, kE2, kP3, 0},     {kConv, 6, 0, 0},        {kConv, 7, 0, 0},       {kUp, kE3, kE2, kU2},
    {kConv, 8, 0, 0},         {kConv, 9, 0, 0},        {kUp, kD2, kE1, kU1},   {kConv, 10, 0, 0},
    {kConv, 11, 0, 0},        {kUp, kD1, kE0, kU0},    {kConv, 12, 0, 0},      {kConv, 13, 0, 0},
    {kConv, 14, 0, 0},
};
constexpr int TensorLevel(int t) {
  return (t == kX || t == kE0a || t == kE0 || t == kU0 || t == kD0a || t == kD0 || t == kOut) ? 0
         : (t == kP1 || t == kE1a || t == kE1 || t == kU1 || t == kD1a || t == kD1)         ? 1
         : (t == kP2 || t == kE2a || t == kE2 || t == kU2 || t == kD2a || t == kD2)         ? 2
                                                                                            : 3;
}
constexpr int TensorChannels(int t) { return t == kX ? kCin : t == kOut ? kCout : kC; }

int PaddedKHost(int cin) { return (9 * cin + 31) / 32 * 32; }

}  // namespace
}  // namespace apgx

struct apgx_prior {
  apgx_prior_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t param_count = 0;
  size_t w_off[apgx::kConvs] = {}, b_off[apgx::kConvs] = {};  // into the flat vector
  size_t pf_off[apgx::kConvs] = {}, pb_off[apgx::kConvs] = {}; // into the bf16 pack
  float* d_params = nullptr;  // fp32 master
  float* d_grads = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  uint16_t* d_pack = nullptr;  // bf16 forward + backward weights of every conv
  uint32_t* d_step = nullptr;
  float* d_loss = nullptr;
  float* d_part = nullptr;     // weight-gradient partial sums
  size_t part_floats = 0;
  // Activations and gradients for up to `cap_batch` samples of gh x gw.
  int cap_batch = 0;
  float* act[apgx::kTensors] = {};
  float* grad[apgx::kTensors] = {};
  float* d_target = nullptr;
  float* d_gtmp = nullptr;     // ReLU-masked gradient scratch
  // Device data set for training.
  float* d_inputs = nullptr;
  float* d_targets = nullptr;
  uint32_t* d_idx = nullptr;
  uint32_t* d_aug = nullptr;
  uint32_t samples = 0;
  hipGraphExec_t step_graph = nullptr;
  int graph_batch = 0;
  ~apgx_prior();
};

namespace apgx {
namespace {

void FreeAll(apgx_prior* p) {
  auto f = [](void* q) { if (q) (void)hipFree(q); };
  f(p->d_params); f(p->d_grads); f(p->d_m); f(p->d_v); f(p->d_pack); f(p->d_step); f(p->d_loss); f(p->d_part);
  for (int t = 0; t < kTensors; ++t) { f(p->act[t]); f(p->grad[t]); p->act[t] = p->grad[t] = nullptr; }
  f(p->d_target); f(p->d_gtmp); f(p->d_inputs); f(p->d_targets); f(p->d_idx); f(p->d_aug);
  p->d_params = p->d_grads = p->d_m = p->d_v = p->d_loss = p->d_part = p->d_target = p->d_gtmp = nullptr;
  p->d_inputs = p->d_targets = nullptr;
  p->d_pack = nullptr;
  p->d_step = p->d_idx = p->d_aug = nullptr;
}

int Dims(const apgx_prior* p, int level, int* h, int* w) {
  *h = static_cast<int>(p->cfg.grid_h) >> level;
  *w = static_cast<int>(p->cfg.grid_w) >> level;
  return APGX_OK;
}

// Activation buffers sized for `batch` samples (reallocated when larger).
int EnsureBatch(apgx_prior* p, int batch) {
  if (batch <= p->cap_batch) return APGX_OK;
  for (int t = 0; t < kTensors; ++t) {
    if (p->act[t]) (void)hipFree(p->act[t]);
    if (p->grad[t]) (void)hipFree(p->grad[t]);
    p->act[t] = p->grad[t] = nullptr;
  }
  if (p->d_target) (void)hipFree(p->d_target);
  if (p->d_gtmp) (void)hipFree(p->d_gtmp);
  if (p->d_part) (void)hipFree(p->d_part);
  if (p->d_idx) (void)hipFree(p->d_idx);
  if (p->d_aug) (void)hipFree(p->d_aug);
  p->d_target = p->d_gtmp = p->d_part = nullptr;
  p->d_idx = p->d_aug = nullptr;
  if (p->step_graph) {
    (void)hipGraphExecDestroy(p->step_graph);
    p->step_graph = nullptr;
  }
  p->cap_batch = 0;
  for (int t = 0; t < kTensors; ++t) {
    int h, w;
    Dims(p, TensorLevel(t), &h, &w);
    const size_t bytes = 4ull * batch * h * w * TensorChannels(t);
    APGX_HIP(hipMalloc(&p->act[t], bytes), "hipMalloc");
    APGX_HIP(hipMalloc(&p->grad[t], bytes), "hipMalloc");
  }
  const size_t fine = 4ull * batch * p->cfg.grid_h * p->cfg.grid_w;
  APGX_HIP(hipMalloc(&p->d_target, fine * kCout), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_gtmp, fine * kC), "hipMalloc");
  const size_t pixels = static_cast<size_t>(batch) * p->cfg.grid_h * p->cfg.grid_w;
  const size_t splits = (pixels + kSplitPixels - 1) / kSplitPixels;
  p->part_floats = splits * 9 * kC * kC;
  APGX_HIP(hipMalloc(&p->d_part, 4 * p->part_floats), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_idx, 4ull * batch), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_aug, 4ull * batch), "hipMalloc");
  p->cap_batch = batch;
  return APGX_OK;
}

template <int CIN, int COUT>
void LaunchConvFwd(hipStream_t s, const float* x, const uint16_t* wt, const float* b, float* y, int n, int h,
                   int w, int relu) {
  const int P = n * h * w;
  hipLaunchKernelGGL((ConvFwd<CIN, COUT>), dim3(Blocks(P, 64)), dim3(256), 0, s, x, wt, b, y, n, h, w, relu);
}

template <int CIN, int COUT>
void LaunchWgrad(hipStream_t s, const float* x, const float* dy, float* part, float* dw, int n, int h, int w) {
  const int P = n * h * w;
  const int splits = (P + kSplitPixels - 1) / kSplitPixels;
  hipLaunchKernelGGL((ConvWgrad<CIN, COUT>), dim3((9 * CIN + 15) / 16, COUT / 16, splits), dim3(64), 0, s, x, dy,
                     part, n, h, w);
  const int count = COUT * 9 * CIN;
  hipLaunchKernelGGL(ReduceSplits, dim3(Blocks(count)), dim3(256), 0, s, part, dw, count, splits);
}

void PackAll(apgx_prior* p) {
  for (int i = 0; i < kConvs; ++i) {
    const ConvSpec& c = kConvSpec[i];
    const int kpf = PaddedKHost(c.cin), kpb = PaddedKHost(c.cout);
    const int count = std::max(c.cout * kpf, c.cin * kpb);
    hipLaunchKernelGGL(PackWeights, dim3(Blocks(count)), dim3(256), 0, p->stream, p->d_params + p->w_off[i],
                       p->d_pack + p->pf_off[i], p->d_pack + p->pb_off[i], c.cin, c.cout, kpf, kpb);
  }
}

void Forward(apgx_prior* p, int n) {
  hipStream_t s = p->stream;
  for (const Op& op : kOps) {
    if (op.kind == kConv) {
      const ConvSpec& c = kConvSpec[op.a];
      int h, w;
      Dims(p, c.level, &h, &w);
      const uint16_t* wt = p->d_pack + p->pf_off[op.a];
      const float* b = p->d_params + p->b_off[op.a];
      if (c.cin == kCin)
        LaunchConvFwd<kCin, kC>(s, p->act[c.in], wt, b, p->act[c.out], n, h, w, c.relu);
      else if (c.cout == kCout)
        LaunchConvFwd<kC, kCout>(s, p->act[c.in], wt, b, p->act[c.out], n, h, w, c.relu);
      else
        LaunchConvFwd<kC, kC>(s, p->act[c.in], wt, b, p->act[c.out], n, h, w, c.relu);
    } else if (op.kind == kPool) {
      int h, w;
      Dims(p, TensorLevel(op.b), &h, &w);
      hipLaunchKernelGGL(AvgPool2, dim3(Blocks(static_cast<long long>(n) * h * w * kC)), dim3(256), 0, s,
                         p->act[op.a], p->act[op.b], n, h, w);
    } else {
      int h, w;
      Dims(p, TensorLevel(op.c), &h, &w);
      hipLaunchKernelGGL(UpAdd, dim3(Blocks(static_cast<long long>(n) * h * w * kC)), dim3(256), 0, s,
                         p->act[op.a], p->act[op.b], p->act[op.c], n, h, w);
    }
  }
}

// Backward from grad[kOut] (already set): fills d_grads for every parameter.
void Backward(apgx_prior* p, int n) {
  hipStream_t s = p->stream;
  const int nops = static_cast<int>(sizeof(kOps) / sizeof(kOps[0]));
  for (int k = nops - 1; k >= 0; --k) {
    const Op& op = kOps[k];
    if (op.kind == kConv) {
      const ConvSpec& c = kConvSpec[op.a];
      int h, w;
      Dims(p, c.level, &h, &w);
      const int P = n * h * w;
      const float* gy = p->grad[c.out];
      if (c.relu) {
        hipLaunchKernelGGL(ReluMask, dim3(Blocks(static_cast<long long>(P) * c.cout)), dim3(256), 0, s,
                           p->act[c.out], p->grad[c.out], p->d_gtmp, P * c.cout);
        gy = p->d_gtmp;
      }
      float* dw = p->d_grads + p->w_off[op.a];
      float* db = p->d_grads + p->b_off[op.a];
      hipLaunchKernelGGL(BiasGrad, dim3(c.cout), dim3(256), 0, s, gy, db, P, c.cout);
      const uint16_t* wb = p->d_pack + p->pb_off[op.a];
      if (c.cin == kCin) {
        LaunchWgrad<kCin, kC>(s, p->act[c.in], gy, p->d_part, dw, n, h, w);
        // The network input needs no gradient.
      } else if (c.cout == kCout) {
        LaunchWgrad<kC, kCout>(s, p->act[c.in], gy, p->d_part, dw, n, h, w);
        LaunchConvFwd<kCout, kC>(s, gy, wb, nullptr, p->grad[c.in], n, h, w, 0);
      } else {
        LaunchWgrad<kC, kC>(s, p->act[c.in], gy, p->d_part, dw, n, h, w);
        LaunchConvFwd<kC, kC>(s, gy, wb, nullptr, p->grad[c.in], n, h, w, 0);
      }
    } else if (op.kind == kPool) {
      int h, w;
      Dims(p, TensorLevel(op.b), &h, &w);
      hipLaunchKernelGGL(AvgPool2Bwd, dim3(Blocks(static_cast<long long>(n) * 4 * h * w * kC)), dim3(256), 0, s,
                         p->grad[op.b], p->grad[op.a], n, h, w);
    } else {
      int h, w;
      Dims(p, TensorLevel(op.c), &h, &w);
      hipLaunchKernelGGL(UpAddBwd, dim3(Blocks(static_cast<long long>(n) * h * w * kC)), dim3(256), 0, s,
                         p->grad[op.c], p->grad[op.a], p->grad[op.b], n, h, w);
    }
  }
}

// Loss + gradient of the output for `n` samples already in act[kX] / d_target.
void LossAndBackward(apgx_prior* p, int n) {
  const int P = n * static_cast<int>(p->cfg.grid_h * p->cfg.grid_w);
  hipLaunchKernelGGL(MseLoss, dim3(1), dim3(256), 0, p->stream, p->act[kOut], p->d_target, p->grad[kOut],
                     p->d_loss, P, static_cast<int>(p->cfg.lattice_layers));
  Backward(p, n);
}

void GatherBatch(apgx_prior* p, const float* inputs, const float* targets, int n) {
  const int h = static_cast<int>(p->cfg.grid_h), w = static_cast<int>(p->cfg.grid_w);
  const int cin = static_cast<int>(2 * p->cfg.lattice_layers), cout = static_cast<int>(p->cfg.lattice_layers);
  hipLaunchKernelGGL(Gather, dim3(Blocks(static_cast<long long>(n) * h * w * kCin)), dim3(256), 0, p->stream,
                     inputs, cin, kCin, p->d_idx, p->d_aug, p->act[kX], n, h, w);
  if (targets)
    hipLaunchKernelGGL(Gather, dim3(Blocks(static_cast<long long>(n) * h * w * kCout)), dim3(256), 0, p->stream,
                       targets, cout, kCout, p->d_idx, p->d_aug, p->d_target, n, h, w);
}

void TrainStep(apgx_prior* p, int n, const apgx_prior_train_options& o) {
  GatherBatch(p, p->d_inputs, p->d_targets, n);
  PackAll(p);
  Forward(p, n);
  LossAndBackward(p, n);
  hipLaunchKernelGGL(BumpStep, dim3(1), dim3(1), 0, p->stream, p->d_step);
  hipLaunchKernelGGL(AdamStep, dim3(Blocks(static_cast<long long>(p->param_count))), dim3(256), 0, p->stream,
                     p->d_params, p->d_grads, p->d_m, p->d_v, p->d_step, o.learning_rate, o.beta1, o.beta2,
                     o.epsilon, static_cast<int>(p->param_count));
}

int Validate(const apgx_prior_config& c) {
  if (c.lattice_layers == 0 || 2 * c.lattice_layers > static_cast<uint32_t>(kCin) ||
      c.lattice_layers > static_cast<uint32_t>(kCout))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_create: lattice_layers must be 1..8");
  if (c.grid_h == 0 || c.grid_w == 0 || c.grid_h % 8 || c.grid_w % 8 || c.grid_h > 4096 || c.grid_w > 4096)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_create: grid sides must be positive multiples of 8");
  return APGX_OK;
}

// Deterministic normal samples (xoshiro-free: splitmix64 + Box-Muller).
struct SplitMix {
  uint64_t s;
  uint64_t Next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double Uniform() { return (static_cast<double>(Next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double Normal() {
    const double u1 = Uniform(), u2 = Uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct DeviceGuard {
  int previous = -1;
  explicit DeviceGuard(int d) {
    if (hipGetDevice(&previous) == hipSuccess && previous != d) (void)hipSetDevice(d);
    else previous = -1;
  }
  ~DeviceGuard() {
    if (previous >= 0) (void)hipSetDevice(previous);
  }
};

}  // namespace
}  // namespace apgx

apgx_prior::~apgx_prior() {
  apgx::DeviceGuard g(device);
  if (step_graph) (void)hipGraphExecDestroy(step_graph);
  apgx::FreeAll(this);
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" int apgx_prior_create(const apgx_prior_config* config, apgx_prior** out) {
  using namespace apgx;
  if (config == nullptr || out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_create: null argument");
  *out = nullptr;
  if (int rc = Validate(*config); rc != APGX_OK) return rc;
  std::unique_ptr<apgx_prior> p(new (std::nothrow) apgx_prior());
  if (!p) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_prior_create: host allocation");
  p->cfg = *config;
  APGX_HIP(hipGetDevice(&p->device), "hipGetDevice");
  APGX_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
  size_t off = 0, pack = 0;
  for (int i = 0; i < kConvs; ++i) {
    const ConvSpec& c = kConvSpec[i];
    p->w_off[i] = off;
    off += static_cast<size_t>(c.cout) * 9 * c.cin;
    p->b_off[i] = off;
    off += c.cout;
    p->pf_off[i] = pack;
    pack += static_cast<size_t>(c.cout) * PaddedKHost(c.cin);
    p->pb_off[i] = pack;
    pack += static_cast<size_t>(c.cin) * PaddedKHost(c.cout);
  }
  p->param_count = off;
  APGX_HIP(hipMalloc(&p->d_params, 4 * off), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_grads, 4 * off), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_m, 4 * off), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_v, 4 * off), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_pack, 2 * pack), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_step, 4), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_loss, 4), "hipMalloc");
  // He-normal weights on the real fan-in (padded input channels get zero
  // weights), zero biases.
  std::vector<float> host(off, 0.f);
  SplitMix rng{config->seed};
  const int real_in = static_cast<int>(2 * config->lattice_layers);
  for (int i = 0; i < kConvs; ++i) {
    const ConvSpec& c = kConvSpec[i];
    const int fan_in = 9 * (c.cin == kCin ? real_in : c.cin);
    const double sd = std::sqrt((c.relu ? 2.0 : 1.0) / fan_in);
    for (int co = 0; co < c.cout; ++co)
      for (int t = 0; t < 9; ++t)
        for (int ci = 0; ci < c.cin; ++ci) {
          const bool live = !(c.cin == kCin && ci >= real_in) &&
                            !(c.cout == kCout && co >= static_cast<int>(config->lattice_layers));
          const double v = rng.Normal() * sd;
          host[p->w_off[i] + (static_cast<size_t>(co) * 9 + t) * c.cin + ci] = live ? static_cast<float>(v) : 0.f;
        }
  }
  APGX_HIP(hipMemcpy(p->d_params, host.data(), 4 * off, hipMemcpyHostToDevice), "parameter upload");
  APGX_HIP(hipMemset(p->d_m, 0, 4 * off), "hipMemset");
  APGX_HIP(hipMemset(p->d_v, 0, 4 * off), "hipMemset");
  APGX_HIP(hipMemset(p->d_grads, 0, 4 * off), "hipMemset");
  APGX_HIP(hipMemset(p->d_step, 0, 4), "hipMemset");
  *out = p.release();
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_prior_release(apgx_prior* p) { delete p; }

extern "C" uint64_t apgx_prior_param_count(const apgx_prior* p) { return p ? p->param_count : 0; }

extern "C" int apgx_prior_get_params(const apgx_prior* p, float* params, uint64_t count) {
  using namespace apgx;
  if (p == nullptr || params == nullptr || count != p->param_count)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_get_params: null argument or wrong count");
  DeviceGuard g(p->device);
  APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  APGX_HIP(hipMemcpy(params, p->d_params, 4 * count, hipMemcpyDeviceToHost), "parameter readback");
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_prior_set_params(apgx_prior* p, const float* params, uint64_t count) {
  using namespace apgx;
  if (p == nullptr || params == nullptr || count != p->param_count)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_set_params: null argument or wrong count");
  DeviceGuard g(p->device);
  APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  APGX_HIP(hipMemcpy(p->d_params, params, 4 * count, hipMemcpyHostToDevice), "parameter upload");
  ClearError();
  return APGX_OK;
}

namespace apgx {
namespace {

// Uploads `n` planar samples and fills the batch identity index list.
int StageSamples(apgx_prior* p, const float* inputs, const float* targets, uint32_t n, float** d_in,
                 float** d_tg) {
  const size_t cells = static_cast<size_t>(p->cfg.grid_h) * p->cfg.grid_w;
  const size_t in_floats = n * cells * 2 * p->cfg.lattice_layers;
  const size_t tg_floats = n * cells * p->cfg.lattice_layers;
  APGX_HIP(hipMalloc(d_in, 4 * in_floats), "hipMalloc");
  APGX_HIP(hipMemcpyAsync(*d_in, inputs, 4 * in_floats, hipMemcpyHostToDevice, p->stream), "input upload");
  if (targets) {
    APGX_HIP(hipMalloc(d_tg, 4 * tg_floats), "hipMalloc");
    APGX_HIP(hipMemcpyAsync(*d_tg, targets, 4 * tg_floats, hipMemcpyHostToDevice, p->stream), "target upload");
  }
  std::vector<uint32_t> idx(n), aug(n, 0);
  for (uint32_t i = 0; i < n; ++i) idx[i] = i;
  APGX_HIP(hipMemcpyAsync(p->d_idx, idx.data(), 4 * n, hipMemcpyHostToDevice, p->stream), "index upload");
  APGX_HIP(hipMemcpyAsync(p->d_aug, aug.data(), 4 * n, hipMemcpyHostToDevice, p->stream), "index upload");
  // The host vectors die at return: wait for the copies.
  APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_prior_forward(apgx_prior* p, const float* inputs, uint32_t n, float* outputs) {
  using namespace apgx;
  if (p == nullptr || (n && (inputs == nullptr || outputs == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_forward: null argument");
  if (n == 0) return APGX_OK;
  DeviceGuard g(p->device);
  if (int rc = EnsureBatch(p, static_cast<int>(n)); rc != APGX_OK) return rc;
  float* d_in = nullptr;
  float* d_tg = nullptr;
  int rc = StageSamples(p, inputs, nullptr, n, &d_in, &d_tg);
  if (rc == APGX_OK) {
    GatherBatch(p, d_in, nullptr, static_cast<int>(n));
    PackAll(p);
    Forward(p, static_cast<int>(n));
    rc = hipGetLastError() == hipSuccess ? APGX_OK : Fail(APGX_E_BACKEND_FAILURE, "apgx_prior_forward: launch failed");
  }
  // Planar outputs: out[s][l][y][x] from act[kOut][s][y][x][l].
  std::vector<float> nhwc;
  if (rc == APGX_OK) {
    const size_t cells = static_cast<size_t>(p->cfg.grid_h) * p->cfg.grid_w;
    nhwc.resize(n * cells * kCout);
    const hipError_t e = hipMemcpyAsync(nhwc.data(), p->act[kOut], 4 * nhwc.size(), hipMemcpyDeviceToHost, p->stream);
    if (e != hipSuccess) rc = HipFail(e, "output readback");
    const hipError_t e2 = hipStreamSynchronize(p->stream);
    if (rc == APGX_OK && e2 != hipSuccess) rc = HipFail(e2, "hipStreamSynchronize");
    const uint32_t L = p->cfg.lattice_layers;
    if (rc == APGX_OK)
      for (uint32_t s = 0; s < n; ++s)
        for (uint32_t l = 0; l < L; ++l)
          for (size_t c = 0; c < cells; ++c) outputs[(s * L + l) * cells + c] = nhwc[(s * cells + c) * kCout + l];
  }
  if (d_in) (void)hipFree(d_in);
  if (rc == APGX_OK) ClearError();
  return rc;
}

extern "C" int apgx_prior_gradients(apgx_prior* p, const float* inputs, const float* targets, uint32_t n,
                                    float* loss, float* grads) {
  using namespace apgx;
  if (p == nullptr || n == 0 || inputs == nullptr || targets == nullptr || grads == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_gradients: null argument or empty batch");
  DeviceGuard g(p->device);
  if (int rc = EnsureBatch(p, static_cast<int>(n)); rc != APGX_OK) return rc;
  float* d_in = nullptr;
  float* d_tg = nullptr;
  int rc = StageSamples(p, inputs, targets, n, &d_in, &d_tg);
  if (rc == APGX_OK) {
    GatherBatch(p, d_in, d_tg, static_cast<int>(n));
    PackAll(p);
    Forward(p, static_cast<int>(n));
    LossAndBackward(p, static_cast<int>(n));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = HipFail(e, "gradient launch");
  }
  if (rc == APGX_OK) {
    hipError_t e = hipMemcpyAsync(grads, p->d_grads, 4 * p->param_count, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, p->d_loss, 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = HipFail(e, "gradient readback");
  }
  if (d_in) (void)hipFree(d_in);
  if (d_tg) (void)hipFree(d_tg);
  if (rc == APGX_OK) ClearError();
  return rc;
}

extern "C" void apgx_prior_train_defaults(apgx_prior_train_options* o) {
  if (o == nullptr) return;
  std::memset(o, 0, sizeof(*o));
  o->epochs = 200;
  o->batch = 32;
  o->learning_rate = 2e-3f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->epsilon = 1e-8f;
  o->augment = 1;
  o->use_graph = 1;
  o->seed = 1;
}

extern "C" int apgx_prior_train(apgx_prior* p, const float* inputs, const float* targets, uint32_t samples,
                                const apgx_prior_train_options* options, apgx_prior_train_report* report) {
  using namespace apgx;
  if (p == nullptr || inputs == nullptr || targets == nullptr || samples == 0)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_train: null argument or empty data set");
  apgx_prior_train_options o;
  apgx_prior_train_defaults(&o);
  if (options) o = *options;
  if (o.batch == 0 || o.epochs == 0) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_train: zero batch or epochs");
  if (o.augment && p->cfg.grid_h != p->cfg.grid_w)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_train: augmentation needs a square grid");
  const auto t0 = std::chrono::steady_clock::now();
  DeviceGuard g(p->device);
  const int batch = static_cast<int>(std::min(o.batch, samples));
  if (int rc = EnsureBatch(p, batch); rc != APGX_OK) return rc;
  if (p->d_inputs) (void)hipFree(p->d_inputs);
  if (p->d_targets) (void)hipFree(p->d_targets);
  p->d_inputs = p->d_targets = nullptr;
  const size_t cells = static_cast<size_t>(p->cfg.grid_h) * p->cfg.grid_w;
  const size_t in_floats = samples * cells * 2 * p->cfg.lattice_layers;
  const size_t tg_floats = samples * cells * p->cfg.lattice_layers;
  APGX_HIP(hipMalloc(&p->d_inputs, 4 * in_floats), "hipMalloc");
  APGX_HIP(hipMalloc(&p->d_targets, 4 * tg_floats), "hipMalloc");
  APGX_HIP(hipMemcpy(p->d_inputs, inputs, 4 * in_floats, hipMemcpyHostToDevice), "data set upload");
  APGX_HIP(hipMemcpy(p->d_targets, targets, 4 * tg_floats, hipMemcpyHostToDevice), "data set upload");
  p->samples = samples;

  hipEvent_t e0 = nullptr, e1 = nullptr;
  APGX_HIP(hipEventCreate(&e0), "hipEventCreate");
  APGX_HIP(hipEventCreate(&e1), "hipEventCreate");
  struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
  // Pinned staging of the per-step batch indices: one slot per step of an
  // epoch, so no copy overwrites a slot an earlier queued copy still reads.
  const uint32_t steps_per_epoch = (samples + batch - 1) / batch;
  uint32_t* pinned = nullptr;
  APGX_HIP(hipHostMalloc(&pinned, 8ull * batch * steps_per_epoch, hipHostMallocDefault), "hipHostMalloc");
  struct PinGuard { uint32_t* q; ~PinGuard() { (void)hipHostFree(q); } } pg{pinned};
  float* d_losses = nullptr;  // per-epoch mean loss (sum of step losses)
  std::vector<float> epoch_loss(o.epochs, 0.f);

  // One training step as a graph: gather, pack, forward, loss, backward, Adam.
  if (o.use_graph && (p->step_graph == nullptr || p->graph_batch != batch)) {
    if (p->step_graph) (void)hipGraphExecDestroy(p->step_graph);
    p->step_graph = nullptr;
    hipGraph_t graph = nullptr;
    APGX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    TrainStep(p, batch, o);
    APGX_HIP(hipStreamEndCapture(p->stream, &graph), "hipStreamEndCapture");
    const hipError_t inst = hipGraphInstantiate(&p->step_graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    APGX_HIP(inst, "hipGraphInstantiate");
    p->graph_batch = batch;
  }
  APGX_HIP(hipMalloc(&d_losses, 4ull * o.epochs * steps_per_epoch), "hipMalloc");
  struct LossGuard { float* q; ~LossGuard() { (void)hipFree(q); } } lg{d_losses};
  SplitMix rng{o.seed};
  std::vector<uint32_t> order(samples);
  APGX_HIP(hipEventRecord(e0, p->stream), "hipEventRecord");
  uint64_t steps = 0;
  for (uint32_t ep = 0; ep < o.epochs; ++ep) {
    for (uint32_t i = 0; i < samples; ++i) order[i] = i;
    for (uint32_t i = samples - 1; i > 0; --i) std::swap(order[i], order[rng.Next() % (i + 1)]);
    // Every slot of the pinned staging buffer is reused once per epoch: wait
    // for the previous epoch's copies before overwriting.
    if (ep) APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    for (uint32_t s = 0; s < steps_per_epoch; ++s) {
      uint32_t* slot = pinned + 2ull * batch * s;
      for (int b = 0; b < batch; ++b) {
        slot[b] = order[(static_cast<uint64_t>(s) * batch + b) % samples];
        slot[batch + b] = o.augment ? static_cast<uint32_t>(rng.Next() & 7u) : 0u;
      }
      APGX_HIP(hipMemcpyAsync(p->d_idx, slot, 4ull * batch, hipMemcpyHostToDevice, p->stream), "index upload");
      APGX_HIP(hipMemcpyAsync(p->d_aug, slot + batch, 4ull * batch, hipMemcpyHostToDevice, p->stream), "index upload");
      if (p->step_graph) {
        APGX_HIP(hipGraphLaunch(p->step_graph, p->stream), "hipGraphLaunch");
      } else {
        TrainStep(p, batch, o);
        APGX_HIP(hipGetLastError(), "train step launch");
      }
      APGX_HIP(hipMemcpyAsync(d_losses + steps, p->d_loss, 4, hipMemcpyDeviceToDevice, p->stream), "loss copy");
      ++steps;
    }
  }
  APGX_HIP(hipEventRecord(e1, p->stream), "hipEventRecord");
  std::vector<float> losses(steps);
  APGX_HIP(hipMemcpyAsync(losses.data(), d_losses, 4 * steps, hipMemcpyDeviceToHost, p->stream), "loss readback");
  APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  float ms = 0.f;
  APGX_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  if (report) {
    std::memset(report, 0, sizeof(*report));
    report->steps = steps;
    report->first_epoch_loss = 0.f;
    report->last_epoch_loss = 0.f;
    for (uint32_t s = 0; s < steps_per_epoch; ++s) {
      report->first_epoch_loss += losses[s] / steps_per_epoch;
      report->last_epoch_loss += losses[steps - steps_per_epoch + s] / steps_per_epoch;
    }
    report->device_ms = ms;
    report->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    report->graph = p->step_graph != nullptr;
  }
  ClearError();
  return APGX_OK;
}

// ---------------------------------------------------------------------------
// Board features, training targets and the decoded history plane.
// ---------------------------------------------------------------------------
namespace apgx {
namespace {

// Blocked cells per (layer, tile): one block per (layer, tile row), threads
// stride the cells of that strip of rows.
template <typename CostT>
__global__ __launch_bounds__(256) void BlockedPerTile(const CostT* __restrict__ cost, uint32_t W, uint32_t H,
                                                      uint32_t tile, uint32_t gh, uint32_t gw,
                                                      uint32_t* __restrict__ counts) {
  const uint32_t l = blockIdx.y, ty = blockIdx.x;
  const uint32_t y0 = ty * tile, y1 = min(H, y0 + tile);
  for (uint32_t i = threadIdx.x; i < (y1 - y0) * W; i += 256) {
    const uint32_t y = y0 + i / W, x = i % W;
    if (cost[(static_cast<uint64_t>(l) * H + y) * W + x] == 0)
      atomicAdd(&counts[(l * gh + ty) * gw + x / tile], 1u);
  }
}

__global__ __launch_bounds__(256) void PinsPerTile(const uint32_t* __restrict__ pins, uint64_t count, uint32_t W,
                                                   uint32_t H, uint32_t tile, uint32_t gh, uint32_t gw,
                                                   uint32_t* __restrict__ counts) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= count) return;
  const uint32_t c = pins[i];
  const uint32_t x = c % W, y = (c / W) % H, l = c / (W * H);
  atomicAdd(&counts[(l * gh + y / tile) * gw + x / tile], 1u);
}

// features[2l][ty][tx] = pins / kPinScale, features[2l+1] = blocked fraction
// (cells outside the lattice on the grid's edge count as blocked).
constexpr float kPinScale = 64.f;
__global__ __launch_bounds__(256) void Features(const uint32_t* __restrict__ pins, const uint32_t* __restrict__ blocked,
                                                uint32_t L, uint32_t gh, uint32_t gw, uint32_t W, uint32_t H,
                                                uint32_t tile, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= L * gh * gw) return;
  const uint32_t tx = i % gw, ty = (i / gw) % gh, l = i / (gw * gh);
  const uint32_t inside = (min(W, (tx + 1) * tile) - min(W, tx * tile)) * (min(H, (ty + 1) * tile) - min(H, ty * tile));
  const float area = static_cast<float>(tile * tile);
  out[(2 * l) * gh * gw + ty * gw + tx] = static_cast<float>(pins[i]) / kPinScale;
  out[(2 * l + 1) * gh * gw + ty * gw + tx] = (static_cast<float>(blocked[i]) + (area - inside)) / area;
}

// targets[l][ty][tx] = log1p(mean history of the tile's lattice cells / unit).
__global__ __launch_bounds__(256) void PoolHistory(const uint32_t* __restrict__ hist, uint32_t W, uint32_t H,
                                                   uint32_t tile, uint32_t gw, float unit,
                                                   float* __restrict__ out) {
  __shared__ double red[256];
  const uint32_t tx = blockIdx.x, ty = blockIdx.y, l = blockIdx.z;
  const uint32_t gh = gridDim.y;
  const uint32_t x0 = tx * tile, y0 = ty * tile;
  const uint32_t x1 = min(W, x0 + tile), y1 = min(H, y0 + tile);
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < tile * tile; i += 256) {
    const uint32_t x = x0 + i % tile, y = y0 + i / tile;
    if (x < x1 && y < y1) s += hist[(static_cast<uint64_t>(l) * H + y) * W + x];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < static_cast<uint32_t>(o)) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = static_cast<double>((x1 - x0) * (y1 - y0));
    out[(l * gh + ty) * gw + tx] = static_cast<float>(log1p(red[0] / n / unit));
  }
}

// history[l][y][x] = min(max_value, round(gain * unit * expm1(max(pred, 0))))
// of the cell's tile; pred is the network output act[kOut] (NHWC, 16 ch).
__global__ __launch_bounds__(256) void DecodeHistory(const float* __restrict__ pred, uint32_t W, uint32_t H,
                                                     uint32_t L, uint32_t tile, uint32_t gw, float unit, float gain,
                                                     uint32_t max_value, uint32_t* __restrict__ hist) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= static_cast<uint64_t>(W) * H * L) return;
  const uint32_t x = i % W, y = (i / W) % H, l = static_cast<uint32_t>(i / (static_cast<uint64_t>(W) * H));
  const float v = pred[((y / tile) * gw + x / tile) * kCout + l];
  const float h = gain * unit * expm1f(fmaxf(v, 0.f));
  hist[i] = h >= static_cast<float>(max_value) ? max_value : static_cast<uint32_t>(lrintf(h));
}

int FeaturesOnDevice(const apgx_lattice* lat, const uint32_t* pins, uint64_t pin_count, uint32_t tile,
                     float* d_out, hipStream_t s) {
  const apgx_lattice_desc& d = lat->desc;
  const uint32_t gh = (d.height + tile - 1) / tile, gw = (d.width + tile - 1) / tile;
  const size_t tiles = static_cast<size_t>(d.layers) * gh * gw;
  uint32_t* counts = nullptr;
  uint32_t* d_pins = nullptr;
  APGX_HIP(hipMalloc(&counts, 8 * tiles), "hipMalloc");
  struct G { uint32_t* a; uint32_t** b; ~G() { (void)hipFree(a); if (*b) (void)hipFree(*b); } } g{counts, &d_pins};
  APGX_HIP(hipMemsetAsync(counts, 0, 8 * tiles, s), "hipMemsetAsync");
  if (pin_count) {
    APGX_HIP(hipMalloc(&d_pins, 4 * pin_count), "hipMalloc");
    APGX_HIP(hipMemcpyAsync(d_pins, pins, 4 * pin_count, hipMemcpyHostToDevice, s), "pin upload");
    hipLaunchKernelGGL(PinsPerTile, dim3(Blocks(static_cast<long long>(pin_count))), dim3(256), 0, s, d_pins,
                       pin_count, d.width, d.height, tile, gh, gw, counts);
  }
  if (d.cost_bytes == 2)
    hipLaunchKernelGGL(BlockedPerTile<uint16_t>, dim3(gh, d.layers), dim3(256), 0, s,
                       static_cast<const uint16_t*>(lat->d_cost), d.width, d.height, tile, gh, gw, counts + tiles);
  else
    hipLaunchKernelGGL(BlockedPerTile<uint32_t>, dim3(gh, d.layers), dim3(256), 0, s,
                       static_cast<const uint32_t*>(lat->d_cost), d.width, d.height, tile, gh, gw, counts + tiles);
  hipLaunchKernelGGL(Features, dim3(Blocks(static_cast<long long>(tiles))), dim3(256), 0, s, counts, counts + tiles,
                     d.layers, gh, gw, d.width, d.height, tile, d_out);
  APGX_HIP(hipGetLastError(), "feature launch");
  APGX_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_prior_features(apgx_lattice* lat, const uint32_t* pins, uint64_t pin_count, uint32_t tile,
                                   float* features) {
  using namespace apgx;
  if (lat == nullptr || features == nullptr || (pin_count && pins == nullptr) || tile == 0)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_features: null argument or zero tile");
  const apgx_lattice_desc& d = lat->desc;
  const uint64_t cells = static_cast<uint64_t>(d.width) * d.height * d.layers;
  for (uint64_t i = 0; i < pin_count; ++i)
    if (pins[i] >= cells) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_features: pin outside the lattice");
  const uint32_t gh = (d.height + tile - 1) / tile, gw = (d.width + tile - 1) / tile;
  const size_t n = 2ull * d.layers * gh * gw;
  std::lock_guard<std::mutex> lease(lat->exec_mutex);
  float* d_out = nullptr;
  APGX_HIP(hipMalloc(&d_out, 4 * n), "hipMalloc");
  int rc = FeaturesOnDevice(lat, pins, pin_count, tile, d_out, lat->stream);
  if (rc == APGX_OK) {
    const hipError_t e = hipMemcpy(features, d_out, 4 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = HipFail(e, "feature readback");
  }
  (void)hipFree(d_out);
  if (rc == APGX_OK) ClearError();
  return rc;
}

extern "C" int apgx_prior_pool_history(const uint32_t* history, uint32_t width, uint32_t height, uint32_t layers,
                                       uint32_t tile, float unit, float* targets) {
  using namespace apgx;
  if (history == nullptr || targets == nullptr || tile == 0 || !(unit > 0.f))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_pool_history: null argument, zero tile or unit");
  const uint64_t cells = static_cast<uint64_t>(width) * height * layers;
  const uint32_t gh = (height + tile - 1) / tile, gw = (width + tile - 1) / tile;
  uint32_t* d_h = nullptr;
  float* d_t = nullptr;
  APGX_HIP(hipMalloc(&d_h, 4 * cells), "hipMalloc");
  struct G { uint32_t* a; float** b; ~G() { (void)hipFree(a); if (*b) (void)hipFree(*b); } } g{d_h, &d_t};
  APGX_HIP(hipMalloc(&d_t, 4ull * layers * gh * gw), "hipMalloc");
  APGX_HIP(hipMemcpy(d_h, history, 4 * cells, hipMemcpyHostToDevice), "history upload");
  hipLaunchKernelGGL(PoolHistory, dim3(gw, gh, layers), dim3(256), 0, nullptr, d_h, width, height, tile, gw, unit, d_t);
  APGX_HIP(hipGetLastError(), "pool launch");
  APGX_HIP(hipMemcpy(targets, d_t, 4ull * layers * gh * gw, hipMemcpyDeviceToHost), "target readback");
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_prior_seed_history(apgx_prior* p, apgx_lattice* lat, const uint32_t* pins, uint64_t pin_count,
                                       const apgx_prior_decode* decode, uint32_t* history) {
  using namespace apgx;
  if (p == nullptr || lat == nullptr || decode == nullptr || history == nullptr || (pin_count && pins == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_seed_history: null argument");
  const apgx_lattice_desc& d = lat->desc;
  const uint32_t tile = decode->tile;
  if (tile == 0 || d.layers != p->cfg.lattice_layers || (d.height + tile - 1) / tile != p->cfg.grid_h ||
      (d.width + tile - 1) / tile != p->cfg.grid_w)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_seed_history: lattice does not match the network grid");
  if (!(decode->unit > 0.f) || !(decode->gain >= 0.f))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_seed_history: unit must be positive, gain non-negative");
  const uint64_t cells = static_cast<uint64_t>(d.width) * d.height * d.layers;
  for (uint64_t i = 0; i < pin_count; ++i)
    if (pins[i] >= cells) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_prior_seed_history: pin outside the lattice");
  DeviceGuard g(p->device);
  if (int rc = EnsureBatch(p, 1); rc != APGX_OK) return rc;
  std::lock_guard<std::mutex> lease(lat->exec_mutex);
  float* d_feat = nullptr;
  uint32_t* d_hist = nullptr;
  APGX_HIP(hipMalloc(&d_feat, 8ull * d.layers * p->cfg.grid_h * p->cfg.grid_w), "hipMalloc");
  struct G { float* a; uint32_t** b; ~G() { (void)hipFree(a); if (*b) (void)hipFree(*b); } } gd{d_feat, &d_hist};
  APGX_HIP(hipMalloc(&d_hist, 4 * cells), "hipMalloc");
  if (int rc = FeaturesOnDevice(lat, pins, pin_count, tile, d_feat, p->stream); rc != APGX_OK) return rc;
  const uint32_t zero = 0;
  APGX_HIP(hipMemcpyAsync(p->d_idx, &zero, 4, hipMemcpyHostToDevice, p->stream), "index upload");
  APGX_HIP(hipMemcpyAsync(p->d_aug, &zero, 4, hipMemcpyHostToDevice, p->stream), "index upload");
  GatherBatch(p, d_feat, nullptr, 1);
  PackAll(p);
  Forward(p, 1);
  hipLaunchKernelGGL(DecodeHistory, dim3(Blocks(static_cast<long long>(cells))), dim3(256), 0, p->stream,
                     p->act[kOut], d.width, d.height, d.layers, tile, p->cfg.grid_w, decode->unit, decode->gain,
                     decode->max_value, d_hist);
  APGX_HIP(hipGetLastError(), "decode launch");
  APGX_HIP(hipMemcpyAsync(history, d_hist, 4 * cells, hipMemcpyDeviceToHost, p->stream), "history readback");
  APGX_HIP(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  ClearError();
  return APGX_OK;
}
