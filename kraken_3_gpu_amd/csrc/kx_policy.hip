// Policy network and evolution strategy on the device: the batched forward
// pass of src/brain/net.js:71-124 (forwardInto + decode) over a population
// with one genome per creature, and the ES generation of src/brain/es.js:67-149
// (genomeAt for every member, update) with the trainer's mulberry32 stream
// (src/sim/train.js:65-72) drawn by index instead of in sequence.
//
// Numerics follow the reference's JS arithmetic: float64 operations on
// float32-stored values, each result rounded where the reference stores into
// a Float32Array. Sums run in the reference's order; a product of two float32
// values is exact in float64, so fma(x, w, s) equals JS's s + x*w. The ES
// kernels compile with contraction off so every float64 operation rounds as
// its JS counterpart does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/kxpbd.h"
#include "kx_internal.h"

namespace {

constexpr int KX_POL_MAX_ROWS = 64;      // hidden and n_out: one lane per row
constexpr int KX_POL_MAX_IN = 1024;
constexpr size_t KX_POL_MAX_LDS = 64 * 1024;

int pfail(int code, const char *fmt, ...) {
  char buf[384];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return kx_error(code, buf);
}

#define PCHK(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return pfail(KX_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct PolDims {
  int n_in, hidden, n_out, nw;
  int b1, b2, b4, b5;   // offsets of net.js:73-77 (b3 = b2 + hidden^2 is never read)
};

PolDims pol_dims(int n_in, int hidden, int n_out) {
  PolDims d;
  d.n_in = n_in;
  d.hidden = hidden;
  d.n_out = n_out;
  d.b1 = n_in * hidden;
  d.b2 = d.b1 + hidden;
  int b3 = d.b2 + hidden * hidden;
  d.b4 = b3 + hidden;
  d.b5 = d.b4 + hidden * n_out;
  d.nw = d.b5 + n_out;
  return d;
}

__device__ __forceinline__ double kx_act(double s) {
  // net.js:81: s > 20 -> 1, s < -20 -> -1, else tanh(s) (NaN passes through)
  return s > 20.0 ? 1.0 : (s < -20.0 ? -1.0 : tanh(s));
}

typedef float kx_v4f __attribute__((ext_vector_type(4)));   // native 16-byte vector (register-resident)
constexpr int KX_POL_V4 = 16;   // genome float4 per lane held in registers (up to 4096 weights)
constexpr int KX_POL_XR = 4;    // input floats per lane held in registers (up to 256 inputs)

// s + sum_i a[i]*b[i] for i = 0 .. n-1 in that order, in float64 (a float32
// product is exact in float64, so each fma is the reference's s += a*b). The
// LDS reads go in batches of 16 ahead of their fmas, so one LDS latency covers
// sixteen steps of the chain.
__device__ __forceinline__ double kx_dot_seq(double s, const float *a, const float *b, int n) {
  int i = 0;
  for (; i + 16 <= n; i += 16) {
    float av[16], bv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      av[u] = a[i + u];
      bv[u] = b[i + u];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) s = fma((double)av[u], (double)bv[u], s);
  }
  for (; i < n; i++) s = fma((double)a[i], (double)b[i], s);
  return s;
}

// The three layers and decode for the creature whose genome (W), inputs (X) are
// staged in LDS; lane j computes row j of each layer.
__device__ __forceinline__ void kx_policy_eval(const PolDims &d, int c, int lane, const float *W, const float *X,
                                               float *H1, float *H2, float *O, float *__restrict__ out,
                                               float *__restrict__ scratch, float *__restrict__ controls, int arms) {
  if (lane < d.hidden) H1[lane] = (float)kx_act(kx_dot_seq((double)W[d.b1 + lane], X, W + lane * d.n_in, d.n_in));
  __syncthreads();
  // layer-2 bias read at b2 + j as net.js:84-86 does (the first weight row)
  if (lane < d.hidden)
    H2[lane] = (float)kx_act(kx_dot_seq((double)W[d.b2 + lane], H1, W + d.b2 + lane * d.hidden, d.hidden));
  __syncthreads();
  if (lane < d.n_out) {
    const float o = (float)kx_act(kx_dot_seq((double)W[d.b5 + lane], H2, W + d.b4 + lane * d.hidden, d.hidden));
    out[(size_t)c * d.n_out + lane] = o;
    O[lane] = o;
  }
  if (scratch && lane < d.hidden) {
    scratch[(size_t)c * 2 * d.hidden + lane] = H1[lane];
    scratch[(size_t)c * 2 * d.hidden + d.hidden + lane] = H2[lane];
  }
  if (controls) {
    // decode (net.js:114-124) into a control-tape record: [0] tone, then
    // shoulder, curl, adhesion per arm; words of absent arms are zero
    __syncthreads();
    float *rec = controls + (size_t)c * KX_TAPE_WORDS;
    if (lane < KX_TAPE_WORDS) {
      float v = 0.f;
      if (lane == 0) {
        v = (float)(0.5 + 0.5 * (double)O[0]);
      } else {
        const int a = (lane - 1) / 3, f = (lane - 1) % 3;
        if (a < arms) {
          const float u = O[1 + a * 3 + f];   // index < 1 + 3*arms <= n_out (host check)
          v = (f == 2) ? (u > 0.f ? u : 0.f) : u;
        }
      }
      rec[lane] = v;
    }
  }
}

// Lane `lane`'s share of creature cc's genome (float4 k = lane + 64 r) and
// inputs, loaded into registers.
__device__ __forceinline__ void kx_pol_fetch(kx_v4f (&gr)[KX_POL_V4], float (&xr)[KX_POL_XR],
                                             const float *__restrict__ genomes, const float *__restrict__ x,
                                             const PolDims &d, int cc, int n4, int lane) {
  const kx_v4f *g4 = reinterpret_cast<const kx_v4f *>(genomes + (size_t)cc * d.nw);
#pragma unroll
  for (int r = 0; r < KX_POL_V4; r++) {
    const int k = lane + 64 * r;
    if (k < n4) gr[r] = g4[k];
  }
  const float *xs = x + (size_t)cc * d.n_in;
#pragma unroll
  for (int r = 0; r < KX_POL_XR; r++) {
    const int k = lane + 64 * r;
    if (k < d.n_in) xr[r] = xs[k];
  }
}

// One wave per creature at a time, looping over creatures c = blockIdx.x,
// + gridDim.x, ... The genome (nW floats), the input row, both hidden
// activations and the outputs are staged in LDS. With `pipe` (nW a multiple
// of 4, at most 64 * KX_POL_V4 float4, n_in <= 64 * KX_POL_XR), each lane
// holds its share of the next creature's genome and inputs in registers: the
// loads are issued before the current creature is evaluated, so HBM reads
// overlap the arithmetic; otherwise the genome is copied through LDS directly.
__global__ __launch_bounds__(64) void kx_k_policy_fwd(PolDims d, int n, const float *__restrict__ genomes,
                                                      const float *__restrict__ x, float *__restrict__ out,
                                                      float *__restrict__ scratch, float *__restrict__ controls,
                                                      int arms, int pipe) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  float *W = lds;
  float *X = W + d.nw;
  float *H1 = X + d.n_in;
  float *H2 = H1 + d.hidden;
  float *O = H2 + d.hidden;   // n_out floats: the outputs decode reads
  int c = blockIdx.x;
  if (c >= n) return;
  if (pipe) {
    const int n4 = d.nw >> 2;
    kx_v4f gr[KX_POL_V4];
    float xr[KX_POL_XR];
    kx_pol_fetch(gr, xr, genomes, x, d, c, n4, lane);
    kx_v4f *W4 = reinterpret_cast<kx_v4f *>(W);
    for (; c < n; c += gridDim.x) {
#pragma unroll
      for (int r = 0; r < KX_POL_V4; r++) {
        const int k = lane + 64 * r;
        if (k < n4) W4[k] = gr[r];
      }
#pragma unroll
      for (int r = 0; r < KX_POL_XR; r++) {
        const int k = lane + 64 * r;
        if (k < d.n_in) X[k] = xr[r];
      }
      __syncthreads();
      if (c + (int)gridDim.x < n) kx_pol_fetch(gr, xr, genomes, x, d, c + gridDim.x, n4, lane);
      kx_policy_eval(d, c, lane, W, X, H1, H2, O, out, scratch, controls, arms);
      __syncthreads();   // LDS is rewritten for the next creature
    }
  } else {
    for (; c < n; c += gridDim.x) {
      const float *g = genomes + (size_t)c * d.nw;
      for (int k = lane; k < d.nw; k += 64) W[k] = g[k];
      const float *xs = x + (size_t)c * d.n_in;
      for (int k = lane; k < d.n_in; k += 64) X[k] = xs[k];
      __syncthreads();
      kx_policy_eval(d, c, lane, W, X, H1, H2, O, out, scratch, controls, arms);
      __syncthreads();
    }
  }
}

// The value of call idx of a mulberry32 stream with state `st` before its
// first call: the state advances by 0x6D2B79F5 per call (src/world/map.js:11-20).
__device__ __forceinline__ double kx_mb32_at(uint32_t st, uint64_t idx) {
  uint32_t a = st + (uint32_t)(idx + 1) * 0x6D2B79F5u;
  uint32_t t = a;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return (double)(t ^ (t >> 14)) * (1.0 / 4294967296.0);
}

// genomeAt for every member of a generation (es.js:67-88 as called by
// train.js:154-158): thread (p, k) draws eps_p[k] (four calls, k-major within
// the pair, pairs in order) and writes both mirrored genomes.
__global__ __launch_bounds__(256) void kx_k_es_genomes(int nw, int pairs, uint32_t st, double sigma,
                                                       const float *__restrict__ theta, float *__restrict__ eps_bank,
                                                       float *__restrict__ genomes) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)pairs * nw) return;
  const int p = (int)(t / nw), k = (int)(t - (int64_t)p * nw);
  const uint64_t base = (uint64_t)t * 4;
  const double r0 = kx_mb32_at(st, base), r1 = kx_mb32_at(st, base + 1);
  const double r2 = kx_mb32_at(st, base + 2), r3 = kx_mb32_at(st, base + 3);
  const float e = (float)(((((r0 + r1) + r2) + r3) - 2.0) * 1.7320508);   // gauss01, es.js:84-86
  eps_bank[t] = e;
  const double th = (double)theta[k];
  genomes[(size_t)(2 * p) * nw + k] = (float)(th + (double)e * sigma);
  genomes[(size_t)(2 * p + 1) * nw + k] = (float)(th + (double)e * -sigma);
}

struct KxEsStep {
  double inv, clip, b1, b2, omb1, omb2, c1, c2, lr, eps;
};

constexpr int KX_ES_BATCH = 64;   // eps loads in flight per thread ahead of the additions

// update (es.js:107-135) per weight: the gradient accumulated in float32 over
// the pairs in order, scaled and clipped, then one Adam step. The additions
// are a chain per weight (each rounds to float32), so the loads run ahead of
// them in batches; the pair advantages sit in LDS.
__global__ __launch_bounds__(64) void kx_k_es_update(int nw, int half, KxEsStep q, const double *__restrict__ diffs,
                                                     const float *__restrict__ eps_bank, float *__restrict__ theta,
                                                     float *__restrict__ m, float *__restrict__ v) {
#pragma clang fp contract(off)
  extern __shared__ double dsh[];   // half advantages
  for (int p = threadIdx.x; p < half; p += blockDim.x) dsh[p] = diffs[p];
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  float g = 0.f;
  int p = 0;
  for (; p + KX_ES_BATCH <= half; p += KX_ES_BATCH) {
    float e[KX_ES_BATCH];
#pragma unroll
    for (int u = 0; u < KX_ES_BATCH; u++) e[u] = eps_bank[(size_t)(p + u) * nw + k];
#pragma unroll
    for (int u = 0; u < KX_ES_BATCH; u++) g = (float)((double)g + dsh[p + u] * (double)e[u]);
  }
  for (; p < half; p++) g = (float)((double)g + dsh[p] * (double)eps_bank[(size_t)p * nw + k]);
  double a = (double)g * q.inv;
  if (a > q.clip) a = q.clip;
  else if (a < -q.clip) a = -q.clip;
  const double gk = (double)(float)a;
  const float mk = (float)(q.b1 * (double)m[k] + q.omb1 * gk);
  const float vk = (float)(q.b2 * (double)v[k] + q.omb2 * gk * gk);
  m[k] = mk;
  v[k] = vk;
  const double step = q.lr * ((double)mk / q.c1) / (sqrt((double)vk / q.c2) + q.eps);
  if (isfinite(step)) theta[k] = (float)((double)theta[k] + step);
}

// The policy inputs of gputrain.js:77-82 for every creature: readProprio's
// floats from the device readout, then the senses (zeros when none are
// given), then the creature's previous outputs.
__global__ __launch_bounds__(256) void kx_k_policy_inputs(int n, int n_in, int n_out, int proprio, int sense_dim,
                                                          const float *__restrict__ state, int stride,
                                                          const float *__restrict__ senses,
                                                          const float *__restrict__ prev, float *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n_in) return;
  const int c = (int)(t / n_in), k = (int)(t - (int64_t)c * n_in);
  float v;
  if (k < proprio) v = state[(size_t)c * stride + k];
  else if (k < proprio + sense_dim) v = senses ? senses[(size_t)c * sense_dim + (k - proprio)] : 0.f;
  else v = prev[(size_t)c * n_out + (k - proprio - sense_dim)];
  x[t] = v;
}

}  // namespace

struct kx_policy {
  int n = 0;
  PolDims d{};
  float *genomes = nullptr, *x = nullptr, *out = nullptr, *scratch = nullptr, *controls = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  int grid = 0;   // workgroups of the forward launch (set at its first launch)
  float *senses = nullptr;
  float *senses_host = nullptr;   // pinned staging of the senses (hipHostMalloc)
  int senses_cap = 0;             // floats of both
  hipEvent_t senses_done = nullptr;   // the last upload from the staging buffer
  // the end of the last kx_policy_drive's work on the policy's buffers
  // (senses, inputs, outputs), on whichever world stream it ran
  hipEvent_t drive_done = nullptr;
  bool drove = false;
};

struct kx_es {
  int nw = 0;
  kx_es_opts o{};
  double sigma = 0, best_ever = -INFINITY;
  int t = 0, have_eps = 0, have_best = 0;
  float *theta = nullptr, *m = nullptr, *v = nullptr, *eps = nullptr, *best = nullptr;
  double *diffs = nullptr;
  hipStream_t st = nullptr;
};

extern "C" int32_t kx_policy_weights(int32_t n_in, int32_t hidden, int32_t n_out) {
  if (n_in <= 0 || hidden <= 0 || n_out <= 0) return 0;
  return pol_dims(n_in, hidden, n_out).nw;
}

extern "C" int kx_policy_create(int32_t n, int32_t n_in, int32_t hidden, int32_t n_out, kx_policy **out) {
  if (!out) return pfail(KX_EINVAL, "null argument");
  *out = nullptr;
  if (n <= 0 || n_in <= 0 || hidden <= 0 || n_out <= 0)
    return pfail(KX_EINVAL, "policy of %d creatures, %d -> %d -> %d", n, n_in, hidden, n_out);
  if (hidden > KX_POL_MAX_ROWS || n_out > KX_POL_MAX_ROWS || n_in > KX_POL_MAX_IN)
    return pfail(KX_EINVAL, "policy %d -> %d -> %d: hidden and outputs at most %d, inputs at most %d", n_in, hidden,
                 n_out, KX_POL_MAX_ROWS, KX_POL_MAX_IN);
  PolDims d = pol_dims(n_in, hidden, n_out);
  size_t lds = (size_t)(d.nw + n_in + 2 * hidden + n_out) * sizeof(float);
  if (lds > KX_POL_MAX_LDS) return pfail(KX_EINVAL, "genome of %d weights exceeds the %zu-byte LDS stage", d.nw, KX_POL_MAX_LDS);
  kx_policy *p = new kx_policy();
  p->n = n;
  p->d = d;
  auto cleanup = [&](int rc) {
    kx_policy_destroy(p);
    return rc;
  };
  if (hipMalloc((void **)&p->genomes, (size_t)n * d.nw * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&p->x, (size_t)n * n_in * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&p->out, (size_t)n * n_out * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&p->scratch, (size_t)n * 2 * hidden * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&p->controls, (size_t)n * KX_TAPE_WORDS * sizeof(float)) != hipSuccess)
    return cleanup(pfail(KX_ENOMEM, "policy buffers for %d creatures", n));
  if (hipStreamCreate(&p->st) != hipSuccess || hipEventCreate(&p->ev0) != hipSuccess ||
      hipEventCreate(&p->ev1) != hipSuccess)
    return cleanup(pfail(KX_EDEVICE, "policy stream"));
  if (hipMemsetAsync(p->genomes, 0, (size_t)n * d.nw * sizeof(float), p->st) != hipSuccess ||
      hipMemsetAsync(p->out, 0, (size_t)n * n_out * sizeof(float), p->st) != hipSuccess ||
      hipMemsetAsync(p->x, 0, (size_t)n * n_in * sizeof(float), p->st) != hipSuccess ||
      hipStreamSynchronize(p->st) != hipSuccess)
    return cleanup(pfail(KX_EDEVICE, "policy clear"));
  *out = p;
  return 0;
}

extern "C" int kx_policy_destroy(kx_policy *p) {
  if (!p) return 0;
  if (p->st) (void)hipStreamSynchronize(p->st);
  if (p->drove) (void)hipEventSynchronize(p->drive_done);   // a drive on a world stream may still read them
  (void)hipFree(p->genomes);
  (void)hipFree(p->x);
  (void)hipFree(p->out);
  (void)hipFree(p->scratch);
  (void)hipFree(p->controls);
  (void)hipFree(p->senses);
  if (p->senses_host) (void)hipHostFree(p->senses_host);
  if (p->senses_done) (void)hipEventDestroy(p->senses_done);
  if (p->drive_done) (void)hipEventDestroy(p->drive_done);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  if (p->st) (void)hipStreamDestroy(p->st);
  delete p;
  return 0;
}

extern "C" int kx_policy_set_genomes(kx_policy *p, const float *genomes) {
  if (!p || !genomes) return pfail(KX_EINVAL, "null argument");
  PCHK(hipMemcpyAsync(p->genomes, genomes, (size_t)p->n * p->d.nw * sizeof(float), hipMemcpyHostToDevice, p->st));
  PCHK(hipStreamSynchronize(p->st));
  return 0;
}

extern "C" int kx_policy_get_genomes(const kx_policy *p, float *genomes) {
  if (!p || !genomes) return pfail(KX_EINVAL, "null argument");
  PCHK(hipMemcpyAsync(genomes, p->genomes, (size_t)p->n * p->d.nw * sizeof(float), hipMemcpyDeviceToHost, p->st));
  PCHK(hipStreamSynchronize(p->st));
  return 0;
}

static int policy_launch(kx_policy *p, int arms, bool want_controls, bool want_scratch,
                         hipStream_t st = nullptr, float *controls_dst = nullptr) {
  const PolDims &d = p->d;
  const size_t lds = (size_t)(d.nw + d.n_in + 2 * d.hidden + d.n_out) * sizeof(float);
  // register pipeline when the genome is whole float4s (hipMalloc bases are
  // 256-byte aligned, so every genome then starts 16-byte aligned) and fits
  const int pipe = (d.nw & 3) == 0 && (d.nw >> 2) <= 64 * KX_POL_V4 && d.n_in <= 64 * KX_POL_XR;
  if (p->grid == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    PCHK(hipGetDevice(&dev));
    PCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kx_k_policy_fwd, 64, lds));
    const long slots = (long)std::max(1, cus) * std::max(1, per_cu);
    // creatures per wave: up to two, each wave prefetching its second while
    // evaluating its first; past two rounds of resident waves, one creature
    // per wave and the hardware's wave scheduling (measured at 4096 and
    // 16384 creatures: 24.0 vs 24.9 us, 69.0 vs 61.5 us)
    long rounds = (p->n + slots - 1) / slots;
    if (rounds > 2) rounds = 1;
    p->grid = (int)((p->n + rounds - 1) / rounds);
  }
  if (!st) st = p->st;
  float *ctl = want_controls ? (controls_dst ? controls_dst : p->controls) : nullptr;
  PCHK(hipEventRecord(p->ev0, st));
  hipLaunchKernelGGL(kx_k_policy_fwd, dim3(p->grid), dim3(64), lds, st, d, p->n, p->genomes, p->x, p->out,
                     want_scratch ? p->scratch : nullptr, ctl, arms, pipe);
  PCHK(hipGetLastError());
  PCHK(hipEventRecord(p->ev1, st));
  return 0;
}

extern "C" int kx_policy_forward(kx_policy *p, const float *x, float *out, float *scratch, float *controls,
                                 int32_t arms) {
  if (!p || !x || !out) return pfail(KX_EINVAL, "null argument");
  if (controls && (arms < 0 || arms > KX_MAX_ARMS || p->d.n_out < 1 + 3 * arms))
    return pfail(KX_EINVAL, "decode of %d arms from %d outputs (at most %d arms)", arms, p->d.n_out, KX_MAX_ARMS);
  const PolDims &d = p->d;
  PCHK(hipMemcpyAsync(p->x, x, (size_t)p->n * d.n_in * sizeof(float), hipMemcpyHostToDevice, p->st));
  int rc = policy_launch(p, arms, controls != nullptr, scratch != nullptr);
  if (rc) return rc;
  PCHK(hipMemcpyAsync(out, p->out, (size_t)p->n * d.n_out * sizeof(float), hipMemcpyDeviceToHost, p->st));
  if (scratch)
    PCHK(hipMemcpyAsync(scratch, p->scratch, (size_t)p->n * 2 * d.hidden * sizeof(float), hipMemcpyDeviceToHost,
                        p->st));
  if (controls)
    PCHK(hipMemcpyAsync(controls, p->controls, (size_t)p->n * KX_TAPE_WORDS * sizeof(float), hipMemcpyDeviceToHost,
                        p->st));
  PCHK(hipStreamSynchronize(p->st));
  PCHK(hipEventElapsedTime(&p->last_ms, p->ev0, p->ev1));
  return 0;
}

extern "C" int kx_policy_drive(kx_policy *p, kx_handle *h, const float *senses, int32_t sense_dim,
                               const kx_law *law, int32_t step) {
  if (!p || !h || !law) return pfail(KX_EINVAL, "null argument");
  if (sense_dim < 0) return pfail(KX_EINVAL, "sense_dim %d", sense_dim);
  const PolDims &d = p->d;
  float *state = nullptr, *tape = nullptr;
  hipStream_t st = nullptr;
  int32_t n_cr = 0, arms = 0;
  const int stride = KX_STATE_EXTRA + 6 * KX_MAX_ARMS + 6;
  // every check before anything is launched (the readout advances the
  // previous mantle velocity, so a refused call must not run it)
  int need = kx_policy_io(h, stride, &state, &tape, &st, &n_cr, &arms, 0);
  if (need < 0) return need;
  const int proprio = 6 * arms + 6;
  if (n_cr != p->n) return pfail(KX_EINVAL, "policy of %d creatures for a world of %d", p->n, n_cr);
  if (proprio + sense_dim + d.n_out != d.n_in)
    return pfail(KX_EINVAL, "inputs: %d proprio + %d senses + %d outputs != %d", proprio, sense_dim, d.n_out, d.n_in);
  if (d.n_out < 1 + 3 * arms) return pfail(KX_EINVAL, "decode of %d arms from %d outputs", arms, d.n_out);
  need = kx_policy_io(h, stride, &state, &tape, &st, &n_cr, &arms, 1);
  if (need < 0) return need;
  // the previous drive may have run on another world's stream: its reads and
  // writes of the shared senses, inputs and outputs come before this one's
  if (!p->drive_done) PCHK(hipEventCreateWithFlags(&p->drive_done, hipEventDisableTiming));
  if (p->drove) PCHK(hipStreamWaitEvent(st, p->drive_done, 0));
  const float *sd = nullptr;
  if (senses && sense_dim > 0) {
    const int nf = p->n * sense_dim;
    if (nf > p->senses_cap) {
      if (p->drove) PCHK(hipEventSynchronize(p->drive_done));
      PCHK(hipStreamSynchronize(st));
      (void)hipFree(p->senses);
      if (p->senses_host) (void)hipHostFree(p->senses_host);
      p->senses = nullptr;
      p->senses_host = nullptr;
      p->senses_cap = 0;
      if (hipMalloc((void **)&p->senses, (size_t)nf * sizeof(float)) != hipSuccess ||
          hipHostMalloc((void **)&p->senses_host, (size_t)nf * sizeof(float), hipHostMallocDefault) != hipSuccess)
        return pfail(KX_ENOMEM, "senses of %d creatures", p->n);
      if (!p->senses_done) PCHK(hipEventCreateWithFlags(&p->senses_done, hipEventDisableTiming));
      p->senses_cap = nf;
    } else if (p->senses_done) {
      PCHK(hipEventSynchronize(p->senses_done));   // the previous upload has left the staging buffer
    }
    // the caller's (pageable) senses go through the pinned staging buffer, so
    // the upload itself is an asynchronous DMA on the world's stream
    std::memcpy(p->senses_host, senses, (size_t)nf * sizeof(float));
    PCHK(hipMemcpyAsync(p->senses, p->senses_host, (size_t)nf * sizeof(float), hipMemcpyHostToDevice, st));
    PCHK(hipEventRecord(p->senses_done, st));
    sd = p->senses;
  }
  const int64_t tot = (int64_t)p->n * d.n_in;
  hipLaunchKernelGGL(kx_k_policy_inputs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, p->n, d.n_in, d.n_out,
                     proprio, sense_dim, state, stride, sd, p->out, p->x);
  PCHK(hipGetLastError());
  int rc = policy_launch(p, arms, true, false, st, tape);
  if (rc) return rc;
  PCHK(hipEventRecord(p->drive_done, st));
  p->drove = true;
  // later work on the policy's own stream (forward, ES generation) touches
  // the same inputs, outputs and genomes: order it after this forward pass
  PCHK(hipStreamWaitEvent(p->st, p->ev1, 0));
  kx_law L = *law;
  L.flags |= KX_LAW_TAPE;
  return kx_rollout(h, &L, step, 1, nullptr);
}

extern "C" int kx_policy_forward_repeat(kx_policy *p, int32_t reps, float *ms) {
  if (!p || reps <= 0) return pfail(KX_EINVAL, "policy repeat of %d", reps);
  float tot = 0.f;
  for (int r = 0; r < reps; r++) {
    int rc = policy_launch(p, 0, false, false);
    if (rc) return rc;
    PCHK(hipStreamSynchronize(p->st));
    float t = 0.f;
    PCHK(hipEventElapsedTime(&t, p->ev0, p->ev1));
    tot += t;
  }
  p->last_ms = tot / reps;
  if (ms) *ms = p->last_ms;
  return 0;
}

extern "C" int kx_policy_last_ms(const kx_policy *p, float *ms) {
  if (!p || !ms) return pfail(KX_EINVAL, "null argument");
  *ms = p->last_ms;
  return 0;
}

extern "C" int kx_es_create(int32_t n_weights, const kx_es_opts *opts, const float *theta, kx_es **out) {
  if (!out || !opts) return pfail(KX_EINVAL, "null argument");
  *out = nullptr;
  if (n_weights <= 0) return pfail(KX_EINVAL, "ES over %d weights", n_weights);
  kx_es_opts o = *opts;
  if (o.pop_size <= 0 || o.pop_size > 8192)
    return pfail(KX_EINVAL, "ES population %d (1 .. 8192: the update keeps up to 4096 pair advantages)", o.pop_size);
  if (o.pop_size % 2) o.pop_size++;   // es.js:37 keeps the population even
  kx_es *e = new kx_es();
  e->nw = n_weights;
  e->o = o;
  e->sigma = o.sigma;
  auto cleanup = [&](int rc) {
    kx_es_destroy(e);
    return rc;
  };
  const size_t wb = (size_t)n_weights * sizeof(float);
  if (hipMalloc((void **)&e->theta, wb) != hipSuccess || hipMalloc((void **)&e->m, wb) != hipSuccess ||
      hipMalloc((void **)&e->v, wb) != hipSuccess || hipMalloc((void **)&e->best, wb) != hipSuccess ||
      hipMalloc((void **)&e->eps, (size_t)(o.pop_size / 2) * wb) != hipSuccess ||
      hipMalloc((void **)&e->diffs, (size_t)(o.pop_size / 2) * sizeof(double)) != hipSuccess)
    return cleanup(pfail(KX_ENOMEM, "ES state of %d weights", n_weights));
  if (hipStreamCreate(&e->st) != hipSuccess) return cleanup(pfail(KX_EDEVICE, "ES stream"));
  bool ok = hipMemsetAsync(e->m, 0, wb, e->st) == hipSuccess && hipMemsetAsync(e->v, 0, wb, e->st) == hipSuccess;
  if (ok && theta) ok = hipMemcpyAsync(e->theta, theta, wb, hipMemcpyHostToDevice, e->st) == hipSuccess;
  else if (ok) ok = hipMemsetAsync(e->theta, 0, wb, e->st) == hipSuccess;
  if (!ok || hipStreamSynchronize(e->st) != hipSuccess) return cleanup(pfail(KX_EDEVICE, "ES init"));
  *out = e;
  return 0;
}

extern "C" int kx_es_destroy(kx_es *e) {
  if (!e) return 0;
  if (e->st) (void)hipStreamSynchronize(e->st);
  (void)hipFree(e->theta);
  (void)hipFree(e->m);
  (void)hipFree(e->v);
  (void)hipFree(e->eps);
  (void)hipFree(e->best);
  (void)hipFree(e->diffs);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return 0;
}

extern "C" int kx_es_generation(kx_es *e, kx_policy *p, uint32_t *rnd_state) {
  if (!e || !p || !rnd_state) return pfail(KX_EINVAL, "null argument");
  if (p->d.nw != e->nw || p->n != e->o.pop_size)
    return pfail(KX_EINVAL, "policy of %d genomes x %d weights for an ES of %d x %d", p->n, p->d.nw, e->o.pop_size,
                 e->nw);
  const int pairs = e->o.pop_size / 2;
  const int64_t tot = (int64_t)pairs * e->nw;
  // the policy's stream waits for nothing else of the ES; order the two by
  // running the generation on the policy's stream after the ES's work
  PCHK(hipStreamSynchronize(e->st));
  hipLaunchKernelGGL(kx_k_es_genomes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, p->st, e->nw, pairs,
                     *rnd_state, e->sigma, e->theta, e->eps, p->genomes);
  PCHK(hipGetLastError());
  PCHK(hipStreamSynchronize(p->st));
  *rnd_state = (uint32_t)(*rnd_state + (uint32_t)((uint64_t)tot * 4) * 0x6D2B79F5u);
  e->have_eps = 1;
  return 0;
}

extern "C" int kx_es_update(kx_es *e, const double *fitness, double *stats4) {
  if (!e || !fitness) return pfail(KX_EINVAL, "null argument");
  if (!e->have_eps) return pfail(KX_EINVAL, "ES update before any generation (kx_es_generation)");
  const kx_es_opts &o = e->o;
  const int pop = o.pop_size, half = pop >> 1;
  // advantages (es.js:96-106), in float64 on the host as the reference does
  double diffs[4096];
  if (half > 4096) return pfail(KX_EINVAL, "ES population %d exceeds 8192", pop);
  double mean = 0;
  for (int p = 0; p < half; p++) {
    diffs[p] = (fitness[p * 2] - fitness[p * 2 + 1]) / 2;
    mean += fitness[p * 2] + fitness[p * 2 + 1];
  }
  mean /= pop;
  double sd = 0;
  for (int p = 0; p < half; p++) sd += diffs[p] * diffs[p];
  sd = std::sqrt(sd / (half > 1 ? half : 1));
  if (sd == 0 || std::isnan(sd)) sd = 1;   // `|| 1`
  KxEsStep q;
  q.inv = 1 / (half * e->sigma * sd);
  q.clip = o.clip;
  e->t++;
  q.b1 = o.beta1;
  q.b2 = o.beta2;
  q.omb1 = 1 - o.beta1;
  q.omb2 = 1 - o.beta2;
  q.c1 = 1 - std::pow(o.beta1, (double)e->t);
  q.c2 = 1 - std::pow(o.beta2, (double)e->t);
  q.lr = o.lr;
  q.eps = o.adam_eps;
  PCHK(hipMemcpyAsync(e->diffs, diffs, (size_t)half * sizeof(double), hipMemcpyHostToDevice, e->st));
  hipLaunchKernelGGL(kx_k_es_update, dim3((e->nw + 63) / 64), dim3(64), (size_t)half * sizeof(double), e->st, e->nw,
                     half, q, e->diffs, e->eps, e->theta, e->m, e->v);
  PCHK(hipGetLastError());
  e->sigma = std::max(o.min_sigma, e->sigma * o.sigma_decay);
  double best = -INFINITY;
  for (int i = 0; i < pop; i++)
    if (fitness[i] > best) best = fitness[i];
  if (best > e->best_ever) {
    e->best_ever = best;
    PCHK(hipMemcpyAsync(e->best, e->theta, (size_t)e->nw * sizeof(float), hipMemcpyDeviceToDevice, e->st));
    e->have_best = 1;
  }
  PCHK(hipStreamSynchronize(e->st));
  if (stats4) {
    stats4[0] = best;
    stats4[1] = mean;
    stats4[2] = sd * 2;
    stats4[3] = e->sigma;
  }
  return 0;
}

extern "C" int kx_es_get(const kx_es *e, float *theta, float *m, float *v, float *best_theta, double *scalars3) {
  if (!e) return pfail(KX_EINVAL, "null argument");
  const size_t wb = (size_t)e->nw * sizeof(float);
  if (theta) PCHK(hipMemcpyAsync(theta, e->theta, wb, hipMemcpyDeviceToHost, e->st));
  if (m) PCHK(hipMemcpyAsync(m, e->m, wb, hipMemcpyDeviceToHost, e->st));
  if (v) PCHK(hipMemcpyAsync(v, e->v, wb, hipMemcpyDeviceToHost, e->st));
  if (best_theta) {
    if (e->have_best) PCHK(hipMemcpyAsync(best_theta, e->best, wb, hipMemcpyDeviceToHost, e->st));
    else std::memset(best_theta, 0, wb);
  }
  PCHK(hipStreamSynchronize(e->st));
  if (scalars3) {
    scalars3[0] = e->sigma;
    scalars3[1] = e->t;
    scalars3[2] = e->best_ever;
  }
  return 0;
}
