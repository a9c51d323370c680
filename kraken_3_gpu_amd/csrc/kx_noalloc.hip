// Allocation-free rollout: kx_rollout with its per-step event words staged in
// a file-scope device array instead of the handle's grow-on-demand events
// buffer, for callers that cannot allocate device memory (code under graph
// capture, tight serving loops, harnesses that forbid it).
//
// The array is one per GPU, held until the process ends. Its address on a
// device is resolved once, with that device current, through
// hipGetSymbolAddress; host code uses only that address (the array's name in
// host code is a host-side stand-in, not a device address). Every call on a
// GPU shares that GPU's array and leaves its last contents there: the words a
// call uses are cleared by a fill kernel first, and calls that overlap on one
// GPU race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kx_internal.h"

// 4 Mi event words (16 MiB): e.g. 4096 creatures x 1024 steps, or 65536 x 64
#define KX_NOALLOC_EVENT_WORDS (1u << 22)
#define KX_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) uint32_t scratch[KX_NOALLOC_EVENT_WORDS];

static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[KX_NOALLOC_MAX_DEVICES];

__global__ void kx_k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// The device address of `scratch` on device `dev` (made current by the caller).
static int scratch_on(int dev, uint32_t **out) {
  if (dev < 0 || dev >= KX_NOALLOC_MAX_DEVICES) return kx_error(KX_ENOTSUP, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return kx_error(KX_EDEVICE, "hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = (uint32_t *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" int kx_rollout_noalloc(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps,
                                  uint32_t *events_host) {
  int dev = 0, ncr = 0;
  hipStream_t st = nullptr;
  int r = kx_handle_basics(h, &dev, &ncr, &st);
  if (r) return r;
  if (!law) return kx_error(KX_EINVAL, "null argument");
  if (nsteps < 0 || step0 < 0) return kx_error(KX_EINVAL, "bad step range");
  if (!events_host || nsteps == 0) return kx_rollout_into(h, law, step0, nsteps, nullptr, events_host);
  const size_t need = (size_t)nsteps * (size_t)ncr;
  if (need > KX_NOALLOC_EVENT_WORDS)
    return kx_error(KX_EINVAL, "kx_rollout_noalloc: nsteps x creatures exceeds the 4194304 event words of its "
                               "static scratch; split the rollout");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return kx_error(KX_EDEVICE, "hipGetDevice failed");
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return kx_error(KX_EDEVICE, "hipSetDevice failed");
  uint32_t *ev = nullptr;
  r = scratch_on(dev, &ev);
  if (!r) {
    const uint32_t n = (uint32_t)need;
    const uint32_t blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
    hipLaunchKernelGGL(kx_k_fill_u32, dim3(blocks), dim3(256), 0, st, ev, n, 0u);
    r = hipGetLastError() == hipSuccess ? 0 : kx_error(KX_EDEVICE, "scratch clear launch failed");
  }
  if (!r) r = kx_rollout_into(h, law, step0, nsteps, ev, events_host);
  if (cur != dev) (void)hipSetDevice(cur);
  return r;
}
