// HBM bandwidth probe (C ABI in include/gpt_amd_wc.h: wc_bandwidth_probe):
// the rate a plain streaming kernel reaches on this device, the practical
// ceiling against which the hop's bandwidth is read.  Not on the operator's
// path; no counterpart in the reference.
//
// kind 0 copies `bytes` from src to dst; kind 1 only reads src and folds it
// with XOR into one 32-bit word per workgroup (the first kProbeBlocks words
// of dst), so that the loads cannot be dropped.  dst holds `bytes` bytes for
// either kind.  16 bytes per lane per
// access, 4 accesses in flight per lane, non-temporal (each byte is touched
// once), a fixed grid of kProbeBlocks workgroups striding over the buffer.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "wc_internal.h"

namespace {

// a native 16-byte vector (the nontemporal builtins take no struct types)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kProbeThreads = 256;
constexpr int kProbeBlocks = 8192;  // 32 per CU on 256 CUs (the fastest of 1024..16384 on MI355X)

__global__ __launch_bounds__(kProbeThreads) void k_probe_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                              long n) {
  const long stride = (long)gridDim.x * kProbeThreads;
  long i = (long)blockIdx.x * kProbeThreads + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const u32x4 a = __builtin_nontemporal_load(src + i);
    const u32x4 b = __builtin_nontemporal_load(src + i + stride);
    const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
    const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + stride);
    __builtin_nontemporal_store(c, dst + i + 2 * stride);
    __builtin_nontemporal_store(d, dst + i + 3 * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_read(const u32x4 *__restrict__ src, long n,
                                                              uint32_t *__restrict__ out) {
  const long stride = (long)gridDim.x * kProbeThreads;
  long i = (long)blockIdx.x * kProbeThreads + threadIdx.x;
  uint32_t x = 0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const u32x4 a = __builtin_nontemporal_load(src + i);
    const u32x4 b = __builtin_nontemporal_load(src + i + stride);
    const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
    const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
    x ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
  }
  for (; i < n; i += stride) {
    const u32x4 a = __builtin_nontemporal_load(src + i);
    x ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  __shared__ uint32_t part[kProbeThreads];
  part[threadIdx.x] = x;
  __syncthreads();
  for (int s = kProbeThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] ^= part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

}  // namespace

extern "C" int wc_bandwidth_probe(int kind, const void *src, void *dst, long bytes, void *stream) {
  if (!src || !dst) WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: null argument");
  if (kind != 0 && kind != 1) WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: kind must be 0 (copy) or 1 (read)");
  if (bytes <= 0 || bytes % 16) WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: bytes must be a positive multiple of 16");
  if (((uintptr_t)src | (uintptr_t)dst) % 16) WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: buffers must be 16-byte aligned");
  if (src == dst) WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: src and dst must differ");
  if (kind == 1 && bytes < (long)sizeof(uint32_t) * kProbeBlocks)
    WC_FAIL(WC_ERR_ARG, "wc_bandwidth_probe: a read probe needs at least %d bytes", (int)sizeof(uint32_t) * kProbeBlocks);
  hipStream_t s = (hipStream_t)stream;
  const long n = bytes / 16;
  if (kind == 0)
    k_probe_copy<<<kProbeBlocks, kProbeThreads, 0, s>>>((const u32x4 *)src, (u32x4 *)dst, n);
  else
    k_probe_read<<<kProbeBlocks, kProbeThreads, 0, s>>>((const u32x4 *)src, n, (uint32_t *)dst);
  WC_LAUNCHED();
  return WC_OK;
}

