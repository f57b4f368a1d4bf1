// An entry point that allocates nothing: the Schur pieces of
// wc_apply_mpc_block with their scratch fields in a fixed device array.
//
// wc_apply_mpc_block keeps its two or three scratch half fields in the
// operator and allocates them (hipMalloc) on first use and when the number of
// right-hand sides grows.  Code under HIP graph capture, serving loops that
// must not allocate and harnesses that forbid it call
// wc_apply_mpc_block_noalloc instead: the same kernels, with the scratch in
// `scratch` below (kNoAllocScratchBytes per GPU, held from the library's
// first use on that GPU until the process ends).  Every call on one GPU
// shares that GPU's array and leaves its contents behind; nothing in it has
// to start at zero (each scratch field is written before it is read), and
// calls that overlap on one GPU (different streams or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wc_internal.h"

// 64 MiB per GPU: two scratch half fields of Vh * nrhs * 12 complex values
// (FP64: Vh * nrhs <= 174762, e.g. 16^3 x 32 with 2 right-hand sides; FP32:
// Vh * nrhs <= 349525, e.g. 24^3 x 48 with one)
static constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

// The array's device address on each GPU, taken once per device with that
// device current.  In host code `scratch` names a host-side stand-in, never
// the device memory: only these pointers go to kernels.
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(void **out) {
  int dev = 0;
  WC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    WC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) WC_FAIL(WC_ERR_HIP, "wc_apply_mpc_block_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return WC_OK;
}

extern "C" int wc_apply_mpc_block_noalloc(wc_operator *op, int parity, int which, int nrhs, void *dst,
                                          const void *src, void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_noalloc: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_noalloc: dst must not alias src");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_noalloc: parity must be 0 or 1");
  if (which < 0 || which > 2) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_noalloc: which must be 0, 1 or 2");
  int rc = check_nrhs(op, nrhs, "wc_apply_mpc_block_noalloc");
  if (rc) return rc;
  // Mpc and Mpc^dag Mpc use two fields (A_D^-1 src, then the Schur
  // intermediate of Mpc^dag Mpc in the same place), Mpc^dag one
  const size_t field = half_bytes(op) * (size_t)nrhs;
  const size_t need = (which == 1 ? 1 : 2) * field;
  if (need > kNoAllocScratchBytes)
    WC_FAIL(WC_ERR_ARG,
            "wc_apply_mpc_block_noalloc: %zu bytes of scratch needed (half volume %ld, %d right-hand sides, "
            "%s), more than the %zu bytes the library reserves per GPU",
            need, op->g.Vh, nrhs, op->prec == WC_DOUBLE ? "FP64" : "FP32", kNoAllocScratchBytes);
  void *base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  char *b = (char *)base;
  void *tC = b, *tD = which == 1 ? nullptr : b + field;
  hipStream_t s = (hipStream_t)stream;
  return op->prec == WC_DOUBLE
             ? mpc_run<double>(op, nrhs, parity, which, dst, src, nullptr, s, tD, tC, tD, nullptr, nullptr, nullptr, 0)
             : mpc_run<float>(op, nrhs, parity, which, dst, src, nullptr, s, tD, tC, tD, nullptr, nullptr, nullptr, 0);
}
