// Per-site device helpers shared by the hop and the site-local kernels:
// packed clover blocks and the spinor loads/stores of the AoSoA layout with
// nrhs right-hand sides innermost.
#pragma once
#include "wc_common.h"

namespace wck {

// packed Hermitian 6x6: slots 0..2 hold the real diagonal in pairs, slots
// 3..17 the strictly-lower entries in row order (1,0),(2,0),(2,1),(3,0),...
__host__ __device__ constexpr int tri(int i, int j) { return 3 + i * (i - 1) / 2 + j; }

// one chiral block (18 slots) of the packed clover of a site
template <typename C> struct CloverBlock {
  C d[3];
  C L[15];
};

// block b of a site whose 36 slots lie at p[k * stride]
template <typename C>
__device__ __forceinline__ void clov_load_at(const C *p, long stride, int b, CloverBlock<C> &B) {
  const C *Ab = p + (long)(b * 18) * stride;
#pragma unroll
  for (int k = 0; k < 3; ++k) B.d[k] = Ab[(long)k * stride];
#pragma unroll
  for (int k = 0; k < 15; ++k) B.L[k] = Ab[(long)(3 + k) * stride];
}
template <typename C>
__device__ __forceinline__ void clov_load(const C *A, long h, int b, CloverBlock<C> &B) {
  clov_load_at(A + aidx(0, h, 36), (long)kBlk, b, B);
}

// o = B v on the 6 entries of one chiral block
template <typename R, typename C>
__device__ __forceinline__ void clov_mul6(const CloverBlock<C> &B, const C *v, C *o) {
  const R d[6] = {B.d[0].x, B.d[0].y, B.d[1].x, B.d[1].y, B.d[2].x, B.d[2].y};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    C acc = cscale(v[i], d[i]);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (j < i) acc = cmadd(acc, B.L[tri(i, j) - 3], v[j]);
      if (j > i) acc = cmaddc(acc, B.L[tri(j, i) - 3], v[j]);
    }
    o[i] = acc;
  }
}

// o = A v for one spinor of a site whose 36 slots lie at p[k * stride]
template <typename R, typename C>
__device__ __forceinline__ void clov_apply_at(const C *p, long stride, const C v[12], C o[12]) {
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    CloverBlock<C> B;
    clov_load_at(p, stride, b, B);
    clov_mul6<R>(B, v + 6 * b, o + 6 * b);
  }
}
// o = A v for one spinor of site h
template <typename R, typename C>
__device__ __forceinline__ void clov_apply(const C *A, long h, const C v[12], C o[12]) {
  clov_apply_at<R>(A + aidx(0, h, 36), (long)kBlk, v, o);
}

// v[j] = A v[j] for RB spinors of a site whose slots lie at p[k * stride]
// (each block loaded once)
template <typename R, int RB, typename C>
__device__ __forceinline__ void clov_apply_inplace_at(const C *p, long stride, C v[RB][12]) {
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    CloverBlock<C> B;
    clov_load_at(p, stride, b, B);
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      C o[6];
      clov_mul6<R>(B, &v[j][6 * b], o);
#pragma unroll
      for (int k = 0; k < 6; ++k) v[j][6 * b + k] = o[k];
    }
  }
}
// the same for site h of packed blocks A
template <typename R, int RB, typename C>
__device__ __forceinline__ void clov_apply_inplace(const C *A, long h, C v[RB][12]) {
  clov_apply_inplace_at<R, RB>(A + aidx(0, h, 36), (long)kBlk, v);
}

// spinor of site h, right-hand side r of N: entry k at aidx(k, h, 12) * N + r,
// stored as S (double2, float2 or the 16-bit c2h) and computed in C
template <typename S, typename C>
__device__ __forceinline__ void load12(const S *f, long h, int r, int N, C v[12]) {
  const S *p = f + aidx_site(h, 12) * N + r;
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = widen(p[(long)k * kBlk * N]);
}
template <typename S, typename C>
__device__ __forceinline__ void store12(S *f, long h, int r, int N, const C v[12]) {
  S *p = f + aidx_site(h, 12) * N + r;
#pragma unroll
  for (int k = 0; k < 12; ++k) p[(long)k * kBlk * N] = narrow<S>(v[k]);
}
// output of a hop: written once, not re-read by the kernel.  big: the field
// exceeds one descriptor's 4 GB; the descriptor is then based at the 64-site
// block of the wave's first active lane (every lane of a wave stores into
// that block)
template <typename S, typename C>
__device__ __forceinline__ void store12_stream(S *f, long Vh, long h, int r, int N, const C v[12],
                                               bool big = false) {
#if WC_NT_STORE
  wc_rsrc rs;
  uint32_t vo;
  if (big) {
    const long blk = (long)__builtin_amdgcn_readfirstlane((int)(h >> 6));
    const long base = blk * (12 * kBlk) * (long)N;
    rs = make_rsrc(f + base, (uint32_t)(12u * kBlk * (uint32_t)N * (uint32_t)sizeof(S)));
    vo = (uint32_t)(aidx_site(h, 12) * N + r - base) * (uint32_t)sizeof(S);
  } else {
    rs = make_rsrc(f, (uint32_t)(12u * (uint32_t)Vh * (uint32_t)N * (uint32_t)sizeof(S)));
    vo = (uint32_t)(aidx_site(h, 12) * N + r) * (uint32_t)sizeof(S);
  }
#pragma unroll
  for (int k = 0; k < 12; ++k)
    bstore_t<S, 2>::go(narrow<S>(v[k]), rs, vo, (uint32_t)(k * kBlk * N * sizeof(S)));
#else
  store12(f, h, r, N, v);
#endif
}

}  // namespace wck
