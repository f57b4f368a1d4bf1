// Halo of a T-slab decomposition: one GPU's share of the EO hop when the
// lattice is cut in t (C ABI in include/gpt_amd_wc.h: wc_halo_pack,
// wc_halo_unpack).
//
// A rank owning the time slices [t0, t0 + Lt) computes the hop of its output
// sites from spinors it owns, except for the t-direction terms of its two
// boundary slices:
//   - an output site x at local t = Lt - 1 needs (1 - s gamma_t) U_t(x)
//     psi(x + t), with psi(x + t) on the t = 0 slice of the rank above;
//   - an output site x at local t = 0 needs (1 + s gamma_t) U_t(x - t)^dag
//     psi(x - t), with psi and U_t(x - t) on the t = Lt - 1 slice of the rank
//     below.
// Both are sent as half spinors (the projection keeps 2 of the 4 spins, the
// lower two are rebuilt by the receiver): pack writes, for the input parity,
//   lo: (1 - s gamma_t) psi(y) of the t = 0 slice, for the rank below;
//   hi: U_t(y)^dag (1 + s gamma_t) psi(y) of the t = Lt - 1 slice, for the rank
//       above (the sender owns that link; the hop coefficient -1/2 and the
//       boundary phase of a wrapping link are folded in as the hop folds them),
// and unpack adds the rebuilt t-terms of the received faces to the boundary
// slices of an output field.  wc_halo_interior is the rest of the rank's hop:
// every term except those two.  On one GPU with periodic t the rank above and
// below is the GPU itself: unpack(lo, hi) of its own pack completes a hop
// whose wrapping t-links were left out.  Grid exchanges these faces inside
// its stencil (lib/cgpt/lib/operators/wilson_clover.h:45 turns on its
// communication / computation overlap).
//
// Face layout: entry k = 3 * spin + colour (spin 0, 1 of the half spinor) of
// face site f (the spatial half-site index, 0 <= f < Lx Ly Lz / 2) and
// right-hand side r at (k * nf + f) * N + r: consecutive lanes read
// consecutive (f, r) pairs.
#include <hip/hip_runtime.h>

#include "wc_hop.h"
#include "wc_internal.h"

namespace {

using namespace wck;

constexpr int kHaloThreads = 64;

// S of the t-direction term in the hop (compute_dir): forward (FB = 0) and
// backward (FB = 1) for the hop (DAG = 0) and its adjoint (DAG = 1)
template <int FB, int DAG> constexpr int halo_sign() { return ((FB == 0) != (DAG != 0)) ? 1 : -1; }

template <typename R, int NL, int DAG>
__global__ __launch_bounds__(kHaloThreads) void k_halo_pack(const HopLaunch a, long nf, typename c2<R>::t *lo,
                                                            typename c2<R>::t *hi) {
  using C = typename c2<R>::t;
  const Geom &g = a.g;
  const int N = a.nrhs;
  const long items = nf * N;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * items) return;
  const int face = q >= items ? 1 : 0;
  if (face) q -= items;
  const long f = q / N;
  const int r = (int)(q - f * N);
  const long hy = (face ? (long)(g.L[3] - 1) * nf : 0) + f;  // input-parity site y
  C psi[12];
  load12((const C *)a.in, hy, r, N, psi);
  C hs[2][3];
  if (face == 0) {
#pragma unroll
    for (int col = 0; col < 3; ++col)
      project<3, halo_sign<0, DAG>()>(psi[col], psi[3 + col], psi[6 + col], psi[9 + col], hs[0][col], hs[1][col]);
  } else {
#pragma unroll
    for (int col = 0; col < 3; ++col)
      project<3, halo_sign<1, DAG>()>(psi[col], psi[3 + col], psi[6 + col], psi[9 + col], hs[0][col], hs[1][col]);
    // U_t(y)^dag is direction 7 of the output site x = y + t, the t = 0
    // site with y's spatial coordinates (half-site f of parity pout)
    int c[4];
    half_coords(g, f, a.pout, c);
    C u[9];
#pragma unroll
    for (int e = 0; e < NL; ++e) u[e] = ((const C *)a.W)[aidx(7 * NL + e, f, 8 * NL)];
    prepare_link<R, NL>(a, 3, 1, c, u);
    C ch[2][3];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        C t = cmul(u[3 * i + 0], hs[s][0]);
        t = cmadd(t, u[3 * i + 1], hs[s][1]);
        ch[s][i] = cmadd(t, u[3 * i + 2], hs[s][2]);
      }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 3; ++i) hs[s][i] = ch[s][i];
  }
  C *dst = face ? hi : lo;
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) dst[((long)(3 * s + i) * nf + f) * N + r] = hs[s][i];
}

// dst (parity pout) += the t-terms of its boundary slices: lo (the rank
// above's t = 0 face) on t = Lt - 1, multiplied by U_t(x); hi (the rank
// below's t = Lt - 1 face, already multiplied) on t = 0
template <typename R, int NL, int DAG>
__global__ __launch_bounds__(kHaloThreads) void k_halo_unpack(const HopLaunch a, long nf,
                                                              const typename c2<R>::t *lo,
                                                              const typename c2<R>::t *hi) {
  using C = typename c2<R>::t;
  const Geom &g = a.g;
  const int N = a.nrhs;
  const long items = nf * N;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * items) return;
  const int face = q >= items ? 1 : 0;  // 0: output slice Lt - 1 from lo; 1: slice 0 from hi
  if (face) q -= items;
  const long f = q / N;
  const int r = (int)(q - f * N);
  const long hx = (face ? 0 : (long)(g.L[3] - 1) * nf) + f;
  const C *src = face ? hi : lo;
  C hs[2][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) hs[s][i] = src[((long)(3 * s + i) * nf + f) * N + r];
  C ch[2][3];
  if (face == 0) {
    int c[4];
    half_coords(g, hx, a.pout, c);
    C u[9];
#pragma unroll
    for (int e = 0; e < NL; ++e) u[e] = ((const C *)a.W)[aidx(6 * NL + e, hx, 8 * NL)];
    prepare_link<R, NL>(a, 3, 0, c, u);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        C t = cmul(u[3 * i + 0], hs[s][0]);
        t = cmadd(t, u[3 * i + 1], hs[s][1]);
        ch[s][i] = cmadd(t, u[3 * i + 2], hs[s][2]);
      }
  } else {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 3; ++i) ch[s][i] = hs[s][i];
  }
  C acc[12];
  load12((const C *)a.out, hx, r, N, acc);
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    acc[col] = cadd(acc[col], ch[0][col]);
    acc[3 + col] = cadd(acc[3 + col], ch[1][col]);
    if (face == 0)
      accum_lower<3, halo_sign<0, DAG>()>(acc[6 + col], acc[9 + col], ch[0][col], ch[1][col]);
    else
      accum_lower<3, halo_sign<1, DAG>()>(acc[6 + col], acc[9 + col], ch[0][col], ch[1][col]);
  }
  store12((C *)a.out, hx, r, N, acc);
}

template <typename R>
int halo_impl(const wc_operator *op, int unpack, int pout, int dag, int nrhs, const void *fld, void *lo, void *hi,
              hipStream_t s) {
  using C = typename c2<R>::t;
  const long nf = op->g.Vh / op->g.L[3];
  HopLaunch a = hop_args<R>(op, nrhs, pout, dag, 0, unpack ? nullptr : fld, unpack ? (void *)fld : nullptr);
  const long items = 2 * nf * nrhs;
  const unsigned grid = (unsigned)((items + kHaloThreads - 1) / kHaloThreads);
#define WC_HALO_LAUNCH(NL_, DAG_)                                                                             \
  do {                                                                                                      \
    if (unpack)                                                                                             \
      k_halo_unpack<R, NL_, DAG_><<<grid, kHaloThreads, 0, s>>>(a, nf, (const C *)lo, (const C *)hi);       \
    else                                                                                                    \
      k_halo_pack<R, NL_, DAG_><<<grid, kHaloThreads, 0, s>>>(a, nf, (C *)lo, (C *)hi);                     \
  } while (0)
  if (op->nl == 9) {
    if (dag) WC_HALO_LAUNCH(9, 1); else WC_HALO_LAUNCH(9, 0);
  } else {
    if (dag) WC_HALO_LAUNCH(6, 1); else WC_HALO_LAUNCH(6, 0);
  }
#undef WC_HALO_LAUNCH
  WC_LAUNCHED();
  return WC_OK;
}

int halo_call(wc_operator *op, int unpack, int pout, int dag, int nrhs, const void *fld, void *lo, void *hi,
              void *stream) {
  const char *fn = unpack ? "wc_halo_unpack" : "wc_halo_pack";
  if (!op || !fld || !lo || !hi) WC_FAIL(WC_ERR_ARG, "%s: null argument", fn);
  if (lo == hi || fld == lo || fld == hi) WC_FAIL(WC_ERR_ARG, "%s: the field and the two faces must not alias", fn);
  if (pout != 0 && pout != 1) WC_FAIL(WC_ERR_ARG, "%s: parity must be 0 or 1", fn);
  if (dag != 0 && dag != 1) WC_FAIL(WC_ERR_ARG, "%s: dag must be 0 or 1", fn);
  if (op->open_bc) WC_FAIL(WC_ERR_ARG, "%s: open boundaries in t leave no halo in t", fn);
  int rc = check_nrhs(op, nrhs, fn);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  return op->prec == WC_DOUBLE ? halo_impl<double>(op, unpack, pout, dag, nrhs, fld, lo, hi, s)
                               : halo_impl<float>(op, unpack, pout, dag, nrhs, fld, lo, hi, s);
}

// dst = the hop of src without the wrapping t-terms (k_hop with
// skip_t_faces), the part of a rank's hop that needs no halo
template <typename R>
int interior_impl(const wc_operator *op, int pout, int dag, int nrhs, void *dst, const void *src, hipStream_t s) {
  HopLaunch a = hop_args<R>(op, nrhs, pout, dag, 0, src, dst);
  a.skip_t_faces = 1;
  WC_HOP(R, a, s);
  WC_LAUNCHED();
  return WC_OK;
}

}  // namespace

extern "C" {

int wc_halo_interior(wc_operator *op, int parity_out, int dag, int nrhs, void *dst, const void *src, void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_halo_interior: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_halo_interior: dst must not alias src");
  if (parity_out != 0 && parity_out != 1) WC_FAIL(WC_ERR_ARG, "wc_halo_interior: parity must be 0 or 1");
  if (dag != 0 && dag != 1) WC_FAIL(WC_ERR_ARG, "wc_halo_interior: dag must be 0 or 1");
  if (op->open_bc) WC_FAIL(WC_ERR_ARG, "wc_halo_interior: open boundaries in t leave no halo in t");
  int rc = check_nrhs(op, nrhs, "wc_halo_interior");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  return op->prec == WC_DOUBLE ? interior_impl<double>(op, parity_out, dag, nrhs, dst, src, s)
                               : interior_impl<float>(op, parity_out, dag, nrhs, dst, src, s);
}

int wc_halo_pack(wc_operator *op, int parity_out, int dag, int nrhs, const void *src, void *face_lo, void *face_hi,
                 void *stream) {
  return halo_call(op, 0, parity_out, dag, nrhs, src, face_lo, face_hi, stream);
}

int wc_halo_unpack(wc_operator *op, int parity_out, int dag, int nrhs, const void *face_lo, const void *face_hi,
                   void *dst, void *stream) {
  return halo_call(op, 1, parity_out, dag, nrhs, dst, (void *)face_lo, (void *)face_hi, stream);
}

}  // extern "C"
