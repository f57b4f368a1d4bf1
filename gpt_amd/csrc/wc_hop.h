// The even-odd hop (Meooe) kernel of the Wilson-clover operator for gfx950,
// with fused epilogues.  Instantiated once per precision (wc_hop_f64.hip,
// wc_hop_f32.hip).
//
// A work item is a site and RB of its right-hand sides.  For each of the 8
// directions an item gathers the
// neighbour spinors of its RB right-hand sides, projects each to a half
// spinor with (1 -+ gamma_mu), multiplies by the double-stored link (9
// complex, or 6 with the third row rebuilt), loaded and scaled once for the
// RB spinors, and accumulates the reconstructed 4-spinors.  Epilogues fuse
// the clover (inverse) block, the Schur complement's axpy and the CG inner
// product, so the normal operator needs no extra passes.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "wc_kernels.h"
#include "wc_site.h"

namespace wck {

// spin projection h = (1 - s gamma_mu) psi restricted to the upper spins, for
// gpt's chiral basis (gamma_mu = [[0, X_mu], [X_mu^dag, 0]]):
//   mu=0: h0 = p0 - s i p3, h1 = p1 - s i p2      mu=1: h0 = p0 + s p3, h1 = p1 - s p2
//   mu=2: h0 = p0 - s i p2, h1 = p1 + s i p3      mu=3: h0 = p0 - s p2, h1 = p1 - s p3
template <int MU, int S, typename C>
__device__ __forceinline__ void project(const C p0, const C p1, const C p2, const C p3, C &h0, C &h1) {
  if (MU == 0) {
    h0 = S > 0 ? csub(p0, cimul(p3)) : cadd(p0, cimul(p3));
    h1 = S > 0 ? csub(p1, cimul(p2)) : cadd(p1, cimul(p2));
  } else if (MU == 1) {
    h0 = S > 0 ? cadd(p0, p3) : csub(p0, p3);
    h1 = S > 0 ? csub(p1, p2) : cadd(p1, p2);
  } else if (MU == 2) {
    h0 = S > 0 ? csub(p0, cimul(p2)) : cadd(p0, cimul(p2));
    h1 = S > 0 ? cadd(p1, cimul(p3)) : csub(p1, cimul(p3));
  } else {
    h0 = S > 0 ? csub(p0, p2) : cadd(p0, p2);
    h1 = S > 0 ? csub(p1, p3) : cadd(p1, p3);
  }
}

// lower spins of (1 - s gamma_mu) psi from chi = U h: -s X_mu^dag chi
//   mu=0: l2 = s i c1, l3 = s i c0      mu=1: l2 = -s c1, l3 = s c0
//   mu=2: l2 = s i c0, l3 = -s i c1     mu=3: l2 = -s c0, l3 = -s c1
template <int MU, int S, typename C>
__device__ __forceinline__ void accum_lower(C &a2, C &a3, const C c0, const C c1) {
  if (MU == 0) {
    a2 = S > 0 ? cadd(a2, cimul(c1)) : csub(a2, cimul(c1));
    a3 = S > 0 ? cadd(a3, cimul(c0)) : csub(a3, cimul(c0));
  } else if (MU == 1) {
    a2 = S > 0 ? csub(a2, c1) : cadd(a2, c1);
    a3 = S > 0 ? cadd(a3, c0) : csub(a3, c0);
  } else if (MU == 2) {
    a2 = S > 0 ? cadd(a2, cimul(c0)) : csub(a2, cimul(c0));
    a3 = S > 0 ? csub(a3, cimul(c1)) : cadd(a3, cimul(c1));
  } else {
    a2 = S > 0 ? csub(a2, c0) : cadd(a2, c0);
    a3 = S > 0 ? csub(a3, c1) : cadd(a3, c1);
  }
}

// the link of direction (mu, fb) as the hop applies it: the third row
// rebuilt for 12-real storage, the hop coefficient (and the boundary phase on
// links that wrap around) folded in
template <typename R, int NL>
__device__ __forceinline__ void prepare_link(const HopLaunch &a, int mu, int fb, const int c[4],
                                             typename c2<R>::t u[9]) {
  using C = typename c2<R>::t;
  const Geom &g = a.g;
  const bool bnd = (fb == 0) ? (c[mu] == g.L[mu] - 1) : (c[mu] == 0);
  if (NL == 6) {
    // third row = conj(row0 x row1) for SU(3)
    u[6] = cconj(csub(cmul(u[1], u[5]), cmul(u[2], u[4])));
    u[7] = cconj(csub(cmul(u[2], u[3]), cmul(u[0], u[5])));
    u[8] = cconj(csub(cmul(u[0], u[4]), cmul(u[1], u[3])));
  }
  const R coef = (R)(mu == 3 ? a.ct : a.cs);
  if (bnd) {
    C f = cmk<C>((R)(a.ph[mu][0]), (R)(fb == 0 ? a.ph[mu][1] : -a.ph[mu][1]));
    f = cscale(f, coef);
#pragma unroll
    for (int e = 0; e < 9; ++e) u[e] = cmul(f, u[e]);
  } else {
#pragma unroll
    for (int e = 0; e < 9; ++e) u[e] = cscale(u[e], coef);
  }
}

// direction (MU, FB): the link is prepared (unless PRE: staged prepared),
// then every spinor is projected (in P: the working precision, or packed
// FP16 for c2h storage), widened, multiplied and accumulated
template <typename R, int NL, int DAG, int MU, int FB, int RB, bool PRE, typename P>
__device__ __forceinline__ void compute_dir(const HopLaunch &a, const int c[4], typename c2<R>::t u[9],
                                            const P psi[RB][12], typename c2<R>::t acc[RB][12]) {
  using C = typename c2<R>::t;
  constexpr int S = ((FB == 0) != (DAG != 0)) ? 1 : -1;
  if constexpr (!PRE) prepare_link<R, NL>(a, MU, FB, c, u);
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    C hs[2][3];
#pragma unroll
    for (int col = 0; col < 3; ++col) {
      P h0, h1;
      project<MU, S>(psi[j][0 + col], psi[j][3 + col], psi[j][6 + col], psi[j][9 + col], h0, h1);
      hs[0][col] = widen(h0);
      hs[1][col] = widen(h1);
    }
    C ch[2][3];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        C t = cmul(u[3 * i + 0], hs[s][0]);
        t = cmadd(t, u[3 * i + 1], hs[s][1]);
        t = cmadd(t, u[3 * i + 2], hs[s][2]);
        ch[s][i] = t;
      }
#pragma unroll
    for (int col = 0; col < 3; ++col) {
      acc[j][0 + col] = cadd(acc[j][0 + col], ch[0][col]);
      acc[j][3 + col] = cadd(acc[j][3 + col], ch[1][col]);
      accum_lower<MU, S>(acc[j][6 + col], acc[j][9 + col], ch[0][col], ch[1][col]);
    }
  }
}

// The 8 directions run as a real loop (not unrolled): the loads of one
// direction cannot be hoisted into another, so the live registers stay at one
// direction's operands plus the accumulators; the other waves on the SIMD
// keep HBM busy meanwhile.  The projection variant is picked by a
// wave-uniform switch.  Neighbour offsets come from uniform strides of the
// half-site index h = ((t*Lz + z)*Ly + y)*Lxh + x/2.  rl[j] is the
// right-hand side a spinor is read for (clamped into range for padded lanes).
//
// LS: the links are staged in LDS, prepared (prepare_link), entry e of
// direction dir of the site at lw[(dir * 9 + e) * ns].
//
// XS: xst (non-null) holds the input spinors of the half-sites xs0 ..
// xs0 + xsn - 1, component k of site xs0 + i at xst[k * xsl + i]; a
// neighbour in that range is read from there (one right-hand side).
template <typename R, int NL, int DAG, int RB, typename SI, bool LS, bool XS>
__device__ __forceinline__ void hop_site(const HopLaunch &a, long h, const int rl[RB], const int cx,
                                         const int cy, const int cz, const int ct,
                                         typename c2<R>::t acc[RB][12], const typename c2<R>::t *lw,
                                         int ns, const typename c2<R>::t *xst, long xs0, int xsn, int xsl) {
  // coordinates as scalars: a select among array elements would become a
  // dynamically indexed (scratch) array
  const int c[4] = {cx, cy, cz, ct};
  using C = typename c2<R>::t;
  // the projection's operand type: packed halves for FP16 storage
  using P = typename std::conditional<std::is_same<SI, c2h>::value, h2v, C>::type;
  const Geom &g = a.g;
  const uint32_t N = (uint32_t)a.nrhs;
  const uint32_t hh = (uint32_t)h;
  // a field beyond one descriptor's 4 GB (a.big): each direction's gathers
  // go through a descriptor based at the t-slice that holds the neighbours
  // (a wave's sites share their t: a 64-site block lies in one slice), so
  // the per-lane offsets stay within a slice
  const wc_rsrc in_all = make_rsrc(a.in, 12u * (uint32_t)g.Vh * N * (uint32_t)sizeof(SI));
  const uint32_t slice = (uint32_t)(g.Vh / g.L[3]);
  // a wave's sites are consecutive and lie in at most two aligned 64-site
  // blocks: its links are addressed from the first active lane's block
  // (unused, and removed by the compiler, when the links are staged)
  const long blk = (long)__builtin_amdgcn_readfirstlane((int)(h >> 6));
  const wc_rsrc rw = make_rsrc((const C *)a.W + blk * (8 * NL * kBlk), 2u * 8u * NL * kBlk * (uint32_t)sizeof(C));
  const uint32_t vw = (uint32_t)(aidx_site(h, 8 * NL) - blk * (8 * NL * kBlk)) * (uint32_t)sizeof(C);
  const uint32_t kstride = (uint32_t)kBlk * N * (uint32_t)sizeof(SI);
  const uint32_t xh = (uint32_t)cx >> 1;
  const uint32_t st1 = (uint32_t)g.Lxh, st2 = st1 * (uint32_t)g.L[1], st3 = st2 * (uint32_t)g.L[2];
#pragma unroll 1
  for (int dir = 0; dir < 8; ++dir) {
    const int mu = dir >> 1, fb = dir & 1;
    if (mu == 3 && a.skip_t_faces && (fb == 0 ? ct == g.L[3] - 1 : ct == 0)) continue;
    uint32_t nb;
    if (mu == 0) {
      const int xn = fb == 0 ? (cx == g.L[0] - 1 ? 0 : cx + 1) : (cx == 0 ? g.L[0] - 1 : cx - 1);
      nb = hh - xh + ((uint32_t)xn >> 1);
    } else {
      const uint32_t stride = mu == 1 ? st1 : (mu == 2 ? st2 : st3);
      const int cm = mu == 1 ? cy : (mu == 2 ? cz : ct);
      const int Lm = mu == 1 ? g.L[1] : (mu == 2 ? g.L[2] : g.L[3]);
      const uint32_t wrap = (uint32_t)(Lm - 1) * stride;
      if (fb == 0)
        nb = (cm == Lm - 1) ? hh - wrap : hh + stride;
      else
        nb = (cm == 0) ? hh + wrap : hh - stride;
    }
    C u[9];
    if constexpr (LS) {
#pragma unroll
      for (int e = 0; e < 9; ++e) u[e] = lw[(dir * 9 + e) * ns];
    } else {
      const uint32_t dsoff = (uint32_t)(dir * NL) * (uint32_t)(kBlk * sizeof(C));
#pragma unroll
      for (int e = 0; e < NL; ++e)
        u[e] = bload<C, WC_LINK_AUX>(rw, vw, dsoff + (uint32_t)(e * kBlk * sizeof(C)));
    }
    P psi[RB][12];
    bool staged = false;
    if constexpr (XS) {
      const uint32_t li = nb - (uint32_t)xs0;
      if (xst != nullptr && li < (uint32_t)xsn) {
#pragma unroll
        for (int k = 0; k < 12; ++k) psi[0][k] = xst[k * xsl + (int)li];
        staged = true;
      }
    }
    if (!staged) {
      wc_rsrc in = in_all;
      uint32_t vs;
      if (a.big) {
        const int tn = mu != 3 ? ct : (fb == 0 ? (ct == g.L[3] - 1 ? 0 : ct + 1) : (ct == 0 ? g.L[3] - 1 : ct - 1));
        const long s0 = (long)__builtin_amdgcn_readfirstlane(tn) * (long)slice;  // first half-site of the slice
        const long base = (s0 >> 6) * (12 * kBlk) * (long)N;                     // its element offset
        in = make_rsrc((const SI *)a.in + base, 12u * slice * N * (uint32_t)sizeof(SI));
        vs = (uint32_t)((long)aidx_site(nb, 12) * N - base);
      } else {
        vs = (uint32_t)aidx_site(nb, 12) * N;
      }
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const uint32_t vo = (vs + (uint32_t)rl[j]) * (uint32_t)sizeof(SI);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const SI v = bload<SI, WC_SPINOR_AUX>(in, vo, (uint32_t)k * kstride);
          if constexpr (std::is_same<SI, c2h>::value)
            psi[j][k] = __builtin_bit_cast(h2v, v);
          else
            psi[j][k] = widen(v);
        }
      }
    }
    switch (dir) {
      case 0: compute_dir<R, NL, DAG, 0, 0, RB, LS>(a, c, u, psi, acc); break;
      case 1: compute_dir<R, NL, DAG, 0, 1, RB, LS>(a, c, u, psi, acc); break;
      case 2: compute_dir<R, NL, DAG, 1, 0, RB, LS>(a, c, u, psi, acc); break;
      case 3: compute_dir<R, NL, DAG, 1, 1, RB, LS>(a, c, u, psi, acc); break;
      case 4: compute_dir<R, NL, DAG, 2, 0, RB, LS>(a, c, u, psi, acc); break;
      case 5: compute_dir<R, NL, DAG, 2, 1, RB, LS>(a, c, u, psi, acc); break;
      case 6: compute_dir<R, NL, DAG, 3, 0, RB, LS>(a, c, u, psi, acc); break;
      default: compute_dir<R, NL, DAG, 3, 1, RB, LS>(a, c, u, psi, acc); break;
    }
  }
}

// Visiting order of the 64-site blocks, seg > 1: the bps blocks of a time
// slice are cut into seg runs of consecutive blocks (whole z-y rows), and a
// run is visited for t = 0, 1, ... before the next run, so the +-t
// neighbours of a block were read a few blocks earlier and are still in L2.
// A bijection of [0, nb) for any 1 < seg <= bps (nb = bps * lt).
__device__ __forceinline__ long block_order(long T, int seg, int bps, int lt) {
  if (seg <= 1) return T;
  // run s holds slice blocks [lo(s), lo(s+1)), lo(s) = s * bps / seg
  const long u = T / lt;
  long s = u * seg / bps;
  while ((s + 1) * bps / seg <= u) ++s;
  const long lo = s * bps / seg, len = (s + 1) * bps / seg - lo;
  const long off = T - (long)lt * lo;
  const long t = off / len;
  return t * bps + lo + (off - t * len);
}

// ls == 2 (spinor-block staging) applies to hops of one right-hand side in
// natural block order whose input is in the working precision, without link
// staging; xs_possible is the compile-time part, xs_on the launch's
template <typename R, int RB, typename SI, bool LS> constexpr bool xs_possible() {
  return !LS && RB == 1 && std::is_same<SI, typename c2<R>::t>::value;
}
__host__ __device__ inline bool xs_on(const HopLaunch &a) {
  return a.shape.ls == 2 && a.nrhs == 1 && a.shape.seg <= 1;
}

// waves per SIMD the register allocation must allow: 4 for the FP32 hop
// with one right-hand side per thread (<= 128 VGPRs), else 2 (<= 256; keeps
// the compiler from parking accumulators in AGPRs at 1 wave)
template <typename R, int RB> constexpr int hop_min_waves() { return (sizeof(R) == 4 && RB == 1) ? 4 : 2; }

// Work items are (site h, right-hand-side group g) with g fastest, g covering
// the RB right-hand sides g*RB .. g*RB+RB-1, one item per thread over the
// whole lattice: consecutive lanes read consecutive right-hand sides of a
// site and then of the next site, so every gather is one contiguous run and
// the sites in flight stay few (their neighbours stay in L2).  EPI_DOT
// writes each (site, right-hand side) product to dots[h * N + r]; dot_blocks
// sums them per 64-site block in site order, so the sums do not depend on
// the launch shape.
template <typename R, int NL, int DAG, int EPI, int RB, typename SI, typename SO, bool LS>
__global__ __launch_bounds__(512, (hop_min_waves<R, RB>())) void k_hop(const HopLaunch a) {
  using C = typename c2<R>::t;
  if (a.halt && *a.halt) return;
  const Geom &g = a.g;
  const int N = a.nrhs;
  const int Gt = (N + RB - 1) / RB;
  const long q0 = xcd_tile(blockIdx.x, gridDim.x) * (long)blockDim.x;
  const long q = q0 + threadIdx.x;
  // items number Vh * Gt < 2^31 (a field of N right-hand sides fits 4 GB):
  // 32-bit unsigned divisions
  const long v = Gt == 1 ? q : (long)((uint32_t)q / (uint32_t)Gt);  // site in visiting order
  // a 64-site block has 64 * Gt items, a whole number of waves: a wave's
  // sites lie in one block, and blocks can be visited in any order
  auto site_of = [&](long vv) {
    return a.shape.seg > 1 ? block_order(vv >> 6, a.shape.seg, a.bps, g.L[3]) * kBlk + (vv & 63) : vv;
  };
  // ls: the workgroup's sites v0 .. v0 + ns - 1 get their links, prepared
  // (prepare_link: 9 entries per direction), and the EPI_CLOV blocks (36
  // entries) copied to LDS once, entry e of site v0 + s at stage[e * ns + s];
  // the Gt items of a site then read them from there instead of each loading
  // them through the texture path and preparing the links again
  extern __shared__ __align__(16) char wc_dyn_lds[];
  C *stage = (C *)wc_dyn_lds;
  const long v0 = Gt == 1 ? q0 : (long)((uint32_t)q0 / (uint32_t)Gt);
  int ns = 0;
  const bool clov_staged = LS && (EPI & EPI_CLOV) && a.clov && !a.clov_roff;
  if constexpr (LS) {
    long v1 = (long)((uint32_t)(q0 + (long)blockDim.x - 1) / (uint32_t)Gt);
    if (v1 > g.Vh - 1) v1 = g.Vh - 1;
    ns = v1 >= v0 ? (int)(v1 - v0 + 1) : 0;
    const int nw = 8 * NL;
    for (int i = threadIdx.x; i < 8 * ns; i += blockDim.x) {
      const int dir = (int)((uint32_t)i / (uint32_t)ns), sidx = i - dir * ns;
      const long hv = site_of(v0 + sidx);
      int cs[4];
      half_coords(g, hv, a.pout, cs);
      C u[9];
#pragma unroll
      for (int e = 0; e < NL; ++e) u[e] = ((const C *)a.W)[aidx(dir * NL + e, hv, nw)];
      prepare_link<R, NL>(a, dir >> 1, dir & 1, cs, u);
#pragma unroll
      for (int e = 0; e < 9; ++e) stage[(dir * 9 + e) * ns + sidx] = u[e];
    }
    if (clov_staged)
      for (int i = threadIdx.x; i < 36 * ns; i += blockDim.x) {
        const int e = (int)((uint32_t)i / (uint32_t)ns), sidx = i - e * ns;
        stage[72 * ns + i] = ((const C *)a.clov)[aidx(e, site_of(v0 + sidx), 36)];
      }
    __syncthreads();
  }
  // spinor-block staging (xs_on): the workgroup's output sites are the
  // half-sites q0 .. q0 + blockDim - 1 (one item per site); their input
  // spinors (the other parity's sites of the same half-site index) are
  // copied to LDS, component k of site q0 + i at stage[k * blockDim + i].
  // The +-x neighbours of a site lie in its row and most +-y neighbours in
  // the adjacent rows, inside that range, so those gathers read LDS and
  // only the rest go through the texture path.
  const C *xst = nullptr;
  int xsn = 0;
  if constexpr (xs_possible<R, RB, SI, LS>()) {
    if (xs_on(a)) {
      const long rem = g.Vh - q0;  // > 0 for every workgroup of the grid; clamped all the same
      xsn = rem <= 0 ? 0 : (rem < (long)blockDim.x ? (int)rem : (int)blockDim.x);
      if ((int)threadIdx.x < xsn) {
        const long hs = q0 + threadIdx.x;
#pragma unroll
        for (int k = 0; k < 12; ++k) stage[k * (int)blockDim.x + (int)threadIdx.x] = ((const C *)a.in)[aidx(k, hs, 12)];
      }
      __syncthreads();
      xst = stage;
    }
  }
  if (v >= g.Vh) return;
  const int rb = (int)(q - v * Gt) * RB;
  const long h = site_of(v);
  const int sv = (int)(v - v0);
  int rl[RB];
#pragma unroll
  for (int j = 0; j < RB; ++j) rl[j] = min(rb + j, N - 1);
  int c[4];
  half_coords(g, h, a.pout, c);
  C acc[RB][12];
#pragma unroll
  for (int j = 0; j < RB; ++j)
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[j][k] = cmk<C>(0, 0);
  hop_site<R, NL, DAG, RB, SI, LS, xs_possible<R, RB, SI, LS>()>(a, h, rl, c[0], c[1], c[2], c[3], acc,
                                                                   stage + sv, ns, xst, q0, xsn, (int)blockDim.x);
  // epilogue
  const bool zb = a.open_bc && (c[3] == 0 || c[3] == g.L[3] - 1);
  if (zb) {
#pragma unroll
    for (int j = 0; j < RB; ++j)
#pragma unroll
      for (int k = 0; k < 12; ++k) acc[j][k] = cmk<C>(0, 0);
  } else if (EPI & EPI_CLOV) {
    if (LS && clov_staged) {
      clov_apply_inplace_at<R, RB>(stage + 72 * ns + sv, (long)ns, acc);
    } else if (a.clov && a.clov_roff) {
      // one set of blocks per right-hand side (a batch of several masses)
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        C o[12];
        clov_apply<R>((const C *)a.clov + a.clov_roff[rl[j]], h, acc[j], o);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[j][k] = o[k];
      }
    } else if (a.clov) {
      clov_apply_inplace<R, RB>((const C *)a.clov, h, acc);
    } else {
#pragma unroll
      for (int j = 0; j < RB; ++j)
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[j][k] = cscale(acc[j][k], (R)a.clov_scalar);
    }
  }
  if (EPI & (EPI_DIAG | EPI_AXPY)) {
    C y[RB][12];
#pragma unroll
    for (int j = 0; j < RB; ++j) load12((const SI *)a.y, h, rl[j], N, y[j]);
    if (EPI & EPI_DIAG) {
      if (!zb) {
        if (a.diag_clov) {
          clov_apply_inplace<R, RB>((const C *)a.diag_clov, h, y);
        } else {
#pragma unroll
          for (int j = 0; j < RB; ++j)
#pragma unroll
            for (int k = 0; k < 12; ++k) y[j][k] = cscale(y[j][k], (R)a.diag);
        }
#pragma unroll
        for (int j = 0; j < RB; ++j)
#pragma unroll
          for (int k = 0; k < 12; ++k) acc[j][k] = cadd(acc[j][k], y[j][k]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < RB; ++j)
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[j][k] = csub(y[j][k], acc[j][k]);
    }
  }
  if (EPI & EPI_DOT) {
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      C v[12];
      load12((const SI *)a.dotv, h, rl[j], N, v);
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 12; ++k)
        s = fma((double)v[k].y, (double)acc[j][k].y, fma((double)v[k].x, (double)acc[j][k].x, s));
      if (rb + j < N) a.dots[h * N + rb + j] = s;
    }
  }
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if (rb + j < N) store12_stream((SO *)a.out, g.Vh, h, rb + j, N, acc[j], a.big != 0);
}

// epilogues instantiated per storage: all of them in the working precision;
// with 16-bit in/y/dotv those the normal operator of the 16-bit inner solve
// uses (CLOV, AXPY to 16-bit, CLOV|AXPY|DOT to FP32) and the bare hop
// (16-bit to 16-bit, wc_apply_block_s16)
template <typename R, typename SI, typename SO> constexpr bool epi_instantiated(int e) {
  using C = typename c2<R>::t;
  if (std::is_same<SI, C>::value && std::is_same<SO, C>::value) return true;
  if (is16<SI>() && std::is_same<SO, SI>::value) return e == 0 || e == EPI_CLOV || e == EPI_AXPY;
  if (is16<SI>() && std::is_same<SO, C>::value) return e == (EPI_CLOV | EPI_AXPY | EPI_DOT);
  return false;
}

template <typename R, int NL, int DAG, int RB, typename SI, typename SO>
static bool hop_epi(const HopLaunch &a, hipStream_t s) {
  using C = typename c2<R>::t;
  const long threads = kBlk * a.shape.tpb;
  const long G = (a.nrhs + RB - 1) / RB;
  const long items = a.g.Vh * G;
  const unsigned grid = (unsigned)((items + threads - 1) / threads);
  // LDS: the staging area (shape.ls), at least the share of a CU that caps
  // the workgroups per CU at shape.occ
  size_t lds = a.shape.occ > 0 ? (size_t)((kLdsPerCU / a.shape.occ) & ~1023) : 0;
  if (a.shape.ls == 1) {
    const size_t st = (size_t)(72 + ((a.epi & EPI_CLOV) && a.clov && !a.clov_roff ? 36 : 0)) *
                      (size_t)stage_sites(threads, G) * sizeof(C);
    if (st > (size_t)kMaxStageBytes) return false;
    lds = std::max(lds, st);
  } else if (xs_possible<R, RB, SI, false>() && xs_on(a)) {
    const size_t st = (size_t)12 * (size_t)threads * sizeof(C);
    if (st > (size_t)kMaxStageBytes) return false;
    lds = std::max(lds, st);
  }
  switch (a.epi) {
#define WC_EPI(E)                                                                        \
  case E:                                                                                \
    if constexpr (epi_instantiated<R, SI, SO>(E)) {                                     \
      if (a.shape.ls == 1)                                                               \
        k_hop<R, NL, DAG, E, RB, SI, SO, true><<<grid, (unsigned)threads, lds, s>>>(a);  \
      else                                                                               \
        k_hop<R, NL, DAG, E, RB, SI, SO, false><<<grid, (unsigned)threads, lds, s>>>(a); \
      return true;                                                                       \
    }                                                                                    \
    return false;
    WC_EPI(0)
    WC_EPI(EPI_CLOV)
    WC_EPI(EPI_DIAG)
    WC_EPI(EPI_AXPY)
    WC_EPI(EPI_CLOV | EPI_AXPY)
    WC_EPI(EPI_CLOV | EPI_AXPY | EPI_DOT)
    WC_EPI(EPI_AXPY | EPI_DOT)
#undef WC_EPI
    default: return false;
  }
}

template <typename R, int NL, int DAG, typename SI, typename SO>
static bool hop_rb(const HopLaunch &a, hipStream_t s) {
  switch (a.shape.rb) {
    case 1: return hop_epi<R, NL, DAG, 1, SI, SO>(a, s);
    case 2:
      if constexpr (max_rb<R>() >= 2) return hop_epi<R, NL, DAG, 2, SI, SO>(a, s);
      return false;
    default: return false;
  }
}

template <typename R, typename SI, typename SO> static bool hop_links(const HopLaunch &a, hipStream_t s) {
  if (a.nl == 9) return a.dag ? hop_rb<R, 9, 1, SI, SO>(a, s) : hop_rb<R, 9, 0, SI, SO>(a, s);
  return a.dag ? hop_rb<R, 6, 1, SI, SO>(a, s) : hop_rb<R, 6, 0, SI, SO>(a, s);
}

template <typename R> bool hop_launch(const HopLaunch &a0, hipStream_t s) {
  using C = typename c2<R>::t;
  HopLaunch a = a0;
  {
    // fields beyond one descriptor's 4 GB: per-slice gather and per-block
    // store descriptors, which need whole 64-site blocks per t-slice and a
    // slice within 4 GB
    const double in_b = a.half_in ? 4.0 : (double)sizeof(C), out_b = a.half_out ? 4.0 : (double)sizeof(C);
    const double e = 12.0 * (double)a.g.Vh * (double)a.nrhs;
    a.big = (e * in_b >= 4294967296.0 || e * out_b >= 4294967296.0) ? 1 : 0;
    if (a.big && ((a.g.Vh / a.g.L[3]) % kBlk != 0 || e / a.g.L[3] * in_b >= 4294967296.0 ||
                  a.shape.ls == 2 || a.skip_t_faces))
      return false;
  }
  const HopShape &sh = a.shape;
  if (sh.tpb < 1 || sh.tpb > kMaxTpb || sh.rb < 1 || sh.rb > max_rb<R>() || sh.occ < 0 || sh.occ > kMaxOcc ||
      a.nrhs < 1 || a.nrhs > 64 || a.g.Vh % kBlk)
    return false;
  if (sh.seg < 1 || (sh.seg > 1 && (a.bps <= 0 || sh.seg > a.bps))) return false;
  // k_hop indexes items in 32 bits
  if ((double)a.g.Vh * (double)((a.nrhs + sh.rb - 1) / sh.rb) >= 2147483648.0) return false;
  if (!a.half_in && !a.half_out) return hop_links<R, C, C>(a, s);
  if constexpr (sizeof(R) == 4) {
    if (a.half_out && a.half_out != a.half_in) return false;
    if (a.half_in == kS16Half) return a.half_out ? hop_links<R, c2h, c2h>(a, s) : hop_links<R, c2h, C>(a, s);
    if (a.half_in == kS16Bf16) return a.half_out ? hop_links<R, c2b, c2b>(a, s) : hop_links<R, c2b, C>(a, s);
  }
  return false;
}

}  // namespace wck
