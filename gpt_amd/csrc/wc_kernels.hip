// HIP kernels of the even-odd Wilson-clover operator for gfx950 (MI355X).
//
// Hot kernel: k_hop, one thread per output half-site.  For each of the 8
// directions it gathers the neighbour spinor (12 complex, SoA so a wave reads
// 1 KB contiguous per component), projects it to a half spinor with
// (1 -+ gamma_mu), multiplies by the double-stored link (9 complex, or 6 with
// the third row rebuilt), and accumulates the reconstructed 4-spinor.
// Epilogues fuse the clover (inverse) block, the Schur complement's axpy and
// the CG inner product, so the normal operator needs no extra passes.
#include <hip/hip_runtime.h>

#include "wc_kernels.h"

namespace wck {

int site_blocks(long Vh) { return (int)((Vh + kSiteThreads - 1) / kSiteThreads); }

// ------------------------------------------------------------ clover blocks
// packed Hermitian 6x6: slots 0..2 hold the real diagonal in pairs, slots
// 3..17 the strictly-lower entries in row order (1,0),(2,0),(2,1),(3,0),...
__host__ __device__ constexpr int tri(int i, int j) { return 3 + i * (i - 1) / 2 + j; }

template <typename R, typename C>
__device__ __forceinline__ void clov_apply(const C *A, long Vh, long h, const C v[12], C o[12]) {
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const C *Ab = A + aidx(b * 18, h, 36);
    R d[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      C dd = Ab[(long)k * kBlk];
      d[2 * k] = dd.x;
      d[2 * k + 1] = dd.y;
    }
    C L[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) L[k] = Ab[(long)(3 + k) * kBlk];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      C acc = cscale(v[6 * b + i], d[i]);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (j < i) acc = cadd(acc, cmul(L[tri(i, j) - 3], v[6 * b + j]));
        if (j > i) acc = cadd(acc, cmulc(L[tri(j, i) - 3], v[6 * b + j]));
      }
      o[6 * b + i] = acc;
    }
  }
}

// spinor of site h, right-hand side r of N: entry k at aidx(k, h, 12) * N + r
template <typename C>
__device__ __forceinline__ void load12(const C *f, long h, int r, int N, C v[12]) {
  const C *p = f + aidx_site(h, 12) * N + r;
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = p[(long)k * kBlk * N];
}
template <typename C>
__device__ __forceinline__ void store12(C *f, long h, int r, int N, const C v[12]) {
  C *p = f + aidx_site(h, 12) * N + r;
#pragma unroll
  for (int k = 0; k < 12; ++k) p[(long)k * kBlk * N] = v[k];
}
// output of a hop: written once, not re-read by this kernel
template <typename C>
__device__ __forceinline__ void store12_stream(C *f, long Vh, long h, int r, int N, const C v[12]) {
#if WC_NT_STORE
  const wc_rsrc rs = make_rsrc(f, (uint32_t)(12u * (uint32_t)Vh * (uint32_t)N * (uint32_t)sizeof(C)));
  const uint32_t vo = (uint32_t)(aidx_site(h, 12) * N + r) * (uint32_t)sizeof(C);
#pragma unroll
  for (int k = 0; k < 12; ++k) bstore_t<C, 2>::go(v[k], rs, vo, (uint32_t)(k * kBlk * N * sizeof(C)));
#else
  store12(f, h, r, N, v);
#endif
}

// ---------------------------------------------------------------- the hop
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

// one direction's operands: the neighbour spinor and the link
template <typename C> struct DirData {
  C psi[12];
  C u[9];
  long pad;
};

// project, multiply, accumulate direction (MU, FB)
template <typename R, int NL, int DAG, int MU, int FB>
__device__ __forceinline__ void compute_dir(const HopLaunch &a, const int c[4],
                                            DirData<typename c2<R>::t> &d, typename c2<R>::t acc[12]) {
  using C = typename c2<R>::t;
  const Geom &g = a.g;
  const bool bnd = (FB == 0) ? (c[MU] == g.L[MU] - 1) : (c[MU] == 0);
  constexpr int S = ((FB == 0) != (DAG != 0)) ? 1 : -1;
  C hs[2][3];
#pragma unroll
  for (int col = 0; col < 3; ++col)
    project<MU, S>(d.psi[0 + col], d.psi[3 + col], d.psi[6 + col], d.psi[9 + col], hs[0][col], hs[1][col]);
  // hop coefficient, with the boundary phase on links that wrap around
  const R coef = (R)(MU == 3 ? a.ct : a.cs);
  if (bnd) {
    C f = cmk<C>((R)(a.ph[MU][0]), (R)(FB == 0 ? a.ph[MU][1] : -a.ph[MU][1]));
    f = cscale(f, coef);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int col = 0; col < 3; ++col) hs[s][col] = cmul(f, hs[s][col]);
  } else {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int col = 0; col < 3; ++col) hs[s][col] = cscale(hs[s][col], coef);
  }
  C *u = d.u;
  if (NL == 6) {
    // third row = conj(row0 x row1) for SU(3)
    u[6] = cconj(csub(cmul(u[1], u[5]), cmul(u[2], u[4])));
    u[7] = cconj(csub(cmul(u[2], u[3]), cmul(u[0], u[5])));
    u[8] = cconj(csub(cmul(u[0], u[4]), cmul(u[1], u[3])));
  }
  C ch[2][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      C t = cmul(u[3 * i + 0], hs[s][0]);
      t = cadd(t, cmul(u[3 * i + 1], hs[s][1]));
      t = cadd(t, cmul(u[3 * i + 2], hs[s][2]));
      ch[s][i] = t;
    }
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    acc[0 + col] = cadd(acc[0 + col], ch[0][col]);
    acc[3 + col] = cadd(acc[3 + col], ch[1][col]);
    accum_lower<MU, S>(acc[6 + col], acc[9 + col], ch[0][col], ch[1][col]);
  }
}

// The 8 directions run as a real loop (not unrolled): the loads of one
// direction cannot be hoisted into another, so the live registers stay at one
// direction's operands plus the accumulator; the other waves on the SIMD keep
// HBM busy meanwhile.  The projection variant is picked by a wave-uniform
// switch.  Neighbour offsets come from uniform strides of the half-site index
// h = ((t*Lz + z)*Ly + y)*Lxh + x/2.
template <typename R, int NL, int DAG>
__device__ __forceinline__ void hop_site(const HopLaunch &a, const typename c2<R>::t *inp,
                                         const typename c2<R>::t *W, long h, int r, int N, const int c[4],
                                         typename c2<R>::t acc[12]) {
  using C = typename c2<R>::t;
  const Geom &g = a.g;
  const uint32_t Vh = (uint32_t)g.Vh;
  const uint32_t hh = (uint32_t)h;
  const wc_rsrc in = make_rsrc(inp, 12u * Vh * (uint32_t)N * (uint32_t)sizeof(C));
  // a wave's sites lie in at most two aligned 64-site blocks (one when N = 1):
  // their links are addressed from the wave's first block, a wave-uniform base
  const long blk = (long)__builtin_amdgcn_readfirstlane((int)(h >> 6));
  const wc_rsrc rw = make_rsrc(W + blk * (8 * NL * kBlk), 2u * 8u * NL * kBlk * (uint32_t)sizeof(C));
  const uint32_t vw = (uint32_t)(aidx_site(h, 8 * NL) - blk * (8 * NL * kBlk)) * (uint32_t)sizeof(C);
  const uint32_t kstride = (uint32_t)(kBlk * N * sizeof(C));
  const uint32_t xh = (uint32_t)c[0] >> 1;
  const uint32_t st1 = (uint32_t)g.Lxh, st2 = st1 * (uint32_t)g.L[1], st3 = st2 * (uint32_t)g.L[2];
#pragma unroll 1
  for (int dir = 0; dir < 8; ++dir) {
    const int mu = dir >> 1, fb = dir & 1;
    uint32_t nb;
    if (mu == 0) {
      const int x = c[0];
      const int xn = fb == 0 ? (x == g.L[0] - 1 ? 0 : x + 1) : (x == 0 ? g.L[0] - 1 : x - 1);
      nb = hh - xh + ((uint32_t)xn >> 1);
    } else {
      const uint32_t stride = mu == 1 ? st1 : (mu == 2 ? st2 : st3);
      const int cm = mu == 1 ? c[1] : (mu == 2 ? c[2] : c[3]);
      const int Lm = g.L[mu];
      const uint32_t wrap = (uint32_t)(Lm - 1) * stride;
      if (fb == 0)
        nb = (cm == Lm - 1) ? hh - wrap : hh + stride;
      else
        nb = (cm == 0) ? hh + wrap : hh - stride;
    }
    DirData<C> d;
    const uint32_t vs = ((uint32_t)aidx_site(nb, 12) * (uint32_t)N + (uint32_t)r) * (uint32_t)sizeof(C);
#pragma unroll
    for (int k = 0; k < 12; ++k) d.psi[k] = bload<C, WC_SPINOR_AUX>(in, vs, (uint32_t)k * kstride);
    const uint32_t dsoff = (uint32_t)(dir * NL) * (uint32_t)(kBlk * sizeof(C));
#pragma unroll
    for (int e = 0; e < NL; ++e)
      d.u[e] = bload<C, WC_LINK_AUX>(rw, vw, dsoff + (uint32_t)(e * kBlk * sizeof(C)));
    switch (dir) {
      case 0: compute_dir<R, NL, DAG, 0, 0>(a, c, d, acc); break;
      case 1: compute_dir<R, NL, DAG, 0, 1>(a, c, d, acc); break;
      case 2: compute_dir<R, NL, DAG, 1, 0>(a, c, d, acc); break;
      case 3: compute_dir<R, NL, DAG, 1, 1>(a, c, d, acc); break;
      case 4: compute_dir<R, NL, DAG, 2, 0>(a, c, d, acc); break;
      case 5: compute_dir<R, NL, DAG, 2, 1>(a, c, d, acc); break;
      case 6: compute_dir<R, NL, DAG, 3, 0>(a, c, d, acc); break;
      default: compute_dir<R, NL, DAG, 3, 1>(a, c, d, acc); break;
    }
  }
}

template <typename R, int NL, int DAG, int EPI>
__global__ __launch_bounds__(kSiteThreads, WC_HOP_MINWAVES) void k_hop(const HopLaunch a) {
  using C = typename c2<R>::t;
  __shared__ double red[kSiteThreads];
  const Geom &g = a.g;
  const long Vh = g.Vh;
  const int N = a.nrhs;
  // thread t -> (site h, right-hand side r); the N threads of a site are
  // adjacent, so they share each link load
  const long t0 = xcd_tile(blockIdx.x, gridDim.x) * kSiteThreads;
  const long t = t0 + threadIdx.x;
  const long h = (N == 1) ? t : t / N;
  const int r = (int)(t - h * N);
  double dot = 0.0;
  if (h < Vh) {
    int c[4];
    half_coords(g, h, a.pout, c);
    const C *in = (const C *)a.in;
    const C *W = (const C *)a.W;
    C acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = cmk<C>(0, 0);
    hop_site<R, NL, DAG>(a, in, W, h, r, N, c, acc);
    // epilogue
    const bool zb = a.open_bc && (c[3] == 0 || c[3] == g.L[3] - 1);
    C res[12];
    if (zb) {
#pragma unroll
      for (int k = 0; k < 12; ++k) res[k] = cmk<C>(0, 0);
    } else if (EPI & EPI_CLOV) {
      if (a.clov) {
        clov_apply<R>((const C *)a.clov, Vh, h, acc, res);
      } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) res[k] = cscale(acc[k], (R)a.clov_scalar);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) res[k] = acc[k];
    }
    if (EPI & (EPI_DIAG | EPI_AXPY)) {
      C y[12];
      load12((const C *)a.y, h, r, N, y);
      if (EPI & EPI_DIAG) {
        if (!zb) {
          C dy[12];
          if (a.diag_clov) {
            clov_apply<R>((const C *)a.diag_clov, Vh, h, y, dy);
          } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) dy[k] = cscale(y[k], (R)a.diag);
          }
#pragma unroll
          for (int k = 0; k < 12; ++k) res[k] = cadd(res[k], dy[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) res[k] = csub(y[k], res[k]);
      }
    }
    if (EPI & EPI_DOT) {
      C v[12];
      load12((const C *)a.dotv, h, r, N, v);
#pragma unroll
      for (int k = 0; k < 12; ++k) dot += (double)v[k].x * (double)res[k].x + (double)v[k].y * (double)res[k].y;
    }
    store12_stream((C *)a.out, Vh, h, r, N, res);
  }
  if (EPI & EPI_DOT) {
    if (N == 1) {
      double s = block_sum<kSiteThreads>(dot, red);
      if (threadIdx.x == 0) a.partial[blockIdx.x] = s;
    } else {
      // per-RHS sums of the block in fixed lane order: partial[block * N + r]
      red[threadIdx.x] = dot;
      __syncthreads();
      if (threadIdx.x < N) {
        double s = 0.0;
        const int r0 = (int)(t0 % N);
        for (int i = ((int)threadIdx.x - r0 + N) % N; i < kSiteThreads; i += N) s += red[i];
        a.partial[(long)blockIdx.x * N + threadIdx.x] = s;
      }
    }
  }
}

template <typename R, int NL, int DAG>
static void hop_epi(const HopLaunch &a, hipStream_t s) {
  const int nb = site_blocks(a.g.Vh * a.nrhs);
  switch (a.epi) {
#define WC_EPI(E) \
  case E: k_hop<R, NL, DAG, E><<<nb, kSiteThreads, 0, s>>>(a); break;
    WC_EPI(0)
    WC_EPI(EPI_CLOV)
    WC_EPI(EPI_DIAG)
    WC_EPI(EPI_AXPY)
    WC_EPI(EPI_CLOV | EPI_AXPY)
    WC_EPI(EPI_CLOV | EPI_AXPY | EPI_DOT)
    WC_EPI(EPI_AXPY | EPI_DOT)
#undef WC_EPI
    default: break;
  }
}

template <typename R> void hop(const HopLaunch &a, hipStream_t s) {
  if (a.nl == 9) {
    if (a.dag) hop_epi<R, 9, 1>(a, s); else hop_epi<R, 9, 0>(a, s);
  } else {
    if (a.dag) hop_epi<R, 6, 1>(a, s); else hop_epi<R, 6, 0>(a, s);
  }
}
template void hop<double>(const HopLaunch &, hipStream_t);
template void hop<float>(const HopLaunch &, hipStream_t);

// ------------------------------------------------------- site-local clover
template <typename R>
__global__ __launch_bounds__(kSiteThreads) void k_site_clover(Geom g, int N, int p, const typename c2<R>::t *A,
                                                             double diag, const typename c2<R>::t *in,
                                                             typename c2<R>::t *out,
                                                             const typename c2<R>::t *y, int open_bc) {
  using C = typename c2<R>::t;
  const long t = xcd_tile(blockIdx.x, gridDim.x) * kSiteThreads + threadIdx.x;
  const long h = (N == 1) ? t : t / N;
  const int r = (int)(t - h * N);
  if (h >= g.Vh) return;
  C v[12], o[12];
  load12(in, h, r, N, v);
  bool zb = false;
  if (open_bc) {
    int c[4];
    half_coords(g, h, p, c);
    zb = (c[3] == 0 || c[3] == g.L[3] - 1);
  }
  if (zb) {
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = cmk<C>(0, 0);
  } else if (A) {
    clov_apply<R>(A, g.Vh, h, v, o);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = cscale(v[k], (R)diag);
  }
  if (y) {
    C yy[12];
    load12(y, h, r, N, yy);
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = csub(yy[k], o[k]);
  }
  store12(out, h, r, N, o);
}

template <typename R>
void site_clover(const Geom &g, int nrhs, int p, const void *blocks, double diag, const void *in, void *out,
                 const void *y, int open_bc, hipStream_t s) {
  using C = typename c2<R>::t;
  k_site_clover<R><<<site_blocks(g.Vh * nrhs), kSiteThreads, 0, s>>>(g, nrhs, p, (const C *)blocks, diag,
                                                                     (const C *)in, (C *)out, (const C *)y,
                                                                     open_bc);
}
template void site_clover<double>(const Geom &, int, int, const void *, double, const void *, void *,
                                  const void *, int, hipStream_t);
template void site_clover<float>(const Geom &, int, int, const void *, double, const void *, void *,
                                 const void *, int, hipStream_t);

// ------------------------------------------------------------------ BLAS-1
static int elem_blocks(long n) { return (int)((n + 255) / 256); }

template <typename R>
__global__ void k_axpy(long n, double ar, double ai, const typename c2<R>::t *x,
                       const typename c2<R>::t *y, typename c2<R>::t *z) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  C xv = x[i], yv = y[i];
  z[i] = cadd(cmul(cmk<C>((R)ar, (R)ai), xv), yv);
}
template <typename R>
void axpy(long n, double ar, double ai, const void *x, const void *y, void *z, hipStream_t s) {
  using C = typename c2<R>::t;
  if (n <= 0) return;
  k_axpy<R><<<elem_blocks(n), 256, 0, s>>>(n, ar, ai, (const C *)x, (const C *)y, (C *)z);
}
template void axpy<double>(long, double, double, const void *, const void *, void *, hipStream_t);
template void axpy<float>(long, double, double, const void *, const void *, void *, hipStream_t);

template <typename R>
__global__ void k_scale(long n, double ar, double ai, const typename c2<R>::t *x, typename c2<R>::t *z) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  z[i] = cmul(cmk<C>((R)ar, (R)ai), x[i]);
}
template <typename R> void scale(long n, double ar, double ai, const void *x, void *z, hipStream_t s) {
  using C = typename c2<R>::t;
  if (n <= 0) return;
  k_scale<R><<<elem_blocks(n), 256, 0, s>>>(n, ar, ai, (const C *)x, (C *)z);
}
template void scale<double>(long, double, double, const void *, void *, hipStream_t);
template void scale<float>(long, double, double, const void *, void *, hipStream_t);

// reductions: fixed grid of kRedBlocks x 256, grid-stride; deterministic
template <typename R>
__global__ __launch_bounds__(256) void k_dot(long n, const typename c2<R>::t *a, const typename c2<R>::t *b,
                                             double *pre, double *pim) {
  __shared__ double red[256 / 64 + 1];
  double sr = 0.0, si = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
    sr += ax * bx + ay * by;
    si += ax * by - ay * bx;
  }
  sr = block_sum<256>(sr, red);
  if (pim) si = block_sum<256>(si, red);
  if (threadIdx.x == 0) {
    pre[blockIdx.x] = sr;
    if (pim) pim[blockIdx.x] = si;
  }
}
template <typename R>
void dot_partials(long n, const void *a, const void *b, double *pre, double *pim, hipStream_t s) {
  using C = typename c2<R>::t;
  k_dot<R><<<kRedBlocks, 256, 0, s>>>(n, (const C *)a, (const C *)b, pre, pim);
}
template void dot_partials<double>(long, const void *, const void *, double *, double *, hipStream_t);
template void dot_partials<float>(long, const void *, const void *, double *, double *, hipStream_t);

template <typename R> void norm2_partials(long n, const void *a, double *pre, hipStream_t s) {
  dot_partials<R>(n, a, a, pre, nullptr, s);
}
template void norm2_partials<double>(long, const void *, double *, hipStream_t);
template void norm2_partials<float>(long, const void *, double *, hipStream_t);

template <typename R>
__global__ __launch_bounds__(256) void k_axpy_norm2(long n, double ar, double ai, const typename c2<R>::t *x,
                                                    const typename c2<R>::t *y, typename c2<R>::t *z,
                                                    double *pre) {
  using C = typename c2<R>::t;
  __shared__ double red[256 / 64 + 1];
  double sr = 0.0;
  const C av = cmk<C>((R)ar, (R)ai);
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    C v = cadd(cmul(av, x[i]), y[i]);
    z[i] = v;
    sr += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  sr = block_sum<256>(sr, red);
  if (threadIdx.x == 0) pre[blockIdx.x] = sr;
}
template <typename R>
void axpy_norm2_partials(long n, double ar, double ai, const void *x, const void *y, void *z, double *pre,
                         hipStream_t s) {
  using C = typename c2<R>::t;
  k_axpy_norm2<R><<<kRedBlocks, 256, 0, s>>>(n, ar, ai, (const C *)x, (const C *)y, (C *)z, pre);
}
template void axpy_norm2_partials<double>(long, double, double, const void *, const void *, void *, double *,
                                          hipStream_t);
template void axpy_norm2_partials<float>(long, double, double, const void *, const void *, void *, double *,
                                         hipStream_t);

__global__ __launch_bounds__(1024) void k_reduce_final(const double *p, int n, double *out) {
  __shared__ double red[1024 / 64 + 1];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += p[i];
  s = block_sum<1024>(s, red);
  if (threadIdx.x == 0) out[0] = s;
}
void reduce_final(const double *partials, int n, double *out, hipStream_t s) {
  k_reduce_final<<<1, 1024, 0, s>>>(partials, n, out);
}

// ------------------------------------------------------------- conversions
__global__ void k_d2f(long n, const double2 *a, float2 *b) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void k_f2d(long n, const float2 *a, double2 *b) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
void convert_d2f(long n, const void *src, void *dst, hipStream_t s) {
  if (n > 0) k_d2f<<<elem_blocks(n), 256, 0, s>>>(n, (const double2 *)src, (float2 *)dst);
}
void convert_f2d(long n, const void *src, void *dst, hipStream_t s) {
  if (n > 0) k_f2d<<<elem_blocks(n), 256, 0, s>>>(n, (const float2 *)src, (double2 *)dst);
}

template <typename R>
__global__ void k_import(Geom g, const typename c2<R>::t *nat, typename c2<R>::t *eo) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.V) return;
  long r = i / g.L[0];
  int x = (int)(i - r * g.L[0]);
  long r2 = r / g.L[1];
  int y = (int)(r - r2 * g.L[1]);
  long r3 = r2 / g.L[2];
  int z = (int)(r2 - r3 * g.L[2]);
  int t = (int)r3;
  int p = (x + y + z + t) & 1;
  long h = i >> 1;
#pragma unroll
  for (int k = 0; k < 12; ++k) eo[(long)p * 12 * g.Vh + aidx(k, h, 12)] = nat[i * 12 + k];
}
template <typename R>
__global__ void k_export(Geom g, const typename c2<R>::t *eo, typename c2<R>::t *nat) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.V) return;
  long r = i / g.L[0];
  int x = (int)(i - r * g.L[0]);
  long r2 = r / g.L[1];
  int y = (int)(r - r2 * g.L[1]);
  long r3 = r2 / g.L[2];
  int z = (int)(r2 - r3 * g.L[2]);
  int t = (int)r3;
  int p = (x + y + z + t) & 1;
  long h = i >> 1;
#pragma unroll
  for (int k = 0; k < 12; ++k) nat[i * 12 + k] = eo[(long)p * 12 * g.Vh + aidx(k, h, 12)];
}
template <typename R> void import_field(const Geom &g, const void *nat, void *eo, hipStream_t s) {
  using C = typename c2<R>::t;
  k_import<R><<<elem_blocks(g.V), 256, 0, s>>>(g, (const C *)nat, (C *)eo);
}
template <typename R> void export_field(const Geom &g, const void *eo, void *nat, hipStream_t s) {
  using C = typename c2<R>::t;
  k_export<R><<<elem_blocks(g.V), 256, 0, s>>>(g, (const C *)eo, (C *)nat);
}
template void import_field<double>(const Geom &, const void *, void *, hipStream_t);
template void import_field<float>(const Geom &, const void *, void *, hipStream_t);
template void export_field<double>(const Geom &, const void *, void *, hipStream_t);
template void export_field<float>(const Geom &, const void *, void *, hipStream_t);

// ----------------------------------------------------------- link set-up
__device__ __forceinline__ long nat_index(const Geom &g, int x, int y, int z, int t) {
  return (((long)t * g.L[2] + z) * g.L[1] + y) * g.L[0] + x;
}

template <typename R>
__global__ void k_build_links(Geom g, const double2 *U, int nl, typename c2<R>::t *W0, typename c2<R>::t *W1) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= 2 * g.Vh) return;
  const int p = (int)(i / g.Vh);
  const long h = i - (long)p * g.Vh;
  C *W = p ? W1 : W0;
  int c[4];
  half_coords(g, h, p, c);
  const long xi = nat_index(g, c[0], c[1], c[2], c[3]);
  for (int mu = 0; mu < 4; ++mu) {
    int m[4] = {c[0], c[1], c[2], c[3]};
    m[mu] = (c[mu] + g.L[mu] - 1) % g.L[mu];
    const long xm = nat_index(g, m[0], m[1], m[2], m[3]);
    const double2 *uf = U + ((long)mu * g.V + xi) * 9;
    const double2 *ub = U + ((long)mu * g.V + xm) * 9;
    for (int e = 0; e < nl; ++e) {
      const int r = e / 3, col = e % 3;
      double2 f = uf[e];
      double2 b = ub[col * 3 + r];  // (U^dag)_{r col} = conj(U_{col r})
      W[aidx(2 * mu * nl + e, h, 8 * nl)] = cmk<C>((R)f.x, (R)f.y);
      W[aidx((2 * mu + 1) * nl + e, h, 8 * nl)] = cmk<C>((R)b.x, (R)(-b.y));
    }
  }
}
template <typename R>
void build_links(const Geom &g, const void *Unat, int nl, void *W0, void *W1, hipStream_t s) {
  using C = typename c2<R>::t;
  k_build_links<R><<<elem_blocks(2 * g.Vh), 256, 0, s>>>(g, (const double2 *)Unat, nl, (C *)W0, (C *)W1);
}
template void build_links<double>(const Geom &, const void *, int, void *, void *, hipStream_t);
template void build_links<float>(const Geom &, const void *, int, void *, void *, hipStream_t);

// ---------------------------------------------------------- clover set-up
struct M3 {
  double2 e[9];
};
__device__ __forceinline__ M3 ld3(const double2 *U, const Geom &g, int mu, const int x[4]) {
  M3 m;
  const long xi = nat_index(g, x[0], x[1], x[2], x[3]);
  const double2 *p = U + ((long)mu * g.V + xi) * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) m.e[k] = p[k];
  return m;
}
// op: 0 = A B, 1 = A B^dag, 2 = A^dag B
__device__ __forceinline__ M3 mm(const M3 &A, const M3 &B, int op) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double2 a = (op == 2) ? cconj(A.e[k * 3 + i]) : A.e[i * 3 + k];
        double2 b = (op == 1) ? cconj(B.e[j * 3 + k]) : B.e[k * 3 + j];
        s = cadd(s, cmul(a, b));
      }
      r.e[i * 3 + j] = s;
    }
  return r;
}
__device__ __forceinline__ M3 msub(const M3 &A, const M3 &B) {
  M3 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.e[k] = csub(A.e[k], B.e[k]);
  return r;
}
__device__ __forceinline__ void shiftc(const Geom &g, const int x[4], int mu, int d, int o[4]) {
  o[0] = x[0]; o[1] = x[1]; o[2] = x[2]; o[3] = x[3];
  o[mu] = (x[mu] + d + g.L[mu]) % g.L[mu];
}

// v(y) = staple_up(y) - staple_down(y) in the (mu, nu) plane:
//   staple_up(y)   = U_nu(y+mu) U_mu(y+nu)^dag U_nu(y)^dag
//   staple_down(y) = U_nu(y-nu+mu)^dag U_mu(y-nu)^dag U_nu(y-nu)
__device__ M3 staple_diff(const double2 *U, const Geom &g, const int y[4], int mu, int nu) {
  int ypm[4], ypn[4], ymn[4], ymnpm[4];
  shiftc(g, y, mu, 1, ypm);
  shiftc(g, y, nu, 1, ypn);
  shiftc(g, y, nu, -1, ymn);
  shiftc(g, ymn, mu, 1, ymnpm);
  M3 up = mm(mm(ld3(U, g, nu, ypm), ld3(U, g, mu, ypn), 1), ld3(U, g, nu, y), 1);
  // (U_mu(y-nu) U_nu(y-nu+mu))^dag U_nu(y-nu)
  M3 t = mm(ld3(U, g, mu, ymn), ld3(U, g, nu, ymnpm), 0);
  M3 dn = mm(t, ld3(U, g, nu, ymn), 2);
  return msub(up, dn);
}

// clover-leaf F_{mu nu}(x) = (Q - Q^dag)/8, Q = U_mu(x) v(x) + v(x-mu) U_mu(x-mu)
// (lib/gpt/qcd/gauge/loops.py:162-174)
__device__ M3 field_strength(const double2 *U, const Geom &g, const int x[4], int mu, int nu) {
  int xm[4];
  shiftc(g, x, mu, -1, xm);
  M3 F = mm(ld3(U, g, mu, x), staple_diff(U, g, x, mu, nu), 0);
  M3 F2 = mm(staple_diff(U, g, xm, mu, nu), ld3(U, g, mu, xm), 0);
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 q = cadd(F.e[i * 3 + j], F2.e[i * 3 + j]);
      double2 qa = cadd(F.e[j * 3 + i], F2.e[j * 3 + i]);
      r.e[i * 3 + j] = cscale(csub(q, cconj(qa)), 0.125);
    }
  return r;
}

template <typename R>
__global__ __launch_bounds__(128) void k_build_clover(Geom g, const double2 *U, CloverParams cp,
                                                      typename c2<R>::t *A0, typename c2<R>::t *A1,
                                                      typename c2<R>::t *I0, typename c2<R>::t *I1,
                                                      int *err) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 128 + threadIdx.x;
  if (i >= g.V) return;
  long r = i / g.L[0];
  int x[4];
  x[0] = (int)(i - r * g.L[0]);
  long r2 = r / g.L[1];
  x[1] = (int)(r - r2 * g.L[1]);
  long r3 = r2 / g.L[2];
  x[2] = (int)(r2 - r3 * g.L[2]);
  x[3] = (int)r3;
  const int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  const long h = i >> 1;
  const int T = g.L[3];
  const bool edge = cp.open_bc && (x[3] == 0 || x[3] == T - 1);
  const bool next = cp.open_bc && cp.cF != 1.0 && (x[3] == 1 || x[3] == T - 2);

  // B[chir][6x6] (full Hermitian, only lower + diag used)
  double2 B[2][36];
#pragma unroll
  for (int k = 0; k < 36; ++k) {
    B[0][k] = make_double2(0.0, 0.0);
    B[1][k] = make_double2(0.0, 0.0);
  }
  if (!edge) {
    int pl = 0;
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu, ++pl) {
        if (cp.coef[pl] == 0.0) continue;
        M3 F = field_strength(U, g, x, mu, nu);
#pragma unroll
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              double2 sg = make_double2(cp.sig[pl][ch][a * 2 + b][0], cp.sig[pl][ch][a * 2 + b][1]);
              sg = cscale(sg, cp.coef[pl]);
#pragma unroll
              for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
                for (int c2_ = 0; c2_ < 3; ++c2_) {
                  int ii = a * 3 + c1, jj = b * 3 + c2_;
                  B[ch][ii * 6 + jj] = cadd(B[ch][ii * 6 + jj], cmul(sg, F.e[c1 * 3 + c2_]));
                }
            }
      }
  }
  double dg = cp.diag;
  if (edge) dg += -0.5 * cp.csw_t;
  if (next) dg += cp.cF - 1.0;
#pragma unroll
  for (int ch = 0; ch < 2; ++ch)
#pragma unroll
    for (int k = 0; k < 6; ++k) B[ch][k * 7].x += dg;

  C *A = p ? A1 : A0;
  C *Ai = p ? I1 : I0;
  const long Vh = g.Vh;
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    // store A (packed)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      A[aidx(ch * 18 + k, h, 36)] = cmk<C>((R)B[ch][(2 * k) * 7].x, (R)B[ch][(2 * k + 1) * 7].x);
#pragma unroll
    for (int ii = 1; ii < 6; ++ii)
#pragma unroll
      for (int jj = 0; jj < ii; ++jj) {
        double2 v = B[ch][ii * 6 + jj];
        A[aidx(ch * 18 + tri(ii, jj), h, 36)] = cmk<C>((R)v.x, (R)v.y);
      }
    // LDL^H: B = L D L^H (L unit lower)
    double d[6];
    double2 L[36];
    double scale = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) scale = fmax(scale, fabs(B[ch][k * 7].x));
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double dj = B[ch][j * 7].x;
#pragma unroll
      for (int k = 0; k < j; ++k) {
        double2 l = L[j * 6 + k];
        dj -= (l.x * l.x + l.y * l.y) * d[k];
      }
      if (!(fabs(dj) > 1e-12 * scale)) {
        *err = 1;
        dj = 1.0;
      }
      d[j] = dj;
#pragma unroll
      for (int ii = j + 1; ii < 6; ++ii) {
        double2 s = B[ch][ii * 6 + j];
#pragma unroll
        for (int k = 0; k < j; ++k) s = csub(s, cscale(cmulc(L[j * 6 + k], L[ii * 6 + k]), d[k]));
        // s - sum L_ik conj(L_jk) d_k  (cmulc gives conj(L_jk) * L_ik)
        L[ii * 6 + j] = cscale(s, 1.0 / dj);
      }
    }
    // M = L^-1 (unit lower)
    double2 M[36];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
#pragma unroll
      for (int ii = j + 1; ii < 6; ++ii) {
        double2 s = L[ii * 6 + j];
#pragma unroll
        for (int k = j + 1; k < ii; ++k) s = cadd(s, cmul(L[ii * 6 + k], M[k * 6 + j]));
        M[ii * 6 + j] = make_double2(-s.x, -s.y);
      }
    }
    // (B^-1)_ij = sum_{k >= i} conj(M_ki) M_kj / d_k  for i >= j (M_kk = 1)
#pragma unroll
    for (int ii = 0; ii < 6; ++ii)
#pragma unroll
      for (int jj = 0; jj <= ii; ++jj) {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = ii; k < 6; ++k) {
          double2 mki = (k == ii) ? make_double2(1.0, 0.0) : M[k * 6 + ii];
          double2 mkj = (k == jj) ? make_double2(1.0, 0.0) : M[k * 6 + jj];
          s = cadd(s, cscale(cmulc(mki, mkj), 1.0 / d[k]));
        }
        if (ii == jj) {
          B[ch][ii * 7] = s;  // B's diagonal is dead: keep the inverse's diagonal there
        } else {
          Ai[aidx(ch * 18 + tri(ii, jj), h, 36)] = cmk<C>((R)s.x, (R)s.y);
        }
      }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      Ai[aidx(ch * 18 + k, h, 36)] = cmk<C>((R)B[ch][(2 * k) * 7].x, (R)B[ch][(2 * k + 1) * 7].x);
  }
}

template <typename R>
void build_clover(const Geom &g, const void *Unat, const CloverParams &cp, void *A0, void *A1, void *Ai0,
                  void *Ai1, int *err_flag, hipStream_t s) {
  using C = typename c2<R>::t;
  k_build_clover<R><<<(int)((g.V + 127) / 128), 128, 0, s>>>(g, (const double2 *)Unat, cp, (C *)A0, (C *)A1,
                                                            (C *)Ai0, (C *)Ai1, err_flag);
}
template void build_clover<double>(const Geom &, const void *, const CloverParams &, void *, void *, void *,
                                   void *, int *, hipStream_t);
template void build_clover<float>(const Geom &, const void *, const CloverParams &, void *, void *, void *,
                                  void *, int *, hipStream_t);

// ------------------------------------------------- per-RHS reductions
// block size: the largest multiple of nrhs not above 256, so that a thread's
// elements (strided by a multiple of nrhs) all belong to one RHS
int rhs_block_threads(int nrhs) { return (256 / nrhs) * nrhs; }

// sum of lds[0..nt) over the entries of RHS threadIdx.x (< nrhs), fixed order
__device__ __forceinline__ double seg_sum(const double *lds, int nt, int nrhs) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nt; i += nrhs) s += lds[i];
  return s;
}

template <typename R>
__global__ __launch_bounds__(256) void k_dot_rhs(long n, int N, const typename c2<R>::t *a,
                                                 const typename c2<R>::t *b, double *pre, double *pim) {
  __shared__ double lr[256], li[256];
  const long stride = (long)gridDim.x * blockDim.x;
  double sr = 0.0, si = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
    sr += ax * bx + ay * by;
    si += ax * by - ay * bx;
  }
  lr[threadIdx.x] = sr;
  li[threadIdx.x] = si;
  __syncthreads();
  if ((int)threadIdx.x < N) {
    pre[(long)blockIdx.x * N + threadIdx.x] = seg_sum(lr, blockDim.x, N);
    if (pim) pim[(long)blockIdx.x * N + threadIdx.x] = seg_sum(li, blockDim.x, N);
  }
}
template <typename R>
void dot_partials_rhs(long n, int nrhs, const void *a, const void *b, double *pre, double *pim, hipStream_t s) {
  using C = typename c2<R>::t;
  k_dot_rhs<R><<<kRedBlocks, rhs_block_threads(nrhs), 0, s>>>(n, nrhs, (const C *)a, (const C *)b, pre, pim);
}
template void dot_partials_rhs<double>(long, int, const void *, const void *, double *, double *, hipStream_t);
template void dot_partials_rhs<float>(long, int, const void *, const void *, double *, double *, hipStream_t);

// out[r] = sum_b partials[b * N + r]; one block per RHS, fixed order
__global__ __launch_bounds__(256) void k_reduce_final_rhs(const double *p, int nb, int N, double *out) {
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += p[(long)b * N + r];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) out[r] = s;
}
void reduce_final_rhs(const double *partials, int nblocks, int nrhs, double *out, hipStream_t s) {
  k_reduce_final_rhs<<<nrhs, 256, 0, s>>>(partials, nblocks, nrhs, out);
}

__global__ void k_acc_f2d(long n, const float2 *lo, double2 *hi) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double2 h = hi[i];
  float2 l = lo[i];
  hi[i] = make_double2(h.x + (double)l.x, h.y + (double)l.y);
}
void accumulate_f2d(long n, const void *lo, void *hi, hipStream_t s) {
  if (n > 0) k_acc_f2d<<<elem_blocks(n), 256, 0, s>>>(n, (const float2 *)lo, (double2 *)hi);
}
template <typename R> __global__ void k_zero(long n, typename c2<R>::t *x) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = cmk<C>(0, 0);
}
template <typename R> void zero(long n, void *x, hipStream_t s) {
  using C = typename c2<R>::t;
  if (n > 0) k_zero<R><<<elem_blocks(n), 256, 0, s>>>(n, (C *)x);
}
template void zero<double>(long, void *, hipStream_t);
template void zero<float>(long, void *, hipStream_t);

template <typename R>
__global__ void k_scale_rhs(long n, int N, const double *f, typename c2<R>::t *x) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const R a = (R)f[i % N];
  C v = x[i];
  x[i] = cmk<C>(a * v.x, a * v.y);
}
template <typename R> void scale_rhs(long n, int nrhs, const double *f, void *x, hipStream_t s) {
  using C = typename c2<R>::t;
  if (n > 0) k_scale_rhs<R><<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (C *)x);
}
template void scale_rhs<double>(long, int, const double *, void *, hipStream_t);
template void scale_rhs<float>(long, int, const double *, void *, hipStream_t);

// ------------------------------------------------------------ CG scalars
// st[j * N + r]: j = 0 c, 1 d, 2 alpha, 3 beta, 4 cp, 5 active
__global__ __launch_bounds__(256) void k_cg_alpha(const double *p, int nb, int N, double *st) {
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += p[(long)b * N + r];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    st[1 * N + r] = s;
    st[2 * N + r] = (st[5 * N + r] != 0.0 && s != 0.0) ? st[0 * N + r] / s : 0.0;
  }
}
__global__ __launch_bounds__(256) void k_cg_beta(const double *p, int nb, int N, double *st) {
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += p[(long)b * N + r];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    st[4 * N + r] = s;
    if (st[5 * N + r] != 0.0) {
      st[3 * N + r] = s / st[0 * N + r];
      st[0 * N + r] = s;
    } else {
      st[3 * N + r] = 0.0;
    }
  }
}
void cg_alpha(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s) {
  k_cg_alpha<<<nrhs, 256, 0, s>>>(partials, nblocks, nrhs, st);
}
void cg_beta(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s) {
  k_cg_beta<<<nrhs, 256, 0, s>>>(partials, nblocks, nrhs, st);
}

// r -= alpha_r mmp; x += alpha_r p; partial |r|^2 per RHS
template <typename R>
__global__ __launch_bounds__(256) void k_cg_update1(long n, int N, const double *st, typename c2<R>::t *r,
                                                    const typename c2<R>::t *mmp, typename c2<R>::t *x,
                                                    const typename c2<R>::t *p, double *pre) {
  using C = typename c2<R>::t;
  __shared__ double lds[256];
  const long stride = (long)gridDim.x * blockDim.x;
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const R a = (R)st[2 * N + (int)(i0 % N)];
  double s = 0.0;
  for (long i = i0; i < n; i += stride) {
    C rv = r[i], mv = mmp[i], xv = x[i], pv = p[i];
    rv = cmk<C>(rv.x - a * mv.x, rv.y - a * mv.y);
    xv = cmk<C>(xv.x + a * pv.x, xv.y + a * pv.y);
    r[i] = rv;
    x[i] = xv;
    s += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < N) pre[(long)blockIdx.x * N + threadIdx.x] = seg_sum(lds, blockDim.x, N);
}
template <typename R>
void cg_update1(long n, int nrhs, const double *st, void *r, const void *mmp, void *x, const void *p,
                double *partials, hipStream_t s) {
  using C = typename c2<R>::t;
  k_cg_update1<R><<<kRedBlocks, rhs_block_threads(nrhs), 0, s>>>(n, nrhs, st, (C *)r, (const C *)mmp, (C *)x,
                                                                 (const C *)p, partials);
}
template void cg_update1<double>(long, int, const double *, void *, const void *, void *, const void *,
                                 double *, hipStream_t);
template void cg_update1<float>(long, int, const double *, void *, const void *, void *, const void *,
                                double *, hipStream_t);

// p = beta_r p + r
template <typename R>
__global__ void k_cg_update2(long n, int N, const double *st, const typename c2<R>::t *r,
                             typename c2<R>::t *p) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const R b = (R)st[3 * N + (int)(i % N)];
  C pv = p[i], rv = r[i];
  p[i] = cmk<C>(b * pv.x + rv.x, b * pv.y + rv.y);
}
template <typename R> void cg_update2(long n, int nrhs, const double *st, const void *r, void *p, hipStream_t s) {
  using C = typename c2<R>::t;
  k_cg_update2<R><<<elem_blocks(n), 256, 0, s>>>(n, nrhs, st, (const C *)r, (C *)p);
}
template void cg_update2<double>(long, int, const double *, const void *, void *, hipStream_t);
template void cg_update2<float>(long, int, const double *, const void *, void *, hipStream_t);

}  // namespace wck
