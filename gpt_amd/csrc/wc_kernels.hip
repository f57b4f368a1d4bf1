// HIP kernels of the even-odd Wilson-clover operator for gfx950 (MI355X),
// other than the hop (wc_hop.h): the site-local clover, BLAS-1 with
// deterministic reductions, the CG scalar and update kernels, layout
// conversions and the link / clover set-up.
#include <hip/hip_runtime.h>

#include "wc_kernels.h"
#include "wc_site.h"

namespace wck {

int site_blocks(long Vh) { return (int)((Vh + kSiteThreads - 1) / kSiteThreads); }

// ------------------------------------------------------- site-local clover
// S: storage of in and out (y, when given, in the same storage)
template <typename R, typename S>
__global__ __launch_bounds__(kSiteThreads) void k_site_clover(Geom g, int N, int p, const typename c2<R>::t *A,
                                                             double diag, const S *in, S *out, const S *y,
                                                             int open_bc, const long *roff) {
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
    clov_apply<R>(roff ? A + roff[r] : A, h, v, o);
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
                 const void *y, int open_bc, hipStream_t s, int half, const long *roff) {
  using C = typename c2<R>::t;
  const int nb = site_blocks(g.Vh * nrhs);
  if constexpr (sizeof(R) == 4) {
    if (half == kS16Half) {
      k_site_clover<R, c2h><<<nb, kSiteThreads, 0, s>>>(g, nrhs, p, (const C *)blocks, diag, (const c2h *)in,
                                                        (c2h *)out, (const c2h *)y, open_bc, roff);
      return;
    }
    if (half == kS16Bf16) {
      k_site_clover<R, c2b><<<nb, kSiteThreads, 0, s>>>(g, nrhs, p, (const C *)blocks, diag, (const c2b *)in,
                                                        (c2b *)out, (const c2b *)y, open_bc, roff);
      return;
    }
  }
  k_site_clover<R, C><<<nb, kSiteThreads, 0, s>>>(g, nrhs, p, (const C *)blocks, diag, (const C *)in, (C *)out,
                                                  (const C *)y, open_bc, roff);
}
template void site_clover<double>(const Geom &, int, int, const void *, double, const void *, void *,
                                  const void *, int, hipStream_t, int, const long *);
template void site_clover<float>(const Geom &, int, int, const void *, double, const void *, void *,
                                 const void *, int, hipStream_t, int, const long *);

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

// ------------------------------------------------------- gauge observables
// U: natural FP64 links [4][V][3][3]; per site, Re tr of the 6 plaquettes
// U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag (mu < nu) and of the 4 links,
// summed per block (fixed grid: deterministic)
__device__ __forceinline__ void su3_load(const double2 *U, long V, int mu, long site, double2 m[9]) {
  const double2 *p = U + ((long)mu * V + site) * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = p[k];
}
__global__ __launch_bounds__(256) void k_gauge_obs(Geom g, const double2 *U, double *pplaq, double *ptrace) {
  __shared__ double red[256 / 64 + 1];
  const long V = 2 * g.Vh;
  double sp = 0.0, st = 0.0;
  for (long x = (long)blockIdx.x * 256 + threadIdx.x; x < V; x += (long)gridDim.x * 256) {
    int c[4];
    long r = x;
    for (int mu = 0; mu < 4; ++mu) {
      c[mu] = (int)(r % g.L[mu]);
      r /= g.L[mu];
    }
    long stride[4];
    stride[0] = 1;
    for (int mu = 1; mu < 4; ++mu) stride[mu] = stride[mu - 1] * g.L[mu - 1];
    auto shift = [&](long site, int mu) {  // site + mu, periodic
      const int cm = (int)((site / stride[mu]) % g.L[mu]);
      return cm == g.L[mu] - 1 ? site - (long)(g.L[mu] - 1) * stride[mu] : site + stride[mu];
    };
    (void)c;
    for (int mu = 0; mu < 4; ++mu) {
      double2 a[9];
      su3_load(U, V, mu, x, a);
      st += a[0].x + a[4].x + a[8].x;
      const long xm = shift(x, mu);
      for (int nu = mu + 1; nu < 4; ++nu) {
        double2 b[9], cc[9], d[9];
        su3_load(U, V, nu, xm, b);               // U_nu(x+mu)
        su3_load(U, V, mu, shift(x, nu), cc);    // U_mu(x+nu)
        su3_load(U, V, nu, x, d);                // U_nu(x)
        // e = a b
        double2 e[9], f[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double2 t = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const double2 u = a[3 * i + k], v = b[3 * k + j];
              t.x = fma(u.x, v.x, fma(-u.y, v.y, t.x));
              t.y = fma(u.x, v.y, fma(u.y, v.x, t.y));
            }
            e[3 * i + j] = t;
          }
        // f = d cc (then Re tr(e f^dag) = Re tr(a b cc^dag d^dag))
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double2 t = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const double2 u = d[3 * i + k], v = cc[3 * k + j];
              t.x = fma(u.x, v.x, fma(-u.y, v.y, t.x));
              t.y = fma(u.x, v.y, fma(u.y, v.x, t.y));
            }
            f[3 * i + j] = t;
          }
        // Re tr(e f^dag) = sum_ij Re(e_ij conj(f_ij))
        double tr = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) tr = fma(e[k].x, f[k].x, fma(e[k].y, f[k].y, tr));
        sp += tr;
      }
    }
  }
  sp = block_sum<256>(sp, red);
  st = block_sum<256>(st, red);
  if (threadIdx.x == 0) {
    pplaq[blockIdx.x] = sp;
    ptrace[blockIdx.x] = st;
  }
}
void gauge_obs_partials(const Geom &g, const void *U, double *pplaq, double *ptrace, hipStream_t s) {
  k_gauge_obs<<<kRedBlocks, 256, 0, s>>>(g, (const double2 *)U, pplaq, ptrace);
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
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    // store A (packed), unless only the inverse is wanted
    if (A) {
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
    pre[(long)threadIdx.x * gridDim.x + blockIdx.x] = seg_sum(lr, blockDim.x, N);
    if (pim) pim[(long)threadIdx.x * gridDim.x + blockIdx.x] = seg_sum(li, blockDim.x, N);
  }
}
template <typename R>
void dot_partials_rhs(long n, int nrhs, const void *a, const void *b, double *pre, double *pim, hipStream_t s) {
  using C = typename c2<R>::t;
  k_dot_rhs<R><<<kRedBlocks, rhs_block_threads(nrhs), 0, s>>>(n, nrhs, (const C *)a, (const C *)b, pre, pim);
}
template void dot_partials_rhs<double>(long, int, const void *, const void *, double *, double *, hipStream_t);
template void dot_partials_rhs<float>(long, int, const void *, const void *, double *, double *, hipStream_t);

// sum of p[r * nb + b] over b in a fixed order (thread t: b = t mod 256,
// then the block tree); p is read contiguously
__device__ __forceinline__ double rhs_total(const double *p, int nb, int r, double *red) {
  const double *q = p + (long)r * nb;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += q[b];
  return block_sum<256>(s, red);
}

// per-(block, RHS) sums of per-site products, sites in order
__global__ void k_dot_blocks(long nb, int N, const double *dots, double *partial, const int *halt) {
  if (halt && *halt) return;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nb * N) return;
  const long b = t / N;
  const int r = (int)(t - b * N);
  const double *p = dots + b * kBlk * N + r;
  double s = 0.0;
  for (int i = 0; i < kBlk; ++i) s += p[(long)i * N];
  partial[(long)r * nb + b] = s;
}
void dot_blocks(const Geom &g, int nrhs, const double *dots, double *partial, hipStream_t s, const int *halt) {
  const long nb = hop_blocks(g);
  k_dot_blocks<<<(unsigned)((nb * nrhs + 255) / 256), 256, 0, s>>>(nb, nrhs, dots, partial, halt);
}

// out[r] = sum_b partials[r * nb + b]; one block per RHS
__global__ __launch_bounds__(256) void k_reduce_final_rhs(const double *p, int nb, int N, double *out) {
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  const double s = rhs_total(p, nb, r, red);
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

// x[i] = a deterministic pseudo-random complex number in [-1, 1)^2 with
// every mantissa bit random (a 64-bit mix of i, splitmix64): tuner input
// with the bit statistics of real spinors.  All-zero input runs the hop up to
// 9 % faster than Gaussian data on MI355X, and not by the same amount for
// every launch shape, so the tuner must not time on zeros.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
template <typename R> __global__ void k_fill_pattern(long n, typename c2<R>::t *x) {
  using C = typename c2<R>::t;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = mix64(2 * (uint64_t)i + 0x9e3779b97f4a7c15ull), b = mix64(2 * (uint64_t)i + 1);
  x[i] = cmk<C>((R)((double)(int64_t)a * 0x1p-63), (R)((double)(int64_t)b * 0x1p-63));
}
template <typename R> void fill_pattern(long n, void *x, hipStream_t s) {
  using C = typename c2<R>::t;
  if (n > 0) k_fill_pattern<R><<<elem_blocks(n), 256, 0, s>>>(n, (C *)x);
}
template void fill_pattern<double>(long, void *, hipStream_t);
template void fill_pattern<float>(long, void *, hipStream_t);

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
__global__ __launch_bounds__(256) void k_cg_alpha(const double *p, int nb, int N, double *st, const int *halt) {
  if (halt && *halt) return;
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  const double s = rhs_total(p, nb, r, red);
  if (threadIdx.x == 0) {
    st[1 * N + r] = s;
    st[2 * N + r] = (st[5 * N + r] != 0.0 && s != 0.0) ? st[0 * N + r] / s : 0.0;
  }
}
__global__ __launch_bounds__(256) void k_cg_beta(const double *p, int nb, int N, double *st, const int *halt) {
  if (halt && *halt) return;
  __shared__ double red[256 / 64 + 1];
  const int r = blockIdx.x;
  const double s = rhs_total(p, nb, r, red);
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
void cg_alpha(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s, const int *halt) {
  k_cg_alpha<<<nrhs, 256, 0, s>>>(partials, nblocks, nrhs, st, halt);
}
void cg_beta(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s, const int *halt) {
  k_cg_beta<<<nrhs, 256, 0, s>>>(partials, nblocks, nrhs, st, halt);
}

// r -= alpha_r mmp; partial |r|^2 per RHS
template <typename R>
__global__ __launch_bounds__(256) void k_cg_update_r(long n, int N, const double *st, typename c2<R>::t *r,
                                                     const typename c2<R>::t *mmp, double *pre,
                                                     const int *halt) {
  using C = typename c2<R>::t;
  if (halt && *halt) return;
  __shared__ double lds[256];
  const long stride = (long)gridDim.x * blockDim.x;
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const R a = (R)st[2 * N + (int)(i0 % N)];
  double s = 0.0;
  for (long i = i0; i < n; i += stride) {
    C rv = r[i];
    const C mv = mmp[i];
    rv = cmk<C>(rv.x - a * mv.x, rv.y - a * mv.y);
    r[i] = rv;
    s += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < N) pre[(long)threadIdx.x * gridDim.x + blockIdx.x] = seg_sum(lds, blockDim.x, N);
}
template <typename R>
void cg_update_r(long n, int nrhs, const double *st, void *r, const void *mmp, double *partials, hipStream_t s,
                 const int *halt) {
  using C = typename c2<R>::t;
  k_cg_update_r<R><<<kRedBlocks, rhs_block_threads(nrhs), 0, s>>>(n, nrhs, st, (C *)r, (const C *)mmp, partials,
                                                                  halt);
}
template void cg_update_r<double>(long, int, const double *, void *, const void *, double *, hipStream_t,
                                  const int *);
template void cg_update_r<float>(long, int, const double *, void *, const void *, double *, hipStream_t,
                                 const int *);

// x += alpha_r p ; p = beta_r p + r ; t = A p (site-local, parity par: the
// clover inverse blocks A, or diag when A is null; zero on an open boundary).
// SP: storage of p and t.
template <typename R, typename SP>
__global__ __launch_bounds__(kSiteThreads) void k_cg_update_xp(Geom g, int N, int par, const double *st,
                                                              const typename c2<R>::t *r, SP *p,
                                                              typename c2<R>::t *x, const typename c2<R>::t *A,
                                                              double diag, SP *t, int open_bc,
                                                              const int *halt, int stage, const long *roff) {
  using C = typename c2<R>::t;
  if (halt && *halt) return;
  const long t0 = xcd_tile(blockIdx.x, gridDim.x) * kSiteThreads;
  const long tid = t0 + threadIdx.x;
  const long h = (N == 1) ? tid : tid / N;
  const int rr = (int)(tid - h * N);
  // stage: the block's sites s0 .. s0 + ns - 1 get their clover blocks copied
  // to LDS once (slot k of site s0 + i at cl[k * ns + i]), read by the N items
  // of each site
  extern __shared__ __align__(16) char upd_lds[];
  C *cl = (C *)upd_lds;
  const long s0 = t0 / N;
  int ns = 0;
  if (stage) {
    long s1 = (t0 + kSiteThreads - 1) / N;
    if (s1 > g.Vh - 1) s1 = g.Vh - 1;
    ns = s1 >= s0 ? (int)(s1 - s0 + 1) : 0;
    for (int i = threadIdx.x; i < 36 * ns; i += kSiteThreads) {
      const int k = i / ns, si = i - k * ns;
      cl[i] = A[aidx(k, s0 + si, 36)];
    }
    __syncthreads();
  }
  if (h >= g.Vh) return;
  const R a = (R)st[2 * N + rr], b = (R)st[3 * N + rr];
  C pv[12], v[12];
  load12(p, h, rr, N, pv);
  load12(x, h, rr, N, v);
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = cmk<C>(fma(a, pv[k].x, v[k].x), fma(a, pv[k].y, v[k].y));
  store12(x, h, rr, N, v);
  load12(r, h, rr, N, v);
#pragma unroll
  for (int k = 0; k < 12; ++k) pv[k] = cmk<C>(fma(b, pv[k].x, v[k].x), fma(b, pv[k].y, v[k].y));
  store12(p, h, rr, N, pv);
  // the clover acts on p as stored (rounded to SP), as a hop reading p sees it
#pragma unroll
  for (int k = 0; k < 12; ++k) pv[k] = widen(narrow<SP>(pv[k]));
  bool zb = false;
  if (open_bc) {
    int c[4];
    half_coords(g, h, par, c);
    zb = (c[3] == 0 || c[3] == g.L[3] - 1);
  }
  if (zb) {
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<C>(0, 0);
  } else if (A) {
    if (stage) clov_apply_at<R>(cl + (h - s0), (long)ns, pv, v);
    else clov_apply<R>(roff ? A + roff[rr] : A, h, pv, v);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cscale(pv[k], (R)diag);
  }
  store12(t, h, rr, N, v);
}
template <typename R>
void cg_update_xp(const Geom &g, int nrhs, int par, const double *st, const void *r, void *p, void *x,
                  const void *blocks, double diag, void *t, int open_bc, hipStream_t s, int half,
                  const int *halt, const long *roff) {
  using C = typename c2<R>::t;
  const int nb = site_blocks(g.Vh * nrhs);
  // with several right-hand sides sharing one set of clover blocks, the
  // blocks are staged (the items of a site would each load them through the
  // texture path)
  const int stage = (blocks && nrhs >= 4 && !roff) ? 1 : 0;
  const size_t lds = stage ? sizeof(C) * 36 * (size_t)stage_sites(kSiteThreads, nrhs) : 0;
  if constexpr (sizeof(R) == 4) {
    if (half == kS16Half) {
      k_cg_update_xp<R, c2h><<<nb, kSiteThreads, lds, s>>>(g, nrhs, par, st, (const C *)r, (c2h *)p, (C *)x,
                                                           (const C *)blocks, diag, (c2h *)t, open_bc, halt,
                                                           stage, roff);
      return;
    }
    if (half == kS16Bf16) {
      k_cg_update_xp<R, c2b><<<nb, kSiteThreads, lds, s>>>(g, nrhs, par, st, (const C *)r, (c2b *)p, (C *)x,
                                                           (const C *)blocks, diag, (c2b *)t, open_bc, halt,
                                                           stage, roff);
      return;
    }
  }
  k_cg_update_xp<R, C><<<nb, kSiteThreads, lds, s>>>(g, nrhs, par, st, (const C *)r, (C *)p, (C *)x,
                                                     (const C *)blocks, diag, (C *)t, open_bc, halt, stage,
                                                     roff);
}
template void cg_update_xp<double>(const Geom &, int, int, const double *, const void *, void *, void *,
                                   const void *, double, void *, int, hipStream_t, int, const int *, const long *);
template void cg_update_xp<float>(const Geom &, int, int, const double *, const void *, void *, void *,
                                  const void *, double, void *, int, hipStream_t, int, const int *, const long *);

// one wave; lane r handles right-hand side r (nrhs <= 64)
__global__ __launch_bounds__(64) void k_iter_check(int N, int mode, double delta2, double *st, double *dat,
                                                   int *ctl, double *hist) {
  if (ctl[kCtlHalt]) return;
  const int r = threadIdx.x;
  const int k = ctl[kCtlIter];
  if (k >= kMaxChunk) {  // the history holds kMaxChunk iterations
    if (r == 0) ctl[kCtlHalt] = 1;
    return;
  }
  const bool act = r < N && st[5 * N + r] != 0.0;
  bool bad = false, trig = false, left = false;
  if (act) {
    const double cp = st[4 * N + r];
    hist[(long)k * N + r] = cp;
    bad = !isfinite(cp);
    if (mode == 0) {
      if (fabs(cp) <= dat[0 * kMaxRhs + r]) st[5 * N + r] = 0.0;
      else left = true;
    } else {
      double mx = dat[1 * kMaxRhs + r];
      if (cp > mx) mx = cp;
      dat[1 * kMaxRhs + r] = mx;
      const double it = dat[4 * kMaxRhs + r] + 1.0;
      dat[4 * kMaxRhs + r] = it;
      const double sc = dat[3 * kMaxRhs + r];
      trig = cp < delta2 * mx || it >= dat[5 * kMaxRhs + r] ||
             dat[2 * kMaxRhs + r] * cp / (sc * sc) <= dat[0 * kMaxRhs + r];
      left = true;
    }
  }
  const unsigned long long badm = __ballot(bad), trigm = __ballot(trig), leftm = __ballot(left);
  if (r == 0) {
    ctl[kCtlIter] = k + 1;
    if (badm) ctl[kCtlErr] = 1 + __builtin_ctzll(badm);
    if (trigm) ctl[kCtlTrig] = 1;
    if (badm || trigm || !leftm || k + 1 >= ctl[kCtlMax]) ctl[kCtlHalt] = 1;
  }
}
void iter_check(int nrhs, int mode, double delta2, double *st, double *dat, int *ctl, double *hist,
                hipStream_t s) {
  k_iter_check<<<1, 64, 0, s>>>(nrhs, mode, delta2, st, dat, ctl, hist);
}

// ------------------------------------------- 16-bit-storage inner solve
__global__ void k_d2f_scaled(long n, int N, const double *f, const double2 *src, float2 *dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double a = f[i % N];
  dst[i] = make_float2((float)(src[i].x * a), (float)(src[i].y * a));
}
void convert_d2f_scaled(long n, int nrhs, const double *f, const void *src, void *dst, hipStream_t s) {
  if (n > 0) k_d2f_scaled<<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (const double2 *)src, (float2 *)dst);
}
template <typename S> __global__ void k_scale_half(long n, int N, const double *f, S *x) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float a = (float)f[i % N];
  const float2 v = widen(x[i]);
  x[i] = narrow<S>(make_float2(a * v.x, a * v.y));
}
void scale_rhs_half(long n, int nrhs, const double *f, void *x, int half, hipStream_t s) {
  if (n <= 0) return;
  if (half == kS16Bf16)
    k_scale_half<c2b><<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (c2b *)x);
  else
    k_scale_half<c2h><<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (c2h *)x);
}
__global__ void k_acc_f2d_scaled(long n, int N, const double *f, const float2 *lo, double2 *hi) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double a = f[i % N];
  const double2 h = hi[i];
  hi[i] = make_double2(fma((double)lo[i].x, a, h.x), fma((double)lo[i].y, a, h.y));
}
void accumulate_f2d_scaled(long n, int nrhs, const double *f, const void *lo, void *hi, hipStream_t s) {
  if (n > 0) k_acc_f2d_scaled<<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (const float2 *)lo, (double2 *)hi);
}
// ------------------------------------------- fused reliable-update passes
// rl = (float) (bs - mp) (mp null: bs) and per-RHS partials of |bs - mp|^2,
// summed as k_dot_rhs sums |x|^2 (same grid, order and roundings)
__global__ __launch_bounds__(256) void k_resid_d2f(long n, int N, const double2 *bs, const double2 *mp,
                                                   float2 *rl, double *pre) {
  __shared__ double lr[256];
  const long stride = (long)gridDim.x * blockDim.x;
  double sr = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 v = bs[i];
    if (mp) {
      const double2 m = mp[i];
      v = make_double2(v.x - m.x, v.y - m.y);
    }
    rl[i] = make_float2((float)v.x, (float)v.y);
    sr += v.x * v.x + v.y * v.y;
  }
  lr[threadIdx.x] = sr;
  __syncthreads();
  if ((int)threadIdx.x < N) pre[(long)threadIdx.x * gridDim.x + blockIdx.x] = seg_sum(lr, blockDim.x, N);
}
void resid_d2f(long n, int nrhs, const void *bs, const void *mp, void *rl, double *partials, hipStream_t s) {
  k_resid_d2f<<<kRedBlocks, rhs_block_threads(nrhs), 0, s>>>(n, nrhs, (const double2 *)bs, (const double2 *)mp,
                                                             (float2 *)rl, partials);
}

// xh += f[r] xl (f null: 1); xl = 0; out = A xh on parity p (A: FP64 blocks,
// or diag when null; zero on an open boundary)
__global__ __launch_bounds__(kSiteThreads) void k_fold_site_clover(Geom g, int N, int p, const double2 *A,
                                                                  double diag, const double *f, float2 *xl,
                                                                  double2 *xh, double2 *out, int open_bc,
                                                                  const long *roff) {
  const long t = xcd_tile(blockIdx.x, gridDim.x) * kSiteThreads + threadIdx.x;
  const long h = (N == 1) ? t : t / N;
  const int r = (int)(t - h * N);
  if (h >= g.Vh) return;
  const double a = f ? f[r] : 1.0;
  double2 v[12], o[12];
  float2 l[12];
  load12(xh, h, r, N, v);
  load12(xl, h, r, N, l);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    v[k] = make_double2(fma((double)l[k].x, a, v[k].x), fma((double)l[k].y, a, v[k].y));
    l[k] = make_float2(0.0f, 0.0f);
  }
  store12(xh, h, r, N, v);
  store12(xl, h, r, N, l);
  bool zb = false;
  if (open_bc) {
    int c[4];
    half_coords(g, h, p, c);
    zb = (c[3] == 0 || c[3] == g.L[3] - 1);
  }
  if (zb) {
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = make_double2(0.0, 0.0);
  } else if (A) {
    clov_apply<double>(roff ? A + roff[r] : A, h, v, o);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = cscale(v[k], diag);
  }
  store12(out, h, r, N, o);
}
void fold_site_clover(const Geom &g, int nrhs, int p, const void *blocks, double diag, const double *f, void *xl,
                      void *xh, void *out, int open_bc, hipStream_t s, const long *roff) {
  k_fold_site_clover<<<site_blocks(g.Vh * nrhs), kSiteThreads, 0, s>>>(
      g, nrhs, p, (const double2 *)blocks, diag, f, (float2 *)xl, (double2 *)xh, (double2 *)out, open_bc,
      roff);
}

// dst = (float) f[r] * src (FP32 fields)
__global__ void k_scale_copy(long n, int N, const double *f, const float2 *src, float2 *dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float a = (float)f[i % N];
  const float2 v = src[i];
  dst[i] = make_float2(a * v.x, a * v.y);
}
void scale_copy_rhs(long n, int nrhs, const double *f, const void *src, void *dst, hipStream_t s) {
  if (n > 0) k_scale_copy<<<elem_blocks(n), 256, 0, s>>>(n, nrhs, f, (const float2 *)src, (float2 *)dst);
}

template <typename S> __global__ void k_f2h(long n, const float2 *src, S *dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = narrow<S>(src[i]);
}
void convert_f2h(long n, const void *src, void *dst, int half, hipStream_t s) {
  if (n <= 0) return;
  if (half == kS16Bf16)
    k_f2h<c2b><<<elem_blocks(n), 256, 0, s>>>(n, (const float2 *)src, (c2b *)dst);
  else
    k_f2h<c2h><<<elem_blocks(n), 256, 0, s>>>(n, (const float2 *)src, (c2h *)dst);
}

}  // namespace wck
