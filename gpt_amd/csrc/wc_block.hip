// Block CG (lib/gpt/algorithms/inverter/block_cg.py:36-149) on the eo2_ne
// normal operator Mpc^dag Mpc for the N right-hand sides of one multi-RHS
// field (C ABI in include/gpt_amd_wc.h: wc_block_cg_eo2_ne,
// wc_solve_eo2_ne_block_cg).
//
// The block recursion works on N x N complex matrices between the fields: the
// Gram matrices D^dag Z and T^dag T come from one pass over two N-RHS
// fields (k_gram: per-workgroup partial sums in a fixed row order, then a
// fixed-order sum over the workgroups), the small algebra (inverse,
// Cholesky factor, products) runs on the host in FP64, and the field updates
// X += D (M C), T = Q - Z M, Q = T S^-1, D = Q + D S^dag are one pass each
// (k_lincomb: a workgroup holds whole rows -- one (site, component) with
// its N right-hand sides -- so an update may write the field it reads).
#include <hip/hip_runtime.h>

#include <complex>
#include <cstring>
#include <vector>

#include "wc_internal.h"

namespace {

using cplx = std::complex<double>;

constexpr int kGramBlocks = 512;  // fixed grid: the same sums in the same order every time
constexpr int kGramTile = 32;     // rows staged in LDS per step
constexpr int kGramPairs = (wck::kMaxRhs * wck::kMaxRhs + 255) / 256;  // (i, j) pairs per thread

template <typename C> __device__ __forceinline__ double2 to_d2(C v) { return make_double2((double)v.x, (double)v.y); }

// part[b * N * N + i * N + j] = sum over the rows of workgroup b of
// conj(A[row][i]) B[row][j]; rows of N values (right-hand side innermost)
template <typename C>
__global__ __launch_bounds__(256) void k_gram(long rows, int N, const C *A, const C *B, double2 *part) {
  extern __shared__ double2 gsh[];
  double2 *sa = gsh, *sb = gsh + kGramTile * N;
  const int NN = N * N;
  const long r0 = rows * blockIdx.x / gridDim.x, r1 = rows * (blockIdx.x + 1) / gridDim.x;
  double2 acc[kGramPairs];
#pragma unroll
  for (int k = 0; k < kGramPairs; ++k) acc[k] = make_double2(0.0, 0.0);
  for (long t0 = r0; t0 < r1; t0 += kGramTile) {
    const int nt = (int)((r1 - t0) < kGramTile ? (r1 - t0) : kGramTile);
    for (int e = threadIdx.x; e < nt * N; e += blockDim.x) {
      sa[e] = to_d2(A[t0 * N + e]);
      sb[e] = to_d2(B[t0 * N + e]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGramPairs; ++k) {
      const int pr = threadIdx.x + k * 256;
      if (pr < NN) {
        const int i = pr / N, j = pr - i * N;
        double2 s = acc[k];
        for (int t = 0; t < nt; ++t) {
          const double2 a = sa[t * N + i], b = sb[t * N + j];
          s.x = fma(a.x, b.x, fma(a.y, b.y, s.x));
          s.y = fma(a.x, b.y, fma(-a.y, b.x, s.y));
        }
        acc[k] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kGramPairs; ++k) {
    const int pr = threadIdx.x + k * 256;
    if (pr < NN) part[(long)blockIdx.x * NN + pr] = acc[k];
  }
}

// out[e] = sum over the workgroups b in order of part[b * NN + e]
__global__ void k_gram_sum(int NN, int nb, const double2 *part, double2 *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  double2 s = make_double2(0.0, 0.0);
  for (int b = 0; b < nb; ++b) {
    const double2 v = part[(long)b * NN + e];
    s.x += v.x;
    s.y += v.y;
  }
  out[e] = s;
}

// out[row][i] = (Y ? Y[row][i] : 0) + sgn * sum_j X[row][j] M[j * N + i];
// a workgroup holds rpb whole rows (rpb * N threads), staged in LDS first, so
// out may be X or Y
template <typename C>
__global__ __launch_bounds__(256) void k_lincomb(long rows, int N, int rpb, const C *X, const double2 *M,
                                                 const C *Y, double sgn, C *out) {
  extern __shared__ double2 lsh[];
  const long row0 = (long)blockIdx.x * rpb;
  const int nr = (int)((rows - row0) < rpb ? (rows - row0) : rpb);
  for (int e = threadIdx.x; e < nr * N; e += blockDim.x) lsh[e] = to_d2(X[row0 * N + e]);
  __syncthreads();
  const int lr = threadIdx.x / N, i = threadIdx.x - lr * N;
  if (lr >= nr) return;
  const double2 *x = lsh + lr * N;
  double2 s = make_double2(0.0, 0.0);
  for (int j = 0; j < N; ++j) {
    const double2 a = x[j], m = M[j * N + i];
    s.x = fma(a.x, m.x, fma(-a.y, m.y, s.x));
    s.y = fma(a.x, m.y, fma(a.y, m.x, s.y));
  }
  const long o = (row0 + lr) * N + i;
  double2 r = make_double2(sgn * s.x, sgn * s.y);
  if (Y) {
    const double2 y = to_d2(Y[o]);
    r.x += y.x;
    r.y += y.y;
  }
  out[o].x = (decltype(out[o].x))r.x;
  out[o].y = (decltype(out[o].y))r.y;
}

// ------------------------------------------- the same on the matrix cores
// FP64 fields with 8 <= N <= 48 right-hand sides: both are tall-skinny
// complex GEMMs (rows x N by N x N, or its transpose), O(N^2) per row, which
// at N = 48 the vector ALUs ran at about 10 TFLOP/s.  v_mfma_f64_16x16x4f64
// lane layout (checked on MI355X with a probe): operand A[i][k] in lane
// i + 16 k, operand B[k][j] in lane j + 16 k, result D[i][j] in lane
// j + 16 (i % 4), register i / 4.  Every MFMA adds its 4 products in a fixed
// order, so the sums are the same on every run.
typedef double wc_v4d __attribute__((ext_vector_type(4)));
constexpr int kGramBlocksMfma = 2048;  // fixed grid, one wave each

__device__ __forceinline__ wc_v4d mfma64(double a, double b, wc_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// part[b * N * N + i * N + j] = sum over the rows of workgroup b of
// conj(A[row][i]) B[row][j]; T = ceil(N / 16) tiles per dimension.  The rows
// go in steps of 4 (an MFMA's K): lane c + 16 q loads row base + q, RHS
// 16 t + c of each field.
template <int T>
__global__ __launch_bounds__(64) void k_gram_mfma(long rows, int N, const double2 *A, const double2 *B,
                                                  double2 *part) {
  const int l = threadIdx.x, c = l & 15, q = l >> 4;
  const long r0 = rows * blockIdx.x / gridDim.x, r1 = rows * (blockIdx.x + 1) / gridDim.x;
  wc_v4d re[T][T], im[T][T];
#pragma unroll
  for (int ti = 0; ti < T; ++ti)
#pragma unroll
    for (int tj = 0; tj < T; ++tj) {
      re[ti][tj] = wc_v4d{0.0, 0.0, 0.0, 0.0};
      im[ti][tj] = wc_v4d{0.0, 0.0, 0.0, 0.0};
    }
  for (long base = r0; base < r1; base += 4) {
    const long row = base + q;
    double2 a[T], b[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int idx = 16 * t + c;
      const bool ok = row < r1 && idx < N;
      a[t] = ok ? A[row * N + idx] : make_double2(0.0, 0.0);
      b[t] = ok ? B[row * N + idx] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int ti = 0; ti < T; ++ti)
#pragma unroll
      for (int tj = 0; tj < T; ++tj) {
        re[ti][tj] = mfma64(a[ti].x, b[tj].x, re[ti][tj]);
        re[ti][tj] = mfma64(a[ti].y, b[tj].y, re[ti][tj]);
        im[ti][tj] = mfma64(a[ti].x, b[tj].y, im[ti][tj]);
        im[ti][tj] = mfma64(-a[ti].y, b[tj].x, im[ti][tj]);
      }
  }
  const long NN = (long)N * N;
#pragma unroll
  for (int ti = 0; ti < T; ++ti)
#pragma unroll
    for (int tj = 0; tj < T; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * ti + q + 4 * r, j = 16 * tj + c;
        if (i < N && j < N) part[(long)blockIdx.x * NN + (long)i * N + j] = make_double2(re[ti][tj][r], im[ti][tj][r]);
      }
}

// out[row][j] = (Y ? Y[row][j] : 0) + sgn * sum_i X[row][i] M[i * N + j]: a
// workgroup of 4 waves stages M in LDS once and covers kLincombRows rows,
// each wave groups of 16 rows in turn.  A wave reads all X values of a group
// before it writes the group's rows, and no other wave touches them, so out
// may be X or Y.  K step s covers i = 4 s .. 4 s + 3: lane c + 16 q takes
// X[row c][4 s + q] and M[4 s + q][16 t + c]; the result row q + 4 r,
// column 16 t + c.
constexpr int kLincombRows = 512;  // rows per workgroup: 8 groups of 16 per wave

template <int T>
__global__ __launch_bounds__(256) void k_lincomb_mfma(long rows, int N, const double2 *X, const double2 *M,
                                                      const double2 *Y, double sgn, double2 *out) {
  extern __shared__ double2 msh[];
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) msh[e] = M[e];
  __syncthreads();
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, c = l & 15, q = l >> 4;
  const int ks = (N + 3) / 4;
  const long wg0 = (long)blockIdx.x * kLincombRows;
  for (int grp = w; grp < kLincombRows / 16; grp += 4) {
    const long row0 = wg0 + 16L * grp;
    if (row0 >= rows) break;
    const int nr = (int)((rows - row0) < 16 ? (rows - row0) : 16);
    wc_v4d re[T], im[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      re[t] = wc_v4d{0.0, 0.0, 0.0, 0.0};
      im[t] = wc_v4d{0.0, 0.0, 0.0, 0.0};
    }
    for (int st = 0; st < ks; ++st) {
      const int i = 4 * st + q;
      const double2 x = (i < N && c < nr) ? X[(row0 + c) * N + i] : make_double2(0.0, 0.0);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int j = 16 * t + c;
        const double2 m = (i < N && j < N) ? msh[i * N + j] : make_double2(0.0, 0.0);
        re[t] = mfma64(x.x, m.x, re[t]);
        re[t] = mfma64(-x.y, m.y, re[t]);
        im[t] = mfma64(x.x, m.y, im[t]);
        im[t] = mfma64(x.y, m.x, im[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = q + 4 * r, j = 16 * t + c;
        if (rr < nr && j < N) {
          const long o = (row0 + rr) * N + j;
          double2 v = make_double2(sgn * re[t][r], sgn * im[t][r]);
          if (Y) {
            const double2 y = Y[o];
            v.x += y.x;
            v.y += y.y;
          }
          out[o] = v;
        }
      }
  }
}

// whether the block recursion's Gram and update kernels run on the matrix cores
template <typename R> inline bool block_mfma(int N) { return sizeof(R) == 8 && N >= 8 && N <= 48; }

// ----------------------------------------------------- host linear algebra
// inverse by Gauss-Jordan elimination with partial pivoting; false if singular
bool inverse(int n, const std::vector<cplx> &a, std::vector<cplx> &inv) {
  std::vector<cplx> m(a);
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(m[(size_t)r * n + c]) > std::abs(m[(size_t)p * n + c])) p = r;
    if (!(std::abs(m[(size_t)p * n + c]) > 0.0)) return false;
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(m[(size_t)p * n + k], m[(size_t)c * n + k]);
        std::swap(inv[(size_t)p * n + k], inv[(size_t)c * n + k]);
      }
    const cplx d = 1.0 / m[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      m[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const cplx f = m[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        m[(size_t)r * n + k] -= f * m[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// G = L L^dag for Hermitian positive definite G (Hermitised first); returns
// C = L^dag (upper triangular), false if G is not positive definite
bool cholesky_upper(int n, const std::vector<cplx> &g, std::vector<cplx> &C) {
  std::vector<cplx> h((size_t)n * n), L((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) h[(size_t)i * n + j] = 0.5 * (g[(size_t)i * n + j] + std::conj(g[(size_t)j * n + i]));
  for (int j = 0; j < n; ++j) {
    double d = h[(size_t)j * n + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(L[(size_t)j * n + k]);
    if (!(d > 0.0)) return false;
    const double ljj = std::sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      cplx s = h[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * std::conj(L[(size_t)j * n + k]);
      L[(size_t)i * n + j] = s / ljj;
    }
  }
  C.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) C[(size_t)i * n + j] = std::conj(L[(size_t)j * n + i]);
  return true;
}

std::vector<cplx> matmul(int n, const std::vector<cplx> &a, const std::vector<cplx> &b) {
  std::vector<cplx> c((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cplx x = a[(size_t)i * n + k];
      for (int j = 0; j < n; ++j) c[(size_t)i * n + j] += x * b[(size_t)k * n + j];
    }
  return c;
}

std::vector<cplx> adjoint(int n, const std::vector<cplx> &a) {
  std::vector<cplx> c((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) c[(size_t)i * n + j] = std::conj(a[(size_t)j * n + i]);
  return c;
}

// device and pinned host space for the N x N matrices of a block solve: the
// Gram result and kSlots coefficient matrices used in turn (at most two
// updates run between host synchronisations, so a slot is never rewritten
// while a copy from it may be pending)
constexpr int kSlots = 4;
struct BlockScratch {
  double2 *part = nullptr, *dmat = nullptr;  // Gram partials; (1 + kSlots) N x N matrices
  cplx *host = nullptr;                      // pinned, (1 + kSlots) N x N
  int next = 0;
  hipStream_t s = nullptr;  // the stream the solve ran on
  ~BlockScratch() {
    (void)wc_dev_free_on(part, s);
    (void)wc_dev_free_on(dmat, s);
    if (host) (void)wc_host_free(host);
  }
};

// Block CG on Mpc^dag Mpc with D = pd for the N right-hand sides in b; x
// holds the guess and gets the solution.  Scratch fields: tmp[0..2] (the
// normal operator), tmp[3], tmp[4], tmp[5], tmp[9].  history (optional,
// maxiter entries): the largest relative |r|^2 of the block per iteration.
template <typename R>
int block_cg_impl(wc_operator *op, int N, int pd, void *x, const void *b, double eps, double eps_abs,
                  long miniter, long maxiter, long *iterations, int *converged, double *history, hipStream_t s) {
  using C = typename c2<R>::t;
  int rc;
  if ((rc = ensure_tmp(op, 10, N, s))) return rc;
  if ((rc = ensure_red(op, N, s))) return rc;
  const long rows = op->g.Vh * 12;
  const long n = rows * N;
  void *T = op->tmp[3], *Q = op->tmp[4], *D = op->tmp[5], *Z = op->tmp[9];
  BlockScratch bs;
  bs.s = s;
  const int NN = N * N;
  const bool mf = block_mfma<R>(N);
  const int ntile = (N + 15) / 16;  // 16 x 16 tiles per dimension
  const int nparts = mf ? kGramBlocksMfma : kGramBlocks;
  WC_HIP(wc_dev_malloc((void **)&bs.part, sizeof(double2) * (size_t)nparts * NN));
  WC_HIP(wc_dev_malloc((void **)&bs.dmat, sizeof(double2) * (size_t)NN * (1 + kSlots)));
  WC_HIP(hipHostMalloc((void **)&bs.host, sizeof(cplx) * (size_t)NN * (1 + kSlots), hipHostMallocDefault));
  *iterations = 0;
  *converged = 0;

  // N x N Gram matrix of two fields to the host, G[i][j] = <A_i, B_j>
  auto gram = [&](const void *A, const void *B, std::vector<cplx> &G) -> int {
    if (mf) {
      const double2 *a2 = (const double2 *)A, *b2 = (const double2 *)B;
      if (ntile == 1) k_gram_mfma<1><<<kGramBlocksMfma, 64, 0, s>>>(rows, N, a2, b2, bs.part);
      else if (ntile == 2) k_gram_mfma<2><<<kGramBlocksMfma, 64, 0, s>>>(rows, N, a2, b2, bs.part);
      else k_gram_mfma<3><<<kGramBlocksMfma, 64, 0, s>>>(rows, N, a2, b2, bs.part);
    } else {
      const size_t lds = sizeof(double2) * 2 * kGramTile * N;
      k_gram<C><<<kGramBlocks, 256, lds, s>>>(rows, N, (const C *)A, (const C *)B, bs.part);
    }
    k_gram_sum<<<(NN + 255) / 256, 256, 0, s>>>(NN, nparts, bs.part, bs.dmat);
    WC_LAUNCHED();
    WC_HIP(hipMemcpyAsync(bs.host, bs.dmat, sizeof(cplx) * NN, hipMemcpyDeviceToHost, s));
    WC_HIP(hipStreamSynchronize(s));
    G.assign(bs.host, bs.host + NN);
    return WC_OK;
  };
  // out = Y + sgn X M (rows of N right-hand sides; M row-major, coefficient
  // of X_j in out_i at M[j][i])
  auto lincomb = [&](const void *X, const std::vector<cplx> &M, const void *Y, double sgn, void *out) -> int {
    const size_t slot = (size_t)NN * (1 + bs.next);
    bs.next = (bs.next + 1) % kSlots;
    memcpy(bs.host + slot, M.data(), sizeof(cplx) * NN);
    WC_HIP(hipMemcpyAsync(bs.dmat + slot, bs.host + slot, sizeof(cplx) * NN, hipMemcpyHostToDevice, s));
    if (mf) {
      const unsigned grid = (unsigned)((rows + kLincombRows - 1) / kLincombRows);
      const size_t lds = sizeof(double2) * (size_t)N * N;
      const double2 *x2 = (const double2 *)X, *y2 = (const double2 *)Y, *m2 = bs.dmat + slot;
      double2 *o2 = (double2 *)out;
      if (ntile == 1) k_lincomb_mfma<1><<<grid, 256, lds, s>>>(rows, N, x2, m2, y2, sgn, o2);
      else if (ntile == 2) k_lincomb_mfma<2><<<grid, 256, lds, s>>>(rows, N, x2, m2, y2, sgn, o2);
      else k_lincomb_mfma<3><<<grid, 256, lds, s>>>(rows, N, x2, m2, y2, sgn, o2);
    } else {
      const int rpb = 256 / N;
      const unsigned grid = (unsigned)((rows + rpb - 1) / rpb);
      k_lincomb<C><<<grid, rpb * N, sizeof(double2) * rpb * N, s>>>(rows, N, rpb, (const C *)X, bs.dmat + slot,
                                                                   (const C *)Y, sgn, (C *)out);
    }
    WC_LAUNCHED();
    return WC_OK;
  };
  // Q C = R (block_cg.py:24-32): C from the Cholesky factor of R^dag R, Q = R C^-1
  auto qc_is_r = [&](void *Qo, const void *Rin, std::vector<cplx> &Cm) -> int {
    std::vector<cplx> G, Ci;
    int rc2;
    if ((rc2 = gram(Rin, Rin, G))) return rc2;
    if (!cholesky_upper(N, G, Cm) || !inverse(N, Cm, Ci))
      WC_FAIL(WC_ERR_NUMERIC,
              "block CG: the residual block is not of full rank (zero or linearly dependent right-hand sides, "
              "or right-hand sides converged at different rates)");
    return lincomb(Rin, Ci, nullptr, 1.0, Qo);
  };

  // |b_i|^2; an all-zero block is solved by x = 0 (block_cg.py:60-64)
  if ((rc = norm2_rhs<R>(op, N, b, op->st, s))) return rc;
  WC_HIP(hipMemcpyAsync(op->host, op->st, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  std::vector<double> ssq(op->host, op->host + N);
  double tot = 0.0;
  for (double v : ssq) tot += v;
  if (tot == 0.0) {
    wck::zero<R>(n, x, s);
    WC_LAUNCHED();
    *converged = 1;
    return WC_OK;
  }
  // T = B - A X ; Q C = T ; D = Q
  if ((rc = mpc_impl<R>(op, N, pd, 2, Z, x, nullptr, s))) return rc;
  wck::axpy<R>(n, -1.0, 0.0, Z, b, T, s);
  WC_LAUNCHED();
  std::vector<cplx> Cm, S, G, Mi;
  if ((rc = qc_is_r(Q, T, Cm))) return rc;
  WC_HIP(hipMemcpyAsync(D, Q, half_bytes(op) * N, hipMemcpyDeviceToDevice, s));
  for (long k = 0; k < maxiter; ++k) {
    // Z = A D ; M = (D^dag Z)^-1
    if ((rc = mpc_impl<R>(op, N, pd, 2, Z, D, nullptr, s))) return rc;
    if ((rc = gram(D, Z, G))) return rc;
    if (!inverse(N, G, Mi)) WC_FAIL(WC_ERR_NUMERIC, "block CG: D^dag A D is singular at iteration %ld", k + 1);
    // X += D (M C) ; T = Q - Z M ; Q S = T ; D = Q + D S^dag ; C = S C
    if ((rc = lincomb(D, matmul(N, Mi, Cm), x, 1.0, x))) return rc;
    if ((rc = lincomb(Z, Mi, Q, -1.0, T))) return rc;
    if ((rc = qc_is_r(Q, T, S))) return rc;
    if ((rc = lincomb(D, adjoint(N, S), Q, 1.0, D))) return rc;
    Cm = matmul(N, S, Cm);
    // |r_b|^2 = (C^dag C)_bb
    double rel = 0.0, ab = 0.0;
    for (int bb = 0; bb < N; ++bb) {
      double rr = 0.0;
      for (int i = 0; i < N; ++i) rr += std::norm(Cm[(size_t)i * N + bb]);
      ab = std::max(ab, rr);
      if (ssq[bb] > 0.0) rel = std::max(rel, rr / ssq[bb]);
    }
    if (!std::isfinite(rel)) WC_FAIL(WC_ERR_NUMERIC, "block CG: residual became %g at iteration %ld", rel, k + 1);
    if (history) history[k] = rel;
    *iterations = k + 1;
    if (k + 1 >= miniter && ((eps_abs > 0.0 && ab <= eps_abs * eps_abs) || rel <= eps * eps)) {
      *converged = 1;
      break;
    }
  }
  WC_HIP(hipStreamSynchronize(s));
  return WC_OK;
}

int check_block_args(wc_operator *op, int parity, int nrhs, const void *a, const void *b, long maxiter,
                     const char *fn) {
  if (!op || !a || !b) WC_FAIL(WC_ERR_ARG, "%s: null argument", fn);
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "%s: parity must be 0 or 1", fn);
  if (maxiter < 0) WC_FAIL(WC_ERR_ARG, "%s: maxiter must be >= 0", fn);
  return check_nrhs(op, nrhs, fn);
}

}  // namespace

extern "C" {

int wc_block_cg_eo2_ne(wc_operator *op, int parity, int nrhs, void *x, const void *b, double eps, double eps_abs,
                       long miniter, long maxiter, long *iterations, int *converged, double *history,
                       void *stream) {
  int rc = check_block_args(op, parity, nrhs, x, b, maxiter, "wc_block_cg_eo2_ne");
  if (rc) return rc;
  long it = 0;
  int cv = 0;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = maybe_tune(op, nrhs, s))) return rc;
  rc = op->prec == WC_DOUBLE
           ? block_cg_impl<double>(op, nrhs, parity, x, b, eps, eps_abs, miniter, maxiter, &it, &cv, history, s)
           : block_cg_impl<float>(op, nrhs, parity, x, b, eps, eps_abs, miniter, maxiter, &it, &cv, history, s);
  if (iterations) *iterations = it;
  if (converged) *converged = cv;
  return rc;
}

int wc_solve_eo2_ne_block_cg(wc_operator *op, int parity, int nrhs, void *dst, const void *src, double eps,
                             double eps_abs, long miniter, long maxiter, long *iterations, int *converged,
                             double *history, void *stream) {
  int rc = check_block_args(op, parity, nrhs, dst, src, maxiter, "wc_solve_eo2_ne_block_cg");
  if (rc) return rc;
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_block_cg: dst must not alias src");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = maybe_tune(op, nrhs, s))) return rc;
  long it = 0;
  int cv = 0;
  auto inner = [&](auto tag, void *px, const void *pb) {
    using R = decltype(tag);
    return block_cg_impl<R>(op, nrhs, parity, px, pb, eps, eps_abs, miniter, maxiter, &it, &cv, history, s);
  };
  rc = op->prec == WC_DOUBLE ? schur_solve<double>(op, nrhs, parity, dst, src, s,
                                                   [&](void *px, const void *pb) { return inner(0.0, px, pb); })
                             : schur_solve<float>(op, nrhs, parity, dst, src, s,
                                                  [&](void *px, const void *pb) { return inner(0.0f, px, pb); });
  if (iterations) *iterations = it;
  if (converged) *converged = cv;
  return rc;
}

}  // extern "C"
