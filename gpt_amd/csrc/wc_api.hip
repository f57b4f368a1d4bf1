// C ABI of the MI355X even-odd Wilson-clover operator (declared in
// include/gpt_amd_wc.h).  Host-side orchestration: operator state, opcode
// dispatch, the Schur-complement compositions and the CG solve loops.
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "wc_internal.h"

thread_local std::string g_err;
std::mutex g_shape_mu;
std::map<ShapeKey, wck::HopShape> g_shape_cache;
std::map<ShapeKey, int> g_mma_cache;
std::map<ShapeKey, wck::HopShape> g_shape16_cache;

// gamma matrices of gpt's chiral basis (lib/gpt/core/gamma.py:28-42)
typedef std::complex<double> cdbl;
static void gamma_mat(int mu, cdbl g[4][4]) {
  const cdbl I(0, 1);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = 0;
  switch (mu) {
    case 0: g[0][3] = I; g[1][2] = I; g[2][1] = -I; g[3][0] = -I; break;
    case 1: g[0][3] = -1; g[1][2] = 1; g[2][1] = 1; g[3][0] = -1; break;
    case 2: g[0][2] = I; g[1][3] = -I; g[2][0] = -I; g[3][1] = I; break;
    default: g[0][2] = 1; g[1][3] = 1; g[2][0] = 1; g[3][1] = 1; break;
  }
}

static wck::CloverParams clover_params(const wc_operator *op) {
  wck::CloverParams cp{};
  const wc_params &p = op->p;
  int pl = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++pl) {
      double c = (nu == 3) ? p.csw_t : p.csw_r / p.xi_0;
      cp.coef[pl] = -0.5 * c;
      cdbl gm[4][4], gn[4][4];
      gamma_mat(mu, gm);
      gamma_mat(nu, gn);
      for (int ch = 0; ch < 2; ++ch)
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            int i = 2 * ch + a, j = 2 * ch + b;
            cdbl s = 0;
            for (int k = 0; k < 4; ++k) s += gm[i][k] * gn[k][j] - gn[i][k] * gm[k][j];
            s *= 0.5;
            cp.sig[pl][ch][a * 2 + b][0] = s.real();
            cp.sig[pl][ch][a * 2 + b][1] = s.imag();
          }
    }
  cp.diag = op->diag;
  cp.csw_t = p.csw_t;
  cp.cF = p.cF;
  cp.open_bc = op->open_bc ? 1 : 0;
  return cp;
}

// The clover-inverse blocks of both parities for another mass than the
// operator's (a slot of a multi-mass batch): the operator's clover term with
// the diagonal of `mass`, built and inverted as build_all builds the
// operator's own, into ai0 / ai1 in the operator's precision (and, if a0 / a1
// are given, the blocks themselves).  Synchronises (checks the pivots).
int build_mass_blocks(const wc_operator *op, double mass, void *ai0, void *ai1, void *a0, void *a1, int *err,
                      hipStream_t s) {
  wck::CloverParams cp = clover_params(op);
  cp.diag = mass + 1.0 + 3.0 * op->p.nu / op->p.xi_0;  // 1/(2 kappa), as set_mass_values
  WC_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  if (op->prec == WC_DOUBLE)
    wck::build_clover<double>(op->g, op->Unat, cp, a0, a1, ai0, ai1, err, s);
  else
    wck::build_clover<float>(op->g, op->Unat, cp, a0, a1, ai0, ai1, err, s);
  WC_LAUNCHED();
  int herr = 0;
  WC_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  if (herr) WC_FAIL(WC_ERR_NUMERIC, "clover block is singular or near-singular at some site (mass %g)", mass);
  return WC_OK;
}

static int build_all(wc_operator *op, hipStream_t s) {
  const Geom &g = op->g;
  if (op->prec == WC_DOUBLE)
    wck::build_links<double>(g, op->Unat, op->nl, op->W[0], op->W[1], s);
  else
    wck::build_links<float>(g, op->Unat, op->nl, op->W[0], op->W[1], s);
  WC_LAUNCHED();
  if (op->has_clover) {
    WC_HIP(hipMemsetAsync(op->err, 0, sizeof(int), s));
    wck::CloverParams cp = clover_params(op);
    if (op->prec == WC_DOUBLE)
      wck::build_clover<double>(g, op->Unat, cp, op->A[0], op->A[1], op->Ai[0], op->Ai[1], op->err, s);
    else
      wck::build_clover<float>(g, op->Unat, cp, op->A[0], op->A[1], op->Ai[0], op->Ai[1], op->err, s);
    WC_LAUNCHED();
    int herr = 0;
    WC_HIP(hipMemcpyAsync(&herr, op->err, sizeof(int), hipMemcpyDeviceToHost, s));
    WC_HIP(hipStreamSynchronize(s));
    if (herr) WC_FAIL(WC_ERR_NUMERIC, "clover block is singular or near-singular at some site (mass %g)", op->mass);
  }
  return WC_OK;
}

static void set_mass_values(wc_operator *op, double mass) {
  op->mass = mass;
  op->diag = mass + 1.0 + 3.0 * op->p.nu / op->p.xi_0;  // 1/(2 kappa)
}

extern "C" {

const char *wc_last_error(void) { return g_err.c_str(); }

int wc_create(const wc_params *p, const void *U, void *stream, wc_operator **out) {
  if (!p || !U || !out) WC_FAIL(WC_ERR_ARG, "wc_create: null argument");
  *out = nullptr;
  for (int mu = 0; mu < 4; ++mu)
    if (p->fdims[mu] < 2 || (p->fdims[mu] & 1))
      WC_FAIL(WC_ERR_ARG, "wc_create: fdims[%d] = %d must be even and >= 2", mu, p->fdims[mu]);
  if (p->precision != WC_DOUBLE && p->precision != WC_SINGLE)
    WC_FAIL(WC_ERR_ARG, "wc_create: unknown precision %d", p->precision);
  if (!(p->xi_0 > 0.0)) WC_FAIL(WC_ERR_ARG, "wc_create: xi_0 must be positive");
  {
    const long V = (long)p->fdims[0] * p->fdims[1] * p->fdims[2] * p->fdims[3];
    if ((V / 2) % 64 != 0)
      WC_FAIL(WC_ERR_ARG, "wc_create: the half volume %ld must be a multiple of 64 (AoSoA blocks)", V / 2);
    const double csz = p->precision == WC_DOUBLE ? 16.0 : 8.0;
    if (12.0 * (V / 2) * csz >= 4294967296.0)
      WC_FAIL(WC_ERR_ARG, "wc_create: a half field of %ld sites exceeds the 4 GB gather window", V / 2);
  }
  const bool open_bc = (p->boundary_phases[3][0] == 0.0 && p->boundary_phases[3][1] == 0.0);
  if (!open_bc && p->cF != 1.0) WC_FAIL(WC_ERR_ARG, "wc_create: cF != 1 needs an open temporal boundary");
  if (open_bc && (p->xi_0 != 1.0 || p->nu != 1.0 || p->csw_r != p->csw_t))
    WC_FAIL(WC_ERR_ARG, "wc_create: open boundary needs the isotropic case with csw_r == csw_t");
  hipStream_t s = (hipStream_t)stream;
  wc_operator *op = new wc_operator();
  op->p = *p;
  op->prec = p->precision;
  op->csz = p->precision == WC_DOUBLE ? 16 : 8;
  op->nl = p->compress_links ? 6 : 9;
  op->open_bc = open_bc;
  op->has_clover = (p->csw_r != 0.0 || p->csw_t != 0.0);
  for (int mu = 0; mu < 4; ++mu) op->g.L[mu] = p->fdims[mu];
  op->g.Lxh = p->fdims[0] / 2;
  op->g.V = (long)p->fdims[0] * p->fdims[1] * p->fdims[2] * p->fdims[3];
  op->g.Vh = op->g.V / 2;
  op->cs = -0.5 * p->nu / p->xi_0;
  op->ct = -0.5;
  {
    const long slice = op->g.Vh / op->g.L[3];
    op->bps = (slice % 64 == 0) ? (int)(slice / 64) : 0;
  }
  set_mass_values(op, p->mass);
  int rc = WC_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == WC_OK) {
      char b[256];
      snprintf(b, sizeof b, "wc_create: %s failed: %s", what, hipGetErrorString(e));
      g_err = b;
      rc = WC_ERR_HIP;
    }
  };
  const size_t Vh = (size_t)op->g.Vh;
  for (int q = 0; q < 2; ++q) {
    chk(wc_dev_malloc(&op->W[q], Vh * 8 * op->nl * op->csz), "wc_dev_malloc(links)");
    if (op->has_clover) {
      chk(wc_dev_malloc(&op->A[q], Vh * 36 * op->csz), "wc_dev_malloc(clover)");
      chk(wc_dev_malloc(&op->Ai[q], Vh * 36 * op->csz), "wc_dev_malloc(clover inverse)");
    }
  }
  chk(wc_dev_malloc(&op->Unat, (size_t)op->g.V * 4 * 9 * 16), "wc_dev_malloc(gauge copy)");
  chk(wc_dev_malloc((void **)&op->err, sizeof(int)), "wc_dev_malloc(flag)");
  if (rc == WC_OK && ensure_red(op, 1, (hipStream_t)stream) != WC_OK) rc = WC_ERR_HIP;
  if (rc != WC_OK) {
    free_op(op);
    return rc;
  }
  if (hipMemcpyAsync(op->Unat, U, (size_t)op->g.V * 4 * 9 * 16, hipMemcpyDefault, s) != hipSuccess) {
    free_op(op);
    WC_FAIL(WC_ERR_HIP, "wc_create: copying the gauge field failed");
  }
  rc = build_all(op, s);
  if (rc != WC_OK) {
    std::string keep = g_err;
    free_op(op);
    g_err = keep;
    return rc;
  }
  *out = op;
  return WC_OK;
}

int wc_update(wc_operator *op, const void *U, void *stream) {
  if (!op || !U) WC_FAIL(WC_ERR_ARG, "wc_update: null argument");
  hipStream_t s = (hipStream_t)stream;
  WC_HIP(hipMemcpyAsync(op->Unat, U, (size_t)op->g.V * 4 * 9 * 16, hipMemcpyDefault, s));
  return build_all(op, s);
}

int wc_set_mass(wc_operator *op, double mass, void *stream) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_set_mass: null operator");
  set_mass_values(op, mass);
  if (!op->has_clover) return WC_OK;
  return build_all(op, (hipStream_t)stream);
}

int wc_destroy(wc_operator *op) {
  free_op(op);
  return WC_OK;
}

long wc_half_volume(const wc_operator *op) { return op ? op->g.Vh : 0; }
int wc_precision(const wc_operator *op) { return op ? op->prec : -1; }

}  // extern "C"

template <typename R>
static int apply_impl(wc_operator *op, int nrhs, int opcode, void *dst, const void *src, int cb, hipStream_t s) {
  const size_t hb = half_bytes(op) * (size_t)nrhs;
  auto half = [&](const void *f, int p) { return (void *)((char *)f + (size_t)p * hb); };
  switch (opcode) {
    case 2001:
    case 2002:
    case 3001:
    case 4001: {  // M, Mdag, Dhop, DhopDag on full fields
      const int dag = (opcode == 2002 || opcode == 4001);
      const bool withdiag = (opcode == 2001 || opcode == 2002);
      for (int p = 0; p < 2; ++p) {
        HopLaunch a = hop_args<R>(op, nrhs, p, dag, withdiag ? wck::EPI_DIAG : 0, half(src, 1 - p), half(dst, p));
        a.diag_clov = op->has_clover ? op->A[p] : nullptr;
        a.y = half(src, p);
        WC_HOP(R, a, s);
      }
      break;
    }
    case 2009:  // Mdiag
      for (int p = 0; p < 2; ++p) site_op<R>(op, nrhs, p, 0, half(src, p), half(dst, p), nullptr, s);
      break;
    case 2010:
    case 2011:
    case 2012:
    case 2013:
    case 2014:
    case 2015:  // Dminus(Dag), Import/Export: identity for Wilson-clover
      WC_HIP(hipMemcpyAsync(dst, src, 2 * hb, hipMemcpyDeviceToDevice, s));
      break;
    case 2003:
    case 2004:
    case 3002:
    case 4002: {
      if (cb != 0 && cb != 1) WC_FAIL(WC_ERR_ARG, "wc_apply: opcode %d needs src_cb 0 or 1", opcode);
      if ((opcode == 3002 || opcode == 4002) && cb != 1)
        WC_FAIL(WC_ERR_ARG, "wc_apply: DhopEO maps odd to even (src_cb must be 1)");
      const int dag = (opcode == 2004 || opcode == 4002);
      HopLaunch a = hop_args<R>(op, nrhs, 1 - cb, dag, 0, src, dst);
      WC_HOP(R, a, s);
      break;
    }
    case 2005:
    case 2006:
    case 2007:
    case 2008:
      if (cb != 0 && cb != 1) WC_FAIL(WC_ERR_ARG, "wc_apply: opcode %d needs src_cb 0 or 1", opcode);
      site_op<R>(op, nrhs, cb, opcode >= 2007, src, dst, nullptr, s);
      break;
    default:
      WC_FAIL(WC_ERR_ARG, "Unknown opcode %d", opcode);
  }
  WC_LAUNCHED();
  return WC_OK;
}

// CG of lib/gpt/algorithms/inverter/cg.py:48-112 on Mpc^dag Mpc, run for nrhs
// independent right-hand sides at once (each its own scalars, the operator
// applied to all of them in one pass).  A right-hand side that converges is
// frozen (alpha = beta = 0 from then on).  iters/conv/r2/b2: nrhs entries
// each (may be null); history: maxiter * nrhs doubles (may be null).
template <typename R>
static int cg_impl(wc_operator *op, int nrhs, int pd, void *x, const void *b, double eps, long maxiter,
                   double *history, long *iters, int *conv, double *r2, double *b2, hipStream_t s) {
  int rc = ensure_tmp(op, 10, nrhs, s);
  if (rc) return rc;
  if ((rc = ensure_red(op, nrhs, s))) return rc;
  const int N = nrhs;
  void *r = op->tmp[3], *p = op->tmp[4], *mmp = op->tmp[5], *aDp = op->tmp[9];
  const long n = op->g.Vh * 12 * N;
  double *st = op->st, *hs = op->host;
  // mmp = A x ; r = b - mmp ; p = r ; c = |p|^2 ; ssq = |b|^2
  if ((rc = mpc_impl<R>(op, N, pd, 2, mmp, x, nullptr, s))) return rc;
  wck::axpy<R>(n, -1.0, 0.0, mmp, b, r, s);
  WC_HIP(hipMemcpyAsync(p, r, half_bytes(op) * N, hipMemcpyDeviceToDevice, s));
  if ((rc = norm2_rhs<R>(op, N, p, st + 0 * N, s))) return rc;  // c
  if ((rc = norm2_rhs<R>(op, N, b, st + 6 * N, s))) return rc;  // |b|^2
  WC_HIP(hipMemcpyAsync(hs, st, sizeof(double) * 7 * N, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  std::vector<double> rsq(N), active(N, 1.0), c0(N);
  int nactive = 0;
  for (int i = 0; i < N; ++i) {
    const double ssq = hs[6 * N + i];
    rsq[i] = eps * eps * ssq;
    c0[i] = hs[i];
    if (iters) iters[i] = 0;
    if (conv) conv[i] = 0;
    if (r2) r2[i] = hs[i];
    if (b2) b2[i] = ssq;
    if (ssq == 0.0) {  // zero source: zero solution, nothing to iterate (cg.py:69-72)
      active[i] = 0.0;
      if (conv) conv[i] = 1;
    } else {
      ++nactive;
    }
  }
  if (nactive < N) {
    for (int i = 0; i < N; ++i) hs[7 * N + i] = active[i];
    WC_HIP(hipMemcpyAsync(st + 7 * N, hs + 7 * N, sizeof(double) * N, hipMemcpyHostToDevice, s));
    wck::scale_rhs<R>(n, N, st + 7 * N, x, s);
  }
  for (int i = 0; i < N; ++i) hs[i] = active[i];
  WC_HIP(hipMemcpyAsync(st + 5 * N, hs, sizeof(double) * N, hipMemcpyHostToDevice, s));
  // A_D^-1 p, kept up to date by the fused x/p update
  site_op<R>(op, N, pd, 1, p, aDp, nullptr, s);
  // Iterations run in chunks (wck::iter_check): the device applies the
  // stopping rule after every iteration and the host replays the residual
  // history of each run, so the result and history do not depend on the
  // chunk lengths.
  IterRun &ir = op->run;
  if ((rc = ensure_run(ir))) return rc;
  for (int i = 0; i < N; ++i) ir.hdat[i] = rsq[i];
  WC_HIP(hipMemcpyAsync(ir.dat, ir.hdat, sizeof(double) * N, hipMemcpyHostToDevice, s));
  std::vector<uintptr_t> key = {0, (uintptr_t)N, (uintptr_t)pd, (uintptr_t)x, (uintptr_t)r, (uintptr_t)p,
                                (uintptr_t)mmp, (uintptr_t)aDp, (uintptr_t)st};
  op_key(op, N, key);
  auto one = [&](hipStream_t q) -> int {
    op->halt_cur = ir.ctl + wck::kCtlHalt;
    const int rc2 = inner_iteration<R>(op, N, pd, r, p, mmp, x, aDp, q);
    op->halt_cur = nullptr;
    if (rc2) return rc2;
    wck::iter_check(N, 0, 0.0, st, ir.dat, ir.ctl, ir.hist, q);
    WC_LAUNCHED();
    return WC_OK;
  };
  const int cap = chunk_cap(op, N, op->chunk);
  std::vector<double> clast(c0);
  long k = 0;
  while (k < maxiter && nactive > 0) {
    // the run ends when the last right-hand side converges
    double need = 0.0;
    for (int i = 0; i < N; ++i) {
      if (active[i] == 0.0) continue;
      const double it = iters_to(clast[i], rsq[i], decay_rate(c0[i], clast[i], k, 0.0));
      need = std::max(need, it < 0.0 ? (double)kFirstRun : it);
    }
    const int K = pow2_floor(std::min(need, (double)(maxiter - k)), cap);
    if ((rc = run_iterations(ir, key, K, N, s, one))) return rc;
    const int done = ir.hctl[wck::kCtlIter];
    for (int j = 0; j < done; ++j) {
      for (int i = 0; i < N; ++i) {
        if (active[i] == 0.0) continue;
        const double cp = ir.hhist[(size_t)j * N + i];
        if (history) history[(k + j) * N + i] = cp;
        if (r2) r2[i] = cp;
        clast[i] = cp;
        if (!std::isfinite(cp)) {
          if (iters) iters[i] = k + j + 1;
          WC_FAIL(WC_ERR_NUMERIC, "CG: residual of right-hand side %d became %g at iteration %ld", i, cp,
                  k + j + 1);
        }
        if (std::fabs(cp) <= rsq[i]) {
          active[i] = 0.0;
          --nactive;
          if (iters) iters[i] = k + j + 1;
          if (conv) conv[i] = 1;
        }
      }
    }
    k += done;
  }
  for (int i = 0; i < N; ++i)
    if (active[i] != 0.0 && iters) iters[i] = k;
  return WC_OK;
}

// dst = L CG(R_ne src) + S src for nrhs right-hand sides (preconditioned.py:32-54)
template <typename R>
static int solve_impl(wc_operator *op, int nrhs, int pd, void *dst, const void *src, double eps, long maxiter,
                      double *history, long *iters, int *conv, double *r2, double *b2, hipStream_t s) {
  return schur_solve<R>(op, nrhs, pd, dst, src, s, [&](void *px, const void *pb) {
    return cg_impl<R>(op, nrhs, pd, px, pb, eps, maxiter, history, iters, conv, r2, b2, s);
  });
}

// Mixed-precision solve with reliable updates for nrhs right-hand sides.
// The inner CG (cg.py's recursion on Mpc^dag Mpc) runs in the inner
// operator's precision; FP64 holds the solution x, the source terms and the
// true residuals.  A reliable update (x += x_inner, r = b - A x recomputed in
// FP64, the search direction kept) happens when a right-hand side's iterated
// |r|^2 has dropped by delta^2 since the last one.  Stopping test, at each
// reliable update: |R src - Mpc y| <= tol |src| -- after the L, S
// reconstruction this is exactly the full system's true residual
// |src - M dst| / |src| (the C rows are solved exactly).
//
// storage16 (FP32 inner operator): the search direction and the Schur
// intermediates are stored as 16-bit floats.  Each right-hand side's inner
// state is then kept in a frame scaled by s_r, chosen at every reliable
// update so that the new residual has unit RMS per component (s_r =
// sqrt(12 Vh / |r_h|^2)); the search direction is carried over by s_new /
// s_old, and the inner solution is folded back with 1 / s.  alpha and beta
// do not change under a common scaling of r and p.
static int mixed_impl(wc_operator *hi, wc_operator *lo, int N, int pd, void *dst, const void *src, double tol,
                      long maxiter, double delta, int storage16, long *iters, int *conv, double *tres, long *nupd,
                      hipStream_t s) {
  const int pc = 1 - pd;
  int rc;
  if ((rc = ensure_tmp(hi, 9, N, s))) return rc;
  if ((rc = ensure_tmp(lo, 8, N, s))) return rc;
  if ((rc = ensure_red(hi, N, s))) return rc;
  if ((rc = ensure_red(lo, N, s))) return rc;
  const size_t hb = half_bytes(hi) * (size_t)N;
  auto half = [&](const void *f, int p) { return (void *)((char *)f + (size_t)p * hb); };
  void *bs = hi->tmp[3], *xh = hi->tmp[4], *rs = hi->tmp[5], *rh = hi->tmp[6], *tC = hi->tmp[7], *mp = hi->tmp[8];
  void *rl = lo->tmp[3], *pl = lo->tmp[4], *ml = lo->tmp[5], *xl = lo->tmp[6], *aDl = lo->tmp[7];
  const long n = hi->g.Vh * 12 * N;
  double *sth = hi->st, *hsh = hi->host, *stl = lo->st, *hsl = lo->host;
  const bool lo_single = lo->prec == WC_SINGLE;
  const int h16 = lo_single ? storage16 : 0;  // kernels' 16-bit storage code
  if (h16) delta = std::max(delta, s16_delta_floor(h16));

  // b_s = R src = src_D - Meooe A_C^-1 src_C ; |src|^2 per RHS
  site_op<double>(hi, N, pc, 1, half(src, pc), tC, nullptr, s);
  {
    HopLaunch a = hop_args<double>(hi, N, pd, 0, wck::EPI_AXPY, tC, bs);
    a.y = half(src, pd);
    WC_HOP(double, a, s);
  }
  // |src|^2 over both parities (a full field is two half fields back to back)
  wck::dot_partials_rhs<double>(2 * n, N, src, src, hi->partials, nullptr, s);
  wck::reduce_final_rhs(hi->partials, wck::kRedBlocks, N, sth + 8 * N, s);
  // guess: y = L^-1 dst = A_D dst_D; with a zero guess (all right-hand
  // sides) the first true residual is b_s itself
  site_op<double>(hi, N, pd, 0, half(dst, pd), xh, nullptr, s);
  if (lo_single) wck::zero<float>(n, xl, s);
  else wck::zero<double>(n, xl, s);
  if ((rc = norm2_rhs<double>(hi, N, xh, sth + 14 * N, s))) return rc;
  WC_HIP(hipMemcpyAsync(hsh + 14 * N, sth + 14 * N, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  bool guess_zero = true;
  for (int i = 0; i < N; ++i) guess_zero = guess_zero && hsh[14 * N + i] == 0.0;
  WC_LAUNCHED();

  // ratio: |r_s|^2 / |r_h|^2 at the last reliable update, to predict from the
  // iterated |r_h|^2 when a right-hand side meets tol (and check it then)
  std::vector<double> src2(N), maxr(N), active(N, 1.0), scale(N, 1.0), ratio(N, 0.0);
  int nactive = N;
  long updates = 0;
  // chunked inner iterations (wck::iter_check mode 1 decides on the device
  // when a reliable update is due); cref / clast: inner |r|^2 at the last
  // update and now, rate: its decay per iteration in the last cycle
  IterRun &ir = lo->run;
  if ((rc = ensure_run(ir))) return rc;
  std::vector<double> cref(N, 0.0), clast(N, 0.0), rate(N, 0.0);
  long since = 0;
  auto upload = [&]() -> int {
    double *d = ir.hdat;
    for (int i = 0; i < N; ++i) {
      d[0 * wck::kMaxRhs + i] = tol * tol * src2[i];
      d[1 * wck::kMaxRhs + i] = maxr[i];
      d[2 * wck::kMaxRhs + i] = ratio[i];
      d[3 * wck::kMaxRhs + i] = scale[i];
      d[4 * wck::kMaxRhs + i] = 0.0;
      d[5 * wck::kMaxRhs + i] = HUGE_VAL;  // no per-RHS iteration limit
    }
    WC_HIP(hipMemcpyAsync(ir.dat, ir.hdat, IterRun::kDatBytes, hipMemcpyHostToDevice, s));
    return WC_OK;
  };
  // scale rows in hi->st: 11 s_new, 12 s_new / s_old, 13 1 / s (for the next fold)
  for (int i = 0; i < N; ++i) hsh[13 * N + i] = 1.0;
  WC_HIP(hipMemcpyAsync(sth + 13 * N, hsh + 13 * N, sizeof(double) * N, hipMemcpyHostToDevice, s));
  // reliable update: y += x_inner / s, x_inner = 0; r_s = b_s - Mpc y,
  // true = |r_s|^2, r_h = Mpc^dag r_s, r_inner = s r_h, c = s^2 |r_h|^2
  auto reliable = [&](long k, bool first) -> int {
    int rc2;
    if (lo_single) {
      // one pass folds x_inner into y and applies A_D^-1 to y (the first
      // piece of Mpc y); r_s = b_s - Mpc y goes straight to FP32 (rl) with
      // its FP64 norm.  Mpc^dag r_s only restarts the inner recursion, which
      // runs in FP32: it is applied by the inner operator (into ml, free
      // between iterations); the true residual and the stopping test stay FP64
      if (first && guess_zero) {
        wck::resid_d2f(n, N, bs, nullptr, rl, hi->partials, s);
      } else {
        void *aD = hi->tmp[0];
        wck::fold_site_clover(hi->g, N, pd, clov_inv(hi, pd), 1.0 / hi->diag,
                              h16 ? sth + 13 * N : nullptr, xl, xh, aD, hi->open_bc ? 1 : 0, s, clov_roff(hi));
        if ((rc2 = mpc_impl<double>(hi, N, pd, 0, mp, xh, nullptr, s, aD))) return rc2;
        wck::resid_d2f(n, N, bs, mp, rl, hi->partials, s);
      }
      wck::reduce_final_rhs(hi->partials, wck::kRedBlocks, N, sth + 9 * N, s);
      if ((rc2 = mpc_impl<float>(lo, N, pd, 1, ml, rl, nullptr, s))) return rc2;
      if ((rc2 = norm2_rhs<float>(lo, N, ml, sth + 10 * N, s))) return rc2;
    } else {
      wck::axpy<double>(n, 1.0, 0.0, xl, xh, xh, s);
      wck::zero<double>(n, xl, s);
      if ((rc2 = mpc_impl<double>(hi, N, pd, 0, mp, xh, nullptr, s))) return rc2;
      wck::axpy<double>(n, -1.0, 0.0, mp, bs, rs, s);
      if ((rc2 = norm2_rhs<double>(hi, N, rs, sth + 9 * N, s))) return rc2;
      if ((rc2 = mpc_impl<double>(hi, N, pd, 1, rh, rs, nullptr, s))) return rc2;
      if ((rc2 = norm2_rhs<double>(hi, N, rh, sth + 10 * N, s))) return rc2;
    }
    WC_HIP(hipMemcpyAsync(hsh + 8 * N, sth + 8 * N, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s));
    WC_HIP(hipStreamSynchronize(s));
    ++updates;
    for (int i = 0; i < N; ++i) {
      src2[i] = hsh[8 * N + i];
      if (active[i] == 0.0) continue;
      const double t2 = hsh[9 * N + i], c = hsh[10 * N + i];
      if (!std::isfinite(t2) || !std::isfinite(c))
        WC_FAIL(WC_ERR_NUMERIC, "mixed CG: true residual of right-hand side %d became %g", i, t2);
      if (tres) tres[i] = src2[i] > 0 ? std::sqrt(t2 / src2[i]) : 0.0;
      ratio[i] = c > 0.0 ? t2 / c : 0.0;
      if (t2 <= tol * tol * src2[i]) {
        active[i] = 0.0;
        --nactive;
        if (iters) iters[i] = k;
        if (conv) conv[i] = 1;
      }
    }
    // the new frame of every right-hand side (1 without 16-bit storage)
    for (int i = 0; i < N; ++i) {
      const double c = hsh[10 * N + i];
      const double sn = (h16 && active[i] != 0.0 && c > 0.0) ? std::sqrt(12.0 * (double)hi->g.Vh / c) : 1.0;
      hsh[11 * N + i] = sn;
      hsh[12 * N + i] = sn / scale[i];
      hsh[13 * N + i] = 1.0 / sn;
      scale[i] = sn;
      maxr[i] = c * sn * sn;
    }
    if (h16) {
      WC_HIP(hipMemcpyAsync(sth + 11 * N, hsh + 11 * N, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
      wck::scale_copy_rhs(n, N, sth + 11 * N, ml, rl, s);
      if (!first) {
        wck::scale_rhs_half(n, N, sth + 12 * N, pl, h16, s);
        wck::scale_rhs_half(n, N, sth + 12 * N, aDl, h16, s);
      }
    } else if (lo_single) {
      WC_HIP(hipMemcpyAsync(rl, ml, half_bytes(lo) * N, hipMemcpyDeviceToDevice, s));
    } else {
      WC_HIP(hipMemcpyAsync(rl, rh, hb, hipMemcpyDeviceToDevice, s));
    }
    // c of the inner recursion (in the new frame) for the still active right-hand sides
    for (int i = 0; i < N; ++i) hsl[i] = active[i] != 0.0 ? maxr[i] : 0.0;
    WC_HIP(hipMemcpyAsync(stl + 0 * N, hsl, sizeof(double) * N, hipMemcpyHostToDevice, s));
    for (int i = 0; i < N; ++i) hsl[5 * N + i] = active[i];
    WC_HIP(hipMemcpyAsync(stl + 5 * N, hsl + 5 * N, sizeof(double) * N, hipMemcpyHostToDevice, s));
    for (int i = 0; i < N; ++i) {
      if (since >= 2) rate[i] = decay_rate(cref[i], clast[i], since, rate[i]);
      cref[i] = clast[i] = maxr[i];
    }
    since = 0;
    return upload();
  };
  for (int i = 0; i < N; ++i) {
    if (iters) iters[i] = 0;
    if (conv) conv[i] = 0;
    if (tres) tres[i] = 0.0;
  }
  if ((rc = reliable(0, true))) return rc;
  // zero sources converge at once (the true residual is 0 <= 0); p = r, A_D^-1 p
  if (h16) {
    wck::convert_f2h(n, rl, pl, h16, s);
    site_op<float>(lo, N, pd, 1, pl, aDl, nullptr, s, h16);
  } else {
    WC_HIP(hipMemcpyAsync(pl, rl, half_bytes(lo) * N, hipMemcpyDeviceToDevice, s));
    if (lo_single) site_op<float>(lo, N, pd, 1, pl, aDl, nullptr, s);
    else site_op<double>(lo, N, pd, 1, pl, aDl, nullptr, s);
  }
  // A reliable update is due when a right-hand side's iterated |r|^2 has
  // fallen below delta^2 times its maximum since the last update, or when
  // it predicts convergence (ratio |r|^2 / s^2 <= tol^2 |src|^2): then the
  // true residual is checked at once rather than at the next delta step.
  const double d2 = delta * delta;
  auto one = [&](hipStream_t q) -> int {
    lo->halt_cur = ir.ctl + wck::kCtlHalt;
    const int rc2 = lo_single ? inner_iteration<float>(lo, N, pd, rl, pl, ml, xl, aDl, q, nullptr, nullptr, h16)
                              : inner_iteration<double>(lo, N, pd, rl, pl, ml, xl, aDl, q);
    lo->halt_cur = nullptr;
    if (rc2) return rc2;
    wck::iter_check(N, 1, d2, stl, ir.dat, ir.ctl, ir.hist, q);
    WC_LAUNCHED();
    return WC_OK;
  };
  std::vector<uintptr_t> key = {1, (uintptr_t)N, (uintptr_t)pd, (uintptr_t)h16, (uintptr_t)rl, (uintptr_t)pl,
                                (uintptr_t)ml, (uintptr_t)xl, (uintptr_t)aDl, (uintptr_t)stl};
  op_key(lo, N, key);
  const int cap = chunk_cap(lo, N, hi->chunk);
  long k = 0;
  while (k < maxiter && nactive > 0) {
    // the run ends at the first right-hand side that triggers an update
    double need = 1e30;
    for (int i = 0; i < N; ++i) {
      if (active[i] == 0.0) continue;
      const double r_ = decay_rate(cref[i], clast[i], since, rate[i]);
      double it = iters_to(clast[i], d2 * maxr[i], r_);
      if (ratio[i] > 0.0) {
        const double it2 = iters_to(clast[i], tol * tol * src2[i] * scale[i] * scale[i] / ratio[i], r_);
        if (it2 >= 0.0 && (it < 0.0 || it2 < it)) it = it2;
      }
      need = std::min(need, it < 0.0 ? (double)kFirstRun : it);
    }
    const int K = pow2_floor(std::min(need, (double)(maxiter - k)), cap);
    if ((rc = run_iterations(ir, key, K, N, s, one))) return rc;
    const int done = ir.hctl[wck::kCtlIter];
    for (int j = 0; j < done; ++j)
      for (int i = 0; i < N; ++i) {
        if (active[i] == 0.0) continue;
        const double cp = ir.hhist[(size_t)j * N + i];
        if (!std::isfinite(cp))
          WC_FAIL(WC_ERR_NUMERIC, "mixed CG: residual of right-hand side %d became %g", i, cp);
        if (cp > maxr[i]) maxr[i] = cp;
        clast[i] = cp;
      }
    k += done;
    since += done;
    // fold the inner solution into the FP64 one, restart its accumulation
    if (ir.hctl[wck::kCtlTrig] && (rc = reliable(k, false))) return rc;
  }
  if (nactive > 0) {
    // not converged within maxiter: fold what the inner solve has and report
    if ((rc = reliable(k, false))) return rc;
    for (int i = 0; i < N; ++i)
      if (active[i] != 0.0 && iters) iters[i] = k;
  }
  if (nupd) *nupd = updates;
  // dst = L y + S src
  site_op<double>(hi, N, pd, 1, xh, half(dst, pd), nullptr, s);
  site_op<double>(hi, N, pc, 1, half(src, pc), tC, nullptr, s);
  {
    HopLaunch a = hop_args<double>(hi, N, pc, 0, wck::EPI_CLOV | wck::EPI_AXPY, half(dst, pd), half(dst, pc));
    a.clov = clov_inv(hi, pc);
    a.y = tC;
    WC_HOP(double, a, s);
  }
  WC_LAUNCHED();
  return WC_OK;
}

// ------------------------------------------------------------- hop tuner
// Times the hop (output parity 0, clover epilogue) of nrhs right-hand sides
// for every valid launch shape and returns the fastest.  Scratch: tmp[0..1];
// the input is a fixed pattern of values in [-1, 1) (no denormals or NaN,
// and not zero: all-zero input runs faster, by an amount that depends on the
// shape), the output is overwritten.
template <typename R>
static int tune_impl(wc_operator *op, int N, int reps, hipStream_t s, wck::HopShape &best, float &best_ms) {
  int rc = ensure_tmp(op, 2, N, s);
  if (rc) return rc;
  const long n = op->g.Vh * 12 * N;
  wck::fill_pattern<R>(n, op->tmp[0], s);
  wck::zero<R>(n, op->tmp[1], s);
  std::vector<wck::HopShape> cand;
  // link staging (ls 1) only pays when a site has several items (more
  // right-hand sides than rb), spinor-block staging (ls 2) is for one
  // right-hand side; the t-marching order (seg > 1) is tried second, on the
  // fastest shape in natural order
  const int occs[] = {0, 2, 4, 8, 16};
  for (int rb = 1; rb <= max_rb_prec(op->prec) && rb <= N; ++rb)
    for (int ls = 0; ls <= 2; ++ls)
      for (int tpb = 1; tpb <= wck::kMaxTpb; tpb *= 2)
        for (int oc : occs) {
          wck::HopShape sh;
          sh.rb = rb;
          sh.tpb = tpb;
          sh.occ = oc;
          sh.ls = ls;
          if ((ls == 1 && N <= rb) || (ls == 2 && N != 1)) continue;
          if (shape_valid(op, sh) && (ls != 2 || stage_fits(op, N, sh))) cand.push_back(sh);
        }
  hipEvent_t e0, e1;
  WC_HIP(hipEventCreate(&e0));
  WC_HIP(hipEventCreate(&e1));
  best_ms = -1.0f;
  auto try_shape = [&](const wck::HopShape &sh) -> int {
    // the hop of the normal operator's first step (clover-inverse epilogue),
    // whose blocks the staging shapes also stage
    HopLaunch a = hop_args<R>(op, N, 0, 0, op->has_clover ? wck::EPI_CLOV : 0, op->tmp[1], op->tmp[0]);
    a.clov = op->has_clover ? op->Ai[0] : nullptr;
    a.shape = sh;
    // shapes this device cannot launch, or whose staging area is too large
    if (!wck::hop<R>(a, s) || hipGetLastError() != hipSuccess) return WC_OK;
    WC_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k) WC_HOP(R, a, s);
    WC_HIP(hipEventRecord(e1, s));
    WC_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    WC_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (best_ms < 0.0f || ms < best_ms) {
      best_ms = ms;
      best = sh;
    }
    return WC_OK;
  };
  for (const wck::HopShape &sh : cand)
    if ((rc = try_shape(sh))) break;
  // t-marching runs of bps/seg blocks: with several right-hand sides a time
  // slice outgrows L2, and shorter runs bring the +-t neighbours of a block
  // back while they are still cached (32^3 x 64, 12 FP64 right-hand sides:
  // 4 runs per slice, 11 % faster than natural order).  The natural-order
  // winner is timed again beside its t-marching variants: its first time is
  // the minimum over many noisy candidates and would be biased low.
  if (!rc && best_ms >= 0.0f && best.ls != 2 && op->bps > 1) {
    const wck::HopShape nat = best;
    best_ms = -1.0f;
    rc = try_shape(nat);
    for (int sg = 2; sg <= 16 && sg <= op->bps && !rc; sg *= 2) {
      wck::HopShape sh = nat;
      sh.seg = sg;
      if (shape_valid(op, sh)) rc = try_shape(sh);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  best_ms /= (float)reps;
  return WC_OK;
}

// ------------------------------------------------- persisted tuner results
// gpt keeps tuned parameters across processes as one JSON file per tuned
// object, .gpt_auto_tune/<sha256(tag)>.json in the working directory, written
// after tuning and read when the object is set up
// (lib/gpt/core/auto_tune.py:20-37,99-136).  The hop tuner does the same:
// one file per (device, lattice, precision, link format, class of
// right-hand-side counts) in the directory WC_TUNE_CACHE (default
// .gpt_amd_auto_tune; empty or 0: nothing read or written), holding
// {"tag": ..., "params": [rb, tpb, occ, seg, ls], "ms": ...}.  A shape read
// back is used only if it is valid for the operator; every shape gives the
// same bits, so a stale file can cost speed, never correctness.
static std::string tune_cache_dir() {
  const char *e = getenv("WC_TUNE_CACHE");
  if (!e) return ".gpt_amd_auto_tune";
  if (!e[0] || (e[0] == '0' && !e[1])) return "";
  return e;
}

// tag and file name: device architecture and CU count, lattice, precision,
// link format, right-hand-side class
static std::string tune_tag(const wc_operator *op, int nrhs) {
  int dev = 0;
  hipDeviceProp_t pr{};
  std::string arch = "unknown";
  int cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess) {
    arch = pr.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    cus = pr.multiProcessorCount;
  }
  char b[256];
  snprintf(b, sizeof b, "hop_%s_cu%d_%dx%dx%dx%d_%s_links%d_rhs%d", arch.c_str(), cus, op->g.L[0], op->g.L[1],
           op->g.L[2], op->g.L[3], op->prec == WC_DOUBLE ? "f64" : "f32", 2 * op->nl, nrhs_class(nrhs));
  std::string t = b;
  for (char &c : t)
    if (!isalnum((unsigned char)c) && c != '_' && c != '-') c = '_';
  return t;
}

// a tuned shape (and, for the 16-bit picks, the form: "form" 0 vector, 1
// matrix cores, -1 not decided) from the file of tag; false when there is
// none or it does not fit the operator
static bool load_tag(const wc_operator *op, int nrhs, const std::string &tag, wck::HopShape &sh, int *form) {
  const std::string dir = tune_cache_dir();
  if (dir.empty()) return false;
  FILE *f = fopen((dir + "/" + tag + ".json").c_str(), "rt");
  if (!f) return false;
  char buf[1024];
  const size_t n = fread(buf, 1, sizeof buf - 1, f);
  fclose(f);
  buf[n] = 0;
  const std::string txt = buf;
  if (txt.find("\"tag\": \"" + tag + "\"") == std::string::npos) return false;
  const size_t at = txt.find("\"params\": [");
  if (at == std::string::npos) return false;
  wck::HopShape x;
  if (sscanf(buf + at, "\"params\": [%d, %d, %d, %d, %d]", &x.rb, &x.tpb, &x.occ, &x.seg, &x.ls) != 5) return false;
  if (!shape_valid(op, x) || x.rb > nrhs || !stage_fits(op, nrhs, x)) return false;
  if (form) {
    const size_t fa = txt.find("\"form\": ");
    if (fa == std::string::npos || sscanf(buf + fa, "\"form\": %d", form) != 1 || *form < -1 || *form > 1)
      return false;
  }
  sh = x;
  return true;
}

static bool load_tuned(const wc_operator *op, int nrhs, wck::HopShape &sh) {
  return load_tag(op, nrhs, tune_tag(op, nrhs), sh, nullptr);
}

static void save_tag(const std::string &tag, const wck::HopShape &sh, float ms, const int *form) {
  const std::string dir = tune_cache_dir();
  if (dir.empty()) return;
  mkdir(dir.c_str(), 0755);  // may exist already
  const std::string fn = dir + "/" + tag + ".json";
  char tmp_name[64];
  snprintf(tmp_name, sizeof tmp_name, ".tmp%ld", (long)getpid());
  const std::string tmp = fn + tmp_name;
  FILE *f = fopen(tmp.c_str(), "wt");
  if (!f) return;  // an unwritable directory only costs the next process a re-tune
  fprintf(f, "{\"tag\": \"%s\", \"params\": [%d, %d, %d, %d, %d], ", tag.c_str(), sh.rb, sh.tpb, sh.occ, sh.seg,
          sh.ls);
  if (form) fprintf(f, "\"form\": %d, ", *form);
  fprintf(f, "\"ms\": %.6g}\n", (double)ms);
  const bool ok = fclose(f) == 0;
  if (!ok || rename(tmp.c_str(), fn.c_str()) != 0) remove(tmp.c_str());
}

static void save_tuned(const wc_operator *op, int nrhs, const wck::HopShape &sh, float ms) {
  save_tag(tune_tag(op, nrhs), sh, ms, nullptr);
}

// the file of the 16-bit picks: the FP32 operator's tag, and whether the
// matrix-core form was a candidate
static std::string s16_tag(const wc_operator *op, int nrhs, bool try_mma) {
  return tune_tag(op, nrhs) + (try_mma ? "_s16_mma" : "_s16");
}

static int tune_hop(wc_operator *op, int nrhs, int reps, hipStream_t s, float *ms_out) {
  wck::HopShape best;
  float ms = 0.0f;
  int rc = op->prec == WC_DOUBLE ? tune_impl<double>(op, nrhs, reps, s, best, ms)
                                 : tune_impl<float>(op, nrhs, reps, s, best, ms);
  if (rc) return rc;
  op->shape[nrhs] = best;
  op->shape_set[nrhs] = true;
  op->shape_user[nrhs] = false;
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape_cache[shape_key(op, nrhs)] = best;
  }
  save_tuned(op, nrhs, best, ms);
  if (ms_out) *ms_out = ms;
  return WC_OK;
}

// Launch shape and form of the hops with 16-bit spinors (FP32 operator),
// once per lattice, link format and class of right-hand-side counts
// (process-wide): the normal operator's first 16-bit hop (clover epilogue,
// FP16 in and out) is timed with the FP32 hop's tuned shape and a few
// others (one or two right-hand sides per thread, 128-512 threads,
// occupancy caps), and, in automatic form
// (op->mma == 2, a multiple of 4 right-hand sides), on the matrix cores.
// The 16-bit hops move a third of the FP32 hops' spinor bytes and are bound
// by the VALUs instead, so their best shape is not always the FP32 one.  All
// candidates give results within the FP16 storage's rounding of each other
// (the vector shapes bitwise the same), so the choice changes speed, not the
// solve's tolerance.  The input is a pseudo-random FP16 pattern (all-zero
// input runs faster, and not by the same amount for every candidate); the
// candidates are timed in two alternating rounds, the better round of each
// counting.  The picks persist like the tuner's shapes (one file per
// lattice, link format, class and whether the matrix-core form was a
// candidate).  Scratch: tmp[0..1].
static int pick_s16(wc_operator *op, int N, hipStream_t s) {
  if (op->prec != WC_SINGLE) return WC_OK;
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    if (g_shape16_cache.count(shape_key(op, N))) return WC_OK;
  }
  const bool try_mma = op->mma == 2 && N >= 4 && N % 4 == 0;
  {
    // an earlier process's picks
    wck::HopShape sh;
    int form = -1;
    if (load_tag(op, N, s16_tag(op, N, try_mma), sh, &form)) {
      std::lock_guard<std::mutex> lk(g_shape_mu);
      g_shape16_cache[shape_key(op, N)] = sh;
      if (try_mma && form >= 0) g_mma_cache[shape_key(op, N)] = form;
      return WC_OK;
    }
  }
  int rc = ensure_tmp(op, 2, N, s);
  if (rc) return rc;
  const long n = op->g.Vh * 12 * N;
  wck::fill_pattern<float>(n, op->tmp[1], s);
  wck::convert_f2h(n, op->tmp[1], op->tmp[0], kS16Half, s);
  wck::zero<float>(n, op->tmp[1], s);
  // candidates: (shape, form)
  std::vector<std::pair<wck::HopShape, int>> cand;
  auto add = [&](const wck::HopShape &sh, int form) {
    if (!shape_valid(op, sh) || !stage_fits(op, N, sh) || sh.rb > N) return;
    for (auto &c : cand)
      if (c.second == form && c.first.rb == sh.rb && c.first.tpb == sh.tpb && c.first.occ == sh.occ &&
          c.first.seg == sh.seg && c.first.ls == sh.ls)
        return;
    cand.push_back({sh, form});
  };
  const wck::HopShape base = shape_for(op, N);
  add(base, 0);
  for (int rb = 1; rb <= 2 && rb <= N; ++rb)
    for (int tpb : {2, 4, 8})
      for (int oc : {0, 4, 8}) {
        wck::HopShape sh;
        sh.rb = rb;
        sh.tpb = tpb;
        sh.occ = oc;
        sh.seg = 1;
        sh.ls = N > rb ? 1 : 0;
        add(sh, 0);
      }
  if (try_mma) add(base, 1);
  hipEvent_t e0, e1;
  WC_HIP(hipEventCreate(&e0));
  WC_HIP(hipEventCreate(&e1));
  std::vector<float> best(cand.size(), -1.0f);
  for (int round = 0; round < 2 && !rc; ++round)
    for (size_t i = 0; i < cand.size() && !rc; ++i) {
      HopLaunch a = hop_args<float>(op, N, 0, 0, op->has_clover ? wck::EPI_CLOV : 0, op->tmp[1], op->tmp[0]);
      a.clov = op->has_clover ? op->Ai[0] : nullptr;
      a.half_in = a.half_out = kS16Half;
      a.shape = cand[i].first;
      a.mma = cand[i].second;
      if (cand[i].second == 1 && !wck::hop_mma_applies(a)) continue;
      // shapes this device cannot launch, or whose staging area is too large
      if (round == 0 && (!wck::hop<float>(a, s) || hipGetLastError() != hipSuccess)) continue;
      if (round == 1 && best[i] < 0.0f) continue;
      auto timed = [&]() -> int {
        WC_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < 3; ++k) WC_HOP(float, a, s);
        WC_HIP(hipEventRecord(e1, s));
        WC_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        WC_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (best[i] < 0.0f || ms < best[i]) best[i] = ms;
        return WC_OK;
      };
      rc = timed();
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  int bv = -1, bm = -1;  // fastest vector shape, fastest candidate overall
  for (size_t i = 0; i < cand.size(); ++i) {
    if (best[i] < 0.0f) continue;
    if (cand[i].second == 0 && (bv < 0 || best[i] < best[bv])) bv = (int)i;
    if (bm < 0 || best[i] < best[bm]) bm = (int)i;
  }
  const wck::HopShape pick = bv >= 0 ? cand[bv].first : base;
  const int form = try_mma && bm >= 0 ? cand[bm].second : -1;
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape16_cache[shape_key(op, N)] = pick;
    if (form >= 0) g_mma_cache[shape_key(op, N)] = form;
  }
  save_tag(s16_tag(op, N, try_mma), pick, bv >= 0 ? best[bv] / 3 : 0.0f, &form);
  return WC_OK;
}

// Solvers tune the hop on first use for lattices where launch shapes matter
// (>= 2^17 site-RHS pairs; WC_AUTOTUNE=0 turns it off), once per lattice,
// precision, link format and class of right-hand-side counts, unless an
// earlier process left its result in the tuner's directory.
int maybe_tune(wc_operator *op, int nrhs, hipStream_t s) {
  const char *env = getenv("WC_AUTOTUNE");
  if (env && env[0] == '0') return WC_OK;
  if (op->g.Vh * (long)nrhs < (1L << 17)) return WC_OK;
  bool have = op->shape_set[nrhs];
  if (!have) {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    have = g_shape_cache.count(shape_key(op, nrhs)) > 0;
  }
  if (have) return pick_s16(op, nrhs, s);
  wck::HopShape sh;
  int rc = WC_OK;
  if (load_tuned(op, nrhs, sh)) {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape_cache[shape_key(op, nrhs)] = sh;
  } else {
    rc = tune_hop(op, nrhs, 3, s, nullptr);
  }
  if (rc == WC_OK) rc = pick_s16(op, nrhs, s);
  return rc;
}

extern "C" {

int wc_tune_hop(wc_operator *op, int nrhs, int reps, int *shape, float *ms_per_call, void *stream) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_tune_hop: null operator");
  int rc = check_nrhs(op, nrhs, "wc_tune_hop");
  if (rc) return rc;
  if (reps < 1) WC_FAIL(WC_ERR_ARG, "wc_tune_hop: reps must be >= 1");
  if ((rc = tune_hop(op, nrhs, reps, (hipStream_t)stream, ms_per_call))) return rc;
  if (shape) {
    shape[0] = op->shape[nrhs].rb;
    shape[1] = op->shape[nrhs].tpb;
    shape[2] = op->shape[nrhs].occ;
    shape[3] = op->shape[nrhs].seg;
    shape[4] = op->shape[nrhs].ls;
  }
  return WC_OK;
}

int wc_set_hop_shape(wc_operator *op, int nrhs, const int *shape) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_set_hop_shape: null operator");
  int rc = check_nrhs(op, nrhs, "wc_set_hop_shape");
  if (rc) return rc;
  if (!shape) {  // back to automatic
    op->shape_set[nrhs] = false;
    op->shape_user[nrhs] = false;
    return WC_OK;
  }
  wck::HopShape sh;
  sh.rb = shape[0];
  sh.tpb = shape[1];
  sh.occ = shape[2];
  sh.seg = shape[3];
  sh.ls = shape[4];
  if (!shape_valid(op, sh) || !stage_fits(op, nrhs, sh))
    WC_FAIL(WC_ERR_ARG,
            "wc_set_hop_shape: rb=%d tpb=%d occ=%d seg=%d ls=%d invalid (rb 1..%d, tpb 1..%d, occ 0..%d, seg 1 or "
            "2..%d, ls 0, 1 or 2 with a staging area of at most %d bytes)",
            sh.rb, sh.tpb, sh.occ, sh.seg, sh.ls, max_rb_prec(op->prec), wck::kMaxTpb, wck::kMaxOcc, op->bps,
            wck::kMaxStageBytes);
  op->shape[nrhs] = sh;
  op->shape_set[nrhs] = true;
  op->shape_user[nrhs] = true;
  return WC_OK;
}

int wc_get_hop_shape(const wc_operator *op, int nrhs, int *shape) {
  if (!op || !shape) WC_FAIL(WC_ERR_ARG, "wc_get_hop_shape: null argument");
  if (nrhs < 1 || nrhs > 64) WC_FAIL(WC_ERR_ARG, "wc_get_hop_shape: nrhs = %d must be in [1, 64]", nrhs);
  const wck::HopShape sh = shape_for(op, nrhs);
  shape[0] = sh.rb;
  shape[1] = sh.tpb;
  shape[2] = sh.occ;
  shape[3] = sh.seg;
  shape[4] = sh.ls;
  return WC_OK;
}

int wc_trim(wc_operator *op) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_trim: null operator");
  wc_device_wait();
  drop_graphs(op->run);
  for (void *q : op->tmp) (void)wc_dev_free_idle(q);
  op->tmp.clear();
  op->tmp_bytes = 0;
  (void)wc_dev_free_idle(op->partials);
  op->partials = nullptr;
  op->npartials = 0;
  (void)wc_dev_free_idle(op->dots);
  op->dots = nullptr;
  op->ndots = 0;
  return WC_OK;
}

int wc_set_iteration_chunk(wc_operator *op, int chunk) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_set_iteration_chunk: null operator");
  if (chunk < 0 || chunk > 32) WC_FAIL(WC_ERR_ARG, "wc_set_iteration_chunk: chunk = %d must be in [0, 32]", chunk);
  op->chunk = chunk;
  return WC_OK;
}

int wc_iteration_runs(const wc_operator *op, long *graph_runs, long *direct_runs) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_iteration_runs: null operator");
  if (graph_runs) *graph_runs = op->run.graph_runs;
  if (direct_runs) *direct_runs = op->run.direct_runs;
  return WC_OK;
}

int wc_apply_block(wc_operator *op, int opcode, int nrhs, void *dst, const void *src, int src_cb, void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_apply: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_apply: dst must not alias src");
  int rc = check_nrhs(op, nrhs, "wc_apply");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  return op->prec == WC_DOUBLE ? apply_impl<double>(op, nrhs, opcode, dst, src, src_cb, s)
                               : apply_impl<float>(op, nrhs, opcode, dst, src, src_cb, s);
}

int wc_apply(wc_operator *op, int opcode, void *dst, const void *src, int src_cb, void *stream) {
  return wc_apply_block(op, opcode, 1, dst, src, src_cb, stream);
}

int wc_apply_mpc_block(wc_operator *op, int parity, int which, int nrhs, void *dst, const void *src,
                       void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc: dst must not alias src");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc: parity must be 0 or 1");
  if (which < 0 || which > 2) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc: which must be 0, 1 or 2");
  int rc = check_nrhs(op, nrhs, "wc_apply_mpc");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  return op->prec == WC_DOUBLE ? mpc_impl<double>(op, nrhs, parity, which, dst, src, nullptr, s)
                               : mpc_impl<float>(op, nrhs, parity, which, dst, src, nullptr, s);
}

int wc_apply_mpc(wc_operator *op, int parity, int which, void *dst, const void *src, void *stream) {
  return wc_apply_mpc_block(op, parity, which, 1, dst, src, stream);
}

int wc_set_hop_mma(wc_operator *op, int mode) {
  if (!op) WC_FAIL(WC_ERR_ARG, "wc_set_hop_mma: null operator");
  if (mode < 0 || mode > 2) WC_FAIL(WC_ERR_ARG, "wc_set_hop_mma: mode %d is not 0, 1 or 2", mode);
  op->mma = mode;
  return WC_OK;
}

int wc_get_hop_mma(const wc_operator *op, int nrhs, int *form) {
  if (!op || !form) WC_FAIL(WC_ERR_ARG, "wc_get_hop_mma: null argument");
  if (nrhs < 1 || nrhs > 64) WC_FAIL(WC_ERR_ARG, "wc_get_hop_mma: nrhs = %d must be in [1, 64]", nrhs);
  HopLaunch a{};
  a.mma = mma_for(op, nrhs);
  a.half_in = a.half_out = kS16Half;
  a.nrhs = nrhs;
  a.g = op->g;
  a.nl = op->nl;
  *form = op->prec == WC_SINGLE && wck::hop_mma_applies(a) ? 1 : 0;
  return WC_OK;
}

// the kernels' 16-bit code of a storage argument of the *_s16 entry points
// (FP32 operator only), in *code
static int s16_code(const wc_operator *op, int storage, const char *fn, int *code) {
  if (op->prec != WC_SINGLE) WC_FAIL(WC_ERR_ARG, "%s: 16-bit storage needs an FP32 operator", fn);
  *code = storage_code(storage);
  if (*code != kS16Half && *code != kS16Bf16)
    WC_FAIL(WC_ERR_ARG, "%s: storage %d is not WC_STORAGE_HALF or WC_STORAGE_BF16", fn, storage);
  return WC_OK;
}

int wc_apply_block_s16(wc_operator *op, int opcode, int nrhs, int storage, void *dst, const void *src, int src_cb,
                       void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_apply_block_s16: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_apply_block_s16: dst must not alias src");
  if (opcode != 2003 && opcode != 2004) WC_FAIL(WC_ERR_ARG, "wc_apply_block_s16: opcode %d is not 2003 or 2004", opcode);
  if (src_cb != 0 && src_cb != 1) WC_FAIL(WC_ERR_ARG, "wc_apply_block_s16: src_cb must be 0 or 1");
  int rc = check_nrhs(op, nrhs, "wc_apply_block_s16");
  if (rc) return rc;
  int code = 0;
  if ((rc = s16_code(op, storage, "wc_apply_block_s16", &code))) return rc;
  hipStream_t s = (hipStream_t)stream;
  HopLaunch a = hop_args<float>(op, nrhs, 1 - src_cb, opcode == 2004, 0, src, dst);
  a.half_in = a.half_out = code;
  a.shape = shape16_for(op, nrhs);
  WC_HOP(float, a, s);
  WC_LAUNCHED();
  return WC_OK;
}

int wc_apply_mpc_block_s16(wc_operator *op, int parity, int nrhs, int storage, void *dst, const void *src,
                           void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_s16: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_s16: dst must not alias src");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_apply_mpc_block_s16: parity must be 0 or 1");
  int rc = check_nrhs(op, nrhs, "wc_apply_mpc_block_s16");
  if (rc) return rc;
  int code = 0;
  if ((rc = s16_code(op, storage, "wc_apply_mpc_block_s16", &code))) return rc;
  // as the inner CG applies it: the last hop also forms Re<src, dst> per
  // 64-site block (unused here)
  return mpc_impl<float>(op, nrhs, parity, 2, dst, src, src, (hipStream_t)stream, nullptr, nullptr, code);
}

int wc_cg_eo2_ne_block(wc_operator *op, int parity, int nrhs, void *x, const void *b, double eps, long maxiter,
                       double *history, long *iterations, int *converged, double *r2, double *b2,
                       void *stream) {
  if (!op || !x || !b) WC_FAIL(WC_ERR_ARG, "wc_cg_eo2_ne: null argument");
  if (x == b) WC_FAIL(WC_ERR_ARG, "wc_cg_eo2_ne: x must not alias b");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_cg_eo2_ne: parity must be 0 or 1");
  if (maxiter < 0) WC_FAIL(WC_ERR_ARG, "wc_cg_eo2_ne: maxiter must be >= 0");
  int rc = check_nrhs(op, nrhs, "wc_cg_eo2_ne");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = maybe_tune(op, nrhs, s))) return rc;
  return op->prec == WC_DOUBLE
             ? cg_impl<double>(op, nrhs, parity, x, b, eps, maxiter, history, iterations, converged, r2, b2, s)
             : cg_impl<float>(op, nrhs, parity, x, b, eps, maxiter, history, iterations, converged, r2, b2, s);
}

int wc_cg_eo2_ne(wc_operator *op, int parity, void *x, const void *b, double eps, long maxiter,
                 double *history, wc_solver_stats *stats, void *stream) {
  long it = 0;
  int cv = 0;
  double r2 = 0, b2 = 0;
  int rc = wc_cg_eo2_ne_block(op, parity, 1, x, b, eps, maxiter, history, &it, &cv, &r2, &b2, stream);
  if (stats) {
    stats->iterations = it;
    stats->converged = cv;
    stats->r2 = r2;
    stats->b2 = b2;
  }
  return rc;
}

int wc_solve_eo2_ne_block(wc_operator *op, int parity, int nrhs, void *dst, const void *src, double eps,
                          long maxiter, double *history, long *iterations, int *converged, double *r2,
                          double *b2, void *stream) {
  if (!op || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne: dst must not alias src");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne: parity must be 0 or 1");
  if (maxiter < 0) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne: maxiter must be >= 0");
  int rc = check_nrhs(op, nrhs, "wc_solve_eo2_ne");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = maybe_tune(op, nrhs, s))) return rc;
  return op->prec == WC_DOUBLE ? solve_impl<double>(op, nrhs, parity, dst, src, eps, maxiter, history, iterations,
                                                    converged, r2, b2, s)
                               : solve_impl<float>(op, nrhs, parity, dst, src, eps, maxiter, history, iterations,
                                                   converged, r2, b2, s);
}

int wc_solve_eo2_ne_mixed(wc_operator *op, wc_operator *inner, int parity, int nrhs, void *dst, const void *src,
                          double tol, long maxiter, double delta, int inner_storage, long *iterations,
                          int *converged, double *true_residual, long *reliable_updates, void *stream) {
  if (!op || !inner || !dst || !src) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: null argument");
  if (dst == src) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: dst must not alias src");
  if (op->prec != WC_DOUBLE) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: the outer operator must be FP64");
  if (op == inner) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: inner must be a separate operator object");
  for (int mu = 0; mu < 4; ++mu)
    if (op->g.L[mu] != inner->g.L[mu]) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: operators of different lattices");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: parity must be 0 or 1");
  if (maxiter < 0 || !(tol > 0.0) || !(delta > 0.0 && delta < 1.0))
    WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: need maxiter >= 0, tol > 0, 0 < delta < 1");
  int rc = check_nrhs(op, nrhs, "wc_solve_eo2_ne_mixed");
  if (rc) return rc;
  if ((rc = maybe_tune(op, nrhs, (hipStream_t)stream))) return rc;
  if ((rc = maybe_tune(inner, nrhs, (hipStream_t)stream))) return rc;
  const int code = storage_code(inner_storage);
  if (code < 0) WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: inner_storage %d unknown", inner_storage);
  if (code && inner->prec != WC_SINGLE)
    WC_FAIL(WC_ERR_ARG, "wc_solve_eo2_ne_mixed: 16-bit storage needs an FP32 inner operator");
  return mixed_impl(op, inner, nrhs, parity, dst, src, tol, maxiter, delta, code, iterations, converged,
                    true_residual, reliable_updates, (hipStream_t)stream);
}

int wc_solve_eo2_ne(wc_operator *op, int parity, void *dst, const void *src, double eps, long maxiter,
                    double *history, wc_solver_stats *stats, void *stream) {
  long it = 0;
  int cv = 0;
  double r2 = 0, b2 = 0;
  int rc = wc_solve_eo2_ne_block(op, parity, 1, dst, src, eps, maxiter, history, &it, &cv, &r2, &b2, stream);
  if (stats) {
    stats->iterations = it;
    stats->converged = cv;
    stats->r2 = r2;
    stats->b2 = b2;
  }
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ BLAS
extern "C" {

static thread_local double *g_red = nullptr;      // device partials + result
static thread_local double *g_red_host = nullptr;  // pinned host copy

static int red_buffers() {
  if (!g_red) {
    WC_HIP(wc_dev_malloc((void **)&g_red, sizeof(double) * (2 * wck::kRedBlocks + 4)));
    WC_HIP(hipHostMalloc((void **)&g_red_host, sizeof(double) * 4, hipHostMallocDefault));
  }
  return WC_OK;
}

#define WC_PREC_CHECK(prec)                                                   \
  if ((prec) != WC_DOUBLE && (prec) != WC_SINGLE) WC_FAIL(WC_ERR_ARG, "unknown precision %d", (int)(prec))

int wc_axpy(int prec, long n, double a_re, double a_im, const void *x, const void *y, void *z, void *stream) {
  WC_PREC_CHECK(prec);
  if (n < 0 || (n > 0 && (!x || !y || !z))) WC_FAIL(WC_ERR_ARG, "wc_axpy: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (prec == WC_DOUBLE) wck::axpy<double>(n, a_re, a_im, x, y, z, s);
  else wck::axpy<float>(n, a_re, a_im, x, y, z, s);
  WC_LAUNCHED();
  return WC_OK;
}

int wc_scale(int prec, long n, double a_re, double a_im, const void *x, void *z, void *stream) {
  WC_PREC_CHECK(prec);
  if (n < 0 || (n > 0 && (!x || !z))) WC_FAIL(WC_ERR_ARG, "wc_scale: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (prec == WC_DOUBLE) wck::scale<double>(n, a_re, a_im, x, z, s);
  else wck::scale<float>(n, a_re, a_im, x, z, s);
  WC_LAUNCHED();
  return WC_OK;
}

int wc_inner_product(int prec, long n, const void *a, const void *b, double *result, void *stream) {
  WC_PREC_CHECK(prec);
  if (n < 0 || !result || (n > 0 && (!a || !b))) WC_FAIL(WC_ERR_ARG, "wc_inner_product: bad arguments");
  int rc = red_buffers();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  double *pre = g_red, *pim = g_red + wck::kRedBlocks, *res = g_red + 2 * wck::kRedBlocks;
  if (prec == WC_DOUBLE) wck::dot_partials<double>(n, a, b, pre, pim, s);
  else wck::dot_partials<float>(n, a, b, pre, pim, s);
  wck::reduce_final(pre, wck::kRedBlocks, res, s);
  wck::reduce_final(pim, wck::kRedBlocks, res + 1, s);
  WC_LAUNCHED();
  WC_HIP(hipMemcpyAsync(g_red_host, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  result[0] = g_red_host[0];
  result[1] = g_red_host[1];
  return WC_OK;
}

int wc_norm2(int prec, long n, const void *x, double *result, void *stream) {
  WC_PREC_CHECK(prec);
  if (n < 0 || !result || (n > 0 && !x)) WC_FAIL(WC_ERR_ARG, "wc_norm2: bad arguments");
  int rc = red_buffers();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  double *pre = g_red, *res = g_red + 2 * wck::kRedBlocks;
  if (prec == WC_DOUBLE) wck::norm2_partials<double>(n, x, pre, s);
  else wck::norm2_partials<float>(n, x, pre, s);
  wck::reduce_final(pre, wck::kRedBlocks, res, s);
  WC_LAUNCHED();
  WC_HIP(hipMemcpyAsync(g_red_host, res, sizeof(double), hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  *result = g_red_host[0];
  return WC_OK;
}

int wc_axpy_norm2(int prec, long n, double a_re, double a_im, const void *x, const void *y, void *z,
                  double *result, void *stream) {
  WC_PREC_CHECK(prec);
  if (n < 0 || !result || (n > 0 && (!x || !y || !z))) WC_FAIL(WC_ERR_ARG, "wc_axpy_norm2: bad arguments");
  int rc = red_buffers();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  double *pre = g_red, *res = g_red + 2 * wck::kRedBlocks;
  if (prec == WC_DOUBLE) wck::axpy_norm2_partials<double>(n, a_re, a_im, x, y, z, pre, s);
  else wck::axpy_norm2_partials<float>(n, a_re, a_im, x, y, z, pre, s);
  wck::reduce_final(pre, wck::kRedBlocks, res, s);
  WC_LAUNCHED();
  WC_HIP(hipMemcpyAsync(g_red_host, res, sizeof(double), hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  *result = g_red_host[0];
  return WC_OK;
}

int wc_convert(int dst_prec, void *dst, int src_prec, const void *src, long n, void *stream) {
  WC_PREC_CHECK(dst_prec);
  WC_PREC_CHECK(src_prec);
  if (n < 0 || (n > 0 && (!dst || !src))) WC_FAIL(WC_ERR_ARG, "wc_convert: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (dst_prec == src_prec) {
    if (dst != src) WC_HIP(hipMemcpyAsync(dst, src, (size_t)n * (dst_prec == WC_DOUBLE ? 16 : 8),
                                          hipMemcpyDeviceToDevice, s));
    return WC_OK;
  }
  if (dst_prec == WC_SINGLE) wck::convert_d2f(n, src, dst, s);
  else wck::convert_f2d(n, src, dst, s);
  WC_LAUNCHED();
  return WC_OK;
}

static int geom_of(const int fdims[4], Geom *g) {
  if (!fdims) WC_FAIL(WC_ERR_ARG, "null fdims");
  for (int mu = 0; mu < 4; ++mu) {
    if (fdims[mu] < 2 || (fdims[mu] & 1)) WC_FAIL(WC_ERR_ARG, "fdims[%d] = %d must be even and >= 2", mu, fdims[mu]);
    g->L[mu] = fdims[mu];
  }
  g->Lxh = fdims[0] / 2;
  g->V = (long)fdims[0] * fdims[1] * fdims[2] * fdims[3];
  g->Vh = g->V / 2;
  if (g->Vh % 64 != 0) WC_FAIL(WC_ERR_ARG, "the half volume %ld must be a multiple of 64", g->Vh);
  return WC_OK;
}

int wc_field_import(const int fdims[4], int prec, const void *natural, void *eo, void *stream) {
  WC_PREC_CHECK(prec);
  Geom g;
  int rc = geom_of(fdims, &g);
  if (rc) return rc;
  if (!natural || !eo || natural == eo) WC_FAIL(WC_ERR_ARG, "wc_field_import: bad pointers");
  hipStream_t s = (hipStream_t)stream;
  if (prec == WC_DOUBLE) wck::import_field<double>(g, natural, eo, s);
  else wck::import_field<float>(g, natural, eo, s);
  WC_LAUNCHED();
  return WC_OK;
}

int wc_gauge_observables(const int fdims[4], const void *U, double *plaquette, double *link_trace, void *stream) {
  Geom g;
  int rc = geom_of(fdims, &g);
  if (rc) return rc;
  if (!U || !plaquette || !link_trace) WC_FAIL(WC_ERR_ARG, "wc_gauge_observables: null argument");
  if ((rc = red_buffers())) return rc;
  hipStream_t s = (hipStream_t)stream;
  double *pp = g_red, *pt = g_red + wck::kRedBlocks, *res = g_red + 2 * wck::kRedBlocks;
  wck::gauge_obs_partials(g, U, pp, pt, s);
  wck::reduce_final(pp, wck::kRedBlocks, res, s);
  wck::reduce_final(pt, wck::kRedBlocks, res + 1, s);
  WC_LAUNCHED();
  WC_HIP(hipMemcpyAsync(g_red_host, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  // gpt's normalisation: averages of Re tr / 3 over the 6 planes and the 4 directions
  *plaquette = g_red_host[0] / (6.0 * 3.0 * (double)g.V);
  *link_trace = g_red_host[1] / (4.0 * 3.0 * (double)g.V);
  return WC_OK;
}

int wc_field_export(const int fdims[4], int prec, const void *eo, void *natural, void *stream) {
  WC_PREC_CHECK(prec);
  Geom g;
  int rc = geom_of(fdims, &g);
  if (rc) return rc;
  if (!natural || !eo || natural == eo) WC_FAIL(WC_ERR_ARG, "wc_field_export: bad pointers");
  hipStream_t s = (hipStream_t)stream;
  if (prec == WC_DOUBLE) wck::export_field<double>(g, eo, natural, s);
  else wck::export_field<float>(g, eo, natural, s);
  WC_LAUNCHED();
  return WC_OK;
}

}  // extern "C"
