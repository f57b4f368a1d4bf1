/*
 * ORACLE / CPU BASELINE ONLY.  Not part of the product: tests compare the
 * GPU solves with it where the Python restatement would be slow, and bench.py
 * times it on the host cores as the stand-in for gpt's Grid CPU path (Grid
 * is not available to build).
 *
 * C/OpenMP restatement of the even-odd Wilson-clover operator and CG:
 *   hop   : lib/gpt/qcd/fermion/reference/wilson_clover.py:182-200
 *           out(x) = -sum_mu c_mu/2 [(1-g_mu) U_mu(x) psi(x+mu)
 *                                   + (1+g_mu) U_mu(x-mu)^dag psi(x-mu)]
 *           with the spin projection (1 -+ g_mu) = [h ; -+X^dag h],
 *           h = psi_up -+ X psi_dn, for gpt's chiral basis
 *           (lib/gpt/core/gamma.py:28-42); the adjoint flips the sign of g_mu
 *           (gamma5 hermiticity)
 *   clover: two 6x6 chiral blocks per site, A or A^-1 (wilson_clover.py:202-220)
 *   CG    : lib/gpt/algorithms/inverter/cg.py:48-112 on the eo2_ne normal
 *           operator Mpc^dag Mpc, Mpc = 1 - Meooe A_C^-1 Meooe A_D^-1
 *           (lib/gpt/algorithms/preconditioner/schur_complement_two.py:87-112)
 *
 * Layout: half-lattice index h = (((t Lz + z) Ly + y) Lx + x) / 2 (x
 * collapsed, as gpt's checkerboarded grids); a spinor is [h][spin 4][colour 3]
 * complex double; links are double-stored per output parity: W[p][h][8][3][3]
 * with dir 2mu = U_mu(x), 2mu+1 = U_mu(x-mu)^dag, boundary phases applied.
 *
 * Pinned against the numpy oracle per site (tests/test_oracle_solvers.py), which
 * in turn reproduces the reference's Grid fingerprints.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cd;

typedef struct {
  int L[4];
  long V, Vh;
  cd *W[2];      /* [Vh][8][9] per output parity */
  long *nbr[2];  /* [Vh][8] neighbour half index (other parity) */
  int *tco[2];   /* time coordinate of each half site */
  cd *blk[2];    /* [Vh][2][36] clover blocks, or NULL */
  cd *iblk[2];   /* [Vh][2][36] inverse blocks, or NULL */
  double diag;   /* 1/(2 kappa) when no clover */
  double c[4];   /* c_mu */
  int open_bc;
} wcc;

/* off-diagonal 2x2 blocks X_mu of gamma_mu = [[0, X],[X^dag, 0]] */
static const cd *gx(int mu) {
  static cd X[4][4];
  static int init = 0;
  if (!init) {
    cd x0[4] = {0, I, I, 0};
    cd x1[4] = {0, -1, 1, 0};
    cd x2[4] = {I, 0, 0, -I};
    cd x3[4] = {1, 0, 0, 1};
    memcpy(X[0], x0, sizeof x0);
    memcpy(X[1], x1, sizeof x1);
    memcpy(X[2], x2, sizeof x2);
    memcpy(X[3], x3, sizeof x3);
    init = 1;
  }
  return X[mu];
}

static void coords_of(const wcc *w, long h, int p, int x[4]) {
  int lxh = w->L[0] / 2;
  long r = h;
  int xh = (int)(r % lxh);
  r /= lxh;
  x[1] = (int)(r % w->L[1]);
  r /= w->L[1];
  x[2] = (int)(r % w->L[2]);
  r /= w->L[2];
  x[3] = (int)r;
  x[0] = 2 * xh + ((x[1] + x[2] + x[3] + p) & 1);
}

static long full_index(const wcc *w, const int x[4]) {
  return ((((long)x[3] * w->L[2] + x[2]) * w->L[1] + x[1]) * w->L[0] + x[0]);
}

/* U: [4][V][3][3] natural order (x fastest), phases[4]; c[4] = c_mu.
 * blocks/iblocks: [V][2][6][6] natural order, or NULL (then diag is used). */
wcc *wcc_create(const int L[4], const double *Ure, const double *phases_re, const double *c,
                const double *blocks, const double *iblocks, double diag, int open_bc) {
  wcc *w = (wcc *)calloc(1, sizeof(wcc));
  const cd *U = (const cd *)Ure;
  const cd *ph = (const cd *)phases_re;
  memcpy(w->L, L, sizeof w->L);
  w->V = (long)L[0] * L[1] * L[2] * L[3];
  w->Vh = w->V / 2;
  for (int mu = 0; mu < 4; ++mu) w->c[mu] = c[mu];
  w->diag = diag;
  w->open_bc = open_bc;
  for (int p = 0; p < 2; ++p) {
    w->W[p] = (cd *)malloc(sizeof(cd) * w->Vh * 72);
    w->nbr[p] = (long *)malloc(sizeof(long) * w->Vh * 8);
    w->tco[p] = (int *)malloc(sizeof(int) * w->Vh);
    if (blocks) {
      w->blk[p] = (cd *)malloc(sizeof(cd) * w->Vh * 72);
      w->iblk[p] = (cd *)malloc(sizeof(cd) * w->Vh * 72);
    }
#pragma omp parallel for
    for (long h = 0; h < w->Vh; ++h) {
      int x[4];
      coords_of(w, h, p, x);
      w->tco[p][h] = x[3];
      long xi = full_index(w, x);
      for (int mu = 0; mu < 4; ++mu) {
        int xp[4], xm[4];
        memcpy(xp, x, sizeof xp);
        memcpy(xm, x, sizeof xm);
        xp[mu] = (x[mu] + 1) % L[mu];
        xm[mu] = (x[mu] - 1 + L[mu]) % L[mu];
        w->nbr[p][h * 8 + 2 * mu] = full_index(w, xp) / 2;
        w->nbr[p][h * 8 + 2 * mu + 1] = full_index(w, xm) / 2;
        const cd *uf = U + ((long)mu * w->V + xi) * 9;
        const cd *ub = U + ((long)mu * w->V + full_index(w, xm)) * 9;
        cd pf = (x[mu] == L[mu] - 1) ? ph[mu] : 1.0;
        cd pb = (xm[mu] == L[mu] - 1) ? conj(ph[mu]) : 1.0;
        cd *wf = w->W[p] + (h * 8 + 2 * mu) * 9;
        cd *wb = w->W[p] + (h * 8 + 2 * mu + 1) * 9;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            wf[i * 3 + j] = pf * uf[i * 3 + j];
            wb[i * 3 + j] = pb * conj(ub[j * 3 + i]);
          }
      }
      if (blocks) {
        memcpy(w->blk[p] + h * 72, (const cd *)blocks + xi * 72, sizeof(cd) * 72);
        memcpy(w->iblk[p] + h * 72, (const cd *)iblocks + xi * 72, sizeof(cd) * 72);
      }
    }
  }
  return w;
}

void wcc_destroy(wcc *w) {
  if (!w) return;
  for (int p = 0; p < 2; ++p) {
    free(w->W[p]);
    free(w->nbr[p]);
    free(w->tco[p]);
    free(w->blk[p]);
    free(w->iblk[p]);
  }
  free(w);
}

long wcc_half_volume(const wcc *w) { return w->Vh; }

/* threads of the OpenMP team a parallel loop of this library runs with */
int wcc_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

static inline void zero_open(const wcc *w, int p, long h, cd *o) {
  if (w->open_bc) {
    int t = w->tco[p][h];
    if (t == 0 || t == w->L[3] - 1)
      for (int k = 0; k < 12; ++k) o[k] = 0;
  }
}

/* out (parity 1-pin) = Meooe in (parity pin); dag = 1 for the adjoint. */
void wcc_meooe(const wcc *w, double *outr, const double *inr, int pin, int dag) {
  int p = 1 - pin;
  cd *out = (cd *)outr;
  const cd *in = (const cd *)inr;
#pragma omp parallel for schedule(static)
  for (long h = 0; h < w->Vh; ++h) {
    cd acc[12] = {0};
    for (int mu = 0; mu < 4; ++mu) {
      const cd *X = gx(mu);
      for (int fb = 0; fb < 2; ++fb) {
        /* forward uses (1 - g), backward (1 + g); the adjoint swaps them */
        double s = (fb == 0) ? 1.0 : -1.0;
        if (dag) s = -s;
        const cd *ps = in + w->nbr[p][h * 8 + 2 * mu + fb] * 12;
        const cd *u = w->W[p] + (h * 8 + 2 * mu + fb) * 9;
        cd hs[2][3], ch[2][3];
        for (int c = 0; c < 3; ++c) {
          hs[0][c] = ps[0 * 3 + c] - s * (X[0] * ps[2 * 3 + c] + X[1] * ps[3 * 3 + c]);
          hs[1][c] = ps[1 * 3 + c] - s * (X[2] * ps[2 * 3 + c] + X[3] * ps[3 * 3 + c]);
        }
        for (int a = 0; a < 2; ++a)
          for (int i = 0; i < 3; ++i)
            ch[a][i] = u[i * 3 + 0] * hs[a][0] + u[i * 3 + 1] * hs[a][1] + u[i * 3 + 2] * hs[a][2];
        double cf = -0.5 * w->c[mu];
        for (int c = 0; c < 3; ++c) {
          /* lower = -s X^dag (U h) */
          cd l2 = -s * (conj(X[0]) * ch[0][c] + conj(X[2]) * ch[1][c]);
          cd l3 = -s * (conj(X[1]) * ch[0][c] + conj(X[3]) * ch[1][c]);
          acc[0 * 3 + c] += cf * ch[0][c];
          acc[1 * 3 + c] += cf * ch[1][c];
          acc[2 * 3 + c] += cf * l2;
          acc[3 * 3 + c] += cf * l3;
        }
      }
    }
    memcpy(out + h * 12, acc, sizeof acc);
    zero_open(w, p, h, out + h * 12);
  }
}

/* out = A in (inv = 0) or A^-1 in (inv = 1) on parity p */
void wcc_mooee(const wcc *w, double *outr, const double *inr, int p, int inv) {
  cd *out = (cd *)outr;
  const cd *in = (const cd *)inr;
  const cd *B = inv ? w->iblk[p] : w->blk[p];
  double sc = inv ? 1.0 / w->diag : w->diag;
#pragma omp parallel for schedule(static)
  for (long h = 0; h < w->Vh; ++h) {
    const cd *v = in + h * 12;
    cd *o = out + h * 12;
    if (B) {
      cd r[12];
      for (int b = 0; b < 2; ++b) {
        const cd *m = B + h * 72 + b * 36;
        for (int i = 0; i < 6; ++i) {
          cd a = 0;
          for (int j = 0; j < 6; ++j) a += m[i * 6 + j] * v[b * 6 + j];
          r[b * 6 + i] = a;
        }
      }
      memcpy(o, r, sizeof r);
    } else {
      for (int k = 0; k < 12; ++k) o[k] = sc * v[k];
    }
    zero_open(w, p, h, o);
  }
}

static double dot_re(const cd *a, const cd *b, long n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += creal(a[i]) * creal(b[i]) + cimag(a[i]) * cimag(b[i]);
  return s;
}

/* Mpc (dag = 0) or Mpc^dag (dag = 1) with D = parity pd; t1, t2: scratch of
 * one half field each. */
static void mpc(const wcc *w, cd *out, const cd *in, int pd, int dag, cd *t1, cd *t2) {
  int pc = 1 - pd;
  long n = w->Vh * 12;
  if (!dag) {
    wcc_mooee(w, (double *)t1, (const double *)in, pd, 1);
    wcc_meooe(w, (double *)t2, (const double *)t1, pd, 0);
    wcc_mooee(w, (double *)t1, (const double *)t2, pc, 1);
    wcc_meooe(w, (double *)out, (const double *)t1, pc, 0);
  } else {
    wcc_meooe(w, (double *)t1, (const double *)in, pd, 1);
    wcc_mooee(w, (double *)t2, (const double *)t1, pc, 1);
    wcc_meooe(w, (double *)t1, (const double *)t2, pc, 1);
    wcc_mooee(w, (double *)out, (const double *)t1, pd, 1);
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) out[i] = in[i] - out[i];
}

void wcc_mpc_ne(const wcc *w, double *outr, const double *inr, int pd, double *scratch) {
  long n = w->Vh * 12;
  cd *s = (cd *)scratch;
  mpc(w, s + 2 * n, (const cd *)inr, pd, 0, s, s + n);
  mpc(w, (cd *)outr, s + 2 * n, pd, 1, s, s + n);
}

/* CG on Mpc^dag Mpc (D = parity pd), same recursion as gpt's inverter.
 * x: initial guess in, solution out; returns iterations (negative if not
 * converged); history (optional, maxiter entries) gets |r|^2 per iteration. */
long wcc_cg_ne(const wcc *w, double *xr, const double *br, int pd, double eps, long maxiter,
               double *history) {
  long n = w->Vh * 12;
  cd *x = (cd *)xr;
  const cd *b = (const cd *)br;
  cd *buf = (cd *)malloc(sizeof(cd) * n * 6);
  cd *p = buf, *mmp = buf + n, *r = buf + 2 * n, *scr = buf + 3 * n;
  wcc_mpc_ne(w, (double *)mmp, (const double *)x, pd, (double *)scr);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    r[i] = b[i] - mmp[i];
    p[i] = r[i];
  }
  double cp = dot_re(p, p, n);
  double ssq = dot_re(b, b, n);
  if (ssq == 0.0) {
    memset(x, 0, sizeof(cd) * n);
    free(buf);
    return 0;
  }
  double rsq = eps * eps * ssq;
  long k;
  for (k = 0; k < maxiter; ++k) {
    double c = cp;
    wcc_mpc_ne(w, (double *)mmp, (const double *)p, pd, (double *)scr);
    double d = dot_re(p, mmp, n);
    double a = c / d;
    double cpn = 0.0;
#pragma omp parallel for reduction(+ : cpn) schedule(static)
    for (long i = 0; i < n; ++i) {
      r[i] = r[i] - a * mmp[i];
      cpn += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
    }
    cp = cpn;
    double bb = cp / c;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      x[i] = x[i] + a * p[i];
      p[i] = bb * p[i] + r[i];
    }
    if (history) history[k] = cp;
    if (cp <= rsq) {
      free(buf);
      return k + 1;
    }
  }
  free(buf);
  return -k;
}
