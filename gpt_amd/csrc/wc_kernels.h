// Internal launcher interface between the C-ABI (wc_api.hip) and the kernels
// (wc_kernels.hip).  Not exported.
#pragma once
#include <hip/hip_runtime.h>

#include "wc_common.h"

namespace wck {

// hop epilogue flags
enum : int {
  EPI_CLOV = 1,  // res = A res       (A = packed blocks of the output parity)
  EPI_DIAG = 2,  // res = res + D y   (D = packed blocks or the scalar diag)
  EPI_AXPY = 4,  // res = y - res
  EPI_DOT = 8,   // dots[h * nrhs + r] = Re <dotv, res> at site h (then dot_blocks)
};

// Launch shape of the hop: a thread handles rb right-hand sides of a site,
// 64 * tpb threads per workgroup, and with occ > 0 each workgroup reserves
// 1/occ of a CU's LDS so that at most occ workgroups share a CU (a smaller
// window of sites in flight, whose neighbours then stay in L2); seg > 1
// visits the 64-site blocks in t-marching order.  Every shape gives bitwise
// the same result.
struct HopShape {
  int rb = 1;
  int tpb = 1;
  int occ = 0;
  int seg = 1;  // > 1: t-marching order, seg runs per time slice (block_order)
  int ls = 0;   // 1: a workgroup stages its sites' links (and EPI_CLOV blocks) in LDS once;
                // 2 (one right-hand side): it stages the input spinors of its own half-site
                // range, where the +-x and most +-y neighbours are
};
constexpr int kLdsPerCU = 160 * 1024;
constexpr int kMaxStageBytes = 80 * 1024;  // largest LDS staging area of one workgroup
// sites a workgroup of `threads` items covers, G items per site
inline long stage_sites(long threads, long G) { return (threads + G - 1) / G + 1; }
constexpr int kMaxOcc = 32;
constexpr int kMaxTpb = 8;  // up to 512 threads per workgroup
// largest RB instantiated for the precision (beyond it the accumulators and
// one direction's operands exceed 256 VGPRs and spill)
template <typename R> constexpr int max_rb() { return sizeof(R) == 8 ? 1 : 2; }

struct HopLaunch {
  Geom g;
  int nrhs;          // right-hand sides per site (innermost index of every field)
  int pout;          // output parity
  int dag;           // adjoint hop (gamma_mu -> -gamma_mu)
  int nl;            // 9 or 6 complex per link
  int epi;           // EPI_* flags
  const void *in;    // half field of parity 1 - pout
  void *out;         // half field of parity pout
  const void *W;     // links of output parity [8][nl][Vh]
  const void *clov;  // EPI_CLOV blocks [36][Vh] (parity pout), or null
  double clov_scalar;  // EPI_CLOV factor when clov is null (no clover term)
  const void *diag_clov;  // EPI_DIAG blocks, or null for the scalar diag
  double diag;       // scalar diagonal when diag_clov is null
  const void *y;     // EPI_DIAG / EPI_AXPY operand (parity pout)
  const void *dotv;  // EPI_DOT left vector (parity pout)
  double *dots;      // EPI_DOT per-(site, right-hand side) products, Vh * nrhs
  double cs;         // spatial hop coefficient -c_s/2
  double ct;         // temporal hop coefficient -1/2
  double ph[4][2];   // boundary phases
  int open_bc;
  int half_in;       // FP32 hops: in, y and dotv stored as 16-bit (kS16Half c2h / kS16Bf16 c2b)
  int half_out;      // FP32 hops: out stored as the same 16-bit type (0: FP32)
  HopShape shape;    // launch shape (1 <= tpb <= kMaxTpb, 1 <= rb <= max_rb, 0 <= occ <= kMaxOcc)
  int bps;           // 64-site blocks per time slice (0: not a whole number; then seg must be 1)
  const int *halt;   // non-null and *halt != 0: the launch does nothing (chunked iterations)
  int skip_t_faces;  // 1: leave out the t-terms through the links that wrap in t (the interior
                     // part of a T-slab rank's hop; the halo kernels add them)
  int mma;           // 1: FP16-storage hops of 4k right-hand sides run on the matrix cores (k_hop_mma)
  const long *clov_roff; // non-null: right-hand side r uses the EPI_CLOV blocks at clov + clov_roff[r]
                         // (device array, complex elements; a batch of several masses), null: one set
  int big;           // set by the launch: in or out exceeds one buffer descriptor's 4 GB, so
                     // gathers go through a descriptor based at the neighbours' t-slice and
                     // stores through one based at the wave's 64-site block
};

// number of blocks a site launch uses for Vh sites
int site_blocks(long Vh);
// number of 64-site blocks (the per-RHS partial count of dot_blocks)
inline int hop_blocks(const Geom &g) { return (int)(g.Vh / 64); }
// partial[r * (Vh / 64) + b] = sum over the 64 sites of block b, in site
// order, of dots[h * nrhs + r]
void dot_blocks(const Geom &g, int nrhs, const double *dots, double *partial, hipStream_t s,
                const int *halt = nullptr);
#ifndef WC_SITE_THREADS
#define WC_SITE_THREADS 64
#endif
constexpr int kSiteThreads = WC_SITE_THREADS;
constexpr int kRedBlocks = 1024;  // fixed grid of the reduction kernels

// returns false (nothing launched) for a shape outside the limits above
template <typename R> bool hop(const HopLaunch &a, hipStream_t s);
// the matrix-core form of the FP16-storage hop: hop_mma_applies says whether
// the launch has that form (a.mma set, c2h input, N a multiple of 4, ...);
// hop_mma launches it (false: not applicable or epilogue not instantiated)
bool hop_mma_applies(const HopLaunch &a);
bool hop_mma(const HopLaunch &a, hipStream_t s);

// out = A in (blocks != null) or diag * in on parity p, for nrhs right-hand
// sides; zeroes the open boundary; optional axpy (out = y - A in).  half:
// FP32 only, in and out stored as 16-bit (kS16Half / kS16Bf16; y must be null).
template <typename R>
// roff non-null: right-hand side r uses the blocks at blocks + roff[r] (device array, complex elements)
void site_clover(const Geom &g, int nrhs, int p, const void *blocks, double diag, const void *in, void *out,
                 const void *y, int open_bc, hipStream_t s, int half = 0, const long *roff = nullptr);

// Per-right-hand-side reductions over fields whose innermost index is the
// RHS (nrhs of them): partials[r * kRedBlocks + b] for b < kRedBlocks, then
// out[r] = sum over b in fixed order.
template <typename R>
void dot_partials_rhs(long n, int nrhs, const void *a, const void *b, double *pre, double *pim, hipStream_t s);
void reduce_final_rhs(const double *partials, int nblocks, int nrhs, double *out, hipStream_t s);
int rhs_block_threads(int nrhs);
// hi[e] += (double) lo[e] (FP32 -> FP64 accumulation of n entries)
void accumulate_f2d(long n, const void *lo, void *hi, hipStream_t s);
// x = 0 for n complex entries of R (a kernel, so it can sit in a captured graph)
template <typename R> void zero(long n, void *x, hipStream_t s);
template <typename R> void fill_pattern(long n, void *x, hipStream_t s);
// x[e] *= f[e % nrhs] for n entries (per-RHS real factors on the device)
template <typename R> void scale_rhs(long n, int nrhs, const double *f, void *x, hipStream_t s);

// BLAS-1 on n complex entries (C2 of R)
template <typename R>
void axpy(long n, double ar, double ai, const void *x, const void *y, void *z, hipStream_t s);
template <typename R> void scale(long n, double ar, double ai, const void *x, void *z, hipStream_t s);
// partial sums (kRedBlocks entries) of Re/Im <a,b> (2 arrays) or |a|^2 (re only)
template <typename R>
void dot_partials(long n, const void *a, const void *b, double *pre, double *pim, hipStream_t s);
template <typename R> void norm2_partials(long n, const void *a, double *pre, hipStream_t s);
template <typename R>
void axpy_norm2_partials(long n, double ar, double ai, const void *x, const void *y, void *z,
                         double *pre, hipStream_t s);
// out[0] = sum of partials (fixed order, one block)
void reduce_final(const double *partials, int n, double *out, hipStream_t s);

// gauge observables of natural FP64 links U[4][V][3][3] (periodic): per block
// of a fixed grid (kRedBlocks), the sum over sites of Re tr of the 6
// plaquettes and of Re tr of the 4 links
void gauge_obs_partials(const Geom &g, const void *U, double *pplaq, double *ptrace, hipStream_t s);

// precision conversion of n complex entries
void convert_d2f(long n, const void *src, void *dst, hipStream_t s);
void convert_f2d(long n, const void *src, void *dst, hipStream_t s);

// natural [V][12] <-> even-odd [2][12][Vh] (both in R)
template <typename R> void import_field(const Geom &g, const void *nat, void *eo, hipStream_t s);
template <typename R> void export_field(const Geom &g, const void *eo, void *nat, hipStream_t s);

// operator set-up from natural FP64 links Unat[4][V][9]
template <typename R>
void build_links(const Geom &g, const void *Unat, int nl, void *W0, void *W1, hipStream_t s);

struct CloverParams {
  double coef[6];          // -csw_{mu nu}/2 for the planes (01,02,03,12,13,23)
  double sig[6][2][4][2];  // chiral 2x2 blocks of sigma_{mu nu}: [plane][chir][a*2+b][re/im]
  double diag;             // 1/(2 kappa)
  double csw_t;
  double cF;
  int open_bc;
};
// writes packed A and A^-1 per parity; err_flag (device int) gets 1 on a
// near-singular block
template <typename R>
// A0 / A1 may be null (only the inverse wanted)
void build_clover(const Geom &g, const void *Unat, const CloverParams &cp, void *A0, void *A1,
                  void *Ai0, void *Ai1, int *err_flag, hipStream_t s);

// CG on nrhs independent right-hand sides with device scalars:
// st[j * nrhs + r], j = 0 c, 1 d, 2 alpha, 3 beta, 4 cp, 5 active (1/0);
// partials[r * nblocks + b].  halt (optional, here and below): a device
// word; while it is non-zero the launch does nothing.
void cg_alpha(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s, const int *halt = nullptr);
void cg_beta(const double *partials, int nblocks, int nrhs, double *st, hipStream_t s, const int *halt = nullptr);
// r -= alpha_r mmp with per-RHS partials of |r|^2 (kRedBlocks blocks)
template <typename R>
void cg_update_r(long n, int nrhs, const double *st, void *r, const void *mmp, double *partials, hipStream_t s,
                 const int *halt = nullptr);
// x += alpha_r p ; p = beta_r p + r ; t = A p with A the clover-inverse
// blocks of parity par (or diag when null), zero on an open boundary.
// half: FP32 only, p and t stored as 16-bit (kS16Half / kS16Bf16).
template <typename R>
void cg_update_xp(const Geom &g, int nrhs, int par, const double *st, const void *r, void *p, void *x,
                  const void *blocks, double diag, void *t, int open_bc, hipStream_t s, int half = 0,
                  const int *halt = nullptr, const long *roff = nullptr);

// Chunked CG iterations: a run of up to K iterations is enqueued (or
// replayed from a captured graph) without the host looking in between.
// After each iteration iter_check decides on the device, from the iterated
// |r|^2 of every right-hand side (st row 4), whether the run goes on; once
// it sets ctl[kCtlHalt] every later kernel of the run returns at once, so the
// state is exactly the one after the deciding iteration.
enum {
  kCtlHalt = 0,   // set: the rest of the run does nothing
  kCtlIter = 1,   // iterations done in this run
  kCtlMax = 2,    // iterations allowed in this run
  kCtlErr = 3,    // 1 + first right-hand side with a non-finite |r|^2, or 0
  kCtlTrig = 4,   // mode 1: a reliable update is due
  kCtlInts = 16,  // ints of the control block (64 B)
};
// dat rows (kMaxRhs entries each): 0 stop limit (mode 0: eps^2 |b|^2;
// mode 1: tol^2 |src|^2 for the predictor), 1 running max of |r|^2 since the
// last reliable update, 2 ratio |true r|^2 / |iterated r|^2 of that update,
// 3 frame scale s, 4 iterations so far, 5 iteration limit
constexpr int kMaxRhs = 64;
constexpr int kDatRows = 6;
constexpr int kMaxChunk = 64;  // iterations per run (rows of hist)
// mode 0 (CG, cg.py:95-112): an active right-hand side whose |r|^2 <= limit
// is made inactive (st row 5 = 0); the run halts when none is left.
// mode 1 (reliable-update trigger): the run halts when any active right-hand
// side has |r|^2 < delta2 * max, reaches its iteration limit, or is
// predicted to meet its tolerance (ratio |r|^2 / s^2 <= limit).
// Every mode: hist[k * nrhs + r] = |r|^2 of iteration k of the run; a
// non-finite |r|^2 halts with kCtlErr; iteration kCtlMax halts.
void iter_check(int nrhs, int mode, double delta2, double *st, double *dat, int *ctl, double *hist,
                hipStream_t s);

// per-right-hand-side scalings of the 16-bit-storage inner solve (n complex
// entries, RHS innermost): dst = (float) (src * f[r]); x16 *= f[r];
// hi += (double) lo * f[r]; dst16 = src.  half: kS16Half or kS16Bf16.
void convert_d2f_scaled(long n, int nrhs, const double *f, const void *src, void *dst, hipStream_t s);
void scale_rhs_half(long n, int nrhs, const double *f, void *x, int half, hipStream_t s);
void accumulate_f2d_scaled(long n, int nrhs, const double *f, const void *lo, void *hi, hipStream_t s);
void convert_f2h(long n, const void *src, void *dst, int half, hipStream_t s);

// fused passes of the mixed solve's reliable update (FP64 hi, FP32 lo):
// rl = (float) (bs - mp) with per-RHS partials of |bs - mp|^2 for
// reduce_final_rhs (mp null: bs); xh += f[r] xl (f null: 1), xl = 0,
// out = A xh on parity p (blocks FP64 or null for diag; open boundary zero);
// dst = (float) f[r] src.
void resid_d2f(long n, int nrhs, const void *bs, const void *mp, void *rl, double *partials, hipStream_t s);
void fold_site_clover(const Geom &g, int nrhs, int p, const void *blocks, double diag, const double *f, void *xl,
                      void *xh, void *out, int open_bc, hipStream_t s, const long *roff = nullptr);
void scale_copy_rhs(long n, int nrhs, const double *f, const void *src, void *dst, hipStream_t s);

}  // namespace wck
