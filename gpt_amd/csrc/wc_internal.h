// Internal host-side state and compositions shared by the C-ABI sources
// (wc_api.hip, wc_batch.hip).  Not exported.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gpt_amd_wc.h"
#include "wc_kernels.h"

using wck::HopLaunch;

extern thread_local std::string g_err;  // message of the last failure (wc_api.hip)

// device memory of the library's buffers: hipMalloc / hipFree behind an
// exact-size cache of freed blocks (wc_alloc.hip)
hipError_t wc_dev_malloc_raw(void **p, size_t bytes);
hipError_t wc_dev_free(const void *p);
// the same for a block whose last users ran on stream s (waits for s only)
hipError_t wc_dev_free_on(const void *p, hipStream_t s);
// a device-wide wait that no graph capture on another thread is cut by; after
// it, wc_dev_free_idle caches blocks without waiting again (teardown of many)
void wc_device_wait();
hipError_t wc_dev_free_idle(const void *p);
// hipHostFree, which may wait for the device, at a moment no capture runs
hipError_t wc_host_free(void *p);
// Frees by this thread between begin and end are held back; end(synced)
// caches them, at once when synced (the stream their last users ran on has
// been synchronized), else through wc_dev_free's wait for the device.
void wc_alloc_defer_begin();
void wc_alloc_defer_end(bool synced);
// a graph capture in progress on this thread (device-wide waits and frees
// of the allocator wait for it to end)
void wc_capture_begin();
void wc_capture_end();
struct CaptureGuard {
  CaptureGuard() { wc_capture_begin(); }
  ~CaptureGuard() { wc_capture_end(); }
};
struct AllocDefer {
  bool synced = false;
  AllocDefer() { wc_alloc_defer_begin(); }
  ~AllocDefer() { wc_alloc_defer_end(synced); }
};
template <typename T> inline hipError_t wc_dev_malloc(T **p, size_t bytes) {
  return wc_dev_malloc_raw((void **)p, bytes);
}

#define WC_FAIL(code, ...)                       \
  do {                                           \
    char _b[512];                                \
    snprintf(_b, sizeof _b, __VA_ARGS__);        \
    g_err = _b;                                  \
    return code;                                 \
  } while (0)

#define WC_HIP(x)                                                                                    \
  do {                                                                                               \
    hipError_t _e = (x);                                                                             \
    if (_e != hipSuccess)                                                                            \
      WC_FAIL(WC_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

// launches are asynchronous: surface launch-configuration errors right away
#define WC_LAUNCHED() WC_HIP(hipGetLastError())

// Control state of chunked CG iterations (see wck::iter_check): the device
// control block, per-RHS data and residual history with pinned host mirrors,
// and the graphs captured for runs of 2, 4, ..., 32 iterations.
struct IterRun {
  static constexpr int kGraphs = 6;  // run lengths 1 << j, j = 1 .. 5
  static constexpr size_t kCtlBytes = sizeof(int) * wck::kCtlInts;
  static constexpr size_t kDatBytes = sizeof(double) * wck::kDatRows * wck::kMaxRhs;
  static constexpr size_t kBytes = kCtlBytes + kDatBytes + sizeof(double) * wck::kMaxChunk * wck::kMaxRhs;
  char *dev = nullptr, *host = nullptr;
  int *ctl = nullptr, *hctl = nullptr;
  double *dat = nullptr, *hdat = nullptr, *hist = nullptr, *hhist = nullptr;
  hipStream_t cap = nullptr;  // capture stream (graphs are launched on the caller's stream)
  hipGraphExec_t exec[kGraphs] = {};
  std::vector<uintptr_t> key;  // what the graphs were captured for
  long graph_runs = 0, direct_runs = 0;
};

inline void drop_graphs(IterRun &ir) {
  for (hipGraphExec_t &e : ir.exec)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  ir.key.clear();
}

// (after wc_device_wait: the run's work has ended)
inline void free_run(IterRun &ir) {
  drop_graphs(ir);
  if (ir.cap) (void)hipStreamDestroy(ir.cap);
  (void)wc_dev_free_idle(ir.dev);
  if (ir.host) (void)wc_host_free(ir.host);
  ir = IterRun();
}

inline int ensure_run(IterRun &ir) {
  if (ir.dev) return WC_OK;
  WC_HIP(wc_dev_malloc((void **)&ir.dev, IterRun::kBytes));
  WC_HIP(hipHostMalloc((void **)&ir.host, IterRun::kBytes, hipHostMallocDefault));
  memset(ir.host, 0, IterRun::kBytes);
  ir.ctl = (int *)ir.dev;
  ir.hctl = (int *)ir.host;
  ir.dat = (double *)(ir.dev + IterRun::kCtlBytes);
  ir.hdat = (double *)(ir.host + IterRun::kCtlBytes);
  ir.hist = (double *)(ir.dev + IterRun::kCtlBytes + IterRun::kDatBytes);
  ir.hhist = (double *)(ir.host + IterRun::kCtlBytes + IterRun::kDatBytes);
  return WC_OK;
}

// Clover-inverse blocks per right-hand side (a batch whose slots have
// different masses): right-hand side r of parity p uses Ai[p] + roff[r]
// (roff: device array of offsets in complex elements; several right-hand
// sides of one mass share a set).  roff null: Ai[p] serves every
// right-hand side.  While an operator's `cov` points here, every
// clover-inverse application of that operator uses these blocks.
struct ClovSets {
  const void *Ai[2] = {nullptr, nullptr};
  const long *roff = nullptr;
};

struct wc_operator {
  wc_params p{};
  Geom g{};
  int prec = WC_DOUBLE;
  int nl = 9;
  size_t csz = 16;  // bytes per complex value
  double mass = 0, diag = 0, cs = 0, ct = -0.5;
  bool has_clover = false, open_bc = false;
  void *W[2] = {nullptr, nullptr};
  void *A[2] = {nullptr, nullptr};
  void *Ai[2] = {nullptr, nullptr};
  void *Unat = nullptr;  // FP64 links [4][V][9], kept to rebuild the clover on set_mass
  int *err = nullptr;    // device flag of the clover build
  std::vector<void *> tmp;      // half-field scratch, tmp_bytes each
  size_t tmp_bytes = 0;
  double *partials = nullptr;   // per-block reduction partials
  size_t npartials = 0;
  double *dots = nullptr;       // per-(site, RHS) products of the hop's EPI_DOT
  size_t ndots = 0;
  double *st = nullptr;         // device scalars (CG state), nst doubles
  double *host = nullptr;       // pinned host mirror, nst doubles
  size_t nst = 0;
  int bps = 0;                  // 64-site blocks per time slice (0: not a whole number)
  wck::HopShape shape[65];      // hop launch shape per right-hand-side count
  bool shape_set[65] = {};      // shape chosen (by the tuner or the caller)
  bool shape_user[65] = {};     // shape set by the caller (wc_set_hop_shape): used for 16-bit hops too
  int chunk = 0;                // iterations per run: 0 automatic, 1 one at a time, else at most this many
  // FP16-storage hops of 4k right-hand sides: 0 vector ALUs, 1 matrix cores,
  // 2 automatic (the faster of the two, timed per lattice and class of
  // right-hand-side counts at first use; the vector form until then)
  int mma = 2;
  const int *halt_cur = nullptr;  // halt word of the run being enqueued (null outside runs)
  IterRun run;                  // chunked iterations of this operator's CG and mixed solves
  const ClovSets *cov = nullptr;  // per-RHS clover inverses in use (a multi-mass batch), or null
};

// the clover-inverse blocks of parity p as the operator applies them now
// (null: the scalar 1 / diag), their per-RHS offsets (null: one set), and
// the clover blocks themselves (never per right-hand side)
inline const void *clov_inv(const wc_operator *op, int p) {
  if (op->cov) return op->cov->Ai[p];
  return op->has_clover ? op->Ai[p] : nullptr;
}
inline const long *clov_roff(const wc_operator *op) { return op->cov ? op->cov->roff : nullptr; }
inline const void *clov_fwd(const wc_operator *op, int p) { return op->has_clover ? op->A[p] : nullptr; }

inline size_t half_bytes(const wc_operator *op) { return (size_t)op->g.Vh * 12 * op->csz; }

// the kernels' 16-bit storage code of an inner_storage argument (-1: unknown)
inline int storage_code(int inner_storage) {
  switch (inner_storage) {
    case WC_STORAGE_NATIVE: return kS16None;
    case WC_STORAGE_HALF: return kS16Half;
    case WC_STORAGE_BF16: return kS16Bf16;
    default: return -1;
  }
}
// Smallest reliable-update interval (residual reduction per inner cycle) of
// a 16-bit storage: FP16 resolves about 5e-4 of a field's scale, BF16 about
// 4e-3; asking one inner cycle for a reduction near that makes the recursion
// stagnate or diverge (BF16 at 2e-2 diverges on an 8^3x16 clover solve;
// 1e-1 needs the fewest iterations and updates on 24^3x48).
inline double s16_delta_floor(int code) {
  return code == kS16Bf16 ? 1e-1 : code == kS16Half ? 1e-2 : 0.0;
}

// n scratch half fields of nrhs right-hand sides each; s: the stream of
// the call (the previous fields' last users ran on it)
inline int ensure_tmp(wc_operator *op, size_t n, int nrhs, hipStream_t s) {
  const size_t bytes = half_bytes(op) * (size_t)nrhs;
  if (bytes > op->tmp_bytes) {
    for (void *q : op->tmp) (void)wc_dev_free_on(q, s);
    op->tmp.clear();
    op->tmp_bytes = bytes;
  }
  while (op->tmp.size() < n) {
    void *q = nullptr;
    WC_HIP(wc_dev_malloc(&q, op->tmp_bytes));
    op->tmp.push_back(q);
  }
  return WC_OK;
}

// reduction partials and CG scalars for nrhs right-hand sides (s as for
// ensure_tmp)
inline int ensure_red(wc_operator *op, int nrhs, hipStream_t s) {
  const size_t np = (size_t)std::max(wck::kRedBlocks, wck::hop_blocks(op->g)) * nrhs;
  if (np > op->npartials) {
    (void)wc_dev_free_on(op->partials, s);
    op->partials = nullptr;
    WC_HIP(wc_dev_malloc((void **)&op->partials, sizeof(double) * np));
    op->npartials = np;
  }
  const size_t nd = (size_t)op->g.Vh * nrhs;
  if (nd > op->ndots) {
    (void)wc_dev_free_on(op->dots, s);
    op->dots = nullptr;
    WC_HIP(wc_dev_malloc((void **)&op->dots, sizeof(double) * nd));
    op->ndots = nd;
  }
  // the scalars (16 per right-hand side, 8 KB at most) are sized for the
  // largest count at once: they never grow, so no pinned block is freed
  // while other threads may capture
  const size_t ns = (size_t)16 * wck::kMaxRhs;
  if (ns > op->nst) {
    (void)wc_dev_free_on(op->st, s);
    if (op->host) (void)wc_host_free(op->host);
    op->st = nullptr;
    op->host = nullptr;
    WC_HIP(wc_dev_malloc((void **)&op->st, sizeof(double) * ns));
    WC_HIP(hipHostMalloc((void **)&op->host, sizeof(double) * ns, hipHostMallocDefault));
    op->nst = ns;
  }
  return WC_OK;
}

inline void free_op(wc_operator *op) {
  if (!op) return;
  wc_device_wait();  // the operator's work has ended
  for (int p = 0; p < 2; ++p) {
    (void)wc_dev_free_idle(op->W[p]);
    (void)wc_dev_free_idle(op->A[p]);
    (void)wc_dev_free_idle(op->Ai[p]);
  }
  (void)wc_dev_free_idle(op->Unat);
  (void)wc_dev_free_idle(op->err);
  for (void *q : op->tmp) (void)wc_dev_free_idle(q);
  (void)wc_dev_free_idle(op->partials);
  (void)wc_dev_free_idle(op->dots);
  (void)wc_dev_free_idle(op->st);
  if (op->host) (void)wc_host_free(op->host);
  free_run(op->run);
  delete op;
}

// ------------------------------------------------------- hop launch shapes
template <typename R> inline int max_rb_of() { return wck::max_rb<R>(); }
inline int max_rb_prec(int prec) { return prec == WC_DOUBLE ? max_rb_of<double>() : max_rb_of<float>(); }

// the LDS staging area of a shape (prepared links, 9 complex per direction,
// and clover blocks of the sites a workgroup covers) fits the per-workgroup
// limit
inline bool stage_fits(const wc_operator *op, int nrhs, const wck::HopShape &sh) {
  if (!sh.ls) return true;
  // ls 2: 12 spinor components per thread (the hop ignores it unless one RHS)
  if (sh.ls == 2) return (size_t)12 * kBlk * sh.tpb * op->csz <= (size_t)wck::kMaxStageBytes;
  const long G = (nrhs + sh.rb - 1) / sh.rb;
  const size_t bytes = (size_t)(72 + 36) * (size_t)wck::stage_sites((long)kBlk * sh.tpb, G) * op->csz;
  return bytes <= (size_t)wck::kMaxStageBytes;
}

// Shape used until the tuner (or the caller) picks one: at most 4
// workgroups per CU; with several right-hand sides, rb per thread with the
// least padding (ties: the larger rb), 256 threads per workgroup and the
// links staged in LDS where the staging area fits.
inline wck::HopShape default_shape(const wc_operator *op, int nrhs) {
  wck::HopShape sh;
  sh.occ = 4;
  if (nrhs <= 1) return sh;
  sh.tpb = 4;
  const int mrb = max_rb_prec(op->prec);
  int best = 1, waste_best = 1 << 30;
  for (int rb = mrb; rb >= 1; --rb) {
    const int waste = ((nrhs + rb - 1) / rb) * rb - nrhs;
    if (waste < waste_best) {
      waste_best = waste;
      best = rb;
    }
  }
  sh.rb = best;
  sh.ls = 1;
  if (!stage_fits(op, nrhs, sh)) sh.ls = 0;
  return sh;
}

inline bool shape_valid(const wc_operator *op, const wck::HopShape &sh) {
  return sh.tpb >= 1 && sh.tpb <= wck::kMaxTpb && sh.rb >= 1 && sh.rb <= max_rb_prec(op->prec) && sh.occ >= 0 &&
         sh.occ <= wck::kMaxOcc && sh.seg >= 1 && (sh.seg == 1 || (op->bps > 0 && sh.seg <= op->bps)) &&
         (sh.ls >= 0 && sh.ls <= 2);
}

// process-wide cache of tuned shapes: (dims, precision, link format, class of
// nrhs: 1, 2-3, 4-7, 8-15, 16-31, 32-64)
typedef std::tuple<int, int, int, int, int, int, int> ShapeKey;
inline int nrhs_class(int nrhs) {
  int c = 1;
  while (2 * c <= nrhs && c < 32) c *= 2;
  return c;
}
extern std::mutex g_shape_mu;
extern std::map<ShapeKey, wck::HopShape> g_shape_cache;
inline ShapeKey shape_key(const wc_operator *op, int nrhs) {
  return ShapeKey(op->g.L[0], op->g.L[1], op->g.L[2], op->g.L[3], op->prec, op->nl, nrhs_class(nrhs));
}

// the shape set for this operator, else the cached shape of the same
// lattice, precision, link format and right-hand-side class (rb kept within
// what nrhs allows), else the default
inline wck::HopShape shape_for(const wc_operator *op, int nrhs) {
  if (nrhs >= 1 && nrhs <= 64 && op->shape_set[nrhs]) return op->shape[nrhs];
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shape_cache.find(shape_key(op, nrhs));
    if (it != g_shape_cache.end()) {
      wck::HopShape sh = it->second;
      if (sh.rb > nrhs) sh.rb = 1;
      if (!stage_fits(op, nrhs, sh)) sh.ls = 0;
      return sh;
    }
  }
  return default_shape(op, nrhs);
}

// launch shape of the hops with 16-bit spinors (FP32 operator): the caller's
// shape if one was set, else the one timed fastest for those hops (pick_s16,
// process-wide per lattice, link format and class of right-hand-side
// counts), else the FP32 hop's
extern std::map<ShapeKey, wck::HopShape> g_shape16_cache;
inline wck::HopShape shape16_for(const wc_operator *op, int nrhs) {
  if (nrhs >= 1 && nrhs <= 64 && op->shape_user[nrhs]) return op->shape[nrhs];
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shape16_cache.find(shape_key(op, nrhs));
    if (it != g_shape16_cache.end()) {
      wck::HopShape sh = it->second;
      if (sh.rb > nrhs) sh.rb = 1;
      if (!stage_fits(op, nrhs, sh)) sh.ls = 0;
      return sh;
    }
  }
  return shape_for(op, nrhs);
}

// the hop form timed faster for the operator's lattice, link format and
// class of right-hand-side counts (process-wide, like the shapes), or -1
extern std::map<ShapeKey, int> g_mma_cache;
inline int mma_for(const wc_operator *op, int nrhs) {
  if (op->mma != 2) return op->mma;
  std::lock_guard<std::mutex> lk(g_shape_mu);
  auto it = g_mma_cache.find(shape_key(op, nrhs));
  return it != g_mma_cache.end() ? it->second : 0;
}

// clover-inverse blocks of both parities (and optionally the blocks) of the
// operator's clover term at another mass (wc_api.hip)
int build_mass_blocks(const wc_operator *op, double mass, void *ai0, void *ai1, void *a0, void *a1, int *err,
                      hipStream_t s);

// ------------------------------------------------------------ compositions
template <typename R>
inline HopLaunch hop_args(const wc_operator *op, int nrhs, int pout, int dag, int epi, const void *in, void *out) {
  HopLaunch a{};
  a.g = op->g;
  a.nrhs = nrhs;
  a.pout = pout;
  a.dag = dag;
  a.nl = op->nl;
  a.epi = epi;
  a.in = in;
  a.out = out;
  a.W = op->W[pout];
  a.cs = op->cs;
  a.ct = op->ct;
  for (int mu = 0; mu < 4; ++mu) {
    a.ph[mu][0] = op->p.boundary_phases[mu][0];
    a.ph[mu][1] = op->p.boundary_phases[mu][1];
  }
  a.open_bc = op->open_bc ? 1 : 0;
  a.diag = op->diag;
  a.clov_scalar = 1.0 / op->diag;
  a.shape = shape_for(op, nrhs);
  a.bps = op->bps;
  a.halt = op->halt_cur;
  a.mma = mma_for(op, nrhs);
  a.clov_roff = clov_roff(op);
  return a;
}

// launch a hop; an invalid shape is an error, never a silent no-op
#define WC_HOP(R, a, s)                                                                        \
  do {                                                                                         \
    if (!wck::hop<R>((a), (s)))                                                                \
      WC_FAIL(WC_ERR_ARG, "hop: launch shape rb=%d tpb=%d occ=%d seg=%d ls=%d is outside the limits", \
              (a).shape.rb, (a).shape.tpb, (a).shape.occ, (a).shape.seg, (a).shape.ls);        \
  } while (0)

// out = A_p in (inv = 0) or A_p^-1 in (inv = 1) [optionally y - that]; the
// inverse follows the per-RHS sets in use (clov_inv), A_p is the operator's
template <typename R>
inline void site_op(const wc_operator *op, int nrhs, int p, int inv, const void *in, void *out, const void *y,
                    hipStream_t s, int half = 0) {
  const void *B = inv ? clov_inv(op, p) : clov_fwd(op, p);
  double sc = inv ? 1.0 / op->diag : op->diag;
  wck::site_clover<R>(op->g, nrhs, p, B, sc, in, out, y, op->open_bc ? 1 : 0, s, half,
                      inv ? clov_roff(op) : nullptr);
}

// Mpc (which 0), Mpc^dag (1) or Mpc^dag Mpc (2) with D = pd on nrhs
// right-hand sides; dotv (optional) makes the last hop write per-block,
// per-RHS partials of Re<dotv, out>.  tD_in (optional, which 0 or 2): a
// field already holding A_D^-1 in.  half (FP32, which 2): in, dotv, tD_in
// and the intermediates are stored as 16-bit c2h, out in FP32.
// Scratch: tmp[0..2] (mpc_impl), or the fields given to mpc_run: tD (which
// 0, 2), tC (always), tY (which 2; may be tD, which is dead by the time tY
// is written); dots and partials only with dotv.
template <typename R>
inline int mpc_run(wc_operator *op, int nrhs, int pd, int which, void *out, const void *in, const void *dotv,
                   hipStream_t s, void *tD, void *tC, void *tY, const void *tD_in, double *dots, double *partials,
                   int half) {
  if (half && (sizeof(R) != 4 || which != 2))
    WC_FAIL(WC_ERR_ARG, "mpc: 16-bit storage is for the FP32 normal operator only");
  const int pc = 1 - pd;
  const void *clovC = clov_inv(op, pc);
  const void *clovD = clov_inv(op, pd);
  const void *src = in;
  void *dst_mpc = (which == 2) ? tY : out;
  if (which == 0 || which == 2) {
    const void *aD = tD_in;
    if (!aD) {
      site_op<R>(op, nrhs, pd, 1, src, tD, nullptr, s, half);
      aD = tD;
    }
    HopLaunch a = hop_args<R>(op, nrhs, pc, 0, wck::EPI_CLOV, aD, tC);
    a.clov = clovC;
    a.half_in = a.half_out = half;
    if (half) a.shape = shape16_for(op, nrhs);
    WC_HOP(R, a, s);
    HopLaunch b = hop_args<R>(op, nrhs, pd, 0, wck::EPI_AXPY | (which == 0 && dotv ? wck::EPI_DOT : 0), tC,
                              dst_mpc);
    b.y = src;
    b.dotv = dotv;
    b.dots = dots;
    b.half_in = b.half_out = half;
    if (half) b.shape = a.shape;
    WC_HOP(R, b, s);
    if (which == 0 && dotv) wck::dot_blocks(op->g, nrhs, dots, partials, s, op->halt_cur);
    src = dst_mpc;
  }
  if (which == 1 || which == 2) {
    HopLaunch a = hop_args<R>(op, nrhs, pc, 1, wck::EPI_CLOV, src, tC);
    a.clov = clovC;
    a.half_in = a.half_out = half;
    if (half) a.shape = shape16_for(op, nrhs);
    WC_HOP(R, a, s);
    HopLaunch b = hop_args<R>(op, nrhs, pd, 1, wck::EPI_CLOV | wck::EPI_AXPY | (dotv ? wck::EPI_DOT : 0), tC,
                              out);
    b.clov = clovD;
    b.y = src;
    b.dotv = dotv;
    b.dots = dots;
    b.half_in = half;  // the product itself stays in the working precision
    if (half) b.shape = a.shape;
    WC_HOP(R, b, s);
    if (dotv) wck::dot_blocks(op->g, nrhs, dots, partials, s, op->halt_cur);
  }
  WC_LAUNCHED();
  return WC_OK;
}

template <typename R>
inline int mpc_impl(wc_operator *op, int nrhs, int pd, int which, void *out, const void *in, const void *dotv,
                    hipStream_t s, const void *tD_in = nullptr, double *partials = nullptr, int half = 0) {
  int rc = ensure_tmp(op, 3, nrhs, s);
  if (rc) return rc;
  if ((rc = ensure_red(op, nrhs, s))) return rc;
  // the operator's partials exist only from here on
  if (!partials) partials = op->partials;
  return mpc_run<R>(op, nrhs, pd, which, out, in, dotv, s, op->tmp[0], op->tmp[1], op->tmp[2], tD_in, op->dots,
                    partials, half);
}

// per-RHS |x|^2 into dev_out[0..nrhs)
template <typename R>
inline int norm2_rhs(wc_operator *op, int nrhs, const void *x, double *dev_out, hipStream_t s) {
  int rc = ensure_red(op, nrhs, s);
  if (rc) return rc;
  const long n = op->g.Vh * 12 * nrhs;
  wck::dot_partials_rhs<R>(n, nrhs, x, x, op->partials, nullptr, s);
  wck::reduce_final_rhs(op->partials, wck::kRedBlocks, nrhs, dev_out, s);
  WC_LAUNCHED();
  return WC_OK;
}

// one iteration of the inner CG on lo: mmp = A p (A_D^-1 p in aD), alpha,
// r -= alpha mmp, beta, x += alpha p, p = beta p + r, aD = A_D^-1 p
template <typename R>
inline int inner_iteration(wc_operator *lo, int N, int pd, void *r, void *p, void *mmp, void *x, void *aD,
                           hipStream_t s, double *st = nullptr, double *partials = nullptr, int half = 0) {
  int rc;
  const long n = lo->g.Vh * 12 * N;
  if (!st) st = lo->st;
  if (!partials) partials = lo->partials;
  const int *halt = lo->halt_cur;
  if ((rc = mpc_impl<R>(lo, N, pd, 2, mmp, p, p, s, aD, partials, half))) return rc;
  wck::cg_alpha(partials, wck::hop_blocks(lo->g), N, st, s, halt);
  wck::cg_update_r<R>(n, N, st, r, mmp, partials, s, halt);
  wck::cg_beta(partials, wck::kRedBlocks, N, st, s, halt);
  wck::cg_update_xp<R>(lo->g, N, pd, st, r, p, x, clov_inv(lo, pd), 1.0 / lo->diag, aD,
                       lo->open_bc ? 1 : 0, s, half, halt, clov_roff(lo));
  WC_LAUNCHED();
  return WC_OK;
}

// ------------------------------------------------- chunked CG iterations
// What a captured run depends on beyond the caller's pointers: the
// operator's links, clover, coefficients and hop launch shape.
inline void op_key(const wc_operator *op, int nrhs, std::vector<uintptr_t> &key) {
  auto bits = [](double d) {
    uint64_t u;
    memcpy(&u, &d, sizeof u);
    return (uintptr_t)u;
  };
  const wck::HopShape sh = shape_for(op, nrhs), s16 = shape16_for(op, nrhs);
  auto tmp = [&](size_t i) { return i < op->tmp.size() ? (uintptr_t)op->tmp[i] : (uintptr_t)0; };
  for (uintptr_t v : {(uintptr_t)op, (uintptr_t)op->W[0], (uintptr_t)op->W[1], (uintptr_t)clov_inv(op, 0),
                      (uintptr_t)clov_inv(op, 1), (uintptr_t)clov_roff(op), tmp(0), tmp(1), tmp(2), (uintptr_t)op->dots, (uintptr_t)op->partials, bits(op->diag), bits(op->cs), bits(op->ct), (uintptr_t)op->open_bc,
                      (uintptr_t)op->has_clover, (uintptr_t)sh.rb, (uintptr_t)sh.tpb, (uintptr_t)sh.occ,
                      (uintptr_t)sh.seg, (uintptr_t)sh.ls, (uintptr_t)op->bps, (uintptr_t)s16.rb, (uintptr_t)s16.tpb,
                      (uintptr_t)s16.occ, (uintptr_t)s16.seg, (uintptr_t)s16.ls, (uintptr_t)mma_for(op, nrhs)})
    key.push_back(v);
  for (int mu = 0; mu < 4; ++mu) {
    key.push_back(bits(op->p.boundary_phases[mu][0]));
    key.push_back(bits(op->p.boundary_phases[mu][1]));
  }
}

// Iterations per run: the operator's setting, or automatically from the
// estimated time of one iteration (HBM bytes of Mpc^dag Mpc and the updates
// at ~5 TB/s): short iterations are dominated by launch and synchronisation
// costs, which a captured run of many iterations removes.
inline int chunk_cap(const wc_operator *op, int nrhs, int setting) {
  if (setting > 0) return std::min(setting, 32);
  const double bytes = (double)op->g.Vh * (double)op->csz * (32.0 * op->nl + 240.0 * nrhs + 144.0);
  const double us = bytes / 5e6;
  return us < 100.0 ? 32 : us < 400.0 ? 16 : 8;
}

// Largest power of two <= n within [1, cap].
inline int pow2_floor(double n, int cap) {
  int k = 1;
  while (2 * k <= cap && 2 * k <= n) k *= 2;
  return k;
}

// Runs up to K (a power of two <= 32) iterations: enqueue(q) enqueues one
// iteration followed by iter_check on stream q (with the operator's halt_cur
// set to the run's halt word).  K > 1 replays a graph captured once per K and
// key.  On return (the stream synchronised) hctl and the first
// hctl[kCtlIter] rows of hhist (nrhs each) hold the run's outcome.
template <typename F>
inline int run_iterations(IterRun &ir, const std::vector<uintptr_t> &key, int K, int nrhs, hipStream_t s,
                          F &&enqueue) {
  int rc;
  // scratch that grows inside the run is freed once the run has ended on s
  // (a device-wide wait inside a capture would break it)
  AllocDefer defer;
  if ((rc = ensure_run(ir))) return rc;
  if (ir.key != key) {
    drop_graphs(ir);
    ir.key = key;
  }
  ir.hctl[wck::kCtlHalt] = 0;
  ir.hctl[wck::kCtlIter] = 0;
  ir.hctl[wck::kCtlMax] = K;
  ir.hctl[wck::kCtlErr] = 0;
  ir.hctl[wck::kCtlTrig] = 0;
  WC_HIP(hipMemcpyAsync(ir.ctl, ir.hctl, IterRun::kCtlBytes, hipMemcpyHostToDevice, s));
  if (K <= 1) {
    if ((rc = enqueue(s))) return rc;
    ++ir.direct_runs;
  } else {
    int j = 0;
    while ((2 << j) < K) ++j;
    if (!ir.exec[j]) {
      if (!ir.cap) WC_HIP(hipStreamCreateWithFlags(&ir.cap, hipStreamNonBlocking));
      hipGraph_t gr = nullptr;
      hipError_t e;
      {
        CaptureGuard cg;
        WC_HIP(hipStreamBeginCapture(ir.cap, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < K && !rc; ++i) rc = enqueue(ir.cap);
        e = hipStreamEndCapture(ir.cap, &gr);
      }
      if (rc || e != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        if (rc) return rc;
        WC_FAIL(WC_ERR_HIP, "capturing a run of %d CG iterations failed: %s", K, hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&ir.exec[j], gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (ei != hipSuccess) {
        ir.exec[j] = nullptr;
        WC_FAIL(WC_ERR_HIP, "instantiating a run of %d CG iterations failed: %s", K, hipGetErrorString(ei));
      }
    }
    WC_HIP(hipGraphLaunch(ir.exec[j], s));
    ++ir.graph_runs;
  }
  WC_HIP(hipMemcpyAsync(ir.hctl, ir.ctl, IterRun::kCtlBytes, hipMemcpyDeviceToHost, s));
  WC_HIP(hipMemcpyAsync(ir.hhist, ir.hist, sizeof(double) * (size_t)K * nrhs, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  defer.synced = true;
  const int done = ir.hctl[wck::kCtlIter];
  if (done < 1 || done > K) WC_FAIL(WC_ERR_HIP, "CG run: %d iterations reported for a run of %d", done, K);
  return WC_OK;
}

// average decay of log |r|^2 per iteration from c_ref to c_now over n
// iterations (negative), or fallback when that is not known
inline double decay_rate(double c_ref, double c_now, long n, double fallback) {
  if (n >= 2 && c_ref > 0.0 && c_now > 0.0) {
    const double r = std::log(c_now / c_ref) / (double)n;
    if (r < 0.0) return r;
  }
  return fallback;
}
// iterations from c_now to target at that rate; -1 when unknown
inline double iters_to(double c_now, double target, double rate) {
  if (!(rate < 0.0) || !(c_now > 0.0) || !(target > 0.0)) return -1.0;
  if (c_now <= target) return 0.0;
  return std::log(target / c_now) / rate;
}
// run length when no prediction is available
constexpr int kFirstRun = 4;

// dst = L x + S src (preconditioned.py:32-54, schur_complement_two.py) with
// x from inner(px, pb), which solves Mpc^dag Mpc x = pb = R_ne src starting
// from the guess px = L^-1 dst = A_D dst_D.  Scratch: tmp[1], tmp[6..8].
template <typename R, typename F>
inline int schur_solve(wc_operator *op, int nrhs, int pd, void *dst, const void *src, hipStream_t s, F &&inner) {
  const int pc = 1 - pd;
  int rc = ensure_tmp(op, 9, nrhs, s);
  if (rc) return rc;
  const size_t hb = half_bytes(op) * (size_t)nrhs;
  auto half = [&](const void *f, int p) { return (void *)((char *)f + (size_t)p * hb); };
  void *tR = op->tmp[6], *pb = op->tmp[7], *px = op->tmp[8];
  void *tC = op->tmp[1];
  // R src = src_D - Meooe A_C^-1 src_C
  site_op<R>(op, nrhs, pc, 1, half(src, pc), tC, nullptr, s);
  {
    HopLaunch a = hop_args<R>(op, nrhs, pd, 0, wck::EPI_AXPY, tC, tR);
    a.y = half(src, pd);
    WC_HOP(R, a, s);
  }
  // R_ne src = Mpc^dag R src
  if ((rc = mpc_impl<R>(op, nrhs, pd, 1, pb, tR, nullptr, s))) return rc;
  // guess L^-1 dst = A_D dst_D
  site_op<R>(op, nrhs, pd, 0, half(dst, pd), px, nullptr, s);
  WC_LAUNCHED();
  if ((rc = inner(px, (const void *)pb))) return rc;
  // dst = L x + S src: dst_D = A_D^-1 x ; dst_C = A_C^-1 src_C - A_C^-1 Meooe dst_D
  site_op<R>(op, nrhs, pd, 1, px, half(dst, pd), nullptr, s);
  site_op<R>(op, nrhs, pc, 1, half(src, pc), tR, nullptr, s);
  {
    HopLaunch a = hop_args<R>(op, nrhs, pc, 0, wck::EPI_CLOV | wck::EPI_AXPY, half(dst, pd), half(dst, pc));
    a.clov = clov_inv(op, pc);
    a.y = tR;
    WC_HOP(R, a, s);
  }
  WC_LAUNCHED();
  return WC_OK;
}

// A half field of nrhs right-hand sides larger than one buffer descriptor's
// 4 GB is addressed per t-slice by the hop (HopLaunch::big): that needs whole
// 64-site blocks per slice and a slice of nrhs right-hand sides within 4 GB.
inline int check_nrhs(const wc_operator *op, int nrhs, const char *fn) {
  if (nrhs < 1 || nrhs > 64) WC_FAIL(WC_ERR_ARG, "%s: nrhs = %d must be in [1, 64]", fn, nrhs);
  if ((double)half_bytes(op) * nrhs >= 4294967296.0) {
    const long slice = op->g.Vh / op->g.L[3];
    if (slice % 64 != 0)
      WC_FAIL(WC_ERR_ARG, "%s: a half field of %d right-hand sides exceeds 4 GB and a t-slice of %ld half-sites "
                          "is not a whole number of 64-site blocks", fn, nrhs, slice);
    if ((double)half_bytes(op) / op->g.L[3] * nrhs >= 4294967296.0)
      WC_FAIL(WC_ERR_ARG, "%s: a t-slice of %d right-hand sides exceeds 4 GB", fn, nrhs);
  }
  return WC_OK;
}


// tunes the hop of nrhs right-hand sides on first use (wc_api.hip)
int maybe_tune(wc_operator *op, int nrhs, hipStream_t s);
