// Continuous-batching mixed-precision eo2_ne solver (C ABI in
// include/gpt_amd_wc.h, wc_batch_*).  N slots share one operator pair (FP64
// outer, inner precision for the CG recursion) and advance together: every
// inner iteration applies Mpc^dag Mpc to all N slots in one pass of the
// multi-RHS kernels.  A request enters a free slot whenever one is free
// (its source terms and initial residual are computed on its own, FP64), and
// leaves when its FP64 true residual meets its tolerance, so slots are
// recycled while the others keep iterating.
//
// Per-slot state: the FP64 solution y and R src in N-RHS fields, the inner
// r, p, x, A_D^-1 p in N-RHS fields of the inner precision, and the CG
// scalars of the slot (st[j * N + slot]).  Reliable updates (y += x, r
// recomputed in FP64 from y, p kept) happen for all iterating slots at once,
// when any slot's iterated residual has fallen by delta^2 or a slot reaches
// its iteration limit; a slot finishes at the reliable update that finds its
// FP64 true residual within its tolerance.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "wc_internal.h"

namespace {

// slot r of an N-RHS half field <-> a single-RHS half field: element e of
// the single field is element e * N + r of the N-RHS one (same AoSoA order)
template <typename CS, typename CD> __device__ __forceinline__ CD cvt(CS v) {
  if constexpr (std::is_same<CS, CD>::value) {
    return v;
  } else {
    CD o;
    o.x = (decltype(o.x))v.x;
    o.y = (decltype(o.y))v.y;
    return o;
  }
}
template <typename CS, typename CD> __global__ void k_to_slot(long n, int N, int slot, const CS *src, CD *dst) {
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e * N + slot] = cvt<CS, CD>(src[e]);
}
// dst slot = narrow(a * src), src FP64 single-RHS, dst FP32 or 16-bit
template <typename CD>
__global__ void k_to_slot_scaled(long n, int N, int slot, double a, const double2 *src, CD *dst) {
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e * N + slot] = narrow<CD>(make_float2((float)(a * src[e].x), (float)(a * src[e].y)));
}
template <typename C> __global__ void k_from_slot(long n, int N, int slot, const C *src, C *dst) {
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e] = src[e * N + slot];
}
template <typename C> __global__ void k_zero_slot(long n, int N, int slot, C *dst) {
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) {
    C z;
    z.x = 0;
    z.y = 0;
    dst[e * N + slot] = z;
  }
}
inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct wc_batch {
  wc_operator *hi = nullptr, *lo = nullptr;
  int N = 0, pd = 1;
  int h16 = 0;  // 16-bit storage code of the inner p and Schur intermediates (kS16*)
  double delta = 0.1;
  hipStream_t s = nullptr;
  // N-RHS half fields: FP64 y (solution of Mpc y = R src), R src, scratch
  void *xh = nullptr, *bs = nullptr, *mp = nullptr, *rs = nullptr, *rh = nullptr;
  // N-RHS half fields in the inner precision
  void *rl = nullptr, *pl = nullptr, *ml = nullptr, *xl = nullptr, *aD = nullptr;
  std::vector<void *> src;  // per slot: the request's full FP64 source (kept for S src)
  // single-RHS half fields for admission and extraction: FP64 (one) and in
  // the inner precision (lone)
  void *one[6] = {}, *lone[2] = {};
  double *st = nullptr, *host = nullptr, *partials = nullptr;  // CG scalars, reduction partials
  // host-side slot state
  std::vector<int> occupied, active, finished, conv;
  std::vector<long> iters, maxiter;
  std::vector<double> tol, src2, maxr, tres;
  // per-slot frame of the 16-bit inner state (1 otherwise) and the ratio
  // |r_s|^2 / |r_h|^2 of the last reliable update (convergence predictor)
  std::vector<double> scale, ratio;
  long updates = 0;
  // chunked iterations (wck::iter_check mode 1): device trigger data to be
  // re-sent after an admission or a reliable update; per slot the inner
  // |r|^2 at its last restart and now, the decay rate per iteration of its
  // last cycle, and the iterations since its restart
  IterRun run;
  bool dirty = true;
  std::vector<double> cref, clast, rate;
  std::vector<long> since;
  // several masses (wc_batch_admit_mass): nsets (default N) sets of
  // clover-inverse blocks of both parities, FP64 for the outer operator (ai_hi) and in the inner
  // precision (ai_lo), set k of parity p at ai_*[p] + k * 36 Vh complex
  // values.  Slot r uses set set_of[r]; the slots of one mass share a set
  // (set_mass: the mass a set holds, NaN before it is built).  While the
  // batch works, the operators use them through their `cov`: roff[r] =
  // set_of[r] * 36 Vh on the device, or one set for all when every occupied
  // slot has the same one.
  bool mm = false;
  int nsets = 0;  // sets a multi-mass batch holds (0: one per slot)
  void *ai_hi[2] = {nullptr, nullptr}, *ai_lo[2] = {nullptr, nullptr};
  long *roff = nullptr;
  int *err = nullptr;
  std::vector<double> mass, set_mass;
  std::vector<int> set_of;
  ClovSets cov_hi, cov_lo;
};

// the clover-inverse set k of both operators, parity p
static const void *set_hi(const wc_batch *b, int p, int k) {
  return (const char *)b->ai_hi[p] + (size_t)k * 36 * b->hi->g.Vh * b->hi->csz;
}
static const void *set_lo(const wc_batch *b, int p, int k) {
  return (const char *)b->ai_lo[p] + (size_t)k * 36 * b->lo->g.Vh * b->lo->csz;
}

// Sets the operators' per-slot clover inverses for the duration of a batch
// call (all slots: slot < 0; one slot's blocks as a single set: slot >= 0)
// and clears them again.  When every occupied slot uses one set, all slots
// run on that set alone: the kernels then stage its blocks per site rather
// than load them per right-hand side (the unoccupied slots' results are not
// used).
struct SlotClover {
  wc_batch *b;
  ClovSets one_hi, one_lo;
  SlotClover(wc_batch *b_, int slot) : b(b_) {
    if (!b->mm) return;
    int k = slot >= 0 ? b->set_of[slot] : -1;
    if (slot < 0) {
      for (int i = 0; i < b->N; ++i) {
        if (!b->occupied[i]) continue;
        if (k < 0) {
          k = b->set_of[i];
        } else if (k != b->set_of[i]) {
          k = -1;
          break;
        }
      }
      if (k < 0) {
        b->hi->cov = &b->cov_hi;
        b->lo->cov = &b->cov_lo;
        return;
      }
    }
    for (int p = 0; p < 2; ++p) {
      one_hi.Ai[p] = set_hi(b, p, k);
      one_lo.Ai[p] = set_lo(b, p, k);
    }
    b->hi->cov = &one_hi;
    b->lo->cov = &one_lo;
  }
  ~SlotClover() {
    if (b->mm) b->hi->cov = b->lo->cov = nullptr;
  }
};

static size_t hbytes(const wc_operator *op) { return half_bytes(op); }

static void free_batch(wc_batch *b) {
  if (!b) return;
  wc_device_wait();  // the batch's work has ended
  for (void *q : {b->xh, b->bs, b->mp, b->rs, b->rh, b->rl, b->pl, b->ml, b->xl, b->aD}) (void)wc_dev_free_idle(q);
  for (void *q : b->src) (void)wc_dev_free_idle(q);
  for (void *q : b->one) (void)wc_dev_free_idle(q);
  for (void *q : b->lone) (void)wc_dev_free_idle(q);
  (void)wc_dev_free_idle(b->st);
  (void)wc_dev_free_idle(b->partials);
  for (int p = 0; p < 2; ++p) {
    (void)wc_dev_free_idle(b->ai_hi[p]);
    (void)wc_dev_free_idle(b->ai_lo[p]);
  }
  (void)wc_dev_free_idle(b->roff);
  (void)wc_dev_free_idle(b->err);
  if (b->host) (void)wc_host_free(b->host);
  free_run(b->run);
  delete b;
}

// the per-slot set offsets to the device (by value: nothing on the host
// has to outlive the copy)
struct SetOffsets {
  long v[wck::kMaxRhs];
};
__global__ void k_set_offsets(SetOffsets o, int N, long *roff) {
  if ((int)threadIdx.x < N) roff[threadIdx.x] = o.v[threadIdx.x];
}
static int send_offsets(wc_batch *b) {
  SetOffsets o{};
  for (int i = 0; i < b->N; ++i) o.v[i] = (long)b->set_of[i] * 36 * b->hi->g.Vh;
  k_set_offsets<<<1, wck::kMaxRhs, 0, b->s>>>(o, b->N, b->roff);
  WC_LAUNCHED();
  return WC_OK;
}

// set k of both operators for `mass` (the operator's own blocks are copied
// when the mass is the operator's and it has a clover term)
static int build_set(wc_batch *b, int k, double mass) {
  wc_operator *hi = b->hi, *lo = b->lo;
  const size_t sh = (size_t)36 * hi->g.Vh * hi->csz, sl = (size_t)36 * lo->g.Vh * lo->csz;
  void *h0 = (void *)set_hi(b, 0, k), *h1 = (void *)set_hi(b, 1, k);
  void *l0 = (void *)set_lo(b, 0, k), *l1 = (void *)set_lo(b, 1, k);
  b->set_mass[k] = NAN;  // until it is built
  int rc;
  if (mass == hi->mass && hi->has_clover && lo->has_clover && mass == lo->mass) {
    WC_HIP(hipMemcpyAsync(h0, hi->Ai[0], sh, hipMemcpyDeviceToDevice, b->s));
    WC_HIP(hipMemcpyAsync(h1, hi->Ai[1], sh, hipMemcpyDeviceToDevice, b->s));
    WC_HIP(hipMemcpyAsync(l0, lo->Ai[0], sl, hipMemcpyDeviceToDevice, b->s));
    WC_HIP(hipMemcpyAsync(l1, lo->Ai[1], sl, hipMemcpyDeviceToDevice, b->s));
  } else {
    if ((rc = build_mass_blocks(hi, mass, h0, h1, nullptr, nullptr, b->err, b->s))) return rc;
    if ((rc = build_mass_blocks(lo, mass, l0, l1, nullptr, nullptr, b->err, b->s))) return rc;
  }
  b->set_mass[k] = mass;
  return WC_OK;
}

// slot's clover inverses for `mass`: the set of another slot of that mass,
// else a set that no other occupied slot uses, built for it (an error when
// the other occupied slots use every set: as many other masses as sets)
static int set_slot_mass(wc_batch *b, int slot, double mass) {
  const int N = b->N, K = (int)b->set_mass.size();
  int k = -1, rc;
  for (int j = 0; j < K && k < 0; ++j)
    if (b->set_mass[j] == mass) k = j;
  if (k < 0) {
    std::vector<char> used(K, 0);
    for (int i = 0; i < N; ++i)
      if (b->occupied[i] && i != slot) used[b->set_of[i]] = 1;
    for (int j = 0; j < K && k < 0; ++j)
      if (!used[j]) k = j;
    if (k < 0)
      WC_FAIL(WC_ERR_ARG, "wc_batch_admit_mass: the batch's %d clover-inverse sets are in use by %d other masses", K,
              K);
    if ((rc = build_set(b, k, mass))) return rc;
  }
  b->mass[slot] = mass;
  if (b->set_of[slot] != k) {
    b->set_of[slot] = k;
    return send_offsets(b);
  }
  return WC_OK;
}

// switch a batch to clover-inverse sets per slot, every slot on set 0 at
// the operator's mass to begin with
static int enable_masses(wc_batch *b) {
  if (b->mm) return WC_OK;
  wc_operator *hi = b->hi, *lo = b->lo;
  const int N = b->N, K = b->nsets > 0 ? b->nsets : N;
  const size_t sh = (size_t)36 * hi->g.Vh * hi->csz, sl = (size_t)36 * lo->g.Vh * lo->csz;
  for (int p = 0; p < 2; ++p) {
    WC_HIP(wc_dev_malloc(&b->ai_hi[p], sh * K));
    WC_HIP(wc_dev_malloc(&b->ai_lo[p], sl * K));
  }
  WC_HIP(wc_dev_malloc(&b->roff, sizeof(long) * N));
  WC_HIP(wc_dev_malloc((void **)&b->err, sizeof(int)));
  b->mass.assign(N, hi->mass);
  b->set_mass.assign(K, NAN);
  b->set_of.assign(N, 0);
  int rc;
  if ((rc = build_set(b, 0, hi->mass))) return rc;
  if ((rc = send_offsets(b))) return rc;
  for (int p = 0; p < 2; ++p) {
    b->cov_hi.Ai[p] = b->ai_hi[p];
    b->cov_lo.Ai[p] = b->ai_lo[p];
  }
  b->cov_hi.roff = b->cov_lo.roff = b->roff;
  b->mm = true;
  return WC_OK;
}

// the operators' scratch at the batch's size, so that no later call of the
// batch grows (frees and re-allocates) it: N-RHS temporaries tmp[0..2] and
// reductions of both operators
static int ensure_scratch(wc_batch *b) {
  int rc;
  if ((rc = ensure_tmp(b->hi, 3, b->N, b->s))) return rc;
  if ((rc = ensure_tmp(b->lo, 3, b->N, b->s))) return rc;
  if ((rc = ensure_red(b->hi, b->N, b->s))) return rc;
  return ensure_red(b->lo, b->N, b->s);
}

// FP64: y += x_inner / s for every slot, x_inner = 0, then r_s = b_s - Mpc y,
// r_h = Mpc^dag r_s, r_inner = s r_h in the slot's new frame s (1 without
// 16-bit storage; p and A_D^-1 p carried over by s_new / s_old); |r_s|^2
// and |r_h|^2 per slot to the host.  Scale rows of st: 11 s_new, 12
// s_new / s_old, 13 1 / s.
static int batch_reliable(wc_batch *b) {
  wc_operator *hi = b->hi, *lo = b->lo;
  const int N = b->N;
  const long n = hi->g.Vh * 12 * N;
  hipStream_t s = b->s;
  int rc;
  const bool single = lo->prec == WC_SINGLE;
  if (single) {
    // one pass folds x_inner into y and applies A_D^-1 (the first piece of
    // Mpc y); r_s = b_s - Mpc y goes straight to FP32 with its FP64 norm.
    // Mpc^dag r_s only restarts the inner recursion, which runs in FP32: it is
    // applied by the inner operator (into ml, free between iterations)
    if ((rc = ensure_tmp(hi, 3, N, s))) return rc;
    void *aD = hi->tmp[0];
    wck::fold_site_clover(hi->g, N, b->pd, clov_inv(hi, b->pd), 1.0 / hi->diag,
                          b->h16 ? b->st + 13 * N : nullptr, b->xl, b->xh, aD, hi->open_bc ? 1 : 0, s,
                          clov_roff(hi));
    if ((rc = mpc_impl<double>(hi, N, b->pd, 0, b->mp, b->xh, nullptr, s, aD, b->partials))) return rc;
    wck::resid_d2f(n, N, b->bs, b->mp, b->rl, b->partials, s);
    wck::reduce_final_rhs(b->partials, wck::kRedBlocks, N, b->st + 8 * N, s);
    if ((rc = mpc_impl<float>(lo, N, b->pd, 1, b->ml, b->rl, nullptr, s, nullptr, b->partials))) return rc;
    wck::dot_partials_rhs<float>(n, N, b->ml, b->ml, b->partials, nullptr, s);
  } else {
    wck::axpy<double>(n, 1.0, 0.0, b->xl, b->xh, b->xh, s);
    wck::zero<double>(n, b->xl, s);
    if ((rc = mpc_impl<double>(hi, N, b->pd, 0, b->mp, b->xh, nullptr, s, nullptr, b->partials))) return rc;
    wck::axpy<double>(n, -1.0, 0.0, b->mp, b->bs, b->rs, s);
    wck::dot_partials_rhs<double>(n, N, b->rs, b->rs, b->partials, nullptr, s);
    wck::reduce_final_rhs(b->partials, wck::kRedBlocks, N, b->st + 8 * N, s);
    if ((rc = mpc_impl<double>(hi, N, b->pd, 1, b->rh, b->rs, nullptr, s, nullptr, b->partials))) return rc;
    wck::dot_partials_rhs<double>(n, N, b->rh, b->rh, b->partials, nullptr, s);
  }
  wck::reduce_final_rhs(b->partials, wck::kRedBlocks, N, b->st + 9 * N, s);
  WC_LAUNCHED();
  WC_HIP(hipMemcpyAsync(b->host + 8 * N, b->st + 8 * N, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  ++b->updates;
  for (int i = 0; i < N; ++i) {
    const double t2 = b->host[8 * N + i], c = b->host[9 * N + i];
    double sn = 1.0;
    if (b->active[i]) {
      if (!std::isfinite(t2) || !std::isfinite(c))
        WC_FAIL(WC_ERR_NUMERIC, "wc_batch: true residual of slot %d became %g", i, t2);
      b->tres[i] = b->src2[i] > 0 ? std::sqrt(t2 / b->src2[i]) : 0.0;
      b->ratio[i] = c > 0.0 ? t2 / c : 0.0;
      if (t2 <= b->tol[i] * b->tol[i] * b->src2[i]) {
        b->active[i] = 0;
        b->finished[i] = 1;
        b->conv[i] = 1;
      } else if (b->iters[i] >= b->maxiter[i]) {
        b->active[i] = 0;
        b->finished[i] = 1;
        b->conv[i] = 0;
      }
      if (b->h16 && b->active[i] && c > 0.0) sn = std::sqrt(12.0 * (double)hi->g.Vh / c);
    }
    b->host[11 * N + i] = sn;
    b->host[12 * N + i] = sn / b->scale[i];
    b->host[13 * N + i] = 1.0 / sn;
    b->scale[i] = sn;
    b->maxr[i] = c * sn * sn;
    if (b->active[i]) {
      if (b->since[i] >= 2) b->rate[i] = decay_rate(b->cref[i], b->clast[i], b->since[i], b->rate[i]);
      b->cref[i] = b->clast[i] = b->maxr[i];
      b->since[i] = 0;
    }
    // c = s^2 |r_h|^2 for the slots that go on, active flags (st rows 0 and 5)
    b->host[i] = b->active[i] ? b->maxr[i] : 0.0;
    b->host[5 * N + i] = b->active[i] ? 1.0 : 0.0;
  }
  if (b->h16) {
    WC_HIP(hipMemcpyAsync(b->st + 11 * N, b->host + 11 * N, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
    wck::scale_copy_rhs(n, N, b->st + 11 * N, b->ml, b->rl, s);
    wck::scale_rhs_half(n, N, b->st + 12 * N, b->pl, b->h16, s);
    wck::scale_rhs_half(n, N, b->st + 12 * N, b->aD, b->h16, s);
  } else if (single) {
    WC_HIP(hipMemcpyAsync(b->rl, b->ml, hbytes(lo) * N, hipMemcpyDeviceToDevice, s));
  } else {
    WC_HIP(hipMemcpyAsync(b->rl, b->rh, hbytes(hi) * N, hipMemcpyDeviceToDevice, s));
  }
  WC_HIP(hipMemcpyAsync(b->st, b->host, sizeof(double) * N, hipMemcpyHostToDevice, s));
  WC_HIP(hipMemcpyAsync(b->st + 5 * N, b->host + 5 * N, sizeof(double) * N, hipMemcpyHostToDevice, s));
  b->dirty = true;
  return WC_OK;
}

extern "C" {

int wc_batch_create(wc_operator *op, wc_operator *inner, int parity, int nslots, double delta, int inner_storage,
                    void *stream, wc_batch **out) {
  if (!op || !inner || !out) WC_FAIL(WC_ERR_ARG, "wc_batch_create: null argument");
  *out = nullptr;
  if (op->prec != WC_DOUBLE) WC_FAIL(WC_ERR_ARG, "wc_batch_create: the outer operator must be FP64");
  if (op == inner) WC_FAIL(WC_ERR_ARG, "wc_batch_create: inner must be a separate operator object");
  for (int mu = 0; mu < 4; ++mu)
    if (op->g.L[mu] != inner->g.L[mu]) WC_FAIL(WC_ERR_ARG, "wc_batch_create: operators of different lattices");
  if (parity != 0 && parity != 1) WC_FAIL(WC_ERR_ARG, "wc_batch_create: parity must be 0 or 1");
  if (!(delta > 0.0 && delta < 1.0)) WC_FAIL(WC_ERR_ARG, "wc_batch_create: need 0 < delta < 1");
  const int code = storage_code(inner_storage);
  if (code < 0) WC_FAIL(WC_ERR_ARG, "wc_batch_create: inner_storage %d unknown", inner_storage);
  if (code && inner->prec != WC_SINGLE)
    WC_FAIL(WC_ERR_ARG, "wc_batch_create: 16-bit storage needs an FP32 inner operator");
  int rc = check_nrhs(op, nslots, "wc_batch_create");
  if (rc) return rc;
  wc_batch *b = new wc_batch();
  b->hi = op;
  b->lo = inner;
  b->N = nslots;
  b->pd = parity;
  b->h16 = code;
  b->delta = std::max(delta, s16_delta_floor(code));
  b->s = (hipStream_t)stream;
  const int N = nslots;
  const size_t hh = hbytes(op) * N, hl = hbytes(inner) * N;
  auto alloc = [&](void **q, size_t bytes) {
    if (rc == WC_OK && wc_dev_malloc(q, bytes) != hipSuccess) {
      g_err = "wc_batch_create: hipMalloc failed";
      rc = WC_ERR_HIP;
    }
  };
  for (void **q : {&b->xh, &b->bs, &b->mp, &b->rs, &b->rh}) alloc(q, hh);
  for (void **q : {&b->rl, &b->pl, &b->ml, &b->xl, &b->aD}) alloc(q, hl);
  b->src.assign(N, nullptr);
  for (int i = 0; i < N; ++i) alloc(&b->src[i], 2 * hbytes(op));
  for (void *&q : b->one) alloc(&q, hbytes(op));
  for (void *&q : b->lone) alloc(&q, hbytes(inner));
  alloc((void **)&b->st, sizeof(double) * 16 * N);
  alloc((void **)&b->partials, sizeof(double) * std::max(wck::kRedBlocks, wck::hop_blocks(op->g)) * N);
  if (rc == WC_OK && hipHostMalloc((void **)&b->host, sizeof(double) * 16 * N, hipHostMallocDefault) != hipSuccess) {
    g_err = "wc_batch_create: hipHostMalloc failed";
    rc = WC_ERR_HIP;
  }
  if (rc == WC_OK) rc = ensure_scratch(b);
  if (rc == WC_OK) rc = maybe_tune(op, N, b->s);
  if (rc == WC_OK) rc = maybe_tune(inner, N, b->s);
  if (rc != WC_OK) {
    std::string keep = g_err;
    free_batch(b);
    g_err = keep;
    return rc;
  }
  // every slot starts empty: zero fields and scalars
  const long n = op->g.Vh * 12 * N;
  hipStream_t s = b->s;
  for (void *q : {b->xh, b->bs, b->mp, b->rs, b->rh}) wck::zero<double>(n, q, s);
  for (void *q : {b->rl, b->pl, b->ml, b->xl, b->aD}) {
    if (inner->prec == WC_SINGLE) wck::zero<float>(n, q, s);
    else wck::zero<double>(n, q, s);
  }
  std::memset(b->host, 0, sizeof(double) * 16 * N);
  for (int i = 0; i < N; ++i) b->host[13 * N + i] = 1.0;  // 1 / s of the empty slots
  if (hipMemcpyAsync(b->st, b->host, sizeof(double) * 16 * N, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    free_batch(b);
    WC_FAIL(WC_ERR_HIP, "wc_batch_create: initialising the slots failed");
  }
  b->occupied.assign(N, 0);
  b->active.assign(N, 0);
  b->finished.assign(N, 0);
  b->conv.assign(N, 0);
  b->iters.assign(N, 0);
  b->maxiter.assign(N, 0);
  b->tol.assign(N, 0.0);
  b->src2.assign(N, 0.0);
  b->maxr.assign(N, 0.0);
  b->scale.assign(N, 1.0);
  b->ratio.assign(N, 0.0);
  b->tres.assign(N, 0.0);
  b->cref.assign(N, 0.0);
  b->clast.assign(N, 0.0);
  b->rate.assign(N, 0.0);
  b->since.assign(N, 0);
  *out = b;
  return WC_OK;
}

int wc_batch_destroy(wc_batch *b) {
  free_batch(b);
  return WC_OK;
}

}  // extern "C"

// admission of one right-hand side; mass (optional) is the slot's own mass
static int admit_impl(wc_batch *b, int slot, const void *src, const void *guess, double tol, long maxiter,
                      const double *mass) {
  if (!b || !src) WC_FAIL(WC_ERR_ARG, "wc_batch_admit: null argument");
  if (slot < 0 || slot >= b->N) WC_FAIL(WC_ERR_ARG, "wc_batch_admit: slot %d out of range [0, %d)", slot, b->N);
  if (b->occupied[slot]) WC_FAIL(WC_ERR_ARG, "wc_batch_admit: slot %d is occupied", slot);
  if (!(tol > 0.0) || maxiter < 0) WC_FAIL(WC_ERR_ARG, "wc_batch_admit: need tol > 0 and maxiter >= 0");
  if (mass && !std::isfinite(*mass)) WC_FAIL(WC_ERR_ARG, "wc_batch_admit_mass: mass %g", *mass);
  wc_operator *hi = b->hi, *lo = b->lo;
  const int N = b->N, pd = b->pd, pc = 1 - pd;
  hipStream_t s = b->s;
  int rc;
  if ((rc = ensure_scratch(b))) return rc;
  // a mass other than the operators' makes the batch one of several masses
  if (mass && *mass != hi->mass && (rc = enable_masses(b))) return rc;
  if (b->mm && (rc = set_slot_mass(b, slot, mass ? *mass : hi->mass))) return rc;
  SlotClover use(b, slot);
  const size_t hb = hbytes(hi);
  const long n1 = hi->g.Vh * 12;
  void *t1 = b->one[0], *bs1 = b->one[1], *y1 = b->one[2], *mp1 = b->one[3], *rs1 = b->one[4], *rh1 = b->one[5];
  void *pl1 = b->lone[0], *aD1 = b->lone[1];
  auto half = [&](const void *f, int p) { return (const void *)((const char *)f + (size_t)p * hb); };
  WC_HIP(hipMemcpyAsync(b->src[slot], src, 2 * hb, hipMemcpyDefault, s));
  // b_s = R src = src_D - Meooe A_C^-1 src_C
  site_op<double>(hi, 1, pc, 1, half(src, pc), t1, nullptr, s);
  {
    HopLaunch a = hop_args<double>(hi, 1, pd, 0, wck::EPI_AXPY, t1, bs1);
    a.y = half(src, pd);
    WC_HOP(double, a, s);
  }
  // y = L^-1 guess = A_D guess_D, or 0 (A_D of the slot's mass)
  if (guess && b->mm && b->mass[slot] != hi->mass) {
    void *aD = nullptr;
    const size_t sh = (size_t)36 * hi->g.Vh * hi->csz;
    WC_HIP(wc_dev_malloc(&aD, sh));
    // the slot's inverse set is written again with the values it holds
    // (the same build), in stream order with every kernel that reads it
    void *h0 = (void *)set_hi(b, 0, b->set_of[slot]), *h1 = (void *)set_hi(b, 1, b->set_of[slot]);
    rc = build_mass_blocks(hi, b->mass[slot], h0, h1, pd == 0 ? aD : nullptr, pd == 1 ? aD : nullptr, b->err, s);
    if (rc == WC_OK)
      wck::site_clover<double>(hi->g, 1, pd, aD, 0.0, half(guess, pd), y1, nullptr, hi->open_bc ? 1 : 0, s);
    const hipError_t e = hipStreamSynchronize(s);
    (void)wc_dev_free_on(aD, s);
    if (rc) return rc;
    if (e != hipSuccess) WC_FAIL(WC_ERR_HIP, "wc_batch_admit: %s", hipGetErrorString(e));
  } else if (guess) {
    site_op<double>(hi, 1, pd, 0, half(guess, pd), y1, nullptr, s);
  } else {
    wck::zero<double>(n1, y1, s);
  }
  // r_s = b_s - Mpc y ; r_h = Mpc^dag r_s
  if ((rc = mpc_impl<double>(hi, 1, pd, 0, mp1, y1, nullptr, s))) return rc;
  wck::axpy<double>(n1, -1.0, 0.0, mp1, bs1, rs1, s);
  if ((rc = mpc_impl<double>(hi, 1, pd, 1, rh1, rs1, nullptr, s))) return rc;
  // |src|^2 (both parities), |r_s|^2, |r_h|^2
  double *st1 = hi->st;
  if ((rc = norm2_rhs<double>(hi, 1, src, st1 + 0, s))) return rc;
  // the full field is two half fields: add the second half's norm
  if ((rc = norm2_rhs<double>(hi, 1, half(src, 1), st1 + 1, s))) return rc;
  if ((rc = norm2_rhs<double>(hi, 1, rs1, st1 + 2, s))) return rc;
  if ((rc = norm2_rhs<double>(hi, 1, rh1, st1 + 3, s))) return rc;
  double h4[4];
  WC_HIP(hipMemcpyAsync(h4, st1, sizeof h4, hipMemcpyDeviceToHost, s));
  WC_HIP(hipStreamSynchronize(s));
  const double s2 = h4[0] + h4[1], t2 = h4[2], c = h4[3];
  if (!std::isfinite(t2) || !std::isfinite(c)) WC_FAIL(WC_ERR_NUMERIC, "wc_batch_admit: initial residual is %g", t2);
  const bool done = (t2 <= tol * tol * s2) || maxiter == 0;
  // the slot's frame: unit RMS per component of the inner residual (16-bit storage)
  const double sn = (b->h16 && !done && c > 0.0) ? std::sqrt(12.0 * (double)hi->g.Vh / c) : 1.0;
  // into the slot: y, b_s; inner r = p = s r_h, x = 0, A_D^-1 p
  k_to_slot<double2, double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)y1, (double2 *)b->xh);
  k_to_slot<double2, double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)bs1, (double2 *)b->bs);
  if (b->h16) {
    k_to_slot_scaled<float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, sn, (const double2 *)rh1, (float2 *)b->rl);
    // p as 16-bit in a single-RHS field, its A_D^-1 from that (as the batch computes it)
    if (b->h16 == kS16Bf16)
      k_to_slot_scaled<c2b><<<nblk(n1), 256, 0, s>>>(n1, 1, 0, sn, (const double2 *)rh1, (c2b *)pl1);
    else
      k_to_slot_scaled<c2h><<<nblk(n1), 256, 0, s>>>(n1, 1, 0, sn, (const double2 *)rh1, (c2h *)pl1);
    site_op<float>(lo, 1, pd, 1, pl1, aD1, nullptr, s, b->h16);
    // 16-bit entries are moved as 4-byte words
    k_to_slot<uint32_t, uint32_t><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const uint32_t *)pl1, (uint32_t *)b->pl);
    k_zero_slot<float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (float2 *)b->xl);
    k_to_slot<uint32_t, uint32_t><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const uint32_t *)aD1, (uint32_t *)b->aD);
  } else if (lo->prec == WC_SINGLE) {
    wck::convert_d2f(n1, rh1, pl1, s);
    site_op<float>(lo, 1, pd, 1, pl1, aD1, nullptr, s);
    k_to_slot<float2, float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const float2 *)pl1, (float2 *)b->rl);
    k_to_slot<float2, float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const float2 *)pl1, (float2 *)b->pl);
    k_zero_slot<float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (float2 *)b->xl);
    k_to_slot<float2, float2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const float2 *)aD1, (float2 *)b->aD);
  } else {
    site_op<double>(lo, 1, pd, 1, rh1, aD1, nullptr, s);
    k_to_slot<double2, double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)rh1, (double2 *)b->rl);
    k_to_slot<double2, double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)rh1, (double2 *)b->pl);
    k_zero_slot<double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (double2 *)b->xl);
    k_to_slot<double2, double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)aD1, (double2 *)b->aD);
  }
  WC_LAUNCHED();
  b->occupied[slot] = 1;
  b->finished[slot] = 0;
  b->conv[slot] = 0;
  b->iters[slot] = 0;
  b->maxiter[slot] = maxiter;
  b->tol[slot] = tol;
  b->src2[slot] = s2;
  b->scale[slot] = sn;
  b->ratio[slot] = c > 0.0 ? t2 / c : 0.0;
  b->maxr[slot] = c * sn * sn;
  b->cref[slot] = b->clast[slot] = b->maxr[slot];
  b->rate[slot] = 0.0;
  b->since[slot] = 0;
  b->dirty = true;
  b->tres[slot] = s2 > 0 ? std::sqrt(t2 / s2) : 0.0;
  b->active[slot] = done ? 0 : 1;
  if (done) {
    b->finished[slot] = 1;
    b->conv[slot] = (t2 <= tol * tol * s2) ? 1 : 0;
  }
  // the slot's inner CG scalars (c = s^2 |r_h|^2, active flag) and its frame (1 / s for the fold)
  double cs[3] = {done ? 0.0 : c * sn * sn, done ? 0.0 : 1.0, 1.0 / sn};
  WC_HIP(hipMemcpyAsync(b->st + 0 * N + slot, &cs[0], sizeof(double), hipMemcpyHostToDevice, s));
  WC_HIP(hipMemcpyAsync(b->st + 5 * N + slot, &cs[1], sizeof(double), hipMemcpyHostToDevice, s));
  WC_HIP(hipMemcpyAsync(b->st + 13 * N + slot, &cs[2], sizeof(double), hipMemcpyHostToDevice, s));
  WC_HIP(hipStreamSynchronize(s));
  return WC_OK;
}

extern "C" {

int wc_batch_set_mass_sets(wc_batch *b, int nsets) {
  if (!b) WC_FAIL(WC_ERR_ARG, "wc_batch_set_mass_sets: null batch");
  if (nsets < 1 || nsets > b->N) WC_FAIL(WC_ERR_ARG, "wc_batch_set_mass_sets: nsets %d not in [1, %d]", nsets, b->N);
  if (b->mm) WC_FAIL(WC_ERR_ARG, "wc_batch_set_mass_sets: the batch already holds several masses");
  b->nsets = nsets;
  return WC_OK;
}

int wc_batch_admit(wc_batch *b, int slot, const void *src, const void *guess, double tol, long maxiter) {
  return admit_impl(b, slot, src, guess, tol, maxiter, nullptr);
}

int wc_batch_admit_mass(wc_batch *b, int slot, double mass, const void *src, const void *guess, double tol,
                        long maxiter) {
  return admit_impl(b, slot, src, guess, tol, maxiter, &mass);
}

int wc_batch_step(wc_batch *b, int iters, int *finished, int *nfinished) {
  if (!b) WC_FAIL(WC_ERR_ARG, "wc_batch_step: null batch");
  if (iters < 0) WC_FAIL(WC_ERR_ARG, "wc_batch_step: iters must be >= 0");
  SlotClover use(b, -1);
  const int N = b->N;
  hipStream_t s = b->s;
  wc_operator *lo = b->lo;
  int rc;
  auto count_active = [&]() {
    int k = 0;
    for (int i = 0; i < N; ++i) k += b->active[i];
    return k;
  };
  // Iterations run in chunks: wck::iter_check (mode 1) decides on the
  // device when a reliable update is due -- a slot's iterated |r|^2 below
  // delta^2 times its maximum since its restart, its iteration limit, or a
  // predicted convergence (ratio |r|^2 / s^2 <= tol^2 |src|^2) -- and the
  // rest of the run then does nothing.
  IterRun &ir = b->run;
  if ((rc = ensure_run(ir))) return rc;
  const double d2 = b->delta * b->delta;
  auto one = [&](hipStream_t q) -> int {
    lo->halt_cur = ir.ctl + wck::kCtlHalt;
    const int rc2 = lo->prec == WC_SINGLE
                        ? inner_iteration<float>(lo, N, b->pd, b->rl, b->pl, b->ml, b->xl, b->aD, q, b->st,
                                                 b->partials, b->h16)
                        : inner_iteration<double>(lo, N, b->pd, b->rl, b->pl, b->ml, b->xl, b->aD, q, b->st,
                                                  b->partials);
    lo->halt_cur = nullptr;
    if (rc2) return rc2;
    wck::iter_check(N, 1, d2, b->st, ir.dat, ir.ctl, ir.hist, q);
    WC_LAUNCHED();
    return WC_OK;
  };
  std::vector<uintptr_t> key = {2, (uintptr_t)N, (uintptr_t)b->pd, (uintptr_t)b->h16, (uintptr_t)b->rl,
                                (uintptr_t)b->pl, (uintptr_t)b->ml, (uintptr_t)b->xl, (uintptr_t)b->aD,
                                (uintptr_t)b->st, (uintptr_t)b->partials};
  op_key(lo, N, key);
  const int cap = chunk_cap(lo, N, b->hi->chunk);
  int it = 0;
  while (it < iters && count_active() > 0) {
    if (b->dirty) {
      double *d = ir.hdat;
      for (int i = 0; i < N; ++i) {
        d[0 * wck::kMaxRhs + i] = b->tol[i] * b->tol[i] * b->src2[i];
        d[1 * wck::kMaxRhs + i] = b->maxr[i];
        d[2 * wck::kMaxRhs + i] = b->ratio[i];
        d[3 * wck::kMaxRhs + i] = b->scale[i];
        d[4 * wck::kMaxRhs + i] = (double)b->iters[i];
        d[5 * wck::kMaxRhs + i] = (double)b->maxiter[i];
      }
      WC_HIP(hipMemcpyAsync(ir.dat, ir.hdat, IterRun::kDatBytes, hipMemcpyHostToDevice, s));
      b->dirty = false;
    }
    // the run ends at the first slot that triggers an update
    double need = 1e30;
    for (int i = 0; i < N; ++i) {
      if (!b->active[i]) continue;
      const double r_ = decay_rate(b->cref[i], b->clast[i], b->since[i], b->rate[i]);
      double t = iters_to(b->clast[i], d2 * b->maxr[i], r_);
      if (b->ratio[i] > 0.0) {
        const double t2 = iters_to(b->clast[i], b->tol[i] * b->tol[i] * b->src2[i] * b->scale[i] * b->scale[i] /
                                                    b->ratio[i], r_);
        if (t2 >= 0.0 && (t < 0.0 || t2 < t)) t = t2;
      }
      if (t < 0.0) t = kFirstRun;
      need = std::min({need, t, (double)(b->maxiter[i] - b->iters[i])});
    }
    const int K = pow2_floor(std::min(need, (double)(iters - it)), cap);
    if ((rc = run_iterations(ir, key, K, N, s, one))) return rc;
    const int done = ir.hctl[wck::kCtlIter];
    for (int j = 0; j < done; ++j)
      for (int i = 0; i < N; ++i) {
        if (!b->active[i]) continue;
        ++b->iters[i];
        ++b->since[i];
        const double cp = ir.hhist[(size_t)j * N + i];
        if (!std::isfinite(cp)) WC_FAIL(WC_ERR_NUMERIC, "wc_batch: residual of slot %d became %g", i, cp);
        if (cp > b->maxr[i]) b->maxr[i] = cp;
        b->clast[i] = cp;
      }
    it += done;
    if (ir.hctl[wck::kCtlTrig] && (rc = batch_reliable(b))) return rc;
  }
  int k = 0;
  for (int i = 0; i < N; ++i) {
    if (finished) finished[i] = b->finished[i];
    k += b->finished[i];
  }
  if (nfinished) *nfinished = k;
  return WC_OK;
}

int wc_batch_extract(wc_batch *b, int slot, void *dst, long *iterations, double *true_residual, int *converged) {
  if (!b || !dst) WC_FAIL(WC_ERR_ARG, "wc_batch_extract: null argument");
  if (slot < 0 || slot >= b->N) WC_FAIL(WC_ERR_ARG, "wc_batch_extract: slot %d out of range", slot);
  if (!b->occupied[slot] || !b->finished[slot]) WC_FAIL(WC_ERR_ARG, "wc_batch_extract: slot %d has no finished solve", slot);
  SlotClover use(b, slot);
  wc_operator *hi = b->hi;
  const int N = b->N, pd = b->pd, pc = 1 - pd;
  hipStream_t s = b->s;
  int rc;
  if ((rc = ensure_scratch(b))) return rc;
  const size_t hb = hbytes(hi);
  const long n1 = hi->g.Vh * 12;
  void *y1 = b->one[2], *t1 = b->one[0];
  auto half = [&](void *f, int p) { return (void *)((char *)f + (size_t)p * hb); };
  k_from_slot<double2><<<nblk(n1), 256, 0, s>>>(n1, N, slot, (const double2 *)b->xh, (double2 *)y1);
  // dst = L y + S src: dst_D = A_D^-1 y ; dst_C = A_C^-1 src_C - A_C^-1 Meooe dst_D
  site_op<double>(hi, 1, pd, 1, y1, half(dst, pd), nullptr, s);
  site_op<double>(hi, 1, pc, 1, half(b->src[slot], pc), t1, nullptr, s);
  {
    HopLaunch a = hop_args<double>(hi, 1, pc, 0, wck::EPI_CLOV | wck::EPI_AXPY, half(dst, pd), half(dst, pc));
    a.clov = clov_inv(hi, pc);
    a.y = t1;
    WC_HOP(double, a, s);
  }
  WC_LAUNCHED();
  WC_HIP(hipStreamSynchronize(s));
  if (iterations) *iterations = b->iters[slot];
  if (true_residual) *true_residual = b->tres[slot];
  if (converged) *converged = b->conv[slot];
  b->occupied[slot] = 0;
  b->finished[slot] = 0;
  return WC_OK;
}

int wc_batch_slots(const wc_batch *b, int *occupied, int *active) {
  if (!b) WC_FAIL(WC_ERR_ARG, "wc_batch_slots: null batch");
  for (int i = 0; i < b->N; ++i) {
    if (occupied) occupied[i] = b->occupied[i];
    if (active) active[i] = b->active[i];
  }
  return WC_OK;
}

}  // extern "C"
