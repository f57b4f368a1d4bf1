/*
 * gpt_amd_wc.h -- C ABI of the MI355X even-odd Wilson-clover operator and
 * its CG solve.
 *
 * This is the drop-in for the fermion-operator FFI that gpt's Python layer
 * calls through cgpt (paths relative to the gpt repository root):
 *   lib/cgpt/lib/operators.cc:34-133      create / update / set_mass /
 *                                          delete / apply_fermion_operator
 *   lib/cgpt/lib/operators/register.h:1-20 the opcode numbers (kept as is)
 *   lib/cgpt/lib/transform.cc:143-246      lattice_axpy, lattice_norm2,
 *                                          lattice_rank_inner_product
 *   lib/cgpt/lib/foundation/transfer.h:27-75  pick / set checkerboard
 * The caller is lib/gpt/qcd/fermion/operator/interface.py:29-103.
 *
 * Conventions
 *   - All pointers to fields are DEVICE pointers owned by the caller; the
 *     operator owns its links, clover blocks and scratch.
 *   - A half field (one parity) is 12 x Vh complex values in blocks of 64
 *     sites: element (k, h) at index ((h/64)*12 + k)*64 + h%64, with
 *     k = 3*spin + colour and half-site h = (((t*Lz + z)*Ly + y)*Lx + x)/2.
 *     A full field is two half fields, even parity first.  Complex values
 *     are (re, im) pairs of double (WC_DOUBLE) or float (WC_SINGLE).  The
 *     half volume Vh must be a multiple of 64 (V a multiple of 128), and a
 *     half field must fit in 4 GB.
 *   - Links given to wc_create/wc_update are FP64 complex, natural order
 *     [4][V][3][3], site index ((t*Lz + z)*Ly + y)*Lx + x, WITHOUT boundary
 *     phases (the operator applies them, lib/gpt/core/covariant.py:28-37).
 *   - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *     asynchronous except those that return a reduction to the host.
 *   - Every call returns WC_OK or an error code; wc_last_error() gives the
 *     message (the Python layer raises RuntimeError with it, mirroring
 *     lib/cgpt/lib/exception.h:23-39).  Not thread safe per operator.
 */
#ifndef GPT_AMD_WC_H
#define GPT_AMD_WC_H

#ifdef __cplusplus
extern "C" {
#endif

#define WC_OK 0
#define WC_ERR_ARG 1     /* bad argument: sizes, parities, aliasing, opcode */
#define WC_ERR_HIP 2     /* a HIP runtime call failed */
#define WC_ERR_NUMERIC 3 /* e.g. a near-singular clover block */

#define WC_DOUBLE 0
#define WC_SINGLE 1

typedef struct wc_operator wc_operator;

/* Parameters of g.qcd.fermion.wilson_clover (lib/gpt/qcd/fermion/wilson.py:63-96,
 * parsed in lib/cgpt/lib/operators/wilson_clover.h:28-39).  kappa is
 * converted to mass = 1/(2 kappa) - 4 by the caller, as wilson.py:78-80. */
typedef struct wc_params {
  int fdims[4];                  /* Lx, Ly, Lz, Lt: all even, >= 2 */
  double mass;                   /* m0 */
  double csw_r, csw_t;           /* clover coefficients (csw_r is divided by xi_0) */
  double cF;                     /* open-boundary improvement coefficient */
  double xi_0, nu;               /* anisotropy */
  int isAnisotropic;
  double boundary_phases[4][2];  /* (re, im) per direction; [3] = 0 -> open boundary */
  int precision;                 /* WC_DOUBLE or WC_SINGLE (link/clover/field storage) */
  int compress_links;            /* 1: store 12 reals per link, rebuild row 2 (SU(3) only) */
} wc_params;

typedef struct wc_solver_stats {
  long iterations;  /* CG iterations (gpt's len(cg.history)) */
  int converged;    /* 1 if |r|^2 <= eps^2 |b|^2 was reached */
  double r2;        /* final iterated |r|^2 */
  double b2;        /* |b|^2 */
} wc_solver_stats;

/* --- operator life cycle (cgpt operators.cc:34-94) ----------------------- */

/* create_fermion_operator (operators.cc:34-57, wilson_clover.h:19-71):
 * builds the double-stored links, the clover blocks A and A^-1. */
int wc_create(const wc_params *p, const void *U, void *stream, wc_operator **out);
/* update_fermion_operator (operators.cc:59-70, implementation.h:77-86) */
int wc_update(wc_operator *op, const void *U, void *stream);
/* set_mass_fermion_operator (operators.cc:72-83, implementation.h:21-29);
 * rebuilds A and A^-1 for the new mass */
int wc_set_mass(wc_operator *op, double mass, void *stream);
/* delete_fermion_operator (operators.cc:85-94) */
int wc_destroy(wc_operator *op);

/* geometry / storage queries */
long wc_half_volume(const wc_operator *op);
int wc_precision(const wc_operator *op);

/* apply_fermion_operator (operators.cc:96-107, operators/unary.h:36-48,
 * register.h:1-20).  Opcodes:
 *   full fields : 2001 M, 2002 Mdag, 2009 Mdiag, 2010/2011 Dminus(Dag) = copy,
 *                 2012-2015 Import/Export = copy, 3001 Dhop, 4001 DhopDag
 *   half fields : 2003 Meooe, 2004 MeooeDag, 2005 Mooee, 2006 MooeeDag,
 *                 2007 MooeeInv, 2008 MooeeInvDag, 3002 DhopEO, 4002 DhopEODag
 * src_cb: parity of the half-field src (0 even, 1 odd); ignored for full
 * fields.  Meooe maps parity src_cb to 1 - src_cb; DhopEO needs src_cb = 1.
 * dst must not alias src (reference operators assert dst != src). */
int wc_apply(wc_operator *op, int opcode, void *dst, const void *src, int src_cb, void *stream);

/* --- BLAS-1 on n complex entries (cgpt transform.cc:143-246) ------------- */
/* lattice_axpy: z = a x + y */
int wc_axpy(int prec, long n, double a_re, double a_im, const void *x, const void *y, void *z,
            void *stream);
/* z = a x */
int wc_scale(int prec, long n, double a_re, double a_im, const void *x, void *z, void *stream);
/* lattice_norm2: |x|^2 accumulated in FP64, result on the host */
int wc_norm2(int prec, long n, const void *x, double *result, void *stream);
/* lattice_rank_inner_product: <a,b> = sum conj(a) b in FP64, result[2] = (re, im) */
int wc_inner_product(int prec, long n, const void *a, const void *b, double *result, void *stream);
/* axpy_norm2 (lib/gpt/core/transform.py:151-153): z = a x + y, returns |z|^2 */
int wc_axpy_norm2(int prec, long n, double a_re, double a_im, const void *x, const void *y, void *z,
                  double *result, void *stream);
/* g.convert between FP64 and FP32 storage (transform.cc:128-141) */
int wc_convert(int dst_prec, void *dst, int src_prec, const void *src, long n, void *stream);

/* natural [V][4][3] <-> even-odd full field (layout plumbing for I/O) */
int wc_field_import(const int fdims[4], int prec, const void *natural, void *eo, void *stream);
int wc_field_export(const int fdims[4], int prec, const void *eo, void *natural, void *stream);

/* --- solvers ----------------------------------------------------------- */
/* CG on the eo2_ne normal operator Mpc^dag Mpc with D = `parity`
 * (lib/gpt/algorithms/inverter/cg.py:48-112 over
 *  lib/gpt/algorithms/preconditioner/schur_complement_two.py:87-112 and
 *  normal_equation.py:44-45).  x: initial guess in, solution out (half
 * field of parity `parity`); b: right-hand side.  history (optional, maxiter
 * doubles) receives |r|^2 per iteration.  Same recursion and stopping rule
 * as gpt: stop when |r|^2 <= eps^2 |b|^2 on the iterated residual. */
int wc_cg_eo2_ne(wc_operator *op, int parity, void *x, const void *b, double eps, long maxiter,
                 double *history, wc_solver_stats *stats, void *stream);

/* inv.preconditioned(pc.eo2_ne(parity), inv.cg) applied to one full-field
 * right-hand side (lib/gpt/algorithms/inverter/preconditioned.py:32-54):
 * dst = L CG(R_ne src, guess L^-1 dst) + S src.  dst holds the initial
 * guess on entry (zero it for gpt's default). */
int wc_solve_eo2_ne(wc_operator *op, int parity, void *dst, const void *src, double eps, long maxiter,
                    double *history, wc_solver_stats *stats, void *stream);

/* the Schur pieces on half fields (schur_complement_two.py:87-148,
 * normal_equation.py:44-48): which = 0 Mpc, 1 Mpc^dag, 2 Mpc^dag Mpc */
int wc_apply_mpc(wc_operator *op, int parity, int which, void *dst, const void *src, void *stream);

/* --- multiple right-hand sides ------------------------------------------
 * gpt's n_rhs > 1 operator (lib/gpt/qcd/fermion/wilson.py:82-86, a 5-d grid
 * with the RHS as the fastest index, operator/base.py:47-60) applies one
 * operator to nrhs fields at once, reusing every link and clover block for
 * all of them.  A field of nrhs right-hand sides keeps the half-field layout
 * with the RHS innermost: element (k, h, r) at
 * (((h/64)*12 + k)*64 + h%64)*nrhs + r.  1 <= nrhs <= 64.
 * The CG runs nrhs independent recursions of cg.py (each its own scalars
 * and stopping test, so iteration counts equal the single-RHS ones); arrays
 * iterations / converged / r2 / b2 hold nrhs entries (each may be NULL),
 * history maxiter * nrhs doubles (may be NULL). */
int wc_apply_block(wc_operator *op, int opcode, int nrhs, void *dst, const void *src, int src_cb, void *stream);
int wc_apply_mpc_block(wc_operator *op, int parity, int which, int nrhs, void *dst, const void *src,
                       void *stream);
int wc_cg_eo2_ne_block(wc_operator *op, int parity, int nrhs, void *x, const void *b, double eps, long maxiter,
                       double *history, long *iterations, int *converged, double *r2, double *b2,
                       void *stream);
int wc_solve_eo2_ne_block(wc_operator *op, int parity, int nrhs, void *dst, const void *src, double eps,
                          long maxiter, double *history, long *iterations, int *converged, double *r2,
                          double *b2, void *stream);

/* wc_apply_mpc_block without allocating.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same kernels
 * and bitwise the same results as wc_apply_mpc_block, with the scratch half
 * fields in a fixed 64 MiB array per GPU (held from the library's first use on
 * that GPU until the process ends): which 0 and 2 need two fields, which 1
 * one, of 12 * Vh * nrhs complex values; a larger need is refused
 * (WC_ERR_ARG) before anything runs.  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams or threads) race. */
int wc_apply_mpc_block_noalloc(wc_operator *op, int parity, int which, int nrhs, void *dst, const void *src,
                               void *stream);

/* Mixed-precision eo2_ne solve with reliable updates (no cgpt counterpart;
 * the precision pattern of lib/gpt/algorithms/inverter/defect_correcting.py
 * and mixed_precision.py, done inside one CG).  `op` is the FP64 operator,
 * `inner` the same operator in the inner precision: a separate operator
 * object (wc_create with precision WC_SINGLE, or WC_DOUBLE).  The CG recursion runs in the inner
 * precision; FP64 keeps the solution, the source terms and the true
 * residuals.  When a right-hand side's iterated |r|^2 has fallen by delta^2
 * since the last reliable update, the inner solution is added to the FP64
 * one and r = b - A x is recomputed in FP64 (the search direction is kept).
 * Stops per right-hand side when the FP64 true residual of the full system
 * satisfies |src - M dst| <= tol |src| (checked at reliable updates).
 * inner_storage WC_STORAGE_HALF (FP32 inner operator only) stores the
 * search direction and the Schur intermediates of the inner recursion as
 * IEEE 16-bit floats (FP32 arithmetic), each right-hand side rescaled to
 * unit RMS at every reliable update (delta is then at least 1e-2: FP16
 * resolves about 5e-4 of a field's scale); WC_STORAGE_BF16 stores them as
 * bfloat16 the same way (delta at least 1e-1: BF16 resolves about 4e-3);
 * WC_STORAGE_NATIVE keeps the inner precision.
 * dst holds the initial guess on entry.  iterations / converged /
 * true_residual: nrhs entries (may be NULL); reliable_updates: one long. */
#define WC_STORAGE_NATIVE 0
#define WC_STORAGE_HALF 2
#define WC_STORAGE_BF16 3
int wc_solve_eo2_ne_mixed(wc_operator *op, wc_operator *inner, int parity, int nrhs, void *dst, const void *src,
                          double tol, long maxiter, double delta, int inner_storage, long *iterations,
                          int *converged, double *true_residual, long *reliable_updates, void *stream);

/* Block CG on Mpc^dag Mpc (lib/gpt/algorithms/inverter/block_cg.py:36-149,
 * which gpt runs in Python on lists of fields): the nrhs right-hand sides of
 * one multi-RHS half field b are solved together over their block Krylov
 * space, with the Cholesky-QR factorisation of the residual block each
 * iteration.  Stops when every right-hand side has |r|^2 <= eps^2 |b|^2, or
 * (eps_abs > 0) |r|^2 <= eps_abs^2 for all, after at least miniter
 * iterations.  x holds the guess and gets the solution.  history (may be
 * NULL; maxiter doubles) gets the largest relative |r|^2 of the block per
 * iteration.  A block whose residuals are not linearly independent (a zero
 * right-hand side beside non-zero ones, or right-hand sides that converge
 * at very different rates) makes the Cholesky factorisation fail, as in the
 * reference: WC_ERR_NUMERIC.  An all-zero block gives x = 0.
 * wc_solve_eo2_ne_block_cg is the same inside the eo2_ne Schur solve
 * (dst, src: full fields), as wc_solve_eo2_ne_block is for CG. */
int wc_block_cg_eo2_ne(wc_operator *op, int parity, int nrhs, void *x, const void *b, double eps, double eps_abs,
                       long miniter, long maxiter, long *iterations, int *converged, double *history, void *stream);
int wc_solve_eo2_ne_block_cg(wc_operator *op, int parity, int nrhs, void *dst, const void *src, double eps,
                             double eps_abs, long miniter, long maxiter, long *iterations, int *converged,
                             double *history, void *stream);

/* Hop launch shapes (no cgpt counterpart: Grid fixes its kernel launch at
 * compile time).  A shape is five ints {rb, tpb, occ, seg, ls}: rb
 * right-hand sides per thread (1 in FP64, 1..2 in FP32), 64 * tpb threads
 * per workgroup (tpb 1..8), occ (0..32) caps the workgroups per CU by
 * reserving LDS (0: no cap), seg > 1 visits the 64-site blocks in
 * t-marching order with seg runs per time slice (needs the time slice to be
 * a whole number of blocks, seg <= that number), and ls = 1 makes a
 * workgroup copy the links (and clover blocks) of its sites to LDS once, so
 * that the several items of a site read them from there (at most 80 KB of
 * LDS per workgroup), and ls = 2, for one right-hand side in natural order
 * (ignored otherwise), makes it copy the input spinors of its own
 * half-site range to LDS, where the +-x and most +-y neighbours of its
 * sites are, so that those gathers read LDS.  Every shape gives bitwise the
 * same result; only the speed differs.
 *
 * wc_tune_hop times the hop of nrhs right-hand sides (reps calls per shape)
 * for every valid shape and keeps the fastest for this operator and, in a
 * process-wide cache, for every operator of the same dimensions, precision
 * and link format, for right-hand-side counts of the same class (1, 2-3,
 * 4-7, 8-15, 16-31, 32-64).  shape (5 ints) and ms_per_call may be NULL.
 * The CG, solve and batch entry points tune on first use when
 * Vh * nrhs >= 2^17, once per class (environment WC_AUTOTUNE=0 turns tuning
 * off).  The tuner leaves seg at 1.  wc_set_hop_shape fixes a shape (NULL:
 * back to automatic); wc_get_hop_shape reports the one in use. */
int wc_tune_hop(wc_operator *op, int nrhs, int reps, int *shape, float *ms_per_call, void *stream);
int wc_set_hop_shape(wc_operator *op, int nrhs, const int *shape);
int wc_get_hop_shape(const wc_operator *op, int nrhs, int *shape);

/* Matrix cores for the FP16-storage hops (no cgpt counterpart: Grid's 5-d
 * multi-RHS operator, lib/cgpt/lib/operators/wilson_clover.h:58-66, runs its
 * stencil on the vector units).  The hops of a mixed solve or batch with
 * WC_STORAGE_HALF on a multiple of 4 right-hand sides can run as
 * v_mfma_f32_16x16x32_f16 products of the FP16 spinors with the FP32 links
 * split into FP16 hi + lo parts (FP32 accumulation).  mode 0: vector ALUs,
 * 1: matrix cores, 2 (the default): the faster of the two, timed when the
 * solvers tune (first use of a lattice and class of right-hand-side counts,
 * Vh * nrhs >= 2^17), the vector form until then.  The two forms differ by
 * the order of their FP32 sums, not bitwise. */
int wc_set_hop_mma(wc_operator *op, int mode);
/* the form an FP16-storage hop of nrhs right-hand sides takes now: 1 matrix
 * cores, 0 vector ALUs */
int wc_get_hop_mma(const wc_operator *op, int nrhs, int *form);

/* The inner operator's pieces with 16-bit storage, as the mixed solve's inner
 * CG runs them (no cgpt counterpart: gpt's inner operator of
 * lib/gpt/algorithms/inverter/mixed_precision.py:23-32 is a plain FP32 one).
 * FP32 operator only; storage WC_STORAGE_HALF or WC_STORAGE_BF16, each
 * complex value two 16-bit floats (re, im) in the layout of a half field.
 *   wc_apply_block_s16     opcode 2003 / 2004 (Meooe / MeooeDag) of the
 *                          16-bit src (parity src_cb) into the 16-bit dst;
 *   wc_apply_mpc_block_s16 dst (FP32) = Mpc^dag Mpc src (16-bit src, Schur
 *                          intermediates stored in 16 bits), D = parity. */
int wc_apply_block_s16(wc_operator *op, int opcode, int nrhs, int storage, void *dst, const void *src, int src_cb,
                       void *stream);
int wc_apply_mpc_block_s16(wc_operator *op, int parity, int nrhs, int storage, void *dst, const void *src,
                           void *stream);

/* CG iteration schedule (no cgpt counterpart: gpt's cg.py reads the
 * residual on the host after every iteration).  The CG, solve, mixed and
 * batch entry points run their iterations in runs of up to `chunk`: a run is
 * one HIP graph (captured once per run length, operator state and field
 * set, then replayed), in which a device-side check after every iteration
 * applies the stopping rule (CG) or the reliable-update trigger (mixed,
 * batch) and turns the rest of the run into no-ops; the host then replays
 * the run's residual history.  Results, iteration counts and histories do
 * not depend on the schedule.  chunk: 0 automatic (run lengths predicted
 * from the residual decay, capped by the iteration's size), 1 one iteration
 * at a time without graphs, 2..32 a fixed cap.  The setting of the outer
 * (FP64) operator governs its mixed solves and batches.
 * wc_iteration_runs reports how many runs this operator's CG and mixed
 * solves (as the inner operator) made as graphs and one at a time. */
int wc_set_iteration_chunk(wc_operator *op, int chunk);
/* Free the operator's solver scratch (half fields, reduction buffers and
 * captured iteration graphs; reallocated on next use), keeping its links and
 * clover.  gpt parks unused operators for reuse (interface.py:23-80); a
 * parked multi-RHS operator would otherwise hold many GB of scratch.  No
 * cgpt counterpart. */
int wc_trim(wc_operator *op);
int wc_iteration_runs(const wc_operator *op, long *graph_runs, long *direct_runs);
/* The library keeps device blocks it frees (operators, scratch, batches) for
 * a later request of the same size, up to an eighth of the device's memory
 * (at most 32 GB): re-creating an operator or batch then allocates nothing,
 * and freeing one does not synchronise the device.  An allocation that
 * fails returns the cached blocks first and tries again.  This returns
 * every cached block of the current device to the runtime now and reports
 * their size.  No cgpt counterpart. */
int wc_release_cached_memory(long *released_bytes);

/* Continuous-batching solver (no cgpt counterpart: gpt solves right-hand
 * sides one after another, matrix_operator.py:288-295).  A wc_batch holds
 * nslots slots of the mixed-precision eo2_ne solve on one operator pair
 * (op FP64, inner a separate operator object in the inner precision, as for
 * wc_solve_eo2_ne_mixed, including inner_storage) and advances all
 * occupied slots together with the multi-RHS kernels.  Requests enter free slots at any time and leave when
 * done, so the batch keeps running while its membership changes.
 *
 *   wc_batch_admit   puts one right-hand side (src: full FP64 field; guess:
 *                    full FP64 field or NULL for zero) into a free slot, with
 *                    its own tolerance (true relative residual of the full
 *                    system) and iteration limit.
 *   wc_batch_step    runs up to iters inner iterations of the occupied slots
 *                    (fewer if none is left iterating); finished[nslots]
 *                    (may be NULL) flags the slots whose solve is done.
 *   wc_batch_extract writes a finished slot's solution dst = L y + S src
 *                    (full FP64 field) with its statistics, and frees the slot.
 *   wc_batch_slots   reports occupied / iterating slots (arrays may be NULL).
 *
 * Calls on one batch, and on its operators, must not overlap; batches on
 * different operator pairs and streams run concurrently. */
typedef struct wc_batch wc_batch;
int wc_batch_create(wc_operator *op, wc_operator *inner, int parity, int nslots, double delta, int inner_storage,
                    void *stream, wc_batch **out);
int wc_batch_destroy(wc_batch *b);
int wc_batch_admit(wc_batch *b, int slot, const void *src, const void *guess, double tol, long maxiter);
/* the same with the slot's own mass (no cgpt counterpart; gpt re-creates or
 * re-masses an operator per mass, interface.py:41-48).  The first mass
 * other than the operators' gives the batch one set of clover-inverse
 * blocks per slot (both parities, FP64 and the inner precision: 36 Vh
 * (16 + inner bytes) x 2 per set), built from the operators' gauge field
 * for a mass; slots of one mass share a set.  From then on every
 * clover-inverse application of the batch uses each slot's set, so requests
 * of any mass on one gauge field share one batch.  A slot admitted with
 * wc_batch_admit then gets the operators' mass. */
int wc_batch_admit_mass(wc_batch *b, int slot, double mass, const void *src, const void *guess, double tol,
                        long maxiter);
/* the number of clover-inverse sets a batch of several masses holds (no
 * cgpt counterpart), 1 <= nsets <= nslots, before its first admission at
 * another mass (default: nslots).  With fewer sets than slots the batch
 * holds at most nsets masses at a time: an admission whose mass would need
 * one more set returns WC_ERR_ARG and leaves the batch as it was, so a
 * caller keeps such a request waiting until a mass leaves. */
int wc_batch_set_mass_sets(wc_batch *b, int nsets);
int wc_batch_step(wc_batch *b, int iters, int *finished, int *nfinished);
int wc_batch_extract(wc_batch *b, int slot, void *dst, long *iterations, double *true_residual, int *converged);
int wc_batch_slots(const wc_batch *b, int *occupied, int *active);

/* --- halo of a T-slab decomposition -------------------------------------
 * Grid exchanges the faces of a cut lattice inside its stencil, with
 * communication overlapped with the interior
 * (lib/cgpt/lib/operators/wilson_clover.h:45, overlapCommsCompute); gpt's
 * FFI has no call for it.  These two calls are the kernels one GPU runs for
 * that exchange when the lattice is cut in t and this operator's lattice is
 * one rank's slab (periodic in t, so on one GPU the rank above and below is
 * the GPU itself).  For the hop (dag 0) or its adjoint (dag 1) of output
 * parity parity_out on nrhs right-hand sides (field layout as for
 * wc_apply_block), with nf = Lx Ly Lz / 2 face sites:
 *   wc_halo_pack reads src (parity 1 - parity_out) and writes
 *     face_lo: (1 -+ gamma_t) psi of the t = 0 slice (for the rank below),
 *     face_hi: U_t^dag (1 +- gamma_t) psi of the t = Lt - 1 slice, hop
 *              coefficient and boundary phase folded in (for the rank above);
 *   wc_halo_unpack adds to dst (parity parity_out) the t-terms that
 *     face_lo (received from above) gives its t = Lt - 1 slice and face_hi
 *     (received from below) its t = 0 slice.
 * A face buffer holds 6 * nf * nrhs complex values of the operator's
 * precision: entry k = 3 * spin + colour of face site f (spatial half-site
 * index) and right-hand side r at (k * nf + f) * nrhs + r.  A hop whose
 * t-links on the last slice are zero, plus wc_halo_unpack of
 * wc_halo_pack(src), is the hop of src.  Not for open boundaries. */
int wc_halo_pack(wc_operator *op, int parity_out, int dag, int nrhs, const void *src, void *face_lo, void *face_hi,
                 void *stream);
int wc_halo_unpack(wc_operator *op, int parity_out, int dag, int nrhs, const void *face_lo, const void *face_hi,
                   void *dst, void *stream);
/* the rest of the rank's hop: dst (parity parity_out) = the hop of src
 * without the t-terms of the t = Lt - 1 slice's forward and the t = 0
 * slice's backward neighbour (those wc_halo_unpack adds), so that
 * wc_halo_interior then wc_halo_unpack of wc_halo_pack(src) is the hop of
 * src (Meooe / MeooeDag, opcodes 2003 / 2004).  Not for open boundaries. */
int wc_halo_interior(wc_operator *op, int parity_out, int dag, int nrhs, void *dst, const void *src, void *stream);

/* Gauge observables of natural-layout FP64 links U[4][V][3][3] (site
 * ((t*Lz + z)*Ly + y)*Lx + x, periodic): the average plaquette
 * (1/(18 V)) sum_x sum_{mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
 * and the average link trace (1/(12 V)) sum_{x,mu} Re tr U_mu(x), as gpt's
 * g.qcd.gauge.plaquette and the NERSC reader's LINK_TRACE check compute them
 * (lib/gpt/qcd/gauge/loops.py, lib/gpt/core/io/nersc_io.py).  Used by the
 * NERSC configuration reader and writer (gpt_amd/io.py). */
int wc_gauge_observables(const int fdims[4], const void *U, double *plaquette, double *link_trace, void *stream);

/* HBM bandwidth probe (no counterpart in the reference; a measurement
 * utility): kind 0 copies `bytes` from src to dst, kind 1 only reads src (and
 * writes a 32 KB XOR fold to dst, so bytes >= 32768); both stream 16 bytes
 * per lane with non-temporal accesses over a fixed grid.  Timed by the
 * caller, it gives the practical bandwidth ceiling of the device next to the
 * 8 TB/s peak.  src and dst: 16-byte aligned device buffers of `bytes`. */
int wc_bandwidth_probe(int kind, const void *src, void *dst, long bytes, void *stream);

/* message of the last failed call on this thread */
const char *wc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
