"""Serving loop for propagator solves on resident gauge fields.

No gpt counterpart: gpt solves right-hand sides one after another
(lib/gpt/core/operator/matrix_operator.py:288-295, container.py:330-357) on
one operator at a time.  Here requests (a gauge field, a mass, 1..12
right-hand sides, a tolerance) arrive at any time and are served by
continuous batching:

* a *group* per resident field owns an FP64 / FP32 operator pair and a
  ``batch`` of solve slots (``wc_batch_*``: the mixed-precision eo2_ne CG
  advancing all occupied slots with the multi-RHS kernels), sized to the
  right-hand sides waiting when it starts (a multiple of 4, at most
  ``nslots``); requests of any mass share it: each slot uses clover-inverse
  blocks for its request's mass (``wc_batch_admit_mass``; slots of one mass
  share one set), so same-field requests are batched whatever their masses;
* right-hand sides enter free slots of their group whenever slots free up,
  while the group keeps iterating, and leave when their FP64 true residual
  |src - M dst| / |src| meets the request's tolerance;
* a new group is created only if its HBM footprint fits the budget left
  (``group_bytes``); otherwise the request waits, and idle groups are
  released (their device blocks stay in the library's cache, so the next
  group of the same lattice and size allocates nothing);
* groups step on their own HIP streams from worker threads (ctypes releases
  the GIL), so the work of different groups overlaps on the device.

Every result is a solve to the request's tolerance, whatever else shares its
batch: alone or in a mix, solutions agree to within the solver tolerance
(the reliable-update points depend on the batch, the converged solutions do
not beyond that).  ``latency`` is arrival to result on the host.
"""
import collections
import concurrent.futures as cf
import ctypes
import time

import numpy as np
import torch

from . import _lib, core
from .core import lattice


class batch:
    """One wc_batch: nslots slots of the mixed-precision eo2_ne solve on the
    operator pair (op FP64, inner in the inner precision).  inner_storage
    "half" (the default for an FP32 inner operator; IEEE FP16) or "bf16"
    keeps the inner search direction and Schur intermediates in 16-bit
    storage.  mass_sets: clover-inverse sets the batch holds once it serves
    several masses (None: one per slot); with fewer, it holds at most that
    many masses at a time (``admit`` of one more raises)."""

    def __init__(self, op, inner, nslots, parity=core.odd, delta=0.01, inner_storage="auto", mass_sets=None):
        self.op, self.inner = op, inner
        self.nslots = int(nslots)
        if inner_storage == "auto":
            inner_storage = "half" if inner.precision is core.single else None
        if inner_storage not in _lib.STORAGE_CODES:
            raise ValueError("inner_storage must be 'auto', None, 'half' or 'bf16'")
        code = _lib.STORAGE_CODES[inner_storage]
        h = ctypes.c_void_p()
        _lib.check(_lib.load().wc_batch_create(op.handle, inner.handle, parity.tag, self.nslots, float(delta), code,
                                               core.stream_ptr(), ctypes.byref(h)))
        self.handle = h.value
        if mass_sets is not None:
            rc = _lib.load().wc_batch_set_mass_sets(self.handle, max(1, min(int(mass_sets), self.nslots)))
            if rc:
                self.close()
                _lib.check(rc)
        self.free = list(range(self.nslots))
        self.slot_of = {}

    def admit(self, src, tol, maxiter, guess=None, mass=None):
        """put src into a free slot; mass: the slot's own mass (None: the
        operators'), so one batch serves several masses"""
        slot = self.free.pop(0)
        gp = guess.ptr() if guess is not None else None
        if mass is None:
            rc = _lib.load().wc_batch_admit(self.handle, slot, src.ptr(), gp, float(tol), int(maxiter))
        else:
            rc = _lib.load().wc_batch_admit_mass(self.handle, slot, float(mass), src.ptr(), gp, float(tol),
                                                 int(maxiter))
        if rc:
            self.free.insert(0, slot)
            _lib.check(rc)
        return slot

    def step(self, iters):
        fin = np.zeros(self.nslots, dtype=np.int32)
        nf = ctypes.c_int()
        _lib.check(_lib.load().wc_batch_step(self.handle, int(iters), fin.ctypes.data, ctypes.byref(nf)))
        return [i for i in range(self.nslots) if fin[i]]

    def extract(self, slot, dst):
        it = ctypes.c_long()
        tres = ctypes.c_double()
        conv = ctypes.c_int()
        _lib.check(_lib.load().wc_batch_extract(self.handle, slot, dst.ptr(), ctypes.byref(it), ctypes.byref(tres),
                                                ctypes.byref(conv)))
        self.free.append(slot)
        self.free.sort()
        return int(it.value), float(tres.value), bool(conv.value)

    def occupied(self):
        occ = np.zeros(self.nslots, dtype=np.int32)
        _lib.check(_lib.load().wc_batch_slots(self.handle, occ.ctypes.data, None))
        return int(occ.sum())

    def close(self):
        if self.handle is not None:
            _lib.load().wc_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class request:
    """A propagator solve: ``sources`` are full FP64 fields (one right-hand
    side each) on the grid of the resident gauge field ``field``."""

    _ids = 0

    def __init__(self, field, mass, sources, tol=1e-10, maxiter=100000):
        request._ids += 1
        self.id = request._ids
        self.field = field
        self.mass = float(mass)
        self.sources = list(sources)
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        n = len(self.sources)
        self.results = [None] * n
        self.iterations = [0] * n
        self.true_residual = [0.0] * n
        self.converged = [False] * n
        self.remaining = n
        self.t_arrival = None
        self.t_done = None

    @property
    def done(self):
        return self.remaining == 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


def group_bytes(fdims, nslots, masses=True, mass_sets=None):
    """HBM held by one group: both operators (links -- 12-real in the inner
    one --, clover and inverse, the FP64 gauge copy each keeps, solver
    scratch), the batch state and, for a batch of several masses, its sets
    of clover inverses (both parities, FP64 and FP32; mass_sets of them, or
    one per slot)."""
    V = int(np.prod(fdims))
    Vh = V // 2
    h64, h32 = Vh * 12 * 16, Vh * 12 * 8
    op64 = 2 * Vh * 8 * 9 * 16 + 4 * Vh * 36 * 16 + V * 4 * 9 * 16 + h64 * (3 * nslots + 9)
    op32 = 2 * Vh * 8 * 6 * 8 + 4 * Vh * 36 * 8 + V * 4 * 9 * 16 + h32 * (3 * nslots + 5)
    state = 5 * h64 * nslots + 5 * h32 * nslots + 2 * h64 * nslots
    nsets = nslots if mass_sets is None else min(nslots, int(mass_sets))
    sets = nsets * 2 * Vh * 36 * (16 + 8) if masses else 0
    return op64 + op32 + state + sets


class _group:
    def __init__(self, server, key, U, mass, stream, nslots):
        self.server = server
        self.key = key
        self.U = U
        self.mass = mass  # the operators' mass (the first request's); slots carry their own
        self.stream = stream
        self.setup_s = None  # wall time of creating the operators and the batch
        self.queue = collections.deque()  # (request, rhs index)
        self.in_slot = {}  # slot -> (request, rhs index)
        self.batch = None
        self.op = self.inner = None
        self.nslots = nslots
        self.bytes = group_bytes(U[0].grid.fdimensions, nslots, mass_sets=server.mass_sets)

    def busy(self):
        return bool(self.queue) or bool(self.in_slot)

    def work(self, iters):
        """admit what fits, run iters inner iterations, hand back finished solves
        (runs on a worker thread, on this group's stream)"""
        from .qcd import fermion

        done = []
        with torch.cuda.stream(self.stream):
            if self.batch is None:
                t0 = time.perf_counter()
                p = dict(self.server.params, mass=self.mass)
                self.op = fermion.wilson_clover(self.U, p)
                # the inner operator only drives the recursion between FP64
                # reliable updates: 12-real links change the speed, not the result
                self.inner = fermion.wilson_clover(self.U, dict(p, compress_links=True),
                                                   precision=self.server.inner_precision)
                if self.server.iteration_chunk is not None:
                    self.op.set_iteration_chunk(self.server.iteration_chunk)
                self.batch = batch(self.op, self.inner, self.nslots, delta=self.server.delta,
                                   inner_storage=self.server.inner_storage, mass_sets=self.server.mass_sets)
                torch.cuda.current_stream().synchronize()
                self.setup_s = time.perf_counter() - t0
            cap = self.server.mass_sets
            while self.queue and self.batch.free:
                req, j = self.queue[0]
                # a batch holds at most mass_sets masses at a time: a new
                # mass waits (in arrival order) until one leaves
                if cap is not None:
                    used = {r.mass for r, _ in self.in_slot.values()}
                    if req.mass not in used and len(used) >= min(cap, self.nslots):
                        break
                self.queue.popleft()
                slot = self.batch.admit(req.sources[j], req.tol, req.maxiter,
                                        mass=None if req.mass == self.mass else req.mass)
                self.in_slot[slot] = (req, j)
            finished = self.batch.step(iters) if self.in_slot else []
            for slot in finished:
                req, j = self.in_slot.pop(slot)
                dst = lattice(self.op.F_grid)
                it, tres, conv = self.batch.extract(slot, dst)
                done.append((req, j, dst, it, tres, conv))
        return done

    def release(self):
        if self.batch is not None:
            self.batch.close()
        for o in (self.inner, self.op):
            if o is not None:
                o.release()
        self.batch = self.op = self.inner = None


class server:
    """Continuous-batching solve server (see the module notes).

    params: operator parameters without the mass (csw_r, csw_t, boundary
    phases, ...); nslots: slots per group; streams: HIP streams the groups
    are spread over; iters_per_step: inner iterations a group runs before the
    scheduler looks again; hbm_budget: bytes the groups may hold (default 80%
    of the device memory free at start, less the library's block cache);
    mass_sets: the clover-inverse sets of a group's batch (at most that many
    masses in a batch at a time; None: one set per slot)."""

    def __init__(self, params, nslots=12, streams=4, iters_per_step=25, hbm_budget=None,
                 inner_precision=core.single, delta=0.01, inner_storage="auto", iteration_chunk=None,
                 min_slots=1, mass_sets=4):
        self.params = {k: v for k, v in dict(params).items() if k not in ("mass", "kappa")}
        self.nslots = int(nslots)
        self.min_slots = max(1, min(int(min_slots), self.nslots))  # smallest batch a group opens with
        self.mass_sets = None if mass_sets is None else max(2, int(mass_sets))
        self.iters_per_step = int(iters_per_step)
        self.inner_precision = inner_precision
        self.delta = float(delta)
        self.inner_storage = inner_storage
        # cap on the inner iterations per captured graph run (None: automatic; 1: no graphs)
        self.iteration_chunk = iteration_chunk
        if hbm_budget is None:
            # 80 % of the free memory, less what the library may keep cached
            # (an eighth of the device, at most 32 GB)
            free, total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free) - min(total // 8, 32 << 30)
        self.hbm_budget = int(hbm_budget)
        self.hbm_used = 0
        self.fields = {}
        self.groups = {}
        self.streams = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]
        self._next_stream = 0
        self.pool = cf.ThreadPoolExecutor(max_workers=len(self.streams))
        self.completed = []
        self.peak_groups = 0
        self.waited_for_memory = 0
        self.setup_times = []  # per group: seconds to create its operators and batch
        self.warmed = set()

    def add_field(self, key, U, warmup=True):
        """keep a gauge field resident: 4 link fields (gpt_amd.qcd.gauge) or
        natural-layout numpy arrays, held in HBM in FP64.  warmup: tune the
        hop launch shapes of this lattice size now, for batches of every size
        class up to nslots, so that no request waits for the tuner"""
        from .qcd import gauge

        if not isinstance(U[0], gauge.link_field):
            fd = list(reversed(np.asarray(U[0]).shape[:4]))
            U = gauge.from_numpy(core.grid(fd, core.double), U)
        self.fields[key] = U
        if warmup:
            self.warm(U)

    def warm(self, U):
        """tune the FP64 and inner operators' hops for this lattice size and
        batches of 1, 2-3, 4-7, ... right-hand sides (up to nslots); the tuned
        shapes are kept process-wide per lattice size, precision, link format
        and class, so it runs once per lattice size"""
        from .qcd import fermion

        fd = tuple(U[0].grid.fdimensions)
        if fd in self.warmed:
            return
        Vh = int(np.prod(fd)) // 2
        p = dict(self.params, mass=0.1)
        ops = [fermion.wilson_clover(U, p),
               fermion.wilson_clover(U, dict(p, compress_links=True), precision=self.inner_precision)]
        n = 1
        while n <= self.nslots:
            # below 2^17 site-RHS pairs solves do not tune (launch shapes matter little)
            if Vh * n >= (1 << 17):
                for op in ops:
                    op.tune_hop(n_rhs=n, reps=2)
            n *= 2
        for op in ops:
            op.release()
        self.warmed.add(fd)

    # one group per resident field: its batch has as many slots as the
    # request that opens it has right-hand sides, rounded up to a multiple of 4
    # (the matrix-core hop's unit) and at most nslots -- a slot costs the same
    # whether or not it is occupied; later requests of the field, of any mass,
    # take slots as they free up
    def _route(self, req):
        key = req.field
        g = self.groups.get(key)
        if g is None:
            U = self.fields[req.field]
            nslots = max(self.min_slots, min(self.nslots, 4 * ((len(req.sources) + 3) // 4)))
            need = group_bytes(U[0].grid.fdimensions, nslots, mass_sets=self.mass_sets)
            if self.hbm_used + need > self.hbm_budget:
                return False
            g = _group(self, key, U, req.mass, self.streams[self._next_stream % len(self.streams)], nslots)
            self._next_stream += 1
            self.groups[key] = g
            self.hbm_used += g.bytes
            self.peak_groups = max(self.peak_groups, len(self.groups))
        g.queue.extend((req, j) for j in range(len(req.sources)))
        return True

    def run(self, schedule, on_done=None):
        """serve requests arriving at the given offsets (seconds from now);
        schedule: list of (offset, request); on_done(request) is called as
        each request completes.  Returns the requests in completion order."""
        schedule = sorted(schedule, key=lambda x: x[0])
        t0 = time.perf_counter()
        pending = collections.deque()
        busy = {}
        i = 0
        order = []
        while True:
            now = time.perf_counter()
            while i < len(schedule) and t0 + schedule[i][0] <= now:
                req = schedule[i][1]
                req.t_arrival = t0 + schedule[i][0]
                pending.append(req)
                i += 1
            waiting = collections.deque()
            while pending:
                req = pending.popleft()
                if not self._route(req):
                    waiting.append(req)
            if waiting:
                self.waited_for_memory += 1
            pending = waiting
            for key, g in self.groups.items():
                if key not in busy and g.busy():
                    busy[key] = self.pool.submit(g.work, self.iters_per_step)
            if not busy:
                if i >= len(schedule) and not pending:
                    break
                if pending:
                    raise RuntimeError("a request needs more HBM than the budget holds")
                time.sleep(max(0.0, min(0.01, t0 + schedule[i][0] - time.perf_counter())))
                continue
            timeout = None
            if i < len(schedule):
                timeout = max(0.0, t0 + schedule[i][0] - time.perf_counter())
            fin, _ = cf.wait(list(busy.values()), timeout=timeout, return_when=cf.FIRST_COMPLETED)
            for key in [k for k, f in busy.items() if f in fin]:
                f = busy.pop(key)
                for req, j, dst, it, tres, conv in f.result():
                    req.results[j] = dst
                    req.iterations[j] = it
                    req.true_residual[j] = tres
                    req.converged[j] = conv
                    req.remaining -= 1
                    if req.done:
                        req.t_done = time.perf_counter()
                        order.append(req)
                        if on_done is not None:
                            on_done(req)
                g = self.groups[key]
                if not g.busy():
                    if g.setup_s is not None:
                        self.setup_times.append(g.setup_s)
                    # hipFree synchronises the device, which a graph capture
                    # running on another worker thread does not survive: free
                    # only once the running steps have ended (their results
                    # are collected on the next pass)
                    if busy:
                        cf.wait(list(busy.values()))
                    g.release()
                    self.hbm_used -= g.bytes
                    del self.groups[key]
        self.completed.extend(order)
        return order

    def close(self):
        for g in self.groups.values():
            g.release()
        self.groups.clear()
        self.pool.shutdown(wait=True)


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
