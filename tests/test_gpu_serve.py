"""Continuous batching (wc_batch_*) and the serving loop (gpt_amd.serve).

Checks, with the oracle's FP64 operator as the referee:
* every solve leaves the batch with a true residual |src - M x| / |src| at
  or below its tolerance, evaluated by the numpy oracle;
* a right-hand side admitted while others are mid-solve, and one solved in a
  batch of others, agree with the same solve done alone to 1e-8 relative
  (both are solves to tol = 1e-10; the reliable-update points depend on the
  batch, so the solutions agree to within a few condition numbers times tol);
* the server serves requests for two resident fields and several masses,
  with an HBM budget small enough that requests must wait for memory;
* requests of different masses on one field share one batch (one group per
  field): each slot carries its own clover inverses, every solve meets its
  tolerance through the oracle's M at its own mass and agrees with its solve
  alone -- with clover, and for plain Wilson fermions; slots of one mass
  share one set, and a batch of fewer sets than slots holds that many
  masses at a time;
* fields of different volumes are served side by side;
* a request served alone and among requests of other fields gives bitwise
  the same solutions.
"""
import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import serve  # noqa: E402

P = dict(csw_r=1.0, csw_t=1.0, cF=1.0, xi_0=1.0, nu=1.0, isAnisotropic=False, boundary_phases=[1, 1, 1, -1])
TOL = 1e-10
AGREE = 1e-8


@pytest.fixture(scope="module")
def setup():
    gr = R.grid([8, 8, 8, 16])
    rng = R.random("serve")
    U = [G.random_links(gr, rng, scale=0.5), G.random_links(gr, rng, scale=0.4)]
    srcs = [rng.vspincolor(gr) for _ in range(8)]
    return gr, U, srcs


def field(arr, grid):
    x = g.lattice(grid)
    x.from_numpy(arr)
    return x


def true_res(ref, x, src):
    return np.linalg.norm(src - ref.M(x)) / np.linalg.norm(src)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def solve_alone(Uh, mass, src_arr, storage="auto"):
    w = g.qcd.fermion.wilson_clover(Uh, dict(P, mass=mass))
    inner = w.converted(g.single)
    b = serve.batch(w, inner, 1, inner_storage=storage)
    s = field(src_arr, w.F_grid)
    b.admit(s, TOL, 10000)
    while not b.step(50):
        pass
    x = g.lattice(w.F_grid)
    it, tres, conv = b.extract(0, x)
    assert conv and tres <= TOL
    b.close()
    return x.to_numpy()


@pytest.mark.parametrize("storage", [None, "half", "bf16"])
def test_batch_slots_and_mid_flight_admission(setup, storage):
    gr, U, srcs = setup
    mass = 0.1
    ref = W.wilson_clover(U[0], mass=mass, **P)
    w = g.qcd.fermion.wilson_clover(U[0], dict(P, mass=mass))
    inner = w.converted(g.single)
    b = serve.batch(w, inner, 4, inner_storage=storage)
    S = [field(s, w.F_grid) for s in srcs[:4]]
    owner = {}
    owner[b.admit(S[0], TOL, 10000)] = 0
    owner[b.admit(S[1], TOL, 10000)] = 1
    assert sorted(owner) == [0, 1]
    assert b.step(7) == []  # both mid-solve
    owner[b.admit(S[2], TOL, 10000)] = 2  # joins while 0 and 1 iterate
    results = {}
    for _ in range(400):
        for sl in b.step(10):
            x = g.lattice(w.F_grid)
            it, tres, conv = b.extract(sl, x)
            assert conv and 0 < tres <= TOL and it > 0
            results[owner.pop(sl)] = x.to_numpy()
            if 3 not in results and 3 not in owner.values():
                # the freed slot is taken by a new right-hand side at once
                slot = b.admit(S[3], TOL, 10000)
                assert slot == sl
                owner[slot] = 3
        if len(results) == 4:
            break
    assert b.occupied() == 0 and sorted(results) == [0, 1, 2, 3]
    for k in range(4):
        assert true_res(ref, results[k], srcs[k]) <= 1.05 * TOL
        assert rel(results[k], solve_alone(U[0], mass, srcs[k], storage)) < AGREE
    b.close()


def test_batch_reports_iteration_limit_and_zero_source(setup):
    gr, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U[0], dict(P, mass=0.1))
    inner = w.converted(g.single)
    b = serve.batch(w, inner, 2)
    z = g.lattice(w.F_grid)
    z[:] = 0
    b.admit(z, TOL, 100)  # zero source: done at admission
    b.admit(field(srcs[0], w.F_grid), TOL, 5)  # cannot converge in 5 iterations
    fin = []
    for _ in range(10):
        fin = b.step(10)
        if len(fin) == 2:
            break
    assert sorted(fin) == [0, 1]
    x = g.lattice(w.F_grid)
    it0, t0, c0 = b.extract(0, x)
    assert c0 and it0 == 0 and np.abs(x.to_numpy()).max() == 0.0
    it1, t1, c1 = b.extract(1, x)
    assert not c1 and it1 >= 5 and t1 > TOL
    with pytest.raises(RuntimeError, match="no finished solve"):
        b.extract(1, x)
    b.close()


def test_server_mixed_fields_masses_and_memory_wait(setup):
    gr, U, srcs = setup
    Uh = {"a": U[0], "b": U[1]}
    nslots = 4
    # room for one group (4 slots of several masses), not for two
    budget = int(1.05 * serve.group_bytes([8, 8, 8, 16], nslots))
    assert 2 * serve.group_bytes([8, 8, 8, 16], nslots) > budget
    srv = serve.server(P, nslots=nslots, streams=2, iters_per_step=10, hbm_budget=budget)
    for k, u in Uh.items():
        srv.add_field(k, u)
    plan = [("a", 0.1, [0, 1, 2]), ("b", 0.15, [3]), ("a", 0.1, [4]), ("a", 0.2, [5, 6]), ("b", 0.15, [7, 0])]
    reqs = []
    schedule = []
    for n, (f, m, idx) in enumerate(plan):
        grid = g.grid([8, 8, 8, 16], g.double)
        r = serve.request(f, m, [field(srcs[i], grid) for i in idx], tol=TOL, maxiter=10000)
        r.idx = idx
        reqs.append(r)
        schedule.append((0.002 * n, r))
    done = srv.run(schedule)
    srv.close()
    assert sorted(r.id for r in done) == sorted(r.id for r in reqs)
    assert srv.peak_groups == 1 and srv.waited_for_memory > 0
    for r in reqs:
        assert all(r.converged) and r.latency > 0
        ref = W.wilson_clover(Uh[r.field], mass=r.mass, **P)
        for j, i in enumerate(r.idx):
            x = r.results[j].to_numpy()
            assert true_res(ref, x, srcs[i]) <= 1.05 * TOL
            assert rel(x, solve_alone(Uh[r.field], r.mass, srcs[i])) < AGREE
    st = serve.latency_stats(done)
    assert st["p50_s"] <= st["p99_s"]


def test_request_alone_equals_request_among_others(setup):
    """determinism, alone vs mixed: a request served on its own and the same
    request served while requests for another field (two masses, sharing a
    batch) run in their own group on another stream give bitwise the same
    solutions and iteration counts (groups share no arithmetic; launch shapes
    and iteration schedules do not change bits)"""
    gr, U, srcs = setup
    Uh = {"a": U[0], "b": U[1]}
    grid = g.grid([8, 8, 8, 16], g.double)

    def serve_plan(plan):
        srv = serve.server(P, nslots=4, streams=3, iters_per_step=7)
        for k, u in Uh.items():
            srv.add_field(k, u)
        reqs = [serve.request(f, m, [field(srcs[i], grid) for i in idx], tol=TOL, maxiter=10000)
                for f, m, idx in plan]
        srv.run([(0.001 * n, r) for n, r in enumerate(reqs)])
        srv.close()
        return reqs

    alone = serve_plan([("a", 0.1, [0, 1])])[0]
    mixed = serve_plan([("b", 0.15, [3, 4, 5]), ("a", 0.1, [0, 1]), ("b", 0.05, [7])])
    among = mixed[1]
    assert all(r.converged for r in mixed) and all(alone.converged)
    assert alone.iterations == among.iterations
    for x, y in zip(alone.results, among.results):
        assert np.array_equal(x.to_numpy(), y.to_numpy())


@pytest.mark.parametrize("csw", [1.0, 0.0])
def test_masses_share_one_batch(setup, csw):
    """three masses admitted into one 4-slot batch (two while the first is
    mid-solve): each solve meets its tolerance at its own mass and agrees
    with the same solve alone on an operator of that mass"""
    gr, U, srcs = setup
    prm = dict(P, csw_r=csw, csw_t=csw)
    masses = [0.1, 0.05, 0.2]
    w = g.qcd.fermion.wilson_clover(U[0], dict(prm, mass=masses[0]))
    inner = g.qcd.fermion.wilson_clover(U[0], dict(prm, mass=masses[0], compress_links=True), precision=g.single)
    b = serve.batch(w, inner, 4)
    S = [field(srcs[k], w.F_grid) for k in range(3)]
    owner = {b.admit(S[0], TOL, 10000): 0}
    assert b.step(5) == []
    owner[b.admit(S[1], TOL, 10000, mass=masses[1])] = 1
    owner[b.admit(S[2], TOL, 10000, mass=masses[2])] = 2
    results = {}
    for _ in range(400):
        for sl in b.step(10):
            x = g.lattice(w.F_grid)
            it, tres, conv = b.extract(sl, x)
            assert conv and 0 < tres <= TOL
            results[owner.pop(sl)] = x.to_numpy()
        if len(results) == 3:
            break
    b.close()
    assert sorted(results) == [0, 1, 2]
    for k in range(3):
        ref = W.wilson_clover(U[0], mass=masses[k], **prm)
        assert true_res(ref, results[k], srcs[k]) <= 1.05 * TOL, k
        wk = g.qcd.fermion.wilson_clover(U[0], dict(prm, mass=masses[k]))
        bk = serve.batch(wk, wk.converted(g.single), 1)
        bk.admit(field(srcs[k], wk.F_grid), TOL, 10000)
        while not bk.step(50):
            pass
        xk = g.lattice(wk.F_grid)
        bk.extract(0, xk)
        bk.close()
        assert rel(results[k], xk.to_numpy()) < AGREE, k


@pytest.mark.parametrize("mma", [False, True])
def test_repeated_masses_share_clover_sets(setup, mma):
    """slots of one mass share a clover-inverse set: three slots at a mass
    other than the operators' (all occupied slots on one set, which is then
    applied as a single set), then slots of the operators' mass, a repeated
    mass and a new one (sets per slot), and slots admitted into freed sets
    as others finish.  Every solve meets its tolerance through the oracle at
    its own mass and agrees with the same solve alone.  mma: the inner
    16-bit hops on the matrix cores (their per-RHS clover epilogue)."""
    gr, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U[0], dict(P, mass=0.1))
    inner = w.converted(g.single)
    if mma:
        inner.set_hop_mma(True)
        assert inner.hop_mma(8)
    b = serve.batch(w, inner, 8, inner_storage="half")
    S = [field(srcs[k], w.F_grid) for k in range(8)]
    first = [0.05, 0.05, 0.05]
    later = [0.1, 0.2, 0.05]
    last = [0.2, 0.15]
    mass_of, owner, results = {}, {}, {}

    def admit(k, m):
        owner[b.admit(S[k], TOL, 10000, mass=m)] = k
        mass_of[k] = m

    for k, m in enumerate(first):
        admit(k, m)
    assert b.step(5) == []
    for k, m in enumerate(later):
        admit(3 + k, m)
    pending = list(enumerate(last, start=6))
    for _ in range(600):
        for sl in b.step(10):
            x = g.lattice(w.F_grid)
            it, tres, conv = b.extract(sl, x)
            assert conv and 0 < tres <= TOL
            results[owner.pop(sl)] = x.to_numpy()
            if pending:
                admit(*pending.pop(0))
        if len(results) == 8:
            break
    b.close()
    assert sorted(results) == list(range(8))
    for k in range(8):
        ref = W.wilson_clover(U[0], mass=mass_of[k], **P)
        assert true_res(ref, results[k], srcs[k]) <= 1.05 * TOL, k
        assert rel(results[k], solve_alone(U[0], mass_of[k], srcs[k])) < AGREE, k


def test_mass_sets_cap(setup):
    """a batch of two clover-inverse sets holds two masses at a time: a third
    mass is refused while both are in use and admitted once one has left;
    a server whose groups hold two sets serves requests of four masses on
    one field, each checked by the oracle at its own mass"""
    gr, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U[0], dict(P, mass=0.1))
    b = serve.batch(w, w.converted(g.single), 4, mass_sets=2)
    S = [field(srcs[k], w.F_grid) for k in range(4)]
    sa = b.admit(S[0], TOL, 10000)
    sb = b.admit(S[1], TOL, 10000, mass=0.05)
    with pytest.raises(RuntimeError):
        b.admit(S[2], TOL, 10000, mass=0.2)
    assert len(b.free) == 2
    owner = {sa: (0, 0.1), sb: (1, 0.05)}
    results = {}
    waiting = [(2, 0.2), (3, 0.05)]
    for _ in range(600):
        for sl in b.step(10):
            x = g.lattice(w.F_grid)
            it, tres, conv = b.extract(sl, x)
            assert conv and 0 < tres <= TOL
            k, m = owner.pop(sl)
            results[k] = (m, x.to_numpy())
            while waiting:
                k2, m2 = waiting[0]
                if m2 not in {mm for _, mm in owner.values()} and len({mm for _, mm in owner.values()}) >= 2:
                    break
                waiting.pop(0)
                owner[b.admit(S[k2], TOL, 10000, mass=m2)] = (k2, m2)
        if len(results) == 4:
            break
    b.close()
    assert sorted(results) == [0, 1, 2, 3]
    for k, (m, x) in results.items():
        ref = W.wilson_clover(U[0], mass=m, **P)
        assert true_res(ref, x, srcs[k]) <= 1.05 * TOL, k

    srv = serve.server(P, nslots=4, streams=2, mass_sets=2)
    srv.add_field("a", U[0], warmup=False)
    reqs = [serve.request("a", m, [field(srcs[k], w.F_grid)], tol=TOL, maxiter=10000)
            for k, m in enumerate([0.1, 0.05, 0.2, 0.15, 0.05])]
    done = srv.run([(0.0, r) for r in reqs])
    srv.close()
    assert len(done) == 5
    for k, r in enumerate(reqs):
        ref = W.wilson_clover(U[0], mass=r.mass, **P)
        assert r.converged[0] and true_res(ref, r.results[0].to_numpy(), srcs[k]) <= 1.05 * TOL, k


def test_server_one_group_per_field_and_mixed_volumes(setup):
    """requests of four masses on one 8^3 x 16 field and of two masses on an
    8^4 field: one group per field (peak 2), every result checked with the
    oracle at its own mass and volume"""
    gr, U, srcs = setup
    gr4 = R.grid([8, 8, 8, 8])
    rng = R.random("serve_8888")
    U4 = G.random_links(gr4, rng, scale=0.5)
    srcs4 = [rng.vspincolor(gr4) for _ in range(3)]
    srv = serve.server(P, nslots=4, streams=2, iters_per_step=10)
    srv.add_field("big", U[0])
    srv.add_field("small", U4)
    g16, g8 = g.grid([8, 8, 8, 16], g.double), g.grid([8, 8, 8, 8], g.double)
    plan = [("big", 0.1, [0, 1], srcs, g16), ("small", 0.08, [0], srcs4, g8), ("big", 0.03, [2], srcs, g16),
            ("big", 0.2, [3, 4], srcs, g16), ("small", 0.15, [1, 2], srcs4, g8), ("big", 0.12, [5], srcs, g16)]
    reqs = []
    for f, m, idx, pool, grid in plan:
        r = serve.request(f, m, [field(pool[i], grid) for i in idx], tol=TOL, maxiter=10000)
        r.idx, r.pool = idx, pool
        reqs.append(r)
    done = srv.run([(0.003 * n, r) for n, r in enumerate(reqs)])
    srv.close()
    assert len(done) == len(reqs) and srv.peak_groups == 2
    assert all(t > 0 for t in srv.setup_times)
    fields = {"big": U[0], "small": U4}
    for r in reqs:
        assert all(r.converged)
        ref = W.wilson_clover(fields[r.field], mass=r.mass, **P)
        for j, i in enumerate(r.idx):
            assert true_res(ref, r.results[j].to_numpy(), r.pool[i]) <= 1.05 * TOL, (r.field, r.mass, i)
