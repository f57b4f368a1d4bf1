"""CG iterations run in captured HIP graphs with the stopping test on the
device (wc_set_iteration_chunk).  The schedule must not change anything:
solutions are bitwise identical, iteration counts, residual histories and
reliable-update counts equal, between one iteration at a time (host test
after every iteration, no graph) and runs of up to 8 or an automatic number
of iterations replayed from graphs.  Each solve replays its graphs many times
with eager work (reliable updates, host synchronisation) in between, and a
graph captured before a change of the operator (set_mass) is not reused
after it."""
import numpy as np
import pytest

from oracle import gauge as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import serve  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack, unpack  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
P = dict(mass=0.05, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
SCHEDULES = [1, 8, 0]


@pytest.fixture(scope="module")
def setup():
    fd = [8, 8, 8, 16]
    U = G.smooth_links(fd, 0.3, 4001)
    rng = np.random.default_rng(4100)
    shape = tuple(reversed(fd))
    srcs = [(rng.standard_normal(shape + (4, 3)) + 1j * rng.standard_normal(shape + (4, 3))) / np.sqrt(2)
            for _ in range(3)]
    srcs[1] = np.zeros_like(srcs[1])  # a zero source in the batch
    return fd, U, srcs


def _dev(a, grid):
    x = g.lattice(grid)
    x.from_numpy(a)
    return x


def test_cg_bits_independent_of_schedule(setup):
    fd, U, srcs = setup
    out = {}
    for chunk in SCHEDULES:
        w = g.qcd.fermion.wilson_clover(U, P, n_rhs=3)
        w.set_iteration_chunk(chunk)
        cg = inv.cg(eps=1e-10, maxiter=2000)
        S = pack([_dev(s, g.grid(fd, g.double)) for s in srcs])
        x = g(w.propagator(inv.preconditioned(pc.eo2_ne(), cg)) * S)
        runs = w.iteration_runs()
        out[chunk] = ([r.to_numpy() for r in unpack(x)], list(cg.iterations), np.array(cg.history))
        print(f"cg chunk {chunk}: iterations {cg.iterations}, runs (graph, direct) {runs}", flush=True)
        if chunk == 1:
            assert runs[0] == 0 and runs[1] >= max(cg.iterations)
        else:
            assert runs[0] >= 1
            # far fewer host round trips than iterations
            assert runs[0] + runs[1] < max(cg.iterations) / 2
        w.release()
    x1, it1, h1 = out[1]
    assert it1[1] == 0 and it1[0] > 20
    for chunk in SCHEDULES[1:]:
        x, it, h = out[chunk]
        assert it == it1
        assert h.shape == h1.shape and np.array_equal(h, h1)
        for a, b in zip(x, x1):
            assert np.array_equal(a, b), chunk


@pytest.mark.parametrize("storage", [None, "half", "bf16"])
def test_mixed_bits_independent_of_schedule(setup, storage):
    fd, U, srcs = setup
    out = {}
    for chunk in SCHEDULES:
        w = g.qcd.fermion.wilson_clover(U, P, n_rhs=3)
        w.set_iteration_chunk(chunk)
        mcg = inv.cg_mixed(eps=1e-10, maxiter=5000, inner_storage=storage)
        S = pack([_dev(s, g.grid(fd, g.double)) for s in srcs])
        x = g(w.propagator(inv.preconditioned(pc.eo2_ne(), mcg)) * S)
        runs = mcg.inner_operator(w).iteration_runs()
        print(f"mixed {storage} chunk {chunk}: iterations {mcg.iterations}, updates {mcg.reliable_updates}, "
              f"runs {runs}", flush=True)
        assert mcg.converged
        if chunk != 1:
            assert runs[0] >= 1
        out[chunk] = ([r.to_numpy() for r in unpack(x)], list(mcg.iterations), mcg.reliable_updates,
                      list(mcg.true_residual))
        mcg.inner_operator(w).release()
        w.release()
    x1, it1, u1, t1 = out[1]
    for chunk in SCHEDULES[1:]:
        x, it, u, t = out[chunk]
        assert it == it1 and u == u1 and t == t1
        for a, b in zip(x, x1):
            assert np.array_equal(a, b), chunk


def test_batch_bits_independent_of_schedule(setup):
    fd, U, srcs = setup
    res = {}
    for chunk in SCHEDULES:
        w = g.qcd.fermion.wilson_clover(U, P)
        w.set_iteration_chunk(chunk)
        inner = w.converted(g.single)
        b = serve.batch(w, inner, 2, inner_storage="half")
        assert b.admit(_dev(srcs[0], w.F_grid), 1e-10, 10000) == 0
        fin = []
        for steps in range(1, 1000):
            fin = b.step(7)  # step lengths that do not divide the run lengths
            if steps == 3:
                assert fin == []
                assert b.admit(_dev(srcs[2], w.F_grid), 1e-10, 10000) == 1  # mid-flight admission
            if len(fin) == 2:
                break
        assert fin == [0, 1]
        xs = []
        for slot in range(2):
            x = g.lattice(w.F_grid)
            it, tres, conv = b.extract(slot, x)
            assert conv and tres <= 1e-10
            xs.append((x.to_numpy(), it, tres))
        b.close()
        res[chunk] = xs
        print(f"batch chunk {chunk}: iterations {[v[1] for v in xs]}", flush=True)
    for chunk in SCHEDULES[1:]:
        for (a, ia, ta), (b_, ib, tb) in zip(res[chunk], res[1]):
            assert ia == ib and ta == tb
            assert np.array_equal(a, b_)


def test_graph_not_reused_after_set_mass(setup):
    import gc

    from gpt_amd.qcd.fermion import operator as O

    fd, U, srcs = setup

    def solve(op):
        slv = op.propagator(inv.preconditioned(pc.eo2_ne(), inv.cg(eps=1e-10, maxiter=2000)))
        return g(slv * _dev(srcs[0], op.F_grid)).to_numpy()

    w = g.qcd.fermion.wilson_clover(U, P)
    w.set_iteration_chunk(8)
    solve(w)  # captures graphs at mass 0.05
    tag = O.operator_tag[w.handle]
    del w
    gc.collect()
    # the parked handle (and its graphs) is reused with set_mass
    w2 = g.qcd.fermion.wilson_clover(U, dict(P, mass=0.1))
    assert O.operator_tag[w2.handle] == tag
    w2.set_iteration_chunk(8)
    a = solve(w2)
    w3 = g.qcd.fermion.wilson_clover(U, dict(P, mass=0.1))
    w3.set_iteration_chunk(1)
    b = solve(w3)
    assert np.array_equal(a, b)
