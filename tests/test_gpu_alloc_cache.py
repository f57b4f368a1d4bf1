"""The library's device blocks are cached by size when freed
(wc_release_cached_memory): an operator released and created again on the
same lattice takes its blocks back from the cache, allocating no new device
memory, and computes the same bits; flushing the cache returns the blocks
to the runtime (the device's free memory grows by what was reported).
Frees on one thread do not cut graph captures of CG runs on another."""
import ctypes

import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import _lib  # noqa: E402

FD = [8, 8, 8, 16]
P = dict(mass=0.1, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])


def release_cache():
    n = ctypes.c_long()
    _lib.check(_lib.load().wc_release_cached_memory(ctypes.byref(n)))
    torch.cuda.synchronize()
    return n.value


def meooe(U, src):
    w = g.qcd.fermion.wilson_clover(U, P)
    s = g.lattice(w.F_grid)
    s.from_numpy(src)
    h = g.pick_checkerboard(g.odd, s)
    out = w.Meooe(h).data.clone()
    torch.cuda.synchronize()
    return w, out


def test_released_operator_blocks_are_reused_then_flushed():
    rng = R.random("alloc_cache")
    U = G.random_links(R.grid(FD), rng, scale=0.5)
    src = rng.vspincolor(R.grid(FD))
    release_cache()
    w, ref = meooe(U, src)
    w.release()  # its blocks go to the cache, not to the runtime
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    w, out = meooe(U, src)  # the same sizes again: served from the cache
    free1, _ = torch.cuda.mem_get_info()
    assert torch.equal(out, ref)
    links_bytes = 2 * (int(np.prod(FD)) // 2) * 8 * 9 * 16
    assert free0 - free1 < links_bytes, (free0, free1)
    w.release()
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    released = release_cache()
    free3, _ = torch.cuda.mem_get_info()
    assert released >= links_bytes
    assert free3 - free2 >= links_bytes
    assert release_cache() == 0


def test_parked_operators_are_destroyed_by_release_pool():
    """an operator dropped without release() parks its handle for reuse
    (gpt's operator pool, interface.py:23-80); release_pool destroys the
    parked handles, and a new operator then computes the same bits"""
    import gc

    from gpt_amd.qcd.fermion import operator as fop

    rng = R.random("alloc_pool")
    U = G.random_links(R.grid(FD), rng, scale=0.5)
    src = rng.vspincolor(R.grid(FD))
    w, ref = meooe(U, src)
    del w
    gc.collect()
    assert fop.release_pool() >= 1
    assert fop.release_pool() == 0
    w, out = meooe(U, src)
    assert torch.equal(out, ref)
    w.release()


def test_frees_on_one_thread_do_not_cut_captures_on_another():
    """one thread solves in fresh batches whose iteration runs are captured as
    graphs (every batch captures anew) while another creates and destroys
    operators (device-wide waits and frees): every solve converges to the
    same bits as the first"""
    import threading

    from gpt_amd import serve
    from oracle import wilson_clover as W

    rng = R.random("alloc_capture")
    U = G.random_links(R.grid(FD), rng, scale=0.5)
    srcs = [rng.vspincolor(R.grid(FD)) for _ in range(2)]
    errors, sols = [], []
    stop = threading.Event()

    def solver():
        try:
            w = g.qcd.fermion.wilson_clover(U, P)
            w.set_iteration_chunk(8)
            inner = w.converted(g.single)
            for _ in range(4):
                b = serve.batch(w, inner, 4)
                for a in srcs:
                    x = g.lattice(w.F_grid)
                    x.from_numpy(a)
                    b.admit(x, 1e-10, 10000)
                done = {}
                for _ in range(2000):
                    for sl in b.step(8):
                        x = g.lattice(w.F_grid)
                        it, tres, conv = b.extract(sl, x)
                        assert conv and tres <= 1e-10
                        done[sl] = x.to_numpy()
                    if len(done) == 2:
                        break
                b.close()
                sols.append([done[k] for k in sorted(done)])
            w.release()
            inner.release()
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(repr(e))
        finally:
            stop.set()

    def churn():
        try:
            while not stop.is_set():
                w = g.qcd.fermion.wilson_clover(U, dict(P, mass=0.2))
                w.release()
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    t1, t2 = threading.Thread(target=solver), threading.Thread(target=churn)
    t1.start()
    t2.start()
    t1.join(timeout=120)
    stop.set()
    t2.join(timeout=60)
    assert not t1.is_alive() and not t2.is_alive()
    assert errors == []
    assert len(sols) == 4
    ref = W.wilson_clover(U, **P)
    for k, a in enumerate(srcs):
        x = sols[0][k]
        assert np.linalg.norm(a - ref.M(x)) / np.linalg.norm(a) <= 1.05e-10
        for s in sols[1:]:
            assert np.array_equal(s[k], x)
