"""The hop tuner's results persist across processes, as gpt's auto-tune cache
does (lib/gpt/core/auto_tune.py:20-37,99-136: one JSON file per tuned object
in .gpt_auto_tune/, written after tuning, read at set-up).

Here the directory is WC_TUNE_CACHE (a temporary one per test).  A solve on
a lattice where launch shapes matter (>= 2^17 site-RHS pairs) tunes once and
writes its file; a file already there for the operator's key is used instead
of tuning (a shape planted in it is the one the operator runs with); an
invalid or foreign file is ignored; an empty WC_TUNE_CACHE writes nothing.
The 16-bit hops' shape and form picks persist the same way.
Results never depend on the shape, so none of this changes a solution.
"""
import json
import os

import numpy as np
import pytest

from oracle import gauge as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
# a lattice no other test tunes in this process (the in-process cache would
# answer first): 16^4 in FP32
FD = [16, 16, 16, 16]
P = dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])


@pytest.fixture(scope="module")
def links():
    return G.smooth_links(FD, 0.3, 8001)


def short_solve(U, n, compress=False):
    """a few CG iterations of n right-hand sides (tunes on first use)"""
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=n, precision=g.single, compress_links=compress)
    rng = np.random.default_rng(8100 + n)
    shape = tuple(reversed(FD)) + (4, 3)
    fs = []
    for _ in range(n):
        x = g.lattice(w.F_grid.single_rhs())
        x.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64))
        fs.append(x)
    g(w.propagator(inv.preconditioned(pc.eo2_ne(), inv.cg(eps=1e-6, maxiter=3))) * pack(fs))
    torch.cuda.synchronize()
    return w


def test_tuned_shape_written_then_reused(links, tmp_path, monkeypatch):
    monkeypatch.setenv("WC_TUNE_CACHE", str(tmp_path))
    monkeypatch.delenv("WC_AUTOTUNE", raising=False)
    w8 = short_solve(links, 8)  # 16^4 / 2 sites x 8 right-hand sides = 2^18 pairs: tuned
    files = sorted(f for f in os.listdir(tmp_path) if "_s16" not in f)  # the FP32 hop's file
    assert len(files) == 1 and files[0].endswith("_rhs8.json"), files
    rec = json.load(open(tmp_path / files[0]))
    assert rec["tag"] == files[0][:-5]
    assert tuple(rec["params"]) == w8.hop_shape(8), (rec, w8.hop_shape(8))
    assert "_f32_links18_" in rec["tag"]

    # plant a different valid shape for the class of 16-31 right-hand sides
    planted = [1, 2, 8, 1, 0]
    tag16 = rec["tag"].replace("_rhs8", "_rhs16")
    with open(tmp_path / (tag16 + ".json"), "w") as f:
        json.dump({"tag": tag16, "params": planted, "ms": 1.0}, f)
    w16 = short_solve(links, 16)
    assert w16.hop_shape(16) == tuple(planted)  # read back instead of tuned
    assert json.load(open(tmp_path / (tag16 + ".json")))["params"] == planted  # not re-tuned, not rewritten


def test_foreign_or_invalid_files_are_ignored(links, tmp_path, monkeypatch):
    monkeypatch.setenv("WC_TUNE_CACHE", str(tmp_path))
    monkeypatch.delenv("WC_AUTOTUNE", raising=False)
    # a file whose tag does not match its name (class 32-64), and one with an
    # invalid shape (tpb 9; 12-real links, class 8-15), are both ignored:
    # the operator tunes and rewrites them
    short_solve(links, 4)  # 2^17 pairs: tuned, gives the tag format
    name = [f for f in os.listdir(tmp_path) if f.endswith("_rhs4.json")][0]
    tag32 = name[:-5].replace("_rhs4", "_rhs32")
    with open(tmp_path / (tag32 + ".json"), "w") as f:
        json.dump({"tag": "something_else", "params": [1, 2, 8, 1, 0], "ms": 1.0}, f)
    tag8c = name[:-5].replace("_rhs4", "_rhs8").replace("_links18_", "_links12_")
    with open(tmp_path / (tag8c + ".json"), "w") as f:
        json.dump({"tag": tag8c, "params": [1, 9, 0, 1, 0], "ms": 1.0}, f)
    w32 = short_solve(links, 32)
    rec = json.load(open(tmp_path / (tag32 + ".json")))
    assert rec["tag"] == tag32 and tuple(rec["params"]) == w32.hop_shape(32)
    w8c = short_solve(links, 8, compress=True)
    rec = json.load(open(tmp_path / (tag8c + ".json")))
    assert rec["tag"] == tag8c and tuple(rec["params"]) == w8c.hop_shape(8) and rec["params"][1] != 9


def mixed16_solve(U, fd, n):
    """a 16-bit mixed solve of n right-hand sides (picks the 16-bit hop shape
    and form on first use); returns its inner operator"""
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=n)
    rng = np.random.default_rng(8200 + n)
    shape = tuple(reversed(fd)) + (4, 3)
    fs = []
    for _ in range(n):
        x = g.lattice(w.F_grid.single_rhs())
        x.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0))
        fs.append(x)
    m = inv.cg_mixed(eps=1e-6, maxiter=2000, inner_storage="half")
    g(w.propagator(inv.preconditioned(pc.eo2_ne(), m)) * pack(fs))
    torch.cuda.synchronize()
    return m.inner_operator(w)


def test_16bit_picks_written_then_reused(tmp_path, monkeypatch):
    """the 16-bit hops' shape and form picks persist too: a solve writes them
    (with the form, 0 vector or 1 matrix cores), and a planted file for
    another class is read back instead of timed (a lattice no other test of
    this process picks for, as picks are also kept process-wide)"""
    fd = [16, 16, 16, 8]
    U = G.smooth_links(fd, 0.3, 8002)
    monkeypatch.setenv("WC_TUNE_CACHE", str(tmp_path))
    monkeypatch.delenv("WC_AUTOTUNE", raising=False)
    mixed16_solve(U, fd, 8)  # 16^3 x 8 / 2 sites x 8 right-hand sides = 2^17 pairs: picked
    names = [f for f in os.listdir(tmp_path) if f.endswith("_rhs8_s16_mma.json")]
    assert len(names) == 1, os.listdir(tmp_path)
    rec = json.load(open(tmp_path / names[0]))
    assert rec["tag"] == names[0][:-5] and rec["form"] in (0, 1) and len(rec["params"]) == 5
    tag16 = rec["tag"].replace("_rhs8", "_rhs16")
    with open(tmp_path / (tag16 + ".json"), "w") as f:
        json.dump({"tag": tag16, "params": rec["params"], "form": 1, "ms": 1.0}, f)
    inner = mixed16_solve(U, fd, 16)
    assert inner.hop_mma(16)  # the planted form, read back
    assert json.load(open(tmp_path / (tag16 + ".json")))["form"] == 1  # not re-timed, not rewritten


def test_empty_setting_writes_nothing(links, tmp_path, monkeypatch):
    monkeypatch.setenv("WC_TUNE_CACHE", "")
    monkeypatch.chdir(tmp_path)
    short_solve(links, 4, compress=True)  # tuned (2^17 pairs), kept in-process only
    assert os.listdir(tmp_path) == []
