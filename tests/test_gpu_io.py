"""NERSC configuration IO on the GPU box: a save / load round trip is exact,
the GPU plaquette reproduces the reference's known answers
(tests/random/simple.py:14-31) and the oracle's plaquette and link trace,
a two-row IEEE32 file written to the format is read with the third row
rebuilt, and a wrong checksum or plaquette is refused."""
import json
import os

import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import io  # noqa: E402
from gpt_amd.qcd import gauge  # noqa: E402


def _links(fd, seed="io", scale=0.5):
    gr = R.grid(fd)
    U = G.random_links(gr, R.random(seed), scale=scale)
    return U, gauge.from_numpy(g.grid(fd, g.double), U)


def test_plaquette_known_answers():
    k = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))["plaquette"]
    rng = R.random(k["seed"])
    for case in k["cases"]:
        U = G.random_links(R.grid(case["fdims"], case["precision"]), rng, case["scale"])
        if case["precision"] != "double":
            continue
        P, L = io.gauge_observables(gauge.from_numpy(g.grid(case["fdims"], g.double), U))
        assert abs(P - case["value"]) ** 2 < case["tol_abs2"], (P, case["value"])
        L_ref = sum(np.trace(u, axis1=-2, axis2=-1).real.mean() for u in U) / 12.0
        assert abs(L - L_ref) < 1e-15


def test_round_trip(tmp_path):
    Un, U = _links([8, 8, 8, 8])
    fn = str(tmp_path / "ckpoint.0000")
    g.save(fn, U, g.format.nersc(label="test", id="gpt_amd", sequence_number=7))
    meta, _ = io.read_header(fn)
    assert meta["ENSEMBLE_LABEL"] == "test" and meta["SEQUENCE_NUMBER"] == "7"
    assert abs(float(meta["PLAQUETTE"]) - G.plaquette(Un)) < 1e-14
    V = g.load(fn)
    for a, b in zip(V, U):
        assert torch.equal(a.data, b.data)
    assert V[0].metadata["DATATYPE"] == "4D_SU3_GAUGE_3x3"


def _write(fn, Un, fd, kind, fp, plaq=None):
    rows = 2 if kind == "4D_SU3_GAUGE" else 3
    m = np.stack([u.reshape(-1, 3, 3) for u in Un], axis=1)[:, :, :rows, :]
    dt = {"IEEE32LITTLE": "<f4", "IEEE64BIG": ">f8"}[fp]
    native = np.ascontiguousarray(m.astype(np.complex128)).view(np.float64).reshape(-1)
    if dt.endswith("f4"):
        native = native.astype(np.float32)
    cs = io.checksum(native)
    P = G.plaquette([u.astype(np.complex128) for u in Un]) if plaq is None else plaq
    hdr = ("BEGIN_HEADER\nDATATYPE = %s\n" % kind + "".join(f"DIMENSION_{i + 1} = {fd[i]}\n" for i in range(4))
           + f"PLAQUETTE = {P:.10f}\nCHECKSUM = {cs:x}\nFLOATING_POINT = {fp}\nEND_HEADER\n")
    with open(fn, "wb") as f:
        f.write(hdr.encode())
        f.write(native.astype(dt).tobytes())
    return native


def test_two_row_single_precision_file(tmp_path):
    fd = [8, 8, 8, 8]
    Un, _ = _links(fd, "io32")
    fn = str(tmp_path / "two_rows")
    _write(fn, Un, fd, "4D_SU3_GAUGE", "IEEE32LITTLE")
    V = g.load(fn)
    for mu in range(4):
        got = V[mu].data.cpu().numpy().reshape(Un[mu].shape)
        assert np.abs(got - Un[mu]).max() < 1e-6  # float32 storage, third row rebuilt


def test_bad_files_are_refused(tmp_path):
    fd = [8, 8, 8, 8]
    Un, _ = _links(fd, "iobad")
    fn = str(tmp_path / "bad_plaquette")
    _write(fn, Un, fd, "4D_SU3_GAUGE_3x3", "IEEE64BIG", plaq=0.123)
    with pytest.raises(ValueError, match="plaquette"):
        g.load(fn)
    fn = str(tmp_path / "bad_checksum")
    _write(fn, Un, fd, "4D_SU3_GAUGE_3x3", "IEEE64BIG")
    raw = bytearray(open(fn, "rb").read())
    raw[-3] ^= 0x10
    open(fn, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        g.load(fn)
