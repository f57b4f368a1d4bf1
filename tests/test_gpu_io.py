"""Configuration IO with the GPU computing the checks.  NERSC: a save / load round trip is
exact, the GPU plaquette reproduces the reference's known answers
(tests/random/simple.py:14-31) and the oracle's plaquette and link trace,
a two-row IEEE32 file written to the format is read with the third row
rebuilt, and a wrong checksum or plaquette is refused.  LIME: gpt's records
round-trip exactly; ILDG files written value by value (64- and 32-bit) are
read back, a corrupted one is refused by its SciDAC checksum.  openQCD: a
file written point by point is read back exactly, a wrong header plaquette is
refused."""
import json
import os
import struct

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
        Ud = gauge.from_numpy(g.grid(case["fdims"], g.double), U)
        P, L = io.gauge_observables(Ud)
        assert abs(P - case["value"]) ** 2 < case["tol_abs2"], (P, case["value"])
        assert g.qcd.gauge.plaquette(Ud) == P and g.qcd.gauge.link_trace(Ud) == L
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


def test_lime_round_trip(tmp_path):
    """g.save(fn, U, g.format.lime()) / g.format.grid_scidac and g.load are
    exact (tests/io/io.py:208-216 asks 1e-14) and write the records gpt writes"""
    Un, U = _links([8, 8, 8, 8], "lime")
    for fmt, data_tag in ((g.format.lime(), "gpt-binary-data"), (g.format.grid_scidac, "ildg-binary-data")):
        fn = str(tmp_path / f"lime.{data_tag}")
        g.save(fn, U, fmt)
        r = io.lime_reader(fn)
        assert list(r.index)[-3:] == [data_tag, "scidac-checksum", "gpt-format"]
        assert r.xml_dict("gpt-format") == {"grid": "[8,8,8,8];double;full", "n": "4",
                                           "otype": "ot_matrix_su_n_fundamental_group(3);none"}
        V = g.load(fn)
        for a, b in zip(V, U):
            assert torch.equal(a.data, b.data)


def _ildg(fn, Un, fd, prec, checksum=True):
    """an ILDG file written value by value: site-major, 4 links, row-major
    3x3; 64-bit big-endian doubles, or 32-bit as the reference stores a
    complex pair (its 8 bytes reversed: big-endian imaginary, then real)"""
    V = int(np.prod(fd))
    links = [u.reshape(V, 3, 3) for u in Un]
    out = bytearray()
    for s in range(V):
        for mu in range(4):
            for z in links[mu][s].reshape(-1):
                out += struct.pack(">dd", z.real, z.imag) if prec == 64 else struct.pack(">ff", z.imag, z.real)
    w = io.lime_writer(fn)
    w.xml("ildg-format", "ildgFormat", {"version": "1.0", "field": "su3gauge", "precision": prec,
                                        "lx": fd[0], "ly": fd[1], "lz": fd[2], "lt": fd[3]})
    w.record("ildg-binary-data", bytes(out))
    if checksum:
        a, b = io.scidac_checksums(bytes(out), len(out) // V)
        w.xml("scidac-checksum", "scidacChecksum", {"version": "1.0", "suma": f"{a:x}", "sumb": f"{b:x}"})
    w.close()
    return bytes(out)


def test_ildg_files(tmp_path):
    fd = [4, 4, 4, 8]
    Un, _ = _links(fd, "ildg")
    for prec, tol in ((64, 0.0), (32, 1e-7)):
        fn = str(tmp_path / f"ildg{prec}")
        _ildg(fn, Un, fd, prec)
        V = g.load(fn)
        for mu in range(4):
            got = V[mu].data.cpu().numpy().reshape(Un[mu].shape)
            assert np.abs(got - Un[mu]).max() <= tol
    fn = str(tmp_path / "ildg_nochecksum")
    _ildg(fn, Un, fd, 64, checksum=False)
    assert torch.equal(g.load(fn)[2].data.cpu(), torch.from_numpy(Un[2].reshape(-1, 3, 3)))
    good = str(tmp_path / "ildg64")
    off, size = io.lime_reader(good).index["ildg-binary-data"]
    raw = bytearray(open(good, "rb").read())
    raw[off + size // 2] ^= 0x01  # one data byte
    fn = str(tmp_path / "ildg_bad")
    open(fn, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        g.load(fn)


def _openqcd(fn, Un, fd, plaq):
    """an openQCD file written point by point: header (T, X, Y, Z, plaquette
    normalised to 3), then per odd point, x3 fastest, U(x,0) U(x-0,0) ...
    U(x,3) U(x-3,3) with openQCD's direction 0 the time direction"""
    X, Y, Z, T = fd
    out = bytearray(struct.pack("<iiiid", T, X, Y, Z, plaq))
    for t in range(T):
        for x in range(X):
            for y in range(Y):
                for z in range(Z):
                    if (t + x + y + z) % 2 == 0:
                        continue
                    c = [x, y, z, t]
                    for mu in (3, 0, 1, 2):
                        b = list(c)
                        b[mu] = (b[mu] - 1) % fd[mu]
                        for p in (c, b):
                            out += Un[mu][p[3], p[2], p[1], p[0]].astype("<c16").tobytes()
    open(fn, "wb").write(bytes(out))


def test_openqcd_files(tmp_path):
    fd = [4, 6, 4, 8]
    Un, _ = _links(fd, "openqcd")
    P = G.plaquette(Un)
    fn = str(tmp_path / "oqcd")
    _openqcd(fn, Un, fd, 3.0 * P)
    V = g.load(fn)
    for mu in range(4):
        assert np.array_equal(V[mu].data.cpu().numpy().reshape(Un[mu].shape), Un[mu])
    assert V[0].metadata["DIMENSION_4"] == "8"
    _openqcd(fn, Un, fd, 3.0 * P + 1e-6)
    with pytest.raises(ValueError, match="plaquette"):
        g.load(fn)
