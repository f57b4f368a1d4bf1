"""The oracle's restatement of gpt's RNG against the reference's known answers
(tests/random/simple.py:14-31 and :66-130)."""
import hashlib

import numpy as np

from oracle import gauge as G
from oracle import rng as R


def test_sha256_matches_fips_digest():
    for msg in [b"", b"abc", bytes(range(55)), bytes(range(56)), bytes(range(200))]:
        h = hashlib.sha256(msg).digest()
        assert R.sha256_words(msg) == [int.from_bytes(h[4 * i:4 * i + 4], "big") for i in range(8)]


def test_plaquette_known_answers(kats):
    k = kats["plaquette"]
    rng = R.random(k["seed"])
    for case in k["cases"]:
        g = R.grid(case["fdims"], case["precision"])
        U = G.random_links(g, rng, case["scale"])
        p = G.plaquette([u.astype(np.complex128) for u in U])
        assert abs(p - case["value"]) ** 2 < case["tol_abs2"]
        for u in U:
            u = u.astype(np.complex128)
            dev = np.abs(np.conj(np.swapaxes(u, -1, -2)) @ u - np.eye(3)).max()
            assert dev < (1e-6 if case["precision"] == "single" else 1e-13)
            assert np.abs(np.linalg.det(u) - 1).max() < (1e-6 if case["precision"] == "single" else 1e-13)


def test_normal_sequence_known_answers(kats):
    k = kats["normal_sequence"]
    rng = R.random(k["seed"])
    g = R.grid(k["fdims"])

    def pick(v):
        return np.array([v[t, z, y, x].real for x, y, z, t in k["sites_xyzt"]])

    v = rng.normal(g)
    assert np.linalg.norm(pick(v) - k["first_fill"]) < k["tol"]
    for _ in range(1000):
        v = rng.normal(g)
    assert np.linalg.norm(pick(v) - k["after_1000_more_fills"]) < k["tol"]


def test_distinct_grids_restart_streams():
    # gpt keys streams by grid object: a second grid of equal dims replays them
    rng = R.random("x")
    a = rng.cnormal(R.grid([4, 4, 4, 4]), (4, 3))
    b = rng.cnormal(R.grid([4, 4, 4, 4]), (4, 3))
    c = rng.cnormal(R.grid([4, 4, 4, 4]), (4, 3))
    assert np.array_equal(a, b) and np.array_equal(b, c)


def test_cnormal_moments():
    rng = R.random("moments")
    v = rng.cnormal(R.grid([8, 8, 8, 8]), (4, 3)).reshape(-1)
    assert abs(np.mean(v)) < 0.02
    assert abs(np.mean(np.abs(v) ** 2) - 2.0) < 0.03  # re and im each unit variance
