"""Gauge helpers of the oracle: the SU(3) generators (lib/gpt/core/object_type/su_n.py:208-240),
the link generators that feed the benchmark configurations (smooth U = exp(i eps H) links and
Haar-random links), gpt's gauge.random links and the clover-leaf field strength.

What fixes each one:
* generators: Hermitian, traceless, Tr(T_a T_b) = delta_ab / 2;
* every link generator: unitarity and det = 1 per link;
* smooth_links: the exponential against scipy's expm (an independent algorithm) and the
  moments of H (tr H^2 = 4 on average for h_a ~ N(0, 1));
* haar_links: the Haar moments of SU(3), E[tr U] = 0, E[|tr U|^2] = 1 and E[(tr U)^3] = 1
  (the one singlet of 3 x 3 x 3; it is 0 for U(3)).  A QR factor without its phases fixed, or
  without the determinant divided out, fails them (checked below);
* field_strength: zero on unit links and anti-Hermitian (the single-plaquette value is in
  tests/test_oracle_operator.py).
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import conventions as cv
from oracle import gauge as G
from oracle import rng as R


def _su3_checks(U, tol=1e-13):
    for u in U:
        m = u.reshape(-1, 3, 3)
        dev = np.abs(np.conj(np.swapaxes(m, 1, 2)) @ m - np.eye(3)).max()
        assert dev < tol, dev
        assert np.abs(np.linalg.det(m) - 1.0).max() < tol


def test_generators():
    T = G.su3_generators()
    assert len(T) == 8
    for a, ta in enumerate(T):
        assert np.abs(ta - ta.conj().T).max() == 0.0
        assert abs(np.trace(ta)) < 1e-15
        for b, tb in enumerate(T):
            assert abs(np.trace(ta @ tb) - 0.5 * (a == b)) < 1e-15, (a, b)


def test_smooth_links():
    fd = [4, 4, 4, 8]
    eps, seed = 0.3, 2001
    U = G.smooth_links(fd, eps, seed)
    _su3_checks(U)
    # the same draws as smooth_links: h ~ N(0, 1) per generator, mu slowest
    shape = cv.fdims_to_shape(fd)
    h = np.random.default_rng(seed).standard_normal((4,) + tuple(shape) + (8,))
    T = np.stack(G.su3_generators())
    H = np.einsum("...a,aij->...ij", h, T)
    trh2 = np.einsum("...ij,...ji->...", H, H).real
    assert abs(trh2.mean() - 4.0) < 0.2  # 8 generators x 1/2, 4096 samples
    for mu in range(4):
        for site in [(0, 0, 0, 0), (3, 1, 2, 3), (7, 3, 3, 0)]:
            ref = scipy.linalg.expm(1j * eps * H[(mu,) + site])
            assert np.abs(U[mu][site] - ref).max() < 1e-14
    # different seeds, different fields
    assert np.abs(G.smooth_links(fd, eps, seed + 1)[0] - U[0]).max() > 0.1


def _haar_moments(m):
    t = np.trace(m, axis1=-2, axis2=-1)
    return t.mean(), (np.abs(t) ** 2).mean(), (t ** 3).mean()


def test_haar_links():
    U = G.haar_links([8, 8, 8, 8], 1001)
    _su3_checks(U)
    m1, m2, m3 = _haar_moments(np.stack(U).reshape(-1, 3, 3))  # 16384 links
    # standard errors: 0.008 (tr U), 0.008 (|tr U|^2), 0.019 ((tr U)^3)
    assert abs(m1) < 0.04 and abs(m2 - 1.0) < 0.05 and abs(m3 - 1.0) < 0.1, (m1, m2, m3)


@pytest.mark.parametrize("mistake", ["no_phase_fix", "no_det"])
def test_haar_moments_catch_planted_mistakes(mistake):
    """the moment test fails for the two plausible mistakes of the generator"""
    r = np.random.default_rng(1001)
    z = (r.standard_normal((16384, 3, 3)) + 1j * r.standard_normal((16384, 3, 3))) / np.sqrt(2.0)
    q, rr = np.linalg.qr(z)
    if mistake == "no_phase_fix":
        det = np.linalg.det(q)
        q[..., :, 0] /= det[..., None]
    else:
        d = np.diagonal(rr, axis1=-2, axis2=-1)
        q = q * (d / np.abs(d))[..., None, :]
    m1, m2, m3 = _haar_moments(q)
    assert abs(m1) >= 0.04 or abs(m2 - 1.0) >= 0.05 or abs(m3 - 1.0) >= 0.1


def test_random_links_are_su3():
    g = R.grid([4, 4, 4, 4])
    for scale in (0.5, 2.0, 10.0):
        _su3_checks(G.random_links(g, R.random("su3"), scale=scale), tol=1e-12)


def test_field_strength_unit_and_antihermitian():
    shape = cv.fdims_to_shape([4, 4, 4, 4])
    Uu = G.unit_links(shape)
    U = G.smooth_links([4, 4, 4, 4], 0.5, 7)
    for mu in range(4):
        for nu in range(mu + 1, 4):
            assert np.abs(G.field_strength(Uu, mu, nu)).max() == 0.0
            F = G.field_strength(U, mu, nu)
            assert np.abs(F + cv.dagger(F)).max() < 1e-15
            assert np.abs(F).max() > 1e-3
            # F_{nu mu} = -F_{mu nu}
            assert np.abs(G.field_strength(U, nu, mu) + F).max() < 1e-14
