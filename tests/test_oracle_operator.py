"""The numpy Wilson-clover oracle against Grid's fingerprints
(tests/qcd/fermion_operators.py:338-382) and against closed forms."""
import cmath

import numpy as np
import pytest

from oracle import conventions as cv
from oracle import gauge as G
from oracle import rng as R
from oracle import wilson_clover as W


@pytest.fixture(scope="module")
def fp_fields(kats):
    k = kats["fingerprints"]
    g = R.grid(k["fdims"])
    rng = R.random(k["seed"])
    U = G.random_links(g, rng)
    src = rng.vspincolor(g)
    dst = rng.vspincolor(g)
    return U, src, dst


@pytest.mark.parametrize("name", ["wilson", "wilson_clover", "wilson_clover_open"])
def test_grid_fingerprints(kats, fp_fields, name):
    k = kats["fingerprints"]
    op = k["operators"][name]
    U, src, dst = fp_fields
    src, dst = src.copy(), dst.copy()
    p = dict(op["params"])
    if p["boundary_phases"][-1] == 0.0:
        for f in (src, dst):
            f[0] = 0.0
            f[-1] = 0.0
    w = W.wilson_clover(U, **p)
    for tag, fn in [("M", w.M), ("Mdiag", w.Mdiag)]:
        X = cv.inner_product(dst, fn(src))
        ref = complex(*op[tag])
        assert abs(X - ref) / abs(X) < k["tol_rel"], (name, tag, X, ref)
    # adjoint consistency <dst, M src> = conj(<src, Mdag dst>) (fermion_operators.py:710-714)
    X = cv.inner_product(dst, w.M(src))
    Y = np.conj(cv.inner_product(src, w.Mdag(dst)))
    assert abs(X - Y) / abs(X) < 1e-13
    # even-odd decomposition M = Mooee + Meooe on each parity (:724-743)
    for p_ in (0, 1):
        s = w._mask(src, p_)
        full = w.M(s)
        eo = w.Mooee(s, p_) + w.Meooe(s, p_)
        assert np.abs(full - eo).max() < 1e-12


def test_free_field_plane_wave():
    """U = 1 with twisted phases: M e^{ipx} chi = [1/(2k) - sum c cos p + i sum c g sin p] e^{ipx} chi."""
    fd = [4, 6, 4, 8]
    shape = cv.fdims_to_shape(fd)
    U = G.unit_links(shape)
    theta = [0.3, -1.1, 2.0, np.pi]
    phases = [cmath.exp(1j * t) for t in theta]
    xi0, nu, m0 = 1.2, 1.7, 0.07
    w = W.wilson_clover(U, mass=m0, csw_r=1.3, csw_t=0.9, xi_0=xi0, nu=nu, boundary_phases=phases)
    n = [1, 2, 3, 1]
    pm = [(2 * np.pi * n[mu] + theta[mu]) / fd[mu] for mu in range(4)]
    x, y, z, t = cv.coords(shape)
    wave = np.exp(1j * (pm[0] * x + pm[1] * y + pm[2] * z + pm[3] * t))
    chi = np.random.default_rng(0).standard_normal((4, 3)) + 0j
    psi = wave[..., None, None] * chi
    c = [nu / xi0] * 3 + [1.0]
    D = (m0 + 1 + 3 * nu / xi0) * np.eye(4, dtype=complex)
    for mu in range(4):
        D += -c[mu] * np.cos(pm[mu]) * np.eye(4) + 1j * c[mu] * np.sin(pm[mu]) * cv.GAMMA[mu]
    expect = wave[..., None, None] * (D @ chi)
    assert np.abs(w.M(psi) - expect).max() < 1e-13


def test_clover_single_plaquette():
    """One non-trivial link U_0(x0) = V = exp(i a T_3) on unit links.  The two
    plaquettes holding that link in the (0,1) plane give leaves V (upper) and
    V^dag (lower): F_01 cancels at x0 and is (V - V^dag)/8 at x0 + y_hat, where
    no other plane is touched, so there A = 1/(2 kappa) - csw/2 sigma_01 F_01."""
    fd = [4, 4, 4, 4]
    shape = cv.fdims_to_shape(fd)
    U = G.unit_links(shape)
    V = G.su3_exp_hermitian(G.su3_generators()[2], 0.4)
    U[0][0, 0, 0, 0] = V
    F = G.field_strength(U, 0, 1)
    assert np.abs(F[0, 0, 0, 0]).max() < 1e-15
    assert np.allclose(F[0, 0, 1, 0], (V - cv.dagger(V)) / 8.0, atol=1e-15)
    csw, m = 1.3, 0.1
    w = W.wilson_clover(U, mass=m, csw_r=csw, csw_t=csw)
    A = w.clover[0, 0, 1, 0].reshape(4, 3, 4, 3)
    expect = -0.5 * csw * np.einsum("ab,ij->aibj", cv.sigma(0, 1), (V - cv.dagger(V)) / 8.0)
    for s_ in range(4):
        for c in range(3):
            expect[s_, c, s_, c] += 4.0 + m
    assert np.abs(A - expect).max() < 1e-15
    A12 = A.reshape(12, 12)
    assert np.abs(A12 - A12.conj().T).max() < 1e-15  # Hermitian
    assert np.abs(A12[:6, 6:]).max() == 0.0  # chiral block diagonal
