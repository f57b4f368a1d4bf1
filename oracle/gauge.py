"""Gauge-field helpers of the oracle (test infrastructure only).

* ``su3_generators``  -- T_a in Gell-Mann order, Tr(T_a T_b) = delta_ab / 2
  (lib/gpt/core/object_type/su_n.py:208-240)
* ``gpt_matrix_exp``  -- gpt's lattice matrix exponential: one global
  power-of-two rescaling from the field's overall norm, Taylor series to
  order 19, then repeated squaring
  (lib/gpt/core/foundation/lattice/matrix/exp.py:167-220)
* ``random_links``    -- g.qcd.gauge.random (lib/gpt/qcd/gauge/create.py:66-71)
* ``plaquette``       -- average of Re tr U_{mu nu} / 3 over sites and planes
  (lib/gpt/qcd/gauge/loops.py:86-160 with L_mu = L_nu = 1)
* ``field_strength``  -- clover-leaf F_{mu nu} (lib/gpt/qcd/gauge/loops.py:162-174)
* ``smooth_links``    -- the seeded synthetic generator U = exp(i eps H),
  H = sum_a h_a T_a with h_a ~ N(0, 1) from numpy's PCG64 (no gpt
  counterpart; used for the large benchmark lattices)

Pinned by: the plaquette known answers of tests/random/simple.py:18-31
(reproduced through ``random_links``); unitarity and det = 1 of every link;
field_strength vanishing on unit links and agreeing with a hand-built
single-plaquette perturbation (tests/test_oracle_gauge.py).
"""
import numpy as np

from . import conventions as cv


def su3_generators():
    gens = []
    n = 3
    for i in range(n):
        for j in range(i + 1, n):
            a = np.zeros((n, n), np.complex128)
            a[i, j] = a[j, i] = 1.0
            gens.append(a)
            a = np.zeros((n, n), np.complex128)
            a[i, j] = -1j
            a[j, i] = 1j
            gens.append(a)
            if j == i + 1:
                a = np.zeros((n, n), np.complex128)
                for k in range(j):
                    a[k, k] = 1.0
                a[j, j] = -j
                gens.append(a)
    return [g / np.sqrt(2.0 * np.trace(g @ g).real) for g in gens]


def gpt_matrix_exp(x):
    """exp of a field of matrices, with gpt's global rescaling rule."""
    x = np.array(x, dtype=np.complex128, copy=True)
    nsites = int(np.prod(x.shape[:-2]))
    n = np.sqrt(np.vdot(x.reshape(-1), x.reshape(-1)).real) / nsites
    ns = 0
    if n > 0.01:
        ns = int(np.log2(n / 0.01))
        x /= 2.0**ns
    ident = np.broadcast_to(np.eye(x.shape[-1], dtype=np.complex128), x.shape)
    o = ident + x
    xn = x.copy()
    nfac = 1.0
    for j in range(2, 20):
        nfac /= j
        xn = xn @ x
        o = o + xn * nfac
    for _ in range(ns):
        o = o @ o
    return o


def random_links(g, rng, scale=1.0):
    """[U_0, U_1, U_2, U_3] drawn exactly as g.qcd.gauge.random(grid, rng, scale)."""
    return [rng.element(g, scale=scale) for _ in range(4)]


def unit_links(shape):
    return [np.broadcast_to(np.eye(3, dtype=np.complex128), tuple(shape) + (3, 3)).copy()
            for _ in range(4)]


def plaquette(U):
    total = 0.0
    count = 0
    for mu in range(4):
        for nu in range(mu):
            p = (U[mu] @ cv.shift(U[nu], mu, 1) @ cv.dagger(cv.shift(U[mu], nu, 1))
                 @ cv.dagger(U[nu]))
            total += np.trace(p, axis1=-2, axis2=-1).real.sum()
            count += 1
    return total / (count * 3.0 * U[0][..., 0, 0].size)


def field_strength(U, mu, nu):
    """F_{mu nu} = (Q - Q^dag) / 8 with Q the sum of the four leaves."""
    staple_up = cv.shift(U[nu], mu, 1) @ cv.dagger(cv.shift(U[mu], nu, 1)) @ cv.dagger(U[nu])
    staple_down = cv.shift(
        cv.dagger(cv.shift(U[nu], mu, 1)) @ cv.dagger(U[mu]) @ U[nu], nu, -1)
    v = staple_up - staple_down
    F = U[mu] @ v + cv.shift(v @ U[mu], mu, -1)
    return 0.125 * (F - cv.dagger(F))


def smooth_links(fdims, eps, seed):
    """U_mu(x) = exp(i eps H_mu(x)), H = sum_a h_a T_a, h_a ~ N(0,1) (PCG64 seed)."""
    shape = cv.fdims_to_shape(fdims)
    r = np.random.default_rng(seed)
    h = r.standard_normal((4,) + tuple(shape) + (8,))
    T = np.stack(su3_generators())
    H = np.einsum("...a,aij->...ij", h, T)
    return [su3_exp_hermitian(H[mu], eps) for mu in range(4)]


def su3_exp_hermitian(H, eps):
    """exp(i eps H) for Hermitian H by eigendecomposition (exact to rounding)."""
    w, v = np.linalg.eigh(H)
    ph = np.exp(1j * eps * w)
    return np.einsum("...ij,...j,...kj->...ik", v, ph, np.conj(v))


def haar_links(fdims, seed):
    """Haar-random SU(3): QR of a complex Gaussian matrix, phases fixed so
    R has a positive diagonal, then det normalised to 1."""
    shape = cv.fdims_to_shape(fdims)
    r = np.random.default_rng(seed)
    z = (r.standard_normal((4,) + tuple(shape) + (3, 3))
         + 1j * r.standard_normal((4,) + tuple(shape) + (3, 3))) / np.sqrt(2.0)
    q, rr = np.linalg.qr(z)
    d = np.diagonal(rr, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(q)
    q[..., :, 0] /= det[..., None]  # det of the phase-fixed Q has modulus 1
    return [q[mu] for mu in range(4)]
