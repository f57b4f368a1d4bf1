"""Conventions of the oracle (lib/gpt/core/gamma.py:28-53, checkerboard.py:23-35,
covariant.py:28-37): the gamma algebra fixes the transcribed matrices (a wrong
sign or a swapped entry breaks anticommutation, Hermiticity or
gamma5 = g0 g1 g2 g3), the chiral block form the kernels' spin projection
relies on, and the shift / parity conventions."""
import itertools

import numpy as np

from oracle import conventions as cv

G = cv.GAMMA
I4 = np.eye(4)


def test_clifford_algebra_and_hermiticity():
    for mu, nu in itertools.product(range(4), repeat=2):
        anti = G[mu] @ G[nu] + G[nu] @ G[mu]
        assert np.abs(anti - 2.0 * (mu == nu) * I4).max() == 0.0, (mu, nu)
    for mu in (0, 1, 2, 3, 5):
        assert np.abs(G[mu] - G[mu].conj().T).max() == 0.0, mu


def test_gamma5_is_the_product_and_anticommutes():
    assert np.abs(G[0] @ G[1] @ G[2] @ G[3] - G[5]).max() == 0.0
    assert np.abs(G[5] - np.diag([1, 1, -1, -1])).max() == 0.0
    for mu in range(4):
        assert np.abs(G[5] @ G[mu] + G[mu] @ G[5]).max() == 0.0


def test_chiral_block_form():
    """gamma_mu = [[0, X_mu], [X_mu^dag, 0]] with the 2x2 blocks the hop's
    spin projection uses, and sigma_{mu nu} block diagonal (the clover term
    splits into two 6x6 chiral blocks)"""
    X = {0: [[0, 1j], [1j, 0]], 1: [[0, -1], [1, 0]], 2: [[1j, 0], [0, -1j]], 3: [[1, 0], [0, 1]]}
    for mu in range(4):
        x = np.array(X[mu], complex)
        assert np.abs(G[mu][:2, :2]).max() == 0.0 and np.abs(G[mu][2:, 2:]).max() == 0.0
        assert np.abs(G[mu][:2, 2:] - x).max() == 0.0
        assert np.abs(G[mu][2:, :2] - x.conj().T).max() == 0.0
    for mu, nu in itertools.combinations(range(4), 2):
        s = cv.sigma(mu, nu)
        assert np.abs(s[:2, 2:]).max() == 0.0 and np.abs(s[2:, :2]).max() == 0.0
        assert np.abs(s - 0.5 * (G[mu] @ G[nu] - G[nu] @ G[mu])).max() == 0.0
        # sigma_{mu nu} is anti-Hermitian and squares to -1 (mu != nu)
        assert np.abs(s + s.conj().T).max() == 0.0
        assert np.abs(s @ s + I4).max() == 0.0


def test_shift_and_parity():
    shape = (4, 2, 6, 8)  # (t, z, y, x)
    f = np.arange(int(np.prod(shape))).reshape(shape)
    x, y, z, t = cv.coords(shape)
    for mu, c in enumerate((x, y, z, t)):
        L = shape[cv.axis(mu)]
        for disp in (1, -1):
            s = cv.shift(f, mu, disp)
            # result[x] = field[x + disp mu_hat], periodic
            idx = [t, z, y, x]
            idx[3 - mu] = (c + disp) % L
            assert np.array_equal(s, f[tuple(idx)]), (mu, disp)
    p = cv.parity(shape)
    assert np.array_equal(p, (x + y + z + t) % 2)
    assert p[0, 0, 0, 0] == 0 and p[0, 0, 0, 1] == 1
    assert cv.fdims_to_shape([8, 6, 2, 4]) == shape
