"""Lattice conventions of gpt, restated for numpy (oracle; test infrastructure).

Field layout used by the oracle: a lattice field is a numpy array whose first
four axes are (t, z, y, x) -- gpt's lexicographic site order with x fastest
(tests/qcd/openqcd_dslash.py:105-113) -- followed by the per-site tensor
indices.  A spinor is [..., spin(4), colour(3)]; a link is [..., 3, 3].
Direction mu = 0, 1, 2, 3 is x, y, z, t, so mu lives on numpy axis 3 - mu.

Pinned by: the gamma algebra (anticommutation, Hermiticity, gamma5 =
g0 g1 g2 g3) checked in tests/test_oracle_conventions.py; the fingerprints of
lib/gpt tests/qcd/fermion_operators.py:371-382 (which fail under any other
gamma basis or shift sign).
"""
import numpy as np

# lib/gpt/core/gamma.py:28-42 (chiral basis, gamma5 = diag(1, 1, -1, -1))
GAMMA = {
    0: np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]], np.complex128),
    1: np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], np.complex128),
    2: np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]], np.complex128),
    3: np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], np.complex128),
    5: np.diag([1, 1, -1, -1]).astype(np.complex128),
}
ID4 = np.eye(4, dtype=np.complex128)


def sigma(mu, nu):
    """sigma_{mu nu} = [gamma_mu, gamma_nu] / 2 (lib/gpt/core/gamma.py:45-53)."""
    a, b = GAMMA[mu], GAMMA[nu]
    return 0.5 * (a @ b - b @ a)


def axis(mu):
    """numpy axis of direction mu for (t, z, y, x)-ordered fields."""
    return 3 - mu


def fdims_to_shape(fdims):
    """gpt grid dims [Lx, Ly, Lz, Lt] -> numpy shape (Lt, Lz, Ly, Lx)."""
    return tuple(reversed(list(fdims)))


def coords(shape):
    """Integer coordinates (x, y, z, t) broadcastable over a (t, z, y, x) shape."""
    t, z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return x, y, z, t


def parity(shape):
    """0 on even sites, 1 on odd: sum of coordinates mod 2
    (lib/gpt/core/checkerboard.py:23-35)."""
    x, y, z, t = coords(shape)
    return (x + y + z + t) % 2


def shift(field, mu, disp):
    """gpt.cshift: result[x] = field[x + disp * mu_hat], periodic."""
    return np.roll(field, -disp, axis=axis(mu))


def spin_mul(gmat, psi):
    """Apply a 4x4 spin matrix to a spinor field [..., 4, 3]."""
    return np.matmul(gmat, psi)


def colour_mul(u, psi):
    """Apply a colour matrix field [..., 3, 3] to a spinor field [..., 4, 3]."""
    return np.matmul(psi, np.swapaxes(u, -1, -2))


def dagger(u):
    return np.conj(np.swapaxes(u, -1, -2))


def inner_product(a, b):
    """<a, b> = sum conj(a) b over all sites and components."""
    return np.vdot(a.reshape(-1), b.reshape(-1))


def norm2(a):
    return float(np.vdot(a.reshape(-1), a.reshape(-1)).real)
