"""The Wilson-clover operator restated in numpy (oracle; test infrastructure).

Follows lib/gpt/qcd/fermion/reference/wilson_clover.py:27-281 term by term:

* parameters (:41-83): mass from kappa, csw_r divided by xi_0, the diagonal
  1/(2 kappa) = m0 + 1 + 3 nu / xi_0, open boundary when the last boundary
  phase is 0 (then cF enters next to the temporal boundaries);
* boundary phases multiply the links on the last slice of each direction
  (lib/gpt/core/covariant.py:28-37);
* clover (:91-142): A = 1/(2 kappa) - 1/2 sum_{mu<nu} c_{mu nu} sigma_{mu nu}
  F_{mu nu}, c = csw_t for planes with t, csw_r / xi_0 otherwise, and its
  per-site inverse;
* Meooe (:182-200): -c_mu/2 [(1 - gamma_mu) U_mu(x) psi(x+mu)
  + (1 + gamma_mu) U_mu(x-mu)^dag psi(x-mu)], c_mu = nu/xi_0 spatial,
  1 temporal; Dhop is the same on both parities (:222-234), M = Dhop + Mooee
  (:236-248), Mdiag = Mooee (:250-262); open boundaries zero the output's
  first and last time slice (lib/gpt/qcd/fermion/boundary_conditions.py:23-31).
* the adjoint uses gamma5-hermiticity: M^dag = g5 M g5, Meooe^dag =
  g5 Meooe g5, (Mooee^-1)^dag = Mooee^-1 (Hermitian clover blocks).

Fields are full-lattice numpy arrays [t, z, y, x, 4, 3]; an even-odd
("checkerboarded") field is a full array that is zero on the other parity.

Pinned by the Grid fingerprints <dst, M src> and <dst, Mdiag src> of
tests/qcd/fermion_operators.py:347-382 (Wilson, anisotropic clover, open-BC
clover; tolerance 1e-13 relative, :527, :906-908) and by the free-field
closed form D(p) = m + sum_mu (1 - cos p_mu) + i sum_mu gamma_mu sin p_mu
(tests/test_oracle_operator.py).
"""
import numpy as np

from . import conventions as cv
from .gauge import field_strength

G5 = cv.GAMMA[5]


class wilson_clover:
    def __init__(self, U, kappa=None, mass=None, csw_r=0.0, csw_t=0.0, xi_0=1.0, nu=1.0, cF=1.0,
                 boundary_phases=(1.0, 1.0, 1.0, 1.0), isAnisotropic=None, use_legacy=False,
                 n_rhs=1):
        shape = U[0].shape[:4]
        self.shape = shape
        self.L = list(reversed(shape))  # [Lx, Ly, Lz, Lt]
        self.boundary_phases = [complex(b) for b in boundary_phases]
        self.U_raw = [np.asarray(u, dtype=np.complex128) for u in U]
        # links with boundary phases on the last slice of each direction
        self.U = []
        for mu in range(4):
            u = np.array(U[mu], dtype=np.complex128, copy=True)
            idx = [slice(None)] * 4
            idx[cv.axis(mu)] = self.L[mu] - 1
            u[tuple(idx)] *= self.boundary_phases[mu]
            self.U.append(u)
        self.Udag = [cv.dagger(u) for u in self.U]

        if kappa is not None:
            assert mass is None
            self.m0 = 1.0 / kappa / 2.0 - 4.0
        else:
            self.m0 = mass
        self.xi_0 = xi_0
        self.csw_r = csw_r / xi_0
        self.csw_t = csw_t
        self.nu = nu
        self.kappa = 1.0 / (2.0 * (self.m0 + 1.0 + 3.0 * nu / xi_0))
        self.open_bc = self.boundary_phases[3] == 0.0
        self.cF = cF
        self.parity = cv.parity(shape)

        self.clover = None
        self.clover_inv = None
        if self.csw_r != 0.0 or self.csw_t != 0.0:
            # the clover leaves use the links as given (no boundary phases),
            # as the reference does (wilson_clover.py:93-105)
            U0 = [np.asarray(u, dtype=np.complex128) for u in U]
            A = np.zeros(shape + (4, 3, 4, 3), dtype=np.complex128)
            for mu in range(4):
                for nu_ in range(mu + 1, 4):
                    cp = self.csw_t if (mu == 3 or nu_ == 3) else self.csw_r
                    F = field_strength(U0, mu, nu_)
                    A += -0.5 * cp * np.einsum("ab,...ij->...aibj", cv.sigma(mu, nu_), F)
            if self.open_bc:
                T = self.L[3]
                val = -0.5 * self.csw_t
                for t in (0, T - 1):
                    A[t] = 0.0
                    for a in range(4):
                        for c in range(3):
                            A[t, ..., a, c, a, c] = val
                if cF != 1.0:
                    for t in (1, T - 2):
                        for a in range(4):
                            for c in range(3):
                                A[t, ..., a, c, a, c] += cF - 1.0
            for a in range(4):
                for c in range(3):
                    A[..., a, c, a, c] += 0.5 / self.kappa
            self.clover = A.reshape(shape + (12, 12))
            self.clover_inv = np.linalg.inv(self.clover)

    # ------------------------------------------------------------ helpers
    def _apply_boundaries(self, dst):
        if self.open_bc:
            dst[0] = 0.0
            dst[self.L[3] - 1] = 0.0
        return dst

    def _mask(self, psi, p):
        out = psi.copy()
        out[self.parity != p] = 0.0
        return out

    def _hop(self, src):
        """Full-lattice hop: sum over mu of both terms, -1/2 and c_mu included."""
        dst = np.zeros_like(src, dtype=np.complex128)
        for mu in range(4):
            cc = 1.0 if mu == 3 else self.nu / self.xi_0
            fwd = cv.colour_mul(self.U[mu], cv.shift(src, mu, 1))
            bwd = cv.shift(cv.colour_mul(self.Udag[mu], src), mu, -1)
            dst += (cc / 2.0) * (cv.spin_mul(cv.GAMMA[mu] - cv.ID4, fwd)
                                 - cv.spin_mul(cv.GAMMA[mu] + cv.ID4, bwd))
        return dst

    def _site12(self, mat, psi):
        s = psi.reshape(psi.shape[:4] + (12,))
        return np.einsum("...ij,...j->...i", mat, s).reshape(psi.shape)

    # ------------------------------------------------------- the operator
    def Meooe(self, src, src_parity):
        """Hop from parity src_parity to the other parity."""
        dst = self._hop(self._mask(src, src_parity))
        dst = self._mask(dst, 1 - src_parity)
        return self._apply_boundaries(dst)

    def MeooeDag(self, src, src_parity):
        return cv.spin_mul(G5, self.Meooe(cv.spin_mul(G5, src), src_parity))

    def Mooee(self, src, p=None):
        if self.clover is not None:
            dst = self._site12(self.clover, src)
        else:
            dst = (0.5 / self.kappa) * src
        if p is not None:
            dst = self._mask(dst, p)
        return self._apply_boundaries(dst)

    def MooeeInv(self, src, p=None):
        if self.clover_inv is not None:
            dst = self._site12(self.clover_inv, src)
        else:
            dst = (2.0 * self.kappa) * src
        if p is not None:
            dst = self._mask(dst, p)
        return self._apply_boundaries(dst)

    MooeeDag = Mooee
    MooeeInvDag = MooeeInv

    def Dhop(self, src):
        return self._apply_boundaries(self._hop(src))

    def DhopDag(self, src):
        return cv.spin_mul(G5, self.Dhop(cv.spin_mul(G5, src)))

    def Mdiag(self, src):
        return self.Mooee(src)

    def M(self, src):
        return self._apply_boundaries(self._hop(src) + self.Mooee(src))

    def Mdag(self, src):
        return cv.spin_mul(G5, self.M(cv.spin_mul(G5, src)))

    __call__ = M

    # packed clover: the two chiral 6x6 blocks (spin 0,1 and spin 2,3)
    def clover_blocks(self, inverse=False):
        A = self.clover_inv if inverse else self.clover
        if A is None:
            return None
        A = A.reshape(self.shape + (4, 3, 4, 3))
        up = A[..., 0:2, :, 0:2, :].reshape(self.shape + (6, 6))
        dn = A[..., 2:4, :, 2:4, :].reshape(self.shape + (6, 6))
        return up, dn
