"""Solvers of the oracle (test infrastructure only).

* ``cg``                  -- lib/gpt/algorithms/inverter/cg.py:48-112, the same
  recursion and stopping rule (|r|^2 <= eps^2 |b|^2 on the iterated
  residual, one history entry per iteration)
* ``eo2_ne``              -- lib/gpt/qcd/fermion/preconditioner/even_odd_sites.py:61-72
  over lib/gpt/algorithms/preconditioner/schur_complement_two.py:58-152 and
  normal_equation.py:35-65 with D = ``parity`` sites (odd by default)
* ``eo1_ne``              -- even_odd_sites.py:35-46 over
  lib/gpt/algorithms/preconditioner/schur_complement_one.py and
  normal_equation.py (N = 1 - DD^-1 DC CC^-1 CD on D)
* ``eo2_kappa_ne``        -- even_odd_sites.py:75-89: schur_complement_two under
  similarity_transformation.py with the scalar V = kappa, then the normal
  equation
* ``preconditioned``      -- lib/gpt/algorithms/inverter/preconditioned.py:32-54
  (dst = L Mpc^-1 R src + S src, initial guess L^-1 dst = DD P_D dst)
* ``defect_correcting``   -- lib/gpt/algorithms/inverter/defect_correcting.py:74-142
* ``block_cg``            -- lib/gpt/algorithms/inverter/block_cg.py:36-149

Pinned by: the 16-value correlator of tests/qcd/fermion_operators.py:193-218
(eo2_ne + CG on the Wilson operator with complex boundary phases); by
brute force on a 2^4 / 4^4 lattice against numpy.linalg.solve of the dense
operator; and by the Schur determinant identity
det M = det DD det CC det Mpc (tests/test_oracle_solvers.py).
``defect_correcting`` is pinned by closed forms on a dense 2^3 x 4 operator:
one outer step with an exact inner inverse, exactly 2^-i true residual after
i steps with an inner inverse that returns half the correction, the zero
source, and a 1e-45 source solved with a single-precision inner inverse
(the normalisation), which copies of the loop with the normalisation
removed fail.  ``block_cg``: see test_block_cg_pinned.
"""
import numpy as np

from . import conventions as cv


def cg(mat, src, psi=None, eps=1e-15, maxiter=1000000, eps_abs=None, miniter=0, history=None):
    """Returns (psi, iterations, converged)."""
    psi = np.zeros_like(src) if psi is None else psi.copy()
    mmp = mat(psi)
    r = src - mmp
    p = r.copy()
    cp = cv.norm2(p)
    ssq = cv.norm2(src)
    if ssq == 0.0:
        return np.zeros_like(src), 0, True
    rsq = eps**2 * ssq
    k = -1
    for k in range(maxiter):
        c = cp
        mmp = mat(p)
        d = cv.inner_product(p, mmp).real
        a = c / d
        r = r - a * mmp
        cp = cv.norm2(r)
        b = cp / c
        psi = psi + a * p
        p = b * p + r
        res = abs(cp)
        if history is not None:
            history.append(res)
        if k + 1 >= miniter:
            if eps_abs is not None and res <= eps_abs**2:
                return psi, k + 1, True
            if res <= rsq:
                return psi, k + 1, True
    return psi, k + 1, False


class eo2_ne:
    """Schur complement two with D = ``parity`` and its normal equation."""

    def __init__(self, op, parity=1):
        self.op = op
        self.pD = parity
        self.pC = 1 - parity

    def _mask(self, x, p):
        return self.op._mask(x, p)

    def Mpc(self, i_d):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.MooeeInv(i_d, pD)
        t = op.Meooe(t, pD)
        t = op.MooeeInv(t, pC)
        t = op.Meooe(t, pC)
        return i_d - t

    def MpcDag(self, i_d):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.MeooeDag(i_d, pD)
        t = op.MooeeInvDag(t, pC)
        t = op.MeooeDag(t, pC)
        t = op.MooeeInvDag(t, pD)
        return i_d - t

    def Mpc_ne(self, i_d):
        return self.MpcDag(self.Mpc(i_d))

    def R(self, src):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.MooeeInv(self._mask(src, pC), pC)
        t = op.Meooe(t, pC)
        return self._mask(src, pD) - t

    def R_ne(self, src):
        return self.MpcDag(self.R(src))

    def L(self, pc_dst):
        op, pD, pC = self.op, self.pD, self.pC
        d = op.MooeeInv(pc_dst, pD)
        c = -op.MooeeInv(op.Meooe(d, pD), pC)
        return d + c

    def L_inv(self, dst):
        return self.op.Mooee(self._mask(dst, self.pD), self.pD)

    def S(self, src):
        return self.op.MooeeInv(self._mask(src, self.pC), self.pC)


class eo1_ne:
    """Schur complement one with D = ``parity`` and its normal equation."""

    def __init__(self, op, parity=1):
        self.op = op
        self.pD = parity
        self.pC = 1 - parity

    def _mask(self, x, p):
        return self.op._mask(x, p)

    def Mpc(self, i_d):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.Meooe(i_d, pD)
        t = op.MooeeInv(t, pC)
        t = op.Meooe(t, pC)
        t = op.MooeeInv(t, pD)
        return i_d - t

    def MpcDag(self, i_d):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.MooeeInvDag(i_d, pD)
        t = op.MeooeDag(t, pD)
        t = op.MooeeInvDag(t, pC)
        t = op.MeooeDag(t, pC)
        return i_d - t

    def Mpc_ne(self, i_d):
        return self.MpcDag(self.Mpc(i_d))

    def R(self, src):
        op, pD, pC = self.op, self.pD, self.pC
        t = op.Meooe(op.MooeeInv(self._mask(src, pC), pC), pC)
        return op.MooeeInv(self._mask(src, pD) - t, pD)

    def R_ne(self, src):
        return self.MpcDag(self.R(src))

    def L(self, pc_dst):
        op, pD, pC = self.op, self.pD, self.pC
        return pc_dst - op.MooeeInv(op.Meooe(pc_dst, pD), pC)

    def L_inv(self, dst):
        return self._mask(dst, self.pD)

    def S(self, src):
        return self.op.MooeeInv(self._mask(src, self.pC), self.pC)


class eo2_kappa_ne(eo2_ne):
    """eo2_ne under the similarity transformation by the scalar kappa:
    Mpc -> kappa Mpc kappa^-1 = Mpc, R -> kappa R, L -> L kappa^-1,
    L^-1 -> kappa L^-1 (similarity_transformation.py:22-60)."""

    def __init__(self, op, parity=1):
        super().__init__(op, parity)
        self.kappa = op.kappa

    def R_ne(self, src):
        return self.MpcDag(self.kappa * self.R(src))

    def L(self, pc_dst):
        return super().L(pc_dst / self.kappa)

    def L_inv(self, dst):
        return self.kappa * super().L_inv(dst)


def preconditioned(op, src, parity=1, guess=None, pc_class=None, **cg_params):
    """Solve M dst = src through eo2_ne (or pc_class) + CG; returns (dst,
    iterations, converged)."""
    pc = (pc_class or eo2_ne)(op, parity)
    pc_src = pc.R_ne(src)
    guess = np.zeros_like(src) if guess is None else guess
    pc_dst0 = pc.L_inv(guess)
    pc_dst, it, conv = cg(pc.Mpc_ne, pc_src, pc_dst0, **cg_params)
    return pc.L(pc_dst) + pc.S(src), it, conv


def defect_correcting(outer_mat, inner_inv, src, eps=1e-15, maxiter=1000000, psi=None):
    """Outer true-residual loop; inner_inv(rhs) solves approximately.
    Returns (psi, outer iterations, converged)."""
    psi = np.zeros_like(src) if psi is None else psi.copy()
    n2src = cv.norm2(src)
    if n2src == 0.0:
        n2src = cv.norm2(outer_mat(psi)) or 1.0
    for i in range(maxiter):
        s = src - outer_mat(psi)
        e = (cv.norm2(s) / n2src) ** 0.5
        if e < eps:
            return psi, i, True
        s = s / n2src**0.5
        psi = psi + inner_inv(s) * n2src**0.5
    return psi, maxiter, False


def _qc_is_r(R):
    """Q C = R with C upper triangular from the Cholesky factor of the
    Hermitised Gram matrix R^dag R (block_cg.py:24-32); returns (Q, C)."""
    G = R.conj().T @ R
    G = 0.5 * (G + G.conj().T)
    L = np.linalg.cholesky(G)  # G = L L^dag
    C = L.conj().T
    return R @ np.linalg.inv(C), C


def block_cg(mat, srcs, eps=1e-15, maxiter=1000000, eps_abs=None, miniter=0, guess=None, history=None):
    """Block CG with Cholesky QR (block_cg.py:36-149) on the columns srcs;
    returns (X, iterations, converged).  history (optional list) gets the
    largest relative |r|^2 of the block per iteration."""
    n = len(srcs)
    shp = srcs[0].shape
    B = np.stack([s.reshape(-1) for s in srcs], axis=1)
    ssq = np.array([np.vdot(B[:, j], B[:, j]).real for j in range(n)])
    if ssq.sum() == 0.0:
        return [np.zeros(shp, B.dtype) for _ in range(n)], 0, True
    X = np.zeros_like(B) if guess is None else np.stack([x.reshape(-1) for x in guess], axis=1)

    def A(Y):
        return np.stack([mat(Y[:, j].reshape(shp)).reshape(-1) for j in range(n)], axis=1)

    Q, C = _qc_is_r(B - A(X))
    D = Q.copy()
    k = -1
    for k in range(maxiter):
        Z = A(D)
        M = np.linalg.inv(D.conj().T @ Z)
        X = X + D @ (M @ C)
        Q, S = _qc_is_r(Q - Z @ M)
        D = Q + D @ S.conj().T
        C = S @ C
        rr = np.real(np.diag(C.conj().T @ C))
        rel, ab = (rr / ssq).max(), rr.max()
        if history is not None:
            history.append(rel)
        if k + 1 >= miniter:
            if (eps_abs is not None and ab <= eps_abs**2) or rel <= eps**2:
                return [X[:, j].reshape(shp) for j in range(n)], k + 1, True
    return [X[:, j].reshape(shp) for j in range(n)], k + 1, False
