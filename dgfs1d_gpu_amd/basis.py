"""DG bases (host set-up).  Same five bases, operator matrices and SVD
significant-mode selection as the reference dgfs1D/basis.py:22-611:

  modal-orthonormal-legendre   basis.py:127-201
  nodal-sem-gll                basis.py:206-329
  nodal-gll                    basis.py:333-421
  modal-modified-karniadakis   basis.py:427-514
  nodal-sem-gll-2              basis.py:521-611

Every operator the reference hands to GiMMiK is exposed as a dense matrix
``A`` with ``out = A @ in`` over the leading (mode) axis, after the same
|x| < 1e-15 filtering (util.py:16-25): ``mat_fwd``, ``mat_bwd``, ``mat_face``,
``mat_deriv``, ``mat_invmass``, ``Ut[m] = U[m].T`` and ``V[m]``.  The H
tensor H[m,a,b] = sum_q B_mq B_aq B_bq w_q is decomposed per m by SVD; the
significant modes are the columns of U with a non-zero sum that are also
rows of V with a non-zero sum (basis.py:168-175).
"""
import numpy as np

from .quadratures import (jac_nodal_basis_at, jac_ortho_basis_at, jacobi, jacobi_diff,
                          nodal_basis_at, ortho_basis_at, zwglj, zwgrjm)

basissect = 'basis'


def filter_tol(mat, tol=1e-15, val=0.0):
    mat = np.array(mat, dtype=np.float64, copy=True)
    mat[np.abs(mat) < tol] = val
    return mat


def svd_modes(T):
    """U[m], V[m] (= diag(s) v), sigModes[m] of each slice T[m] (basis.py:168-175)."""
    K = T.shape[0]
    U, V, sig = [None] * K, [None] * K, [None] * K
    for m in range(K):
        u, s, v = np.linalg.svd(T[m])
        U[m] = filter_tol(u)
        V[m] = filter_tol(np.dot(np.diag(s), v))
        nnzU = np.where(np.sum(U[m], axis=0) != 0)[0]
        nnzV = np.where(np.sum(V[m], axis=1) != 0)[0]
        sig[m] = np.intersect1d(nnzU, nnzV)
    return U, V, sig


class Basis:
    basis_kind = None

    def upwind_ops(self, sigma=-1.0):
        """Upwind summation-by-parts derivative operators of the nodal-sem-gll
        basis (reference basis.py:303-312): Vo from the QR factorisation of
        the orthonormal basis at the nodes, S = Vo diag(0, .., 0, sigma) Vo^T,
        D+- = D +- S P^-1 / 2 with P = diag(w); returned in the 'ux = A u'
        orientation (D+^T, D-^T), as the reference's GiMMiK kernels take them."""
        if self.basis_kind != 'nodal-sem-gll':
            raise ValueError('the upwind operators are defined for nodal-sem-gll only')
        V = ortho_basis_at(self.K, self.z).T
        Vo, _ = np.linalg.qr(V)
        sig = np.zeros(self.Nq)
        sig[-1] = sigma
        S = Vo @ np.diag(sig) @ Vo.T
        invP = np.diag(1.0 / self.w)
        Dp = self.Dn + 0.5 * (S @ invP)
        Dm = self.Dn - 0.5 * (S @ invP)
        return Dp.T, Dm.T

    def limiter_ops(self):
        """(cell-average operator, slope operator of the linear part) of the
        nodal bases (reference basis.py:289-298, :408-417, :598-607):
        V = orthonormal basis at the nodes, A = V[:, :1] inv(V)[:1],
        Lin = V[:, :2] inv(V)[:2], DrLin = D^T Lin with D the nodal
        derivative matrix (d/dr; multiply by invjac for d/dx)."""
        Dn = getattr(self, 'Dn', None)
        if Dn is None or self.Nq != self.K:
            raise ValueError('the linear limiter needs a nodal basis (%s)' % self.basis_kind)
        V = ortho_basis_at(self.K, self.z).T
        invV = np.linalg.inv(V)
        A = V[:, :1] @ invV[:1, :]
        Lin = V[:, :2] @ invV[:2, :]
        return A, Dn.T @ Lin

    def __init__(self, cfg):
        self.cfg = cfg
        self.K = cfg.lookupint(basissect, 'K')
        if self.K < 1:
            raise ValueError('basis::K must be >= 1')

    def _finish(self, mat_fwd, mat_bwd, mat_face, mat_deriv, mat_invmass, U, V, sig):
        self.mat_fwd = filter_tol(mat_fwd)
        self.mat_bwd = filter_tol(mat_bwd)
        self.mat_face = filter_tol(mat_face)
        self.mat_deriv = filter_tol(mat_deriv)
        self.mat_invmass = filter_tol(mat_invmass)
        self.U, self.V, self.sigModes = U, V, sig
        self.Ut = [filter_tol(u.T) for u in U]
        self.Vm = [filter_tol(v) for v in V]
        # face reconstruction rows used by the fused transport kernel
        self.BL = self.mat_face[0, :].copy()
        self.BR = self.mat_face[-1, :].copy()

    @property
    def invmass_is_identity(self):
        return np.array_equal(self.mat_invmass, np.eye(self.K))

    def collision_tasks(self, Ne):
        """(out_slice, f_terms, g_terms) for every (m, r in sigModes[m], e):
        f = sum_a U[m][a,r] u[a,e], g = sum_b V[m][r,b] u[b,e] -> mode m, element e
        (std/std.py:352-356)."""
        tasks = []
        for m in range(self.K):
            for r in self.sigModes[m]:
                fa = [(a, self.Ut[m][r, a]) for a in range(self.K) if self.Ut[m][r, a] != 0.0]
                gb = [(b, self.Vm[m][r, b]) for b in range(self.K) if self.Vm[m][r, b] != 0.0]
                for e in range(Ne):
                    tasks.append((m * Ne + e, [(a * Ne + e, c) for a, c in fa],
                                  [(b * Ne + e, c) for b, c in gb]))
        return tasks


class ModalOrthonormalLegendre(Basis):
    basis_kind = 'modal-orthonormal-legendre'

    def __init__(self, cfg):
        super().__init__(cfg)
        K = self.K
        self.Nq = cfg.lookupordefault(basissect, 'Nq', int(np.ceil(1.5 * K)))
        self.z, self.w = zwgrjm(self.Nq, 0.0, 0.0)
        B = ortho_basis_at(K, self.z)
        self.B = B
        self.M = np.einsum('mq,q,lq->ml', B, self.w, B)
        self.invM = np.linalg.inv(self.M)
        if not np.allclose(self.M, np.eye(K)):
            raise ValueError('Insufficient quadrature points')
        self.fwdTransMat = np.einsum('mq,q->mq', B, self.w)
        Nqr = cfg.lookupint(basissect, 'Nqr')
        self.interpMat = ortho_basis_at(K, np.linspace(-1, 1, Nqr))
        H = np.einsum('mq,aq,bq,q->mab', B, B, B, self.w)
        invM_H = np.einsum('ml,mab->lab', self.invM, H)
        U, V, sig = svd_modes(invM_H)
        D = jac_ortho_basis_at(K, self.z)
        Sx = np.einsum('lq,q,mq->ml', B, self.w, D)
        self.Nqf = 2
        zqf, _ = zwglj(2, 0.0, 0.0)
        Bqf = ortho_basis_at(K, zqf)
        self.gLD, self.gRD = Bqf[:, 0].copy(), Bqf[:, -1].copy()
        self._finish(self.fwdTransMat, B.T, Bqf.T, Sx.T, self.invM.T, U, V, sig)


class NodalSemGLL(Basis):
    basis_kind = 'nodal-sem-gll'

    def _nodal_common(self, cfg):
        K = self.K
        self.Nq = K
        self.z, self.w = zwglj(K, 0.0, 0.0)
        self.B = nodal_basis_at(K, self.z, self.z).T
        self.M = np.einsum('mq,q,lq->ml', self.B, self.w, self.B)
        self.invM = np.linalg.inv(self.M)
        self.fwdTransMat = self.B
        Nqr = cfg.lookupint(basissect, 'Nqr')
        zr, _ = zwglj(Nqr, 0.0, 0.0)
        self.interpMat = nodal_basis_at(K, self.z, zr).T
        self.H = np.einsum('mq,aq,bq,q->mab', self.B, self.B, self.B, self.w)
        self.Dn = jac_nodal_basis_at(K, self.z, self.z)
        self.Nqf = 2
        zqf, _ = zwglj(2, 0.0, 0.0)
        self.Bqf = nodal_basis_at(K, self.z, zqf).T

    def __init__(self, cfg):
        super().__init__(cfg)
        self._nodal_common(cfg)
        invM_H = np.einsum('ml,mab->lab', self.invM, self.H)
        U, V, sig = svd_modes(invM_H)
        Sx = self.Dn
        self.derivMat = Sx.T
        self.gLD = self.Bqf[:, 0] / self.w[0]
        self.gRD = self.Bqf[:, -1] / self.w[-1]
        self._finish(self.fwdTransMat, self.B.T, self.Bqf.T, Sx.T, self.B.T, U, V, sig)


class NodalGLL(Basis):
    basis_kind = 'nodal-gll'

    def __init__(self, cfg):
        super().__init__(cfg)
        K = self.K
        self.Nq = K
        self.z, self.w = zwglj(K, 0.0, 0.0)
        self.B = ortho_basis_at(K, self.z).T
        self.invM = self.B @ self.B.T
        self.M = np.linalg.inv(self.invM)
        I = np.eye(K)
        self.fwdTransMat = I
        Nqr = cfg.lookupint(basissect, 'Nqr')
        zr, _ = zwglj(Nqr, 0.0, 0.0)
        self.interpMat = nodal_basis_at(K, self.z, zr).T
        H = np.einsum('mq,aq,bq->mab', I, I, I)     # mass-lumped (basis.py:368)
        U, V, sig = svd_modes(H)
        Sx = jac_nodal_basis_at(K, self.z, self.z)
        self.Dn = Sx
        self.derivMat = Sx.T
        self.Nqf = K
        self.gLD, self.gRD = self.invM[0, :].copy(), self.invM[-1, :].copy()
        self._finish(I, I, I.T, Sx.T, I, U, V, sig)


class ModalModifiedKarniadakis(Basis):
    basis_kind = 'modal-modified-karniadakis'

    @staticmethod
    def modified_basis_at(order, pts):
        pts = np.asarray(pts, dtype=np.float64)
        jp = jacobi(order - 1, 1.0, 1.0, pts)
        rows = [0.5 * (1 - pts)] + [jp[p] * 0.25 * (1 - pts ** 2) for p in range(order - 2)]
        return np.vstack(rows + [0.5 * (1 + pts)])

    @staticmethod
    def jac_modified_basis_at(order, pts):
        pts = np.asarray(pts, dtype=np.float64)
        jp = jacobi(order - 1, 1.0, 1.0, pts)
        djp = jacobi_diff(order - 1, 1.0, 1.0, pts)
        rows = [-0.5 * np.ones(len(pts))]
        rows += [djp[p] * 0.25 * (1.0 - pts ** 2) - jp[p] * 0.5 * pts for p in range(order - 2)]
        return np.vstack(rows + [0.5 * np.ones(len(pts))])

    def __init__(self, cfg):
        super().__init__(cfg)
        K = self.K
        self.Nq = cfg.lookupordefault(basissect, 'Nq', int(np.ceil(1.5 * K)))
        self.z, self.w = zwgrjm(self.Nq, 0.0, 0.0)
        B = self.modified_basis_at(K, self.z)
        self.B = B
        self.M = np.einsum('mq,q,lq->ml', B, self.w, B)
        self.invM = np.linalg.inv(self.M)
        self.fwdTransMat = np.einsum('ml,lq,q->mq', self.invM, B, self.w)
        Nqr = cfg.lookupint(basissect, 'Nqr')
        self.interpMat = self.modified_basis_at(K, np.linspace(-1, 1, Nqr))
        H = np.einsum('mq,aq,bq,q->mab', B, B, B, self.w)
        U, V, sig = svd_modes(H)          # the reference decomposes H itself (basis.py:473-479)
        D = self.jac_modified_basis_at(K, self.z)
        Sx = np.einsum('lq,q,mq->ml', B, self.w, D)
        self.Nqf = 2
        zqf, _ = zwglj(2, 0.0, 0.0)
        Bqf = self.modified_basis_at(K, zqf)
        self.gLD, self.gRD = Bqf[:, 0].copy(), Bqf[:, -1].copy()
        self._finish(self.fwdTransMat, B.T, Bqf.T, Sx.T, self.invM.T, U, V, sig)


class NodalSemGLL2(NodalSemGLL):
    basis_kind = 'nodal-sem-gll-2'

    def __init__(self, cfg):
        Basis.__init__(self, cfg)
        self._nodal_common(cfg)
        U, V, sig = svd_modes(self.H)     # invM_H = H (basis.py:554-555)
        Sx = np.matmul(self.M, self.Dn.T).T
        self.derivMat = Sx.T
        self.gLD = self.Bqf[:, 0].copy()
        self.gRD = self.Bqf[:, -1].copy()
        self._finish(self.fwdTransMat, self.B.T, self.Bqf.T, Sx.T, self.invM.T, U, V, sig)
