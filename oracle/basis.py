"""Oracle: DG basis operators (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates dgfs1D/basis.py with the oracle's own quadratures:
  modal-orthonormal-legendre :127-201, nodal-sem-gll :206-329,
  nodal-gll :333-421, modal-modified-karniadakis :427-514,
  nodal-sem-gll-2 :521-611.
Operators are returned as the matrices the reference gives GiMMiK
(util.py:20-25, out = A @ in, entries |x| < 1e-15 zeroed).
"""
import numpy as np

from . import quadratures as Q


def _ft(a, tol=1e-15):
    a = np.array(a, dtype=np.float64, copy=True)
    a[np.abs(a) < tol] = 0.0
    return a


def _svd(T):
    K = T.shape[0]
    U, V, sig = [], [], []
    for m in range(K):
        u, s, v = np.linalg.svd(T[m])
        Um, Vm = _ft(u), _ft(np.diag(s) @ v)
        nu = np.where(Um.sum(axis=0) != 0)[0]
        nv = np.where(Vm.sum(axis=1) != 0)[0]
        U.append(Um)
        V.append(Vm)
        sig.append(np.intersect1d(nu, nv))
    return U, V, sig


class OracleBasis:
    def __init__(self, kind, K, Nq=None, Nqr=8):
        self.kind, self.K = kind, K
        getattr(self, '_' + kind.replace('-', '_'))(K, Nq, Nqr)
        self.BL = self.face[0].copy()
        self.BR = self.face[-1].copy()

    def _modal_orthonormal_legendre(self, K, Nq, Nqr):
        Nq = Nq or int(np.ceil(1.5 * K))
        z, w = Q.zwgrjm(Nq, 0.0, 0.0)
        B = Q.ortho_basis_at(K, z)
        M = np.einsum('mq,q,lq->ml', B, w, B)
        invM = np.linalg.inv(M)
        H = np.einsum('mq,aq,bq,q->mab', B, B, B, w)
        self.U, self.V, self.sig = _svd(np.einsum('ml,mab->lab', invM, H))
        D = Q.jac_ortho_basis_at(K, z)
        Sx = np.einsum('lq,q,mq->ml', B, w, D)
        Bqf = Q.ortho_basis_at(K, np.array([-1.0, 1.0]))
        self.gLD, self.gRD = Bqf[:, 0].copy(), Bqf[:, -1].copy()
        self.z, self.Nq = z, Nq
        self.fwdT = np.einsum('mq,q->mq', B, w)
        self.fwd, self.bwd, self.face = _ft(self.fwdT), _ft(B.T), _ft(Bqf.T)
        self.deriv, self.invmass = _ft(Sx.T), _ft(invM.T)
        self.interp = Q.ortho_basis_at(K, np.linspace(-1, 1, Nqr))

    def _nodal(self, K, Nqr):
        z, w = Q.zwglj(K, 0.0, 0.0)
        B = Q.nodal_basis_at(K, z, z).T
        M = np.einsum('mq,q,lq->ml', B, w, B)
        self.z, self.w, self.B, self.M, self.invM, self.Nq = z, w, B, M, np.linalg.inv(M), K
        self.H = np.einsum('mq,aq,bq,q->mab', B, B, B, w)
        self.D = Q.jac_nodal_basis_at(K, z, z)
        self.Bqf = Q.nodal_basis_at(K, z, np.array([-1.0, 1.0])).T
        zr, _ = Q.zwglj(Nqr, 0.0, 0.0)
        self.interp = Q.nodal_basis_at(K, z, zr).T
        self.fwdT = B

    def _nodal_sem_gll(self, K, Nq, Nqr):
        self._nodal(K, Nqr)
        self.U, self.V, self.sig = _svd(np.einsum('ml,mab->lab', self.invM, self.H))
        self.gLD, self.gRD = self.Bqf[:, 0] / self.w[0], self.Bqf[:, -1] / self.w[-1]
        self.fwd, self.bwd, self.face = _ft(self.B), _ft(self.B.T), _ft(self.Bqf.T)
        self.deriv, self.invmass = _ft(self.D.T), _ft(self.B.T)

    def _nodal_sem_gll_2(self, K, Nq, Nqr):
        self._nodal(K, Nqr)
        self.U, self.V, self.sig = _svd(self.H)
        Sx = (self.M @ self.D.T).T
        self.gLD, self.gRD = self.Bqf[:, 0].copy(), self.Bqf[:, -1].copy()
        self.fwd, self.bwd, self.face = _ft(self.B), _ft(self.B.T), _ft(self.Bqf.T)
        self.deriv, self.invmass = _ft(Sx.T), _ft(self.invM.T)

    def _nodal_gll(self, K, Nq, Nqr):
        z, w = Q.zwglj(K, 0.0, 0.0)
        B = Q.ortho_basis_at(K, z).T
        invM = B @ B.T
        I = np.eye(K)
        self.U, self.V, self.sig = _svd(np.einsum('mq,aq,bq->mab', I, I, I))
        D = Q.jac_nodal_basis_at(K, z, z)
        self.gLD, self.gRD = invM[0, :].copy(), invM[-1, :].copy()
        self.z, self.Nq = z, K
        self.fwdT = I
        self.fwd, self.bwd, self.face, self.deriv, self.invmass = I, I, I, _ft(D.T), I
        zr, _ = Q.zwglj(Nqr, 0.0, 0.0)
        self.interp = Q.nodal_basis_at(K, z, zr).T

    @staticmethod
    def _mod(K, x):
        jp = Q.jacobi(K - 1, 1.0, 1.0, x)
        return np.vstack([0.5 * (1 - x)] + [jp[p] * 0.25 * (1 - x ** 2) for p in range(K - 2)]
                         + [0.5 * (1 + x)])

    @staticmethod
    def _dmod(K, x):
        jp = Q.jacobi(K - 1, 1.0, 1.0, x)
        dj = Q.jacobi_diff(K - 1, 1.0, 1.0, x)
        return np.vstack([-0.5 * np.ones(len(x))]
                         + [dj[p] * 0.25 * (1 - x ** 2) - jp[p] * 0.5 * x for p in range(K - 2)]
                         + [0.5 * np.ones(len(x))])

    def _modal_modified_karniadakis(self, K, Nq, Nqr):
        Nq = Nq or int(np.ceil(1.5 * K))
        z, w = Q.zwgrjm(Nq, 0.0, 0.0)
        B = self._mod(K, z)
        M = np.einsum('mq,q,lq->ml', B, w, B)
        invM = np.linalg.inv(M)
        self.U, self.V, self.sig = _svd(np.einsum('mq,aq,bq,q->mab', B, B, B, w))
        Sx = np.einsum('lq,q,mq->ml', B, w, self._dmod(K, z))
        Bqf = self._mod(K, np.array([-1.0, 1.0]))
        self.gLD, self.gRD = Bqf[:, 0].copy(), Bqf[:, -1].copy()
        self.z, self.Nq = z, Nq
        self.fwdT = np.einsum('ml,lq,q->mq', invM, B, w)
        self.fwd, self.bwd, self.face = _ft(self.fwdT), _ft(B.T), _ft(Bqf.T)
        self.deriv, self.invmass = _ft(Sx.T), _ft(invM.T)
        self.interp = self._mod(K, np.linspace(-1, 1, Nqr))
