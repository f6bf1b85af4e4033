"""ORACLE -- test infrastructure only.  FP64 dense restatement of the operators.

For small n the genomic relationship operator of src/reml.cpp is formed explicitly:
  Zc = Z - 1 gm'                      (Z = low+high bit dosages, gm the float32 genotype means)
  K  = Zc diag(gw) Zc'
  ldet(v)     = diag(rr) K diag(rr) v + tau0 v                       (H0_ldet_mv, :52-94)
  quadform(v) = P diag(rr) K diag(rr) P v + tau0 v, P = I - Q Q'     (H0_quadform_mv, :8-50)
and solves / log-determinants are exact (Cholesky / symmetric eigendecomposition).
"""
from __future__ import annotations

import numpy as np


def dosage_of_kmat(kmat: np.ndarray) -> np.ndarray:
    """kmat bits (m, 2n) -> dosage matrix Z (n, m) float64."""
    k = np.asarray(kmat, dtype=np.float64)
    return (k[:, 0::2] + k[:, 1::2]).T.copy()


class Dense:
    def __init__(self, Z: np.ndarray, gm, gw, rr, Q=None):
        self.Z = np.asarray(Z, dtype=np.float64)
        self.n, self.m = self.Z.shape
        self.gm = np.asarray(gm, dtype=np.float64)
        self.gw = np.asarray(gw, dtype=np.float64)
        self.rr = np.asarray(rr, dtype=np.float64)
        self.Zc = self.Z - self.gm[None, :]
        self.K = (self.Zc * self.gw[None, :]) @ self.Zc.T
        if Q is None:
            self.Q = np.zeros((self.n, 0))
        else:
            Qa = np.asarray(Q, dtype=np.float64)
            self.Q = Qa[:, None] if Qa.ndim == 1 else Qa
        self.P = np.eye(self.n) - self.Q @ self.Q.T
        self.A_ldet0 = self.rr[:, None] * self.K * self.rr[None, :]
        self.A_quad0 = self.P @ self.A_ldet0 @ self.P

    def A(self, kind: str, tau0: float) -> np.ndarray:
        base = self.A_quad0 if kind == "quadform" else self.A_ldet0
        return base + tau0 * np.eye(self.n)

    def op(self, kind: str, tau0: float, V) -> np.ndarray:
        V = np.asarray(V, dtype=np.float64)
        return self.A(kind, tau0) @ V

    def rmatvec(self, V, weighted=False):
        out = self.Zc.T @ np.asarray(V, dtype=np.float64)
        return self.gw[:, None] * out if weighted else out

    def matvec(self, U):
        return self.Zc @ np.asarray(U, dtype=np.float64)

    def solve(self, kind: str, tau: float, B) -> np.ndarray:
        """(A0 + tau I)^-1 B, the quantity L_Solve approximates with sigma = tau - tau0."""
        return np.linalg.solve(self.A(kind, tau), np.asarray(B, dtype=np.float64))

    def quad_logm(self, kind: str, tau: float, V) -> float:
        """sum_k v_k' log(A0 + tau I) v_k -- the exact value SLQ estimates for fixed probes."""
        w, U = np.linalg.eigh(self.A(kind, tau))
        C = U.T @ np.asarray(V, dtype=np.float64)
        return float(np.sum((C ** 2) * np.log(w)[:, None]))

    def logdet(self, kind: str, tau: float) -> float:
        sign, ld = np.linalg.slogdet(self.A(kind, tau))
        return float(ld)


def reml_loglik(Z, gm, gw, X, y, h2: float, rr=None) -> float:
    """Exact restricted log-likelihood (up to a constant) at h2 with sigma^2 profiled out:
    -(1/2)[(n-c) log(y'Py) + log|H| + log|X'H^-1X|], H = Zc W Zc' + tau R, tau = (1-h2)/h2."""
    Z = np.asarray(Z, np.float64)
    n = Z.shape[0]
    Zc = Z - np.asarray(gm, np.float64)[None, :]
    K = (Zc * np.asarray(gw, np.float64)[None, :]) @ Zc.T
    R = np.ones(n) if rr is None else 1.0 / np.asarray(rr, np.float64) ** 2
    tau = (1 - h2) / h2
    H = K + tau * np.diag(R)
    X = np.asarray(X, np.float64)
    X = X[:, None] if X.ndim == 1 else X
    c = X.shape[1]
    Hi = np.linalg.inv(H)
    XtHiX = X.T @ Hi @ X
    P = Hi - Hi @ X @ np.linalg.solve(XtHiX, X.T @ Hi)
    yPy = float(y @ P @ y)
    return -0.5 * ((n - c) * np.log(yPy) + np.linalg.slogdet(H)[1] + np.linalg.slogdet(XtHiX)[1])
