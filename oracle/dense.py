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


def sldf_criterion(D: "Dense", y_proj, Q, V, h2: float, tau0: float):
    """REML_criterion (src/reml.cpp:457-489) evaluated exactly in FP64: yPy and Q'H^-1Q by dense
    solves, and the SLQ log-determinant replaced by its limit n/nV sum_k v_k' log(A + sigma I) v_k
    for the same probes V.  Returns (slogLL, vg, ve, lrt)."""
    n = D.n
    c = Q.shape[1]
    tau = (1 - h2) / h2
    yp = np.asarray(y_proj, np.float64)
    x = D.solve("quadform", tau, yp)
    yPy = float(yp @ x)
    ySy = float(yp @ yp)
    vg = yPy / (n - c)
    ve = vg * tau
    ldetH = D.quad_logm("ldet", tau, V) * n / V.shape[1]
    XQ = D.solve("ldet", tau, np.asarray(Q, np.float64))
    logdet = np.linalg.slogdet(np.asarray(Q, np.float64).T @ XQ)[1]
    lrt = n * np.log(ySy / yPy) - ldetH
    return (n - c) * np.log(ySy / yPy) - logdet - ldetH, vg, ve, lrt


def sldf_gss(D: "Dense", y_proj, Q, V, s2min=1e-4, s2max=0.7, tol=1e-4):
    """REML_GSS (src/reml.cpp:493-547) in FP64 over sldf_criterion; returns the last evaluated
    point (vg, ve, lrt, h2) exactly as the reference does."""
    invphi = (np.sqrt(5) - 1) / 2
    invphi2 = (3 - np.sqrt(5)) / 2
    a, b = s2min, s2max
    h = b - a
    tau0 = (1 - s2max) / s2max
    nitr = 1 if h < tol else int(np.ceil(np.log(tol / h) / np.log(invphi)))
    c, d = a + invphi2 * h, a + invphi * h
    fc, *_ = sldf_criterion(D, y_proj, Q, V, c, tau0)
    fd, vg, ve, lrt = sldf_criterion(D, y_proj, Q, V, d, tau0)
    last = (vg, ve, lrt)
    for _ in range(nitr):
        if fc > fd:
            b, d, fd = d, c, fc
            h *= invphi
            c = a + invphi2 * h
            fc, *last = sldf_criterion(D, y_proj, Q, V, c, tau0)
        else:
            a, c, fc = c, d, fd
            h *= invphi
            d = a + invphi * h
            fd, *last = sldf_criterion(D, y_proj, Q, V, d, tau0)
    vg, ve, lrt = last
    return vg, ve, lrt, vg / (vg + ve)


def lmm_gstat_fake(D: "Dense", Py, tau1: float, V0):
    """Quadratic forms of the --fake_geno branch (src/reml.cpp:900-917) in FP64 with exact solves,
    for uncentred fake dosages V0 (n x k): the columns centred by their float mean (one float
    divide of the exact dosage sum), xPy = x'Py, v = P R^-1/2 x, xPx = v'(A_quad0 + tau1 I)^-1 v,
    xSx = v'v."""
    V0 = np.asarray(V0, np.float32)
    n = V0.shape[0]
    mu = (V0.astype(np.float64).sum(axis=0).astype(np.float32) / np.float32(n)).astype(np.float32)
    X = (V0 - mu[None, :]).astype(np.float64)
    xPy = X.T @ np.asarray(Py, dtype=np.float64)
    V = D.P @ (D.rr[:, None] * X)
    xPx = np.sum(V * D.solve("quadform", tau1, V), axis=0)
    xSx = np.sum(V * V, axis=0)
    return xPy, xPx, xSx


def lmm_gstat(D: "Dense", Py, tau1: float, snps, window_size: int):
    """Quadratic forms of --lmm (src/reml.cpp:920-980, real-genotype branch) in FP64 with exact
    solves: xPy = x'Py, xPx = v'(A_quad0 + tau1 I)^-1 v, xSx = v'v with v = P R^-1/2 (x - gm),
    xBx = (v'v - z'(Zw'Zw + tau1 diag(1/gw))^-1 z) / tau1 with Zw the projected, scaled window of
    window_size SNPs around the SNP and z = Zw'v."""
    snps = np.asarray(snps, dtype=np.int64)
    m, w = D.m, int(window_size)
    X = D.Zc[:, snps]
    xPy = X.T @ np.asarray(Py, dtype=np.float64)
    V = D.P @ (D.rr[:, None] * X)
    xPx = np.sum(V * D.solve("quadform", tau1, V), axis=0)
    xSx = np.sum(V * V, axis=0)
    xBx = np.zeros(len(snps))
    for i, j in enumerate(snps):
        wds = 0 if j < w // 2 else (m - w if j > m - w // 2 else j - w // 2)
        Zw = D.P @ (D.rr[:, None] * D.Zc[:, wds:wds + w])
        with np.errstate(divide="ignore"):
            ridge = tau1 / D.gw[wds:wds + w]
        G = Zw.T @ Zw + np.diag(ridge)
        z = Zw.T @ V[:, i]
        xBx[i] = (xSx[i] - z @ np.linalg.solve(G, z)) / tau1
    return xPy, xPx, xBx, xSx
