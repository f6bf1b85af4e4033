"""ORACLE -- test infrastructure only.  FP64 dense restatement of the operators.

For small n the genomic relationship operator of src/reml.cpp is formed explicitly:
  Zc = Z - 1 gm'                      (Z = low+high bit dosages, gm the float32 genotype means)
  K  = Zc diag(gw) Zc'
  ldet(v)     = diag(rr) K diag(rr) v + tau0 v                       (H0_ldet_mv, :52-94)
  quadform(v) = P diag(rr) K diag(rr) P v + tau0 v, P = I - Q Q'     (H0_quadform_mv, :8-50)
and solves / log-determinants are exact: each operator kind is eigendecomposed once in FP64 and
every shift A0 + tau I is applied through that decomposition.
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
        self._eig = {}

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

    def eig(self, kind: str):
        """Eigendecomposition (w, U) of the unshifted operator A0 of `kind`, computed once: every
        shift tau then costs only products with U, since A0 + tau I = U diag(w + tau) U'."""
        if kind not in self._eig:
            base = self.A_quad0 if kind == "quadform" else self.A_ldet0
            self._eig[kind] = np.linalg.eigh(base)
        return self._eig[kind]

    def solve(self, kind: str, tau: float, B) -> np.ndarray:
        """(A0 + tau I)^-1 B, the quantity L_Solve approximates with sigma = tau - tau0."""
        w, U = self.eig(kind)
        B = np.asarray(B, dtype=np.float64)
        return U @ ((U.T @ B) / (w + tau).reshape((-1,) + (1,) * (B.ndim - 1)))

    def quad_logm(self, kind: str, tau: float, V) -> float:
        """sum_k v_k' log(A0 + tau I) v_k -- the exact value SLQ estimates for fixed probes."""
        w, U = self.eig(kind)
        C = U.T @ np.asarray(V, dtype=np.float64)
        C = C.reshape(C.shape[0], -1)
        return float(np.sum((C ** 2) * np.log(w + tau)[:, None]))

    def logdet(self, kind: str, tau: float) -> float:
        w, _ = self.eig(kind)
        return float(np.sum(np.log(w + tau)))


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


def lmc_inputs(D: "Dense", y, u, e) -> np.ndarray:
    """The Lanczos block of LMC_REML (src/reml.cpp:788-813) in FP64 from the same float draws:
    V = (I - QQ') [R^-1/2 y, R^-1/2 Zc (sqrt(gw) u), e], with sqrt(gw) u a float product as in
    the reference (u: m x n_MC normals, e: n x n_MC normals)."""
    gw32 = np.asarray(D.gw, np.float32)
    us = (np.asarray(u, np.float32) * np.sqrt(gw32)[:, None]).astype(np.float64)
    V = np.concatenate([np.asarray(y, np.float64)[:, None], D.Zc @ us, np.asarray(e, np.float64)], axis=1)
    k = us.shape[1]
    V[:, :1 + k] *= D.rr[:, None]
    return D.P @ V


def lmc_first_order(D: "Dense", V, h2: float, expect: bool = False):
    """REML_first_order (src/reml.cpp:696-728) in FP64 with exact solves of the projected system
    (A_quad0 + tau I) X = V, tau = (1 - h2)/h2:
      Py_0 = P x_0, Py_k = P (x_k + sqrt(tau) x_{n_MC+k}),  e = Py,  u = sqrt(gw) Zc' R^-1/2 Py,
      f = log[(|u_0|^2/|e_0|^2) / (sum_k |u_k|^2 / sum_k |e_k|^2)].
    With expect=True the MC ratio is replaced by its expectation tr(A M)/tr(M), M the projected
    inverse (I-QQ')(A_quad0 + tau I)^-1 (I-QQ'), for which the root of f is the REML optimum.
    Returns (f, Py_0) with Py_0 already scaled by R^-1/2 (the Py of the fit)."""
    tau = (1 - h2) / h2
    k = (V.shape[1] - 1) // 2
    X = D.solve("quadform", tau, V)
    Py = np.concatenate([X[:, :1], X[:, 1:1 + k] + np.sqrt(tau) * X[:, 1 + k:]], axis=1)
    Py = D.P @ Py
    eB = Py
    Pr = D.rr[:, None] * Py
    uB = np.sqrt(D.gw)[:, None] * (D.Zc.T @ Pr)
    ratio_y = np.sum(uB[:, 0] ** 2) / np.sum(eB[:, 0] ** 2)
    if expect:
        w, U = D.eig("quadform")
        M = D.P @ (U * (1.0 / (w + tau))[None, :]) @ U.T @ D.P
        ratio_mc = float(np.sum(D.A_ldet0 * M)) / float(np.trace(M))
    else:
        ratio_mc = np.sum(uB[:, 1:] ** 2) / np.sum(eB[:, 1:] ** 2)
    return float(np.log(ratio_y / ratio_mc)), Pr[:, 0]


def lmc_secant(f, s2max: float = 0.7, tol: float = 0.002, max_iter: int = 9):
    """The secant iterations of LMC_REML on log((1-h2)/h2) (src/reml.cpp:821-853): start at
    s2max/2, then s2max/4 or s2max by the sign of f, update
    lt = (lt1 f2 - lt2 f1)/(f2 - f1) clamped to [-9, 9], stop when |lt - lt2| < tol or after
    max_iter evaluations in all (the reference: tol 0.002, 9).  In double.  Returns (h2, the
    value f returned at it, evaluations)."""
    s2 = s2max / 2
    lt1 = np.log((1 - s2) / s2)
    r1 = f(s2)
    s2 = s2max / 4 if r1[0] < 0 else s2max
    lt2 = np.log((1 - s2) / s2)
    r2 = f(s2)
    evals = 2
    for _ in range(3, max_iter + 1):
        lt = (lt1 * r2[0] - lt2 * r1[0]) / (r2[0] - r1[0])
        lt = min(max(lt, -9.0), 9.0)
        if abs(lt - lt2) < tol:
            break
        s2 = 1 / (np.exp(lt) + 1)
        r1, r2 = r2, f(s2)
        evals += 1
        lt1, lt2 = lt2, lt
    return s2, r2, evals


def lmc_reml(D: "Dense", y, u, e, s2max: float = 0.7, tol: float = 0.002, max_iter: int = 9, expect=False):
    """LMC_REML's variance components (src/reml.cpp:730-862) in FP64 over the same MC draws:
    h2 from lmc_secant, vg = y'Py/(n - c) at it (:856), ve = vg (1-h2)/h2.  Returns
    dict(h2, vg, ve, Py, evals)."""
    V = lmc_inputs(D, y, u, e)
    h2, (fv, Py), evals = lmc_secant(lambda s: lmc_first_order(D, V, s, expect), s2max, tol, max_iter)
    c = D.Q.shape[1]
    vg = float(np.asarray(y, np.float64) @ Py) / (D.n - c)
    return dict(h2=float(h2), vg=vg, ve=vg * (1 - h2) / h2, Py=Py, evals=evals, f=fv)


def lmc_fit(Z, gm, gw, rvec, X, y, rng_seed=0, n_MC=None, s2max: float = 0.7, tol: float = 0.002,
            max_iter: int = 9):
    """Every output of LMC_REML (src/reml.cpp:730-882) in FP64 with exact solves, over the float MC
    draws of the reference's stream (mt19937(seed + 54321), src/reml.cpp:788-801): h2, vg, ve and
    Py from lmc_reml; snp_blup = gw Zc' Py (:866); blue = (X'X)^-1 X' (y - Zc snp_blup - tau1 R Py)
    (:870-877).  Q is the FP64 orthonormal basis of R^-1/2 X.  n_MC defaults to the reference's
    rule max(min(4e9/n^2, 30), 3) (src/option.cpp:414)."""
    import oracle
    y = np.asarray(y, np.float64)
    n = len(y)
    m = np.asarray(gm).shape[0]
    if n_MC is None:
        n_MC = max(min(int(4e9 / n / n), 30), 3)
    rv = np.ones(n) if rvec is None else np.asarray(rvec, np.float64)
    rr = 1.0 / np.sqrt(rv)
    X = np.asarray(X, np.float64).reshape(n, -1)
    Q, R = np.linalg.qr(rr[:, None] * X)
    if np.min(np.abs(np.diag(R))) <= 1e-6 * np.max(np.abs(np.diag(R))):
        raise ValueError("Fixed-effects design matrix is not of full rank.")
    D = Dense(Z, gm, gw, rr, Q)
    u, e = oracle.lmc_samples(rng_seed, n, m, n_MC)
    r = lmc_reml(D, y, u, e, s2max=s2max, tol=tol, max_iter=max_iter)
    Py = r["Py"]
    snp_blup = D.rmatvec(Py[:, None], weighted=True)[:, 0]
    tau1 = (1 - r["h2"]) / r["h2"]
    resid = y - (D.matvec(snp_blup) + tau1 * rv * Py)
    blue = np.linalg.solve(X.T @ X, X.T @ resid)
    return dict(r, snp_blup=snp_blup, blue=blue, n_MC=n_MC)


def iter_weighting_reml(kmat, gm, rvec, X, y, maf, persnp=None, window_size=20, sig_chisq=1e6, indep_r2=0.5,
                        rng_seed=0, n_MC=None, s2max: float = 0.7, tol: float = 0.002, max_iter: int = 9):
    """--iter_weighting (src/option.cpp:417-528) with LMC fits in FP64 (lmc_fit): the restated
    driver of oracle/reml.py (float per-SNP variances, clumping, window-mean logistic weights as
    the reference computes them) over exact fits on the same MC draws."""
    import oracle
    from oracle import reml as oreml
    Z = dosage_of_kmat(kmat)

    def fit(sldf, pw, Xf):
        if sldf:
            raise ValueError("the FP64 driver restates the LMC fits (lrt = False)")
        return lmc_fit(Z, gm, oracle.gw(gm, pw), rvec, Xf, y, rng_seed, n_MC, s2max, tol, max_iter)

    return oreml.iter_weighting_reml(kmat, gm, rvec, X, y, maf, persnp=persnp, window_size=window_size,
                                     sig_chisq=sig_chisq, indep_r2=indep_r2, lrt=False, rng_seed=rng_seed,
                                     n_MC=n_MC, s2max=s2max, fit=fit)


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
