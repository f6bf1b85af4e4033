"""ORACLE -- test infrastructure only.  SLDF_REML and LMC_REML (src/reml.cpp:457-882)
restated over the FP32 CPU primitives of liboracle.so (operators, L_Seed, L_Solve, RNG).

Scalars follow the reference's FP32 (numpy float32) where the reference computes in float.
The tridiagonal eigendecomposition (Eigen's SelfAdjointEigenSolver in float, src/reml.cpp:641-646)
uses scipy's LAPACK routine in FP64, and the covariate QR uses numpy (FP64): neither library is
available in the reference's own form here (parity of those two steps is at the rounding level
of their results, not bit-exact).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import eigh_tridiagonal

import oracle

f32 = np.float32


def vg_from_py(y, Py, c: int) -> float:
    """Vg = y'Py / (n - c) (src/reml.cpp:856)."""
    y = np.asarray(y, np.float64)
    return float(np.dot(y, np.asarray(Py, np.float64)) / (len(y) - c))


def covariate_qr(X: np.ndarray, rr: np.ndarray):
    """Q (n x c, float32) and R with Q R = diag(rr) X (src/reml.cpp:583-588)."""
    A = rr.astype(np.float64)[:, None] * np.asarray(X, np.float64)
    Q, R = np.linalg.qr(A)
    if np.min(np.abs(np.diag(R))) <= 1e-6 * np.max(np.abs(np.diag(R))):
        raise ValueError("Fixed-effects design matrix is not of full rank.")
    return Q.astype(np.float32), R


def slq_spectrum(seed: "oracle.Seed", k: int):
    """Eigenvalues D and first eigenvector components W of probe k's Jacobi matrix."""
    J = seed.steps
    d = seed.delta[k, :J].astype(np.float64)
    e = seed.beta[k, 1:J].astype(np.float64)
    D, Wv = eigh_tridiagonal(d, e)
    return D, Wv[0, :]


def _engine(kmat, engine):
    """The operator engine of the drivers: oracle.Model over the kmat bits (the C++ restatement),
    or engine(gm, gw, rr, Q, subset_size) -- e.g. the MKL panel engine of oracle.cpu_panels."""
    if engine is not None:
        return engine
    return lambda gm, gw, rr, Q, ss: oracle.Model(kmat, gm, gw, rr, Q, ss)


def SLDF_REML(kmat, gm, gw, rvec, X, y, rng_seed=0, n_V=30, tol_L=5e-4, s2max=0.7, s2min=1e-4,
              tol_VC=1e-4, subset_size=1000, engine=None):
    """Returns dict(vg, ve, lrt, h2, Py, snp_blup, blue, var_blue, steps, evals)."""
    n = len(y)
    rvec = np.ones(n, np.float32) if rvec is None else np.asarray(rvec, np.float32)
    rr = (1.0 / np.sqrt(rvec.astype(np.float64))).astype(np.float32)
    X = np.asarray(X, np.float64)
    X = X[:, None] if X.ndim == 1 else X
    c = X.shape[1]
    Q, qrR = covariate_qr(X, rr)
    yp = (rr * np.asarray(y, np.float32)).astype(np.float32)
    yp = (yp - Q @ (Q.T @ yp)).astype(np.float32)
    V = oracle.rademacher(rng_seed, n, n_V)
    mdl = _engine(kmat, engine)(gm, gw, rr, Q, subset_size)
    tau0 = f32((1 - f32(s2max)) / f32(s2max))
    seed_y = mdl.L_Seed("quadform", float(tau0), yp, tol_L)
    seed_Q = mdl.L_Seed("ldet", float(tau0), Q, tol_L)
    seed_V = mdl.L_Seed("ldet", float(tau0), V, tol_L)
    spec = [slq_spectrum(seed_V, k) for k in range(n_V)]
    ySy = float(np.dot(yp.astype(np.float64), yp))
    evals = [0]

    def criterion(ss):
        evals[0] += 1
        tau = f32((1 - f32(ss)) / f32(ss))
        sigma = f32(tau - tau0)
        x = seed_y.L_Solve(yp, float(sigma))[:, 0]
        yPy = float(np.dot(yp.astype(np.float64), x))
        vg = f32(yPy / (n - c))
        ve = f32(vg * tau)
        ldetH = sum(float(np.sum(W ** 2 * np.log(D + sigma))) for D, W in spec) * n / n_V
        lrt = n * math.log(ySy / yPy) - ldetH
        XQ = seed_Q.L_Solve(Q, float(sigma))
        QtX = Q.T.astype(np.float64) @ XQ
        logdet = np.linalg.slogdet(QtX)[1]
        return (n - c) * math.log(ySy / yPy) - logdet - ldetH, vg, ve, lrt

    invphi = 0.618033988749894848204586834365638117720309179805762862135
    invphi2 = 0.381966011250105151795413165634361882279690820194237137864
    a, b = f32(s2min), f32(s2max)
    h = f32(b - a)
    nitr = 1 if h < tol_VC else int(math.ceil(math.log(f32(tol_VC) / h) / math.log(invphi)))
    cc = f32(a + invphi2 * h)
    dd = f32(a + invphi * h)
    fc, vg, ve, lrt = criterion(cc)
    fd, vg, ve, lrt = criterion(dd)
    for _ in range(nitr):
        if fc > fd:
            b = dd
            dd = cc
            fd = fc
            h = f32(h * invphi)
            cc = f32(a + invphi2 * h)
            fc, vg, ve, lrt = criterion(cc)
        else:
            a = cc
            cc = dd
            fc = fd
            h = f32(h * invphi)
            dd = f32(a + invphi * h)
            fd, vg, ve, lrt = criterion(dd)
    tau1 = f32(ve / vg)
    Py = seed_y.L_Solve(yp, float(tau1 - tau0))[:, 0]
    Py = (Py - Q @ (Q.T @ Py)).astype(np.float32)
    Py = (rr * Py).astype(np.float32)
    snp_blup = mdl.rmatvec(Py, weighted=True)[:, 0]
    XQ = seed_Q.L_Solve(Q, float(tau1 - tau0)).astype(np.float64)
    HinvX = rr.astype(np.float64)[:, None] * (XQ @ qrR)
    XtHX = X.T @ HinvX
    blue = np.linalg.solve(XtHX, HinvX.T @ np.asarray(y, np.float64))
    var_blue = np.linalg.inv(XtHX) * float(vg)
    return dict(vg=float(vg), ve=float(ve), lrt=float(lrt), h2=float(vg / (vg + ve)), Py=Py,
                snp_blup=snp_blup, blue=blue, var_blue=var_blue,
                steps=(seed_y.steps, seed_Q.steps, seed_V.steps), evals=evals[0])


def LMC_REML(kmat, gm, gw, rvec, X, y, rng_seed=0, n_MC=None, tol_L=5e-4, s2max=0.7, subset_size=1000,
             engine=None):
    """Returns dict(vg, ve, h2, Py, snp_blup, blue, steps, evals)."""
    n = len(y)
    m = len(gm)
    if n_MC is None:  # src/option.cpp:414
        n_MC = max(min(int(4e9 / n / n), 30), 3)
    rvec = np.ones(n, np.float32) if rvec is None else np.asarray(rvec, np.float32)
    rr = (1.0 / np.sqrt(rvec.astype(np.float64))).astype(np.float32)
    X = np.asarray(X, np.float64)
    X = X[:, None] if X.ndim == 1 else X
    c = X.shape[1]
    Q, _ = covariate_qr(X, rr)
    gw = np.asarray(gw, np.float32)
    mdl = _engine(kmat, engine)(gm, gw, rr, Q, subset_size)
    u, e = oracle.lmc_samples(rng_seed, n, m, n_MC)
    u = (u * np.sqrt(gw)[:, None]).astype(np.float32)
    V = np.zeros((n, 1 + 2 * n_MC), np.float32)
    V[:, 0] = y
    V[:, 1:1 + n_MC] = mdl.matvec(u)
    V[:, 1 + n_MC:] = e
    V[:, :1 + n_MC] *= rr[:, None]
    V = (V - Q @ (Q.T @ V)).astype(np.float32)
    tau0 = f32((1 - f32(s2max)) / f32(s2max))
    seed = mdl.L_Seed("quadform", float(tau0), V, 10 * tol_L)
    evals = [0]
    state = {}

    def first_order(ss):
        evals[0] += 1
        tau = f32((1 - f32(ss)) / f32(ss))
        sigma = f32(tau - tau0)
        PV = seed.L_Solve(V, float(sigma))
        Py = np.empty((n, 1 + n_MC), np.float32)
        Py[:, 0] = PV[:, 0]
        Py[:, 1:] = PV[:, 1:1 + n_MC] + np.sqrt(tau) * PV[:, 1 + n_MC:]
        Py = (Py - Q @ (Q.T @ Py)).astype(np.float32)
        eB = Py.astype(np.float64)
        Py = (Py * rr[:, None]).astype(np.float32)
        uB = mdl.rmatvec(Py).astype(np.float64) * np.sqrt(gw.astype(np.float64))[:, None]
        state["Py"] = Py
        return math.log((np.sum(uB[:, 0] ** 2) / np.sum(eB[:, 0] ** 2)) /
                        (np.sum(uB[:, 1:] ** 2) / np.sum(eB[:, 1:] ** 2)))

    s2 = f32(s2max / 2)
    lt1 = math.log((1 - s2) / s2)
    f1 = first_order(s2)
    s2 = f32(s2max / 4) if f1 < 0 else f32(s2max)
    lt2 = math.log((1 - s2) / s2)
    f2 = first_order(s2)
    for _ in range(3, 10):
        lt = (lt1 * f2 - lt2 * f1) / (f2 - f1)
        lt = min(max(lt, -9.0), 9.0)
        if abs(lt - lt2) < 0.002:
            break
        s2 = f32(1 / (math.exp(lt) + 1))
        f1 = f2
        f2 = first_order(s2)
        lt1, lt2 = lt2, lt
    Py = state["Py"][:, 0].copy()
    vg = vg_from_py(y, Py, c)
    tau1 = float((1 - s2) / s2)
    ve = tau1 * vg
    snp_blup = mdl.rmatvec(Py, weighted=True)[:, 0]
    resid = np.asarray(y, np.float64) - (mdl.matvec(snp_blup)[:, 0] + tau1 * rvec * Py)
    blue = np.linalg.solve(X.T @ X, X.T @ resid)
    return dict(vg=vg, ve=ve, h2=float(s2), Py=Py, snp_blup=snp_blup, blue=blue, steps=seed.steps,
                evals=evals[0], n_MC=n_MC)


def iter_weighting_reml(kmat, gm, rvec, X, y, maf, persnp=None, window_size=20, sig_chisq=1e6, indep_r2=0.5,
                        lrt=False, rng_seed=0, n_V=30, n_MC=None, tol_L=5e-4, s2max=0.7, fit=None):
    """--iter_weighting (src/option.cpp:417-528) over the restated REML drivers: equal-weight fit,
    2 maf (1-maf) b^2 normalised, clumping by the largest variance above sig_chisq with the
    indep_r2 residual rule, window-mean logistic weights, refit with the clumped SNPs as
    covariates.  Returns the final fit's dict plus persnp and sig_snps.  fit(sldf, persnp, X)
    replaces the FP32 restated drivers (dense.iter_weighting_reml passes FP64 fits)."""
    n = len(y)
    m = len(gm)
    Z = (np.asarray(kmat)[:, 0::2].astype(np.float32) + np.asarray(kmat)[:, 1::2]).T  # n x m dosages
    base = np.ones(m, np.float32) if persnp is None else np.asarray(persnp, np.float32)

    def fit32(sldf, pw, Xf):
        gw = oracle.gw(gm, pw)
        if sldf:
            return SLDF_REML(kmat, gm, gw, rvec, Xf, y, rng_seed=rng_seed, n_V=n_V, tol_L=tol_L, s2max=s2max)
        return LMC_REML(kmat, gm, gw, rvec, Xf, y, rng_seed=rng_seed, n_MC=n_MC, tol_L=tol_L, s2max=s2max)

    fit = fit or fit32

    if m < window_size + 1:
        pw = base / np.float32(base.mean())
        r = fit(lrt, pw, X)
        return dict(r, persnp=pw, sig_snps=np.zeros(0, np.int64))
    r0 = fit(lrt and n < 5000, np.ones(m, np.float32), X)
    b = np.asarray(r0["snp_blup"], np.float32)
    mf = np.asarray(maf, np.float32)
    snpvar = np.float32(2.0) * mf * (np.float32(1.0) - mf) * b * b
    snpvar = snpvar / np.float32(snpvar.astype(np.float64).mean())
    tmp = snpvar.copy()
    sig, cols = [], []
    while True:
        j = int(np.argmax(tmp))
        if not tmp[j] > sig_chisq:
            break
        tmp[j] = 0
        z = Z[:, j].astype(np.float64)
        yc = z - float(np.float32(z.sum() / n))
        if not sig:
            take = True
        else:
            S = np.stack(cols, axis=1)
            res = yc - S @ np.linalg.lstsq(S, yc, rcond=None)[0]
            take = (res @ res) / (yc @ yc) > indep_r2
        if take:
            sig.append(j)
            cols.append(yc)
    w, half = window_size, window_size // 2
    pre = np.concatenate([[0.0], np.cumsum(snpvar.astype(np.float64))])
    for i in range(m):
        a = 0 if i < half else (m - w - 1 if i > m - half - 1 else i - half)
        v = np.float32((pre[a + w + 1] - pre[a]) / (w + 1))
        tmp[i] = np.float32(0.001) / (np.float32(0.001) + np.float32(0.999) * np.exp(v * np.float32(-0.5)))
    tmp[sig] = 0
    pw = base * tmp
    pw = pw / np.float32(pw.astype(np.float64).mean())
    Xn = np.concatenate([np.asarray(X, np.float64).reshape(n, -1)] + [c[:, None] for c in cols], axis=1)
    r = fit(lrt, pw, Xn)
    blup = np.array(r["snp_blup"], np.float32)
    c = np.asarray(X).reshape(n, -1).shape[1]
    for a, j in enumerate(sig):
        blup[j] = r["blue"][c + a]
    return dict(r, snp_blup=blup, blue=np.asarray(r["blue"])[:c], persnp=pw, sig_snps=np.array(sig, np.int64))
