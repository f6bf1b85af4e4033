"""ORACLE -- test infrastructure only.  FP64 dense REML log-likelihood of the variance-component
model and its derivatives, for checking the device AI-REML and the adjoint (derivative) path.

Model (that of SLDF_REML, src/reml.cpp:549-690, with free per-SNP variances):
    V = Zc diag(d) Zc' + s2e R,     Zc = Z - 1 gm',   R = diag(rvec)
    l(d, s2e) = -1/2 [ log|V| + log|X'V^-1X| + y'Py ],   P = V^-1 - V^-1 X (X'V^-1 X)^-1 X'V^-1
d_j = sigma_g^2 gw_j for the usual two-component model.  Closed forms (no reference counterpart,
the reference has no derivatives):
    dl/dd_j  = -1/2 [ zc_j' P zc_j - (zc_j' P y)^2 ]
    dl/ds2e  = -1/2 [ tr(P R) - y'P R P y ]
    AI_kl    =  1/2 y'P V_k P V_l P y                      (V_g = Zc W Zc', V_e = R)
"""
from __future__ import annotations

import numpy as np


def _factor(Zc, d, s2e, rvec, X):
    Zc = np.asarray(Zc, np.float64)
    n = Zc.shape[0]
    V = (Zc * np.asarray(d, np.float64)[None, :]) @ Zc.T + s2e * np.diag(np.asarray(rvec, np.float64))
    L = np.linalg.cholesky(V)
    X = np.asarray(X, np.float64).reshape(n, -1)
    Vi = np.linalg.inv(V)
    Vi = 0.5 * (Vi + Vi.T)
    XtViX = X.T @ Vi @ X
    P = Vi - Vi @ X @ np.linalg.solve(XtViX, X.T @ Vi)
    P = 0.5 * (P + P.T)
    return L, XtViX, P


def reml_loglik(Zc, d, s2e, rvec, X, y) -> float:
    """l(d, s2e) up to the constant -(n-c)/2 log(2 pi)."""
    L, XtViX, P = _factor(Zc, d, s2e, rvec, X)
    y = np.asarray(y, np.float64)
    return float(-0.5 * (2.0 * np.sum(np.log(np.diag(L))) + np.linalg.slogdet(XtViX)[1] + y @ P @ y))


def reml_grad(Zc, d, s2e, rvec, X, y):
    """(dl/dd (m,), dl/ds2e) by the closed forms above."""
    _, _, P = _factor(Zc, d, s2e, rvec, X)
    Zc = np.asarray(Zc, np.float64)
    y = np.asarray(y, np.float64)
    r = np.asarray(rvec, np.float64)
    Py = P @ y
    zPy = Zc.T @ Py
    zPz = np.einsum("ij,ij->j", Zc, P @ Zc)
    gd = -0.5 * (zPz - zPy ** 2)
    ge = -0.5 * (np.sum(np.diag(P) * r) - Py @ (r * Py))
    return gd, float(ge)


def two_component(Zc, gw, vg, ve, rvec, X, y):
    """Score (dl/dvg, dl/dve), AI matrix and the expected information 1/2 tr(P V_k P V_l) of the
    model V = vg Zc diag(gw) Zc' + ve R."""
    Zc = np.asarray(Zc, np.float64)
    gw = np.asarray(gw, np.float64)
    _, _, P = _factor(Zc, vg * gw, ve, rvec, X)
    y = np.asarray(y, np.float64)
    Vg = (Zc * gw[None, :]) @ Zc.T
    Ve = np.diag(np.asarray(rvec, np.float64))
    Py = P @ y
    Vs = (Vg, Ve)
    score = np.array([-0.5 * (np.sum(P * Vk) - Py @ Vk @ Py) for Vk in Vs])
    AI = np.array([[0.5 * (Py @ Vk @ P @ Vl @ Py) for Vl in Vs] for Vk in Vs])
    EI = np.array([[0.5 * np.sum((P @ Vk) * (P @ Vl).T) for Vl in Vs] for Vk in Vs])
    return score, AI, EI


def ai_reml(Zc, gw, rvec, X, y, vg0, ve0, tol=1e-10, max_iter=100):
    """Exact AI-REML in FP64 (Newton with the AI matrix, step halving for positivity).
    Returns (vg, ve, iterations)."""
    vg, ve = float(vg0), float(ve0)
    for it in range(1, max_iter + 1):
        s, AI, _ = two_component(Zc, gw, vg, ve, rvec, X, y)
        step = np.linalg.solve(AI, s)
        lam = 1.0
        while not (vg + lam * step[0] > 0 and ve + lam * step[1] > 0):
            lam *= 0.5
        vg1, ve1 = vg + lam * step[0], ve + lam * step[1]
        rel = max(abs(vg1 - vg) / vg, abs(ve1 - ve) / ve)
        vg, ve = vg1, ve1
        if rel < tol:
            return vg, ve, it
    return vg, ve, max_iter


def window_weights_to_gw(base_gw, win, w):
    """Per-SNP weights from per-window weights w (persnp_j = w[win_j]), normalised by the mean of
    persnp as src/reml.cpp:622 does: gw_j = base_gw_j * persnp_j / mean(persnp)."""
    persnp = np.asarray(w, np.float64)[np.asarray(win)]
    return np.asarray(base_gw, np.float64) * persnp / persnp.mean()


def hutch_two_component(Zc, gw, vg, ve, rvec, X, y, V=None):
    """two_component in the R^-1/2-whitened model with the two traces replaced by the probe
    estimates n/nV sum_v v'P A v and n/nV sum_v v'P v (V: n x nV, E[vv'] = I/n; None = exact
    traces).  This is the score the device AI-REML evaluates.  Returns (score, AI, terms)."""
    Zc = np.asarray(Zc, np.float64)
    n = Zc.shape[0]
    r = np.asarray(rvec, np.float64)
    rr = 1.0 / np.sqrt(r)
    gw = np.asarray(gw, np.float64)
    A = (rr[:, None] * Zc * gw[None, :]) @ (Zc.T * rr[None, :])
    Xw = np.asarray(X, np.float64).reshape(n, -1) * rr[:, None]
    yw = rr * np.asarray(y, np.float64)
    Vw = vg * A + ve * np.eye(n)
    Vi = np.linalg.inv(Vw)
    P = Vi - Vi @ Xw @ np.linalg.solve(Xw.T @ Vi @ Xw, Xw.T @ Vi)
    P = 0.5 * (P + P.T)
    Py = P @ yw
    if V is None:
        trPA, trP = float(np.sum(P * A)), float(np.trace(P))
    else:
        V = np.asarray(V, np.float64)
        PV = P @ V
        w = n / V.shape[1]
        trPA = w * float(np.sum(PV * (A @ V)))
        trP = w * float(np.sum(PV * V))
    APy = A @ Py
    score = np.array([-0.5 * (trPA - Py @ APy), -0.5 * (trP - Py @ Py)])
    AI = 0.5 * np.array([[APy @ P @ APy, APy @ P @ Py], [APy @ P @ Py, Py @ P @ Py]])
    return score, AI, {"trPA": trPA, "trP": trP, "yPAPy": float(Py @ APy), "yPPy": float(Py @ Py)}


def hutch_ai_reml(Zc, gw, rvec, X, y, V, vg0, ve0, tol=1e-10, max_iter=100):
    """AI-REML iterations on hutch_two_component's score (exact AI), the FP64 counterpart of the
    device AI-REML for the same probes.  Returns (vg, ve, iterations)."""
    vg, ve = float(vg0), float(ve0)
    for it in range(1, max_iter + 1):
        s, AI, _ = hutch_two_component(Zc, gw, vg, ve, rvec, X, y, V)
        step = np.linalg.solve(AI, s)
        lam = 1.0
        while not (vg + lam * step[0] > 0 and ve + lam * step[1] > 0):
            lam *= 0.5
        vg1, ve1 = vg + lam * step[0], ve + lam * step[1]
        rel = max(abs(vg1 - vg) / vg, abs(ve1 - ve) / ve)
        vg, ve = vg1, ve1
        if rel < tol:
            return vg, ve, it
    return vg, ve, max_iter
