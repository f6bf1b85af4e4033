"""Pins of the FP64 variance-component oracle (oracle/vc.py) by finite differences, the
observed/expected-information identity and the profile-likelihood optimum.  CPU only."""
import numpy as np
import pytest
from scipy.optimize import minimize_scalar

from oracle import dense, vc


def _small(seed=0, n=150, m=80, c=2):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, m)
    Z = rng.binomial(2, p, size=(n, m)).astype(np.float64)
    gm = Z.mean(0)
    Zc = Z - gm
    gw = 1.0 / (gm * (2 - gm) / 2) / m
    rvec = rng.uniform(0.5, 2.0, n)
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(c - 1)])
    beta = rng.normal(size=m) * np.sqrt(0.6 * gw)
    y = X @ rng.normal(size=c) + Zc @ beta + rng.normal(size=n) * np.sqrt(0.4 * rvec)
    return Z, gm, Zc, gw, rvec, X, y


def test_gradient_matches_central_differences():
    Z, gm, Zc, gw, rvec, X, y = _small()
    d = 0.7 * gw * (1 + 0.3 * np.random.default_rng(1).random(gw.size))
    s2e = 0.45
    gd, ge = vc.reml_grad(Zc, d, s2e, rvec, X, y)
    for j in [0, 7, 33, 79]:
        h = 1e-4 * d[j]
        dp, dm = d.copy(), d.copy()
        dp[j] += h
        dm[j] -= h
        fd = (vc.reml_loglik(Zc, dp, s2e, rvec, X, y) - vc.reml_loglik(Zc, dm, s2e, rvec, X, y)) / (2 * h)
        assert abs(fd - gd[j]) <= 1e-5 * max(1.0, abs(gd[j]))
    h = 1e-4 * s2e
    fd = (vc.reml_loglik(Zc, d, s2e + h, rvec, X, y) - vc.reml_loglik(Zc, d, s2e - h, rvec, X, y)) / (2 * h)
    assert abs(fd - ge) <= 1e-6 * max(1.0, abs(ge))


def test_two_component_score_is_chain_of_per_snp_gradient():
    Z, gm, Zc, gw, rvec, X, y = _small(seed=2)
    vg, ve = 0.8, 0.5
    score, AI, EI = vc.two_component(Zc, gw, vg, ve, rvec, X, y)
    gd, ge = vc.reml_grad(Zc, vg * gw, ve, rvec, X, y)
    assert np.isclose(score[0], np.sum(gw * gd), rtol=1e-10, atol=1e-12)
    assert np.isclose(score[1], ge, rtol=1e-10, atol=1e-12)


def test_ai_is_mean_of_observed_and_expected_information():
    """For V linear in theta: -d2l = 2 AI - EI, so AI = (I_obs + EI)/2 with I_obs from central
    differences of the score."""
    Z, gm, Zc, gw, rvec, X, y = _small(seed=3)
    th = np.array([0.9, 0.6])
    _, AI, EI = vc.two_component(Zc, gw, th[0], th[1], rvec, X, y)
    Iobs = np.zeros((2, 2))
    for k in range(2):
        h = 1e-5 * th[k]
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        sp, _, _ = vc.two_component(Zc, gw, tp[0], tp[1], rvec, X, y)
        sm, _, _ = vc.two_component(Zc, gw, tm[0], tm[1], rvec, X, y)
        Iobs[:, k] = -(sp - sm) / (2 * h)
    assert np.allclose(0.5 * (Iobs + EI), AI, rtol=1e-6, atol=1e-9)


def test_ai_reml_reaches_profile_likelihood_optimum():
    Z, gm, Zc, gw, rvec, X, y = _small(seed=4, n=200, m=120)
    vg, ve, it = vc.ai_reml(Zc, gw, rvec, X, y, 1.0, 1.0)
    assert it < 50
    rr = 1 / np.sqrt(rvec)
    res = minimize_scalar(lambda h2: -dense.reml_loglik(Z, gm, gw, X, y, h2, rr=rr), bounds=(1e-3, 0.999),
                          method="bounded", options={"xatol": 1e-10})
    assert abs(vg / (vg + ve) - res.x) < 1e-6
    score, _, _ = vc.two_component(Zc, gw, vg, ve, rvec, X, y)
    assert np.all(np.abs(score) < 1e-8 * (1 + np.abs(score).max()))


def test_window_weights_unit_is_identity():
    gw = np.random.default_rng(5).random(250)
    win = np.arange(250) // 100
    assert np.array_equal(vc.window_weights_to_gw(gw, win, np.ones(3)), gw)
    w = np.array([1.0, 2.0, 4.0])
    got = vc.window_weights_to_gw(gw, win, w)
    persnp = w[win]
    assert np.allclose(got, gw * persnp / persnp.mean())


def test_hutch_score_with_identity_probes_is_exact():
    Z, gm, Zc, gw, rvec, X, y = _small(seed=6)
    n = Zc.shape[0]
    s0, AI0, _ = vc.two_component(Zc, gw, 0.7, 0.6, rvec, X, y)
    s1, AI1, t1 = vc.hutch_two_component(Zc, gw, 0.7, 0.6, rvec, X, y, None)
    s2, AI2, t2 = vc.hutch_two_component(Zc, gw, 0.7, 0.6, rvec, X, y, np.eye(n))
    assert np.allclose(s1, s0, rtol=1e-9, atol=1e-10) and np.allclose(AI1, AI0, rtol=1e-9)
    assert np.allclose(s2, s0, rtol=1e-9, atol=1e-10) and np.allclose(AI2, AI0, rtol=1e-9)
    vg, ve, _ = vc.ai_reml(Zc, gw, rvec, X, y, 1.0, 1.0)
    vg2, ve2, _ = vc.hutch_ai_reml(Zc, gw, rvec, X, y, np.eye(n), 1.0, 1.0)
    assert abs(vg2 - vg) < 1e-8 * vg and abs(ve2 - ve) < 1e-8 * ve
