"""Derivatives of the REML log-likelihood on the GPU: the adjoint kernel behind the autograd
harness against central finite differences of the FP64 dense likelihood, the device score /
average-information matrix against their FP64 closed forms, and AI-REML against the FP64
iteration on the same probes."""
import numpy as np
import pytest
import torch

import oracle
from oracle import vc
from tests.problems import make_phenotype, make_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gpu):
    """n=600 with identity probes (exact traces), weights R, two covariates, 9 windows of 100."""
    p = make_problem(n=600, m=900, seed=21, weights=True, c=2, missing=0.0)
    y = make_phenotype(p, 0.5, n_causal=100, seed=4)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec)
    return p, y, g


def test_adjoint_kernel_against_dense(small):
    """out_j = sum_k coef_k (Zc'R^-1/2 a_k)_j (Zc'R^-1/2 b_k)_j, FP64 dense reference."""
    p, y, g = small
    rng = np.random.default_rng(0)
    A = rng.normal(size=(p.n, 5)).astype(np.float32)
    B = rng.normal(size=(p.n, 5)).astype(np.float32)
    coef = rng.normal(size=5)
    Zc = p.Z - p.gm.astype(np.float64)
    za = Zc.T @ (p.rr[:, None].astype(np.float64) * A)
    zb = Zc.T @ (p.rr[:, None].astype(np.float64) * B)
    want = (za * zb) @ coef
    for prec, tol in (("fp32", 2e-5), ("bf16x2", 1e-4)):
        got = g.adjoint(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), coef, precision=prec).cpu().numpy()
        assert rel_err(got, want) < tol, prec


def _harness(small, w_np, sg2, s2e, probes, tol=2e-6):
    from slemm_amd.autodiff import RemlProblem, loglik
    p, y, g = small
    prob = RemlProblem(g, p.X, y, rvec=p.rvec, probes=probes, tol=tol)
    win = torch.arange(p.m, device="cuda") // 100
    w = torch.tensor(w_np, dtype=torch.float64, device="cuda", requires_grad=True)
    a = torch.tensor(sg2, dtype=torch.float64, device="cuda", requires_grad=True)
    b = torch.tensor(s2e, dtype=torch.float64, device="cuda", requires_grad=True)
    ll = loglik(prob, w, a, b, win)
    ll.backward()
    return prob, ll.item(), w.grad.cpu().numpy(), float(a.grad), float(b.grad)


def test_autograd_against_finite_differences(small):
    """dl/dw (9 window weights), dl/dsigma_g^2, dl/dsigma_e^2 from the adjoint vs central
    differences (relative step 1e-4) of the exact FP64 likelihood."""
    p, y, g = small
    rng = np.random.default_rng(5)
    w0 = rng.uniform(0.5, 1.5, 9)
    sg2, s2e = 0.6, 0.5
    prob, ll, gw_, gs, ge = _harness(small, w0, sg2, s2e, torch.eye(p.n, device="cuda"))
    base = prob.base_gw.astype(np.float64)
    win = np.arange(p.m) // 100
    Zc = p.Z - p.gm.astype(np.float64)

    def f(w, a, b):
        return vc.reml_loglik(Zc, a * vc.window_weights_to_gw(base, win, w), b, p.rvec, p.X, y)

    ref = f(w0, sg2, s2e)
    assert abs(ll - ref) < 1e-4 * abs(ref), (ll, ref)
    fd_w = np.zeros(9)
    for k in range(9):
        h = 1e-4 * w0[k]
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        fd_w[k] = (f(wp, sg2, s2e) - f(wm, sg2, s2e)) / (2 * h)
    fd_a = (f(w0, sg2 * (1 + 1e-4), s2e) - f(w0, sg2 * (1 - 1e-4), s2e)) / (2e-4 * sg2)
    fd_b = (f(w0, sg2, s2e * (1 + 1e-4)) - f(w0, sg2, s2e * (1 - 1e-4))) / (2e-4 * s2e)
    msg = f"w {gw_} vs {fd_w}; sg2 {gs} vs {fd_a}; s2e {ge} vs {fd_b}"
    assert np.max(np.abs(gw_ - fd_w)) < 2e-3 * np.max(np.abs(fd_w)), msg
    assert abs(gs - fd_a) < 2e-3 * abs(fd_a), msg
    assert abs(ge - fd_b) < 2e-3 * abs(fd_b), msg


def test_device_score_and_ai_exact_traces(small, gpu):
    """slemm_gpu_reml_score with identity probes against the FP64 closed forms; the harness's
    dl/dsigma^2 at unit window weights is the same score."""
    p, y, g = small
    Zc = p.Z - p.gm.astype(np.float64)
    g.set_model(None, p.rvec)
    _, gw, _ = g.model()
    vg, ve = 0.9, 0.45
    eye = torch.eye(p.n, device="cuda")
    score, AI = g.reml_score(p.X, y, vg, ve, probes=eye, tol_L=2e-6)
    s_ref, AI_ref, terms = vc.hutch_two_component(Zc, gw, vg, ve, p.rvec, p.X, y, None)
    scale = np.array([terms["trPA"] + terms["yPAPy"], terms["trP"] + terms["yPPy"]])
    assert np.all(np.abs(score - s_ref) < 2e-4 * scale), (score, s_ref)
    assert rel_err(AI, AI_ref) < 1e-3, (AI, AI_ref)
    _, _, _, gs, ge = _harness(small, np.ones(9), vg, ve, eye)
    assert np.all(np.abs(np.array([gs, ge]) - s_ref) < 2e-4 * scale), (gs, ge, s_ref)


def test_harness_matches_device_score_stochastic(gpu):
    """64 Rademacher probes at n=4000: the harness (adjoint) and slemm_gpu_reml_score evaluate the
    same trace estimator on the same probes."""
    from slemm_amd.autodiff import RemlProblem, loglik
    p = make_problem(n=4000, m=5000, seed=31, c=1, missing=0.0)
    y = make_phenotype(p, 0.4, n_causal=300, seed=9)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model()
    gen = torch.Generator(device="cuda").manual_seed(3)
    V = torch.where(torch.rand((p.n, 64), generator=gen, device="cuda") < 0.5, -1.0, 1.0) / np.sqrt(p.n)
    vy = float(np.var(y))
    vg, ve = 0.5 * vy, 0.6 * vy
    prob = RemlProblem(g, p.X, y, probes=V, tol=1e-6)
    win = torch.arange(p.m, device="cuda") // 100
    w = torch.ones(p.m // 100, dtype=torch.float64, device="cuda", requires_grad=True)
    a = torch.tensor(vg, dtype=torch.float64, device="cuda", requires_grad=True)
    b = torch.tensor(ve, dtype=torch.float64, device="cuda", requires_grad=True)
    loglik(prob, w, a, b, win).backward()
    g.set_model()  # the harness left sigma_g^2-scaled weights in the handle
    score, _ = g.reml_score(p.X, y, vg, ve, probes=V, tol_L=1e-6)
    Zc = p.Z - p.gm.astype(np.float64)
    _, gw, _ = g.model()
    s_ref, _, terms = vc.hutch_two_component(Zc, gw, vg, ve, np.ones(p.n), p.X, y, V.double().cpu().numpy())
    scale = np.array([terms["trPA"] + terms["yPAPy"], terms["trP"] + terms["yPPy"]])
    got = np.array([float(a.grad), float(b.grad)])
    assert np.all(np.abs(got - s_ref) < 3e-4 * scale), (got, s_ref)
    assert np.all(np.abs(score - s_ref) < 3e-4 * scale), (score, s_ref)
    # window gradients sum to zero against the weights (the mean normalisation is scale-free)
    assert abs(float(w.grad @ w.detach())) < 1e-6 * float(torch.abs(w.grad).sum()) + 1e-9


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_ai_reml_against_fp64_iteration(gpu, prec):
    """AI-REML on the device vs the FP64 dense AI iteration on the same 30 probes."""
    p = make_problem(n=1500, m=2000, seed=3)
    y = make_phenotype(p, 0.5)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model()
    got = g.AI_REML(p.X, y, n_V=30, tol_L=1e-5, tol_VC=1e-6, max_iter=30, precision=prec)
    _, gw, _ = g.model()
    Zc = p.Z - p.gm.astype(np.float64)
    V = oracle.rademacher(0, p.n, 30)
    vg, ve, _ = vc.hutch_ai_reml(Zc, gw, np.ones(p.n), p.X, y, V, got["history"][0, 0], got["history"][0, 1])
    h2 = vg / (vg + ve)
    d = {"h2": abs(got["h2"] - h2) / h2, "vg": abs(got["vg"] - vg) / vg, "ve": abs(got["ve"] - ve) / ve}
    msg = f"gpu {got['h2']:.7f} {got['vg']:.6f} {got['ve']:.6f} fp64 {h2:.7f} {vg:.6f} {ve:.6f} rel {d}"
    print(msg)
    assert got["iterations"] < 30, msg
    assert max(d.values()) < 1e-4, msg
    last = got["history"][-1]
    assert abs(last[2]) < 1e-2 * abs(last[4]) * vg and abs(last[3]) < 1e-2 * abs(last[6]) * ve
