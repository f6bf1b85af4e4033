"""GPU Krylov engine (L_Seed / L_Solve / SLQ) against exact dense FP64 results and the FP32
CPU restatement of src/reml.cpp:246-451."""
import numpy as np
import pytest
import torch

import oracle
from tests.problems import make_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(gpu):
    p = make_problem(n=1200, m=1500, seed=23, weights=True, c=2)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec, p.Q)
    return p, g


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_lseed_lsolve_against_dense_solve(prob, prec):
    p, g = prob
    D = p.dense()
    tau0 = (1 - 0.7) / 0.7
    rng = np.random.default_rng(1)
    y = rng.normal(size=p.n).astype(np.float32)
    yp = (y - p.Q @ (p.Q.T @ y)).astype(np.float32)
    B = np.concatenate([yp[:, None], p.Q], axis=1)
    kinds = ["quadform"] + ["ldet"] * p.Q.shape[1]
    groups = [0] + [1] * p.Q.shape[1]
    s = g.L_Seed("ldet", tau0, torch.from_numpy(B).cuda(), tol=1e-7, maxit=400, keep_cols=B.shape[1],
                 precision=prec, col_kind=kinds, col_group=groups)
    for h2 in (0.1, 0.3, 0.6):
        tau = (1 - h2) / h2
        X = s.L_Solve(torch.from_numpy(B[:, :1]).cuda(), tau - tau0, tol=1e-6, col0=0).cpu().numpy()
        ref = D.solve("quadform", tau, B[:, :1])
        assert rel_err(X, ref) < (1e-4 if prec == "fp32" else 5e-4)
        XQ = s.L_Solve(torch.from_numpy(p.Q).cuda(), tau - tau0, tol=1e-6, col0=1).cpu().numpy()
        refQ = D.solve("ldet", tau, p.Q)
        assert rel_err(XQ, refQ) < (1e-4 if prec == "fp32" else 5e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_converged_groups_leave_the_block(prob, prec, monkeypatch):
    """[y, E, 30 probes] in three groups that stop at different steps (E: two columns in the span
    of three eigenvectors of the operator, done in a few steps): once a group has converged the
    seed runs the operator on the remaining columns only (33 -> 31 columns cuts the int8 digit
    tiles from 7 to 6 and the FP32 operand from 48 to 32 columns).  Step counts, the tridiagonal
    coefficients and the solves agree with the seed that keeps every column in the block
    (SLEMM_LSEED_COMPACT=0), and the solves with the dense FP64 ones."""
    p, g = prob
    D = p.dense()
    tau0 = (1 - 0.7) / 0.7
    rng = np.random.default_rng(4)
    y = rng.normal(size=p.n).astype(np.float32)
    yp = (y - p.Q @ (p.Q.T @ y)).astype(np.float32)
    w, U = D.eig("ldet")
    c = 2
    E = U[:, -3:] @ rng.normal(size=(3, c))
    E /= np.linalg.norm(E, axis=0)
    nv = 30
    V = oracle.rademacher(3, p.n, nv)
    B = np.ascontiguousarray(np.concatenate([yp[:, None], E, V], axis=1).astype(np.float32))
    kinds = ["quadform"] + ["ldet"] * (c + nv)
    groups = [0] + [1] * c + [2] * nv
    seeds = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("SLEMM_LSEED_COMPACT", mode)
        seeds[mode] = g.L_Seed("ldet", tau0, torch.from_numpy(B).cuda(), tol=1e-6, maxit=400,
                               keep_cols=1 + c, precision=prec, col_kind=kinds, col_group=groups)
    a, b = seeds["1"], seeds["0"]
    cs = a.col_steps
    assert int(cs[1]) < min(int(cs[0]), int(cs[1 + c])), cs  # E stops first
    assert np.abs(a.col_steps.astype(int) - b.col_steps.astype(int)).max() <= 1
    ba, da = a.coeffs()
    bb, db = b.coeffs()
    k = min(5, int(cs.min()))
    assert rel_err(da[:, :k], db[:, :k]) < 1e-5
    assert rel_err(ba[:, :k + 1], bb[:, :k + 1]) < 1e-5
    tol = 1e-4 if prec == "fp32" else 5e-4
    for h2 in (0.2, 0.5):
        tau = (1 - h2) / h2
        ref = D.solve("quadform", tau, B[:, :1])
        for s in (a, b):
            X = s.L_Solve(torch.from_numpy(B[:, :1]).cuda(), tau - tau0, tol=1e-6, col0=0).cpu().numpy()
            assert rel_err(X, ref) < tol
        la = a.SLQ_LDet(tau - tau0, col0=1 + c, nv=nv)
        lb = b.SLQ_LDet(tau - tau0, col0=1 + c, nv=nv)
        assert abs(la - lb) <= 5e-5 * abs(lb), (la, lb)


def test_lseed_matches_cpu_restatement(prob):
    """Same inputs, same stopping rule: step counts agree and the tridiagonal coefficients match
    the FP32 CPU restatement to FP32 rounding (the first steps, before rounding compounds)."""
    p, g = prob
    tau0 = (1 - 0.7) / 0.7
    V = oracle.rademacher(0, p.n, 8)
    mdl = p.model()
    ref = mdl.L_Seed("ldet", tau0, V, tol=5e-4)
    s = g.L_Seed("ldet", tau0, torch.from_numpy(np.ascontiguousarray(V)).cuda(), tol=5e-4, keep_cols=0,
                 precision="fp32")
    assert abs(s.steps - ref.steps) <= 1
    beta, delta = s.coeffs()
    k = min(5, ref.steps)
    assert rel_err(delta[:, :k], ref.delta[:, :k]) < 1e-4
    assert rel_err(beta[:, :k + 1], ref.beta[:, :k + 1]) < 1e-4


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_slq_against_exact_quadrature(prob, prec):
    """For fixed probes, SLQ with converged Lanczos equals sum_k v_k' log(A + sigma I) v_k."""
    p, g = prob
    D = p.dense()
    tau0 = (1 - 0.7) / 0.7
    nv = 6
    V = oracle.rademacher(3, p.n, nv)
    s = g.L_Seed("ldet", tau0, torch.from_numpy(np.ascontiguousarray(V)).cuda(), tol=1e-6, keep_cols=0,
                 precision=prec, maxit=500)
    for h2 in (0.2, 0.5):
        tau = (1 - h2) / h2
        est = s.SLQ_LDet(tau - tau0)
        exact = D.quad_logm("ldet", tau, V) * p.n / nv
        assert abs(est - exact) / abs(exact) < (1e-6 if prec == "fp32" else 2e-5)


def test_seed_outlives_its_genotype_handle(gpu):
    """A seed keeps its own device and basis: L_Solve still works, and destroying it leaves no
    pending HIP error behind, after the genotype handle it was built on is gone; a new handle is
    built and used normally afterwards.  (The seed once read its device through the handle at
    destroy time, which read freed memory when the handle went first.)"""
    from slemm_amd import _lib
    p = make_problem(n=700, m=900, seed=4)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model()
    tau0 = (1 - 0.7) / 0.7
    rng = np.random.default_rng(2)
    b = torch.from_numpy(rng.normal(size=(p.n, 1)).astype(np.float32)).cuda()
    s = g.L_Seed("ldet", tau0, b, tol=1e-6)
    x0 = s.L_Solve(b, 0.1).cpu().numpy()
    s._owner = None          # drop the keep-alive reference: the handle goes first
    g.close()
    x1 = s.L_Solve(b, 0.1).cpu().numpy()
    assert np.array_equal(x0, x1)
    lib = _lib.load()
    _lib.check(lib.slemm_gpu_lseed_destroy(s._h))
    s._h = None
    g2 = gpu.Geno.from_bed(p.rows, p.n)
    g2.set_model()
    out = g2.lnr_op_select("ldet", tau0, b, precision="bf16x2").cpu().numpy()
    ref = p.dense().op("ldet", tau0, b.cpu().numpy())
    assert rel_err(out, ref) < 3e-4
