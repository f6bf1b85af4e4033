"""Batched traits on individual subsets: every trait's model terms, operator and Lanczos solves
must equal what a handle holding only that trait's individuals computes (the reference's own
one-trait-per-run setting, src/input.cpp:389-454), and the FP64 dense operator on the subset."""
import numpy as np
import pytest
import torch

from oracle.dense import Dense
from tests.problems import make_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(gpu):
    p = make_problem(n=2000, m=1500, seed=41, missing=0.01)
    rng = np.random.default_rng(5)
    sizes = [700, 1200, 1900]
    masks = np.zeros((3, p.n), bool)
    for t, s in enumerate(sizes):
        idx = rng.choice(np.arange(2, p.n), size=s - 2, replace=False)
        masks[t, idx] = True
        masks[t, :2] = True  # individuals 0/1 make every SNP polymorphic in every subset
    g = gpu.Geno.from_bed(p.rows, p.n)
    tr = gpu.Traits(g, masks)
    subs = []
    for t in range(3):
        keep = np.flatnonzero(masks[t])
        gs = gpu.Geno.from_bed(p.rows, p.n, keep_idx=keep)
        gs.set_model(None, None, np.full((len(keep), 1), 1 / np.sqrt(len(keep)), np.float32))
        subs.append((keep, gs))
    return p, masks, g, tr, subs


def test_trait_model_terms_bit_exact(batch):
    p, masks, g, tr, subs = batch
    nt, mono = tr.info()
    assert list(nt) == [int(m.sum()) for m in masks] and list(mono) == [0, 0, 0]
    for t, (keep, gs) in enumerate(subs):
        gm, gw, pc = tr.model(t)
        pc1, _, _, _, gm1 = gs.counts()
        _, gw1, _ = gs.model()
        assert np.array_equal(pc, pc1) and np.array_equal(gm, gm1) and np.array_equal(gw, gw1)


def test_batched_operator_equals_subset_handles(batch):
    p, masks, g, tr, subs = batch
    rng = np.random.default_rng(9)
    per = 5
    col_trait = np.repeat(np.arange(3), per)
    kinds = ["quadform", "ldet", "ldet", "quadform", "ldet"] * 3
    V = np.zeros((p.n, 3 * per), np.float32)
    for k, t in enumerate(col_trait):
        V[masks[t], k] = rng.normal(size=masks[t].sum())
    tau0 = 0.37
    out = tr.lnr_op_select(col_trait, kinds, tau0, torch.from_numpy(V).cuda()).cpu().numpy()
    for t, (keep, gs) in enumerate(subs):
        cols = np.flatnonzero(col_trait == t)
        Vt = torch.from_numpy(np.ascontiguousarray(V[np.ix_(keep, cols)])).cuda()
        ref = gs.lnr_op_select("ldet", tau0, Vt, col_kind=[kinds[k] for k in cols]).cpu().numpy()
        got = out[np.ix_(keep, cols)]
        assert rel_err(got, ref) < 1e-5, t
        outside = np.flatnonzero(~masks[t])
        assert np.all(out[np.ix_(outside, cols)] == 0.0)
        # FP64 dense operator of the subset
        Zs = p.Z[keep]
        gm, gw, _ = tr.model(t)
        Q = np.full((len(keep), 1), 1 / np.sqrt(len(keep)))
        D = Dense(Zs, gm, gw, np.ones(len(keep)), Q)
        for j, k in enumerate(cols):
            want = D.op(kinds[k], tau0, V[keep, k][:, None])[:, 0]
            assert rel_err(got[:, j], want) < 3e-4, (t, k)


def test_batched_seed_solves(batch):
    """One batched seed over [y, probes] of every trait; L_Solve at another shift vs dense."""
    p, masks, g, tr, subs = batch
    rng = np.random.default_rng(2)
    col_trait, kinds, groups, cols = [], [], [], []
    for t in range(3):
        for j in range(3):
            v = np.zeros(p.n, np.float32)
            v[masks[t]] = rng.normal(size=masks[t].sum())
            cols.append(v)
            col_trait.append(t)
            kinds.append("quadform" if j == 0 else "ldet")
            groups.append(2 * t + (0 if j == 0 else 1))
    B = torch.from_numpy(np.stack(cols, 1)).cuda()
    tau0, sigma = 0.5, 0.8
    seed = tr.L_Seed(col_trait, kinds, tau0, B, tol=1e-6, col_group=groups)
    X = seed.L_Solve(B, sigma, tol=1e-7).cpu().numpy()
    Bn = B.cpu().numpy()
    for t, (keep, gs) in enumerate(subs):
        gm, gw, _ = tr.model(t)
        Q = np.full((len(keep), 1), 1 / np.sqrt(len(keep)))
        D = Dense(p.Z[keep], gm, gw, np.ones(len(keep)), Q)
        for k in np.flatnonzero(np.array(col_trait) == t):
            want = D.solve(kinds[k], tau0 + sigma, Bn[keep, k][:, None])[:, 0]
            assert rel_err(X[keep, k], want) < 2e-4, (t, k)
            assert np.all(X[~masks[t], k] == 0.0)
