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


def test_device_individual_subset_equals_host_compaction(batch):
    p, masks, g, tr, subs = batch
    keep, gs = subs[0]
    gk = g.keep(keep)
    assert (gk.n, gk.m) == (len(keep), p.m)
    assert np.array_equal(gk.download(), gs.download())
    gk.set_model()
    gs2 = gs  # the transposed copy agrees too: same matvec
    U = torch.from_numpy(np.random.default_rng(1).normal(size=(p.m, 2)).astype(np.float32)).cuda()
    gs2.set_model()
    assert np.array_equal(gk.matvec(U, precision="fp32").cpu().numpy(), gs2.matvec(U, precision="fp32").cpu().numpy())
    gs.set_model(None, None, np.full((len(keep), 1), 1 / np.sqrt(len(keep)), np.float32))


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


def test_batched_seed_equals_subset_seeds(batch):
    """The batched Lanczos seed (every trait's y, intercept and probe columns in one system)
    against the seeds that subset handles run on the same columns: the first Lanczos
    coefficients agree to rounding (later ones are not comparable: without re-orthogonalisation
    FP32 Lanczos amplifies rounding once Ritz values converge), and so do what the drivers use --
    the replayed solves and the SLQ log-determinant."""
    import oracle
    p, masks, g, tr, subs = batch
    rng = np.random.default_rng(21)
    cols, col_trait, kinds, groups = [], [], [], []
    per_trait = []
    for t, (keep, gs) in enumerate(subs):
        nt = len(keep)
        y = rng.normal(size=nt).astype(np.float32)
        y -= y.mean()
        q = np.full(nt, 1 / np.sqrt(nt), np.float32)
        V = oracle.rademacher(t, nt, 4)
        block = np.column_stack([y, q, V]).astype(np.float32)
        per_trait.append(block)
        for j in range(block.shape[1]):
            v = np.zeros(p.n, np.float32)
            v[keep] = block[:, j]
            cols.append(v)
            col_trait.append(t)
            kinds.append("quadform" if j == 0 else "ldet")
            groups.append(3 * t + min(j, 2))
    B = torch.from_numpy(np.stack(cols, 1)).cuda()
    seed = tr.L_Seed(col_trait, kinds, 0.43, B, tol=5e-4, col_group=groups)
    beta, delta = seed.coeffs()
    X = seed.L_Solve(B, 0.6).cpu().numpy()
    off = 0
    for t, (keep, gs) in enumerate(subs):
        blk = per_trait[t]
        w = blk.shape[1]
        Bt = torch.from_numpy(blk).cuda()
        s1 = gs.L_Seed("ldet", 0.43, Bt, tol=5e-4, col_kind=["quadform"] + ["ldet"] * (w - 1),
                       col_group=[0, 1] + [2] * (w - 2))
        b1, d1 = s1.coeffs()
        assert np.allclose(delta[off:off + w, :4], d1[:, :4], rtol=3e-5), t
        assert np.allclose(beta[off:off + w, :5], b1[:, :5], rtol=3e-5), t
        X1 = s1.L_Solve(Bt, 0.6).cpu().numpy()
        assert rel_err(X[keep, off:off + w], X1) < 2e-5, t
        l_b = seed.SLQ_LDet(0.6, col0=off + 2, nv=w - 2) * len(keep) / p.n
        l_1 = s1.SLQ_LDet(0.6, col0=2, nv=w - 2)
        assert abs(l_b - l_1) <= 1e-5 * abs(l_1), (t, l_b, l_1)
        off += w


def test_plan_choice(batch, gpu):
    """The automatic plan: random subsets (every 16-column tile spans nearly all individuals, 1.6x
    the one-at-a-time work here) run one trait at a time; nested leading subsets batch at 30
    probes (1.2x) but not at 20 (1.76x: tiles that mix traits span the largest one)."""
    p, masks, g, tr, subs = batch
    plan, ratio = tr.plan(30)
    assert plan == "sequential" and ratio > 1.25, (plan, ratio)
    lead = np.zeros_like(masks)
    for t, s in enumerate(masks.sum(axis=1)):
        lead[t, :s] = True
    tl = gpu.Traits(g, lead)
    plan, ratio = tl.plan(30)
    assert plan == "batched" and ratio <= 1.25, (plan, ratio)
    assert tl.plan(20)[0] == "sequential"  # 20 probes: the mixed 16-column tiles cost 1.76x
    tl.set_plan("batched")
    assert tl.plan(20)[0] == "batched"
    with pytest.raises(ValueError):
        tl.set_plan("parallel")
    tl.close()


@pytest.mark.parametrize("plan", ["batched", "sequential"])
def test_batched_sldf_equals_per_subset_runs(batch, plan, gpu):
    """SLDF_REML for all traits, batched and one trait at a time, vs the single-trait driver on
    each subset handle with the same probes (same generator over the subset), and vs the FP64
    criterion.  One trait at a time runs the single-trait driver on a device-compacted handle and
    must match the host-compacted handle's run exactly.  At the reference's settings
    (tol_VC = 1e-4, tol_L = 5e-4) the search's last evaluation point on the flat top of the
    criterion differs between runs whose criteria differ in the 6th digit, by up to about 1e-3 in
    h2 at n ~ 1e3, which is the tolerance here; test_batched_vc_to_1e_4 runs both to 1e-6."""
    import oracle
    from oracle.dense import sldf_gss
    p, masks, g, tr, subs = batch
    rng = np.random.default_rng(13)
    Y = np.zeros((3, p.n), np.float32)
    for t in range(3):
        keep = np.flatnonzero(masks[t])
        idx = rng.choice(p.m, size=150, replace=False)
        gv = p.Z[np.ix_(keep, idx)] @ rng.normal(size=150)
        h2 = [0.2, 0.5, 0.7][t]
        e = rng.normal(size=len(keep)) * np.sqrt(gv.var() * (1 - h2) / h2)
        Y[t, keep] = (gv + e + 5.0).astype(np.float32)
        Y[t, ~masks[t]] = 123.0  # ignored
    tr.set_plan(plan)
    got = tr.SLDF_REML(Y, n_V=20)
    tr.set_plan("auto")
    assert got["plan"] == plan
    if plan == "sequential":  # largest subset first: the reused handle shrinks between traits
        rev = gpu.Traits(g, masks[::-1].copy())
        rev.set_plan("sequential")
        back = rev.SLDF_REML(Y[::-1].copy(), n_V=20)
        rev.close()
        assert np.array_equal(back["h2"][::-1], got["h2"]) and np.array_equal(back["vg"][::-1], got["vg"])
    for t, (keep, gs) in enumerate(subs):
        ref = gs.SLDF_REML(np.ones((len(keep), 1), np.float32), Y[t, keep], n_V=20)
        if plan == "sequential":
            assert got["h2"][t] == ref["h2"] and got["vg"][t] == ref["vg"], (t, got["h2"][t], ref["h2"])
        nt = len(keep)
        gm, gw, _ = tr.model(t)
        Q = np.full((nt, 1), 1 / np.sqrt(nt))
        yp = Y[t, keep].astype(np.float64)
        yp = (yp - yp.mean()).astype(np.float32)
        _, _, _, h2_64 = sldf_gss(Dense(p.Z[keep], gm, gw, np.ones(nt), Q), yp, Q, oracle.rademacher(0, nt, 20))
        msg = (f"trait {t}: batch {got['h2'][t]:.6f} {got['vg'][t]:.5f} single {ref['h2']:.6f} {ref['vg']:.5f} "
               f"fp64 {h2_64:.6f}")
        assert abs(got["h2"][t] - ref["h2"]) <= 1e-3, msg
        assert abs(got["h2"][t] - h2_64) <= 1e-3, msg
        assert abs(got["vg"][t] - ref["vg"]) <= 1e-2 * ref["vg"], msg
        assert np.all(np.abs(got["steps"][t] - np.array(ref["steps"])) <= 1), msg


def _trait_phenotypes(p, masks, seed=13):
    rng = np.random.default_rng(seed)
    Y = np.zeros((len(masks), p.n), np.float32)
    for t in range(len(masks)):
        keep = np.flatnonzero(masks[t])
        idx = rng.choice(p.m, size=150, replace=False)
        gv = p.Z[np.ix_(keep, idx)] @ rng.normal(size=150)
        h2 = [0.2, 0.5, 0.7][t % 3]
        e = rng.normal(size=len(keep)) * np.sqrt(gv.var() * (1 - h2) / h2)
        Y[t, keep] = (gv + e + 5.0).astype(np.float32)
        Y[t, ~masks[t]] = 123.0  # ignored
    return Y


@pytest.mark.parametrize("plan", ["batched", "sequential"])
def test_batched_vc_to_1e_4(batch, plan):
    """Both traits plans with the golden-section search and the Lanczos solves run to 1e-6,
    against the FP64 dense criterion of each trait's subset (exact solves, exact quadrature for the
    same 20 probes) maximised to 1e-6: Vg, Ve and h2 of every trait within 1e-4 relative."""
    import oracle
    from oracle.dense import sldf_gss
    p, masks, g, tr, subs = batch
    Y = _trait_phenotypes(p, masks)
    tr.set_plan(plan)
    got = tr.SLDF_REML(Y, n_V=20, tol_L=1e-6, tol_VC=1e-6)
    tr.set_plan("auto")
    assert got["plan"] == plan
    print(plan, [float(v) for v in got["h2"]], [float(v) for v in got["vg"]], got["steps"].tolist())
    for t, (keep, gs) in enumerate(subs):
        nt = len(keep)
        gm, gw, _ = tr.model(t)
        Q = np.full((nt, 1), 1 / np.sqrt(nt))
        yp = Y[t, keep].astype(np.float64)
        yp = (yp - yp.mean()).astype(np.float32)
        vg, ve, _, h2 = sldf_gss(Dense(p.Z[keep], gm, gw, np.ones(nt), Q), yp, Q, oracle.rademacher(0, nt, 20),
                                 tol=1e-6)
        d = {"vg": abs(got["vg"][t] - vg) / vg, "ve": abs(got["ve"][t] - ve) / ve, "h2": abs(got["h2"][t] - h2) / h2}
        msg = f"{plan} trait {t} (n {nt}): {d}"
        print(msg)
        assert max(d.values()) < 1e-4, msg


def test_graph_replayed_steps_equal_eager_steps(batch, monkeypatch):
    """Lanczos steps after the first replay HIP graphs of the operator and update kernels; every
    replayed step (coefficients) and the replayed solves must equal the eager launches exactly,
    for the batched operator and the single-model operator."""
    p, masks, g, tr, subs = batch
    rng = np.random.default_rng(4)
    cols, col_trait = [], []
    for t in range(3):
        for _ in range(2):
            v = np.zeros(p.n, np.float32)
            v[masks[t]] = rng.normal(size=masks[t].sum())
            cols.append(v)
            col_trait.append(t)
    B = torch.from_numpy(np.stack(cols, 1)).cuda()
    keep, gs = subs[1]
    Bs = torch.from_numpy(rng.normal(size=(len(keep), 3)).astype(np.float32)).cuda()
    runs = {}
    for mode in ("1", "0", "1"):
        monkeypatch.setenv("SLEMM_LANCZOS_GRAPH", mode)
        s1 = tr.L_Seed(col_trait, ["quadform", "ldet"] * 3, 0.4, B, tol=1e-6)
        s2 = gs.L_Seed("quadform", 0.4, Bs, tol=1e-6)
        res = (s1.coeffs(), s1.L_Solve(B, 0.3).cpu().numpy(), s2.coeffs(), s2.L_Solve(Bs, 0.3).cpu().numpy())
        runs.setdefault(mode, []).append(res)
        s1.close()
        s2.close()
    eager = runs["0"][0]
    for res in runs["1"]:
        for a, b in zip((res[0][0], res[0][1], res[1], res[2][0], res[2][1], res[3]),
                        (eager[0][0], eager[0][1], eager[1], eager[2][0], eager[2][1], eager[3])):
            assert np.array_equal(a, b)


def test_prefix_subsets_skip_zero_work(gpu):
    """Subsets that end early (individuals [0, n_t)): the block products skip reduction chunks and
    output rows past each column group's last member; results must still equal subset handles.
    32 columns per trait, so column groups hold one or two traits."""
    p = make_problem(n=2500, m=1300, seed=57)
    sizes = [700, 1300, 2500]
    masks = np.zeros((3, p.n), bool)
    for t, s in enumerate(sizes):
        masks[t, :s] = True
    g = gpu.Geno.from_bed(p.rows, p.n)
    tr = gpu.Traits(g, masks)
    rng = np.random.default_rng(3)
    per = 32
    col_trait = np.repeat(np.arange(3), per)
    kinds = (["quadform"] + ["ldet"] * (per - 1)) * 3
    V = np.zeros((p.n, 3 * per), np.float32)
    for k, t in enumerate(col_trait):
        V[masks[t], k] = rng.normal(size=masks[t].sum())
    out = tr.lnr_op_select(col_trait, kinds, 0.3, torch.from_numpy(V).cuda()).cpu().numpy()
    for t, s in enumerate(sizes):
        keep = np.arange(s)
        gs = g.keep(keep)
        gs.set_model(None, None, np.full((s, 1), 1 / np.sqrt(s), np.float32))
        cols = np.flatnonzero(col_trait == t)
        ref = gs.lnr_op_select("ldet", 0.3, torch.from_numpy(np.ascontiguousarray(V[np.ix_(keep, cols)])).cuda(),
                               col_kind=[kinds[k] for k in cols]).cpu().numpy()
        assert rel_err(out[np.ix_(keep, cols)], ref) < 1e-5, t
        assert np.all(out[s:, cols] == 0.0)
