"""GPU REML drivers (SLDF_REML and LMC_REML, src/reml.cpp:457-882) against the CPU restatement
run on the same inputs, probes and MC draws, and against the exact FP64 REML likelihood."""
import numpy as np
import pytest

from oracle import reml as oreml
from oracle.dense import reml_loglik
from tests.problems import make_phenotype, make_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    p = make_problem(n=1500, m=2000, seed=3)
    y = make_phenotype(p, 0.5)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model()
    return p, y, g


@pytest.fixture(scope="module")
def case_weighted(gpu):
    p = make_problem(n=1300, m=1700, seed=8, weights=True, c=3, persnp=True)
    y = make_phenotype(p, 0.3, seed=2)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(p.persnp, p.rvec)
    return p, y, g


@pytest.fixture(scope="module")
def sldf_fp64(case):
    """The SLDF criterion evaluated exactly in FP64 (dense solves, exact quadrature for the same
    probes) and maximised with the same golden-section schedule."""
    import oracle
    from oracle.dense import sldf_gss
    p, y, g = case
    Q, _ = oreml.covariate_qr(p.X, p.rr)
    yp = (p.rr * y).astype(np.float32)
    yp = (yp - Q @ (Q.T @ yp)).astype(np.float32)
    V = oracle.rademacher(0, p.n, 30)
    vg, ve, lrt, h2 = sldf_gss(p.dense(), yp, Q, V)
    return dict(vg=vg, ve=ve, lrt=lrt, h2=h2)


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_sldf_against_fp64_oracle(case, sldf_fp64, prec):
    p, y, g = case
    got = g.SLDF_REML(p.X, y, n_V=30, precision=prec)
    ref = sldf_fp64
    msg = f"gpu {got['h2']:.7f} {got['vg']:.6f} {got['ve']:.6f} fp64 {ref['h2']:.7f} {ref['vg']:.6f} {ref['ve']:.6f}"
    # The search stops after 21 evaluations with a bracket (0.7 - 1e-4) * 0.618^19 = 7.5e-5 wide
    # and reports vg / (vg + ve) at its last evaluation point.  On the flat top of the criterion,
    # criteria that differ in the 7th digit take different last comparisons (measured: FP64
    # 0.54083, FP32 GPU 0.54089, split-BF16 0.54095, the same with Lanczos tolerance 1e-5), so
    # h2 is compared at twice the search tolerance.
    assert abs(got["h2"] - ref["h2"]) < 2e-4, msg
    assert abs(got["vg"] - ref["vg"]) / ref["vg"] < 5e-4, msg
    assert abs(got["ve"] - ref["ve"]) / ref["ve"] < 5e-4, msg
    assert got["evals"] == 21


def _fp64_tight(p, y, nV=30):
    """sldf_gss run to 1e-6 over the FP64 dense criterion (exact solves, exact v'log(A)v of the
    same probes): the point every variance-component assertion below is measured from."""
    from oracle.dense import sldf_gss
    import oracle
    Q, _ = oreml.covariate_qr(p.X, p.rr)
    yp = (p.rr * y).astype(np.float32)
    yp = (yp - Q @ (Q.T @ yp)).astype(np.float32)
    V = oracle.rademacher(0, p.n, nV)
    vg, ve, _, h2 = sldf_gss(p.dense(), yp, Q, V, tol=1e-6)
    return dict(vg=vg, ve=ve, h2=h2)


def _rel(a, b):
    return {k: abs(a[k] - b[k]) / abs(b[k]) for k in ("vg", "ve", "h2")}


def test_sldf_tight_search_vc_precision(case):
    """The default path with the golden-section search run to tol 1e-6 (30 evaluations) and the
    Lanczos solves to 1e-6, so that neither rounds the answer, against the FP64 dense evaluation
    of the criterion.  Both precisions to 1e-4 relative in Vg, Ve and h2.  The criterion is
    compared in double and y'Py comes from the phenotype seed's FP64 tridiagonal, so it is smooth
    in h2; with y'Py formed from the FP32-rounded L_Solve solution the rounding jitter, about
    (n - c) 1e-9 in criterion units, exceeded the differences the last search steps compare on
    the flat top and left 5e-5 (n=1500) to 2e-4 (n=5000) in h2."""
    p, y, g = case
    ref = _fp64_tight(p, y)
    got = {prec: g.SLDF_REML(p.X, y, n_V=30, tol_L=1e-6, tol_VC=1e-6, precision=prec) for prec in ("fp32", "bf16x2")}
    d64 = {prec: _rel(got[prec], ref) for prec in got}
    msg = f"vs fp64 {d64}"
    print(msg)
    assert all(r["evals"] == 30 for r in got.values()), msg
    assert max(d64["bf16x2"].values()) < 1e-4, msg
    assert max(d64["fp32"].values()) < 1e-4, msg


def test_sldf_reference_replay_criterion(case, monkeypatch):
    """SLEMM_CRITERION_FP32=1 replays the reference's arithmetic: y'Py from the L_Solve solution
    and the criterion rounded to float before each comparison (src/reml.cpp:477-489).  The
    search then stops short of the FP64 optimum by the float resolution of a criterion of
    magnitude ~1e3-1e5 (measured 1.9e-4 / 1.4e-4 / 1.5e-4 in Vg / Ve / h2 at n=1500), which is
    why the default compares in double."""
    p, y, g = case
    ref = _fp64_tight(p, y)
    monkeypatch.setenv("SLEMM_CRITERION_FP32", "1")
    got = g.SLDF_REML(p.X, y, n_V=30, tol_L=1e-6, tol_VC=1e-6, precision="bf16x2")
    d = _rel(got, ref)
    print(f"replay vs fp64: {d}")
    assert got["evals"] == 30
    assert max(d.values()) < 5e-4, d


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_sldf_matches_restatement(case, prec):
    """Against the FP32 CPU restatement (same probes, same schedule).  Its FP32 Lanczos scalars
    perturb the flat criterion near the optimum, so agreement is at the golden-section bracket
    scale (a few 1e-4 on h2), not at the 1e-4 relative of the FP64 comparison above."""
    p, y, g = case
    ref = oreml.SLDF_REML(p.kmat, p.gm, p.gw, None, p.X, y, n_V=30)
    got = g.SLDF_REML(p.X, y, n_V=30, precision=prec)
    assert abs(got["h2"] - ref["h2"]) < 5e-4
    assert abs(got["vg"] - ref["vg"]) / ref["vg"] < 2e-3
    assert abs(got["ve"] - ref["ve"]) / ref["ve"] < 2e-3
    assert abs(got["lrt"] - ref["lrt"]) < 1e-2 * max(1.0, abs(ref["lrt"]))
    assert rel_err(got["blue"], ref["blue"]) < 1e-4
    assert rel_err(got["Py"], ref["Py"]) < 5e-3
    assert rel_err(got["snp_blup"], ref["snp_blup"]) < 5e-3
    assert got["evals"] == ref["evals"] == 21


def test_sldf_near_exact_reml_optimum(case):
    """The SLQ-based optimum sits within the probe noise of the exact FP64 REML argmax."""
    p, y, g = case
    got = g.SLDF_REML(p.X, y, n_V=60, precision="bf16x2")
    hs = np.linspace(0.3, 0.8, 51)
    ll = [reml_loglik(p.Z, p.gm, p.gw, p.X, y, h) for h in hs]
    assert abs(got["h2"] - hs[int(np.argmax(ll))]) < 0.03


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_lmc_matches_restatement(case, prec):
    p, y, g = case
    ref = oreml.LMC_REML(p.kmat, p.gm, p.gw, None, p.X, y)
    got = g.LMC_REML(p.X, y, precision=prec)
    assert got["n_MC"] == ref["n_MC"]
    assert abs(got["h2"] - ref["h2"]) < 2e-3
    assert abs(got["vg"] - ref["vg"]) / ref["vg"] < 5e-3
    assert rel_err(got["blue"], ref["blue"]) < 1e-3
    assert rel_err(got["snp_blup"], ref["snp_blup"]) < 1e-2


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_lmc_against_fp64_oracle(case, prec):
    """LMC_REML (src/reml.cpp:730-882) against its FP64 restatement with exact solves on the same
    MC draws (oracle.dense.lmc_reml).  With the secant run to 1e-6 both land on the root of
    REML_first_order: Vg, Ve and h2 to 1e-4 relative.  With the reference's schedule (0.002, at
    most 9 evaluations) both take the same number of evaluations and report nearby iterates."""
    import oracle
    from oracle.dense import lmc_reml
    p, y, g = case
    D = p.dense()
    n_MC = max(min(int(4e9 / p.n / p.n), 30), 3)
    u, e = oracle.lmc_samples(0, p.n, p.m, n_MC)
    ref = lmc_reml(D, y, u, e, tol=1e-6, max_iter=40)
    got = g.LMC_REML(p.X, y, n_MC=n_MC, tol_L=1e-6, secant_tol=1e-6, secant_max_iter=40, precision=prec)
    d = _rel(got, ref)
    msg = f"{prec}: gpu h2 {got['h2']:.7f} fp64 {ref['h2']:.7f} rel {d} evals {got['evals']} / {ref['evals']}"
    print(msg)
    assert max(d.values()) < 1e-4, msg
    assert rel_err(got["Py"], ref["Py"]) < 1e-4, msg
    ref9 = lmc_reml(D, y, u, e)
    got9 = g.LMC_REML(p.X, y, n_MC=n_MC, tol_L=1e-6, precision=prec)
    assert got9["evals"] == ref9["evals"], (got9["evals"], ref9["evals"])
    assert abs(got9["h2"] - ref9["h2"]) < 1e-3, (got9["h2"], ref9["h2"])


def test_sldf_weighted_covariates(case_weighted):
    p, y, g = case_weighted
    ref = oreml.SLDF_REML(p.kmat, p.gm, p.gw, p.rvec, p.X, y, n_V=30)
    got = g.SLDF_REML(p.X, y, n_V=30, precision="bf16x2")
    assert abs(got["h2"] - ref["h2"]) < 5e-4
    assert abs(got["vg"] - ref["vg"]) / ref["vg"] < 2e-3
    assert rel_err(got["blue"], ref["blue"]) < 1e-3
    assert rel_err(got["var_blue"], ref["var_blue"]) < 1e-3


def test_rank_deficient_covariates_raise(case):
    p, y, g = case
    from slemm_amd._lib import SlemmError, ERR_RANK
    X = np.concatenate([p.X, 2 * p.X], axis=1)
    with pytest.raises(SlemmError) as ei:
        g.SLDF_REML(X, y, n_V=5)
    assert ei.value.code == ERR_RANK


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_lmm_gstat_against_fp64(gpu, prec):
    """--lmm quadratic forms (src/reml.cpp:920-980) after an LMC fit, against the FP64 dense
    evaluation with exact solves; SNPs include both ends of the chromosome (clamped windows)."""
    from oracle.dense import Dense, lmm_gstat
    p = make_problem(n=1200, m=1500, seed=12, weights=True, c=2)
    y = make_phenotype(p, 0.4, seed=3)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec)
    fit = g.LMC_REML(p.X, y, n_MC=10, precision=prec)
    h2 = np.float32(fit["h2"])
    tau1 = float((np.float32(1) - h2) / h2)
    snps = np.array([0, 7, 99, 100, 640, 1399, 1400, 1499])
    w = 200
    got = g.lmm_gstat(p.X, fit["Py"], float(h2), snps, window_size=w, precision=prec)
    D = Dense(p.Z, p.gm, p.gw, p.rr, p.Q)
    xPy, xPx, xBx, xSx = lmm_gstat(D, fit["Py"], tau1, snps, w)
    assert rel_err(got["xPy"], xPy) < 1e-5
    assert rel_err(got["xSx"], xSx) < 1e-5
    assert rel_err(got["xPx"], xPx) < 1e-4, (got["xPx"], xPx)
    assert rel_err(got["xBx"], xBx) < 1e-5, (got["xBx"], xBx)
    assert np.all(got["xBx"] > got["xPx"])


def test_lmm_gstat_fake_against_fp64(gpu):
    """--fake_geno quadratic forms (src/reml.cpp:886-918): the fit's random stream continued after
    its MC samples gives the same fake markers as the oracle's draws (xSx to 1e-5 would fail for
    any other marker), and xPy, xPx, xSx agree with the FP64 dense evaluation."""
    import oracle
    from oracle.dense import Dense, lmm_gstat_fake
    p = make_problem(n=1200, m=1500, seed=12, weights=True, c=2)
    y = make_phenotype(p, 0.4, seed=3)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec)
    seed, nmc, k = 7, 10, 9
    fit = g.LMC_REML(p.X, y, rng_seed=seed, n_MC=nmc)
    h2 = np.float32(fit["h2"])
    tau1 = float((np.float32(1) - h2) / h2)
    got = g.lmm_gstat_fake(p.X, fit["Py"], float(h2), k, rng_seed=seed, n_MC=nmc)
    V0, _ = oracle.fake_geno(seed, p.n, p.m, nmc, k)
    D = Dense(p.Z, p.gm, p.gw, p.rr, p.Q)
    xPy, xPx, xSx = lmm_gstat_fake(D, fit["Py"], tau1, V0)
    assert rel_err(got["xSx"], xSx) < 1e-5
    assert rel_err(got["xPy"], xPy) < 1e-5
    assert rel_err(got["xPx"], xPx) < 1e-4


def test_grammar_gamma_pipeline(gpu):
    """LMC fit -> gstat quadratic forms on 30 sampled SNPs -> correction factor -> GRAMMAR-gamma
    tests of every SNP (bin/slemm_gamma.py:163-174, 306-325), against the dense restatement."""
    from oracle import gwas as ogwas
    from oracle.dense import lmm_gstat
    from slemm_amd import gwas
    p = make_problem(n=1200, m=1500, seed=12, weights=True, c=2)
    y = make_phenotype(p, 0.4, seed=3)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec)
    fit = g.LMC_REML(p.X, y, n_MC=10)
    # the reference's own SNP sample (src/reml.cpp:920-925), from the library (snps=None)
    gst = g.lmm_gstat(p.X, fit["Py"], fit["h2"], None, window_size=200, num_qf_markers=30)
    assert np.array_equal(gst["snps"], ogwas.lmm_sample(p.gm, p.n, 30))
    h2 = np.float32(fit["h2"])
    want = lmm_gstat(p.dense(), fit["Py"], float((np.float32(1) - h2) / h2), gst["snps"], 200)
    for key, w, tol in zip(("xPy", "xPx", "xBx", "xSx"), want, (1e-5, 1e-4, 1e-5, 1e-5)):
        assert rel_err(gst[key], w) < tol, key
    res = gwas.grammar_gamma(g, fit["Py"], gst, snp_var=fit["vg"], mac=20)
    xy, xx = ogwas.gamma_scores(p.Z, fit["Py"], p.rr, p.Q, mac=20)
    cf = ogwas.mean_wo_outliers(gst["xSx"].astype(np.float64) / gst["xPx"].astype(np.float64))
    assert abs(res["cfactor"] - cf) < 1e-12
    beta, se, chisq, pv = ogwas.gamma_stats(xy, xx, cf, fit["vg"])
    ok = np.isfinite(chisq)
    assert np.array_equal(ok, np.isfinite(res["CHISQ"])) and ok.sum() > 1000
    assert rel_err(res["BETA"][ok], beta[ok]) < 3e-4
    assert rel_err(res["SE"][ok], se[ok]) < 3e-4
    assert np.max(np.abs(res["P"][ok] - pv[ok])) < 1e-3


def test_iter_weighting_against_restatement(gpu):
    """--iter_weighting (src/option.cpp:417-528) vs the restated driver over the CPU REML: same
    clumped SNPs (a duplicated large-effect SNP enters once), weights and final estimates."""
    import oracle
    from oracle import bedio
    from oracle.dense import dosage_of_kmat
    p = make_problem(n=1500, m=1200, seed=31)
    codes = p.codes.copy()
    codes[41] = codes[40]
    rows = bedio.encode_codes(codes)
    kmat = bedio.kmat_of_codes(codes)
    _, _, _, _, gm = oracle.counts(kmat)
    Z = dosage_of_kmat(kmat)
    rng = np.random.default_rng(8)
    y = (Z[:, 40] * 3.0 + Z[:, 700] * 2.5 + Z @ rng.normal(scale=0.05, size=p.m) + rng.normal(size=p.n))
    y = y.astype(np.float32)
    maf = oracle.maf(kmat)
    X = np.ones((p.n, 1), np.float32)
    ref = oreml.iter_weighting_reml(kmat, gm, None, X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    g = gpu.Geno.from_bed(rows, p.n)
    got = g.iter_weighting_REML(X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    assert list(got["sig_snps"]) == list(ref["sig_snps"]) and 700 in got["sig_snps"]
    assert (40 in got["sig_snps"]) != (41 in got["sig_snps"])
    assert rel_err(got["persnp"], ref["persnp"]) < 2e-2
    assert abs(got["h2"] - ref["h2"]) < 2e-3
    sig = got["sig_snps"]
    assert rel_err(got["snp_blup"][sig], ref["snp_blup"][sig]) < 1e-2


def test_iter_weighting_against_fp64(gpu):
    """--iter_weighting against the same driver over exact FP64 LMC fits on the same MC draws
    (dense.iter_weighting_reml), with error-variance weights and a second covariate: the same
    clumped SNPs, and the weights, h2, Vg, the BLUPs and the clumped SNPs' estimates within 1e-4
    (the bar for variance components).  The Lanczos seed runs to 1e-6 (tol_L = 1e-7: LMC seeds at
    10 tol_L, src/reml.cpp:810), so that f(h2) is the exact solves' value to FP32 rounding and the
    secant, with the reference's stop, takes the same evaluations as the FP64 driver; at the
    reference's tol_L the seed's 5e-3 residuals move the first fit's secant path (5 evaluations
    instead of 7 here) and the weights by 6e-3."""
    import oracle
    from oracle import bedio
    from oracle import dense as odense
    from oracle.dense import dosage_of_kmat
    p = make_problem(n=1500, m=1200, seed=31, weights=True, c=2)
    codes = p.codes.copy()
    codes[41] = codes[40]
    rows = bedio.encode_codes(codes)
    kmat = bedio.kmat_of_codes(codes)
    _, _, _, _, gm = oracle.counts(kmat)
    Z = dosage_of_kmat(kmat)
    rng = np.random.default_rng(8)
    y = (Z[:, 40] * 3.0 + Z[:, 700] * 2.5 + Z @ rng.normal(scale=0.05, size=p.m) + rng.normal(size=p.n))
    y = y.astype(np.float32)
    maf = oracle.maf(kmat)
    X = np.asarray(p.X, np.float32)
    ref = odense.iter_weighting_reml(kmat, gm, p.rvec, X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    g = gpu.Geno.from_bed(rows, p.n)
    got = g.iter_weighting_REML(X, y, maf, rvec=p.rvec, window_size=20, sig_chisq=8.0, n_MC=8, tol_L=1e-7)
    sig = got["sig_snps"]
    errs = dict(persnp=rel_err(got["persnp"], ref["persnp"]), h2=abs(got["h2"] - ref["h2"]) / ref["h2"],
                vg=abs(got["vg"] - ref["vg"]) / ref["vg"], blup=rel_err(got["snp_blup"], ref["snp_blup"]),
                sig_blup=rel_err(got["snp_blup"][sig], ref["snp_blup"][sig]), blue=rel_err(got["blue"], ref["blue"]))
    msg = " ".join(f"{k} {v:.2e}" for k, v in errs.items())
    print("iter_weighting vs FP64:", msg)
    assert list(sig) == list(ref["sig_snps"]) and 700 in sig and (40 in sig) != (41 in sig), msg
    assert errs["persnp"] < 1e-4 and errs["h2"] < 1e-4 and errs["vg"] < 1e-4, msg
    assert errs["blup"] < 1e-4 and errs["sig_blup"] < 1e-4 and errs["blue"] < 1e-4, msg


@pytest.mark.parametrize("n,m", [(900, 1300), (5001, 700)])
def test_gwa_window_scan_against_fp64(gpu, n, m):
    """Windowed score tests of slemm_gwa step 2 (bin/slemm_gwa.py:322-407) vs the FP64
    restatement: overlapping blocks of model SNPs, every SNP of the range tested, MAC filter.
    n = 5001 splits the block Gram products into two row chunks plus one leftover row."""
    from oracle import gwas as ogwas
    from slemm_amd import gwas
    p = make_problem(n=n, m=m, seed=15, weights=True, c=2, missing=0.0)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec)
    rng = np.random.default_rng(3)
    y = rng.normal(size=p.n).astype(np.float32)
    ms = np.sort(rng.choice(p.m, size=500, replace=False))
    tau, ct, mac = 1.8, 500, 15.0
    for bs, (t0, t1) in ((100, (0, p.m - 1)), (64, (200, min(900, p.m - 1)))):
        xy, xx = g.gwa_window(p.X, y, ms, bs, t0, t1, tau, ct, mac)
        rxy, rxx = ogwas.gwa_window_scores(p.Z, y, p.rr, p.Q, ms, bs, t0, t1, tau, ct, mac)
        assert np.array_equal(np.isnan(xx), np.isnan(rxx)) and np.isfinite(rxx).sum() > 0.8 * (t1 - t0 + 1)
        ok = np.isfinite(rxx)
        assert rel_err(xy[ok], rxy[ok]) < 1e-5
        assert rel_err(xx[ok], rxx[ok]) < 1e-6
    st = gwas.gwa_window_stats(xy, xx, 1.1, 0.5)
    assert np.all(np.isfinite(st["P"][ok])) and np.all((st["P"][ok] >= 0) & (st["P"][ok] <= 1))
