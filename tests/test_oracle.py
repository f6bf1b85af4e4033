"""Pins of the oracle (CPU, no GPU): known answers, brute force, closed forms, the FP64 dense
oracle, the libstdc++ RNG definitions and the reference's example outputs."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import bedio, gwas
from oracle import reml as oreml
from oracle.dense import Dense, dosage_of_kmat, reml_loglik
from tests.problems import make_phenotype, make_problem, rel_err

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def _build(oracle_lib):
    return oracle_lib


# ------------------------------------------------------------------------------ bed decode ----
def test_decode_known_answer_byte():
    # 0b11_10_01_00: individuals 0..3 carry codes 00, 01 (missing), 10 (het), 11 -> dosages 0,1,1,2
    bed = bedio.MAGIC + bytes([0b11100100])
    kmat = oracle.decode_bed(bed, 4, 1)
    assert kmat.tolist() == [[0, 0, 1, 0, 0, 1, 1, 1]]
    assert dosage_of_kmat(kmat)[:, 0].tolist() == [0, 1, 1, 2]


def test_decode_matches_numpy_codec_with_masks_and_order():
    rng = np.random.default_rng(0)
    n, m = 37, 23
    codes = rng.integers(0, 4, size=(m, n)).astype(np.uint8)
    rows = bedio.encode_codes(codes)
    bindi = (rng.random(n) < 0.7).astype(np.uint8)
    bmarker = (rng.random(m) < 0.8).astype(np.uint8)
    kept = np.flatnonzero(bmarker)
    order = rng.permutation(len(kept))
    kmat = oracle.decode_bed(bedio.bed_bytes(rows), n, m, bindi, bmarker, order)
    want = bedio.kmat_of_codes(codes[kept][order][:, bindi.astype(bool)])
    assert np.array_equal(kmat, want)


def test_decode_truncated_bed_raises():
    rows = bedio.encode_codes(np.zeros((5, 9), np.uint8))
    with pytest.raises(ValueError, match="problem with the BED file"):
        oracle.decode_bed(bedio.bed_bytes(rows)[:-3], 9, 5)


# ------------------------------------------------------------------------------ counts ----------
def test_counts_brute_force_and_frequencies():
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 4, size=(50, 333)).astype(np.uint8)
    kmat = bedio.kmat_of_codes(codes)
    pc, n0, n1, n2, gm = oracle.counts(kmat)
    assert np.array_equal(pc, kmat.sum(axis=1))
    assert np.array_equal(n2, (codes == 3).sum(axis=1))
    assert np.array_equal(n1, ((codes == 1) | (codes == 2)).sum(axis=1))
    assert np.array_equal(n0, (codes == 0).sum(axis=1))
    assert np.array_equal(gm, (pc.astype(np.float32) / np.float32(333)).astype(np.float32))
    f = pc.astype(np.float32) / np.float32(666)
    assert np.array_equal(oracle.maf(kmat), np.abs((f > 0.5).astype(np.float32) - f))


def test_gw_closed_form():
    gm = np.array([0.1, 0.5, 1.0, 1.9], np.float32)
    ps = np.array([1.0, 2.0, 3.0, 2.0], np.float32)
    v = (gm * (np.float32(2) - gm)) / np.float32(2)
    want = (np.float32(1) / v) / np.float32(4) * (ps / np.float32(2.0))
    assert np.allclose(oracle.gw(gm, ps), want, rtol=1e-7, atol=0)
    want1 = (np.float32(1) / v) / np.float32(4)
    assert np.array_equal(oracle.gw(gm), want1.astype(np.float32))


# ------------------------------------------------------------------------------ operators -------
@pytest.fixture(scope="module")
def small():
    p = make_problem(n=403, m=517, seed=4, weights=True, c=2)
    return p, p.model(subset_size=100), p.dense()


def test_operators_vs_dense(small):
    p, mdl, D = small
    V = np.random.default_rng(3).normal(size=(p.n, 5)).astype(np.float32)
    for kind in ("ldet", "quadform"):
        assert rel_err(mdl.op(kind, 0.43, V), D.op(kind, 0.43, V)) < 2e-5
    assert rel_err(mdl.rmatvec(V), D.rmatvec(V)) < 2e-5
    assert rel_err(mdl.rmatvec(V, weighted=True), D.rmatvec(V, weighted=True)) < 2e-5
    U = np.random.default_rng(4).normal(size=(p.m, 3)).astype(np.float32)
    assert rel_err(mdl.matvec(U), D.matvec(U)) < 2e-5


def test_centring_identity(small):
    """Zc'v = Z'v - gm (1'v) and Zc u = Z u - 1 (gm'u): the identities the kernels' epilogues use."""
    p, mdl, D = small
    v = np.random.default_rng(5).normal(size=p.n)
    u = np.random.default_rng(6).normal(size=p.m)
    assert np.allclose(D.Zc.T @ v, p.Z.T @ v - p.gm.astype(np.float64) * v.sum(), rtol=1e-12, atol=1e-9)
    assert np.allclose(D.Zc @ u, p.Z @ u - np.dot(p.gm.astype(np.float64), u), rtol=1e-12, atol=1e-9)


def test_quadform_is_projected_ldet(small):
    p, mdl, D = small
    P = np.eye(p.n) - p.Q.astype(np.float64) @ p.Q.T.astype(np.float64)
    assert np.allclose(D.A("quadform", 0.0), P @ D.A("ldet", 0.0) @ P, atol=1e-10)


def test_lseed_lsolve_vs_dense_solve(small):
    p, mdl, D = small
    B = np.random.default_rng(7).normal(size=(p.n, 2)).astype(np.float32)
    tau0 = 0.4286
    s = mdl.L_Seed("ldet", tau0, B, tol=1e-6)
    assert s.converged
    for tau in (0.5, 2.0, 9.0):
        X = s.L_Solve(B, tau - tau0, tol=1e-6)
        assert rel_err(X, D.solve("ldet", tau, B)) < 1e-4
    # Jacobi matrix of a converged Lanczos run: its extreme Ritz values bracket A's spectrum ends
    lam = np.linalg.eigvalsh(D.A("ldet", tau0))
    T = np.diag(s.delta[0]) + np.diag(s.beta[0, 1:s.steps], 1) + np.diag(s.beta[0, 1:s.steps], -1)
    ritz = np.linalg.eigvalsh(T.astype(np.float64))
    assert lam[0] - 1e-4 <= ritz[0] and ritz[-1] <= lam[-1] + 1e-3 * lam[-1]
    assert abs(ritz[-1] - lam[-1]) / lam[-1] < 1e-3


def test_slq_vs_exact_log(small):
    p, mdl, D = small
    V = oracle.rademacher(1, p.n, 4)
    tau0 = 0.4286
    s = mdl.L_Seed("ldet", tau0, V, tol=1e-6)
    for tau in (0.6, 3.0):
        est = sum(float(np.sum(W ** 2 * np.log(Dv + tau - tau0)))
                  for Dv, W in (oreml.slq_spectrum(s, k) for k in range(4)))
        exact = D.quad_logm("ldet", tau, V)
        assert abs(est - exact) / abs(exact) < 1e-4


# ------------------------------------------------------------------------------ RNG -------------
def _mt_raw(seed, count):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    return bg.random_raw(count).astype(np.uint64)


def test_rademacher_is_mt19937_uniform_float():
    """std::mt19937(seed+54321) + uniform_real_distribution<float>(0,1): u = float(x) / 2^32 from
    one 32-bit draw (generate_canonical<float,24>); v = -1 if u < 0.5 (src/reml.cpp:597-605)."""
    n, nv, seed = 50, 3, 11
    x = _mt_raw(seed + 54321, n * nv)
    u = x.astype(np.float32) / np.float32(4294967296.0)
    u = np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u)
    signs = np.where(u < 0.5, -1.0, 1.0).reshape(nv, n).T
    V = oracle.rademacher(seed, n, nv)
    assert np.array_equal(np.sign(V), signs)
    assert np.allclose(np.abs(V), np.float32(1) / np.sqrt(np.float32(n)), rtol=0, atol=1e-7)


def test_lmc_samples_are_mt19937_marsaglia_polar():
    """normal_distribution<float> in libstdc++: Marsaglia polar method on pairs of
    2u-1 draws, caching the second deviate (src/reml.cpp:795-809)."""
    n, m, nmc, seed = 7, 9, 2, 5
    u_o, e_o = oracle.lmc_samples(seed, n, m, nmc)
    raw = _mt_raw(seed + 54321, 4 * (m + n) * nmc + 64)
    pos = [0]

    def unif():
        v = np.float32(raw[pos[0]]) / np.float32(4294967296.0)
        pos[0] += 1
        return v

    out = []
    saved = None
    while len(out) < (m + n) * nmc:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            x = np.float32(2.0) * unif() - np.float32(1.0)
            y = np.float32(2.0) * unif() - np.float32(1.0)
            r2 = x * x + y * y
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = np.sqrt(np.float32(-2.0) * np.log(r2) / r2)
        saved = x * mult
        out.append(y * mult)
    want = np.array(out, np.float32)
    got = np.concatenate([u_o.T.ravel(), e_o.T.ravel()])
    assert np.allclose(got, want, rtol=2e-6, atol=1e-7)


class _Libstdcxx:
    """libstdc++'s distributions over a raw mt19937 stream (numpy's MT19937 with the legacy
    init_genrand seeding): generate_canonical<float, 24> takes one 32-bit draw, <double, 53> two
    (low word first); normal_distribution<float> is Marsaglia's polar method keeping the second
    deviate; uniform_real_distribution<float>(a, b) is U * (b - a) + a in float;
    binomial_distribution<int>(t, p) with t * min(p, 1-p) < 8 is Devroye's second waiting-time
    method (Non-Uniform Random Variate Generation, 1986, ch. X.4) on min(p, 1-p), flipped to t - x
    when p > 0.5."""

    def __init__(self, seed, count):
        self.raw = _mt_raw(seed, count)
        self.pos = 0
        self.saved = None

    def u32(self):
        v = int(self.raw[self.pos])
        self.pos += 1
        return v

    def canon_f(self):
        v = np.float32(self.u32()) / np.float32(4294967296.0)
        return np.nextafter(np.float32(1), np.float32(0)) if v >= 1 else v

    def canon_d(self):
        x0 = float(self.u32())
        s = x0 + float(self.u32()) * 4294967296.0
        v = s / 18446744073709551616.0
        return math.nextafter(1.0, 0.0) if v >= 1 else v

    def normal_f(self):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = np.float32(2.0) * self.canon_f() - np.float32(1.0)
            y = np.float32(2.0) * self.canon_f() - np.float32(1.0)
            r2 = x * x + y * y
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = np.sqrt(np.float32(-2.0) * np.log(r2) / r2)
        self.saved = x * mult
        return y * mult

    def uniform_f(self, a, b):
        a, b = np.float32(a), np.float32(b)
        return self.canon_f() * (b - a) + a

    def binomial(self, t, p):
        p = float(p)
        p12 = p if p <= 0.5 else 1.0 - p
        q = -math.log(1.0 - p12)
        x, s = 0, 0.0
        while True:
            if x == t:
                break
            s += -math.log(1.0 - self.canon_d()) / (t - x)
            x += 1
            if not s <= q:
                x -= 1
                break
        return t - x if p12 != p else x


def test_fake_geno_is_libstdcxx_stream():
    """--fake_geno markers (src/reml.cpp:886-899): the LMC stream's n_MC*(m+n) normals replayed,
    then per marker one uniform_real<float>(0.01, 0.99) frequency and n binomial(2, p) dosages."""
    n, m, nmc, k, seed = 23, 11, 2, 14, 3
    V, f = oracle.fake_geno(seed, n, m, nmc, k)
    st = _Libstdcxx(seed + 54321, 8 * nmc * (m + n) + 4 * n * k + 64)
    for _ in range(nmc * (m + n)):
        st.normal_f()
    for j in range(k):
        a1 = st.uniform_f(0.01, 0.99)
        assert f[j] == a1, (j, f[j], a1)
        col = [st.binomial(2, a1) for _ in range(n)]
        assert V[:, j].tolist() == col, j
    assert np.any(f > 0.5) and np.any(f < 0.5)  # both branches of the flip
    assert set(np.unique(V)) <= {0.0, 1.0, 2.0}


# ------------------------------------------------------------------------------ GEBV ------------
def test_gebv_exact():
    rng = np.random.default_rng(9)
    n, m = 61, 40
    codes = rng.integers(0, 4, size=(m, n)).astype(np.uint8)
    rows = bedio.encode_codes(codes)
    bm = (rng.random(m) < 0.8).astype(np.uint8)
    bf = (rng.random(m) < 0.4).astype(np.uint8)
    alpha = rng.normal(size=m)
    got = oracle.gebv(bedio.bed_bytes(rows), n, m, bm, bf, alpha)
    z = bedio.dosage_of_codes(codes).astype(np.int64)
    z = np.where(bf[:, None] == 1, 2 - z, z)
    want = np.zeros(n)
    for j in range(m):  # same order as the reference: SNP by SNP, FP64
        if bm[j]:
            want += z[j] * alpha[j]
    assert np.array_equal(got, want)


# ------------------------------------------------------------------------------ REML drivers ----
def test_sldf_restatement_near_exact_optimum():
    p = make_problem(n=900, m=1200, seed=12)
    y = make_phenotype(p, 0.4, seed=1)
    r = oreml.SLDF_REML(p.kmat, p.gm, p.gw, None, p.X, y, n_V=50)
    hs = np.linspace(0.1, 0.8, 71)
    ll = [reml_loglik(p.Z, p.gm, p.gw, p.X, y, h) for h in hs]
    assert abs(r["h2"] - hs[int(np.argmax(ll))]) < 0.03
    assert r["evals"] == 21  # 2 + ceil(ln(1e-4/0.6999)/ln 0.618) golden-section evaluations
    # with a single intercept the BLUE is the GLS mean, and Vg = y'Py/(n-c) holds at the optimum
    assert abs(r["vg"] - oreml.vg_from_py(y, r["Py"], 1)) / r["vg"] < 1e-3


def _sldf_inputs(p, y):
    Q, _ = oreml.covariate_qr(p.X, p.rr)
    yp = (p.rr * np.asarray(y, np.float32)).astype(np.float32)
    yp = (yp - Q @ (Q.T @ yp)).astype(np.float32)
    return Q, yp


@pytest.fixture(scope="module")
def weighted_small():
    """n=240 with error-variance weights R, per-SNP weights and two covariates, so that R^-1/2,
    the (n - c) factor and the covariate log-determinant all enter the criterion."""
    p = make_problem(n=240, m=400, seed=21, weights=True, c=2, persnp=True)
    y = make_phenotype(p, 0.45, n_causal=80, seed=5)
    return p, y


def test_sldf_criterion_is_twice_reml_loglik_plus_const(weighted_small):
    """With exact log-determinants -- the n unit vectors as probes, for which
    n/n_V sum_k e_k' log(A + sigma I) e_k = log|A + sigma I| -- the FP64 SLDF criterion of
    src/reml.cpp:457-489 equals 2 x the restricted log-likelihood plus a constant in h2:
    log|H| = log|H0| + log|R|, X'H^-1 X = R_x' Q'H0^-1 Q R_x, and the projected operator's
    y'Py is the REML y'Py.  A dropped term, a sign or an n in place of n - c breaks the
    constancy.  The criterion is compared on a grid of h2 that spans the search range."""
    from oracle.dense import sldf_criterion
    p, y = weighted_small
    Q, yp = _sldf_inputs(p, y)
    D = p.dense()
    I = np.eye(p.n)
    tau0 = (1 - 0.7) / 0.7
    hs = np.linspace(0.05, 0.69, 17)
    diff = []
    for h in hs:
        crit, vg, ve, lrt = sldf_criterion(D, yp, Q, I, h, tau0)
        ll = reml_loglik(p.Z, p.gm, p.gw, p.X, y, h, rr=p.rr)
        diff.append(crit - 2.0 * ll)
        # Vg = y'Py / (n - c) with the REML projection (src/reml.cpp:479), Ve = Vg tau
        tau = (1 - h) / h
        H = (p.rr[:, None] ** -1) * (D.A_ldet0 + tau * I) * (p.rr[None, :] ** -1)
        Hi = np.linalg.inv(H)
        X = p.X
        Pm = Hi - Hi @ X @ np.linalg.solve(X.T @ Hi @ X, X.T @ Hi)
        yd = np.asarray(y, np.float64)
        # (y_p and Q are float32 inputs, as the reference keeps them: 1e-8 relative in y'Py)
        assert abs(vg - yd @ Pm @ yd / (p.n - p.X.shape[1])) < 1e-6 * vg
        assert abs(ve - vg * tau) < 1e-12 * ve
    diff = np.asarray(diff)
    scale = np.max(np.abs(diff))
    print("criterion - 2 loglik over the grid: spread", np.ptp(diff), "scale", scale)
    assert np.ptp(diff) < 1e-6 * max(scale, 1.0), diff
    # the constant itself: (n - c) log(y_p'y_p) + log|R| + 2 log|R_x|
    _, Rx = oreml.covariate_qr(p.X, p.rr)
    const = ((p.n - p.X.shape[1]) * np.log(float(yp.astype(np.float64) @ yp))
             + np.sum(np.log(1.0 / p.rr.astype(np.float64) ** 2)) + 2 * np.sum(np.log(np.abs(np.diag(Rx)))))
    assert abs(diff[0] - const) < 1e-8 * max(abs(const), 1.0)


def test_sldf_gss_finds_the_exact_reml_argmax(weighted_small):
    """REML_GSS restated in FP64 (src/reml.cpp:493-547) over the exact criterion (identity
    probes) lands on the bounded argmax of the exact REML likelihood within the bracket width,
    and returns the variance components of its last evaluation point."""
    from scipy.optimize import minimize_scalar
    from oracle.dense import sldf_criterion, sldf_gss
    p, y = weighted_small
    Q, yp = _sldf_inputs(p, y)
    D = p.dense()
    I = np.eye(p.n)
    opt = minimize_scalar(lambda h: -reml_loglik(p.Z, p.gm, p.gw, p.X, y, h, rr=p.rr),
                          bounds=(1e-4, 0.7), method="bounded", options={"xatol": 1e-9})
    for tol in (1e-4, 1e-6):
        vg, ve, lrt, h2 = sldf_gss(D, yp, Q, I, tol=tol)
        assert abs(h2 - opt.x) < 1.5 * tol, (h2, opt.x, tol)
        _, vg_at, ve_at, lrt_at = sldf_criterion(D, yp, Q, I, h2, (1 - 0.7) / 0.7)
        assert abs(vg - vg_at) < 1e-9 * vg and abs(ve - ve_at) < 1e-9 * ve
    # stochastic probes move the optimum only by the probe noise of the log-determinant
    vgs, ves, _, h2s = sldf_gss(D, yp, Q, oracle.rademacher(0, p.n, 30), tol=1e-6)
    assert abs(h2s - opt.x) < 0.05


def test_sldf_restatement_tracks_fp64_criterion():
    """The FP32 restatement of SLDF_REML against the pinned FP64 criterion on the same probes
    and schedule: its FP32 Lanczos scalars and float criterion perturb only the flat top."""
    from oracle.dense import sldf_gss
    p = make_problem(n=600, m=900, seed=12)
    y = make_phenotype(p, 0.4, seed=1)
    Q, yp = _sldf_inputs(p, y)
    r = oreml.SLDF_REML(p.kmat, p.gm, p.gw, None, p.X, y, n_V=30)
    vg, ve, _, h2 = sldf_gss(p.dense(), yp, Q, oracle.rademacher(0, p.n, 30))
    assert abs(r["h2"] - h2) < 1e-3, (r["h2"], h2)
    assert abs(r["vg"] - vg) / vg < 3e-3 and abs(r["ve"] - ve) / ve < 3e-3


def test_lmc_fp64_expected_ratio_root_is_reml_argmax(weighted_small):
    """With the MC ratio of REML_first_order replaced by its expectation tr(A M)/tr(M), the first
    order condition of src/reml.cpp:696-728 is the REML stationarity condition in tau, so the
    FP64 secant (src/reml.cpp:821-853, run to 1e-10) lands on the argmax of the exact REML
    likelihood.  The y part (R^-1/2 scaling, projections, sqrt(gw) Zc'), the secant update and
    vg = y'Py/(n-c) are what this checks."""
    from scipy.optimize import minimize_scalar
    from oracle.dense import lmc_reml
    p, y = weighted_small
    D = p.dense()
    u, e = oracle.lmc_samples(0, p.n, p.m, 3)
    r = lmc_reml(D, y, u, e, tol=1e-10, max_iter=60, expect=True)
    opt = minimize_scalar(lambda h: -reml_loglik(p.Z, p.gm, p.gw, p.X, y, h, rr=p.rr),
                          bounds=(1e-3, 0.99), method="bounded", options={"xatol": 1e-10})
    assert abs(r["h2"] - opt.x) < 1e-6, (r["h2"], opt.x)
    assert abs(r["f"]) < 1e-8
    # vg at the optimum is the profile REML estimate y'Py/(n - c)
    tau = (1 - r["h2"]) / r["h2"]
    H = (D.A_ldet0 + tau * np.eye(p.n)) / (p.rr[:, None] * p.rr[None, :])
    Hi = np.linalg.inv(H)
    X = p.X
    Pm = Hi - Hi @ X @ np.linalg.solve(X.T @ Hi @ X, X.T @ Hi)
    yd = np.asarray(y, np.float64)
    assert abs(r["vg"] - yd @ Pm @ yd / (p.n - X.shape[1])) < 1e-6 * r["vg"]


def test_lmc_fp64_mc_ratio_estimates_its_expectation(weighted_small):
    """The MC half of REML_first_order (src/reml.cpp:788-813, 715-727) on 400 draws from the
    libstdc++ stream estimates tr(A M)/tr(M) (the draws have covariance R^-1/2 Zc W Zc' R^-1/2 + tau
    I after projection): the two forms of f agree to the MC error, at two heritabilities."""
    from oracle.dense import lmc_first_order, lmc_inputs
    p, y = weighted_small
    D = p.dense()
    u, e = oracle.lmc_samples(5, p.n, p.m, 400)
    V = lmc_inputs(D, y, u, e)
    for h in (0.3, 0.6):
        f_mc, _ = lmc_first_order(D, V, h)
        f_ex, _ = lmc_first_order(D, V, h, expect=True)
        assert abs(f_mc - f_ex) < 0.015, (h, f_mc, f_ex)


def test_lmc_restatement_tracks_fp64_oracle():
    """The FP32 restatement of LMC_REML against the pinned FP64 LMC on the same MC draws and the
    reference's secant schedule (0.002, at most 9 evaluations)."""
    from oracle.dense import lmc_reml
    p = make_problem(n=600, m=900, seed=13)
    y = make_phenotype(p, 0.5, seed=2)
    r = oreml.LMC_REML(p.kmat, p.gm, p.gw, None, p.X, y)
    u, e = oracle.lmc_samples(0, p.n, p.m, r["n_MC"])
    ref = lmc_reml(p.dense(), y, u, e)
    assert r["evals"] == ref["evals"]
    assert abs(r["h2"] - ref["h2"]) < 2e-4, (r["h2"], ref["h2"])
    assert abs(r["vg"] - ref["vg"]) / ref["vg"] < 1e-3


def test_lmc_restatement_near_exact_optimum():
    p = make_problem(n=900, m=1200, seed=13)
    y = make_phenotype(p, 0.5, seed=2)
    r = oreml.LMC_REML(p.kmat, p.gm, p.gw, None, p.X, y)
    hs = np.linspace(0.1, 0.9, 81)
    ll = [reml_loglik(p.Z, p.gm, p.gw, p.X, y, h) for h in hs]
    assert r["n_MC"] == 30
    assert abs(r["h2"] - hs[int(np.argmax(ll))]) < 0.06


# ------------------------------------------------------------------------------ example outputs -
@pytest.fixture(scope="module")
def ex10k():
    d = np.load(GOLDEN / "ex10k_reml.npz")
    meta = json.loads((GOLDEN / "ex10k_meta.json").read_text())
    return d, meta


def test_example_vg_from_py(ex10k):
    d, meta = ex10k
    vg = oreml.vg_from_py(d["y"], d["Py"], meta["c"])
    assert abs(vg - meta["Vg"]) / meta["Vg"] < 5e-6
    assert abs(meta["Vg"] / (meta["Vg"] + meta["Ve"]) - meta["h2"]) < 1e-6
    assert abs(float(np.sum(d["Py"].astype(np.float64)))) < 1e-3 * float(np.abs(d["Py"]).sum()) / 1e3
    assert meta["pop_size"] == len(d["y"])


def test_example_gstat_statistics(ex10k):
    _, meta = ex10k
    for row in meta["gstat"]:
        chisq = gwas.gstat_chisq(float(row["xPy"]), float(row["xPx"]), meta["Vg"])
        assert abs(chisq - float(row["ChiSq"])) / float(row["ChiSq"]) < 2e-5
        assert abs(gwas.one_df_chisq_pval(float(row["ChiSq"])) - float(row["Pval"])) < 1e-5
        assert gwas.mac_of_maf(float(row["MAF"]), meta["pop_size"]) == int(row["MAC"])


def test_example_gamma_statistics(ex10k):
    _, meta = ex10k
    hdr = meta["gamma_header"]
    ib, isd, ic, ip = (hdr.index(k) for k in ("BETA", "SE", "CHISQ", "P"))
    for r in meta["gamma_rows"]:
        beta, se, chisq, pv = float(r[ib]), float(r[isd]), float(r[ic]), float(r[ip])
        assert abs((beta / se) ** 2 - chisq) <= 1e-5 * max(chisq, 1e-3)
        assert abs(math.erfc(math.sqrt(chisq / 2)) - pv) < 1e-6


def test_example_maf_are_allele_fractions(ex10k):
    _, meta = ex10k
    two_n = 2 * meta["pop_size"]
    for f in meta["maf"]:
        k = round(f * two_n)
        assert abs(f - k / two_n) < 1e-6


# ---------------------------------------------------------------------- LMM quadratic forms ----
def test_gstat_window_form_is_woodbury_and_whole_genome_window_is_xPx():
    """xBx = (v'v - z'(Zw'Zw + tau1 D^-1)^-1 z)/tau1 equals v'(Zw D Zw' + tau1 I)^-1 v (Woodbury),
    and with the window widened to every SNP it is xPx; windows at both ends hold their SNP."""
    from oracle.dense import lmm_gstat
    p = make_problem(n=300, m=160, seed=21, weights=True, c=2)
    D = p.dense()
    tau1 = 1.7
    Py = np.random.default_rng(5).normal(size=p.n)
    snps = np.array([0, 3, 15, 16, 80, 143, 144, 159])
    w = 32
    xPy, xPx, xBx, xSx = lmm_gstat(D, Py, tau1, snps, w)
    assert np.allclose(xPy, D.Zc[:, snps].T @ Py)
    for i, j in enumerate(snps):
        wds = 0 if j < w // 2 else (p.m - w if j > p.m - w // 2 else j - w // 2)
        assert wds <= j < wds + w and 0 <= wds and wds + w <= p.m
        Zw = D.P @ (D.rr[:, None] * D.Zc[:, wds:wds + w])
        v = D.P @ (D.rr * D.Zc[:, j])
        ref = v @ np.linalg.solve(Zw @ np.diag(D.gw[wds:wds + w]) @ Zw.T + tau1 * np.eye(p.n), v)
        assert abs(xBx[i] - ref) < 1e-9 * abs(ref)
        assert abs(xSx[i] - v @ v) < 1e-9 * (v @ v)
    _, xPx_all, xBx_all, _ = lmm_gstat(D, Py, tau1, snps, p.m)
    assert np.allclose(xBx_all, xPx_all, rtol=1e-9)
    assert np.allclose(xPx_all, xPx, rtol=1e-12)
    assert np.all(xBx > xPx)  # a window explains less of v than the whole genome


# ----------------------------------------------------------------------- iterative weighting ----
def _iterw_problem():
    p = make_problem(n=500, m=240, seed=31)
    codes = p.codes.copy()
    codes[41] = codes[40]  # SNP 41 duplicates SNP 40
    rows = bedio.encode_codes(codes)
    kmat = bedio.kmat_of_codes(codes)
    pc, _, _, _, gm = oracle.counts(kmat)
    Z = dosage_of_kmat(kmat)
    rng = np.random.default_rng(8)
    y = (Z[:, 40] * 3.0 + Z[:, 150] * 2.5 + Z @ rng.normal(scale=0.05, size=p.m) + rng.normal(size=p.n))
    return p, codes, rows, kmat, gm, y.astype(np.float32)


def test_iter_weighting_restatement_rules():
    p, codes, rows, kmat, gm, y = _iterw_problem()
    maf = oracle.maf(kmat)
    X = np.ones((p.n, 1))
    r = oreml.iter_weighting_reml(kmat, gm, None, X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    sig = list(r["sig_snps"])
    assert 150 in sig and (40 in sig) != (41 in sig)  # one copy of the duplicated SNP, not both
    assert np.all(r["persnp"][sig] == 0) and abs(float(r["persnp"].mean()) - 1) < 1e-6
    assert np.all(r["persnp"] >= 0)
    # disabled when the window is wider than the genome: one fit with persnp / mean(persnp)
    ps = np.random.default_rng(2).uniform(0.5, 2.0, p.m).astype(np.float32)
    r2 = oreml.iter_weighting_reml(kmat, gm, None, X, y, maf, persnp=ps, window_size=p.m, n_MC=8)
    r3 = oreml.LMC_REML(kmat, gm, oracle.gw(gm, ps / np.float32(ps.mean())), None, X, y, n_MC=8)
    assert r2["h2"] == r3["h2"] and len(r2["sig_snps"]) == 0


def test_lmc_fit_blup_and_blue_are_the_mixed_model_solutions():
    """dense.lmc_fit's BLUP and BLUE (src/reml.cpp:866-877) at its h2 against the mixed-model
    solutions formed directly: with H = Zc W Zc' + tau R, b = (X'H^-1 X)^-1 X'H^-1 y (GLS) and
    a = W Zc' H^-1 (y - X b); Py = H^-1 (y - X b) and vg = y'Py/(n - c).  A tau in place of 1/tau,
    a dropped R or an unweighted Zc' fails these."""
    from oracle import dense as odense
    p = make_problem(n=300, m=250, seed=44, weights=True, c=2, persnp=True)
    y = make_phenotype(p, 0.5, n_causal=60, seed=9)
    r = odense.lmc_fit(p.Z, p.gm, p.gw, p.rvec, p.X, y, rng_seed=3, n_MC=6)
    tau = (1 - r["h2"]) / r["h2"]
    Zc = p.Z - p.gm[None, :].astype(np.float64)
    H = (Zc * np.asarray(p.gw, np.float64)[None, :]) @ Zc.T + tau * np.diag(np.asarray(p.rvec, np.float64))
    X = np.asarray(p.X, np.float64)
    Hi_X = np.linalg.solve(H, X)
    b = np.linalg.solve(X.T @ Hi_X, Hi_X.T @ y.astype(np.float64))
    Py = np.linalg.solve(H, y.astype(np.float64) - X @ b)
    a = np.asarray(p.gw, np.float64) * (Zc.T @ Py)
    assert rel_err(r["Py"], Py) < 1e-9
    assert rel_err(r["snp_blup"], a) < 1e-9
    assert rel_err(r["blue"], b) < 1e-9
    assert abs(r["vg"] - float(y.astype(np.float64) @ Py) / (p.n - X.shape[1])) < 1e-9 * r["vg"]
    assert r["n_MC"] == 6 and r["evals"] <= 9


def test_iter_weighting_fp64_driver_tracks_restatement():
    """The FP64 --iter_weighting driver (exact LMC fits on the same draws) against the FP32
    restatement: same clumped SNPs, weights at FP32-fit scale, and with the weighting disabled
    exactly one FP64 fit with persnp / mean(persnp)."""
    from oracle import dense as odense
    p, codes, rows, kmat, gm, y = _iterw_problem()
    maf = oracle.maf(kmat)
    X = np.ones((p.n, 1))
    r32 = oreml.iter_weighting_reml(kmat, gm, None, X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    r64 = odense.iter_weighting_reml(kmat, gm, None, X, y, maf, window_size=20, sig_chisq=8.0, n_MC=8)
    assert list(r64["sig_snps"]) == list(r32["sig_snps"])
    assert rel_err(r64["persnp"], r32["persnp"]) < 2e-2
    assert abs(r64["h2"] - r32["h2"]) < 5e-3
    ps = np.random.default_rng(2).uniform(0.5, 2.0, p.m).astype(np.float32)
    r2 = odense.iter_weighting_reml(kmat, gm, None, X, y, maf, persnp=ps, window_size=p.m, n_MC=8)
    r3 = odense.lmc_fit(dosage_of_kmat(kmat), gm, oracle.gw(gm, ps / np.float32(ps.mean())), None, X, y, n_MC=8)
    assert r2["h2"] == r3["h2"] and len(r2["sig_snps"]) == 0


# ------------------------------------------------------------------------ windowed GWA scan ----
def test_gwa_blocks_known_layout():
    """Blocks of 8 over 20 model SNPs at odd positions: step 4, bound at block position 6."""
    ms = np.arange(1, 40, 2)
    blocks, bounds = gwas.gwa_blocks(ms, 8, 0, 40)
    assert [list(b) for b in blocks] == [list(ms[0:8]), list(ms[4:12]), list(ms[8:16]), list(ms[12:20])]
    assert bounds.tolist() == [[0, 13], [13, 21], [21, 29], [29, 41]]
    b2, _ = gwas.gwa_blocks(ms[:18], 8, 0, 40)  # (18 - 4) % 4 != 0: a last block ending at the end
    assert list(b2[-1]) == list(ms[10:18])


def test_gwa_window_score_is_woodbury_form():
    """x'x - z'(Zb'Zb + D)^-1 z = x'(I + Zb D^-1 Zb')^-1 x for one block holding every model SNP."""
    p = make_problem(n=220, m=60, seed=41, weights=True, c=2)
    ms = np.arange(0, 60, 2)
    tau, ct = 1.3, 30
    xy, xx = gwas.gwa_window_scores(p.Z, np.ones(p.n), p.rr, p.Q, ms, len(ms), 0, 59, tau, ct)
    mu = (p.Z[:, ms].sum(axis=0) / p.n).astype(np.float32)
    Zb = (p.Z[:, ms] - mu) * p.rr[:, None]
    Zb -= p.Q @ (p.Q.T @ Zb)
    D = ct * tau * ((2 - mu) * mu / 2).astype(np.float64)
    M = np.eye(p.n) + Zb @ np.diag(1 / D) @ Zb.T
    for j in (1, 7, 30, 59):
        x = p.Z[:, j] - np.float32(p.Z[:, j].sum()) / np.float32(p.n)
        v = x * p.rr
        v -= p.Q @ (p.Q.T @ v)
        assert abs(xx[j] - v @ np.linalg.solve(M, v) / tau) < 1e-9 * abs(xx[j])


def test_panel_engine_matches_dense_and_cxx_restatement():
    """The MKL panel engine of the CPU baseline (oracle.cpu_panels, per-subset float panels and
    two sgemms as src/reml.cpp:28-94) against the FP64 dense operator, and its Lanczos seed and
    SLDF / LMC drivers against the C++ restatement on the same inputs."""
    from oracle.cpu_panels import PanelModel
    p = make_problem(n=700, m=1900, seed=33, weights=True, c=2)
    y = make_phenotype(p, 0.4, seed=4)
    D = p.dense()
    eng = PanelModel(p.rows, p.n, p.gm, p.gw, p.rr, p.Q, subset_size=500)
    V = np.random.default_rng(1).normal(size=(p.n, 5)).astype(np.float32)
    U = np.random.default_rng(2).normal(size=(p.m, 3)).astype(np.float32)
    for kind in ("ldet", "quadform"):
        assert rel_err(eng.op(kind, 0.43, V), D.op(kind, 0.43, V)) < 2e-5
    assert rel_err(eng.rmatvec(V, weighted=True), D.rmatvec(V, weighted=True)) < 2e-5
    assert rel_err(eng.matvec(U), D.matvec(U)) < 2e-5
    s = eng.L_Seed("ldet", 0.43, V, tol=1e-6)
    ref = p.model(500).L_Seed("ldet", 0.43, V, tol=1e-6)
    assert abs(s.steps - ref.steps) <= 1
    X = s.L_Solve(V, 0.2)
    assert rel_err(X, D.solve("ldet", 0.63, V)) < 1e-4
    panel = lambda gm, gw, rr, Q, ss: PanelModel(p.rows, p.n, gm, gw, rr, Q, ss)
    a = oreml.SLDF_REML(p.kmat, p.gm, p.gw, p.rvec, p.X, y, n_V=20)
    b = oreml.SLDF_REML(None, p.gm, p.gw, p.rvec, p.X, y, n_V=20, engine=panel)
    assert abs(a["h2"] - b["h2"]) < 5e-4 and a["evals"] == b["evals"]
    a = oreml.LMC_REML(p.kmat, p.gm, p.gw, p.rvec, p.X, y)
    b = oreml.LMC_REML(None, p.gm, p.gw, p.rvec, p.X, y, engine=panel)
    assert abs(a["h2"] - b["h2"]) < 2e-3
