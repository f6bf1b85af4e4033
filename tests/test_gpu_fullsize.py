"""GPU checks at the benchmark's full sizes, through properties that do not need the whole bed on
the host: bit-exact counts and FP64 GEBV on a downloaded block of SNPs (with the effects zero
outside it, the reference's SNP-order sum over all SNPs equals the sum over the block), the
GEBV checksum sum_i GEBV_i = sum_j alpha_j popcnt_j, the operator's symmetry and its split-BF16
result against the independent FP32 kernel at K = 65 (64 probes + 1 RHS, the tail hi/lo layout),
Zc'V of a block of SNPs against FP64 in both precisions,
the centring identity Zc'1 = popcnt - n gm, and the GWAS score terms of a block of SNPs against FP64.
Shapes: n = 100,000 x m = 500,000 and n = 1,000,000 x m = 60,000."""
import numpy as np
import pytest
import torch

import oracle
from oracle import bedio

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(gpu):
    n, m = 100_000, 500_000
    g = gpu.Geno.synthesize(n, m, seed=2024, maf_lo=0.05, maf_hi=0.5, missing=0.01)
    g.set_model(None, None, np.full((n, 1), 1.0 / np.sqrt(n), np.float32))
    yield g, n, m
    g.close()


def _gebv_block_check(g, n, m, j0, cnt, seed):
    rng = np.random.default_rng(seed)
    alpha = np.zeros(m)
    a_blk = rng.normal(size=cnt) * np.exp2(rng.integers(-20, 20, size=cnt))
    a_blk[5::13] = -0.0
    alpha[j0:j0 + cnt] = a_blk
    flip = np.zeros(m, np.uint8)
    flip[j0:j0 + cnt] = rng.random(cnt) < 0.3
    got = g.gebv(torch.from_numpy(alpha).cuda(), flip).cpu().numpy()
    rows = g.download(j0, cnt)
    want = oracle.gebv(bedio.bed_bytes(rows), n, cnt, np.ones(cnt, np.uint8), flip[j0:j0 + cnt], a_blk)
    assert np.array_equal(got.view(np.uint64), np.asarray(want, np.float64).view(np.uint64))


def test_c2_counts_block_bit_exact(c2):
    g, n, m = c2
    pc, n0, n1, n2, gm = g.counts()
    j0, cnt = 311_007, 300
    codes = bedio.decode_codes(g.download(j0, cnt), n)
    kmat = bedio.kmat_of_codes(codes)
    opc, on0, on1, on2, ogm = oracle.counts(kmat)
    assert np.array_equal(pc[j0:j0 + cnt], opc) and np.array_equal(n2[j0:j0 + cnt], on2)
    assert np.array_equal(n1[j0:j0 + cnt], on1) and np.array_equal(n0[j0:j0 + cnt], on0)
    assert np.array_equal(gm[j0:j0 + cnt].view(np.uint32), np.asarray(ogm, np.float32).view(np.uint32))


def test_c2_gebv_block_bit_exact_and_checksum(c2):
    g, n, m = c2
    _gebv_block_check(g, n, m, 123_457, 2_000, seed=1)
    # checksum over all SNPs: sum_i sum_j alpha_j z_ij = sum_j alpha_j popcnt_j
    pc = g.counts()[0]
    alpha = np.random.default_rng(2).normal(size=m)
    got = g.gebv(torch.from_numpy(alpha).cuda()).cpu().numpy()
    want = float(np.dot(alpha, pc.astype(np.float64)))
    assert abs(got.sum() - want) <= 1e-9 * np.abs(alpha).dot(pc)
    # the one-pass GEBV over all 500,000 SNPs, with flips, against the exact FP64 scores
    flip = (np.random.default_rng(3).random(m) < 0.3).astype(np.uint8)
    a = torch.from_numpy(alpha).cuda()
    exact = g.gebv(a, flip).cpu().numpy()
    fast = g.gebv_fast(a, flip).cpu().numpy()
    err = float(np.max(np.abs(fast - exact)) / np.max(np.abs(exact)))
    print(f"C2 fast GEBV: max |fast - exact| / max |exact| = {err:.2e}")
    assert err < 2e-6


def test_c2_operator_k65_symmetry_and_fp32_agreement(c2):
    """A = R^-1/2 Zc W Zc' R^-1/2 + tau0 I on 64 Rademacher probes + 1 RHS: u'(Av) = v'(Au), and
    the split-BF16 operator within 5e-5 (relative Frobenius) of the FP32 kernel's (measured 7e-6;
    element-wise the FP32 path is the less accurate of the two at this n, see the next test)."""
    g, n, m = c2
    gen = torch.Generator(device="cuda").manual_seed(3)
    V = torch.where(torch.rand((n, 65), generator=gen, device="cuda") < 0.5, -1.0, 1.0) / np.sqrt(n)
    V[:, 0] = torch.randn(n, generator=gen, device="cuda")
    tau0 = 0.4286
    A16 = g.lnr_op_select("ldet", tau0, V, precision="bf16x2").double()
    A32 = g.lnr_op_select("ldet", tau0, V, precision="fp32").double()
    assert ((A16 - A32).norm() / A32.norm()).item() < 5e-5
    Vd = V.double()
    M = Vd.t() @ A16  # (u_k' A v_l), scaled by the Cauchy-Schwarz bound sqrt(u'Au v'Av)
    d = torch.diagonal(M)
    assert torch.all(d > 0)
    asym = ((M - M.t()).abs() / torch.sqrt(d[:, None] * d[None, :])).max().item()
    assert asym < 1e-4, asym


def test_c2_pass1_block_against_fp64(c2):
    """Zc'V at K = 65 for a block of 300 SNPs over all 100,000 individuals against FP64 on the
    host.  The default path (int8 digits rendered from the FP32 operand, exact int32 sums) to
    FP32 level: 2e-5 max-relative per column, 1e-5 relative Frobenius (with digits rendered from
    the BF16 hi + lo pair it was 1.2e-4 / 3.8e-6); the plain FP32 kernel's sums of 100,000 terms
    drift further (measured 1.2e-3 max-relative, 1.0e-5 Frobenius), bounded here at 5e-3 / 3e-5."""
    g, n, m = c2
    gen = torch.Generator(device="cuda").manual_seed(3)
    V = torch.where(torch.rand((n, 65), generator=gen, device="cuda") < 0.5, -1.0, 1.0) / np.sqrt(n)
    V[:, 0] = torch.randn(n, generator=gen, device="cuda")
    j0, cnt = 311_007, 300
    codes = bedio.decode_codes(g.download(j0, cnt), n)
    Z = ((codes & 1) + (codes >> 1)).astype(np.float64)
    gm = g.counts()[4][j0:j0 + cnt].astype(np.float64)
    want = (Z - gm[:, None]) @ V.double().cpu().numpy()
    # default path: int8 digits of the FP32 operand with exact int32 sums; plain FP32 MFMA path
    for prec, mx, fr in (("bf16x2", 2e-5, 1e-5), ("fp32", 5e-3, 3e-5)):
        got = g.rmatvec(V, precision=prec).double().cpu().numpy()[j0:j0 + cnt]
        e = np.abs(got - want)
        col = np.max(e.max(0) / np.abs(want).max(0))
        frob = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"C2 Zc'V {prec}: max-relative per column {col:.2e}, relative Frobenius {frob:.2e}")
        assert col < mx, (prec, col)
        assert frob < fr, (prec, frob)


def test_c2_centring_identity(c2):
    """Zc'1 = sum_i (z_ij - gm_j) = popcnt_j - n gm_j with the float gm (src/reml.cpp:182): the
    exact-centring epilogue reproduces it for a constant column in both precisions."""
    g, n, m = c2
    pc, _, _, _, gm = g.counts()
    want = pc.astype(np.float64) - n * gm.astype(np.float64)
    one = torch.ones((n, 1), device="cuda")
    for prec in ("bf16x2", "fp32"):
        z = g.rmatvec(one, precision=prec).double().cpu().numpy()[:, 0]
        assert np.max(np.abs(z - want) - 1e-6 * np.abs(want)) < 1e-4, prec


def test_c2_gwas_score_block(c2):
    g, n, m = c2
    rng = np.random.default_rng(4)
    Py = rng.normal(size=n).astype(np.float32)
    xy, xx = g.gwas_score(torch.from_numpy(Py).cuda(), skip_mac=0.0)
    xy = xy.double().cpu().numpy()
    xx = xx.double().cpu().numpy()
    j0, cnt = 77_777, 200
    codes = bedio.decode_codes(g.download(j0, cnt), n)
    Z = ((codes & 1) + (codes >> 1)).astype(np.float64)
    gm = g.counts()[4][j0:j0 + cnt].astype(np.float64)
    Xc = Z.T - gm[None, :]
    wxy = Xc.T @ Py.astype(np.float64)
    Xp = Xc - Xc.mean(axis=0, keepdims=True)  # Q = 1/sqrt(n): the projection removes the mean
    wxx = np.sum(Xp * Xp, axis=0)
    assert np.max(np.abs(xy[j0:j0 + cnt] - wxy)) <= 3e-4 * np.max(np.abs(wxy))
    assert np.max(np.abs(xx[j0:j0 + cnt] - wxx) / wxx) < 1e-4


def test_c5_gebv_block_bit_exact(gpu):
    n, m = 1_000_000, 60_000
    g = gpu.Geno.synthesize(n, m, seed=7, maf_lo=0.01, maf_hi=0.5, missing=0.0)
    try:
        _gebv_block_check(g, n, m, 40_001, 300, seed=5)
        alpha = np.random.default_rng(6).normal(size=m)
        a = torch.from_numpy(alpha).cuda()
        exact = g.gebv(a).cpu().numpy()
        fast = g.gebv_fast(a).cpu().numpy()
        err = float(np.max(np.abs(fast - exact)) / np.max(np.abs(exact)))
        print(f"C5 fast GEBV: max |fast - exact| / max |exact| = {err:.2e}")
        assert err < 2e-6
    finally:
        g.close()


def test_squared_pass_splits_past_4m_individuals(gpu, monkeypatch):
    """GWAS score terms with error-variance weights run a squared-dosage pass (dosages up to 4)
    whose reduction runs over individuals.  Its int32 sums stay exact only over at most 2^22 rows
    per reduction split, so at n = 4.3M individuals the split count is raised to at least 2 even
    when the workgroup target asks for one split (SLEMM_I8_TARGET=1); the launch refuses any
    longer split.  64 SNPs against FP64 on the host."""
    n, m = 4_300_000, 64
    g = gpu.Geno.synthesize(n, m, seed=11, maf_lo=0.05, maf_hi=0.5, missing=0.0)
    try:
        rng = np.random.default_rng(6)
        rvec = rng.uniform(0.5, 2.0, n).astype(np.float32)
        rr = (1.0 / np.sqrt(rvec.astype(np.float64))).astype(np.float32)
        q = (rr.astype(np.float64) / np.linalg.norm(rr.astype(np.float64))).astype(np.float32)
        g.set_model(None, rvec, q[:, None])
        _, _, rr_dev = g.model()
        Py = rng.normal(size=n).astype(np.float32)
        monkeypatch.setenv("SLEMM_I8_TARGET", "1")
        xy, xx = g.gwas_score(torch.from_numpy(Py).cuda(), skip_mac=0.0)
        xy, xx = xy.double().cpu().numpy(), xx.double().cpu().numpy()
        codes = bedio.decode_codes(g.download(0, m), n)
        gm = g.counts()[4].astype(np.float64)
        r = rr_dev.astype(np.float64)
        qd = q.astype(np.float64)
        wxy, wxx = np.zeros(m), np.zeros(m)
        for j in range(m):
            xc = ((codes[j] & 1) + (codes[j] >> 1)).astype(np.float64) - gm[j]
            wxy[j] = xc @ Py.astype(np.float64)
            v = r * xc
            wxx[j] = v @ v - (qd @ v) ** 2
        assert np.max(np.abs(xy - wxy)) <= 3e-4 * np.max(np.abs(wxy))
        assert np.max(np.abs(xx - wxx) / wxx) < 1e-4, np.max(np.abs(xx - wxx) / wxx)
    finally:
        g.close()


def test_c5_sldf_compacted_seed_equals_full_block(gpu, monkeypatch):
    """SLDF-REML at n = 1,000,000 x m = 60,000 with 100 probes: the block holds y, the intercept
    and the probes (102 columns); the intercept's system stops after a couple of steps, after
    which the seed runs the operator on the 101 columns still running (19 int8 digit tiles in
    one launch instead of 21 in two), and on y alone once the probes stop.  The fit equals the
    one that keeps all 102 columns in the block (SLEMM_LSEED_COMPACT=0) to the rounding of the
    reduction splits."""
    n, m = 1_000_000, 60_000
    g = gpu.Geno.synthesize(n, m, seed=5, maf_lo=0.01, maf_hi=0.5, missing=0.0)
    try:
        g.set_model()
        rng = np.random.default_rng(8)
        beta = np.zeros(m, np.float32)
        beta[rng.choice(m, size=2000, replace=False)] = rng.normal(size=2000)
        gv = g.matvec(torch.from_numpy(beta[:, None]).cuda(), precision="fp32")[:, 0].double().cpu().numpy()
        ve = gv.var() * (1 - 0.4) / 0.4
        y = (gv + rng.normal(scale=np.sqrt(ve), size=n) + 5.0).astype(np.float32)
        X = np.ones((n, 1), np.float32)
        fits = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("SLEMM_LSEED_COMPACT", mode)
            fits[mode] = g.SLDF_REML(X, y, rng_seed=0, n_V=100)
        a, b = fits["1"], fits["0"]
        print("steps", a["steps"], "h2", a["h2"], b["h2"])
        assert a["steps"] == b["steps"] and a["steps"][1] < min(a["steps"][0], a["steps"][2])
        for k in ("vg", "ve", "h2"):
            assert abs(a[k] - b[k]) <= 1e-5 * abs(b[k]), (k, a[k], b[k])
        assert np.max(np.abs(a["snp_blup"] - b["snp_blup"])) <= 1e-4 * np.max(np.abs(b["snp_blup"]))
    finally:
        g.close()
