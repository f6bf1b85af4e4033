"""GPU parity of the genotype engine against the oracle (bit-exact integer work) and the FP64
dense oracle (floating point).  Every call goes through the C ABI of libslemm_gpu.so."""
import numpy as np
import pytest
import torch

import oracle
from oracle import bedio
from tests.problems import make_problem, rel_err

pytestmark = pytest.mark.gpu

# relative-error bounds (max |err| / max |ref|) per precision, stated against FP64 dense
TOL = {"fp32": 2e-5, "bf16x2": 3e-4}
# int8 block products: digits from the FP32 operand (23-bit fixed point per column) and exact int32
# sums, against the FP64 dense result -- FP32-level accuracy
I8_TOL = 2e-5


def _geno(gpu, p, keep=None, snps=None, n_raw=None):
    return gpu.Geno.from_bed(p.rows, p.n if n_raw is None else n_raw, keep, snps)


def test_decode_bit_exact_with_keep_and_order(gpu):
    p = make_problem(n=1037, m=611, seed=3, missing=0.02)
    rng = np.random.default_rng(7)
    keep = np.sort(rng.choice(p.n, size=801, replace=False))
    order = rng.permutation(p.m)[:517]
    g = _geno(gpu, p, keep, order)
    got = bedio.decode_codes(g.download(), len(keep))
    bindi = np.zeros(p.n, np.uint8)
    bindi[keep] = 1
    kmat = oracle.decode_bed(bedio.bed_bytes(p.rows), p.n, p.m, bindi, None, order)
    want = (kmat[:, 0::2] | (kmat[:, 1::2] << 1)).astype(np.uint8)
    assert np.array_equal(got, want)
    # padding codes beyond n are zero
    full = g.download()
    assert np.all((full[:, -1] >> (2 * (len(keep) % 4))) == 0) if len(keep) % 4 else True


@pytest.mark.parametrize("kind", ["all", "prefix", "half", "sparse", "mixed", "one", "last"])
def test_device_keep_bit_exact(gpu, kind):
    """Individuals kept on the device (handle to handle, LDS-staged chunks of 4,096 kept
    individuals, a global gather where they are sparser than one in eight) against the oracle's
    decode of the same bed; the transposed copy through the product with the host-compacted
    handle (exact in FP32: integer dosages)."""
    # the sparse case spans more than 32 KiB of source bytes per chunk: the global gather
    p = make_problem(n=140001 if kind == "sparse" else 9001, m=64 if kind == "sparse" else 300, seed=17,
                     missing=0.02)
    rng = np.random.default_rng(11)
    keep = {"all": np.arange(p.n), "prefix": np.arange(5000),
            "half": np.sort(rng.choice(p.n, 4500, replace=False)),
            "sparse": np.sort(rng.choice(p.n, 3000, replace=False)),
            "mixed": np.concatenate([np.arange(0, 4200), np.arange(4200, p.n, 25)]),
            "one": np.array([4097]), "last": np.arange(p.n - 3, p.n)}[kind]
    g = _geno(gpu, p)
    gk = g.keep(keep)
    want = bedio.decode_codes(p.rows, p.n)[:, keep]
    assert np.array_equal(bedio.decode_codes(gk.download(), len(keep)), want)
    raw = gk.download()
    if len(keep) % 4:
        assert np.all((raw[:, -1] >> (2 * (len(keep) % 4))) == 0)
    gh = _geno(gpu, p, keep)
    assert all(np.array_equal(a, b) for a, b in zip(gk.counts(), gh.counts()))
    if len(keep) < 100:  # too few individuals for a model without monomorphic SNPs
        return
    gk.set_model(), gh.set_model()
    U = torch.from_numpy(rng.integers(-3, 4, size=(p.m, 2)).astype(np.float32)).cuda()
    assert np.array_equal(gk.matvec(U, precision="fp32").cpu().numpy(), gh.matvec(U, precision="fp32").cpu().numpy())


def test_counts_bit_exact(gpu):
    p = make_problem(n=2051, m=733, seed=5, missing=0.03)
    g = _geno(gpu, p)
    pc, n0, n1, n2, gm = g.counts()
    opc, on0, on1, on2, ogm = oracle.counts(p.kmat)
    assert np.array_equal(pc, opc) and np.array_equal(n0, on0)
    assert np.array_equal(n1, on1) and np.array_equal(n2, on2)
    assert np.array_equal(gm.view(np.uint32), ogm.view(np.uint32))  # IEEE f32 divide, bit-exact


def test_model_terms_bit_exact(gpu):
    p = make_problem(n=600, m=500, seed=9, weights=True)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    gm, gw, rr = g.model()
    assert np.array_equal(gw.view(np.uint32), p.gw.view(np.uint32))
    assert np.array_equal(rr, (1.0 / np.sqrt(p.rvec.astype(np.float64))).astype(np.float32)) or rel_err(rr, p.rr) < 1e-7


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
@pytest.mark.parametrize("t", [1, 7, 33, 66])
def test_op_ldet_and_quadform(gpu, prec, t):
    p = make_problem(n=913, m=1201, seed=11, weights=True, c=2)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    D = p.dense()
    rng = np.random.default_rng(t)
    V = rng.normal(size=(p.n, t)).astype(np.float32)
    Vd = torch.from_numpy(V).cuda()
    tau0 = 0.4286
    for kind in ("ldet", "quadform"):
        out = g.lnr_op_select(kind, tau0, Vd, precision=prec).cpu().numpy()
        ref = D.op(kind, tau0, V)
        assert rel_err(out, ref) < TOL[prec], (kind, prec, t, rel_err(out, ref))
    # mixed batch: one pass serves both kinds
    kinds = ["quadform" if k % 3 == 0 else "ldet" for k in range(t)]
    out = g.lnr_op_select("ldet", tau0, Vd, precision=prec, col_kind=kinds).cpu().numpy()
    for k, kd in enumerate(kinds):
        ref = D.op(kd, tau0, V[:, k:k + 1])[:, 0]
        assert rel_err(out[:, k], ref) < TOL[prec]


BP_VARIANTS = 10  # block-product kernel variants selectable with SLEMM_BP_VARIANT


@pytest.mark.parametrize("t", [1, 3, 5, 12, 16, 33, 65, 66, 100, 161])
def test_block_product_variants(gpu, t, monkeypatch):
    """Every split-BF16 block-product variant at one reduction split and at several, for both
    split layouts (hi and lo halves of narrow blocks or of the last tile as columns, or as two
    blocks), against the FP64 dense operator and the transposed product."""
    p = make_problem(n=2600, m=1700, seed=19, weights=True, c=2)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    D = p.dense()
    rng = np.random.default_rng(100 + t)
    V = rng.normal(size=(p.n, t)).astype(np.float32)
    Vd = torch.from_numpy(V).cuda()
    tau0 = 0.37
    ref = D.op("ldet", tau0, V)
    zref = D.rmatvec(V, weighted=True)
    monkeypatch.setenv("SLEMM_I8_NARROW", "0")  # all blocks on the split-BF16 kernels
    monkeypatch.setenv("SLEMM_I8_WIDE", "0")
    for hl in ("1", "0"):
        monkeypatch.setenv("SLEMM_HL_COLUMNS", hl)
        for v in range(BP_VARIANTS):
            for target in ("1", "2048"):
                monkeypatch.setenv("SLEMM_BP_VARIANT", str(v))
                monkeypatch.setenv("SLEMM_BP_TARGET", target)
                out = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
                assert rel_err(out, ref) < TOL["bf16x2"], (hl, v, target, rel_err(out, ref))
                zt = g.weighted_rmatvec(Vd, precision="bf16x2").cpu().numpy()
                assert rel_err(zt, zref) < TOL["bf16x2"], (hl, v, target, rel_err(zt, zref))


@pytest.mark.parametrize("t", [1, 2, 3, 4, 5, 6])
def test_int8_narrow_blocks(gpu, t, monkeypatch):
    """Narrow blocks (1 to 5 vectors) on the int8 kernels, at one reduction split and at several
    (including an odd number of super-chunks per split), against the FP64 dense operator, the
    transposed product and the product, and against the split-BF16 kernels on the same operand;
    a zero column and a column whose largest magnitude is a power of two exercise the scale
    choice.  t = 6 takes the split-BF16 path either way."""
    p = make_problem(n=2600, m=1700, seed=23, weights=True, c=2)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    D = p.dense()
    rng = np.random.default_rng(200 + t)
    V = rng.normal(size=(p.n, t)).astype(np.float32)
    if t >= 3:
        V[:, 1] = 0.0
        V[:, 2] *= 4.0 / np.abs(V[:, 2]).max()
    U = rng.normal(size=(p.m, t)).astype(np.float32)
    Vd, Ud = torch.from_numpy(V).cuda(), torch.from_numpy(U).cuda()
    tau0 = 0.37
    ref = D.op("ldet", tau0, V)
    zref = D.rmatvec(V, weighted=True)
    uref = D.matvec(U)
    monkeypatch.setenv("SLEMM_I8_NARROW", "0")
    base = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
    monkeypatch.setenv("SLEMM_I8_NARROW", "1")
    for target in ("1", "6", "100000"):
        monkeypatch.setenv("SLEMM_I8_TARGET", target)
        out = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
        # digits rendered from the FP32 operand (23 bits) and exact int32 sums: FP32-level error
        tol = I8_TOL if t <= 5 else TOL["bf16x2"]
        assert rel_err(out, ref) < tol, (target, rel_err(out, ref))
        assert rel_err(out, base) < TOL["bf16x2"], (target, rel_err(out, base))
        zt = g.weighted_rmatvec(Vd, precision="bf16x2").cpu().numpy()
        assert rel_err(zt, zref) < tol, (target, rel_err(zt, zref))
        if t >= 3:
            assert np.all(np.abs(zt[:, 1]) <= 1e-6 * np.abs(zt).max())
        z = g.matvec(Ud, precision="bf16x2").cpu().numpy()
        assert rel_err(z, uref) < tol, (target, rel_err(z, uref))


@pytest.mark.parametrize("t", [7, 9, 16, 17, 21, 33, 40, 51, 65, 66, 100, 102])
def test_int8_wide_blocks(gpu, t, monkeypatch):
    """Blocks of more than 5 vectors on the int8 wide kernel (16-vector tiles of three digit
    tiles, a tail tile for up to 5 trailing vectors, column groups of two vector tiles; an odd
    number of vector tiles -- t = 17, 21, 40, 51, 102 -- runs the pairs and the last tile with the
    tail as two launches), over the operand layouts the split-BF16 path uses (tail columns for
    t = 16 q + r with r <= 8, separate hi and lo blocks otherwise; 6 to 8 vectors take the wide
    kernel only without hi/lo columns), at one reduction split and at several, against the FP64
    dense operator, the transposed product and the product, against the split-BF16 kernels, and
    bit for bit against the first form of the wide kernel (exact integer sums either way)."""
    if t <= 8:
        monkeypatch.setenv("SLEMM_HL_COLUMNS", "0")
    p = make_problem(n=2600, m=1700, seed=31, weights=True, c=2)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    D = p.dense()
    rng = np.random.default_rng(300 + t)
    V = rng.normal(size=(p.n, t)).astype(np.float32)
    V[:, 3] = 0.0
    V[:, 4] *= 1e-3
    V[:, 5] *= 1e-33  # its own scale 2^-109: the digit rendering must not overflow
    U = rng.normal(size=(p.m, t)).astype(np.float32)
    Vd, Ud = torch.from_numpy(V).cuda(), torch.from_numpy(U).cuda()
    tau0 = 0.37
    ref = D.op("ldet", tau0, V)
    zref = D.rmatvec(V, weighted=True)
    uref = D.matvec(U)
    monkeypatch.setenv("SLEMM_I8_WIDE", "0")
    base = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
    monkeypatch.setenv("SLEMM_I8_WIDE", "1")
    for target in ("1", "20", "100000"):
        monkeypatch.setenv("SLEMM_I8_TARGET", target)
        out = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
        monkeypatch.setenv("SLEMM_I8_WIDE_FORM", "1")
        first = g.lnr_op_select("ldet", tau0, Vd, precision="bf16x2").cpu().numpy()
        monkeypatch.delenv("SLEMM_I8_WIDE_FORM")
        if t > 8:
            assert np.array_equal(out, first), (target, rel_err(out, first))
        tol = I8_TOL
        for k in range(t):  # every column against its own scale (column 5 is 1e-33 of the others)
            assert rel_err(out[:, k], ref[:, k]) < tol, (target, k, rel_err(out[:, k], ref[:, k]))
        assert rel_err(out, base) < TOL["bf16x2"], (target, rel_err(out, base))
        zt = g.weighted_rmatvec(Vd, precision="bf16x2").cpu().numpy()
        assert rel_err(zt, zref) < tol, (target, rel_err(zt, zref))
        assert np.all(zt[:, 3] == 0.0)
        assert rel_err(zt[:, 4], zref[:, 4]) < tol
        assert rel_err(zt[:, 5], zref[:, 5]) < tol, rel_err(zt[:, 5], zref[:, 5])
        z = g.matvec(Ud, precision="bf16x2").cpu().numpy()
        assert rel_err(z, uref) < tol, (target, rel_err(z, uref))


def test_block_kernel_selection(gpu, monkeypatch):
    """The kernel family per block width: int8 narrow for 1-5 vectors, split-BF16 for 6-8 (hi and lo
    halves as columns), the one-tile int8 kernel for 9-16, the general int8 wide kernel beyond, FP32
    kernels at FP32 precision; the switches move blocks back to the split-BF16 kernels."""
    p = make_problem(n=700, m=300, seed=3)
    g = _geno(gpu, p)
    want = {1: "int8 narrow", 5: "int8 narrow", 6: "bf16x2", 8: "bf16x2", 9: "int8 one tile",
            16: "int8 one tile", 17: "int8 wide", 65: "int8 wide", 101: "int8 wide"}
    for t, k in want.items():
        assert g.block_kernel(t) == k, t
        assert g.block_kernel(t, "fp32") == "fp32", t
    monkeypatch.setenv("SLEMM_I8_NARROW", "0")
    monkeypatch.setenv("SLEMM_I8_WIDE", "0")
    assert {g.block_kernel(t) for t in want} == {"bf16x2"}


@pytest.mark.parametrize("t", [3, 16, 33])
def test_int8_autotune(gpu, t):
    """The int8 autotuner times every split target (and the narrow kernel's row tiles) per pass,
    keeps the fastest on the handle, and the operator on the tuned settings still matches the FP64
    dense operator to the int8 bound."""
    p = make_problem(n=2600, m=1700, seed=41, weights=True, c=1)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    D = p.dense()
    at = g.autotune(t, reps=2)
    assert at["kernel"] == g.block_kernel(t)
    times = [v for v in at["table_ms"].values()]
    assert len(times) == (20 if t <= 5 else 10) and min(times) > 0, at
    for pas in ("pass1", "pass2"):
        name, target, ms = at[pas]
        assert target in (1024, 2048, 4096, 8192, 16384), at
        key = f"{pas}/" + (name.split()[-1] + "/" if t <= 5 else "") + f"{target}"
        assert at["table_ms"][key] == ms == min(v for k, v in at["table_ms"].items() if k.startswith(pas)), at
    rng = np.random.default_rng(400 + t)
    V = rng.normal(size=(p.n, t)).astype(np.float32)
    U = rng.normal(size=(p.m, t)).astype(np.float32)
    out = g.lnr_op_select("ldet", 0.37, torch.from_numpy(V).cuda(), precision="bf16x2").cpu().numpy()
    assert rel_err(out, D.op("ldet", 0.37, V)) < I8_TOL
    z = g.matvec(torch.from_numpy(U).cuda(), precision="bf16x2").cpu().numpy()
    assert rel_err(z, D.matvec(U)) < I8_TOL


@pytest.mark.parametrize("t", [1, 3, 17])
def test_int8_nonfinite_column(gpu, t, monkeypatch):
    """A NaN or infinite element makes every product of its column NaN on the int8 kernels, as on
    the split-BF16 ones; the other columns are unaffected."""
    p = make_problem(n=1100, m=900, seed=37)
    g = _geno(gpu, p)
    g.set_model(None, None, p.Q)
    V = np.random.default_rng(5).normal(size=(p.n, t)).astype(np.float32)
    V[17, 0] = np.nan
    if t > 1:
        V[3, t - 1] = np.inf
    Vd = torch.from_numpy(V).cuda()
    out = {}
    for i8 in ("1", "0"):
        monkeypatch.setenv("SLEMM_I8_NARROW", i8)
        monkeypatch.setenv("SLEMM_I8_WIDE", i8)
        out[i8] = g.rmatvec(Vd, precision="bf16x2").cpu().numpy()
    bad = [0] + ([t - 1] if t > 1 else [])
    good = [k for k in range(t) if k not in bad]
    for i8 in out:
        assert np.all(np.isnan(out[i8][:, bad])), i8
        assert np.all(np.isfinite(out[i8][:, good])), i8
    if good:
        assert rel_err(out["1"][:, good], out["0"][:, good]) < TOL["bf16x2"]


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_rmatvec_matvec(gpu, prec):
    p = make_problem(n=1500, m=2100, seed=13)
    g = _geno(gpu, p)
    g.set_model(None, None, p.Q)
    D = p.dense()
    rng = np.random.default_rng(2)
    V = rng.normal(size=(p.n, 5)).astype(np.float32)
    U = rng.normal(size=(p.m, 3)).astype(np.float32)
    zt = g.rmatvec(torch.from_numpy(V).cuda(), precision=prec).cpu().numpy()
    assert rel_err(zt, D.rmatvec(V)) < TOL[prec]
    wzt = g.weighted_rmatvec(torch.from_numpy(V).cuda(), precision=prec).cpu().numpy()
    assert rel_err(wzt, D.rmatvec(V, weighted=True)) < TOL[prec]
    z = g.matvec(torch.from_numpy(U).cuda(), precision=prec).cpu().numpy()
    assert rel_err(z, D.matvec(U)) < TOL[prec]


def test_centring_of_constant_vector(gpu):
    """Zc' 1 = popcnt - n gm is zero up to the rounding of gm: the rank-1 centring epilogue must
    not lose it to cancellation (the intercept column of Q hits exactly this case)."""
    p = make_problem(n=4099, m=300, seed=17)
    g = _geno(gpu, p)
    g.set_model(None, None, p.Q)
    one = torch.full((p.n, 1), 1.0 / np.sqrt(p.n), dtype=torch.float32, device="cuda")
    D = p.dense()
    ref = D.rmatvec(one.cpu().numpy())
    for prec in ("fp32", "bf16x2"):
        got = g.rmatvec(one, precision=prec).cpu().numpy()
        # the column is removed as its mean and re-enters only through the exact term
        # xbar (popcnt - n gm): the result is as accurate as that FP64 product
        assert np.max(np.abs(got - ref)) < 1e-6 + 1e-4 * np.max(np.abs(ref)), prec


def test_synthesized_bed_consistent(gpu):
    g = gpu.Geno.synthesize(3000, 700, seed=5, missing=0.01)
    rows = g.download()
    codes = bedio.decode_codes(rows, 3000)
    kmat = bedio.kmat_of_codes(codes)
    pc, n0, n1, n2, gm = g.counts()
    opc, on0, on1, on2, ogm = oracle.counts(kmat)
    assert np.array_equal(pc, opc) and np.array_equal(n2, on2) and np.array_equal(n1, on1)
    f = pc / (2 * 3000.0)
    assert 0.04 < f.mean() < 0.3
    miss = np.mean(codes == 1)
    assert 0.005 < miss < 0.02
    # the transposed copy agrees with the SNP-major one: matvec vs dense
    g.set_model()
    from oracle.dense import Dense, dosage_of_kmat
    D = Dense(dosage_of_kmat(kmat), ogm, oracle.gw(ogm), np.ones(3000))
    U = np.random.default_rng(0).normal(size=(700, 2)).astype(np.float32)
    z = g.matvec(torch.from_numpy(U).cuda(), precision="fp32").cpu().numpy()
    assert rel_err(z, D.matvec(U)) < TOL["fp32"]


def test_synthesized_bed_given_frequencies(gpu):
    """Per-SNP allele frequencies passed in: dosage frequency within 5 sd of p, codes 00/10/11
    only (no missing), and the device counts equal the oracle's on the downloaded rows."""
    n, m = 20000, 400
    p = np.random.default_rng(8).beta(0.8, 0.8, size=m).astype(np.float32)
    p[:3] = [0.0, 1.0, 0.5]
    g = gpu.Geno.synthesize_p(n, p, seed=3)
    codes = bedio.decode_codes(g.download(), n)
    assert not np.any(codes == 1)
    pc, n0, n1, n2, gm = g.counts()
    opc, *_ = oracle.counts(bedio.kmat_of_codes(codes))
    assert np.array_equal(pc, opc)
    f = pc / (2.0 * n)
    sd = np.sqrt(p * (1 - p) / (2 * n)) + 1e-9
    assert np.all(np.abs(f - p) <= 5 * sd + 1e-6)
    assert f[0] == 0.0 and f[1] == 1.0


@pytest.mark.parametrize("n,m,flips", [(3001, 800, True), (257, 1, False), (1024, 1537, False), (700, 2048, True)])
def test_gebv_bit_exact(gpu, n, m, flips):
    """FP64 GEBV in the reference's SNP order, compared bit for bit (signed zeros included):
    alphas with exact cancellations, -0.0 and a wide exponent range; flips on a third of the SNPs."""
    p = make_problem(n=n, m=m, seed=19 + m, missing=0.01)
    rng = np.random.default_rng(4)
    alpha = rng.normal(size=p.m) * np.exp2(rng.integers(-30, 30, size=p.m))
    alpha[1::7] = -alpha[0:-1:7][: alpha[1::7].size]
    alpha[3::11] = -0.0
    flip = (rng.random(p.m) < 0.3).astype(np.uint8) if flips else None
    g = _geno(gpu, p)
    got = g.gebv(torch.from_numpy(alpha).cuda(), flip).cpu().numpy()
    fl = flip if flips else np.zeros(p.m, np.uint8)
    want = oracle.gebv(bedio.bed_bytes(p.rows), p.n, p.m, np.ones(p.m, np.uint8), fl, alpha)
    assert np.array_equal(got.view(np.uint64), np.asarray(want, np.float64).view(np.uint64))


@pytest.mark.parametrize("n,m,flips", [(3001, 800, True), (257, 1, False), (1024, 1537, False), (700, 2048, True)])
def test_gebv_fast_against_exact(gpu, n, m, flips, monkeypatch):
    """The one-pass GEBV against the exact FP64 sum of the reference's order, with flipped SNPs and
    missing codes: FP32 accuracy on the int8 narrow kernel (bound 2e-6 of the largest score), the
    split-BF16 operand's 16 bits on the BF16 kernels (3e-5)."""
    p = make_problem(n=n, m=m, seed=11 + m, missing=0.03)
    g = _geno(gpu, p)
    rng = np.random.default_rng(m)
    alpha = rng.normal(size=m)
    flip = (rng.random(m) < 0.3).astype(np.uint8) if flips else None
    codes = bedio.decode_codes(p.rows, p.n).astype(np.int64)
    z = (codes & 1) + (codes >> 1)
    if flips:
        z = np.where(flip[:, None] == 1, 2 - z, z)
    want = (alpha[:, None] * z).sum(axis=0)
    a = torch.from_numpy(alpha).cuda()
    exact = g.gebv(a, flip).cpu().numpy()
    assert np.allclose(exact, want, rtol=0, atol=1e-9 * np.abs(want).max())
    fast = g.gebv_fast(a, flip).cpu().numpy()
    assert rel_err(fast, exact) < 2e-6, rel_err(fast, exact)
    monkeypatch.setenv("SLEMM_I8_NARROW", "0")
    fast_bf = g.gebv_fast(a, flip).cpu().numpy()
    assert rel_err(fast_bf, exact) < 3e-5, rel_err(fast_bf, exact)


@pytest.mark.parametrize("weights", [False, True])
def test_gwas_score_terms(gpu, weights):
    """xy = x_c'Py and xx = ||(I-QQ')R^-1/2 x_c||^2 per SNP (bin/slemm_gamma.py:306-319)."""
    from oracle import gwas
    p = make_problem(n=2311, m=1400, seed=29, weights=weights, c=2, missing=0.01)
    g = _geno(gpu, p)
    g.set_model(None, p.rvec, p.Q)
    Py = np.random.default_rng(3).normal(size=p.n).astype(np.float32)
    mac = 20.0
    xy_ref, xx_ref = gwas.gamma_scores(p.Z, Py, p.rr, p.Q, mac)
    from slemm_amd import _lib as L
    import ctypes as C
    dPy = torch.from_numpy(Py).cuda()
    xy = torch.empty(p.m, dtype=torch.float32, device="cuda")
    xx = torch.empty(p.m, dtype=torch.float32, device="cuda")
    L.check(L.load().slemm_gpu_gwas_score(g.handle, C.c_void_p(dPy.data_ptr()), mac, C.c_void_p(xy.data_ptr()),
                                          C.c_void_p(xx.data_ptr()), None))
    xy, xx = xy.cpu().numpy(), xx.cpu().numpy()
    keep = ~np.isnan(xy_ref)
    assert np.array_equal(np.isnan(xy), ~keep)
    assert rel_err(xy[keep], xy_ref[keep]) < 3e-4
    assert np.max(np.abs(xx[keep] - xx_ref[keep]) / xx_ref[keep]) < 1e-4
