"""Host statistics of the association layer (slemm_amd/gwas.py) against the oracle and the
reference's example outputs (examples/expected/slemm/10k.gstat.txt, 10k.gamma.txt)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import gwas as ogwas
from slemm_amd import gwas

GOLDEN = Path(__file__).resolve().parent / "golden"


def _meta():
    return json.loads((GOLDEN / "ex10k_meta.json").read_text())


def test_trimmed_mean_matches_restatement_with_outliers():
    rng = np.random.default_rng(0)
    for size in (3, 30, 31, 300):
        x = rng.normal(1.0, 0.1, size)
        x[: max(1, size // 10)] *= 25.0  # outliers above the upper fence
        assert abs(gwas.mean_wo_outliers(x) - ogwas.mean_wo_outliers(x)) < 1e-12


def test_example_gstat_columns_and_correction_factor():
    meta = _meta()
    rows = meta["gstat"]
    gst = {k: np.array([float(r[k]) for r in rows], np.float32) for k in ("xPy", "xPx", "xBx", "xSx")}
    tab = gwas.gstat_table(gst, meta["Vg"])
    want = np.array([float(r["ChiSq"]) for r in rows])
    assert np.max(np.abs(tab["ChiSq"] - want) / want) < 2e-5
    assert np.max(np.abs(tab["Pval"] - np.array([float(r["Pval"]) for r in rows]))) < 1e-5
    cf = gwas.correction_factor(gst)
    ratio = gst["xSx"].astype(np.float64) / gst["xPx"].astype(np.float64)
    assert abs(cf - ogwas.mean_wo_outliers(ratio)) < 1e-12
    # the factor lies inside the range of the ratios, and the window form is at least the
    # whole-genome form on every example row (a window explains less than all SNPs)
    assert ratio.min() <= cf <= ratio.max()
    assert np.all(gst["xBx"] >= gst["xPx"])


def test_example_gamma_pvalues():
    meta = _meta()
    hdr = meta["gamma_header"]
    ic, ip = hdr.index("CHISQ"), hdr.index("P")
    chisq = np.array([float(r[ic]) for r in meta["gamma_rows"]])
    pv = np.array([float(r[ip]) for r in meta["gamma_rows"]])
    assert np.max(np.abs(gwas.chisq_pval_1df(chisq) - pv)) < 1e-6
    assert gwas.chisq_pval_1df(np.array([0.0, -1.0]))[0] == 1.0


def test_gstat_snp_sampler_respects_mac():
    gm = np.array([0.001, 0.5, 1.0, 1.99, 0.02, 1.5], np.float32)
    s = gwas.sample_gstat_snps(gm, 1000, 10)
    assert set(s.tolist()) == {1, 2, 5} and math.isclose(gm[4] * 1000, 20.0)


def test_oracle_glibc_rand_matches_libc():
    """The oracle's TYPE_3 restatement against this C library's own rand() after srand(1)."""
    import ctypes
    from oracle import gwas as ogwas
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = [libc.rand() for _ in range(5000)]
    assert ogwas.glibc_rand(5000) == want


@pytest.fixture(scope="module")
def shuffle_prog(tmp_path_factory):
    """std::random_shuffle compiled here with g++ (libstdc++ 11): no srand, as in the reference.
    Reads N and N floats gm (one per line) and n, prints the reference's sample order: the
    filter of src/reml.cpp:922-923 and the shuffle of :924."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("shuffle")
    src = d / "shuffle.cpp"
    src.write_text(
        "#include <algorithm>\n#include <cstdio>\n#include <vector>\n"
        "int main() { long N, n; if (scanf(\"%ld %ld\", &N, &n) != 2) return 1;\n"
        "  std::vector<int> v; for (long i = 0; i < N; ++i) { float g; if (scanf(\"%f\", &g) != 1) return 1;\n"
        "    if (g * n > 20 && (2 - g) * n > 20) v.push_back((int)i); }\n"
        "  std::random_shuffle(v.begin(), v.end());\n"
        "  for (int x : v) printf(\"%d\\n\", x); return 0; }\n")
    exe = d / "shuffle"
    subprocess.run(["g++", "-std=c++11", "-O1", "-w", str(src), "-o", str(exe)], check=True)

    def run(gm, n):
        inp = f"{len(gm)} {n}\n" + "\n".join("%.9g" % float(g) for g in gm) + "\n"
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
        return np.array([int(x) for x in out.split()], np.int64)
    return run


def test_lmm_sample_matches_compiled_random_shuffle(shuffle_prog):
    """The reference's --lmm sample (src/reml.cpp:920-925) three ways -- the oracle's Python
    restatement, the library's C++ (slemm_gpu_lmm_sample, host only) and libstdc++'s own
    random_shuffle on glibc's rand() -- on genotype means with SNPs right at the MAC-20 edges."""
    from oracle import gwas as ogwas
    rng = np.random.default_rng(3)
    n = 1000
    gm = rng.uniform(0.0, 2.0, 4000).astype(np.float32)
    gm[:6] = [0.02, 1.98, 0.0201, 1.9799, 0.0, 2.0]   # 20 and 1980 of 2000 alleles: excluded
    want = shuffle_prog(gm, n)
    assert want.size > 3500 and not set(want.tolist()) & {0, 1, 4, 5}
    assert np.array_equal(ogwas.lmm_sample(gm, n, len(gm)), want)
    assert np.array_equal(gwas.sample_gstat_snps(gm, n, len(gm)), want)
    # the first k of the shuffled order, as V keeps min(num_qf_markers, N) columns (:927)
    assert np.array_equal(gwas.sample_gstat_snps(gm, n, 30), want[:30])


def _gamma_cols_of_rows(rows):
    """BETA and SE of example rows as float32, CHISQ and P recomputed from them."""
    from scipy.stats import chi2
    beta = np.array([np.float32(r[6]) for r in rows], np.float32)
    se = np.array([np.float32(r[7]) for r in rows], np.float32)
    chisq = np.square(beta / se)
    return dict(BETA=beta, SE=se, CHISQ=chisq, P=chi2.sf(chisq, 1))


def test_example_gamma_chain_is_float32_bit_exact():
    """10k.gamma.txt: CHISQ = (BETA/SE)^2 in float32 and P = chi2.sf(CHISQ, 1) reproduce the text
    of every fixture row digit for digit (bin/slemm_gamma.py:321-325)."""
    rows = _meta()["gamma_rows"]
    cols = _gamma_cols_of_rows(rows)
    for k, r in enumerate(rows):
        assert str(cols["BETA"][k]) == r[6] and str(cols["SE"][k]) == r[7]
        assert str(cols["CHISQ"][k]) == r[8], (k, cols["CHISQ"][k], r[8])
        assert str(cols["P"][k]) == r[9], (k, cols["P"][k], r[9])


def test_assoc_columns_float32_against_restatement_and_closed_form():
    rng = np.random.default_rng(5)
    xy = rng.normal(size=500).astype(np.float32) * 30
    xx = rng.uniform(50, 3000, size=500).astype(np.float32)
    xy[7] = xx[7] = np.nan
    cf, sv = 1.0873412, 0.31
    got = gwas.assoc_columns(xy, xx, cf, sv)
    beta, se, chisq, pv = ogwas.gamma_stats(xy, xx, cf, sv)
    for a, b in ((got["BETA"], beta), (got["SE"], se), (got["CHISQ"], chisq), (got["P"], pv)):
        assert np.array_equal(a, b, equal_nan=True)
    assert got["BETA"].dtype == np.float32 and got["P"].dtype == np.float64
    ok = np.isfinite(xx)
    # (BETA/SE)^2 = xy^2 cf / (xx snp_var), in double
    want = xy[ok].astype(np.float64) ** 2 * cf / (xx[ok].astype(np.float64) * sv)
    assert np.max(np.abs(got["CHISQ"][ok] - want) / want) < 2e-6
    assert np.isnan(got["P"][7])


def test_write_assoc_reproduces_example_lines(tmp_path):
    """The GRAMMAR-gamma text writer (bin/slemm_gamma.py:327-361) on a .bim holding the example's
    variant columns writes the example's lines byte for byte, for the whole range and a sub-range."""
    meta = _meta()
    rows = meta["gamma_rows"]
    bim = tmp_path / "x.bim"
    bim.write_text("".join("\t".join(r[:6]) + "\n" for r in rows))
    cols = _gamma_cols_of_rows(rows)
    out = tmp_path / "x.gamma.txt"
    assert gwas.write_assoc(out, bim, cols) == len(rows)
    want = ["\t".join(meta["gamma_header"])] + ["\t".join(r) for r in rows]
    assert out.read_text().splitlines() == want
    b, e = 50, 120
    sub = {k: v[b:e + 1] for k, v in cols.items()}
    assert gwas.write_assoc(out, bim, sub, begin=b, end=e) == e - b + 1
    assert out.read_text().splitlines() == [want[0]] + want[1 + b:2 + e]


def test_write_assoc_headers(tmp_path):
    cols = dict(BETA=np.float32([0.5, np.nan]), SE=np.float32([0.25, np.nan]), CHISQ=np.float32([4.0, np.nan]),
                P=np.array([0.0455, np.nan]))
    pvar = tmp_path / "a.pvar"
    pvar.write_text("##fileformat\n#CHROM\tPOS\tID\tREF\tALT\n1\t10\trs1\tA\tG\n1\t20\trs2\tC\tT\n")
    gwas.write_assoc(tmp_path / "o1", pvar, cols)
    assert (tmp_path / "o1").read_text().splitlines() == [
        "#CHROM\tPOS\tID\tREF\tALT\tBETA\tSE\tCHISQ\tP", "1\t10\trs1\tA\tG\t0.5\t0.25\t4.0\t0.0455",
        "1\t20\trs2\tC\tT\tnan\tnan\tnan\tnan"]
    five = tmp_path / "b.txt"
    five.write_text("1 rs1 10 A G\n1 rs2 20 C T\n")
    gwas.write_assoc(tmp_path / "o2", five, cols)
    assert (tmp_path / "o2").read_text().splitlines()[0] == "CHR\tSNP\tBP\tA1\tA2\tBETA\tSE\tCHISQ\tP"


def test_write_gstat_reproduces_example_columns(tmp_path):
    """The .gstat.txt writer (src/output.cpp:85-120) given the example's printed values writes the
    same id, Chr, Pos, MAC, MAF, HWEP, quadratic forms and window size, and ChiSq/Pval that agree
    with the printed ones to their 6 digits (the inputs are themselves rounded to 6)."""
    meta = _meta()
    lines = meta["gstat_lines"]
    f = [l.split("\t") for l in lines[1:]]
    gst = {k: np.array([np.float32(r[6 + i]) for r in f]) for i, k in enumerate(("xPy", "xPx", "xBx", "xSx"))}
    out = tmp_path / "x.gstat.txt"
    gwas.write_gstat(out, [r[0] for r in f], [r[1] for r in f], [int(r[2]) for r in f],
                     [np.float32(r[4]) for r in f], [float(r[5]) for r in f], meta["pop_size"],
                     int(f[0][10]), np.float32(meta["Vg"]), gst)
    got = [l.split("\t") for l in out.read_text().splitlines()]
    assert "\t".join(got[0]) == lines[0]
    for g, w in zip(got[1:], f):
        assert g[:11] == w[:11], (g, w)
        assert abs(float(g[11]) - float(w[11])) <= 2e-5 * float(w[11])
        assert abs(float(g[12]) - float(w[12])) <= 2e-5
    # the scripts' readers of the same file
    src = tmp_path / "ex.gstat.txt"
    src.write_text("\n".join(lines) + "\n")
    r = [float(w[9]) / float(w[7]) for w in f]
    assert gwas.gamma_factor_from_gstat(src) == ogwas.mean_wo_outliers(r)
    cf, bs = gwas.gwa_factor_from_gstat(src)
    assert bs == int(f[-1][10])
    assert abs(cf - sum(float(w[8]) / float(w[7]) for w in f) / len(f)) < 1e-15


def test_mac_of_maf_float_product_double_sum():
    # the + 0.5 is a double sum: 16777215 + 0.5 in float would round to 16777216
    assert gwas.mac_of_maf(0.5, 16777215) == 16777215 == ogwas.mac_of_maf(0.5, 16777215)
    for row in _meta()["gstat"]:
        assert gwas.mac_of_maf(float(row["MAF"]), _meta()["pop_size"]) == int(row["MAC"])


def test_write_gstat_fake_rows(tmp_path):
    """--fake_geno rows of .gstat.txt (src/output.cpp:101-106): the marker index -(j+1) printed
    after "fake", NA in the columns a fake marker does not have."""
    gst = dict(xPy=np.float32([-70.8334, 12.5]), xPx=np.float32([172.206, 99.0]), xSx=np.float32([1727.29, 800.0]))
    out = tmp_path / "f.gstat.txt"
    gwas.write_gstat_fake(out, 1.5, gst)
    lines = out.read_text().splitlines()
    assert lines[1].split("\t")[:6] == ["fake-1", "NA", "NA", "NA", "NA", "NA"]
    f = lines[2].split("\t")
    assert f[0] == "fake-2" and f[8] == "NA" and f[10] == "NA" and f[6] == "12.5" and f[9] == "800"
    chisq = np.float32(12.5) * np.float32(12.5) / (np.float32(1.5) * np.float32(99.0))
    assert f[11] == "%g" % float(chisq)
