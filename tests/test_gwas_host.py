"""Host statistics of the association layer (slemm_amd/gwas.py) against the oracle and the
reference's example outputs (examples/expected/slemm/10k.gstat.txt, 10k.gamma.txt)."""
import json
import math
from pathlib import Path

import numpy as np

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
    s = gwas.sample_gstat_snps(gm, 1000, 10, seed=1)
    assert set(s.tolist()) == {1, 2, 5} and math.isclose(gm[4] * 1000, 20.0)
