"""Association statistics on top of the engine: the --lmm gstat table (src/output.cpp:85-120) and
the GRAMMAR-gamma scan with its outlier-trimmed correction factor (bin/slemm_gamma.py:90-105,
163-174, 306-325).  Host arithmetic in numpy; the per-SNP quadratic forms and score terms come
from the GPU (`Geno.lmm_gstat`, `Geno.gwas_score`)."""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

from . import gpu


def chisq_pval_1df(x) -> np.ndarray:
    """Upper tail of the 1-df chi-square, erfc(sqrt(x/2)), 1 for x <= 0 (src/chisq.h:14-18)."""
    x = np.asarray(x, np.float64)
    out = np.ones_like(x)
    pos = x > 0
    out[pos] = [math.erfc(math.sqrt(v * 0.5)) for v in x[pos]]
    return out


def gstat_table(gst: dict, vg: float) -> dict:
    """Columns of .gstat.txt computed from the quadratic forms: ChiSq = xPy^2 / (vg xPx) in
    float32 (src/output.cpp:105, 115) and its p-value."""
    xPy = np.asarray(gst["xPy"], np.float32)
    xPx = np.asarray(gst["xPx"], np.float32)
    chisq = (xPy * xPy) / (np.float32(vg) * xPx)
    return dict(gst, ChiSq=chisq, Pval=chisq_pval_1df(chisq))


def mean_wo_outliers(x) -> float:
    """Mean of the values strictly inside the 1.5-IQR fences around the quartiles (numpy's
    default linear percentiles), as bin/slemm_gamma.py:90-105."""
    x = np.asarray(x, np.float64)
    q1, q3 = np.percentile(x, 25), np.percentile(x, 75)
    lo, hi = q1 - 1.5 * (q3 - q1), q3 + 1.5 * (q3 - q1)
    inside = x[(x > lo) & (x < hi)]
    return float(inside.sum() / inside.size)


def correction_factor(gst: dict) -> float:
    """GRAMMAR-gamma correction factor: trimmed mean of xSx / xPx over the sampled SNPs
    (bin/slemm_gamma.py:163-174)."""
    return mean_wo_outliers(np.asarray(gst["xSx"], np.float64) / np.asarray(gst["xPx"], np.float64))


def gwa_correction_factor(gst: dict) -> float:
    """Correction factor of the windowed scan: the plain mean of xBx / xPx over the sampled SNPs
    (bin/slemm_gwa.py:141-157)."""
    r = np.asarray(gst["xBx"], np.float64) / np.asarray(gst["xPx"], np.float64)
    return float(r.sum() / r.size)


def gwa_window_stats(xy, xx, cfactor: float, snp_var: float) -> dict:
    """BETA, SE, CHISQ, P of the windowed scan from its score terms (bin/slemm_gwa.py:408-412)."""
    xy = np.asarray(xy, np.float64)
    xx = np.asarray(xx, np.float64)
    beta = xy / xx * cfactor
    se = np.sqrt((snp_var * cfactor) / xx)
    chisq = np.square(beta / se)
    pval = np.full_like(chisq, np.nan)
    ok = np.isfinite(chisq)
    pval[ok] = chisq_pval_1df(chisq[ok])
    return dict(BETA=beta, SE=se, CHISQ=chisq, P=pval)


def grammar_gamma(geno: "gpu.Geno", Py: np.ndarray, gst: dict, snp_var: float, mac: float = 0.0,
                  cfactor: Optional[float] = None) -> dict:
    """Single-SNP GRAMMAR-gamma tests of every SNP of the handle (bin/slemm_gamma.py:306-325):
    score terms xy = x_c'Py and xx = ||(I-QQ')R^-1/2 x_c||^2 on the GPU (the handle's model: R
    and Q of the fit), then BETA = xy/xx * cf, SE = sqrt(snp_var cf / xx), CHISQ = (BETA/SE)^2
    and its p-value.  SNPs with minor-allele count at or below `mac` are NaN."""
    cf = correction_factor(gst) if cfactor is None else float(cfactor)
    py = torch.as_tensor(np.ascontiguousarray(Py, np.float32)).cuda()
    xy, xx = geno.gwas_score(py, skip_mac=mac)
    xy = xy.double().cpu().numpy()
    xx = xx.double().cpu().numpy()
    beta = xy / xx * cf
    se = np.sqrt((snp_var * cf) / xx)
    chisq = np.square(beta / se)
    pval = np.full_like(chisq, np.nan)
    ok = np.isfinite(chisq)
    pval[ok] = chisq_pval_1df(chisq[ok])
    return dict(BETA=beta, SE=se, CHISQ=chisq, P=pval, cfactor=cf)


def sample_gstat_snps(gm: np.ndarray, n: int, k: int, seed: int = 0) -> np.ndarray:
    """k SNPs with minor-allele count above 20 (src/reml.cpp:923-925), drawn uniformly without
    replacement.  The reference shuffles with std::random_shuffle (rand()), whose order is not
    reproduced; the returned indices are sorted."""
    gm = np.asarray(gm, np.float32)
    ok = np.nonzero((gm * n > 20) & ((2 - gm) * n > 20))[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(ok, size=min(k, ok.size), replace=False))
