"""Association statistics on top of the engine: the --lmm gstat table and its text file
(src/output.cpp:85-120), the GRAMMAR-gamma scan with its outlier-trimmed correction factor
(bin/slemm_gamma.py:90-105, 163-174, 306-361) and the statistics and text file of the windowed scan
(bin/slemm_gwa.py:138-158, 427-466).  Host arithmetic in numpy with the reference's types (float32
statistics, Python-float factors, scipy's chi2.sf for the scripts' P); the per-SNP quadratic forms
and score terms come from the GPU (`Geno.lmm_gstat`, `Geno.gwas_score`, `Geno.gwa_window`)."""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import torch
from scipy.stats import chi2

from . import gpu

M_SQRT1_2 = 0.7071067811865476  # 1/sqrt(2) as a double


def chisq_pval_1df(x) -> np.ndarray:
    """Upper tail of the 1-df chi-square as the C++ side computes it, erfc(sqrt(x) * M_SQRT1_2)
    in double, 1 for x <= 0 (src/chisq.h:14-18)."""
    x = np.asarray(x, np.float64)
    out = np.ones_like(x)
    pos = x > 0
    out[pos] = [math.erfc(math.sqrt(v) * M_SQRT1_2) for v in x[pos]]
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
    lo, hi = q1 - (q3 - q1) * 1.5, q3 + (q3 - q1) * 1.5
    rsum, rcnt = 0.0, 0.0
    for v in x[(x > lo) & (x < hi)].tolist():  # left-to-right double sum, as the script's loop
        rsum += v
        rcnt += 1.0
    return rsum / rcnt


def correction_factor(gst: dict) -> float:
    """GRAMMAR-gamma correction factor: trimmed mean of xSx / xPx over the sampled SNPs
    (bin/slemm_gamma.py:163-174)."""
    return mean_wo_outliers(np.asarray(gst["xSx"], np.float64) / np.asarray(gst["xPx"], np.float64))


def gwa_correction_factor(gst: dict) -> float:
    """Correction factor of the windowed scan: the plain mean of xBx / xPx over the sampled SNPs
    (bin/slemm_gwa.py:141-157)."""
    r = np.asarray(gst["xBx"], np.float64) / np.asarray(gst["xPx"], np.float64)
    return float(r.sum() / r.size)


def assoc_columns(xy, xx, cfactor: float, snp_var: float) -> dict:
    """BETA, SE, CHISQ, P from float32 score terms as bin/slemm_gamma.py:321-325 and
    bin/slemm_gwa.py:428-431 compute them: the factors are Python floats, so BETA = xy/xx*cf,
    SE = sqrt((snp_var*cf)/xx) and CHISQ = (BETA/SE)^2 are float32 operations (the double
    snp_var*cf rounded to float32 once), and P = chi2.sf(CHISQ, 1) in float64.  NaN terms (SNPs
    under the MAC filter) stay NaN."""
    xy = np.asarray(xy, np.float32)
    xx = np.asarray(xx, np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        beta = xy / xx * np.float32(cfactor)
        se = np.sqrt(np.float32(float(snp_var) * float(cfactor)) / xx)
        chisq = np.square(beta / se)
    return dict(BETA=beta, SE=se, CHISQ=chisq, P=chi2.sf(chisq, 1))


def gwa_window_stats(xy, xx, cfactor: float, snp_var: float) -> dict:
    """BETA, SE, CHISQ, P of the windowed scan from its score terms (bin/slemm_gwa.py:428-431;
    `xx` already divided by tau, :427)."""
    return assoc_columns(xy, xx, cfactor, snp_var)


def grammar_gamma(geno: "gpu.Geno", Py: np.ndarray, gst: dict, snp_var: float, mac: float = 0.0,
                  cfactor: Optional[float] = None) -> dict:
    """Single-SNP GRAMMAR-gamma tests of every SNP of the handle (bin/slemm_gamma.py:306-325):
    score terms xy = x_c'Py and xx = ||(I-QQ')R^-1/2 x_c||^2 on the GPU (the handle's model: R
    and Q of the fit), then BETA, SE, CHISQ and P as `assoc_columns` computes them.  SNPs with
    minor-allele count at or below `mac` are NaN."""
    cf = correction_factor(gst) if cfactor is None else float(cfactor)
    py = torch.as_tensor(np.ascontiguousarray(Py, np.float32)).cuda()
    xy, xx = geno.gwas_score(py, skip_mac=mac)
    out = assoc_columns(xy.cpu().numpy(), xx.cpu().numpy(), cf, snp_var)
    out["cfactor"] = cf
    return out


def sample_gstat_snps(gm: np.ndarray, n: int, k: int) -> np.ndarray:
    """The --lmm SNP sample of src/reml.cpp:920-925: k SNPs with minor-allele count above 20 in
    the order std::random_shuffle leaves them with rand() in its initial state, as the reference
    draws them (the library's C++ restatement, slemm_gpu_lmm_sample).  Fewer than k come back
    when fewer SNPs pass the filter."""
    import ctypes as C
    from . import _lib as L
    g = np.ascontiguousarray(gm, np.float32)
    out = np.zeros(max(int(k), 1), np.int64)
    cnt = C.c_int()
    L.check(L.load().slemm_gpu_lmm_sample(g.ctypes.data_as(C.POINTER(C.c_float)), g.size, int(n), int(k),
                                          out.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(cnt)))
    return out[:cnt.value].copy()


def _fmt_g(v) -> str:
    """A number as a C++ ostream prints it by default (%g, 6 significant digits)."""
    return "%g" % float(v)


def mac_of_maf(maf, pop_size: int) -> int:
    """MAC column of .gstat.txt, int(maf*2*pop_size + 0.5) with a float product and a double
    sum (src/output.cpp:112)."""
    f = np.float32(np.float32(maf) * np.float32(2)) * np.float32(pop_size)
    return int(float(f) + 0.5)


def write_gstat(path, variants: Sequence[str], chrom: Sequence[str], pos: Sequence[int], maf, hwep,
                pop_size: int, window_size: int, vg: float, gst: dict) -> None:
    """The .gstat.txt of --lmm (src/output.cpp:85-120): one line per tested SNP with its id,
    chromosome, position, MAC, MAF, HWE p-value, xPy, xPx, xBx, xSx, the window size and
    ChiSq = xPy^2/(vg xPx) in float with its p-value, numbers printed as an ostream does.  `maf`
    and `hwep` are the caller's per-SNP values (the HWE test is not part of this library: None
    writes NA)."""
    xPy = np.asarray(gst["xPy"], np.float32)
    xPx = np.asarray(gst["xPx"], np.float32)
    xBx = np.asarray(gst["xBx"], np.float32)
    xSx = np.asarray(gst["xSx"], np.float32)
    maf = np.asarray(maf, np.float32)
    vg32 = np.float32(vg)
    with np.errstate(divide="ignore", invalid="ignore"):
        chisq = (xPy * xPy) / (vg32 * xPx)
    pval = chisq_pval_1df(chisq)
    with open(path, "w") as f:
        f.write("variant\tChr\tPos\tMAC\tMAF\tHWEP\txPy\txPx\txBx\txSx\tWindowSize\tChiSq\tPval\n")
        for j, vid in enumerate(variants):
            hw = "NA" if hwep is None else _fmt_g(np.float32(hwep[j]))
            f.write("\t".join([str(vid), str(chrom[j]), str(int(pos[j])), str(mac_of_maf(maf[j], pop_size)),
                               _fmt_g(maf[j]), hw, _fmt_g(xPy[j]), _fmt_g(xPx[j]), _fmt_g(xBx[j]),
                               _fmt_g(xSx[j]), str(int(window_size)), _fmt_g(chisq[j]), _fmt_g(pval[j])]) + "\n")


def write_gstat_fake(path, vg: float, gst: dict) -> None:
    """The .gstat.txt of --lmm --fake_geno (src/output.cpp:101-106): rows fake-1, fake-2, ... with
    NA for Chr, Pos, MAC, MAF, HWEP, xBx and WindowSize."""
    xPy = np.asarray(gst["xPy"], np.float32)
    xPx = np.asarray(gst["xPx"], np.float32)
    xSx = np.asarray(gst["xSx"], np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        chisq = (xPy * xPy) / (np.float32(vg) * xPx)
    pval = chisq_pval_1df(chisq)
    with open(path, "w") as f:
        f.write("variant\tChr\tPos\tMAC\tMAF\tHWEP\txPy\txPx\txBx\txSx\tWindowSize\tChiSq\tPval\n")
        for j in range(xPy.size):
            f.write(f"fake-{j + 1}\tNA\tNA\tNA\tNA\tNA\t" + "\t".join(
                [_fmt_g(xPy[j]), _fmt_g(xPx[j]), "NA", _fmt_g(xSx[j]), "NA", _fmt_g(chisq[j]), _fmt_g(pval[j])]) + "\n")


def read_gstat(path) -> dict:
    """Columns of a .gstat.txt as the scripts read them: whitespace-split fields, numbers as
    Python floats (bin/slemm_gamma.py:163-172, bin/slemm_gwa.py:141-156)."""
    with open(path) as f:
        next(f)
        rows = [line.strip().split() for line in f]
    return {"xPx": [float(r[7]) for r in rows], "xBx": [float(r[8]) for r in rows],
            "xSx": [float(r[9]) for r in rows], "WindowSize": [int(r[10]) for r in rows]}


def gamma_factor_from_gstat(path) -> float:
    """GRAMMAR-gamma correction factor from a .gstat.txt: trimmed mean of xSx/xPx over its rows,
    ratios of the printed values in double (bin/slemm_gamma.py:163-174)."""
    g = read_gstat(path)
    return mean_wo_outliers([y / x for x, y in zip(g["xPx"], g["xSx"])])


def gwa_factor_from_gstat(path):
    """Correction factor and block size of the windowed scan from a .gstat.txt: the plain mean of
    xBx/xPx accumulated left to right in double, and the WindowSize column of its last row
    (bin/slemm_gwa.py:138-158)."""
    g = read_gstat(path)
    sum_r, num_r = 0.0, 0
    for x, y in zip(g["xPx"], g["xBx"]):
        sum_r += y / x
        num_r += 1
    return sum_r / num_r, g["WindowSize"][-1]


def write_assoc(path, variant_file, cols: dict, begin: int = 0, end: Optional[int] = None) -> int:
    """Text output of the GRAMMAR-gamma and windowed scans (bin/slemm_gamma.py:327-361,
    bin/slemm_gwa.py:434-466): the variant file's last '#' header line (a .pvar) extended with
    BETA, SE, CHISQ, P, else the 6-column (.bim with CM) or 5-column header; then each variant
    line with index in [begin, end], stripped, with the four values appended as str() of their
    numpy types (float32 statistics, float64 P).  Returns the number of lines written."""
    beta, se, chisq, pval = cols["BETA"], cols["SE"], cols["CHISQ"], cols["P"]
    end = (1 << 62) if end is None else end
    written = 0
    with open(variant_file) as fp, open(path, "w") as out:
        header, first = None, None
        for line in fp:
            if line[0] == "#":
                header = line
            else:
                first = line
                break
        if first is None:
            raise ValueError(f"{variant_file}: no variant lines")
        if header is not None:
            out.write(header.strip() + "\tBETA\tSE\tCHISQ\tP\n")
        elif len(first.split()) == 6:
            out.write("CHR\tSNP\tCM\tBP\tA1\tA2\tBETA\tSE\tCHISQ\tP\n")
        else:
            out.write("CHR\tSNP\tBP\tA1\tA2\tBETA\tSE\tCHISQ\tP\n")

        def emit(line, k):
            out.write(line.strip() + "\t" + str(beta[k]) + "\t" + str(se[k]) + "\t" + str(chisq[k]) + "\t"
                      + str(pval[k]) + "\n")

        if begin == 0:
            emit(first, 0)
            written += 1
        idx = 0
        for line in fp:
            idx += 1
            if idx > end:
                break
            if idx >= begin:
                emit(line, idx - begin)
                written += 1
    return written

