"""ORACLE -- test infrastructure only.  GWAS statistics of the REML path restated in numpy:
the gstat quadratic-form statistic of src/output.cpp:98-117, the GRAMMAR-gamma score test
of bin/slemm_gamma.py:90-105, :296-325, and the --lmm SNP sample of src/reml.cpp:920-925
(glibc's rand() and libstdc++'s std::random_shuffle restated in Python)."""
from __future__ import annotations

import math

import numpy as np


def glibc_rand(count: int, seed: int = 1) -> list:
    """The first `count` outputs of glibc's rand() after srand(seed) -- seed 1 is also its state
    when srand is never called, as in the reference.  Published definition of the TYPE_3
    generator: r_0 = seed, r_i = 16807 r_{i-1} mod (2^31 - 1) (i < 31), r_i = r_{i-31}
    (31 <= i < 34), r_i = r_{i-31} + r_{i-3} mod 2^32 (i >= 34), output k = r_{k+344} >> 1."""
    r = [seed if seed else 1]
    for i in range(1, 31):
        r.append((16807 * r[-1]) % 2147483647)
    for i in range(31, 34):
        r.append(r[i - 31])
    for i in range(34, 344 + count):
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
    return [x >> 1 for x in r[344:344 + count]]


def random_shuffle(v, rand_stream) -> list:
    """libstdc++'s std::random_shuffle(first, last) (bits/stl_algo.h, GCC 11): for each i from 1
    to N-1, j = rand() % (i + 1), swap v[i] and v[j] when they differ."""
    v = list(v)
    it = iter(rand_stream)
    for i in range(1, len(v)):
        j = next(it) % (i + 1)
        if i != j:
            v[i], v[j] = v[j], v[i]
    return v


def lmm_sample(gm, n: int, k: int) -> np.ndarray:
    """src/reml.cpp:920-925: indices i with gm_i*n > 20 and (2 - gm_i)*n > 20 in float32,
    std::random_shuffle'd with rand() in its initial state, first min(k, N)."""
    gm = np.asarray(gm, np.float32)
    fn = np.float32(n)
    ok = np.flatnonzero((gm * fn > np.float32(20)) & ((np.float32(2) - gm) * fn > np.float32(20)))
    v = random_shuffle(ok.tolist(), glibc_rand(max(len(ok) - 1, 0)))
    return np.asarray(v[:k], np.int64)


def one_df_chisq_pval(x: float) -> float:
    """getOneDfChisqPval (src/chisq.h:14-18)."""
    return 1.0 if x <= 0 else math.erfc(math.sqrt(x) * math.sqrt(0.5))


def gstat_chisq(xPy: float, xPx: float, vg: float) -> float:
    """ChiSq column of .gstat.txt (src/output.cpp:105,115), float arithmetic."""
    xPy, xPx, vg = np.float32(xPy), np.float32(xPx), np.float32(vg)
    return float((xPy * xPy) / (vg * xPx))


def mac_of_maf(maf: float, pop_size: int) -> int:
    """MAC column of .gstat.txt (src/output.cpp:112): maf*2*pop_size is a float product, the
    + 0.5 a double sum."""
    prod = np.float32(np.float32(maf) * np.float32(2)) * np.float32(pop_size)
    return int(float(prod) + 0.5)


def mean_wo_outliers(x) -> float:
    """Correction factor: mean of x inside the 1.5-IQR fences (bin/slemm_gamma.py:90-105)."""
    x = list(x)
    q1 = np.percentile(x, 25)
    q3 = np.percentile(x, 75)
    lower = q1 - (q3 - q1) * 1.5
    upper = q3 + (q3 - q1) * 1.5
    rsum, rcnt = 0.0, 0.0
    for v in x:
        if lower < v < upper:
            rsum += v
            rcnt += 1
    return rsum / rcnt


def gamma_scores(Z: np.ndarray, Py: np.ndarray, rr: np.ndarray, Q: np.ndarray, mac: float = 0.0):
    """Per-SNP xy = x_c'Py and xx = ||(I-QQ')R^-1/2 x_c||^2 (bin/slemm_gamma.py:306-319) for the
    dosage matrix Z (n x m); SNPs at or below the MAC filter get NaN."""
    Z = np.asarray(Z, np.float64)
    n, m = Z.shape
    xy = np.empty(m)
    xx = np.empty(m)
    Py = np.asarray(Py, np.float64)
    rr = np.asarray(rr, np.float64)
    Q = np.asarray(Q, np.float64).reshape(n, -1)
    for j in range(m):
        x = Z[:, j].copy()
        s = x.sum()
        if s <= mac or 2 * n - s <= mac:
            xy[j] = xx[j] = np.nan
            continue
        x -= s / n
        xy[j] = x @ Py
        x *= rr
        x -= Q @ (Q.T @ x)
        xx[j] = x @ x
    return xy, xx


def gamma_stats(xy, xx, cfactor: float, snp_var: float):
    """BETA, SE, CHISQ, P of bin/slemm_gamma.py:321-325: xy and xx are float32 arrays there and
    cfactor, snp_var Python floats, so every step but the double product snp_var*cfactor is a
    float32 operation; P = chi2.sf in float64."""
    from scipy.stats import chi2

    xy = np.asarray(xy, np.float32)
    xx = np.asarray(xx, np.float32)
    cf32 = np.float32(cfactor)
    with np.errstate(divide="ignore", invalid="ignore"):
        beta = xy / xx * cf32
        serr = np.sqrt(np.float32(float(snp_var) * float(cfactor)) / xx)
        chisq = np.square(beta / serr)
    return beta, serr, chisq, chi2.sf(chisq, 1)


def gwa_blocks(model_snps, block_size: int, test_begin: int, test_end: int):
    """Blocks of model SNPs and the tested range [lo, hi) of each (bin/slemm_gwa.py:322-345)."""
    ksnp = np.asarray(model_snps, np.int64)
    K = len(ksnp)
    b = block_size if block_size <= K else int(K / 2) * 2
    step = int(b / 2)
    ext = int(step / 2)
    nb = int((K - step) / step)
    blocks = [ksnp[i * step:i * step + b] for i in range(nb)]
    if (K - step) % step != 0:
        blocks.append(ksnp[K - b:K])
    bounds = np.empty((len(blocks), 2), np.int64)
    bounds[0, 0] = test_begin
    bounds[0, 1] = blocks[0][b - ext]
    for i in range(1, len(blocks)):
        bounds[i, 0] = bounds[i - 1, 1]
        bounds[i, 1] = blocks[i][b - ext]
    bounds[-1, 1] = test_end + 1
    return blocks, bounds


def gwa_window_scores(Z, y, rr, Q, model_snps, block_size, test_begin, test_end, tau, model_snp_ct, mac=0.0):
    """Score terms of the windowed scan (bin/slemm_gwa.py:347-407) in FP64: per block
    H = Zb'Zb + model_snp_ct tau diag((2 - p) p / 2); per tested SNP xy = x'y and
    xx = (x'x - z'H^-1 z) / tau, x centred (float32 mean), scaled by rr, projected off Q."""
    Z = np.asarray(Z, np.float64)
    n = Z.shape[0]
    rr = np.asarray(rr, np.float64)
    Q = np.asarray(Q, np.float64).reshape(n, -1)
    y = np.asarray(y, np.float64)
    blocks, bounds = gwa_blocks(model_snps, block_size, test_begin, test_end)
    t = test_end - test_begin + 1
    xy = np.full(t, np.nan)
    xx = np.full(t, np.nan)

    def prep(cols):
        mu = np.float32(cols.sum(axis=0)) / np.float32(n)  # float32 mean of exact integer sums
        A = (cols - mu.astype(np.float64)[None, :]) * rr[:, None]
        return A - Q @ (Q.T @ A), mu

    for blk, (lo, hi) in zip(blocks, bounds):
        Zb, mu = prep(Z[:, blk])
        rv = ((np.float32(2.0) - mu) * mu / np.float32(2.0)).astype(np.float64)
        H = Zb.T @ Zb + np.diag(model_snp_ct * tau * rv)
        for j in range(max(lo, test_begin), min(hi, test_end + 1)):
            x = Z[:, j]
            s = x.sum()
            if s <= mac or 2 * n - s <= mac:
                continue
            xc = x - float(np.float32(s) / np.float32(n))
            xy[j - test_begin] = xc @ y
            v = xc * rr
            v -= Q @ (Q.T @ v)
            z = Zb.T @ v
            xx[j - test_begin] = (v @ v - z @ np.linalg.solve(H, z)) / tau
    return xy, xx
