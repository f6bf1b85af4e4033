"""Benchmark of the stochastic-Lanczos REML hot path on one MI355X.

Workload (default): synthetic PLINK bed with n=100,000 individuals and m=500,000 SNPs
(MAF ~ U(0.05,0.5), Binomial(2,p) under HWE), generated directly in HBM; block of 64
Rademacher probes + 1 phenotype right-hand side (K = 65).  One step = one call of the REML
operator Zc W Zc' (lnr_op_select, src/reml.cpp:96-118) on the whole block in split-BF16 on
MFMA: two passes over the packed bed (Z'V then Z U).

Metric: bed-matvec throughput in GB/s of algorithmic bytes, per pass
    m*ceil(n/4) + n*K*4 + m*K*4
(two passes per step), with the plain FP32 path timed beside it, plus REML time-to-converge
when the driver is available.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA
INT8_PEAK_TOPS = 5000.0     # dense int8 MFMA
FP32_PEAK_TFLOPS = 157.3


# Memory-side bytes per launch of the split-BF16 block product (byte-table pipelined kernel at
# K=65, narrow-block kernel at K=1; 2048-workgroup grids) measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE on this workload (fetch counter scaled x2.0, write x1.0 against a 1 GiB device copy);
# commands and calibration in profiles/blockprod_c2_k65_hbm_traffic.md.  FETCH_SIZE includes
# Infinity Cache hits.
MEASURED_TRAFFIC = {(100_000, 500_000, 65): {"pass1": 15.14e9 + 0.321e9, "pass2": 17.44e9 + 0.193e9},
                    (100_000, 500_000, 1): {"pass1": 12.64e9 + 0.064e9, "pass2": 12.71e9 + 0.039e9}}
# The same for the int8 kernels (wide kernel at K=65: two column groups, each reading the codes;
# narrow kernel at K=1); profiles/blockprod_int8_c2_hbm_traffic.md.
I8_MEASURED_TRAFFIC = {(100_000, 500_000, 65): {"pass1": 27.62e9 + 0.398e9, "pass2": 32.85e9 + 0.293e9},
                       (100_000, 500_000, 1): {"pass1": 12.59e9 + 0.012e9, "pass2": 12.59e9 + 0.009e9}}


def measured_traffic(n, m, K, kernel_substr, timeout=300, env_extra=None):
    """HBM bytes per launch of the dominant block-product kernel, measured live: two child runs
    of this script's --only-op mode under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`
    (one counter per run), each scaled by what the same counter reads for a 1 GiB device copy in
    the same run (--copy-calibration); kernel_substr: "|"-separated names of the kernel.  Returns {pass1, pass2, fetch_factor, write_factor} in bytes,
    or {"error": ...} when the profiler is missing or fails (then the line reports traffic null)."""
    import glob
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    if shutil.which("rocprofv3") is None:
        return {"error": "rocprofv3 not found"}
    env = dict(os.environ, TMPDIR="/tmp", **(env_extra or {}))
    vals = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="slemm_traffic_", dir="/tmp")
        cmd = ["rocprofv3", "--pmc", counter, "-d", d, "-o", "t", "--", sys.executable, os.path.abspath(__file__),
               "--only-op", "--copy-calibration", "--steps", "1", "--warmup", "0", "--no-autotune",
               "--n", str(n), "--m", str(m), "--probes", str(K - 1)]
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, timeout=timeout, capture_output=True)
            dbs = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
            if not dbs:
                return {"error": f"rocprofv3 --pmc {counter} wrote no results"}
            con = sqlite3.connect(dbs[0])
            rows = con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                               "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,)).fetchall()
            con.close()
        except Exception as e:  # reported, never fatal to the bench line
            return {"error": f"{counter}: {e!r}"}
        finally:
            shutil.rmtree(d, ignore_errors=True)
        copies = [v for _, kn, v in rows if "copyBuffer" in kn]
        kern = [v for _, kn, v in rows if any(k in kn for k in kernel_substr.split("|"))]
        if not copies or len(kern) < 2:
            return {"error": f"{counter}: copy or block-product dispatches missing"}
        factor = float(1 << 30) / (max(copies) * 1024.0)  # the counter reads KiB
        if not 0.25 < factor < 8.0:
            return {"error": f"{counter}: implausible copy calibration factor {factor:.3f}"}
        vals[counter] = (factor, [kern[0] * 1024.0 * factor, kern[1] * 1024.0 * factor])
    (ff, fetch), (wf, write) = vals["FETCH_SIZE"], vals["WRITE_SIZE"]
    return {"pass1": fetch[0] + write[0], "pass2": fetch[1] + write[1], "fetch_factor": round(ff, 4),
            "write_factor": round(wf, 4), "fetch_bytes": [round(x) for x in fetch],
            "write_bytes": [round(x) for x in write]}


def issue_factor(K: int) -> float:
    """MFMA columns issued per algorithmic column of a K-vector split-BF16 block: a hi and a lo
    tile per 16 vectors, except that a last tile of at most 8 vectors holds its hi and lo halves
    as columns (one tile), as do blocks of at most 8 vectors."""
    q, r = divmod(K, 16)
    tiles = 2 * q + (0 if r == 0 else (1 if r <= 8 else 2))
    return 16 * tiles / K


def i8_issue_factor(K: int) -> float:
    """Int8 MFMA columns issued per algorithmic column: three digit columns per vector, in tiles
    of 16 vectors (48 columns), and the up to 5 vectors after the last full tile in one 16-column
    tail tile (blocks of 1 to 5 vectors: one tile)."""
    if K <= 5:
        return 16.0 / K
    q, r = divmod(K, 16)
    tiles = 3 * q + 1 if 1 <= r <= 5 else 3 * ((K + 15) // 16)
    return 16.0 * tiles / K


class bf16_mfma_kernels:
    """Context: run the split-BF16 operand on the BF16 MFMA kernels (FP32 accumulation) instead of
    the int8 ones (SLEMM_I8_NARROW=0, SLEMM_I8_WIDE=0, read by the library per call)."""

    def __enter__(self):
        self.saved = {k: os.environ.get(k) for k in ("SLEMM_I8_NARROW", "SLEMM_I8_WIDE")}
        os.environ["SLEMM_I8_NARROW"] = "0"
        os.environ["SLEMM_I8_WIDE"] = "0"
        return self

    def __exit__(self, *exc):
        for k, v in self.saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        return False


def i8_on(g, K: int) -> bool:
    """Whether a K-vector block runs on the int8 kernels, as the library selects it now (blocks of
    6 to 8 vectors stay on the BF16 kernels with hi and lo halves as columns; SLEMM_I8_NARROW /
    SLEMM_I8_WIDE switch the others)."""
    return g.block_kernel(K).startswith("int8")


def i8_target_env(tune) -> dict:
    """SLEMM_I8_TARGET / SLEMM_I8_RT that reproduce an int8 autotune result in a child process."""
    if not tune or not tune.get("kernel", "").startswith("int8"):
        return {}
    env = {"SLEMM_I8_TARGET": f"{tune['pass1'][1]},{tune['pass2'][1]}"}
    return env


def pass_bytes(n, m, K):
    return m * ((n + 3) // 4) + n * K * 4 + m * K * 4


def kernel_ms(lib, pas, prec):
    tot, cnt = C.c_double(), C.c_int()
    lib.slemm_gpu_timing_read(pas, prec, C.byref(tot), C.byref(cnt))
    return tot.value / max(cnt.value, 1), cnt.value


def lmc_timing(g, X, y, h2_sldf):
    """LMC_REML (src/reml.cpp:730-882, the estimator `slemm --reml` runs by default) on the same
    fit: n_MC as src/option.cpp:414, the reference's secant schedule, both precisions; wall time
    from the resident bed, Lanczos steps, secant evaluations and h2 beside SLDF's."""
    import torch
    out = {}
    for prec in ("bf16x2", "fp32"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = g.LMC_REML(X, y, precision=prec)
        torch.cuda.synchronize()
        out[prec] = {"seconds": round(time.perf_counter() - t0, 3), "n_MC": r["n_MC"], "K": 1 + 2 * r["n_MC"],
                     "lanczos_steps": int(r["steps"]), "secant_evals": int(r["evals"]),
                     "h2": round(float(r["h2"]), 6), "vg": float(r["vg"]), "ve": float(r["ve"])}
    out["h2_sldf_same_phenotype"] = round(h2_sldf, 6)
    out["speedup_bf16x2_over_fp32"] = round(out["fp32"]["seconds"] / out["bf16x2"]["seconds"], 2)
    return out


def _panel_engine(rows, n):
    from oracle.cpu_panels import PanelModel
    return lambda gm, gw, rr, Q, ss: PanelModel(rows, n, gm, gw, rr, Q, ss)


def cpu_op_seconds(rows, n, gm, gw, ts, reps=1):
    """Per-call seconds of the reference-structured CPU operator (oracle.cpu_panels: per-subset
    float panels + two MKL sgemms) over the given SNP rows, for each column count in ts."""
    import numpy as np
    from oracle.cpu_panels import PanelModel
    eng = PanelModel(rows, n, gm, gw, np.ones(n, np.float32), np.full((n, 1), 1 / np.sqrt(n), np.float32))
    rng = np.random.default_rng(0)
    out = {}
    for t in ts:
        V = rng.choice([-1.0, 1.0], size=(n, t)).astype(np.float32) / np.float32(np.sqrt(n))
        eng.op("ldet", 0.4286, V)  # page-in, thread pool
        t0 = time.perf_counter()
        for _ in range(reps):
            eng.op("ldet", 0.4286, V)
        out[t] = (time.perf_counter() - t0) / reps
    return out


def c1_run(n=5000, m=50_000, nv=30, n_causal=1000, h2=0.5, cpu=True, seed=3):
    """The reference's CPU-runnable workload: n=5,000, m=50,000, MAF~U(0.05,0.5), 1% missing, one
    trait at h2=0.5 on 1,000 causal SNPs, 30 Rademacher probes.  SLDF-REML and LMC-REML on the GPU
    (both precisions) and, as the reported CPU baseline, the same drivers on host cores over the
    reference-structured panel engine (per-subset float panels, two MKL sgemms each), end to end."""
    import numpy as np
    import torch
    from slemm_amd import gpu
    from oracle import reml as oreml

    g = gpu.Geno.synthesize(n, m, seed=seed, maf_lo=0.05, maf_hi=0.5, missing=0.01)
    g.set_model()
    y, _, _, rh2 = simulate_phenotype(g, n, m, n_causal, h2, seed + 1)
    X = np.ones((n, 1), np.float32)
    out = {"workload": f"n={n} m={m}, 1% missing, h2={h2} on {n_causal} causal SNPs, {nv} probes",
           "realised_h2": round(rh2, 5)}
    for prec in ("bf16x2", "fp32"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv, precision=prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        q = g.LMC_REML(X, y, precision=prec)
        torch.cuda.synchronize()
        out[f"gpu_{prec}"] = {"sldf_s": round(t1 - t0, 3), "sldf_h2": round(float(r["h2"]), 6),
                              "lmc_s": round(time.perf_counter() - t1, 3), "lmc_h2": round(float(q["h2"]), 6)}
    if cpu:
        rows = g.download(0, m)
        gm, gw, _ = g.model()
        eng = _panel_engine(rows, n)
        t0 = time.perf_counter()
        a = oreml.SLDF_REML(None, gm, gw, None, X, y, n_V=nv, engine=eng)
        t1 = time.perf_counter()
        b = oreml.LMC_REML(None, gm, gw, None, X, y, engine=eng)
        t2 = time.perf_counter()
        out["cpu_reference_structure"] = {
            "sldf_s": round(t1 - t0, 2), "sldf_h2": round(float(a["h2"]), 6), "lmc_s": round(t2 - t1, 2),
            "lmc_h2": round(float(b["h2"]), 6), "threads": torch.get_num_threads(),
            "engine": "FP32 per-subset float panels (1000 SNPs) + two MKL sgemms per subset (torch CPU), "
                      "restated L_Seed/L_Solve/GSS/secant (oracle.cpu_panels, oracle.reml)"}
        gpu_s = out["gpu_bf16x2"]
        out["speedup_gpu_over_cpu"] = {"sldf": round((t1 - t0) / gpu_s["sldf_s"], 1),
                                       "lmc": round((t2 - t1) / gpu_s["lmc_s"], 1)}
    g.close()
    return out


def cpu_baseline(g, n, m, K, m_sample, reps, reml=None):
    """The reference-structured CPU operator (oracle.cpu_panels: per call, 1000-SNP float panels of
    centred dosages and two MKL sgemms per panel, src/reml.cpp:28-94) on the first m_sample SNPs of
    the same bed over all n individuals, scaled linearly to m (the operator's cost is linear in m).
    With the GPU fit's Lanczos step counts (reml), also the SLDF seed systems' CPU seconds
    extrapolated as per-call time x steps."""
    import torch

    rows = g.download(0, m_sample)
    gm, gw, _ = g.model()
    ts = sorted({K, 1, K - 1})
    per = cpu_op_seconds(rows, n, gm[:m_sample], gw[:m_sample], ts, reps)
    scale = m / m_sample
    t_full = per[K] * scale
    res = {
        "value": round(2 * pass_bytes(n, m, K) / t_full / 1e9, 3),
        "unit": "GB/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": (f"FP32 ldet operator with the reference's structure (per call: 1000-SNP float panels of "
                   f"centred dosages, two MKL sgemms each, src/reml.cpp:28-94; torch CPU) on the first "
                   f"{m_sample} SNPs of the same bed, n={n}, K={K}, {reps} calls of {per[K]:.3f} s, scaled "
                   f"linearly to m={m} ({t_full:.1f} s per call)"),
        "per_call_s_extrapolated": {str(t): round(v * scale, 2) for t, v in per.items()},
    }
    if reml is not None:
        st = reml["bf16x2"]["lanczos_steps"]  # y, Q, probes of the GPU fit
        ext = (st[0] * per[1] + st[1] * per[1] + st[2] * per[K - 1]) * scale
        res["sldf_seed_s_extrapolated"] = {
            "seconds": round(ext, 1),
            "how": (f"per-call CPU time x the GPU fit's Lanczos steps (y {st[0]}, Q {st[1]}, probes {st[2]}), "
                    f"three separate seed systems as the reference runs them; extrapolated, not run"),
            "gpu_seconds_whole_fit": reml["bf16x2"]["seconds"],
            "ratio": round(ext / reml["bf16x2"]["seconds"], 1)}
    return res


def small_k_passes(g, lib, L, n, m, ks=(1, 3, 5, 7, 16), reps=3):
    """HBM-bound regime (K*P <~ 39): Z'V and Z U passes at small K (GEBV K=1, GWAS K=1+c+1,
    LMC K=1+2*3, ...), kernel time from the library's HIP events; plus the FP64 GEBV kernel."""
    import torch

    out = {}
    for K in ks:
        V = torch.randn(n, K, device="cuda")
        U = torch.randn(m, K, device="cuda")
        at = g.autotune(K, reps=2)
        g.rmatvec(V)
        g.matvec(U)
        torch.cuda.synchronize()
        lib.slemm_gpu_timing_reset()
        lib.slemm_gpu_timing(1)
        for _ in range(reps):
            g.rmatvec(V)
            g.matvec(U)
        torch.cuda.synchronize()
        lib.slemm_gpu_timing(0)
        p1, _ = kernel_ms(lib, 0, L.PREC_BF16X2)
        p2, _ = kernel_ms(lib, 1, L.PREC_BF16X2)
        lib.slemm_gpu_timing_reset()
        b = pass_bytes(n, m, K)
        out[f"K={K}"] = {"pass1_ms": round(p1, 3), "pass2_ms": round(p2, 3),
                         "pass1_GB/s": round(b / p1 / 1e6, 1), "pass2_GB/s": round(b / p2 / 1e6, 1),
                         "pass1_hbm_frac": round(b / p1 / 1e6 / HBM_PEAK_GBS, 3),
                         "pass2_hbm_frac": round(b / p2 / 1e6 / HBM_PEAK_GBS, 3),
                         "kernels": [at["pass1"][:2], at["pass2"][:2]]}
        trf = (I8_MEASURED_TRAFFIC if i8_on(g, K) else MEASURED_TRAFFIC).get((n, m, K))
        if trf:
            out[f"K={K}"]["traffic_bytes"] = trf
    alpha = torch.randn(m, dtype=torch.float64, device="cuda")
    g.gebv(alpha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.gebv(alpha)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    gb = m * ((n + 3) // 4) + m * 8 + n * 8
    out["gebv_fp64"] = {"ms": round(dt * 1e3, 3), "GB/s": round(gb / dt / 1e9, 1),
                        "hbm_frac": round(gb / dt / 1e9 / HBM_PEAK_GBS, 3)}
    exact = g.gebv(alpha)
    g.gebv_fast(alpha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fast = g.gebv_fast(alpha)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    out["gebv_fast"] = {"ms": round(dt * 1e3, 3), "GB/s": round(gb / dt / 1e9, 1),
                        "hbm_frac": round(gb / dt / 1e9 / HBM_PEAK_GBS, 3),
                        "max_rel_err_vs_exact": float((fast - exact).abs().max() / exact.abs().max()),
                        "how": "one pass of the 1-vector block product (int8 narrow kernel) + FP64 constant"}
    return out


def gemm_ceiling(M=16384, N=16384, Kd=8192, reps=5):
    """Practical dense BF16 MFMA rate on this GPU: the vendor GEMM (hipBLASLt through torch) on a
    large square problem, for reading the block product's issued rate against (the published
    2.5 PFLOP/s assumes peak clocks that dense MFMA work does not sustain)."""
    import torch

    a = torch.randn(M, Kd, device="cuda", dtype=torch.bfloat16)
    b = torch.randn(Kd, N, device="cuda", dtype=torch.bfloat16)
    for _ in range(2):
        c = a @ b
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        c = a @ b
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del a, b, c
    torch.cuda.empty_cache()
    out = {"gemm": f"{M}x{N}x{Kd} bf16 torch.matmul (hipBLASLt)", "TFLOP/s": round(2.0 * M * N * Kd / dt / 1e12, 1)}
    try:  # the same shape in int8 (int32 results) through torch._int_mm
        a = torch.randint(-128, 128, (M, Kd), device="cuda", dtype=torch.int8)
        b = torch.randint(-128, 128, (Kd, N), device="cuda", dtype=torch.int8)
        for _ in range(2):
            c = torch._int_mm(a, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            c = torch._int_mm(a, b)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        del a, b, c
        torch.cuda.empty_cache()
        out["int8_gemm"] = f"{M}x{N}x{Kd} int8 torch._int_mm"
        out["int8_TOP/s"] = round(2.0 * M * N * Kd / dt / 1e12, 1)
    except Exception as e:  # reported, never fatal
        out["int8_error"] = repr(e)
    return out


MFMA_NAMES = ["i8_16x16x64", "i8_32x32x32", "bf16_16x16x32", "bf16_32x32x16",
              "i8_16x16x64_random", "i8_32x32x32_random", "bf16_16x16x32_random", "bf16_32x32x16_random"]


def mfma_ceiling(lib, L, configs, scale=1):
    """Register-only MFMA issue rate of this GPU (slemm_gpu_mfma_ceiling): for each (instruction,
    independent accumulator chains per wave, waves per SIMD) the tera-operations per second and
    the shader clock over the loop.  scale multiplies the loop length (1: a few ms per run; 20:
    about 100 ms, long enough for the power management to settle the clock)."""
    import ctypes as C
    rows = []
    for instr, chains, wps in configs:
        tops, cr, ms = C.c_double(), C.c_double(), C.c_double()
        iters = (20000 if instr % 4 in (0, 2) else 10000) * scale
        L.check(lib.slemm_gpu_mfma_ceiling(instr, chains, wps, iters, C.byref(tops), C.byref(cr), C.byref(ms)))
        rows.append({"instr": MFMA_NAMES[instr], "chains": chains, "waves_per_simd": wps, "iters": iters,
                     "TOP/s": round(tops.value, 1), "clock_GHz": round(cr.value * 0.1, 3), "ms": round(ms.value, 3)})
    return rows


def mfma_ceiling_summary(lib, L):
    """The best measured rate of the int8 and BF16 16x16 instructions over 4-8 chains and 2-4 waves
    per SIMD (the block products' regime), with fixed small operands and with random ones (the
    block products' bit switching, which lowers the clock under the power limit)."""
    rows = mfma_ceiling(lib, L, [(i, c, w) for i in (0, 2) for c in (4, 8) for w in (2, 4)])
    rows += mfma_ceiling(lib, L, [(i, 8, 4) for i in (4, 6)], scale=20)
    best = {}
    for r in rows:
        if r["instr"] not in best or r["TOP/s"] > best[r["instr"]]["TOP/s"]:
            best[r["instr"]] = r
    return best


def simulate_phenotype(g, n, m, n_causal, h2, seed):
    """y = Zc beta + e + 10, beta ~ N(0,1) on n_causal random SNPs, var(e) set for h2."""
    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    beta = np.zeros(m, np.float32)
    beta[rng.choice(m, size=n_causal, replace=False)] = rng.normal(size=n_causal).astype(np.float32)
    gv = g.matvec(torch.from_numpy(beta).cuda(), precision="fp32")[:, 0].double().cpu().numpy()
    ve = gv.var() * (1 - h2) / h2
    y = (gv + rng.normal(scale=np.sqrt(ve), size=n) + 10.0).astype(np.float32)
    return y, beta, gv, float(gv.var() / (gv.var() + ve))


def reml_time_to_converge(g, n, m, nv, y, realised_h2, n_causal, h2):
    """SLDF-REML (src/reml.cpp:549-690) on the resident bed for a phenotype simulated from it.
    Time from the resident bed to the final variance components, split-BF16 and FP32; plus the
    AI-REML driver (average-information Newton on the same model and probes) in split-BF16."""
    import numpy as np
    import torch

    X = np.ones((n, 1), np.float32)
    out = {"workload": f"SLDF-REML, {nv} Rademacher probes, simulated h2={h2} with {n_causal} causal SNPs",
           "realised_h2": round(realised_h2, 5)}
    res = {}
    for prec in ("bf16x2", "fp32"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv, precision=prec)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res[prec] = r
        out[prec] = {"seconds": round(wall, 3), "seconds_seed": round(r["seconds_seed"], 3),
                     "lanczos_steps": list(r["steps"]), "h2": round(float(r["h2"]), 6),
                     "vg": float(r["vg"]), "ve": float(r["ve"]), "lrt": float(r["lrt"])}
    out["speedup_bf16x2_over_fp32"] = round(out["fp32"]["seconds"] / out["bf16x2"]["seconds"], 2)
    out["vc_rel_diff_bf16x2_vs_fp32"] = {
        k: float(abs(res["bf16x2"][k] - res["fp32"][k]) / abs(res["fp32"][k])) for k in ("vg", "ve", "h2")}
    # the same split-BF16 operand on the BF16 MFMA kernels (FP32 accumulation) instead of int8
    with bf16_mfma_kernels():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv, precision="bf16x2")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    out["bf16x2_on_bf16_mfma"] = {"seconds": round(wall, 3), "seconds_seed": round(r["seconds_seed"], 3),
                                  "lanczos_steps": list(r["steps"]), "h2": round(float(r["h2"]), 6),
                                  "vg": float(r["vg"]), "ve": float(r["ve"]),
                                  "speedup_over_fp32": round(out["fp32"]["seconds"] / wall, 2),
                                  "vc_rel_diff_vs_fp32": {k: float(abs(r[k] - res["fp32"][k]) / abs(res["fp32"][k]))
                                                          for k in ("vg", "ve", "h2")}}
    # the same fits replaying the reference's criterion arithmetic (SLEMM_CRITERION_FP32=1: y'Py
    # from the L_Solve solution, criterion rounded to float before each comparison)
    saved = os.environ.get("SLEMM_CRITERION_FP32")
    os.environ["SLEMM_CRITERION_FP32"] = "1"
    try:
        c32 = {}
        for prec in ("bf16x2", "fp32"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv, precision=prec)
            torch.cuda.synchronize()
            c32[prec] = (time.perf_counter() - t0, r)
    finally:
        if saved is None:
            os.environ.pop("SLEMM_CRITERION_FP32", None)
        else:
            os.environ["SLEMM_CRITERION_FP32"] = saved
    out["criterion_reference_replay"] = {
        prec: {"seconds": round(sec, 3), "h2": round(float(r["h2"]), 6), "vg": float(r["vg"]), "ve": float(r["ve"]),
               "criterion_evals": int(r["evals"])} for prec, (sec, r) in c32.items()}
    out["criterion_reference_replay"]["vc_rel_diff_bf16x2_vs_fp32"] = {
        k: float(abs(c32["bf16x2"][1][k] - c32["fp32"][1][k]) / abs(c32["fp32"][1][k])) for k in ("vg", "ve", "h2")}
    out["lmc"] = lmc_timing(g, X, y, float(res["bf16x2"]["h2"]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = g.AI_REML(X, y, rng_seed=0, n_V=nv, precision="bf16x2")
    torch.cuda.synchronize()
    out["ai_reml_bf16x2"] = {"seconds": round(time.perf_counter() - t0, 3), "iterations": int(r["iterations"]),
                             "lanczos_steps": list(r["steps"]), "h2": round(float(r["h2"]), 6),
                             "vg": float(r["vg"]), "ve": float(r["ve"])}
    return out, res["bf16x2"]


def next_rows_run(g, n, m, y, sl, loader_m=40_000):
    """The rows either side of the hot path, timed on the resident bed and the split-BF16 SLDF fit
    of the REML run: the --lmm quadratic forms of 30 sampled SNPs with the reference's 1,000-SNP
    windows (src/reml.cpp:920-980), the windowed scan of slemm_gwa step 2 over the first 50,000 SNPs
    with every SNP a model SNP and 1,000-SNP blocks (bin/slemm_gwa.py:322-431), --iter_weighting
    with the reference's defaults (src/option.cpp:417-528); then, on a synthetic fileset of n
    individuals x loader_m SNPs (random codes) written to a temporary directory, the streaming .bed
    loader (src/input.cpp:343-386) and --pred end to end (src/pred.cpp:4-151).  FP64 rates count
    the window systems' dominant products (Zw'Zw per window or block, Zb'x per tested SNP)."""
    import os
    import shutil
    import tempfile

    import numpy as np
    import torch
    from slemm_amd import gpu, gwas

    out = {}
    X = np.ones((n, 1), np.float32)
    Py, h2, vg, ve = sl["Py"], float(sl["h2"]), float(sl["vg"]), float(sl["ve"])
    pc, _, _, _, gm = g.counts()

    w = 1000
    snps = gwas.sample_gstat_snps(gm, n, 30)  # the reference's sample (src/reml.cpp:920-925)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gst = g.lmm_gstat(X, Py, h2, snps, window_size=w)
    dt = time.perf_counter() - t0
    out["lmm_gstat"] = {"snps": int(snps.size), "window": w, "seconds": round(dt, 3),
                        "fp64_TFLOP/s_window_systems": round(snps.size * float(n) * w * w / dt / 1e12, 2),
                        "correction_factor_xSx_over_xPx": round(gwas.correction_factor(gst), 4)}
    print(f"# next rows: --lmm gstat {dt:.2f} s", file=sys.stderr, flush=True)

    bs, t_end = 1000, min(m, 50_000) - 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xy, xx = g.gwa_window(X, Py, np.arange(m), bs, 0, t_end, ve / vg, m, 20.0)
    dt = time.perf_counter() - t0
    nt = t_end + 1
    nblk = -(-nt // bs) + 1
    out["gwa_window"] = {"tested_snps": nt, "block_size": bs, "seconds": round(dt, 3),
                         "snps_per_s": round(nt / dt, 1),
                         "fp64_TFLOP/s": round((nblk * float(n) * bs * bs + 2.0 * nt * n * bs) / dt / 1e12, 2),
                         "finite": int(np.isfinite(xx).sum())}
    print(f"# next rows: windowed scan {dt:.2f} s", file=sys.stderr, flush=True)

    f = pc.astype(np.float64) / (2.0 * n)
    maf = np.where(f > 0.5, 1.0 - f, f).astype(np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = g.iter_weighting_REML(X, y, maf)
    dt = time.perf_counter() - t0
    out["iter_weighting"] = {"seconds": round(dt, 3), "h2": round(float(r["h2"]), 6),
                             "clumped_snps": int(r["sig_snps"].size)}
    print(f"# next rows: --iter_weighting {dt:.2f} s", file=sys.stderr, flush=True)
    g.set_model()

    try:
        tmp = tempfile.mkdtemp(prefix="slemm_bench_")
    except OSError as e:
        out["bed_loader"] = out["pred"] = {"error": repr(e)}
        return out
    try:
        nb = (n + 3) // 4
        prefix = os.path.join(tmp, "syn")
        rng = np.random.default_rng(5)
        with open(prefix + ".bed", "wb") as fh:
            fh.write(bytes([0x6C, 0x1B, 0x01]))
            step = max(1, (64 << 20) // nb)
            for j0 in range(0, loader_m, step):
                fh.write(rng.integers(0, 256, size=(min(step, loader_m - j0), nb), dtype=np.uint8).tobytes())
        with open(prefix + ".fam", "w") as fh:
            fh.writelines(f"f{i} i{i} 0 0 0 -9\n" for i in range(n))
        with open(prefix + ".bim", "w") as fh:
            fh.writelines(f"1\tsnp{j}\t0\t{j + 1}\tA\tG\n" for j in range(loader_m))
        eff = rng.normal(size=loader_m) * 1e-3
        with open(prefix + ".est.csv", "w") as fh:
            fh.write("SNP,CHR,POS,A1,A2,MAF,HWEP,x,y,EFFECT\n")
            fh.writelines(f"snp{j},1,{j + 1},{'G' if j % 3 == 0 else 'A'},x,0,0,0,0,{float(eff[j])!r}\n"
                          for j in range(loader_m))
        file_bytes = 3 + loader_m * nb
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = gpu.Geno.load_bed(prefix + ".bed", n, loader_m)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        h.close()
        out["bed_loader"] = {"n": n, "m": loader_m, "file_bytes": file_bytes, "seconds": round(dt, 3),
                             "GB/s": round(file_bytes / dt / 1e9, 2),
                             "note": "file just written (in the page cache): host memory -> HBM, both layouts built"}
        print(f"# next rows: bed loader {dt:.2f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        npred, mused = gpu.pred_plink(prefix + ".est.csv", prefix, prefix + ".gebv.csv")
        dt = time.perf_counter() - t0
        out["pred"] = {"individuals": int(npred), "snps": int(mused), "seconds": round(dt, 3),
                       "note": "estimate/fam/bim parsing, bed streaming, FP64 GEBV and IID,GEBV output"}
    except OSError as e:  # e.g. no room for the 1 GB file: the two file rows are reported missing
        out.setdefault("bed_loader", {"error": repr(e)})
        out.setdefault("pred", {"error": repr(e)})
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def differentiation_run(g, n, m, y, vg, ve, nv=64, win_size=100):
    """dl/dsigma_g^2, dl/dsigma_e^2 and dl/d(window weights) of the REML log-likelihood through the
    autograd harness: window weights = windowed mean of a marginal chi-square scan (100-SNP
    windows) normalised to mean 1; forward = Lanczos seed + CG solves + SLQ on [y, nv probes],
    backward = one adjoint pass over the bed.  The harness's dl/dsigma^2 is checked against the
    device REML score on the same probes."""
    import numpy as np
    import torch
    from slemm_amd.autodiff import RemlProblem, loglik

    dev = torch.device("cuda")
    X = np.ones((n, 1), np.float32)
    # seeded marginal scan: chi2_j = (x_c'y_c)^2 / (x_c'x_c var(y))
    g.set_model(None, None, np.full((n, 1), 1 / np.sqrt(n), np.float32))
    yc = torch.from_numpy(y - y.mean()).to(dev)
    xy, xx = g.gwas_score(yc, skip_mac=0.0)
    chi2 = (xy.double() ** 2 / xx.double() / float(np.var(y))).nan_to_num(0.0)
    nwin = -(-m // win_size)
    win = torch.arange(m, device=dev) // win_size
    wsum = torch.zeros(nwin, dtype=torch.float64, device=dev).index_add_(0, win, chi2)
    wcnt = torch.bincount(win, minlength=nwin).double()
    wmean = wsum / wcnt
    w0 = (wmean / wmean.mean()).clamp_min(1e-3)
    gen = torch.Generator(device=dev).manual_seed(17)
    V = torch.where(torch.rand((n, nv), generator=gen, device=dev) < 0.5, -1.0, 1.0) / np.sqrt(n)
    prob = RemlProblem(g, X, y, probes=V, tol=1e-5)
    # the first forward/backward pays for the workspaces of the 2(nv+1)-column adjoint block and
    # the first launches (timed apart as first_call_s); the timed pair is the second
    times = []
    for _ in range(2):
        w = w0.clone().requires_grad_(True)
        a = torch.tensor(vg, dtype=torch.float64, device=dev, requires_grad=True)
        b = torch.tensor(ve, dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ll = loglik(prob, w, a, b, win)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ll.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t1))
    # the device score on the same probes, with the same window weights as persnp
    g.set_model(w0[win].float().cpu().numpy(), None, None)
    score, _ = g.reml_score(X, y, vg, ve, probes=V, tol_L=1e-5)
    got = np.array([float(a.grad), float(b.grad)])
    g.set_model()
    return {"workload": f"n={n} m={m}, {nwin} windows of {win_size} SNPs, {nv} probes + y",
            "forward_s": round(times[1][0], 3), "backward_s": round(times[1][1], 3),
            "first_call_s": {"forward": round(times[0][0], 3), "backward": round(times[0][1], 3)},
            "loglik": float(ll.detach()), "dl_dsigma_g2": got[0], "dl_dsigma_e2": got[1],
            "dl_dw_norm": float(torch.linalg.vector_norm(w.grad)),
            "device_score": [float(score[0]), float(score[1])],
            "score_rel_diff": [float(abs(got[k] - score[k]) / max(abs(score[k]), 1e-30)) for k in range(2)],
            "euler_check_w_dot_grad": float(w.grad @ w0)}


def scale_run(n, m, nv, n_causal, h2, seed=5, cpu=True):
    """Whole-pipeline run at biobank shape: synthetic bed with allele frequencies drawn from
    Beta(0.8, 0.8) and kept where MAF >= 0.01, phenotype y = Zc beta + e with beta ~ N(0,1) on
    n_causal SNPs and var(e) set for h2; then SLDF-REML with nv probes (src/reml.cpp:549-690),
    GEBV of every individual from the SNP BLUPs (src/pred.cpp:112-130) and the GRAMMAR-gamma
    score terms of every SNP (bin/slemm_gamma.py:306-319).  Wall-clock seconds per stage."""
    import numpy as np
    import torch
    from slemm_amd import gpu

    rng = np.random.default_rng(seed)
    p = rng.beta(0.8, 0.8, size=4 * m)
    p = p[np.minimum(p, 1 - p) >= 0.01][:m]
    assert p.size == m
    out = {"workload": (f"n={n} m={m}, MAF~Beta(0.8,0.8)>=0.01, h2={h2}, {n_causal} causal, "
                        f"SLDF-REML with {nv} probes + GEBV + GWAS score")}
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    g = gpu.Geno.synthesize_p(n, p, seed=seed)
    torch.cuda.synchronize()
    out["generate_s"] = round(time.perf_counter() - t_all, 3)
    g.set_model(None, None, None)
    y, beta, gv, rh2 = simulate_phenotype(g, n, m, n_causal, h2, seed + 1)
    X = np.ones((n, 1), np.float32)
    out["realised_h2"] = round(rh2, 5)
    at = g.autotune(nv + 1, reps=2)  # the SLDF blocks: nv probes + 1 RHS
    out["autotune"] = {"K": nv + 1, "kernel": at["kernel"], "pass1": at["pass1"], "pass2": at["pass2"]}

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv, precision="bf16x2")
    torch.cuda.synchronize()
    out["reml"] = {"seconds": round(time.perf_counter() - t0, 3), "seconds_seed": round(r["seconds_seed"], 3),
                   "lanczos_steps": list(r["steps"]), "criterion_evals": r["evals"],
                   "h2": round(float(r["h2"]), 5), "vg": float(r["vg"]), "ve": float(r["ve"])}

    out["lmc"] = lmc_timing(g, X, y, float(r["h2"]))
    if cpu:  # reference-structured CPU operator on one 1000-SNP panel, extrapolated per call x steps
        rows = g.download(0, 1000)
        gm, gw, _ = g.model()
        per = cpu_op_seconds(rows, n, gm[:1000], gw[:1000], sorted({1, nv}), reps=1)
        st = out["reml"]["lanczos_steps"]
        out["cpu_sldf_seed_s_extrapolated"] = {
            "seconds": round((st[0] * per[1] + st[1] * per[1] + st[2] * per[nv]) * m / 1000, 1),
            "per_call_s": {str(t): round(v * m / 1000, 2) for t, v in per.items()},
            "how": "one 1000-SNP panel timed on the host cores, x m/1000, x the GPU fit's Lanczos steps"}
    alpha = torch.from_numpy(r["snp_blup"].astype(np.float64)).cuda()
    g.gebv(alpha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gebv = g.gebv(alpha)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["gebv"] = {"ms": round(dt * 1e3, 3),
                   "GB/s": round((m * ((n + 3) // 4) + 8 * (m + n)) / dt / 1e9, 1),
                   "corr_with_true_genetic_value": round(float(np.corrcoef(gebv.cpu().numpy(), gv)[0, 1]), 4)}
    g.gebv_fast(alpha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fast = g.gebv_fast(alpha)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["gebv_fast"] = {"ms": round(dt * 1e3, 3),
                        "GB/s": round((m * ((n + 3) // 4) + 8 * (m + n)) / dt / 1e9, 1),
                        "max_rel_err_vs_exact": float((fast - gebv).abs().max() / gebv.abs().max())}

    Py = torch.from_numpy(r["Py"]).cuda()
    g.gwas_score(Py)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xy, xx = g.gwas_score(Py)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s2 = (xy.double() ** 2 / xx.double()).cpu().numpy()   # chi-square up to the gamma factor
    causal = beta != 0
    out["gwas_score"] = {"ms": round(dt * 1e3, 3), "snps": m,
                         "score_enrichment_causal_over_null":
                             round(float(np.nanmean(s2[causal]) / np.nanmean(s2[~causal])), 3)}
    torch.cuda.synchronize()
    out["total_s"] = round(time.perf_counter() - t_all, 3)
    g.close()
    return out


def traits_run(n, m, T, nv, seed=9, fp32=True, random_subsets=False):
    """Batching: T traits on individual subsets of one shared bed (subset sizes log-uniform in
    [n/20, n], h2 ~ U(0.1, 0.7), 1,000 causal SNPs each), SLDF-REML with nv probes + 1 RHS + the
    intercept per trait.  Batched: all T*(nv+2) columns in one block product per pass, one HIP
    graph per Lanczos step.  Baselines: the same driver one trait at a time on per-subset handles,
    split-BF16 and plain FP32.  Subsets are the first n_t individuals (individuals are
    exchangeable in the synthetic bed), which lets the batched products skip the work past each
    column group's last member; random_subsets draws each trait's individuals at random instead
    (every column group then spans nearly all individuals)."""
    import numpy as np
    import torch
    from slemm_amd import gpu

    rng = np.random.default_rng(seed)
    sizes = np.sort(np.exp(rng.uniform(np.log(n / 20), np.log(n), size=T)).astype(np.int64))[::-1]
    h2s = rng.uniform(0.1, 0.7, size=T)
    g = gpu.Geno.synthesize(n, m, seed=seed + 1, maf_lo=0.05, maf_hi=0.5, missing=0.0)
    g.set_model()
    U = np.zeros((m, T), np.float32)
    for t in range(T):
        U[rng.choice(m, size=1000, replace=False), t] = rng.normal(size=1000)
    gv = g.matvec(torch.from_numpy(U).cuda(), precision="fp32").double().cpu().numpy()
    Y = np.zeros((T, n), np.float32)
    masks = np.zeros((T, n), bool)
    members = []
    for t in range(T):
        s = int(sizes[t])
        idx = np.sort(rng.choice(n, size=s, replace=False)) if random_subsets else np.arange(s)
        members.append(idx)
        masks[t, idx] = True
        gt = gv[idx, t] - gv[idx, t].mean()
        ve = gt.var() * (1 - h2s[t]) / h2s[t]
        Y[t, idx] = (gt + rng.normal(scale=np.sqrt(ve), size=s) + 10.0).astype(np.float32)
    out = {"workload": (f"n={n} m={m}, {T} traits on {'random' if random_subsets else 'leading'} subsets of "
                        f"{int(sizes.min())}..{int(sizes.max())} individuals (sum {int(sizes.sum())}), "
                        f"SLDF-REML with {nv} probes + RHS + intercept each"),
           "columns": int(T * (nv + 2))}
    at = g.autotune(T * (nv + 2), reps=2)  # the batched block: every trait's columns
    out["autotune_batched"] = {"K": T * (nv + 2), "kernel": at["kernel"], "pass1": at["pass1"], "pass2": at["pass2"]}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tr = gpu.Traits(g, masks)
    torch.cuda.synchronize()
    out["traits_create_s"] = round(time.perf_counter() - t0, 3)
    plan, ratio = tr.plan(nv)
    t0 = time.perf_counter()
    ra = tr.SLDF_REML(Y, n_V=nv)  # the library's automatic plan
    torch.cuda.synchronize()
    out["traits_api_auto"] = {"plan": plan, "work_ratio_batched_over_one_at_a_time": round(ratio, 3),
                              "seconds": round(time.perf_counter() - t0, 3),
                              "seconds_seed": round(float(ra["seconds_seed"]), 3)}
    # both plans of the library, timed alike (the other one run explicitly)
    plans = {plan: (out["traits_api_auto"]["seconds"], ra)}
    other = "sequential" if plan == "batched" else "batched"
    tr.set_plan(other)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ro = tr.SLDF_REML(Y, n_V=nv)
    torch.cuda.synchronize()
    plans[other] = (round(time.perf_counter() - t0, 3), ro)
    tr.set_plan("auto")
    r = plans["batched"][1]
    out["batched_bf16x2"] = {"seconds": plans["batched"][0], "seconds_seed": round(float(r["seconds_seed"]), 3)}
    out["one_at_a_time_bf16x2"] = {"seconds": plans["sequential"][0],
                                   "seconds_seed": round(float(plans["sequential"][1]["seconds_seed"]), 3),
                                   "note": "the library's plan: each trait compacted on the device, then SLDF"}
    out["h2_max_abs_diff_auto_vs_batched"] = float(np.max(np.abs(ra["h2"] - r["h2"])))
    out["h2_max_abs_diff_plans"] = float(np.max(np.abs(plans["batched"][1]["h2"] - plans["sequential"][1]["h2"])))
    out["batched_bf16x2"]["lanczos_steps_max"] = int(r["steps"].max())
    tr.close()
    seq = {}
    for prec in (("bf16x2", "fp32") if fp32 else ("bf16x2",)):
        h2 = np.zeros(T)
        tot = 0.0
        for t in range(T):
            s = int(sizes[t])
            gk = g.keep(members[t])
            gk.set_model()
            gk.autotune(nv + 1, reps=2)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rt = gk.SLDF_REML(np.ones((s, 1), np.float32), Y[t, members[t]], n_V=nv, precision=prec)
            torch.cuda.synchronize()
            tot += time.perf_counter() - t0
            h2[t] = rt["h2"]
            gk.close()
        seq[prec] = h2
        out[f"sequential_{prec}"] = {"seconds": round(tot, 3),
                                     "note": "per-subset handles built outside the timing, one SLDF each"}
    out["speedup_batched_over_sequential_bf16x2"] = round(out["sequential_bf16x2"]["seconds"] /
                                                         out["batched_bf16x2"]["seconds"], 2)
    out["speedup_auto_over_sequential_bf16x2"] = round(out["sequential_bf16x2"]["seconds"] /
                                                      out["traits_api_auto"]["seconds"], 2)
    out["speedup_auto_over_other_plan"] = round(max(out["batched_bf16x2"]["seconds"],
                                                    out["one_at_a_time_bf16x2"]["seconds"]) /
                                                out["traits_api_auto"]["seconds"], 2)
    if fp32:
        out["speedup_batched_over_sequential_fp32"] = round(out["sequential_fp32"]["seconds"] /
                                                           out["batched_bf16x2"]["seconds"], 2)
    out["h2_max_abs_diff_batched_vs_sequential"] = float(np.max(np.abs(r["h2"] - seq["bf16x2"])))
    out["h2_mean_abs_error_vs_simulated"] = float(np.mean(np.abs(r["h2"] - h2s)))
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--probes", type=int, default=64)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--cpu-snps", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reml", action="store_true")
    ap.add_argument("--no-smallk", action="store_true")
    ap.add_argument("--no-diff", action="store_true", help="skip the autograd/adjoint run")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the rows either side of the path (--lmm, windowed scan, --iter_weighting, loader, --pred)")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the vendor-GEMM MFMA ceiling")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM byte counters)")
    ap.add_argument("--causal", type=int, default=5000)
    ap.add_argument("--h2", type=float, default=0.3)
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the n=1M m=60k REML + GEBV + GWAS pipeline run")
    ap.add_argument("--no-traits", action="store_true",
                    help="skip the 32-trait batched run (n=200k, m=300k)")
    ap.add_argument("--traits-n", type=int, default=200_000)
    ap.add_argument("--traits-m", type=int, default=300_000)
    ap.add_argument("--traits-T", type=int, default=32)
    ap.add_argument("--scale-n", type=int, default=1_000_000)
    ap.add_argument("--scale-m", type=int, default=60_000)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two child rocprofv3 --pmc runs that measure the dominant kernel's HBM bytes")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the n=5k m=50k SLDF/LMC run (GPU, and CPU end to end unless --no-cpu)")
    ap.add_argument("--mfma-sweep", action="store_true",
                    help="print the register-only MFMA issue rates (4 instructions x 1-8 chains x 1-4 waves per SIMD) and exit")
    ap.add_argument("--only-sldf", action="store_true",
                    help="time SLDF-REML alone (one untimed fit, then --steps fits) and print a line per fit")
    ap.add_argument("--only-op", action="store_true",
                    help="run only the headline operator steps (for profiling) and print their timing")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine's path runs on a single MI355X "
              "(SNP-sharded multi-GPU operation is a design note only); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from slemm_amd import _lib as L
    from slemm_amd import gpu

    lib = L.load()
    if args.mfma_sweep:
        torch.cuda.init()
        for r in mfma_ceiling(lib, L, [(i, c, w) for i in range(8) for c in (1, 2, 4, 8) for w in (1, 2, 4)]):
            print(json.dumps(r), flush=True)
        for r in mfma_ceiling(lib, L, [(i, c, 4) for i in range(8) for c in (4, 8)], scale=20):
            print(json.dumps(dict(r, long=True)), flush=True)
        return
    n, m, nv = args.n, args.m, args.probes
    K = nv + 1
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")

    t0 = time.perf_counter()
    g = gpu.Geno.synthesize(n, m, seed=2024, maf_lo=0.05, maf_hi=0.5, missing=0.0)
    torch.cuda.synchronize()
    t_load = time.perf_counter() - t0
    g.set_model(None, None, np.full((n, 1), 1.0 / np.sqrt(n), np.float32))

    if args.only_sldf:
        # SLDF-REML alone, as reml_time_to_converge runs it: one untimed fit, then --steps timed
        # fits (for a kernel trace of the time to converge)
        y, _, _, rh2 = simulate_phenotype(g, n, m, args.causal, args.h2, 11)
        X = np.ones((n, 1), np.float32)
        g.autotune(nv + 2, reps=2)
        g.SLDF_REML(X, y, rng_seed=0, n_V=nv)
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = g.SLDF_REML(X, y, rng_seed=0, n_V=nv)
            torch.cuda.synchronize()
            print(json.dumps({"workload": f"SLDF-REML n={n} m={m}, {nv} probes", "seconds": round(time.perf_counter() - t0, 4),
                              "seconds_seed": round(r["seconds_seed"], 4), "lanczos_steps": list(r["steps"]),
                              "h2": round(float(r["h2"]), 6)}), flush=True)
        return

    # block: column 0 = phenotype-like RHS (projected operator), 1..nv = Rademacher probes; stored
    # column-major (n x K, leading dimension n) as the C ABI and the reference's Eigen blocks hold it
    gen = torch.Generator(device=dev).manual_seed(7)
    V = torch.empty((K, n), dtype=torch.float32, device=dev).t()
    V[:, 0] = torch.randn(n, generator=gen, device=dev)
    V[:, 1:] = torch.where(torch.rand((n, nv), generator=gen, device=dev) < 0.5, -1.0, 1.0) / np.sqrt(n)
    kinds = ["quadform"] + ["ldet"] * nv
    tau0 = (1 - 0.7) / 0.7

    def step(prec):
        return g.lnr_op_select("ldet", tau0, V, precision=prec, col_kind=kinds)

    if args.copy_calibration:  # 1 GiB device-to-device copy: the byte counters' reference kernel
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tune = None if args.no_autotune else g.autotune(K, reps=2)
    i8 = i8_on(g, K)

    for _ in range(args.warmup):
        step("bf16x2")
    torch.cuda.synchronize()
    lib.slemm_gpu_timing_reset()
    lib.slemm_gpu_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step("bf16x2")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    lib.slemm_gpu_timing(0)
    k1, _ = kernel_ms(lib, 0, L.PREC_BF16X2)
    k2, _ = kernel_ms(lib, 1, L.PREC_BF16X2)
    lib.slemm_gpu_timing_reset()
    if args.only_op:
        print(json.dumps({"workload": f"operator K={K}, n={n}, m={m}", "ms_per_step": round(dt * 1e3, 3),
                          "pass1_ms": round(k1, 3), "pass2_ms": round(k2, 3)}), flush=True)
        return

    # the same operator on the BF16 MFMA kernels (split-BF16 operand, FP32 accumulation)
    bf = None
    if i8:
        with bf16_mfma_kernels():
            tune_bf = None if args.no_autotune else g.autotune(K, reps=2)
            for _ in range(max(args.warmup, 1)):
                step("bf16x2")
            torch.cuda.synchronize()
            lib.slemm_gpu_timing(1)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step("bf16x2")
            torch.cuda.synchronize()
            dtb = (time.perf_counter() - t0) / args.steps
            lib.slemm_gpu_timing(0)
            b1, _ = kernel_ms(lib, 0, L.PREC_BF16X2)
            b2, _ = kernel_ms(lib, 1, L.PREC_BF16X2)
            lib.slemm_gpu_timing_reset()
        bf = {"ms_per_step": round(dtb * 1e3, 3), "GB/s": round(2 * pass_bytes(n, m, K) / dtb / 1e9, 2),
              "pass1_ms": round(b1, 3), "pass2_ms": round(b2, 3),
              "kernels": None if tune_bf is None else [tune_bf["pass1"][:2], tune_bf["pass2"][:2]]}

    # plain FP32 path beside it
    step("fp32")
    torch.cuda.synchronize()
    lib.slemm_gpu_timing(1)
    t0 = time.perf_counter()
    for _ in range(args.fp32_steps):
        step("fp32")
    torch.cuda.synchronize()
    dt32 = (time.perf_counter() - t0) / args.fp32_steps
    lib.slemm_gpu_timing(0)
    f1, _ = kernel_ms(lib, 0, L.PREC_FP32)
    f2, _ = kernel_ms(lib, 1, L.PREC_FP32)
    lib.slemm_gpu_timing_reset()

    reml = diff = nxt = None
    if not args.no_reml:
        y, _, _, rh2 = simulate_phenotype(g, n, m, args.causal, args.h2, 11)
        reml, sl = reml_time_to_converge(g, n, m, nv, y, rh2, args.causal, args.h2)
        if not args.no_diff:
            diff = differentiation_run(g, n, m, y, float(sl["vg"]), float(sl["ve"]), nv=nv)
        if not args.no_next:
            print("# next rows: --lmm, windowed scan, --iter_weighting, loader, --pred", file=sys.stderr, flush=True)
            try:
                nxt = next_rows_run(g, n, m, y, sl)
            except Exception as e:  # reported, never fatal to the bench line
                nxt = {"error": repr(e)}
                g.set_model()
    smallk = None if args.no_smallk else small_k_passes(g, lib, L, n, m)

    ceiling = None if args.no_ceiling else gemm_ceiling()
    measured_ceiling = mfma_ceiling_summary(lib, L)
    traffic = None if args.no_traffic else measured_traffic(n, m, K, "k_bp_i8w|k_bp_i8x" if i8 else "k_bp_",
                                                               env_extra=i8_target_env(tune))
    dom_pass = "pass2" if k2 >= k1 else "pass1"
    traffic_bytes = (traffic or {}).get(dom_pass)
    pb = pass_bytes(n, m, K)
    value = 2 * pb / dt / 1e9
    flops = 2.0 * n * m * K
    dom_ms = max(k1, k2)
    dom = "pass2 Z U (individual-major)" if k2 >= k1 else "pass1 Z'V (SNP-major)"
    ach_tf = flops / (dom_ms * 1e-3) / 1e12

    result = {
        "metric": "REML time-to-converge + bed-matvec HBM GB/s, n=100k m=500k, split-BF16 vs FP32",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("i8x3 (FP32 operand rendered per vector as three int8 digits of a 23-bit fixed-point value "
                  "relative to the vector's power-of-two scale, genotype dosages exact in int8, exact int32 "
                  "accumulation)" if i8 else
                  "bf16x2 (genotypes exact in BF16, vector block split hi+lo, FP32 accumulate)"),
        "data": "synthetic bed generated in HBM (MAF~U(0.05,0.5), HWE), Rademacher probes",
        "config": {"workload": f"Zc W Zc' operator on a K={K} block (64 probes + 1 RHS)",
                   "n": n, "m": m, "K": K, "bed_bytes": m * ((n + 3) // 4)},
        "kernel_ms": {"bf16x2_pass1": round(k1, 3), "bf16x2_pass2": round(k2, 3),
                      "fp32_pass1": round(f1, 3), "fp32_pass2": round(f2, 3)},
        "fp32_path": {"ms_per_step": round(dt32 * 1e3, 3), "GB/s": round(2 * pb / dt32 / 1e9, 2),
                      "speedup_bf16x2_over_fp32": round(dt32 / dt, 2)},
        "bf16_mfma_kernels": (None if bf is None else
                              dict(bf, speedup_over_fp32=round(dt32 * 1e3 / bf["ms_per_step"], 2))),
        "bed_generate_s": round(t_load, 3),
        "autotune": tune,
        "reml_time_to_converge": reml,
        "differentiation": diff,
        "next_rows": nxt,
        "small_k_passes": smallk,
        "roofline": ({
            "bound": "mfma",
            "kernel": f"int8 wide block product (k_bp_i8x, three fixed-point digits per vector), {dom}",
            "achieved": round(ach_tf, 2),
            "peak": INT8_PEAK_TOPS,
            "peak_kind": "dense int8 MFMA, 5.0 POP/s; achieved counts the product's 2 n m K operations",
            "unit": "TFLOP/s",
            "frac": round(ach_tf / INT8_PEAK_TOPS, 4),
            "traffic": None if traffic_bytes is None else round(traffic_bytes),
            "traffic_measured": traffic,
            "algorithmic_per_launch": {"flop": flops, "bytes": pb},
            "hbm_GB/s_of_kernel": round(pb / (dom_ms * 1e-3) / 1e9, 1),
            "mfma_issue_factor": round(i8_issue_factor(K), 3),
            "issued_TOP/s": round(ach_tf * i8_issue_factor(K), 1),
            "issued_frac_of_measured_ceiling": round(ach_tf * i8_issue_factor(K) /
                                                     measured_ceiling["i8_16x16x64"]["TOP/s"], 3),
            "issued_frac_of_measured_ceiling_random_operands": round(
                ach_tf * i8_issue_factor(K) / measured_ceiling["i8_16x16x64_random"]["TOP/s"], 3),
            "measured_mfma_ceiling": measured_ceiling,
            "vendor_gemm": ceiling,
        } if i8 else {
            "bound": "mfma",
            "kernel": f"split-BF16 block product ({(tune or {}).get('pass2' if k2 >= k1 else 'pass1', ['pipe'])[0]} "
                      f"variant), {dom}",
            "achieved": round(ach_tf, 2),
            "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(ach_tf / BF16_PEAK_TFLOPS, 4),
            "traffic": None if traffic_bytes is None else round(traffic_bytes),
            "traffic_measured": traffic,
            "algorithmic_per_launch": {"flop": flops, "bytes": pb},
            "hbm_GB/s_of_kernel": round(pb / (dom_ms * 1e-3) / 1e9, 1),
            "mfma_issue_factor": round(issue_factor(K), 3),
            "issued_TFLOP/s": round(ach_tf * issue_factor(K), 1),
            "issued_frac_of_measured_ceiling": round(ach_tf * issue_factor(K) /
                                                     measured_ceiling["bf16_16x16x32"]["TOP/s"], 3),
            "issued_frac_of_measured_ceiling_random_operands": round(
                ach_tf * issue_factor(K) / measured_ceiling["bf16_16x16x32_random"]["TOP/s"], 3),
            "measured_mfma_ceiling": measured_ceiling,
            "vendor_gemm": ceiling,
        }),
    }
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(g, n, m, K, args.cpu_snps, reps=2, reml=reml)
        except Exception as e:  # reported, never fatal
            result["cpu_baseline"] = {"value": None, "error": repr(e)}
    else:
        result["cpu_baseline"] = None
    if not args.no_c1:
        print("# C1: n=5000 m=50000 SLDF/LMC, GPU and CPU", file=sys.stderr, flush=True)
        try:
            result["c1"] = c1_run(cpu=not args.no_cpu)
        except Exception as e:  # reported, never fatal to the bench line
            result["c1"] = {"error": repr(e)}
    if not args.no_scale:
        g.close()
        del V
        torch.cuda.empty_cache()
        print(f"# scale pipeline n={args.scale_n} m={args.scale_m}", file=sys.stderr, flush=True)
        try:
            result["scale_pipeline"] = scale_run(args.scale_n, args.scale_m, 100, 2000, 0.4, cpu=not args.no_cpu)
        except Exception as e:  # reported, never fatal to the bench line
            result["scale_pipeline"] = {"error": repr(e)}
            torch.cuda.empty_cache()
    if not args.no_traits:
        if args.no_scale:
            g.close()
            del V
        torch.cuda.empty_cache()
        print(f"# batched traits n={args.traits_n} m={args.traits_m} T={args.traits_T}", file=sys.stderr, flush=True)
        try:
            result["batched_traits"] = traits_run(args.traits_n, args.traits_m, args.traits_T, 32)
        except Exception as e:  # reported, never fatal to the bench line
            result["batched_traits"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        try:
            result["batched_traits_random_subsets"] = traits_run(args.traits_n, args.traits_m, args.traits_T, 32,
                                                                 fp32=False, random_subsets=True)
        except Exception as e:  # reported, never fatal to the bench line
            result["batched_traits_random_subsets"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
