"""ctypes binding of ``libslemm_gpu.so`` (the C ABI declared in ``include/slemm_gpu.h``).

The library is the only compute path: importing this module fails loudly if it has not been
built (``python -c "import __graft_entry__ as g; g.build()"``); nothing falls back to a CPU
implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libslemm_gpu.so"

OK, ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_STATE, ERR_RANK, ERR_DATA, ERR_NUMERIC = range(8)
OP_LDET, OP_QUADFORM = 0, 1
PREC_FP32, PREC_BF16X2 = 0, 1
TRAITS_AUTO, TRAITS_BATCHED, TRAITS_SEQUENTIAL = 0, 1, 2

c_i64 = C.c_int64
c_i64p = C.POINTER(C.c_int64)
c_i32p = C.POINTER(C.c_int32)
c_intp = C.POINTER(C.c_int)
c_fp = C.POINTER(C.c_float)
c_dp = C.POINTER(C.c_double)
c_u8p = C.POINTER(C.c_uint8)
c_vp = C.c_void_p


class RemlStats(C.Structure):
    _fields_ = [
        ("vg", C.c_float), ("ve", C.c_float), ("h2", C.c_float), ("lrt", C.c_float),
        ("lanczos_steps", C.c_int * 3), ("criterion_evals", C.c_int),
        ("seconds_seed", C.c_double), ("seconds_total", C.c_double),
    ]


# name -> (restype, argtypes); every exported symbol of include/slemm_gpu.h
SIGNATURES = {
    "slemm_gpu_last_error": (C.c_char_p, []),
    "slemm_gpu_version": (C.c_int, []),
    "slemm_gpu_device_count": (C.c_int, [c_intp]),
    "slemm_gpu_geno_create": (C.c_int, [c_u8p, c_i64, c_i64, c_i64, c_i64p, c_i64, c_i64p, c_i64, C.POINTER(c_vp)]),
    "slemm_gpu_geno_subset": (C.c_int, [c_vp, c_i64p, c_i64, C.POINTER(c_vp)]),
    "slemm_gpu_geno_keep": (C.c_int, [c_vp, c_i64p, c_i64, C.POINTER(c_vp)]),
    "slemm_gpu_geno_synthesize": (C.c_int, [c_i64, c_i64, C.c_uint64, C.c_float, C.c_float, C.c_float, C.POINTER(c_vp)]),
    "slemm_gpu_geno_synthesize_p": (C.c_int, [c_i64, c_i64, C.c_uint64, C.POINTER(C.c_float), C.c_float, C.POINTER(c_vp)]),
    "slemm_gpu_geno_destroy": (C.c_int, [c_vp]),
    "slemm_gpu_geno_dims": (C.c_int, [c_vp, c_i64p, c_i64p]),
    "slemm_gpu_geno_download": (C.c_int, [c_vp, c_i64, c_i64, c_u8p]),
    "slemm_gpu_geno_counts": (C.c_int, [c_vp, c_i32p, c_i32p, c_i32p, c_i32p, c_fp]),
    "slemm_gpu_set_model": (C.c_int, [c_vp, c_fp, c_fp, c_fp, c_i64, C.c_int]),
    "slemm_gpu_get_model": (C.c_int, [c_vp, c_fp, c_fp, c_fp]),
    "slemm_gpu_op": (C.c_int, [c_vp, C.c_int, c_intp, C.c_float, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int, c_vp]),
    "slemm_gpu_op_noalloc": (C.c_int, [c_vp, C.c_int, c_intp, C.c_float, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int,
                                       c_vp]),
    "slemm_gpu_zt": (C.c_int, [c_vp, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int, C.c_int, C.c_int, c_vp]),
    "slemm_gpu_z": (C.c_int, [c_vp, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int, c_vp]),
    "slemm_gpu_gebv": (C.c_int, [c_vp, c_vp, c_u8p, c_vp, c_vp]),
    "slemm_gpu_gebv_fast": (C.c_int, [c_vp, c_vp, c_u8p, c_vp, c_vp]),
    "slemm_gpu_gwas_score": (C.c_int, [c_vp, c_vp, C.c_float, c_vp, c_vp, c_vp]),
    "slemm_gpu_lseed": (C.c_int, [c_vp, C.c_int, c_intp, c_intp, C.c_float, c_vp, c_i64, C.c_int, C.c_float,
                                  C.c_int, C.c_int, C.c_int, C.c_int, c_vp, C.POINTER(c_vp)]),
    "slemm_gpu_lseed_info": (C.c_int, [c_vp, c_intp, c_intp, c_intp]),
    "slemm_gpu_lseed_col_steps": (C.c_int, [c_vp, c_intp]),
    "slemm_gpu_lseed_coeffs": (C.c_int, [c_vp, c_dp, c_dp]),
    "slemm_gpu_lseed_destroy": (C.c_int, [c_vp]),
    "slemm_gpu_lsolve": (C.c_int, [c_vp, C.c_int, C.c_int, c_vp, c_i64, C.c_float, C.c_float, c_vp, c_i64, c_vp]),
    "slemm_gpu_lseed_spectrum": (C.c_int, [c_vp, C.c_int, c_dp, c_dp]),
    "slemm_gpu_slq_ldet": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_double, c_dp]),
    "slemm_gpu_sldf_reml": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.c_float, C.c_float, C.c_int, C.c_int, c_fp, c_fp, c_fp, c_fp,
                                      C.POINTER(RemlStats)]),
    "slemm_gpu_autotune": (C.c_int, [c_vp, C.c_int, C.c_int, c_intp, c_intp, c_dp]),
    "slemm_gpu_block_kernel": (C.c_int, [c_vp, C.c_int, C.c_int, c_intp]),
    "slemm_gpu_mfma_ceiling": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "slemm_gpu_timing": (C.c_int, [C.c_int]),
    "slemm_gpu_timing_read": (C.c_int, [C.c_int, C.c_int, c_dp, c_intp]),
    "slemm_gpu_timing_reset": (C.c_int, []),
    "slemm_gpu_lmc_reml": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_float, C.c_int, C.c_int, c_fp, c_fp, c_fp, C.POINTER(RemlStats)]),
    "slemm_gpu_lmc_reml_ex": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, c_fp, c_fp, c_fp,
                                        C.POINTER(RemlStats)]),
    "slemm_gpu_lmc_draws": (C.c_int, [C.c_int, c_i64, c_i64, C.c_int, c_fp, c_fp]),
    "slemm_gpu_rademacher_probes": (C.c_int, [C.c_int, c_i64, C.c_int, c_fp]),
    "slemm_gpu_set_gw": (C.c_int, [c_vp, c_fp]),
    "slemm_gpu_geno_load_bed": (C.c_int, [C.c_char_p, c_i64, c_i64, c_i64p, c_i64, c_i64p, c_i64, C.POINTER(c_vp)]),
    "slemm_gpu_pred_plink": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, c_i64p, c_i64p]),
    "slemm_gpu_gwa_window": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, c_i64p, C.c_int, C.c_int, c_i64, c_i64,
                                       C.c_float, c_i64, C.c_float, c_fp, c_fp]),
    "slemm_gpu_iter_weighting_reml": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, c_fp, c_fp, c_fp, C.c_int,
                                                C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                                c_fp, c_fp, c_fp, c_fp, c_i64p, C.POINTER(C.c_int),
                                                C.POINTER(RemlStats)]),
    "slemm_gpu_lmm_gstat": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_float, C.POINTER(C.c_int64), C.c_int,
                                      C.c_int, C.c_float, C.c_int, C.c_int, c_fp, c_fp, c_fp, c_fp]),
    "slemm_gpu_lmm_sample": (C.c_int, [c_fp, c_i64, c_i64, C.c_int, c_i64p, C.POINTER(C.c_int)]),
    "slemm_gpu_lmm_gstat_fake": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_float, C.c_int, C.c_int, C.c_int,
                                           C.c_float, C.c_int, C.c_int, c_fp, c_fp, c_fp]),
    "slemm_gpu_traits_create": (C.c_int, [c_vp, C.c_int, c_u8p, C.POINTER(c_vp)]),
    "slemm_gpu_traits_destroy": (C.c_int, [c_vp]),
    "slemm_gpu_traits_info": (C.c_int, [c_vp, c_i64p, c_i64p]),
    "slemm_gpu_traits_model": (C.c_int, [c_vp, C.c_int, c_fp, c_fp, c_i32p]),
    "slemm_gpu_traits_op": (C.c_int, [c_vp, c_intp, c_intp, C.c_float, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int,
                                      c_vp]),
    "slemm_gpu_traits_set_plan": (C.c_int, [c_vp, C.c_int]),
    "slemm_gpu_traits_plan": (C.c_int, [c_vp, C.c_int, c_intp, c_dp]),
    "slemm_gpu_traits_sldf_reml": (C.c_int, [c_vp, c_fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                             C.c_float, C.c_int, C.c_int, c_fp, c_fp, c_fp, c_fp, c_intp, c_dp]),
    "slemm_gpu_traits_lseed": (C.c_int, [c_vp, c_intp, c_intp, c_intp, C.c_float, c_vp, c_i64, C.c_int, C.c_float,
                                         C.c_int, C.c_int, C.c_int, C.c_int, c_vp, C.POINTER(c_vp)]),
    "slemm_gpu_adjoint": (C.c_int, [c_vp, c_vp, c_i64, c_vp, c_i64, C.c_int, c_dp, C.c_int, c_vp, C.c_int, c_vp]),
    "slemm_gpu_reml_score": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_double, C.c_double, c_vp, c_i64,
                                       C.c_int, C.c_int, C.c_float, C.c_int, c_dp, c_dp]),
    "slemm_gpu_ai_reml": (C.c_int, [c_vp, c_fp, c_i64, C.c_int, c_fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, c_fp, c_fp,
                                    c_fp, c_fp, c_dp, C.POINTER(RemlStats)]),
}

_lib = None


class SlemmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[slemm_gpu status {code}] {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load the library (once) and attach the ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: the HIP engine has not been built "
            "(run __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != OK:
        msg = load().slemm_gpu_last_error().decode(errors="replace")
        raise SlemmError(status, msg)
