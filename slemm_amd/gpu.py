"""Python mirror of the reference's operator and Krylov interfaces over the gfx950 engine.

Names, argument meaning and error behaviour follow src/reml.cpp: ``lnr_op_select`` (:96-118),
``rmatvec`` (:159-195), ``weighted_rmatvec`` (:120-157), ``matvec`` (:197-233),
``L_Seed`` (:246-339), ``L_Solve`` (:369-431), ``SLQ_LDet`` (:440-451), and the GEBV loop of
src/pred.cpp:112-130.  Matrices are torch float32 tensors of shape (rows, cols) on the GPU;
internally they are handed to the C ABI column-major (a contiguous (cols, rows) buffer), the
layout of the reference's Eigen ``MatrixXf``.  PyTorch is only the device-memory carrier.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L

PRECISIONS = {"fp32": L.PREC_FP32, "bf16x2": L.PREC_BF16X2}
OPS = {"ldet": L.OP_LDET, "quadform": L.OP_QUADFORM}


def _prec(p) -> int:
    if isinstance(p, str):
        if p not in PRECISIONS:
            raise ValueError(f"unknown precision {p!r}")
        return PRECISIONS[p]
    return int(p)


def _colmajor(x: torch.Tensor, rows: int) -> torch.Tensor:
    """(rows, t) or (rows,) float32 CUDA tensor -> contiguous (t, rows) buffer."""
    if x.dim() == 1:
        x = x[:, None]
    if x.shape[0] != rows:
        raise ValueError(f"expected {rows} rows, got {tuple(x.shape)}")
    if not x.is_cuda or x.dtype != torch.float32:
        raise TypeError("expected a float32 CUDA tensor")
    return x.t().contiguous()


def _ptr(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _np_ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Geno:
    """Device-resident packed genotypes (both layouts) plus the operator's model terms."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        lib = L.load()
        n, m = C.c_int64(), C.c_int64()
        L.check(lib.slemm_gpu_geno_dims(self._h, C.byref(n), C.byref(m)))
        self.n, self.m = n.value, m.value

    # ------------------------------------------------------------------ construction
    @classmethod
    def from_bed(cls, bed_rows: np.ndarray, n_raw: int, keep_idx: Optional[Sequence[int]] = None,
                 snp_idx: Optional[Sequence[int]] = None) -> "Geno":
        """bed_rows: uint8 array (m_raw, ceil(n_raw/4)) -- the .bed body after the magic bytes."""
        lib = L.load()
        rows = np.ascontiguousarray(bed_rows, dtype=np.uint8)
        if rows.ndim != 2:
            raise ValueError("bed_rows must be 2-D (SNPs x bytes)")
        keep = None if keep_idx is None else np.ascontiguousarray(keep_idx, dtype=np.int64)
        snps = None if snp_idx is None else np.ascontiguousarray(snp_idx, dtype=np.int64)
        h = C.c_void_p()
        L.check(lib.slemm_gpu_geno_create(
            _np_ptr(rows, C.c_uint8), n_raw, rows.shape[0], rows.shape[1],
            _np_ptr(keep, C.c_int64), 0 if keep is None else len(keep),
            _np_ptr(snps, C.c_int64), 0 if snps is None else len(snps), C.byref(h)))
        return cls(h)

    @classmethod
    def load_bed(cls, bed_path: str, n_raw: int, m_raw: int, keep_idx: Optional[Sequence[int]] = None,
                 snp_idx: Optional[Sequence[int]] = None) -> "Geno":
        """Handle streamed from a PLINK .bed file (src/input.cpp:343-386 without host copies)."""
        h = C.c_void_p()
        keep = None if keep_idx is None else np.ascontiguousarray(keep_idx, dtype=np.int64)
        snps = None if snp_idx is None else np.ascontiguousarray(snp_idx, dtype=np.int64)
        L.check(L.load().slemm_gpu_geno_load_bed(
            str(bed_path).encode(), n_raw, m_raw,
            _np_ptr(keep, C.c_int64), 0 if keep is None else len(keep),
            _np_ptr(snps, C.c_int64), 0 if snps is None else len(snps), C.byref(h)))
        return cls(h)

    @classmethod
    def synthesize(cls, n: int, m: int, seed: int = 1, maf_lo: float = 0.05, maf_hi: float = 0.5,
                   missing: float = 0.0) -> "Geno":
        lib = L.load()
        h = C.c_void_p()
        L.check(lib.slemm_gpu_geno_synthesize(n, m, seed, maf_lo, maf_hi, missing, C.byref(h)))
        return cls(h)

    @classmethod
    def synthesize_p(cls, n: int, p: np.ndarray, seed: int = 1, missing: float = 0.0) -> "Geno":
        """Synthetic genotypes with the given per-SNP A2 allele frequencies (len(p) SNPs)."""
        lib = L.load()
        pf = np.ascontiguousarray(p, dtype=np.float32)
        h = C.c_void_p()
        L.check(lib.slemm_gpu_geno_synthesize_p(n, len(pf), seed, _np_ptr(pf, C.c_float), missing,
                                                C.byref(h)))
        return cls(h)

    def subset(self, snp_idx: Sequence[int]) -> "Geno":
        lib = L.load()
        idx = np.ascontiguousarray(snp_idx, dtype=np.int64)
        h = C.c_void_p()
        L.check(lib.slemm_gpu_geno_subset(self._h, _np_ptr(idx, C.c_int64), len(idx), C.byref(h)))
        return Geno(h)

    def keep(self, keep_idx: Sequence[int]) -> "Geno":
        """New handle with the given individuals (strictly increasing), compacted on the device."""
        idx = np.ascontiguousarray(keep_idx, dtype=np.int64)
        h = C.c_void_p()
        L.check(L.load().slemm_gpu_geno_keep(self._h, _np_ptr(idx, C.c_int64), len(idx), C.byref(h)))
        return Geno(h)

    def close(self) -> None:
        if self._h:
            L.check(L.load().slemm_gpu_geno_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # ------------------------------------------------------------------ inspection
    def download(self, j0: int = 0, cnt: Optional[int] = None) -> np.ndarray:
        cnt = self.m - j0 if cnt is None else cnt
        out = np.zeros((cnt, (self.n + 3) // 4), dtype=np.uint8)
        L.check(L.load().slemm_gpu_geno_download(self._h, j0, cnt, _np_ptr(out, C.c_uint8)))
        return out

    def counts(self):
        """(popcnt, n0, n1, n2, gm) host arrays -- src/input.cpp:12-21, src/reml.cpp:612-621."""
        m = self.m
        pc, n0, n1, n2 = (np.zeros(m, np.int32) for _ in range(4))
        gm = np.zeros(m, np.float32)
        L.check(L.load().slemm_gpu_geno_counts(self._h, _np_ptr(pc, C.c_int32), _np_ptr(n0, C.c_int32),
                                               _np_ptr(n1, C.c_int32), _np_ptr(n2, C.c_int32),
                                               _np_ptr(gm, C.c_float)))
        return pc, n0, n1, n2, gm

    def set_model(self, persnp: Optional[np.ndarray] = None, rvec: Optional[np.ndarray] = None,
                  Q: Optional[np.ndarray] = None) -> None:
        ps = None if persnp is None else np.ascontiguousarray(persnp, dtype=np.float32)
        rv = None if rvec is None else np.ascontiguousarray(rvec, dtype=np.float32)
        c = 0
        q = None
        if Q is not None:
            Qa = np.asarray(Q, dtype=np.float32)
            if Qa.ndim == 1:
                Qa = Qa[:, None]
            c = Qa.shape[1]
            q = np.asfortranarray(Qa)
        L.check(L.load().slemm_gpu_set_model(self._h, _np_ptr(ps, C.c_float), _np_ptr(rv, C.c_float),
                                             _np_ptr(q, C.c_float), self.n, c))

    def model(self):
        gm = np.zeros(self.m, np.float32)
        gw = np.zeros(self.m, np.float32)
        rr = np.zeros(self.n, np.float32)
        L.check(L.load().slemm_gpu_get_model(self._h, _np_ptr(gm, C.c_float), _np_ptr(gw, C.c_float),
                                             _np_ptr(rr, C.c_float)))
        return gm, gw, rr

    def set_gw(self, gw) -> None:
        """Replaces the operator's per-SNP weights W (after set_model) with gw (m,)."""
        g = np.ascontiguousarray(gw.detach().cpu().numpy() if torch.is_tensor(gw) else gw, dtype=np.float32)
        if g.shape != (self.m,):
            raise ValueError(f"gw must have shape ({self.m},)")
        L.check(L.load().slemm_gpu_set_gw(self._h, _np_ptr(g, C.c_float)))

    # ------------------------------------------------------------------ operators
    def lnr_op_select(self, lnr, tau0: float, v: torch.Tensor, precision="bf16x2",
                      col_kind: Optional[Sequence] = None) -> torch.Tensor:
        """src/reml.cpp:96-118 for all columns of v (n, t) in one batch; returns (n, t)."""
        if isinstance(lnr, str):
            if lnr not in OPS:
                raise ValueError("Unknown linear operator in lnr_op_select()")
            kind = OPS[lnr]
        else:
            kind = int(lnr)
        vc = _colmajor(v, self.n)
        t = vc.shape[0]
        out = torch.empty_like(vc)
        ck = None
        if col_kind is not None:
            ck_arr = (C.c_int * t)(*[OPS[x] if isinstance(x, str) else int(x) for x in col_kind])
            ck = C.cast(ck_arr, C.POINTER(C.c_int))
        L.check(L.load().slemm_gpu_op(self._h, kind, ck, tau0, _ptr(vc), self.n, t, _ptr(out), self.n,
                                      _prec(precision), _stream()))
        return out.t()

    def lnr_op_select_noalloc(self, lnr, tau0: float, v: torch.Tensor, out: torch.Tensor, precision="bf16x2",
                              col_kind: Optional[Sequence] = None) -> torch.Tensor:
        """lnr_op_select through slemm_gpu_op_noalloc: no device allocation inside the call, so the
        caller passes both column-major buffers, v and out, as contiguous (t, n) tensors."""
        kind = OPS[lnr] if isinstance(lnr, str) else int(lnr)
        if v.dim() != 2 or v.shape[1] != self.n or not v.is_contiguous() or v.dtype != torch.float32:
            raise ValueError("v must be a contiguous float32 (t, n) tensor")
        if out.shape != v.shape or not out.is_contiguous() or out.dtype != torch.float32:
            raise ValueError("out must be a contiguous float32 tensor shaped like v")
        t = v.shape[0]
        ck = None
        if col_kind is not None:
            ck = C.cast((C.c_int * t)(*[OPS[x] if isinstance(x, str) else int(x) for x in col_kind]),
                        C.POINTER(C.c_int))
        L.check(L.load().slemm_gpu_op_noalloc(self._h, kind, ck, tau0, _ptr(v), self.n, t, _ptr(out), self.n,
                                              _prec(precision), _stream()))
        return out

    def rmatvec(self, v: torch.Tensor, precision="bf16x2", weighted: bool = False,
                scale_rr: bool = False) -> torch.Tensor:
        """(Z-m)'v (src/reml.cpp:159-195); W(Z-m)'v when weighted (:120-157).  (n,t) -> (m,t)."""
        vc = _colmajor(v, self.n)
        t = vc.shape[0]
        out = torch.empty((t, self.m), dtype=torch.float32, device=vc.device)
        L.check(L.load().slemm_gpu_zt(self._h, _ptr(vc), self.n, t, _ptr(out), self.m, int(weighted),
                                      int(scale_rr), _prec(precision), _stream()))
        return out.t()

    def weighted_rmatvec(self, v: torch.Tensor, precision="bf16x2") -> torch.Tensor:
        return self.rmatvec(v, precision, weighted=True)

    def matvec(self, u: torch.Tensor, precision="bf16x2") -> torch.Tensor:
        """(Z-m)u (src/reml.cpp:197-233).  (m,t) -> (n,t)."""
        uc = _colmajor(u, self.m)
        t = uc.shape[0]
        out = torch.empty((t, self.n), dtype=torch.float32, device=uc.device)
        L.check(L.load().slemm_gpu_z(self._h, _ptr(uc), self.m, t, _ptr(out), self.n, _prec(precision),
                                     _stream()))
        return out.t()

    def gebv(self, alpha: torch.Tensor, flip: Optional[np.ndarray] = None) -> torch.Tensor:
        """sum_j alpha_j z_ij over raw dosages, z -> 2-z where flip (src/pred.cpp:112-130)."""
        if alpha.dtype != torch.float64 or not alpha.is_cuda or alpha.numel() != self.m:
            raise TypeError("alpha must be a float64 CUDA tensor of length m")
        a = alpha.contiguous()
        out = torch.empty(self.n, dtype=torch.float64, device=a.device)
        fl = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
        L.check(L.load().slemm_gpu_gebv(self._h, _ptr(a), _np_ptr(fl, C.c_uint8), _ptr(out), _stream()))
        return out

    def gebv_fast(self, alpha: torch.Tensor, flip: Optional[np.ndarray] = None) -> torch.Tensor:
        """The same GEBV at FP32 accuracy in one pass on the block-product kernels."""
        if alpha.dtype != torch.float64 or not alpha.is_cuda or alpha.numel() != self.m:
            raise TypeError("alpha must be a float64 CUDA tensor of length m")
        a = alpha.contiguous()
        out = torch.empty(self.n, dtype=torch.float64, device=a.device)
        fl = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
        L.check(L.load().slemm_gpu_gebv_fast(self._h, _ptr(a), _np_ptr(fl, C.c_uint8), _ptr(out), _stream()))
        return out

    def adjoint(self, A: torch.Tensor, B: torch.Tensor, coef, scale_rr: bool = True,
                precision="bf16x2") -> torch.Tensor:
        """out_j = sum_k coef_k (Zc'A)_jk (Zc'B)_jk (rows of A, B scaled by r_rsqrt first when
        scale_rr); A, B (n, t) float32 CUDA tensors.  Returns (m,) float64."""
        Ac = _colmajor(A, self.n)
        Bc = _colmajor(B, self.n)
        t = Ac.shape[0]
        if Bc.shape[0] != t:
            raise ValueError("A and B must have the same number of columns")
        cf = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        if cf.size != t:
            raise ValueError("need one coefficient per column")
        out = torch.empty(self.m, dtype=torch.float64, device=Ac.device)
        L.check(L.load().slemm_gpu_adjoint(self._h, _ptr(Ac), self.n, _ptr(Bc), self.n, t,
                                           _np_ptr(cf, C.c_double), int(scale_rr), _ptr(out), _prec(precision),
                                           _stream()))
        return out

    def gwas_score(self, Py: torch.Tensor, skip_mac: float = 20.0):
        """Per-SNP score terms of the GRAMMAR-gamma scan (bin/slemm_gamma.py:306-319):
        xy_j = x_cj' Py and xx_j = ||(I - QQ') R^-1/2 x_cj||^2 with the model's rvec and Q;
        SNPs whose minor-allele count is below skip_mac come back as NaN."""
        if Py.dtype != torch.float32 or not Py.is_cuda or Py.numel() != self.n:
            raise TypeError("Py must be a float32 CUDA tensor of length n")
        p = Py.contiguous()
        xy = torch.empty(self.m, dtype=torch.float32, device=p.device)
        xx = torch.empty(self.m, dtype=torch.float32, device=p.device)
        L.check(L.load().slemm_gpu_gwas_score(self._h, _ptr(p), skip_mac, _ptr(xy), _ptr(xx), _stream()))
        return xy, xx

    BLOCK_KERNELS = ("fp32", "bf16x2", "int8 narrow", "int8 one tile", "int8 wide")

    def block_kernel(self, t: int, precision="bf16x2") -> str:
        """Kernel family a t-vector operator block at this precision runs on."""
        k = C.c_int(-1)
        L.check(L.load().slemm_gpu_block_kernel(self._h, t, _prec(precision), C.byref(k)))
        return self.BLOCK_KERNELS[k.value]

    def autotune(self, t: int, reps: int = 2) -> dict:
        """Times the block-product settings for a t-column operator and keeps the fastest: the
        int8 kernels' split target (and narrow row tiles), or the split-BF16 kernel variants."""
        v = (C.c_int * 2)()
        tg = (C.c_int * 2)()
        kernel = self.block_kernel(t)
        nv = 10
        ms = (C.c_double * (2 + 6 * nv))(*([float("nan")] * (2 + 6 * nv)))
        L.check(L.load().slemm_gpu_autotune(self._h, t, reps, v, tg, ms))
        if kernel.startswith("int8"):
            nrt = 2 if kernel == "int8 narrow" else 1
            targets = [1024, 2048, 4096, 8192, 16384]
            table = {}
            for p in range(2):
                for ri in range(nrt):
                    for ti in range(5):
                        key = f"pass{p + 1}/" + (f"rt{[4, 8][ri]}/" if nrt == 2 else "") + f"{targets[ti]}"
                        table[key] = round(ms[2 + p * 3 * nv + ri * 5 + ti], 3)
            name = (lambda p: f"{kernel} rt{v[p]}") if nrt == 2 else (lambda p: kernel)
            return {"kernel": kernel, "pass1": (name(0), tg[0], round(ms[0], 3)),
                    "pass2": (name(1), tg[1], round(ms[1], 3)), "table_ms": table}
        names = ["reg", "lds4", "lds2", "reg8", "regl", "pipe", "persist", "pipel", "piper16", "piper32"]
        table = {f"pass{p + 1}/{names[vv]}/{[512, 1024, 2048][ti]}": round(ms[2 + p * 3 * nv + vv * 3 + ti], 3)
                 for p in range(2) for vv in range(nv) for ti in range(3)}
        return {"kernel": kernel, "pass1": (names[v[0]], tg[0], round(ms[0], 3)),
                "pass2": (names[v[1]], tg[1], round(ms[1], 3)), "table_ms": table}

    # ------------------------------------------------------------------ REML drivers
    def SLDF_REML(self, X: np.ndarray, y: np.ndarray, rng_seed: int = 0, n_V: int = 30, tol_L: float = 5e-4,
                  s2max: float = 0.7, s2min: float = 1e-4, tol_VC: float = 1e-4, precision="bf16x2",
                  verbose: bool = False) -> dict:
        """SLDF_REML (src/reml.cpp:549-690); persnp and rvec come from set_model."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        Py = np.zeros(self.n, np.float32)
        blup = np.zeros(self.m, np.float32)
        blue = np.zeros(c, np.float32)
        vb = np.zeros((c, c), np.float32, order="F")
        st = L.RemlStats()
        L.check(L.load().slemm_gpu_sldf_reml(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), rng_seed, n_V, tol_L, s2max,
            s2min, tol_VC, _prec(precision), int(verbose), _np_ptr(Py, C.c_float), _np_ptr(blup, C.c_float),
            _np_ptr(blue, C.c_float), _np_ptr(vb, C.c_float), C.byref(st)))
        return dict(vg=st.vg, ve=st.ve, h2=st.h2, lrt=st.lrt, Py=Py, snp_blup=blup, blue=blue, var_blue=vb,
                    steps=tuple(st.lanczos_steps), evals=st.criterion_evals, seconds_seed=st.seconds_seed,
                    seconds_total=st.seconds_total)

    def LMC_REML(self, X: np.ndarray, y: np.ndarray, rng_seed: int = 0, n_MC: Optional[int] = None,
                 tol_L: float = 5e-4, s2max: float = 0.7, s2min: float = 1e-4, precision="bf16x2",
                 verbose: bool = False, secant_tol: float = 0.002, secant_max_iter: int = 9) -> dict:
        """LMC_REML (src/reml.cpp:730-882, without --lmm); n_MC default as src/option.cpp:414.
        secant_tol / secant_max_iter: the secant's stop (the reference's 0.002 and 9 by default)."""
        if n_MC is None:
            n_MC = max(min(int(4e9 / self.n / self.n), 30), 3)
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        Py = np.zeros(self.n, np.float32)
        blup = np.zeros(self.m, np.float32)
        blue = np.zeros(c, np.float32)
        st = L.RemlStats()
        L.check(L.load().slemm_gpu_lmc_reml_ex(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), rng_seed, n_MC, tol_L, s2max,
            s2min, secant_tol, secant_max_iter, _prec(precision), int(verbose), _np_ptr(Py, C.c_float),
            _np_ptr(blup, C.c_float), _np_ptr(blue, C.c_float), C.byref(st)))
        return dict(vg=st.vg, ve=st.ve, h2=st.h2, Py=Py, snp_blup=blup, blue=blue, steps=st.lanczos_steps[0],
                    evals=st.criterion_evals, n_MC=n_MC, seconds_seed=st.seconds_seed,
                    seconds_total=st.seconds_total)

    def iter_weighting_REML(self, X: np.ndarray, y: np.ndarray, maf: np.ndarray, persnp: Optional[np.ndarray] = None,
                            rvec: Optional[np.ndarray] = None, window_size: int = 20, sig_chisq: float = 1e6,
                            indep_r2: float = 0.5, lrt: bool = False, rng_seed: int = 0, n_V: int = 30,
                            n_MC: Optional[int] = None, tol_L: float = 5e-4, s2max: float = 0.7, s2min: float = 1e-4,
                            tol_VC: float = 1e-4, precision="bf16x2", verbose: bool = False) -> dict:
        """--iter_weighting (src/option.cpp:417-528); defaults as src/option.cpp:52-76, 303, 414."""
        if n_MC is None:
            n_MC = max(min(int(4e9 / self.n / self.n), 30), 3)
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        mf = np.ascontiguousarray(maf, np.float32)
        ps = None if persnp is None else np.ascontiguousarray(persnp, np.float32)
        rv = None if rvec is None else np.ascontiguousarray(rvec, np.float32)
        Py = np.zeros(self.n, np.float32)
        blup = np.zeros(self.m, np.float32)
        blue = np.zeros(c, np.float32)
        w = np.zeros(self.m, np.float32)
        sig = np.zeros(self.m, np.int64)
        nsig = C.c_int()
        st = L.RemlStats()
        L.check(L.load().slemm_gpu_iter_weighting_reml(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), _np_ptr(ps, C.c_float),
            _np_ptr(rv, C.c_float), _np_ptr(mf, C.c_float), window_size, sig_chisq, indep_r2, int(lrt), rng_seed,
            n_V, n_MC, tol_L, s2max, s2min, tol_VC, _prec(precision), int(verbose), _np_ptr(Py, C.c_float),
            _np_ptr(blup, C.c_float), _np_ptr(blue, C.c_float), _np_ptr(w, C.c_float), _np_ptr(sig, C.c_int64),
            C.byref(nsig), C.byref(st)))
        return dict(vg=st.vg, ve=st.ve, h2=st.h2, Py=Py, snp_blup=blup, blue=blue, persnp=w,
                    sig_snps=sig[:nsig.value].copy())

    def gwa_window(self, X: np.ndarray, y: np.ndarray, model_snps: np.ndarray, block_size: int, test_begin: int,
                   test_end: int, tau: float, model_snp_ct: int, mac: float = 0.0):
        """Windowed score terms of slemm_gwa step 2 (bin/slemm_gwa.py:322-407): xy, xx for the
        SNPs test_begin..test_end (inclusive)."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        ms = np.ascontiguousarray(model_snps, np.int64)
        t = test_end - test_begin + 1
        xy = np.zeros(t, np.float32)
        xx = np.zeros(t, np.float32)
        L.check(L.load().slemm_gpu_gwa_window(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), _np_ptr(ms, C.c_int64), ms.size,
            block_size, test_begin, test_end, tau, model_snp_ct, mac, _np_ptr(xy, C.c_float),
            _np_ptr(xx, C.c_float)))
        return xy, xx

    def lmm_gstat_fake(self, X: np.ndarray, Py: np.ndarray, h2: float, k: int, rng_seed: int = 0, n_MC: int = 10,
                       tol_L: float = 5e-4, precision="bf16x2", verbose: bool = False) -> dict:
        """--fake_geno quadratic forms (src/reml.cpp:886-918) after an LMC_REML fit with the same
        rng_seed and n_MC: xPy, xPx, xSx of k simulated markers (xBx is not defined for them)."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        pv = np.ascontiguousarray(Py, np.float32)
        out = {key: np.zeros(int(k), np.float32) for key in ("xPy", "xPx", "xSx")}
        L.check(L.load().slemm_gpu_lmm_gstat_fake(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(pv, C.c_float), h2, int(k), int(rng_seed), int(n_MC),
            tol_L, _prec(precision), int(verbose), *[_np_ptr(out[key], C.c_float) for key in ("xPy", "xPx", "xSx")]))
        return out

    def lmm_sample(self, num_qf_markers: int = 30) -> np.ndarray:
        """The reference's --lmm SNP sample (src/reml.cpp:920-925) on this handle's genotype means."""
        from .gwas import sample_gstat_snps
        gm, _, _ = self.model()
        return sample_gstat_snps(gm, self.n, num_qf_markers)

    def lmm_gstat(self, X: np.ndarray, Py: np.ndarray, h2: float, snps: Optional[np.ndarray] = None,
                  window_size: int = 1000, tol_L: float = 5e-4, precision="bf16x2", verbose: bool = False,
                  num_qf_markers: int = 30) -> dict:
        """Quadratic forms of --lmm (src/reml.cpp:884-982): xPy, xPx, xBx, xSx for the SNPs
        `snps` after a fit with heritability h2 that returned Py; snps=None takes the
        reference's own sample of num_qf_markers SNPs (returned under "snps")."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        pv = np.ascontiguousarray(Py, np.float32)
        if snps is None:
            idx = None
            k = int(self.lmm_sample(num_qf_markers).size)
        else:
            idx = np.ascontiguousarray(snps, np.int64)
            k = idx.size
        out = {key: np.zeros(k, np.float32) for key in ("xPy", "xPx", "xBx", "xSx")}
        L.check(L.load().slemm_gpu_lmm_gstat(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(pv, C.c_float), h2,
            _np_ptr(idx, C.c_int64), k, window_size, tol_L, _prec(precision), int(verbose),
            *[_np_ptr(out[key], C.c_float) for key in ("xPy", "xPx", "xBx", "xSx")]))
        out["snps"] = self.lmm_sample(num_qf_markers) if snps is None else idx.copy()
        return out

    def reml_score(self, X: np.ndarray, y: np.ndarray, vg: float, ve: float,
                   probes: Optional[torch.Tensor] = None, n_V: int = 30, rng_seed: int = 0,
                   tol_L: float = 5e-4, precision="bf16x2"):
        """(dl/dvg, dl/dve) and the AI matrix [[gg, ge], [ge, ee]] at (vg, ve); probes (n, n_V)
        float32 CUDA with E[vv'] = I/n, or Rademacher probes as SLDF_REML draws them."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        pc = None
        if probes is not None:
            pc = _colmajor(probes, self.n)
            n_V = pc.shape[0]
        score = np.zeros(2)
        ai = np.zeros(3)
        L.check(L.load().slemm_gpu_reml_score(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), vg, ve, _ptr(pc), self.n, n_V,
            rng_seed, tol_L, _prec(precision), _np_ptr(score, C.c_double), _np_ptr(ai, C.c_double)))
        return score, np.array([[ai[0], ai[1]], [ai[1], ai[2]]])

    def AI_REML(self, X: np.ndarray, y: np.ndarray, rng_seed: int = 0, n_V: int = 30, tol_L: float = 5e-4,
                s2max: float = 0.7, vg0: float = 0.0, ve0: float = 0.0, tol_VC: float = 1e-4,
                max_iter: int = 20, precision="bf16x2", verbose: bool = False) -> dict:
        """Average-information REML on (vg, ve) for the SLDF_REML model (no reference counterpart)."""
        Xf = np.asfortranarray(np.asarray(X, np.float32).reshape(self.n, -1))
        c = Xf.shape[1]
        yv = np.ascontiguousarray(y, np.float32)
        Py = np.zeros(self.n, np.float32)
        blup = np.zeros(self.m, np.float32)
        blue = np.zeros(c, np.float32)
        vb = np.zeros((c, c), np.float32, order="F")
        hist = np.zeros((max_iter, 7))
        st = L.RemlStats()
        L.check(L.load().slemm_gpu_ai_reml(
            self._h, _np_ptr(Xf, C.c_float), self.n, c, _np_ptr(yv, C.c_float), rng_seed, n_V, tol_L, s2max, vg0,
            ve0, tol_VC, max_iter, _prec(precision), int(verbose), _np_ptr(Py, C.c_float),
            _np_ptr(blup, C.c_float), _np_ptr(blue, C.c_float), _np_ptr(vb, C.c_float),
            _np_ptr(hist, C.c_double), C.byref(st)))
        it = st.criterion_evals
        return dict(vg=st.vg, ve=st.ve, h2=st.h2, Py=Py, snp_blup=blup, blue=blue, var_blue=vb,
                    iterations=it, history=hist[:it], steps=tuple(st.lanczos_steps),
                    seconds_seed=st.seconds_seed, seconds_total=st.seconds_total)

    # ------------------------------------------------------------------ Krylov engine
    def L_Seed(self, lnr, tau0: float, B: torch.Tensor, tol: float = 5e-4, maxit: int = 600,
               keep_cols: Optional[int] = None, precision="bf16x2", col_kind=None, col_group=None,
               verbose: bool = False) -> "LSeed":
        kind = OPS[lnr] if isinstance(lnr, str) else int(lnr)
        Bc = _colmajor(B, self.n)
        t = Bc.shape[0]
        keep = t if keep_cols is None else keep_cols
        ck = cg = None
        if col_kind is not None:
            ck = C.cast((C.c_int * t)(*[OPS[x] if isinstance(x, str) else int(x) for x in col_kind]),
                        C.POINTER(C.c_int))
        if col_group is not None:
            cg = C.cast((C.c_int * t)(*[int(x) for x in col_group]), C.POINTER(C.c_int))
        h = C.c_void_p()
        L.check(L.load().slemm_gpu_lseed(self._h, kind, ck, cg, tau0, _ptr(Bc), self.n, t, tol, maxit, keep,
                                         _prec(precision), int(verbose), _stream(), C.byref(h)))
        return LSeed(h, self.n, t, owner=self)


class LSeed:
    """struct lseed (src/reml.cpp:240-244) kept on the device.  `owner` (the Geno or Traits the
    seed was built on) is kept alive as long as the seed, so handles are released seed first."""

    def __init__(self, handle, n: int, t: int, owner=None):
        self._h = handle
        self._owner = owner
        self.n, self.t = n, t
        lib = L.load()
        tt, steps, conv = C.c_int(), C.c_int(), C.c_int()
        L.check(lib.slemm_gpu_lseed_info(self._h, C.byref(tt), C.byref(steps), C.byref(conv)))
        self.steps, self.converged = steps.value, bool(conv.value)
        cs = np.zeros(t, dtype=np.int32)
        L.check(lib.slemm_gpu_lseed_col_steps(self._h, cs.ctypes.data_as(C.POINTER(C.c_int))))
        self.col_steps = cs

    def coeffs(self):
        """beta (t, steps+1), delta (t, steps) as in struct lseed."""
        J = self.steps
        beta = np.zeros((J + 1, self.t), dtype=np.float64)
        delta = np.zeros((max(J, 1), self.t), dtype=np.float64)
        L.check(L.load().slemm_gpu_lseed_coeffs(self._h, _np_ptr(beta, C.c_double), _np_ptr(delta, C.c_double)))
        return beta.T.copy(), delta[:J].T.copy()

    def L_Solve(self, B: torch.Tensor, sigma: float, tol: float = 1e-6, col0: int = 0) -> torch.Tensor:
        Bc = _colmajor(B, self.n)
        t2 = Bc.shape[0]
        X = torch.empty_like(Bc)
        L.check(L.load().slemm_gpu_lsolve(self._h, col0, t2, _ptr(Bc), self.n, sigma, tol, _ptr(X), self.n,
                                          _stream()))
        return X.t()

    def spectrum(self, k: int):
        J = int(self.col_steps[k])
        D = np.zeros(J)
        W = np.zeros(J)
        L.check(L.load().slemm_gpu_lseed_spectrum(self._h, k, _np_ptr(D, C.c_double), _np_ptr(W, C.c_double)))
        return D, W

    def SLQ_LDet(self, sigma: float, col0: int = 0, nv: Optional[int] = None) -> float:
        nv = self.t - col0 if nv is None else nv
        out = C.c_double()
        L.check(L.load().slemm_gpu_slq_ldet(self._h, col0, nv, sigma, C.byref(out)))
        return out.value

    def close(self):
        if self._h:
            L.check(L.load().slemm_gpu_lseed_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Traits:
    """Several traits on individual subsets of one Geno, batched through one block product per
    pass (no reference counterpart; the reference analyses one trait per run).  masks: (T, n)
    booleans, row t = members of trait t's subset."""

    def __init__(self, geno: Geno, masks: np.ndarray):
        m8 = np.ascontiguousarray(np.asarray(masks, dtype=bool).astype(np.uint8))
        if m8.ndim != 2 or m8.shape[1] != geno.n:
            raise ValueError(f"masks must have shape (T, {geno.n})")
        self.geno = geno
        self.T = m8.shape[0]
        self.masks = m8.astype(bool)
        h = C.c_void_p()
        L.check(L.load().slemm_gpu_traits_create(geno.handle, self.T, _np_ptr(m8, C.c_uint8), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            L.check(L.load().slemm_gpu_traits_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nt = np.zeros(self.T, np.int64)
        mono = np.zeros(self.T, np.int64)
        L.check(L.load().slemm_gpu_traits_info(self._h, _np_ptr(nt, C.c_int64), _np_ptr(mono, C.c_int64)))
        return nt, mono

    def model(self, t: int):
        m = self.geno.m
        gm = np.zeros(m, np.float32)
        gw = np.zeros(m, np.float32)
        pc = np.zeros(m, np.int32)
        L.check(L.load().slemm_gpu_traits_model(self._h, t, _np_ptr(gm, C.c_float), _np_ptr(gw, C.c_float),
                                                _np_ptr(pc, C.c_int32)))
        return gm, gw, pc

    def lnr_op_select(self, col_trait: Sequence[int], col_kind: Optional[Sequence], tau0: float, v: torch.Tensor,
                      precision="bf16x2") -> torch.Tensor:
        """Batched lnr_op_select (src/reml.cpp:96-118): column k of v (n, t) belongs to trait
        col_trait[k] and gets that trait's operator of kind col_kind[k]."""
        vc = _colmajor(v, self.geno.n)
        t = vc.shape[0]
        ct = (C.c_int * t)(*[int(x) for x in col_trait])
        ck = None
        if col_kind is not None:
            ck = (C.c_int * t)(*[OPS[x] if isinstance(x, str) else int(x) for x in col_kind])
        out = torch.empty_like(vc)
        L.check(L.load().slemm_gpu_traits_op(self._h, ct, ck, tau0, _ptr(vc), self.geno.n, t, _ptr(out),
                                             self.geno.n, _prec(precision), _stream()))
        return out.t()

    def L_Seed(self, col_trait: Sequence[int], col_kind: Optional[Sequence], tau0: float, B: torch.Tensor,
               tol: float = 5e-4, maxit: int = 600, keep_cols: Optional[int] = None, col_group=None,
               precision="bf16x2", verbose: bool = False) -> "LSeed":
        Bc = _colmajor(B, self.geno.n)
        t = Bc.shape[0]
        keep = t if keep_cols is None else keep_cols
        ct = (C.c_int * t)(*[int(x) for x in col_trait])
        ck = None if col_kind is None else (C.c_int * t)(*[OPS[x] if isinstance(x, str) else int(x)
                                                           for x in col_kind])
        cg = None if col_group is None else (C.c_int * t)(*[int(x) for x in col_group])
        h = C.c_void_p()
        L.check(L.load().slemm_gpu_traits_lseed(self._h, ct, ck, cg, tau0, _ptr(Bc), self.geno.n, t, tol, maxit,
                                                keep, _prec(precision), int(verbose), _stream(), C.byref(h)))
        return LSeed(h, self.geno.n, t, owner=self)

    PLANS = {"auto": L.TRAITS_AUTO, "batched": L.TRAITS_BATCHED, "sequential": L.TRAITS_SEQUENTIAL}

    def set_plan(self, plan: str) -> None:
        """"auto" (default), "batched" or "sequential" (one trait at a time on compacted handles)."""
        if plan not in self.PLANS:
            raise ValueError(f"unknown plan {plan!r}")
        L.check(L.load().slemm_gpu_traits_set_plan(self._h, self.PLANS[plan]))

    def plan(self, n_V: int = 30) -> tuple:
        """(plan the next SLDF_REML with n_V probes runs, batched / one-at-a-time work ratio)."""
        pl = C.c_int(-1)
        r = C.c_double(0)
        L.check(L.load().slemm_gpu_traits_plan(self._h, n_V, C.byref(pl), C.byref(r)))
        return {v: k for k, v in self.PLANS.items()}[pl.value], r.value

    def SLDF_REML(self, y: np.ndarray, rng_seed: int = 0, n_V: int = 30, tol_L: float = 5e-4, s2max: float = 0.7,
                  s2min: float = 1e-4, tol_VC: float = 1e-4, precision="bf16x2", verbose: bool = False) -> dict:
        """SLDF_REML (src/reml.cpp:549-690) for all traits, on the handle's plan; y (T, n)."""
        yv = np.ascontiguousarray(np.asarray(y, np.float32).reshape(self.T, self.geno.n))
        vg, ve, h2, lrt = (np.zeros(self.T, np.float32) for _ in range(4))
        steps = np.zeros((self.T, 3), np.int32)
        secs = np.zeros(2)
        plan, _ = self.plan(n_V)
        L.check(L.load().slemm_gpu_traits_sldf_reml(
            self._h, _np_ptr(yv, C.c_float), rng_seed, n_V, tol_L, s2max, s2min, tol_VC, _prec(precision),
            int(verbose), _np_ptr(vg, C.c_float), _np_ptr(ve, C.c_float), _np_ptr(h2, C.c_float),
            _np_ptr(lrt, C.c_float), _np_ptr(steps, C.c_int), _np_ptr(secs, C.c_double)))
        return dict(vg=vg, ve=ve, h2=h2, lrt=lrt, steps=steps, seconds_seed=secs[0], seconds_total=secs[1],
                    plan=plan)


def rademacher_probes(rng_seed: int, n: int, n_V: int) -> np.ndarray:
    """The Rademacher probes of SLDF_REML (src/reml.cpp:597-605): n x n_V float32, column-major."""
    B = np.zeros((n, n_V), np.float32, order="F")
    L.check(L.load().slemm_gpu_rademacher_probes(rng_seed, n, n_V, _np_ptr(B, C.c_float)))
    return B


def lmc_draws(rng_seed: int, m: int, n: int, n_MC: int) -> tuple:
    """The Monte Carlo draws of LMC_REML (src/reml.cpp:788-801): u (m x n_MC) then e (n x n_MC),
    column-major float32, from mt19937(rng_seed + 54321) through normal_distribution<float>."""
    u = np.zeros((m, n_MC), np.float32, order="F")
    e = np.zeros((n, n_MC), np.float32, order="F")
    L.check(L.load().slemm_gpu_lmc_draws(rng_seed, m, n, n_MC, _np_ptr(u, C.c_float), _np_ptr(e, C.c_float)))
    return u, e


def pred_plink(estimate_file: str, bfile_prefix: str, output_file: str) -> tuple:
    """--pred (src/pred.cpp:4-151): writes IID,GEBV for the fileset; returns (n, SNPs used)."""
    n = C.c_int64()
    m = C.c_int64()
    L.check(L.load().slemm_gpu_pred_plink(str(estimate_file).encode(), str(bfile_prefix).encode(),
                                          str(output_file).encode(), C.byref(n), C.byref(m)))
    return n.value, m.value
