"""ORACLE -- test infrastructure only.

CPU restatement of the reference's stochastic-Lanczos REML path, used as the checker by
``tests/``, by ``__graft_entry__.smoke()`` and as the CPU baseline of ``bench.py``.  The product
(``slemm_amd``) never imports this package.

Parts:
  * ``liboracle.so`` (``slemm_oracle.cpp``): FP32 restatement of src/input.cpp bed decode,
    src/weight.cpp / src/reml.cpp counts, the operators and Krylov engine of src/reml.cpp, the
    libstdc++ RNG draws and the GEBV loop of src/pred.cpp.
  * :mod:`oracle.dense`: FP64 dense oracle (explicit Zc W Zc', exact solves and log-dets).
  * :mod:`oracle.reml`: the SLDF / LMC REML drivers restated over the primitives above.
  * :mod:`oracle.bedio`: PLINK bed writer/reader in numpy with the reference's bit order.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
i64 = C.c_int64


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "orc_decode_bed": (C.c_int, [u8p, i64, i64, i64, u8p, u8p, i64p, i64, u8p]),
        "orc_counts": (None, [u8p, i64, i64, i32p, i32p, i32p, i32p, f32p]),
        "orc_maf": (None, [u8p, i64, i64, f32p]),
        "orc_gw": (None, [f32p, f32p, i64, f32p]),
        "orc_op": (None, [C.c_int, u8p, i64, i64, f32p, f32p, f32p, C.c_float, f32p, C.c_int, f32p, C.c_int,
                          C.c_int, f32p]),
        "orc_rmatvec": (None, [u8p, i64, i64, f32p, f32p, f32p, C.c_int, C.c_int, f32p]),
        "orc_matvec": (None, [u8p, i64, i64, f32p, f32p, C.c_int, C.c_int, f32p]),
        "orc_lseed": (C.c_void_p, [C.c_int, u8p, i64, i64, f32p, f32p, f32p, C.c_float, f32p, C.c_int, f32p,
                                   C.c_int, C.c_float, C.c_int, C.c_int]),
        "orc_lseed_steps": (C.c_int, [C.c_void_p]),
        "orc_lseed_converged": (C.c_int, [C.c_void_p]),
        "orc_lseed_coeffs": (None, [C.c_void_p, f32p, f32p]),
        "orc_lseed_free": (None, [C.c_void_p]),
        "orc_lsolve": (None, [C.c_void_p, f32p, C.c_float, C.c_float, f32p]),
        "orc_rademacher": (None, [C.c_int, i64, C.c_int, f32p]),
        "orc_lmc_samples": (None, [C.c_int, i64, i64, C.c_int, f32p, f32p]),
        "orc_gebv": (C.c_int, [u8p, i64, i64, i64, u8p, u8p, f64p, f64p]),
        "orc_fake_geno": (None, [C.c_int, i64, i64, C.c_int, C.c_int, f32p, f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _cm(a, rows):
    """(rows, t) array -> Fortran-ordered float32 (column-major, ld = rows)."""
    a = np.asarray(a, dtype=np.float32)
    if a.ndim == 1:
        a = a[:, None]
    assert a.shape[0] == rows
    return np.asfortranarray(a)


# ----------------------------------------------------------------------------- input ------
def decode_bed(bed: bytes | np.ndarray, n_raw: int, m_raw: int, bindi=None, bmarker=None, order=None) -> np.ndarray:
    """kmat bits (m_keep, 2*n_keep) uint8 -- src/input.cpp:343-386, :446-454."""
    lib = load()
    buf = np.frombuffer(bytes(bed), dtype=np.uint8) if not isinstance(bed, np.ndarray) else bed.astype(np.uint8)
    bi = np.ones(n_raw, np.uint8) if bindi is None else np.ascontiguousarray(bindi, np.uint8)
    bm = np.ones(m_raw, np.uint8) if bmarker is None else np.ascontiguousarray(bmarker, np.uint8)
    m_ret = int(bm.sum())
    od = np.arange(m_ret, dtype=np.int64) if order is None else np.ascontiguousarray(order, np.int64)
    n_keep = int(bi.sum())
    out = np.ones((len(od), 2 * n_keep), dtype=np.uint8)
    rc = lib.orc_decode_bed(_p(buf, C.c_uint8), len(buf), n_raw, m_raw, _p(bi, C.c_uint8), _p(bm, C.c_uint8),
                            _p(od, C.c_int64), len(od), _p(out, C.c_uint8))
    if rc != 0:
        raise ValueError("Error: problem with the BED file. Has the FAM/BIM file been changed?")
    return out


def counts(kmat: np.ndarray):
    lib = load()
    m, w = kmat.shape
    n = w // 2
    km = np.ascontiguousarray(kmat, np.uint8)
    pc, n0, n1, n2 = (np.zeros(m, np.int32) for _ in range(4))
    gm = np.zeros(m, np.float32)
    lib.orc_counts(_p(km, C.c_uint8), m, n, _p(pc, C.c_int32), _p(n0, C.c_int32), _p(n1, C.c_int32),
                   _p(n2, C.c_int32), _p(gm, C.c_float))
    return pc, n0, n1, n2, gm


def maf(kmat: np.ndarray) -> np.ndarray:
    lib = load()
    m, w = kmat.shape
    km = np.ascontiguousarray(kmat, np.uint8)
    out = np.zeros(m, np.float32)
    lib.orc_maf(_p(km, C.c_uint8), m, w // 2, _p(out, C.c_float))
    return out


def gw(gm: np.ndarray, persnp=None) -> np.ndarray:
    lib = load()
    g = _f32(gm)
    ps = _f32(persnp)
    out = np.zeros(len(g), np.float32)
    lib.orc_gw(_p(g, C.c_float), _p(ps, C.c_float), len(g), _p(out, C.c_float))
    return out


# ------------------------------------------------------------------------- operators ------
class Model:
    """The operator's inputs exactly as src/reml.cpp passes them (kmat, gm, gw, rr, Q)."""

    def __init__(self, kmat, gm, gw_, rr, Q=None, subset_size=1000):
        self.kmat = np.ascontiguousarray(kmat, np.uint8)
        self.m, w = self.kmat.shape
        self.n = w // 2
        self.gm = _f32(gm)
        self.gw = _f32(gw_)
        self.rr = _f32(rr)
        if Q is None:
            self.Q, self.c = np.zeros((self.n, 0), np.float32, order="F"), 0
        else:
            Qa = np.asarray(Q, np.float32)
            if Qa.ndim == 1:
                Qa = Qa[:, None]
            self.Q, self.c = np.asfortranarray(Qa), Qa.shape[1]
        self.subset_size = subset_size

    def op(self, kind: str, tau0: float, V: np.ndarray) -> np.ndarray:
        lib = load()
        Vc = _cm(V, self.n)
        t = Vc.shape[1]
        out = np.zeros((self.n, t), np.float32, order="F")
        lib.orc_op(1 if kind == "quadform" else 0, _p(self.kmat, C.c_uint8), self.n, self.m, _p(self.gm, C.c_float),
                   _p(self.gw, C.c_float), _p(self.rr, C.c_float), tau0, _p(self.Q, C.c_float), self.c,
                   _p(Vc, C.c_float), t, self.subset_size, _p(out, C.c_float))
        return out

    def rmatvec(self, V, weighted=False) -> np.ndarray:
        lib = load()
        Vc = _cm(V, self.n)
        t = Vc.shape[1]
        out = np.zeros((self.m, t), np.float32, order="F")
        lib.orc_rmatvec(_p(self.kmat, C.c_uint8), self.n, self.m, _p(self.gm, C.c_float),
                        _p(self.gw, C.c_float) if weighted else None, _p(Vc, C.c_float), t, self.subset_size,
                        _p(out, C.c_float))
        return out

    def matvec(self, U) -> np.ndarray:
        lib = load()
        Uc = _cm(U, self.m)
        t = Uc.shape[1]
        out = np.zeros((self.n, t), np.float32, order="F")
        lib.orc_matvec(_p(self.kmat, C.c_uint8), self.n, self.m, _p(self.gm, C.c_float), _p(Uc, C.c_float), t,
                       self.subset_size, _p(out, C.c_float))
        return out

    def L_Seed(self, kind: str, tau0: float, B, tol=5e-4, maxit=600) -> "Seed":
        lib = load()
        Bc = _cm(B, self.n)
        t = Bc.shape[1]
        h = lib.orc_lseed(1 if kind == "quadform" else 0, _p(self.kmat, C.c_uint8), self.n, self.m,
                          _p(self.gm, C.c_float), _p(self.gw, C.c_float), _p(self.rr, C.c_float), tau0,
                          _p(self.Q, C.c_float), self.c, _p(Bc, C.c_float), t, tol, maxit, self.subset_size)
        return Seed(h, self.n, t)


class Seed:
    """struct lseed of src/reml.cpp:240-244 (FP32, CPU)."""

    def __init__(self, h, n, t):
        self._h, self.n, self.t = h, n, t
        lib = load()
        self.steps = lib.orc_lseed_steps(h)
        self.converged = bool(lib.orc_lseed_converged(h))
        beta = np.zeros((self.steps + 1, t), np.float32)
        delta = np.zeros((max(self.steps, 1), t), np.float32)
        lib.orc_lseed_coeffs(h, _p(beta, C.c_float), _p(delta, C.c_float))
        self.beta = beta.T.copy()              # t x (steps+1)
        self.delta = delta[: self.steps].T.copy()  # t x steps

    def L_Solve(self, B, sigma, tol=1e-6) -> np.ndarray:
        Bc = _cm(B, self.n)
        assert Bc.shape[1] == self.t
        X = np.zeros((self.n, self.t), np.float32, order="F")
        load().orc_lsolve(self._h, _p(Bc, C.c_float), sigma, tol, _p(X, C.c_float))
        return X

    def __del__(self):
        try:
            if self._h:
                load().orc_lseed_free(self._h)
                self._h = None
        except Exception:
            pass


# ------------------------------------------------------------------------------ RNG -------
def rademacher(seed: int, n: int, nv: int) -> np.ndarray:
    V = np.zeros((n, nv), np.float32, order="F")
    load().orc_rademacher(seed, n, nv, _p(V, C.c_float))
    return V


def lmc_samples(seed: int, n: int, m: int, nmc: int):
    u = np.zeros((m, nmc), np.float32, order="F")
    e = np.zeros((n, nmc), np.float32, order="F")
    load().orc_lmc_samples(seed, n, m, nmc, _p(u, C.c_float), _p(e, C.c_float))
    return u, e


def fake_geno(seed: int, n: int, m: int, nmc: int, k: int):
    """Uncentred fake-marker dosages (n x k) and their frequencies (k) of --fake_geno."""
    V = np.zeros((n, k), np.float32, order="F")
    f = np.zeros(k, np.float32)
    load().orc_fake_geno(seed, n, m, nmc, k, _p(V, C.c_float), _p(f, C.c_float))
    return V, f


# ------------------------------------------------------------------------------ GEBV ------
def gebv(bed: bytes, indi_num: int, snp_num: int, bmarker, bflip, alpha) -> np.ndarray:
    lib = load()
    buf = np.frombuffer(bytes(bed), dtype=np.uint8)
    bm = np.ascontiguousarray(bmarker, np.uint8)
    bf = np.ascontiguousarray(bflip, np.uint8)
    al = np.ascontiguousarray(alpha, np.float64)
    out = np.zeros(indi_num, np.float64)
    rc = lib.orc_gebv(_p(buf, C.c_uint8), len(buf), indi_num, snp_num, _p(bm, C.c_uint8), _p(bf, C.c_uint8),
                      _p(al, C.c_double), _p(out, C.c_double))
    if rc == -2:
        raise ValueError("No SNPs found both in estimate file and in genotype file.")
    if rc != 0:
        raise ValueError("Error: problem with the BED file ... has the FAM/BIM file been changed?")
    return out
