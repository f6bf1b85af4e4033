"""ORACLE -- test infrastructure only.  PLINK .bed encoding in numpy.

Bit convention of the reference (src/input.cpp:354-357, :369-377): SNP-major rows of
ceil(n/4) bytes after the magic bytes 6C 1B 01; individual i uses bits 2(i%4) (low) and
2(i%4)+1 (high) of byte i//4.  Codes: 00 hom A1 (dosage 0), 01 missing (the reference sums
the bits, so dosage 1), 10 het (1), 11 hom A2 (2).
"""
from __future__ import annotations

import numpy as np

MAGIC = bytes([0x6C, 0x1B, 0x01])

# A2 count -> 2-bit code (low bit, high bit) = 00, 10 (binary value 2), 11 (value 3)
_CODE_OF_DOSAGE = np.array([0, 2, 3], dtype=np.uint8)
CODE_MISSING = 1


def encode_codes(codes: np.ndarray) -> np.ndarray:
    """codes (m, n) in {0,1,2,3} -> packed rows (m, ceil(n/4)) uint8; padding bits zero."""
    m, n = codes.shape
    nb = (n + 3) // 4
    pad = np.zeros((m, nb * 4), dtype=np.uint8)
    pad[:, :n] = codes
    q = pad.reshape(m, nb, 4)
    return (q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)).astype(np.uint8)


def decode_codes(rows: np.ndarray, n: int) -> np.ndarray:
    """packed rows (m, ceil(n/4)) -> codes (m, n)."""
    r = np.asarray(rows, dtype=np.uint8)
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    codes = (r[:, :, None] >> sh[None, None, :]) & 3
    return codes.reshape(r.shape[0], -1)[:, :n]


def dosage_of_codes(codes: np.ndarray) -> np.ndarray:
    return ((codes & 1) + (codes >> 1)).astype(np.int8)


def codes_of_dosage(d: np.ndarray, missing_mask=None) -> np.ndarray:
    c = _CODE_OF_DOSAGE[np.asarray(d, dtype=np.int64)]
    if missing_mask is not None:
        c = np.where(missing_mask, CODE_MISSING, c).astype(np.uint8)
    return c


def bed_bytes(rows: np.ndarray) -> bytes:
    return MAGIC + np.ascontiguousarray(rows, dtype=np.uint8).tobytes()


def kmat_of_codes(codes: np.ndarray) -> np.ndarray:
    """codes (m, n) -> the reference's kmat bits (m, 2n): [low, high] per individual."""
    m, n = codes.shape
    out = np.empty((m, 2 * n), dtype=np.uint8)
    out[:, 0::2] = codes & 1
    out[:, 1::2] = codes >> 1
    return out


def synth_codes(n: int, m: int, seed: int, maf_lo=0.05, maf_hi=0.5, missing=0.0):
    """Synthetic genotypes: p_j ~ U(lo, hi), A2 count ~ Bin(2, p_j) under HWE, `missing`
    fraction coded 01.  Returns (codes (m, n) uint8, p (m,))."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(maf_lo, maf_hi, size=m)
    d = rng.binomial(2, p[:, None], size=(m, n))
    miss = rng.random((m, n)) < missing if missing > 0 else None
    return codes_of_dosage(d, miss), p
