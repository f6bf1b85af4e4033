"""Seeded small problems shared by the CPU and GPU tests."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import oracle
from oracle import bedio
from oracle.dense import Dense, dosage_of_kmat


@dataclass
class Problem:
    n: int
    m: int
    codes: np.ndarray   # (m, n) 2-bit codes
    rows: np.ndarray    # packed bed rows (m, ceil(n/4))
    kmat: np.ndarray    # reference kmat bits (m, 2n)
    Z: np.ndarray       # dosages (n, m) float64
    gm: np.ndarray
    gw: np.ndarray
    rr: np.ndarray
    rvec: np.ndarray | None
    Q: np.ndarray       # (n, c) float32 orthonormal basis of R^-1/2 X
    X: np.ndarray       # (n, c) covariates
    persnp: np.ndarray | None

    def dense(self) -> Dense:
        return Dense(self.Z, self.gm, self.gw, self.rr, self.Q)

    def model(self, subset_size=1000) -> "oracle.Model":
        return oracle.Model(self.kmat, self.gm, self.gw, self.rr, self.Q, subset_size)


def make_problem(n=700, m=900, seed=1, missing=0.01, c=1, weights=False, persnp=False) -> Problem:
    rng = np.random.default_rng(seed + 1000)
    codes, _ = bedio.synth_codes(n, m, seed, missing=missing)
    # guarantee polymorphic SNPs (monomorphic ones have infinite weight in the reference too)
    codes[:, 0] = 3
    codes[:, 1] = 0
    rows = bedio.encode_codes(codes)
    kmat = bedio.kmat_of_codes(codes)
    pc, n0, n1, n2, gm = oracle.counts(kmat)
    ps = rng.uniform(0.5, 2.0, m).astype(np.float32) if persnp else None
    gw = oracle.gw(gm, ps)
    rvec = rng.uniform(0.5, 2.0, n).astype(np.float32) if weights else None
    rr = (1.0 / np.sqrt(rvec)).astype(np.float32) if weights else np.ones(n, np.float32)
    X = np.ones((n, 1), np.float64)
    if c > 1:
        X = np.concatenate([X, rng.normal(size=(n, c - 1))], axis=1)
    Qd, _ = np.linalg.qr(rr[:, None].astype(np.float64) * X)
    Q = Qd.astype(np.float32)
    return Problem(n, m, codes, rows, kmat, dosage_of_kmat(kmat), gm, gw, rr, rvec, Q, X, ps)


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def make_phenotype(p: Problem, h2=0.5, n_causal=200, seed=0) -> np.ndarray:
    """y = Z beta + e with beta ~ N(0, 1) on n_causal SNPs and var(e) set for the target h2."""
    rng = np.random.default_rng(seed + 77)
    idx = rng.choice(p.m, size=min(n_causal, p.m), replace=False)
    beta = rng.normal(size=len(idx))
    g = p.Z[:, idx] @ beta
    g -= g.mean()
    ve = g.var() * (1 - h2) / h2
    e = rng.normal(scale=np.sqrt(ve), size=p.n)
    if p.rvec is not None:
        e *= np.sqrt(p.rvec)
    return (g + e + 3.0).astype(np.float32)
