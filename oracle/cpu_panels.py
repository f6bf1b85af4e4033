"""ORACLE -- test infrastructure and bench.py's CPU baseline only.  The reference's operator
pipeline on host cores, in the structure src/reml.cpp runs it:

* per operator call, each subset of `subset_size` SNPs is expanded from the packed bed into an
  n x subset float panel of centred dosages z - gm (src/reml.cpp:28-47: the per-call expansion of
  the vector<bool> genotype bits into tmp_mat1 / tmp_mat2);
* the operator then takes two single-precision GEMMs per subset, T = panel' V and
  out += panel (W T) (src/reml.cpp:36, 79), on MKL through torch's CPU backend (the reference
  reaches MKL's sgemm through Eigen with EIGEN_USE_MKL_ALL);
* L_Seed / L_Solve (src/reml.cpp:246-339, 369-431) in FP32 numpy over that operator, with the
  recurrences of oracle.Model's C++ restatement, and SLDF_REML / LMC_REML from oracle.reml on top.

`PanelModel` has the interface of oracle.Model (op, rmatvec, matvec, L_Seed), so the restated
drivers run on it unchanged.
"""
from __future__ import annotations

import numpy as np
import torch

# dosage (low bit + high bit) of each 2-bit code, for the 4 individuals of a packed byte
_LUT = np.array([[((b >> (2 * k)) & 1) + ((b >> (2 * k + 1)) & 1) for k in range(4)] for b in range(256)],
                np.float32)


class PanelModel:
    """rows: packed SNP-major bed rows (m, ceil(n/4)) uint8, the reference's bit order."""

    def __init__(self, rows: np.ndarray, n: int, gm, gw, rr, Q=None, subset_size: int = 1000):
        self.rows = np.ascontiguousarray(rows, np.uint8)
        self.m = self.rows.shape[0]
        self.n = int(n)
        self.gm = np.asarray(gm, np.float32)
        self.gw = np.asarray(gw, np.float32)
        self.rr = np.asarray(rr, np.float32)
        if Q is None:
            self.Q = np.zeros((self.n, 0), np.float32)
        else:
            Qa = np.asarray(Q, np.float32)
            self.Q = Qa[:, None] if Qa.ndim == 1 else Qa
        self.c = self.Q.shape[1]
        self.subset_size = int(subset_size)

    # ----------------------------------------------------------------------- panels -------
    def _panel(self, j0: int, j1: int) -> torch.Tensor:
        """(j1 - j0, n) float32 panel of centred dosages of SNPs j0..j1-1 (rows: SNPs)."""
        d = _LUT[self.rows[j0:j1]].reshape(j1 - j0, -1)[:, :self.n]
        d -= self.gm[j0:j1, None]
        return torch.from_numpy(d)

    def _zt(self, V: np.ndarray, weighted: bool) -> np.ndarray:
        Vt = torch.from_numpy(np.ascontiguousarray(V, np.float32))
        out = np.empty((self.m, Vt.shape[1]), np.float32)
        for j0 in range(0, self.m, self.subset_size):
            j1 = min(self.m, j0 + self.subset_size)
            T = self._panel(j0, j1) @ Vt
            if weighted:
                T *= torch.from_numpy(self.gw[j0:j1, None])
            out[j0:j1] = T.numpy()
        return out

    def rmatvec(self, V, weighted=False) -> np.ndarray:
        V = np.asarray(V, np.float32)
        return self._zt(V[:, None] if V.ndim == 1 else V, weighted)

    def matvec(self, U) -> np.ndarray:
        U = np.asarray(U, np.float32)
        U = U[:, None] if U.ndim == 1 else U
        out = torch.zeros((self.n, U.shape[1]), dtype=torch.float32)
        for j0 in range(0, self.m, self.subset_size):
            j1 = min(self.m, j0 + self.subset_size)
            out += self._panel(j0, j1).t() @ torch.from_numpy(np.ascontiguousarray(U[j0:j1]))
        return out.numpy()

    def op(self, kind: str, tau0: float, V: np.ndarray) -> np.ndarray:
        """H0_ldet_mv (src/reml.cpp:52-94) or H0_quadform_mv (:8-50) over float panels."""
        V = np.asarray(V, np.float32)
        V = V[:, None] if V.ndim == 1 else V
        X = V
        if kind == "quadform" and self.c:
            X = (X - self.Q @ (self.Q.T @ X)).astype(np.float32)
        X = (self.rr[:, None] * X).astype(np.float32)
        Xt = torch.from_numpy(np.ascontiguousarray(X))
        out = torch.zeros_like(Xt)
        for j0 in range(0, self.m, self.subset_size):
            j1 = min(self.m, j0 + self.subset_size)
            P = self._panel(j0, j1)
            T = (P @ Xt) * torch.from_numpy(self.gw[j0:j1, None])
            out += P.t() @ T
        out = (self.rr[:, None] * out.numpy()).astype(np.float32)
        if kind == "quadform" and self.c:  # loo = ZPv + tau0 v - Q(Q'ZPv) (src/reml.cpp:49)
            out = (out - self.Q @ (self.Q.T @ out)).astype(np.float32)
        return (out + np.float32(tau0) * V).astype(np.float32)

    # ----------------------------------------------------------------------- Krylov -------
    def L_Seed(self, kind: str, tau0: float, B, tol=5e-4, maxit=600) -> "PanelSeed":
        """L_Seed (src/reml.cpp:246-339) in FP32, all columns in lock step."""
        f32 = np.float32
        B = np.asarray(B, np.float32)
        B = B[:, None] if B.ndim == 1 else B
        t = B.shape[1]
        Bn = np.sqrt(np.sum(B * B, axis=0, dtype=np.float32)).astype(np.float32)
        beta = [Bn.copy()]
        delta, U = [], [(B / Bn[None, :]).astype(np.float32)]
        rho = Bn.copy()
        gamma_prev = omega_prev = None
        j, cnvg = 0, False
        while j < maxit - 1:
            tmp = self.op(kind, tau0, U[j])
            if j > 0:
                tmp = (tmp - U[j - 1] * beta[j][None, :]).astype(np.float32)
            d = np.sum(U[j] * tmp, axis=0, dtype=np.float32).astype(np.float32)
            tmp = (tmp - U[j] * d[None, :]).astype(np.float32)
            b1 = np.sqrt(np.sum(tmp * tmp, axis=0, dtype=np.float32)).astype(np.float32)
            U.append((tmp / b1[None, :]).astype(np.float32))
            delta.append(d)
            beta.append(b1)
            gamma = (f32(1) / d) if j == 0 else (f32(1) / (d - omega_prev / gamma_prev))
            gamma = gamma.astype(np.float32)
            omega = ((b1 * gamma) * (b1 * gamma)).astype(np.float32)
            rho = (-b1 * gamma * rho).astype(np.float32)
            gamma_prev, omega_prev = gamma, omega
            j += 1
            res = np.max(np.sqrt(np.sum((U[j] * rho[None, :]) ** 2, axis=0, dtype=np.float32)) / Bn)
            if res <= tol:
                cnvg = True
                break
        return PanelSeed(U, np.array(beta, np.float32).T.copy(), np.array(delta, np.float32).T.copy(), j, cnvg)


class PanelSeed:
    """struct lseed (src/reml.cpp:240-244): the basis U, beta (t x steps+1), delta (t x steps)."""

    def __init__(self, U, beta, delta, steps, converged):
        self.U, self.beta, self.delta, self.steps, self.converged = U, beta, delta, steps, converged
        self.t = beta.shape[0]

    def L_Solve(self, B, sigma, tol=1e-6) -> np.ndarray:
        """L_Solve (src/reml.cpp:369-431) in FP32: the shifted CG replayed on the stored basis."""
        f32 = np.float32
        B = np.asarray(B, np.float32)
        B = B[:, None] if B.ndim == 1 else B
        Bn = np.sqrt(np.sum(B * B, axis=0, dtype=np.float32)).astype(np.float32)
        P = B.copy()
        X = np.zeros_like(B)
        rho = Bn.copy()
        gamma_prev = omega_prev = None
        for j in range(len(self.U) - 1):
            dj = (self.delta[:, j] + f32(sigma)).astype(np.float32)
            gamma = (f32(1) / dj) if j == 0 else (f32(1) / (dj - omega_prev / gamma_prev))
            gamma = gamma.astype(np.float32)
            b1 = self.beta[:, j + 1]
            omega = ((b1 * gamma) * (b1 * gamma)).astype(np.float32)
            rho = (-b1 * gamma * rho).astype(np.float32)
            X = (X + P * gamma[None, :]).astype(np.float32)
            R = (self.U[j + 1] * rho[None, :]).astype(np.float32)
            P = (R + P * omega[None, :]).astype(np.float32)
            gamma_prev, omega_prev = gamma, omega
            res = np.max(np.sqrt(np.sum(R * R, axis=0, dtype=np.float32)) / Bn)
            if res <= tol:
                break
        return X
