"""Differentiable REML log-likelihood over the device engine (thin torch.autograd harness).

Model (that of SLDF_REML, src/reml.cpp:549-690, with free per-SNP variances d):
    V = Zc diag(d) Zc' + s2e R,   l(d, s2e) = -1/2 [log|V| + log|X'V^-1X| + y'Py] (+ const)
With A_d = R^-1/2 Zc diag(d) Zc' R^-1/2, Pi = I - QQ' (QR0 = R^-1/2 X) and
Hq = Pi A_d Pi + s2e I (the QUADFORM operator of src/reml.cpp:8-50 with the weights set to d):
    l = -1/2 [log|Hq| - c log s2e + log|X~'X~| + sum log R + y~' Pi Hq^-1 Pi y~],   y~ = R^-1/2 y
    dl/dd_j = -1/2 [z~_j' P z~_j - (z~_j' P y~)^2],   z~_j = R^-1/2 zc_j
    dl/ds2e = -1/2 [tr(P) - ||P y~||^2],               P = Pi Hq^-1 Pi
The forward pass runs one Lanczos seed on [y~, probes] (L_Seed), the CG solves (L_Solve) and
the SLQ log-determinant (SLQ_LDet).  The backward pass is the adjoint kernel: one pass over the
bed for (Zc'a)o(Zc'b) with a = [P v, P y~], b = [v, P y~], where traces are n/nV sum_v v'Mv over
the probe columns (E[v v'] = I/n; identity columns give exact traces).

Per-window weights w enter through the normalisation of src/reml.cpp:622,
d_j = sigma_g^2 base_gw_j w[win_j] / mean(w[win]), which torch differentiates.
No reference counterpart: the reference computes no derivatives.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import gpu


class RemlProblem:
    """Phenotype, covariates and probes bound to a Geno whose rvec is already set."""

    def __init__(self, geno: "gpu.Geno", X: np.ndarray, y: np.ndarray, rvec: Optional[np.ndarray] = None,
                 probes: Optional[torch.Tensor] = None, n_V: int = 64, seed: int = 0, tol: float = 1e-6,
                 precision: str = "bf16x2"):
        self.g = geno
        n = geno.n
        self.n = n
        self.tol = tol
        self.precision = precision
        rv = np.ones(n) if rvec is None else np.asarray(rvec, np.float64)
        self.rvec = rv
        rr = 1.0 / np.sqrt(rv)
        Xw = np.asarray(X, np.float64).reshape(n, -1) * rr[:, None]
        Q, R0 = np.linalg.qr(Xw)
        self.c = Q.shape[1]
        geno.set_model(None, rv.astype(np.float32), Q.astype(np.float32))
        _, self.base_gw, _ = geno.model()
        self.const = float(2.0 * np.sum(np.log(np.abs(np.diag(R0)))) + np.sum(np.log(rv)))
        yw = rr * np.asarray(y, np.float64)
        yw -= Q @ (Q.T @ yw)
        dev = torch.device("cuda")
        self.Q = torch.from_numpy(Q).to(dev)
        self.y = torch.from_numpy(yw).to(dev)
        if probes is None:
            gen = torch.Generator(device=dev).manual_seed(seed)
            v = torch.where(torch.rand((n, n_V), generator=gen, device=dev) < 0.5, -1.0, 1.0)
            probes = v / np.sqrt(n)
        self.V = probes.to(dev, torch.float32).contiguous()
        self.nV = self.V.shape[1]

    def project(self, X: torch.Tensor) -> torch.Tensor:
        X = X.double()
        return X - self.Q @ (self.Q.T @ X)


class RemlLogLik(torch.autograd.Function):
    """l(d, s2e); d (m,) and s2e () float64 tensors.  Forward on the engine, backward by the
    adjoint kernel."""

    @staticmethod
    def forward(ctx, d: torch.Tensor, s2e: torch.Tensor, prob: RemlProblem):
        g = prob.g
        g.set_gw(d)
        s2 = float(s2e)
        B = torch.cat([prob.y.float()[:, None], prob.V], dim=1)
        seed = g.L_Seed("quadform", s2, B, tol=prob.tol, keep_cols=B.shape[1], precision=prob.precision)
        try:
            S = seed.L_Solve(B, 0.0, tol=prob.tol)
            ldet = seed.SLQ_LDet(0.0, col0=1, nv=prob.nV)
        finally:
            seed.close()
        PS = prob.project(S)
        Py = PS[:, 0]
        PV = PS[:, 1:]
        yPy = float(prob.y @ Py)
        ll = -0.5 * (ldet - prob.c * np.log(s2) + prob.const + yPy)
        ctx.prob = prob
        ctx.save_for_backward(PV.float(), Py.float(), torch.tensor(s2))
        return torch.tensor(ll, dtype=torch.float64, device=d.device)

    @staticmethod
    def backward(ctx, grad_out):
        PV, Py, _ = ctx.saved_tensors
        prob = ctx.prob
        w = prob.n / prob.nV
        A = torch.cat([PV, Py[:, None]], dim=1)
        B = torch.cat([prob.V, Py[:, None]], dim=1)
        coef = np.full(prob.nV + 1, -0.5 * w)
        coef[-1] = 0.5
        gd = prob.g.adjoint(A, B, coef, scale_rr=True, precision=prob.precision)
        trP = w * float(torch.sum(PV.double() * prob.V.double()))
        gs = -0.5 * (trP - float(Py.double() @ Py.double()))
        go = grad_out.to(torch.float64)
        return go * gd.to(go.device), go * torch.tensor(gs, dtype=torch.float64, device=go.device), None


def window_gw(base_gw: torch.Tensor, win: torch.Tensor, w: torch.Tensor, sigma_g2: torch.Tensor) -> torch.Tensor:
    """d_j = sigma_g^2 base_gw_j w[win_j] / mean(w[win]) (persnp normalisation of src/reml.cpp:622)."""
    persnp = w[win]
    return sigma_g2 * base_gw * persnp / persnp.mean()


def loglik(prob: RemlProblem, w: torch.Tensor, sigma_g2: torch.Tensor, s2e: torch.Tensor,
           win: torch.Tensor) -> torch.Tensor:
    """l as a differentiable function of the window weights and both variance components."""
    base = torch.as_tensor(prob.base_gw, dtype=torch.float64, device=w.device)
    d = window_gw(base, win, w, sigma_g2)
    return RemlLogLik.apply(d, s2e, prob)
