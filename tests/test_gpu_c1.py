"""Variance components at the reference's CPU-runnable workload: n=5,000 individuals, m=50,000
SNPs, MAF ~ U(0.05, 0.5) under HWE, 1% missing (code 01, dosage 1), an intercept, h2 = 0.5 on
1,000 causal SNPs, 30 Rademacher probes.  SLDF_REML (src/reml.cpp:549-690) and LMC_REML
(src/reml.cpp:730-882) on the GPU, both precisions, against their FP64 dense restatements on
the same probes and MC draws (oracle.dense: sldf_gss, lmc_reml), with every tolerance that
would round the answer run to 1e-6.  Bar: 1e-4 relative in Vg, Ve and h2."""
import numpy as np
import pytest

from oracle import reml as oreml
from tests.problems import make_phenotype, make_problem, rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module")
def c1(gpu):
    import oracle
    from oracle.dense import sldf_gss
    p = make_problem(n=5000, m=50000, seed=31, missing=0.01)
    y = make_phenotype(p, 0.5, n_causal=1000, seed=3)
    D = p.dense()
    Q, _ = oreml.covariate_qr(p.X, p.rr)
    yp = (p.rr * y).astype(np.float32)
    yp = (yp - Q @ (Q.T @ yp)).astype(np.float32)
    vg, ve, _, h2 = sldf_gss(D, yp, Q, oracle.rademacher(0, p.n, 30), tol=1e-6)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model()
    print(f"C1 FP64 SLDF optimum h2 {h2:.7f} vg {vg:.6f} ve {ve:.6f}", flush=True)
    return p, y, D, g, dict(vg=vg, ve=ve, h2=h2)


def _rel(a, b):
    return {k: abs(a[k] - b[k]) / abs(b[k]) for k in ("vg", "ve", "h2")}


@pytest.mark.parametrize("prec", ["bf16x2", "fp32"])
def test_c1_sldf_vc_against_fp64(c1, prec):
    p, y, D, g, ref = c1
    got = g.SLDF_REML(p.X, y, n_V=30, tol_L=1e-6, tol_VC=1e-6, precision=prec)
    d = _rel(got, ref)
    msg = f"C1 SLDF {prec}: h2 {got['h2']:.7f} rel {d} steps {got['steps']} {got['seconds_total']:.2f}s"
    print(msg)
    assert got["evals"] == 30, msg
    assert max(d.values()) < 1e-4, msg


@pytest.mark.parametrize("prec", ["bf16x2", "fp32"])
def test_c1_lmc_vc_against_fp64(c1, prec):
    import oracle
    from oracle.dense import lmc_reml
    p, y, D, g, _ = c1
    n_MC = max(min(int(4e9 / p.n / p.n), 30), 3)  # src/option.cpp:414: 30 at n = 5,000
    u, e = oracle.lmc_samples(0, p.n, p.m, n_MC)
    ref = lmc_reml(D, y, u, e, tol=1e-6, max_iter=40)
    got = g.LMC_REML(p.X, y, n_MC=n_MC, tol_L=1e-6, secant_tol=1e-6, secant_max_iter=40, precision=prec)
    d = _rel(got, ref)
    msg = f"C1 LMC {prec}: h2 {got['h2']:.7f} fp64 {ref['h2']:.7f} rel {d} evals {got['evals']}/{ref['evals']}"
    print(msg)
    assert max(d.values()) < 1e-4, msg
    assert rel_err(got["Py"], ref["Py"]) < 1e-4, msg
