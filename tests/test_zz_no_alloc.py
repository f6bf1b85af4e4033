"""slemm_gpu_op_noalloc (the operator without device allocation, workspaces in one static
scratch array per GPU) against slemm_gpu_op: the same kernels, so the same results bit for bit,
on two calls in a row with the second on a smaller input, for the int8 wide (65 columns) and
narrow (1 and 3 columns), the BF16 hi/lo-in-columns (7) and the plain FP32 paths, with mixed
operator kinds and covariates; the handle's own buffers are untouched; an oversized call is
refused before any work."""
import ctypes as C

import numpy as np
import pytest
import torch

from tests.problems import make_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(gpu):
    p = make_problem(n=3000, m=4000, seed=41, weights=True, c=2)
    g = gpu.Geno.from_bed(p.rows, p.n)
    g.set_model(None, p.rvec, p.Q)
    return p, g


def _block(n, t, seed):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.normal(size=(t, n)).astype(np.float32)).cuda()


@pytest.mark.parametrize("prec", ["bf16x2", "fp32"])
@pytest.mark.parametrize("t1,t2", [(65, 3), (7, 1), (16, 5)])
def test_noalloc_equals_op_two_calls(prob, prec, t1, t2):
    p, g = prob
    tau0 = (1 - 0.7) / 0.7
    for call, t in enumerate((t1, t2)):
        v = _block(p.n, t, 100 * t + call)
        kinds = ["quadform" if k % 3 == 0 else "ldet" for k in range(t)]
        want = g.lnr_op_select("ldet", tau0, v.t(), precision=prec, col_kind=kinds).t().contiguous()
        out = torch.full_like(v, float("nan"))
        g.lnr_op_select_noalloc("ldet", tau0, v, out, precision=prec, col_kind=kinds)
        torch.cuda.synchronize()
        assert torch.equal(out, want), (prec, t, float((out - want).abs().max()))
        if call == 0:  # and against FP64 once per shape
            D = p.dense()
            ref = np.concatenate([D.op(kinds[k], tau0, v[k].cpu().numpy()[:, None]) for k in range(t)], axis=1)
            assert rel_err(out.cpu().numpy().T, ref) < (3e-4 if prec == "bf16x2" else 1e-4)


def test_noalloc_leaves_handle_buffers_and_refuses_oversized_calls(prob, gpu):
    from slemm_amd import _lib
    p, g = prob
    tau0 = 0.5
    v = _block(p.n, 33, 7)
    before = g.lnr_op_select("ldet", tau0, v.t()).t().contiguous()
    out = torch.empty_like(v)
    g.lnr_op_select_noalloc("ldet", tau0, v, out)
    after = g.lnr_op_select("ldet", tau0, v.t()).t().contiguous()
    assert torch.equal(before, after) and torch.equal(out, before)
    # 200 columns over 300,000 SNPs: the pass-2 operand alone is 250 MB
    big = gpu.Geno.synthesize(20000, 300000, seed=3)
    big.set_model()
    vb = _block(big.n, 200, 9)
    ob = torch.zeros_like(vb)
    with pytest.raises(_lib.SlemmError, match="no-allocation|scratch"):
        big.lnr_op_select_noalloc("ldet", tau0, vb, ob)
    torch.cuda.synchronize()
    assert torch.count_nonzero(ob) == 0
    big.close()
