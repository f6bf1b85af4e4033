"""File-level entry points: the streaming .bed loader (src/input.cpp:343-386) and --pred
(src/pred.cpp:4-151), against the in-memory upload and an exact numpy GEBV."""
import numpy as np
import pytest

from oracle import bedio
from tests.problems import make_problem

pytestmark = pytest.mark.gpu

MAGIC = bytes([0x6C, 0x1B, 0x01])


def test_load_bed_equals_in_memory_upload(gpu, tmp_path):
    p = make_problem(n=1037, m=611, seed=3, missing=0.02)
    path = tmp_path / "g.bed"
    path.write_bytes(MAGIC + p.rows.tobytes())
    keep = np.sort(np.random.default_rng(1).choice(p.n, size=800, replace=False))
    order = np.r_[10:50, 5, 300:320, 610, 0:3, 100:101]  # runs, singles, the last SNP
    for k, o in ((None, None), (keep, None), (keep, order)):
        a = gpu.Geno.load_bed(path, p.n, p.m, k, o)
        b = gpu.Geno.from_bed(p.rows, p.n, k, o)
        assert np.array_equal(a.download(), b.download())
        assert all(np.array_equal(x, y) for x, y in zip(a.counts(), b.counts()))


def test_load_bed_many_chunks(gpu, tmp_path):
    """More rows than one 65,535-row staging chunk: the double-buffered pipeline reuses buffers."""
    n, m = 64, 140_000
    rows = np.random.default_rng(7).integers(0, 256, size=(m, n // 4), dtype=np.uint8)
    path = tmp_path / "big.bed"
    path.write_bytes(MAGIC + rows.tobytes())
    a = gpu.Geno.load_bed(path, n, m)
    b = gpu.Geno.from_bed(rows, n)
    assert np.array_equal(a.download(), b.download())


def test_load_bed_rejects_foreign_and_short_files(gpu, tmp_path):
    p = make_problem(n=101, m=37, seed=4)
    bad = tmp_path / "bad.bed"
    bad.write_bytes(bytes([0x6C, 0x1B, 0x00]) + p.rows.tobytes())  # individual-major magic
    with pytest.raises(Exception, match="BED"):
        gpu.Geno.load_bed(bad, p.n, p.m)
    short = tmp_path / "short.bed"
    short.write_bytes(MAGIC + p.rows.tobytes()[:-3])
    with pytest.raises(Exception, match="BED"):
        gpu.Geno.load_bed(short, p.n, p.m)
    # SNPs before the truncation are still readable, as the reference reads up to the last one
    g = gpu.Geno.load_bed(short, p.n, p.m, None, np.arange(p.m - 1))
    assert g.download().shape[0] >= p.m - 1


def test_pred_writes_reference_gebv(gpu, tmp_path):
    p = make_problem(n=503, m=400, seed=9, missing=0.02)
    pre = tmp_path / "fs"
    (tmp_path / "fs.bed").write_bytes(MAGIC + p.rows.tobytes())
    (tmp_path / "fs.fam").write_text("".join(f"F{i} I{i} 0 0 0 -9\n" for i in range(p.n)))
    (tmp_path / "fs.bim").write_text("".join(f"1 s{j} 0 {1000 + j} A C\n" for j in range(p.m)))
    rng = np.random.default_rng(2)
    used = np.sort(rng.choice(p.m - 20, size=250, replace=False))  # the last 20 SNPs have no effect
    flip = rng.random(p.m) < 0.3
    alpha = rng.normal(size=p.m)
    lines = ["SNP,Chr,Pos,Allele1,Allele2,MAF,HWE_Pval,Group,Weight,Effect"]
    for j in used:
        lines.append(f"s{j},1,{1000 + j},{'C' if flip[j] else 'A'},x,0.1,0.5,NULL,1,{float(alpha[j])!r}")
    lines.append("s0,1,1000,A,C,0.1,0.5,NULL,1,")  # empty effect: skipped
    lines.append("short,row")
    est = tmp_path / "est.csv"
    est.write_text("\n".join(lines) + "\n")
    out = tmp_path / "gebv.csv"
    n, m = gpu.pred_plink(est, pre, out)
    assert (n, m) == (p.n, len(used))  # the later s0 row has an empty effect and is skipped
    want = np.zeros(p.n)
    for j in used:
        z = (p.codes[j] & 1) + (p.codes[j] >> 1)
        want += (2 - z if flip[j] else z).astype(np.float64) * alpha[j]
    rows = out.read_text().splitlines()
    assert rows[0] == "IID,GEBV" and len(rows) == p.n + 1
    got = np.array([float(r.split(",")[1]) for r in rows[1:]])
    assert [r.split(",")[0] for r in rows[1:]] == [f"I{i}" for i in range(p.n)]
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))  # 6 significant digits
    assert all(r.split(",")[1] == "%g" % w for r, w in zip(rows[1:], want))
