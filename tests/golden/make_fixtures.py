"""Writes the golden fixtures of tests/golden/ from the data files the reference ships with its
worked example (examples/data and examples/expected/slemm of jiang18/slemm v0.90.1).

Only data are taken (phenotypes, the run's outputs); the example's genotype file 10k.bed is not
in the reference tree, so these fixtures pin the output-level identities of the REML path, not
a rerun:
  * 10k.reml.vc.csv / 10k.reml.py.txt / 10k.slemm.csv: Vg = y'Py/(n-c) (src/reml.cpp:856),
    h2 = Vg/(Vg+Ve), sum(Py) ~ 0 (projection, src/reml.cpp:718), BLUE intercept;
  * 10k.gstat.txt: ChiSq = xPy^2/(Vg xPx) (src/output.cpp:105,115), MAC = int(MAF*2n+0.5) (:112),
    Pval = erfc(sqrt(ChiSq/2)) (src/chisq.h:14-18);
  * 10k.gamma.txt: CHISQ = (BETA/SE)^2 in float32 and P = chi2.sf(CHISQ, 1) (bin/slemm_gamma.py:
    321-325), the text of each line (:327-361);
  * 10k.gstat.txt lines as written (src/output.cpp:85-120);
  * 10k.reml.snp.csv: MAF values are k/(2n) (src/weight.cpp:12-14).
Run from the repository root:  python tests/golden/make_fixtures.py [reference_root]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def main(ref_root: str = "/root/reference") -> None:
    ex = Path(ref_root) / "examples"
    exp = ex / "expected" / "slemm"

    # phenotype (IID,QT) and Py (IID Py intercept R)
    ph = {}
    for line in (ex / "data" / "10k.slemm.csv").read_text().splitlines()[1:]:
        iid, qt = line.split(",")
        ph[iid] = float(qt)
    iids, py = [], []
    for line in (exp / "10k.reml.py.txt").read_text().splitlines()[1:]:
        f = line.split()
        iids.append(f[0])
        py.append(float(f[1]))
    y = np.array([ph[i] for i in iids], dtype=np.float64)
    vc = dict(l.split(",") for l in (exp / "10k.reml.vc.csv").read_text().splitlines())
    blue = (exp / "10k.reml.blue.csv").read_text().splitlines()[1].split(",")
    np.savez_compressed(HERE / "ex10k_reml.npz", y=y.astype(np.float32), Py=np.array(py, np.float32))
    meta = {"pop_size": int(vc["pop_size"]), "num_markers": int(vc["num_markers"]), "Vg": float(vc["Vg"]),
            "Ve": float(vc["Ve"]), "h2": float(vc["h2"]), "LRT": float(vc["LRT"]), "pval": float(vc["pval"]),
            "blue_intercept": float(blue[1]), "c": 1}

    # gstat rows
    gs = [l.split("\t") for l in (exp / "10k.gstat.txt").read_text().splitlines()]
    hdr, rows = gs[0], gs[1:]
    col = {h: i for i, h in enumerate(hdr)}
    meta["gstat"] = [{k: r[col[k]] for k in ("variant", "MAC", "MAF", "xPy", "xPx", "xBx", "xSx", "ChiSq", "Pval")}
                     for r in rows]
    meta["gstat_lines"] = (exp / "10k.gstat.txt").read_text().splitlines()

    # first rows of the GRAMMAR-gamma output
    gam = (exp / "10k.gamma.txt").read_text().splitlines()
    ghdr = gam[0].split("\t")
    meta["gamma_header"] = ghdr
    meta["gamma_rows"] = [r.split("\t") for r in gam[1:201]]

    # MAF column of the SNP effect file
    snp = (exp / "10k.reml.snp.csv").read_text().splitlines()
    meta["snp_header"] = snp[0].split(",")
    meta["maf"] = [float(l.split(",")[5]) for l in snp[1:2001]]
    meta["effect_rows"] = [l.split(",") for l in snp[1:21]]
    (HERE / "ex10k_meta.json").write_text(json.dumps(meta, indent=1))
    print("wrote", HERE / "ex10k_reml.npz", HERE / "ex10k_meta.json")


if __name__ == "__main__":
    main(*sys.argv[1:])
