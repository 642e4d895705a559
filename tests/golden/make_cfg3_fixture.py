"""Writes tests/golden/cfg3_mgwr_n3000.npz: the oracle's GWRMultiscale result
on the multiscale-cosines workload at n = 3000 (k = 6, bisquare adaptive CV
for every parameter, CVR threshold 1e-5, predictors 1..5 centred, no hat
matrix).  The oracle takes a few minutes at this size, so the GPU test
(tests/test_gpu_workloads.py) reads the result from this fixture; the inputs
are regenerated from the seeded generator and checked against the stored
fingerprint.

Run from the repository root: python tests/golden/make_cfg3_fixture.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from libgwmodel_amd import synth  # noqa: E402
from oracle import pyoracle  # noqa: E402

N = 3000
EXTENT = 1000.0 * np.sqrt(N / 50_000)   # the point density of the n = 50,000 workload


def inputs():
    return synth.multiscale_cosines(n=N, extent=EXTENT)


def fingerprint(d):
    return np.array([d["coords"].sum(), d["X"].sum(), d["y"].sum(), (d["y"] ** 2).sum()])


def main():
    d = inputs()
    k = d["X"].shape[1]
    r = pyoracle.mgwr_run(d["coords"], d["X"], d["y"], [pyoracle.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k,
                          [pyoracle.CRIT_CV] * k, [0] + [1] * (k - 1), [1e-5] * k, has_hat=False, crit_type=0,
                          crit_threshold=1e-5)
    diag = np.array([r["diagnostic"][key] for key in ["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare",
                                                        "RSquareAdjust"]])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg3_mgwr_n3000.npz")
    np.savez_compressed(out, fingerprint=fingerprint(d), bandwidths=r["bandwidths"], betas=r["betas"], diag=diag,
                        rss=np.array([r["rss"]]), iterations=np.array([r["iterations"]]))
    print("wrote", out, r["bandwidths"], r["iterations"])


if __name__ == "__main__":
    main()
