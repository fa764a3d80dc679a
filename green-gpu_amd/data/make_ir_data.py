#!/usr/bin/env python3
"""Regenerate the committed tau <-> bosonic-Matsubara transforms used by the
synthetic workloads (see ir_grid.py):

* beta = 50, Lambda = 1e3 (the small CPU-runnable system): this package's
  own IR basis and sampling points;
* beta = 100, Lambda = 1e4 (the large systems): the tau grid of the
  reference's green-grids file ir/1e4.h5 (108 sampling points + endpoints),
  as recovered from the reference's GW regression fixture by
  tests/golden/make_gw_grid.py, with transforms from this package's basis on
  that grid (ir_grid.transforms_on_grid)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import ir_grid  # noqa: E402


def main():
    d = ir_grid.make_transforms(50.0, 1e3)
    ref = np.load(os.path.join(ROOT, "tests", "golden", "fixtures", "gw_grid_beta100_lambda1e4.npz"))
    e = ir_grid.transforms_on_grid(100.0, 1e4, ref["tau"])
    for (beta, lam), x in (((50.0, 1e3), d), ((100.0, 1e4), e)):
        name = os.path.join(HERE, f"ir_beta{int(beta)}_lambda{int(lam)}.npz")
        np.savez(name, **x)
        print(name, "nts", x["tau"].size, "nw_b", x["T_wt"].shape[0])


if __name__ == "__main__":
    main()
