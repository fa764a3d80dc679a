#!/usr/bin/env python3
"""Regenerate the committed tau <-> bosonic-Matsubara transforms used by the
synthetic workloads: beta = 50, Lambda = 1e3 (the small CPU-runnable system)
and beta = 100, Lambda = 1e4 (the large systems).  See ir_grid.py."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import ir_grid  # noqa: E402

for beta, lam in ((50.0, 1e3), (100.0, 1e4)):
    d = ir_grid.make_transforms(beta, lam)
    name = os.path.join(HERE, f"ir_beta{int(beta)}_lambda{int(lam)}.npz")
    np.savez(name, **d)
    print(name, "nts", d["tau"].size, "nw_b", d["T_wt"].shape[0])
