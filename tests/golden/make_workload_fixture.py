#!/usr/bin/env python3
"""Oracle results at the benchmark workloads' per-pair dimensions on a two-k
mesh, so the GPU parity tests cover the production GEMM shapes (tile edges,
K-batches over Q, the K > 64 split-bf16 path at nao = 104) without running
the oracle at test time.  beta = 100, IR Lambda = 1e4 (nts = 110, nw_b = 99),
ns = 2, FP64 Cholesky iteration.

Inputs are regenerated from their seeds by the test: G from
``synth.make_green(2, nao, 2, 100.0, tau, seed_g)``, integrals from
``synth.make_integrals(2, NQ, nao, seed_v)`` (the device generator gives the
same values).  Stored: Sigma(tau) at a few tau indices (the full tensor is
10-40 MB), the HF exchange K for D = -G(beta), max |Sigma| and max |K|.

    python tests/golden/make_workload_fixture.py                # nao 52, NQ 300 (nk = 64 workload)
    python tests/golden/make_workload_fixture.py --nao 104 --NQ 600 --tau 0,109
    python tests/golden/make_workload_fixture.py --nao 104 --NQ 420 --tau 55
"""
import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

MESH, NS, BETA, LAM = (2, 1, 1), 2, 100.0, 1e4
SEED_G, SEED_V = 7, 8


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nao", type=int, default=52)
    ap.add_argument("--NQ", type=int, default=300)
    ap.add_argument("--tau", default="0,27,55,82,109", help="tau indices of Sigma to store")
    a = ap.parse_args()
    out = os.path.join(HERE, "fixtures", f"gw_workload_nao{a.nao}_nq{a.NQ}_nk2.npz")
    tau_idx = np.array([int(x) for x in a.tau.split(",")])
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gw_oracle as O
    synth = _load("synth", os.path.join(ROOT, "green-gpu_amd", "synth.py"))
    tau, T_wt, T_tw = synth.load_ir(BETA, LAM)
    nk = MESH[0] * MESH[1] * MESH[2]
    G, D = synth.make_green(nk, a.nao, NS, BETA, tau, SEED_G)
    V = synth.make_integrals(nk, a.NQ, a.nao, SEED_V)
    osym = O.symmetry_none(MESH, a.nao, a.NQ)
    t = time.time()
    Sigma = O.gw_solve(G, V, osym, T_wt, T_tw, solver="cholesky")
    K = O.hf_exchange(D, V, osym, NS)
    np.savez_compressed(out, mesh=np.array(MESH), nao=a.nao, NQ=a.NQ, ns=NS, beta=BETA, lam=LAM,
                        seed_g=SEED_G, seed_v=SEED_V, tau_idx=tau_idx, sigma=Sigma[tau_idx], K=K,
                        sigma_abs_max=np.abs(Sigma).max(), k_abs_max=np.abs(K).max())
    print(f"wrote {out} in {time.time() - t:.0f}s: max |Sigma| {np.abs(Sigma).max():.4g}, "
          f"max |K| {np.abs(K).max():.4g}")


if __name__ == "__main__":
    main()
