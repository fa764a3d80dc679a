#!/usr/bin/env python3
"""Oracle results at the bench workload's per-pair dimensions (nao = 52,
NQ = 300, ns = 2, beta = 100, IR Lambda = 1e4: nts = 110, nw_b = 99) on a
two-k mesh, so the GPU parity tests cover the production GEMM shapes (tile
edges at M = 300 and N = 2704, K-batches over 300 Q) without running the
oracle (about a minute here) at test time.

Inputs are regenerated from their seeds by the test: G from
``synth.make_green(2, 52, 2, 100.0, tau, 7)``, integrals from
``synth.make_integrals(2, 300, 52, 8)``.  Stored: Sigma(tau) of one FP64
Cholesky iteration at five tau indices (the full tensor is 9.5 MB) and the
HF exchange K for D = -G(beta).

Run:  python tests/golden/make_workload_fixture.py
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fixtures", "gw_workload_nao52_nq300_nk2.npz")

MESH, NAO, NQ, NS, BETA, LAM = (2, 1, 1), 52, 300, 2, 100.0, 1e4
SEED_G, SEED_V = 7, 8
TAU_IDX = np.array([0, 27, 55, 82, 109])


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gw_oracle as O
    synth = _load("synth", os.path.join(ROOT, "green-gpu_amd", "synth.py"))
    tau, T_wt, T_tw = synth.load_ir(BETA, LAM)
    nk = MESH[0] * MESH[1] * MESH[2]
    G, D = synth.make_green(nk, NAO, NS, BETA, tau, SEED_G)
    V = synth.make_integrals(nk, NQ, NAO, SEED_V)
    osym = O.symmetry_none(MESH, NAO, NQ)
    Sigma = O.gw_solve(G, V, osym, T_wt, T_tw, solver="cholesky")
    K = O.hf_exchange(D, V, osym, NS)
    np.savez_compressed(OUT, mesh=np.array(MESH), nao=NAO, NQ=NQ, ns=NS, beta=BETA, lam=LAM,
                        seed_g=SEED_G, seed_v=SEED_V, tau_idx=TAU_IDX, sigma=Sigma[TAU_IDX], K=K,
                        sigma_abs_max=np.abs(Sigma).max())
    print(f"wrote {OUT}: max |Sigma| {np.abs(Sigma).max():.4g}, max |K| {np.abs(K).max():.4g}")


if __name__ == "__main__":
    main()
