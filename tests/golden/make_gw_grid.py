#!/usr/bin/env python3
"""Derive the imaginary-time grid and tau <-> Omega transforms of the
reference's GW regression (test/cu_solver_test.cpp:226-288) from its own data.

The regression runs one GW iteration on ``GW/data.h5:/G_tau`` with
green-grids' sparse-IR file ``ir/1e4.h5`` (beta = 100, 108 fermionic
coefficients, cu_solver_test.cpp:231-234,262-267), which is not shipped with
the reference.  Its ``G_tau`` is the Hartree-Fock Green's function of the
``HF/Fock-k`` / ``HF/S-k`` matrices of ``GW/input.h5``:
in the generalised eigenbasis F C = S C e the projection C^H S G(tau) S C is
diagonal, with entries -d_m(tau), d_m(tau) = exp(-x_m tau) / (1 + exp(-beta x_m)).
So the grid is recovered from the data alone:

1. x_m = ln(d_m(0) / d_m(beta)) / beta for each mode (no chemical potential
   needed);
2. tau_i = -ln(d_m(tau_i) / d_m(0)) / x_m, taking per point the mode whose
   values carry the smallest relative round-off;
3. the grid is symmetric about beta/2 to the recovery's accuracy
   (asserted), and is symmetrised exactly.

The transforms T_wt (nw_b x (nts-2)) and T_tw (nts x nw_b) on that grid come
from this package's own IR construction (``ir_grid.transforms_on_grid``,
Lambda = 1e4).  Output: ``fixtures/gw_grid_beta100_lambda1e4.npz``.

Run:  python tests/golden/make_gw_grid.py   (reads only fixtures/ref_fixtures.npz)
"""
import importlib.util
import os

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FIX = os.path.join(HERE, "fixtures")
BETA, LAM = 100.0, 1e4


def _ir_grid():
    spec = importlib.util.spec_from_file_location("ir_grid", os.path.join(ROOT, "green-gpu_amd", "ir_grid.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def mode_occupations(fx):
    """d_m(tau) for every (spin, IBZ k, orbital) mode: (n_modes, nts)."""
    G = fx["gw_G_tau"]                       # (nts, ns, ink, nao, nao)
    F, S = fx["Fock_k"], fx["S_k"]           # (ns, nk, nao, nao)
    ibz = fx["sym_k_ibz2bz"]
    out = []
    for s in range(G.shape[1]):
        for ik, k in enumerate(ibz):
            Fk = 0.5 * (F[s, k] + F[s, k].conj().T)
            Sk = 0.5 * (S[s, k] + S[s, k].conj().T)
            _, C = sla.eigh(Fk, Sk)
            g = np.einsum("ai,ab,tbc,cd,dj->tij", C.conj(), Sk, G[:, s, ik], Sk, C)
            diag = np.einsum("tii->ti", g)
            off = np.abs(g - diag[:, :, None] * np.eye(g.shape[-1])).max()
            assert off < 1e-12, f"G_tau is not diagonal in the HF eigenbasis ({off:.1e})"
            out.extend(-diag.real.T)
    return np.array(out)


def recover_tau(d, beta=BETA):
    ulp = np.finfo(np.float64).eps
    x = np.log(d[:, 0] / d[:, -1]) / beta
    with np.errstate(divide="ignore", invalid="ignore"):
        est = -np.log(np.abs(d) / d[:, :1]) / x[:, None]
        err = ulp * (1 / np.abs(d) + 1 / d[:, :1] + 1 / np.abs(d[:, -1:])) / np.abs(x)[:, None]
    err = np.where(d > 0, err, np.inf)
    j = np.argmin(err, axis=0)
    tau = est[j, np.arange(d.shape[1])]
    asym = np.abs(tau + tau[::-1] - beta).max()
    assert asym < 1e-8, f"recovered grid is not symmetric ({asym:.1e})"
    tau = 0.5 * (tau + beta - tau[::-1])
    tau[0], tau[-1] = 0.0, beta
    return tau, asym, float(err[j, np.arange(d.shape[1])].max())


def main():
    fx = dict(np.load(os.path.join(FIX, "ref_fixtures.npz")))
    tau, asym, err = recover_tau(mode_occupations(fx))
    print(f"nts = {tau.size}, asymmetry before symmetrising {asym:.1e}, round-off bound {err:.1e}")
    tr = _ir_grid().transforms_on_grid(BETA, LAM, tau)
    print(f"nw_b = {tr['n_boson'].size}, IR basis size {tr['T_wt'].shape}")
    np.savez_compressed(os.path.join(FIX, "gw_grid_beta100_lambda1e4.npz"), tau=tau, n_boson=tr["n_boson"],
                        T_wt=tr["T_wt"], T_tw=tr["T_tw"], beta=BETA, lam=LAM)


if __name__ == "__main__":
    main()
