"""FP64 products emulated on the int8 matrix cores (precision "i8x5": five
signed 7-bit slices per element with a power-of-two scale per row / column,
exact int32 slice products up to slice diagonal 5, FP64 recombination),
through the C ABI, against the FP64 oracle.

The splitting keeps 35 bits of every element relative to its row (column)
maximum and drops the slice pairs beyond diagonal 5, about 2e-10 of max|C|
per product on Gaussian operands; a GW iteration chains six products and the
Dyson solve (with cancellation in the sums); measured 4e-10 .. 9e-9 of
max |Sigma| (the strongly coupled two-k fixture systems amplify most), so
Sigma and K are held to 2e-8 of their maximum -- three orders of magnitude
tighter than complex64 arithmetic on the same systems.  The plans emulate
every contraction site here (the built-in choice emulates only the Sigma
P.Y1 product)."""
import os

import numpy as np
import pytest

from conftest import ROOT, max_abs

pytestmark = pytest.mark.gpu

RTOL = 2e-8


CONTRACTION_SITES = ("p0_gv", "p0_gvh", "p0_contract", "sig_vg", "sig_py", "sig_contract")


def emulate_everywhere(ggpu, kern):
    v = [-1] * len(ggpu.SITES)
    for name in CONTRACTION_SITES:
        v[ggpu.SITES.index(name)] = 200
    kern.set_variants(v)
    assert all(kern.variant_name(kern.get_variants()[ggpu.SITES.index(n)]) == "int8" for n in CONTRACTION_SITES)


def gw_system(ggpu, synth, mesh, nao, NQ, ns, seed, beta=50.0, lam=1e3):
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, T_wt, T_tw = synth.load_ir(beta, lam)
    G, D = synth.make_green(sym.nk, nao, ns, beta, tau, seed)
    V = synth.make_integrals(sym.nk, NQ, nao, seed + 17)
    return sym, T_wt, T_tw, G, D, V


@pytest.mark.parametrize("case", [
    dict(mesh=(2, 1, 1), nao=3, NQ=5, ns=2, solver="lu"),
    dict(mesh=(3, 1, 1), nao=17, NQ=70, ns=1, solver="cholesky"),
    dict(mesh=(2, 2, 1), nao=13, NQ=40, ns=2, solver="cholesky"),
])
def test_gw_int8_emulation_small(ggpu, synth, oracle, case):
    mesh, nao, NQ, ns, solver = case["mesh"], case["nao"], case["NQ"], case["ns"], case["solver"]
    sym, T_wt, T_tw, G, D, V = gw_system(ggpu, synth, mesh, nao, NQ, ns, 91)
    print(f"i8x5 GW {case}", flush=True)
    k = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, precision="i8x5", solver=solver)
    k.load_integrals(V)
    emulate_everywhere(ggpu, k)
    Sig = k.solve(G)
    osym = oracle.symmetry_none(mesh, nao, NQ)
    Sig_or, parts = oracle.gw_solve(G, V, osym, T_wt, T_tw, solver=solver, return_parts=True)
    scale = np.abs(Sig_or).max()
    err = max_abs(Sig, Sig_or)
    P0 = k.polarization(G, 0, 0)
    perr = max_abs(P0, parts["P0"][0]) / np.abs(parts["P0"][0]).max()
    print(f"i8x5: Sigma max-abs {err:.3e} (rel {err / scale:.2e}), P0 rel {perr:.2e}", flush=True)
    assert err < RTOL * scale and perr < RTOL


def test_gw_int8_emulation_symmetric_fixture(ggpu, oracle, fx):
    """The reference's GW regression system (27 k, 4 IBZ k and q, stored
    conj / trans pairs, U_q rotations) against Sigma_tau (1e-6, the
    reference's own bar) and the oracle."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "fixtures", "gw_grid_beta100_lambda1e4.npz"))
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    G, V, ref = fx["gw_G_tau"], fx["gw_VQ_0"], fx["gw_Sigma_tau"]
    k = ggpu.GwKernel(nao, ns, NQ, g["T_wt"], g["T_tw"], sym, precision="i8x5", solver="lu")
    k.load_integrals(V)
    emulate_everywhere(ggpu, k)
    Sig = k.solve(G)
    err_ref = max_abs(Sig, ref)
    Sig_or = oracle.gw_solve(G, V, sym, g["T_wt"], g["T_tw"], solver="lu")
    err = max_abs(Sig, Sig_or)
    print(f"i8x5 fixture: vs Sigma_tau {err_ref:.2e}, vs oracle {err:.2e}", flush=True)
    assert err_ref < 1e-6 and err < RTOL * np.abs(Sig_or).max()


@pytest.mark.parametrize("nao,NQ", [(52, 300), (104, 600)])
def test_int8_emulation_at_workload_shapes(ggpu, synth, nao, NQ):
    """Sigma and K at the benchmark workloads' per-pair shapes (nk = 64 and
    nk = 216 systems on a two-k mesh) against the committed oracle values."""
    f = np.load(os.path.join(ROOT, "tests", "golden", "fixtures", f"gw_workload_nao{nao}_nq{NQ}_nk2.npz"))
    mesh = tuple(int(x) for x in f["mesh"])
    ns, beta = int(f["ns"]), float(f["beta"])
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, T_wt, T_tw = synth.load_ir(beta, float(f["lam"]))
    G, D = synth.make_green(sym.nk, nao, ns, beta, tau, int(f["seed_g"]))
    print(f"i8x5 workload shape nao={nao} NQ={NQ}", flush=True)
    k = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, precision="i8x5", solver="cholesky")
    k.synth_integrals(int(f["seed_v"]), 0.1)
    emulate_everywhere(ggpu, k)
    S = k.solve(G)[f["tau_idx"]]
    k.close()
    h = ggpu.HfKernel(nao, ns, NQ, sym, precision="i8x5")
    h.synth_integrals(int(f["seed_v"]), 0.1)
    v = [-1] * len(ggpu.SITES)
    v[ggpu.SITES.index("hf")] = 200
    h.set_variants(v)
    K = h.exchange(D)
    h.close()
    scale, kscale = float(f["sigma_abs_max"]), float(np.abs(f["K"]).max())
    err, errk = max_abs(S, f["sigma"]), max_abs(K, f["K"])
    print(f"i8x5: Sigma max-abs {err:.3e} (|Sigma| {scale:.3g}, rel {err / scale:.2e}), "
          f"K {errk:.3e} (rel {errk / kscale:.2e})", flush=True)
    assert err < RTOL * scale and errk < RTOL * kscale


def test_hf_int8_emulation_fixture(ggpu, oracle, fx):
    """Exchange products emulated on int8 with FP64 K sums, J and Ewald: the
    reference's HF fixture to 1e-8 and the oracle to 2e-8 of max |K|; E_HF
    to 2e-9 relative."""
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    h = ggpu.HfKernel(nao, ns, NQ, sym, madelung=float(fx["madelung"]), S_k=S_ibz, precision="i8x5")
    h.load_integrals(fx["hf_VQ_0"])
    v = [-1] * len(ggpu.SITES)
    v[ggpu.SITES.index("hf")] = 200            # emulate the exchange products
    h.set_variants(v)
    F = h.solve(dm)
    F_or = oracle.hf_solve(dm, fx["hf_VQ_0"], sym, ns, S_ibz, float(fx["madelung"]))
    Kf = oracle.hf_exchange(dm, fx["hf_VQ_0"], sym, ns)
    err = max_abs(F, F_or)
    E, E_or = h.energy(), oracle.hf_energy(dm, F_or, sym)
    print(f"i8x5 HF fixture: F max-abs {err:.2e}, E_HF {abs(E - E_or):.2e}", flush=True)
    assert max_abs(-F, fx["hf_Sigma1"]) < 1e-8
    assert err < RTOL * np.abs(Kf).max()
    assert abs(E - E_or) < 2e-9 * abs(E_or)


def test_int8_autotune_chooses_per_site(ggpu, synth, oracle):
    """In the i8x5 mode the autotuner times the int8 emulation against the
    FP64 kernel at each contraction site and keeps the faster; both are
    FP64-accurate, so the tuned iteration stays within the bound."""
    mesh, nao, NQ, ns = (2, 2, 1), 13, 40, 2
    sym, T_wt, T_tw, G, D, V = gw_system(ggpu, synth, mesh, nao, NQ, ns, 93)
    k = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, precision="i8x5", solver="cholesky")
    k.load_integrals(V)
    rep = k.autotune(G, rounds=1)
    for name in CONTRACTION_SITES:
        assert sorted(rep["sites"][name]["ms"]) == ["bk8_w3", "int8"]
        assert rep["sites"][name]["chosen"] == min(rep["sites"][name]["ms"], key=rep["sites"][name]["ms"].get)
    Sig = k.solve(G)
    ref = oracle.gw_solve(G, V, oracle.symmetry_none(mesh, nao, NQ), T_wt, T_tw, solver="cholesky")
    assert max_abs(Sig, ref) < RTOL * np.abs(ref).max()
