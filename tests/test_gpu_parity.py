"""Parity of the gfx950 path (through the C ABI) with the CPU oracle and with
the reference's own regression fixture.

Tolerances: the plain FP64 path must match the FP64 oracle to 1e-10 max-abs
on Sigma(tau) and K(k) (and on the intermediates P0, P(Omega)); the HF
fixture is checked at the reference's own 1e-8 (test/cu_solver_test.cpp:95).
"""
import numpy as np
import pytest

from conftest import max_abs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def gw_system(ggpu, synth, mesh, nao, NQ, ns, beta, lam, seed):
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, T_wt, T_tw = synth.load_ir(beta, lam)
    G, D = synth.make_green(sym.nk, nao, ns, beta, tau, seed)
    V = synth.make_integrals(sym.nk, NQ, nao, seed + 17)
    return sym, tau, T_wt, T_tw, G, D, V


# ---------------------------------------------------------------------------
# HF
# ---------------------------------------------------------------------------
def test_hf_reference_fixture(ggpu, oracle, fx):
    """The reference's HF regression (test/cu_solver_test.cpp:34-96): full
    symmetry (27 k, 4 IBZ), stored k-pair integrals, result * -1 == Sigma1."""
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    V = fx["hf_VQ_0"]
    print(f"HF fixture: nk={sym.nk} ink={sym.ink} nao={nao} NQ={NQ} ns={ns}", flush=True)
    k = ggpu.HfKernel(nao, ns, NQ, sym, madelung=float(fx["madelung"]), S_k=S_ibz)
    k.load_integrals(V)
    F = k.solve(dm)
    ref = fx["hf_Sigma1"]
    assert max_abs(-F, ref) < 1e-8
    F_or = oracle.hf_solve(dm, V, sym, ns, S_ibz, float(fx["madelung"]))
    assert max_abs(F, F_or) < 1e-12
    K = k.exchange(dm)
    assert max_abs(K, oracle.hf_exchange(dm, V, sym, ns)) < 1e-12


@pytest.mark.parametrize("ns", [1, 2])
def test_hf_exchange_synthetic(ggpu, synth, oracle, ns):
    """HF exchange at the small workload's shape (nk=8, nao=26, NQ=100)."""
    sym, tau, T_wt, T_tw, G, D, V = gw_system(ggpu, synth, (2, 2, 2), 26, 100, ns, 50.0, 1e3, 100 + ns)
    print(f"HF synthetic ns={ns}", flush=True)
    k = ggpu.HfKernel(26, ns, 100, sym)
    k.load_integrals(V)
    K = k.exchange(D)
    osym = oracle.symmetry_none((2, 2, 2), 26, 100)
    K_or = oracle.hf_exchange(D, V, osym, ns)
    assert max_abs(K, K_or) < TOL


# ---------------------------------------------------------------------------
# GW
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", [
    dict(mesh=(2, 2, 1), nao=3, NQ=5, ns=2, solver="lu"),
    dict(mesh=(3, 1, 1), nao=5, NQ=37, ns=1, solver="cholesky"),
    dict(mesh=(2, 1, 1), nao=17, NQ=70, ns=2, solver="cholesky"),
])
def test_gw_small(ggpu, synth, oracle, case):
    mesh, nao, NQ, ns, solver = case["mesh"], case["nao"], case["NQ"], case["ns"], case["solver"]
    sym, tau, T_wt, T_tw, G, D, V = gw_system(ggpu, synth, mesh, nao, NQ, ns, 50.0, 1e3, 7)
    print(f"GW small {case}", flush=True)
    kern = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver=solver)
    kern.load_integrals(V)
    osym = oracle.symmetry_none(mesh, nao, NQ)
    Sig_or, parts = oracle.gw_solve(G, V, osym, T_wt, T_tw, solver=solver, return_parts=True)
    for iq in range(sym.inq):
        P0 = kern.polarization(G, iq, 0)
        assert max_abs(P0, parts["P0"][iq]) < TOL
        Pw = kern.polarization(G, iq, 1)
        assert max_abs(Pw, parts["Pw"][iq]) < TOL
    Sig = kern.solve(G)
    assert max_abs(Sig, Sig_or) < TOL


def test_gw_symmetry_fixture_system(ggpu, synth, oracle, fx):
    """GW on the reference fixture's symmetric system (27 k, 4 IBZ k and q,
    non-trivial U_k and U_q, conj/trans pair storage), synthetic G on the IBZ."""
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    tau, T_wt, T_tw = synth.load_ir(100.0, 1e4)
    G, D = synth.make_green(sym.ink, nao, ns, 100.0, tau, 5)
    V = fx["gw_VQ_0"]
    print(f"GW symmetry fixture: nk={sym.nk} ink={sym.ink} NQ={NQ}", flush=True)
    kern = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver="lu")
    kern.load_integrals(V)
    Sig = kern.solve(G)
    Sig_or = oracle.gw_solve(G, V, sym, T_wt, T_tw, solver="lu")
    assert max_abs(Sig, Sig_or) < TOL


def test_gw_small_workload(ggpu, synth, oracle):
    """The reference's CPU-runnable workload: nk=2x2x2, nao=26, NQ=100, ns=1,
    beta=50, IR Lambda=1e3; plain FP64 GW iteration vs the oracle at 1e-10."""
    sym, tau, T_wt, T_tw, G, D, V = gw_system(ggpu, synth, (2, 2, 2), 26, 100, 1, 50.0, 1e3, 1000)
    print("GW small workload nk=8 nao=26 NQ=100 ns=1", flush=True)
    kern = ggpu.GwKernel(26, 1, 100, T_wt, T_tw, sym, solver="lu")
    kern.load_integrals(V)
    Sig = kern.solve(G)
    osym = oracle.symmetry_none((2, 2, 2), 26, 100)
    Sig_or = oracle.gw_solve(G, V, osym, T_wt, T_tw, solver="lu")
    err = max_abs(Sig, Sig_or)
    print(f"max-abs Sigma error {err:.3e} (|Sigma|max {np.abs(Sig_or).max():.3e})", flush=True)
    assert err < TOL
