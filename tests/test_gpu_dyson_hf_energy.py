"""GPU checks of the Dyson solvers' semantics, the HF energy and the
intermediates read back after an iteration, through the C ABI.

* LU (getrf / getrs with partial pivoting, cugw_qpt.cu:266-322) returns the
  oracle's LU solution where 1 - P0(Omega) is indefinite, and Cholesky
  reports GGPU_ERR_NOT_POSITIVE_DEFINITE there (a status code, not a fault);
* E_HF of the device (FP64 traces, star weights) equals the oracle's
  definition on the reference's scalar and two-component HF fixtures;
* ggpu_gw_last_polarization returns the intermediates of the iteration's
  last q without recomputing them.
"""
import numpy as np
import pytest

from conftest import max_abs

pytestmark = pytest.mark.gpu


def indefinite_system(oracle, synth):
    """G whose tau >= beta/2 half is negated makes P0 positive; integrals
    scaled so that the largest eigenvalue of herm P0(Omega) is 3, so
    1 - P0(Omega) has eigenvalues near -2 (indefinite, nonsingular)."""
    mesh, nao, NQ, ns = (2, 1, 1), 6, 20, 2
    tau, T_wt, T_tw = synth.load_ir(50.0, 1e3)
    G, _ = synth.make_green(2, nao, ns, 50.0, tau, 11)
    G[tau.size // 2:] *= -1.0
    V = synth.make_integrals(2, NQ, nao, 12)
    osym = oracle.symmetry_none(mesh, nao, NQ)
    P0w = oracle.tau_to_omega(oracle.gw_p0(oracle.g_full(G, osym), V, osym, 0, ns, tau.size), T_wt)
    lam = max(np.linalg.eigvalsh(oracle.herm(P0w[w])).max() for w in range(P0w.shape[0]))
    assert lam > 0
    V = V * np.sqrt(3.0 / lam)
    return mesh, nao, NQ, ns, tau, T_wt, T_tw, G, V, osym


def test_lu_solves_indefinite_dyson_and_cholesky_reports_it(ggpu, synth, oracle):
    mesh, nao, NQ, ns, tau, T_wt, T_tw, G, V, osym = indefinite_system(oracle, synth)
    sym = ggpu.mesh_symmetry_none(mesh)
    Sig_or, parts = oracle.gw_solve(G, V, osym, T_wt, T_tw, solver="lu", return_parts=True)
    eig = min(np.linalg.eigvalsh(np.eye(NQ) - oracle.herm(P)).min() for P in parts["P0w"][0])
    print(f"smallest eigenvalue of 1 - P0(Omega) over Omega: {eig:.3f}", flush=True)
    assert eig < -0.5
    lu = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver="lu")
    lu.load_integrals(V)
    for iq in range(sym.inq):
        Pw = lu.polarization(G, iq, 1)
        scale = np.abs(parts["Pw"][iq]).max()
        assert max_abs(Pw, parts["Pw"][iq]) < 1e-10 * max(1.0, scale)
    Sig = lu.solve(G)
    scale = np.abs(Sig_or).max()
    err = max_abs(Sig, Sig_or)
    print(f"LU: Sigma max-abs {err:.2e} (|Sigma| {scale:.3g})", flush=True)
    assert err < 1e-10 * max(1.0, scale)
    ch = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver="cholesky")
    ch.load_integrals(V)
    with pytest.raises(ggpu.GgpuError) as e:
        ch.solve(G)
    assert e.value.code == 4 and "not positive definite" in str(e.value)


def test_lu_matches_cholesky_on_positive_definite_systems(ggpu, synth, oracle):
    """On a physical G (1 - P0 Hermitian positive definite) the pivoted LU and
    the Cholesky solve agree to FP64 rounding, at NQ above the 32-column
    block (several panels, trailing updates and row interchanges)."""
    mesh, nao, NQ, ns = (2, 1, 1), 9, 75, 2
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, T_wt, T_tw = synth.load_ir(50.0, 1e3)
    G, _ = synth.make_green(2, nao, ns, 50.0, tau, 21)
    V = synth.make_integrals(2, NQ, nao, 22)
    out = {}
    for solver in ("lu", "cholesky"):
        k = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver=solver)
        k.load_integrals(V)
        out[solver] = k.solve(G)
    scale = np.abs(out["lu"]).max()
    assert max_abs(out["lu"], out["cholesky"]) < 1e-12 * scale
    ref = oracle.gw_solve(G, V, oracle.symmetry_none(mesh, nao, NQ), T_wt, T_tw, solver="lu")
    assert max_abs(out["lu"], ref) < 1e-10


def test_hf_energy_matches_oracle_on_fixtures(ggpu, oracle, fx):
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    k = ggpu.HfKernel(nao, ns, NQ, sym, madelung=float(fx["madelung"]), S_k=S_ibz)
    k.load_integrals(fx["hf_VQ_0"])
    with pytest.raises(ggpu.GgpuError) as e:
        k.energy()
    assert e.value.code == 5
    F = k.solve(dm)
    F_or = oracle.hf_solve(dm, fx["hf_VQ_0"], sym, ns, S_ibz, float(fx["madelung"]))
    E, E_or = k.energy(), oracle.hf_energy(dm, F_or, sym)
    print(f"E_HF device {E:.15f}, oracle {E_or:.15f}", flush=True)
    assert abs(E - E_or) < 1e-12 * abs(E_or)
    # k shares: the energies of the shares sum to the total
    parts = []
    for r in range(3):
        k.set_k_share(r, 3)
        k.solve(dm)
        parts.append(k.energy())
    assert abs(sum(parts) - E) < 1e-12 * abs(E)


def test_hf_energy_x2c_fixture(ggpu, oracle):
    import os
    from conftest import ROOT
    f = dict(np.load(os.path.join(ROOT, "tests", "golden", "fixtures", "ref_fixtures_x2c.npz")))
    nao, nso, NQ = int(f["params_nao"]), int(f["params_nso"]), int(f["params_NQ"])
    sym = oracle.symmetry_from_fixture(f, NQ)
    S_ibz = f["S_k"][:, sym.k_reduced_to_full]
    dm = f["hf_G_tau"][-1]
    k = ggpu.HfKernel(nao, 1, NQ, sym, madelung=float(f["madelung"]), S_k=S_ibz, nso=nso)
    k.load_integrals(f["hf_VQ_0"])
    k.solve(dm)
    F_or = oracle.hf_solve_x2c(dm, f["hf_VQ_0"], sym, S_ibz, float(f["madelung"]))
    E_or = oracle.hf_energy(dm, F_or, sym)
    assert abs(k.energy() - E_or) < 1e-12 * abs(E_or)


def test_last_polarization_reads_back_the_last_q(ggpu, synth, oracle):
    mesh, nao, NQ, ns = (3, 1, 1), 7, 33, 2
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, T_wt, T_tw = synth.load_ir(50.0, 1e3)
    G, _ = synth.make_green(3, nao, ns, 50.0, tau, 31)
    V = synth.make_integrals(3, NQ, nao, 32)
    k = ggpu.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, solver="cholesky")
    k.load_integrals(V)
    with pytest.raises(ggpu.GgpuError) as e:
        k.last_polarization(0)
    assert e.value.code == 5
    k.solve(G)
    got = {w: k.last_polarization(w) for w in (0, 1, 2)}
    assert all(iq == sym.inq - 1 for iq, _ in got.values())
    for w in (0, 1, 2):
        assert np.array_equal(got[w][1], k.polarization(G, sym.inq - 1, w))
    _, parts = oracle.gw_solve(G, V, oracle.symmetry_none(mesh, nao, NQ), T_wt, T_tw, solver="cholesky",
                               return_parts=True)
    assert max_abs(got[0][1], parts["P0"][-1]) < 1e-10
    assert max_abs(got[2][1], parts["Pt"][-1]) < 1e-10
    # a pair-batch limit gives a partial P0 (fewer pairs) and is undone by 0
    k.set_pair_batch_limit(1)
    k.solve(G)
    k.set_pair_batch_limit(0)
    k.solve(G)
    assert np.array_equal(k.last_polarization(0)[1], got[0][1])
