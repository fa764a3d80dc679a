"""Pin the oracle against the reference's own regression data (CPU only):
the HF fixture ``HF/data.h5:/result/Sigma1`` and the ``Fock-k`` symmetry
round trip of test/cu_symmetry_test.cu."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT, max_abs


def test_hf_fixture_sigma1(oracle, fx):
    """test/cu_solver_test.cpp:34-96: one HF step on the 27-k / 4-IBZ system
    with the stored (conj/trans-reduced) k-pair integrals; the solver output
    times -1 equals the stored Sigma1 to 1e-8."""
    NQ, ns = int(fx["params_NQ"]), int(fx["params_ns"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    F = oracle.hf_solve(dm, fx["hf_VQ_0"], sym, ns, S_ibz, float(fx["madelung"]))
    err = max_abs(-F, fx["hf_Sigma1"])
    print(f"HF fixture: oracle vs Sigma1 max-abs {err:.2e}")
    assert err < 1e-8


def test_hf_fixture_is_sensitive(oracle, fx):
    """A dropped Ewald term or a wrong exchange prefactor must fail the
    fixture check above (guards against a vacuous comparison)."""
    NQ, ns = int(fx["params_NQ"]), int(fx["params_ns"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    V = fx["hf_VQ_0"]
    J = oracle.hf_direct(dm, V, sym, ns)
    K = oracle.hf_exchange(dm, V, sym, ns)
    E = oracle.hf_ewald(dm, S_ibz, float(fx["madelung"]), ns)
    ref = fx["hf_Sigma1"]
    # hf_ewald returns the term with its sign (-c * madelung * S D S)
    assert max_abs(-(J + K + E), ref) < 1e-8
    assert max_abs(-(J + K), ref) > 1e-6           # Ewald term dropped
    assert max_abs(-(J + K - E), ref) > 1e-6       # Ewald sign flipped
    assert max_abs(-(J + 2 * K + E), ref) > 1e-6   # exchange prefactor doubled
    assert max_abs(-(J + K + 2 * E), ref) > 1e-6   # ns=1 Ewald factor used for ns=2


def test_fock_symmetry_roundtrip(oracle, fx):
    """test/cu_symmetry_test.cu:37-129: rotating the IBZ Fock matrices with
    U_k (and TR conjugation) reproduces the full-BZ Fock matrices; the
    reference accepts 1e-4 (cu_symmetry_test.cu:405)."""
    NQ = int(fx["params_NQ"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    F_full = fx["Fock_k"]
    F_ibz = F_full[:, sym.k_reduced_to_full]
    F_rot = oracle.ibz_to_full(F_ibz, sym)
    err = max_abs(F_rot, F_full)
    print(f"Fock-k rotation: max-abs {err:.2e}")
    assert err < 1e-4
    # identity on the IBZ representatives themselves
    assert max_abs(F_rot[:, sym.k_reduced_to_full], F_ibz) < 1e-13


GRID = os.path.join(ROOT, "tests", "golden", "fixtures", "gw_grid_beta100_lambda1e4.npz")


def gw_fixture_inputs(oracle, fx):
    g = np.load(GRID)
    NQ = int(fx["params_NQ"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    return sym, g["T_wt"], g["T_tw"], fx["gw_G_tau"], fx["gw_VQ_0"], fx["gw_Sigma_tau"]


def test_gw_grid_recovered_from_fixture(fx):
    """The committed tau grid is what tests/golden/make_gw_grid.py derives
    from the fixture's own G_tau (an exact HF Green's function)."""
    spec = importlib.util.spec_from_file_location("make_gw_grid", os.path.join(ROOT, "tests", "golden", "make_gw_grid.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    tau, asym, err = mg.recover_tau(mg.mode_occupations(fx))
    g = np.load(GRID)
    assert tau.size == fx["gw_G_tau"].shape[0] == 110
    assert np.array_equal(tau, g["tau"])
    assert asym < 1e-8 and err < 1e-8
    assert np.all(np.diff(tau) > 0)
    assert g["T_wt"].shape == (g["n_boson"].size, 108) and g["T_tw"].shape == (110, g["n_boson"].size)


@pytest.mark.parametrize("solver", ["lu", "cholesky"])
def test_gw_fixture_sigma_tau(oracle, fx, solver):
    """test/cu_solver_test.cpp:226-288: one GW iteration on the fixture
    system (27 k, 4 IBZ k and q, point-group U_k / U_q, stored k-pair
    integrals) reproduces ``GW/data.h5:/result/Sigma_tau`` within the
    reference's 1e-6; with the recovered grid the agreement is ~1e-10."""
    sym, T_wt, T_tw, G, V, ref = gw_fixture_inputs(oracle, fx)
    S = oracle.gw_solve(G, V, sym, T_wt, T_tw, solver=solver)
    err = max_abs(S, ref)
    print(f"GW fixture ({solver}): oracle vs Sigma_tau max-abs {err:.2e} (|Sigma| max {np.abs(ref).max():.2e})")
    assert err < 1e-6
    assert err < 1e-8


def test_gw_fixture_is_sensitive(oracle, fx):
    """Plausible mistakes move Sigma far beyond the fixture's tolerance:
    transposed integrals, U_q P U_q^dagger with the TR conjugation
    flags inverted or without the rotation at all, and a lost P0 mirror half."""
    sym, T_wt, T_tw, G, V, ref = gw_fixture_inputs(oracle, fx)
    solve = lambda s=sym, v=V, ttw=T_tw: oracle.gw_solve(G, v, s, T_wt, ttw, solver="lu")
    assert max_abs(solve(v=np.ascontiguousarray(np.swapaxes(V, -1, -2))), ref) > 1e-6
    # the fixture has no TR q-points: conjugating every rotated P must show
    all_conj = oracle.Symmetry(**{**sym.__dict__, "q_tr_conj": 1 - np.asarray(sym.q_tr_conj)})
    assert max_abs(solve(s=all_conj), ref) > 1e-6
    eye = np.broadcast_to(np.eye(sym.U_q.shape[-1]), sym.U_q.shape)
    no_rot = oracle.Symmetry(**{**sym.__dict__, "U_q": eye})
    assert max_abs(solve(s=no_rot), ref) > 1e-6
    T_half = T_tw.copy()
    T_half[T_tw.shape[0] // 2:] = 0
    assert max_abs(solve(ttw=T_half), ref) > 1e-6


X2C = os.path.join(ROOT, "tests", "golden", "fixtures", "ref_fixtures_x2c.npz")


@pytest.fixture(scope="module")
def fx2c():
    return dict(np.load(X2C))


def test_hf_x2c_fixture_sigma1(oracle, fx2c):
    """test/cu_solver_test.cpp:331-334 (HF_X2C): two-component HF on one
    k-point (nao 10, nso 20, NQ 154); solver output times -1 equals Sigma1."""
    sym = oracle.symmetry_from_fixture(fx2c, int(fx2c["params_NQ"]))
    S_ibz = fx2c["S_k"][:, sym.k_reduced_to_full]
    F = oracle.hf_solve_x2c(fx2c["hf_G_tau"][-1], fx2c["hf_VQ_0"], sym, S_ibz, float(fx2c["madelung"]))
    err = max_abs(-F, fx2c["hf_Sigma1"])
    print(f"HF X2C fixture: max-abs {err:.2e}")
    assert err < 1e-8


def test_gw_x2c_fixture_sigma_tau(oracle, fx2c):
    """test/cu_solver_test.cpp:317-328 (GW_X2C): two-component GW with the
    beta = 100 / Lambda = 1e4 grid recovered from the scalar GW fixture."""
    g = np.load(GRID)
    sym = oracle.symmetry_from_fixture(fx2c, int(fx2c["params_NQ"]))
    S = oracle.gw_solve_x2c(fx2c["gw_G_tau"], fx2c["gw_VQ_0"], sym, g["T_wt"], g["T_tw"], solver="lu")
    err = max_abs(S, fx2c["gw_Sigma_tau"])
    print(f"GW X2C fixture: max-abs {err:.2e} (|Sigma| max {np.abs(fx2c['gw_Sigma_tau']).max():.2e})")
    assert err < 1e-6
    assert err < 1e-8


def test_x2c_fixtures_are_sensitive(oracle, fx2c, monkeypatch):
    """Without the spin-block swap of G(beta - tau), or with K_ba = K_ab
    instead of its adjoint, the X2C results leave the fixture tolerances."""
    sym = oracle.symmetry_from_fixture(fx2c, int(fx2c["params_NQ"]))
    g = np.load(GRID)
    G, V = fx2c["gw_G_tau"], fx2c["gw_VQ_0"]
    nao = G.shape[-1] // 2
    Gf2 = oracle.g_full(G, sym)[:, 0]
    G4 = np.ascontiguousarray(np.swapaxes(oracle.spinor_blocks(Gf2, nao, oracle.X2C_BLOCKS), 1, 2))
    S4, _ = oracle._gw_core(G4, G4, 4, G.shape[0], sym, V, g["T_wt"], g["T_tw"], "lu", False)
    S = np.zeros_like(G)
    for ss, (a, b) in enumerate(oracle.X2C_BLOCKS):
        S[:, 0, :, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] = S4[:, ss]
    assert max_abs(S, fx2c["gw_Sigma_tau"]) > 1e-6
    K = oracle.hf_exchange_x2c(fx2c["hf_G_tau"][-1], fx2c["hf_VQ_0"], sym)
    assert max_abs(K[0, :, nao:, :nao], K[0, :, :nao, nao:]) > 1e-6
    assert max_abs(K[0, :, nao:, :nao], np.conj(np.swapaxes(K[0, :, :nao, nao:], -1, -2))) < 1e-14
