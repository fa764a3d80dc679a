"""Pin the oracle with what the mathematics fixes: the Lindhard bubble, the
symmetry-reduced vs unreduced calculation on the reference's symmetric
fixture system, the Dyson residual, linearity and spin bookkeeping."""
import numpy as np

from conftest import max_abs


def fermi(e, beta):
    return 1.0 / (1.0 + np.exp(beta * e))


def test_lindhard_one_orbital(oracle, synth):
    """nao = NQ = 1: P0(q, i Omega) must equal the Lindhard function
    pref/nk sum_k |v|^2 (n(e_{k+q}) - n(e_k)) / (i Omega + e_k - e_{k+q})
    (per spin; ns = 2 identical spins give pref = -1 each).  The reference
    evaluates P0 only for tau < beta/2 and mirrors it (cugw_qpt.cu:184-191),
    which is exact when e(k) = e(-k) and |v| is k-independent, as here."""
    beta, lam = 50.0, 1e3
    tau, T_wt, T_tw = synth.load_ir(beta, lam)
    nts = tau.size
    mesh = (3, 1, 1)
    sym = oracle.symmetry_none(mesh, 1, 1)
    eps = np.array([-0.7, 0.35, 0.35])                  # e(k) = e(-k) on the 3-point mesh
    f = synth.fermi_weight(eps, tau, beta)              # (nts, 3)
    ns = 2
    G = np.zeros((nts, ns, 3, 1, 1), dtype=complex)
    for s in range(ns):
        G[:, s, :, 0, 0] = -f
    rng = np.random.default_rng(9)
    V = 0.3 * np.exp(2j * np.pi * rng.uniform(size=(6, 1, 1, 1)))   # |v| = 0.3, random phases
    for k in range(3):                                   # diagonal pairs are Hermitian (real for 1x1)
        p = k * (k + 1) // 2 + k
        V[p] = 0.3
    Gf = oracle.g_full(G, sym)
    n_boson = np.load(__import__("os").path.join(synth.HERE, "data", "ir_beta50_lambda1000.npz"))["n_boson"]
    Om = 2 * np.pi * n_boson / beta
    for q in range(3):
        P0 = oracle.gw_p0(Gf, V, sym, q, ns, nts)
        Pw = oracle.tau_to_omega(P0, T_wt)[:, 0, 0]
        ref = np.zeros_like(Om, dtype=complex)
        for k in range(3):
            k1 = sym.k1_from_k2q[k, q]
            v = oracle.v_pair(V, sym, k, k1)[0, 0, 0]
            ea, eb = eps[k], eps[k1]
            if ea == eb:            # Omega = 0 limit: -beta n (1 - n); zero otherwise
                ref += ns * (-1.0) / 3 * abs(v) ** 2 * np.where(
                    Om == 0, beta * fermi(ea, beta) * (1 - fermi(ea, beta)), 0.0)
            else:
                ref += ns * (-1.0) / 3 * abs(v) ** 2 * (fermi(eb, beta) - fermi(ea, beta)) / (1j * Om + ea - eb)
        assert max_abs(Pw, ref) < 1e-10 * max(1.0, np.abs(ref).max())


def ibz_green_from_fock(fx, sym, tau, beta, synth):
    F = fx["Fock_k"][:, sym.k_reduced_to_full]            # (ns, ink, nao, nao)
    F = 0.5 * (F + np.conj(np.swapaxes(F, -1, -2)))
    e, W = np.linalg.eigh(F)
    f = synth.fermi_weight(e, tau, beta)                    # (ns, ink, nts, nao)
    G = -np.einsum("skai,skti,skbi->tskab", W, f, np.conj(W))
    return np.ascontiguousarray(G), np.ascontiguousarray(-G[-1])


def test_symmetry_reduced_matches_unreduced(oracle, synth, fx):
    """GW and HF on the reference's symmetric fixture (27 k, 4 IBZ k and q,
    point-group U_k / U_q, conj/trans pair storage) must agree with the same
    system unfolded to the full BZ without symmetry - this checks the G
    rotation, the pair unfolding and the q-rotation U_q P U_q^dagger.  The
    tolerance is the symmetry floor of the stored integrals (the reference
    accepts 1e-5 for its analogous no-symm / full-symm check,
    test/cu_solver_test.cpp:108-224)."""
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    beta = 50.0
    tau, T_wt, T_tw = synth.load_ir(beta, 1e3)
    # G from the fixture's own Fock matrices on the IBZ, so that G respects
    # the little group of every IBZ point (a random G would not)
    G, D = ibz_green_from_fock(fx, sym, tau, beta, synth)
    V = fx["gw_VQ_0"]
    Sig_sym = oracle.gw_solve(G, V, sym, T_wt, T_tw, solver="lu")

    nk = sym.nk
    Gfull = np.ascontiguousarray(np.transpose(oracle.g_full(G, sym), (2, 1, 0, 3, 4)))   # (nts, ns, nk)
    pairs = [(k1, k2) for k1 in range(nk) for k2 in range(k1 + 1)]
    V_all = np.stack([oracle.v_pair(V, sym, k1, k2) for k1, k2 in pairs])
    k1m, k2m = oracle.mesh_maps(fx["sym_k_mesh_scaled"])
    npair = len(pairs)
    ar = np.arange(nk)
    none = oracle.Symmetry(
        nk=nk, ink=nk, nq=nk, inq=nk, k_full_to_reduced=ar, k_reduced_to_full=ar,
        k_tr_conj=np.zeros(nk, int), k_stars=[np.array([k]) for k in ar],
        q_full_to_reduced=ar, q_reduced_to_full=ar, q_tr_conj=np.zeros(nk, int),
        q_stars=[np.array([q]) for q in ar], k1_from_k2q=k1m, k2_from_k1q=k2m,
        U_k=np.broadcast_to(np.eye(nao), (nk, nao, nao)), U_q=np.broadcast_to(np.eye(NQ), (nk, NQ, NQ)),
        conj_kpair=np.arange(npair), trans_kpair=np.arange(npair), irre_pos=np.arange(npair),
        num_kpair_stored=npair)
    Sig_full = oracle.gw_solve(Gfull, V_all, none, T_wt, T_tw, solver="lu")
    scale = np.abs(Sig_full).max()
    err = max_abs(Sig_sym, Sig_full[:, :, sym.k_reduced_to_full])
    print(f"Sigma: symmetric vs unreduced max-abs {err:.2e} (|Sigma| max {scale:.2e})")
    assert err < 1e-5 * max(1.0, scale)

    dm = D                                   # (ns, ink, nao, nao)
    K_sym = oracle.hf_exchange(dm, V, sym, ns)
    K_full = oracle.hf_exchange(oracle.ibz_to_full(dm, sym), V_all, none, ns)
    errk = max_abs(K_sym, K_full[:, sym.k_reduced_to_full])
    print(f"K: symmetric vs unreduced max-abs {errk:.2e}")
    assert errk < 1e-5


def test_dyson_residual_and_solvers(oracle, synth):
    rng = np.random.default_rng(3)
    n, nw = 40, 5
    A = rng.normal(size=(nw, n, n)) + 1j * rng.normal(size=(nw, n, n))
    P0 = -(A @ np.conj(np.swapaxes(A, 1, 2))) / n                  # Hermitian negative definite
    Pc = oracle.gw_dyson(P0, "cholesky")
    Pl = oracle.gw_dyson(P0, "lu")
    assert max_abs(Pc, Pl) < 1e-12
    eye = np.eye(n)
    for w in range(nw):
        assert max_abs((eye - P0[w]) @ Pc[w], P0[w]) < 1e-12
        # P = (1 - P0)^-1 P0 has eigenvalues lambda / (1 - lambda) in (-1, 0]
        ev = np.linalg.eigvalsh(oracle.herm(Pc[w]))
        assert ev.max() < 1e-12 and ev.min() > -1.0


def test_linearity_and_spin_bookkeeping(oracle, synth):
    mesh, nao, NQ = (2, 1, 1), 4, 7
    sym = oracle.symmetry_none(mesh, nao, NQ)
    tau, T_wt, T_tw = synth.load_ir(50.0, 1e3)
    G1, D1 = synth.make_green(2, nao, 1, 50.0, tau, 5)
    V = synth.make_integrals(2, NQ, nao, 6)
    # K is linear in the density matrix
    _, D2 = synth.make_green(2, nao, 1, 50.0, tau, 7)
    K = lambda D: oracle.hf_exchange(D, V, sym, 1)
    assert max_abs(K(D1 + 2.5 * D2), K(D1) + 2.5 * K(D2)) < 1e-12
    # ns = 1 with pref -2 equals ns = 2 with two identical spins (pref -1 each)
    G2 = np.concatenate([G1, G1], axis=1)
    S1 = oracle.gw_solve(G1, V, sym, T_wt, T_tw, solver="cholesky")
    S2 = oracle.gw_solve(G2, V, sym, T_wt, T_tw, solver="cholesky")
    assert max_abs(S2[:, 0], S1[:, 0]) < 1e-12 and max_abs(S2[:, 1], S1[:, 0]) < 1e-12
    # no interaction, no self-energy
    Z = oracle.gw_solve(G1, np.zeros_like(V), sym, T_wt, T_tw, solver="lu")
    assert np.abs(Z).max() == 0.0
    # P0(tau) is Hermitian and mirror-symmetric
    P0 = oracle.gw_p0(oracle.g_full(G1, sym), V, sym, 1, 1, tau.size)
    assert max_abs(P0, oracle.herm(P0)) < 1e-12
    assert max_abs(P0, P0[::-1]) == 0.0


def test_hf_energy_one_orbital(oracle):
    """One k, one orbital, one auxiliary vector: J, K and the Ewald term are
    scalars (J = v^2 sum_s d_s, K_s = px v^2 d_s, Ewald_s = -c m s^2 d_s with
    px = -1, c = 1 for ns = 2 and px = c = -1/2, 1/2 for ns = 1), and the
    two-electron energy is E = 1/2 sum_s d_s F_s."""
    v, m, s_ov = 0.3, 0.5, 1.1
    sym = oracle.symmetry_none((1, 1, 1), 1, 1)
    V = np.full((1, 1, 1, 1), v, dtype=complex)
    for ns, d in ((2, (0.7, 0.4)), (1, (1.3,))):
        dm = np.array(d, dtype=complex).reshape(ns, 1, 1, 1)
        S = np.full((ns, 1, 1, 1), s_ov, dtype=complex)
        F = oracle.hf_solve(dm, V, sym, ns, S, m)
        px, c = (-1.0, 1.0) if ns == 2 else (-0.5, 0.5)
        hand = sum(0.5 * ds * (v * v * sum(d) + px * v * v * ds - c * m * s_ov ** 2 * ds) for ds in d)
        assert abs(oracle.hf_energy(dm, F, sym) - hand) < 1e-15


def test_hf_energy_star_weights(oracle, fx):
    """On the reference's HF fixture (27 k, 4 IBZ k): the IBZ sum with star
    weights equals the direct sum over all 27 k of the rotated density and
    Fock matrices, and an unweighted IBZ sum does not."""
    NQ, ns = int(fx["params_NQ"]), int(fx["params_ns"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    S_ibz = fx["S_k"][:, sym.k_reduced_to_full]
    dm = fx["hf_G_tau"][-1]
    F = oracle.hf_solve(dm, fx["hf_VQ_0"], sym, ns, S_ibz, float(fx["madelung"]))
    E = oracle.hf_energy(dm, F, sym)
    Df, Ff = oracle.ibz_to_full(dm, sym), oracle.ibz_to_full(F, sym)
    E_full = 0.5 * np.einsum("skij,skji->", Df, Ff).real / sym.nk
    print(f"E_HF {E:.12f}, full-zone sum {E_full:.12f}")
    assert abs(E - E_full) < 1e-10 * abs(E_full)
    E_flat = 0.5 * np.einsum("skij,skji->", dm, F).real / sym.nk
    assert abs(E_flat - E_full) > 1e-3 * abs(E_full)
