"""CPU FP64 oracle for the HF-exchange and GW self-energy path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The shipped library never calls it.

This is a numpy restatement (complex128, LAPACK/BLAS through numpy/scipy) of
what the reference green-gpu plugin computes.  Every function names the
reference lines it follows.  All tensors are row-major, with the reference's
host layouts:

* ``G`` / ``Sigma``: (nts, ns, ink, nao, nao)   (gw_gpu_kernel.cpp:163-205)
* density matrix / Fock: (ns, ink, nao, nao)     (hf_gpu_kernel.cpp:66-104)
* stored integrals: (num_kpair_stored, NQ, nao, nao), pair k1 >= k2
  (df_integral_t.h:68-144)
* ``T_wt``: (nw_b, nts-2), ``T_tw``: (nts, nw_b)  (cugw_qpt.cu:329-358)

Pinning (what checks this oracle independently of its own formulas):

* ``hf_solve``, ``ibz_to_full``, ``v_pair``: the reference's HF regression
  fixture ``HF/data.h5:/result/Sigma1`` at 1e-8 (test/cu_solver_test.cpp:34-96)
  and the ``Fock-k`` symmetry-rotation fixture (test/cu_symmetry_test.cu:37-129).
* ``gw_*`` contractions: a replay of the reference's exact cuBLAS call
  sequence (operation flags, leading dimensions, strides) through a
  column-major BLAS emulator (tests/test_oracle_blas_replay.py); brute-force
  per-element loops on tiny shapes; the analytic Lindhard bubble for a
  one-orbital model; and symmetry-reduced vs unreduced runs on the fixture
  system (the q-rotation U_q P U_q^dagger).
* ``gw_dyson``: LAPACK semantics (residual of (I-P0) P = P0, Cholesky vs LU).

* ``gw_solve`` end to end: the reference's GW regression fixture
  ``GW/data.h5:/result/Sigma_tau`` (test/cu_solver_test.cpp:226-315) on the
  tau grid recovered from that fixture (tests/golden/make_gw_grid.py), to
  ~1e-10 against the reference's 1e-6 (tests/test_oracle_fixtures.py); the
  two-component variant on ``GW_X2C/data.h5``.
* ``hf_energy``: the reference defines no energy (green-mbpt does), so its
  definition here is fixed by the mathematics: a hand-computed one-orbital
  system (J, K and the Ewald term as scalars) and the symmetry-reduced sum
  with star weights against the direct full-Brillouin-zone sum on the HF
  fixture (tests/test_oracle_physics.py).
"""
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla


# ----------------------------------------------------------------------------
# Symmetry tables (the plain data the reference uploads: cu_symmetry.h:41-66,
# plus the k-pair tables of df_integral_t.h:68-86)
# ----------------------------------------------------------------------------
@dataclass
class Symmetry:
    nk: int
    ink: int
    nq: int
    inq: int
    k_full_to_reduced: np.ndarray
    k_reduced_to_full: np.ndarray
    k_tr_conj: np.ndarray
    k_stars: list
    q_full_to_reduced: np.ndarray
    q_reduced_to_full: np.ndarray
    q_tr_conj: np.ndarray
    q_stars: list
    k1_from_k2q: np.ndarray          # (nk, nq): k1 = k2 + q
    k2_from_k1q: np.ndarray          # (nk, nq): k2 = k1 - q
    U_k: np.ndarray                  # (nk, nao, nao) AO representation
    U_q: np.ndarray                  # (nq, NQ, NQ) aux-basis representation for P0
    conj_kpair: np.ndarray           # (nk(nk+1)/2,)
    trans_kpair: np.ndarray          # (nk(nk+1)/2,)
    irre_pos: np.ndarray             # (nk(nk+1)/2,) position in stored list, -1 if not stored
    num_kpair_stored: int = field(default=0)


def mesh_maps(mesh_scaled):
    """k1 = k2 + q and k2 = k1 - q on a Monkhorst-Pack mesh given in fractional
    coordinates (green-symmetry k_q_map, used at gw_gpu_kernel.cpp:54-61)."""
    m = np.asarray(mesh_scaled, dtype=np.float64)
    nk = m.shape[0]

    def find(v):
        d = m - v[None, :]
        d -= np.round(d)
        idx = np.nonzero(np.all(np.abs(d) < 1e-6, axis=1))[0]
        assert idx.size == 1
        return int(idx[0])

    k1 = np.empty((nk, nk), dtype=np.int64)
    k2 = np.empty((nk, nk), dtype=np.int64)
    for a in range(nk):
        for q in range(nk):
            k1[a, q] = find(m[a] + m[q])
            k2[a, q] = find(m[a] - m[q])
    return k1, k2


def symmetry_from_fixture(fx, NQ):
    """Build the tables from the reference's ``input.h5`` symmetry group."""
    nk = int(fx["params_nk"])
    ibz2bz = fx["sym_k_ibz2bz"]
    bz2ibz = fx["sym_k_bz2ibz"]
    pos = {int(v): i for i, v in enumerate(ibz2bz)}
    k_f2r = np.array([pos[int(v)] for v in bz2ibz], dtype=np.int64)
    qibz2bz = fx["sym_q_ibz2bz"]
    qpos = {int(v): i for i, v in enumerate(qibz2bz)}
    q_f2r = np.array([qpos[int(v)] for v in fx["sym_q_bz2ibz"]], dtype=np.int64)
    k_stars = [np.array(fx[f"sym_k_star_{i}"], dtype=np.int64) for i in range(int(fx["sym_k_n_stars"]))]
    q_stars = [np.array(fx[f"sym_q_star_{i}"], dtype=np.int64) for i in range(int(fx["sym_q_n_stars"]))]
    k1, k2 = mesh_maps(fx["sym_k_mesh_scaled"])
    npair = nk * (nk + 1) // 2
    irre = np.full(npair, -1, dtype=np.int64)
    for i, p in enumerate(fx["pairs_kpair_irre_list"]):
        irre[int(p)] = i
    return Symmetry(
        nk=nk, ink=len(ibz2bz), nq=nk, inq=len(qibz2bz),
        k_full_to_reduced=k_f2r, k_reduced_to_full=np.array(ibz2bz, dtype=np.int64),
        k_tr_conj=np.array(fx["sym_k_tr_conj"], dtype=np.int64), k_stars=k_stars,
        q_full_to_reduced=q_f2r, q_reduced_to_full=np.array(qibz2bz, dtype=np.int64),
        q_tr_conj=np.array(fx["sym_q_tr_conj"], dtype=np.int64), q_stars=q_stars,
        k1_from_k2q=k1, k2_from_k1q=k2,
        U_k=np.array(fx["sym_k_transform_ao"]), U_q=np.array(fx["sym_q_transform_p0"])[:, :NQ, :NQ],
        conj_kpair=np.array(fx["pairs_conj_pairs_list"], dtype=np.int64),
        trans_kpair=np.array(fx["pairs_trans_pairs_list"], dtype=np.int64),
        irre_pos=irre, num_kpair_stored=int(fx["pairs_num_kpair_stored"]))


def symmetry_none(mesh, nao, NQ):
    """Tables for a mesh with no symmetry reduction: every k is its own star,
    U = I, no time-reversal pairs, every k1 >= k2 pair stored in pair order."""
    n1, n2, n3 = mesh
    nk = n1 * n2 * n3
    g = np.stack(np.meshgrid(np.arange(n1) / n1, np.arange(n2) / n2, np.arange(n3) / n3,
                             indexing="ij"), axis=-1).reshape(nk, 3)
    k1, k2 = mesh_maps(g)
    npair = nk * (nk + 1) // 2
    ar = np.arange(nk, dtype=np.int64)
    return Symmetry(
        nk=nk, ink=nk, nq=nk, inq=nk,
        k_full_to_reduced=ar.copy(), k_reduced_to_full=ar.copy(),
        k_tr_conj=np.zeros(nk, np.int64), k_stars=[np.array([k]) for k in range(nk)],
        q_full_to_reduced=ar.copy(), q_reduced_to_full=ar.copy(),
        q_tr_conj=np.zeros(nk, np.int64), q_stars=[np.array([q]) for q in range(nk)],
        k1_from_k2q=k1, k2_from_k1q=k2,
        U_k=np.broadcast_to(np.eye(nao, dtype=np.complex128), (nk, nao, nao)).copy(),
        U_q=np.broadcast_to(np.eye(NQ, dtype=np.complex128), (nk, NQ, NQ)).copy(),
        conj_kpair=np.arange(npair, dtype=np.int64), trans_kpair=np.arange(npair, dtype=np.int64),
        irre_pos=np.arange(npair, dtype=np.int64), num_kpair_stored=npair)


# ----------------------------------------------------------------------------
# Integral pair unfolding: df_integral_t.h:68-86 (index), :146-166 (v_type),
# :175-235 (symmetrize), :237-247 (wrap)
# ----------------------------------------------------------------------------
def pair_index(k1, k2):
    return k1 * (k1 + 1) // 2 + k2 if k1 >= k2 else k2 * (k2 + 1) // 2 + k1


def v_pair(V_stored, sym, k1, k2):
    """V^Q(k1,k2)[i][j], (NQ, nao, nao), unfolded as df_integral_t::symmetrize."""
    idx = pair_index(k1, k2)
    sign = 1 if k1 >= k2 else -1
    if sym.conj_kpair[idx] != idx:
        kind, idx = "conj", int(sym.conj_kpair[idx])
    elif sym.trans_kpair[idx] != idx:
        kind, idx = "trans", int(sym.trans_kpair[idx])
    else:
        kind = "direct"
    pos = int(sym.irre_pos[idx])
    assert pos >= 0, "pair not stored"
    M = V_stored[pos]
    if sign < 0:                                   # per-Q conjugate transpose
        M = np.conj(np.swapaxes(M, -1, -2))
    if kind == "conj":
        M = np.conj(M)
    elif kind == "trans":
        M = np.swapaxes(M, -1, -2)
    return np.ascontiguousarray(M)


# ----------------------------------------------------------------------------
# k-point symmetry rotation: cu_symmetry.cu:182-233 (value_AO semantics)
# ----------------------------------------------------------------------------
def value_ao(X_ibz, sym, k_full):
    """U_k X(ibz(k)) U_k^dagger, complex-conjugated for a time-reversal image.
    ``X_ibz`` is indexed (..., ink, n, n)."""
    U = sym.U_k[k_full]
    Y = U @ X_ibz[..., sym.k_full_to_reduced[k_full], :, :] @ U.conj().T
    return np.conj(Y) if sym.k_tr_conj[k_full] else Y


def ibz_to_full(X_ibz, sym):
    """(..., ink, n, n) -> (..., nk, n, n) with value_ao for every full k."""
    return np.stack([value_ao(X_ibz, sym, k) for k in range(sym.nk)], axis=-3)


# ----------------------------------------------------------------------------
# Hartree-Fock: hf_gpu_kernel.cpp:66-209, cuhf_utils.cu:90-159
# ----------------------------------------------------------------------------
def hf_exchange(dm, V_stored, sym, ns):
    """K_s(k) = (px/nk) sum_{k2,Q} V^Q(k,k2) D_s(k2) V^Q(k,k2)^dagger for IBZ k,
    px = -0.5 (ns=1) or -1 (ns=2)  (cuhf_utils.cu:108-134, :136-159)."""
    if ns not in (1, 2):
        raise ValueError('Invalid value of "ns"')
    dm_fbz = ibz_to_full(dm, sym)                    # (ns, nk, nao, nao)
    px = -0.5 if ns == 1 else -1.0
    nao = dm.shape[-1]
    K = np.zeros((ns, sym.ink, nao, nao), dtype=np.complex128)
    for ik in range(sym.ink):
        k = int(sym.k_reduced_to_full[ik])
        for k2 in range(sym.nk):
            V = v_pair(V_stored, sym, k, k2)             # (NQ, nao, nao)
            VH = np.conj(np.swapaxes(V, -1, -2))
            for s in range(ns):
                K[s, ik] += (px / sym.nk) * np.einsum("Qia,ab,Qbj->ij", V, dm_fbz[s, k2], VH, optimize=True)
    return K


def hf_direct(dm, V_stored, sym, ns):
    """J(k) = sum_Q V^Q(k,k) (1/nk) sum_{s,k'} Tr[V^Q(k',k') D_s(k')]
    (hf_gpu_kernel.cpp:141-186)."""
    nao = dm.shape[-1]
    upper = 0
    for kp in range(sym.nk):
        V = v_pair(V_stored, sym, kp, kp)
        for s in range(ns):
            D = value_ao(dm[s], sym, kp)
            upper = upper + np.einsum("Qab,ba->Q", V, D)
    upper = upper / sym.nk
    J = np.zeros((ns, sym.ink, nao, nao), dtype=np.complex128)
    for s in range(ns):
        for ik in range(sym.ink):
            k = int(sym.k_reduced_to_full[ik])
            J[s, ik] = np.einsum("Q,Qij->ij", upper, v_pair(V_stored, sym, k, k))
    return J


def hf_ewald(dm, S_k, madelung, ns):
    """-c * madelung * S D S, c = 1 (ns=2) or 0.5 (ns=1) (hf_gpu_kernel.cpp:188-201)."""
    c = 1.0 if ns == 2 else 0.5
    return -c * madelung * (S_k @ dm @ S_k)


def hf_solve(dm, V_stored, sym, ns, S_k, madelung):
    """Fock = J + K - Ewald on the IBZ (hf_gpu_kernel.cpp:66-104)."""
    return (hf_exchange(dm, V_stored, sym, ns) + hf_direct(dm, V_stored, sym, ns)
            + hf_ewald(dm, S_k, madelung, ns))


def hf_energy(dm, F, sym):
    """Two-electron HF energy (1/nk) sum_{s, k in IBZ} w_k Re Tr[D_s(k) F_s(k)] / 2
    with the star weights w_k = |star(k)|, i.e. the sum over the full
    Brillouin zone (Tr[D F] is invariant under the unitary k rotation and only
    conjugated under time reversal).  F = J + K - Ewald as hf_solve returns
    it; with ns = 1 the density is the spin-summed one (J uses it whole, the
    exchange takes half of it), so the same formula holds.  The reference
    returns only the Fock matrix (hf_gpu_kernel.cpp:66-104) and defines no
    energy; this definition is pinned by a hand-computed one-orbital system
    and by the reduced-vs-full-zone sums on the HF fixture."""
    w = np.array([len(s) for s in sym.k_stars], dtype=np.float64)
    return float(0.5 * np.einsum("k,skij,skji->", w, dm, F).real / sym.nk)


# ----------------------------------------------------------------------------
# GW: cugw_utils.cu:103-202 (q loop), cugw_qkpt.cu:202-338 (contractions),
# cugw_qpt.cu:174-358 (scale/mirror, IR transforms, Dyson)
# ----------------------------------------------------------------------------
def check_gw_dims(nts, nw_b):
    if nts % 2 != 0:
        raise ValueError("nts must be even")                   # gw_gpu_kernel.h:67
    if nts < nw_b:
        raise ValueError("Nt is not greater than or equal to Nw_b")   # cugw_qpt.cu:79


def g_full(G, sym):
    """G(tau) on the full BZ, (nk, ns, nts, nao, nao), from the IBZ store
    (cugw_utils.cu:204-238 with cu_symmetry.cu:182-233; the -tau copy of
    cuda_common.cu:57-63 is the reversed tau index)."""
    Gt = np.transpose(G, (1, 0, 2, 3, 4))            # (ns, nts, ink, a, b)
    return np.stack([value_ao(Gt, sym, k) for k in range(sym.nk)], axis=0)


def p0_pair(Gm, G1, V, pref):
    """Contribution of one (k, k+q) pair to P0(tau)[Q'][Q] for the given
    tau slices: pref * Tr[G(k,beta-tau) V^{Q'} G(k+q,tau) V^{Q dagger}].
    ``Gm``, ``G1``: (nt, nao, nao).  Follows the three strided-batched GEMMs
    of cugw_qkpt.cu:202-234."""
    X1 = np.matmul(Gm[:, None], V[None])                      # (t, Q', n, m) = Gm V^Q'
    X2 = np.matmul(G1[:, None], np.conj(np.swapaxes(V, -1, -2))[None])   # (t, Q, m, n)
    nt, NQ, nao = X1.shape[0], X1.shape[1], X1.shape[2]
    A = X1.reshape(nt, NQ, nao * nao)
    B = np.swapaxes(X2, -1, -2).reshape(nt, NQ, nao * nao)   # [Q][n][m]
    return pref * np.matmul(A, np.swapaxes(B, -1, -2))


# tau points per contraction chunk: every tau is an independent product, so
# chunking changes no result and bounds the intermediates' memory at large
# nao / NQ
TAU_CHUNK = 16


def gw_p0(Gf, V_stored, sym, q_ir, ns, nts, Gm_src=None):
    """P0(q, tau) for all nts points: k-sum over every full-BZ k (k-stars of the
    IBZ), t < nts/2 computed, scaled by 1/nk and mirrored
    (cugw_utils.cu:121-150, cugw_qkpt.cu:202-246, cugw_qpt.cu:174-191).
    ``Gm_src`` supplies the G(k, beta - tau) factor (default ``Gf``); the
    two-component path passes the spin-block-swapped copy."""
    pref = -2.0 if ns == 1 else -1.0
    Gm_src = Gf if Gm_src is None else Gm_src
    half = nts // 2
    NQ = None
    P0 = None
    for ik in range(sym.ink):
        for k in sym.k_stars[ik]:
            k = int(k)
            k1 = int(sym.k1_from_k2q[k, q_ir])
            V = v_pair(V_stored, sym, k, k1)
            if P0 is None:
                NQ = V.shape[0]
                P0 = np.zeros((nts, NQ, NQ), dtype=np.complex128)
            for s in range(ns):
                for c0 in range(0, half, TAU_CHUNK):
                    t = np.arange(c0, min(half, c0 + TAU_CHUNK))
                    P0[t] += p0_pair(Gm_src[k, s, nts - 1 - t], Gf[k1, s, t], V, pref)
    P0[:half] *= 1.0 / sym.nk
    P0[nts - 1 - np.arange(half)] = P0[:half]
    return P0


def herm(M):
    """(M + M^dagger)/2 with an exactly real diagonal (cuda_common.cu:249-265)."""
    return 0.5 * (M + np.conj(np.swapaxes(M, -1, -2)))


def tau_to_omega(P0, T_wt):
    """P0(Omega) = sum_{t=1}^{nts-2} T_wt[Omega][t-1] P0(tau_t)  (cugw_qpt.cu:329-339)."""
    nt, NQ = P0.shape[0], P0.shape[1]
    return (T_wt @ P0[1:-1].reshape(nt - 2, NQ * NQ)).reshape(-1, NQ, NQ)


def omega_to_tau(Pw, T_tw):
    """P(tau) = sum_Omega T_tw[tau][Omega] P(Omega) over all tau (cugw_qpt.cu:341-358)."""
    nw, NQ = Pw.shape[0], Pw.shape[1]
    return (T_tw @ Pw.reshape(nw, NQ * NQ)).reshape(-1, NQ, NQ)


def gw_dyson(P0w, solver="LU"):
    """P(Omega) = (I - herm P0)^-1 herm P0 for every Omega; the LU variant
    re-hermitises the solution (cugw_qpt.cu:223-322, cu_compute_Pq.cu:44-63)."""
    P0h = herm(P0w)
    NQ = P0h.shape[-1]
    out = np.empty_like(P0h)
    eye = np.eye(NQ)
    for w in range(P0h.shape[0]):
        A = eye - P0h[w]
        if solver.lower() == "cholesky":
            c = sla.cho_factor(A, lower=True)
            out[w] = sla.cho_solve(c, P0h[w])
        else:
            lu = sla.lu_factor(A)
            out[w] = sla.lu_solve(lu, P0h[w])
    if solver.lower() != "cholesky":
        out = herm(out)
    return out


def q_rotate(Pt, sym, q_deg):
    """P on a q in the star of the IBZ q: U_q P U_q^dagger, complex-conjugated
    for a time-reversal image (cugw_utils.cu:240-269, cugw_qkpt.cu:248-338)."""
    U = sym.U_q[q_deg]
    Pr = U @ Pt @ U.conj().T
    return np.conj(Pr) if sym.q_tr_conj[q_deg] else Pr


def sigma_pair(Pt, G1, V):
    """-sum_{Q,Q'} P[t][Q'][Q] (V^{Q'} G(k-q,t) V^{Q dagger}) for all t:
    (nt, nao, nao).  Follows the GEMM chain of cugw_qkpt.cu:288-333."""
    nt, NQ, nao = G1.shape[0], V.shape[0], V.shape[1]
    Y1 = np.matmul(V[None], G1[:, None])                      # (t, Q', i, n)
    Y2 = np.matmul(np.swapaxes(Pt, -1, -2), Y1.reshape(nt, NQ, nao * nao))   # (t, Q, (i, n))
    Y2 = np.swapaxes(Y2.reshape(nt, NQ, nao, nao), 1, 2)      # (t, i, Q, m)
    VH = np.conj(np.swapaxes(V, -1, -2))                      # V^{Q dagger}[m][j]
    return -np.matmul(Y2.reshape(nt, nao, NQ * nao), VH.reshape(NQ * nao, nao))


def _gw_core(Gf, Gm_src, ns, nts, sym, V_stored, T_wt, T_tw, solver, return_parts):
    """Sigma(tau) in the (nts, ns, ink, nao, nao) layout for full-BZ G blocks
    ``Gf`` (nk, ns, nts, nao, nao) (cugw_utils.cu:103-202)."""
    nao = Gf.shape[-1]
    Sigma = np.zeros((nts, ns, sym.ink, nao, nao), dtype=np.complex128)
    parts = {"P0": [], "P0w": [], "Pw": [], "Pt": []}
    for iq in range(sym.inq):
        q_ir = int(sym.q_reduced_to_full[iq])
        P0 = gw_p0(Gf, V_stored, sym, q_ir, ns, nts, Gm_src)
        P0w = tau_to_omega(P0, T_wt)
        Pw = gw_dyson(P0w, solver)
        Pt = omega_to_tau(Pw, T_tw)
        if return_parts:
            parts["P0"].append(P0)
            parts["P0w"].append(P0w)
            parts["Pw"].append(Pw)
            parts["Pt"].append(Pt)
        for ik in range(sym.ink):
            k = int(sym.k_reduced_to_full[ik])
            for q_deg in sym.q_stars[iq]:
                q_deg = int(q_deg)
                k1 = int(sym.k2_from_k1q[k, q_deg])
                V = v_pair(V_stored, sym, k, k1)
                Pq = q_rotate(Pt, sym, q_deg)
                for s in range(ns):
                    for c0 in range(0, nts, TAU_CHUNK):
                        t = slice(c0, min(nts, c0 + TAU_CHUNK))
                        Sigma[t, s, ik] += sigma_pair(Pq[t], Gf[k1, s, t], V)
    Sigma /= sym.nk
    return Sigma, parts


def _check_transforms(nts, T_wt, T_tw):
    nw_b = T_wt.shape[0]
    check_gw_dims(nts, nw_b)
    assert T_wt.shape == (nw_b, nts - 2) and T_tw.shape == (nts, nw_b)


def gw_solve(G, V_stored, sym, T_wt, T_tw, solver="LU", return_parts=False):
    """One GW iteration: Sigma(tau) (nts, ns, ink, nao, nao) from G(tau) of the
    same shape (gw_gpu_kernel.cpp:163-205, cugw_utils.cu:103-202)."""
    nts, ns = G.shape[0], G.shape[1]
    _check_transforms(nts, T_wt, T_tw)
    Gf = g_full(G, sym)
    Sigma, parts = _gw_core(Gf, Gf, ns, nts, sym, V_stored, T_wt, T_tw, solver, return_parts)
    return (Sigma, parts) if return_parts else Sigma


# ----------------------------------------------------------------------------
# Two-component (X2C) variants: gw_gpu_kernel.cpp:458-551, cugw_qkpt.cu:381-403,
# hf_gpu_kernel.cpp:229-338, cuda_common.cu:176-205
# ----------------------------------------------------------------------------
# Pseudo-spin order of the spinor blocks (row block, column block): aa, bb, ab, ba
X2C_BLOCKS = ((0, 0), (1, 1), (0, 1), (1, 0))


def _x2c_check(sym):
    # the rotation of a spinor matrix onto a time-reversed k-point (value_AO's
    # spin flip) lives in green-symmetry and is not restated here
    if np.any(np.asarray(sym.k_tr_conj) != 0):
        raise ValueError("x2c: time-reversal images in the k-point tables are not supported")


def spinor_blocks(X, nao, blocks):
    """(..., 2 nao, 2 nao) -> (..., len(blocks), nao, nao) on a new axis -3."""
    return np.stack([X[..., a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] for a, b in blocks], axis=-3)


def gw_solve_x2c(G, V_stored, sym, T_wt, T_tw, solver="LU", return_parts=False):
    """Two-component GW: G, Sigma (nts, 1, ink, nso, nso), nso = 2 nao.  The
    scalar contractions run on pseudo-spins (aa, bb, ab, ba) with prefactor -1;
    G(k, beta - tau) enters with its spin blocks swapped (the ab pseudo-spin
    pairs G_ba(beta - tau) with G_ab(tau)), and every Sigma block is
    accumulated from its own pseudo-spin (gw_gpu_kernel.cpp:525-551,
    cugw_qkpt.cu:381-403)."""
    _x2c_check(sym)
    nts, nso = G.shape[0], G.shape[-1]
    assert G.shape[1] == 1 and nso % 2 == 0
    nao = nso // 2
    _check_transforms(nts, T_wt, T_tw)
    Gf2 = g_full(G, sym)[:, 0]                          # (nk, nts, nso, nso)
    G4 = np.ascontiguousarray(np.swapaxes(spinor_blocks(Gf2, nao, X2C_BLOCKS), 1, 2))   # (nk, 4, nts, n, n)
    Gm4 = np.ascontiguousarray(G4[:, [0, 1, 3, 2]])
    S4, parts = _gw_core(G4, Gm4, 4, nts, sym, V_stored, T_wt, T_tw, solver, return_parts)
    Sigma = np.zeros_like(G)
    for ss, (a, b) in enumerate(X2C_BLOCKS):
        Sigma[:, 0, :, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] = S4[:, ss]
    return (Sigma, parts) if return_parts else Sigma


def _exchange_fbz(dm_fbz, V_stored, sym, px):
    """K[s, ik] = (px/nk) sum_{k2,Q} V^Q(k,k2) D_s(k2) V^Q(k,k2)^dagger from
    full-BZ density blocks (ns', nk, n, n)."""
    ns, n = dm_fbz.shape[0], dm_fbz.shape[-1]
    K = np.zeros((ns, sym.ink, n, n), dtype=np.complex128)
    for ik in range(sym.ink):
        k = int(sym.k_reduced_to_full[ik])
        for k2 in range(sym.nk):
            V = v_pair(V_stored, sym, k, k2)
            VH = np.conj(np.swapaxes(V, -1, -2))
            for s in range(ns):
                K[s, ik] += (px / sym.nk) * np.einsum("Qia,ab,Qbj->ij", V, dm_fbz[s, k2], VH, optimize=True)
    return K


def hf_exchange_x2c(dm, V_stored, sym):
    """Two-component exchange (1, ink, nso, nso): the aa, bb and ab blocks are
    the scalar exchange of the matching density blocks with prefactor -1
    (pseudo-spin 3), and K_ba = K_ab^dagger (hf_gpu_kernel.cpp:229-249,
    cuda_common.cu:176-205)."""
    _x2c_check(sym)
    nso = dm.shape[-1]
    nao = nso // 2
    dm3 = spinor_blocks(ibz_to_full(dm[0], sym), nao, X2C_BLOCKS[:3])   # (nk, 3, n, n)
    K3 = _exchange_fbz(np.swapaxes(dm3, 0, 1), V_stored, sym, -1.0)     # (3, ink, n, n)
    K = np.zeros((1, sym.ink, nso, nso), dtype=np.complex128)
    for ss, (a, b) in enumerate(X2C_BLOCKS[:3]):
        K[0, :, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] = K3[ss]
    K[0, :, nao:, :nao] = np.conj(np.swapaxes(K3[2], -1, -2))
    return K


def hf_solve_x2c(dm, V_stored, sym, S_k, madelung):
    """Two-component Fock = J + K - Ewald (hf_gpu_kernel.cpp:253-327):
    J from the aa + bb density on both diagonal blocks; the Ewald term uses
    the AO overlap S_aa on both sides of every density block (aa, bb, ab;
    ba as the adjoint of ab)."""
    _x2c_check(sym)
    nso = dm.shape[-1]
    nao = nso // 2
    F = hf_exchange_x2c(dm, V_stored, sym)
    upper = 0
    for kp in range(sym.nk):
        V = v_pair(V_stored, sym, kp, kp)
        D = value_ao(dm[0], sym, kp)
        upper = upper + np.einsum("Qab,ba->Q", V, D[:nao, :nao] + D[nao:, nao:])
    upper = upper / sym.nk
    for ik in range(sym.ink):
        k = int(sym.k_reduced_to_full[ik])
        J = np.einsum("Q,Qij->ij", upper, v_pair(V_stored, sym, k, k))
        F[0, ik, :nao, :nao] += J
        F[0, ik, nao:, nao:] += J
        S_aa = S_k[0, ik, :nao, :nao]
        for a, b in X2C_BLOCKS[:3]:
            buf = madelung * S_aa @ dm[0, ik, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] @ S_aa
            F[0, ik, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] -= buf
            if a != b:
                F[0, ik, nao:, :nao] -= np.conj(buf.T)
    return F


# ----------------------------------------------------------------------------
# Algorithmic FLOP model (complex MAC = 8 FLOP, minimal contraction set, U = I)
# ----------------------------------------------------------------------------
def gw_flops(nk, nao, NQ, ns, nts, nw_b, nq=None):
    nq = nk if nq is None else nq
    c = 16.0 * NQ * nao ** 3 + 8.0 * NQ * NQ * nao * nao
    p0 = nq * nk * ns * (nts // 2) * c
    sig = nq * nk * ns * nts * c
    ir = nq * 8.0 * NQ * NQ * (nw_b * (nts - 2) + nts * nw_b)
    dyson = nq * nw_b * (4.0 / 3.0 + 8.0) * NQ ** 3
    return {"p0": p0, "sigma": sig, "ir": ir, "dyson": dyson, "total": p0 + sig + ir + dyson}


def hf_flops(ink, nk, nao, NQ, ns):
    return ink * nk * ns * 16.0 * NQ * nao ** 3
