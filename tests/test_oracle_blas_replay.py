"""Pin the oracle's closed forms against the reference's own call sequences.

The reference computes every contraction through cuBLAS/cuSOLVER on
row-major buffers, so its mathematics is fixed by the operation flags,
leading dimensions and strides at each call site.  This file replays those
calls literally - a column-major BLAS emulator applied to flat buffers - and
checks the oracle's row-major closed forms against the results.  A wrong
index order, a missing conjugation or a transposed operand in the oracle
fails here even though the oracle's own formulas are not re-typed.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import max_abs


# ---------------------------------------------------------------------------
# Column-major BLAS/LAPACK semantics on flat complex buffers
# ---------------------------------------------------------------------------
def _mat(buf, off, rows, cols, ld):
    """Column-major rows x cols matrix starting at buf[off] with leading dim ld."""
    idx = off + np.arange(rows)[:, None] + ld * np.arange(cols)[None, :]
    return buf[idx]


def _op(buf, off, op, rows_op, cols_op, ld):
    if op == "N":
        return _mat(buf, off, rows_op, cols_op, ld)
    M = _mat(buf, off, cols_op, rows_op, ld).T
    return np.conj(M) if op == "C" else M


def gemm(opA, opB, m, n, k, alpha, A, oA, lda, B, oB, ldb, beta, C, oC, ldc):
    a = _op(A, oA, opA, m, k, lda)
    b = _op(B, oB, opB, k, n, ldb)
    idx = oC + np.arange(m)[:, None] + ldc * np.arange(n)[None, :]
    C[idx] = alpha * (a @ b) + (beta * C[idx] if beta != 0 else 0)


def gemm_sb(opA, opB, m, n, k, alpha, A, oA, lda, sA, B, oB, ldb, sB, beta, C, oC, ldc, sC, batch):
    for i in range(batch):
        gemm(opA, opB, m, n, k, alpha, A, oA + i * sA, lda, B, oB + i * sB, ldb, beta, C, oC + i * sC, ldc)


def geam(opA, m, n, A, lda, C, ldc):
    """C = op(A) (alpha = 1, beta = 0)."""
    idx = np.arange(m)[:, None] + ldc * np.arange(n)[None, :]
    C[idx] = _op(A, 0, opA, m, n, lda)


def conj_inplace(buf):
    """RSCAL(alpha = -1) on the imaginary lane (stride 2 over the doubles)."""
    buf[:] = np.conj(buf)


def rand_c(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ns", [1, 2])
def test_p0_call_sequence(oracle, ns):
    """cugw_qkpt.cu:115-130 (upload_coulomb_v_first) and :202-246
    (compute_first_tau_contraction, write_P0) for one (k, k+q) pair."""
    rng = np.random.default_rng(1)
    nao, naux, nt = 3, 5, 6
    nao2, nauxnao, nauxnao2, naux2 = nao * nao, naux * nao, naux * nao * nao, naux * naux
    V = rand_c(rng, naux, nao, nao)
    Gk = rand_c(rng, ns, nt, nao, nao)       # G(k, tau) of the k point
    Gk1 = rand_c(rng, ns, nt, nao, nao)      # G(k+q, tau)
    pref = -2.0 if ns == 1 else -1.0

    V_Qpm = V.ravel().copy()
    V_pmQ = np.zeros(nauxnao2, complex)
    geam("T", naux, nao2, V_Qpm, nao2, V_pmQ, naux)
    conj_inplace(V_Qpm)
    g_smtij = Gk[:, ::-1].ravel().copy()     # G(k, -tau): reversed tau index (cuda_common.cu:57-63)
    g_stij = Gk1.ravel().copy()
    P0 = np.zeros(nt * naux2, complex)
    X1 = np.zeros(nauxnao2, complex)
    X2 = np.zeros(nauxnao2, complex)
    Ploc = np.zeros(naux2, complex)
    for s in range(ns):
        for t in range(nt // 2):
            st = s * nt + t
            gemm_sb("N", "N", nao * naux, nao, nao, 1, V_pmQ, 0, nauxnao, 0, g_smtij, st * nao2, nao, nao2,
                    0, X1, 0, nauxnao, nauxnao2, 1)
            gemm_sb("T", "N", nao, nauxnao, nao, 1, g_stij, st * nao2, nao, nao2, V_Qpm, 0, nao, 0,
                    0, X2, 0, nao, nauxnao2, 1)
            gemm_sb("T", "T", naux, naux, nao2, pref, X2, 0, nao2, nauxnao2, X1, 0, naux, nauxnao2,
                    0, Ploc, 0, naux, naux2, 1)
            P0[t * naux2:(t + 1) * naux2] += Ploc
    P0 = P0.reshape(nt, naux, naux)
    half = nt // 2
    ref = 0
    for s in range(ns):
        ref = ref + oracle.p0_pair(Gk[s, nt - 1 - np.arange(half)], Gk1[s, :half], V, pref)
    assert max_abs(P0[:half], ref) < 1e-12


@pytest.mark.parametrize("tr", [0, 1])
def test_sigma_call_sequence(oracle, tr):
    """cugw_qkpt.cu:162-173 (upload_coulomb_v_second) and :248-338
    (compute_second_tau_contraction with U), cugw_utils.cu:240-269 operand
    choice for non-TR / TR q."""
    rng = np.random.default_rng(2)
    nao, naux, nt = 3, 4, 5
    nao2, nauxnao, nauxnao2, naux2 = nao * nao, naux * nao, naux * nao * nao, naux * naux
    V = rand_c(rng, naux, nao, nao)
    G1 = rand_c(rng, nt, nao, nao)
    P = rand_c(rng, nt, naux, naux)
    Uq = np.linalg.qr(rand_c(rng, naux, naux))[0]

    V_Qim = V.ravel().copy()
    V_nPj = np.zeros(nauxnao2, complex)
    geam("C", nauxnao, nao, V_Qim, nao, V_nPj, nauxnao)
    U_store, Uc_store = Uq.ravel().copy(), np.conj(Uq).ravel().copy()
    U, U_conj = (Uc_store, U_store) if tr else (U_store, Uc_store)
    Pq = (np.conj(P) if tr else P).ravel().copy()
    g = G1.ravel().copy()
    Y1 = np.zeros(nauxnao2, complex)
    Y2 = np.zeros(nauxnao2, complex)
    sig = np.zeros(nt * nao2, complex)
    for t in range(nt):
        gemm_sb("N", "N", nao, nauxnao, nao, 1, g, t * nao2, nao, nao2, V_Qim, 0, nao, 0, 0, Y1, 0, nao,
                nauxnao2, 1)
        gemm_sb("N", "T", naux, nao2, naux, 1, U, 0, naux, 0, Y1, 0, nao2, nauxnao2, 0, Y2, 0, naux, nauxnao2, 1)
        gemm_sb("N", "N", naux, nao2, naux, 1, Pq, t * naux2, naux, naux2, Y2, 0, naux, nauxnao2, 0, Y1, 0,
                naux, nauxnao2, 1)
        gemm_sb("T", "N", naux, nao2, naux, 1, U_conj, 0, naux, 0, Y1, 0, naux, nauxnao2, 0, Y2, 0, naux,
                nauxnao2, 1)
        gemm_sb("N", "N", nao, nao, nauxnao, -1, V_nPj, 0, nao, 0, Y2, 0, nauxnao, nauxnao2, 0, sig, t * nao2,
                nao, nao2, 1)
    sig = sig.reshape(nt, nao, nao)

    class S:
        U_q = Uq[None]
        q_tr_conj = np.array([tr])
    ref = oracle.sigma_pair(oracle.q_rotate(P, S, 0), G1, V)
    assert max_abs(sig, ref) < 1e-12


def test_hf_call_sequence(oracle):
    """cuhf_utils.cu:90-134: conj-transposed V, X = V D, (Q,ia) -> (ia,Q),
    Y = X V^dagger, F += pref * w * Y; against the oracle on a 1-k mesh."""
    rng = np.random.default_rng(3)
    nao, NQ = 4, 6
    nao2, NQnao2 = nao * nao, NQ * nao * nao
    Vm = rand_c(rng, NQ, nao, nao)
    D = rand_c(rng, nao, nao)
    V = Vm.ravel().copy()
    VaQj = np.zeros(NQnao2, complex)
    geam("C", NQ * nao, nao, V, nao, VaQj, NQ * nao)
    X = np.zeros(NQnao2, complex)
    gemm_sb("N", "N", nao, NQ * nao, nao, 1, D.ravel().copy(), 0, nao, nao2, V, 0, nao, NQnao2, 0, X, 0, nao,
            NQnao2, 1)
    XiaQ = np.zeros(NQnao2, complex)
    geam("T", NQ, nao2, X, nao2, XiaQ, NQ)
    Y = np.zeros(nao2, complex)
    gemm_sb("N", "N", nao, nao, NQ * nao, 1, VaQj, 0, nao, NQnao2, XiaQ, 0, NQ * nao, NQnao2, 0, Y, 0, nao, nao2, 1)
    for ns in (1, 2):
        px = -0.5 if ns == 1 else -1.0
        sym = oracle.symmetry_none((1, 1, 1), nao, NQ)
        K = oracle.hf_exchange(np.broadcast_to(D, (ns, 1, nao, nao)), Vm[None], sym, ns)
        assert max_abs(K[0, 0], px * Y.reshape(nao, nao)) < 1e-12


@pytest.mark.parametrize("tr", [0, 1])
def test_symmetry_rotation_call_sequence(oracle, tr):
    """cu_symmetry.cu:182-233 transform_k_ao_device: GEMM(C,N) then GEMM(N,N)
    with row-major U, then conjugation for a time-reversal image."""
    rng = np.random.default_rng(4)
    nao, nb = 3, 4
    U = np.linalg.qr(rand_c(rng, nao, nao))[0]
    Gi = rand_c(rng, nb, nao, nao)
    inp = Gi.ravel().copy()
    work = np.zeros_like(inp)
    out = np.zeros_like(inp)
    Uf = U.ravel().copy()
    gemm_sb("C", "N", nao, nao, nao, 1, Uf, 0, nao, 0, inp, 0, nao, nao * nao, 0, work, 0, nao, nao * nao, nb)
    gemm_sb("N", "N", nao, nao, nao, 1, work, 0, nao, nao * nao, Uf, 0, nao, 0, 0, out, 0, nao, nao * nao, nb)
    if tr:
        conj_inplace(out)

    class S:
        U_k = U[None]
        k_full_to_reduced = np.array([0])
        k_tr_conj = np.array([tr])
    ref = oracle.value_ao(Gi[:, None], S, 0)
    assert max_abs(out.reshape(nb, nao, nao), ref) < 1e-12


def test_ir_call_sequence(oracle):
    """cugw_qpt.cu:329-358: tau -> Omega over interior tau, Omega -> tau over all tau."""
    rng = np.random.default_rng(5)
    naux, nt, nw = 3, 8, 5
    naux2 = naux * naux
    P0 = rand_c(rng, nt, naux, naux)
    Twt = rand_c(rng, nw, nt - 2)
    Ttw = rand_c(rng, nt, nw)
    buf = P0.ravel().copy()
    Pw = np.zeros(nw * naux2, complex)
    gemm("N", "N", naux2, nw, nt - 2, 1, buf, naux2, naux2, Twt.ravel().copy(), 0, nt - 2, 0, Pw, 0, naux2)
    assert max_abs(Pw.reshape(nw, naux, naux), oracle.tau_to_omega(P0, Twt)) < 1e-12
    Pt = np.zeros(nt * naux2, complex)
    gemm("N", "N", naux2, nt, nw, 1, Pw, 0, naux2, Ttw.ravel().copy(), 0, nw, 0, Pt, 0, naux2)
    assert max_abs(Pt.reshape(nt, naux, naux), oracle.omega_to_tau(Pw.reshape(nw, naux, naux), Ttw)) < 1e-12


@pytest.mark.parametrize("solver", ["cholesky", "lu"])
def test_dyson_call_sequence(oracle, solver):
    """cugw_qpt.cu:223-322 + cu_compute_Pq.cu:44-63: hermitise, 1 - P0 written
    row-major, cuSOLVER potrf(LOWER)/potrs or getrf/getrs on the column-major
    view of the same buffers, LU result re-hermitised."""
    rng = np.random.default_rng(6)
    n, nw = 6, 3
    M = rand_c(rng, nw, n, n)
    M = -(M @ np.conj(np.swapaxes(M, 1, 2))) * 0.1 + 0.01 * rand_c(rng, nw, n, n)   # near-NSD, not Hermitian
    out = np.empty_like(M)
    for w in range(nw):
        rm = M[w].copy()
        h = np.empty_like(rm)
        for i in range(n):                              # hermitian_symmetrize, cuda_common.cu:249-265
            h[i, i] = rm[i, i].real
            for j in range(i):
                re = 0.5 * (rm[i, j].real + rm[j, i].real)
                im = 0.5 * (rm[i, j].imag - rm[j, i].imag)
                h[i, j] = re + 1j * im
                h[j, i] = re - 1j * im
        one_minus = np.eye(n) - h                       # set_up_one_minus_P (row-major)
        A_cm = one_minus.T                              # column-major view of the buffer
        B_cm = h.T
        if solver == "cholesky":
            L = np.linalg.cholesky(A_cm)
            X_cm = sla.cho_solve((L, True), B_cm)
        else:
            X_cm = sla.lu_solve(sla.lu_factor(A_cm), B_cm)
        X = X_cm.T                                      # back to the row-major buffer
        if solver == "lu":
            X = 0.5 * (X + np.conj(X.T))
        out[w] = X
    assert max_abs(out, oracle.gw_dyson(M, solver)) < 1e-12
