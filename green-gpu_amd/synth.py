"""Synthetic inputs of the benchmark workloads.

* H_s(k) = W diag(eps) W^dagger with W from the QR of a complex Gaussian
  matrix and eps ~ U[-2, 2] Ha, chemical potential 0;
* G_s(k, tau) = -W diag(f(eps, tau)) W^dagger, f = e^{-eps tau} / (1 + e^{-beta eps})
  (evaluated in the overflow-free form for eps < 0);
* D = -G(beta), the Fermi occupation matrix;
* V^Q_ij(k1, k2), stored for k1 >= k2 in pair order k1(k1+1)/2 + k2, with
  Re, Im ~ N(0, sigma^2) from a counter-based generator keyed by the element's
  global index (the device kernel synth_integrals_kernel uses the same
  formulas), and the diagonal pairs Hermitised per Q.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_ir(beta, lam):
    """(tau, T_wt, T_tw) of the committed transforms (data/make_ir_data.py)."""
    d = np.load(os.path.join(HERE, "data", f"ir_beta{int(beta)}_lambda{int(lam)}.npz"))
    return d["tau"], d["T_wt"], d["T_tw"]


def fermi_weight(eps, tau, beta):
    eps = np.asarray(eps, dtype=np.float64)
    tau = np.asarray(tau, dtype=np.float64)
    e = eps[..., None, :]
    t = tau[:, None]
    pos = e >= 0
    ep = np.where(pos, e, 0.0)
    en = np.where(pos, 0.0, e)
    fp = np.exp(-ep * t) / (1.0 + np.exp(-beta * ep))
    fn = np.exp(en * (beta - t)) / (1.0 + np.exp(beta * en))
    return np.where(pos, fp, fn)          # (..., ntau, nao)


def make_green(nk, nao, ns, beta, tau, seed):
    """G (nts, ns, nk, nao, nao) and D = -G(beta) (ns, nk, nao, nao)."""
    rng = np.random.default_rng(seed)
    Z = rng.normal(size=(ns, nk, nao, nao)) + 1j * rng.normal(size=(ns, nk, nao, nao))
    W, _ = np.linalg.qr(Z)
    eps = rng.uniform(-2.0, 2.0, size=(ns, nk, nao))
    f = fermi_weight(eps, tau, beta)                       # (ns, nk, nts, nao)
    # G[t,s,k][a][b] = -sum_m f[s,k,t][m] W[a][m] conj(W[b][m]): for each (s, k)
    # one real GEMM of the weights (nts x nao) with the outer products
    # P[m][(a, b)] = W[a][m] conj(W[b][m]) viewed as real (nao x 2 nao^2)
    nts = f.shape[2]
    G = np.empty((nts, ns, nk, nao, nao), dtype=np.complex128)
    for s in range(ns):
        for k in range(nk):
            Wk = W[s, k]
            P = np.ascontiguousarray((Wk.T[:, :, None] * np.conj(Wk.T)[:, None, :]).reshape(nao, nao * nao))
            Gk = np.dot(f[s, k], P.view(np.float64)).view(np.complex128)    # (nts, nao^2)
            G[:, s, k] = -Gk.reshape(nts, nao, nao)
    D = -G[-1]
    return np.ascontiguousarray(G), np.ascontiguousarray(D)


_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)
_MS = np.uint64(0xD1B54A32D192ED03)


def _splitmix64(z):
    z = z + _M1
    z = (z ^ (z >> np.uint64(30))) * _M2
    z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def gauss_c(seed, e, sigma):
    with np.errstate(over="ignore"):
        h1 = _splitmix64(np.uint64(seed) * _MS + np.asarray(e, dtype=np.uint64))
        h2 = _splitmix64(h1)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rad = sigma * np.sqrt(-2.0 * np.log(u1))
    ang = 6.283185307179586 * u2
    return rad * np.cos(ang) + 1j * rad * np.sin(ang)


def pair_of(p):
    """(k1, k2) of pair index p = k1(k1+1)/2 + k2, k1 >= k2."""
    k1 = int((np.sqrt(8.0 * p + 1.0) - 1.0) // 2)
    while (k1 + 1) * (k1 + 2) // 2 <= p:
        k1 += 1
    while k1 * (k1 + 1) // 2 > p:
        k1 -= 1
    return k1, p - k1 * (k1 + 1) // 2


def make_integrals(nk, NQ, nao, seed, sigma=0.1, pairs=None):
    """Stored pairs (npairs, NQ, nao, nao) for identity pair storage; ``pairs``
    selects a subset of stored positions (default: all nk(nk+1)/2)."""
    if pairs is None:
        pairs = range(nk * (nk + 1) // 2)
    pairs = list(pairs)
    out = np.empty((len(pairs), NQ, nao, nao), dtype=np.complex128)
    ij = np.arange(nao * nao, dtype=np.int64)
    ji = (ij % nao) * nao + ij // nao
    for n, p in enumerate(pairs):
        k1, k2 = pair_of(p)
        base = (p * NQ + np.arange(NQ, dtype=np.int64))[:, None] * (nao * nao)
        v = gauss_c(seed, (base + ij[None, :]).astype(np.uint64), sigma)
        if k1 == k2:
            t = gauss_c(seed, (base + ji[None, :]).astype(np.uint64), sigma)
            v = 0.5 * (v + np.conj(t))
        out[n] = v.reshape(NQ, nao, nao)
    return out
