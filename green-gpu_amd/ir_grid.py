"""Sparse-IR-style imaginary-time / bosonic-Matsubara transforms.

green-mbpt hands the GW kernel two dense matrices from green-grids
(``transformer_t::Tnt_BF`` and ``Ttn_FB``, consumed at gw_gpu_kernel.cpp:303
and cugw_qpt.cu:329-358):

* ``T_wt`` (nw_b x (nts-2)): bosonic quantity on the interior fermionic tau
  sampling points -> bosonic Matsubara frequencies;
* ``T_tw`` (nts x nw_b): bosonic Matsubara frequencies -> every tau point,
  including tau = 0 and tau = beta.

The green-grids IR files are not available here, so this module builds an
intermediate-representation basis from first principles (Shinaoka et al.,
PRB 96, 035147 (2017); Chikano et al., CPC 240, 181 (2019)):

1. singular value expansion of the logistic kernel
   K(tau, omega) = exp(-tau omega) / (1 + exp(-beta omega)),
   discretised with composite Gauss-Legendre rules graded towards the
   kernel's boundary layers (tau -> 0, beta and omega -> 0);
2. basis functions u_l(tau) by Nystrom interpolation,
   u_l(tau) = s_l^-1 int K(tau, omega) v_l(omega) domega;
3. tau sampling points = the roots of the first discarded function u_L
   (L points, symmetric about beta/2), plus the two endpoints;
4. Matsubara transforms in closed form,
   int_0^beta e^{i Omega tau} K(tau, omega) dtau = tanh(beta omega / 2) / (omega - i Omega)
   for bosonic Omega;
5. T_wt = Uhat(Omega) U(tau_interior)^-1, T_tw = U(tau_all) Uhat(Omega)^+ .

These matrices are inputs of the GW kernel: parity only needs the same
matrices on both sides, and the construction only has to be a sound
tau <-> Omega transform (so that P0(i Omega) comes out Hermitian negative
semi-definite and the Dyson matrix 1 - P0 stays positive definite).
"""
import numpy as np


def _graded_gauss(a, b, n_per, levels, toward):
    """Composite Gauss-Legendre nodes on [a, b], segments halving towards
    ``toward`` ('a', 'b' or 'both')."""
    xg, wg = np.polynomial.legendre.leggauss(n_per)
    edges = [0.0, 1.0]
    # geometric grading in [0,1]
    pts = [2.0 ** (-j) for j in range(levels, 0, -1)]
    if toward == "a":
        e = [0.0] + pts + [1.0]
    elif toward == "b":
        e = [0.0] + [1.0 - p for p in reversed(pts)] + [1.0]
    else:
        half = [0.5 * p for p in pts]
        e = [0.0] + half + [0.5] + [1.0 - h for h in reversed(half)] + [1.0]
    edges = np.unique(np.array(e))
    xs, ws = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        lo_, hi_ = a + (b - a) * lo, a + (b - a) * hi
        xs.append(0.5 * (hi_ - lo_) * xg + 0.5 * (hi_ + lo_))
        ws.append(0.5 * (hi_ - lo_) * wg)
    return np.concatenate(xs), np.concatenate(ws)


def _kernel(x, y, lam):
    """Logistic kernel in x = 2 tau/beta - 1, y = omega/omega_max (stable form)."""
    X = x[:, None]
    Y = y[None, :]
    out = np.empty((x.size, y.size))
    pos = Y >= 0
    Yp = np.where(pos, Y, 0.0)
    Yn = np.where(pos, 0.0, -Y)
    kp = np.exp(-0.5 * lam * (X + 1.0) * Yp) / (1.0 + np.exp(-lam * Yp))
    kn = np.exp(-0.5 * lam * (1.0 - X) * Yn) / (1.0 + np.exp(-lam * Yn))
    out[:] = np.where(pos, kp, kn)
    return out


class IRBasis:
    def __init__(self, beta, lam, eps=1e-12, n_per=24):
        self.beta = float(beta)
        self.lam = float(lam)
        self.wmax = lam / beta
        lv = int(np.ceil(np.log2(lam))) + 3
        xr, wxr = _graded_gauss(0.0, 1.0, n_per, lv, "b")     # x in [0,1]
        self.x = np.concatenate([-xr[::-1], xr])
        self.wx = np.concatenate([wxr[::-1], wxr])
        yr, wyr = _graded_gauss(0.0, 1.0, n_per, lv, "a")     # y in [0,1]
        self.y = np.concatenate([-yr[::-1], yr])
        self.wy = np.concatenate([wyr[::-1], wyr])
        A = np.sqrt(self.wx)[:, None] * _kernel(self.x, self.y, lam) * np.sqrt(self.wy)[None, :]
        U, s, Vt = np.linalg.svd(A, full_matrices=False)
        L = int(np.sum(s / s[0] > eps))
        if L % 2:                          # keep nts = L + 2 even
            L -= 1
        self.L = L
        self.s = s[: L + 1]
        self.Vw = (Vt[: L + 1].T * np.sqrt(self.wy)[:, None])   # sqrt(w_j) v_l(y_j)

    def u(self, tau):
        """u_l(tau), l = 0..L (includes the first discarded function), (ntau, L+1)."""
        x = 2.0 * np.asarray(tau, dtype=np.float64) / self.beta - 1.0
        return _kernel(x, self.y, self.lam) @ self.Vw / self.s[None, :]

    def uhat_boson(self, n):
        """int_0^beta e^{i Omega_n tau} u_l(tau) dtau for bosonic Omega_n = 2 pi n / beta,
        with u_l normalised as in ``u``."""
        om = 2.0 * np.pi * np.asarray(n, dtype=np.float64) / self.beta
        w = self.wmax * self.y
        K = np.tanh(0.5 * self.beta * w)[None, :] / (w[None, :] - 1j * om[:, None])
        return K @ self.Vw / self.s[None, :]

    def tau_sampling(self):
        """Roots of u_L in (0, beta): bisection on a fine grid."""
        grid = np.concatenate([
            self.beta * 0.5 * (1 + _graded_gauss(-1.0, 1.0, 64, int(np.ceil(np.log2(self.lam))) + 6, "both")[0])
        ])
        grid = np.unique(np.clip(grid, 0.0, self.beta))
        f = self.u(grid)[:, self.L]
        roots = []
        for i in range(grid.size - 1):
            if f[i] == 0.0:
                roots.append(grid[i])
            elif f[i] * f[i + 1] < 0:
                a, b, fa = grid[i], grid[i + 1], f[i]
                for _ in range(80):
                    m = 0.5 * (a + b)
                    fm = self.u(np.array([m]))[0, self.L]
                    if fa * fm <= 0:
                        b = m
                    else:
                        a, fa = m, fm
                roots.append(0.5 * (a + b))
        roots = np.array(sorted(roots))
        if roots.size != self.L:
            raise RuntimeError(f"found {roots.size} sampling points, expected {self.L}")
        # u_L has definite parity under tau -> beta - tau, so its roots are
        # symmetric; the Nystrom evaluation of u_L carries noise ~ eps / s_L,
        # so the roots of the lower half are mirrored exactly (the GW kernel
        # relies on tau_{nts-1-t} = beta - tau_t, cugw_qpt.cu:184-191).
        lower = roots[: self.L // 2]
        return np.concatenate([lower, self.beta - lower[::-1]])

    def boson_frequencies(self):
        """Bosonic Matsubara sampling indices: n = 0 plus +-n at the sign changes
        of the first discarded function uhat_L(i Omega_n) (the choice of the
        sparse-ir sampling; Li et al., PRB 101, 035144 (2020))."""
        part = np.real if self.L % 2 == 0 else np.imag
        nmax = int(4 * self.lam) + 16
        vals = []
        for lo in range(0, nmax, 2048):
            n = np.arange(lo, min(lo + 2048, nmax))
            vals.append(part(self.uhat_boson(n)[:, self.L]))
        f = np.concatenate(vals)
        ch = np.nonzero(np.sign(f[:-1]) * np.sign(f[1:]) < 0)[0]
        pos = []
        for i in ch:           # the index of the two with the smaller |f|
            j = int(i) if abs(f[i]) < abs(f[i + 1]) else int(i) + 1
            if j > 0 and (not pos or j > pos[-1]):
                pos.append(j)
        # The sign changes alone give fewer than L + 1 rows; refine the largest
        # (relative) gaps until the fit for L coefficients is overdetermined.
        pts = [0] + pos
        while 2 * (len(pts) - 1) + 1 < self.L + 1:
            g = [(np.log((pts[i + 1] + 1.0) / (pts[i] + 1.0)), i) for i in range(len(pts) - 1)
                 if pts[i + 1] - pts[i] > 1]
            _, i = max(g)
            mid = int(round(np.sqrt((pts[i] + 1.0) * (pts[i + 1] + 1.0)) - 1.0))
            mid = min(max(mid, pts[i] + 1), pts[i + 1] - 1)
            pts.insert(i + 1, mid)
        pos = np.array(pts[1:], dtype=np.int64)
        return np.concatenate([-pos[::-1], [0], pos])


def transforms_on_grid(beta, lam, tau, eps=1e-14):
    """T_wt / T_tw for a given symmetric tau grid (endpoints included), e.g.
    the sampling points of another IR basis of the same Lambda.

    The bosonic coefficients are fitted by least squares from the nts-2
    interior points (which may outnumber the L basis functions) and from the
    Matsubara sampling frequencies; any basis that represents the bosonic
    functions to round-off gives the same tau -> tau map up to that
    round-off, whatever its own sampling points are."""
    tau = np.asarray(tau, dtype=np.float64)
    if tau.ndim != 1 or tau.size % 2 or tau[0] != 0.0 or tau[-1] != beta:
        raise ValueError("tau grid must have an even number of points from 0 to beta")
    if np.abs(tau + tau[::-1] - beta).max() > 1e-12 * beta:
        raise ValueError("tau grid must be symmetric about beta/2")
    b = IRBasis(beta, lam, eps)
    L = b.L
    if tau.size - 2 < L:
        raise ValueError(f"{tau.size - 2} interior tau points cannot fix {L} coefficients")
    n = b.boson_frequencies()
    U_int = b.u(tau[1:-1])[:, :L]
    U_all = b.u(tau)[:, :L]
    Uhat = b.uhat_boson(n)[:, :L]
    T_wt = Uhat @ np.linalg.pinv(U_int)
    T_tw = U_all @ np.linalg.pinv(Uhat)
    return {"beta": b.beta, "lam": b.lam, "tau": tau, "n_boson": n,
            "T_wt": T_wt.astype(np.complex128), "T_tw": T_tw.astype(np.complex128),
            "singular_values": b.s}


def make_transforms(beta, lam, eps=1e-12):
    """Return dict(tau (nts,), n_boson (nw_b,), T_wt (nw_b, nts-2), T_tw (nts, nw_b))."""
    b = IRBasis(beta, lam, eps)
    L = b.L
    tau_s = b.tau_sampling()
    tau = np.concatenate([[0.0], tau_s, [b.beta]])
    n = b.boson_frequencies()
    if n.size > tau.size:
        n = n[(n.size - tau.size) // 2 + 1: n.size - (n.size - tau.size) // 2 - 1]
    U_int = b.u(tau_s)[:, :L]                    # (L, L)
    U_all = b.u(tau)[:, :L]                      # (L+2, L)
    Uhat = b.uhat_boson(n)[:, :L]                # (nw, L)
    T_wt = Uhat @ np.linalg.inv(U_int)           # (nw, L) = (nw_b, nts-2)
    T_tw = U_all @ np.linalg.pinv(Uhat)          # (nts, nw_b)
    return {"beta": b.beta, "lam": b.lam, "tau": tau, "n_boson": n,
            "T_wt": T_wt.astype(np.complex128), "T_tw": T_tw.astype(np.complex128),
            "singular_values": b.s}
