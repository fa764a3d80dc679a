"""Python binding of libgreen_gpu_amd.so (the gfx950 HF/GW library).

The binding mirrors the reference plugin's kernel objects
(src/green/gpu/hf_gpu_kernel.h, gw_gpu_kernel.h): an HF kernel whose
``solve(dm)`` returns the Fock contribution and a GW kernel whose
``solve(G, Sigma)`` overwrites Sigma(tau).  Every call goes through the C ABI
declared in include/green_gpu_amd.h; there is no Python or CPU fallback - if
the shared library is missing, importing the kernels raises.

Tensor layouts are the reference's host layouts (complex128, row-major):
G, Sigma (nts, ns, ink, nso, nso); dm, S_k, Fock (ns, ink, nso, nso);
stored integrals (num_kpair_stored, NQ, nao, nao).  nso = nao for the
scalar kernels; nso = 2 nao (with ns = 1) selects the two-component (X2C)
kernels, as custom_hf_kernel / custom_gw_kernel do (gpu_factory.h:47-89).
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgreen_gpu_amd.so")

PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "i8x5": 3, "i8x4": 4}
SOLVER = {"cholesky": 0, "lu": 1}
RESIDENCE = {"device": 0, "host": 1, "generated": 2, "reader": 3}
# ggpu_pair_reader: int (*)(void* user, int64_t pos, double* dst)
PAIR_READER = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_double))


class GgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run green-gpu_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, u64, f64, sz = C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_size_t
        pd = C.POINTER(C.c_double)
        L.ggpu_last_error.restype = C.c_char_p
        L.ggpu_device_count.argtypes = [C.POINTER(C.c_int)]
        L.ggpu_device_memory.argtypes = [C.POINTER(sz), C.POINTER(sz)]
        L.ggpu_hf_plan_create.argtypes = [C.POINTER(vp), i32, i32, i32, i32, C.POINTER(SymmetryC)]
        L.ggpu_hf_set_integral_residence.argtypes = [vp, i32]
        L.ggpu_hf_set_k_share.argtypes = [vp, i32, i32]
        L.ggpu_hf_set_variants.argtypes = [vp, C.POINTER(C.c_int)]
        L.ggpu_hf_set_precision.argtypes = [vp, i32]
        L.ggpu_hf_load_integrals.argtypes = [vp, pd]
        L.ggpu_hf_set_integral_reader.argtypes = [vp, PAIR_READER, vp]
        L.ggpu_hf_synth_integrals.argtypes = [vp, u64, f64]
        L.ggpu_hf_exchange.argtypes = [vp, pd, pd]
        L.ggpu_hf_solve.argtypes = [vp, pd, pd, f64, pd]
        L.ggpu_hf_energy.argtypes = [vp, pd]
        L.ggpu_hf_exchange_device_noalloc.argtypes = [vp, vp, vp]
        L.ggpu_hf_plan_destroy.argtypes = [vp]
        L.ggpu_hf_plan_destroy.restype = None
        L.ggpu_gw_plan_create.argtypes = [C.POINTER(vp), i32, i32, i32, i32, i32, i32, C.POINTER(SymmetryC),
                                          pd, pd, i32, i32, sz]
        L.ggpu_gw_set_integral_residence.argtypes = [vp, i32]
        L.ggpu_gw_streamed_bytes.argtypes = [vp, pd]
        L.ggpu_gw_generated_bytes.argtypes = [vp, pd]
        L.ggpu_gw_staging_ms.argtypes = [vp, pd]
        L.ggpu_gw_set_q_share.argtypes = [vp, i32, i32]
        L.ggpu_gw_load_integrals.argtypes = [vp, pd]
        L.ggpu_gw_synth_integrals.argtypes = [vp, u64, f64]
        L.ggpu_gw_iterate.argtypes = [vp, pd, pd]
        L.ggpu_gw_iterate_device.argtypes = [vp, vp, vp]
        L.ggpu_gw_iterate_many.argtypes = [C.POINTER(vp), i32, C.POINTER(vp), C.POINTER(vp)]
        L.ggpu_gw_polarization.argtypes = [vp, pd, i32, i32, pd]
        L.ggpu_gw_last_polarization.argtypes = [vp, i32, C.POINTER(C.c_int), pd]
        L.ggpu_gw_set_pair_batch_limit.argtypes = [vp, i32]
        L.ggpu_gw_last_timings.argtypes = [vp, pd]
        L.ggpu_gw_set_profiling.argtypes = [vp, i32]
        L.ggpu_gw_site_stats.argtypes = [vp, i32, pd, pd, C.POINTER(C.c_int64)]
        L.ggpu_gw_set_graph.argtypes = [vp, i32]
        L.ggpu_gw_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), i32, C.POINTER(vp), C.POINTER(vp)]
        L.ggpu_gw_batch_run.argtypes = [vp, pd]
        L.ggpu_gw_batch_captures.argtypes = [vp, C.POINTER(C.c_int)]
        L.ggpu_gw_batch_destroy.argtypes = [vp]
        L.ggpu_gw_batch_destroy.restype = None
        L.ggpu_gw_graph_captures.argtypes = [vp, C.POINTER(C.c_int)]
        L.ggpu_gw_autotune.argtypes = [vp, pd, i32, i32]
        L.ggpu_gw_autotune_device.argtypes = [vp, vp, i32, i32]
        L.ggpu_gw_autotune_report.argtypes = [vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), pd, pd]
        L.ggpu_gw_get_variants.argtypes = [vp, C.POINTER(C.c_int)]
        L.ggpu_gw_set_integral_reader.argtypes = [vp, PAIR_READER, vp]
        L.ggpu_gw_set_variants.argtypes = [vp, C.POINTER(C.c_int)]
        L.ggpu_gw_plan_destroy.argtypes = [vp]
        L.ggpu_gw_plan_destroy.restype = None
        _lib = L
    return _lib


EXPORTED = ["ggpu_last_error", "ggpu_device_count", "ggpu_device_memory",
            "ggpu_hf_plan_create", "ggpu_hf_set_integral_residence", "ggpu_hf_set_k_share", "ggpu_hf_set_variants",
            "ggpu_hf_set_precision",
            "ggpu_hf_load_integrals", "ggpu_hf_set_integral_reader",
            "ggpu_hf_synth_integrals",
            "ggpu_hf_exchange", "ggpu_hf_exchange_device_noalloc", "ggpu_hf_solve", "ggpu_hf_energy",
            "ggpu_hf_plan_destroy",
            "ggpu_gw_plan_create", "ggpu_gw_set_integral_residence", "ggpu_gw_streamed_bytes",
            "ggpu_gw_generated_bytes", "ggpu_gw_staging_ms", "ggpu_gw_set_q_share",
            "ggpu_gw_load_integrals", "ggpu_gw_synth_integrals", "ggpu_gw_set_integral_reader",
            "ggpu_gw_iterate", "ggpu_gw_iterate_device", "ggpu_gw_iterate_many", "ggpu_gw_polarization",
            "ggpu_gw_last_polarization", "ggpu_gw_set_pair_batch_limit",
            "ggpu_gw_last_timings", "ggpu_gw_set_profiling", "ggpu_gw_site_stats",
            "ggpu_gw_set_graph", "ggpu_gw_graph_captures",
            "ggpu_gw_batch_create", "ggpu_gw_batch_run", "ggpu_gw_batch_captures", "ggpu_gw_batch_destroy",
            "ggpu_gw_autotune", "ggpu_gw_autotune_device", "ggpu_gw_autotune_report",
            "ggpu_gw_get_variants", "ggpu_gw_set_variants",
            "ggpu_gw_plan_destroy"]

# Launch variants of the contraction GEMMs (ids as the library numbers them):
# FP64 / FP32 LDS k-depth and waves-per-SIMD register allocation, split-bf16
# workgroup tiles; in the int8 modes the emulation with the mode's slices
# ("int8") or with five slices ("int8x5", offered to four-slice plans).
VARIANTS_MFMA = {8: "bk8", 16: "bk16", 83: "bk8_w3", 84: "bk8_w4", 200: "int8", 201: "int8x5"}
VARIANTS_BF16 = {0: "64x64", 1: "128x64", 4: "64x128", 5: "160x64", 6: "96x64_w3", 7: "64x96_w3", 8: "64x64_w3",
                 9: "64x64_astat", 10: "128x128_presplit", 11: "128x128_presplit_w8",
                 12: "128x128_presplit_w16", 13: "128x192_presplit_w8"}

SITES = ["other", "p0_gv", "p0_gvh", "p0_contract", "sig_vg", "sig_py", "sig_contract",
         "ir", "dyson", "hf", "rotate"]


def _check(rc):
    if rc != 0:
        raise GgpuError(rc, lib().ggpu_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c128(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


# ---------------------------------------------------------------------------
# Symmetry tables
# ---------------------------------------------------------------------------
class SymmetryC(C.Structure):
    _fields_ = [("nk", C.c_int32), ("ink", C.c_int32), ("nq", C.c_int32), ("inq", C.c_int32)] + [
        (n, C.c_void_p) for n in (
            "k_full_to_reduced", "k_reduced_to_full", "k_tr_conj", "k_star_offsets", "k_star_members",
            "q_full_to_reduced", "q_reduced_to_full", "q_tr_conj", "q_star_offsets", "q_star_members",
            "k1_from_k2q", "k2_from_k1q", "U_k", "U_q", "conj_kpair", "trans_kpair", "irre_pos")
    ] + [("num_kpair_stored", C.c_int64)]


@dataclass
class Symmetry:
    """Plain symmetry tables (the fields of green-symmetry the path uses)."""
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
    k1_from_k2q: np.ndarray
    k2_from_k1q: np.ndarray
    U_k: object = None
    U_q: object = None
    conj_kpair: object = None
    trans_kpair: object = None
    irre_pos: object = None
    num_kpair_stored: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @classmethod
    def like(cls, other):
        """Copy the tables of any object exposing the same attribute names."""
        names = [f for f in cls.__dataclass_fields__ if not f.startswith("_")]
        return cls(**{n: getattr(other, n, None) for n in names})

    def to_c(self):
        keep = []

        def i64(a):
            a = np.ascontiguousarray(a, dtype=np.int64)
            keep.append(a)
            return a.ctypes.data

        def c128(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.complex128)
            keep.append(a)
            return a.ctypes.data

        ko = np.concatenate([[0], np.cumsum([len(s) for s in self.k_stars])]).astype(np.int64)
        qo = np.concatenate([[0], np.cumsum([len(s) for s in self.q_stars])]).astype(np.int64)
        c = SymmetryC()
        c.nk, c.ink, c.nq, c.inq = int(self.nk), int(self.ink), int(self.nq), int(self.inq)
        c.k_full_to_reduced = i64(self.k_full_to_reduced)
        c.k_reduced_to_full = i64(self.k_reduced_to_full)
        c.k_tr_conj = i64(self.k_tr_conj)
        c.k_star_offsets = i64(ko)
        c.k_star_members = i64(np.concatenate([np.asarray(s) for s in self.k_stars]))
        c.q_full_to_reduced = i64(self.q_full_to_reduced)
        c.q_reduced_to_full = i64(self.q_reduced_to_full)
        c.q_tr_conj = i64(self.q_tr_conj)
        c.q_star_offsets = i64(qo)
        c.q_star_members = i64(np.concatenate([np.asarray(s) for s in self.q_stars]))
        c.k1_from_k2q = i64(self.k1_from_k2q)
        c.k2_from_k1q = i64(self.k2_from_k1q)
        c.U_k = c128(self.U_k)
        c.U_q = c128(self.U_q)
        if self.irre_pos is not None:
            c.conj_kpair = i64(self.conj_kpair)
            c.trans_kpair = i64(self.trans_kpair)
            c.irre_pos = i64(self.irre_pos)
            c.num_kpair_stored = int(self.num_kpair_stored)
        else:
            c.num_kpair_stored = int(self.nk * (self.nk + 1) // 2)
        self._keep = keep
        return c


def mesh_symmetry_none(mesh):
    """Tables of a Monkhorst-Pack mesh (n1, n2, n3) with no symmetry
    reduction: k = (i1 n2 + i2) n3 + i3, k1 = k2 + q and k2 = k1 - q
    component-wise modulo the mesh, U = I, every pair k1 >= k2 stored."""
    n1, n2, n3 = (int(x) for x in mesh)
    nk = n1 * n2 * n3
    i = np.arange(nk)
    c = np.stack([i // (n2 * n3), (i // n3) % n2, i % n3], axis=1)
    n = np.array([n1, n2, n3])

    def idx(v):
        v = v % n
        return (v[..., 0] * n2 + v[..., 1]) * n3 + v[..., 2]

    k1 = idx(c[:, None, :] + c[None, :, :])
    k2 = idx(c[:, None, :] - c[None, :, :])
    ar = np.arange(nk, dtype=np.int64)
    return Symmetry(nk=nk, ink=nk, nq=nk, inq=nk,
                    k_full_to_reduced=ar, k_reduced_to_full=ar, k_tr_conj=np.zeros(nk, np.int64),
                    k_stars=[[k] for k in range(nk)],
                    q_full_to_reduced=ar, q_reduced_to_full=ar, q_tr_conj=np.zeros(nk, np.int64),
                    q_stars=[[q] for q in range(nk)],
                    k1_from_k2q=k1.astype(np.int64), k2_from_k1q=k2.astype(np.int64),
                    num_kpair_stored=nk * (nk + 1) // 2)


def iterate_many(kernels, G_ptrs, Sigma_ptrs):
    """One GW iteration for several independent systems (GwKernel objects of
    any sizes) on device-resident G / Sigma; the systems overlap on the GPU."""
    n = len(kernels)
    if not (n == len(G_ptrs) == len(Sigma_ptrs)) or n == 0:
        raise ValueError("one G and one Sigma pointer per kernel")
    arr = lambda xs: (C.c_void_p * n)(*[C.c_void_p(x) for x in xs])
    _check(lib().ggpu_gw_iterate_many(arr([k._h.value for k in kernels]), n, arr(G_ptrs), arr(Sigma_ptrs)))


_hip = None


class GwBatch:
    """Several GwKernel plans (integrals resident in HBM) replayed as one HIP
    graph launch; G_ptrs / Sigma_ptrs are device pointers, fixed here."""

    def __init__(self, kernels, G_ptrs, Sigma_ptrs):
        n = len(kernels)
        if not (n == len(G_ptrs) == len(Sigma_ptrs)):
            raise ValueError("one G and one Sigma per kernel")
        self._kernels = list(kernels)          # the plans must outlive the batch
        plans = (C.c_void_p * n)(*[k._h.value for k in kernels])
        gp = (C.c_void_p * n)(*G_ptrs)
        sp = (C.c_void_p * n)(*Sigma_ptrs)
        h = C.c_void_p()
        _check(lib().ggpu_gw_batch_create(C.byref(h), plans, n, gp, sp))
        self._h = h

    def run(self):
        """One replay of every system's iteration; returns its device time (ms)."""
        ms = C.c_double()
        _check(lib().ggpu_gw_batch_run(self._h, C.byref(ms)))
        return ms.value

    def captures(self):
        n = C.c_int()
        _check(lib().ggpu_gw_batch_captures(self._h, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().ggpu_gw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _hip_runtime():
    """The HIP runtime libgreen_gpu_amd.so itself is linked against (already
    loaded by the library; dlopen resolves the same soname to it).  Device
    buffers handed to the library must come from this runtime: a second HIP
    runtime in the process (torch's bundled one, for instance) keeps its own
    view of the device and does not interoperate with it."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        H.hipMalloc.argtypes = [C.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, C.c_int]
        H.hipMemset.argtypes = [vp, C.c_int, sz]
        H.hipDeviceSynchronize.argtypes = []
        H.hipGetErrorString.argtypes = [C.c_int]
        H.hipGetErrorString.restype = C.c_char_p
        _hip = H
    return _hip


def _hip_check(rc, what):
    if rc != 0:
        raise GgpuError(rc, f"{what}: {_hip_runtime().hipGetErrorString(rc).decode()}")


_H2D, _D2H, _D2D = 1, 2, 3


class DeviceArray:
    """A device buffer holding a copy of a numpy array (dtype and shape kept
    on the host side).  ``ptr`` is what the *_device entry points take."""

    def __init__(self, shape, dtype=np.complex128, zero=False):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _hip_check(_hip_runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._p = p
        if zero:
            _hip_check(_hip_runtime().hipMemset(p, 0, self.nbytes), "hipMemset")
            synchronize()

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _hip_check(_hip_runtime().hipMemcpy(d._p, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        synchronize()     # the library's streams do not wait on the null stream
        return d

    @classmethod
    def empty_like(cls, other, zero=False):
        return cls(other.shape, other.dtype, zero=zero)

    @property
    def ptr(self):
        if self._p is None:
            raise ValueError("device buffer already freed")
        return self._p.value

    def copy_from(self, other):
        """Device-to-device copy of a buffer of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _hip_check(_hip_runtime().hipMemcpy(self._p, other._p, self.nbytes, _D2D), "hipMemcpy D2D")
        synchronize()

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _hip_check(_hip_runtime().hipMemcpy(out.ctypes.data, self._p, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self._p is not None and self._p.value:
            _hip_runtime().hipFree(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _hip_check(_hip_runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_count():
    n = C.c_int(0)
    _check(lib().ggpu_device_count(C.byref(n)))
    return n.value


def device_memory():
    f, t = C.c_size_t(0), C.c_size_t(0)
    _check(lib().ggpu_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value



def _pair_reader(read_pair, pair_elems):
    """ctypes ggpu_pair_reader around read_pair(pos) -> NQ * nao^2 complex
    values of stored pair pos (a Python exception becomes status 1)."""
    def cb(user, pos, dst):
        try:
            out = np.ctypeslib.as_array(dst, shape=(2 * pair_elems,)).view(np.complex128)
            out[:] = np.asarray(read_pair(int(pos)), dtype=np.complex128).reshape(-1)
            return 0
        except Exception:
            return 1
    return PAIR_READER(cb)


# ---------------------------------------------------------------------------
# Kernels
# ---------------------------------------------------------------------------
class HfKernel:
    """HF kernel on the GPU (reference: scalar_hf_gpu_kernel, hf_gpu_kernel.h:67-84;
    x2c_hf_gpu_kernel, hf_gpu_kernel.h:87-100, when nso = 2 nao)."""

    def __init__(self, nao, ns, NQ, sym, madelung=0.0, S_k=None, nso=None, integrals="device", precision="fp64"):
        self.nao, self.ns, self.NQ = int(nao), int(ns), int(NQ)
        self.nso = self.nao if nso is None else int(nso)
        self.sym = sym if isinstance(sym, Symmetry) else Symmetry.like(sym)
        self.madelung = float(madelung)
        self.S_k = None if S_k is None else _c128(S_k, self.mat_shape)
        h = C.c_void_p()
        self._csym = self.sym.to_c()
        _check(lib().ggpu_hf_plan_create(C.byref(h), self.nao, self.nso, self.ns, self.NQ, C.byref(self._csym)))
        self._h = h
        _check(lib().ggpu_hf_set_integral_residence(self._h, RESIDENCE[integrals]))
        _check(lib().ggpu_hf_set_precision(self._h, PRECISION[precision]))

    @property
    def mat_shape(self):
        return (self.ns, self.sym.ink, self.nso, self.nso)

    def load_integrals(self, V_stored):
        V = _c128(V_stored)
        _check(lib().ggpu_hf_load_integrals(self._h, _dptr(V)))

    def set_integral_reader(self, read_pair):
        """READER residence: stored pair pos comes from read_pair(pos) (NQ x
        nao x nao complex) whenever a pair batch needs it."""
        self._reader = _pair_reader(read_pair, self.NQ * self.nao * self.nao)
        _check(lib().ggpu_hf_set_integral_reader(self._h, self._reader, None))

    def synth_integrals(self, seed, sigma=0.1):
        _check(lib().ggpu_hf_synth_integrals(self._h, int(seed), float(sigma)))

    def set_k_share(self, rank, size):
        """Compute only the rows of IBZ k = rank, rank + size, ...
        (cuhf_utils.cu:140); the other rows are zero."""
        _check(lib().ggpu_hf_set_k_share(self._h, int(rank), int(size)))

    def set_variants(self, variants):
        """Launch variant per call site (only site "hf"; 200: int8 emulation
        with the mode's slices, 201: with five slices, 83: FP64 kernel,
        -1: built-in) for the exchange products."""
        v = (C.c_int * len(SITES))(*[int(x) for x in variants])
        _check(lib().ggpu_hf_set_variants(self._h, v))

    def exchange(self, dm):
        dm = _c128(dm, self.mat_shape)
        K = np.empty_like(dm)
        _check(lib().ggpu_hf_exchange(self._h, _dptr(dm), _dptr(K)))
        return K

    def solve(self, dm):
        """J + K - Ewald for the IBZ density matrix (hf_gpu_kernel.cpp:66-104)."""
        dm = _c128(dm, self.mat_shape)
        S = self.S_k if self.S_k is not None else np.zeros_like(dm)
        F = np.empty_like(dm)
        _check(lib().ggpu_hf_solve(self._h, _dptr(dm), _dptr(S), self.madelung, _dptr(F)))
        return F

    def energy(self):
        """Two-electron HF energy of the last solve: (1/nk) sum_{s, k in IBZ}
        w_k Re Tr[D_s(k) F_s(k)] / 2 with star weights w_k (FP64 on the device)."""
        e = C.c_double()
        _check(lib().ggpu_hf_energy(self._h, C.byref(e)))
        return e.value

    def close(self):
        if getattr(self, "_h", None):
            lib().ggpu_hf_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class GwKernel:
    """GW kernel on the GPU (reference: scalar_gw_gpu_kernel, gw_gpu_kernel.h:139-172;
    x2c_gw_gpu_kernel, gw_gpu_kernel.h:169-206, when nso = 2 nao)."""

    def __init__(self, nao, ns, NQ, T_wt, T_tw, sym, precision="fp64", solver="lu", work_bytes=0, nso=None,
                 integrals="device"):
        T_wt = _c128(T_wt)
        T_tw = _c128(T_tw)
        self.nao, self.ns, self.NQ = int(nao), int(ns), int(NQ)
        self.nso = self.nao if nso is None else int(nso)
        self.nw_b, self.nts = T_wt.shape[0], T_tw.shape[0]
        self.sym = sym if isinstance(sym, Symmetry) else Symmetry.like(sym)
        h = C.c_void_p()
        self._csym = self.sym.to_c()
        _check(lib().ggpu_gw_plan_create(C.byref(h), self.nao, self.nso, self.ns, self.NQ, self.nts, self.nw_b,
                                         C.byref(self._csym), _dptr(T_wt), _dptr(T_tw),
                                         PRECISION[precision], SOLVER[solver.lower()], int(work_bytes)))
        self._h = h
        self.precision = precision
        _check(lib().ggpu_gw_set_integral_residence(self._h, RESIDENCE[integrals]))

    @property
    def shape(self):
        return (self.nts, self.ns, self.sym.ink, self.nso, self.nso)

    def load_integrals(self, V_stored):
        V = _c128(V_stored)
        _check(lib().ggpu_gw_load_integrals(self._h, _dptr(V)))

    def set_integral_reader(self, read_pair):
        """READER residence: stored pair pos comes from read_pair(pos) (NQ x
        nao x nao complex) whenever a pair batch needs it."""
        self._reader = _pair_reader(read_pair, self.NQ * self.nao * self.nao)
        _check(lib().ggpu_gw_set_integral_reader(self._h, self._reader, None))

    def synth_integrals(self, seed, sigma=0.1):
        _check(lib().ggpu_gw_synth_integrals(self._h, int(seed), float(sigma)))

    def solve(self, G, Sigma=None):
        """One GW iteration; returns (and fills, if given) Sigma(tau)."""
        G = _c128(G, self.shape)
        out = np.empty(self.shape, dtype=np.complex128) if Sigma is None else Sigma
        if out.dtype != np.complex128 or not out.flags.c_contiguous or out.shape != self.shape:
            raise ValueError("Sigma must be a C-contiguous complex128 array of shape %s" % (self.shape,))
        _check(lib().ggpu_gw_iterate(self._h, _dptr(G), _dptr(out)))
        return out

    def solve_device(self, G_ptr, Sigma_ptr):
        """Iteration on device-resident G / Sigma (raw device pointers)."""
        _check(lib().ggpu_gw_iterate_device(self._h, C.c_void_p(G_ptr), C.c_void_p(Sigma_ptr)))

    def polarization(self, G, iq, which):
        G = _c128(G, self.shape)
        n = {0: self.nts, 1: self.nw_b, 2: self.nts}[which]
        out = np.empty((n, self.NQ, self.NQ), dtype=np.complex128)
        _check(lib().ggpu_gw_polarization(self._h, _dptr(G), int(iq), int(which), _dptr(out)))
        return out

    def last_polarization(self, which):
        """(iq, array): P0(tau) (which 0), P(Omega) (1) or P(tau) (2) of the last
        q the last iteration processed, read back without recomputation."""
        n = {0: self.nts, 1: self.nw_b, 2: self.nts}[which]
        out = np.empty((n, self.NQ, self.NQ), dtype=np.complex128)
        iq = C.c_int()
        _check(lib().ggpu_gw_last_polarization(self._h, int(which), C.byref(iq), _dptr(out)))
        return iq.value, out

    def set_pair_batch_limit(self, batches):
        """Run only the first `batches` pair batches of every q's P0 and Sigma
        chains (0: all) - a partial iteration for warm-up and profiling."""
        _check(lib().ggpu_gw_set_pair_batch_limit(self._h, int(batches)))

    def streamed_bytes(self):
        """Integral bytes copied host -> device by the last iteration."""
        b = C.c_double()
        _check(lib().ggpu_gw_streamed_bytes(self._h, C.byref(b)))
        return b.value

    def generated_bytes(self):
        """Integral bytes generated on the device by the last iteration."""
        b = C.c_double()
        _check(lib().ggpu_gw_generated_bytes(self._h, C.byref(b)))
        return b.value

    def staging_ms(self):
        """Copy-stream time (ms) of the last iteration's integral staging."""
        b = C.c_double()
        _check(lib().ggpu_gw_staging_ms(self._h, C.byref(b)))
        return b.value

    def set_q_share(self, rank, size):
        """Run only IBZ q = rank, rank + size, ... (cugw_utils.cu:115); the
        Sigma of all shares sums to the full Sigma."""
        _check(lib().ggpu_gw_set_q_share(self._h, int(rank), int(size)))

    def timings(self):
        t = np.zeros(6)
        _check(lib().ggpu_gw_last_timings(self._h, _dptr(t)))
        return dict(zip(["total", "stage_G", "p0", "ir_dyson", "sigma", "stage_sigma"], t))

    def set_profiling(self, on=True):
        _check(lib().ggpu_gw_set_profiling(self._h, 1 if on else 0))

    def site_stats(self):
        """{site: (ms, algorithmic FLOPs, launches)} of the last profiled iteration."""
        out = {}
        for i, name in enumerate(SITES):
            ms, fl, n = C.c_double(), C.c_double(), C.c_int64()
            _check(lib().ggpu_gw_site_stats(self._h, i, C.byref(ms), C.byref(fl), C.byref(n)))
            if n.value:
                out[name] = (ms.value, fl.value, n.value)
        return out

    def set_graph(self, on=True):
        """Capture the iteration into a HIP graph on the next solve and replay
        it afterwards (integrals in HBM, streamed from host memory or
        generated on the device; not with a pair reader)."""
        _check(lib().ggpu_gw_set_graph(self._h, 1 if on else 0))

    def graph_captures(self):
        n = C.c_int()
        _check(lib().ggpu_gw_graph_captures(self._h, C.byref(n)))
        return n.value

    def variant_name(self, v):
        """Readable name of a launch variant id in this plan's precision."""
        names = VARIANTS_BF16 if self.precision == "bf16x3" else VARIANTS_MFMA
        return "default" if v == -1 else names.get(v, str(v))

    def autotune(self, G, rounds=2, batches=1):
        """Time every launch variant of the six contraction sites on the first
        q of this plan's share and keep the fastest per site (results are
        unchanged).  Returns {site: {"chosen": name, "ms": {name: ms}}}."""
        G = _c128(G, self.shape)
        _check(lib().ggpu_gw_autotune(self._h, _dptr(G), int(rounds), int(batches)))
        return self.autotune_report()

    def autotune_device(self, G_ptr, rounds=2, batches=1):
        _check(lib().ggpu_gw_autotune_device(self._h, C.c_void_p(G_ptr), int(rounds), int(batches)))
        return self.autotune_report()

    def autotune_report(self):
        n = C.c_int()
        cap = 16
        ids = (C.c_int * cap)()
        ms = np.full(cap * len(SITES), -1.0)
        wall = C.c_double()
        _check(lib().ggpu_gw_autotune_report(self._h, cap, C.byref(n), ids, _dptr(ms), C.byref(wall)))
        chosen = self.get_variants()
        out = {"wall_ms": wall.value, "sites": {}}
        for si, name in enumerate(SITES):
            times = {self.variant_name(ids[c]): float(ms[c * len(SITES) + si])
                     for c in range(min(n.value, cap)) if ms[c * len(SITES) + si] >= 0}
            if times:
                out["sites"][name] = {"chosen": self.variant_name(chosen[si]), "ms": times}
        return out

    def get_variants(self):
        v = (C.c_int * len(SITES))()
        _check(lib().ggpu_gw_get_variants(self._h, v))
        return list(v)

    def set_variants(self, variants):
        v = (C.c_int * len(SITES))(*[int(x) for x in variants])
        _check(lib().ggpu_gw_set_variants(self._h, v))

    def close(self):
        if getattr(self, "_h", None):
            lib().ggpu_gw_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
