"""The C ABI: the library loads, exports every function include/green_gpu_amd.h
declares, and rejects the inputs the reference rejects - as status codes
with the reference's messages, before any device work (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "green_gpu_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ggpu_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol(ggpu):
    L = ggpu.lib()
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(L, n), f"{n} declared but not exported"
    assert sorted(ggpu.EXPORTED) == names


def test_no_cpu_fallback_in_binding(ggpu):
    """The Python binding only forwards to the shared library."""
    src = open(os.path.join(ROOT, "green-gpu_amd", "__init__.py")).read()
    assert "oracle" not in src and "gw_oracle" not in src


def small_tables(ggpu, mesh=(2, 1, 1)):
    return ggpu.mesh_symmetry_none(mesh)


def ir(synth):
    return synth.load_ir(50.0, 1e3)


def expect_error(fn, code, text):
    import green_gpu_amd as g
    with pytest.raises(g.GgpuError) as e:
        fn()
    assert e.value.code == code
    assert text in str(e.value)


def test_gw_odd_nts_rejected(ggpu, synth):
    """gw_gpu_kernel.h:67: nts must be even."""
    tau, T_wt, T_tw = ir(synth)
    sym = small_tables(ggpu)
    Twt = np.ascontiguousarray(T_wt[:, :-1])
    Ttw = np.ascontiguousarray(T_tw[:-1])
    expect_error(lambda: ggpu.GwKernel(4, 1, 6, Twt, Ttw, sym), 1, "nts must be even")


def test_gw_nts_below_nwb_rejected(ggpu, synth):
    """cugw_qpt.cu:79: Nt is not greater than or equal to Nw_b."""
    sym = small_tables(ggpu)
    nts, nw = 6, 8
    expect_error(lambda: ggpu.GwKernel(4, 1, 6, np.zeros((nw, nts - 2)), np.zeros((nts, nw)), sym), 1,
                 "Nt is not greater than or equal to Nw_b")


def test_bad_spin_count_rejected(ggpu, synth):
    """cuhf_utils.cu:65-71: Invalid value of "ns"."""
    tau, T_wt, T_tw = ir(synth)
    sym = small_tables(ggpu)
    expect_error(lambda: ggpu.HfKernel(4, 3, 6, sym), 1, 'Invalid value of "ns"')
    expect_error(lambda: ggpu.GwKernel(4, 3, 6, T_wt, T_tw, sym), 1, "ns must be 1 or 2")


def test_bad_spinor_dimensions_rejected(ggpu, synth):
    """nso must be nao (scalar) or 2 nao (X2C, gpu_factory.h:47-89 pick the
    X2C kernels when nso != nao); X2C runs with ns = 1."""
    tau, T_wt, T_tw = ir(synth)
    sym = small_tables(ggpu)
    expect_error(lambda: ggpu.HfKernel(4, 1, 6, sym, nso=6), 1, "nso must be nao or 2 nao")
    expect_error(lambda: ggpu.GwKernel(4, 1, 6, T_wt, T_tw, sym, nso=5), 1, "nso must be nao or 2 nao")
    expect_error(lambda: ggpu.HfKernel(4, 2, 6, sym, nso=8), 1, "needs ns = 1")
    expect_error(lambda: ggpu.GwKernel(4, 2, 6, T_wt, T_tw, sym, nso=8), 1, "needs ns = 1")
    tr = small_tables(ggpu)
    tr.k_tr_conj = np.array([0, 1])
    expect_error(lambda: ggpu.GwKernel(4, 1, 6, T_wt, T_tw, tr, nso=8), 1, "time-reversal")


def test_bad_symmetry_tables_rejected(ggpu, synth):
    tau, T_wt, T_tw = ir(synth)
    sym = small_tables(ggpu)
    sym.k_full_to_reduced = np.array([0, 5])
    expect_error(lambda: ggpu.HfKernel(4, 1, 6, sym), 1, "out of range")
    sym = small_tables(ggpu)
    sym.k_stars = [[0], []]
    expect_error(lambda: ggpu.HfKernel(4, 1, 6, sym), 1, "star")


def test_unknown_solver_and_null_args(ggpu, synth):
    L = ggpu.lib()
    tau, T_wt, T_tw = ir(synth)
    sym = small_tables(ggpu)
    cs = sym.to_c()
    h = C.c_void_p()
    T_wt = np.ascontiguousarray(T_wt)
    T_tw = np.ascontiguousarray(T_tw)
    pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = L.ggpu_gw_plan_create(C.byref(h), 4, 4, 1, 6, T_tw.shape[0], T_wt.shape[0], C.byref(cs),
                               pd(T_wt), pd(T_tw), 0, 7, 0)
    assert rc == 1 and b"unknown linear solver" in L.ggpu_last_error()
    rc = L.ggpu_gw_plan_create(C.byref(h), 4, 4, 1, 6, T_tw.shape[0], T_wt.shape[0], None,
                               pd(T_wt), pd(T_tw), 0, 0, 0)
    assert rc == 1
    rc = L.ggpu_gw_iterate(None, None, None)
    assert rc == 1 and b"NULL" in L.ggpu_last_error()
