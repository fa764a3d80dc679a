"""The no-allocation HF exchange (ggpu_hf_exchange_device_noalloc), called
through the C ABI on device buffers, against the allocating entry point
ggpu_hf_exchange on the same plan: two calls in a row, the second on a
smaller system (the static per-GPU scratch keeps the first call's partial
sums, which must not leak into the second), scalar ns = 2, ns = 1 and a
two-component plan, FP64 to 1e-12 of max |K| (the partial slices may be
fewer than the allocating path's, so the summation order can differ)."""
import ctypes as C

import numpy as np
import pytest

from conftest import max_abs

pytestmark = pytest.mark.gpu


def noalloc(ggpu, kern, dm):
    L = ggpu.lib()
    f = L.ggpu_hf_exchange_device_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    dd = ggpu.DeviceArray.from_numpy(np.ascontiguousarray(dm))
    Kd = ggpu.DeviceArray.empty_like(dd)
    rc = f(kern._h, C.c_void_p(dd.ptr), C.c_void_p(Kd.ptr))
    assert rc == 0, L.ggpu_last_error().decode()
    out = Kd.numpy()
    dd.free()
    Kd.free()
    return out


@pytest.mark.parametrize("first,second", [
    (((2, 2, 2), 26, 100, 2), ((2, 1, 1), 13, 40, 2)),
    (((2, 2, 1), 20, 64, 1), ((2, 1, 1), 7, 20, 1)),
])
def test_noalloc_exchange_matches_allocating_entry_point(ggpu, synth, first, second):
    for i, (mesh, nao, NQ, ns) in enumerate((first, second)):
        sym = ggpu.mesh_symmetry_none(mesh)
        tau, _, _ = synth.load_ir(50.0, 1e3)
        _, D = synth.make_green(sym.nk, nao, ns, 50.0, tau, 300 + i)
        k = ggpu.HfKernel(nao, ns, NQ, sym)
        k.synth_integrals(400 + i, 0.1)
        ref = k.exchange(D)
        got = noalloc(ggpu, k, D)
        print(f"call {i}: mesh={mesh} nao={nao} NQ={NQ} ns={ns}: max-abs {max_abs(got, ref):.2e}", flush=True)
        assert max_abs(got, ref) <= 1e-12 * np.abs(ref).max()
        k.close()


def test_noalloc_exchange_two_component(ggpu, synth):
    mesh, nao, NQ = (2, 1, 1), 6, 18
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, _, _ = synth.load_ir(50.0, 1e3)
    _, D = synth.make_green(sym.nk, 2 * nao, 1, 50.0, tau, 501)
    k = ggpu.HfKernel(nao, 1, NQ, sym, nso=2 * nao)
    k.synth_integrals(502, 0.1)
    ref = k.exchange(D)
    assert max_abs(noalloc(ggpu, k, D), ref) <= 1e-12 * np.abs(ref).max()


def test_noalloc_exchange_refuses_staged_integrals(ggpu, synth):
    mesh, nao, NQ, ns = (2, 1, 1), 5, 12, 2
    sym = ggpu.mesh_symmetry_none(mesh)
    tau, _, _ = synth.load_ir(50.0, 1e3)
    _, D = synth.make_green(sym.nk, nao, ns, 50.0, tau, 601)
    k = ggpu.HfKernel(nao, ns, NQ, sym, integrals="generated")
    k.synth_integrals(602, 0.1)
    L = ggpu.lib()
    f = L.ggpu_hf_exchange_device_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    dd = ggpu.DeviceArray.from_numpy(np.ascontiguousarray(D))
    Kd = ggpu.DeviceArray.empty_like(dd)
    assert f(k._h, C.c_void_p(dd.ptr), C.c_void_p(Kd.ptr)) == 5
    assert "resident in HBM" in L.ggpu_last_error().decode()
    dd.free()
    Kd.free()
