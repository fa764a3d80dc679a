"""The C++ adapter for green-mbpt (include/green/gpu/gpu_factory.h): compiled
against the green-* stand-ins in tests/adapter/stubs and HDF5, then run on
the reference's symmetric HF and GW fixture system (27 k, 4 IBZ, conj/trans
pair storage).  It must read green-mbpt's chunked integral files back bit for
bit, hand the symmetry tables to the C ABI unchanged, and - without a GPU -
stop with the library's device error in both kernels (with one, its Fock
matrices are checked against the reference's Sigma1 and the GW kernel's
Sigma(tau), through custom_gw_kernel with a transformer_t carrying the
Lambda = 1e4 transforms on the grid recovered from the fixture, against the
reference's Sigma_tau, with the integrals read whole and read chunk by chunk
per pair batch under cuda_low_cpu_memory)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, max_abs

HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")
LIB = os.path.join(ROOT, "green-gpu_amd", "libgreen_gpu_amd.so")

pytestmark = pytest.mark.skipif(
    not (os.path.exists(os.path.join(HDF5, "include", "hdf5_hl.h")) and shutil.which("g++") and os.path.exists(LIB)),
    reason="needs g++, HDF5 headers/libraries (HDF5_ROOT) and the built library")


def write_inputs(d, fx, oracle, chunk=37):
    NQ, ns, nao = int(fx["params_NQ"]), int(fx["params_ns"]), int(fx["params_nao"])
    sym = oracle.symmetry_from_fixture(fx, NQ)
    with open(os.path.join(d, "sizes.txt"), "w") as f:
        f.write(f"nao {nao}\nnso {nao}\nns {ns}\nNQ {NQ}\nchunk {chunk}\nmadelung {float(fx['madelung'])!r}\n")
    rows = {"k_f2r": sym.k_full_to_reduced, "k_r2f": sym.k_reduced_to_full, "k_tr": sym.k_tr_conj,
            "q_f2r": sym.q_full_to_reduced, "q_r2f": sym.q_reduced_to_full, "q_tr": sym.q_tr_conj,
            "k1": np.asarray(sym.k1_from_k2q).ravel(), "k2": np.asarray(sym.k2_from_k1q).ravel(),
            "conj": sym.conj_kpair, "trans": sym.trans_kpair, "irre": sym.irre_pos,
            "nstored": [sym.num_kpair_stored]}
    for i, st in enumerate(sym.k_stars):
        rows[f"k_star_{i}"] = st
    for i, st in enumerate(sym.q_stars):
        rows[f"q_star_{i}"] = st
    with open(os.path.join(d, "tables.txt"), "w") as f:
        for k, v in rows.items():
            f.write(k + " " + " ".join(str(int(x)) for x in np.asarray(v).ravel()) + "\n")
    np.ascontiguousarray(sym.U_k, dtype=np.complex128).tofile(os.path.join(d, "U_k.bin"))
    np.ascontiguousarray(sym.U_q, dtype=np.complex128).tofile(os.path.join(d, "U_q.bin"))
    np.ascontiguousarray(fx["hf_VQ_0"]).tofile(os.path.join(d, "V.bin"))
    np.ascontiguousarray(fx["S_k"][:, sym.k_reduced_to_full]).tofile(os.path.join(d, "S.bin"))
    np.ascontiguousarray(fx["hf_G_tau"][-1]).tofile(os.path.join(d, "dm.bin"))
    os.makedirs(os.path.join(d, "df"), exist_ok=True)
    # GW: the fixture's G, integrals and the bosonic transforms on its grid
    g = np.load(os.path.join(ROOT, "tests", "golden", "fixtures", "gw_grid_beta100_lambda1e4.npz"))
    T_tw, T_wt = np.ascontiguousarray(g["T_tw"], np.complex128), np.ascontiguousarray(g["T_wt"], np.complex128)
    with open(os.path.join(d, "gw_sizes.txt"), "w") as f:
        f.write(f"nts {T_tw.shape[0]}\nnw_b {T_tw.shape[1]}\nchunk {int(fx['gw_chunk_size'])}\n")
    T_tw.tofile(os.path.join(d, "T_tw.bin"))
    T_wt.tofile(os.path.join(d, "T_wt.bin"))
    np.ascontiguousarray(fx["gw_VQ_0"]).tofile(os.path.join(d, "gw_V.bin"))
    np.ascontiguousarray(fx["gw_G_tau"]).tofile(os.path.join(d, "gw_G.bin"))
    os.makedirs(os.path.join(d, "df_gw"), exist_ok=True)
    return sym


def build(tmp):
    """HDF5's directory may hold an older C++ runtime than the one the HIP
    runtime needs, so only the HDF5 libraries are exposed (as links in a
    private directory), never the whole directory."""
    exe = os.path.join(tmp, "adapter_check")
    libdir = os.path.join(tmp, "lib")
    os.makedirs(libdir, exist_ok=True)
    src = os.path.join(HDF5, "lib")
    for f in os.listdir(src):
        if f.startswith(("libhdf5.so", "libhdf5_hl.so", "libz.so")):
            os.symlink(os.path.join(src, f), os.path.join(libdir, f))
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "adapter", "stubs"),
           "-I", os.path.join(HDF5, "include"), os.path.join(ROOT, "tests", "adapter", "adapter_check.cpp"),
           "-o", exe, LIB, os.path.join(libdir, "libhdf5_hl.so"), os.path.join(libdir, "libhdf5.so"),
           f"-Wl,-rpath,{os.path.dirname(LIB)}", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


REF_DF = "/root/reference/test/data/HF/df_hf_int"


def run_adapter(d, fx, oracle):
    write_inputs(d, fx, oracle)
    exe = build(d)
    args = [exe, d] + ([REF_DF] if os.path.exists(os.path.join(REF_DF, "meta.h5")) else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    return r.stdout


def test_adapter_on_hf_fixture(tmp_path, fx, oracle):
    d = str(tmp_path)
    run_adapter(d, fx, oracle)
    ref = os.path.join(d, "V_ref.bin")
    if os.path.exists(ref):
        # the reader on the reference's own integral files gives the stored
        # pairs the fixture extraction (h5dump text) gave
        V = np.fromfile(ref, dtype=np.complex128).reshape(fx["hf_VQ_0"].shape)
        assert np.array_equal(V, fx["hf_VQ_0"])


@pytest.mark.gpu
def test_adapter_hf_fixture_on_device(tmp_path, fx, oracle):
    """With a device, custom_hf_kernel solves the fixture: -F == Sigma1, and
    custom_gw_kernel's callback (LU, FP64) reproduces the reference's GW
    Sigma_tau (its bar 1e-6; the FP64 path is within ~1e-10) and the
    oracle."""
    d = str(tmp_path)
    out = run_adapter(d, fx, oracle)
    assert "hf kernel: solved on the device" in out and "gw kernel: solved on the device" in out
    assert "gw kernel (chunked reads): solved on the device" in out
    F = np.fromfile(os.path.join(d, "F.bin"), dtype=np.complex128).reshape(fx["hf_Sigma1"].shape)
    assert max_abs(-F, fx["hf_Sigma1"]) < 1e-8
    S = np.fromfile(os.path.join(d, "Sigma.bin"), dtype=np.complex128).reshape(fx["gw_Sigma_tau"].shape)
    g = np.load(os.path.join(ROOT, "tests", "golden", "fixtures", "gw_grid_beta100_lambda1e4.npz"))
    NQ = int(fx["params_NQ"])
    S_or = oracle.gw_solve(fx["gw_G_tau"], fx["gw_VQ_0"], oracle.symmetry_from_fixture(fx, NQ), g["T_wt"],
                           g["T_tw"], solver="lu")
    err_ref, err_or = max_abs(S, fx["gw_Sigma_tau"]), max_abs(S, S_or)
    print(f"adapter GW: vs Sigma_tau {err_ref:.2e}, vs oracle {err_or:.2e}", flush=True)
    assert err_ref < 1e-6 and err_or < 1e-9
    # cuda_low_cpu_memory: the same Sigma with the pairs read chunk by chunk
    S_chunks = np.fromfile(os.path.join(d, "Sigma_chunks.bin"), dtype=np.complex128).reshape(S.shape)
    assert np.array_equal(S_chunks, S)
