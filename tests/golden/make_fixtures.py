#!/usr/bin/env python3
"""Convert the reference's HDF5 regression fixtures into small ``.npy`` files.

The reference's Catch2 regression (test/cu_solver_test.cpp:34-96, :226-288,
:343-412) reads these HDF5 files with green-h5pp.  This container has no
h5py, so the datasets are read through the ``h5dump`` text dump of the
HDF5 command-line tools (present only in the build container, at
``/opt/conda/bin/h5dump``) and parsed here.  The GPU box never runs this
script; it only loads the committed ``.npy`` outputs.

What is extracted (all paths relative to ``test/data`` of the reference):

* ``HF/input.h5`` (byte-identical to ``GW/input.h5``): dimensions, Madelung
  constant, overlap ``S-k`` and ``Fock-k`` on the full k-mesh, and every
  symmetry table green-symmetry reads (k/q maps, stars, AO and aux-basis
  representation matrices, k-pair conj/trans/irreducible lists).
* ``HF/data.h5``: ``G_tau`` and the expected HF ``result/Sigma1``.
* ``HF/df_hf_int``: the stored k-pair integral chunk ``VQ_0.h5:/0`` and
  ``meta.h5:chunk_size``.
* ``GW/data.h5`` + ``GW/df_int``: ``G_tau``, expected ``result/Sigma_tau``
  and the integral chunk used by the GW regression.
* the two-component (X2C) system of ``HF_X2C`` / ``GW_X2C`` (one k-point,
  nao = 10, nso = 20, NQ = 154): the same tables, ``Sigma1`` and
  ``Sigma_tau`` (test/cu_solver_test.cpp:317-336), into
  ``ref_fixtures_x2c.npz``.

Run:  python tests/golden/make_fixtures.py  (writes tests/golden/fixtures/)
"""
import os
import re
import subprocess
import sys

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"
REF = "/root/reference/test/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")

_num = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|[-+]?inf|nan")


def h5read(path, dset):
    """Return (array, is_compound) for dataset ``dset`` of HDF5 file ``path``."""
    txt = subprocess.run([H5DUMP, "-d", dset, "-y", "-w", "0", "-m", "%.17g", path],
                         check=True, capture_output=True, text=True).stdout
    head, _, body = txt.partition("DATA {")
    body = body.split("ATTRIBUTE")[0]
    compound = "H5T_COMPOUND" in head
    is_float = "IEEE_F" in head
    m = re.search(r"DATASPACE\s+(SCALAR|SIMPLE \{ \( ([^)]*) \))", head)
    shape = () if m.group(1) == "SCALAR" else tuple(int(x) for x in m.group(2).split(","))
    vals = [float(x) for x in _num.findall(body)]
    n = int(np.prod(shape)) if shape else 1
    if compound:
        arr = np.array(vals, dtype=np.float64)
        assert arr.size == 2 * n, (path, dset, arr.size, n)
        arr = arr.view(np.complex128).reshape(shape)
    else:
        arr = np.array(vals, dtype=np.float64 if is_float else np.int64)
        assert arr.size == n, (path, dset, arr.size, n)
        arr = arr.reshape(shape)
    return arr


def complex_attr(a):
    """Datasets tagged ``__complex__`` store (re, im) in a trailing axis of 2."""
    assert a.shape[-1] == 2
    return np.ascontiguousarray(a).view(np.complex128)[..., 0]


def complex_interleaved(a):
    """Integral chunks store complex numbers as interleaved (re, im) doubles in
    the last axis, read by the reference through a ``double*`` cast of a
    complex128 buffer (df_integral_t.h:136-144), so the last axis halves."""
    assert a.shape[-1] % 2 == 0
    return np.ascontiguousarray(a).view(np.complex128)


def read_system(inp):
    """Dimensions, overlap/Fock matrices and every symmetry table of an
    ``input.h5``."""
    out = {}
    for key in ("nao", "nso", "NQ", "ns", "nk"):
        out["params_" + key] = h5read(inp, "/params/" + key)
    out["madelung"] = h5read(inp, "/HF/madelung")
    out["S_k"] = complex_attr(h5read(inp, "/HF/S-k"))
    out["Fock_k"] = complex_attr(h5read(inp, "/HF/Fock-k"))
    for kq in ("k", "q"):
        for name in ("bz2ibz", "ibz2bz", "tr_conj", "mesh_scaled", "n_stars"):
            out[f"sym_{kq}_{name}"] = h5read(inp, f"/symmetry/{kq}/{name}")
        nst = int(out[f"sym_{kq}_n_stars"])
        for i in range(nst):
            out[f"sym_{kq}_star_{i}"] = h5read(inp, f"/symmetry/{kq}/stars/{i}")
    out["sym_k_ink"] = h5read(inp, "/symmetry/k/ink")
    out["sym_q_inq"] = h5read(inp, "/symmetry/q/inq")
    out["sym_k_transform_ao"] = h5read(inp, "/symmetry/k/k_sym_transform_ao")
    out["sym_q_transform_p0"] = h5read(inp, "/symmetry/q/k_sym_transform_p0")
    for name in ("conj_pairs_list", "trans_pairs_list", "kpair_irre_list", "kpair_idx",
                 "num_kpair_stored"):
        out["pairs_" + name] = h5read(inp, "/symmetry/pairs/" + name)
    return out


def read_run(out, prefix, data, integrals, result):
    out[prefix + "_G_tau"] = h5read(data, "/G_tau")
    out[prefix + "_" + result.split("/")[-1]] = h5read(data, result)
    out[prefix + "_chunk_size"] = h5read(os.path.join(integrals, "meta.h5"), "/chunk_size")
    out[prefix + "_VQ_0"] = complex_interleaved(h5read(os.path.join(integrals, "VQ_0.h5"), "/0"))


def write(name, out):
    np.savez_compressed(os.path.join(OUT, name), **out)
    for k, v in out.items():
        print(f"{k:28s} {str(v.dtype):10s} {v.shape}")


def main():
    os.makedirs(OUT, exist_ok=True)
    # scalar (ns = 2) system: HF/input.h5 and GW/input.h5 are byte-identical
    out = read_system(os.path.join(REF, "HF/input.h5"))
    read_run(out, "hf", os.path.join(REF, "HF/data.h5"), os.path.join(REF, "HF/df_hf_int"), "/result/Sigma1")
    read_run(out, "gw", os.path.join(REF, "GW/data.h5"), os.path.join(REF, "GW/df_int"), "/result/Sigma_tau")
    write("ref_fixtures.npz", out)
    # two-component (X2C) system: HF_X2C and GW_X2C share input.h5 and integrals
    x2c = read_system(os.path.join(REF, "HF_X2C/input.h5"))
    read_run(x2c, "hf", os.path.join(REF, "HF_X2C/data.h5"), os.path.join(REF, "HF_X2C/df_hf_int"), "/result/Sigma1")
    read_run(x2c, "gw", os.path.join(REF, "GW_X2C/data.h5"), os.path.join(REF, "GW_X2C/df_hf_int"),
             "/result/Sigma_tau")
    write("ref_fixtures_x2c.npz", x2c)


if __name__ == "__main__":
    sys.exit(main())
