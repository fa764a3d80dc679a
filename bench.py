#!/usr/bin/env python3
"""GW-iteration benchmark on one MI355X.

A step is one full GW iteration (P0 bubble for every q and k, tau -> Omega,
Dyson solve, Omega -> tau, Sigma contraction, Sigma staging) on synthetic
inputs already resident in HBM.  Default workload: nk = 4x4x4, nao = 52,
NQ = 300, ns = 2, beta = 100, sparse-IR Lambda = 1e4 (the largest of the
reference benchmark shapes whose integrals fit one GPU).

Every requested precision mode runs the same iteration on the same inputs;
the FP64 result is the reference for each mode's Sigma error.  Prints one
JSON line (rank 0).  `value` is the algorithmic TFLOP/s of the iteration
(complex multiply-add = 8 FLOP, minimal contraction set) of the fastest mode
whose Sigma stays within 1e-6 max-abs of FP64; every mode's time, rate,
speedup and error are listed under "modes".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GW iteration wall-time (s) + TFLOP/s, mixed vs FP64 on MI355X at Σ err≤1e-6"
PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0}      # TFLOP/s, MI355X dense matrix
# Matrix-core FLOPs executed per algorithmic FLOP (8 per complex multiply-add):
# FP64 forms 3 real products per complex one (Gauss), FP32 4, split-bf16 x3
# 4 real products of 3 bf16 products each.
MFMA_PER_ALG = {"fp64": 0.75, "fp32": 1.0, "bf16x3": 3.0}
HBM_PEAK = 8000.0                                           # GB/s

WORKLOADS = {
    # the reference's small CPU-runnable system
    "gw_nk8_nao26_nq100": dict(mesh=(2, 2, 2), nao=26, NQ=100, ns=1, beta=50.0, lam=1e3),
    # the plain-FP64 / mixed-precision single-GPU system
    "gw_nk64_nao52_nq300": dict(mesh=(4, 4, 4), nao=52, NQ=300, ns=2, beta=100.0, lam=1e4),
    # six independent systems of uneven size in one batched call
    "gw_batch6_nk27": dict(mesh=(3, 3, 3), ns=2, beta=100.0, lam=1e4,
                           systems=[(26, 100), (40, 160), (52, 220), (68, 280), (80, 340), (104, 420)]),
    # the large system: 2.4 TB of stored pairs, so the integrals are generated
    # per batch on the device; a bounded sample of q (one GPU's q share) is
    # timed and the iteration extrapolated (every q costs the same without
    # symmetry)
    "gw_nk216_nao104_nq600": dict(mesh=(6, 6, 6), nao=104, NQ=600, ns=2, beta=100.0, lam=1e4,
                                  integrals="generated", q_share="0/216"),
}
DEFAULT_WORKLOAD = "gw_nk64_nao52_nq300"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gw_flops(nk, nao, NQ, ns, nts, nw_b):
    c = 16.0 * NQ * nao ** 3 + 8.0 * NQ * NQ * nao * nao
    p0 = nk * nk * ns * (nts // 2) * c
    sig = nk * nk * ns * nts * c
    ir = nk * 8.0 * NQ * NQ * (nw_b * (nts - 2) + nts * nw_b)
    dyson = nk * nw_b * (4.0 / 3.0 + 8.0) * NQ ** 3
    return {"p0": p0, "sigma": sig, "ir": ir, "dyson": dyson, "total": p0 + sig + ir + dyson}


def mode_roofline(mode, sites):
    """Dominant GEMM site of one precision mode against that mode's matrix
    peak: algorithmic TFLOP/s, matrix-core TFLOP/s executed, and the
    fraction of the peak the latter is."""
    name, (ms, flops, n) = max(sites.items(), key=lambda kv: kv[1][0])
    alg = flops / (ms * 1e-3) / 1e12
    mfma = alg * MFMA_PER_ALG[mode]
    peak = PEAK[{"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[mode]]
    return {"kernel": f"gemm site {name}", "achieved_algorithmic": round(alg, 2), "achieved_mfma": round(mfma, 2),
            "peak": peak, "unit": "TFLOP/s", "frac_mfma": round(mfma / peak, 4)}


def cpu_baseline(w, G, V_pairs, T_wt, T_tw, sym, npairs=2):
    """Oracle (numpy/OpenBLAS) on a bounded sample of the same iteration:
    q = 0, P0 and Sigma contributions of `npairs` k-points, one Dyson solve.
    Returns the sustained FLOP rate; the iteration time is extrapolated."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gw_oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        cores = int(os.environ.get("OMP_NUM_THREADS", "1"))
    nao, NQ, ns = w["nao"], w["NQ"], w["ns"]
    nts = T_tw.shape[0]
    half = nts // 2
    q = 0
    flops = 0.0
    t0 = time.perf_counter()
    P0 = np.zeros((nts, NQ, NQ), dtype=np.complex128)
    for k in range(npairs):
        k1 = int(sym.k1_from_k2q[k, q])
        V = V_pairs[(k, k1)]
        for s in range(ns):
            Gm = G[nts - 1 - np.arange(half), s, k]
            G1 = G[:half, s, k1]
            P0[:half] += O.p0_pair(Gm, G1, V, -1.0 if ns == 2 else -2.0)
            flops += half * (16.0 * NQ * nao ** 3 + 8.0 * NQ * NQ * nao * nao)
    P0[nts - 1 - np.arange(half)] = P0[:half]
    P0w = O.tau_to_omega(P0, T_wt)
    Pw = O.gw_dyson(P0w, "cholesky")
    Pt = O.omega_to_tau(Pw, T_tw)
    nw_b = T_wt.shape[0]
    flops += 8.0 * NQ * NQ * (nw_b * (nts - 2) + nts * nw_b) + nw_b * (4.0 / 3.0 + 8.0) * NQ ** 3
    for k in range(npairs):
        k1 = int(sym.k2_from_k1q[k, q])
        V = V_pairs[(k, k1)]
        for s in range(ns):
            O.sigma_pair(Pt, G[:, s, k1], V)
            flops += nts * (16.0 * NQ * nao ** 3 + 8.0 * NQ * NQ * nao * nao)
    dt = time.perf_counter() - t0
    return flops / dt, dt, cores, flops


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64,fp32,bf16x3",
                    help="comma-separated modes; fp64 always runs (the error reference)")
    ap.add_argument("--solver", default="cholesky", choices=["cholesky", "lu"])
    ap.add_argument("--integrals", default=None, choices=["host", "device", "generated"],
                    help="host: pinned host memory, streamed per pair batch (overlapped); device: resident in "
                         "HBM; generated: synthesised per batch on the device (default: the workload's, else host)")
    ap.add_argument("--q-share", default=None,
                    help="R/S: time only IBZ q = R, R+S, ... (one GPU's share of a q-sharded iteration) and "
                         "extrapolate the iteration (default: the workload's, else 0/1)")
    ap.add_argument("--no-hf", action="store_true", help="skip the HF build")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--cpu-pairs", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch variants instead of autotuning each plan before the warmup")
    ap.add_argument("--tune-rounds", type=int, default=2)
    ap.add_argument("--graph-modes", default=None,
                    help="batched workload: modes (comma-separated) also timed as one captured graph launch "
                         "(default: every mode, when the integrals are resident in HBM)")
    ap.add_argument("--graph", action="store_true",
                    help="time the iteration as one captured HIP graph replay (needs --integrals device); "
                         "per-site times then come from one profiled eager iteration before the timed steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 1 GiB device-to-device copy kernel (scale for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import __graft_entry__ as ge
    pkg = ge._load_package()
    import green_gpu_amd.synth as synth

    w = WORKLOADS[a.workload]
    a.integrals = a.integrals or w.get("integrals", "host")
    if a.graph and a.integrals != "device":
        sys.exit("--graph replays a captured iteration and needs --integrals device")
    if "systems" in w:
        return run_batch(a, w, pkg, synth)
    sym = pkg.mesh_symmetry_none(w["mesh"])
    nk, nao, NQ, ns = sym.nk, w["nao"], w["NQ"], w["ns"]
    tau, T_wt, T_tw = synth.load_ir(w["beta"], w["lam"])
    nts, nw_b = tau.size, T_wt.shape[0]
    q_rank, q_size = (int(x) for x in (a.q_share or w.get("q_share", "0/1")).split("/"))
    if not (0 <= q_rank < q_size):
        sys.exit("--q-share R/S needs 0 <= R < S")
    q_run = len(range(q_rank, sym.inq, q_size))
    log(f"workload {a.workload}: nk={nk} nao={nao} NQ={NQ} ns={ns} nts={nts} nw_b={nw_b}; "
        f"q share {q_rank}/{q_size}: {q_run} of {sym.inq} q; integrals {a.integrals}")
    t = time.perf_counter()
    G, _ = synth.make_green(nk, nao, ns, w["beta"], tau, 1000)
    log(f"G generated in {time.perf_counter() - t:.1f}s")
    Gd = pkg.DeviceArray.from_numpy(G)
    seed, sigma = 1017, 0.1
    fl = {k: v * q_run / sym.inq for k, v in gw_flops(nk, nao, NQ, ns, nts, nw_b).items()}
    modes = [m.strip() for m in a.precision.split(",")]
    if "fp64" not in modes:
        modes = ["fp64"] + modes
    runs, sigmas, p0s = {}, {}, {}
    for mode in modes:
        kern = pkg.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, precision=mode, solver=a.solver, integrals=a.integrals)
        kern.synth_integrals(seed, sigma)
        kern.set_q_share(q_rank, q_size)
        Sd = pkg.DeviceArray.empty_like(Gd)
        pkg.synchronize()
        log(f"[{mode}] inputs resident in HBM")
        tune = None
        if not a.no_autotune:
            rep = kern.autotune_device(Gd.ptr, rounds=a.tune_rounds)
            tune = {"wall_ms": round(rep["wall_ms"], 1),
                    "chosen": {k: v["chosen"] for k, v in rep["sites"].items()},
                    "candidate_ms": {k: {c: round(t, 3) for c, t in v["ms"].items()}
                                     for k, v in rep["sites"].items()}}
            log(f"[{mode}] autotuned in {rep['wall_ms'] / 1e3:.2f}s: {tune['chosen']}")
        for i in range(a.warmup):
            t = time.perf_counter()
            kern.solve_device(Gd.ptr, Sd.ptr)
            pkg.synchronize()
            log(f"[{mode}] warmup step {i}: {time.perf_counter() - t:.2f}s")
        kern.set_profiling(True)
        if a.graph:
            # per-site times from one profiled eager iteration, then the
            # graph is captured (untimed) and replayed in the timed steps
            kern.solve_device(Gd.ptr, Sd.ptr)
            sites = kern.site_stats()
            kern.set_profiling(False)
            kern.set_graph(True)
            t = time.perf_counter()
            kern.solve_device(Gd.ptr, Sd.ptr)
            pkg.synchronize()
            log(f"[{mode}] graph captured and replayed once: {time.perf_counter() - t:.2f}s")
        pkg.synchronize()
        t0 = time.perf_counter()
        for i in range(a.steps):
            kern.solve_device(Gd.ptr, Sd.ptr)
            pkg.synchronize()
            log(f"[{mode}] step {i} done")
        dt = (time.perf_counter() - t0) / a.steps
        kern.set_profiling(False)
        if not a.graph:
            sites = kern.site_stats()
        runs[mode] = {"s": dt, "sites": sites, "phases": kern.timings(),
                      "streamed": kern.streamed_bytes(), "generated": kern.generated_bytes(), "autotune": tune,
                      "staging_ms": kern.staging_ms()}
        sigmas[mode] = Sd.numpy()
        Sd.free()
        # P0(tau) of the first IBZ q, for the per-mode P0 error
        p0s[mode] = kern.polarization(G, 0, 0)
        kern.close()
        del kern
    S64 = sigmas["fp64"]
    scale = float(np.abs(S64).max())
    for mode in modes:
        r = runs[mode]
        err = float(np.abs(sigmas[mode] - S64).max())
        r["sigma_max_abs_err"] = err
        r["sigma_rel_err"] = err / scale
        r["tflops"] = fl["total"] / r["s"] / 1e12
        p0_scale = float(np.abs(p0s["fp64"]).max())
        r["p0_max_abs_err"] = float(np.abs(p0s[mode] - p0s["fp64"]).max())
        r["p0_rel_err"] = r["p0_max_abs_err"] / p0_scale
        r["speedup_vs_fp64"] = runs["fp64"]["s"] / r["s"]
        log(f"[{mode}] {r['s']:.2f}s  {r['tflops']:.1f} TFLOP/s  Sigma max-abs err {err:.2e} (|Sigma| max {scale:.2e})")
    if a.calibrate_copy:
        src = pkg.DeviceArray.from_numpy(np.ones(1 << 27))            # 1 GiB
        dst = pkg.DeviceArray.empty_like(src)
        dst.copy_from(src)
        log("calibration copy: 1073741824 bytes read, 1073741824 bytes written")
        src.free()
        dst.free()

    # HF build (exchange K + direct J) on the same system in every mode, the
    # exchange products in the mode's arithmetic and K's partial sums, J and
    # the Ewald-free Fock matrix in FP64: D = -G(beta), synthetic integrals
    # with the same seed
    hf = None
    if not a.no_hf:
        D = -np.ascontiguousarray(G[-1])                                   # (ns, ink, nao, nao)
        hf_fl = float(nk) * nk * ns * 16.0 * NQ * nao ** 3
        hf_modes, F64 = {}, None
        for mode in modes:
            hk = pkg.HfKernel(nao, ns, NQ, sym, integrals="generated" if a.integrals == "generated" else "device",
                              precision=mode)
            hk.synth_integrals(seed, sigma)
            if q_run == sym.inq:
                hk.solve(D)                                                # warm-up
            t = time.perf_counter()
            F = hk.solve(D)
            hf_s = time.perf_counter() - t
            hk.close()
            # E_HF-like scalar, accumulated in FP64 on the host (the reference defines none)
            e_hf = float(0.5 * np.real(np.einsum("skab,skba->", D, F)) / nk)
            if mode == "fp64":
                F64, e64 = F, e_hf
            kerr = float(np.abs(F - F64).max())
            hf_modes[mode] = {"seconds": round(hf_s, 4), "exchange_tflops": round(hf_fl / hf_s / 1e12, 2),
                              "e_hf_like": e_hf, "k_max_abs_err": kerr,
                              # relative to max |F|: F - F_fp64 is K's error (J is FP64 in every mode)
                              "k_rel_err": kerr / float(np.abs(F64).max()),
                              "e_hf_abs_err": abs(e_hf - e64), "e_hf_rel_err": abs(e_hf - e64) / abs(e64)}
            log(f"[hf {mode}] {hf_s:.3f}s  {hf_fl / hf_s / 1e12:.1f} TFLOP/s (exchange FLOPs)  "
                f"K err {kerr:.2e}  E_HF err {abs(e_hf - e64):.2e}")
        h64 = hf_modes["fp64"]
        hf = {"seconds": h64["seconds"], "exchange_tflops": h64["exchange_tflops"], "flop": hf_fl,
              "e_hf_like": h64["e_hf_like"], "precision": "fp64", "includes": "K + J, host-to-host",
              "modes": hf_modes}

    # headline: the fastest mode whose Sigma stays within 1e-6 of FP64
    ok = [m for m in modes if runs[m]["sigma_max_abs_err"] <= 1e-6]
    head = min(ok, key=lambda m: runs[m]["s"])
    r = runs[head]
    dtype = {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[head]
    dom_name, (dom_ms, dom_fl, dom_n) = max(r["sites"].items(), key=lambda kv: kv[1][0])
    achieved = dom_fl / (dom_ms * 1e-3) / 1e12
    # HBM bytes per launch of that kernel from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes of this workload and precision (null when not profiled)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(a.workload, {}).get(head, {}).get(dom_name)
    roofline = {"bound": "mfma", "kernel": f"gemm site {dom_name}", "achieved": round(achieved, 3),
                "peak": PEAK[dtype], "unit": "TFLOP/s", "frac": round(achieved / PEAK[dtype], 4),
                "traffic": None if traffic is None else round(traffic), "launches": dom_n,
                "avg_launch_ms": round(dom_ms / dom_n, 4)}
    out = {
        "metric": METRIC, "value": round(r["tflops"], 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(r["s"] * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded random H(k) -> G(tau); Gaussian V, sigma 0.1)",
        "config": {"workload": a.workload, "nk": nk, "nao": nao, "NQ": NQ, "ns": ns, "nts": nts,
                   "nw_b": nw_b, "beta": w["beta"], "ir_lambda": w["lam"], "precision": head,
                   "solver": a.solver, "integrals": a.integrals, "q_share": f"{q_rank}/{q_size}",
                   "graph": bool(a.graph)},
        "gw_iteration_s": round(r["s"], 3),
        "gw_flop_per_iteration": fl["total"],
        "modes": {m: {"gw_iteration_s": round(runs[m]["s"], 3), "tflops": round(runs[m]["tflops"], 2),
                      "speedup_vs_fp64": round(runs[m]["speedup_vs_fp64"], 3),
                      "sigma_max_abs_err": runs[m]["sigma_max_abs_err"],
                      "sigma_rel_err": runs[m]["sigma_rel_err"],
                      "p0_max_abs_err": runs[m]["p0_max_abs_err"],
                      "p0_rel_err": runs[m]["p0_rel_err"],
                      "phases_ms": {k: round(v, 1) for k, v in runs[m]["phases"].items()},
                      "sites_ms": {k: round(v[0], 1) for k, v in runs[m]["sites"].items()},
                      "autotune": runs[m]["autotune"],
                      "roofline": mode_roofline(m, runs[m]["sites"])} for m in modes},
        "sigma_abs_max": scale,
        "integrals": {"residence": a.integrals,
                      "h2d_gb_per_iteration": round(r["streamed"] / 1e9, 2),
                      "h2d_gb_per_s_average": round(r["streamed"] / 1e9 / r["s"], 2),
                      # rate of the copies (or generation) while the copy stream runs them
                      "staging_s": round(r["staging_ms"] / 1e3, 3),
                      "staging_gb_per_s": (round((r["streamed"] + r["generated"]) / 1e9 / (r["staging_ms"] / 1e3), 2)
                                           if r["staging_ms"] > 0 else None),
                      "generated_gb_per_iteration": round(r["generated"] / 1e9, 2)},
        "phases_ms": {k: round(v, 1) for k, v in r["phases"].items()},
        "sites": {k: {"ms": round(v[0], 1), "tflops": round(v[1] / (v[0] * 1e-3) / 1e12, 2), "launches": v[2]}
                  for k, v in r["sites"].items()},
        "roofline": roofline,
        "hf": hf,
    }
    if q_run < sym.inq:
        out["q_share"] = {"rank": q_rank, "size": q_size, "q_timed": q_run, "inq": sym.inq,
                          "extrapolated_iteration_s": {m: round(runs[m]["s"] * sym.inq / q_run, 1) for m in modes},
                          "note": "step = this share's q-points; all q cost the same without symmetry"}
    if not a.no_cpu and a.workload == DEFAULT_WORKLOAD:
        pairs = {}
        for k in range(a.cpu_pairs):
            for k1 in (int(sym.k1_from_k2q[k, 0]), int(sym.k2_from_k1q[k, 0])):
                lo, hi = min(k, k1), max(k, k1)
                p = hi * (hi + 1) // 2 + lo
                S = synth.make_integrals(nk, NQ, nao, seed, sigma, pairs=[p])[0]
                pairs[(k, k1)] = S if k >= k1 else np.conj(np.swapaxes(S, -1, -2))
        rate, cdt, cores, cfl = cpu_baseline(w, G, pairs, T_wt, T_tw, sym, a.cpu_pairs)
        out["cpu_baseline"] = {
            "value": round(rate / 1e12, 5), "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": f"oracle (numpy/OpenBLAS, FP64) on q=0: P0 and Sigma of {a.cpu_pairs} k-points + one "
                      f"Dyson solve ({cfl:.3g} FLOP in {cdt:.1f}s); extrapolated iteration "
                      f"{fl['total'] / rate:.0f}s"}
    print(json.dumps(out))


def run_batch(a, w, pkg, synth):
    """Several independent systems per step through ggpu_gw_iterate_many (one
    stream per system), against the same systems run one after another."""
    sym = pkg.mesh_symmetry_none(w["mesh"])
    nk, ns = sym.nk, w["ns"]
    tau, T_wt, T_tw = synth.load_ir(w["beta"], w["lam"])
    nts, nw_b = tau.size, T_wt.shape[0]
    modes = [m.strip() for m in a.precision.split(",")]
    if "fp64" not in modes:
        modes = ["fp64"] + modes
    Gd = []
    for i, (nao, NQ) in enumerate(w["systems"]):
        G, _ = synth.make_green(nk, nao, ns, w["beta"], tau, 2000 + i)
        Gd.append(pkg.DeviceArray.from_numpy(G))
    fl = sum(gw_flops(nk, nao, NQ, ns, nts, nw_b)["total"] for nao, NQ in w["systems"])
    log(f"workload {a.workload}: {len(w['systems'])} systems, nk={nk}, {fl:.3g} FLOP per step")
    res, sig = {}, {}
    for mode in modes:
        kerns, Sd = [], []
        for i, (nao, NQ) in enumerate(w["systems"]):
            k = pkg.GwKernel(nao, ns, NQ, T_wt, T_tw, sym, precision=mode, solver=a.solver,
                             work_bytes=16 << 30, integrals=a.integrals)
            k.synth_integrals(3000 + i, 0.1)
            kerns.append(k)
            Sd.append(pkg.DeviceArray.empty_like(Gd[i]))
        gp, sp = [g.ptr for g in Gd], [x.ptr for x in Sd]
        tune_ms = 0.0
        if not a.no_autotune:
            for i, k in enumerate(kerns):
                tune_ms += k.autotune_device(gp[i], rounds=a.tune_rounds)["wall_ms"]
            log(f"[{mode}] {len(kerns)} plans autotuned in {tune_ms / 1e3:.2f}s")
        for _ in range(a.warmup):
            pkg.iterate_many(kerns, gp, sp)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            pkg.iterate_many(kerns, gp, sp)
            log(f"[{mode}] batched step done")
        batched = (time.perf_counter() - t0) / a.steps
        t0 = time.perf_counter()
        for i, k in enumerate(kerns):
            k.solve_device(gp[i], sp[i])
            log(f"[{mode}] system {i} alone done")
        sequential = time.perf_counter() - t0
        log(f"[{mode}] batched {batched:.2f}s, one after another {sequential:.2f}s")
        res[mode] = {"batched_s": batched, "sequential_s": sequential, "tflops": fl / batched / 1e12,
                     "autotune_ms": tune_ms}
        sig[mode] = [x.numpy() for x in Sd]
        if a.integrals == "device" and (a.graph_modes is None or mode in a.graph_modes.split(",")):
            # all systems as one captured HIP graph (one launch per step)
            b = pkg.GwBatch(kerns, gp, sp)
            t = time.perf_counter()
            b.run()
            log(f"[{mode}] batch graph captured and replayed once: {time.perf_counter() - t:.2f}s")
            t0 = time.perf_counter()
            for _ in range(a.steps):
                b.run()
                log(f"[{mode}] graph step done")
            res[mode]["graph_s"] = (time.perf_counter() - t0) / a.steps
            res[mode]["graph_equals_eager"] = all(np.array_equal(x.numpy(), y) for x, y in zip(Sd, sig[mode]))
            b.close()
            log(f"[{mode}] one graph launch {res[mode]['graph_s']:.2f}s, "
                f"equal to the eager result: {res[mode]['graph_equals_eager']}")
        for x in Sd:
            x.free()
        for k in kerns:
            k.close()
    for mode in modes:
        err = 0.0
        for i in range(len(Gd)):
            err = max(err, float(np.abs(sig[mode][i] - sig["fp64"][i]).max()))
        res[mode]["sigma_max_abs_err"] = err
    ok = [m for m in modes if res[m]["sigma_max_abs_err"] <= 1e-6]
    head = min(ok, key=lambda m: res[m]["batched_s"])
    r = res[head]
    out = {
        "metric": METRIC, "value": round(r["tflops"], 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(r["batched_s"] * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[head],
        "data": "synthetic (seeded random H(k) -> G(tau); Gaussian V, sigma 0.1)",
        "config": {"workload": a.workload, "nk": nk, "ns": ns, "systems": w["systems"], "nts": nts, "nw_b": nw_b,
                   "beta": w["beta"], "ir_lambda": w["lam"], "precision": head, "solver": a.solver,
                   "integrals": a.integrals},
        "gw_flop_per_step": fl,
        "modes": {m: {"batched_s": round(v["batched_s"], 3), "sequential_s": round(v["sequential_s"], 3),
                      "batching_speedup": round(v["sequential_s"] / v["batched_s"], 3),
                      "tflops": round(v["tflops"], 2),
                      "speedup_vs_fp64": round(res["fp64"]["batched_s"] / v["batched_s"], 3),
                      "sigma_max_abs_err": v["sigma_max_abs_err"],
                      "autotune_ms": round(v["autotune_ms"], 1),
                      **({"graph_s": round(v["graph_s"], 3), "graph_equals_eager": v["graph_equals_eager"]}
                         if "graph_s" in v else {})} for m, v in res.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
