"""Benchmark: chem cells/s per dt_react on one MI355X (learned map + BDF fallback, and plain GPU BDF).

Workload (one "step" = one ReactorBase::react call over the whole box, inputs resident in HBM):
drm19 CH4/air flame-like box, 64^3 = 262144 cells of uneven stiffness (30 % cold unburnt at
300 K, 20 % flame front with radicals, 50 % burnt products), Gaussian external sources at 5 %
of |rhoY|/dt and |rhoH|/dt, dt_react = 5e-6 s, 1 atm, rtol = atol = 1e-10, cvode.max_order = 4
(the value the reference's production inputs use), FP64 throughout. The box is restored from a
pristine device copy before each react call, so every step integrates the same states.

With a trained network (pelelmex_amd/surrogate_drm19.npz) the metric value is the learned map +
FP64 trust check + BDF fallback path's throughput; the plain GPU BDF is timed on the same box in
the same run (bdf_value) and the surrogate result is compared with it on every cell.

Prints one JSON line. `roofline` is the integration kernel's algorithmic FP64 rate
(pelelmex_amd.roofline counts, kernel time from HIP events on the launch stream) against the
MI355X FP64 peak; `cpu_baseline` is the CPU oracle (OpenMP over cells, all host threads) on a
seeded random sample of the same box.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "chem cells/s per dt_react: surrogate+fallback vs GPU BDF at max|ΔT|/T≤1e-4"
ROOT = os.path.dirname(os.path.abspath(__file__))
PROFILE_DIR = os.path.join(ROOT, "profiles")


def kernel_source_digest():
    """sha256 over the device sources the react kernel is built from (ties a profile to a build)."""
    import hashlib
    csrc = os.path.join(ROOT, "pelelmex_amd", "csrc")
    h = hashlib.sha256()
    for dirpath, _, files in sorted(os.walk(csrc)):
        for f in sorted(files):
            if f.endswith((".h", ".hip")):
                p = os.path.join(dirpath, f)
                h.update(os.path.relpath(p, csrc).encode())
                h.update(open(p, "rb").read())
    return h.hexdigest()[:16]


def profiled_traffic(workload_key):
    """HBM bytes per launch of the react kernel from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/*.md) of this workload and these exact kernel sources; None otherwise."""
    if not os.path.isdir(PROFILE_DIR):
        return None, None
    digest = kernel_source_digest()
    for f in sorted(os.listdir(PROFILE_DIR)):
        if not f.endswith(".md"):
            continue
        kv = {}
        for line in open(os.path.join(PROFILE_DIR, f)):
            if ": " in line and not line.startswith(("|", "#", "`")):
                k, v = line.rstrip("\n").split(": ", 1)
                kv[k.strip()] = v.strip()
        if kv.get("workload_key") == workload_key and kv.get("source_digest") == digest and \
                "traffic_bytes_per_launch" in kv:
            return float(kv["traffic_bytes_per_launch"]), f
    return None, None


def copy_probe(mb, dev):
    """One plain device copy of `mb` MiB (its byte counters calibrate a profile's FETCH / WRITE sizes)."""
    if not mb:
        return
    import torch
    a = torch.empty(mb * (1 << 20) // 8, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def host_cpu():
    """CPU model, logical CPUs and OMP_NUM_THREADS of the host the baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1, os.environ.get("OMP_NUM_THREADS")


def cpu_baseline(box, sample, seed=7, what="the same 64^3 box", max_order=4):
    """CPU oracle on `sample` random cells of the box. Returns (cpu_baseline object, (cell
    indices, oracle result)) - the result also checks the GPU's on those cells."""
    from oracle import pyoracle as po
    from pelelmex_amd import workloads as W
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(box["T"].size, sample, replace=False))
    sub = W.subsample(box, idx)
    m = po.Mech(box["mech"])
    model, ncpu, omp = host_cpu()
    nthreads = int(omp or 0) or ncpu
    t0 = time.perf_counter()
    out = m.react(sub["rY"], sub["rYsrc"], sub["T"], sub["rH"], sub["rHsrc"], sub["mask"], sub["dt"],
                  max_order=max_order, nthreads=nthreads)
    dt = time.perf_counter() - t0
    base = {"value": sample / dt, "unit": "cells/s", "cores": nthreads, "kind": "port",
            "cpu_model": model, "nproc": ncpu, "OMP_NUM_THREADS": omp,
            "sample": f"{sample} cells (seeded random order) of {what}, CPU oracle BDF (C, -O2, OpenMP "
                      f"dynamic over cells on {nthreads} threads), {dt:.1f} s, failures "
                      f"{int(np.sum(out['status'] != 0))}"}
    return base, (idx, out)


def gpu_vs_oracle(gpu_state, ref):
    """Parity of a GPU react result (device tensors rY [N, ...], T, FC over the whole box) with
    the CPU oracle's result on the cells it integrated: bit-exactness and the worst errors."""
    idx, out = ref
    N = out["rY"].shape[0]
    gY = np.ascontiguousarray(gpu_state["rY"].reshape(N, -1).cpu().numpy()[:, idx])  # summed like the oracle's
    gT = gpu_state["T"].reshape(-1).cpu().numpy()[idx]
    gF = gpu_state["FC"].reshape(-1).cpu().numpy()[idx]
    act = out["status"] == 0
    Yg = gY / gY.sum(0)
    Yo = out["rY"] / out["rY"].sum(0)
    return {"cells": int(idx.size), "reference": "CPU oracle (same CVODE-semantics BDF, rtol = atol = 1e-10)",
            "bit_exact": bool(np.array_equal(gY, out["rY"]) and np.array_equal(gT, out["T"]) and
                              np.array_equal(gF, out["FC"])),
            "max_dT_rel": float((np.abs(gT - out["T"]) / out["T"])[act].max()),
            "max_dY": float(np.abs(Yg - Yo)[:, act].max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=32, help="hipbdf.lanes_per_cell (0: autotune)")
    ap.add_argument("--max-order", type=int, default=4)
    ap.add_argument("--cpu-sample", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true", help="measure the plain GPU BDF only")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload only: skip the other workloads measured after it")
    ap.add_argument("--surrogate", default=None, help="trained network (.npz); default: the packaged one")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one plain device copy of this size (calibrates byte counters in a profile)")
    ap.add_argument("--waves", type=int, default=0,
                    help="dodecane_128 / syn53_128: hipbdf.waves_per_simd (with --block; 0: the autotuner chooses)")
    ap.add_argument("--block", type=int, default=0, help="dodecane_128 / syn53_128: hipbdf.block (with --waves)")
    ap.add_argument("--workload", default="drm19_64",
                    choices=["drm19_64", "dodecane_128", "syn53_128", "amr_mixed", "amr_dodecane", "h2_32"],
                    help="dodecane_128: plain GPU BDF on n-dodecane, 128^3 cells at 60 atm in eight "
                         "262144-cell launches (one step = all eight); syn53_128: the same on the synthetic "
                         "53-species mechanism; amr_mixed / amr_dodecane: learned map + batched fallback over "
                         "mixed AMR boxes in one hipGraph (drm19 / n-dodecane, about 1.5 M cells on three levels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X; the library has no multi-GPU path (each GPU would run its own "
              "boxes independently).", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("dodecane_128", "syn53_128"):
        print(json.dumps(bench_chunks(args, "dodecane_lu" if args.workload == "dodecane_128" else "syn53")), flush=True)
        return
    if args.workload in ("amr_mixed", "amr_dodecane"):
        print(json.dumps(bench_amr(args, "dodecane_lu" if args.workload == "amr_dodecane" else "drm19")), flush=True)
        return
    if args.workload == "h2_32":
        print(json.dumps(bench_h2(args)), flush=True)
        return

    import torch
    from pelelmex_amd import roofline, surrogate, workloads
    from pelelmex_amd.reactor import Reactor, box_tensors

    box = workloads.ch4_air_flame_box(nx=args.nx, ny=args.nx, nz=args.nx)
    ncells = box["T"].size
    dev = torch.device("cuda:0")
    t = box_tensors(box, dev)
    pristine = {k: t[k].clone() for k in ("rY", "T", "rH")}
    keys = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": args.max_order,
            "cvode.max_substeps": 10000, "cvode.solve_type": "denseAJ_direct"}
    if args.lanes:
        keys["hipbdf.lanes_per_cell"] = args.lanes
    else:
        keys["hipbdf.autotune"] = 1

    def measure(r):
        """W untimed + K timed react calls on the pristine box; (ms per call, last stats, state)."""
        def step():
            for k in pristine:
                t[k].copy_(pristine[k])
            nfail, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"],
                                  0.0)
            if nfail:
                raise RuntimeError(f"{nfail} cells failed to integrate")
            return r.stats()
        for _ in range(args.warmup):
            step()
        ref_T = t["T"].clone() if args.warmup > 0 else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kstats = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        det = bool(torch.equal(ref_T, t["T"])) if ref_T is not None else None
        return ms, kstats, det, {k: t[k].clone() for k in ("rY", "T", "FC")}

    copy_probe(args.copy_probe_mb, dev)

    r = Reactor.create("ReactorCvode", "drm19", keys)
    r.init(2, 1)
    ms_bdf, kstats, deterministic, bdf_state = measure(r)
    r.close()
    bdf_value = ncells / (ms_bdf / 1e3)

    kernel_ms = float(np.mean([s["kernel_ms"] for s in kstats]))
    st = kstats[-1]
    F = roofline.flops("drm19", st)
    achieved = F / (kernel_ms / 1e3) / 1e12
    counts = roofline.static_counts("drm19")
    workload_key = f"drm19 {args.nx}^3 lanes {st['lanes_per_cell']} max_order {args.max_order}"
    traffic, traffic_src = profiled_traffic(workload_key)

    # learned map + BDF fallback on the same box
    sur = None
    wpath = args.surrogate or surrogate.WEIGHTS
    if not args.no_surrogate and os.path.exists(wpath):
        model = surrogate.load(wpath)
        skeys = dict(keys)
        if "trust_rtol" in model:
            skeys["surrogate.trust_rtol"] = float(model["trust_rtol"])
            skeys["surrogate.trust_atol"] = float(model["trust_atol"])
        rs = Reactor.create("ReactorCvode", "drm19", skeys)
        rs.init(2, 1)
        surrogate.attach(rs, model)
        ms_sur, sstats, _, sur_state = measure(rs)
        rs.close()
        # the same path with every increment zero (source step alone, no network pass): the
        # predictor the network is measured against, under the same trust check and tolerances
        rz = Reactor.create("ReactorCvode", "drm19", dict(skeys, **{"surrogate.zero_increment": 1}))
        rz.init(2, 1)
        surrogate.attach(rz, model)
        ms_zero, zstats, _, zero_state = measure(rz)
        rz.close()
        z2 = zstats[-1]
        bdf_fc = bdf_state["FC"].reshape(-1).double()

        def work_removed(state):  # share of the BDF's right-hand-side evaluations the accepted cells carry
            acc = state["FC"].reshape(-1) == 1.0
            return float(bdf_fc[acc].sum() / bdf_fc.sum())
        s2 = sstats[-1]
        N = bdf_state["rY"].shape[0]
        dT = ((sur_state["T"] - bdf_state["T"]).abs() / bdf_state["T"]).max().item()
        Ys = sur_state["rY"].reshape(N, -1) / sur_state["rY"].reshape(N, -1).sum(0)
        Yb = bdf_state["rY"].reshape(N, -1) / bdf_state["rY"].reshape(N, -1).sum(0)
        sur = {"value": ncells / (ms_sur / 1e3), "ms_per_step": ms_sur,
               "accepted": s2["sur_accepted"] / ncells, "rejected_to_bdf": s2["sur_rejected"] / ncells,
               "trust_rtol": skeys.get("surrogate.trust_rtol"), "trust_atol": skeys.get("surrogate.trust_atol"),
               "network_and_trust_ms": s2["surrogate_ms"], "fallback_bdf_ms": s2["kernel_ms"],
               "audit_ms": s2["audit_ms"], "audited": s2["sur_audited"], "audit_max_dT_rel": s2["audit_max_dT_rel"],
               "audit_max_dY": s2["audit_max_dY"],
               "max_dT_rel_vs_bdf_all_cells": dT,
               "max_dY_vs_bdf_all_cells": (Ys - Yb).abs().max().item(),
               "speedup_vs_bdf": bdf_value and (ncells / (ms_sur / 1e3)) / bdf_value,
               "bdf_work_removed": work_removed(sur_state),
               "weights": os.path.relpath(wpath, ROOT),
               "zero_increment_predictor": {
                   "value": ncells / (ms_zero / 1e3), "ms_per_step": ms_zero,
                   "accepted": z2["sur_accepted"] / ncells, "bdf_work_removed": work_removed(zero_state),
                   "max_dT_rel_vs_bdf_all_cells":
                       ((zero_state["T"] - bdf_state["T"]).abs() / bdf_state["T"]).max().item(),
                   "speedup_vs_bdf": (ncells / (ms_zero / 1e3)) / bdf_value}}

    value = sur["value"] if sur else bdf_value
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sur["ms_per_step"] if sur else ms_bdf,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"drm19 CH4/air flame-like box {args.nx}^3 = {ncells} cells, dt_react 5e-6 s, "
                               "1 atm, 5% Gaussian sources",
                   "mechanism": "drm19", "cells": ncells, "dt_react": box["dt"], "rtol": 1e-10, "atol": 1e-10,
                   "max_order": args.max_order, "lanes_per_cell": st["lanes_per_cell"], "seed": 2002,
                   "cell_order": "hipbdf.cost_order = 1: FC carried from the previous step (as PeleLMeX's functC) "
                                 "orders the work queue; the first (warm-up) step starts from FC = 0",
                   "path": "learned map (BF16 MFMA) + FP64 trust check + GPU BDF fallback" if sur
                   else "plain GPU BDF"},
        "bdf_value": bdf_value,
        "bdf_ms_per_step": ms_bdf,
        "surrogate": sur,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": roofline.MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / roofline.MI355X_FP64_TFLOPS, "traffic": traffic,
                     "traffic_source": (f"profiles/{traffic_src} (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a "
                                        "1 GiB device copy)") if traffic else "no profile of these kernel sources",
                     "kernel": "react_kernel<21,84,%d,%d> (plain BDF over the whole box, %d-thread blocks)" % (
                         st["lanes_per_cell"], st["waves_per_simd"], st["block_threads"]),
                     "kernel_ms": kernel_ms, "flops_per_launch": F,
                     "hbm_algorithmic_bytes": counts["bytes_cell"] * ncells,
                     "hbm_algorithmic_GBs": counts["bytes_cell"] * ncells / (kernel_ms / 1e3) / 1e9},
        "integrator": {"steps_per_cell": st["nst"] / st["cells"], "rhs_per_cell": st["nfe"] / st["cells"],
                       "jac_per_cell": st["nje"] / st["cells"], "lu_per_cell": st["nsetups"] / st["cells"],
                       "newton_per_cell": st["nni"] / st["cells"], "err_fail_per_cell": st["netf"] / st["cells"],
                       "deterministic_repeat": deterministic},
    }
    if not args.no_cpu:
        line["cpu_baseline"], ref = cpu_baseline(box, min(args.cpu_sample, ncells), max_order=args.max_order)
        line["parity"] = gpu_vs_oracle(bdf_state, ref)
    if not args.no_extra:
        torch.cuda.empty_cache()
        line["other_workloads"] = other_workloads(args)
    print(json.dumps(line), flush=True)


def other_workloads(args):
    """The other workloads, measured after the headline one in the same run (their own metrics;
    not part of `value`): the H2/air 32^3 box, the n-dodecane 128^3 box at 60 atm in eight
    launches, and the learned map + batched fallback over mixed AMR boxes for drm19 and
    n-dodecane. One timed step each after one warm-up (three for the small H2 box), parity on a
    CPU-oracle sample where the run has one. A workload that fails is reported as such."""
    import copy
    import torch
    runs = [("box_kernels_drm19_64", lambda a: bench_box_kernels(a), 20),
            ("h2_32", lambda a: bench_h2(a), 3),
            ("dodecane_128", lambda a: bench_chunks(a, "dodecane_lu"), 1),
            ("amr_mixed", lambda a: bench_amr(a, "drm19"), 1),
            ("amr_dodecane", lambda a: bench_amr(a, "dodecane_lu"), 1)]
    out = {}
    for name, fn, steps in runs:
        a = copy.copy(args)
        a.workload, a.steps, a.warmup, a.waves, a.block, a.copy_probe_mb = name, steps, 1, 0, 0, 0
        a.cpu_sample = min(args.cpu_sample, 4096)
        print(f"other workload {name}", file=sys.stderr, flush=True)
        try:
            ln = fn(a)
        except Exception as e:  # reported, not raised: the headline line stands on its own
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            continue
        finally:
            torch.cuda.empty_cache()
        if name.startswith("box_kernels"):
            out[name] = ln
            continue
        keep = {k: ln[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "warmup", "bdf_value",
                                   "failed_cells", "parity") if k in ln}
        keep["workload"] = ln["config"].get("workload")
        keep["roofline_frac"] = ln["roofline"]["frac"]
        if "cpu_baseline" in ln:
            keep["cpu_baseline"] = {k: ln["cpu_baseline"][k] for k in ("value", "unit", "cores", "sample")}
        if ln.get("surrogate"):
            keep["surrogate"] = {k: ln["surrogate"].get(k) for k in ("value", "speedup_vs_bdf", "accepted",
                                                                      "audit_max_dT_rel")}
        if "integrator" in ln and "failed_cells_per_step" in ln["integrator"]:
            keep["failed_cells_per_step"] = ln["integrator"]["failed_cells_per_step"]
        out[name] = keep
    return out


def bench_box_kernels(args):
    """The box kernels around the chemistry on the 64^3 drm19 flame box (MKS arrays as PeleLMeX
    holds them): reactionRateRhoY, setTemperature (from the stored T), floorSpecies, I_R and
    getHeatRelease. Each is timed over `steps` calls after two warm-up calls with HIP events on
    the launch stream, so a call's time includes its launch and the entry point's own stream
    synchronization; bytes are the algorithmic state traffic per cell (doubles read + written)
    and the rate is set against the 8 TB/s HBM peak (the first two evaluate the rates / the
    species thermo per cell and are bound by that, the others by HBM)."""
    import torch
    from pelelmex_amd import roofline, workloads
    from pelelmex_amd.reactor import Reactor, box_tensors
    box = workloads.ch4_air_flame_box()
    n = box["T"].size
    N = box["rY"].shape[0]
    dev = torch.device("cuda:0")
    t = box_tensors(box, dev)
    rY = (t["rY"] * 1e3).contiguous()
    rH = (t["rH"] * 0.1).contiguous()
    T = t["T"].clone()
    rho = rY.sum(0).contiguous()
    r = Reactor.create("ReactorCvode", "drm19", {"ode.rtol": 1e-10, "ode.atol": 1e-10})
    r.init(2, 1)
    rYdot = torch.empty_like(rY)
    HR = torch.empty_like(T)
    IR = torch.empty_like(rY)
    F = torch.zeros_like(rY)
    rY_old = rY.clone()
    ops = [("reactionRateRhoY", lambda: r.reaction_rate_rhoY(rY, rH, T, rYdot=rYdot), (2 * N + 2) * 8),
           ("setTemperature", lambda: r.set_temperature(rho, rY, rH, T), (N + 4) * 8),
           ("floorSpecies", lambda: r.floor_species(rY, rho, rH, T), (2 * N + 3) * 8),
           ("I_R", lambda: r.reaction_source(rY_old, rY, F, 1e-6, IR=IR), 4 * N * 8),
           ("getHeatRelease", lambda: r.heat_release(T, rYdot, HR=HR), (N + 2) * 8)]
    out = {"workload": "drm19 CH4/air flame-like box 64^3 = 262144 cells, MKS state arrays",
           "calls_timed": args.steps}
    for name, fn, bpc in ops:
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        gbs = bpc * n / (ms / 1e3) / 1e9
        out[name] = {"ms_per_call": ms, "cells_per_s": n / (ms / 1e3), "bytes_per_cell": bpc, "GBs": gbs,
                     "frac_of_hbm_peak": gbs / roofline.MI355X_HBM_GBS}
    r.close()
    return out


def bench_h2(args):
    """Plain GPU BDF on the H2/air LiDryer box (32^3 = 32768 cells, phi ~ U[0.5, 2], burn fraction
    ~ U[0, 1], radicals 1e-8..1e-3, T ~ U[1000, 1800] K, 1 atm, dt_react 1e-6 s, no sources,
    rtol = atol = 1e-10). One step = one react over the box restored from a pristine copy (FC
    carried); the CPU oracle integrates the whole box for the baseline."""
    import torch
    from pelelmex_amd import roofline, workloads
    from pelelmex_amd.reactor import Reactor, box_tensors
    box = workloads.h2_air_box()
    n = box["T"].size
    t = box_tensors(box, torch.device("cuda:0"))
    pristine = {k: t[k].clone() for k in ("rY", "T", "rH")}
    r = Reactor.create("ReactorCvode", "LiDryer", {"ode.rtol": 1e-10, "ode.atol": 1e-10,
                                                   "cvode.max_order": args.max_order, "cvode.max_substeps": 10000})
    r.init(2, 1)

    def step():
        for k in pristine:
            t[k].copy_(pristine[k])
        nfail, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0)
        if nfail:
            raise RuntimeError(f"{nfail} cells failed to integrate")
        return r.stats()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    r.close()
    st = stats[-1]
    kernel_ms = float(np.mean([x["kernel_ms"] for x in stats]))
    F = roofline.flops("LiDryer", st)
    achieved = F / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "chem cells/s per dt_react: plain GPU BDF, H2/air LiDryer",
        "value": n / (ms / 1e3), "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "LiDryer H2/air box 32^3 = 32768 cells, dt_react 1e-6 s, 1 atm, no sources",
                   "mechanism": "LiDryer", "cells": n, "dt_react": box["dt"], "rtol": 1e-10, "atol": 1e-10,
                   "max_order": args.max_order, "lanes_per_cell": st["lanes_per_cell"],
                   "waves_per_simd": st["waves_per_simd"], "seed": 1001},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": roofline.MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / roofline.MI355X_FP64_TFLOPS, "traffic": None,
                     "traffic_source": "not profiled for this workload",
                     "kernel": "react_kernel<9,21,%d,%d>" % (st["lanes_per_cell"], st["waves_per_simd"]),
                     "kernel_ms": kernel_ms, "flops_per_launch": F},
        "integrator": {"steps_per_cell": st["nst"] / st["cells"], "rhs_per_cell": st["nfe"] / st["cells"]},
    }
    if not args.no_cpu:
        line["cpu_baseline"], ref = cpu_baseline(box, n, what="the same 32^3 box (every cell)",
                                                 max_order=args.max_order)
        line["parity"] = gpu_vs_oracle(t, ref)
    return line


def bench_amr(args, mech="drm19"):
    """Learned map + batched BDF fallback over mixed AMR boxes in one hipGraph: boxes of
    three levels with edge 16, 32 or 64 drawn uniformly until about 1.5 M cells, per-box dt_react
    from {1e-6, 2e-6, 4e-6} s (seed 5005 + box). mech "dodecane_lu": the n-dodecane 60 atm states
    (60 atm, 10 % masked) with the 6 x 512 network; mech "drm19": flame-like drm19 states with 5 %
    Gaussian sources and the 4 x 256 network. One step = restoring every box's state from its
    pristine copy (FC carries over, as PeleLMeX's functC does, so from the second step the cells are
    handed out by decreasing cost) and one replay of the captured pr_react_batch_enqueue over all
    boxes. The plain batched BDF, captured the same way, is timed on the same boxes, and the learned
    result is compared with it on every cell."""
    import torch
    from pelelmex_amd import roofline, surrogate, workloads
    from pelelmex_amd.reactor import Reactor, box_tensors
    dev = torch.device("cuda:0")
    rng = np.random.Generator(np.random.Philox(5005))
    boxes, total = [], 0
    while total < 1_500_000:
        e = int(rng.choice([16, 32, 64]))
        lev = int(rng.integers(0, 3))
        dt = float(rng.choice([1e-6, 2e-6, 4e-6]))
        if mech == "dodecane_lu":
            b = workloads.dodecane_box(nx=e, ny=e, nz=e, seed=5005 + len(boxes), dt=dt)
        else:
            b = workloads.ch4_air_flame_box(nx=e, ny=e, nz=e, seed=5005 + len(boxes), dt=dt)
        b["level"] = lev
        boxes.append(b)
        total += e ** 3
    base = [box_tensors(b, dev) for b in boxes]
    print(f"amr ({mech}): {len(boxes)} boxes, {total} cells", file=sys.stderr, flush=True)
    wname = "surrogate_dodecane_amr.npz" if mech == "dodecane_lu" else "surrogate_drm19_amr.npz"
    wpath = args.surrogate or os.path.join(ROOT, "pelelmex_amd", wname)
    model = surrogate.load(wpath) if os.path.exists(wpath) else None
    lanes = 64 if mech == "dodecane_lu" else 32
    keys = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": args.max_order, "cvode.max_substeps": 10000,
            "cvode.solve_type": "denseAJ_direct", "hipbdf.lanes_per_cell": lanes}

    def timed(with_model):
        skeys = dict(keys)
        if with_model and "trust_rtol" in model:  # a network whose calibration found no tolerance keeps the defaults
            skeys["surrogate.trust_rtol"] = float(model["trust_rtol"])
            skeys["surrogate.trust_atol"] = float(model["trust_atol"])
        r = Reactor.create("ReactorCvode", mech, skeys)
        r.init(2, 1)
        if with_model:
            surrogate.attach(r, model)
        work = []
        for b, t0 in zip(boxes, base):
            t = {k: v.clone() for k, v in t0.items()}
            t["dt"] = b["dt"]
            work.append(t)

        def restore():  # FC carries over, as PeleLMeX's functC does between advances
            for t, t0 in zip(work, base):
                for k in ("rY", "T", "rH"):
                    t[k].copy_(t0[k])

        restore()
        r.react_batch(work)  # eager: sizes every buffer the capture uses
        restore()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            r.react_batch(work, enqueue_only=True)
        for _ in range(args.warmup):
            restore()
            g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            restore()
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        nfail = r.collect()   # cells of the last replay that failed (CVODE's error-test limit), kept at
        st = r.stats()        # their last accepted state and reported
        st["failed_cells"] = int(nfail)
        r.close()
        out = {k: [t[k].clone() for t in work] for k in ("rY", "T")}
        return ms, st, out

    ms_bdf, st_bdf, o_bdf = timed(False)
    F = roofline.flops(mech, st_bdf)
    achieved = F / (ms_bdf / 1e3) / 1e12
    counts = roofline.static_counts(mech)
    sizes = {e: sum(1 for b in boxes if b["dims"][0] == e) for e in (16, 32, 64)}
    sur = None
    if model is not None:
        ms_sur, st_sur, o_sur = timed(True)
        act = [torch.as_tensor(b["mask"] != -1, device=dev).reshape(t.shape) for b, t in zip(boxes, o_bdf["T"])]
        dT = max(((a - b).abs() / b)[m].max().item() for a, b, m in zip(o_sur["T"], o_bdf["T"], act))
        sur = {"value": total / (ms_sur / 1e3), "ms_per_step": ms_sur,
               "accepted": st_sur["sur_accepted"] / total, "rejected_to_bdf": st_sur["sur_rejected"] / total,
               "audited": st_sur["sur_audited"], "audit_max_dT_rel": st_sur["audit_max_dT_rel"],
               "max_dT_rel_vs_bdf_all_cells": dT, "speedup_vs_bdf": ms_bdf / ms_sur,
               "failed_cells": st_sur["failed_cells"],
               "network": os.path.relpath(wpath, ROOT) + " (%s)" % " x ".join(str(x) for x in model["dims"][1:-1])}
    value = sur["value"] if sur else total / (ms_bdf / 1e3)
    what = "n-dodecane (dodecane_lu), 60 atm, 10% masked" if mech == "dodecane_lu" else "drm19, 5% sources"
    line = {
        "metric": f"chem cells/s per dt_react: surrogate+fallback over mixed AMR boxes in one hipGraph ({mech})",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sur["ms_per_step"] if sur else ms_bdf, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{what}; {len(boxes)} boxes on 3 levels (edges 16/32/64: {sizes[16]}/{sizes[32]}/"
                               f"{sizes[64]}), {total} cells, per-box dt_react in {{1e-6, 2e-6, 4e-6}} s",
                   "mechanism": mech, "cells": total, "boxes": len(boxes), "max_order": args.max_order,
                   "lanes_per_cell": lanes, "seed": "5005 + box",
                   "path": ("one hipGraph: features per box, one network pass, trust check per box, one persistent "
                            "BDF launch over every box's rejected cells, 1% audit") if sur else
                           "one hipGraph: one persistent BDF launch over every box (no network shipped)"},
        "bdf_value": total / (ms_bdf / 1e3), "bdf_ms_per_step": ms_bdf,
        "failed_cells": st_bdf["failed_cells"],
        "surrogate": sur,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": roofline.MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / roofline.MI355X_FP64_TFLOPS, "traffic": None,
                     "traffic_source": "not profiled for this workload",
                     "kernel": "react_kernel<%d,%d,%d,*> (plain batched BDF, whole graph replay)" % (
                         counts["N"], counts["R"], lanes),
                     "flops_per_step": F},
    }
    return line


def bench_chunks(args, mech="dodecane_lu"):
    """Plain GPU BDF on n-dodecane (dodecane_lu, 53 species, 253 reactions; or the
    synthetic syn53 of the same size with mech="syn53"), a 128^3 box at 60 atm with 10 % of cells
    masked, dt_react 2e-6 s, as eight 64^3 chunks of 262144 cells (one react launch each). One step
    = the eight launches over the whole box, each chunk restored from a pristine device copy first.
    Lanes per cell, waves per SIMD and block size are chosen by the reactor's autotuner on the first
    (warm-up) call, or fixed by --waves / --block."""
    import torch
    from pelelmex_amd import roofline, workloads
    from pelelmex_amd.reactor import Reactor, box_tensors
    dev = torch.device("cuda:0")
    gen_chunk = workloads.dodecane_chunk if mech == "dodecane_lu" else workloads.syn53_chunk
    nchunk = 8
    chunks, pristine, host0 = [], [], None
    for c in range(nchunk):
        b = gen_chunk(c)
        if c == 0:
            host0 = b
        t = box_tensors(b, dev)
        chunks.append((b, t))
        pristine.append({k: t[k].clone() for k in ("rY", "T", "rH")})
    ncells = sum(b["T"].size for b, _ in chunks)
    active = sum(int((b["mask"] != -1).sum()) for b, _ in chunks)
    copy_probe(args.copy_probe_mb, dev)
    keys = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": args.max_order, "cvode.max_substeps": 10000,
            "cvode.solve_type": "denseAJ_direct"}
    if args.waves:
        keys["hipbdf.waves_per_simd"] = args.waves
        if args.block:
            keys["hipbdf.block"] = args.block
    else:
        keys["hipbdf.autotune"] = 1
    r = Reactor.create("ReactorCvode", mech, keys)
    r.init(2, 1)

    failed = {}

    def launch(c):
        # A cell whose integration fails (CVODE's error-test failure limit, hit by a few igniting
        # 60 atm cells at rtol = atol = 1e-10; the oracle fails the same cells) keeps its last
        # accepted state; the count is reported, as PeleLMeX's react return value would carry it.
        b, t = chunks[c]
        for k in pristine[c]:
            t[k].copy_(pristine[c][k])
        nfail, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], b["dt"], 0.0)
        failed[c] = nfail
        return r.stats()

    print(f"{mech}: {ncells} cells generated and resident; warm-up (autotune) on chunk 0", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        launch(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats = [launch(c) for c in range(nchunk)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    r.close()
    tot = {k: sum(s[k] for s in stats) for k in ("nfe", "nje", "nsetups", "nni", "nst", "netf", "cells")}
    kernel_ms = sum(s["kernel_ms"] for s in stats)
    F = roofline.flops(mech, tot)
    achieved = F / (kernel_ms / 1e3) / 1e12
    counts = roofline.static_counts(mech)
    st = stats[-1]
    workload_key = f"{mech} 128^3 lanes {st['lanes_per_cell']} waves {st['waves_per_simd']} max_order {args.max_order}"
    traffic, traffic_src = profiled_traffic(workload_key)
    what = ("n-dodecane dodecane_lu (53 species, 253 reactions)" if mech == "dodecane_lu"
            else "syn53 (synthetic: 53 species, 268 reactions)")
    line = {
        "metric": "chem cells/s per dt_react: plain GPU BDF, " + ("n-dodecane" if mech == "dodecane_lu"
                                                                  else "synthetic 53-species mechanism"),
        "value": ncells / (ms / 1e3),
        "unit": "cells/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": what + " 128^3 = 2097152 cells in 8 launches of "
                               "262144, 60 atm, dt_react 2e-6 s, 10% masked, mixture fraction ~ Beta(0.5, 2), "
                               "T ~ U[700, 2400] K",
                   "mechanism": mech, "cells": ncells, "active_cells": active, "dt_react": host0["dt"],
                   "rtol": 1e-10, "atol": 1e-10, "max_order": args.max_order,
                   "lanes_per_cell": st["lanes_per_cell"], "waves_per_simd": st["waves_per_simd"],
                   "block_threads": st["block_threads"], "seed": "3003 + chunk",
                   "path": "plain GPU BDF (" + ("fixed variant" if args.waves else "autotuned") + ")"},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": roofline.MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / roofline.MI355X_FP64_TFLOPS, "traffic": traffic,
                     "traffic_source": (f"profiles/{traffic_src} (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a "
                                        "1 GiB device copy)") if traffic else "no profile of these kernel sources",
                     "kernel": "react_kernel<53,%d,%d,%d> (%d-thread blocks)" % (
                         counts["R"], st["lanes_per_cell"], st["waves_per_simd"], st["block_threads"]),
                     "launches_per_step": nchunk, "kernel_ms_per_step": kernel_ms,
                     "kernel_ms_avg_launch": kernel_ms / nchunk, "flops_per_step": F,
                     "traffic_per": "launch", "hbm_algorithmic_bytes_per_launch": counts["bytes_cell"] * ncells // nchunk,
                     "hbm_algorithmic_GBs": counts["bytes_cell"] * ncells / (kernel_ms / 1e3) / 1e9},
        "integrator": {"steps_per_cell": tot["nst"] / tot["cells"], "rhs_per_cell": tot["nfe"] / tot["cells"],
                       "jac_per_cell": tot["nje"] / tot["cells"], "lu_per_cell": tot["nsetups"] / tot["cells"],
                       "newton_per_cell": tot["nni"] / tot["cells"], "err_fail_per_cell": tot["netf"] / tot["cells"],
                       "failed_cells_per_step": int(sum(failed.values())),
                       "failed_note": "cells ending in CVODE's error-test failure limit (status -3), the same "
                                      "cells as the oracle's; they keep their last accepted state"},
    }
    if not args.no_cpu:
        line["cpu_baseline"], ref = cpu_baseline(host0, min(args.cpu_sample, 65536), max_order=args.max_order,
                                                 what="the first 64^3 chunk (masked cells count as cells)")
        line["parity"] = gpu_vs_oracle(chunks[0][1], ref)   # chunk 0's tensors hold its last result
        line["parity"]["box"] = "the first 64^3 chunk, the CPU sample's cells"
    return line


if __name__ == "__main__":
    main()
