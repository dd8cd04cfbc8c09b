#!/usr/bin/env python3
"""Benchmark: 3-D Sedov blast, 256^3, FP64, PPM + HLLC + RK3, CFL 0.3,
outflow boundaries, on one MI355X through libplutomx.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]

One step = one full time step of the reference's main loop for this path
(boundaries + 3 RK stages + cons->prim + CFL reduction + NextTimeStep), all
on the device with the state resident in HBM.  Prints ONE JSON line.

Also reported:
  roofline      dominant kernel (the fused stage kernel): algorithmic bytes
                per launch / its mean duration, measured with HIP events on
                the library's stream over the timed steps;
  cpu_baseline  the unmodified reference CPU build (oracle/_ref, compiled
                from the reference sources) timed on this host on a bounded
                sample, single core.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pluto-sirocco_amd"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import setups  # noqa: E402

METRIC = "Mcell-updates/s per RK step; surrogate vs plain cooling at <=1e-3 T error"
HBM_PEAK_GBS = 8000.0
# HBM bytes per stage-kernel launch at 256^3, from two rocprofv3 --pmc passes
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a plain copy of known
# size reads on this device), mean over the three RK3 stage launches; the
# commands and per-launch values are in profiles/stage_kernel_sedov256.md.
TRAFFIC_BYTES_PER_LAUNCH_256 = 3.768e9
TRAFFIC_SOURCE = ("profiles/stage_kernel_sedov256.md: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                  "of this kernel (a profile value, not measured in this run)")


def algorithmic_bytes_per_cell(rk: int) -> list:
    """Per interior cell and RK stage s: 8*NVAR*(4 + [s >= 2]) bytes (read V
    and U, write U and V, plus read U^n from stage 2 on)."""
    nvar = 5
    return [8 * nvar * (4 + (1 if s >= 2 else 0)) for s in range(1, rk + 1)]


def cpu_baseline(n_sample: int = 128, steps: int = 3) -> dict | None:
    """Time the reference CPU build on a bounded sample of the workload."""
    from oracle import refrun
    cfg = setups.c2_sedov(n_sample)
    V0 = setups.sedov(cfg, seed=1234)
    binary = refrun.reference_binary(cfg.reference_problem())
    if binary.exists():
        r = refrun.run_reference(cfg, V0, steps, time_only=True)
        kind, wall, nst = "reference", r["wall"], r["steps_timed"]
    else:  # plain-C restatement of the same algorithm
        t0 = time.perf_counter()
        refrun.run_oracle(cfg, V0, steps)
        kind, wall, nst = "port", time.perf_counter() - t0, steps
    value = cfg.ncells * nst / wall / 1e6
    return {"value": round(value, 4), "unit": "Mcell-updates/s", "cores": 1, "kind": kind,
            "sample": f"Sedov {n_sample}^3 PPM+HLLC+RK3, {nst} steps, 1 process on "
                      f"{os.cpu_count()} host cores ({wall:.2f} s)"}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_all_cores(n_sample: int = 128, steps: int = 3) -> dict | None:
    """The reference CPU build on every host core this run may use: one
    independent serial process per core on the same bounded sample (there
    is no MPI here), aggregate cell-updates over the wall time of the
    slowest (the throughput analogue of a sweep of independent runs)."""
    import concurrent.futures as cf
    from oracle import refrun
    cfg = setups.c2_sedov(n_sample)
    if not refrun.reference_binary(cfg.reference_problem()).exists():
        return None
    V0 = setups.sedov(cfg, seed=1234)
    cores = min(os.cpu_count() or 1, 16)   # the GPU's share of a shared host
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        runs = list(ex.map(lambda _: refrun.run_reference(cfg, V0, steps, time_only=True),
                           range(cores)))
    wall = max(r["wall"] for r in runs)
    nst = runs[0]["steps_timed"]
    return {"value": round(cfg.ncells * nst * cores / wall / 1e6, 4), "unit": "Mcell-updates/s",
            "cores": cores, "kind": "reference", "cpu": cpu_model(),
            "sample": f"Sedov {n_sample}^3 PPM+HLLC+RK3, {nst} steps, {cores} concurrent "
                      f"processes (slowest {wall:.2f} s)"}


# 4 -> 64 -> 64 -> 64 -> 1: forward 2 (4*64 + 64*64 + 64*64 + 64) = 17,024
# FLOP per sample; backward adds the weight and input gradients (the first
# layer's weights only): 2 (4*64) + 4 (64*64) * 2 + 4 * 64 = 33,536
MLP_FWD_FLOP_PER_SAMPLE = 2 * (4 * 64 + 64 * 64 + 64 * 64 + 64)
MLP_TRAIN_FLOP_PER_SAMPLE = MLP_FWD_FLOP_PER_SAMPLE + 2 * 4 * 64 + 4 * 64 * 64 * 2 + 4 * 64


def kh_leg(n: int = 256, steps: int = 100, warmup: int = 10) -> dict:
    """The 2-D Kelvin-Helmholtz workload (256^2, PLM van Leer + HLLC + RK2,
    CFL 0.4, periodic) for its stated 100 steps, graph-replayed: a 65,536-
    cell grid is launch-latency bound (23.6 MB of algorithmic traffic per
    step, 3 us at the HBM peak), so the replay is what a driver would use."""
    import plutomx
    cfg = setups.c1_kh(n)
    V0 = setups.kelvin_helmholtz(cfg, seed=1234)
    with plutomx.Hydro(cfg) as h:
        h.set_prim(V0)
        h.set_time(0.0, cfg.first_dt)
        h.run(warmup, graph=True)
        h.sync()
        out = {}
        for graph in (True, False):
            t0 = time.perf_counter()
            h.run(steps, graph=graph)
            h.sync()
            el = time.perf_counter() - t0
            out["graph" if graph else "eager"] = {
                "Mcell_updates_per_s": round(cfg.ncells * steps / el / 1e6, 2),
                "ms_per_step": round(1e3 * el / steps, 4)}
    out["workload"] = (f"2-D KH {n}^2, PLM(van Leer)+HLLC+RK2, CFL 0.4, periodic, "
                       f"{steps} steps (the metric's CPU-runnable configuration)")
    return out


def wind_cpu_baseline(nr: int = 256, nt: int = 128, steps: int = 4) -> dict | None:
    """The compiled reference on a bounded sample of the C3 wind (a quarter
    of its cells, the same physics); the plain-C restatement where the
    reference build is absent."""
    from oracle import refrun
    cfg = setups.c3_wind(nr, nt, setups.COOL_PLAIN)
    V0 = setups.wind_ic(cfg)
    binary = refrun.reference_binary(cfg.reference_problem())
    if binary.exists():
        r = refrun.run_reference(cfg, V0, steps, time_only=True)
        kind, wall, nst = "reference", r["wall"], r["steps_timed"]
    else:
        t0 = time.perf_counter()
        refrun.run_wind_oracle(cfg, V0, steps)
        kind, wall, nst = "port", time.perf_counter() - t0, steps
    return {"value": round(cfg.ncells * nst / wall / 1e6, 4), "unit": "Mcell-updates/s",
            "cores": 1, "kind": kind,
            "sample": f"wind {nr}x{nt}, PLM+HLLC+RK2, line force, plain cooling, {nst} steps, "
                      f"1 process ({wall:.2f} s)"}


def c5_leg(steps: int, mlp=None) -> dict:
    """64 uneven wind grids advanced as one batch (graph replay, autotuned
    tiles), with the surrogate cooling when a trained network is given."""
    import plutomx
    cfgs = setups.c5_batch()
    with plutomx.WindBatch(cfgs) as b:
        for w, c in zip(b.grids, cfgs):
            w.set_prim(setups.wind_ic(c))
            w.set_time(0.0, c.first_dt, 0)
        b.set_step(0)
        if mlp is not None:
            b.set_surrogate(mlp)
        tune = b.autotune()
        b.run(2, graph=True)
        b.sync()
        t0 = time.perf_counter()
        b.run(steps, graph=True)
        b.sync()
        el = time.perf_counter() - t0
        cells = b.cells
    cooling = "surrogate cooling with plain fallback" if mlp is not None else "plain cooling"
    return {"grids": len(cfgs), "cells": cells,
            "Mcell_updates_per_s": round(cells * steps / el / 1e6, 2),
            "ms_per_step": round(1e3 * el / steps, 3), "autotune": tune,
            "workload": "64 wind grids 128..768 x 64..384, L_x 1e36..1e38, "
                        f"k 0.1..0.4, {cooling}, graph replay, {steps} steps"}


def wind_legs(steps: int = 200, cpu: bool = True, c3_only: bool = False) -> dict:
    """The wind workloads: C3 (512 x 256 spherical wind, plain sub-cycled
    cooling), C4 (the same with the MLP surrogate, trained here on the GPU
    on 2e6 samples of the plain integrator), the cooling step alone on
    cells drawn over the surrogate's training hull (where cooling is stiff),
    and C5 (64 uneven grids advanced as one batch).  c3_only: the C3 leg
    alone (for counter passes, which serialise every kernel)."""
    import plutomx
    out = {}
    cfg = setups.c3_wind(512, 256, setups.COOL_PLAIN)
    V0 = setups.wind_ic(cfg)

    def run_wind(mlp=None):
        # the main loop replayed as a captured HIP graph (what a driver
        # runs), then 20 eager steps with per-launch events for the split
        # into stage and cooling time
        with plutomx.Wind(cfg) as w:
            w.set_prim(V0)
            w.set_time(0.0, cfg.first_dt, 0)
            if mlp is not None:
                w.set_surrogate(mlp)
            w.run(2, graph=True)
            w.sync()
            w.surrogate_stats(reset=True)
            t0 = time.perf_counter()
            w.run(steps, graph=True)
            w.sync()
            el = time.perf_counter() - t0
            st = w.surrogate_stats()
            T = w.get_prim()
            w.timing(True, reset=True)
            w.run(20)
            w.sync()
            tm = w.timing(False)
        return {"Mcell_updates_per_s": round(cfg.ncells * steps / el / 1e6, 2),
                "ms_per_step": round(1e3 * el / steps, 4),
                "stage_ms_per_rk_stage": round(tm["stage_ms"] / max(tm["stage_launches"], 1), 4),
                "cool_ms_per_call": round(tm["cool_ms"] / max(tm["cool_launches"], 1), 4),
                "surrogate": st}, T

    plain, Tp = run_wind()
    out["c3_plain"] = dict(plain, workload=f"wind 512x256, PLM+HLLC+RK2, line force, plain "
                                            f"cooling, {steps} steps, graph replay")
    if cpu:
        out["c3_plain"]["cpu_baseline"] = wind_cpu_baseline()
    if c3_only:
        return out
    # the same wind with the production problem's set-up (cv_idl boundaries
    # and floors, MULTID shock flattening, characteristic limiting, entropy
    # switch, HLL), 20 steps: from this start state the set-up reaches a zone
    # whose Cash-Karp integration exceeds its sub-step limit near step 90
    # (the oracle stops there too), after which dt collapses
    prod_cfg = dataclasses.replace(setups.cvidl_wind(512, 256, dfloor=1.e-20),
                                   shock_flattening=True, char_limiting=True,
                                   entropy_switch=True, solver=setups.HLL)
    with plutomx.Wind(prod_cfg) as w:
        w.set_prim(V0)
        w.set_time(0.0, prod_cfg.first_dt, 0)
        w.run(2)
        w.sync()
        w.timing(True, reset=True)
        psteps = min(steps, 20)
        t0 = time.perf_counter()
        w.run(psteps)
        w.sync()
        el = time.perf_counter() - t0
        tm = w.timing(False)
    out["c3_production_setup"] = {
        "Mcell_updates_per_s": round(prod_cfg.ncells * psteps / el / 1e6, 2),
        "ms_per_step": round(1e3 * el / psteps, 4),
        "stage_ms_per_rk_stage": round(tm["stage_ms"] / max(tm["stage_launches"], 1), 4),
        "workload": "C3 with cv_idl boundaries and floors, SHOCK_FLATTENING MULTID, "
                    f"CHAR_LIMITING, ENTROPY_SWITCH ALWAYS, HLL, {psteps} steps"}
    # output every 5 steps (V + the 23 user variables, 30 MB, read in place
    # from pinned memory): no output, the snapshot collected at once
    # (blocking), and collected after the next 5 steps were queued (its
    # transfer overlapping them)
    def output_run(mode, nblk=4, per=5):
        with plutomx.Wind(prod_cfg) as w:
            w.set_prim(V0)
            w.set_time(0.0, prod_cfg.first_dt, 0)
            w.run(2)
            w.sync()
            t0 = time.perf_counter()
            for _ in range(nblk):
                if mode == "blocking":
                    w.snapshot(user_vars=True)
                    w.snapshot_wait(copy=False)
                    w.run(per)
                elif mode == "overlapped":
                    w.snapshot(user_vars=True)
                    w.run(per)
                    w.snapshot_wait(copy=False)
                else:
                    w.run(per)
            w.sync()
            return (time.perf_counter() - t0) / (nblk * per)
    ms = {mode: 1e3 * output_run(mode) for mode in ("none", "blocking", "overlapped")}
    out["c3_production_output"] = {
        "ms_per_step_no_output": round(ms["none"], 4),
        "ms_per_step_blocking": round(ms["blocking"], 4),
        "ms_per_step_overlapped": round(ms["overlapped"], 4),
        "workload": "C3 production set-up, a snapshot of V + 23 user variables "
                    "(30 MB) every 5 steps, 20 steps"}
    m = plutomx.Mlp(seed=11)
    t0 = time.perf_counter()
    m.sample(cfg, 2_000_000, seed=5)
    t1 = time.perf_counter()
    hist = m.train(20000, batch=8192, lr0=3e-3, lr1=1e-5)
    t2 = time.perf_counter()
    surr, Ts = run_wind(m)
    out["c4_surrogate"] = dict(surr, workload="C3 with the surrogate cooling (cells outside the "
                                              "training hull take the plain integrator)",
                               training={"samples": 2_000_000, "sample_s": round(t1 - t0, 2),
                                         "steps": 20000, "batch": 8192,
                                         "train_s": round(t2 - t1, 2),
                                         "final_loss": float(np.mean(hist[-50:])),
                                         # forward + backward FLOPs of the 4-64-64-64-1
                                         # network per sample (dense layers only, the first
                                         # layer unpadded), over the wall time of the steps
                                         # (training kernel, Adam update and launches)
                                         "TFLOP_per_s_wall": round(
                                             MLP_TRAIN_FLOP_PER_SAMPLE * 8192 * 20000
                                             / (t2 - t1) / 1e12, 3)},
                               speedup_vs_plain=round(plain["ms_per_step"] / surr["ms_per_step"], 3))
    # cooling alone over the training hull (dt 1e-2..1e2 t_cool)
    m.sample(cfg, 131072, seed=99)
    X, Y, nev, aux = m.get_dataset(131072)
    ok = np.isfinite(Y)
    rho, p0, xi, dt = (aux[ok, q].copy() for q in range(4))
    res = {}
    for mode in (0, 1, 0, 1):
        res[mode] = m.cool_cells(cfg, rho, p0, xi, dt, mode)
    e = np.abs(res[1][0] / res[0][0] - 1.0)
    out["cooling_on_training_hull"] = {
        "cells": int(ok.sum()), "plain_ms": round(res[0][2], 3),
        "surrogate_ms": round(res[1][2], 3), "speedup": round(res[0][2] / res[1][2], 2),
        "frac_T_err_le_1e-3": float(np.mean(e <= 1e-3)), "T_err_p99": float(np.quantile(e, 0.99)),
        "stats": res[1][1], "plain_radiat_calls": res[0][1]["radiat_calls"]}
    # C5: batched parameter sweep, with the surrogate cooling (the network
    # trained above; the hull does not depend on L_x or k) and with the
    # plain integrator for comparison
    for name, mlp in (("c5_batch", m), ("c5_batch_plain", None)):
        out[name] = c5_leg(steps, mlp)
    m.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay a captured HIP graph")
    ap.add_argument("--no-wind", action="store_true", help="skip the wind workloads")
    ap.add_argument("--wind-only", action="store_true",
                    help="only the wind workloads (for profiling them)")
    ap.add_argument("--wind-steps", type=int, default=200)
    ap.add_argument("--wind-c3-only", action="store_true",
                    help="with --wind-only: the C3 leg alone (for counter passes)")
    ap.add_argument("--wind-c5-only", action="store_true",
                    help="with --wind-only: the C5 batch with plain cooling alone "
                         "(for counter passes)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the hydro step shards by spatial "
              "domain decomposition (see the design notes), which is not built here.",
              file=sys.stderr)
        return 2

    import plutomx
    if args.wind_only and args.wind_c5_only:
        print(json.dumps({"wind": {"c5_batch_plain": c5_leg(args.wind_steps)}}), flush=True)
        return 0
    if args.wind_only:
        print(json.dumps({"wind": wind_legs(args.wind_steps, cpu=not args.no_cpu,
                                            c3_only=args.wind_c3_only)}), flush=True)
        return 0
    if args.copy_probe:
        nbytes = args.copy_probe * 1 << 20
        ms = plutomx.copy_probe(nbytes, reps=5)
        rt = plutomx.copy_probe(nbytes, reps=5, runtime=True)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms,
                          "GBps_read_plus_write": 2 * nbytes / ms / 1e6,
                          "runtime_memcpy_GBps": 2 * nbytes / rt / 1e6}), flush=True)
        return 0
    cfg = setups.c2_sedov(args.n)
    V0 = setups.sedov(cfg, seed=1234)
    h = plutomx.Hydro(cfg)
    h.set_prim(V0)
    h.set_time(0.0, cfg.first_dt)

    h.run(args.warmup, graph=args.graph)
    h.sync()

    h.stage_timing(True, reset=True)
    t0 = time.perf_counter()
    h.run(args.steps)
    h.sync()
    t1 = time.perf_counter()
    stage_ms, launches = h.stage_timing(False, reset=True)
    t, dt, inv, flags = h.get_time()
    # the same step with the state handed over in host memory (upload V,
    # one step, download V): what a driver pays that keeps Vc on the host
    Vh = h.get_prim()
    t2 = time.perf_counter()
    for _ in range(2):
        h.set_prim(Vh)
        h.run(1)
        Vh = h.get_prim()
    host_ms = 1e3 * (time.perf_counter() - t2) / 2
    h.close()

    copy_ms = plutomx.copy_probe(1 << 30, reps=5)
    rt_copy_ms = plutomx.copy_probe(1 << 30, reps=5, runtime=True)
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = cfg.ncells * args.steps / elapsed / 1e6

    per_stage = algorithmic_bytes_per_cell(cfg.rk)
    bytes_per_launch = cfg.ncells * sum(per_stage) / len(per_stage)
    mean_launch_ms = stage_ms / max(launches, 1)
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    traffic = TRAFFIC_BYTES_PER_LAUNCH_256 if args.n == 256 else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": TRAFFIC_SOURCE if traffic else None,
                "kernel": "stage_kernel (fused PPM+HLLC stage, all directions)",
                "mean_launch_ms": round(mean_launch_ms, 4),
                "algorithmic_bytes_per_launch": bytes_per_launch,
                "plain_copy_GBps": round(2 * (1 << 30) / copy_ms / 1e6, 1),
                "runtime_copy_GBps": round(2 * (1 << 30) / rt_copy_ms / 1e6, 1)}

    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Sedov IC)",
           "config": {"workload": f"3-D Sedov {args.n}^3, PPM+HLLC+RK3, CFL 0.3, outflow",
                      "cells": cfg.ncells, "sim_time_after": t, "dt_next": dt,
                      "c2p_flags": flags},
           "roofline": roofline,
           "host_roundtrip": {"ms_per_step": round(host_ms, 3),
                              "Mcell_updates_per_s": round(cfg.ncells / host_ms / 1e3, 2),
                              "what": "set_prim + 1 step + get_prim of the full state from "
                                      "pageable host memory (not the headline value)"},
           "cpu_baseline": None if args.no_cpu else cpu_baseline(),
           "cpu_baseline_all_cores": None if args.no_cpu else cpu_baseline_all_cores(),
           "kh_256": kh_leg()}
    if not args.no_wind:
        out["wind"] = wind_legs(args.wind_steps, cpu=not args.no_cpu)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
