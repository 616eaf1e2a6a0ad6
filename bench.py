"""Throughput of the MI355X FDM engine on the 512x512x1 thin film with a hole.

Workload: 512x512x1 cells of 2x2x5 nm, circular non-magnetic hole of radius
64 cells at the centre (249,252 active of 262,144 cells), Ms 8.6e5 A/m,
A 1.3e-11 J/m, alpha 0.02, B = (10 mT, 0, 0), seeded (42) random-unit m0,
exchange + Newell demag + Zeeman, Dormand-Prince RK45 with atol 1e-5,
dt0 1e-13, dt in [1e-18, 1e-10], headroom 0.8.

One step = one accepted RK45 step of the whole grid (6 RHS evaluations with
FSAL, rejected trials included in the time).  The steps run in the engine's
device-resident loop (fullmag_fdm_backend_advance); all state is in HBM when
the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cpu-seconds S]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s (RK45, exch+demag) at 512²×1, fp32 vs fp64 at stated error"
NX = NY = 512
DX = DY = 2e-9
DZ = 5e-9
MS, AEX, ALPHA = 8.6e5, 1.3e-11, 0.02
MU0 = 4e-7 * np.pi
HEXT = (0.01 / MU0, 0.0, 0.0)
HBM_PEAK_GBS = 8000.0


def workload():
    from fullmag_amd.backend import hole_mask, random_unit_vectors
    mask = hole_mask(NX, NY, 1, DX, DY, 64 * DX)
    m0 = random_unit_vectors(42, NX * NY)
    m0[mask == 0] = 0.0
    return mask, m0


def make_backend(precision, mask, m0):
    from fullmag_amd.backend import FdmBackend, FdmPlan
    plan = FdmPlan(NX, NY, 1, DX, DY, DZ, MS, AEX, ALPHA, m0=m0, precision=precision, integrator="dp45",
                   h_ext=HEXT, mask=mask, atol=1e-5, dt_min=1e-18, dt_max=1e-10, headroom=0.8)
    return FdmBackend(plan)


def kernel_bytes(s):
    """Algorithmic HBM bytes of one launch of each kernel kind at stage 2
    (scalar size s), thin film: spectrum = 3 (nx+1) ny complex values,
    tensor = 4 (nx+1)(ny+1) reals, N cells.  Kind 6 (stage 2 inverse joined
    with stage 3 forward) does not re-read k2, which it produced itself."""
    n = NX * NY
    spec = 3 * (NX + 1) * NY * 2 * s
    tens = 4 * (NX + 1) * (NY + 1) * s
    return {
        0: 3 * s * n * 3 + spec,          # x-forward: m0 + k1 in, y out, spectrum out
        1: 2 * spec + tens,               # y pass: spectrum in/out + tensor
        4: spec + 3 * s * n * 2 + n,      # x-inverse + local: spectrum + y in, mask in, k out
        6: 2 * spec + 3 * s * n * 5 + n,  # joined: spectrum in/out, y2 + m0 + k1 in, mask, k2 + y3 out
    }


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# on this workload (FP32, stage 2; FETCH_SIZE scaled by 2.0, WRITE_SIZE by
# 1.0, the factors a plain 256 MiB device copy gives); the passes and their
# commands are recorded in profiles/film512_dp45_hbm_traffic.md
MEASURED_TRAFFIC_FP32 = {0: 16.01e6, 1: 17.53e6, 4: 13.71e6, 6: 29.87e6}

# launches of each kind in one DP45 trial step (stages 2..7)
LAUNCHES_PER_STEP = {0: 1, 1: 6, 4: 1, 6: 5}
NAMES = {0: "k_fwd_x", 1: "k_yf", 4: "k_inv_x", 6: "k_xx"}


def step_bytes(s):
    """Algorithmic HBM bytes of one accepted DP45 step (6 RHS with FSAL):
    stage combinations read m0 and 15 k-vectors from HBM in total (the k of
    the previous stage is produced in the same launch) and write 6 states;
    6 x (x-forward spectrum write, y pass read+write+tensor, x-inverse
    spectrum read, state read, mask read, k write); the error norm reads 5
    k-vectors."""
    n = NX * NY
    spec = 3 * (NX + 1) * NY * 2 * s
    tens = 4 * (NX + 1) * (NY + 1) * s
    return (6 + 15) * 3 * s * n + 6 * 3 * s * n + 24 * spec + 6 * tens + 6 * (6 * s * n + n) + 15 * s * n


def run_gpu(precision, mask, m0, steps, warmup):
    be = make_backend(precision, mask, m0)
    dt = 1e-13
    r = be.advance(float("inf"), warmup, dt) if warmup else {"suggested_next_dt": dt}
    dt = r["suggested_next_dt"] or dt
    c0 = be.counters()
    t0 = time.perf_counter()
    r = be.advance(float("inf"), steps, dt)
    wall = time.perf_counter() - t0
    c1 = be.counters()
    assert r["steps_taken"] == steps, r
    kt = be.time_kernels(reps=50)
    m = be.field("m")
    be.close()
    return {
        "wall": wall,
        "ms_per_step": 1e3 * wall / steps,
        "value": NX * NY * steps / wall / 1e6,
        "rejected": c1["rejected"] - c0["rejected"],
        "sim_time_s": c1["time"] - c0["time"],
        "kernel_us": kt,
        "norm_dev": float(np.abs(np.linalg.norm(m[mask == 1], axis=1) - 1.0).max()),
    }


def accuracy(mask, m0, t_end=20e-12):
    """max |m_fp32 - m_fp64| at the same simulated time (each precision with
    its own adaptive dt sequence)."""
    out = {}
    for p in ("single", "double"):
        be = make_backend(p, mask, m0)
        r = be.advance(t_end, 10 ** 7, 1e-13)
        out[p] = (be.field("m"), r["steps_taken"], r["time_seconds"])
        be.close()
    dm = float(np.abs(out["single"][0] - out["double"][0]).max())
    return {"t_end_s": t_end, "max_dm_fp32_vs_fp64": dm, "steps_fp32": out["single"][1],
            "steps_fp64": out["double"][1]}


def cpu_baseline(mask, m0, seconds):
    """The FP64 CPU oracle (restatement of the reference's fdm_cpu_reference
    RK45 path) on the same workload for a bounded number of accepted steps."""
    from oracle.oracle import Oracle, Problem, num_threads
    orc = Oracle(Problem(NX, NY, 1, DX, DY, DZ, MS, AEX, ALPHA, h_ext=HEXT, mask=mask, atol=1e-5))
    m = orc.prepare(m0)
    t, dt = 0.0, 1e-13
    t, rep = orc.step("rk45", m, t, dt)  # first step also evaluates k1 (no FSAL yet)
    dt = rep["dt_next"]
    steps, t0 = 0, time.perf_counter()
    while True:
        t, rep = orc.step("rk45", m, t, dt)
        dt = rep["dt_next"]
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 200:
            break
    return {"value": NX * NY * steps / el / 1e6, "unit": "Mcell-steps/s", "cores": num_threads(), "kind": "port",
            "sample": f"{steps} accepted RK45 steps (after 1 untimed) of the same 512x512x1 film, FP64, "
                      f"{el:.1f} s wall"}


def copy_calibration(nbytes=256 << 20, reps=4):
    """Plain device-to-device copies of a known size: the hardware counters
    of these dispatches (KiB fetched / written per launch) against the known
    byte counts give the scale applied to the engine kernels' counters."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 plain device copies of 256 MiB (scale for FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported by this engine "
              "(a single 512x512x1 system runs on one device)", file=sys.stderr)
        sys.exit(2)
    from fullmag_amd.backend import is_available
    if not is_available():
        raise SystemExit("no gfx9 GPU visible to libfullmag_fdm.so")
    mask, m0 = workload()
    r32 = run_gpu("single", mask, m0, a.steps, a.warmup)
    r64 = run_gpu("double", mask, m0, a.steps, a.warmup)
    kb = kernel_bytes(4)
    kt = r32["kernel_us"]
    # dominant kernel: most device time per trial step
    dom = max(LAUNCHES_PER_STEP, key=lambda k: kt[k] * LAUNCHES_PER_STEP[k])
    us = kt[dom]
    achieved = kb[dom] / (us * 1e-6) / 1e9
    roof = {"bound": "hbm", "kernel": NAMES[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC_FP32.get(dom),
            "bytes_per_launch": kb[dom], "avg_launch_us": round(us, 3),
            "per_kernel_us": {NAMES[k]: round(kt[k], 3) for k in LAUNCHES_PER_STEP},
            "launches_per_step": {NAMES[k]: v for k, v in LAUNCHES_PER_STEP.items()},
            "controller_us": round(kt[5], 3)}
    sb32 = step_bytes(4)
    out = {
        "metric": METRIC,
        "value": round(r32["value"], 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "thin film 512x512x1, 2x2x5 nm cells, r=64-cell hole, exch+Newell demag+Zeeman "
                               "(10 mT x), DP45 atol 1e-5, seed-42 random m0",
                   "cells": NX * NY, "active_cells": int(mask.sum()), "integrator": "dp45",
                   "parallelism": "single GPU"},
        "fp64": {"value": round(r64["value"], 2), "ms_per_step": round(r64["ms_per_step"], 5),
                 "rejected": r64["rejected"]},
        "fp32_over_fp64": round(r32["value"] / r64["value"], 3),
        "rejected_fp32": r32["rejected"],
        "sim_time_fp32_s": r32["sim_time_s"],
        "norm_deviation_fp32": r32["norm_dev"],
        "step_bytes_model_fp32": sb32,
        "step_effective_GBs_fp32": round(sb32 / (r32["ms_per_step"] * 1e-3) / 1e9, 1),
        "roofline": roof,
    }
    if not a.no_accuracy:
        out["accuracy"] = accuracy(mask, m0)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mask, m0, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
