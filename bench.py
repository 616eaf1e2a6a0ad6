"""Throughput of the MI355X FDM engine (one JSON line).

Default workload (film512): 512x512x1 cells of 2x2x5 nm, circular
non-magnetic hole of radius 64 cells at the centre (249,252 active of
262,144 cells), Ms 8.6e5 A/m, A 1.3e-11 J/m, alpha 0.02, B = (10 mT, 0, 0),
seeded (42) random-unit m0, exchange + Newell demag + Zeeman,
Dormand-Prince RK45 with atol 1e-5, dt0 1e-13, dt in [1e-18, 1e-10],
headroom 0.8.

--workload granular3d: 256x256x64 cells of 3 nm cubes, Ms 1e6 A/m,
A 1.5e-11 J/m, Ku 5e5 J/m^3 with per-grain axes (512 seeded Voronoi grains),
alpha 0.1, B ramp 0 -> 1 T along z over 1 ns, same integrator settings.

One step = one accepted RK45 step of the whole grid (6 RHS evaluations with
FSAL, rejected trials included in the time).  The steps run in the engine's
device-resident loop (fullmag_fdm_backend_advance); all state is in HBM when
the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload film512|granular3d]
                    [--no-cpu] [--cpu-seconds S] [--no-accuracy]
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
METRIC_3D = "Mcell-steps/s (RK45, exch+demag+anisotropy+ramp) at 256²×64, fp32 vs fp64 at stated error"
HBM_PEAK_GBS = 8000.0

# Measured HBM traffic per launch: the rocprofv3 FETCH_SIZE / WRITE_SIZE
# summaries under profiles/ (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0, the
# factors of a plain 256 MiB device copy), read from the "total MB" column of
# the FP32 stage-2 instantiation of each kernel kind.  They are the values
# recorded when those profiles were taken; bench.py cannot read hardware
# counters itself.
TRAFFIC_PROFILES = {"film512": "film512_dp45_hbm_traffic.md", "granular3d": "granular3d_dp45_profile.md"}
TRAFFIC_ROWS = {0: [("k_fwd_x<float", ", 1>`")], 1: [("k_yf8<float", ">`"), ("k_yf<float", ">`")], 2: [("k_zf<float", ">`")],
                4: [("k_inv_x<float", ">`")], 6: [("k_xxp<float", ", 2>`"), ("k_xx<float", ", 2>`")]}


def recorded_traffic(workload, kind):
    """(bytes per launch, source) from the committed profile summary, or
    (None, reason)."""
    name = TRAFFIC_PROFILES.get(workload)
    rows = TRAFFIC_ROWS.get(kind)
    if not name or not rows:
        return None, "no recorded profile"
    path = os.path.join(ROOT, "profiles", name)
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None, f"profiles/{name} missing"
    for row in rows:
        for ln in lines:
            cells = [c.strip() for c in ln.strip().strip("|").split("|")]
            if cells and cells[0].startswith("`" + row[0]) and cells[0].endswith(row[1]) and "KiB" not in cells[0]:
                try:
                    return float(cells[-1]) * 1e6, (f"profiles/{name} (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE of "
                                                    f"{cells[0].strip('`')}, recorded)")
                except ValueError:
                    continue
    return None, f"no row for {rows[0][0]} in profiles/{name}"


NAMES_FILM = {0: "k_fwd_x", 1: "k_yf8 (FP32 film y pass; k_yf in FP64)", 4: "k_inv_x (final stage)",
              6: "k_xxp (FP32 film joined x pass; k_xx in FP64)"}
NAMES_3D = {0: "k_fwd_x", 1: "k_y (forward)", 2: "k_z", 3: "k_y (inverse)", 4: "k_inv_x (final stage)", 6: "k_xx"}


class Workload:
    """One benchmark problem: plan arguments, launch counts per DP45 trial
    step and the algorithmic bytes of each kernel kind."""

    def __init__(self, name):
        from fullmag_amd import workloads
        self.name = name
        if name == "film512":
            self.kw = workloads.film_with_hole(42, 512)
            self.metric = METRIC
            self.desc = ("thin film 512x512x1, 2x2x5 nm cells, r=64-cell hole, exch+Newell demag+Zeeman "
                         "(10 mT x), DP45 atol 1e-5, seed-42 random m0")
        elif name == "granular3d":
            self.kw = workloads.granular_block()
            self.metric = METRIC_3D
            self.desc = ("3-D granular block 256x256x64 of 3 nm cubes, 512 Voronoi grains with per-grain "
                         "uniaxial axes (Ku 5e5), exch+Newell demag, B ramp 0->1 T along z over 1 ns, "
                         "DP45 atol 1e-5, seed-7 random m0")
        else:
            raise SystemExit(f"unknown workload {name}")
        self.nx, self.ny, self.nz = self.kw["nx"], self.kw["ny"], self.kw["nz"]
        self.n = self.nx * self.ny * self.nz
        self.film = self.nz == 1
        mask = self.kw.get("mask")
        self.active = int(mask.sum()) if mask is not None else self.n
        self.names = NAMES_FILM if self.film else NAMES_3D
        # launches of each kind in one DP45 trial step (stages 2..7, stage s
        # joined with s+1 for s < 7)
        self.launches = {0: 1, 1: 6, 4: 1, 6: 5} if self.film else {0: 1, 1: 6, 2: 6, 3: 6, 4: 1, 6: 5}

    def backend(self, precision):
        from fullmag_amd.backend import FdmBackend, FdmPlan
        return FdmBackend(FdmPlan(precision=precision, **self.kw))

    def kernel_bytes(self, s):
        """Algorithmic HBM bytes of one launch of each kernel kind at stage 2
        (scalar size s).  Spectrum with ny rows = 3 nz (nx+1) ny complex,
        with the padded 2ny rows twice that; film tensor 4 (nx+1)(ny+1)
        reals, 3-D tensor 6 (nx+1)(ny+1)(nz+1).  The local pass reads the
        stage state, the mask (1 B/cell, if any) and per-cell axes (if any)
        and writes k; kind 6 also reads m0 and k1 for stage 3 but not k2,
        which it produced itself."""
        nx, ny, nz, n = self.nx, self.ny, self.nz, self.n
        half = 3 * nz * (nx + 1) * ny * 2 * s
        full = 2 * half
        mask = n if self.kw.get("mask") is not None else 0
        axes = 3 * s * n if self.kw.get("axis_field") is not None else 0
        local_in = 3 * s * n + mask + axes
        out = {0: 9 * s * n + half, 4: half + local_in + 3 * s * n,
               6: 2 * half + local_in + 3 * s * n + 9 * s * n}
        if self.film:
            out[1] = 2 * half + 4 * (nx + 1) * (ny + 1) * s
        else:
            out[1] = half + full
            out[2] = 2 * full + 6 * (nx + 1) * (ny + 1) * (nz + 1) * s
            out[3] = full + half
        return out

    def step_bytes(self, s):
        """Algorithmic HBM bytes of one accepted DP45 step (FSAL, no
        rejections): the launches of one trial step, plus the stage-7
        combination's extra k reads (k1, k3..k5 beyond the stage-2 form) and
        the error norm's 5 k-vectors, minus nothing else."""
        kb = self.kernel_bytes(s)
        n = self.n
        tot = sum(kb[k] * c for k, c in self.launches.items())
        # kind 6 and 0 are priced with one or two k reads; the stage
        # combinations of DP45 read 20 k in total, 5 of them produced in the
        # same launch: 15 HBM k reads against 1 (kind 0) + 5 (kind 6) priced
        tot += (15 - 6) * 3 * s * n
        tot += 5 * 3 * s * n  # error norm in the final stage
        return tot


def run_gpu(wl, precision, steps, warmup):
    be = wl.backend(precision)
    dt = 1e-13
    r = be.advance(float("inf"), warmup, dt) if warmup else {"suggested_next_dt": dt}
    dt = r["suggested_next_dt"] or dt
    c0 = be.counters()
    t0 = time.perf_counter()
    r = be.advance(float("inf"), steps, dt)
    wall = time.perf_counter() - t0
    c1 = be.counters()
    assert r["steps_taken"] == steps, r
    kt = be.time_kernels(reps=20 if wl.n > 10 ** 6 else 50)
    ktt = be.time_trial_kernels(trials=10 if wl.n > 10 ** 6 else 50)
    m = be.field("m")
    be.close()
    mask = wl.kw.get("mask")
    act = m if mask is None else m[mask == 1]
    return {
        "wall": wall,
        "ms_per_step": 1e3 * wall / steps,
        "value": wl.n * steps / wall / 1e6,
        "rejected": c1["rejected"] - c0["rejected"],
        "sim_time_s": c1["time"] - c0["time"],
        "kernel_us": kt,
        "trial_kernel_us": ktt,
        "norm_dev": float(np.abs(np.linalg.norm(act, axis=1) - 1.0).max()),
    }


def accuracy(wl, horizons, oracle_horizon):
    """Error of the FP32 engine against FP64 at stated times.

    matched_dt: the FP64 engine runs with its own controller, and its
    accepted dt sequence is replayed (atol 1: every step accepted, so every
    dt is taken as given) by the FP32 engine, by the FP64 engine from m0
    perturbed by 1e-12 (the control: how fast a roundoff-sized difference
    grows on this trajectory) and, up to `oracle_horizon`, by the FP64 CPU
    oracle (the parity statement of the FP64 form); max_i |dm| and the
    relative total-energy error at each horizon.
    independent_dt: each precision with its own adaptive dt sequence."""
    from fullmag_amd.backend import FdmBackend, FdmPlan
    kw = dict(wl.kw)
    be64 = FdmBackend(FdmPlan(precision="double", **kw))
    dts, ref, dt = [], [], 1e-13
    for h in horizons:
        while True:
            r = be64.advance(h, 1, dt)
            if r["steps_taken"] == 0:
                break
            dts.append(r["dt_seconds"])
            dt = r["suggested_next_dt"] or dt
            if r["time_seconds"] >= h:
                break
        ref.append((len(dts), be64.field("m"), r["total_energy_joules"], r["time_seconds"]))
    be64.close()

    def replay(precision, m0, upto):
        b = FdmBackend(FdmPlan(**{**kw, "precision": precision, "atol": 1.0, "m0": m0}))
        out, done = [], 0
        for nsteps, _, _, _ in ref[:upto]:
            st = None
            for d in dts[done:nsteps]:
                st = b.step(d)
            done = nsteps
            out.append((b.field("m"), st["total_energy_joules"]))
        b.close()
        return out

    rng = np.random.default_rng(5)
    mp = kw["m0"] + 1e-12 * rng.standard_normal(kw["m0"].shape)
    r32 = replay("single", kw["m0"], len(ref))
    rpt = replay("double", mp, len(ref))
    rows = []
    for i, (nsteps, m64, e64, t) in enumerate(ref):
        rows.append({"t_s": horizons[i], "steps": nsteps,
                     "fp32_vs_fp64_max_dm": float(np.abs(r32[i][0] - m64).max()),
                     "fp32_vs_fp64_energy_rel": float(abs(r32[i][1] - e64) / abs(e64)),
                     "fp64_perturbed_1e-12_vs_fp64_max_dm": float(np.abs(rpt[i][0] - m64).max())})
    # the FP64 engine against the CPU oracle on the same dt sequence
    k = max(i for i, h in enumerate(horizons) if h <= oracle_horizon) if oracle_horizon >= horizons[0] else -1
    if k >= 0:
        from oracle.oracle import Oracle, Problem
        orc = Oracle(Problem(kw["nx"], kw["ny"], kw["nz"], kw["dx"], kw["dy"], kw["dz"], kw["ms"], kw["a_ex"],
                             kw["alpha"], h_ext=kw.get("h_ext"), mask=kw.get("mask"), atol=1.0))
        m = orc.prepare(kw["m0"])
        t, done = 0.0, 0
        t0 = time.perf_counter()
        for i in range(k + 1):
            for d in dts[done:ref[i][0]]:
                t, rep = orc.step("rk45", m, t, d)
            done = ref[i][0]
            rows[i]["fp64_vs_oracle_max_dm"] = float(np.abs(m - ref[i][1]).max())
            rows[i]["fp64_vs_oracle_energy_rel"] = float(abs(rep["e_total"] - ref[i][2]) / abs(rep["e_total"]))
        rows[k]["oracle_cpu_s"] = round(time.perf_counter() - t0, 1)
    indep = []
    runs = {}
    for p in ("single", "double"):
        be = wl.backend(p)
        dt, res = 1e-13, []
        for t_end in horizons:
            r = be.advance(t_end, 10 ** 7, dt)
            dt = r["suggested_next_dt"] or dt
            res.append((be.field("m"), r["total_energy_joules"]))
        runs[p] = res
        be.close()
    for i, t_end in enumerate(horizons):
        (m32, e32), (m64, e64) = runs["single"][i], runs["double"][i]
        indep.append({"t_s": t_end, "max_dm_fp32_vs_fp64": float(np.abs(m32 - m64).max()),
                      "energy_rel_err_fp32_vs_fp64": float(abs(e32 - e64) / abs(e64))})
    return {"dt_source": "the FP64 engine's own DP45 controller (atol 1e-5); replays take each dt as given",
            "matched_dt": rows, "independent_dt": indep}


def physical_cores():
    """Physical cores of the host (distinct core, socket pairs of lscpu -p),
    or None when lscpu is unavailable."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {tuple(ln.split(",")) for ln in out.splitlines() if ln and not ln.startswith("#")}
    return len(pairs) or None


def cpu_baseline(wl, seconds):
    """The FP64 CPU oracle (restatement of the reference's fdm_cpu_reference
    RK45 path) on the same workload for a bounded number of accepted steps,
    with all host threads and with one thread."""
    from oracle.oracle import Oracle, Problem, num_threads, set_num_threads
    kw = wl.kw
    orc = Oracle(Problem(kw["nx"], kw["ny"], kw["nz"], kw["dx"], kw["dy"], kw["dz"], kw["ms"], kw["a_ex"],
                         kw["alpha"], h_ext=kw.get("h_ext"), mask=kw.get("mask"), atol=kw["atol"]))
    m = orc.prepare(kw["m0"])
    t, dt = 0.0, 1e-13
    t, rep = orc.step("rk45", m, t, dt)  # first step also evaluates k1 (no FSAL yet)
    dt = rep["dt_next"]

    def timed(limit_s, max_steps):
        nonlocal t, dt
        steps, t0 = 0, time.perf_counter()
        while True:
            t, rep = orc.step("rk45", m, t, dt)
            dt = rep["dt_next"]
            steps += 1
            el = time.perf_counter() - t0
            if el >= limit_s or steps >= max_steps:
                return steps, el

    threads = num_threads()
    phys = physical_cores()
    steps, el = timed(seconds, 200)
    set_num_threads(1)
    steps1, el1 = timed(seconds / 3, 50)
    set_num_threads(threads)
    return {"value": wl.n * steps / el / 1e6, "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} accepted RK45 steps (after 1 untimed) of the same {wl.nx}x{wl.ny}x{wl.nz} "
                      f"problem, FP64, {el:.1f} s wall on {threads} threads",
            "host_physical_cores": phys,
            "threads_note": "OpenMP threads = the process's CPU share on the GPU host (OMP_NUM_THREADS); "
                            "host_physical_cores counts the whole host",
            "single_thread_value": wl.n * steps1 / el1 / 1e6,
            "single_thread_sample": f"{steps1} further steps on 1 thread, {el1:.1f} s wall"}


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


def adjoint_bench(steps, fd_params, fd_h=1e-3):
    """Config-5 waveguide: forward cost, gradient cost (forward + reverse
    sweep) and, optionally, central differences of the FP64 forward on
    seeded parameters (fd_params - 1 design cells and the pulse amplitude)."""
    from fullmag_amd.adjoint import WaveguideProblem
    rng = np.random.default_rng(2024)
    base = WaveguideProblem(steps=steps, window=min(0.5e-9, steps * 50e-15 / 2))
    design = rng.uniform(6e5, 1e6, size=base.design_shape())
    amp = 5e-3
    out = {"workload": "waveguide 256x64x2 of 5 nm cells, 64x64 design Ms map U[6e5,1e6], sinc pulse 5 mT "
                       f"20 GHz on 8 columns, RK4 dt 50 fs x {steps} steps, J = mean m_z^2 over the last 8 "
                       "columns in the last 0.5 ns",
           "cells": base.n}
    grads = {}
    for prec in ("single", "double"):
        pr = WaveguideProblem(steps=steps, window=base.window, precision=prec)
        warm = WaveguideProblem(steps=2, window=base.dt, precision=prec)
        warm.value_and_grad(design, amp)  # module load, LDS limits, graph capture
        t0 = time.perf_counter()
        j_f, be = pr.run(design, amp)
        t_fwd = time.perf_counter() - t0
        be.close()
        t0 = time.perf_counter()
        j, gd, ga = pr.value_and_grad(design, amp)
        t_grad = time.perf_counter() - t0
        grads[prec] = (j, gd, ga)
        # default record budget (half the free HBM) holds the whole run
        # here (7 (nstage - 1) N values per step: 55 GB FP32, 110 GB FP64);
        # a 1-byte budget keeps only the last sqrt(N) segment and recomputes
        # every other one from its checkpoint
        sq = WaveguideProblem(steps=steps, window=base.window, precision=prec, record_budget_bytes=1)
        t0 = time.perf_counter()
        j_s, gd_s, ga_s = sq.value_and_grad(design, amp)
        t_sq = time.perf_counter() - t0
        out[prec] = {"objective": j, "forward_s": round(t_fwd, 3), "gradient_s": round(t_grad, 3),
                     "gradient_over_forward": round(t_grad / t_fwd, 3),
                     "forward_Mcell_steps_per_s": round(pr.n * steps / t_fwd / 1e6, 2),
                     "checkpoint_interval": int(np.ceil(np.sqrt(steps))),
                     "record": "default budget (half the free device memory): every segment recorded by the forward",
                     "recompute_gradient_s": round(t_sq, 3),
                     "recompute_gradient_over_forward": round(t_sq / t_fwd, 3),
                     "recompute_note": "last segment recorded, the other sqrt(N) segments recomputed",
                     "recompute_equals_recorded": bool(j_s == j and np.array_equal(gd_s, gd) and ga_s == ga)}
        if prec == "single":
            st = WaveguideProblem(steps=steps, window=base.window, precision=prec, beta_state_precision=True)
            grads["single_beta32"] = st.value_and_grad(design, amp)
    if fd_params:
        pr = WaveguideProblem(steps=steps, window=base.window, precision="double")
        dy, dx = pr.design_shape()
        cells = rng.choice(dy * dx, size=fd_params - 1, replace=False)
        fd, adj32, adj64, adj32b = [], [], [], []
        for f in cells:
            iy, ix = int(f // dx), int(f % dx)
            p, m = design.copy(), design.copy()
            p[iy, ix] *= 1 + fd_h
            m[iy, ix] *= 1 - fd_h
            jp, b = pr.run(p, amp)
            b.close()
            jm, b = pr.run(m, amp)
            b.close()
            fd.append((jp - jm) / (2 * fd_h * design[iy, ix]))
            adj32.append(grads["single"][1][iy, ix])
            adj64.append(grads["double"][1][iy, ix])
            adj32b.append(grads["single_beta32"][1][iy, ix])
        jp, b = pr.run(design, amp * (1 + fd_h))
        b.close()
        jm, b = pr.run(design, amp * (1 - fd_h))
        b.close()
        fd.append((jp - jm) / (2 * fd_h * amp))
        adj32.append(grads["single"][2])
        adj64.append(grads["double"][2])
        adj32b.append(grads["single_beta32"][2])
        fd, adj32, adj64, adj32b = map(np.array, (fd, adj32, adj64, adj32b))
        sc = np.abs(fd).max()
        out["finite_differences"] = {
            "params": fd_params, "h_rel": fd_h, "fd_from": "FP64 engine forward",
            "max_rel_err_fp64_adjoint": float(np.abs(adj64 - fd).max() / sc),
            "max_rel_err_fp32_adjoint": float(np.abs(adj32 - fd).max() / sc),
            "grad_scale": float(sc),
            "per_param_rel_err_fp32": [float(v) for v in np.abs(adj32 - fd) / np.abs(fd)],
            "per_param_rel_err_fp64": [float(v) for v in np.abs(adj64 - fd) / np.abs(fd)],
            "max_rel_err_fp32_adjoint_fp32_beta": float(np.abs(adj32b - fd).max() / sc),
            "per_param_rel_err_fp32_fp32_beta": [float(v) for v in np.abs(adj32b - fd) / np.abs(fd)],
            "fp32_note": "FP32 adjoint: field cotangent convolved in FP64 (default); *_fp32_beta: in FP32",
            "fd": [float(v) for v in fd]}
    return out


def batch_bench(count, t_end, sequential):
    """Config-4 sweep: `count` independent FP32 systems (sizes drawn by seed
    from 64x64x1, 128x64x2, 256x128x1, 256x256x4 cells of 4 nm; Ms, A, alpha
    drawn per system; random m0), each integrated to t_end with its own
    DP45 dt, all in one batch (groups of equal shape share launches)."""
    from fullmag_amd.backend import FdmBackend, FdmBatch, FdmPlan
    from fullmag_amd.workloads import batch_sweep
    plans = [FdmPlan(**kw) for kw in batch_sweep(count=count, seed=4)]
    cells = np.array([p.nx * p.ny * p.nz for p in plans])
    with FdmBatch(plans) as b:
        groups = b.groups()
        b.advance(1e-14, 2, 1e-13)  # warm-up: graph capture, module load
        t0 = time.perf_counter()
        wall, steps = b.advance(t_end + 1e-14, 10 ** 7, 1e-13)
        total = time.perf_counter() - t0
    out = {"systems": count, "groups": groups, "t_end_s": t_end, "wall_s": round(total, 4),
           "cells_total": int(cells.sum()), "accepted_steps_total": int(steps.sum()),
           "value": float((cells * steps).sum() / total / 1e6),
           "per_system_time_s": {"p50": float(np.percentile(wall, 50)), "p99": float(np.percentile(wall, 99)),
                                 "max": float(wall.max())},
           "steps_per_system": {"min": int(steps.min()), "p50": int(np.median(steps)), "max": int(steps.max())}}
    if sequential:
        seq = 0.0
        for p in plans:
            with FdmBackend(p) as be:
                be.advance(1e-14, 2, 1e-13)
                t0 = time.perf_counter()
                be.advance(t_end + 1e-14, 10 ** 7, 1e-13)
                seq += time.perf_counter() - t0
        out["sequential_wall_s"] = round(seq, 4)
        out["sequential_Mcell_steps_per_s"] = float((cells * steps).sum() / seq / 1e6)
        out["batch_speedup"] = round(seq / total, 3)
    return out


def roofline(wl, r32):
    """Launch times inside trial steps (each kernel reading what the one
    before it wrote, as in the timed stepping loop: HIP events around every
    launch of whole trial steps run as separate launches) against the
    algorithmic bytes of the same launch; the back-to-back repetitions of one
    kernel (inputs still in L2) are reported beside them as isolated_us."""
    kb = wl.kernel_bytes(4)
    # the final x pass in a trial step runs at the last stage, whose error
    # norm also reads 5 k-vectors
    kb4 = kb[4] + 5 * 3 * 4 * wl.n
    kt = r32["trial_kernel_us"]
    ki = r32["kernel_us"]
    # dominant kernel: most device time per trial step
    dom = max(wl.launches, key=lambda k: kt[k] * wl.launches[k])
    us = kt[dom]
    bdom = kb4 if dom == 4 else kb[dom]
    achieved = bdom / (us * 1e-6) / 1e9
    per = {}
    for k in wl.launches:
        b = kb4 if k == 4 else kb[k]
        per[wl.names[k]] = {"us": round(kt[k], 3), "isolated_us": round(ki[k], 3), "launches_per_step": wl.launches[k],
                            "algorithmic_bytes": b, "GBs": round(b / (kt[k] * 1e-6) / 1e9, 1)}
    return {"bound": "hbm", "kernel": wl.names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": recorded_traffic(wl.name, dom)[0], "traffic_source": recorded_traffic(wl.name, dom)[1],
            "bytes_per_launch": bdom, "avg_launch_us": round(us, 3), "isolated_launch_us": round(ki[dom], 3),
            "timing": "launches inside trial steps (fullmag_fdm_backend_time_trial_kernels; stage 2, final x pass "
                      "at the last stage), HIP events on the engine's stream, less the events' own cost per launch",
            "trial_step_us": round(kt[7], 3), "event_cost_us": round(kt[8], 3),
            "per_kernel": per, "controller_us": round(kt[5], 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="film512", choices=["film512", "granular3d", "adjoint", "batch"])
    ap.add_argument("--batch-systems", type=int, default=96)
    ap.add_argument("--batch-t-end", type=float, default=10e-12)
    ap.add_argument("--no-sequential", action="store_true", help="batch workload: skip the one-at-a-time baseline")
    ap.add_argument("--adj-steps", type=int, default=20000)
    ap.add_argument("--fd-params", type=int, default=16, help="adjoint workload: parameters checked by FD (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--oracle-horizon", type=float, default=1e-12,
                    help="film: simulated time over which the CPU oracle replays the FP64 engine's dt sequence")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 plain device copies of 256 MiB (scale for FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported by this engine "
              "(one system runs on one device)", file=sys.stderr)
        sys.exit(2)
    from fullmag_amd.backend import is_available
    if not is_available():
        raise SystemExit("no gfx9 GPU visible to libfullmag_fdm.so")
    if a.workload == "batch":
        r = batch_bench(a.batch_systems, a.batch_t_end, not a.no_sequential)
        out = {"metric": "aggregate Mcell-steps/s of a batched sweep (RK45 each with own dt, fp32)",
               "value": round(r.pop("value"), 2), "unit": "Mcell-steps/s", "n_gpus": 1,
               "steps": r["accepted_steps_total"], "warmup": 2, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
               "config": {"workload": f"{a.batch_systems} independent systems, sizes by seed from 64x64x1, "
                                      "128x64x2, 256x128x1, 256x256x4 cells of 4 nm, Ms U[6e5,1.2e6], A U[1e-11,"
                                      "2e-11], alpha U[0.01,0.5], random m0, exch+demag, DP45 atol 1e-5, "
                                      f"to t = {a.batch_t_end:g} s", "parallelism": "single GPU"}}
        out.update(r)
        print(json.dumps(out), flush=True)
        return
    if a.workload == "adjoint":
        r = adjoint_bench(a.adj_steps, a.fd_params)
        out = {"metric": "adjoint gradient cost / forward cost (RK4, 256x64x2, fp32)",
               "value": r["single"]["gradient_over_forward"], "unit": "x", "n_gpus": 1, "steps": a.adj_steps,
               "warmup": 0, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
               "data": "synthetic", "config": {"workload": r.pop("workload"), "cells": r.pop("cells"),
                                               "parallelism": "single GPU"}}
        out.update(r)
        print(json.dumps(out), flush=True)
        return
    wl = Workload(a.workload)
    r32 = run_gpu(wl, "single", a.steps, a.warmup)
    r64 = run_gpu(wl, "double", a.steps, a.warmup)
    sb32 = wl.step_bytes(4)
    out = {
        "metric": wl.metric,
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
        "config": {"workload": wl.desc, "cells": wl.n, "active_cells": wl.active, "integrator": "dp45",
                   "parallelism": "single GPU"},
        "fp64": {"value": round(r64["value"], 2), "ms_per_step": round(r64["ms_per_step"], 5),
                 "rejected": r64["rejected"]},
        "fp32_over_fp64": round(r32["value"] / r64["value"], 3),
        "rejected_fp32": r32["rejected"],
        "sim_time_fp32_s": r32["sim_time_s"],
        "norm_deviation_fp32": r32["norm_dev"],
        "step_bytes_model_fp32": sb32,
        "step_effective_GBs_fp32": round(sb32 / (r32["ms_per_step"] * 1e-3) / 1e9, 1),
        "roofline": roofline(wl, r32),
    }
    if not a.no_accuracy:
        out["accuracy"] = accuracy(wl, [1e-12, 5e-12, 20e-12] if wl.film else [1e-12, 5e-12],
                                   a.oracle_horizon if (wl.film and not a.no_cpu) else 0.0)
    if not a.no_cpu and wl.film:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
