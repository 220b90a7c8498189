#!/usr/bin/env python3
"""Benchmark of the MI355X MHD step: one "step" is one full RK time step
(4 stages of ConsToPrim + MHD_Residual + Source_Integrate_Tcheck + ghost fill,
then TVDlimit, TCheck, ghost fill) over the stratified 256x512x512 box
(vertical 256, dx = 16 km), synthetic seeded state, inputs resident in HBM.

Prints one JSON line.  value = interior cells x steps / wall time
(Mcell-updates/s) of the plain FP64 step.  fp32_mixed: the same workload in
the FP32-mixed form (FP32 rho, M, B, p and stencils; FP64 e, e_int recovery,
dt).  precision_study: both forms run 100 steps from one state of the
3 km/s box; per-field relative L2 / Linf error of the mixed form against the
FP64 form (which matches the oracle to 1e-12) and the Mcell/s of both.
roofline: the dominant kernel's algorithmic bytes per launch / its mean launch
time from HIP events on the library's stream during the timed steps.
cpu_baseline: the oracle (C restatement, FP64, OpenMP) on a bounded sample of
the same workload on this host's cores.
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

METRIC = "Mcell-updates/s per RK step + % HBM peak; FP64 plain vs FP32-mixed at stated error"
HBM_PEAK_GBS = 8000.0
MODEL_BYTES = {"fp64": 1400.0, "fp32": 788.0}   # algorithmic HBM bytes per cell-update

# algorithmic bytes per unit of each kernel; unit = interior cell unless noted
KERNEL_BYTES = {
    "cons_to_prim": (112.0, "buffer"),    # read rho,M,e 40; write M,e 32 + p,T,ne,rhoi,amb 40
    # MHD_Residual after a Source_Integrate_Tcheck that recorded Res := 0:
    # read U 64 + p 8, write Res 64 (the zeros are not read)
    "residual": (136.0, "interior"),
    # Source_Integrate_Tcheck: read U, U0, Res 192; write U 64 (Res := 0 recorded)
    "source_integrate": (256.0, "interior"),
    "save_cons": (128.0, "interior"),
    "tcheck": (80.0, "interior"),
    "tvd_sweep_x": (136.0, "interior"),   # read U 64 + p 8; write U 64
    "tvd_sweep_y": (136.0, "interior"),
    "tvd_sweep_z": (136.0, "interior"),
    "eos_p": (72.0, "buffer"),
}
# fused path: the stage kernel reads U_in 64 + p 8 + U0 64 and writes
# U_out 64 + p 8 = 208 B; stage 1 reads U0 as its input (144 B), stage 4
# writes no pressure (200 B).  The timers average over the four launches of
# a step, so the bytes per launch are the mean over the stages.
STAGE_BYTES = {"fp64": (144.0, 208.0, 208.0, 200.0), "fp32": (80.0, 116.0, 116.0, 112.0)}
KERNEL_BYTES_FUSED = dict(KERNEL_BYTES, tvd_sweep_z=(144.0, "interior"))   # + new p (epilogue)
# drop-in sequence with the deferred stage: MHD_Residual forms ConsToPrim's
# pressure and the speeds for dt (read U 64, write p 8 per buffer cell; the
# other EOS outputs wait for a reader) and Source_Integrate_Tcheck runs the
# stage kernel from the conservative state (read U 64 + p 8 + U0 64, write
# U 64 = 200 B; at stage 1 U0 is U itself, 136 B); TCheck runs as the
# epilogue of TVDlimit's dim-2 sweep (136 B)
KERNEL_BYTES_DROPIN = dict(KERNEL_BYTES, cons_to_prim=(72.0, "buffer"))
STAGE_BYTES_DROPIN = (136.0, 200.0, 200.0, 200.0)
# FP32-mixed form: a cell is rho, M, B in FP32 (28 B) + e in FP64 (8 B) = 36 B, p 4 B
KERNEL_BYTES_MX = {
    "tvd_sweep_x": (76.0, "interior"),    # read U 36 + p 4; write U 36
    "tvd_sweep_y": (76.0, "interior"),
    "tvd_sweep_z": (80.0, "interior"),    # + new p 4
    "eos_p": (40.0, "buffer"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """model name, sockets, cores per socket, threads per core and the CPUs
    this process may run on"""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    info["usable_cpus"] = len(os.sched_getaffinity(0))
    return info


def oracle_run(table, phys, box, steps, threads, count=False):
    """the oracle (C restatement of the reference, FP64, OpenMP on `threads`
    threads) from the seeded state of box: one warm-up step, then `steps`
    timed steps.  Returns the oracle, its run descriptor, the dt of every
    step and the seconds per timed step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from muram_main_amd import abi, config, eos_table, initial
    from orc import Oracle
    tab = eos_table.load_tables(table)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    Oracle.lib().orc_set_threads(threads)
    o = Oracle(g, phys, table)
    r = config.run_desc()
    o.set_state(U)
    del U
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    if count:
        o.branch_count(1, abi.BRANCH_TIE)
    dts = []
    o.step(r, config.DT_RAD_OFF, 1)          # warm-up
    dts.append(r.dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(r, config.DT_RAD_OFF, 1)
        dts.append(r.dt)
    sec = (time.perf_counter() - t0) / steps
    return o, r, dts, sec


def gpu_parity(table, wl, o, ro, dts, path):
    """the GPU step (path: fused or sequence) from the same seeded state for
    as many steps as the oracle ran, with the limiter branch accounting on:
    per-field errors of the final state against the oracle's, the largest
    relative dt difference over the steps and the branch mismatch"""
    from muram_main_amd import abi, config, initial, eos_table
    os.environ["MRM_STEP_PATH"] = path
    try:
        U = initial.make_state(eos_table.load_tables(table), wl.box)
        m, r = make_grid(wl, table, U, abi.PREC_FP64)
    finally:
        os.environ["MRM_STEP_PATH"] = "fused"
    del U
    m.branch_count(True, abi.BRANCH_TIE)
    ddt = 0.0
    for dt_o in dts:
        m.step(r, config.DT_RAD_OFF, 1)
        ddt = max(ddt, abs(r.dt - dt_o) / dt_o)
    A = m.get_state()
    bg = m.branch_counts()
    m.close()
    B = o.state(0)
    errs, l2 = {}, {}
    inner = (slice(2, -2),) * 3
    for v, name in enumerate(FIELDS):
        a, b = A[v][inner], B[v][inner]
        d = a - b
        errs[name] = float(f"{float(np.max(np.abs(d)) / max(np.max(np.abs(b)), 1e-300)):.3e}")
        l2[name] = float(f"{float(np.sqrt(np.sum(d * d)) / max(np.sqrt(np.sum(b * b)), 1e-300)):.3e}")
    del A
    bo = o.branch_counts()
    return {"path": path, "steps": len(dts), "cells": "interior",
            "field_rel_err_max": errs, "max_rel_err": max(errs.values()),
            "field_rel_err_l2": l2, "max_rel_err_l2": max(l2.values()),
            "dt_rel_diff_max": float(f"{ddt:.3e}"), "branch_mismatch": abi.branch_mismatch(bg, bo),
            "branch_tie": abi.BRANCH_TIE, "branch_counts": abi.branch_dict(bo)["counts"]}


def cpu_baseline(table, wl, steps=3, parity_paths=("fused", "sequence")):
    """Oracle FP64 step (C restatement of the reference, OpenMP) on the same
    workload: all usable host cores on the full box (one warm-up, `steps`
    timed steps, branch accounting on), then one core on a 256x128x128
    sample of it; the GPU paths from the same state are compared with the
    oracle's state after the same steps (parity columns)."""
    from muram_main_amd import initial
    box = wl.box
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", cpu["usable_cpus"]))
    o, r, dts, sec = oracle_run(table, wl.phys, box, steps, threads, count=True)
    ncell = box.gsize[0] * box.gsize[1] * box.gsize[2]
    log(f"cpu baseline: {threads} threads, {sec:.2f} s/step")
    parity = [gpu_parity(table, wl, o, r, dts, p) for p in parity_paths]
    del o
    sub = initial.BoxSpec(gsize=(box.gsize[0], 128, 128), dx=box.dx, v_rms=box.v_rms, seed=box.seed)
    o1, _, _, sec1 = oracle_run(table, wl.phys, sub, steps, 1)
    del o1
    n1 = sub.gsize[0] * sub.gsize[1] * sub.gsize[2]
    log(f"cpu baseline: 1 thread, {sec1:.2f} s/step on {sub.gsize}")
    return dict(value=ncell / sec / 1e6, unit="Mcell-updates/s", cores=threads, kind="port",
                sample=f"oracle FP64 step on the same {box.gsize[0]}x{box.gsize[1]}x{box.gsize[2]} seeded state, "
                       f"{steps} timed steps after 1 warm-up, {threads} OpenMP threads",
                s_per_step=sec, host=cpu,
                one_core={"value": n1 / sec1 / 1e6, "unit": "Mcell-updates/s", "cores": 1,
                          "s_per_step": sec1,
                          "sample": f"{sub.gsize[0]}x{sub.gsize[1]}x{sub.gsize[2]} box of the same distributions, "
                                    f"{steps} timed steps after 1 warm-up"},
                parity=parity)


def timed_run(m, r, steps, warmup):
    """warm-up steps, then `steps` timed steps with the per-kernel timers of
    the volume kernels on (the surface kernels' timers are left off: each
    event record idles the GPU for a few microseconds)."""
    import ctypes as C
    import torch
    from muram_main_amd import abi, config
    for _ in range(warmup):
        m.step(r, config.DT_RAD_OFF, 1)
    m.synchronize()
    lib = m.lib
    lib.mrm_timers(m.h, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.step(r, config.DT_RAD_OFF, 1)
    m.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = (C.c_double * abi.NTIMER)()
    cnt = (C.c_long * abi.NTIMER)()
    lib.mrm_get_timers(m.h, ms, cnt)
    lib.mrm_timers(m.h, 0)
    kernels = {}
    for i in range(abi.NTIMER):
        if cnt[i]:
            kernels[lib.mrm_timer_name(i).decode()] = dict(total_ms=ms[i], launches=cnt[i], mean_ms=ms[i] / cnt[i])
    return wall, kernels


def roofline(kernels, table, ncell, nbuf, traffic_key, stage_bytes=None):
    """the dominant kernel's algorithmic bytes per launch over its mean launch
    time (HIP events on the library's stream), next to its HBM traffic per
    launch from the PMC counters in profiles/traffic.json"""
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    if dom == "stage_fused" and stage_bytes:
        per, unit = sum(stage_bytes) / len(stage_bytes), "interior"
    else:
        per, unit = table.get(dom, (0.0, "interior"))
    units = nbuf if unit == "buffer" else ncell
    bytes_launch = per * units
    achieved = bytes_launch / (kernels[dom]["mean_ms"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath)).get(traffic_key, {}).get(dom)
        if tr:
            traffic = tr.get("bytes_per_launch")
    out = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
               bytes_per_launch=bytes_launch, bytes_per_unit=per, units=units)
    if dom == "stage_fused" and stage_bytes:
        out["bytes_per_unit_by_stage"] = list(stage_bytes)
    if traffic:
        out["traffic_over_algorithmic"] = round(traffic / bytes_launch, 3)
    return out


def kernel_summary(kernels, wall):
    return {k: {"mean_ms": round(v["mean_ms"], 4), "launches": v["launches"],
                "share": round(v["total_ms"] / (wall * 1e3), 4)} for k, v in kernels.items()}


def make_grid(wl, table, U, precision):
    from muram_main_amd import config
    from muram_main_amd.mhd import MhdGrid
    g = config.grid_desc(wl.box.gsize, wl.box.dx)
    m = MhdGrid(g, wl.phys, precision=precision)
    m.eos_init(table)
    m.set_state(U)
    r = config.run_desc()
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    return m, r


FIELDS = ("rho", "M0", "M1", "M2", "e", "B0", "B1", "B2")


def _errs(A, B, sl):
    """per-field relative L2 (||a-b|| / ||b||) and Linf (max|a-b| / max|b|) over the slice"""
    out = {}
    for v, name in enumerate(FIELDS):
        a, b = A[sl][v], B[sl][v]
        d = a - b
        l2 = float(np.sqrt(np.sum(d * d)) / max(np.sqrt(np.sum(b * b)), 1e-300))
        li = float(np.max(np.abs(d)) / max(np.max(np.abs(b)), 1e-300))
        out[name] = {"l2": float(f"{l2:.3e}"), "linf": float(f"{li:.3e}")}
    return out


def _eint_stats(U):
    """min internal energy e - M^2/(2 rho) over the interior and the count of cells with e_int <= 0."""
    sl = (slice(2, -2),) * 3
    ei = U[4][sl] - 0.5 * (U[1][sl] ** 2 + U[2][sl] ** 2 + U[3][sl] ** 2) / U[0][sl]
    return {"min_eint": float(ei.min()), "n_nonpositive": int(np.count_nonzero(~(ei > 0)))}


def precision_study(wl, table, tab, nsteps, control=True, positivity=False):
    """FP64 and FP32-mixed forms from one state, nsteps steps each; per-field
    errors of the mixed state against the FP64 one over the interior, its
    lower and upper halves, and (control) the FP64 form started from the
    FP32-rounded state: the step's own sensitivity to storage rounding."""
    from muram_main_amd import abi, config, initial
    t = time.perf_counter()
    U = initial.make_state(tab, wl.box)
    log(f"precision study state {wl.box.gsize} in {time.perf_counter() - t:.1f}s")
    out = {"workload": wl.name, "steps": nsteps}
    states, dts = {}, {}
    ncell = wl.box.gsize[0] * wl.box.gsize[1] * wl.box.gsize[2]
    runs = [("fp64", abi.PREC_FP64), ("fp32_mixed", abi.PREC_FP32_MIXED)]
    if control:
        runs.append(("control", abi.PREC_FP64))
    for name, prec in runs:
        if name == "control":
            for v in (0, 1, 2, 3, 5, 6, 7):
                U[v] = U[v].astype(np.float32).astype(np.float64)
        m, r = make_grid(wl, table, U, prec)
        m.synchronize()
        t0 = time.perf_counter()
        m.step(r, config.DT_RAD_OFF, nsteps)
        m.synchronize()
        wall = time.perf_counter() - t0
        if name != "control":
            out[f"{name}_mcells"] = round(ncell * nsteps / wall / 1e6, 2)
        states[name] = m.get_state()
        dts[name] = r.dt
        m.close()
        log(f"precision study {name}: {nsteps} steps in {wall:.2f}s")
    del U
    n0 = wl.box.gsize[0]
    inner = (slice(2, -2),) * 2
    sl_all = (slice(None),) + inner + (slice(2, 2 + n0),)
    sl_lo = (slice(None),) + inner + (slice(2, 2 + n0 // 2),)
    sl_hi = (slice(None),) + inner + (slice(2 + n0 // 2, 2 + n0),)
    A, B = states["fp32_mixed"], states["fp64"]
    out["speedup"] = round(out["fp32_mixed_mcells"] / out["fp64_mcells"], 3)
    out["dt_rel_diff"] = abs(dts["fp32_mixed"] / dts["fp64"] - 1.0)
    out["errors"] = _errs(A, B, sl_all)
    out["max_l2"] = max(v["l2"] for v in out["errors"].values())
    out["max_linf"] = max(v["linf"] for v in out["errors"].values())
    out["errors_lower_half"] = _errs(A, B, sl_lo)
    out["errors_upper_half"] = _errs(A, B, sl_hi)
    out["max_l2_lower_half"] = max(v["l2"] for v in out["errors_lower_half"].values())
    if control:
        out["control_fp64_from_rounded_state"] = _errs(states["control"], B, sl_all)
        out["control_max_l2"] = max(v["l2"] for v in out["control_fp64_from_rounded_state"].values())
        out["control_lower_half"] = _errs(states["control"], B, sl_lo)
        out["control_max_l2_lower_half"] = max(v["l2"] for v in out["control_lower_half"].values())
        out["control_upper_half"] = _errs(states["control"], B, sl_hi)
    if positivity:
        out["eint_fp64"] = _eint_stats(B)
        out["eint_fp32_mixed"] = _eint_stats(A)
    out["reference"] = "FP64 form of this library (parity with the FP64 oracle <= 1e-12)"
    return out


def storage_floor_study(table, tab, nsteps=100, gsize=(256, 64, 64)):
    """The box512_v3 distributions on a 256 x 64 x 64 crop, nsteps steps: the
    FP32-mixed form, the FP64 form from the FP32-rounded state (control_init)
    and the FP64 form with rho, M and B rounded to FP32 after every step
    (control_each: FP32 storage with FP64 arithmetic), each against the FP64
    form; relative L2 per field over the lower and upper halves.  control_each
    is the floor that any FP32-storage scheme meets on this configuration;
    control_each_stage rounds after every RK stage and after TCheck as well,
    where the FP32-mixed form stores its state (the reference's stage loop
    through the drop-in operations, FP64 arithmetic throughout)."""
    from muram_main_amd import abi, config, initial
    from muram_main_amd.mhd import MhdGrid
    wl0 = config.workloads()["box512_v3"]
    box = initial.BoxSpec(gsize=gsize, dx=wl0.box.dx, v_rms=wl0.box.v_rms, seed=wl0.box.seed)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)

    def rnd(V):
        V = V.copy()
        for v in (0, 1, 2, 3, 5, 6, 7):
            V[v] = V[v].astype(np.float32).astype(np.float64)
        return V

    def run(U0, prec, each=False):
        m = MhdGrid(g, wl0.phys, prec)
        m.eos_init(table)
        r = config.run_desc()
        m.set_state(U0)
        m.exchange_grid()
        m.SetBoundaryConditions(r, 0)
        for _ in range(nsteps):
            m.step(r, config.DT_RAD_OFF, 1)
            if each:
                m.set_state(rnd(m.get_state()))
        out = m.get_state()
        m.close()
        return out

    def run_staged(U0, round_stages):
        """the reference's stage loop (solver.C:109-365) through the drop-in
        operations, with rho, M and B rounded to FP32 after each stage's
        update and after TCheck (round_stages), i.e. where the FP32-mixed
        form stores its state"""
        m = MhdGrid(g, wl0.phys, abi.PREC_FP64)
        m.eos_init(table)
        r = config.run_desc()
        m.set_state(U0)
        m.exchange_grid()
        m.SetBoundaryConditions(r, 0)
        for _ in range(nsteps):
            m.SaveCons()
            cfl = 0.0
            for stage in range(1, 5):
                m.ConsToPrim()
                if stage == 1:
                    m.Adjust_Valf_Max(r)
                dt_mhd = m.MHD_Residual(r)
                if stage == 1:
                    cfl = m.Sync_timestep(r, config.DT_RAD_OFF, dt_mhd)
                m.Source_Integrate_Tcheck(r, stage)
                if round_stages:
                    m.set_state(rnd(m.get_state()))
                m.exchange_grid()
                m.SetBoundaryConditions(r, stage)
            if cfl <= r.CFL_tvd:
                m.TVDlimit(r, r.dt)
            else:
                m.TVDlimit(r, 0.5 * r.dt)
                m.exchange_grid()
                m.TVDlimit(r, 0.5 * r.dt)
            m.TCheck(r)
            if round_stages:
                m.set_state(rnd(m.get_state()))
            m.exchange_grid()
            m.SetBoundaryConditions(r, 5)
            r.iteration += 1
            r.globiter += 1
            r.time += r.dt
        out = m.get_state()
        m.close()
        return out

    t = time.perf_counter()
    ref = run(U, abi.PREC_FP64)
    res = {"mixed": run(U, abi.PREC_FP32_MIXED), "control_init": run(rnd(U), abi.PREC_FP64),
           "control_each": run(rnd(U), abi.PREC_FP64, each=True),
           "control_each_stage": run_staged(rnd(U), True)}
    # the Python stage loop itself, unrounded, against mrm_step (a consistency check)
    seq = run_staged(U, False)
    n0 = gsize[0]
    inner = (slice(2, -2),) * 2
    sl = {"lower_half": (slice(None),) + inner + (slice(2, 2 + n0 // 2),),
          "upper_half": (slice(None),) + inner + (slice(2 + n0 // 2, 2 + n0),)}
    out = {"workload": f"box512_v3 distributions on {gsize[0]}x{gsize[1]}x{gsize[2]}", "steps": nsteps,
           "stage_loop_vs_mrm_step_max_l2": max(e["l2"] for e in _errs(seq, ref, (slice(None),) + inner +
                                                                        (slice(2, 2 + n0),)).values())}
    for h, s_ in sl.items():
        out[h] = {name: {f: e["l2"] for f, e in _errs(X, ref, s_).items()} for name, X in res.items()}
        out[h + "_max_l2"] = {name: max(v.values()) for name, v in out[h].items()}
    log(f"storage floor study in {time.perf_counter() - t:.1f}s")
    return out


def local_direction(U, inner, rng):
    """random interior direction scaled by each cell's own magnitudes: rho and
    e relative, momentum as rho times the rms velocity, B by its rms"""
    d = np.zeros_like(U)
    rho = U[0][inner]
    v_rms = np.sqrt(np.mean((U[1][inner] / rho) ** 2 + (U[2][inner] / rho) ** 2 + (U[3][inner] / rho) ** 2) / 3)
    b_rms = np.sqrt(np.mean(U[5][inner] ** 2 + U[6][inner] ** 2 + U[7][inner] ** 2) / 3)
    shape = rho.shape
    d[0][inner] = rho * rng.standard_normal(shape)
    for q in (1, 2, 3):
        d[q][inner] = rho * v_rms * rng.standard_normal(shape)
    d[4][inner] = U[4][inner] * rng.standard_normal(shape)
    for q in (5, 6, 7):
        d[q][inner] = b_rms * rng.standard_normal(shape)
    return d


def device_free_bytes():
    """free device memory of the current GPU from the HIP runtime the library
    runs on (hipMemGetInfo)"""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    if rc != 0:
        raise RuntimeError(f"hipMemGetInfo failed ({rc})")
    return free.value


def adjoint_study(table, tab, steps=50, every=10, ndir=2, hs=(1e-9,), precision=None, ref_grad=None):
    """Adjoint of J = sum v_vertical^2 over the interior layer 64 of the 128^3
    box after `steps` steps, checkpointed every `every` steps: forward,
    recorded-forward and backward wall times, HBM in use, and central
    differences along random interior directions (dt held fixed), one list of
    relative differences per finite-difference step h in hs.  With several
    steps, each direction's estimate is also taken at the step where it
    agrees best with the next smaller one (the plateau between the limiter's
    kinks at larger steps and rounding at smaller ones).  precision: the
    context's (FP64 by default; in an FP32-mixed context the forward is the
    mixed step and the gradient is compared with ref_grad, the FP64
    context's).  Returns (summary, gradient)."""
    from muram_main_amd import abi, config, initial
    prec = abi.PREC_FP64 if precision is None else precision
    wl = config.workloads()["adjoint128"]
    U = initial.make_state(tab, wl.box)
    free0 = device_free_bytes()
    m, r = make_grid(wl, table, U, prec)
    p_bc, s_bc = m.bc()
    m.step(r, config.DT_RAD_OFF, 1)
    dt = 0.5 * r.dt
    L = 2 + 64
    inner = (slice(2, -2),) * 3

    def run(U0, record=False):
        m.set_state(U0)
        m.set_bc(p_bc, s_bc)
        rr = config.run_desc()
        rr.autostep = 0
        rr.dt = dt
        m.exchange_grid()
        m.SetBoundaryConditions(rr, 0)
        m.synchronize()
        t0 = time.perf_counter()
        if record:
            m.adjoint_forward(rr, config.DT_RAD_OFF, steps, every)
        else:
            m.step(rr, config.DT_RAD_OFF, steps)
        m.synchronize()
        return m.get_state(), time.perf_counter() - t0

    def J(UN):
        v = UN[1, 2:-2, 2:-2, L] / UN[0, 2:-2, 2:-2, L]
        return float(np.sum(v * v))

    _, t_plain = run(U)
    UN, t_rec = run(U, record=True)
    seed = np.zeros_like(UN)
    v0, rho = UN[1, 2:-2, 2:-2, L], UN[0, 2:-2, 2:-2, L]
    seed[1, 2:-2, 2:-2, L] = 2.0 * v0 / rho ** 2
    seed[0, 2:-2, 2:-2, L] = -2.0 * v0 ** 2 / rho ** 3
    t0 = time.perf_counter()
    grad = m.adjoint_backward(seed, fold_initial_ghosts=True)
    t_back = time.perf_counter() - t0
    free1 = device_free_bytes()
    rng = np.random.default_rng(11)
    # the forward is smooth only between limiter switches, which lie within
    # 1e-10 of many points (relative near ties): differences at 1e-6..1e-8
    # straddle switches and disagree among themselves, so the steps are
    # 1e-8 and below (relative to each cell's magnitudes)
    errs = {h: [] for h in hs}
    plateau = []
    for _ in range(ndir):
        d = local_direction(U, inner, rng)
        ad = float(np.sum(grad[(slice(None),) + inner] * d[(slice(None),) + inner]))
        fds = []
        for h in hs:
            fd = (J(run(U + h * d)[0]) - J(run(U - h * d)[0])) / (2 * h)
            fds.append(fd)
            errs[h].append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
        if len(hs) > 1:
            k = min(range(len(hs) - 1), key=lambda i: abs(fds[i] - fds[i + 1]))
            plateau.append(abs(fds[k] - ad) / max(abs(fds[k]), abs(ad), 1e-300))
    m.close()
    out = {"workload": wl.name, "steps": steps, "checkpoint_every": every, "dt_s": dt,
           "forward_s": round(t_plain, 4), "forward_recorded_s": round(t_rec, 4), "backward_s": round(t_back, 4),
           "backward_over_forward": round(t_back / t_plain, 2),
           "hbm_in_use_bytes": int(free0 - free1)}
    if ndir > 0:
        out["fd_step"] = list(hs) if len(hs) > 1 else hs[0]
        out["fd_rel_diff"] = ({f"{h:g}": [float(f"{e:.3e}") for e in errs[h]] for h in hs} if len(hs) > 1
                              else [float(f"{e:.3e}") for e in errs[hs[0]]])
        if plateau:
            out["fd_rel_diff_plateau"] = [float(f"{e:.3e}") for e in plateau]
            out["fd_plateau_max"] = float(f"{max(plateau):.3e}")
    if ref_grad is not None:
        # per field: max|a - b| / max|b| and ||a - b|| / ||b|| over the
        # interior; along random directions: |a.d - b.d| / sum|b * d|
        fe, f2 = {}, {}
        for q, name in enumerate(FIELDS):
            a, b = grad[q][inner], ref_grad[q][inner]
            den, n2 = float(np.abs(b).max()), float(np.linalg.norm(b))
            fe[name] = float(f"{(np.abs(a - b).max() / den if den > 0 else 0.0):.3e}")
            f2[name] = float(f"{(np.linalg.norm(a - b) / n2 if n2 > 0 else 0.0):.3e}")
        out["grad_rel_diff_vs_fp64_adjoint"] = {"linf": fe, "l2": f2}
        rng2 = np.random.default_rng(12)
        dd = []
        for _ in range(8):
            d = local_direction(U, inner, rng2)
            ta = grad[(slice(None),) + inner] * d[(slice(None),) + inner]
            tb = ref_grad[(slice(None),) + inner] * d[(slice(None),) + inner]
            dd.append(float(f"{abs(float(ta.sum()) - float(tb.sum())) / float(np.abs(tb).sum()):.3e}"))
        out["directional_rel_diff_vs_fp64_adjoint"] = dd
    out["note"] = ("FP64 adjoint; dt, TVD branch, p_bc/s_bc held at the forward's values" if prec == abi.PREC_FP64 else
                   "FP32-mixed context: mixed forward and mixed checkpoints, segments replayed with the mixed step, "
                   "FP64 transposes at the widened mixed states; dt, TVD branch, p_bc/s_bc held")
    return out, grad


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box512")
    ap.add_argument("--precision", default="both", choices=["fp64", "fp32", "both"])
    ap.add_argument("--error-steps", type=int, default=100,
                    help="steps of the FP64 vs FP32-mixed precision study (0: skip)")
    ap.add_argument("--error-workload", default="box512_v3")
    ap.add_argument("--adjoint-steps", type=int, default=50,
                    help="steps of the 128^3 adjoint study (0: skip)")
    ap.add_argument("--corona-steps", type=int, default=10,
                    help="steps of the coronal low-beta FP64 vs FP32-mixed run (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timed oracle steps of the CPU baseline (after one warm-up)")
    ap.add_argument("--adjoint-dirs", type=int, default=16,
                    help="random directions of the adjoint's finite-difference check")
    ap.add_argument("--no-drop-in", dest="drop_in", action="store_false",
                    help="skip the timed drop-in sequence sub-line")
    ap.add_argument("--table-n", type=int, default=512)
    ap.add_argument("--path", default="fused", choices=["fused", "sequence"])
    ap.add_argument("--va-adjust", type=float, default=None,
                    help="override param_va_adjust of the workload (1: Adjust_Valf_Max's adaptive limiter)")
    ap.add_argument("--eta", type=float, default=None,
                    help="param_eta (resistivity, cm^2/s); default: the workload's (0)")
    ap.add_argument("--spitzer", type=float, default=None,
                    help="param_spitzer (Spitzer heat conduction); default: the workload's (0)")
    ap.add_argument("--divb", action="store_true",
                    help="Clean_divB on (divB_switch 1, at most 5 passes, error target 0.2)")
    ap.add_argument("--stage-zwin", default=None, choices=["0", "1", "2", "3", "4"],
                    help="stage kernel: 4 three wave groups, the EOS in its own (default), 3 two wave groups, "
                         "2 fluxes shared through LDS in one group, 1 register window, 0 re-read")
    ap.add_argument("--fma", default=None, choices=["0", "1"])
    ap.add_argument("--calibrate", action="store_true",
                    help="run a 512 MiB device copy first (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    os.environ["MRM_STEP_PATH"] = args.path
    if args.stage_zwin is not None:
        os.environ["MRM_STAGE_ZWIN"] = args.stage_zwin
    if args.fma is not None:
        os.environ["MRM_FMA"] = args.fma
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from muram_main_amd import abi, config, eos_table, initial

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.calibrate:
        x = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
    wl = config.workloads()[args.workload]
    if args.va_adjust is not None:
        wl.phys.params[config.I_PARAM_VA_ADJUST] = args.va_adjust
    if args.divb:
        wl.phys.divB[:] = [1, 5, 0.2]
    if args.eta is not None:
        wl.phys.params[config.I_PARAM_ETA] = args.eta
    if args.spitzer is not None:
        wl.phys.params[config.I_PARAM_SPITZER] = args.spitzer
    box = wl.box
    t = time.perf_counter()
    spec = eos_table.TableSpec(n_eps=args.table_n, n_lr=args.table_n, n_lp3=200, n_s3=150)
    table = eos_table.generate(spec)
    tab = eos_table.load_tables(table)
    log(f"EOS table {args.table_n}x{args.table_n} in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    U = initial.make_state(tab, box)
    log(f"state {box.gsize} in {time.perf_counter() - t:.1f}s")
    ncell = box.gsize[0] * box.gsize[1] * box.gsize[2]
    forms = {"fp64": ["fp64"], "fp32": ["fp32"], "both": ["fp64", "fp32"]}[args.precision]
    res_forms = {}
    dt_s = None
    for form in forms:
        prec = abi.PREC_FP64 if form == "fp64" else abi.PREC_FP32_MIXED
        m, r = make_grid(wl, table, U, prec)
        wall, kernels = timed_run(m, r, args.steps, args.warmup)
        nbuf = m.bufsize
        m.close()
        value = ncell * args.steps / wall / 1e6
        if form == "fp32":
            kb = KERNEL_BYTES_MX
        else:
            kb = KERNEL_BYTES if args.path == "sequence" else KERNEL_BYTES_FUSED
        res_forms[form] = dict(
            value=round(value, 2), ms_per_step=round(wall / args.steps * 1e3, 3),
            hbm_frac_model=round(value * 1e6 * MODEL_BYTES[form] / (HBM_PEAK_GBS * 1e9), 4),
            roofline=roofline(kernels, kb, ncell, nbuf, form, STAGE_BYTES[form]),
            kernels=kernel_summary(kernels, wall), dt_s=r.dt)
        log(f"{form}: {res_forms[form]['ms_per_step']} ms/step, {value:.1f} Mcell/s")
    fp64_fma = None
    if args.path == "fused" and "fp64" in forms and args.fma is None:
        # the same step from the contracted build of the fused kernels
        os.environ["MRM_FMA"] = "1"
        try:
            m, r = make_grid(wl, table, U, abi.PREC_FP64)
        finally:
            os.environ.pop("MRM_FMA", None)
        wall, kernels = timed_run(m, r, args.steps, args.warmup)
        nbuf = m.bufsize
        m.close()
        value = ncell * args.steps / wall / 1e6
        fp64_fma = dict(value=round(value, 2), ms_per_step=round(wall / args.steps * 1e3, 3),
                        roofline=roofline(kernels, KERNEL_BYTES_FUSED, ncell, nbuf, "fp64", STAGE_BYTES["fp64"]),
                        kernels=kernel_summary(kernels, wall),
                        note="fused kernels built with FMA contraction (MRM_FMA=1); within 5e-12 relative L2 "
                             "of the oracle after 10 box512 steps instead of 1e-12")
        log(f"fp64 contracted: {fp64_fma['ms_per_step']} ms/step, {value:.1f} Mcell/s")
    drop_in = None
    if args.path == "fused" and "fp64" in forms and args.drop_in:
        # the per-operation sequence the unchanged solver.C drives through the shim
        os.environ["MRM_STEP_PATH"] = "sequence"
        try:
            m, r = make_grid(wl, table, U, abi.PREC_FP64)
        finally:
            os.environ["MRM_STEP_PATH"] = "fused"
        nseq = max(1, args.steps // 2)
        wall, kernels = timed_run(m, r, nseq, max(1, args.warmup // 2))
        nbuf = m.bufsize
        m.close()
        value = ncell * nseq / wall / 1e6
        drop_in = dict(value=round(value, 2), ms_per_step=round(wall / nseq * 1e3, 3), steps=nseq,
                       hbm_frac_model=round(value * 1e6 * MODEL_BYTES["fp64"] / (HBM_PEAK_GBS * 1e9), 4),
                       roofline=roofline(kernels, KERNEL_BYTES_DROPIN, ncell, nbuf, "fp64_sequence",
                                         STAGE_BYTES_DROPIN),
                       kernels=kernel_summary(kernels, wall),
                       path="sequence: one C-ABI call per reference operation (solver.C:109-365)")
        log(f"drop-in sequence: {drop_in['ms_per_step']} ms/step, {value:.1f} Mcell/s")
    del U
    main_form = forms[0]
    R = res_forms[main_form]
    res = {
        "metric": METRIC, "value": R["value"], "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": R["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32+f64",
        "data": "synthetic (seeded stratified box, generated EOS table)",
        "config": {"workload": wl.name, "gsize": list(box.gsize), "dx_cm": box.dx[0],
                   "precision": "fp64 plain" if main_form == "fp64" else "fp32-mixed",
                   "steps_timed": args.steps, "dt_s": R["dt_s"],
                   "path": args.path, "va_adjust": wl.phys.params[config.I_PARAM_VA_ADJUST],
                   "divb": bool(wl.phys.divB[0] > 0), "eta": wl.phys.params[config.I_PARAM_ETA],
                   "spitzer": wl.phys.params[config.I_PARAM_SPITZER],
                   "stage_kernel": {"4": "three wave groups (k_stage_w, EOS group)", "3": "two wave groups (k_stage_w)",
                                    "2": "one group, shared fluxes (k_stage_f)", "1": "register window (k_stage)",
                                    "0": "re-read (k_stage)"}.get(os.environ.get("MRM_STAGE_ZWIN", "4"), "?"),
                   "fma": os.environ.get("MRM_FMA", "0")},
        "hbm_frac_model": R["hbm_frac_model"],
        "roofline": R["roofline"],
        "kernels": R["kernels"],
    }
    if drop_in:
        res["drop_in"] = drop_in
    if fp64_fma:
        res["fp64_contracted"] = fp64_fma
    if "fp32" in res_forms and main_form == "fp64":
        F = res_forms["fp32"]
        res["fp32_mixed"] = {k: F[k] for k in ("value", "ms_per_step", "hbm_frac_model", "roofline", "kernels")}
        res["fp32_mixed"]["speedup_vs_fp64"] = round(F["value"] / R["value"], 3)
    if args.error_steps > 0 and "fp32" in res_forms:
        res["precision_study"] = precision_study(config.workloads()[args.error_workload], table, tab,
                                                 args.error_steps)
        # the FP32-storage floor of the same distributions (a crop, where the
        # every-step storage rounding of the FP64 form is affordable)
        res["precision_study"]["storage_floor_crop"] = storage_floor_study(table, tab, args.error_steps)
    if args.corona_steps > 0 and "fp32" in res_forms:
        res["corona_study"] = precision_study(config.workloads()["corona"], table, tab, args.corona_steps,
                                              control=True, positivity=True)
    if args.adjoint_steps > 0:
        t = time.perf_counter()
        hs = (1e-8, 3e-9, 1e-9, 3e-10)
        res["adjoint_study"], g64 = adjoint_study(table, tab, steps=args.adjoint_steps, ndir=args.adjoint_dirs,
                                                  hs=hs)
        # the same horizon in an FP32-mixed context (timing, HBM, gradient
        # against the FP64 context's)
        from muram_main_amd import abi as _abi
        res["adjoint_study_fp32_mixed"], _ = adjoint_study(table, tab, steps=args.adjoint_steps, ndir=0,
                                                           precision=_abi.PREC_FP32_MIXED, ref_grad=g64)
        del g64
        # a shorter horizon; the limiter's near ties (about 1e6 per step at
        # margins below 1e-10) make the forward piecewise smooth at any
        # horizon, so the steps stay small
        res["adjoint_study_10"], _ = adjoint_study(table, tab, steps=10, every=5, ndir=args.adjoint_dirs, hs=hs)
        log(f"adjoint study in {time.perf_counter() - t:.1f}s")
    if not args.no_cpu:
        t = time.perf_counter()
        res["cpu_baseline"] = cpu_baseline(table, wl, steps=args.cpu_steps)
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s")
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
