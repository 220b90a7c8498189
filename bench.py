#!/usr/bin/env python3
"""Benchmark of the MI355X MHD step: one "step" is one full RK time step
(4 stages of ConsToPrim + MHD_Residual + Source_Integrate_Tcheck + ghost fill,
then TVDlimit, TCheck, ghost fill) over the stratified 256x512x512 box
(vertical 256, dx = 16 km), synthetic seeded state, inputs resident in HBM.

Prints one JSON line.  value = interior cells x steps / wall time (Mcell-updates/s).
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
    "residual": (200.0, "interior"),      # read U 64 + p 8 + Res 64; write Res 64
    "source_integrate": (320.0, "interior"),  # read U, Res, U0 192; write U, Res 128
    "save_cons": (128.0, "interior"),
    "tcheck": (80.0, "interior"),
    "tvd_sweep_x": (136.0, "interior"),   # read U 64 + p 8; write U 64
    "tvd_sweep_y": (136.0, "interior"),
    "tvd_sweep_z": (136.0, "interior"),
    # fused path: stage kernel reads U_in 64 + p 8 + U0 64, writes U_out 64 + p 8
    "stage_fused": (208.0, "interior"),
    "eos_p": (72.0, "buffer"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(table, phys, box_gsize, dx, steps=3):
    """Oracle FP64 step on a bounded sample box with the same distributions."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from muram_main_amd import config, eos_table, initial
    from orc import Oracle
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=box_gsize, dx=dx, seed=1)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    o = Oracle(g, phys, table)
    r = config.run_desc()
    o.set_state(U)
    del U
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    o.step(r, config.DT_RAD_OFF, 1)          # warm-up
    t0 = time.perf_counter()
    o.step(r, config.DT_RAD_OFF, steps)
    dt = time.perf_counter() - t0
    ncell = box_gsize[0] * box_gsize[1] * box_gsize[2]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=ncell * steps / dt / 1e6, unit="Mcell-updates/s", cores=threads, kind="port",
                sample=f"oracle FP64 step, {box_gsize[0]}x{box_gsize[1]}x{box_gsize[2]} box of the same "
                       f"distributions, {steps} steps after 1 warm-up, {threads} OpenMP threads",
                s_per_step=dt / steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box512")
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--table-n", type=int, default=512)
    ap.add_argument("--path", default="fused", choices=["fused", "sequence"])
    ap.add_argument("--stage-zwin", default=None, choices=["0", "1"])
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
    from muram_main_amd.mhd import MhdGrid

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.calibrate:
        x = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
    wl = config.workloads()[args.workload]
    box = wl.box
    t = time.perf_counter()
    spec = eos_table.TableSpec(n_eps=args.table_n, n_lr=args.table_n, n_lp3=200, n_s3=150)
    table = eos_table.generate(spec)
    tab = eos_table.load_tables(table)
    log(f"EOS table {args.table_n}x{args.table_n} in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    U = initial.make_state(tab, box)
    log(f"state {box.gsize} in {time.perf_counter() - t:.1f}s")
    g = config.grid_desc(box.gsize, box.dx)
    m = MhdGrid(g, wl.phys)
    m.eos_init(table)
    m.set_state(U)
    del U
    r = config.run_desc()
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    for _ in range(args.warmup):
        m.step(r, config.DT_RAD_OFF, 1)
    m.synchronize()
    log(f"warm-up done, dt = {r.dt:.4e} s")

    lib = m.lib
    lib.mrm_timers(m.h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        m.step(r, config.DT_RAD_OFF, 1)
    m.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes as C
    ms = (C.c_double * abi.NTIMER)()
    cnt = (C.c_long * abi.NTIMER)()
    lib.mrm_get_timers(m.h, ms, cnt)
    lib.mrm_timers(m.h, 0)
    kernels = {}
    for i in range(abi.NTIMER):
        if cnt[i]:
            kernels[lib.mrm_timer_name(i).decode()] = dict(total_ms=ms[i], launches=cnt[i],
                                                           mean_ms=ms[i] / cnt[i])
    ncell = box.gsize[0] * box.gsize[1] * box.gsize[2]
    nbuf = m.bufsize
    value = ncell * args.steps / wall / 1e6
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    per, unit = KERNEL_BYTES.get(dom, (0.0, "interior"))
    units = nbuf if unit == "buffer" else ncell
    bytes_launch = per * units
    achieved = bytes_launch / (kernels[dom]["mean_ms"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath)).get(dom)
        if tr:
            traffic = tr.get("bytes_per_launch")
    roof = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                bytes_per_launch=bytes_launch, bytes_per_unit=per, units=units)
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded stratified box, generated EOS table)",
        "config": {"workload": wl.name, "gsize": list(box.gsize), "dx_cm": box.dx[0],
                   "precision": args.precision, "steps_timed": args.steps, "dt_s": r.dt,
                   "path": args.path, "stage_zwin": os.environ.get("MRM_STAGE_ZWIN", "1"),
                   "fma": os.environ.get("MRM_FMA", "1")},
        "hbm_frac_model": round(value * 1e6 * MODEL_BYTES[args.precision] / (HBM_PEAK_GBS * 1e9), 4),
        "roofline": roof,
        "kernels": {k: {"mean_ms": round(v["mean_ms"], 4), "launches": v["launches"],
                        "share": round(v["total_ms"] / (wall * 1e3), 4)} for k, v in kernels.items()},
    }
    m.close()
    if not args.no_cpu:
        t = time.perf_counter()
        res["cpu_baseline"] = cpu_baseline(table, wl.phys, (256, 128, 128), box.dx)
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s")
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
