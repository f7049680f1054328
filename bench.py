#!/usr/bin/env python3
"""Benchmark of the FINFLO hot path on one MI355X.

A step is one solver iteration of the finest-level CYCLES sequence over the block (STATIM,
TURBDI, RESIDI, IMPSDI with LU-SGS, FIN + limiters), with every array resident in HBM.
Prints one JSON line with the throughput, the roofline of the dominant kernel (algorithmic
bytes per cell over its event-timed launch duration) and a CPU baseline timed on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from finflo_amd.cases import flat_plate_block   # noqa: E402
from finflo_amd.layout import default_params    # noqa: E402

IMAX, JMAX, KMAX = 192, 96, 96
MACH, REYNOLDS = 0.2, 5.0e6
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory, unique) FP64 bytes per cell of each kernel: every field the
# kernel must read or write, once (neighbour values come from the same arrays).
KERNEL_BYTES_PER_CELL = {
    # flux state (RO U V W P C DRDP DRDH), viscous state (VIS EPS2 VIST CP CH TEMP), VOL,
    # 3 owned faces x (A, nx, ny, nz), D out (5)
    "k_residual_tile": 8 * (8 + 6 + 1 + 12 + 5),
    "k_residual": 8 * (8 + 6 + 1 + 12 + 5),
    # D in (5), state (U V W C RO VOL VIS VIST DTL), 3 owned faces x 4, D out (5)
    "k_pred_tiles": 8 * (5 + 9 + 12 + 5),
    "k_pred_plane": 8 * (5 + 9 + 12 + 5),
    # the same plus CP (DRO/CP at the end of IMPS)
    "k_corr_tiles": 8 * (5 + 10 + 12 + 5),
    "k_corr_plane": 8 * (5 + 9 + 12 + 5),
    # 3 faces x 4, VOL, 11 state fields, D in, D and DTL out
    "k_time3_copr": 8 * (12 + 1 + 11 + 5 + 6),
    # every cell incl. ghosts: 5 solved variables in, 12 derived fields out
    "k_state": 8 * (5 + 12),
}
ALL_CELL_KERNELS = {"k_state"}          # launched over ghosts too
# whole iteration, laminar ideal gas, every field touched once per phase (residual; backward
# sweep; forward sweep + update): 88 doubles = 704 B per cell-update
ITER_BYTES_PER_CELL = 704


def workload_params():
    prm = default_params()
    c = np.sqrt(prm.gamma * prm.rgas * prm.frstem)
    prm.frsvel = MACH * c
    # Re = rho U L / mu with L = 1 m: scale Sutherland's viscosity (RMULTV) to reach it
    t = prm.frstem
    mu_suth = prm.visu0 * t * np.sqrt(t) / (t + prm.tsu0)
    prm.rmult = prm.frsden * prm.frsvel / REYNOLDS / mu_suth
    prm.frsvis = mu_suth * prm.rmult
    prm.cfl = prm.cfll = 5.0
    return prm


def make_workload(seed=0):
    prm = workload_params()
    blk = flat_plate_block(IMAX, JMAX, KMAX, prm, first=1e-6, ratio=1.15, seed=seed, amp=0.005)
    return prm, blk


def cpu_baseline(prm, blk, seconds=20.0):
    """Time the reference's own Fortran (oracle/_ref) on the host, single thread, on a
    bounded sample of the same workload: whole iterations of the full block."""
    try:
        from oracle.oracle_py import OracleBlock, RefBlock, have_ref
    except Exception as e:   # the oracle is optional for the bench
        return {"value": None, "unit": "cell-updates/s", "cores": 1, "kind": "none",
                "sample": f"oracle unavailable: {e}"}
    kind = "reference" if have_ref() else "port"
    cls = RefBlock if have_ref() else OracleBlock
    b = blk.copy()
    o = cls(prm, b)
    ncell = blk.imax * blk.jmax * blk.kmax
    n = 0
    t0 = time.perf_counter()
    while True:
        o.cycle()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or el / n * (n + 1) > 1.5 * seconds:
            break
    return {"value": ncell * n / el, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": f"{n} full iterations of the {blk.imax}x{blk.jmax}x{blk.kmax} workload "
                      f"block ({el:.1f} s, one thread)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(blocks would shard across GPUs; this build measures one GPU)", file=sys.stderr)
        sys.exit(2)

    from finflo_amd.ffhip import Context

    prm, blk = make_workload()
    ncell = IMAX * JMAX * KMAX
    print(f"[bench] workload {IMAX}x{JMAX}x{KMAX} ({ncell} cells), building context", file=sys.stderr, flush=True)
    ctx = Context(prm)
    ctx.register(blk)
    ctx.upload()
    ctx.cycle(ncycles=args.warmup)
    ctx.finish()
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    # timed region: K iterations, device-resident, bracketed by stream synchronisation
    ctx.finish()
    t0 = time.perf_counter()
    resi = ctx.cycle(ncycles=args.steps)
    ctx.finish()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    value = ncell * args.steps / el
    print(f"[bench] {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)

    # the same K iterations again with a HIP event pair around every launch on the
    # library's stream: per-kernel durations for the roofline (events add host overhead,
    # so they stay out of the timed region above)
    ctx.profile(True)
    ctx.reset_stats()
    ctx.cycle(ncycles=args.steps)
    ctx.finish()
    ctx.profile(False)

    kstats = {}
    for name in ("k_residual_tile", "k_residual", "k_pred_tiles", "k_corr_tiles",
                 "k_pred_plane", "k_corr_plane", "k_time3_copr", "k_state", "k_limits",
                 "k_fin", "k_zero", "k_save_old", "k_divcp", "k_turb_lam", "k_resl2_part",
                 "k_resl2_final"):
        n, ms = ctx.kernel_stats(name)
        if n:
            kstats[name] = {"launches": n, "ms_total": ms, "ms_per_step": ms / args.steps}
    # dominant kernel: the largest share of the step among those with a byte model
    dom = max((k for k in kstats if k in KERNEL_BYTES_PER_CELL), key=lambda k: kstats[k]["ms_total"])
    ks = kstats[dom]
    avg_ms = ks["ms_total"] / ks["launches"]
    cells_per_launch = blk.ntot if dom in ALL_CELL_KERNELS else ncell
    achieved = KERNEL_BYTES_PER_CELL[dom] * cells_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "avg_launch_ms": avg_ms,
                "bytes_per_cell": KERNEL_BYTES_PER_CELL[dom], "cells_per_launch": cells_per_launch}
    sweep_ms = sum(kstats.get(k, {}).get("ms_total", 0.0)
                   for k in ("k_pred_plane", "k_corr_plane", "k_pred_tiles", "k_corr_tiles")) / args.steps
    iter_bw = ITER_BYTES_PER_CELL * ncell / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(prm, blk, args.cpu_seconds)

    line = {
        "metric": "cell-updates/s per MI355X (residual+LU-SGS)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"flat plate {IMAX}x{JMAX}x{KMAX} single block, laminar "
                               f"ideal gas M={MACH} Re={REYNOLDS:g}, Roe+FLUXDO, LU-SGS "
                               f"CFL=5, finest level only",
                   "cells": ncell, "iteration": "STATIM+TURBDI+RESIDI+IMPSDI+FIN",
                   "lusgs_sweep_ms_per_step": sweep_ms,
                   "iteration_hbm_GBps_at_704B_per_cell": iter_bw,
                   "density_resl2": float(np.sqrt(resi[0]) / ncell)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kstats,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
