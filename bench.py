#!/usr/bin/env python3
"""Benchmark of the FINFLO hot path on one MI355X.

Default workload: the 192x96x96 flat-plate block with k-omega SST and a 3-level multigrid
V-cycle (CYCLES with MGRID=3: the finest-level iteration STATIM, TURBDI, RESIDI, IMPSDI with
LU-SGS, FIN + limiters, then the FAS down-loop over the coarse levels and the
prolongation).  A step is one V-cycle; cell-updates count the finest level's cells once per
V-cycle.  Every array is resident in HBM before the timed region starts.

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

from finflo_amd.cases import (config2, config4, flat_plate_hierarchy, flow_params,  # noqa: E402
                               plate_bcs, plate_wall)
from finflo_amd.layout import default_params        # noqa: E402

IMAX, JMAX, KMAX = 192, 96, 96
MACH = 0.2
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # k-omega SST (ITURB=6), Re = 5e6 per plate length, MGRID = 3
    "sst-mg": dict(iturb=6, nlev=3, reynolds=5.0e6,
                   label="k-omega SST, 3-level multigrid V-cycle"),
    # laminar ideal gas on the finest level only (MGRID = 1)
    "laminar": dict(iturb=0, nlev=1, reynolds=5.0e6,
                    label="laminar ideal gas, finest level only"),
    # k-omega SST on the finest level only (MGRID = 1)
    "sst": dict(iturb=6, nlev=1, reynolds=5.0e6,
                label="k-omega SST, finest level only"),
}

# Algorithmic (compulsory, unique) FP64 bytes per cell of each kernel: every field the
# kernel must read or write, once (neighbour values come from the same arrays).
RECORD_FIELDS = 48      # values per cell in each staged LU-SGS record (ffhip.hip NRF)
KERNEL_BYTES_PER_CELL = {
    0: {
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
        "k_pred_km": 8 * (5 + 9 + 12 + 5),
        "k_corr_km": 8 * (5 + 10 + 12 + 5),
        # 3 faces x 4, VOL, 11 state fields, D in, D and DTL out
        "k_time3_copr": 8 * (12 + 1 + 11 + 5 + 6),
        # staged LU-SGS: the pass reads D (5), U V W C RO VOL VIS VIST DTL CP, A1..A3 and
        # their normals (12) and writes the two 48-double records; the sweeps read a record
        # (48) and write (predictor) or read and write (corrector) the 5 increments
        "k_sweep_prep": 8 * (5 + 10 + 12 + 2 * RECORD_FIELDS),
        "k_pred_st": 8 * (RECORD_FIELDS + 5),
        "k_corr_st": 8 * (RECORD_FIELDS + 5 + 5),
        # every cell incl. ghosts: 5 solved variables in, 12 derived fields out
        "k_state": 8 * (5 + 12),
    },
    6: {
        # flux state (RO U V W P C DRDP DRDH RK REPS), viscous (VIS EPS2 VIST CP CH TEMP),
        # FUN1, VOL, 3 owned faces x 4, sources (PTUR, SEPS in); SRK, SEPS, D (7) out
        "k_residual_kw": 8 * (10 + 6 + 1 + 1 + 12 + 2 + 9),
        # D in (7), state (U V W C RO VOL VIS VIST DTL RK DRDP DRDH), 3 faces x 4, D out (7)
        "k_pred_tiles": 8 * (7 + 12 + 12 + 7),
        # the same plus CP
        "k_corr_tiles": 8 * (7 + 13 + 12 + 7),
        "k_pred_km": 8 * (7 + 12 + 12 + 7),
        "k_corr_km": 8 * (7 + 13 + 12 + 7),
        # staged LU-SGS (7 increments; RK DRDP DRDH read by the pass as well)
        "k_sweep_prep": 8 * (7 + 13 + 12 + 2 * RECORD_FIELDS),
        "k_pred_st": 8 * (RECORD_FIELDS + 7),
        "k_corr_st": 8 * (RECORD_FIELDS + 7 + 7),
    },
}
# Compulsory bytes per cell of the three phases of one finest-level iteration, with S state
# scalars, G geometry scalars, R unknowns and X stored eddy viscosity per cell (FP64):
#   P1 residual R(W)                     8 [(S + G) + R + X]
#   P2 backward sweep (PREDGS + DIAGGS)  8 [(R + S + G + X) + R]
#   P3 forward sweep (CORRGS) + FIN      8 [(R + S + G + X) + S]
# laminar S=5 G=16 R=5 X=0 -> 208 + 248 + 248 = 704 B; k-omega S=7 G=17 R=7 X=1 ->
# 256 + 312 + 312 = 880 B.  A kernel is credited with the bytes of the phase it carries;
# passes that a fused design would not need (state, record staging, ...) are credited none.
PHASE_BYTES = {0: (208, 248, 248), 6: (256, 312, 312)}
PHASE_OF = {"k_residual_kw": 0, "k_residual_kw_ns": 0, "k_residual_tile": 0,
            "k_residual_tile_ns": 0, "k_residual": 0,
            "k_pred_st": 1, "k_pred_tiles": 1, "k_pred_km": 1, "k_pred_plane": 1,
            "k_corr_st": 2, "k_corr_tiles": 2, "k_corr_km": 2, "k_corr_plane": 2}
ALL_CELL_KERNELS = {"k_state", "k_save_old", "k_limits", "k_limits_mg", "k_forcing"}
# whole finest-level iteration, every field touched once per phase (residual; backward
# sweep; forward sweep + update): laminar 88 doubles = 704 B, k-omega 110 doubles = 880 B
ITER_BYTES_PER_CELL = {0: 704, 6: 880}

KERNELS = ("k_residual_kw", "k_residual_tile", "k_residual", "k_pred_tiles", "k_corr_tiles",
           "k_sweep_prep", "k_pred_st", "k_corr_st", "k_wall", "k_wall_forces", "k_temder",
           "k_residual_kw_ns", "k_residual_tile_ns",
           "k_pred_km", "k_corr_km", "k_pred_plane", "k_corr_plane", "k_time3_copr", "k_state", "k_turb_lam",
           "k_turb_sst", "k_grad_zero", "k_derfun", "k_extend", "k_crosko", "k_kw_round",
           "k_max_final", "k_fun1_prod", "k_limits", "k_limits_mg", "k_fin", "k_zero", "k_fill_slots",
           "k_save_old", "k_divcp", "k_resl2_part", "k_resl2_final", "k_restrict",
           "k_restrict_ghosts", "k_forcing", "k_prolong", "k_ddadi", "k_ddadi_end", "k_reflec")


def make_workload(name="sst-mg", seed=0):
    """The flat plate of the workload and its no-slip wall (the J-low face of every level):
    sst-mg is the 192x96x96 k-omega SST plate with 3 multigrid levels, sst the same on its
    finest level only; laminar is the same grid, laminar, finest level only."""
    if name in ("sst-mg", "sst"):
        return config2(seed=seed, nlev=WORKLOADS[name]["nlev"])
    prm = flow_params(iturb=0, mach=MACH, reynolds=WORKLOADS[name]["reynolds"], cfl=5.0)
    blocks = flat_plate_hierarchy(IMAX, JMAX, KMAX, 1, prm, first=1e-6, ratio=1.15,
                                  seed=seed, amp=0.005)
    return prm, blocks, [plate_wall(b, prm) for b in blocks]


def run_steps(ctx, nlev, n, graph=False):
    """n V-cycles (iterations with one level); with graph, the first n - 1 replay one captured
    HIP graph and the last one runs eagerly to return the level-1 residual sums."""
    eager = (lambda k: ctx.vcycle(nlev, ncycles=k)) if nlev > 1 else (lambda k: ctx.cycle(ncycles=k))
    if graph and n > 1:
        ctx.vcycle_graph(nlev, ncycles=n - 1)
        return eager(1)
    return eager(n)


def _cpu_worker(prm, blocks, walls, seconds, use_ref, q):
    """One process of the CPU baseline: whole V-cycles (or iterations) of the workload on
    this process's copy-on-write view of the blocks until `seconds` have passed."""
    from oracle.oracle_py import OracleBlock, OracleHierarchy, RefBlock
    if len(blocks) > 1:
        o = OracleHierarchy(prm, blocks)
        for lv, (wp, patches) in zip(o.levels, walls):
            lv.set_walls(wp, patches)
        step = o.vcycle
    else:
        o = (RefBlock if use_ref else OracleBlock)(prm, blocks[0])
        o.set_walls(*walls[0])
        step = o.cycle
    step()        # untimed: first touch of this process's copy of the blocks
    n = 0
    t0 = time.perf_counter()
    while True:
        step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or el / n * (n + 1) > 1.5 * seconds:
            break
    q.put((n, el))


def converge(ctx, nlev, ncell, max_cycles, chunk=10, periodic=False, dims=None):
    """V-cycles (iterations with one level) from the current device state until the density
    residual RESI(31) = sqrt(sum DRO^2)/N (SUBIN ns3c.f:10248) has fallen 8 decades below
    that of the first cycle, or max_cycles.  The residual is read every `chunk` cycles (one
    stream synchronisation each); the wall-clock time includes them.  periodic: the block's
    K faces are connected to each other and exchanged after every cycle."""
    if periodic:
        step = lambda k: ctx.vcycle_blocks([1], nlev, ncycles=k)
    elif nlev > 1:
        step = lambda k: ctx.vcycle(nlev, ncycles=k)
    else:
        step = lambda k: ctx.cycle(ncycles=k)
    ctx.finish()
    t0 = time.perf_counter()
    r0 = float(np.sqrt(step(1)[0]) / ncell)
    n, hist, t8 = 1, [(1, r0)], None
    while n < max_cycles:
        r = float(np.sqrt(step(chunk)[0]) / ncell)
        n += chunk
        hist.append((n, r))
        if not np.isfinite(r):
            break
        if r <= 1e-8 * r0:
            t8 = time.perf_counter() - t0
            break
    el = time.perf_counter() - t0
    last = hist[-1][1]
    dm = ctx.dromax()
    cell = None
    if dims is not None:
        # IXERR (1-based, two ghost layers) as the interior (i, j, k) of the largest change
        isz, il = dims[0] + 4, (dims[0] + 4) * (dims[1] + 4)
        q = dm[1] - 1
        cell = [q % isz - 1, (q % il) // isz - 1, q // il - 1]
    return {"cycles": n, "seconds": el, "first_resi": r0, "last_resi": last,
            "decades": float(np.log10(r0 / last)) if last > 0 and np.isfinite(last) else None,
            "time_to_8_decades": t8, "dromax_last_cycle": dm[0], "dromax_cell": cell,
            "history": hist[:: max(1, len(hist) // 40)] + [hist[-1]]}


def cpu_baseline(prm, blocks, walls, seconds=20.0):
    """Time the CPU path on the host cores on a bounded sample of the same workload: whole
    V-cycles (or iterations) of the full block with its wall.  First one process alone (the
    per-core rate), then one independent process per available core (at most 16) at once
    (the whole-host rate, which is `value`).  One level: the reference's own Fortran
    (oracle/_ref) when built, else the C restatement; multigrid: the C restatement (the
    reference's multigrid transfers are not compiled here).  Runs before the GPU is touched,
    so the forked workers inherit no device state."""
    import multiprocessing as mp
    try:
        from oracle.oracle_py import have_ref, oracle
        oracle()
    except Exception as e:   # the oracle is optional for the bench
        return {"value": None, "unit": "cell-updates/s", "cores": 0, "kind": "none",
                "sample": f"oracle unavailable: {e}"}
    use_ref = len(blocks) == 1 and have_ref()
    kind = "reference" if use_ref else "port"
    ncell = blocks[0].imax * blocks[0].jmax * blocks[0].kmax
    ncores = max(1, min(len(os.sched_getaffinity(0)), 16))
    fork = mp.get_context("fork")

    def run(nproc, secs):
        q = fork.Queue()
        ps = [fork.Process(target=_cpu_worker, args=(prm, blocks, walls, secs, use_ref, q))
              for _ in range(nproc)]
        for p in ps:
            p.start()
        res = [q.get() for _ in ps]
        for p in ps:
            p.join()
        return res

    one = run(1, seconds / 2)
    allc = run(ncores, seconds)
    per_core = sum(ncell * n / el for n, el in one)
    host = sum(ncell * n / el for n, el in allc)
    what = "V-cycles" if len(blocks) > 1 else "iterations"
    b0 = blocks[0]
    return {"value": host, "unit": "cell-updates/s", "cores": ncores, "kind": kind,
            "per_core": per_core, "nproc_host": os.cpu_count(),
            "sample": f"{ncores} independent processes at once, each running whole {what} of "
                      f"the {b0.imax}x{b0.jmax}x{b0.kmax} workload with its wall for about "
                      f"{seconds:.0f} s after one untimed step ({sum(n for n, _ in allc)} timed in "
                      f"total); per_core: one "
                      f"process alone, {one[0][0]} {what} in {one[0][1]:.1f} s; "
                      f"{'reference Fortran' if use_ref else 'C restatement'} built "
                      f"-O3 -march=x86-64-v3 -ffp-contract=off"}


def committed_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed counter summary
    (profiles/hbm_traffic_*.md: FETCH_SIZE and WRITE_SIZE passes of this bench, each scaled
    by what the counter reads for a copy kernel of known size), or None."""
    tag = {"sst-mg": "sst_mg3", "laminar": "laminar"}.get(workload)
    if tag is None:
        return None, None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"hbm_traffic_{tag}_{IMAX}x{JMAX}x{KMAX}.md")
    if not os.path.exists(path):
        return None, None
    src = os.path.relpath(path, os.path.dirname(path) + "/..")
    lines = [ln for ln in open(path) if ln.startswith("| `")]
    # the launch-label rows (label, launches, traffic MB/launch, ...): every launch of the
    # label on every level, as the roofline averages them
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells[0].strip("`") == kernel:
            return float(cells[2]) * 1e6, src
    # per-kernel rows (kernel, launches, FETCH_SIZE, WRITE_SIZE, traffic MB/launch)
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells[0].strip("`").startswith(kernel + "<"):
            return float(cells[4]) * 1e6, src
    return None, None


def copy_probe(mib):
    """A known-size HBM copy for counter calibration: an elementwise kernel that reads
    MIB MiB and writes MIB MiB (y = x * 1), launched three times."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "launches": 3}))


def multiblock_bench(args):
    """The multi-block cases over the wall bump, with one-to-one connections exchanged on the
    device after every multi-block cycle:
    * config3: 24 blocks (about 7.6M cells), k-omega SST, 3 multigrid levels;
    * config5: 64 blocks (about 40M cells) around a swept bump, k-omega SST with the EARSM,
      4 multigrid levels; the chosen form adds the learned closure, trained first on the
      exact EARSM samples of the first blocks.
    Times K multi-block cycles: the plain FP64 path with eager launches, one block after the
    other; the same replayed from the captured multi-block graph (parallel per-block
    branches, where the library captures one); the same with the blocks side by side on the
    library's stream pool (FF_OPT_BLOCK_STREAMS); and the chosen form (FP32 LU-SGS records,
    autotuned tile shapes, the blocks on the stream pool; with the closure on config5).
    Prints one JSON line; `value` is the chosen form's."""
    from finflo_amd.cases import config3, config5
    from finflo_amd.ffhip import Context
    t0 = time.perf_counter()
    five = args.workload == "config5"
    if five:
        # the closure's sample arrays are torch device tensors: bring up torch's device
        # context before the library's (it reported no device when started after it)
        import torch
        torch.zeros(1, device="cuda")
    prm, blocks, cons, walls, bcs = (config5 if five else config3)()
    nlev = len(blocks[0])
    nb = len(blocks)
    ncell = sum(lv[0].imax * lv[0].jmax * lv[0].kmax for lv in blocks)
    print(f"[bench] {args.workload}: {nb} blocks, {ncell} cells, {nlev} levels, built in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx = Context(prm)
    for n, lv in enumerate(blocks, start=1):
        for m, b in enumerate(lv, start=1):
            ctx.register(b, level=m, block=n)
            ctx.upload(level=m, block=n)
            if walls[n - 1][m - 1] is not None:
                ctx.wall_register(*walls[n - 1][m - 1], level=m, block=n)
            ctx.bc_register(bcs[n - 1][m - 1], level=m, block=n)
    for n in range(1, nb + 1):
        ctx.con_register(cons[n - 1], block=n)
        for m in range(1, nlev + 1):
            ctx.ghostext(level=m, block=n)
    ctx.halo_exchange(what="geometry")
    ctx.halo_exchange(what="flow")
    ids = list(range(1, nb + 1))
    closure = None
    if five:
        # the learned closure: exact-EARSM samples of the first eight blocks after one
        # exact cycle, normalisation, Adam steps on the device
        import torch
        from finflo_amd.mlp import init_params
        ctx.vcycle_blocks(ids, nlev, ncycles=1)
        xs, ys = [], []
        for n in range(1, 9):
            ctx.turbdi(block=n)
            x, y = ctx.earsm_samples(block=n)
            xs.append(x)
            ys.append(y)
        x, y = torch.cat(xs), torch.cat(ys)
        ctx.mlp_fit_norm(x, y)
        ctx.mlp_set(init_params(5))
        t1 = time.perf_counter()
        loss = ctx.mlp_train(x, y, steps=2000, batch=8192, lr=3e-3, seed=11)
        closure = {"samples": len(x), "train_steps": 2000, "train_s": time.perf_counter() - t1,
                   "loss_first": float(loss[0]), "loss_last": float(loss[-1])}
        print(f"[bench] closure trained on {len(x)} samples: loss {loss[0]:.3g} -> {loss[-1]:.3g}",
              file=sys.stderr, flush=True)
    res = {}
    # the library captures at most 3 blocks x 3 levels per graph branch (ff_vcycle_blocks_graph)
    can_graph = (nb + 7) // 8 * nlev <= 9
    forms = [("fp64 eager", "fp64", False, "eager", "exact"),
             ("fp64 graph", "fp64", False, "graph", "exact"),
             ("fp64 streams", "fp64", False, "streams", "exact"),
             ("chosen form", "mixed", True, "streams", "mlp" if five else "exact")]
    for label, prec, tune, launch, earsm in forms:
        if launch == "graph" and not can_graph:
            res[label] = {"skipped": f"{nb} blocks of {nlev} levels exceed the library's graph "
                                     "capture limit (3 blocks x 3 levels per branch)"}
            continue
        ctx.precision(prec)
        ctx.autotune(tune)
        ctx.block_streams(launch == "streams")
        if five:
            ctx.earsm_mode(earsm)
        run = (lambda k: ctx.vcycle_blocks_graph(ids, nlev, ncycles=k)) if launch == "graph" else \
              (lambda k: ctx.vcycle_blocks(ids, nlev, ncycles=k))
        ctx.vcycle_blocks(ids, nlev, ncycles=max(1, args.warmup))   # tunes, allocates
        run(1)
        ctx.finish()
        t1 = time.perf_counter()
        run(args.steps)
        ctx.finish()
        el = time.perf_counter() - t1
        res[label] = {"ms_per_step": el * 1e3 / args.steps, "value": ncell * args.steps / el}
        print(f"[bench] {args.workload} {label}: {el * 1e3 / args.steps:.2f} ms per cycle",
              file=sys.stderr, flush=True)
    r = ctx.vcycle_blocks(ids, nlev, ncycles=1)
    base = res["fp64 eager"]
    tuned = {}
    for n in ids[:4]:
        t = ctx.tuned_shape(level=1, block=n)
        if t:
            tuned[f"block{n}"] = {"shape": t[0], "ms": t[1]}
    line = {
        "metric": "cell-updates/s per MI355X (residual+LU-SGS), multi-block chosen form",
        "value": res["chosen form"]["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["chosen form"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 with FP32 LU-SGS records" + (", bf16 closure" if five else ""),
        "chosen_form_launch": "eager launches, blocks side by side on the stream pool",
        "data": "synthetic",
        "config": {"workload": f"{nb}-block {'swept ' if five else ''}bump case, {ncell} cells, "
                               f"k-omega SST{' + EARSM' if five else ''}, M={0.6 if five else 0.5} "
                               f"Re={'1e7' if five else '1e6'}, {nlev}-level multigrid, one-to-one "
                               "connections exchanged on the device; a step is one multi-block cycle",
                   "blocks": nb, "cells": ncell, "forms": res,
                   "speedup_chosen_over_fp64_eager": base["ms_per_step"] / res["chosen form"]["ms_per_step"],
                   "tuned_shapes": tuned, "closure": closure,
                   "density_resl2": float(np.sqrt(r[0]) / ncell)},
    }
    print(json.dumps(line))
    ctx.close()


def earsm_mlp_bench(args):
    """The learned EARSM closure on a 256x128x128 block (4.19M cells, k-omega SST with
    EARSM): exact-EARSM samples of every cell, training (Adam minibatches), inference of
    every cell, and EARSM1 through the closure against the exact EARSM1, timed with HIP
    events and compared field by field.  Prints one JSON line."""
    import torch
    from finflo_amd.ffhip import Context
    ni, nj, nk = 256, 128, 128
    prm, blk, _ = config4(seed=4)
    ncell = ni * nj * nk
    print(f"[bench] earsm-mlp {ni}x{nj}x{nk} ({ncell} cells), building context", file=sys.stderr, flush=True)
    ctx = Context(prm)
    ctx.register(blk)
    ctx.upload()
    ctx.save_old()
    ctx.statim()
    ctx.turbdi()                                   # exact EARSM1
    ctx.finish()
    ctx.download(fields=["B11", "B12", "B13", "B22", "B23", "B33"])
    exact = np.stack([blk.interior(f).ravel().copy() for f in ("B11", "B12", "B13", "B22", "B23", "B33")], 1)
    x, y = ctx.earsm_samples()
    ctx.mlp_fit_norm(x, y)
    from finflo_amd.mlp import init_params
    ctx.mlp_set(init_params(5))
    steps, batch = args.steps * 100, 8192
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = ctx.mlp_train(x, y, steps=steps, batch=batch, lr=3e-3, seed=11)
    train_s = time.perf_counter() - t0
    print(f"[bench] {steps} training steps in {train_s:.2f} s, loss {loss[0]:.3g} -> {loss[-1]:.3g}",
          file=sys.stderr, flush=True)
    # event-timed kernels: training steps, inference, EARSM1 both ways
    ctx.profile(True)
    ctx.reset_stats()
    ctx.mlp_train(x, y, steps=args.steps, batch=batch, lr=1e-5, seed=12)
    st_grad = ctx.kernel_stats("k_mlp_grad")
    for _ in range(args.warmup):
        ctx.mlp_infer(x)
    ctx.reset_stats()
    for _ in range(args.steps):
        yp = ctx.mlp_infer(x)
    ctx.finish()
    st = {"k_mlp_infer": ctx.kernel_stats("k_mlp_infer"), "k_mlp_grad": st_grad}
    ctx.earsm_mode("mlp")
    ctx.reset_stats()
    for _ in range(args.steps):
        ctx.turbdi()
    ctx.finish()
    st["k_turb_sst without EARSM1"] = ctx.kernel_stats("k_turb_sst")
    st["k_earsm_mlp"] = ctx.kernel_stats("k_earsm_mlp")
    ctx.download(fields=["B11", "B12", "B13", "B22", "B23", "B33"])
    got = np.stack([blk.interior(f).ravel() for f in ("B11", "B12", "B13", "B22", "B23", "B33")], 1)
    ctx.earsm_mode("exact")
    ctx.reset_stats()
    for _ in range(args.steps):
        ctx.turbdi()
    ctx.finish()
    st["k_turb_sst with EARSM1"] = ctx.kernel_stats("k_turb_sst")
    ctx.profile(False)
    nin, msin = st["k_mlp_infer"]
    infer_ms = msin / nin
    flops_cell = 2 * (5 * 64 + 64 * 64 + 64 * 3)
    tflops = flops_cell * ncell / (infer_ms * 1e-3) / 1e12
    ng, msg = st["k_mlp_grad"]
    # the training kernels' flops per sample: forward, the two input-gradient products and
    # the three weight-gradient products
    grad_tflops = 3 * flops_cell * batch / (msg / ng * 1e-3) / 1e12
    rel = [float(np.linalg.norm(got[:, q] - exact[:, q]) / np.linalg.norm(exact[:, q])) for q in range(6)]
    yt = y.cpu().numpy()
    logerr = np.abs(np.log(yp.cpu().numpy() / yt)).mean(0).tolist()
    avg = {k: v[1] / max(v[0], 1) for k, v in st.items()}
    line = {
        "metric": "learned EARSM closure: inference TFLOP/s (BF16 MFMA) and b_ij error against the exact EARSM",
        "value": tflops, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "bf16 (FP32 accumulate)", "data": "synthetic",
        "config": {"workload": f"3-D Gaussian bump {ni}x{nj}x{nk} single block, k-omega SST + "
                               "EARSM, M=0.3 Re=2e6, velocity noise 5 %, training samples = "
                               "every cell's exact EARSM",
                   "cells": ncell, "train_steps": steps, "train_batch": batch,
                   "train_seconds": train_s, "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "kernel": "k_mlp_infer", "avg_launch_ms": infer_ms,
                     "flops_per_cell": flops_cell},
        "training_tflops": grad_tflops,
        "mean_abs_log_error_beta4_beta6_beta9": logerr,
        "bij_relative_l2_error": dict(zip(("B11", "B12", "B13", "B22", "B23", "B33"), rel)),
        "earsm_pass_ms": {
            "exact EARSM1 (k_turb_sst with it minus without it)":
                avg["k_turb_sst with EARSM1"] - avg["k_turb_sst without EARSM1"],
            "learned closure (k_earsm_mlp)": avg["k_earsm_mlp"],
            "k_turb_sst with EARSM1": avg["k_turb_sst with EARSM1"],
            "k_turb_sst without EARSM1": avg["k_turb_sst without EARSM1"]},
    }
    print(json.dumps(line))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["earsm-mlp", "config3", "config5"],
                    default="sst-mg")
    ap.add_argument("--implicit", choices=["lusgs", "ddadi"], default="lusgs",
                    help="the implicit stage: LU-SGS (BLKS%%LUSGS = 1, default) or DDADI (0)")
    ap.add_argument("--lusgs", choices=["tiles", "kmarch", "planes", "staged", "staged1"], default=None,
                    help="LU-SGS sweep form (default: the library's)")
    ap.add_argument("--graph", type=int, choices=[0, 1], default=0,
                    help="replay the steps from a captured HIP graph")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cfl", type=float, default=None,
                    help="CFL number of every level (default: the workload's, 5)")
    ap.add_argument("--kbc", choices=["ext", "periodic"], default="ext",
                    help="the K faces in the convergence run: free stream (EXT) or connected "
                         "to each other (periodic span)")
    ap.add_argument("--outer", choices=["ext", "inout"], default="ext",
                    help="the I faces of the plate: free stream (EXT, GHOSTEXT at start-up) or "
                         "a subsonic inlet (INLOUT 4) and a pressure outlet (INLOUT 3)")
    ap.add_argument("--converge-cycles", type=int, default=2000,
                    help="after the timing, run from the initial state until an 8-decade drop "
                         "of the density residual or this many cycles (0: skip; default 2000)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead of the bench: run a plain device copy of MIB MiB three times "
                         "(the known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE "
                         "counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(blocks would shard across GPUs; this build measures one GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "earsm-mlp":
        earsm_mlp_bench(args)
        return
    if args.workload in ("config3", "config5"):
        multiblock_bench(args)
        return

    from finflo_amd.ffhip import Context

    w = WORKLOADS[args.workload]
    nlev, iturb = w["nlev"], w["iturb"]
    prm, blocks, walls = make_workload(args.workload)
    if args.cfl:
        prm.cfl = prm.cfll = args.cfl
    prm.ddadi = 1 if args.implicit == "ddadi" else 0
    ncell = IMAX * JMAX * KMAX
    cpu = None
    if not args.no_cpu:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(prm, blocks, walls, args.cpu_seconds)
        print(f"[bench] CPU baseline {cpu['value']} cell-updates/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    print(f"[bench] workload {args.workload} {IMAX}x{JMAX}x{KMAX} ({ncell} cells, {nlev} "
          f"level(s)), building context", file=sys.stderr, flush=True)
    ctx = Context(prm)
    if args.lusgs:
        ctx.lusgs_mode(args.lusgs)
    ctx.register_levels(blocks)
    def start_state():
        # the initial state of every level, the plate's wall (BOTSKE patches and REFLEC) and
        # the free stream in the ghost cells of the other faces (GHOSTEXT at start-up)
        for m in range(1, nlev + 1):
            ctx.upload(level=m)
            ctx.wall_register(*walls[m - 1], level=m)
            ctx.wall_sync(0, level=m)
            bcs = plate_bcs(blocks[m - 1], outer=args.outer, prm=prm)
            if args.kbc == "periodic":
                bcs = [bp for bp in bcs if bp.iface not in (3, 6)]
            ctx.bc_register(bcs, level=m)
            ctx.ghostext(level=m)
        ctx.finish()

    start_state()
    resi0 = run_steps(ctx, nlev, args.warmup, args.graph)
    ctx.finish()
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    # timed region: K steps, device-resident, bracketed by stream synchronisation
    ctx.finish()
    t0 = time.perf_counter()
    resi = run_steps(ctx, nlev, args.steps, args.graph)
    ctx.finish()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    value = ncell * args.steps / el
    print(f"[bench] {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    if not np.all(np.isfinite(resi)):
        print(f"[bench] non-finite residual norms {resi}", file=sys.stderr, flush=True)
        sys.exit(1)

    # the same K steps in the mixed form (the LU-SGS records in FP32), for the record; the
    # headline value stays the FP64 path's
    mixed = None
    if nlev > 1 or iturb == 0:
        ctx.precision("mixed")
        run_steps(ctx, nlev, max(1, args.warmup))
        ctx.finish()
        t0 = time.perf_counter()
        rm = run_steps(ctx, nlev, args.steps, args.graph)
        ctx.finish()
        elm = time.perf_counter() - t0
        mixed = {"ms_per_step": elm * 1e3 / args.steps, "value": ncell * args.steps / elm,
                 "speedup_over_fp64": el / elm, "finite": bool(np.all(np.isfinite(rm)))}
        ctx.precision("fp64")
        print(f"[bench] mixed form: {mixed['ms_per_step']:.3f} ms per step", file=sys.stderr, flush=True)

    # the same K steps again with a HIP event pair around every launch on the library's
    # stream: per-kernel durations for the roofline (events add host overhead, so they stay
    # out of the timed region above)
    ctx.profile(True)
    ctx.reset_stats()
    run_steps(ctx, nlev, args.steps)
    ctx.finish()
    ctx.profile(False)

    cells = {m: blocks[m - 1].imax * blocks[m - 1].jmax * blocks[m - 1].kmax
             for m in range(1, nlev + 1)}
    ntot = {m: blocks[m - 1].ntot for m in range(1, nlev + 1)}
    kstats = {}
    for name in KERNELS:
        n, ms = ctx.kernel_stats(name)
        if not n:
            continue
        ent = {"launches": n, "ms_total": ms, "ms_per_step": ms / args.steps}
        ncells = 0
        for m in range(1, nlev + 1):
            nm, msm = ctx.kernel_stats(f"{name}@{m}")
            if nm:
                ent[f"level{m}"] = {"launches": nm, "ms_total": msm}
                ncells += nm * (ntot[m] if name in ALL_CELL_KERNELS else cells[m])
        ent["cells_total"] = ncells
        kstats[name] = ent
    models = KERNEL_BYTES_PER_CELL[iturb]
    # dominant kernel: the largest share of the step among the kernels that carry a phase of
    # the iteration; its algorithmic bytes per launch are its phase's compulsory bytes per
    # cell times the mean cells per launch over all its launches (every level), which is
    # what a kernel trace averages over
    dom = max((k for k in kstats if k in PHASE_OF), key=lambda k: kstats[k]["ms_total"])
    ks = kstats[dom]
    avg_ms = ks["ms_total"] / ks["launches"]
    cells_per_launch = ks["cells_total"] / ks["launches"]
    bpc = PHASE_BYTES[iturb][PHASE_OF[dom]]
    achieved = bpc * cells_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, tsrc = committed_traffic(args.workload, dom)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "phase": ("P1 residual", "P2 backward sweep", "P3 forward sweep + update")[PHASE_OF[dom]],
                "avg_launch_ms": avg_ms, "bytes_per_cell": bpc,
                "cells_per_launch": cells_per_launch,
                "design_bytes_per_cell": models.get(dom),
                "design_model_GBps": (models[dom] * cells_per_launch / (avg_ms * 1e-3) / 1e9
                                      if dom in models else None)}
    if traffic:
        roofline["traffic_source"] = tsrc
        roofline["traffic_over_algorithmic"] = traffic / (bpc * cells_per_launch)
        if dom in models:
            roofline["traffic_over_design_model"] = traffic / (models[dom] * cells_per_launch)
    l1 = ks.get("level1")
    if l1:
        roofline["level1_GBps"] = bpc * cells[1] / (l1["ms_total"] / l1["launches"] * 1e-3) / 1e9
    # every phase kernel against its phase bytes, for the record
    roofline["phase_kernels_GBps"] = {
        k: PHASE_BYTES[iturb][PHASE_OF[k]] * kstats[k]["cells_total"] / (kstats[k]["ms_total"] * 1e-3) / 1e9
        for k in kstats if k in PHASE_OF}
    sweep_ms = sum(kstats.get(k, {}).get("ms_total", 0.0)
                   for k in ("k_pred_plane", "k_corr_plane", "k_pred_tiles", "k_corr_tiles",
                             "k_pred_km", "k_corr_km", "k_sweep_prep", "k_pred_st",
                             "k_corr_st")) / args.steps
    # the whole LU-SGS (with the staged form: the record pass and both sweeps) against the
    # bytes the two sweeps must move without any staging (the tile-form byte models)
    corr = next((kstats[k] for k in ("k_corr_st", "k_corr_tiles", "k_corr_km", "k_corr_plane")
                 if k in kstats), None)
    lusgs_gbps = None
    if corr and sweep_ms > 0:
        lusgs_bytes = (models["k_pred_tiles"] + models["k_corr_tiles"]) * corr["cells_total"]
        lusgs_gbps = lusgs_bytes / args.steps / (sweep_ms * 1e-3) / 1e9
    iter_bw = ITER_BYTES_PER_CELL[iturb] * ncell / (ms_per_step * 1e-3) / 1e9

    conv = None
    if args.converge_cycles > 0:
        print(f"[bench] convergence from the initial state, up to {args.converge_cycles} "
              f"{'V-cycles' if nlev > 1 else 'iterations'}", file=sys.stderr, flush=True)
        start_state()
        if args.kbc == "periodic":
            b0 = blocks[0]
            ctx.con_register([(3, 1, b0.imax, 1, b0.jmax, 1, 6, 1, 1, 0, 1, 1),
                              (6, 1, b0.imax, 1, b0.jmax, 1, 3, 1, 1, 0, 1, 1)])
            ctx.halo_exchange(what="geometry")
            ctx.halo_exchange(what="flow")
        conv = converge(ctx, nlev, ncell, args.converge_cycles, periodic=args.kbc == "periodic",
                        dims=(IMAX, JMAX, KMAX))
        conv["k_faces"] = args.kbc
        conv["i_faces"] = "inlet 4 / outlet 3" if args.outer == "inout" else "ext"
        print(f"[bench] convergence: {conv['decades']} decades in {conv['cycles']} cycles, "
              f"{conv['seconds']:.2f} s", file=sys.stderr, flush=True)

    line = {
        "metric": "cell-updates/s per MI355X (residual+LU-SGS) and time to 8-decade residual drop",
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
        "config": {"workload": f"flat plate {IMAX}x{JMAX}x{KMAX} single block, {w['label']}, "
                               f"M={MACH} Re={w['reynolds']:g}, Roe+FLUXDO, "
                               f"{'DDADI' if prm.ddadi else 'LU-SGS'} CFL=5, "
                               f"{'inlet/outlet on the I faces' if args.outer == 'inout' else 'free stream on the outer faces'}; "
                               f"a step is one {'V-cycle' if nlev > 1 else 'iteration'}",
                   "cells": ncell, "levels": nlev,
                   "launch": "HIP graph replay of K-1 steps + 1 eager step (residual sums)"
                   if args.graph else "eager launches",
                   "lusgs_sweep_ms_per_step": sweep_ms,
                   "lusgs_compulsory_GBps": lusgs_gbps,
                   f"iteration_hbm_GBps_at_{ITER_BYTES_PER_CELL[iturb]}B_per_cell": iter_bw,
                   "density_resl2_after_warmup": float(np.sqrt(resi0[0]) / ncell),
                   "density_resl2": float(np.sqrt(resi[0]) / ncell),
                   "mixed_precision_form": mixed,
                   "time_to_8_decade_drop": conv["time_to_8_decades"] if conv else None,
                   "convergence": conv},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kstats,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
