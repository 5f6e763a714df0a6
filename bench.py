"""Benchmark: planned FP32 map2tod + tod2map on one MI355X.

Workload (single-TOD full scale): ndet=2048 x nt=262144 samples (5.4e8), constant-elevation
raster x = 4096 + 3000 tri(t/40000) + dx, y = 500 + 0.01 t + dy (detector offsets uniform in a
disk of radius 300 px), alpha = theta_det + 2 pi t / 1e6, on a 3x4096x8192 global map held as a
sparse LocalPixelization of the 64x64 cells the scan touches; map values are a Gaussian random
field with k^-2 power per component, TOD values N(0,1). Synthetic data, generated on the GPU.

A step is one planned map2tod plus one planned tod2map over the whole TOD (inputs resident in
HBM). value = samples per second through the step, in Gsamp/s.

Prints one JSON line (rank 0). Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_SAMPLE = 16      # 12 B pointing + 4 B TOD
CELL_BYTES = 3 * 64 * 64 * 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(torch, gmm, synth, dev, ndet, nt):
    c = synth.CFG1
    t0 = time.time()
    xp, _ = synth.raster_xpointing_torch(ndet, nt, seed=1, device=dev, **c["scan"])
    torch.cuda.synchronize()
    log(f"[bench] pointing {tuple(xp.shape)} generated in {time.time() - t0:.1f}s")

    ny, nx = c["ny"], c["nx"]
    t0 = time.time()
    pre = gmm.PointingPrePlan(xp, ny, nx, periodic_xcoord=False)
    plan = gmm.PointingPlan(pre, xp)
    torch.cuda.synchronize()
    log(f"[bench] plan nmt={plan.plan_nmt} nchunks={plan.nchunks} nmlines={plan.nmlines} "
        f"overhead={pre.overhead:.3f} built in {time.time() - t0:.2f}s")

    cells = np.unique(plan.get_plan_mt() & np.uint64(0xFFFFF))
    cy, cx = (cells >> np.uint64(10)).astype(np.int64), (cells & np.uint64(1023)).astype(np.int64)
    lp = gmm.LocalPixelization.make_cells(ny, nx, list(zip(cy.tolist(), cx.tolist())), periodic_xcoord=False)
    ncells = len(cells)

    grf = synth.grf_map_torch(ny, nx, seed=2, device=dev)
    g = grf.reshape(3, ny // 64, 64, nx // 64, 64).permute(1, 3, 0, 2, 4)
    lmap = g[torch.from_numpy(cy).to(dev), torch.from_numpy(cx).to(dev)].contiguous().reshape(-1)
    del grf, g
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    tod = torch.randn((ndet, nt), generator=gen, device=dev, dtype=torch.float32)
    torch.cuda.synchronize()
    log(f"[bench] local map: {ncells} cells ({3 * lp.npix * 4 / 1e6:.0f} MB)")
    return dict(xp=xp, pre=pre, plan=plan, lp=lp, lmap=lmap, tod=tod, ncells=ncells, ny=ny, nx=nx)


def time_plan_construction(torch, gmm, w, reps=3):
    xp, pre = w["xp"], w["pre"]
    buf = torch.empty(pre.plan_nbytes, dtype=torch.uint8, device=xp.device)
    tmp = torch.empty(pre.plan_constructor_tmp_nbytes, dtype=torch.uint8, device=xp.device)
    t_pre, t_plan = [], []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = gmm.PointingPrePlan(xp, w["ny"], w["nx"], periodic_xcoord=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gmm.PointingPlan(p, xp, buf=buf, tmp_buf=tmp)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_pre.append(t1 - t0)
        t_plan.append(t2 - t1)
    return 1e3 * float(np.median(t_pre)), 1e3 * float(np.median(t_plan))


def cpu_baseline(w, ndet_sample, torch):
    """Oracle (the CPU restatement of reference_map2tod/tod2map) on the host cores, on the first
    ndet_sample detectors of the same workload."""
    from oracle import pyoracle as orc
    nt = w["tod"].shape[1]
    xp = w["xp"][:, :ndet_sample, :].contiguous().cpu().numpy()
    lmap = w["lmap"].cpu().numpy()
    tod = w["tod"][:ndet_sample].contiguous().cpu().numpy()
    lp = w["lp"]
    pix = orc.Pix(w["ny"], w["nx"], lp.cell_offsets_cpu, lp.ystride, lp.polstride, False)
    n = ndet_sample * nt
    t0 = time.perf_counter()
    _, e1 = orc.map2tod(lmap, xp, pix)
    t1 = time.perf_counter()
    _, e2 = orc.tod2map(tod, xp, pix)
    t2 = time.perf_counter()
    assert e1 == 0 and e2 == 0
    st = n / (t2 - t0) / 1e9
    nthr = orc.max_threads()
    t3 = time.perf_counter()
    orc.map2tod(lmap, xp, pix, nthreads=nthr)
    t4 = time.perf_counter()
    orc.tod2map(tod, xp, pix, nthreads=nthr)
    t5 = time.perf_counter()
    mt = n / (t5 - t3) / 1e9
    log(f"[bench] cpu baseline: 1 thread {st:.4f} Gsamp/s ({t2 - t0:.1f}s); {nthr} threads {mt:.4f} Gsamp/s "
        f"(map2tod {t4 - t3:.1f}s, tod2map {t5 - t4:.1f}s)")
    return dict(value=round(st, 5), unit="Gsamp/s", cores=1, kind="port",
                sample=f"first {ndet_sample} of 2048 detectors x {nt} samples ({n} samples), map2tod + tod2map, "
                       f"single thread (oracle/oracle.c)",
                multicore=dict(value=round(mt, 5), cores=nthr, threads_kind="OpenMP, tod2map with private maps"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ndet", type=int, default=2048)
    ap.add_argument("--nt", type=int, default=262144)
    ap.add_argument("--cpu-ndet", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import gpu_mm_amd as gmm
    from gpu_mm_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    w = build_workload(torch, gmm, synth, dev, args.ndet, args.nt)
    nsamp = args.ndet * args.nt
    lmap = gmm.LocalMap(w["lp"], w["lmap"])
    out_map = gmm.LocalMap(w["lp"], torch.zeros_like(w["lmap"]))
    tod_out = torch.empty_like(w["tod"])
    plan, xp, tod = w["plan"], w["xp"], w["tod"]

    gmm.set_errflag_check(False)

    def step():
        gmm.map2tod(tod_out, lmap, xp, plan)
        gmm.tod2map(out_map, tod, xp, plan)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: exactly K steps.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    plan.check_errflags("bench")

    # Per-operator kernel time with HIP events on the launch stream (torch's current stream).
    stream = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    for i in range(args.steps):
        ev[i][0].record(stream)
        gmm.map2tod(tod_out, lmap, xp, plan)
        ev[i][1].record(stream)
        ev[i][2].record(stream)
        gmm.tod2map(out_map, tod, xp, plan)
        ev[i][3].record(stream)
    torch.cuda.synchronize()
    plan.check_errflags("bench")
    t_m2t = float(np.mean([e[0].elapsed_time(e[1]) for e in ev])) * 1e-3
    t_t2m = float(np.mean([e[2].elapsed_time(e[3]) for e in ev])) * 1e-3

    ncells = w["ncells"]
    bytes_m2t = BYTES_PER_SAMPLE * nsamp + CELL_BYTES * ncells
    bytes_t2m = BYTES_PER_SAMPLE * nsamp + 2 * CELL_BYTES * ncells
    ops = {
        "map2tod": dict(seconds=t_m2t, bytes=bytes_m2t),
        "tod2map": dict(seconds=t_t2m, bytes=bytes_t2m),
    }
    for k, v in ops.items():
        v["gsamp_s"] = nsamp / v["seconds"] / 1e9
        v["gb_s"] = v["bytes"] / v["seconds"] / 1e9
        v["frac_hbm"] = v["bytes"] / v["seconds"] / HBM_PEAK
        log(f"[bench] {k}: {v['seconds'] * 1e3:.3f} ms  {v['gsamp_s']:.1f} Gsamp/s  "
            f"{v['gb_s']:.0f} GB/s  ({100 * v['frac_hbm']:.1f}% of 8 TB/s)")
    dom = max(ops, key=lambda k: ops[k]["seconds"])

    traffic = None
    tfile = os.path.join(REPO, "profiles", f"traffic_{dom}_cfg1.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    if args.copy_ref:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t_pre, t_plan = time_plan_construction(torch, gmm, w)
    log(f"[bench] PointingPrePlan {t_pre:.2f} ms, PointingPlan {t_plan:.2f} ms")

    cpu = None if args.no_cpu else cpu_baseline(w, min(args.cpu_ndet, args.ndet), torch)

    ms_per_step = 1e3 * t_total / args.steps
    value = nsamp / (t_total / args.steps) / 1e9
    a = ops[dom]["bytes"] / ops[dom]["seconds"]
    line = {
        "metric": "Gsamp/s map2tod+tod2map & %HBM peak, chosen vs plain FP32 at <=1e-4 rel err",
        "value": round(value, 3),
        "unit": "Gsamp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated raster pointing, GRF map, N(0,1) TOD)",
        "config": {
            "workload": "planned map2tod+tod2map, plain FP32, single TOD 2048x262144 on sparse 3x4096x8192 map",
            "ndet": args.ndet, "nt": args.nt, "nsamp": nsamp, "cells": ncells, "plan_nmt": int(plan.plan_nmt),
            "chunks": int(plan.nchunks), "form": "plain FP32 (chosen form)",
        },
        "ops": {k: {"ms": round(v["seconds"] * 1e3, 4), "gsamp_s": round(v["gsamp_s"], 2),
                    "gb_s": round(v["gb_s"], 1), "frac_hbm": round(v["frac_hbm"], 4)} for k, v in ops.items()},
        "plan_ms": {"preplan": round(t_pre, 3), "plan": round(t_plan, 3)},
        "roofline": {
            "kernel": f"planned_{dom}_kernel",
            "bound": "hbm",
            "achieved": round(a / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(a / HBM_PEAK, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": int(ops[dom]["bytes"]),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
