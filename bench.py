#!/usr/bin/env python3
"""Photon-transport benchmark: cell-crossings/s on one MI355X.

A step is one kp_forward launch over the whole workload (every packet of
every band: emission, march, scattering, peel-off).  The timed region
covers exactly --steps launches with inputs resident in HBM, bracketed by
device synchronisation; kernel time is also taken with HIP events on the
launch stream.  One JSON line is printed.

The headline value is the mixed-precision path (FP32 block-local transport,
FP64 optical depth and tallies); the plain FP64 path is timed on the same
workload, and the Stokes-image error of both is reported against an FP64
reference run with an independent packet range.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="config1")
    p.add_argument("--packets", type=float, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    return p.parse_args()


def algorithmic_bytes(cnt, n_incl):
    """B_alg = 20 B per crossing (4 B kappa + 16 B FP64 tally read-modify-write;
    peel-off crossings read kappa only: 4 B) + 48 B per peel-off pixel update."""
    main, peel, events, packets = cnt[0], cnt[1], cnt[2], cnt[7]
    peel_px = (events + packets) * n_incl
    return 20.0 * main + 4.0 * peel + 48.0 * peel_px


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    from kratos_polrad_amd import _abi, configs
    from kratos_polrad_amd.engine import Engine, RunSpec

    if args.workload == "config1":
        n = int(args.packets) or 1_000_000
        w = configs.config1(n_packets=n, precision=_abi.KP_F32)
    else:
        raise SystemExit("unknown workload")
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    stream = torch.cuda.current_stream()
    n_incl = len(w.run.incl_deg)

    def timed(prec, steps, warmup, offset0=0):
        run = RunSpec(**{**w.run.__dict__, "precision": prec})
        out = eng.alloc_outputs(run)
        for k in range(warmup):
            run.packet_offset = offset0 + 10**12 + k * run.total_packets
            eng.forward(w.source, run, out)
        torch.cuda.synchronize()
        for v in out.values():
            v.zero_()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        for k in range(steps):
            run.packet_offset = offset0 + k * run.total_packets
            eng.forward(w.source, run, out)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = ev0.elapsed_time(ev1) / 1e3
        cnt = out["counters"].cpu().numpy().astype(np.float64)
        img = out["image"].cpu().numpy() if "image" in out else None
        return wall, kern, cnt, img

    wall32, kern32, cnt32, img32 = timed(_abi.KP_F32, args.steps, args.warmup)
    wall64, kern64, cnt64, img64 = timed(_abi.KP_F64, args.steps, args.warmup)
    cross32 = cnt32[0] + cnt32[1]
    cross64 = cnt64[0] + cnt64[1]
    value = cross32 / wall32
    b_alg = algorithmic_bytes(cnt32, n_incl) / args.steps
    achieved = b_alg / (kern32 / args.steps)

    # Stokes-image error of each path against an independent FP64 run
    def img_err(img, ref):
        a = img[..., 0].sum(axis=(0, 1))
        b = ref[..., 0].sum(axis=(0, 1))
        return float(np.abs(a - b).sum() / np.abs(b).sum())

    _, _, _, ref = timed(_abi.KP_F64, args.steps, 0, offset0=10**13)
    err32 = img_err(img32, ref)
    err64 = img_err(img64, ref)

    cpu = None
    if not args.no_cpu:
        import oracle

        n_cpu = max(1000, int(2e5 * args.cpu_seconds / 15.0))
        run = RunSpec(**{**w.run.__dict__, "precision": _abi.KP_F64})
        run.bands = [type(run.bands[0])(**{**run.bands[0].__dict__, "n_packets": n_cpu})]
        cores = len(os.sched_getaffinity(0))
        cores = min(cores, 16)
        t0 = time.perf_counter()
        o = oracle.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, run, n_threads=cores)
        dt = time.perf_counter() - t0
        c = o["counters"].astype(np.float64)
        cpu = {"value": float((c[0] + c[1]) / dt), "unit": "cell-crossings/s", "cores": cores,
               "kind": "port",
               "sample": f"{n_cpu} packets of the same workload, FP64 oracle, {dt:.1f} s"}

    res = {
        "metric": "cell-crossings/s at fixed Stokes-image error vs FP64",
        "value": float(value),
        "unit": "cell-crossings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transport / f64 tau+tallies",
        "data": "synthetic",
        "config": {"workload": w.name, "packets": w.run.total_packets,
                   "blocks": w.mesh.n_blocks, "cells": w.mesh.n_cells,
                   "n_incl": n_incl, "image_n": w.run.image_n},
        "fp64_path": {"value": float(cross64 / wall64), "ms_per_step": 1e3 * wall64 / args.steps,
                      "image_err_vs_fp64": err64},
        "mixed_path_image_err_vs_fp64": err32,
        "speedup_mixed_vs_fp64": float((cross32 / wall32) / (cross64 / wall64)),
        "crossings_per_step": {"main": cnt32[0] / args.steps, "peel": cnt32[1] / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
