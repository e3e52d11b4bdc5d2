"""Benchmark of the residual-and-update hot path on one MI355X.

One step = one pseudo-time iteration (main.F90:48-211 with Kchld = 1): spectral radius,
BC fills, reconstruction + inviscid (and viscous/SST) fluxes, LU-SGS, update of the state
and the dual-time history, over all blocks of the workload.  The state lives in HBM for the
whole timed region.

Prints ONE JSON line (rank 0) with the throughput in Mcell*iter/s, the HBM roofline of the
dominant kernel (HIP-event timing on the library's own stream) and the CPU oracle timed on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of work of each kernel (FP64 unless stated): every input the
# kernel needs read once and every output written once ("Roofline").
ALG_BYTES = {
    # per interior cell: D(5) r/w, V(1:4)+PP of the cell, 3 faces x SD(3), Rds(3), Dtm, Vol, F1
    "k_lusgs_sweep<double,true>": ("cell", 8 * (5 + 5 + 5 + 9 + 3 + 3)),
    "k_lusgs_sweep<double,false>": ("cell", 8 * (5 + 5 + 5 + 9 + 3 + 3)),
    "k_lusgs_sweep<float,true>": ("cell", 4 * (5 + 5 + 3) + 8 * (5 + 9 + 3)),
    "k_lusgs_sweep<float,false>": ("cell", 4 * (5 + 5 + 3) + 8 * (5 + 9 + 3)),
    # per face: cell V(1:4), PP, T, rad, Alagm (each cell once), SD(3)+Grad, flux out (5)
    "k_muscl_roe<1>": ("face1", 8 * (8 + 4 + 5)),
    "k_muscl_roe<2>": ("face2", 8 * (8 + 4 + 5)),
    "k_muscl_roe<3>": ("face3", 8 * (8 + 4 + 5)),
    # per cell: V(1:4), PP, T, Rmiu, Vol, 3 faces SD(3) in; Rds(3), Dtm out
    "k_spectral_radius": ("cell", 8 * (7 + 1 + 9 + 4)),
    # per cell: D(5) in, 3 directions x 2 faces x flux(5) (faces shared -> 3 x 5), D(5) out
    "k_accumulate": ("cell", 8 * (5 + 15 + 5)),
    # per cell: D(5), V(5), Wt0(5), Wt1(5), Rds(3), Dtm, Vol in; D(5), F1 out
    "k_lusgs_rhs": ("cell", 8 * (5 + 5 + 5 + 5 + 3 + 1 + 1 + 5 + 1)),
    # per cell: D(7), V(7), PP, rad, (DQ 9, Dst, Rmiu when turbulent) in; V(7), PP, T, Rmiu out
    "k_update_v": ("cell", 8 * (7 + 7 + 1 + 1 + 7 + 1 + 1)),
    "k_update_u": ("cell", 8 * (7 + 7 + 7 + 7)),
}


def build_workload(name):
    from trans_amd import cases, solver
    if name == "bump":
        blocks, org, glob = cases.bump_channel()
        desc = "bump channel 129x65x9, Euler, Roe + MUSCL, LU-SGS CFL 5, FP64"
    elif name == "flatplate":
        blocks, org, glob = cases.flat_plate()
        desc = "flat plate 257x129x5 in 8 uneven blocks, RANS-SST, AUSMPW+ + MDCD, LU-SGS, FP64"
    else:
        raise SystemExit(f"unknown workload {name}")
    prm = solver.prepare_case(blocks, org, glob)
    return blocks, prm, desc


def units(blocks):
    u = {"cell": 0, "face1": 0, "face2": 0, "face3": 0}
    for b in blocks:
        n1, n2, n3 = b.ni - 1, b.nj - 1, b.nk - 1
        u["cell"] += n1 * n2 * n3
        u["face1"] += (n1 + 1) * n2 * n3
        u["face2"] += n1 * (n2 + 1) * n3
        u["face3"] += n1 * n2 * (n3 + 1)
    return u


def cpu_baseline(blocks, prm, budget_s=12.0):
    """The CPU oracle (single thread) on the same workload, a bounded number of iterations."""
    from oracle.oracle import OracleCase
    oc = OracleCase(blocks, prm)
    ki = oc.iterate(1, 1)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        ki = oc.iterate(ki, 1)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    return {"value": cells * n / dt / 1e6, "unit": "Mcell*iter/s", "cores": 1, "kind": "port",
            "sample": f"{n} iterations of the same workload after 1 warm-up iteration, "
                      f"oracle/trans_oracle.c single-threaded, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=os.environ.get("TRANS_BENCH_WORKLOAD", "bump"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup+steps without the CPU leg and print nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "only (see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    from trans_amd.solver import Solver
    blocks, prm, desc = build_workload(args.workload)
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    sv = Solver(blocks, prm, ki=1)
    print(f"[bench] {desc}: {cells} cells, warmup {args.warmup}", flush=True)
    sv.iterate(args.warmup, want_resid=False)
    sv.sync()
    sv.profile(True)
    sv.sync()
    t0 = time.perf_counter()
    res = sv.iterate(args.steps, want_resid=True)
    sv.sync()
    dt = time.perf_counter() - t0
    stats = sv.profile_stats()
    sv.profile(False)
    print(f"[bench] {args.steps} steps in {dt:.3f} s, DRSa(block 1) = {res[0]['drsa']:.6e}", flush=True)
    if args.profile_only:
        return
    u = units(blocks)
    # dominant kernel: largest total device time
    name, (ms, n) = max(stats.items(), key=lambda kv: kv[1][0])
    avg_s = ms / 1e3 / n
    unit, per = ALG_BYTES.get(name, ("cell", None))
    roof = None
    if per is not None:
        alg = per * u[unit]
        ach = alg / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "kernel": name,
                "avg_launch_us": round(avg_s * 1e6, 3), "alg_bytes_per_launch": alg}
    value = cells * args.steps / dt / 1e6
    cpu = None if args.no_cpu else cpu_baseline(blocks, prm)
    out = {
        "metric": "Mcell*iter/s (residual + LU-SGS update, FP64 residual)",
        "value": round(value, 3), "unit": "Mcell*iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated grid, uniform start)",
        "config": {"workload": desc, "cells": cells, "blocks": len(blocks)},
        "roofline": roof, "cpu_baseline": cpu,
        "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
