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
METRIC = "Mcell\u00b7iter/s (RANS-SST, FP64 residual); solves/hr + p50/p99 to 1e-8 drop"
UNIT = "Mcell\u00b7iter/s"
# compulsory HBM bytes per cell of one fused iteration, FP64: 8*(N_in + N_geom + N_out) + 2*8*N_eq
ITER_BYTES_RANS = 8 * (10 + 20 + 10) + 2 * 8 * 7    # 432
ITER_BYTES_EULER = 8 * (7 + 17 + 7) + 2 * 8 * 5    # 328
# one explicit stage with k-omega, FP64: state, p, T, mu_t in and out, geometry, and the stage's
# saved state Wt2 (7) and the residual D (5) in
ITER_BYTES_RK_STAGE = 8 * (10 + 20 + 10) + 8 * (7 + 5)  # 416


def alg_bytes(name, neq, fp32, dual_shift=True):
    """Algorithmic bytes per unit of work of a kernel: every input it needs read once and
    every output written once, in the precision it is stored in.  Returns (unit, bytes).
    dual_shift: the update carries the dual-time history shift (every deck here)."""
    d = 8
    r = 4 if fp32 else 8
    turb = neq == 7
    table = {
        # V(1:4), PP, T, Rmiu, Vol, SD (9 face values per cell) in; Rds(3), Dtm out
        "k_spectral_radius": ("cell", d * (4 + 1 + 1 + 1 + 1 + 9 + 3 + 1) + (4 * 3 if fp32 else 0)),
        # V(1:4[,6:7]) in; Q(3[,5]) out
        "k_prim_cells": ("cell", d * ((6 if turb else 4) + (5 if turb else 3))),
        # Q(3[,5]), T, Alagm(3), SD(9), Vol in; DQ(12[,18]), shock out
        "k_gradients": ("cell", d * ((5 if turb else 3) + 1 + 3 + 9 + 1 + (18 if turb else 12) + 1)),
        # per face of direction d: cell Q(5), V(1), T, Rmiu, Dst, Alagm, Vol, DQ(18) once; SD(3); GV(6) out
        "k_visc_face<1>": ("face1", d * (6 + 1 + 1 + 1 + 1 + 1 + 18 + 3 + 6)),
        "k_visc_face<2>": ("face2", d * (6 + 1 + 1 + 1 + 1 + 1 + 18 + 3 + 6)),
        "k_visc_face<3>": ("face3", d * (6 + 1 + 1 + 1 + 1 + 1 + 18 + 3 + 6)),
        # per face: cell Q(3[,5]), V(1), PP, T, rad, Alagm, shock once; SD(3), Grad; G(5[,7]) out
        "k_conv_face<1>": ("face1", d * ((6 if turb else 4) + 5 + 4 + (7 if turb else 5))),
        "k_conv_face<2>": ("face2", d * ((6 if turb else 4) + 5 + 4 + (7 if turb else 5))),
        "k_conv_face<3>": ("face3", d * ((6 if turb else 4) + 5 + 4 + (7 if turb else 5))),
        # ghost cells of the three layers beyond each face: V(1:4[,6:7]) in, Q(3[,5]) out
        "k_prim_ghost": ("ghost", d * ((6 if turb else 4) + (5 if turb else 3))),
        # per cell: G 3 x 7, GV 3 x 6 (viscous), V(1,3,4), Yc, Zc, Vol; D(7) out
        "k_accumulate_visc": ("cell", d * (21 + 18 + 6 + 7)),
        # D(5), V(5), Wt0(5), Wt1(5), Rds(3), Dtm, Vol in; D(5), F1 out
        "k_lusgs_rhs": ("cell", d * (5 + 5 + 5 + 5 + 3 + 1 + 1 + 5 + 1)),
        # D(6:7), V(1,6,7), Wt0, Wt1 (2 each), DQ(15), Rmiu, T, Dst, Dtm, Rds(3), Vol in; D(6:7), F1t out
        "k_sst_rhs": ("cell", d * (2 + 3 + 2 + 2 + 15 + 1 + 1 + 1 + 1 + 3 + 1 + 2 + 1)),
        # hyperplane-major inputs in the operator's precision: D(neq) in/out, stept/F1 and
        # cvs/F1 (mean flow), (stept, cvs, F1t for k-omega), 1/rho and u, v, w, PP, face means
        # SD(9), Rds(3)
        "k_lusgs_line<fwd>": ("cell", r * (2 * neq + 2 + (3 if turb else 0) + 4 + 1 + 9 + 3)),
        "k_lusgs_line<bwd>": ("cell", r * (2 * neq + 2 + (3 if turb else 0) + 4 + 1 + 9 + 3)),
        # tile wavefront: the same inputs and increments as the line march, every input read
        # once (the mean-flow and k-omega tiles share 1/rho, u, v, w, SD means and Rds)
        "k_lusgs_tile<fwd>": ("cell", r * (2 * neq + (5 if turb else 2) + 4 + 1 + 9 + 3)),
        "k_lusgs_tile<bwd>": ("cell", r * (2 * neq + (5 if turb else 2) + 4 + 1 + 9 + 3)),
        # D(5), V(5), Wt0(5), Wt1(5), Rds(3), Dtm, Vol in; 17 sweep inputs out in the operator's precision
        "k_lusgs_rhs_w": ("cell", d * (5 + 5 + 5 + 5 + 3 + 1 + 1) + r * 17),
        "k_sst_rhs_w": ("cell", d * (2 + 3 + 2 + 2 + 15 + 1 + 1 + 1 + 1 + 3 + 1) + r * 3),
        "k_unskew_w": ("cell", (r + d) * neq),
        # D(neq), Dtm, Vol, F1(,F1t), V(1:4), PP, SD(9), Rds(3) in the natural layout
        "k_lusgs_sweep<fwd>": ("cell", 2 * neq * r + d * (1 + 1 + (2 if turb else 1) + 4 + 1 + 9) + 3 * r),
        "k_lusgs_sweep<bwd>": ("cell", 2 * neq * r + d * (1 + 1 + (2 if turb else 1) + 4 + 1 + 9) + 3 * r),
        # D(7), V(7), PP, rad, DQ(9), Dst in; V(7), PP, T, Rmiu out; with the dual-time shift
        # fused into the last sub-iteration (Kind_Dual 1, 2) also Wt0(7), Wt1(7) in and Wt0, Wt1,
        # Wt2 out
        "k_update_v": ("cell", d * (7 + 7 + 1 + 1 + 9 + 1 + 7 + 1 + 1 + 1 + (35 if dual_shift else 0))),
        # V(7), Wt0(7) in; Wt0(7), Wt1(7) out
        "k_update_u": ("cell", d * 28),
    }
    return table.get(name, ("cell", None))


def build_workload(name, implicit_fp32=0, scale=1.0, back_pressure=1.2, deck=()):
    """(blocks, prepared params, description) of a named workload; `deck` holds KEY=VALUE
    overrides of original.org values (OrgDeck field names) applied before preparation."""
    from trans_amd import cases, solver
    if name == "bump":
        blocks, org, glob = cases.bump_channel()
        desc = "bump channel 129x65x9, Euler, Roe + MUSCL, LU-SGS CFL 5"
    elif name == "flatplate":
        blocks, org, glob = cases.flat_plate()
        desc = "flat plate 257x129x5 in 8 uneven blocks, RANS-SST, AUSMPW+ + MDCD, LU-SGS CFL 5"
    elif name == "plate8m":
        # the same deck at the 8M-cell scale: 1025x513x17 points in 16 uneven blocks
        # (wall-normal stretching 1.0157 so that 512 cells from y+ = 1 reach the plate length)
        blocks, org, glob = cases.flat_plate(nblk=16, ni_total=1025, nj=513, nk=17, lo=33, hi=97, ratio=1.0157)
        desc = "flat plate 1025x513x17 (8.4M cells) in 16 uneven blocks, RANS-SST, AUSMPW+ + MDCD, LU-SGS CFL 5"
    elif name == "compressor":
        blocks, org, glob = cases.compressor_sector(scale=scale, back_pressure=back_pressure)
        ncell = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
        desc = (f"compressor-rotor sector, 16 blocks ({ncell / 1e6:.2f}M cells), 1200 rad/s, periodic 10 deg, "
                f"outlet {back_pressure:g} x inlet P0, RANS-SST, Roe + MUSCL, LU-SGS CFL 5")
    elif name == "cascade":
        blocks, org, glob = cases.cascade(scale=scale)
        ncell = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
        desc = (f"blunt-trailing-edge linear cascade, 12 blocks ({ncell / 1e6:.2f}M cells), exit Mach 0.7, "
                f"Re_c 1e6, SST-DES, AUSMPW+ + MDCD, RK3 at CFL 0.8 with a global time step")
    else:
        raise SystemExit(f"unknown workload {name}")
    for kv in deck:
        key, _, val = kv.partition("=")
        if not hasattr(org, key):
            raise SystemExit(f"unknown deck value {key}")
        setattr(org, key, type(getattr(org, key))(float(val)) if isinstance(getattr(org, key), int)
                else float(val))
        desc += f", {key}={val}"
    prm = solver.prepare_case(blocks, org, glob, implicit_fp32=implicit_fp32,
                              device_geometry=name in ("plate8m", "compressor", "cascade"))
    if name == "cascade":
        # Kind_Dual 0: one global step, CFL 0.8 in the cell that limits it
        prm.dt_lusgs_main = cases.global_time_step(blocks, prm, org.CFL_Sub_end)
    return blocks, prm, desc


def wake_stats(blocks, prm, steps, desc):
    """Time-averaged mean and RMS statistics (Output/Average_Field.F90, 23 fields) over the
    second half of `steps` iterations, on the batched path and on the plain drop-in path from
    the same start; prints their largest relative difference per block and the throughput of
    both."""
    import numpy as np
    from trans_amd.mesh import sl
    from trans_amd.solver import Solver
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    ki_aver = steps // 2 + 1
    out, stats = {}, []
    for plain in (False, True):
        sv = Solver(blocks, prm, ki=1)
        if plain:
            sv.set_tuning(-3, 1, 0, 16, 16, 8)
        sv.stats_enable(ki_aver, ki_aver)
        t0 = time.perf_counter()
        res = sv.iterate(steps)
        sv.sync()
        dt = time.perf_counter() - t0
        tag = "plain" if plain else "batched"
        out[tag] = {"seconds": round(dt, 3), "Mcell_iter_per_s": round(cells * steps / dt / 1e6, 3),
                    "drsa_block1": res[0]["drsa"]}
        print(f"[wake] {tag}: {steps} steps in {dt:.2f} s", flush=True)
        st = []
        for n, b in enumerate(blocks):
            a = sv.get_stats(n)[:, sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)]
            st.append(a.astype(np.float64, copy=True))
        stats.append(st)
        sv.close()
    worst = []
    for a, b in zip(*stats):
        den = np.maximum(np.max(np.abs(b), axis=(1, 2, 3)), 1e-300)
        worst.append(float(np.max(np.max(np.abs(a - b), axis=(1, 2, 3)) / den)))
    out.update({"workload": desc, "steps": steps, "statistics_from_iteration": ki_aver,
                "stats_relerr_per_block": worst, "stats_relerr_max": max(worst)})
    print(json.dumps({"wake_stats": out}), flush=True)


def serving_leg(n_req=16, seed=20261202, target=0.5, max_iter=4000, mean_gap_s=1.0,
                budget_gib=96, max_active=8, chunk=50):
    """Continuous-batching serving of a Poisson stream of steady compressor-sector solves.

    Requests draw a grid scale of the compressor-rotor sector (0.49M, 2.0M or 8.13M cells) and
    an outlet back-pressure ratio in [0.9, 1.3] (Pouthub, BC/BC_Outlet.F90:49-70) from the
    seed and arrive with exponential gaps; each runs the sector's RANS-SST deck (Roe + MUSCL,
    LU-SGS, FP32 implicit operator, CFL ramp 1 -> 5) to a residual drop of `target` (or
    `max_iter`).  The server admits them by HBM footprint (budget `budget_gib`) into at most
    `max_active` resident slots and interleaves their iterations.  Latency = completion -
    submission.  One prepared case per grid size serves every request of that size: a request
    differs only in its outlet pressure."""
    import numpy as np
    from trans_amd import capi, cases, serving, solver
    rng = np.random.Generator(np.random.MT19937(seed))
    scales = [0.397, 0.63, 1.0]
    reqs = []
    t = 0.0
    for _ in range(n_req):
        t += rng.exponential(mean_gap_s)
        reqs.append((t, int(rng.integers(0, 3)), float(rng.uniform(0.9, 1.3))))
    built, ncell = {}, {}
    for q in sorted({r[1] for r in reqs}):
        blocks, org, glob = cases.compressor_sector(scale=scales[q], back_pressure=1.0)
        built[q] = (blocks, solver.prepare_case(blocks, org, glob, implicit_fp32=1, device_geometry=True))
        ncell[q] = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    print(f"[serving] cases prepared: {sorted(ncell.values())} cells", file=sys.stderr, flush=True)

    def params(q, bp):
        p = capi.TransParams.from_buffer_copy(built[q][1])
        p.pouthub = built[q][1].pouthub * bp  # built at Pouthub = 1.0 x P0in
        return p

    srv = serving.Server(budget_gib << 30, max_active=max_active, chunk=chunk)
    ids = []
    t0 = time.perf_counter()
    nxt = 0
    unfinished = 0
    while nxt < n_req or unfinished:
        now = time.perf_counter() - t0
        while nxt < n_req and reqs[nxt][0] <= now:
            _, q, bp = reqs[nxt]
            ids.append(srv.submit(built[q][0], params(q, bp), target, max_iter))
            nxt += 1
        due = nxt < n_req and reqs[nxt][0] - now < 0.01
        left = srv.progress(wait=not due)
        if left < unfinished or (nxt == n_req and left == 0):
            # one line per completion: the solves submitted and the ones still queued or running
            print(f"[serving] {nxt} submitted, {left} queued or running, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        unfinished = left
        if not unfinished and nxt < n_req:
            time.sleep(max(0.0, min(0.01, reqs[nxt][0] - (time.perf_counter() - t0))))
    sts = [srv.poll(i) for i in ids]
    srv.close()
    lat = np.array([s["t_done"] - s["t_submit"] for s in sts])
    span = max(s["t_done"] for s in sts) - min(s["t_submit"] for s in sts)
    cell_iters = sum(ncell[q] * s["iterations"] for (_, q, _), s in zip(reqs, sts))
    return {"requests": n_req, "converged": sum(s["state"] == 2 for s in sts), "target_drop": target,
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default"),
            "solves_per_hr": round(n_req / span * 3600, 1), "p50_s": round(float(np.percentile(lat, 50)), 3),
            "p99_s": round(float(np.percentile(lat, 99)), 3), "makespan_s": round(span, 3),
            "mcell_iter_per_s": round(cell_iters / span / 1e6, 2),
            "cells": [ncell[q] for (_, q, _) in reqs],
            "back_pressure": [round(bp, 4) for (_, _, bp) in reqs],
            "iterations": [s["iterations"] for s in sts],
            "iterations_to_target": [s["iterations_to_target"] for s in sts],
            "workload": f"Poisson stream (mean gap {mean_gap_s} s, seed {seed}) of compressor-rotor sector solves "
                        f"(RANS-SST, Roe + MUSCL, LU-SGS CFL ramp 1 -> 5, FP64 residual / FP32 implicit), grids "
                        f"0.49M / 2.0M / 8.13M cells, outlet back pressure U[0.9, 1.3] x inlet P0, to a {target:g} "
                        f"residual drop (cap {max_iter}); budget {budget_gib} GiB, {max_active} slots, "
                        f"{chunk}-iteration chunks"}


def units(blocks):
    u = {"cell": 0, "face1": 0, "face2": 0, "face3": 0, "ghost": 0}
    for b in blocks:
        n1, n2, n3 = b.ni - 1, b.nj - 1, b.nk - 1
        u["cell"] += n1 * n2 * n3
        u["ghost"] += 6 * (n2 * n3 + n1 * n3 + n1 * n2)
        u["face1"] += (n1 + 1) * n2 * n3
        u["face2"] += n1 * (n2 + 1) * n3
        u["face3"] += n1 * n2 * (n3 + 1)
    return u


def cpu_baseline(blocks, prm, budget_s=12.0):
    """The CPU oracle on the same workload for a bounded number of iterations: OpenMP over
    blocks with as many threads as blocks (capped by OMP_NUM_THREADS), then one thread."""
    from oracle.oracle import OracleCase, lib
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)

    def run(threads, budget):
        lib().ora_set_threads(threads)
        oc = OracleCase(blocks, prm)
        ki = oc.iterate(1, 1)  # warm
        n, t0 = 0, time.perf_counter()
        while True:
            ki = oc.iterate(ki, 1)
            n += 1
            if time.perf_counter() - t0 > budget or n >= 200:
                break
        dt = time.perf_counter() - t0
        return cells * n / dt / 1e6, n, dt

    # the process's CPU share (OMP_NUM_THREADS; 16 on the GPU hosts here), at most one thread per
    # block: the oracle parallelises over blocks, which the 16 uneven blocks of the compressor
    # sector balance to within a few per cent (largest block 524K of 8.13M cells)
    share = int(os.environ.get("OMP_NUM_THREADS", "16"))
    threads = max(1, min(share, len(blocks), os.cpu_count() or 1))
    cells_b = [(b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks]
    balance = sum(cells_b) / threads / max(cells_b) if threads > 1 else 1.0
    v, n, dt = run(threads, budget_s)
    v1, n1, dt1 = run(1, budget_s / 2)
    return {"value": round(v, 4), "unit": UNIT, "cores": threads, "kind": "port",
            "physical_cores_host": physical_cores(), "cpu_share": share,
            "sample": f"{n} iterations of the same workload after 1 warm-up iteration, "
                      f"oracle/trans_oracle*.c with OpenMP over blocks on {threads} threads (CPU share "
                      f"{share}, block balance {balance:.2f}), {dt:.1f} s; single thread: {v1:.4f} {UNIT} "
                      f"({n1} iterations, {dt1:.1f} s)"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/hbm_traffic.json, counters scaled by a plain-copy calibration), or None.
    Names match on the kernel name and on its direction: the sweep direction (fwd = a `true`
    template argument) or the face direction (the first template argument)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        tab = json.load(fh).get(workload, {})
    kb, _, ka = kernel.partition("<")
    ka = ka.rstrip(">")
    for name, rec in tab.items():
        nb, _, na = name.partition("<")
        args = [a.strip() for a in na.rstrip(">").split(",")] if na else []
        if nb != kb:
            continue
        if ka in ("fwd", "bwd"):
            ok = ("true" in args) == (ka == "fwd")
        else:
            ok = not ka or (bool(args) and args[0] == ka)
        if ok:
            return rec["bytes"]
    return None


def _cfg_name(r):
    s = f"sweep{r['sweep']}/streams{r['streams']}/graph{r['graph']}"
    return s + (f"/tile{r['tile_a']}x{r['tile_b']}/chunk{r['chunk']}" if r["sweep"] == 3 else "")


def timed_run(blocks, prm, steps, warmup, prof_steps, tune=0, force=None):
    """Warm up, let the library's autotuner pick the launch configuration (`tune` timed
    iterations per candidate, state restored afterwards; 0 keeps the default), time `steps`
    iterations without instrumentation, then profile `prof_steps` more with HIP events on the
    library's stream."""
    from trans_amd.solver import Solver
    sv = Solver(blocks, prm, ki=1)
    if force is not None:
        sv.set_tuning(*force)
    sv.iterate(warmup, want_resid=False)
    sv.sync()
    timed_run.tuning = None
    if tune > 0:
        chosen, table = sv.autotune(iters=tune)
        timed_run.tuning = {"chosen": {k: chosen[k] for k in ("sweep", "streams", "graph", "tile_a", "tile_b",
                                                              "chunk")},
                            "ms_per_iter": {_cfg_name(r): round(r["ms_per_iter"], 4) for r in table}}
        # one untimed iteration captures the chosen configuration's graph outside the timed region
        sv.iterate(1, want_resid=False)
        sv.sync()
    t0 = time.perf_counter()
    res = sv.iterate(steps, want_resid=True)
    sv.sync()
    dt = time.perf_counter() - t0
    stats = {}
    if prof_steps > 0:
        # per-kernel durations on one stream without graph replay, so that no kernel's
        # duration includes time it shared the device with the helper streams' kernels
        t = sv.tuning()
        if t["sweep"] != -3:
            sv.set_tuning(t["sweep"], 1, 0, t["tile_a"], t["tile_b"], t["chunk"])
        sv.profile(True)
        sv.iterate(prof_steps, want_resid=False)
        sv.sync()
        stats = sv.profile_stats()
        sv.profile(False)
    sv.close()
    return dt, res, stats


def global_drsa(res, blocks):
    """sqrt(sum_b N_b DRSa_b^2 / sum_b N_b) over the blocks."""
    n = [(b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks]
    return (sum(nb * r["drsa"] ** 2 for nb, r in zip(n, res)) / sum(n)) ** 0.5


def converge(blocks, prm, desc, max_iter, drop=1e-8, chunk=100):
    """Iterate to `drop` x the first iteration's global DRSa (time-to-convergence)."""
    from trans_amd.solver import Solver
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    sv = Solver(blocks, prm, ki=1)
    t0 = time.perf_counter()
    r0 = global_drsa(sv.iterate(1), blocks)
    it, r = 1, r0
    print(f"[converge] {desc}: iteration 1 DRSa {r0:.4e}", flush=True)
    while it < max_iter and r > drop * r0:
        n = min(chunk, max_iter - it)
        r = global_drsa(sv.iterate(n), blocks)
        it += n
        print(f"[converge] iteration {it} DRSa {r:.4e} drop {r / r0:.3e} "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
    dt = time.perf_counter() - t0
    sv.close()
    print(json.dumps({"converge": {"iterations": it, "seconds": round(dt, 3), "drop": r / r0,
                                   "reached": r <= drop * r0,
                                   "Mcell_iter_per_s": round(cells * it / dt / 1e6, 3)}}))


def physical_cores():
    """Physical cores of the host (distinct (core, socket) pairs of lscpu), or None."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        return len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.CalledProcessError):
        return None


def roofline_of(stats, blocks, neq, fp32, workload):
    """Roofline of the dominant kernel (largest device time) from the HIP-event profile."""
    u = units(blocks)
    name, (ms, n) = max(stats.items(), key=lambda kv: kv[1][0])
    avg_s = ms / 1e3 / n
    unit, per = alg_bytes(name, neq, fp32)
    if per is None:
        return None
    alg = per * u[unit]
    ach = alg / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": measured_traffic(workload, name),
            "kernel": name, "avg_launch_us": round(avg_s * 1e6, 3), "alg_bytes_per_launch": alg,
            "alg_bytes_per_unit": per, "unit_of_work": unit, "units_per_launch": u[unit]}


def kernel_table(stats, blocks, neq, fp32, prof_steps):
    u = units(blocks)
    kern = {}
    for k, (kms, kn) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        un, pb = alg_bytes(k, neq, fp32)
        avg = kms / kn / 1e3
        kern[k] = {"us": round(avg * 1e6, 2), "per_step": round(kn / prof_steps, 3),
                   "GBs": round(pb * u[un] / avg / 1e9, 1) if pb else None}
    return kern


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--workload", default=os.environ.get("TRANS_BENCH_WORKLOAD", "compressor"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serving", type=int, default=16, metavar="N",
                    help="requests of the serving leg (0 skips it; 64 is the full serving workload)")
    ap.add_argument("--serving-target", type=float, default=0.5, metavar="DROP",
                    help="residual drop each served solve runs to")
    ap.add_argument("--serving-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-fp32", action="store_true", help="skip the second-precision leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain drop-in leg")
    ap.add_argument("--tune", type=int, default=3, metavar="ITERS",
                    help="autotuner iterations per launch configuration before timing (0: default configuration)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup+steps of the primary precision only and print the kernel table")
    ap.add_argument("--sweep", default=None, metavar="S[,STREAMS,GRAPH,TA,TB,CHUNK]",
                    help="with --profile-only: force a launch configuration (trans_hip_set_tuning)")
    ap.add_argument("--implicit-fp32", type=int, default=None, choices=(0, 1),
                    help="precision of the implicit operator of the primary leg (default: the workload's)")
    ap.add_argument("--converge", type=int, default=0,
                    help="instead of the timed run: iterate up to N times to a 1e-8 drop of the "
                         "global DRSa and print the history")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="compressor workload: grid scale (1 = 8.13M cells, 0.63 = 2.0M, 0.397 = 0.49M)")
    ap.add_argument("--back-pressure", type=float, default=1.2,
                    help="compressor workload: outlet static pressure / inlet total pressure")
    ap.add_argument("--deck", action="append", default=[], metavar="KEY=VALUE",
                    help="override an original.org value of the workload's deck (repeatable)")
    ap.add_argument("--wake-stats", type=int, default=0, metavar="N",
                    help="instead of the timed run: N iterations with mean/RMS statistics over the second "
                         "half, batched path against the plain drop-in path")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.serving_child:
        print(json.dumps(serving_leg(args.serving, target=args.serving_target)), flush=True)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU sharding of the "
                                   "blocks is a design only", "n_gpus": args.gpus}))
        sys.exit(2)

    # the compressor sector (8M cells) is quoted with the FP32 implicit operator and the FP64
    # residual; the flat plate with both in FP64
    primary32 = args.implicit_fp32 if args.implicit_fp32 is not None else int(args.workload == "compressor")
    blocks, prm, desc = build_workload(args.workload, primary32, args.scale, args.back_pressure, args.deck)
    cells = sum((b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks)
    if args.converge:
        converge(blocks, prm, desc, args.converge)
        return
    if args.wake_stats:
        wake_stats(blocks, prm, args.wake_stats, desc)
        return
    neq = 7 if (prm.if_euler == 0 and prm.if_turb_org == 1) else 5
    print(f"[bench] {desc}: {cells} cells, warmup {args.warmup}, steps {args.steps}", flush=True)
    prof_steps = min(args.steps, 20)
    force = None
    if args.sweep:
        v = [int(x) for x in args.sweep.split(",")]
        force = tuple(v + [4, 1, 16, 16, 8][len(v) - 1:])[:6]
    dt, res, stats = timed_run(blocks, prm, args.steps, args.warmup, prof_steps,
                               0 if args.profile_only else args.tune, force)
    tuning = timed_run.tuning
    tag = "FP32 implicit" if primary32 else "FP64"
    print(f"[bench] {tag}: {args.steps} steps in {dt:.3f} s, DRSa(block 1) = {res[0]['drsa']:.6e}", flush=True)
    if args.profile_only:
        for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0]):
            print(f"[bench] {k:28s} {v[0] / v[1] * 1e3:9.2f} us x {v[1]}")
        return
    value = cells * args.steps / dt / 1e6
    roof = roofline_of(stats, blocks, neq, bool(primary32), args.workload)
    kern = kernel_table(stats, blocks, neq, bool(primary32), prof_steps)
    other = None
    if not args.no_fp32:
        b2, p2, _ = build_workload(args.workload, 1 - primary32, args.scale, args.back_pressure, args.deck)
        dt2, res2, _ = timed_run(b2, p2, args.steps, args.warmup, 0, args.tune)
        other = {"implicit": "fp32" if not primary32 else "fp64",
                 "value": round(cells * args.steps / dt2 / 1e6, 3),
                 "ms_per_step": round(dt2 / args.steps * 1e3, 4),
                 "drsa_block1": res2[0]["drsa"], "drsa_block1_primary": res[0]["drsa"],
                 "tuning": (timed_run.tuning or {}).get("chosen")}
        print(f"[bench] {other['implicit']} implicit: {args.steps} steps in {dt2:.3f} s", flush=True)
        del b2
    plain = None
    if not args.no_plain:
        # the same workload as a plain drop-in: every routine launched once per block, the sweeps
        # one launch per hyperplane per block, one stream, no graph (trans_hip_set_tuning -3)
        ns = min(args.steps, 10)
        dtp, _, _ = timed_run(blocks, prm, ns, args.warmup, 0, 0, (-3, 1, 0, 16, 16, 8))
        vp = cells * ns / dtp / 1e6
        plain = {"value": round(vp, 3), "ms_per_step": round(dtp / ns * 1e3, 3), "steps": ns,
                 "speedup_of_value": round(value / vp, 2),
                 "what": "per-block launches of every routine, per-plane LU-SGS launches per block, "
                         "one stream, no HIP graph; same precision as value"}
        print(f"[bench] plain drop-in: {ns} steps in {dtp:.3f} s ({vp:.1f} Mcell*iter/s)", flush=True)
    cpu = None if args.no_cpu else cpu_baseline(blocks, prm)
    serve = None
    if args.serving > 0:
        # in a child process: it spreads the resident solves' streams over 16 hardware queues
        # (HIP's default is 4), a per-process setting that must precede HIP's initialisation
        import subprocess
        env = dict(os.environ, GPU_MAX_HW_QUEUES="16")
        cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--serving-child",
                             "--serving", str(args.serving), "--serving-target", str(args.serving_target)],
                            env=env, stdout=subprocess.PIPE, text=True)
        if cp.returncode != 0:
            raise SystemExit(f"serving leg failed ({cp.returncode})")
        serve = json.loads(cp.stdout.strip().splitlines()[-1])
        print(f"[bench] serving: {serve['requests']} solves in {serve['makespan_s']} s", flush=True)
    # whole-iteration figure: compulsory bytes of one fused iteration per cell (state in and
    # out once, geometry once, one round trip of the increment between the two sweeps)
    b_it = (ITER_BYTES_RANS - (2 * 4 * 7 if primary32 else 0)) if neq == 7 else ITER_BYTES_EULER
    if prm.if_rk == 1:
        b_it = 3 * ITER_BYTES_RK_STAGE  # three Runge-Kutta stages per iteration
    it_roof = {"alg_bytes_per_cell": b_it, "achieved": round(value * 1e6 * b_it / 1e9, 2),
               "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(value * 1e6 * b_it / 1e9 / HBM_PEAK_GBS, 5)}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated grid, uniform start)",
        "config": {"workload": desc, "cells": cells, "blocks": len(blocks),
                   "residual": "fp64", "implicit": "fp32" if primary32 else "fp64"},
        "roofline": roof, "iteration_roofline": it_roof, "cpu_baseline": cpu, "other_precision": other,
        "plain_dropin": plain, "kernels": kern, "tuning": tuning, "serving": serve,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
