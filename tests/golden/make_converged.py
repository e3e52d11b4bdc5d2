"""Writes converged.json: the oracle's converged quantities for the converged-solution tests.

Run from the repository root (CPU only; about three minutes on 8 cores):
    python tests/golden/make_converged.py

Each case iterates the oracle (oracle/trans_oracle*.c, FP64) from the uniform start, in the
order of main.F90:48-211, until the global DRSa (cell-count weighted RMS of the blocks'
DRSa) falls to `drop` times its first-iteration value or `max_iter` iterations have run,
checking every 50 iterations, and records the iteration count reached together with the
monitored quantities: inlet and outlet mass flow (Output/OUTPUT_Residual_NS.F90:37-75: the
planes i = 4 of the inlet blocks and i = NI - 4 of the outlet blocks, summed over blocks) and mass-averaged absolute-frame total pressure and total
temperature (the total-pressure formula of Output/Average_Field.F90:66 with the absolute
velocity), the total-pressure ratio, the total-temperature ratio and the isentropic
efficiency, and the last global DRSa.  The GPU tests iterate the device
path for exactly the recorded count and compare.
"""
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleCase, lib  # noqa: E402
from trans_amd import cases, solver  # noqa: E402

CHECK = 50


def build(name):
    """(blocks, params, inlet planes, outlet planes, drop, max_iter) of a named case."""
    if name == "sector_euler":
        # the compressor-rotor sector at 1/16 scale, inviscid, back pressure 0.9 x P0in
        blocks, org, glob = cases.compressor_sector(scale=1 / 16, back_pressure=0.9)
        org.IF_Euler = 1
        drop, max_iter = 1e-8, 20000
    elif name == "sector_rans":
        # the same sector with the SST deck at CFL 2: its residual levels off near 2e-7 of the
        # first iteration's (limiter and wall-function noise), so it runs a fixed 12000 iterations
        blocks, org, glob = cases.compressor_sector(scale=1 / 16, back_pressure=0.9)
        org.CFL_Sub_end = 2.0
        drop, max_iter = 1e-8, 12000
    elif name == "bump65":
        blocks, org, glob = cases.bump_channel(65, 33, 9)
        drop, max_iter = 1e-8, 12000
    elif name == "bump_2000":
        # the 129x65x9 bump channel for its specified 2000 iterations
        blocks, org, glob = cases.bump_channel()
        drop, max_iter = 0.0, 2000
    else:
        raise KeyError(name)
    prm = solver.prepare_case(blocks, org, glob)
    # the monitor planes of OUTPUT_Residual_NS: i = 4 of the inlet blocks, i = NI - 4 of the outlet
    # blocks (faces between interior cells, away from the boundary ghosts)
    inl = [(n, 1, 4) for n, b in enumerate(blocks) if any(r.bctype in (2, 5) and r.face == 1 for r in b.bcs)]
    out = [(n, 1, b.ni - 4) for n, b in enumerate(blocks) if any(r.bctype in (2, 7) and r.face == 2 for r in b.bcs)]
    return blocks, prm, inl, out, drop, max_iter


CASES = ("sector_euler", "sector_rans", "bump65", "bump_2000")


def global_drsa(res, blocks):
    n = [(b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks]
    return math.sqrt(sum(nb * r["drsa"] ** 2 for nb, r in zip(n, res)) / sum(n))


def quantities(integrals, inl, out, gamma=1.4):
    """Monitored quantities from a plane-integral function integrals(n, dir, index)."""
    def plane(planes):
        rec = [integrals(*p) for p in planes]
        m = sum(r["mass_flow"] for r in rec)
        p0 = sum(r["mass_flow"] * r["p0_abs_mass_avg"] for r in rec) / m
        t0 = sum(r["mass_flow"] * r["t0_abs_mass_avg"] for r in rec) / m
        return m, p0, t0
    mi, p0i, t0i = plane(inl)
    mo, p0o, t0o = plane(out)
    pi, tau = p0o / p0i, t0o / t0i
    # a passive duct changes T0 only by its numerical error: no efficiency to speak of
    eta = (pi ** ((gamma - 1.0) / gamma) - 1.0) / (tau - 1.0) if abs(tau - 1.0) > 1e-3 else None
    return {"mass_flow_in": mi, "mass_flow_out": mo, "p0_in": p0i, "p0_out": p0o, "t0_in": t0i, "t0_out": t0o,
            "pressure_ratio": pi, "temperature_ratio": tau, "efficiency": eta}


def run_oracle(name, threads=8):
    blocks, prm, inl, out, drop, max_iter = build(name)
    lib().ora_set_threads(threads)
    oc = OracleCase(blocks, prm)
    ki = oc.iterate(1, 1)
    r0 = r = global_drsa(oc.resid(), blocks)
    it = 1
    while it < max_iter and not (r <= drop * r0):
        n = min(CHECK - it % CHECK if it % CHECK else CHECK, max_iter - it)
        ki = oc.iterate(ki, n)
        it += n
        r = global_drsa(oc.resid(), blocks)
    q = quantities(oc.plane_integrals, inl, out)
    q.update(iterations=it, drop=r / r0, drsa0=r0, drsa=r)
    return q


def main():
    res = {}
    for name in CASES:
        t0 = time.time()
        res[name] = run_oracle(name)
        print(name, f"{time.time() - t0:.1f} s", res[name], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "converged.json"), "w") as fh:
        json.dump(res, fh, indent=1)


if __name__ == "__main__":
    main()
