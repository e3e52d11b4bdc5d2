"""The blunt-trailing-edge linear cascade (SST-DES, AUSMPW+ with MDCD, explicit three-stage
Runge-Kutta with one global time step, all FP64) on the device against the oracle, and its
time-averaged wake statistics on the batched path against the plain drop-in path.

Case: cases.cascade at 1/8 scale (12 blocks, 37 632 cells, 8 spanwise cells): pitchwise and
spanwise periods as translated interfaces (PerType 0), the flat leading- and trailing-edge
faces and the blade surfaces as walls, total-pressure inlet, static-pressure outlet; the global
step is 0.8 times the smallest local step of the start.  A seeded +-1 % perturbation of the
uniform start reaches the limiter, DES and entropy branches.

On the perturbed start a stage changes the interior density by 0.2 to 5 % of itself, so the
stage's increments are compared as well as the state.

Bars: gradients and residual D(1:7) from identical post-BC inputs within 1e-12; each
Runge-Kutta stage from identical inputs within 1e-12 on the state, p, T, mu_t, Wt2 and the
norms (argmax locations exact) and within 1e-10 on the increment of the state; whole
iterations within 1e-6; the mean and RMS statistics of the batched path within 1e-9 of the
plain path's (per-block launches, per-plane k-omega sweeps), whose k-omega sweeps round
their fused multiply-adds differently.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu

NORMS = ("drsa", "dua", "drsk", "duk", "drso", "duo")
ARGS = ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo")


def _cascade(scale=1 / 8, nk=8, seed=20261203):
    blocks, org, glob = cases.cascade(scale=scale, nk=nk)
    prm = solver.prepare_case(blocks, org, glob)
    prm.dt_lusgs_main = cases.global_time_step(blocks, prm, org.CFL_Sub_end)
    perturb(blocks, prm, seed=seed, turb=True)
    return blocks, prm


def _compare_state(sv, oc, blocks, tol, names):
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in names:
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < tol), (n, nm, e)


def test_cascade_single_residual():
    """Gradients (ghosts included, after BC(0)) and D(1:7) from identical post-BC(-1) states:
    the translated periodic interfaces, the flat leading/trailing-edge walls, DES blending."""
    print("case: cascade 1/8 scale, residual from identical post-BC inputs", flush=True)
    blocks, prm = _cascade()
    oc = OracleCase(blocks, prm)
    oc.spectral_radius(1)
    oc.bc(1, -1)
    for n, b in enumerate(blocks):
        for nm in ("v", "pp", "t", "rmiu"):
            b[nm][...] = oc[n][nm]
    sv = Solver(blocks, prm, ki=1)
    sv.spectral_radius()
    sv.space()
    oc.residual(1)
    for n, b in enumerate(blocks):
        dq = sv.get_field(n, capi.FIELD_DQDXYZ)
        e = np.array([np.max(np.abs(dq[c] - oc[n]["DQdxyz"][c])) / max(np.max(np.abs(oc[n]["DQdxyz"][c])), 1e-300)
                      for c in range(dq.shape[0])])
        assert np.all(e < 1e-12), (n, "DQdxyz", e)
        d = sv.get_field(n, capi.FIELD_D)
        e = relerr(interior(d, b), interior(oc[n]["D"], b))
        assert np.all(e < 1e-12), (n, "D", e)
    sv.close()


def test_cascade_rk_stages():
    print("case: cascade 1/8 scale, 12 blocks, RK3 stages 1-3 from identical inputs", flush=True)
    blocks, prm = _cascade()
    assert prm.kind_dual == 0 and prm.if_rk == 1 and prm.kind_hybrid == 3
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    for stage in (1, 2, 3):
        print("stage", stage, flush=True)
        if stage > 1:
            for n in range(len(blocks)):
                sv.set_state(n, oc[n])
        v0 = [oc[n]["v"].copy() for n in range(len(blocks))]
        oc.residual(1)
        sv.residual()
        ores = oc.rk_update(1, stage)
        gres = sv.rk_update(stage)
        _compare_state(sv, oc, blocks, 1e-12, ("v", "pp", "t", "rmiu", "wt2"))
        for n, b in enumerate(blocks):
            dg = interior(sv.get_state(n)["v"], b) - interior(v0[n], b)
            do = interior(oc[n]["v"], b) - interior(v0[n], b)
            e = relerr(dg, do)
            assert np.all(np.max(np.abs(do), axis=(1, 2, 3)) > 0.0)
            assert np.all(e < 1e-10), (n, "increment", e)
        for n in range(len(blocks)):
            for key in NORMS:
                assert abs(gres[n][key] - ores[n][key]) <= 1e-12 * abs(ores[n][key]), (n, key)
            for key in ARGS:
                assert gres[n][key] == ores[n][key], (n, key)
    sv.close()


def test_cascade_iterations():
    n_iter = 4
    print(f"case: cascade 1/8 scale, {n_iter} RK3 iterations through trans_hip_iterate", flush=True)
    blocks, prm = _cascade()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.iterate(1, n_iter)
    gres = sv.iterate(n_iter)
    ores = oc.resid()
    _compare_state(sv, oc, blocks, 1e-6, ("v",))
    for n in range(len(blocks)):
        for key in ("drsa", "drsk", "drso"):
            assert abs(gres[n][key] - ores[n][key]) <= 1e-6 * ores[n][key], (n, key)
    sv.close()


def test_cascade_wake_statistics_batched_vs_plain():
    n_iter, ki_aver = 24, 8
    print(f"case: cascade 1/8 scale, {n_iter} RK3 iterations, statistics from iteration {ki_aver}, "
          "batched vs plain drop-in", flush=True)
    stats = []
    blocks = None
    for plain in (False, True):
        blocks, prm = _cascade()
        sv = Solver(blocks, prm, ki=1)
        if plain:
            sv.set_tuning(-3, 1, 0, 16, 16, 8)
            assert sv.tuning()["sweep"] == -3
        sv.stats_enable(ki_aver, ki_aver)
        sv.iterate(n_iter, want_resid=False)
        stats.append([sv.get_stats(n) for n in range(len(blocks))])
        sv.close()
    for n, (a, b) in enumerate(zip(*stats)):
        a, b = interior(a, blocks[n]), interior(b, blocks[n])
        assert np.all(np.isfinite(a)) and np.any(a != 0.0)
        e = relerr(a, b)
        print(n, "stats relerr max", float(np.max(e)), flush=True)
        assert np.all(np.asarray(e) < 1e-9), (n, e)
