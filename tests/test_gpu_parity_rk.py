"""GPU parity of the Runge-Kutta update (IF_RK = 1) and of the hyperplane LU-SGS sweep
against the CPU oracle.

Cases: the three-block turbulent flat plate (MDCD + AUSMPW+, k-omega SST, seeded +-1 %
perturbation) and the inviscid bump channel (MUSCL + Roe), both at CFL 0.8 so that the
explicit stages stay stable.  With k-omega a stage runs the mean-flow stage, the k-omega
LU-SGS on the new state and the k/omega update; without it the stage kernel does the whole
update.

Bars: one stage from identical inputs within 1e-12 relative (state, p, T, mu_t, Wt2 and the
three norms, argmax locations exact), as for the LU-SGS update; whole iterations within
1e-6 (the transcendental last-bit differences of the residual amplified through limiters, as
for the LU-SGS path); FP32 k-omega sweeps within 1e-5 of the k and omega increments.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12
NORMS = ("drsa", "dua", "drsk", "duk", "drso", "duo")
ARGS = ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo")


def _rk(org):
    org.IF_RK = 1
    org.CFL_Sub_start = org.CFL_Sub_end = 0.8


def _plate(seed=20261202):
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    _rk(org)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=seed, turb=True)
    return blocks, prm


def _bump(seed=20261202):
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    _rk(org)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=seed)
    return blocks, prm


def _compare_state(sv, oc, blocks, tol, names=("v", "pp", "t", "rmiu")):
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in names:
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            print(n, nm, e, flush=True)
            assert np.all(np.asarray(e) < tol), (n, nm, e)


def _compare_norms(gres, ores, tol):
    for n in range(len(ores)):
        for key in NORMS:
            assert abs(gres[n][key] - ores[n][key]) <= tol * abs(ores[n][key]), (n, key)
        for key in ARGS:
            assert gres[n][key] == ores[n][key], (n, key)


def test_rk_stages_turbulent_plate():
    """Each of the three stages from identical inputs (the device takes the oracle's state,
    Wt2 included, before stages 2 and 3): 1e-12.  Without the re-sync the last-bit
    differences of one stage reach ~5e-12 in k after the next one (k-omega sweep and
    limiter), which the iteration test below covers at its own bar."""
    print("case: flat plate 3 blocks, RK3 stages 1-3 with k-omega", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    for stage in (1, 2, 3):
        print("stage", stage, flush=True)
        if stage > 1:
            for n in range(len(blocks)):
                sv.set_state(n, oc[n])
        oc.residual(1)
        sv.residual()
        ores = oc.rk_update(1, stage)
        gres = sv.rk_update(stage)
        _compare_state(sv, oc, blocks, TOL, ("v", "pp", "t", "rmiu", "wt2"))
        _compare_norms(gres, ores, TOL)
    sv.close()


def test_rk_stage_euler_bump():
    print("case: bump 33x17x5 Euler, RK3 stage 1 (single-kernel stage)", flush=True)
    blocks, prm = _bump()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    ores = oc.rk_update(1, 1)
    gres = sv.rk_update(1)
    _compare_state(sv, oc, blocks, TOL, ("v", "pp", "t", "wt2"))
    _compare_norms(gres, ores, TOL)
    sv.close()


@pytest.mark.parametrize("case", ["plate", "bump"])
def test_rk_iterations(case):
    n_iter = 5
    print(f"case: {case}, {n_iter} RK3 iterations through trans_hip_iterate", flush=True)
    blocks, prm = _plate() if case == "plate" else _bump()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    assert oc.iterate(1, n_iter) == 1 + n_iter
    gres = sv.iterate(n_iter)
    ores = oc.resid()
    _compare_state(sv, oc, blocks, 1e-6, ("v",))
    for n in range(len(blocks)):
        for key in ("drsa", "drsk", "drso"):
            assert abs(gres[n][key] - ores[n][key]) <= 1e-6 * ores[n][key], (n, key)
    sv.close()


def test_rk_fp32_kw_sweeps():
    """implicit_fp32 with RK: the mean-flow stage stays FP64 (1e-12); the FP32 k-omega sweeps
    give k and omega increments within 1e-5 of the FP64 oracle's."""
    print("case: flat plate, RK3 stage 1 with FP32 k-omega sweeps", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    v0 = [b["v"].copy() for b in blocks]
    p32 = capi.TransParams.from_buffer_copy(prm)
    p32.implicit_fp32 = 1
    sv = Solver(blocks, p32, ki=1)
    oc.residual(1)
    sv.residual()
    oc.rk_update(1, 1)
    sv.rk_update(1)
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        e = relerr(interior(st["v"], b, slice(0, 5)), interior(oc[n]["v"], b, slice(0, 5)))
        assert np.all(e < TOL), (n, e)
        dv_o = interior(oc[n]["v"], b, slice(5, 7)) - interior(v0[n], b, slice(5, 7))
        dv_g = interior(st["v"], b, slice(5, 7)) - interior(v0[n], b, slice(5, 7))
        e = relerr(dv_g, dv_o)
        print(n, "FP32 k-omega increment relerr", e, flush=True)
        assert np.all(e < 1e-5)
    sv.close()


@pytest.mark.parametrize("rk", [0, 1])
def test_hyperplane_sweep(monkeypatch, rk):
    """TRANS_HIP_LINE_SWEEP=0 selects the hyperplane wavefront (the path of blocks whose
    cross-sections exceed one workgroup): one LU-SGS update, or one RK stage with its k-omega
    sweeps, from identical inputs at 1e-12."""
    print(f"case: flat plate, hyperplane sweeps, IF_RK={rk}", flush=True)
    monkeypatch.setenv("TRANS_HIP_LINE_SWEEP", "0")
    blocks, prm = _plate()
    prm.if_rk = rk
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    if rk:
        ores, gres = oc.rk_update(1, 1), sv.rk_update(1)
    else:
        ores, gres = oc.lusgs_update(1), sv.lusgs_update()
    _compare_state(sv, oc, blocks, TOL)
    _compare_norms(gres, ores, TOL)
    sv.close()


@pytest.mark.parametrize("rk", [0, 1])
def test_plane_launch_sweep(rk):
    """The sweep as one launch per hyperplane (the choice for blocks whose cross sections
    exceed a line-march workgroup), selected through trans_hip_set_tuning: one LU-SGS update,
    or one RK stage with its k-omega sweeps, from identical inputs at 1e-12."""
    print(f"case: flat plate, per-plane sweep launches, IF_RK={rk}", flush=True)
    blocks, prm = _plate()
    prm.if_rk = rk
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    sv.set_tuning(-2, 4, 0)
    assert sv.tuning()["sweep"] == -2
    oc.residual(1)
    sv.residual()
    if rk:
        ores, gres = oc.rk_update(1, 1), sv.rk_update(1)
    else:
        ores, gres = oc.lusgs_update(1), sv.lusgs_update()
    _compare_state(sv, oc, blocks, TOL)
    _compare_norms(gres, ores, TOL)
    sv.close()
