"""GPU parity of the plain drop-in schedule (trans_hip_set_tuning sweep -3).

The plain schedule is the denominator of the batched path's speedup: every routine is one
launch per block (spectral radius, BC fills, gradients, the face kernels of each direction,
accumulation, the LU-SGS right-hand sides, one launch per hyperplane per block for each
sweep, the update), on one stream, without graph replay.  It must compute what the batched
schedule computes.  Bars: one iteration (BC, residual, LU-SGS, update) within 1e-12 of the
batched schedule and, on the sector, of the oracle with the norms' argmax locations exact;
the per-plane sweep entry (one LU-SGS + update from the oracle's post-BC residual) within
1e-12 of the oracle; ten iterations within 1e-6 of the oracle and of the batched schedule
(the per-plane sweep places its contracted multiply-adds differently from the batched sweep,
and over ten iterations the omega limiters amplify that last-bit difference to ~1e-7); the FP32 operator within 1e-5 of the FP64 increment.  Decks: the rotating
compressor sector at 1/16 scale (Roe + MUSCL) and the SST flat plate (AUSMPW+ + MDCD), both
perturbed by the seed whose oracle comparisons hold in the other GPU parity tests; an RK3
deck of the flat plate on the plain schedule against the oracle and the batched schedule.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _post_bc(blocks, prm):
    oc = OracleCase(blocks, prm)
    oc.spectral_radius(1)
    oc.bc(1, -1)
    for n, b in enumerate(blocks):
        for nm in ("v", "pp", "t", "rmiu"):
            b[nm][...] = oc[n][nm]
    return blocks, prm


def _case(name, seed=20261202, post_bc=True):
    if name == "sector":
        blocks, org, glob = cases.compressor_sector(scale=1 / 16, crossed=True)
    else:
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=seed, turb=True)
    return _post_bc(blocks, prm) if post_bc else (blocks, prm)


def _check(blocks, sv, oc, gres, ores, tol=TOL):
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < tol), (n, nm, e)
        for key in ("drsa", "dua", "drsk", "duk", "drso", "duo"):
            assert abs(gres[n][key] - ores[n][key]) <= tol * abs(ores[n][key]), (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)


class _Batched:
    """Device state of a solver, indexable like the oracle case."""
    def __init__(self, sv, n):
        self.st = [sv.get_state(q) for q in range(n)]

    def __getitem__(self, q):
        return self.st[q]


@pytest.mark.parametrize("name", ["sector", "plate"])
def test_plain_one_iteration(name):
    print(f"case: {name}, plain drop-in schedule, one iteration vs batched and oracle", flush=True)
    blocks, prm = _case(name, post_bc=False)
    pl = Solver(blocks, prm, ki=1)
    pl.set_tuning(-3, 1, 0)
    t = pl.tuning()
    assert (t["sweep"], t["streams"], t["graph"]) == (-3, 1, 0)
    bt = Solver(blocks, prm, ki=1)
    pres = pl.iterate(1)
    bres = bt.iterate(1)
    _check(blocks, pl, _Batched(bt, len(blocks)), pres, bres)
    if name == "sector":
        oc = OracleCase(blocks, prm)
        oc.iterate(1, 1)
        _check(blocks, pl, oc, pres, oc.resid())
    pl.close()
    bt.close()


@pytest.mark.parametrize("name", ["sector", "plate"])
def test_plain_plane_sweep_update(name):
    print(f"case: {name}, plain drop-in schedule, one LU-SGS update", flush=True)
    blocks, prm = _case(name)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    sv.set_tuning(-3, 1, 0)
    oc.residual(1)
    sv.spectral_radius()
    sv.space()
    ores = oc.lusgs_update(1)
    gres = sv.lusgs_update()
    _check(blocks, sv, oc, gres, ores)
    sv.close()


@pytest.mark.parametrize("name", ["sector", "plate"])
def test_plain_ten_iterations(name):
    print(f"case: {name}, plain drop-in schedule, 10 iterations vs oracle and batched", flush=True)
    blocks, prm = _case(name, post_bc=False)
    oc = OracleCase(blocks, prm)
    pl = Solver(blocks, prm, ki=1)
    pl.set_tuning(-3, 1, 0)
    bt = Solver(blocks, prm, ki=1)
    oc.iterate(1, 10)
    pres = pl.iterate(10)
    bres = bt.iterate(10)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        e = relerr(interior(pl.get_state(n)["v"], b), interior(oc[n]["v"], b))
        assert np.all(e < 1e-6), (n, e)
        e = relerr(interior(pl.get_state(n)["v"], b), interior(bt.get_state(n)["v"], b))
        assert np.all(e < 1e-6), (n, "plain vs batched", e)
        for key in ("drsa", "drsk", "drso"):
            assert abs(pres[n][key] - ores[n][key]) <= 1e-6 * ores[n][key], (n, key)
            assert abs(pres[n][key] - bres[n][key]) <= 1e-6 * bres[n][key], (n, key)
    pl.close()
    bt.close()


def test_plain_fp32_operator():
    print("case: sector, plain drop-in schedule, FP32 operator", flush=True)
    blocks, prm = _case("sector")
    oc = OracleCase(blocks, prm)
    v0 = [b["v"].copy() for b in blocks]
    p32 = capi.TransParams.from_buffer_copy(prm)
    p32.implicit_fp32 = 1
    sv = Solver(blocks, p32, ki=1)
    sv.set_tuning(-3, 1, 0)
    oc.residual(1)
    sv.spectral_radius()
    sv.space()
    oc.lusgs_update(1)
    sv.lusgs_update()
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_state(n)["v"], b) - interior(v0[n], b),
                   interior(oc[n]["v"], b) - interior(v0[n], b))
        assert np.all(e < 1e-5), (n, e)
    sv.close()


def test_plain_rk_iteration():
    """An RK3 deck (SST flat plate, CFL 0.8) on the plain schedule: its stages are per-block
    launches and its k-omega sweeps per-plane launches per block.  One whole iteration (three
    stages without re-synchronising the state) within 1e-6 of the oracle and of the batched
    schedule, the whole-iteration bar of the RK tests: the k limiter amplifies the sweeps'
    last-bit differences to about 2e-9 in k over the three stages; the mean flow stays within
    1e-12."""
    print("case: flat plate 2 blocks, RK3, plain schedule, one iteration", flush=True)
    blocks, org, glob = cases.flat_plate(nblk=2, ni_total=33, nj=17, nk=4, lo=9, hi=25)
    org.IF_RK = 1
    org.CFL_Sub_start = org.CFL_Sub_end = 0.8
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=20261202, turb=True)
    oc = OracleCase(blocks, prm)
    oc.iterate(1, 1)
    ores = oc.resid()
    states = []
    for plain in (True, False):
        sv = Solver(blocks, prm, ki=1)
        if plain:
            sv.set_tuning(-3, 1, 0)
            assert sv.tuning()["sweep"] == -3
        gres = sv.iterate(1)
        states.append([sv.get_state(n)["v"] for n in range(len(blocks))])
        for n, b in enumerate(blocks):
            e = relerr(interior(states[-1][n], b), interior(oc[n]["v"], b))
            assert np.all(e[:5] < TOL) and np.all(e < 1e-6), (plain, n, e)
            for key in ("drsa", "drsk", "drso"):
                assert abs(gres[n][key] - ores[n][key]) <= 1e-6 * ores[n][key], (plain, n, key)
        sv.close()
    for n, b in enumerate(blocks):
        assert np.all(relerr(interior(states[0][n], b), interior(states[1][n], b)) < 1e-6)
