"""GPU tests of the restart records and the time statistics.

Bars: the library's restart file is byte-identical to trans_amd/records.py's encoding of the
device state (which Fortran reads back, tests/test_records.py); a restart read restores V,
Rmiu, Wt0 and Wt1 bitwise and forms p and T as the oracle's restatement of Input/Init.F90
does, bitwise (no transcendental functions); the Average_Fields statistics match the oracle
within 1e-12 relative from identical inputs (the total pressure uses pow, whose last bit may
differ); the statistics file round-trips bitwise.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase, lib
from trans_amd import capi, cases, records, solver
from trans_amd.mesh import O, sl
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu


def _plate(seed=20261202):
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=seed, turb=True)
    return blocks, prm


def _inner(a, b):
    I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
    return a[..., I, J, K]


def test_restart_write_read(tmp_path):
    print("case: flat plate, 6 iterations, restart write, read into a fresh context", flush=True)
    blocks, prm = _plate()
    sv = Solver(blocks, prm, ki=1)
    sv.iterate(6, want_resid=False)
    kd = prm.kind_dual
    for n, b in enumerate(blocks):
        p = tmp_path / f"blk{n}.dat"
        sv.write_restart(n, 6, p)
        st = sv.get_state(n)
        want = tmp_path / f"want{n}.dat"
        records.write_restart(want, 6, _inner(st["v"], b), _inner(st["rmiu"], b), _inner(st["wt0"], b),
                              _inner(st["wt1"], b), kd)
        assert p.read_bytes() == want.read_bytes(), n
    # read into a context created from the initial state
    b2, p2 = _plate()
    sv2 = Solver(b2, p2, ki=1)
    for n, b in enumerate(blocks):
        assert sv2.read_restart(n, tmp_path / f"blk{n}.dat") == 6
        st, st2 = sv.get_state(n), sv2.get_state(n)
        for nm in ("v", "rmiu") + (("wt0",) if kd >= 1 else ()) + (("wt1",) if kd == 2 else ()):
            assert np.array_equal(_inner(st2[nm], b), _inner(st[nm], b)), (n, nm)
        # p and T as Input/Init.F90 forms them from V
        oc = OracleCase([b2[n]], p2)
        oc[0]["v"][...] = st2["v"]
        lib().ora_pp_t_from_v(__import__("ctypes").byref(p2), __import__("ctypes").byref(oc.ob[0]))
        assert np.array_equal(_inner(st2["pp"], b), _inner(oc[0]["pp"], b)), n
        assert np.array_equal(_inner(st2["t"], b), _inner(oc[0]["t"], b)), n
    # and the restarted solve continues like the original one (p, T formed differently, so
    # the last bits of the following iterations differ)
    sv2.set_iteration(7)
    r1 = sv.iterate(3)
    r2 = sv2.iterate(3)
    for n in range(len(blocks)):
        assert abs(r1[n]["drsa"] - r2[n]["drsa"]) <= 1e-9 * r1[n]["drsa"], n
    sv.close()
    sv2.close()


def test_average_fields_against_oracle(tmp_path):
    """Five Average_Fields calls (KI_aver = 3, KI_statis = 4) on states that change between
    calls, from identical inputs (state and gradients of the same residual evaluation)."""
    print("case: flat plate, Average_Fields at KI 3..7 on perturbed states", flush=True)
    import ctypes as C
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    sv.stats_enable(3, 4)
    S = [np.zeros((23,) + b.shape(), order="F") for b in blocks]
    rng = np.random.Generator(np.random.MT19937(99))
    for ki in range(3, 8):
        for n, b in enumerate(blocks):
            for nm in ("v", "pp"):
                oc[n][nm][...] *= 1 + 0.01 * rng.uniform(-1, 1, oc[n][nm].shape)
            sv.set_state(n, {nm: oc[n][nm] for nm in ("v", "pp", "t", "rmiu", "wt0", "wt1", "wt2")})
        oc.residual(1)
        sv.residual()
        sv.set_iteration(ki)
        sv.average_fields()
        for n in range(len(blocks)):
            lib().ora_average_fields(C.byref(prm), ki, 3, 4, C.byref(oc.ob[n]), S[n].ctypes.data_as(C.POINTER(C.c_double)))
    for n, b in enumerate(blocks):
        g = sv.get_stats(n)
        for q in range(23):
            e = relerr(_inner(g[q], b), _inner(S[n][q], b))
            assert e < 1e-12, (n, capi.STAT_NAMES[q], e)
        # the FLfld file round trip
        p = tmp_path / f"blk{n}.sta"
        sv.write_stats(n, b["xc"], b["yc"], b["zc"], p)
        r = records.read_stats(p, b.ni, b.nj, b.nk)
        assert np.array_equal(r["stats"], _inner(g, b))
        assert np.array_equal(r["xc"], _inner(b["xc"], b)) and np.array_equal(r["yc"], _inner(b["yc"], b))
    # read back into a fresh context
    sv2 = Solver(blocks, prm, ki=1)
    sv2.stats_enable(3, 4)
    for n, b in enumerate(blocks):
        sv2.read_stats(n, tmp_path / f"blk{n}.sta")
        assert np.array_equal(_inner(sv2.get_stats(n), b), _inner(sv.get_stats(n), b))
    sv.close()
    sv2.close()


def test_iterate_updates_statistics():
    """trans_hip_iterate runs Average_Fields itself from KI_aver on (after each outer
    iteration, before KI advances), also when the iteration replays as a HIP graph."""
    print("case: flat plate, 8 iterations with statistics from KI 4", flush=True)
    blocks, prm = _plate()
    sv = Solver(blocks, prm, ki=1)
    sv.stats_enable(4, 6)
    sv.iterate(8, want_resid=False)
    ref = Solver(blocks, prm, ki=1)
    ref.stats_enable(4, 6)
    for _ in range(8):
        ref.iterate(1, want_resid=False)  # its own update at KI >= 4
    for n, b in enumerate(blocks):
        g, r = sv.get_stats(n), ref.get_stats(n)
        assert np.array_equal(_inner(g, b), _inner(r, b))
        assert np.all(np.isfinite(_inner(g, b))) and np.any(_inner(g[16], b) > 0)
    sv.close()
    ref.close()
