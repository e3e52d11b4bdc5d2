"""GPU tests of the serving entry points (trans_srv_*): continuous batching of steady solves.

Requests are bump-channel Euler solves (MUSCL + Roe, LU-SGS, CFL 5) of three grid sizes and
freestream Mach numbers, each to a residual drop of 0.1 (a few hundred iterations) or an
iteration cap.  Bars, all exact (bitwise):
- each solve's final state and iteration counts are the same alone and in a mix;
- a served solve equals the plain trans_hip_iterate path run for the same iteration count;
- a context's allocation equals the footprint the server admits it by.
"""
import numpy as np
import pytest

from trans_amd import capi, cases, serving, solver
from trans_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = [(33, 17, 5, 0.5), (41, 21, 5, 0.5), (25, 13, 5, 0.55)]
TARGET, MAXIT, CHUNK = 0.1, 600, 16
BIG = 1 << 36


def _bump(ni, nj, nk, mach):
    blocks, org, glob = cases.bump_channel(ni, nj, nk)
    org.XM = mach
    return blocks, solver.prepare_case(blocks, org, glob)


def _alone(case, chunk=CHUNK):
    srv = serving.Server(BIG, max_active=1, chunk=chunk)
    blocks, prm = _bump(*case)
    rid = srv.submit(blocks, prm, TARGET, MAXIT)
    srv.run()
    st, res = srv.poll(rid), srv.result(rid)
    srv.close()
    return st, res


def _same(res_a, res_b):
    for a, b in zip(res_a, res_b):
        for nm in a:
            assert np.array_equal(a[nm], b[nm]), nm


def test_same_alone_and_in_the_mix():
    print("case: three bump solves alone, then together (3 slots)", flush=True)
    alone = [_alone(c) for c in CASES]
    for st, _ in alone:
        print("alone", st["state_name"], st["iterations"], st["iterations_to_target"], flush=True)
        assert st["state"] == capi.SRV_CONVERGED
        assert st["iterations_to_target"] <= st["iterations"] < st["iterations_to_target"] + CHUNK
    srv = serving.Server(BIG, max_active=3, chunk=CHUNK)
    ids = [srv.submit(*_bump(*c), TARGET, MAXIT) for c in CASES]
    srv.run()
    sts = [srv.poll(i) for i in ids]
    # all three were resident together: each was admitted before any finished
    assert max(s["t_admit"] for s in sts) < min(s["t_done"] for s in sts)
    for (st0, res0), st, rid in zip(alone, sts, ids):
        assert st["state"] == capi.SRV_CONVERGED
        for key in ("iterations", "iterations_to_target", "drsa_first", "drsa_last"):
            assert st[key] == st0[key], key
        _same(srv.result(rid), res0)
    srv.close()


def test_served_solve_equals_plain_iterations():
    print("case: served bump 33x17x5 vs trans_hip_iterate for the same count", flush=True)
    st, res = _alone(CASES[0])
    blocks, prm = _bump(*CASES[0])
    sv = Solver(blocks, prm, ki=1)
    sv.iterate(st["iterations"], want_resid=False)
    _same([sv.get_state(0)], res)
    sv.close()


def test_admission_by_footprint():
    """A budget that holds the first two solves but not all three: the third waits for
    one of them to leave the device, and still gives its stand-alone result."""
    print("case: budget for two of three solves", flush=True)
    fps = [serving.footprint(*_bump(*c)) for c in CASES]
    budget = fps[0] + fps[1] + min(fps[2], fps[0], fps[1]) // 2
    assert budget < sum(fps)
    srv = serving.Server(budget, max_active=3, chunk=CHUNK)
    ids = [srv.submit(*_bump(*c), TARGET, MAXIT) for c in CASES]
    srv.progress(wait=False)
    assert [srv.poll(i)["state"] for i in ids] == [capi.SRV_RUNNING, capi.SRV_RUNNING, capi.SRV_QUEUED]
    srv.run()
    sts = [srv.poll(i) for i in ids]
    assert sts[2]["t_admit"] >= min(sts[0]["t_done"], sts[1]["t_done"])
    _same(srv.result(ids[2]), _alone(CASES[2])[1])
    srv.close()


def test_cancel_queued_and_running():
    print("case: cancel a queued and a running solve", flush=True)
    srv = serving.Server(BIG, max_active=1, chunk=CHUNK)
    a = srv.submit(*_bump(*CASES[0]), TARGET, MAXIT)
    b = srv.submit(*_bump(*CASES[1]), TARGET, MAXIT)
    srv.progress(wait=False)
    assert srv.poll(a)["state"] == capi.SRV_RUNNING and srv.poll(b)["state"] == capi.SRV_QUEUED
    srv.cancel(b)
    assert srv.poll(b)["state"] == capi.SRV_CANCELLED and srv.poll(b)["iterations"] == 0
    srv.progress(wait=True)
    srv.cancel(a)
    st = srv.poll(a)
    assert st["state"] == capi.SRV_CANCELLED and st["iterations"] % CHUNK == 0 and st["iterations"] > 0
    assert np.all(np.isfinite(srv.result(a)[0]["v"]))
    # the slot is free again
    c = srv.submit(*_bump(*CASES[2]), TARGET, MAXIT)
    srv.run()
    assert srv.poll(c)["state"] == capi.SRV_CONVERGED
    srv.close()


def test_footprint_equals_allocation():
    import ctypes as C
    for case in CASES:
        blocks, prm = _bump(*case)
        sv = Solver(blocks, prm)
        got = C.c_size_t()
        assert sv.lib.trans_hip_device_bytes(sv.ctx, C.byref(got)) == 0
        assert got.value == serving.footprint(blocks, prm)
        sv.close()


SECTORS = [(1 / 16, 0.9), (1 / 8, 1.3)]
S_TARGET, S_MAXIT = 0.5, 240


def _sector(scale, bp):
    blocks, org, glob = cases.compressor_sector(scale=scale, back_pressure=bp)
    return blocks, solver.prepare_case(blocks, org, glob, implicit_fp32=1)


def test_compressor_family_alone_and_in_the_mix():
    """Two compressor-sector solves of the serving workload's family (RANS-SST, FP32 implicit
    operator, 2K and 16K cells, outlet back pressure 0.9 and 1.3): final state and iteration
    counts bitwise the same alone and resident together."""
    print("case: compressor sectors 1/16 (bp 0.9) and 1/8 (bp 1.3), alone then together", flush=True)
    alone = []
    for sc in SECTORS:
        srv = serving.Server(BIG, max_active=1, chunk=CHUNK)
        rid = srv.submit(*_sector(*sc), S_TARGET, S_MAXIT)
        srv.run()
        alone.append((srv.poll(rid), srv.result(rid)))
        srv.close()
        print("alone", sc, alone[-1][0]["state_name"], alone[-1][0]["iterations"], flush=True)
    srv = serving.Server(BIG, max_active=2, chunk=CHUNK)
    ids = [srv.submit(*_sector(*sc), S_TARGET, S_MAXIT) for sc in SECTORS]
    srv.run()
    sts = [srv.poll(i) for i in ids]
    assert max(s["t_admit"] for s in sts) < min(s["t_done"] for s in sts)
    for (st0, res0), st, rid in zip(alone, sts, ids):
        assert st["state"] in (capi.SRV_CONVERGED, capi.SRV_MAX_ITER)
        for key in ("state", "iterations", "iterations_to_target", "drsa_first", "drsa_last"):
            assert st[key] == st0[key], key
        _same(srv.result(rid), res0)
    srv.close()
