"""Benchmark configurations at their full size on the device against the oracle.

- The turbulent flat plate of the FP64/FP32 benchmark: 8 blocks of uneven i-extent drawn from
  the seed, 257x129x5 points (131 072 cells), MDCD + AUSMPW+, SST, LU-SGS, from the uniform
  start: the first update within 1e-12 of the oracle (norms' argmax locations exact) and five
  iterations within 1e-6, with the FP64 operator; the FP32 operator's state within 1e-5 after
  the first update.
- The compressor-rotor sector of the headline: 16 blocks, 8 126 464 cells, 1200 rad/s,
  periodic interfaces, FP64 residual / FP32 implicit operator, one iteration from the uniform
  start against the oracle (OpenMP over blocks): the residual of that iteration within 1e-12,
  the state within 1e-6 of the FP64 oracle (the FP32 operator's increment rounding), the norms
  within 1e-5.

The flat plate starts uniform, so the first update is exact where the flow has not yet been
disturbed; the tests compare every cell anyway.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase, lib
from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import interior, relerr

pytestmark = pytest.mark.gpu


def test_flat_plate_full_size():
    print("case: flat plate 257x129x5 in 8 uneven blocks (131072 cells), FP64 operator", flush=True)
    blocks, org, glob = cases.flat_plate()
    prm = solver.prepare_case(blocks, org, glob)
    lib().ora_set_threads(8)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    ki = oc.iterate(1, 1)
    gres = sv.iterate(1)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < 1e-12), (n, nm, e)
        for key in ("drsa", "drsk", "drso"):
            assert abs(gres[n][key] - ores[n][key]) <= 1e-12 * ores[n][key], (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)
    oc.iterate(ki, 4)
    gres = sv.iterate(4)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_state(n)["v"], b), interior(oc[n]["v"], b))
        assert np.all(e < 1e-6), (n, e)
    sv.close()
    p32 = capi.TransParams.from_buffer_copy(prm)
    p32.implicit_fp32 = 1
    sv = Solver(blocks, p32, ki=1)
    oc = OracleCase(blocks, prm)
    oc.iterate(1, 1)
    sv.iterate(1)
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_state(n)["v"], b), interior(oc[n]["v"], b))
        assert np.all(e < 1e-5), (n, e)
    sv.close()


def test_compressor_sector_full_size_one_iteration():
    print("case: compressor sector 16 blocks, 8126464 cells, one iteration FP32 implicit vs FP64 oracle",
          flush=True)
    blocks, org, glob = cases.compressor_sector()
    prm = solver.prepare_case(blocks, org, glob, implicit_fp32=1, device_geometry=True)
    p64 = capi.TransParams.from_buffer_copy(prm)
    p64.implicit_fp32 = 0
    sv = Solver(blocks, prm, ki=1)
    sv.residual()
    d_dev = [sv.get_field(n, capi.FIELD_D) for n in range(len(blocks))]
    sv.close()
    sv = Solver(blocks, prm, ki=1)
    gres = sv.iterate(1)
    lib().ora_set_threads(16)
    oc = OracleCase(blocks, p64)
    oc.residual(1)
    for n, b in enumerate(blocks):
        e = relerr(interior(d_dev[n], b), interior(oc[n]["D"], b))
        assert np.all(e < 1e-12), (n, "D", e)
    del d_dev
    oc.lusgs_update(1)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_state(n)["v"], b), interior(oc[n]["v"], b))
        assert np.all(e < 1e-6), (n, e)
        for key in ("drsa", "drsk", "drso"):
            assert abs(gres[n][key] - ores[n][key]) <= 1e-5 * ores[n][key], (n, key)
    sv.close()
