"""Error paths of the C-ABI on the device: each call returns its code (the reference's NaN
STOP as TRANS_ENONFINITE, invalid arguments as TRANS_EINVAL) instead of computing on, and the
context stays usable afterwards.  No case makes a kernel fault: the NaN case computes on NaN
values, the others are rejected on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

from trans_amd import capi, cases, solver
from trans_amd.solver import Solver, TransError

pytestmark = pytest.mark.gpu


def _bump():
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    return blocks, solver.prepare_case(blocks, org, glob)


def test_nan_state_stops_like_the_reference_and_recovers():
    print("case: bump 33x17x5, one NaN density, then the clean state again", flush=True)
    blocks, prm = _bump()
    sv = Solver(blocks, prm, ki=1)
    clean = sv.get_state(0)
    bad = {k: v.copy() for k, v in clean.items()}
    bad["v"][0, 10, 8, 4] = np.nan
    sv.set_state(0, bad)
    with pytest.raises(TransError) as ei:
        sv.iterate(1)
    assert ei.value.rc == capi.TRANS_ENONFINITE
    assert "NaN" in str(ei.value)
    sv.set_state(0, clean)
    res = sv.iterate(2)
    assert np.isfinite(res[0]["drsa"]) and res[0]["drsa"] > 0.0
    sv.close()


def test_invalid_arguments_are_rejected_before_launch():
    print("case: bump 33x17x5, invalid arguments", flush=True)
    blocks, prm = _bump()
    sv = Solver(blocks, prm, ki=1)
    lib = sv.lib
    # a state view with the wrong component count
    st = sv.get_state(0)
    v5 = np.asfortranarray(st["v"][:5])
    view = capi.TransStateView()
    keep = {}
    for nm in ("v", "pp", "t", "rmiu", "wt0", "wt1", "wt2"):
        keep[nm] = v5 if nm == "v" else np.asfortranarray(st[nm])
        setattr(view, nm, capi.fview(keep[nm]))
    assert lib.trans_hip_set_state(sv.ctx, 0, C.byref(view)) == capi.TRANS_EINVAL
    # a block index out of range, a Runge-Kutta stage out of range, a monitor plane out of range
    assert lib.trans_hip_set_state(sv.ctx, 5, C.byref(view)) == capi.TRANS_EINVAL
    res = (capi.TransResid * 1)()
    assert lib.trans_hip_rk_update(sv.ctx, 4, res) == capi.TRANS_EINVAL
    out = capi.TransPlaneIntegrals()
    assert lib.trans_hip_plane_integrals(sv.ctx, 0, 1, 99, C.byref(out)) == capi.TRANS_EINVAL
    assert lib.trans_hip_plane_integrals(sv.ctx, 0, 4, 3, C.byref(out)) == capi.TRANS_EINVAL
    # parameters fixed at creation
    q = capi.TransParams.from_buffer_copy(prm)
    q.implicit_fp32 = 1
    assert lib.trans_hip_set_params(sv.ctx, C.byref(q)) == capi.TRANS_EINVAL
    # a tuning entry outside the documented ranges
    t = capi.TransTuneEntry(7, 1, 0, 0.0, 16, 16, 8)
    assert lib.trans_hip_set_tuning(sv.ctx, C.byref(t)) == capi.TRANS_EINVAL
    t = capi.TransTuneEntry(3, 1, 0, 0.0, 32, 16, 8)  # 512 lines per tile
    assert lib.trans_hip_set_tuning(sv.ctx, C.byref(t)) == capi.TRANS_EINVAL
    # the context still iterates and matches a fresh one bitwise
    sv.iterate(3, want_resid=False)
    ref = Solver(blocks, prm, ki=1)
    ref.iterate(3, want_resid=False)
    assert np.array_equal(sv.get_state(0)["v"], ref.get_state(0)["v"])
    ref.close()
    sv.close()
