"""Dual time with several sub-iterations (Kchld > 1) on the device against the oracle.

With Kchld = 3 the last sub-iteration's update carries the dual-time history shift
(Kind_Dual 1: Wt0 <- V; Kind_Dual 2: Wt1 <- Wt0, Wt0 <- V), the two before it read the
unshifted Wt0 and Wt1, and the BDF term 3V - 4Wt0 + Wt1 over 2 DT_LUSGS_Main enters every
sub-iteration's right-hand side; a physical step DT_LUSGS_Main = 0.05 makes it matter.  Case:
the perturbed 33x17x5 Euler bump and the three-block SST flat plate, four outer iterations
through trans_hip_iterate.  Bar: V, Wt0 and Wt1 within 1e-6 of the oracle, the bar of every
multi-iteration test (last-bit transcendental differences amplified through the limiters),
and within 1e-8 after the first outer iteration (its three sub-iterations amplify them to
about 2e-10 on the bump); a shift at the wrong sub-iteration moves Wt0 by the sub-iteration's
increment, about 1e-3 of the state.
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu


def _case(name, kind_dual):
    if name == "bump":
        blocks, org, glob = cases.bump_channel(33, 17, 5)
        turb = False
    else:
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
        turb = True
    org.Kind_Dual, org.Kchld, org.DT_LUSGS_Main = kind_dual, 3, 0.05
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=20261202, turb=turb)
    return blocks, prm


@pytest.mark.parametrize("name", ["bump", "plate"])
@pytest.mark.parametrize("kind_dual", [1, 2])
def test_dual_time_subiterations(name, kind_dual):
    print(f"case: {name}, Kind_Dual={kind_dual}, Kchld=3, DT_LUSGS_Main=0.05", flush=True)
    blocks, prm = _case(name, kind_dual)
    assert prm.kchld == 3 and prm.kind_dual == kind_dual
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    ki = 1
    for it, tol in ((1, 1e-8), (3, 1e-6)):
        ki = oc.iterate(ki, it)
        sv.iterate(it, want_resid=False)
        for n, b in enumerate(blocks):
            st = sv.get_state(n)
            for nm, key in (("v", "v"), ("wt0", "wt0"), ("wt1", "wt1")):
                e = relerr(interior(st[nm], b), interior(oc[n][key], b))
                assert np.all(np.asarray(e) < tol), (name, kind_dual, it, n, nm, e)
    # the history really moved: Wt0 differs from Wt1 after the shifts
    st = sv.get_state(0)
    assert not np.array_equal(interior(st["wt0"], blocks[0]), interior(st["wt1"], blocks[0]))
    sv.close()
