"""GPU tests of the launch autotuner: every configuration it tries (sweep variant, helper
streams, HIP graph) computes the same iteration (one update from the same state within 1e-12
relative: the sweep variants contract multiply-adds into FMAs in different places, so they may
differ in the last bits; six iterations within 1e-6, the bar of the other multi-iteration
tests), it restores the state and KI bitwise, and the configuration it keeps is its fastest."""
import numpy as np
import pytest

from trans_amd import cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "plate":
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    else:
        blocks, org, glob = cases.bump_channel(33, 17, 5)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=5, turb=name == "plate")
    return blocks, prm


@pytest.mark.parametrize("name", ["plate", "bump"])
def test_autotune_keeps_results(name):
    print(f"case: {name}, autotune then 6 iterations vs the default configuration", flush=True)
    blocks, prm = _case(name)
    a = Solver(blocks, prm, ki=1)
    a.iterate(2, want_resid=False)
    before = [a.get_state(n)["v"].copy() for n in range(len(blocks))]
    chosen, table = a.autotune(iters=3)
    print("chosen", chosen, "of", len(table), flush=True)
    for row in table:
        print(row, flush=True)
    # the tried set covers both sweep kernels, both stream modes and both graph modes
    assert {r["sweep"] >= 0 for r in table} == {True, False}
    assert {r["streams"] for r in table} == {1, 4} and {r["graph"] for r in table} == {0, 1}
    best = min(table, key=lambda r: r["ms_per_iter"])
    assert (chosen["sweep"], chosen["streams"], chosen["graph"]) == (best["sweep"], best["streams"], best["graph"])
    # the state and KI are restored
    for n in range(len(blocks)):
        assert np.array_equal(a.get_state(n)["v"], before[n])
    b = Solver(blocks, prm, ki=1)
    b.iterate(2, want_resid=False)
    b1 = b.iterate(1)
    v1 = [b.get_state(n)["v"].copy() for n in range(len(blocks))]
    rb = b.iterate(5)
    ra = a.iterate(6)
    for n, blk in enumerate(blocks):
        assert relerr(interior(a.get_state(n)["v"], blk), interior(b.get_state(n)["v"], blk)).max() < 1e-6, n
        assert abs(ra[n]["drsa"] - rb[n]["drsa"]) <= 1e-6 * rb[n]["drsa"]
    # every configuration it tried, applied to a fresh context, gives the same iterations
    for row in table:
        c = Solver(blocks, prm, ki=1)
        c.iterate(2, want_resid=False)
        c.set_tuning(row["sweep"], row["streams"], row["graph"], row["tile_a"], row["tile_b"], row["chunk"])
        c1 = c.iterate(1)
        for n, blk in enumerate(blocks):
            e = relerr(interior(c.get_state(n)["v"], blk), interior(v1[n], blk))
            assert np.all(e < 1e-12), (row, n, e)
            assert abs(c1[n]["drsa"] - b1[n]["drsa"]) <= 1e-12 * b1[n]["drsa"], row
        rc = c.iterate(5)
        for n, blk in enumerate(blocks):
            e = relerr(interior(c.get_state(n)["v"], blk), interior(b.get_state(n)["v"], blk))
            assert np.all(e < 1e-6), (row, n, e)
            assert abs(rc[n]["drsa"] - rb[n]["drsa"]) <= 1e-6 * rb[n]["drsa"], row
        c.close()
    a.close()
    b.close()


def test_autotune_with_preconditioning():
    """With IF_PRECONDITION = 1 only the hyperplane sweep carries the transforms: the tuner
    tries it alone (streams and graphs), keeps it, the tuned context iterates as an untuned
    one, and asking for another sweep fails."""
    print("case: flat plate M=0.1, IF_Precondition=1, autotune", flush=True)
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25, mach=0.1)
    org.IF_Precondition = 1
    org.CFL_Sub_start = org.CFL_Sub_end = 0.3
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=5, turb=True)
    a = Solver(blocks, prm, ki=1)
    assert a.tuning()["sweep"] == -1
    chosen, table = a.autotune(iters=1)
    assert {r["sweep"] for r in table} == {-1} and chosen["sweep"] == -1
    assert {r["streams"] for r in table} == {1, 4} and {r["graph"] for r in table} == {0, 1}
    with pytest.raises(Exception):
        a.set_tuning(1, 4, 0)
    b = Solver(blocks, prm, ki=1)
    ra, rb = a.iterate(1), b.iterate(1)
    for n, blk in enumerate(blocks):
        e = relerr(interior(a.get_state(n)["v"], blk), interior(b.get_state(n)["v"], blk))
        assert np.all(e < 1e-12), (n, e)
        assert abs(ra[n]["drsa"] - rb[n]["drsa"]) <= 1e-12 * rb[n]["drsa"]
    a.close()
    b.close()
