"""GPU parity of the inviscid (Euler) path against the CPU oracle.

Every stage of one pseudo-time iteration of the bump-channel case (Roe + MUSCL/ALIMITER,
LU-SGS, far-field and symmetry BCs) is compared after the same seeded perturbation.
Bar: FP64 residual and increments within 1e-12 relative (max-norm per component).
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import capi
from trans_amd.mesh import sl
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr, small_bump

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(ni=33, nj=17, nk=5, seed=20261202):
    blocks, org, glob, prm = small_bump(ni, nj, nk)
    perturb(blocks, prm, seed=seed)
    return blocks, prm


def test_spectral_radius_and_bc():
    print("case: bump 33x17x5, spectral radius + BC(-1)", flush=True)
    blocks, prm = _case()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.spectral_radius(1)
    sv.spectral_radius()
    b = blocks[0]
    rds = sv.get_field(0, capi.FIELD_RDS)
    dtm = sv.get_field(0, capi.FIELD_DTM)
    assert np.all(relerr(interior(rds, b), interior(oc[0]["Rds"], b)) < TOL)
    assert relerr(interior(dtm, b), interior(oc[0]["Dtm"], b)) < TOL
    oc.bc(1, -1)
    sv.bc(-1)
    st = sv.get_state(0)
    # every ghost written by the oracle must match bit-for-bit-ish; compare whole arrays
    for nm, key in (("v", "v"), ("pp", "pp"), ("t", "t")):
        g, o = st[nm], oc[0][key]
        assert np.max(np.abs(g - o)) <= 1e-14 * max(np.max(np.abs(o)), 1.0), nm
    sv.close()


def test_single_residual():
    print("case: bump 33x17x5, single residual", flush=True)
    blocks, prm = _case()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    d = sv.get_field(0, capi.FIELD_D)
    b = blocks[0]
    err = relerr(interior(d, b, slice(0, 5)), interior(oc[0]["D"], b, slice(0, 5)))
    print("residual relerr", err)
    assert np.all(err < TOL)
    sv.close()


def test_lusgs_update_and_norms():
    print("case: bump 33x17x5, LU-SGS + update", flush=True)
    blocks, prm = _case()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    ores = oc.lusgs_update(1)[0]
    gres = sv.lusgs_update()[0]
    b = blocks[0]
    st = sv.get_state(0)
    for nm in ("v", "pp", "t"):
        e = relerr(interior(st[nm], b), interior(oc[0][nm], b))
        print(nm, e)
        assert np.all(np.asarray(e) < TOL), nm
    assert abs(gres["drsa"] - ores["drsa"]) <= TOL * ores["drsa"]
    assert abs(gres["dua"] - ores["dua"]) <= TOL * ores["dua"]
    assert (gres["ima"], gres["jma"], gres["kma"]) == (ores["ima"], ores["jma"], ores["kma"])
    sv.close()


def test_twenty_iterations_full_c1():
    """Twenty iterations of the full 129x65x9 bump channel, GPU vs oracle.

    The residual path is bit-identical per evaluation; across iterations the last-bit
    differences of exp() in the positivity update and pow() in Sutherland's law (device
    math library vs glibc) are amplified by the limiter during the transient, so the state
    is compared at the converged-quantity bar of 1e-6."""
    print("case: bump 129x65x9, 20 iterations", flush=True)
    from helpers import small_bump as sb
    blocks, org, glob, prm = sb(129, 65, 9)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    ki = oc.iterate(1, 20)
    gres = sv.iterate(20)[0]
    ores = oc.resid()[0]
    b = blocks[0]
    st = sv.get_state(0)
    e = relerr(interior(st["v"], b, slice(0, 5)), interior(oc[0]["v"], b, slice(0, 5)))
    print("state relerr after 20 iterations", e, "drsa", gres["drsa"], ores["drsa"])
    assert np.all(e < 1e-6)
    assert abs(gres["drsa"] - ores["drsa"]) <= 1e-6 * ores["drsa"]
    sv.close()
