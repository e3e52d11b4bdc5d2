"""Input checks of the C-ABI that need no device: the sizing call (trans_hip_footprint) runs
trans_hip_create's own checks without allocating.  Each case is a deck or block table the
library must reject with a return code (no CPU fallback, no partial context), next to the
same inputs corrected, which it must accept."""
import ctypes as C

import pytest

from trans_amd import capi, cases, mesh, serving, solver
from trans_amd.solver import TransError


def _bump():
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    return blocks, solver.prepare_case(blocks, org, glob)


def _rc(blocks, prm, descs=None, n=None):
    lib = capi.load_library()
    if descs is None:
        descs, keep = serving._descs(blocks)
    out = C.c_size_t()
    return lib.trans_hip_footprint(C.byref(prm), len(blocks) if n is None else n, descs, C.byref(out)), out.value


def test_accepts_a_valid_case():
    blocks, prm = _bump()
    rc, nbytes = _rc(blocks, prm)
    assert rc == 0 and nbytes > 0


@pytest.mark.parametrize("ni,nj,nk", [(3, 17, 5), (33, 3, 5), (33, 17, 2)])
def test_rejects_blocks_below_the_stencil(ni, nj, nk):
    """Fewer than 3 cells along i or j, or 2 along k, cannot hold the reference's stencils."""
    blocks, org, glob = cases.bump_channel(ni, nj, nk)
    with pytest.raises(TransError):
        serving.footprint(blocks, solver.prepare_case(blocks, org, glob))


def test_rejects_a_block_beyond_32_bit_indexing():
    """The kernels' schedule tables and per-block index arithmetic are 32-bit: a block of more
    than 2^28 padded cells is refused at creation (here 1100^3 points, about 1.4e9 cells)."""
    blocks, prm = _bump()
    descs, keep = serving._descs(blocks)
    descs[0].ni = descs[0].nj = descs[0].nk = 1100
    rc, _ = _rc(blocks, prm, descs)
    assert rc == capi.TRANS_EINVAL
    descs[0].ni, descs[0].nj, descs[0].nk = blocks[0].ni, blocks[0].nj, blocks[0].nk
    assert _rc(blocks, prm, descs)[0] == 0


def test_rejects_an_empty_block_table():
    blocks, prm = _bump()
    rc, _ = _rc(blocks, prm, n=0)
    assert rc == capi.TRANS_EINVAL


def test_rejects_interface_targets_out_of_range():
    blocks, org, glob = cases.flat_plate(nblk=2, ni_total=33, nj=17, nk=4, lo=9, hi=25)
    prm = solver.prepare_case(blocks, org, glob)
    assert _rc(blocks, prm)[0] == 0
    descs, keep = serving._descs(blocks)
    q = [n for n, c in enumerate(blocks[0].patches) if c["bctype"] == 9][0]
    for field, bad in (("target_block", 2), ("target_block", -1), ("target_conn", 99)):
        pt = C.cast(descs[0].patches, C.POINTER(capi.TransPatch))[q]
        good = getattr(pt, field)
        setattr(pt, field, bad)
        rc, _ = _rc(blocks, prm, descs)
        setattr(pt, field, good)
        assert rc == capi.TRANS_EINVAL, (field, bad)
    assert _rc(blocks, prm, descs)[0] == 0


@pytest.mark.parametrize("field,value", [("if_rk", 2), ("kind_spatial", 5), ("kind_roe", 0), ("cfl_sub_step", 0),
                                         ("kind_rotmod", 1), ("if_precondition", 2)])
def test_rejects_deck_values_outside_the_reference(field, value):
    blocks, prm = _bump()
    q = capi.TransParams.from_buffer_copy(prm)
    setattr(q, field, value)
    rc, _ = _rc(blocks, q)
    assert rc in (capi.TRANS_EINVAL, capi.TRANS_EUNSUPPORTED), (field, value, rc)
