"""The C-ABI library loads and exports every function include/trans_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from trans_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "trans_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(trans_(?:hip|srv|geom|tecplot)_\w+)\s*\(", src)))


def test_header_matches_python_binding():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes_match_header():
    # the ctypes mirror must have the C layout: compare against a compiled probe
    import subprocess, tempfile
    probe = r'''
#include <stdio.h>
#include <stddef.h>
#include "trans_hip.h"
int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(trans_params),
 sizeof(trans_patch), sizeof(trans_block_desc), sizeof(trans_fview), sizeof(trans_geom_view),
 sizeof(trans_state_view), sizeof(trans_resid), sizeof(trans_plane_integrals), sizeof(trans_srv_config),
 sizeof(trans_srv_request), sizeof(trans_srv_status), sizeof(trans_geom_metrics_out)); return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(probe)
        exe = os.path.join(d, "p")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [ctypes.sizeof(t) for t in (capi.TransParams, capi.TransPatch, capi.TransBlockDesc,
                                       capi.TransFView, capi.TransGeomView, capi.TransStateView,
                                       capi.TransResid, capi.TransPlaneIntegrals, capi.TransSrvConfig,
                                       capi.TransSrvRequest, capi.TransSrvStatus, capi.TransGeomMetricsOut)]
    assert got == want


def test_footprint_without_a_device():
    """trans_hip_footprint runs trans_hip_create's sizing without touching the device: it
    works here, grows with the cell count, and adds the viscous face fluxes and the k-omega
    sweep inputs for a turbulent case."""
    from trans_amd import cases, serving, solver
    sizes = []
    for ni in (33, 65):
        blocks, org, glob = cases.bump_channel(ni, 17, 5)
        sizes.append(serving.footprint(blocks, solver.prepare_case(blocks, org, glob)))
    assert 0 < sizes[0] < sizes[1]
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    org.IF_Euler, org.IF_Turb_Org = 0, 1
    assert serving.footprint(blocks, solver.prepare_case(blocks, org, glob)) > sizes[0]


def test_deck_checks_without_a_device():
    """The library's deck checks run in the sizing call: RCCorrection mode 2 (which only
    writes f_r2, read by no routine) is accepted, mode 1 (unset index) is rejected, and
    IDDES / Suzen and low-Mach preconditioning are accepted for the SST model."""
    from trans_amd import cases, serving, solver
    from trans_amd.solver import TransError
    import pytest
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    org.IF_Euler, org.IF_Turb_Org = 0, 1
    prm = solver.prepare_case(blocks, org, glob)
    for field, value, ok in (("kind_rotmod", 2, True), ("kind_rotmod", 1, False), ("if_precondition", 1, True),
                             ("if_precondition", 2, False),
                             ("kind_subhy", 5, True), ("kind_subhy", 4, False), ("if_comppress", 1, True)):
        q = capi.TransParams.from_buffer_copy(prm)
        q.kind_hybrid = 3
        setattr(q, field, value)
        if ok:
            assert serving.footprint(blocks, q) > 0, field
        else:
            with pytest.raises(TransError):
                serving.footprint(blocks, q)
