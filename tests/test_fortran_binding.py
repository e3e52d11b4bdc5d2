"""The Fortran host binding (trans_amd/fortran/trans_hip_mod.f90).

CPU: the bind(C) derived types have the byte sizes of the C structs of include/trans_hip.h
(the module is compiled and a probe program prints c_sizeof of each type).
GPU: the Fortran demo driver runs a box case through the C-ABI and checks freestream
preservation and the damping of a density bump on its own (it prints "box_demo ok").
"""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

from trans_amd import build as tb
from trans_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "trans_amd", "fortran")


@pytest.mark.skipif(not os.path.exists(tb.FC) or not os.path.exists(tb.OUT),
                    reason="no Fortran compiler in this image, or the library is not built")
def test_bind_c_sizes_match_c_structs():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "abi_sizes")
        subprocess.run([tb.FC, "-module-dir", d, "-o", exe, os.path.join(FDIR, "trans_hip_mod.f90"),
                        os.path.join(FDIR, "abi_sizes.f90"), "-L" + tb.HERE, "-ltrans_hip",
                        "-Wl,-rpath," + tb.HERE], check=True, cwd=d)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    sizes = {ln.split()[0]: int(ln.split()[1]) for ln in out.strip().splitlines()}
    want = {"trans_params": capi.TransParams, "trans_patch": capi.TransPatch,
            "trans_block_desc": capi.TransBlockDesc, "trans_fview": capi.TransFView,
            "trans_geom_view": capi.TransGeomView, "trans_state_view": capi.TransStateView,
            "trans_resid": capi.TransResid, "trans_plane_integrals": capi.TransPlaneIntegrals,
            "trans_srv_config": capi.TransSrvConfig, "trans_srv_request": capi.TransSrvRequest,
            "trans_srv_status": capi.TransSrvStatus, "trans_geom_metrics_out": capi.TransGeomMetricsOut}
    for name, ty in want.items():
        assert sizes[name] == C.sizeof(ty), (name, sizes[name], C.sizeof(ty))


@pytest.mark.gpu
def test_fortran_host_runs_box_case():
    exe = tb.FORTRAN_DEMO
    if not os.path.exists(exe):
        pytest.skip("Fortran demo not built (no Fortran compiler at build time)")
    print("case: Fortran host, 33x9x5 box, freestream + density bump", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "box_demo ok" in r.stdout
