"""The LDS-tiled face-flux + accumulation passes (k_resid_visc, k_resid_conv; TRANS_HIP_FUSED=1)
against the default path (face kernels of each direction through G and GV, then
k_accumulate_visc).

The tiled passes evaluate every face flux with the face kernels' own expressions and sum each
cell's faces in the accumulation kernel's order, so D must be bitwise equal; the other GPU
parity tests hold the default path to the oracle.  Cases: the rotating
compressor sector (MUSCL + Roe, SST, periodic and crossed interfaces, tile sizes smaller than
the workgroup's), the SST flat plate (MDCD + AUSMPW+), the Euler bump (MUSCL + Roe), each
perturbed; the k planes split into chunks of 3 so that chunk starts and ends are exercised.
"""
import numpy as np
import pytest

from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import perturb

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "sector":
        blocks, org, glob = cases.compressor_sector(scale=1 / 8, crossed=True)
        turb = True
    elif name == "plate":
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=97, nj=33, nk=9, lo=17, hi=49)
        turb = True
    else:
        blocks, org, glob = cases.bump_channel(65, 17, 9)
        turb = False
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=20261202, turb=turb)
    return blocks, prm


@pytest.mark.parametrize("name", ["sector", "plate", "bump"])
def test_fused_residual_bitwise(monkeypatch, name):
    print(f"case: {name}, residual tiled passes vs face kernels", flush=True)
    monkeypatch.setenv("TRANS_HIP_RESID_KCHUNK", "3")
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("TRANS_HIP_FUSED", fused)
        blocks, prm = _case(name)
        sv = Solver(blocks, prm, ki=1)
        sv.residual()
        out.append([sv.get_field(n, capi.FIELD_D) for n in range(len(blocks))])
        sv.close()
    for n, (a, b) in enumerate(zip(*out)):
        assert np.array_equal(a, b), (name, n, float(np.max(np.abs(a - b))))


def test_fused_iterations_bitwise(monkeypatch):
    """Five whole iterations of the sector with the default chunking: the states agree bitwise."""
    print("case: sector, 5 iterations tiled passes vs face kernels", flush=True)
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("TRANS_HIP_FUSED", fused)
        blocks, prm = _case("sector")
        sv = Solver(blocks, prm, ki=1)
        sv.iterate(5, want_resid=False)
        out.append([sv.get_state(n)["v"] for n in range(len(blocks))])
        sv.close()
    for n, (a, b) in enumerate(zip(*out)):
        assert np.array_equal(a, b), n
