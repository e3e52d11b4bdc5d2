"""HIP-graph replay of the outer iteration against eager launches.

One iteration is captured once its kernel arguments no longer change (CFL ramp done,
turbulence switched on) and replayed after that; TRANS_HIP_GRAPH=0 at context creation
keeps every iteration eager.  The two contexts run the same kernels in the same order, so
their states must agree bit for bit after every call, with downloads and other host work
between the replays.
"""
import numpy as np
import pytest

from trans_amd import cases, solver
from trans_amd.solver import Solver

from helpers import perturb

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == "plate":
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    elif kind == "plate_rk":
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
        org.IF_RK, org.CFL_Sub_start, org.CFL_Sub_end = 1, 0.8, 0.8
    else:
        blocks, org, glob = cases.bump_channel(33, 17, 5)
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=3, turb=kind != "bump")
    return blocks, prm


@pytest.mark.parametrize("kind", ["plate", "plate_rk", "bump"])
def test_graph_replay_matches_eager(monkeypatch, kind):
    print(f"case: {kind}, graph replay vs eager", flush=True)
    blocks, prm = _case(kind)
    monkeypatch.setenv("TRANS_HIP_GRAPH", "0")
    eager = Solver(blocks, prm, ki=1)
    monkeypatch.setenv("TRANS_HIP_GRAPH", "1")
    graph = Solver(blocks, prm, ki=1)
    for chunk in (1, 2, 3, 1, 4):
        re = eager.iterate(chunk)
        rg = graph.iterate(chunk)
        assert re == rg, chunk
        for n in range(len(blocks)):
            se, sg = eager.get_state(n), graph.get_state(n)
            for nm in ("v", "pp", "t", "rmiu", "wt0", "wt1"):
                assert np.array_equal(se[nm], sg[nm]), (chunk, n, nm)
    eager.close()
    graph.close()
