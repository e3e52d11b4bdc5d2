"""The blunt-trailing-edge cascade generator (CPU): block layout, boundary topology, and the
interface maps as geometry.

For every interface row, the donor face's points must be the receiver face's points moved by
one of the cascade's periods: nothing for an interface inside the passage, one pitch along y
for the pitchwise period, one span depth along z for the spanwise period (a translation, so
PerType 0 and no rotation of the donor's vectors).  The flat leading- and trailing-edge faces
and the blade surfaces must be walls, and the oracle must advance the reduced case through
two Runge-Kutta iterations with finite states.
"""
import numpy as np

from oracle.oracle import OracleCase
from trans_amd import cases, solver
from trans_amd.mesh import O


def _face_points(b, face, a0, a1, b0, b1):
    d = (face - 1) // 2
    side = (face - 1) % 2
    idx = [slice(None)] * 3
    idx[d] = O + (1 if side == 0 else [b.ni, b.nj, b.nk][d])
    t = [x for x in range(3) if x != d]
    idx[t[0]] = slice(O + a0, O + a1 + 2)
    idx[t[1]] = slice(O + b0, O + b1 + 2)
    pts = np.stack([b.xx[tuple(idx)].ravel(), b.yy[tuple(idx)].ravel(), b.zz[tuple(idx)].ravel()], axis=1)
    return pts[np.lexsort(pts.T[::-1])]


def test_full_size_layout():
    blocks, org, glob = cases.cascade()
    cells = [(b.ni - 1) * (b.nj - 1) * (b.nk - 1) for b in blocks]
    assert len(blocks) == 12 and sum(cells) == 19267584
    assert len(set(cells)) > 4  # uneven sizes
    assert org.IF_RK == 1 and org.Kind_Dual == 0 and org.Kind_Hybrid == 3
    assert org.Kind_Spatial == 3 and org.Kind_Roe == 3


def test_interfaces_are_the_cascade_periods():
    blocks, org, glob = cases.cascade(scale=1 / 8, nk=8)
    from trans_amd.params import make_params
    c = 1.0e6 / make_params(org, glob).ref
    pitch, depth = 0.8 * c, 0.2 * c
    shifts = {"none": 0, "pitch": 0, "span": 0}
    for n, b in enumerate(blocks):
        for r in b.bcs:
            if r.bctype != 9:
                continue
            assert r.pertype == 0 and r.lcross == 0
            dn = blocks[r.target_blk - 1]
            rd = dn.bcs[r.target_conn - 1]
            assert rd.bctype == 9 and rd.target_blk == n + 1
            p = _face_points(b, r.face, r.a0, r.a1, r.b0, r.b1)
            q = _face_points(dn, rd.face, rd.a0, rd.a1, rd.b0, rd.b1)
            assert p.shape == q.shape
            found = None
            for name, t in (("none", (0, 0, 0)), ("pitch", (0, pitch, 0)), ("pitch", (0, -pitch, 0)),
                            ("span", (0, 0, depth)), ("span", (0, 0, -depth))):
                moved = p + np.array(t)
                moved = moved[np.lexsort(moved.T[::-1])]
                if np.allclose(moved, q, rtol=0.0, atol=1e-9 * c):
                    found = name
                    break
            assert found is not None, (n, r)
            shifts[found] += 1
    assert all(v > 0 for v in shifts.values()), shifts


def test_edges_and_surfaces_are_walls():
    blocks, org, glob = cases.cascade(scale=1 / 8, nk=8)
    walls = [(n, r.face) for n, b in enumerate(blocks) for r in b.bcs if r.bctype == 1]
    # leading-edge face (upstream band imax), trailing-edge face (first wake band imin),
    # suction surface (lower passage blocks jmin), pressure surface (upper passage blocks jmax)
    assert (0, 2) in walls and (6, 1) in walls
    assert sum(1 for _, f in walls if f == 3) == 2 and sum(1 for _, f in walls if f == 4) == 2
    assert len(walls) == 6
    assert [n for n, b in enumerate(blocks) for r in b.bcs if r.bctype == 5] == [0, 1]
    assert [n for n, b in enumerate(blocks) for r in b.bcs if r.bctype == 7] == [10, 11]


def test_oracle_advances_reduced_cascade():
    blocks, org, glob = cases.cascade(scale=1 / 8, nk=4)
    prm = solver.prepare_case(blocks, org, glob)
    prm.dt_lusgs_main = cases.global_time_step(blocks, prm, org.CFL_Sub_end)
    assert prm.dt_lusgs_main > 0.0
    oc = OracleCase(blocks, prm)
    assert oc.iterate(1, 2) == 3
    for n in range(len(blocks)):
        assert np.all(np.isfinite(oc[n]["v"]))
    assert any(r["drsa"] > 0.0 for r in oc.resid())
