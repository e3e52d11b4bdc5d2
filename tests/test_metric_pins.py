"""Pins of the host geometry metrics (trans_amd/mesh.py, Geom/CalcDst.F90:160-632) against
the mathematics of the grid; the device metrics equal these bitwise (tests/test_gpu_geometry.py).

* Affine grid (points A (x_i, y_j, z_k) + x0 with a sheared, stretched, rotated A and uneven
  spacings): every cell is a parallelepiped spanned by a = A e_x dx, b = A e_y dy, c = A e_z dz,
  so the i-face vector is b x c, the j-face vector c x a, the k-face vector a x b (half the
  cross product of a face's diagonals is its exact area vector), Grad its length, the volume
  a . (b x c), the centre the mean of the eight corners, Alagm the distances |a|, |b|, |c|
  between opposite face centres.  A swapped diagonal, a lost factor 1/2 or 1/3, an index
  shifted by one cell fails them.
* Curved grids (the bump channel, a compressor-sector block, the cascade's blade block): each
  cell's outward face vectors sum to zero (the faces are bilinear and close the cell), and the
  i-face vectors of a block's boundary agree with the neighbouring block's across an interface.
"""
import numpy as np

from trans_amd import cases, mesh
from trans_amd.mesh import O, BCRow, Block, sl


def _affine_block(ni=7, nj=6, nk=5):
    rng = np.random.Generator(np.random.MT19937(5))
    x1 = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, ni - 1))])
    y1 = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, nj - 1))])
    z1 = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, nk - 1))])
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    S = np.array([[1.3, 0.4, -0.2], [0.0, 0.8, 0.3], [0.0, 0.0, 1.1]])  # shear and stretch, det > 0
    A = R @ S
    x0 = np.array([2.0, -1.0, 0.5])
    Xi, Yj, Zk = np.meshgrid(x1, y1, z1, indexing="ij")
    P = np.einsum("ab,bijk->aijk", A, np.stack([Xi, Yj, Zk])) + x0[:, None, None, None]
    bcs = [BCRow(f, 0, 0, 0, 0, 3) for f in range(1, 7)]
    b = Block(ni, nj, nk, bcs)
    cases._points(b, P[0], P[1], P[2])
    mesh.build_connectivity([b])
    mesh.attach_wall_panels([b])
    mesh.compute_geometry([b])
    return b, A, x0, (x1, y1, z1)


def test_affine_grid_metrics_exact():
    b, A, x0, (x1, y1, z1) = _affine_block()
    NI1, NJ1, NK1 = b.ni - 1, b.nj - 1, b.nk - 1
    sd, grad, vol, alagm = b["sd"], b["grad"], b["vol"], b["alagm"]
    C = [b["xc"], b["yc"], b["zc"]]
    ex, ey, ez = A[:, 0], A[:, 1], A[:, 2]
    for i in range(1, NI1 + 1):
        for j in range(1, NJ1 + 1):
            for k in range(1, NK1 + 1):
                dx, dy, dz = x1[i] - x1[i - 1], y1[j] - y1[j - 1], z1[k] - z1[k - 1]
                a, bb, c = ex * dx, ey * dy, ez * dz
                I, J, K = i + O, j + O, k + O
                for d, want in ((0, np.cross(bb, c)), (1, np.cross(c, a)), (2, np.cross(a, bb))):
                    got = np.array([sd[d + 3 * x][I, J, K] for x in range(3)])
                    assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.linalg.norm(want)), (i, j, k, d)
                    assert abs(grad[d][I, J, K] - np.linalg.norm(want)) <= 1e-13 * np.linalg.norm(want)
                assert abs(vol[I, J, K] - np.dot(a, np.cross(bb, c))) <= 1e-13 * abs(np.dot(a, np.cross(bb, c)))
                xc = A @ np.array([0.5 * (x1[i - 1] + x1[i]), 0.5 * (y1[j - 1] + y1[j]), 0.5 * (z1[k - 1] + z1[k])]) + x0
                got = np.array([C[x][I, J, K] for x in range(3)])
                assert np.allclose(got, xc, rtol=0, atol=1e-13 * np.linalg.norm(xc))
                for d, L in ((0, a), (1, bb), (2, c)):
                    assert abs(alagm[d][I, J, K] - np.linalg.norm(L)) <= 1e-13 * np.linalg.norm(L)


def _closure(b):
    NI1, NJ1, NK1 = b.ni - 1, b.nj - 1, b.nk - 1
    sd, vol = b["sd"], b["vol"]
    I, J, K = sl(1, NI1), sl(1, NJ1), sl(1, NK1)
    I1, J1, K1 = sl(2, NI1 + 1), sl(2, NJ1 + 1), sl(2, NK1 + 1)
    worst = 0.0
    for x in range(3):
        s = (sd[0 + 3 * x][I1, J, K] - sd[0 + 3 * x][I, J, K] + sd[1 + 3 * x][I, J1, K] - sd[1 + 3 * x][I, J, K]
             + sd[2 + 3 * x][I, J, K1] - sd[2 + 3 * x][I, J, K])
        scale = np.maximum.reduce([np.abs(sd[d + 3 * y][I, J, K]) for d in range(3) for y in range(3)])
        worst = max(worst, float(np.max(np.abs(s) / scale)))
    return worst


def test_cells_are_closed_on_curved_grids():
    blocks, org, glob = cases.bump_channel(33, 17, 5)
    mesh.build_connectivity(blocks)
    mesh.attach_wall_panels(blocks)
    mesh.compute_geometry(blocks)
    assert _closure(blocks[0]) < 1e-12
    blocks, org, glob = cases.compressor_sector(scale=1 / 16, crossed=True)
    mesh.build_connectivity(blocks, nblade=glob.nblade, npassage=glob.Npassage)
    mesh.attach_wall_panels(blocks)
    mesh.compute_geometry(blocks)
    for b in blocks:
        assert _closure(b) < 1e-12
    blocks, org, glob = cases.cascade(scale=1 / 16, nk=4)
    mesh.build_connectivity(blocks)
    mesh.attach_wall_panels(blocks)
    mesh.compute_geometry(blocks)
    for b in blocks:
        assert _closure(b) < 1e-12


def test_interface_face_vectors_match_across_blocks():
    """The flat plate's blocks meet at i-faces: the last i-face vectors of a block equal the
    first of the next (same points, same orientation)."""
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    mesh.build_connectivity(blocks)
    mesh.attach_wall_panels(blocks)
    mesh.compute_geometry(blocks)
    for a, b in zip(blocks[:-1], blocks[1:]):
        J, K = sl(1, a.nj - 1), sl(1, a.nk - 1)
        for x in range(3):
            assert np.array_equal(a["sd"][0 + 3 * x][O + a.ni, J, K], b["sd"][0 + 3 * x][O + 1, J, K])
