"""Generated block pairs for the interface pins: every orientation of a donor block against a
receiver block, and a periodic annular sector.

Nothing here restates the reference's index algebra.  The deck rows are written the way a
deck author writes them from the geometry alone (which local axis of the neighbour runs
along which of mine, and in which sense), and the expected ghost cells come from the global
point grid the blocks were cut from.
"""
import itertools
import math

import numpy as np

from trans_amd import cases
from trans_amd.mesh import O, BCRow, Block

# receiver R: 4 x 4 x 3 cells; donor T: 3 x 4 x 3 cells of the global grid, glued to R's imax
NR, NT, NY, NZ = 4, 3, 4, 3


def global_points():
    """Union point grid (NR+NT+1) x (NY+1) x (NZ+1), slightly sheared so that no two axes
    are alike and every cell centre is distinct."""
    x1 = np.array([0.0, 0.3, 0.55, 0.9, 1.2, 1.45, 1.8, 2.05])[:NR + NT + 1]
    y1 = np.array([0.0, 0.2, 0.45, 0.65, 0.95])[:NY + 1]
    z1 = np.array([0.0, 0.35, 0.6, 1.0])[:NZ + 1]
    X, Y, Z = np.meshgrid(x1, y1, z1, indexing="ij")
    X = X + 0.07 * Y - 0.04 * Z
    Z = Z + 0.05 * X
    return X, Y, Z


def global_centres(X, Y, Z, gi, gj, gk):
    """Centre of global cell (gi, gj, gk) (0-based) as the mean of its 8 corners."""
    out = []
    for A in (X, Y, Z):
        s = 0.0
        for a, b, c in itertools.product((0, 1), repeat=3):
            s = s + A[gi + a, gj + b, gk + c]
        out.append(s / 8.0)
    return out


ORIENTATIONS = [(perm, flips) for perm in itertools.permutations(range(3))
                for flips in itertools.product((0, 1), repeat=3)]


def _tangential(d):
    """Tangential local axes of a face whose normal is axis d (0-based), lower first."""
    return [a for a in range(3) if a != d]


def orientation_pair(perm, flips, bctype_other=3):
    """Blocks R and T (T stored with local axis m = global axis perm[m], reversed where
    flips[m]), their interface rows written from the geometry, and the maps needed to check
    the ghosts.  Returns (blocks, info)."""
    X, Y, Z = global_points()
    R = Block(NR + 1, NY + 1, NZ + 1, [])
    cases._points(R, X[:NR + 1], Y[:NR + 1], Z[:NR + 1])
    Tg = [A[NR:] for A in (X, Y, Z)]
    Tl = []
    for A in Tg:
        B = np.transpose(A, perm)
        for m in range(3):
            if flips[m]:
                B = np.flip(B, axis=m)
        Tl.append(np.ascontiguousarray(B))
    ncell_g = [NT, NY, NZ]  # T's cells along global axes
    nloc = [ncell_g[perm[m]] for m in range(3)]
    T = Block(nloc[0] + 1, nloc[1] + 1, nloc[2] + 1, [])
    cases._points(T, *Tl)
    # T's face on global imin: the local axis carrying global axis 0, at its min (or max if reversed)
    dT = perm.index(0)
    faceT = 2 * dT + (2 if flips[dT] else 1)
    tT = _tangential(dT)
    # R -> T: R's tangential axes are global y (1) then z (2), both increasing
    img = {perm[t]: t for t in tT}  # global axis -> T local tangential axis
    lcross_R = 1 if img[1] == tT[1] else 0
    rev_R = [flips[img[1]], flips[img[2]]]
    # T -> R: T's tangential axes tT[0], tT[1]; lcross when T's first one runs along global z
    lcross_T = 1 if perm[tT[0]] == 2 else 0
    rev_T = [flips[tT[0]], flips[tT[1]]]
    # rows: R faces 1..6 with imax = interface (row 2); T's interface face is its row faceT
    R.bcs = [BCRow(f, 0, 0, 0, 0, bctype_other) for f in range(1, 7)]
    R.bcs[1] = BCRow(2, 0, 0, 0, 0, 9, 0, 2, faceT, lcross_R, rev_R[0], rev_R[1])
    T.bcs = [BCRow(f, 0, 0, 0, 0, bctype_other) for f in range(1, 7)]
    T.bcs[faceT - 1] = BCRow(faceT, 0, 0, 0, 0, 9, 0, 1, 2, lcross_T, rev_T[0], rev_T[1])

    def t_local_to_global(c):
        """T local cell index (0-based, may lie outside the block) -> global cell index."""
        g = [0, 0, 0]
        for m in range(3):
            gm = c[m] if not flips[m] else nloc[m] - 1 - c[m]
            g[perm[m]] = gm
        g[0] += NR
        return g

    info = dict(X=X, Y=Y, Z=Z, perm=perm, flips=flips, nloc=nloc, dT=dT, faceT=faceT,
                t_local_to_global=t_local_to_global,
                rows=dict(R=(lcross_R, *rev_R), T=(lcross_T, *rev_T)))
    return [R, T], info


def expected_ghosts(blocks, info):
    """For each block, a list of (ghost cell index in reference numbering, global cell
    (gi, gj, gk), receiver axis -> global axis map) of every interface ghost of layers 1-3."""
    R, T = blocks
    out = {0: [], 1: []}
    # R's imax ghosts: reference cell (NR + L, j, k), L = 1..3, is global cell (NR - 1 + L, j-1, k-1)
    for L in (1, 2, 3):
        for k in range(1, NZ + 1):
            for j in range(1, NY + 1):
                out[0].append(((NR + L, j, k), (NR - 1 + L, j - 1, k - 1), L))
    dT = info["dT"]
    nloc = info["nloc"]
    ranges = [range(1, nloc[m] + 1) for m in range(3)]
    lo_face = info["faceT"] % 2 == 1
    for L in (1, 2, 3):
        ranges_l = list(ranges)
        ranges_l[dT] = [1 - L] if lo_face else [nloc[dT] + L]
        for c in itertools.product(*ranges_l):
            g = info["t_local_to_global"]([c[0] - 1, c[1] - 1, c[2] - 1])
            out[1].append((tuple(c), tuple(g), L))
    return out


def annular_sector(ni=7, nj=6, nk=8, pitch_deg=10.0, r_hub=0.6, r_tip=1.0, length=0.8, theta0_deg=20.0,
                   twist=0.15, nsplit_k=2):
    """Annular sector about the x axis (rotation axis of the rotating frame) spanning one pitch,
    cut into `nsplit_k` blocks along the circumferential index k, with a plain interface between
    them and the periodic interface across the pitch (BC/BC_P2P.F90 with rotateAngle =
    PerType * 2 pi / nblade, Input/PLOT3DReadInit.F90:155-160).  Grid lines in theta follow a
    twist in x so that the sector faces are not planes of constant theta.  Returns
    (blocks, point_fn) where point_fn(i, j, k) (0-based, any real k) gives the points of the
    global grid and its periodic continuation, for the geometric checks."""
    pitch = math.radians(pitch_deg)
    th0 = math.radians(theta0_deg)
    nblade = int(round(360.0 / pitch_deg))
    x1 = np.linspace(0.0, length, ni)
    r1 = r_hub + (r_tip - r_hub) * (np.linspace(0.0, 1.0, nj) ** 1.2)

    def point_fn(i, j, k):
        x = x1[i]
        r = r1[j]
        th = th0 + twist * x + pitch * (k / (nk - 1))
        return x, r * math.cos(th), r * math.sin(th)

    kcuts = np.linspace(0, nk - 1, nsplit_k + 1).round().astype(int)
    blocks = []
    for s in range(nsplit_k):
        k0, k1 = kcuts[s], kcuts[s + 1]
        nkb = k1 - k0 + 1
        X = np.empty((ni, nj, nkb)); Y = np.empty_like(X); Z = np.empty_like(X)
        for i in range(ni):
            for j in range(nj):
                for kk in range(nkb):
                    X[i, j, kk], Y[i, j, kk], Z[i, j, kk] = point_fn(i, j, k0 + kk)
        bcs = [BCRow(1, 0, 0, 0, 0, 2), BCRow(2, 0, 0, 0, 0, 2), BCRow(3, 0, 0, 0, 0, 1),
               BCRow(4, 0, 0, 0, 0, 1, 1)]
        # kmin: periodic to the last block's kmax (receiver at theta - pitch: donor rotated by +pitch
        # brings it back, rotateAngle = +pitch) or plain to the previous block
        if s == 0:
            bcs.append(BCRow(5, 0, 0, 0, 0, 9, 0, nsplit_k, 6, 0, 0, 0, 1))
        else:
            bcs.append(BCRow(5, 0, 0, 0, 0, 9, 0, s, 6))
        if s == nsplit_k - 1:
            bcs.append(BCRow(6, 0, 0, 0, 0, 9, 0, 1, 5, 0, 0, 0, -1))
        else:
            bcs.append(BCRow(6, 0, 0, 0, 0, 9, 0, s + 2, 5))
        b = Block(ni, nj, nkb, bcs)
        cases._points(b, X, Y, Z)
        b.k0 = int(k0)
        blocks.append(b)
    return blocks, point_fn, nblade


def sector_cell_centre(point_fn, i, j, k):
    """Centre of the cell with 0-based corner indices (i, j, k) of the sector's global grid,
    k extended periodically beyond the sector by the point function itself."""
    s = np.zeros(3)
    for a, b, c in itertools.product((0, 1), repeat=3):
        s += np.array(point_fn(i + a, j + b, k + c))
    return s / 8.0
