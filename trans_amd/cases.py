"""Grid and deck generators for the benchmark configurations.

Each generator returns (blocks, OrgDeck, GlobalInfo) with point coordinates filled in and
the BC rows the reference would read from grd/dimension_blkNNN_m.dat.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import O, BCRow, Block
from .params import GlobalInfo, OrgDeck


def _points(b: Block, x, y, z):
    """Store point coordinates (NI,NJ,NK arrays, 1-based in the reference) into b."""
    shp = b.shape()
    for nm, a in (("xx", x), ("yy", y), ("zz", z)):
        full = np.zeros(shp, order="F")
        full[O + 1:O + 1 + b.ni, O + 1:O + 1 + b.nj, O + 1:O + 1 + b.nk] = a
        setattr(b, nm, full)


def bump_channel(ni=129, nj=65, nk=9, beta=2.0):
    """Channel 3c x 1c x 0.25c with a 10 %-thick circular-arc bump of chord c centred at
    x = 1.5c on the lower wall; uniform in i and k, tanh clustering toward the bump in j."""
    t, c = 0.1, 1.0
    R = (c * c / 4.0 + t * t) / (2.0 * t)
    x1 = np.linspace(0.0, 3.0, ni)
    yb = np.where(np.abs(x1 - 1.5) <= 0.5,
                  np.sqrt(np.maximum(R * R - (x1 - 1.5) ** 2, 0.0)) - (R - t), 0.0)
    eta = np.linspace(0.0, 1.0, nj)
    s = 1.0 + np.tanh(beta * (eta - 1.0)) / math.tanh(beta)
    z1 = np.linspace(0.0, 0.25, nk)
    X = np.broadcast_to(x1[:, None, None], (ni, nj, nk))
    Y = yb[:, None, None] + (1.0 - yb[:, None, None]) * s[None, :, None]
    Y = np.broadcast_to(Y, (ni, nj, nk))
    Z = np.broadcast_to(z1[None, None, :], (ni, nj, nk))
    bcs = [BCRow(1, 0, 0, 0, 0, 2), BCRow(2, 0, 0, 0, 0, 2),
           BCRow(3, 0, 0, 0, 0, 3), BCRow(4, 0, 0, 0, 0, 3),
           BCRow(5, 0, 0, 0, 0, 3), BCRow(6, 0, 0, 0, 0, 3)]
    b = Block(ni, nj, nk, bcs)
    _points(b, X, Y, Z)
    org = OrgDeck(XM=0.5, Tinf=288.0, Kind_Spatial=2, Kind_Roe=1, IF_Euler=1, IF_Turb_Org=0,
                  IF_RK=0, Kind_Dual=2, Kchld=1, DT_LUSGS_Main=1.0e10, CFL_Sub_start=5.0,
                  CFL_Sub_end=5.0, CFL_Sub_step=1, IF_Precondition=0, Kind_RotMod=0, MA=2000,
                  IF_NS_Uniform=1)
    glob = GlobalInfo(nblocks=1, P0in=101325.0 * (1 + 0.2 * 0.25) ** 3.5, T0in=288.0,
                      Pouthub=101325.0)
    return [b], org, glob


def split_sizes(total_cells: int, nblk: int, lo: int, hi: int, seed: int):
    """Point extents n_b in [lo, hi] with sum(n_b - 1) = total_cells, drawn by rejection
    from a 64-bit Mersenne twister seeded with `seed`."""
    rng = np.random.Generator(np.random.MT19937(seed))
    while True:
        n = rng.integers(lo, hi + 1, size=nblk)
        if int((n - 1).sum()) == total_cells:
            return [int(v) for v in n]


def flat_plate(nblk=8, ni_total=257, nj=129, nk=5, seed=20261202, re_l=5.0e6, mach=0.3,
               ratio=1.1, yplus=1.0, lo=17, hi=65, top=2):
    """Turbulent flat plate split into `nblk` blocks of uneven i-extent.

    Inflow through the total-pressure inlet (BC 5) at the first block's imin, static-pressure
    outlet (BC 7) at the last block's imax, far field (BC 2) at jmax.  The far-field BC is
    kept off the min faces: its outward normal there is -2 times the unit normal
    (BC/BC_far.F90:38), which doubles the reflection of the normal velocity and makes an
    imin far-field inflow diverge after ~100 iterations.  `top` is the BC kind at jmax
    (2 far field, 3 symmetry).

    The plate starts at x = 0 and ends at x = L (the block row covers a short
    symmetry run-in of 1/8 of the length); wall-normal geometric stretching at `ratio`
    with the first cell height from y+ ~ 1 at the plate end (Cf ~ 0.0027).  Block point
    extents are drawn from [lo, hi]."""
    from .params import make_params
    sizes = split_sizes(ni_total - 1, nblk, lo, hi, seed)
    org = OrgDeck(XM=mach, Tinf=288.0, Kind_Spatial=3, Kind_Roe=3, IF_Euler=0, IF_Turb_Org=1,
                  Kind_Model=1, IF_RK=0, Kind_Dual=2, Kchld=1, DT_LUSGS_Main=1.0e10,
                  CFL_Sub_start=5.0, CFL_Sub_end=5.0, CFL_Sub_step=1, KItm=1,
                  gama_max=0.012, gama_min=0.001, Prod_Lmt=20.0, VisRatio=1.0, FSTI=0.01,
                  IF_NS_Uniform=1, MA=20000)
    # freestream static state of the nondimensionalisation: rho = 1.2 kg/m^3, T = Tinf
    p_inf = 1.2 * 287.0 * org.Tinf
    stag = 1.0 + 0.2 * mach * mach
    glob = GlobalInfo(nblocks=nblk, P0in=p_inf * stag ** 3.5, T0in=org.Tinf * stag, Pouthub=p_inf)
    ref = make_params(org, glob).ref
    L = re_l / ref  # grid length of the plate so that Re_L = re_l
    cf = 0.0027
    dy1 = yplus * L / (re_l * math.sqrt(cf / 2.0))
    # wall-normal distribution: geometric growth up to a cap, domain height ~ L
    ys = [0.0]
    d = dy1
    for _ in range(nj - 1):
        ys.append(ys[-1] + d)
        d *= ratio
    y1 = np.array(ys)
    y1 = y1 * (L / y1[-1]) if y1[-1] > L else y1
    x_start = -L / 8.0
    x1 = np.linspace(x_start, L, ni_total)
    z1 = np.linspace(0.0, 0.1 * L, nk)
    blocks = []
    i0 = 0
    for ib, n in enumerate(sizes):
        xs = x1[i0:i0 + n]
        X = np.broadcast_to(xs[:, None, None], (n, nj, nk))
        Y = np.broadcast_to(y1[None, :, None], (n, nj, nk))
        Z = np.broadcast_to(z1[None, None, :], (n, nj, nk))
        bcs = []
        # imin: total-pressure inlet on the first block
        if ib == 0:
            bcs.append(BCRow(1, 0, 0, 0, 0, 5))
        else:
            bcs.append(BCRow(1, 0, 0, 0, 0, 9, 0, ib, 2))  # to previous block's imax (row 2)
        # imax: static-pressure outlet on the last block
        if ib == nblk - 1:
            bcs.append(BCRow(2, 0, 0, 0, 0, 7))
        else:
            bcs.append(BCRow(2, 0, 0, 0, 0, 9, 0, ib + 2, 1))
        # jmin: symmetry ahead of the leading edge, wall on the plate
        xc = 0.5 * (xs[:-1] + xs[1:])
        wall_cells = np.nonzero(xc > 0.0)[0]
        if len(wall_cells) == 0:
            bcs.append(BCRow(3, 0, 0, 0, 0, 3))
        elif len(wall_cells) == n - 1:
            bcs.append(BCRow(3, 0, 0, 0, 0, 1))
        else:
            first = int(wall_cells[0]) + 1
            bcs.append(BCRow(3, 1, first - 1, 0, 0, 3))
            bcs.append(BCRow(3, first, n - 1, 0, 0, 1))
        bcs.append(BCRow(4, 0, 0, 0, 0, top))
        bcs.append(BCRow(5, 0, 0, 0, 0, 3))
        bcs.append(BCRow(6, 0, 0, 0, 0, 3))
        b = Block(n, nj, nk, bcs)
        _points(b, X, Y, Z)
        blocks.append(b)
        i0 += n - 1
    # fix the interface rows: BC row numbers of imin/imax are 1 and 2 in every block
    return blocks, org, glob
