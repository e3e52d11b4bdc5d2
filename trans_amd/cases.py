"""Grid and deck generators for the benchmark configurations.

Each generator returns (blocks, OrgDeck, GlobalInfo) with point coordinates filled in and
the BC rows the reference would read from grd/dimension_blkNNN_m.dat.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

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


# --------------------------------------------------------------------------------------
# multi-block assembly from boxes of one global structured grid

@dataclass
class BoxSpec:
    """A block cut from a global structured grid: point index ranges lo..hi along the global
    axes, stored with local axis m running along global axis perm[m] (reversed where
    flips[m])."""
    lo: tuple
    hi: tuple
    perm: tuple = (0, 1, 2)
    flips: tuple = (0, 0, 0)

    def ncell_local(self):
        return [self.hi[self.perm[m]] - self.lo[self.perm[m]] for m in range(3)]

    def local_cells(self, m, g0, g1):
        """Local 1-based cell range along local axis m of the global cells g0..g1-1."""
        n = self.hi[self.perm[m]] - self.lo[self.perm[m]]
        a, b = g0 - self.lo[self.perm[m]], g1 - self.lo[self.perm[m]]
        return (a + 1, b) if not self.flips[m] else (n - b + 1, n - a)


def _face_rect(spec, m, side):
    """Global plane (axis, index) and tangential rectangle of local face (m, side)."""
    g = spec.perm[m]
    at_hi = (side == 1) != bool(spec.flips[m])
    c = spec.hi[g] if at_hi else spec.lo[g]
    tang = [t for t in range(3) if t != m]
    rect = {spec.perm[t]: (spec.lo[spec.perm[t]], spec.hi[spec.perm[t]]) for t in tang}
    return g, c, tang, rect


def assemble_blocks(specs, point_fn, dims, physical_bc, periodic_axis=None, periodic_ok=None,
                    periodic_turn=True):
    """Blocks and their BC rows (grd/dimension_blkNNN_m.dat, Input/Read_BlockDimensions.F90:61-74)
    for boxes that tile the global grid of dims = (NI, NJ, NK) cells conformingly, or part of
    it (cells no box covers are solid).

    Faces inside the grid become interface patches (BC 9) against every block sharing the
    plane; a face inside the grid that no other block shares bounds a solid region and takes
    physical_bc(axis, index, rect) -> (bctype, subbctype), as do faces on the grid's boundary,
    unless the axis is in `periodic_axis` (one axis or a tuple) and periodic_ok(rect) holds:
    then the face is a periodic interface to the blocks on the opposite boundary.  With
    periodic_turn (an annular sector) it is PerType +1 at index 0, whose ghosts lie one pitch
    below and take the donor turned back by the pitch, -1 at the top; without it (a linear
    cascade, a spanwise period) PerType 0, a plain copy of the translated donor.
    lcross and lrevA/lrevB are written from the geometry: lcross when the receiver's first
    tangential axis runs along the donor's second, lrev where a receiver axis runs against its
    donor counterpart.  point_fn(I, J, K) gives the global point coordinates."""
    nb = len(specs)
    if periodic_axis is None:
        per_axes = ()
    elif isinstance(periodic_axis, int):
        per_axes = (periodic_axis,)
    else:
        per_axes = tuple(periodic_axis)
    if periodic_ok is None:
        def periodic_ok(rect):
            return True
    faces = []  # per block: list of (m, side, g, c, tang, rect)
    for s in specs:
        fl = []
        for m in range(3):
            for side in (0, 1):
                fl.append((m, side) + _face_rect(s, m, side))
        faces.append(fl)

    def overlap(r1, r2):
        out = {}
        for g in r1:
            a, b = max(r1[g][0], r2[g][0]), min(r1[g][1], r2[g][1])
            if b <= a:
                return None
            out[g] = (a, b)
        return out

    def area(r):
        a = 1
        for (lo_, hi_) in r.values():
            a *= hi_ - lo_
        return a

    patches = [[] for _ in range(nb)]  # (m, side, rect, kind, data)
    for ib, s in enumerate(specs):
        for (m, side, g, c, tang, rect) in faces[ib]:
            if 0 < c < dims[g]:
                covered = 0
                for jb, s2 in enumerate(specs):
                    if jb == ib:
                        continue
                    for (m2, side2, g2, c2, tang2, rect2) in faces[jb]:
                        if g2 == g and c2 == c and (m2, side2) is not None:
                            ov = overlap(rect, rect2)
                            if ov and ((s.lo[g] == c) != (s2.lo[g] == c)):
                                patches[ib].append((m, side, ov, "iface", (jb, m2, side2, 0)))
                                covered += area(ov)
                if covered == 0:
                    bt, sub = physical_bc(g, c, rect)
                    patches[ib].append((m, side, rect, "phys", (bt, sub)))
                elif covered != area(rect):
                    raise ValueError(f"block {ib} face {(m, side)}: partly covered by other blocks")
            elif g in per_axes and periodic_ok(rect):
                opp = dims[g] - c
                pert = 0 if not periodic_turn else (1 if c == 0 else -1)
                for jb, s2 in enumerate(specs):
                    for (m2, side2, g2, c2, tang2, rect2) in faces[jb]:
                        if g2 == g and c2 == opp:
                            ov = overlap(rect, rect2)
                            if ov:
                                patches[ib].append((m, side, ov, "iface", (jb, m2, side2, pert)))
            else:
                bt, sub = physical_bc(g, c, rect)
                patches[ib].append((m, side, rect, "phys", (bt, sub)))
    # rows and their numbering
    rows = [[] for _ in range(nb)]
    index = {}
    for ib, s in enumerate(specs):
        for n, (m, side, rect, kind, data) in enumerate(patches[ib]):
            index[(ib, m, side, tuple(sorted(rect.items())))] = n
    blocks = []
    for ib, s in enumerate(specs):
        nloc = s.ncell_local()
        for (m, side, rect, kind, data) in patches[ib]:
            tang = [t for t in range(3) if t != m]
            (a0, a1), (b0, b1) = (s.local_cells(t, *rect[s.perm[t]]) for t in tang)
            face = 2 * m + 1 + side
            if kind == "phys":
                rows[ib].append(BCRow(face, a0, a1, b0, b1, data[0], data[1]))
                continue
            jb, m2, side2, pert = data
            s2 = specs[jb]
            conn = index[(jb, m2, side2, tuple(sorted(rect.items())))]
            tang2 = [t for t in range(3) if t != m2]
            gr = [s.perm[t] for t in tang]
            gd = [s2.perm[t] for t in tang2]
            lcross = 1 if gr[0] == gd[1] else 0
            rev = []
            for q in range(2):
                t2 = tang2[gd.index(gr[q])]
                rev.append(1 if bool(s.flips[tang[q]]) != bool(s2.flips[t2]) else 0)
            rows[ib].append(BCRow(face, a0, a1, b0, b1, 9, 0, jb + 1, conn + 1, lcross, rev[0], rev[1], pert))
        b = Block(nloc[0] + 1, nloc[1] + 1, nloc[2] + 1, rows[ib])
        # points in local order
        gi = [np.arange(s.lo[g], s.hi[g] + 1) for g in range(3)]
        G = np.meshgrid(*gi, indexing="ij")
        XYZ = point_fn(*G)
        loc = []
        for A in XYZ:
            B = np.transpose(A, s.perm)
            for m in range(3):
                if s.flips[m]:
                    B = np.flip(B, axis=m)
            loc.append(np.ascontiguousarray(B))
        _points(b, *loc)
        b.spec = s
        blocks.append(b)
    return blocks


def _two_sided(n, first):
    """n+1 points on [0, 1] clustered at both ends, first cell of width ~ `first` (tanh)."""
    eta = np.linspace(0.0, 1.0, n + 1)

    def dist(beta):
        return 0.5 * (1.0 + np.tanh(beta * (2.0 * eta - 1.0)) / math.tanh(beta))
    lo, hi = 1e-6, 20.0
    if dist(lo)[1] <= first:
        return eta
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if dist(mid)[1] > first:
            lo = mid
        else:
            hi = mid
    return dist(0.5 * (lo + hi))


# NACA 65-series 010 thickness ordinates (x/c, y/c in %), Abbott & von Doenhoff, Theory of
# Wing Sections, appendix I; half thickness, 10 % thick section
_NACA65_X = [0, .5, .75, 1.25, 2.5, 5, 7.5, 10, 15, 20, 25, 30, 35, 40, 45, 50, 55, 60, 65, 70, 75, 80, 85,
             90, 95, 100]
_NACA65_Y = [0, .772, .932, 1.169, 1.574, 2.177, 2.647, 3.040, 3.666, 4.143, 4.503, 4.760, 4.924, 4.996,
             4.963, 4.812, 4.530, 4.146, 3.682, 3.156, 2.584, 1.987, 1.385, 0.810, 0.306, 0.0]


def naca65_half_thickness(xi, t_max):
    """Half thickness (fraction of chord) of a NACA-65-style section of maximum thickness t_max
    at chordwise positions xi in [0, 1] (monotone cubic interpolation of the 65-010 table)."""
    from scipy.interpolate import PchipInterpolator
    f = PchipInterpolator(np.array(_NACA65_X) / 100.0, np.array(_NACA65_Y) / 100.0 * (t_max / 0.10))
    return f(np.clip(xi, 0.0, 1.0))


def compressor_sector(scale=1.0, nblade=36, rpm=1200.0 * 60.0 / (2.0 * math.pi), stagger_deg=30.0,
                      camber_deg=20.0, t_max=0.06, r_tip=0.25, hub_tip=0.6, chord=0.05, back_pressure=1.2,
                      crossed=False, mach=0.5, seg=None, span_cells=None, pitch_cells=None):
    """Analytic annular compressor-rotor sector: one blade with NACA-65-style thickness on a
    circular-arc camber line (stagger `stagger_deg` from the axis, camber `camber_deg`), hub/tip
    radius ratio `hub_tip`, `nblade` blades so the periodic sector spans 360/nblade degrees,
    rotating about x at `rpm` (dimension_global.dat's omega in rpm, Input/Read_GlobalInfo.F90:52);
    inlet total pressure / temperature 101325 Pa / 288 K, outlet static pressure at the hub
    `back_pressure` times the inlet total pressure.

    H-type topology in 16 blocks: axial segments upstream (U), two over the blade (B1, B2), two
    downstream (D1, D2); U and D2 span hub to casing, B1, B2 and D1 are split at mid span;
    every segment is split at mid pitch.  Between the blades' periodic lines upstream and
    downstream the pitch halves meet through the periodic interface (PerType +-1,
    Input/PLOT3DReadInit.F90:155-160); over the blade the k faces are the blade's suction and
    pressure surfaces (walls rotating with the frame, SubBCType 0); the hub rotates (SubBCType
    0), the casing is stationary (SubBCType 1: ghost velocity omega x r in the relative frame,
    BC/BC_Wall.F90:41-90).  Inlet at U's imin (BC 5), outlet at D2's imax (BC 7) spanning hub to
    tip so that its radial-equilibrium march covers the full span.  The deck ramps the CFL
    number from 1 to 5 over the first 50 iterations (Time/SpectralRadius.F90:159-160): from the
    uniform start the full-scale grid diverges within ten iterations at CFL 5, and runs at CFL
    2 or with the ramp.

    Cells at scale 1: segments 48, 128, 128, 128, 64 along x, 128 radial, 128 circumferential:
    blocks of 49x129x65 (2), 129x65x65 (12) and 65x129x65 (2) points, 8,126,464 cells.
    `crossed` stores the tip-side pitch-max block of B2 with its i and k axes exchanged and k
    reversed, so that its interfaces need lcross and lrev.  Returns (blocks, OrgDeck,
    GlobalInfo): Roe + MUSCL, RANS-SST, LU-SGS at CFL 5 (after the ramp), dual time with one
    sub-iteration."""
    from .params import GlobalInfo, OrgDeck
    if seg is None:
        seg = [max(3, int(round(n * scale))) for n in (48, 128, 128, 128, 64)]
    nj = span_cells or max(8, 2 * int(round(64 * scale)))
    nk = pitch_cells or max(8, 2 * int(round(64 * scale)))
    nj -= nj % 2
    nk -= nk % 2
    pitch = 2.0 * math.pi / nblade
    r_hub = hub_tip * r_tip
    r_mid = 0.5 * (r_hub + r_tip)
    gam = math.radians(stagger_deg)
    phi = math.radians(camber_deg)
    cx = chord * math.cos(gam)  # axial chord
    x_le = chord
    x_te = x_le + cx
    x_in, x_out = 0.0, x_te + 2.0 * chord
    # camber line in the unrolled (x, s) plane, s = r_mid theta: circular arc of camber phi on
    # a chord at stagger gam, leading edge at (x_le, 0), turning toward the axis
    R = chord / (2.0 * math.sin(phi / 2.0))
    h0 = math.sqrt(R * R - chord * chord / 4.0)

    def camber(xi):
        """(x, s) of the camber line and the half thickness (in s) at chord fraction xi."""
        u = xi * chord
        yc = np.sqrt(R * R - (u - chord / 2.0) ** 2) - h0
        x = x_le + u * math.cos(gam) + yc * math.sin(gam)
        s = -u * math.sin(gam) + yc * math.cos(gam)
        return x, s

    # axial stations: segments U | B1 B2 | D1 D2 with clustering at the leading and trailing edges
    nU, nB1, nB2, nD1, nD2 = seg
    xs_u = x_in + (x_le - x_in) * np.sin(np.linspace(0.0, math.pi / 2, nU + 1))
    xi_b = _two_sided(nB1 + nB2, 0.25 / (nB1 + nB2))
    xs_d = x_te + (x_out - x_te) * (1.0 - np.cos(np.linspace(0.0, math.pi / 2, nD1 + nD2 + 1)))
    # camber line sampled at the blade's axial stations: invert x(xi) by bisection on xi
    xb_target = x_le + xi_b * (x_te - x_le)
    xi_of_x = np.empty_like(xb_target)
    for q, xt in enumerate(xb_target):
        lo, hi = 0.0, 1.0
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if camber(mid)[0] < xt:
                lo = mid
            else:
                hi = mid
        xi_of_x[q] = 0.5 * (lo + hi)
    xi_of_x[0], xi_of_x[-1] = 0.0, 1.0
    _, s_b = camber(xi_of_x)
    th_b = s_b / r_mid
    tau_b = 2.0 * naca65_half_thickness(xi_of_x, t_max) * chord / r_mid  # angular thickness
    # straight extensions of the camber line upstream and downstream (its end slopes)
    d = 1e-6
    sl_le = (camber(d)[1] - camber(0.0)[1]) / (camber(d)[0] - camber(0.0)[0]) / r_mid
    sl_te = (camber(1.0)[1] - camber(1.0 - d)[1]) / (camber(1.0)[0] - camber(1.0 - d)[0]) / r_mid
    xs = np.concatenate([xs_u, xb_target[1:], xs_d[1:]])
    th_lo = np.concatenate([th_b[0] + sl_le * (xs_u - x_le), th_b[1:] + 0.5 * tau_b[1:],
                            th_b[-1] + sl_te * (xs_d[1:] - x_te)])
    th_hi = np.concatenate([th_b[0] + sl_le * (xs_u - x_le) + pitch, th_b[1:] - 0.5 * tau_b[1:] + pitch,
                            th_b[-1] + sl_te * (xs_d[1:] - x_te) + pitch])
    NI = len(xs) - 1
    # radial and circumferential distributions (first cells ~ y+ 1 at the walls)
    re_m = 1.2 * mach * math.sqrt(1.4 * 287.0 * 288.0) / 1.8e-5
    rr = r_hub + (r_tip - r_hub) * _two_sided(nj, max(2.0 / re_m / (r_tip - r_hub) * 50, 0.2 / nj))
    eta_k = _two_sided(nk, max(2.0 / re_m / (pitch * r_mid) * 50, 0.2 / nk))

    def point_fn(I, J, K):
        x = xs[I]
        r = rr[J]
        th = th_lo[I] + (th_hi[I] - th_lo[I]) * eta_k[K]
        return x, r * np.cos(th), r * np.sin(th)

    iU, iB1, iB2, iD1 = nU, nU + nB1, nU + nB1 + nB2, nU + nB1 + nB2 + nD1
    assert NI == iD1 + nD2
    jm, km = nj // 2, nk // 2
    specs = []
    for (i0, i1, split) in ((0, iU, False), (iU, iB1, True), (iB1, iB2, True), (iB2, iD1, True), (iD1, NI, False)):
        for (k0, k1) in ((0, km), (km, nk)):
            spans = ((0, jm), (jm, nj)) if split else ((0, nj),)
            for (j0, j1) in spans:
                specs.append(BoxSpec((i0, j0, k0), (i1, j1, k1)))
    if crossed:
        # B2, tip half, pitch-max half: local i = global k reversed, local k = global i
        q = [n for n, s in enumerate(specs) if s.lo == (iB1, jm, km)][0]
        specs[q] = BoxSpec(specs[q].lo, specs[q].hi, perm=(2, 1, 0), flips=(1, 0, 0))

    def physical_bc(g, c, rect):
        if g == 0:
            return (5, 0) if c == 0 else (7, 0)
        if g == 1:
            return (1, 0) if c == 0 else (1, 1)
        return (1, 0)  # blade surfaces

    def periodic_ok(rect):
        i0, i1 = rect[0]
        return i1 <= iU or i0 >= iB2

    blocks = assemble_blocks(specs, point_fn, (NI, nj, nk), physical_bc, periodic_axis=2, periodic_ok=periodic_ok)
    org = OrgDeck(XM=mach, Tinf=288.0, Kind_Spatial=2, Kind_Roe=1, IF_Euler=0, IF_Turb_Org=1, Kind_Model=1,
                  IF_RK=0, Kind_Dual=2, Kchld=1, DT_LUSGS_Main=1.0e10, CFL_Sub_start=1.0, CFL_Sub_end=5.0,
                  CFL_Sub_step=50, KItm=1, Prod_Lmt=20.0, VisRatio=1.0, FSTI=0.01, IF_NS_Uniform=1, MA=20000)
    glob = GlobalInfo(nblocks=len(blocks), P0in=101325.0, T0in=288.0, Pouthub=back_pressure * 101325.0,
                      nblade=nblade, Npassage=1, omega_rpm=(rpm, 0.0, 0.0))
    return blocks, org, glob


def cascade(scale=1.0, nk=None, mach=0.7, re_c=1.0e6, t_max=0.06, t_le=0.005, t_te=0.02, pitch_c=0.8,
            depth_c=0.2, x_up_c=1.0, x_down_c=2.0, dy_wall_c=1.0e-4, r0_c=100.0, seg=None, cfl=0.8):
    """Linear cascade of symmetric blades with a blunt trailing edge, extruded in span.

    Section: thickness t(xi) = t_le + (t_te - t_le) xi + NACA-65-style thickness of maximum
    t_max (fractions of chord c), so the trailing edge is a flat face of height t_te c (and the
    leading edge one of height t_le c); zero stagger and camber, axial inflow.  Pitch
    `pitch_c` c, span depth `depth_c` c with a spanwise period, the pitchwise period a
    translation (PerType 0); the blade row sits at y = r0_c c above the x axis so that the
    outlet's radial-equilibrium march (BC/BC_Outlet.F90:49-103, along j = pitch) adds only
    O(pitch / r0) to the outlet pressure.

    H topology on the global cell box NI x NJ x NK, i streamwise, j pitchwise, k spanwise: the
    pitch is split into the band j < nS, which continues the blade's thickness upstream (height
    t_le) and downstream (height t_te) and is solid over the blade, and the passage j >= nS from
    the blade's upper surface to the next blade's lower surface.  The band's upstream end is the
    leading-edge face, its downstream start the trailing-edge face (walls), the passage's j
    faces over the blade are the blade surfaces (walls).  Blocks: upstream band and passage (2),
    blade passage split twice in i and once in j (4), wake band and passage in three streamwise
    segments (6): 12 blocks of uneven sizes, 19.27M cells at scale 1 (NK = 64).

    Deck: SST-DES (Kind_Hybrid 3, Kind_SubHY 0), AUSMPW+ with MDCD, explicit SSP-RK3 (IF_RK 1)
    with one global time step (Kind_Dual 0: Dtm = DT_LUSGS_Main, Time/SpectralRadius.F90:146-176)
    set on the host to `cfl` times the smallest local step of the uniform start (see
    global_time_step), all FP64.  Inlet total pressure / temperature of the exit Mach number
    `mach` at rho = 1.2 kg/m^3 and 288 K, outlet static pressure that of the uniform start,
    Re_c = re_c."""
    from .params import make_params
    if seg is None:
        seg = [max(2, int(round(n * scale))) for n in (64, 96, 96, 96, 128, 160)]
    nU, nB1, nB2, nW1, nW2, nW3 = seg
    nS = max(4, 2 * int(round(16 * scale)))
    nP = max(8, 2 * int(round(224 * scale)))
    NK = nk or max(4, int(round(64 * scale)))
    nB, nW = nB1 + nB2, nW1 + nW2 + nW3
    NI, NJ = nU + nB + nW, nS + nP
    org = OrgDeck(XM=mach, Tinf=288.0, Kind_Spatial=3, Kind_Roe=3, IF_Euler=0, IF_Turb_Org=1, Kind_Model=1,
                  IF_RK=1, Kind_Dual=0, Kchld=1, DT_LUSGS_Main=1.0, CFL_Sub_start=cfl, CFL_Sub_end=cfl,
                  CFL_Sub_step=1, KItm=1, gama_max=0.012, gama_min=0.001, Prod_Lmt=20.0, VisRatio=1.0,
                  FSTI=0.01, Kind_Hybrid=3, Kind_SubHY=0, IF_NS_Uniform=1, MA=5000)
    p_inf = 1.2 * 287.0 * org.Tinf
    stag = 1.0 + 0.2 * mach * mach
    glob = GlobalInfo(nblocks=12, P0in=p_inf * stag ** 3.5, T0in=org.Tinf * stag, Pouthub=p_inf)
    c = re_c / make_params(org, glob).ref  # chord in grid units: Re_c = ref * c
    pitch, depth, r0 = pitch_c * c, depth_c * c, r0_c * c

    def thick(xi):
        return c * (t_le + (t_te - t_le) * xi + 2.0 * naca65_half_thickness(xi, t_max))

    xs_u = -x_up_c * c * (1.0 - np.sin(np.linspace(0.0, math.pi / 2, nU + 1)))
    xs_b = c * _two_sided(nB, min(0.5 / nB, 0.004))
    xs_w = c + x_down_c * c * (1.0 - np.cos(np.linspace(0.0, math.pi / 2, nW + 1)))
    xs = np.concatenate([xs_u, xs_b[1:], xs_w[1:]])
    xi = np.clip(xs / c, 0.0, 1.0)
    tt = thick(xi)
    eta_s = _two_sided(nS, min(0.5 / nS, dy_wall_c / t_te))
    eta_p = _two_sided(nP, dy_wall_c / (pitch_c - t_te))
    zs = np.linspace(0.0, depth, NK + 1)
    assert len(xs) == NI + 1

    def point_fn(I, J, K):
        t = tt[I]
        Jb = np.minimum(J, nS)
        Jp = np.maximum(J - nS, 0)
        y = np.where(J <= nS, -0.5 * t + t * eta_s[Jb], 0.5 * t + (pitch - t) * eta_p[Jp])
        return xs[I], r0 + y, zs[K]

    i1, i2, i3 = nU, nU + nB1, nU + nB
    w1, w2 = i3 + nW1, i3 + nW1 + nW2
    jm = nS + nP // 2
    specs = [BoxSpec((0, 0, 0), (i1, nS, NK)), BoxSpec((0, nS, 0), (i1, NJ, NK))]
    for (a, b_) in ((i1, i2), (i2, i3)):
        specs += [BoxSpec((a, nS, 0), (b_, jm, NK)), BoxSpec((a, jm, 0), (b_, NJ, NK))]
    for (a, b_) in ((i3, w1), (w1, w2), (w2, NI)):
        specs += [BoxSpec((a, 0, 0), (b_, nS, NK)), BoxSpec((a, nS, 0), (b_, NJ, NK))]

    def physical_bc(g, cc, rect):
        if g == 0 and cc == 0:
            return (5, 0)
        if g == 0 and cc == NI:
            return (7, 0)
        return (1, 0)  # leading- and trailing-edge faces, blade surfaces

    def periodic_ok(rect):
        if 0 in rect and 2 in rect:  # a j face: periodic up- and downstream of the blade
            i0, i1_ = rect[0]
            return i1_ <= i1 or i0 >= i3
        return True  # k faces: the spanwise period

    blocks = assemble_blocks(specs, point_fn, (NI, NJ, NK), physical_bc, periodic_axis=(1, 2),
                             periodic_ok=periodic_ok, periodic_turn=False)
    glob.nblocks = len(blocks)
    return blocks, org, glob


def global_time_step(blocks, prm, cfl):
    """cfl times the smallest local time step vol / (Rds1 + Rds2 + Rds3) of the current state
    (Time/SpectralRadius.F90:40-138 without preconditioning): the global step of a Kind_Dual 0
    deck.  Needs the geometry and the initial state of prepare_case."""
    from .mesh import O
    dt = math.inf
    for b in blocks:
        A = b.arrays
        I = slice(O + 1, O + b.ni)
        J = slice(O + 1, O + b.nj)
        K = slice(O + 1, O + b.nk)
        I1 = slice(O + 2, O + b.ni + 1)
        J1 = slice(O + 2, O + b.nj + 1)
        K1 = slice(O + 2, O + b.nk + 1)
        v = A["v"]
        den = v[0][I, J, K]
        u = [v[m][I, J, K] / den for m in (1, 2, 3)]
        pp = A["pp"][I, J, K]
        aa = np.sqrt(1.4 * pp / den)
        tt = A["t"][I, J, K]
        rmiul = (1.0 + prm.csthlnd) / (tt * prm.csth1 + prm.csthlnd) * (prm.csth1 * tt) ** 1.5
        rmiut = A["rmiu"][I, J, K]
        uvmu = (rmiul + rmiut + 1.0e-20) * 4.0 / 3.0
        uvmupr = (rmiul / 0.72 + rmiut / 0.9 + 1.0e-20) * 1.4
        pr = np.maximum(uvmu, uvmupr) / den
        vol = A["vol"][I, J, K]
        rsum = 0.0
        for d, (nxt) in enumerate(((I1, J, K), (I, J1, K), (I, J, K1))):
            s = [0.5 * (A["sd"][d + 3 * x][I, J, K] + A["sd"][d + 3 * x][nxt]) for x in range(3)]
            st = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
            un = u[0] * s[0] + u[1] * s[1] + u[2] * s[2]
            rsum = rsum + np.abs(un) + np.sqrt(st) * aa + 2.0 * st / vol * pr / prm.ref
        dt = min(dt, float(np.min(vol / rsum)))
    return cfl * dt
