"""Host-side block, connectivity and geometry preparation.

This is the host half of the drop-in: it turns the reference's deck rows and grids into
the arrays the C-ABI takes, following the reference's own pre-processing so that every
array means what the reference's does:

* connectivity -- PLOT3D_Read_And_Init (Input/PLOT3DReadInit.F90:1-349)
* metrics      -- GEOM_CalcDST (Geom/CalcDst.F90:28-719)
* ghost geometry at physical BCs -- GeomTransform (Geom/startup.F90:396-508)
* ghost geometry across interfaces -- GeomTransform1 (Geom/startup.F90:510-755)
* initial field -- INPUT_Inititializes (Input/Init.F90:94-169)

Arrays use the reference layout: Fortran order, component index first, and bounds
-2:N+2 in each direction (N = points), so A[c, i+2, j+2, k+2] is A(c+1, i, j, k).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi

O = 2  # offset of index 0 in a -2:N+2 array


def sl(lo: int, hi: int) -> slice:
    """Slice covering reference indices lo..hi of a -2:N+2 axis."""
    return slice(lo + O, hi + O + 1)


@dataclass
class BCRow:
    """One BC line of grd/dimension_blkNNN_m.dat (Input/Read_BlockDimensions.F90:61-74)."""
    face: int          # 1..6 = imin, imax, jmin, jmax, kmin, kmax
    a0: int
    a1: int
    b0: int
    b1: int
    bctype: int
    subbctype: int = 0
    target_blk: int = 0  # 1-based, 0 = self
    target_conn: int = 0  # 1-based
    lcross: int = 0
    lrev_a: int = 0
    lrev_b: int = 0
    pertype: int = 0


@dataclass
class Block:
    ni: int
    nj: int
    nk: int
    bcs: list
    xx: np.ndarray = None  # point coordinates, (NI+5, NJ+5, NK+5)
    yy: np.ndarray = None
    zz: np.ndarray = None
    walls: list = field(default_factory=list)  # faces (1..6) marked as walls
    arrays: dict = field(default_factory=dict)
    patches: list = field(default_factory=list)  # list of dicts (ConnectivityStruct)

    @property
    def cells(self) -> int:
        return (self.ni - 1) * (self.nj - 1) * (self.nk - 1)

    def shape(self):
        return (self.ni + 5, self.nj + 5, self.nk + 5)

    def alloc(self, name: str, ncomp: int = 1, fill: float = 0.0) -> np.ndarray:
        shp = self.shape() if ncomp == 1 else (ncomp,) + self.shape()
        a = np.full(shp, fill, dtype=np.float64, order="F")
        self.arrays[name] = a
        return a

    def __getitem__(self, name):
        return self.arrays[name]


# --------------------------------------------------------------------------------------
# connectivity (Input/PLOT3DReadInit.F90:19-348)

def _prange(b: Block, r: BCRow):
    NX, NY, NZ = b.ni, b.nj, b.nk
    a0, a1, b0, b1 = r.a0, r.a1, r.b0, r.b1
    if r.face in (1, 2):
        a0 = a0 or 1; a1 = a1 or NY - 1; b0 = b0 or 1; b1 = b1 or NZ - 1
        I = 1 if r.face == 1 else NX
        return (I, a0, b0, I, a1, b1)
    if r.face in (3, 4):
        a0 = a0 or 1; a1 = a1 or NX - 1; b0 = b0 or 1; b1 = b1 or NZ - 1
        J = 1 if r.face == 3 else NY
        return (a0, J, b0, a1, J, b1)
    if r.face in (5, 6):
        a0 = a0 or 1; a1 = a1 or NX - 1; b0 = b0 or 1; b1 = b1 or NY - 1
        K = 1 if r.face == 5 else NZ
        return (a0, b0, K, a1, b1, K)
    raise ValueError("face must be 1..6")


def build_connectivity(blocks: list, nblade: int = 1, npassage: int = 1):
    """Fill Block.patches with the ConnectivityStruct fields, one per BC row."""
    pr = [[_prange(b, r) for r in b.bcs] for b in blocks]
    for ib, b in enumerate(blocks):
        b.patches = []
        for ic, r in enumerate(b.bcs):
            Imin, Jmin, Kmin, Imax, Jmax, Kmax = pr[ib][ic]
            c = dict(bctype=r.bctype, subbctype=r.subbctype, pertype=r.pertype,
                     point_start=[Imin, Jmin, Kmin], point_end=[Imax, Jmax, Kmax],
                     lcross=r.lcross, lreva=r.lrev_a, lrevb=r.lrev_b)
            c["rotate_angle"] = 0.0
            if r.pertype != 0:
                c["rotate_angle"] = float(npassage * r.pertype) * 2 * math.pi / float(nblade)
            # the face number fixes the direction (a one-cell-wide patch has two constant indices)
            d = (r.face + 1) // 2
            if [Imin, Jmin, Kmin][d - 1] != [Imax, Jmax, Kmax][d - 1]:
                raise ValueError("patch must be a constant-index face")
            c["ijkdirect"] = d
            mn = [Imin, Jmin, Kmin][d - 1]
            c["left_or_right"] = 1 if mn == 1 else 2
            dims = [b.ni, b.nj, b.nk]
            ijkadd = [0, 0, 0]; c1 = [0, 0, 0]; c2 = [0, 0, 0]; e1 = [0, 0, 0]; e2 = [0, 0, 0]
            if mn == 1:
                ijkadd[d - 1] = 1; c1[d - 1] = 1; c2[d - 1] = 2; e1[d - 1] = -2; e2[d - 1] = -3
            elif mn == dims[d - 1]:
                ijkadd[d - 1] = -1; c1[d - 1] = -2; c2[d - 1] = -3; e1[d - 1] = 1; e2[d - 1] = 2
            else:
                raise ValueError("patch not on a block face")
            # Fortran integer division truncates toward zero (PLOT3DReadInit.F90:299-300)
            c["ijkadd"] = ijkadd
            c["cijkadd"] = [int((a - 1) / 2) for a in ijkadd]
            c["ceijkadd"] = [int((-a - 1) / 2) for a in ijkadd]
            c["cijkadd1"], c["cijkadd2"] = c1, c2
            c["ceijkadd1"], c["ceijkadd2"] = e1, e2
            tr = [0, 0, 0]; tr[d - 1] = 1
            rev = [0, 0, 0]
            others = [x for x in (1, 2, 3) if x != d]
            rev[others[0] - 1] = r.lrev_a
            rev[others[1] - 1] = r.lrev_b
            c["transform"], c["rev_matrix"] = tr, rev
            c["transform_t"] = [0, 0, 0]
            c["target_block"] = (r.target_blk or (ib + 1)) - 1
            c["target_conn"] = max(r.target_conn - 1, 0)
            if r.bctype == 9:
                t = pr[c["target_block"]][c["target_conn"]]
                tt = [0, 0, 0]
                if t[0] == t[3]:
                    tt[0] = 1
                elif t[1] == t[4]:
                    tt[1] = 1
                elif t[2] == t[5]:
                    tt[2] = 1
                c["transform_t"] = tt
                c["tplijk"] = [abs(t[0] - t[3]) + 1, abs(t[1] - t[4]) + 1, abs(t[2] - t[5]) + 1]
            b.patches.append(c)


def ijk_start_end(i, j, k, tl, tt, lcross):
    """Geom/startup.F90:1-25 -- orientation map of a block interface."""
    ft1, ft2, ft3 = tl
    gt1, gt2, gt3 = tt
    lc1 = 1 - lcross
    temp1 = ft1 * i + ft2 * j + ft3 * k
    temp2 = ft1 * (j * lc1 + k * lcross) + ft2 * (i * lc1 + k * lcross) + ft3 * (i * lc1 + j * lcross)
    temp3 = ft1 * (j * lcross + k * lc1) + ft2 * (i * lcross + k * lc1) + ft3 * (i * lcross + j * lc1)
    return (gt1 * temp1 + (gt2 + gt3) * temp2,
            gt2 * temp1 + gt1 * temp2 + gt3 * temp3,
            gt3 * temp1 + (gt1 + gt2) * temp3)


def patch_structs(b: Block):
    arr = (capi.TransPatch * max(len(b.patches), 1))()
    for n, c in enumerate(b.patches):
        p = arr[n]
        p.bctype, p.subbctype, p.pertype = c["bctype"], c["subbctype"], c["pertype"]
        p.point_start = capi.I3(*c["point_start"])
        p.point_end = capi.I3(*c["point_end"])
        p.ijkdirect, p.left_or_right = c["ijkdirect"], c["left_or_right"]
        for nm in ("ijkadd", "cijkadd", "cijkadd1", "cijkadd2", "ceijkadd", "ceijkadd1",
                   "ceijkadd2", "transform", "transform_t", "rev_matrix"):
            setattr(p, nm, capi.I3(*c[nm]))
        p.target_block, p.target_conn = c["target_block"], c["target_conn"]
        p.lcross = c["lcross"]
        p.rotate_angle = c["rotate_angle"]
    return arr


# --------------------------------------------------------------------------------------
# metrics (Geom/CalcDst.F90:28-719)

def compute_geometry(blocks: list, tiny: float = 1e-20, device: bool = False):
    """Metrics of every block; device=True computes the O(N) metrics (trans_geom_metrics)
    and the wall distances (trans_geom_wall_distance) with the HIP library, bitwise the same
    as here (thtc up to the device atan2)."""
    for b in blocks:
        b.wall_distance_on_device = device
        if device:
            _metrics_device(b, tiny)
        else:
            _metrics(b, tiny)
    for b in blocks:
        _ghost_geometry_physical(b)
    _ghost_geometry_interfaces(blocks)
    for b in blocks:
        _markwall(b)


def _metrics(b: Block, tiny: float):
    NI, NJ, NK = b.ni, b.nj, b.nk
    NI1, NJ1, NK1 = NI - 1, NJ - 1, NK - 1
    X = np.stack([b.xx, b.yy, b.zz])  # (3, NI+5, NJ+5, NK+5)
    xc = b.alloc("xc"); yc = b.alloc("yc"); zc = b.alloc("zc")
    C = [xc, yc, zc]
    I, J, K = sl(1, NI1), sl(1, NJ1), sl(1, NK1)
    I1, J1, K1 = sl(2, NI), sl(2, NJ), sl(2, NK)
    for c in range(3):
        Xc = X[c]
        C[c][I, J, K] = (Xc[I1, J1, K1] + Xc[I1, J1, K] + Xc[I1, J, K1] + Xc[I1, J, K] +
                         Xc[I, J1, K1] + Xc[I, J1, K] + Xc[I, J, K1] + Xc[I, J, K]) / 8.0
    thtc = b.alloc("thtc"); rad = b.alloc("rad")
    thtc[I, J, K] = np.arctan2(zc[I, J, K], yc[I, J, K])
    rad[I, J, K] = np.sqrt(yc[I, J, K] ** 2 + zc[I, J, K] ** 2)

    sd = b.alloc("sd", 9); grad = b.alloc("grad", 3)
    # orientation check at (1,1,1) (CalcDst.F90:198-210)
    xv = np.array([C[c][3, 3, 3] - C[c][3 + 1, 3, 3] for c in range(3)]) * -1.0
    r1 = X[:, 1 + O, 2 + O, 2 + O] - X[:, 1 + O, 1 + O, 1 + O]
    r2 = X[:, 1 + O, 1 + O, 2 + O] - X[:, 1 + O, 2 + O, 1 + O]
    sv = np.cross(r1, r2) * 0.5
    dl = -1.0 if float(np.dot(xv, sv)) < 0 else 1.0

    def face_vec(dirn, R1, R2, tgt):
        S = np.empty((3,) + R1.shape[1:])
        S[0] = (R1[1] * R2[2] - R1[2] * R2[1]) * 0.5 * dl
        S[1] = (R1[2] * R2[0] - R1[0] * R2[2]) * 0.5 * dl
        S[2] = (R1[0] * R2[1] - R1[1] * R2[0]) * 0.5 * dl
        for x in range(3):
            sd[(dirn - 1) + 3 * x][tgt] = S[x]
        grad[dirn - 1][tgt] = np.sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + tiny)

    # i faces I=1..NI, J=1..NJ1, K=1..NK1
    Ia = sl(1, NI)
    face_vec(1, X[:, Ia, J1, K1] - X[:, Ia, J, K], X[:, Ia, J, K1] - X[:, Ia, J1, K], (Ia, J, K))
    Ja = sl(1, NJ)
    face_vec(2, X[:, I, Ja, K1] - X[:, I1, Ja, K], X[:, I1, Ja, K1] - X[:, I, Ja, K], (I, Ja, K))
    Ka = sl(1, NK)
    face_vec(3, X[:, I1, J, Ka] - X[:, I, J1, Ka], X[:, I1, J1, Ka] - X[:, I, J, Ka], (I, J, Ka))
    for x in range(3):
        sd[0 + 3 * x][O + 0, J, K] = sd[0 + 3 * x][O + 1, J, K]
        sd[1 + 3 * x][I, O + 0, K] = sd[1 + 3 * x][I, O + 1, K]
        sd[2 + 3 * x][I, J, O + 0] = sd[2 + 3 * x][I, J, O + 1]
    grad[0][O + 0, J, K] = grad[0][O + 1, J, K]
    grad[1][I, O + 0, K] = grad[1][I, O + 1, K]
    grad[2][I, J, O + 0] = grad[2][I, J, O + 1]

    # face-to-face lengths (CalcDst.F90:289-325)
    alagm = b.alloc("alagm", 3)

    def fc(i0, i1, j0, j1, k0, k1):
        return [0.25 * (X[c][i0, j0, k0] + X[c][i1, j1, k0] + X[c][i0, j0, k1] + X[c][i1, j1, k1])
                for c in range(3)]
    w = fc(I, I, J, J1, K, K1); e = fc(I1, I1, J, J1, K, K1)
    alagm[0][I, J, K] = np.sqrt(sum((w[c] - e[c]) ** 2 for c in range(3)))
    s = fc(I, I1, J, J, K, K1); n = fc(I, I1, J1, J1, K, K1)
    alagm[1][I, J, K] = np.sqrt(sum((s[c] - n[c]) ** 2 for c in range(3)))
    bb = fc(I, I1, J, J1, K, K); t = fc(I, I1, J, J1, K1, K1)
    alagm[2][I, J, K] = np.sqrt(sum((bb[c] - t[c]) ** 2 for c in range(3)))

    # boundary face centres stored at ghost index 0 / N (CalcDst.F90:332-382)
    for c in range(3):
        Xc = X[c]
        C[c][O + 0, J, K] = (Xc[O + 1, J, K] + Xc[O + 1, J1, K] + Xc[O + 1, J, K1] + Xc[O + 1, J1, K1]) / 4.0
        C[c][O + NI, J, K] = (Xc[O + NI, J, K] + Xc[O + NI, J1, K] + Xc[O + NI, J, K1] + Xc[O + NI, J1, K1]) / 4.0
        C[c][I, O + 0, K] = (Xc[I, O + 1, K] + Xc[I1, O + 1, K] + Xc[I, O + 1, K1] + Xc[I1, O + 1, K1]) / 4.0
        C[c][I, O + NJ, K] = (Xc[I, O + NJ, K] + Xc[I1, O + NJ, K] + Xc[I, O + NJ, K1] + Xc[I1, O + NJ, K1]) / 4.0
        C[c][I, J, O + 0] = (Xc[I, J, O + 1] + Xc[I1, J, O + 1] + Xc[I, J1, O + 1] + Xc[I1, J1, O + 1]) / 4.0
        C[c][I, J, O + NK] = (Xc[I, J, O + NK] + Xc[I1, J, O + NK] + Xc[I, J1, O + NK] + Xc[I1, J1, O + NK]) / 4.0

    # cell length scales (CalcDst.F90:430-476)
    aleng = b.alloc("aleng")
    cc = np.stack(C)

    def dd(di, dj, dk):
        sh = cc[:, sl(1 + di, NI1 + di), sl(1 + dj, NJ1 + dj), sl(1 + dk, NK1 + dk)] - cc[:, I, J, K]
        return np.sqrt((sh ** 2).sum(axis=0))
    dw, de, ds, dn, dlo, du = dd(-1, 0, 0), dd(1, 0, 0), dd(0, -1, 0), dd(0, 1, 0), dd(0, 0, -1), dd(0, 0, 1)
    cand = [dw, de, ds, dn, dlo, du]
    if NI1 == 2:
        cand = [ds, dn, dlo, du]
    if NJ1 == 2:
        cand = [dw, de, dlo, du]
    if NK1 == 2:
        cand = [dw, de, ds, dn]
    aleng[I, J, K] = np.maximum.reduce(cand)
    aseng = b.alloc("aseng")
    aseng[I, J, K] = np.minimum.reduce(cand)

    # volumes (CalcDst.F90:481-515)
    vol = b.alloc("vol")
    om = np.zeros((NI1, NJ1, NK1))
    XR = X[:, I1, J1, K1] - X[:, I, J1, K1]
    om += sum(XR[x] * sd[0 + 3 * x][I, J, K] for x in range(3))
    XR = X[:, I1, J1, K1] - X[:, I1, J, K1]
    om += sum(XR[x] * sd[1 + 3 * x][I, J, K] for x in range(3))
    XR = X[:, I1, J1, K1] - X[:, I1, J1, K]
    om += sum(XR[x] * sd[2 + 3 * x][I, J, K] for x in range(3))
    vol[I, J, K] = (1.0 / 3.0) * om
    if np.any(vol[I, J, K] < tiny):
        raise ValueError("negative cell volume")

    dst = b.alloc("dst", fill=0.0)
    _wall_distance(b)

    # first ghost layer copies (CalcDst.F90:579-608) and SD extension (:611-632)
    vol[O + 0, J, K] = vol[O + 1, J, K]; vol[O + NI, J, K] = vol[O + NI1, J, K]
    dst[O + 0, J, K] = dst[O + 1, J, K]; dst[O + NI, J, K] = dst[O + NI1, J, K]
    alagm[0][O + 0, J, K] = alagm[0][O + 1, J, K]; alagm[0][O + NI, J, K] = alagm[0][O + NI1, J, K]
    vol[I, O + 0, K] = vol[I, O + 1, K]; vol[I, O + NJ, K] = vol[I, O + NJ1, K]
    dst[I, O + 0, K] = dst[I, O + 1, K]; dst[I, O + NJ, K] = dst[I, O + NJ1, K]
    alagm[1][I, O + 0, K] = alagm[1][I, O + 1, K]; alagm[1][I, O + NJ, K] = alagm[1][I, O + NJ1, K]
    vol[I, J, O + 0] = vol[I, J, O + 1]; vol[I, J, O + NK] = vol[I, J, O + NK1]
    dst[I, J, O + 0] = dst[I, J, O + 1]; dst[I, J, O + NK] = dst[I, J, O + NK1]
    alagm[2][I, J, O + 0] = alagm[2][I, J, O + 1]; alagm[2][I, J, O + NK] = alagm[2][I, J, O + NK1]
    for x in range(3):
        sd[0 + 3 * x][O + 0, sl(1, NJ), sl(1, NK)] = sd[0 + 3 * x][O + 2, sl(1, NJ), sl(1, NK)]
        sd[0 + 3 * x][O + NI + 1, sl(1, NJ), sl(1, NK)] = sd[0 + 3 * x][O + NI1, sl(1, NJ), sl(1, NK)]
        sd[1 + 3 * x][sl(1, NI), O + 0, sl(1, NK)] = sd[1 + 3 * x][sl(1, NI), O + 2, sl(1, NK)]
        sd[1 + 3 * x][sl(1, NI), O + NJ + 1, sl(1, NK)] = sd[1 + 3 * x][sl(1, NI), O + NJ1, sl(1, NK)]
        sd[2 + 3 * x][sl(1, NI), sl(1, NJ), O + 0] = sd[2 + 3 * x][sl(1, NI), sl(1, NJ), O + 2]
        sd[2 + 3 * x][sl(1, NI), sl(1, NJ), O + NK + 1] = sd[2 + 3 * x][sl(1, NI), sl(1, NJ), O + NK1]


def _metrics_device(b: Block, tiny: float):
    """_metrics with the per-cell work on the device; the wall distance and its ghost copies
    follow as in _metrics."""
    import ctypes as C
    NI, NJ, NK = b.ni, b.nj, b.nk
    NI1, NJ1, NK1 = NI - 1, NJ - 1, NK - 1
    names = ("xc", "yc", "zc", "thtc", "rad", "aleng", "aseng", "vol", "sd", "grad", "alagm")
    ncomp = {"sd": 9, "grad": 3, "alagm": 3}
    arrs = {nm: b.alloc(nm, ncomp.get(nm, 1)) for nm in names}
    pts = [np.asfortranarray(a, dtype=np.float64) for a in (b.xx, b.yy, b.zz)]
    DP = C.POINTER(C.c_double)
    out = capi.TransGeomMetricsOut(**{nm: arrs[nm].ctypes.data_as(DP) for nm in names})
    lib = capi.load_library()
    rc = lib.trans_geom_metrics(NI, NJ, NK, *(p.ctypes.data_as(DP) for p in pts), tiny, C.byref(out))
    if rc == capi.TRANS_EINVAL:
        raise ValueError("negative cell volume")
    if rc != 0:
        raise RuntimeError(f"trans_geom_metrics failed ({rc})")
    dst = b.alloc("dst", fill=0.0)
    _wall_distance(b)
    J, K, I = sl(1, NJ1), sl(1, NK1), sl(1, NI1)
    dst[O + 0, J, K] = dst[O + 1, J, K]; dst[O + NI, J, K] = dst[O + NI1, J, K]
    dst[I, O + 0, K] = dst[I, O + 1, K]; dst[I, O + NJ, K] = dst[I, O + NJ1, K]
    dst[I, J, O + 0] = dst[I, J, O + 1]; dst[I, J, O + NK] = dst[I, J, O + NK1]


def _wall_panels(b: Block):
    """Centres of the wall faces of a block (the panels grd/Surf.dat would list)."""
    out = []
    X = (b.xx, b.yy, b.zz)
    for c in b.patches:
        if c["bctype"] != 1:
            continue
        d = c["ijkdirect"]
        ps, pe = c["point_start"], c["point_end"]
        idx = [sl(ps[0], pe[0]), sl(ps[1], pe[1]), sl(ps[2], pe[2])]
        # panel corners: along tangential dirs cells a..b use points a..b+1
        lo = [ps[0], ps[1], ps[2]]
        hi = [pe[0], pe[1], pe[2]]
        for t in range(3):
            if t != d - 1:
                hi[t] = pe[t] + 1
        cen = []
        for c3 in X:
            sub = c3[sl(lo[0], hi[0]), sl(lo[1], hi[1]), sl(lo[2], hi[2])]
            sub = np.squeeze(sub, axis=d - 1)
            cen.append(0.25 * (sub[:-1, :-1] + sub[1:, :-1] + sub[:-1, 1:] + sub[1:, 1:]))
        out.append(np.stack([q.ravel() for q in cen], axis=1))
        del idx
    return out


def _wall_distance(b: Block):
    """Minimum distance from cell centres to wall panel centres (CalcDst.F90:121-158)."""
    NI1, NJ1, NK1 = b.ni - 1, b.nj - 1, b.nk - 1
    I, J, K = sl(1, NI1), sl(1, NJ1), sl(1, NK1)
    panels = getattr(b, "all_wall_panels", None)
    dst = b.arrays["dst"]
    if panels is None or len(panels) == 0:
        dst[I, J, K] = 1.0e3  # no wall anywhere: distance is never used
        return
    pts = np.stack([b.arrays[n][I, J, K].ravel(order="F") for n in ("xc", "yc", "zc")], axis=1)
    if getattr(b, "wall_distance_on_device", False):
        dst[I, J, K] = wall_distance_device(pts, panels).reshape((NI1, NJ1, NK1), order="F")
        return
    best = np.full(pts.shape[0], 1.0e20)
    chunk = 4096
    for s in range(0, pts.shape[0], chunk):
        p = pts[s:s + chunk]
        for q in range(0, panels.shape[0], 8192):
            pan = panels[q:q + 8192]
            d2 = ((p[:, None, :] - pan[None, :, :]) ** 2).sum(axis=2).min(axis=1)
            best[s:s + chunk] = np.minimum(best[s:s + chunk], d2)
    dst[I, J, K] = np.sqrt(np.maximum(best, 1e-20)).reshape((NI1, NJ1, NK1), order="F")


def wall_distance_device(cells: np.ndarray, panels: np.ndarray) -> np.ndarray:
    """cells (n, 3), panels (m, 3) -> distance of each cell to the nearest panel centre,
    computed by the HIP library (raises if it is not built or the device fails)."""
    import ctypes as C
    lib = capi.load_library()
    c = np.ascontiguousarray(np.asarray(cells, dtype=np.float64).T)
    p = np.ascontiguousarray(np.asarray(panels, dtype=np.float64).T)
    out = np.empty(c.shape[1])
    DP = C.POINTER(C.c_double)
    rc = lib.trans_geom_wall_distance(c.ctypes.data_as(DP), c.shape[1], p.ctypes.data_as(DP), p.shape[1],
                                      out.ctypes.data_as(DP))
    if rc != 0:
        raise RuntimeError(f"trans_geom_wall_distance failed ({rc})")
    return out


def _layers(c):
    return [(c["cijkadd"], c["ceijkadd"]), (c["cijkadd1"], c["ceijkadd1"]), (c["cijkadd2"], c["ceijkadd2"])]


def _face_cells(c):
    ps, pe = c["point_start"], c["point_end"]
    for k in range(ps[2], pe[2] + 1):
        for j in range(ps[1], pe[1] + 1):
            for i in range(ps[0], pe[0] + 1):
                yield i, j, k


def _ghost_geometry_physical(b: Block):
    """Geom/startup.F90:396-508 (vectorised over the patch)."""
    A = b.arrays
    for c in b.patches:
        if c["bctype"] == 9:
            continue
        d = c["ijkdirect"]
        ps, pe = c["point_start"], c["point_end"]
        rng = [np.arange(ps[t], pe[t] + 1) for t in range(3)]
        ii, jj, kk = np.meshgrid(*rng, indexing="ij")

        def at(off):
            return (ii + off[0] + O, jj + off[1] + O, kk + off[2] + O)
        s0, s1 = at(c["cijkadd"]), at(c["cijkadd1"])
        e0, e1, e2 = at(c["ceijkadd"]), at(c["ceijkadd1"]), at(c["ceijkadd2"])
        for e in (e0, e1, e2):
            A["vol"][e] = A["vol"][s0]
            A["dst"][e] = A["dst"][s0]
            A["alagm"][d - 1][e] = A["alagm"][d - 1][s0]
        for nm in ("rad", "thtc"):
            a = A[nm]
            a[e0] = 2.0 * a[s0] - a[s1]
            a[e1] = 2.0 * a[e0] - a[s0]
            a[e2] = 2.0 * a[e1] - a[e0]
            if c["bctype"] == 5:
                a[e0] = a[s0]; a[e1] = a[s0]; a[e2] = a[s0]
        sa = c["ijkadd"]
        src = at(sa)
        dstf = (ii - sa[0] + O, jj - sa[1] + O, kk - sa[2] + O)
        here = at((0, 0, 0))
        for x in range(3):
            comp = A["sd"][(d - 1) + 3 * x]
            comp[dstf] = 2.0 * comp[here] - comp[src]


def _ghost_geometry_interfaces(blocks: list):
    """Geom/startup.F90:510-755 with the FP64 payload of the promoted build."""
    for b in blocks:
        for c in b.patches:
            if c["bctype"] != 9:
                continue
            t = blocks[c["target_block"]]
            tc = t.patches[c["target_conn"]]
            A, TA = b.arrays, t.arrays
            td = tc["ijkdirect"]
            tps = tc["point_start"]
            out_nt = float(2 * tc["left_or_right"] - 3)
            d = c["ijkdirect"]
            ps, pe = c["point_start"], c["point_end"]
            ii_add = pe[0] - ps[0] + 2
            jj_add = pe[1] - ps[1] + 2
            kk_add = pe[2] - ps[2] + 2
            rm = c["rev_matrix"]
            dcita = c["rotate_angle"]
            sint, cost = math.sin(dcita), math.cos(dcita)
            out_n = float(2 * c["left_or_right"] - 3)
            for (i, j, k) in _face_cells(c):
                iL, jL, kL = i - ps[0] + 1, j - ps[1] + 1, k - ps[2] + 1
                iL = rm[0] * ii_add + (1 - 2 * rm[0]) * iL
                jL = rm[1] * jj_add + (1 - 2 * rm[1]) * jL
                kL = rm[2] * kk_add + (1 - 2 * rm[2]) * kL
                Is, Js, Ks = ijk_start_end(iL, jL, kL, c["transform"], c["transform_t"], c["lcross"])
                # target cell (ibuf, jbuf, kbuf) in target numbering
                ti, tj, tk = Is + tps[0] - 1, Js + tps[1] - 1, Ks + tps[2] - 1
                for (sadd, _), (_, eadd) in zip(_layers(tc), _layers(c)):
                    sI = (ti + sadd[0] + O, tj + sadd[1] + O, tk + sadd[2] + O)
                    eI = (i + eadd[0] + O, j + eadd[1] + O, k + eadd[2] + O)
                    A["vol"][eI] = TA["vol"][sI]
                    A["dst"][eI] = TA["dst"][sI]
                    A["alagm"][d - 1][eI] = TA["alagm"][td - 1][sI]
                    A["rad"][eI] = TA["rad"][sI]
                    A["thtc"][eI] = TA["thtc"][sI] - dcita
                if c["pertype"] != 0:
                    ad, ad1, ad2 = c["cijkadd"], c["cijkadd1"], c["cijkadd2"]
                    e0 = (i + c["ceijkadd"][0] + O, j + c["ceijkadd"][1] + O, k + c["ceijkadd"][2] + O)
                    e1 = (i + c["ceijkadd1"][0] + O, j + c["ceijkadd1"][1] + O, k + c["ceijkadd1"][2] + O)
                    e2 = (i + c["ceijkadd2"][0] + O, j + c["ceijkadd2"][1] + O, k + c["ceijkadd2"][2] + O)
                    s1 = (i + ad[0] + O, j + ad[1] + O, k + ad[2] + O)
                    # (startup.F90:716) mixes the layer-1 i/k offsets with the layer-2 j offset
                    s2 = (i + ad1[0] + O, j + ad2[1] + O, k + ad1[2] + O)
                    for nm in ("thtc", "rad"):
                        a = A[nm]
                        a[e0] = 2.0 * a[s1] - a[s2]
                        a[e1] = 2.0 * a[e0] - a[s1]
                        a[e2] = 2.0 * a[e1] - a[e0]
                # face vector behind the interface
                tsa = tc["ijkadd"]
                tf = (ti + tsa[0] + O, tj + tsa[1] + O, tk + tsa[2] + O)
                sdv = [TA["sd"][(td - 1) + 3 * x][tf] * out_nt for x in range(3)]
                sa = c["ijkadd"]
                fI = (i - sa[0] + O, j - sa[1] + O, k - sa[2] + O)
                A["sd"][(d - 1) + 0][fI] = -out_n * sdv[0]
                SDy, SDz = -out_n * sdv[1], -out_n * sdv[2]
                A["sd"][(d - 1) + 3][fI] = cost * SDy + sint * SDz
                A["sd"][(d - 1) + 6][fI] = -sint * SDy + cost * SDz
                if c["pertype"] != 0:
                    src = (i + sa[0] + O, j + sa[1] + O, k + sa[2] + O)
                    here = (i + O, j + O, k + O)
                    for x in range(3):
                        comp = A["sd"][(d - 1) + 3 * x]
                        comp[fI] = 2.0 * comp[here] - comp[src]


def _markwall(b: Block):
    """MARKWALL arrays (Input/Read_BlockDimensions.F90:146-283) from the wall faces."""
    NI1, NJ1, NK1 = b.ni - 1, b.nj - 1, b.nk - 1
    shapes = [(NJ1, NK1), (NJ1, NK1), (NI1, NK1), (NI1, NK1), (NI1, NJ1), (NI1, NJ1)]
    mw = [np.zeros(s, dtype=np.int32, order="F") for s in shapes]
    for c in b.patches:
        if c["bctype"] != 1:
            continue
        d = c["ijkdirect"]
        f = 2 * (d - 1) + (c["left_or_right"] - 1)
        ps, pe = c["point_start"], c["point_end"]
        t = [x for x in range(3) if x != d - 1]
        mw[f][ps[t[0]] - 1:pe[t[0]], ps[t[1]] - 1:pe[t[1]]] = 1
    b.markwall = mw


def attach_wall_panels(blocks: list):
    """Every block's wall distance is measured to all walls of all blocks (Surf.dat)."""
    pans = []
    for b in blocks:
        pans += _wall_panels(b)
    allp = np.concatenate(pans, axis=0) if pans else np.zeros((0, 3))
    for b in blocks:
        b.all_wall_panels = allp


# --------------------------------------------------------------------------------------
# initial state (Input/Init.F90:94-169)

def initial_state(blocks: list, prm, q_dst: float = 0.0, if_ns_uniform: int = 1):
    """Uniform start with the reference's wall-distance k/omega profile.

    Ghost cells are filled from the nearest interior cell.  The reference leaves them
    at zero (Input/Read_BlockDimensions.F90:416), which its symmetry BC never repairs
    for the density of the second ghost layer (BC/BC_Symm.F90:68-75).
    """
    om = prm.omega1
    for b in blocks:
        NI1, NJ1, NK1 = b.ni - 1, b.nj - 1, b.nk - 1
        I, J, K = sl(1, NI1), sl(1, NJ1), sl(1, NK1)
        V = b.alloc("v", 7); PP = b.alloc("pp"); T = b.alloc("t"); RM = b.alloc("rmiu")
        yc, zc, rad, dst = b["yc"][I, J, K], b["zc"][I, J, K], b["rad"][I, J, K], b["dst"][I, J, K]
        V[0][I, J, K] = 1.0
        V[1][I, J, K] = 1.0
        V[2][I, J, K] = om * zc
        V[3][I, J, K] = -om * yc
        ptemp = 1.0 / (1.4 * prm.xm * prm.xm)
        PP[I, J, K] = ptemp
        T[I, J, K] = prm.rxm2 * ptemp
        RM[I, J, K] = prm.rmiuf
        if if_ns_uniform == 0 and q_dst > 0:
            coff = 4.6 / q_dst
            rat = 1.0 - np.exp(-coff * dst)
            V[1][I, J, K] = np.where(dst < q_dst, V[1][I, J, K] * rat, V[1][I, J, K])
        V[4][I, J, K] = (PP[I, J, K] / 0.4 + 0.5 * (V[1][I, J, K] ** 2 + V[2][I, J, K] ** 2 +
                                                   V[3][I, J, K] ** 2) - 0.5 * (om * rad) ** 2)
        if prm.kind_model in (1, 2):
            C1, C2 = 45.8, 1.68
            smax = C2 / (2.0 * C1)
            tmax = -C1 * smax * smax + C2 * smax
            akin = np.maximum(prm.rkf, -C1 * dst ** 2 + C2 * dst)
            v3d = 100.0 * akin / tmax
            aoin = np.maximum(-12444.0 * dst + 0.54, akin / v3d)
            V[5][I, J, K] = akin
            V[6][I, J, K] = aoin
        VV = V[1][I, J, K] ** 2 + V[2][I, J, K] ** 2 + V[3][I, J, K] ** 2
        PP[I, J, K] = (1.4 - 1.0) * V[4][I, J, K] - 0.5 * (1.4 - 1.0) * (
            VV / V[0][I, J, K] - (om * rad) ** 2 * V[0][I, J, K])
        T[I, J, K] = prm.rxm2 * PP[I, J, K] / V[0][I, J, K]
        # ghosts: nearest interior cell
        ix = np.clip(np.arange(-2, b.ni + 3), 1, NI1) + O
        jx = np.clip(np.arange(-2, b.nj + 3), 1, NJ1) + O
        kx = np.clip(np.arange(-2, b.nk + 3), 1, NK1) + O
        for a in (V, PP, T, RM):
            a[...] = a[..., ix, :, :][..., :, jx, :][..., :, :, kx]
        b.alloc("wt0", 7)[...] = V
        b.alloc("wt1", 7)[...] = V
        b.alloc("wt2", 7)


# --------------------------------------------------------------------------------------
# views for the C-ABI

def geom_view(b: Block) -> capi.TransGeomView:
    g = capi.TransGeomView()
    A = b.arrays
    for nm in ("sd", "grad", "alagm", "vol", "rad", "dst", "aleng", "aseng", "yc", "zc", "thtc"):
        setattr(g, nm, capi.fview(A[nm]))
    import ctypes as C
    arr = (C.POINTER(C.c_int) * 6)()
    for f in range(6):
        arr[f] = b.markwall[f].ctypes.data_as(C.POINTER(C.c_int))
    g.markwall = arr
    return g


def state_view(b: Block) -> capi.TransStateView:
    s = capi.TransStateView()
    A = b.arrays
    for nm in ("v", "pp", "t", "rmiu", "wt0", "wt1", "wt2"):
        setattr(s, nm, capi.fview(A[nm]))
    return s
