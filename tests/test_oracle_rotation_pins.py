"""Pins of the oracle's rotating-frame terms and of its block-interface maps for every
orientation and for the periodic (rotated) interface.  Like test_oracle_pins, each check
comes from the mathematics or from the grid the blocks were cut from, not from the oracle's
own formulas:

* Roe flux with the rothalpy term (Space/CONVECT_Roe_Original.F90:79-80,283-301);
* Coriolis and centrifugal source (Space/DIFFUSION_Centered.F90:811-822);
* Update_V energy with -rho (omega r)^2 / 2 (Time/TIME_Update_V.F90:622-688);
* wall ghost velocity SubBCType * omega x r (BC/BC_Wall.F90:41-90);
* inlet ghost without absolute swirl (BC/BC_Inlet.F90:38-107);
* outlet radial equilibrium of a rigid rotation (BC/BC_Outlet.F90:49-70);
* BC_P2P with IJKStartEnd for all 48 orientations of the donor block (BC/BC_P2P.F90:76-139,
  Geom/startup.F90:1-25, Input/PLOT3DReadInit.F90:183-348);
* the periodic interface: a swirl-free field and the rotation of the gradient tensor
  (BC/BC_P2P.F90:86-95,140-181).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle import OracleCase, lib
from trans_amd import cases, solver
from trans_amd.mesh import O, BCRow, Block, sl
from trans_amd.params import GlobalInfo, OrgDeck, make_params

import interface_cases as ic

RPM = 2400.0


def _rbox(ni=7, nj=6, nk=5, bcs=None, omega_rpm=RPM, **deck):
    """Box off the rotation axis: x in [0, 0.6], y in [0.5, 1.1] (geometric), z in [0.2, 0.5]."""
    x1 = np.linspace(0.0, 0.6, ni)
    y1 = 0.5 + np.concatenate([[0.0], np.cumsum(0.1 * 1.1 ** np.arange(nj - 1))])
    z1 = np.linspace(0.2, 0.5, nk)
    if bcs is None:
        bcs = [BCRow(1, 0, 0, 0, 0, 2), BCRow(2, 0, 0, 0, 0, 2), BCRow(3, 0, 0, 0, 0, 3),
               BCRow(4, 0, 0, 0, 0, 3), BCRow(5, 0, 0, 0, 0, 3), BCRow(6, 0, 0, 0, 0, 3)]
    b = Block(ni, nj, nk, bcs)
    X, Y, Z = np.meshgrid(x1, y1, z1, indexing="ij")
    cases._points(b, X, Y, Z)
    org = OrgDeck(**{**dict(XM=0.5, Tinf=288.0, Kind_Spatial=2, Kind_Roe=1, IF_Euler=1), **deck})
    p_inf = 1.2 * 287.0 * 288.0
    stag = 1.0 + 0.2 * 0.25
    glob = GlobalInfo(nblocks=1, P0in=p_inf * stag ** 3.5, T0in=288.0 * stag, Pouthub=p_inf,
                      omega_rpm=(omega_rpm, 0.0, 0.0))
    prm = solver.prepare_case([b], org, glob)
    return b, prm, (x1, y1, z1)


def _edge_centres(e):
    return 0.5 * (e[:-1] + e[1:])


def test_omega_from_rpm():
    """Input/ParameterSet.F90:87: omega = rpm 2 pi Lref / (60 Vref), Vref = M sqrt(1.4 R Tinf)."""
    b, prm, _ = _rbox()
    vref = 0.5 * math.sqrt(1.4 * 287.0 * 288.0)
    assert prm.omega1 == pytest.approx(RPM * 2.0 * math.pi / 60.0 / vref, rel=1e-15)
    assert prm.omega1 > 0.1


# ---------------------------------------------------------------------------------------
# Roe flux with the rothalpy term

STATES = [(1.1, 0.8, 0.1, -0.05, 1 / (1.4 * 0.25)), (0.7, -0.2, 0.4, 0.3, 2.0)]
FACES = [np.array([0.3, 0.1, 0.05]), np.array([-0.02, 0.5, 0.0])]


def _roe(prm, L_, R_, S, rl, rr):
    out = (C.c_double * 5)()
    sav = float(np.sqrt((S ** 2).sum()))
    (rhoL, uL, vL, wL, pL), (rhoR, uR, vR, wR, pR) = L_, R_
    TL, TR = pL / (prm.ppf * rhoL), pR / (prm.ppf * rhoR)
    lib().ora_roe_common(C.byref(prm), pL, pR, TL, TR, uL, uR, vL, vR, wL, wR, *S, sav, rl, rr, out)
    return np.array(out[:])


@pytest.mark.parametrize("st", STATES)
@pytest.mark.parametrize("S", FACES)
@pytest.mark.parametrize("wr", [0.0, 0.35, 1.7])
def test_roe_rothalpy_consistency(st, S, wr):
    """Equal states at radius r (omega r = wr): the flux is the physical flux whose energy
    carries the rothalpy h + |w|^2/2 - (omega r)^2/2."""
    prm = make_params(OrgDeck(XM=0.5), GlobalInfo())
    rho, u, v, w, p = st
    n = S / np.linalg.norm(S)
    un = u * n[0] + v * n[1] + w * n[2]
    want = np.array([rho * un, rho * u * un + p * n[0], rho * v * un + p * n[1], rho * w * un + p * n[2],
                     rho * un * (1.4 * p / 0.4 / rho + 0.5 * (u * u + v * v + w * w) - 0.5 * wr * wr)])
    np.testing.assert_allclose(_roe(prm, st, st, S, wr, wr), -want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("S", FACES)
def test_roe_rothalpy_shifts_energy_by_mass_flux(S):
    """Different states of equal density at one radius: the rothalpy term shifts only the
    energy flux, by -(omega r)^2 / 2 times the mass flux (the speed of sound of the Roe
    average, am2 = (g-1)|h - |u|^2/2 + (omega r)^2/2|, does not move)."""
    prm = make_params(OrgDeck(XM=0.5), GlobalInfo())
    L_ = (1.05, 0.9, 0.1, -0.2, 2.7)
    R_ = (1.05, 0.6, -0.15, 0.05, 3.1)
    f0 = _roe(prm, L_, R_, S, 0.0, 0.0)
    for wr in (0.4, 1.3):
        f = _roe(prm, L_, R_, S, wr, wr)
        # (the Roe average's speed of sound moves by rounding only: (omega r)^2/2 enters and leaves)
        np.testing.assert_allclose(f[:4], f0[:4], rtol=1e-13, atol=1e-14)
        assert f[4] == pytest.approx(f0[4] - 0.5 * wr * wr * f0[0], rel=1e-12, abs=1e-13)


def test_convect_roe_passes_omega_times_face_radius():
    """CONVECT_Roe_Original.F90:79-80: the flux routine gets rLo = rccl * omega: a uniform
    relative flow on the box gives a momentum residual of zero and an energy residual equal to
    the change of (omega r)^2 / 2 times the mass flux through each cell's faces."""
    b, prm, edges = _rbox()
    rho, u, v, w, p = 1.0, 0.7, 0.05, -0.1, 1 / (1.4 * 0.25)
    V = b["v"]
    V[0], V[1], V[2], V[3] = rho, rho * u, rho * v, rho * w
    b["pp"][...] = p
    b["t"][...] = prm.rxm2 * p / rho
    oc = OracleCase([b], prm)
    oc.call("ora_spaceinterp", 0, 0)
    oc.call("ora_convect_muscl", 0)
    oc[0]["D"][...] = 0.0
    oc.call("ora_convect_roe", 0)
    D = oc[0]["D"]
    I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
    for L in range(4):
        assert np.max(np.abs(D[L][I, J, K])) < 1e-12
    # energy: with equal states on both sides the dissipation vanishes and the face's energy
    # flux carries -(omega rL)^2/4 - (omega rR)^2/4 per unit mass flux, rL, rR the reconstructed
    # face radii; the enthalpy part cancels over each closed cell
    om = prm.omega1
    rl, rr = oc[0]["rccl"], oc[0]["rccr"]
    sd = b["sd"]
    want = np.zeros_like(D[4][I, J, K])
    for d in range(3):
        di, dj, dk = (d == 0), (d == 1), (d == 2)
        S = [sd[d + 3 * x] for x in range(3)]
        mflux = rho * (u * S[0] + v * S[1] + w * S[2])  # through face (i, j, k) of direction d
        e = 0.25 * om * om * (rl[d] ** 2 + rr[d] ** 2) * mflux
        lo = e[I, J, K]
        hi = e[sl(1 + di, b.ni - 1 + di), sl(1 + dj, b.nj - 1 + dj), sl(1 + dk, b.nk - 1 + dk)]
        want += hi - lo
    assert np.max(np.abs(want)) > 1e-6
    np.testing.assert_allclose(D[4][I, J, K], want, rtol=1e-9, atol=1e-13)


# ---------------------------------------------------------------------------------------
# Coriolis and centrifugal source

def test_rotating_source_closed_form():
    """Uniform relative flow: every diffusive flux vanishes and DIFFUSION_Centered adds
    rho (2 omega w + omega^2 y) Vol to the y momentum and rho (-2 omega v + omega^2 z) Vol to
    the z momentum (the relative-frame Coriolis and centrifugal forces for rotation about x)."""
    b, prm, edges = _rbox(IF_Euler=0)
    rho, u, v, w = 1.2, 0.6, 0.25, -0.15
    V = b["v"]
    V[0], V[1], V[2], V[3] = rho, rho * u, rho * v, rho * w
    b["rmiu"][...] = 0.0
    oc = OracleCase([b], prm)
    oc.call("ora_spaceinterp", 0, 0)
    oc.call("ora_diff_dqdxyz", 0, 0)
    oc[0]["D"][...] = 0.0
    oc.call("ora_diffusion", 0)
    D = oc[0]["D"]
    I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
    om = prm.omega1
    yc = _edge_centres(edges[1])[None, :, None]
    zc = _edge_centres(edges[2])[None, None, :]
    vol = b["vol"][I, J, K]
    for L in (0, 1, 4):
        assert np.max(np.abs(D[L][I, J, K])) < 1e-14
    np.testing.assert_allclose(D[2][I, J, K], rho * (2 * om * w + om * om * yc) * vol, rtol=1e-12)
    np.testing.assert_allclose(D[3][I, J, K], rho * (-2 * om * v + om * om * zc) * vol, rtol=1e-12)


# ---------------------------------------------------------------------------------------
# Update_V at omega != 0

@pytest.mark.parametrize("turb", [0, 1])
def test_update_v_rotating_energy(turb):
    """Time/TIME_Update_V.F90:622-688: the increments of rho, u and p do not involve omega
    (the primitive pressure increment is that of the non-rotating energy), and the new energy
    is p / 0.4 + rho |w|^2 / 2 - rho (omega rad)^2 / 2; the k-omega update (which recovers p
    from rho E with the same term, :749-750) then equals the non-rotating one."""
    deck = dict(IF_Euler=0, IF_Turb_Org=1, Kind_Model=1) if turb else {}
    runs = []
    for rpm in (0.0, RPM):
        b, prm, _ = _rbox(omega_rpm=rpm, **deck)
        rng = np.random.Generator(np.random.MT19937(5))
        oc = OracleCase([b], prm)
        V = oc[0]["v"]
        rad = b["rad"]
        I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
        # the same primitive state for both runs
        rho = 1.0 + 0.05 * rng.uniform(-1, 1, rad.shape)
        uvw = [0.6 + 0.1 * rng.uniform(-1, 1, rad.shape) for _ in range(3)]
        p = 2.8 * (1 + 0.05 * rng.uniform(-1, 1, rad.shape))
        V[0] = rho
        for c in range(3):
            V[1 + c] = rho * uvw[c]
        V[4] = p / 0.4 + 0.5 * rho * sum(q * q for q in uvw) - 0.5 * (prm.omega1 * rad) ** 2 * rho
        oc[0]["pp"][...] = p
        oc[0]["t"][...] = prm.rxm2 * p / rho
        if turb:
            oc[0]["DQdxyz"][:9] = 0.3 * rng.uniform(-1, 1, oc[0]["DQdxyz"][:9].shape)
        D = oc[0]["D"]
        D[:7, I, J, K] = 2e-3 * rng.uniform(-1, 1, D[:7, I, J, K].shape)
        oc.call("ora_update_v", 0, 1, turb)
        runs.append((oc, prm, b))
    (o0, p0, b0), (o1, p1, b1) = runs
    I, J, K = sl(1, b1.ni - 1), sl(1, b1.nj - 1), sl(1, b1.nk - 1)
    for c in range(4):
        np.testing.assert_array_equal(o1[0]["v"][c][I, J, K], o0[0]["v"][c][I, J, K])
    np.testing.assert_array_equal(o1[0]["pp"][I, J, K], o0[0]["pp"][I, J, K])
    np.testing.assert_array_equal(o1[0]["t"][I, J, K], o0[0]["t"][I, J, K])
    V = o1[0]["v"]
    rho = V[0][I, J, K]
    ke = 0.5 * (V[1] ** 2 + V[2] ** 2 + V[3] ** 2)[I, J, K] / rho
    wr2 = (p1.omega1 * b1["rad"][I, J, K]) ** 2
    np.testing.assert_allclose(V[4][I, J, K], o1[0]["pp"][I, J, K] / 0.4 + ke - 0.5 * wr2 * rho, rtol=1e-13)
    np.testing.assert_allclose(V[4][I, J, K] - o0[0]["v"][4][I, J, K], -0.5 * wr2 * rho, rtol=1e-9)
    if turb:
        for c in (5, 6):
            np.testing.assert_allclose(V[c][I, J, K], o0[0]["v"][c][I, J, K], rtol=1e-12)
        np.testing.assert_allclose(o1[0]["rmiu"][I, J, K], o0[0]["rmiu"][I, J, K], rtol=1e-12)


# ---------------------------------------------------------------------------------------
# wall, inlet and outlet ghosts in the rotating frame

def test_rotating_wall_ghost_velocity():
    """BC/BC_Wall.F90:41-90: a SubBCType-1 wall (stationary in the absolute frame) gives its
    three ghost layers the relative velocity of the wall, (0, omega z, -omega y) at the point
    halfway between the first cell's centre and the face centre (the layer-0 ghost holds the
    face centre); a SubBCType-0 wall (rotating with the frame) gives zero."""
    bcs = [BCRow(1, 0, 0, 0, 0, 2), BCRow(2, 0, 0, 0, 0, 2), BCRow(3, 0, 0, 0, 0, 1, 0),
           BCRow(4, 0, 0, 0, 0, 1, 1), BCRow(5, 0, 0, 0, 0, 3), BCRow(6, 0, 0, 0, 0, 3)]
    b, prm, edges = _rbox(bcs=bcs, IF_Euler=0)
    oc = OracleCase([b], prm)
    oc.bc(1, -1)
    V = oc[0]["v"]
    om = prm.omega1
    I, K = sl(1, b.ni - 1), sl(1, b.nk - 1)
    zc = _edge_centres(edges[2])[None, :]
    y_last = _edge_centres(edges[1])[-1]
    y_mid = 0.5 * (y_last + edges[1][-1])            # jmax: first cell NJ1 and the face
    for L in (0, 1, 2):
        g = O + b.nj + L
        rho = V[0][I, g, K]
        np.testing.assert_allclose(V[1][I, g, K], 0.0, atol=0)
        np.testing.assert_allclose(V[2][I, g, K] / rho, np.broadcast_to(om * zc, rho.shape), rtol=1e-13)
        np.testing.assert_allclose(V[3][I, g, K] / rho, -om * y_mid * np.ones_like(rho), rtol=1e-13)
        g0 = O - L  # jmin, SubBCType 0
        for c in (1, 2, 3):
            assert np.all(V[c][I, g0, K] == 0.0)


def test_inlet_ghost_has_no_absolute_swirl():
    """BC/BC_Inlet.F90:40-96: the ghost's relative velocity is the axial isentropic speed plus
    (0, omega z, -omega y) of the first interior cell, and its energy carries -rho (omega rad)^2/2
    with the ghost's radius; layers 1 and 2 copy layer 0."""
    bcs = [BCRow(1, 0, 0, 0, 0, 5), BCRow(2, 0, 0, 0, 0, 7), BCRow(3, 0, 0, 0, 0, 3),
           BCRow(4, 0, 0, 0, 0, 3), BCRow(5, 0, 0, 0, 0, 3), BCRow(6, 0, 0, 0, 0, 3)]
    b, prm, edges = _rbox(bcs=bcs)
    # interior at the freestream static pressure: ghost Mach 0.5, rho 1, |u_abs| = 1
    oc = OracleCase([b], prm)
    oc[0]["pp"][...] = 1 / (1.4 * 0.25)
    oc.bc(1, -1)
    V = oc[0]["v"]
    om = prm.omega1
    J, K = sl(1, b.nj - 1), sl(1, b.nk - 1)
    yc = _edge_centres(edges[1])[:, None]
    zc = _edge_centres(edges[2])[None, :]
    rad_g = b["rad"][O + 0, J, K]
    np.testing.assert_allclose(rad_g, np.sqrt(yc ** 2 + zc ** 2), rtol=1e-14)  # ghost keeps the cell's radius
    for g in (O + 0, O - 1, O - 2):
        rho = V[0][g, J, K]
        np.testing.assert_allclose(rho, 1.0, rtol=1e-12)
        np.testing.assert_allclose(V[1][g, J, K], 1.0, rtol=1e-12)
        np.testing.assert_allclose(V[2][g, J, K], om * zc * np.ones_like(rho), rtol=1e-12)
        np.testing.assert_allclose(V[3][g, J, K], -om * yc * np.ones_like(rho), rtol=1e-12)
        ke = 0.5 * (1.0 + (om * zc) ** 2 + (om * yc) ** 2)
        np.testing.assert_allclose(V[4][g, J, K], 1 / (1.4 * 0.25) / 0.4 + ke - 0.5 * (om * rad_g) ** 2,
                                   rtol=1e-12)


def test_outlet_radial_equilibrium_of_rigid_rotation():
    """BC/BC_Outlet.F90:49-70: with an absolute swirl of rigid rotation Omega r and uniform T,
    the ghost pressure marches from Pouthub at the patch's first j as
    p_j = p_(j-1) (1 + Omega^2 gamma M^2 / T r_(j-1) (r_j - r_(j-1))), with the ghosts' radii
    (the interior's, extrapolated along i) -- the discrete radial equilibrium dp/dr = rho V_t^2/r
    with the density of the previous ghost."""
    bcs = [BCRow(1, 0, 0, 0, 0, 2), BCRow(2, 0, 0, 0, 0, 7), BCRow(3, 0, 0, 0, 0, 3),
           BCRow(4, 0, 0, 0, 0, 3), BCRow(5, 0, 0, 0, 0, 3), BCRow(6, 0, 0, 0, 0, 3)]
    b, prm, edges = _rbox(bcs=bcs, nj=9)
    om = prm.omega1
    Om = 0.45  # absolute swirl rate
    yc, zc = b["yc"], b["zc"]
    V = b["v"]
    T0 = 1.3
    # relative velocity = absolute (0, -Om z, Om y) minus frame (0, -om z, om y)
    V[0] = 1.0
    V[1] = 0.4
    V[2] = -(Om - om) * zc
    V[3] = (Om - om) * yc
    b["t"][...] = T0
    oc = OracleCase([b], prm)
    oc.bc(1, -1)
    P = oc[0]["pp"]
    kappa = Om * Om * prm.rxm2 / T0
    g = O + b.ni  # first ghost layer
    r = np.sqrt(_edge_centres(edges[1])[:, None] ** 2 + _edge_centres(edges[2])[None, :] ** 2)
    want = np.empty_like(r)
    want[0] = prm.pouthub
    for j in range(1, r.shape[0]):
        want[j] = want[j - 1] * (1.0 + kappa * r[j - 1] * (r[j] - r[j - 1]))
    got = P[g, sl(1, b.nj - 1), sl(1, b.nk - 1)]
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert np.all(np.diff(got, axis=0) > 0)
    for L in (1, 2):  # outer layers copy the first
        np.testing.assert_array_equal(P[g + L, sl(1, b.nj - 1), sl(1, b.nk - 1)], got)


# ---------------------------------------------------------------------------------------
# BC_P2P + IJKStartEnd for every orientation of the donor block

def _orient_case(perm, flips):
    blocks, info = ic.orientation_pair(perm, flips)
    org = OrgDeck(XM=0.5, Tinf=288.0, Kind_Spatial=2, Kind_Roe=1, IF_Euler=1)
    prm = solver.prepare_case(blocks, org, GlobalInfo(nblocks=2))
    for n, b in enumerate(blocks):
        V = b["v"]
        V[0] = 1.0 + n
        V[1], V[2], V[3] = b["xc"], b["yc"], b["zc"]
    return blocks, info, prm


def _seed_rds(oc, blocks, info):
    """Rds of each interior cell tagged with the global axis (1, 2, 3) of its local direction."""
    perm = info["perm"]
    for n, b in enumerate(blocks):
        for d in range(3):
            oc[n]["Rds"][d] = float(d + 1) if n == 0 else float(perm[d] + 1)


@pytest.mark.parametrize("perm,flips", ic.ORIENTATIONS)
def test_interface_halo_all_orientations(perm, flips):
    """For each of the 48 ways to store the donor block (axis permutation and reversals), the
    deck rows a deck author writes from the geometry (lcross when my first tangential axis runs
    along the neighbour's second; lrev where the senses differ) make every ghost of all three
    layers, on both sides, hold the cell of the global grid it overlaps, and the layer-0 spectral
    radii arrive in the receiver's directions."""
    blocks, info, prm = _orient_case(perm, flips)
    oc = OracleCase(blocks, prm)
    _seed_rds(oc, blocks, info)
    oc.bc(1, -1)
    exp = ic.expected_ghosts(blocks, info)
    X, Y, Z = info["X"], info["Y"], info["Z"]
    nchk = 0
    for n in (0, 1):
        V = oc[n]["v"]
        for (c, g, L) in exp[n]:
            idx = tuple(x + O for x in c)
            want = ic.global_centres(X, Y, Z, *g)
            got = [V[1][idx], V[2][idx], V[3][idx]]
            assert np.allclose(got, want, rtol=0, atol=1e-14), (n, c, g, L, got, want)
            if L == 1:
                # the receiver's direction a gets the donor's value along the same global axis
                rds = [oc[n]["Rds"][d][idx] for d in range(3)]
                want_r = [1.0, 2.0, 3.0] if n == 0 else [float(info["perm"][d] + 1) for d in range(3)]
                assert rds == want_r, (n, c, rds, want_r)
            nchk += 1
    assert nchk == 3 * (ic.NY * ic.NZ) * 2


def test_orientation_rows_cover_every_combination():
    """The 48 orientations use all eight (lcross, lrevA, lrevB) rows on both sides."""
    seen_r, seen_t = set(), set()
    for perm, flips in ic.ORIENTATIONS:
        _, info = ic.orientation_pair(perm, flips)
        seen_r.add(info["rows"]["R"])
        seen_t.add(info["rows"]["T"])
    assert len(seen_r) == 8 and len(seen_t) == 8


def test_interface_ghost_geometry_all_orientations():
    """GeomTransform1 (Geom/startup.F90:510-755) through the same maps: every interface ghost of
    every orientation gets the volume of the global cell it overlaps."""
    for perm, flips in ic.ORIENTATIONS[::5]:
        blocks, info, prm = _orient_case(perm, flips)
        exp = ic.expected_ghosts(blocks, info)
        X, Y, Z = info["X"], info["Y"], info["Z"]
        ref = Block(ic.NR + ic.NT + 1, ic.NY + 1, ic.NZ + 1,
                    [BCRow(f, 0, 0, 0, 0, 3) for f in range(1, 7)])
        cases._points(ref, X, Y, Z)
        solver.prepare_case([ref], OrgDeck(XM=0.5), GlobalInfo(nblocks=1))
        for n, b in enumerate(blocks):
            for (c, g, L) in exp[n]:
                idx = tuple(x + O for x in c)
                gidx = tuple(x + 1 + O for x in g)
                assert b["vol"][idx] == pytest.approx(ref["vol"][gidx], rel=1e-12), (perm, flips, n, c)


# ---------------------------------------------------------------------------------------
# periodic interface of an annular sector

def _sector(rpm=RPM, nsplit_k=2):
    blocks, point_fn, nblade = ic.annular_sector(nsplit_k=nsplit_k)
    org = OrgDeck(XM=0.5, Tinf=288.0, Kind_Spatial=2, Kind_Roe=1, IF_Euler=0, IF_Turb_Org=0)
    glob = GlobalInfo(nblocks=len(blocks), nblade=nblade, Npassage=1, omega_rpm=(rpm, 0.0, 0.0),
                      P0in=101325.0, T0in=288.0, Pouthub=101325.0)
    prm = solver.prepare_case(blocks, org, glob)
    return blocks, point_fn, prm


def test_periodic_rotate_angle_from_deck():
    """Input/PLOT3DReadInit.F90:155-160: rotateAngle = Npassage * PerType * 2 pi / nblade."""
    blocks, _, _ = _sector()
    pa = [c["rotate_angle"] for c in blocks[0].patches if c["pertype"]]
    pb = [c["rotate_angle"] for c in blocks[-1].patches if c["pertype"]]
    assert pa == [pytest.approx(2 * math.pi / 36)] and pb == [pytest.approx(-2 * math.pi / 36)]


def test_periodic_halo_of_swirl_free_field():
    """A field without absolute swirl (relative velocity -omega x r, axisymmetric pressure and
    density) is the same in every passage, so the periodic ghosts must hold it at their own
    positions: relative velocity (u, omega z_g, -omega y_g) with (y_g, z_g) the ghost cell's centre
    in the sector's periodic continuation -- the donor's (v, w) turned by the pitch."""
    blocks, point_fn, prm = _sector()
    om = prm.omega1
    for b in blocks:
        V = b["v"]
        rho = 1.0 + 0.3 * b["rad"]
        V[0] = rho
        V[1] = rho * (0.5 + b["xc"])
        V[2] = rho * om * b["zc"]
        V[3] = -rho * om * b["yc"]
        b["pp"][...] = 2.0 + b["rad"] ** 2
    oc = OracleCase(blocks, prm)
    oc.bc(1, -1)
    nk_tot = sum(b.nk - 1 for b in blocks)
    checked = 0
    for n, b in enumerate(blocks):
        V = oc[n]["v"]
        for side in (0, 1):
            if (n == 0 and side == 1) or (n == len(blocks) - 1 and side == 0) or (0 < n < len(blocks) - 1):
                continue
            for L in (1, 2, 3):
                kloc = 1 - L if side == 0 else b.nk - 1 + L
                kg = b.k0 + kloc - 1  # 0-based global cell index, may leave 0..nk_tot-1
                assert kg < 0 or kg >= nk_tot
                for j in range(1, b.nj):
                    for i in range(1, b.ni):
                        xg, yg, zg = ic.sector_cell_centre(point_fn, i - 1, j - 1, kg)
                        idx = (i + O, j + O, kloc + O)
                        rho = V[0][idx]
                        assert rho == pytest.approx(1.0 + 0.3 * math.hypot(yg, zg), rel=1e-12)
                        assert V[1][idx] / rho == pytest.approx(0.5 + xg, rel=1e-12)
                        assert V[2][idx] / rho == pytest.approx(om * zg, rel=1e-11, abs=1e-13)
                        assert V[3][idx] / rho == pytest.approx(-om * yg, rel=1e-11, abs=1e-13)
                        assert oc[n]["pp"][idx] == pytest.approx(2.0 + yg * yg + zg * zg, rel=1e-12)
                        checked += 1
    assert checked == 2 * 3 * (blocks[0].ni - 1) * (blocks[0].nj - 1)


def _rot(a):
    """Rotation of the (y, z) components that BC_P2P applies for rotateAngle a."""
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, s], [0, -s, c]])


def test_periodic_gradient_rotation():
    """BC/BC_P2P.F90:140-181 (sweep 0): the ghost's velocity gradient is the donor's turned by
    the pitch, G' = R G R^T, and the gradients of T, k and omega are turned as vectors, R g --
    except one entry: the reference writes dw/dz as cos^2 wz + sin^2 vy + sin cos (vz + wy), where
    R G R^T has -sin cos (vz + wy) (a sign slip in BC_P2P.F90:176).  Reproduced as written."""
    blocks, point_fn, prm = _sector()
    rng = np.random.Generator(np.random.MT19937(3))
    oc = OracleCase(blocks, prm)
    src = []
    for n in range(len(blocks)):
        DQ = oc[n]["DQdxyz"]
        DQ[...] = rng.uniform(-1, 1, DQ.shape)
        src.append(DQ.copy())
    oc.bc(1, 0)
    b0, bl = blocks[0], blocks[-1]
    cases_chk = [(0, 0, bl, len(blocks) - 1, blocks[0].patches[4]["rotate_angle"]),
                 (len(blocks) - 1, 1, b0, 0, blocks[-1].patches[5]["rotate_angle"])]
    slip = 0.0
    for n, side, donor, nd, ang in cases_chk:
        b = blocks[n]
        R = _rot(ang)
        for j in range(1, b.nj):
            for i in range(1, b.ni):
                kr = 0 if side == 0 else b.nk
                kd = donor.nk - 1 if side == 0 else 1
                G = src[nd][:9, i + O, j + O, kd + O].reshape(3, 3)  # row: u, v, w; column: x, y, z
                got = oc[n]["DQdxyz"][:, i + O, j + O, kr + O]
                want = R @ G @ R.T
                s, c = math.sin(ang), math.cos(ang)
                quirk = want.copy()
                quirk[2, 2] += 2 * s * c * (G[1, 2] + G[2, 1])
                np.testing.assert_allclose(got[:9].reshape(3, 3), quirk, rtol=1e-13, atol=1e-14)
                slip = max(slip, abs(got[8] - want[2, 2]))
                for q in range(3):  # grad T, grad k, grad omega
                    g = src[nd][9 + 3 * q:12 + 3 * q, i + O, j + O, kd + O]
                    np.testing.assert_allclose(got[9 + 3 * q:12 + 3 * q], R @ g, rtol=1e-13, atol=1e-14)
    assert slip > 1e-3  # the test tells the written entry from the rotated one


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_plane_totals_absolute_frame(direction):
    """Swirl-free absolute flow (U, 0, 0) seen in the rotating frame, V_rel = (U, omega z,
    -omega y) from the cell centres: through any plane the absolute-frame totals are those of
    the axial speed alone, p0 = p (1 + 0.2 U^2 / a^2)^3.5, T0 = T (1 + 0.2 U^2 / a^2), while
    the relative-frame totals carry the blade speed (omega r)^2 of every face.  The face
    centre is the same crossed volume weighting as the face state, which reproduces a linear
    field at the face, so the absolute swirl cancels to rounding.  At omega = 0 both frames
    give the same bits."""
    b, prm, edges = _rbox()
    om = prm.omega1
    rho, U, p = 1.1, 0.6, 1.0 / (1.4 * 0.25)
    yc = _edge_centres(edges[1])[None, :, None] * np.ones((b.ni - 1, 1, b.nk - 1))
    zc = _edge_centres(edges[2])[None, None, :] * np.ones((b.ni - 1, b.nj - 1, 1))
    I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
    V = b["v"]
    V[...] = 0.0
    V[0][...] = rho
    V[0][I, J, K] = rho
    vr, wr = om * zc, -om * yc
    V[1][...] = rho * U
    V[2][I, J, K] = rho * vr
    V[3][I, J, K] = rho * wr
    b["pp"][...] = p
    b["t"][...] = prm.rxm2 * p / rho
    oc = OracleCase([b], prm)
    idx = 3
    r = oc.plane_integrals(0, direction, idx)
    a2 = 1.4 * p / rho
    if direction == 1:  # the axial planes carry the mass flow
        assert r["p0_abs_mass_avg"] == pytest.approx(p * (1 + 0.2 * U * U / a2) ** 3.5, rel=1e-13)
        assert r["t0_abs_mass_avg"] == pytest.approx(prm.rxm2 * p / rho * (1 + 0.2 * U * U / a2), rel=1e-13)
        fac = 1 + 0.2 * (U * U + vr[0] ** 2 + wr[0] ** 2) / a2
        area = np.diff(edges[1])[:, None] * np.diff(edges[2])[None, :]  # dm = rho U dA
        want = np.sum(p * fac ** 3.5 * area) / np.sum(area)
        assert r["p0_mass_avg"] == pytest.approx(want, rel=1e-12)
        assert r["p0_mass_avg"] > r["p0_abs_mass_avg"] * (1 + 1e-3)
    b0, prm0, _ = _rbox(omega_rpm=0.0)
    for nm in ("v", "pp", "t"):
        b0[nm][...] = b[nm]
    r0 = OracleCase([b0], prm0).plane_integrals(0, direction, idx)
    assert r0["p0_abs_mass_avg"] == r0["p0_mass_avg"] and r0["t0_abs_mass_avg"] == r0["t0_mass_avg"]
