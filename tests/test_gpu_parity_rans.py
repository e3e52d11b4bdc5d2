"""GPU parity of the RANS-SST path against the CPU oracle.

Case: a small turbulent flat plate of three blocks of uneven i-extent (block interfaces
through BC_P2P, total-pressure inlet, static-pressure outlet, a wall with a symmetry run-in,
far field at jmax, symmetry side planes), with MDCD reconstruction + AUSMPW+ and the k-omega
SST model, perturbed by a seeded +-1 % field (k, omega and mu_t included) so that every
limiter and switch is exercised.

Bars: ghost layers within 1e-14 per component (last-bit pow/sin/cos differences); gradients,
shock sensor and residual D(1:7) from identical inputs within 1e-12 relative (max-norm per
component; the device evaluates pow/tanh/exp in its own math library, whose last-bit
differences from glibc stay below 1e-14 relative in one evaluation); one LU-SGS + update
within 1e-12; ten iterations within 1e-6 (the same last-bit differences amplified through
the limiters of the update, as for the inviscid path).
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase
from trans_amd import capi, cases, solver
from trans_amd.solver import Solver

from helpers import interior, perturb, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _plate(spatial=3, roe=3, seed=20261202, flim2=0.0, **kw):
    args = dict(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    args.update(kw)
    blocks, org, glob = cases.flat_plate(**args)
    org.Kind_Spatial = spatial
    org.Kind_Roe = roe
    org.F_limiter2 = flim2
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=seed, turb=True)
    return blocks, prm


def _percomp(g, o):
    """max |g - o| / max |o| per component (or overall for a scalar field)."""
    if g.ndim == 4:
        return np.array([np.max(np.abs(g[c] - o[c])) / max(np.max(np.abs(o[c])), 1e-300)
                         for c in range(g.shape[0])])
    return np.max(np.abs(g - o)) / max(np.max(np.abs(o)), 1e-300)


def test_bc_interfaces_and_ghosts():
    """All ghost layers after BC(-1), per component.  Interface copies, walls, symmetry and
    the outlet are plain data movement and arithmetic; the far-field and inlet states call
    pow/sin/cos, whose device results may differ from glibc in the last bit, hence 1e-14 of
    each component's magnitude."""
    print("case: flat plate 3 blocks 48x32x3 cells, BC(-1) all ghosts", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.spectral_radius(1)
    sv.spectral_radius()
    oc.bc(1, -1)
    sv.bc(-1)
    for n in range(len(blocks)):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = _percomp(st[nm], oc[n][nm])
            assert np.all(e <= 1e-14), (n, nm, e)
    sv.close()


def _post_bc_case(spatial, roe, flim2=0.0):
    """Oracle spectral radius + BC(-1) on the perturbed case; the blocks then carry the
    oracle's ghost layers, so that device and oracle residuals start from identical bits."""
    blocks, prm = _plate(spatial, roe, flim2=flim2)
    if roe == 2:
        # eddy viscosities over six decades put Fscheme (the rotated Roe blending) across
        # its whole range instead of at 1
        rng = np.random.Generator(np.random.MT19937(53))
        for b in blocks:
            b["rmiu"][...] = prm.ref * 10.0 ** rng.uniform(-8, -2, b["rmiu"].shape)
    oc = OracleCase(blocks, prm)
    oc.spectral_radius(1)
    oc.bc(1, -1)
    for n, b in enumerate(blocks):
        for nm in ("v", "pp", "t", "rmiu"):
            b[nm][...] = oc[n][nm]
    return blocks, prm, oc


@pytest.mark.parametrize("spatial,roe,flim2", [(3, 3, 0.0), (2, 1, 0.0), (3, 1, 0.0), (2, 3, 0.0), (1, 1, 0.0),
                                               (1, 3, 0.0), (4, 3, 0.0), (4, 3, 2.0), (4, 1, 2.0), (2, 2, 0.0),
                                               (3, 2, 0.0)])
def test_single_residual(spatial, roe, flim2):
    """One residual evaluation (gradients, shock sensor, BC(0), viscous and k-omega diffusion,
    reconstruction, inviscid and k-omega fluxes) from identical post-BC states.

    Identical inputs matter for MUSCL: ALIMITER switches off when its arguments change sign
    (factor |sgn a + sgn b| / 2), so a last-bit difference in a ghost value from a far-field
    pow() can move a face state by 2/3 of the other argument.  The BC test above bounds
    those ghost differences separately."""
    print(f"case: flat plate, residual, Kind_Spatial={spatial} Kind_Roe={roe} F_limiter2={flim2}", flush=True)
    blocks, prm, oc = _post_bc_case(spatial, roe, flim2)
    sv = Solver(blocks, prm, ki=1)
    sv.spectral_radius()
    sv.space()
    oc.residual(1)  # BC(-1) is idempotent on an already filled halo
    for n, b in enumerate(blocks):
        dq = sv.get_field(n, capi.FIELD_DQDXYZ)
        e = relerr(interior(dq, b), interior(oc[n]["DQdxyz"], b))
        assert np.all(e < TOL), (n, "DQdxyz", e)
        # ghost layers written by BC(0)
        e = _percomp(dq, oc[n]["DQdxyz"])
        assert np.all(e < TOL), (n, "DQdxyz ghosts", e)
        sh = sv.get_field(n, capi.FIELD_SHOCK)
        assert relerr(interior(sh, b), interior(oc[n]["shock"], b)) < TOL
        if roe == 2:
            fs = sv.get_field(n, capi.FIELD_FSCHEME)
            e = relerr(interior(fs, b), interior(oc[n]["Fscheme"], b))
            print(n, "Fscheme relerr", e, "range", interior(fs, b).min(), interior(fs, b).max())
            assert e < TOL, (n, "Fscheme", e)
        d = sv.get_field(n, capi.FIELD_D)
        e = relerr(interior(d, b), interior(oc[n]["D"], b))
        print(n, "D relerr", e)
        assert np.all(e < TOL), (n, "D", e)
    sv.close()


def test_composite_residual_mdcd():
    """BC(-1) + residual on the device against the oracle for the configured scheme
    (MDCD + AUSMPW+), whose smooth nonlinear weights keep last-bit ghost differences at the
    last-bit level."""
    print("case: flat plate, composite residual MDCD + AUSMPW+", flush=True)
    blocks, prm = _plate(3, 3)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    for n, b in enumerate(blocks):
        d = sv.get_field(n, capi.FIELD_D)
        e = relerr(interior(d, b), interior(oc[n]["D"], b))
        print(n, "D relerr", e)
        assert np.all(e < TOL), (n, "D", e)
    sv.close()


def test_lusgs_update_and_norms():
    print("case: flat plate, residual + LU-SGS (mean + k-omega) + update", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    ores = oc.lusgs_update(1)
    gres = sv.lusgs_update()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            print(n, nm, e)
            assert np.all(np.asarray(e) < TOL), (n, nm)
        for key in ("drsa", "dua", "drsk", "duk", "drso", "duo"):
            assert abs(gres[n][key] - ores[n][key]) <= TOL * abs(ores[n][key]), (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)
    sv.close()


def test_ten_iterations():
    print("case: flat plate, 10 iterations", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.iterate(1, 10)
    gres = sv.iterate(10)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        e = relerr(interior(st["v"], b), interior(oc[n]["v"], b))
        print(n, "state relerr after 10 iterations", e)
        assert np.all(e < 1e-6)
        for key in ("drsa", "drsk", "drso"):
            assert abs(gres[n][key] - ores[n][key]) <= 1e-6 * ores[n][key], (n, key)
    sv.close()


def test_fp32_implicit_operator():
    """FP32 LU-SGS sweeps (residual and update stay FP64): after one update the state
    differs from the FP64 oracle by no more than 1e-5 of each component's increment scale
    -- the sweep's rounding is ~1e-7 relative per operation and does not compound beyond
    the ~100 steps of one wavefront."""
    print("case: flat plate, FP32 implicit operator", flush=True)
    blocks, prm = _plate()
    oc = OracleCase(blocks, prm)
    v0 = [b["v"].copy() for b in blocks]
    p32 = capi.TransParams.from_buffer_copy(prm)
    p32.implicit_fp32 = 1
    sv = Solver(blocks, p32, ki=1)
    oc.residual(1)
    sv.residual()
    oc.lusgs_update(1)
    sv.lusgs_update()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        dv_o = interior(oc[n]["v"], b) - interior(v0[n], b)
        dv_g = interior(st["v"], b) - interior(v0[n], b)
        e = relerr(dv_g, dv_o)
        print(n, "FP32 increment relerr", e)
        assert np.all(e < 1e-5)
    sv.close()


@pytest.mark.parametrize("rk", [0, 1])
def test_wdp_kw_model(rk):
    """WD+ k-omega (Kind_Model = 2): unblended diffusion sigmas, WD+ sources with the
    vortex-stretching function, and the WD+ limiter of the LU-SGS or RK update, one update
    from identical inputs at 1e-12 with norms and argmax locations exact."""
    print(f"case: flat plate, WD+ k-omega, IF_RK={rk}", flush=True)
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    org.Kind_Model = 2
    if rk:
        org.IF_RK, org.CFL_Sub_start, org.CFL_Sub_end = 1, 0.8, 0.8
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=11, turb=True)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_field(n, capi.FIELD_D), b), interior(oc[n]["D"], b))
        assert np.all(e < TOL), (n, "D", e)
    if rk:
        ores, gres = oc.rk_update(1, 1), sv.rk_update(1)
    else:
        ores, gres = oc.lusgs_update(1), sv.lusgs_update()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < TOL), (n, nm, e)
        for key in ("drsa", "dua", "drsk", "duk", "drso", "duo"):
            assert abs(gres[n][key] - ores[n][key]) <= TOL * abs(ores[n][key]), (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)
    sv.close()


@pytest.mark.parametrize("sub,rk", [(0, 0), (1, 0), (2, 0), (3, 1), (5, 0), (5, 1)])
def test_des_kw_model(sub, rk):
    """SST-DES/DDES (Kind_Hybrid = 3, Kind_SubHY = sub): the k destruction multiplied by
    Fdes, including the reference's carry of Vomega and arg1 from the previous cell of its
    loop (the oracle carries them through its sequential loop; the device recomputes them
    from the previous cell).  The DES length scale aLeng is shrunk so that Fdes > 1 in most
    cells.  Kind_SubHY = 5 is IDDES (fd, fe, the min/max cell lengths).  One LU-SGS update
    or RK stage from identical inputs at 1e-12."""
    print(f"case: flat plate, DES Kind_SubHY={sub}, IF_RK={rk}", flush=True)
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    org.Kind_Hybrid, org.Kind_SubHY = 3, sub
    if rk:
        org.IF_RK, org.CFL_Sub_start, org.CFL_Sub_end = 1, 0.8, 0.8
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=13, turb=True)
    for b in blocks:
        b["aleng"][...] *= 1e-9
        b["aseng"][...] *= 1e-9
    oc = OracleCase(blocks, prm)
    # the multiplier is active: the RANS sources differ
    p0 = capi.TransParams.from_buffer_copy(prm)
    p0.kind_hybrid = 0
    o0 = OracleCase(blocks, p0)
    oc.residual(1)
    o0.residual(1)
    sv = Solver(blocks, prm, ki=1)
    sv.residual()
    if rk:
        ores, gres = oc.rk_update(1, 1), sv.rk_update(1)
        o0.rk_update(1, 1)
    else:
        ores, gres = oc.lusgs_update(1), sv.lusgs_update()
        o0.lusgs_update(1)
    changed = np.mean([np.mean(interior(oc[n]["v"], b)[5] != interior(o0[n]["v"], b)[5])
                       for n, b in enumerate(blocks)])
    print("fraction of cells whose k differs from RANS", changed, flush=True)
    assert changed > 0.5
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < TOL), (n, nm, e)
        for key in ("drsa", "dua", "drsk", "duk", "drso", "duo"):
            assert abs(gres[n][key] - ores[n][key]) <= TOL * abs(ores[n][key]), (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)
    sv.close()


@pytest.mark.parametrize("hybrid,sub", [(0, 0), (3, 1), (3, 5)])
def test_suzen_compressibility(hybrid, sub):
    """IF_Comppress = 1 (Suzen's correction with the turbulent Mach number), with RANS,
    DDES (Kind_SubHY = 1) or IDDES (5): one LU-SGS update from identical inputs at 1e-12."""
    print(f"case: flat plate, IF_Comppress=1, Kind_Hybrid={hybrid}, Kind_SubHY={sub}", flush=True)
    blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
    org.IF_Comppress, org.Kind_Hybrid, org.Kind_SubHY = 1, hybrid, sub
    prm = solver.prepare_case(blocks, org, glob)
    perturb(blocks, prm, seed=17, turb=True)
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, prm, ki=1)
    oc.residual(1)
    sv.residual()
    ores, gres = oc.lusgs_update(1), sv.lusgs_update()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        for nm in ("v", "pp", "t", "rmiu"):
            e = relerr(interior(st[nm], b), interior(oc[n][nm], b))
            assert np.all(np.asarray(e) < TOL), (n, nm, e)
        for key in ("drsa", "dua", "drsk", "duk", "drso", "duo"):
            assert abs(gres[n][key] - ores[n][key]) <= TOL * abs(ores[n][key]), (n, key)
        for key in ("ima", "jma", "kma", "imk", "jmk", "kmk", "imo", "jmo", "kmo"):
            assert gres[n][key] == ores[n][key], (n, key)
    sv.close()


@pytest.mark.parametrize("spatial,roe,rk,fp32,nit", [(2, 1, 0, 0, 1), (2, 2, 0, 0, 1), (3, 3, 0, 0, 4),
                                                     (2, 1, 1, 0, 0), (2, 1, 0, 1, 0)])
def test_low_mach_preconditioning(spatial, roe, rk, fp32, nit):
    """IF_PRECONDITION = 1 at M = 0.1, CFL 0.3: the preconditioned spectral radius and Roe
    eigenvalues (original and rotated), the transforms of the LU-SGS right-hand side and
    neighbour terms (hyperplane sweep) and the conservative update.  Residual, Rds and Dtm, and
    one LU-SGS update from identical inputs at 1e-12 (FP32 sweeps: the increment within 1e-5);
    then `nit` whole iterations from the start within 1e-8.  The reference's preconditioned
    LU-SGS diverges on this case after 3 (Roe) or more (AUSMPW+) iterations -- its PrecondDW
    is not the identity at (U_r/c)^2 = 1 -- and explicit RK with the preconditioned time step
    diverges within its first iteration, so the RK case stops at the residual.  On the way
    the Roe cases amplify rounding ~1e4-1e6 per iteration (the oracle built with FMA
    contraction differs from itself by 1e-7 after 2 iterations, and the rotated flux's
    velocity-jump switch flips), so they are iterated once; AUSMPW+ stays within 1e-11
    over 4 iterations."""
    print(f"case: flat plate M=0.1, IF_Precondition=1, Kind_Spatial={spatial}, Kind_Roe={roe}, IF_RK={rk}, "
          f"fp32={fp32}", flush=True)

    def case():
        blocks, org, glob = cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25, mach=0.1)
        org.IF_Precondition, org.Kind_Spatial, org.Kind_Roe = 1, spatial, roe
        org.CFL_Sub_start = org.CFL_Sub_end = 0.3
        if rk:
            org.IF_RK, org.CFL_Sub_start, org.CFL_Sub_end = 1, 0.8, 0.8
        prm = solver.prepare_case(blocks, org, glob)
        perturb(blocks, prm, seed=20261202, turb=True)
        p_dev = capi.TransParams.from_buffer_copy(prm)
        p_dev.implicit_fp32 = fp32
        return blocks, prm, p_dev

    blocks, prm, p_dev = case()
    oc = OracleCase(blocks, prm)
    v0 = [b["v"].copy() for b in blocks]
    sv = Solver(blocks, p_dev, ki=1)
    oc.residual(1)
    sv.residual()
    for n, b in enumerate(blocks):
        for nm, fld in (("D", capi.FIELD_D), ("Rds", capi.FIELD_RDS), ("Dtm", capi.FIELD_DTM)):
            e = np.asarray(relerr(interior(sv.get_field(n, fld), b), interior(oc[n][nm], b)))
            print(n, nm, e)
            # D(6:7) (not preconditioned) in the first cells off the inlet is a small difference
            # of large k-omega fluxes through ghosts set with pow/exp: last-bit differences there
            # reach ~1e-11 of max|D| at M = 0.1
            bar = np.array([TOL] * 5 + [1e-9] * 2) if nm == "D" else TOL
            assert np.all(e < bar), (n, nm, e)
    if rk:
        sv.close()
        return
    ores, gres = oc.lusgs_update(1), sv.lusgs_update()
    for n, b in enumerate(blocks):
        st = sv.get_state(n)
        if fp32:
            dv_o = interior(oc[n]["v"], b) - interior(v0[n], b)
            e = relerr(interior(st["v"], b) - interior(v0[n], b), dv_o)
            print(n, "FP32 increment relerr", e)
            assert np.all(e < 1e-5), (n, e)
            continue
        # the k-omega part inherits the inlet-cell D(6:7) differences (bar 1e-9, as above)
        for nm, bar in (("v", np.array([TOL] * 5 + [1e-9] * 2)), ("pp", TOL), ("t", TOL), ("rmiu", 1e-9)):
            e = np.asarray(relerr(interior(st[nm], b), interior(oc[n][nm], b)))
            print(n, nm, e)
            assert np.all(e < bar), (n, nm, e)
        for key, bar in (("drsa", TOL), ("dua", TOL), ("drsk", 1e-9), ("duk", 1e-9), ("drso", 1e-9), ("duo", 1e-9)):
            assert abs(gres[n][key] - ores[n][key]) <= bar * abs(ores[n][key]), (n, key)
    sv.close()
    if not nit:
        return
    blocks, prm, p_dev = case()
    oc = OracleCase(blocks, prm)
    sv = Solver(blocks, p_dev, ki=1)
    oc.iterate(1, nit)
    gres = sv.iterate(nit)
    ores = oc.resid()
    for n, b in enumerate(blocks):
        e = relerr(interior(sv.get_state(n)["v"], b), interior(oc[n]["v"], b))
        print(n, f"state relerr after {nit} iterations", e)
        assert np.all(e < 1e-8), (n, e)
        assert abs(gres[n]["drsa"] - ores[n]["drsa"]) <= 1e-8 * ores[n]["drsa"], n
    sv.close()
