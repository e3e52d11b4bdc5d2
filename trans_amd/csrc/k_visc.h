// Viscous / RANS-SST residual kernels: Green-Gauss gradients with the shock sensor,
// viscous and k-omega diffusion face fluxes, MDCD / MUSCL reconstruction with the AUSMPW+
// or Roe flux and the k-omega upwind flux, and the ordered owner-computes accumulation.
// Face averages of SPACEINTERP_GetSurfCenterVars are recomputed where they are needed
// instead of being stored (the reference keeps 48 doubles per cell of VL/VR).
#pragma once
#include "dev_common.h"
#include "k_space.h"

// Cell value set used by face averages: u, v, w, T, k/rho, omega/rho.
struct CellQ {
    double u, v, w, t, k, o;
};

__device__ __forceinline__ CellQ load_q(const Fields& f, long c, bool turb)
{
    const long tot = f.tot;
    CellQ q;
    const double den = f.V[c];
    q.u = f.V[tot + c] / den;
    q.v = f.V[2 * tot + c] / den;
    q.w = f.V[3 * tot + c] / den;
    q.t = f.T[c];
    if (turb) {
        q.k = f.V[5 * tot + c] / den;
        q.o = f.V[6 * tot + c] / den;
    } else {
        q.k = 0.0;
        q.o = 0.0;
    }
    return q;
}

// The same values from Q, where the residual formed them once per cell before the gradient
// pass (k_prim_cells for the interior cells, k_prim_ghost for the ghost layers): the same
// divisions, so the same bits.
__device__ __forceinline__ CellQ load_qc(const Fields& f, long c, bool turb)
{
    const long tot = f.tot;
    CellQ q;
    q.u = f.Q[c];
    q.v = f.Q[tot + c];
    q.w = f.Q[2 * tot + c];
    q.t = f.T[c];
    if (turb) {
        q.k = f.Q[3 * tot + c];
        q.o = f.Q[4 * tot + c];
    } else {
        q.k = 0.0;
        q.o = 0.0;
    }
    return q;
}

__device__ __forceinline__ void store_q(const Fields& f, long c, const CellQ& q, bool turb)
{
    const long tot = f.tot;
    f.Q[c] = q.u;
    f.Q[tot + c] = q.v;
    f.Q[2 * tot + c] = q.w;
    if (turb) {
        f.Q[3 * tot + c] = q.k;
        f.Q[4 * tot + c] = q.o;
    }
}

// Q of the interior cells.
__global__ void __launch_bounds__(256) k_prim_cells(int if_turb, Fields f, const BlkInfo* __restrict__ blks,
                                                    const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k);
    store_q(f, c, load_q(f, c, if_turb != 0), if_turb != 0);
}

// Q of the three ghost layers beyond each face of a block (the cells the face stencils reach
// outside the interior), after BC(-1) has filled them.  Index space per block: direction d,
// side, layer, then the face's interior cells with the lower tangential index fastest.
__device__ __forceinline__ long ghost_slab_cells(const BlkInfo& b)
{
    const long NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    return 6 * (NJ1 * NK1 + NI1 * NK1 + NI1 * NJ1);
}

__global__ void __launch_bounds__(256) k_prim_ghost(int if_turb, Fields f, const BlkInfo* __restrict__ blks,
                                                    const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    // 32-bit index arithmetic: a block's ghost slab count fits (checked at creation)
    int lin = s.base + (int)threadIdx.x;
    if (lin >= (int)ghost_slab_cells(b)) return;
    const int area[3] = {NJ1 * NK1, NI1 * NK1, NI1 * NJ1};
    int d = 0;
    while (lin >= 6 * area[d]) {
        lin -= 6 * area[d];
        d++;
    }
    const int side = lin / (3 * area[d]);
    const int layer = (lin / area[d]) % 3;
    const int t = lin % area[d];
    const int n[3] = {NI1, NJ1, NK1};
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;  // tangential axes, ta fastest
    int ijk[3];
    ijk[ta] = 1 + t % n[ta];
    ijk[tb] = 1 + t / n[ta];
    ijk[d] = side ? n[d] + 1 + layer : -layer;
    const long c = cidx(b, ijk[0], ijk[1], ijk[2]);
    store_q(f, c, load_q(f, c, if_turb != 0), if_turb != 0);
}

// face value between cell m (below) and p (above) in direction d (SPACEINTERP)
__device__ __forceinline__ CellQ face_q(const Fields& f, int d, long cm, long cp, const CellQ& qm,
                                        const CellQ& qp, bool turb)
{
    const double* Al = f.Alagm + (long)d * f.tot;
    const double am = Al[cm], ap = Al[cp];
    CellQ r;
    r.u = face_avg(am, ap, qm.u, qp.u);
    r.v = face_avg(am, ap, qm.v, qp.v);
    r.w = face_avg(am, ap, qm.w, qp.w);
    r.t = face_avg(am, ap, qm.t, qp.t);
    if (turb) {
        r.k = face_avg(am, ap, qm.k, qp.k);
        r.o = face_avg(am, ap, qm.o, qp.o);
    } else {
        r.k = 0.0;
        r.o = 0.0;
    }
    return r;
}

// ---------------------------------------------------------------------------------------
// Space/DIFF_DqDxyz.F90:25-78, :232-265 (interior cells) and the shock sensor :300-384.
// The ghost-layer gradients of the reference are overwritten by BC(0) and are not formed.
template <bool FSCHEME>
__global__ void __launch_bounds__(256) k_gradients(trans_params p, int if_turb, Fields f,
                                                   const BlkInfo* __restrict__ blks,
                                                   const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k);
    const long tot = f.tot;
    const bool turb = if_turb != 0;
    const long st[3] = {1, (long)b.nx, (long)b.nx * b.ny};
    const double volinv = 1.0 / f.Vol[c];
    // Green-Gauss sums fp_i*S_i+ - fm_i*S_i- over i = 1, 2, 3, added left to right as the
    // reference writes them, in two passes of three variables (u, v, w, then T, k/rho,
    // omega/rho) so that nine sums and one direction's face values are live at a time.
    // Variable Q component qc of cell c+o: Q(qc), or T for qc < 0.
    auto qv = [&](int qc, long cc) { return qc < 0 ? f.T[cc] : f.Q[(long)qc * tot + cc]; };
    auto gg3 = [&](const int qc[3], int nv, double g[9]) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double* Al = f.Alagm + (long)d * tot;
            const double am = Al[c - st[d]], a0 = Al[c], ap = Al[c + st[d]];
            double fm[3], fp[3];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                if (v >= nv) break;
                const double q0 = qv(qc[v], c);
                fm[v] = face_avg(am, a0, qv(qc[v], c - st[d]), q0);
                fp[v] = face_avg(a0, ap, q0, qv(qc[v], c + st[d]));
            }
#pragma unroll
            for (int x = 0; x < 3; x++) {
                const double sp = f.SD[(long)(3 * x + d) * tot + c + st[d]], sm = f.SD[(long)(3 * x + d) * tot + c];
#pragma unroll
                for (int v = 0; v < 3; v++) {
                    if (v >= nv) break;
                    g[3 * v + x] = d == 0 ? fp[v] * sp - fm[v] * sm : g[3 * v + x] + fp[v] * sp - fm[v] * sm;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 9; q++)
            if (q < 3 * nv) g[q] = volinv * g[q];
    };
    // u, v, w: their gradients, the shock sensor and the strain and vorticity magnitudes
    double gu[9];
    {
        const int qc[3] = {0, 1, 2};
        gg3(qc, 3, gu);
    }
#pragma unroll
    for (int q = 0; q < 9; q++) f.DQ[(long)q * tot + c] = gu[q];
    const double DuDx = gu[0], DuDy = gu[1], DuDz = gu[2], DvDx = gu[3], DvDy = gu[4], DvDz = gu[5];
    const double DwDx = gu[6], DwDy = gu[7], DwDz = gu[8];
    const double W12 = (DuDy - DvDx) / 2., W13 = (DuDz - DwDx) / 2., W23 = (DvDz - DwDy) / 2.;
    const double Vort = sqrt(2. * (2. * W12 * W12 + 2. * W13 * W13 + 2. * W23 * W23));
    const double DivU = sqrt(DuDx * DuDx + DvDy * DvDy + DwDz * DwDz);
    f.shock[c] = DivU * DivU / fmax(DivU * DivU + Vort * Vort, 1.0);
    double Skl = 0.0;
    if (FSCHEME) {
        const double S11 = (DuDx + DuDx) / 2., S22 = (DvDy + DvDy) / 2., S33 = (DwDz + DwDz) / 2.;
        const double S12 = (DuDy + DvDx) / 2., S13 = (DuDz + DwDx) / 2., S23 = (DwDy + DvDz) / 2.;
        Skl = sqrt(2. * (1. * S11 * S11 + 1. * S22 * S22 + 1. * S33 * S33 + 2. * S12 * S12 + 2. * S13 * S13 +
                         2. * S23 * S23));
    }
    // T (and k/rho, omega/rho)
    double gt[9];
    {
        const int qc[3] = {-1, 3, 4};
        gg3(qc, turb ? 3 : 1, gt);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) f.DQ[(long)(9 + q) * tot + c] = gt[q];
    if (turb)
#pragma unroll
        for (int q = 3; q < 9; q++) f.DQ[(long)(9 + q) * tot + c] = gt[q];
    if (FSCHEME) {
        // DES blending Fscheme (Space/DIFF_DqDxyz.F90:295-381; CH1 = 3, CH2 = 1, CH3 = 2,
        // Cdes_epsi = 0.61, Cdes_omeg = 0.78, cmu = 0.09), read by the rotated Roe flux only
        const double Den = f.V[c], RmiuT = f.Rmiu[c];
        const double RMIUL = sutherland(p, f.T[c]);
        const double vK = f.V[5 * tot + c] / Den, vOmega = f.V[6 * tot + c] / Den;
        const double DLTxyz = f.aLeng[c];
        const double tau = 1.0;
        double gko[6];
#pragma unroll
        for (int q = 0; q < 6; q++) gko[q] = turb ? gt[3 + q] : f.DQ[(long)(12 + q) * tot + c];
        const double Dist = f.Dst[c];
        const double dTKO = gko[0] * gko[3] + gko[1] * gko[4] + gko[2] * gko[5];
        const double CDKO = fmax(2.0 * Den * p.sigo2 / vOmega * (dTKO), 1.0e-10);
        double arg1 = fmax(sqrt(vK) / p.beta_star / vOmega / Dist / p.ref,
                           500.0 * RMIUL / Den / Dist / Dist / vOmega / p.ref / p.ref);
        arg1 = fmin(arg1, 4.0 * Den * p.sigo2 * vK / CDKO / Dist / Dist);
        const double a2 = arg1 * arg1;
        const double FunC1 = tanh(a2 * a2);
        const double aLgrd = ((1. - FunC1) * 0.61 + FunC1 * 0.78) * DLTxyz;
        const double B0 = 2. * Vort * fmax(Vort, Skl) / fmax((Vort * Vort + Skl * Skl) / 2., p.alscale * 0.1);
        const double g0 = tanh(B0 * B0 * B0 * B0);
        // X**0.5 of the reference is the (correctly rounded) square root, as compilers evaluate it
        const double sok = fmax(sqrt((Vort * Vort + Skl * Skl) / 2.), 0.1 / tau);
        const double alturb = sqrt((RmiuT + RMIUL) / Den / (p.ref * sok * pow(0.09, 1.5)));
        const double A0 = 1. * fmax(((aLgrd / alturb / g0) - 0.1), 0.0);
        f.Fscheme[c] = fmax(1. * tanh(pow(A0, 3.)), p.f_limiter1);
    }
}

// Wall mark of face (i,j,k) in direction DIR (MARKWALL arrays, Input/Read_BlockDimensions.F90:146-283)
__device__ __forceinline__ bool wall_face(const int* __restrict__ m, const BlkInfo& b, int DIR, int i, int j,
                                          int k)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    if (DIR == 1) {
        if (i == 1) return m[(j - 1) + NJ1 * (k - 1)] == 1;
        if (i == b.ni) return m[NJ1 * NK1 + (j - 1) + NJ1 * (k - 1)] == 1;
    } else if (DIR == 2) {
        const int o = 2 * NJ1 * NK1;
        if (j == 1) return m[o + (i - 1) + NI1 * (k - 1)] == 1;
        if (j == b.nj) return m[o + NI1 * NK1 + (i - 1) + NI1 * (k - 1)] == 1;
    } else {
        const int o = 2 * NJ1 * NK1 + 2 * NI1 * NK1;
        if (k == 1) return m[o + (i - 1) + NI1 * (j - 1)] == 1;
        if (k == b.nk) return m[o + NI1 * NJ1 + (i - 1) + NI1 * (j - 1)] == 1;
    }
    return false;
}

// face gradient: volume-weighted cell gradients + face-centred correction
// (Space/DIFFUSION_Centered.F90:72-105 and its siblings)
__device__ __forceinline__ void face_grad(const Fields& f, const double* Sd, long cf, long st, int q0,
                                          long cm, long c0, double fpp, double fnn, double qp, double qn,
                                          double all, double alr, double alsinv, double out[3])
{
    const long tot = f.tot;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        const double g1 = f.DQ[(long)(q0 + x) * tot + cm], g0 = f.DQ[(long)(q0 + x) * tot + c0];
        const double gl = all * g1 + alr * g0;
        const double* S = Sd + (long)x * 3 * tot;
        out[x] = alsinv * (-S[cf + st] * (fpp - qp) - S[cf - st] * (qn - fnn) + S[cf] * (qp - qn)) + gl;
    }
}

// Viscous flux (Space/DIFFUSION_Centered.F90) and, when turbulent, the k-omega diffusion flux
// (Space/TURBTWO_DIFFUSION_Centered.F90, SST) through one face of direction DIR.
// out[0..3]: Fx, Fy, Fz, Fe; out[4..5]: Fk, Fo when turbulent.  (i, j, k): the face's index,
// the cell on its high side.
template <int DIR>
__device__ __forceinline__ void visc_face_flux(const trans_params& p, bool turb, const Fields& f,
                                               const BlkInfo& b, int i, int j, int k, double out[6])
{
    const long st = DIR == 1 ? 1 : (DIR == 2 ? (long)b.nx : (long)b.nx * b.ny);
    const long c0 = cidx(b, i, j, k);  // cell on the high side of face f
    const long cm = c0 - st;
    const long tot = f.tot;
    const bool wall = wall_face(f.mw + b.mwoff, b, DIR, i, j, k);
    // cells f-2, f-1, f, f+1 and the face values at faces f-1, f, f+1
    const CellQ qm2 = load_qc(f, c0 - 2 * st, turb);
    const CellQ qm1 = load_qc(f, cm, turb);
    const CellQ q0 = load_qc(f, c0, turb);
    const CellQ qp1 = load_qc(f, c0 + st, turb);
    const int d = DIR - 1;
    const CellQ fnn = face_q(f, d, c0 - 2 * st, cm, qm2, qm1, turb);
    const CellQ f0 = face_q(f, d, cm, c0, qm1, q0, turb);
    const CellQ fpp = face_q(f, d, c0, c0 + st, q0, qp1, turb);
    const double* Sd = f.SD + (long)d * tot;  // component (d, x) at Sd[3*x*tot]

    double alsinv = 1.0 / (f.Vol[c0] + f.Vol[cm]);
    double all = f.Vol[cm] * alsinv;
    double alr = f.Vol[c0] * alsinv;
    if (wall) alsinv = 0.0;
    double gu[3], gv[3], gw[3], gt[3];
    face_grad(f, Sd, c0, st, 0, cm, c0, fpp.u, fnn.u, q0.u, qm1.u, all, alr, alsinv, gu);
    face_grad(f, Sd, c0, st, 3, cm, c0, fpp.v, fnn.v, q0.v, qm1.v, all, alr, alsinv, gv);
    face_grad(f, Sd, c0, st, 6, cm, c0, fpp.w, fnn.w, q0.w, qm1.w, all, alr, alsinv, gw);
    face_grad(f, Sd, c0, st, 9, cm, c0, fpp.t, fnn.t, q0.t, qm1.t, all, alr, alsinv, gt);
    double gk[3], go[3];
    if (turb) {
        face_grad(f, Sd, c0, st, 12, cm, c0, fpp.k, fnn.k, q0.k, qm1.k, all, alr, alsinv, gk);
        face_grad(f, Sd, c0, st, 15, cm, c0, fpp.o, fnn.o, q0.o, qm1.o, all, alr, alsinv, go);
    }
    const double* Al = f.Alagm + (long)d * tot;
    alsinv = 1.0 / (Al[c0] + Al[cm]);
    all = Al[c0] * alsinv;
    alr = Al[cm] * alsinv;
    const double RmiuL = sutherland(p, f0.t);
    double RmiuTL = all * f.Rmiu[cm] + alr * f.Rmiu[c0];
    if (wall) RmiuTL = 0.0;
    const double NMD = -2. / 3.;
    const double DuDxL = gu[0], DuDyL = gu[1], DuDzL = gu[2];
    const double DvDxL = gv[0], DvDyL = gv[1], DvDzL = gv[2];
    const double DwDxL = gw[0], DwDyL = gw[1], DwDzL = gw[2];
    const double TxxLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DuDxL);
    const double TyyLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DvDyL);
    const double TzzLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DwDzL);
    const double TxyLam = RmiuL * (DuDyL + DvDxL);
    const double TyzLam = RmiuL * (DvDzL + DwDyL);
    const double TzxLam = RmiuL * (DuDzL + DwDxL);
    const double TxxTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DuDxL);
    const double TyyTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DvDyL);
    const double TzzTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DwDzL);
    const double TxyTur = RmiuTL * (DuDyL + DvDxL);
    const double TyzTur = RmiuTL * (DvDzL + DwDyL);
    const double TzxTur = RmiuTL * (DuDzL + DwDxL);
    const double Txx = TxxLam + TxxTur, Tyy = TyyLam + TyyTur, Tzz = TzzLam + TzzTur;
    const double Txy = TxyLam + TxyTur, Tyz = TyzLam + TyzTur, Tzx = TzxLam + TzxTur;
    const double prl = 0.72, prte = 0.9;
    const double cp = 1.4 * p.ppf / (1.4 - 1.0);
    const double akmu = cp * (RmiuL / prl + RmiuTL / prte);
    const double ucc = f0.u, vcc = f0.v, wcc = f0.w;
    const double Phix = akmu * gt[0] + ucc * Txx + vcc * Txy + wcc * Tzx;
    const double Phiy = akmu * gt[1] + ucc * Txy + vcc * Tyy + wcc * Tyz;
    const double Phiz = akmu * gt[2] + ucc * Tzx + vcc * Tyz + wcc * Tzz;
    const double Dx = Sd[c0], Dy = Sd[3 * tot + c0], Dz = Sd[6 * tot + c0];
    out[0] = (Txx * Dx + Txy * Dy + Tzx * Dz) / p.ref;
    out[1] = (Txy * Dx + Tyy * Dy + Tyz * Dz) / p.ref;
    out[2] = (Tzx * Dx + Tyz * Dy + Tzz * Dz) / p.ref;
    out[3] = (Phix * Dx + Phiy * Dy + Phiz * Dz) / p.ref;
    if (!turb) return;
    // k-omega diffusion (Space/TURBTWO_DIFFUSION_Centered.F90:40-214), same Alagm weights
    const double TTL = all * f.T[cm] + alr * f.T[c0];
    const double RmiuLt = (1.0 + p.csthlnd) / (TTL + p.csthlnd) * pow(TTL, 1.5);
    const double DEN0 = f.V[c0];
    const double TKK0 = q0.k, TOO0 = q0.o;
    const double DEN1 = f.V[cm];
    const double TKK1 = qm1.k, TOO1 = qm1.o;
    const double Den = all * DEN1 + alr * DEN0;
    const double TKK = all * TKK1 + alr * TKK0;
    const double TOO = all * TOO1 + alr * TOO0;
    const double dTKO = gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2];
    const double CDKO = fmax(2.0 * Den * p.sigo2 / TOO * (dTKO), 1.0e-10);
    const double distfc = alr * f.Dst[c0] + all * f.Dst[cm];
    double arg1 = fmax(sqrt(TKK) / p.beta_star / TOO / distfc / p.ref,
                       500.0 * RmiuLt / Den / distfc / distfc / TOO / p.ref / p.ref);
    arg1 = fmin(arg1, 4.0 * Den * p.sigo2 * TKK / CDKO / distfc / distfc);
    const double a2 = arg1 * arg1;
    const double fc1c = tanh(a2 * a2);
    // WD+ (Kind_Model = 2, :168-170) takes the unblended sigmas
    const double SigkAB = p.kind_model == 2 ? p.sigk1 : fc1c * p.sigk1 + (1. - fc1c) * p.sigk2;
    const double SigoAB = p.kind_model == 2 ? p.sigo1 : fc1c * p.sigo1 + (1. - fc1c) * p.sigo2;
    const double vmute = RmiuLt + RmiuTL * SigkAB;
    const double vmuoq = RmiuLt + RmiuTL * SigoAB;
    out[4] = (vmute * (gk[0] * Dx + gk[1] * Dy + gk[2] * Dz)) / p.ref;
    out[5] = (vmuoq * (go[0] * Dx + go[1] * Dy + go[2] * Dz)) / p.ref;
}

// One thread per face of direction DIR: GV[DIR-1] receives visc_face_flux.
template <int DIR>
__global__ void __launch_bounds__(256, 3) k_visc_face(trans_params p, int if_turb, Fields f,
                                                   const BlkInfo* __restrict__ blks,
                                                   const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int n0 = DIR == 1 ? b.ni : b.ni - 1;
    const int n1 = DIR == 2 ? b.nj : b.nj - 1;
    const int n2 = DIR == 3 ? b.nk : b.nk - 1;
    const int lin = s.base + threadIdx.x;
    int i, j, k;
    if (DIR == 1) {
        if (lin >= n0 * n1 * n2) return;
        decode3(lin, n0, n1, i, j, k);
    } else if (!decode_face_tile<DIR>(lin, n0, n1, n2, i, j, k)) {
        return;
    }
    double o[6];
    visc_face_flux<DIR>(p, if_turb != 0, f, b, i, j, k, o);
    const long c0 = cidx(b, i, j, k), tot = f.tot;
    double* G = f.GV[DIR - 1];
#pragma unroll
    for (int q = 0; q < 4; q++) G[(long)q * tot + c0] = o[q];
    if (if_turb) {
        G[4 * tot + c0] = o[4];
        G[5 * tot + c0] = o[5];
    }
}

// ---------------------------------------------------------------------------------------
// Space/CONVECT_MDCD.F90:41-88 -- one reconstruction from the six values a_2..a3.
__device__ __forceinline__ double mdcd_rec(double gama, double v_2, double v_1, double v0, double v1,
                                           double v2, double v3)
{
    const double beta = sqrt(13.0 / 12.0), epsm = 1.0e-6;
    const double aw0 = 1.5 * (0.0463783 + gama);
    const double aw1 = 0.5 - 1.5 * 0.0463783 + 4.5 * gama;
    const double aw2 = 0.5 - 1.5 * 0.0463783 - 4.5 * gama;
    const double aw3 = 1.5 * (0.0463783 - gama);
    const double I0 = sqr(0.5 * v_2 - 2.0 * v_1 + 1.5 * v0) + sqr(beta * v_2 - 2.0 * beta * v_1 + beta * v0);
    const double I1 = sqr(-0.5 * v_1 + 0.0 * v0 + 0.5 * v1) + sqr(beta * v_1 - 2.0 * beta * v0 + beta * v1);
    const double I2 = sqr(-1.5 * v0 + 2.0 * v1 - 0.5 * v2) + sqr(beta * v0 - 2.0 * beta * v1 + beta * v2);
    double I3 = sqr(2.5 * v1 - 4.0 * v2 + 1.5 * v3) + sqr(beta * v1 - 2.0 * beta * v2 + beta * v3);
    I3 = fmax(fmax(fmax(I0, I1), I2), I3);
    const double ak0 = aw0 / sqr(epsm + I0);
    const double ak1 = aw1 / sqr(epsm + I1);
    const double ak2 = aw2 / sqr(epsm + I2);
    const double ak3 = aw3 / sqr(epsm + I3);
    const double w0 = ak0 / (ak0 + ak1 + ak2 + ak3);
    const double w1 = ak1 / (ak0 + ak1 + ak2 + ak3);
    const double w2 = ak2 / (ak0 + ak1 + ak2 + ak3);
    const double w3 = ak3 / (ak0 + ak1 + ak2 + ak3);
    const double qq0 = 2. / 6. * v_2 - 7. / 6. * v_1 + 11. / 6. * v0;
    const double qq1 = -1. / 6. * v_1 + 5. / 6. * v0 + 2. / 6. * v1;
    const double qq2 = 2. / 6. * v0 + 5. / 6. * v1 - 1. / 6. * v2;
    const double qq3 = 11. / 6. * v1 - 7. / 6. * v2 + 1. / 3. * v3;
    return w0 * qq0 + w1 * qq1 + w2 * qq2 + w3 * qq3;
}

// Space/CONVECT_AUSMPW_Plus.F90:204-317 (the 3/16 coefficient is an integer division = 0)
__device__ __forceinline__ void ausmpw_common(const trans_params& p, double PLo, double PRo, double TLo,
                                              double TRo, double ULo, double URo, double VLo, double VRo,
                                              double WLo, double WRo, double sav1, double sav2, double sav3,
                                              double sav, double out[5])
{
    const double rcpcv = p.ppf, tiny = p.tiny;
    const double DENLo = PLo / (rcpcv * TLo);
    const double DENRo = PRo / (rcpcv * TRo);
    const double s1 = sav1 / sav, s2 = sav2 / sav, s3 = sav3 / sav;
    const double ga = 1.4, ga1 = 1.4 - 1.0;
    const double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo;
    const double VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    const double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo;
    const double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo;
    const double ULn = (s1 * ULo + s2 * VLo + s3 * WLo);
    const double URn = (s1 * URo + s2 * VRo + s3 * WRo);
    const double astar_L = sqrt(2 * ga1 * HLo / (ga + 1));
    const double astar_R = sqrt(2 * ga1 * HRo / (ga + 1));
    const double aas_L = sqr(astar_L) / fmax(sqrt(VAMLo), astar_L);
    const double aas_R = sqr(astar_R) / fmax(sqrt(VAMRo), astar_R);
    const double Acc = fmin(aas_L, aas_R);
    const double MaL = ULn / Acc, MaR = URn / Acc;
    const double c316 = 0.0;
    double MaL_plus, pL_plus, MaR_neg, pR_neg;
    if (fabs(MaL) > 1.0) {
        MaL_plus = 0.5 * (MaL + fabs(MaL));
        pL_plus = 0.5 * (MaL + fabs(MaL)) / (MaL + tiny);
    } else {
        MaL_plus = 0.25 * sqr(MaL + 1);
        pL_plus = 0.25 * sqr(MaL + 1) * (2 - MaL) + c316 * MaL * sqr(MaL * MaL - 1);
    }
    if (fabs(MaR) > 1.0) {
        MaR_neg = 0.5 * (MaR - fabs(MaR));
        pR_neg = 0.5 * (MaR - fabs(MaR)) / (MaR + tiny);
    } else {
        MaR_neg = (-1) * 0.25 * sqr(MaR - 1);
        pR_neg = 0.25 * sqr(MaR - 1) * (2 + MaR) - c316 * MaR * sqr(MaR * MaR - 1);
    }
    const double pS = pL_plus * PLo + pR_neg * PRo;
    const double Mas = MaL_plus + MaR_neg;
    const double mn = fmin(PLo / PRo, PRo / PLo);
    const double wP = 1.0 - (mn * mn) * mn;
    const double fL = fabs(MaL) < 1 ? PLo / (pS + tiny) - 1 : 0;
    const double fR = fabs(MaR) < 1 ? PRo / (pS + tiny) - 1 : 0;
    double MMaL_plus, MMaR_neg;
    if (Mas >= 0) {
        MMaL_plus = MaL_plus + MaR_neg * ((1 - wP) * (1 + fR) - fL);
        MMaR_neg = MaR_neg * wP * (1 + fR);
    } else {
        MMaL_plus = MaL_plus * wP * (1 + fL);
        MMaR_neg = MaR_neg + MaL_plus * ((1 - wP) * (1 + fL) - fR);
    }
    out[0] = (-1) * MMaL_plus * Acc * DENLo - MMaR_neg * Acc * DENRo;
    out[1] = (-1) * MMaL_plus * Acc * DENLo * ULo - MMaR_neg * Acc * DENRo * URo - s1 * (pL_plus * PLo + pR_neg * PRo);
    out[2] = (-1) * MMaL_plus * Acc * DENLo * VLo - MMaR_neg * Acc * DENRo * VRo - s2 * (pL_plus * PLo + pR_neg * PRo);
    out[3] = (-1) * MMaL_plus * Acc * DENLo * WLo - MMaR_neg * Acc * DENRo * WRo - s3 * (pL_plus * PLo + pR_neg * PRo);
    out[4] = (-1) * MMaL_plus * Acc * DENLo * HLo - MMaR_neg * Acc * DENRo * HRo;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Convective flux through one face of direction DIR: reconstruction (SPATIAL 1 = CONVECT_MUSCL2,
// 2 = CONVECT_MUSCL, 3 = CONVECT_MDCD, 4 = CONVECT_MDCDHYB), inviscid flux (ROE 1 =
// CONVECT_Roe_Original, 2 = CONVECT_Roe_Rotated, 3 = CONVECT_AUSMPW_Plus) and, when turbulent,
// TURBTWO_CONVECT_MUSCL + TURBTWO_CONVECT_Upwind.  out[0..4]: flux*|S| (the rotated flux is
// area-weighted already), out[5..6]: the area-weighted k, omega fluxes.
template <int DIR, int SPATIAL, int ROE>
__device__ __forceinline__ void conv_face_flux(const trans_params& p, bool turb, const Fields& f,
                                               const BlkInfo& b, int i, int j, int k, double out[7])
{
    const long st = DIR == 1 ? 1 : (DIR == 2 ? (long)b.nx : (long)b.nx * b.ny);
    const long c0 = cidx(b, i, j, k);
    const long tot = f.tot;
    const int d = DIR - 1;
    const double* Al = f.Alagm + (long)d * tot;
    double qL[6], qR[6];
    if (SPATIAL == 3 || SPATIAL == 4) {
        // six cells f-3..f+2; the shock sensor's ghost values are those of the nearest
        // interior cell (Space/DIFF_DqDxyz.F90:388-397)
        double sv[6][6];
        double sensor = 0.0;
        const int pos[3] = {i, j, k};
        const int nmax[3] = {b.ni - 1, b.nj - 1, b.nk - 1};
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const long c = c0 + (long)(n - 3) * st;
            sv[0][n] = f.Q[c];
            sv[1][n] = f.Q[tot + c];
            sv[2][n] = f.Q[2 * tot + c];
            sv[3][n] = f.PP[c];
            sv[4][n] = f.T[c];
            sv[5][n] = f.rad[c];
            int cp[3] = {pos[0], pos[1], pos[2]};
            cp[d] += n - 3;
            const long cs = cidx(b, clampi(cp[0], 1, nmax[0]), clampi(cp[1], 1, nmax[1]), clampi(cp[2], 1, nmax[2]));
            const double sh = f.shock[cs];
            sensor = n == 0 ? sh : fmax(sensor, sh);
        }
        if (SPATIAL == 3 && DIR == 3) sensor = fmax(sensor, 0.0);
        const double gama = p.gama_min + sensor * (p.gama_max - p.gama_min);
#pragma unroll
        for (int L = 0; L < 5; L++) {
            qL[L] = mdcd_rec(gama, sv[L][0], sv[L][1], sv[L][2], sv[L][3], sv[L][4], sv[L][5]);
            qR[L] = mdcd_rec(gama, sv[L][5], sv[L][4], sv[L][3], sv[L][2], sv[L][1], sv[L][0]);
        }
        if (SPATIAL == 4 && sensor >= p.f_limiter2) {
            // Space/CONVECT_MDCDHYB.F90: half-slope ALIMITER from cell values near shocks
#pragma unroll
            for (int L = 0; L < 5; L++) {
                qL[L] = sv[L][2] + 0.5 * alimiter(sv[L][3] - sv[L][2], sv[L][2] - sv[L][1]);
                qR[L] = sv[L][3] - 0.5 * alimiter(sv[L][3] - sv[L][2], sv[L][4] - sv[L][3]);
            }
        }
        if (SPATIAL == 4 || DIR == 3) {
            // MDCDHYB leaves all face radii, and MDCD the k-face radii (CONVECT_MDCD.F90:450-521
            // stands outside its loop), at the face average
            qL[5] = face_avg(Al[c0 - st], Al[c0], sv[5][2], sv[5][3]);
            qR[5] = qL[5];
        } else {
            qL[5] = mdcd_rec(gama, sv[5][0], sv[5][1], sv[5][2], sv[5][3], sv[5][4], sv[5][5]);
            qR[5] = mdcd_rec(gama, sv[5][5], sv[5][4], sv[5][3], sv[5][2], sv[5][1], sv[5][0]);
        }
    } else {
        const Prim qm2 = load_prim(f, c0 - 2 * st);
        const Prim qm1 = load_prim(f, c0 - st);
        const Prim q0 = load_prim(f, c0);
        const Prim qp1 = load_prim(f, c0 + st);
        const double am2 = Al[c0 - 2 * st], am1 = Al[c0 - st], a0 = Al[c0], ap1 = Al[c0 + st];
#pragma unroll
        for (int L = 0; L < 6; L++) {
            const double avm = face_avg(am2, am1, qm2.q[L], qm1.q[L]);
            const double av0 = face_avg(am1, a0, qm1.q[L], q0.q[L]);
            const double avp = face_avg(a0, ap1, q0.q[L], qp1.q[L]);
            if (SPATIAL == 1) {  // Space/CONVECT_MUSCL2.F90
                qL[L] = qm1.q[L] + muscl2(av0 - qm1.q[L], qm1.q[L] - avm);
                qR[L] = q0.q[L] - muscl2(q0.q[L] - av0, avp - q0.q[L]);
            } else {  // Space/CONVECT_MUSCL.F90
                qL[L] = qm1.q[L] + alimiter(av0 - qm1.q[L], qm1.q[L] - avm);
                qR[L] = q0.q[L] - alimiter(q0.q[L] - av0, avp - q0.q[L]);
            }
        }
    }
    const double* S = f.SD + (long)d * tot;
    const double sav = f.Grad[(long)d * tot + c0];
    // face eddy viscosity (Space/SPACEINTERP_GetSurfCenterVars.F90, VL(d,8)) and Alagm of the
    // preconditioned Roe eigenvalues
    const double RmiuT = p.if_precondition && turb ? face_avg(Al[c0 - st], Al[c0], f.Rmiu[c0 - st], f.Rmiu[c0]) : 0.0;
    const double AlagmR = Al[c0];
    double fl[5];
    if (ROE == 2) {  // Space/CONVECT_Roe_Rotated.F90: area-weighted flux, scattered as it is
        // Funscheme from the two cells; the first ghost layer of Fscheme is the nearest
        // interior cell (Space/DIFF_DqDxyz.F90:399-404)
        int pm[3] = {i, j, k};
        pm[d] -= 1;
        const int n3[3] = {b.ni - 1, b.nj - 1, b.nk - 1};
        const long ca = cidx(b, clampi(pm[0], 1, n3[0]), clampi(pm[1], 1, n3[1]), clampi(pm[2], 1, n3[2]));
        const long cb = cidx(b, clampi(i, 1, n3[0]), clampi(j, 1, n3[1]), clampi(k, 1, n3[2]));
        const double Funscheme = (f.Fscheme[ca] + f.Fscheme[cb]) / 2.;
        roe_rotated_face(p, qL[3], qR[3], qL[4], qR[4], qL[0], qR[0], qL[1], qR[1], qL[2], qR[2], S[c0],
                         S[3 * tot + c0], S[6 * tot + c0], sav, Funscheme, qL[5] * p.omega1, qR[5] * p.omega1, RmiuT,
                         AlagmR, fl);
#pragma unroll
        for (int L = 0; L < 5; L++) out[L] = fl[L];
    } else if (ROE == 1)
        roe_common(p, qL[3], qR[3], qL[4], qR[4], qL[0], qR[0], qL[1], qR[1], qL[2], qR[2], S[c0], S[3 * tot + c0],
                   S[6 * tot + c0], sav, qL[5] * p.omega1, qR[5] * p.omega1, RmiuT, AlagmR, fl);
    else
        ausmpw_common(p, qL[3], qR[3], qL[4], qR[4], qL[0], qR[0], qL[1], qR[1], qL[2], qR[2], S[c0],
                      S[3 * tot + c0], S[6 * tot + c0], sav, fl);
    if (ROE != 2)
#pragma unroll
        for (int L = 0; L < 5; L++) out[L] = fl[L] * sav;
    if (!turb) return;
    // k/rho, omega/rho: face averages + ALIMITER reconstruction (Space/TURBTWO_CONVECT_MUSCL.F90)
    double tL[2], tR[2];
    {
        const long cm2 = c0 - 2 * st, cm1 = c0 - st, cp1 = c0 + st;
        const double am2 = Al[cm2], am1 = Al[cm1], a0 = Al[c0], ap1 = Al[cp1];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double* Qq = f.Q + (long)(3 + q) * tot;  // k/rho, omega/rho
            const double vm2 = Qq[cm2];
            const double vm1 = Qq[cm1];
            const double v0 = Qq[c0];
            const double vp1 = Qq[cp1];
            const double avm = face_avg(am2, am1, vm2, vm1);
            const double av0 = face_avg(am1, a0, vm1, v0);
            const double avp = face_avg(a0, ap1, v0, vp1);
            tL[q] = vm1 + alimiter(av0 - vm1, vm1 - avm);
            tR[q] = v0 - alimiter(v0 - av0, avp - v0);
        }
    }
    // Space/TURBTWO_CONVECT_Upwind.F90:58-115
    const double rcpcv = p.ppf;
    const double DENL = qL[3] / (rcpcv * qL[4]), DENR = qR[3] / (rcpcv * qR[4]);
    const double sav1 = S[c0], sav2 = S[3 * tot + c0], sav3 = S[6 * tot + c0];
    const double ulnormal = (sav1 * qL[0] + sav2 * qL[1] + sav3 * qL[2]);
    const double urnormal = (sav1 * qR[0] + sav2 * qR[1] + sav3 * qR[2]);
    const double rulnormal = DENL * ulnormal, rurnormal = DENR * urnormal;
    const double small = 1.e-10;
    const double eigTKO = 0.5 * fmax(fabs(ulnormal), fabs(urnormal));
    const double TKL = fmax(tL[0], small), TKR = fmax(tR[0], small);
    const double dTK = -0.5 * (rulnormal * TKL + rurnormal * TKR - eigTKO * (DENR * TKR - DENL * TKL));
    const double TOL = fmax(tL[1], small), TOR = fmax(tR[1], small);
    const double dTO = -0.5 * (rulnormal * TOL + rurnormal * TOR - eigTKO * (DENR * TOR - DENL * TOL));
    out[5] = dTK;
    out[6] = dTO;
}

// One thread per face of direction DIR: G[DIR-1] receives conv_face_flux.
template <int DIR, int SPATIAL, int ROE>
__global__ void __launch_bounds__(256) k_conv_face(trans_params p, int if_turb, Fields f,
                                                   const BlkInfo* __restrict__ blks,
                                                   const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int n0 = DIR == 1 ? b.ni : b.ni - 1;
    const int n1 = DIR == 2 ? b.nj : b.nj - 1;
    const int n2 = DIR == 3 ? b.nk : b.nk - 1;
    const int lin = s.base + threadIdx.x;
    int i, j, k;
    if (DIR == 1) {
        if (lin >= n0 * n1 * n2) return;
        decode3(lin, n0, n1, i, j, k);
    } else if (!decode_face_tile<DIR>(lin, n0, n1, n2, i, j, k)) {
        return;
    }
    double o[7];
    conv_face_flux<DIR, SPATIAL, ROE>(p, if_turb != 0, f, b, i, j, k, o);
    const long c0 = cidx(b, i, j, k), tot = f.tot;
    double* G = f.G[DIR - 1];
#pragma unroll
    for (int L = 0; L < 5; L++) G[(long)L * tot + c0] = o[L];
    if (if_turb) {
        G[5 * tot + c0] = o[5];
        G[6 * tot + c0] = o[6];
    }
}

// ---------------------------------------------------------------------------------------
// D in the reference's accumulation order: D = 0; diffusion i, j, k (own face subtracted,
// next face added); rotating-frame source; convection i, j, k.  D(6:7): k-omega diffusion
// i, j, k then k-omega upwind i, j, k (main.F90:92-133).
__global__ void __launch_bounds__(256) k_accumulate_visc(trans_params p, int if_euler, int if_turb, Fields f,
                                                         const BlkInfo* __restrict__ blks,
                                                         const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k);
    const long tot = f.tot;
    const long st[3] = {1, (long)b.nx, (long)b.nx * b.ny};
    double dd[7];
    dd[0] = 0.0;
#pragma unroll
    for (int L = 1; L < 5; L++) {
        double x = 0.0;
        if (!if_euler)
#pragma unroll
            for (int dir = 0; dir < 3; dir++) {
                const double* G = f.GV[dir] + (long)(L - 1) * tot;
                x = x - G[c];
                x = x + G[c + st[dir]];
            }
        dd[L] = x;
    }
    if (!if_euler) {
        const double RoW = p.omega1;
        const double v1 = f.V[c];
        const double ss3 = 2.0 * RoW * f.V[3 * tot + c] + RoW * f.Yc[c] * p.omega1 * v1;
        const double ss4 = -2.0 * RoW * f.V[2 * tot + c] + RoW * f.Zc[c] * p.omega1 * v1;
        dd[2] = dd[2] + ss3 * f.Vol[c];
        dd[3] = dd[3] + ss4 * f.Vol[c];
    }
#pragma unroll
    for (int L = 0; L < 5; L++) {
        double x = dd[L];
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const double* G = f.G[dir] + (long)L * tot;
            x = x - G[c];
            x = x + G[c + st[dir]];
        }
        f.D[(long)L * tot + c] = x;
    }
    if (!if_turb) {
        f.D[5 * tot + c] = 0.0;
        f.D[6 * tot + c] = 0.0;
        return;
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        double x = 0.0;
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const double* G = f.GV[dir] + (long)(4 + q) * tot;
            x = x - G[c];
            x = x + G[c + st[dir]];
        }
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const double* G = f.G[dir] + (long)(5 + q) * tot;
            x = x - G[c];
            x = x + G[c + st[dir]];
        }
        f.D[(long)(5 + q) * tot + c] = x;
    }
}

// ---------------------------------------------------------------------------------------
// LDS-tiled face fluxes and accumulation in two passes (an alternative to the face kernels
// through G and GV plus k_accumulate_visc, selected with TRANS_HIP_FUSED=1; measured slower on
// the MI355X, see the design notes).  A workgroup owns an RTI x RTJ (i, j) tile of cells of one
// block and marches k through a chunk of planes.  At each plane the fluxes of the tile's i and
// j faces (its cells' low and high faces) are formed once into LDS, flat over the workgroup's
// threads; each thread forms the flux of its cell's high k face in registers and keeps it as
// the next plane's low k face.  Each cell then runs its sums in the reference order (i, j, k;
// own face subtracted, next face added).  k_resid_visc forms the viscous faces and leaves each
// cell's viscous sums with the rotating-frame source in D (D(1) = 0); k_resid_conv forms the
// convective faces and continues the sums from D.  Every flux is the face kernels' own
// expression, every sum the accumulation kernel's order, and the partial sums pass through HBM
// unrounded (FP64), so D is bitwise that of the face-kernel path.  A chunk starts by forming
// the low k faces of its first plane.
constexpr int RTI = 64, RTJ = 4;
constexpr int RT_NI = (RTI + 1) * RTJ;          // i-face slots: (fi, fj) at fj * (RTI + 1) + fi
constexpr int RT_NF = RT_NI + RTI * (RTJ + 1);  // then j-face slots: (fi, fj) at RT_NI + fj * RTI + fi

__global__ void __launch_bounds__(256, 2) k_resid_visc(trans_params p, int if_turb, Fields f,
                                                    const BlkInfo* __restrict__ blks,
                                                    const Seg* __restrict__ seg, int kchunk)
{
    __shared__ double lds[6 * RT_NF];
    const Seg sg = seg[blockIdx.x];
    const BlkInfo b = blks[sg.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int ntx = (NI1 + RTI - 1) / RTI, nty = (NJ1 + RTJ - 1) / RTJ;
    const int tx = sg.base % ntx, ty = (sg.base / ntx) % nty, tz = sg.base / (ntx * nty);
    const int i0 = tx * RTI + 1, j0 = ty * RTJ + 1, k0 = tz * kchunk + 1;
    const int k1 = min(k0 + kchunk - 1, NK1);
    const int ni_t = min(RTI, NI1 - i0 + 1), nj_t = min(RTJ, NJ1 - j0 + 1);
    const int nfi = (ni_t + 1) * nj_t, nf = nfi + ni_t * (nj_t + 1);
    const int t = threadIdx.x;
    const int ci = t % RTI, cj = t / RTI;
    const bool own = ci < ni_t && cj < nj_t;
    const int i = i0 + ci, j = j0 + cj;
    const bool turb = if_turb != 0;
    const long tot = f.tot;
    const int si = cj * (RTI + 1) + ci, sj = RT_NI + cj * RTI + ci;
    double gv_lo[6] = {0, 0, 0, 0, 0, 0};
    if (own) visc_face_flux<3>(p, turb, f, b, i, j, k0, gv_lo);
    for (int k = k0; k <= k1; k++) {
        for (int q = t; q < nf; q += 256) {
            double o[6];
            int slot;
            if (q < nfi) {
                const int fi = q % (ni_t + 1), fj = q / (ni_t + 1);
                visc_face_flux<1>(p, turb, f, b, i0 + fi, j0 + fj, k, o);
                slot = fj * (RTI + 1) + fi;
            } else {
                const int fi = (q - nfi) % ni_t, fj = (q - nfi) / ni_t;
                visc_face_flux<2>(p, turb, f, b, i0 + fi, j0 + fj, k, o);
                slot = RT_NI + fj * RTI + fi;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) lds[c * RT_NF + slot] = o[c];
            if (turb) {
                lds[4 * RT_NF + slot] = o[4];
                lds[5 * RT_NF + slot] = o[5];
            }
        }
        double gv_hi[6] = {0, 0, 0, 0, 0, 0};
        if (own) visc_face_flux<3>(p, turb, f, b, i, j, k + 1, gv_hi);
        __syncthreads();
        if (own) {
            const long c = cidx(b, i, j, k);
            double x[7];
            x[0] = 0.0;
#pragma unroll
            for (int L = 1; L < 5; L++) {
                const double* V = lds + (L - 1) * RT_NF;
                double y = 0.0;
                y = y - V[si];
                y = y + V[si + 1];
                y = y - V[sj];
                y = y + V[sj + RTI];
                y = y - gv_lo[L - 1];
                y = y + gv_hi[L - 1];
                x[L] = y;
            }
            const double RoW = p.omega1;
            const double v1 = f.V[c];
            const double ss3 = 2.0 * RoW * f.V[3 * tot + c] + RoW * f.Yc[c] * p.omega1 * v1;
            const double ss4 = -2.0 * RoW * f.V[2 * tot + c] + RoW * f.Zc[c] * p.omega1 * v1;
            x[2] = x[2] + ss3 * f.Vol[c];
            x[3] = x[3] + ss4 * f.Vol[c];
#pragma unroll
            for (int L = 0; L < 5; L++) f.D[(long)L * tot + c] = x[L];
            if (turb)
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const double* V = lds + (4 + q) * RT_NF;
                    double y = 0.0;
                    y = y - V[si];
                    y = y + V[si + 1];
                    y = y - V[sj];
                    y = y + V[sj + RTI];
                    y = y - gv_lo[4 + q];
                    y = y + gv_hi[4 + q];
                    f.D[(long)(5 + q) * tot + c] = y;
                }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 6; c++) gv_lo[c] = gv_hi[c];
    }
}

template <int SPATIAL, int ROE>
__global__ void __launch_bounds__(256, 2) k_resid_conv(trans_params p, int if_euler, int if_turb, Fields f,
                                                    const BlkInfo* __restrict__ blks,
                                                    const Seg* __restrict__ seg, int kchunk)
{
    __shared__ double lds[7 * RT_NF];
    const Seg sg = seg[blockIdx.x];
    const BlkInfo b = blks[sg.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int ntx = (NI1 + RTI - 1) / RTI, nty = (NJ1 + RTJ - 1) / RTJ;
    const int tx = sg.base % ntx, ty = (sg.base / ntx) % nty, tz = sg.base / (ntx * nty);
    const int i0 = tx * RTI + 1, j0 = ty * RTJ + 1, k0 = tz * kchunk + 1;
    const int k1 = min(k0 + kchunk - 1, NK1);
    const int ni_t = min(RTI, NI1 - i0 + 1), nj_t = min(RTJ, NJ1 - j0 + 1);
    const int nfi = (ni_t + 1) * nj_t, nf = nfi + ni_t * (nj_t + 1);
    const int t = threadIdx.x;
    const int ci = t % RTI, cj = t / RTI;
    const bool own = ci < ni_t && cj < nj_t;
    const int i = i0 + ci, j = j0 + cj;
    const bool turb = if_turb != 0, visc = if_euler == 0;
    const long tot = f.tot;
    const int si = cj * (RTI + 1) + ci, sj = RT_NI + cj * RTI + ci;
    double g_lo[7] = {0, 0, 0, 0, 0, 0, 0};
    if (own) conv_face_flux<3, SPATIAL, ROE>(p, turb, f, b, i, j, k0, g_lo);
    for (int k = k0; k <= k1; k++) {
        for (int q = t; q < nf; q += 256) {
            double o[7];
            int slot;
            if (q < nfi) {
                const int fi = q % (ni_t + 1), fj = q / (ni_t + 1);
                conv_face_flux<1, SPATIAL, ROE>(p, turb, f, b, i0 + fi, j0 + fj, k, o);
                slot = fj * (RTI + 1) + fi;
            } else {
                const int fi = (q - nfi) % ni_t, fj = (q - nfi) / ni_t;
                conv_face_flux<2, SPATIAL, ROE>(p, turb, f, b, i0 + fi, j0 + fj, k, o);
                slot = RT_NI + fj * RTI + fi;
            }
#pragma unroll
            for (int c = 0; c < 5; c++) lds[c * RT_NF + slot] = o[c];
            if (turb) {
                lds[5 * RT_NF + slot] = o[5];
                lds[6 * RT_NF + slot] = o[6];
            }
        }
        double g_hi[7] = {0, 0, 0, 0, 0, 0, 0};
        if (own) conv_face_flux<3, SPATIAL, ROE>(p, turb, f, b, i, j, k + 1, g_hi);
        __syncthreads();
        if (own) {
            const long c = cidx(b, i, j, k);
#pragma unroll
            for (int L = 0; L < 5; L++) {
                const double* G = lds + L * RT_NF;
                double y = visc ? f.D[(long)L * tot + c] : 0.0;
                y = y - G[si];
                y = y + G[si + 1];
                y = y - G[sj];
                y = y + G[sj + RTI];
                y = y - g_lo[L];
                y = y + g_hi[L];
                f.D[(long)L * tot + c] = y;
            }
            if (turb) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const double* G = lds + (5 + q) * RT_NF;
                    double y = visc ? f.D[(long)(5 + q) * tot + c] : 0.0;
                    y = y - G[si];
                    y = y + G[si + 1];
                    y = y - G[sj];
                    y = y + G[sj + RTI];
                    y = y - g_lo[5 + q];
                    y = y + g_hi[5 + q];
                    f.D[(long)(5 + q) * tot + c] = y;
                }
            } else {
                f.D[5 * tot + c] = 0.0;
                f.D[6 * tot + c] = 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 7; c++) g_lo[c] = g_hi[c];
    }
}

// ---------------------------------------------------------------------------------------
// Time/TURBTWO_TIME_LUSGS_SST.F90:60-415: k-omega right-hand side, SST sources (RANS,
// IF_Comppress = 0, f_r1 = 1) and the diagonal F1t = 1 + sum(Rds)*stept/Vol + (Cneg)*stept.
// SST blending F1 = tanh(arg^4) of cell c as the SST source forms it
// (Time/TURBTWO_TIME_LUSGS_SST.F90:263-270)
__device__ __forceinline__ double sst_fc1(const trans_params& p, const Fields& f, long c)
{
    const long tot = f.tot;
    const double Ref = p.ref;
    const double ARo = f.V[c], ARTk = f.V[5 * tot + c], ARTO = f.V[6 * tot + c];
    const double* Q = f.DQ;
    const double cdif = Q[12 * tot + c] * Q[15 * tot + c] + Q[13 * tot + c] * Q[16 * tot + c] +
                        Q[14 * tot + c] * Q[17 * tot + c];
    const double RmiuL = sutherland(p, f.T[c]);
    const double DistN = f.Dst[c];
    const double CDKO = fmax(2.0 * ARo * p.sigo2 / (ARTO / ARo) * cdif, 1.0e-10);
    double arg1 = fmax(sqrt(ARTk / ARo) / p.beta_star / (ARTO / ARo) / DistN / Ref,
                       500.0 * RmiuL / DistN / DistN / ARTO / Ref / Ref);
    arg1 = fmin(arg1, 4.0 * p.sigo2 * ARTk / CDKO / DistN / DistN);
    const double a2 = arg1 * arg1;
    return tanh(a2 * a2);
}

// DES/DDES shielding function FSST (Time/Hybrid_function.F90:1-39): 0 for Kind_SubHY 0
// (Strelets DES), Menter's F1 for 1 and 3, F2 for 2
__device__ __forceinline__ double fsst_des(const trans_params& p, double Den, double vK, double vOmega, double Dist,
                                           double RmiuL, double CDiff)
{
    const double Ref = p.ref;
    if (p.kind_subhy == 1 || p.kind_subhy == 3) {
        const double arg1 = sqrt(vK) / (0.09 * vOmega * Dist * Ref);
        const double arg2 = 500. * RmiuL / (Den * Dist * Dist * vOmega * Ref * Ref);
        const double arga = fmax(arg1, arg2);
        const double temp = fmax(Den * CDiff * Ref, 1.e-10);
        const double argb = 4. * Den * 0.856 * vK / (temp * Dist * Dist);
        const double arg = fmin(arga, argb);
        return tanh(arg * arg * arg * arg);
    }
    if (p.kind_subhy == 2) {
        const double arg1 = 2. * sqrt(vK) / (0.09 * vOmega * Dist * Ref);
        const double arg2 = 500. * RmiuL / (Den * Dist * Dist * vOmega * Ref * Ref);
        const double arg = fmax(arg1, arg2);
        return tanh(arg * arg);
    }
    return 0.;
}

// IDDES functions (Time/Hybrid_function.F90:42-105; kappa = 0.41, Ct = 1.87, Cl = 5,
// Module_Global.F90:126).  S = max(|grad u|, 1e-10) (Frobenius norm of the 9 gradients).
__device__ __forceinline__ double iddes_fd(double Den, double RmiuT, double aDst, double aLmax, double Sn, double Ref)
{
    const double alpha1 = 0.25 - aDst / aLmax;
    const double fb = fmin(2. * exp(-9.0 * alpha1 * alpha1), 1.);
    const double rdt = (RmiuT / Den) / (0.41 * 0.41 * (aDst * aDst) * Sn * Ref);
    const double fdt = 1. - tanh((8. * rdt) * (8. * rdt) * (8. * rdt));
    return fmax(1. - fdt, fb);
}

__device__ __forceinline__ double iddes_fe(double Den, double RmiuT, double RmiuL, double aDst, double aLmax, double Sn,
                                           double Ref)
{
    const double alpha2 = 0.25 - aDst / aLmax;
    const double fe1 = alpha2 >= 0. ? 2. * exp(-11.09 * alpha2 * alpha2) : 2. * exp(-9. * alpha2 * alpha2);
    const double rdt = (RmiuT / Den) / (0.41 * 0.41 * (aDst * aDst) * Sn * Ref);
    const double ct2 = 1.87 * 1.87;
    const double ft = tanh((ct2 * rdt) * (ct2 * rdt) * (ct2 * rdt));
    const double rdl = (RmiuL / Den) / (0.41 * 0.41 * (aDst * aDst) * Sn * Ref);
    const double fl = tanh(pow(5.00 * 5.00 * rdl, 10.));
    const double fe2 = 1. - fmax(ft, fl);
    return fe2 * fmax((fe1 - 1.), 0.);
}

// The k-omega right-hand side d6, d7 and the diagonal F1t of cell (i, j, k) = c.
__device__ __forceinline__ void sst_rhs_cell(const trans_params& p, const Fields& f, const BlkInfo& b, int i, int j,
                                             int k, long c, double& d6_out, double& d7_out, double& f1t_out)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1;
    const long tot = f.tot;
    const double dtm = p.dt_lusgs_main, Ref = p.ref;
    const double vol = f.Vol[c];
    double d6 = f.D[5 * tot + c] / vol - (3.0 * f.V[5 * tot + c] - 4.0 * f.Wt0[5 * tot + c] + f.Wt1[5 * tot + c]) * 0.5 / dtm;
    double d7 = f.D[6 * tot + c] / vol - (3.0 * f.V[6 * tot + c] - 4.0 * f.Wt0[6 * tot + c] + f.Wt1[6 * tot + c]) * 0.5 / dtm;
    const double* Q = f.DQ;
    const double dudx = Q[c], dudy = Q[tot + c], dudz = Q[2 * tot + c];
    const double dvdx = Q[3 * tot + c], dvdy = Q[4 * tot + c], dvdz = Q[5 * tot + c];
    const double dwdx = Q[6 * tot + c], dwdy = Q[7 * tot + c], dwdz = Q[8 * tot + c];
    const double dtkdx = Q[12 * tot + c], dtkdy = Q[13 * tot + c], dtkdz = Q[14 * tot + c];
    const double dtodx = Q[15 * tot + c], dtody = Q[16 * tot + c], dtodz = Q[17 * tot + c];
    const double Div = dudx + dvdy + dwdz;
    const double Votg = sqr(dudy - dvdx) + sqr(dvdz - dwdy) + sqr(dwdx - dudz);
    const double prod = (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sqr(dudy + dvdx) + sqr(dvdz + dwdy) +
                         sqr(dwdx + dudz) - 2.0 / 3.0 * Div * Div);
    const double cdif = dtkdx * dtodx + dtkdy * dtody + dtkdz * dtodz;
    double tprod = fmin(prod, Votg) * f.Rmiu[c] / Ref;
    const double ARo = f.V[c], ARTk = f.V[5 * tot + c], ARTO = f.V[6 * tot + c];
    const double tprodl1 = p.prod_lmt * p.beta_star * ARTk * ARTO / ARo * Ref;
    tprod = fmin(tprod, tprodl1);
    double CnegTk, CnegTO;
    if (p.kind_model == 2) {  // WD+ (Time/TURBTWO_TIME_LUSGS_SST.F90:418-547)
        double SourceP = tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
        double SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) - p.beta_star * ARTO / ARo * Ref;
        d6 = d6 + (SourceP + SourceN1 * ARTk);
        CnegTk = fabs(SourceN1);
        tprod = fmin(prod, Votg) / Ref;
        // W(II,JJ) W(JJ,KK) S(KK,II) summed with II outermost, zero entries included
        const double w12 = 0.5 * (dudy - dvdx), w13 = 0.5 * (dudz - dwdx), w23 = 0.5 * (dvdz - dwdy);
        const double W[3][3] = {{0., w12, w13}, {-w12, 0., w23}, {-w13, -w23, 0.}};
        const double s12 = 0.5 * (dudy + dvdx), s13 = 0.5 * (dudz + dwdx), s23 = 0.5 * (dvdz + dwdy);
        const double S[3][3] = {{0.5 * (dudx + dudx), s12, s13},
                                {s12, 0.5 * (dvdy + dvdy), s23},
                                {s13, s23, 0.5 * (dwdz + dwdz)}};
        double WWSg = 0.;
#pragma unroll
        for (int II = 0; II < 3; II++)
#pragma unroll
            for (int JJ = 0; JJ < 3; JJ++)
#pragma unroll
                for (int KK = 0; KK < 3; KK++) WWSg = WWSg + W[II][JJ] * W[JJ][KK] * S[KK][II];
        const double aK_omg = fabs(WWSg / pow(0.09 * ARTO / ARo * Ref, 3.));
        const double FBeta = (1. + 85. * aK_omg) / (1. + 100. * aK_omg);
        const double Beta_06 = p.beta1 * FBeta;
        const double dltCD = cdif <= 0. ? 0. : 0.125;
        SourceP = p.gama1 * ARo * tprod - 2.0 / 3.0 * p.gama1 * fmin(Div, 0.0) * ARTO;
        SourceN1 = -2.0 / 3.0 * p.gama1 * fmax(Div, 0.0) - Beta_06 * ARTO / ARo * Ref;
        const double SourceCross = dltCD * ARo * ARo / ARTO / Ref * cdif;
        d7 = d7 + (SourceP + SourceN1 * ARTO + SourceCross);
        CnegTO = 2.0 * fabs(SourceN1);
    } else {  // SST (:139-415)
        const double arg1 = sst_fc1(p, f, c);
        double Fdes = 1.0;
        double alIDDES = 0.0;  // IDDES length scale (Kind_SubHY 5)
        if (p.kind_hybrid == 3) {
            // DES/DDES destruction multiplier (:192-214).  The reference reads Vomega and arg1
            // there before assigning them in the cell, i.e. the values left by the previous cell
            // of its K, J, I loop (for a block's first cell they are unset: its own are taken)
            const double RmiuL = sutherland(p, f.T[c]);  // mean-flow form (:185)
            const double aLt = sqrt(ARTk / ARo) / (p.beta_star * ARTO / ARo * Ref);
            long cp = c;
            if (i > 1)
                cp = cidx(b, i - 1, j, k);
            else if (j > 1)
                cp = cidx(b, NI1, j - 1, k);
            else if (k > 1)
                cp = cidx(b, NI1, NJ1, k - 1);
            const double Vomega_prev = f.V[6 * tot + cp] / f.V[cp];
            const double arg1_prev = cp == c ? arg1 : sst_fc1(p, f, cp);
            const double Cdiff_DDES_S = (2. * 0.856 / Vomega_prev) / Ref * cdif;
            const double Fsst_DES = fsst_des(p, ARo, ARTk / ARo, ARTO / ARo, f.Dst[c], RmiuL, Cdiff_DDES_S);
            const double aLgrd = ((1. - arg1_prev) * 0.61 + arg1_prev * 0.78) * f.aLeng[c];
            Fdes = fmax(aLt * (1. - Fsst_DES) / aLgrd, 1.);
            if (p.kind_subhy == 5) {  // IDDES (:218-247); its k destruction replaces Fdes's below
                const double aLmax = f.aLeng[c], aDst = f.Dst[c], aLmin = f.aSeng[c];
                const double Cdes = (1. - arg1_prev) * 0.61 + arg1_prev * 0.78;
                const double dlt_IDDES = fmin(fmax(fmax(0.15 * aLmax, 0.15 * aDst), aLmin), aLmax);
                const double aLgrd_i = Cdes * dlt_IDDES;
                double Sn = dudx * dudx + dudy * dudy + dudz * dudz + dvdx * dvdx + dvdy * dvdy + dvdz * dvdz +
                            dwdx * dwdx + dwdy * dwdy + dwdz * dwdz;
                Sn = fmax(sqrt(Sn), 1.e-10);
                const double RmiuT = f.Rmiu[c];
                const double fd = iddes_fd(ARo, RmiuT, aDst, aLmax, Sn, Ref);
                const double fe = iddes_fe(ARo, RmiuT, RmiuL, aDst, aLmax, Sn, Ref);
                alIDDES = (1. + fe) * fd * aLt + (1. - fd) * aLgrd_i;
            }
        }
        // IF_Comppress 0 (:282-330) or 1, Suzen's compressibility correction with the turbulent
        // Mach number Mt = sqrt(k) / a and Alf1 = 1, Alf2 = 0.4, Alf3 = 0.2 (:331-396)
        const bool suzen = p.if_comppress == 1;
        double Mt = 0.;
        if (suzen) {
            const double AA = sqrt(1.4 * f.PP[c] / ARo);
            Mt = sqrt(ARTk / ARo) / AA;
        }
        const double GamaAB = p.gama1 * arg1 + p.gama2 * (1.0 - arg1);
        const double BetaAB = p.beta1 * arg1 + p.beta2 * (1.0 - arg1);
        double SourceP, SourceN1;
        if (suzen) {
            SourceP = (1. - (1. - arg1) * 0.4 * Mt * Mt) * tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
            SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) -
                       (1. + (1. - arg1) * (1.0 - 0.2) * Mt * Mt) * p.beta_star * ARTO / ARo * Ref * Fdes;
        } else {
            SourceP = tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
            SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) - p.beta_star * ARTO / ARo * Ref * Fdes;
        }
        SourceP = SourceP * 1.0;  // f_r1 = 1 (Kind_RotMod = 0)
        double Source = SourceP + SourceN1 * ARTk;
        CnegTk = fabs(SourceN1);
        if (p.kind_hybrid == 3 && p.kind_subhy == 5) {  // (:299-311, :354-366)
            SourceN1 = suzen ? -2.0 / 3.0 * fmax(Div, 0.0) -
                                   (1. + (1. - arg1) * (1.0 - 0.2) * Mt * Mt) * sqrt(ARTk / ARo) / alIDDES
                             : -2.0 / 3.0 * fmax(Div, 0.0) - sqrt(ARTk / ARo) / alIDDES;
            Source = SourceP + SourceN1 * ARTk;
            CnegTk = fabs(-2.0 / 3.0 * fmax(Div, 0.0) - 1.5 * sqrt(ARTk / ARo) / alIDDES);
        }
        d6 = d6 + Source;
        tprod = fmin(prod, Votg) / Ref;
        if (suzen) {
            SourceP = (GamaAB + (1. - arg1) * 0.4 * Mt * Mt) * ARo * tprod - 2.0 / 3.0 * GamaAB * fmin(Div, 0.0) * ARTO;
            SourceN1 = -2.0 / 3.0 * GamaAB * fmax(Div, 0.0) -
                       (1. - (1. - arg1) * (1.0 - 0.2) * Mt * Mt) * BetaAB * ARTO / ARo * Ref;
        } else {
            SourceP = GamaAB * ARo * tprod - 2.0 / 3.0 * GamaAB * fmin(Div, 0.0) * ARTO;
            SourceN1 = -2.0 / 3.0 * GamaAB * fmax(Div, 0.0) - BetaAB * ARTO / ARo * Ref;
        }
        SourceP = SourceP * 1.0;
        Source = SourceP + SourceN1 * ARTO;
        double SourceCross = 2.0 * (1 - arg1) * ARo * p.sigo2 / (ARTO / ARo) * cdif / Ref;
        SourceCross = copysign(1.0, SourceCross) * fmin(fabs(SourceCross), fabs(SourceP) * 0.5);
        Source = Source + SourceCross;
        d7 = d7 + Source;
        CnegTO = 2.0 * fabs(SourceN1);
    }
    const double Dt = f.Dtm[c];
    const double stept = 2.0 * Dt / (3.0 * Dt / dtm + 2.0);
    const double f1t = 1. + (f.Rds[c] + f.Rds[tot + c] + f.Rds[2 * tot + c]) * (stept / vol) + (CnegTk + CnegTO) * stept;
    d6_out = d6;
    d7_out = d7;
    f1t_out = f1t;
}

__global__ void __launch_bounds__(256) k_sst_rhs(trans_params p, Fields f, const BlkInfo* __restrict__ blks,
                                                 const Seg* __restrict__ seg)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k);
    const long tot = f.tot;
    double d6, d7, f1t;
    sst_rhs_cell(p, f, b, i, j, k, c, d6, d7, f1t);
    if (has_w(f)) {  // hyperplane-major inputs of the line-march sweep
        const long w = widx(b, f.march, i, j, k);
        wput(f, W_F1T, w, f1t);
        wput(f, W_D + 5, w, d6);
        wput(f, W_D + 6, w, d7);
    } else {
        f.F1t[c] = f1t;
        f.D[5 * tot + c] = d6;
        f.D[6 * tot + c] = d7;
    }
}

// The same into the hyperplane-major buffer (march axis j) tile by tile (see k_lusgs_rhs_w).
__global__ void __launch_bounds__(256) k_sst_rhs_w(trans_params p, Fields f, const BlkInfo* __restrict__ blks,
                                                   const Seg* __restrict__ seg)
{
    __shared__ double sv[3][WT * WT];
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int ti = threadIdx.x % WT, tj = threadIdx.x / WT;
    const TileCell tc = wtile_cell(s, b, ti, tj);
    if (tc.ok) {
        double d6, d7, f1t;
        sst_rhs_cell(p, f, b, tc.i, tc.j, tc.k, cidx(b, tc.i, tc.j, tc.k), d6, d7, f1t);
        sv[0][threadIdx.x] = f1t;
        sv[1][threadIdx.x] = d6;
        sv[2][threadIdx.x] = d7;
    }
    __syncthreads();
    int di, dj;
    wtile_diag(threadIdx.x, di, dj);
    const TileCell oc = wtile_cell(s, b, di, dj);
    if (!oc.ok) return;
    const long w = widx(b, 1, oc.i, oc.j, oc.k);
    wput(f, W_F1T, w, sv[0][dj * WT + di]);
    wput(f, W_D + 5, w, sv[1][dj * WT + di]);
    wput(f, W_D + 6, w, sv[2][dj * WT + di]);
}
