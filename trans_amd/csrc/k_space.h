// Residual kernels: spectral radius / local time step, reconstruction + inviscid flux,
// and the owner-computes accumulation of face fluxes into D.
#pragma once
#include "dev_common.h"

// ---------------------------------------------------------------------------------------
// Time/SpectralRadius.F90:47-180 -- one thread per interior cell.
// Rds(3) and the local pseudo time step Dtm; grid velocity zero, IF_Local_Main = 0 and
// IF_Local_Sub = 1 (Input/ReadOrg.F90:155-156).  With preconditioning (:91-120) the acoustic
// part of each direction uses the preconditioned speed sqrt(alf^2 |u|^2 + ub c^2) and the
// convective part |u.S| (1 - alf), ub raised to that direction's viscous limit.
__global__ void __launch_bounds__(256) k_spectral_radius(trans_params p, int ki, Fields f,
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
    const long sx = 1, sy = b.nx, sz = (long)b.nx * b.ny;
    const long tot = f.tot;
    const double Den = f.V[c];
    const double Vxi = f.V[tot + c] / Den;
    const double Vyi = f.V[2 * tot + c] / Den;
    const double Vzi = f.V[3 * tot + c] / Den;
    const double AAc = sqrt(1.4 * f.PP[c] / Den);
    const double RmiuL = sutherland(p, f.T[c]);
    const double RmiuTL = f.Rmiu[c];
    const double uvmu = (RmiuL + RmiuTL + 1.0e-20) * 4.0 / 3.0;
    const double uvmupr = (RmiuL / 0.72 + RmiuTL / 0.9 + 1.0e-20) * 1.4;
    const double pruvmu = fmax(uvmu, uvmupr) / Den;
    const double vol = f.Vol[c];
    double rds[3];
    const long step[3] = {sx, sy, sz};
    double alf[3] = {0.0, 0.0, 0.0}, AAcd[3] = {AAc, AAc, AAc};
    if (p.if_precondition) {
        const double UVW2 = Vxi * Vxi + Vyi * Vyi + Vzi * Vzi;
        const double ub = precond_para(p, UVW2, AAc * AAc, Den, 0.0, 0.0, 1.0);
        const double RmiuAll = RmiuL + RmiuTL;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double r = RmiuAll / Den / p.ref / f.Alagm[(long)d * tot + c] / AAc;
            const double ubd = fmax(ub, r * r);
            alf[d] = (1.0 - ubd) / 2.0;
            AAcd[d] = sqrt(alf[d] * alf[d] * UVW2 + ubd * AAc * AAc);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double* S = f.SD + (long)d * tot;  // component (d, x) at S[3*x*tot]
        double ax = 0.5 * (S[c] + S[c + step[d]]);
        double ay = 0.5 * (S[3 * tot + c] + S[3 * tot + c + step[d]]);
        double az = 0.5 * (S[6 * tot + c] + S[6 * tot + c + step[d]]);
        double st = ax * ax + ay * ay + az * az;
        double un = Vxi * ax + Vyi * ay + Vzi * az + 0.0;
        rds[d] = fabs(un) * (1.0 - alf[d]) + sqrt(st) * AAcd[d] + 2.0 * st / vol * pruvmu / p.ref;
        f.Rds[(long)d * tot + c] = rds[d];
    }
    if (f.Rdsf) {
#pragma unroll
        for (int d = 0; d < 3; d++) f.Rdsf[(long)d * tot + c] = (float)rds[d];
    }
    const double dtm_local = vol / (rds[0] + rds[1] + rds[2]);
    double dtm = p.dt_lusgs_main;
    double cfl = p.cfl_sub_start +
                 (p.cfl_sub_end - p.cfl_sub_start) * (double)(ki - p.ki_ini) / (double)p.cfl_sub_step;
    cfl = fmin(cfl, p.cfl_sub_end);
    if (p.kind_dual != 0) dtm = cfl * dtm_local;
    f.Dtm[c] = dtm;
}

// ---------------------------------------------------------------------------------------
// Space/UTIL_MUSCL.F90:2-18
__device__ __forceinline__ double alimiter(double var1, double var2)
{
    const double epsm = 1.0e-7;
    return (var1 * (var2 * var2 + 2.0 * epsm * epsm) + var2 * (2.0 * var1 * var1 + epsm * epsm)) /
           (2.0 * var1 * var1 - var1 * var2 + 2.0 * var2 * var2 + 3.0 * epsm * epsm) * 0.5 *
           fabs(copysign(1.0, var1) + copysign(1.0, var2));
}

// Space/UTIL_MUSCL.F90:21-35 (MUSCL2): smoothness-weighted blend of the two slopes.
__device__ __forceinline__ double muscl2(double var1, double var2)
{
    const double epsm = 1.e-6;
    const double smooth = (2. * var1 * var2 + epsm) / (var1 * var1 + var2 * var2 + epsm);
    return ((1. + smooth) * var1 + (1. - smooth) * var2) * smooth / 2.;
}

// Preconditioned Roe eigenvalues (Space/CONVECT_Roe_Original.F90:319-332 and
// Space/CONVECT_Roe_Rotated.F90:669-680): ub from the Roe-mean state at the Roe-mean
// temperature with the face eddy viscosity RmiuT and AlagmR, alf = (1 - ub)/2, the acoustic
// pair |un (1 - alf) +/- sqrt((alf un)^2 + ub c^2)|, no entropy fix.
__device__ __forceinline__ void roe_precond_eigs(const trans_params& p, double vm2, double am2, double rm, double un,
                                                 double RmiuT, double AlagmR, double& eig1, double& eig2,
                                                 double& eig3, double& alf, double& ubetac2, double& amrev)
{
    const double TTL = am2 / 1.4 / p.ppf;
    ubetac2 = precond_para(p, vm2, am2, rm, TTL, RmiuT, AlagmR);
    alf = (1.0 - ubetac2) / 2.0;
    const double unrev = un * (1.0 - alf);
    amrev = sqrt((alf * un) * (alf * un) + ubetac2 * am2);
    eig1 = fabs(un);
    eig2 = fabs(unrev + amrev);
    eig3 = fabs(unrev - amrev);
}

// Space/CONVECT_Roe_Rotated.F90:597-752 (that file's own Roe_Common): adds the Roe
// dissipation through the sub-face (sav1, sav2, sav3), |.| = sav, weighted by Funscheme, to
// the area-weighted flux acc.  RmiuT, AlagmR: face eddy viscosity and Alagm of the
// preconditioned eigenvalues.
__device__ __forceinline__ void roe_rot_common(const trans_params& p, double PLo, double PRo, double TLo,
                                               double TRo, double ULo, double URo, double VLo, double VRo,
                                               double WLo, double WRo, double sav1, double sav2, double sav3,
                                               double sav, double Funscheme, double rLo, double rRo,
                                               double RmiuT, double AlagmR, double acc[5])
{
    const double rcpcv = p.ppf;
    const double DENLo = PLo / (rcpcv * TLo), DENRo = PRo / (rcpcv * TRo);
    const double ga = 1.4, ga1 = 0.4;
    const double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo, VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    const double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    const double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    const double rrorl = sqrt(DENRo / DENLo), rrorlp1 = 1.0 + rrorl;
    const double rm = sqrt(DENRo * DENLo);
    const double um = (ULo + URo * rrorl) / rrorlp1, vm = (VLo + VRo * rrorl) / rrorlp1;
    const double wm = (WLo + WRo * rrorl) / rrorlp1, hm = (HLo + HRo * rrorl) / rrorlp1;
    const double vm2 = um * um + vm * vm + wm * wm;
    const double rccm = (rLo + rRo * rrorl) / rrorlp1;
    const double am2 = ga1 * fabs(hm - 0.5 * vm2 + 0.5 * (rccm * rccm));
    const double am = sqrt(am2);
    const double sav1n = sav1 / sav, sav2n = sav2 / sav, sav3n = sav3 / sav;
    const double unormaln = (sav1n * um + sav2n * vm + sav3n * wm);
    const double anormaln = am;
    double eig1 = fabs(unormaln), eig2 = fabs(unormaln + anormaln), eig3 = fabs(unormaln - anormaln);
    double alf = 0.0, ubetac2 = 1.0, amrev = am;
    if (p.if_precondition)
        roe_precond_eigs(p, vm2, am2, rm, unormaln, RmiuT, AlagmR, eig1, eig2, eig3, alf, ubetac2, amrev);
    const double dunormaln = (sav1n * (URo - ULo) + sav2n * (VRo - VLo) + sav3n * (WRo - WLo));
    const double astar = (eig2 + eig3) / 2.0;
    const double Mstar = (eig2 - eig3) / 2.0 / amrev;
    const double dW1 = DENRo - DENLo, dU1 = URo - ULo, dU2 = VRo - VLo, dU3 = WRo - WLo, dP1 = PRo - PLo;
    const double dW2 = rm * dU1 + dW1 * um, dW3 = rm * dU2 + dW1 * vm, dW4 = rm * dU3 + dW1 * wm;
    const double dW5 = dP1 / ga1 + 0.5 * vm2 * dW1 + rm * um * dU1 + rm * vm * dU2 + rm * wm * dU3;
    const double dUroe = Mstar * dunormaln + (astar - (1 - 2.0 * alf) * eig1 - alf * unormaln * Mstar) *
                                                 (PRo - PLo) / rm / (ubetac2 * am2);
    const double dProe = Mstar * (PRo - PLo) + (astar - eig1 + alf * unormaln * Mstar) * rm * dunormaln;
    const double r1[5] = {eig1 * dW1 + dUroe * rm, eig1 * dW2 + dUroe * rm * um + dProe * sav1n,
                          eig1 * dW3 + dUroe * rm * vm + dProe * sav2n, eig1 * dW4 + dUroe * rm * wm + dProe * sav3n,
                          eig1 * dW5 + dUroe * rm * hm + dProe * unormaln};
#pragma unroll
    for (int L = 0; L < 5; L++) acc[L] = acc[L] + 0.5 * r1[L] * sav * Funscheme;
}

// Space/CONVECT_Roe_Rotated.F90:84-176 for one face: the central flux through the face, then
// the Roe dissipation through the face (velocity jump at or below Tiny), or through its parts
// along and across the velocity jump.  The grid-velocity term vibc of :169 is never set in the
// routine and is taken as 0.  out = area-weighted flux.
__device__ __forceinline__ void roe_rotated_face(const trans_params& p, double PLo, double PRo, double TLo,
                                                 double TRo, double ULo, double URo, double VLo, double VRo,
                                                 double WLo, double WRo, double s1, double s2, double s3,
                                                 double grad, double Funscheme, double rLo, double rRo,
                                                 double RmiuT, double AlagmR, double out[5])
{
    const double rcpcv = p.ppf;
    const double DENLo = PLo / (rcpcv * TLo), DENRo = PRo / (rcpcv * TRo);
    const double ga = 1.4, ga1 = 0.4;
    const double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo, VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    const double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    const double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    double sav = grad;
    double sav1n = s1 / sav, sav2n = s2 / sav, sav3n = s3 / sav;
    const double ulnormaln = (sav1n * ULo + sav2n * VLo + sav3n * WLo);
    const double urnormaln = (sav1n * URo + sav2n * VRo + sav3n * WRo);
    const double rulnormaln = DENLo * ulnormaln, rurnormaln = DENRo * urnormaln;
    const double vibc = 0.0;
    out[0] = -0.5 * (rulnormaln + rurnormaln) * sav;
    out[1] = -0.5 * (rulnormaln * ULo + rurnormaln * URo + sav1n * (PLo + PRo)) * sav;
    out[2] = -0.5 * (rulnormaln * VLo + rurnormaln * VRo + sav2n * (PLo + PRo)) * sav;
    out[3] = -0.5 * (rulnormaln * WLo + rurnormaln * WRo + sav3n * (PLo + PRo)) * sav;
    out[4] = -0.5 * (rulnormaln * HLo + rurnormaln * HRo - vibc * (PLo + PRo)) * sav;
    const double DeltaU = sqrt((URo - ULo) * (URo - ULo) + (VRo - VLo) * (VRo - VLo) + (WRo - WLo) * (WRo - WLo));
    if (DeltaU <= p.tiny) {
        roe_rot_common(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, s1, s2, s3, grad, Funscheme, rLo, rRo,
                       RmiuT, AlagmR, out);
    } else {
        sav1n = (URo - ULo) / DeltaU;
        sav2n = (VRo - VLo) / DeltaU;
        sav3n = (WRo - WLo) / DeltaU;
        sav = sav1n * s1 + sav2n * s2 + sav3n * s3;
        sav1n = copysign(1.0, sav) * sav1n;
        sav2n = copysign(1.0, sav) * sav2n;
        sav3n = copysign(1.0, sav) * sav3n;
        sav = fabs(sav);
        const double a1 = sav1n * sav, a2 = sav2n * sav, a3 = sav3n * sav;
        const double b1 = s1 - a1, b2 = s2 - a2, b3 = s3 - a3;
        roe_rot_common(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, a1, a2, a3,
                       fmax(sqrt(a1 * a1 + a2 * a2 + a3 * a3), p.tiny), Funscheme, rLo, rRo, RmiuT, AlagmR, out);
        roe_rot_common(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, b1, b2, b3,
                       fmax(sqrt(b1 * b1 + b2 * b2 + b3 * b3), p.tiny), Funscheme, rLo, rRo, RmiuT, AlagmR, out);
    }
}

// Space/CONVECT_Roe_Original.F90:242-407.  Returns the flux per unit area (droR..dreR).
// RmiuT, AlagmR: face eddy viscosity and Alagm of the preconditioned eigenvalues.
__device__ __forceinline__ void roe_common(const trans_params& p, double PLo, double PRo,
                                           double TLo, double TRo, double ULo, double URo,
                                           double VLo, double VRo, double WLo, double WRo,
                                           double sav1, double sav2, double sav3, double sav,
                                           double rLo, double rRo, double RmiuT, double AlagmR, double out[5])
{
    const double rcpcv = p.ppf;
    const double DENLo = PLo / (rcpcv * TLo);
    const double DENRo = PRo / (rcpcv * TRo);
    const double ga = 1.4, ga1 = 1.4 - 1.0;
    const double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo;
    const double VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    const double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    const double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    const double rrorl = sqrt(DENRo / DENLo);
    const double rrorlp1 = 1.0 + rrorl;
    const double rm = sqrt(DENRo * DENLo);
    const double um = (ULo + URo * rrorl) / rrorlp1;
    const double vm = (VLo + VRo * rrorl) / rrorlp1;
    const double wm = (WLo + WRo * rrorl) / rrorlp1;
    const double hm = (HLo + HRo * rrorl) / rrorlp1;
    const double vm2 = um * um + vm * vm + wm * wm;
    const double rccm = (rLo + rrorl * rRo) / rrorlp1;
    const double am2 = ga1 * fabs(hm - 0.5 * vm2 + 0.5 * (rccm * rccm));
    const double am = sqrt(am2);
    const double s1 = sav1 / sav, s2 = sav2 / sav, s3 = sav3 / sav;
    const double un = (s1 * um + s2 * vm + s3 * wm);
    double eig1, eig2, eig3, alf = 0.0, ubetac2 = 1.0, amrev = am;
    if (p.if_precondition) {
        roe_precond_eigs(p, vm2, am2, rm, un, RmiuT, AlagmR, eig1, eig2, eig3, alf, ubetac2, amrev);
    } else {
        eig1 = fabs(un);
        eig2 = fabs(un + am);
        eig3 = fabs(un - am);
        if (eig1 < 0.1) eig1 = (eig1 * eig1 + 0.1 * 0.1) / 0.2;
        if (eig2 < 0.1) eig2 = (eig2 * eig2 + 0.1 * 0.1) / 0.2;
        if (eig3 < 0.1) eig3 = (eig3 * eig3 + 0.1 * 0.1) / 0.2;
    }
    const double dun = (s1 * (URo - ULo) + s2 * (VRo - VLo) + s3 * (WRo - WLo));
    const double astar = (eig2 + eig3) / 2.0;
    const double Mstar = (eig2 - eig3) / 2.0 / amrev;
    const double dW1 = DENRo - DENLo;
    const double dU1 = URo - ULo, dU2 = VRo - VLo, dU3 = WRo - WLo;
    const double dP1 = PRo - PLo;
    const double dW2 = rm * dU1 + dW1 * um;
    const double dW3 = rm * dU2 + dW1 * vm;
    const double dW4 = rm * dU3 + dW1 * wm;
    const double dW5 = dP1 / ga1 + 0.5 * vm2 * dW1 + rm * um * dU1 + rm * vm * dU2 + rm * wm * dU3;
    double dUroe, dProe;
    if (p.if_precondition) {
        dUroe = Mstar * dun + (astar - (1 - 2.0 * alf) * eig1 - alf * un * Mstar) * (PRo - PLo) / rm / (ubetac2 * am2);
        dProe = Mstar * (PRo - PLo) + (astar - eig1 + alf * un * Mstar) * rm * dun;
    } else {
        dUroe = Mstar * dun + (astar - 1.0 * eig1) * (PRo - PLo) / rm / (1.0 * am2);
        dProe = Mstar * (PRo - PLo) + (astar - eig1) * rm * dun;
    }
    const double droR1 = eig1 * dW1 + dUroe * rm;
    const double drxR1 = eig1 * dW2 + dUroe * rm * um + dProe * s1;
    const double dryR1 = eig1 * dW3 + dUroe * rm * vm + dProe * s2;
    const double drzR1 = eig1 * dW4 + dUroe * rm * wm + dProe * s3;
    const double dreR1 = eig1 * dW5 + dUroe * rm * hm + dProe * un;
    const double uln = (s1 * ULo + s2 * VLo + s3 * WLo);
    const double urn = (s1 * URo + s2 * VRo + s3 * WRo);
    const double ruln = DENLo * uln;
    const double rurn = DENRo * urn;
    out[0] = -0.5 * (ruln + rurn - droR1);
    out[1] = -0.5 * (ruln * ULo + rurn * URo + s1 * (PLo + PRo) - drxR1);
    out[2] = -0.5 * (ruln * VLo + rurn * VRo + s2 * (PLo + PRo) - dryR1);
    out[3] = -0.5 * (ruln * WLo + rurn * WRo + s3 * (PLo + PRo) - drzR1);
    out[4] = -0.5 * (ruln * HLo + rurn * HRo - dreR1);
}

// Primitive set of a cell: u, v, w, p, T and the radius.
struct Prim {
    double q[6];
};

// u, v, w from Q (formed once per cell per residual: the same divisions), p, T, rad
__device__ __forceinline__ Prim load_prim(const Fields& f, long c)
{
    Prim r;
    const long tot = f.tot;
    r.q[0] = f.Q[c];
    r.q[1] = f.Q[tot + c];
    r.q[2] = f.Q[2 * tot + c];
    r.q[3] = f.PP[c];
    r.q[4] = f.T[c];
    r.q[5] = f.rad[c];
    return r;
}

// Face value of SPACEINTERP_GetSurfCenterVars (Space/SPACEINTERP_GetSurfCenterVars.F90:27-56):
// Alagm-weighted average of cells g-1 (qm) and g (qp).
__device__ __forceinline__ double face_avg(double am, double ap, double qm, double qp)
{
    const double alsinv = 1.0 / (ap + am);
    return (ap * qm + am * qp) * alsinv;
}
