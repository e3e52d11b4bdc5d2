// Ghost-cell fills: physical BCs (wall, far field, symmetry, inlet, outlet) and the
// block-interface / periodic halo copy, all in one launch over a table of BC patches.
// Patches write disjoint ghost cells and read only interior cells (or ghosts they wrote
// themselves), so the reference's sequential patch order (BC/Boundary.F90:22-47) does not
// change any result.
#pragma once
#include "dev_common.h"

struct DPatch {
    trans_patch p;
    int blk;    // owning block
    int tpatch; // global index of the target patch (interfaces)
    int count;  // face cells of the patch
    int pad;
};

struct BcSeg {
    int patch;
    int base;
};

__device__ __forceinline__ long gidx(const BlkInfo& b, int i, int j, int k)
{
    return cidx(b, i, j, k);
}

// BC/BC_far.F90:135-207
__device__ __forceinline__ void farbond(double XM, double P2, double RR, double Vxi, double Vyi,
                                        double Vzi, double DNx, double DNy, double DNz, double VxF,
                                        double VyF, double VzF, double& PP0, double& RR0, double& VX0,
                                        double& VY0, double& VZ0, double& dir)
{
    const double pGam = 1.4;
    const double AF = 1. / XM;
    const double AE = sqrt(pGam * P2 / RR);
    const double QNE = Vxi * DNx + Vyi * DNy + Vzi * DNz;
    const double QNF = VxF * DNx + VyF * DNy + VzF * DNz;
    const double RF = QNF - 5. * AF;
    const double RE = QNE + 5. * AE;
    const double QN = 0.5 * (RE + RF);
    const double AS = 0.1 * (RE - RF);
    if (QN > -AS && QN < 0.) {
        const double SON = 1. / pGam / (XM * XM);
        VX0 = (VxF - QNF * DNx) + QN * DNx;
        VY0 = (VyF - QNF * DNy) + QN * DNy;
        VZ0 = (VzF - QNF * DNz) + QN * DNz;
        RR0 = pow(AS * AS / SON / pGam, 2.5);
        PP0 = AS * AS * RR0 / pGam;
        dir = -1.;
    }
    if (QN >= 0. && QN < AS) {
        const double SON = P2 / pow(RR, pGam);
        VX0 = (Vxi - QNE * DNx) + QN * DNx;
        VY0 = (Vyi - QNE * DNy) + QN * DNy;
        VZ0 = (Vzi - QNE * DNz) + QN * DNz;
        RR0 = pow(AS * AS / SON / pGam, 2.5);
        PP0 = AS * AS * RR0 / pGam;
        dir = 1.;
    }
    if (QN <= -AS) {
        VX0 = VxF;
        VY0 = VyF;
        VZ0 = VzF;
        RR0 = 1.0;
        PP0 = 1. / pGam / XM / XM;
        dir = -1.;
    }
    if (QN >= AS) {
        VX0 = Vxi;
        VY0 = Vyi;
        VZ0 = Vzi;
        RR0 = RR;
        PP0 = P2;
        dir = 1.;
    }
}

// Geom/startup.F90:1-25
__device__ __forceinline__ void ijk_start_end(int I, int J, int K, const int* tl, const int* tt,
                                              int lcross, int& Is, int& Js, int& Ks)
{
    const int ft1 = tl[0], ft2 = tl[1], ft3 = tl[2];
    const int gt1 = tt[0], gt2 = tt[1], gt3 = tt[2];
    const int lc1 = 1 - lcross;
    const int t1 = ft1 * I + ft2 * J + ft3 * K;
    const int t2 = ft1 * (J * lc1 + K * lcross) + ft2 * (I * lc1 + K * lcross) + ft3 * (I * lc1 + J * lcross);
    const int t3 = ft1 * (J * lcross + K * lc1) + ft2 * (I * lcross + K * lc1) + ft3 * (I * lcross + J * lc1);
    Is = gt1 * t1 + (gt2 + gt3) * t2;
    Js = gt2 * t1 + gt1 * t2 + gt3 * t3;
    Ks = gt3 * t1 + (gt1 + gt2) * t3;
}

__device__ __forceinline__ const int* lay_src(const trans_patch& c, int lay)
{
    return lay == 0 ? c.cijkadd : (lay == 1 ? c.cijkadd1 : c.cijkadd2);
}
__device__ __forceinline__ const int* lay_dst(const trans_patch& c, int lay)
{
    return lay == 0 ? c.ceijkadd : (lay == 1 ? c.ceijkadd1 : c.ceijkadd2);
}

__global__ void __launch_bounds__(128) k_bc(trans_params p, int if_turb, int sweep, Fields f,
                                            const BlkInfo* __restrict__ blks,
                                            const DPatch* __restrict__ pt, const BcSeg* __restrict__ seg)
{
    const BcSeg sg = seg[blockIdx.x];
    const DPatch& dp = pt[sg.patch];
    const int lin = sg.base + threadIdx.x;
    const trans_patch& c = dp.p;
    const int n0 = c.point_end[0] - c.point_start[0] + 1;
    const int n1 = c.point_end[1] - c.point_start[1] + 1;
    const int n2 = c.point_end[2] - c.point_start[2] + 1;
    if (lin >= n0 * n1 * n2) return;
    const BlkInfo b = blks[dp.blk];
    const long tot = f.tot;
    const int i = c.point_start[0] + lin % n0;
    const int j = c.point_start[1] + (lin / n0) % n1;
    const int k = c.point_start[2] + lin / (n0 * n1);
    const int* ad = c.cijkadd;
    const int* e = c.ceijkadd;

    if (sweep == 0) {
        if (c.bctype == 9) {
            const DPatch& tp = pt[dp.tpatch];
            const BlkInfo br = blks[tp.blk];
            const trans_patch& bc = tp.p;
            const int IIadd = c.point_end[0] - c.point_start[0] + 2;
            const int JJadd = c.point_end[1] - c.point_start[1] + 2;
            const int KKadd = c.point_end[2] - c.point_start[2] + 2;
            int iL = i - c.point_start[0] + 1, jL = j - c.point_start[1] + 1, kL = k - c.point_start[2] + 1;
            iL = c.rev_matrix[0] * IIadd + (1 - 2 * c.rev_matrix[0]) * iL;
            jL = c.rev_matrix[1] * JJadd + (1 - 2 * c.rev_matrix[1]) * jL;
            kL = c.rev_matrix[2] * KKadd + (1 - 2 * c.rev_matrix[2]) * kL;
            int ib, jb, kb;
            ijk_start_end(iL, jL, kL, c.transform, c.transform_t, c.lcross, ib, jb, kb);
            ib += bc.point_start[0] - 1;
            jb += bc.point_start[1] - 1;
            kb += bc.point_start[2] - 1;
            const long src = gidx(br, ib + bc.cijkadd[0], jb + bc.cijkadd[1], kb + bc.cijkadd[2]);
            const long dst = gidx(b, i + e[0], j + e[1], k + e[2]);
            const double sint = sin(c.rotate_angle), cost = cos(c.rotate_angle);
            const double cos2t = cost * cost, sin2t = sint * sint, sct = sint * cost;
            const double* q = f.DQ;
            double* o = f.DQ;
            const double ugx = q[src], ugy = q[tot + src], ugz = q[2 * tot + src];
            const double vgx = q[3 * tot + src], vgy = q[4 * tot + src], vgz = q[5 * tot + src];
            const double wgx = q[6 * tot + src], wgy = q[7 * tot + src], wgz = q[8 * tot + src];
            o[dst] = ugx;
            o[tot + dst] = cost * ugy + sint * ugz;
            o[2 * tot + dst] = -sint * ugy + cost * ugz;
            o[3 * tot + dst] = cost * vgx + sint * wgx;
            o[4 * tot + dst] = cos2t * vgy + sct * (vgz + wgy) + sin2t * wgz;
            o[5 * tot + dst] = cos2t * vgz + sct * (-vgy + wgz) - sin2t * wgy;
            o[6 * tot + dst] = -sint * vgx + cost * wgx;
            o[7 * tot + dst] = -sin2t * vgz + sct * (-vgy + wgz) + cos2t * wgy;
            o[8 * tot + dst] = sin2t * vgy + sct * (vgz + wgy) + cos2t * wgz;
            for (int L = 9; L < 18; L += 3) {
                o[(long)L * tot + dst] = q[(long)L * tot + src];
                const double gy = q[(long)(L + 1) * tot + src], gz = q[(long)(L + 2) * tot + src];
                o[(long)(L + 1) * tot + dst] = cost * gy + sint * gz;
                o[(long)(L + 2) * tot + dst] = -sint * gy + cost * gz;
            }
        } else if (c.bctype == 1 || c.bctype == 2 || c.bctype == 3 || c.bctype == 5 || c.bctype == 7) {
            const long src = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
            const long dst = gidx(b, i + e[0], j + e[1], k + e[2]);
            for (int L = 0; L < 18; L++) f.DQ[(long)L * tot + dst] = f.DQ[(long)L * tot + src];
        }
        return;
    }
    if (sweep != -1) return;

    switch (c.bctype) {
    case 1: {  // BC/BC_Wall.F90:41-120
        long is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        const double ztemp = 0.5 * (f.Zc[is] + f.Zc[il]);
        const double ytemp = 0.5 * (f.Yc[is] + f.Yc[il]);
        const double vibc = 0.0;
        const double vjbc = (double)c.subbctype * p.omega1 * ztemp;
        const double vkbc = -(double)c.subbctype * p.omega1 * ytemp;
        for (int lay = 0; lay < 3; lay++) {
            const int* sa = lay_src(c, lay);
            const int* ea = lay_dst(c, lay);
            const long s2 = gidx(b, i + sa[0], j + sa[1], k + sa[2]);
            const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            const double pp = f.PP[s2], tt = f.T[s2];
            f.PP[d2] = pp;
            f.T[d2] = tt;
            const double den = pp / (tt + p.tiny) * p.rxm2;
            f.V[d2] = den;
            f.V[tot + d2] = vibc * den;
            f.V[2 * tot + d2] = vjbc * den;
            f.V[3 * tot + d2] = vkbc * den;
            if (lay == 0)
                for (int d = 0; d < 3; d++) f.Rds[(long)d * tot + d2] = f.Rds[(long)d * tot + s2];
            if (lay == 0 && f.Rdsf)
                for (int d = 0; d < 3; d++) f.Rdsf[(long)d * tot + d2] = f.Rdsf[(long)d * tot + s2];
        }
        if (if_turb) {
            const double Dn = f.Dst[il];
            const double Den = f.V[il];
            const double TT = f.T[il];
            double Rmiuw = sutherland(p, TT);
            const double Omegaw = 60.0 * Rmiuw / (Den * 0.075 * Dn * Dn * p.ref * p.ref);
            Rmiuw = f.Rmiu[is];
            for (int lay = 0; lay < 3; lay++) {
                const int* ea = lay_dst(c, lay);
                const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
                f.V[5 * tot + d2] = 0.0;
                f.V[6 * tot + d2] = Omegaw;
                f.Rmiu[d2] = Rmiuw;
            }
        }
        break;
    }
    case 2: {  // BC/BC_far.F90:40-118
        const int dn = c.ijkdirect - 1;
        const double out_n = (double)(2 * c.left_or_right - 4);
        long is = gidx(b, i + c.ijkadd[0], j + c.ijkadd[1], k + c.ijkadd[2]);
        const double g = f.Grad[(long)dn * tot + is];
        const double* S = f.SD + (long)dn * tot;
        const double Sout1 = S[is] / g, Sout2 = S[3 * tot + is] / g, Sout3 = S[6 * tot + is] / g;
        const double Dnx = Sout1 * out_n, Dny = Sout2 * out_n, Dnz = Sout3 * out_n;
        is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const double Den = f.V[is];
        const double Vxi = f.V[tot + is] / Den, Vyi = f.V[2 * tot + is] / Den, Vzi = f.V[3 * tot + is] / Den;
        const double P2 = f.PP[is];
        double PP0 = 0, Den0 = 0, Vx0 = 0, Vy0 = 0, Vz0 = 0, dir = 0;
        farbond(p.xm, P2, Den, Vxi, Vyi, Vzi, Dnx, Dny, Dnz, p.vxf, p.vyf, p.vzf, PP0, Den0, Vx0, Vy0,
                Vz0, dir);
        const long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        const double den = 2 * Den0 - Den;
        f.V[il] = den;
        f.V[tot + il] = 2.0 * Vx0 - Vxi;
        f.V[2 * tot + il] = 2.0 * Vy0 - Vyi;
        f.V[3 * tot + il] = 2.0 * Vz0 - Vzi;
        const double pg = 2.0 * PP0 - P2;
        f.PP[il] = pg;
        f.T[il] = p.rxm2 * pg / den;
        for (int lay = 1; lay < 3; lay++) {
            const int* ea = lay_dst(c, lay);
            const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            for (int L = 0; L < 5; L++) f.V[(long)L * tot + d2] = f.V[(long)L * tot + il];
            f.PP[d2] = f.PP[il];
            f.T[d2] = f.T[il];
        }
        if (if_turb) {
            if (dir > 0.0) {
                f.V[5 * tot + il] = f.V[5 * tot + is];
                f.V[6 * tot + il] = f.V[6 * tot + is];
                f.Rmiu[il] = f.Rmiu[is];
            } else {
                const double TT = f.T[il];
                const double Dn = f.V[il];
                const double RmiuFar = sutherland(p, TT);
                const double rm = 2.0 * p.visratio * RmiuFar - f.Rmiu[is];
                f.Rmiu[il] = rm;
                const double v6 = 2.0 * 1.5 * p.fsti * p.fsti * Dn - f.V[5 * tot + is];
                f.V[5 * tot + il] = v6;
                f.V[6 * tot + il] = v6 / (rm + p.tiny) * Dn;
            }
            for (int lay = 1; lay < 3; lay++) {
                const int* ea = lay_dst(c, lay);
                const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
                f.V[5 * tot + d2] = f.V[5 * tot + il];
                f.V[6 * tot + d2] = f.V[6 * tot + il];
                f.Rmiu[d2] = f.Rmiu[il];
            }
        }
        break;
    }
    case 3: {  // BC/BC_Symm.F90:42-110 (layer 1 keeps its density)
        const int dn = c.ijkdirect - 1;
        long is = gidx(b, i + c.ijkadd[0], j + c.ijkadd[1], k + c.ijkadd[2]);
        const double g = f.Grad[(long)dn * tot + is];
        const double* S = f.SD + (long)dn * tot;
        const double Sout1 = S[is] / g, Sout2 = S[3 * tot + is] / g, Sout3 = S[6 * tot + is] / g;
        is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const double den = f.V[is];
        const double Vc1 = f.V[tot + is] / den * Sout1 + f.V[2 * tot + is] / den * Sout2 +
                           f.V[3 * tot + is] / den * Sout3;
        for (int lay = 0; lay < 3; lay++) {
            const int* sa = lay_src(c, lay);
            const int* ea = lay_dst(c, lay);
            const long s2 = gidx(b, i + sa[0], j + sa[1], k + sa[2]);
            const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            const double r1 = f.V[s2];
            if (lay != 1) f.V[d2] = r1;
            f.V[tot + d2] = f.V[tot + s2] - 2 * Sout1 * Vc1 * r1;
            f.V[2 * tot + d2] = f.V[2 * tot + s2] - 2 * Sout2 * Vc1 * r1;
            f.V[3 * tot + d2] = f.V[3 * tot + s2] - 2 * Sout3 * Vc1 * r1;
            f.V[4 * tot + d2] = f.V[4 * tot + s2];
            f.PP[d2] = f.PP[s2];
            f.T[d2] = f.T[s2];
            if (if_turb) {
                f.V[5 * tot + d2] = f.V[5 * tot + s2];
                f.V[6 * tot + d2] = f.V[6 * tot + s2];
                f.Rmiu[d2] = f.Rmiu[s2];
            }
        }
        break;
    }
    case 5: {  // BC/BC_Inlet.F90:38-107
        const long is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        double PsIn = f.PP[is];
        PsIn = fmin(PsIn, 0.999 * p.p0in);
        const double comm = fmax(1.0, pow(p.p0in / PsIn, 1.0 / 3.5));
        const double mach = sqrt((comm - 1.0) * 2 / 0.4);
        const double TsIn = p.t0in / comm;
        const double Roin = p.rxm2 * PsIn / TsIn;
        const double asonic = sqrt(1.4 * PsIn / Roin);
        const double Vabs = mach * asonic;
        const double VxInlet = Vabs;
        const double VyInlet = p.omega1 * (f.Zc[is]);
        const double VzInlet = -p.omega1 * f.Yc[is];
        const double v2 = Roin * VxInlet, v3 = Roin * VyInlet, v4 = Roin * VzInlet;
        f.V[il] = Roin;
        f.V[tot + il] = v2;
        f.V[2 * tot + il] = v3;
        f.V[3 * tot + il] = v4;
        f.V[4 * tot + il] = PsIn / 0.4 + 0.5 * (v2 * v2 + v3 * v3 + v4 * v4) / Roin -
                            0.5 * sqr(p.omega1 * f.rad[il]) * Roin;
        f.PP[il] = PsIn;
        f.T[il] = TsIn;
        for (int d = 0; d < 3; d++) f.Rds[(long)d * tot + il] = f.Rds[(long)d * tot + is];
        if (f.Rdsf)
            for (int d = 0; d < 3; d++) f.Rdsf[(long)d * tot + il] = f.Rdsf[(long)d * tot + is];
        for (int lay = 1; lay < 3; lay++) {
            const int* ea = lay_dst(c, lay);
            const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            for (int L = 0; L < 5; L++) f.V[(long)L * tot + d2] = f.V[(long)L * tot + il];
            f.PP[d2] = PsIn;
            f.T[d2] = TsIn;
        }
        if (if_turb) {
            const double TT = f.T[il];
            const double RmiuIn = sutherland(p, TT);
            const double rm = p.visratio * RmiuIn;
            f.Rmiu[il] = rm;
            const double Den = f.V[il];
            const double v6 = 1.5 * p.fsti * p.fsti * Den;
            f.V[5 * tot + il] = v6;
            f.V[6 * tot + il] = v6 / (rm + p.tiny) * Den;
            for (int lay = 1; lay < 3; lay++) {
                const int* ea = lay_dst(c, lay);
                const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
                f.V[5 * tot + d2] = f.V[5 * tot + il];
                f.V[6 * tot + d2] = f.V[6 * tot + il];
                f.Rmiu[d2] = f.Rmiu[il];
            }
        }
        break;
    }
    case 7: {  // outlet ghosts come from k_bc_outlet; the layer-0 Rds copy is here
        const long is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        for (int d = 0; d < 3; d++) f.Rds[(long)d * tot + il] = f.Rds[(long)d * tot + is];
        if (f.Rdsf)
            for (int d = 0; d < 3; d++) f.Rdsf[(long)d * tot + il] = f.Rdsf[(long)d * tot + is];
        break;
    }
    case 9: {  // BC/BC_P2P.F90:76-139
        const DPatch& tp = pt[dp.tpatch];
        const BlkInfo br = blks[tp.blk];
        const trans_patch& bc = tp.p;
        const int IIadd = c.point_end[0] - c.point_start[0] + 2;
        const int JJadd = c.point_end[1] - c.point_start[1] + 2;
        const int KKadd = c.point_end[2] - c.point_start[2] + 2;
        int iL = i - c.point_start[0] + 1, jL = j - c.point_start[1] + 1, kL = k - c.point_start[2] + 1;
        iL = c.rev_matrix[0] * IIadd + (1 - 2 * c.rev_matrix[0]) * iL;
        jL = c.rev_matrix[1] * JJadd + (1 - 2 * c.rev_matrix[1]) * jL;
        kL = c.rev_matrix[2] * KKadd + (1 - 2 * c.rev_matrix[2]) * kL;
        int ib, jb, kb;
        ijk_start_end(iL, jL, kL, c.transform, c.transform_t, c.lcross, ib, jb, kb);
        ib += bc.point_start[0] - 1;
        jb += bc.point_start[1] - 1;
        kb += bc.point_start[2] - 1;
        const double sint = sin(c.rotate_angle), cost = cos(c.rotate_angle);
        int pr0, pr1, pr2;
        ijk_start_end(1, 2, 3, c.transform, c.transform_t, c.lcross, pr0, pr1, pr2);
        for (int lay = 0; lay < 3; lay++) {
            const int* sa = lay_src(bc, lay);
            const int* ea = lay_dst(c, lay);
            const long src = gidx(br, ib + sa[0], jb + sa[1], kb + sa[2]);
            const long dst = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            f.V[dst] = f.V[src];
            f.V[tot + dst] = f.V[tot + src];
            const double txy = f.V[2 * tot + src], txz = f.V[3 * tot + src];
            f.V[2 * tot + dst] = cost * txy + sint * txz;
            f.V[3 * tot + dst] = -sint * txy + cost * txz;
            for (int L = 4; L < 7; L++) f.V[(long)L * tot + dst] = f.V[(long)L * tot + src];
            f.PP[dst] = f.PP[src];
            f.T[dst] = f.T[src];
            f.Rmiu[dst] = f.Rmiu[src];
            f.shock[dst] = f.shock[src];
            if (lay == 0) {
                const double s0 = f.Rds[src], s1 = f.Rds[tot + src], s2 = f.Rds[2 * tot + src];
                f.Rds[(long)(pr0 - 1) * tot + dst] = s0;
                f.Rds[(long)(pr1 - 1) * tot + dst] = s1;
                f.Rds[(long)(pr2 - 1) * tot + dst] = s2;
                if (f.Rdsf) {
                    f.Rdsf[(long)(pr0 - 1) * tot + dst] = (float)s0;
                    f.Rdsf[(long)(pr1 - 1) * tot + dst] = (float)s1;
                    f.Rdsf[(long)(pr2 - 1) * tot + dst] = (float)s2;
                }
            }
        }
        break;
    }
    default:
        break;
    }
}

// ---------------------------------------------------------------------------------------
// BC/BC_Outlet.F90:49-107: static-pressure outlet with the radial-equilibrium march along j
// from the hub.  One workgroup per (i,k) line of an outlet patch: the march inputs are staged
// in LDS by all lanes, lane 0 runs the pressure recurrence, and all lanes write the ghosts
// (the layer-0 Rds copy of the same routine is done by k_bc, after the spectral radius).
// It reads interior cells and geometry only.
constexpr int OUTLET_MAXJ = 512;

struct OutLine {
    int patch;  // index into the device patch table
    int i, k;
};

__global__ void __launch_bounds__(256) k_bc_outlet(trans_params p, int if_turb, Fields f,
                                                   const BlkInfo* __restrict__ blks,
                                                   const DPatch* __restrict__ pt,
                                                   const OutLine* __restrict__ lines)
{
    __shared__ double s_pp[OUTLET_MAXJ], s_t[OUTLET_MAXJ], s_v1[OUTLET_MAXJ], s_v3[OUTLET_MAXJ],
        s_v4[OUTLET_MAXJ], s_rad[OUTLET_MAXJ], s_sin[OUTLET_MAXJ], s_cos[OUTLET_MAXJ];
    const OutLine ln = lines[blockIdx.x];
    const DPatch& dp = pt[ln.patch];
    const trans_patch& c = dp.p;
    const BlkInfo b = blks[dp.blk];
    const long tot = f.tot;
    const int* ad = c.cijkadd;
    const int* e = c.ceijkadd;
    const int j0 = c.point_start[1];
    const int nj = c.point_end[1] - j0 + 1;
    const int i = ln.i, k = ln.k;
    for (int jj = threadIdx.x; jj < nj; jj += blockDim.x) {
        const int j = j0 + jj;
        const long is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        s_t[jj] = f.T[is];
        s_v1[jj] = f.V[is];
        s_v3[jj] = f.V[2 * tot + is];
        s_v4[jj] = f.V[3 * tot + is];
        s_rad[jj] = f.rad[il];
        s_sin[jj] = sin(f.thtc[il]);
        s_cos[jj] = cos(f.thtc[il]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double pp = p.pouthub;
        s_pp[0] = pp;
        for (int jj = 1; jj < nj; jj++) {
            // ghost state of the previous j (density, y/z momentum) as the reference has it
            const double den = pp / s_t[jj - 1] * p.rxm2;
            const double ratio = den / s_v1[jj - 1];
            const double gv3 = s_v3[jj - 1] * ratio, gv4 = s_v4[jj - 1] * ratio;
            const double rcs1 = s_rad[jj], rcs2 = s_rad[jj - 1];
            const double RoVr2 = sqr(gv4 / den * s_cos[jj - 1] - s_sin[jj - 1] * gv3 / den + p.omega1 * rcs2);
            pp = pp + fabs((rcs1 - rcs2) * den * RoVr2) / (rcs2 + p.tiny);
            s_pp[jj] = pp;
        }
    }
    __syncthreads();
    for (int jj = threadIdx.x; jj < nj; jj += blockDim.x) {
        const int j = j0 + jj;
        const long is = gidx(b, i + ad[0], j + ad[1], k + ad[2]);
        const long il = gidx(b, i + e[0], j + e[1], k + e[2]);
        const double tt = s_t[jj], pp = s_pp[jj];
        f.T[il] = tt;
        f.PP[il] = pp;
        const double v1 = pp / tt * p.rxm2;
        f.V[il] = v1;
        const double den_ratio = v1 / s_v1[jj];
        f.V[tot + il] = f.V[tot + is] * den_ratio;
        f.V[2 * tot + il] = s_v3[jj] * den_ratio;
        f.V[3 * tot + il] = s_v4[jj] * den_ratio;
        for (int lay = 1; lay < 3; lay++) {
            const int* ea = lay_dst(c, lay);
            const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
            for (int L = 0; L < 4; L++) f.V[(long)L * tot + d2] = f.V[(long)L * tot + il];
            f.PP[d2] = pp;
            f.T[d2] = tt;
        }
        if (if_turb) {
            f.V[5 * tot + il] = f.V[5 * tot + is];
            f.V[6 * tot + il] = f.V[6 * tot + is];
            f.Rmiu[il] = f.Rmiu[is];
            for (int lay = 1; lay < 3; lay++) {
                const int* ea = lay_dst(c, lay);
                const long d2 = gidx(b, i + ea[0], j + ea[1], k + ea[2]);
                f.V[5 * tot + d2] = f.V[5 * tot + il];
                f.V[6 * tot + d2] = f.V[6 * tot + il];
                f.Rmiu[d2] = f.Rmiu[il];
            }
        }
    }
}
