/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see trans_oracle.h).
 *
 * Plain-C FP64 restatement of the reference's hot-path routines.  Each function names
 * the reference file and lines it follows.  Loop order, operation order and the
 * reference's quirks are kept on purpose (they are what parity is measured against).
 *
 * Every function is checked against a closed form, a limit, an invariant or brute force in
 * tests/test_oracle_pins.py and tests/test_oracle_rotation_pins.py.  Those of the reference's
 * routines that compile unmodified from its own sources are also checked bit for bit against
 * that build (oracle/_ref, tests/test_oracle_ref.py): ora_precond_para and ora_precond_dw,
 * whose coefficients no closed form fixes, ora_alimiter, ora_muscl2, ora_ausmpw_common,
 * ora_fsst, ora_iddes_fd, ora_iddes_fe, and the whole-block ora_convect_muscl,
 * ora_convect_muscl2, ora_convect_mdcd, ora_convect_mdcdhyb, ora_turb_convect_muscl,
 * ora_convect_ausmpw, ora_turb_convect_upwind, ora_update_u and ora_rk_save.
 */
#include "trans_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#include "oracle_idx.h"

double ora_sqr(double x) { return x * x; }
#define sqr ora_sqr
#define sutherland ora_sutherland

/* main.F90:59-62 */
int ora_if_turb(const trans_params* p, int ki)
{
    return (p->if_euler == 0 && p->if_turb_org == 1 && (ki >= p->kitm || p->kind_hybrid > 0));
}

/* Sutherland law as used by the mean flow (Time/SpectralRadius.F90:59). */
double ora_sutherland(const trans_params* p, double tt)
{
    return (1.0 + p->csthlnd) / (tt * p->csth1 + p->csthlnd) * pow(p->csth1 * tt, 1.5);
}

/* ------------------------------------------------------------------------------------ */
/* Time/SpectralRadius.F90:1-180 (preconditioning off; vibn=vjbn=vkbn=0,
 * Input/Read_BlockDimensions.F90:420; IF_Local_Main=0, IF_Local_Sub=1, Input/ReadOrg.F90:155-156) */
void ora_spectral_radius(const trans_params* p, int ki, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double Den = V(1, I, J, K);
                double Vxi = V(2, I, J, K) / Den;
                double Vyi = V(3, I, J, K) / Den;
                double Vzi = V(4, I, J, K) / Den;
                double AAc = sqrt(1.4 * PP(I, J, K) / Den);
                double TT = T(I, J, K);
                double RmiuL = sutherland(p, TT);
                double RmiuTL = RMIU(I, J, K);
                double uvmu = (RmiuL + RmiuTL + 1.0e-20) * 4.0 / 3.0;
                double uvmupr = (RmiuL / 0.72 + RmiuTL / 0.9 + 1.0e-20) * 1.4;
                double pruvmu = fmax(uvmu, uvmupr) / Den;

                double ARX1 = 0.5 * (SD(1, 1, I, J, K) + SD(1, 1, I + 1, J, K));
                double ARY1 = 0.5 * (SD(1, 2, I, J, K) + SD(1, 2, I + 1, J, K));
                double ARZ1 = 0.5 * (SD(1, 3, I, J, K) + SD(1, 3, I + 1, J, K));
                double ST1sq = ARX1 * ARX1 + ARY1 * ARY1 + ARZ1 * ARZ1;
                double ARX2 = 0.5 * (SD(2, 1, I, J, K) + SD(2, 1, I, J + 1, K));
                double ARY2 = 0.5 * (SD(2, 2, I, J, K) + SD(2, 2, I, J + 1, K));
                double ARZ2 = 0.5 * (SD(2, 3, I, J, K) + SD(2, 3, I, J + 1, K));
                double ST2sq = ARX2 * ARX2 + ARY2 * ARY2 + ARZ2 * ARZ2;
                double ARX3 = 0.5 * (SD(3, 1, I, J, K) + SD(3, 1, I, J, K + 1));
                double ARY3 = 0.5 * (SD(3, 2, I, J, K) + SD(3, 2, I, J, K + 1));
                double ARZ3 = 0.5 * (SD(3, 3, I, J, K) + SD(3, 3, I, J, K + 1));
                double ST3sq = ARX3 * ARX3 + ARY3 * ARY3 + ARZ3 * ARZ3;
                /* vibc = vjbc = vkbc = 0.5*(0+0) */
                double UU = Vxi * ARX1 + Vyi * ARY1 + Vzi * ARZ1 + 0.0;
                double VV = Vxi * ARX2 + Vyi * ARY2 + Vzi * ARZ2 + 0.0;
                double WW = Vxi * ARX3 + Vyi * ARY3 + Vzi * ARZ3 + 0.0;
                double vol = VOL(I, J, K);
                double alf[3] = {0.0, 0.0, 0.0}, AAcd[3] = {AAc, AAc, AAc};
                if (p->if_precondition) { /* :91-120 */
                    double UVW2 = Vxi * Vxi + Vyi * Vyi + Vzi * Vzi;
                    double ub = ora_precond_para(p, UVW2, AAc * AAc, Den, 0.0, 0.0, 1.0);
                    double RmiuAll = RmiuL + RmiuTL;
                    for (int d = 0; d < 3; d++) {
                        double r = RmiuAll / Den / p->ref / ALAGM(d + 1, I, J, K) / AAc;
                        double ubd = fmax(ub, r * r);
                        alf[d] = (1.0 - ubd) / 2.0;
                        AAcd[d] = sqrt(alf[d] * alf[d] * UVW2 + ubd * AAc * AAc);
                    }
                }
                RDS(1, I, J, K) = fabs(UU) * (1.0 - alf[0]) + sqrt(ST1sq) * AAcd[0] +
                                  2.0 * ST1sq / vol * pruvmu / p->ref;
                RDS(2, I, J, K) = fabs(VV) * (1.0 - alf[1]) + sqrt(ST2sq) * AAcd[1] +
                                  2.0 * ST2sq / vol * pruvmu / p->ref;
                RDS(3, I, J, K) = fabs(WW) * (1.0 - alf[2]) + sqrt(ST3sq) * AAcd[2] +
                                  2.0 * ST3sq / vol * pruvmu / p->ref;
                double Dtm_local = vol / (RDS(1, I, J, K) + RDS(2, I, J, K) + RDS(3, I, J, K));
                /* IF_local_Main == 0 branch */
                DTM(I, J, K) = p->dt_lusgs_main;
                double cfl_sub = p->cfl_sub_start + (p->cfl_sub_end - p->cfl_sub_start) *
                                                        (double)(ki - p->ki_ini) /
                                                        (double)p->cfl_sub_step;
                cfl_sub = fmin(cfl_sub, p->cfl_sub_end);
                if (p->kind_dual != 0) DTM(I, J, K) = cfl_sub * Dtm_local; /* IF_local_Sub == 1 */
            }
}

/* ------------------------------------------------------------------------------------ */
/* BC/BC_FARBOND (BC/BC_far.F90:135-207) */
static void farbond(double XM, double P2, double RR, double Vxi, double Vyi, double Vzi,
                    double DNx, double DNy, double DNz, double VxF, double VyF, double VzF,
                    double* PP0, double* RR0, double* VX0, double* VY0, double* VZ0, double* dir)
{
    const double pGam = 1.4;
    double AF = 1. / XM;
    double AE = sqrt(pGam * P2 / RR);
    double QNE = Vxi * DNx + Vyi * DNy + Vzi * DNz;
    double QNF = VxF * DNx + VyF * DNy + VzF * DNz;
    double RF = QNF - 5. * AF;
    double RE = QNE + 5. * AE;
    double QN = 0.5 * (RE + RF);
    double AS = 0.1 * (RE - RF);
    if (QN > -AS && QN < 0.) {
        double SON = 1. / pGam / (XM * XM);
        double QTX = VxF - QNF * DNx, QTY = VyF - QNF * DNy, QTZ = VzF - QNF * DNz;
        *VX0 = QTX + QN * DNx;
        *VY0 = QTY + QN * DNy;
        *VZ0 = QTZ + QN * DNz;
        *RR0 = pow(AS * AS / SON / pGam, 2.5);
        *PP0 = AS * AS * *RR0 / pGam;
        *dir = -1.;
    }
    if (QN >= 0. && QN < AS) {
        double SON = P2 / pow(RR, pGam);
        double QTX = Vxi - QNE * DNx, QTY = Vyi - QNE * DNy, QTZ = Vzi - QNE * DNz;
        *VX0 = QTX + QN * DNx;
        *VY0 = QTY + QN * DNy;
        *VZ0 = QTZ + QN * DNz;
        *RR0 = pow(AS * AS / SON / pGam, 2.5);
        *PP0 = AS * AS * *RR0 / pGam;
        *dir = 1.;
    }
    if (QN <= -AS) {
        *VX0 = VxF;
        *VY0 = VyF;
        *VZ0 = VzF;
        *RR0 = 1.0;
        *PP0 = 1. / pGam / XM / XM;
        *dir = -1.;
    }
    if (QN >= AS) {
        *VX0 = Vxi;
        *VY0 = Vyi;
        *VZ0 = Vzi;
        *RR0 = RR;
        *PP0 = P2;
        *dir = 1.;
    }
}

/* Geom/startup.F90:1-25 */
void ora_ijk_start_end(int I, int J, int K, const int tl[3], const int tt[3], int lcross,
                       int out[3])
{
    int ft1 = tl[0], ft2 = tl[1], ft3 = tl[2];
    int gt1 = tt[0], gt2 = tt[1], gt3 = tt[2];
    int lcross1 = 1 - lcross;
    int temp1 = ft1 * I + ft2 * J + ft3 * K;
    int temp2 = ft1 * (J * lcross1 + K * lcross) + ft2 * (I * lcross1 + K * lcross) +
                ft3 * (I * lcross1 + J * lcross);
    int temp3 = ft1 * (J * lcross + K * lcross1) + ft2 * (I * lcross + K * lcross1) +
                ft3 * (I * lcross + J * lcross1);
    out[0] = gt1 * temp1 + (gt2 + gt3) * temp2;
    out[1] = gt2 * temp1 + gt1 * temp2 + gt3 * temp3;
    out[2] = gt3 * temp1 + (gt1 + gt2) * temp3;
}

/* BC/BC_Wall.F90:2-135 */
static void bc_wall(const trans_params* p, int if_turb, ora_block* b, const trans_patch* c,
                    int sweep)
{
    const int *ad = c->cijkadd, *ad1 = c->cijkadd1, *ad2 = c->cijkadd2;
    const int *e = c->ceijkadd, *e1 = c->ceijkadd1, *e2 = c->ceijkadd2;
    for (int k = c->point_start[2]; k <= c->point_end[2]; k++)
        for (int j = c->point_start[1]; j <= c->point_end[1]; j++)
            for (int i = c->point_start[0]; i <= c->point_end[0]; i++) {
                if (sweep == -1) {
                    int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                    int il = i + e[0], jl = j + e[1], kl = k + e[2];
                    double ztemp = 0.5 * (ZC(is, js, ks) + ZC(il, jl, kl));
                    double ytemp = 0.5 * (YC(is, js, ks) + YC(il, jl, kl));
                    double vibc = 0.0;
                    double vjbc = (double)c->subbctype * p->omega1 * ztemp;
                    double vkbc = -(double)c->subbctype * p->omega1 * ytemp;
                    for (int lay = 0; lay < 3; lay++) {
                        const int* sa = lay == 0 ? ad : (lay == 1 ? ad1 : ad2);
                        const int* ea = lay == 0 ? e : (lay == 1 ? e1 : e2);
                        is = i + sa[0]; js = j + sa[1]; ks = k + sa[2];
                        il = i + ea[0]; jl = j + ea[1]; kl = k + ea[2];
                        PP(il, jl, kl) = PP(is, js, ks);
                        T(il, jl, kl) = T(is, js, ks);
                        V(1, il, jl, kl) = PP(il, jl, kl) / (T(il, jl, kl) + p->tiny) * p->rxm2;
                        double den = V(1, il, jl, kl);
                        V(2, il, jl, kl) = vibc * den;
                        V(3, il, jl, kl) = vjbc * den;
                        V(4, il, jl, kl) = vkbc * den;
                        if (lay == 0)
                            for (int d = 1; d <= 3; d++) RDS(d, il, jl, kl) = RDS(d, is, js, ks);
                    }
                    if (if_turb) {
                        is = i + ad[0]; js = j + ad[1]; ks = k + ad[2];
                        il = i + e[0]; jl = j + e[1]; kl = k + e[2];
                        double Dn = DST(il, jl, kl);
                        double Den = V(1, il, jl, kl);
                        double TT = T(il, jl, kl);
                        double Rmiuw = sutherland(p, TT);
                        double Omegaw = 60.0 * Rmiuw / (Den * 0.075 * Dn * Dn * p->ref * p->ref);
                        Rmiuw = RMIU(is, js, ks);
                        for (int lay = 0; lay < 3; lay++) {
                            const int* ea = lay == 0 ? e : (lay == 1 ? e1 : e2);
                            il = i + ea[0]; jl = j + ea[1]; kl = k + ea[2];
                            V(6, il, jl, kl) = 0.0;
                            V(7, il, jl, kl) = Omegaw;
                            RMIU(il, jl, kl) = Rmiuw;
                        }
                    }
                } else if (sweep == 0) {
                    int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                    int il = i + e[0], jl = j + e[1], kl = k + e[2];
                    for (int L = 1; L <= 18; L++) DQ(L, il, jl, kl) = DQ(L, is, js, ks);
                }
            }
}

/* BC/BC_far.F90:1-131 */
static void bc_far(const trans_params* p, int if_turb, ora_block* b, const trans_patch* c,
                   int sweep)
{
    const int *ad = c->cijkadd, *e = c->ceijkadd, *e1 = c->ceijkadd1, *e2 = c->ceijkadd2;
    const int* sa = c->ijkadd;
    const int dir_ = c->ijkdirect;
    const double out_n = (double)(2 * c->left_or_right - 4);
    for (int k = c->point_start[2]; k <= c->point_end[2]; k++)
        for (int j = c->point_start[1]; j <= c->point_end[1]; j++)
            for (int i = c->point_start[0]; i <= c->point_end[0]; i++) {
                if (sweep == -1) {
                    int is = i + sa[0], js = j + sa[1], ks = k + sa[2];
                    double g = GRAD(dir_, is, js, ks);
                    double Sout1 = SD(dir_, 1, is, js, ks) / g;
                    double Sout2 = SD(dir_, 2, is, js, ks) / g;
                    double Sout3 = SD(dir_, 3, is, js, ks) / g;
                    double Dnx = Sout1 * out_n, Dny = Sout2 * out_n, Dnz = Sout3 * out_n;
                    is = i + ad[0]; js = j + ad[1]; ks = k + ad[2];
                    double Den = V(1, is, js, ks);
                    double Vxi = V(2, is, js, ks) / Den;
                    double Vyi = V(3, is, js, ks) / Den;
                    double Vzi = V(4, is, js, ks) / Den;
                    double P2 = PP(is, js, ks);
                    double PP0 = 0, Den0 = 0, Vx0 = 0, Vy0 = 0, Vz0 = 0, dir = 0;
                    farbond(p->xm, P2, Den, Vxi, Vyi, Vzi, Dnx, Dny, Dnz, p->vxf, p->vyf, p->vzf,
                            &PP0, &Den0, &Vx0, &Vy0, &Vz0, &dir);
                    int il = i + e[0], jl = j + e[1], kl = k + e[2];
                    V(1, il, jl, kl) = 2 * Den0 - V(1, is, js, ks);
                    double den = V(1, il, jl, kl);
                    V(2, il, jl, kl) = 2.0 * Vx0 - Vxi; /* velocity stored in the momentum slot */
                    V(3, il, jl, kl) = 2.0 * Vy0 - Vyi;
                    V(4, il, jl, kl) = 2.0 * Vz0 - Vzi;
                    PP(il, jl, kl) = 2.0 * PP0 - P2;
                    T(il, jl, kl) = p->rxm2 * PP(il, jl, kl) / den;
                    for (int lay = 1; lay < 3; lay++) {
                        const int* ea = lay == 1 ? e1 : e2;
                        int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                        for (int L = 1; L <= 5; L++) V(L, i2, j2, k2) = V(L, il, jl, kl);
                        PP(i2, j2, k2) = PP(il, jl, kl);
                        T(i2, j2, k2) = T(il, jl, kl);
                    }
                    if (if_turb) {
                        if (dir > 0.0) {
                            V(6, il, jl, kl) = V(6, is, js, ks);
                            V(7, il, jl, kl) = V(7, is, js, ks);
                            RMIU(il, jl, kl) = RMIU(is, js, ks);
                        } else {
                            double TT = T(il, jl, kl);
                            double Dn = V(1, il, jl, kl);
                            double RmiuFar = sutherland(p, TT);
                            RMIU(il, jl, kl) = 2.0 * p->visratio * RmiuFar - RMIU(is, js, ks);
                            V(6, il, jl, kl) = 2.0 * 1.5 * p->fsti * p->fsti * Dn - V(6, is, js, ks);
                            V(7, il, jl, kl) = V(6, il, jl, kl) / (RMIU(il, jl, kl) + p->tiny) * Dn;
                        }
                        for (int lay = 1; lay < 3; lay++) {
                            const int* ea = lay == 1 ? e1 : e2;
                            int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                            V(6, i2, j2, k2) = V(6, il, jl, kl);
                            V(7, i2, j2, k2) = V(7, il, jl, kl);
                            RMIU(i2, j2, k2) = RMIU(il, jl, kl);
                        }
                    }
                } else if (sweep == 0) {
                    int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                    int il = i + e[0], jl = j + e[1], kl = k + e[2];
                    for (int L = 1; L <= 18; L++) DQ(L, il, jl, kl) = DQ(L, is, js, ks);
                }
            }
}

/* BC/BC_Symm.F90:1-131.  Layer 1 does not receive the density (:68-75). */
static void bc_symm(const trans_params* p, int if_turb, ora_block* b, const trans_patch* c,
                    int sweep)
{
    (void)p;
    const int *ad = c->cijkadd, *ad1 = c->cijkadd1, *ad2 = c->cijkadd2;
    const int *e = c->ceijkadd, *e1 = c->ceijkadd1, *e2 = c->ceijkadd2;
    const int* sa = c->ijkadd;
    const int dir_ = c->ijkdirect;
    for (int k = c->point_start[2]; k <= c->point_end[2]; k++)
        for (int j = c->point_start[1]; j <= c->point_end[1]; j++)
            for (int i = c->point_start[0]; i <= c->point_end[0]; i++) {
                if (sweep == -1) {
                    int is = i + sa[0], js = j + sa[1], ks = k + sa[2];
                    double g = GRAD(dir_, is, js, ks);
                    double Sout1 = SD(dir_, 1, is, js, ks) / g;
                    double Sout2 = SD(dir_, 2, is, js, ks) / g;
                    double Sout3 = SD(dir_, 3, is, js, ks) / g;
                    is = i + ad[0]; js = j + ad[1]; ks = k + ad[2];
                    double den = V(1, is, js, ks);
                    double Vxi = V(2, is, js, ks) / den;
                    double Vyi = V(3, is, js, ks) / den;
                    double Vzi = V(4, is, js, ks) / den;
                    double Vc1 = Vxi * Sout1 + Vyi * Sout2 + Vzi * Sout3;
                    for (int lay = 0; lay < 3; lay++) {
                        const int* sl = lay == 0 ? ad : (lay == 1 ? ad1 : ad2);
                        const int* ea = lay == 0 ? e : (lay == 1 ? e1 : e2);
                        is = i + sl[0]; js = j + sl[1]; ks = k + sl[2];
                        int il = i + ea[0], jl = j + ea[1], kl = k + ea[2];
                        if (lay != 1) V(1, il, jl, kl) = V(1, is, js, ks);
                        V(2, il, jl, kl) = V(2, is, js, ks) - 2 * Sout1 * Vc1 * V(1, is, js, ks);
                        V(3, il, jl, kl) = V(3, is, js, ks) - 2 * Sout2 * Vc1 * V(1, is, js, ks);
                        V(4, il, jl, kl) = V(4, is, js, ks) - 2 * Sout3 * Vc1 * V(1, is, js, ks);
                        V(5, il, jl, kl) = V(5, is, js, ks);
                        PP(il, jl, kl) = PP(is, js, ks);
                        T(il, jl, kl) = T(is, js, ks);
                    }
                    if (if_turb) {
                        for (int lay = 0; lay < 3; lay++) {
                            const int* sl = lay == 0 ? ad : (lay == 1 ? ad1 : ad2);
                            const int* ea = lay == 0 ? e : (lay == 1 ? e1 : e2);
                            is = i + sl[0]; js = j + sl[1]; ks = k + sl[2];
                            int il = i + ea[0], jl = j + ea[1], kl = k + ea[2];
                            V(6, il, jl, kl) = V(6, is, js, ks);
                            V(7, il, jl, kl) = V(7, is, js, ks);
                            RMIU(il, jl, kl) = RMIU(is, js, ks);
                        }
                    }
                } else if (sweep == 0) {
                    int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                    int il = i + e[0], jl = j + e[1], kl = k + e[2];
                    for (int L = 1; L <= 18; L++) DQ(L, il, jl, kl) = DQ(L, is, js, ks);
                }
            }
}

/* BC/BC_Inlet.F90:1-127 */
static void bc_inlet(const trans_params* p, int if_turb, ora_block* b, const trans_patch* c,
                     int sweep)
{
    const int *ad = c->cijkadd, *e = c->ceijkadd, *e1 = c->ceijkadd1, *e2 = c->ceijkadd2;
    for (int k = c->point_start[2]; k <= c->point_end[2]; k++)
        for (int j = c->point_start[1]; j <= c->point_end[1]; j++)
            for (int i = c->point_start[0]; i <= c->point_end[0]; i++) {
                int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                int il = i + e[0], jl = j + e[1], kl = k + e[2];
                if (sweep == -1) {
                    double PsIn = PP(is, js, ks);
                    PsIn = fmin(PsIn, 0.999 * p->p0in);
                    double comm = fmax(1.0, pow(p->p0in / PsIn, 1.0 / 3.5));
                    double mach = sqrt((comm - 1.0) * 2 / 0.4);
                    double TsIn = p->t0in / comm;
                    double Roin = p->rxm2 * PsIn / TsIn;
                    double asonic = sqrt(1.4 * PsIn / Roin);
                    double Vabs = mach * asonic;
                    double VxInlet = Vabs;
                    double VyInlet = p->omega1 * (ZC(is, js, ks));
                    double VzInlet = -p->omega1 * YC(is, js, ks);
                    V(1, il, jl, kl) = Roin;
                    V(2, il, jl, kl) = Roin * VxInlet;
                    V(3, il, jl, kl) = Roin * VyInlet;
                    V(4, il, jl, kl) = Roin * VzInlet;
                    V(5, il, jl, kl) = PsIn / 0.4 +
                                       0.5 * (sqr(V(2, il, jl, kl)) + sqr(V(3, il, jl, kl)) +
                                              sqr(V(4, il, jl, kl))) / Roin -
                                       0.5 * sqr(p->omega1 * RAD(il, jl, kl)) * Roin;
                    PP(il, jl, kl) = PsIn;
                    T(il, jl, kl) = TsIn;
                    for (int d = 1; d <= 3; d++) RDS(d, il, jl, kl) = RDS(d, is, js, ks);
                    for (int lay = 1; lay < 3; lay++) {
                        const int* ea = lay == 1 ? e1 : e2;
                        int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                        for (int L = 1; L <= 5; L++) V(L, i2, j2, k2) = V(L, il, jl, kl);
                        PP(i2, j2, k2) = PsIn;
                        T(i2, j2, k2) = TsIn;
                    }
                    if (if_turb) {
                        double TT = T(il, jl, kl);
                        double RmiuIn = sutherland(p, TT);
                        RMIU(il, jl, kl) = p->visratio * RmiuIn;
                        double Den = V(1, il, jl, kl);
                        V(6, il, jl, kl) = 1.5 * p->fsti * p->fsti * Den;
                        V(7, il, jl, kl) = V(6, il, jl, kl) / (RMIU(il, jl, kl) + p->tiny) * Den;
                        for (int lay = 1; lay < 3; lay++) {
                            const int* ea = lay == 1 ? e1 : e2;
                            int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                            V(6, i2, j2, k2) = V(6, il, jl, kl);
                            V(7, i2, j2, k2) = V(7, il, jl, kl);
                            RMIU(i2, j2, k2) = RMIU(il, jl, kl);
                        }
                    }
                } else if (sweep == 0) {
                    for (int L = 1; L <= 18; L++) DQ(L, il, jl, kl) = DQ(L, is, js, ks);
                }
            }
}

/* BC/BC_Outlet.F90:1-111 (radial equilibrium marches in j from Jbgn) */
static void bc_outlet(const trans_params* p, int if_turb, ora_block* b, const trans_patch* c,
                      int sweep)
{
    const int *ad = c->cijkadd, *e = c->ceijkadd, *e1 = c->ceijkadd1, *e2 = c->ceijkadd2;
    const int Jbgn = c->point_start[1];
    for (int k = c->point_start[2]; k <= c->point_end[2]; k++)
        for (int j = c->point_start[1]; j <= c->point_end[1]; j++)
            for (int i = c->point_start[0]; i <= c->point_end[0]; i++) {
                int is = i + ad[0], js = j + ad[1], ks = k + ad[2];
                int il = i + e[0], jl = j + e[1], kl = k + e[2];
                if (sweep == 0) {
                    for (int L = 1; L <= 18; L++) DQ(L, il, jl, kl) = DQ(L, is, js, ks);
                } else if (sweep == -1) {
                    T(il, jl, kl) = T(is, js, ks);
                    if (j == Jbgn) {
                        PP(il, jl, kl) = p->pouthub;
                    } else {
                        double rcs1 = RAD(il, jl, kl);
                        double rcs2 = RAD(il, jl - 1, kl);
                        double sinr = sin(THTC(il, jl - 1, kl));
                        double cosr = cos(THTC(il, jl - 1, kl));
                        double den = V(1, il, jl - 1, kl);
                        double RoVr2 = sqr(V(4, il, jl - 1, kl) / den * cosr -
                                           sinr * V(3, il, jl - 1, kl) / den + p->omega1 * rcs2);
                        PP(il, jl, kl) = PP(il, jl - 1, kl) +
                                         fabs((rcs1 - rcs2) * den * RoVr2) / (rcs2 + p->tiny);
                    }
                    V(1, il, jl, kl) = PP(il, jl, kl) / T(il, jl, kl) * p->rxm2;
                    double den_ratio = V(1, il, jl, kl) / V(1, is, js, ks);
                    V(2, il, jl, kl) = V(2, is, js, ks) * den_ratio;
                    V(3, il, jl, kl) = V(3, is, js, ks) * den_ratio;
                    V(4, il, jl, kl) = V(4, is, js, ks) * den_ratio;
                    for (int d = 1; d <= 3; d++) RDS(d, il, jl, kl) = RDS(d, is, js, ks);
                    for (int lay = 1; lay < 3; lay++) {
                        const int* ea = lay == 1 ? e1 : e2;
                        int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                        for (int L = 1; L <= 4; L++) V(L, i2, j2, k2) = V(L, il, jl, kl);
                        PP(i2, j2, k2) = PP(il, jl, kl);
                        T(i2, j2, k2) = T(il, jl, kl);
                    }
                    if (if_turb) {
                        V(6, il, jl, kl) = V(6, is, js, ks);
                        V(7, il, jl, kl) = V(7, is, js, ks);
                        RMIU(il, jl, kl) = RMIU(is, js, ks);
                        for (int lay = 1; lay < 3; lay++) {
                            const int* ea = lay == 1 ? e1 : e2;
                            int i2 = i + ea[0], j2 = j + ea[1], k2 = k + ea[2];
                            V(6, i2, j2, k2) = V(6, il, jl, kl);
                            V(7, i2, j2, k2) = V(7, il, jl, kl);
                            RMIU(i2, j2, k2) = RMIU(il, jl, kl);
                        }
                    }
                }
            }
}

/* BC/BC_P2P.F90:1-186: block-interface / periodic halo copy through the orientation map. */
static void bc_p2p(ora_block* b, const trans_patch* c, ora_block* br, int sweep)
{
    const trans_patch* bc = &br->patches[c->target_conn];
    const int Ibgn = c->point_start[0], Jbgn = c->point_start[1], Kbgn = c->point_start[2];
    const int Iend = c->point_end[0], Jend = c->point_end[1], Kend = c->point_end[2];
    const int Tibgn = bc->point_start[0], Tjbgn = bc->point_start[1], Tkbgn = bc->point_start[2];
    const int IIadd = Iend - Ibgn + 2, JJadd = Jend - Jbgn + 2, KKadd = Kend - Kbgn + 2;
    const int* rm = c->rev_matrix;
    double dstage = c->rotate_angle;
    double sint = sin(dstage), cost = cos(dstage);
    double cos2t = cost * cost, sin2t = sint * sint, sct = sint * cost;
    const long sxr = (long)br->ni + 5, syr = (long)br->nj + 5;
#define IXR(i, j, k) ((long)((i) + 2) + sxr * ((long)((j) + 2) + syr * (long)((k) + 2)))
    for (int k = Kbgn; k <= Kend; k++)
        for (int j = Jbgn; j <= Jend; j++)
            for (int i = Ibgn; i <= Iend; i++) {
                int iL = i - Ibgn + 1, jL = j - Jbgn + 1, kL = k - Kbgn + 1;
                iL = rm[0] * IIadd + (1 - 2 * rm[0]) * iL;
                jL = rm[1] * JJadd + (1 - 2 * rm[1]) * jL;
                kL = rm[2] * KKadd + (1 - 2 * rm[2]) * kL;
                int buf[3];
                ora_ijk_start_end(iL, jL, kL, c->transform, c->transform_t, c->lcross, buf);
                int ibuf = buf[0] + Tibgn - 1, jbuf = buf[1] + Tjbgn - 1, kbuf = buf[2] + Tkbgn - 1;
                if (sweep == -1) {
                    for (int lay = 0; lay < 3; lay++) {
                        const int* sa = lay == 0 ? bc->cijkadd : (lay == 1 ? bc->cijkadd1 : bc->cijkadd2);
                        const int* ea = lay == 0 ? c->ceijkadd : (lay == 1 ? c->ceijkadd1 : c->ceijkadd2);
                        int is = ibuf + sa[0], js = jbuf + sa[1], ks = kbuf + sa[2];
                        int il = i + ea[0], jl = j + ea[1], kl = k + ea[2];
                        long src = IXR(is, js, ks), dst = IX(il, jl, kl);
                        b->V[0 + 7 * dst] = br->V[0 + 7 * src];
                        b->V[1 + 7 * dst] = br->V[1 + 7 * src];
                        double txy = br->V[2 + 7 * src], txz = br->V[3 + 7 * src];
                        b->V[2 + 7 * dst] = cost * txy + sint * txz;
                        b->V[3 + 7 * dst] = -sint * txy + cost * txz;
                        for (int L = 4; L < 7; L++) b->V[L + 7 * dst] = br->V[L + 7 * src];
                        b->PP[dst] = br->PP[src];
                        b->T[dst] = br->T[src];
                        b->Rmiu[dst] = br->Rmiu[src];
                        b->shock[dst] = br->shock[src];
                        if (lay == 0) {
                            double spec[3] = {br->Rds[0 + 3 * src], br->Rds[1 + 3 * src],
                                              br->Rds[2 + 3 * src]};
                            int one23[3] = {1, 2, 3}, perm[3];
                            ora_ijk_start_end(one23[0], one23[1], one23[2], c->transform,
                                              c->transform_t, c->lcross, perm);
                            b->Rds[perm[0] - 1 + 3 * dst] = spec[0];
                            b->Rds[perm[1] - 1 + 3 * dst] = spec[1];
                            b->Rds[perm[2] - 1 + 3 * dst] = spec[2];
                        }
                    }
                } else if (sweep == 0) {
                    int is = ibuf + bc->cijkadd[0], js = jbuf + bc->cijkadd[1], ks = kbuf + bc->cijkadd[2];
                    int il = i + c->ceijkadd[0], jl = j + c->ceijkadd[1], kl = k + c->ceijkadd[2];
                    const double* q = &br->DQdxyz[18 * IXR(is, js, ks)];
                    double* o = &b->DQdxyz[18 * IX(il, jl, kl)];
                    double ugx = q[0], ugy = q[1], ugz = q[2], vgx = q[3], vgy = q[4], vgz = q[5];
                    double wgx = q[6], wgy = q[7], wgz = q[8];
                    o[0] = ugx;
                    o[1] = cost * ugy + sint * ugz;
                    o[2] = -sint * ugy + cost * ugz;
                    o[3] = cost * vgx + sint * wgx;
                    o[4] = cos2t * vgy + sct * (vgz + wgy) + sin2t * wgz;
                    o[5] = cos2t * vgz + sct * (-vgy + wgz) - sin2t * wgy;
                    o[6] = -sint * vgx + cost * wgx;
                    o[7] = -sin2t * vgz + sct * (-vgy + wgz) + cos2t * wgy;
                    o[8] = sin2t * vgy + sct * (vgz + wgy) + cos2t * wgz;
                    for (int LL = 4; LL <= 6; LL++) {
                        int L = (LL - 1) * 3; /* 0-based start of the triple */
                        o[L] = q[L];
                        double gy = q[L + 1], gz = q[L + 2];
                        o[L + 1] = cost * gy + sint * gz;
                        o[L + 2] = -sint * gy + cost * gz;
                    }
                }
            }
#undef IXR
}

/* BC/Boundary.F90:1-48 */
void ora_bc(const trans_params* p, int if_turb, ora_block* blocks, int nblk, int sweep)
{
    for (int ib = 0; ib < nblk; ib++) {
        ora_block* b = &blocks[ib];
        for (int n = 0; n < b->npatch; n++) {
            const trans_patch* c = &b->patches[n];
            switch (c->bctype) {
            case 1: bc_wall(p, if_turb, b, c, sweep); break;
            case 2: bc_far(p, if_turb, b, c, sweep); break;
            case 3: bc_symm(p, if_turb, b, c, sweep); break;
            case 5: bc_inlet(p, if_turb, b, c, sweep); break;
            case 7: bc_outlet(p, if_turb, b, c, sweep); break;
            case 9: bc_p2p(b, c, &blocks[c->target_block], sweep); break;
            default: break;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Space/SPACEINTERP_GetSurfCenterVars.F90:1-232 */
void ora_spaceinterp(const trans_params* p, int if_turb, ora_block* b)
{
    (void)p;
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ilo = di ? 0 : 1, ihi = di ? NI + 1 : NI1;
        const int jlo = dj ? 0 : 1, jhi = dj ? NJ + 1 : NJ1;
        const int klo = dk ? 0 : 1, khi = dk ? NK + 1 : NK1;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    const int im = i - di, jm = j - dj, km = k - dk;
                    double Den1 = V(1, im, jm, km);
                    double UUU1 = V(2, im, jm, km) / Den1;
                    double VVV1 = V(3, im, jm, km) / Den1;
                    double WWW1 = V(4, im, jm, km) / Den1;
                    double PPP1 = PP(im, jm, km), TTT1 = T(im, jm, km);
                    double Den = V(1, i, j, k);
                    double UUU = V(2, i, j, k) / Den;
                    double VVV = V(3, i, j, k) / Den;
                    double WWW = V(4, i, j, k) / Den;
                    double PPP = PP(i, j, k), TTT = T(i, j, k);
                    double a0 = ALAGM(d, i, j, k), a1 = ALAGM(d, im, jm, km);
                    double alsinv = 1.0 / (a0 + a1);
                    VL(d, 1, i, j, k) = (a0 * UUU1 + a1 * UUU) * alsinv;
                    VL(d, 2, i, j, k) = (a0 * VVV1 + a1 * VVV) * alsinv;
                    VL(d, 3, i, j, k) = (a0 * WWW1 + a1 * WWW) * alsinv;
                    VL(d, 4, i, j, k) = (a0 * PPP1 + a1 * PPP) * alsinv;
                    VL(d, 5, i, j, k) = (a0 * TTT1 + a1 * TTT) * alsinv;
                    RCCL(d, i, j, k) = (a0 * RAD(im, jm, km) + a1 * RAD(i, j, k)) * alsinv;
                    RCCR(d, i, j, k) = RCCL(d, i, j, k);
                    for (int L = 1; L <= 5; L++) VR(d, L, i, j, k) = VL(d, L, i, j, k);
                }
    }
    if (!if_turb) return;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ilo = di ? 0 : 1, ihi = di ? NI + 1 : NI1;
        const int jlo = dj ? 0 : 1, jhi = dj ? NJ + 1 : NJ1;
        const int klo = dk ? 0 : 1, khi = dk ? NK + 1 : NK1;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    const int im = i - di, jm = j - dj, km = k - dk;
                    double Den1 = V(1, im, jm, km);
                    double Tkk1 = V(6, im, jm, km) / Den1;
                    double TOO1 = V(7, im, jm, km) / Den1;
                    double RmiuT1 = RMIU(im, jm, km);
                    double Den = V(1, i, j, k);
                    double Tkk = V(6, i, j, k) / Den;
                    double TOO = V(7, i, j, k) / Den;
                    double RmiuT = RMIU(i, j, k);
                    double a0 = ALAGM(d, i, j, k), a1 = ALAGM(d, im, jm, km);
                    double alsinv = 1.0 / (a0 + a1);
                    VL(d, 6, i, j, k) = (a0 * Tkk1 + a1 * Tkk) * alsinv;
                    VL(d, 7, i, j, k) = (a0 * TOO1 + a1 * TOO) * alsinv;
                    VL(d, 8, i, j, k) = (a0 * RmiuT1 + a1 * RmiuT) * alsinv;
                    for (int L = 6; L <= 8; L++) VR(d, L, i, j, k) = VL(d, L, i, j, k);
                }
    }
}

/* Space/UTIL_MUSCL.F90:2-18.  SIGN(1.0, x) follows the sign bit, as copysign does. */
double ora_alimiter(double var1, double var2)
{
    const double one = 1.0, epsm = 1.0e-7;
    return (var1 * (var2 * var2 + 2.0 * epsm * epsm) + var2 * (2.0 * var1 * var1 + epsm * epsm)) /
           (2.0 * var1 * var1 - var1 * var2 + 2.0 * var2 * var2 + 3.0 * epsm * epsm) * 0.5 *
           fabs(copysign(one, var1) + copysign(one, var2));
}

/* Space/UTIL_MUSCL.F90:21-35 (MUSCL2, the smoothness-weighted slope of CONVECT_MUSCL2). */
double ora_muscl2(double var1, double var2)
{
    const double epsm = 1.e-6;
    double smooth = (2. * var1 * var2 + epsm) / (var1 * var1 + var2 * var2 + epsm);
    return ((1. + smooth) * var1 + (1. - smooth) * var2) * smooth / 2.;
}

typedef double (*ora_limiter)(double, double);

/* Space/CONVECT_MUSCL.F90:40-323 (ALIMITER) and Space/CONVECT_MUSCL2.F90:30-290 (MUSCL2):
 * the same reconstruction from the face averages with a different limiter. */
static void convect_muscl_lim(ora_block* b, ora_limiter lim)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int i = 0; i <= NI + 1; i++)
        for (int j = 0; j <= NJ + 1; j++)
            for (int k = 0; k <= NK + 1; k++)
                for (int L = 1; L <= 5; L++) {
                    VR(1, L, i, j, k) = VL(1, L, i, j, k);
                    VR(2, L, i, j, k) = VL(2, L, i, j, k);
                    VR(3, L, i, j, k) = VL(3, L, i, j, k);
                }
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ilo = di ? 0 : 1, ihi = di ? NI : NI1;
        const int jlo = dj ? 0 : 1, jhi = dj ? NJ : NJ1;
        const int klo = dk ? 0 : 1, khi = dk ? NK : NK1;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    const int ip = i + di, jp = j + dj, kp = k + dk;
                    double q[5];
                    q[0] = V(2, i, j, k) / V(1, i, j, k);
                    q[1] = V(3, i, j, k) / V(1, i, j, k);
                    q[2] = V(4, i, j, k) / V(1, i, j, k);
                    q[3] = PP(i, j, k);
                    q[4] = T(i, j, k);
                    for (int L = 1; L <= 5; L++) {
                        double DR = VL(d, L, ip, jp, kp) - q[L - 1];
                        double DL = q[L - 1] - VR(d, L, i, j, k);
                        VL(d, L, ip, jp, kp) = q[L - 1] + lim(DR, DL);
                        VR(d, L, i, j, k) = q[L - 1] - lim(DL, DR);
                    }
                    double rrr = RAD(i, j, k);
                    double DR = RCCL(d, ip, jp, kp) - rrr;
                    double DL = rrr - RCCR(d, i, j, k);
                    RCCL(d, ip, jp, kp) = rrr + lim(DR, DL);
                    RCCR(d, i, j, k) = rrr - lim(DL, DR);
                }
    }
}

void ora_convect_muscl(const trans_params* p, ora_block* b)
{
    (void)p;
    convect_muscl_lim(b, ora_alimiter);
}

void ora_convect_muscl2(const trans_params* p, ora_block* b)
{
    (void)p;
    convect_muscl_lim(b, ora_muscl2);
}

/* Common/UTIL_Precond_Funcs.F90:6-30 (getPrecondPara): (U_r/c)^2 of the low-Mach
 * preconditioning, K = 3, Kc = 1, with the viscous limit. */
double ora_precond_para(const trans_params* p, double VM2, double AM2, double RR, double TT, double RMIUT,
                        double ALAGMR)
{
    const double Kp = 3.0, Kc = 1.0;
    double ubetac2 = fmin(fmax(VM2 / AM2, Kp * p->xm * p->xm), 1.0);
    double RmiuL = (1.0 + p->csthlnd) / (TT * p->csth1 + p->csthlnd) * pow(p->csth1 * TT, 1.5);
    double RmiuAll = RmiuL + RMIUT;
    double r = RmiuAll / RR / p->ref / ALAGMR;
    ubetac2 = fmax(ubetac2, Kc * (r * r) / AM2);
    return ubetac2;
}

/* Common/UTIL_Precond_Funcs.F90:86-130 (PrecondDW): the conservative increment a[5] through
 * the preconditioning transform (momentum components unchanged). */
void ora_precond_dw(const trans_params* p, double a[5], double pRo, double pVx, double pVy, double pVz, double pE,
                    double AAc, double ubetac2, double pPP, double pTT)
{
    (void)pPP;
    double pU2 = pVx * pVx + pVy * pVy + pVz * pVz;
    double pUAU = pVx * a[1] + pVy * a[2] + pVz * a[3];
    a[4] = a[4] - (pUAU) - (pE / 1.4 - 1.2 * (pU2)) * a[0];
    double rCp = 1.4 / 0.4 * p->ppf * pRo;
    double a11 = AAc * AAc * ubetac2;
    double a15 = -0.4 * ubetac2 / rCp;
    double b11 = 1.0 * a11 / rCp;
    double b15 = (1.0 + 0.4 * ubetac2) / rCp;
    double AinP = a11 * a[0] + a15 * a[4];
    double AinT = b11 * a[0] + b15 * a[4];
    a[0] = AinP / p->ppf / pTT - pRo * AinT / pTT;
    a[4] = AinP / 0.4 + (pUAU);
}

/* The preconditioning transform of a cell's increment with the cell's own state
 * (Time/TIME_LUSGS.F90:113-160 and the neighbour terms :235-246 and mirrors). */
static void precond_cell(const trans_params* p, ora_block* b, int i, int j, int k, double a[5])
{
    double pRo = V(1, i, j, k), roinv = 1.0 / pRo;
    double pVx = V(2, i, j, k) * roinv, pVy = V(3, i, j, k) * roinv, pVz = V(4, i, j, k) * roinv;
    double pE = V(5, i, j, k) * roinv;
    double pPP = PP(i, j, k), pTT = T(i, j, k);
    double AAc = sqrt(1.4 * pPP * roinv);
    double UVW2 = pVx * pVx + pVy * pVy + pVz * pVz;
    double ALL = fmin(fmin(ALAGM(1, i, j, k), ALAGM(2, i, j, k)), ALAGM(3, i, j, k));
    double ubetac2 = ora_precond_para(p, UVW2, AAc * AAc, pRo, pTT, RMIU(i, j, k), ALL);
    ora_precond_dw(p, a, pRo, pVx, pVy, pVz, pE, AAc, ubetac2, pPP, pTT);
}

/* Space/CONVECT_Roe_Original.F90:242-407 (CONVECT_Roe_Common) without the face eddy
 * viscosity and Alagm, which only its preconditioned branch reads. */
void ora_roe_common(const trans_params* p, double PLo, double PRo, double TLo, double TRo,
                    double ULo, double URo, double VLo, double VRo, double WLo, double WRo,
                    double sav1, double sav2, double sav3, double sav, double rLo, double rRo,
                    double out[5])
{
    ora_roe_common_pc(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, sav1, sav2, sav3, sav, rLo, rRo, 0.0,
                      1.0, out);
}

/* Space/CONVECT_Roe_Original.F90:242-407 (CONVECT_Roe_Common) with the face eddy viscosity
 * RmiuT = VL(d,8) and AlagmR = Alagm(d) that the preconditioned branch (:319-332,
 * IF_PRECONDITION = 1, no entropy fix) reads. */
void ora_roe_common_pc(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                       double URo, double VLo, double VRo, double WLo, double WRo, double sav1, double sav2,
                       double sav3, double sav, double rLo, double rRo, double RmiuT, double AlagmR, double out[5])
{
    const double rcpcv = p->ppf;
    double DENLo = PLo / (rcpcv * TLo);
    double DENRo = PRo / (rcpcv * TRo);
    const double ga = 1.4, ga1 = 1.4 - 1.0;
    double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo;
    double VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    double acl = sqrt(ga * PLo / DENLo);
    double acr = sqrt(ga * PRo / DENRo);
    double rrorl = sqrt(DENRo / DENLo);
    double rrorlp1 = 1.0 + rrorl;
    double rm = sqrt(DENRo * DENLo);
    double um = (ULo + URo * rrorl) / rrorlp1;
    double vm = (VLo + VRo * rrorl) / rrorlp1;
    double wm = (WLo + WRo * rrorl) / rrorlp1;
    double hm = (HLo + HRo * rrorl) / rrorlp1;
    double vm2 = um * um + vm * vm + wm * wm;
    double rccm = (rLo + rrorl * rRo) / rrorlp1;
    double am2 = ga1 * fabs(hm - 0.5 * vm2 + 0.5 * (rccm * rccm));
    double am = sqrt(am2);
    double sav1n = sav1 / sav, sav2n = sav2 / sav, sav3n = sav3 / sav;
    double unormaln = (sav1n * um + sav2n * vm + sav3n * wm);
    double anormaln = am;
    (void)acl; (void)acr;
    double eig1, eig2, eig3, alf, ubetac2, amrev;
    if (p->if_precondition) { /* :319-332 */
        double TTL = am2 / ga / p->ppf;
        ubetac2 = ora_precond_para(p, vm2, am2, rm, TTL, RmiuT, AlagmR);
        alf = (1.0 - ubetac2) / 2.0;
        double unrev = unormaln * (1.0 - alf);
        amrev = sqrt((alf * unormaln) * (alf * unormaln) + ubetac2 * am2);
        eig1 = fabs(unormaln);
        eig2 = fabs(unrev + amrev);
        eig3 = fabs(unrev - amrev);
    } else {
        eig1 = fabs(unormaln);
        eig2 = fabs(unormaln + anormaln);
        eig3 = fabs(unormaln - anormaln);
        if (eig1 < 0.1) eig1 = (eig1 * eig1 + 0.1 * 0.1) / 0.2;
        if (eig2 < 0.1) eig2 = (eig2 * eig2 + 0.1 * 0.1) / 0.2;
        if (eig3 < 0.1) eig3 = (eig3 * eig3 + 0.1 * 0.1) / 0.2;
        alf = 0.0;
        ubetac2 = 1.0;
        amrev = am;
    }
    double dunormaln = (sav1n * (URo - ULo) + sav2n * (VRo - VLo) + sav3n * (WRo - WLo));
    double astar = (eig2 + eig3) / 2.0;
    double Mstar = (eig2 - eig3) / 2.0 / amrev;
    double dW1 = DENRo - DENLo;
    double dU1 = URo - ULo, dU2 = VRo - VLo, dU3 = WRo - WLo;
    double dP1 = PRo - PLo;
    double dW2 = rm * dU1 + dW1 * um;
    double dW3 = rm * dU2 + dW1 * vm;
    double dW4 = rm * dU3 + dW1 * wm;
    double dW5 = dP1 / ga1 + 0.5 * vm2 * dW1 + rm * um * dU1 + rm * vm * dU2 + rm * wm * dU3;
    double dUroe = Mstar * dunormaln +
                   (astar - (1 - 2.0 * alf) * eig1 - alf * unormaln * Mstar) * (PRo - PLo) / rm /
                       (ubetac2 * am2);
    double dProe = Mstar * (PRo - PLo) + (astar - eig1 + alf * unormaln * Mstar) * rm * dunormaln;
    double droR1 = eig1 * dW1 + dUroe * rm;
    double drxR1 = eig1 * dW2 + dUroe * rm * um + dProe * sav1n;
    double dryR1 = eig1 * dW3 + dUroe * rm * vm + dProe * sav2n;
    double drzR1 = eig1 * dW4 + dUroe * rm * wm + dProe * sav3n;
    double dreR1 = eig1 * dW5 + dUroe * rm * hm + dProe * unormaln;
    double ulnormaln = (sav1n * ULo + sav2n * VLo + sav3n * WLo);
    double urnormaln = (sav1n * URo + sav2n * VRo + sav3n * WRo);
    double rulnormaln = DENLo * ulnormaln;
    double rurnormaln = DENRo * urnormaln;
    out[0] = -0.5 * (rulnormaln + rurnormaln - droR1);
    out[1] = -0.5 * (rulnormaln * ULo + rurnormaln * URo + sav1n * (PLo + PRo) - drxR1);
    out[2] = -0.5 * (rulnormaln * VLo + rurnormaln * VRo + sav2n * (PLo + PRo) - dryR1);
    out[3] = -0.5 * (rulnormaln * WLo + rurnormaln * WRo + sav3n * (PLo + PRo) - drzR1);
    out[4] = -0.5 * (rulnormaln * HLo + rurnormaln * HRo - dreR1);
}

/* Space/CONVECT_Roe_Original.F90:1-236 */
void ora_convect_roe(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    double f[5];
                    double sav = GRAD(d, i, j, k);
                    ora_roe_common_pc(p, VL(d, 4, i, j, k), VR(d, 4, i, j, k), VL(d, 5, i, j, k),
                                      VR(d, 5, i, j, k), VL(d, 1, i, j, k), VR(d, 1, i, j, k),
                                      VL(d, 2, i, j, k), VR(d, 2, i, j, k), VL(d, 3, i, j, k),
                                      VR(d, 3, i, j, k), SD(d, 1, i, j, k), SD(d, 2, i, j, k),
                                      SD(d, 3, i, j, k), sav, RCCL(d, i, j, k) * p->omega1,
                                      RCCR(d, i, j, k) * p->omega1, VL(d, 8, i, j, k), ALAGM(d, i, j, k), f);
                    for (int L = 1; L <= 5; L++) {
                        D(L, i, j, k) = D(L, i, j, k) - f[L - 1] * sav;
                        D(L, i - di, j - dj, k - dk) = D(L, i - di, j - dj, k - dk) + f[L - 1] * sav;
                    }
                }
    }
}

/* Space/CONVECT_Roe_Rotated.F90:597-752 (its own Roe_Common; the preconditioned eigenvalues
 * of :669-680 read RmiuT and AlagmR): adds the Roe dissipation through the sub-face (sav1,
 * sav2, sav3), |.| = sav, weighted by Funscheme, to the area-weighted flux acc[5]. */
static void roe_rot_common(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                           double URo, double VLo, double VRo, double WLo, double WRo, double sav1, double sav2,
                           double sav3, double sav, double Funscheme, double rLo, double rRo, double RmiuT,
                           double AlagmR, double acc[5])
{
    double rcpcv = p->ppf;
    double DENLo = PLo / (rcpcv * TLo), DENRo = PRo / (rcpcv * TRo);
    double ga = 1.4, ga1 = 0.4;
    double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo, VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    double rrorl = sqrt(DENRo / DENLo), rrorlp1 = 1.0 + rrorl;
    double rm = sqrt(DENRo * DENLo);
    double um = (ULo + URo * rrorl) / rrorlp1, vm = (VLo + VRo * rrorl) / rrorlp1;
    double wm = (WLo + WRo * rrorl) / rrorlp1, hm = (HLo + HRo * rrorl) / rrorlp1;
    double vm2 = um * um + vm * vm + wm * wm;
    double rccm = (rLo + rRo * rrorl) / rrorlp1;
    double am2 = ga1 * fabs(hm - 0.5 * vm2 + 0.5 * (rccm * rccm));
    double am = sqrt(am2);
    double sav1n = sav1 / sav, sav2n = sav2 / sav, sav3n = sav3 / sav;
    double unormaln = (sav1n * um + sav2n * vm + sav3n * wm);
    double anormaln = am;
    double eig1, eig2, eig3, alf, ubetac2, amrev;
    if (p->if_precondition) {
        double TTL = am2 / ga / p->ppf;
        ubetac2 = ora_precond_para(p, vm2, am2, rm, TTL, RmiuT, AlagmR);
        alf = (1.0 - ubetac2) / 2.0;
        double unrev = unormaln * (1.0 - alf);
        amrev = sqrt((alf * unormaln) * (alf * unormaln) + ubetac2 * am2);
        eig1 = fabs(unormaln);
        eig2 = fabs(unrev + amrev);
        eig3 = fabs(unrev - amrev);
    } else {
        eig1 = fabs(unormaln);
        eig2 = fabs(unormaln + anormaln);
        eig3 = fabs(unormaln - anormaln);
        alf = 0.0;
        ubetac2 = 1.0;
        amrev = am;
    }
    double dunormaln = (sav1n * (URo - ULo) + sav2n * (VRo - VLo) + sav3n * (WRo - WLo));
    double astar = (eig2 + eig3) / 2.0;
    double Mstar = (eig2 - eig3) / 2.0 / amrev;
    double dW1 = DENRo - DENLo, dU1 = URo - ULo, dU2 = VRo - VLo, dU3 = WRo - WLo, dP1 = PRo - PLo;
    double dW2 = rm * dU1 + dW1 * um, dW3 = rm * dU2 + dW1 * vm, dW4 = rm * dU3 + dW1 * wm;
    double dW5 = dP1 / ga1 + 0.5 * vm2 * dW1 + rm * um * dU1 + rm * vm * dU2 + rm * wm * dU3;
    double dUroe = Mstar * dunormaln + (astar - (1 - 2.0 * alf) * eig1 - alf * unormaln * Mstar) * (PRo - PLo) / rm /
                                           (ubetac2 * am2);
    double dProe = Mstar * (PRo - PLo) + (astar - eig1 + alf * unormaln * Mstar) * rm * dunormaln;
    double r1[5];
    r1[0] = eig1 * dW1 + dUroe * rm;
    r1[1] = eig1 * dW2 + dUroe * rm * um + dProe * sav1n;
    r1[2] = eig1 * dW3 + dUroe * rm * vm + dProe * sav2n;
    r1[3] = eig1 * dW4 + dUroe * rm * wm + dProe * sav3n;
    r1[4] = eig1 * dW5 + dUroe * rm * hm + dProe * unormaln;
    for (int L = 0; L < 5; L++) acc[L] = acc[L] + 0.5 * r1[L] * sav * Funscheme;
}

/* Space/CONVECT_Roe_Rotated.F90:84-176 for one face: central flux through the face, then
 * the Roe dissipation through the face (velocity jump below Tiny), or through its parts
 * along and across the velocity jump.  The grid-velocity term vibc of :169 is never set in
 * the routine; it is taken as 0.  out = area-weighted flux droR..dreR. */
void ora_roe_rotated_face(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                          double URo, double VLo, double VRo, double WLo, double WRo, double s1, double s2,
                          double s3, double grad, double Funscheme, double rLo, double rRo, double out[5])
{
    ora_roe_rotated_face_pc(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, s1, s2, s3, grad, Funscheme, rLo,
                            rRo, 0.0, 1.0, out);
}

/* As ora_roe_rotated_face, with the face eddy viscosity RmiuT = VL(d,8) and AlagmR = Alagm(d)
 * of the preconditioned branch (Space/CONVECT_Roe_Rotated.F90:112-113). */
void ora_roe_rotated_face_pc(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                             double URo, double VLo, double VRo, double WLo, double WRo, double s1, double s2,
                             double s3, double grad, double Funscheme, double rLo, double rRo, double RmiuT,
                             double AlagmR, double out[5])
{
    double rcpcv = p->ppf;
    double DENLo = PLo / (rcpcv * TLo), DENRo = PRo / (rcpcv * TRo);
    double ga = 1.4, ga1 = 0.4;
    double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo, VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo - 0.5 * (rLo * rLo);
    double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo - 0.5 * (rRo * rRo);
    double sav = grad;
    double sav1n = s1 / sav, sav2n = s2 / sav, sav3n = s3 / sav;
    double ulnormaln = (sav1n * ULo + sav2n * VLo + sav3n * WLo);
    double urnormaln = (sav1n * URo + sav2n * VRo + sav3n * WRo);
    double rulnormaln = DENLo * ulnormaln, rurnormaln = DENRo * urnormaln;
    const double vibc = 0.0;
    out[0] = -0.5 * (rulnormaln + rurnormaln) * sav;
    out[1] = -0.5 * (rulnormaln * ULo + rurnormaln * URo + sav1n * (PLo + PRo)) * sav;
    out[2] = -0.5 * (rulnormaln * VLo + rurnormaln * VRo + sav2n * (PLo + PRo)) * sav;
    out[3] = -0.5 * (rulnormaln * WLo + rurnormaln * WRo + sav3n * (PLo + PRo)) * sav;
    out[4] = -0.5 * (rulnormaln * HLo + rurnormaln * HRo - vibc * (PLo + PRo)) * sav;
    double DeltaU = sqrt((URo - ULo) * (URo - ULo) + (VRo - VLo) * (VRo - VLo) + (WRo - WLo) * (WRo - WLo));
    if (DeltaU <= p->tiny) {
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
        double a1 = sav1n * sav, a2 = sav2n * sav, a3 = sav3n * sav;
        double b1 = s1 - a1, b2 = s2 - a2, b3 = s3 - a3;
        sav = fmax(sqrt(a1 * a1 + a2 * a2 + a3 * a3), p->tiny);
        roe_rot_common(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, a1, a2, a3, sav, Funscheme, rLo, rRo,
                       RmiuT, AlagmR, out);
        sav = fmax(sqrt(b1 * b1 + b2 * b2 + b3 * b3), p->tiny);
        roe_rot_common(p, PLo, PRo, TLo, TRo, ULo, URo, VLo, VRo, WLo, WRo, b1, b2, b3, sav, Funscheme, rLo, rRo,
                       RmiuT, AlagmR, out);
    }
}

/* Space/CONVECT_Roe_Rotated.F90:1-595: the scatter of CONVECT_Roe_Original with the rotated
 * face flux and Funscheme = (Fscheme of the two cells)/2. */
void ora_convect_roe_rotated(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    double f[5];
                    double Funscheme = (FSCHEME(i - di, j - dj, k - dk) + FSCHEME(i, j, k)) / 2.;
                    ora_roe_rotated_face_pc(p, VL(d, 4, i, j, k), VR(d, 4, i, j, k), VL(d, 5, i, j, k),
                                            VR(d, 5, i, j, k), VL(d, 1, i, j, k), VR(d, 1, i, j, k),
                                            VL(d, 2, i, j, k), VR(d, 2, i, j, k), VL(d, 3, i, j, k),
                                            VR(d, 3, i, j, k), SD(d, 1, i, j, k), SD(d, 2, i, j, k),
                                            SD(d, 3, i, j, k), GRAD(d, i, j, k), Funscheme,
                                            RCCL(d, i, j, k) * p->omega1, RCCR(d, i, j, k) * p->omega1,
                                            VL(d, 8, i, j, k), ALAGM(d, i, j, k), f);
                    for (int L = 1; L <= 5; L++) {
                        D(L, i, j, k) = D(L, i, j, k) - f[L - 1];
                        D(L, i - di, j - dj, k - dk) = D(L, i - di, j - dj, k - dk) + f[L - 1];
                    }
                }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Time/TIME_LUSGS.F90:696-768 (LUSGS_ADQ, gamma = 1.4, grid velocity 0) */
void ora_lusgs_adq(double pVx, double pVy, double pVz, double aSDx, double aSDy, double aSDz,
                   double pPP, double roinv, const double ain[5], double dif[5])
{
    const double pGam = 1.4;
    double AinRo = ain[0], AinRx = ain[1], AinRy = ain[2], AinRz = ain[3], AinRe = ain[4];
    double VP = pVx * aSDx + pVy * aSDy + pVz * aSDz + 0.0;
    double Ut2 = pVx * pVx + pVy * pVy + pVz * pVz;
    const double aSDt = 0.0;
    double Gm1 = pGam - 1.0, Gm2 = pGam - 2.0;
    double AA2 = pGam * pPP * roinv;
    double gUt2 = 0.5 * Gm1 * Ut2;
    dif[0] = AinRo * aSDt + AinRx * aSDx + AinRy * aSDy + AinRz * aSDz + AinRe * 0.0;
    dif[1] = AinRo * (aSDx * gUt2 - pVx * VP) + AinRx * (aSDt + VP - aSDx * Gm2 * pVx) +
             AinRy * (aSDy * pVx - aSDx * Gm1 * pVy) + AinRz * (aSDz * pVx - aSDx * Gm1 * pVz) +
             AinRe * (aSDx * Gm1);
    dif[2] = AinRo * (aSDy * gUt2 - pVy * VP) + AinRx * (aSDx * pVy - aSDy * Gm1 * pVx) +
             AinRy * (aSDt + VP - aSDy * Gm2 * pVy) + AinRz * (aSDz * pVy - aSDy * Gm1 * pVz) +
             AinRe * (aSDy * Gm1);
    dif[3] = AinRo * (aSDz * gUt2 - pVz * VP) + AinRx * (aSDx * pVz - aSDz * Gm1 * pVx) +
             AinRy * (aSDy * pVz - aSDz * Gm1 * pVy) + AinRz * (aSDt + VP - aSDz * Gm2 * pVz) +
             AinRe * (aSDz * Gm1);
    double GEt1 = 0.5 * Gm2 * Ut2 - AA2 / Gm1;
    double GEt2 = 0.5 * Ut2 + AA2 / Gm1;
    dif[4] = AinRo * VP * GEt1 + AinRx * (aSDx * GEt2 - Gm1 * pVx * VP) +
             AinRy * (aSDy * GEt2 - Gm1 * pVy * VP) + AinRz * (aSDz * GEt2 - Gm1 * pVz * VP) +
             AinRe * (aSDt + pGam * VP);
}

/* Neighbour term of the sweeps (Time/TIME_LUSGS.F90:190-256 and mirrors):
 * 0.5*(A_nb dQ_nb + sgn*Rds_nb dQ_nb), A built from the neighbour's state and the mean of
 * its two faces in direction d. */
static void lusgs_nb(const trans_params* p, ora_block* b, int d, int in, int jn, int kn, int sgn, double out[5])
{
    const int di = d == 1, dj = d == 2, dk = d == 3;
    double ain[5];
    for (int L = 0; L < 5; L++) ain[L] = D(L + 1, in, jn, kn);
    double pRo = V(1, in, jn, kn);
    double roinv = 1.0 / pRo;
    double pVx = V(2, in, jn, kn) * roinv;
    double pVy = V(3, in, jn, kn) * roinv;
    double pVz = V(4, in, jn, kn) * roinv;
    /* faces of the neighbour cell in direction d: (in,jn,kn) and (in+di,...) */
    double aSDx = 0.5 * (SD(d, 1, in + di, jn + dj, kn + dk) + SD(d, 1, in, jn, kn));
    double aSDy = 0.5 * (SD(d, 2, in + di, jn + dj, kn + dk) + SD(d, 2, in, jn, kn));
    double aSDz = 0.5 * (SD(d, 3, in + di, jn + dj, kn + dk) + SD(d, 3, in, jn, kn));
    double dif[5];
    ora_lusgs_adq(pVx, pVy, pVz, aSDx, aSDy, aSDz, PP(in, jn, kn), roinv, ain, dif);
    if (p->if_precondition) precond_cell(p, b, in, jn, kn, dif); /* :235-246 and mirrors */
    double RDSt = RDS(d, in, jn, kn);
    for (int L = 0; L < 5; L++) out[L] = 0.5 * (dif[L] + (double)sgn * RDSt * ain[L]);
}

/* Time/TIME_LUSGS.F90:2-688 (with the preconditioning transforms when IF_PRECONDITION = 1) */
void ora_lusgs(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    const double dtm = p->dt_lusgs_main;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++)
                for (int L = 1; L <= 5; L++)
                    D(L, I, J, K) = D(L, I, J, K) / VOL(I, J, K) -
                                    (3.0 * V(L, I, J, K) - 4.0 * WT0(L, I, J, K) + WT1(L, I, J, K)) *
                                        0.5 / dtm;
    if (p->if_precondition) /* :113-160 */
        for (int K = 1; K <= NK1; K++)
            for (int J = 1; J <= NJ1; J++)
                for (int I = 1; I <= NI1; I++) {
                    double a[5];
                    for (int L = 0; L < 5; L++) a[L] = D(L + 1, I, J, K);
                    precond_cell(p, b, I, J, K, a);
                    for (int L = 0; L < 5; L++) D(L + 1, I, J, K) = a[L];
                }
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                F(1, I, J, K) = 1.0 + (RDS(1, I, J, K) + RDS(2, I, J, K) + RDS(3, I, J, K)) *
                                          (stept / VOL(I, J, K));
            }
    /* L sweep */
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                double cvs = stept / VOL(I, J, K);
                double dI[5] = {0, 0, 0, 0, 0}, dJ[5] = {0, 0, 0, 0, 0}, dK[5] = {0, 0, 0, 0, 0};
                if (I != 1) lusgs_nb(p, b, 1, I - 1, J, K, +1, dI);
                if (J != 1) lusgs_nb(p, b, 2, I, J - 1, K, +1, dJ);
                if (K != 1) lusgs_nb(p, b, 3, I, J, K - 1, +1, dK);
                double f1 = F(1, I, J, K);
                for (int L = 1; L <= 5; L++)
                    D(L, I, J, K) = (D(L, I, J, K) * stept + (dI[L - 1] + dJ[L - 1] + dK[L - 1]) * cvs) / f1;
            }
    /* U sweep */
    for (int K = NK1; K >= 1; K--)
        for (int J = NJ1; J >= 1; J--)
            for (int I = NI1; I >= 1; I--) {
                double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                double cvs = stept / VOL(I, J, K);
                double dI[5] = {0, 0, 0, 0, 0}, dJ[5] = {0, 0, 0, 0, 0}, dK[5] = {0, 0, 0, 0, 0};
                if (I != NI1) lusgs_nb(p, b, 1, I + 1, J, K, -1, dI);
                if (J != NJ1) lusgs_nb(p, b, 2, I, J + 1, K, -1, dJ);
                if (K != NK1) lusgs_nb(p, b, 3, I, J, K + 1, -1, dK);
                double f1 = F(1, I, J, K);
                for (int L = 1; L <= 5; L++)
                    D(L, I, J, K) = D(L, I, J, K) - (dI[L - 1] + dJ[L - 1] + dK[L - 1]) * cvs / f1;
            }
}

/* WD+ limiter of the k-omega update (Time/TIME_Update_V.F90:821-858,
 * Time/TIME_RK_V_Update_New.F90:195-230): omega floor from the production, k cap, then
 * mu_t from a second omega floor on the deviatoric strain that is not stored back. */
static void wdp_limit(const trans_params* p, ora_block* b, int I, int J, int K, double pRo, double RmiuL,
                      double prod, double ARTk, double ARTO)
{
    double dudx = DQ(1, I, J, K), dudy = DQ(2, I, J, K), dudz = DQ(3, I, J, K);
    double dvdx = DQ(4, I, J, K), dvdy = DQ(5, I, J, K), dvdz = DQ(6, I, J, K);
    double dwdx = DQ(7, I, J, K), dwdy = DQ(8, I, J, K), dwdz = DQ(9, I, J, K);
    double Sii = dudx - (dudx + dvdy + dwdz) / 3.;
    double Sjj = dvdy - (dudx + dvdy + dwdz) / 3.;
    double Skk = dwdz - (dudx + dvdy + dwdz) / 3.;
    double Sij = (dudy + dvdx) / 2.0, Sik = (dudz + dwdx) / 2.0, Sjk = (dvdz + dwdy) / 2.0;
    double Skl = sqrt(1.0 * Sii * Sii + 1.0 * Sjj * Sjj + 1.0 * Skk * Skk + 2.0 * Sij * Sij + 2.0 * Sik * Sik +
                      2.0 * Sjk * Sjk);
    /* X**0.5 is the square root (compilers evaluate the reference's **0.5 so) */
    ARTO = fmax(ARTO, sqrt(pRo * fabs(prod) / (p->prod_lmt * p->beta_star)) / p->ref);
    ARTk = fmin(ARTk, 100000.0 * RmiuL * ARTO / pRo);
    V(6, I, J, K) = ARTk;
    V(7, I, J, K) = ARTO;
    ARTO = fmax(ARTO, 7. / 8. * pRo * sqrt(Skl * Skl * 2. / 0.09) / p->ref);
    RMIU(I, J, K) = ARTk / ARTO * pRo;
}

/* ------------------------------------------------------------------------------------ */
/* Time/TIME_Update_V.F90:1-917 (non-preconditioned path; the smoothing passes only write
 * to the scratch CC and are omitted) */
void ora_update_v(const trans_params* p, int ki, int if_turb, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    double DRSa = 0.0, DUa = 0.0;
    int Ima = 0, Jma = 0, Kma = 0;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DR = fabs(D(1, I, J, K) / V(1, I, J, K));
                if (DUa <= DR) { DUa = DR; Ima = I; Jma = J; Kma = K; }
                DRSa = DRSa + DR * DR;
            }
    DRSa = sqrt(DRSa / NI1 / NJ1 / NK1);
    double DRSk = 0.0, DUk = 0.0, DRSo = 0.0, DUo = 0.0;
    int Imk = 0, Jmk = 0, Kmk = 0, Imo = 0, Jmo = 0, Kmo = 0;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DRk = fabs(D(6, I, J, K) / V(1, I, J, K));
                if (DUk <= DRk) { DUk = DRk; Imk = I; Jmk = J; Kmk = K; }
                DRSk = DRSk + DRk * DRk;
            }
    DRSk = sqrt(DRSk / NI1 / NJ1 / NK1);
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DRo = fabs(D(7, I, J, K) / V(1, I, J, K));
                if (DUo <= DRo) { DUo = DRo; Imo = I; Jmo = J; Kmo = K; }
                DRSo = DRSo + DRo * DRo;
            }
    DRSo = sqrt(DRSo / NI1 / NJ1 / NK1);

    const double om = p->omega1;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double AinRo = D(1, I, J, K), AinRx = D(2, I, J, K), AinRy = D(3, I, J, K);
                double AinRz = D(4, I, J, K), AinRe = D(5, I, J, K);
                double pRo = V(1, I, J, K);
                double roinv = 1.0 / pRo;
                if (p->if_precondition) { /* Time/TIME_Update_V.F90:522-606 */
                    double pRx = V(2, I, J, K) + AinRx, pRy = V(3, I, J, K) + AinRy;
                    double pRz = V(4, I, J, K) + AinRz;
                    double pRe = fabs(V(5, I, J, K) + AinRe);
                    pRo = fabs(pRo + AinRo);
                    /* the kinetic energy divides by the density before the update */
                    double DV = 0.5 * (pRx * pRx + pRy * pRy + pRz * pRz) * roinv;
                    double pPP = (pRe - DV) * (1.4 - 1.0);
                    if (pPP <= 0.0) {
                        pPP = fabs(pPP);
                        pRe = pPP / (1.4 - 1.0) + DV;
                    }
                    /* T from the pressure before the update */
                    double pTT = p->rxm2 * PP(I, J, K) / pRo;
                    V(1, I, J, K) = pRo;
                    V(2, I, J, K) = pRx;
                    V(3, I, J, K) = pRy;
                    V(4, I, J, K) = pRz;
                    V(5, I, J, K) = pRe;
                    PP(I, J, K) = pPP;
                    T(I, J, K) = pTT;
                    goto turb;
                }
                double pVx = V(2, I, J, K) * roinv;
                double pVy = V(3, I, J, K) * roinv;
                double pVz = V(4, I, J, K) * roinv;
                double AinX = (AinRx - pVx * AinRo) * roinv;
                double AinY = (AinRy - pVy * AinRo) * roinv;
                double AinZ = (AinRz - pVz * AinRo) * roinv;
                double AinP = (1.4 - 1.0) * (AinRe - 0.5 * AinRo * (pVx * pVx + pVy * pVy + pVz * pVz) -
                                             pRo * (AinX * pVx + AinY * pVy + AinZ * pVz));
                if (AinRo >= 0.)
                    pRo = pRo + AinRo;
                else
                    pRo = pRo * exp(AinRo / (pRo + 1.e-10));
                pRo = fabs(pRo);
                if (AinP >= 0.)
                    PP(I, J, K) = PP(I, J, K) + AinP;
                else
                    PP(I, J, K) = PP(I, J, K) * exp(AinP / (PP(I, J, K) + 1.e-10));
                PP(I, J, K) = fabs(PP(I, J, K));
                T(I, J, K) = p->rxm2 * PP(I, J, K) / pRo;
                pVx = pVx + AinX;
                pVy = pVy + AinY;
                pVz = pVz + AinZ;
                V(1, I, J, K) = pRo;
                V(2, I, J, K) = pRo * pVx;
                V(3, I, J, K) = pRo * pVy;
                V(4, I, J, K) = pRo * pVz;
                double DV = 0.5 * (pVx * pVx + pVy * pVy + pVz * pVz);
                V(5, I, J, K) = PP(I, J, K) / (1.4 - 1.0) + DV * pRo -
                                0.5 * sqr(om * RAD(I, J, K)) * pRo;

            turb:
                if (!(p->if_euler == 0 && if_turb == 1 && ki >= p->kitm)) continue;

                double AinRTk = D(6, I, J, K), AinRTO = D(7, I, J, K);
                double ARTk = V(6, I, J, K), ARTO = V(7, I, J, K);
                if (AinRTk >= 0.0)
                    ARTk = ARTk + AinRTk;
                else
                    ARTk = ARTk * exp(AinRTk / (ARTk + 1.e-10));
                ARTk = fmax(ARTk, 1.e-10);
                if (AinRTO >= 0.0)
                    ARTO = ARTO + AinRTO;
                else
                    ARTO = ARTO * exp(AinRTO / (ARTO + 1.e-10));
                ARTO = fmax(ARTO, 1.e-10);
                pRo = V(1, I, J, K);
                double VV = sqr(V(2, I, J, K)) + sqr(V(3, I, J, K)) + sqr(V(4, I, J, K));
                double PPP = (1.4 - 1.0) * V(5, I, J, K) -
                             0.5 * (1.4 - 1.0) * (VV / pRo - sqr(om * RAD(I, J, K)) * pRo);
                double TT = p->rxm2 * PPP / pRo;
                double RmiuL = sutherland(p, TT);
                double DistN = DST(I, J, K);
                double dudx = DQ(1, I, J, K), dudy = DQ(2, I, J, K), dudz = DQ(3, I, J, K);
                double dvdx = DQ(4, I, J, K), dvdy = DQ(5, I, J, K), dvdz = DQ(6, I, J, K);
                double dwdx = DQ(7, I, J, K), dwdy = DQ(8, I, J, K), dwdz = DQ(9, I, J, K);
                double Div = dudx + dvdy + dwdz;
                double prod = (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sqr(dudy + dvdx) +
                               sqr(dvdz + dwdy) + sqr(dwdx + dudz) - 2.0 / 3.0 * Div * Div);
                if (p->kind_model == 1) {
                    double arg2 = 2.0 * sqrt(ARTk / pRo) / p->beta_star / (ARTO / pRo) / DistN / p->ref;
                    arg2 = fmax(arg2, 500.0 * RmiuL / DistN / DistN / ARTO / p->ref / p->ref);
                    arg2 = tanh(arg2 * arg2);
                    double Sii = dudx, Sjj = dvdy, Skk = dwdz;
                    double Sij = (dudy + dvdx) / 2.0, Sik = (dudz + dwdx) / 2.0, Sjk = (dvdz + dwdy) / 2.0;
                    double Skl = sqrt(2.0 * (1.0 * Sii * Sii + 1.0 * Sjj * Sjj + 1.0 * Skk * Skk +
                                             2.0 * Sij * Sij + 2.0 * Sik * Sik + 2.0 * Sjk * Sjk));
                    double Cmua = fmin(1.0, p->a1 * ARTO / pRo / ((Skl)*arg2 + 1.e-10) * p->ref);
                    /* X**0.5 is the square root (compilers evaluate the reference's **0.5 so) */
                    ARTO = fmax(ARTO, sqrt(pRo * Cmua * fabs(prod) / (p->prod_lmt * p->beta_star)) / p->ref);
                    ARTk = fmin(ARTk, 100000.0 * RmiuL * ARTO / pRo / Cmua);
                    V(6, I, J, K) = ARTk;
                    V(7, I, J, K) = ARTO;
                    RMIU(I, J, K) = Cmua * ARTk / ARTO * pRo;
                } else if (p->kind_model == 2) {
                    wdp_limit(p, b, I, J, K, pRo, RmiuL, prod, ARTk, ARTO);
                }
            }
    b->res.drsa = DRSa; b->res.dua = DUa; b->res.ima = Ima; b->res.jma = Jma; b->res.kma = Kma;
    b->res.drsk = DRSk; b->res.duk = DUk; b->res.imk = Imk; b->res.jmk = Jmk; b->res.kmk = Kmk;
    b->res.drso = DRSo; b->res.duo = DUo; b->res.imo = Imo; b->res.jmo = Jmo; b->res.kmo = Kmo;
}

/* Time/TIME_Update_U.F90:1-45 */
void ora_update_u(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    int MLtemp = TRANS_ML;
    for (int L = 1; L <= MLtemp; L++)
        for (int J = 1; J <= NJ1; J++)
            for (int K = 1; K <= NK1; K++)
                for (int I = 1; I <= NI1; I++) {
                    if (p->kind_dual == 1) {
                        WT0(L, I, J, K) = V(L, I, J, K);
                    } else if (p->kind_dual == 2) {
                        WT1(L, I, J, K) = WT0(L, I, J, K);
                        WT0(L, I, J, K) = V(L, I, J, K);
                    }
                }
}

/* ------------------------------------------------------------------------------------ */
/* main.F90:66-133 */
void ora_residual(const trans_params* p, int ki, ora_block* blocks, int nblk)
{
    /* Blocks are independent between the two BC sweeps; OpenMP over blocks (the BC
     * sweeps read only interior cells of other blocks and write only own ghosts). */
    const int if_turb = ora_if_turb(p, ki);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < nblk; ib++) ora_spectral_radius(p, ki, &blocks[ib]);
    ora_bc(p, if_turb, blocks, nblk, -1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < nblk; ib++) {
        ora_spaceinterp(p, if_turb, &blocks[ib]);
        if (p->if_euler == 0) ora_diff_dqdxyz(p, if_turb, &blocks[ib]);
    }
    ora_bc(p, if_turb, blocks, nblk, 0);
    /* RCCorrection: Kind_RotMod = 0 leaves f_r1 = 1 */
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < nblk; ib++) {
        ora_block* b = &blocks[ib];
        long n = 7L * (b->ni + 5) * (b->nj + 5) * (b->nk + 5);
        memset(b->D, 0, sizeof(double) * (size_t)n);
        if (p->if_euler == 0) ora_diffusion(p, b);
        if (if_turb) ora_turb_diffusion(p, b);
        if (p->kind_spatial == 1) ora_convect_muscl2(p, b);
        if (p->kind_spatial == 2) ora_convect_muscl(p, b);
        if (p->kind_spatial == 3) ora_convect_mdcd(p, b);
        if (p->kind_spatial == 4) ora_convect_mdcdhyb(p, b);
        if (if_turb) ora_turb_convect_muscl(p, b);
        if (p->kind_roe == 1) ora_convect_roe(p, b);
        if (p->kind_roe == 2) ora_convect_roe_rotated(p, b);
        if (p->kind_roe == 3) ora_convect_ausmpw(p, b);
        if (if_turb) ora_turb_convect_upwind(p, b);
    }
}

/* main.F90:137-151 */
void ora_lusgs_update(const trans_params* p, int ki, ora_block* blocks, int nblk)
{
    const int if_turb = ora_if_turb(p, ki);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < nblk; ib++) {
        ora_lusgs(p, &blocks[ib]);
        if (if_turb) ora_turb_lusgs_sst(p, &blocks[ib]);
        ora_update_v(p, ki, if_turb, &blocks[ib]);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Time/TIME_RK_Save.F90:14-27: the state at the start of the Runge-Kutta step. */
void ora_rk_save(int itr_chld, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    if (itr_chld != 1) return;
    for (int L = 1; L <= TRANS_ML; L++)
        for (int J = 1; J <= NJ1; J++)
            for (int K = 1; K <= NK1; K++)
                for (int I = 1; I <= NI1; I++) WT2(L, I, J, K) = V(L, I, J, K);
}

/* Time/TIME_RK.F90:24-91: norms of the density residual D(1)/rho before the stage (last
 * maximum wins), then the stage of the three-stage TVD Runge-Kutta scheme on rho..rho*E. */
void ora_rk(int itr_chld, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    double DRSa = 0.0, DUa = 0.0;
    int Ima = 0, Jma = 0, Kma = 0;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DR = fabs(D(1, I, J, K) / V(1, I, J, K));
                if (DUa <= DR) { DUa = DR; Ima = I; Jma = J; Kma = K; }
                DRSa = DRSa + DR * DR;
            }
    DRSa = sqrt(DRSa / NI1 / NJ1 / NK1);
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                if (itr_chld == 1)
                    for (int L = 1; L <= 5; L++)
                        V(L, I, J, K) = V(L, I, J, K) + DTM(I, J, K) * D(L, I, J, K) / VOL(I, J, K);
                if (itr_chld == 2)
                    for (int L = 1; L <= 5; L++)
                        V(L, I, J, K) = 0.75 * WT2(L, I, J, K) + 0.25 * V(L, I, J, K) +
                                        0.25 * DTM(I, J, K) * D(L, I, J, K) / VOL(I, J, K);
                if (itr_chld == 3)
                    for (int L = 1; L <= 5; L++)
                        V(L, I, J, K) = 1.0 / 3.0 * WT2(L, I, J, K) + 2.0 / 3.0 * V(L, I, J, K) +
                                        2.0 / 3.0 * DTM(I, J, K) * D(L, I, J, K) / VOL(I, J, K);
            }
    b->res.drsa = DRSa; b->res.dua = DUa; b->res.ima = Ima; b->res.jma = Jma; b->res.kma = Kma;
}

/* Time/TIME_RK_V_Update_New.F90:75-260: pressure and temperature from the new conserved
 * variables (no rotation term), the positivity-damped k/omega update with the SST limiter
 * (its own constants: p = 0.4*rho*E - 0.2*|rho u|^2/rho and the k-omega Sutherland form
 * without Csth1), then :265-355 the k and omega norms of the LU-SGS increments. */
void ora_rk_update_v(const trans_params* p, int ki, int if_turb, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double VV = V(2, I, J, K) * V(2, I, J, K) + V(3, I, J, K) * V(3, I, J, K) +
                            V(4, I, J, K) * V(4, I, J, K);
                PP(I, J, K) = (1.4 - 1.0) * V(5, I, J, K) - 0.5 * (1.4 - 1.0) * VV / V(1, I, J, K);
                T(I, J, K) = p->rxm2 * PP(I, J, K) / V(1, I, J, K);
                if (!(p->if_euler == 0 && if_turb == 1 && ki >= p->kitm)) continue;
                double AinRTk = D(6, I, J, K), AinRTO = D(7, I, J, K);
                double ARTk = V(6, I, J, K), ARTO = V(7, I, J, K);
                if (AinRTk >= 0.0)
                    ARTk = ARTk + AinRTk;
                else
                    ARTk = ARTk * exp(AinRTk / (ARTk + 1.e-10));
                ARTk = fmax(ARTk, 1.e-10);
                if (AinRTO >= 0.0)
                    ARTO = ARTO + AinRTO;
                else
                    ARTO = ARTO * exp(AinRTO / (ARTO + 1.e-10));
                ARTO = fmax(ARTO, 1.e-10);
                double pRo = V(1, I, J, K);
                VV = sqr(V(2, I, J, K)) + sqr(V(3, I, J, K)) + sqr(V(4, I, J, K));
                double PPP = 0.4 * V(5, I, J, K) - 0.2 * VV / pRo;
                double TT = p->rxm2 * PPP / pRo;
                double RmiuL = (1.0 + p->csthlnd) / (TT + p->csthlnd) * pow(TT, 1.5);
                double DistN = DST(I, J, K);
                double dudx = DQ(1, I, J, K), dudy = DQ(2, I, J, K), dudz = DQ(3, I, J, K);
                double dvdx = DQ(4, I, J, K), dvdy = DQ(5, I, J, K), dvdz = DQ(6, I, J, K);
                double dwdx = DQ(7, I, J, K), dwdy = DQ(8, I, J, K), dwdz = DQ(9, I, J, K);
                double Div = dudx + dvdy + dwdz;
                double prod = (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sqr(dudy + dvdx) +
                               sqr(dvdz + dwdy) + sqr(dwdx + dudz) - 2.0 / 3.0 * Div * Div);
                if (p->kind_model == 1) {
                    double arg2 = 2.0 * sqrt(ARTk / pRo) / p->beta_star / (ARTO / pRo) / DistN / p->ref;
                    arg2 = fmax(arg2, 500.0 * RmiuL / DistN / DistN / ARTO / p->ref / p->ref);
                    arg2 = tanh(arg2 * arg2);
                    double Sii = dudx, Sjj = dvdy, Skk = dwdz;
                    double Sij = (dudy + dvdx) / 2.0, Sik = (dudz + dwdx) / 2.0, Sjk = (dvdz + dwdy) / 2.0;
                    double Skl = sqrt(2.0 * (1.0 * Sii * Sii + 1.0 * Sjj * Sjj + 1.0 * Skk * Skk +
                                             2.0 * Sij * Sij + 2.0 * Sik * Sik + 2.0 * Sjk * Sjk));
                    double Cmua = fmin(1.0, p->a1 * ARTO / pRo / ((Skl)*arg2 + 1.e-10) * p->ref);
                    /* X**0.5 is the square root (compilers evaluate the reference's **0.5 so) */
                    ARTO = fmax(ARTO, sqrt(pRo * Cmua * fabs(prod) / (p->prod_lmt * p->beta_star)) / p->ref);
                    ARTk = fmin(ARTk, 100000.0 * RmiuL * ARTO / pRo / Cmua);
                    V(6, I, J, K) = ARTk;
                    V(7, I, J, K) = ARTO;
                    RMIU(I, J, K) = Cmua * ARTk / ARTO * pRo;
                } else if (p->kind_model == 2) {
                    wdp_limit(p, b, I, J, K, pRo, RmiuL, prod, ARTk, ARTO);
                }
            }
    double DRSk = 0.0, DUk = 0.0, DRSo = 0.0, DUo = 0.0;
    int Imk = 0, Jmk = 0, Kmk = 0, Imo = 0, Jmo = 0, Kmo = 0;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DRk = fabs(D(6, I, J, K) / V(1, I, J, K));
                if (DUk <= DRk) { DUk = DRk; Imk = I; Jmk = J; Kmk = K; }
                DRSk = DRSk + DRk * DRk;
            }
    DRSk = sqrt(DRSk / NI1 / NJ1 / NK1);
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DRo = fabs(D(7, I, J, K) / V(1, I, J, K));
                if (DUo <= DRo) { DUo = DRo; Imo = I; Jmo = J; Kmo = K; }
                DRSo = DRSo + DRo * DRo;
            }
    DRSo = sqrt(DRSo / NI1 / NJ1 / NK1);
    b->res.drsk = DRSk; b->res.duk = DUk; b->res.imk = Imk; b->res.jmk = Jmk; b->res.kmk = Kmk;
    b->res.drso = DRSo; b->res.duo = DUo; b->res.imo = Imo; b->res.jmo = Jmo; b->res.kmo = Kmo;
}

/* main.F90:158-170 */
void ora_rk_update(const trans_params* p, int ki, int itr_chld, ora_block* blocks, int nblk)
{
    const int if_turb = ora_if_turb(p, ki);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < nblk; ib++) {
        ora_rk_save(itr_chld, &blocks[ib]);
        ora_rk(itr_chld, &blocks[ib]);
        if (if_turb) ora_turb_lusgs_sst(p, &blocks[ib]);
        ora_rk_update_v(p, ki, if_turb, &blocks[ib]);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Plane integrals.  Face state: the volume weighting of Output/OUTPUT_Residual_NS.F90:19-32
 * (k outer, then the other in-plane index); p0 = p (1 + 0.2 M^2)^3.5 as
 * Output/Average_Field.F90:60-66 forms it per cell, T0 = T (1 + 0.2 M^2), with the stored
 * relative velocity; out[4], out[5] the same with the absolute velocity u, v - omega z_f,
 * w + omega y_f (face centre by the same weighting; rotation omega about x). */
void ora_plane_integrals(ora_block* b, int dir, int index, double omega, double out[6])
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    const int n0 = dir == 1 ? NJ1 : NI1;  /* fast in-plane index */
    const int n1 = dir == 3 ? NJ1 : NK1;  /* slow in-plane index */
    double msum = 0.0, area = 0.0, psum = 0.0, tsum = 0.0, pasum = 0.0, tasum = 0.0;
    for (int q1 = 1; q1 <= n1; q1++)
        for (int q0 = 1; q0 <= n0; q0++) {
            int i, j, k, im, jm, km;
            if (dir == 1) { i = index; j = q0; k = q1; }
            else if (dir == 2) { i = q0; j = index; k = q1; }
            else { i = q0; j = q1; k = index; }
            im = i - (dir == 1); jm = j - (dir == 2); km = k - (dir == 3);
            double wei1 = VOL(im, jm, km), wei2 = VOL(i, j, k), wei = wei1 + wei2;
            double ro = (wei1 * V(1, i, j, k) + wei2 * V(1, im, jm, km)) / wei;
            double ru = (wei1 * V(2, i, j, k) + wei2 * V(2, im, jm, km)) / wei;
            double rv = (wei1 * V(3, i, j, k) + wei2 * V(3, im, jm, km)) / wei;
            double rw = (wei1 * V(4, i, j, k) + wei2 * V(4, im, jm, km)) / wei;
            double pf = (wei1 * PP(i, j, k) + wei2 * PP(im, jm, km)) / wei;
            double tf = (wei1 * T(i, j, k) + wei2 * T(im, jm, km)) / wei;
            double s1 = SD(dir, 1, i, j, k), s2 = SD(dir, 2, i, j, k), s3 = SD(dir, 3, i, j, k);
            double dm = ru * s1 + rv * s2 + rw * s3;
            double u = ru / ro, v = rv / ro, w = rw / ro;
            double m2 = (u * u + v * v + w * w) / (1.4 * pf / ro);
            double fac = 1.0 + 0.2 * m2;
            msum = msum + dm;
            area = area + sqrt(s1 * s1 + s2 * s2 + s3 * s3);
            psum = psum + pf * pow(fac, 3.5) * dm;
            tsum = tsum + tf * fac * dm;
            double yf = (wei1 * YC(i, j, k) + wei2 * YC(im, jm, km)) / wei;
            double zf = (wei1 * ZC(i, j, k) + wei2 * ZC(im, jm, km)) / wei;
            double va = v - omega * zf, wa = w + omega * yf;
            double fa = 1.0 + 0.2 * ((u * u + va * va + wa * wa) / (1.4 * pf / ro));
            pasum = pasum + pf * pow(fa, 3.5) * dm;
            tasum = tasum + tf * fa * dm;
        }
    out[0] = msum;
    out[1] = area;
    out[2] = psum / msum;
    out[3] = tsum / msum;
    out[4] = pasum / msum;
    out[5] = tasum / msum;
}

/* Threads an OpenMP region of the oracle uses. */
int ora_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ora_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* main.F90:48-211 (no output) */
void ora_iterate(const trans_params* p, int* ki, int n_iter, ora_block* blocks, int nblk)
{
    for (int it = 0; it < n_iter; it++) {
        int nsub = p->kind_dual != 0 ? p->kchld : 1;
        if (p->if_rk == 1) nsub = 3;
        for (int s = 0; s < nsub; s++) {
            ora_residual(p, *ki, blocks, nblk);
            if (p->if_rk == 1)
                ora_rk_update(p, *ki, s + 1, blocks, nblk);
            else
                ora_lusgs_update(p, *ki, blocks, nblk);
        }
        *ki += 1;
        if (p->kind_dual != 0)
            for (int ib = 0; ib < nblk; ib++) ora_update_u(p, &blocks[ib]);
    }
}
