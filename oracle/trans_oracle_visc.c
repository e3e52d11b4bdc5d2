/*
 * CPU oracle, viscous / RANS-SST part -- TEST INFRASTRUCTURE ONLY (see trans_oracle.h).
 *
 * Plain-C FP64 restatement of the reference's gradient, viscous-flux, MDCD, AUSMPW+,
 * k-omega and SST routines, in the reference's loop order and operation order.
 * Integer powers x**2 / x**3 / x**4 are evaluated as repeated products
 * (x*x, (x*x)*x, (x*x)*(x*x)), the usual expansion of constant integer exponents.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_idx.h"
#include "trans_oracle.h"

#define sqr ora_sqr
#define sutherland ora_sutherland
double ora_sqr(double x);
double ora_sutherland(const trans_params* p, double tt);

static double pow4(double x)
{
    double x2 = x * x;
    return x2 * x2;
}

/* ------------------------------------------------------------------------------------ */
/* Green-Gauss gradient of one face-average component c (VL(.,c,..)) for cell (i,j,k),
 * component xyz (1..3): Space/DIFF_DqDxyz.F90:30-37. */
static double gg(ora_block* b, int c, int x, int i, int j, int k)
{
    double disfe = VL(1, c, i + 1, j, k), disfw = VL(1, c, i, j, k);
    double disfn = VL(2, c, i, j + 1, k), disfs = VL(2, c, i, j, k);
    double disft = VL(3, c, i, j, k + 1), disfb = VL(3, c, i, j, k);
    return disfe * SD(1, x, i + 1, j, k) - disfw * SD(1, x, i, j, k) + disfn * SD(2, x, i, j + 1, k) -
           disfs * SD(2, x, i, j, k) + disft * SD(3, x, i, j, k + 1) - disfb * SD(3, x, i, j, k);
}

static void grad_cell(ora_block* b, int i, int j, int k, int turb)
{
    double volinv = 1.0 / VOL(i, j, k);
    static const int comp[4] = {1, 2, 3, 5};
    for (int q = 0; q < 4; q++)
        for (int x = 1; x <= 3; x++) DQ(3 * q + x, i, j, k) = volinv * gg(b, comp[q], x, i, j, k);
    if (turb)
        for (int q = 0; q < 2; q++)
            for (int x = 1; x <= 3; x++) DQ(13 + 3 * q + x - 1, i, j, k) = volinv * gg(b, 6 + q, x, i, j, k);
}

/* Space/DIFF_DqDxyz.F90:1-408 */
void ora_diff_dqdxyz(const trans_params* p, int if_turb, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int k = 1; k <= NK1; k++)
        for (int j = 1; j <= NJ1; j++)
            for (int i = 1; i <= NI1; i++) {
                double volinv = 1.0 / VOL(i, j, k);
                static const int comp[4] = {1, 2, 3, 5};
                for (int q = 0; q < 4; q++)
                    for (int x = 1; x <= 3; x++) DQ(3 * q + x, i, j, k) = volinv * gg(b, comp[q], x, i, j, k);
            }
    /* first ghost layer (overwritten by BC(0) wherever a BC patch covers the face) */
    for (int k = 1; k <= NK1; k++)
        for (int j = 1; j <= NJ1; j++)
            for (int i = 0; i <= NI; i += NI) grad_cell(b, i, j, k, 0);
    for (int k = 1; k <= NK1; k++)
        for (int j = 0; j <= NJ; j += NJ)
            for (int i = 1; i <= NI1; i++) grad_cell(b, i, j, k, 0);
    for (int k = 0; k <= NK; k += NK)
        for (int j = 1; j <= NJ1; j++)
            for (int i = 1; i <= NI1; i++) grad_cell(b, i, j, k, 0);
    if (if_turb)
        for (int k = 1; k <= NK1; k++)
            for (int j = 1; j <= NJ1; j++)
                for (int i = 1; i <= NI1; i++) {
                    double volinv = 1.0 / VOL(i, j, k);
                    for (int q = 0; q < 2; q++)
                        for (int x = 1; x <= 3; x++)
                            DQ(13 + 3 * q + x - 1, i, j, k) = volinv * gg(b, 6 + q, x, i, j, k);
                }
    /* DES blending Fscheme and shock sensor (:295-384; CH1 = 3, CH2 = 1, CH3 = 2,
     * Cdes_epsi = 0.61, Cdes_omeg = 0.78, cmu = 0.09 from Module_Global.F90:64-124).
     * Fscheme feeds the rotated Roe flux. */
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double DuDx = DQ(1, I, J, K), DuDy = DQ(2, I, J, K), DuDz = DQ(3, I, J, K);
                double DvDx = DQ(4, I, J, K), DvDy = DQ(5, I, J, K), DvDz = DQ(6, I, J, K);
                double DwDx = DQ(7, I, J, K), DwDy = DQ(8, I, J, K), DwDz = DQ(9, I, J, K);
                double W12 = (DuDy - DvDx) / 2., W13 = (DuDz - DwDx) / 2., W23 = (DvDz - DwDy) / 2.;
                double S11 = (DuDx + DuDx) / 2., S22 = (DvDy + DvDy) / 2., S33 = (DwDz + DwDz) / 2.;
                double S12 = (DuDy + DvDx) / 2., S13 = (DuDz + DwDx) / 2., S23 = (DwDy + DvDz) / 2.;
                double Vort = sqrt(2. * (2. * W12 * W12 + 2. * W13 * W13 + 2. * W23 * W23));
                double Skl = sqrt(2. * (1. * S11 * S11 + 1. * S22 * S22 + 1. * S33 * S33 + 2. * S12 * S12 +
                                        2. * S13 * S13 + 2. * S23 * S23));
                double DivU = sqrt(DuDx * DuDx + DvDy * DvDy + DwDz * DwDz);
                double Den = V(1, I, J, K), RmiuT = RMIU(I, J, K), TT = T(I, J, K);
                double RMIUL = (1. + p->csthlnd) / (TT * p->csth1 + p->csthlnd) * pow(p->csth1 * TT, 1.5);
                double vK = V(6, I, J, K) / Den, vOmega = V(7, I, J, K) / Den;
                double DLTxyz = ALENG(I, J, K);
                double tau = 1.0;
                double dkdx = DQ(13, I, J, K), dkdy = DQ(14, I, J, K), dkdz = DQ(15, I, J, K);
                double dodx = DQ(16, I, J, K), dody = DQ(17, I, J, K), dodz = DQ(18, I, J, K);
                double Dist = DST(I, J, K);
                double dTKO = dkdx * dodx + dkdy * dody + dkdz * dodz;
                double CDKO = fmax(2.0 * Den * p->sigo2 / vOmega * (dTKO), 1.0e-10);
                double arg1 = fmax(sqrt(vK) / p->beta_star / vOmega / Dist / p->ref,
                                   500.0 * RMIUL / Den / Dist / Dist / vOmega / p->ref / p->ref);
                arg1 = fmin(arg1, 4.0 * Den * p->sigo2 * vK / CDKO / Dist / Dist);
                double FunC1 = tanh(pow4(arg1));
                double aLgrd = ((1. - FunC1) * 0.61 + FunC1 * 0.78) * DLTxyz;
                double B0 = 2. * Vort * fmax(Vort, Skl) / fmax((Vort * Vort + Skl * Skl) / 2., p->alscale * 0.1);
                double g0 = tanh(B0 * B0 * B0 * B0);
                /* X**0.5 is the square root (compilers evaluate the reference's **0.5 so) */
                double sok = fmax(sqrt((Vort * Vort + Skl * Skl) / 2.), 0.1 / tau);
                double alturb = sqrt((RmiuT + RMIUL) / Den / (p->ref * sok * pow(0.09, 1.5)));
                double A0 = 1. * fmax(((aLgrd / alturb / g0) - 0.1), 0.0);
                FSCHEME(I, J, K) = fmax(1. * tanh(pow(A0, 3.)), p->f_limiter1);
                SHOCK(I, J, K) = DivU * DivU / fmax(DivU * DivU + Vort * Vort, 1.0);
            }
    /* Fscheme's first ghost layer (:399-404) */
    for (int k = -2; k <= NK + 2; k++)
        for (int j = -2; j <= NJ + 2; j++) FSCHEME(0, j, k) = FSCHEME(1, j, k);
    for (int k = -2; k <= NK + 2; k++)
        for (int i = -2; i <= NI + 2; i++) FSCHEME(i, 0, k) = FSCHEME(i, 1, k);
    for (int j = -2; j <= NJ + 2; j++)
        for (int i = -2; i <= NI + 2; i++) FSCHEME(i, j, 0) = FSCHEME(i, j, 1);
    for (int k = -2; k <= NK + 2; k++)
        for (int j = -2; j <= NJ + 2; j++) FSCHEME(NI, j, k) = FSCHEME(NI1, j, k);
    for (int k = -2; k <= NK + 2; k++)
        for (int i = -2; i <= NI + 2; i++) FSCHEME(i, NJ, k) = FSCHEME(i, NJ1, k);
    for (int j = -2; j <= NJ + 2; j++)
        for (int i = -2; i <= NI + 2; i++) FSCHEME(i, j, NK) = FSCHEME(i, j, NK1);
    /* ghost extension (:388-397) */
    for (int L = 1; L <= 3; L++) {
        for (int k = -2; k <= NK + 2; k++)
            for (int j = -2; j <= NJ + 2; j++) SHOCK(1 - L, j, k) = SHOCK(1, j, k);
        for (int k = -2; k <= NK + 2; k++)
            for (int i = -2; i <= NI + 2; i++) SHOCK(i, 1 - L, k) = SHOCK(i, 1, k);
        for (int j = -2; j <= NJ + 2; j++)
            for (int i = -2; i <= NI + 2; i++) SHOCK(i, j, 1 - L) = SHOCK(i, j, 1);
        for (int k = -2; k <= NK + 2; k++)
            for (int j = -2; j <= NJ + 2; j++) SHOCK(NI1 + L, j, k) = SHOCK(NI1, j, k);
        for (int k = -2; k <= NK + 2; k++)
            for (int i = -2; i <= NI + 2; i++) SHOCK(i, NJ1 + L, k) = SHOCK(i, NJ1, k);
        for (int j = -2; j <= NJ + 2; j++)
            for (int i = -2; i <= NI + 2; i++) SHOCK(i, j, NK1 + L) = SHOCK(i, j, NK1);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Face gradient of one quantity in direction d at face (i,j,k): volume-weighted cell
 * gradients plus the face-centred correction (Space/DIFFUSION_Centered.F90:83-105). */
static void face_grad(ora_block* b, int d, int i, int j, int k, int dq0, int vlc, double qp, double qn,
                      double all, double alr, double alsinv, double out[3])
{
    const int di = d == 1, dj = d == 2, dk = d == 3;
    const int im = i - di, jm = j - dj, km = k - dk;
    const int ip = i + di, jp = j + dj, kp = k + dk;
    double disfpp = VL(d, vlc, ip, jp, kp);
    double disfnn = VL(d, vlc, im, jm, km);
    for (int x = 1; x <= 3; x++) {
        double g1 = DQ(dq0 + x - 1, im, jm, km), g0 = DQ(dq0 + x - 1, i, j, k);
        double gl = all * g1 + alr * g0;
        out[x - 1] = alsinv * (-SD(d, x, ip, jp, kp) * (disfpp - qp) - SD(d, x, im, jm, km) * (qn - disfnn) +
                               SD(d, x, i, j, k) * (qp - qn)) +
                     gl;
    }
}

static int wall_face(ora_block* b, int d, int i, int j, int k)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1;
    if (d == 1) {
        if (i == 1 && b->mw[0] && b->mw[0][(j - 1) + NJ1 * (k - 1)] == 1) return 1;
        if (i == NI && b->mw[1] && b->mw[1][(j - 1) + NJ1 * (k - 1)] == 1) return 1;
    } else if (d == 2) {
        if (j == 1 && b->mw[2] && b->mw[2][(i - 1) + NI1 * (k - 1)] == 1) return 1;
        if (j == NJ && b->mw[3] && b->mw[3][(i - 1) + NI1 * (k - 1)] == 1) return 1;
    } else {
        if (k == 1 && b->mw[4] && b->mw[4][(i - 1) + NI1 * (j - 1)] == 1) return 1;
        if (k == NK && b->mw[5] && b->mw[5][(i - 1) + NI1 * (j - 1)] == 1) return 1;
    }
    return 0;
}

/* Space/DIFFUSION_Centered.F90:1-827 */
void ora_diffusion(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    const double NMD = -2. / 3.;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    const int im = i - di, jm = j - dj, km = k - dk;
                    int wall = wall_face(b, d, i, j, k);
                    double alsinv = 1.0 / (VOL(i, j, k) + VOL(im, jm, km));
                    double all = VOL(im, jm, km) * alsinv;
                    double alr = VOL(i, j, k) * alsinv;
                    if (wall) alsinv = 0.0;
                    double gu[3], gv[3], gw[3], gt[3];
                    face_grad(b, d, i, j, k, 1, 1, V(2, i, j, k) / V(1, i, j, k),
                              V(2, im, jm, km) / V(1, im, jm, km), all, alr, alsinv, gu);
                    face_grad(b, d, i, j, k, 4, 2, V(3, i, j, k) / V(1, i, j, k),
                              V(3, im, jm, km) / V(1, im, jm, km), all, alr, alsinv, gv);
                    face_grad(b, d, i, j, k, 7, 3, V(4, i, j, k) / V(1, i, j, k),
                              V(4, im, jm, km) / V(1, im, jm, km), all, alr, alsinv, gw);
                    face_grad(b, d, i, j, k, 10, 5, T(i, j, k), T(im, jm, km), all, alr, alsinv, gt);
                    alsinv = 1.0 / (ALAGM(d, i, j, k) + ALAGM(d, im, jm, km));
                    all = ALAGM(d, i, j, k) * alsinv;
                    alr = ALAGM(d, im, jm, km) * alsinv;
                    double TTL = VL(d, 5, i, j, k);
                    double RmiuL = sutherland(p, TTL);
                    double RmiuTL = all * RMIU(im, jm, km) + alr * RMIU(i, j, k);
                    if (wall) RmiuTL = 0.0;
                    double DuDxL = gu[0], DuDyL = gu[1], DuDzL = gu[2];
                    double DvDxL = gv[0], DvDyL = gv[1], DvDzL = gv[2];
                    double DwDxL = gw[0], DwDyL = gw[1], DwDzL = gw[2];
                    double TxxLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DuDxL);
                    double TyyLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DvDyL);
                    double TzzLam = RmiuL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DwDzL);
                    double TxyLam = RmiuL * (DuDyL + DvDxL);
                    double TyzLam = RmiuL * (DvDzL + DwDyL);
                    double TzxLam = RmiuL * (DuDzL + DwDxL);
                    double TxxTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DuDxL);
                    double TyyTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DvDyL);
                    double TzzTur = RmiuTL * (NMD * (DuDxL + DvDyL + DwDzL) + 2.0 * DwDzL);
                    double TxyTur = RmiuTL * (DuDyL + DvDxL);
                    double TyzTur = RmiuTL * (DvDzL + DwDyL);
                    double TzxTur = RmiuTL * (DuDzL + DwDxL);
                    double Txx = TxxLam + TxxTur, Tyy = TyyLam + TyyTur, Tzz = TzzLam + TzzTur;
                    double Txy = TxyLam + TxyTur, Tyz = TyzLam + TyzTur, Tzx = TzxLam + TzxTur;
                    const double prl = 0.72, prte = 0.9;
                    double cp = 1.4 * p->ppf / (1.4 - 1.0);
                    double akmu = cp * (RmiuL / prl + RmiuTL / prte);
                    double ucc = VL(d, 1, i, j, k), vcc = VL(d, 2, i, j, k), wcc = VL(d, 3, i, j, k);
                    double Phix = akmu * gt[0] + ucc * Txx + vcc * Txy + wcc * Tzx;
                    double Phiy = akmu * gt[1] + ucc * Txy + vcc * Tyy + wcc * Tyz;
                    double Phiz = akmu * gt[2] + ucc * Tzx + vcc * Tyz + wcc * Tzz;
                    double Dx = SD(d, 1, i, j, k), Dy = SD(d, 2, i, j, k), Dz = SD(d, 3, i, j, k);
                    double F[5];
                    F[0] = 0.0;
                    F[1] = (Txx * Dx + Txy * Dy + Tzx * Dz) / p->ref;
                    F[2] = (Txy * Dx + Tyy * Dy + Tyz * Dz) / p->ref;
                    F[3] = (Tzx * Dx + Tyz * Dy + Tzz * Dz) / p->ref;
                    F[4] = (Phix * Dx + Phiy * Dy + Phiz * Dz) / p->ref;
                    for (int L = 1; L <= 5; L++) {
                        D(L, i, j, k) = D(L, i, j, k) - F[L - 1];
                        D(L, im, jm, km) = D(L, im, jm, km) + F[L - 1];
                    }
                }
    }
    /* rotating-frame source (:811-822) */
    for (int k = 1; k <= NK1; k++)
        for (int j = 1; j <= NJ1; j++)
            for (int i = 1; i <= NI1; i++) {
                double RoW = p->omega1;
                double ss3 = 2.0 * RoW * V(4, i, j, k) + RoW * YC(i, j, k) * p->omega1 * V(1, i, j, k);
                double ss4 = -2.0 * RoW * V(3, i, j, k) + RoW * ZC(i, j, k) * p->omega1 * V(1, i, j, k);
                D(3, i, j, k) = D(3, i, j, k) + ss3 * VOL(i, j, k);
                D(4, i, j, k) = D(4, i, j, k) + ss4 * VOL(i, j, k);
            }
}

/* Space/TURBTWO_DIFFUSION_Centered.F90:1-555 (SST, Kind_model = 1) */
void ora_turb_diffusion(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    const int im = i - di, jm = j - dj, km = k - dk;
                    int wall = wall_face(b, d, i, j, k);
                    double alsinv = 1. / (VOL(i, j, k) + VOL(im, jm, km));
                    double all = VOL(im, jm, km) * alsinv;
                    double alr = VOL(i, j, k) * alsinv;
                    if (wall) alsinv = 0.0;
                    double gk[3], go[3];
                    face_grad(b, d, i, j, k, 13, 6, V(6, i, j, k) / V(1, i, j, k),
                              V(6, im, jm, km) / V(1, im, jm, km), all, alr, alsinv, gk);
                    face_grad(b, d, i, j, k, 16, 7, V(7, i, j, k) / V(1, i, j, k),
                              V(7, im, jm, km) / V(1, im, jm, km), all, alr, alsinv, go);
                    alsinv = 1.0 / (ALAGM(d, i, j, k) + ALAGM(d, im, jm, km));
                    all = ALAGM(d, i, j, k) * alsinv;
                    alr = ALAGM(d, im, jm, km) * alsinv;
                    double TT = T(i, j, k), TT1 = T(im, jm, km);
                    double TTL = all * TT1 + alr * TT;
                    double RmiuL = (1.0 + p->csthlnd) / (TTL + p->csthlnd) * pow(TTL, 1.5);
                    double RmiuTL = all * RMIU(im, jm, km) + alr * RMIU(i, j, k);
                    if (wall) RmiuTL = 0.0;
                    double SigkAB, SigoAB;
                    if (p->kind_model == 2) { /* WD+ (:168-170): unblended sigmas */
                        SigkAB = p->sigk1;
                        SigoAB = p->sigo1;
                    } else {
                        double DEN0 = V(1, i, j, k);
                        double TKK0 = V(6, i, j, k) / DEN0, TOO0 = V(7, i, j, k) / DEN0;
                        double DEN1 = V(1, im, jm, km);
                        double TKK1 = V(6, im, jm, km) / DEN1, TOO1 = V(7, im, jm, km) / DEN1;
                        double Den = all * DEN1 + alr * DEN0;
                        double TKK = all * TKK1 + alr * TKK0;
                        double TOO = all * TOO1 + alr * TOO0;
                        double dTKO = gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2];
                        double CDKO = fmax(2.0 * Den * p->sigo2 / TOO * (dTKO), 1.0e-10);
                        double distfc = alr * DST(i, j, k) + all * DST(im, jm, km);
                        double arg1 = fmax(sqrt(TKK) / p->beta_star / TOO / distfc / p->ref,
                                           500.0 * RmiuL / Den / distfc / distfc / TOO / p->ref / p->ref);
                        arg1 = fmin(arg1, 4.0 * Den * p->sigo2 * TKK / CDKO / distfc / distfc);
                        double fc1c = tanh(pow4(arg1));
                        SigkAB = fc1c * p->sigk1 + (1. - fc1c) * p->sigk2;
                        SigoAB = fc1c * p->sigo1 + (1. - fc1c) * p->sigo2;
                    }
                    double vmute = RmiuL + RmiuTL * SigkAB;
                    double vmuoq = RmiuL + RmiuTL * SigoAB;
                    double Dx = SD(d, 1, i, j, k), Dy = SD(d, 2, i, j, k), Dz = SD(d, 3, i, j, k);
                    double Fk = (vmute * (gk[0] * Dx + gk[1] * Dy + gk[2] * Dz)) / p->ref;
                    double Fo = (vmuoq * (go[0] * Dx + go[1] * Dy + go[2] * Dz)) / p->ref;
                    D(6, i, j, k) = D(6, i, j, k) - Fk;
                    D(6, im, jm, km) = D(6, im, jm, km) + Fk;
                    D(7, i, j, k) = D(7, i, j, k) - Fo;
                    D(7, im, jm, km) = D(7, im, jm, km) + Fo;
                }
    }
}

/* ------------------------------------------------------------------------------------ */
/* One MDCD reconstruction from six values s[0..5] (cells f-3..f+2 for the left state;
 * the right state passes them mirrored): Space/CONVECT_MDCD.F90:41-88. */
static double mdcd_rec(double gama, double v_2, double v_1, double v0, double v1, double v2, double v3)
{
    const double beta = sqrt(13.0 / 12.0), epsm = 1.0e-6;
    double aw0 = 1.5 * (0.0463783 + gama);
    double aw1 = 0.5 - 1.5 * 0.0463783 + 4.5 * gama;
    double aw2 = 0.5 - 1.5 * 0.0463783 - 4.5 * gama;
    double aw3 = 1.5 * (0.0463783 - gama);
    double ISK[4];
    ISK[0] = sqr(0.5 * v_2 - 2.0 * v_1 + 1.5 * v0) + sqr(beta * v_2 - 2.0 * beta * v_1 + beta * v0);
    ISK[1] = sqr(-0.5 * v_1 + 0.0 * v0 + 0.5 * v1) + sqr(beta * v_1 - 2.0 * beta * v0 + beta * v1);
    ISK[2] = sqr(-1.5 * v0 + 2.0 * v1 - 0.5 * v2) + sqr(beta * v0 - 2.0 * beta * v1 + beta * v2);
    ISK[3] = sqr(2.5 * v1 - 4.0 * v2 + 1.5 * v3) + sqr(beta * v1 - 2.0 * beta * v2 + beta * v3);
    ISK[3] = fmax(fmax(fmax(ISK[0], ISK[1]), ISK[2]), ISK[3]);
    double ak0 = aw0 / sqr(epsm + ISK[0]);
    double ak1 = aw1 / sqr(epsm + ISK[1]);
    double ak2 = aw2 / sqr(epsm + ISK[2]);
    double ak3 = aw3 / sqr(epsm + ISK[3]);
    double w0 = ak0 / (ak0 + ak1 + ak2 + ak3);
    double w1 = ak1 / (ak0 + ak1 + ak2 + ak3);
    double w2 = ak2 / (ak0 + ak1 + ak2 + ak3);
    double w3 = ak3 / (ak0 + ak1 + ak2 + ak3);
    double qq0 = 2. / 6. * v_2 - 7. / 6. * v_1 + 11. / 6. * v0;
    double qq1 = -1. / 6. * v_1 + 5. / 6. * v0 + 2. / 6. * v1;
    double qq2 = 2. / 6. * v0 + 5. / 6. * v1 - 1. / 6. * v2;
    double qq3 = 11. / 6. * v1 - 7. / 6. * v2 + 1. / 3. * v3;
    return w0 * qq0 + w1 * qq1 + w2 * qq2 + w3 * qq3;
}

double ora_mdcd_rec(double gama, const double s[6], int right)
{
    if (!right) return mdcd_rec(gama, s[0], s[1], s[2], s[3], s[4], s[5]);
    return mdcd_rec(gama, s[5], s[4], s[3], s[2], s[1], s[0]);
}

/* Space/CONVECT_MDCDHYB.F90:1-477: the MDCD reconstruction of u, v, w, p, T with the
 * six-cell shock maximum as sensor in all three directions, switched to a half-slope
 * ALIMITER reconstruction from cell values where the sensor reaches F_limiter2; the face
 * radii keep the SPACEINTERP averages. */
double ora_alimiter(double var1, double var2);
void ora_convect_mdcdhyb(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? b->ni : NI1, jhi = dj ? b->nj : NJ1, khi = dk ? b->nk : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    double q[5][6], sensor = 0.0;
                    for (int n = 0; n < 6; n++) {
                        const int ii = i + di * (n - 3), jj = j + dj * (n - 3), kk = k + dk * (n - 3);
                        q[0][n] = V(2, ii, jj, kk) / V(1, ii, jj, kk);
                        q[1][n] = V(3, ii, jj, kk) / V(1, ii, jj, kk);
                        q[2][n] = V(4, ii, jj, kk) / V(1, ii, jj, kk);
                        q[3][n] = PP(ii, jj, kk);
                        q[4][n] = T(ii, jj, kk);
                        const double sh = SHOCK(ii, jj, kk);
                        sensor = n == 0 ? sh : fmax(sensor, sh);
                    }
                    const double gama = p->gama_min + sensor * (p->gama_max - p->gama_min);
                    for (int L = 1; L <= 5; L++) {
                        const double* v = q[L - 1];
                        VL(d, L, i, j, k) = mdcd_rec(gama, v[0], v[1], v[2], v[3], v[4], v[5]);
                        if (sensor >= p->f_limiter2)
                            VL(d, L, i, j, k) = v[2] + 0.5 * ora_alimiter(v[3] - v[2], v[2] - v[1]);
                        VR(d, L, i, j, k) = mdcd_rec(gama, v[5], v[4], v[3], v[2], v[1], v[0]);
                        if (sensor >= p->f_limiter2)
                            VR(d, L, i, j, k) = v[3] - 0.5 * ora_alimiter(v[3] - v[2], v[4] - v[3]);
                    }
                }
    }
}

/* Space/CONVECT_MDCD.F90:1-668.  The k-direction radius block stands outside its loop
 * (:450-521) and writes one ghost corner only: k-face radii keep the face averages. */
void ora_convect_mdcd(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    double s[6][6];
                    double sh[6];
                    for (int n = 0; n < 6; n++) {
                        int ii = i + di * (n - 3), jj = j + dj * (n - 3), kk = k + dk * (n - 3);
                        s[0][n] = V(2, ii, jj, kk) / V(1, ii, jj, kk);
                        s[1][n] = V(3, ii, jj, kk) / V(1, ii, jj, kk);
                        s[2][n] = V(4, ii, jj, kk) / V(1, ii, jj, kk);
                        s[3][n] = PP(ii, jj, kk);
                        s[4][n] = T(ii, jj, kk);
                        s[5][n] = RAD(ii, jj, kk);
                        sh[n] = SHOCK(ii, jj, kk);
                    }
                    double sensor = fmax(fmax(fmax(fmax(fmax(sh[0], sh[1]), sh[2]), sh[3]), sh[4]), sh[5]);
                    if (d == 3) sensor = fmax(sensor, 0.0);
                    double gama = p->gama_min + sensor * (p->gama_max - p->gama_min);
                    if (d != 3) {
                        RCCL(d, i, j, k) = ora_mdcd_rec(gama, s[5], 0);
                        RCCR(d, i, j, k) = ora_mdcd_rec(gama, s[5], 1);
                    }
                    for (int L = 1; L <= 5; L++) {
                        VL(d, L, i, j, k) = ora_mdcd_rec(gama, s[L - 1], 0);
                        VR(d, L, i, j, k) = ora_mdcd_rec(gama, s[L - 1], 1);
                    }
                }
    }
}

/* Space/TURBTWO_CONVECT_MUSCL.F90:1-145 */
void ora_turb_convect_muscl(const trans_params* p, ora_block* b)
{
    (void)p;
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int i = 0; i <= NI + 1; i++)
        for (int j = 0; j <= NJ + 1; j++)
            for (int k = 0; k <= NK + 1; k++)
                for (int L = 6; L <= 7; L++) {
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
                    for (int L = 6; L <= 7; L++) {
                        double TKK = V(L, i, j, k) / V(1, i, j, k);
                        double DR = VL(d, L, ip, jp, kp) - TKK;
                        double DL = TKK - VR(d, L, i, j, k);
                        VL(d, L, ip, jp, kp) = TKK + ora_alimiter(DR, DL);
                        VR(d, L, i, j, k) = TKK - ora_alimiter(DL, DR);
                    }
                }
    }
}

/* Space/CONVECT_AUSMPW_Plus.F90:204-317 (3/16 is an integer division = 0) */
void ora_ausmpw_common(const trans_params* p, double PLo, double PRo, double TLo, double TRo,
                       double ULo, double URo, double VLo, double VRo, double WLo, double WRo,
                       double sav1, double sav2, double sav3, double sav, double out[5])
{
    const double rcpcv = p->ppf, tiny = p->tiny;
    double DENLo = PLo / (rcpcv * TLo);
    double DENRo = PRo / (rcpcv * TRo);
    double sav1n = sav1 / sav, sav2n = sav2 / sav, sav3n = sav3 / sav;
    const double ga = 1.4, ga1 = 1.4 - 1.0;
    double VAMLo = ULo * ULo + VLo * VLo + WLo * WLo;
    double VAMRo = URo * URo + VRo * VRo + WRo * WRo;
    double HLo = PLo * ga / (DENLo * ga1) + 0.5 * VAMLo;
    double HRo = PRo * ga / (DENRo * ga1) + 0.5 * VAMRo;
    double ULn = (sav1n * ULo + sav2n * VLo + sav3n * WLo);
    double URn = (sav1n * URo + sav2n * VRo + sav3n * WRo);
    double astar_L = sqrt(2 * ga1 * HLo / (ga + 1));
    double astar_R = sqrt(2 * ga1 * HRo / (ga + 1));
    double aas_L = sqr(astar_L) / fmax(sqrt(VAMLo), astar_L);
    double aas_R = sqr(astar_R) / fmax(sqrt(VAMRo), astar_R);
    double Acc = fmin(aas_L, aas_R);
    double MaL = ULn / Acc, MaR = URn / Acc;
    const double c316 = 0.0; /* 3/16 in integer arithmetic */
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
    double pS = pL_plus * PLo + pR_neg * PRo;
    double Mas = MaL_plus + MaR_neg;
    double mn = fmin(PLo / PRo, PRo / PLo);
    double wP = 1.0 - (mn * mn) * mn;
    double fL = fabs(MaL) < 1 ? PLo / (pS + tiny) - 1 : 0;
    double fR = fabs(MaR) < 1 ? PRo / (pS + tiny) - 1 : 0;
    double MMaL_plus, MMaR_neg;
    if (Mas >= 0) {
        MMaL_plus = MaL_plus + MaR_neg * ((1 - wP) * (1 + fR) - fL);
        MMaR_neg = MaR_neg * wP * (1 + fR);
    } else {
        MMaL_plus = MaL_plus * wP * (1 + fL);
        MMaR_neg = MaR_neg + MaL_plus * ((1 - wP) * (1 + fL) - fR);
    }
    out[0] = (-1) * MMaL_plus * Acc * DENLo - MMaR_neg * Acc * DENRo;
    out[1] = (-1) * MMaL_plus * Acc * DENLo * ULo - MMaR_neg * Acc * DENRo * URo - sav1n * (pL_plus * PLo + pR_neg * PRo);
    out[2] = (-1) * MMaL_plus * Acc * DENLo * VLo - MMaR_neg * Acc * DENRo * VRo - sav2n * (pL_plus * PLo + pR_neg * PRo);
    out[3] = (-1) * MMaL_plus * Acc * DENLo * WLo - MMaR_neg * Acc * DENRo * WRo - sav3n * (pL_plus * PLo + pR_neg * PRo);
    out[4] = (-1) * MMaL_plus * Acc * DENLo * HLo - MMaR_neg * Acc * DENRo * HRo;
}

/* Space/CONVECT_AUSMPW_Plus.F90:1-201 */
void ora_convect_ausmpw(const trans_params* p, ora_block* b)
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
                    ora_ausmpw_common(p, VL(d, 4, i, j, k), VR(d, 4, i, j, k), VL(d, 5, i, j, k),
                                      VR(d, 5, i, j, k), VL(d, 1, i, j, k), VR(d, 1, i, j, k),
                                      VL(d, 2, i, j, k), VR(d, 2, i, j, k), VL(d, 3, i, j, k),
                                      VR(d, 3, i, j, k), SD(d, 1, i, j, k), SD(d, 2, i, j, k),
                                      SD(d, 3, i, j, k), sav, f);
                    for (int L = 1; L <= 5; L++) {
                        D(L, i, j, k) = D(L, i, j, k) - f[L - 1] * sav;
                        D(L, i - di, j - dj, k - dk) = D(L, i - di, j - dj, k - dk) + f[L - 1] * sav;
                    }
                }
    }
}

/* Space/TURBTWO_CONVECT_Upwind.F90:2-125 */
void ora_turb_convect_upwind(const trans_params* p, ora_block* b)
{
    const int NI = b->ni, NJ = b->nj, NK = b->nk;
    const int NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    for (int d = 1; d <= 3; d++) {
        const int di = d == 1, dj = d == 2, dk = d == 3;
        const int ihi = di ? NI : NI1, jhi = dj ? NJ : NJ1, khi = dk ? NK : NK1;
        for (int k = 1; k <= khi; k++)
            for (int j = 1; j <= jhi; j++)
                for (int i = 1; i <= ihi; i++) {
                    double UL = VL(d, 1, i, j, k), UR = VR(d, 1, i, j, k);
                    double VLl = VL(d, 2, i, j, k), VRr = VR(d, 2, i, j, k);
                    double WL = VL(d, 3, i, j, k), WR = VR(d, 3, i, j, k);
                    double PL = VL(d, 4, i, j, k), PR = VR(d, 4, i, j, k);
                    double TL = VL(d, 5, i, j, k), TR = VR(d, 5, i, j, k);
                    double rcpcv = p->ppf;
                    double DENL = PL / (rcpcv * TL), DENR = PR / (rcpcv * TR);
                    double sav1 = SD(d, 1, i, j, k), sav2 = SD(d, 2, i, j, k), sav3 = SD(d, 3, i, j, k);
                    double ulnormal = (sav1 * UL + sav2 * VLl + sav3 * WL);
                    double urnormal = (sav1 * UR + sav2 * VRr + sav3 * WR);
                    double rulnormal = DENL * ulnormal, rurnormal = DENR * urnormal;
                    const double small = 1.e-10;
                    double eigTKO = 0.5 * fmax(fabs(ulnormal), fabs(urnormal));
                    double TKL = fmax(VL(d, 6, i, j, k), small), TKR = fmax(VR(d, 6, i, j, k), small);
                    double dTK = -0.5 * (rulnormal * TKL + rurnormal * TKR - eigTKO * (DENR * TKR - DENL * TKL));
                    double TOL = fmax(VL(d, 7, i, j, k), small), TOR = fmax(VR(d, 7, i, j, k), small);
                    double dTO = -0.5 * (rulnormal * TOL + rurnormal * TOR - eigTKO * (DENR * TOR - DENL * TOL));
                    D(6, i, j, k) = D(6, i, j, k) - dTK;
                    D(6, i - di, j - dj, k - dk) = D(6, i - di, j - dj, k - dk) + dTK;
                    D(7, i, j, k) = D(7, i, j, k) - dTO;
                    D(7, i - di, j - dj, k - dk) = D(7, i - di, j - dj, k - dk) + dTO;
                }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Time/Hybrid_function.F90:1-39 (FSST): DES/DDES shielding function */
double ora_fsst(const trans_params* p, double Den, double vK, double vOmega, double Dist, double RmiuL,
                double CDiff)
{
    const double Ref = p->ref;
    double Fsst = 0.;
    if (p->kind_subhy == 0) {
        Fsst = 0.;
    } else if (p->kind_subhy == 1 || p->kind_subhy == 3) {
        double arg1 = sqrt(vK) / (0.09 * vOmega * Dist * Ref);
        double arg2 = 500. * RmiuL / (Den * Dist * Dist * vOmega * Ref * Ref);
        double arga = fmax(arg1, arg2);
        double temp = fmax(Den * CDiff * Ref, 1.e-10);
        double argb = 4. * Den * 0.856 * vK / (temp * Dist * Dist);
        double arg = fmin(arga, argb);
        Fsst = tanh(arg * arg * arg * arg);
    } else if (p->kind_subhy == 2) {
        double arg1 = 2. * sqrt(vK) / (0.09 * vOmega * Dist * Ref);
        double arg2 = 500. * RmiuL / (Den * Dist * Dist * vOmega * Ref * Ref);
        double arg = fmax(arg1, arg2);
        Fsst = tanh(arg * arg);
    }
    return Fsst;
}

/* Time/Hybrid_function.F90:42-68 (fd_function), kappa = 0.41 (Module_Global.F90:126) */
double ora_iddes_fd(double Den1, double RmiuT1, double aDst1, double aLmax1, const double g[9], double Ref)
{
    double alpha1 = 0.25 - aDst1 / aLmax1;
    double fb = fmin(2. * exp(-9.0 * alpha1 * alpha1), 1.);
    double Sij_square = g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3] + g[4] * g[4] + g[5] * g[5] +
                        g[6] * g[6] + g[7] * g[7] + g[8] * g[8];
    Sij_square = sqrt(Sij_square);
    Sij_square = fmax(Sij_square, 1.e-10);
    const double akapa = 0.41;
    double rdt = (RmiuT1 / Den1) / (akapa * akapa * (aDst1 * aDst1) * Sij_square * Ref);
    double fdt = 1. - tanh((8. * rdt) * (8. * rdt) * (8. * rdt));
    return fmax((1. - fdt), fb);
}

/* Time/Hybrid_function.F90:72-105 (fe_function), Ct = 1.87, Cl = 5 */
double ora_iddes_fe(double Den2, double RmiuT2, double RMIUL2, double aDst2, double aLmax2, const double g[9],
                    double Ref)
{
    double alpha2 = 0.25 - aDst2 / aLmax2;
    double fe1;
    if (alpha2 >= 0.)
        fe1 = 2. * exp(-11.09 * alpha2 * alpha2);
    else
        fe1 = 2. * exp(-9. * alpha2 * alpha2);
    double Sij_square = g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3] + g[4] * g[4] + g[5] * g[5] +
                        g[6] * g[6] + g[7] * g[7] + g[8] * g[8];
    Sij_square = sqrt(Sij_square);
    Sij_square = fmax(Sij_square, 1.e-10);
    const double akapa = 0.41, Ct = 1.87, Cl = 5.00;
    double rdt = (RmiuT2 / Den2) / (akapa * akapa * (aDst2 * aDst2) * Sij_square * Ref);
    double ft = tanh((Ct * Ct * rdt) * (Ct * Ct * rdt) * (Ct * Ct * rdt));
    double rdl = (RMIUL2 / Den2) / (akapa * akapa * (aDst2 * aDst2) * Sij_square * Ref);
    double fl = tanh(pow(Cl * Cl * rdl, 10.));
    double fe2 = 1. - fmax(ft, fl);
    return fe2 * fmax((fe1 - 1.), 0.);
}

/* Time/TURBTWO_TIME_LUSGS_SST.F90:2-415 (and :418-547 for WD+): right-hand side, sources
 * (SST with the DES/DDES multiplier of Kind_Hybrid = 3, or WD+; f_r1 = 1) and the diagonal
 * F(1:2) of the k-omega LU-SGS. */
void ora_sst_sources(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    const double dtm = p->dt_lusgs_main, Ref = p->ref;
    /* Vomega and arg1 of the previous cell: the DES block (:192-214) reads both before this
     * cell assigns them.  They are unset for the first cell of a call (a local of the
     * reference routine); the first cell takes its own. */
    double Vomega = 0., arg1_prev = 0.;
    int have_prev = 0;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++)
                for (int L = 6; L <= 7; L++)
                    D(L, I, J, K) = D(L, I, J, K) / VOL(I, J, K) -
                                    (3.0 * V(L, I, J, K) - 4.0 * WT0(L, I, J, K) + WT1(L, I, J, K)) * 0.5 / dtm;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double dudx = DQ(1, I, J, K), dudy = DQ(2, I, J, K), dudz = DQ(3, I, J, K);
                double dvdx = DQ(4, I, J, K), dvdy = DQ(5, I, J, K), dvdz = DQ(6, I, J, K);
                double dwdx = DQ(7, I, J, K), dwdy = DQ(8, I, J, K), dwdz = DQ(9, I, J, K);
                double dtkdx = DQ(13, I, J, K), dtkdy = DQ(14, I, J, K), dtkdz = DQ(15, I, J, K);
                double dtodx = DQ(16, I, J, K), dtody = DQ(17, I, J, K), dtodz = DQ(18, I, J, K);
                double Div = dudx + dvdy + dwdz;
                double Votg = sqr(dudy - dvdx) + sqr(dvdz - dwdy) + sqr(dwdx - dudz);
                double prod = (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sqr(dudy + dvdx) +
                               sqr(dvdz + dwdy) + sqr(dwdx + dudz) - 2.0 / 3.0 * Div * Div);
                double cdif = dtkdx * dtodx + dtkdy * dtody + dtkdz * dtodz;
                if (p->kind_model == 2) { /* WD+ (Time/TURBTWO_TIME_LUSGS_SST.F90:418-547) */
                    double W[3][3], S[3][3];
                    W[0][0] = 0.; W[1][1] = 0.; W[2][2] = 0.;
                    W[0][1] = 0.5 * (dudy - dvdx);
                    W[0][2] = 0.5 * (dudz - dwdx);
                    W[1][2] = 0.5 * (dvdz - dwdy);
                    W[1][0] = -W[0][1];
                    W[2][0] = -W[0][2];
                    W[2][1] = -W[1][2];
                    S[0][0] = 0.5 * (dudx + dudx);
                    S[1][1] = 0.5 * (dvdy + dvdy);
                    S[2][2] = 0.5 * (dwdz + dwdz);
                    S[0][1] = 0.5 * (dudy + dvdx);
                    S[0][2] = 0.5 * (dudz + dwdx);
                    S[1][2] = 0.5 * (dvdz + dwdy);
                    S[1][0] = S[0][1];
                    S[2][0] = S[0][2];
                    S[2][1] = S[1][2];
                    double tprod = fmin(prod, Votg) * RMIU(I, J, K) / Ref;
                    double ARo = V(1, I, J, K), ARTk = V(6, I, J, K), ARTO = V(7, I, J, K);
                    double tprodl1 = p->prod_lmt * p->beta_star * ARTk * ARTO / ARo * Ref;
                    tprod = fmin(tprod, tprodl1);
                    double SourceP = tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
                    double SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) - p->beta_star * ARTO / ARo * Ref;
                    double Source = SourceP + SourceN1 * ARTk;
                    double CnegTk = fabs(SourceN1);
                    D(6, I, J, K) = D(6, I, J, K) + Source;
                    tprod = fmin(prod, Votg) / Ref;
                    double WWSg = 0.;
                    for (int II = 0; II < 3; II++)
                        for (int JJ = 0; JJ < 3; JJ++)
                            for (int KK = 0; KK < 3; KK++) WWSg = WWSg + W[II][JJ] * W[JJ][KK] * S[KK][II];
                    double aK_omg = fabs(WWSg / pow(0.09 * ARTO / ARo * Ref, 3.));
                    double FBeta = (1. + 85. * aK_omg) / (1. + 100. * aK_omg);
                    double Beta_06 = p->beta1 * FBeta;
                    double dltCD = cdif <= 0. ? 0. : 0.125;
                    SourceP = p->gama1 * ARo * tprod - 2.0 / 3.0 * p->gama1 * fmin(Div, 0.0) * ARTO;
                    SourceN1 = -2.0 / 3.0 * p->gama1 * fmax(Div, 0.0) - Beta_06 * ARTO / ARo * Ref;
                    Source = SourceP + SourceN1 * ARTO;
                    double SourceCross = dltCD * ARo * ARo / ARTO / Ref * cdif;
                    Source = Source + SourceCross;
                    double CnegTO = 2.0 * fabs(SourceN1);
                    D(7, I, J, K) = D(7, I, J, K) + Source;
                    double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                    F(1, I, J, K) = 1. + (RDS(1, I, J, K) + RDS(2, I, J, K) + RDS(3, I, J, K)) * (stept / VOL(I, J, K)) +
                                    (CnegTk + CnegTO) * stept;
                    F(2, I, J, K) = F(1, I, J, K);
                    continue;
                }
                double tprod = fmin(prod, Votg) * RMIU(I, J, K) / Ref;
                double ARo = V(1, I, J, K), ARTk = V(6, I, J, K), ARTO = V(7, I, J, K);
                double tprodl1 = p->prod_lmt * p->beta_star * ARTk * ARTO / ARo * Ref;
                tprod = fmin(tprod, tprodl1);
                double TT = T(I, J, K);
                double RmiuL = sutherland(p, TT);
                double DistN = DST(I, J, K);
                double CDKO = fmax(2.0 * ARo * p->sigo2 / (ARTO / ARo) * cdif, 1.0e-10);
                double arg1 = fmax(sqrt(ARTk / ARo) / p->beta_star / (ARTO / ARo) / DistN / Ref,
                                   500.0 * RmiuL / DistN / DistN / ARTO / Ref / Ref);
                arg1 = fmin(arg1, 4.0 * p->sigo2 * ARTk / CDKO / DistN / DistN);
                arg1 = tanh(pow4(arg1));
                double Fdes = 1.0, alIDDES = 0.0;
                if (p->kind_hybrid == 3) { /* :192-247 */
                    if (!have_prev) {
                        Vomega = ARTO / ARo;
                        arg1_prev = arg1;
                    }
                    double aLt = sqrt(ARTk / ARo) / (p->beta_star * ARTO / ARo * Ref);
                    double Cdiff_DDES_S = (2. * 0.856 / Vomega) / Ref * cdif;
                    double Den = ARo, vk = ARTk / ARo;
                    Vomega = ARTO / ARo;
                    double Fsst_DES = ora_fsst(p, Den, vk, Vomega, DistN, RmiuL, Cdiff_DDES_S);
                    double aLgrd = ((1. - arg1_prev) * 0.61 + arg1_prev * 0.78) * ALENG(I, J, K);
                    Fdes = fmax(aLt * (1. - Fsst_DES) / aLgrd, 1.);
                    if (p->kind_subhy == 5) { /* IDDES, Cw = 0.15 */
                        double aLmax = ALENG(I, J, K), aDst = DST(I, J, K), aLmin = ASENG(I, J, K);
                        double Cdes = (1. - arg1_prev) * 0.61 + arg1_prev * 0.78;
                        double dlt_IDDES = fmin(fmax(fmax(0.15 * aLmax, 0.15 * aDst), aLmin), aLmax);
                        aLgrd = Cdes * dlt_IDDES;
                        const double g9[9] = {dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz};
                        double fd = ora_iddes_fd(Den, RMIU(I, J, K), aDst, aLmax, g9, Ref);
                        double fe = ora_iddes_fe(Den, RMIU(I, J, K), RmiuL, aDst, aLmax, g9, Ref);
                        alIDDES = (1. + fe) * fd * aLt + (1. - fd) * aLgrd;
                    }
                }
                arg1_prev = arg1; /* arg1 of this cell, read by the next one */
                Vomega = ARTO / ARo;
                have_prev = 1;
                double AA = sqrt(1.4 * PP(I, J, K) / ARo); /* :271-273 */
                double Mt = sqrt(ARTk / ARo) / AA;
                const double Alf1 = 1.0, Alf2 = 0.4, Alf3 = 0.2; /* :75-76 */
                double SourceP, SourceN1, Source, CnegTk, GamaAB, BetaAB;
                if (p->if_comppress == 0) { /* :282-330 */
                    SourceP = tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
                    SourceP = SourceP * b->f_r1[IX(I, J, K)];
                    SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) - p->beta_star * ARTO / ARo * Ref * Fdes;
                    Source = SourceP + SourceN1 * ARTk;
                    CnegTk = fabs(SourceN1);
                    if (p->kind_hybrid == 3 && p->kind_subhy == 5) { /* :299-311 */
                        SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) - sqrt(ARTk / ARo) / alIDDES;
                        Source = SourceP + SourceN1 * ARTk;
                        CnegTk = fabs(-2.0 / 3.0 * fmax(Div, 0.0) - 1.5 * sqrt(ARTk / ARo) / alIDDES);
                    }
                    D(6, I, J, K) = D(6, I, J, K) + Source;
                    tprod = fmin(prod, Votg) / Ref;
                    GamaAB = p->gama1 * arg1 + p->gama2 * (1.0 - arg1);
                    BetaAB = p->beta1 * arg1 + p->beta2 * (1.0 - arg1);
                    SourceP = GamaAB * ARo * tprod - 2.0 / 3.0 * GamaAB * fmin(Div, 0.0) * ARTO;
                    SourceP = SourceP * b->f_r1[IX(I, J, K)];
                    SourceN1 = -2.0 / 3.0 * GamaAB * fmax(Div, 0.0) - BetaAB * ARTO / ARo * Ref;
                } else { /* Suzen (:331-396) */
                    SourceP = (1. - (1. - arg1) * Alf2 * Mt * Mt) * tprod - 2.0 / 3.0 * fmin(Div, 0.0) * ARTk;
                    SourceP = SourceP * b->f_r1[IX(I, J, K)];
                    SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) -
                               (1. + (1. - arg1) * (Alf1 - Alf3) * Mt * Mt) * p->beta_star * ARTO / ARo * Ref * Fdes;
                    Source = SourceP + SourceN1 * ARTk;
                    CnegTk = fabs(SourceN1);
                    if (p->kind_hybrid == 3 && p->kind_subhy == 5) { /* :354-366 */
                        SourceN1 = -2.0 / 3.0 * fmax(Div, 0.0) -
                                   (1. + (1. - arg1) * (Alf1 - Alf3) * Mt * Mt) * sqrt(ARTk / ARo) / alIDDES;
                        Source = SourceP + SourceN1 * ARTk;
                        CnegTk = fabs(-2.0 / 3.0 * fmax(Div, 0.0) - 1.5 * sqrt(ARTk / ARo) / alIDDES);
                    }
                    D(6, I, J, K) = D(6, I, J, K) + Source;
                    tprod = fmin(prod, Votg) / Ref;
                    GamaAB = p->gama1 * arg1 + p->gama2 * (1.0 - arg1);
                    BetaAB = p->beta1 * arg1 + p->beta2 * (1.0 - arg1);
                    SourceP = (GamaAB + (1. - arg1) * Alf2 * Mt * Mt) * ARo * tprod -
                              2.0 / 3.0 * GamaAB * fmin(Div, 0.0) * ARTO;
                    SourceP = SourceP * b->f_r1[IX(I, J, K)];
                    SourceN1 = -2.0 / 3.0 * GamaAB * fmax(Div, 0.0) -
                               (1. - (1. - arg1) * (Alf1 - Alf3) * Mt * Mt) * BetaAB * ARTO / ARo * Ref;
                }
                Source = SourceP + SourceN1 * ARTO;
                double SourceCross = 2.0 * (1 - arg1) * ARo * p->sigo2 / (ARTO / ARo) * cdif / Ref;
                SourceCross = copysign(1.0, SourceCross) * fmin(fabs(SourceCross), fabs(SourceP) * 0.5);
                Source = Source + SourceCross;
                D(7, I, J, K) = D(7, I, J, K) + Source;
                double CnegTO = 2.0 * fabs(SourceN1);
                double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                F(1, I, J, K) = 1. + (RDS(1, I, J, K) + RDS(2, I, J, K) + RDS(3, I, J, K)) * (stept / VOL(I, J, K)) +
                                (CnegTk + CnegTO) * stept;
                F(2, I, J, K) = F(1, I, J, K);
            }
}

/* Time/TURBTWO_TIME_LUSGS_SST.F90:2-775: sources, then the sweeps (:550-766), whose flux
 * Jacobian reduces to the normal velocity. */
void ora_turb_lusgs_sst(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    const double dtm = p->dt_lusgs_main;
    ora_sst_sources(p, b);
    for (int pass = 0; pass < 2; pass++) {
        const int fwd = pass == 0;
        for (int K0 = 1; K0 <= NK1; K0++)
            for (int J0 = 1; J0 <= NJ1; J0++)
                for (int I0 = 1; I0 <= NI1; I0++) {
                    int I = fwd ? I0 : NI1 + 1 - I0, J = fwd ? J0 : NJ1 + 1 - J0, K = fwd ? K0 : NK1 + 1 - K0;
                    double stept = 2.0 * DTM(I, J, K) / (3.0 * DTM(I, J, K) / dtm + 2.0);
                    double cvs = stept / VOL(I, J, K);
                    double tk[3] = {0, 0, 0}, to[3] = {0, 0, 0};
                    for (int d = 1; d <= 3; d++) {
                        const int di = d == 1, dj = d == 2, dk = d == 3;
                        int at_edge = fwd ? ((d == 1 && I == 1) || (d == 2 && J == 1) || (d == 3 && K == 1))
                                          : ((d == 1 && I == NI1) || (d == 2 && J == NJ1) || (d == 3 && K == NK1));
                        if (at_edge) continue;
                        int sgn = fwd ? -1 : 1;
                        int in = I + sgn * di, jn = J + sgn * dj, kn = K + sgn * dk;
                        double pRo = V(1, in, jn, kn);
                        double roinv = 1.0 / pRo;
                        double pVx = V(2, in, jn, kn) * roinv, pVy = V(3, in, jn, kn) * roinv, pVz = V(4, in, jn, kn) * roinv;
                        double aSDx = 0.5 * (SD(d, 1, in + di, jn + dj, kn + dk) + SD(d, 1, in, jn, kn));
                        double aSDy = 0.5 * (SD(d, 2, in + di, jn + dj, kn + dk) + SD(d, 2, in, jn, kn));
                        double aSDz = 0.5 * (SD(d, 3, in + di, jn + dj, kn + dk) + SD(d, 3, in, jn, kn));
                        double VP = pVx * aSDx + pVy * aSDy + pVz * aSDz;
                        double DifRTk = D(6, in, jn, kn) * VP, DifRTO = D(7, in, jn, kn) * VP;
                        double RDSt = RDS(d, in, jn, kn);
                        if (fwd) {
                            tk[d - 1] = 0.5 * (DifRTk + RDSt * D(6, in, jn, kn));
                            to[d - 1] = 0.5 * (DifRTO + RDSt * D(7, in, jn, kn));
                        } else {
                            tk[d - 1] = 0.5 * (DifRTk - RDSt * D(6, in, jn, kn));
                            to[d - 1] = 0.5 * (DifRTO - RDSt * D(7, in, jn, kn));
                        }
                    }
                    if (fwd) {
                        D(6, I, J, K) = (D(6, I, J, K) * stept + (tk[0] + tk[1] + tk[2]) * cvs) / F(1, I, J, K);
                        D(7, I, J, K) = (D(7, I, J, K) * stept + (to[0] + to[1] + to[2]) * cvs) / F(2, I, J, K);
                    } else {
                        D(6, I, J, K) = D(6, I, J, K) - (tk[0] + tk[1] + tk[2]) * cvs / F(1, I, J, K);
                        D(7, I, J, K) = D(7, I, J, K) - (to[0] + to[1] + to[2]) * cvs / F(2, I, J, K);
                    }
                }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Output/Average_Field.F90:46-263 (Average_Fields at KI, one block).  S(23, cells) in the
 * FLfld record order: U V W R M Mu k P0 P T uu vv ww uv uw vw, uurms vvrms wwrms uvrms vwrms
 * uwrms pprms; component index fastest, like the other oracle arrays. */
void ora_average_fields(const trans_params* p, int ki, int ki_aver, int ki_statis, ora_block* b, double* S)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    const double Ref = p->ref;
#define STAT(q, i, j, k) S[(q) + 23 * IX(i, j, k)]
    for (int J = 1; J <= NJ1; J++)
        for (int K = 1; K <= NK1; K++)
            for (int I = 1; I <= NI1; I++) {
                double Den = V(1, I, J, K), RmiuT = RMIU(I, J, K), vK = V(6, I, J, K) / Den;
                double vab[16];
                vab[0] = V(2, I, J, K) / Den;
                vab[1] = V(3, I, J, K) / Den;
                vab[2] = V(4, I, J, K) / Den;
                vab[3] = Den;
                double QQ = sqrt(vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2]);
                double AA = sqrt(1.4 * PP(I, J, K) / Den);
                vab[4] = QQ / AA;
                vab[5] = RMIU(I, J, K);
                vab[6] = V(6, I, J, K) / Den;
                vab[7] = PP(I, J, K) * pow(1.0 + 0.2 * vab[4] * vab[4], 3.5) / p->ppf;
                vab[8] = PP(I, J, K);
                vab[9] = T(I, J, K);
                double dudx = DQ(1, I, J, K), dudy = DQ(2, I, J, K), dudz = DQ(3, I, J, K);
                double dvdx = DQ(4, I, J, K), dvdy = DQ(5, I, J, K), dvdz = DQ(6, I, J, K);
                double dwdx = DQ(7, I, J, K), dwdy = DQ(8, I, J, K), dwdz = DQ(9, I, J, K);
                double S0 = -2.0 / 3.0 * (dudx + dvdy + dwdz);
                double upup = RmiuT / Ref * (S0 + 2.0 * dudx) - 2.0 / 3.0 * Den * vK;
                double vpvp = RmiuT / Ref * (S0 + 2.0 * dvdy) - 2.0 / 3.0 * Den * vK;
                double wpwp = RmiuT / Ref * (S0 + 2.0 * dwdz) - 2.0 / 3.0 * Den * vK;
                double upvp = RmiuT / Ref * (dudy + dvdx);
                double upwp = RmiuT / Ref * (dudz + dwdx);
                double vpwp = RmiuT / Ref * (dvdz + dwdy);
                vab[10] = -upup;
                vab[11] = -vpvp;
                vab[12] = -wpwp;
                vab[13] = -upvp;
                vab[14] = -upwp;
                vab[15] = -vpwp;
                if (ki == ki_aver) {
                    for (int q = 0; q < 16; q++) STAT(q, I, J, K) = vab[q];
                } else {
                    double Time = (double)(ki - ki_aver);
                    for (int q = 0; q < 16; q++) STAT(q, I, J, K) = (STAT(q, I, J, K) * Time + vab[q]) / (Time + 1.);
                }
            }
    if (ki >= ki_statis) {
        double Time = (double)(ki - ki_statis);
        for (int K = 1; K <= NK1; K++)
            for (int J = 1; J <= NJ1; J++)
                for (int I = 1; I <= NI1; I++) {
                    double Den = V(1, I, J, K);
                    double v1 = V(2, I, J, K) / Den, v2 = V(3, I, J, K) / Den, v3 = V(4, I, J, K) / Den;
                    double v5 = PP(I, J, K);
                    double Ub = STAT(0, I, J, K), Vb = STAT(1, I, J, K), Wb = STAT(2, I, J, K), Pb = STAT(8, I, J, K);
                    STAT(16, I, J, K) = (STAT(16, I, J, K) * Time + (v1 - Ub) * (v1 - Ub)) / (Time + 1.);
                    STAT(17, I, J, K) = (STAT(17, I, J, K) * Time + (v2 - Vb) * (v2 - Vb)) / (Time + 1.);
                    STAT(18, I, J, K) = (STAT(18, I, J, K) * Time + (v3 - Wb) * (v3 - Wb)) / (Time + 1.);
                    STAT(19, I, J, K) = (STAT(19, I, J, K) * Time + (v1 - Ub) * (v2 - Vb)) / (Time + 1.);
                    STAT(21, I, J, K) = (STAT(21, I, J, K) * Time + (v1 - Ub) * (v3 - Wb)) / (Time + 1.);
                    STAT(20, I, J, K) = (STAT(20, I, J, K) * Time + (v2 - Vb) * (v3 - Wb)) / (Time + 1.);
                    STAT(22, I, J, K) = (STAT(22, I, J, K) * Time + (v5 - Pb) * (v5 - Pb)) / (Time + 1.);
                }
    }
#undef STAT
}

/* Input/Init.F90:146-160: p and T from V after a restart read */
void ora_pp_t_from_v(const trans_params* p, ora_block* b)
{
    const int NI1 = b->ni - 1, NJ1 = b->nj - 1, NK1 = b->nk - 1;
    for (int K = 1; K <= NK1; K++)
        for (int J = 1; J <= NJ1; J++)
            for (int I = 1; I <= NI1; I++) {
                double VV = V(2, I, J, K) * V(2, I, J, K) + V(3, I, J, K) * V(3, I, J, K) + V(4, I, J, K) * V(4, I, J, K);
                double wr = p->omega1 * b->rad[IX(I, J, K)];
                PP(I, J, K) = (1.4 - 1.0) * V(5, I, J, K) - 0.5 * (1.4 - 1.0) * (VV / V(1, I, J, K) + -(wr * wr) * V(1, I, J, K));
                T(I, J, K) = p->rxm2 * PP(I, J, K) / V(1, I, J, K);
            }
}
