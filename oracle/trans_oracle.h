/*
 * CPU oracle for the TRANS residual-and-update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the reference's
 * Fortran routines, kept in the reference's own array layout (component index first,
 * column-major) and loop order so that it can be read side by side with the reference.
 * It is the checker for the HIP library and the CPU baseline of bench.py; nothing in
 * the product (trans_amd/, include/) calls it.
 *
 * Layout: every array of a block spans -2:N+2 in each direction (N = points), i.e.
 * N+5 entries, with i fastest after the component index.  Component arrays put the
 * component first: V(L,i,j,k) is V[(L-1) + 7*IX(i,j,k)].
 *
 * Every routine is pinned against closed forms, invariants or brute force
 * (tests/test_oracle_pins.py; the rotating frame and all 48 orientations of the
 * block-interface copy in tests/test_oracle_rotation_pins.py).  The routines whose reference
 * source compiles unmodified (preconditioning, limiters, hybrid RANS/LES functions, the
 * convective reconstructions, the AUSMPW+ and k-omega upwind fluxes, the dual-time shift, the
 * RK save) are also compared bit for bit with that build of the reference (oracle/_ref,
 * tests/test_oracle_ref.py).
 */
#ifndef TRANS_ORACLE_H
#define TRANS_ORACLE_H

#include "../include/trans_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_block {
    int ni, nj, nk; /* points */
    int npatch;
    const trans_patch* patches;
    /* state */
    double *V, *PP, *T, *Rmiu;         /* V: 7 comps */
    /* geometry */
    double *SD, *Grad, *Alagm;         /* 9, 3, 3 comps */
    double *Vol, *rad, *Dst, *aLeng, *aSeng, *Yc, *Zc, *thtc; /* scalars */
    const int* mw[6];                  /* MARKWALLi0, im, j0, jm, k0, km */
    /* work */
    double *VL, *VR;                   /* (3,8): comp (d-1)+3*(L-1) */
    double *rccl, *rccr;               /* 3 */
    double *D, *F;                     /* 7 */
    double *Rds;                       /* 3 */
    double *Dtm;
    double *Wt0, *Wt1, *Wt2;           /* 7 */
    double *DQdxyz;                    /* 18 */
    double *shock, *Fscheme, *f_r1;
    trans_resid res;
} ora_block;

/* IF_turb as main.F90:59-62 derives it for iteration ki. */
int ora_if_turb(const trans_params* p, int ki);
double ora_fsst(const trans_params* p, double Den, double vK, double vOmega, double Dist, double RmiuL,
                double CDiff);
void ora_sst_sources(const trans_params* p, ora_block* b);
void ora_average_fields(const trans_params* p, int ki, int ki_aver, int ki_statis, ora_block* b, double* S);
void ora_pp_t_from_v(const trans_params* p, ora_block* b);
double ora_iddes_fd(double Den1, double RmiuT1, double aDst1, double aLmax1, const double g[9], double Ref);
double ora_iddes_fe(double Den2, double RmiuT2, double RMIUL2, double aDst2, double aLmax2, const double g[9],
                    double Ref);

void ora_spectral_radius(const trans_params* p, int ki, ora_block* b);
void ora_bc(const trans_params* p, int if_turb, ora_block* blocks, int nblk, int sweep);
void ora_spaceinterp(const trans_params* p, int if_turb, ora_block* b);
void ora_convect_muscl(const trans_params* p, ora_block* b);
void ora_convect_muscl2(const trans_params* p, ora_block* b);
double ora_muscl2(double var1, double var2);
void ora_convect_roe(const trans_params* p, ora_block* b);
void ora_convect_roe_rotated(const trans_params* p, ora_block* b);
void ora_roe_rotated_face(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                          double URo, double VLo, double VRo, double WLo, double WRo, double s1, double s2,
                          double s3, double grad, double Funscheme, double rLo, double rRo, double out[5]);
void ora_roe_rotated_face_pc(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                             double URo, double VLo, double VRo, double WLo, double WRo, double s1, double s2,
                             double s3, double grad, double Funscheme, double rLo, double rRo, double RmiuT,
                             double AlagmR, double out[5]);
void ora_lusgs(const trans_params* p, ora_block* b);
void ora_update_v(const trans_params* p, int ki, int if_turb, ora_block* b);
void ora_update_u(const trans_params* p, ora_block* b);

void ora_diff_dqdxyz(const trans_params* p, int if_turb, ora_block* b);
void ora_diffusion(const trans_params* p, ora_block* b);
void ora_turb_diffusion(const trans_params* p, ora_block* b);
void ora_convect_mdcd(const trans_params* p, ora_block* b);
void ora_convect_mdcdhyb(const trans_params* p, ora_block* b);
void ora_turb_convect_muscl(const trans_params* p, ora_block* b);
void ora_convect_ausmpw(const trans_params* p, ora_block* b);
void ora_turb_convect_upwind(const trans_params* p, ora_block* b);
void ora_turb_lusgs_sst(const trans_params* p, ora_block* b);
double ora_mdcd_rec(double gama, const double s[6], int right);
void ora_ausmpw_common(const trans_params* p, double PLo, double PRo, double TLo, double TRo,
                       double ULo, double URo, double VLo, double VRo, double WLo, double WRo,
                       double sav1, double sav2, double sav3, double sav, double out[5]);

/* Composite steps over all blocks, in main.F90 order. */
void ora_residual(const trans_params* p, int ki, ora_block* blocks, int nblk);
void ora_lusgs_update(const trans_params* p, int ki, ora_block* blocks, int nblk);
/* Third-order Runge-Kutta stage itr_chld (1..3) with the k-omega LU-SGS (IF_RK = 1). */
void ora_rk_save(int itr_chld, ora_block* b);
void ora_rk(int itr_chld, ora_block* b);
void ora_rk_update_v(const trans_params* p, int ki, int if_turb, ora_block* b);
void ora_rk_update(const trans_params* p, int ki, int itr_chld, ora_block* blocks, int nblk);
/* Threads an OpenMP region of the oracle uses (blocks run in parallel). */
int ora_num_threads(void);
void ora_set_threads(int n);
/* n outer iterations starting at *ki; *ki is advanced. */
void ora_iterate(const trans_params* p, int* ki, int n_iter, ora_block* blocks, int nblk);

/* Plane integrals of include/trans_hip.h: out = {mass_flow, area, p0_mass_avg, t0_mass_avg}. */
void ora_plane_integrals(ora_block* b, int dir, int index, double omega, double out[6]);

/* Reference functions exposed for unit tests. */
double ora_alimiter(double var1, double var2);
void ora_roe_common(const trans_params* p, double PLo, double PRo, double TLo, double TRo,
                    double ULo, double URo, double VLo, double VRo, double WLo, double WRo,
                    double sav1, double sav2, double sav3, double sav, double rLo, double rRo,
                    double out[5]);
void ora_roe_common_pc(const trans_params* p, double PLo, double PRo, double TLo, double TRo, double ULo,
                       double URo, double VLo, double VRo, double WLo, double WRo, double sav1, double sav2,
                       double sav3, double sav, double rLo, double rRo, double RmiuT, double AlagmR, double out[5]);
double ora_precond_para(const trans_params* p, double VM2, double AM2, double RR, double TT, double RMIUT,
                        double ALAGMR);
void ora_precond_dw(const trans_params* p, double a[5], double pRo, double pVx, double pVy, double pVz, double pE,
                    double AAc, double ubetac2, double pPP, double pTT);
void ora_lusgs_adq(double pVx, double pVy, double pVz, double aSDx, double aSDy,
                   double aSDz, double pPP, double roinv, const double ain[5], double dif[5]);
void ora_ijk_start_end(int i, int j, int k, const int tl[3], const int tt[3], int lcross,
                       int out[3]);

#ifdef __cplusplus
}
#endif
#endif
