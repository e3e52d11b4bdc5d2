/* Array accessors of the oracle (reference layout, bounds -2:N+2) -- test infrastructure. */
#pragma once
/* Index of (i,j,k) in a -2:N+2 array of block b (N = points). */
#define SX ((long)b->ni + 5)
#define SY ((long)b->nj + 5)
#define IX(i, j, k) ((long)((i) + 2) + SX * ((long)((j) + 2) + SY * (long)((k) + 2)))

#define V(L, i, j, k) b->V[(L)-1 + 7 * IX(i, j, k)]
#define D(L, i, j, k) b->D[(L)-1 + 7 * IX(i, j, k)]
#define F(L, i, j, k) b->F[(L)-1 + 7 * IX(i, j, k)]
#define WT0(L, i, j, k) b->Wt0[(L)-1 + 7 * IX(i, j, k)]
#define WT1(L, i, j, k) b->Wt1[(L)-1 + 7 * IX(i, j, k)]
#define WT2(L, i, j, k) b->Wt2[(L)-1 + 7 * IX(i, j, k)]
#define PP(i, j, k) b->PP[IX(i, j, k)]
#define T(i, j, k) b->T[IX(i, j, k)]
#define RMIU(i, j, k) b->Rmiu[IX(i, j, k)]
#define SD(d, x, i, j, k) b->SD[(d)-1 + 3 * ((x)-1) + 9 * IX(i, j, k)]
#define GRAD(d, i, j, k) b->Grad[(d)-1 + 3 * IX(i, j, k)]
#define ALAGM(d, i, j, k) b->Alagm[(d)-1 + 3 * IX(i, j, k)]
#define VOL(i, j, k) b->Vol[IX(i, j, k)]
#define RAD(i, j, k) b->rad[IX(i, j, k)]
#define DST(i, j, k) b->Dst[IX(i, j, k)]
#define YC(i, j, k) b->Yc[IX(i, j, k)]
#define ZC(i, j, k) b->Zc[IX(i, j, k)]
#define THTC(i, j, k) b->thtc[IX(i, j, k)]
#define VL(d, L, i, j, k) b->VL[(d)-1 + 3 * ((L)-1) + 24 * IX(i, j, k)]
#define VR(d, L, i, j, k) b->VR[(d)-1 + 3 * ((L)-1) + 24 * IX(i, j, k)]
#define RCCL(d, i, j, k) b->rccl[(d)-1 + 3 * IX(i, j, k)]
#define RCCR(d, i, j, k) b->rccr[(d)-1 + 3 * IX(i, j, k)]
#define RDS(d, i, j, k) b->Rds[(d)-1 + 3 * IX(i, j, k)]
#define DTM(i, j, k) b->Dtm[IX(i, j, k)]
#define DQ(L, i, j, k) b->DQdxyz[(L)-1 + 18 * IX(i, j, k)]
#define SHOCK(i, j, k) b->shock[IX(i, j, k)]
#define FSCHEME(i, j, k) b->Fscheme[IX(i, j, k)]
#define F_R1(i, j, k) b->f_r1[IX(i, j, k)]
#define ALENG(i, j, k) b->aLeng[IX(i, j, k)]
#define ASENG(i, j, k) b->aSeng[IX(i, j, k)]

