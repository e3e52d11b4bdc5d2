// Three-stage TVD Runge-Kutta update (IF_RK = 1, main.F90:158-170): the stage on the mean
// flow, the k-omega LU-SGS on the stage's state, and the state update with its norms.
#pragma once
#include "dev_common.h"
#include "k_time.h"

// Time/TIME_RK_Save.F90:14-27 (stage 1 keeps the state of all seven components in Wt2) and
// Time/TIME_RK.F90:24-84: the norm of D(1)/rho before the stage, then the stage
//   1: V = V + Dtm*D/Vol
//   2: V = 3/4 Wt2 + 1/4 V + 1/4 Dtm*D/Vol
//   3: V = 1/3 Wt2 + 2/3 V + 2/3 Dtm*D/Vol
// on rho..rho*E.  `fused` (no k-omega this stage) also does the rest of
// Time/TIME_RK_V_Update_New.F90: p and T of the new state (:94-97) and the k and omega
// norms of D(6:7) (:265-316).  Otherwise, with `wout`, the new state and the step sizes the
// k-omega line-march sweep reads go to the hyperplane-major buffer when there is one (k_rk_w
// writes them tile by tile instead).
__global__ void __launch_bounds__(256) k_rk_stage(trans_params p, int stage, int fused, int wout, Fields f,
                                                  const BlkInfo* __restrict__ blks,
                                                  const Seg* __restrict__ seg, NormPart* part)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    const long tot = f.tot;
    double dr[3] = {0.0, 0.0, 0.0};
    int arg = -1;
    if (lin < NI1 * NJ1 * NK1) {
        int i, j, k;
        decode3(lin, NI1, NJ1, i, j, k);
        arg = lin;
        const long c = cidx(b, i, j, k);
        if (stage == 1)
            for (int L = 0; L < TRANS_ML; L++) f.Wt2[(long)L * tot + c] = f.V[(long)L * tot + c];
        dr[0] = fabs(f.D[c] / f.V[c]);
        const double dtm = f.Dtm[c], vol = f.Vol[c];
        double vn[5];
#pragma unroll
        for (int L = 0; L < 5; L++) {
            const long e = (long)L * tot + c;
            const double v = f.V[e], d = f.D[e];
            if (stage == 1)
                vn[L] = v + dtm * d / vol;
            else if (stage == 2)
                vn[L] = 0.75 * f.Wt2[e] + 0.25 * v + 0.25 * dtm * d / vol;
            else
                vn[L] = 1.0 / 3.0 * f.Wt2[e] + 2.0 / 3.0 * v + 2.0 / 3.0 * dtm * d / vol;
            f.V[e] = vn[L];
        }
        if (fused) {
            const double VV = vn[1] * vn[1] + vn[2] * vn[2] + vn[3] * vn[3];
            const double pp = (1.4 - 1.0) * vn[4] - 0.5 * (1.4 - 1.0) * VV / vn[0];
            f.PP[c] = pp;
            f.T[c] = p.rxm2 * pp / vn[0];
            dr[1] = fabs(f.D[5 * tot + c] / vn[0]);
            dr[2] = fabs(f.D[6 * tot + c] / vn[0]);
        } else if (wout && has_w(f)) {
            // step sizes as Time/TURBTWO_TIME_LUSGS_SST.F90:567-568 forms them
            const double dtl = p.dt_lusgs_main;
            const double stept = 2.0 * dtm / (3.0 * dtm / dtl + 2.0);
            const long w = widx(b, f.march, i, j, k);
            wput(f, W_STEPT, w, stept);
            wput(f, W_CVS, w, stept / vol);
#pragma unroll
            for (int L = 0; L < 4; L++) wput(f, W_V + L, w, L == 0 ? 1.0 / vn[0] : vn[L] * (1.0 / vn[0]));
#pragma unroll
            for (int d = 0; d < 3; d++) wput(f, W_RDS + d, w, f.Rds[(long)d * tot + c]);
        }
    }
    norm_reduce(dr, arg, part, 0);
}

// The k-omega sweep's inputs of an explicit stage into the hyperplane-major buffer (march axis
// j) tile by tile (see k_lusgs_rhs_w): the step sizes of Time/TURBTWO_TIME_LUSGS_SST.F90:567-568,
// 1/rho and the velocities of the stage's new state, and Rds, the values k_rk_stage writes
// one element per cache line otherwise.
__global__ void __launch_bounds__(256) k_rk_w(trans_params p, Fields f, const BlkInfo* __restrict__ blks,
                                              const Seg* __restrict__ seg)
{
    __shared__ double sv[9][WT * WT];
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int ti = threadIdx.x % WT, tj = threadIdx.x / WT;
    const TileCell tc = wtile_cell(s, b, ti, tj);
    const long tot = f.tot;
    if (tc.ok) {
        const long c = cidx(b, tc.i, tc.j, tc.k);
        const double dtm = f.Dtm[c], vol = f.Vol[c];
        const double dtl = p.dt_lusgs_main;
        const double stept = 2.0 * dtm / (3.0 * dtm / dtl + 2.0);
        sv[0][threadIdx.x] = stept;
        sv[1][threadIdx.x] = stept / vol;
        const double v0 = f.V[c];
        sv[2][threadIdx.x] = 1.0 / v0;
#pragma unroll
        for (int L = 1; L < 4; L++) sv[2 + L][threadIdx.x] = f.V[(long)L * tot + c] * (1.0 / v0);
#pragma unroll
        for (int d = 0; d < 3; d++) sv[6 + d][threadIdx.x] = f.Rds[(long)d * tot + c];
    }
    __syncthreads();
    int di, dj;
    wtile_diag(threadIdx.x, di, dj);
    const TileCell oc = wtile_cell(s, b, di, dj);
    if (!oc.ok) return;
    const long w = widx(b, 1, oc.i, oc.j, oc.k);
    const int q = dj * WT + di;
    wput(f, W_STEPT, w, sv[0][q]);
    wput(f, W_CVS, w, sv[1][q]);
#pragma unroll
    for (int L = 0; L < 4; L++) wput(f, W_V + L, w, sv[2 + L][q]);
#pragma unroll
    for (int d = 0; d < 3; d++) wput(f, W_RDS + d, w, sv[6 + d][q]);
}

// Time/TIME_RK_V_Update_New.F90:75-260 after the k-omega LU-SGS: p and T of the new state
// (no rotation term), the positivity-damped k/omega update with the SST limiter, which here
// takes p = 0.4*rho*E - 0.2*|rho u|^2/rho and the k-omega Sutherland form without Csth1
// (:126-132), and the k and omega norms of the increments (:265-316).  The density norm of the
// stage stays in part.
__global__ void __launch_bounds__(256) k_rk_update_v(trans_params p, int ki, int if_turb, Fields f,
                                                     const BlkInfo* __restrict__ blks,
                                                     const Seg* __restrict__ seg, NormPart* part)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    const long tot = f.tot;
    double dr[3] = {0.0, 0.0, 0.0};
    int arg = -1;
    if (lin < NI1 * NJ1 * NK1) {
        int i, j, k;
        decode3(lin, NI1, NJ1, i, j, k);
        arg = lin;
        const long c = cidx(b, i, j, k);
        const double v1 = f.V[c], v2 = f.V[tot + c], v3 = f.V[2 * tot + c], v4 = f.V[3 * tot + c];
        const double v5 = f.V[4 * tot + c];
        const double VV = v2 * v2 + v3 * v3 + v4 * v4;
        const double pp = (1.4 - 1.0) * v5 - 0.5 * (1.4 - 1.0) * VV / v1;
        f.PP[c] = pp;
        f.T[c] = p.rxm2 * pp / v1;
        if (p.if_euler == 0 && if_turb == 1 && ki >= p.kitm) {
            const double PPP = 0.4 * v5 - 0.2 * (sqr(v2) + sqr(v3) + sqr(v4)) / v1;
            const double TT = p.rxm2 * PPP / v1;
            const double RmiuL = (1.0 + p.csthlnd) / (TT + p.csthlnd) * pow(TT, 1.5);
            sst_kw_update(p, f, c, v1, RmiuL);
        }
        dr[1] = fabs(f.D[5 * tot + c] / v1);
        dr[2] = fabs(f.D[6 * tot + c] / v1);
    }
    norm_reduce(dr, arg, part, 1);
}
