// Implicit update kernels: LU-SGS right-hand side, diagonal, the two sweeps as
// diagonal-hyperplane wavefronts, the state update with its residual norms, and the
// dual-time history shift.
#pragma once
#include "dev_common.h"

// The implicit operator (LU-SGS sweeps) may contract multiply-adds into FMAs: its result
// only has to reach the update within the parity bar (it also runs in FP32), while the
// residual kernels keep the oracle's exact operation sequence (-ffp-contract=off).
#define TRANS_SWEEP_CONTRACT _Pragma("clang fp contract(fast)")

// ---------------------------------------------------------------------------------------
// Time/TIME_LUSGS.F90:60-108: BDF2 right-hand side and the scalar diagonal F1, then with
// preconditioning the right-hand side through the cell's transform (:113-160).  With the
// line-march sweep the right-hand side and the state the sweep needs go to the
// hyperplane-major buffer (the sweep re-forms F1 from Rds and cvs); otherwise D, F1 and Stp.
// The hyperplane-major buffer with march axis j, written or read through LDS: a workgroup owns
// a WT x WT tile of cells (i, j) at one k (schedule entry Seg{block, tile index}).  A diagonal
// i + j = const of the tile is a run of consecutive slots (consecutive lines of one
// hyperplane), so the tile moves to or from the buffer diagonal by diagonal in whole cache
// lines instead of one element per line.
constexpr int WT = 16;

struct TileCell {
    int i, j, k;
    bool ok;
};

__device__ __forceinline__ TileCell wtile_cell(const Seg& s, const BlkInfo& b, int ti, int tj)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1;
    const int ntx = (NI1 + WT - 1) / WT, nty = (NJ1 + WT - 1) / WT;
    const int t = s.base;
    const int tx = t % ntx, ty = (t / ntx) % nty;
    TileCell c{tx * WT + ti + 1, ty * WT + tj + 1, t / (ntx * nty) + 1, false};
    c.ok = c.i <= NI1 && c.j <= NJ1;
    return c;
}

// Tile position of the q-th cell in diagonal order (diagonals ti + tj ascending, then ti).
__device__ __forceinline__ void wtile_diag(int q, int& ti, int& tj)
{
    int dd = 0, c0 = 0;
    for (; dd < 2 * WT - 2; dd++) {
        const int n = dd < WT ? dd + 1 : 2 * WT - 1 - dd;
        if (q < c0 + n) break;
        c0 += n;
    }
    ti = (dd < WT ? 0 : dd - (WT - 1)) + (q - c0);
    tj = dd - ti;
}

// Time/TIME_LUSGS.F90:60-108 for cell c: right-hand side (through the preconditioning
// transform when on), stept, cvs and the diagonal F1.
__device__ __forceinline__ void lusgs_rhs_cell(const trans_params& p, const Fields& f, long c, double rhs[5],
                                               double& stept, double& cvs, double& f1)
{
    const long tot = f.tot;
    const double dtm = p.dt_lusgs_main;
    const double vol = f.Vol[c];
#pragma unroll
    for (int L = 0; L < 5; L++) {
        const long e = (long)L * tot + c;
        rhs[L] = f.D[e] / vol - (3.0 * f.V[e] - 4.0 * f.Wt0[e] + f.Wt1[e]) * 0.5 / dtm;
    }
    if (p.if_precondition) precond_cell<double>(p, f, c, rhs);
    const double Dt = f.Dtm[c];
    stept = 2.0 * Dt / (3.0 * Dt / dtm + 2.0);
    cvs = stept / vol;
    f1 = 1.0 + (f.Rds[c] + f.Rds[tot + c] + f.Rds[2 * tot + c]) * cvs;
}

// The 17 values the line-march and tile sweeps read from the hyperplane-major buffer, written
// by the right-hand side, with their components: D(1:5), stept, cvs, stept/F1, cvs/F1, 1/rho,
// u, v, w, p, Rds(1:3).
constexpr int RHS_NW = 17;
__device__ __forceinline__ int rhs_w_comp(int q)
{
    return q < 5 ? W_D + q : (q < 9 ? (q == 5 ? W_STEPT : (q == 6 ? W_CVS : (q == 7 ? W_SA : W_SB)))
                                    : (q < 13 ? W_V + (q - 9) : (q == 13 ? W_PP : W_RDS + (q - 14))));
}

__device__ __forceinline__ void lusgs_rhs_wvals(const trans_params& p, const Fields& f, long c, double v[RHS_NW])
{
    const long tot = f.tot;
    double stept, cvs, f1;
    lusgs_rhs_cell(p, f, c, v, stept, cvs, f1);
    v[5] = stept;
    v[6] = cvs;
    const double rf1 = 1.0 / f1;
    v[7] = stept * rf1;
    v[8] = cvs * rf1;
    const double roinv = 1.0 / f.V[c];
    v[9] = roinv;
#pragma unroll
    for (int L = 1; L < 4; L++) v[9 + L] = f.V[(long)L * tot + c] * roinv;
    v[13] = f.PP[c];
#pragma unroll
    for (int d = 0; d < 3; d++) v[14 + d] = f.Rds[(long)d * tot + c];
}

// The right-hand side into the hyperplane-major buffer (march axis j), tile by tile.  R is the
// buffer's element type (Wf for the FP32 operator, W otherwise): the values are rounded to it
// before the LDS stage, which halves the stage for the FP32 operator (more workgroups per CU).
template <typename R>
__global__ void __launch_bounds__(256) k_lusgs_rhs_w(trans_params p, Fields f, const BlkInfo* __restrict__ blks,
                                                     const Seg* __restrict__ seg)
{
    __shared__ R sv[RHS_NW][WT * WT];
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int ti = threadIdx.x % WT, tj = threadIdx.x / WT;
    const TileCell tc = wtile_cell(s, b, ti, tj);
    if (tc.ok) {
        double v[RHS_NW];
        lusgs_rhs_wvals(p, f, cidx(b, tc.i, tc.j, tc.k), v);
#pragma unroll
        for (int q = 0; q < RHS_NW; q++) sv[q][threadIdx.x] = (R)v[q];
    }
    __syncthreads();
    int di, dj;
    wtile_diag(threadIdx.x, di, dj);
    const TileCell oc = wtile_cell(s, b, di, dj);
    if (!oc.ok) return;
    const long w = widx(b, 1, oc.i, oc.j, oc.k);
    R* const Wd = std::is_same<R, float>::value ? (R*)f.Wf : (R*)f.W;
#pragma unroll
    for (int q = 0; q < RHS_NW; q++) Wd[(long)rhs_w_comp(q) * f.wtot + w] = sv[q][dj * WT + di];
}

__global__ void __launch_bounds__(256) k_lusgs_rhs(trans_params p, Fields f,
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
    if (has_w(f)) {
        double v[RHS_NW];
        lusgs_rhs_wvals(p, f, c, v);
        const long w = widx(b, f.march, i, j, k);
#pragma unroll
        for (int q = 0; q < RHS_NW; q++) wput(f, rhs_w_comp(q), w, v[q]);
    } else {
        double rhs[5], stept, cvs, f1;
        lusgs_rhs_cell(p, f, c, rhs, stept, cvs, f1);
#pragma unroll
        for (int L = 0; L < 5; L++) f.D[(long)L * tot + c] = rhs[L];
        f.F1[c] = f1;
        f.Stp[c] = stept;
        f.Stp[tot + c] = cvs;
    }
}

// Face-mean vectors 0.5*(S_d(f+1) + S_d(f)) of every interior cell (geometry; the LU-SGS
// flux Jacobians of Time/TIME_LUSGS.F90:190-256 use them for the neighbour cell), also in
// hyperplane-major order when the line-march sweep is used.
__global__ void __launch_bounds__(256) k_face_means(Fields f, const BlkInfo* __restrict__ blks,
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
    const long w = has_w(f) ? widx(b, f.march, i, j, k) : 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const long o = (long)(d + 3 * q) * tot;
            const double m = 0.5 * (f.SD[o + c + st[d]] + f.SD[o + c]);
            f.SDm[o + c] = m;
            if (has_w(f)) wput(f, W_SDM + d + 3 * q, w, m);
        }
}

// Increments D(c0+1..c0+neq) from the hyperplane-major buffer (march axis j) back to D, tile
// by tile: read diagonal by diagonal, written in the natural layout.
template <typename R>
__global__ void __launch_bounds__(256) k_unskew_w(int c0, int neq, Fields f, const R* __restrict__ Wd,
                                                  const BlkInfo* __restrict__ blks, const Seg* __restrict__ seg)
{
    __shared__ R sv[7][WT * WT];
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    int di, dj;
    wtile_diag(threadIdx.x, di, dj);
    const TileCell ic = wtile_cell(s, b, di, dj);
    if (ic.ok) {
        const long w = widx(b, 1, ic.i, ic.j, ic.k);
        for (int L = 0; L < neq; L++) sv[L][dj * WT + di] = Wd[(long)(c0 + L) * f.wtot + w];
    }
    __syncthreads();
    const int ti = threadIdx.x % WT, tj = threadIdx.x / WT;
    const TileCell oc = wtile_cell(s, b, ti, tj);
    if (!oc.ok) return;
    const long c = cidx(b, oc.i, oc.j, oc.k);
    for (int L = 0; L < neq; L++) f.D[(long)(c0 + L) * f.tot + c] = (double)sv[L][threadIdx.x];
}

// Increments D(c0+1..c0+neq) of the line-march sweep back to D in the natural layout (the
// state updates read D).
template <typename R>
__global__ void __launch_bounds__(256) k_unskew(int c0, int neq, Fields f, const R* __restrict__ Wd,
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
    const long w = widx(b, f.march, i, j, k);
    for (int L = c0; L < c0 + neq; L++) f.D[(long)L * f.tot + c] = (double)Wd[(long)L * f.wtot + w];
}

// Time/TIME_LUSGS.F90:696-768 (LUSGS_ADQ) -- flux-Jacobian times increment, gamma = 1.4, static
// grid.  The reference multiplies out the 5x5 Jacobian; this is the same product in its
// factored form (the linearised flux through S): with VP = u.S, the mass-flux increment
// S.dm and the pressure increment dp = (gamma-1) (dE - u.dm + |u|^2/2 drho),
//   dif0 = S.dm,  dif_i = u_i S.dm + VP (dm_i - u_i drho) + S_i dp,
//   dif4 = H S.dm + VP (dE + dp - H drho),  H = c^2/(gamma-1) + |u|^2/2,
// term by term equal to the reference's coefficients, in about half the operations.
// The direction-independent part of the product for one cell and increment, then the
// product through one face vector S.
template <typename R>
struct AdqPrep {
    R u[3], dm[3];
    R dp;   // pressure increment
    R H;    // total enthalpy
    R w[3]; // dm_i - u_i drho
    R q4;   // dE + dp - H drho
};

template <typename R>
__device__ __forceinline__ AdqPrep<R> adq_prep(R pVx, R pVy, R pVz, R pPP, R roinv, const R ain[5])
{
    TRANS_SWEEP_CONTRACT
    const R Gm1 = R(1.4 - 1.0);
    AdqPrep<R> a;
    a.u[0] = pVx;
    a.u[1] = pVy;
    a.u[2] = pVz;
    a.dm[0] = ain[1];
    a.dm[1] = ain[2];
    a.dm[2] = ain[3];
    const R uDm = pVx * ain[1] + pVy * ain[2] + pVz * ain[3];
    const R hU2 = R(0.5) * (pVx * pVx + pVy * pVy + pVz * pVz);
    a.dp = Gm1 * (ain[4] - uDm + hU2 * ain[0]);
    a.H = R(1.4) * pPP * roinv * (R(1) / Gm1) + hU2;
#pragma unroll
    for (int q = 0; q < 3; q++) a.w[q] = a.dm[q] - a.u[q] * ain[0];
    a.q4 = ain[4] + a.dp - a.H * ain[0];
    return a;
}

template <typename R>
__device__ __forceinline__ void adq_dir(const AdqPrep<R>& a, R aSDx, R aSDy, R aSDz, R dif[5])
{
    TRANS_SWEEP_CONTRACT
    const R VP = a.u[0] * aSDx + a.u[1] * aSDy + a.u[2] * aSDz;
    const R Sm = a.dm[0] * aSDx + a.dm[1] * aSDy + a.dm[2] * aSDz;
    dif[0] = Sm;
    dif[1] = a.u[0] * Sm + VP * a.w[0] + aSDx * a.dp;
    dif[2] = a.u[1] * Sm + VP * a.w[1] + aSDy * a.dp;
    dif[3] = a.u[2] * Sm + VP * a.w[2] + aSDz * a.dp;
    dif[4] = a.H * Sm + VP * a.q4;
}

template <typename R>
__device__ __forceinline__ void lusgs_adq(R pVx, R pVy, R pVz, R aSDx, R aSDy, R aSDz, R pPP,
                                          R roinv, const R ain[5], R dif[5])
{
    adq_dir<R>(adq_prep<R>(pVx, pVy, pVz, pPP, roinv, ain), aSDx, aSDy, aSDz, dif);
}

// Neighbour contribution 0.5*(A_nb dQ_nb +/- Rds_nb dQ_nb) of the sweeps
// (Time/TIME_LUSGS.F90:190-256 and the mirrored blocks).  `cn` is the neighbour cell and
// `st` the stride of direction d; A uses the mean of the neighbour's two d-faces.  With
// preconditioning A_nb dQ_nb passes through the neighbour's transform (:235-246).
template <typename R>
__device__ __forceinline__ void lusgs_nb(const trans_params& p, const Fields& f, const R* __restrict__ Dq, int d,
                                         long cn, long st, R sgn, R acc[5])
{
    TRANS_SWEEP_CONTRACT
    const long tot = f.tot;
    R ain[5];
#pragma unroll
    for (int L = 0; L < 5; L++) ain[L] = Dq[(long)L * tot + cn];
    const double pRo = f.V[cn];
    const double roinv = 1.0 / pRo;
    const double pVx = f.V[tot + cn] * roinv;
    const double pVy = f.V[2 * tot + cn] * roinv;
    const double pVz = f.V[3 * tot + cn] * roinv;
    const double* S = f.SD + (long)d * tot;
    const double aSDx = 0.5 * (S[cn + st] + S[cn]);
    const double aSDy = 0.5 * (S[3 * tot + cn + st] + S[3 * tot + cn]);
    const double aSDz = 0.5 * (S[6 * tot + cn + st] + S[6 * tot + cn]);
    R dif[5];
    lusgs_adq<R>(R(pVx), R(pVy), R(pVz), R(aSDx), R(aSDy), R(aSDz), R(f.PP[cn]), R(roinv), ain, dif);
    if (p.if_precondition) precond_cell<R>(p, f, cn, dif);
    R rdst;
    if constexpr (sizeof(R) == 4)
        rdst = f.Rdsf[(long)d * tot + cn];
    else
        rdst = f.Rds[(long)d * tot + cn];
#pragma unroll
    for (int L = 0; L < 5; L++) acc[L] = R(0.5) * (dif[L] + sgn * rdst * ain[L]);
}

// k-omega neighbour term 0.5*(V.S dQ +/- Rds dQ) (Time/TURBTWO_TIME_LUSGS_SST.F90:578-620).
template <typename R>
__device__ __forceinline__ void turb_nb(const Fields& f, const R* __restrict__ Dq, int d, long cn, long st,
                                        R sgn, R acc[2])
{
    TRANS_SWEEP_CONTRACT
    const long tot = f.tot;
    const double roinv = 1.0 / f.V[cn];
    const R pVx = R(f.V[tot + cn] * roinv), pVy = R(f.V[2 * tot + cn] * roinv), pVz = R(f.V[3 * tot + cn] * roinv);
    const double* S = f.SD + (long)d * tot;
    const R aSDx = R(0.5 * (S[cn + st] + S[cn]));
    const R aSDy = R(0.5 * (S[3 * tot + cn + st] + S[3 * tot + cn]));
    const R aSDz = R(0.5 * (S[6 * tot + cn + st] + S[6 * tot + cn]));
    const R VP = pVx * aSDx + pVy * aSDy + pVz * aSDz;
    R rdst;
    if constexpr (sizeof(R) == 4)
        rdst = f.Rdsf[(long)d * tot + cn];
    else
        rdst = f.Rds[(long)d * tot + cn];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const R dd = Dq[(long)(5 + q) * tot + cn];
        acc[q] = R(0.5) * (dd * VP + sgn * rdst * dd);
    }
}

// One workgroup per block walks the diagonal hyperplanes s = (i-1)+(j-1)+(k-1); every cell of
// a hyperplane depends only on the previous one, so this order performs exactly the
// reference's lexicographic Gauss-Seidel operations (Time/TIME_LUSGS.F90:172-436 forward,
// :442-675 backward; NEQ = 7 adds Time/TURBTWO_TIME_LUSGS_SST.F90:550-766).
// R = double or float (FP32 implicit operator); Dq is D or its FP32 copy.  NEQ = 2 runs the
// k-omega sweeps alone (the Runge-Kutta path, main.F90:160-163).
// One cell (i, j, k) of the hyperplane sweeps: its increment from the neighbours' on the
// previous hyperplane (natural layout).
template <typename R, bool FORWARD, int NEQ>
__device__ __forceinline__ void plane_cell(const trans_params& p, const Fields& f, R* Dq, const BlkInfo& b, int i,
                                           int j, int k)
{
    TRANS_SWEEP_CONTRACT
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const long tot = f.tot;
    const long st3[3] = {1, (long)b.nx, (long)b.nx * b.ny};
    const double dtm = p.dt_lusgs_main;
    const long c = cidx(b, i, j, k);
    const double Dt = f.Dtm[c];
    const double stept = 2.0 * Dt / (3.0 * Dt / dtm + 2.0);
    const double cvs = stept / f.Vol[c];
    R dn[3][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
    R tn[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    const int pos[3] = {i, j, k};
    const int nmax[3] = {NI1, NJ1, NK1};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const bool edge = FORWARD ? pos[d] == 1 : pos[d] == nmax[d];
        if (edge) continue;
        const long cn = FORWARD ? c - st3[d] : c + st3[d];
        if (NEQ != 2) lusgs_nb<R>(p, f, Dq, d, cn, st3[d], FORWARD ? R(1) : R(-1), dn[d]);
        if (NEQ != 5) turb_nb<R>(f, Dq, d, cn, st3[d], FORWARD ? R(1) : R(-1), tn[d]);
    }
    if (NEQ != 2) {
        const R rf1 = R(1) / R(f.F1[c]);  // one division per cell, five multiplications
#pragma unroll
        for (int L = 0; L < 5; L++) {
            const long e = (long)L * tot + c;
            if (FORWARD)
                Dq[e] = (Dq[e] * R(stept) + (dn[0][L] + dn[1][L] + dn[2][L]) * R(cvs)) * rf1;
            else
                Dq[e] = Dq[e] - (dn[0][L] + dn[1][L] + dn[2][L]) * R(cvs) * rf1;
        }
    }
    if (NEQ != 5) {
        const R rf1t = R(1) / R(f.F1t[c]);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const long e = (long)(5 + q) * tot + c;
            if (FORWARD)
                Dq[e] = (Dq[e] * R(stept) + (tn[0][q] + tn[1][q] + tn[2][q]) * R(cvs)) * rf1t;
            else
                Dq[e] = Dq[e] - (tn[0][q] + tn[1][q] + tn[2][q]) * R(cvs) * rf1t;
        }
    }
}

// Whole sweep of a block in one workgroup, hyperplane after hyperplane (the fallback when
// nothing else fits; one workgroup per block).
template <typename R, bool FORWARD, int NEQ>
__global__ void __launch_bounds__(1024) k_lusgs_sweep(trans_params p, Fields f, R* Dq,
                                                      const BlkInfo* __restrict__ blks)
{
    const BlkInfo b = blks[blockIdx.x];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int nplanes = NI1 + NJ1 + NK1 - 2;
    const int nlines = NJ1 * NK1;
    for (int s0 = 0; s0 < nplanes; s0++) {
        const int s = FORWARD ? s0 : nplanes - 1 - s0;
        for (int ln = threadIdx.x; ln < nlines; ln += blockDim.x) {
            const int j = ln % NJ1 + 1, k = ln / NJ1 + 1;
            const int i = s - (j - 1) - (k - 1) + 1;
            if (i < 1 || i > NI1) continue;
            plane_cell<R, FORWARD, NEQ>(p, f, Dq, b, i, j, k);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Line-march wavefront of the same sweeps.  The block is cut into lines along a march axis
// MARCH (j when the (i,k) cross-section fits one workgroup).  At step t the thread of line
// (a,b) updates the cell of its line on hyperplane s = (i-1)+(j-1)+(k-1).  After updating its
// increment a cell evaluates the three terms its downstream neighbours need,
// 0.5*(A_d dQ +/- Rds_d dQ) (exactly the expressions of Time/TIME_LUSGS.F90:190-256 the
// neighbour would evaluate), keeps the march-axis term in registers for its own next cell and
// hands the two others to the neighbouring lines through ping-pong LDS.  Incoming terms are
// summed in i, j, k order as the reference does.  One LDS-only barrier per step; the next
// cell's inputs are loaded while the current one is computed.
//
// The k-omega sweeps of Time/TURBTWO_TIME_LUSGS_SST.F90:550-766 (flux Jacobian: the normal
// velocity V.S of the neighbour, no grid velocity term; diagonal F1t) read only their own
// increments, so they run as their own wavefront: NEQ = 2 sweeps components 6 and 7 alone,
// in a second row of workgroups beside the mean-flow ones (NEQ = 5) when both are swept.
__host__ __device__ constexpr int sweep_c0(int neq) { return neq == 2 ? 5 : 0; }

template <typename S, int NEQ>
struct LineRaw {
    S d[NEQ];
    S stept, cvs, f1;  // f1: F1t (k-omega); the mean-flow F1 is re-formed from rds and cvs
    S v[4], pp;
    S sdm[3][3];  // face mean of SD in each direction
    S rds[3];
};

// A wave-uniform pointer made opaque to the compiler (readfirstlane), so that a load from it
// at a 32-bit byte offset compiles to the scalar-base + vector-offset form instead of a 64-bit
// address per lane and component.
template <typename T>
__device__ __forceinline__ T* sbase(T* p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

// Element w of a component array in global memory; w * sizeof(T) fits 32 bits (checked at
// creation).
#define TRANS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ldg(const T* base, unsigned w)
{
    const TRANS_GLOBAL char* b = (const TRANS_GLOBAL char*)base;
    return *(const TRANS_GLOBAL T*)(b + w * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stg(T* base, unsigned w, T v)
{
    TRANS_GLOBAL char* b = (TRANS_GLOBAL char*)base;
    *(TRANS_GLOBAL T*)(b + w * (unsigned)sizeof(T)) = v;
}

// Scalar bases of the sweep's components in the hyperplane-major buffer (W for the FP64
// operator, Wf for the FP32 one; R is the element type).
template <typename R, int NEQ>
struct LineSrc {
    R* d[NEQ];
    const R *stept, *cvs, *f1, *v[4], *pp, *sdm[3][3], *rds[3];
    __device__ __forceinline__ LineSrc(const Fields& f, R* Dq)
    {
        constexpr int C0 = sweep_c0(NEQ);
        const size_t wt = (size_t)f.wtot;
        const R* W = Dq;  // the increments share the buffer with the other inputs
#pragma unroll
        for (int L = 0; L < NEQ; L++) d[L] = sbase(Dq + (C0 + L) * wt);
        // the mean flow reads stept / F1 and cvs / F1 in their places
        stept = sbase(W + (C0 == 0 ? W_SA : W_STEPT) * wt);
        cvs = sbase(W + (C0 == 0 ? W_SB : W_CVS) * wt);
        f1 = sbase(W + W_F1T * wt);
#pragma unroll
        for (int L = 0; L < 4; L++) v[L] = sbase(W + (W_V + L) * wt);
        pp = sbase(W + W_PP * wt);
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
#pragma unroll
            for (int q = 0; q < 3; q++) sdm[dd][q] = sbase(W + (W_SDM + dd + 3 * q) * wt);
            rds[dd] = sbase(W + (W_RDS + dd) * wt);
        }
    }
};

// Inputs of one cell of the sweep from the hyperplane-major buffer (slot w).
template <typename R, int NEQ>
__device__ __forceinline__ void line_load(const LineSrc<R, NEQ>& src, unsigned w, LineRaw<R, NEQ>& x)
{
    constexpr bool MF = sweep_c0(NEQ) == 0;
#pragma unroll
    for (int L = 0; L < NEQ; L++) x.d[L] = ldg(src.d[L], w);
    x.stept = ldg(src.stept, w);
    x.cvs = ldg(src.cvs, w);
    x.f1 = MF ? R(1) : ldg(src.f1, w);
#pragma unroll
    for (int L = 0; L < 4; L++) x.v[L] = ldg(src.v[L], w);
    x.pp = MF ? ldg(src.pp, w) : R(0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int q = 0; q < 3; q++) x.sdm[d][q] = ldg(src.sdm[d][q], w);
        x.rds[d] = ldg(src.rds[d], w);
    }
}

__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One block's wavefront for NEQ = 5 (mean flow) or 2 (k-omega).  Lines are indexed
// a-fastest over the two axes other than MARCH (a < b).  LDS `buf`: 2 parities x 2 directions
// x NEQ rows of nlines + 2 values; slot nlines of a row is a sink for lanes with nothing to
// hand on, slot nlines + 1 stays zero for lanes with no upstream line.  `sink` is this
// workgroup's slot range past the data of each component (Fields::wsink).
//
// The step loop is free of wave-level branches: lanes outside their line's window compute on
// whatever their slot holds and send their results to the sinks.  Every step issues the same
// memory operations, which keeps the compiler's memory-counter waits precise across the loop
// (the next cell's loads stay in flight while the current cell is computed).
template <typename R, bool FORWARD, int NEQ, int MARCH>
__device__ __forceinline__ void line_sweep(const Fields& f, R* __restrict__ Dq, const BlkInfo& b, unsigned sink,
                                           R* __restrict__ buf)
{
    TRANS_SWEEP_CONTRACT
    static_assert(NEQ == 5 || NEQ == 2, "mean flow or k-omega");
    constexpr bool MF = NEQ == 5;
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int Nm = MARCH == 0 ? NI1 : (MARCH == 1 ? NJ1 : NK1);
    const int Na = MARCH == 0 ? NJ1 : NI1;
    const int Nb = MARCH == 2 ? NJ1 : NK1;
    const int nlines = Na * Nb;
    const int row = nlines + 2;
    const int zero = nlines + 1;
    const int nplanes = NI1 + NJ1 + NK1 - 2;
    const int ln = threadIdx.x;
    const bool active = ln < nlines;
    const int ia = active ? ln % Na + 1 : 1, ib = active ? ln / Na + 1 : 1;
    const R sgn = FORWARD ? R(1) : R(-1);
    // LDS slots this lane writes (the sink when there is no downstream line) and reads
    const int wa = active && (FORWARD ? ia < Na : ia > 1) ? (FORWARD ? ln + 1 : ln - 1) : nlines;
    const int wb = active && (FORWARD ? ib < Nb : ib > 1) ? (FORWARD ? ln + Na : ln - Na) : nlines;
    const int ra = !active ? nlines : ((FORWARD ? ia > 1 : ia < Na) ? ln : zero);
    const int rb = !active ? nlines : ((FORWARD ? ib > 1 : ib < Nb) ? ln : zero);
    const int start = FORWARD ? (ia - 1) + (ib - 1) : (Na - ia) + (Nb - ib);
    for (int q = ln; q < 4 * NEQ; q += blockDim.x) buf[q * row + zero] = R(0);
    __syncthreads();
    // hyperplane-major slot of the lane's line on hyperplane s is lw + s*nlines; padding
    // lanes read the last line's slots
    const unsigned lw = (unsigned)(b.woff + (active ? ln : nlines - 1));
    const unsigned my_sink = sink + (unsigned)ln;
    const LineSrc<R, NEQ> src(f, Dq);
    R tm[NEQ];  // term from the previous cell of this line
#pragma unroll
    for (int L = 0; L < NEQ; L++) tm[L] = R(0);
    // one step of the wavefront on cell `cur`, loading the next cell into `nxt`; the loop
    // below alternates two buffers so that loads stay in flight across a whole step
    auto step = [&](const int t, const LineRaw<R, NEQ>& cur, LineRaw<R, NEQ>& nxt) {
        const int par = t & 1;
        const int rel = t - start;
        const bool act = active && rel >= 0 && rel < Nm;
        const int s = FORWARD ? t : nplanes - 1 - t;
        const int sn = FORWARD ? min(s + 1, nplanes - 1) : max(s - 1, 0);
        line_load<R, NEQ>(src, lw + (unsigned)sn * (unsigned)nlines, nxt);
        const unsigned wst = act ? lw + (unsigned)s * (unsigned)nlines : my_sink;
        const double stept = cur.stept;
        const double cvs = cur.cvs;
        // mean flow: stept and cvs arrive divided by F1 (Time/TIME_LUSGS.F90:92-108, formed in
        // k_lusgs_rhs); k-omega: its own diagonal F1t
        const R rf1 = MF ? R(1) : R(1) / R(cur.f1);
        R dq[NEQ];
#pragma unroll
        for (int L = 0; L < NEQ; L++) {
            const R sa = buf[((par * 2 + 0) * NEQ + L) * row + ra];
            const R sb = buf[((par * 2 + 1) * NEQ + L) * row + rb];
            const R sm = rel == 0 ? R(0) : tm[L];
            // incoming terms summed in i, j, k order
            R sum;
            if constexpr (MARCH == 0)
                sum = sm + sa + sb;
            else if constexpr (MARCH == 1)
                sum = sa + sm + sb;
            else
                sum = sa + sb + sm;
            if constexpr (MF) {
                if (FORWARD)
                    dq[L] = R(cur.d[L]) * R(stept) + sum * R(cvs);
                else
                    dq[L] = R(cur.d[L]) - sum * R(cvs);
            } else {
                if (FORWARD)
                    dq[L] = (R(cur.d[L]) * R(stept) + sum * R(cvs)) * rf1;
                else
                    dq[L] = R(cur.d[L]) - sum * R(cvs) * rf1;
            }
            stg(src.d[L], wst, dq[L]);
        }
        const double roinv = double(cur.v[0]);
        const R pVx = R(cur.v[1]), pVy = R(cur.v[2]), pVz = R(cur.v[3]);
        const int npar = par ^ 1;
        AdqPrep<R> prep;
        if constexpr (MF) prep = adq_prep<R>(pVx, pVy, pVz, R(cur.pp), R(roinv), dq);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R o[NEQ];
            const R sx = R(cur.sdm[d][0]), sy = R(cur.sdm[d][1]), sz = R(cur.sdm[d][2]);
            const R rd = R(cur.rds[d]);
            if constexpr (MF) {
                R dif[5];
                adq_dir<R>(prep, sx, sy, sz, dif);
#pragma unroll
                for (int L = 0; L < 5; L++) o[L] = R(0.5) * (dif[L] + sgn * rd * dq[L]);
            } else {
                const R VP = pVx * sx + pVy * sy + pVz * sz;
#pragma unroll
                for (int L = 0; L < NEQ; L++) o[L] = R(0.5) * (dq[L] * VP + sgn * rd * dq[L]);
            }
            if (d == MARCH) {
#pragma unroll
                for (int L = 0; L < NEQ; L++) tm[L] = o[L];
            } else {
                const bool to_a = d == (MARCH == 0 ? 1 : 0);
                const int slot = to_a ? wa : wb;
                const int dir = to_a ? 0 : 1;
#pragma unroll
                for (int L = 0; L < NEQ; L++) buf[((npar * 2 + dir) * NEQ + L) * row + slot] = o[L];
            }
        }
        lds_barrier();
    };
    LineRaw<R, NEQ> bufA, bufB;
    line_load<R, NEQ>(src, lw + (unsigned)(FORWARD ? 0 : nplanes - 1) * (unsigned)nlines, bufA);
    for (int t = 0; t < nplanes; t += 2) {
        step(t, bufA, bufB);
        if (t + 1 < nplanes) step(t + 1, bufB, bufA);
    }
}

// NEQ = 7: blockIdx.y 0 sweeps the mean flow and 1 the k-omega equations of block
// blockIdx.x; NEQ = 5 or 2: one row of workgroups.  Dq is W (FP64 operator) or Wf.
template <typename R, bool FORWARD, int NEQ, int MAXT, int MARCH>
__global__ void __launch_bounds__(MAXT) k_lusgs_line(Fields f, R* Dq, const BlkInfo* __restrict__ blks)
{
    extern __shared__ unsigned char lds_raw[];
    R* buf = reinterpret_cast<R*>(lds_raw);
    const BlkInfo b = blks[blockIdx.x];
    const unsigned sink = (unsigned)(f.wsink + (long)(blockIdx.x * gridDim.y + blockIdx.y) * MAXT);
    if constexpr (NEQ == 7) {
        if (blockIdx.y == 0)
            line_sweep<R, FORWARD, 5, MARCH>(f, Dq, b, sink, buf);
        else
            line_sweep<R, FORWARD, 2, MARCH>(f, Dq, b, sink, buf);
    } else {
        line_sweep<R, FORWARD, NEQ, MARCH>(f, Dq, b, sink, buf);
    }
}

// One hyperplane of every block per launch, on the hyperplane-major buffer of the line march
// (march axis j): the host launches the planes in order, so the launch boundary is the
// barrier, and workgroups cover every block's lines, so there is no limit on a block's cross
// section.  A cell forms the terms of its three upstream neighbours itself from their slots on
// the previous plane (the same expressions, in the same order, as the line march evaluates
// them on the neighbour's side), sums them in i, j, k order and stores its increment in place.
// Consecutive lines hold consecutive slots, so every load is coalesced.
struct PlaneBlk {
    int nplanes;  // NI1 + NJ1 + NK1 - 2
    int na, nb;   // lines: i (fastest) by k
};

template <typename R, int NEQ>
__device__ __forceinline__ void wplane_term(const Fields& f, const R* __restrict__ W, size_t wt, unsigned wn, int d,
                                            R sgn, R o[NEQ])
{
    TRANS_SWEEP_CONTRACT
    constexpr bool MF = NEQ != 2;
    constexpr int C0 = sweep_c0(NEQ);
    R dq[NEQ];
#pragma unroll
    for (int L = 0; L < NEQ; L++) dq[L] = ldg(W + (C0 + L) * wt, wn);
    const double roinv = double(ldg(W + W_V * wt, wn));
    const R pVx = ldg(W + (W_V + 1) * wt, wn), pVy = ldg(W + (W_V + 2) * wt, wn), pVz = ldg(W + (W_V + 3) * wt, wn);
    const R sx = ldg(W + (W_SDM + d) * wt, wn), sy = ldg(W + (W_SDM + d + 3) * wt, wn),
            sz = ldg(W + (W_SDM + d + 6) * wt, wn);
    const R rd = ldg(W + (W_RDS + d) * wt, wn);
    if constexpr (NEQ == 7) {
        R dif[5];
        lusgs_adq<R>(pVx, pVy, pVz, sx, sy, sz, ldg(W + W_PP * wt, wn), R(roinv), dq, dif);
#pragma unroll
        for (int L = 0; L < 5; L++) o[L] = R(0.5) * (dif[L] + sgn * rd * dq[L]);
        const R VP = pVx * sx + pVy * sy + pVz * sz;
#pragma unroll
        for (int L = 5; L < 7; L++) o[L] = R(0.5) * (dq[L] * VP + sgn * rd * dq[L]);
    } else if constexpr (MF) {
        R dif[5];
        lusgs_adq<R>(pVx, pVy, pVz, sx, sy, sz, ldg(W + W_PP * wt, wn), R(roinv), dq, dif);
#pragma unroll
        for (int L = 0; L < 5; L++) o[L] = R(0.5) * (dif[L] + sgn * rd * dq[L]);
    } else {
        const R VP = pVx * sx + pVy * sy + pVz * sz;
#pragma unroll
        for (int L = 0; L < NEQ; L++) o[L] = R(0.5) * (dq[L] * VP + sgn * rd * dq[L]);
    }
}

template <typename R, bool FORWARD, int NEQ>
__global__ void __launch_bounds__(256) k_lusgs_wplane(Fields f, R* __restrict__ W, const BlkInfo* __restrict__ blks,
                                                      const Seg* __restrict__ seg, const PlaneBlk* __restrict__ pb,
                                                      int s0)
{
    TRANS_SWEEP_CONTRACT
    const Seg sg = seg[blockIdx.x];
    const PlaneBlk q = pb[sg.blk];
    if (s0 >= q.nplanes) return;
    const int ln = sg.base + threadIdx.x;
    const int nlines = q.na * q.nb;
    if (ln >= nlines) return;
    const BlkInfo b = blks[sg.blk];
    const int s = FORWARD ? s0 : q.nplanes - 1 - s0;
    const int i = ln % q.na + 1, k = ln / q.na + 1;
    const int j = s - (i - 1) - (k - 1) + 1;
    const int NJ1 = b.nj - 1;
    if (j < 1 || j > NJ1) return;
    constexpr bool MF = NEQ != 2;
    constexpr int C0 = sweep_c0(NEQ);
    const size_t wt = (size_t)f.wtot;
    const unsigned w = (unsigned)(b.woff + (long)s * nlines + ln);
    const unsigned wp = FORWARD ? w - (unsigned)nlines : w + (unsigned)nlines;  // same line, previous plane
    const R sgn = FORWARD ? R(1) : R(-1);
    const int pos[3] = {i, j, k};
    const int nmax[3] = {q.na, NJ1, q.nb};
    // upstream neighbours on the previous plane: i -/+ 1 (line -/+ 1), j -/+ 1 (same line),
    // k -/+ 1 (line -/+ na)
    const int dl[3] = {FORWARD ? -1 : 1, 0, FORWARD ? -q.na : q.na};
    R sum[NEQ];
#pragma unroll
    for (int L = 0; L < NEQ; L++) sum[L] = R(0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const bool edge = FORWARD ? pos[d] == 1 : pos[d] == nmax[d];
        R o[NEQ];
        if (edge) {
#pragma unroll
            for (int L = 0; L < NEQ; L++) o[L] = R(0);
        } else {
            wplane_term<R, NEQ>(f, W, wt, wp + (unsigned)dl[d], d, sgn, o);
        }
#pragma unroll
        for (int L = 0; L < NEQ; L++) sum[L] = d == 0 ? o[L] : sum[L] + o[L];
    }
    const double stept = ldg(W + W_STEPT * wt, w);
    const double cvs = ldg(W + W_CVS * wt, w);
    if constexpr (MF) {
        // F1 as Time/TIME_LUSGS.F90:92-108 forms it (k_lusgs_rhs evaluates the same expression)
        const double r0 = ldg(W + W_RDS * wt, w), r1 = ldg(W + (W_RDS + 1) * wt, w), r2 = ldg(W + (W_RDS + 2) * wt, w);
        const R rf1 = R(1) / R(1.0 + (double(r0) + double(r1) + double(r2)) * cvs);
#pragma unroll
        for (int L = 0; L < 5; L++) {
            R* a = W + (C0 + L) * wt;
            const R dv = ldg(a, w);
            stg(a, w, FORWARD ? (dv * R(stept) + sum[L] * R(cvs)) * rf1 : dv - sum[L] * R(cvs) * rf1);
        }
    }
    if constexpr (NEQ != 5) {
        constexpr int T0 = NEQ == 7 ? 5 : 0;
        const R rf1t = R(1) / ldg(W + W_F1T * wt, w);
#pragma unroll
        for (int L = T0; L < NEQ; L++) {
            R* a = W + (C0 + L) * wt;
            const R dv = ldg(a, w);
            stg(a, w, FORWARD ? (dv * R(stept) + sum[L] * R(cvs)) * rf1t : dv - sum[L] * R(cvs) * rf1t);
        }
    }
}

// Copy D(1:ncomp) of interior cells between the FP64 and FP32 arrays.
template <typename A, typename B>
__global__ void __launch_bounds__(256) k_convert(int ncomp, int c0, long tot, const A* __restrict__ src,
                                                 B* __restrict__ dst, const BlkInfo* __restrict__ blks,
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
    for (int L = c0; L < c0 + ncomp; L++) dst[(long)L * tot + c] = (B)src[(long)L * tot + c];
}

// ---------------------------------------------------------------------------------------
// Time/TIME_Update_V.F90:109-146 (norms of the increment, last maximum wins) and
// :502-690 (positivity-damped update of rho, p, T, velocities, rho*E), the SST k/omega
// update and limiter :709-817.  Per-workgroup partial norms go to `part`.
struct NormPart {
    double sum[3];
    double mx[3];
    int arg[3];
    int pad;
};

__device__ __forceinline__ void norm_combine(double& s, double& m, int& a, double s2, double m2, int a2)
{
    s += s2;
    if (m2 > m || (m2 == m && a2 > a)) {
        m = m2;
        a = a2;
    }
}

// k/omega update with positivity damping and the SST (or WD+) limiter on eddy viscosity
// (Time/TIME_Update_V.F90:709-858, Time/TIME_RK_V_Update_New.F90:104-230): increments from
// D(6:7), new k, omega and Rmiu written back.  pRo is the updated density and RmiuL the
// laminar viscosity, which the two callers evaluate in their own forms.
__device__ __forceinline__ void sst_kw_update(const trans_params& p, const Fields& f, long c, double pRo,
                                              double RmiuL)
{
    const long tot = f.tot;
    const double AinRTk = f.D[5 * tot + c], AinRTO = f.D[6 * tot + c];
    double ARTk = f.V[5 * tot + c], ARTO = f.V[6 * tot + c];
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
    const double DistN = f.Dst[c];
    const double* Q = f.DQ;
    const double dudx = Q[c], dudy = Q[tot + c], dudz = Q[2 * tot + c];
    const double dvdx = Q[3 * tot + c], dvdy = Q[4 * tot + c], dvdz = Q[5 * tot + c];
    const double dwdx = Q[6 * tot + c], dwdy = Q[7 * tot + c], dwdz = Q[8 * tot + c];
    const double Div = dudx + dvdy + dwdz;
    const double prod = (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sqr(dudy + dvdx) +
                         sqr(dvdz + dwdy) + sqr(dwdx + dudz) - 2.0 / 3.0 * Div * Div);
    if (p.kind_model == 1) {
        double arg2 = 2.0 * sqrt(ARTk / pRo) / p.beta_star / (ARTO / pRo) / DistN / p.ref;
        arg2 = fmax(arg2, 500.0 * RmiuL / DistN / DistN / ARTO / p.ref / p.ref);
        arg2 = tanh(arg2 * arg2);
        const double Sii = dudx, Sjj = dvdy, Skk = dwdz;
        const double Sij = (dudy + dvdx) / 2.0, Sik = (dudz + dwdx) / 2.0, Sjk = (dvdz + dwdy) / 2.0;
        const double Skl = sqrt(2.0 * (1.0 * Sii * Sii + 1.0 * Sjj * Sjj + 1.0 * Skk * Skk +
                                       2.0 * Sij * Sij + 2.0 * Sik * Sik + 2.0 * Sjk * Sjk));
        const double Cmua = fmin(1.0, p.a1 * ARTO / pRo / ((Skl)*arg2 + 1.e-10) * p.ref);
        // X**0.5 of the reference is the (correctly rounded) square root, as compilers evaluate it
        ARTO = fmax(ARTO, sqrt(pRo * Cmua * fabs(prod) / (p.prod_lmt * p.beta_star)) / p.ref);
        ARTk = fmin(ARTk, 100000.0 * RmiuL * ARTO / pRo / Cmua);
        f.V[5 * tot + c] = ARTk;
        f.V[6 * tot + c] = ARTO;
        f.Rmiu[c] = Cmua * ARTk / ARTO * pRo;
    } else if (p.kind_model == 2) {  // WD+ (Time/TIME_Update_V.F90:821-858)
        const double Sii = dudx - (dudx + dvdy + dwdz) / 3.;
        const double Sjj = dvdy - (dudx + dvdy + dwdz) / 3.;
        const double Skk = dwdz - (dudx + dvdy + dwdz) / 3.;
        const double Sij = (dudy + dvdx) / 2.0, Sik = (dudz + dwdx) / 2.0, Sjk = (dvdz + dwdy) / 2.0;
        const double Skl = sqrt(1.0 * Sii * Sii + 1.0 * Sjj * Sjj + 1.0 * Skk * Skk + 2.0 * Sij * Sij +
                                2.0 * Sik * Sik + 2.0 * Sjk * Sjk);
        // X**0.5 of the reference is the (correctly rounded) square root, as compilers evaluate it
        ARTO = fmax(ARTO, sqrt(pRo * fabs(prod) / (p.prod_lmt * p.beta_star)) / p.ref);
        ARTk = fmin(ARTk, 100000.0 * RmiuL * ARTO / pRo);
        f.V[5 * tot + c] = ARTk;
        f.V[6 * tot + c] = ARTO;
        ARTO = fmax(ARTO, 7. / 8. * pRo * sqrt(Skl * Skl * 2. / 0.09) / p.ref);
        f.Rmiu[c] = ARTk / ARTO * pRo;
    }
}

// Workgroup reduction of sum(dr^2) and max(dr) (last index wins ties) of the norms q >= q0
// into part[blockIdx.x]; the entries q < q0 are left as they are.  All threads call it.
__device__ __forceinline__ void norm_reduce(const double dr[3], int arg, NormPart* part, int q0)
{
    // within each 64-lane wave a butterfly of cross-lane exchanges, then the waves' results
    // through LDS (one barrier); the max keeps the last index among equal values in any order
    constexpr int NW = 256 / 64;
    __shared__ double ssum[3][NW];
    __shared__ double smax[3][NW];
    __shared__ int sarg[3][NW];
    double sm[3], mx[3];
    int ag[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        sm[q] = dr[q] * dr[q];
        mx[q] = arg >= 0 ? dr[q] : -1.0;
        ag[q] = arg;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double s2 = __shfl_xor(sm[q], off, 64);
            const double m2 = __shfl_xor(mx[q], off, 64);
            const int a2 = __shfl_xor(ag[q], off, 64);
            norm_combine(sm[q], mx[q], ag[q], s2, m2, a2);
        }
    const int w = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            ssum[q][w] = sm[q];
            smax[q][w] = mx[q];
            sarg[q][w] = ag[q];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        NormPart* o = part + blockIdx.x;
        for (int q = q0; q < 3; q++) {
            double s = ssum[q][0], m = smax[q][0];
            int a = sarg[q][0];
            for (int v = 1; v < NW; v++) norm_combine(s, m, a, ssum[q][v], smax[q][v], sarg[q][v]);
            o->sum[q] = s;
            o->mx[q] = m;
            o->arg[q] = a;
        }
        o->pad = 0;
    }
}

// shift: 0, or the Kind_Dual of a fused dual-time history shift (TIME_Update_U, as
// k_update_u) of the updated state, for the last sub-iteration of an outer iteration.
__global__ void __launch_bounds__(256) k_update_v(trans_params p, int ki, int if_turb, Fields f,
                                                  const BlkInfo* __restrict__ blks,
                                                  const Seg* __restrict__ seg, NormPart* part, int shift)
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
        const double v1 = f.V[c];
        dr[0] = fabs(f.D[c] / v1);
        dr[1] = fabs(f.D[5 * tot + c] / v1);
        dr[2] = fabs(f.D[6 * tot + c] / v1);

        const double AinRo = f.D[c], AinRx = f.D[tot + c], AinRy = f.D[2 * tot + c];
        const double AinRz = f.D[3 * tot + c], AinRe = f.D[4 * tot + c];
        double pRo = v1;
        const double roinv = 1.0 / pRo;
        const double rad = f.rad[c];
        double vn[5];
        if (p.if_precondition) {
            // Time/TIME_Update_V.F90:522-606: the increment added to the conservative state; the
            // kinetic energy divides by the density before the update, and T is formed from the
            // pressure before the update
            const double pRx = f.V[tot + c] + AinRx, pRy = f.V[2 * tot + c] + AinRy, pRz = f.V[3 * tot + c] + AinRz;
            double pRe = fabs(f.V[4 * tot + c] + AinRe);
            pRo = fabs(pRo + AinRo);
            const double DV = 0.5 * (pRx * pRx + pRy * pRy + pRz * pRz) * roinv;
            double pPP = (pRe - DV) * (1.4 - 1.0);
            if (pPP <= 0.0) {
                pPP = fabs(pPP);
                pRe = pPP / (1.4 - 1.0) + DV;
            }
            f.T[c] = p.rxm2 * f.PP[c] / pRo;
            f.PP[c] = pPP;
            vn[0] = pRo;
            vn[1] = pRx;
            vn[2] = pRy;
            vn[3] = pRz;
            vn[4] = pRe;
        } else {
            double pVx = f.V[tot + c] * roinv;
            double pVy = f.V[2 * tot + c] * roinv;
            double pVz = f.V[3 * tot + c] * roinv;
            const double AinX = (AinRx - pVx * AinRo) * roinv;
            const double AinY = (AinRy - pVy * AinRo) * roinv;
            const double AinZ = (AinRz - pVz * AinRo) * roinv;
            const double AinP = (1.4 - 1.0) * (AinRe - 0.5 * AinRo * (pVx * pVx + pVy * pVy + pVz * pVz) -
                                               pRo * (AinX * pVx + AinY * pVy + AinZ * pVz));
            if (AinRo >= 0.)
                pRo = pRo + AinRo;
            else
                pRo = pRo * exp(AinRo / (pRo + 1.e-10));
            pRo = fabs(pRo);
            double pp = f.PP[c];
            if (AinP >= 0.)
                pp = pp + AinP;
            else
                pp = pp * exp(AinP / (pp + 1.e-10));
            pp = fabs(pp);
            f.PP[c] = pp;
            f.T[c] = p.rxm2 * pp / pRo;
            pVx = pVx + AinX;
            pVy = pVy + AinY;
            pVz = pVz + AinZ;
            const double DV = 0.5 * (pVx * pVx + pVy * pVy + pVz * pVz);
            vn[0] = pRo;
            vn[1] = pRo * pVx;
            vn[2] = pRo * pVy;
            vn[3] = pRo * pVz;
            vn[4] = pp / (1.4 - 1.0) + DV * pRo - 0.5 * sqr(p.omega1 * rad) * pRo;
        }
#pragma unroll
        for (int L = 0; L < 5; L++) f.V[(long)L * tot + c] = vn[L];

        if (p.if_euler == 0 && if_turb == 1 && ki >= p.kitm) {
            const double VV = sqr(vn[1]) + sqr(vn[2]) + sqr(vn[3]);
            const double PPP = (1.4 - 1.0) * vn[4] - 0.5 * (1.4 - 1.0) * (VV / pRo - sqr(p.omega1 * rad) * pRo);
            const double TT = p.rxm2 * PPP / pRo;
            sst_kw_update(p, f, c, pRo, sutherland(p, TT));
        }
        if (shift) {
#pragma unroll
            for (int L = 0; L < TRANS_ML; L++) {
                const long e = (long)L * tot + c;
                if (shift == 2) f.Wt1[e] = f.Wt0[e];
                f.Wt0[e] = f.V[e];
            }
        }
    }
    norm_reduce(dr, arg, part, 0);
}

// Final per-block norms from the partials of the block's workgroups [w0, w1).
__global__ void k_norm_final(const NormPart* __restrict__ part, const int* __restrict__ wrange,
                             const BlkInfo* __restrict__ blks, trans_resid* out)
{
    const int blk = blockIdx.x;
    const BlkInfo b = blks[blk];
    const int w0 = wrange[blk], w1 = wrange[blk + 1];
    __shared__ NormPart sp[256];
    NormPart acc;
    for (int q = 0; q < 3; q++) {
        acc.sum[q] = 0.0;
        acc.mx[q] = -1.0;
        acc.arg[q] = -1;
    }
    for (int w = w0 + threadIdx.x; w < w1; w += blockDim.x)
        for (int q = 0; q < 3; q++)
            norm_combine(acc.sum[q], acc.mx[q], acc.arg[q], part[w].sum[q], part[w].mx[q], part[w].arg[q]);
    sp[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 3; q++)
                norm_combine(sp[threadIdx.x].sum[q], sp[threadIdx.x].mx[q], sp[threadIdx.x].arg[q],
                             sp[threadIdx.x + w].sum[q], sp[threadIdx.x + w].mx[q],
                             sp[threadIdx.x + w].arg[q]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
        trans_resid r;
        double* rs[3] = {&r.drsa, &r.drsk, &r.drso};
        double* rm[3] = {&r.dua, &r.duk, &r.duo};
        int* ri[3] = {&r.ima, &r.imk, &r.imo};
        int* rj[3] = {&r.jma, &r.jmk, &r.jmo};
        int* rk[3] = {&r.kma, &r.kmk, &r.kmo};
        for (int q = 0; q < 3; q++) {
            *rs[q] = sqrt(sp[0].sum[q] / NI1 / NJ1 / NK1);
            *rm[q] = sp[0].mx[q];
            const int a = sp[0].arg[q];
            *ri[q] = a % NI1 + 1;
            *rj[q] = (a / NI1) % NJ1 + 1;
            *rk[q] = a / (NI1 * NJ1) + 1;
        }
        out[blk] = r;
    }
}

// Time/TIME_Update_U.F90:64-92
__global__ void __launch_bounds__(256) k_update_u(int kind_dual, Fields f,
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
    for (int L = 0; L < TRANS_ML; L++) {
        const long e = (long)L * tot + c;
        if (kind_dual == 1) {
            f.Wt0[e] = f.V[e];
        } else if (kind_dual == 2) {
            f.Wt1[e] = f.Wt0[e];
            f.Wt0[e] = f.V[e];
        }
    }
}
