// Device-side layout and helpers shared by all kernels.
//
// Every field of every block lives in one device buffer per field, component-major:
//   F[c * tot + blk.off + (i+2) + nx*((j+2) + ny*(k+2))],   i,j,k in -2..N+2 (N = points),
//   nx = N_i + 5 rounded up to a multiple of 16, blk.off = 13 mod 16, tot a multiple of 16
// so that i is the contiguous (coalesced) index and all blocks can be addressed by one
// kernel launch through a per-workgroup schedule table (Seg).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/trans_hip.h"

struct BlkInfo {
    int ni, nj, nk;  // points
    int nx, ny, nz;  // padded extents: row length (see ROW_ALIGN below), N_j + 5, N_k + 5
    long off;        // element offset of the block inside each field buffer
    int npatch;
    int patch0;      // first entry of this block in the device patch table
    long mwoff;      // offset of the block's wall marks in Fields::mw
    long woff;       // offset of the block in the hyperplane-major sweep buffer Fields::W
};

// One workgroup's share of an index space: block `blk`, linear indices base..base+255.
struct Seg {
    int blk;
    int base;
};

struct Fields {
    long tot;  // component stride
    double* V;        // 7
    double* PP;
    double* T;
    double* Rmiu;
    double* SD;       // 9: (dir-1) + 3*(xyz-1)
    double* SDm;      // 9: mean of a cell's two faces of direction dir, same component order
    double* Grad;     // 3
    double* Alagm;    // 3
    double* Vol;
    double* rad;
    double* Dst;
    double* aLeng;
    double* aSeng;    // smallest neighbour distance (IDDES)
    double* Yc;
    double* Zc;
    double* thtc;
    double* D;        // 7
    double* Rds;      // 3
    double* Dtm;
    double* Wt0;      // 7
    double* Wt1;      // 7
    double* Wt2;      // 7
    double* DQ;       // 18
    double* Q;        // 5: u, v, w, k/rho, omega/rho (the divisions of the face stencils, formed once per residual)
    double* shock;
    double* Fscheme;  // DES blending of the rotated Roe dissipation (zero for inviscid runs)
    double* F1;       // LU-SGS diagonal
    double* Stp;      // 2: dual-time pseudo step stept and stept/Vol of the LU-SGS
    double* F1t;      // k-omega LU-SGS diagonal
    double* G[3];     // area-weighted convective face fluxes per direction, 7 comps each
    double* GV[3];    // viscous face fluxes per direction: x, y, z momentum, energy, k, omega
    float* Ds;        // FP32 copy of D for the FP32 implicit operator (7)
    float* Rdsf;      // FP32 Rds (3)
    int* mw;          // wall marks per block: i0 (NJ1*NK1), im, j0 (NI1*NK1), jm, k0 (NI1*NJ1), km
    // Inputs and increments of the line-march LU-SGS in hyperplane-major order (see widx):
    // component stride wtot, components W_* below; W for the FP64 operator, Wf (FP32) for
    // the FP32 one.  Both null when the hyperplane fallback sweep is used.
    double* W;
    float* Wf;
    long wtot;        // component stride: the slots of all blocks, then the sinks
    long wsink;       // first sink slot (idle lanes of the line sweep store there)
    int march;        // march axis of the line sweep (0 i, 1 j, 2 k)
};

// components of Fields::W / Fields::Wf (the line-march mean-flow sweep reads the step sizes
// already divided by its diagonal F1 = 1 + sum(Rds)*cvs; the per-plane sweep re-forms F1 from
// W_RDS and W_CVS)
constexpr int W_D = 0;      // 7: right-hand side, then the increment
constexpr int W_STEPT = 7;  // stept
constexpr int W_CVS = 8;    // stept / Vol
constexpr int W_F1T = 9;    // k-omega diagonal
constexpr int W_V = 10;     // 4: 1/rho, u, v, w (u = (rho u) * (1/rho), as the sweep forms them)
constexpr int W_PP = 14;    // pressure
constexpr int W_RDS = 15;   // 3: spectral radii
constexpr int W_SDM = 18;   // 9: face-mean vectors (geometry)
constexpr int W_SA = 27;    // stept / F1 (mean flow)
constexpr int W_SB = 28;    // cvs / F1 (mean flow)
constexpr int W_NCOMP = 29;

// Whether the sweep inputs live in the hyperplane-major buffer.
__device__ __forceinline__ bool has_w(const Fields& f) { return f.W != nullptr || f.Wf != nullptr; }

// Component `comp` of the hyperplane-major buffer at slot w (rounded to FP32 for the FP32
// operator, whose sweep computes in FP32 anyway).
__device__ __forceinline__ void wput(const Fields& f, int comp, long w, double v)
{
    if (f.Wf)
        f.Wf[(long)comp * f.wtot + w] = (float)v;
    else
        f.W[(long)comp * f.wtot + w] = v;
}

// Rows are padded to a multiple of 16 elements and every block's offset is 13 modulo 16, so
// interior cell i = 1 of every row sits on a multiple of 16 elements (128 bytes in FP64): the
// 16- and 64-wide row segments the kernels read start on cache-line boundaries.
constexpr int ROW_ALIGN = 16;

__host__ __device__ __forceinline__ long cidx(const BlkInfo& b, int i, int j, int k)
{
    return b.off + (long)(i + 2) + (long)b.nx * ((long)(j + 2) + (long)b.ny * (long)(k + 2));
}

// Decode a linear index over an ni x nj x nk box starting at (1,1,1).
__device__ __forceinline__ void decode3(int lin, int ni, int nj, int& i, int& j, int& k)
{
    i = lin % ni;
    int r = lin / ni;
    j = r % nj;
    k = r / nj;
    i += 1;
    j += 1;
    k += 1;
}

// Faces of one block in FTA x FTB tiles of the plane spanned by i and the stencil direction
// DIR (2: the (i, j) plane, then k; 3: the (i, k) plane, then j), 256 per workgroup: the
// neighbours a face stencil reads along DIR are read by the same workgroup, so they come from
// the L2 of one XCD instead of being fetched again by workgroups that the dispatcher put on
// other XCDs, while each wave still reads one row of 64 consecutive i.  Returns false for the
// padding of partial tiles.
#ifndef TRANS_FTA
#define TRANS_FTA 64
#define TRANS_FTB 4
#endif
constexpr int FTA = TRANS_FTA, FTB = TRANS_FTB;
template <int DIR>
__device__ __forceinline__ bool decode_face_tile(int lin, int n0, int n1, int n2, int& i, int& j, int& k)
{
    const int tile = lin / (FTA * FTB), w = lin % (FTA * FTB);
    const int ti = w % FTA, tm = w / FTA;
    const int ntx = (n0 + FTA - 1) / FTA;
    const int tx = tile % ntx;
    i = tx * FTA + ti + 1;
    if (DIR == 2) {
        const int nty = (n1 + FTB - 1) / FTB;
        j = (tile / ntx) % nty * FTB + tm + 1;
        k = tile / (ntx * nty) + 1;
        return i <= n0 && j <= n1;
    }
    const int ntz = (n2 + FTB - 1) / FTB;
    k = (tile / ntx) % ntz * FTB + tm + 1;
    j = tile / (ntx * ntz) + 1;
    return i <= n0 && k <= n2;
}

// Hyperplane-major index of interior cell (i,j,k) for the line-march sweeps.  Lines run
// along the march axis m; the two other axes a < b number the lines a-fastest.  Cells of one
// hyperplane s = (i-1)+(j-1)+(k-1) are stored contiguously by line, so the lanes of a wave,
// which hold consecutive lines on the same hyperplane at every step, read consecutive
// addresses (in the natural layout they would be nx-1 elements apart).
__device__ __forceinline__ long widx(const BlkInfo& b, int march, int i, int j, int k)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1;
    const int Na = march == 0 ? NJ1 : NI1;
    const int Nb = march == 2 ? NJ1 : b.nk - 1;
    const int a = march == 0 ? j : i;
    const int bb = march == 2 ? j : k;
    const long s = (long)(i - 1) + (j - 1) + (k - 1);
    return b.woff + s * (long)Na * Nb + (a - 1) + (long)Na * (bb - 1);
}

__device__ __forceinline__ double sutherland(const trans_params& p, double tt)
{
    return (1.0 + p.csthlnd) / (tt * p.csth1 + p.csthlnd) * pow(p.csth1 * tt, 1.5);
}

__device__ __forceinline__ double sqr(double x) { return x * x; }

// Common/UTIL_Precond_Funcs.F90:6-30 (getPrecondPara): (U_r/c)^2 of the low-Mach
// preconditioning, min(max(M^2, 3 M_inf^2), 1), raised to the viscous limit
// ((mu + mu_t) / rho / Re / Alagm)^2 / c^2.
__device__ __forceinline__ double precond_para(const trans_params& p, double VM2, double AM2, double RR, double TT,
                                               double RMIUT, double ALAGMR)
{
    const double ub = fmin(fmax(VM2 / AM2, 3.0 * p.xm * p.xm), 1.0);
    const double RmiuAll = sutherland(p, TT) + RMIUT;
    const double r = RmiuAll / RR / p.ref / ALAGMR;
    return fmax(ub, 1.0 * (r * r) / AM2);
}

// Common/UTIL_Precond_Funcs.F90:86-130 (PrecondDW): a conservative increment a[5] through
// the preconditioning transform; the momentum components pass unchanged.
template <typename R>
__device__ __forceinline__ void precond_dw(const trans_params& p, R a[5], R pRo, R pVx, R pVy, R pVz, R pE, R AAc,
                                           R ubetac2, R pTT)
{
    const R pU2 = pVx * pVx + pVy * pVy + pVz * pVz;
    const R pUAU = pVx * a[1] + pVy * a[2] + pVz * a[3];
    a[4] = a[4] - (pUAU) - (pE / R(1.4) - R(1.2) * (pU2)) * a[0];
    const R rCp = R(1.4 / 0.4) * R(p.ppf) * pRo;
    const R a11 = AAc * AAc * ubetac2;
    const R a15 = R(-0.4) * ubetac2 / rCp;
    const R b11 = R(1) * a11 / rCp;
    const R b15 = (R(1) + R(0.4) * ubetac2) / rCp;
    const R AinP = a11 * a[0] + a15 * a[4];
    const R AinT = b11 * a[0] + b15 * a[4];
    a[0] = AinP / R(p.ppf) / pTT - pRo * AinT / pTT;
    a[4] = AinP / R(0.4) + (pUAU);
}

// The transform with the state of cell c (Time/TIME_LUSGS.F90:113-160 for the right-hand
// side, :235-246 and the mirrored blocks for the neighbour terms): own velocity, energy,
// p, T, eddy viscosity and the smallest of the cell's three Alagm.
template <typename R>
__device__ __forceinline__ void precond_cell(const trans_params& p, const Fields& f, long c, R a[5])
{
    const long tot = f.tot;
    const double pRo = f.V[c], roinv = 1.0 / pRo;
    const double pVx = f.V[tot + c] * roinv, pVy = f.V[2 * tot + c] * roinv, pVz = f.V[3 * tot + c] * roinv;
    const double pE = f.V[4 * tot + c] * roinv;
    const double pPP = f.PP[c], pTT = f.T[c];
    const double AAc = sqrt(1.4 * pPP * roinv);
    const double UVW2 = pVx * pVx + pVy * pVy + pVz * pVz;
    const double ALL = fmin(fmin(f.Alagm[c], f.Alagm[tot + c]), f.Alagm[2 * tot + c]);
    const double ub = precond_para(p, UVW2, AAc * AAc, pRo, pTT, f.Rmiu[c], ALL);
    precond_dw<R>(p, a, R(pRo), R(pVx), R(pVy), R(pVz), R(pE), R(AAc), R(ub), R(pTT));
}
