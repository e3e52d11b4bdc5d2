// Device-side layout and helpers shared by all kernels.
//
// Every field of every block lives in one device buffer per field, component-major:
//   F[c * tot + blk.off + (i+2) + nx*((j+2) + ny*(k+2))],   i,j,k in -2..N+2 (N = points)
// so that i is the contiguous (coalesced) index and all blocks can be addressed by one
// kernel launch through a per-workgroup schedule table (Seg).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/trans_hip.h"

struct BlkInfo {
    int ni, nj, nk;  // points
    int nx, ny, nz;  // padded extents (N+5)
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

// components of Fields::W / Fields::Wf (the mean-flow diagonal F1 = 1 + sum(Rds)*cvs is
// re-formed by the sweep from W_RDS and W_CVS)
constexpr int W_D = 0;      // 7: right-hand side, then the increment
constexpr int W_STEPT = 7;  // stept
constexpr int W_CVS = 8;    // stept / Vol
constexpr int W_F1T = 9;    // k-omega diagonal
constexpr int W_V = 10;     // 4: rho, rho*u, rho*v, rho*w
constexpr int W_PP = 14;    // pressure
constexpr int W_RDS = 15;   // 3: spectral radii
constexpr int W_SDM = 18;   // 9: face-mean vectors (geometry)
constexpr int W_NCOMP = 27;

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

__device__ __forceinline__ long cidx(const BlkInfo& b, int i, int j, int k)
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
