// Geometry pre-processing on the device: the brute-force wall distance of
// Geom/CalcDst.F90:121-158 (every cell centre against every wall panel centre).
#pragma once
#include <hip/hip_runtime.h>

// Panels are staged through LDS in tiles; each thread keeps the running minimum of the
// squared distance of WD_CPT cells, so a panel read from LDS serves WD_CPT cells.  The
// squared distance is formed as the reference forms it, ((dx*dx + dy*dy) + dz*dz) with
// d = cell - panel, and the minimum is exact, so the result does not depend on the tiling.
constexpr int WD_TILE = 512;
constexpr int WD_CPT = 4;

__global__ void __launch_bounds__(256) k_wall_distance(const double* __restrict__ cell, long ncell,
                                                       const double* __restrict__ pan, long npan,
                                                       double* __restrict__ d2min)
{
    __shared__ double sx[WD_TILE], sy[WD_TILE], sz[WD_TILE];
    const long base = ((long)blockIdx.x * blockDim.x) * WD_CPT + threadIdx.x;
    double cx[WD_CPT], cy[WD_CPT], cz[WD_CPT], best[WD_CPT];
#pragma unroll
    for (int q = 0; q < WD_CPT; q++) {
        const long c = base + (long)q * blockDim.x;  // cells of one thread are blockDim apart: coalesced loads
        const bool ok = c < ncell;
        cx[q] = ok ? cell[c] : 0.0;
        cy[q] = ok ? cell[ncell + c] : 0.0;
        cz[q] = ok ? cell[2 * ncell + c] : 0.0;
        best[q] = 1.0e20;
    }
    for (long t0 = 0; t0 < npan; t0 += WD_TILE) {
        const int nt = (int)((npan - t0) < WD_TILE ? (npan - t0) : WD_TILE);
        __syncthreads();
        for (int q = threadIdx.x; q < nt; q += blockDim.x) {
            sx[q] = pan[t0 + q];
            sy[q] = pan[npan + t0 + q];
            sz[q] = pan[2 * npan + t0 + q];
        }
        __syncthreads();
        for (int p = 0; p < nt; p++) {
            const double px = sx[p], py = sy[p], pz = sz[p];
#pragma unroll
            for (int q = 0; q < WD_CPT; q++) {
                const double dx = cx[q] - px, dy = cy[q] - py, dz = cz[q] - pz;
                const double ddd = dx * dx + dy * dy + dz * dz;
                best[q] = ddd < best[q] ? ddd : best[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < WD_CPT; q++) {
        const long c = base + (long)q * blockDim.x;
        if (c < ncell) d2min[c] = best[q];
    }
}
