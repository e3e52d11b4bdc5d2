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

// ---------------------------------------------------------------------------------------
// O(N) metrics of one block (Geom/CalcDst.F90:170-632): cell centres and boundary face
// centres, face area vectors and areas, face-to-face lengths, neighbour distances, volumes
// and the first-ghost copies.  Arrays use the block's padded layout (bounds -2:N+2 on each
// axis, component first); every expression keeps the host restatement's operation order,
// so the results are bitwise equal to it (the angle thtc comes from the device atan2).
struct GeomGrid {
    int ni, nj, nk;  // points
    long nx, ny;     // padded extents ni+5, nj+5
    __host__ __device__ long at(int i, int j, int k) const { return (i + 2) + nx * ((j + 2) + ny * (long)(k + 2)); }
};

struct GeomOut {
    double *xc, *yc, *zc, *thtc, *rad, *aleng, *aseng, *vol;
    double* sd;     // 9 components first: (d-1) + 3*x
    double *grad, *alagm;  // 3 components first
};

// centres: interior cells and the boundary face centres at index 0 / N
__global__ void __launch_bounds__(256) k_geom_centres(GeomGrid g, const double* __restrict__ X,
                                                      const double* __restrict__ Y, const double* __restrict__ Z,
                                                      GeomOut o)
{
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    const long n0 = ni + 1, n1 = nj + 1;
    const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= n0 * n1 * (nk + 1)) return;
    const int i = (int)(lin % n0), j = (int)((lin / n0) % n1), k = (int)(lin / (n0 * n1));
    const bool ii = i >= 1 && i <= ni - 1, ji = j >= 1 && j <= nj - 1, ki = k >= 1 && k <= nk - 1;
    const double* P[3] = {X, Y, Z};
    double* C[3] = {o.xc, o.yc, o.zc};
    const long c = g.at(i, j, k);
    if (ii && ji && ki) {
        for (int q = 0; q < 3; q++) {
            const double* A = P[q];
            C[q][c] = (A[g.at(i + 1, j + 1, k + 1)] + A[g.at(i + 1, j + 1, k)] + A[g.at(i + 1, j, k + 1)] +
                       A[g.at(i + 1, j, k)] + A[g.at(i, j + 1, k + 1)] + A[g.at(i, j + 1, k)] + A[g.at(i, j, k + 1)] +
                       A[g.at(i, j, k)]) /
                      8.0;
        }
        const double yc = o.yc[c], zc = o.zc[c];
        o.thtc[c] = atan2(zc, yc);
        o.rad[c] = sqrt(yc * yc + zc * zc);
        return;
    }
    if ((i == 0 || i == ni) && ji && ki) {
        const int f = i == 0 ? 1 : ni;
        for (int q = 0; q < 3; q++) {
            const double* A = P[q];
            C[q][c] = (A[g.at(f, j, k)] + A[g.at(f, j + 1, k)] + A[g.at(f, j, k + 1)] + A[g.at(f, j + 1, k + 1)]) / 4.0;
        }
    } else if ((j == 0 || j == nj) && ii && ki) {
        const int f = j == 0 ? 1 : nj;
        for (int q = 0; q < 3; q++) {
            const double* A = P[q];
            C[q][c] = (A[g.at(i, f, k)] + A[g.at(i + 1, f, k)] + A[g.at(i, f, k + 1)] + A[g.at(i + 1, f, k + 1)]) / 4.0;
        }
    } else if ((k == 0 || k == nk) && ii && ji) {
        const int f = k == 0 ? 1 : nk;
        for (int q = 0; q < 3; q++) {
            const double* A = P[q];
            C[q][c] = (A[g.at(i, j, f)] + A[g.at(i + 1, j, f)] + A[g.at(i, j + 1, f)] + A[g.at(i + 1, j + 1, f)]) / 4.0;
        }
    }
}

// face area vectors S = (R1 x R2) / 2 * dl and areas |S| (plus tiny under the root) of the
// faces of direction DIR; dl orients them by the first cell's i direction (:198-210)
template <int DIR>
__global__ void __launch_bounds__(256) k_geom_faces(GeomGrid g, const double* __restrict__ X,
                                                    const double* __restrict__ Y, const double* __restrict__ Z,
                                                    GeomOut o, double tiny)
{
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    const int n0 = DIR == 1 ? ni : ni - 1, n1 = DIR == 2 ? nj : nj - 1, n2 = DIR == 3 ? nk : nk - 1;
    const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= (long)n0 * n1 * n2) return;
    const int i = (int)(lin % n0) + 1, j = (int)((lin / n0) % n1) + 1, k = (int)(lin / ((long)n0 * n1)) + 1;
    const double* P[3] = {X, Y, Z};
    double dl;
    {
        double xv[3], r1[3], r2[3];
        for (int q = 0; q < 3; q++) {
            const double* C = q == 0 ? o.xc : (q == 1 ? o.yc : o.zc);
            xv[q] = (C[g.at(1, 1, 1)] - C[g.at(2, 1, 1)]) * -1.0;
            r1[q] = P[q][g.at(1, 2, 2)] - P[q][g.at(1, 1, 1)];
            r2[q] = P[q][g.at(1, 1, 2)] - P[q][g.at(1, 2, 1)];
        }
        const double sv[3] = {(r1[1] * r2[2] - r1[2] * r2[1]) * 0.5, (r1[2] * r2[0] - r1[0] * r2[2]) * 0.5,
                              (r1[0] * r2[1] - r1[1] * r2[0]) * 0.5};
        dl = xv[0] * sv[0] + xv[1] * sv[1] + xv[2] * sv[2] < 0 ? -1.0 : 1.0;
    }
    double R1[3], R2[3];
    for (int q = 0; q < 3; q++) {
        const double* A = P[q];
        if (DIR == 1) {
            R1[q] = A[g.at(i, j + 1, k + 1)] - A[g.at(i, j, k)];
            R2[q] = A[g.at(i, j, k + 1)] - A[g.at(i, j + 1, k)];
        } else if (DIR == 2) {
            R1[q] = A[g.at(i, j, k + 1)] - A[g.at(i + 1, j, k)];
            R2[q] = A[g.at(i + 1, j, k + 1)] - A[g.at(i, j, k)];
        } else {
            R1[q] = A[g.at(i + 1, j, k)] - A[g.at(i, j + 1, k)];
            R2[q] = A[g.at(i + 1, j + 1, k)] - A[g.at(i, j, k)];
        }
    }
    const double S0 = (R1[1] * R2[2] - R1[2] * R2[1]) * 0.5 * dl;
    const double S1 = (R1[2] * R2[0] - R1[0] * R2[2]) * 0.5 * dl;
    const double S2 = (R1[0] * R2[1] - R1[1] * R2[0]) * 0.5 * dl;
    const long c = g.at(i, j, k);
    o.sd[(DIR - 1) + 9 * c] = S0;
    o.sd[(DIR - 1) + 3 + 9 * c] = S1;
    o.sd[(DIR - 1) + 6 + 9 * c] = S2;
    o.grad[(DIR - 1) + 3 * c] = sqrt(S0 * S0 + S1 * S1 + S2 * S2 + tiny);
}

__device__ __forceinline__ double geom_fc(const GeomGrid& g, const double* A, int i0, int i1, int j0, int j1, int k0,
                                          int k1)
{
    return 0.25 * (A[g.at(i0, j0, k0)] + A[g.at(i1, j1, k0)] + A[g.at(i0, j0, k1)] + A[g.at(i1, j1, k1)]);
}

// face-to-face lengths Alagm (:289-325), neighbour distances aleng/aseng (:430-476) and
// volumes (:481-515) of the interior cells
__global__ void __launch_bounds__(256) k_geom_cells(GeomGrid g, const double* __restrict__ X,
                                                    const double* __restrict__ Y, const double* __restrict__ Z,
                                                    GeomOut o)
{
    const int NI1 = g.ni - 1, NJ1 = g.nj - 1, NK1 = g.nk - 1;
    const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= (long)NI1 * NJ1 * NK1) return;
    const int i = (int)(lin % NI1) + 1, j = (int)((lin / NI1) % NJ1) + 1, k = (int)(lin / ((long)NI1 * NJ1)) + 1;
    const double* P[3] = {X, Y, Z};
    const long c = g.at(i, j, k);
    {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int q = 0; q < 3; q++) {
            const double* A = P[q];
            const double w = geom_fc(g, A, i, i, j, j + 1, k, k + 1), e = geom_fc(g, A, i + 1, i + 1, j, j + 1, k, k + 1);
            const double s = geom_fc(g, A, i, i + 1, j, j, k, k + 1), n = geom_fc(g, A, i, i + 1, j + 1, j + 1, k, k + 1);
            const double b = geom_fc(g, A, i, i + 1, j, j + 1, k, k), t = geom_fc(g, A, i, i + 1, j, j + 1, k + 1, k + 1);
            s0 = s0 + (w - e) * (w - e);
            s1 = s1 + (s - n) * (s - n);
            s2 = s2 + (b - t) * (b - t);
        }
        o.alagm[0 + 3 * c] = sqrt(s0);
        o.alagm[1 + 3 * c] = sqrt(s1);
        o.alagm[2 + 3 * c] = sqrt(s2);
    }
    {
        const double* C[3] = {o.xc, o.yc, o.zc};
        auto dd = [&](int di, int dj, int dk) {
            const long cn = g.at(i + di, j + dj, k + dk);
            const double a = C[0][cn] - C[0][c], b = C[1][cn] - C[1][c], e = C[2][cn] - C[2][c];
            return sqrt(a * a + b * b + e * e);
        };
        const double dw = dd(-1, 0, 0), de = dd(1, 0, 0), ds = dd(0, -1, 0), dn = dd(0, 1, 0), dlo = dd(0, 0, -1),
                     du = dd(0, 0, 1);
        double cand[6] = {dw, de, ds, dn, dlo, du};
        int nc = 6;
        if (NI1 == 2) {
            cand[0] = ds, cand[1] = dn, cand[2] = dlo, cand[3] = du;
            nc = 4;
        }
        if (NJ1 == 2) {
            cand[0] = dw, cand[1] = de, cand[2] = dlo, cand[3] = du;
            nc = 4;
        }
        if (NK1 == 2) {
            cand[0] = dw, cand[1] = de, cand[2] = ds, cand[3] = dn;
            nc = 4;
        }
        double mx = cand[0], mn = cand[0];
        for (int q = 1; q < nc; q++) {
            mx = fmax(mx, cand[q]);
            mn = fmin(mn, cand[q]);
        }
        o.aleng[c] = mx;
        o.aseng[c] = mn;
    }
    {
        const long p = g.at(i + 1, j + 1, k + 1);
        const long pi = g.at(i, j + 1, k + 1), pj = g.at(i + 1, j, k + 1), pk = g.at(i + 1, j + 1, k);
        double si = 0.0, sj = 0.0, sk = 0.0;
        for (int x = 0; x < 3; x++) {
            const double* A = P[x];
            si = si + (A[p] - A[pi]) * o.sd[0 + 3 * x + 9 * c];
            sj = sj + (A[p] - A[pj]) * o.sd[1 + 3 * x + 9 * c];
            sk = sk + (A[p] - A[pk]) * o.sd[2 + 3 * x + 9 * c];
        }
        double om = 0.0;
        om = om + si;
        om = om + sj;
        om = om + sk;
        o.vol[c] = (1.0 / 3.0) * om;
    }
}

// first-ghost copies of Vol and Alagm (:579-608), Grad's face-0 copy, and the SD extension
// to faces 0 and N+1 (:611-632); every read is of an interior entry no thread writes
__global__ void __launch_bounds__(256) k_geom_ghosts(GeomGrid g, GeomOut o)
{
    const int NI = g.ni, NJ = g.nj, NK = g.nk, NI1 = NI - 1, NJ1 = NJ - 1, NK1 = NK - 1;
    const long n0 = NI + 2, n1 = NJ + 2;
    const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= n0 * n1 * (NK + 2)) return;
    const int i = (int)(lin % n0), j = (int)((lin / n0) % n1), k = (int)(lin / (n0 * n1));
    const bool ii = i >= 1 && i <= NI1, ji = j >= 1 && j <= NJ1, ki = k >= 1 && k <= NK1;
    const long c = g.at(i, j, k);
    if (ji && ki && (i == 0 || i == NI)) {
        const long s = g.at(i == 0 ? 1 : NI1, j, k);
        o.vol[c] = o.vol[s];
        o.alagm[0 + 3 * c] = o.alagm[0 + 3 * s];
        if (i == 0) o.grad[0 + 3 * c] = o.grad[0 + 3 * s];
    }
    if (ii && ki && (j == 0 || j == NJ)) {
        const long s = g.at(i, j == 0 ? 1 : NJ1, k);
        o.vol[c] = o.vol[s];
        o.alagm[1 + 3 * c] = o.alagm[1 + 3 * s];
        if (j == 0) o.grad[1 + 3 * c] = o.grad[1 + 3 * s];
    }
    if (ii && ji && (k == 0 || k == NK)) {
        const long s = g.at(i, j, k == 0 ? 1 : NK1);
        o.vol[c] = o.vol[s];
        o.alagm[2 + 3 * c] = o.alagm[2 + 3 * s];
        if (k == 0) o.grad[2 + 3 * c] = o.grad[2 + 3 * s];
    }
    const bool jw = j >= 1 && j <= NJ, kw = k >= 1 && k <= NK, iw = i >= 1 && i <= NI;
    if (jw && kw && (i == 0 || i == NI + 1)) {
        const long s = g.at(i == 0 ? 2 : NI1, j, k);
        for (int x = 0; x < 3; x++) o.sd[0 + 3 * x + 9 * c] = o.sd[0 + 3 * x + 9 * s];
    }
    if (iw && kw && (j == 0 || j == NJ + 1)) {
        const long s = g.at(i, j == 0 ? 2 : NJ1, k);
        for (int x = 0; x < 3; x++) o.sd[1 + 3 * x + 9 * c] = o.sd[1 + 3 * x + 9 * s];
    }
    if (iw && jw && (k == 0 || k == NK + 1)) {
        const long s = g.at(i, j, k == 0 ? 2 : NK1);
        for (int x = 0; x < 3; x++) o.sd[2 + 3 * x + 9 * c] = o.sd[2 + 3 * x + 9 * s];
    }
}
