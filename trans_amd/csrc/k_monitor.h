// Plane integrals for monitoring: mass flow, area and mass-averaged total pressure and
// temperature over one constant-index face plane of a block.
#pragma once
#include "dev_common.h"

// One workgroup: each thread sums a strided share of the plane's faces, then a fixed-order
// tree reduction, so repeated calls give identical bits.  Face state and totals as
// ora_plane_integrals restates them (Output/OUTPUT_Residual_NS.F90:19-32,
// Output/Average_Field.F90:60-66), and the same totals with the absolute velocity (rotation
// omega about x).  out = {mass_flow, area, p0_mass_avg, t0_mass_avg, p0_abs, t0_abs}.
__global__ void __launch_bounds__(256) k_plane_integrals(Fields f, BlkInfo b, int dir, int index, double omega,
                                                         double* out)
{
    constexpr int NQ = 6;
    __shared__ double sh[NQ][256];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int n0 = dir == 1 ? NJ1 : NI1;
    const int n1 = dir == 3 ? NJ1 : NK1;
    const long tot = f.tot;
    double acc[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < n0 * n1; q += blockDim.x) {
        const int q0 = q % n0 + 1, q1 = q / n0 + 1;
        const int i = dir == 1 ? index : q0;
        const int j = dir == 1 ? q0 : (dir == 2 ? index : q1);
        const int k = dir == 3 ? index : q1;
        const long c = cidx(b, i, j, k);
        const long cm = cidx(b, i - (dir == 1), j - (dir == 2), k - (dir == 3));
        const double wei1 = f.Vol[cm], wei2 = f.Vol[c], wei = wei1 + wei2;
        const double ro = (wei1 * f.V[c] + wei2 * f.V[cm]) / wei;
        const double ru = (wei1 * f.V[tot + c] + wei2 * f.V[tot + cm]) / wei;
        const double rv = (wei1 * f.V[2 * tot + c] + wei2 * f.V[2 * tot + cm]) / wei;
        const double rw = (wei1 * f.V[3 * tot + c] + wei2 * f.V[3 * tot + cm]) / wei;
        const double pf = (wei1 * f.PP[c] + wei2 * f.PP[cm]) / wei;
        const double tf = (wei1 * f.T[c] + wei2 * f.T[cm]) / wei;
        const double* S = f.SD + (long)(dir - 1) * tot;
        const double s1 = S[c], s2 = S[3 * tot + c], s3 = S[6 * tot + c];
        const double dm = ru * s1 + rv * s2 + rw * s3;
        const double u = ru / ro, v = rv / ro, w = rw / ro;
        const double m2 = (u * u + v * v + w * w) / (1.4 * pf / ro);
        const double fac = 1.0 + 0.2 * m2;
        acc[0] += dm;
        acc[1] += sqrt(s1 * s1 + s2 * s2 + s3 * s3);
        acc[2] += pf * pow(fac, 3.5) * dm;
        acc[3] += tf * fac * dm;
        const double yf = (wei1 * f.Yc[c] + wei2 * f.Yc[cm]) / wei;
        const double zf = (wei1 * f.Zc[c] + wei2 * f.Zc[cm]) / wei;
        const double va = v - omega * zf, wa = w + omega * yf;
        const double fa = 1.0 + 0.2 * ((u * u + va * va + wa * wa) / (1.4 * pf / ro));
        acc[4] += pf * pow(fa, 3.5) * dm;
        acc[5] += tf * fa * dm;
    }
    for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0];
        out[2] = sh[2][0] / sh[0][0];
        out[3] = sh[3][0] / sh[0][0];
        out[4] = sh[4][0] / sh[0][0];
        out[5] = sh[5][0] / sh[0][0];
    }
}
