// Kernels of the restart records and the time statistics: interior cells of one block
// packed into or out of the reference's record order, the state's p and T recomputed after a
// restart read, and the running mean / RMS fields of Average_Fields.
#pragma once
#include "dev_common.h"

// Record order of OUTPUT_SaveFields (Output/OUTPUT_SaveField.F90:19-25): component
// outermost, then I, J, with K fastest.  One thread per (component, cell), K fastest, so the
// record side is written / read coalesced.
__global__ void __launch_bounds__(256) k_pack_record(double* __restrict__ field, long tot, int ncomp, BlkInfo b,
                                                     double* __restrict__ rec, int unpack)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const long ncell = (long)NI1 * NJ1 * NK1;
    const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= ncell * ncomp) return;
    const int L = (int)(lin / ncell);
    long r = lin - (long)L * ncell;
    const int k = (int)(r % NK1) + 1;
    r /= NK1;
    const int j = (int)(r % NJ1) + 1;
    const int i = (int)(r / NJ1) + 1;
    const long c = cidx(b, i, j, k);
    if (unpack)
        field[(long)L * tot + c] = rec[lin];
    else
        rec[lin] = field[(long)L * tot + c];
}

// p and T from the conservative state after a restart read (Input/Init.F90:146-160):
// p = (1.4-1) rho E - 0.5 (1.4-1) (|rho u|^2 / rho - (omega r)^2 rho), T = gamma M^2 p / rho
__global__ void __launch_bounds__(256) k_pp_t_from_v(trans_params p, Fields f, BlkInfo b)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k), tot = f.tot;
    const double v1 = f.V[c], v2 = f.V[tot + c], v3 = f.V[2 * tot + c], v4 = f.V[3 * tot + c];
    const double VV = v2 * v2 + v3 * v3 + v4 * v4;
    const double wr = p.omega1 * f.rad[c];
    const double pp = (1.4 - 1.0) * f.V[4 * tot + c] - 0.5 * (1.4 - 1.0) * (VV / v1 + -(wr * wr) * v1);
    f.PP[c] = pp;
    f.T[c] = p.rxm2 * pp / v1;
}

// Average_Fields (Output/Average_Field.F90:46-263) for one block.  S has TRANS_NSTAT
// components in the reference's FLfld record order: U V W R M Mu k P0 P T uu vv ww uv uw vw
// (means; the last six are the modelled Reynolds stresses) then uurms vvrms wwrms uvrms vwrms
// uwrms pprms.  first: KI == KI_aver (means := instantaneous values); otherwise the running
// mean with weight tm = KI - KI_aver.  rms: KI >= KI_statis, weight ts = KI - KI_statis, about
// the means just updated.
__global__ void __launch_bounds__(256) k_average_fields(trans_params p, Fields f, double* __restrict__ S,
                                                        long stot, const BlkInfo* __restrict__ blks,
                                                        const Seg* __restrict__ seg, int first, double tm,
                                                        int rms, double ts)
{
    const Seg s = seg[blockIdx.x];
    const BlkInfo b = blks[s.blk];
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int lin = s.base + threadIdx.x;
    if (lin >= NI1 * NJ1 * NK1) return;
    int i, j, k;
    decode3(lin, NI1, NJ1, i, j, k);
    const long c = cidx(b, i, j, k), tot = f.tot;
    const double Den = f.V[c];
    const double RmiuT = f.Rmiu[c];
    const double vK = f.V[5 * tot + c] / Den;
    double v[16];
    v[0] = f.V[tot + c] / Den;
    v[1] = f.V[2 * tot + c] / Den;
    v[2] = f.V[3 * tot + c] / Den;
    v[3] = Den;
    const double QQ = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double AA = sqrt(1.4 * f.PP[c] / Den);
    v[4] = QQ / AA;
    v[5] = RmiuT;
    v[6] = f.V[5 * tot + c] / Den;
    v[7] = f.PP[c] * pow(1.0 + 0.2 * v[4] * v[4], 3.5) / p.ppf;
    v[8] = f.PP[c];
    v[9] = f.T[c];
    const double* Q = f.DQ;
    const double dudx = Q[c], dudy = Q[tot + c], dudz = Q[2 * tot + c];
    const double dvdx = Q[3 * tot + c], dvdy = Q[4 * tot + c], dvdz = Q[5 * tot + c];
    const double dwdx = Q[6 * tot + c], dwdy = Q[7 * tot + c], dwdz = Q[8 * tot + c];
    const double S0 = -2.0 / 3.0 * (dudx + dvdy + dwdz);
    const double Ref = p.ref;
    v[10] = -(RmiuT / Ref * (S0 + 2.0 * dudx) - 2.0 / 3.0 * Den * vK);
    v[11] = -(RmiuT / Ref * (S0 + 2.0 * dvdy) - 2.0 / 3.0 * Den * vK);
    v[12] = -(RmiuT / Ref * (S0 + 2.0 * dwdz) - 2.0 / 3.0 * Den * vK);
    v[13] = -(RmiuT / Ref * (dudy + dvdx));
    v[14] = -(RmiuT / Ref * (dudz + dwdx));
    v[15] = -(RmiuT / Ref * (dvdz + dwdy));
    double m[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        double* a = S + (long)q * stot + c;
        m[q] = first ? v[q] : (*a * tm + v[q]) / (tm + 1.);
        *a = m[q];
    }
    if (!rms) return;
    const double du = v[0] - m[0], dv = v[1] - m[1], dw = v[2] - m[2], dp = v[8] - m[8];
    const double dev[7] = {du * du, dv * dv, dw * dw, du * dv, dv * dw, du * dw, dp * dp};
#pragma unroll
    for (int q = 0; q < 7; q++) {
        double* a = S + (long)(16 + q) * stot + c;
        *a = (*a * ts + dev[q]) / (ts + 1.);
    }
}
