// Tecplot ASCII output of one block's instantaneous field: the point coordinates and the
// cell-centred variables of Output/OUTPUT_InstField.F90:150-370 (its kind_field = 0 branch),
// in Tecplot's documented ASCII format (ordered zone, BLOCK data packing, the field variables
// CELLCENTERED) instead of the closed TecIO binary calls.  The variables are computed on the
// device from the resident state; the host formats them.
#pragma once

#define TEC_NCELLVAR 31

static const char* const tec_cell_names[TEC_NCELLVAR] = {
    "density", "U", "V", "W", "Uabs", "Vabs", "Wabs", "Ma", "Mabs", "Ps", "P0rel", "P0", "Ts", "T0rel", "T0", "Ox",
    "Oy", "Oz", "Om", "Q", "Yita4", "Rmiu", "k", "shock", "fr", "uu", "vv", "ww", "uv", "uw", "vw"};

// One thread per interior cell (i fastest).  out[q * ncell + n] for the variables in the
// order of tec_cell_names, with the reference's dimensional scales (density * Roref,
// velocities * Vref, pressures * Pref, temperatures * Tinf; the rest as stored) and its
// formulas, quirks included:
//   - Mabs is |c| of the absolute velocity c = (u, v - omega Zc, w + omega Yc), not divided by
//     the speed of sound (OUTPUT_InstField.F90:232-240);
//   - the absolute total pressure is Ps (1 + Mabs^2) * 3.5, a product, and the absolute total
//     temperature Ts (1 + Mabs^2) (:262-266);
//   - fr is the rotation-curvature factor f_r1, identically 1 without a rotation-curvature
//     model (Kind_RotMod = 0);
//   - the Reynolds-stress columns use rho*k (V(6)) where the comment reads k (:398-403).
// The velocity gradients are DQdxyz(1:9) as the last residual evaluation left them (zero
// before the first residual of an inviscid run).
__global__ void __launch_bounds__(256) k_field_vars(Fields f, BlkInfo b, double omega, double ro_ref, double v_ref,
                                                    double p_ref, double t_ref, double* __restrict__ out)
{
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const long ncell = (long)NI1 * NJ1 * NK1;
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ncell) return;
    const int i = (int)(n % NI1) + 1, j = (int)((n / NI1) % NJ1) + 1, k = (int)(n / ((long)NI1 * NJ1)) + 1;
    const long c = cidx(b, i, j, k);
    const long tot = f.tot;
    double r[TEC_NCELLVAR];
    const double den = f.V[c];
    const double uu = f.V[tot + c] / den, vv = f.V[2 * tot + c] / den, ww = f.V[3 * tot + c] / den;
    const double uabs = uu, vabs = vv - omega * f.Zc[c], wabs = ww + omega * f.Yc[c];
    const double pp = f.PP[c], tt = f.T[c];
    const double asonic = sqrt(1.4 * pp / den);
    const double ma = sqrt(uu * uu + vv * vv + ww * ww) / asonic;
    const double mabs = sqrt(uabs * uabs + vabs * vabs + wabs * wabs);
    const double crel = 1.0 + 0.2 * ma * ma;
    const double cabs = 1.0 + mabs * mabs;
    r[0] = den * ro_ref;
    r[1] = uu * v_ref;
    r[2] = vv * v_ref;
    r[3] = ww * v_ref;
    r[4] = uabs * v_ref;
    r[5] = vabs * v_ref;
    r[6] = wabs * v_ref;
    r[7] = ma;
    r[8] = mabs;
    r[9] = pp * p_ref;
    r[10] = pp * pow(crel, 3.5) * p_ref;
    r[11] = pp * (cabs * 3.5) * p_ref;
    r[12] = tt * t_ref;
    r[13] = tt * crel * t_ref;
    r[14] = tt * cabs * t_ref;
    double g[9];
#pragma unroll
    for (int q = 0; q < 9; q++) g[q] = f.DQ[(long)q * tot + c];
    // rows u, v, w; columns d/dx, d/dy, d/dz
    double S[3][3], W[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int e = 0; e < 3; e++) {
            S[a][e] = 0.5 * (g[3 * a + e] + g[3 * e + a]);
            W[a][e] = 0.5 * (g[3 * a + e] - g[3 * e + a]);
        }
    const double ox = W[2][1], oy = W[0][2], oz = W[1][0];
    r[15] = ox;
    r[16] = oy;
    r[17] = oz;
    r[18] = 2.0 * sqrt(ox * ox + oy * oy + oz * oz);
    // the reference's summation orders: first index fastest
    double qv = 0.0;
    for (int e = 0; e < 3; e++)
        for (int a = 0; a < 3; a++) qv = qv + W[a][e] * W[a][e] - S[a][e] * S[a][e];
    r[19] = 0.5 * qv;
    double y4 = 0.0;
    for (int h = 0; h < 3; h++)
        for (int e = 0; e < 3; e++)
            for (int a = 0; a < 3; a++) y4 = y4 + S[a][e] * W[e][h] * W[h][a];
    r[20] = y4;
    const double mut = f.Rmiu[c], rk = f.V[5 * tot + c];
    r[21] = mut;
    r[22] = rk / den;
    r[23] = f.shock[c];
    r[24] = 1.0;
    const double div = g[0] + g[4] + g[8];
    r[25] = (2.0 * g[0] - 2.0 * div / 3.0) * mut - 2.0 * rk / 3.0;
    r[26] = (2.0 * g[4] - 2.0 * div / 3.0) * mut - 2.0 * rk / 3.0;
    r[27] = (2.0 * g[8] - 2.0 * div / 3.0) * mut - 2.0 * rk / 3.0;
    r[28] = (g[1] + g[3]) * mut;
    r[29] = (g[2] + g[6]) * mut;
    r[30] = (g[5] + g[7]) * mut;
#pragma unroll
    for (int q = 0; q < TEC_NCELLVAR; q++) out[(long)q * ncell + n] = r[q];
}

// Writes `count` values, five to a line, as %.10e (11 significant digits).
static bool tec_values(FILE* fp, const double* v, size_t count)
{
    for (size_t q = 0; q < count; q++)
        if (std::fprintf(fp, (q % 5 == 4 || q + 1 == count) ? "%.10e\n" : "%.10e ", v[q]) < 0) return false;
    return true;
}

extern "C" {

int trans_tecplot_write_ascii(const char* path, const char* title, const char* zone, int ni, int nj, int nk,
                              const double* x, const double* y, const double* z, int ncellvar,
                              const char* const* cellvar_names, const double* cellvars)
{
    if (!path || !x || !y || !z || ni < 2 || nj < 2 || nk < 2 || ncellvar < 0 || (ncellvar > 0 && !cellvars) ||
        (ncellvar > 0 && !cellvar_names))
        return TRANS_EINVAL;
    for (int q = 0; q < ncellvar; q++)
        if (!cellvar_names[q] || std::strchr(cellvar_names[q], '"')) return TRANS_EINVAL;
    if ((title && std::strchr(title, '"')) || (zone && std::strchr(zone, '"'))) return TRANS_EINVAL;
    FILE* fp = std::fopen(path, "w");
    if (!fp) return TRANS_EINVAL;
    bool ok = std::fprintf(fp, "TITLE = \"%s\"\nVARIABLES = \"X\" \"Y\" \"Z\"", title ? title : "") > 0;
    for (int q = 0; q < ncellvar && ok; q++) ok = std::fprintf(fp, " \"%s\"", cellvar_names[q]) > 0;
    ok = ok && std::fprintf(fp, "\nZONE T=\"%s\", I=%d, J=%d, K=%d, ZONETYPE=ORDERED, DATAPACKING=BLOCK", zone ? zone : "",
                            ni, nj, nk) > 0;
    if (ok && ncellvar > 0)
        ok = (ncellvar == 1 ? std::fprintf(fp, ", VARLOCATION=([4]=CELLCENTERED)")
                            : std::fprintf(fp, ", VARLOCATION=([4-%d]=CELLCENTERED)", 3 + ncellvar)) > 0;
    ok = ok && std::fprintf(fp, "\n") > 0;
    const size_t np = (size_t)ni * nj * nk, nc = (size_t)(ni - 1) * (nj - 1) * (nk - 1);
    ok = ok && tec_values(fp, x, np) && tec_values(fp, y, np) && tec_values(fp, z, np);
    for (int q = 0; q < ncellvar && ok; q++) ok = tec_values(fp, cellvars + (size_t)q * nc, nc);
    if (std::fclose(fp) != 0) ok = false;
    return ok ? 0 : TRANS_EINVAL;
}

int trans_hip_write_tecplot(trans_ctx* c, int blk, const trans_fview* x, const trans_fview* y, const trans_fview* z,
                            const double* scales, const char* path)
{
    if (!c || !x || !y || !z || !path || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    const HostBlock& h = c->hb[blk];
    for (const trans_fview* v : {x, y, z})
        if (!v->p || v->ncomp != 1 || v->lo[0] > 1 || v->lo[1] > 1 || v->lo[2] > 1 || v->hi[0] < h.ni ||
            v->hi[1] < h.nj || v->hi[2] < h.nk)
            return fail(c, TRANS_EINVAL, "point coordinate views must cover 1..NI, 1..NJ, 1..NK");
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    const double* s = scales ? scales : one;
    const long ncell = h.ncell();
    const size_t np = (size_t)h.ni * h.nj * h.nk;
    std::vector<double> cv((size_t)TEC_NCELLVAR * ncell), pts(3 * np);
    double* d = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMalloc(&d, cv.size() * sizeof(double)));
    hipLaunchKernelGGL(k_field_vars, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, c->stream, c->f,
                       blk_info(c, blk), c->p.omega1, s[0], s[1], s[2], s[3], d);
    hipError_t e = hipGetLastError();
    int rc = e != hipSuccess ? fail(c, TRANS_EHIP, hipGetErrorString(e)) : d2h(c, cv.data(), d, cv.size() * sizeof(double));
    (void)hipFree(d);
    if (rc) return rc;
    for (int q = 0; q < 3; q++) {
        const trans_fview* v = q == 0 ? x : (q == 1 ? y : z);
        size_t o = (size_t)q * np;
        for (int k = 1; k <= h.nk; k++)
            for (int j = 1; j <= h.nj; j++)
                for (int i = 1; i <= h.ni; i++) pts[o++] = view_at(*v, i, j, k);
    }
    char zone[32];
    std::snprintf(zone, sizeof(zone), "block %d", blk + 1);
    rc = trans_tecplot_write_ascii(path, "TRANS instantaneous field", zone, h.ni, h.nj, h.nk, pts.data(), pts.data() + np,
                                   pts.data() + 2 * np, TEC_NCELLVAR, tec_cell_names, cv.data());
    return rc ? fail(c, TRANS_EINVAL, std::string("cannot write ") + path) : 0;
}

}  // extern "C"
