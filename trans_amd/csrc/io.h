// Restart records (Output/OUTPUT_SaveField.F90, Input/Init.F90:22-46, 77-91, 146-160) and the
// time statistics of Average_Fields (Output/Average_Field.F90:3-263, :264-281 for the FLfld
// record).  Files are Fortran unformatted sequential: every record is framed by its byte
// count as a 4-byte integer before and after it (records must stay below 2 GiB).
#include <cstdint>

static bool rec_write(FILE* fp, const void* data, size_t bytes)
{
    if (bytes > 0x7fffffffu) return false;
    const int32_t m = (int32_t)bytes;
    return fwrite(&m, 4, 1, fp) == 1 && (bytes == 0 || fwrite(data, 1, bytes, fp) == bytes) && fwrite(&m, 4, 1, fp) == 1;
}

static bool rec_read(FILE* fp, void* data, size_t bytes)
{
    int32_t m = -1, t = -1;
    if (fread(&m, 4, 1, fp) != 1 || (size_t)m != bytes) return false;
    if (bytes && fread(data, 1, bytes, fp) != bytes) return false;
    return fread(&t, 4, 1, fp) == 1 && t == m;
}

static BlkInfo blk_info(const trans_ctx* c, int blk)
{
    const HostBlock& h = c->hb[blk];
    return BlkInfo{h.ni, h.nj, h.nk, h.nx, h.ny, h.nz, h.off, 0, 0, 0, 0};
}

// field (ncomp components, stride tot) of block blk <-> host vector in record order
static int pack_block(trans_ctx* c, int blk, double* field, long tot, int ncomp, std::vector<double>& host,
                      bool unpack)
{
    const size_t n = (size_t)c->hb[blk].ncell() * ncomp;
    if (!unpack) host.resize(n);
    else if (host.size() != n) return fail(c, TRANS_EINVAL, "record size mismatch");
    double* d = nullptr;
    HIPCHK(c, hipMalloc(&d, n * sizeof(double)));
    int rc = 0;
    if (unpack) rc = h2d(c, d, host.data(), n * sizeof(double));
    if (!rc) {
        hipLaunchKernelGGL(k_pack_record, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, field, tot, ncomp,
                           blk_info(c, blk), d, unpack ? 1 : 0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, TRANS_EHIP, hipGetErrorString(e));
    }
    if (!rc && !unpack) rc = d2h(c, host.data(), d, n * sizeof(double));
    if (!rc && unpack) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, TRANS_EHIP, hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

extern "C" {

int trans_hip_write_restart(trans_ctx* c, int blk, int ki, const char* path)
{
    if (!c || !path || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(c, TRANS_EINVAL, std::string("cannot open ") + path);
    const int32_t k32 = ki;
    bool ok = rec_write(fp, &k32, 4);
    std::vector<double> buf;
    int rc = 0;
    struct Item {
        double* a;
        int n;
    } items[4] = {{c->f.V, 7}, {c->f.Rmiu, 1}, {c->f.Wt0, 7}, {c->f.Wt1, 7}};
    const int nitem = c->p.kind_dual == 1 ? 3 : (c->p.kind_dual == 2 ? 4 : 2);
    for (int q = 0; q < nitem && ok && !rc; q++) {
        rc = pack_block(c, blk, items[q].a, c->tot, items[q].n, buf, false);
        if (!rc) ok = rec_write(fp, buf.data(), buf.size() * sizeof(double));
    }
    if (std::fclose(fp) != 0) ok = false;
    if (rc) return rc;
    return ok ? 0 : fail(c, TRANS_EINVAL, std::string("write failed: ") + path);
}

int trans_hip_read_restart(trans_ctx* c, int blk, const char* path, int* ki)
{
    if (!c || !path || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(c, TRANS_EINVAL, std::string("cannot open ") + path);
    int32_t k32 = 0;
    bool ok = rec_read(fp, &k32, 4);
    const size_t ncell = (size_t)c->hb[blk].ncell();
    std::vector<double> buf;
    int rc = 0;
    struct Item {
        double* a;
        int n;
    } items[4] = {{c->f.V, 7}, {c->f.Rmiu, 1}, {c->f.Wt0, 7}, {c->f.Wt1, 7}};
    const int nitem = c->p.kind_dual == 1 ? 3 : (c->p.kind_dual == 2 ? 4 : 2);
    for (int q = 0; q < nitem && ok && !rc; q++) {
        buf.resize(ncell * items[q].n);
        ok = rec_read(fp, buf.data(), buf.size() * sizeof(double));
        if (ok) rc = pack_block(c, blk, items[q].a, c->tot, items[q].n, buf, true);
    }
    std::fclose(fp);
    if (rc) return rc;
    if (!ok) return fail(c, TRANS_EINVAL, std::string("not a restart record of this block: ") + path);
    const long n = (long)ncell;
    hipLaunchKernelGGL(k_pp_t_from_v, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->p, c->f,
                       blk_info(c, blk));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ki) *ki = k32;
    return 0;
}

int trans_hip_stats_enable(trans_ctx* c, int ki_aver, int ki_statis)
{
    if (!c) return TRANS_EINVAL;
    if (ki_aver < 0) {
        c->stats_on = false;
        return 0;
    }
    if (!c->d_stats) {
        int rc = dalloc(c, &c->d_stats, (size_t)TRANS_NSTAT * c->tot);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->stats_on = true;
    c->ki_aver = ki_aver;
    c->ki_statis = ki_statis;
    return 0;
}

// Average_Fields at the context's KI (main.F90:185-188 calls it once KI >= KI_aver)
static int average_fields_async(trans_ctx* c)
{
    if (!c->stats_on || c->ki < c->ki_aver) return 0;
    const int first = c->ki == c->ki_aver;
    const int rms = c->ki >= c->ki_statis;
    KLAUNCH(c, "k_average_fields", k_average_fields, dim3(c->seg_cell.n), dim3(WG), c->p, c->f, c->d_stats, c->tot,
            c->d_blk, c->seg_cell.d, first, (double)(c->ki - c->ki_aver), rms, (double)(c->ki - c->ki_statis));
    return launch_check(c, "average_fields");
}

int trans_hip_average_fields(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    if (!c->stats_on) return fail(c, TRANS_EINVAL, "statistics are not enabled");
    return average_fields_async(c);
}

int trans_hip_get_stats(trans_ctx* c, int blk, const trans_fview* out)
{
    if (!c || !out || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    if (!c->d_stats) return fail(c, TRANS_EINVAL, "statistics are not enabled");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return copy_view(c, blk, *out, c->d_stats, TRANS_NSTAT, false);
}

int trans_hip_set_stats(trans_ctx* c, int blk, const trans_fview* in)
{
    if (!c || !in || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    if (!c->d_stats) return fail(c, TRANS_EINVAL, "statistics are not enabled");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return copy_view(c, blk, *in, c->d_stats, TRANS_NSTAT, true);
}

// Reads one component of a host view at (i, j, k) (0 outside the view)
static double view_at(const trans_fview& v, int i, int j, int k)
{
    if (!v.p || i < v.lo[0] || i > v.hi[0] || j < v.lo[1] || j > v.hi[1] || k < v.lo[2] || k > v.hi[2]) return 0.0;
    const long n0 = v.hi[0] - v.lo[0] + 1, n1 = v.hi[1] - v.lo[1] + 1;
    return v.p[(long)v.ncomp * ((i - v.lo[0]) + n0 * ((j - v.lo[1]) + n1 * (long)(k - v.lo[2])))];
}

int trans_hip_write_stats(trans_ctx* c, int blk, const trans_fview* xc, const trans_fview* yc, const trans_fview* zc,
                          const char* path)
{
    if (!c || !path || !xc || !yc || !zc || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    if (!c->d_stats) return fail(c, TRANS_EINVAL, "statistics are not enabled");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> st;
    int rc = pack_block(c, blk, c->d_stats, c->tot, TRANS_NSTAT, st, false);
    if (rc) return rc;
    const HostBlock& h = c->hb[blk];
    const int NI1 = h.ni - 1, NJ1 = h.nj - 1, NK1 = h.nk - 1;
    const size_t ncell = (size_t)h.ncell();
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(c, TRANS_EINVAL, std::string("cannot open ") + path);
    bool ok = true;
    double r[3 + TRANS_NSTAT];
    for (int k = 1; k <= NK1 && ok; k++)
        for (int j = 1; j <= NJ1 && ok; j++)
            for (int i = 1; i <= NI1 && ok; i++) {
                r[0] = view_at(*xc, i, j, k);
                r[1] = view_at(*yc, i, j, k);
                r[2] = view_at(*zc, i, j, k);
                const size_t cell = ((size_t)(i - 1) * NJ1 + (j - 1)) * NK1 + (k - 1);  // record order of st
                for (int q = 0; q < TRANS_NSTAT; q++) r[3 + q] = st[(size_t)q * ncell + cell];
                ok = rec_write(fp, r, sizeof(r));
            }
    if (std::fclose(fp) != 0) ok = false;
    return ok ? 0 : fail(c, TRANS_EINVAL, std::string("write failed: ") + path);
}

int trans_hip_read_stats(trans_ctx* c, int blk, const char* path)
{
    if (!c || !path || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    if (!c->d_stats) return fail(c, TRANS_EINVAL, "statistics are not enabled");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const HostBlock& h = c->hb[blk];
    const int NI1 = h.ni - 1, NJ1 = h.nj - 1, NK1 = h.nk - 1;
    const size_t ncell = (size_t)h.ncell();
    std::vector<double> st((size_t)TRANS_NSTAT * ncell);
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(c, TRANS_EINVAL, std::string("cannot open ") + path);
    bool ok = true;
    double r[3 + TRANS_NSTAT];
    for (int k = 1; k <= NK1 && ok; k++)
        for (int j = 1; j <= NJ1 && ok; j++)
            for (int i = 1; i <= NI1 && ok; i++) {
                ok = rec_read(fp, r, sizeof(r));
                const size_t cell = ((size_t)(i - 1) * NJ1 + (j - 1)) * NK1 + (k - 1);
                for (int q = 0; q < TRANS_NSTAT && ok; q++) st[(size_t)q * ncell + cell] = r[3 + q];
            }
    std::fclose(fp);
    if (!ok) return fail(c, TRANS_EINVAL, std::string("not a statistics record of this block: ") + path);
    return pack_block(c, blk, c->d_stats, c->tot, TRANS_NSTAT, st, true);
}

}  // extern "C"
