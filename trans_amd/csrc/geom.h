// Host side of the device geometry pre-processing (k_geom.h).
#include <mutex>

// Staging memory of trans_geom_wall_distance_static: the cell centres (3 ncell), the panel
// centres (3 npanel) and the squared minima (ncell) of one call, in that order.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code never passes the
// name `scratch` to a kernel or a copy (on the host it is only a stand-in for the symbol):
// it uses the per-device address from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[TRANS_WD_STATIC_DOUBLES];

namespace {
constexpr int WD_MAX_DEVICES = 64;
std::mutex wd_scratch_mu;
double* wd_scratch_ptr[WD_MAX_DEVICES] = {};

// Address of the current device's copy of `scratch`, looked up once per device with that
// device current.  nullptr on failure.
double* wd_scratch_for_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WD_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(wd_scratch_mu);
    if (!wd_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        wd_scratch_ptr[dev] = (double*)p;
    }
    return wd_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int trans_geom_wall_distance_static(const double* cells, long ncell, const double* panels, long npanel, double* dst)
{
    if (!cells || !dst || ncell < 0 || npanel < 0 || (npanel > 0 && !panels)) return TRANS_EINVAL;
    // 4 ncell + 3 npanel doubles must fit the static staging array
    if (ncell > TRANS_WD_STATIC_DOUBLES || npanel > TRANS_WD_STATIC_DOUBLES ||
        4 * ncell + 3 * npanel > TRANS_WD_STATIC_DOUBLES)
        return TRANS_ESCRATCH;
    if (ncell == 0) return 0;
    if (npanel == 0) {  // no wall anywhere: the distance is never used
        for (long c = 0; c < ncell; c++) dst[c] = 1.0e3;
        return 0;
    }
    double* base = wd_scratch_for_current_device();
    if (!base) return TRANS_EHIP;
    double* dc = base;                // 3 ncell
    double* dp = base + 3 * ncell;    // 3 npanel
    double* dd = dp + 3 * npanel;     // ncell, every entry written by the kernel
    // the null stream: no stream is created (creating one allocates)
    const hipStream_t s = 0;
    if (hipMemcpyAsync(dc, cells, 3 * ncell * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dp, panels, 3 * npanel * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        return TRANS_EHIP;
    const long per = 256L * WD_CPT;
    hipLaunchKernelGGL(k_wall_distance, dim3((unsigned)((ncell + per - 1) / per)), dim3(256), 0, s, dc, ncell, dp,
                       npanel, dd);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(dst, dd, ncell * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return TRANS_EHIP;
    for (long c = 0; c < ncell; c++) dst[c] = std::sqrt(std::max(dst[c], 1.0e-20));
    return 0;
}

int trans_geom_wall_distance(const double* cells, long ncell, const double* panels, long npanel, double* dst)
{
    if (!cells || !dst || ncell < 0 || npanel < 0 || (npanel > 0 && !panels)) return TRANS_EINVAL;
    if (ncell == 0) return 0;
    if (npanel == 0) {  // no wall anywhere: the distance is never used
        for (long c = 0; c < ncell; c++) dst[c] = 1.0e3;
        return 0;
    }
    hipStream_t s = nullptr;
    double *dc = nullptr, *dp = nullptr, *dd = nullptr;
    int rc = 0;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && !rc) rc = TRANS_EHIP;
        return e == hipSuccess;
    };
    if (ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && ok(hipMalloc(&dc, 3 * ncell * sizeof(double))) &&
        ok(hipMalloc(&dp, 3 * npanel * sizeof(double))) && ok(hipMalloc(&dd, ncell * sizeof(double))) &&
        ok(hipMemcpyAsync(dc, cells, 3 * ncell * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(dp, panels, 3 * npanel * sizeof(double), hipMemcpyHostToDevice, s))) {
        const long per = 256L * WD_CPT;
        hipLaunchKernelGGL(k_wall_distance, dim3((unsigned)((ncell + per - 1) / per)), dim3(256), 0, s, dc, ncell, dp,
                           npanel, dd);
        if (ok(hipGetLastError()) && ok(hipMemcpyAsync(dst, dd, ncell * sizeof(double), hipMemcpyDeviceToHost, s)))
            ok(hipStreamSynchronize(s));
    }
    if (dc) (void)hipFree(dc);
    if (dp) (void)hipFree(dp);
    if (dd) (void)hipFree(dd);
    if (s) (void)hipStreamDestroy(s);
    if (rc) return rc;
    // sqrt of the clamped minimum (the reference clamps with tiny = 1e-20 between surfaces)
    for (long c = 0; c < ncell; c++) dst[c] = std::sqrt(std::max(dst[c], 1.0e-20));
    return 0;
}

int trans_geom_metrics(int ni, int nj, int nk, const double* xx, const double* yy, const double* zz, double tiny,
                       const trans_geom_metrics_out* out)
{
    if (ni < 3 || nj < 3 || nk < 3 || !xx || !yy || !zz || !out) return TRANS_EINVAL;
    double* const host[11] = {out->xc,   out->yc,    out->zc, out->thtc, out->rad, out->aleng,
                              out->aseng, out->vol, out->sd,  out->grad, out->alagm};
    const int ncomp[11] = {1, 1, 1, 1, 1, 1, 1, 1, 9, 3, 3};
    for (double* h : host)
        if (!h) return TRANS_EINVAL;
    const GeomGrid g{ni, nj, nk, (long)ni + 5, (long)nj + 5};
    const size_t n = (size_t)(ni + 5) * (nj + 5) * (nk + 5);
    hipStream_t s = nullptr;
    double* dx = nullptr;  // points (3n), then the outputs in the order of `host`
    int rc = 0;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && !rc) rc = TRANS_EHIP;
        return e == hipSuccess;
    };
    size_t total = 3 * n;
    for (int q = 0; q < 11; q++) total += (size_t)ncomp[q] * n;
    if (ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && ok(hipMalloc(&dx, total * sizeof(double)))) {
        double* dev[11];
        size_t off = 3 * n;
        for (int q = 0; q < 11; q++) {
            dev[q] = dx + off;
            off += (size_t)ncomp[q] * n;
        }
        // outputs start as the caller's arrays, so entries no routine sets keep their values
        bool up = ok(hipMemcpyAsync(dx, xx, n * sizeof(double), hipMemcpyHostToDevice, s)) &&
                  ok(hipMemcpyAsync(dx + n, yy, n * sizeof(double), hipMemcpyHostToDevice, s)) &&
                  ok(hipMemcpyAsync(dx + 2 * n, zz, n * sizeof(double), hipMemcpyHostToDevice, s));
        for (int q = 0; q < 11 && up; q++)
            up = ok(hipMemcpyAsync(dev[q], host[q], (size_t)ncomp[q] * n * sizeof(double), hipMemcpyHostToDevice, s));
        if (up) {
            const GeomOut o{dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6], dev[7], dev[8], dev[9], dev[10]};
            const double *X = dx, *Y = dx + n, *Z = dx + 2 * n;
            auto blocks = [](long m) { return dim3((unsigned)((m + 255) / 256)); };
            hipLaunchKernelGGL(k_geom_centres, blocks((long)(ni + 1) * (nj + 1) * (nk + 1)), dim3(256), 0, s, g, X, Y,
                               Z, o);
            hipLaunchKernelGGL(k_geom_faces<1>, blocks((long)ni * (nj - 1) * (nk - 1)), dim3(256), 0, s, g, X, Y, Z, o,
                               tiny);
            hipLaunchKernelGGL(k_geom_faces<2>, blocks((long)(ni - 1) * nj * (nk - 1)), dim3(256), 0, s, g, X, Y, Z, o,
                               tiny);
            hipLaunchKernelGGL(k_geom_faces<3>, blocks((long)(ni - 1) * (nj - 1) * nk), dim3(256), 0, s, g, X, Y, Z, o,
                               tiny);
            hipLaunchKernelGGL(k_geom_cells, blocks((long)(ni - 1) * (nj - 1) * (nk - 1)), dim3(256), 0, s, g, X, Y, Z,
                               o);
            hipLaunchKernelGGL(k_geom_ghosts, blocks((long)(ni + 2) * (nj + 2) * (nk + 2)), dim3(256), 0, s, g, o);
            bool down = ok(hipGetLastError());
            for (int q = 0; q < 11 && down; q++)
                down = ok(
                    hipMemcpyAsync(host[q], dev[q], (size_t)ncomp[q] * n * sizeof(double), hipMemcpyDeviceToHost, s));
            if (down) ok(hipStreamSynchronize(s));
        }
    }
    if (dx) (void)hipFree(dx);
    if (s) (void)hipStreamDestroy(s);
    if (rc) return rc;
    // Geom/CalcDst.F90:481-515 stops on a non-positive volume
    for (int k = 1; k <= nk - 1; k++)
        for (int j = 1; j <= nj - 1; j++)
            for (int i = 1; i <= ni - 1; i++)
                if (out->vol[g.at(i, j, k)] < tiny) return TRANS_EINVAL;
    return 0;
}

}  // extern "C"
