// Host side of the device geometry pre-processing (k_geom.h).
extern "C" {

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

}  // extern "C"
