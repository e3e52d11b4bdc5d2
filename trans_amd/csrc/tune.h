// Launch-configuration autotuner.  The candidates change how an iteration is scheduled, never
// what it computes: every LU-SGS sweep variant visits the cells in an order that respects the
// same dependencies and performs the same operations per cell.  The sweeps contract
// multiply-adds into FMAs, and the compiler places them differently in each kernel, so the
// variants agree within rounding (the GPU tests hold one update to 1e-12), not bitwise; a
// run that must be reproducible bit for bit fixes the configuration with set_tuning.
//   - the LU-SGS sweep: the line march along each axis whose cross sections fit one
//     workgroup (i, j or k), the tile wavefront with several tile shapes and progress chunks,
//     one launch per hyperplane, or the one-workgroup hyperplane kernel;
//   - the helper streams (face kernels of the three directions concurrently) or one stream;
//   - replaying the iteration as a HIP graph or launching it eagerly.
// Each candidate runs `iters` timed iterations from the same saved state (after one untimed
// iteration that also captures its graph); the fastest is kept and the state restored.

static void apply_streams(trans_ctx* c, bool single)
{
    for (int q = 0; q < 3; q++) c->aux[q] = (single || !c->aux_own[q]) ? c->stream : c->aux_own[q];
    drop_graph(c);
}

extern "C" {

int trans_hip_autotune(trans_ctx* c, int iters, trans_tune_entry* out, int max_out, int* n_out)
{
    if (!c || iters < 1) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Fields& f = c->f;
    const long n = c->tot;
    // the state an iteration reads and writes
    struct Part {
        double* a;
        int ncomp;
    } parts[7] = {{f.V, 7}, {f.PP, 1}, {f.T, 1}, {f.Rmiu, 1}, {f.Wt0, 7}, {f.Wt1, 7}, {f.Wt2, 7}};
    size_t total = 0;
    for (const Part& p : parts) total += (size_t)p.ncomp * n;
    double* save = nullptr;
    HIPCHK(c, hipMalloc(&save, total * sizeof(double)));
    auto copy_state = [&](bool restore) -> int {
        size_t o = 0;
        for (const Part& p : parts) {
            const size_t b = (size_t)p.ncomp * n * sizeof(double);
            HIPCHK(c, hipMemcpyAsync(restore ? (void*)p.a : (void*)(save + o), restore ? (void*)(save + o) : (void*)p.a,
                                     b, hipMemcpyDeviceToDevice, c->stream));
            o += (size_t)p.ncomp * n;
        }
        return 0;
    };
    const int ki0 = c->ki;
    const bool stats0 = c->stats_on;
    const bool graphs0 = c->graphs;
    // tuning leaves the plain drop-in schedule for the batched candidates
    const bool plain0 = c->plain;
    c->plain = false;
    const int sweep0 = plain0 ? -2 : cur_sweep(c);
    const int ta0 = c->tile_a, tb0 = c->tile_b, ch0 = c->tile_chunk;
    const bool single0 = c->aux[0] == c->stream;
    c->stats_on = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = copy_state(false);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(c, TRANS_EHIP, "hipEventCreate");
    std::vector<trans_tune_entry> res;
    // tile wavefront candidates: tile shape (lines along i x along k) and progress chunk; they
    // are timed with helper streams and graph replay only
    const int tshape[4][2] = {{16, 16}, {32, 8}, {64, 4}, {128, 2}};
    const int tchunk[4] = {2, 4, 8, 16};
    struct Cand {
        int sweep, ta, tb, chunk;
    };
    std::vector<Cand> cands;
    for (int sw : {1, 2, 0, -2, -1}) cands.push_back(Cand{sw, ta0, tb0, ch0});
    for (const auto& sh : tshape)
        for (int ch : tchunk) cands.push_back(Cand{3, sh[0], sh[1], ch});
    for (size_t ci = 0; ci < cands.size() && !rc; ci++) {
        const int sw = cands[ci].sweep;
        if (sw >= 0 && sw <= 2 && !march_fits(c, sw, nullptr)) continue;
        if (c->p.if_precondition && sw != -1) continue;  // the transform lives in the hyperplane sweep
        c->tile_a = cands[ci].ta;
        c->tile_b = cands[ci].tb;
        c->tile_chunk = cands[ci].chunk;
        if ((rc = set_sweep(c, sw, true))) break;
        for (int st = 0; st < 2 && !rc; st++) {
            if (st == 0 && !c->aux_own[0]) continue;  // single-stream context
            if (sw == 3 && st == 1 && c->aux_own[0]) continue;
            apply_streams(c, st == 1);
            for (int g = 0; g < 2 && !rc; g++) {
                if (sw == 3 && g == 0) continue;
                c->graphs = g == 1;
                drop_graph(c);
                if ((rc = copy_state(true))) break;
                c->ki = ki0;
                if ((rc = trans_hip_iterate(c, 1, nullptr))) break;  // untimed: captures the graph
                if ((rc = copy_state(true))) break;
                c->ki = ki0;
                HIPCHK(c, hipEventRecord(e0, c->stream));
                if ((rc = trans_hip_iterate(c, iters, nullptr))) break;
                HIPCHK(c, hipEventRecord(e1, c->stream));
                HIPCHK(c, hipEventSynchronize(e1));
                if ((rc = check_sweep_flag(c))) break;
                float ms = 0.f;
                HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
                res.push_back(trans_tune_entry{sw, st == 0 ? 4 : 1, g, (double)ms / iters, c->tile_a, c->tile_b,
                                               c->tile_chunk});
            }
        }
    }
    // keep the fastest (the configuration before tuning if nothing ran)
    int best = -1;
    for (size_t q = 0; q < res.size(); q++)
        if (best < 0 || res[q].ms_per_iter < res[(size_t)best].ms_per_iter) best = (int)q;
    int rc2 = 0;
    if (best >= 0 && !rc) {
        c->tile_a = res[(size_t)best].tile_a;
        c->tile_b = res[(size_t)best].tile_b;
        c->tile_chunk = res[(size_t)best].chunk;
        rc2 = set_sweep(c, res[(size_t)best].sweep, true);
        apply_streams(c, res[(size_t)best].streams == 1);
        c->graphs = res[(size_t)best].graph != 0;
    } else {
        c->tile_a = ta0;
        c->tile_b = tb0;
        c->tile_chunk = ch0;
        rc2 = set_sweep(c, sweep0, true);
        apply_streams(c, single0);
        c->graphs = graphs0;
        c->plain = plain0;
    }
    if (!rc2) rc2 = copy_state(true);
    if (!rc2) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc2 = fail(c, TRANS_EHIP, hipGetErrorString(e));
    }
    drop_graph(c);
    c->ki = ki0;
    c->stats_on = stats0;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    if (rc) return rc;
    if (rc2) return rc2;
    if (out)
        for (int q = 0; q < (int)res.size() && q < max_out; q++) out[q] = res[(size_t)q];
    if (n_out) *n_out = (int)res.size();
    return best >= 0 ? 0 : fail(c, TRANS_EINVAL, "no configuration could be timed");
}

int trans_hip_set_tuning(trans_ctx* c, const trans_tune_entry* t)
{
    if (!c || !t) return TRANS_EINVAL;
    if (t->sweep < -3 || t->sweep > 3) return fail(c, TRANS_EINVAL, "sweep must be -3, -2, -1, 0, 1, 2 or 3");
    if (t->sweep == -3) {  // the plain drop-in: per-block launches, per-plane sweeps, one stream, no graph
        if (c->p.if_precondition)
            return fail(c, TRANS_EUNSUPPORTED, "the plain drop-in runs decks without preconditioning");
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int rc = set_sweep(c, -2, true);
        if (rc) return rc;
        c->plain = true;
        apply_streams(c, true);
        c->graphs = false;
        drop_graph(c);
        return 0;
    }
    c->plain = false;
    if (t->streams != 1 && t->streams != 4) return fail(c, TRANS_EINVAL, "streams must be 1 or 4");
    if (t->streams == 4 && !c->aux_own[0]) return fail(c, TRANS_EINVAL, "this context has one stream");
    if (t->sweep == 3 && (t->tile_a < 1 || t->tile_b < 1 || t->tile_a * t->tile_b > 256 || t->chunk < 1))
        return fail(c, TRANS_EINVAL, "tile shape must have 1..256 lines and chunk >= 1");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int cur = cur_sweep(c);
    const bool reshape = t->sweep == 3 && (t->tile_a != c->tile_a || t->tile_b != c->tile_b);
    if (t->sweep == 3) {
        c->tile_a = t->tile_a;
        c->tile_b = t->tile_b;
        c->tile_chunk = t->chunk;
    }
    if (t->sweep != cur || reshape) {
        int rc = set_sweep(c, t->sweep, true);
        if (rc) return rc;
    }
    apply_streams(c, t->streams == 1);
    c->graphs = t->graph != 0;
    drop_graph(c);
    return 0;
}

int trans_hip_get_tuning(const trans_ctx* c, trans_tune_entry* out)
{
    if (!c || !out) return TRANS_EINVAL;
    out->sweep = cur_sweep(c);
    out->streams = c->aux[0] == c->stream ? 1 : 4;
    out->graph = c->graphs ? 1 : 0;
    out->ms_per_iter = 0.0;
    out->tile_a = c->tile_a;
    out->tile_b = c->tile_b;
    out->chunk = c->tile_chunk;
    return 0;
}

}  // extern "C"
