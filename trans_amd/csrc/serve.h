// Serving: continuous batching of independent steady solves on one device.
//
// Each admitted request owns a context (its state stays resident in HBM) and advances in
// chunks of iterations queued on that context's one non-blocking stream, so the chunks of
// all admitted requests overlap on the device: streams are spread over the hardware queues
// in creation order, and one stream per solve puts different solves on different queues
// (GPU_MAX_HW_QUEUES, 4 by default, caps how many run side by side).  Requests wait in arrival order and are
// admitted first-fit while their footprint (trans_hip_footprint) fits the budget left by the
// resident ones and a slot is free; a finished request leaves the device at once (its final
// state is kept on the host), which admits the next ones.  The reference runs one case per
// process and has no counterpart of this layer.
//
// The global residual of a request is DRSa = sqrt(sum_b N_b DRSa_b^2 / sum_b N_b) over its
// blocks (N_b cells), the drop is DRSa / DRSa(first iteration).  Every iteration's per-block
// norms are copied into pinned host memory, so the iteration that first reaches the target
// is known exactly although convergence is checked once per chunk.  A request's iterations
// depend only on its own inputs and on the chunk length, so its result is the same alone and
// in any mix.
#include <chrono>
#include <cmath>

namespace {

struct SrvReq {
    int id = 0;
    int state = TRANS_SRV_QUEUED;
    int error = 0;
    trans_params p{};
    std::vector<trans_block_desc> desc;
    std::vector<std::vector<trans_patch>> patches;
    // host copies of the submitted views (geometry until admission, state until released)
    std::vector<std::vector<double>> geom_buf, state_buf;
    std::vector<std::vector<int>> mw_buf;
    std::vector<trans_geom_view> geom;
    std::vector<trans_state_view> st;
    std::vector<double> ncell;
    size_t footprint = 0;
    trans_ctx* ctx = nullptr;
    double target = 0.0;
    int max_iter = 0;
    int iters = 0, iters_target = 0;
    double drsa0 = 0.0, drsa_last = 0.0;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int inflight = 0;                 // iterations of the queued chunk
    trans_resid* ring = nullptr;      // pinned: chunk x nblk norms of the queued chunk
    hipEvent_t ev = nullptr;          // end of the queued chunk
    long enq_seq = 0;                 // order in which chunks were queued
};

}  // namespace

struct trans_srv {
    trans_srv_config cfg{};
    std::vector<SrvReq*> reqs;  // by id - 1
    size_t used = 0;
    int active = 0;
    long seq = 0;
    std::chrono::steady_clock::time_point t0;
    std::string err;
};

static int srv_fail(trans_srv* s, int code, const std::string& msg)
{
    if (s) s->err = msg;
    return code;
}

static double srv_now(const trans_srv* s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

// Dense host copy of a view (component fastest, then i, j, k), re-pointed at the copy.
static void stage_view(trans_fview& v, std::vector<double>& buf)
{
    if (!v.p) return;
    size_t n = (size_t)v.ncomp;
    for (int d = 0; d < 3; d++) n *= (size_t)std::max(0, v.hi[d] - v.lo[d] + 1);
    buf.assign(v.p, v.p + n);
    v.p = buf.data();
}

// Copies the overlap of two views with the same component count.
static int copy_between(const trans_fview& src, const trans_fview& dst)
{
    if (!src.p || !dst.p) return 0;
    if (src.ncomp != dst.ncomp) return TRANS_EINVAL;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = std::max(src.lo[d], dst.lo[d]);
        hi[d] = std::min(src.hi[d], dst.hi[d]);
        if (hi[d] < lo[d]) return 0;
    }
    const long sn0 = src.hi[0] - src.lo[0] + 1, sn1 = src.hi[1] - src.lo[1] + 1;
    const long dn0 = dst.hi[0] - dst.lo[0] + 1, dn1 = dst.hi[1] - dst.lo[1] + 1;
    const int nc = src.ncomp;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                const long si = (long)nc * ((i - src.lo[0]) + sn0 * ((j - src.lo[1]) + sn1 * (long)(k - src.lo[2])));
                const long di = (long)nc * ((i - dst.lo[0]) + dn0 * ((j - dst.lo[1]) + dn1 * (long)(k - dst.lo[2])));
                for (int c = 0; c < nc; c++) dst.p[di + c] = src.p[si + c];
            }
    return 0;
}

static void srv_free_device(trans_srv* s, SrvReq* r)
{
    if (r->ctx) {
        trans_hip_destroy(r->ctx);
        r->ctx = nullptr;
        s->used -= r->footprint;
        s->active -= 1;
    }
    if (r->ring) (void)hipHostFree(r->ring);
    r->ring = nullptr;
    if (r->ev) (void)hipEventDestroy(r->ev);
    r->ev = nullptr;
    r->inflight = 0;
}

// Final state to the host copies, then the device is given back.
static void srv_retire(trans_srv* s, SrvReq* r, int state, int error)
{
    if (r->ctx && error == 0)
        for (size_t b = 0; b < r->st.size(); b++) {
            const int rc = trans_hip_get_state(r->ctx, (int)b, &r->st[b]);
            if (rc) {
                state = TRANS_SRV_FAILED;
                error = rc;
                break;
            }
        }
    if (state == TRANS_SRV_FAILED && r->ctx) s->err = trans_hip_last_error(r->ctx);
    srv_free_device(s, r);
    r->state = state;
    r->error = error;
    r->t_done = srv_now(s);
}

static int srv_admit(trans_srv* s, SrvReq* r)
{
    trans_ctx* c = nullptr;
    int rc = create_ctx(&r->p, (int)r->desc.size(), r->desc.data(), &c, false, true);
    r->ctx = c;
    if (!rc)
        for (size_t b = 0; b < r->desc.size() && !rc; b++) {
            rc = trans_hip_set_geometry(c, (int)b, &r->geom[b]);
            if (!rc) rc = trans_hip_set_state(c, (int)b, &r->st[b]);
        }
    if (!rc) rc = trans_hip_set_iteration(c, 1);
    const size_t nring = (size_t)s->cfg.chunk * r->desc.size();
    if (!rc && hipHostMalloc((void**)&r->ring, nring * sizeof(trans_resid), hipHostMallocDefault) != hipSuccess)
        rc = TRANS_ENOMEM;
    if (!rc && hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess) rc = TRANS_EHIP;
    s->used += r->footprint;
    s->active += 1;
    r->t_admit = srv_now(s);
    if (rc) {
        s->err = c ? trans_hip_last_error(c) : "context creation failed";
        srv_free_device(s, r);
        r->state = TRANS_SRV_FAILED;
        r->error = rc;
        r->t_done = r->t_admit;
        return rc;
    }
    r->state = TRANS_SRV_RUNNING;
    r->geom_buf.clear();  // the geometry now lives on the device
    r->mw_buf.clear();
    r->geom.clear();
    return 0;
}

// Queues the next chunk: iterations, each followed by an asynchronous copy of its norms.
static int srv_enqueue(trans_srv* s, SrvReq* r)
{
    trans_ctx* c = r->ctx;
    const int nblk = (int)r->desc.size();
    const int n = std::min(s->cfg.chunk, r->max_iter - r->iters);
    for (int it = 0; it < n; it++) {
        int rc = trans_hip_iterate(c, 1, nullptr);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(r->ring + (size_t)it * nblk, c->d_resid, nblk * sizeof(trans_resid),
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipEventRecord(r->ev, c->stream));
    r->inflight = n;
    r->enq_seq = s->seq++;
    return 0;
}

// Reads a finished chunk's norms; returns true when the request is done.
static bool srv_collect(trans_srv* s, SrvReq* r)
{
    const int nblk = (int)r->desc.size();
    double ntot = 0.0;
    for (int b = 0; b < nblk; b++) ntot += r->ncell[b];
    for (int it = 0; it < r->inflight; it++) {
        const trans_resid* q = r->ring + (size_t)it * nblk;
        double acc = 0.0;
        for (int b = 0; b < nblk; b++) acc += r->ncell[b] * q[b].drsa * q[b].drsa;
        const double g = std::sqrt(acc / ntot);
        const int n = r->iters + it + 1;
        if (!(g == g)) {
            r->iters = n;
            s->err = "request " + std::to_string(r->id) + ": residual is NaN at iteration " + std::to_string(n);
            srv_retire(s, r, TRANS_SRV_FAILED, TRANS_ENONFINITE);
            return true;
        }
        if (n == 1) r->drsa0 = g;
        r->drsa_last = g;
        if (!r->iters_target && g <= r->target * r->drsa0) r->iters_target = n;
    }
    r->iters += r->inflight;
    r->inflight = 0;
    if (r->iters_target) {
        srv_retire(s, r, TRANS_SRV_CONVERGED, 0);
        return true;
    }
    if (r->iters >= r->max_iter) {
        srv_retire(s, r, TRANS_SRV_MAX_ITER, 0);
        return true;
    }
    return false;
}

static SrvReq* srv_get(trans_srv* s, int id)
{
    if (!s || id < 1 || id > (int)s->reqs.size()) return nullptr;
    return s->reqs[id - 1];
}

extern "C" {

int trans_srv_create(const trans_srv_config* cfg, trans_srv** out)
{
    if (!cfg || !out) return TRANS_EINVAL;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->chunk < 1 || cfg->hbm_budget == 0) return TRANS_EINVAL;
    trans_srv* s = new trans_srv();
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return 0;
}

const char* trans_srv_last_error(const trans_srv* s) { return s ? s->err.c_str() : "null server"; }

int trans_srv_submit(trans_srv* s, const trans_srv_request* q, int* id)
{
    if (!s || !q || !id || q->nblk <= 0 || !q->blocks || !q->geom || !q->state) return TRANS_EINVAL;
    if (!(q->target_drop > 0.0) || q->max_iter < 1) return srv_fail(s, TRANS_EINVAL, "target_drop > 0 and max_iter >= 1");
    size_t fp = 0;
    int rc = trans_hip_footprint(&q->params, q->nblk, q->blocks, &fp);
    if (rc) return srv_fail(s, rc, "request rejected by the solver's checks");
    if (fp > s->cfg.hbm_budget)
        return srv_fail(s, TRANS_ENOMEM, "request needs " + std::to_string(fp) + " device bytes, budget is " +
                                             std::to_string(s->cfg.hbm_budget));
    SrvReq* r = new SrvReq();
    r->id = (int)s->reqs.size() + 1;
    r->p = q->params;
    r->footprint = fp;
    r->target = q->target_drop;
    r->max_iter = q->max_iter;
    const int nblk = q->nblk;
    r->desc.assign(q->blocks, q->blocks + nblk);
    r->patches.resize(nblk);
    r->geom.assign(q->geom, q->geom + nblk);
    r->st.assign(q->state, q->state + nblk);
    r->geom_buf.resize((size_t)nblk * 11);
    r->state_buf.resize((size_t)nblk * 7);
    r->mw_buf.resize((size_t)nblk * 6);
    for (int b = 0; b < nblk; b++) {
        trans_block_desc& d = r->desc[b];
        r->patches[b].assign(d.patches, d.patches + d.npatch);
        d.patches = r->patches[b].data();
        r->ncell.push_back((double)(d.ni - 1) * (d.nj - 1) * (d.nk - 1));
        trans_geom_view& g = r->geom[b];
        trans_fview* gv[11] = {&g.sd,    &g.grad,  &g.alagm, &g.vol, &g.rad, &g.dst,
                               &g.aleng, &g.aseng, &g.yc,    &g.zc,  &g.thtc};
        for (int k = 0; k < 11; k++) stage_view(*gv[k], r->geom_buf[(size_t)b * 11 + k]);
        const size_t fsz[6] = {(size_t)(d.nj - 1) * (d.nk - 1), (size_t)(d.nj - 1) * (d.nk - 1),
                               (size_t)(d.ni - 1) * (d.nk - 1), (size_t)(d.ni - 1) * (d.nk - 1),
                               (size_t)(d.ni - 1) * (d.nj - 1), (size_t)(d.ni - 1) * (d.nj - 1)};
        for (int f = 0; f < 6; f++)
            if (g.markwall[f]) {
                std::vector<int>& m = r->mw_buf[(size_t)b * 6 + f];
                m.assign(g.markwall[f], g.markwall[f] + fsz[f]);
                g.markwall[f] = m.data();
            }
        trans_state_view& v = r->st[b];
        trans_fview* sv[7] = {&v.v, &v.pp, &v.t, &v.rmiu, &v.wt0, &v.wt1, &v.wt2};
        for (int k = 0; k < 7; k++) stage_view(*sv[k], r->state_buf[(size_t)b * 7 + k]);
    }
    r->t_submit = srv_now(s);
    s->reqs.push_back(r);
    *id = r->id;
    return 0;
}

int trans_srv_progress(trans_srv* s, int wait, int* n_unfinished)
{
    if (!s) return TRANS_EINVAL;
    for (int pass = 0; pass < 2; pass++) {
        // 1. finished chunks
        for (SrvReq* r : s->reqs) {
            if (r->state != TRANS_SRV_RUNNING || r->inflight == 0) continue;
            const hipError_t q = hipEventQuery(r->ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                s->err = std::string("request ") + std::to_string(r->id) + ": " + hipGetErrorString(q);
                srv_retire(s, r, TRANS_SRV_FAILED, TRANS_EHIP);
                continue;
            }
            srv_collect(s, r);
        }
        // 2. admissions, in arrival order, first fit
        for (SrvReq* r : s->reqs) {
            if (r->state != TRANS_SRV_QUEUED) continue;
            if (s->active >= s->cfg.max_active) break;
            if (s->used + r->footprint > s->cfg.hbm_budget) continue;
            srv_admit(s, r);
        }
        // 3. next chunks
        for (SrvReq* r : s->reqs) {
            if (r->state != TRANS_SRV_RUNNING || r->inflight != 0) continue;
            const int rc = srv_enqueue(s, r);
            if (rc) {
                s->err = trans_hip_last_error(r->ctx);
                srv_retire(s, r, TRANS_SRV_FAILED, rc);
            }
        }
        // 4. optionally block until the oldest queued chunk ends, then go round once more
        if (!wait || pass == 1) break;
        SrvReq* oldest = nullptr;
        for (SrvReq* r : s->reqs)
            if (r->state == TRANS_SRV_RUNNING && r->inflight && (!oldest || r->enq_seq < oldest->enq_seq)) oldest = r;
        if (!oldest) break;
        (void)hipEventSynchronize(oldest->ev);
    }
    if (n_unfinished) {
        int n = 0;
        for (SrvReq* r : s->reqs) n += r->state == TRANS_SRV_QUEUED || r->state == TRANS_SRV_RUNNING;
        *n_unfinished = n;
    }
    return 0;
}

int trans_srv_poll(trans_srv* s, int id, trans_srv_status* out)
{
    SrvReq* r = srv_get(s, id);
    if (!r || !out) return TRANS_EINVAL;
    out->state = r->state;
    out->error = r->error;
    out->iterations = r->iters;
    out->iterations_to_target = r->iters_target;
    out->drsa_first = r->drsa0;
    out->drsa_last = r->drsa_last;
    out->t_submit = r->t_submit;
    out->t_admit = r->t_admit;
    out->t_done = r->t_done;
    out->footprint = r->footprint;
    return 0;
}

int trans_srv_cancel(trans_srv* s, int id)
{
    SrvReq* r = srv_get(s, id);
    if (!r) return TRANS_EINVAL;
    if (r->state == TRANS_SRV_QUEUED) {
        r->state = TRANS_SRV_CANCELLED;
        r->t_done = srv_now(s);
    } else if (r->state == TRANS_SRV_RUNNING) {
        if (r->ctx) (void)hipStreamSynchronize(r->ctx->stream);
        if (r->inflight) srv_collect(s, r);  // norms of the chunk that ran
        if (r->state == TRANS_SRV_RUNNING) srv_retire(s, r, TRANS_SRV_CANCELLED, 0);
    }
    return 0;
}

int trans_srv_result(trans_srv* s, int id, int blk, const trans_state_view* out)
{
    SrvReq* r = srv_get(s, id);
    if (!r || !out || blk < 0 || blk >= (int)r->desc.size()) return TRANS_EINVAL;
    if (r->state == TRANS_SRV_QUEUED || r->state == TRANS_SRV_RUNNING)
        return srv_fail(s, TRANS_EINVAL, "request " + std::to_string(id) + " has not finished");
    if (r->state_buf.empty()) return srv_fail(s, TRANS_EINVAL, "request " + std::to_string(id) + " was released");
    const trans_state_view& a = r->st[blk];
    const trans_fview* src[7] = {&a.v, &a.pp, &a.t, &a.rmiu, &a.wt0, &a.wt1, &a.wt2};
    const trans_fview* dst[7] = {&out->v, &out->pp, &out->t, &out->rmiu, &out->wt0, &out->wt1, &out->wt2};
    for (int k = 0; k < 7; k++)
        if (copy_between(*src[k], *dst[k])) return srv_fail(s, TRANS_EINVAL, "view has the wrong component count");
    return 0;
}

int trans_srv_release(trans_srv* s, int id)
{
    SrvReq* r = srv_get(s, id);
    if (!r) return TRANS_EINVAL;
    if (r->state == TRANS_SRV_QUEUED || r->state == TRANS_SRV_RUNNING)
        return srv_fail(s, TRANS_EINVAL, "request " + std::to_string(id) + " has not finished");
    r->state_buf.clear();
    r->st.clear();
    return 0;
}

int trans_srv_destroy(trans_srv* s)
{
    if (!s) return 0;
    for (SrvReq* r : s->reqs) {
        if (r->ctx) (void)hipStreamSynchronize(r->ctx->stream);
        srv_free_device(s, r);
        delete r;
    }
    delete s;
    return 0;
}

}  // extern "C"
