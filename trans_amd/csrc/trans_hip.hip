// C-ABI of the MI355X residual-and-update hot path (include/trans_hip.h).
//
// Unity build: the kernels live in the headers included below; this file owns the
// context, the device layout, the schedule tables and the launch sequence that replaces
// the body of the reference's pseudo-time loop (main.F90:65-197).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <climits>
#include "dev_common.h"
#include "k_bc.h"
#include "k_geom.h"
#include "k_io.h"
#include "k_monitor.h"
#include "k_space.h"
#include "k_rk.h"
#include "k_time.h"
#include "k_tile.h"
#include "k_visc.h"

namespace {

constexpr int WG = 256;  // threads per workgroup of the cell/face kernels

struct HostBlock {
    int ni, nj, nk;
    int nx, ny, nz;
    long off;
    long ncell() const { return (long)(ni - 1) * (nj - 1) * (nk - 1); }
};

struct SegTable {
    Seg* d = nullptr;
    int n = 0;
    std::vector<int> wrange;  // per-block first workgroup (size nblk+1)
};

}  // namespace

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

struct trans_ctx {
    trans_params p{};
    // kernel profiler (HIP events on the context's stream)
    bool prof = false;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<long> prof_n;
    int ki = 1;
    std::vector<HostBlock> hb;
    std::vector<trans_patch> patches;      // all blocks' patches, concatenated
    std::vector<int> patch0;               // first patch of each block
    long tot = 0;
    Fields f{};
    BlkInfo* d_blk = nullptr;
    DPatch* d_patch = nullptr;
    BcSeg* d_bcseg = nullptr;
    int n_bcseg = 0;
    OutLine* d_outl = nullptr;  // (patch, i, k) lines of the outlet patches
    int n_outl = 0;
    SegTable seg_cell, seg_face[3];
    SegTable seg_wtile;  // WT x WT (i, j) tiles at each k: the tiled writers of the sweep buffer
    SegTable seg_ghost;  // the three ghost layers beyond each block face (k_prim_ghost)
    // face fluxes + accumulation: 0 the face kernels through G and GV, then k_accumulate_visc;
    // 1 the LDS-tiled viscous and convective passes (k_resid_visc, k_resid_conv) on RTI x RTJ
    // (i, j) tiles marching chunks of rt_kchunk k planes (TRANS_HIP_FUSED=1 at creation)
    SegTable seg_rtile;
    int rt_kchunk = 8;
    int fused = 0;
    int* d_wrange = nullptr;
    NormPart* d_part = nullptr;
    trans_resid* d_resid = nullptr;
    double* d_mon = nullptr;  // 6 doubles: result of k_plane_integrals
    hipStream_t stream = nullptr;
    // helper streams for independent kernels of one call (forked from and joined back
    // into `stream` inside the call, so that calls stay ordered on `stream`)
    hipStream_t aux[3] = {nullptr, nullptr, nullptr};
    hipStream_t aux_own[3] = {nullptr, nullptr, nullptr};  // created helper streams (null: single-stream context)
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
    // one outer iteration captured as a HIP graph and replayed while the iteration's kernel
    // arguments stay the same (set_params and profiling drop it)
    bool graphs = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int march = 1;         // march axis of the line sweep (0 i, 1 j, 2 k)
    int max_lines = 0;     // largest line count over the blocks for that axis
    int line_threads = 0;  // workgroup size of the line-march sweep (0: hyperplane fallback)
    bool plane_launch = false;  // hyperplane sweep as one launch per plane (else one workgroup per block)
    SegTable seg_plane;         // workgroups of the per-plane launches: (block, first line)
    // plain drop-in (sweep -3, a baseline): every routine launched once per block, the sweeps one
    // launch per hyperplane per block, one stream, no graph
    bool plain = false;
    std::vector<int> bc_wrange, outl_wrange;  // per-block ranges of the BC and outlet-line tables
    PlaneBlk* d_pblk = nullptr;
    int max_planes = 0;
    void* wbuf = nullptr;  // the sweep buffer W / Wf (reallocated when the march axis changes)
    size_t wbytes = 0;
    // tile wavefront of the sweeps (sweep mode 3, march axis j): tile tables for NEQ = 7, 5
    // and 2 (mean flow and k-omega, mean flow alone, k-omega alone), the tiles' edge buffers,
    // the item and progress counters of both sweeps, and the error flag of the bounded waits
    bool tile_mode = false;
    int tile_a = 16, tile_b = 16, tile_chunk = 8;
    int tile_grid_k[3] = {0, 0, 0};  // persistent grid of the tile kernel for tables 0, 1, 2
    SweepTile* d_tiles[3] = {nullptr, nullptr, nullptr};
    int ntiles[3] = {0, 0, 0};
    void* tile_edge = nullptr;
    int* d_tctr = nullptr;
    size_t tile_bytes = 0;
    int* d_err = nullptr;
    std::vector<long> mwoff;  // per-block offset of the wall marks (size nblk+1)
    std::vector<void*> allocs;
    // time statistics (Average_Fields): TRANS_NSTAT components, allocated when enabled
    double* d_stats = nullptr;
    bool stats_on = false;
    int ki_aver = 0, ki_statis = 0;
    size_t bytes = 0;  // device bytes allocated by the context
    bool dry = false;  // sizing run of trans_hip_footprint: no HIP call, no allocation
    std::string err;
};

static hipEvent_t prof_event(trans_ctx* c)
{
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

static void prof_flush(trans_ctx* c)
{
    if (c->prof_pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (ProfRec& r : c->prof_pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        size_t q = 0;
        while (q < c->prof_names.size() && c->prof_names[q] != r.name) q++;
        if (q == c->prof_names.size()) {
            c->prof_names.push_back(r.name);
            c->prof_ms.push_back(0.0);
            c->prof_n.push_back(0);
        }
        c->prof_ms[q] += ms;
        c->prof_n[q] += 1;
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_pending.clear();
}

// Records HIP events around one kernel launch when profiling is on.
struct ProfScope {
    trans_ctx* c;
    hipStream_t s;
    ProfRec r{};
    ProfScope(trans_ctx* cc, const char* name, hipStream_t st) : c(cc), s(st)
    {
        if (!c->prof) return;
        r.name = name;
        r.a = prof_event(c);
        r.b = prof_event(c);
        (void)hipEventRecord(r.a, s);
    }
    ~ProfScope()
    {
        if (!c->prof) return;
        (void)hipEventRecord(r.b, s);
        c->prof_pending.push_back(r);
        if (c->prof_pending.size() >= 4096) prof_flush(c);
    }
};

#define KLAUNCH_LDS(c, name, kern, grid, blk, lds, ...)                    \
    do {                                                                   \
        ProfScope ps_(c, name, (c)->stream);                               \
        hipLaunchKernelGGL(kern, grid, blk, lds, (c)->stream, __VA_ARGS__); \
    } while (0)

// launch on helper stream `st` (between fork and join)
#define KLAUNCH_ON(c, st, name, kern, grid, blk, ...)               \
    do {                                                            \
        ProfScope ps_(c, name, st);                                 \
        hipLaunchKernelGGL(kern, grid, blk, 0, st, __VA_ARGS__);    \
    } while (0)

#define KLAUNCH(c, name, kern, grid, blk, ...)                          \
    do {                                                                 \
        ProfScope ps_(c, name, (c)->stream);                             \
        hipLaunchKernelGGL(kern, grid, blk, 0, (c)->stream, __VA_ARGS__); \
    } while (0)

static int fail(trans_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

// Helper streams 0..n-1 wait for the work queued so far on the context's stream.
static void fork_aux(trans_ctx* c, int n)
{
    if (c->aux[0] == c->stream) return;  // single-stream context
    (void)hipEventRecord(c->ev_fork, c->stream);
    for (int q = 0; q < n; q++) (void)hipStreamWaitEvent(c->aux[q], c->ev_fork, 0);
}

// The context's stream waits for the work queued on helper streams 0..n-1.
static void join_aux(trans_ctx* c, int n)
{
    if (c->aux[0] == c->stream) return;
    for (int q = 0; q < n; q++) {
        (void)hipEventRecord(c->ev_join[q], c->aux[q]);
        (void)hipStreamWaitEvent(c->stream, c->ev_join[q], 0);
    }
}

static void drop_graph(trans_ctx* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

#define HIPCHK(ctx, x)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(ctx, TRANS_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// Copies ordered on the context's stream and complete on return (host memory may be
// pageable).  The context's streams are non-blocking, so nothing here may use the null
// stream: work of other contexts on the device is never waited for.
static int h2d(trans_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (c->dry) return 0;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int d2h(trans_ctx* c, void* dst, const void* src, size_t bytes)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

template <typename T>
static int dalloc(trans_ctx* c, T** out, size_t n)
{
    void* p = nullptr;
    if (n == 0) n = 1;
    if (c->dry) {  // sizing run: count only
        c->bytes += n * sizeof(T);
        *out = nullptr;
        return 0;
    }
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) return fail(c, TRANS_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    // stream-ordered clear: the context's streams do not synchronise with the null stream
    if (hipMemsetAsync(p, 0, n * sizeof(T), c->stream) != hipSuccess) return fail(c, TRANS_EHIP, "hipMemsetAsync");
    c->allocs.push_back(p);
    c->bytes += n * sizeof(T);
    *out = (T*)p;
    return 0;
}

static SegTable make_segs(trans_ctx* c, int kind)  // kind: 0 cells, 1..3 faces of dir
{
    std::vector<Seg> segs;
    SegTable t;
    t.wrange.push_back(0);
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        long n0 = h.ni - 1 + (kind == 1), n1 = h.nj - 1 + (kind == 2), n2 = h.nk - 1 + (kind == 3);
        long cnt = n0 * n1 * n2;
        // faces of directions j and k run in FTA x FTB tiles of the (i, j) or (i, k) plane
        // (decode_face_tile): whole tiles, padding included
        if (kind == 2) cnt = (long)FTA * FTB * ((n0 + FTA - 1) / FTA) * ((n1 + FTB - 1) / FTB) * n2;
        if (kind == 3) cnt = (long)FTA * FTB * ((n0 + FTA - 1) / FTA) * ((n2 + FTB - 1) / FTB) * n1;
        for (long base = 0; base < cnt; base += WG) segs.push_back(Seg{b, (int)base});
        t.wrange.push_back((int)segs.size());
    }
    t.n = (int)segs.size();
    if (dalloc(c, &t.d, segs.size()) == 0)
        (void)h2d(c, t.d, segs.data(), segs.size() * sizeof(Seg));
    return t;
}

// Schedule of the tiled sweep-buffer kernels: per block, ceil(NI1/WT) x ceil(NJ1/WT) tiles at
// each k, entry Seg{block, tile index}.
static SegTable make_wtile_segs(trans_ctx* c)
{
    std::vector<Seg> segs;
    SegTable t;
    t.wrange.push_back(0);
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        const int nt = ((h.ni - 1 + WT - 1) / WT) * ((h.nj - 1 + WT - 1) / WT) * (h.nk - 1);
        for (int q = 0; q < nt; q++) segs.push_back(Seg{b, q});
        t.wrange.push_back((int)segs.size());
    }
    t.n = (int)segs.size();
    if (dalloc(c, &t.d, segs.size()) == 0)
        (void)h2d(c, t.d, segs.data(), segs.size() * sizeof(Seg));
    return t;
}

// Schedule of k_prim_ghost: per block, ghost_slab_cells in 256-cell workgroups.
static SegTable make_ghost_segs(trans_ctx* c)
{
    std::vector<Seg> segs;
    SegTable t;
    t.wrange.push_back(0);
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        const long n1 = h.ni - 1, n2 = h.nj - 1, n3 = h.nk - 1;
        const long cnt = 6 * (n2 * n3 + n1 * n3 + n1 * n2);
        for (long base = 0; base < cnt; base += WG) segs.push_back(Seg{b, (int)base});
        t.wrange.push_back((int)segs.size());
    }
    t.n = (int)segs.size();
    if (dalloc(c, &t.d, segs.size()) == 0)
        (void)h2d(c, t.d, segs.data(), segs.size() * sizeof(Seg));
    return t;
}

// Schedule of k_resid_visc and k_resid_conv: per block, ceil(NI1/RTI) x ceil(NJ1/RTJ) tiles times
// ceil(NK1/kchunk) chunks of k planes, entry Seg{block, tile index}.
static SegTable make_rtile_segs(trans_ctx* c, int kchunk)
{
    std::vector<Seg> segs;
    SegTable t;
    t.wrange.push_back(0);
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        const int nt = ((h.ni - 1 + RTI - 1) / RTI) * ((h.nj - 1 + RTJ - 1) / RTJ) * ((h.nk - 1 + kchunk - 1) / kchunk);
        for (int q = 0; q < nt; q++) segs.push_back(Seg{b, q});
        t.wrange.push_back((int)segs.size());
    }
    t.n = (int)segs.size();
    if (dalloc(c, &t.d, segs.size()) == 0)
        (void)h2d(c, t.d, segs.data(), segs.size() * sizeof(Seg));
    return t;
}

extern "C" {

const char* trans_hip_last_error(const trans_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static int check_params(trans_ctx* c, const trans_params* p)
{
    if (p->if_precondition != 0 && p->if_precondition != 1) return fail(c, TRANS_EINVAL, "IF_PRECONDITION must be 0 or 1");
    // RCCorrection (Space/RC_Correction.F90): mode 2 only writes f_r2, which no routine reads,
    // so it leaves the state as mode 0 does; mode 1 reads omega(M) with M unset (:107)
    if (p->kind_rotmod != 0 && p->kind_rotmod != 2)
        return fail(c, TRANS_EUNSUPPORTED, "Kind_RotMod must be 0 or 2 (mode 1 reads an unset index)");
    if (p->cfl_sub_step < 1) return fail(c, TRANS_EINVAL, "CFL_Sub_step must be >= 1");
    if (p->if_rk != 0 && p->if_rk != 1) return fail(c, TRANS_EINVAL, "IF_RK must be 0 or 1");
    if (p->kind_spatial < 1 || p->kind_spatial > 4)
        return fail(c, TRANS_EUNSUPPORTED, "Kind_Spatial must be 1 (MUSCL2), 2 (MUSCL), 3 (MDCD) or 4 (MDCDHYB)");
    if (p->kind_roe < 1 || p->kind_roe > 3)
        return fail(c, TRANS_EUNSUPPORTED, "Kind_Roe must be 1 (Roe), 2 (rotated Roe) or 3 (AUSMPW+)");
    if (p->if_euler == 0 && p->if_turb_org == 1) {
        if (p->kind_model != 1 && p->kind_model != 2)
            return fail(c, TRANS_EUNSUPPORTED, "Kind_Model must be 1 (SST) or 2 (WD+ k-omega)");
        if (p->kind_hybrid < 0) return fail(c, TRANS_EINVAL, "Kind_Hybrid must be >= 0");
        if (p->kind_hybrid == 3 && (p->kind_subhy < 0 || (p->kind_subhy > 3 && p->kind_subhy != 5)))
            return fail(c, TRANS_EUNSUPPORTED, "Kind_SubHY must be 0-3 (DES, DDES) or 5 (IDDES)");
        if (p->if_comppress != 0 && p->if_comppress != 1)
            return fail(c, TRANS_EUNSUPPORTED, "IF_Comppress must be 0 or 1 (Suzen)");
    }
    return 0;
}

// Whether the line-march sweep can run along axis m (0 i, 1 j, 2 k): every block's cross
// section (the lines) fits one workgroup and the LDS rows of the march.  *mx: the largest.
static bool march_fits(const trans_ctx* c, int m, int* mx_out)
{
    const size_t esz = c->p.implicit_fp32 ? sizeof(float) : sizeof(double);
    const size_t neq = (c->p.if_euler == 0 && c->p.if_turb_org == 1) ? 7 : 5;
    int mx = 0;
    for (const HostBlock& h : c->hb) {
        const int n[3] = {h.ni - 1, h.nj - 1, h.nk - 1};
        mx = std::max(mx, n[(m + 1) % 3] * n[(m + 2) % 3]);
    }
    if (mx_out) *mx_out = mx;
    return mx <= 1024 && 4 * std::min(neq, (size_t)5) * (mx + 2) * esz <= 128 * 1024;
}

// Per-block descriptors (with each block's first slot in the sweep buffer) to the device.
static int upload_blk_table(trans_ctx* c)
{
    const int nblk = (int)c->hb.size();
    std::vector<BlkInfo> bi(nblk);
    long wt = 0;
    for (int b = 0; b < nblk; b++) {
        const HostBlock& h = c->hb[b];
        const int n[3] = {h.ni - 1, h.nj - 1, h.nk - 1};
        bi[b] = BlkInfo{h.ni, h.nj, h.nk, h.nx, h.ny, h.nz, h.off, c->patch0[b + 1] - c->patch0[b], c->patch0[b],
                        c->mwoff[b], c->wbytes ? wt : 0};
        wt += (long)(n[0] + n[1] + n[2] - 2) * n[(c->march + 1) % 3] * n[(c->march + 2) % 3];
    }
    if (!c->d_blk) return 0;  // set before the table exists: create uploads it
    return h2d(c, c->d_blk, bi.data(), nblk * sizeof(BlkInfo));
}

// The sweep in use: 0, 1, 2 line march along that axis, 3 tile wavefront, -2 one launch per
// hyperplane, -1 one workgroup per block.
static int cur_sweep(const trans_ctx* c)
{
    if (c->plain) return -3;
    if (c->tile_mode) return 3;
    return c->line_threads ? c->march : (c->plane_launch ? -2 : -1);
}

static void free_tiles(trans_ctx* c)
{
    if (!c->dry) {
        for (SweepTile*& t : c->d_tiles)
            if (t) (void)hipFree(t);
        if (c->tile_edge) (void)hipFree(c->tile_edge);
        if (c->d_tctr) (void)hipFree(c->d_tctr);
    }
    for (SweepTile*& t : c->d_tiles) t = nullptr;
    for (int& n : c->ntiles) n = 0;
    c->tile_edge = nullptr;
    c->d_tctr = nullptr;
    c->bytes -= c->tile_bytes;
    c->tile_bytes = 0;
}

// Tile tables of the wavefront sweep for the current tile shape (tile_a x tile_b lines):
// table 0 holds the mean-flow and k-omega tiles, 1 the mean-flow ones, 2 the k-omega ones,
// each in the order of the tiles' diagonals (upstream tiles first).  Edge buffers are laid
// out once for table 0; the other tables use the same offsets.
static int build_tiles(trans_ctx* c)
{
    const int TA = c->tile_a, TB = c->tile_b;
    if (TA < 1 || TB < 1 || TA * TB > TILE_MAXT) return fail(c, TRANS_EINVAL, "tile shape must have 1..256 lines");
    const size_t esz = c->p.implicit_fp32 ? sizeof(float) : sizeof(double);
    struct Item {
        int diag, blk, neq, ta, tb;
    };
    std::vector<Item> all;
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        const int nta = (h.ni - 1 + TA - 1) / TA, ntb = (h.nk - 1 + TB - 1) / TB;
        for (int tb = 0; tb < ntb; tb++)
            for (int ta = 0; ta < nta; ta++)
                for (int neq : {5, 2}) all.push_back(Item{ta + tb, b, neq, ta, tb});
    }
    std::stable_sort(all.begin(), all.end(), [](const Item& x, const Item& y) {
        if (x.diag != y.diag) return x.diag < y.diag;
        if (x.blk != y.blk) return x.blk < y.blk;
        return x.neq > y.neq;
    });
    // edge offsets (elements), per (block, kind, ta, tb), from the full table
    std::map<std::tuple<int, int, int, int>, std::pair<long, long>> eoff;
    long etot = 0;
    for (const Item& it : all) {
        const HostBlock& h = c->hb[it.blk];
        const int na = std::min(TA, h.ni - 1 - it.ta * TA), nb = std::min(TB, h.nk - 1 - it.tb * TB);
        const long ea = etot;
        etot += (long)it.neq * (h.nj - 1) * nb;
        const long eb = etot;
        etot += (long)it.neq * (h.nj - 1) * na;
        eoff[std::make_tuple(it.blk, it.neq, it.ta, it.tb)] = std::make_pair(ea, eb);
    }
    int ntmax = 0;
    std::vector<SweepTile> tab[3];
    for (int w = 0; w < 3; w++) {
        std::map<std::tuple<int, int, int, int>, int> index;
        for (const Item& it : all) {
            if ((w == 1 && it.neq != 5) || (w == 2 && it.neq != 2)) continue;
            const HostBlock& h = c->hb[it.blk];
            SweepTile t{};
            t.blk = it.blk;
            t.neq = it.neq;
            t.a0 = it.ta * TA;
            t.b0 = it.tb * TB;
            t.na = std::min(TA, h.ni - 1 - t.a0);
            t.nb = std::min(TB, h.nk - 1 - t.b0);
            const auto& e = eoff[std::make_tuple(it.blk, it.neq, it.ta, it.tb)];
            t.ea = e.first;
            t.eb = e.second;
            index[std::make_tuple(it.blk, it.neq, it.ta, it.tb)] = (int)tab[w].size();
            tab[w].push_back(t);
        }
        for (SweepTile& t : tab[w]) {
            const int ta = t.a0 / TA, tb = t.b0 / TB;
            auto find = [&](int a, int b) {
                auto q = index.find(std::make_tuple(t.blk, t.neq, a, b));
                return q == index.end() ? -1 : q->second;
            };
            t.lo_a = find(ta - 1, tb);
            t.hi_a = find(ta + 1, tb);
            t.lo_b = find(ta, tb - 1);
            t.hi_b = find(ta, tb + 1);
        }
        c->ntiles[w] = (int)tab[w].size();
        ntmax = std::max(ntmax, c->ntiles[w]);
    }
    const size_t ctr_n = 2 + 2 * (size_t)ntmax;
    const size_t bytes = 3 * (size_t)ntmax * sizeof(SweepTile) + (size_t)std::max(etot, 1L) * esz + ctr_n * sizeof(int);
    c->tile_bytes = bytes;
    c->bytes += bytes;
    if (c->dry) return 0;
    for (int w = 0; w < 3; w++) {
        HIPCHK(c, hipMalloc(&c->d_tiles[w], std::max<size_t>(1, tab[w].size()) * sizeof(SweepTile)));
        if (!tab[w].empty()) {
            int rc = h2d(c, c->d_tiles[w], tab[w].data(), tab[w].size() * sizeof(SweepTile));
            if (rc) return rc;
        }
    }
    HIPCHK(c, hipMalloc(&c->tile_edge, (size_t)std::max(etot, 1L) * esz));
    HIPCHK(c, hipMemsetAsync(c->tile_edge, 0, (size_t)std::max(etot, 1L) * esz, c->stream));
    HIPCHK(c, hipMalloc(&c->d_tctr, ctr_n * sizeof(int)));
    HIPCHK(c, hipMemsetAsync(c->d_tctr, 0, ctr_n * sizeof(int), c->stream));
    // persistent grid: as many tile workgroups as the device holds at once
    int dev = 0, cus = 0, occ = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int w = 0; w < 3; w++) {
        const void* kern;
        if (c->p.implicit_fp32)
            kern = w == 0 ? (const void*)k_lusgs_tile<float, true, 0>
                          : (w == 1 ? (const void*)k_lusgs_tile<float, true, 5> : (const void*)k_lusgs_tile<float, true, 2>);
        else
            kern = w == 0 ? (const void*)k_lusgs_tile<double, true, 0>
                          : (w == 1 ? (const void*)k_lusgs_tile<double, true, 5> : (const void*)k_lusgs_tile<double, true, 2>);
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, TILE_MAXT, 0));
        c->tile_grid_k[w] = std::max(1, std::min(TILE_GRID_MAX, std::max(1, occ) * std::max(1, cus)));
    }
    return 0;
}

// Selects the LU-SGS sweep: m = 0, 1, 2 the line march along that axis (hyperplane-major
// buffer W or Wf, reallocated here), m = 3 the tile wavefront (march axis j), m = -2 one
// launch per hyperplane, m = -1 the hyperplane kernel on the natural layout.
static int set_sweep(trans_ctx* c, int m, bool upload)
{
    int mx = 0;
    // the low-Mach preconditioning transforms live in the natural-layout hyperplane sweep
    if (c->p.if_precondition && m != -1)
        return fail(c, TRANS_EUNSUPPORTED, "with IF_PRECONDITION = 1 only the hyperplane sweep (-1) runs");
    if (m < -2 || m > 3) return fail(c, TRANS_EINVAL, "sweep must be -2, -1, 0, 1, 2 or 3");
    const bool planes = m == -2;  // per-plane launches on the hyperplane-major buffer, march axis j
    const bool tiles = m == 3;    // tile wavefront on the hyperplane-major buffer, march axis j
    if (planes || tiles) m = 1;
    if (!planes && !tiles && m >= 0 && !march_fits(c, m, &mx))
        return fail(c, TRANS_EINVAL, "the line sweep does not fit this axis");
    c->plane_launch = planes;
    c->tile_mode = tiles;
    free_tiles(c);
    if (c->wbuf) {
        if (!c->dry) (void)hipFree(c->wbuf);
        c->wbuf = nullptr;
    }
    c->bytes -= c->wbytes;
    c->wbytes = 0;
    Fields& f = c->f;
    f.W = nullptr;
    f.Wf = nullptr;
    c->line_threads = 0;
    if (m >= 0) {
        c->march = m;
        c->max_lines = mx;
        c->line_threads = (planes || tiles) ? 0 : (mx + 63) / 64 * 64;
        // hyperplane-major sweep buffer: nplanes x nlines slots per block, then per
        // component a sink of 1024 slots for each of the 2 workgroups of every block
        long wt = 0;
        for (const HostBlock& h : c->hb) {
            const int n[3] = {h.ni - 1, h.nj - 1, h.nk - 1};
            wt += (long)(n[0] + n[1] + n[2] - 2) * n[(m + 1) % 3] * n[(m + 2) % 3];
        }
        f.wsink = wt;
        // (the tile wavefront's persistent workgroups have TILE_MAXT sink slots each)
        wt += tiles ? std::max((long)c->hb.size() * 2 * 1024, (long)TILE_GRID_MAX * TILE_MAXT)
                    : (long)c->hb.size() * 2 * 1024;
        // the sweep addresses a component with 32-bit byte offsets
        if (wt * (long)sizeof(double) >= (1L << 32)) {
            c->line_threads = 0;
            c->plane_launch = false;
            c->tile_mode = false;
            return fail(c, TRANS_EUNSUPPORTED, "too many cells for the line-march sweep buffer");
        }
        f.wtot = wt;
        f.march = m;
        const size_t bytes = (size_t)W_NCOMP * wt * (c->p.implicit_fp32 ? sizeof(float) : sizeof(double));
        if (!c->dry) {
            hipError_t e = hipMalloc(&c->wbuf, bytes);
            if (e != hipSuccess) return fail(c, TRANS_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            HIPCHK(c, hipMemsetAsync(c->wbuf, 0, bytes, c->stream));
        }
        c->wbytes = bytes;
        c->bytes += bytes;
        if (c->p.implicit_fp32)
            f.Wf = (float*)c->wbuf;
        else
            f.W = (double*)c->wbuf;
    }
    if (tiles) {
        int rc = build_tiles(c);
        if (rc) return rc;
    }
    drop_graph(c);
    if (!upload) return 0;
    int rc = upload_blk_table(c);
    if (rc || m < 0 || c->dry) return rc;
    // the geometry part of the new buffer (face means W_SDM, written once per geometry upload)
    hipLaunchKernelGGL(k_face_means, dim3(c->seg_cell.n), dim3(WG), 0, c->stream, c->f, c->d_blk, c->seg_cell.d);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// dry: sizing run only.  single_stream: the helper streams are the context's own stream, so
// that every launch of the context is ordered on one stream (the serving layer's contexts:
// one stream each spreads the resident solves over the device's hardware queues).
static int create_ctx(const trans_params* params, int nblk, const trans_block_desc* blocks, trans_ctx** out,
                      bool dry, bool single_stream = false)
{
    if (!params || !blocks || !out || nblk <= 0) return TRANS_EINVAL;
    trans_ctx* c = new trans_ctx();
    *out = c;
    c->dry = dry;
    int rc = check_params(c, params);
    if (rc) return rc;
    c->p = *params;
    long off = 0;
    for (int b = 0; b < nblk; b++) {
        const trans_block_desc& d = blocks[b];
        if (d.ni < 4 || d.nj < 4 || d.nk < 3) return fail(c, TRANS_EINVAL, "block too small");
        HostBlock h;
        h.ni = d.ni; h.nj = d.nj; h.nk = d.nk;
        h.nx = (d.ni + 5 + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN; h.ny = d.nj + 5; h.nz = d.nk + 5;
        // schedule tables and the per-block index arithmetic of the kernels are 32-bit
        if ((long)h.nx * h.ny * h.nz > (long)INT_MAX / 8)
            return fail(c, TRANS_EINVAL, "block too large (more than 2^28 padded cells)");
        // interior i = 1 (element off + 3 of each row) on a multiple of ROW_ALIGN
        h.off = (off + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN + (ROW_ALIGN - 3);
        off = h.off + (long)h.nx * h.ny * h.nz;
        c->hb.push_back(h);
        c->patch0.push_back((int)c->patches.size());
        for (int q = 0; q < d.npatch; q++) c->patches.push_back(d.patches[q]);
    }
    c->patch0.push_back((int)c->patches.size());
    c->tot = (off + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN;
    for (int b = 0; b < nblk; b++)
        for (int q = c->patch0[b]; q < c->patch0[b + 1]; q++) {
            const trans_patch& pt = c->patches[q];
            if (pt.bctype == 9 && (pt.target_block < 0 || pt.target_block >= nblk ||
                                   pt.target_conn < 0 ||
                                   pt.target_conn >= c->patch0[pt.target_block + 1] - c->patch0[pt.target_block]))
                return fail(c, TRANS_EINVAL, "interface target out of range");
        }
    if (!dry) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (int q = 0; q < 3; q++) {
            if (single_stream) {
                c->aux[q] = c->stream;
                continue;
            }
            HIPCHK(c, hipStreamCreateWithFlags(&c->aux_own[q], hipStreamNonBlocking));
            c->aux[q] = c->aux_own[q];
            HIPCHK(c, hipEventCreateWithFlags(&c->ev_join[q], hipEventDisableTiming));
        }
        if (!single_stream) HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    }
    {
        const char* env = std::getenv("TRANS_HIP_GRAPH");
        c->graphs = !(env && env[0] == '0');
    }

    Fields& f = c->f;
    f.tot = c->tot;
    const long n = c->tot;
    if ((rc = dalloc(c, &f.V, 7 * n)) || (rc = dalloc(c, &f.PP, n)) || (rc = dalloc(c, &f.T, n)) ||
        (rc = dalloc(c, &f.Rmiu, n)) || (rc = dalloc(c, &f.SD, 9 * n)) || (rc = dalloc(c, &f.Grad, 3 * n)) ||
        (rc = dalloc(c, &f.Alagm, 3 * n)) || (rc = dalloc(c, &f.Vol, n)) || (rc = dalloc(c, &f.rad, n)) ||
        (rc = dalloc(c, &f.Dst, n)) || (rc = dalloc(c, &f.aLeng, n)) || (rc = dalloc(c, &f.aSeng, n)) || (rc = dalloc(c, &f.Yc, n)) ||
        (rc = dalloc(c, &f.Zc, n)) || (rc = dalloc(c, &f.thtc, n)) || (rc = dalloc(c, &f.D, 7 * n)) ||
        (rc = dalloc(c, &f.Rds, 3 * n)) || (rc = dalloc(c, &f.Dtm, n)) || (rc = dalloc(c, &f.Wt0, 7 * n)) ||
        (rc = dalloc(c, &f.Wt1, 7 * n)) || (rc = dalloc(c, &f.Wt2, 7 * n)) || (rc = dalloc(c, &f.DQ, 18 * n)) ||
        (rc = dalloc(c, &f.Q, 5 * n)) || (rc = dalloc(c, &f.shock, n)) || (rc = dalloc(c, &f.Fscheme, n)) || (rc = dalloc(c, &f.F1, n)) || (rc = dalloc(c, &f.F1t, n)) ||
        (rc = dalloc(c, &f.SDm, 9 * n)) || (rc = dalloc(c, &f.Stp, 2 * n)) ||
        (rc = dalloc(c, &f.G[0], 7 * n)) || (rc = dalloc(c, &f.G[1], 7 * n)) || (rc = dalloc(c, &f.G[2], 7 * n)))
        return rc;
    if (c->p.if_euler == 0) {
        if ((rc = dalloc(c, &f.GV[0], 6 * n)) || (rc = dalloc(c, &f.GV[1], 6 * n)) ||
            (rc = dalloc(c, &f.GV[2], 6 * n)))
            return rc;
    }
    if (c->p.implicit_fp32) {
        if ((rc = dalloc(c, &f.Ds, 7 * n)) || (rc = dalloc(c, &f.Rdsf, 3 * n))) return rc;
    }
    c->mwoff.push_back(0);
    for (int b = 0; b < nblk; b++) {
        const HostBlock& h = c->hb[b];
        const long a = (long)(h.nj - 1) * (h.nk - 1), bb = (long)(h.ni - 1) * (h.nk - 1),
                   cc = (long)(h.ni - 1) * (h.nj - 1);
        c->mwoff.push_back(c->mwoff.back() + 2 * (a + bb + cc));
    }
    if ((rc = dalloc(c, &f.mw, c->mwoff.back()))) return rc;


    // BC patch table and its schedule
    std::vector<DPatch> dp;
    std::vector<BcSeg> bs;
    for (int b = 0; b < nblk; b++)
        for (int q = c->patch0[b]; q < c->patch0[b + 1]; q++) {
            DPatch d{};
            d.p = c->patches[q];
            d.blk = b;
            d.tpatch = d.p.bctype == 9 ? c->patch0[d.p.target_block] + d.p.target_conn : 0;
            const int n0 = d.p.point_end[0] - d.p.point_start[0] + 1;
            const int n1 = d.p.point_end[1] - d.p.point_start[1] + 1;
            const int n2 = d.p.point_end[2] - d.p.point_start[2] + 1;
            d.count = n0 * n1 * n2;
            if (d.p.bctype == 7 && d.p.ijkdirect == 2)
                return fail(c, TRANS_EUNSUPPORTED, "outlet on a j face");
            dp.push_back(d);
        }
    for (int q = 0; q < (int)dp.size(); q++)
        for (int base = 0; base < dp[q].count; base += 128) bs.push_back(BcSeg{q, base});
    std::vector<OutLine> ol;
    for (int q = 0; q < (int)dp.size(); q++) {
        const trans_patch& pt = dp[q].p;
        if (pt.bctype != 7) continue;
        if (pt.point_end[1] - pt.point_start[1] + 1 > OUTLET_MAXJ)
            return fail(c, TRANS_EUNSUPPORTED, "outlet patch longer than " + std::to_string(OUTLET_MAXJ) + " in j");
        for (int k = pt.point_start[2]; k <= pt.point_end[2]; k++)
            for (int i = pt.point_start[0]; i <= pt.point_end[0]; i++) ol.push_back(OutLine{q, i, k});
    }
    // per-block ranges of the BC segments and outlet lines (both are in block order)
    c->bc_wrange.assign(nblk + 1, 0);
    c->outl_wrange.assign(nblk + 1, 0);
    for (const BcSeg& q : bs) c->bc_wrange[dp[q.patch].blk + 1]++;
    for (const OutLine& q : ol) c->outl_wrange[dp[q.patch].blk + 1]++;
    for (int b = 0; b < nblk; b++) {
        c->bc_wrange[b + 1] += c->bc_wrange[b];
        c->outl_wrange[b + 1] += c->outl_wrange[b];
    }
    c->n_outl = (int)ol.size();
    if (c->n_outl) {
        if ((rc = dalloc(c, &c->d_outl, ol.size()))) return rc;
        if ((rc = h2d(c, c->d_outl, ol.data(), ol.size() * sizeof(OutLine)))) return rc;
    }
    c->n_bcseg = (int)bs.size();
    if ((rc = dalloc(c, &c->d_patch, dp.size())) || (rc = dalloc(c, &c->d_bcseg, bs.size()))) return rc;
    if (!dp.empty() && (rc = h2d(c, c->d_patch, dp.data(), dp.size() * sizeof(DPatch)))) return rc;
    if (!bs.empty() && (rc = h2d(c, c->d_bcseg, bs.data(), bs.size() * sizeof(BcSeg)))) return rc;

    {
        // march along j (lanes on i: coalesced) unless the (i,k) sections do not fit a workgroup;
        // TRANS_HIP_LINE_SWEEP=0 in the environment, or preconditioning, selects the hyperplane
        // sweep instead
        const char* env = std::getenv("TRANS_HIP_LINE_SWEEP");
        const bool lines_ok = !(env && env[0] == '0') && !c->p.if_precondition;
        const int order[3] = {1, 2, 0};
        int m = -1;
        for (int q = 0; q < 3 && lines_ok && m < 0; q++)
            if (march_fits(c, order[q], nullptr)) m = order[q];
        if (lines_ok && m < 0) m = 3;  // blocks too large for a line-march workgroup: the tile wavefront
        if ((rc = set_sweep(c, m, false))) return rc;
    }
    if ((rc = dalloc(c, &c->d_blk, nblk))) return rc;
    if ((rc = upload_blk_table(c))) return rc;
    c->seg_cell = make_segs(c, 0);
    for (int d = 0; d < 3; d++) c->seg_face[d] = make_segs(c, d + 1);
    c->seg_wtile = make_wtile_segs(c);
    c->seg_ghost = make_ghost_segs(c);
    {
        const char* env = std::getenv("TRANS_HIP_FUSED");
        if (env && (env[0] == '0' || env[0] == '1')) c->fused = env[0] - '0';
        const char* kc = std::getenv("TRANS_HIP_RESID_KCHUNK");
        if (kc && std::atoi(kc) > 0) c->rt_kchunk = std::atoi(kc);
        c->seg_rtile = make_rtile_segs(c, c->rt_kchunk);
    }
    {
        // per-plane sweep launches: lines (i, k) of every block, march axis j
        std::vector<PlaneBlk> pb;
        std::vector<Seg> ps;
        c->seg_plane.wrange.push_back(0);
        for (int b = 0; b < nblk; b++) {
            const HostBlock& h = c->hb[b];
            const int n[3] = {h.ni - 1, h.nj - 1, h.nk - 1};
            PlaneBlk q{n[0] + n[1] + n[2] - 2, n[0], n[2]};
            pb.push_back(q);
            c->max_planes = std::max(c->max_planes, q.nplanes);
            for (int base = 0; base < q.na * q.nb; base += WG) ps.push_back(Seg{b, base});
            c->seg_plane.wrange.push_back((int)ps.size());
        }
        c->seg_plane.n = (int)ps.size();
        if ((rc = dalloc(c, &c->seg_plane.d, ps.size())) || (rc = dalloc(c, &c->d_pblk, pb.size()))) return rc;
        if ((rc = h2d(c, c->seg_plane.d, ps.data(), ps.size() * sizeof(Seg))) ||
            (rc = h2d(c, c->d_pblk, pb.data(), pb.size() * sizeof(PlaneBlk))))
            return rc;
    }
    if ((rc = dalloc(c, &c->d_wrange, nblk + 1))) return rc;
    if ((rc = h2d(c, c->d_wrange, c->seg_cell.wrange.data(), (nblk + 1) * sizeof(int)))) return rc;
    if ((rc = dalloc(c, &c->d_part, c->seg_cell.n)) || (rc = dalloc(c, &c->d_resid, nblk)) ||
        (rc = dalloc(c, &c->d_mon, 6)) || (rc = dalloc(c, &c->d_err, 1)))
        return rc;
    if (!dry) HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int trans_hip_create(const trans_params* params, int nblk, const trans_block_desc* blocks,
                     trans_ctx** out)
{
    return create_ctx(params, nblk, blocks, out, false);
}

int trans_hip_footprint(const trans_params* params, int nblk, const trans_block_desc* blocks, size_t* bytes)
{
    if (!bytes) return TRANS_EINVAL;
    trans_ctx* c = nullptr;
    const int rc = create_ctx(params, nblk, blocks, &c, true);
    *bytes = c ? c->bytes : 0;
    trans_hip_destroy(c);
    return rc;
}

int trans_hip_device_bytes(const trans_ctx* c, size_t* bytes)
{
    if (!c || !bytes) return TRANS_EINVAL;
    *bytes = c->bytes;
    return 0;
}

int trans_hip_destroy(trans_ctx* c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (ProfRec& r : c->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->wbuf) (void)hipFree(c->wbuf);
    free_tiles(c);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (int q = 0; q < 3; q++) {
        if (c->aux_own[q]) (void)hipStreamDestroy(c->aux_own[q]);
        if (c->ev_join[q]) (void)hipEventDestroy(c->ev_join[q]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int trans_hip_set_params(trans_ctx* c, const trans_params* p)
{
    if (!c || !p) return TRANS_EINVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (p->implicit_fp32 != c->p.implicit_fp32)
        return fail(c, TRANS_EINVAL, "implicit_fp32 is fixed at context creation");
    if (p->if_euler != c->p.if_euler || p->if_turb_org != c->p.if_turb_org)
        return fail(c, TRANS_EINVAL, "IF_Euler and IF_Turb_Org are fixed at context creation");
    if (p->if_precondition != c->p.if_precondition)
        return fail(c, TRANS_EINVAL, "IF_PRECONDITION is fixed at context creation (it selects the sweep)");
    c->p = *p;
    drop_graph(c);
    return 0;
}

int trans_hip_set_iteration(trans_ctx* c, int ki)
{
    if (!c) return TRANS_EINVAL;
    c->ki = ki;
    return 0;
}

// ---------------------------------------------------------------------------------------
// host <-> device transfers through Fortran-layout views

static int copy_view(trans_ctx* c, int blk, const trans_fview& v, double* dbase, int ncomp, bool upload)
{
    if (!v.p) return 0;
    if (v.ncomp != ncomp) return fail(c, TRANS_EINVAL, "view has the wrong component count");
    const HostBlock& h = c->hb[blk];
    // clip the view to the device extent -2..N+2
    int lo[3], hi[3];
    const int dims[3] = {h.ni, h.nj, h.nk};
    for (int d = 0; d < 3; d++) {
        lo[d] = std::max(v.lo[d], -2);
        hi[d] = std::min(v.hi[d], dims[d] + 2);
        if (hi[d] < lo[d]) return 0;
    }
    const long vn0 = v.hi[0] - v.lo[0] + 1, vn1 = v.hi[1] - v.lo[1] + 1;
    const long blkn = (long)h.nx * h.ny * h.nz;
    std::vector<double> stage((size_t)ncomp * blkn);
    for (int cc = 0; cc < ncomp; cc++) {
        double* dst = dbase + (long)cc * c->tot + h.off;
        // (an upload also reads the block first: the view may cover only part of it)
        int rc = d2h(c, stage.data() + (size_t)cc * blkn, dst, blkn * sizeof(double));
        if (rc) return rc;
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    const long vi = cc + (long)ncomp * ((i - v.lo[0]) + vn0 * ((j - v.lo[1]) + vn1 * (long)(k - v.lo[2])));
                    const long di = (long)cc * blkn + (i + 2) + (long)h.nx * ((j + 2) + (long)h.ny * (k + 2));
                    if (upload)
                        stage[di] = v.p[vi];
                    else
                        v.p[vi] = stage[di];
                }
        if (upload && (rc = h2d(c, dst, stage.data() + (size_t)cc * blkn, blkn * sizeof(double)))) return rc;
    }
    return 0;
}

int trans_hip_set_geometry(trans_ctx* c, int blk, const trans_geom_view* g)
{
    if (!c || !g || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Fields& f = c->f;
    int rc;
    if ((rc = copy_view(c, blk, g->sd, f.SD, 9, true)) || (rc = copy_view(c, blk, g->grad, f.Grad, 3, true)) ||
        (rc = copy_view(c, blk, g->alagm, f.Alagm, 3, true)) || (rc = copy_view(c, blk, g->vol, f.Vol, 1, true)) ||
        (rc = copy_view(c, blk, g->rad, f.rad, 1, true)) || (rc = copy_view(c, blk, g->dst, f.Dst, 1, true)) ||
        (rc = copy_view(c, blk, g->aleng, f.aLeng, 1, true)) || (rc = copy_view(c, blk, g->aseng, f.aSeng, 1, true)) || (rc = copy_view(c, blk, g->yc, f.Yc, 1, true)) ||
        (rc = copy_view(c, blk, g->zc, f.Zc, 1, true)) || (rc = copy_view(c, blk, g->thtc, f.thtc, 1, true)))
        return rc;
    // wall marks (Input/Read_BlockDimensions.F90:146-283): faces i0, im, j0, jm, k0, km
    const HostBlock& h = c->hb[blk];
    const long fsz[6] = {(long)(h.nj - 1) * (h.nk - 1), (long)(h.nj - 1) * (h.nk - 1),
                         (long)(h.ni - 1) * (h.nk - 1), (long)(h.ni - 1) * (h.nk - 1),
                         (long)(h.ni - 1) * (h.nj - 1), (long)(h.ni - 1) * (h.nj - 1)};
    std::vector<int> mw;
    for (int q = 0; q < 6; q++) {
        if (g->markwall[q])
            mw.insert(mw.end(), g->markwall[q], g->markwall[q] + fsz[q]);
        else
            mw.insert(mw.end(), (size_t)fsz[q], 0);
    }
    if ((rc = h2d(c, f.mw + c->mwoff[blk], mw.data(), mw.size() * sizeof(int)))) return rc;
    hipLaunchKernelGGL(k_face_means, dim3(c->seg_cell.n), dim3(WG), 0, c->stream, c->f, c->d_blk, c->seg_cell.d);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int trans_hip_set_state(trans_ctx* c, int blk, const trans_state_view* s)
{
    if (!c || !s || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Fields& f = c->f;
    int rc;
    if ((rc = copy_view(c, blk, s->v, f.V, 7, true)) || (rc = copy_view(c, blk, s->pp, f.PP, 1, true)) ||
        (rc = copy_view(c, blk, s->t, f.T, 1, true)) || (rc = copy_view(c, blk, s->rmiu, f.Rmiu, 1, true)) ||
        (rc = copy_view(c, blk, s->wt0, f.Wt0, 7, true)) || (rc = copy_view(c, blk, s->wt1, f.Wt1, 7, true)) ||
        (rc = copy_view(c, blk, s->wt2, f.Wt2, 7, true)))
        return rc;
    return 0;
}

int trans_hip_get_state(trans_ctx* c, int blk, const trans_state_view* s)
{
    if (!c || !s || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Fields& f = c->f;
    int rc;
    if ((rc = copy_view(c, blk, s->v, f.V, 7, false)) || (rc = copy_view(c, blk, s->pp, f.PP, 1, false)) ||
        (rc = copy_view(c, blk, s->t, f.T, 1, false)) || (rc = copy_view(c, blk, s->rmiu, f.Rmiu, 1, false)) ||
        (rc = copy_view(c, blk, s->wt0, f.Wt0, 7, false)) || (rc = copy_view(c, blk, s->wt1, f.Wt1, 7, false)) ||
        (rc = copy_view(c, blk, s->wt2, f.Wt2, 7, false)))
        return rc;
    return 0;
}

int trans_hip_get_field(trans_ctx* c, int blk, int which, const trans_fview* out)
{
    if (!c || !out || blk < 0 || blk >= (int)c->hb.size()) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Fields& f = c->f;
    switch (which) {
    case TRANS_FIELD_D: return copy_view(c, blk, *out, f.D, 7, false);
    case TRANS_FIELD_RDS: return copy_view(c, blk, *out, f.Rds, 3, false);
    case TRANS_FIELD_DTM: return copy_view(c, blk, *out, f.Dtm, 1, false);
    case TRANS_FIELD_DQDXYZ: return copy_view(c, blk, *out, f.DQ, 18, false);
    case TRANS_FIELD_SHOCK: return copy_view(c, blk, *out, f.shock, 1, false);
    case TRANS_FIELD_FSCHEME: return copy_view(c, blk, *out, f.Fscheme, 1, false);
    default: return fail(c, TRANS_EINVAL, "unknown field");
    }
}

// ---------------------------------------------------------------------------------------
// stages

static int if_turb_now(const trans_ctx* c)
{
    const trans_params& p = c->p;
    return (p.if_euler == 0 && p.if_turb_org == 1 && (c->ki >= p.kitm || p.kind_hybrid > 0)) ? 1 : 0;
}

static int launch_check(trans_ctx* c, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, TRANS_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

int trans_hip_spectral_radius(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    KLAUNCH(c, "k_spectral_radius", k_spectral_radius, dim3(c->seg_cell.n), dim3(WG), c->p, c->ki, c->f,
                       c->d_blk, c->seg_cell.d);
    return launch_check(c, "k_spectral_radius");
}

int trans_hip_bc(trans_ctx* c, int sweep)
{
    if (!c || (sweep != -1 && sweep != 0)) return TRANS_EINVAL;
    if (c->n_bcseg == 0) return 0;
    KLAUNCH(c, "k_bc", k_bc, dim3(c->n_bcseg), dim3(128), c->p, if_turb_now(c), sweep, c->f,
                       c->d_blk, c->d_patch, c->d_bcseg);
    if (sweep == -1 && c->n_outl)
        KLAUNCH(c, "k_bc_outlet", k_bc_outlet, dim3(c->n_outl), dim3(256), c->p, if_turb_now(c), c->f, c->d_blk,
                c->d_patch, c->d_outl);
    return launch_check(c, "k_bc");
}

}  // extern "C"

// Face kernels of direction d+1 on helper stream d: the three directions are independent
// until k_accumulate_visc, which runs after the join.
template <int SPATIAL, int ROE>
static void launch_conv(trans_ctx* c, int if_turb)
{
    KLAUNCH_ON(c, c->aux[0], "k_conv_face<1>", (k_conv_face<1, SPATIAL, ROE>), dim3(c->seg_face[0].n), dim3(WG),
               c->p, if_turb, c->f, c->d_blk, c->seg_face[0].d);
    KLAUNCH_ON(c, c->aux[1], "k_conv_face<2>", (k_conv_face<2, SPATIAL, ROE>), dim3(c->seg_face[1].n), dim3(WG),
               c->p, if_turb, c->f, c->d_blk, c->seg_face[1].d);
    KLAUNCH_ON(c, c->aux[2], "k_conv_face<3>", (k_conv_face<3, SPATIAL, ROE>), dim3(c->seg_face[2].n), dim3(WG),
               c->p, if_turb, c->f, c->d_blk, c->seg_face[2].d);
}

template <int SPATIAL>
static void launch_conv_roe(trans_ctx* c, int if_turb)
{
    if (c->p.kind_roe == 1)
        launch_conv<SPATIAL, 1>(c, if_turb);
    else if (c->p.kind_roe == 2)
        launch_conv<SPATIAL, 2>(c, if_turb);
    else
        launch_conv<SPATIAL, 3>(c, if_turb);
}

// The LDS-tiled passes for the schemes the workloads use (MUSCL or MDCD with Roe or AUSMPW+);
// false when this scheme runs the face kernels.
static bool launch_fused(trans_ctx* c, int if_turb)
{
    const int sp = c->p.kind_spatial, ro = c->p.kind_roe;
    if (!((sp == 2 || sp == 3) && (ro == 1 || ro == 3))) return false;
    if (c->p.if_euler == 0)
        KLAUNCH(c, "k_resid_visc", k_resid_visc, dim3(c->seg_rtile.n), dim3(256), c->p, if_turb, c->f, c->d_blk,
                c->seg_rtile.d, c->rt_kchunk);
#define RC(S, R)                                                                                                   \
    KLAUNCH(c, "k_resid_conv", (k_resid_conv<S, R>), dim3(c->seg_rtile.n), dim3(256), c->p, c->p.if_euler, if_turb, \
            c->f, c->d_blk, c->seg_rtile.d, c->rt_kchunk)
    if (sp == 2 && ro == 1) RC(2, 1);
    else if (sp == 2 && ro == 3) RC(2, 3);
    else if (sp == 3 && ro == 1) RC(3, 1);
    else RC(3, 3);
#undef RC
    return true;
}

extern "C" {

int trans_hip_space(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    const int if_turb = if_turb_now(c);
    const bool visc = c->p.if_euler == 0;
    int rc;
    // Q (u, v, w, k/rho, omega/rho) of every cell the gradient and face stencils read, formed
    // once: k_prim_cells for the interior cells, k_prim_ghost for the ghost layers
    KLAUNCH(c, "k_prim_cells", k_prim_cells, dim3(c->seg_cell.n), dim3(WG), if_turb, c->f, c->d_blk,
            c->seg_cell.d);
    KLAUNCH(c, "k_prim_ghost", k_prim_ghost, dim3(c->seg_ghost.n), dim3(WG), if_turb, c->f, c->d_blk,
            c->seg_ghost.d);
    if (visc) {
        if (c->p.kind_roe == 2)
            KLAUNCH(c, "k_gradients", k_gradients<true>, dim3(c->seg_cell.n), dim3(WG), c->p, if_turb, c->f,
                    c->d_blk, c->seg_cell.d);
        else
            KLAUNCH(c, "k_gradients", k_gradients<false>, dim3(c->seg_cell.n), dim3(WG), c->p, if_turb, c->f,
                    c->d_blk, c->seg_cell.d);
        if ((rc = launch_check(c, "k_gradients"))) return rc;
    }
    if (visc && (rc = trans_hip_bc(c, 0))) return rc;
    if (c->fused && launch_fused(c, if_turb)) return launch_check(c, "space");
    fork_aux(c, 3);
    if (visc) {
        KLAUNCH_ON(c, c->aux[0], "k_visc_face<1>", k_visc_face<1>, dim3(c->seg_face[0].n), dim3(WG), c->p, if_turb,
                   c->f, c->d_blk, c->seg_face[0].d);
        KLAUNCH_ON(c, c->aux[1], "k_visc_face<2>", k_visc_face<2>, dim3(c->seg_face[1].n), dim3(WG), c->p, if_turb,
                   c->f, c->d_blk, c->seg_face[1].d);
        KLAUNCH_ON(c, c->aux[2], "k_visc_face<3>", k_visc_face<3>, dim3(c->seg_face[2].n), dim3(WG), c->p, if_turb,
                   c->f, c->d_blk, c->seg_face[2].d);
    }
    if (c->p.kind_spatial == 1)
        launch_conv_roe<1>(c, if_turb);
    else if (c->p.kind_spatial == 2)
        launch_conv_roe<2>(c, if_turb);
    else if (c->p.kind_spatial == 3)
        launch_conv_roe<3>(c, if_turb);
    else
        launch_conv_roe<4>(c, if_turb);
    join_aux(c, 3);
    KLAUNCH(c, "k_accumulate_visc", k_accumulate_visc, dim3(c->seg_cell.n), dim3(WG), c->p, c->p.if_euler,
            if_turb, c->f, c->d_blk, c->seg_cell.d);
    return launch_check(c, "space");
}

int trans_hip_residual(trans_ctx* c)
{
    int rc;
    // The outlet march reads interior cells only and writes only its own ghosts, so it runs
    // beside the spectral radius and the other patches' BC (which copies its Rds ghosts).
    const bool outl = c->n_outl > 0;
    if (outl) {
        fork_aux(c, 1);
        KLAUNCH_ON(c, c->aux[0], "k_bc_outlet", k_bc_outlet, dim3(c->n_outl), dim3(256), c->p, if_turb_now(c), c->f,
                   c->d_blk, c->d_patch, c->d_outl);
    }
    if ((rc = trans_hip_spectral_radius(c))) return rc;
    if (c->n_bcseg)
        KLAUNCH(c, "k_bc", k_bc, dim3(c->n_bcseg), dim3(128), c->p, if_turb_now(c), -1, c->f, c->d_blk, c->d_patch,
                c->d_bcseg);
    if (outl) join_aux(c, 1);
    if ((rc = launch_check(c, "k_bc"))) return rc;
    if ((rc = trans_hip_space(c))) return rc;
    return 0;
}

}  // extern "C"

template <typename R, int NEQ, int MAXT, int MARCH>
static void launch_lines(trans_ctx* c, R* dq)
{
    const int nblk = (int)c->hb.size();
    // LDS: 2 parities x 2 directions x (5 or 2) equations x (lines + 2) values
    const size_t lds = (size_t)4 * (NEQ == 2 ? 2 : 5) * (c->max_lines + 2) * sizeof(R);
    const dim3 grid(nblk, NEQ == 7 ? 2 : 1);
    const bool kw = NEQ == 2;  // k-omega sweeps alone
    KLAUNCH_LDS(c, kw ? "k_turb_line<fwd>" : "k_lusgs_line<fwd>", (k_lusgs_line<R, true, NEQ, MAXT, MARCH>), grid,
                dim3(c->line_threads), lds, c->f, dq, c->d_blk);
    KLAUNCH_LDS(c, kw ? "k_turb_line<bwd>" : "k_lusgs_line<bwd>", (k_lusgs_line<R, false, NEQ, MAXT, MARCH>), grid,
                dim3(c->line_threads), lds, c->f, dq, c->d_blk);
}

template <typename R, int NEQ, int MAXT>
static void launch_lines(trans_ctx* c, R* dq)
{
    if (c->march == 0)
        launch_lines<R, NEQ, MAXT, 0>(c, dq);
    else if (c->march == 1)
        launch_lines<R, NEQ, MAXT, 1>(c, dq);
    else
        launch_lines<R, NEQ, MAXT, 2>(c, dq);
}

template <typename R, int NEQ>
static void launch_sweeps(trans_ctx* c, R* dq)
{
    const int nblk = (int)c->hb.size();
    if (c->line_threads) {
        // the register budget per lane follows the workgroup size
        if (c->line_threads <= 256)
            launch_lines<R, NEQ, 256>(c, dq);
        else if (c->line_threads <= 512)
            launch_lines<R, NEQ, 512>(c, dq);
        else
            launch_lines<R, NEQ, 1024>(c, dq);
    } else if (c->tile_mode) {
        const int w = NEQ == 7 ? 0 : (NEQ == 5 ? 1 : 2);
        const int nt = c->ntiles[w];
        const int ntmax = std::max(c->ntiles[0], std::max(c->ntiles[1], c->ntiles[2]));
        const int nctr = 2 + 2 * ntmax;
        const int grid = std::max(1, std::min(c->tile_grid_k[w], nt));
        R* edge = (R*)c->tile_edge;
        KLAUNCH(c, "k_tile_reset", k_tile_reset, dim3((nctr + 255) / 256), dim3(256), c->d_tctr, nctr);
        constexpr int KIND = NEQ == 7 ? 0 : NEQ;
        KLAUNCH(c, "k_lusgs_tile<fwd>", (k_lusgs_tile<R, true, KIND>), dim3(grid), dim3(TILE_MAXT), c->f, dq,
                c->d_blk, c->d_tiles[w], nt, c->d_tctr, c->d_tctr + 2, edge, c->d_err, c->tile_chunk,
                TILE_SPIN_LIMIT);
        KLAUNCH(c, "k_lusgs_tile<bwd>", (k_lusgs_tile<R, false, KIND>), dim3(grid), dim3(TILE_MAXT), c->f, dq,
                c->d_blk, c->d_tiles[w], nt, c->d_tctr + 1, c->d_tctr + 2 + ntmax, edge, c->d_err, c->tile_chunk,
                TILE_SPIN_LIMIT);
    } else if (c->plane_launch) {
        const bool kw = NEQ == 2;
        for (int s0 = 0; s0 < c->max_planes; s0++)
            KLAUNCH(c, kw ? "k_turb_wplane<fwd>" : "k_lusgs_wplane<fwd>", (k_lusgs_wplane<R, true, NEQ>),
                    dim3(c->seg_plane.n), dim3(WG), c->f, dq, c->d_blk, c->seg_plane.d, c->d_pblk, s0);
        for (int s0 = 0; s0 < c->max_planes; s0++)
            KLAUNCH(c, kw ? "k_turb_wplane<bwd>" : "k_lusgs_wplane<bwd>", (k_lusgs_wplane<R, false, NEQ>),
                    dim3(c->seg_plane.n), dim3(WG), c->f, dq, c->d_blk, c->seg_plane.d, c->d_pblk, s0);
    } else {
        const bool kw = NEQ == 2;
        KLAUNCH(c, kw ? "k_turb_sweep<fwd>" : "k_lusgs_sweep<fwd>", (k_lusgs_sweep<R, true, NEQ>), dim3(nblk),
                dim3(1024), c->p, c->f, dq, c->d_blk);
        KLAUNCH(c, kw ? "k_turb_sweep<bwd>" : "k_lusgs_sweep<bwd>", (k_lusgs_sweep<R, false, NEQ>), dim3(nblk),
                dim3(1024), c->p, c->f, dq, c->d_blk);
    }
}

extern "C" {

// LU-SGS (Time/TIME_LUSGS.F90, Time/TURBTWO_TIME_LUSGS_SST.F90) and Update_V (main.F90:135-150)
static int lusgs_update_async(trans_ctx* c, int shift = 0)
{
    const int nblk = (int)c->hb.size();
    const int if_turb = if_turb_now(c);
    // the sweep buffer along j moves through LDS tiles (whole cache lines per diagonal)
    const bool wtiled = (c->f.W || c->f.Wf) && c->f.march == 1;
    if (wtiled && c->f.Wf)
        KLAUNCH(c, "k_lusgs_rhs_w", k_lusgs_rhs_w<float>, dim3(c->seg_wtile.n), dim3(WG), c->p, c->f, c->d_blk,
                c->seg_wtile.d);
    else if (wtiled)
        KLAUNCH(c, "k_lusgs_rhs_w", k_lusgs_rhs_w<double>, dim3(c->seg_wtile.n), dim3(WG), c->p, c->f, c->d_blk,
                c->seg_wtile.d);
    else
        KLAUNCH(c, "k_lusgs_rhs", k_lusgs_rhs, dim3(c->seg_cell.n), dim3(WG), c->p, c->f, c->d_blk, c->seg_cell.d);
    if (if_turb) {
        if (wtiled)
            KLAUNCH(c, "k_sst_rhs_w", k_sst_rhs_w, dim3(c->seg_wtile.n), dim3(WG), c->p, c->f, c->d_blk,
                    c->seg_wtile.d);
        else
            KLAUNCH(c, "k_sst_rhs", k_sst_rhs, dim3(c->seg_cell.n), dim3(WG), c->p, c->f, c->d_blk, c->seg_cell.d);
    }
    const int neq = if_turb ? 7 : 5;
    if (c->f.W || c->f.Wf) {
        // line-march sweeps on the hyperplane-major buffer, increments back to D after
        if (c->p.implicit_fp32) {
            if (if_turb)
                launch_sweeps<float, 7>(c, c->f.Wf);
            else
                launch_sweeps<float, 5>(c, c->f.Wf);
            if (wtiled)
                KLAUNCH(c, "k_unskew_w", k_unskew_w<float>, dim3(c->seg_wtile.n), dim3(WG), 0, neq, c->f,
                        (const float*)c->f.Wf, c->d_blk, c->seg_wtile.d);
            else
                KLAUNCH(c, "k_unskew", k_unskew<float>, dim3(c->seg_cell.n), dim3(WG), 0, neq, c->f,
                        (const float*)c->f.Wf, c->d_blk, c->seg_cell.d);
        } else {
            if (if_turb)
                launch_sweeps<double, 7>(c, c->f.W);
            else
                launch_sweeps<double, 5>(c, c->f.W);
            if (wtiled)
                KLAUNCH(c, "k_unskew_w", k_unskew_w<double>, dim3(c->seg_wtile.n), dim3(WG), 0, neq, c->f,
                        (const double*)c->f.W, c->d_blk, c->seg_wtile.d);
            else
                KLAUNCH(c, "k_unskew", k_unskew<double>, dim3(c->seg_cell.n), dim3(WG), 0, neq, c->f,
                        (const double*)c->f.W, c->d_blk, c->seg_cell.d);
        }
    } else if (c->p.implicit_fp32) {
        KLAUNCH(c, "k_convert<double,float>", (k_convert<double, float>), dim3(c->seg_cell.n), dim3(WG), neq, 0,
                c->tot, (const double*)c->f.D, c->f.Ds, c->d_blk, c->seg_cell.d);
        if (if_turb)
            launch_sweeps<float, 7>(c, c->f.Ds);
        else
            launch_sweeps<float, 5>(c, c->f.Ds);
        KLAUNCH(c, "k_convert<float,double>", (k_convert<float, double>), dim3(c->seg_cell.n), dim3(WG), neq, 0,
                c->tot, (const float*)c->f.Ds, c->f.D, c->d_blk, c->seg_cell.d);
    } else {
        if (if_turb)
            launch_sweeps<double, 7>(c, c->f.D);
        else
            launch_sweeps<double, 5>(c, c->f.D);
    }
    KLAUNCH(c, "k_update_v", k_update_v, dim3(c->seg_cell.n), dim3(WG), c->p, c->ki, if_turb, c->f, c->d_blk,
            c->seg_cell.d, c->d_part, shift);
    KLAUNCH(c, "k_norm_final", k_norm_final, dim3(nblk), dim3(256), c->d_part, c->d_wrange, c->d_blk, c->d_resid);
    return launch_check(c, "lusgs_update");
}

// Runge-Kutta stage itr_chld (main.F90:158-170): TIME_RK_Save and TIME_RK, then with k-omega
// TURBTWO_TIME_LUSGS_SST on the stage's state and TIME_RK_V_Update_New.
static int rk_update_async(trans_ctx* c, int stage)
{
    const int nblk = (int)c->hb.size();
    const int if_turb = if_turb_now(c);
    // the sweep buffer along j moves through LDS tiles (whole cache lines per diagonal)
    const bool wtiled = (c->f.W || c->f.Wf) && c->f.march == 1;
    KLAUNCH(c, "k_rk_stage", k_rk_stage, dim3(c->seg_cell.n), dim3(WG), c->p, stage, if_turb ? 0 : 1, wtiled ? 0 : 1,
            c->f, c->d_blk, c->seg_cell.d, c->d_part);
    if (if_turb) {
        if (wtiled) {
            KLAUNCH(c, "k_rk_w", k_rk_w, dim3(c->seg_wtile.n), dim3(WG), c->p, c->f, c->d_blk, c->seg_wtile.d);
            KLAUNCH(c, "k_sst_rhs_w", k_sst_rhs_w, dim3(c->seg_wtile.n), dim3(WG), c->p, c->f, c->d_blk,
                    c->seg_wtile.d);
        } else {
            KLAUNCH(c, "k_sst_rhs", k_sst_rhs, dim3(c->seg_cell.n), dim3(WG), c->p, c->f, c->d_blk, c->seg_cell.d);
        }
        if (c->f.W || c->f.Wf) {
            if (c->p.implicit_fp32) {
                launch_sweeps<float, 2>(c, c->f.Wf);
                if (wtiled)
                    KLAUNCH(c, "k_unskew_w", k_unskew_w<float>, dim3(c->seg_wtile.n), dim3(WG), 5, 2, c->f,
                            (const float*)c->f.Wf, c->d_blk, c->seg_wtile.d);
                else
                    KLAUNCH(c, "k_unskew", k_unskew<float>, dim3(c->seg_cell.n), dim3(WG), 5, 2, c->f,
                            (const float*)c->f.Wf, c->d_blk, c->seg_cell.d);
            } else {
                launch_sweeps<double, 2>(c, c->f.W);
                if (wtiled)
                    KLAUNCH(c, "k_unskew_w", k_unskew_w<double>, dim3(c->seg_wtile.n), dim3(WG), 5, 2, c->f,
                            (const double*)c->f.W, c->d_blk, c->seg_wtile.d);
                else
                    KLAUNCH(c, "k_unskew", k_unskew<double>, dim3(c->seg_cell.n), dim3(WG), 5, 2, c->f,
                            (const double*)c->f.W, c->d_blk, c->seg_cell.d);
            }
        } else if (c->p.implicit_fp32) {
            KLAUNCH(c, "k_convert<double,float>", (k_convert<double, float>), dim3(c->seg_cell.n), dim3(WG), 2, 5,
                    c->tot, (const double*)c->f.D, c->f.Ds, c->d_blk, c->seg_cell.d);
            launch_sweeps<float, 2>(c, c->f.Ds);
            KLAUNCH(c, "k_convert<float,double>", (k_convert<float, double>), dim3(c->seg_cell.n), dim3(WG), 2, 5,
                    c->tot, (const float*)c->f.Ds, c->f.D, c->d_blk, c->seg_cell.d);
        } else {
            launch_sweeps<double, 2>(c, c->f.D);
        }
        KLAUNCH(c, "k_rk_update_v", k_rk_update_v, dim3(c->seg_cell.n), dim3(WG), c->p, c->ki, if_turb, c->f,
                c->d_blk, c->seg_cell.d, c->d_part);
    }
    KLAUNCH(c, "k_norm_final", k_norm_final, dim3(nblk), dim3(256), c->d_part, c->d_wrange, c->d_blk, c->d_resid);
    return launch_check(c, "rk_update");
}

// The error flag of the tile wavefront's bounded waits (read once the stream is idle).
static int check_sweep_flag(trans_ctx* c)
{
    if (!c->d_err) return 0;
    int e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!e) return 0;
    HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return fail(c, TRANS_ETIMEOUT, "a bounded wait of the tile-wavefront LU-SGS sweep expired; results are invalid");
}

static int fetch_resid(trans_ctx* c, trans_resid* per_block)
{
    HIPCHK(c, hipMemcpyAsync(per_block, c->d_resid, c->hb.size() * sizeof(trans_resid), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (int rc = check_sweep_flag(c)) return rc;
    for (size_t b = 0; b < c->hb.size(); b++)
        if (per_block[b].drsa != per_block[b].drsa)
            return fail(c, TRANS_ENONFINITE, "residual of block " + std::to_string(b + 1) + " is NaN");
    return 0;
}

int trans_hip_rk_update(trans_ctx* c, int itr_chld, trans_resid* per_block)
{
    if (!c) return TRANS_EINVAL;
    if (itr_chld < 1 || itr_chld > 3) return fail(c, TRANS_EINVAL, "Runge-Kutta stage must be 1, 2 or 3");
    int rc = rk_update_async(c, itr_chld);
    if (rc) return rc;
    return per_block ? fetch_resid(c, per_block) : 0;
}

int trans_hip_lusgs_update(trans_ctx* c, trans_resid* per_block)
{
    if (!c) return TRANS_EINVAL;
    int rc = lusgs_update_async(c);
    if (rc) return rc;
    return per_block ? fetch_resid(c, per_block) : 0;
}

int trans_hip_update_u(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    KLAUNCH(c, "k_update_u", k_update_u, dim3(c->seg_cell.n), dim3(WG), c->p.kind_dual, c->f, c->d_blk,
                       c->seg_cell.d);
    return launch_check(c, "k_update_u");
}

// The device work of one outer iteration (main.F90:65-197 without I/O), queued only.
static int average_fields_async(trans_ctx* c);  // io.h

}  // extern "C"

// The plain drop-in of main.F90:65-197 (a baseline, selected by set_tuning with sweep -3): each
// routine is one launch per block, as a straight port of the reference's per-block calls would
// run, and the LU-SGS sweeps (mean flow and k-omega, or the k-omega sweeps of an RK stage) are
// one launch per hyperplane per block.
template <typename R>
static void plain_sweeps(trans_ctx* c, R* W, int neq)
{
    if (neq == 2) {
        for (int b = 0; b < (int)c->hb.size(); b++) {
            const HostBlock& h = c->hb[b];
            const int np = (h.ni - 1) + (h.nj - 1) + (h.nk - 1) - 2;
            const int g0 = c->seg_plane.wrange[b], ng = c->seg_plane.wrange[b + 1] - g0;
            for (int s0 = 0; s0 < np; s0++)
                KLAUNCH(c, "k_turb_wplane<fwd>", (k_lusgs_wplane<R, true, 2>), dim3(ng), dim3(WG), c->f, W, c->d_blk,
                        c->seg_plane.d + g0, c->d_pblk, s0);
            for (int s0 = 0; s0 < np; s0++)
                KLAUNCH(c, "k_turb_wplane<bwd>", (k_lusgs_wplane<R, false, 2>), dim3(ng), dim3(WG), c->f, W,
                        c->d_blk, c->seg_plane.d + g0, c->d_pblk, s0);
        }
        return;
    }
    for (int b = 0; b < (int)c->hb.size(); b++) {
        const HostBlock& h = c->hb[b];
        const int np = (h.ni - 1) + (h.nj - 1) + (h.nk - 1) - 2;
        const int g0 = c->seg_plane.wrange[b], ng = c->seg_plane.wrange[b + 1] - g0;
        for (int fwd = 1; fwd >= 0; fwd--)
            for (int s0 = 0; s0 < np; s0++) {
                if (neq == 7 && fwd)
                    KLAUNCH(c, "k_lusgs_wplane<fwd>", (k_lusgs_wplane<R, true, 7>), dim3(ng), dim3(WG), c->f, W,
                            c->d_blk, c->seg_plane.d + g0, c->d_pblk, s0);
                else if (neq == 7)
                    KLAUNCH(c, "k_lusgs_wplane<bwd>", (k_lusgs_wplane<R, false, 7>), dim3(ng), dim3(WG), c->f, W,
                            c->d_blk, c->seg_plane.d + g0, c->d_pblk, s0);
                else if (fwd)
                    KLAUNCH(c, "k_lusgs_wplane<fwd>", (k_lusgs_wplane<R, true, 5>), dim3(ng), dim3(WG), c->f, W,
                            c->d_blk, c->seg_plane.d + g0, c->d_pblk, s0);
                else
                    KLAUNCH(c, "k_lusgs_wplane<bwd>", (k_lusgs_wplane<R, false, 5>), dim3(ng), dim3(WG), c->f, W,
                            c->d_blk, c->seg_plane.d + g0, c->d_pblk, s0);
            }
    }
}

template <int SPATIAL, int ROE>
static void plain_conv(trans_ctx* c, int if_turb, int b)
{
    const SegTable* t = c->seg_face;
    KLAUNCH(c, "k_conv_face<1>", (k_conv_face<1, SPATIAL, ROE>), dim3(t[0].wrange[b + 1] - t[0].wrange[b]), dim3(WG),
            c->p, if_turb, c->f, c->d_blk, t[0].d + t[0].wrange[b]);
    KLAUNCH(c, "k_conv_face<2>", (k_conv_face<2, SPATIAL, ROE>), dim3(t[1].wrange[b + 1] - t[1].wrange[b]), dim3(WG),
            c->p, if_turb, c->f, c->d_blk, t[1].d + t[1].wrange[b]);
    KLAUNCH(c, "k_conv_face<3>", (k_conv_face<3, SPATIAL, ROE>), dim3(t[2].wrange[b + 1] - t[2].wrange[b]), dim3(WG),
            c->p, if_turb, c->f, c->d_blk, t[2].d + t[2].wrange[b]);
}

extern "C" {

static int plain_body(trans_ctx* c)
{
    const int nblk = (int)c->hb.size();
    const int if_turb = if_turb_now(c);
    const bool visc = c->p.if_euler == 0;
    const int neq = if_turb ? 7 : 5;
    const bool rk = c->p.if_rk == 1;
    const int nsub = rk ? 3 : (c->p.kind_dual != 0 ? c->p.kchld : 1);
    auto cells = [&](const SegTable& t, int b) { return dim3(t.wrange[b + 1] - t.wrange[b]); };
    for (int s = 0; s < nsub; s++) {
        for (int b = 0; b < nblk; b++)
            KLAUNCH(c, "k_spectral_radius", k_spectral_radius, cells(c->seg_cell, b), dim3(WG), c->p, c->ki, c->f,
                    c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
        for (int b = 0; b < nblk; b++) {
            if (c->bc_wrange[b + 1] > c->bc_wrange[b])
                KLAUNCH(c, "k_bc", k_bc, dim3(c->bc_wrange[b + 1] - c->bc_wrange[b]), dim3(128), c->p, if_turb, -1, c->f,
                        c->d_blk, c->d_patch, c->d_bcseg + c->bc_wrange[b]);
            if (c->outl_wrange[b + 1] > c->outl_wrange[b])
                KLAUNCH(c, "k_bc_outlet", k_bc_outlet, dim3(c->outl_wrange[b + 1] - c->outl_wrange[b]), dim3(256), c->p,
                        if_turb, c->f, c->d_blk, c->d_patch, c->d_outl + c->outl_wrange[b]);
        }
        for (int b = 0; b < nblk; b++) {
            KLAUNCH(c, "k_prim_cells", k_prim_cells, cells(c->seg_cell, b), dim3(WG), if_turb, c->f, c->d_blk,
                    c->seg_cell.d + c->seg_cell.wrange[b]);
            KLAUNCH(c, "k_prim_ghost", k_prim_ghost, cells(c->seg_ghost, b), dim3(WG), if_turb, c->f, c->d_blk,
                    c->seg_ghost.d + c->seg_ghost.wrange[b]);
            if (visc && c->p.kind_roe == 2)
                KLAUNCH(c, "k_gradients", k_gradients<true>, cells(c->seg_cell, b), dim3(WG), c->p, if_turb, c->f,
                        c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
            else if (visc)
                KLAUNCH(c, "k_gradients", k_gradients<false>, cells(c->seg_cell, b), dim3(WG), c->p, if_turb, c->f,
                        c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
        }
        if (visc) {
            for (int b = 0; b < nblk; b++)
                if (c->bc_wrange[b + 1] > c->bc_wrange[b])
                    KLAUNCH(c, "k_bc", k_bc, dim3(c->bc_wrange[b + 1] - c->bc_wrange[b]), dim3(128), c->p, if_turb, 0,
                            c->f, c->d_blk, c->d_patch, c->d_bcseg + c->bc_wrange[b]);
            for (int b = 0; b < nblk; b++) {
                KLAUNCH(c, "k_visc_face<1>", k_visc_face<1>, cells(c->seg_face[0], b), dim3(WG), c->p, if_turb, c->f,
                        c->d_blk, c->seg_face[0].d + c->seg_face[0].wrange[b]);
                KLAUNCH(c, "k_visc_face<2>", k_visc_face<2>, cells(c->seg_face[1], b), dim3(WG), c->p, if_turb, c->f,
                        c->d_blk, c->seg_face[1].d + c->seg_face[1].wrange[b]);
                KLAUNCH(c, "k_visc_face<3>", k_visc_face<3>, cells(c->seg_face[2], b), dim3(WG), c->p, if_turb, c->f,
                        c->d_blk, c->seg_face[2].d + c->seg_face[2].wrange[b]);
            }
        }
        for (int b = 0; b < nblk; b++) {
            const int sp = c->p.kind_spatial, ro = c->p.kind_roe;
            if (sp == 2 && ro == 1) plain_conv<2, 1>(c, if_turb, b);
            else if (sp == 2 && ro == 3) plain_conv<2, 3>(c, if_turb, b);
            else if (sp == 3 && ro == 1) plain_conv<3, 1>(c, if_turb, b);
            else if (sp == 3 && ro == 3) plain_conv<3, 3>(c, if_turb, b);
            else return fail(c, TRANS_EUNSUPPORTED, "the plain drop-in runs MUSCL or MDCD with Roe or AUSMPW+");
        }
        for (int b = 0; b < nblk; b++)
            KLAUNCH(c, "k_accumulate_visc", k_accumulate_visc, cells(c->seg_cell, b), dim3(WG), c->p, c->p.if_euler,
                    if_turb, c->f, c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
        if (rk) {  // main.F90:158-170: stage s + 1 of the three-stage Runge-Kutta update
            for (int b = 0; b < nblk; b++)
                KLAUNCH(c, "k_rk_stage", k_rk_stage, cells(c->seg_cell, b), dim3(WG), c->p, s + 1, if_turb ? 0 : 1, 1,
                        c->f, c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b], c->d_part + c->seg_cell.wrange[b]);
            if (if_turb) {
                for (int b = 0; b < nblk; b++)
                    KLAUNCH(c, "k_sst_rhs", k_sst_rhs, cells(c->seg_cell, b), dim3(WG), c->p, c->f, c->d_blk,
                            c->seg_cell.d + c->seg_cell.wrange[b]);
                if (c->p.implicit_fp32)
                    plain_sweeps<float>(c, c->f.Wf, 2);
                else
                    plain_sweeps<double>(c, c->f.W, 2);
                for (int b = 0; b < nblk; b++) {
                    if (c->p.implicit_fp32)
                        KLAUNCH(c, "k_unskew", k_unskew<float>, cells(c->seg_cell, b), dim3(WG), 5, 2, c->f,
                                (const float*)c->f.Wf, c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
                    else
                        KLAUNCH(c, "k_unskew", k_unskew<double>, cells(c->seg_cell, b), dim3(WG), 5, 2, c->f,
                                (const double*)c->f.W, c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b]);
                    KLAUNCH(c, "k_rk_update_v", k_rk_update_v, cells(c->seg_cell, b), dim3(WG), c->p, c->ki, if_turb,
                            c->f, c->d_blk, c->seg_cell.d + c->seg_cell.wrange[b], c->d_part + c->seg_cell.wrange[b]);
                }
            }
            KLAUNCH(c, "k_norm_final", k_norm_final, dim3(nblk), dim3(256), c->d_part, c->d_wrange, c->d_blk,
                    c->d_resid);
            continue;
        }
        for (int b = 0; b < nblk; b++) {
            if (c->f.Wf)
                KLAUNCH(c, "k_lusgs_rhs_w", k_lusgs_rhs_w<float>, cells(c->seg_wtile, b), dim3(WG), c->p, c->f,
                        c->d_blk, c->seg_wtile.d + c->seg_wtile.wrange[b]);
            else
                KLAUNCH(c, "k_lusgs_rhs_w", k_lusgs_rhs_w<double>, cells(c->seg_wtile, b), dim3(WG), c->p, c->f,
                        c->d_blk, c->seg_wtile.d + c->seg_wtile.wrange[b]);
            if (if_turb)
                KLAUNCH(c, "k_sst_rhs_w", k_sst_rhs_w, cells(c->seg_wtile, b), dim3(WG), c->p, c->f, c->d_blk,
                        c->seg_wtile.d + c->seg_wtile.wrange[b]);
        }
        if (c->p.implicit_fp32)
            plain_sweeps<float>(c, c->f.Wf, neq);
        else
            plain_sweeps<double>(c, c->f.W, neq);
        const bool fuse = c->p.kind_dual == 1 || c->p.kind_dual == 2;
        for (int b = 0; b < nblk; b++) {
            if (c->p.implicit_fp32)
                KLAUNCH(c, "k_unskew_w", k_unskew_w<float>, cells(c->seg_wtile, b), dim3(WG), 0, neq, c->f,
                        (const float*)c->f.Wf, c->d_blk, c->seg_wtile.d + c->seg_wtile.wrange[b]);
            else
                KLAUNCH(c, "k_unskew_w", k_unskew_w<double>, cells(c->seg_wtile, b), dim3(WG), 0, neq, c->f,
                        (const double*)c->f.W, c->d_blk, c->seg_wtile.d + c->seg_wtile.wrange[b]);
            KLAUNCH(c, "k_update_v", k_update_v, cells(c->seg_cell, b), dim3(WG), c->p, c->ki, if_turb, c->f, c->d_blk,
                    c->seg_cell.d + c->seg_cell.wrange[b], c->d_part + c->seg_cell.wrange[b],
                    fuse && s == nsub - 1 ? c->p.kind_dual : 0);
        }
        KLAUNCH(c, "k_norm_final", k_norm_final, dim3(nblk), dim3(256), c->d_part, c->d_wrange, c->d_blk, c->d_resid);
    }
    if (rk && c->p.kind_dual != 0)  // main.F90:192-197 after the Runge-Kutta stages
        for (int b = 0; b < nblk; b++)
            KLAUNCH(c, "k_update_u", k_update_u, cells(c->seg_cell, b), dim3(WG), c->p.kind_dual, c->f, c->d_blk,
                    c->seg_cell.d + c->seg_cell.wrange[b]);
    return launch_check(c, "plain iteration");
}

static int iteration_body(trans_ctx* c)
{
    int rc;
    if (c->plain) return plain_body(c);
    const int nsub = c->p.if_rk == 1 ? 3 : (c->p.kind_dual != 0 ? c->p.kchld : 1);
    // with LU-SGS the dual-time shift of the last sub-iteration is fused into its update
    const bool fuse = c->p.if_rk != 1 && (c->p.kind_dual == 1 || c->p.kind_dual == 2);
    for (int s = 0; s < nsub; s++) {
        if ((rc = trans_hip_residual(c))) return rc;
        if ((rc = c->p.if_rk == 1 ? rk_update_async(c, s + 1)
                                  : lusgs_update_async(c, fuse && s == nsub - 1 ? c->p.kind_dual : 0)))
            return rc;
    }
    if (c->p.kind_dual != 0 && !fuse)
        if ((rc = trans_hip_update_u(c))) return rc;
    return 0;
}

// KI enters the kernels only through the CFL ramp of the local step (constant once the
// ramp has reached a non-decreasing end value, or without dual time) and the turbulence
// switch (constant from KItm on), so from then on every iteration launches the same
// kernels with the same arguments.
static bool iteration_is_static(const trans_ctx* c)
{
    const trans_params& p = c->p;
    const bool ramp = p.kind_dual == 0 || (p.cfl_sub_end >= p.cfl_sub_start && c->ki - p.ki_ini >= p.cfl_sub_step);
    const bool turb = !(p.if_euler == 0 && p.if_turb_org == 1) || p.kind_hybrid > 0 || c->ki >= p.kitm;
    return ramp && turb;
}

// Capture one iteration into c->gexec; on any failure the iteration runs eagerly instead.
static int capture_iteration(trans_ctx* c)
{
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
    const int rc = iteration_body(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return -1;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return -1;
    }
    c->graph = g;
    c->gexec = x;
    return 0;
}

int trans_hip_iterate(trans_ctx* c, int n_iter, trans_resid* per_block)
{
    if (!c || n_iter < 0) return TRANS_EINVAL;
    int rc;
    for (int it = 0; it < n_iter; it++) {
        const bool replay = c->graphs && !c->plain && !c->prof && iteration_is_static(c);
        if (replay && !c->gexec && capture_iteration(c) != 0) c->graphs = false;  // eager from now on
        if (replay && c->gexec) {
            HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        } else {
            drop_graph(c);
            if ((rc = iteration_body(c))) return rc;
        }
        if (c->stats_on && (rc = average_fields_async(c))) return rc;  // main.F90:185-188
        c->ki += 1;
    }
    return per_block ? fetch_resid(c, per_block) : 0;
}

int trans_hip_profile(trans_ctx* c, int on)
{
    if (!c) return TRANS_EINVAL;
    prof_flush(c);
    c->prof = on != 0;
    if (on) {
        c->prof_names.clear();
        c->prof_ms.clear();
        c->prof_n.clear();
    }
    return 0;
}

int trans_hip_profile_count(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    prof_flush(c);
    return (int)c->prof_names.size();
}

int trans_hip_profile_get(trans_ctx* c, int idx, char* name, int name_len, double* total_ms, long* launches)
{
    if (!c || idx < 0) return TRANS_EINVAL;
    prof_flush(c);
    if (idx >= (int)c->prof_names.size()) return TRANS_EINVAL;
    if (name && name_len > 0) {
        std::strncpy(name, c->prof_names[idx].c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = c->prof_ms[idx];
    if (launches) *launches = c->prof_n[idx];
    return 0;
}

int trans_hip_plane_integrals(trans_ctx* c, int blk, int dir, int index, trans_plane_integrals* out)
{
    if (!c || !out) return TRANS_EINVAL;
    if (blk < 0 || blk >= (int)c->hb.size()) return fail(c, TRANS_EINVAL, "block out of range");
    if (dir < 1 || dir > 3) return fail(c, TRANS_EINVAL, "dir must be 1, 2 or 3");
    const HostBlock& h = c->hb[blk];
    const int n = dir == 1 ? h.ni : (dir == 2 ? h.nj : h.nk);
    if (index < 1 || index > n) return fail(c, TRANS_EINVAL, "face index out of range");
    BlkInfo b{h.ni, h.nj, h.nk, h.nx, h.ny, h.nz, h.off, 0, 0, 0, 0};
    KLAUNCH(c, "k_plane_integrals", k_plane_integrals, dim3(1), dim3(256), c->f, b, dir, index, c->p.omega1,
            c->d_mon);
    double r[6];
    HIPCHK(c, hipMemcpyAsync(r, c->d_mon, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->mass_flow = r[0];
    out->area = r[1];
    out->p0_mass_avg = r[2];
    out->t0_mass_avg = r[3];
    out->p0_abs_mass_avg = r[4];
    out->t0_abs_mass_avg = r[5];
    return 0;
}

int trans_hip_sync(trans_ctx* c)
{
    if (!c) return TRANS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_sweep_flag(c);
}

}  // extern "C"

#include "serve.h"
#include "io.h"
#include "tune.h"
#include "geom.h"
#include "tecplot.h"
