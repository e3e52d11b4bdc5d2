// LU-SGS sweeps (Time/TIME_LUSGS.F90:172-436 forward, :442-675 backward, and the k-omega
// sweeps of Time/TURBTWO_TIME_LUSGS_SST.F90:550-766) as a tile wavefront over many CUs.
//
// A block's lines along j (march axis j of the hyperplane-major buffer, lines (i, k)) are cut
// into tiles of na x nb lines; one workgroup sweeps one tile, lane (ia, ib) marching its line
// as the line-march kernel does: at global step t it updates its cell on hyperplane
// s = (i-1)+(j-1)+(k-1) (forward) and hands the terms 0.5*(A dQ +/- Rds dQ) of its downstream
// neighbours on through LDS.  Terms that cross a tile boundary go through the tile's edge
// buffer in global memory instead, and the downstream tile reads them one step later.
//
// Scheduling: every tile of every block (mean flow and k-omega separately, they are
// independent) is one work item, in an order where a tile's upstream tiles come first (sorted
// by tile diagonal; the backward sweep takes the items in reverse).  Workgroups take items from
// an atomic counter, so an item's upstream items have always been taken by running
// workgroups, whatever the grid size: no wait can block on a workgroup that has not started.
// A tile publishes its progress (global steps done) every `chunk` steps; before a chunk a tile
// waits until its upstream tiles have published the steps the chunk reads.  Every wait is
// bounded: past `spin_limit` polls the workgroup raises the error flag and goes on (the host
// reports the failure), so the kernel always ends.
//
// Hand-off without cache maintenance: the edge terms and the progress counters are the only
// data one workgroup reads from another during the kernel, and they move as agent-scope
// atomics (loads and stores that bypass the non-coherent per-XCD L2, `sc1`).  A tile waits for
// its own edge stores to complete (s_waitcnt) before the barrier after which it publishes
// progress; the consumer polls the counter, then issues its edge loads.  No L2 write-back or
// invalidate is needed, which an agent-scope release/acquire fence would do on every chunk.
//
// Per cell the arithmetic is the line march's, in the same order; the result is the
// reference's lexicographic Gauss-Seidel up to the FMA rounding of the implicit operator.
#pragma once
#include "k_time.h"

struct SweepTile {
    int blk;
    int neq;          // 5 mean flow, 2 k-omega
    int a0, b0;       // first line of the tile along a (i) and b (k), 0-based
    int na, nb;       // lines of the tile along a and b
    int lo_a, hi_a;   // tiles of the same block and kind before / after along a (-1: none)
    int lo_b, hi_b;   // ... along b
    long ea, eb;      // offsets of the tile's outgoing edge buffers: terms crossing its a / b edge
};

constexpr int TILE_MAXT = 256;       // lanes of a tile workgroup (na * nb <= TILE_MAXT)
constexpr int TILE_GRID_MAX = 2048;  // most persistent workgroups (their sink slots are reserved)
constexpr int TILE_SPIN_LIMIT = 1 << 22;  // polls of one bounded wait (~0.2 s) before the error flag

// Zeroes the item counters and progress counters of both sweeps (no memset: the sweep is
// captured into the iteration's graph).
__global__ void k_tile_reset(int* __restrict__ ctr, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) ctr[q] = 0;
}

__device__ __forceinline__ int ld_prog(const int* p)
{
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Edge terms: coherent across the device's L2 caches (see above).  Element e of one tile's edge
// buffer (NEQ x NJ1 x tile width elements) at a wave-uniform base, so the address is the
// scalar-base + 32-bit vector-offset form, with no 64-bit multiply per term.
template <typename R>
__device__ __forceinline__ R ld_edge(const R* base, unsigned e)
{
    TRANS_GLOBAL R* p = (TRANS_GLOBAL R*)((TRANS_GLOBAL char*)const_cast<R*>(base) + e * (unsigned)sizeof(R));
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename R>
__device__ __forceinline__ void st_edge(R* base, unsigned e, R v)
{
    TRANS_GLOBAL R* p = (TRANS_GLOBAL R*)((TRANS_GLOBAL char*)base + e * (unsigned)sizeof(R));
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The edge terms a lane reads at one step (across its tile's upstream a and b edges).
template <typename R, int NEQ>
struct EdgeIn {
    R a[NEQ];
    R b[NEQ];
};

template <typename R, bool FORWARD, int NEQ>
__device__ __forceinline__ void tile_sweep(const Fields& f, R* __restrict__ Dq, const BlkInfo& b,
                                           const SweepTile& tl, int me, const SweepTile* __restrict__ tiles,
                                           R* __restrict__ edge, int* __restrict__ prog, int* __restrict__ err,
                                           int chunk, int spin_limit, unsigned sink, R* __restrict__ buf)
{
    TRANS_SWEEP_CONTRACT
    static_assert(NEQ == 5 || NEQ == 2, "mean flow or k-omega");
    constexpr bool MF = NEQ == 5;
    const int NI1 = b.ni - 1, NJ1 = b.nj - 1, NK1 = b.nk - 1;
    const int Na = NI1, Nb = NK1;  // march axis j: lines (i, k), i fastest
    const int nlines = Na * Nb;
    const int nplanes = NI1 + NJ1 + NK1 - 2;
    constexpr int row = TILE_MAXT + 2;
    constexpr int sinkslot = TILE_MAXT, zero = TILE_MAXT + 1;
    const int ln = threadIdx.x;
    const int ntl = tl.na * tl.nb;
    const bool active = ln < ntl;
    const int ia = active ? ln % tl.na : 0, ib = active ? ln / tl.na : 0;  // within the tile
    const int a = tl.a0 + ia + 1, bb = tl.b0 + ib + 1;                       // line, 1-based
    const R sgn = FORWARD ? R(1) : R(-1);
    // downstream neighbours: inside the tile through LDS, across its edge through the edge buffer
    const bool dn_a = active && (FORWARD ? a < Na : a > 1);
    const bool dn_b = active && (FORWARD ? bb < Nb : bb > 1);
    const bool edge_out_a = dn_a && (FORWARD ? ia == tl.na - 1 : ia == 0);
    const bool edge_out_b = dn_b && (FORWARD ? ib == tl.nb - 1 : ib == 0);
    const int wa = dn_a && !edge_out_a ? (FORWARD ? ln + 1 : ln - 1) : sinkslot;
    const int wb = dn_b && !edge_out_b ? (FORWARD ? ln + tl.na : ln - tl.na) : sinkslot;
    // upstream neighbours
    const bool up_a = active && (FORWARD ? a > 1 : a < Na);
    const bool up_b = active && (FORWARD ? bb > 1 : bb < Nb);
    const bool edge_in_a = up_a && (FORWARD ? ia == 0 : ia == tl.na - 1);
    const bool edge_in_b = up_b && (FORWARD ? ib == 0 : ib == tl.nb - 1);
    const int ra = !active ? sinkslot : (up_a && !edge_in_a ? ln : zero);
    const int rb = !active ? sinkslot : (up_b && !edge_in_b ? ln : zero);
    const int up_tile_a = FORWARD ? tl.lo_a : tl.hi_a;
    const int up_tile_b = FORWARD ? tl.lo_b : tl.hi_b;
    // edge buffers: the tile's own outgoing ones, and those of its upstream tiles it reads
    // (element (L, j, lane along the edge) at L * NJ1 * width + j * width + lane; tiles in one
    // row along a share nb, in one column along b share na)
    R* const out_a = sbase(edge + tl.ea);
    R* const out_b = sbase(edge + tl.eb);
    const R* const in_a = sbase(up_tile_a >= 0 ? edge + tiles[up_tile_a].ea : edge);
    const R* const in_b = sbase(up_tile_b >= 0 ? edge + tiles[up_tile_b].eb : edge);
    const unsigned estr_a = (unsigned)NJ1 * (unsigned)tl.nb, estr_b = (unsigned)NJ1 * (unsigned)tl.na;
    const int start = FORWARD ? (a - 1) + (bb - 1) : (Na - a) + (Nb - bb);
    // the tile's steps
    const int t0 = FORWARD ? tl.a0 + tl.b0 : (Na - (tl.a0 + tl.na)) + (Nb - (tl.b0 + tl.nb));
    const int t1 = t0 + (tl.na - 1) + (tl.nb - 1) + NJ1 - 1;
    for (int q = ln; q < 4 * NEQ; q += blockDim.x) buf[q * row + zero] = R(0);
    __syncthreads();
    const unsigned lw = (unsigned)(b.woff + (long)((a - 1) + Na * (bb - 1)));
    const unsigned my_sink = sink + (unsigned)ln;
    const LineSrc<R, NEQ> src(f, Dq);
    R tm[NEQ];
#pragma unroll
    for (int L = 0; L < NEQ; L++) tm[L] = R(0);
    // edge terms of step t: every step of a chunk is covered by the chunk's wait, so the terms
    // of the next step are loaded one step ahead inside a chunk (at a chunk's first step they
    // are loaded after its wait)
    auto load_edges = [&](const int t, EdgeIn<R, NEQ>& e) {
        const int rel = t - start;
        const bool act = active && rel >= 0 && rel < NJ1;
        const unsigned jrow = act ? (unsigned)rel : 0u;
        if (edge_in_a && act) {
            const unsigned e0 = __umul24(jrow, (unsigned)tl.nb) + (unsigned)ib;
#pragma unroll
            for (int L = 0; L < NEQ; L++) e.a[L] = ld_edge(in_a, e0 + (unsigned)L * estr_a);
        }
        if (edge_in_b && act) {
            const unsigned e0 = __umul24(jrow, (unsigned)tl.na) + (unsigned)ia;
#pragma unroll
            for (int L = 0; L < NEQ; L++) e.b[L] = ld_edge(in_b, e0 + (unsigned)L * estr_b);
        }
    };
    auto step = [&](const int t, const LineRaw<R, NEQ>& cur, LineRaw<R, NEQ>& nxt, EdgeIn<R, NEQ>& ec,
                    EdgeIn<R, NEQ>& en) {
        // chunk start: wait until the upstream tiles have published the steps this chunk reads
        if ((t - t0) % chunk == 0) {
            if (threadIdx.x == 0) {
                const int need = t + chunk - 1;  // steps < need done upstream
                int spins = 0;
                for (;;) {
                    bool ok = true;
                    if (up_tile_a >= 0) {
                        const SweepTile& u = tiles[up_tile_a];
                        const int u0 = FORWARD ? u.a0 + u.b0 : (Na - (u.a0 + u.na)) + (Nb - (u.b0 + u.nb));
                        const int u1 = u0 + (u.na - 1) + (u.nb - 1) + NJ1 - 1;
                        ok = ok && ld_prog(prog + up_tile_a) >= min(need, u1 + 1);
                    }
                    if (up_tile_b >= 0) {
                        const SweepTile& u = tiles[up_tile_b];
                        const int u0 = FORWARD ? u.a0 + u.b0 : (Na - (u.a0 + u.na)) + (Nb - (u.b0 + u.nb));
                        const int u1 = u0 + (u.na - 1) + (u.nb - 1) + NJ1 - 1;
                        ok = ok && ld_prog(prog + up_tile_b) >= min(need, u1 + 1);
                    }
                    if (ok) break;
                    if (++spins > spin_limit || ld_prog(err) != 0) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
            load_edges(t, ec);
        }
        const int par = t & 1;
        const int rel = t - start;
        const bool act = active && rel >= 0 && rel < NJ1;
        const int s = FORWARD ? t : nplanes - 1 - t;
        const int sn = FORWARD ? min(s + 1, nplanes - 1) : max(s - 1, 0);
        // the next step's inputs, loaded by every lane: the slot of the lane's line on hyperplane
        // sn exists in the block's area whether or not the line has a cell there (a lane off its
        // line's range computes on what it loaded and stores to its sink slot).  Loads that no
        // branch skips keep the compiler's memory-counter waits at one step's count, so they
        // stay in flight while the current cell is computed.
        line_load<R, NEQ>(src, lw + (unsigned)sn * (unsigned)nlines, nxt);
        if (t + 1 <= t1 && (t + 1 - t0) % chunk != 0) load_edges(t + 1, en);
        const unsigned wst = act ? lw + (unsigned)s * (unsigned)nlines : my_sink;
        const int jrow = act ? rel : 0;  // j - 1 of the lane's cell
        const double stept = cur.stept;
        const double cvs = cur.cvs;
        const R rf1 = MF ? R(1) : R(1) / R(cur.f1);
        R dq[NEQ];
#pragma unroll
        for (int L = 0; L < NEQ; L++) {
            R sa = buf[((par * 2 + 0) * NEQ + L) * row + ra];
            R sb = buf[((par * 2 + 1) * NEQ + L) * row + rb];
            if (edge_in_a && act) sa = ec.a[L];
            if (edge_in_b && act) sb = ec.b[L];
            const R sm = rel == 0 ? R(0) : tm[L];
            const R sum = sa + sm + sb;  // i, j, k order
            if constexpr (MF) {
                if (FORWARD)
                    dq[L] = R(cur.d[L]) * R(stept) + sum * R(cvs);
                else
                    dq[L] = R(cur.d[L]) - sum * R(cvs);
            } else {
                if (FORWARD)
                    dq[L] = (R(cur.d[L]) * R(stept) + sum * R(cvs)) * rf1;
                else
                    dq[L] = R(cur.d[L]) - sum * R(cvs) * rf1;
            }
            stg(src.d[L], wst, dq[L]);
        }
        const double roinv = double(cur.v[0]);
        const R pVx = R(cur.v[1]), pVy = R(cur.v[2]), pVz = R(cur.v[3]);
        const int npar = par ^ 1;
        AdqPrep<R> prep;
        if constexpr (MF) prep = adq_prep<R>(pVx, pVy, pVz, R(cur.pp), R(roinv), dq);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R o[NEQ];
            const R sx = R(cur.sdm[d][0]), sy = R(cur.sdm[d][1]), sz = R(cur.sdm[d][2]);
            const R rd = R(cur.rds[d]);
            if constexpr (MF) {
                R dif[5];
                adq_dir<R>(prep, sx, sy, sz, dif);
#pragma unroll
                for (int L = 0; L < 5; L++) o[L] = R(0.5) * (dif[L] + sgn * rd * dq[L]);
            } else {
                const R VP = pVx * sx + pVy * sy + pVz * sz;
#pragma unroll
                for (int L = 0; L < NEQ; L++) o[L] = R(0.5) * (dq[L] * VP + sgn * rd * dq[L]);
            }
            if (d == 1) {
#pragma unroll
                for (int L = 0; L < NEQ; L++) tm[L] = o[L];
            } else {
                const int dir = d == 0 ? 0 : 1;
                const int slot = d == 0 ? wa : wb;
#pragma unroll
                for (int L = 0; L < NEQ; L++) buf[((npar * 2 + dir) * NEQ + L) * row + slot] = o[L];
                if (d == 0 && edge_out_a && act) {
                    const unsigned e0 = __umul24((unsigned)jrow, (unsigned)tl.nb) + (unsigned)ib;
#pragma unroll
                    for (int L = 0; L < NEQ; L++) st_edge(out_a, e0 + (unsigned)L * estr_a, o[L]);
                }
                if (d == 2 && edge_out_b && act) {
                    const unsigned e0 = __umul24((unsigned)jrow, (unsigned)tl.na) + (unsigned)ia;
#pragma unroll
                    for (int L = 0; L < NEQ; L++) st_edge(out_b, e0 + (unsigned)L * estr_b, o[L]);
                }
            }
        }
        lds_barrier();
        // chunk end: publish the steps done once every lane's edge stores have completed
        if ((t - t0 + 1) % chunk == 0 || t == t1) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(prog + me, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    LineRaw<R, NEQ> bufA, bufB;
    EdgeIn<R, NEQ> edA, edB;
    {
        const int s = FORWARD ? t0 : nplanes - 1 - t0;
        line_load<R, NEQ>(src, lw + (unsigned)s * (unsigned)nlines, bufA);
        bufB = bufA;  // defined values for lanes that load nothing at the next step
    }
    for (int t = t0; t <= t1; t += 2) {
        step(t, bufA, bufB, edA, edB);
        if (t + 1 <= t1) step(t + 1, bufB, bufA, edB, edA);
    }
}

// Persistent tile-wavefront sweep over all blocks.  ctr[0] is the item counter, prog[] the
// per-tile progress of this sweep.  KIND 0: the table holds mean-flow and k-omega items; 5 or
// 2: only that kind (an instantiation without the other path, whose registers and LDS it
// would otherwise reserve).
template <typename R, bool FORWARD, int KIND>
__global__ void __launch_bounds__(TILE_MAXT) k_lusgs_tile(Fields f, R* Dq, const BlkInfo* __restrict__ blks,
                                                          const SweepTile* __restrict__ tiles, int ntiles,
                                                          int* __restrict__ ctr, int* __restrict__ prog,
                                                          R* __restrict__ edge, int* __restrict__ err, int chunk,
                                                          int spin_limit)
{
    __shared__ R buf[2 * 2 * (KIND == 2 ? 2 : 5) * (TILE_MAXT + 2)];
    __shared__ int s_item;
    const unsigned sink = (unsigned)(f.wsink + (long)blockIdx.x * TILE_MAXT);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(ctr, 1);
        __syncthreads();
        const int q = s_item;
        __syncthreads();
        if (q >= ntiles) break;
        const int me = FORWARD ? q : ntiles - 1 - q;
        const SweepTile tl = tiles[me];
        const BlkInfo b = blks[tl.blk];
        if (KIND != 2 && (KIND == 5 || tl.neq == 5))
            tile_sweep<R, FORWARD, 5>(f, Dq, b, tl, me, tiles, edge, prog, err, chunk, spin_limit, sink, buf);
        else if (KIND != 5)
            tile_sweep<R, FORWARD, 2>(f, Dq, b, tl, me, tiles, edge, prog, err, chunk, spin_limit, sink, buf);
        __syncthreads();
    }
}
