// Immersed-boundary kernels: Peskin 4-point spreading as a deterministic
// per-tile gather, interpolation, multi-direct forcing and node motion.
//
// Spreading without float atomics: every node is binned into the lattice tiles
// its 4x4 stencil touches; one workgroup per tile sorts its node list by
// global node index (capsule-major, the reference's loop order) and every
// cell-thread sums the contributions of the listed nodes in that order.  A
// cell's force is therefore the same sequential sum the reference forms
// (coupling/immersed_boundary.cpp:44-80 at one thread), independent of the
// launch and of how nodes were appended to the lists.
#include <algorithm>
#include <type_traits>

#include "sf_internal.h"

namespace sf {

// Peskin phi from the 1025-entry table (coupling/delta_function.h:45-54).
__device__ __forceinline__ double phi_lut(const double* t, double r) {
    double ar = fabs(r);
    if (ar >= 2.0) return 0.0;
    double idx = ar * (1024 * 0.5);
    int i = static_cast<int>(idx);
    if (i >= 1024) return 0.0;
    double frac = idx - static_cast<double>(i);
    return t[i] * (1.0 - frac) + t[i + 1] * frac;
}

// Node binning, the first stage of a spread: every node's record, and its
// reference in each lattice tile its 4x4 stencil touches.
struct BinArgs {
    NodeSpread* rec;        // per node
    int* tile_cnt;          // per tile: nodes binned
    int* tile_list;         // per tile: TILE_MAXL node indices
    NodeSpread* tile_rec;   // per tile: copies of its first TILE_RECS records
    int* counters;          // [1]: tile list overflow; [3]: touched-tile count
    int nx, ny, ntx;
    int2* act;              // touched-tile list (tile, -1) built here, or null (built after binning)
    int ntiles;
};

// Arc length element of node k: half the distances to its ring neighbours
// ring = (previous, next) node index (immersed_boundary.cpp:44-60), minimum
// image over nx.
__device__ __forceinline__ double node_arc(const double* x, const double* y, int2 ring, double px, double py,
                                           int nx) {
    const double Lx = static_cast<double>(nx);
    const double dp = min_image_ds(px, py, x[ring.x], y[ring.x], Lx);
    const double dn = min_image_ds(x[ring.y], y[ring.y], px, py, Lx);
    return 0.5 * (dp + dn);
}

// The tiles a record's stencil touches, as up to 8 candidates: candidate c
// is the (c & 3)-th distinct tile column of the stencil columns i0-1 .. i0+2
// (each wrapped once, in column order) in the (c >> 2)-th tile row of its rows
// j0-1 .. j0+2 within the lattice, or -1.  A record with a finite position is
// binned even when its force is not finite (it then adds nothing).  Registers
// only: the distinct columns are kept in four scalars.
__device__ __forceinline__ void node_tiles(const NodeSpread& s, int nx, int ny, int ntx, int (&t)[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c) t[c] = -1;
    if (!finite_position(s.px, s.py)) return;
    const int i0 = s.i0, j0 = s.j0;
    if (j0 < -2 || j0 >= ny + 2) return;
    const int ylo = j0 - 1 < 0 ? 0 : j0 - 1;
    const int yhi = j0 + 2 > ny - 1 ? ny - 1 : j0 + 2;
    if (ylo > yhi) return;
    const int ty0 = ylo / TILE_Y, ty1 = yhi / TILE_Y;
    int u0 = -1, u1 = -1, u2 = -1, u3 = -1, nt = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        int jx = i0 - 1 + d;
        if (jx < 0) jx += nx;
        if (jx >= nx) jx -= nx;
        if (jx < 0 || jx >= nx) continue;
        const int tx = jx / TILE_X;
        if (tx == u0 || tx == u1 || tx == u2 || tx == u3) continue;
        if (nt == 0) u0 = tx;
        else if (nt == 1) u1 = tx;
        else if (nt == 2) u2 = tx;
        else u3 = tx;
        ++nt;
    }
    const int u[4] = {u0, u1, u2, u3};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ty = ty0 + (c >> 2);
        if ((c & 3) < nt && ty <= ty1) t[c] = ty * ntx + u[c & 3];
    }
}

// Appends node s.pad to each candidate tile's list (and its record to the
// tile's copies).  Every lane of the wavefront calls it (t all -1 for lanes
// without a node); the slots are claimed per wavefront (wave_claim), every
// candidate's atomics issued before any store waits on one.
__device__ __forceinline__ void bin_node(const BinArgs& b, const int (&t)[8], const NodeSpread& s) {
    int lead[8], base[8], slot[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) wave_claim(b.tile_cnt, t[c], lead[c], base[c]);
#pragma unroll
    for (int c = 0; c < 8; ++c) slot[c] = wave_slot(lead[c], base[c]);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (t[c] < 0) continue;
        if (slot[c] < TILE_MAXL) b.tile_list[static_cast<size_t>(t[c]) * TILE_MAXL + slot[c]] = s.pad;
        else atomicOr(&b.counters[1], 1);
        if (slot[c] < TILE_RECS) b.tile_rec[static_cast<size_t>(t[c]) * TILE_RECS + slot[c]] = s;
    }
    if (b.act) {
        // The claim that opened a tile (slot 0) lists it as touched, one list
        // atomic per workgroup (at most 256 threads, all of which get here);
        // the gather reads the tile's count once the binning is over (entry
        // .y = -1).
        __shared__ int s_cnt[4], s_base;
        const int lane = threadIdx.x & 63, wv = static_cast<int>(threadIdx.x >> 6);
        int fresh = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) fresh += t[c] >= 0 && slot[c] == 0;
        int incl = fresh;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        if (lane == 63) s_cnt[wv] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) total += s_cnt[w];
            s_base = total ? atomicAdd(&b.counters[3], total) : 0;
        }
        __syncthreads();
        int pos = s_base + incl - fresh;
        for (int w = 0; w < wv; ++w) pos += s_cnt[w];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (t[c] < 0 || slot[c] != 0) continue;
            if (pos < b.ntiles) b.act[pos] = make_int2(t[c], -1);
            ++pos;
        }
    }
}

// ---------------------------------------------------------------------------
// The node part of MDF correction it, one thread per node k: records the
// iteration's decision (mdf_decide; node 0) and, when the correction is
// applied, counts it (node 0) and adds it to the node's step total, and to its
// head total while it <= MDF_HEAD.
__device__ __forceinline__ void mdf_node_apply(MdfDev* md, double* total, double* head, const NodeSpread* rec,
                                               int K, int it, double tol, int k) {
    if (k >= K) return;
    const int dec = mdf_decide(md, it, tol);
    if (k == 0) md->cby[it] = dec;
    if (dec) return;
    if (k == 0) md->applied += 1;
    total[2 * k] += rec[k].fx;
    total[2 * k + 1] += rec[k].fy;
    if (it <= MDF_HEAD) {
        head[2 * k] += rec[k].fx;
        head[2 * k + 1] += rec[k].fy;
    }
}

// Node records + tile binning (mode 0: the node's own force; with md set it
// also clears the step's MDF state -- decisions, residuals, totals -- ahead of
// the first correction).  mode 1 is MDF correction it's node part
// (mdf_node_apply): the MDF interpolation has already written the node's
// record with the correction.
// arc: the arc length elements at the current positions, or null (computed
// here from the ring neighbours).
__device__ __forceinline__ void node_prep_body(const double* x, const double* y, const double* fx, const double* fy,
                            const int2* ring, const double* arc, NodeSpread* rec, int K, int nx, int ny,
                            int ntx, int nty, int* tile_cnt, int* tile_list, NodeSpread* tile_rec,
                            int* counters, int mode, MdfDev* md, double tol, int it,
                            double* total, double* head, int2* act, int bx) {
    const int k = bx * blockDim.x + threadIdx.x;
    if (mode == 0) {
        if (md) {
            const int nthr = gridDim.x * blockDim.x;
            if (k == 0) { md->conv = 0; md->applied = 0; md->conv_it = 0; }
            for (int i = k; i < MDF_RES_LD; i += nthr) md->cby[i] = 0;
            for (int i = k; i < MDF_RES_SUB * (MDF_MAX_CORR + 2); i += nthr)
                md->res[i / (MDF_MAX_CORR + 2)][i % (MDF_MAX_CORR + 2)] = 0ull;
            for (int i = k; i < 2 * K; i += nthr) { total[i] = 0.0; head[i] = 0.0; }
        }
        // binning: every lane of the wavefront takes part in the slot claims.
        // Without act the touched-tile list is built after binning, from zero;
        // with it, the binning appends to a list the previous step's close
        // (or launch_bins_close) emptied.
        if (k == 0 && !act) counters[3] = 0;
        int t[8];
        NodeSpread s{};
        if (k < K) {
            const double px = x[k], py = y[k];
            const double ds = arc ? arc[k] : node_arc(x, y, ring[k], px, py, nx);
            s = node_record(px, py, fx[k], fy[k], ds, k, ny);
            rec[k] = s;
            node_tiles(s, nx, ny, ntx, t);
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) t[c] = -1;
        }
        bin_node(BinArgs{rec, tile_cnt, tile_list, tile_rec, counters, nx, ny, ntx, act, ntx * nty}, t, s);
        return;
    }
    if (mode == 1) mdf_node_apply(md, total, head, rec, K, it, tol, k);
}

__global__ void k_node_prep(const double* x, const double* y, const double* fx, const double* fy,
                            const int2* ring, const double* arc, NodeSpread* rec, int K, int nx, int ny,
                            int ntx, int nty, int* tile_cnt, int* tile_list, NodeSpread* tile_rec,
                            int* counters, int mode, MdfDev* md, double tol, int it,
                            double* total, double* head, int2* act) {
    node_prep_body(x, y, fx, fy, ring, arc, rec, K, nx, ny, ntx, nty, tile_cnt, tile_list, tile_rec, counters, mode, md, tol, it, total, head, act, blockIdx.x);
}

// Gather: one 256-thread workgroup per TILE_X x TILE_Y tile.
// mode 0: F = sum;  mode 1: F_out = F_in + sum  (MDF correction).
// Tiles no node touches are skipped when the output buffer already holds zeros
// there (dirty flags per buffer), so the dense write covers touched tiles only.
constexpr int STAGE_MAX = 64;      // node records staged in LDS (lists up to this length)
constexpr int CELL_MAXPC = 16;     // contributions one cell collects in LDS before falling back
constexpr int NO_ROW = -0x40000000;

// Unwrapped stencil column (lo .. lo+3) whose single periodic wrap lands on cx,
// or NO_ROW if none does.
__device__ __forceinline__ int stencil_column(int cx, int lo, int nx) {
    if (cx - lo >= 0 && cx - lo <= 3) return cx;
    if (cx - nx - lo >= 0 && cx - nx - lo <= 3) return cx - nx;
    if (cx + nx - lo >= 0 && cx + nx - lo <= 3) return cx + nx;
    return NO_ROW;
}

// rest_tpw(mode) tiles per wavefront, one per lane of its first lanes.  Tiles no
// node touches: the wavefront clears (or, for an MDF correction, copies from
// the input) each one that needs it -- those whose output still holds non-zero
// values, or whose input does; after convergence (skip) every tile counts as
// untouched.  Touched tiles (mode 0, which runs once per step after the
// binning): appended to act as (tile, node count), one atomic per wavefront.
// The pass after the binning (mode 0) mostly clears: a whole wavefront of
// tiles each, so that few wavefronts take part in the touched-tile list.  MDF
// corrections (mode 1) copy whole tiles once converged: 16 tiles each, for more
// wavefronts in flight.
constexpr int REST_TPW0 = 64;
constexpr int REST_TPW1 = 16;
__host__ __device__ constexpr int rest_tpw(int mode) { return mode == 0 ? REST_TPW0 : REST_TPW1; }
__host__ __device__ constexpr int rest_tiles_per_block(int mode) { return 4 * rest_tpw(mode); }

// After convergence (residual mode) the remaining corrections add nothing, so
// the force both buffers of the correction ping-pong must end with is the one
// the last applied correction left.  Rather than copying it forward at every
// remaining iteration, the iterations before the last do nothing and the last
// one copies it once into the buffer that does not hold it: the result
// equals the copied-forward chain's.  Convergence detected at iteration j
// leaves the force in the input buffer of iteration j, which is the input
// buffer of the last iteration when (last - j) is even and its output
// buffer otherwise (the copy then runs output -> input).
template <bool FP32>
__device__ __forceinline__ void spread_rest_body(const int* tile_cnt, const void* Fin, void* Fout,
                                                     const uint8_t* dirty_in, uint8_t* dirty_out,
                                                     int2* act, int* counters, int nx, int ny,
                                                     int ntx, int ntiles, int mode,
                                                     const MdfDev* md, int bx, int it = 0, int last_it = 0,
                                                     double tol = 0.0, int dec_given = -1) {
    using T = typename std::conditional<FP32, float, double>::type;
    // dec_given >= 0: the decision, known to the caller (the correction loop's
    // tail, which skips the iterations after the one that stopped it)
    const int dec = dec_given >= 0 ? dec_given : (mode == 1 && md ? mdf_decide(md, it, tol) : 0);
    const bool skip = dec != 0;
    if (skip && last_it > 0) {
        if (it < last_it) return;
        if (((last_it - dec) & 1) != 0) {
            // the force is in this iteration's output buffer: copy it back
            const void* t = Fin;
            Fin = Fout;
            Fout = const_cast<void*>(t);
            const uint8_t* d = dirty_in;
            dirty_in = dirty_out;
            dirty_out = const_cast<uint8_t*>(d);
        }
    }
    const int lane = threadIdx.x & 63;
    const int tpw = rest_tpw(mode);
    const int wave_tile0 = (bx * static_cast<int>(blockDim.x >> 6) + static_cast<int>(threadIdx.x >> 6)) * tpw;
    const int tile = lane < tpw ? wave_tile0 + lane : ntiles;   // the other lanes hold none
    const int cnt = tile < ntiles ? tile_cnt[tile] : 0;
    bool need = false, copy = false;
    if (tile < ntiles && (skip || cnt == 0)) {
        copy = mode == 1 && dirty_in[tile];
        need = copy || dirty_out[tile];
    }
    if (mode == 0 && act) {
        // the touched-tile list, when the binning did not make it: one list
        // atomic per workgroup (at most 256 threads); the wavefronts' touched
        // counts meet in LDS (mode 0 has no early exit)
        __shared__ int s_cnt[4], s_base;
        const int wv = static_cast<int>(threadIdx.x >> 6);
        const unsigned long long touched = __ballot(cnt > 0);
        if (lane == 0) s_cnt[wv] = __popcll(touched);
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) total += s_cnt[w];
            s_base = total ? atomicAdd(&counters[3], total) : 0;
        }
        __syncthreads();
        int base = s_base;
        for (int w = 0; w < wv; ++w) base += s_cnt[w];
        if (cnt > 0) {
            const int pos = base + __popcll(touched & ((1ull << lane) - 1ull));
            if (pos < ntiles) act[pos] = make_int2(tile, cnt);
        }
    }
    unsigned long long todo = __ballot(need);
    const unsigned long long cmask = __ballot(copy);
    // The wavefront clears or copies its tiles two at a time; every lane issues
    // all its loads of both tiles before its stores.  A copy reads the input
    // buffer and writes the output buffer, which is always a different one.
    using T2 = typename std::conditional<FP32, float2, double2>::type;
    constexpr int PER = TILE_X * TILE_Y / 64;    // cells per lane and tile
    const T2* fin = static_cast<const T2*>(Fin);
    T2* fout = static_cast<T2*>(Fout);
    while (todo) {
        int tl[2] = {0, 0};
        bool cps[2] = {false, false};
        int nt = 0;
        for (; nt < 2 && todo; ++nt) {
            const int b = __ffsll(static_cast<long long>(todo)) - 1;
            todo &= todo - 1;
            tl[nt] = wave_tile0 + b;
            cps[nt] = (cmask >> b) & 1ull;
        }
        T2 v[2][PER];
        size_t at[2][PER];
        bool in[2][PER];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int tx0 = (tl[t] % ntx) * TILE_X, ty0 = (tl[t] / ntx) * TILE_Y;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int c = lane + 64 * i;
                const int cx = tx0 + (c % TILE_X), cy = ty0 + (c / TILE_X);
                in[t][i] = t < nt && cx < nx && cy < ny;
                at[t][i] = static_cast<size_t>(cy) * nx + cx;
                v[t][i].x = T(0);
                v[t][i].y = T(0);
                if (in[t][i] && cps[t]) v[t][i] = fin[at[t][i]];
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < PER; ++i)
                if (in[t][i]) fout[at[t][i]] = v[t][i];
        if (lane == 0)
            for (int t = 0; t < nt; ++t) dirty_out[tl[t]] = cps[t] ? 1 : 0;
    }
}

// Touched tiles: a persistent grid walks the list of tiles the node binning
// touched (act[0 .. counters[3])), one 256-thread workgroup per tile and one
// thread per cell of its TILE_X x TILE_Y cells.
// mode 0: F = sum;  mode 1: F_out = F_in + sum  (MDF correction it, unless
// the loop has stopped, mdf_decide).  last: the tile counts are cleared.
template <bool FP32>
__device__ __forceinline__ void spread_body(const NodeSpread* rec, const NodeSpread* tile_rec,
                                                int* tile_cnt, const int* tile_list, const int2* act,
                                                const int* counters, const double* lut,
                                                const void* Fin, void* Fout, uint8_t* dirty_out,
                                                int nx, int ny, int ntx, int ntiles, int mode,
                                                int last, const MdfDev* md, int bx, int gx, int it = 0,
                                                double tol = 0.0) {
    // the delta table is read through the caches: staged in LDS it held the
    // spread to 7 workgroups per CU, and each tile is short, latency-bound work
    const double* tab = lut;
    __shared__ double sfx[STAGE_MAX], sfy[STAGE_MAX], sds[STAGE_MAX];
    // each staged node's stencil weights: phi(px - (i0 - 1 + d)), phi(py - (j0 - 1 + d))
    __shared__ double swx[STAGE_MAX][4], swy[STAGE_MAX][4];
    __shared__ int si0[STAGE_MAX], sj0[STAGE_MAX], skey[STAGE_MAX];
    __shared__ int ccnt[TILE_X * TILE_Y];
    __shared__ int cslot_raw[TILE_X * TILE_Y * CELL_MAXPC / 2];
    unsigned short* cslot = reinterpret_cast<unsigned short*>(cslot_raw);
    int* list = cslot_raw;   // the long-list path keeps its sorted list here
    using T = typename std::conditional<FP32, float, double>::type;
    const int t = threadIdx.x;
    // each list entry is read one tile ahead, the first one beside the count
    // (the list holds ntiles entries; those past the count are never used)
    int2 te_next = act[bx < ntiles ? bx : 0];
    const int nact = min(counters[3], ntiles);
    const bool skip = mode == 1 && md && mdf_decide(md, it, tol) != 0;
    if (skip) {
        // converged MDF: nothing to add (the rest kernel copied F_in)
        if (last)
            for (int i = bx * blockDim.x + t; i < nact; i += gx * blockDim.x) tile_cnt[act[i].x] = 0;
        return;
    }
    // workgroups past the touched-tile count have nothing to do
    if (bx >= nact) return;
    for (int ai = bx; ai < nact; ai += gx) {
        const int2 te = te_next;
        if (ai + gx < ntiles) te_next = act[ai + gx];
        const int tile = te.x;
        __syncthreads();          // the previous tile's LDS use is over
        int L = te.y >= 0 ? te.y : tile_cnt[tile];   // -1: listed by the binning
        if (L > TILE_MAXL) L = TILE_MAXL;
        const int tx0 = (tile % ntx) * TILE_X, ty0 = (tile / ntx) * TILE_Y;
        const int cx = tx0 + (t % TILE_X);
        const int cy = ty0 + (t / TILE_X);
        const bool inside = cx < nx && cy < ny;
        const size_t n = static_cast<size_t>(cy) * nx + cx;
        const int* tl = tile_list + static_cast<size_t>(tile) * TILE_MAXL;
        double ax = 0.0, ay = 0.0;
        if (mode == 1 && inside) {
            ax = static_cast<const T*>(Fin)[2 * n];
            ay = static_cast<const T*>(Fin)[2 * n + 1];
        }
        // Staging: four threads per record; thread d of a record computes the
        // record's stencil weights of column and row d (the table is read here
        // only), thread 0 also stages the record.  A weight below is then the
        // product phi(px - ix) * phi(py - iy) of the same two table values.
        auto stage_rec = [&](int i, int d, const NodeSpread& s) {
            const bool valid = s.valid != 0;
            if (d == 0) {
                sfx[i] = s.fx; sfy[i] = s.fy; sds[i] = s.ds;
                si0[i] = s.i0;
                sj0[i] = valid ? s.j0 : NO_ROW;   // invalid records touch no row
                skey[i] = s.pad;                  // the node index
            }
            // a valid record's position is finite
            swx[i][d] = valid ? phi_lut(tab, s.px - static_cast<double>(s.i0 - 1 + d)) : 0.0;
            swy[i][d] = valid ? phi_lut(tab, s.py - static_cast<double>(s.j0 - 1 + d)) : 0.0;
        };
        auto stage = [&](int i, int d, int k) { stage_rec(i, d, rec[k]); };
        auto contribution = [&](int i, double& w, int& ok) {
            ok = 0;
            w = 0.0;
            const int dyc = cy - (sj0[i] - 1);
            if (dyc < 0 || dyc > 3) return;
            const int lo = si0[i] - 1;
            const int ix = stencil_column(cx, lo, nx);
            if (ix == NO_ROW) return;
            w = swx[i][ix - lo] * swy[i][dyc];
            ok = w > 1e-15;
        };

        if (L <= STAGE_MAX) {
            // stage the records, then enumerate node x stencil pairs in parallel
            // into per-cell lists
            // mode 0 reads the record copies the binning left in the tile
            // (one dependent load less); MDF corrections change the forces
            // after binning, so mode 1 reads the records themselves
            if (mode == 0 && L <= TILE_RECS) {
                const NodeSpread* tr = tile_rec + static_cast<size_t>(tile) * TILE_RECS;
                for (int i = t; i < 4 * L; i += 256) stage_rec(i >> 2, i & 3, tr[i >> 2]);
            } else {
                for (int i = t; i < 4 * L; i += 256) stage(i >> 2, i & 3, tl[i >> 2]);
            }
            ccnt[t] = 0;
            __syncthreads();
            if (t == 0) {
                if (last) tile_cnt[tile] = 0;
                dirty_out[tile] = 1;
            }
            for (int i = t; i < 16 * L; i += 256) {
                const int p = i >> 4, s = i & 15;
                const int j0 = sj0[p];
                if (j0 == NO_ROW) continue;
                const int iy = j0 - 1 + (s >> 2);
                if (iy < 0 || iy >= ny) continue;
                const int ly = iy - ty0;
                if (ly < 0 || ly >= TILE_Y) continue;
                const int ix = si0[p] - 1 + (s & 3);
                int jx = ix;
                if (jx < 0) jx += nx;
                if (jx >= nx) jx -= nx;
                const int lx = jx - tx0;
                if (lx < 0 || lx >= TILE_X) continue;
                const double w = swx[p][s & 3] * swy[p][s >> 2];
                if (!(w > 1e-15)) continue;
                const int cell = ly * TILE_X + lx;
                const int slot = atomicAdd(&ccnt[cell], 1);
                if (slot < CELL_MAXPC) cslot[cell * CELL_MAXPC + slot] = static_cast<unsigned short>(p);
            }
            __syncthreads();
            const int m = ccnt[t];
            if (inside && m <= CELL_MAXPC) {
                // this cell's contributions in global node order (insertion sort)
                unsigned short* my = cslot + t * CELL_MAXPC;
                for (int a = 1; a < m; ++a) {
                    const unsigned short v = my[a];
                    const int kv = skey[v];
                    int b = a - 1;
                    while (b >= 0 && skey[my[b]] > kv) { my[b + 1] = my[b]; --b; }
                    my[b + 1] = v;
                }
                for (int j = 0; j < m; ++j) {
                    const int p = my[j];
                    double w; int ok;
                    contribution(p, w, ok);
                    ax += sfx[p] * w * sds[p];
                    ay += sfy[p] * w * sds[p];
                }
            } else if (inside) {
                // more contributions than slots: walk the staged records in node order
                int lastk = -1;
                for (;;) {
                    int best = -1, bestk = 0x7fffffff;
                    for (int i = 0; i < L; ++i) {
                        const int k = skey[i];
                        if (k <= lastk || k >= bestk) continue;
                        double w; int ok;
                        contribution(i, w, ok);
                        if (ok) { best = i; bestk = k; }
                    }
                    if (best < 0) break;
                    double w; int ok;
                    contribution(best, w, ok);
                    ax += sfx[best] * w * sds[best];
                    ay += sfy[best] * w * sds[best];
                    lastk = bestk;
                }
            }
        } else {
            // long list: sort it, then every cell scans it in chunks
            int P = 1;
            while (P < L) P <<= 1;
            for (int i = t; i < P; i += 256) list[i] = i < L ? tl[i] : 0x7fffffff;
            __syncthreads();
            if (t == 0) {
                if (last) tile_cnt[tile] = 0;
                dirty_out[tile] = 1;
            }
            for (int k = 2; k <= P; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = t; i < P; i += 256) {
                        const int l = i ^ j;
                        if (l > i) {
                            const int a = list[i], b = list[l];
                            const bool up = (i & k) == 0;
                            if ((a > b) == up) { list[i] = b; list[l] = a; }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int base = 0; base < L; base += STAGE_MAX) {
                const int m = min(STAGE_MAX, L - base);
                for (int i = t; i < 4 * m; i += 256) stage(i >> 2, i & 3, list[base + (i >> 2)]);
                __syncthreads();
                if (inside) {
                    for (int i = 0; i < m; ++i) {
                        double w; int ok;
                        contribution(i, w, ok);
                        if (ok) {
                            ax += sfx[i] * w * sds[i];
                            ay += sfy[i] * w * sds[i];
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (inside) {
            static_cast<T*>(Fout)[2 * n] = static_cast<T>(ax);
            static_cast<T*>(Fout)[2 * n + 1] = static_cast<T>(ay);
        }
    }
}

// The spread of a step's node forces in one launch: workgroups [0, nrest)
// clear the untouched tiles that still hold force (spread_rest_body; the
// binning made the touched-tile list), the rest gather the touched tiles
// (spread_body).  The two parts write disjoint tiles.  The tile counts stay
// set through the launch (the untouched part reads them): the step's closing
// interpolation clears them, or launch_bins_close.
template <bool FP32>
__global__ __launch_bounds__(256) void k_spread_step(const NodeSpread* rec, const NodeSpread* tile_rec,
                                                     int* tile_cnt, const int* tile_list, const int2* act,
                                                     int* counters, const double* lut, void* F, uint8_t* dirty,
                                                     int nx, int ny, int ntx, int ntiles, int nrest) {
    const int b = blockIdx.x;
    if (b < nrest) {
        spread_rest_body<FP32>(tile_cnt, F, F, dirty, dirty, nullptr, counters, nx, ny, ntx, ntiles, 0, nullptr, b);
        return;
    }
    spread_body<FP32>(rec, tile_rec, tile_cnt, tile_list, act, counters, lut, F, F, dirty, nx, ny, ntx, ntiles, 0,
                      0, nullptr, b - nrest, static_cast<int>(gridDim.x) - nrest);
}

// MDF correction it in one launch: workgroups [0, nrest) the untouched tiles
// (spread_rest_body), [nrest, nrest + nspread) the touched ones (spread_body),
// the rest the node part (mdf_node_apply, one node per thread).  The three
// parts write disjoint data and each takes the iteration's decision itself.
// The tile counts stay (the rest part reads them): the statistics fold clears
// them.
template <bool FP32>
__global__ __launch_bounds__(256) void k_mdf_apply(const NodeSpread* rec, int* tile_cnt, const int* tile_list,
                                                   int2* act, int* counters, const double* lut, const void* Fin,
                                                   void* Fout, const uint8_t* dirty_in, uint8_t* dirty_out,
                                                   int nx, int ny, int ntx, int ntiles, MdfDev* md, double tol,
                                                   int it, int last_it, double* total, double* head, int K,
                                                   int nrest, int nspread) {
    const int b = blockIdx.x;
    if (b < nrest)
        spread_rest_body<FP32>(tile_cnt, Fin, Fout, dirty_in, dirty_out, act, counters, nx, ny, ntx, ntiles, 1, md,
                               b, it, last_it, tol);
    else if (b < nrest + nspread)
        spread_body<FP32>(rec, nullptr, tile_cnt, tile_list, act, counters, lut, Fin, Fout, dirty_out, nx, ny, ntx,
                          ntiles, 1, 0, md, b - nrest, nspread, it, tol);
    else
        mdf_node_apply(md, total, head, rec, K, it, tol, (b - nrest - nspread) * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Velocity at a stencil cell from the populations and a force field.
// force_mode 0: raw force (the MDF's computeMacroscopic on F so far);
// force_mode 1: the collision force (raw + body force, capped).
struct CellVelArgs {
    const void* f;
    const void* F;
    int nx, ny, N;
    int walls_x, walls_y;
    const uint8_t* flags;
    int generic_flags;
    int has_body;
    double bfx, bfy, fmax, fmax2;
};

// FLUID cells take the body force (INLET and OUTLET cells do not).
__device__ __forceinline__ bool cell_fluid(const CellVelArgs& a, int x, int y) {
    return !a.generic_flags || a.flags[y * a.nx + x] == SF_FLUID;
}

__device__ __forceinline__ bool cell_solid(const CellVelArgs& a, int x, int y) {
    if (a.generic_flags) return a.flags[y * a.nx + x] == SF_SOLID;
    return (a.walls_y && (y == 0 || y == a.ny - 1)) || (a.walls_x && (x == 0 || x == a.nx - 1));
}

template <bool FP32>
__device__ __forceinline__ void cell_velocity(const CellVelArgs& a, int x, int y, int force_mode,
                                              double& rho, double& ux, double& uy) {
    if (cell_solid(a, x, y)) { rho = 0.0; ux = 0.0; uy = 0.0; return; }
    const int n = y * a.nx + x;
    const size_t N = static_cast<size_t>(a.N);
    if constexpr (!FP32) {
        const double* f = static_cast<const double*>(a.f);
        double r = 0.0, vx = 0.0, vy = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            double fi = f[q * N + n];
            r += fi;
            vx += fi * static_cast<double>(dcx(q));
            vy += fi * static_cast<double>(dcy(q));
        }
        const double* Fp = static_cast<const double*>(a.F) + 2 * static_cast<size_t>(n);
        double fx = Fp[0], fy = Fp[1];
        if (force_mode == 1) {
            if (a.has_body && cell_fluid(a, x, y)) { fx += a.bfx; fy += a.bfy; }
            double m2 = fx * fx + fy * fy;
            if (m2 > a.fmax2) { double sc = a.fmax / sqrt(m2); fx *= sc; fy *= sc; }
        }
        rho = r;
        if (r > 1e-15) { ux = (vx + 0.5 * fx) / r; uy = (vy + 0.5 * fy) / r; }
        else { ux = 0.0; uy = 0.0; }
    } else {
        const float* f = static_cast<const float*>(a.f);
        float d[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) d[q] = f[q * N + n];
        const float drho = ((d[0] + (d[1] + d[3])) + (d[2] + d[4])) + ((d[5] + d[7]) + (d[6] + d[8]));
        const float jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
        const float jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
        const float* Fp = static_cast<const float*>(a.F) + 2 * static_cast<size_t>(n);
        float fx = Fp[0], fy = Fp[1];
        if (force_mode == 1) {
            if (a.has_body && cell_fluid(a, x, y)) { fx += static_cast<float>(a.bfx); fy += static_cast<float>(a.bfy); }
            float m2 = fx * fx + fy * fy;
            if (m2 > static_cast<float>(a.fmax2)) { float sc = static_cast<float>(a.fmax) / sqrtf(m2); fx *= sc; fy *= sc; }
        }
        const float r = 1.0f + drho;
        rho = r;
        ux = (jx + 0.5f * fx) / r;
        uy = (jy + 0.5f * fy) / r;
    }
}

// Interpolation at a point (immersed_boundary.cpp:191-229) with LPN lanes
// per node (16, 8 or 4): lane l handles the 16 / LPN stencil cells
// l * 16 / LPN .., cell c being (iy, ix) = (c / 4, c % 4); the terms are summed
// in the reference's loop order (iy outer, ix inner) by the group's first
// lane, so the result is the reference's sequential sum whatever LPN is.
constexpr int INTERP_NODES = 16;   // nodes per 256-thread workgroup at 16 lanes per node
__host__ __device__ constexpr int interp_nodes(int lpn) { return 256 / lpn; }

template <bool FP32, int LPN>
__device__ __forceinline__ void interp_group(const CellVelArgs& a, const double* lut, double px,
                                             double py, int force_mode, int l, double* tx,
                                             double* ty, int* use, double& Ux, double& Uy) {
    // tx/ty/use: this node's 16-entry slots in LDS
    bool ok = !(isnan(px) || isnan(py) || isinf(px) || isinf(py));
    int ix0 = 0, iy0 = 0;
    if (ok) {
        ix0 = static_cast<int>(floor(px));
        iy0 = static_cast<int>(floor(py));
        if (iy0 < -2 || iy0 >= a.ny + 2) ok = false;
    }
    constexpr int CPL = 16 / LPN;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
        const int cell = l * CPL + j;
        int flag = 0;
        double vx = 0.0, vy = 0.0;
        if (ok) {
            const int ix = ix0 - 1 + (cell & 3), iy = iy0 - 1 + (cell >> 2);
            int jx = ix;
            if (jx < 0) jx += a.nx;
            if (jx >= a.nx) jx -= a.nx;
            if (iy >= 0 && iy < a.ny) {
                const double w = phi_lut(lut, px - static_cast<double>(ix)) *
                                 phi_lut(lut, py - static_cast<double>(iy));
                if (w > 1e-15) {
                    double r, ux, uy;
                    cell_velocity<FP32>(a, jx, iy, force_mode, r, ux, uy);
                    vx = ux * w;
                    vy = uy * w;
                    flag = 1;
                }
            }
        }
        tx[cell] = vx; ty[cell] = vy; use[cell] = flag;
    }
    __syncthreads();
    Ux = 0.0; Uy = 0.0;
    if (l == 0) {
        for (int m = 0; m < 16; ++m)
            if (use[m]) { Ux += tx[m]; Uy += ty[m]; }
    }
}

// MDF correction forces (immersed_boundary.cpp:245-284): for every node,
// dF = 2 rho_l (V_prev - U(X)) with U from the populations and the raw force
// spread so far; written into the node's spread record.
// Also records the no-slip residual max_k |V_k - U(X_k)| of this iteration
// (md->res[it], an order-independent max) and, when u0 is set, the slip
// V - U per node.  After convergence (residual mode) it does nothing.
template <bool FP32, int LPN = 16>
__device__ __forceinline__ void mdf_interp_body(const CellVelArgs& a, const double* lut, const double* x,
                                                const double* y, const double* vx, const double* vy,
                                                NodeSpread* rec, int K, MdfDev* md, int it, double* u0,
                                                int bx) {
    constexpr int NPW = interp_nodes(LPN);
    __shared__ double tx[16 * NPW], ty[16 * NPW];
    __shared__ int use[16 * NPW];
    __shared__ double sres[NPW];
    if (it > 1 && md->cby[it - 1]) return;   // the loop stopped at an earlier iteration
    const int g = threadIdx.x / LPN, l = threadIdx.x % LPN;
    const int k = bx * NPW + g;
    const bool live = k < K;
    const double px = live ? x[k] : 0.0, py = live ? y[k] : 0.0;
    double Ux, Uy;
    interp_group<FP32, LPN>(a, lut, px, py, 0, l, tx + 16 * g, ty + 16 * g, use + 16 * g, Ux, Uy);
    if (l == 0) {
        double res = 0.0;
        if (live) {
            const double dux = vx[k] - Ux, duy = vy[k] - Uy;
            double rl = 1.0;
            int ix = static_cast<int>(px + 0.5);
            const int iy = static_cast<int>(py + 0.5);
            if (ix < 0) ix += a.nx;
            if (ix >= a.nx) ix -= a.nx;
            if (iy >= 0 && iy < a.ny) {
                double r, u, v;
                cell_velocity<FP32>(a, ix, iy, 0, r, u, v);
                rl = r;
            }
            const double two = 2.0 * rl;
            // the whole record, valid where the correction is finite
            rec[k] = node_record(px, py, dux * two, duy * two, rec[k].ds, k, a.ny);
            if (u0) { u0[2 * k] = dux; u0[2 * k + 1] = duy; }
            res = sqrt(dux * dux + duy * duy);
        }
        sres[g] = res;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < NPW; ++i) m = (sres[i] > m || sres[i] != sres[i]) ? sres[i] : m;
        // non-negative doubles order like their bit patterns (NaN above all)
        atomicMax(&md->res[bx % MDF_RES_SUB][it], static_cast<unsigned long long>(__double_as_longlong(m)));
    }
}

// XCD-aware order of the node-parallel interpolations.  Workgroup b runs on
// XCD b mod 8; with band = ceil(blocks / 8) it takes node block
// (b mod 8) * band + b / 8, so each XCD walks one contiguous eighth of the
// nodes.  A capsule's nodes (and the lattice rows around them) are then read
// through one XCD's L2 instead of two or three: the stencils of neighbouring
// nodes share most of their cache lines.  band = 0 keeps blockIdx order.
__device__ __forceinline__ int band_block(int b, int band) {
    return band > 0 ? (b % 8) * band + b / 8 : b;
}

template <bool FP32, int LPN>
__global__ __launch_bounds__(256) void k_mdf_interp(CellVelArgs a, const double* lut, const double* x,
                                                    const double* y, const double* vx, const double* vy,
                                                    NodeSpread* rec, int K, MdfDev* md, int it,
                                                    double* u0, int band) {
    const int bx = band_block(blockIdx.x, band);
    if (bx * interp_nodes(LPN) >= K) return;       // padding workgroup of the banded grid
    mdf_interp_body<FP32, LPN>(a, lut, x, y, vx, vy, rec, K, md, it, u0, bx);
}

// ---------------------------------------------------------------------------
// The correction loop's tail: iterations first .. last of a residual-controlled
// MDF step in one launch.  Once the loop has stopped, the separate launches of
// the remaining iterations would each cost their dispatch while doing nothing
// but carry the decision forward; here the launch finds the decision at its
// start and leaves.  Each iteration that still runs is the separate launches'
// work -- the interpolation (mdf_interp_body) over node blocks, then the
// correction (spread_rest_body, spread_body, mdf_node_apply) -- with the
// same arguments, so results are the same bit for bit; a grid barrier stands
// where the kernel boundary stood.  The grid holds only as many workgroups as
// are resident at once (the launcher sizes it), which the barrier needs.
//
// The barrier: the last workgroup to arrive resets the count and advances the
// generation; the others wait for the generation to move, for at most `spin`
// polls (then bit 3 of *flag is raised and the workgroup leaves: the results
// are void and the host reports the failure).  The agent-scope release before
// arriving and acquire after leaving make each workgroup's global writes
// visible to the others across the XCDs' L2 caches.  bar[0] (count) and
// bar[32] (generation) sit on their own cache lines; the count is 0 between
// barriers.
__device__ bool grid_barrier(unsigned* bar, int* flag, unsigned spin) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* count = bar;
        unsigned* gen = bar + 32;
        const unsigned g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        int ok = 1;
        if (arrived == gridDim.x) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned n = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                __builtin_amdgcn_s_sleep(2);
                if (++n > spin) { ok = 0; break; }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (!ok) atomicOr(flag, 8);
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

struct TailArgs {
    CellVelArgs a;                       // a.F: set per iteration
    const double *lut, *x, *y, *vx, *vy;
    NodeSpread* rec;
    int K, nblk_interp;                  // nodes; node blocks of the interpolation
    int* tile_cnt;
    const int* tile_list;
    int2* act;
    int* counters;
    void *FP, *FQ;                       // the correction ping-pong: first reads FP, writes FQ
    uint8_t *dP, *dQ;                    // their tile flags
    int nx, ny, ntx, ntiles, nrest;
    MdfDev* md;
    double tol;
    int first, last;
    double *total, *head;
    unsigned* bar;
    unsigned spin;
};

template <bool FP32, int LPN>
__global__ __launch_bounds__(256) void k_mdf_tail(TailArgs t) {
    const int G = static_cast<int>(gridDim.x);
    const int nnode = (t.K + 255) / 256;
    const int napply = t.nrest + G + nnode;
    for (int it = t.first; it <= t.last; ++it) {
        const bool odd = ((it - t.first) & 1) != 0;
        void* Fin = odd ? t.FQ : t.FP;
        void* Fout = odd ? t.FP : t.FQ;
        uint8_t* din = odd ? t.dQ : t.dP;
        uint8_t* dout = odd ? t.dP : t.dQ;
        // written before this launch, or before the last barrier
        const int dec = __hip_atomic_load(&t.md->cby[it - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (dec) {
            // stopped: the decision holds for the remaining iterations, and
            // the last iteration's copy keeps both buffers of the ping-pong
            // equal
            if (blockIdx.x == 0 && threadIdx.x == 0)
                for (int j = it; j <= t.last; ++j) t.md->cby[j] = dec;
            const bool lodd = ((t.last - t.first) & 1) != 0;
            void* Lin = lodd ? t.FQ : t.FP;
            void* Lout = lodd ? t.FP : t.FQ;
            uint8_t* Ldin = lodd ? t.dQ : t.dP;
            uint8_t* Ldout = lodd ? t.dP : t.dQ;
            if (dec >= 2) {
                // A correction ran this step, so both buffers hold force in the
                // touched tiles only (each correction copied or cleared every
                // other tile): copying those tiles, one workgroup each, is the
                // whole of the last iteration's copy.  The force is in the last
                // iteration's input buffer when (last - dec) is even, else in
                // its output buffer.
                using T2 = typename std::conditional<FP32, float2, double2>::type;
                const bool back = ((t.last - dec) & 1) != 0;
                const T2* src = static_cast<const T2*>(back ? Lout : Lin);
                T2* dst = static_cast<T2*>(back ? Lin : Lout);
                uint8_t* ddst = back ? Ldin : Ldout;
                const int nact = min(t.counters[3], t.ntiles);
                for (int i = blockIdx.x; i < nact; i += G) {
                    const int tile = t.act[i].x;
                    const int cx = (tile % t.ntx) * TILE_X + static_cast<int>(threadIdx.x) % TILE_X;
                    const int cy = (tile / t.ntx) * TILE_Y + static_cast<int>(threadIdx.x) / TILE_X;
                    if (threadIdx.x < TILE_X * TILE_Y && cx < t.nx && cy < t.ny) {
                        const size_t n = static_cast<size_t>(cy) * t.nx + cx;
                        dst[n] = src[n];
                    }
                    if (threadIdx.x == 0) ddst[tile] = 1;
                }
            } else {
                // no correction ran: the other buffer may hold an older step's
                // force anywhere (spread_rest_body at last, as its own launch
                // would run it)
                for (int b = blockIdx.x; b < t.nrest; b += G)
                    spread_rest_body<FP32>(t.tile_cnt, Lin, Lout, Ldin, Ldout, t.act, t.counters, t.nx, t.ny,
                                           t.ntx, t.ntiles, 1, t.md, b, t.last, t.last, t.tol, dec);
            }
            return;
        }
        CellVelArgs a = t.a;
        a.F = Fin;
        for (int b = blockIdx.x; b < t.nblk_interp; b += G) {
            mdf_interp_body<FP32, LPN>(a, t.lut, t.x, t.y, t.vx, t.vy, t.rec, t.K, t.md, it, nullptr, b);
            __syncthreads();   // the body's LDS is reused by the next block
        }
        if (!grid_barrier(t.bar, t.counters + 2, t.spin)) return;
        for (int b = blockIdx.x; b < napply; b += G) {
            if (b < t.nrest)
                spread_rest_body<FP32>(t.tile_cnt, Fin, Fout, din, dout, t.act, t.counters, t.nx, t.ny, t.ntx,
                                       t.ntiles, 1, t.md, b, it, t.last, t.tol);
            else if (b < t.nrest + G)
                spread_body<FP32>(t.rec, nullptr, t.tile_cnt, t.tile_list, t.act, t.counters, t.lut, Fin, Fout, dout,
                                  t.nx, t.ny, t.ntx, t.ntiles, 1, 0, t.md, b - t.nrest, G, it, t.tol);
            else
                mdf_node_apply(t.md, t.total, t.head, t.rec, t.K, it, t.tol,
                               (b - t.nrest - G) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x));
        }
        if (!grid_barrier(t.bar, t.counters + 2, t.spin)) return;
    }
}

// The warm-start model's inputs of every node (mdf_node_features), and the
// plain first correction dF1 (the reference frame of the model's output).
__global__ void k_mdf_features(FeatArgs a, float* feat, const NodeSpread* rec, double* f1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.K) return;
    f1[2 * k] = rec[k].fx;
    f1[2 * k + 1] = rec[k].fy;
    mdf_node_features(a, k, feat + static_cast<size_t>(k) * MLP_IN);
}

// Clears the per-step MDF state of a batch member (graph-safe: a kernel, not a
// memset).
__device__ __forceinline__ void mdf_begin_body(MdfDev* md, double* total, double* head, int K, int bx) {
    const int i = bx * blockDim.x + threadIdx.x;
    if (i == 0) { md->conv = 0; md->applied = 0; md->conv_it = 0; }
    if (i < MDF_RES_LD) md->cby[i] = 0;
    if (i < MDF_RES_SUB * (MDF_MAX_CORR + 2)) md->res[i / (MDF_MAX_CORR + 2)][i % (MDF_MAX_CORR + 2)] = 0ull;
    if (i < 2 * K) { total[i] = 0.0; head[i] = 0.0; }
}


// The model's targets: the sum T of the step's first MDF_HEAD plain
// corrections in the frame of the first one, dF1: T = g_par dF1 +
// g_perp rot90(dF1), clamped to [-MDF_GAIN_MAX, MDF_GAIN_MAX] (unit gain where
// dF1 vanishes).  Predicting more of the plain iteration in one shot (the
// total of four corrections, a gain near 2.2) overshoots the modes of stiffer
// local coupling -- nearby membranes, where one correction already moves the
// interpolated velocity further -- and the coupled run diverges within a few
// hundred steps; the two-correction head stays stable and the plain
// corrections that follow take the residual down to the tolerance.
__device__ __forceinline__ void gains_of(double tx, double ty, double ax, double ay, float& gp,
                                         float& gq) {
    const double a = ax * ax + ay * ay;
    if (!(a > 1e-40)) { gp = 1.f; gq = 0.f; return; }
    const double p = (tx * ax + ty * ay) / a, q = (ax * ty - ay * tx) / a;
    gp = static_cast<float>(fmin(fmax(p, -MDF_GAIN_MAX), MDF_GAIN_MAX));
    gq = static_cast<float>(fmin(fmax(q, -MDF_GAIN_MAX), MDF_GAIN_MAX));
}

// End of the MDF stage: shift the per-node slip history, append sample rows
// (inputs, gains of the total correction applied) and accumulate the step
// statistics.
__global__ void k_mdf_record(MdfDev* md, const double* total, const double* head, const double* u0,
                             const double* f1, float* hist, const float* feat, float* ds,
                             int64_t ds_cap, int K, int hist_on) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (ds) {
        const int64_t row = md->ds_rows + k;
        if (md->ds_rows + K <= ds_cap) {
            float* r = ds + row * (MLP_IN + MLP_OUT);
            const float* f = feat + static_cast<size_t>(k) * MLP_IN;
            for (int j = 0; j < MLP_IN; ++j) r[j] = f[j];
            float gp, gq;
            gains_of(head[2 * k], head[2 * k + 1], f1[2 * k], f1[2 * k + 1], gp, gq);
            r[MLP_IN] = gp;
            r[MLP_IN + 1] = gq;
        }
    }
    if (hist_on) {
        for (int h = MLP_HIST - 1; h > 0; --h) {
            hist[(static_cast<size_t>(h) * K + k) * 2] = hist[(static_cast<size_t>(h - 1) * K + k) * 2];
            hist[(static_cast<size_t>(h) * K + k) * 2 + 1] = hist[(static_cast<size_t>(h - 1) * K + k) * 2 + 1];
        }
        hist[2 * static_cast<size_t>(k)] = static_cast<float>(u0[2 * k]);
        hist[2 * static_cast<size_t>(k) + 1] = static_cast<float>(u0[2 * k + 1]);
    }
}

__device__ __forceinline__ void mdf_finish_body(MdfDev* md, int ncorr, int probe, int K, int collect,
                                                int64_t ds_cap) {
    const int dec = md->cby[ncorr];
    md->conv = dec != 0;
    md->conv_it = dec;
    const int it_final = dec ? dec : (probe ? ncorr + 1 : ncorr);
    const double r = mdf_res(md, it_final);
    md->sum_applied += md->applied;
    md->sum_final += r;
    md->max_final = (r > md->max_final || r != r) ? r : md->max_final;
    md->steps += 1.0;
    if (collect && md->ds_rows + K <= ds_cap) md->ds_rows += K;
}

// The statistics fold (one thread) and the tile counts of the step cleared
// (every thread), which the MDF corrections leave set.
__global__ void k_mdf_finish(MdfDev* md, int ncorr, int probe, int K, int collect, int64_t ds_cap, int* tile_cnt,
                             int ntiles, int* act_cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        mdf_finish_body(md, ncorr, probe, K, collect, ds_cap);
        *act_cnt = 0;
    }
    if (t < ntiles) tile_cnt[t] = 0;
}

// interpolateVelocity (immersed_boundary.cpp:174-189) with the post-step
// velocity u = macro(f_{n+1}, F_n).
// With move set (no rigid capsule present) the node is also advanced here:
// moveNodes (capsule.cpp:287-296) and the y-clamp (simulation.cpp:665-675).
struct MoveArgs {
    int move;
    double dt;
    int periodic_nx, clamp_y;
    double ymin, ymax;
};

__device__ __forceinline__ void move_node(const MoveArgs& m, double& px, double& py, double ux,
                                          double uy) {
    px = px + ux * m.dt;
    py = py + uy * m.dt;
    if (m.periodic_nx > 0 && isfinite(px)) {
        // repeated shifts as the reference does; a bounded count keeps a
        // runaway coordinate from looping (it is then folded with fmod)
        const double Lx = static_cast<double>(m.periodic_nx);
        if (fabs(px) > 64.0 * Lx) px = fmod(px, Lx);
        while (px < 0.0) px += Lx;
        while (px >= Lx) px -= Lx;
    }
    if (m.clamp_y) {
        if (py < m.ymin) py = m.ymin;
        if (py > m.ymax) py = m.ymax;
    }
}

template <bool FP32, int LPN = 16>
__device__ __forceinline__ void interp_body(CellVelArgs a, const double* lut, double* x,
                                                double* y, double* vx, double* vy, int K, MoveArgs mv, int bx,
                                                const CellFill& cf = CellFill{}) {
    constexpr int NPW = interp_nodes(LPN);
    __shared__ double tx[16 * NPW], ty[16 * NPW];
    __shared__ int use[16 * NPW];
    const int g = threadIdx.x / LPN, l = threadIdx.x % LPN;
    const int k = bx * NPW + g;
    const bool live = k < K;
    double px = live ? x[k] : 0.0, py = live ? y[k] : 0.0;
    double Ux, Uy;
    interp_group<FP32, LPN>(a, lut, px, py, 1, l, tx + 16 * g, ty + 16 * g, use + 16 * g, Ux, Uy);
    if (live && l == 0) {
        vx[k] = Ux;
        vy[k] = Uy;
        if (mv.move) {
            move_node(mv, px, py, Ux, Uy);
            x[k] = px;
            y[k] = py;
        }
    }
    if (cf.cnt) {
        // the next step's contact cell list, from the moved positions: the
        // workgroup's nodes claim their slots together from its first
        // wavefront, in node order, so a run of consecutive nodes in one cell
        // shares an atomic across the whole workgroup
        static_assert(NPW <= 64, "one wavefront claims for the workgroup");
        __shared__ int s_cell[NPW];
        if (l == 0)
            s_cell[g] = live ? contact_cell(px, py, cf.xorg, cf.Lx, cf.cdx, cf.cdy, cf.ybot, cf.ncx, cf.ncy) : -1;
        __syncthreads();
        if (threadIdx.x < 64) {
            const int i = static_cast<int>(threadIdx.x);
            contact_cell_claim(cf.cnt, cf.list, cf.counters, i < NPW ? s_cell[i] : -1, bx * NPW + i);
        }
    }
}

// What the interpolation that ends a step also does: clear the bins the
// step's spread left open (tile counts, touched-tile count; tile_cnt null:
// none) and fold the step's MDF statistics (mdf_finish_body; md null: none).
struct MdfFold {
    MdfDev* md;
    int ncorr, probe, K, collect;
    int64_t ds_cap;
    int* tile_cnt;
    int ntiles;
    int* act_cnt;
};

__device__ __forceinline__ void mdf_fold(const MdfFold& f) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x, nthr = gridDim.x * blockDim.x;
    if (f.tile_cnt) {
        for (int t = gid; t < f.ntiles; t += nthr) f.tile_cnt[t] = 0;
        if (gid == 0) *f.act_cnt = 0;
    }
    if (gid == 0 && f.md) mdf_finish_body(f.md, f.ncorr, f.probe, f.K, f.collect, f.ds_cap);
}

template <bool FP32, int LPN>
__global__ __launch_bounds__(256) void k_interp(CellVelArgs a, const double* lut, double* x,
                                                double* y, double* vx, double* vy, int K, MoveArgs mv,
                                                int band, MdfFold fold, CellFill cf) {
    if (fold.md || fold.tile_cnt) mdf_fold(fold);
    const int bx = band_block(blockIdx.x, band);
    if (bx * interp_nodes(LPN) >= K) return;       // padding workgroup of the banded grid
    interp_body<FP32, LPN>(a, lut, x, y, vx, vy, K, mv, bx, cf);
}

// moveAllNodes + y-clamp (capsule.cpp:267-297, simulation.cpp:665-675); one
// wavefront per capsule so rigid capsules can project their velocities.
__global__ void k_move(double* x, double* y, double* vx, double* vy, const int* off,
                       const sf_capsule_desc* desc, int ncaps, double dt, int periodic_nx,
                       int clamp_y, double ymin, double ymax) {
    const int cap = blockIdx.x;
    if (cap >= ncaps) return;
    const int o = off[cap], n = off[cap + 1] - o;
    const int lane = threadIdx.x;
    if (desc[cap].is_rigid) {
        // centroid with periodic minimum image about node 0 (capsule.cpp:159-173)
        __shared__ double ctr[2], vcm[2], om;
        if (lane == 0) {
            double sx, sy;
            if (periodic_nx > 0) {
                const double Lx = static_cast<double>(periodic_nx);
                const double rx = x[o], ry = y[o];
                sx = rx; sy = ry;
                for (int i = 1; i < n; ++i) {
                    double dx = x[o + i] - rx, dy = y[o + i] - ry;
                    if (dx > 0.5 * Lx) dx -= Lx;
                    if (dx < -0.5 * Lx) dx += Lx;
                    sx += rx + dx; sy += ry + dy;
                }
            } else {
                sx = 0.0; sy = 0.0;
                for (int i = 0; i < n; ++i) { sx += x[o + i]; sy += y[o + i]; }
            }
            ctr[0] = sx / static_cast<double>(n); ctr[1] = sy / static_cast<double>(n);
            double cx = 0.0, cy = 0.0;
            for (int i = 0; i < n; ++i) { cx += vx[o + i]; cy += vy[o + i]; }
            const double inv = 1.0 / n;
            cx *= inv; cy *= inv;
            double num = 0.0, den = 0.0;
            for (int i = 0; i < n; ++i) {
                const double rx = x[o + i] - ctr[0], ry = y[o + i] - ctr[1];
                num += rx * vy[o + i] - ry * vx[o + i];
                den += rx * rx + ry * ry;
            }
            vcm[0] = cx; vcm[1] = cy;
            om = (den > 1e-10) ? num / den : 0.0;
        }
        __syncthreads();
        for (int i = lane; i < n; i += blockDim.x) {
            const double rx = x[o + i] - ctr[0], ry = y[o + i] - ctr[1];
            vx[o + i] = vcm[0] - om * ry;
            vy[o + i] = vcm[1] + om * rx;
        }
        __syncthreads();
    }
    const MoveArgs mv{1, dt, periodic_nx, clamp_y, ymin, ymax};
    for (int i = lane; i < n; i += blockDim.x) {
        double px = x[o + i], py = y[o + i];
        move_node(mv, px, py, vx[o + i], vy[o + i]);
        x[o + i] = px;
        y[o + i] = py;
    }
}

static CellVelArgs vel_args(Ctx& c, const void* f, const void* F) {
    CellVelArgs a{};
    a.f = f; a.F = F;
    a.nx = c.nx; a.ny = c.ny; a.N = c.N;
    a.walls_x = c.walls_x; a.walls_y = c.walls_y;
    a.flags = c.flags; a.generic_flags = c.generic_flags;
    a.has_body = (c.p.body_force_x != 0.0 || c.p.body_force_y != 0.0);
    a.bfx = c.p.body_force_x; a.bfy = c.p.body_force_y;
    a.fmax = c.p.max_lattice_force; a.fmax2 = a.fmax * a.fmax;
    return a;
}

// workgroups per XCD band of the node-parallel interpolations (K nodes, lpn
// lanes per node)
static int interp_band(int K, int lpn) {
    const int blocks = (K + interp_nodes(lpn) - 1) / interp_nodes(lpn);
    return (blocks + 7) / 8;
}

// k_interp at lpn lanes per node (an interpolation with mv.move = 0 only
// writes the velocities)
static void launch_interp_lpn(Ctx& c, int lpn, const CellVelArgs& a, double* vx, double* vy, const MoveArgs& mv,
                              const MdfFold& fold, const CellFill& cf = CellFill{}) {
    const int band = interp_band(c.K, lpn);
    const dim3 grid(8 * band);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, c.stream, a, c.phi, c.x, c.y, vx, vy, c.K, mv, band, fold,
                           cf);
    };
    if (c.fp32) {
        if (lpn == 4) go(k_interp<true, 4>);
        else if (lpn == 8) go(k_interp<true, 8>);
        else go(k_interp<true, 16>);
    } else {
        if (lpn == 4) go(k_interp<false, 4>);
        else if (lpn == 8) go(k_interp<false, 8>);
        else go(k_interp<false, 16>);
    }
}

// Bin the nodes and spread their forces into F[Fcur] (overwrite); mdf: the
// binning also clears the step's MDF state.  The bins stay open (tile counts,
// touched-tile list) for the step's MDF corrections until the step's closing
// interpolation or launch_bins_close clears them.
void launch_spread(Ctx& c, bool mdf) {
    if (c.K == 0) {
        launch_fill_zero(c, c.F[c.Fcur], (c.fp32 ? 8 : 16) * static_cast<size_t>(c.N));
        return;
    }
    launch_bins_close(c);   // a binning no interpolation closed (the autotuner's repeats)
    // the membrane kernel left the arc lengths of the current positions
    const double* arc = c.arc_fresh ? c.arc : nullptr;
    hipLaunchKernelGGL(k_node_prep, dim3((c.K + 255) / 256), dim3(256), 0, c.stream, c.x, c.y,
                       c.fx, c.fy, c.ring, arc, c.spread, c.K, c.nx, c.ny, c.ntx, c.nty,
                       c.tile_cnt, c.tile_list, c.tile_rec, c.counters, 0, mdf ? c.mdf : nullptr, 0.0, 0,
                       mdf ? c.mdf_total : nullptr, mdf ? c.mdf_head : nullptr, c.tile_act);
    c.bins_open = true;
    void* F = c.F[c.Fcur];
    uint8_t* d = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    const int nrest = (c.ntiles + rest_tiles_per_block(0) - 1) / rest_tiles_per_block(0);
    const int nspread = c.spread_grid > 0 ? std::min(c.ntiles, c.spread_grid) : c.ntiles;
    const dim3 g(nrest + nspread);
    if (c.fp32)
        hipLaunchKernelGGL((k_spread_step<true>), g, dim3(256), 0, c.stream, c.spread, c.tile_rec, c.tile_cnt,
                           c.tile_list, c.tile_act, c.counters, c.phi, F, d, c.nx, c.ny, c.ntx, c.ntiles, nrest);
    else
        hipLaunchKernelGGL((k_spread_step<false>), g, dim3(256), 0, c.stream, c.spread, c.tile_rec, c.tile_cnt,
                           c.tile_list, c.tile_act, c.counters, c.phi, F, d, c.nx, c.ny, c.ntx, c.ntiles, nrest);
}

// Clear bins a spread left open when no closing interpolation follows.
void launch_bins_close(Ctx& c) {
    if (!c.bins_open) return;
    if (c.tile_cnt) launch_fill_int(c, c.tile_cnt, c.ntiles, 0);
    launch_fill_int(c, c.counters + 3, 1, 0);
    c.bins_open = false;
}

// MDF correction it of last_it: the correction on top of F[Fcur] into the
// buffer that is neither current nor previous, which becomes current; the
// previous contents become F[Fpart].
void launch_mdf_apply(Ctx& c, int it, int last_it) {
    const void* Fin = c.F[c.Fcur];
    const int out = 3 - c.Fcur - c.Fprev;
    void* Fout = c.F[out];
    const uint8_t* din = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    uint8_t* dout = c.tile_dirty + static_cast<size_t>(out) * c.ntiles;
    const int nrest = (c.ntiles + rest_tiles_per_block(1) - 1) / rest_tiles_per_block(1);
    const int nspread = c.spread_grid > 0 ? std::min(c.ntiles, c.spread_grid) : c.ntiles;
    const int nnode = (c.K + 255) / 256;
    const dim3 g(nrest + nspread + nnode);
    if (c.fp32)
        hipLaunchKernelGGL((k_mdf_apply<true>), g, dim3(256), 0, c.stream, c.spread, c.tile_cnt, c.tile_list,
                           c.tile_act, c.counters, c.phi, Fin, Fout, din, dout, c.nx, c.ny, c.ntx, c.ntiles, c.mdf,
                           c.mdf_tol, it, last_it, c.mdf_total, c.mdf_head, c.K, nrest, nspread);
    else
        hipLaunchKernelGGL((k_mdf_apply<false>), g, dim3(256), 0, c.stream, c.spread, c.tile_cnt, c.tile_list,
                           c.tile_act, c.counters, c.phi, Fin, Fout, din, dout, c.nx, c.ny, c.ntx, c.ntiles, c.mdf,
                           c.mdf_tol, it, last_it, c.mdf_total, c.mdf_head, c.K, nrest, nspread);
    c.Fpart = c.Fcur;
    c.Fcur = out;
}

// Workgroups of the correction loop's tail: one per CU, as long as one fits
// next to nothing else (the launch follows the step's other kernels in
// stream order, so the whole grid is resident and the grid barrier can
// complete).  0: the tail is not available (the loop stays one launch pair per
// iteration).  Sized outside any stream capture (sf_mdf_set_control).
int mdf_tail_size(Ctx& c) {
    if (c.mdf_tail_grid >= 0) return c.mdf_tail_grid;
    int cus = 0, per = 0;
    c.mdf_tail_grid = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess) return 0;
    auto occ = [&](auto kern) { return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0); };
    hipError_t e;
    if (c.fp32) e = c.interp_lpn == 4 ? occ(k_mdf_tail<true, 4>) : c.interp_lpn == 8 ? occ(k_mdf_tail<true, 8>)
                                                                                     : occ(k_mdf_tail<true, 16>);
    else e = c.interp_lpn == 4 ? occ(k_mdf_tail<false, 4>) : c.interp_lpn == 8 ? occ(k_mdf_tail<false, 8>)
                                                                               : occ(k_mdf_tail<false, 16>);
    if (e == hipSuccess && per >= 1) c.mdf_tail_grid = cus;
    return c.mdf_tail_grid;
}

bool mdf_tail_ready(const Ctx& c) { return c.mdf_bar && c.mdf_tail_grid > 0; }

// MDF iterations first .. last in one launch (k_mdf_tail); rotates the force
// roles as that many launch_mdf_apply calls do.
void launch_mdf_tail(Ctx& c, int first, int last) {
    const int out = 3 - c.Fcur - c.Fprev;
    TailArgs t{};
    t.a = vel_args(c, c.f[c.fcur], nullptr);
    t.lut = c.phi; t.x = c.x; t.y = c.y; t.vx = c.vx; t.vy = c.vy;
    t.rec = c.spread;
    t.K = c.K;
    t.nblk_interp = (c.K + interp_nodes(c.interp_lpn) - 1) / interp_nodes(c.interp_lpn);
    t.tile_cnt = c.tile_cnt; t.tile_list = c.tile_list; t.act = c.tile_act; t.counters = c.counters;
    t.FP = c.F[c.Fcur];
    t.FQ = c.F[out];
    t.dP = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    t.dQ = c.tile_dirty + static_cast<size_t>(out) * c.ntiles;
    t.nx = c.nx; t.ny = c.ny; t.ntx = c.ntx; t.ntiles = c.ntiles;
    t.nrest = (c.ntiles + rest_tiles_per_block(1) - 1) / rest_tiles_per_block(1);
    t.md = c.mdf;
    t.tol = c.mdf_tol;
    t.first = first;
    t.last = last;
    t.total = c.mdf_total;
    t.head = c.mdf_head;
    t.bar = c.mdf_bar;
    t.spin = 1u << 20;
    const dim3 g(c.mdf_tail_grid);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, dim3(256), 0, c.stream, t); };
    if (c.fp32) {
        if (c.interp_lpn == 4) go(k_mdf_tail<true, 4>);
        else if (c.interp_lpn == 8) go(k_mdf_tail<true, 8>);
        else go(k_mdf_tail<true, 16>);
    } else {
        if (c.interp_lpn == 4) go(k_mdf_tail<false, 4>);
        else if (c.interp_lpn == 8) go(k_mdf_tail<false, 8>);
        else go(k_mdf_tail<false, 16>);
    }
    for (int it = first; it <= last; ++it) {
        const int o = 3 - c.Fcur - c.Fprev;
        c.Fpart = c.Fcur;
        c.Fcur = o;
    }
}

// MDF iteration it (1-based; MDF_MAX_CORR + 1 at most, the last being the
// residual probe after the final correction).
void launch_mdf_interp(Ctx& c, int it) {
    if (c.K == 0) return;
    CellVelArgs a = vel_args(c, c.f[c.fcur], c.F[c.Fcur]);
    const int lpn = c.interp_lpn;
    const int band = interp_band(c.K, lpn);
    const dim3 grid(8 * band);
    double* u0 = it == 1 && c.mdf_hist_on ? c.mdf_u0 : nullptr;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, c.stream, a, c.phi, c.x, c.y, c.vx, c.vy, c.spread, c.K,
                           c.mdf, it, u0, band);
    };
    if (c.fp32) {
        if (lpn == 4) go(k_mdf_interp<true, 4>);
        else if (lpn == 8) go(k_mdf_interp<true, 8>);
        else go(k_mdf_interp<true, 16>);
    } else {
        if (lpn == 4) go(k_mdf_interp<false, 4>);
        else if (lpn == 8) go(k_mdf_interp<false, 8>);
        else go(k_mdf_interp<false, 16>);
    }
}


FeatArgs feat_args(const Ctx& c) {
    return FeatArgs{c.x, c.y, c.fx, c.fy, c.off, c.owner, c.mdf_u0, c.mdf_hist, c.mdf_dmin, c.K,
                    c.periodic_x ? static_cast<double>(c.nx) : 0.0, c.p.periodic_y ? -1e30 : 0.5,
                    c.p.periodic_y ? 1e30 : static_cast<double>(c.ny) - 1.5, 2.0 * c.p.rep_r_cut};
}

void launch_mdf_features(Ctx& c) {
    hipLaunchKernelGGL(k_mdf_features, dim3((c.K + 255) / 256), dim3(256), 0, c.stream, feat_args(c), c.mdf_feat,
                       c.spread, c.mdf_f1);
}

void launch_mdf_record(Ctx& c, int ncorr) {
    const bool collect = c.ds_buf != nullptr;
    if (collect || c.mdf_hist_on)
        hipLaunchKernelGGL(k_mdf_record, dim3((c.K + 255) / 256), dim3(256), 0, c.stream, c.mdf,
                           c.mdf_total, c.mdf_head, c.mdf_u0, c.mdf_f1, c.mdf_hist, c.mdf_feat, c.ds_buf,
                           c.ds_cap, c.K,
                           c.mdf_hist_on);
    // the statistics fold and the tile counts ride on the step's closing
    // interpolation (launch_interp_move), or launch_mdf_fold when none follows
    c.mdf_fold_pending = true;
    c.mdf_fold_ncorr = ncorr;
}

void launch_mdf_fold(Ctx& c) {
    if (!c.mdf_fold_pending) return;
    hipLaunchKernelGGL(k_mdf_finish, dim3((c.ntiles + 255) / 256), dim3(256), 0, c.stream, c.mdf,
                       c.mdf_fold_ncorr, c.mdf_probe, c.K, c.ds_buf != nullptr ? 1 : 0, c.ds_cap, c.tile_cnt,
                       c.ntiles, c.counters + 3);
    c.mdf_fold_pending = false;
    c.bins_open = false;   // the tile counts and the touched-tile count are cleared
}

// Capsule recirculation in x over a non-periodic fluid (Simulation::step 10a,
// simulation.cpp:687-700): a capsule whose centroid -- the plain node mean,
// since capsules carry no periodic length there (capsule.cpp:168-172) -- has
// left [0, nx) moves rigidly by -nx or +nx.  One thread per capsule sums its
// nodes in index order, as the reference's centroid does.
__global__ void k_recirculate_x(double* x, const int* off, int ncaps, double Lx) {
    const int cap = blockIdx.x * blockDim.x + threadIdx.x;
    if (cap >= ncaps) return;
    const int o = off[cap], n = off[cap + 1] - o;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[o + i];
    const double cx = s / static_cast<double>(n);
    double shift = 0.0;
    if (cx >= Lx) shift = -Lx;
    else if (cx < 0.0) shift = Lx;
    if (shift != 0.0)
        for (int i = 0; i < n; ++i) x[o + i] += shift;
}

void launch_interp_probe(Ctx& c, int lpn, double* vx, double* vy, const CellFill& cf) {
    if (c.K == 0) return;
    const MoveArgs mv{0, c.p.dt, 0, 0, 0.0, 0.0};
    launch_interp_lpn(c, lpn, vel_args(c, c.f[c.fcur], c.F[c.Fcur]), vx, vy, mv, MdfFold{}, cf);
}

void launch_interp_move(Ctx& c) {
    c.arc_fresh = false;   // the nodes move
    if (c.K == 0) return;
    const CellVelArgs a = vel_args(c, c.f[c.fcur], c.F[c.Fcur]);
    const MoveArgs mv{c.any_rigid ? 0 : 1, c.p.dt, c.periodic_x ? c.nx : 0, c.p.periodic_y ? 0 : 1,
                      1.5, static_cast<double>(c.ny) - 1.5};
    MdfFold fold{};
    if (c.mdf_fold_pending) {
        fold = MdfFold{c.mdf, c.mdf_fold_ncorr, c.mdf_probe, c.K, c.ds_buf != nullptr ? 1 : 0, c.ds_cap};
        c.mdf_fold_pending = false;
    }
    if (c.bins_open) {
        fold.tile_cnt = c.tile_cnt;
        fold.ntiles = c.ntiles;
        fold.act_cnt = c.counters + 3;
        c.bins_open = false;
    }
    launch_interp_lpn(c, c.interp_lpn, a, c.vx, c.vy, mv, fold, contact_fill_ahead(c, mv.move != 0));
    if (c.any_rigid)
        hipLaunchKernelGGL(k_move, dim3(c.ncaps), dim3(64), 0, c.stream, c.x, c.y, c.vx, c.vy, c.off,
                           c.desc, c.ncaps, c.p.dt, c.periodic_x ? c.nx : 0, c.p.periodic_y ? 0 : 1,
                           1.5, static_cast<double>(c.ny) - 1.5);
    if (c.p.capsule_periodic_x && !c.periodic_x)
        hipLaunchKernelGGL(k_recirculate_x, dim3((c.ncaps + 255) / 256), dim3(256), 0, c.stream, c.x,
                           c.off, c.ncaps, static_cast<double>(c.nx));
}


// ---------------------------------------------------------------------------
// Batched launches over the members of an ensemble (map[b] = member and its
// own workgroup index), for the steady step: node forces spread once, no MDF.
// mode, md, tol, it, total, head: node_prep_body's MDF arguments (mode 1: an
// MDF correction; md null in mode 0)
struct PrepPack {
    const double *x, *y, *fx, *fy;
    const int2* ring;
    NodeSpread* rec;
    int K, nx, ny, ntx, nty;
    int* tile_cnt;
    int* tile_list;
    NodeSpread* tile_rec;
    int* counters;
    int mode;
    MdfDev* md;
    double tol;
    int it;
    double *total, *head;
};
struct RestPack {
    const int* tile_cnt;
    const void* Fin;
    void* Fout;
    const uint8_t* din;
    uint8_t* dout;
    int2* act;
    int* counters;
    int nx, ny, ntx, ntiles;
    int mode;
    const MdfDev* md;
    int it, last_it;
    double tol;
};
struct SpreadPack {
    const NodeSpread* rec;
    const NodeSpread* tile_rec;
    int* tile_cnt;
    const int* tile_list;
    const int2* act;
    const int* counters;
    const double* lut;
    const void* Fin;
    void* Fout;
    uint8_t* dout;
    int nx, ny, ntx, ntiles, gx;
    int mode, last;
    const MdfDev* md;
    int it;
    double tol;
};
struct MdfBeginPack {
    MdfDev* md;
    double *total, *head;
    int K;
};
struct MdfInterpPack {
    CellVelArgs a;
    const double* lut;
    const double *x, *y, *vx, *vy;
    NodeSpread* rec;
    int K;
    MdfDev* md;
    int it;
};
struct MdfFinishPack {
    MdfDev* md;
    int ncorr, probe, K;
    int* tile_cnt;
    int ntiles;
};

size_t batch_pack_size(int kind) {
    switch (kind) {
    case BK_NODE_PREP: case BK_MDF_PREP: return sizeof(PrepPack);
    case BK_SPREAD_REST: case BK_MDF_REST: return sizeof(RestPack);
    case BK_SPREAD: case BK_MDF_SPREAD: return sizeof(SpreadPack);
    case BK_MDF_BEGIN: return sizeof(MdfBeginPack);
    case BK_MDF_INTERP: return sizeof(MdfInterpPack);
    case BK_MDF_FINISH: return sizeof(MdfFinishPack);
    default: return 0;
    }
}
struct InterpPack {
    CellVelArgs a;
    MoveArgs mv;
    const double* lut;
    double *x, *y, *vx, *vy;
    int K;
};

__global__ __launch_bounds__(256) void kb_node_prep(const PrepPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const PrepPack p = packs[w.x];
    node_prep_body(p.x, p.y, p.fx, p.fy, p.ring, nullptr, p.rec, p.K, p.nx, p.ny, p.ntx, p.nty,
                   p.tile_cnt, p.tile_list, p.tile_rec, p.counters, p.mode, p.md, p.tol, p.it, p.total,
                   p.head, nullptr, w.y);
}

template <bool FP32>
__global__ __launch_bounds__(256) void kb_spread_rest(const RestPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const RestPack p = packs[w.x];
    spread_rest_body<FP32>(p.tile_cnt, p.Fin, p.Fout, p.din, p.dout, p.act, p.counters, p.nx, p.ny,
                           p.ntx, p.ntiles, p.mode, p.md, w.y, p.it, p.last_it, p.tol);
}

template <bool FP32>
__global__ __launch_bounds__(256) void kb_spread(const SpreadPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const SpreadPack p = packs[w.x];
    spread_body<FP32>(p.rec, p.tile_rec, p.tile_cnt, p.tile_list, p.act, p.counters, p.lut, p.Fin,
                      p.Fout, p.dout, p.nx, p.ny, p.ntx, p.ntiles, p.mode, p.last, p.md, w.y, p.gx, p.it, p.tol);
}

__global__ __launch_bounds__(256) void kb_mdf_begin(const MdfBeginPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const MdfBeginPack p = packs[w.x];
    mdf_begin_body(p.md, p.total, p.head, p.K, w.y);
}

template <bool FP32>
__global__ __launch_bounds__(256) void kb_mdf_interp(const MdfInterpPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const MdfInterpPack p = packs[w.x];
    mdf_interp_body<FP32>(p.a, p.lut, p.x, p.y, p.vx, p.vy, p.rec, p.K, p.md, p.it, nullptr, w.y);
}

// one workgroup per member; its first thread folds the step's MDF statistics
// and all its threads clear the member's tile counts
__global__ __launch_bounds__(256) void kb_mdf_finish(const MdfFinishPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const MdfFinishPack p = packs[w.x];
    if (threadIdx.x == 0) mdf_finish_body(p.md, p.ncorr, p.probe, p.K, 0, 0);
    for (int t = threadIdx.x; t < p.ntiles; t += blockDim.x) p.tile_cnt[t] = 0;
}

template <bool FP32>
__global__ __launch_bounds__(256) void kb_interp(const InterpPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const InterpPack p = packs[w.x];
    interp_body<FP32>(p.a, p.lut, p.x, p.y, p.vx, p.vy, p.K, p.mv, w.y);
}

// Workgroups of one member's spread gather in a batch: each walks the touched
// tiles with this stride; the rest exit at once (touched tiles are far fewer
// than tiles).
constexpr int BATCH_SPREAD_GRID = 64;

// The arguments of launch_spread(c, 0, last) (no capsules: not batched).
void batch_items_spread(Ctx& c, BatchItem* it, bool last) {
    batch_put(it[0], (c.K + 255) / 256,
              PrepPack{c.x, c.y, c.fx, c.fy, c.ring, c.spread, c.K, c.nx, c.ny, c.ntx, c.nty,
                       c.tile_cnt, c.tile_list, c.tile_rec, c.counters, 0, nullptr, 0.0, 0, nullptr, nullptr});
    const void* Fin = c.F[c.Fcur];
    void* Fout = c.F[c.Fcur];
    const uint8_t* din = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    uint8_t* dout = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    batch_put(it[1], (c.ntiles + rest_tiles_per_block(0) - 1) / rest_tiles_per_block(0),
              RestPack{c.tile_cnt, Fin, Fout, din, dout, c.tile_act, c.counters, c.nx, c.ny, c.ntx,
                       c.ntiles, 0, nullptr, 0, 0, 0.0});
    const int gx = std::min(c.ntiles, BATCH_SPREAD_GRID);
    batch_put(it[2], gx,
              SpreadPack{c.spread, c.tile_rec, c.tile_cnt, c.tile_list, c.tile_act, c.counters, c.phi,
                         Fin, Fout, dout, c.nx, c.ny, c.ntx, c.ntiles, gx, 0, last ? 1 : 0, nullptr, 0, 0.0});
}

// the clear of a member's MDF state at the start of its step (a single
// context clears it in its binning pass)
void batch_items_mdf_begin(Ctx& c, BatchItem& it) {
    const int n = std::max(2 * c.K, MDF_RES_SUB * (MDF_MAX_CORR + 2));
    batch_put(it, (n + 255) / 256, MdfBeginPack{c.mdf, c.mdf_total, c.mdf_head, c.K});
}

// launch_mdf_interp(c, i)'s arguments (no model: no first-slip record)
void batch_items_mdf_interp(Ctx& c, int i, BatchItem& it) {
    batch_put(it, (c.K + INTERP_NODES - 1) / INTERP_NODES,
              MdfInterpPack{vel_args(c, c.f[c.fcur], c.F[c.Fcur]), c.phi, c.x, c.y, c.vx, c.vy, c.spread,
                            c.K, c.mdf, i});
}

// MDF correction i of ncorr as four launches: the interpolation, the node part
// (node_prep mode 1), the untouched tiles and the touched tiles -- the parts
// launch_mdf_apply runs as one; rotates the force roles as it does
void batch_items_mdf_correction(Ctx& c, int i, int ncorr, BatchItem* it) {
    batch_items_mdf_interp(c, i, it[0]);
    batch_put(it[1], (c.K + 255) / 256,
              PrepPack{c.x, c.y, c.fx, c.fy, c.ring, c.spread, c.K, c.nx, c.ny, c.ntx, c.nty,
                       c.tile_cnt, c.tile_list, c.tile_rec, c.counters, 1, c.mdf, c.mdf_tol, i,
                       c.mdf_total, c.mdf_head});
    const void* Fin = c.F[c.Fcur];
    const int out = 3 - c.Fcur - c.Fprev;
    void* Fout = c.F[out];
    const uint8_t* din = c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles;
    uint8_t* dout = c.tile_dirty + static_cast<size_t>(out) * c.ntiles;
    batch_put(it[2], (c.ntiles + rest_tiles_per_block(1) - 1) / rest_tiles_per_block(1),
              RestPack{c.tile_cnt, Fin, Fout, din, dout, c.tile_act, c.counters, c.nx, c.ny, c.ntx,
                       c.ntiles, 1, c.mdf, i, ncorr, c.mdf_tol});
    const int gx = std::min(c.ntiles, BATCH_SPREAD_GRID);
    batch_put(it[3], gx,
              SpreadPack{c.spread, c.tile_rec, c.tile_cnt, c.tile_list, c.tile_act, c.counters, c.phi,
                         Fin, Fout, dout, c.nx, c.ny, c.ntx, c.ntiles, gx, 1, 0, c.mdf, i, c.mdf_tol});
    c.Fpart = c.Fcur;
    c.Fcur = out;
}

// launch_mdf_record's statistics (no model, no sample collection)
void batch_items_mdf_finish(Ctx& c, int ncorr, BatchItem& it) {
    batch_put(it, 1, MdfFinishPack{c.mdf, ncorr, c.mdf_probe, c.K, c.tile_cnt, c.ntiles});
}

// a slot this member takes no part in: no workgroups, a pack of the slot's size
void batch_items_none(int kind, BatchItem& it) {
    it.blocks = 0;
    it.pack.assign(batch_pack_size(kind), 0);
}

// The arguments of launch_interp_move (deformable capsules only).
void batch_items_interp(Ctx& c, BatchItem* it) {
    const CellVelArgs a = vel_args(c, c.f[c.fcur], c.F[c.Fcur]);
    const MoveArgs mv{1, c.p.dt, c.periodic_x ? c.nx : 0, c.p.periodic_y ? 0 : 1, 1.5,
                      static_cast<double>(c.ny) - 1.5};
    batch_put(it[0], (c.K + INTERP_NODES - 1) / INTERP_NODES,
              InterpPack{a, mv, c.phi, c.x, c.y, c.vx, c.vy, c.K});
}

void batch_launch_ibm(int kind, const BatchVariant& v, const void* packs, const int2* map,
                      int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    const dim3 g(nblocks), b(256);
    switch (kind) {
    case BK_NODE_PREP:
        hipLaunchKernelGGL(kb_node_prep, g, b, 0, s, static_cast<const PrepPack*>(packs), map);
        break;
    case BK_SPREAD_REST:
        if (v.fp32) hipLaunchKernelGGL(kb_spread_rest<true>, g, b, 0, s, static_cast<const RestPack*>(packs), map);
        else hipLaunchKernelGGL(kb_spread_rest<false>, g, b, 0, s, static_cast<const RestPack*>(packs), map);
        break;
    case BK_SPREAD:
        if (v.fp32) hipLaunchKernelGGL(kb_spread<true>, g, b, 0, s, static_cast<const SpreadPack*>(packs), map);
        else hipLaunchKernelGGL(kb_spread<false>, g, b, 0, s, static_cast<const SpreadPack*>(packs), map);
        break;
    case BK_INTERP:
        if (v.fp32) hipLaunchKernelGGL(kb_interp<true>, g, b, 0, s, static_cast<const InterpPack*>(packs), map);
        else hipLaunchKernelGGL(kb_interp<false>, g, b, 0, s, static_cast<const InterpPack*>(packs), map);
        break;
    case BK_MDF_BEGIN:
        hipLaunchKernelGGL(kb_mdf_begin, g, b, 0, s, static_cast<const MdfBeginPack*>(packs), map);
        break;
    case BK_MDF_INTERP:
        if (v.fp32) hipLaunchKernelGGL(kb_mdf_interp<true>, g, b, 0, s, static_cast<const MdfInterpPack*>(packs), map);
        else hipLaunchKernelGGL(kb_mdf_interp<false>, g, b, 0, s, static_cast<const MdfInterpPack*>(packs), map);
        break;
    case BK_MDF_PREP:
        hipLaunchKernelGGL(kb_node_prep, g, b, 0, s, static_cast<const PrepPack*>(packs), map);
        break;
    case BK_MDF_REST:
        if (v.fp32) hipLaunchKernelGGL(kb_spread_rest<true>, g, b, 0, s, static_cast<const RestPack*>(packs), map);
        else hipLaunchKernelGGL(kb_spread_rest<false>, g, b, 0, s, static_cast<const RestPack*>(packs), map);
        break;
    case BK_MDF_SPREAD:
        if (v.fp32) hipLaunchKernelGGL(kb_spread<true>, g, b, 0, s, static_cast<const SpreadPack*>(packs), map);
        else hipLaunchKernelGGL(kb_spread<false>, g, b, 0, s, static_cast<const SpreadPack*>(packs), map);
        break;
    case BK_MDF_FINISH:
        hipLaunchKernelGGL(kb_mdf_finish, g, b, 0, s, static_cast<const MdfFinishPack*>(packs), map);
        break;
    default:
        break;
    }
}

}  // namespace sf
