// Batched ensembles: many independent channels of uneven size and capsule
// count advanced together.  Every kernel of the step is one launch whose grid
// holds the workgroups of all members (a per-launch map sends workgroup b to
// (member, that member's own workgroup index)), and six steps -- the period
// of the buffer rotation -- are captured as one hipGraph.  Each member runs
// exactly the single-context kernel bodies on its own arguments, so its
// results are bit-identical to stepping it alone.
//
// The reference has no batching API; each member corresponds to one
// Simulation stepped by simulation.cpp:470-787.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "sf_internal.h"

namespace sf {

constexpr int BATCH_PERIOD = 6;   // steps after which every member's buffer roles repeat

// Buffer roles of a member between steps.
struct Roles {
    int fcur, gcur, Fcur, Fprev, Fpart, cell_cur;
    bool operator==(const Roles& o) const {
        return fcur == o.fcur && gcur == o.gcur && Fcur == o.Fcur && Fprev == o.Fprev &&
               Fpart == o.Fpart && cell_cur == o.cell_cur;
    }
    bool operator!=(const Roles& o) const { return !(*this == o); }
};

static Roles roles_of(const Ctx& c) { return Roles{c.fcur, c.gcur, c.Fcur, c.Fprev, c.Fpart, c.cell_cur}; }

static void set_roles(Ctx& c, const Roles& r) {
    c.fcur = r.fcur; c.gcur = r.gcur; c.Fcur = r.Fcur; c.Fprev = r.Fprev; c.Fpart = r.Fpart;
    c.cell_cur = r.cell_cur;
}

// A member is at the batched (steady) step when its next collision takes the
// previous step's force and, with a scalar, the previous step's scalar step is
// pending and its concentration is the one its populations hold -- the state
// in which one_step fuses the scalar step into the collide-stream.
static bool steady_member(const Ctx& c) {
    return c.shift_next == SHIFT_PREV && (c.nsp == 0 || (c.adr_pending && !c.c_from_store));
}

// MDF corrections a member applies per step (mdf_corrections in sf_api.hip)
static int member_corrections(const Ctx& c) {
    const int n = c.mdf_tol > 0.0 ? c.mdf_max_corr : c.p.ibm_iterations - 1;
    return c.K > 0 && n > 0 ? n : 0;
}

// The launches of one batched step, in one_step's order: slot k launches the
// kernel kinds[k] over every member's items in that slot.  The membrane and
// cell-sort slots (0 and 2) share one launch.  With multi-direct forcing
// (C corrections at most over the members) the spread is followed by C
// corrections of four launches each, an optional residual probe and the
// statistics fold; members with fewer corrections, or none, take no part in
// the slots beyond theirs.
struct StepLayout {
    int C = 0;        // corrections of the member that applies the most
    bool probe = false;
    std::vector<int> kinds;
    int lbm = 0;      // slot of the collide-stream
    void build(int C_, bool probe_) {
        C = C_;
        probe = probe_;
        kinds = {BK_MEMBRANE, BK_CELLS_FILL, BK_CELLS_SORT, BK_CONTACT};
        if (C > 0) kinds.push_back(BK_MDF_BEGIN);
        kinds.insert(kinds.end(), {BK_NODE_PREP, BK_SPREAD_REST, BK_SPREAD});
        for (int i = 0; i < C; ++i) kinds.insert(kinds.end(), {BK_MDF_INTERP, BK_MDF_PREP, BK_MDF_REST, BK_MDF_SPREAD});
        if (probe) kinds.push_back(BK_MDF_INTERP);
        if (C > 0) kinds.push_back(BK_MDF_FINISH);
        lbm = static_cast<int>(kinds.size());
        kinds.insert(kinds.end(), {BK_LBM, BK_INTERP});
    }
    int slots() const { return static_cast<int>(kinds.size()); }
};

// The items of one steady step of member c in the layout's slots; advances
// c's buffer roles as one_step does (force-buffer rotation, membrane, contact,
// spread and the MDF corrections, collide-stream, interpolation + move).
static void record_step(Ctx& c, const StepLayout& L, int bx, int by, int pair, int flat, BatchItem* it) {
    const int spare = 3 - c.Fcur - c.Fprev;
    c.Fprev = c.Fcur;
    c.Fcur = spare;
    batch_items_membrane(c, it + BK_MEMBRANE);
    batch_items_contact(c, it + BK_CELLS_FILL);
    int k = BK_CONTACT + 1;
    const int ncorr = member_corrections(c);
    if (L.C > 0) {
        if (ncorr > 0) batch_items_mdf_begin(c, it[k]);
        else batch_items_none(BK_MDF_BEGIN, it[k]);
        ++k;
    }
    batch_items_spread(c, it + k, ncorr == 0);
    k += 3;
    for (int i = 1; i <= L.C; ++i, k += 4) {
        if (i <= ncorr) batch_items_mdf_correction(c, i, ncorr, it + k);
        else for (int j = 0; j < 4; ++j) batch_items_none(L.kinds[k + j], it[k + j]);
    }
    if (L.probe) {
        if (ncorr > 0 && c.mdf_probe) batch_items_mdf_interp(c, ncorr + 1, it[k]);
        else batch_items_none(BK_MDF_INTERP, it[k]);
        ++k;
    }
    if (L.C > 0) {
        if (ncorr > 0) batch_items_mdf_finish(c, ncorr, it[k]);
        else batch_items_none(BK_MDF_FINISH, it[k]);
        ++k;
    }
    batch_items_lbm(c, bx, by, pair, flat, ncorr > 0 ? SHIFT_PARTIAL : SHIFT_PREV, it + k);
    batch_items_interp(c, it + k + 1);
    c.shift_next = SHIFT_PREV;
    c.step_count++;
}

static size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

// Workgroup map of one batched launch: members' workgroups back to back, or
// (banded) arranged so that the workgroups dispatched to each of the 8 XCDs
// (b % 8) take one contiguous eighth of that sequence.
static std::vector<int2> make_map(const std::vector<int>& blocks, bool banded) {
    std::vector<int2> flat;
    for (size_t m = 0; m < blocks.size(); ++m)
        for (int b = 0; b < blocks[m]; ++b) flat.push_back(make_int2(static_cast<int>(m), b));
    if (!banded || flat.empty()) return flat;
    const size_t T = flat.size(), per = (T + 7) / 8;
    std::vector<int2> map(8 * per, make_int2(-1, 0));
    for (size_t b = 0; b < map.size(); ++b) {
        const size_t t = (b % 8) * per + b / 8;
        if (t < T) map[b] = flat[t];
    }
    return map;
}

__global__ void kb_member_flags(int* const* ctr, int M, int* out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) out[m] = (ctr[m][1] & 1) | ((ctr[m][2] & 3) << 1);
}

__global__ void kb_clear_capped(int* const* ctr, int M) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) ctr[m][0] = 0;
}

static void launch_kind(int k, const BatchVariant& v, const void* packs, const int2* map, int nblocks,
                        hipStream_t s) {
    if (k <= BK_CONTACT) batch_launch_membrane(k, v, packs, map, nblocks, s);
    else if (k == BK_LBM) batch_launch_lbm(v, packs, map, nblocks, s);
    else batch_launch_ibm(k, v, packs, map, nblocks, s);
}

}  // namespace sf

using namespace sf;

struct sf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<Ctx*> mem;
    BatchVariant v;
    int banded = 1;
    bool graph_mode = true;

    // launch plan for one period, from the members' roles at phase 0
    bool planned = false;
    std::vector<int64_t> gens;
    std::vector<Roles> roles[BATCH_PERIOD + 1];   // roles before step s of the period
    int phase = 0;
    unsigned char* dpacks = nullptr;
    StepLayout layout;
    std::vector<size_t> off[BATCH_PERIOD];   // per slot
    std::vector<int2*> dmap;
    std::vector<int> nblk;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;

    // overflow flags of the members
    int** dctr = nullptr;
    int* dflags = nullptr;
    int* hflags = nullptr;
};

static void drop_graph(sf_batch& B) {
    // a launch of the executable graph may still be in flight on the stream
    if (B.gexec) hipStreamSynchronize(B.stream);
    if (B.gexec) hipGraphExecDestroy(B.gexec);
    if (B.graph) hipGraphDestroy(B.graph);
    B.gexec = nullptr;
    B.graph = nullptr;
    B.graph_valid = false;
}

static void drop_plan(sf_batch& B) {
    drop_graph(B);
    if (B.dpacks) hipFree(B.dpacks);
    B.dpacks = nullptr;
    for (int2* m : B.dmap)
        if (m) hipFree(m);
    B.dmap.clear();
    B.nblk.clear();
    B.planned = false;
}

// Members the batched step covers (the checks name the first one that fails).
static int validate(sf_batch& B) {
    const Ctx& c0 = *B.mem[0];
    for (size_t m = 0; m < B.mem.size(); ++m) {
        const Ctx& c = *B.mem[m];
        const std::string who = "batch member " + std::to_string(m) + ": ";
        if (!c.lattice_ready) return fail(SF_E_STATE, who + "populations not set");
        if (c.device != B.device) return fail(SF_E_ARG, who + "on another device");
        if (c.fp32 != c0.fp32) return fail(SF_E_ARG, who + "precision differs from member 0");
        if (c.p.collision_model != c0.p.collision_model)
            return fail(SF_E_ARG, who + "collision model differs from member 0");
        if (!c.periodic_x || c.generic_flags)
            return fail(SF_E_ARG, who + "batches take x-periodic plain channels (no obstacle flags)");
        if (c.nobs > 0 || c.p.use_interpolated_bb)
            return fail(SF_E_ARG, who + "obstacles and interpolated bounce-back are not batched");
        if (member_corrections(c) > 0) {
            if (c.nsp > 0)
                return fail(SF_E_ARG, who + "multi-direct forcing with a scalar is not batched");
            if ((c.mdf_warm && c.model) || c.mdf_hist_on || c.ds_buf)
                return fail(SF_E_ARG, who + "the learned warm start and sample collection are not batched");
        }
        if (c.ncaps == 0 || c.any_rigid)
            return fail(SF_E_ARG, who + "batched members carry deformable capsules");
    }
    return SF_OK;
}

static int build_plan(sf_batch& B) {
    drop_plan(B);
    const int M = static_cast<int>(B.mem.size());
    B.v.fp32 = B.mem[0]->fp32 ? 1 : 0;
    B.v.op = B.mem[0]->p.collision_model;
    int maxn = 0;
    for (Ctx* c : B.mem) maxn = std::max(maxn, c->max_nodes);
    B.v.maxn = maxn <= 64 ? 64 : 256;
    B.v.adr = 0;
    for (Ctx* c : B.mem) B.v.adr |= c->nsp > 0 ? 1 : 0;

    int C = 0;
    bool probe = false;
    for (Ctx* c : B.mem) {
        C = std::max(C, member_corrections(*c));
        probe = probe || (member_corrections(*c) > 0 && c->mdf_probe);
    }
    B.layout.build(C, probe);
    const int S = B.layout.slots();
    std::vector<BatchItem> items(static_cast<size_t>(BATCH_PERIOD) * M * S);
    std::vector<int64_t> steps0(M);
    for (int m = 0; m < M; ++m) steps0[m] = B.mem[m]->step_count;
    for (int s = 0; s <= BATCH_PERIOD; ++s) B.roles[s].assign(M, Roles{});
    for (int s = 0; s < BATCH_PERIOD; ++s)
        for (int m = 0; m < M; ++m) {
            B.roles[s][m] = roles_of(*B.mem[m]);
            record_step(*B.mem[m], B.layout, B.v.bx, B.v.by, B.v.pair, B.v.flat,
                        &items[(static_cast<size_t>(s) * M + m) * S]);
        }
    for (int m = 0; m < M; ++m) {
        B.roles[BATCH_PERIOD][m] = roles_of(*B.mem[m]);
        B.mem[m]->step_count = steps0[m];
        if (B.roles[BATCH_PERIOD][m] != B.roles[0][m]) {
            set_roles(*B.mem[m], B.roles[0][m]);
            return fail(SF_E_STATE, "batch: buffer rotation did not close over the period");
        }
    }
    auto item = [&](int s, int m, int k) -> const BatchItem& {
        return items[(static_cast<size_t>(s) * M + m) * S + k];
    };
    // pack layout: [step][slot] -> M packs of one size
    size_t total = 0;
    for (int s = 0; s < BATCH_PERIOD; ++s) B.off[s].assign(S, 0);
    for (int s = 0; s < BATCH_PERIOD; ++s)
        for (int k = 0; k < S; ++k) {
            const size_t sz = item(s, 0, k).pack.size();
            for (int m = 0; m < M; ++m) {
                if (item(s, m, k).pack.size() != sz || item(s, m, k).blocks != item(0, m, k).blocks)
                    return fail(SF_E_STATE, "batch: members disagree on a launch's layout");
            }
            B.off[s][k] = total;
            total += align256(sz * M);
        }
    std::vector<unsigned char> host(total, 0);
    for (int s = 0; s < BATCH_PERIOD; ++s)
        for (int k = 0; k < S; ++k) {
            const size_t sz = item(s, 0, k).pack.size();
            for (int m = 0; m < M; ++m)
                if (sz) std::memcpy(host.data() + B.off[s][k] + m * sz, item(s, m, k).pack.data(), sz);
        }
    B.dmap.assign(S, nullptr);
    B.nblk.assign(S, 0);
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&B.dpacks), std::max<size_t>(total, 256)));
    SF_HIP(hipMemcpy(B.dpacks, host.data(), total, hipMemcpyHostToDevice));
    for (int k = 0; k < S; ++k) {
        if (k == BK_CELLS_SORT) continue;     // its workgroups follow the membrane ones
        std::vector<int> blocks(M);
        for (int m = 0; m < M; ++m) blocks[m] = item(0, m, k).blocks;
        std::vector<int2> map;
        if (k == BK_MEMBRANE) {
            // each member's membrane workgroups, then its sort workgroups: the two
            // kinds are dispatched side by side across the batch
            for (int m = 0; m < M; ++m) {
                for (int b = 0; b < blocks[m]; ++b) map.push_back(make_int2(m, b));
                for (int b = 0; b < item(0, m, BK_CELLS_SORT).blocks; ++b) map.push_back(make_int2(M + m, b));
            }
        } else {
            map = make_map(blocks, k == B.layout.lbm && B.banded);
        }
        B.nblk[k] = static_cast<int>(map.size());
        if (map.empty()) continue;
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&B.dmap[k]), map.size() * sizeof(int2)));
        SF_HIP(hipMemcpy(B.dmap[k], map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    B.gens.resize(M);
    for (int m = 0; m < M; ++m) B.gens[m] = B.mem[m]->gen;
    B.phase = 0;
    B.planned = true;
    return SF_OK;
}

// One step: the contact cell fill, then the membrane forces and the cell sorts
// in one launch (the sort needs only node positions), then the rest in order.
static void launch_step(sf_batch& B, int s) {
    const int M = static_cast<int>(B.mem.size());
    auto run = [&](int k) {
        launch_kind(B.layout.kinds[k], B.v, B.dpacks + B.off[s][k], B.dmap[k], B.nblk[k], B.stream);
    };
    run(BK_CELLS_FILL);
    batch_launch_membrane_sort(B.v, B.dpacks + B.off[s][BK_MEMBRANE], B.dpacks + B.off[s][BK_CELLS_SORT],
                               B.dmap[BK_MEMBRANE], B.nblk[BK_MEMBRANE], M, B.stream);
    for (int k = BK_CONTACT; k < B.layout.slots(); ++k) run(k);
}

// The plan still describes every member: same buffers, parameters and roles.
static bool plan_current(const sf_batch& B) {
    if (!B.planned) return false;
    for (size_t m = 0; m < B.mem.size(); ++m)
        if (B.mem[m]->gen != B.gens[m] || roles_of(*B.mem[m]) != B.roles[B.phase][m]) return false;
    return true;
}

// Work queued on the members' own streams is done before the batch starts.
static int wait_members(sf_batch& B) {
    for (Ctx* c : B.mem) {
        launch_bins_close(*c);      // the batched binning starts from empty tiles
        drop_contact_prefill(*c);   // and the batched cell fill from empty cells
        SF_HIP(hipStreamSynchronize(c->stream));
    }
    return SF_OK;
}

// Later work on the members' own streams waits for the batch.  The batched
// steps leave their touched-tile counts set: a member's own next spread
// clears them first.
static int release_members(sf_batch& B) {
    SF_HIP(hipEventRecord(B.ev, B.stream));
    for (Ctx* c : B.mem) {
        SF_HIP(hipStreamWaitEvent(c->stream, B.ev, 0));
        if (c->tile_cnt) c->bins_open = true;
    }
    return SF_OK;
}

static int check_flags(sf_batch& B) {
    const int M = static_cast<int>(B.mem.size());
    hipLaunchKernelGGL(kb_member_flags, dim3((M + 255) / 256), dim3(256), 0, B.stream, B.dctr, M, B.dflags);
    SF_HIP(hipMemcpyAsync(B.hflags, B.dflags, M * sizeof(int), hipMemcpyDeviceToHost, B.stream));
    SF_HIP(hipStreamSynchronize(B.stream));
    for (int m = 0; m < M; ++m) {
        const int f = B.hflags[m];
        const std::string who = "batch member " + std::to_string(m) + ": ";
        if (f & 1) return fail(SF_E_CAPACITY, who + "IBM tile node list overflowed (more than 1024 nodes near one 32x8 tile)");
        if (f & 2) return fail(SF_E_CAPACITY, who + "contact cell list overflowed (more than 32 nodes in one cell)");
        if (f & 4) return fail(SF_E_CAPACITY, who + "contact cell grid exceeded its capacity");
    }
    return SF_OK;
}

extern "C" {

int sf_batch_create(sf_batch** out, sf_ctx* const* members, int n) {
    if (!out || !members || n < 1) return fail(SF_E_ARG, "batch needs at least one member");
    *out = nullptr;
    for (int m = 0; m < n; ++m)
        for (int j = 0; j < m; ++j)
            if (!members[m] || members[m] == members[j])
                return fail(SF_E_ARG, "batch members must be distinct contexts");
    if (!members[0]) return fail(SF_E_ARG, "null member");
    auto B = new sf_batch();
    for (int m = 0; m < n; ++m) B->mem.push_back(&members[m]->c);
    B->device = B->mem[0]->device;
    auto cleanup = [&](int rc) {
        sf_batch_destroy(B);
        return rc;
    };
    if (hipSetDevice(B->device) != hipSuccess) return cleanup(fail(SF_E_HIP, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&B->ev, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(SF_E_HIP, "batch stream"));
    std::vector<int*> ctr(n);
    for (int m = 0; m < n; ++m) ctr[m] = B->mem[m]->counters;
    if (hipMalloc(reinterpret_cast<void**>(&B->dctr), n * sizeof(int*)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&B->dflags), n * sizeof(int)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&B->hflags), n * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMemcpy(B->dctr, ctr.data(), n * sizeof(int*), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(SF_E_HIP, "batch buffers"));
    *out = B;
    return SF_OK;
}

int sf_batch_destroy(sf_batch* b) {
    if (!b) return SF_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_plan(*b);
    if (b->dctr) hipFree(b->dctr);
    if (b->dflags) hipFree(b->dflags);
    if (b->hflags) hipHostFree(b->hflags);
    if (b->ev) hipEventDestroy(b->ev);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return SF_OK;
}

int sf_batch_step(sf_batch* b, int nsteps) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    if (nsteps < 0) return fail(SF_E_ARG, "negative step count");
    sf_batch& B = *b;
    int rc = validate(B);
    if (rc != SF_OK) return rc;
    SF_HIP(hipSetDevice(B.device));
    if (nsteps == 0) return SF_OK;
    // the batched step is the steady one (previous-step shift force); a
    // member's first step after initialisation runs on its own first
    // (members with a scalar may need a second such step after a host write)
    for (;;) {
        bool steady = true;
        for (Ctx* c : B.mem) steady = steady && steady_member(*c);
        if (steady || nsteps == 0) break;
        for (Ctx* c : B.mem) ctx_one_step(*c);
        --nsteps;
    }
    if (nsteps == 0) {
        SF_HIP(hipGetLastError());
        return SF_OK;
    }
    rc = wait_members(B);
    if (rc != SF_OK) return rc;
    if (!plan_current(B)) {
        rc = build_plan(B);
        if (rc != SF_OK) return rc;
    }
    while (nsteps > 0) {
        if (B.graph_mode && B.phase == 0 && nsteps >= BATCH_PERIOD) {
            if (!B.graph_valid) {
                SF_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < BATCH_PERIOD; ++s) launch_step(B, s);
                SF_HIP(hipStreamEndCapture(B.stream, &B.graph));
                SF_HIP(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
                B.graph_valid = true;
            }
            SF_HIP(hipGraphLaunch(B.gexec, B.stream));
            for (Ctx* c : B.mem) c->step_count += BATCH_PERIOD;
            nsteps -= BATCH_PERIOD;
            continue;
        }
        launch_step(B, B.phase);
        const int next = B.phase + 1;
        for (size_t m = 0; m < B.mem.size(); ++m) {
            set_roles(*B.mem[m], B.roles[next][m]);
            B.mem[m]->step_count++;
        }
        B.phase = next % BATCH_PERIOD;
        --nsteps;
    }
    SF_HIP(hipGetLastError());
    rc = release_members(B);
    if (rc != SF_OK) return rc;
    return check_flags(B);
}

int sf_batch_synchronize(sf_batch* b) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    SF_HIP(hipStreamSynchronize(b->stream));
    return SF_OK;
}

int sf_batch_set_graph_mode(sf_batch* b, int on) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    b->graph_mode = on != 0;
    if (!on) drop_graph(*b);
    return SF_OK;
}

int sf_batch_set_lbm_launch(sf_batch* b, int bx, int by, int xcd_banded) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    if (bx < 1 || by < 1 || bx * by > 256 || (bx * by) % 64 != 0)
        return fail(SF_E_ARG, "launch shape must hold 64..256 threads in whole wavefronts");
    b->v.bx = bx;
    b->v.by = by;
    b->banded = xcd_banded ? 1 : 0;
    b->planned = false;
    return SF_OK;
}

int sf_batch_set_lbm_flat(sf_batch* b, int on) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    b->v.flat = on ? 1 : 0;
    b->planned = false;
    return SF_OK;
}

int sf_batch_autotune(sf_batch* b, int reps, char* report, int report_len) {
    if (!b) return fail(SF_E_ARG, "null pointer");
    sf_batch& B = *b;
    int rc = validate(B);
    if (rc != SF_OK) return rc;
    SF_HIP(hipSetDevice(B.device));
    if (reps < 1) reps = 10;
    rc = wait_members(B);
    if (rc != SF_OK) return rc;
    const int M = static_cast<int>(B.mem.size());
    static const int shapes[][2] = {{64, 4}, {64, 2}, {64, 1}, {128, 2}, {128, 1},
                                    {256, 1}, {32, 8}, {32, 4}};
    hipEvent_t e0, e1;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    std::string rep;
    double best = 1e30;
    int bbx = B.v.bx, bby = B.v.by, bband = B.banded, bpair = B.v.pair, bflat = B.v.flat;
    BatchVariant v = B.v;
    v.fp32 = B.mem[0]->fp32 ? 1 : 0;
    v.op = B.mem[0]->p.collision_model;
    v.adr = 0;
    for (Ctx* c : B.mem) v.adr |= c->nsp > 0 ? 1 : 0;
    for (const auto& sh : shapes)
    for (int flat = 0; flat <= 1; ++flat)
    for (int pair = 0; pair <= 1; ++pair) {
        // flat runs depend only on the thread count: the one-row shapes cover them
        if (flat && sh[1] != 1) continue;
        // the collide-stream of the members' current roles (the launch only
        // reads the source populations, so repeats leave the state as it is)
        std::vector<BatchItem> it(M);
        std::vector<int> blocks(M);
        for (int m = 0; m < M; ++m) {
            Ctx& c = *B.mem[m];
            const int f0 = c.fcur, g0 = c.gcur;
            const bool store0 = c.c_from_store, pend0 = c.adr_pending;
            batch_items_lbm(c, sh[0], sh[1], pair, flat, member_corrections(c) > 0 ? SHIFT_PARTIAL : SHIFT_PREV,
                            &it[m]);
            c.fcur = f0;
            c.gcur = g0;
            c.c_from_store = store0;
            c.adr_pending = pend0;
            blocks[m] = it[m].blocks;
        }
        const size_t sz = it[0].pack.size();
        std::vector<unsigned char> host(sz * M);
        for (int m = 0; m < M; ++m) std::memcpy(host.data() + m * sz, it[m].pack.data(), sz);
        unsigned char* dp = nullptr;
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&dp), host.size()));
        SF_HIP(hipMemcpy(dp, host.data(), host.size(), hipMemcpyHostToDevice));
        v.bx = sh[0];
        v.by = sh[1];
        for (int band = 0; band <= 1; ++band) {
            const std::vector<int2> map = make_map(blocks, band != 0);
            int2* dm = nullptr;
            SF_HIP(hipMalloc(reinterpret_cast<void**>(&dm), map.size() * sizeof(int2)));
            SF_HIP(hipMemcpy(dm, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
            const int nb = static_cast<int>(map.size());
            batch_launch_lbm(v, dp, dm, nb, B.stream);   // warm
            SF_HIP(hipEventRecord(e0, B.stream));
            for (int r = 0; r < reps; ++r) batch_launch_lbm(v, dp, dm, nb, B.stream);
            SF_HIP(hipEventRecord(e1, B.stream));
            SF_HIP(hipEventSynchronize(e1));
            SF_HIP(hipGetLastError());
            float ms = 0.f;
            SF_HIP(hipEventElapsedTime(&ms, e0, e1));
            SF_HIP(hipFree(dm));
            const double per = ms / reps;
            char line[112];
            std::snprintf(line, sizeof(line), "%dx%d %s%s%s %.4f ms\n", sh[0], sh[1],
                          band ? "xcd-banded" : "linear", pair ? " 2 cells/lane" : "",
                          flat ? " flat" : "", per);
            rep += line;
            if (per < best) { best = per; bbx = sh[0]; bby = sh[1]; bband = band; bpair = pair; bflat = flat; }
        }
        SF_HIP(hipFree(dp));
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // the repeated launches added to the capped-cell counts
    hipLaunchKernelGGL(kb_clear_capped, dim3((M + 255) / 256), dim3(256), 0, B.stream, B.dctr, M);
    SF_HIP(hipGetLastError());
    rc = release_members(B);
    if (rc != SF_OK) return rc;
    SF_HIP(hipStreamSynchronize(B.stream));
    B.v.bx = bbx;
    B.v.pair = bpair;
    B.v.by = bby;
    B.v.flat = bflat;
    B.banded = bband;
    B.planned = false;
    char line[112];
    std::snprintf(line, sizeof(line), "chosen %dx%d %s%s%s\n", bbx, bby, bband ? "xcd-banded" : "linear",
                  bpair ? " 2 cells/lane" : "", bflat ? " flat" : "");
    rep += line;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return SF_OK;
}

int sf_batch_info(sf_batch* b, int64_t* out4) {
    if (!b || !out4) return fail(SF_E_ARG, "null pointer");
    int64_t cells = 0, nodes = 0, wg = 0;
    for (Ctx* c : b->mem) {
        cells += c->N;
        nodes += c->K;
        wg += static_cast<int64_t>((c->nx + b->v.bx - 1) / b->v.bx) * ((c->ny + b->v.by - 1) / b->v.by);
    }
    out4[0] = static_cast<int64_t>(b->mem.size());
    out4[1] = cells;
    out4[2] = nodes;
    out4[3] = wg;
    return SF_OK;
}

}  // extern "C"
