// Internal declarations shared by the HIP translation units of libsoftflow_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/softflow.h"
#include "sf_geometry.h"

namespace sf {

// D2Q9 direction tables in the reference's order (lbm/lattice.h:9-31).
__host__ __device__ constexpr int dcx(int q) {
    return q == 1 || q == 5 || q == 8 ? 1 : (q == 3 || q == 6 || q == 7 ? -1 : 0);
}
__host__ __device__ constexpr int dcy(int q) {
    return q == 2 || q == 5 || q == 6 ? 1 : (q == 4 || q == 7 || q == 8 ? -1 : 0);
}
__host__ __device__ constexpr int dopp(int q) {
    return q == 0 ? 0 : (q <= 4 ? ((q + 1) % 4) + 1 : ((q - 3) % 4) + 5);
}
__host__ __device__ constexpr double dw(int q) {
    return q == 0 ? 4.0 / 9.0 : (q <= 4 ? 1.0 / 9.0 : 1.0 / 36.0);
}

// Tile of lattice cells owned by one workgroup of the spread gather.
constexpr int TILE_X = 32;
constexpr int TILE_Y = 8;
constexpr int TILE_MAXL = 1024;    // node references per tile
constexpr int TILE_RECS = 64;      // node records copied into the tile (the spread's short lists)
constexpr int CELL_MAXC = 32;      // nodes per contact cell

// Shift-force source of the collision velocity (see Ctx::shift_next).
enum Shift : int { SHIFT_ZERO = 0, SHIFT_PREV = 1, SHIFT_PARTIAL = 2, SHIFT_STORED = 3 };

// Per-node contribution to a spread: position, force, arc length, the
// stencil origin and (pad) the node index.  One record per node,
// capsule-major.
struct NodeSpread {
    double px, py, fx, fy, ds;
    int i0, j0, valid, pad;
};

__device__ __forceinline__ bool finite_position(double px, double py) {
    return !(isnan(px) || isnan(py) || isinf(px) || isinf(py));
}

// Node k's spread record: position, force, arc length, stencil origin; valid
// when the position and force are finite and the stencil rows can reach the
// lattice.
__device__ __forceinline__ NodeSpread node_record(double px, double py, double fx, double fy, double ds, int k,
                                                  int ny) {
    NodeSpread s;
    s.px = px;
    s.py = py;
    s.fx = fx;
    s.fy = fy;
    s.ds = ds;
    const bool pos_ok = finite_position(px, py);
    int i0 = 0, j0 = 0;
    bool ok = pos_ok && !(isnan(fx) || isnan(fy));
    if (pos_ok) {
        i0 = static_cast<int>(floor(px));
        j0 = static_cast<int>(floor(py));
        if (j0 < -2 || j0 >= ny + 2) ok = false;
    }
    s.i0 = i0;
    s.j0 = j0;
    s.valid = ok ? 1 : 0;
    s.pad = k;
    return s;
}

// Slot claims aggregated over a wavefront (1-D workgroups of whole
// wavefronts; every lane of the wavefront takes part).  Each lane with a key
// >= 0 claims one slot of the counter cnt[key].  A run of consecutive lanes
// with the same key shares one atomic: its first lane adds the run's length,
// and the run's lanes take consecutive slots in lane order.  Consecutive nodes
// of a membrane mostly fall into the same bins, so a wavefront issues a few
// atomics instead of one per lane.  leader/base: the lane's run leader and
// the value its atomic returned (0 for lanes without a key); the slot is
// base + lane - leader, read with wave_slot once the atomics are issued.
__device__ __forceinline__ void wave_claim(int* cnt, int key, int& leader, int& base) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const bool start = lane == 0 || key != prev;
    const unsigned long long starts = __ballot(start);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0 .. lane
    leader = 63 - __clzll(static_cast<long long>(starts & upto));
    const unsigned long long after = starts & ~upto;
    const int next = after ? __ffsll(static_cast<long long>(after)) - 1 : 64;
    base = 0;
    if (start && key >= 0) base = atomicAdd(&cnt[key], next - lane);
}

__device__ __forceinline__ int wave_slot(int leader, int base) {
    return __shfl(base, leader) + (static_cast<int>(threadIdx.x & 63) - leader);
}

// The contact cell of a position: cells of cdx x cdy from (xorg, ybot),
// x wrapped over Lx (> 0) and both indices clamped into the grid; a
// non-finite position goes to cell 0.
__device__ __forceinline__ int contact_cell(double px0, double py, double xorg, double Lx, double cdx,
                                            double cdy, double ybot, int ncx, int ncy) {
    if (isnan(px0) || isnan(py) || isinf(px0) || isinf(py)) return 0;
    double px = px0 - xorg;
    if (Lx > 0.0) {
        px = fmod(px, Lx);
        if (px < 0.0) px += Lx;
    }
    int ix = static_cast<int>(floor(px / cdx));
    int iy = static_cast<int>(floor((py - ybot) / cdy));
    if (ix < 0) ix = 0;
    if (ix >= ncx) ix = ncx - 1;
    if (iy < 0) iy = 0;
    if (iy >= ncy) iy = ncy - 1;
    return iy * ncx + ix;
}

// Listing nodes in the contact cells: node k into cell ci (< 0: none).  Every
// lane of the wavefront calls it (wave_claim); a cell's list is in claim order
// until the cell sort puts it in node order.
__device__ __forceinline__ void contact_cell_claim(int* cnt, int* list, int* counters, int ci, int k) {
    int lead, base;
    wave_claim(cnt, ci, lead, base);
    const int slot = wave_slot(lead, base);
    if (ci < 0) return;
    if (slot < CELL_MAXC) list[static_cast<size_t>(ci) * CELL_MAXC + slot] = k;
    else atomicOr(&counters[2], 1);
}

// The contact cell list's fill done by the step's closing interpolation, at
// the positions it moves the nodes to (cnt null: none).
struct CellFill {
    int* cnt;
    int* list;
    int* counters;
    double xorg, Lx, cdx, cdy, ybot;
    int ncx, ncy;
};

// Distance between two points with the minimum image over Lx (> 0).
__device__ __forceinline__ double min_image_ds(double ax, double ay, double bx, double by, double Lx) {
    double dx = ax - bx, dy = ay - by;   // immersed_boundary.cpp:16-23
    if (Lx > 0.0) {
        if (dx > 0.5 * Lx) dx -= Lx;
        if (dx < -0.5 * Lx) dx += Lx;
    }
    return sqrt(dx * dx + dy * dy);
}

// Multi-direct forcing control.  Fixed mode (tol <= 0) is the reference's
// ibm_iterations - 1 corrections; residual mode stops as soon as the no-slip
// residual max_k |V_k - U(X_k)| is at most tol, after at most max_corr
// corrections.  The first correction may come from a learned model.
constexpr int MDF_MAX_CORR = 32;
// model inputs: slip (2), node force (2), curvature (1), distance to the
// nearest other capsule (1), distance to the nearest wall (1), 5 x 2 slip history
constexpr int MLP_IN = 17;
constexpr int MLP_OUT = 2;
constexpr int MLP_HIST = 5;
constexpr int MDF_HEAD = 2;           // plain corrections the model's output stands for
constexpr double MDF_GAIN_MAX = 8.0;  // |gain| bound of the targets; predictions are held to [0, 4] x [-2, 2]

// Per-node inputs of the warm-start model, in FP32: slip V - U at the first
// correction, the node's membrane + contact force, the signed discrete
// curvature (turning angle / arc length), the distances to the nearest node of
// another capsule and to the nearest wall plane (both capped at 2 r_cut: the
// Peskin stencils of close neighbours overlap and a uniform correction then
// moves the node's velocity more), and the slips of the last MLP_HIST steps.
// The history holds measured slips, never the model's own outputs, so a
// prediction cannot feed back into later inputs.
struct FeatArgs {
    const double *x, *y, *fx, *fy;
    const int *off, *owner;
    const double* u0;       // slip at the first correction, 2 per node
    const float* hist;      // MLP_HIST x K x 2 slips of the previous steps
    const float* dmin;      // distance to the nearest node of another capsule
    int K;
    double Lx;              // > 0: x periodic over Lx
    double ybot, ytop, dcap;
};

__device__ __forceinline__ void mdf_node_features(const FeatArgs& a, int k, float* f) {
    const int cap = a.owner[k];
    const int o = a.off[cap], n = a.off[cap + 1] - o, i = k - o;
    const int kp = o + (i - 1 + n) % n, kn = o + (i + 1) % n;
    double ex0 = a.x[k] - a.x[kp], ey0 = a.y[k] - a.y[kp];
    double ex1 = a.x[kn] - a.x[k], ey1 = a.y[kn] - a.y[k];
    if (a.Lx > 0.0) {
        ex0 -= a.Lx * rint(ex0 / a.Lx);
        ex1 -= a.Lx * rint(ex1 / a.Lx);
    }
    const double th = atan2(ex0 * ey1 - ey0 * ex1, ex0 * ex1 + ey0 * ey1);
    const double ds = 0.5 * (sqrt(ex0 * ex0 + ey0 * ey0) + sqrt(ex1 * ex1 + ey1 * ey1));
    f[0] = static_cast<float>(a.u0[2 * k]);
    f[1] = static_cast<float>(a.u0[2 * k + 1]);
    f[2] = static_cast<float>(a.fx[k]);
    f[3] = static_cast<float>(a.fy[k]);
    f[4] = static_cast<float>(ds > 0.0 ? th / ds : 0.0);
    f[5] = fminf(a.dmin[k], static_cast<float>(a.dcap));
    f[6] = static_cast<float>(fmin(fmax(fmin(a.y[k] - a.ybot, a.ytop - a.y[k]), 0.0), a.dcap));
    for (int h = 0; h < MLP_HIST; ++h) {
        f[7 + 2 * h] = a.hist[(static_cast<size_t>(h) * a.K + k) * 2];
        f[8 + 2 * h] = a.hist[(static_cast<size_t>(h) * a.K + k) * 2 + 1];
    }
}

// The residual maximum of an iteration is kept in MDF_RES_SUB sub-slots on
// their own cache lines (workgroup b updates sub-slot b mod MDF_RES_SUB), so
// the interpolation's workgroups do not all queue on one word; mdf_res folds
// them.
constexpr int MDF_RES_SUB = 8;
constexpr int MDF_RES_LD = 48;   // iterations per sub-slot row (>= MDF_MAX_CORR + 2, 128-byte rows)
struct MdfDev {
    int conv;                    // the residual reached tol this step (set by the statistics fold)
    int applied;                 // corrections spread this step
    int conv_it;                 // iteration at which it converged (0: not)
    int pad;
    // the decision taken after iteration it (mdf_decide): the iteration at
    // which the loop stopped, 0 while it runs on.  Slot it is written while
    // iteration it is decided and read by the launches of later iterations,
    // so no launch reads a slot it writes.
    int cby[MDF_RES_LD];
    // max residual (double bits) per iteration and sub-slot
    alignas(128) unsigned long long res[MDF_RES_SUB][MDF_RES_LD];
    double sum_applied;          // accumulated over steps since the last read
    double sum_final, max_final;
    double steps;
    int64_t ds_rows;             // rows written to the sample set
};

// The residual maximum of iteration it (non-negative doubles order like their
// bit patterns, NaN above all).
__device__ __forceinline__ double mdf_res(const MdfDev* md, int it) {
    unsigned long long m = 0ull;
#pragma unroll
    for (int s = 0; s < MDF_RES_SUB; ++s) m = md->res[s][it] > m ? md->res[s][it] : m;
    return __longlong_as_double(static_cast<long long>(m));
}

// The decision after iteration it (1-based), as every launch of that
// iteration takes it: the iteration at which the loop stopped -- carried
// forward from it - 1, or it itself when its residual is within tolerance
// (residual mode) -- or 0 when correction it is applied.  Each launch computes
// it from slot it - 1 and the residuals of iteration it, which earlier
// launches completed, so launches of one iteration agree without waiting on
// each other.
__device__ __forceinline__ int mdf_decide(const MdfDev* md, int it, double tol) {
    const int prev = md->cby[it - 1];
    if (prev) return prev;
    return (tol > 0.0 && mdf_res(md, it) <= tol) ? it : 0;
}

constexpr int MLP_W = 64;           // hidden width (two hidden layers)
constexpr int MLP_P = MLP_W * MLP_IN + MLP_W + MLP_W * MLP_W + MLP_W + MLP_OUT * MLP_W + MLP_OUT;
constexpr int MLP_NORM = 2 * MLP_IN + 2 * MLP_OUT;   // in_mean, in_scale, out_mean, out_scale
constexpr int MLP_PACKED = 160 * 64;                 // lane-major copy of the weights for inference

// The warm-start network: x -> GELU(W1 x^ + b1) -> GELU(W2 . + b2) -> W3 . + b3,
// with x^ = (x - in_mean) / in_scale and y = out_mean + out_scale * out.
// Parameters are one FP32 block: W1[W][IN], b1[W], W2[W][W], b2[W], W3[OUT][W], b3[OUT].
struct Mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;             // MLP_P
    float* norm = nullptr;               // MLP_NORM
    float* packed = nullptr;             // MLP_PACKED: params + norm in inference order, kept current
    float* adam_m = nullptr;             // MLP_P
    float* adam_v = nullptr;             // MLP_P
    int64_t adam_t = 0;
    bool norm_set = false;
    float* partial = nullptr;            // per-block gradient partials
    int partial_blocks = 0;
    double* acc = nullptr;               // device loss accumulator
};

struct Timer {
    const char* name;
    hipEvent_t a, b;
    double total_ms;
    int launches;
};

// Completion of output snapshots, signalled by a host function the stream runs
// after a snapshot's copies, so that a thread waiting for one makes no HIP call
// (the stepping thread may be capturing a graph on the same stream meanwhile).
struct SnapSync {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int64_t> issued, done;   // tickets enqueued and not yet waited for / completed
};

struct Ctx {
    sf_params p{};
    int device = 0;
    int nx = 0, ny = 0, N = 0;
    bool periodic_x = false, periodic_y = false;
    bool generic_flags = false;          // flags differ from the plain channel
    bool walls_y = true, walls_x = false;
    bool inlet_outlet = false;           // INLET / OUTLET cells present: Zou-He after each collide-stream
    bool fp32 = false;
    hipStream_t stream = nullptr;

    // lattice
    uint8_t* flags = nullptr;            // N
    uint16_t* nbm = nullptr;             // N: solid-neighbour bits and own type (generic flags)
    void* f[2] = {nullptr, nullptr};     // 9*N populations (double, or float f-w)
    int fcur = 0;
    void* F[3] = {nullptr, nullptr, nullptr};   // raw spread force, (x,y) pairs
    int Fcur = 0, Fprev = 1, Fpart = 2;
    double* stored_rho = nullptr;        // macro override for one collision
    double* stored_u = nullptr;          // 2N
    int shift_next = SHIFT_ZERO;
    int pending_shift = SHIFT_ZERO;      // shift chosen by the IBM stage (sf_stage)

    // scalar
    int nsp = 0;
    void* g[4][2] = {};                  // 9*N, SoA planes
    int gcur = 0;
    double* Cstore[4] = {};
    bool c_from_store = true;
    // the scalar step of the last LBM step has not run yet: it rides along with
    // the next LBM launch, or is flushed before anyone reads the scalar
    bool adr_pending = false;

    // capsules
    int ncaps = 0, K = 0, max_nodes = 0;
    bool any_rigid = false;
    int* off = nullptr;
    int* owner = nullptr;
    int2* ring = nullptr;                // per node: (previous, next) node of its ring
    // per node: the spread's arc length element at the current positions,
    // written by the membrane kernel (valid while arc_fresh)
    double* arc = nullptr;
    bool arc_fresh = false;
    sf_capsule_desc* desc = nullptr;
    int* has_prev = nullptr;
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr;
    double *fx = nullptr, *fy = nullptr;
    double *L0 = nullptr, *th0 = nullptr, *ds0 = nullptr, *Lprev = nullptr;
    NodeSpread* spread = nullptr;        // K

    // IBM tiles
    double* phi = nullptr;               // 1025-entry table
    int ntx = 0, nty = 0, ntiles = 0;
    int* tile_cnt = nullptr;             // zero between steps (the last spread resets it)
    int* tile_list = nullptr;
    NodeSpread* tile_rec = nullptr;      // ntiles x TILE_RECS record copies
    int2* tile_act = nullptr;            // (tile, node count) of the tiles touched this step (counters[3])
    uint8_t* tile_dirty = nullptr;       // [3][ntiles]: tile of force buffer b holds non-zeros

    // obstacles (contact forces) and interpolated bounce-back links
    std::vector<uint8_t> h_flags;        // host copy of the cell types
    std::vector<Obstacle> h_obs;         // polygon vertices: offsets into h_obs_v until uploaded
    std::vector<double> h_obs_v;
    int nobs = 0;
    Obstacle* obs = nullptr;             // device copies, vertices pointing into obs_v
    double* obs_v = nullptr;
    bool ibb_dirty = false;              // links must be rebuilt before the next step
    int ibb_groups = 0;                  // cells owning links, in the reference's (y, x) order
    int ibb_par = 0;                     // the first ibb_par groups touch no other group's slots
    int* ibb_cell = nullptr;             // cell of each group
    int* ibb_off = nullptr;              // groups' link ranges (ibb_groups + 1)
    int* ibb_q = nullptr;                // link direction into the solid
    int* ibb_up = nullptr;               // upstream cell for q < 1/2, -1: halfway fallback
    double* ibb_qf = nullptr;            // wall distance along the link, in [0, 1]

    // output snapshots: staging on the device, probe cells, tickets
    double* snap_dev = nullptr;
    size_t snap_cap = 0;                 // doubles
    int nprobes = 0;
    int* probe_cells = nullptr;
    SnapSync* snap_sync = nullptr;       // tickets completed by host callbacks on the stream
    int64_t snap_next = 1;

    // repulsion cells
    int ncx = 0, ncy = 0;
    double cdx = 1.0, cdy = 1.0, xorg = 0.0, Lx = 0.0;
    int* cell_cnt = nullptr;             // two count arrays, alternating by step
    int* cell_list = nullptr;
    double4* cell_pay = nullptr;         // (x, y, vx, vy) of each listed node, in list order
    int* cell_own = nullptr;             // capsule of each listed node
    int cell_cur = 0;
    int* cell_built = nullptr;           // count array filled ahead this step (launch_contact_fill)
    int* cell_other = nullptr;           // the other step's count array, cleared by the sort

    // multi-direct forcing control, history and the sample set
    double mdf_tol = 0.0;
    int mdf_max_corr = 0;
    int mdf_probe = 0;
    int mdf_tail_from = 6;               // residual mode: iterations from here on run in one launch (0: never)
    int mdf_tail_grid = -1;              // its workgroups (-1: not sized yet, 0: unavailable)
    unsigned* mdf_bar = nullptr;         // its grid barrier (count, generation)
    bool mdf_warm = false;
    MdfDev* mdf = nullptr;               // device
    double* mdf_total = nullptr;         // 2K: corrections spread this step, per node
    double* mdf_head = nullptr;          // 2K: sum of the first MDF_HEAD corrections this step
    double* mdf_u0 = nullptr;            // 2K: slip V - U at the first correction
    int* warm_grp = nullptr;             // capsule groups of the fused warm start (warm_ngroups + 1)
    int warm_ngroups = 0;
    double* mdf_f1 = nullptr;            // 2K: the plain first correction dF1
    float* mdf_dmin = nullptr;           // K: distance to the nearest other capsule's node
    float* mdf_hist = nullptr;           // MLP_HIST x 2K: slips of the previous steps
    float* mdf_feat = nullptr;           // K x MLP_IN
    int mdf_hist_on = 0;                 // history is being recorded
    // the step's MDF statistics fold and tile-count clear wait for its
    // closing interpolation
    bool mdf_fold_pending = false;
    bool bins_open = false;              // a spread's tile counts / touched-tile list await clearing
    bool cells_prefilled = false;        // the contact cell counts of cell_cur hold the current positions
    int mdf_fold_ncorr = 0;
    float* ds_buf = nullptr;             // sample rows (MLP_IN + MLP_OUT floats)
    int64_t ds_cap = 0;
    Mlp* model = nullptr;                // borrowed from the caller

    // counters: [0] capped cells, [1] tile overflow, [2] cell overflow, [3] touched tiles
    int* counters = nullptr;
    int* host_counters = nullptr;        // pinned (64 ints; the last 32 take the lagged copy)
    int* ctr_host = nullptr;             // host_counters + 32: the copy enqueued after sf_step
    hipEvent_t ctr_ev = nullptr;         // that copy has landed
    bool ctr_pending = false;

    // graph
    bool graph_mode = false;
    // captured step graphs, one per (buffer roles at capture, length): a graph
    // replays only from the roles it was captured with, and the rotation
    // cycles through a few such sets
    struct GraphEntry {
        int key[6];                      // fcur, gcur, Fcur, Fprev, cell_cur, cells_prefilled
        int len;                         // steps in the graph (1..6)
        int out[7];                      // fcur, gcur, Fcur, Fprev, Fpart, cell_cur, cells_prefilled after them
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    int64_t gen = 0;                     // bumped whenever buffers, parameters or launch shapes change
    int64_t graph_captures = 0;          // step graphs captured and instantiated so far

    // timing
    bool timing = false;
    std::vector<Timer> timers;

    int64_t step_count = 0;
    bool lattice_ready = false;

    // collide-stream launch shape (chosen by sf_autotune)
    int lbm_bx = 64, lbm_by = 4, lbm_remap = 1, lbm_persist = 0;
    // two cells per lane where the paired collide-stream applies (lbm_pair_ok):
    // 0 never, 1 channels, 2 also generic cell types
    int lbm_pair = 1;
    int spread_grid = 0;                 // workgroups of the spread gather (0: one per tile; else persistent)
    int interp_lpn = 16;                 // lanes per node of the interpolations (16, 8 or 4)
};

// Error reporting shared by the C-ABI translation units: sets the message
// returned by sf_last_error() and returns code.
int fail(int code, const std::string& msg);

#define SF_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::sf::fail(SF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// -- launchers (defined in the .hip files) --------------------------------
void launch_lbm(Ctx& c, int shift_mode, bool fuse_adr);
// the same launch with an explicit shape and no change of buffer roles
void launch_lbm_cfg(Ctx& c, int shift_mode, bool fuse_adr, int bx, int by, int remap, int persist,
                    int pair);
bool lbm_pair_ok(const Ctx& c);
void launch_adr(Ctx& c, int Fidx);
void launch_macro(Ctx& c, double* rho, double* ux, double* uy, double* Fx, double* Fy);
void launch_macro_cells(Ctx& c, const int* cells, int count, double* out);
void launch_sum(Ctx& c, const double* v, int64_t n, double* ws, double* out);
// workgroups of the field-statistics reductions (fixed: the summation order
// depends on it)
constexpr int STATS_BLOCKS = 256;
void launch_field_stats(Ctx& c, double* dev_out);
void launch_field_stats_fused(Ctx& c, double* part, double* dev_out);
void launch_global_stats(Ctx& c, double rho0, double* dev_out);
// interpolated bounce-back on the populations the last collide-stream wrote
void launch_ibb(Ctx& c);
int build_ibb_links(Ctx& c);
void launch_membrane(Ctx& c);
void launch_repulsion(Ctx& c);
// Where the contact cell list applies (x-periodic, more than the all-pairs
// limit of nodes) a step runs launch_contact_fill, then launch_membrane_sort
// (membrane forces and the cell sort in one launch), then launch_repulsion.
bool contact_uses_cells(const Ctx& c);
void launch_contact_fill(Ctx& c);
CellFill contact_fill_ahead(Ctx& c, bool move);
void drop_contact_prefill(Ctx& c);
void launch_membrane_sort(Ctx& c);
// node binning and the spread of the node forces into F[Fcur] (last: no MDF
// corrections follow; mdf: the binning also clears the step's MDF state)
void launch_spread(Ctx& c, bool mdf = false);
void launch_bins_close(Ctx& c);
void launch_mdf_interp(Ctx& c, int it);
FeatArgs feat_args(const Ctx& c);
void launch_mdf_features(Ctx& c);
// history and sample rows; leaves the statistics fold pending for the step's
// closing interpolation (launch_mdf_fold runs it on its own)
void launch_mdf_record(Ctx& c, int ncorr);
void launch_mdf_fold(Ctx& c);
// MDF correction it of last_it after its interpolation: the decision, the
// correction spread (touched tiles) and the untouched tiles' copies in one
// launch; rotates the force roles as launch_spread mode 1 does
void launch_mdf_apply(Ctx& c, int it, int last_it);
// MDF iterations first .. last (residual mode) in one launch: the tail of the
// correction loop, which leaves at once when the loop has stopped
void launch_mdf_tail(Ctx& c, int first, int last);
int mdf_tail_size(Ctx& c);
bool mdf_tail_ready(const Ctx& c);
void launch_mdf_warm(Ctx& c);
std::vector<int> warm_groups(const int32_t* off, int ncaps);
void launch_interp_move(Ctx& c);
// The step's interpolation at lpn lanes per node without the move, writing
// the velocities into vx/vy (the autotuner's timing run).
void launch_interp_probe(Ctx& c, int lpn, double* vx, double* vy, const CellFill& cf = CellFill{});
void launch_fill_int(Ctx& c, int* p, int n, int v);
void launch_fill_zero(Ctx& c, void* p, size_t bytes);
void probe_math(int op, int n, const double* a, const double* b, double* out);
// one step of a single context, launched on its own stream (simulation.cpp:470-787 order)
void ctx_one_step(Ctx& c);

// -- batched ensembles ------------------------------------------------------
// One batched launch runs the workgroups of every member of an ensemble in a
// single grid: map[b] = (member, workgroup index within that member's own
// launch), member -1 for padding workgroups.  Each member contributes its
// single-context kernel arguments (the "pack") at its current buffer roles.
enum BatchKernel : int {
    BK_MEMBRANE = 0,
    BK_CELLS_FILL,
    BK_CELLS_SORT,
    BK_CONTACT,
    BK_NODE_PREP,
    BK_SPREAD_REST,
    BK_SPREAD,
    BK_LBM,
    BK_INTERP,
    BK_COUNT,
    // multi-direct forcing (a step holds begin, the corrections' four launches
    // each, an optional residual probe and the statistics fold)
    BK_MDF_BEGIN = BK_COUNT,
    BK_MDF_INTERP,
    BK_MDF_PREP,
    BK_MDF_REST,
    BK_MDF_SPREAD,
    BK_MDF_FINISH
};

struct BatchItem {
    int blocks = 0;
    std::vector<unsigned char> pack;
};

// Template choices shared by all members of a batch.
struct BatchVariant {
    int fp32 = 1;
    int op = SF_BGK;
    int maxn = 64;          // membrane LDS capacity (64 or 256 nodes)
    int bx = 64, by = 4;    // collide-stream workgroup shape
    int pair = 1;           // two cells per lane for the members where it applies
    int adr = 0;            // some member carries a scalar: the collide-stream runs its lagged step
    int flat = 0;           // collide-stream workgroups take runs of the row-major cell index
};

template <typename P>
inline void batch_put(BatchItem& it, int blocks, const P& p) {
    it.blocks = blocks;
    it.pack.resize(sizeof(P));
    std::memcpy(it.pack.data(), &p, sizeof(P));
}

// Items of one steady step, in launch order.  Each call advances the member's
// buffer roles exactly as the single-context launcher does.
void batch_items_membrane(Ctx& c, BatchItem* it);          // BK_MEMBRANE
void batch_items_contact(Ctx& c, BatchItem* it);           // BK_CELLS_FILL .. BK_CONTACT
void batch_items_spread(Ctx& c, BatchItem* it, bool last); // BK_NODE_PREP .. BK_SPREAD
void batch_items_mdf_begin(Ctx& c, BatchItem& it);         // BK_MDF_BEGIN
void batch_items_mdf_interp(Ctx& c, int i, BatchItem& it); // BK_MDF_INTERP (the residual probe)
void batch_items_mdf_correction(Ctx& c, int i, int ncorr, BatchItem* it);   // BK_MDF_INTERP .. BK_MDF_SPREAD
void batch_items_mdf_finish(Ctx& c, int ncorr, BatchItem& it);              // BK_MDF_FINISH
void batch_items_none(int kind, BatchItem& it);            // a slot the member takes no part in
size_t batch_pack_size(int kind);
void batch_items_lbm(Ctx& c, int bx, int by, int pair, int flat, int shift, BatchItem* it);   // BK_LBM
void batch_items_interp(Ctx& c, BatchItem* it);            // BK_INTERP
// launch kind over nblocks workgroups of the map
void batch_launch_membrane(int kind, const BatchVariant& v, const void* packs, const int2* map,
                           int nblocks, hipStream_t s);
// membrane forces and cell sorts in one launch: map entries (m, b) with m < M are
// membrane workgroups of member m, (M + m, b) sort workgroups of member m
void batch_launch_membrane_sort(const BatchVariant& v, const void* mem_packs, const void* sort_packs,
                                const int2* map, int nblocks, int M, hipStream_t s);
void batch_launch_ibm(int kind, const BatchVariant& v, const void* packs, const int2* map,
                      int nblocks, hipStream_t s);
void batch_launch_lbm(const BatchVariant& v, const void* packs, const int2* map, int nblocks,
                      hipStream_t s);

}  // namespace sf

struct sf_ctx {
    sf::Ctx c;
};

struct sf_mlp {
    sf::Mlp m;
};
