// C-ABI entry points, device state management and the per-step schedule.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "sf_internal.h"
#include "sf_math.h"

namespace sf {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

__global__ void k_fill_int(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_zero(uint32_t* p, size_t n4) {
    const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < n4) p[i] = 0u;
}

// ---- Allocation-free field statistics ---------------------------------------
// A file-scope device array holds the partial sums and the results of
// sf_field_stats_noalloc: 5 doubles per reduction workgroup, then the 5
// results.  Each device holds its own copy; its address comes from
// hipGetSymbolAddress once per device, with that device current, and only that
// address is handed to kernels and copies (the array's name is a host-side
// stand-in).  Calls on one device share the array: calls that overlap on one
// device race.  Nothing in it needs clearing: every partial and every result is
// written before it is read.
constexpr int NOALLOC_SCRATCH_DOUBLES = 5 * STATS_BLOCKS + 8;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the scratch array's device address on the current device (device id dev)
static int scratch_on_device(int dev, double** out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(SF_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return fail(SF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SF_OK;
}

void launch_fill_int(Ctx& c, int* p, int n, int v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, c.stream, p, n, v);
}
void launch_fill_zero(Ctx& c, void* p, size_t bytes) {
    const size_t n4 = bytes / 4;     // buffers cleared this way hold 4-byte words
    if (n4 == 0) return;
    hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((n4 + 255) / 256)), dim3(256), 0,
                       c.stream, static_cast<uint32_t*>(p), n4);
}

// Scalar concentration from the distributions (computeConcentration,
// advection_diffusion.cpp:212-223), SoA planes in this build.
template <typename T>
__global__ void k_csum(const T* g, int N, double* C) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) s += static_cast<double>(g[static_cast<size_t>(q) * N + n]);
    C[n] = s;
}

// Streaming copy for the achievable-bandwidth reference: 16 bytes per thread,
// nontemporal both ways, as the collide-stream moves its populations.
typedef float copy_v4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_nt(const copy_v4* a, copy_v4* b, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ void k_probe(int op, int n, const double* a, const double* b, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0;
    if (op == 0) r = sqrt(a[i]);
    else if (op == 1) r = a[i] / b[i];
    else if (op == 2) r = atan2(a[i], b[i]);
    else if (op == 3) r = pow_int_cr(a[i], static_cast<int>(b[i]));
    out[i] = r;
}

// ---------------------------------------------------------------------------
// event timing of named stages
struct TimingPool {
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> live;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> freelist;
    std::map<std::string, std::pair<double, long>> acc;
};
static std::map<const Ctx*, TimingPool>& pools() {
    static std::map<const Ctx*, TimingPool> p;
    return p;
}

struct Scope {
    Ctx& c;
    TimingPool::Rec rec{};
    bool on;
    Scope(Ctx& cc, const char* name) : c(cc), on(cc.timing) {
        if (!on) return;
        TimingPool& P = pools()[&c];
        if (P.freelist.empty()) {
            hipEvent_t a, b;
            hipEventCreate(&a);
            hipEventCreate(&b);
            P.freelist.push_back({a, b});
        }
        auto pr = P.freelist.back();
        P.freelist.pop_back();
        rec = {name, pr.first, pr.second};
        hipEventRecord(rec.a, c.stream);
    }
    ~Scope() {
        if (!on) return;
        hipEventRecord(rec.b, c.stream);
        pools()[&c].live.push_back(rec);
    }
};

static void harvest_timing(Ctx& c) {
    auto it = pools().find(&c);
    if (it == pools().end()) return;
    TimingPool& P = it->second;
    for (auto& r : P.live) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& e = P.acc[r.name];
        e.first += ms;
        e.second += 1;
        P.freelist.push_back({r.a, r.b});
    }
    P.live.clear();
}

// ---------------------------------------------------------------------------
// One step in simulation.cpp:470-787 order.
static void step_forces(Ctx& c) {
    if (c.ncaps == 0) return;
    // the membrane kernel writes the arc lengths of the current positions
    c.arc_fresh = c.arc != nullptr;
    if (contact_uses_cells(c)) {
        // the cell sort needs only node positions: it shares a launch with the
        // membrane forces
        { Scope s(c, "contact_fill"); launch_contact_fill(c); }
        { Scope s(c, "membrane+sort"); launch_membrane_sort(c); }
    } else {
        Scope s(c, "membrane");
        launch_membrane(c);
    }
    { Scope s(c, "repulsion"); launch_repulsion(c); }
}

// MDF corrections this step may apply: residual mode allows up to
// mdf_max_corr, fixed mode applies the reference's ibm_iterations - 1.
static int mdf_corrections(const Ctx& c) {
    return c.mdf_tol > 0.0 ? c.mdf_max_corr : c.p.ibm_iterations - 1;
}

static int step_ibm(Ctx& c) {
    int shift;
    const int ncorr = mdf_corrections(c);
    const bool mdf = ncorr > 0 && c.K > 0;
    launch_mdf_fold(c);   // a step staged without its closing interpolation
    { Scope s(c, "spread"); launch_spread(c, mdf); }
    if (ncorr > 0) {
        if (c.K > 0) {
            const bool learned = c.mdf_warm && c.model;
            // residual mode: the iterations from mdf_tail_from on run in one
            // launch, which leaves at once when the loop has already stopped
            const int tail = c.mdf_tol > 0.0 && c.mdf_tail_from >= 2 && ncorr - c.mdf_tail_from >= 1 &&
                                     mdf_tail_ready(c)
                                 ? c.mdf_tail_from
                                 : 0;
            for (int it = 1; it <= ncorr; ++it) {
                if (it == tail) {
                    Scope s(c, "mdf_tail");
                    launch_mdf_tail(c, it, ncorr);
                    break;
                }
                { Scope s(c, "mdf_interp"); launch_mdf_interp(c, it); }
                if (it == 1 && c.mdf_hist_on) {
                    Scope s(c, "mdf_warm_start");
                    if (learned) launch_mdf_warm(c);
                    else launch_mdf_features(c);
                }
                { Scope s(c, "mdf_apply"); launch_mdf_apply(c, it, ncorr); }
            }
            if (c.mdf_probe) { Scope s(c, "mdf_probe"); launch_mdf_interp(c, ncorr + 1); }
            launch_mdf_record(c, ncorr);
            shift = SHIFT_PARTIAL;
        } else {
            shift = SHIFT_ZERO;   // computeMacroscopic with F = 0
        }
    } else {
        shift = c.shift_next;
    }
    return shift;
}

// Run a scalar step that is still pending (it needs this step's populations
// and the collision force of the step that produced them, F[Fcur]).
static void flush_adr(Ctx& c) {
    if (c.nsp > 0 && c.adr_pending) {
        Scope s(c, "adr");
        launch_adr(c, c.Fcur);
        c.adr_pending = false;
    }
}

static void one_step(Ctx& c) {
    // The scalar step n-1 runs inside LBM step n when the collision velocity
    // of step n is the one that scalar step needs: macro(f_n, F_{n-1}), i.e.
    // the shift force is the previous step's collision force.
    const bool fuse_ok = c.nsp > 0 && mdf_corrections(c) <= 0 && c.shift_next == SHIFT_PREV;
    if (!fuse_ok) flush_adr(c);
    // rotate the raw-force buffers: this step's force goes into the spare one
    const int spare = 3 - c.Fcur - c.Fprev;
    c.Fprev = c.Fcur;
    c.Fcur = spare;
    if (c.ncaps == 0) launch_fill_int(c, c.counters, 1, 0);   // else the membrane kernel clears it
    step_forces(c);
    const int shift = step_ibm(c);
    const bool fuse = fuse_ok && c.adr_pending;
    { Scope s(c, fuse ? "lbm+adr" : "lbm"); launch_lbm(c, shift, fuse); }
    if (c.nsp > 0) {
        if (mdf_corrections(c) <= 0) {
            c.adr_pending = true;
        } else {
            Scope s(c, "adr");
            launch_adr(c, c.Fcur);
            c.adr_pending = false;
        }
    }
    if (c.ncaps > 0) { Scope s(c, "interp_move"); launch_interp_move(c); }
    c.shift_next = SHIFT_PREV;
    c.step_count++;
}

void ctx_one_step(Ctx& c) { one_step(c); }

// An executable graph may still be queued or running on the stream (sf_step
// returns once the launches are enqueued).  The runtime keeps per-launch state
// that refers to the executable graph until the launch retires, so it is
// destroyed only after the stream has drained.
static void destroy_graph_entry(Ctx& c, Ctx::GraphEntry& g, bool drained) {
    if (!drained) hipStreamSynchronize(c.stream);
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
}

static void invalidate_graph(Ctx& c) {
    c.gen++;
    if (c.graphs.empty()) return;
    hipStreamSynchronize(c.stream);
    for (Ctx::GraphEntry& g : c.graphs) destroy_graph_entry(c, g, true);
    c.graphs.clear();
}

constexpr size_t MAX_GRAPHS = 24;

// Graph replay: the step's buffer rotation has period <= 6 (two population
// buffers, a permutation of three force buffers), so six steps are captured
// as one graph whose end state maps buffers back onto the start state.  A
// shorter remainder (a run cut at an output step, say) is a graph of its own
// length, which leaves the roles rotated as the same steps run eagerly would.
static int run_graph(Ctx& c, int nsteps) {
    const int period = 6;
    launch_mdf_fold(c);   // never inside a captured step
    launch_bins_close(c);
    while (nsteps > 0) {
        if (c.shift_next != SHIFT_PREV ||
            (c.nsp > 0 && (c.c_from_store || (mdf_corrections(c) <= 0 && !c.adr_pending)))) {
            one_step(c);
            --nsteps;
            continue;
        }
        const int len = std::min(nsteps, period);
        const int f0 = c.fcur, g0 = c.gcur, a0 = c.Fcur, b0 = c.Fprev;
        const int key[6] = {f0, g0, a0, b0, c.cell_cur, c.cells_prefilled ? 1 : 0};
        // a graph bakes in the buffer roles it was captured with: replay the
        // one captured from these roles, or capture it
        Ctx::GraphEntry* hit = nullptr;
        for (Ctx::GraphEntry& g : c.graphs)
            if (g.len == len && std::equal(key, key + 6, g.key)) hit = &g;
        if (!hit) {
            Ctx::GraphEntry g{};
            std::copy(key, key + 6, g.key);
            g.len = len;
            SF_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < len; ++s) one_step(c);
            hipError_t e = hipStreamEndCapture(c.stream, &g.graph);
            c.step_count -= len;   // counted again below
            if (e != hipSuccess) return fail(SF_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            if (len == period && (c.fcur != f0 || c.gcur != g0 || c.Fcur != a0 || c.Fprev != b0)) {
                hipGraphDestroy(g.graph);
                return fail(SF_E_STATE, "step buffer rotation did not close over the graph period");
            }
            const int out[7] = {c.fcur, c.gcur, c.Fcur, c.Fprev, c.Fpart, c.cell_cur, c.cells_prefilled ? 1 : 0};
            std::copy(out, out + 7, g.out);
            e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                hipGraphDestroy(g.graph);
                SF_HIP(e);
            }
            if (c.graphs.size() >= MAX_GRAPHS) {
                destroy_graph_entry(c, c.graphs.front(), false);
                c.graphs.erase(c.graphs.begin());
            }
            c.graphs.push_back(g);
            c.graph_captures++;
            hit = &c.graphs.back();
        }
        SF_HIP(hipGraphLaunch(hit->exec, c.stream));
        c.fcur = hit->out[0]; c.gcur = hit->out[1]; c.Fcur = hit->out[2];
        c.Fprev = hit->out[3]; c.Fpart = hit->out[4]; c.cell_cur = hit->out[5];
        c.cells_prefilled = hit->out[6] != 0;
        c.step_count += len;
        nsteps -= len;
        c.c_from_store = false;
    }
    return SF_OK;
}

static int counters_verdict(Ctx& c) {
    if (c.host_counters[1] & 1) return fail(SF_E_CAPACITY, "IBM tile node list overflowed (more than 1024 nodes near one 32x8 tile)");
    if (c.host_counters[2] & 1) return fail(SF_E_CAPACITY, "contact cell list overflowed (more than 32 nodes in one cell)");
    if (c.host_counters[2] & 2) return fail(SF_E_CAPACITY, "contact cell grid exceeded its capacity");
    if (c.host_counters[2] & 8) return fail(SF_E_HIP, "MDF correction loop: a grid barrier timed out");
    return SF_OK;
}

static int poll_counters(Ctx& c, bool wait) {
    if (!c.ctr_pending) return SF_OK;
    if (wait) {
        SF_HIP(hipEventSynchronize(c.ctr_ev));
    } else {
        const hipError_t q = hipEventQuery(c.ctr_ev);
        if (q == hipErrorNotReady) return SF_OK;
        SF_HIP(q);
    }
    c.ctr_pending = false;
    for (int k = 0; k < 4; ++k) c.host_counters[k] = c.ctr_host[k];
    return counters_verdict(c);
}

// The overflow flags are sticky on the device.  After each sf_step their copy
// is enqueued behind the steps and read once it has landed: at the next call
// that finds it complete, or at a synchronisation (sync = true waits for it).
// Short sf_step calls (stepping between output steps) therefore do not drain
// the stream each time.
static int check_counters(Ctx& c, bool sync) {
    if (!c.ctr_ev) SF_HIP(hipEventCreateWithFlags(&c.ctr_ev, hipEventDisableTiming));
    if (!sync) {
        SF_HIP(hipMemcpyAsync(c.ctr_host, c.counters, 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        SF_HIP(hipEventRecord(c.ctr_ev, c.stream));
        c.ctr_pending = true;
    }
    return poll_counters(c, sync);
}

static double phi_exact_host(double r) {       // coupling/delta_function.h:18-26
    double ar = std::abs(r);
    if (ar <= 1.0) return 0.125 * (3.0 - 2.0 * ar + std::sqrt(1.0 + 4.0 * ar - 4.0 * ar * ar));
    if (ar <= 2.0) return 0.125 * (5.0 - 2.0 * ar - std::sqrt(-7.0 + 12.0 * ar - 4.0 * ar * ar));
    return 0.0;
}

template <typename T>
static int dmalloc(T** p, size_t bytes) {
    SF_HIP(hipMalloc(reinterpret_cast<void**>(p), bytes));
    return SF_OK;
}

// Per-cell mask for the collide-stream with generic flags: bit q when the
// neighbour in direction q, both axes wrapped, is SOLID; the cell's own type
// in bits 9-11.
static int build_nbm(Ctx& c) {
    if (!c.generic_flags) return SF_OK;
    const int nx = c.nx, ny = c.ny;
    std::vector<uint16_t> m(static_cast<size_t>(c.N));
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            unsigned v = static_cast<unsigned>(c.h_flags[static_cast<size_t>(y) * nx + x]) << 9;
            for (int q = 0; q < 9; ++q) {
                const int xw = (x + dcx(q) + nx) % nx, yw = (y + dcy(q) + ny) % ny;
                if (c.h_flags[static_cast<size_t>(yw) * nx + xw] == SF_SOLID) v |= 1u << q;
            }
            m[static_cast<size_t>(y) * nx + x] = static_cast<uint16_t>(v);
        }
    if (!c.nbm) SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.nbm), m.size() * sizeof(uint16_t)));
    SF_HIP(hipMemcpy(c.nbm, m.data(), m.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    return SF_OK;
}

// InterpolatedBounceBack::initialize (lbm/interpolated_bounce_back.cpp:7-53):
// a link for every direction from a non-SOLID cell into an in-bounds SOLID
// cell, its wall distance the smallest over the obstacles of a 20-step
// bisection along the link (1 when no obstacle is met).  Groups (a cell's
// links) that read a slot another group writes -- the upstream slot of a link
// nearer than half a cell, written by the opposite link of the upstream cell
// -- go to the serial tail, in the reference's cell order.
int build_ibb_links(Ctx& c) {
    invalidate_graph(c);
    for (int** p : {&c.ibb_cell, &c.ibb_off, &c.ibb_q, &c.ibb_up})
        if (*p) { hipFree(*p); *p = nullptr; }
    if (c.ibb_qf) { hipFree(c.ibb_qf); c.ibb_qf = nullptr; }
    c.ibb_groups = c.ibb_par = 0;
    const int nx = c.nx, ny = c.ny;
    const std::vector<uint8_t>& fl = c.h_flags;
    std::vector<Obstacle> obs = c.h_obs;
    for (Obstacle& o : obs)
        if (o.kind == OBS_POLYGON) o.v = c.h_obs_v.data() + static_cast<size_t>(o.p[0]);
    struct Link { int q, up; double qf; };
    std::vector<int> gcell, goff{0};
    std::vector<Link> links;
    std::vector<int> group_of(static_cast<size_t>(c.N), -1);
    std::vector<uint16_t> dirs;   // per group: bit q set when the group has a link in direction q
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int n = y * nx + x;
            if (fl[n] == SF_SOLID) continue;
            const size_t first = links.size();
            uint16_t mask = 0;
            for (int q = 1; q < 9; ++q) {
                const int xn = x + dcx(q), yn = y + dcy(q);
                if (xn < 0 || xn >= nx || yn < 0 || yn >= ny) continue;
                if (fl[static_cast<size_t>(yn) * nx + xn] != SF_SOLID) continue;
                const double x0 = static_cast<double>(x), y0 = static_cast<double>(y);
                const double dx = static_cast<double>(dcx(q)), dy = static_cast<double>(dcy(q));
                double min_q = 1.0;
                for (const Obstacle& o : obs) {
                    double lo = 0.0, hi = 1.0;
                    for (int it = 0; it < 20; ++it) {
                        const double mid = 0.5 * (lo + hi);
                        if (obstacle_contains(o, x0 + mid * dx, y0 + mid * dy)) hi = mid;
                        else lo = mid;
                    }
                    if (hi < min_q) min_q = hi;
                }
                int up = -1;
                if (min_q < 0.5) {
                    int xu = x - dcx(q);
                    const int yu = y - dcy(q);
                    if (c.periodic_x) xu = ((xu % nx) + nx) % nx;
                    if (xu >= 0 && xu < nx && yu >= 0 && yu < ny) {
                        up = yu * nx + xu;
                        if (fl[up] == SF_SOLID)
                            return fail(SF_E_ARG, "interpolated bounce-back: the fluid gap at cell (" +
                                                      std::to_string(x) + ", " + std::to_string(y) +
                                                      ") is one cell wide (its upstream cell is SOLID)");
                    }
                }
                links.push_back({q, up, min_q});
                mask |= static_cast<uint16_t>(1u << q);
            }
            if (links.size() > first) {
                group_of[n] = static_cast<int>(gcell.size());
                gcell.push_back(n);
                goff.push_back(static_cast<int>(links.size()));
                dirs.push_back(mask);
            }
        }
    const int G = static_cast<int>(gcell.size());
    if (G == 0) return SF_OK;
    std::vector<uint8_t> coupled(G, 0);
    for (int g = 0; g < G; ++g)
        for (int l = goff[g]; l < goff[g + 1]; ++l) {
            const Link& L = links[l];
            if (L.qf >= 0.5 || L.up < 0) continue;
            const int w = group_of[L.up];
            // the upstream cell's link in the opposite direction writes slot (up, q)
            if (w >= 0 && (dirs[w] >> dopp(L.q) & 1)) coupled[g] = coupled[w] = 1;
        }
    std::vector<int> cell2, off2{0}, q2, up2;
    std::vector<double> qf2;
    int par = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int g = 0; g < G; ++g) {
            if (coupled[g] != pass) continue;
            cell2.push_back(gcell[g]);
            for (int l = goff[g]; l < goff[g + 1]; ++l) {
                q2.push_back(links[l].q);
                up2.push_back(links[l].up);
                qf2.push_back(links[l].qf);
            }
            off2.push_back(static_cast<int>(q2.size()));
            if (pass == 0) ++par;
        }
    const size_t L = q2.size();
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ibb_cell), G * sizeof(int)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ibb_off), (G + 1) * sizeof(int)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ibb_q), L * sizeof(int)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ibb_up), L * sizeof(int)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ibb_qf), L * sizeof(double)));
    SF_HIP(hipMemcpy(c.ibb_cell, cell2.data(), G * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.ibb_off, off2.data(), (G + 1) * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.ibb_q, q2.data(), L * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.ibb_up, up2.data(), L * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.ibb_qf, qf2.data(), L * sizeof(double), hipMemcpyHostToDevice));
    c.ibb_groups = G;
    c.ibb_par = par;
    return SF_OK;
}

static void free_caps(Ctx& c) {
    drop_contact_prefill(c);
    void* ptrs[] = {c.off, c.owner, c.ring, c.arc, c.desc, c.has_prev, c.x, c.y, c.vx, c.vy, c.fx, c.fy,
                    c.L0, c.th0, c.ds0, c.Lprev, c.spread, c.mdf_total, c.mdf_u0, c.mdf_hist,
                    c.mdf_feat, c.mdf_f1, c.mdf_dmin, c.mdf_head, c.warm_grp};
    for (void* p : ptrs) if (p) hipFree(p);
    c.mdf_total = nullptr; c.mdf_u0 = nullptr; c.mdf_hist = nullptr; c.mdf_feat = nullptr;
    c.mdf_f1 = nullptr; c.mdf_dmin = nullptr; c.mdf_head = nullptr;
    c.warm_grp = nullptr;
    c.warm_ngroups = 0;
    c.off = c.owner = c.has_prev = nullptr;
    c.ring = nullptr;
    c.arc = nullptr;
    c.arc_fresh = false;
    c.desc = nullptr;
    c.x = c.y = c.vx = c.vy = c.fx = c.fy = c.L0 = c.th0 = c.ds0 = c.Lprev = nullptr;
    c.spread = nullptr;
    c.ncaps = c.K = 0;
}

}  // namespace sf

using namespace sf;

extern "C" {

const char* sf_last_error(void) { return g_err.c_str(); }

int sf_device_count(int* n) {
    if (!n) return fail(SF_E_ARG, "null pointer");
    SF_HIP(hipGetDeviceCount(n));
    return SF_OK;
}

int sf_create(sf_ctx** out, int device, const sf_params* p) {
    if (!out || !p) return fail(SF_E_ARG, "null pointer");
    *out = nullptr;
    if (p->nx < 3 || p->ny < 3) return fail(SF_E_ARG, "domain must be at least 3x3");
    if (p->boundary_type != SF_PERIODIC && p->boundary_type != SF_CLOSED && p->boundary_type != SF_INLET_OUTLET)
        return fail(SF_E_ARG, "unknown boundary type");
    if (p->boundary_type == SF_INLET_OUTLET && !(p->outlet_density > 0.0))
        return fail(SF_E_ARG, "outlet density must be positive");
    if (p->tau <= 0.5) return fail(SF_E_ARG, "tau must exceed 0.5");
    if (p->collision_model < 0 || p->collision_model > 2) return fail(SF_E_ARG, "unknown collision model");
    if (p->precision != SF_FP64 && p->precision != SF_FP32) return fail(SF_E_ARG, "precision must be SF_FP64 or SF_FP32");
    if (p->scalar_enabled && (p->n_species < 1 || p->n_species > 4)) return fail(SF_E_ARG, "1..4 scalar species supported");
    if (p->ibm_iterations < 1 || p->ibm_iterations > MDF_MAX_CORR + 1)
        return fail(SF_E_ARG, "ibm_iterations must be in 1..33");
    int ndev = 0;
    SF_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SF_E_ARG, "no such HIP device");
    SF_HIP(hipSetDevice(device));
    auto* h = new sf_ctx();
    Ctx& c = h->c;
    c.p = *p;
    c.device = device;
    c.nx = p->nx; c.ny = p->ny; c.N = p->nx * p->ny;
    c.periodic_x = p->boundary_type == SF_PERIODIC;
    c.periodic_y = p->periodic_y != 0;
    c.walls_y = !c.periodic_y;
    c.walls_x = p->boundary_type == SF_CLOSED;
    c.fp32 = p->precision == SF_FP32;
    const size_t N = c.N;
    const size_t es = c.fp32 ? 4 : 8;
    int rc = SF_OK;
    auto chk = [&](int r) { if (r != SF_OK && rc == SF_OK) rc = r; };
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(SF_E_HIP, "stream creation failed");
    }
    chk(dmalloc(&c.flags, N));
    for (int b = 0; b < 2; ++b) chk(dmalloc(&c.f[b], 9 * N * es));
    for (int b = 0; b < 3; ++b) chk(dmalloc(&c.F[b], 2 * N * es));
    chk(dmalloc(&c.stored_rho, N * 8));
    chk(dmalloc(&c.stored_u, 2 * N * 8));
    chk(dmalloc(&c.counters, 64 * sizeof(int)));
    chk(dmalloc(&c.phi, 1025 * 8));
    chk(dmalloc(&c.mdf, sizeof(MdfDev)));
    chk(dmalloc(&c.mdf_bar, 64 * sizeof(unsigned)));
    if (rc != SF_OK) { sf_destroy(h); return rc; }
    SF_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.host_counters), 64 * sizeof(int)));
    {   // the allocation-free entry point's scratch address on this device
        double* sp = nullptr;
        rc = scratch_on_device(c.device, &sp);
        if (rc != SF_OK) { sf_destroy(h); return rc; }
    }
    c.ctr_host = c.host_counters + 32;
    // flags: the channel of LBMSolver::initialize (lbm_solver.cpp:130-158);
    // with inlet/outlet, INLET and OUTLET columns between the walls
    std::vector<uint8_t> fl(N, SF_FLUID);
    for (int y = 0; y < c.ny; ++y)
        for (int x = 0; x < c.nx; ++x)
            if ((c.walls_y && (y == 0 || y == c.ny - 1)) || (c.walls_x && (x == 0 || x == c.nx - 1)))
                fl[static_cast<size_t>(y) * c.nx + x] = SF_SOLID;
    if (p->boundary_type == SF_INLET_OUTLET) {
        for (int y = 1; y < c.ny - 1; ++y) {
            fl[static_cast<size_t>(y) * c.nx] = SF_INLET;
            fl[static_cast<size_t>(y) * c.nx + c.nx - 1] = SF_OUTLET;
        }
        c.inlet_outlet = true;
        c.generic_flags = true;
    }
    if (c.walls_x && c.periodic_y) {
        // periodic y clears both rows after the closed columns are set
        // (lbm_solver.cpp:146-158): the four corners are FLUID
        for (int x : {0, c.nx - 1})
            for (int y : {0, c.ny - 1}) fl[static_cast<size_t>(y) * c.nx + x] = SF_FLUID;
        c.generic_flags = true;
    }
    c.h_flags = fl;
    SF_HIP(hipMemcpy(c.flags, fl.data(), N, hipMemcpyHostToDevice));
    if (build_nbm(c) != SF_OK) { sf_destroy(h); return SF_E_HIP; }
    for (int b = 0; b < 3; ++b) SF_HIP(hipMemset(c.F[b], 0, 2 * N * es));
    SF_HIP(hipMemset(c.counters, 0, 64 * sizeof(int)));
    SF_HIP(hipMemset(c.mdf, 0, sizeof(MdfDev)));
    SF_HIP(hipMemset(c.mdf_bar, 0, 64 * sizeof(unsigned)));
    c.Fcur = 0; c.Fprev = 1; c.Fpart = 2;
    double tab[1025];
    for (int i = 0; i <= 1024; ++i) tab[i] = phi_exact_host(2.0 * static_cast<double>(i) / 1024);
    SF_HIP(hipMemcpy(c.phi, tab, sizeof(tab), hipMemcpyHostToDevice));
    c.ntx = (c.nx + TILE_X - 1) / TILE_X;
    c.nty = (c.ny + TILE_Y - 1) / TILE_Y;
    c.ntiles = c.ntx * c.nty;
    // contact cell grid (cell_list.h:47-70)
    const double rcut = p->rep_r_cut;
    const double ybot = 0.5, ytop = static_cast<double>(c.ny) - 1.5;
    c.ncy = std::max(1, static_cast<int>(std::floor((ytop - ybot) / rcut)));
    c.cdy = (ytop - ybot) / static_cast<double>(c.ncy);
    if (c.periodic_x) {
        c.ncx = std::max(1, static_cast<int>(std::floor(static_cast<double>(c.nx) / rcut)));
        c.cdx = static_cast<double>(c.nx) / static_cast<double>(c.ncx);
    } else {
        c.ncx = static_cast<int>(std::floor((static_cast<double>(c.nx) + 4.0 * rcut) / rcut)) + 2;
        c.cdx = rcut;
    }
    if (p->scalar_enabled) {
        c.nsp = p->n_species;
        for (int s = 0; s < c.nsp; ++s) {
            for (int b = 0; b < 2; ++b) {
                if (dmalloc(&c.g[s][b], 9 * N * es) != SF_OK) { sf_destroy(h); return SF_E_HIP; }
                SF_HIP(hipMemset(c.g[s][b], 0, 9 * N * es));
            }
            if (dmalloc(&c.Cstore[s], N * 8) != SF_OK) { sf_destroy(h); return SF_E_HIP; }
            SF_HIP(hipMemset(c.Cstore[s], 0, N * 8));
        }
        c.c_from_store = true;
    }
    c.shift_next = SHIFT_ZERO;
    if (p->use_interpolated_bb) {
        const int r = build_ibb_links(c);
        if (r != SF_OK) { sf_destroy(h); return r; }
    }
    *out = h;
    return SF_OK;
}

int sf_destroy(sf_ctx* h) {
    if (!h) return SF_OK;
    Ctx& c = h->c;
    hipSetDevice(c.device);
    if (c.stream) hipStreamSynchronize(c.stream);
    invalidate_graph(c);
    free_caps(c);
    void* ptrs[] = {c.flags, c.f[0], c.f[1], c.F[0], c.F[1], c.F[2], c.stored_rho, c.stored_u,
                    c.counters, c.phi, c.tile_cnt, c.tile_list, c.tile_act, c.tile_rec, c.tile_dirty, c.cell_cnt, c.cell_list,
                    c.cell_pay, c.cell_own, c.mdf, c.mdf_bar, c.ds_buf, c.obs, c.obs_v, c.ibb_cell, c.ibb_off,
                    c.ibb_q, c.ibb_up, c.ibb_qf, c.snap_dev, c.probe_cells, c.nbm};
    delete c.snap_sync;     // the stream has drained: no host function is pending
    c.snap_sync = nullptr;
    for (void* p : ptrs) if (p) hipFree(p);
    for (int s = 0; s < 4; ++s) {
        for (int b = 0; b < 2; ++b) if (c.g[s][b]) hipFree(c.g[s][b]);
        if (c.Cstore[s]) hipFree(c.Cstore[s]);
    }
    if (c.host_counters) hipHostFree(c.host_counters);
    if (c.ctr_ev) hipEventDestroy(c.ctr_ev);
    auto it = pools().find(&c);
    if (it != pools().end()) {
        for (auto& pr : it->second.freelist) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
        for (auto& r : it->second.live) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
        pools().erase(it);
    }
    if (c.stream) hipStreamDestroy(c.stream);
    delete h;
    return SF_OK;
}

int sf_set_flags(sf_ctx* h, const int32_t* flags) {
    if (!h || !flags) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    std::vector<uint8_t> fl(c.N);
    bool channel = true, io = false;
    for (int y = 0; y < c.ny; ++y)
        for (int x = 0; x < c.nx; ++x) {
            const size_t n = static_cast<size_t>(y) * c.nx + x;
            const int v = flags[n];
            if (v != SF_FLUID && v != SF_SOLID && v != SF_INLET && v != SF_OUTLET)
                return fail(SF_E_ARG, "device path supports FLUID, SOLID, INLET and OUTLET cells");
            fl[n] = static_cast<uint8_t>(v);
            const bool wall = (c.walls_y && (y == 0 || y == c.ny - 1)) || (c.walls_x && (x == 0 || x == c.nx - 1));
            if ((v == SF_SOLID) != wall || v == SF_INLET || v == SF_OUTLET) channel = false;
            io |= v == SF_INLET || v == SF_OUTLET;
        }
    SF_HIP(hipMemcpy(c.flags, fl.data(), c.N, hipMemcpyHostToDevice));
    c.h_flags = fl;
    c.generic_flags = !channel;
    c.inlet_outlet = io;
    invalidate_graph(c);
    if (build_nbm(c) != SF_OK) return SF_E_HIP;
    if (c.p.use_interpolated_bb) return build_ibb_links(c);
    return SF_OK;
}

int sf_set_obstacles(sf_ctx* h, int count, const int32_t* kinds, const int32_t* offsets,
                     const double* coords) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    if (count < 0) return fail(SF_E_ARG, "negative obstacle count");
    if (count > 0 && (!kinds || !offsets || !coords)) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    std::vector<Obstacle> obs(count);
    std::vector<double> verts;
    for (int i = 0; i < count; ++i) {
        const int b = offsets[i], e = offsets[i + 1], n = e - b;
        if (b < 0 || n < 0 || (i == 0 && b != 0)) return fail(SF_E_ARG, "obstacle offsets must start at 0 and not decrease");
        Obstacle& o = obs[i];
        o = Obstacle{};
        o.kind = kinds[i];
        if (o.kind == OBS_CIRCLE) {
            if (n != 3) return fail(SF_E_ARG, "a circle takes (cx, cy, r)");
            for (int k = 0; k < 3; ++k) o.p[k] = coords[b + k];
        } else if (o.kind == OBS_RECT) {
            if (n != 4) return fail(SF_E_ARG, "a rectangle takes (x0, y0, x1, y1)");
            // rect_obstacle.h:10-12
            o.p[0] = std::min(coords[b], coords[b + 2]);
            o.p[1] = std::min(coords[b + 1], coords[b + 3]);
            o.p[2] = std::max(coords[b], coords[b + 2]);
            o.p[3] = std::max(coords[b + 1], coords[b + 3]);
        } else if (o.kind == OBS_POLYGON) {
            if (n < 6 || n % 2) return fail(SF_E_ARG, "a polygon takes at least 3 (x, y) vertices");
            o.nv = n / 2;
            o.p[0] = static_cast<double>(verts.size());   // offset, resolved below
            verts.insert(verts.end(), coords + b, coords + e);
        } else {
            return fail(SF_E_ARG, "unknown obstacle kind");
        }
    }
    if (c.obs) { hipFree(c.obs); c.obs = nullptr; }
    if (c.obs_v) { hipFree(c.obs_v); c.obs_v = nullptr; }
    c.h_obs = obs;
    c.h_obs_v = verts;
    c.nobs = 0;
    if (count > 0) {
        if (!verts.empty()) {
            SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.obs_v), verts.size() * sizeof(double)));
            SF_HIP(hipMemcpy(c.obs_v, verts.data(), verts.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        std::vector<Obstacle> dev = obs;
        for (Obstacle& o : dev) {
            obstacle_bound(o, o.kind == OBS_POLYGON ? verts.data() + static_cast<size_t>(o.p[0]) : nullptr);
            if (o.kind == OBS_POLYGON) o.v = c.obs_v + static_cast<size_t>(o.p[0]);
        }
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.obs), dev.size() * sizeof(Obstacle)));
        SF_HIP(hipMemcpy(c.obs, dev.data(), dev.size() * sizeof(Obstacle), hipMemcpyHostToDevice));
        c.nobs = count;
    }
    invalidate_graph(c);
    if (c.p.use_interpolated_bb) return build_ibb_links(c);
    return SF_OK;
}

int sf_set_populations(sf_ctx* h, const double* f) {
    if (!h || !f) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    // a scalar step still pending needs the populations it was due with
    flush_adr(c);
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t N = c.N;
    if (!c.fp32) {
        SF_HIP(hipMemcpy(c.f[c.fcur], f, 9 * N * 8, hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(c.f[1 - c.fcur], f, 9 * N * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> d(9 * N);
        for (int q = 0; q < 9; ++q)
            for (size_t n = 0; n < N; ++n) d[q * N + n] = static_cast<float>(f[q * N + n] - dw(q));
        SF_HIP(hipMemcpy(c.f[c.fcur], d.data(), 9 * N * 4, hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(c.f[1 - c.fcur], d.data(), 9 * N * 4, hipMemcpyHostToDevice));
    }
    c.lattice_ready = true;
    return SF_OK;
}

int sf_get_populations(sf_ctx* h, double* f) {
    if (!h || !f) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t N = c.N;
    if (!c.fp32) {
        SF_HIP(hipMemcpy(f, c.f[c.fcur], 9 * N * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> d(9 * N);
        SF_HIP(hipMemcpy(d.data(), c.f[c.fcur], 9 * N * 4, hipMemcpyDeviceToHost));
        for (int q = 0; q < 9; ++q)
            for (size_t n = 0; n < N; ++n) f[q * N + n] = static_cast<double>(d[q * N + n]) + dw(q);
    }
    return SF_OK;
}

int sf_get_macroscopic(sf_ctx* h, double* rho, double* ux, double* uy, double* Fx, double* Fy) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    const size_t N = c.N;
    double* buf = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&buf), 5 * N * 8));
    launch_macro(c, buf, buf + N, buf + 2 * N, buf + 3 * N, buf + 4 * N);
    hipError_t e = hipStreamSynchronize(c.stream);
    double* outs[5] = {rho, ux, uy, Fx, Fy};
    for (int k = 0; k < 5 && e == hipSuccess; ++k)
        if (outs[k]) e = hipMemcpy(outs[k], buf + k * N, N * 8, hipMemcpyDeviceToHost);
    hipFree(buf);
    SF_HIP(e);
    return SF_OK;
}

int sf_set_macroscopic(sf_ctx* h, const double* rho, const double* ux, const double* uy) {
    if (!h || !rho || !ux || !uy) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));   // an enqueued step may still read the stored state
    std::vector<double> u(2 * static_cast<size_t>(c.N));
    for (int n = 0; n < c.N; ++n) { u[2 * n] = ux[n]; u[2 * n + 1] = uy[n]; }
    SF_HIP(hipMemcpy(c.stored_rho, rho, c.N * 8, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.stored_u, u.data(), 2 * c.N * 8, hipMemcpyHostToDevice));
    c.shift_next = SHIFT_STORED;
    return SF_OK;
}

int sf_get_force_state(sf_ctx* h, double* F) {
    if (!h || !F) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t n2 = 2 * static_cast<size_t>(c.N);
    if (c.fp32) {
        std::vector<float> t(n2);
        SF_HIP(hipMemcpy(t.data(), c.F[c.Fcur], n2 * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n2; ++i) F[i] = t[i];
    } else {
        SF_HIP(hipMemcpy(F, c.F[c.Fcur], n2 * 8, hipMemcpyDeviceToHost));
    }
    return SF_OK;
}

int sf_set_force_state(sf_ctx* h, const double* F) {
    if (!h || !F) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t n2 = 2 * static_cast<size_t>(c.N);
    if (c.fp32) {
        std::vector<float> t(n2);
        for (size_t i = 0; i < n2; ++i) t[i] = static_cast<float>(F[i]);
        SF_HIP(hipMemcpy(c.F[c.Fcur], t.data(), n2 * 4, hipMemcpyHostToDevice));
    } else {
        SF_HIP(hipMemcpy(c.F[c.Fcur], F, n2 * 8, hipMemcpyHostToDevice));
    }
    // the buffer may now hold non-zeros anywhere
    if (c.tile_dirty)
        SF_HIP(hipMemset(c.tile_dirty + static_cast<size_t>(c.Fcur) * c.ntiles, 1, c.ntiles));
    c.shift_next = SHIFT_PREV;
    invalidate_graph(c);
    return SF_OK;
}

int sf_set_capsules(sf_ctx* h, int ncaps, const int32_t* off, const sf_capsule_desc* desc,
                    const double* x, const double* y, const double* vx, const double* vy,
                    const double* rest_len, const double* rest_angle, const double* rest_ds,
                    const double* prev_len, const int32_t* has_prev) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    invalidate_graph(c);
    free_caps(c);
    if (ncaps == 0) return SF_OK;
    if (ncaps < 0 || !off || !desc || !x || !y || !vx || !vy || !rest_len || !prev_len || !has_prev)
        return fail(SF_E_ARG, "null capsule array");
    if (off[0] != 0) return fail(SF_E_ARG, "offsets must start at 0");
    int maxn = 0;
    for (int i = 0; i < ncaps; ++i) {
        const int n = off[i + 1] - off[i];
        if (n < 1) return fail(SF_E_ARG, "capsule with no nodes");
        if (n > 256) return fail(SF_E_ARG, "capsules with more than 256 nodes are not supported");
        if (desc[i].use_helfrich && (!rest_angle || !rest_ds)) return fail(SF_E_ARG, "Helfrich capsule without rest angles");
        maxn = std::max(maxn, n);
    }
    c.any_rigid = false;
    for (int i = 0; i < ncaps; ++i) c.any_rigid |= desc[i].is_rigid != 0;
    if (c.nx < 4) return fail(SF_E_ARG, "capsules need nx >= 4");
    const int K = off[ncaps];
    c.ncaps = ncaps; c.K = K; c.max_nodes = maxn;
    std::vector<int> owner(K);
    std::vector<int2> ring(K);
    for (int i = 0; i < ncaps; ++i) {
        const int o = off[i], n = off[i + 1] - o;
        for (int k = o; k < o + n; ++k) {
            owner[k] = i;
            ring[k] = make_int2(o + (k - o - 1 + n) % n, o + (k - o + 1) % n);
        }
    }
    std::vector<double> zeros(K, 0.0);
    int rc = SF_OK;
    auto up = [&](double** d, const double* src) {
        if (rc != SF_OK) return;
        rc = dmalloc(d, K * 8);
        if (rc == SF_OK && hipMemcpy(*d, src ? src : zeros.data(), K * 8, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(SF_E_HIP, "capsule upload failed");
    };
    up(&c.x, x); up(&c.y, y); up(&c.vx, vx); up(&c.vy, vy);
    up(&c.fx, nullptr); up(&c.fy, nullptr);
    up(&c.L0, rest_len); up(&c.th0, rest_angle); up(&c.ds0, rest_ds); up(&c.Lprev, prev_len);
    if (rc != SF_OK) return rc;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.off), (ncaps + 1) * sizeof(int)));
    SF_HIP(hipMemcpy(c.off, off, (ncaps + 1) * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.owner), K * sizeof(int)));
    SF_HIP(hipMemcpy(c.owner, owner.data(), K * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ring), K * sizeof(int2)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.arc), K * sizeof(double)));
    SF_HIP(hipMemcpy(c.ring, ring.data(), K * sizeof(int2), hipMemcpyHostToDevice));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.desc), ncaps * sizeof(sf_capsule_desc)));
    SF_HIP(hipMemcpy(c.desc, desc, ncaps * sizeof(sf_capsule_desc), hipMemcpyHostToDevice));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.has_prev), ncaps * sizeof(int)));
    SF_HIP(hipMemcpy(c.has_prev, has_prev, ncaps * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.spread), K * sizeof(NodeSpread)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_total), 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_u0), 2 * static_cast<size_t>(K) * sizeof(double)));
    {
        const std::vector<int> grp = warm_groups(off, ncaps);
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.warm_grp), grp.size() * sizeof(int)));
        SF_HIP(hipMemcpy(c.warm_grp, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice));
        c.warm_ngroups = static_cast<int>(grp.size()) - 1;
    }
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_f1), 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_dmin), static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_head), 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMemset(c.mdf_head, 0, 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMemset(c.mdf_dmin, 0, static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_hist), MLP_HIST * 2 * static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.mdf_feat), MLP_IN * static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMemset(c.mdf_total, 0, 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMemset(c.mdf_u0, 0, 2 * static_cast<size_t>(K) * sizeof(double)));
    SF_HIP(hipMemset(c.mdf_hist, 0, MLP_HIST * 2 * static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMemset(c.mdf_feat, 0, MLP_IN * static_cast<size_t>(K) * sizeof(float)));
    SF_HIP(hipMemset(c.mdf, 0, sizeof(MdfDev)));
    if (!c.tile_cnt) {
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.tile_cnt), c.ntiles * sizeof(int)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.tile_list), static_cast<size_t>(c.ntiles) * TILE_MAXL * sizeof(int)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.tile_act), static_cast<size_t>(c.ntiles) * sizeof(int2)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.tile_rec), static_cast<size_t>(c.ntiles) * TILE_RECS * sizeof(NodeSpread)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.tile_dirty), 3 * static_cast<size_t>(c.ntiles)));
        SF_HIP(hipMemset(c.tile_cnt, 0, c.ntiles * sizeof(int)));
        // the force buffers may hold anything from before capsules existed
        SF_HIP(hipMemset(c.tile_dirty, 1, 3 * static_cast<size_t>(c.ntiles)));
    }
    if (!c.cell_cnt && K > 100) {
        const size_t ncells = static_cast<size_t>(c.ncx) * c.ncy;
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.cell_cnt), 2 * ncells * sizeof(int)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.cell_list), ncells * CELL_MAXC * sizeof(int)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.cell_pay), ncells * CELL_MAXC * sizeof(double4)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.cell_own), ncells * CELL_MAXC * sizeof(int)));
        SF_HIP(hipMemset(c.cell_cnt, 0, 2 * ncells * sizeof(int)));
        c.cell_cur = 0;
    }
    return SF_OK;
}

int sf_get_nodes(sf_ctx* h, double* x, double* y, double* vx, double* vy, double* fx, double* fy) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t b = static_cast<size_t>(c.K) * 8;
    if (b == 0) return SF_OK;
    if (x) SF_HIP(hipMemcpy(x, c.x, b, hipMemcpyDeviceToHost));
    if (y) SF_HIP(hipMemcpy(y, c.y, b, hipMemcpyDeviceToHost));
    if (vx) SF_HIP(hipMemcpy(vx, c.vx, b, hipMemcpyDeviceToHost));
    if (vy) SF_HIP(hipMemcpy(vy, c.vy, b, hipMemcpyDeviceToHost));
    if (fx) SF_HIP(hipMemcpy(fx, c.fx, b, hipMemcpyDeviceToHost));
    if (fy) SF_HIP(hipMemcpy(fy, c.fy, b, hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_set_nodes(sf_ctx* h, const double* x, const double* y, const double* vx, const double* vy) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t b = static_cast<size_t>(c.K) * 8;
    if (b == 0) return SF_OK;
    c.arc_fresh = false;
    if (x || y) drop_contact_prefill(c);
    if (x) SF_HIP(hipMemcpy(c.x, x, b, hipMemcpyHostToDevice));
    if (y) SF_HIP(hipMemcpy(c.y, y, b, hipMemcpyHostToDevice));
    if (vx) SF_HIP(hipMemcpy(c.vx, vx, b, hipMemcpyHostToDevice));
    if (vy) SF_HIP(hipMemcpy(c.vy, vy, b, hipMemcpyHostToDevice));
    return SF_OK;
}

int sf_get_membrane_state(sf_ctx* h, double* prev_len, int32_t* has_prev) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    if (c.K == 0) return SF_OK;
    if (prev_len) SF_HIP(hipMemcpy(prev_len, c.Lprev, c.K * 8, hipMemcpyDeviceToHost));
    if (has_prev) SF_HIP(hipMemcpy(has_prev, c.has_prev, c.ncaps * sizeof(int), hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_set_scalar(sf_ctx* h, int sp, const double* g, const double* C) {
    if (!h || !g || !C) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (sp < 0 || sp >= c.nsp) return fail(SF_E_ARG, "no such scalar species");
    SF_HIP(hipSetDevice(c.device));
    flush_adr(c);
    SF_HIP(hipStreamSynchronize(c.stream));
    const size_t N = c.N;
    if (!c.fp32) {
        std::vector<double> t(9 * N);
        for (size_t n = 0; n < N; ++n) for (int q = 0; q < 9; ++q) t[q * N + n] = g[n * 9 + q];
        for (int b = 0; b < 2; ++b) SF_HIP(hipMemcpy(c.g[sp][b], t.data(), 9 * N * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(9 * N);
        for (size_t n = 0; n < N; ++n) for (int q = 0; q < 9; ++q) t[q * N + n] = static_cast<float>(g[n * 9 + q]);
        for (int b = 0; b < 2; ++b) SF_HIP(hipMemcpy(c.g[sp][b], t.data(), 9 * N * 4, hipMemcpyHostToDevice));
    }
    SF_HIP(hipMemcpy(c.Cstore[sp], C, N * 8, hipMemcpyHostToDevice));
    c.c_from_store = true;
    invalidate_graph(c);
    return SF_OK;
}

int sf_get_scalar(sf_ctx* h, int sp, double* C, double* g) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (sp < 0 || sp >= c.nsp) return fail(SF_E_ARG, "no such scalar species");
    SF_HIP(hipSetDevice(c.device));
    flush_adr(c);
    const size_t N = c.N;
    if (C) {
        if (c.c_from_store) {
            SF_HIP(hipStreamSynchronize(c.stream));
            SF_HIP(hipMemcpy(C, c.Cstore[sp], N * 8, hipMemcpyDeviceToHost));
        } else {
            double* buf = nullptr;
            SF_HIP(hipMalloc(reinterpret_cast<void**>(&buf), N * 8));
            if (c.fp32)
                hipLaunchKernelGGL(k_csum<float>, dim3((c.N + 255) / 256), dim3(256), 0, c.stream,
                                   static_cast<const float*>(c.g[sp][c.gcur]), c.N, buf);
            else
                hipLaunchKernelGGL(k_csum<double>, dim3((c.N + 255) / 256), dim3(256), 0, c.stream,
                                   static_cast<const double*>(c.g[sp][c.gcur]), c.N, buf);
            hipError_t e = hipStreamSynchronize(c.stream);
            if (e == hipSuccess) e = hipMemcpy(C, buf, N * 8, hipMemcpyDeviceToHost);
            hipFree(buf);
            SF_HIP(e);
        }
    }
    if (g) {
        SF_HIP(hipStreamSynchronize(c.stream));
        if (!c.fp32) {
            std::vector<double> t(9 * N);
            SF_HIP(hipMemcpy(t.data(), c.g[sp][c.gcur], 9 * N * 8, hipMemcpyDeviceToHost));
            for (size_t n = 0; n < N; ++n) for (int q = 0; q < 9; ++q) g[n * 9 + q] = t[q * N + n];
        } else {
            std::vector<float> t(9 * N);
            SF_HIP(hipMemcpy(t.data(), c.g[sp][c.gcur], 9 * N * 4, hipMemcpyDeviceToHost));
            for (size_t n = 0; n < N; ++n) for (int q = 0; q < 9; ++q) g[n * 9 + q] = t[q * N + n];
        }
    }
    return SF_OK;
}

int sf_step(sf_ctx* h, int nsteps) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (!c.lattice_ready) return fail(SF_E_STATE, "populations not set (call sf_set_populations first)");
    if (nsteps < 0) return fail(SF_E_ARG, "negative step count");
    SF_HIP(hipSetDevice(c.device));
    const bool use_graph = c.graph_mode && !c.timing && (c.ncaps == 0 || c.periodic_x);
    if (use_graph) {
        int rc = run_graph(c, nsteps);
        if (rc != SF_OK) return rc;
    } else {
        for (int s = 0; s < nsteps; ++s) one_step(c);
    }
    SF_HIP(hipGetLastError());
    int rc = check_counters(c, false);
    if (c.timing) {
        // the per-launch events are read once every launch has completed
        if (rc == SF_OK) rc = poll_counters(c, true);
        SF_HIP(hipStreamSynchronize(c.stream));
        harvest_timing(c);
    }
    return rc;
}

int sf_set_graph_mode(sf_ctx* h, int on) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    h->c.graph_mode = on != 0;
    if (!on) invalidate_graph(h->c);
    return SF_OK;
}

int sf_synchronize(sf_ctx* h) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    SF_HIP(hipStreamSynchronize(h->c.stream));
    return poll_counters(h->c, true);
}

int sf_stage(sf_ctx* h, int stage) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    switch (stage) {
    case SF_STAGE_FORCES: {
        const int spare = 3 - c.Fcur - c.Fprev;
        c.Fprev = c.Fcur;
        c.Fcur = spare;
        if (c.ncaps == 0) launch_fill_int(c, c.counters, 1, 0);
        step_forces(c);
        break;
    }
    case SF_STAGE_IBM: c.pending_shift = step_ibm(c); break;
    case SF_STAGE_LBM: flush_adr(c); launch_lbm(c, c.pending_shift, false); break;
    case SF_STAGE_SCALAR: if (c.nsp) launch_adr(c, c.Fcur); c.adr_pending = false; break;
    case SF_STAGE_MOVE:
        if (c.ncaps) launch_interp_move(c);
        c.shift_next = SHIFT_PREV;
        c.step_count++;
        break;
    default: return fail(SF_E_ARG, "unknown stage");
    }
    SF_HIP(hipGetLastError());
    const int rc = check_counters(c, false);
    return rc != SF_OK ? rc : poll_counters(c, true);
}

int sf_last_capped(sf_ctx* h, int32_t* capped) {
    if (!h || !capped) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipMemcpyAsync(c.host_counters, c.counters, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SF_HIP(hipStreamSynchronize(c.stream));
    *capped = c.host_counters[0];
    return SF_OK;
}

int sf_field_stats(sf_ctx* h, double* out5) {
    if (!h || !out5) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    double* buf = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&buf), (8 + 3 * static_cast<size_t>(c.N) + 5 * 256) * 8));
    launch_field_stats(c, buf);
    hipError_t e = hipStreamSynchronize(c.stream);
    if (e == hipSuccess) e = hipMemcpy(out5, buf, 5 * 8, hipMemcpyDeviceToHost);
    hipFree(buf);
    SF_HIP(e);
    return SF_OK;
}

int sf_field_stats_noalloc(sf_ctx* h, double* out5) {
    if (!h || !out5) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    // the reduction walks the lattice with a fixed grid; its partials fill the
    // scratch array whatever the lattice size, and cells are indexed by int
    if (static_cast<int64_t>(c.nx) * c.ny > INT32_MAX)
        return fail(SF_E_ARG, "lattice larger than 2^31 - 1 cells");
    static_assert(5 * STATS_BLOCKS + 5 <= NOALLOC_SCRATCH_DOUBLES, "scratch holds partials and results");
    SF_HIP(hipSetDevice(c.device));
    double* ptr = nullptr;
    const int rc = scratch_on_device(c.device, &ptr);
    if (rc != SF_OK) return rc;
    double* part = ptr;
    double* res = ptr + 5 * STATS_BLOCKS;
    launch_field_stats_fused(c, part, res);
    SF_HIP(hipGetLastError());
    SF_HIP(hipMemcpyAsync(out5, res, 5 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SF_HIP(hipStreamSynchronize(c.stream));
    return SF_OK;
}

int sf_global_stats(sf_ctx* h, double rho0, double* out11) {
    if (!h || !out11) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    double* buf = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&buf), (16 + 3 * static_cast<size_t>(c.N) + 11 * 256) * 8));
    launch_global_stats(c, rho0, buf);
    hipError_t e = hipStreamSynchronize(c.stream);
    if (e == hipSuccess) e = hipMemcpy(out11, buf, 11 * 8, hipMemcpyDeviceToHost);
    hipFree(buf);
    SF_HIP(e);
    return SF_OK;
}

int sf_host_alloc(void** p, uint64_t bytes) {
    if (!p) return fail(SF_E_ARG, "null pointer");
    *p = nullptr;
    if (bytes == 0) return SF_OK;
    SF_HIP(hipHostMalloc(p, bytes, hipHostMallocDefault));
    return SF_OK;
}

int sf_host_free(void* p) {
    if (p) SF_HIP(hipHostFree(p));
    return SF_OK;
}

int sf_set_probes(sf_ctx* h, int n, const int32_t* cells) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    if (n < 0 || (n > 0 && !cells)) return fail(SF_E_ARG, "probe cells");
    Ctx& c = h->c;
    for (int i = 0; i < n; ++i)
        if (cells[i] < 0 || cells[i] >= c.N) return fail(SF_E_ARG, "probe cell outside the lattice");
    SF_HIP(hipSetDevice(c.device));
    SF_HIP(hipStreamSynchronize(c.stream));
    if (c.probe_cells) { hipFree(c.probe_cells); c.probe_cells = nullptr; }
    c.nprobes = 0;
    if (n > 0) {
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.probe_cells), n * sizeof(int)));
        SF_HIP(hipMemcpy(c.probe_cells, cells, n * sizeof(int), hipMemcpyHostToDevice));
        c.nprobes = n;
    }
    return SF_OK;
}

int sf_snapshot_async(sf_ctx* h, const sf_snapshot* s, int64_t* ticket) {
    if (!h || !s || !ticket) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    for (int sp = c.nsp; sp < 4; ++sp)
        if (s->C[sp]) return fail(SF_E_ARG, "no such scalar species");
    if (s->probes && c.nprobes == 0) return fail(SF_E_ARG, "no probe cells set");
    SF_HIP(hipSetDevice(c.device));
    const size_t N = c.N;
    const bool fields = s->rho || s->ux || s->uy || s->Fx || s->Fy;
    bool anyC = s->stats && c.nsp > 0;
    for (int sp = 0; sp < c.nsp; ++sp) anyC |= s->C[sp] != nullptr;
    // staging: macro 5N | C nsp*N | probes | stats (16 + 3N + 11*256) | sum scratch 256 + 4
    const size_t o_C = 5 * N, o_pr = o_C + c.nsp * N, o_st = o_pr + 3 * static_cast<size_t>(c.nprobes);
    const size_t o_sum = o_st + 16 + 3 * N + 11 * 256, need = o_sum + 260;
    if (need > c.snap_cap) {
        SF_HIP(hipStreamSynchronize(c.stream));
        if (c.snap_dev) hipFree(c.snap_dev);
        c.snap_dev = nullptr;
        c.snap_cap = 0;
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.snap_dev), need * 8));
        c.snap_cap = need;
    }
    double* st = c.snap_dev;
    if (fields) launch_macro(c, st, st + N, st + 2 * N, st + 3 * N, st + 4 * N);
    if (anyC) {
        flush_adr(c);
        for (int sp = 0; sp < c.nsp; ++sp) {
            double* Cd = st + o_C + sp * N;
            if (c.c_from_store)
                SF_HIP(hipMemcpyAsync(Cd, c.Cstore[sp], N * 8, hipMemcpyDeviceToDevice, c.stream));
            else if (c.fp32)
                hipLaunchKernelGGL(k_csum<float>, dim3((c.N + 255) / 256), dim3(256), 0, c.stream,
                                   static_cast<const float*>(c.g[sp][c.gcur]), c.N, Cd);
            else
                hipLaunchKernelGGL(k_csum<double>, dim3((c.N + 255) / 256), dim3(256), 0, c.stream,
                                   static_cast<const double*>(c.g[sp][c.gcur]), c.N, Cd);
        }
    }
    if (s->probes) launch_macro_cells(c, c.probe_cells, c.nprobes, st + o_pr);
    if (s->stats) {
        launch_global_stats(c, s->rho0, st + o_st);
        double* mass = st + o_st + 11;
        if (c.nsp > 0) launch_sum(c, st + o_C, static_cast<int64_t>(c.nsp) * c.N, st + o_sum, mass);
        else SF_HIP(hipMemsetAsync(mass, 0, 8, c.stream));
    }
    auto d2h = [&](double* dst, const double* src, size_t count) -> hipError_t {
        if (!dst || count == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, count * 8, hipMemcpyDeviceToHost, c.stream);
    };
    double* outs[5] = {s->rho, s->ux, s->uy, s->Fx, s->Fy};
    for (int k = 0; k < 5; ++k) SF_HIP(d2h(outs[k], st + k * N, N));
    for (int sp = 0; sp < c.nsp; ++sp) SF_HIP(d2h(s->C[sp], st + o_C + sp * N, N));
    const size_t K = c.K;
    SF_HIP(d2h(s->x, c.x, K)); SF_HIP(d2h(s->y, c.y, K));
    SF_HIP(d2h(s->vx, c.vx, K)); SF_HIP(d2h(s->vy, c.vy, K));
    SF_HIP(d2h(s->fx, c.fx, K)); SF_HIP(d2h(s->fy, c.fy, K));
    SF_HIP(d2h(s->probes, st + o_pr, 3 * static_cast<size_t>(c.nprobes)));
    SF_HIP(d2h(s->stats, st + o_st, 12));
    if (!c.snap_sync) c.snap_sync = new SnapSync();
    struct Note { SnapSync* s; int64_t t; };
    Note* note = new Note{c.snap_sync, c.snap_next};
    {
        std::lock_guard<std::mutex> lk(c.snap_sync->mu);
        c.snap_sync->issued.push_back(note->t);
    }
    hipError_t e = hipLaunchHostFunc(c.stream, [](void* p) {
        Note* n = static_cast<Note*>(p);
        {
            std::lock_guard<std::mutex> lk(n->s->mu);
            n->s->done.push_back(n->t);
        }
        n->s->cv.notify_all();
        delete n;
    }, note);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(c.snap_sync->mu);
        c.snap_sync->issued.pop_back();
        delete note;
        SF_HIP(e);
    }
    *ticket = c.snap_next++;
    return SF_OK;
}

int sf_snapshot_wait(sf_ctx* h, int64_t ticket) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    SnapSync* s = h->c.snap_sync;
    if (!s) return fail(SF_E_ARG, "unknown snapshot ticket");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = std::find(s->issued.begin(), s->issued.end(), ticket);
    if (it == s->issued.end()) return fail(SF_E_ARG, "unknown snapshot ticket");
    s->cv.wait(lk, [&] { return std::find(s->done.begin(), s->done.end(), ticket) != s->done.end(); });
    s->issued.erase(std::find(s->issued.begin(), s->issued.end(), ticket));
    s->done.erase(std::find(s->done.begin(), s->done.end(), ticket));
    return SF_OK;
}

int sf_step_count(sf_ctx* h, int64_t* n) {
    if (!h || !n) return fail(SF_E_ARG, "null pointer");
    *n = h->c.step_count;
    return SF_OK;
}

int sf_graph_captures(sf_ctx* h, int64_t* n) {
    if (!h || !n) return fail(SF_E_ARG, "null pointer");
    *n = h->c.graph_captures;
    return SF_OK;
}

int sf_set_timing(sf_ctx* h, int on) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    h->c.timing = on != 0;
    pools()[&h->c].acc.clear();
    return SF_OK;
}

int sf_kernel_times(sf_ctx* h, int max, const char** names, double* ms, int* count) {
    if (!h || !count) return fail(SF_E_ARG, "null pointer");
    auto& P = pools()[&h->c];
    int i = 0;
    static std::vector<std::string> keep;
    keep.clear();
    for (auto& kv : P.acc) keep.push_back(kv.first);
    for (auto& kv : P.acc) {
        if (i < max) {
            names[i] = keep[i].c_str();
            ms[i] = kv.second.second ? kv.second.first / kv.second.second : 0.0;
        }
        ++i;
    }
    *count = i;
    return SF_OK;
}

int sf_autotune(sf_ctx* h, int reps, char* report, int report_len) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (!c.lattice_ready) return fail(SF_E_STATE, "populations not set");
    SF_HIP(hipSetDevice(c.device));
    if (reps < 1) reps = 10;
    // the steady-state launch: previous-step shift force, scalar riding along
    const bool fuse = c.nsp > 0 && mdf_corrections(c) <= 0;
    const int shift = mdf_corrections(c) > 0 ? SHIFT_PARTIAL : SHIFT_PREV;
    // every candidate must fit the kernel's __launch_bounds__(256); the last
    // entries are persistent grids (workgroups per XCD), XCD-banded only
    static const int shapes[][3] = {{64, 4, 0}, {64, 2, 0}, {64, 1, 0}, {128, 2, 0}, {128, 1, 0},
                                    {256, 1, 0}, {32, 8, 0}, {32, 4, 0}, {256, 1, 64},
                                    {256, 1, 128}, {256, 1, 256}, {128, 2, 128}, {64, 4, 128}};
    hipEvent_t e0, e1;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    std::string rep;
    if (c.K > 0 && c.tile_cnt) {
        // The spread's persistent grid: node binning and the gather of the
        // current node forces into the spare force buffer, which the next
        // step overwrites (its tile flags stay exact); the records and tile
        // lists are rebuilt by the next step's binning.
        static const int grids[] = {512, 1024, 2048, 4096, 8192, 0};
        const int keep_F = c.Fcur, keep_grid = c.spread_grid;
        const bool keep_arc = c.arc_fresh;
        c.arc_fresh = false;   // the positions may have moved since the arc lengths were taken
        c.Fcur = 3 - c.Fcur - c.Fprev;
        double sbest = 1e30;
        int sgrid = keep_grid;
        for (const int gsz : grids) {
            c.spread_grid = gsz;
            launch_spread(c);   // warm
            SF_HIP(hipEventRecord(e0, c.stream));
            for (int r = 0; r < reps; ++r) launch_spread(c);
            SF_HIP(hipEventRecord(e1, c.stream));
            SF_HIP(hipEventSynchronize(e1));
            SF_HIP(hipGetLastError());
            float ms = 0.f;
            SF_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double per = ms / reps;
            char line[96];
            std::snprintf(line, sizeof(line), "spread grid %d %.4f ms\n", gsz > 0 ? std::min(gsz, c.ntiles) : c.ntiles,
                          per);
            rep += line;
            if (per < sbest) { sbest = per; sgrid = gsz; }
        }
        launch_bins_close(c);
        c.Fcur = keep_F;
        c.arc_fresh = keep_arc;
        c.spread_grid = sgrid;
        // lanes per node of the interpolations, timed without the move and
        // into scratch velocities; where the step's interpolation also fills
        // the next step's contact cells, the timing fills scratch cells too
        // (cleared before each run)
        double* scratch_v = nullptr;
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&scratch_v), 2 * static_cast<size_t>(c.K) * sizeof(double)));
        CellFill cf{};
        const int ncells = c.ncx * c.ncy;
        if (contact_uses_cells(c) && c.cell_cnt && !c.any_rigid) {
            int* sc = nullptr;
            SF_HIP(hipMalloc(reinterpret_cast<void**>(&sc),
                             (static_cast<size_t>(ncells) * (1 + CELL_MAXC) + 64) * sizeof(int)));
            SF_HIP(hipMemset(sc, 0, (static_cast<size_t>(ncells) * (1 + CELL_MAXC) + 64) * sizeof(int)));
            cf = CellFill{sc, sc + ncells, sc + static_cast<size_t>(ncells) * (1 + CELL_MAXC), 0.0,
                          static_cast<double>(c.nx), c.cdx, c.cdy, 0.5, c.ncx, c.ncy};
        }
        auto probe = [&](int lpn) {
            if (cf.cnt) launch_fill_int(c, cf.cnt, ncells, 0);
            launch_interp_probe(c, lpn, scratch_v, scratch_v + c.K, cf);
        };
        // three rounds over the candidates, each candidate's best round
        // counts (a first launch of a kernel can be slow)
        const int lpns[3] = {16, 8, 4};
        double per_lpn[3] = {1e30, 1e30, 1e30};
        for (int round = 0; round < 3; ++round) {
            for (int i = 0; i < 3; ++i) {
                probe(lpns[i]);   // warm
                SF_HIP(hipEventRecord(e0, c.stream));
                for (int r = 0; r < reps; ++r) probe(lpns[i]);
                SF_HIP(hipEventRecord(e1, c.stream));
                SF_HIP(hipEventSynchronize(e1));
                SF_HIP(hipGetLastError());
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, e0, e1));
                per_lpn[i] = std::min(per_lpn[i], static_cast<double>(ms) / reps);
            }
        }
        double ibest = 1e30;
        int ilpn = c.interp_lpn;
        for (int i = 0; i < 3; ++i) {
            char line[96];
            std::snprintf(line, sizeof(line), "interpolation %d lanes/node %.4f ms\n", lpns[i], per_lpn[i]);
            rep += line;
            if (per_lpn[i] < ibest) { ibest = per_lpn[i]; ilpn = lpns[i]; }
        }
        SF_HIP(hipFree(scratch_v));
        if (cf.cnt) SF_HIP(hipFree(cf.cnt));
        c.interp_lpn = ilpn;
    }
    double best = 1e30;
    int bbx = c.lbm_bx, bby = c.lbm_by, brm = c.lbm_remap, bps = c.lbm_persist, bpr = c.lbm_pair;
    const int pair_max = lbm_pair_ok(c) ? 1 : 0;
    SF_HIP(hipGetLastError());
    for (const auto& sh : shapes) {
        if (sh[0] * sh[1] > 256) continue;
        for (int remap = sh[2] ? 1 : 0; remap <= 1; ++remap) {
            for (int pair_on = 0; pair_on <= pair_max; ++pair_on) {
                const int pair = pair_on ? (c.generic_flags ? 2 : 1) : 0;
                launch_lbm_cfg(c, shift, fuse, sh[0], sh[1], remap, sh[2], pair);   // warm
                SF_HIP(hipEventRecord(e0, c.stream));
                for (int r = 0; r < reps; ++r) launch_lbm_cfg(c, shift, fuse, sh[0], sh[1], remap, sh[2], pair);
                SF_HIP(hipEventRecord(e1, c.stream));
                SF_HIP(hipEventSynchronize(e1));
                SF_HIP(hipGetLastError());
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, e0, e1));
                const double per = ms / reps;
                const char* cl = pair ? " 2 cells/lane" : "";
                char line[112];
                if (sh[2])
                    std::snprintf(line, sizeof(line), "%dx%d persistent %d/xcd%s %.4f ms\n", sh[0], sh[1], sh[2],
                                  cl, per);
                else
                    std::snprintf(line, sizeof(line), "%dx%d %s%s %.4f ms\n", sh[0], sh[1],
                                  remap ? "xcd-banded" : "linear", cl, per);
                rep += line;
                if (per < best) { best = per; bbx = sh[0]; bby = sh[1]; brm = remap; bps = sh[2]; bpr = pair; }
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    SF_HIP(hipGetLastError());
    launch_fill_int(c, c.counters, 1, 0);
    c.lbm_bx = bbx; c.lbm_by = bby; c.lbm_remap = brm; c.lbm_persist = bps; c.lbm_pair = bpr;
    invalidate_graph(c);
    char line[160];
    const char* cl = (bpr && pair_max) ? " 2 cells/lane" : "";
    char sg[80] = "";
    if (c.K > 0 && c.tile_cnt)
        std::snprintf(sg, sizeof(sg), ", spread grid %d, interpolation %d lanes/node",
                      c.spread_grid > 0 ? std::min(c.spread_grid, c.ntiles) : c.ntiles, c.interp_lpn);
    if (bps)
        std::snprintf(line, sizeof(line), "chosen %dx%d persistent %d/xcd%s%s\n", bbx, bby, bps, cl, sg);
    else
        std::snprintf(line, sizeof(line), "chosen %dx%d %s%s%s\n", bbx, bby, brm ? "xcd-banded" : "linear", cl,
                      sg);
    rep += line;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    SF_HIP(hipStreamSynchronize(c.stream));
    return SF_OK;
}

int sf_set_lbm_launch(sf_ctx* h, int bx, int by, int xcd_banded, int persist) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    if (bx < 1 || by < 1 || bx * by > 256 || (bx * by) % 64 != 0)
        return fail(SF_E_ARG, "launch shape must hold 64..256 threads in whole wavefronts");
    if (persist < 0) return fail(SF_E_ARG, "persistent workgroups per XCD must be >= 0");
    Ctx& c = h->c;
    c.lbm_bx = bx; c.lbm_by = by; c.lbm_remap = (xcd_banded || persist) ? 1 : 0;
    c.lbm_persist = persist;
    invalidate_graph(c);
    return SF_OK;
}

int sf_set_capsule_launch(sf_ctx* h, int spread_grid, int interp_lanes) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    if (spread_grid < 0) return fail(SF_E_ARG, "spread grid must be >= 0");
    if (interp_lanes != 16 && interp_lanes != 8 && interp_lanes != 4)
        return fail(SF_E_ARG, "interpolation lanes per node must be 16, 8 or 4");
    Ctx& c = h->c;
    c.spread_grid = spread_grid;
    c.interp_lpn = interp_lanes;
    invalidate_graph(c);
    return SF_OK;
}

int sf_set_lbm_pairs(sf_ctx* h, int on) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    c.lbm_pair = on ? 2 : 0;
    invalidate_graph(c);
    return SF_OK;
}

int sf_mdf_set_control(sf_ctx* h, double tol, int max_corrections, int probe) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (tol > 0.0 && (max_corrections < 1 || max_corrections > MDF_MAX_CORR))
        return fail(SF_E_ARG, "max_corrections must be in 1..32");
    if (!(tol == tol)) return fail(SF_E_ARG, "tolerance is NaN");
    c.mdf_tol = tol > 0.0 ? tol : 0.0;
    c.mdf_max_corr = tol > 0.0 ? max_corrections : 0;
    c.mdf_probe = probe ? 1 : 0;
    SF_HIP(hipSetDevice(c.device));
    mdf_tail_size(c);
    invalidate_graph(c);
    return SF_OK;
}

int sf_mdf_set_tail(sf_ctx* h, int first) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    if (first != 0 && (first < 2 || first > MDF_MAX_CORR)) return fail(SF_E_ARG, "first must be 0 or in 2..32");
    Ctx& c = h->c;
    c.mdf_tail_from = first;
    invalidate_graph(c);
    return SF_OK;
}

int sf_mdf_stats(sf_ctx* h, double* out4) {
    if (!h || !out4) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    launch_mdf_fold(c);   // a staged step whose closing interpolation has not run
    SF_HIP(hipStreamSynchronize(c.stream));
    MdfDev md{};
    SF_HIP(hipMemcpy(&md, c.mdf, sizeof(MdfDev), hipMemcpyDeviceToHost));
    const double n = md.steps;
    out4[0] = n > 0 ? md.sum_applied / n : 0.0;
    out4[1] = n > 0 ? md.sum_final / n : 0.0;
    out4[2] = md.max_final;
    out4[3] = n;
    md.sum_applied = md.sum_final = md.max_final = md.steps = 0.0;
    SF_HIP(hipMemcpy(c.mdf, &md, sizeof(MdfDev), hipMemcpyHostToDevice));
    return SF_OK;
}

int sf_mdf_collect(sf_ctx* h, int64_t capacity) {
    if (!h || capacity < 0) return fail(SF_E_ARG, "bad arguments");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    launch_mdf_fold(c);
    SF_HIP(hipStreamSynchronize(c.stream));
    if (c.ds_buf) hipFree(c.ds_buf);
    c.ds_buf = nullptr;
    c.ds_cap = 0;
    MdfDev md{};
    SF_HIP(hipMemcpy(&md, c.mdf, sizeof(MdfDev), hipMemcpyDeviceToHost));
    md.ds_rows = 0;
    SF_HIP(hipMemcpy(c.mdf, &md, sizeof(MdfDev), hipMemcpyHostToDevice));
    if (capacity > 0) {
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&c.ds_buf),
                         static_cast<size_t>(capacity) * (MLP_IN + MLP_OUT) * sizeof(float)));
        c.ds_cap = capacity;
    }
    c.mdf_hist_on = (c.mdf_warm || c.ds_buf) ? 1 : 0;
    invalidate_graph(c);
    return SF_OK;
}

int sf_mdf_samples(sf_ctx* h, int64_t* rows, float* rows_out) {
    if (!h || !rows) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    SF_HIP(hipSetDevice(c.device));
    launch_mdf_fold(c);
    SF_HIP(hipStreamSynchronize(c.stream));
    MdfDev md{};
    SF_HIP(hipMemcpy(&md, c.mdf, sizeof(MdfDev), hipMemcpyDeviceToHost));
    *rows = c.ds_buf ? md.ds_rows : 0;
    if (rows_out && c.ds_buf && md.ds_rows > 0)
        SF_HIP(hipMemcpy(rows_out, c.ds_buf, static_cast<size_t>(md.ds_rows) * (MLP_IN + MLP_OUT) * sizeof(float),
                         hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_mdf_set_model(sf_ctx* h, sf_mlp* net, int warm) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Ctx& c = h->c;
    if (net && net->m.device != c.device) return fail(SF_E_ARG, "network and context are on different devices");
    if (warm && !net) return fail(SF_E_ARG, "warm start needs a network");
    SF_HIP(hipSetDevice(c.device));
    if (net) SF_HIP(hipStreamSynchronize(net->m.stream));
    c.model = net ? &net->m : nullptr;
    c.mdf_warm = warm != 0;
    c.mdf_hist_on = (c.mdf_warm || c.ds_buf) ? 1 : 0;
    invalidate_graph(c);
    return SF_OK;
}

int sf_probe_math(int op, int n, const double* a, const double* b, double* out) {
    if (n <= 0 || !a || !out || (op != 0 && !b)) return fail(SF_E_ARG, "bad probe arguments");
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&da), n * 8));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&db), n * 8));
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&dout), n * 8));
    SF_HIP(hipMemcpy(da, a, n * 8, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(db, b ? b : a, n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, nullptr, op, n, da, db, dout);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost);
    hipFree(da); hipFree(db); hipFree(dout);
    SF_HIP(e);
    return SF_OK;
}

int sf_copy_ceiling(int device, uint64_t bytes, int reps, double* gbps) {
    if (!gbps) return fail(SF_E_ARG, "null pointer");
    if (bytes < (1ull << 20) || bytes > (8ull << 30) || bytes % 16 != 0)
        return fail(SF_E_ARG, "bytes must be a multiple of 16 in [1 MiB, 8 GiB]");
    if (reps < 1 || reps > 1000) return fail(SF_E_ARG, "reps must be in 1..1000");
    SF_HIP(hipSetDevice(device));
    const int64_t n = static_cast<int64_t>(bytes / 16);
    const int64_t nblk = (n + 255) / 256;
    if (nblk > 0x7fffffffLL) return fail(SF_E_ARG, "copy too large for one grid");
    copy_v4 *a = nullptr, *b = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&a), bytes));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b), bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t st = nullptr;
    if (e == hipSuccess) e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, st);
    float ms = 0.f;
    if (e == hipSuccess) {
        const dim3 g(static_cast<unsigned>(nblk)), blk(256);
        hipLaunchKernelGGL(k_copy_nt, g, blk, 0, st, a, b, n);   // warm
        hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_nt, g, blk, 0, st, a, b, n);
        hipEventRecord(e1, st);
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    hipFree(a);
    if (b) hipFree(b);
    SF_HIP(e);
    *gbps = ms > 0.f ? 2.0 * static_cast<double>(bytes) * reps / (static_cast<double>(ms) * 1e-3) / 1e9 : 0.0;
    return SF_OK;
}

}  // extern "C"
