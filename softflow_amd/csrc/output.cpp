// Output writers over asynchronous device snapshots (see output.h).
#include "output.h"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "engine.h"

namespace softflow {

namespace {

// "%.6e" appended to s (the same bytes fprintf writes)
inline void put_e(std::string& s, double v) {
    char b[40];
    const int n = std::snprintf(b, sizeof(b), "%.6e", v);
    s.append(b, static_cast<size_t>(n));
}

inline void put_i(std::string& s, int v) {
    char b[16];
    const int n = std::snprintf(b, sizeof(b), "%d", v);
    s.append(b, static_cast<size_t>(n));
}

inline void put_c(std::string& s, char c) { s.push_back(c); }

// rows [0, rows) formatted by fn(row, out) on up to 16 threads, written in order
template <typename Fn>
void emit_rows(FILE* fp, int rows, Fn fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int T = static_cast<int>(std::min(16u, hw == 0 ? 1u : hw));
    if (rows < 64) T = 1;
    T = std::max(1, std::min(T, rows));
    std::vector<std::string> parts(T);
    auto run = [&](int t) {
        const int r0 = static_cast<int>(static_cast<int64_t>(rows) * t / T);
        const int r1 = static_cast<int>(static_cast<int64_t>(rows) * (t + 1) / T);
        std::string& s = parts[t];
        for (int r = r0; r < r1; ++r) fn(r, s);
    };
    if (T == 1) {
        run(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 1; t < T; ++t) th.emplace_back(run, t);
        run(0);
        for (auto& x : th) x.join();
    }
    for (const std::string& s : parts) std::fwrite(s.data(), 1, s.size(), fp);
}

inline void fput(FILE* fp, const std::string& s) { std::fwrite(s.data(), 1, s.size(), fp); }

// central differences with clamped neighbours (vtk_writer.cpp:39-73)
struct Grad {
    double dux_dx, duy_dy, dux_dy, duy_dx;
};
inline Grad gradient(const Frame& f, int x, int y) {
    const int nx = f.nx, ny = f.ny;
    const int xp = std::min(x + 1, nx - 1), xm = std::max(x - 1, 0);
    const int yp = std::min(y + 1, ny - 1), ym = std::max(y - 1, 0);
    auto U = [&](int i, int j) { return f.ux[static_cast<size_t>(j) * nx + i]; };
    auto V = [&](int i, int j) { return f.uy[static_cast<size_t>(j) * nx + i]; };
    Grad g;
    g.dux_dx = (U(xp, y) - U(xm, y)) / static_cast<double>(xp - xm);
    g.duy_dy = (V(x, yp) - V(x, ym)) / static_cast<double>(yp - ym);
    g.dux_dy = (U(x, yp) - U(x, ym)) / static_cast<double>(yp - ym);
    g.duy_dx = (V(xp, y) - V(xm, y)) / static_cast<double>(xp - xm);
    return g;
}

inline double vorticity(const Frame& f, int x, int y) {
    const Grad g = gradient(f, x, y);
    return g.duy_dx - g.dux_dy;
}

inline double strain_rate(const Frame& f, int x, int y) {
    const Grad g = gradient(f, x, y);
    const double s11 = g.dux_dx, s22 = g.duy_dy, s12 = 0.5 * (g.dux_dy + g.duy_dx);
    return std::sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
}

// A capsule's polygon for VTK: nodes made contiguous across the periodic seam,
// split into two clipped copies when the polygon leaves [0, Lx)
// (vtk_writer.cpp:211-289).
struct Poly {
    int cap;
    std::vector<Vec2d> pts;
};

std::vector<Poly> polygons(const std::vector<Capsule>& caps, int periodic_nx) {
    std::vector<Poly> out;
    out.reserve(caps.size() * 2);
    const double Lx = static_cast<double>(periodic_nx);
    for (int c = 0; c < static_cast<int>(caps.size()); ++c) {
        const Capsule& cp = caps[c];
        const int n = cp.numNodes();
        std::vector<Vec2d> uw(n);
        uw[0] = cp.nodePosition(0);
        for (int k = 1; k < n; ++k) {
            if (periodic_nx > 0) {
                Vec2d d = cp.nodePosition(k) - cp.nodePosition(k - 1);
                if (d.x > 0.5 * Lx) d.x -= Lx;
                if (d.x < -0.5 * Lx) d.x += Lx;
                uw[k] = uw[k - 1] + d;
            } else {
                uw[k] = cp.nodePosition(k);
            }
        }
        if (periodic_nx <= 0) {
            out.push_back({c, std::move(uw)});
            continue;
        }
        double lo = uw[0].x, hi = uw[0].x;
        for (const Vec2d& p : uw) {
            if (p.x < lo) lo = p.x;
            if (p.x > hi) hi = p.x;
        }
        const bool right = hi >= Lx, left = lo < 0.0;
        if (!right && !left) {
            out.push_back({c, std::move(uw)});
            continue;
        }
        auto clampx = [&](double v) { return v < 0.0 ? 0.0 : (Lx < v ? Lx : v); };
        std::vector<Vec2d> a(uw.size()), b(uw.size());
        const double shift = right ? -Lx : Lx;
        for (size_t k = 0; k < uw.size(); ++k) {
            a[k] = Vec2d{clampx(uw[k].x), uw[k].y};
            b[k] = Vec2d{clampx(uw[k].x + shift), uw[k].y};
        }
        out.push_back({c, std::move(a)});
        out.push_back({c, std::move(b)});
    }
    return out;
}

Vec2d mean_velocity(const Capsule& c) {
    Vec2d v{0.0, 0.0};
    const int n = c.numNodes();
    for (int k = 0; k < n; ++k) {
        const Vec2d w = c.nodeVelocity(k);
        v.x += w.x;
        v.y += w.y;
    }
    v.x /= n;
    v.y /= n;
    return v;
}

Vec2d total_force(const Capsule& c) {
    Vec2d F{0.0, 0.0};
    for (int k = 0; k < c.numNodes(); ++k) {
        const Vec2d f = c.nodeForce(k);
        F.x += f.x;
        F.y += f.y;
    }
    return F;
}

}  // namespace

bool ParticleFilter::passes(int id, int type_id, int group, double cx, double cy) const {
    auto in = [](const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    switch (filter_type) {
    case ALL: return true;
    case BY_ID: return in(selected_ids, id);
    case BY_TYPE: return in(selected_types, type_id);
    case BY_GROUP: return in(selected_groups, group);
    case BY_REGION: return cx >= region_xmin && cx <= region_xmax && cy >= region_ymin && cy <= region_ymax;
    case COMBINED:
        if (!selected_types.empty() && !in(selected_types, type_id)) return false;
        if (!selected_ids.empty() && !in(selected_ids, id)) return false;
        if (!selected_groups.empty() && !in(selected_groups, group)) return false;
        return !(cx < region_xmin || cx > region_xmax || cy < region_ymin || cy > region_ymax);
    }
    return true;
}

OutputManager::OutputManager(const SimulationParams& p, int nsp)
    : ctx_(nullptr), dir_(p.output_dir), nx_(p.nx), ny_(p.ny), nsp_(nsp),
      periodic_nx_(p.fluid.boundary_type == BoundaryType::PERIODIC ? p.nx : 0), dt_(p.dt),
      rho0_(p.fluid.rho0), vtk_every_(p.vtk_dump_every), csv_every_(p.csv_dump_every),
      probe_every_(p.probe_dump_every), stats_every_(p.stats_dump_every),
      output_interval_(p.output_interval), legacy_(p.vtk_format == "legacy"), has_scalar_(nsp > 0) {
    cfg_.output_dir = dir_;
    cfg_.vtk_format = p.vtk_format;
    csv_.filename = "particle_data.csv";
    csv_.format = "csv";
    thread_ = std::thread(&OutputManager::worker, this);
}

OutputManager::~OutputManager() {
    try {
        drain();
    } catch (...) {
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    for (FILE* fp : {stats_fp_, probe_fp_, csv_.fp}) if (fp) std::fclose(fp);
    for (auto& e : extra_csv_) if (e->fp) std::fclose(e->fp);
    freeSlots();
}

void OutputManager::attach(sf_ctx* ctx) {
    drain();
    ctx_ = ctx;
    freeSlots();
    probes_dirty_ = true;
}

int OutputManager::vtkInterval() const { return vtk_every_ > 0 ? vtk_every_ : output_interval_; }
int OutputManager::csvInterval() const { return csv_every_ > 0 ? csv_every_ : output_interval_; }

bool OutputManager::due(int64_t s) const {
    const int v = vtkInterval(), c = csvInterval();
    return (v > 0 && s % v == 0) || (c > 0 && s % c == 0) ||
           (probe_every_ > 0 && !cfg_.probes.empty() && s % probe_every_ == 0) ||
           (stats_every_ > 0 && s % stats_every_ == 0);
}

int64_t OutputManager::nextDue(int64_t from) const {
    int64_t best = -1;
    auto consider = [&](int iv) {
        if (iv <= 0) return;
        const int64_t n = ((from + iv - 1) / iv) * iv;
        if (best < 0 || n < best) best = n;
    };
    consider(vtkInterval());
    consider(csvInterval());
    if (!cfg_.probes.empty()) consider(probe_every_);
    consider(stats_every_);
    return best;
}

void OutputManager::setOutputConfig(const OutputConfig& c) {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    // simulation.cpp:971-994: fields, filters, extra CSVs and probes; the
    // directory and intervals stay those of the parameters
    cfg_.fluid_fields = c.fluid_fields;
    cfg_.particle_vtk_fields = c.particle_vtk_fields;
    cfg_.vtk_format = c.vtk_format;
    csv_.fields = c.csv_fields;
    csv_.filter = c.csv_filter;
    cfg_.csv_fields = c.csv_fields;
    cfg_.csv_filter = c.csv_filter;
    for (auto& e : extra_csv_) if (e->fp) std::fclose(e->fp);
    extra_csv_.clear();
    for (const ExtraCSVConfig& e : c.extra_csv) {
        auto w = std::make_unique<CsvState>();
        w->filename = e.filename;
        w->format = "csv";
        w->append = e.append;
        w->fields = e.fields;
        w->filter = e.filter;
        extra_csv_.push_back(std::move(w));
    }
    cfg_.extra_csv = c.extra_csv;
    for (const FluidProbe& p : c.probes) {
        cfg_.probes.push_back(p);
        probe_header_ = false;
    }
    probes_dirty_ = true;
}

void OutputManager::setVTKOutput(int every, const std::string& format, const FluidOutputFields& ff,
                                 const ParticleVTKFields& pf) {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    vtk_every_ = every;
    cfg_.vtk_format = format;
    cfg_.fluid_fields = ff;
    cfg_.particle_vtk_fields = pf;
}

void OutputManager::setCSVOutput(int every, const std::string& format, const ParticleOutputFields& fields,
                                 const ParticleFilter& filter, bool append) {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    csv_every_ = every;
    if (csv_.fp) std::fclose(csv_.fp);
    csv_ = CsvState{};
    csv_.filename = "particle_data.csv";
    csv_.format = format;
    csv_.append = append;
    csv_.fields = fields;
    csv_.filter = filter;
}

void OutputManager::addExtraCSVOutput(const std::string& filename, const ParticleOutputFields& fields,
                                      const ParticleFilter& filter) {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    auto w = std::make_unique<CsvState>();
    w->filename = filename;
    w->format = "csv";
    w->fields = fields;
    w->filter = filter;
    extra_csv_.push_back(std::move(w));
}

void OutputManager::addFluidProbe(int i, int j, const std::string& label) {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    cfg_.probes.push_back(FluidProbe{i, j, label});
    probe_header_ = false;
    probes_dirty_ = true;
}

void OutputManager::uploadProbes() {
    std::vector<int32_t> cells;
    for (const FluidProbe& p : cfg_.probes) {
        const int x = std::clamp(p.i, 0, nx_ - 1), y = std::clamp(p.j, 0, ny_ - 1);
        cells.push_back(y * nx_ + x);
    }
    const int rc = sf_set_probes(ctx_, static_cast<int>(cells.size()), cells.data());
    if (rc != SF_OK) throw std::runtime_error(std::string("sf_set_probes: ") + sf_last_error());
    probes_dirty_ = false;
}

void OutputManager::freeSlots() {
    for (Slot& s : slots_) {
        for (double** p : {&s.fields, &s.nodes, &s.probes})
            if (*p) { sf_host_free(*p); *p = nullptr; }
        s.nfields = s.nnodes = s.nprobes = 0;
        s.busy = false;
    }
}

void OutputManager::capture(int64_t step, const std::vector<Capsule>& caps, const std::vector<int>& flags) {
    if (!ctx_) throw std::runtime_error("OutputManager: no device context");
    const int s = static_cast<int>(step);
    const int v = vtkInterval(), c = csvInterval();
    auto fr = std::make_unique<Frame>();
    fr->step = s;
    fr->time = static_cast<double>(step) * dt_;
    fr->vtk = v > 0 && step % v == 0;
    fr->csv = c > 0 && step % c == 0;
    fr->probes = probe_every_ > 0 && !cfg_.probes.empty() && step % probe_every_ == 0;
    fr->stats = stats_every_ > 0 && step % stats_every_ == 0;
    if (!(fr->vtk || fr->csv || fr->probes || fr->stats)) return;
    fr->nx = nx_;
    fr->ny = ny_;
    fr->nsp = nsp_;
    const bool need_nodes = fr->vtk || fr->csv || fr->stats;
    if (need_nodes) {
        fr->caps = caps;
        for (Capsule& cp : fr->caps) cp.sys_ = nullptr;
    }
    if (fr->vtk) fr->flags = flags;
    if (fr->probes && probes_dirty_) uploadProbes();

    // a free pinned buffer set (the writer thread frees one when done with it)
    int slot = -1;
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !error_.empty() || !slots_[0].busy || !slots_[1].busy; });
        if (!error_.empty()) throw std::runtime_error("output writer: " + error_);
        slot = !slots_[0].busy ? 0 : 1;
        slots_[slot].busy = true;
    }
    Slot& sl = slots_[slot];
    const size_t N = static_cast<size_t>(nx_) * ny_;
    size_t K = 0;
    for (const Capsule& cp : caps) K += static_cast<size_t>(cp.numNodes());
    auto grow = [&](double*& p, size_t& have, size_t need) {
        if (need <= have) return;
        if (p) sf_host_free(p);
        p = nullptr;
        have = 0;
        void* q = nullptr;
        if (sf_host_alloc(&q, need * sizeof(double)) != SF_OK)
            throw std::runtime_error(std::string("sf_host_alloc: ") + sf_last_error());
        p = static_cast<double*>(q);
        have = need;
    };
    sf_snapshot snap{};
    snap.rho0 = rho0_;
    if (fr->vtk) {
        grow(sl.fields, sl.nfields, (5 + static_cast<size_t>(nsp_)) * N);
        double* b = sl.fields;
        fr->rho = b; fr->ux = b + N; fr->uy = b + 2 * N; fr->Fx = b + 3 * N; fr->Fy = b + 4 * N;
        snap.rho = fr->rho; snap.ux = fr->ux; snap.uy = fr->uy; snap.Fx = fr->Fx; snap.Fy = fr->Fy;
        for (int sp = 0; sp < nsp_; ++sp) {
            fr->C[sp] = b + (5 + sp) * N;
            snap.C[sp] = fr->C[sp];
        }
    }
    if (need_nodes && K > 0) {
        grow(sl.nodes, sl.nnodes, 6 * K);
        double* b = sl.nodes;
        fr->x = b; fr->y = b + K; fr->vx = b + 2 * K; fr->vy = b + 3 * K; fr->fx = b + 4 * K; fr->fy = b + 5 * K;
        snap.x = fr->x; snap.y = fr->y; snap.vx = fr->vx; snap.vy = fr->vy; snap.fx = fr->fx; snap.fy = fr->fy;
    }
    if (fr->probes) {
        grow(sl.probes, sl.nprobes, 3 * cfg_.probes.size());
        fr->probe_vals = sl.probes;
        snap.probes = sl.probes;
    }
    if (fr->stats) snap.stats = fr->stat_vals;
    fr->slot = slot;
    int rc = sf_snapshot_async(ctx_, &snap, &fr->ticket);
    if (rc != SF_OK) {
        std::lock_guard<std::mutex> lk(mu_);
        slots_[slot].busy = false;
        throw std::runtime_error(std::string("sf_snapshot_async: ") + sf_last_error());
    }
    if (fr->vtk) {
        recorded_steps_.push_back(fr->step);
        recorded_times_.push_back(fr->time);
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        queue_.push_back(std::move(fr));
        ++inflight_;
    }
    cv_.notify_all();
}

void OutputManager::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return inflight_ == 0 || !error_.empty(); });
    if (!error_.empty()) {
        const std::string e = error_;
        error_.clear();
        throw std::runtime_error("output writer: " + e);
    }
}

void OutputManager::worker() {
    for (;;) {
        std::unique_ptr<Frame> fr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;
            fr = std::move(queue_.front());
            queue_.pop_front();
        }
        std::string err;
        try {
            if (sf_snapshot_wait(ctx_, fr->ticket) != SF_OK)
                throw std::runtime_error(std::string("sf_snapshot_wait: ") + sf_last_error());
            write(*fr);
        } catch (const std::exception& e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (fr->slot >= 0) slots_[fr->slot].busy = false;
            --inflight_;
            if (!err.empty() && error_.empty()) error_ = err;
        }
        cv_.notify_all();
    }
}

void OutputManager::ensureDir(const std::string& d) {
    if (std::find(made_dirs_.begin(), made_dirs_.end(), d) != made_dirs_.end()) return;
    mkdir(d.c_str(), 0755);
    made_dirs_.push_back(d);
}

void OutputManager::write(Frame& f) {
    // node data into the capsule copies
    size_t k = 0;
    for (Capsule& cp : f.caps)
        for (int i = 0; i < cp.n_; ++i, ++k) {
            cp.pos_[i] = Vec2d{f.x[k], f.y[k]};
            cp.vel_[i] = Vec2d{f.vx[k], f.vy[k]};
            cp.frc_[i] = Vec2d{f.fx[k], f.fy[k]};
        }
    ensureDir(dir_);
    // simulation.cpp:742-767: VTK, CSV (and the extra CSVs), probes, statistics
    if (f.vtk) {
        writeFluidVtk(f);
        writeCapsulesVtk(f);
    }
    if (f.csv) {
        writeCsv(csv_, f);
        for (auto& e : extra_csv_) writeCsv(*e, f);
    }
    if (f.probes) writeProbes(f);
    if (f.stats) writeStats(f);
}

// GlobalStatsWriter::writeTimestep (global_stats_writer.cpp:41-124)
void OutputManager::writeStats(const Frame& f) {
    if (!stats_fp_) stats_fp_ = std::fopen((dir_ + "/global_stats.csv").c_str(), "a");
    if (!stats_fp_) return;
    if (!stats_header_) {
        std::fputs("timestep,time,mean_rho,max_velocity,mean_velocity,total_kinetic_energy,"
                   "n_active_particles,mean_deformation,max_deformation,min_rho,max_rho",
                   stats_fp_);
        if (has_scalar_) std::fputs(",total_dissolved_mass,total_particle_mass,total_adsorbed_mass", stats_fp_);
        std::fputs("\n", stats_fp_);
        stats_header_ = true;
    }
    const double* g = f.stat_vals;
    const double n = g[0];
    double mean_def = 0.0, max_def = 0.0;
    for (const Capsule& cp : f.caps) {
        const double D = cp.deformationIndex();
        mean_def += D;
        if (D > max_def) max_def = D;
    }
    const int ncaps = static_cast<int>(f.caps.size());
    if (ncaps > 0) mean_def /= ncaps;
    std::string s;
    put_i(s, f.step); put_c(s, ',');
    put_e(s, f.time); put_c(s, ',');
    put_e(s, n > 0 ? g[1] / n : 0.0); put_c(s, ',');
    put_e(s, std::sqrt(g[4])); put_c(s, ',');
    put_e(s, n > 0 ? g[5] / n : 0.0); put_c(s, ',');
    put_e(s, g[6]); put_c(s, ',');
    put_i(s, ncaps); put_c(s, ',');
    put_e(s, mean_def); put_c(s, ',');
    put_e(s, max_def); put_c(s, ',');
    put_e(s, n > 0 ? g[2] : 1e30); put_c(s, ',');
    put_e(s, n > 0 ? g[3] : -1e30);
    if (has_scalar_) {
        // no particle reservoirs or adsorption on this path: their masses are zero
        put_c(s, ','); put_e(s, g[11]);
        put_c(s, ','); put_e(s, 0.0);
        put_c(s, ','); put_e(s, 0.0);
    }
    s.push_back('\n');
    fput(stats_fp_, s);
    std::fflush(stats_fp_);
}

// FluidProbeWriter::writeTimestep (fluid_probe_writer.cpp:27-68)
void OutputManager::writeProbes(const Frame& f) {
    if (!probe_fp_) probe_fp_ = std::fopen((dir_ + "/fluid_probes.csv").c_str(), "a");
    if (!probe_fp_) return;
    std::string s;
    if (!probe_header_) {
        s += "timestep,time";
        for (const FluidProbe& p : cfg_.probes) {
            const std::string& l = p.label;
            s += "," + l + "_rho," + l + "_ux," + l + "_uy," + l + "_p";
        }
        s += "\n";
        probe_header_ = true;
    }
    put_i(s, f.step); put_c(s, ',');
    put_e(s, f.time);
    for (size_t i = 0; i < cfg_.probes.size(); ++i) {
        const double r = f.probe_vals[3 * i], u = f.probe_vals[3 * i + 1], v = f.probe_vals[3 * i + 2];
        put_c(s, ','); put_e(s, r);
        put_c(s, ','); put_e(s, u);
        put_c(s, ','); put_e(s, v);
        put_c(s, ','); put_e(s, r / 3.0);
    }
    s.push_back('\n');
    fput(probe_fp_, s);
    std::fflush(probe_fp_);
}

// CSVWriter::writeTimestep (csv_writer.cpp:23-146)
void OutputManager::writeCsv(CsvState& w, const Frame& f) {
    if (!w.fp) w.fp = std::fopen((dir_ + "/" + w.filename).c_str(), w.append ? "a" : "w");
    if (!w.fp) return;
    const char sep = (w.format == "dat" || w.format == "DAT") ? '\t' : ',';
    const ParticleOutputFields& F = w.fields;
    std::string s;
    if (!w.header) {
        s += "timestep"; s += sep; s += "time"; s += sep; s += "capsule_id";
        auto col = [&](const char* name) { s += sep; s += name; };
        if (F.type) col("type");
        if (F.group) col("group");
        if (F.com_position) { col("cx"); col("cy"); }
        if (F.com_velocity) { col("vx"); col("vy"); }
        if (F.force) { col("fx"); col("fy"); }
        if (F.diameter) col("diameter");
        if (F.deformation) col("deformation_index");
        if (F.area_volume) { col("area"); col("area_ratio"); col("perimeter"); col("perimeter_ratio"); }
        if (F.orientation) col("orientation_angle");
        if (F.angular_vel) col("angular_velocity");
        s += "\n";
        w.header = true;
    }
    for (const Capsule& cp : f.caps) {
        const Vec2d cen = cp.centroid();
        const int id = cp.getId(), tid = cp.getType(), group = 0;
        if (!w.filter.passes(id, tid, group, cen.x, cen.y)) continue;
        const Vec2d vel = mean_velocity(cp);
        const int n = cp.numNodes();
        put_i(s, f.step); s += sep; put_e(s, f.time); s += sep; put_i(s, id);
        if (F.type) { s += sep; put_i(s, tid); }
        if (F.group) { s += sep; put_i(s, group); }
        if (F.com_position) { s += sep; put_e(s, cen.x); s += sep; put_e(s, cen.y); }
        if (F.com_velocity) { s += sep; put_e(s, vel.x); s += sep; put_e(s, vel.y); }
        if (F.force) {
            const Vec2d T = total_force(cp);
            s += sep; put_e(s, T.x); s += sep; put_e(s, T.y);
        }
        if (F.diameter) { s += sep; put_e(s, 2.0 * cp.effectiveRadius()); }
        if (F.deformation) { s += sep; put_e(s, cp.deformationIndex()); }
        if (F.area_volume) {
            const double A = cp.area(), P = cp.perimeter(), r = cp.effectiveRadius();
            const double A0 = PI * r * r, P0 = 2.0 * PI * r;
            s += sep; put_e(s, A);
            s += sep; put_e(s, A0 > 1e-15 ? A / A0 : 0.0);
            s += sep; put_e(s, P);
            s += sep; put_e(s, P0 > 1e-15 ? P / P0 : 0.0);
        }
        if (F.orientation) {
            double Ixx = 0, Iyy = 0, Ixy = 0;
            for (int k = 0; k < n; ++k) {
                const Vec2d p = cp.nodePosition(k);
                const double dx = p.x - cen.x, dy = p.y - cen.y;
                Ixx += dy * dy;
                Iyy += dx * dx;
                Ixy -= dx * dy;
            }
            s += sep; put_e(s, 0.5 * std::atan2(2.0 * Ixy, Ixx - Iyy));
        }
        if (F.angular_vel) {
            double num = 0, den = 0;
            for (int k = 0; k < n; ++k) {
                const Vec2d p = cp.nodePosition(k), v = cp.nodeVelocity(k);
                const double dx = p.x - cen.x, dy = p.y - cen.y;
                num += dx * (v.y - vel.y) - dy * (v.x - vel.x);
                den += dx * dx + dy * dy;
            }
            s += sep; put_e(s, den > 1e-15 ? num / den : 0.0);
        }
        s.push_back('\n');
    }
    fput(w.fp, s);
    std::fflush(w.fp);
}

// VTKWriter::writeFluidField / writeFluidFieldLegacy (vtk_writer.cpp:75-202, 547-679)
void OutputManager::writeFluidVtk(const Frame& f) {
    const int nx = f.nx, ny = f.ny;
    const FluidOutputFields& FF = cfg_.fluid_fields;
    std::string path;
    if (legacy_) {
        path = dir_ + "/";
    } else {
        ensureDir(dir_ + "/fluid");
        ensureDir(dir_ + "/particles");
        path = dir_ + "/fluid/";
    }
    char name[64];
    std::snprintf(name, sizeof(name), legacy_ ? "fluid_%06d.vtk" : "fluid_%06d.vti", f.step);
    FILE* fp = std::fopen((path + name).c_str(), "w");
    if (!fp) return;
    const bool L = legacy_;
    auto sized = [&](const char* xml_name, const char* legacy_name, bool vector) {
        // an array header in either format
        if (L) {
            if (vector) std::fprintf(fp, "VECTORS %s double\n", legacy_name);
            else std::fprintf(fp, "SCALARS %s double 1\nLOOKUP_TABLE default\n", legacy_name);
        } else if (vector) {
            std::fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n", xml_name);
        } else {
            std::fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", xml_name);
        }
    };
    auto close = [&]() { if (!L) std::fputs("</DataArray>\n", fp); };
    auto at = [&](int x, int y) { return static_cast<size_t>(y) * nx + x; };
    auto scalar_rows = [&](auto value) {
        emit_rows(fp, ny, [&](int y, std::string& s) {
            for (int x = 0; x < nx; ++x) { put_e(s, value(x, y)); s.push_back('\n'); }
        });
    };
    if (L) {
        std::fprintf(fp, "# vtk DataFile Version 3.0\nSoftFlow fluid field step %d\nASCII\n"
                         "DATASET STRUCTURED_POINTS\nDIMENSIONS %d %d 1\nORIGIN 0 0 0\nSPACING 1 1 1\n"
                         "POINT_DATA %d\n", f.step, nx, ny, nx * ny);
    } else {
        std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" "
                         "byte_order=\"LittleEndian\">\n<ImageData WholeExtent=\"0 %d 0 %d 0 0\" "
                         "Origin=\"0 0 0\" Spacing=\"1 1 1\">\n<Piece Extent=\"0 %d 0 %d 0 0\">\n"
                         "<PointData Scalars=\"density\" Vectors=\"velocity\">\n",
                     nx - 1, ny - 1, nx - 1, ny - 1);
    }
    auto fluid = [&](int x, int y) { return f.flags[at(x, y)] == static_cast<int>(CellType::FLUID); };
    if (FF.density) {
        sized("density", "density", false);
        scalar_rows([&](int x, int y) { return f.rho[at(x, y)]; });
        close();
    }
    if (FF.velocity) {
        sized("velocity", "velocity", true);
        emit_rows(fp, ny, [&](int y, std::string& s) {
            for (int x = 0; x < nx; ++x) {
                put_e(s, f.ux[at(x, y)]); s.push_back(' ');
                put_e(s, f.uy[at(x, y)]); s += " 0\n";
            }
        });
        close();
    }
    if (FF.velocity_mag) {
        sized("velocity_magnitude", "velocity_magnitude", false);
        scalar_rows([&](int x, int y) {
            const double u = f.ux[at(x, y)], v = f.uy[at(x, y)];
            return std::sqrt(u * u + v * v);
        });
        close();
    }
    if (FF.pressure) {
        sized("pressure", "pressure", false);
        scalar_rows([&](int x, int y) { return fluid(x, y) ? f.rho[at(x, y)] / 3.0 : 0.0; });
        close();
        sized("pressure_fluctuation", "pressure_fluctuation", false);
        scalar_rows([&](int x, int y) { return fluid(x, y) ? (f.rho[at(x, y)] - 1.0) / 3.0 : 0.0; });
        close();
    }
    if (FF.vorticity) {
        sized("vorticity", "vorticity", false);
        scalar_rows([&](int x, int y) { return vorticity(f, x, y); });
        close();
    }
    if (FF.strain_rate) {
        sized("strain_rate", "strain_rate", false);
        scalar_rows([&](int x, int y) { return strain_rate(f, x, y); });
        close();
    }
    if (FF.ibm_force) {
        sized("ibm_force", "ibm_force", true);
        emit_rows(fp, ny, [&](int y, std::string& s) {
            for (int x = 0; x < nx; ++x) {
                put_e(s, f.Fx[at(x, y)]); s.push_back(' ');
                put_e(s, f.Fy[at(x, y)]); s += " 0\n";
            }
        });
        close();
    }
    if (FF.node_type) {
        if (L) std::fputs("SCALARS node_type int 1\nLOOKUP_TABLE default\n", fp);
        else std::fputs("<DataArray type=\"Int32\" Name=\"node_type\" format=\"ascii\">\n", fp);
        emit_rows(fp, ny, [&](int y, std::string& s) {
            for (int x = 0; x < nx; ++x) { put_i(s, f.flags[at(x, y)]); s.push_back('\n'); }
        });
        close();
    }
    if (FF.concentration && has_scalar_) {
        for (int sp = 0; sp < f.nsp; ++sp) {
            char nm[32];
            if (f.nsp == 1) std::snprintf(nm, sizeof(nm), "concentration");
            else std::snprintf(nm, sizeof(nm), "concentration_%d", sp);
            sized(nm, nm, false);
            scalar_rows([&](int x, int y) { return f.C[sp][at(x, y)]; });
            close();
        }
    }
    if (!L) std::fputs("</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n", fp);
    std::fflush(fp);
    std::fclose(fp);
}

// VTKWriter::writeCapsules / writeCapsulesLegacy (vtk_writer.cpp:230-466, 681-929).
// With the scalar on, the per-node coverage and per-capsule reservoir columns
// carry the values of a run without particle chemistry: coverage, released and
// absorbed 0, reservoir -1 (untracked).
void OutputManager::writeCapsulesVtk(const Frame& f) {
    if (f.caps.empty()) return;
    const std::vector<Poly> polys = polygons(f.caps, periodic_nx_);
    int total = 0;
    for (const Poly& p : polys) total += static_cast<int>(p.pts.size());
    const int npoly = static_cast<int>(polys.size());
    char name[64];
    std::snprintf(name, sizeof(name), legacy_ ? "/particles_%06d.vtk" : "/particles/particles_%06d.vtp", f.step);
    FILE* fp = std::fopen((dir_ + name).c_str(), "w");
    if (!fp) return;
    std::string s;
    auto node_vec = [&](bool force) {
        for (const Poly& p : polys) {
            const Capsule& cp = f.caps[p.cap];
            for (int k = 0; k < cp.numNodes(); ++k) {
                const Vec2d v = force ? cp.nodeForce(k) : cp.nodeVelocity(k);
                put_e(s, v.x); s.push_back(' '); put_e(s, v.y); s += " 0\n";
            }
        }
    };
    if (legacy_) {
        char h[128];
        std::snprintf(h, sizeof(h), "# vtk DataFile Version 3.0\nSoftFlow particles step %d\nASCII\n"
                                    "DATASET POLYDATA\nPOINTS %d double\n", f.step, total);
        s += h;
        for (const Poly& p : polys)
            for (const Vec2d& q : p.pts) { put_e(s, q.x); s.push_back(' '); put_e(s, q.y); s += " 0\n"; }
        std::snprintf(h, sizeof(h), "POLYGONS %d %d\n", npoly, npoly + total);
        s += h;
        int off = 0;
        for (const Poly& p : polys) {
            const int n = static_cast<int>(p.pts.size());
            put_i(s, n);
            for (int k = 0; k < n; ++k) { s.push_back(' '); put_i(s, off + k); }
            s.push_back('\n');
            off += n;
        }
        std::snprintf(h, sizeof(h), "POINT_DATA %d\n", total);
        s += h;
        s += "SCALARS capsule_id int 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys)
            for (size_t k = 0; k < p.pts.size(); ++k) { put_i(s, p.cap); s.push_back('\n'); }
        s += "SCALARS capsule_type int 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys)
            for (size_t k = 0; k < p.pts.size(); ++k) { put_i(s, f.caps[p.cap].getType()); s.push_back('\n'); }
        s += "VECTORS node_velocity double\n";
        node_vec(false);
        s += "VECTORS node_force double\n";
        node_vec(true);
        if (has_scalar_) {
            s += "SCALARS gamma double 1\nLOOKUP_TABLE default\n";
            for (const Poly& p : polys)
                for (int k = 0; k < f.caps[p.cap].numNodes(); ++k) { put_e(s, 0.0); s.push_back('\n'); }
        }
        std::snprintf(h, sizeof(h), "CELL_DATA %d\n", npoly);
        s += h;
        s += "SCALARS capsule_id_cell int 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_i(s, p.cap); s.push_back('\n'); }
        s += "SCALARS cell_type int 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_i(s, f.caps[p.cap].getType()); s.push_back('\n'); }
        s += "SCALARS deformation_index double 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_e(s, f.caps[p.cap].deformationIndex()); s.push_back('\n'); }
        s += "VECTORS particle_velocity double\n";
        for (const Poly& p : polys) {
            const Vec2d v = mean_velocity(f.caps[p.cap]);
            put_e(s, v.x); s.push_back(' '); put_e(s, v.y); s += " 0\n";
        }
        s += "VECTORS particle_force double\n";
        for (const Poly& p : polys) {
            const Vec2d T = total_force(f.caps[p.cap]);
            put_e(s, T.x); s.push_back(' '); put_e(s, T.y); s += " 0\n";
        }
        s += "SCALARS particle_speed double 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_e(s, mean_velocity(f.caps[p.cap]).norm()); s.push_back('\n'); }
        s += "SCALARS particle_area double 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_e(s, f.caps[p.cap].area()); s.push_back('\n'); }
        s += "SCALARS particle_perimeter double 1\nLOOKUP_TABLE default\n";
        for (const Poly& p : polys) { put_e(s, f.caps[p.cap].perimeter()); s.push_back('\n'); }
        if (has_scalar_) {
            const char* cols[4] = {"scalar_released", "scalar_absorbed", "Mp", "gamma_mean"};
            const double vals[4] = {0.0, 0.0, -1.0, 0.0};
            for (int c = 0; c < 4; ++c) {
                s += "SCALARS "; s += cols[c]; s += " double 1\nLOOKUP_TABLE default\n";
                for (size_t i = 0; i < polys.size(); ++i) { put_e(s, vals[c]); s.push_back('\n'); }
            }
        }
        fput(fp, s);
        std::fflush(fp);
        std::fclose(fp);
        return;
    }
    char h[256];
    s += "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\">\n<PolyData>\n";
    std::snprintf(h, sizeof(h), "<Piece NumberOfPoints=\"%d\" NumberOfVerts=\"0\" NumberOfLines=\"0\" "
                                "NumberOfStrips=\"0\" NumberOfPolys=\"%d\">\n", total, npoly);
    s += h;
    s += "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (const Poly& p : polys)
        for (const Vec2d& q : p.pts) { put_e(s, q.x); s.push_back(' '); put_e(s, q.y); s += " 0\n"; }
    s += "</DataArray>\n</Points>\n<Polys>\n<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    int off = 0;
    for (const Poly& p : polys) {
        const int n = static_cast<int>(p.pts.size());
        for (int k = 0; k < n; ++k) { put_i(s, off + k); s.push_back(' '); }
        s.push_back('\n');
        off += n;
    }
    s += "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    off = 0;
    for (const Poly& p : polys) { off += static_cast<int>(p.pts.size()); put_i(s, off); s.push_back(' '); }
    s += "\n</DataArray>\n</Polys>\n<PointData>\n";
    const ParticleVTKFields& PF = cfg_.particle_vtk_fields;
    if (PF.particle_id) {
        s += "<DataArray type=\"Int32\" Name=\"capsule_id\" format=\"ascii\">\n";
        for (const Poly& p : polys)
            for (size_t k = 0; k < p.pts.size(); ++k) { put_i(s, p.cap); s.push_back(' '); }
        s += "\n</DataArray>\n";
    }
    if (PF.particle_type) {
        s += "<DataArray type=\"Int32\" Name=\"capsule_type\" format=\"ascii\">\n";
        for (const Poly& p : polys)
            for (size_t k = 0; k < p.pts.size(); ++k) { put_i(s, f.caps[p.cap].getType()); s.push_back(' '); }
        s += "\n</DataArray>\n";
    }
    if (PF.velocity) {
        s += "<DataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        node_vec(false);
        s += "</DataArray>\n";
    }
    if (PF.force) {
        s += "<DataArray type=\"Float64\" Name=\"force\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        node_vec(true);
        s += "</DataArray>\n";
    }
    if (has_scalar_) {
        s += "<DataArray type=\"Float64\" Name=\"gamma\" format=\"ascii\">\n";
        for (const Poly& p : polys)
            for (int k = 0; k < f.caps[p.cap].numNodes(); ++k) { put_e(s, 0.0); s.push_back('\n'); }
        s += "</DataArray>\n";
    }
    s += "</PointData>\n<CellData>\n<DataArray type=\"Int32\" Name=\"cell_capsule_id\" format=\"ascii\">\n";
    for (const Poly& p : polys) { put_i(s, p.cap); s.push_back(' '); }
    s += "\n</DataArray>\n<DataArray type=\"Int32\" Name=\"cell_capsule_type\" format=\"ascii\">\n";
    for (const Poly& p : polys) { put_i(s, f.caps[p.cap].getType()); s.push_back(' '); }
    s += "\n</DataArray>\n<DataArray type=\"Float64\" Name=\"deformation_index\" format=\"ascii\">\n";
    for (const Poly& p : polys) { put_e(s, f.caps[p.cap].deformationIndex()); s.push_back(' '); }
    s += "\n</DataArray>\n<DataArray type=\"Float64\" Name=\"capsule_speed\" format=\"ascii\">\n";
    for (const Poly& p : polys) { put_e(s, mean_velocity(f.caps[p.cap]).norm()); s.push_back(' '); }
    s += "\n</DataArray>\n";
    if (has_scalar_) {
        const char* cols[4] = {"Mp", "gamma_mean", "cumulative_released", "cumulative_absorbed"};
        const double vals[4] = {-1.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < 4; ++c) {
            s += "<DataArray type=\"Float64\" Name=\""; s += cols[c]; s += "\" format=\"ascii\">\n";
            for (size_t i = 0; i < polys.size(); ++i) { put_e(s, vals[c]); s.push_back(' '); }
            s += "\n</DataArray>\n";
        }
    }
    s += "</CellData>\n</Piece>\n</PolyData>\n</VTKFile>\n";
    fput(fp, s);
    std::fflush(fp);
    std::fclose(fp);
}

// VTKWriter::writePVDFiles (vtk_writer.cpp:473-542)
void OutputManager::writePvd() {
    if (recorded_steps_.empty()) return;
    auto one = [&](const std::string& path, const char* pattern) {
        FILE* fp = std::fopen(path.c_str(), "w");
        if (!fp) return;
        std::fputs("<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n<Collection>\n", fp);
        for (size_t d = 0; d < recorded_steps_.size(); ++d) {
            char file[64];
            std::snprintf(file, sizeof(file), pattern, recorded_steps_[d]);
            std::fprintf(fp, "<DataSet timestep=\"%.6e\" file=\"%s\"/>\n", recorded_times_[d], file);
        }
        std::fputs("</Collection>\n</VTKFile>\n", fp);
        std::fclose(fp);
    };
    ensureDir(dir_);
    if (legacy_) {
        one(dir_ + "/fluid.pvd", "fluid_%06d.vtk");
        one(dir_ + "/particles.pvd", "particles_%06d.vtk");
    } else {
        ensureDir(dir_ + "/fluid");
        ensureDir(dir_ + "/particles");
        one(dir_ + "/fluid/fluid.pvd", "fluid_%06d.vti");
        one(dir_ + "/particles/particles.pvd", "particles_%06d.vtp");
    }
}

void OutputManager::finalize() {
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    writePvd();
    for (FILE** fp : {&stats_fp_, &probe_fp_, &csv_.fp})
        if (*fp) { std::fclose(*fp); *fp = nullptr; }
    for (auto& e : extra_csv_)
        if (e->fp) { std::fclose(e->fp); e->fp = nullptr; }
}

}  // namespace softflow
