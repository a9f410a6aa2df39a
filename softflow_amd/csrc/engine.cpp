// Host engine: reference stepping API over libsoftflow_hip's C-ABI.
#include "engine.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>

namespace softflow {

namespace {
constexpr int Q = 9;
constexpr Real W[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                       1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr Real CS2 = 1.0 / 3.0;
constexpr Real CS4 = CS2 * CS2;

// lbm/lattice.h:35-39
inline Real feq(int q, Real rho, Real ux, Real uy) {
    Real eu = static_cast<Real>(CX[q]) * ux + static_cast<Real>(CY[q]) * uy;
    Real uu = ux * ux + uy * uy;
    return W[q] * rho * (1.0 + eu / CS2 + (eu * eu) / (2.0 * CS4) - uu / (2.0 * CS2));
}
}  // namespace

// ============================================================================
LatticeField::LatticeField(int nx, int ny)
    : nx_(nx), ny_(ny), N_(nx * ny),
      f_(static_cast<size_t>(nx) * ny * Q, 0.0), rho_(static_cast<size_t>(nx) * ny, 0.0),
      ux_(rho_.size(), 0.0), uy_(rho_.size(), 0.0), Fx_(rho_.size(), 0.0), Fy_(rho_.size(), 0.0),
      flags_(rho_.size(), static_cast<int>(CellType::FLUID)) {}

void LatticeField::pull() {
    if (stale_ && sim_) {
        sim_->pullField(*this);
        refreshShadows();
    }
}

std::vector<Real>& LatticeField::viewArray(int kind) {
    switch (kind) {
    case VIEW_F: return f_;
    case VIEW_RHO: return rho_;
    case VIEW_UX: return ux_;
    case VIEW_UY: return uy_;
    case VIEW_FX: return Fx_;
    case VIEW_FY: return Fy_;
    default: throw std::invalid_argument("LatticeField: unknown view");
    }
}

Real* LatticeField::leaseView(int kind) {
    std::vector<Real>& a = viewArray(kind);
    pull();
    if (leases_[kind]++ == 0) shadow_[kind] = a;
    return a.data();
}

void LatticeField::releaseView(int kind) {
    if (kind < 0 || kind >= VIEW_COUNT || leases_[kind] == 0) return;
    // writes made through the last view of this array still count
    if (!stale_ && viewArray(kind) != shadow_[kind]) {
        if (kind == VIEW_F) f_dirty_ = true;
        else if (kind <= VIEW_UY) macro_dirty_ = true;
    }
    if (--leases_[kind] == 0) std::vector<Real>().swap(shadow_[kind]);
}

bool LatticeField::anyLease() const {
    for (int k = 0; k < VIEW_COUNT; ++k)
        if (leases_[k]) return true;
    return false;
}

// Writes through leased views since the last step or refresh (the mirror is
// current whenever a lease is held, see Simulation::afterAdvance).
void LatticeField::collectViewWrites() {
    if (stale_) return;
    for (int k = 0; k < VIEW_COUNT; ++k) {
        if (!leases_[k]) continue;
        std::vector<Real>& a = viewArray(k);
        if (a == shadow_[k]) continue;
        if (k == VIEW_F) f_dirty_ = true;
        else if (k <= VIEW_UY) macro_dirty_ = true;
        shadow_[k] = a;
    }
}

void LatticeField::refreshShadows() {
    for (int k = 0; k < VIEW_COUNT; ++k)
        if (leases_[k]) shadow_[k] = viewArray(k);
}
void LatticeField::set_f(int x, int y, int q, Real v) {
    touch_f();
    f_[static_cast<size_t>(q) * N_ + idx(x, y)] = v;
}
void LatticeField::setCellType(int x, int y, CellType ct) {
    flags_[idx(x, y)] = static_cast<int>(ct);
    flags_dirty_ = true;
}
void LatticeField::setRho(int x, int y, Real v) { touch_macro(); rho_[idx(x, y)] = v; }
void LatticeField::setUx(int x, int y, Real v) { touch_macro(); ux_[idx(x, y)] = v; }
void LatticeField::setUy(int x, int y, Real v) { touch_macro(); uy_[idx(x, y)] = v; }

// lattice_field.cpp:75-90
void LatticeField::setEquilibrium(Real rho0, Real ux0, Real uy0) {
    touch_f();
    touch_macro();
    for (int n = 0; n < N_; ++n) {
        rho_[n] = rho0; ux_[n] = ux0; uy_[n] = uy0;
        for (int q = 0; q < Q; ++q) f_[static_cast<size_t>(q) * N_ + n] = feq(q, rho0, ux0, uy0);
    }
}
// lattice_field.cpp:92-113
void LatticeField::initializeEquilibriumAt(int x, int y) {
    touch_f();
    const int n = idx(x, y);
    for (int q = 0; q < Q; ++q) f_[static_cast<size_t>(q) * N_ + n] = feq(q, rho_[n], ux_[n], uy_[n]);
}
void LatticeField::initializeAllEquilibrium() {
    touch_f();
    for (int n = 0; n < N_; ++n)
        for (int q = 0; q < Q; ++q) f_[static_cast<size_t>(q) * N_ + n] = feq(q, rho_[n], ux_[n], uy_[n]);
}
// lattice_field.cpp:28-73, on the host mirror
void LatticeField::computeMacroscopic() {
    touch_macro();
    for (int n = 0; n < N_; ++n) {
        if (flags_[n] == static_cast<int>(CellType::SOLID) || flags_[n] == static_cast<int>(CellType::EMPTY)) {
            rho_[n] = ux_[n] = uy_[n] = 0.0;
            continue;
        }
        Real r = 0.0, vx = 0.0, vy = 0.0;
        for (int q = 0; q < Q; ++q) {
            Real fi = f_[static_cast<size_t>(q) * N_ + n];
            r += fi;
            vx += fi * static_cast<Real>(CX[q]);
            vy += fi * static_cast<Real>(CY[q]);
        }
        rho_[n] = r;
        if (r > 1e-15) { ux_[n] = (vx + 0.5 * Fx_[n]) / r; uy_[n] = (vy + 0.5 * Fy_[n]) / r; }
        else { ux_[n] = 0.0; uy_[n] = 0.0; }
    }
}

// ============================================================================
// Capsule rest geometry: membrane/capsule.cpp:7-141
Capsule::Capsule(int id, Vec2d center, Real radius, int num_nodes, const MembraneParams& p, int type)
    : id_(id), type_(type), n_(num_nodes > 0 ? num_nodes : computeOptimalNodes(radius)), params_(p) {
    pos_.resize(n_);
    vel_.assign(n_, Vec2d{0.0, 0.0});
    frc_.assign(n_, Vec2d{0.0, 0.0});
    rest_len_.resize(n_);
    auto shape_pt = [&](Real th) -> Vec2d {
        if (p.shape == CapsuleShape::ELLIPSE || p.shape == CapsuleShape::FIBER)
            return {radius * std::cos(th), radius * p.aspect_ratio * std::sin(th)};
        if (p.shape == CapsuleShape::BICONCAVE) {
            Real r = radius * (1.0 - p.indent_depth * std::sin(th) * std::sin(th));
            return {r * std::cos(th), r * std::sin(th)};
        }
        return {radius * std::cos(th), radius * std::sin(th)};
    };
    if (p.shape == CapsuleShape::CIRCLE) {
        for (int i = 0; i < n_; ++i) {
            Vec2d r = shape_pt(2.0 * PI * i / n_);
            pos_[i] = Vec2d{center.x + r.x, center.y + r.y};
        }
    } else {
        const int M = 256 * n_;
        std::vector<Vec2d> pts(M + 1);
        std::vector<Real> cum(M + 1, 0.0);
        for (int m = 0; m <= M; ++m) {
            pts[m] = shape_pt(2.0 * PI * m / M);
            if (m > 0) cum[m] = cum[m - 1] + (pts[m] - pts[m - 1]).norm();
        }
        const Real total = cum[M];
        int m = 0;
        for (int i = 0; i < n_; ++i) {
            const Real target = total * i / n_;
            while (m < M - 1 && cum[m + 1] < target) ++m;
            const Real seg = cum[m + 1] - cum[m];
            const Real w = (seg > 1e-15) ? (target - cum[m]) / seg : 0.0;
            Vec2d r = pts[m] + (pts[m + 1] - pts[m]) * w;
            pos_[i] = Vec2d{center.x + r.x, center.y + r.y};
        }
    }
    for (int i = 0; i < n_; ++i) rest_len_[i] = (pos_[(i + 1) % n_] - pos_[i]).norm();
    prev_len_ = rest_len_;
    has_prev_ = false;
    rest_area_ = area();
    rest_perimeter_ = perimeter();
    wlc_kBT_p_ = p.wlc_kBT_p;
    wlc_k_pow_ = p.wlc_k_pow;
    if (p.model == MembraneModel::WLC && n_ > 0) {
        Real L0 = rest_len_[0];
        Real Lmax = p.wlc_L_max_ratio * L0;
        Real x0 = L0 / Lmax;
        if (wlc_kBT_p_ <= 0.0) {
            Real denom = 1.0 / (2.0 * (1.0 - x0) * (1.0 - x0) * (1.0 - x0)) + 1.0;
            wlc_kBT_p_ = p.G_s * Lmax / (L0 * denom);
        }
        if (wlc_k_pow_ <= 0.0) {
            Real Fr = wlc_kBT_p_ * (1.0 / (4.0 * (1.0 - x0) * (1.0 - x0)) - 0.25 + x0);
            wlc_k_pow_ = Fr * L0;
        }
    }
    if (p.use_helfrich_bending) {
        rest_angle_.resize(n_);
        rest_ds_.resize(n_);
        for (int i = 0; i < n_; ++i) {
            Vec2d ep = minImageDiff(pos_[i], pos_[(i - 1 + n_) % n_]);
            Vec2d en = minImageDiff(pos_[(i + 1) % n_], pos_[i]);
            rest_ds_[i] = 0.5 * (ep.norm() + en.norm());
            rest_angle_[i] = (p.kappa_0 != 0.0) ? p.kappa_0 * rest_ds_[i] : std::atan2(ep.cross(en), ep.dot(en));
        }
    }
}

Vec2d Capsule::minImageDiff(const Vec2d& a, const Vec2d& b) const {
    Vec2d d = a - b;
    if (periodic_nx_ > 0) {
        Real Lx = static_cast<Real>(periodic_nx_);
        if (d.x > 0.5 * Lx) d.x -= Lx;
        if (d.x < -0.5 * Lx) d.x += Lx;
    }
    return d;
}

void Capsule::setNodeVelocity(int k, const Vec2d& v) {
    vel_.at(k) = v;
    if (sys_) sys_->dirty_ = true;
}

Vec2d Capsule::centroid() const {                    // capsule.cpp:159-173
    if (periodic_nx_ > 0 && n_ > 0) {
        Vec2d ref = pos_[0], s = ref;
        for (int i = 1; i < n_; ++i) s = s + (ref + minImageDiff(pos_[i], ref));
        return s / static_cast<Real>(n_);
    }
    Vec2d c{0.0, 0.0};
    for (int i = 0; i < n_; ++i) c = c + pos_[i];
    return c / static_cast<Real>(n_);
}

Real Capsule::area() const {                         // capsule.cpp:175-195
    Real A = 0.0;
    if (periodic_nx_ > 0 && n_ > 0) {
        std::vector<Vec2d> u(n_);
        u[0] = pos_[0];
        for (int i = 1; i < n_; ++i) u[i] = u[i - 1] + minImageDiff(pos_[i], pos_[i - 1]);
        for (int i = 0; i < n_; ++i) A += u[i].cross(u[(i + 1) % n_]);
        return 0.5 * std::abs(A);
    }
    for (int i = 0; i < n_; ++i) A += pos_[i].cross(pos_[(i + 1) % n_]);
    return 0.5 * std::abs(A);
}

Real Capsule::perimeter() const {                    // capsule.cpp:197-204
    Real P = 0.0;
    for (int i = 0; i < n_; ++i) P += minImageDiff(pos_[(i + 1) % n_], pos_[i]).norm();
    return P;
}

void Capsule::secondMoments(Real& mxx, Real& myy, Real& mxy) const {   // capsule.cpp:210-246
    std::vector<Vec2d> p(n_);
    p[0] = pos_[0];
    for (int i = 1; i < n_; ++i) p[i] = p[i - 1] + minImageDiff(pos_[i], pos_[i - 1]);
    Real A = 0.0, cx = 0.0, cy = 0.0;
    for (int i = 0; i < n_; ++i) {
        const Vec2d& a = p[i];
        const Vec2d& b = p[(i + 1) % n_];
        Real cr = a.x * b.y - b.x * a.y;
        A += cr; cx += (a.x + b.x) * cr; cy += (a.y + b.y) * cr;
    }
    A *= 0.5;
    mxx = myy = mxy = 0.0;
    if (std::abs(A) < 1e-15) return;
    cx /= (6.0 * A);
    cy /= (6.0 * A);
    for (int i = 0; i < n_; ++i) {
        Real x0 = p[i].x - cx, y0 = p[i].y - cy;
        const Vec2d& q = p[(i + 1) % n_];
        Real x1 = q.x - cx, y1 = q.y - cy;
        Real cr = x0 * y1 - x1 * y0;
        mxx += (x0 * x0 + x0 * x1 + x1 * x1) * cr;
        myy += (y0 * y0 + y0 * y1 + y1 * y1) * cr;
        mxy += (x0 * y1 + 2.0 * x0 * y0 + 2.0 * x1 * y1 + x1 * y0) * cr;
    }
    const Real sg = (A > 0.0) ? 1.0 : -1.0;
    mxx *= sg / 12.0; myy *= sg / 12.0; mxy *= sg / 24.0;
}

Real Capsule::deformationIndex() const {             // capsule.cpp:248-257
    if (n_ < 3) return 0.0;
    Real mxx, myy, mxy;
    secondMoments(mxx, myy, mxy);
    Real tr = 0.5 * (mxx + myy);
    Real dd = std::sqrt(0.25 * (mxx - myy) * (mxx - myy) + mxy * mxy);
    Real J1 = tr + dd, J2 = std::max(tr - dd, 0.0);
    Real s1 = std::sqrt(J1), s2 = std::sqrt(J2);
    return (s1 + s2 > 1e-15) ? (s1 - s2) / (s1 + s2) : 0.0;
}

Real Capsule::inclinationAngle() const {             // capsule.cpp:259-265
    if (n_ < 3) return 0.0;
    Real mxx, myy, mxy;
    secondMoments(mxx, myy, mxy);
    return 0.5 * std::atan2(2.0 * mxy, mxx - myy);
}

sf_capsule_desc Capsule::desc() const {
    sf_capsule_desc d{};
    d.model = static_cast<int>(params_.model);
    d.use_helfrich = params_.use_helfrich_bending ? 1 : 0;
    d.is_rigid = params_.is_rigid ? 1 : 0;
    d.type = type_;
    d.k_stretch = params_.k_stretch; d.G_s = params_.G_s; d.C_skalak = params_.C_skalak;
    d.k_bend = params_.k_bend; d.k_area = params_.k_area; d.k_perimeter = params_.k_perimeter;
    d.gamma_visc = params_.gamma_visc; d.eta_membrane = params_.eta_membrane;
    d.wlc_L_max_ratio = params_.wlc_L_max_ratio; d.wlc_kBT_p = wlc_kBT_p_; d.wlc_k_pow = wlc_k_pow_;
    d.rest_area = rest_area_; d.rest_perimeter = rest_perimeter_;
    return d;
}

// ============================================================================
void CapsuleSystem::addCapsule(Vec2d center, Real radius, int num_nodes, const MembraneParams& p,
                               int type) {
    pull();
    caps_.emplace_back(next_id_++, center, radius, num_nodes, p, type);
    if (periodic_nx_ > 0) caps_.back().setPeriodicX(periodic_nx_);
    for (auto& c : caps_) c.sys_ = this;
    dirty_ = true;
}
void CapsuleSystem::setPeriodicX(int nx) {
    periodic_nx_ = nx;
    for (auto& c : caps_) c.setPeriodicX(nx);
}
int CapsuleSystem::totalNodes() const {
    int t = 0;
    for (const auto& c : caps_) t += c.numNodes();
    return t;
}
void CapsuleSystem::pull() {
    if (stale_ && sim_) sim_->pullNodes(*this);
}

// ============================================================================
AdvectionDiffusion::AdvectionDiffusion(int nx, int ny, const ScalarParams& p)
    : nx_(nx), ny_(ny), nsp_(p.n_species) {
    g_.assign(nsp_, std::vector<Real>(static_cast<size_t>(nx) * ny * Q, 0.0));
    C_.assign(nsp_, std::vector<Real>(static_cast<size_t>(nx) * ny, 0.0));
}
void AdvectionDiffusion::pull() {
    if (stale_ && sim_) {
        sim_->pullScalar(*this);
        refreshShadows();
    }
}

Real* AdvectionDiffusion::leaseView(int species) {
    if (species < 0 || species >= nsp_) return nullptr;
    pull();
    if (leases_.empty()) { leases_.assign(nsp_, 0); shadow_.assign(nsp_, {}); }
    if (leases_[species]++ == 0) shadow_[species] = C_[species];
    return C_[species].data();
}

void AdvectionDiffusion::releaseView(int species) {
    if (leases_.empty() || species < 0 || species >= nsp_ || leases_[species] == 0) return;
    if (!stale_ && C_[species] != shadow_[species]) dirty_ = true;
    if (--leases_[species] == 0) std::vector<Real>().swap(shadow_[species]);
}

bool AdvectionDiffusion::anyLease() const {
    for (int n : leases_)
        if (n) return true;
    return false;
}

// A write into a concentration view sets C; the populations keep theirs (the
// reference's collision relaxes g towards w C(1 + 3 e.u) with the array's C).
void AdvectionDiffusion::collectViewWrites() {
    if (stale_) return;
    for (size_t s = 0; s < leases_.size(); ++s) {
        if (!leases_[s] || C_[s] == shadow_[s]) continue;
        dirty_ = true;
        shadow_[s] = C_[s];
    }
}

void AdvectionDiffusion::refreshShadows() {
    for (size_t s = 0; s < leases_.size(); ++s)
        if (leases_[s]) shadow_[s] = C_[s];
}
void AdvectionDiffusion::initialize(Real c0) {        // advection_diffusion.cpp:46-57
    stale_ = false;
    for (int s = 0; s < nsp_; ++s)
        for (size_t n = 0; n < C_[s].size(); ++n) {
            C_[s][n] = c0;
            for (int q = 0; q < Q; ++q) g_[s][n * Q + q] = W[q] * c0;
        }
    dirty_ = true;
}
void AdvectionDiffusion::setRegion(int x0, int y0, int x1, int y1, Real c, int s) {   // :59-70
    if (s < 0 || s >= nsp_) return;
    pull();
    for (int y = std::max(0, y0); y < std::min(ny_, y1); ++y)
        for (int x = std::max(0, x0); x < std::min(nx_, x1); ++x) {
            const size_t n = static_cast<size_t>(y) * nx_ + x;
            C_[s][n] = c;
            for (int q = 0; q < Q; ++q) g_[s][n * Q + q] = W[q] * c;
        }
    dirty_ = true;
}
void AdvectionDiffusion::setPoint(int x, int y, Real c, int s) {                      // :72-80
    if (x < 0 || x >= nx_ || y < 0 || y >= ny_ || s < 0 || s >= nsp_) return;
    pull();
    const size_t n = static_cast<size_t>(y) * nx_ + x;
    C_[s][n] = c;
    for (int q = 0; q < Q; ++q) g_[s][n * Q + q] = W[q] * c;
    dirty_ = true;
}
void AdvectionDiffusion::setField(const Real* C, int s) {
    if (s < 0 || s >= nsp_) return;
    pull();
    for (size_t n = 0; n < C_[s].size(); ++n) {
        C_[s][n] = C[n];
        for (int q = 0; q < Q; ++q) g_[s][n * Q + q] = W[q] * C[n];
    }
    dirty_ = true;
}
Real AdvectionDiffusion::getConcentration(int x, int y, int s) {
    if (x < 0 || x >= nx_ || y < 0 || y >= ny_ || s < 0 || s >= nsp_) return 0.0;
    pull();
    return C_[s][static_cast<size_t>(y) * nx_ + x];
}
Real* AdvectionDiffusion::concentrationData(int s) {
    if (s < 0 || s >= nsp_) return nullptr;
    pull();
    return C_[s].data();
}

// ============================================================================
Simulation::Simulation(const SimulationParams& p) : params_(p) {
    if (p.nx < 3 || p.ny < 3) throw std::invalid_argument("Simulation: domain must be at least 3x3");
    field_ = std::make_unique<LatticeField>(p.nx, p.ny);
    field_->sim_ = this;
    caps_.sim_ = this;
    if (p.fluid.boundary_type == BoundaryType::PERIODIC) caps_.setPeriodicX(p.nx);   // simulation.cpp:129-133
    if (p.scalar.enabled) {
        adr_ = std::make_unique<AdvectionDiffusion>(p.nx, p.ny, p.scalar);
        adr_->sim_ = this;
    }
    out_ = std::make_unique<OutputManager>(p, adr_ ? adr_->getNumSpecies() : 0);
}

Simulation::~Simulation() {
    out_.reset();            // the writer thread waits on the context's copies
    if (ctx_) sf_destroy(ctx_);
}

void Simulation::output(int64_t step) {
    if (out_ && out_->due(step)) out_->capture(step, caps_.caps_, field_->flags_);
}

void Simulation::finalize() {
    if (out_ && ctx_) out_->finalize();
}

void Simulation::setOutputConfig(const OutputConfig& config) { out_->setOutputConfig(config); }

void Simulation::setVTKOutput(int dump_every, const std::string& format, const FluidOutputFields& ff,
                              const ParticleVTKFields& pf) {
    params_.vtk_dump_every = dump_every;
    out_->setVTKOutput(dump_every, format, ff, pf);
}

void Simulation::setCSVOutput(int dump_every, const std::string& format, const ParticleOutputFields& fields,
                              const ParticleFilter& filter, bool append) {
    params_.csv_dump_every = dump_every;
    out_->setCSVOutput(dump_every, format, fields, filter, append);
}

void Simulation::addExtraCSVOutput(const std::string& filename, int dump_every, const ParticleOutputFields& fields,
                                   const ParticleFilter& filter) {
    (void)dump_every;   // the extra writers follow the CSV interval (simulation.cpp:1015-1023)
    out_->addExtraCSVOutput(filename, fields, filter);
}

void Simulation::addFluidProbe(int i, int j, const std::string& label) { out_->addFluidProbe(i, j, label); }

void Simulation::check(int rc, const char* what) const {
    if (rc != SF_OK) throw std::runtime_error(std::string(what) + ": " + sf_last_error());
}

CircleObstacle::CircleObstacle(Real cx, Real cy, Real radius) {
    o_.kind = sf::OBS_CIRCLE;
    o_.p[0] = cx; o_.p[1] = cy; o_.p[2] = radius;
}

RectObstacle::RectObstacle(Real x0, Real y0, Real x1, Real y1) {   // rect_obstacle.h:10-12
    o_.kind = sf::OBS_RECT;
    o_.p[0] = std::min(x0, x1); o_.p[1] = std::min(y0, y1);
    o_.p[2] = std::max(x0, x1); o_.p[3] = std::max(y0, y1);
}

PolygonObstacle::PolygonObstacle(const std::vector<Vec2d>& vertices) {
    if (vertices.size() < 3) throw std::invalid_argument("PolygonObstacle: at least 3 vertices");
    o_.kind = sf::OBS_POLYGON;
    o_.nv = static_cast<int>(vertices.size());
    for (const Vec2d& v : vertices) { verts_.push_back(v.x); verts_.push_back(v.y); }
}

std::vector<double> Obstacle::coords() const {
    if (o_.kind == sf::OBS_CIRCLE) return {o_.p[0], o_.p[1], o_.p[2]};
    if (o_.kind == sf::OBS_RECT) return {o_.p[0], o_.p[1], o_.p[2], o_.p[3]};
    return verts_;
}

std::vector<int32_t> ChannelBuilder::cellTypes(int nx, int ny, BoundaryType bt) const {
    std::vector<int32_t> fl(static_cast<size_t>(nx) * ny);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            CellType ct = CellType::FLUID;
            if (bottom_ && y == 0) ct = CellType::SOLID;
            if (top_ && y == ny - 1) ct = CellType::SOLID;
            if (bt == BoundaryType::INLET_OUTLET) {
                if (x == 0 && ct == CellType::FLUID) ct = CellType::INLET;
                if (x == nx - 1 && ct == CellType::FLUID) ct = CellType::OUTLET;
            }
            for (const auto& o : obstacles_)
                if (o->contains(static_cast<Real>(x), static_cast<Real>(y))) { ct = CellType::SOLID; break; }
            fl[static_cast<size_t>(y) * nx + x] = static_cast<int32_t>(ct);
        }
    return fl;
}

void Simulation::setChannelBuilder(const ChannelBuilder& b) { builder_ = std::make_unique<ChannelBuilder>(b); }

void Simulation::addCapsule(Vec2d center, Real radius, int num_nodes, const MembraneParams& mp, int type) {
    caps_.addCapsule(center, radius, num_nodes, mp, type);
}

// Simulation::addCapsuleRandom: engine/simulation.cpp:220-274
int Simulation::addCapsuleRandom(int count, Real x0, Real y0, Real x1, Real y1, Real rmin, Real rmax,
                                 int num_nodes, const MembraneParams& mp, int type, int seed,
                                 Real min_gap, int max_attempts) {
    std::mt19937 rng(static_cast<unsigned>(seed));
    std::uniform_real_distribution<Real> dx(x0, x1), dy(y0, y1), dr(rmin, rmax);
    int placed = 0, attempts = 0;
    const int total = (max_attempts > 0) ? max_attempts : count * 500;
    while (placed < count && attempts < total) {
        Real cx = dx(rng);
        Real cy = dy(rng);
        Real r = dr(rng);
        bool overlap = false;
        for (int c = 0; c < caps_.numCapsules(); ++c) {
            Vec2d oc = caps_[c].centroid();
            Real orr = caps_[c].effectiveRadius();
            if ((Vec2d{cx, cy} - oc).norm() < r + orr + min_gap) { overlap = true; break; }
        }
        if (builder_) {
            if (cy - r < builder_->wallBottom() + min_gap || cy + r > builder_->wallTop() - min_gap) overlap = true;
            for (const auto& o : builder_->getObstacles())
                if (o->signedDistance(cx, cy) < r + min_gap) { overlap = true; break; }
        }
        if (!overlap) {
            int nn = (num_nodes > 0) ? num_nodes : Capsule::computeOptimalNodes(r);
            caps_.addCapsule({cx, cy}, r, nn, mp, type);
            placed++;
        }
        attempts++;
    }
    std::cout << "Placed " << placed << "/" << count << " capsules (" << attempts << " attempts)" << std::endl;
    return placed;
}

void Simulation::initialize() {
    const SimulationParams& P = params_;
    // features of SimulationParams outside the device path
    if ((P.fluid.gravity_x != 0.0 || P.fluid.gravity_y != 0.0))
        throw std::invalid_argument("Simulation: gravity is not implemented on the MI355X path");
    if (P.shan_chen.enabled)
        throw std::invalid_argument("Simulation: Shan-Chen multiphase is not implemented on the MI355X path");
    if (P.lubrication.enabled || P.adhesion.enabled)
        throw std::invalid_argument("Simulation: lubrication/adhesion are not implemented on the MI355X path");

    sf_params sp{};
    sp.nx = P.nx; sp.ny = P.ny; sp.dt = P.dt; sp.ibm_iterations = P.ibm_iterations;
    sp.tau = P.fluid.tau; sp.rho0 = P.fluid.rho0;
    sp.boundary_type = static_cast<int>(P.fluid.boundary_type);
    sp.body_force_x = P.fluid.body_force_x; sp.body_force_y = P.fluid.body_force_y;
    // lbm_solver.cpp:22-24: regularized wins over MRT, MRT over BGK
    int model = 0;
    if (P.fluid.collision_model == CollisionModel::REGULARIZED) model = SF_REGULARIZED;
    else if (P.fluid.use_mrt || P.fluid.collision_model == CollisionModel::MRT) model = SF_MRT;
    sp.collision_model = model;
    sp.mrt_se = P.fluid.mrt_se; sp.mrt_s_eps = P.fluid.mrt_s_eps; sp.mrt_sq = P.fluid.mrt_sq;
    sp.periodic_y = P.fluid.periodic_y ? 1 : 0;
    sp.max_lattice_force = P.max_lattice_force;
    sp.capsule_periodic_x = P.fluid.capsule_periodic_x ? 1 : 0;
    sp.rep_epsilon = P.repulsion.epsilon; sp.rep_sigma = P.repulsion.sigma; sp.rep_r_cut = P.repulsion.r_cut;
    sp.rep_power = P.repulsion.power; sp.rep_damping_normal = P.repulsion.damping_normal;
    sp.rep_friction_coeff = P.repulsion.friction_coeff;
    sp.scalar_enabled = P.scalar.enabled ? 1 : 0;
    sp.n_species = P.scalar.n_species;
    for (int s = 0; s < std::min(4, P.scalar.n_species); ++s) {
        sp.diffusivity[s] = s < static_cast<int>(P.scalar.diffusivity.size()) ? P.scalar.diffusivity[s] : 0.01;
        sp.inlet_concentration[s] = s < static_cast<int>(P.scalar.inlet_concentration.size()) ? P.scalar.inlet_concentration[s] : 0.0;
    }
    sp.scalar_periodic_y = P.scalar.periodic_y ? 1 : 0;
    sp.precision = static_cast<int>(P.precision);
    sp.top_wall_velocity = P.fluid.top_wall_velocity;
    sp.bottom_wall_velocity = P.fluid.bottom_wall_velocity;
    sp.inlet_velocity = P.fluid.inlet_velocity;
    sp.outlet_density = P.fluid.outlet_density;
    // links are only built from a channel builder's geometry (simulation.cpp:440-442)
    sp.use_interpolated_bb = (P.fluid.use_interpolated_bb && builder_) ? 1 : 0;

    if (ctx_) {
        out_->attach(nullptr);
        sf_destroy(ctx_);
        ctx_ = nullptr;
    }
    check(sf_create(&ctx_, P.device, &sp), "sf_create");
    out_->attach(ctx_);
    check(sf_set_graph_mode(ctx_, P.use_graph ? 1 : 0), "sf_set_graph_mode");

    // cell types: the channel builder's (simulation.cpp:431-434), then
    // LBMSolver::initialize's rows and columns over them (lbm_solver.cpp:130-158)
    LatticeField& F = *field_;
    const int nx = P.nx, ny = P.ny;
    const int SOLID = static_cast<int>(CellType::SOLID), FLUID = static_cast<int>(CellType::FLUID);
    if (builder_) F.flags_ = builder_->cellTypes(nx, ny, P.fluid.boundary_type);
    else std::fill(F.flags_.begin(), F.flags_.end(), FLUID);
    auto at = [&](int x, int y) -> int32_t& { return F.flags_[static_cast<size_t>(y) * nx + x]; };
    for (int x = 0; x < nx; ++x) at(x, 0) = at(x, ny - 1) = SOLID;
    if (P.fluid.boundary_type == BoundaryType::INLET_OUTLET) {
        for (int y = 1; y < ny - 1; ++y) {
            at(0, y) = static_cast<int>(CellType::INLET);
            at(nx - 1, y) = static_cast<int>(CellType::OUTLET);
        }
    } else if (P.fluid.boundary_type == BoundaryType::CLOSED) {
        for (int y = 0; y < ny; ++y) at(0, y) = at(nx - 1, y) = SOLID;
    }
    if (P.fluid.periodic_y)
        for (int x = 0; x < nx; ++x) at(x, 0) = at(x, ny - 1) = FLUID;
    check(sf_set_flags(ctx_, F.flags_.data()), "sf_set_flags");
    if (builder_ && !builder_->getObstacles().empty()) {
        std::vector<int32_t> kinds, offs{0};
        std::vector<double> coords;
        for (const auto& o : builder_->getObstacles()) {
            kinds.push_back(o->kind());
            const std::vector<double> c = o->coords();
            coords.insert(coords.end(), c.begin(), c.end());
            offs.push_back(static_cast<int32_t>(coords.size()));
        }
        check(sf_set_obstacles(ctx_, static_cast<int>(kinds.size()), kinds.data(), offs.data(), coords.data()),
              "sf_set_obstacles");
    }
    // equilibrium at rest (lattice_field.cpp:75-90)
    F.stale_ = false;
    for (int n = 0; n < F.N_; ++n) {
        F.rho_[n] = P.fluid.rho0; F.ux_[n] = 0.0; F.uy_[n] = 0.0;
        F.Fx_[n] = 0.0; F.Fy_[n] = 0.0;
        for (int q = 0; q < Q; ++q) F.f_[static_cast<size_t>(q) * F.N_ + n] = feq(q, P.fluid.rho0, 0.0, 0.0);
    }
    check(sf_set_populations(ctx_, F.f_.data()), "sf_set_populations");
    F.f_dirty_ = F.macro_dirty_ = F.flags_dirty_ = false;
    F.stale_ = true;   // rho/u as the device computes them (computeMacroscopic with F = 0)
    if (adr_) adr_->initialize(0.0);
    if (P.fluid.boundary_type == BoundaryType::PERIODIC) caps_.setPeriodicX(P.nx);
    caps_.dirty_ = true;
    step_count_ = 0;
    pushDirty();
    check(sf_mdf_set_control(ctx_, mdf_tol_, mdf_max_, mdf_probe_ ? 1 : 0), "sf_mdf_set_control");
    check(sf_mdf_set_tail(ctx_, mdf_tail_), "sf_mdf_set_tail");
    if (mdf_collect_ > 0) check(sf_mdf_collect(ctx_, mdf_collect_), "sf_mdf_collect");
    if (mdf_model_) check(sf_mdf_set_model(ctx_, mdf_model_->handle(), mdf_warm_ ? 1 : 0), "sf_mdf_set_model");
    std::cout << "Simulation initialized: " << nx << "x" << ny << " lattice, " << caps_.caps_.size()
              << " capsules (" << caps_.totalNodes() << " membrane nodes)" << std::endl;
}

void Simulation::pushDirty() {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    LatticeField& F = *field_;
    F.collectViewWrites();
    if (adr_) adr_->collectViewWrites();
    if (F.flags_dirty_) { check(sf_set_flags(ctx_, F.flags_.data()), "sf_set_flags"); F.flags_dirty_ = false; }
    if (F.f_dirty_) { check(sf_set_populations(ctx_, F.f_.data()), "sf_set_populations"); }
    if (F.f_dirty_ || F.macro_dirty_) {
        // the next collision reads the stored rho/u, as the reference's does
        check(sf_set_macroscopic(ctx_, F.rho_.data(), F.ux_.data(), F.uy_.data()), "sf_set_macroscopic");
    }
    F.f_dirty_ = F.macro_dirty_ = false;
    if (caps_.dirty_) {
        auto& cs = caps_.caps_;
        const int nc = static_cast<int>(cs.size());
        std::vector<int32_t> off(nc + 1, 0), hp(nc, 0);
        std::vector<sf_capsule_desc> desc(nc);
        for (int c = 0; c < nc; ++c) {
            off[c + 1] = off[c] + cs[c].n_;
            desc[c] = cs[c].desc();
            hp[c] = cs[c].has_prev_ ? 1 : 0;
        }
        const int K = off[nc];
        std::vector<double> x(K), y(K), vx(K), vy(K), L0(K), th(K, 0.0), ds(K, 0.0), Lp(K);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < cs[c].n_; ++i) {
                const int k = off[c] + i;
                x[k] = cs[c].pos_[i].x; y[k] = cs[c].pos_[i].y;
                vx[k] = cs[c].vel_[i].x; vy[k] = cs[c].vel_[i].y;
                L0[k] = cs[c].rest_len_[i]; Lp[k] = cs[c].prev_len_[i];
                if (!cs[c].rest_angle_.empty()) { th[k] = cs[c].rest_angle_[i]; ds[k] = cs[c].rest_ds_[i]; }
            }
        check(sf_set_capsules(ctx_, nc, off.data(), desc.data(), x.data(), y.data(), vx.data(), vy.data(),
                              L0.data(), th.data(), ds.data(), Lp.data(), hp.data()),
              "sf_set_capsules");
        caps_.dirty_ = false;
    }
    if (adr_ && adr_->dirty_) {
        for (int s = 0; s < adr_->nsp_; ++s)
            check(sf_set_scalar(ctx_, s, adr_->g_[s].data(), adr_->C_[s].data()), "sf_set_scalar");
        adr_->dirty_ = false;
    }
    if (!pending_force_.empty()) {
        // after the stored macroscopic field: the next collision uses this force instead
        check(sf_set_force_state(ctx_, pending_force_.data()), "sf_set_force_state");
        pending_force_.clear();
    }
}

void Simulation::pullField(LatticeField& F) {
    if (!ctx_) return;
    F.stale_ = false;
    check(sf_get_populations(ctx_, F.f_.data()), "sf_get_populations");
    check(sf_get_macroscopic(ctx_, F.rho_.data(), F.ux_.data(), F.uy_.data(), F.Fx_.data(), F.Fy_.data()),
          "sf_get_macroscopic");
}

void Simulation::pullNodes(CapsuleSystem& cs) {
    cs.stale_ = false;
    if (!ctx_ || cs.caps_.empty()) return;
    const int K = cs.totalNodes();
    const int nc = static_cast<int>(cs.caps_.size());
    std::vector<double> x(K), y(K), vx(K), vy(K), fx(K), fy(K), Lp(K);
    std::vector<int32_t> hp(nc);
    check(sf_get_nodes(ctx_, x.data(), y.data(), vx.data(), vy.data(), fx.data(), fy.data()), "sf_get_nodes");
    check(sf_get_membrane_state(ctx_, Lp.data(), hp.data()), "sf_get_membrane_state");
    int k = 0;
    for (int c = 0; c < nc; ++c) {
        Capsule& cp = cs.caps_[c];
        for (int i = 0; i < cp.n_; ++i, ++k) {
            cp.pos_[i] = {x[k], y[k]};
            cp.vel_[i] = {vx[k], vy[k]};
            cp.frc_[i] = {fx[k], fy[k]};
            cp.prev_len_[i] = Lp[k];
        }
        cp.has_prev_ = hp[c] != 0;
    }
}

void Simulation::pullScalar(AdvectionDiffusion& a) {
    a.stale_ = false;
    if (!ctx_) return;
    for (int s = 0; s < a.nsp_; ++s) check(sf_get_scalar(ctx_, s, a.C_[s].data(), a.g_[s].data()), "sf_get_scalar");
}

// After the device advanced: host mirrors are stale; a held field view is
// refreshed at once so it keeps showing the live state.
void Simulation::afterAdvance() {
    field_->stale_ = true;
    caps_.stale_ = true;
    if (adr_) adr_->stale_ = true;
    if (field_->anyLease()) field_->pull();
    if (adr_ && adr_->anyLease()) adr_->pull();
}

void Simulation::step() {
    pushDirty();
    check(sf_step(ctx_, 1), "sf_step");
    afterAdvance();
    if (params_.enable_stability_checks && params_.stability_check_interval > 0 &&
        step_count_ % params_.stability_check_interval == 0)
        checkStability();
    output(step_count_);
    intervalCheckpoint(step_count_);
    if (cb_) cb_(*this, static_cast<int>(step_count_));
    step_count_++;
}

void Simulation::markAdvanced(int steps) {
    afterAdvance();
    step_count_ += steps;
}

void Simulation::run(int num_steps) {
    if (cb_) {
        for (int s = 0; s < num_steps; ++s) step();
        finalize();
        return;
    }
    int done = 0;
    while (done < num_steps) {
        // advance in chunks that end where the reference checks stability
        int chunk = num_steps - done;
        const int iv = params_.stability_check_interval;
        const bool checks = params_.enable_stability_checks && iv > 0;
        if (checks) {
            const int64_t r = step_count_ % iv;
            const int to_check = static_cast<int>(r == 0 ? 1 : iv - r + 1);
            chunk = std::min(chunk, to_check);
        }
        const int civ = params_.checkpoint_interval;
        if (civ > 0) {
            // end the chunk on the next step whose index is a multiple of the interval
            const int64_t r = step_count_ % civ;
            chunk = std::min(chunk, static_cast<int>(r == 0 ? 1 : civ - r + 1));
        }
        // and on the next step index at which a writer is due
        const int64_t due = out_->nextDue(step_count_);
        if (due >= 0) chunk = std::min(chunk, static_cast<int>(due - step_count_ + 1));
        pushDirty();
        check(sf_step(ctx_, chunk), "sf_step");
        afterAdvance();
        step_count_ += chunk;
        if (checks && (step_count_ - 1) % iv == 0) checkStability();
        output(step_count_ - 1);
        intervalCheckpoint(step_count_ - 1);
        done += chunk;
    }
    check(sf_synchronize(ctx_), "sf_synchronize");   // capacity overflows of the last calls
    finalize();
}

Real Simulation::getMaxSpeed() {
    pushDirty();
    double s[5];
    check(sf_field_stats(ctx_, s), "sf_field_stats");
    return s[3];
}

std::vector<double> Simulation::fieldStats() {
    pushDirty();
    std::vector<double> s(5);
    check(sf_field_stats(ctx_, s.data()), "sf_field_stats");
    return s;
}

int Simulation::lastCappedNodes() {
    int32_t c = 0;
    check(sf_last_capped(ctx_, &c), "sf_last_capped");
    return c;
}

// simulation.cpp:820-890, from FP64 device reductions (no copy of the field)
bool Simulation::checkStability() {
    pushDirty();
    const Real rho0 = params_.fluid.rho0;
    double g[11];
    check(sf_global_stats(ctx_, rho0, g), "sf_global_stats");
    const bool nan = g[7] > 0.0;
    const bool dwarn = std::abs(g[8] - rho0) / rho0 > params_.max_density_deviation;
    const Real worst_vel = std::sqrt(g[4]);
    const bool vwarn = worst_vel > params_.max_velocity;
    if (nan)
        std::cerr << "*** STABILITY WARNING: NaN/Inf detected at step " << step_count_ << " in "
                  << static_cast<long long>(g[7]) << " cells" << std::endl;
    if (dwarn)
        std::cerr << "*** STABILITY WARNING: density deviation at step " << step_count_ << " - rho = " << g[8]
                  << " (rho0 = " << rho0 << ", deviation = " << std::abs(g[8] - rho0) / rho0 * 100.0 << "%)"
                  << std::endl;
    if (vwarn)
        std::cerr << "*** STABILITY WARNING: high velocity at step " << step_count_ << " - |u| = " << worst_vel
                  << " (limit = " << params_.max_velocity << ", Ma = " << worst_vel / std::sqrt(1.0 / 3.0) << ")"
                  << std::endl;
    const int capped = lastCappedNodes();
    if (capped > 0)
        std::cerr << "*** STABILITY NOTICE: max_lattice_force cap rescaled " << capped << " fluid nodes at step "
                  << step_count_ << " (cap = " << params_.max_lattice_force << ")" << std::endl;
    return !(nan || dwarn || vwarn);
}

std::vector<std::pair<std::string, double>> Simulation::globalStats() {
    pushDirty();
    double g[11];
    check(sf_global_stats(ctx_, params_.fluid.rho0, g), "sf_global_stats");
    const double n = g[0];
    std::vector<std::pair<std::string, double>> out;
    out.emplace_back("timestep", static_cast<double>(step_count_));
    out.emplace_back("time", static_cast<double>(step_count_) * params_.dt);
    out.emplace_back("mean_rho", n > 0 ? g[1] / n : 0.0);
    out.emplace_back("max_velocity", std::sqrt(g[4]));
    out.emplace_back("mean_velocity", n > 0 ? g[5] / n : 0.0);
    out.emplace_back("total_kinetic_energy", g[6]);
    std::vector<Capsule>& cs = caps_.capsules();
    double mean_def = 0.0, max_def = 0.0;
    for (Capsule& c : cs) {
        const double D = c.deformationIndex();
        mean_def += D;
        if (D > max_def) max_def = D;
    }
    if (!cs.empty()) mean_def /= static_cast<double>(cs.size());
    out.emplace_back("n_active_particles", static_cast<double>(cs.size()));
    out.emplace_back("mean_deformation", mean_def);
    out.emplace_back("max_deformation", max_def);
    out.emplace_back("min_rho", n > 0 ? g[2] : 1e30);
    out.emplace_back("max_rho", n > 0 ? g[3] : -1e30);
    if (adr_) {
        // every cell and species, as the writer sums them
        double mass = 0.0;
        for (int s = 0; s < adr_->nsp_; ++s) {
            const Real* C = adr_->concentrationData(s);
            for (int k = 0; k < params_.nx * params_.ny; ++k) mass += C[k];
        }
        out.emplace_back("total_dissolved_mass", mass);
        out.emplace_back("total_particle_mass", 0.0);
        out.emplace_back("total_adsorbed_mass", 0.0);
    }
    out.emplace_back("nonfinite_cells", g[7]);
    return out;
}

// LBMSolver::setFluidRegion: lbm_solver.cpp:339-350
void Simulation::setFluidRegion(int x0, int y0, int x1, int y1, Real rho, Real ux, Real uy) {
    LatticeField& F = *field_;
    F.touch_f();
    F.touch_macro();
    for (int y = std::max(0, y0); y < std::min(params_.ny, y1); ++y)
        for (int x = std::max(0, x0); x < std::min(params_.nx, x1); ++x) {
            if (F.getCellType(x, y) == CellType::SOLID) continue;
            F.setRho(x, y, rho); F.setUx(x, y, ux); F.setUy(x, y, uy);
            F.initializeEquilibriumAt(x, y);
        }
}

void Simulation::setScalarRegion(int x0, int y0, int x1, int y1, Real c, int species) {
    if (adr_) adr_->setRegion(x0, y0, x1, y1, c, species);
}

std::string Simulation::autotune(int reps) {
    pushDirty();
    std::vector<char> buf(4096, 0);
    check(sf_autotune(ctx_, reps, buf.data(), static_cast<int>(buf.size())), "sf_autotune");
    return std::string(buf.data());
}

void Simulation::setLbmLaunch(int bx, int by, bool xcd_banded, int persist) {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    check(sf_set_lbm_launch(ctx_, bx, by, xcd_banded ? 1 : 0, persist), "sf_set_lbm_launch");
}

void Simulation::setCapsuleLaunch(int spread_grid, int interp_lanes) {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    check(sf_set_capsule_launch(ctx_, spread_grid, interp_lanes), "sf_set_capsule_launch");
}

void Simulation::setLbmPairs(bool on) {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    check(sf_set_lbm_pairs(ctx_, on ? 1 : 0), "sf_set_lbm_pairs");
}

void Simulation::setMdfControl(Real tol, int max_corrections, bool probe) {
    mdf_tol_ = tol;
    mdf_max_ = max_corrections;
    mdf_probe_ = probe;
    if (ctx_) check(sf_mdf_set_control(ctx_, tol, max_corrections, probe ? 1 : 0), "sf_mdf_set_control");
}

void Simulation::setMdfTail(int first) {
    mdf_tail_ = first;
    if (ctx_) check(sf_mdf_set_tail(ctx_, first), "sf_mdf_set_tail");
}

std::vector<double> Simulation::mdfStats() {
    std::vector<double> out(4, 0.0);
    if (ctx_) check(sf_mdf_stats(ctx_, out.data()), "sf_mdf_stats");
    return out;
}

void Simulation::collectMdfSamples(int64_t capacity) {
    mdf_collect_ = capacity;
    if (ctx_) check(sf_mdf_collect(ctx_, capacity), "sf_mdf_collect");
}

int64_t Simulation::mdfSampleCount() {
    int64_t n = 0;
    if (ctx_) check(sf_mdf_samples(ctx_, &n, nullptr), "sf_mdf_samples");
    return n;
}

std::vector<float> Simulation::mdfSamples() {
    int64_t n = mdfSampleCount();
    std::vector<float> out(static_cast<size_t>(n) * (SF_MLP_IN + SF_MLP_OUT));
    if (n > 0) check(sf_mdf_samples(ctx_, &n, out.data()), "sf_mdf_samples");
    return out;
}

void Simulation::setMdfModel(MdfModel* model, bool warm) {
    if (warm && !model) throw std::invalid_argument("Simulation: warm start needs a model");
    mdf_model_ = model;
    mdf_warm_ = warm;
    if (ctx_) check(sf_mdf_set_model(ctx_, model ? model->handle() : nullptr, warm ? 1 : 0), "sf_mdf_set_model");
}

// ---------------------------------------------------------------------------
static void mlp_check(int rc, const char* what) {
    if (rc != SF_OK) throw std::runtime_error(std::string(what) + ": " + sf_last_error());
}

MdfModel::MdfModel(int device, uint64_t seed) { mlp_check(sf_mlp_create(&net_, device, seed), "sf_mlp_create"); }
MdfModel::~MdfModel() { if (net_) sf_mlp_destroy(net_); }

std::vector<double> MdfModel::train(Simulation& sim, int epochs, int batch, double lr, uint64_t seed) {
    if (!sim.deviceContext()) throw std::runtime_error("MdfModel: the simulation is not initialized");
    std::vector<double> loss(std::max(epochs, 0));
    mlp_check(sf_mlp_train(net_, sim.deviceContext(), epochs, batch, static_cast<float>(lr), seed,
                           loss.data()), "sf_mlp_train");
    return loss;
}

std::vector<float> MdfModel::params() const {
    int np = 0, nn = 0;
    sf_mlp_num_params(&np, &nn);
    std::vector<float> p(np);
    mlp_check(sf_mlp_get_params(net_, p.data(), nullptr), "sf_mlp_get_params");
    return p;
}

std::vector<float> MdfModel::normalization() const {
    int np = 0, nn = 0;
    sf_mlp_num_params(&np, &nn);
    std::vector<float> n(nn);
    mlp_check(sf_mlp_get_params(net_, nullptr, n.data()), "sf_mlp_get_params");
    return n;
}

void MdfModel::setParams(const std::vector<float>& p) {
    int np = 0, nn = 0;
    sf_mlp_num_params(&np, &nn);
    if (static_cast<int>(p.size()) != np) throw std::invalid_argument("MdfModel: wrong parameter count");
    mlp_check(sf_mlp_set_params(net_, p.data(), nullptr), "sf_mlp_set_params");
}

void MdfModel::setNormalization(const std::vector<float>& n) {
    int np = 0, nn = 0;
    sf_mlp_num_params(&np, &nn);
    if (static_cast<int>(n.size()) != nn) throw std::invalid_argument("MdfModel: wrong normalisation size");
    mlp_check(sf_mlp_set_params(net_, nullptr, n.data()), "sf_mlp_set_params");
}

std::vector<float> MdfModel::predict(const std::vector<float>& x) const {
    if (x.size() % SF_MLP_IN) throw std::invalid_argument("MdfModel: inputs must be n x 17");
    const int n = static_cast<int>(x.size() / SF_MLP_IN);
    std::vector<float> y(static_cast<size_t>(n) * SF_MLP_OUT);
    mlp_check(sf_mlp_predict(net_, n, x.data(), y.data()), "sf_mlp_predict");
    return y;
}

std::pair<double, std::vector<float>> MdfModel::lossGrad(const std::vector<float>& x,
                                                         const std::vector<float>& t) const {
    if (x.size() % SF_MLP_IN || t.size() != x.size() / SF_MLP_IN * SF_MLP_OUT)
        throw std::invalid_argument("MdfModel: inputs must be n x 17 and targets n x 2");
    const int n = static_cast<int>(x.size() / SF_MLP_IN);
    int np = 0, nn = 0;
    sf_mlp_num_params(&np, &nn);
    std::vector<float> g(np);
    double loss = 0.0;
    mlp_check(sf_mlp_loss_grad(net_, n, x.data(), t.data(), &loss, g.data()), "sf_mlp_loss_grad");
    return {loss, g};
}

void Simulation::setTiming(bool on) { if (ctx_) check(sf_set_timing(ctx_, on ? 1 : 0), "sf_set_timing"); }

void Simulation::synchronize() { if (ctx_) check(sf_synchronize(ctx_), "sf_synchronize"); }

int64_t Simulation::graphCaptures() {
    int64_t n = 0;
    if (ctx_) check(sf_graph_captures(ctx_, &n), "sf_graph_captures");
    return n;
}

std::vector<std::pair<std::string, double>> Simulation::kernelTimes() {
    std::vector<std::pair<std::string, double>> out;
    if (!ctx_) return out;
    const char* names[32];
    double ms[32];
    int n = 0;
    check(sf_kernel_times(ctx_, 32, names, ms, &n), "sf_kernel_times");
    for (int i = 0; i < std::min(n, 32); ++i) out.emplace_back(names[i], ms[i]);
    return out;
}

// ---------------------------------------------------------------------------
// Checkpoints in the reference's .sfck layout, version 4 (io/checkpoint.cpp:10-58
// documents it): header, f (SoA), rho, ux, uy, cell types, per capsule id,
// type, node count, positions, velocities, area, perimeter and periodic nx,
// adhesion bonds (none: adhesion is refused at initialize()), then the scalar
// block.  After it comes an extension block the reference's loader never reads:
// "SFXT", its version, the raw IBM force of the last step and, per capsule, the
// Kelvin-Voigt state.  A checkpoint without it restarts as the reference does:
// the next collision takes rho and u from the file.
namespace {

template <typename T>
void put(std::ofstream& o, const T& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <typename T>
void put_n(std::ofstream& o, const T* p, size_t n) { o.write(reinterpret_cast<const char*>(p), n * sizeof(T)); }
template <typename T>
bool get(std::ifstream& i, T& v) { i.read(reinterpret_cast<char*>(&v), sizeof(T)); return i.good(); }
template <typename T>
bool get_n(std::ifstream& i, T* p, size_t n) {
    i.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
    return i.good();
}
bool skip(std::ifstream& i, size_t bytes) {
    i.seekg(static_cast<std::streamoff>(bytes), std::ios::cur);
    return i.good();
}

constexpr int SFCK_VERSION = 4;
constexpr int SFXT_VERSION = 1;

}  // namespace

bool Simulation::saveCheckpoint(const std::string& filename) {
    return writeCheckpoint(filename, static_cast<int>(step_count_));
}

// Simulation::step writes checkpoint_<step>.sfck after the step whose index is
// a positive multiple of checkpoint_interval (simulation.cpp:773-779).
void Simulation::intervalCheckpoint(int64_t step) {
    const int iv = params_.checkpoint_interval;
    if (iv > 0 && step > 0 && step % iv == 0)
        writeCheckpoint(params_.output_dir + "/checkpoint_" + std::to_string(step) + ".sfck",
                        static_cast<int>(step));
}

bool Simulation::writeCheckpoint(const std::string& filename, int step) {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    pushDirty();
    std::ofstream out(filename, std::ios::binary);
    if (!out.is_open()) {
        std::cerr << "Checkpoint: cannot open " << filename << " for writing\n";
        return false;
    }
    LatticeField& fl = *field_;
    const int nx = params_.nx, ny = params_.ny;
    const size_t N = static_cast<size_t>(nx) * ny;
    std::vector<Capsule>& cs = caps_.capsules();
    const int ncaps = static_cast<int>(cs.size());
    out.write("SFCK", 4);
    put(out, SFCK_VERSION);
    put(out, step);
    put(out, nx);
    put(out, ny);
    put(out, ncaps);
    put_n(out, fl.fData(), 9 * N);
    put_n(out, fl.rhoData(), N);
    put_n(out, fl.uxData(), N);
    put_n(out, fl.uyData(), N);
    put_n(out, fl.cellTypeData(), N);
    for (Capsule& c : cs) {
        put(out, c.getId());
        put(out, c.getType());
        put(out, c.numNodes());
        for (const Vec2d& p : c.pos_) { put(out, p.x); put(out, p.y); }
        for (const Vec2d& v : c.vel_) { put(out, v.x); put(out, v.y); }
        put(out, c.area());
        put(out, c.perimeter());
        put(out, c.getPeriodicX());
    }
    put(out, 0);                                  // adhesion bonds
    const int has_scalar = adr_ ? 1 : 0;
    put(out, has_scalar);
    if (adr_) {
        put(out, adr_->nsp_);
        for (int s = 0; s < adr_->nsp_; ++s) {
            const Real* C = adr_->concentrationData(s);   // pulls g and C
            put_n(out, adr_->g_[s].data(), 9 * N);
            put_n(out, C, N);
        }
        put(out, 0);                              // capsule release / absorb trackers
        put(out, 0);                              // particle masses
        put(out, 0);                              // surface coverage
    }
    out.write("SFXT", 4);
    put(out, SFXT_VERSION);
    std::vector<double> F(2 * N);
    check(sf_get_force_state(ctx_, F.data()), "sf_get_force_state");
    put_n(out, F.data(), F.size());
    for (Capsule& c : cs) {
        put(out, c.has_prev_ ? 1 : 0);
        put_n(out, c.prev_len_.data(), c.prev_len_.size());
    }
    out.close();
    if (!out) {
        std::cerr << "Checkpoint: writing " << filename << " failed\n";
        return false;
    }
    std::cout << "Checkpoint saved: step " << step << ", " << ncaps << " capsules"
              << (has_scalar ? ", scalar state (v4)" : "") << " -> " << filename << std::endl;
    return true;
}

bool Simulation::loadCheckpoint(const std::string& filename) {
    if (!ctx_) throw std::runtime_error("Simulation: initialize() has not been called");
    std::ifstream in(filename, std::ios::binary);
    if (!in.is_open()) {
        std::cerr << "Checkpoint: cannot open " << filename << " for reading\n";
        return false;
    }
    auto bad = [&](const std::string& why) {
        std::cerr << "Checkpoint: " << why << " (" << filename << ")\n";
        return false;
    };
    char magic[4];
    in.read(magic, 4);
    if (!in.good() || std::strncmp(magic, "SFCK", 4) != 0) return bad("invalid magic (not a SoftFlow checkpoint)");
    int version = 0, step = 0, nx = 0, ny = 0, ncaps = 0;
    if (!get(in, version)) return bad("truncated header");
    if (version < 2 || version > 4) return bad("unsupported version " + std::to_string(version) + " (expected 2, 3, or 4)");
    if (version == 2)
        std::cerr << "Checkpoint: loading version 2 file - scalar transport state will be reset "
                     "to initial values (no scalar data in v2 files)\n";
    if (!get(in, step) || !get(in, nx) || !get(in, ny) || !get(in, ncaps)) return bad("truncated header");
    if (nx != params_.nx || ny != params_.ny)
        return bad("domain mismatch: file has " + std::to_string(nx) + "x" + std::to_string(ny));
    std::vector<Capsule>& cs = caps_.capsules();
    if (ncaps != static_cast<int>(cs.size()))
        return bad("capsule count mismatch: file has " + std::to_string(ncaps));
    // read everything first: a short or mismatched file changes nothing
    const size_t N = static_cast<size_t>(nx) * ny;
    std::vector<Real> f(9 * N), rho(N), ux(N), uy(N);
    std::vector<int> flags(N);
    if (!get_n(in, f.data(), f.size()) || !get_n(in, rho.data(), N) || !get_n(in, ux.data(), N) ||
        !get_n(in, uy.data(), N) || !get_n(in, flags.data(), N))
        return bad("truncated lattice state");
    std::vector<std::vector<Vec2d>> pos(ncaps), vel(ncaps);
    for (int c = 0; c < ncaps; ++c) {
        int id = 0, tid = 0, nn = 0, pnx = 0;
        if (!get(in, id) || !get(in, tid) || !get(in, nn)) return bad("truncated capsule record");
        if (nn != cs[c].numNodes())
            return bad("node count mismatch for capsule " + std::to_string(c) + ": file has " + std::to_string(nn));
        pos[c].resize(nn);
        vel[c].resize(nn);
        for (Vec2d& p : pos[c]) if (!get(in, p.x) || !get(in, p.y)) return bad("truncated capsule record");
        for (Vec2d& v : vel[c]) if (!get(in, v.x) || !get(in, v.y)) return bad("truncated capsule record");
        Real area = 0.0, perim = 0.0;
        if (!get(in, area) || !get(in, perim) || !get(in, pnx)) return bad("truncated capsule record");
    }
    int nbonds = 0;
    if (!get(in, nbonds)) return bad("truncated bond count");
    // bonds are {int, int, int, int, Real, Real}; adhesion is not on the device path
    if (nbonds > 0 && !skip(in, static_cast<size_t>(nbonds) * (4 * sizeof(int) + 2 * sizeof(Real))))
        return bad("truncated bonds");
    std::vector<std::vector<Real>> g, C;
    bool scalar_loaded = false;
    if (version >= 3) {
        int has_scalar = 0;
        if (!get(in, has_scalar)) return bad("truncated scalar block");
        if (has_scalar) {
            int nsp = 0;
            if (!get(in, nsp)) return bad("truncated scalar block");
            if (adr_ && nsp != adr_->nsp_) return bad("scalar species count mismatch: file has " + std::to_string(nsp));
            g.assign(nsp, std::vector<Real>(9 * N));
            C.assign(nsp, std::vector<Real>(N));
            for (int s = 0; s < nsp; ++s)
                if (!get_n(in, g[s].data(), 9 * N) || !get_n(in, C[s].data(), N)) return bad("truncated scalar block");
            int n_tracked = 0;
            if (!get(in, n_tracked) || (n_tracked > 0 && !skip(in, 2 * static_cast<size_t>(n_tracked) * sizeof(Real))))
                return bad("truncated scalar trackers");
            if (version >= 4) {
                int n_mp = 0, n_gamma = 0;
                if (!get(in, n_mp) || (n_mp > 0 && !skip(in, static_cast<size_t>(n_mp) * sizeof(Real))) ||
                    !get(in, n_gamma) || (n_gamma > 0 && !skip(in, static_cast<size_t>(n_gamma) * sizeof(Real))))
                    return bad("truncated scalar trackers");
            }
            if (adr_) scalar_loaded = true;
            else std::cerr << "Checkpoint: file contains scalar state but sim has no scalar solver - "
                              "scalar state will be ignored\n";
        }
    }
    // the extension block, when present
    std::vector<double> force;
    std::vector<int> has_prev;
    std::vector<std::vector<Real>> prev_len;
    char ext[4];
    in.read(ext, 4);
    if (in.good() && std::strncmp(ext, "SFXT", 4) == 0) {
        int ev = 0;
        if (!get(in, ev) || ev != SFXT_VERSION) return bad("unsupported extension block");
        force.resize(2 * N);
        if (!get_n(in, force.data(), force.size())) return bad("truncated extension block");
        has_prev.resize(ncaps);
        prev_len.resize(ncaps);
        for (int c = 0; c < ncaps; ++c) {
            prev_len[c].resize(cs[c].numNodes());
            if (!get(in, has_prev[c]) || !get_n(in, prev_len[c].data(), prev_len[c].size()))
                return bad("truncated extension block");
        }
    }
    // apply: the host mirrors are current (pulled above), then overwritten and pushed
    LatticeField& fl = *field_;
    fl.pull();
    fl.f_ = std::move(f);
    fl.rho_ = std::move(rho);
    fl.ux_ = std::move(ux);
    fl.uy_ = std::move(uy);
    if (fl.flags_ != flags) {
        fl.flags_ = std::move(flags);
        fl.flags_dirty_ = true;
    }
    fl.f_dirty_ = fl.macro_dirty_ = true;
    for (int c = 0; c < ncaps; ++c) {
        cs[c].pos_ = pos[c];
        cs[c].vel_ = vel[c];
        if (!force.empty()) {
            cs[c].has_prev_ = has_prev[c] != 0;
            cs[c].prev_len_ = prev_len[c];
        }
    }
    caps_.dirty_ = true;
    if (scalar_loaded) {
        adr_->pull();
        for (int s = 0; s < adr_->nsp_; ++s) {
            adr_->g_[s] = g[s];
            adr_->C_[s] = C[s];
        }
        adr_->dirty_ = true;
    }
    pending_force_ = std::move(force);
    step_count_ = step;
    pushDirty();
    std::cout << "Checkpoint loaded: step " << step << ", " << ncaps << " capsules <- " << filename
              << std::endl;
    return true;
}

// ---------------------------------------------------------------------------
// Ensemble

Ensemble::Ensemble(const std::vector<Simulation*>& members) : members_(members) {
    if (members_.empty()) throw std::runtime_error("Ensemble: no members");
    std::vector<sf_ctx*> ctx;
    for (Simulation* s : members_) {
        if (!s || !s->initialized()) throw std::runtime_error("Ensemble: members must be initialised simulations");
        ctx.push_back(s->deviceContext());
    }
    check(sf_batch_create(&batch_, ctx.data(), static_cast<int>(ctx.size())), "sf_batch_create");
}

Ensemble::~Ensemble() {
    if (batch_) sf_batch_destroy(batch_);
}

void Ensemble::check(int rc, const char* what) const {
    if (rc != SF_OK) throw std::runtime_error(std::string(what) + ": " + sf_last_error());
}

void Ensemble::run(int num_steps) {
    for (Simulation* s : members_) s->pushDirty();
    check(sf_batch_step(batch_, num_steps), "sf_batch_step");
    for (Simulation* s : members_) s->markAdvanced(num_steps);
}

void Ensemble::synchronize() { check(sf_batch_synchronize(batch_), "sf_batch_synchronize"); }

std::string Ensemble::autotune(int reps) {
    for (Simulation* s : members_) s->pushDirty();
    std::vector<char> buf(4096, 0);
    check(sf_batch_autotune(batch_, reps, buf.data(), static_cast<int>(buf.size())), "sf_batch_autotune");
    return std::string(buf.data());
}

void Ensemble::setLbmLaunch(int bx, int by, bool xcd_banded) {
    check(sf_batch_set_lbm_launch(batch_, bx, by, xcd_banded ? 1 : 0), "sf_batch_set_lbm_launch");
}

void Ensemble::setLbmFlat(bool on) { check(sf_batch_set_lbm_flat(batch_, on ? 1 : 0), "sf_batch_set_lbm_flat"); }

void Ensemble::setGraphMode(bool on) { check(sf_batch_set_graph_mode(batch_, on ? 1 : 0), "sf_batch_set_graph_mode"); }

std::vector<int64_t> Ensemble::info() {
    std::vector<int64_t> out(4, 0);
    check(sf_batch_info(batch_, out.data()), "sf_batch_info");
    return out;
}

}  // namespace softflow
