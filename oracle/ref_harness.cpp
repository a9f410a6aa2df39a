// Driver for the reference engine's own hot-path components, built from the
// reference sources in place into oracle/_ref/libsfref.so (see Makefile).
//
// TEST INFRASTRUCTURE ONLY: used to certify the oracle restatement (sfo.cpp),
// to write golden vectors and, built with the reference's Release flags
// (libsfref_release.so), as the CPU baseline that bench.py times.  The built
// libraries travel with the tree to the GPU machine for that baseline; the
// reference sources do not.  The product never loads either library.
//
// engine/simulation.cpp cannot be compiled here: it includes the CMake-
// generated <softflow/build_info.h>.  This driver therefore composes the
// reference's LBMSolver, ImmersedBoundary, CapsuleSystem, RepulsionForce and
// AdvectionDiffusion in the order of Simulation::step (simulation.cpp:470-787)
// and restates Simulation::addCapsuleRandom (simulation.cpp:220-274).
#include "sfo.h"

#include "core/parameters.h"
#include "coupling/immersed_boundary.h"
#include "geometry/channel_builder.h"
#include "geometry/polygon_obstacle.h"
#include "io/csv_writer.h"
#include "io/fluid_probe_writer.h"
#include "io/global_stats_writer.h"
#include "io/vtk_writer.h"
#include "lbm/advection_diffusion.h"
#include "lbm/lbm_solver.h"
#include "membrane/capsule_system.h"
#include "membrane/repulsion.h"

#include <omp.h>

#include <cstring>
#include <memory>
#include <random>

using namespace softflow;

struct sfr_sim {
    SimulationParams params;
    std::unique_ptr<LBMSolver> lbm;
    CapsuleSystem caps;
    ImmersedBoundary ibm;
    std::unique_ptr<RepulsionForce> rep;
    std::unique_ptr<AdvectionDiffusion> adr;
    std::vector<std::shared_ptr<Obstacle>> obs;     // in the order added
    std::vector<int> obs_kind;
    int capped = 0;
    // the reference's writers (simulation.cpp:176-191)
    std::unique_ptr<VTKWriter> vtk;
    std::unique_ptr<CSVWriter> csv;
    std::unique_ptr<FluidProbeWriter> probes;
    std::unique_ptr<GlobalStatsWriter> stats;
};

static SimulationParams to_params(const sfo_params* p) {
    SimulationParams s;
    s.nx = p->nx; s.ny = p->ny; s.dt = p->dt; s.ibm_iterations = p->ibm_iterations;
    s.fluid.tau = p->tau; s.fluid.rho0 = p->rho0;
    s.fluid.boundary_type = static_cast<BoundaryType>(p->boundary_type);
    s.fluid.body_force_x = p->body_force_x; s.fluid.body_force_y = p->body_force_y;
    s.fluid.gravity_x = p->gravity_x; s.fluid.gravity_y = p->gravity_y;
    s.fluid.apply_gravity_to_fluid = p->apply_gravity_to_fluid != 0;
    s.fluid.apply_gravity_to_capsules = p->apply_gravity_to_capsules != 0;
    s.fluid.collision_model = static_cast<CollisionModel>(p->collision_model);
    s.fluid.use_mrt = p->use_mrt != 0;
    s.fluid.mrt_se = p->mrt_se; s.fluid.mrt_s_eps = p->mrt_s_eps; s.fluid.mrt_sq = p->mrt_sq;
    s.fluid.periodic_y = p->periodic_y != 0;
    s.max_lattice_force = p->max_lattice_force;
    s.repulsion.epsilon = p->rep_epsilon; s.repulsion.sigma = p->rep_sigma;
    s.repulsion.r_cut = p->rep_r_cut; s.repulsion.power = p->rep_power;
    s.repulsion.damping_normal = p->rep_damping_normal;
    s.repulsion.friction_coeff = p->rep_friction_coeff;
    s.scalar.enabled = p->scalar_enabled != 0;
    s.scalar.n_species = p->n_species;
    s.scalar.diffusivity.assign(p->diffusivity, p->diffusivity + p->n_species);
    s.scalar.inlet_concentration.assign(p->inlet_concentration,
                                        p->inlet_concentration + p->n_species);
    s.scalar.periodic_y = p->scalar_periodic_y != 0;
    s.fluid.top_wall_velocity = p->top_wall_velocity;
    s.fluid.bottom_wall_velocity = p->bottom_wall_velocity;
    s.fluid.inlet_velocity = p->inlet_velocity;
    s.fluid.outlet_density = p->outlet_density;
    s.fluid.use_interpolated_bb = p->use_interpolated_bb != 0;
    s.fluid.capsule_periodic_x = p->capsule_periodic_x != 0;
    return s;
}

// Simulation::step 6b (simulation.cpp:588-609), restated: simulation.cpp is
// not compiled here.
static int cap_forces(Real* Fx, Real* Fy, int N, Real Fxy_max) {
    const Real Fxy_max2 = Fxy_max * Fxy_max;
    int capped = 0;
    for (int n = 0; n < N; ++n) {
        Real fx = Fx[n], fy = Fy[n];
        Real fmag2 = fx * fx + fy * fy;
        if (fmag2 > Fxy_max2) {
            Real scale = Fxy_max / std::sqrt(fmag2);
            Fx[n] *= scale;
            Fy[n] *= scale;
            ++capped;
        }
    }
    return capped;
}

// Simulation::step 10 (simulation.cpp:665-675), restated.
static void clamp_y(CapsuleSystem& caps, int ny) {
    Real y_min = 1.5;
    Real y_max = static_cast<Real>(ny) - 1.5;
    for (int c = 0; c < caps.numCapsules(); ++c)
        for (auto& p : caps[c].positions()) {
            if (p.y < y_min) p.y = y_min;
            if (p.y > y_max) p.y = y_max;
        }
}

// Simulation::step 10a (simulation.cpp:687-700), restated over the
// reference's own Capsule::centroid.
static int recirculate_x(CapsuleSystem& caps, int nx) {
    const Real Lx = static_cast<Real>(nx);
    int moved = 0;
    for (int c = 0; c < caps.numCapsules(); ++c) {
        const Vec2d cen = caps[c].centroid();
        Real shift = 0.0;
        if (cen.x >= Lx) shift = -Lx;
        else if (cen.x < 0.0) shift = Lx;
        if (shift != 0.0) {
            for (auto& p : caps[c].positions()) p.x += shift;
            ++moved;
        }
    }
    return moved;
}

static MembraneParams to_mem(const sfo_membrane* m) {
    MembraneParams r;
    r.model = static_cast<MembraneModel>(m->model);
    r.k_stretch = m->k_stretch; r.G_s = m->G_s; r.C_skalak = m->C_skalak;
    r.wlc_L_max_ratio = m->wlc_L_max_ratio; r.wlc_kBT_p = m->wlc_kBT_p; r.wlc_k_pow = m->wlc_k_pow;
    r.k_bend = m->k_bend; r.use_helfrich_bending = m->use_helfrich != 0; r.kappa_0 = m->kappa_0;
    r.k_area = m->k_area; r.k_perimeter = m->k_perimeter; r.gamma_visc = m->gamma_visc;
    r.eta_membrane = m->eta_membrane; r.viscosity_ratio = m->viscosity_ratio;
    r.density = m->density; r.shape = static_cast<CapsuleShape>(m->shape);
    r.aspect_ratio = m->aspect_ratio; r.indent_depth = m->indent_depth;
    r.is_rigid = m->is_rigid != 0;
    return r;
}

extern "C" {

sfr_sim* sfr_create(const sfo_params* p) {
    omp_set_num_threads(1);   // the oracle's summation order is the one-thread order
    auto* s = new sfr_sim();
    s->params = to_params(p);
    s->lbm = std::make_unique<LBMSolver>(s->params);
    s->rep = std::make_unique<RepulsionForce>(s->params.repulsion);
    if (s->params.fluid.boundary_type == BoundaryType::PERIODIC) {
        s->rep->setPeriodicX(static_cast<Real>(s->params.nx));
        s->caps.setPeriodicX(s->params.nx);
    }
    if (s->params.scalar.enabled)
        s->adr = std::make_unique<AdvectionDiffusion>(s->params.nx, s->params.ny, s->params.scalar);
    return s;
}

void sfr_destroy(sfr_sim* s) { delete s; }

// OpenMP threads for the reference's parallel regions from here on (the
// CPU-baseline timing; every comparison runs at the default single thread).
void sfr_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void sfr_add_capsule(sfr_sim* s, double cx, double cy, double radius, int num_nodes,
                     const sfo_membrane* m, int type) {
    s->caps.addCapsule(Vec2d{cx, cy}, radius, num_nodes, to_mem(m), type);
}

int sfr_add_capsule_random(sfr_sim* s, int count, double x0, double y0, double x1, double y1,
                           double rmin, double rmax, int num_nodes, const sfo_membrane* m,
                           int type, int seed, double min_gap, int max_attempts) {
    MembraneParams mp = to_mem(m);
    std::mt19937 rng(static_cast<unsigned>(seed));
    std::uniform_real_distribution<Real> dx(x0, x1), dy(y0, y1), dr(rmin, rmax);
    int placed = 0, attempts = 0;
    int total = (max_attempts > 0) ? max_attempts : count * 500;
    const Real wb = 0.0, wt = static_cast<Real>(s->params.ny - 1);   // ChannelBuilder walls
    while (placed < count && attempts < total) {
        Real cx = dx(rng);
        Real cy = dy(rng);
        Real r = dr(rng);
        bool overlap = false;
        for (int c = 0; c < s->caps.numCapsules(); ++c) {
            Vec2d oc = s->caps[c].centroid();
            Real orr = s->caps[c].effectiveRadius();
            if ((Vec2d{cx, cy} - oc).norm() < r + orr + min_gap) { overlap = true; break; }
        }
        if (cy - r < wb + min_gap || cy + r > wt - min_gap) overlap = true;
        for (const auto& o : s->obs)
            if (o->signedDistance(cx, cy) < r + min_gap) { overlap = true; break; }
        if (!overlap) {
            int nn = (num_nodes > 0) ? num_nodes : Capsule::computeOptimalNodes(r);
            s->caps.addCapsule({cx, cy}, r, nn, mp, type);
            placed++;
        }
        attempts++;
    }
    return placed;
}

void sfr_add_obstacle(sfr_sim* s, int kind, const double* c, int n) {
    if (kind == 0) s->obs.push_back(std::make_shared<CircleObstacle>(c[0], c[1], c[2]));
    else if (kind == 1) s->obs.push_back(std::make_shared<RectObstacle>(c[0], c[1], c[2], c[3]));
    else {
        std::vector<Vec2d> v;
        for (int i = 0; i + 1 < n; i += 2) v.push_back(Vec2d{c[i], c[i + 1]});
        s->obs.push_back(std::make_shared<PolygonObstacle>(v));
    }
    s->obs_kind.push_back(kind);
}

// The reference's writers on this simulation's state, set up as the
// Simulation constructor does (simulation.cpp:173-191).  options: bit 0 legacy
// VTK, bit 1 every CSV column, bit 2 the optional fluid fields (vorticity,
// strain rate, IBM force).  probes: n (i, j) pairs labelled p0, p1, ...
void sfr_output_open(sfr_sim* s, const char* dir, int options, int nprobes, const int* ij) {
    const std::string d(dir);
    s->vtk = std::make_unique<VTKWriter>(d);
    if (s->params.fluid.boundary_type == BoundaryType::PERIODIC) s->vtk->setPeriodicX(s->params.nx);
    if (options & 1) s->vtk->setLegacyMode(true);
    if (s->adr) s->vtk->setAdvectionDiffusion(s->adr.get());
    if (options & 4) {
        FluidOutputFields ff;
        ff.vorticity = ff.strain_rate = ff.ibm_force = true;
        s->vtk->setFluidFields(ff);
    }
    s->csv = std::make_unique<CSVWriter>(d);
    if (options & 2) {
        ParticleOutputFields pf;
        pf.force = pf.diameter = pf.group = pf.deformation = pf.area_volume = true;
        pf.orientation = pf.angular_vel = true;
        s->csv->setFields(pf);
    }
    s->probes = std::make_unique<FluidProbeWriter>(d);
    for (int i = 0; i < nprobes; ++i) s->probes->addProbe(ij[2 * i], ij[2 * i + 1], "p" + std::to_string(i));
    s->stats = std::make_unique<GlobalStatsWriter>(d);
}

// Simulation::step's output block (simulation.cpp:736-767) for the writers in mask
// (bit 0 VTK, 1 CSV, 2 probes, 3 statistics)
void sfr_output_write(sfr_sim* s, int step, double time, int mask) {
    LatticeField& f = s->lbm->field();
    if (mask & 1) {
        s->vtk->writeFluidField(f, step);
        s->vtk->writeCapsules(s->caps, step);
        s->vtk->recordTimestep(step, time);
    }
    if (mask & 2) s->csv->writeTimestep(s->caps, step, time);
    if (mask & 4) s->probes->writeTimestep(f, step, time);
    if (mask & 8) s->stats->writeTimestep(f, s->caps, step, time, s->adr.get());
}

// Simulation::finalize (simulation.cpp:801-812)
void sfr_output_close(sfr_sim* s) {
    if (s->vtk) s->vtk->writePVDFiles();
    if (s->csv) s->csv->close();
    if (s->probes) s->probes->close();
    if (s->stats) s->stats->close();
}

double sfr_obstacle_distance(const sfr_sim* s, int i, double x, double y, double* normal) {
    if (normal) { Vec2d nn = s->obs[i]->normalAt(x, y); normal[0] = nn.x; normal[1] = nn.y; }
    return s->obs[i]->signedDistance(x, y);
}

// Simulation::initialize (simulation.cpp:429-442): the channel builder's cell
// types, then the solver's rows and columns, then the IBB links.  The builder
// takes circles and rectangles; polygons have no builder call and are marked
// SOLID here the way applyToField marks its obstacles.
void sfr_initialize(sfr_sim* s) {
    if (!s->obs.empty()) {
        ChannelBuilder b(s->params.nx, s->params.ny);
        b.addWalls();
        for (size_t i = 0; i < s->obs.size(); ++i) {
            if (s->obs_kind[i] == 0) {
                auto* o = static_cast<CircleObstacle*>(s->obs[i].get());
                b.addCirclePillar(o->getCx(), o->getCy(), o->getRadius());
            } else if (s->obs_kind[i] == 1) {
                auto* o = static_cast<RectObstacle*>(s->obs[i].get());
                b.addRectPillar(o->getX0(), o->getY0(), o->getX1(), o->getY1());
            }
        }
        b.setBoundaryType(s->params.fluid.boundary_type);
        LatticeField& f = s->lbm->field();
        b.applyToField(f);
        for (size_t i = 0; i < s->obs.size(); ++i)
            if (s->obs_kind[i] == 2)
                for (int y = 0; y < s->params.ny; ++y)
                    for (int x = 0; x < s->params.nx; ++x)
                        if (s->obs[i]->contains(static_cast<Real>(x), static_cast<Real>(y)))
                            f.setCellType(x, y, CellType::SOLID);
    }
    s->lbm->initialize();
    if (s->params.fluid.use_interpolated_bb) s->lbm->setObstacles(s->obs);
    if (s->adr) s->adr->initialize(0.0);
    if (s->params.fluid.boundary_type == BoundaryType::PERIODIC) s->caps.setPeriodicX(s->params.nx);
}

void sfr_scalar_initialize(sfr_sim* s, double c0) { if (s->adr) s->adr->initialize(c0); }
void sfr_set_scalar_point(sfr_sim* s, int x, int y, double c, int sp) {
    if (s->adr) s->adr->setPoint(x, y, c, sp);
}

void sfr_step(sfr_sim* s, int nsteps) {
    const SimulationParams& P = s->params;
    for (int it = 0; it < nsteps; ++it) {
        LatticeField& field = s->lbm->field();
        field.clearForces();
        s->caps.clearAllForces();
        s->caps.computeAllMembraneForces();
        s->rep->computeAll(s->caps, 0.5, static_cast<Real>(P.ny) - 1.5);
        for (const auto& o : s->obs) s->rep->computeObstacleRepulsion(s->caps, *o);
        if (P.fluid.apply_gravity_to_capsules &&
            (P.fluid.gravity_x != 0.0 || P.fluid.gravity_y != 0.0))
            for (int c = 0; c < s->caps.numCapsules(); ++c)
                s->caps[c].applyBuoyancyForce(P.fluid.gravity_x, P.fluid.gravity_y, P.fluid.rho0);
        if (P.ibm_iterations > 1) s->ibm.multiDirectForcing(s->caps, field, P.ibm_iterations);
        else s->ibm.spreadForces(s->caps, field);
        const Real bfx = P.fluid.body_force_x, bfy = P.fluid.body_force_y;
        const Real gx = P.fluid.gravity_x, gy = P.fluid.gravity_y;
        const bool hb = (bfx != 0.0 || bfy != 0.0);
        const bool hg = P.fluid.apply_gravity_to_fluid && (gx != 0.0 || gy != 0.0);
        const int N = field.size();
        if (hb || hg)
            for (int n = 0; n < N; ++n)
                if (field.cellTypeData()[n] == CellType::FLUID) {
                    if (hb) { field.FxData()[n] += bfx; field.FyData()[n] += bfy; }
                    if (hg) {
                        Real rho = field.rhoData()[n];
                        field.FxData()[n] += rho * gx; field.FyData()[n] += rho * gy;
                    }
                }
        s->capped = cap_forces(field.FxData(), field.FyData(), N, P.max_lattice_force);
        s->lbm->step();
        if (s->adr) s->adr->step(field);
        s->ibm.interpolateVelocity(field, s->caps);
        s->caps.moveAllNodes(P.dt);
        if (!P.fluid.periodic_y) clamp_y(s->caps, P.ny);
        if (P.fluid.capsule_periodic_x && P.fluid.boundary_type != BoundaryType::PERIODIC)
            recirculate_x(s->caps, P.nx);
    }
}

int sfr_num_capsules(const sfr_sim* s) { return s->caps.numCapsules(); }
int sfr_total_nodes(const sfr_sim* s) { return s->caps.totalNodes(); }
void sfr_capsule_offsets(const sfr_sim* s, int* off) {
    off[0] = 0;
    for (int c = 0; c < s->caps.numCapsules(); ++c) off[c + 1] = off[c] + s->caps[c].numNodes();
}
void sfr_get_nodes(const sfr_sim* s, double* x, double* y, double* vx, double* vy, double* fx,
                   double* fy) {
    int k = 0;
    for (int c = 0; c < s->caps.numCapsules(); ++c) {
        const Capsule& cp = s->caps[c];
        for (int i = 0; i < cp.numNodes(); ++i, ++k) {
            if (x) x[k] = cp.nodePosition(i).x;
            if (y) y[k] = cp.nodePosition(i).y;
            if (vx) vx[k] = cp.nodeVelocity(i).x;
            if (vy) vy[k] = cp.nodeVelocity(i).y;
            if (fx) fx[k] = cp.nodeForce(i).x;
            if (fy) fy[k] = cp.nodeForce(i).y;
        }
    }
}
void sfr_set_nodes(sfr_sim* s, const double* x, const double* y, const double* vx,
                   const double* vy) {
    int k = 0;
    for (int c = 0; c < s->caps.numCapsules(); ++c) {
        Capsule& cp = s->caps[c];
        for (int i = 0; i < cp.numNodes(); ++i, ++k) {
            if (x) cp.positions()[i].x = x[k];
            if (y) cp.positions()[i].y = y[k];
            if (vx) cp.velocities()[i].x = vx[k];
            if (vy) cp.velocities()[i].y = vy[k];
        }
    }
}
void sfr_get_flags(const sfr_sim* s, int* flags) {
    const LatticeField& f = s->lbm->field();
    const int nx = s->params.nx, ny = s->params.ny;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) flags[y * nx + x] = static_cast<int>(f.getCellType(x, y));
}
void sfr_set_flags(sfr_sim* s, const int* flags) {
    LatticeField& f = s->lbm->field();
    const int nx = s->params.nx, ny = s->params.ny;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) f.setCellType(x, y, static_cast<CellType>(flags[y * nx + x]));
    s->lbm->rebuildBoundaryNodeLists();
}
void sfr_get_field(const sfr_sim* s, double* rho, double* ux, double* uy, double* fx,
                   double* fy) {
    const LatticeField& f = s->lbm->field();
    size_t N = f.size();
    if (rho) std::memcpy(rho, f.rhoData(), N * 8);
    if (ux) std::memcpy(ux, f.uxData(), N * 8);
    if (uy) std::memcpy(uy, f.uyData(), N * 8);
    if (fx) std::memcpy(fx, f.FxData(), N * 8);
    if (fy) std::memcpy(fy, f.FyData(), N * 8);
}
void sfr_get_f(const sfr_sim* s, double* out) {
    const LatticeField& f = s->lbm->field();
    std::memcpy(out, f.fData(), static_cast<size_t>(f.size()) * 9 * 8);
}
void sfr_get_scalar(const sfr_sim* s, double* C, int sp) {
    if (s->adr) std::memcpy(C, s->adr->concentrationData(sp), static_cast<size_t>(s->params.nx) * s->params.ny * 8);
}
int sfr_last_capped(const sfr_sim* s) { return s->capped; }

// a write through the reference's concentration() view: its own array
void sfr_set_concentration(sfr_sim* s, const double* C, int sp) {
    if (!s->adr || sp < 0 || sp >= s->params.scalar.n_species) return;
    std::memcpy(s->adr->concentrationData(sp), C,
                sizeof(Real) * static_cast<size_t>(s->params.nx) * s->params.ny);
}

int sfr_cap_forces(double* Fx, double* Fy, int N, double max_force) {
    return cap_forces(Fx, Fy, N, max_force);
}

void sfr_clamp_nodes_y(double* y, int K, int ny) {
    // one capsule holding the K nodes, clamped by the step's restatement
    CapsuleSystem cs;
    MembraneParams mp;
    cs.addCapsule(Vec2d{0.0, 0.0}, 1.0, K > 2 ? K : 3, mp);
    auto& pos = cs[0].positions();
    for (int k = 0; k < K; ++k) pos[k] = Vec2d{0.0, y[k]};
    clamp_y(cs, ny);
    for (int k = 0; k < K; ++k) y[k] = pos[k].y;
}

int sfr_recirculate_x(double* x, const double* y, const int* off, int ncaps, int nx) {
    CapsuleSystem cs;
    MembraneParams mp;
    for (int c = 0; c < ncaps; ++c) {
        const int n = off[c + 1] - off[c];
        cs.addCapsule(Vec2d{0.0, 0.0}, 1.0, n, mp);
        auto& pos = cs[c].positions();
        for (int i = 0; i < n; ++i) pos[i] = Vec2d{x[off[c] + i], y[off[c] + i]};
    }
    const int moved = recirculate_x(cs, nx);
    for (int c = 0; c < ncaps; ++c) {
        auto& pos = cs[c].positions();
        for (int i = 0; i < off[c + 1] - off[c]; ++i) x[off[c] + i] = pos[i].x;
    }
    return moved;
}
double sfr_capsule_geometry(const sfr_sim* s, int c, int what) {
    const Capsule& cp = s->caps[c];
    switch (what) {
    case 0: return cp.area();
    case 1: return cp.perimeter();
    case 2: return cp.centroid().x;
    case 3: return cp.centroid().y;
    case 4: return cp.deformationIndex();
    case 7: return cp.inclinationAngle();
    }
    return 0.0;
}
double sfr_phi(double r) { return PeskinDelta::phi(r); }
void sfr_membrane_forces_only(sfr_sim* s) {
    s->caps.clearAllForces();
    s->caps.computeAllMembraneForces();
}

}  // extern "C"
