// Host engine of the MI355X build: the reference's stepping API
// (engine/simulation.h, lbm/lattice_field.h, membrane/capsule*.h) over the
// C-ABI of libsoftflow_hip.  Device state is authoritative; the host objects
// are mirrors refreshed on access and pushed back when modified.
#pragma once
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/softflow.h"
#include "output.h"
#include "sf_geometry.h"

namespace softflow {

using Real = double;
constexpr Real PI = 3.14159265358979323846;

struct Vec2d {
    Real x = 0.0, y = 0.0;
    Vec2d() = default;
    Vec2d(Real x_, Real y_) : x(x_), y(y_) {}
    Vec2d operator+(const Vec2d& o) const { return {x + o.x, y + o.y}; }
    Vec2d operator-(const Vec2d& o) const { return {x - o.x, y - o.y}; }
    Vec2d operator*(Real s) const { return {x * s, y * s}; }
    Vec2d operator/(Real s) const { return {x / s, y / s}; }
    Real dot(const Vec2d& o) const { return x * o.x + y * o.y; }
    Real cross(const Vec2d& o) const { return x * o.y - y * o.x; }
    Real norm() const { return std::sqrt(x * x + y * y); }
};

enum class CellType : int { FLUID = 0, SOLID = 1, INLET = 2, OUTLET = 3, EMPTY = 4 };
enum class BoundaryType : int { INLET_OUTLET = 0, PERIODIC = 1, CLOSED = 2 };
enum class CollisionModel : int { BGK = 0, MRT = 1, REGULARIZED = 2 };
enum class MembraneModel : int { HOOKEAN = 0, NEO_HOOKEAN = 1, SKALAK = 2, WLC = 3 };
enum class CapsuleShape : int { CIRCLE = 0, ELLIPSE = 1, BICONCAVE = 2, FIBER = 3 };
enum class Precision : int { FP64 = 0, FP32 = 1 };

// Parameter structs: the fields of core/parameters.h that the hot path reads,
// with the reference's defaults, plus the fields the Python layer sets.
struct FluidParams {
    Real tau = 0.8, rho0 = 1.0;
    Real inlet_velocity = 0.02, outlet_density = 1.0;
    BoundaryType boundary_type = BoundaryType::INLET_OUTLET;
    Real body_force_x = 0.0, body_force_y = 0.0;
    Real gravity_x = 0.0, gravity_y = 0.0;
    bool apply_gravity_to_fluid = true, apply_gravity_to_capsules = true;
    CollisionModel collision_model = CollisionModel::BGK;
    bool use_mrt = false;
    Real mrt_se = 1.4, mrt_s_eps = 1.4, mrt_sq = 1.4;
    bool periodic_y = false;
    bool capsule_periodic_x = false;
    Real top_wall_velocity = 0.0, bottom_wall_velocity = 0.0;
    bool use_interpolated_bb = false;
    bool viscosity_contrast = false;
    int viscosity_update_interval = 10;
};

struct MembraneParams {
    MembraneModel model = MembraneModel::HOOKEAN;
    Real k_stretch = 0.1, G_s = 0.1, C_skalak = 1.0;
    Real wlc_L_max_ratio = 2.2, wlc_kBT_p = 0.0, wlc_k_pow = 0.0;
    Real k_bend = 0.005;
    bool use_helfrich_bending = false;
    Real kappa_0 = 0.0;
    Real k_area = 0.5, k_perimeter = 0.1, gamma_visc = 0.01, eta_membrane = 0.05;
    Real viscosity_ratio = 1.0, density = 1.0;
    CapsuleShape shape = CapsuleShape::CIRCLE;
    Real aspect_ratio = 1.0, indent_depth = 0.4;
    bool is_rigid = false;
};

struct RepulsionParams {
    Real epsilon = 0.05, sigma = 1.0, r_cut = 3.5;
    int power = 4;
    Real damping_normal = 0.0, friction_coeff = 0.0;
};

struct LubricationParams { bool enabled = false; Real h_threshold = 2.0, h_min = 0.1; };
// Off the hot path (refused at initialize() when enabled); the fields and
// defaults are core/parameters.h's, so a front end can fill them.
struct AdhesionParams {
    bool enabled = false;
    Real k_on = 0.001, k_off = 0.01, k_bond = 0.05, d_bond = 2.0, F_crit = 0.01;
    int max_bonds_per_node = 3;
    bool use_catch_slip = false;
    Real k_off_catch = 0.05, F_catch = 0.02, k_off_slip = 0.001, F_slip = 0.01;
    bool wall_adhesion = false;
    Real wall_k_on = 0.001, wall_k_off = 0.01, wall_k_bond = 0.05, wall_receptor_spacing = 2.0;
    std::vector<std::vector<bool>> adhesion_matrix;
};
struct ShanChenParams {
    bool enabled = false;
    Real G = -5.0, rho0_sc = 1.0;
    int eos_type = 0;
    Real cs_a = 1.0, cs_b = 4.0, cs_T = 0.06, cs_R = 1.0;
    Real wetting_angle = 90.0;
    int n_components = 1;
    Real G_12 = -5.0, tau_2 = 0.8;
};

struct ScalarParams {
    bool enabled = false;
    int n_species = 1;
    std::vector<Real> diffusivity = {0.01};
    std::vector<Real> inlet_concentration = {0.0};
    bool periodic_y = false;
};

struct SimulationParams {
    int nx = 200, ny = 50;
    Real dt = 1.0;
    int num_steps = 10000, output_interval = 100;
    int ibm_iterations = 1;
    FluidParams fluid;
    MembraneParams membrane;
    RepulsionParams repulsion;
    LubricationParams lubrication;
    AdhesionParams adhesion;
    ShanChenParams shan_chen;
    ScalarParams scalar;
    std::string output_dir = "output";
    std::string output_format = "vtk";     // "vtk" or "csv" (parameters.h)
    std::string vtk_format = "ascii";      // "ascii" (XML .vti/.vtp) or "legacy" (.vtk)
    int checkpoint_interval = 0, metrics_interval = 0;
    int vtk_dump_every = 0, csv_dump_every = 0, probe_dump_every = 10, stats_dump_every = 100;
    bool enable_stability_checks = true;
    Real max_density_deviation = 0.1, max_velocity = 0.1;
    int stability_check_interval = 1000;
    bool enable_profiling = true;
    uint64_t rng_seed = 0;
    Real max_lattice_force = 0.01;
    // MI355X build: storage/collide precision of the lattice (FP64 reproduces
    // the reference's arithmetic; FP32 stores f - w_q and collides in FP32)
    Precision precision = Precision::FP64;
    bool use_graph = true;
    int device = 0;
    Real kinematicViscosity() const { return (fluid.tau - 0.5) / 3.0; }
};

class Simulation;
class CapsuleSystem;

// Host mirror of the device lattice (lbm/lattice_field.h:18-122).
class LatticeField {
public:
    LatticeField(int nx, int ny);
    int getNx() const { return nx_; }
    int getNy() const { return ny_; }
    int size() const { return N_; }
    int idx(int x, int y) const { return y * nx_ + x; }

    Real* fData() { pull(); return f_.data(); }
    Real* rhoData() { pull(); return rho_.data(); }
    Real* uxData() { pull(); return ux_.data(); }
    Real* uyData() { pull(); return uy_.data(); }
    Real* FxData() { pull(); return Fx_.data(); }
    Real* FyData() { pull(); return Fy_.data(); }
    const int* cellTypeData() const { return flags_.data(); }

    Real f(int x, int y, int q) { pull(); return f_[static_cast<size_t>(q) * N_ + idx(x, y)]; }
    void set_f(int x, int y, int q, Real v);
    Real getRho(int x, int y) { pull(); return rho_[idx(x, y)]; }
    Real getUx(int x, int y) { pull(); return ux_[idx(x, y)]; }
    Real getUy(int x, int y) { pull(); return uy_[idx(x, y)]; }
    CellType getCellType(int x, int y) const { return static_cast<CellType>(flags_[idx(x, y)]); }
    void setCellType(int x, int y, CellType ct);
    void setRho(int x, int y, Real v);
    void setUx(int x, int y, Real v);
    void setUy(int x, int y, Real v);
    void setEquilibrium(Real rho0, Real ux0, Real uy0);
    void initializeEquilibriumAt(int x, int y);
    void initializeAllEquilibrium();
    void computeMacroscopic();

    // Writable numpy views (python/bindings.cpp:313-355 hands out zero-copy
    // views of the live arrays).  Here a view aliases the host mirror; while
    // one is held (leased), the mirror is refreshed after every step so the
    // view shows the current state, and writes into it are found by comparison
    // with a shadow copy and pushed to the device before the next step:
    // populations as a host write of f, rho/ux/uy as the stored macroscopic
    // state the next collision reads.  Writes into the force views stay on the
    // host, as the reference's are cleared at the start of the next step.
    enum ViewKind { VIEW_F = 0, VIEW_RHO, VIEW_UX, VIEW_UY, VIEW_FX, VIEW_FY, VIEW_COUNT };
    Real* leaseView(int kind);
    void releaseView(int kind);
    bool anyLease() const;

private:
    friend class Simulation;
    void pull();
    void touch_f() { pull(); f_dirty_ = true; }
    void touch_macro() { pull(); macro_dirty_ = true; }
    std::vector<Real>& viewArray(int kind);
    void collectViewWrites();
    void refreshShadows();
    int nx_, ny_, N_;
    std::vector<Real> f_, rho_, ux_, uy_, Fx_, Fy_;
    std::vector<int> flags_;
    Simulation* sim_ = nullptr;
    bool stale_ = false;
    bool f_dirty_ = false, macro_dirty_ = false, flags_dirty_ = false;
    int leases_[VIEW_COUNT] = {};
    std::vector<Real> shadow_[VIEW_COUNT];
};

// Capsule with rest geometry built as membrane/capsule.cpp:7-141.
class Capsule {
public:
    static int computeOptimalNodes(Real radius, Real target_ds = 0.75) {
        int n = static_cast<int>(std::ceil(2.0 * PI * radius / target_ds));
        return n > 12 ? n : 12;
    }
    Capsule(int id, Vec2d center, Real radius, int num_nodes, const MembraneParams& params,
            int type = 0);

    int numNodes() const { return n_; }
    int getId() const { return id_; }
    int getType() const { return type_; }
    Vec2d nodePosition(int k) const { return pos_[k]; }
    Vec2d nodeVelocity(int k) const { return vel_[k]; }
    Vec2d nodeForce(int k) const { return frc_[k]; }
    void setNodeVelocity(int k, const Vec2d& v);
    Vec2d centroid() const;
    Real area() const;
    Real perimeter() const;
    Real deformationIndex() const;
    Real inclinationAngle() const;
    Real effectiveRadius() const { return std::sqrt(rest_area_ / PI); }
    void setPeriodicX(int nx) { periodic_nx_ = nx; }
    int getPeriodicX() const { return periodic_nx_; }
    MembraneModel getMembraneModel() const { return params_.model; }
    const std::vector<Vec2d>& positions() const { return pos_; }

private:
    friend class Simulation;
    friend class CapsuleSystem;
    friend class OutputManager;
    Vec2d minImageDiff(const Vec2d& a, const Vec2d& b) const;
    void secondMoments(Real& mxx, Real& myy, Real& mxy) const;
    sf_capsule_desc desc() const;

    int id_, type_, n_;
    MembraneParams params_;
    std::vector<Vec2d> pos_, vel_, frc_;
    std::vector<Real> rest_len_, rest_angle_, rest_ds_, prev_len_;
    bool has_prev_ = false;
    Real rest_area_ = 0.0, rest_perimeter_ = 0.0;
    Real wlc_kBT_p_ = 0.0, wlc_k_pow_ = 0.0;
    int periodic_nx_ = 0;
    CapsuleSystem* sys_ = nullptr;
};

class CapsuleSystem {
public:
    void addCapsule(Vec2d center, Real radius, int num_nodes, const MembraneParams& params,
                    int type = 0);
    void setPeriodicX(int nx);
    int numCapsules() { pull(); return static_cast<int>(caps_.size()); }
    int totalNodes() const;
    Capsule& operator[](int i) { pull(); return caps_.at(i); }
    std::vector<Capsule>& capsules() { pull(); return caps_; }

private:
    friend class Simulation;
    friend class Capsule;
    void pull();
    std::vector<Capsule> caps_;
    int next_id_ = 0, periodic_nx_ = 0;
    Simulation* sim_ = nullptr;
    bool stale_ = false, dirty_ = false;
};

// Scalar transport mirror (lbm/advection_diffusion.h:17-99).
class AdvectionDiffusion {
public:
    AdvectionDiffusion(int nx, int ny, const ScalarParams& p);
    void initialize(Real c0 = 0.0);
    void setRegion(int x0, int y0, int x1, int y1, Real c, int species = 0);
    void setPoint(int x, int y, Real c, int species = 0);
    Real getConcentration(int x, int y, int species = 0);
    Real* concentrationData(int species = 0);
    // bulk form of setPoint over the whole lattice (C[y*nx+x], g = w_q C)
    void setField(const Real* C, int species = 0);
    int getNx() const { return nx_; }
    int getNy() const { return ny_; }
    int getNumSpecies() const { return nsp_; }
    // Writable concentration views, leased as LatticeField's are: writes reach
    // the device (with g) before the next step.
    Real* leaseView(int species);
    void releaseView(int species);
    bool anyLease() const;

private:
    friend class Simulation;
    void pull();
    void collectViewWrites();
    void refreshShadows();
    int nx_, ny_, nsp_;
    std::vector<std::vector<Real>> g_, C_;
    Simulation* sim_ = nullptr;
    bool stale_ = false, dirty_ = true;
    std::vector<int> leases_;
    std::vector<std::vector<Real>> shadow_;
};

// Obstacles (geometry/obstacle.h, circle_obstacle.h, rect_obstacle.h,
// polygon_obstacle.h).  The geometry is the one the device path evaluates in
// its contact forces (sf_geometry.h), so placement, cell flags and forces agree.
class Obstacle {
public:
    virtual ~Obstacle() = default;
    bool contains(Real x, Real y) const { return sf::obstacle_contains(geo(), x, y); }
    Real signedDistance(Real x, Real y) const { return sf::obstacle_signed_distance(geo(), x, y); }
    Vec2d nearestPoint(Real x, Real y) const {
        const sf::ONearest q = sf::obstacle_nearest_point(geo(), x, y);
        return {q.x, q.y};
    }
    Vec2d normalAt(Real x, Real y) const {
        const sf::ONearest q = sf::obstacle_normal(geo(), x, y);
        return {q.x, q.y};
    }
    // the C-ABI description: kind and coordinates (sf_set_obstacles)
    int kind() const { return o_.kind; }
    std::vector<double> coords() const;

protected:
    sf::Obstacle geo() const {
        sf::Obstacle o = o_;
        o.v = verts_.data();
        return o;
    }
    sf::Obstacle o_{};
    std::vector<double> verts_;
};

class CircleObstacle : public Obstacle {
public:
    CircleObstacle(Real cx, Real cy, Real radius);
    Real getCx() const { return o_.p[0]; }
    Real getCy() const { return o_.p[1]; }
    Real getRadius() const { return o_.p[2]; }
};

class RectObstacle : public Obstacle {
public:
    RectObstacle(Real x0, Real y0, Real x1, Real y1);
    Real getX0() const { return o_.p[0]; }
    Real getY0() const { return o_.p[1]; }
    Real getX1() const { return o_.p[2]; }
    Real getY1() const { return o_.p[3]; }
};

class PolygonObstacle : public Obstacle {
public:
    explicit PolygonObstacle(const std::vector<Vec2d>& vertices);
};

// Walls, boundary type and obstacles of the channel (geometry/channel_builder.h).
class ChannelBuilder {
public:
    ChannelBuilder(int nx, int ny) : nx_(nx), ny_(ny) {}
    ChannelBuilder& addWalls() { bottom_ = top_ = true; return *this; }
    ChannelBuilder& addCirclePillar(Real cx, Real cy, Real radius) {
        obstacles_.push_back(std::make_shared<CircleObstacle>(cx, cy, radius));
        return *this;
    }
    ChannelBuilder& addRectPillar(Real x0, Real y0, Real x1, Real y1) {
        obstacles_.push_back(std::make_shared<RectObstacle>(x0, y0, x1, y1));
        return *this;
    }
    // no counterpart in the reference's builder, which holds polygons only
    // through its obstacle list: a polygonal pillar
    ChannelBuilder& addPolygonPillar(const std::vector<Vec2d>& vertices) {
        obstacles_.push_back(std::make_shared<PolygonObstacle>(vertices));
        return *this;
    }
    ChannelBuilder& setBoundaryType(BoundaryType bt) { bt_ = bt; return *this; }
    BoundaryType getBoundaryType() const { return bt_; }
    Real wallBottom() const { return 0.0; }
    Real wallTop() const { return static_cast<Real>(ny_ - 1); }
    bool hasBottomWall() const { return bottom_; }
    bool hasTopWall() const { return top_; }
    int getNx() const { return nx_; }
    int getNy() const { return ny_; }
    const std::vector<std::shared_ptr<Obstacle>>& getObstacles() const { return obstacles_; }
    // ChannelBuilder::applyToField (channel_builder.cpp:6-36) on an nx x ny field
    std::vector<int32_t> cellTypes(int nx, int ny, BoundaryType bt) const;

private:
    int nx_, ny_;
    bool bottom_ = false, top_ = false;
    BoundaryType bt_ = BoundaryType::INLET_OUTLET;
    std::vector<std::shared_ptr<Obstacle>> obstacles_;
};

class Simulation;

// The learned warm start of multi-direct forcing (no counterpart in the
// reference): a 17 -> 64 -> 64 -> 2 GELU network on the device, trained on
// the (inputs, converged correction) rows a Simulation collects.
class MdfModel {
public:
    explicit MdfModel(int device = 0, uint64_t seed = 1);
    ~MdfModel();
    MdfModel(const MdfModel&) = delete;
    MdfModel& operator=(const MdfModel&) = delete;

    // Adam over the simulation's sample rows; returns the mean loss per epoch.
    std::vector<double> train(Simulation& sim, int epochs, int batch, double lr, uint64_t seed);
    std::vector<float> params() const;
    std::vector<float> normalization() const;
    void setParams(const std::vector<float>& p);
    void setNormalization(const std::vector<float>& n);
    // y[n][2] for x[n][17]
    std::vector<float> predict(const std::vector<float>& x) const;
    // (loss, gradient) on host samples x[n][17], t[n][2]
    std::pair<double, std::vector<float>> lossGrad(const std::vector<float>& x,
                                                   const std::vector<float>& t) const;
    sf_mlp* handle() const { return net_; }

private:
    sf_mlp* net_ = nullptr;
};

// Simulation (engine/simulation.h:37-246): construct, add capsules,
// initialize, step/run, accessors.
class Simulation {
public:
    explicit Simulation(const SimulationParams& params);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void setChannelBuilder(const ChannelBuilder& b);
    void addCapsule(Vec2d center, Real radius, int num_nodes, const MembraneParams& mp, int type = 0);
    int addCapsuleRandom(int count, Real x0, Real y0, Real x1, Real y1, Real rmin, Real rmax,
                         int num_nodes, const MembraneParams& mp, int type = 0, int seed = 12345,
                         Real min_gap = 1.0, int max_attempts = 0);
    void initialize();
    void step();
    void run(int num_steps);
    // Simulation::finalize (simulation.cpp:801-812): outstanding output
    // written, .pvd collections, writers closed; run() ends with it
    void finalize();

    // output configuration (simulation.h:180-200)
    void setOutputConfig(const OutputConfig& config);
    void setVTKOutput(int dump_every, const std::string& format, const FluidOutputFields& fluid_fields,
                      const ParticleVTKFields& particle_fields);
    void setCSVOutput(int dump_every, const std::string& format, const ParticleOutputFields& fields,
                      const ParticleFilter& filter, bool append = true);
    void addExtraCSVOutput(const std::string& filename, int dump_every, const ParticleOutputFields& fields,
                           const ParticleFilter& filter);
    void addFluidProbe(int i, int j, const std::string& label);

    int currentStep() const { return static_cast<int>(step_count_); }
    Real getMaxSpeed();
    int lastCappedNodes();
    bool checkStability();
    std::vector<double> fieldStats();
    // the quantities of GlobalStatsWriter::writeTimestep (io/global_stats_writer.cpp:41-124),
    // from FP64 device reductions; scalar masses from the scalar mirror
    std::vector<std::pair<std::string, double>> globalStats();
    LatticeField& field() { return *field_; }
    CapsuleSystem& capsules() { return caps_; }
    AdvectionDiffusion* advectionDiffusion() { return adr_.get(); }
    const SimulationParams& params() const { return params_; }
    void setStepCallback(std::function<void(Simulation&, int)> cb) { cb_ = std::move(cb); }
    void setFluidRegion(int x0, int y0, int x1, int y1, Real rho, Real ux, Real uy);
    void setScalarRegion(int x0, int y0, int x1, int y1, Real c, int species = 0);
    bool initialized() const { return ctx_ != nullptr; }
    std::vector<std::pair<std::string, double>> kernelTimes();
    void setTiming(bool on);
    void synchronize();
    int64_t graphCaptures();
    std::string autotune(int reps = 10);
    void setLbmLaunch(int bx, int by, bool xcd_banded, int persist = 0);
    void setLbmPairs(bool on);
    // spread grid (0: one workgroup per tile) and interpolation lanes per node
    void setCapsuleLaunch(int spread_grid, int interp_lanes);

    // multi-direct forcing control and the learned warm start (extensions)
    void setMdfControl(Real tol, int max_corrections, bool probe);
    // residual mode: iterations from `first` on in one launch (0: one launch pair each)
    void setMdfTail(int first);
    std::vector<double> mdfStats();
    void collectMdfSamples(int64_t capacity);
    int64_t mdfSampleCount();
    std::vector<float> mdfSamples();
    void setMdfModel(MdfModel* model, bool warm);
    sf_ctx* deviceContext() const { return ctx_; }

    // Checkpoint::save / load (io/checkpoint.cpp:88-369): the .sfck format,
    // version 4, from and into the device state, plus an extension block for
    // bit-exact resume.  false (with a message on stderr) on failure, as the
    // reference returns.
    bool saveCheckpoint(const std::string& filename);
    bool loadCheckpoint(const std::string& filename);
    void setCurrentStep(int step) { step_count_ = step; }

    // device <-> host synchronisation (used by the mirrors)
    void pullField(LatticeField& f);
    void pullNodes(CapsuleSystem& cs);
    void pullScalar(AdvectionDiffusion& a);

private:
    friend class Ensemble;
    void pushDirty();
    void markAdvanced(int steps);   // device state moved on: host mirrors are stale
    void afterAdvance();
    bool writeCheckpoint(const std::string& filename, int step);
    void intervalCheckpoint(int64_t step);
    void check(int rc, const char* what) const;
    SimulationParams params_;
    std::unique_ptr<LatticeField> field_;
    CapsuleSystem caps_;
    std::unique_ptr<AdvectionDiffusion> adr_;
    std::unique_ptr<ChannelBuilder> builder_;
    std::function<void(Simulation&, int)> cb_;
    sf_ctx* ctx_ = nullptr;
    int64_t step_count_ = 0;
    // MDF settings kept for (re)initialisation
    Real mdf_tol_ = 0.0;
    int mdf_max_ = 0;
    bool mdf_probe_ = false;
    int mdf_tail_ = 6;
    int64_t mdf_collect_ = 0;
    MdfModel* mdf_model_ = nullptr;
    bool mdf_warm_ = false;
    // raw force of the last step read from a checkpoint, set after the other state
    std::vector<double> pending_force_;
    std::unique_ptr<OutputManager> out_;
    void output(int64_t step);       // the writers due after step index `step`
};

// A batched ensemble (no counterpart in the reference): independent
// simulations of uneven size advanced together, one launch per kernel of the
// step over all members, six steps per captured graph.  Each member ends
// bit-identical to running it alone.  The simulations must be initialised
// x-periodic channels with deformable capsules, one IBM pass and no scalar,
// sharing precision and collision model; they stay usable on their own.
class Ensemble {
public:
    explicit Ensemble(const std::vector<Simulation*>& members);
    ~Ensemble();
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    void run(int num_steps);
    void synchronize();
    std::string autotune(int reps = 10);
    void setLbmLaunch(int bx, int by, bool xcd_banded);
    void setLbmFlat(bool on);
    void setGraphMode(bool on);
    // members, lattice cells, membrane nodes, collide-stream workgroups
    std::vector<int64_t> info();
    int size() const { return static_cast<int>(members_.size()); }

private:
    void check(int rc, const char* what) const;
    std::vector<Simulation*> members_;
    sf_batch* batch_ = nullptr;
};

}  // namespace softflow
