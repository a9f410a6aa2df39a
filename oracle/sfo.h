/* Oracle C-ABI: the CPU restatement of SoftFlow's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product path (softflow_amd/).
 *
 * The arithmetic follows the reference engine (ramc77/SoftFlow) operation by
 * operation, in IEEE double with no FMA contraction, single-threaded, so that
 * it reproduces the reference's bits when both are built with
 * -O2 -ffp-contract=off (the reference's OpenMP paths at 1 thread).
 */
#ifndef SFO_H
#define SFO_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    nx, ny;
    double dt;
    int    ibm_iterations;
    /* fluid */
    double tau, rho0;
    int    boundary_type;      /* 0 inlet/outlet, 1 periodic, 2 closed */
    double body_force_x, body_force_y;
    double gravity_x, gravity_y;
    int    apply_gravity_to_fluid, apply_gravity_to_capsules;
    int    collision_model;    /* 0 BGK, 1 MRT, 2 regularized */
    int    use_mrt;
    double mrt_se, mrt_s_eps, mrt_sq;
    int    periodic_y;
    double max_lattice_force;
    /* repulsion */
    double rep_epsilon, rep_sigma, rep_r_cut;
    int    rep_power;
    double rep_damping_normal, rep_friction_coeff;
    /* scalar transport */
    int    scalar_enabled, n_species;
    double diffusivity[4];
    double inlet_concentration[4];
    int    scalar_periodic_y;
    /* moving walls: x velocity of the solid rows y = ny-1 and y = 0 */
    double top_wall_velocity, bottom_wall_velocity;
    /* Zou-He inlet/outlet (boundary_type 0) */
    double inlet_velocity, outlet_density;
    /* Bouzidi links towards the obstacles (parameters.h:89) */
    int    use_interpolated_bb;
    /* capsules recirculate in x over a non-periodic fluid (simulation.cpp:687-700) */
    int    capsule_periodic_x;
} sfo_params;

typedef struct {
    int    model;              /* 0 Hookean, 1 neo-Hookean, 2 Skalak, 3 WLC */
    double k_stretch, G_s, C_skalak;
    double wlc_L_max_ratio, wlc_kBT_p, wlc_k_pow;
    double k_bend;
    int    use_helfrich;
    double kappa_0;
    double k_area, k_perimeter, gamma_visc, eta_membrane;
    double viscosity_ratio, density;
    int    shape;              /* 0 circle, 1 ellipse, 2 biconcave, 3 fiber */
    double aspect_ratio, indent_depth;
    int    is_rigid;
} sfo_membrane;

typedef struct sfo_sim sfo_sim;

void   sfo_default_params(sfo_params* p);
void   sfo_default_membrane(sfo_membrane* m);

sfo_sim* sfo_create(const sfo_params* p);
void   sfo_destroy(sfo_sim* s);
void   sfo_add_capsule(sfo_sim* s, double cx, double cy, double radius,
                       int num_nodes, const sfo_membrane* m, int type);
int    sfo_add_capsule_random(sfo_sim* s, int count, double x0, double y0,
                              double x1, double y1, double rmin, double rmax,
                              int num_nodes, const sfo_membrane* m, int type,
                              int seed, double min_gap, int max_attempts);
/* An obstacle of the channel builder (geometry/{circle,rect,polygon}_obstacle.h):
 * kind 0 circle (cx, cy, r), 1 rectangle (x0, y0, x1, y1), 2 polygon (x, y)
 * per vertex; n = number of coordinates.  Added before sfo_initialize, which
 * marks the cells it contains SOLID (channel_builder.cpp:27-33). */
void   sfo_add_obstacle(sfo_sim* s, int kind, const double* coords, int n);
/* Obstacle::signedDistance / normalAt (out[2]) of obstacle i. */
double sfo_obstacle_distance(const sfo_sim* s, int i, double x, double y, double* normal);
/* The interpolated bounce-back links built by sfo_initialize: count, and
 * (x, y, q) + wall distance of each when the arrays are not NULL. */
int    sfo_ibb_links(const sfo_sim* s, int* xyq, double* qf);
void   sfo_initialize(sfo_sim* s);
void   sfo_scalar_initialize(sfo_sim* s, double c0);
void   sfo_set_scalar_point(sfo_sim* s, int x, int y, double c, int species);
void   sfo_step(sfo_sim* s, int nsteps);

int    sfo_num_capsules(const sfo_sim* s);
int    sfo_total_nodes(const sfo_sim* s);
void   sfo_capsule_offsets(const sfo_sim* s, int* off /* ncaps+1 */);
void   sfo_get_nodes(const sfo_sim* s, double* x, double* y, double* vx,
                     double* vy, double* fx, double* fy);
void   sfo_set_nodes(sfo_sim* s, const double* x, const double* y,
                     const double* vx, const double* vy);
void   sfo_get_field(const sfo_sim* s, double* rho, double* ux, double* uy,
                     double* fx, double* fy);
void   sfo_get_f(const sfo_sim* s, double* f /* 9*N, q-major */);
void   sfo_set_f(sfo_sim* s, const double* f);
void   sfo_set_macro(sfo_sim* s, const double* rho, const double* ux,
                     const double* uy);
void   sfo_get_flags(const sfo_sim* s, int* flags);
void   sfo_set_flags(sfo_sim* s, const int* flags);
void   sfo_get_scalar(const sfo_sim* s, double* C, int species);
/* Overwrite the concentration array only (what a write through the
 * reference's concentration() view does, python/bindings.cpp:435-440): the
 * next collision relaxes g towards w C (1 + 3 e.u) with it. */
void   sfo_set_concentration(sfo_sim* s, const double* C, int species);
int    sfo_last_capped(const sfo_sim* s);
/* The step's force cap (simulation.cpp:588-609) on N cells: returns the
 * number of cells rescaled. */
int    sfo_cap_forces(double* Fx, double* Fy, int N, double max_force);
/* The step's node clamp into [1.5, ny - 1.5] (simulation.cpp:665-675). */
void   sfo_clamp_nodes_y(double* y, int K, int ny);
/* The step's capsule recirculation in x (simulation.cpp:687-700): a capsule
 * whose centroid (plain node mean) left [0, nx) moves rigidly by +-nx.
 * off[ncaps+1] are the capsules' node offsets; returns the capsules moved. */
int    sfo_recirculate_x(double* x, const double* y, const int* off, int ncaps, int nx);
double sfo_max_speed(const sfo_sim* s);
double sfo_capsule_geometry(const sfo_sim* s, int c, int what);
void   sfo_rest_lengths(const sfo_sim* s, double* out);

/* single-operation probes used by unit tests */
double sfo_phi(double r);
double sfo_phi_exact(double r);
void   sfo_collide_cell(int model, double tau, double se, double seps,
                        double sq, double* f9, double rho, double ux,
                        double uy, double Fx, double Fy);
void   sfo_membrane_forces_only(sfo_sim* s);
/* RepulsionForce::pairForce (repulsion.cpp:6-38) for the sim's contact
 * parameters; out[2] = force on a. */
void   sfo_pair_force(const sfo_sim* s, const double* pa, const double* va,
                      const double* pb, const double* vb, double* out);

/* MDF control of the device engine (no reference counterpart): tol > 0 stops
 * the corrections once max_k |V_k - U(X_k)| <= tol (at most max_corrections);
 * stats out4: mean corrections, mean and max final residual, steps. */
void   sfo_set_mdf_control(sfo_sim* s, double tol, int max_corrections, int probe);
void   sfo_mdf_stats(sfo_sim* s, double* out4);

/* The device engine's warm-start network restated in FP64 (no reference
 * counterpart): 17 -> width -> width -> 2, GELU. */
int    sfo_mlp_num_params(int width);
void   sfo_mlp_predict(const double* P, const double* norm, int width, int n,
                       const double* x, double* y);
double sfo_mlp_loss_grad(const double* P, const double* norm, int width, int n,
                         const double* x, const double* t, double* grad);

#ifdef __cplusplus
}
#endif
#endif
