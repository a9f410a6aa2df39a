/* softflow.h -- C-ABI of the MI355X (gfx950) per-timestep engine.
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success or a
 * negative SF_E* code, and leaves a message for sf_last_error().  Host buffers
 * are borrowed only for the duration of the call; device state is owned by the
 * context.  Layouts at this boundary are the reference engine's:
 *   populations  f[q*N + y*nx + x], q in D2Q9 order (lattice.h:22-27), as
 *                left by LBMSolver::step (post-stream, post-bounce-back) in
 *                every non-SOLID cell; SOLID cells keep what was last written
 *                there (the reference's pull stream copies neighbour values
 *                into them, which no stage reads: their rho and u are 0);
 *   fields       rho/ux/uy/Fx/Fy[y*nx + x];
 *   scalar       g[n*9 + q] (advection_diffusion.h:149) and C[n];
 *   nodes        capsule-major flat arrays with CSR offsets off[ncaps+1].
 *
 * Citations name the reference interface each entry point replaces
 * (ramc77/SoftFlow, path:line).
 */
#ifndef SOFTFLOW_H
#define SOFTFLOW_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    SF_OK = 0,
    SF_E_ARG = -1,        /* bad argument / unsupported configuration */
    SF_E_HIP = -2,        /* HIP runtime error */
    SF_E_STATE = -3,      /* call out of order (e.g. step before set_lattice) */
    SF_E_CAPACITY = -4    /* a fixed-capacity device list overflowed */
};

enum { SF_FP64 = 0, SF_FP32 = 1 };            /* lattice storage / collide precision */
enum { SF_BGK = 0, SF_MRT = 1, SF_REGULARIZED = 2 };
enum { SF_INLET_OUTLET = 0, SF_PERIODIC = 1, SF_CLOSED = 2 };
enum { SF_FLUID = 0, SF_SOLID = 1, SF_INLET = 2, SF_OUTLET = 3, SF_EMPTY = 4 };

/* Subset of SimulationParams (core/parameters.h:279-342) that the step reads. */
typedef struct {
    int    nx, ny;
    double dt;
    int    ibm_iterations;          /* parameters.h:286 */
    double tau, rho0;
    int    boundary_type;           /* SF_PERIODIC, SF_CLOSED or SF_INLET_OUTLET */
    double body_force_x, body_force_y;
    int    collision_model;         /* SF_BGK / SF_MRT / SF_REGULARIZED */
    double mrt_se, mrt_s_eps, mrt_sq;
    int    periodic_y;
    double max_lattice_force;       /* parameters.h:337 */
    double rep_epsilon, rep_sigma, rep_r_cut;
    int    rep_power;
    double rep_damping_normal, rep_friction_coeff;
    int    scalar_enabled, n_species;
    double diffusivity[4];
    double inlet_concentration[4];
    int    scalar_periodic_y;
    int    precision;               /* SF_FP64 or SF_FP32 */
    double top_wall_velocity;       /* moving walls (lbm/moving_wall.cpp): x velocity of the */
    double bottom_wall_velocity;    /* solid rows y = ny-1 and y = 0 */
    double inlet_velocity;          /* SF_INLET_OUTLET (lbm/zou_he_boundary.cpp): peak of the */
    double outlet_density;          /* parabolic inlet profile; outlet density */
    int    use_interpolated_bb;     /* parameters.h:89: Bouzidi links towards the obstacles */
    int    capsule_periodic_x;      /* parameters.h: capsules recirculate in x over a non-periodic
                                       fluid (Simulation::step 10a, simulation.cpp:687-700) */
} sf_params;

/* Per-capsule constants resolved on the host (rest geometry and WLC constants
 * as computed by Capsule::Capsule, membrane/capsule.cpp:7-141). */
typedef struct {
    int    model;                   /* 0 Hookean, 1 neo-Hookean, 2 Skalak, 3 WLC */
    int    use_helfrich;
    int    is_rigid;
    int    type;
    double k_stretch, G_s, C_skalak, k_bend;
    double k_area, k_perimeter, gamma_visc, eta_membrane;
    double wlc_L_max_ratio, wlc_kBT_p, wlc_k_pow;
    double rest_area, rest_perimeter;
} sf_capsule_desc;

typedef struct sf_ctx sf_ctx;

const char* sf_last_error(void);
int  sf_device_count(int* n);

/* Context = one Simulation's device state (simulation.h:37, Simulation ctor). */
int  sf_create(sf_ctx** ctx, int device, const sf_params* p);
int  sf_destroy(sf_ctx* ctx);

/* Cell types (LatticeField::cellTypeData, lattice_field.h:87).  flags[N]:
 * FLUID, SOLID, INLET, OUTLET (EMPTY is refused). */
int  sf_set_flags(sf_ctx* ctx, const int32_t* flags);

/* Obstacles (geometry/obstacle.h and its three shapes), as the channel builder
 * holds them (ChannelBuilder::getObstacles, channel_builder.h:52-54): they act
 * in the contact forces (RepulsionForce::computeObstacleRepulsion,
 * membrane/repulsion.cpp:180-236, after the wall images, in list order) and,
 * with use_interpolated_bb, place the wall of every link from a non-SOLID cell
 * into a SOLID cell (InterpolatedBounceBack::initialize,
 * lbm/interpolated_bounce_back.cpp:7-53).  They do not change the cell types:
 * the caller sets the SOLID cells (ChannelBuilder::applyToField) through
 * sf_set_flags.  The links are rebuilt from the current cell types whenever the
 * obstacles or the cell types change.
 *   kinds[i]: SF_OBS_CIRCLE   coords (cx, cy, r)            (circle_obstacle.h)
 *             SF_OBS_RECT     coords (x0, y0, x1, y1)       (rect_obstacle.h)
 *             SF_OBS_POLYGON  coords (x, y) per vertex, >= 3 (polygon_obstacle.h)
 *   obstacle i takes coords[offsets[i] .. offsets[i+1]).  count 0 clears them.
 * A link whose wall lies nearer than half a cell and whose upstream cell is
 * SOLID (a fluid gap one cell wide) is refused with SF_E_ARG. */
enum { SF_OBS_CIRCLE = 0, SF_OBS_RECT = 1, SF_OBS_POLYGON = 2 };
int  sf_set_obstacles(sf_ctx* ctx, int count, const int32_t* kinds,
                      const int32_t* offsets, const double* coords);

/* Populations in the reference layout (LatticeField::fData, lattice_field.h:41). */
int  sf_set_populations(sf_ctx* ctx, const double* f);
int  sf_get_populations(sf_ctx* ctx, double* f);

/* Macroscopic arrays as LatticeField holds them after a step
 * (computeMacroscopic, lattice_field.cpp:28-73) and the force field of the
 * last step (Fx/Fy after body force and cap, simulation.cpp:541-609).
 * Any pointer may be NULL. */
int  sf_get_macroscopic(sf_ctx* ctx, double* rho, double* ux, double* uy,
                        double* Fx, double* Fy);
/* Overwrite the stored rho/ux/uy that the next collision reads (the
 * reference's setRho/setUx/setUy, lattice_field.h:94-96). */
int  sf_set_macroscopic(sf_ctx* ctx, const double* rho, const double* ux,
                        const double* uy);

/* Capsules (CapsuleSystem::addCapsule, capsule_system.cpp:5-9): node arrays
 * of K = off[ncaps] entries; prev_len/has_prev carry the Kelvin-Voigt state
 * (capsule.h:121-122). rest_angle/rest_ds may be NULL without Helfrich. */
int  sf_set_capsules(sf_ctx* ctx, int ncaps, const int32_t* off,
                     const sf_capsule_desc* desc, const double* x,
                     const double* y, const double* vx, const double* vy,
                     const double* rest_len, const double* rest_angle,
                     const double* rest_ds, const double* prev_len,
                     const int32_t* has_prev);
int  sf_get_nodes(sf_ctx* ctx, double* x, double* y, double* vx, double* vy,
                  double* fx, double* fy);
int  sf_set_nodes(sf_ctx* ctx, const double* x, const double* y,
                  const double* vx, const double* vy);
int  sf_get_membrane_state(sf_ctx* ctx, double* prev_len, int32_t* has_prev);
/* Checkpoint extension (no counterpart in the reference, whose .sfck files keep
 * f, rho/u, flags and nodes, io/checkpoint.cpp:88-190): the raw IBM force of the
 * last step, (Fx, Fy) pairs per cell, from which the next collision's velocity
 * is computed.  Setting it makes the next step use it instead of a stored
 * macroscopic field, so a run resumed from a checkpoint continues bit for bit. */
int  sf_get_force_state(sf_ctx* ctx, double* F);
int  sf_set_force_state(sf_ctx* ctx, const double* F);

/* Scalar transport (AdvectionDiffusion, advection_diffusion.h:17-99). */
int  sf_set_scalar(sf_ctx* ctx, int species, const double* g, const double* C);
int  sf_get_scalar(sf_ctx* ctx, int species, double* C, double* g);

/* Simulation::step (simulation.cpp:470-787) nsteps times, in-scope stages. */
int  sf_step(sf_ctx* ctx, int nsteps);
/* Replay the step as a captured HIP graph (1) or launch kernel by kernel (0). */
int  sf_set_graph_mode(sf_ctx* ctx, int on);
int  sf_synchronize(sf_ctx* ctx);
/* Step graphs captured and instantiated by this context so far (no
 * counterpart in the reference): a timed run whose count does not change
 * replayed graphs only. */
int  sf_graph_captures(sf_ctx* ctx, int64_t* n);

/* Stage-level entry points of one step, in simulation.cpp:470-787 order.
 * SF_STAGE_FORCES  = clearAllForces + computeAllMembraneForces + repulsion
 * SF_STAGE_IBM     = spreadForces or multiDirectForcing (+ body force, cap)
 * SF_STAGE_LBM     = LBMSolver::step
 * SF_STAGE_SCALAR  = AdvectionDiffusion::step
 * SF_STAGE_MOVE    = interpolateVelocity + moveAllNodes + y-clamp */
enum { SF_STAGE_FORCES = 0, SF_STAGE_IBM = 1, SF_STAGE_LBM = 2,
       SF_STAGE_SCALAR = 3, SF_STAGE_MOVE = 4 };
int  sf_stage(sf_ctx* ctx, int stage);

/* last_capped_nodes_ (simulation.cpp:588-609). */
int  sf_last_capped(sf_ctx* ctx, int32_t* capped);
/* FP64 device reductions over FLUID cells of the current fields:
 * out[0] = sum rho, out[1] = sum rho*ux, out[2] = sum rho*uy,
 * out[3] = max |u| (getMaxSpeed, simulation.cpp:892-907),
 * out[4] = number of non-finite cells (checkStability, :820-890). */
int  sf_field_stats(sf_ctx* ctx, double* out5);
/* The same five values as sf_field_stats, computed without allocating:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid
 * allocation).  The reduction computes each cell's rho and u in place and
 * keeps its partial sums in a fixed device array of 5 * 256 + 8 doubles
 * (10.3 KiB) per device, held until the process ends; lattices of up to
 * 2^31 - 1 cells, larger ones are refused with SF_E_ARG.  Results are
 * identical to sf_field_stats'.  Calls that overlap on one device share that
 * array and race: serialise them. */
int  sf_field_stats_noalloc(sf_ctx* ctx, double* out5);
/* FP64 device reductions over the non-solid cells for GlobalStatsWriter
 * (io/global_stats_writer.cpp:55-80) and checkStability (simulation.cpp:820-890):
 * out[0] cells, [1] sum rho, [2] min rho, [3] max rho, [4] max |u|^2,
 * [5] sum |u|, [6] sum rho |u|^2 / 2, [7] non-finite cells (left out of the
 * rest), [8] the density farthest from rho0, [9] sum rho ux, [10] sum rho uy. */
int  sf_global_stats(sf_ctx* ctx, double rho0, double* out11);
int  sf_step_count(sf_ctx* ctx, int64_t* n);

/* Output snapshots for the writers (the reference's writers read LatticeField,
 * CapsuleSystem and AdvectionDiffusion on the host between steps,
 * simulation.cpp:736-771; this is how they get the same data from the device).
 * sf_snapshot_async enqueues on the context's stream, behind the steps already
 * enqueued, the requested state into the caller's host buffers and returns a
 * ticket at once; steps enqueued afterwards leave the copies intact.  Buffers
 * from sf_host_alloc (pinned) keep the copies off the host's critical path.
 * sf_snapshot_wait blocks until a ticket's copies have landed (the buffers must
 * not be reused before).  Any pointer may be NULL.
 *   rho..Fy   N each, as sf_get_macroscopic gives them;
 *   C[s]      N each, the concentration of species s (NULL for the others);
 *   x..fy     K each, as sf_get_nodes gives them;
 *   probes    3 per probe cell (rho, ux, uy at the cells of sf_set_probes,
 *             FluidProbeWriter::writeTimestep, fluid_probe_writer.cpp:46-68);
 *   stats     12 values: the 11 of sf_global_stats, then the concentration
 *             summed over every cell and species (global_stats_writer.cpp:86-97). */
typedef struct {
    double *rho, *ux, *uy, *Fx, *Fy;
    double *C[4];
    double *x, *y, *vx, *vy, *fx, *fy;
    double *probes;
    double *stats;
    double  rho0;
} sf_snapshot;
int  sf_host_alloc(void** p, uint64_t bytes);
int  sf_host_free(void* p);
int  sf_set_probes(sf_ctx* ctx, int n, const int32_t* cells);
int  sf_snapshot_async(sf_ctx* ctx, const sf_snapshot* s, int64_t* ticket);
int  sf_snapshot_wait(sf_ctx* ctx, int64_t ticket);

/* Per-kernel timing of the last sf_step call, measured with HIP events on the
 * context's stream: names[i] (static strings) and mean ms per launch. */
int  sf_kernel_times(sf_ctx* ctx, int max, const char** names, double* ms,
                     int* count);
int  sf_set_timing(sf_ctx* ctx, int on);

/* Autotuner (no counterpart in the reference): times the collide-stream
 * kernel in place over launch shapes (block width x height, XCD-banded or
 * linear workgroup order, one-shot or a persistent grid) on the current state
 * -- the kernel only reads its source buffers, so repeated launches leave the
 * state unchanged -- and keeps the fastest.  report (may be NULL) receives one
 * line per candidate. */
int  sf_autotune(sf_ctx* ctx, int reps, char* report, int report_len);
/* Fix the collide-stream launch shape instead (bx*by a multiple of 64, <= 256);
 * persist > 0 launches 8 * persist workgroups that walk the XCD bands. */
int  sf_set_lbm_launch(sf_ctx* ctx, int bx, int by, int xcd_banded, int persist);
/* Two cells per lane in the collide-stream (no counterpart in the reference).
 * The paired kernel applies to FP32 lattices, x periodic without x walls, no
 * moving walls, nx even.  By default it runs on channels and one cell per lane
 * runs elsewhere; on = 1 also gives it generic cell types, on = 0 turns it
 * off.  sf_autotune times both forms and keeps the faster.  Results are
 * identical either way. */
int  sf_set_lbm_pairs(sf_ctx* ctx, int on);
/* Capsule-side launch choices (no counterpart in the reference): the spread
 * gather's grid (0: one workgroup per lattice tile, else that many persistent
 * workgroups walking the touched tiles) and the lanes per node of the
 * interpolations (16, 8 or 4).  sf_autotune times both; results are
 * identical for every choice. */
int  sf_set_capsule_launch(sf_ctx* ctx, int spread_grid, int interp_lanes);

/* ---- Multi-direct forcing control and the learned warm start ------------
 * (no counterpart in the reference; its MDF is a fixed iteration count,
 * immersed_boundary.cpp:231-286, which stays the default).
 *
 * Residual mode (tol > 0): after each MDF interpolation the no-slip residual
 * r = max_k |V_k - U(X_k)| is taken on the device; corrections stop once
 * r <= tol, after at most max_corrections.  tol <= 0 keeps the reference's
 * ibm_iterations - 1 corrections.  probe = 1 measures r once more after the
 * final correction (for statistics).  Residual mode runs MDF even when
 * ibm_iterations is 1. */
int  sf_mdf_set_control(sf_ctx* ctx, double tol, int max_corrections, int probe);
/* Residual mode: MDF iterations first .. max_corrections run in one launch
 * (default first = 6; 0: every iteration is its own launch pair).  That
 * launch leaves at once when the loop has stopped before first, so a cap well
 * above the corrections a step needs costs one launch instead of two per
 * unused iteration; the iterations it runs give the same results bit for bit.
 * No counterpart in the reference (an extension of sf_mdf_set_control). */
int  sf_mdf_set_tail(sf_ctx* ctx, int first);
/* out4: mean corrections per step, mean and max final residual, steps; since
 * the last call (which resets them). */
int  sf_mdf_stats(sf_ctx* ctx, double* out4);
/* Record (inputs, converged correction) rows of every node at every step into
 * a device sample set of up to capacity rows (0 frees it).  A row holds
 * SF_MLP_IN inputs then SF_MLP_OUT targets, FP32. */
enum { SF_MLP_IN = 17, SF_MLP_OUT = 2, SF_MLP_WIDTH = 64 };
int  sf_mdf_collect(sf_ctx* ctx, int64_t capacity);
/* rows written so far; rows_out (may be NULL) receives them. */
int  sf_mdf_samples(sf_ctx* ctx, int64_t* rows, float* rows_out);

/* The warm-start network: SF_MLP_IN -> 64 -> 64 -> SF_MLP_OUT with GELU,
 * FP32, with input/output normalisation (in_mean[17], in_scale[17],
 * out_mean[2], out_scale[2]).  Parameters: W1[64][17], b1[64], W2[64][64],
 * b2[64], W3[2][64], b3[2]. */
typedef struct sf_mlp sf_mlp;
int  sf_mlp_create(sf_mlp** net, int device, uint64_t seed);
int  sf_mlp_destroy(sf_mlp* net);
int  sf_mlp_num_params(int* n_params, int* n_norm);
int  sf_mlp_get_params(sf_mlp* net, float* params, float* norm);
int  sf_mlp_set_params(sf_mlp* net, const float* params, const float* norm);
/* y[n][2] = network(x[n][17]): out_mean + out_scale * net((x - in_mean) / in_scale). */
int  sf_mlp_predict(sf_mlp* net, int n, const float* x, float* y);
/* Loss L = 1/(2n) sum |out - t^|^2 in normalised units and its gradient with
 * respect to the parameters, for host samples x[n][17], t[n][2]. */
int  sf_mlp_loss_grad(sf_mlp* net, int n, const float* x, const float* t, double* loss,
                      float* grad);
/* Adam (beta 0.9/0.999) over the context's sample set: epochs passes of
 * shuffled mini-batches; fits the normalisation first if none was set.
 * epoch_loss[epochs] (may be NULL) receives the mean loss of each pass. */
int  sf_mlp_train(sf_mlp* net, sf_ctx* ctx, int epochs, int batch, float lr, uint64_t seed,
                  double* epoch_loss);
/* Use net for the first MDF correction of every step (warm = 1); the plain
 * corrections follow as needed.  net is borrowed and must outlive the use. */
int  sf_mdf_set_model(sf_ctx* ctx, sf_mlp* net, int warm);

/* Batched ensembles (no counterpart in the reference, whose members would each
 * be one Simulation stepped on its own, simulation.cpp:470-787).  A batch
 * borrows n contexts on one device and advances all of them with one launch
 * per kernel of the step, each launch covering the workgroups of every member;
 * six steps (the period of the buffer rotation) are captured as one hipGraph.
 * Every member's results are bit-identical to sf_step on that context alone.
 * Members must share precision and collision model and be x-periodic plain
 * channels (no obstacle flags) with deformable capsules.  Members may carry an
 * advection-diffusion scalar (its lagged step rides in the batched
 * collide-stream, as in sf_step) or apply multi-direct forcing, fixed or
 * residual-controlled (each member its own count and tolerance), but not both,
 * and not the learned warm start or sample collection.  The contexts must outlive the batch; calls on a member
 * between batch steps are allowed (the batch re-plans when a member's buffers
 * or roles changed). */
typedef struct sf_batch sf_batch;
int  sf_batch_create(sf_batch** out, sf_ctx* const* members, int n);
int  sf_batch_destroy(sf_batch* b);
int  sf_batch_step(sf_batch* b, int nsteps);
int  sf_batch_synchronize(sf_batch* b);
/* 0: launch every step directly; 1 (default): replay captured six-step graphs. */
int  sf_batch_set_graph_mode(sf_batch* b, int on);
/* Autotuner: times the batched collide-stream over workgroup shapes and
 * linear or XCD-banded orders of the members' workgroups (repeated launches
 * leave the state unchanged) and keeps the fastest; report may be NULL. */
int  sf_batch_autotune(sf_batch* b, int reps, char* report, int report_len);
int  sf_batch_set_lbm_launch(sf_batch* b, int bx, int by, int xcd_banded);
/* on = 1: the batched collide-stream's workgroups take consecutive runs of
 * bx * by cells (pairs) of each member's row-major cell index, across row ends,
 * instead of bx x by rectangles, so rows narrower than a workgroup leave no
 * lanes idle; results are identical either way (the autotuner times both). */
int  sf_batch_set_lbm_flat(sf_batch* b, int on);
/* out[0] members, out[1] lattice cells, out[2] membrane nodes, out[3] workgroups
 * of the batched collide-stream launch. */
int  sf_batch_info(sf_batch* b, int64_t* out4);

/* Device math probe used by the tests: op 0 sqrt, 1 divide (a/b), 2 atan2(a,b),
 * 3 pow(a, b) with integer b as used by the contact force. */
int  sf_probe_math(int op, int n, const double* a, const double* b, double* out);

/* Achievable HBM bandwidth on this device (no counterpart in the reference):
 * copies `bytes` between two device buffers it allocates, 16 bytes per thread
 * with nontemporal loads and stores, reps times after one warm launch, and
 * writes (bytes read + bytes written) / time in GB/s to *gbps.  bytes: a
 * multiple of 16 in [1 MiB, 8 GiB]. */
int  sf_copy_ceiling(int device, uint64_t bytes, int reps, double* gbps);

#ifdef __cplusplus
}
#endif
#endif
