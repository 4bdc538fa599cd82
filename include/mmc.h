/*
 * mmc.h — C-ABI boundary of the gfx950 beam-lattice compression solver.
 *
 * The reference has no FFI: its boundary is the C++ class meta::Simulator
 * (projects/metamaterial-compression/industry-grade/include/sim.hpp:35-54),
 * whose private hot path is objective()/minimize()/commit()/state()
 * (sim.hpp:48-53, src/sim.cpp:136-212).  This header is the thin C layer that
 * a C++20 Simulator facade with the reference's public signatures calls into.
 * Plain pointers and sizes only; every entry point returns an int status
 * (MMC_OK = 0) and never throws; the message of the last failure on a context
 * is returned by mmc_last_error().
 *
 * Paths below abbreviate projects/metamaterial-compression/industry-grade/ as MC/.
 */
#ifndef MMC_H
#define MMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
  MMC_OK = 0,
  MMC_ERR_ARG = 1,         /* bad argument / size mismatch                       */
  MMC_ERR_HIP = 2,         /* HIP runtime failure (message names the call)       */
  MMC_ERR_STATE = 3,       /* e.g. commit without trial states (MC/src/sim.cpp:212) */
  MMC_ERR_EQUILIBRIUM = 4, /* equilibrium failed after bisection (MC/src/sim.cpp:412-418) */
  MMC_ERR_CAPACITY = 5     /* a device buffer bound was exceeded                 */
};

/* ---- precision policies -------------------------------------------------- */
enum {
  MMC_FP64 = 0,  /* every element operation in FP64 (the reference's arithmetic) */
  MMC_MIXED = 1  /* FP32 rotational element math; FP64 positions, strains, return
                    maps, plastic state, L-BFGS vectors/dots and line search     */
};

/* ---- lattice families ---------------------------------------------------- */
enum {
  MMC_LATTICE_BCC = 0,   /* graded BCC accordion lattice of MC/src/sim.cpp:51-132 */
  MMC_LATTICE_CUBIC = 1  /* simple cubic lattice, one element per cell edge       */
};

/* Material parameters: MaterialConfig, MC/include/sim.hpp:16. */
typedef struct mmc_material {
  double density, young, nu, yield, hardening, bend_hardening, torsion_hardening;
} mmc_material;

/* Solver configuration: SimulationConfig, MC/include/sim.hpp:17-26, same field
 * meaning and defaults (mmc_config_defaults), followed by additive fields that
 * have no counterpart in the reference. */
typedef struct mmc_config {
  int32_t cells_x, cells_y, cells_z;
  int32_t elements_per_strut;
  double cell_size, min_radius, max_radius, joint_scale, trigger_scale;
  double imperfection, member_imperfection;
  mmc_material material;
  double gravity, shear_correction, contact_stiffness, contact_margin, centering_stiffness;
  int32_t contact_samples, load_steps, unload_steps, max_iterations;
  int32_t history, max_line_search, max_adaptive_refinements, fps;
  int32_t frames, threads;
  double max_strain, step_time, translation_scale, max_translation_step, max_rotation_step;
  double gradient_rms_tolerance, gradient_max_tolerance, armijo, min_line_step;
  double penetration_tolerance;
  uint64_t seed;
  /* additive: */
  int32_t lattice;            /* MMC_LATTICE_*                                         */
  int32_t relative_escape;    /* 1: accept rms/max(1e-6,ts*max(1,F)) < relative_tolerance
                                 as the reference does (MC/src/sim.cpp:177); 0: off     */
  double relative_tolerance;  /* 1.5e-3 in the reference                                */
  double position_sigma;      /* cubic: Gaussian node jitter std (m); 0 selects the
                                 reference's corner-node imperfection (sim.cpp:63-69)   */
  int32_t radius_mode;        /* cubic: 0 linear grading min->max along y; 1 per-beam
                                 uniform in [min_radius, max_radius]                    */
  int32_t line_search;        /* Armijo energy test: 0 difference of total energies (the
                                 reference, MC/src/sim.cpp:179); 1 sum of per-beam,
                                 per-node and per-contact-shape energy differences (same
                                 test in exact arithmetic, resolves decreases far below
                                 the round-off of the total); 2 the per-term test of 1,
                                 and when the change is within 1e-6 |E| the approximate-
                                 Wolfe slope test g(z + a p).p <= (2 armijo - 1) g(z).p
                                 of Hager & Zhang (SIAM J. Optim. 16, 2005): reaches
                                 gradient tolerances below the energy's resolution      */
} mmc_config;

/* Fills *cfg with the reference defaults of MC/include/sim.hpp:16-26 (BCC lattice,
 * relative escape on). */
void mmc_config_defaults(mmc_config* cfg);

/* A lattice handed across the boundary.  Caller-owned; read only during the call.
 * Node-major AoS for host convenience; the device stores SoA. */
typedef struct mmc_lattice {
  int32_t n_nodes, n_beams;
  const double* x;        /* 3*n_nodes base positions (Node::x, sim.hpp:27)           */
  const double* q;        /* 4*n_nodes base orientations w,x,y,z                      */
  const double* mass;     /* n_nodes lumped masses (distributeMass, sim.cpp:134)       */
  const uint8_t* locked;  /* 6*n_nodes DOF locks (sim.cpp:134)                          */
  const int32_t* beam_a;  /* n_beams                                                  */
  const int32_t* beam_b;  /* n_beams                                                  */
  const double* L0;       /* n_beams rest lengths                                      */
  const double* radius;   /* n_beams                                                   */
  const double* area;     /* n_beams  pi r^2                                           */
  const double* inertia;  /* n_beams  pi r^4 / 4  (J = 2 I)                             */
  const double* e0;       /* 3*n_beams material directors                               */
  const double* plastic;  /* 7*n_beams PlasticState epsp, alpha, rotp[3], alphaB, alphaT */
  int64_t n_connected;    /* node pairs excluded from self-contact (sim.cpp:115,117)    */
  const uint64_t* connected; /* pairKey(a,b) = min<<32 | max (MC/include/core.hpp:47)   */
  const int32_t* node_layer;  /* n_nodes Node::layer (may be NULL)                        */
  const int32_t* beam_layer;  /* n_beams Beam::layer (may be NULL)                        */
  const int32_t* beam_family; /* n_beams Beam::family (may be NULL)                       */
  double width, height, depth; /* specimen envelope (sim.cpp:56)                          */
} mmc_lattice;

/* Objective record: Objective/ObjectiveDiag, MC/include/sim.hpp:31-32, plus the
 * unlocked-DOF gradient norms used by the convergence test (sim.cpp:44-45). */
typedef struct mmc_objective_out {
  double energy, elastic, plastic, contact, reaction, max_pen;
  double rms, max;
  int32_t contacts, finite;
} mmc_objective_out;

/* OptStats, MC/include/sim.hpp:33. */
typedef struct mmc_opt_stats {
  int32_t iterations, backtracks;
  double rms, max, energy;
  int32_t converged, evaluations;
} mmc_opt_stats;

/* Per-frame record: byte layout of DiskDiagnostics, MC/src/main.cpp:24-31
 * (FrameDiagnostics, MC/include/sim.hpp:7-13). */
typedef struct mmc_frame {
  double time, platen_y, displacement, engineering_strain, engineering_stress, reaction_force;
  double loading_work, recovered_work, dissipated_work, specific_energy_absorption;
  double elastic_energy, plastic_dissipation, contact_energy, residual_rms, residual_max,
      max_penetration;
  double mean_plastic_strain, max_plastic_strain, mean_plastic_rotation, max_plastic_rotation,
      permanent_set;
  int32_t active_contacts, plastic_beams, plastic_hinges, iterations;
  uint32_t converged, loading;
} mmc_frame;

/* Per-beam trial record: TrialBeam, MC/include/sim.hpp:30, flattened to 17 doubles:
 * next[7], strain, stress, moment[3], elastic, dissipation, axialPlastic,
 * bendPlastic, torsionPlastic (flags as 0/1). */
#define MMC_TRIAL_STRIDE 17

/* One body of a captured frame: the byte layout of DiskSnapshot,
 * MC/src/main.cpp:19-22, which is the layout of Snapshot, MC/include/sim.hpp:7
 * (position, orientation w,x,y,z, velocity, angular velocity, half extents,
 * render material, id, BodyKind, module class, ring, slot). */
typedef struct mmc_snapshot {
  double p[3], q[4], v[3], w[3], half[3];
  int32_t material, id, kind, module_class, ring, slot;
} mmc_snapshot;
/* Rigid machine bodies that close every captured frame (MC/src/sim.cpp:297-307):
 * a lattice of N nodes and B beams captures B + N + MMC_MACHINE_BODIES bodies. */
#define MMC_MACHINE_BODIES 7

/* Plastic-state and geometry reductions of diagnostics(), MC/src/sim.cpp:339-370. */
typedef struct mmc_state_stats {
  double sum_axial, max_axial, sum_rot, max_rot, y_min, y_max;
  int32_t plastic_beams, plastic_hinges;
} mmc_state_stats;

/* ==== single-lattice context (the hot path) ================================ */
typedef struct mmc_ctx mmc_ctx;

/* Uploads the lattice SoA to device 0's current HIP device and sizes every work
 * buffer.  Replaces the setup half of Simulator::Simulator (MC/src/sim.cpp:49). */
int mmc_create(const mmc_lattice* lat, const mmc_config* cfg, int32_t precision, mmc_ctx** out);
void mmc_destroy(mmc_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed mmc_create when ctx
 * is NULL).  Never NULL. */
const char* mmc_last_error(const mmc_ctx* ctx);

/* Simulator::objective(z, top, trials), MC/src/sim.cpp:138-172.  z and grad are
 * 6*n_nodes node-major (DOF order tx,ty,tz,rx,ry,rz); grad may be NULL.  With
 * trials != 0 the per-beam trial states stay on the device for mmc_commit and
 * mmc_read_trial. */
int mmc_objective(mmc_ctx* ctx, const double* z, double top, int32_t trials,
                  mmc_objective_out* out, double* grad);

/* Simulator::minimize(top, z, final), MC/src/sim.cpp:174-211, started from z = 0
 * as advanceTo does (sim.cpp:398).  The minimiser, its final objective and the
 * trial states stay on the device; z_out (6*n_nodes, may be NULL) receives z. */
int mmc_minimize(mmc_ctx* ctx, double top, mmc_opt_stats* stats, mmc_objective_out* final_out,
                 double* z_out);

/* Simulator::commit(z, final), MC/src/sim.cpp:212: base pose <- state(z),
 * plastic <- trial.next.  Fails with MMC_ERR_STATE if no trial states exist. */
int mmc_commit(mmc_ctx* ctx);

/* Reads the committed node poses (3N, 4N) and beam plastic states (7B); any may be NULL. */
int mmc_read_state(mmc_ctx* ctx, double* x, double* q, double* plastic);
/* Reads the last trial records (MMC_TRIAL_STRIDE * n_beams). */
int mmc_read_trial(mmc_ctx* ctx, double* trial);
/* Device reductions over the committed state for diagnostics (sim.cpp:339-370). */
int mmc_read_stats(mmc_ctx* ctx, mmc_state_stats* out);

/* Simulator::capture(top, prevX, prevR), MC/src/sim.cpp:215-310, computed on the
 * device from the committed state: per lattice, one strut body per beam (beam
 * order), one joint body per node (node order), then the seven machine bodies
 * of the platen at tops[b].  Lattice b takes part when mask[b] != 0 (all when
 * mask is NULL); the bodies of the lattices taking part follow one another in
 * lattice order in out (B_b + N_b + MMC_MACHINE_BODIES each).  With
 * velocities[b] != 0 the bodies carry the velocities of the reference's
 * prevX/prevR form against the poses stored by the last mmc_capture_mark
 * ((x - x_prev) / dt and quatLog(normalized(R conj(R_prev))) / dt, dt =
 * max(1e-6, step_time)); with velocities NULL or velocities[b] == 0 they carry
 * zeros (the reference's first frame, prevX = prevR = nullptr). */
int mmc_capture(mmc_ctx* ctx, const double* tops, const int32_t* mask, const int32_t* velocities,
                mmc_snapshot* out);
/* Stores the committed node poses of every lattice as the previous poses of
 * the next mmc_capture (the copies of nodes_[i].x and .R that advanceTo keeps
 * before each minimisation, MC/src/sim.cpp:391-397); on the device. */
int mmc_capture_mark(mmc_ctx* ctx);

/* Kernel timer (no reference counterpart): HIP events recorded on the context
 * stream around every launch of one objective kernel (MMC_K_*) while enabled.
 * mmc_kernel_timer(ctx, id) starts (and clears) it, id < 0 stops it;
 * mmc_kernel_timer_read returns the launch count and the mean duration. */
enum { MMC_K_STATE = 0, MMC_K_FILL = 2, MMC_K_CONTACT = 4, MMC_K_BEAM = 5, MMC_K_GATHER = 6,
       MMC_K_FINAL = 7, MMC_K_DIRECTION = 8, MMC_K_PAIR = 9,
       /* the four kernels of one device-loop trial (mmc_time_kernel only) */
       MMC_K_DSTATE_G = 10, MMC_K_ELEM_G = 11, MMC_K_GATHER_G = 12, MMC_K_DECIDE_G = 13,
       /* a plain device copy of 16 MiB between two buffers of its own, the
          known-size reference for HBM traffic counters (mmc_time_kernel only) */
       MMC_K_COPY = 14 };
int mmc_kernel_timer(mmc_ctx* ctx, int32_t kernel);
int mmc_kernel_timer_read(mmc_ctx* ctx, int64_t* launches, double* mean_ms);
/* Objective evaluations performed by the context since creation. */
int64_t mmc_evaluations(const mmc_ctx* ctx);
/* Mean duration of one objective kernel (MMC_K_STATE, _CONTACT, _BEAM,
 * _GATHER) over `launches` back-to-back launches on the context stream at the
 * state of the last evaluation, timed with HIP events (no reference
 * counterpart).  MMC_K_DSTATE_G, _ELEM_G, _GATHER_G and _DECIDE_G time one
 * kernel of the device loop (after an mmc_minimize*) on a scratch copy of the
 * control blocks set to a trial (a fresh L-BFGS direction for DSTATE_G, a
 * backtrack for the others, with a full history for GATHER_G and DECIDE_G);
 * they write only the loop's scratch buffers, so the solver state is left as
 * it was. */
int mmc_time_kernel(mmc_ctx* ctx, int32_t kernel, int32_t launches, double* mean_ms);
/* Launch shape the device loop's tuner chose for the lattices of the
 * context's latest graph capture (no reference counterpart; the tuner runs at
 * the first capture over a set of each size class, 1, 2, 3-4, 5-8, ...
 * lattices, and reuses its choice for later sets of that class): blocks per
 * lattice of the decision kernel (1, 2, 4, 8 or 16), or 0 when the decision
 * runs in the gather kernel's last block per lattice, with the mean times
 * (us) it measured for each candidate (decide_us[6]: gather + decision at the splits 1, 2, 4, 8
 * and 16, then the gather with the decision fused; zeros before the first
 * minimisation or with MMC_TUNE=0).  The shape changes no result.  Either
 * pointer may be NULL. */
int mmc_loop_shape(mmc_ctx* ctx, int32_t* decide_blocks, double* decide_us);

/* Simulator::beamPotential(x, R), MC/src/sim.cpp:468-509, evaluated on the device
 * against the committed plastic state.  x is 3N, q is 4N (w,x,y,z). */
int mmc_beam_potential(mmc_ctx* ctx, const double* x, const double* q, double* out);

/* The same beam potential for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference.  x (3N) and q (4N) are DEVICE arrays in planes
 * (x[d*N + i]; q[0..3][N] = w, x, y, z) and energy is one DEVICE double; the
 * two kernels are queued on `stream` (a hipStream_t; NULL: the context's own
 * stream) and the call returns without waiting.  The reduction's block
 * partials live in a static device array of 65,536 doubles (512 KiB per GPU,
 * held until the process ends), so contexts of up to 16,777,216 beams;
 * larger ones get MMC_ERR_CAPACITY.  No hipMalloc/hipFree or other allocating
 * call is made.  Calls that overlap on one GPU share that array and race.
 * The device current at mmc_create must be current. */
int mmc_beam_potential_noalloc(mmc_ctx* ctx, const double* x, const double* q, double* energy, void* stream);

/* ==== batched contexts (no reference counterpart) ===========================
 * Many independent lattices in one context: one concatenated SoA, every launch
 * covering all of them, each instance with its own control block, platen,
 * material and schedule constants.  Per-instance results are those of the
 * single-lattice entry points above on that lattice alone. */

/* n lattices with one config each (cfgs[i].history etc. may differ). */
int mmc_create_batch(const mmc_lattice* lattices, const mmc_config* cfgs, int32_t n, int32_t precision,
                     mmc_ctx** out);
/* A batched context whose lattices are built on the device (no reference
 * counterpart as an entry point; the construction is Simulator::build and
 * distributeMass, MC/src/sim.cpp:49-135, for the BCC family, and the simple
 * cubic family of mmc_lattice_build): lattice i from cfgs[i] exactly as
 * mmc_lattice_build builds it, generated in the context's device buffers with
 * its beam incidence, lumped masses, locks and exclusion adjacency; no host
 * lattice exists.  Integer data (topology, layers, families, locks, exclusion
 * pairs) and whatever is made of sums, products, quotients and square roots
 * equal mmc_lattice_build's bit for bit; the second moments pi r^4 / 4 and the
 * Gaussian jitter of the cubic family (position_sigma > 0) use the device's
 * pow, log, sin and cos, within a few units in the last place of the host's. */
int mmc_create_generated(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_ctx** out);
/* Lattice b of a context: node and beam counts, exclusion pairs (each pair
 * counted once), envelope, and total mass (for generated lattices that of
 * distributeMass, MC/src/sim.cpp:134: beams in order, then joints; for
 * lattices handed to mmc_create* the sum of their node masses). */
typedef struct mmc_lattice_info {
  int32_t n_nodes, n_beams;
  int64_t n_exclusions;
  double width, height, depth, mass;
} mmc_lattice_info;
int mmc_lattice_info_get(mmc_ctx* ctx, int32_t b, mmc_lattice_info* out);
/* Reads lattice b's construction back in mmc_lattice's node-major layout
 * (local node indices): x 3N, mass N, locked 6N, beam_a, beam_b, L0, radius,
 * area, inertia B, e0 3B, node_layer N, beam_layer and beam_family B, and the
 * sorted exclusion adjacency (adj_off N + 1, adj 2 * n_exclusions).  Any
 * pointer may be NULL. */
int mmc_read_lattice(mmc_ctx* ctx, int32_t b, double* x, double* mass, uint8_t* locked, int32_t* beam_a,
                     int32_t* beam_b, double* L0, double* radius, double* area, double* inertia, double* e0,
                     int32_t* node_layer, int32_t* beam_layer, int32_t* beam_family, int32_t* adj_off,
                     int32_t* adj);
/* Number of lattices held by ctx (1 for mmc_create). */
int mmc_instances(const mmc_ctx* ctx);
/* mmc_minimize for every instance b with active[b] != 0 (all when active is
 * NULL) at platen tops[b], from z = 0.  stats and final_out have n entries;
 * inactive instances' entries describe their unchanged state. */
int mmc_minimize_batch(mmc_ctx* ctx, const double* tops, const int32_t* active, mmc_opt_stats* stats,
                       mmc_objective_out* final_out);
/* mmc_commit for the instances with mask[b] != 0 (all when mask is NULL). */
int mmc_commit_batch(mmc_ctx* ctx, const int32_t* mask);
/* mmc_read_stats of every instance (out has n entries). */
int mmc_read_stats_batch(mmc_ctx* ctx, mmc_state_stats* out);
/* Per-lattice progression of a batch (no reference counterpart; the batched
 * driver's alternative to rounds of mmc_minimize_batch, so that no lattice
 * waits for the slowest one of a round).  mmc_batch_start begins the solve of
 * every lattice b with mask[b] != 0 at platen tops[b] from z = 0 (a lattice
 * whose solve is in flight is an error); the solves in flight go on.
 * mmc_batch_advance runs the device loop until at least min_done of the
 * lattices in flight (or all of them) have ended their solves, and sets
 * running[b] (n entries) for those still solving.  mmc_batch_finish evaluates
 * the lattices with mask[b] != 0, whose solves must have ended, at their
 * iterates (stats[b], final_out[b]; trial states for mmc_commit_batch) and
 * takes them out of flight.  Each lattice's results equal those of
 * mmc_minimize_batch.  Needs line_search 0 and the graph-replayed loop. */
int mmc_batch_start(mmc_ctx* ctx, const double* tops, const int32_t* mask);
int mmc_batch_advance(mmc_ctx* ctx, int32_t min_done, int32_t* running);
int mmc_batch_finish(mmc_ctx* ctx, const int32_t* mask, mmc_opt_stats* stats, mmc_objective_out* final_out);

/* ==== host lattice construction (no device work) ========================== */
/* Builds the lattice the facade would build for cfg (MC/src/sim.cpp:49-135:
 * build + distributeMass, or the simple cubic family) without touching a GPU.
 * *view points into storage owned by *handle until mmc_lattice_free(handle). */
int mmc_lattice_build(const mmc_config* cfg, mmc_lattice* view, void** handle);
void mmc_lattice_free(void* handle);

/* ==== Simulator facade (MC/include/sim.hpp:35-54) exposed for C callers ===== */
typedef struct mmc_sim mmc_sim;

/* Simulator::Simulator(cfg): builds the lattice on the host (BCC or cubic),
 * distributes mass, sets the locks and creates the device context. */
int mmc_sim_create(const mmc_config* cfg, int32_t precision, mmc_sim** out);
void mmc_sim_destroy(mmc_sim* sim);
const char* mmc_sim_last_error(const mmc_sim* sim);
/* nodes().size(), beams().size(), initialHeight(), mass(), area(). */
int mmc_sim_info(const mmc_sim* sim, int32_t* n_nodes, int32_t* n_beams, double* height,
                 double* mass, double* area);
/* The host lattice the facade built (same layout as mmc_lattice; arrays sized by
 * mmc_sim_info).  Any pointer may be NULL. */
int mmc_sim_lattice(const mmc_sim* sim, double* x, double* mass, uint8_t* locked, int32_t* beam_a,
                    int32_t* beam_b, double* L0, double* radius, double* e0);
/* Simulator::run(), MC/src/sim.cpp:374-467: tare + load + unload schedule with
 * adaptive bisection.  Writes up to cap frames; *n_frames gets the count.
 * With verbose != 0 prints the reference's "[equilibrium]" line per step. */
int mmc_sim_run(mmc_sim* sim, mmc_frame* frames, int32_t cap, int32_t* n_frames, int32_t verbose);
/* Stepwise form of mmc_sim_run: mmc_sim_begin resets the schedule (tare, load,
 * unload; work accumulators as in MC/src/sim.cpp:379-387); each
 * mmc_sim_advance solves the next scheduled increment (with its bisection
 * sub-steps) and writes the accepted frames.  *done becomes 1 after the last
 * unload increment. */
int mmc_sim_begin(mmc_sim* sim);
int mmc_sim_advance(mmc_sim* sim, mmc_frame* frames, int32_t cap, int32_t* n_frames, int32_t* done);
/* Simulator::rigidRotationTest / gradientCheck, MC/src/sim.cpp:511-577. */
int mmc_sim_rigid_rotation_test(mmc_sim* sim, double* error, int32_t* pass);
int mmc_sim_gradient_check(mmc_sim* sim, double* error, int32_t* pass);
/* The device context under the facade (for objective-level callers). */
mmc_ctx* mmc_sim_context(mmc_sim* sim);

/* ==== batched Simulator facade (no reference counterpart) ==================
 * n lattices, each with the schedule of mmc_sim_run (its own config: lattice,
 * material, radii, schedule), solved together on one device context. */
typedef struct mmc_batch mmc_batch;
int mmc_batch_create(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_batch** out);
/* The same with the lattices built on the device (mmc_create_generated):
 * no host lattice is built or uploaded. */
int mmc_batch_create_generated(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_batch** out);
void mmc_batch_destroy(mmc_batch* batch);
/* First failure message (a lattice whose equilibrium failed), or the creation
 * error when batch is NULL.  Never NULL. */
const char* mmc_batch_last_error(const mmc_batch* batch);
/* Lattice i: nodes, beams, initial height, mass, reference area. */
int mmc_batch_info(const mmc_batch* batch, int32_t i, int32_t* n_nodes, int32_t* n_beams, double* height,
                   double* mass, double* area);
/* Runs every lattice's schedule.  Lattice b's frames go to frames[b * cap ..]
 * (at most cap), n_frames[b] gets its frame count and status[b] MMC_OK or
 * MMC_ERR_EQUILIBRIUM (that lattice stopped at a failed equilibrium, like
 * mmc_sim_run; the others continue).  With verbose != 0 a progress line goes
 * to stderr about every 10 s.  Returns MMC_OK unless the device fails. */
int mmc_batch_run(mmc_batch* batch, mmc_frame* frames, int32_t cap, int32_t* n_frames, int32_t* status,
                  int32_t verbose);
/* The device context of the batch (with MMC_BATCH_SHARDS > 1 the batch is
 * split over that many contexts, lattice b in context b mod shards, each run
 * by a host thread of its own on its own stream; this returns the first). */
mmc_ctx* mmc_batch_context(mmc_batch* batch);
/* Objective evaluations of the batch's contexts since creation. */
int64_t mmc_batch_evaluations(const mmc_batch* batch);
/* Frame observer of mmc_batch_run (no reference counterpart; the batched form
 * of SimulationResult::frames, MC/src/sim.cpp:436): after every committed
 * increment, fn(lattice, frame, bodies, count, user) receives that lattice's
 * frame as Simulator::capture gives it (B + N + MMC_MACHINE_BODIES bodies,
 * velocities against the lattice's previous frame, none on its first),
 * captured on the device for all lattices committed together; `bodies` is
 * valid during the call only.  fn NULL (the default) captures nothing. */
typedef void (*mmc_batch_frame_fn)(int32_t lattice, int32_t frame, const mmc_snapshot* bodies, int64_t count,
                                   void* user);
int mmc_batch_set_frame_observer(mmc_batch* batch, mmc_batch_frame_fn fn, void* user);

/* ==== manufacturing export ================================================= */
/* What the STL audit found (the fields of the reference's .validation.txt). */
typedef struct mmc_stl_report {
  double voxel;          /* m */
  double signed_volume;  /* m^3, after orientation */
  int32_t nx, ny, nz;    /* grid points */
  int32_t watertight;    /* no boundary and no non-manifold edge */
  int64_t triangles, unique_edges, initial_boundary_edges, repaired_loops, boundary_edges, nonmanifold_edges;
} mmc_stl_report;
/* Simulator::exportSTL, MC/src/sim.cpp:645-888: the union of the beams'
 * capsules (radius[k]) and the joints' spheres (joint_radius) at the lattice's
 * positions x, sampled on the reference's voxel grid (voxel from
 * radial_segments) on the device, its zero set extracted on the device, audited
 * and written as binary STL (mm) at path, with <path stem>.validation.txt.
 * Uses only n_nodes, n_beams, x, beam_a, beam_b and radius of the lattice.
 * Errors carry the reference's messages (mmc_stl_last_error). */
int mmc_export_stl(const mmc_lattice* lattice, double joint_radius, int32_t radial_segments, const char* path,
                   mmc_stl_report* report);
const char* mmc_stl_last_error(void);
/* The cell triangulation of the export (no reference counterpart): for each of
 * the 256 inside/outside patterns of a cell's corners (corner c at
 * (c & 1, c >> 1 & 1, c >> 2 & 1), inside = bit c), counts[case] triangles and
 * tris[case][10][3] their edge ids (edge 4a + m joins the m-th corner without
 * bit a to that corner with bit a).  Host only. */
void mmc_stl_case_table(int32_t* counts, int8_t* tris);
/* Simulator::exportSTL of the facade's lattice (initial positions). */
int mmc_sim_export_stl(mmc_sim* sim, const char* path, int32_t radial_segments, mmc_stl_report* report);

#ifdef __cplusplus
}
#endif
#endif /* MMC_H */
