/*
 * oracle.h — C API of the FP64 CPU oracle (test infrastructure only).
 *
 * The oracle is a single-threaded FP64 restatement of the reference solver
 * (projects/metamaterial-compression/industry-grade/src/sim.cpp).  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it, and
 * only as the checker.  The product library never links it.
 *
 * The same API, with the prefix ref_ instead of orc_, is exported by
 * oracle/_ref/libmmcref.so, built from the reference's own sources by
 * oracle/Makefile; tests drive both through one Python wrapper.
 */
#ifndef MMC_ORACLE_H
#define MMC_ORACLE_H
#include "../include/mmc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc* orc_create(const mmc_config* cfg);
/* Builds a solver around a given lattice geometry (x, q, beams, connected,
 * layers, envelope); masses and locks are recomputed as distributeMass does. */
orc* orc_create_lattice(const mmc_config* cfg, const mmc_lattice* lat);
void orc_destroy(orc* h);
const char* orc_last_error(orc* h);
void orc_info(orc* h, int32_t* n_nodes, int32_t* n_beams, int64_t* n_connected, double* env /*width,height,depth,area,mass,top0*/);
void orc_lattice(orc* h, double* x, double* q, double* mass, uint8_t* locked, int32_t* a, int32_t* b,
                 double* L0, double* radius, double* area, double* inertia, double* e0,
                 double* plastic, uint64_t* connected, int32_t* node_layer, int32_t* beam_layer,
                 int32_t* beam_family);
double orc_objective(orc* h, const double* z, double top, int32_t trials, mmc_objective_out* out,
                     double* grad, double* trial);
/* minimize from the given z (in/out); keeps the final objective for orc_commit. */
int orc_minimize(orc* h, double top, double* z, mmc_opt_stats* st, mmc_objective_out* fin);
int orc_commit(orc* h);
int orc_run(orc* h, mmc_frame* frames, int32_t cap, int32_t* n_frames, int32_t verbose);
/* Observer called by orc_run after every accepted frame with the frame and the
 * number of frames so far; a nonzero return ends the run there (rc 0). */
void orc_set_frame_callback(orc* h, int (*cb)(const mmc_frame* frame, int32_t n_frames, void* user), void* user);
double orc_beam_potential(orc* h, const double* x, const double* q);
/* Energy change z0 -> z1 at platen top: out[0] summed per term (the
 * line_search 1 form), out[1] as the difference of total energies. */
void orc_energy_change(orc* h, const double* z0, const double* z1, double top, double* out);
/* (reference harness only: ref_export_stl(h, path, segments), the reference's
 * Simulator::exportSTL, for the STL goldens) */
int orc_rigid_rotation_test(orc* h, double* err);
int orc_gradient_check(orc* h, double* err);
void orc_read_state(orc* h, double* x, double* q, double* plastic);
/* Overwrites committed plastic states (7*B) — for objective checks at yielded states. */
void orc_set_plastic(orc* h, const double* plastic);
/* Overwrites the committed base pose (x 3N, q 4N). */
void orc_set_pose(orc* h, const double* x, const double* q);

/* Simulator::capture(top, prevX, prevR) (MC/src/sim.cpp:215-310) of the
 * committed state: B + N + MMC_MACHINE_BODIES bodies into out; prev_x (3N) and
 * prev_q (4N) give the velocities, zero when NULL. */
void orc_capture(orc* h, double top, const double* prev_x, const double* prev_q, mmc_snapshot* out);

/* Closed-form helpers exposed for known-answer tests. */
void orc_scalar_return(double total, double pc, double ac, double k, double y, double H, double* out7);
void orc_vector_return(const double* t, const double* pc, double ac, double k, double y, double H, double* out12);
void orc_closest_segments(const double* p1, const double* q1, const double* p2, const double* q2, double* out8);
void orc_quat_exp(const double* v, double* q);
void orc_quat_log(const double* q, double* v);
void orc_jl_inv_t(const double* p, const double* v, double* out);

#ifdef __cplusplus
}
#endif
#endif
