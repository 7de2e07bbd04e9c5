/*
 * qmi.h -- C-ABI of libqmi.so, the MI355X (gfx950) engine for Q's qdyn MD step.
 *
 * The reference exposes this path as a process and file interface: a
 * window is one `qdyn <input.inp>` run (argv rule src/q6/md.f90:2568-2577)
 * that reads the .inp (md.f90:2528-3262), the Qprep topology
 * (topo.f90:529-1082), the .fep file (qatom.f90:253-548, 874-1745) and an
 * optional restart, and writes the per-state energy file (.en,
 * nrgy.f90:94-119, md.f90:4156-4159), the final/backup restart
 * (md.f90:15452-15470) and the stdout log (md.f90:15292-15438).  Each entry
 * point below replaces one part of that process; a caller (ctypes, or the
 * qmi-qdyn command) binds them with plain pointers and sizes.
 *
 * Conventions: status 0 = ok, negative = error (text in qmi_last_error());
 * the library never calls exit().  Atom indices are 0-based.  Buffers are
 * owned by the caller; `cap` is the element capacity.  One engine per host
 * thread; calls on one engine are serialised by the caller.
 */
#ifndef QMI_H
#define QMI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qmi_engine qmi_engine;

#define QMI_ABI_VERSION 1

/* Energy layout of qmi_get_energies: the `energies` type of nrgy.f90:43-49
 * flattened in declaration order: potential, kinetic, LRF, p(bond, angle,
 * torsion, improper), w(4), q(4), pp(el, vdw), pw(2), ww(2), qx(2),
 * restraint(total, fix, shell, protein, solvent_radial, water_pol). */
#define QMI_N_ENERGIES 29
/* Per-state layout of qmi_get_q_energies: the `q_energies` type of
 * nrgy.f90:51-58 (lambda, total, q bond/angle/torsion/improper, qx el/vdw,
 * qq el/vdw, qp el/vdw, qw el/vdw, restraint) -- the 15 doubles of an .en
 * record. */
#define QMI_N_Q_ENERGIES 15

/* pair-list kinds for qmi_get_pairlist (md.f90:5175 nbpp, 6684 nbpw,
 * 8233 nbww, 7875 nbqp, 8066 nbqw) */
enum { QMI_LIST_PP = 0, QMI_LIST_PW = 1, QMI_LIST_WW = 2, QMI_LIST_QP = 3, QMI_LIST_QW = 4 };

/* Version of this interface; bumped on any layout change. No reference
 * counterpart. */
int qmi_abi_version(void);

/* Create an engine on HIP device `hip_device`.  Replaces the process start of
 * qdyn (qdyn.f90:64-90). */
int qmi_engine_create(int hip_device, qmi_engine** out);

/* Add one window: reads `inp_path` as qdyn reads its last argument
 * (md.f90:2528-3262); file names inside it are relative to its directory.
 * `seed_override` >= 0 replaces [MD] random_seed.  The window's outputs go
 * where the .inp names them.  Replaces initialize/topology/prep_coord/
 * get_fep/prep_sim/init_shake/make_nbqqlist/shrink_topology and the
 * Maxwell start (qdyn.f90:94-116). */
int qmi_window_add(qmi_engine* e, const char* inp_path, long long seed_override, int* window_id);

/* Where window w writes its run log (the stdout of qdyn, write_out
 * md.f90:15292-15438): a file path, "-" for this process's stdout, or NULL
 * for none.  Default: the .inp path with the extension .log.  Call before
 * qmi_prepare. */
int qmi_window_set_log(qmi_engine* e, int w, const char* path);

/* Upload all windows to the GPU, build the static Q-atom lists, pick tiles
 * and capture the step graph.  Replaces distribute_nonbonds
 * (md.f90:1044-1361) and the set-up half of md_run (md.f90:3917-4000). */
int qmi_prepare(qmi_engine* e);

/* Advance every window by `nsteps` MD steps (each window stops at its own
 * [MD] steps), writing its .en/log at the intervals of its .inp and its
 * backup restart after every step that is a multiple of 1000; the step loop
 * of md_run (md.f90:4003-4182).  With `finalize` != 0 the final pair lists,
 * energies, summary and restart of md_run (md.f90:4207-4213) are produced
 * for windows that have completed.  A window whose SHAKE does not converge
 * (or whose neighbour rows overflow) stops at that step, as die does
 * (md.f90:400-455): its log gets the termination record and its restart
 * file the state it stopped in; the other windows run on and are finalized,
 * then the call returns -60 (-61) naming the failed windows. */
int qmi_run(qmi_engine* e, long long nsteps, int finalize);

/* Evaluate pair lists (if `rebuild_lists`) and the potential energy and
 * gradient at the current coordinates without moving atoms: make_pair_lists
 * + pot_energy (md.f90:3716, 13278). */
int qmi_evaluate(qmi_engine* e, int rebuild_lists);

/* Energies of window w from the last energy evaluation (QMI_N_ENERGIES
 * doubles) -- E of md.f90 (type energies, nrgy.f90:43-49). */
int qmi_get_energies(qmi_engine* e, int w, double* out, int cap);

/* Per-state Q energies, nstates x QMI_N_Q_ENERGIES doubles -- EQ(:)
 * (qatom.f90:150, nrgy.f90:51-58). Returns the number of states. */
int qmi_get_q_energies(qmi_engine* e, int w, double* out, int cap);

/* Coordinates / velocities / gradient of window w, 3N doubles in Fortran
 * atom order (x, v, d of md.f90:222-225). */
int qmi_get_coords(qmi_engine* e, int w, double* xyz, long long cap);
int qmi_get_velocities(qmi_engine* e, int w, double* v, long long cap);
int qmi_get_forces(qmi_engine* e, int w, double* d, long long cap);
/* Replace coordinates / velocities of window w (3N doubles). */
int qmi_set_state(qmi_engine* e, int w, const double* xyz, const double* v, long long n3);

/* Pair list of window w from the last list build, as 1-based topology atom
 * pairs (i, j) and LJ code (1, 2, 3): the entries of nbpp/nbpw/nbqp, the
 * O-O molecule pairs of nbww (j = O atom), and the O atoms of nbqw (i = 0).
 * Pairs are unordered; `*n` returns the total count even when it exceeds
 * `cap`. */
int qmi_get_pairlist(qmi_engine* e, int w, int kind, int32_t* i, int32_t* j, int16_t* ljcod,
                     long long cap, long long* n);

/* Size of window w: natom, nat_solute, nwat, nstates, nqat (5 ints), and
 * with cap >= 6 its status: 0 running, 1 stopped by a SHAKE failure, 2
 * stopped because a neighbour row overflowed.  A failed window stops at the
 * failing step while the other windows of the engine run on; qmi_run then
 * returns an error naming every failed window. */
int qmi_window_info(qmi_engine* e, int w, int32_t* info, int cap);

/* Write window w's restart file (write_xfin, md.f90:15452-15470). */
int qmi_write_restart(qmi_engine* e, int w, const char* path);

/* The autotuner (on by default) times candidate launch shapes of the Q-tile,
 * pair-row, LRF-build and bonded kernels on the engine's own windows during
 * qmi_prepare and keeps the fastest; the candidates do not change results.
 * Call before qmi_prepare.  No reference counterpart. */
int qmi_set_autotune(qmi_engine* e, int on);

/* Choices of the autotuner: Q tiles per wave, threads per block of the row,
 * LRF-build and bonded kernels, whether blocks are grouped by XCD (0/1),
 * the bonded mode (0: per-term slots and a per-atom gather, 1: atomics, 2:
 * LDS sums per tile of terms), the number of source parts
 * of the LRF build, then the measured microseconds of the first six chosen
 * configurations (13 values).  Returns the number written. */
int qmi_tuning(qmi_engine* e, double* out, int cap);

/* Turn per-step HIP-event timing of the kernel families on or off for the
 * following qmi_run calls (no reference counterpart; for measurement). */
int qmi_set_timing(qmi_engine* e, int on);

/* Device time per kernel family of the last qmi_run, in microseconds per
 * step (no reference counterpart).  Order: pair-list + LRF build,
 * nonbonded tiles, bonded + sequence restraints + water polarisation, LRF
 * evaluation + atom restraints + leapfrog + SHAKE + thermostat, the
 * water-water tile kernel alone, wall-clock total.  All but
 * the last are zero unless qmi_set_timing was on.  Returns the number of
 * values written. */
int qmi_timings(qmi_engine* e, double* out, int cap);

/* Average device time of one launch of a step kernel in microseconds, over
 * `reps` launches in a row on the engine's stream between two HIP events (no
 * reference counterpart; for measurement).  which 0: the water-water force
 * kernel.  Call between qmi_run calls while the windows are still running:
 * the launches add to the gradient, so the engine then evaluates the
 * gradient and energies again at the current coordinates, as qmi_evaluate
 * does without rebuilding lists. */
int qmi_time_kernel(qmi_engine* e, int which, int reps, double* us);

/* Last error message of the calling thread. */
const char* qmi_last_error(void);

/* ---- Overlap predictor of the learned lambda schedule (no reference
 * counterpart: the reference's lambda schedules are written by hand).  An MLP
 * n_in -> n_hidden (SiLU) -> n_hidden (SiLU) -> 1 (sigmoid), trained and run on
 * the GPU; BF16 weights with an FP32 master copy, FP32 accumulation.
 * Parameters are ordered W1 (n_hidden x n_in, row-major), b1, W2 (n_hidden x
 * n_hidden), b2, W3 (1 x n_hidden), b3.  The loss is the mean squared error
 * of the output against targets in [0, 1]. */
typedef struct qmi_mlp qmi_mlp;
int qmi_mlp_create(int hip_device, int n_in, int n_hidden, unsigned long long seed, qmi_mlp** out);
long long qmi_mlp_num_params(qmi_mlp* m);
int qmi_mlp_get_params(qmi_mlp* m, float* params, long long cap);
int qmi_mlp_set_params(qmi_mlp* m, const float* params, long long n);
/* outputs y[batch] for inputs x[batch x n_in] */
int qmi_mlp_forward(qmi_mlp* m, const float* x, int batch, float* y);
/* loss and gradient (num_params floats, may be NULL) for one batch */
int qmi_mlp_grad(qmi_mlp* m, const float* x, const float* t, int batch, float* loss, float* grad);
/* Adam training on n samples, shuffled by `seed`, for `epochs` passes of
 * minibatches of `batch`; final_loss = mean loss over the last pass */
int qmi_mlp_train(qmi_mlp* m, const float* x, const float* t, int n, int epochs, int batch, float lr,
                  unsigned long long seed, float* final_loss);
/* The forward pass of qmi_mlp_forward on device buffers, allocating nothing:
 * d_x [batch x n_in] and d_y [batch] are device pointers of the model's
 * device, the kernels are queued on `stream` (a hipStream_t; NULL = the
 * model's own stream) and the call returns without waiting, so it can be
 * captured into a HIP graph.  No counterpart in the reference; it is for
 * callers that cannot allocate (code under graph capture, serving loops).
 * The activations go to one fixed scratch array per device (4 MiB), so
 * batch x (4 n_hidden + 1) must not exceed 1048576 floats (2047 rows for
 * n_hidden 128; larger batches are refused), and calls that overlap on one
 * device race on it. */
int qmi_mlp_forward_noalloc(qmi_mlp* m, const float* d_x, int batch, float* d_y, void* stream);
const char* qmi_mlp_last_error(void);
void qmi_mlp_destroy(qmi_mlp* m);

void qmi_engine_destroy(qmi_engine* e);

#ifdef __cplusplus
}
#endif
#endif
