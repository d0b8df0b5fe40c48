/*
 * fibrenet.h - C ABI of libfibrenet_amd, the MI355X (gfx950) engine for
 * FibreNet's particle-dynamics step.
 *
 * The reference has no FFI: its device seam is the C++ class OpenclManager
 * (OpenclManager.h:48-69), driven by SimulationManager through buffers
 * addressed by int id, kernels addressed by name, and an in-order queue.  The
 * functions below are that seam as a plain C ABI (opaque handle, plain
 * pointers and sizes, no C++ or torch types), so that the C++ facade in
 * fibrenet_manager.hpp, a ctypes/cgo/JNI binding, or FibreNet's own
 * SimulationManager can drive the HIP engine.  Each declaration names the
 * reference interface it replaces.
 *
 * Conventions: every function returns FN_OK (0) or a negative FN_ERR_* code
 * and never throws; fn_last_error() gives the message of the last failure on
 * a handle.  One host thread per handle; all launches are asynchronous on the
 * handle's HIP stream and fn_finish() / the blocking reads and writes are the
 * synchronisation points.
 */
#ifndef FIBRENET_H
#define FIBRENET_H

#include "fibrenet_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FN_OK 0
#define FN_ERR_ARG -1          /* invalid argument or state                     */
#define FN_ERR_HIP -2          /* HIP runtime failure                           */
#define FN_ERR_MISMATCH -3     /* length / record size differs from the buffer  */
#define FN_ERR_EMPTY -4        /* empty source, as CreateBuffer/WriteBuffer -2  */
#define FN_ERR_UNSUPPORTED -5  /* a configuration the engine does not run       */
#define FN_ERR_CAPACITY -6     /* contact list grew past its reserved capacity  */
#define FN_ERR_LOAD -7         /* the input or model file is rejected           */

/* Arithmetic of an engine instance. */
#define FN_PREC_F64 0          /* fp64 state and arithmetic (parity build)      */
#define FN_PREC_F32 1          /* fp32 forces/velocities, fp64 positions         */

/* Record kinds of a buffer (the reference's element type T of CreateBuffer<T>). */
#define FN_BUF_BEAD 1          /* fn_bead        (Bead)                         */
#define FN_BUF_BEAD_INFO 2     /* fn_bead_info   (BeadInfo)                     */
#define FN_BUF_SPRING 3        /* fn_spring      (SpringStructure)              */
#define FN_BUF_ANGLE 4         /* fn_angle       (AngleBond)                    */
#define FN_BUF_PAIR 5          /* fn_pair        (Pair)                         */
#define FN_BUF_INT 6           /* int            (grid and list heads)          */
#define FN_BUF_GRID_CELL 7     /* fn_grid_cell   (GridCell)                     */
#define FN_BUF_GRID_CELL_NH 8  /* fn_grid_cell_nh(GridCellNh)                   */
#define FN_BUF_PARTICLE_NH 9   /* fn_particle_nh (ParticleNh)                   */

/* Kinds of a kernel argument (OpenclManager::SetKernelArg<T>). */
#define FN_ARG_BUFFER 1        /* value = const int* buffer id                  */
#define FN_ARG_INT 2           /* value = const int*                            */
#define FN_ARG_SETTINGS 3      /* value = const fn_settings*, copied by value   */

typedef struct fn_engine fn_engine;
typedef struct fn_model fn_model;
typedef struct fn_sim fn_sim;
typedef void (*fn_log_fn)(const char *message, void *user);

/* Counters of the engine's last rebuild and run. */
typedef struct fn_stats {
    long long steps;               /* steps advanced by fn_run_steps / launches */
    long long rebuilds;            /* grid + contact-list rebuilds              */
    long long list_entries;        /* two-sided contact-list entries (=2 pairs) */
    long long list_capacity;       /* reserved entries                          */
    long long candidate_entries;   /* two-sided candidate entries (reach+skin)  */
    int candidate_builds;          /* candidate builds during the last run      */
    int hashed_beads;              /* beads inside the grid at the last rebuild */
    int max_cell_occupancy;        /* the reference drops beads beyond 64       */
    int cells_over_capacity;       /* cells that the reference would truncate   */
    int max_list_length;           /* longest two-sided list of one bead        */
    int broken_springs;            /* springs with status 0                     */
    int img_aliases_pos;           /* 1: one array holds position and image     */
} fn_stats;

/* Library version string. */
const char *fn_version(void);

/* OpenclManager::Initialize (OpenclManager.cpp:5-85): opens HIP device
 * `device` (no interactive platform/device prompt) with one stream. */
int fn_create(int device, int precision, fn_engine **out);
/* OpenclManager::Close (OpenclManager.cpp:318-320); frees all device memory. */
int fn_destroy(fn_engine *eng);
/* Message of the last failure on `eng` (checkErr, OpenclManager.cpp:335-409). */
const char *fn_last_error(fn_engine *eng);
/* Log sink replacing log_stream (log_stream.h:10-67); NULL restores stderr. */
int fn_set_log_callback(fn_engine *eng, fn_log_fn fn, void *user);

/* OpenclManager::RegisterKernel (OpenclManager.cpp:87-113): every name must be
 * one of FibreNet's ten step kernels; the engine's HIP code is prebuilt, so
 * there is no source text to compile. */
int fn_register(fn_engine *eng, const char *const *names, int n);

/* OpenclManager::CreateBuffer<T>(std::vector<T>&, flags) (OpenclManager.h:121-146)
 * and CreateBuffer<T>(int, flags) (OpenclManager.h:148-174).  `host` holds
 * n_elems records of `kind` (NULL: zero-filled); the engine pads to a multiple
 * of 64 records as the reference does.  *id receives the buffer id. */
int fn_buffer_create(fn_engine *eng, int kind, int elem_size, int n_elems, const void *host, int *id);
/* OpenclManager::WriteBuffer<T> (OpenclManager.h:100-119): blocking write of
 * n_padded records; FN_ERR_MISMATCH when n_padded or elem_size differ. */
int fn_buffer_write(fn_engine *eng, int id, const void *host, int n_padded, int elem_size);
/* OpenclManager::ReadBuffer (OpenclManager.cpp:174-201): blocking read of the
 * padded length x record size bytes into `host`.  Scratch buffers of the
 * reference's grid and list (GridCell, ParticleNh, list heads) are not
 * materialised and return FN_ERR_UNSUPPORTED; use the fn_export_* calls. */
int fn_buffer_read(fn_engine *eng, int id, void *host);
/* GetBufferSize / GetBufferElements / GetBufferLength / GetBufferStructureSize
 * (OpenclManager.cpp:130-172); any pointer may be NULL. */
int fn_buffer_info(fn_engine *eng, int id, int *size_bytes, int *elements, int *length, int *struct_size);

/* OpenclManager::SetKernelArg<T> (OpenclManager.h:74-98): binds argument
 * `index` of kernel `kernel` following the argument contract of
 * SimulationManager.cpp:1236-1304; settings are captured by value. */
int fn_set_arg(fn_engine *eng, const char *kernel, int index, int arg_kind, const void *value, int size);
/* OpenclManager::SetKernelGlobalWorkRange (OpenclManager.cpp:299-316). */
int fn_set_range(fn_engine *eng, const char *kernel, int range);
/* OpenclManager::ExecuteKernel (OpenclManager.cpp:219-285): enqueues the
 * engine's HIP implementation of the named kernel with its bound arguments. */
int fn_launch(fn_engine *eng, const char *kernel);
/* OpenclManager::RunKernels (OpenclManager.cpp:203-217): waits for the stream. */
int fn_finish(fn_engine *eng);

/* Fast path of SimulationManager::RunSimulationLoop (SimulationManager.cpp:151-181):
 * n_steps steps with a rebuild at the first step and every grid_update_freq
 * steps, as fused hipGraph-captured launches.  run_springs / run_angles mirror
 * the caller's springs_num > 0 / angles_num > 0 tests.  Same results as the
 * per-kernel launches; returns after the work is enqueued and checked.  The
 * contact search rows and list are sized from the longest rows and lists
 * seen; a run with none seen since the state was written keeps its start
 * state and is repeated with grown capacities if it overflows them.  Any
 * other overflow returns FN_ERR_CAPACITY: the run is invalid and the
 * capacities have been grown for the next one. */
int fn_run_steps(fn_engine *eng, int n_steps, int grid_update_freq, int run_springs, int run_angles);

/* Launch options of the fast path (no reference counterpart):
 *   "chunk"  n    steps captured per hipGraph (rounded to grid_update_freq)
 *   "chain"  1/0  bead kernel that evaluates chain bonds (spring (i, i+1),
 *                 angle (i-1, i, i+1), numbered along the chain) in-wave,
 *                 used when every bond follows that layout; also the
 *                 one-launch adjoint step of fn_adjoint_backward
 *   "chain_layout" (read only) 1 when every bond follows the chain layout
 *   "verlet" 0/1/2 keep the contact candidates (pairs within reach + skin)
 *                 between rebuilds while every bead stays within skin/2 of
 *                 its position at their build: never, always, or auto
 *                 (default: every 16th run tries it and keeps it while it
 *                 saves candidate builds).  An ensemble keeps them with one
 *                 skin for all networks (from the largest listing
 *                 distance), and only while that reach stays below half of
 *                 every periodic network's box
 *   "staged" 1/0  neighbour search reads each block's neighbourhood from LDS
 *   "skin"   n    the skin in permille of the largest listing distance (300)
 *   "graphs" 1/0  launch through captured hipGraphs or directly; a new
 *                 engine starts with 0 when the environment variable
 *                 FIBRENET_GRAPHS is "0" (e.g. under a profiler that cannot
 *                 follow graph launches), else with 1
 *   "block"  n    threads per block of the bead kernels: 128, 256 or 512
 * All options give bit-identical results; they change only the speed. */
int fn_set_option(fn_engine *eng, const char *name, int value);
int fn_get_option(fn_engine *eng, const char *name, int *value);

/* Kernel timing of the fast path (no reference counterpart): with timing on,
 * fn_run_steps records HIP events around each launch slot of every step
 * (spring terms, angle terms, bead kernel; with the chain bead kernel the
 * first two slots are empty) and around every grid rebuild on the engine's
 * stream - as event nodes of the captured graphs (the stream is synchronised
 * after each graph to read them) or around direct launches when graphs are
 * off; fn_get_timing sums them since timing was switched on. */
typedef struct fn_timing {
    int steps;                     /* steps measured                           */
    int rebuilds;                  /* rebuild sequences measured               */
    double step_ms;                /* summed durations of the step launches    */
    double rebuild_ms;             /* summed durations of the rebuilds         */
    double spring_ms;              /* of step_ms: the spring-term kernel       */
    double angle_ms;               /*             the angle-term kernel        */
    double bead_ms;                /*             the bead kernel              */
} fn_timing;
int fn_set_timing(fn_engine *eng, int on);
int fn_get_timing(fn_engine *eng, fn_timing *out);

/* Device-to-device copy of `bytes` (rounded up to 16) by a plain streaming
 * kernel (k_copy_probe) on the engine's stream, between two buffers
 * allocated for the call: the reference point that scales hardware traffic
 * counters to bytes (no reference counterpart).  *ms receives its duration. */
int fn_copy_probe(fn_engine *eng, long long bytes, double *ms);

/* ---- batched ensemble (no reference counterpart) ----
 * Independent networks, each with its own records and settings (box, grid,
 * pair table, integrator constants), advanced together: every fn_run_steps
 * step is one launch over all of them, and contacts never cross networks.
 * Bond indices are local to their network.  The ensemble uses the fast path
 * only (fn_run_steps); fn_get_stats and fn_export_* cover all networks with
 * global bead indices (network k's beads follow networks 0..k-1). */
typedef struct fn_network {
    const fn_bead *beads;          /* n_beads records                          */
    const fn_bead_info *info;      /* n_beads records                          */
    int n_beads;
    const fn_spring *springs;      /* may be NULL when n_springs is 0          */
    int n_springs;
    const fn_angle *angles;        /* may be NULL when n_angles is 0           */
    int n_angles;
    const fn_pair *pairs;          /* the network's pair table (<= 32 entries) */
    int n_pairs;
    const fn_settings *settings;   /* box, grid and integrator of the network  */
} fn_network;
/* Opens device `device` and loads the networks into one engine. */
int fn_batch_create(int device, int precision, int n_networks, const fn_network *nets, fn_engine **out);
/* Records of network `network`: kind FN_BUF_BEAD, FN_BUF_SPRING or
 * FN_BUF_ANGLE (bond indices local to the network). */
int fn_batch_read(fn_engine *eng, int network, int kind, void *out);
/* New records for every network of the ensemble, of the sizes loaded by
 * fn_batch_create (beads, bonds, pair entries, grid cells per network);
 * runs continue from them (a fit updating parameters between runs). */
int fn_batch_load(fn_engine *eng, int n_networks, const fn_network *nets);

/* ---- autotuner (no reference counterpart) ----
 * Runs n_steps steps exactly as fn_run_steps does (bitwise the same
 * results) while timing the settings that cannot change a result, one knob
 * at a time: the chain bead kernel, the LDS-staged search, kept contact
 * candidates and their skin, the bead kernels' block size and the chunk
 * length of the captured graphs (options "chain", "staged", "verlet",
 * "skin", "block", "chunk"; in that order; an ensemble tunes all but
 * "staged").  Each value runs one chunk to
 * capture its graphs and one timed chunk; the fastest value of a knob is
 * kept before the next knob is tried, and the chosen settings stay set
 * (fn_get_option reads them).  A run too short for every trial tries what
 * fits.  trials (capacity entries, may be NULL) receives each timing;
 * *n_trials their number. */
enum { FN_TUNE_CHUNK = 0, FN_TUNE_CHAIN = 1, FN_TUNE_STAGED = 2, FN_TUNE_VERLET = 3, FN_TUNE_BLOCK = 4, FN_TUNE_SKIN = 5 };
typedef struct fn_tune_trial {
    int knob;               /* FN_TUNE_*                                       */
    int value;              /* the option's value during the timed chunk       */
    double ms_per_step;     /* device time of the timed chunk per step         */
} fn_tune_trial;
int fn_autotune(fn_engine *eng, int n_steps, int grid_update_freq, int run_springs, int run_angles,
                fn_tune_trial *trials, int capacity, int *n_trials);

/* ---- adjoint (no reference counterpart) ----
 * Reverse-mode derivative of a fast-path run with respect to the force
 * scales of every spring, angle and contact pair type and to the initial
 * positions and velocity_old.  fn_adjoint_forward runs the steps exactly as
 * fn_run_steps does (same results; with kept candidates, option "verlet" 1,
 * they are built afresh at its first step and kept across its segments; a
 * run whose search rows or contact list overflow is repeated from its start
 * with the grown capacities) and records, per sample step, the x-force
 * summed over the flagged beads
 * (a clamp's reaction force) together with a checkpoint of the state every
 * checkpoint_every steps.  fn_adjoint_backward takes dL/d(sample) and
 * recomputes each segment from its checkpoint (bitwise the same steps, so
 * the contact lists and bond statuses of the forward run are reproduced
 * and held fixed), then runs the adjoint steps through it.  Derivatives
 * are fp64 in both precisions.  A scale sigma multiplies the force of one
 * bond (or of every contact of one pair type) and is 1 in the run: for a
 * spring, dL/dk = dL/dsigma / k.  On a batched ensemble every network is
 * sampled on its own: samples and dloss hold n_networks x n_samples values
 * (network-major), d_pair is indexed like the concatenated pair tables, and
 * forward and backward are each one launch per kernel for all networks.
 * Any call that changes the engine's state in between (runs, writes,
 * launches, a new "skin") makes fn_adjoint_backward fail until the next
 * fn_adjoint_forward. */
typedef struct fn_adjoint_spec {
    int n_steps;            /* as fn_run_steps                                 */
    int grid_update_freq;
    int run_springs;
    int run_angles;
    int checkpoint_every;   /* a multiple of grid_update_freq; 0: automatic     */
} fn_adjoint_spec;
/* group[n_beads]: nonzero for the sampled beads; sample_steps[n_samples]:
 * strictly increasing step indices in [0, n_steps); samples[n_samples]
 * (times n_networks on an ensemble) receives the sum over the group (of
 * each network) of the force x of those steps. */
int fn_adjoint_forward(fn_engine *eng, const fn_adjoint_spec *spec, const int *group, const int *sample_steps,
                       int n_samples, double *samples);
/* dloss[n_samples] (times n_networks) = dL/d(sample).  Outputs (any may be NULL):
 * d_spring[n_springs], d_angle[n_angles], d_pair[n_pairs] scale derivatives;
 * d_pos0[3 * n_beads], d_vel0[3 * n_beads] (x, y, z per bead) derivatives
 * with respect to the positions and velocity_old at the start of the run.
 * Leaves the engine in the state fn_adjoint_forward ended in. */
int fn_adjoint_backward(fn_engine *eng, const double *dloss, double *d_spring, double *d_angle, double *d_pair,
                        double *d_pos0, double *d_vel0);

/* Debug exports for parity checks (no reference counterpart). */
/* Cell index of every bead at the last rebuild (-1: outside the grid). */
int fn_export_cells(fn_engine *eng, int *cell_of_bead);
/* Unordered contact pairs (i<j) of the current list, sorted, as 2*n ints. */
int fn_export_pairs(fn_engine *eng, int *pairs, long long capacity, long long *n_pairs);
/* Step index at which each spring broke (-1: intact or broken before upload). */
int fn_export_breaks(fn_engine *eng, int *break_step);
int fn_get_stats(fn_engine *eng, fn_stats *out);

/* ---- loaders (SimulationManager::ReadInputFile, SimulationManager.cpp:1153-1171) ---- */

/* Parses an input file and its model (and restart) file into host records,
 * applying modif_bead / modif_spring; no device is touched.  Paths inside the
 * input file are resolved relative to its directory when not absolute. */
int fn_model_load(const char *input_path, fn_model **out);
int fn_model_free(fn_model *m);
const char *fn_model_error(void);
/* Counts: beads, springs, angles, pair-table entries, cells. */
int fn_model_counts(const fn_model *m, int *n_beads, int *n_springs, int *n_angles, int *n_pairs, int *n_cells);
/* Run parameters: run, output_freq, grid_update_freq, minimize steps/substeps/reset. */
int fn_model_params(const fn_model *m, int *run, int *output_freq, int *grid_update_freq, int *min_steps,
                    int *min_substeps, int *min_reset);
/* Copies the records of `kind` (FN_BUF_BEAD, _BEAD_INFO, _SPRING, _ANGLE,
 * _PAIR, _GRID_CELL_NH) or the settings (kind 0) into `dst`. */
int fn_model_copy(const fn_model *m, int kind, void *dst);

/* ---- simulation host (SimulationManager) ---- */

/* SimulationManager::Initialize (SimulationManager.cpp:8-51) without stdin:
 * loads `input_path`, uploads through the buffer API and binds every kernel
 * argument as ReadKernelData does (SimulationManager.cpp:1180-1307). */
int fn_sim_create(int device, int precision, const char *input_path, fn_sim **out);
int fn_sim_destroy(fn_sim *sim);
fn_engine *fn_sim_engine(fn_sim *sim);
/* Buffer id of a role: kind FN_BUF_BEAD, _BEAD_INFO, _SPRING, _ANGLE, _PAIR. */
int fn_sim_buffer_id(fn_sim *sim, int kind);
/* SimulationManager::RunSimulationLoop (SimulationManager.cpp:151-181).
 * fast = 0: per-kernel launches with a finish every step, as the reference;
 * fast = 1: the fused hipGraph path (fn_run_steps). */
int fn_sim_run_loop(fn_sim *sim, int loop_size, int fast);
/* SimulationManager::MinimizeSystemEnergy (SimulationManager.cpp:111-143). */
int fn_sim_minimize(fn_sim *sim, int fast);
/* SimulationManager::Run (SimulationManager.cpp:54-103): MinimizeSystemEnergy
 * when the input asks for minimize_steps > 0, a DumpData frame at step 0,
 * then RunSimulationLoop chunks of output_freq steps (the last one shorter),
 * each followed by DumpData, and the final DumpRestartData.  Messages go to
 * the fn_sim_set_log_callback sink (stderr without one). */
int fn_sim_run(fn_sim *sim, int fast);
/* SimulationManager::DumpData (SimulationManager.cpp:1937-2061): appends one
 * frame of each output_data field of the exported bead types to
 * <output_file>_coord.xyz / _vel.xyz / _force.xyz (the files are opened for
 * appending at the first call, as the reference does), and writes
 * <output_file>_restart.res when progress / run passes the next of 0.25,
 * 0.5, 0.75, ... (one stage per call, float arithmetic). */
int fn_sim_dump(fn_sim *sim, int progress);
/* SimulationManager::DumpRestartData (SimulationManager.cpp:1855-1928) into
 * `path` (NULL: <output_file>_restart.res). */
int fn_sim_write_restart(fn_sim *sim, const char *path, int progress);
/* Log sink of the simulation host (replacing log_stream, log_stream.h:10-67);
 * it also receives the engine's messages.  NULL restores stderr. */
int fn_sim_set_log_callback(fn_sim *sim, fn_log_fn fn, void *user);
/* Message of the last failure of a fn_sim_* call (or of its engine). */
const char *fn_sim_error(fn_sim *sim);
/* The output_file name of the input (relative names resolved against the
 * input file's directory), the stem of the dump and restart files. */
const char *fn_sim_output_file(fn_sim *sim);

/* ---- output formats (host only; no device is touched) ---- */

/* Fields of an XYZ frame (output_data position / velocity / force). */
#define FN_OUT_POSITION 0
#define FN_OUT_VELOCITY 1
#define FN_OUT_FORCE 2
/* DumpRestartData's text (SimulationManager.cpp:1894-1924) of `beads` (the
 * model's bead count of records, e.g. read back by fn_buffer_read) into
 * `path`, truncated: the bead count, "FRAME progress dt Lx Ly Lz", then per
 * bead its type name, position, image, velocity and force at the stream's
 * default 6 significant digits.  The format is lossy as the reference's is:
 * 6 digits, no velocity_old or force_old (the restart reader sets
 * velocity_old = velocity and force_old = 0) and no bond statuses. */
int fn_model_write_restart(const fn_model *m, const fn_bead *beads, int progress, const char *path);
/* One DumpData frame (SimulationManager.cpp:2002-2048) of `field` for the
 * beads of the exported types, appended to `path`: their count,
 * "FRAME progress dt Lx Ly Lz", then "type x y z" per bead, in fixed
 * notation with 5 decimals. */
int fn_model_write_xyz(const fn_model *m, const fn_bead *beads, int progress, int field, const char *path);

/* ---- checkpoints and observables (no reference counterpart) ---- */

/* Bit-exact binary checkpoint of the engine's dynamic state, in place of the
 * lossy restart text: every bead's position, image, predicted velocity,
 * velocity_old, force and force_old in the engine's own precision, the bond
 * statuses, the spring break steps and the step counter, plus a header of
 * the precision and record counts.  fn_checkpoint_load restores it into an
 * engine holding the same network (same precision, bead, spring and angle
 * counts); runs then continue bit for bit as the saved engine's would.
 * FN_ERR_MISMATCH when the file belongs to another network or build. */
int fn_checkpoint_save(fn_engine *eng, const char *path);
int fn_checkpoint_load(fn_engine *eng, const char *path);
/* The x-force of the current step summed over the beads flagged in
 * group[n_beads] (a clamp's reaction force, as Run's benchmark log records
 * it), per network: out[n_networks].  A fixed-order reduction on the device. */
int fn_group_force(fn_engine *eng, const int *group, double *out);
/* fn_group_force for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); it
 * has no counterpart in the reference.  Same arguments and the same result,
 * bit for bit: the same reduction kernels, with the flagged bead indices, the
 * per-network offsets, the partial sums and the result in a fixed device
 * workspace that the library holds per GPU (about 1 MiB, for at most 2^17
 * flagged beads over at most 256 networks; FN_ERR_CAPACITY beyond that).
 * No hipMalloc, hipFree or other allocating device call is made.  Every call
 * on one GPU shares that GPU's workspace, so calls that overlap on one GPU
 * (from several host threads) race; the workspace keeps the last call's
 * contents, and every call overwrites what it reads. */
int fn_group_force_noalloc(fn_engine *eng, const int *group, double *out);

/* ---- synthetic inputs (no reference counterpart) ---- */

/* Places n_fibres straight fibres of n_beads beads (spacing apart) with
 * uniformly random starts and directions; `finite` resamples fibres leaving
 * [0,L]^3 and, with d_min > 0, fibres coming closer than d_min to an already
 * placed bead.  out receives n_fibres*n_beads*3 unwrapped coordinates;
 * *tries the number of fibres sampled (FN_ERR_CAPACITY past max_tries). */
int fn_gen_fibres(unsigned long long seed, int n_fibres, int n_beads, double lx, double ly, double lz, int finite,
                  double spacing, double d_min, long long max_tries, double *out, long long *tries);

#ifdef __cplusplus
}
#endif

#endif /* FIBRENET_H */
