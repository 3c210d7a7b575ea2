/*
 * temfetd.h -- C ABI of the MI355X FETD forward time-stepping library
 * (libtemfetd.so).  All floating point is FP64; all index arrays use the
 * reference's 1-based vertex/edge/tet numbering; 2-D arrays are row-major
 * [item][component] (the Fortran caller passes the transposes it already
 * holds).
 *
 * Every function returns 0 on success and a negative code on error;
 * tf_last_error() gives the message.  A context is single-threaded and not
 * re-entrant (like the reference's module state); distinct contexts are
 * independent.  The context owns all device memory.
 *
 * The reference's seams that these entry points replace:
 *   model_mesh        src/FETD_FWD.f90:112-282    -> tf_set_mesh, tf_get_topology
 *   model_ctl         src/FETD_FWD.f90:284-586    -> tf_set_materials, tf_set_source,
 *                                                    tf_set_receivers, tf_set_schedule
 *   mesh_process      src/FETD_FWD.f90:588-732    -> tf_assemble, tf_get_matrices
 *   mkl_add_3 +
 *   pardiso_direct_1  src/FETD_FWD.f90:2006-2039,
 *                     src/FETD_UTILITY.f90:879-936 -> tf_factor, tf_factor_csr
 *   pardiso_direct_2  src/FETD_UTILITY.f90:938-957 -> tf_solve
 *   pardiso_direct_3  src/FETD_UTILITY.f90:959-975 -> tf_release
 *   time_solver /
 *   adap_euler        src/FETD_FWD.f90:1174-1464  -> tf_run
 *   solution_output   src/FETD_FWD.f90:1513-1623  -> rows returned by tf_run
 */
#ifndef TEMFETD_H
#define TEMFETD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct tf_ctx tf_ctx;

/* status codes */
#define TF_OK 0
#define TF_ERR_ARG (-1)
#define TF_ERR_STATE (-2)
#define TF_ERR_HIP (-3)
#define TF_ERR_NUMERIC (-4)
#define TF_ERR_INPUT (-5)

/* Create a context on HIP device `device`. */
int tf_create(tf_ctx **ctx, int device);
void tf_destroy(tf_ctx *ctx);
/* Message of the last failed call on ctx; with ctx == NULL, of the last
 * failed tf_create on this thread. */
const char *tf_last_error(const tf_ctx *ctx);

/* model_mesh (src/FETD_FWD.f90:112-282): vertex coordinates [nN][3] and
 * tets [nT][4] (1-based vertex ids).  Builds the global edge numbering
 * (lexicographic rank of (min,max) vertex pairs), tet_edg/tet_drc, edge
 * lengths and the upper CSR pattern shared by all system matrices. */
int tf_set_mesh(tf_ctx *ctx, int nN, const double *nod_cor, int nT, const int *tet_vex);
int tf_get_sizes(const tf_ctx *ctx, int *nE, int *nnzU);
/* Copies (any pointer may be NULL): edg_vex [nE][2], tet_edg [nT][6],
 * tet_drc [nT][6], ia [nE+1], ja [nnzU], all 1-based like the reference. */
int tf_get_topology(const tf_ctx *ctx, int *edg_vex, int *tet_edg, int *tet_drc, int *ia, int *ja);

/* model_ctl (src/FETD_FWD.f90:347-364): per-tet conductivity and 1/mu. */
int tf_set_materials(tf_ctx *ctx, const double *sigma, const double *inv_mu);
/* Anisotropic media (model_ctl 'anisotropic:' src/FETD_FWD.f90:368-410;
 * f_kind_ani/e_kind_ani :853-912): per tet a row-major 3x3 conductivity
 * tensor [nT][9] and the 3x3 curl weight [nT][9] the reference puts between
 * the curls (the entrywise reciprocal of its permeability tensor, zero
 * entries kept, src/FETD_FWD.f90:621-633).  Replaces tf_set_materials until
 * that is called again. */
int tf_set_materials_tensor(tf_ctx *ctx, const double *sigma9, const double *nu9);
/* Source wire (src/FETD_FWD.f90:418-448, sys_right :1466-1511): the nseg
 * segments [nseg][2] (1-based vertex ids) in file order; signs are +1 for all
 * segments but the last (-1).  source_type 1, pulse_type 1 (step-on),
 * 2 (impulse) or 3 (step-off).  Call repeatedly with source index isrc =
 * 0..nsrc-1 to define nsrc sources that share one waveform; nsrc > 1 runs as
 * a multi-RHS block solve. */
int tf_set_source(tf_ctx *ctx, int isrc, int nseg, const int *seg_vex, int source_type,
                  int pulse_type, double pulse_width, double current);
/* Number of sources (1..64) time-stepped together on one factorisation; the
 * doubling check then uses the largest per-source error.  Call before
 * tf_set_source.  No reference counterpart (one source per reference run). */
int tf_set_num_sources(tf_ctx *ctx, int nsrc);
/* Receivers (src/FETD_FWD.f90:500-520, p_in_tet :1009-1049): [nR][3]
 * coordinates; rec_tet_out (nullable) receives the 1-based containing tet. */
int tf_set_receivers(tf_ctx *ctx, int nR, const double *xyz, double rec_charac, int *rec_tet_out);
/* Controller settings (src/FETD_FWD.f90:418-470): sp_division, sp_double,
 * sp_dblesize, time_maximum. */
int tf_set_schedule(tf_ctx *ctx, int sp_division, int sp_double, int sp_dblesize,
                    double time_maximum);

/* mesh_process (src/FETD_FWD.f90:588-732): element matrices on the GPU,
 * gathered into the CSR values in the reference's COO order; receiver
 * interpolation vectors. */
int tf_assemble(tf_ctx *ctx);
/* Upper-CSR values [nnzU] of sigma-mass and curl-curl (nullable). */
int tf_get_matrices(const tf_ctx *ctx, double *mass_sigma, double *curl);
/* Receiver vectors n_arr, d_arr as [nR][6][3] (nullable). */
int tf_get_receiver_vectors(const tf_ctx *ctx, double *n_arr, double *d_arr);

/* Symbolic analysis, once per mesh pattern (the reference redoes it in every
 * phase-12 call).  Optional: tf_factor runs it on first use.  leaf_size is the
 * largest vertex set the nested dissection leaves undivided; <= 0 keeps the
 * default (192). */
int tf_analyze(tf_ctx *ctx, int leaf_size);
/* Tuning knobs of the block-solve tile kernels (no reference counterpart;
 * set by the autotuner, temf3dt_amd/autotune.py).  Keys:
 *   "mt_units"  work units a tile launch aims for (default 8192, >= 64)
 *   "mt_chmin"  minimum inner chunk of a split tile (default 512, multiple of 16)
 *   "mt_ks_n"   k-steps per stage of the N-form kernels (2 or 4, default 2)
 *   "mt_ks_t"   k-steps per stage of the T-form kernels (2 or 4, default 2)
 *   "mt_wide"   levels of the elimination tree with at most this many fronts
 *               use 64-row tiles, the others 32-row tiles (default 256, >= 0)
 *   "graphs"    1 (default): regular steps of tf_run (sources off) replay a
 *               captured hipGraph of right-hand side, solve and history copy;
 *               0: every step is launched kernel by kernel.  Bit-identical.
 *   "overlap"   1 (default): tf_run factors the doubling check's P(2 dt) on a
 *               second stream while the steps before the check run, and
 *               tf_adjoint its sweep's next dt while the current one is in
 *               use; 0: at the point of use.  Bit-identical.
 *   "s_lt"      1 (default): single-source contexts also keep the row-major
 *               copy of L21 that multi-source contexts make after each
 *               factorisation, and the single solve's backward products read
 *               it row-wise; 0: they read L column-wise.  Equal to rounding.
 *   "rhs_chain" 1 (default): regular steps of a multi-source tf_run (even
 *               source count, no checkpointing) form Msigma (4 e2 - e3) as
 *               4 Msigma e2 - Msigma e3, the second product kept from the step
 *               before; 0: the history combination first.  Equal to rounding.
 *   "keep_states" 1 (default): a taped tf_run keeps the solution of every
 *               node while the kept states take at most a quarter of the
 *               device memory and a fifth of it stays free, and tf_adjoint
 *               reads them instead of recomputing them from the checkpoints;
 *               0: every state is recomputed.  Bit-identical.
 *   "keep_factors" a taped tf_run keeps up to this many of the factors
 *               it would overwrite (besides the two slots) while a tenth of
 *               the device memory stays free, and tf_adjoint uses them
 *               instead of factoring P(dt) again (default 8; 0: none).
 *               Bit-identical.
 *   "f_band"    the factorisation's trailing-update tiles are walked in
 *               bands of this many 64-row tile rows, column by column inside
 *               a band, so that the tiles running at once share their operand
 *               panels in L2 (default 16; 0: row by row; 0 .. 1024).
 *               Bit-identical.
 *   "indefinite" 0 (default): a factorisation with a negative pivot fails
 *               with TF_ERR_NUMERIC (P is SPD in exact arithmetic; the
 *               reference's SPD factorisation would fail there and be redone
 *               as symmetric indefinite, src/FETD_UTILITY.f90:919-935);
 *               1: the unpivoted LDL^T is accepted as it is.
 * Results do not depend on the knobs beyond floating-point summation order. */
int tf_set_tuning(tf_ctx *ctx, const char *key, int value);
/* Kernel timing with HIP events on the context's stream (no reference
 * counterpart).  tf_set_profiling(ctx, 1) clears the totals and starts
 * recording an event pair around every kernel launch of the factorisation,
 * the solves and the time step; tf_get_profile returns the number of
 * launches and their summed duration (ms) in one category. */
#define TF_PROF_TILE 0           /* block-solve tile products (k_mtile) */
#define TF_PROF_TILE_RED 1       /* their partial-sum reductions */
#define TF_PROF_PULL 2           /* block-solve forward pull */
#define TF_PROF_FACTOR_UPDATE 3  /* trailing updates of the factorisation */
#define TF_PROF_FACTOR_OTHER 4   /* rest of the factorisation */
#define TF_PROF_SOLVE1 5         /* single right-hand-side solve kernels */
#define TF_PROF_STEP 6           /* right-hand side: SpMV, history, source */
#define TF_PROF_NCAT 7
int tf_set_profiling(tf_ctx *ctx, int on);
int tf_get_profile(tf_ctx *ctx, int category, int64_t *launches, double *ms);
/* Factor statistics: nnz(L) incl. diagonal, dense-front flops, supernodes. */
int tf_factor_stats(const tf_ctx *ctx, int64_t *nnzL, double *flops, int *nsuper, int *nlevels);
/* Storage of the supernodal factor (no reference counterpart; sizes for
 * memory planning and the solve's byte count): the struct row indices of all
 * fronts (sum of each front's rows below its own columns), the doubles of one
 * factor slot's L storage and of the update stack. */
int tf_factor_layout(const tf_ctx *ctx, int64_t *struct_rows, int64_t *l_doubles, int64_t *u_doubles);
/* Pivots of the last factorisation (the counterpart of PARDISO's inertia
 * report, iparm(22)/iparm(23), and of the mtype 2 failure test of
 * pardiso_direct_1, src/FETD_UTILITY.f90:919-935): the number of negative
 * pivots of D and the smallest and largest |d|.  Every factorisation scans
 * its pivots; tf_factor, tf_factor_csr, tf_run, tf_run_models and tf_adjoint
 * return TF_ERR_NUMERIC on a zero or non-finite pivot, and on a negative one
 * unless tuning key "indefinite" is 1. */
int tf_factor_pivots(const tf_ctx *ctx, int64_t *n_negative, double *min_abs, double *max_abs);

/* mkl_add_3 + pardiso_direct_1: P = 3*Msigma + 2*dt*K formed on the GPU and
 * factored as LDL^T into slot 0 or 1. */
int tf_factor(tf_ctx *ctx, int slot, double dt);
/* pardiso_direct_1 with caller values: upper CSR (1-based) whose pattern is
 * the assembled one (ia/ja are checked against it). */
int tf_factor_csr(tf_ctx *ctx, int slot, int n, const int *ia, const int *ja, const double *va);
/* pardiso_direct_2: x = P^-1 b for nrhs right-hand sides, host buffers in
 * edge order, [nrhs][n]; nrhs > 1 runs the block solve, 64 at a time. */
int tf_solve(tf_ctx *ctx, int slot, int nrhs, const double *b, double *x);
/* No reference counterpart.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tf_solve for one right-hand side with b and x as device
 * pointers (n doubles each, edge order), which allocates nothing.  Its work
 * vectors live in a file-scope device array of TF_NOALLOC_SCRATCH_BYTES per
 * GPU (held until the process ends), so it serves systems whose work vectors
 * fit: n + (n + struct rows of the factor) + 2 n doubles, each rounded up to
 * 32, at most 8,388,608 doubles (tf_factor_layout gives the struct rows;
 * the 1.1M-unknown plain 3-D mesh fits, the 1.6M-unknown batched mesh does
 * not).  Larger systems return
 * TF_ERR_ARG.  The slot must have been factored by tf_factor or
 * tf_factor_csr.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (from two threads or contexts) race.  Returns after the
 * solve has finished. */
#define TF_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int tf_solve_noalloc(tf_ctx *ctx, int slot, const double *b, double *x);
/* Device-resident solve timing for the autotuner (no reference counterpart):
 * nrhs (1..64) deterministic right-hand sides already in HBM, one warm-up
 * solve, then reps solves timed with HIP events; *ms_per_solve receives the
 * average.  In a context of independent models (tf_set_models) it times one
 * solve of every model's current factor (nrhs = 1, slot ignored) after a
 * tf_run_models. */
int tf_time_solve(tf_ctx *ctx, int slot, int nrhs, int reps, double *ms_per_solve);
/* pardiso_direct_3 */
int tf_release(tf_ctx *ctx, int slot);

/* time_solver/adap_euler: runs the whole BDF2 loop on the GPU and returns
 * the receiver rows solution_output would write: step, t, dt, and
 * fields [row][src][rec][9] (ex ey ez dBx/dt dBy/dt dBz/dt bx by bz).
 * *nrows receives the number of rows; rows beyond max_rows are counted but
 * not stored.  max_steps > 0 caps the loop (bounded runs). */
int tf_run(tf_ctx *ctx, int max_steps, int max_rows, int *nrows, int *row_step, double *row_t,
           double *row_dt, double *fields);
/* Adjoint (no reference counterpart; the reference has no gradient).
 * tf_set_checkpointing(ctx, k) with k > 0 makes the following tf_run record
 * a tape of its solves and checkpoint the two BDF2 history states every k
 * steps (k = 0 turns it off).  tf_adjoint then returns, for
 *   J = sum over rows, sources, receivers, components of w * fields
 * (w laid out like tf_run's fields: [nrows][src][rec][9], nrows = the rows of
 * the taped run), grad_logsigma[nT] = dJ/d(log sigma_t) per tet: a
 * backward-in-time sweep that takes each node's forward state (kept by the
 * taped run, or recomputed from its checkpoint segment), solves the adjoint
 * systems with the same LDL^T factors (P is symmetric; kept by the run or
 * factored again) and reduces -lambda^T M1_t (3 y - 4 u + v) per element.
 * What the run keeps (tuning keys "keep_states", "keep_factors") stays on
 * the device until the next tf_run, tf_set_checkpointing(ctx, 0) or
 * tf_destroy, so tf_adjoint can be called again with other weights.  The
 * controller's keep/change decisions are those of the taped run
 * (differentiate with them frozen; see tf_set_replay).  Step-on and impulse
 * sources (pulse types 1, 2). */
int tf_set_checkpointing(tf_ctx *ctx, int every);
int tf_adjoint(tf_ctx *ctx, int nrows, const double *w, double *grad_logsigma);
/* Independent models in one context (no reference counterpart: the
 * reference runs one model per job).  Concatenate the models' meshes in
 * tf_set_mesh (each model's vertices one block, blocks in model order; tets in
 * model order), their per-tet materials in tf_set_materials and their
 * receivers in tf_set_receivers; tf_set_models(ctx, M, ntet[M], nrec[M])
 * (after tf_set_mesh, before tf_set_receivers) splits tets and receivers by
 * model, and tf_set_source(ctx, m, ...) gives model m's wire.  The system is
 * block diagonal: one tree of fronts per model, one solve per step for all of
 * them.  tf_run_models steps every model with its own adaptive controller
 * (time, dt, keep/change decisions and factor slot, src/FETD_FWD.f90:1229-1464)
 * until each reaches time_maximum (max_steps > 0 caps the steps); with
 * use_graphs regular steps with the sources off replay captured hipGraphs.
 * tf_get_model_rows returns model m's rows like tf_run ([row][rec][9]). */
int tf_set_models(tf_ctx *ctx, int nmodels, const int *ntet, const int *nrec);
int tf_run_models(tf_ctx *ctx, int max_steps, int use_graphs);
int tf_get_model_rows(tf_ctx *ctx, int model, int max_rows, int *nrows, int *row_step, double *row_t,
                      double *row_dt, double *fields);
/* Replay the controller's decisions instead of computing them: keep[i] is
 * the keep (1) / change (0) outcome of the i-th doubling check (as logged by
 * tf_get_checks of an earlier run); the checks' solves and factorisations
 * still run.  nchecks = 0 restores the error-checked controller
 * (src/FETD_FWD.f90:1380-1417).  No reference counterpart: it makes runs with
 * different source sets follow one schedule. */
int tf_set_replay(tf_ctx *ctx, int nchecks, const int *keep);
/* Per-tet field slices: the .cntl key vtkout (src/FETD_FWD.f90:476-484) with
 * vtk_field_save (:1659-1711) and solution_output's slice trigger
 * (:1642-1656).  tf_set_slices gives the slice times before tf_run; during
 * the run every output row, while a slice is pending, evaluates E and curl E
 * at each tet's centroid and accumulates B += curl E * dt (curl E here is
 * +curl, the opposite sign of the receivers' dB/dt; the reference writes
 * absolute values).  A pending slice is taken at the first row with
 * t >= its time.  tf_get_slice copies slice i as fields9[nT][9] (E, curl E,
 * B) with the step and time of that row; it fails for a slice the run did
 * not reach.  Single-source contexts (tf_run) only; nslice = 0 turns slices
 * off. */
int tf_set_slices(tf_ctx *ctx, int nslice, const double *times);
int tf_get_slice(tf_ctx *ctx, int i, double *fields9, int *step, double *t);
/* Doubling-check log of the last tf_run: step, err, keep flag. */
int tf_get_checks(const tf_ctx *ctx, int max, int *n, int *step, double *err, int *keep);
/* Timers of the last tf_run in seconds: assemble, factor, solve, spmv+rhs,
 * readout, loop wall, analyze, and the number of factorisations. */
int tf_get_timers(const tf_ctx *ctx, double *t8);

#ifdef __cplusplus
}
#endif
#endif
