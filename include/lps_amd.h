/* lps_amd.h - C ABI of the MI355X-native LPS-ABC time loop.
 *
 * The reference solver (src/main.cpp) has no in-process API: its boundary is the command line
 * `PD2_LPS_ABC <prefix> <jobname>`, the input files the preprocessor exports and the .energy/.verify
 * outputs. This header is the thin C layer the C++ host (the drop-in PD2_LPS_ABC program) calls; a
 * problem handed over here is exactly what src/main.cpp:178-293 reads from those files. Each entry
 * point names the part of src/main.cpp it replaces.
 *
 * Conventions: plain pointers and sizes, no C++ types. Every call returns 0 on success or a
 * negative LPS_ERR_* code; lps_last_error() then holds the message. No exception crosses the ABI.
 * A context owns all its device memory and one HIP stream; calls on one context are not
 * thread-safe (use one host thread per context). Host arrays are copied at creation; the caller
 * keeps ownership.
 */
#ifndef LPS_AMD_H
#define LPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPS_OK 0
#define LPS_ERR_ARG -1         /* invalid argument or inconsistent problem */
#define LPS_ERR_HIP -2         /* HIP runtime error (no device, out of memory, launch failure) */
#define LPS_ERR_HORIZON -3     /* a horizon holds more than 128 nodes (kd_tree_interface.h:96-99) */
#define LPS_ERR_NONFINITE -4   /* NaN/Inf in energy or max|u| at an output step (main.cpp:1126-1133) */
#define LPS_ERR_CONSISTENCY -5 /* a FATAL material consistency check (main.cpp:310-322, 496-504) */

/* ABC dilatation method (src/main.cpp:38, 765-812). */
#define LPS_THETA_IDW 0
#define LPS_THETA_GTHETA 1

/* One problem: the contents of the job files (Appendix A of the file formats). Index arrays are the
 * values after the reader's (int)(x + 0.5) rounding (src/helpers/utils.h:52-80), so a -1 pad in
 * .iClouds arrives as 0. */
typedef struct lps_problem {
  int32_t node_no;          /* .data1[0] NodeNo */
  int32_t constrains_no;    /* .data1[4] constrainsNo (ABC rows) */
  int32_t max_cloud_size;   /* .data1[5] maxCloudSize */
  double E, nu, rho, dt;    /* .data2[0], [1], [2], [6] */
  double delta, dx, dy;     /* .data2[7] (horizon R), [8], [9] */
  const double *nodes;      /* node_no x 2, row-major (.nodes) */
  const int32_t *bctypes;   /* 2 * node_no (.bctypes) */
  const double *icvalues;   /* 2 * node_no, (x, y) interleaved (.icvalues) */
  const double *icdvalues;  /* 2 * node_no (.icdvalues) */
  const int32_t *inf_nodes; /* constrains_no (.infNodesIndex) */
  const int32_t *iclouds;   /* constrains_no x max_cloud_size (.iClouds) */
  const double *ci_nodes_u; /* 2 constrains_no x 4 max_cloud_size (.ciNodesU) */
  const double *ci_nodes_v; /* 2 constrains_no x 4 max_cloud_size (.ciNodesV) */
  const double *ci_nodes_theta; /* constrains_no x 2 max_cloud_size (.ciNodesTheta); NULL for IDW */
  int32_t prenotch_no;      /* .data1[2] PrenotchNo (0: no notch) */
  const double *prenotches; /* prenotch_no x 4 (x0 y0 x1 y1), row-major (.prenotch): every bond
                               whose segment meets one of them is cut (src/main.cpp:386-402) */
} lps_problem;

typedef struct lps_options {
  int32_t theta_method;  /* LPS_THETA_GTHETA (reference default) or LPS_THETA_IDW */
  int32_t abc_enabled;   /* 1: ABCS_STEP == 2 (reference default); 0: ABCS_STEP == -1 */
  int32_t exact;         /* 1: per-bond arithmetic in the reference's operation order, no
                            contraction (bit-compatible with a no-FMA CPU build); 0: fast path */
  int32_t use_graph;     /* 1: replay the 30-step inner loop as a captured HIP graph */
  int32_t device;        /* HIP device ordinal */
  double horizon_radius; /* > 0: horizon search radius; 0: delta + dx/2 + 1e-5 (src/main.cpp:352) */
  int32_t variant;       /* fast-path node kernels: 0 autotune at creation (time each candidate
                            on the problem, keep the fastest per pass), 1 generic tile kernels,
                            2 unrolled 68-bond stencil, 3 register-blocked 68-bond stencil (2 and
                            3 when the 68-bond stencil applies, else generic) */
  int32_t concurrent;    /* fast path: 1 runs the ABC rows' kernels (G_theta, G_u, G_v) on a
                            second stream beside the tile kernels; 2 also the partial tiles'
                            kernels on a third stream beside the full tiles'; 3 runs the ABC rows
                            as extra workgroups of the dilatation launch, one stream; 4 as 1,
                            with the "deep" full tiles (full 3 x 3 neighbourhood, interior nodes
                            only in the 24 x 24 window) computing dilatation and force in one
                            launch while the other tiles' dilatation runs on a third stream; 0
                            the row kernels after the tile kernels, one stream; -1 (the default)
                            2 when the partial tiles are at least a tenth of the tiles (batches
                            of small domains), else 1. Every schedule gives the same bits */
} lps_options;

/* One output record: the quantities src/main.cpp:986-1062 computes every 30 steps. */
typedef struct lps_diag {
  int64_t step;
  double t, energy, energy_k, energy_p, px, py, max_u, max_v, max_a, theta_max, theta_min;
  double ub_rms, vb_rms, de_rel;
  int32_t max_u_node, max_v_node; /* original node ids */
} lps_diag;

typedef struct lps_ctx lps_ctx;

/* Fills *opt with the reference's configuration (G_theta, ABC on, fast path, graphs on). */
void lps_default_options(lps_options *opt);

/* Setup of src/main.cpp:178-684: validate and upload the problem, rebuild the horizons on the GPU
 * (replaces build_clouds, src/helpers/kd_tree_interface.h:41-126, called at src/main.cpp:352-354),
 * compute m_i, m_ref and the ABC override (:408-480), the LPS constants (:485-504), the output
 * exclusion radius (:518-535), the initial state (:624-634) and the theta history (:664-682). */
int32_t lps_create(const lps_problem *problem, const lps_options *options, lps_ctx **out);

/* Back to step 0: the initial conditions, theta history and first prediction of lps_create, without
 * rebuilding the horizons. */
int32_t lps_reset(lps_ctx *ctx);

/* New initial conditions (.icvalues / .icdvalues contents, 2 node_no each, the problem's node order)
 * followed by lps_reset: the horizons and constants of lps_create are kept. No counterpart in the
 * reference, which reads the initial conditions once per process (src/main.cpp:243-250, 624-634).
 * Single-problem contexts only (LPS_ERR_ARG on a batch). */
int32_t lps_set_initial(lps_ctx *ctx, const double *ic, const double *icd);

/* Checkpoint and restart. No counterpart in the reference, whose time loop (src/main.cpp:716-1166)
 * always starts from the initial conditions. lps_checkpoint_save writes the state the next step
 * reads (U^n, U^{n-1}, V^n, V^{n-1}, A^n, the prediction, theta^n, theta^{n-1}), the step count
 * and E0 of every instance to a binary file. lps_checkpoint_load restores such a file into a
 * context created from the same problem(s) and options; stepping on then gives the same bits as
 * the uninterrupted run. A file of another problem, layout or option set is refused (LPS_ERR_ARG)
 * before the context is touched. Neither works with a learned closure attached (LPS_ERR_ARG).
 * Both block until the device is done. */
int32_t lps_checkpoint_save(lps_ctx *ctx, const char *path);
int32_t lps_checkpoint_load(lps_ctx *ctx, const char *path);

/* Batches. No counterpart in the reference, which runs one job per process: many independent
 * problems (own radius, nu, dt, ABC operators, initial conditions; the same dx, dy and delta)
 * advanced together, every launch of a step covering all of them. Each problem is set up as
 * lps_create would (GPU horizons, m_i, tile check, theta history, prediction), then the instances
 * are packed into one slot space. Batches run the tile path only: every problem must be a lattice
 * with the 68-bond stencil (LPS_ERR_ARG otherwise); options->variant 2 selects the unrolled node
 * kernels, anything else the register-blocked ones. lps_step, lps_reset, lps_set_timing,
 * lps_pass_times apply as for one problem; lps_run, lps_diagnostics, lps_fetch and lps_horizons
 * return LPS_ERR_ARG on a batch. */
int32_t lps_create_batch(const lps_problem *problems, int32_t count, const lps_options *options,
                         lps_ctx **out);
int32_t lps_instance_count(const lps_ctx *ctx);
/* lps_run for every instance: records[k * count + i] is output record k of instance i (E0, dE/E0,
 * time and ABC-row RMS per instance). Returns the number of output steps, or LPS_ERR_NONFINITE
 * (the message names the instance). Works on a single-problem context too (count 1). */
int32_t lps_run_batch(lps_ctx *ctx, int32_t nsteps, lps_diag *records, int32_t max_records);
/* The output block of every instance at the current step (count records). Tile path only. */
int32_t lps_diagnostics_batch(lps_ctx *ctx, lps_diag *out);
/* lps_fetch of one instance, in that problem's node order. */
int32_t lps_fetch_instance(lps_ctx *ctx, int32_t instance, double *u, double *v, double *a,
                           double *theta, double *m);

/* Advance nsteps steps of the time loop (src/main.cpp:718-1164, steps 1-4 and 7). Stream-ordered;
 * returns without waiting for the device. */
int32_t lps_step(lps_ctx *ctx, int32_t nsteps);

/* Time loop with the output block: advance nsteps; every step s with s % 30 == 0 produces a record
 * (src/main.cpp:981-1123), E0 being the first record's energy. Up to max_records records are
 * written to records (may be NULL). Returns the number of records produced, or LPS_ERR_NONFINITE
 * at the first record whose energy or max|u| is NaN/Inf (the state then stays at that step).
 * Returns once the device has finished the steps. */
int32_t lps_run(lps_ctx *ctx, int32_t nsteps, lps_diag *records, int32_t max_records);

/* The output-block quantities of the current state (src/main.cpp:986-1062); de_rel uses the E0 of
 * the context (0 before the first record). Blocks until the device is done. */
int32_t lps_diagnostics(lps_ctx *ctx, lps_diag *out);

/* Current state in the original node order: u, v, a interleaved (2 node_no each), theta and the
 * weighted volume m (node_no each). Any pointer may be NULL. Blocks until done. */
int32_t lps_fetch(lps_ctx *ctx, double *u, double *v, double *a, double *theta, double *m);

/* Horizons built on the GPU, in original node ids and order: row_ptr (node_no + 1 entries) and
 * cols (self first, then ascending distance, ties by ascending id). Call with cols == NULL to get
 * the entry count. Returns the entry count (>= 0) or an error. */
int64_t lps_horizons(lps_ctx *ctx, int64_t *row_ptr, int32_t *cols);

/* Completed steps, m_ref and the exclusion radius of the context. */
int64_t lps_step_index(const lps_ctx *ctx);
double lps_m_ref(const lps_ctx *ctx);

/* Average device time (ms) per step of the kernels launched by lps_step since timing was switched
 * on, measured with HIP events on the context's stream: out[0] dilatation pass (exact path with
 * the ABC rows; fast path the tile kernels, and the IDW override), out[1] the force pass's main
 * kernel (exact path: nodes and ABC rows; fast path: the selected tile kernel), out[2] the rest of
 * the step (fast path: the partial tiles' force kernel, the wait for the ABC rows' G_u / G_v, which
 * run on a second stream beside the tile kernels, and the ABC prediction), out[3] whole step.
 * lps_set_timing(ctx, 1) records four events per step, lps_set_timing(ctx, 2) only the two around
 * the force pass's main kernel (out[1]; the others read 0), lps_set_timing(ctx, 3) only two around
 * the learned closure's network kernel on its stream (out[1]; a closure must be attached), 0
 * switches timing off. Timed steps
 * replay 30-step graphs with the events as external event-record nodes; where the runtime cannot
 * time such events, they run as separate launches. */
int32_t lps_set_timing(lps_ctx *ctx, int32_t on);
int32_t lps_pass_times(lps_ctx *ctx, double *out4);

/* Which kernels the context runs: the tile path (lattice inputs whose stencil reproduces every
 * GPU-built horizon) or the exact CSR path, with the layout. */
const char *lps_describe(const lps_ctx *ctx);

/* Tile-path layout of the context: out[0] tiles, out[1] full tiles (every window cell a node with
 * m == m_ref; the register-blocked kernels take them), out[2] partial tiles, out[3] nodes in the
 * full tiles, out[4] nodes (all zeros on the exact path). */
int32_t lps_tile_info(const lps_ctx *ctx, int64_t *out5);

const char *lps_last_error(const lps_ctx *ctx);
const char *lps_global_error(void); /* message of a failed lps_create */
void lps_destroy(lps_ctx *ctx);

/* ------------------------------------------------------------------------------------------------
 * Learned ABC closure. No counterpart in the reference, whose ABC is the least-squares update
 * G_u / G_v (src/main.cpp:883-974); the closure adds a learned correction to that update at every
 * ABC row, and the least-squares update stays what it corrects and what a correction of zero
 * gives back bit for bit.
 *
 * Network: 1600 inputs -> 256 -> 256 -> 256 (ReLU) -> 4. Layers 1-3 run on the matrix cores with
 * BF16 operands and FP32 accumulation; the output layer is FP32. Inputs of an ABC row, in the
 * lattice frame that maps the row's outward normal into the 0-45 degree sector: the last 8 steps
 * of u and v at the 49 lattice offsets |o|^2 <= 16 around the node (zero where no node is), u
 * divided by s_u = max |u| and v by s_v = max |v| over those inputs, then nu' = (nu - 0.3) / 0.15,
 * the canonical normal (n'_x, n'_y) and the depth (R - |x|) / delta. Outputs, in the same frame:
 * the corrections of U^{n+1} and V^{n+1} in units of s_u sigma_u and s_v sigma_v.
 *
 * Samples: inputs (1600 BF16 values, as uint16 bit patterns) and aux (8 floats: the 4 targets in
 * the units above without sigma, the weight, s_u, s_v, and the row it came from).
 * ---------------------------------------------------------------------------------------------- */
#define LPS_CLOSURE_IN 1600
#define LPS_CLOSURE_HIDDEN 256
#define LPS_CLOSURE_OUT 4
typedef struct lps_closure lps_closure;

/* Number of FP32 parameters: W1 [256][1600], b1 [256], W2 [256][256], b2, W3 [256][256], b3,
 * W4 [4][256], b4 [4], in that order, weights [out][in]. */
int64_t lps_closure_param_count(void);
/* A network on a device, initialised from seed (He-normal hidden layers, output layer near 0). */
int32_t lps_closure_create(int32_t device, uint64_t seed, lps_closure **out);
void lps_closure_destroy(lps_closure *clo);
const char *lps_closure_error(const lps_closure *clo);
int32_t lps_closure_get_params(lps_closure *clo, float *params);
/* Replaces the parameters and clears the optimizer state. */
int32_t lps_closure_set_params(lps_closure *clo, const float *params);
/* Output scales sigma_u, sigma_v (the targets' spread; the network learns targets / sigma). */
int32_t lps_closure_set_scales(lps_closure *clo, float sigma_u, float sigma_v);
int32_t lps_closure_get_scales(lps_closure *clo, float *sigma2);

/* Forward pass over n rows of inputs: out is n x 4 (units of sigma). */
int32_t lps_closure_forward(lps_closure *clo, const uint16_t *x, int64_t n, float *out);
/* lps_closure_forward without device allocation. No counterpart in the reference; it is for
 * callers that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc):
 * the same kernel and the same bits as lps_closure_forward, with the rows staged in a fixed area
 * of 6.6 MB that the library holds on each GPU until the process ends, instead of a hipMalloc /
 * hipFree pair per call. n is at most 2048 rows (LPS_ERR_ARG beyond). All calls on one GPU share
 * that area: calls that overlap on one GPU race. */
int32_t lps_closure_forward_noalloc(lps_closure *clo, const uint16_t *x, int64_t n, float *out);
/* One Adam step (lr; beta 0.9 / 0.999, eps 1e-8) on a batch of n rows (n a multiple of 64) given
 * by the host; loss = sum_r w_r |f(x_r) - t_r / sigma|^2 / n. grads (param_count floats, may be
 * NULL) receives the gradient of that loss before the update. lr = 0 leaves the weights. */
int32_t lps_closure_train_batch(lps_closure *clo, const uint16_t *x, const float *target,
                                const float *weight, int32_t n, float lr, float *loss,
                                float *grads);

/* The sample store on the closure's device. */
int32_t lps_closure_dataset_reserve(lps_closure *clo, int64_t capacity);
int64_t lps_closure_dataset_count(lps_closure *clo);
int32_t lps_closure_dataset_clear(lps_closure *clo);
int32_t lps_closure_dataset_append(lps_closure *clo, const uint16_t *x, const float *aux,
                                   int64_t n);
int32_t lps_closure_dataset_read(lps_closure *clo, int64_t first, int64_t n, uint16_t *x,
                                 float *aux);
/* nsteps Adam steps on batches of `batch` rows (a multiple of 64) drawn with replacement from
 * samples [first, first + count); the step size follows a cosine from lr0 to lr1. input_noise > 0
 * adds Gaussian noise of that standard deviation to the drawn history inputs (robustness to
 * perturbed histories). loss_curve (nsteps floats, may be NULL) receives each step's batch loss. */
int32_t lps_closure_train(lps_closure *clo, int32_t nsteps, int32_t batch, float lr0, float lr1,
                          int64_t first, int64_t count, uint64_t seed, float input_noise,
                          float *loss_curve);
/* Over samples [first, first + count): out[0] = sum w |f - t/sigma|^2 (sigma units),
 * out[1] = sum w |t/sigma|^2 (the least-squares ABC alone), out[2] = sum w. */
int32_t lps_closure_evaluate(lps_closure *clo, int64_t first, int64_t count, double *out3);

/* Teacher-forced samples from a reflection-free run. truth: a context on a larger domain of the
 * same origin-centred lattice (the fast path); trunc: the truncated problem whose ABC rows are
 * emulated (its nodes lie on truth's lattice). Runs truth for nsteps steps from its current state
 * with an empty history; after each step n + 1 with n + 1 >= first_step and
 * (n + 1) % sample_every == 0, every row of trunc whose hash (seed, step, row) falls under keep
 * and whose s_u >= amp_min appends one sample: its inputs at step n from truth's history, and as
 * target truth's U^{n+1}, V^{n+1} at the row's node minus trunc's G_u / G_v applied to truth's
 * U^n, U^{n-1}, V^n, V^{n-1} (weight 1). A sample whose target exceeds y_max in magnitude (a wave
 * front rising faster than its history shows) is skipped. Samples beyond the capacity are dropped
 * (the count still grows). Errors are reported on truth. */
int32_t lps_closure_collect(lps_closure *clo, lps_ctx *truth, const lps_problem *trunc,
                            int32_t nsteps, int32_t first_step, int32_t sample_every, double keep,
                            double amp_min, double y_max, uint64_t seed);

/* Closed-loop samples (the next round of a DAgger-style training). trunc: a single fast-path
 * context with a closure attached (it corrects trunc's ABC rows as it steps); truth: a context on a
 * larger domain of the same lattice, without ABC rows, at the same step and time step. Both advance
 * nsteps steps together; after each step n + 1 selected as in lps_closure_collect, every kept row
 * whose s_u >= amp_min appends one sample: as inputs the history the closure read at step n (the
 * truncated run's own, with its reflections and the closure's past corrections), as target truth's
 * U^{n+1}, V^{n+1} at the row's node minus trunc's G_u / G_v applied to trunc's U^n, U^{n-1}, V^n,
 * V^{n-1}: the correction that would put the truncated run back on truth. Errors are reported on
 * trunc. */
int32_t lps_closure_collect_closed(lps_closure *clo, lps_ctx *truth, lps_ctx *trunc,
                                   int32_t nsteps, int32_t first_step, int32_t sample_every,
                                   double keep, double amp_min, double y_max, uint64_t seed);

/* Applies clo at every ABC row of ctx (fast path only; a batch uses each instance's nu and
 * radius) from the next step on; the history starts empty (zeros) and is cleared by lps_reset.
 * clo must outlive the attachment. Replaces an earlier attachment. */
int32_t lps_closure_attach(lps_ctx *ctx, lps_closure *clo);
int32_t lps_closure_detach(lps_ctx *ctx);
/* Factor on the correction per instance (count = lps_instance_count; 1 after attach; 0 gives the
 * least-squares ABC). */
int32_t lps_closure_set_gain(lps_ctx *ctx, const float *gain);
/* Amplitude gate per instance (count = lps_instance_count; 0 after attach): no correction at a
 * row whose s_u (max |u| over its inputs) is below min_su, the range a closure was not trained on
 * (lps_closure_collect's amp_min). */
int32_t lps_closure_set_gate(lps_ctx *ctx, const float *min_su);
/* Fallback guard. At every output record (every 30 steps, lps_run / lps_run_batch) an instance
 * whose energy rose by more than tol * E0 since the previous record is switched to the
 * least-squares ABC (gain 0) for the rest of the run: absorbing boundaries take energy out, they do
 * not add it. tol >= 0 sets the tolerance (0: no guard; 1e-3 after attach); a negative tol keeps
 * it. fallback_step (count = lps_instance_count, may be NULL) receives the step of each
 * instance's fallback, -1 where the closure is still on. lps_reset and lps_closure_set_gain put
 * the gains back. */
int32_t lps_closure_guard(lps_ctx *ctx, double tol, int64_t *fallback_step);
/* Inputs and scales of ctx's ABC rows from the history of the last completed step (n_abc x 1600
 * BF16 bits, n_abc x 2), and the correction that step applied (n_abc x 8: 4 outputs, s_u, s_v,
 * 0, 0); any pointer may be NULL. */
int32_t lps_closure_inspect(lps_ctx *ctx, uint16_t *x, float *scale2, float *applied8);

/* ------------------------------------------------------------------------------------------------
 * Least-squares ABC operators on the GPU: the per-row solves of the reference's preprocessor
 * (G_u / G_v preprocessor/main.cpp:415-637, G_theta :657-785, the dgelsd solve :368-402) for every
 * ABC row at once. The caller gives the wave modes (the dispersion relation, :199-356) and the rows'
 * clouds (:850-902); the errors of these calls are reported by lps_global_error(). */
typedef struct lps_abc_modes {
  int32_t count;           /* wave modes, in the order of the system's rows (4 rows per mode) */
  const double *a1, *a2;   /* wave vector (count each) */
  const double *om_p, *om_s; /* P and S angular frequencies */
  const double *g_p, *g_s; /* P and S polarisations, count x 2 */
  const double *s_r, *s_i; /* dilatation factor S, real and imaginary parts, count x 2 */
} lps_abc_modes;

typedef struct lps_abc_rows {
  int32_t count;           /* ABC rows */
  int32_t max_cloud;       /* cloud slots per row (maxCloudSize) */
  const double *xy;        /* count x 2: the ABC node */
  const double *tau;       /* count: its outward angle */
  const int32_t *cloud_len; /* count: cloud size, 1 .. max_cloud */
  const double *cloud_xy;  /* count x max_cloud x 2: the cloud nodes (the first cloud_len valid) */
} lps_abc_rows;

/* G_u, G_v (2 count x 4 max_cloud each, the .ciNodesU / .ciNodesV layout) and, when gth is not
 * NULL, G_theta (count x 2 max_cloud, .ciNodesTheta) for time step dt; theta_coeff = 2 / m_ref.
 * Minimum-norm least squares with the singular values s_i <= rcond s_1 dropped, as dgelsd
 * (rcond 1e-14 in the reference); as in dgelsd, rcond <= 0 or >= 1 means the machine epsilon
 * 2^-53, and a NaN rcond is refused (LPS_ERR_ARG). rank (count x 3, may be NULL) receives the number of singular
 * values kept for G_u, G_v, G_theta, or -1 where the SVD did not converge. */
int32_t lps_abc_operators(int32_t device, const lps_abc_rows *rows, const lps_abc_modes *modes,
                          double dt, double theta_coeff, double rcond, double *gu, double *gv,
                          double *gth, int32_t *rank);
/* The least-squares systems of rows first .. first + count - 1 for one operator (kind 0 G_u,
 * 1 G_v, 2 G_theta): A (count x m x n, each column-major, m = 4 modes, n = 4 or 2 max_cloud) and
 * the right-hand sides B (count x m x nrhs, nrhs 2 or 1). */
int32_t lps_abc_system(int32_t device, const lps_abc_rows *rows, const lps_abc_modes *modes,
                       double dt, double theta_coeff, int32_t kind, int32_t first, int32_t count,
                       double *A, double *B);
/* The batched solver under lps_abc_operators on caller systems: for each of batch problems,
 * X (n x nrhs) = minimum-norm least-squares solution of A (m x n) X = B (m x nrhs), all
 * column-major and contiguous per problem, singular values s_i <= rcond s_1 dropped. rank (batch,
 * may be NULL) as in lps_abc_operators. */
int32_t lps_lstsq_batched(int32_t device, int32_t m, int32_t n, int32_t nrhs, int32_t batch,
                          const double *A, const double *B, double rcond, double *X,
                          int32_t *rank);

#ifdef __cplusplus
}
#endif
#endif
