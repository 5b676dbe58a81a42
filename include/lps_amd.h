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
                            second stream beside the tile kernels; 0 on the one stream */
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
 * Timing is enabled with lps_set_timing(ctx, 1); timed steps run as separate launches, without
 * the HIP graph (HIP cannot time events recorded inside a graph). */
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

#ifdef __cplusplus
}
#endif
#endif
