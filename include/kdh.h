/*
 * kdh.h -- C ABI of the MI355X scan-alignment library (libkdhip.so).
 *
 * The library replaces the GPU side of the reference's aligner plugins
 * ICPCUDA3D / CTICPCUDA3D (rvp-group/kd_slam2).  A C++ adapter that mirrors the
 * reference's TreeAlignerBase_ / ICP_ / CTICP_ interface (registered as the
 * configurable types ICPHIP3D / CTICPHIP3D) calls these entry points; see
 * kd_slam2_amd/adapter/.  Only plain pointers and sizes cross this boundary.
 *
 * Conventions
 *  - Every call returns KDH_OK (0) or a negative KDH_ERR_* code and never
 *    exits the process (the reference's CUDA_CHECK exits,
 *    kd_slam2/src/lib/kd_slam/cuda/cuda_common.h:26-35).  kdh_last_error()
 *    returns the message of the last failing call on that context.
 *  - Poses cross the boundary as 3x4 row-major [R | t] arrays (float for the
 *    FP32 per-iteration cache the reference forms in updateCache, double for
 *    the FP64 state the device loop keeps).
 *  - Quadratic forms come back as full symmetric row-major FP64 matrices:
 *    6x6 for ICP, 12x12 [H_pp H_pv; H_pv^T H_vv] for CT-ICP, as assembled by
 *    the reference (icp/icp_ws_item_.h:36-44, cticp/cticp_cpu_impl_.h:84-93).
 *  - Calls on one context are serialised by the context; contexts are
 *    independent.  One context owns one HIP stream on one device.
 */
#ifndef KDH_H_
#define KDH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KDH_OK 0
#define KDH_ERR_INVALID (-1)   /* bad argument or misuse */
#define KDH_ERR_HIP (-2)       /* HIP runtime error */
#define KDH_ERR_NOMEM (-3)     /* device or host allocation failed */
#define KDH_ERR_BUDGET (-4)    /* request exceeds the serving memory budget */
#define KDH_ERR_NOT_FOUND (-5) /* unknown request / handle */
#define KDH_ERR_CANCELLED (-6) /* request cancelled before it completed */

typedef struct kdh_ctx kdh_ctx;
typedef struct kdh_tree kdh_tree;
typedef struct kdh_server kdh_server;

/* Memory layout of the reference's host node Node_<Point3fTraits>
 * (kd_slam2/src/lib/kd_slam/tree/node_.h:12-43: TimestampInfo base, then
 * _mean, _direction, _idx_start, _idx_end, _idx_left, _idx_right); 64 bytes.
 * A TreeCPU_ node array can be passed to kdh_tree_upload as is. */
typedef struct kdh_node3f {
  double dts_mean;
  double dts_covariance;
  float mean[3];
  float direction[3];
  uint64_t idx_start;
  uint64_t idx_end;
  int32_t idx_left;
  int32_t idx_right;
} kdh_node3f;

/* ICPParams_ / CTICPParams_ (icp/tree_aligner_base_.h:12-21,
 * cticp/cticp_.h:14-18) plus the score termination criterion
 * (icp/icp_stats_.h:85-102) and the Jacobian re-centring switch. */
typedef struct kdh_params {
  float kernel_threshold;      /* 0.1 */
  float damping;               /* 1 */
  float max_error;             /* 1 */
  float min_normal_cos;        /* cos(25 deg) */
  int32_t min_good_measurements; /* 10 */
  int32_t max_iterations;      /* 20 */
  float velocity_damping;      /* 100 (CT only) */
  int32_t use_term_crit;       /* 0: run max_iterations */
  float term_d_pert;           /* 1e-3 */
  float term_d_score;          /* 1e-3 */
  int32_t recenter;            /* 1: residual and Jacobian evaluated in frame-
                                  local FP32 coordinates about an anchor at the
                                  scan centroid, mapped back in FP64 (exact in
                                  real arithmetic; no cancellation far from the
                                  origin); 0: the reference's world-coordinate
                                  formulas.  Correspondences, gates and inlier
                                  decisions are the reference's in both. */
} kdh_params;

/* ICPStats_ (icp/icp_stats_.h:10-20). */
typedef struct kdh_stats {
  int32_t num_inliers;
  int32_t num_outliers;
  int32_t num_bad;
  int32_t num_err_normal;
  int32_t num_err_fn;
  int32_t pad;
  double chi2;
  double chi2_kernelized;
  double pert_pose_norm;
  double pert_vel_norm;
} kdh_stats;

/* Pose prior of TreeAlignerBase_::setPosePrior (tree_aligner_base_.h:127-133). */
typedef struct kdh_prior {
  double X_ref[12];
  double Z[12];
  double info[36];
} kdh_prior;

/* Fills p with the reference defaults. */
void kdh_params_default(kdh_params* p);

/* ---- context ------------------------------------------------------------
 * Replaces the CUDA library's load-time device set-up
 * (kd_slam2/src/lib/kd_slam_cuda_init.cu:3-6). */
int kdh_ctx_create(int device, kdh_ctx** out);
int kdh_ctx_destroy(kdh_ctx* ctx);
const char* kdh_last_error(const kdh_ctx* ctx);

/* ---- trees --------------------------------------------------------------
 * Replaces Tree_CUDAfromCPU_ (kd_slam2/src/lib/kd_slam/tree/tree_cuda_impl_.h:198-222)
 * and the TreeCUDA_ destructor (:31-40).  nodes/leaf_indices are the
 * TreeCPU_ arrays _nodes_ptr / _leaves_indices_ptr (tree_.h:27-33).  The tree
 * is re-laid out for the device (32-byte node records, packed leaf records)
 * and stays resident until kdh_tree_free. */
int kdh_tree_upload(kdh_ctx* ctx, const kdh_node3f* nodes, int64_t n_nodes,
                    const uint64_t* leaf_indices, int64_t n_leaves, kdh_tree** out);
int kdh_tree_free(kdh_ctx* ctx, kdh_tree* tree);
int kdh_tree_info(const kdh_tree* tree, int64_t* n_nodes, int64_t* n_leaves, int64_t* device_bytes);

/* ---- tree maintenance on the device ---------------------------------------
 * Replaces TreeCUDA_::applyVelocitiesInPlace / treeApplyVelocities_kernel
 * (tree/tree_cuda_impl_.h:138-160): every node's mean and direction warped by
 * R(w dts) and v dts with the node's own (float) _dts_mean, leaves skipped
 * when skip_leaves (Node_::applyVelocities, tree/node_impl_.h:86-92).
 * v = [t; w], 6 floats.  The re-centring anchor follows the root. */
int kdh_tree_apply_velocities(kdh_ctx* ctx, kdh_tree* tree, const float v[6], int skip_leaves);

/* Replaces TreeCUDA_::applyIsometryInPlace / treeApplyIsometry_kernel
 * (tree/tree_cuda_impl_.h:162-173; applyIsometry_ on a node, tree/node_.h:127-137):
 * mean <- iso * mean, direction <- R * direction.  iso: 3x4 row-major. */
int kdh_tree_apply_isometry(kdh_ctx* ctx, kdh_tree* tree, const float iso[12]);

/* Replaces TreeCUDA_::extractCompressedLeaves() and (v) (tree/tree_cuda_impl_.h:43-136):
 * per leaf, in leaf-index order, mean then direction (out: 6 * n_leaves
 * floats), warped by v when v is not NULL. */
int kdh_tree_extract_leaves(kdh_ctx* ctx, const kdh_tree* tree, const float* v, float* out);

/* Parity hook: every node's mean and direction as the device holds them
 * (out: 6 * n_nodes floats).  No reference counterpart. */
int kdh_tree_read_nodes(kdh_ctx* ctx, const kdh_tree* tree, float* out);

/* TreeCUDA_::clone (Tree_CUDAfromCUDA_, tree/tree_cuda_impl_.h:176-195;
 * tree/tree_cuda_.h:57): a device-to-device deep copy of a resident tree
 * (every record, the stamps and, for device-built trees, the full records,
 * leaf list and points).  Free it with kdh_tree_free. */
int kdh_tree_clone(kdh_ctx* ctx, const kdh_tree* src, kdh_tree** out);
/* TreeCUDA_::copyNodes (tree/tree_cuda_impl_.h:60-69; tree/tree_cuda_.h:43):
 * overwrites dst's node and leaf records (and full records) with src's, on
 * the device; KDH_ERR_INVALID on a size mismatch, as the reference's
 * "size mismatch".  The restore of the bundle-adjustment pre-deskew
 * (optimizer/optimizer_proc_solver_action_.h:39-58, map/map_.h:88-102). */
int kdh_tree_copy_nodes(kdh_ctx* ctx, const kdh_tree* src, kdh_tree* dst);

/* ---- one quadratic form (the plain per-call path) -------------------------
 * Replaces ICP_CUDA_::_buildQuadraticForm (icp/icp_cuda_impl_.cuh:95-132):
 * for every leaf of `moving`, descend `fixed`, apply the gates, and reduce
 * the point-to-plane normal equations.  Xm, Xf, Xmf are the FP32
 * QuadraticTermCache (icp/icp_.h:97-102) the caller formed in updateCache
 * (icp/icp_impl_.h:16-21).  stats_mode as in quadraticTerm (outliers get
 * weight 0).  H: 36 doubles, b: 6 doubles. */
int kdh_icp_build(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving,
                  const float Xm[12], const float Xf[12], const float Xmf[12],
                  const kdh_params* params, int stats_mode,
                  double H[36], double b[6], kdh_stats* stats);

/* Replaces CTICP_CUDA_::_buildQuadraticForm (cticp/cticp_cuda_impl_.cuh:163-223):
 * as kdh_icp_build with per-leaf deskew by the moving velocities v ([t; w],
 * cticp/cticp_impl_.h:118-121) and the 12-DoF pose + velocity system.
 * H: 144 doubles, b: 12 doubles. */
int kdh_cticp_build(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving,
                    const float Xm[12], const float Xf[12], const float Xmf[12],
                    const float v[6], const kdh_params* params, int stats_mode,
                    double H[144], double b[12], kdh_stats* stats);

/* The form of kdh_icp_build (ct = 0) or kdh_cticp_build (ct = 1) at an FP64
 * state instead of an FP32 cache: the cache is rounded once from (X, Xf, v),
 * as kdh_align does every iteration, and re-centring (params->recenter)
 * anchors at the FP64 state.  X, Xf: 3x4 row-major; v: 6 doubles (CT only,
 * may be NULL for ct = 0).  No reference counterpart: it is the form one
 * iteration of the device loop builds, exposed for callers that keep their
 * pose in FP64 and for precision checks far from the origin. */
int kdh_build_at(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct,
                 const double X[12], const double Xf[12], const double v[6],
                 const kdh_params* params, int stats_mode,
                 double* H, double* b, kdh_stats* stats);

/* Parity hook: the fixed-tree leaf index each moving leaf descends to
 * (-1 = num_bad) and the number of internal nodes visited, at the given
 * cache.  The reference has no counterpart; it exposes the discrete part of
 * quadraticTerm (icp_impl_.h:113-121, cticp_impl_.h:121-131) for checking.
 * leaf_out / depth_out hold n_leaves(moving) int32 each (depth_out may be NULL). */
int kdh_correspondences(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct,
                        const float Xm[12], const float Xf[12], const float Xmf[12],
                        const float v[6], const kdh_params* params,
                        int32_t* leaf_out, int32_t* depth_out);

/* ---- whole alignment on the device ----------------------------------------
 * TreeAlignerBase_::compute (icp/tree_aligner_base_.h:100-115) with the
 * Gauss-Newton loop, the FP64 damped solve and the SE(3) update resident on
 * the device and the iterations replayed from a HIP graph.  X0/v0: initial
 * moving pose (3x4) and velocities (CT), Xf: fixed pose.  prior may be NULL.
 * stats_log receives one entry per executed iteration (params->max_iterations
 * entries, or 1 in stats mode); *result is compute()'s return value (i on
 * early stop, -i on a failed round, max_iterations otherwise).
 * cache_log (optional, 3 x 12 floats per iteration) receives the FP32
 * Xm, Xf, Xmf each iteration used, for parity checks. */
int kdh_align(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct,
              const double X0[12], const double v0[6], const double Xf[12],
              const kdh_params* params, const kdh_prior* prior, int stats_mode,
              double X_out[12], double v_out[6], kdh_stats* stats_log, int* n_log,
              int* result, float* cache_log);

/* kdh_align without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same loop,
 * kernels and results as kdh_align (no cache log); its pair state and partial
 * rows live in a fixed device array of the library (1.6 MB per GPU, held
 * until the process ends), so a moving tree of more than
 * kdh_align_noalloc_max_leaves() leaves (1,048,576) is refused with
 * KDH_ERR_INVALID.  Calls on one GPU share that array: the library serialises
 * them per device (they would race otherwise); plain launches, no graph. */
int kdh_align_noalloc(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct,
                      const double X0[12], const double v0[6], const double Xf[12],
                      const kdh_params* params, const kdh_prior* prior, int stats_mode,
                      double X_out[12], double v_out[6], kdh_stats* stats_log, int* n_log, int* result);
int kdh_align_noalloc_max_leaves(void);  /* at the default 256-leaf tiles (half at 128) */

/* ---- many alignments in one launch -------------------------------------- */
typedef struct kdh_pair {
  const kdh_tree* fixed;
  const kdh_tree* moving;
  int32_t ct;
  int32_t forest;              /* 0: a plain pair.  k > 0: this pair and the
                                  next k pairs are one alignment against a
                                  fixed forest (ICP_::_fixed_forest +
                                  makeFixedEntry, icp/icp_impl_.h:53-71,
                                  icp/icp_.h:90-95): the next k pairs give only
                                  their fixed tree and Xf; they must name the
                                  same moving tree and ct.  H, b and the stats
                                  are summed over the entries each round. */
  double X0[12];
  double v0[6];
  double Xf[12];
  kdh_params params;
  const kdh_prior* prior;      /* may be NULL */
} kdh_pair;

typedef struct kdh_pair_result {
  double X[12];
  double v[6];
  kdh_stats last;              /* the aligner's stats of the last round */
  kdh_stats entry;             /* this pair's own fixed tree's share of the
                                  last round (FixedEntry::stats); equal to
                                  last for a plain pair.  No pose norms. */
  int32_t iterations;
  int32_t result;
} kdh_pair_result;

/* Runs every pair (or forest) to completion: one tile launch and one finish
 * launch per Gauss-Newton iteration over the (pair, tile) list, captured as
 * a HIP graph.  Each result is bit-identical to running it alone through
 * kdh_batch_align; forest entries' results repeat their leader's pose. */
int kdh_batch_align(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, kdh_pair_result* results);

/* A batch prepared once and run many times with new initial states: the
 * loop-closure verifier's shape, the same candidate set re-verified from new
 * guesses (the batched form of the sequential candidate loop,
 * slam/slam_proc_loops_impl_.h:95-115).  kdh_batch_prepare validates and
 * uploads the pairs as kdh_batch_align does and keeps them on the device with
 * their own buffers; the trees must outlive the batch.  kdh_batch_run replays
 * one graph holding the whole run -- the states initialised on the device
 * from X0 (n_pairs x 12, row-major 3x4; NULL keeps the previous run's, first
 * the pairs' own) and v0 (n_pairs x 6, CT pairs; NULL likewise), every
 * iteration, the results written back to pinned host memory -- so a run is one
 * graph launch and one synchronisation.  Results are bit-identical to
 * kdh_batch_align on the same pairs and initial states (forest entries take
 * their leader's X0 / v0).  The graph follows the context's launch
 * configuration at run time.  Runs of one batch, and calls on its context,
 * are serialised by the context. */
typedef struct kdh_batch kdh_batch;
int kdh_batch_prepare(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, kdh_batch** out);
int kdh_batch_run(kdh_batch* batch, const double* X0, const double* v0, kdh_pair_result* results);
int kdh_batch_release(kdh_batch* batch);

/* kdh_batch_run in two halves, so that a caller keeps the device busy while
 * it collects one run's results and prepares the next run's inputs: the
 * reference has no counterpart (its verifier runs one candidate at a time);
 * it is for serving loops over a prepared batch.  kdh_batch_submit enqueues
 * one run from X0 / v0 (as kdh_batch_run) and returns at once; at most two
 * runs are in flight, each with its own pinned inputs and results.
 * kdh_batch_collect waits for the oldest submitted run and writes its
 * results.  Runs execute in submission order on the context's stream, each
 * from its own initial states, so each result equals kdh_batch_run's on the
 * same inputs bit for bit.  kdh_batch_run is refused while runs are in
 * flight, and kdh_batch_submit in profiling mode. */
int kdh_batch_submit(kdh_batch* batch, const double* X0, const double* v0);
int kdh_batch_collect(kdh_batch* batch, kdh_pair_result* results);

/* One quadratic form per pair (or forest) at its X0 / v0 / Xf, all in one
 * tile launch and one finish launch: the per-factor builds of a bundle
 * adjustment batched (MultiViewICPFactor_::compute calls buildQuadraticForm
 * per factor, map/multiview_icp_factor_impl_.h:10-62).  For a rigid factor
 * the two-pose blocks follow from J_B = -J_A (icp/icp_impl_.h:44-50):
 * H_AA = H_BB = H, H_AB = -H, b_B = -b.  The cache is rounded from the FP64
 * state as kdh_build_at does; stats_mode as in quadraticTerm.
 * Outputs per pair i: H + 144 i (D x D row-major, D = 6 or 12), b + 12 i,
 * stats[i].  Forest entries repeat their leader's summed form. */
int kdh_batch_build(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, int stats_mode,
                    double* H, double* b, kdh_stats* stats);

/* ---- CT dual form: bundle-adjustment factors over two CT states -----------
 *
 * One CT factor links a moving cloud A (pose X_moving, velocities v_moving)
 * and a fixed cloud B (pose X_fixed, velocities v_fixed).  Besides the CT
 * form of A (H 12x12, b 12, as kdh_build_ct) the dual build gives the terms
 * of B's velocities, taken at each fixed leaf's own stamp (CTICP_::
 * quadraticTermDual, cticp/cticp_impl_.h:185-293; CTICP_CPU_::
 * _buildQuadraticFormDual, cticp/cticp_cpu_impl_.h:129-205), which
 * MultiViewCTICPFactor_::compute scatters into the four-variable factor
 * (map/multiview_cticp_factor_impl_.h:8-80). */
typedef struct kdh_dual_pair {
  const kdh_tree* fixed;       /* cloud B: its leaves' dts_mean are the stamps of J_vel_B */
  const kdh_tree* moving;      /* cloud A */
  double X_moving[12];         /* moving_state.X */
  double v_moving[6];          /* moving_state.velocities [t; w] */
  double X_fixed[12];          /* fixed_state.X */
  double v_fixed[6];           /* fixed_state.velocities [t; w] */
  kdh_params params;
} kdh_dual_pair;

typedef struct kdh_dual_form {
  double H[144];                        /* 12x12 row-major: CTICP_::H */
  double b[12];                         /* CTICP_::b */
  double H_dual_vel_B[36];              /* 6x6: sum g J_vB^T J_vB */
  double b_dual_vel_B[6];               /* sum g J_vB^T e */
  double H_dual_cross_pose_vel_B[36];   /* 6x6: sum g J_pose^T J_vB (row: pose) */
  double H_dual_cross_vel_AB[36];       /* 6x6: sum g J_vA^T J_vB (row: vel A) */
  kdh_stats stats;
} kdh_dual_form;

/* The dual forms of n_factors factors in one tile launch and one finish
 * launch (the per-factor buildQuadraticFormDual calls of a bundle adjustment,
 * batched).  The cache is rounded from the FP64 states as kdh_build_at does;
 * params.recenter as there.  stats_mode: inlier terms weigh zero and outlier
 * terms keep their kernel weight, in every block (the dual weight rule,
 * cticp/cticp_impl_.h:240-248).  A factor whose moving tree has no leaves
 * gets an all-zero form. */
int kdh_batch_build_dual(kdh_ctx* ctx, int n_factors, const kdh_dual_pair* factors, int stats_mode,
                         kdh_dual_form* out);

/* ---- serving: continuous batching ----------------------------------------
 * An engine thread runs steps of Gauss-Newton iterations over every in-flight
 * alignment (per iteration one tile + one finish launch per aligner kind;
 * four iterations per step while no request waits, one while one does),
 * retires finished alignments and admits queued requests between steps while
 * slots and the HBM budget allow.  Results are bit-identical to running the request
 * alone.  Maps are uploaded by the caller (kdh_tree_upload on a context of the
 * same device) and declared resident; moving trees travel with the request. */
typedef struct kdh_request {
  const kdh_tree* fixed;            /* resident map */
  const kdh_node3f* moving_nodes;   /* host moving tree (Node_ array), copied at submit */
  int64_t n_moving_nodes;
  const uint64_t* moving_leaves;
  int64_t n_moving_leaves;
  int32_t ct;                       /* 0 rigid ICP, 1 CT-ICP */
  int32_t pad;
  double X0[12];
  double v0[6];
  double Xf[12];
  kdh_params params;
  const kdh_prior* prior;           /* may be NULL; copied at submit */
} kdh_request;

typedef struct kdh_request_result {
  kdh_pair_result r;
  double latency_us;                /* submit -> result visible on the host */
  double queue_us;                  /* submit -> admission */
  int32_t status;                   /* KDH_OK, KDH_ERR_CANCELLED or an error */
  int32_t pad;
} kdh_request_result;

typedef struct kdh_server_info {
  int64_t bytes_budget;
  int64_t bytes_resident;
  int64_t bytes_inflight;
  int32_t inflight;
  int32_t queued;
  int64_t steps;                    /* engine steps (one to four iterations each) */
  int64_t completed;
} kdh_server_info;

int kdh_server_create(int device, int64_t budget_bytes, int32_t max_inflight, kdh_server** out);
/* Stops the engine, fails every unfinished request with KDH_ERR_CANCELLED,
 * wakes the threads blocked in kdh_server_wait and returns once they have
 * left; no other call on the server may start after destroy has begun. */
int kdh_server_destroy(kdh_server* server);
const char* kdh_server_last_error(kdh_server* server);
/* Counts a resident map against the budget (KDH_ERR_BUDGET if it cannot fit). */
int kdh_server_add_resident(kdh_server* server, const kdh_tree* map);
int kdh_server_remove_resident(kdh_server* server, const kdh_tree* map);
/* Queues a request; KDH_ERR_BUDGET if it can never fit beside the resident maps. */
int kdh_server_submit(kdh_server* server, const kdh_request* request, int64_t* id);
/* Non-blocking: *done = 1 and the result (consumed) once finished. */
int kdh_server_poll(kdh_server* server, int64_t id, int* done, kdh_request_result* out);
/* Blocks until finished (timeout_s <= 0: no timeout); consumes the result. */
int kdh_server_wait(kdh_server* server, int64_t id, double timeout_s, kdh_request_result* out);
/* Cancels a queued or in-flight request (its result status is KDH_ERR_CANCELLED). */
int kdh_server_cancel(kdh_server* server, int64_t id);
int kdh_server_info_get(kdh_server* server, kdh_server_info* info);

/* Timing hook for benchmarks: average duration in microseconds of the tile
 * kernel launches of the last kdh_align / kdh_batch_align call on ctx, from
 * HIP events recorded on the context stream. */
int kdh_last_kernel_time(kdh_ctx* ctx, double* avg_us, int* n_launches);

/* Profiling switch: when on, kdh_align / kdh_batch_align launch without a
 * graph and record HIP events around the tile kernel (run 4 times back to
 * back -- it is idempotent -- to amortise dispatch) and around the finish
 * kernel; kdh_last_kernel_breakdown then reports both averages (us per
 * iteration).  Off by default. */
int kdh_set_profiling(kdh_ctx* ctx, int on);
int kdh_last_kernel_breakdown(kdh_ctx* ctx, double* tile_us, double* finish_us);

/* Tuning knobs (project autotuner): workgroups per launch (0 = one per tile)
 * and LDS-staged top-of-tree nodes (0/256/512/1024/2048). */
int kdh_set_launch_config(kdh_ctx* ctx, int grid, int lds_nodes);
/* Tuning knob (project autotuner): moving leaves per lane of the tile kernel,
 * 1 (256-thread workgroups) or 2 (128 threads, two descents in flight per
 * lane).  A tile is 256 leaves either way; results are the same bits in any
 * batch for one setting (the per-tile FP32 sums differ in order between
 * settings).  The dual form always runs 1. */
int kdh_set_leaves_per_lane(kdh_ctx* ctx, int leaves_per_lane);
/* Tuning knob (project autotuner): moving leaves per tile, 256 (default) or
 * 128 (128-thread workgroups, one leaf per lane; twice the workgroups and
 * partial rows).  No reference counterpart (the reference's CUDA reduction
 * has its own fixed block size, cuda/reduce_impl_.h:10-78).  Decisions are
 * the same bits at either size; H, b and the statistics differ in the FP32
 * summation order within a tile (the per-tile sums), each size the same bits
 * alone or in any batch.  A batch prepared at one size is refused by
 * kdh_batch_run at the other.  128-leaf tiles take one leaf per lane. */
int kdh_set_tile_leaves(kdh_ctx* ctx, int tile_leaves);
/* Tuning knob (project autotuner): kdh_batch_align replays its iterations on
 * 1 stream, or on 2 -- the pairs split into two halves of about equal tile
 * count, the second half's chain started one tile launch after the first's,
 * so that one half's finish stage runs beside the other half's tile launch.
 * Results are the same bits either way. */
int kdh_set_batch_streams(kdh_ctx* ctx, int streams);

#ifdef __cplusplus
}
#endif

#endif /* KDH_H_ */
