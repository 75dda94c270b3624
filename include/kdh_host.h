/*
 * kdh_host.h -- host-side input preparation of libkdhip.so: the PCA KD-tree
 * builder whose output both the device path and the CPU aligners consume.
 *
 * Replaces TreeGenerator_<NodePoint3f>::makeTree
 * (kd_slam2/src/lib/kd_slam/tree/tree_generator_impl_.h:9-19) with its
 * parameters (tree_generator_.h:18-25).  Callers that already hold a
 * reference TreeCPU_ pass its arrays to kdh_tree_upload (kdh.h) directly and
 * do not need this builder.
 */
#ifndef KDH_HOST_H_
#define KDH_HOST_H_

#include <stdint.h>

#include "kdh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TreeGenerator_ parameters (tree_generator_.h:18-25) plus the reference's
 * FlatLeafPolicy_::parallel_level (label_policy_base_.h:48). */
typedef struct kdh_tree_gen_params {
  int32_t propagate_normal;          /* false */
  int32_t tree_min_good_leaves;      /* 10 */
  int32_t leaf_min_points;           /* 5 */
  int32_t leaf_max_points;           /* 30 */
  float leaf_density;                /* 100 */
  float leaf_min_normal_incidence;   /* 0.1 */
  float leaf_eigenvalue_threshold;   /* 1e-3 (unused by the reference policy too) */
  float leaf_normal_eigenratio;      /* 0.1 */
  int32_t parallel_level;            /* 0: single thread; k: split the top k levels over threads */
  int32_t label_mode;                /* 0: the canonical build, FP64 moments in a fixed order, Jacobi
                                        eigen-solver, order-keeping split (host = device, bit for bit);
                                        1: the reference's own arithmetic (label_policy_base_impl_.h:41-95,
                                        node_impl_.h:12-26): FP32 StableAdder_ moments, E[xx^T] - mu mu^T,
                                        SelfAdjointEigenSolver::computeDirect, the two-ended splitRange.
                                        Host builder only. */
} kdh_tree_gen_params;

typedef struct kdh_host_tree kdh_host_tree;

void kdh_tree_gen_params_default(kdh_tree_gen_params* p);

/* Builds, prunes and BFS-orders the tree of n_points points (xyz: 3 floats
 * each; stamps: one double each, or NULL for unstamped points).  *out is set
 * to NULL (and KDH_OK returned) where makeTree returns nullptr: fewer than
 * leaf_min_points points or fewer than tree_min_good_leaves good leaves. */
int kdh_host_tree_build(const float* xyz, const double* stamps, int64_t n_points,
                        const kdh_tree_gen_params* params, kdh_host_tree** out);
int kdh_host_tree_info(const kdh_host_tree* t, int64_t* n_nodes, int64_t* n_leaves, double* ts_start);
/* Node array (Node_<Point3fTraits> layout) and leaf-index list, valid until
 * kdh_host_tree_free; ready for kdh_tree_upload. */
const kdh_node3f* kdh_host_tree_nodes(const kdh_host_tree* t);
const uint64_t* kdh_host_tree_leaves(const kdh_host_tree* t);
/* The points in the builder's permuted order (3 floats each). */
const float* kdh_host_tree_points(const kdh_host_tree* t);
int kdh_host_tree_free(kdh_host_tree* t);

/* ---- tree generation on the device ----------------------------------------
 * Replaces TreeGenerator_::makeTree (tree/tree_generator_impl_.h:9-19) plus
 * the device upload (Tree_CUDAfromCPU_, tree/tree_cuda_impl_.h:198-222): the
 * tree is built, pruned and breadth-first ordered on the GPU and stays there,
 * bit-identical to kdh_host_tree_build's (same moments order, eigen-solver,
 * labelling and split rule).  xyz: 3 floats per point; stamps: one double per
 * point or NULL; both host pointers, or device pointers of this context's
 * device when inputs_on_device.  *out is NULL (and KDH_OK returned) where
 * makeTree returns nullptr.  params.parallel_level is ignored.  Free with
 * kdh_tree_free. */
int kdh_tree_build(kdh_ctx* ctx, const float* xyz, const double* stamps, int64_t n_points,
                   const kdh_tree_gen_params* params, int inputs_on_device, kdh_tree** out);

/* A device-built tree's full Node_ records (n_nodes), leaf list (n_leaves),
 * permuted points (3 floats each) and stamps (kdh_tree_info gives the
 * counts); any output may be NULL.  TreeCPU_'s nodes / leaves / points
 * (tree/tree_cpu_.h).  KDH_ERR_INVALID for a tree made by kdh_tree_upload. */
int kdh_tree_download(kdh_ctx* ctx, const kdh_tree* tree, kdh_node3f* nodes, uint64_t* leaves, float* xyz,
                      double* stamps);

#ifdef __cplusplus
}
#endif

#endif /* KDH_HOST_H_ */
