// Tree maintenance on the device: the in-place deskew and rigid transform of
// a resident tree, and the compressed-leaf extraction.
//
// Reference counterparts (tree/tree_cuda_impl_.h):
//   extractCompressedLeaves_kernel / ...Vel_kernel   :43-136
//   treeApplyVelocities_kernel, applyVelocitiesInPlace :138-160
//   treeApplyIsometry_kernel, applyIsometryInPlace     :162-173
// Node_::applyVelocities (tree/node_impl_.h:86-92) warps mean and direction
// by R(w dts) and v dts with the node's own (float) _dts_mean,
// GeometryTraits_::applyVelocities (utils/geometry_3d_impl_.h:63-72);
// applyIsometry_ on a node (tree/node_.h:127-137) maps the mean by the
// isometry and the direction by its rotation.  Both act here on the device
// layout of kdh_tree_upload: the 32-byte node records and the 32-byte leaf
// records (copies of the leaf nodes' mean and direction).  All elementwise
// and HBM bound: 32 bytes read and written per record.
#include <hip/hip_runtime.h>

#include "../../include/kdh.h"
#include "kdh_capi_internal.h"
#include "kdh_term.h"

#pragma clang fp contract(off)

using namespace kdh;

namespace {

constexpr int kOpThreads = 256;

struct Vel6 {
  float v[6];
};

struct Iso12 {
  float m[12];  // 3x4 row-major [R | t]
};

__device__ __forceinline__ void warp(V3f& mean, V3f& dir, float dts, const Vel6& V) {
  const V3f t_delta = v3(V.v[0] * dts, V.v[1] * dts, V.v[2] * dts);
  const V3f r_delta = v3(V.v[3] * dts, V.v[4] * dts, V.v[5] * dts);
  const Rodrigues rd = rodrigues_f(r_delta);
  mean = mul(rd.R, mean) + t_delta;
  dir = mul(rd.R, dir);
}

// One thread per node record, then per leaf record.  A node is a leaf when
// both links are -1 (Node_::isLeaf, tree/node_.h:54-63).
// A tree built on the device also holds the full Node_ records
// (kdh_tree_download); `full` (or null) gets the same mean and direction as
// the compact record, so both stay identical after an in-place op.
__device__ __forceinline__ void store_full(kdh_node3f* full, int i, V3f mean, V3f dir) {
  if (!full) return;
  kdh_node3f& n = full[i];
  n.mean[0] = mean.x;
  n.mean[1] = mean.y;
  n.mean[2] = mean.z;
  n.direction[0] = dir.x;
  n.direction[1] = dir.y;
  n.direction[2] = dir.z;
}

// root_out (pinned host memory mapped for the device): the root record's new
// mean, the moving tree's re-centring anchor, read by the host after the
// stream synchronisation with no copy of its own.
__device__ __forceinline__ void store_root(float* root_out, V3f mean) {
  root_out[0] = mean.x;
  root_out[1] = mean.y;
  root_out[2] = mean.z;
}

__global__ __launch_bounds__(kOpThreads) void apply_velocities_kernel(float4* nodes, float4* leaves,
                                                                     const float* node_dts, kdh_node3f* full,
                                                                     int n_nodes, int n_leaves, Vel6 V,
                                                                     int skip_leaves, float* root_out) {
  const int i = blockIdx.x * kOpThreads + threadIdx.x;
  if (i < n_nodes) {
    float4 a = nodes[2 * i], b = nodes[2 * i + 1];
    const bool leaf = __float_as_int(b.z) < 0 && __float_as_int(b.w) < 0;
    V3f mean = v3(a.x, a.y, a.z), dir = v3(a.w, b.x, b.y);
    if (!(leaf && skip_leaves)) {
      warp(mean, dir, node_dts[i], V);
      nodes[2 * i] = make_float4(mean.x, mean.y, mean.z, dir.x);
      nodes[2 * i + 1] = make_float4(dir.y, dir.z, b.z, b.w);
      store_full(full, i, mean, dir);
    }
    if (i == 0) store_root(root_out, mean);
  }
  if (i < n_leaves && !skip_leaves) {
    float4 a = leaves[2 * i], b = leaves[2 * i + 1];
    V3f mean = v3(a.x, a.y, a.z), dir = v3(a.w, b.x, b.y);
    warp(mean, dir, b.z, V);  // the leaf record carries its node's dts
    leaves[2 * i] = make_float4(mean.x, mean.y, mean.z, dir.x);
    leaves[2 * i + 1] = make_float4(dir.y, dir.z, b.z, b.w);
  }
}

__global__ __launch_bounds__(kOpThreads) void apply_isometry_kernel(float4* nodes, float4* leaves,
                                                                   kdh_node3f* full, int n_nodes, int n_leaves,
                                                                   Iso12 T, float* root_out) {
  const int i = blockIdx.x * kOpThreads + threadIdx.x;
  const IsoF X = iso_from_34<float>(T.m);
  if (i < n_nodes) {
    float4 a = nodes[2 * i], b = nodes[2 * i + 1];
    const V3f mean = apply_iso(X, v3(a.x, a.y, a.z)), dir = apply_rot(X, v3(a.w, b.x, b.y));
    nodes[2 * i] = make_float4(mean.x, mean.y, mean.z, dir.x);
    nodes[2 * i + 1] = make_float4(dir.y, dir.z, b.z, b.w);
    store_full(full, i, mean, dir);
    if (i == 0) store_root(root_out, mean);
  }
  if (i < n_leaves) {
    float4 a = leaves[2 * i], b = leaves[2 * i + 1];
    const V3f mean = apply_iso(X, v3(a.x, a.y, a.z)), dir = apply_rot(X, v3(a.w, b.x, b.y));
    leaves[2 * i] = make_float4(mean.x, mean.y, mean.z, dir.x);
    leaves[2 * i + 1] = make_float4(dir.y, dir.z, b.z, b.w);
  }
}

// out: per leaf (mean, direction), optionally warped by V with the leaf's dts.
__global__ __launch_bounds__(kOpThreads) void extract_leaves_kernel(float* out, const float4* leaves, int n_leaves,
                                                                   Vel6 V, int warp_leaves) {
  const int i = blockIdx.x * kOpThreads + threadIdx.x;
  if (i >= n_leaves) return;
  const float4 a = leaves[2 * i], b = leaves[2 * i + 1];
  V3f mean = v3(a.x, a.y, a.z), dir = v3(a.w, b.x, b.y);
  if (warp_leaves) warp(mean, dir, b.z, V);
  float* o = out + 6 * (size_t)i;
  o[0] = mean.x;
  o[1] = mean.y;
  o[2] = mean.z;
  o[3] = dir.x;
  o[4] = dir.y;
  o[5] = dir.z;
}

__global__ __launch_bounds__(kOpThreads) void read_nodes_kernel(float* out, const float4* nodes, int n_nodes) {
  const int i = blockIdx.x * kOpThreads + threadIdx.x;
  if (i >= n_nodes) return;
  const float4 a = nodes[2 * i], b = nodes[2 * i + 1];
  float* o = out + 6 * (size_t)i;
  o[0] = a.x;
  o[1] = a.y;
  o[2] = a.z;
  o[3] = a.w;
  o[4] = b.x;
  o[5] = b.y;
}

int blocks_for(int n) { return (n + kOpThreads - 1) / kOpThreads; }

// The pinned root buffer of the context, made on first use.
int ensure_root_buffer(kdh_ctx* ctx) {
  if (ctx->h_root) return KDH_OK;
  hipError_t e = hipHostMalloc((void**)&ctx->h_root, 4 * sizeof(float), hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&ctx->h_root_dev, ctx->h_root, 0);
  if (e != hipSuccess) {
    if (ctx->h_root) hipHostFree(ctx->h_root);
    ctx->h_root = ctx->h_root_dev = nullptr;
    return kdh_hip_fail(ctx, e, "pinned root buffer");
  }
  return KDH_OK;
}

// The call returns once the op has run: the root record's mean, which the
// kernel wrote into the pinned buffer, is the re-centring anchor of a moving
// tree.
int finish_op(kdh_ctx* ctx, kdh_tree* t) {
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "tree op");
  if (t->dt.n_nodes > 0) t->root_mean = v3(ctx->h_root[0], ctx->h_root[1], ctx->h_root[2]);
  return KDH_OK;
}

}  // namespace

int kdh_tree_apply_velocities(kdh_ctx* ctx, kdh_tree* tree, const float v[6], int skip_leaves) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!tree || !v) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_apply_velocities: null argument");
  if (tree->device != ctx->device) return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  const int n = std::max(tree->dt.n_nodes, tree->dt.n_leaves);
  if (n == 0) return KDH_OK;
  int rc = ensure_root_buffer(ctx);
  if (rc) return rc;
  Vel6 V;
  for (int i = 0; i < 6; ++i) V.v[i] = v[i];
  hipLaunchKernelGGL(apply_velocities_kernel, dim3(blocks_for(n)), dim3(kOpThreads), 0, ctx->stream,
                     (float4*)tree->dt.nodes, (float4*)tree->dt.leaves, (const float*)tree->node_dts,
                     tree->full_nodes, tree->dt.n_nodes, tree->dt.n_leaves, V, skip_leaves ? 1 : 0,
                     ctx->h_root_dev);
  e = hipGetLastError();
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "apply_velocities_kernel");
  return finish_op(ctx, tree);
}

int kdh_tree_apply_isometry(kdh_ctx* ctx, kdh_tree* tree, const float iso[12]) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!tree || !iso) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_apply_isometry: null argument");
  if (tree->device != ctx->device) return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  const int n = std::max(tree->dt.n_nodes, tree->dt.n_leaves);
  if (n == 0) return KDH_OK;
  int rc = ensure_root_buffer(ctx);
  if (rc) return rc;
  Iso12 T;
  for (int i = 0; i < 12; ++i) T.m[i] = iso[i];
  hipLaunchKernelGGL(apply_isometry_kernel, dim3(blocks_for(n)), dim3(kOpThreads), 0, ctx->stream,
                     (float4*)tree->dt.nodes, (float4*)tree->dt.leaves, tree->full_nodes, tree->dt.n_nodes,
                     tree->dt.n_leaves, T, ctx->h_root_dev);
  e = hipGetLastError();
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "apply_isometry_kernel");
  return finish_op(ctx, tree);
}

static int ensure_staging(kdh_ctx* ctx, size_t n_floats, float** dbuf, size_t* cap, const char* what) {
  if (n_floats > *cap) {
    if (*dbuf) hipFree(*dbuf);
    *dbuf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc((void**)dbuf, n_floats * sizeof(float));
    if (e != hipSuccess) return kdh_fail(ctx, KDH_ERR_NOMEM, std::string(what) + ": hipMalloc");
    *cap = n_floats;
  }
  return KDH_OK;
}

int kdh_tree_extract_leaves(kdh_ctx* ctx, const kdh_tree* tree, const float* v, float* out) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!tree || !out) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_extract_leaves: null argument");
  if (tree->device != ctx->device) return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  const int n = tree->dt.n_leaves;
  if (n == 0) return KDH_OK;
  int rc = ensure_staging(ctx, (size_t)n * 6, &ctx->d_extract, &ctx->extract_cap, "kdh_tree_extract_leaves");
  if (rc) return rc;
  Vel6 V;
  for (int i = 0; i < 6; ++i) V.v[i] = v ? v[i] : 0.0f;
  hipLaunchKernelGGL(extract_leaves_kernel, dim3(blocks_for(n)), dim3(kOpThreads), 0, ctx->stream, ctx->d_extract,
                     tree->dt.leaves, n, V, v ? 1 : 0);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, ctx->d_extract, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_tree_extract_leaves");
  return KDH_OK;
}

int kdh_tree_read_nodes(kdh_ctx* ctx, const kdh_tree* tree, float* out) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!tree || !out) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_read_nodes: null argument");
  if (tree->device != ctx->device) return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  const int n = tree->dt.n_nodes;
  if (n == 0) return KDH_OK;
  int rc = ensure_staging(ctx, (size_t)n * 6, &ctx->d_extract, &ctx->extract_cap, "kdh_tree_read_nodes");
  if (rc) return rc;
  hipLaunchKernelGGL(read_nodes_kernel, dim3(blocks_for(n)), dim3(kOpThreads), 0, ctx->stream, ctx->d_extract,
                     tree->dt.nodes, n);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, ctx->d_extract, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_tree_read_nodes");
  return KDH_OK;
}

// ---- backup and restore of a resident tree ----------------------------------
// TreeCUDA_::clone / Tree_CUDAfromCUDA_ (tree/tree_cuda_impl_.h:176-195,
// tree_cuda_.h:57) and copyNodes (tree_cuda_impl_.h:60-69, tree_cuda_.h:43):
// the bundle-adjustment pre-deskew keeps a clone of every keyframe tree and
// restores the nodes from it before each re-deskew
// (optimizer/optimizer_proc_solver_action_.h:39-58, map/map_.h:88-102).
// A tree is one device allocation here, so a clone is one device-to-device
// copy of it with every section pointer moved by the same offset.

static kdh_tree* rebased_copy(const kdh_tree* src, void* d) {
  kdh_tree* t = new kdh_tree(*src);
  const char* o = (const char*)src->dmem;
  char* n = (char*)d;
  auto mv = [&](const void* p) -> char* { return p ? n + ((const char*)p - o) : nullptr; };
  t->dmem = d;
  t->dt.nodes = (const float4*)mv(src->dt.nodes);
  t->dt.leaves = (const float4*)mv(src->dt.leaves);
  t->node_dts = (float*)mv(src->node_dts);
  t->dt.node_dts = t->node_dts;
  t->full_nodes = (kdh_node3f*)mv(src->full_nodes);
  t->full_leaves = (uint64_t*)mv(src->full_leaves);
  t->full_xyz = (float*)mv(src->full_xyz);
  t->full_stamps = (double*)mv(src->full_stamps);
  return t;
}

int kdh_tree_clone(kdh_ctx* ctx, const kdh_tree* src, kdh_tree** out) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!src || !out) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_clone: null argument");
  if (src->device != ctx->device) return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  void* d = nullptr;
  e = hipMalloc(&d, src->bytes ? src->bytes : 1);
  if (e != hipSuccess) return kdh_fail(ctx, KDH_ERR_NOMEM, "kdh_tree_clone: hipMalloc");
  if (src->bytes) e = hipMemcpyAsync(d, src->dmem, src->bytes, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    hipFree(d);
    return kdh_hip_fail(ctx, e, "kdh_tree_clone: copy");
  }
  *out = rebased_copy(src, d);
  return KDH_OK;
}

int kdh_tree_copy_nodes(kdh_ctx* ctx, const kdh_tree* src, kdh_tree* dst) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!src || !dst) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_copy_nodes: null argument");
  if (src->device != ctx->device || dst->device != ctx->device)
    return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  if (src->dt.n_nodes != dst->dt.n_nodes || src->dt.n_leaves != dst->dt.n_leaves)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_copy_nodes: size mismatch");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  const size_t nb = (size_t)src->dt.n_nodes * 32, lb = (size_t)src->dt.n_leaves * 32;
  // node and leaf records (the leaf records repeat the leaf nodes' mean and
  // direction), then the full records when both trees carry them
  if (nb) e = hipMemcpyAsync((void*)dst->dt.nodes, src->dt.nodes, nb, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && lb)
    e = hipMemcpyAsync((void*)dst->dt.leaves, src->dt.leaves, lb, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && nb && src->node_dts && dst->node_dts)
    e = hipMemcpyAsync(dst->node_dts, src->node_dts, (size_t)src->dt.n_nodes * sizeof(float),
                       hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && nb && src->full_nodes && dst->full_nodes)
    e = hipMemcpyAsync(dst->full_nodes, src->full_nodes, (size_t)src->dt.n_nodes * sizeof(kdh_node3f),
                       hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_tree_copy_nodes");
  dst->root_mean = src->root_mean;
  return KDH_OK;
}
