// kdh_align_noalloc: the device-resident Gauss-Newton loop of kdh_align with
// no allocation at all -- for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It has no counterpart in the reference.
//
// The loop's device memory beyond the caller's trees -- the pair state and
// the per-tile partial rows -- lives in `scratch`, a file-scope device array
// of fixed size: every GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in symbol); it takes the array's device address
// with hipGetSymbolAddress once per device, on that device, at the entry
// point's first call there, and uses that address from then on.
//
// Calls on one device share its array and its last contents: every call
// rewrites the pair state and every partial row it reads, so nothing has to
// start at zero.  Calls that would overlap on one device are serialised by a
// per-device lock held from the state upload to the result read-back.
// Launches are plain (no graph: instantiating one allocates).
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <string>

#include "../../include/kdh.h"
#include "kdh_capi_internal.h"
#include "kdh_kernels.h"

using namespace kdh;

namespace {

constexpr int kNoAllocMaxTiles = 4096;  // 1,048,576 moving leaves
constexpr size_t kStateOff = 0;
constexpr size_t kPartialsOff = (sizeof(PairState) + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kPartialsOff + sizeof(float) * (size_t)kNoAllocMaxTiles * kCtSlots;
constexpr int kMaxDevices = 64;

}  // namespace

// The loop's scratch memory: one PairState, then kNoAllocMaxTiles partial rows
// of up to kCtSlots floats (1.6 MB in all).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_table_mu;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_device_mu[kMaxDevices];

// The device address of this device's `scratch` (the device is current).
int scratch_address(kdh_ctx* ctx, void** out) {
  std::lock_guard<std::mutex> lk(g_table_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return kdh_hip_fail(ctx, e, "kdh_align_noalloc: hipGetSymbolAddress(scratch)");
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return KDH_OK;
}

}  // namespace

int kdh_align_noalloc(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct, const double X0[12],
                      const double v0[6], const double Xf[12], const kdh_params* params, const kdh_prior* prior,
                      int stats_mode, double X_out[12], double v_out[6], kdh_stats* stats_log, int* n_log,
                      int* result) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!X0 || !Xf || !params) return kdh_fail(ctx, KDH_ERR_INVALID, "null argument");
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_align_noalloc: device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "hipSetDevice");
  void* base = nullptr;
  int rc = scratch_address(ctx, &base);
  if (rc) return rc;
  PairState ps;
  rc = kdh_fill_pair(ctx, ps, fixed, moving, ct, *params, stats_mode, 0);
  if (rc) return rc;
  const int nt = ps.n_tiles;
  if (nt > kNoAllocMaxTiles)
    return kdh_fail(ctx, KDH_ERR_INVALID,
                    "kdh_align_noalloc: " + std::to_string(moving->dt.n_leaves) + " moving leaves exceed the "
                    "scratch maximum of " + std::to_string(kNoAllocMaxTiles * ctx->tile_leaves) +
                    " at this tile size (use kdh_align)");
  memcpy(ps.X, X0, sizeof(ps.X));
  memcpy(ps.Xf, Xf, sizeof(ps.Xf));
  for (int i = 0; i < 6; ++i) ps.v[i] = (ct && v0) ? v0[i] : 0.0;
  kdh_setup_prior(ps, prior);
  kdh_cache_from_state(ps);
  PairState* d_state = reinterpret_cast<PairState*>(static_cast<unsigned char*>(base) + kStateOff);
  float* d_partials = reinterpret_cast<float*>(static_cast<unsigned char*>(base) + kPartialsOff);
  std::lock_guard<std::mutex> dk(g_device_mu[ctx->device]);
  if (nt == 0) {
    // No moving leaves: every round fails the min_good gate, as the reference does.
    memset(&ps.stats_log[0], 0, sizeof(DevStats));
    ps.iter = 1;
    ps.status = params->min_good_measurements > 0 ? 2 : 1;
  } else {
    e = hipMemcpyAsync(d_state, &ps, kStateBytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_align_noalloc: state upload");
    const int iters = ps.sp.max_iterations;
    e = hipEventRecord(ctx->ev[0], ctx->stream);
    for (int it = 0; it < iters && e == hipSuccess; ++it) {
      rc = kdh_launch(ctx, d_state, nullptr, nt, ct, nullptr, nullptr, d_partials, 1, 3);
      if (rc) return rc;
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&ps, d_state, sizeof(PairState), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_align_noalloc: device loop");
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->last_kernel_us = iters ? 1000.0 * ms / iters : 0.0;
    ctx->last_tile_us = ctx->last_finish_us = 0.0;
    ctx->last_launches = iters;
  }
  if (X_out) memcpy(X_out, ps.X, sizeof(ps.X));
  if (v_out) memcpy(v_out, ps.v, sizeof(ps.v));
  const int n = ps.iter;
  if (stats_log) memcpy(stats_log, ps.stats_log, sizeof(DevStats) * n);
  if (n_log) *n_log = n;
  if (result) *result = kdh_result_code(ps, params->max_iterations);
  return KDH_OK;
}

int kdh_align_noalloc_max_leaves(void) { return kNoAllocMaxTiles * kTile; }
