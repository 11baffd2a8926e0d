// capi_internal.hpp -- state and helpers shared by the C ABI translation
// units (capi.hip: contexts, matrices, fit; capi_grad.hip: gradients;
// capi_batch.hip: batched plans).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <set>
#include <string>
#include <vector>

#include "cpptraj_rmsd.h"
#include "rmsd_kernels.hpp"

namespace rmsd_capi {

constexpr int TILE = 64;      // frames per tile side (all pairwise kernels)
constexpr int GROUP = 8;      // tile rows per band of the grouped schedule
constexpr int NXCD = 8;       // MI355X: 8 XCDs, workgroups dealt round-robin
// Frames averaged into the mean-structure template of the MFMA-reduced forms.
constexpr int TEMPLATE_SAMPLES = 256;
// Standard deviations covered by the probabilistic terms of the error bound.
constexpr double BOUND_LAMBDA = 6.0;
constexpr double DEFAULT_TOLERANCE = 1e-4;  // Angstrom

// What rmsd_ctx_error_report can say about the last pairwise call:
// kind 0 = FP64 (or no-fit) arithmetic, 1 = bounded mixed precision (stats:
// {largest kept bound as double bits, pairs recomputed in FP64}), 2 = mixed
// precision with the bound switched off.
struct ErrorReport {
  int kind = 0;
  unsigned long long* stats = nullptr;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct TileSchedule {
  int nbr = -1, nbc = -1, tri = -1, tc = -1;
  int2* dev = nullptr;
  int ntiles = 0;
};

}  // namespace rmsd_capi

struct rmsd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<std::string, rmsd_capi::DevBuf> bufs;
  rmsd_capi::TileSchedule sched;
  // Stage timing: events recorded on the context's stream around the per-frame
  // stage, the main kernel and the epilogue stage of the last call.
  bool profiling = false;
  int avgcorr_method = 0;  // RMSD_AVGCORR_*
  bool events_made = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int marks = 0;
  // Batch plans bound to this context; freed with it if still alive.
  std::set<rmsd_batch*> plans;
  // Per-residue tables of the last rmsd_perres_nofit call (host image of the
  // device copy in scratch "perres.table"), re-uploaded only when they change.
  std::vector<unsigned char> perres_blob;
  // Error bound of the mixed-precision pairwise forms: tolerance (Angstrom;
  // 0 = off, < 0 = report only) and what the last call reported.
  double tol = rmsd_capi::DEFAULT_TOLERANCE;
  rmsd_capi::ErrorReport report;
  // This device's allocation-free scratch (capi_noalloc.hip), looked up when
  // the context is created.
  void* noalloc = nullptr;
};

namespace rmsd_capi {

// Records an error message on the context and returns `code`.
int fail(rmsd_ctx* c, int code, const std::string& msg);
int hip_fail(rmsd_ctx* c, hipError_t e, const char* where);

#define HIP_TRY(ctx, expr)                                   \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) return rmsd_capi::hip_fail((ctx), e_, #expr); \
  } while (0)

// Stage-timing event k on the context's stream (when profiling is on).
void mark(rmsd_ctx* c, int k);
// Grow-only named scratch buffer of the context.
void* scratch(rmsd_ctx* c, const char* name, size_t bytes, int* status);
int check_frames(rmsd_ctx* c, const rmsd_frames* f, const char* what);
FrameSetDev to_dev(const rmsd_frames& f);
std::vector<int2> grouped_tiles(int nbr, int nbc, int tri, int tr = TILE, int tc = TILE);
int64_t round_up(int64_t x, int64_t m);
// Coefficients of the per-pair error bound for a precision and chunk count.
BoundCoef rmsd_bound_coef(int prec, int KB);
// Host split of the tolerance for keep_limits (tol < 0: report only).
rmsd_dev::KeepSplit rmsd_keep_split(double tol);
// Frees a batch plan (capi_batch.hip); rmsd_ctx_destroy frees leftovers.
void batch_free(rmsd_batch* b);
// Address of the allocation-free scratch on `device` (current), looked up once
// per device (capi_noalloc.hip).
int noalloc_init_device(int device, void** ptr, std::string* err);

}  // namespace rmsd_capi
