// capi_noalloc.hip -- rmsd_fit_to_ref_noalloc: the streaming fit of
// rmsd_fit_to_ref for callers that cannot allocate device memory (code under
// HIP graph capture, tight per-frame loops, harnesses that forbid hipMalloc).
//
// The fit's working memory (the prepared reference, its constants, the padded
// target masses, 16 moments and 15 transform values per frame) lives in one
// file-scope device array per GPU, `scratch`, of RMSD_NOALLOC_SCRATCH_BYTES.
// rmsd_ctx_create looks up its address once per device (hipGetSymbolAddress
// with that device current) and host code only ever passes that address to
// kernels.  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU (two streams, two threads) race.  Nothing here calls hipMalloc,
// hipFree, hipMemset or hipMemcpy; the kernels are the ones rmsd_fit_to_ref
// launches, so the results are the same bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "capi_internal.hpp"

using namespace rmsd_capi;

constexpr size_t NOALLOC_DOUBLES = RMSD_NOALLOC_SCRATCH_BYTES / sizeof(double);
// Host code names it ::scratch: inside namespace rmsd_capi the bare name is the
// context's scratch-buffer function.
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

// Padded target masses: w[k] = mass[k] for k < n, 0 up to ldr (the fit
// kernels read whole quads).  A kernel instead of a memset + copy, so the
// sequence is also valid under stream capture.
__global__ void k_pad_mass(double* w, const double* mass, int n, int ldr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ldr) w[k] = k < n ? mass[k] : 0.0;
}

// Layout of one call in the scratch array, each piece starting on a 256-byte
// boundary as the context's own scratch buffers do: reference SoA + its
// translation (3 ldr + 8), constants (8), padded masses (ldr), moments (16 per
// frame), transforms (15 per frame, coordinate update only).
struct Layout {
  int64_t ref, consts, wpad, moments, xform, end;  // offsets in doubles
};

int64_t up32(int64_t d) { return (d + 31) & ~int64_t(31); }

Layout layout(int ldr, int64_t nf, int fit, int modify) {
  Layout l;
  l.ref = 0;
  l.consts = up32(3 * (int64_t)ldr + 8);
  l.wpad = l.consts + 32;
  l.moments = up32(l.wpad + ldr);
  l.xform = up32(l.moments + (fit ? 16 * nf : 0));
  l.end = l.xform + (fit && modify != RMSD_MODIFY_NONE ? 15 * nf : 0);
  return l;
}

}  // namespace

namespace rmsd_capi {

// Looks up this translation unit's `scratch` on the current device (called by
// rmsd_ctx_create with `device` current); returns the per-device address.
int noalloc_init_device(int device, void** ptr, std::string* err) {
  if (device < 0 || device >= MAX_DEVICES) {
    *err = "device index beyond the allocation-free table";
    return RMSD_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return RMSD_ERR_DEVICE;
    }
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return RMSD_OK;
}

}  // namespace rmsd_capi

extern "C" {

int rmsd_fit_to_ref_noalloc(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                            int ref_frame, int fit, int use_mass, int modify, double* rmsd,
                            double* U, double* tvec) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, frames, "frames")) != RMSD_OK) return st;
  if ((st = check_frames(ctx, ref, "ref")) != RMSD_OK) return st;
  if (ref->nsel != frames->nsel)
    return fail(ctx, RMSD_ERR_ARG, "reference and target select different atom counts");
  if (ref_frame < 0 || ref_frame >= ref->nframes)
    return fail(ctx, RMSD_ERR_ARG, "reference frame out of range");
  if (use_mass && (!frames->mass || !ref->mass))
    return fail(ctx, RMSD_ERR_ARG, "use_mass needs target and reference masses");
  if (modify < RMSD_MODIFY_NONE || modify > RMSD_MODIFY_TRANS)
    return fail(ctx, RMSD_ERR_ARG, "unknown modify mode");
  if (modify != RMSD_MODIFY_NONE && frames->fidx)
    return fail(ctx, RMSD_ERR_ARG, "coordinate update needs frames in storage order (fidx NULL)");
  if (!rmsd && frames->nframes > 0) return fail(ctx, RMSD_ERR_ARG, "rmsd is NULL");
  if (frames->nframes == 0) return RMSD_OK;
  if (!ctx->noalloc)
    return fail(ctx, RMSD_ERR_DEVICE, "allocation-free scratch not set up on this device");
  const int n = frames->nsel;
  const int ldr = (int)round_up(n, 4);
  const int64_t nf = frames->nframes;
  const Layout lay = layout(ldr, nf, fit, modify);
  const int64_t need = lay.end;
  if (need > (int64_t)NOALLOC_DOUBLES)
    return fail(ctx, RMSD_ERR_ARG,
                "rmsd_fit_to_ref_noalloc: " + std::to_string(nf) + " frames of " +
                    std::to_string(n) + " atoms need " + std::to_string(need * 8) +
                    " bytes of scratch, more than RMSD_NOALLOC_SCRATCH_BYTES (" +
                    std::to_string((long long)RMSD_NOALLOC_SCRATCH_BYTES) +
                    "); split the frames into smaller calls");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  double* base = static_cast<double*>(ctx->noalloc);
  double* refbuf = base + lay.ref;
  double* ref_trans = refbuf + 3 * ldr;
  double* consts = base + lay.consts;
  double* wpad = base + lay.wpad;
  double* moments = base + lay.moments;
  double* xform = base + lay.xform;
  const FrameSetDev F = to_dev(*frames);
  const FrameSetDev R = to_dev(*ref);
  if (use_mass) {
    hipLaunchKernelGGL(k_pad_mass, dim3((ldr + 255) / 256), dim3(256), 0, ctx->stream, wpad,
                       frames->mass, n, ldr);
    HIP_TRY(ctx, hipGetLastError());
  }
  HIP_TRY(ctx, rmsd_launch_ref_prep(R, ref_frame, use_mass ? ref->mass : nullptr,
                                    use_mass ? wpad : nullptr, fit ? 1 : 0, refbuf, ref_trans,
                                    consts, ctx->stream));
  if (!fit) {
    NofitArgs a;
    a.fs = F;
    a.mass = use_mass ? wpad : nullptr;
    a.ref = refbuf;
    HIP_TRY(ctx, rmsd_launch_nofit_frames(a, consts, rmsd, ctx->stream));
    return RMSD_OK;
  }
  FitArgs fa;
  fa.fs = F;
  fa.mass = use_mass ? wpad : nullptr;
  fa.ref = refbuf;
  fa.consts = consts;
  fa.moments = moments;
  HIP_TRY(ctx, rmsd_launch_fit_moments(fa, ctx->stream));
  SolveArgs sa;
  sa.moments = moments;
  sa.consts = consts;
  sa.ref_trans = ref_trans;
  sa.nframes = (int)nf;
  sa.fit = 1;
  sa.clamp = 1;
  sa.rmsd = rmsd;
  sa.U = U;
  sa.tvec = tvec;
  sa.xform = modify != RMSD_MODIFY_NONE ? xform : nullptr;
  HIP_TRY(ctx, rmsd_launch_fit_solve(sa, ctx->stream));
  if (modify != RMSD_MODIFY_NONE) {
    ApplyArgs ap;
    ap.xyz = const_cast<float*>(frames->xyz);
    ap.stride = frames->stride;
    ap.ld = frames->ld;
    ap.natom = frames->natom;
    ap.nframes = frames->nframes;
    ap.xform = xform;
    HIP_TRY(ctx, rmsd_launch_apply(ap, modify == RMSD_MODIFY_ROT_TRANS ? 1 : 2, ctx->stream));
  }
  return RMSD_OK;
}

}  // extern "C"
