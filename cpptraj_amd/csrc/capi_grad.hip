// capi_grad.hip -- C ABI of the gradients (rmsd_fit_backward,
// rmsd_pairwise_backward): rotations recomputed in FP64, envelope-theorem
// gradients (k_grad.hip).
#include "capi_internal.hpp"

using namespace rmsd_capi;

extern "C" {

int rmsd_fit_backward(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                      int ref_frame, int use_mass, const double* grad_rmsd, double* grad_xyz) {
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
  if (frames->nframes == 0) return RMSD_OK;
  if (!grad_rmsd || !grad_xyz) return fail(ctx, RMSD_ERR_ARG, "gradient buffers are NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int n = frames->nsel;
  const int ldr = (int)round_up(n, 4);
  const int64_t nf = frames->nframes;
  double* refbuf = (double*)scratch(ctx, "fit.ref", sizeof(double) * (3 * ldr + 8), &st);
  if (!refbuf) return st;
  double* ref_trans = refbuf + 3 * ldr;
  double* consts = (double*)scratch(ctx, "fit.consts", sizeof(double) * 8, &st);
  if (!consts) return st;
  double* wpad = nullptr;
  if (use_mass) {
    wpad = (double*)scratch(ctx, "fit.mass", sizeof(double) * ldr, &st);
    if (!wpad) return st;
    HIP_TRY(ctx, hipMemsetAsync(wpad, 0, sizeof(double) * ldr, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(wpad, frames->mass, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                ctx->stream));
  }
  double* moments = (double*)scratch(ctx, "fit.moments", sizeof(double) * 16 * nf, &st);
  if (!moments) return st;
  double* xform = (double*)scratch(ctx, "fit.xform", sizeof(double) * 15 * nf, &st);
  if (!xform) return st;
  double* rm = (double*)scratch(ctx, "fit.rmsd", sizeof(double) * nf, &st);
  if (!rm) return st;
  const FrameSetDev F = to_dev(*frames);
  const FrameSetDev R = to_dev(*ref);
  // Recompute the per-frame rotations (moments + QCP) rather than keeping
  // them from the forward call.
  HIP_TRY(ctx, rmsd_launch_ref_prep(R, ref_frame, use_mass ? ref->mass : nullptr, wpad, 1, refbuf,
                                    ref_trans, consts, ctx->stream));
  FitArgs fa;
  fa.fs = F;
  fa.mass = wpad;
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
  sa.rmsd = rm;
  sa.U = nullptr;
  sa.tvec = nullptr;
  sa.xform = xform;
  HIP_TRY(ctx, rmsd_launch_fit_solve(sa, ctx->stream));
  HIP_TRY(ctx, rmsd_launch_fit_grad(F, wpad, refbuf, xform, rm, grad_rmsd, consts, grad_xyz,
                                    ctx->stream));
  return RMSD_OK;
}

int rmsd_pairwise_backward(rmsd_ctx* ctx, const rmsd_frames* A, const rmsd_frames* B,
                           int use_mass, const double* grad_out, double* grad_a, double* grad_b) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, A, "A")) != RMSD_OK) return st;
  if ((st = check_frames(ctx, B, "B")) != RMSD_OK) return st;
  if (A->nsel != B->nsel)
    return fail(ctx, RMSD_ERR_ARG, "the two frame sets select different atom counts");
  if (use_mass && !A->mass) return fail(ctx, RMSD_ERR_ARG, "use_mass needs masses");
  if (!grad_out) return fail(ctx, RMSD_ERR_ARG, "grad_out is NULL");
  const int FA = A->nframes, FB = B->nframes, n = A->nsel;
  if (FA == 0 || FB == 0) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const double* w = use_mass ? A->mass : nullptr;
  double* XA = (double*)scratch(ctx, "grad.XA", sizeof(double) * 3 * (size_t)n * FA, &st);
  if (!XA) return st;
  double* XB = (double*)scratch(ctx, "grad.XB", sizeof(double) * 3 * (size_t)n * FB, &st);
  if (!XB) return st;
  double* gA = (double*)scratch(ctx, "grad.gA", sizeof(double) * FA, &st);
  if (!gA) return st;
  double* gB = (double*)scratch(ctx, "grad.gB", sizeof(double) * FB, &st);
  if (!gB) return st;
  double* cw = (double*)scratch(ctx, "grad.cw", sizeof(double) * 10 * (size_t)FA * FB, &st);
  if (!cw) return st;
  double* consts = (double*)scratch(ctx, "pair.consts", sizeof(double) * 16, &st);
  if (!consts) return st;
  SetupArgs s{};
  s.fs = to_dev(*A);
  s.smass = w;
  s.total_mass = consts;
  HIP_TRY(ctx, rmsd_launch_setup(s, ctx->stream));
  HIP_TRY(ctx, rmsd_launch_center(to_dev(*A), w, XA, gA, ctx->stream));
  HIP_TRY(ctx, rmsd_launch_center(to_dev(*B), w, XB, gB, ctx->stream));
  HIP_TRY(ctx, rmsd_launch_pair_grad(XA, XB, gA, gB, w, n, FA, FB, consts, grad_out, cw, grad_a,
                                     grad_b, ctx->stream));
  return RMSD_OK;
}

}  // extern "C"
