// capi.hip -- the C ABI of include/cpptraj_rmsd.h: contexts, pairwise
// matrices, cluster cache, auto-correlation, fit and transform (gradients in
// capi_grad.hip, batched plans in capi_batch.hip).
//
// Host-side orchestration only: argument checks, scratch management, tile
// schedules and the launch sequence of the kernels in k_*.hip.  Every entry
// point enqueues work on the context's stream and returns without waiting
// (except rmsd_ctx_synchronize).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "capi_internal.hpp"

namespace rmsd_capi {

void mark(rmsd_ctx* c, int k) {
  if (!c->profiling || !c->events_made) return;
  (void)hipEventRecord(c->ev[k], c->stream);
  c->marks |= 1 << k;
}

int fail(rmsd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(rmsd_ctx* c, hipError_t e, const char* where) {
  return fail(c, RMSD_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}


// Grow-only named scratch buffer.  A buffer that has to grow is released only
// after the stream has drained, since queued kernels may still read it.
void* scratch(rmsd_ctx* c, const char* name, size_t bytes, int* status) {
  DevBuf& b = c->bufs[name];
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return b.p;
  if (b.p) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  const size_t rounded = (bytes + 255) & ~size_t(255);
  if (hipMalloc(&b.p, rounded) != hipSuccess) {
    b.p = nullptr;
    *status = fail(c, RMSD_ERR_ALLOC, std::string("cannot allocate scratch '") + name + "' of " +
                                          std::to_string(rounded) + " bytes");
    return nullptr;
  }
  b.bytes = rounded;
  return b.p;
}

int check_frames(rmsd_ctx* c, const rmsd_frames* f, const char* what) {
  if (!f) return fail(c, RMSD_ERR_ARG, std::string(what) + " is NULL");
  if (f->nframes < 0 || f->nsel < 1 || f->natom < 1)
    return fail(c, RMSD_ERR_ARG, std::string(what) + ": need nframes >= 0, nsel >= 1, natom >= 1");
  if (!f->xyz && f->nframes > 0) return fail(c, RMSD_ERR_ARG, std::string(what) + ": xyz is NULL");
  if (!f->sel && f->nsel > f->natom)
    return fail(c, RMSD_ERR_ARG, std::string(what) + ": nsel > natom without a selection");
  const int64_t need = f->ld ? (2 * f->ld + f->natom) : 3 * (int64_t)f->natom;
  if (f->ld && f->ld < f->natom) return fail(c, RMSD_ERR_ARG, std::string(what) + ": ld < natom");
  if (f->nframes > 1 && f->stride < need)
    return fail(c, RMSD_ERR_ARG, std::string(what) + ": frame stride smaller than one frame");
  return RMSD_OK;
}

FrameSetDev to_dev(const rmsd_frames& f) {
  FrameSetDev d;
  d.xyz = f.xyz;
  d.stride = f.stride;
  d.ld = f.ld;
  d.nframes = f.nframes;
  d.sel = f.sel;
  d.nsel = f.nsel;
  d.fidx = f.fidx;
  return d;
}

// Grouped tile order: bands of GROUP tile rows, each band swept column by
// column, so that consecutive tiles reuse the band's row blocks and one
// column block.  Tiles are tr x tc frames; triangle schedules keep only tiles
// that hold a pair (row < column).
std::vector<int2> grouped_tiles(int nbr, int nbc, int tri, int tr, int tc) {
  std::vector<int2> t;
  for (int b0 = 0; b0 < nbr; b0 += GROUP) {
    const int b1 = std::min(nbr, b0 + GROUP);
    const int j0 = tri ? (b0 * tr) / tc : 0;
    for (int j = j0; j < nbc; ++j)
      for (int i = b0; i < b1; ++i)
        if (!tri || i * tr < j * tc + tc - 1) t.push_back(make_int2(i, j));
  }
  return t;
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

namespace {
float f_ru(double x) {
  float f = (float)x;
  return (double)f < x ? nextafterf(f, INFINITY) : f;
}
float f_rd(double x) {
  float f = (float)x;
  return (double)f > x ? nextafterf(f, -INFINITY) : f;
}
}  // namespace

rmsd_dev::KeepSplit rmsd_keep_split(double tol) {
  rmsd_dev::KeepSplit k{};
  if (!(tol > 0.0) || tol == INFINITY) {  // report only: keep sqrt(eps)+ and up
    k.inv4t = 0.0f;
    k.c8 = 0.0f;
    k.h1 = k.h2 = k.e16 = k.tol = INFINITY;
    return k;
  }
  const double a = 1.0 + 0x1p-23;
  // h1 = 16 A (h1 2^-23 = 1.9e-6 taken from tol below it); for tiny tol at
  // most half of tol.
  k.h1 = f_rd(fmin(16.0, tol * 0x1p22));
  const double tol1 = tol - (double)k.h1 * 0x1p-23;
  k.inv4t = f_ru(1.0 / (4.0 * tol1));
  k.c8 = f_ru(8.0 * tol1 * tol1);
  // Above h1 up to h2 = 256 A when eps a h1 / (2 h1^2 - eps) <= tol - h2 2^-23.
  k.h2 = 256.0f;
  const double tol2 = tol - 256.0 * 0x1p-23;
  const double h1 = k.h1;
  k.e16 = tol2 > 0.0 ? f_rd(2.0 * tol2 * h1 * h1 / (a * h1 + tol2) * (1.0 - 0x1p-40)) : -1.0f;
  k.tol = f_ru(tol);
  return k;
}

BoundCoef rmsd_bound_coef(int prec, int KB) {
  BoundCoef k{};
  const double sqrt3 = 1.7320508075688772;
  // FP64 reassembly S = P + u_j + u_i^T + C and the QCP solve: a few roundings
  // of numbers up to E0 in size.
  k.e0 = 64.0 * 0x1p-53;
  if (prec == PREC_BF16X3) {
    // D = hi + lo + e with hi, lo bfloat16 (8-bit significands): |lo| <= 2^-8 |D|,
    // |e| <= 2^-16 |D| (and the float step, 2^-24).  The contraction drops
    // lo_i lo_j: bounded deterministically by Cauchy-Schwarz on the per-frame
    // ||lo|| (this term is biased, lo_i = lo_j for identical frames).
    k.drop = 1.0;
    // The residual terms e_i D_j + D_i e_j: independent zero-mean roundings,
    // each |term| <= 2^-16 |D_ik| |D_jk| (1 + 2^-8), standard deviation
    // <= that / sqrt(3); summed over atoms the deviation is at most
    // 2^-16 / sqrt(3) ||D_i,c||_4 ||D_j,d||_4 per element.
    k.rep = BOUND_LAMBDA * 2.0 * 0x1p-16 * (1.0 + 0x1p-8) / sqrt3;
    // FP32 accumulation: 3 MFMAs per 32-atom chunk.  Measured on gfx950
    // (v_mfma_f32_16x16x32_bf16): each product is rounded to nearest, ties
    // symmetric, on a grid of 2^-26 times the accumulator's binade (at most
    // 2^-26 |acc|), and the sum is rounded to float.  Per MFMA 32 such errors
    // of at most half the grid, 2^-27 m (m bounds |acc| by Cauchy-Schwarz,
    // ||D_i,c|| ||D_j,d||), and one of half a float ulp, 2^-24 m.
    const double per = sqrt(32.0) * 0x1p-27 / sqrt3 + 0x1p-24 / sqrt3;
    k.acc = BOUND_LAMBDA * sqrt(3.0 * KB) * per;
  } else if (prec == PREC_F32) {
    // D rounded to float: |D - float(D)| <= 2^-24 |D|, bounded deterministically
    // with the per-frame residual norm l: |e_i^T D_j + D_i^T e_j| <= l_i ||D_j|| + ||D_i|| l_j.
    k.mix = 1.0;
    // FP32 accumulation: 4 MFMAs per 16-atom chunk, each summing 4 products
    // (rounded to float) and the accumulator: 5 roundings of 2^-24 m.
    const double per = sqrt(5.0) * 0x1p-24 / sqrt3;
    k.acc = BOUND_LAMBDA * sqrt(4.0 * KB) * per;
  }
  return k;
}

}  // namespace rmsd_capi

using namespace rmsd_capi;

namespace {

int tiles_for(rmsd_ctx* c, int FA, int FB, int tri, int tc, const int2** tiles, int* ntiles) {
  const int nbr = (FA + TILE - 1) / TILE, nbc = (FB + tc - 1) / tc;
  TileSchedule& s = c->sched;
  if (s.nbr != nbr || s.nbc != nbc || s.tri != tri || s.tc != tc) {
    std::vector<int2> h = grouped_tiles(nbr, nbc, tri, TILE, tc);
    int st = RMSD_OK;
    s.nbr = s.nbc = s.tri = -1;
    int2* d = (int2*)scratch(c, "tiles", h.size() * sizeof(int2), &st);
    if (!d) return st;
    HIP_TRY(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice,
                              c->stream));
    // The host vector dies here: make the copy complete first.
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    s.nbr = nbr;
    s.nbc = nbc;
    s.tri = tri;
    s.tc = tc;
    s.dev = d;
    s.ntiles = (int)h.size();
  }
  *tiles = s.dev;
  *ntiles = s.ntiles;
  return RMSD_OK;
}

// Per-set operands of the pairwise contraction.
struct Packed {
  void* data;
  double* g;
  double* u;
  double* nrm = nullptr;  // MFMA forms: per-frame bound norms (4 doubles)
  double* cen = nullptr;  // MFMA forms: per-frame centroids (4 doubles)
};

int prepare_set(rmsd_ctx* c, const FrameSetDev& fs, const double* cmass, const double* smass,
                int prec, int Kp, const double* T, const double* gT, const double* M,
                const char* tag, Packed* out) {
  const int CH = chunk_atoms(prec);
  // Padded to whole row and column tiles: 64 and rmsd_pairwise_tile_cols.
  const int tc = rmsd_pairwise_tile_cols(prec);
  const int Fp = (int)round_up(fs.nframes, std::lcm(TILE, tc));
  const size_t esz = prec == PREC_F64 ? 8 : (prec == PREC_F32 ? 4 : 2);
  const size_t bytes = (size_t)Fp * Kp * 3 * esz * chunk_parts(prec);
  int st = RMSD_OK;
  std::string n(tag);
  out->data = scratch(c, (n + ".packed").c_str(), bytes, &st);
  if (!out->data) return st;
  out->g = (double*)scratch(c, (n + ".g").c_str(), sizeof(double) * Fp, &st);
  if (!out->g) return st;
  out->u = (double*)scratch(c, (n + ".u").c_str(), sizeof(double) * 9 * Fp, &st);
  if (!out->u) return st;
  if (prec != PREC_F64) {
    out->nrm = (double*)scratch(c, (n + ".nrm").c_str(), sizeof(double) * 4 * Fp, &st);
    if (!out->nrm) return st;
    out->cen = (double*)scratch(c, (n + ".cen").c_str(), sizeof(double) * 4 * Fp, &st);
    if (!out->cen) return st;
  }
  PrepArgs a;
  a.fs = fs;
  a.cmass = cmass;
  a.smass = smass;
  a.prec = prec;
  a.Kp = Kp;
  a.KB = Kp / CH;
  a.Fp = Fp;
  a.packed = out->data;
  a.g = out->g;
  a.u = out->u;
  a.T = T;
  a.gT = gT;
  a.nrm = out->nrm;
  a.cen = out->cen;
  a.total_mass = M;
  if (prec != PREC_F64) a.bound = rmsd_bound_coef(prec, a.KB);
  HIP_TRY(c, rmsd_launch_prep(a, c->stream));
  return RMSD_OK;
}

// The pairwise matrix over frame sets A (rows) and B (columns).  tri: B is A
// and only i < j is produced.  amass / bmass centre A / B; wmass (the target
// masses) weights every covariance term.
int pairwise_impl(rmsd_ctx* c, const FrameSetDev& A, const FrameSetDev& B, int tri,
                  const double* amass, const double* bmass, const double* wmass, int mode,
                  int prec, float* out, int rows_are_ref) {
  const int FA = A.nframes, FB = B.nframes;
  if (FA == 0 || FB == 0 || (tri && FA < 2)) return RMSD_OK;
  int st = RMSD_OK;
  const int2* tiles = nullptr;
  int ntiles = 0;
  const int tc = mode == RMSD_MODE_NOFIT ? TILE : rmsd_pairwise_tile_cols(prec);
  if ((st = tiles_for(c, FA, FB, tri, tc, &tiles, &ntiles)) != RMSD_OK) return st;
  const int per_xcd = (ntiles + NXCD - 1) / NXCD;
  double* consts = (double*)scratch(c, "pair.consts", sizeof(double) * 16, &st);
  if (!consts) return st;
  double* M = consts;        // total weight
  double* P = consts + 1;    // template Gram matrix (9)
  double* gT = consts + 10;  // template norm

  c->marks = 0;
  c->report = {0, nullptr};
  mark(c, 0);
  if (mode == RMSD_MODE_NOFIT) {
    SetupArgs s{};
    s.fs = A;
    s.smass = wmass;
    s.total_mass = M;
    HIP_TRY(c, rmsd_launch_setup(s, c->stream));
    mark(c, 1);
    HIP_TRY(c, rmsd_launch_pair_nofit_tiles(A, B, wmass, M, tiles, ntiles, per_xcd, NXCD, tri, out,
                                            c->stream));
    mark(c, 2);
    mark(c, 3);
    return RMSD_OK;
  }

  const int CH = chunk_atoms(prec);
  const int Kp = (int)round_up(A.nsel, CH);
  const bool mixed = prec != PREC_F64;
  double* T = nullptr;
  if (mixed) {
    T = (double*)scratch(c, "pair.template", sizeof(double) * 3 * Kp, &st);
    if (!T) return st;
  }
  SetupArgs s{};
  s.fs = B;  // template: first column (reference) frame
  s.f0 = 0;
  s.cmass = bmass;
  s.smass = wmass;
  s.Kp = Kp;
  s.T = T;
  s.P = P;
  s.gT = gT;
  s.total_mass = M;
  HIP_TRY(c, rmsd_launch_setup(s, c->stream));
  if (mixed) {
    // Mean structure of evenly spaced reference frames as the template.
    const int nsamp = std::min(TEMPLATE_SAMPLES, B.nframes);
    double* part = (double*)scratch(c, "pair.tmplpart", sizeof(double) * 3 * Kp * nsamp, &st);
    if (!part) return st;
    HIP_TRY(c, rmsd_launch_template_mean(s, nsamp, part, c->stream));
  }

  Packed pa, pb;
  if ((st = prepare_set(c, A, amass, wmass, prec, Kp, T, gT, M, "A", &pa)) != RMSD_OK) return st;
  if (tri) {
    pb = pa;
  } else if ((st = prepare_set(c, B, bmass, wmass, prec, Kp, T, gT, M, "B", &pb)) != RMSD_OK) {
    return st;
  }
  PairArgs p;
  p.A = pa.data;
  p.B = pb.data;
  p.KB = Kp / CH;
  p.gA = pa.g;
  p.gB = pb.g;
  p.uA = pa.u;
  p.uB = pb.u;
  p.P = P;
  p.total_mass = M;
  p.tiles = tiles;
  p.ntiles = ntiles;
  p.per_xcd = per_xcd;
  p.nxcd = NXCD;
  p.FA = FA;
  p.FB = FB;
  p.tri = tri;
  p.clamp = mode == RMSD_MODE_FIT ? 1 : 0;
  p.out = out;
  (void)rows_are_ref;
  const bool bounded = mixed && c->tol != 0.0;
  const WaveLayout lay = rmsd_pairwise_layout(prec);
  const int nw = lay.NWR * lay.NWC;
  unsigned long long* stats = nullptr;
  unsigned int* nflag = nullptr;
  int* wlist = nullptr;
  unsigned int* masks = nullptr;
  if (bounded) {
    // The list of flagged waves and their lane masks are written (and read)
    // only for flagged waves: no clearing.
    wlist = (int*)scratch(c, "pair.wlist", sizeof(int) * (size_t)ntiles * nw, &st);
    if (!wlist) return st;
    masks = (unsigned int*)scratch(c, "pair.masks",
                                   sizeof(unsigned int) * (size_t)ntiles * nw * 64, &st);
    if (!masks) return st;
    // {largest kept bound, pairs recomputed, (listed waves, work counter)}
    stats = (unsigned long long*)scratch(c, "pair.stats", 3 * sizeof(unsigned long long), &st);
    if (!stats) return st;
    nflag = (unsigned int*)(stats + 2);
    // A fill kernel, not a memset: the sequence stays valid under graph capture.
    HIP_TRY(c, rmsd_launch_fill_int((int*)stats, 6, 0, c->stream));
    p.nrmA = pa.nrm;
    p.nrmB = pb.nrm;
    p.ke0 = rmsd_bound_coef(prec, p.KB).e0;
    p.keep = rmsd_keep_split(c->tol);  // < 0: report only
    p.wlist = wlist;
    p.nflag = nflag;
    p.masks = masks;
    p.stats = stats;
  }
  c->report = {mixed ? (bounded ? 1 : 2) : 0, stats};
  mark(c, 1);
  HIP_TRY(c, rmsd_launch_pairwise(prec, p, c->stream));
  mark(c, 2);
  if (bounded && c->tol > 0.0) {
    FallbackArgs f;
    f.A = A;
    f.B = B;
    f.smass = wmass;
    f.cenA = pa.cen;
    f.cenB = pb.cen;
    f.gA = pa.g;
    f.gB = pb.g;
    f.total_mass = M;
    f.tiles = tiles;
    f.ntiles = ntiles;
    f.wlist = wlist;
    f.masks = masks;
    f.FA = FA;
    f.FB = FB;
    f.tri = tri;
    f.clamp = p.clamp;
    f.out = out;
    f.stats = stats;
    HIP_TRY(c, rmsd_launch_pair_fallback(f, lay, nflag, c->stream));
  }
  mark(c, 3);
  return RMSD_OK;
}

}  // namespace

extern "C" {

int rmsd_ctx_create(rmsd_ctx** ctx, int device, void* stream) {
  if (!ctx) return RMSD_ERR_ARG;
  *ctx = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RMSD_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return RMSD_ERR_DEVICE;
  rmsd_ctx* c = new rmsd_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  std::string err;
  if (noalloc_init_device(device, &c->noalloc, &err) != RMSD_OK) {
    fprintf(stderr, "rmsd_ctx_create: %s\n", err.c_str());
    delete c;
    return RMSD_ERR_DEVICE;
  }
  *ctx = c;
  return RMSD_OK;
}

void rmsd_ctx_destroy(rmsd_ctx* ctx) {
  if (!ctx) return;
  while (!ctx->plans.empty()) batch_free(*ctx->plans.begin());
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  if (ctx->events_made)
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ctx->ev[k]);
  delete ctx;
}

const char* rmsd_last_error(const rmsd_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int rmsd_ctx_synchronize(rmsd_ctx* ctx) {
  if (!ctx) return RMSD_ERR_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RMSD_OK;
}

int rmsd_ctx_set_profiling(rmsd_ctx* ctx, int enable) {
  if (!ctx) return RMSD_ERR_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (enable && !ctx->events_made) {
    for (int k = 0; k < 4; ++k) HIP_TRY(ctx, hipEventCreate(&ctx->ev[k]));
    ctx->events_made = true;
  }
  ctx->profiling = enable != 0;
  ctx->marks = 0;
  return RMSD_OK;
}

int rmsd_ctx_stage_ms(rmsd_ctx* ctx, double* ms) {
  if (!ctx || !ms) return RMSD_ERR_ARG;
  if (!ctx->profiling || ctx->marks != 15)
    return fail(ctx, RMSD_ERR_ARG, "no profiled call recorded (rmsd_ctx_set_profiling)");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
  for (int k = 0; k < 3; ++k) {
    float t = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev[k], ctx->ev[k + 1]));
    ms[k] = (double)t;
  }
  return RMSD_OK;
}

int64_t rmsd_ctx_scratch_bytes(const rmsd_ctx* ctx) {
  if (!ctx) return 0;
  int64_t s = 0;
  for (const auto& kv : ctx->bufs) s += (int64_t)kv.second.bytes;
  return s;
}

int rmsd_ctx_set_tolerance(rmsd_ctx* ctx, double tol) {
  if (!ctx) return RMSD_ERR_ARG;
  if (tol != tol) return fail(ctx, RMSD_ERR_ARG, "tolerance is NaN");
  ctx->tol = tol;
  return RMSD_OK;
}

int rmsd_ctx_error_report(rmsd_ctx* ctx, double* max_bound, int64_t* n_fp64) {
  if (!ctx || !max_bound || !n_fp64) return RMSD_ERR_ARG;
  *max_bound = 0.0;
  *n_fp64 = 0;
  if (ctx->report.kind == 2) {
    *max_bound = -1.0;  // mixed precision with the bound switched off: unknown
    return RMSD_OK;
  }
  if (ctx->report.kind != 1) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  unsigned long long h[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(h, ctx->report.stats, sizeof(h), hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  double b;
  memcpy(&b, &h[0], sizeof(b));
  *max_bound = b;
  *n_fp64 = (int64_t)h[1];
  return RMSD_OK;
}

int rmsd_pairwise(rmsd_ctx* ctx, const rmsd_frames* tgt, const rmsd_frames* ref, int mode,
                  int use_mass, int precision, float* out) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, tgt, "tgt")) != RMSD_OK) return st;
  if (ref && (st = check_frames(ctx, ref, "ref")) != RMSD_OK) return st;
  if (ref && ref->nsel != tgt->nsel)
    return fail(ctx, RMSD_ERR_ARG, "reference and target select different atom counts");
  if (mode != RMSD_MODE_FIT && mode != RMSD_MODE_NOFIT && mode != RMSD_MODE_QUAT)
    return fail(ctx, RMSD_ERR_ARG, "unknown mode");
  if (precision < RMSD_PREC_FP64 || precision > RMSD_PREC_BF16X3)
    return fail(ctx, RMSD_ERR_ARG, "unknown precision");
  if (use_mass && (!tgt->mass || (ref && !ref->mass)))
    return fail(ctx, RMSD_ERR_ARG, "use_mass needs masses for every frame set");
  if (!out && tgt->nframes > 0) return fail(ctx, RMSD_ERR_ARG, "out is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const FrameSetDev T = to_dev(*tgt);
  const double* tm = use_mass ? tgt->mass : nullptr;
  if (!ref) return pairwise_impl(ctx, T, T, 1, tm, tm, tm, mode, precision, out, 1);
  const FrameSetDev R = to_dev(*ref);
  const double* rm = use_mass ? ref->mass : nullptr;
  // Rows are targets, columns references: element (ntgt, nref) at ntgt*nref_total + nref.
  return pairwise_impl(ctx, T, R, 0, tm, rm, tm, mode, precision, out, 0);
}

int rmsd_matrix_autocorr(rmsd_ctx* ctx, const float* tri, int n, double* ct) {
  if (!ctx) return RMSD_ERR_ARG;
  if (n < 0 || (n > 0 && (!tri && n > 1)) || (n > 0 && !ct))
    return fail(ctx, RMSD_ERR_ARG, "autocorr: need n >= 0 and non-NULL buffers");
  if (n == 0) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rows = 0, st = RMSD_OK;
  const int nchunks = rmsd_autocorr_chunks(n, &rows);
  double* part = (double*)scratch(ctx, "corr.part", sizeof(double) * (size_t)nchunks * n, &st);
  if (!part) return st;
  HIP_TRY(ctx, rmsd_launch_autocorr(tri, n, part, ct, ctx->stream));
  return RMSD_OK;
}

int rmsd_cluster_cache(rmsd_ctx* ctx, const rmsd_frames* frames, int nofit, int use_mass,
                       int precision, double weight, float* out) {
  if (!ctx) return RMSD_ERR_ARG;
  if (weight != 1.0)
    return fail(ctx, RMSD_ERR_ARG, "metric weights other than 1 are not supported");
  return rmsd_pairwise(ctx, frames, nullptr, nofit ? RMSD_MODE_NOFIT : RMSD_MODE_FIT, use_mass,
                       precision, out);
}

int rmsd_fit_to_ref(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                    int ref_frame, int fit, int use_mass, int modify, double* rmsd, double* U,
                    double* tvec) {
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
  const FrameSetDev F = to_dev(*frames);
  const FrameSetDev R = to_dev(*ref);
  ctx->marks = 0;
  mark(ctx, 0);
  HIP_TRY(ctx, rmsd_launch_ref_prep(R, ref_frame, use_mass ? ref->mass : nullptr, wpad, fit ? 1 : 0,
                                    refbuf, ref_trans, consts, ctx->stream));
  if (!fit) {
    NofitArgs a;
    a.fs = F;
    a.mass = wpad;
    a.ref = refbuf;
    mark(ctx, 1);
    HIP_TRY(ctx, rmsd_launch_nofit_frames(a, consts, rmsd, ctx->stream));
    mark(ctx, 2);
    mark(ctx, 3);
    return RMSD_OK;
  }
  double* moments = (double*)scratch(ctx, "fit.moments", sizeof(double) * 16 * nf, &st);
  if (!moments) return st;
  double* xform = nullptr;
  if (modify != RMSD_MODIFY_NONE) {
    xform = (double*)scratch(ctx, "fit.xform", sizeof(double) * 15 * nf, &st);
    if (!xform) return st;
  }
  FitArgs fa;
  fa.fs = F;
  fa.mass = wpad;
  fa.ref = refbuf;
  fa.consts = consts;
  fa.moments = moments;
  mark(ctx, 1);
  HIP_TRY(ctx, rmsd_launch_fit_moments(fa, ctx->stream));
  mark(ctx, 2);
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
  sa.xform = xform;
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
  mark(ctx, 3);
  return RMSD_OK;
}

int rmsd_fit_pairs(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* refs, int fit,
                   int use_mass, int modify, double* rmsd, double* U, double* tvec) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, frames, "frames")) != RMSD_OK) return st;
  if ((st = check_frames(ctx, refs, "refs")) != RMSD_OK) return st;
  if (refs->nsel != frames->nsel)
    return fail(ctx, RMSD_ERR_ARG, "reference and target select different atom counts");
  if (refs->nframes < frames->nframes)
    return fail(ctx, RMSD_ERR_ARG, "fewer reference frames than target frames");
  if (frames->nsel < 1 && frames->nframes > 0)
    return fail(ctx, RMSD_ERR_ARG, "no atoms selected");
  if (use_mass && (!frames->mass || !refs->mass))
    return fail(ctx, RMSD_ERR_ARG, "use_mass needs target and reference masses");
  if (modify < RMSD_MODIFY_NONE || modify > RMSD_MODIFY_TRANS)
    return fail(ctx, RMSD_ERR_ARG, "unknown modify mode");
  if (modify != RMSD_MODIFY_NONE && (frames->fidx || !fit))
    return fail(ctx, RMSD_ERR_ARG,
                "coordinate update needs a fit and frames in storage order (fidx NULL)");
  if (!rmsd && frames->nframes > 0) return fail(ctx, RMSD_ERR_ARG, "rmsd is NULL");
  if (frames->nframes == 0) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t nf = frames->nframes;
  double* W = (double*)scratch(ctx, "pairfit.W", sizeof(double), &st);
  if (!W) return st;
  double* xform = nullptr;
  if (modify != RMSD_MODIFY_NONE) {
    xform = (double*)scratch(ctx, "pairfit.xform", sizeof(double) * 15 * nf, &st);
    if (!xform) return st;
  }
  ctx->marks = 0;
  mark(ctx, 0);
  HIP_TRY(ctx, rmsd_launch_mass_total(use_mass ? frames->mass : nullptr, frames->nsel, W,
                                      ctx->stream));
  mark(ctx, 1);
  PairFitArgs a;
  a.T = to_dev(*frames);
  a.R = to_dev(*refs);
  a.tmass = use_mass ? frames->mass : nullptr;
  a.rmass = use_mass ? refs->mass : nullptr;
  a.W = W;
  a.fit = fit ? 1 : 0;
  a.clamp = 1;
  a.rmsd = rmsd;
  a.U = fit ? U : nullptr;
  a.tvec = fit ? tvec : nullptr;
  a.xform = xform;
  HIP_TRY(ctx, rmsd_launch_fit_pairs(a, ctx->stream));
  mark(ctx, 2);
  if (modify != RMSD_MODIFY_NONE) {
    // every pair is solved before any target is rewritten (stream order), so
    // references may share the targets' storage ('previous')
    ApplyArgs ap;
    ap.xyz = const_cast<float*>(frames->xyz);
    ap.stride = frames->stride;
    ap.ld = frames->ld;
    ap.natom = frames->natom;
    ap.nframes = frames->nframes;
    ap.xform = xform;
    HIP_TRY(ctx, rmsd_launch_apply(ap, modify == RMSD_MODIFY_ROT_TRANS ? 1 : 2, ctx->stream));
  }
  mark(ctx, 3);
  return RMSD_OK;
}

int rmsd_rotate_ucell(rmsd_ctx* ctx, const double* U, int nframes, double* ucell) {
  if (!ctx) return RMSD_ERR_ARG;
  if (nframes < 0 || (nframes > 0 && (!U || !ucell)))
    return fail(ctx, RMSD_ERR_ARG, "rotate_ucell: need nframes >= 0 and non-NULL U and ucell");
  if (nframes == 0) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, rmsd_launch_rotate_ucell(U, nframes, ucell, ctx->stream));
  return RMSD_OK;
}

int rmsd_transform(rmsd_ctx* ctx, const rmsd_frames* frames, const int32_t* atoms, int natoms,
                   const double* U, const double* tvec, int mode) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, frames, "frames")) != RMSD_OK) return st;
  if (mode != RMSD_MODIFY_ROT_TRANS && mode != RMSD_MODIFY_TRANS)
    return fail(ctx, RMSD_ERR_ARG, "transform mode must be ROT_TRANS or TRANS");
  if (!tvec || (mode == RMSD_MODIFY_ROT_TRANS && !U))
    return fail(ctx, RMSD_ERR_ARG, "transform needs tvec (and U for ROT_TRANS)");
  if (atoms && natoms < 0) return fail(ctx, RMSD_ERR_ARG, "natoms < 0");
  if (frames->nframes == 0 || (atoms && natoms == 0)) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ApplyArgs ap;
  ap.xyz = const_cast<float*>(frames->xyz);
  ap.stride = frames->stride;
  ap.ld = frames->ld;
  ap.natom = frames->natom;
  ap.nframes = frames->nframes;
  ap.xform = nullptr;
  ap.atoms = atoms;
  ap.natoms = natoms;
  ap.U = mode == RMSD_MODIFY_ROT_TRANS ? U : nullptr;
  ap.tvec = tvec;
  HIP_TRY(ctx, rmsd_launch_apply(ap, mode == RMSD_MODIFY_ROT_TRANS ? 1 : 2, ctx->stream));
  return RMSD_OK;
}

// Per-residue no-fit RMSD.  The residue table is validated here, cut into
// pieces whose LDS image fits one workgroup, and uploaded as one blob:
//   res_off (nres+1 int) | tgt (L int) | ref (L int) | pieces (P int4) |
//   tmass (L double) | rmass (L double)
int rmsd_perres_nofit(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                      int ref_frame, int nres, const int32_t* res_off, const int32_t* tgt_atoms,
                      const int32_t* ref_atoms, const double* tmass, const double* rmass,
                      int use_mass, int center, int modify, const double* U, const double* tvec,
                      double* out) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if (!frames || !ref) return fail(ctx, RMSD_ERR_ARG, "frames or ref is NULL");
  if (frames->nframes < 0 || frames->natom < 1 || (!frames->xyz && frames->nframes > 0))
    return fail(ctx, RMSD_ERR_ARG, "frames: need nframes >= 0, natom >= 1 and xyz");
  if (ref->natom < 1 || !ref->xyz) return fail(ctx, RMSD_ERR_ARG, "ref: need natom >= 1 and xyz");
  for (const rmsd_frames* f : {frames, ref}) {
    if (f->ld && f->ld < f->natom) return fail(ctx, RMSD_ERR_ARG, "ld < natom");
    const int64_t need = f->ld ? (2 * f->ld + f->natom) : 3 * (int64_t)f->natom;
    if (f->nframes > 1 && f->stride < need)
      return fail(ctx, RMSD_ERR_ARG, "frame stride smaller than one frame");
  }
  if (ref_frame < 0 || ref_frame >= ref->nframes)
    return fail(ctx, RMSD_ERR_ARG, "reference frame out of range");
  if (nres < 0 || (nres > 0 && (!res_off || !tgt_atoms || !ref_atoms)))
    return fail(ctx, RMSD_ERR_ARG, "residue table incomplete");
  if (use_mass && (!tmass || !rmass))
    return fail(ctx, RMSD_ERR_ARG, "use_mass needs target and reference masses");
  if (modify < RMSD_MODIFY_NONE || modify > RMSD_MODIFY_TRANS)
    return fail(ctx, RMSD_ERR_ARG, "unknown modify mode");
  if (modify != RMSD_MODIFY_NONE && (!tvec || (modify == RMSD_MODIFY_ROT_TRANS && !U)))
    return fail(ctx, RMSD_ERR_ARG, "modify needs tvec (and U for ROT_TRANS)");
  if (nres == 0 || frames->nframes == 0) return RMSD_OK;
  if (!out) return fail(ctx, RMSD_ERR_ARG, "out is NULL");
  if (res_off[0] != 0) return fail(ctx, RMSD_ERR_ARG, "res_off[0] must be 0");
  for (int r = 0; r < nres; ++r)
    if (res_off[r + 1] < res_off[r]) return fail(ctx, RMSD_ERR_ARG, "res_off is decreasing");
  const int L = res_off[nres];
  for (int k = 0; k < L; ++k) {
    if (tgt_atoms[k] < 0 || tgt_atoms[k] >= frames->natom)
      return fail(ctx, RMSD_ERR_ARG, "target atom index out of range");
    if (ref_atoms[k] < 0 || ref_atoms[k] >= ref->natom)
      return fail(ctx, RMSD_ERR_ARG, "reference atom index out of range");
  }
  // Pieces: runs of whole residues that fit PERRES_LDS_PIECE at one frame.  A
  // piece stages its target atom span, or gathers its own atoms when the span
  // is more than twice their number (lo = -1); a residue that would switch a
  // run from one form to the other starts a new piece.
  std::vector<int4> pieces;
  auto staged = [](int span, int ne) { return span <= 2 * ne + 64 ? span : -1; };
  for (int r = 0; r < nres;) {
    int lo = INT32_MAX, hi = -1, r1 = r;
    bool span_form = true;
    while (r1 < nres) {
      int l2 = lo, h2 = hi;
      for (int k = res_off[r1]; k < res_off[r1 + 1]; ++k) {
        l2 = std::min(l2, tgt_atoms[k]);
        h2 = std::max(h2, tgt_atoms[k] + 1);
      }
      const int ne = res_off[r1 + 1] - res_off[r];
      const int span = h2 > l2 ? h2 - l2 : 0;
      const bool sf = staged(span, ne) >= 0;
      const size_t b = rmsd_perres_lds_bytes(r1 + 1 - r, ne, sf ? span : ne, 1);
      if (r1 > r && (b > PERRES_LDS_PIECE || sf != span_form || r1 - r >= PERRES_RES_PIECE))
        break;
      if (r1 == r && b > PERRES_LDS_MAX)
        return fail(ctx, RMSD_ERR_ARG, "residue " + std::to_string(r) + " has too many atoms (" +
                                           std::to_string(ne) + ") for per-residue RMSD");
      lo = l2;
      hi = h2;
      span_form = sf;
      ++r1;
    }
    if (hi < 0) lo = hi = 0;  // residues without atoms
    pieces.push_back(span_form ? make_int4(r, r1, lo, hi) : make_int4(r, r1, -1, -1));
    r = r1;
  }
  // Frames per batch: enough (frame, residue) items for the workgroup's lanes,
  // within PERRES_LDS_BATCH (one frame always runs).
  int maxnr = 1;
  for (const int4& p : pieces) maxnr = std::max(maxnr, p.y - p.x);
  int nfb = std::max(1, PERRES_THREADS / maxnr);
  auto lds_for = [&](int nf) {
    size_t m = 0;
    for (const int4& p : pieces)
      m = std::max(m, rmsd_perres_lds_bytes(p.y - p.x, res_off[p.y] - res_off[p.x],
                                            p.z < 0 ? res_off[p.y] - res_off[p.x] : p.w - p.z, nf));
    return m;
  };
  if (pieces.size() > 65535) return fail(ctx, RMSD_ERR_ARG, "residue table too large");
  while (nfb > 1 && lds_for(nfb) > PERRES_LDS_BATCH) --nfb;
  const size_t lds = lds_for(nfb);
  // Table blob (16-byte aligned sections).
  const size_t o_off = 0;
  const size_t o_tgt = round_up(o_off + sizeof(int32_t) * (nres + 1), 16);
  const size_t o_ref = round_up(o_tgt + sizeof(int32_t) * L, 16);
  const size_t o_pc = round_up(o_ref + sizeof(int32_t) * L, 16);
  const size_t o_tm = round_up(o_pc + sizeof(int4) * pieces.size(), 16);
  const size_t o_rm = round_up(o_tm + (use_mass ? sizeof(double) * L : 0), 16);
  const size_t bytes = round_up(o_rm + (use_mass ? sizeof(double) * L : 0), 16);
  std::vector<unsigned char> blob(bytes, 0);
  memcpy(blob.data() + o_off, res_off, sizeof(int32_t) * (nres + 1));
  memcpy(blob.data() + o_tgt, tgt_atoms, sizeof(int32_t) * L);
  memcpy(blob.data() + o_ref, ref_atoms, sizeof(int32_t) * L);
  memcpy(blob.data() + o_pc, pieces.data(), sizeof(int4) * pieces.size());
  if (use_mass) {
    memcpy(blob.data() + o_tm, tmass, sizeof(double) * L);
    memcpy(blob.data() + o_rm, rmass, sizeof(double) * L);
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  unsigned char* dtab = (unsigned char*)scratch(ctx, "perres.table", bytes, &st);
  if (!dtab) return st;
  if (blob != ctx->perres_blob) {
    // The previous table may still be in flight from the host image.
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->perres_blob.swap(blob);
    HIP_TRY(ctx, hipMemcpyAsync(dtab, ctx->perres_blob.data(), bytes, hipMemcpyHostToDevice,
                                ctx->stream));
  }
  PerresArgs a;
  a.xyz = frames->xyz;
  a.stride = frames->stride;
  a.ld = frames->ld;
  a.fidx = frames->fidx;
  a.nframes = frames->nframes;
  a.ref = to_dev(*ref);
  a.ref_frame = ref_frame;
  a.nres = nres;
  a.res_off = (const int32_t*)(dtab + o_off);
  a.tgt_atoms = (const int32_t*)(dtab + o_tgt);
  a.ref_atoms = (const int32_t*)(dtab + o_ref);
  a.tmass = use_mass ? (const double*)(dtab + o_tm) : nullptr;
  a.rmass = use_mass ? (const double*)(dtab + o_rm) : nullptr;
  a.use_mass = use_mass ? 1 : 0;
  a.center = center ? 1 : 0;
  a.U = modify == RMSD_MODIFY_ROT_TRANS ? U : nullptr;
  a.tvec = modify != RMSD_MODIFY_NONE ? tvec : nullptr;
  a.modify = modify == RMSD_MODIFY_ROT_TRANS ? 1 : (modify == RMSD_MODIFY_TRANS ? 2 : 0);
  a.pieces = (const int4*)(dtab + o_pc);
  a.nfb = nfb;
  a.out = out;
  ctx->marks = 0;
  mark(ctx, 0);
  mark(ctx, 1);
  HIP_TRY(ctx, rmsd_launch_perres(a, (int)pieces.size(), lds, ctx->stream));
  mark(ctx, 2);
  mark(ctx, 3);
  return RMSD_OK;
}

int rmsd_avgcorr_windows(int nframes, int maxwindow, int offset, int32_t* windows, int cap) {
  if (offset < 1) offset = 1;
  int wmax = nframes;
  if (maxwindow != -1) {
    wmax = maxwindow + 1;
    if (wmax > nframes) wmax = nframes;
  }
  const int start = 1 + offset;
  int total = (wmax - start) / offset;
  if (total % offset > 0) ++total;
  if (total < 0) return -1;  // not enough frames
  int n = 0;
  if (windows && n < cap) windows[n] = 1;
  ++n;
  for (int ws = start; ws < wmax; ws += offset) {
    if (windows && n < cap) windows[n] = ws;
    ++n;
  }
  return n;
}

int rmsd_avgcorr_set_method(rmsd_ctx* ctx, int method) {
  if (!ctx) return RMSD_ERR_ARG;
  if (method != RMSD_AVGCORR_AUTO && method != RMSD_AVGCORR_RUNNING &&
      method != RMSD_AVGCORR_PREFIX)
    return fail(ctx, RMSD_ERR_ARG, "unknown rmsavgcorr method " + std::to_string(method));
  ctx->avgcorr_method = method;
  return RMSD_OK;
}

int rmsd_avgcorr(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref, int ref_frame,
                 int use_mass, int maxwindow, int offset, double* ct, double* sd) {
  if (!ctx) return RMSD_ERR_ARG;
  int st;
  if ((st = check_frames(ctx, frames, "frames")) != RMSD_OK) return st;
  if (ref) {
    if ((st = check_frames(ctx, ref, "ref")) != RMSD_OK) return st;
    if (ref->nsel != frames->nsel)
      return fail(ctx, RMSD_ERR_ARG, "reference and target select different atom counts");
    if (ref_frame < 0 || ref_frame >= ref->nframes)
      return fail(ctx, RMSD_ERR_ARG, "reference frame out of range");
    if (use_mass && !ref->mass) return fail(ctx, RMSD_ERR_ARG, "use_mass needs reference masses");
  }
  if (use_mass && !frames->mass) return fail(ctx, RMSD_ERR_ARG, "use_mass needs target masses");
  const int nw = rmsd_avgcorr_windows(frames->nframes, maxwindow, offset, nullptr, 0);
  if (nw < 1) return fail(ctx, RMSD_ERR_ARG, "Not enough frames to perform calculation.");
  if (!ct || !sd) return fail(ctx, RMSD_ERR_ARG, "ct or sd is NULL");
  // Method: the prefix form needs its memory ('first' also one reference
  // per window).
  const int64_t F = frames->nframes;
  const int k = 3 * frames->nsel;
  const bool first = ref == nullptr;
  const int T = rmsd_avgcorr_tile(first);
  const int ntile = (int)((F + 1 + T - 1) / T);
  const size_t prefix_bytes =
      sizeof(double) * ((size_t)(F + 1) * k + (size_t)rmsd_avgcorr_chunks(F) * k + 2 * k +
                        (size_t)(F + 2) * 13 + (size_t)nw * ntile * 4 +
                        (first ? (size_t)nw * (k + 13) : 0));
  int method = ctx->avgcorr_method;
  if (method == RMSD_AVGCORR_AUTO)
    method = prefix_bytes <= AVGCORR_PREFIX_MAX ? RMSD_AVGCORR_PREFIX : RMSD_AVGCORR_RUNNING;
  if (method == RMSD_AVGCORR_PREFIX) {
    if (prefix_bytes > AVGCORR_PREFIX_MAX)
      return fail(ctx, RMSD_ERR_ARG, "rmsavgcorr: the prefix method needs " +
                                         std::to_string(prefix_bytes) + " bytes, over " +
                                         std::to_string(AVGCORR_PREFIX_MAX));
  } else {
    // LDS: the running sum, and the reference in 'first' mode (FP64 [3][nsel]).
    const size_t per_atom = sizeof(double) * (ref ? 3 : 6);
    if (per_atom * (size_t)frames->nsel > AVGCORR_LDS_MAX)
      return fail(ctx, RMSD_ERR_ARG, "rmsavgcorr: at most " +
                                         std::to_string(AVGCORR_LDS_MAX / per_atom) +
                                         " selected atoms" + (ref ? "" : " with 'first'"));
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<int32_t> ws(nw);
  rmsd_avgcorr_windows(frames->nframes, maxwindow, offset, ws.data(), nw);
  int32_t* dws = (int32_t*)scratch(ctx, "avgcorr.windows", sizeof(int32_t) * nw, &st);
  if (!dws) return st;
  HIP_TRY(ctx, hipMemcpyAsync(dws, ws.data(), sizeof(int32_t) * nw, hipMemcpyHostToDevice,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // ws is a local buffer
  ctx->marks = 0;
  mark(ctx, 0);
  double* refbuf = nullptr;
  double* consts = nullptr;
  const int ldr = (int)round_up(frames->nsel, 4);
  if (ref) {
    refbuf = (double*)scratch(ctx, "avgcorr.ref", sizeof(double) * (3 * ldr + 8), &st);
    if (!refbuf) return st;
    consts = (double*)scratch(ctx, "avgcorr.consts", sizeof(double) * 8, &st);
    if (!consts) return st;
    HIP_TRY(ctx, rmsd_launch_ref_prep(to_dev(*ref), ref_frame, use_mass ? ref->mass : nullptr,
                                      use_mass ? frames->mass : nullptr, 1, refbuf,
                                      refbuf + 3 * ldr, consts, ctx->stream));
  }
  if (method == RMSD_AVGCORR_RUNNING) {
    AvgCorrArgs a;
    a.fs = to_dev(*frames);
    a.mass = use_mass ? frames->mass : nullptr;
    a.ref = refbuf;
    a.consts = consts;
    a.ldr = ref ? ldr : 0;
    a.windows = dws;
    a.nwin = nw;
    a.ct = ct;
    a.sd = sd;
    mark(ctx, 1);
    HIP_TRY(ctx, rmsd_launch_avgcorr(a, ref == nullptr, ctx->stream));
    mark(ctx, 2);
    mark(ctx, 3);
    return RMSD_OK;
  }
  // Prefix form: window map and the (row tile, column tile) pairs that hold
  // a window diagonal.
  std::vector<int32_t> wmap((size_t)F + 1, -1);
  for (int i = 0; i < nw; ++i) wmap[ws[i]] = i;
  std::vector<int32_t> cnt((size_t)F + 2, 0);  // cnt[w + 1] = windows <= w
  for (int64_t w = 0; w <= F; ++w) cnt[w + 1] = cnt[w] + (wmap[w] >= 0 ? 1 : 0);
  std::vector<int2> tiles;
  for (int ti = 0; ti < ntile; ++ti)
    for (int tj = 0; tj <= ti; ++tj) {
      const int64_t base = (int64_t)T * (ti - tj);
      const int64_t lo = std::max<int64_t>(1, base - (T - 1)), hi = std::min<int64_t>(F, base + T - 1);
      if (lo <= hi && cnt[hi + 1] > cnt[lo]) tiles.push_back(make_int2(ti, tj));
    }
  AvgPrefixArgs a;
  a.fs = to_dev(*frames);
  a.mass = use_mass ? frames->mass : nullptr;
  a.ref = refbuf;
  a.consts = consts;
  a.ldr = ldr;
  a.k = k;
  a.chunk = (double*)scratch(ctx, "avgcorr.chunk", sizeof(double) * rmsd_avgcorr_chunks(F) * k, &st);
  if (!a.chunk) return st;
  a.mu = (double*)scratch(ctx, "avgcorr.mu", sizeof(double) * k, &st);
  if (!a.mu) return st;
  a.P = (double*)scratch(ctx, "avgcorr.prefix", sizeof(double) * (size_t)(F + 1) * k, &st);
  if (!a.P) return st;
  a.rows = (double*)scratch(ctx, "avgcorr.rows", sizeof(double) * (size_t)(F + 2) * 13, &st);
  if (!a.rows) return st;
  int32_t* dmap = (int32_t*)scratch(ctx, "avgcorr.wmap", sizeof(int32_t) * (size_t)(F + 1), &st);
  if (!dmap) return st;
  int2* dtiles = (int2*)scratch(ctx, "avgcorr.tiles", sizeof(int2) * tiles.size(), &st);
  if (!dtiles) return st;
  a.part = (double*)scratch(ctx, "avgcorr.part", sizeof(double) * (size_t)nw * ntile * 4, &st);
  if (!a.part) return st;
  HIP_TRY(ctx, hipMemcpyAsync(dmap, wmap.data(), sizeof(int32_t) * wmap.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(dtiles, tiles.data(), sizeof(int2) * tiles.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // host vectors are local
  a.wmap = dmap;
  a.tiles = dtiles;
  a.ntiles = (int)tiles.size();
  a.ntile = ntile;
  a.tile = T;
  a.yref = nullptr;
  a.wconst = nullptr;
  if (first) {
    double* wc = (double*)scratch(ctx, "avgcorr.consts", sizeof(double) * 8, &st);
    if (!wc) return st;
    a.consts = wc;
    HIP_TRY(ctx, rmsd_launch_avgcorr_wsum(a, wc, ctx->stream));
    a.yref = (double*)scratch(ctx, "avgcorr.yref", sizeof(double) * (size_t)nw * k, &st);
    if (!a.yref) return st;
    a.wconst = (double*)scratch(ctx, "avgcorr.wconst", sizeof(double) * (size_t)nw * 13, &st);
    if (!a.wconst) return st;
  }
  a.windows = dws;
  a.nwin = nw;
  a.ct = ct;
  a.sd = sd;
  HIP_TRY(ctx, rmsd_launch_avgcorr_prefix(a, ctx->stream));
  mark(ctx, 1);
  HIP_TRY(ctx, rmsd_launch_avgcorr_tiles(a, ctx->stream));
  mark(ctx, 2);
  HIP_TRY(ctx, rmsd_launch_avgcorr_final(a, ctx->stream));
  mark(ctx, 3);
  return RMSD_OK;
}

int rmsd_symm(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref, int ref_frame,
              int ngroups, const int32_t* group_off, const int32_t* group_atoms, int fit,
              int use_mass, double* rmsd, int32_t* amap, double* U, double* tvec) {
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
  if (ngroups < 0 || (ngroups > 0 && (!group_off || !group_atoms)))
    return fail(ctx, RMSD_ERR_ARG, "symmetric groups: need group_off and group_atoms");
  if (!rmsd && frames->nframes > 0) return fail(ctx, RMSD_ERR_ARG, "rmsd is NULL");
  const int n = frames->nsel;
  // Group table: offsets, atoms, each atom's entry, and the scratch offsets
  // of groups too large for the register solver.
  std::vector<int32_t> pos((size_t)n, -1), bigoff((size_t)std::max(ngroups, 1), -1);
  int64_t bigstride = 0;
  if (ngroups > 0 && group_off[0] != 0)
    return fail(ctx, RMSD_ERR_ARG, "symmetric groups: group_off[0] must be 0");
  for (int g = 0; g < ngroups; ++g) {
    const int b = group_off[g], e = group_off[g + 1];
    if (e - b < 1 || e - b > SYMM_GROUP_MAX)
      return fail(ctx, RMSD_ERR_ARG, "symmetric group " + std::to_string(g) + " has " +
                                         std::to_string(e - b) + " atoms (1.." +
                                         std::to_string(SYMM_GROUP_MAX) + " allowed)");
    for (int k = b; k < e; ++k) {
      const int a = group_atoms[k];
      if (a < 0 || a >= n)
        return fail(ctx, RMSD_ERR_ARG, "symmetric group atom out of range of the selection");
      if (pos[a] >= 0)
        return fail(ctx, RMSD_ERR_ARG, "selected atom " + std::to_string(a) +
                                           " is in more than one symmetric group");
      pos[a] = k;
    }
    if (e - b > SYMM_GROUP_REG) {
      bigoff[g] = (int32_t)bigstride;
      bigstride += (int64_t)(e - b) * (e - b) + (6 * (e - b) + 7) / 8;
    }
  }
  const int S = ngroups > 0 ? group_off[ngroups] : 0;
  if (frames->nframes == 0) return RMSD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t nf = frames->nframes;
  const int ldr = (int)round_up(n, 4);
  // Groups in the device table are ordered by size (stable): a wave takes
  // consecutive groups of one frame, which then share one size.  Every group
  // is assigned on its own, so the order changes no result.
  std::vector<int> order((size_t)ngroups);
  for (int g = 0; g < ngroups; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return group_off[x + 1] - group_off[x] < group_off[y + 1] - group_off[y];
  });
  std::vector<int32_t> goff((size_t)ngroups + 1, 0), gat;
  std::vector<int32_t> bigo((size_t)std::max(ngroups, 1), -1);
  gat.reserve(S);
  for (int g = 0; g < ngroups; ++g) {
    const int src = order[g];
    for (int k = group_off[src]; k < group_off[src + 1]; ++k) {
      pos[group_atoms[k]] = (int32_t)gat.size();
      gat.push_back(group_atoms[k]);
    }
    goff[g + 1] = (int32_t)gat.size();
    bigo[g] = bigoff[src];
  }
  std::vector<int32_t> blob;
  blob.reserve((size_t)(ngroups + 1) + S + n + ngroups);
  blob.insert(blob.end(), goff.begin(), goff.end());
  blob.insert(blob.end(), gat.begin(), gat.end());
  blob.insert(blob.end(), pos.begin(), pos.end());
  blob.insert(blob.end(), bigo.begin(), bigo.begin() + ngroups);
  int32_t* dblob = (int32_t*)scratch(ctx, "symm.table", sizeof(int32_t) * blob.size(), &st);
  if (!dblob) return st;
  HIP_TRY(ctx, hipMemcpyAsync(dblob, blob.data(), sizeof(int32_t) * blob.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // blob is a local buffer
  double* refbuf = (double*)scratch(ctx, "symm.ref", sizeof(double) * (3 * ldr + 8), &st);
  if (!refbuf) return st;
  double* ref_trans = refbuf + 3 * ldr;
  double* consts = (double*)scratch(ctx, "symm.consts", sizeof(double) * 8, &st);
  if (!consts) return st;
  double* wpad = nullptr;
  if (use_mass) {
    wpad = (double*)scratch(ctx, "symm.mass", sizeof(double) * ldr, &st);
    if (!wpad) return st;
    HIP_TRY(ctx, hipMemsetAsync(wpad, 0, sizeof(double) * ldr, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(wpad, frames->mass, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                ctx->stream));
  }
  int* gmap = (int*)scratch(ctx, "symm.gmap", sizeof(int) * (size_t)(nf * std::max(S, 1)), &st);
  if (!gmap) return st;
  double* big = nullptr;
  if (bigstride > 0) {
    // Frames of one assignment launch: as many as SYMM_BIG_MAX bytes hold.
    const int64_t chunk =
        std::min<int64_t>(nf, std::max<int64_t>(1, (int64_t)(SYMM_BIG_MAX / (sizeof(double) * bigstride))));
    big = (double*)scratch(ctx, "symm.big", sizeof(double) * (size_t)(chunk * bigstride), &st);
    if (!big) return st;
  }
  const FrameSetDev F = to_dev(*frames);
  ctx->marks = 0;
  mark(ctx, 0);
  HIP_TRY(ctx, rmsd_launch_ref_prep(to_dev(*ref), ref_frame, use_mass ? ref->mass : nullptr, wpad,
                                    fit ? 1 : 0, refbuf, ref_trans, consts, ctx->stream));
  double* U0 = nullptr;
  double* t0 = nullptr;
  double* moments = nullptr;
  if (fit) {
    // Initial fit of the frames as given: U and Trans for the cost matrices,
    // moments for the remapped fit.
    moments = (double*)scratch(ctx, "symm.moments", sizeof(double) * 16 * nf, &st);
    if (!moments) return st;
    double* fit0 = (double*)scratch(ctx, "symm.fit0", sizeof(double) * 16 * nf, &st);
    if (!fit0) return st;
    U0 = fit0;
    t0 = fit0 + 9 * nf;
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
    sa.rmsd = fit0 + 15 * nf;
    sa.U = U0;
    sa.tvec = t0;
    sa.xform = nullptr;
    HIP_TRY(ctx, rmsd_launch_fit_solve(sa, ctx->stream));
  }
  mark(ctx, 1);
  SymmArgs a;
  a.fs = F;
  a.mass = use_mass ? frames->mass : nullptr;
  a.ref = refbuf;
  a.ldr = ldr;
  a.ref_trans = ref_trans;
  a.fit = fit ? 1 : 0;
  a.U0 = U0;
  a.t0 = t0;
  a.moments = moments;
  a.consts = consts;
  a.ngroups = ngroups;
  a.goff = dblob;
  a.gatoms = dblob + (ngroups + 1);
  a.ngatoms = S;
  a.pos = dblob + (ngroups + 1) + S;
  a.bigoff = dblob + (ngroups + 1) + S + n;
  a.bigstride = bigstride;
  a.big = big;
  a.f0 = 0;
  a.f1 = nf;
  a.gmap = gmap;
  a.rmsd = rmsd;
  a.amap = amap;
  a.U = fit ? U : nullptr;
  a.tvec = fit ? tvec : nullptr;
  HIP_TRY(ctx, rmsd_launch_symm(a, ctx->stream));
  mark(ctx, 2);
  mark(ctx, 3);
  return RMSD_OK;
}

}  // extern "C"
