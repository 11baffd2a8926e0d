// capi.hip -- implementation of the C ABI declared in include/cpptraj_rmsd.h.
//
// Host-side orchestration only: argument checks, scratch management, tile
// schedules and the launch sequence of the kernels in k_*.hip.  Every entry
// point enqueues work on the context's stream and returns without waiting
// (except rmsd_ctx_synchronize).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "cpptraj_rmsd.h"
#include "rmsd_kernels.hpp"

namespace {

constexpr int TILE = 64;      // frames per tile side (all pairwise kernels)
constexpr int GROUP = 8;      // tile rows per band of the grouped schedule
constexpr int NXCD = 8;       // MI355X: 8 XCDs, workgroups dealt round-robin

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct TileSchedule {
  int nbr = -1, nbc = -1, tri = -1, tc = -1;
  int2* dev = nullptr;
  int ntiles = 0;
};

}  // namespace

struct rmsd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<std::string, DevBuf> bufs;
  TileSchedule sched;
  // Stage timing: events recorded on the context's stream around the per-frame
  // stage, the main kernel and the epilogue stage of the last call.
  bool profiling = false;
  bool events_made = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int marks = 0;
  // Batch plans bound to this context; freed with it if still alive.
  std::set<rmsd_batch*> plans;
};

namespace {
void batch_free(rmsd_batch* b);
}  // namespace

namespace {

void mark(rmsd_ctx* c, int k) {
  if (!c->profiling || !c->events_made) return;
  (void)hipEventRecord(c->ev[k], c->stream);
  c->marks |= 1 << k;
}

}  // namespace

namespace {

int fail(rmsd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(rmsd_ctx* c, hipError_t e, const char* where) {
  return fail(c, RMSD_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
  } while (0)

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
std::vector<int2> grouped_tiles(int nbr, int nbc, int tri, int tr = TILE, int tc = TILE) {
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

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Per-set operands of the pairwise contraction.
struct Packed {
  void* data;
  double* g;
  double* u;
};

int prepare_set(rmsd_ctx* c, const FrameSetDev& fs, const double* cmass, const double* smass,
                int prec, int Kp, const double* T, const double* gT, const char* tag,
                Packed* out) {
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
  double* T = nullptr;
  if (prec != PREC_F64) {
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

  Packed pa, pb;
  if ((st = prepare_set(c, A, amass, wmass, prec, Kp, T, gT, "A", &pa)) != RMSD_OK) return st;
  if (tri) {
    pb = pa;
  } else if ((st = prepare_set(c, B, bmass, wmass, prec, Kp, T, gT, "B", &pb)) != RMSD_OK) {
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
  mark(c, 1);
  HIP_TRY(c, rmsd_launch_pairwise(prec, p, c->stream));
  mark(c, 2);
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

// ---------------------------------------------------------------------------
// Batched matrices.
struct rmsd_batch {
  rmsd_ctx* ctx = nullptr;
  int mode = 0, prec = 0, use_mass = 0, flags = 0;
  std::vector<rmsd_frames> sets;
  std::vector<float*> outs;
  void* mem = nullptr;  // every device buffer of the plan, one allocation
  // per set (device pointers into mem)
  std::vector<void*> packed;
  std::vector<double*> g, u, consts, T;
  std::vector<int> Kp;
  std::vector<int2*> nofit_tiles;
  std::vector<int> nofit_ntiles;
  // batched per-set stage: one setup and one prep launch for all sets
  SetupArgs* dsetup = nullptr;
  PrepArgs* dprep = nullptr;
  int* dfbase = nullptr;
  int nactive = 0, total_frames = 0;
  // batched kernel
  PairArgs* dargs = nullptr;
  int4* dtiles[2] = {nullptr, nullptr};  // per tile side (64, 32)
  int* dqbeg[2] = {nullptr, nullptr};
  int maxq[2] = {0, 0};
  int ntiles[2] = {0, 0};
  int* counters = nullptr;
  int nq = 8;
  int variant = 0;
  double tune_ms[4] = {-1, -1, -1, -1};
  // graph
  hipStream_t gstream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

// Relative cost of one tile: its atom chunks plus the epilogue (QCP over the
// tile's pairs) in units of one chunk, measured on the 50k x 1200 matrix.
double tile_cost(int prec, int KB) {
  const double epi = prec == PREC_BF16X3 ? 6.0 : 4.0;
  return KB + epi;
}

// Tile queues for tile side S: sets sorted by cost per tile (descending, ties
// by set index), each set's triangle tiles in grouped order, the concatenated
// list cut into nq contiguous runs of equal total cost.
void build_queues(const rmsd_batch& b, int tr, int tc, std::vector<int4>* tiles,
                  std::vector<int>* qbeg, int* maxq) {
  const int ns = (int)b.sets.size();
  const int CH = chunk_atoms(b.prec);
  std::vector<int> order(ns);
  for (int i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b.Kp[x] / CH > b.Kp[y] / CH;
  });
  std::vector<double> cost;
  tiles->clear();
  for (int s : order) {
    const int F = b.sets[s].nframes;
    if (F < 2) continue;
    const int nbr = (F + tr - 1) / tr, nbc = (F + tc - 1) / tc;
    const double c = tile_cost(b.prec, b.Kp[s] / CH);
    for (int2 t : grouped_tiles(nbr, nbc, 1, tr, tc)) {
      tiles->push_back(make_int4(s, t.x, t.y, 0));
      cost.push_back(c);
    }
  }
  double total = 0.0;
  for (double c : cost) total += c;
  qbeg->assign(b.nq + 1, 0);
  double run = 0.0;
  int q = 1;
  for (size_t t = 0; t < cost.size() && q < b.nq; ++t) {
    run += cost[t];
    while (q < b.nq && run >= total * q / b.nq) (*qbeg)[q++] = (int)t + 1;
  }
  while (q <= b.nq) (*qbeg)[q++] = (int)cost.size();
  (*qbeg)[b.nq] = (int)cost.size();
  *maxq = 0;
  for (int k = 0; k < b.nq; ++k) *maxq = std::max(*maxq, (*qbeg)[k + 1] - (*qbeg)[k]);
}

// Enqueues one pass of the batch (per-set preparation, then the tiles) with
// kernel variant v.  Launches only: safe under stream capture.
int batch_enqueue(rmsd_batch* b, int v, hipStream_t st) {
  rmsd_ctx* c = b->ctx;
  const int ns = (int)b->sets.size();
  if (b->mode == RMSD_MODE_NOFIT) {
    for (int s = 0; s < ns; ++s) {
      const rmsd_frames& f = b->sets[s];
      if (f.nframes < 2) continue;
      const FrameSetDev fs = to_dev(f);
      const double* m = b->use_mass ? f.mass : nullptr;
      double* M = b->consts[s];
      SetupArgs su{};
      su.fs = fs;
      su.smass = m;
      su.total_mass = M;
      HIP_TRY(c, rmsd_launch_setup(su, st));
      const int nt = b->nofit_ntiles[s];
      HIP_TRY(c, rmsd_launch_pair_nofit_tiles(fs, fs, m, M, b->nofit_tiles[s], nt,
                                              (nt + NXCD - 1) / NXCD, NXCD, 1, b->outs[s], st));
    }
    return RMSD_OK;
  }
  // Every set's template / total weight, then every set's frames: two launches.
  HIP_TRY(c, rmsd_launch_setup_batch(b->dsetup, b->nactive, st));
  HIP_TRY(c, rmsd_launch_prep_batch(b->prec, b->dprep, b->dfbase, b->nactive, b->total_frames,
                                    st));
  mark(c, 1);
  const int sh = v >> 1;
  BatchArgs ba;
  ba.sets = b->dargs;
  ba.tiles = b->dtiles[sh];
  ba.qbeg = b->dqbeg[sh];
  ba.nq = b->nq;
  ba.maxq = b->maxq[sh];
  ba.counters = b->counters;
  if (b->ntiles[sh] == 0) return RMSD_OK;
  if (v & 1) HIP_TRY(c, rmsd_launch_batch_reset(ba, st));
  HIP_TRY(c, rmsd_launch_pairwise_batch(b->prec, v, ba, st));
  return RMSD_OK;
}

int batch_capture(rmsd_batch* b) {
  rmsd_ctx* c = b->ctx;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  HIP_TRY(c, hipStreamBeginCapture(b->gstream, hipStreamCaptureModeRelaxed));
  const bool prof = c->profiling;
  c->profiling = false;  // no event records inside the graph
  int st = batch_enqueue(b, b->variant, b->gstream);
  c->profiling = prof;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->gstream, &g);
  if (st != RMSD_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  b->graph = g;
  HIP_TRY(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return RMSD_OK;
}

int batch_run_on(rmsd_batch* b, int v) {
  rmsd_ctx* c = b->ctx;
  c->marks = 0;
  mark(c, 0);
  if (b->exec && v == b->variant) {
    mark(c, 1);
    HIP_TRY(c, hipEventRecord(b->ev_in, c->stream));
    HIP_TRY(c, hipStreamWaitEvent(b->gstream, b->ev_in, 0));
    HIP_TRY(c, hipGraphLaunch(b->exec, b->gstream));
    HIP_TRY(c, hipEventRecord(b->ev_out, b->gstream));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, b->ev_out, 0));
  } else {
    int st = batch_enqueue(b, v, c->stream);
    if (st != RMSD_OK) return st;
    if (b->mode == RMSD_MODE_NOFIT) mark(c, 1);
  }
  mark(c, 2);
  mark(c, 3);
  return RMSD_OK;
}

void batch_free(rmsd_batch* b) {
  if (!b) return;
  if (b->ctx) {
    b->ctx->plans.erase(b);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
  }
  if (b->gstream) (void)hipStreamSynchronize(b->gstream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  if (b->mem) (void)hipFree(b->mem);
  delete b;
}

}  // namespace

extern "C" {

int rmsd_batch_create(rmsd_ctx* ctx, const rmsd_frames* sets, int nsets, int mode, int use_mass,
                      int precision, float* const* outs, int flags, rmsd_batch** plan) {
  if (!ctx || !plan) return RMSD_ERR_ARG;
  *plan = nullptr;
  if (nsets < 0 || (nsets > 0 && (!sets || !outs)))
    return fail(ctx, RMSD_ERR_ARG, "sets and outs must hold nsets entries");
  if (mode != RMSD_MODE_FIT && mode != RMSD_MODE_NOFIT && mode != RMSD_MODE_QUAT)
    return fail(ctx, RMSD_ERR_ARG, "unknown mode");
  if (precision < RMSD_PREC_FP64 || precision > RMSD_PREC_BF16X3)
    return fail(ctx, RMSD_ERR_ARG, "unknown precision");
  if (flags & ~(RMSD_BATCH_GRAPH | RMSD_BATCH_AUTOTUNE))
    return fail(ctx, RMSD_ERR_ARG, "unknown batch flags");
  int st;
  for (int s = 0; s < nsets; ++s) {
    if ((st = check_frames(ctx, &sets[s], "set")) != RMSD_OK) return st;
    if (use_mass && !sets[s].mass) return fail(ctx, RMSD_ERR_ARG, "use_mass needs masses for every set");
    if (!outs[s] && sets[s].nframes > 1) return fail(ctx, RMSD_ERR_ARG, "an output is NULL");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  rmsd_batch* b = new rmsd_batch();
  b->ctx = ctx;
  ctx->plans.insert(b);
  b->mode = mode;
  b->prec = mode == RMSD_MODE_NOFIT ? RMSD_PREC_FP64 : precision;
  b->use_mass = use_mass;
  b->flags = flags;
  b->sets.assign(sets, sets + nsets);
  b->outs.assign(outs, outs + nsets);
  const int CH = chunk_atoms(b->prec);
  // Frames padded to whole tiles of every batch shape (rows 64, columns up to
  // rmsd_pairwise_tile_cols).
  const int pad = std::lcm(TILE, rmsd_pairwise_tile_cols(b->prec));
  const size_t esz = b->prec == PREC_F64 ? 8 : (b->prec == PREC_F32 ? 4 : 2);
  // Layout of the single allocation.
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t off = 0;
  std::vector<size_t> o_packed(nsets), o_g(nsets), o_u(nsets), o_c(nsets), o_T(nsets), o_nt(nsets);
  std::vector<std::vector<int2>> nt_host(nsets);
  b->Kp.resize(nsets);
  for (int s = 0; s < nsets; ++s) {
    const int F = sets[s].nframes;
    const int Fp = (int)round_up(F, pad);
    b->Kp[s] = (int)round_up(sets[s].nsel, CH);
    o_c[s] = off;
    off += al(sizeof(double) * 16);
    if (F < 2) continue;
    if (mode == RMSD_MODE_NOFIT) {
      const int nb = (F + TILE - 1) / TILE;
      nt_host[s] = grouped_tiles(nb, nb, 1);
      o_nt[s] = off;
      off += al(sizeof(int2) * nt_host[s].size());
      continue;
    }
    o_packed[s] = off;
    off += al((size_t)Fp * b->Kp[s] * 3 * esz * chunk_parts(b->prec));
    o_g[s] = off;
    off += al(sizeof(double) * Fp);
    o_u[s] = off;
    off += al(sizeof(double) * 9 * Fp);
    o_T[s] = off;
    off += al(sizeof(double) * 3 * b->Kp[s]);
  }
  std::vector<int4> th[2];
  std::vector<int> qh[2];
  size_t o_args = 0, o_tiles[2] = {0, 0}, o_q[2] = {0, 0}, o_cnt = 0;
  size_t o_setup = 0, o_prep = 0, o_fbase = 0;
  int nactive = 0;
  for (int s = 0; s < nsets; ++s) nactive += (mode != RMSD_MODE_NOFIT && sets[s].nframes >= 2);
  if (mode != RMSD_MODE_NOFIT) {
    o_setup = off;
    off += al(sizeof(SetupArgs) * std::max(1, nactive));
    o_prep = off;
    off += al(sizeof(PrepArgs) * std::max(1, nactive));
    o_fbase = off;
    off += al(sizeof(int) * (nactive + 1));
    for (int sh = 0; sh < 2; ++sh) {
      int tr = 0, tc = 0;
      rmsd_batch_tile_shape(b->prec, 2 * sh, &tr, &tc);
      build_queues(*b, tr, tc, &th[sh], &qh[sh], &b->maxq[sh]);
      b->ntiles[sh] = (int)th[sh].size();
      o_tiles[sh] = off;
      off += al(sizeof(int4) * std::max<size_t>(1, th[sh].size()));
      o_q[sh] = off;
      off += al(sizeof(int) * (b->nq + 1));
    }
    o_args = off;
    off += al(sizeof(PairArgs) * std::max(1, nsets));
    o_cnt = off;
    off += al(sizeof(int) * b->nq);
  }
  if (hipMalloc(&b->mem, std::max<size_t>(off, 256)) != hipSuccess) {
    b->mem = nullptr;
    batch_free(b);
    return fail(ctx, RMSD_ERR_ALLOC, "cannot allocate " + std::to_string(off) + " bytes for the batch");
  }
  char* base = (char*)b->mem;
  b->packed.assign(nsets, nullptr);
  b->g.assign(nsets, nullptr);
  b->u.assign(nsets, nullptr);
  b->consts.assign(nsets, nullptr);
  b->T.assign(nsets, nullptr);
  b->nofit_tiles.assign(nsets, nullptr);
  b->nofit_ntiles.assign(nsets, 0);
  std::vector<PairArgs> args(nsets);
  for (int s = 0; s < nsets; ++s) {
    b->consts[s] = (double*)(base + o_c[s]);
    if (sets[s].nframes < 2) continue;
    if (mode == RMSD_MODE_NOFIT) {
      b->nofit_tiles[s] = (int2*)(base + o_nt[s]);
      b->nofit_ntiles[s] = (int)nt_host[s].size();
      continue;
    }
    b->packed[s] = base + o_packed[s];
    b->g[s] = (double*)(base + o_g[s]);
    b->u[s] = (double*)(base + o_u[s]);
    b->T[s] = b->prec == PREC_F64 ? nullptr : (double*)(base + o_T[s]);
    PairArgs& p = args[s];
    memset(&p, 0, sizeof(p));
    p.A = p.B = b->packed[s];
    p.KB = b->Kp[s] / CH;
    p.gA = p.gB = b->g[s];
    p.uA = p.uB = b->u[s];
    p.P = b->consts[s] + 1;
    p.total_mass = b->consts[s];
    p.FA = p.FB = sets[s].nframes;
    p.tri = 1;
    p.clamp = mode == RMSD_MODE_FIT ? 1 : 0;
    p.out = outs[s];
  }
  int rc = RMSD_OK;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (rc != RMSD_OK || bytes == 0) return;
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "batch upload");
  };
  for (int s = 0; s < nsets; ++s)
    if (b->nofit_tiles[s]) up(b->nofit_tiles[s], nt_host[s].data(), sizeof(int2) * nt_host[s].size());
  if (mode != RMSD_MODE_NOFIT) {
    // Per-set setup / prep arguments of the sets with pairs, and the prefix of
    // their padded frame counts (one prep workgroup per padded frame).
    std::vector<SetupArgs> su;
    std::vector<PrepArgs> pa;
    std::vector<int> fbase(1, 0);
    for (int s = 0; s < nsets; ++s) {
      const rmsd_frames& f = sets[s];
      if (f.nframes < 2) continue;
      const double* m = use_mass ? f.mass : nullptr;
      double* M = b->consts[s];
      SetupArgs x{};
      x.fs = to_dev(f);
      x.f0 = 0;
      x.cmass = m;
      x.smass = m;
      x.Kp = b->Kp[s];
      x.T = b->T[s];
      x.P = M + 1;
      x.gT = M + 10;
      x.total_mass = M;
      su.push_back(x);
      PrepArgs a;
      memset(&a, 0, sizeof(a));
      a.fs = to_dev(f);
      a.cmass = m;
      a.smass = m;
      a.prec = b->prec;
      a.Kp = b->Kp[s];
      a.KB = b->Kp[s] / CH;
      a.Fp = (int)round_up(f.nframes, pad);
      a.packed = b->packed[s];
      a.g = b->g[s];
      a.u = b->u[s];
      a.T = b->T[s];
      a.gT = M + 10;
      pa.push_back(a);
      fbase.push_back(fbase.back() + a.Fp);
    }
    b->nactive = (int)su.size();
    b->total_frames = fbase.back();
    b->dsetup = (SetupArgs*)(base + o_setup);
    b->dprep = (PrepArgs*)(base + o_prep);
    b->dfbase = (int*)(base + o_fbase);
    up(b->dsetup, su.data(), sizeof(SetupArgs) * su.size());
    up(b->dprep, pa.data(), sizeof(PrepArgs) * pa.size());
    up(b->dfbase, fbase.data(), sizeof(int) * fbase.size());
  }
  if (mode != RMSD_MODE_NOFIT) {
    b->dargs = (PairArgs*)(base + o_args);
    up(b->dargs, args.data(), sizeof(PairArgs) * nsets);
    for (int sh = 0; sh < 2; ++sh) {
      b->dtiles[sh] = (int4*)(base + o_tiles[sh]);
      b->dqbeg[sh] = (int*)(base + o_q[sh]);
      up(b->dtiles[sh], th[sh].data(), sizeof(int4) * th[sh].size());
      up(b->dqbeg[sh], qh[sh].data(), sizeof(int) * qh[sh].size());
    }
    b->counters = (int*)(base + o_cnt);
  }
  if (rc != RMSD_OK) {
    batch_free(b);
    return rc;
  }
  // Tuner: time every kernel variant on the plan's own data, keep the fastest.
  if ((flags & RMSD_BATCH_AUTOTUNE) && mode != RMSD_MODE_NOFIT) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      batch_free(b);
      return fail(ctx, RMSD_ERR_DEVICE, "hipEventCreate");
    }
    double best = 1e300;
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    for (int v = 0; v < rmsd_batch_variant_count() && rc == RMSD_OK; ++v) {
      rc = batch_enqueue(b, v, ctx->stream);  // warm-up
      double ms = 1e300;
      for (int r = 0; r < 2 && rc == RMSD_OK; ++r) {
        float t = 0.f;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = RMSD_ERR_DEVICE;
        if (rc == RMSD_OK) rc = batch_enqueue(b, v, ctx->stream);
        if (rc == RMSD_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                              hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
          rc = fail(ctx, RMSD_ERR_DEVICE, "batch tuner timing");
        ms = std::min(ms, (double)t);
      }
      b->tune_ms[v] = ms;
      if (ms < best) {
        best = ms;
        b->variant = v;
      }
    }
    ctx->profiling = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != RMSD_OK) {
      batch_free(b);
      return rc;
    }
  }
  if (flags & RMSD_BATCH_GRAPH) {
    if (hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
      batch_free(b);
      return fail(ctx, RMSD_ERR_DEVICE, "cannot create the batch stream / events");
    }
    if ((rc = batch_capture(b)) != RMSD_OK) {
      batch_free(b);
      return rc;
    }
  }
  *plan = b;
  return RMSD_OK;
}

int rmsd_batch_run(rmsd_batch* plan) {
  if (!plan) return RMSD_ERR_ARG;
  HIP_TRY(plan->ctx, hipSetDevice(plan->ctx->device));
  return batch_run_on(plan, plan->variant);
}

int rmsd_batch_set_variant(rmsd_batch* plan, int variant) {
  if (!plan) return RMSD_ERR_ARG;
  if (variant < 0 || variant >= rmsd_batch_variant_count())
    return fail(plan->ctx, RMSD_ERR_ARG, "unknown batch kernel variant");
  HIP_TRY(plan->ctx, hipSetDevice(plan->ctx->device));
  if (variant == plan->variant) return RMSD_OK;
  plan->variant = variant;
  if (plan->gstream) return batch_capture(plan);
  return RMSD_OK;
}

int rmsd_batch_info(const rmsd_batch* plan, int64_t* info, double* tune_ms) {
  if (!plan) return RMSD_ERR_ARG;
  if (info) {
    const int sh = plan->variant >> 1;
    info[0] = plan->variant;
    int tr = TILE, tc = TILE;
    if (plan->mode != RMSD_MODE_NOFIT) rmsd_batch_tile_shape(plan->prec, plan->variant, &tr, &tc);
    info[1] = tr;
    int64_t nt = 0;
    if (plan->mode == RMSD_MODE_NOFIT)
      for (int k : plan->nofit_ntiles) nt += k;
    else
      nt = plan->ntiles[sh];
    info[2] = nt;
    info[3] = plan->exec != nullptr;
    info[4] = plan->nq;
    info[5] = plan->mode == RMSD_MODE_NOFIT ? 0 : plan->maxq[sh];
  }
  if (tune_ms)
    for (int v = 0; v < 4; ++v) tune_ms[v] = plan->tune_ms[v];
  return RMSD_OK;
}

void rmsd_batch_destroy(rmsd_batch* plan) { batch_free(plan); }

int rmsd_pairwise_batch(rmsd_ctx* ctx, const rmsd_frames* sets, int nsets, int mode, int use_mass,
                        int precision, float* const* outs) {
  rmsd_batch* b = nullptr;
  int st = rmsd_batch_create(ctx, sets, nsets, mode, use_mass, precision, outs, 0, &b);
  if (st != RMSD_OK) return st;
  st = rmsd_batch_run(b);
  rmsd_batch_destroy(b);  // waits for the pass to finish before freeing
  return st;
}

}  // extern "C"
