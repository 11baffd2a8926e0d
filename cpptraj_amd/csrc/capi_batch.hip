// capi_batch.hip -- C ABI of batched plans (rmsd_batch_*): many frame sets in
// one graph-captured pass of the setup, prep and batched contraction kernels.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>

#include "capi_internal.hpp"

using namespace rmsd_capi;

struct rmsd_batch {
  rmsd_ctx* ctx = nullptr;
  int mode = 0, prec = 0, use_mass = 0, flags = 0;
  std::vector<rmsd_frames> sets;
  std::vector<float*> outs;
  void* mem = nullptr;  // every device buffer of the plan, one allocation
  // per set (device pointers into mem)
  std::vector<void*> packed;
  std::vector<double*> g, u, consts, T;
  std::vector<double*> nrm, cen, part;  // MFMA forms: bound norms, centroids, template samples
  std::vector<int> Kp;
  std::vector<int2*> nofit_tiles;
  std::vector<int> nofit_ntiles;
  // batched per-set stage: one setup and one prep launch for all sets
  SetupArgs* dsetup = nullptr;
  PrepArgs* dprep = nullptr;
  int* dfbase = nullptr;
  int nactive = 0, total_frames = 0;
  int nquad = 0, quad_frames = 0;  // prep: the leading sets that allow quads
  // mean-structure templates (MFMA forms): sample prefix over the active sets
  int* dsbase = nullptr;
  int total_samples = 0, maxk = 0;
  // error bound (MFMA forms, context tolerance != 0 at creation): room to list
  // every wave of either shape's tile list and its 64 lane masks, {largest
  // kept bound, pairs recomputed, (listed waves, work counter)}, and the
  // per-set arguments of the FP64 recomputation
  bool bounded = false;
  double tol = 0.0;
  int* twlist = nullptr;
  unsigned int* tnflag = nullptr;
  unsigned int* tmasks = nullptr;
  unsigned long long* stats = nullptr;
  FallbackArgs* dfall = nullptr;
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

constexpr size_t QUEUE_BLOCK = 64;  // tiles per block dealt to a queue

// Tile queues for tile side S: sets sorted by atom chunks per tile (the cost
// of a tile; descending, ties by set index), each set's triangle tiles in
// grouped order, and the concatenated list dealt to the nq queues (one per
// XCD) in blocks of QUEUE_BLOCK consecutive tiles, round-robin.  Every queue
// then holds the same mix of heavy and light tiles, with no model of their
// cost, while the tiles a queue's concurrent workgroups run stay neighbours in
// one set (sharing row and column blocks in that XCD's L2).  Contiguous runs of
// equal modelled cost, one per queue, measured 99.3 / 84.3 ms (grid /
// persistent form) against 77.7 / 82.2 ms on the 64-trajectory batch.
void build_queues(const rmsd_batch& b, int tr, int tc, std::vector<int4>* tiles,
                  std::vector<int>* qbeg, int* maxq) {
  const int ns = (int)b.sets.size();
  const int CH = chunk_atoms(b.prec);
  std::vector<int> order(ns);
  for (int i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b.Kp[x] / CH > b.Kp[y] / CH;
  });
  tiles->clear();
  for (int s : order) {
    const int F = b.sets[s].nframes;
    if (F < 2) continue;
    const int nbr = (F + tr - 1) / tr, nbc = (F + tc - 1) / tc;
    for (int2 t : grouped_tiles(nbr, nbc, 1, tr, tc)) tiles->push_back(make_int4(s, t.x, t.y, 0));
  }
  std::vector<int4> dealt;
  dealt.reserve(tiles->size());
  qbeg->assign(b.nq + 1, 0);
  for (int q = 0; q < b.nq; ++q) {
    (*qbeg)[q] = (int)dealt.size();
    for (size_t b0 = (size_t)q * QUEUE_BLOCK; b0 < tiles->size(); b0 += (size_t)b.nq * QUEUE_BLOCK)
      for (size_t t = b0; t < std::min(tiles->size(), b0 + QUEUE_BLOCK); ++t)
        dealt.push_back((*tiles)[t]);
  }
  (*qbeg)[b.nq] = (int)dealt.size();
  *tiles = dealt;
  // w: the tile's position in the list (its flag of the error bound).
  for (size_t t = 0; t < tiles->size(); ++t) (*tiles)[t].w = (int)t;
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
  // Every set's template / total weight (the mean structure of its samples
  // for the MFMA forms), then every set's frames.
  HIP_TRY(c, rmsd_launch_setup_batch(b->dsetup, b->nactive, st));
  if (b->prec != PREC_F64)
    HIP_TRY(c, rmsd_launch_template_mean_batch(b->dsetup, b->dsbase, b->nactive,
                                               b->total_samples, b->maxk, st));
  HIP_TRY(c, rmsd_launch_prep_batch(b->prec, b->dprep, b->dfbase, b->nactive, b->nquad,
                                    b->quad_frames, b->total_frames, st));
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
  const WaveLayout lay = rmsd_batch_layout(b->prec, v);
  if (b->bounded)  // a fill kernel (no memset nodes in the captured graph)
    HIP_TRY(c, rmsd_launch_fill_int((int*)b->stats, 6, 0, st));
  if (v & 1) HIP_TRY(c, rmsd_launch_batch_reset(ba, st));
  HIP_TRY(c, rmsd_launch_pairwise_batch(b->prec, v, ba, st));
  if (b->bounded && b->tol > 0.0)  // tol < 0: report only
    HIP_TRY(c, rmsd_launch_pair_fallback_batch(b->dfall, b->dtiles[sh], b->ntiles[sh], lay,
                                               b->twlist, b->tnflag, b->tmasks, st));
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
  c->report = {b->prec == PREC_F64 ? 0 : (b->bounded ? 1 : 2), b->stats};
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

}  // namespace

void rmsd_capi::batch_free(rmsd_batch* b) {
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
  // those of shape 0).
  int tr0 = 0, tc0 = 0;
  rmsd_batch_tile_shape(b->prec, 0, &tr0, &tc0);
  const int pad = std::lcm(TILE, tc0);
  const size_t esz = b->prec == PREC_F64 ? 8 : (b->prec == PREC_F32 ? 4 : 2);
  // Layout of the single allocation.
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t off = 0;
  std::vector<size_t> o_packed(nsets), o_g(nsets), o_u(nsets), o_c(nsets), o_T(nsets), o_nt(nsets);
  std::vector<size_t> o_nrm(nsets), o_cen(nsets), o_part(nsets);
  std::vector<int> nsamp(nsets, 0);
  const bool mixed = mode != RMSD_MODE_NOFIT && b->prec != PREC_F64;
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
    if (mixed) {
      nsamp[s] = std::min(TEMPLATE_SAMPLES, F);
      o_nrm[s] = off;
      off += al(sizeof(double) * 4 * Fp);
      o_cen[s] = off;
      off += al(sizeof(double) * 4 * Fp);
      o_part[s] = off;
      off += al(sizeof(double) * 3 * b->Kp[s] * nsamp[s]);
    }
  }
  std::vector<int4> th[2];
  std::vector<int> qh[2];
  size_t o_args = 0, o_tiles[2] = {0, 0}, o_q[2] = {0, 0}, o_cnt = 0;
  size_t o_setup = 0, o_prep = 0, o_fbase = 0, o_sbase = 0, o_wlist = 0, o_stats = 0;
  size_t o_fall = 0, o_masks = 0;
  size_t nwaves = 0;
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
    if (mixed) {
      o_sbase = off;
      off += al(sizeof(int) * (nactive + 1));
      for (int sh = 0; sh < 2; ++sh) {
        const WaveLayout L = rmsd_batch_layout(b->prec, 2 * sh);
        nwaves = std::max<size_t>(nwaves, (size_t)b->ntiles[sh] * L.NWR * L.NWC);
      }
      nwaves = std::max<size_t>(nwaves, 1);
      o_wlist = off;
      off += al(sizeof(int) * nwaves);
      o_masks = off;
      off += al(sizeof(unsigned int) * 64 * nwaves);
      o_stats = off;
      off += al(3 * sizeof(unsigned long long));
      o_fall = off;
      off += al(sizeof(FallbackArgs) * std::max(1, nsets));
    }
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
  b->nrm.assign(nsets, nullptr);
  b->cen.assign(nsets, nullptr);
  b->part.assign(nsets, nullptr);
  b->bounded = mixed && ctx->tol != 0.0;
  b->tol = ctx->tol;
  if (mixed) {
    b->twlist = (int*)(base + o_wlist);
    b->tmasks = (unsigned int*)(base + o_masks);
    b->stats = (unsigned long long*)(base + o_stats);
    b->tnflag = (unsigned int*)(b->stats + 2);
    b->dfall = (FallbackArgs*)(base + o_fall);
  }
  std::vector<FallbackArgs> fall(nsets);
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
    if (mixed) {
      b->nrm[s] = (double*)(base + o_nrm[s]);
      b->cen[s] = (double*)(base + o_cen[s]);
      b->part[s] = (double*)(base + o_part[s]);
    }
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
    if (b->bounded) {
      p.nrmA = p.nrmB = b->nrm[s];
      p.ke0 = rmsd_bound_coef(b->prec, p.KB).e0;
      p.keep = rmsd_keep_split(b->tol);
      p.wlist = b->twlist;
      p.nflag = b->tnflag;
      p.masks = b->tmasks;
      p.stats = b->stats;
      FallbackArgs& f = fall[s];
      memset(&f, 0, sizeof(f));
      f.A = f.B = to_dev(sets[s]);
      f.smass = use_mass ? sets[s].mass : nullptr;
      f.cenA = f.cenB = b->cen[s];
      f.gA = f.gB = b->g[s];
      f.total_mass = b->consts[s];
      f.FA = f.FB = sets[s].nframes;
      f.tri = 1;
      f.clamp = p.clamp;
      f.out = outs[s];
      f.stats = b->stats;
    }
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
    std::vector<int> fbase(1, 0), sbase(1, 0);
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
      x.nsamp = nsamp[s];
      x.part = b->part[s];
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
      a.nrm = b->nrm[s];
      a.cen = b->cen[s];
      a.total_mass = M;
      a.bound = rmsd_bound_coef(b->prec, a.KB);
      pa.push_back(a);
      fbase.push_back(fbase.back() + a.Fp);
      sbase.push_back(sbase.back() + nsamp[s]);
      b->maxk = std::max(b->maxk, 3 * b->Kp[s]);
    }
    // Prep order: the sets that allow 16-byte loads of four atoms first (their
    // launch compiles the quad path; the rest run an atom-by-atom kernel).
    {
      std::vector<PrepArgs> q, r;
      for (const PrepArgs& x : pa) (quad_layout(x.fs) ? q : r).push_back(x);
      b->nquad = (int)q.size();
      pa = q;
      pa.insert(pa.end(), r.begin(), r.end());
      fbase.assign(1, 0);
      for (const PrepArgs& x : pa) fbase.push_back(fbase.back() + x.Fp);
      b->quad_frames = fbase[b->nquad];
    }
    b->nactive = (int)su.size();
    b->total_frames = fbase.back();
    b->dsetup = (SetupArgs*)(base + o_setup);
    b->dprep = (PrepArgs*)(base + o_prep);
    b->dfbase = (int*)(base + o_fbase);
    up(b->dsetup, su.data(), sizeof(SetupArgs) * su.size());
    up(b->dprep, pa.data(), sizeof(PrepArgs) * pa.size());
    up(b->dfbase, fbase.data(), sizeof(int) * fbase.size());
    if (mixed) {
      b->total_samples = sbase.back();
      b->dsbase = (int*)(base + o_sbase);
      up(b->dsbase, sbase.data(), sizeof(int) * sbase.size());
      if (b->bounded) up(b->dfall, fall.data(), sizeof(FallbackArgs) * nsets);
    }
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
