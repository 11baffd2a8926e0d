// k_pairwise.hip -- all-pairs best-fit RMSD on MFMA (gfx950).
//
// Replaces the O(F^2 N) loop nest of Analysis_Rms2d::Calculate_2D
// (src/Analysis_Rms2d.cpp:247-287) and Cluster::MetricArray::calcFrameDistances
// (src/Cluster/MetricArray.cpp:776-797).  The 3x3 covariances of all frame pairs
// are one contraction  [3F x N] x [N x 3F]: for a 16x16 block of frame pairs and
// a coordinate pair (c, d) the block  S_cd(i, j) = sum_k X_i[k][c] Y_j[k][d]  is
// one chain of MFMA accumulations over the atoms k.  Each wave owns
// (16*WR) x (16*WC) frame pairs and all nine (c, d) blocks, so every lane ends
// the contraction holding the complete 3x3 covariance of each of its pairs and
// runs the FP64 QCP epilogue in registers (no cross-lane shuffle).
//
// Operands come from the fragment-packed layout written by k_prep.hip: for a
// block of 16 frames, one "chunk" of atoms and one coordinate, the 64 lanes of a
// wave read 64 consecutive 16-byte pieces (one dwordx4 load per lane, 1 KiB per
// wave, fully coalesced).  Lane l holds frame (l % 16) of the block and atoms
// (l / 16) * CHUNK/4 ... of the chunk, so A (rows) and B (columns) share one
// atom order and the contraction is an exact permutation of the atom sum.
//
// Precisions:
//   PREC_F64   - centred, sqrt(w)-scaled FP64 coordinates, v_mfma_f64_16x16x4f64,
//                FP64 accumulation.  8 atoms per chunk.
//   PREC_F32   - per-frame pre-aligned deviations D from a template, FP32,
//                v_mfma_f32_16x16x4f32, FP32 accumulation.  16 atoms per chunk.
//   PREC_BF16X3- the same deviations split as hi + lo bfloat16,
//                hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16, FP32
//                accumulation.  32 atoms per chunk.
// For the two MFMA-reduced forms the full covariance is reassembled in FP64:
//   S(i,j) = P + u_j + u_i^T + D_i^T D_j,  P = T^T T,  u_f = T^T D_f
// (T the template, all per-frame terms FP64), then E0 and QCP run in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) double f64x4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

template <int PREC> struct PrecTraits;
template <> struct PrecTraits<PREC_F64> {
  typedef f64x4 acc_t;
  static constexpr int PARTS = 1;
};
template <> struct PrecTraits<PREC_F32> {
  typedef f32x4 acc_t;
  static constexpr int PARTS = 1;
};
template <> struct PrecTraits<PREC_BF16X3> {
  typedef f32x4 acc_t;
  static constexpr int PARTS = 2;
};

// One 16-byte piece per lane (cached: row and column blocks are re-read by
// neighbouring workgroups through L2).
__device__ __forceinline__ u32x4 ld16(const u32x4* p) { return *gp(p); }

template <int PREC>
__device__ __forceinline__ void mma(typename PrecTraits<PREC>::acc_t& acc, const u32x4* a,
                                    const u32x4* b);

template <>
__device__ __forceinline__ void mma<PREC_F64>(f64x4& acc, const u32x4* a, const u32x4* b) {
  const double2 av = __builtin_bit_cast(double2, a[0]);
  const double2 bv = __builtin_bit_cast(double2, b[0]);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.y, acc, 0, 0, 0);
}

template <>
__device__ __forceinline__ void mma<PREC_F32>(f32x4& acc, const u32x4* a, const u32x4* b) {
  const f32x4 av = __builtin_bit_cast(f32x4, a[0]);
  const f32x4 bv = __builtin_bit_cast(f32x4, b[0]);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc, 0, 0, 0);
}

template <>
__device__ __forceinline__ void mma<PREC_BF16X3>(f32x4& acc, const u32x4* a, const u32x4* b) {
  const bf16x8 ahi = __builtin_bit_cast(bf16x8, a[0]);
  const bf16x8 alo = __builtin_bit_cast(bf16x8, a[1]);
  const bf16x8 bhi = __builtin_bit_cast(bf16x8, b[0]);
  const bf16x8 blo = __builtin_bit_cast(bf16x8, b[1]);
  // Small terms first so that they are not absorbed by the large one.
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc, 0, 0, 0);
}

template <int PREC>
__device__ __forceinline__ double acc_elem(const typename PrecTraits<PREC>::acc_t& a, int v) {
  return (double)a[v];
}

// Fragment pointer for (frame block fb, chunk kb, coordinate c, part p).
__device__ __forceinline__ const u32x4* frag_ptr(const u32x4* base, int64_t fb, int kb,
                                                 int KB, int c, int p, int PARTS) {
  return base + ((((fb * KB + kb) * 3 + c) * PARTS + p) << 6);
}

// ---- FP64 epilogue --------------------------------------------------------------
// Reassembles each pair's 3x3 covariance (plus the FP64 template terms for the
// MFMA-reduced forms), forms E0 and runs QCP pair by pair in registers, with
// the per-frame terms read from memory (gA / uA / nrmA for row frames, gB / uB
// / nrmB for column frames).  (Grouping several pairs' Newton iterations, and
// staging the tile's per-frame terms in LDS, were both measured slower on the
// 50k-frame matrix.)
// Error bound (MFMA-reduced forms with a flag array).  One eps bounds the
// pairs of a lane's column j with its 8 rows (the largest row coefficient times
// j's, plus the FP64 term at the largest E0), and keep_limits turns it into a
// float interval of RMSDs whose stated bound is within tol: per pair two
// compares.  Pairs outside it are recorded in the lane's mask.  The 8 rows
// (4 (lane / 16) + v of two 16-frame blocks starting on a multiple of 32) are
// the same in every tile shape with WR = 2, so which pairs are recomputed --
// and every result -- does not depend on the shape.  Computed before the
// contraction, so that the latency of its loads overlaps the first chunk's.
template <int WC>
struct BoundLimits {
  bool on;
  float klo[WC], khi[WC], eps[WC];
};

template <int PREC, int WR, int WC>
__device__ __forceinline__ BoundLimits<WC> bound_limits(const PairArgs& args, int64_t fbA,
                                                        int64_t fbB, int lane) {
  static_assert(PREC == PREC_F64 || WR == 2, "the bound's row groups assume WR = 2");
  const double M = gp(args.total_mass)[0];
  const auto gA = gp(args.gA);
  const auto gB = gp(args.gB);
  BoundLimits<WC> L;
  float* klo = L.klo;
  float* khi = L.khi;
  float* eps = L.eps;
  L.on = PREC != PREC_F64 && args.wlist != nullptr && M >= RMSD_SMALL;
  if (L.on) {
    double csr = 0.0, gr = 0.0;
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = (fbA + r) * 16 + 4 * (lane >> 4) + v;
        if (i < args.FA) {
          csr = fmax(csr, gp(args.nrmA)[4 * i]);
          gr = fmax(gr, gA[i]);
        }
      }
#pragma unroll
    for (int q = 0; q < WC; ++q) {
      const int64_t j = (fbB + q) * 16 + (lane & 15);
      double csc = 0.0, gc = 0.0;
      if (j < args.FB) {
        csc = gp(args.nrmB)[4 * j];
        gc = gB[j];
      }
      eps[q] = __double2float_ru(csr * csc + args.ke0 / M * 0.5 * (gr + gc));
      rmsd_dev::keep_limits(eps[q], args.keep, &klo[q], &khi[q]);
    }
  }
  return L;
}

template <int PREC, int WR, int WC>
__device__ __forceinline__ void pair_epilogue(const PairArgs& args,
                                              typename PrecTraits<PREC>::acc_t (&acc)[WR][WC][3][3],
                                              int64_t fbA, int64_t fbB, int lane, int tile_id,
                                              int wave, int nwaves, const BoundLimits<WC>& L) {
  const auto gA = gp(args.gA);
  const auto gB = gp(args.gB);
  const auto uA = gp(args.uA);
  const auto uB = gp(args.uB);
  const double M = gp(args.total_mass)[0];
  const bool bounded = L.on;
  const float* klo = L.klo;
  const float* khi = L.khi;
  const float* eps = L.eps;
  unsigned int bits = 0;
  float kmin[WC], kmax[WC], omax[WC];
#pragma unroll
  for (int q = 0; q < WC; ++q) {
    kmin[q] = INFINITY;
    kmax[q] = 0.0f;
    omax[q] = -1.0f;  // >= 0 once a pair of the column is outside [klo, khi]
  }
  const double invM = 1.0 / M;
#pragma unroll
  for (int q = 0; q < WC; ++q) {
    const int64_t j = (fbB + q) * 16 + (lane & 15);  // column frame
    // uj: the column's template terms plus the constant P, added once per column
    double gj = 0.0, uj[9];
    const bool jok = j < args.FB;
    if (jok) {
      gj = gB[j];
      if (PREC != PREC_F64) {
#pragma unroll
        for (int k = 0; k < 9; ++k) uj[k] = gp(args.P)[k] + uB[9 * j + k];
      }
    }
#pragma unroll
    for (int r = 0; r < WR; ++r) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // Accumulator layout (checked on gfx950): FP32 16x16 blocks hold row
        // 4*(lane/16) + v in item v; the FP64 16x16x4 block holds row
        // (lane/16) + 4*v.  Columns are lane % 16 in both.
        const int rr = PREC == PREC_F64 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
        const int64_t i = (fbA + r) * 16 + rr;  // row frame
        if (!(jok && i < args.FA && (!args.tri || i < j))) continue;
        double S[9];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) S[3 * c + d] = acc_elem<PREC>(acc[r][q][c][d], v);
        if (PREC != PREC_F64) {
          const auto ui = uA + 9 * i;
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
              S[3 * c + d] += uj[3 * c + d] + ui[3 * d + c];
        }
        // Row frames are the reference in the triangle layout (the reference
        // stores element (nref, ntgt) with ntgt > nref) and the target in the
        // full layout; QCP takes R = sum w t r^T.
        double R[9];
        if (args.tri) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) R[3 * c + d] = S[3 * d + c];
        } else {
#pragma unroll
          for (int k = 0; k < 9; ++k) R[k] = S[k];
        }
        const double E0 = 0.5 * (gA[i] + gj);
        const int64_t idx = args.tri ? rmsd_dev::tri_index(args.FA, i, j) : i * args.FB + j;
        const float res = M < RMSD_SMALL ? -1.0f
                          : PREC == PREC_F64
                              ? (float)rmsd_dev::qcp_rmsd(R, E0, M, args.clamp)
                              : (float)rmsd_dev::qcp_rmsd_fast(R, E0, invM, args.clamp);
        gp(args.out)[idx] = res;
        if (bounded) {
          if (res >= klo[q] && res <= khi[q]) {  // NaN fails: recomputed in FP64
            kmin[q] = fminf(kmin[q], res);
            kmax[q] = fmaxf(kmax[q], res);
          } else {
            bits |= 1u << ((q * WR + r) * 4 + v);
            omax[q] = fmaxf(omax[q], res);
          }
        }
      }
    }
  }
  if (bounded) {
    // Flag and masks: one store per lane of a wave with a pair over tol.
    const int64_t w = (int64_t)tile_id * nwaves + wave;
    if (__ballot(bits != 0)) {
      gp(args.masks)[w * 64 + lane] = bits;
      if (lane == 0) gp(args.wlist)[atomicAdd(args.nflag, 1u)] = (int)w;
    }
    // The largest stated bound of the kept pairs (<= tol by construction).
    // Report only (tol = infinity): also the pairs below sqrt(eps)+.
    float b = 0.0f;
#pragma unroll
    for (int q = 0; q < WC; ++q) {
      if (kmax[q] > 0.0f)
        b = fmaxf(b, fminf(rmsd_dev::bound_kept(eps[q], kmin[q], kmax[q]), args.keep.tol));
      if (omax[q] >= 0.0f && args.keep.tol == INFINITY)
        b = fmaxf(b, rmsd_dev::bound_any(eps[q], omax[q]));
    }
    const unsigned int bw = rmsd_dev::wave_max_u32(__float_as_uint(b));  // b >= 0
    if (lane == 0 && bw != 0u)
      atomicMax(args.stats, (unsigned long long)__double_as_longlong((double)__uint_as_float(bw)));
  }
}

// ---- FP64 recomputation of flagged pairs -------------------------------------------
// The contraction lists every wave with a pair over the tolerance (wlist,
// nflag[0] entries) and leaves that wave's lane masks.  Here one wave
// recomputes one 16 x 16 block of frame pairs of a listed wave on
// v_mfma_f64_16x16x4f64: operands straight from the frames as stored, centred
// with the per-frame centroids of k_prep (the row side weighted), nine FP64
// accumulator blocks over all atoms, then the FP64 QCP of the block's flagged
// pairs.  Flagged pairs come in dense blocks (pairs of frames close to each
// other: near the diagonal of a trajectory's matrix), so whole blocks are
// computed; blocks without a flagged pair are skipped.  Waves take blocks from
// a counter (nflag[1]), so a fixed grid drains the list whatever its length.
constexpr int FALLBACK_THREADS = 256;
constexpr int FALLBACK_GRID = 2048;

// Four consecutive atoms k, ..., k + 3 of row frame fi and column frame fj
// (zero beyond the selection or the frame count) and their weights.
struct FallbackQuad {
  float x[4][3], y[4][3];
  double w[4];
};

// LAY != 0: both frame sets allow 16-byte loads of four atoms (quad_layout),
// used wherever the four atoms lie within the selection (k is a multiple of 4).
template <int LAY>
__device__ __forceinline__ void fallback_fetch(const FallbackArgs& a, int64_t fi, bool iok,
                                               int64_t fj, bool jok, int k, FallbackQuad& f) {
  if (LAY != 0 && k + 3 < a.A.nsel) {
    float x[3][4], y[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int s = 0; s < 4; ++s) x[c][s] = y[c][s] = 0.0f;
    if (iok) load_quad(a.A, fi, k >> 2, LAY, x[0], x[1], x[2]);
    if (jok) load_quad(a.B, fj, k >> 2, LAY, y[0], y[1], y[2]);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      f.w[s] = a.smass ? gp(a.smass)[k + s] : 1.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        f.x[s][c] = x[c][s];
        f.y[s][c] = y[c][s];
      }
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    f.x[s][0] = f.x[s][1] = f.x[s][2] = f.y[s][0] = f.y[s][1] = f.y[s][2] = 0.0f;
    f.w[s] = 0.0;
    if (k + s < a.A.nsel) {
      f.w[s] = a.smass ? gp(a.smass)[k + s] : 1.0;
      if (iok) load_atom(a.A, fi, k + s, f.x[s]);
      if (jok) load_atom(a.B, fj, k + s, f.y[s]);
    }
  }
}

template <int LAY>
__device__ __forceinline__ void fallback_block(const FallbackArgs& a, int2 tile, int wave,
                                               int blk, WaveLayout L, int lane,
                                               const unsigned int* wmask) {
  const int r = blk % L.WR, q = blk / L.WR;
  const int wr = wave / L.NWC, wc = wave % L.NWC;
  const int tile_r = 16 * L.WR * L.NWR, tile_c = 16 * L.WC * L.NWC;
  const int64_t i0 = (int64_t)tile.x * tile_r + 16 * (wr * L.WR + r);
  const int64_t j0 = (int64_t)tile.y * tile_c + 16 * (wc * L.WC + q);
  // This lane's pairs in the FP64 block: row (lane / 16) + 4 v, column
  // lane % 16.  Their flags sit in the contraction's FP32 layout (row
  // 4 (lane / 16) + v): lane 16 v + lane % 16, bit 4 blk + lane / 16.
  unsigned int fl = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    fl |= ((wmask[16 * v + (lane & 15)] >> (4 * blk + (lane >> 4))) & 1u) << v;
  if (__ballot(fl != 0) == 0) return;
  const int64_t fi = i0 + (lane & 15), fj = j0 + (lane & 15);
  const bool iok = fi < a.FA, jok = fj < a.FB;
  double ci[3] = {0.0, 0.0, 0.0}, cj[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (iok) ci[c] = gp(a.cenA)[4 * fi + c];
    if (jok) cj[c] = gp(a.cenB)[4 * fj + c];
  }
  f64x4 acc[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[c][d] = f64x4{0.0, 0.0, 0.0, 0.0};
  // Atoms in steps of 16: in sub-step s lane l feeds atom k0 + 4 (l / 16) + s
  // of row frame i0 + l % 16 (A, 16 x 4) and of column frame j0 + l % 16
  // (B, 4 x 16), so each lane reads four consecutive atoms per step.  The next
  // step's atoms are fetched before this step's 36 MFMAs.
  const int nsel = a.A.nsel;
  const int kq = 4 * (lane >> 4);
  FallbackQuad nxt;
  fallback_fetch<LAY>(a, fi, iok, fj, jok, kq, nxt);
  for (int k0 = 0; k0 < nsel; k0 += 16) {
    const FallbackQuad cur = nxt;
    fallback_fetch<LAY>(a, fi, iok, fj, jok, k0 + 16 + kq, nxt);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double av[3], bv[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        av[c] = ((double)cur.x[s][c] - ci[c]) * cur.w[s];
        bv[c] = (double)cur.y[s][c] - cj[c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          acc[c][d] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bv[d], acc[c][d], 0, 0, 0);
    }
  }
  const double M = a.total_mass[0];
  unsigned long long nrec = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    nrec += __popcll(__ballot((fl >> v) & 1u));
    if (!((fl >> v) & 1u)) continue;
    const int64_t i = i0 + (lane >> 4) + 4 * v, j = fj;
    double R[9];
    // As in the epilogue: rows are the reference in the triangle layout.
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) R[3 * c + d] = a.tri ? acc[d][c][v] : acc[c][d][v];
    const double E0 = 0.5 * (gp(a.gA)[i] + gp(a.gB)[j]);
    const int64_t idx = a.tri ? rmsd_dev::tri_index(a.FA, i, j) : i * a.FB + j;
    gp(a.out)[idx] = (float)rmsd_dev::qcp_rmsd(R, E0, M, a.clamp);
  }
  if (lane == 0) atomicAdd(a.stats + 1, nrec);
}

// fallback_block with four-atom loads where both frame sets allow them (the
// layout is uniform: a property of the sets).
__device__ __forceinline__ void fallback_any(const FallbackArgs& a, int2 tile, int wave, int blk,
                                             WaveLayout L, int lane, const unsigned int* wmask) {
  const int la = quad_layout(a.A);
  const int lay = la == quad_layout(a.B) ? la : 0;
  if (lay == 1)
    fallback_block<1>(a, tile, wave, blk, L, lane, wmask);
  else if (lay == 2)
    fallback_block<2>(a, tile, wave, blk, L, lane, wmask);
  else
    fallback_block<0>(a, tile, wave, blk, L, lane, wmask);
}

// Next 16 x 16 block for this wave (the same for every lane).
__device__ __forceinline__ unsigned int fallback_next(unsigned int* nflag, int lane) {
  unsigned int b = 0;
  if (lane == 0) b = atomicAdd(nflag + 1, 1u);
  return __shfl(b, 0, 64);
}

__global__ __launch_bounds__(FALLBACK_THREADS, 2) void k_pair_fallback(FallbackArgs a, WaveLayout L,
                                                                    unsigned int* nflag) {
  const int lane = threadIdx.x & 63;
  const int nw = L.NWR * L.NWC, nb = L.WR * L.WC;
  const unsigned int total = nflag[0] * (unsigned int)nb;
  for (unsigned int b = fallback_next(nflag, lane); b < total; b = fallback_next(nflag, lane)) {
    const int w = a.wlist[b / nb];
    fallback_any(a, a.tiles[w / nw], w % nw, (int)(b % nb), L, lane, a.masks + (int64_t)w * 64);
  }
}

__global__ __launch_bounds__(FALLBACK_THREADS, 2) void k_pair_fallback_batch(
    const FallbackArgs* fsets, const int4* tiles, WaveLayout L, const int* wlist,
    unsigned int* nflag, const unsigned int* masks) {
  const int lane = threadIdx.x & 63;
  const int nw = L.NWR * L.NWC, nb = L.WR * L.WC;
  const unsigned int total = nflag[0] * (unsigned int)nb;
  for (unsigned int b = fallback_next(nflag, lane); b < total; b = fallback_next(nflag, lane)) {
    const int w = wlist[b / nb];
    const int4 tl = tiles[w / nw];
    fallback_any(fsets[tl.x], make_int2(tl.y, tl.z), w % nw, (int)(b % nb), L, lane,
                 masks + (int64_t)w * 64);
  }
}

// One 64 x 64 tile of frame pairs: contraction over all atom chunks, then
// the FP64 epilogue.  Shared by the single-matrix and the batched kernels.
template <int PREC, int WR, int WC, int NWR, int NWC>
__device__ __forceinline__ void pair_tile(const PairArgs& args, int2 tile, int tile_id) {
  typedef PrecTraits<PREC> T;
  typedef typename T::acc_t acc_t;
  constexpr int PARTS = T::PARTS;
  constexpr int TILE_R = 16 * WR * NWR;
  constexpr int TILE_C = 16 * WC * NWC;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int wr = wave / NWC, wc = wave % NWC;
  const int64_t fbA = (int64_t)tile.x * (TILE_R / 16) + wr * WR;  // first 16-frame row block
  const int64_t fbB = (int64_t)tile.y * (TILE_C / 16) + wc * WC;  // first 16-frame col block

  const u32x4* A = reinterpret_cast<const u32x4*>(args.A);
  const u32x4* B = reinterpret_cast<const u32x4*>(args.B);
  const int KB = args.KB;

  acc_t acc[WR][WC][3][3];
#pragma unroll
  for (int r = 0; r < WR; ++r)
#pragma unroll
    for (int q = 0; q < WC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[r][q][c][d] = acc_t{0, 0, 0, 0};

  u32x4 fa[WR][3][PARTS], fb[WC][3][PARTS];
  auto load = [&](int kb) {
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
          fa[r][c][p] = ld16(frag_ptr(A, fbA + r, kb, KB, c, p, PARTS) + lane);
#pragma unroll
    for (int q = 0; q < WC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
          fb[q][c][p] = ld16(frag_ptr(B, fbB + q, kb, KB, c, p, PARTS) + lane);
  };

  load(0);
  const BoundLimits<WC> bl = bound_limits<PREC, WR, WC>(args, fbA, fbB, lane);
  for (int kb = 0; kb < KB; ++kb) {
    u32x4 ca[WR][3][PARTS], cb[WC][3][PARTS];
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p) ca[r][c][p] = fa[r][c][p];
#pragma unroll
    for (int q = 0; q < WC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p) cb[q][c][p] = fb[q][c][p];
    if (kb + 1 < KB) load(kb + 1);
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int q = 0; q < WC; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) mma<PREC>(acc[r][q][c][d], ca[r][c], cb[q][d]);
  }

  pair_epilogue<PREC, WR, WC>(args, acc, fbA, fbB, lane, tile_id, wave, NWR * NWC, bl);
}

// ---- Operands shared through LDS ---------------------------------------------------
// One 64 x 64 tile of frame pairs per 4-wave workgroup (2 x 2 waves of 32 x 32
// pairs).  Per chunk of atoms the workgroup needs NF fragments of 1 KiB (4 row
// and 4 column frame blocks x 3 coordinates x PARTS).  They go from L2 straight
// into LDS with global_load_lds_dwordx4 (64 lanes x 16 B = one fragment per
// instruction, no register staging), so each fragment crosses the vector memory
// path once per workgroup rather than once per wave that uses it: 111 instead
// of 227 bytes per split-BF16 MFMA.  One 48 KiB LDS stage and at most 256
// registers per wave let two workgroups share a CU (two waves per SIMD): while
// one waits at a barrier or runs its FP64 epilogue, the other's MFMAs run.
// Per chunk: own DMA landed + barrier (the chunk is in LDS), each wave reads
// its 24 fragments, barrier (the stage is free again), DMA of the next chunk,
// then the chunk's MFMAs from registers.  Chunk order and product order per
// pair are those of pair_tile, so the results are bit-identical to the
// register-tile kernels' (the batched path keeps those).

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// 6:4, lgkmcnt 11:8); a counter not named is not waited for.
constexpr int WAIT_VMCNT0 = (7 << 4) | (15 << 8);
constexpr int WAIT_LGKMCNT0 = 15 | (3 << 14) | (7 << 4);

// Fragments of one chunk in the LDS stage of k_pairwise_shared.
template <int PREC> constexpr int shared_frags() { return 2 * 4 * 3 * PrecTraits<PREC>::PARTS; }

// One 64 x 64 tile (tile_id: its position in the tile list) by the calling
// workgroup of 256 threads; lds: the workgroup's stage of shared_frags<PREC>()
// fragments.  Every wave of the workgroup calls it for the same tile.
template <int PREC>
__device__ __forceinline__ void shared_tile(const PairArgs& args, int2 tile, int tile_id,
                                            u32x4 (*lds)[64]) {
  typedef PrecTraits<PREC> T;
  typedef typename T::acc_t acc_t;
  constexpr int PARTS = T::PARTS;
  constexpr int WR = 2, WC = 2, NWC = 2;
  constexpr int NF = shared_frags<PREC>();  // fragments per chunk and workgroup
  constexpr int LPW = NF / 4;               // DMA instructions per wave and chunk
  // Uniform wave index: the DMA and LDS addresses stay scalar.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int wr = wave / NWC, wc = wave % NWC;
  const int64_t fb0A = (int64_t)tile.x * 4, fb0B = (int64_t)tile.y * 4;
  const int64_t fbA = fb0A + wr * WR, fbB = fb0B + wc * WC;
  const u32x4* A = reinterpret_cast<const u32x4*>(args.A);
  const u32x4* B = reinterpret_cast<const u32x4*>(args.B);
  const int KB = args.KB;

  // LDS fragment f = ((side * 4 + blk) * 3 + c) * PARTS + p; wave v fetches
  // f = v * LPW ... v * LPW + LPW - 1.
  auto fetch = [&](int kb) {
#pragma unroll
    for (int i = 0; i < LPW; ++i) {
      const int f = wave * LPW + i;
      const int p = f % PARTS, c = (f / PARTS) % 3, blk = (f / (3 * PARTS)) % 4;
      const int side = f / (12 * PARTS);
      const u32x4* src =
          frag_ptr(side ? B : A, (side ? fb0B : fb0A) + blk, kb, KB, c, p, PARTS) + lane;
      __builtin_amdgcn_global_load_lds((const void*)src,
                                       (__attribute__((address_space(3))) void*)&lds[f][0], 16, 0,
                                       0);
    }
  };

  acc_t acc[WR][WC][3][3];
#pragma unroll
  for (int r = 0; r < WR; ++r)
#pragma unroll
    for (int q = 0; q < WC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[r][q][c][d] = acc_t{0, 0, 0, 0};

  fetch(0);
  const BoundLimits<WC> bl = bound_limits<PREC, WR, WC>(args, fbA, fbB, lane);
  for (int kb = 0; kb < KB; ++kb) {
    // A bare s_barrier after explicit waits (a __syncthreads fence would also
    // wait for what it need not); the asm's memory clobber keeps the compiler
    // from moving LDS accesses across it.
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);  // this wave's DMA of chunk kb landed
    asm volatile("s_barrier" ::: "memory");   // ... and every wave's
    u32x4 ca[WR][3][PARTS], cb[WC][3][PARTS];
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p) ca[r][c][p] = lds[((wr * WR + r) * 3 + c) * PARTS + p][lane];
#pragma unroll
    for (int q = 0; q < WC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
          cb[q][c][p] = lds[((4 + wc * WC + q) * 3 + c) * PARTS + p][lane];
    __builtin_amdgcn_s_waitcnt(WAIT_LGKMCNT0);  // this wave's reads done
    asm volatile("s_barrier" ::: "memory");     // ... and every wave's: stage free
    if (kb + 1 < KB) fetch(kb + 1);
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
      for (int q = 0; q < WC; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) mma<PREC>(acc[r][q][c][d], ca[r][c], cb[q][d]);
  }
  pair_epilogue<PREC, WR, WC>(args, acc, fbA, fbB, lane, tile_id, wave, 4, bl);
}

template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_pairwise_shared(
    PairArgs args) {
  __shared__ u32x4 lds[shared_frags<PREC>()][64];
  // XCD-aware tile pick, as in k_pairwise.
  const int w = blockIdx.x;
  const int t = (w % args.nxcd) * args.per_xcd + (w / args.nxcd);
  if (t >= args.ntiles) return;  // the whole workgroup, before any barrier
  shared_tile<PREC>(args, args.tiles[t], t, lds);
}

template <int PREC, int WR, int WC, int NWR, int NWC>
__global__ __launch_bounds__(64 * NWR * NWC) void k_pairwise(PairArgs args) {
  // XCD-aware tile pick: the hardware deals workgroups round-robin over the
  // XCDs, so XCD x receives ids x, x+NXCD, ...; give each XCD one contiguous
  // run of the (grouped) tile list so that concurrently running workgroups on
  // one XCD share row and column blocks in its L2.
  const int w = blockIdx.x;
  const int t = (w % args.nxcd) * args.per_xcd + (w / args.nxcd);
  if (t >= args.ntiles) return;
  pair_tile<PREC, WR, WC, NWR, NWC>(args, args.tiles[t], t);
}

// ---- batched matrices ---------------------------------------------------------------
// Tiles of many independent matrices (one PairArgs per set) in nq queues, one
// per XCD, each a contiguous, cost-balanced run of the cost-sorted tile list.
// Grid form: workgroup w takes position w / nq of queue w % nq (the hardware
// hands out workgroups in order, so each queue drains most expensive first).
template <int PREC, int WR, int WC, int NWR, int NWC>
__global__ __launch_bounds__(64 * NWR * NWC) void k_pairwise_batch(BatchArgs b) {
  const int q = blockIdx.x % b.nq;
  const int t = b.qbeg[q] + blockIdx.x / b.nq;
  if (t >= b.qbeg[q + 1]) return;
  const int4 tl = b.tiles[t];
  const PairArgs args = b.sets[tl.x];
  pair_tile<PREC, WR, WC, NWR, NWC>(args, make_int2(tl.y, tl.z), tl.w);
}

// Persistent form: as many workgroups as are resident; each drains its own
// queue through an atomic position counter, then steals from the others.
// Every workgroup leaves once every queue is exhausted.  The counters must be
// zero at launch (rmsd_launch_batch_reset).
template <int PREC, int WR, int WC, int NWR, int NWC>
__global__ __launch_bounds__(64 * NWR * NWC) void k_pairwise_batch_persist(BatchArgs b) {
  __shared__ int slot;
  const int q0 = blockIdx.x % b.nq;
  for (int k = 0; k < b.nq; ++k) {
    const int q = (q0 + k) % b.nq;
    const int beg = b.qbeg[q], len = b.qbeg[q + 1] - beg;
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) slot = atomicAdd(b.counters + q, 1);
      __syncthreads();
      const int pos = __builtin_amdgcn_readfirstlane(slot);
      if (pos >= len) break;
      const int4 tl = b.tiles[beg + pos];
      const PairArgs args = b.sets[tl.x];
      pair_tile<PREC, WR, WC, NWR, NWC>(args, make_int2(tl.y, tl.z), tl.w);
    }
  }
}

// The batched forms of k_pairwise_shared (grid and persistent, as above).
template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_pairwise_shared_batch(BatchArgs b) {
  __shared__ u32x4 lds[shared_frags<PREC>()][64];
  const int q = blockIdx.x % b.nq;
  const int t = b.qbeg[q] + blockIdx.x / b.nq;
  if (t >= b.qbeg[q + 1]) return;
  const int4 tl = b.tiles[t];
  // by reference: the fields are read where they are used (scalar loads)
  const PairArgs& args = b.sets[tl.x];
  shared_tile<PREC>(args, make_int2(tl.y, tl.z), tl.w, lds);
}

template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_pairwise_shared_batch_persist(BatchArgs b) {
  __shared__ u32x4 lds[shared_frags<PREC>()][64];
  __shared__ int slot;
  const int q0 = blockIdx.x % b.nq;
  for (int k = 0; k < b.nq; ++k) {
    const int q = (q0 + k) % b.nq;
    const int beg = b.qbeg[q], len = b.qbeg[q + 1] - beg;
    for (;;) {
      // also orders the previous tile's last LDS reads before the next DMA
      __syncthreads();
      if (threadIdx.x == 0) slot = atomicAdd(b.counters + q, 1);
      __syncthreads();
      const int pos = __builtin_amdgcn_readfirstlane(slot);
      if (pos >= len) break;
      const int4 tl = b.tiles[beg + pos];
      const PairArgs& args = b.sets[tl.x];
      shared_tile<PREC>(args, make_int2(tl.y, tl.z), tl.w, lds);
    }
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.  Tile shapes: FP64 waves own 16x32 pairs (72 FP64 accumulators),
// FP32 and split-BF16 waves 32x48 pairs (216 FP32 accumulators); a workgroup
// covers 64 x 64 (FP64) or 64 x 96 frame pairs.
template <int PREC, int WR, int WC, int NWR, int NWC>
static hipError_t launch_tiles(const PairArgs& a, hipStream_t s) {
  const int nblocks = a.per_xcd * a.nxcd;
  hipLaunchKernelGGL((k_pairwise<PREC, WR, WC, NWR, NWC>), dim3(nblocks), dim3(64 * NWR * NWC), 0,
                     s, a);
  return hipGetLastError();
}

// Pair tile of the single-matrix kernel: 64 row frames x pairwise_tile_cols(prec)
// column frames.  FP32 waves own 32 x 48 pairs of a register-tile kernel (fewer
// operand bytes per MFMA than 32 x 32: 1 % faster on the 50k-frame matrix);
// split-BF16 waves 32 x 32 pairs of k_pairwise_shared's 64 x 64 tiles.
int rmsd_pairwise_tile_cols(int prec) { return prec == PREC_F32 ? 96 : 64; }

// The <WR, WC, NWR, NWC> of the launches below.
WaveLayout rmsd_pairwise_layout(int prec) {
  switch (prec) {
    case PREC_F64: return WaveLayout{1, 2, 4, 2};
    case PREC_F32: return WaveLayout{2, 3, 2, 2};
    default: return WaveLayout{2, 2, 2, 2};
  }
}

WaveLayout rmsd_batch_layout(int prec, int variant) {
  if ((variant >> 1) == 0) return rmsd_pairwise_layout(prec);
  return prec == PREC_F64 ? WaveLayout{1, 2, 2, 1} : WaveLayout{2, 2, 1, 1};
}

template <int PREC>
static hipError_t launch_shared(const PairArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_pairwise_shared<PREC>), dim3(a.per_xcd * a.nxcd), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t rmsd_launch_pairwise(int prec, const PairArgs& a, hipStream_t s) {
  switch (prec) {
    case PREC_F64: return launch_tiles<PREC_F64, 1, 2, 4, 2>(a, s);
    case PREC_F32: return launch_tiles<PREC_F32, 2, 3, 2, 2>(a, s);
    case PREC_BF16X3: return launch_shared<PREC_BF16X3>(a, s);
    default: return hipErrorInvalidValue;
  }
}

// Batched launches.  Variant v: tile shape v / 2 (shape 0: the single-matrix
// kernels' tiles, 64 x 64 for FP64 and split-BF16 (k_pairwise_shared), 64 x 96
// for FP32; shape 1: 32 x 32 pair
// tiles of one or two waves), grid form for even v, persistent form for odd v.
// The tile list must be built for the shape (rmsd_batch_tile_shape).
template <int PREC, int WR, int WC, int NWR, int NWC>
static hipError_t launch_batch_shape(const BatchArgs& b, bool persist, hipStream_t s) {
  constexpr int NT = 64 * NWR * NWC;
  if (!persist) {
    const int64_t grid = (int64_t)b.nq * b.maxq;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_pairwise_batch<PREC, WR, WC, NWR, NWC>), dim3((unsigned)grid), dim3(NT),
                       0, s, b);
    return hipGetLastError();
  }
  int dev = 0, ncu = 256, bpc = 1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &bpc, reinterpret_cast<const void*>(&k_pairwise_batch_persist<PREC, WR, WC, NWR, NWC>), NT, 0);
  if (e != hipSuccess) return e;
  if (bpc < 1) bpc = 1;
  int64_t grid = (int64_t)ncu * bpc;
  const int64_t ntiles = (int64_t)b.nq * b.maxq;
  if (grid > ntiles) grid = ntiles;
  if (grid < b.nq) grid = b.nq;
  hipLaunchKernelGGL((k_pairwise_batch_persist<PREC, WR, WC, NWR, NWC>), dim3((unsigned)grid),
                     dim3(NT), 0, s, b);
  return hipGetLastError();
}

// A wave's pairs must fit the 32 bits of a lane mask.
static bool fallback_fits(WaveLayout L) { return L.WR * L.WC * 4 <= 32; }

hipError_t rmsd_launch_pair_fallback(const FallbackArgs& a, WaveLayout L, unsigned int* nflag,
                                     hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  if (!fallback_fits(L)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_fallback, dim3(FALLBACK_GRID), dim3(FALLBACK_THREADS), 0, s, a, L,
                     nflag);
  return hipGetLastError();
}

hipError_t rmsd_launch_pair_fallback_batch(const FallbackArgs* fsets, const int4* tiles,
                                           int ntiles, WaveLayout L, const int* wlist,
                                           unsigned int* nflag, const unsigned int* masks,
                                           hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (!fallback_fits(L)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_fallback_batch, dim3(FALLBACK_GRID), dim3(FALLBACK_THREADS), 0, s,
                     fsets, tiles, L, wlist, nflag, masks);
  return hipGetLastError();
}

hipError_t rmsd_launch_fill_int(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

int rmsd_batch_variant_count() { return 4; }

void rmsd_batch_tile_shape(int prec, int variant, int* rows, int* cols) {
  if ((variant >> 1) == 0) {
    *rows = 64;
    *cols = rmsd_pairwise_tile_cols(prec);
  } else {
    *rows = *cols = 32;
  }
}

hipError_t rmsd_launch_batch_reset(const BatchArgs& b, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, b.counters, b.nq, 0);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_batch_shared(const BatchArgs& b, bool persist, hipStream_t s) {
  if (!persist) {
    const int64_t grid = (int64_t)b.nq * b.maxq;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_pairwise_shared_batch<PREC>), dim3((unsigned)grid), dim3(256), 0, s, b);
    return hipGetLastError();
  }
  int dev = 0, ncu = 256, bpc = 1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &bpc, reinterpret_cast<const void*>(&k_pairwise_shared_batch_persist<PREC>), 256, 0);
  if (e != hipSuccess) return e;
  if (bpc < 1) bpc = 1;
  int64_t grid = (int64_t)ncu * bpc;
  const int64_t ntiles = (int64_t)b.nq * b.maxq;
  if (grid > ntiles) grid = ntiles;
  if (grid < b.nq) grid = b.nq;
  hipLaunchKernelGGL((k_pairwise_shared_batch_persist<PREC>), dim3((unsigned)grid), dim3(256), 0, s,
                     b);
  return hipGetLastError();
}

hipError_t rmsd_launch_pairwise_batch(int prec, int variant, const BatchArgs& b, hipStream_t s) {
  if (variant < 0 || variant >= rmsd_batch_variant_count()) return hipErrorInvalidValue;
  const bool persist = variant & 1;
  const int shape = variant >> 1;
  switch (prec) {
    case PREC_F64:
      return shape == 0 ? launch_batch_shape<PREC_F64, 1, 2, 4, 2>(b, persist, s)
                        : launch_batch_shape<PREC_F64, 1, 2, 2, 1>(b, persist, s);
    case PREC_F32:
      return shape == 0 ? launch_batch_shape<PREC_F32, 2, 3, 2, 2>(b, persist, s)
                        : launch_batch_shape<PREC_F32, 2, 2, 1, 1>(b, persist, s);
    case PREC_BF16X3:
      return shape == 0 ? launch_batch_shared<PREC_BF16X3>(b, persist, s)
                        : launch_batch_shape<PREC_BF16X3, 2, 2, 1, 1>(b, persist, s);
    default: return hipErrorInvalidValue;
  }
}
