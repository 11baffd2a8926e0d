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
__device__ __forceinline__ u32x4 ld16(const u32x4* p) { return *p; }

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
// MFMA-reduced forms), forms E0 and runs QCP pair by pair in registers.  The
// per-frame terms are read from gA/uA (row frames, indexed i - rbase) and
// gB/uB (column frames, j - cbase): global memory or an LDS copy.  (Grouping
// several pairs' Newton iterations was measured slower on the 50k-frame
// matrix.)
template <int PREC, int WR, int WC>
__device__ __forceinline__ void pair_epilogue(const PairArgs& args,
                                              typename PrecTraits<PREC>::acc_t (&acc)[WR][WC][3][3],
                                              int64_t fbA, int64_t fbB, int lane,
                                              const double* gA, const double* uA, int64_t rbase,
                                              const double* gB, const double* uB, int64_t cbase,
                                              int tile_id) {
  const double M = args.total_mass[0];
  // Error bound (MFMA-reduced forms with a flag array): per lane, whether a
  // pair went over the tolerance, the kept pair with the largest eps / r
  // (compared by cross-multiplication: no division per pair) and the largest
  // kept r; for tol = infinity (report only) also the bounds of the pairs
  // that fail r^2 > eps.
  const bool bounded = PREC != PREC_F64 && args.flags != nullptr && M >= RMSD_SMALL;
  const double ke0M = bounded ? args.ke0 / M : 0.0;
  bool over = false;
  double qe = 0.0, qr = 1.0, rmax = 0.0, bover = 0.0;
  double Pm[9];
  if (PREC != PREC_F64) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Pm[k] = args.P[k];
  }
#pragma unroll
  for (int q = 0; q < WC; ++q) {
    const int64_t j = (fbB + q) * 16 + (lane & 15);  // column frame
    double gj = 0.0, uj[9], csj = 0.0;
    const bool jok = j < args.FB;
    if (jok) {
      gj = gB[j - cbase];
      if (PREC != PREC_F64) {
#pragma unroll
        for (int k = 0; k < 9; ++k) uj[k] = uB[9 * (j - cbase) + k];
        if (bounded) csj = args.nrmB[4 * j];
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
          const double* ui = uA + 9 * (i - rbase);
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
              S[3 * c + d] += Pm[3 * c + d] + uj[3 * c + d] + ui[3 * d + c];
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
        const double E0 = 0.5 * (gA[i - rbase] + gj);
        const int64_t idx = args.tri ? rmsd_dev::tri_index(args.FA, i, j) : i * args.FB + j;
        const float r = M < RMSD_SMALL ? -1.0f : (float)rmsd_dev::qcp_rmsd(R, E0, M, args.clamp);
        args.out[idx] = r;
        if (bounded) {
          const double rr = (double)r;
          const double eps = args.nrmA[4 * i] * csj + ke0M * E0;
          if (rmsd_dev::bound_keep(eps, rr, args.tol)) {
            if (eps * qr > qe * rr) {
              qe = eps;
              qr = rr;
            }
            rmax = fmax(rmax, rr);
          } else {
            over = true;  // NaN included: recomputed in FP64
            if (args.tol == INFINITY) bover = fmax(bover, rmsd_dev::bound_value(eps, rr));
          }
        }
      }
    }
  }
  if (bounded) {
    // One flag store and one atomic per wave.
    if (__ballot(over) && lane == 0) args.flags[tile_id] = 1;
    double bmax = fmax(qe / qr * (1.0 + 0x1p-23) + rmax * 0x1p-23, bover);
    for (int off = 32; off > 0; off >>= 1) bmax = fmax(bmax, __shfl_xor(bmax, off, 64));
    if (lane == 0 && bmax > 0.0)
      atomicMax(args.stats, (unsigned long long)__double_as_longlong(bmax));
  }
}

// ---- FP64 recomputation of flagged pairs -------------------------------------------
// One workgroup per tile of the schedule; tiles without a flag leave at once.
// The bound is evaluated again from the stored float RMSD (the epilogue
// decided on the same float with the same test, so both agree) and the pairs
// over the tolerance are listed in LDS; then each thread recomputes whole
// pairs, one at a time, summing over the atoms in FP64 from the frames as
// stored, centred with the per-frame centroids of k_prep, and solves them with
// the same FP64 QCP.  Consecutive threads take consecutive pairs of a tile row:
// the row frame's atoms are one broadcast load, the column frames' lines stay
// in the vector cache across the atom loop.
constexpr int FALLBACK_THREADS = 256;
constexpr int FALLBACK_MAX_PAIRS = 64 * 96;  // largest tile of any schedule

__device__ __forceinline__ void centred_atom(const FrameSetDev& fs, int64_t f, int k,
                                             const double* cen, const double* w, double x[3]) {
  float p[3];
  load_atom(fs, f, k, p);
  const double s = w ? sqrt(w[k]) : 1.0;
  x[0] = ((double)p[0] - cen[0]) * s;
  x[1] = ((double)p[1] - cen[1]) * s;
  x[2] = ((double)p[2] - cen[2]) * s;
}

__device__ __forceinline__ void fallback_tile(const FallbackArgs& a, int2 tile, int tile_r,
                                              int tile_c) {
  __shared__ unsigned short list[FALLBACK_MAX_PAIRS];
  __shared__ int nlist;
  const double M = a.total_mass[0];
  const double ke0M = a.ke0 / M;
  const int npairs = tile_r * tile_c;
  if (threadIdx.x == 0) nlist = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < npairs; p += FALLBACK_THREADS) {
    const int64_t i = (int64_t)tile.x * tile_r + p / tile_c;
    const int64_t j = (int64_t)tile.y * tile_c + p % tile_c;
    if (i < a.FA && j < a.FB && (!a.tri || i < j)) {
      const int64_t idx = a.tri ? rmsd_dev::tri_index(a.FA, i, j) : i * a.FB + j;
      const double eps = a.nrmA[4 * i] * a.nrmB[4 * j] + ke0M * 0.5 * (a.gA[i] + a.gB[j]);
      if (!rmsd_dev::bound_keep(eps, (double)a.out[idx], a.tol))
        list[atomicAdd(&nlist, 1)] = (unsigned short)p;
    }
  }
  __syncthreads();
  const int n = nlist;
  for (int q = threadIdx.x; q < n; q += FALLBACK_THREADS) {
    const int p = list[q];
    const int64_t i = (int64_t)tile.x * tile_r + p / tile_c;
    const int64_t j = (int64_t)tile.y * tile_c + p % tile_c;
    const double* ci = a.cenA + 4 * i;
    const double* cj = a.cenB + 4 * j;
    double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < a.A.nsel; ++k) {
      double x[3], y[3];
      centred_atom(a.A, i, k, ci, a.smass, x);
      centred_atom(a.B, j, k, cj, a.smass, y);
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) S[3 * c + d] += x[c] * y[d];
    }
    double R[9];
    // As in the epilogue: rows are the reference in the triangle layout.
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) R[3 * c + d] = a.tri ? S[3 * d + c] : S[3 * c + d];
    const double E0 = 0.5 * (a.gA[i] + a.gB[j]);
    const int64_t idx = a.tri ? rmsd_dev::tri_index(a.FA, i, j) : i * a.FB + j;
    a.out[idx] = (float)rmsd_dev::qcp_rmsd(R, E0, M, a.clamp);
  }
  if (threadIdx.x == 0 && n) atomicAdd(a.stats + 1, (unsigned long long)n);
}

__global__ __launch_bounds__(FALLBACK_THREADS) void k_pair_fallback(FallbackArgs a) {
  const int t = blockIdx.x;
  if (t >= a.ntiles || a.flags[t] == 0) return;
  fallback_tile(a, a.tiles[t], a.tile_r, a.tile_c);
}

__global__ __launch_bounds__(FALLBACK_THREADS) void k_pair_fallback_batch(const FallbackArgs* fsets,
                                                             const int4* tiles, int ntiles,
                                                             int tile_r, int tile_c,
                                                             const int* flags) {
  const int t = blockIdx.x;
  if (t >= ntiles || flags[t] == 0) return;
  const int4 tl = tiles[t];
  fallback_tile(fsets[tl.x], make_int2(tl.y, tl.z), tile_r, tile_c);
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

  pair_epilogue<PREC, WR, WC>(args, acc, fbA, fbB, lane, args.gA, args.uA, 0, args.gB, args.uB, 0,
                              tile_id);
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
// column frames.  FP32 and split-BF16 waves own 32 x 48 pairs (fewer operand
// bytes per MFMA than 32 x 32; measured 3 % (split-BF16) and 1 % (FP32) faster
// on the 50k-frame matrix).
int rmsd_pairwise_tile_cols(int prec) { return prec == PREC_F64 ? 64 : 96; }

hipError_t rmsd_launch_pairwise(int prec, const PairArgs& a, hipStream_t s) {
  switch (prec) {
    case PREC_F64: return launch_tiles<PREC_F64, 1, 2, 4, 2>(a, s);
    case PREC_F32: return launch_tiles<PREC_F32, 2, 3, 2, 2>(a, s);
    case PREC_BF16X3: return launch_tiles<PREC_BF16X3, 2, 3, 2, 2>(a, s);
    default: return hipErrorInvalidValue;
  }
}

// Batched launches.  Variant v: tile shape v / 2 (shape 0: the single-matrix
// tiles, 64 x 64 for FP64 and 64 x 96 for FP32 / BF16; shape 1: 32 x 32 pair
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

hipError_t rmsd_launch_pair_fallback(const FallbackArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  if (a.tile_r * a.tile_c > FALLBACK_MAX_PAIRS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_fallback, dim3(a.ntiles), dim3(FALLBACK_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t rmsd_launch_pair_fallback_batch(const FallbackArgs* fsets, const int4* tiles,
                                           int ntiles, int tile_r, int tile_c, const int* flags,
                                           hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (tile_r * tile_c > FALLBACK_MAX_PAIRS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_fallback_batch, dim3(ntiles), dim3(FALLBACK_THREADS), 0, s, fsets,
                     tiles, ntiles, tile_r, tile_c, flags);
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
      return shape == 0 ? launch_batch_shape<PREC_BF16X3, 2, 3, 2, 2>(b, persist, s)
                        : launch_batch_shape<PREC_BF16X3, 2, 2, 1, 1>(b, persist, s);
    default: return hipErrorInvalidValue;
  }
}
