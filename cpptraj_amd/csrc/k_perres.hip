// k_perres.hip -- per-residue no-fit RMSD of the rmsd action ('perres').
//
//   k_perres  Action_Rmsd::DoAction's per-residue loop (src/Action_Rmsd.cpp:
//             398-413): for every frame and residue, the residue's atoms of the
//             frame (optionally as the fit's coordinate update leaves them, in
//             FP64: Frame::Trans_Rot_Trans, src/Frame.h:572-581, or the
//             translate-only update) against the same residue of the uncentred
//             reference frame, both optionally centred (Frame::CenterOnOrigin,
//             src/Frame.cpp:1043-1055), then Frame::RMSD_NoFit
//             (src/Frame.cpp:1279-1307) with the target masses.
//
// Layout.  The residue list is cut by the host into pieces: runs of at most
// PERRES_RES_PIECE whole residues whose LDS image (per entry: mass, reference
// atom, offset; per residue: offsets and reference centre; the target atoms
// of a frame, twice) fits PERRES_LDS_PIECE bytes (a single larger residue
// gets a piece of its own, up to PERRES_LDS_MAX).  A piece stages either the
// contiguous span of target atoms it touches, in the frames' own layout (AoS
// or SoA rows), or, when its atoms are scattered (a span more than twice its
// atom count), just its own atoms (gather, AoS order).
//
// Schedule.  Workgroups are one wavefront and own one piece; they walk over
// batches of `nfb` frames, enough that the batch's (frame, residue) items fill
// the lanes.  Frames and their coordinate updates travel global -> LDS by
// LDS-DMA (global_load_lds_dword: no registers held, a whole batch in flight),
// double buffered: batch i+1 is requested before batch i is computed, and the
// arithmetic touches only LDS.  One lane computes one (frame, residue),
// summing atom by atom in the reference's order with FP64 contraction off.
// HBM traffic is the frames' atoms once (12 B per atom and frame) plus 8 B per
// result (plus 120 B per frame of coordinate update).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_kernels.hpp"

namespace {

constexpr double kSmall = 0.00000000000001;  // Constants::SMALL, src/Constants.h:35
// s_waitcnt vmcnt(0) in the gfx9 encoding (expcnt, lgkmcnt not waited).
constexpr int kWaitVm0 = (7 << 4) | (15 << 8);

typedef __attribute__((address_space(3))) void lds_void;

// q / d for 0 <= q < 2^20 and a uniform d >= 1: float estimate, corrected.
__device__ __forceinline__ int udiv_small(int q, int d, float inv) {
  int t = (int)((float)q * inv);
  t -= (t * d > q) ? 1 : 0;
  t += ((t + 1) * d <= q) ? 1 : 0;
  return t;
}

// One dword global -> LDS per active lane: lane i of the wave writes
// lds_base + 4 i (lds_base uniform, in M0).  Issued as inline asm so that the
// compiler's wait insertion does not treat it as a write to every LDS buffer
// (it would then wait for it before each LDS read, serialising the transfer
// of the next batch with the arithmetic of this one).  Completion is awaited
// explicitly (s_waitcnt vmcnt(0)) before a buffer is read; vmcnt retires in
// order, so the compiler's own waits for its loads stay correct.
__device__ __forceinline__ void dma_dword(const void* src, void* lds_base) {
  const unsigned m0 = (unsigned)(size_t)(lds_void*)lds_base;
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off"
               :
               : "v"(src), "s"(m0)
               : "m0", "memory");
}

// Global address of float q of a batch: LAYOUT 0 = AoS span (floats 3 lo ..
// 3 hi of each frame), 1 = SoA span (rows c of atoms lo .. hi), 2 = gather
// (the piece's own atoms, AoS order; atom indices from LDS).  LDS keeps the
// batch in q order.  `mine` is lane fs's stored frame of the batch: no global
// load happens here, so the DMAs of a batch never wait for one another.
// Called with the whole wavefront active (the shuffle reads lane fs).
template <int LAYOUT>
__device__ __forceinline__ const float* src_of(const PerresArgs& a, int64_t f0, int mine, int q,
                                               int fspan, float inv_fspan, int lo, int span,
                                               float inv_span, const int* gat) {
  const int fs = udiv_small(q, fspan, inv_fspan), w = q - fs * fspan;
  const int64_t stored = a.fidx ? (int64_t)__shfl(mine, fs) : f0 + fs;
  const float* fr = a.xyz + (int64_t)stored * a.stride;
  if (LAYOUT == 0) return fr + 3 * (int64_t)lo + w;
  if (LAYOUT == 1) {
    const int c = udiv_small(w, span, inv_span);
    return fr + c * a.ld + lo + (w - c * span);
  }
  const int k = w / 3, c = w - 3 * k;
  const int at = gat[k];
  return a.ld ? fr + c * a.ld + at : fr + 3 * (int64_t)at + c;
}

template <int LAYOUT>
__device__ __forceinline__ void dma_frames(const PerresArgs& a, float* fx, int64_t f0, int mine,
                                           int total, int fspan, int lo, int span, const int* gat) {
  const float inv_fspan = 1.0f / (float)fspan, inv_span = 1.0f / (float)span;
  for (int base = 0; base < total; base += PERRES_THREADS) {  // uniform trip count
    const int q = base + (int)threadIdx.x;
    const float* src = src_of<LAYOUT>(a, f0, mine, q < total ? q : total - 1, fspan, inv_fspan, lo,
                                      span, inv_span, gat);
    if (q < total) dma_dword(src, fx + base);
  }
}

// Span pieces: per frame (and row), contiguous runs of 64 floats, so an
// address is a uniform base plus the lane's offset.  `mine` holds lane fs's
// stored frame when there is a frame list (read with v_readlane, which needs
// no particular lanes active).
template <bool SOA>
__device__ __forceinline__ void dma_span(const PerresArgs& a, float* fx, int64_t f0, int nf,
                                         int mine, int fspan, int lo, int span) {
  const int lane = (int)threadIdx.x;
  for (int fs = 0; fs < nf; ++fs) {
    const int64_t stored = a.fidx ? (int64_t)__builtin_amdgcn_readlane(mine, fs) : f0 + fs;
    const float* fr = a.xyz + stored * a.stride;
    float* dst = fx + fs * fspan;
    if (SOA) {
      for (int c = 0; c < 3; ++c) {
        const float* row = fr + c * a.ld + lo;
        for (int a0 = 0; a0 < span; a0 += PERRES_THREADS)
          if (a0 + lane < span) dma_dword(row + a0 + lane, dst + c * span + a0);
      }
    } else {
      const float* row = fr + 3 * (int64_t)lo;
      for (int a0 = 0; a0 < fspan; a0 += PERRES_THREADS)
        if (a0 + lane < fspan) dma_dword(row + a0 + lane, dst + a0);
    }
  }
}

// Requests batch f0 (nf frames) into one buffer: the frames' staged floats
// and, with a coordinate update, U (9 doubles) and t1, t2 (6) per frame.
template <int MOD>
__device__ __forceinline__ void issue_batch(const PerresArgs& a, int layout, float* fx, double* xf,
                                            int64_t f0, int nf, int fspan, int lo, int span,
                                            const int* gat) {
  // With a frame list: the stored frame of batch frame `lane` (nf <= one
  // wavefront), read once per batch, first, so that waiting for it never
  // waits for this batch's DMAs as well.
  const int lane = (int)threadIdx.x;
  const int mine = a.fidx ? a.fidx[f0 + (lane < nf ? lane : 0)] : 0;
  __builtin_amdgcn_s_waitcnt(kWaitVm0);  // only this load is outstanding here
  if (MOD) {
    const int nd = 30 * nf;  // dwords
    for (int base = 0; base < nd; base += PERRES_THREADS) {
      const int q = base + (int)threadIdx.x;
      if (q < nd) {
        const int fs = q / 30, d = q - 30 * fs;
        const int64_t f = f0 + fs;
        const float* tv = (const float*)(a.tvec + 6 * f);
        // translate-only updates carry no U: its slots repeat tvec (unread)
        const float* src = d >= 18 ? tv + (d - 18) : (MOD == 1 ? (const float*)(a.U + 9 * f) + d : tv);
        dma_dword(src, (float*)xf + base);
      }
    }
  }
  const int total = nf * fspan;
  if (total > 0) {
    if (layout == 0) dma_span<false>(a, fx, f0, nf, mine, fspan, lo, span);
    else if (layout == 1) dma_span<true>(a, fx, f0, nf, mine, fspan, lo, span);
    else dma_frames<2>(a, fx, f0, mine, total, fspan, lo, span, gat);
  }
}

// Target atom t of a staged frame (SOA: rows of `span`), after the fit's
// update in FP64 with the reference's operation order.
template <int MOD, bool SOA>
__device__ __forceinline__ void moved(const float* frm, int t, int span, const double (&X)[15],
                                      double& x, double& y, double& z) {
#pragma clang fp contract(off)
  if (SOA) {
    x = (double)frm[t];
    y = (double)frm[span + t];
    z = (double)frm[2 * span + t];
  } else {
    x = (double)frm[3 * t];
    y = (double)frm[3 * t + 1];
    z = (double)frm[3 * t + 2];
  }
  if (MOD == 1) {
    const double u = x + X[9], v = y + X[10], w = z + X[11];
    x = u * X[0] + v * X[1] + w * X[2] + X[12];
    y = u * X[3] + v * X[4] + w * X[5] + X[13];
    z = u * X[6] + v * X[7] + w * X[8] + X[14];
  } else if (MOD == 2) {
    x = x + (X[9] + X[12]);
    y = y + (X[10] + X[13]);
    z = z + (X[11] + X[14]);
  }
}

struct PieceLds {
  const double* tm;    // [ne] target masses (1 when unweighted)
  const double* cref;  // [3 nr] reference residue centres
  const float* rref;   // [3 ne] reference atoms
  const int* toff;     // [ne] staged atom of each entry
  const int* roff;     // [nr + 1] entry offsets of the piece's residues
};

// The (frame, residue) items of one staged batch.
template <int MOD, bool CENTER, bool SOA>
__device__ __forceinline__ void compute_batch(const PerresArgs& a, const PieceLds& P,
                                              const float* fx, const double* xf, int64_t f0,
                                              int nf, int nr, int r0, int fspan, int span) {
#pragma clang fp contract(off)
  for (int it = threadIdx.x; it < nf * nr; it += PERRES_THREADS) {
    const int fs = it / nr, r = it - fs * nr;
    double X[15];
    if (MOD)
#pragma unroll
      for (int k = 0; k < 15; ++k) X[k] = xf[15 * fs + k];
    const float* frm = fx + fs * fspan;
    const int b = P.roff[r], e = P.roff[r + 1], n = e - b;
    double ct[3] = {0.0, 0.0, 0.0};
    if (CENTER) {
      double s[3] = {0.0, 0.0, 0.0}, total = 0.0;
#pragma unroll 2
      for (int k = b; k < e; ++k) {
        double x, y, z;
        moved<MOD, SOA>(frm, P.toff[k], span, X, x, y, z);
        if (a.use_mass) {
          const double m = P.tm[k];
          s[0] += x * m;
          s[1] += y * m;
          s[2] += z * m;
          total += m;
        } else {
          s[0] += x;
          s[1] += y;
          s[2] += z;
        }
      }
      if (a.use_mass) {
        if (total != 0.0)
          for (int d = 0; d < 3; ++d) ct[d] = s[d] / total;
      } else {
        for (int d = 0; d < 3; ++d) ct[d] = s[d] / (double)n;
      }
    }
    const double crx = P.cref[3 * r], cry = P.cref[3 * r + 1], crz = P.cref[3 * r + 2];
    double acc = 0.0, total = 0.0;
#pragma unroll 2
    for (int k = b; k < e; ++k) {
      double x, y, z;
      moved<MOD, SOA>(frm, P.toff[k], span, X, x, y, z);
      const double dx = ((double)P.rref[3 * k] - crx) - (x - ct[0]);
      const double dy = ((double)P.rref[3 * k + 1] - cry) - (y - ct[1]);
      const double dz = ((double)P.rref[3 * k + 2] - crz) - (z - ct[2]);
      const double m = P.tm[k];
      total += m;
      acc += m * (dx * dx + dy * dy + dz * dz);
    }
    double val;
    if (total < kSmall) val = -1.0;
    else if (acc < 0.0) val = 0.0;
    else val = sqrt(acc / total);
    a.out[(f0 + fs) * a.nres + r0 + r] = val;
  }
}

template <int MOD, bool CENTER>
__global__ __launch_bounds__(PERRES_THREADS) void k_perres(PerresArgs a) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char lds[];
  // Piece: residues [r0, r1); staged target atoms [lo, hi) (span form) or,
  // with lo < 0, the piece's own atoms in entry order (gather form).
  const int4 pc = a.pieces[blockIdx.y];
  const int r0 = pc.x, nr = pc.y - pc.x;
  const int e0 = a.res_off[r0], ne = a.res_off[pc.y] - e0;
  const bool gather = pc.z < 0;
  const int lo = gather ? 0 : pc.z, span = gather ? ne : pc.w - pc.z;
  const int layout = gather ? 2 : (a.ld ? 1 : 0);
  const int nfb = a.nfb, fspan = 3 * span;
  // LDS: xf[2][nfb][15] double | tm[ne] double | cref[3 nr] double |
  //      rref[3 ne] float | toff[ne] int | gat[ne] int | roff[nr + 1] int |
  //      fx[2][nfb][3 span] float
  double* xf0 = (double*)lds;
  double* tm = xf0 + 2 * 15 * nfb;
  double* cref = tm + ne;
  float* rref = (float*)(cref + 3 * nr);
  int* toff = (int*)(rref + 3 * ne);
  int* gat = toff + ne;
  int* roff = gat + ne;
  float* fx0 = (float*)(roff + nr + 1);

  const float* rfr = a.ref.xyz + frame_of(a.ref, a.ref_frame) * a.ref.stride;
  for (int k = threadIdx.x; k < ne; k += PERRES_THREADS) {
    const int q = a.ref_atoms[e0 + k];
    for (int c = 0; c < 3; ++c) rref[3 * k + c] = a.ref.ld ? rfr[c * a.ref.ld + q] : rfr[3 * q + c];
    tm[k] = a.use_mass ? a.tmass[e0 + k] : 1.0;
    const int at = a.tgt_atoms[e0 + k];
    toff[k] = gather ? k : at - lo;
    gat[k] = at;
  }
  for (int r = threadIdx.x; r <= nr; r += PERRES_THREADS) roff[r] = a.res_off[r0 + r] - e0;
  __syncthreads();
  // Reference residue centres (CenterOnOrigin with the reference masses).
  for (int r = threadIdx.x; r < nr; r += PERRES_THREADS) {
    double c[3] = {0.0, 0.0, 0.0};
    if (CENTER) {
      const int b = roff[r], n = roff[r + 1] - b;
      double s[3] = {0.0, 0.0, 0.0}, total = 0.0;
      for (int k = 0; k < n; ++k) {
        const double m = a.use_mass ? a.rmass[e0 + b + k] : 1.0;
        for (int d = 0; d < 3; ++d) s[d] += a.use_mass ? (double)rref[3 * (b + k) + d] * m
                                                       : (double)rref[3 * (b + k) + d];
        total += m;
      }
      if (a.use_mass) {
        if (total != 0.0)
          for (int d = 0; d < 3; ++d) c[d] = s[d] / total;
      } else {
        for (int d = 0; d < 3; ++d) c[d] = s[d] / (double)n;
      }
    }
    for (int d = 0; d < 3; ++d) cref[3 * r + d] = c[d];
  }
  const PieceLds P{tm, cref, rref, toff, roff};

  const int64_t step = (int64_t)gridDim.x * nfb;
  auto batch_size = [&](int64_t g) {
    return (int)((a.nframes - g) < nfb ? (a.nframes - g) : nfb);
  };
  int64_t f0 = (int64_t)blockIdx.x * nfb;
  int buf = 0;
  if (f0 < a.nframes)
    issue_batch<MOD>(a, layout, fx0, xf0, f0, batch_size(f0), fspan, lo, span, gat);
  for (; f0 < a.nframes; f0 += step, buf ^= 1) {
    // This wave's requests for batch f0 have landed (and its LDS writes of
    // the set-up above are visible); the workgroup is this one wave.
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    const int64_t g = f0 + step;
    if (g < a.nframes)  // the next batch travels during this one's arithmetic
      issue_batch<MOD>(a, layout, fx0 + (buf ^ 1) * nfb * fspan, xf0 + (buf ^ 1) * 15 * nfb, g,
                       batch_size(g), fspan, lo, span, gat);
    const float* fx = fx0 + buf * nfb * fspan;
    const double* xf = xf0 + buf * 15 * nfb;
    const int nf = batch_size(f0);
    if (layout == 1) compute_batch<MOD, CENTER, true>(a, P, fx, xf, f0, nf, nr, r0, fspan, span);
    else compute_batch<MOD, CENTER, false>(a, P, fx, xf, f0, nf, nr, r0, fspan, span);
  }
}

}  // namespace

size_t rmsd_perres_lds_bytes(int nres_piece, int nentries, int span, int nfb) {
  return sizeof(double) * (2 * 15 * (size_t)nfb + (size_t)nentries + 3 * (size_t)nres_piece) +
         sizeof(float) * 3 * (size_t)nentries + sizeof(int) * (2 * (size_t)nentries + nres_piece + 1) +
         2 * sizeof(float) * 3 * (size_t)span * (size_t)nfb;
}

hipError_t rmsd_launch_perres(const PerresArgs& a, int npieces, size_t lds_bytes, hipStream_t s) {
  if (a.nframes == 0 || npieces == 0) return hipSuccess;
  if (lds_bytes > PERRES_LDS_MAX || npieces > 65535) return hipErrorInvalidValue;
  const void* fn;
  switch (a.modify * 2 + (a.center ? 1 : 0)) {
    case 0: fn = reinterpret_cast<const void*>(&k_perres<0, false>); break;
    case 1: fn = reinterpret_cast<const void*>(&k_perres<0, true>); break;
    case 2: fn = reinterpret_cast<const void*>(&k_perres<1, false>); break;
    case 3: fn = reinterpret_cast<const void*>(&k_perres<1, true>); break;
    case 4: fn = reinterpret_cast<const void*>(&k_perres<2, false>); break;
    case 5: fn = reinterpret_cast<const void*>(&k_perres<2, true>); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  const int64_t batches = (a.nframes + a.nfb - 1) / a.nfb;
  // Enough workgroups to fill the chip a few times over; each walks batches.
  const int64_t want = (8192 + npieces - 1) / npieces;
  const unsigned gx = (unsigned)(batches < want ? batches : want);
  void* args[] = {const_cast<PerresArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(gx, npieces), dim3(PERRES_THREADS), args, lds_bytes, s);
}
