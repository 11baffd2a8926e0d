// k_fit.hip -- streaming fit of many frames to one reference (the 'rmsd' action).
//
// Replaces the per-frame body of Action_Rmsd::DoAction (src/Action_Rmsd.cpp:
// 361-417) for a batch of frames and a fixed reference ('first', 'ref' and
// 'refindex' modes of ReferenceAction, src/ReferenceAction.h:73-84):
//   k_ref_prep      gather + centre the reference once (ReferenceAction::
//                   SelectRefAtoms, src/ReferenceAction.cpp:155-169)
//   k_fit_moments   one HBM pass over each frame: FP64 sums w x, w x y^T and
//                   w |x|^2 (the reference's centring + covariance + E0 passes,
//                   src/Frame.cpp:1141-1209, folded into one pass)
//   k_fit_solve     per frame FP64 QCP: RMSD, U (Kabsch convention), Trans
//   k_nofit_frames  Frame::RMSD_NoFit (src/Frame.cpp:1279-1307) per frame
//   k_apply         Frame::Trans_Rot_Trans (src/Frame.h:572-581) or the
//                   translate-only update of 'norotate' on all atoms
//
// Single-pass centring: with the reference centred (sum_k w_k y_k = h, not
// necessarily 0 when target and reference masses differ) the covariance of the
// centred target is  R = sum w x y^T - c h^T  and  sum w |x-c|^2 =
// sum w |x|^2 - W |c|^2.  Every frame is first shifted by its first selected
// atom (an exact identity for both formulas) so that the sums carry the
// molecule's extent only, not its absolute position.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

constexpr int FIT_THREADS = 256;

typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float fcoord(const FrameSetDev& fs, int64_t f, int k, int c) {
  const int a = fs.sel ? fs.sel[k] : k;
  const float* fr = fs.xyz + frame_of(fs, f) * fs.stride;
  return fs.ld ? fr[c * fs.ld + a] : fr[3 * a + c];
}

// ---- reference ---------------------------------------------------------------
// ref_out: SoA [3][ldr] FP64 (ldr = nsel rounded up to 4, zero padded).
// consts: [0..2] sum w y, [3] sum w |y|^2, [4] W = sum w (sequential sum, so
// that it rounds exactly as the reference's running total).
__global__ __launch_bounds__(FIT_THREADS) void k_ref_prep(FrameSetDev fs, int f, const double* rmass,
                                                          const double* tmass, int center, int ldr,
                                                          double* ref_out, double* ref_trans,
                                                          double* consts) {
  __shared__ double red[8 * (FIT_THREADS / 64)];
  double v[4] = {0, 0, 0, 0};
  if (center) {
    for (int k = threadIdx.x; k < fs.nsel; k += blockDim.x) {
      const double m = rmass ? rmass[k] : 1.0;
      v[0] += (double)fcoord(fs, f, k, 0) * m;
      v[1] += (double)fcoord(fs, f, k, 1) * m;
      v[2] += (double)fcoord(fs, f, k, 2) * m;
      v[3] += m;
    }
    rmsd_dev::block_sum<4>(v, red);
  }
  double cen[3] = {0, 0, 0};
  if (center) {
    const double tot = rmass ? v[3] : (double)fs.nsel;
    if (tot != 0.0) {
      cen[0] = v[0] / tot;
      cen[1] = v[1] / tot;
      cen[2] = v[2] / tot;
    }
  }
  double h[4] = {0, 0, 0, 0};
  for (int k = threadIdx.x; k < ldr; k += blockDim.x) {
    double y[3] = {0, 0, 0};
    if (k < fs.nsel) {
      y[0] = (double)fcoord(fs, f, k, 0) - cen[0];
      y[1] = (double)fcoord(fs, f, k, 1) - cen[1];
      y[2] = (double)fcoord(fs, f, k, 2) - cen[2];
      const double w = tmass ? tmass[k] : 1.0;
      h[0] += w * y[0];
      h[1] += w * y[1];
      h[2] += w * y[2];
      h[3] += w * (y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    }
    ref_out[k] = y[0];
    ref_out[ldr + k] = y[1];
    ref_out[2 * ldr + k] = y[2];
  }
  rmsd_dev::block_sum<4>(h, red);
  if (threadIdx.x == 0) {
    consts[0] = h[0];
    consts[1] = h[1];
    consts[2] = h[2];
    consts[3] = h[3];
    double W = 0.0;
    if (tmass) {
      for (int k = 0; k < fs.nsel; ++k) W += tmass[k];
    } else {
      W = (double)fs.nsel;
    }
    consts[4] = W;
    ref_trans[0] = cen[0];
    ref_trans[1] = cen[1];
    ref_trans[2] = cen[2];
  }
}

// ---- moments -----------------------------------------------------------------
template <bool WEIGHTED>
__device__ __forceinline__ void accum(double (&v)[13], double x0, double x1, double x2, double w,
                                      double y0, double y1, double y2) {
  double wx0 = x0, wx1 = x1, wx2 = x2;
  if (WEIGHTED) {
    wx0 = w * x0;
    wx1 = w * x1;
    wx2 = w * x2;
  }
  v[0] += wx0;
  v[1] += wx1;
  v[2] += wx2;
  v[3] += wx0 * y0; v[4] += wx0 * y1; v[5] += wx0 * y2;
  v[6] += wx1 * y0; v[7] += wx1 * y1; v[8] += wx1 * y2;
  v[9] += wx2 * y0; v[10] += wx2 * y1; v[11] += wx2 * y2;
  v[12] += wx0 * x0 + wx1 * x1 + wx2 * x2;
}

// Vector path: SoA rows, no selection, 16-byte aligned rows.
template <bool WEIGHTED>
__global__ __launch_bounds__(FIT_THREADS) void k_fit_moments_soa4(FitArgs a, int ldr) {
  __shared__ double red[13 * (FIT_THREADS / 64)];
  const int64_t f = blockIdx.x;
  const float* fr = a.fs.xyz + frame_of(a.fs, f) * a.fs.stride;
  const int64_t ld = a.fs.ld;
  const int n = a.fs.nsel;
  const double s0 = fr[0], s1 = fr[ld], s2 = fr[2 * ld];
  double v[13];
#pragma unroll
  for (int i = 0; i < 13; ++i) v[i] = 0.0;
  const int n4 = n >> 2;
  // Frames are read once: non-temporal loads keep the reference and the
  // weights resident in L2.
  const f32x4* X = reinterpret_cast<const f32x4*>(fr);
  const f32x4* Y = reinterpret_cast<const f32x4*>(fr + ld);
  const f32x4* Z = reinterpret_cast<const f32x4*>(fr + 2 * ld);
  const double2* r0 = reinterpret_cast<const double2*>(a.ref);
  const double2* r1 = reinterpret_cast<const double2*>(a.ref + ldr);
  const double2* r2 = reinterpret_cast<const double2*>(a.ref + 2 * ldr);
  const double2* mw = reinterpret_cast<const double2*>(a.mass);
  for (int q = threadIdx.x; q < n4; q += blockDim.x) {
    const f32x4 x = __builtin_nontemporal_load(X + q);
    const f32x4 y = __builtin_nontemporal_load(Y + q);
    const f32x4 z = __builtin_nontemporal_load(Z + q);
    const double2 ya = r0[2 * q], yb = r0[2 * q + 1];
    const double2 za = r1[2 * q], zb = r1[2 * q + 1];
    const double2 wa = r2[2 * q], wb = r2[2 * q + 1];
    double2 ma = {1.0, 1.0}, mb = {1.0, 1.0};
    if (WEIGHTED) {
      ma = mw[2 * q];
      mb = mw[2 * q + 1];
    }
    accum<WEIGHTED>(v, x.x - s0, y.x - s1, z.x - s2, ma.x, ya.x, za.x, wa.x);
    accum<WEIGHTED>(v, x.y - s0, y.y - s1, z.y - s2, ma.y, ya.y, za.y, wa.y);
    accum<WEIGHTED>(v, x.z - s0, y.z - s1, z.z - s2, mb.x, yb.x, zb.x, wb.x);
    accum<WEIGHTED>(v, x.w - s0, y.w - s1, z.w - s2, mb.y, yb.y, zb.y, wb.y);
  }
  for (int k = 4 * n4 + threadIdx.x; k < n; k += blockDim.x) {
    const double w = WEIGHTED ? a.mass[k] : 1.0;
    accum<WEIGHTED>(v, (double)fr[k] - s0, (double)fr[ld + k] - s1, (double)fr[2 * ld + k] - s2, w,
                    a.ref[k], a.ref[ldr + k], a.ref[2 * ldr + k]);
  }
  rmsd_dev::block_sum<13>(v, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < 13; ++i) a.moments[13 * f + i] = v[i];
}

// Generic path: any layout, optional selection.
template <bool WEIGHTED>
__global__ __launch_bounds__(FIT_THREADS) void k_fit_moments_any(FitArgs a, int ldr) {
  __shared__ double red[13 * (FIT_THREADS / 64)];
  const int64_t f = blockIdx.x;
  const double s0 = fcoord(a.fs, f, 0, 0), s1 = fcoord(a.fs, f, 0, 1), s2 = fcoord(a.fs, f, 0, 2);
  double v[13];
#pragma unroll
  for (int i = 0; i < 13; ++i) v[i] = 0.0;
  for (int k = threadIdx.x; k < a.fs.nsel; k += blockDim.x) {
    const double w = WEIGHTED ? a.mass[k] : 1.0;
    float p[3];
    load_atom(a.fs, f, k, p);
    accum<WEIGHTED>(v, (double)p[0] - s0, (double)p[1] - s1, (double)p[2] - s2, w, a.ref[k],
                    a.ref[ldr + k], a.ref[2 * ldr + k]);
  }
  rmsd_dev::block_sum<13>(v, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < 13; ++i) a.moments[13 * f + i] = v[i];
}

// Wave-per-frame path: the centred reference (and the weights) sit in LDS for
// the whole launch, each wave reduces one frame at a time and walks frames
// f = wave id, + total waves, ...  A frame's 16 sums are combined with the
// butterfly of wave_reduce16, so there is no block barrier per frame; HBM and
// the vector memory pipe carry the frames only (12 B per atom), LDS serves the
// reference (24 or 32 B per atom).  VEC: SoA rows, no selection, 16-byte
// aligned rows, four atoms per lane per step.  Also stores the per-frame shift.
constexpr int WAVE_FIT_THREADS = 1024;
// Frames a wave streams side by side on the VEC path (1 or 2; two measured
// 5 % faster on 1e6 x 3000 frames).
constexpr int kFitFramesPerWave = 2;

template <bool VEC, bool WEIGHTED>
__global__ __launch_bounds__(WAVE_FIT_THREADS) void k_fit_moments_wave(FitArgs a, int ldr,
                                                                       double* shift) {
  extern __shared__ double sref[];  // [3][ldr] centred reference, then [ldr] weights
  const int n = a.fs.nsel;
  for (int i = threadIdx.x; i < 3 * ldr; i += blockDim.x) sref[i] = a.ref[i];
  if (WEIGHTED)
    for (int i = threadIdx.x; i < ldr; i += blockDim.x) sref[3 * ldr + i] = i < n ? a.mass[i] : 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t ld = a.fs.ld;
  const double* Y0 = sref;
  const double* Y1 = sref + ldr;
  const double* Y2 = sref + 2 * ldr;
  const double* Wt = sref + 3 * ldr;
  for (int64_t f = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       f < a.fs.nframes; f += nw) {
    const float* fr = a.fs.xyz + frame_of(a.fs, f) * a.fs.stride;
    const double s0 = fcoord(a.fs, f, 0, 0), s1 = fcoord(a.fs, f, 0, 1), s2 = fcoord(a.fs, f, 0, 2);
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = 0.0;
    double (&m)[13] = *reinterpret_cast<double(*)[13]>(v);
    int k0 = 0;
    if (VEC) {
      const int n4 = n >> 2;
      const f32x4* X = reinterpret_cast<const f32x4*>(fr);
      const f32x4* Y = reinterpret_cast<const f32x4*>(fr + ld);
      const f32x4* Z = reinterpret_cast<const f32x4*>(fr + 2 * ld);
#pragma unroll 4
      for (int q = lane; q < n4; q += 64) {
        const f32x4 x = __builtin_nontemporal_load(X + q);
        const f32x4 y = __builtin_nontemporal_load(Y + q);
        const f32x4 z = __builtin_nontemporal_load(Z + q);
        const int k = 4 * q;
        accum<WEIGHTED>(m, x.x - s0, y.x - s1, z.x - s2, WEIGHTED ? Wt[k] : 1.0, Y0[k], Y1[k], Y2[k]);
        accum<WEIGHTED>(m, x.y - s0, y.y - s1, z.y - s2, WEIGHTED ? Wt[k + 1] : 1.0, Y0[k + 1],
                        Y1[k + 1], Y2[k + 1]);
        accum<WEIGHTED>(m, x.z - s0, y.z - s1, z.z - s2, WEIGHTED ? Wt[k + 2] : 1.0, Y0[k + 2],
                        Y1[k + 2], Y2[k + 2]);
        accum<WEIGHTED>(m, x.w - s0, y.w - s1, z.w - s2, WEIGHTED ? Wt[k + 3] : 1.0, Y0[k + 3],
                        Y1[k + 3], Y2[k + 3]);
      }
      k0 = 4 * n4;
    }
#pragma unroll 4
    for (int k = k0 + lane; k < n; k += 64) {
      float p[3];
      load_atom(a.fs, f, k, p);
      accum<WEIGHTED>(m, (double)p[0] - s0, (double)p[1] - s1, (double)p[2] - s2,
                      WEIGHTED ? Wt[k] : 1.0, Y0[k], Y1[k], Y2[k]);
    }
    const double t = rmsd_dev::wave_reduce16(v);
    const int idx = lane >> 2;
    if ((lane & 3) == 0 && idx < 13) a.moments[13 * f + idx] = t;
    if (lane == 0) {
      shift[3 * f] = s0;
      shift[3 * f + 1] = s1;
      shift[3 * f + 2] = s2;
    }
  }
}

// Two frames per wave (VEC layout only): the frames f and f + nwaves stream
// side by side through the same atom blocks, so a wave keeps twice the loads
// in flight and one LDS read of the reference serves both frames.
constexpr int WAVE2_FIT_THREADS = 512;

// One frame per wave for AoS frames without a selection (cpptraj's own
// layout): a lane takes four consecutive atoms, twelve floats, as three
// 16-byte loads, so consecutive lanes read consecutive 48 bytes.  The
// reference sits in LDS as twelve arrays over atom quads (component d of atom
// 4q + i at [(3i + d) Q + q]), masses as four, so each LDS read is contiguous
// across the wave; a lane reading Y_d[4q + i] from the plain [3][ldr] table
// would stride 32 bytes and conflict on the banks.  The last n mod 4 atoms
// read the reference from global memory.  The frame loads are plain, not
// non-temporal: each of a lane's three 16-byte loads touches the cache lines
// of all three (lanes are 48 bytes apart), and only the vector L1 keeping
// them lets the second and third load hit (1.34 against 1.80 ms for 200,000
// frames x 3,000 atoms with the non-temporal hint).
template <bool WEIGHTED>
__global__ __launch_bounds__(WAVE_FIT_THREADS) void k_fit_moments_aos4(FitArgs a, int ldr,
                                                                       double* shift) {
  extern __shared__ double sref[];
  const int n = a.fs.nsel;
  const int Q = n >> 2;
  for (int t = threadIdx.x; t < 12 * Q; t += blockDim.x) {
    const int r = t / Q, q = t - r * Q;
    const int i = r / 3, d = r - 3 * i;
    sref[t] = a.ref[(size_t)d * ldr + 4 * q + i];
  }
  if (WEIGHTED)
    for (int t = threadIdx.x; t < 4 * Q; t += blockDim.x) {
      const int i = t / Q, q = t - i * Q;
      sref[12 * Q + t] = a.mass[4 * q + i];
    }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const double* Yq = sref;
  const double* Wq = sref + 12 * Q;
  for (int64_t f = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       f < a.fs.nframes; f += nw) {
    const float* fr = a.fs.xyz + frame_of(a.fs, f) * a.fs.stride;
    const double s0 = fr[0], s1 = fr[1], s2 = fr[2];
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = 0.0;
    double (&m)[13] = *reinterpret_cast<double(*)[13]>(v);
    const f32x4* V = reinterpret_cast<const f32x4*>(fr);
#pragma unroll 1  // as in k_fit_moments_wave2 (here 1-2 % faster than by two)
    for (int q = lane; q < Q; q += 64) {
      const f32x4 p0 = V[3 * q];
      const f32x4 p1 = V[3 * q + 1];
      const f32x4 p2 = V[3 * q + 2];
      accum<WEIGHTED>(m, p0.x - s0, p0.y - s1, p0.z - s2, WEIGHTED ? Wq[q] : 1.0, Yq[q],
                      Yq[Q + q], Yq[2 * Q + q]);
      accum<WEIGHTED>(m, p0.w - s0, p1.x - s1, p1.y - s2, WEIGHTED ? Wq[Q + q] : 1.0,
                      Yq[3 * Q + q], Yq[4 * Q + q], Yq[5 * Q + q]);
      accum<WEIGHTED>(m, p1.z - s0, p1.w - s1, p2.x - s2, WEIGHTED ? Wq[2 * Q + q] : 1.0,
                      Yq[6 * Q + q], Yq[7 * Q + q], Yq[8 * Q + q]);
      accum<WEIGHTED>(m, p2.y - s0, p2.z - s1, p2.w - s2, WEIGHTED ? Wq[3 * Q + q] : 1.0,
                      Yq[9 * Q + q], Yq[10 * Q + q], Yq[11 * Q + q]);
    }
    for (int k = 4 * Q + lane; k < n; k += 64)
      accum<WEIGHTED>(m, (double)fr[3 * k] - s0, (double)fr[3 * k + 1] - s1,
                      (double)fr[3 * k + 2] - s2, WEIGHTED ? a.mass[k] : 1.0, a.ref[k],
                      a.ref[ldr + k], a.ref[2 * ldr + k]);
    const double t = rmsd_dev::wave_reduce16(v);
    const int idx = lane >> 2;
    if ((lane & 3) == 0 && idx < 13) a.moments[13 * f + idx] = t;
    if (lane == 0) {
      shift[3 * f] = s0;
      shift[3 * f + 1] = s1;
      shift[3 * f + 2] = s2;
    }
  }
}

template <bool WEIGHTED>
__global__ __launch_bounds__(WAVE2_FIT_THREADS) void k_fit_moments_wave2(FitArgs a, int ldr,
                                                                         double* shift) {
  extern __shared__ double sref[];
  const int n = a.fs.nsel;
  for (int i = threadIdx.x; i < 3 * ldr; i += blockDim.x) sref[i] = a.ref[i];
  if (WEIGHTED)
    for (int i = threadIdx.x; i < ldr; i += blockDim.x) sref[3 * ldr + i] = i < n ? a.mass[i] : 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t ld = a.fs.ld;
  const double* Y0 = sref;
  const double* Y1 = sref + ldr;
  const double* Y2 = sref + 2 * ldr;
  const double* Wt = sref + 3 * ldr;
  const int n4 = n >> 2;
  for (int64_t f = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       f < a.fs.nframes; f += 2 * nw) {
    const int64_t g = f + nw;
    const bool two = g < a.fs.nframes;
    const int64_t g2 = two ? g : f;  // a lone last frame is read twice, stored once
    const float* fa = a.fs.xyz + frame_of(a.fs, f) * a.fs.stride;
    const float* fb = a.fs.xyz + frame_of(a.fs, g2) * a.fs.stride;
    const double sa0 = fa[0], sa1 = fa[ld], sa2 = fa[2 * ld];
    const double sb0 = fb[0], sb1 = fb[ld], sb2 = fb[2 * ld];
    double va[16], vb[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) va[i] = vb[i] = 0.0;
    double (&ma)[13] = *reinterpret_cast<double(*)[13]>(va);
    double (&mb)[13] = *reinterpret_cast<double(*)[13]>(vb);
    const f32x4* XA = reinterpret_cast<const f32x4*>(fa);
    const f32x4* YA = reinterpret_cast<const f32x4*>(fa + ld);
    const f32x4* ZA = reinterpret_cast<const f32x4*>(fa + 2 * ld);
    const f32x4* XB = reinterpret_cast<const f32x4*>(fb);
    const f32x4* YB = reinterpret_cast<const f32x4*>(fb + ld);
    const f32x4* ZB = reinterpret_cast<const f32x4*>(fb + 2 * ld);
// One 16-byte step per loop trip: unrolling by two measured 8 % slower
    // (5.36-5.41 vs 5.82-5.86 TB/s on 500,000 x 3,000; 241 vs 167 registers).
#pragma unroll 1
    for (int q = lane; q < n4; q += 64) {
      const f32x4 xa = __builtin_nontemporal_load(XA + q);
      const f32x4 ya = __builtin_nontemporal_load(YA + q);
      const f32x4 za = __builtin_nontemporal_load(ZA + q);
      const f32x4 xb = __builtin_nontemporal_load(XB + q);
      const f32x4 yb = __builtin_nontemporal_load(YB + q);
      const f32x4 zb = __builtin_nontemporal_load(ZB + q);
      const int k = 4 * q;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double r0 = Y0[k + e], r1 = Y1[k + e], r2 = Y2[k + e];
        const double w = WEIGHTED ? Wt[k + e] : 1.0;
        accum<WEIGHTED>(ma, xa[e] - sa0, ya[e] - sa1, za[e] - sa2, w, r0, r1, r2);
        accum<WEIGHTED>(mb, xb[e] - sb0, yb[e] - sb1, zb[e] - sb2, w, r0, r1, r2);
      }
    }
    for (int k = 4 * n4 + lane; k < n; k += 64) {
      const double r0 = Y0[k], r1 = Y1[k], r2 = Y2[k];
      const double w = WEIGHTED ? Wt[k] : 1.0;
      accum<WEIGHTED>(ma, (double)fa[k] - sa0, (double)fa[ld + k] - sa1,
                      (double)fa[2 * ld + k] - sa2, w, r0, r1, r2);
      accum<WEIGHTED>(mb, (double)fb[k] - sb0, (double)fb[ld + k] - sb1,
                      (double)fb[2 * ld + k] - sb2, w, r0, r1, r2);
    }
    const double ta = rmsd_dev::wave_reduce16(va);
    const double tb = rmsd_dev::wave_reduce16(vb);
    const int idx = lane >> 2;
    if ((lane & 3) == 0 && idx < 13) {
      a.moments[13 * f + idx] = ta;
      if (two) a.moments[13 * g + idx] = tb;
    }
    if (lane == 0) {
      shift[3 * f] = sa0;
      shift[3 * f + 1] = sa1;
      shift[3 * f + 2] = sa2;
      if (two) {
        shift[3 * g] = sb0;
        shift[3 * g + 1] = sb1;
        shift[3 * g + 2] = sb2;
      }
    }
  }
}

// LDS bytes of the wave-per-frame path, 0 when the reference does not fit.
static inline size_t wave_fit_lds(int ldr, bool weighted) {
  const size_t b = sizeof(double) * (size_t)ldr * (weighted ? 4 : 3);
  return b <= 160 * 1024 ? b : 0;
}

// Per-frame shift used by the moment kernels (first selected atom).
__global__ void k_fit_shift(FrameSetDev fs, double* shift) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= fs.nframes) return;
  shift[3 * f] = fcoord(fs, f, 0, 0);
  shift[3 * f + 1] = fcoord(fs, f, 0, 1);
  shift[3 * f + 2] = fcoord(fs, f, 0, 2);
}

// ---- solve ---------------------------------------------------------------------
// xform (15 doubles per frame): U (9), t1 = Trans (3), t2 = refTrans (3).
__global__ void k_fit_solve(SolveArgs a, const double* shift) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nframes) return;
  const double* m = a.moments + 13 * f;
  const double* K = a.consts;
  const double W = K[4];
  double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double t1[3] = {0, 0, 0};
  double r;
  if (W < RMSD_SMALL) {
    r = -1.0;
  } else {
    // centroid in shifted coordinates, then the centred moments
    const double c0 = m[0] / W, c1 = m[1] / W, c2 = m[2] / W;
    double R[9];
    R[0] = m[3] - c0 * K[0]; R[1] = m[4] - c0 * K[1]; R[2] = m[5] - c0 * K[2];
    R[3] = m[6] - c1 * K[0]; R[4] = m[7] - c1 * K[1]; R[5] = m[8] - c1 * K[2];
    R[6] = m[9] - c2 * K[0]; R[7] = m[10] - c2 * K[1]; R[8] = m[11] - c2 * K[2];
    const double sx = m[12] - W * (c0 * c0 + c1 * c1 + c2 * c2);
    const double E0 = 0.5 * (sx + K[3]);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(R);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    r = rmsd_dev::rmsd_from_lambda(E0, lam, W, a.clamp);
    rmsd_dev::qcp_rotation(R, lam, U);
    t1[0] = -(c0 + shift[3 * f]);
    t1[1] = -(c1 + shift[3 * f + 1]);
    t1[2] = -(c2 + shift[3 * f + 2]);
  }
  a.rmsd[f] = r;
  if (a.U)
    for (int k = 0; k < 9; ++k) a.U[9 * f + k] = U[k];
  if (a.tvec) {
    for (int k = 0; k < 3; ++k) {
      a.tvec[6 * f + k] = t1[k];
      a.tvec[6 * f + 3 + k] = a.ref_trans[k];
    }
  }
  if (a.xform) {
    double* x = a.xform + 15 * f;
    for (int k = 0; k < 9; ++k) x[k] = U[k];
    for (int k = 0; k < 3; ++k) {
      x[9 + k] = t1[k];
      x[12 + k] = a.ref_trans[k];
    }
  }
}

// ---- pairs: frame f against reference frame f ---------------------------------
// The 'reftraj' / 'previous' reference modes: every frame has its own
// reference, so one wave per pair (sums reduced within the wave) centres the
// reference (as k_ref_prep),
// then takes the reference's sums (w y, w |y|^2) and the target's shifted
// moments (as the moment kernels) in one pass over both frames, and solves as
// k_fit_solve.  No fit: RMSD_NoFit in the reference's operation order.
__global__ void k_mass_total(const double* mass, int nsel, double* W) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  if (mass) {
    for (int k = 0; k < nsel; ++k) s += mass[k];
  } else {
    s = (double)nsel;
  }
  *W = s;
}

template <bool WEIGHTED, int LAY>
__global__ __launch_bounds__(FIT_THREADS) void k_fit_pairs(PairFitArgs a) {
  // one pair per wave, reductions within the wave (no barriers)
  const int lane = threadIdx.x & 63;
  const int64_t f = (int64_t)blockIdx.x * (FIT_THREADS / 64) + (threadIdx.x >> 6);
  if (f >= a.T.nframes) return;
  const int n = a.T.nsel;
  const double W = a.W[0];
  if (!a.fit) {
#pragma clang fp contract(off)
    double v[1] = {0.0};
    for (int k = lane; k < n; k += 64) {
      const double dx = (double)fcoord(a.R, f, k, 0) - (double)fcoord(a.T, f, k, 0);
      const double dy = (double)fcoord(a.R, f, k, 1) - (double)fcoord(a.T, f, k, 1);
      const double dz = (double)fcoord(a.R, f, k, 2) - (double)fcoord(a.T, f, k, 2);
      const double w = WEIGHTED ? a.tmass[k] : 1.0;
      v[0] += (w * (dx * dx + dy * dy + dz * dz));
    }
    v[0] = rmsd_dev::wave_sum(v[0]);
    if (lane == 0) {
      double r;
      if (W < RMSD_SMALL) r = -1.0;
      else if (v[0] < 0) r = 0.0;
      else r = sqrt(v[0] / W);
      a.rmsd[f] = r;
    }
    return;
  }
  // reference centroid (centring masses); with LAY != 0 four atoms per lane
  // and step from 16-byte loads, then the remaining atoms one by one
  double c4[4] = {0, 0, 0, 0};
  auto cen_add = [&](int k, float r0, float r1, float r2) {
    const double m = a.rmass ? a.rmass[k] : 1.0;
    c4[0] += (double)r0 * m;
    c4[1] += (double)r1 * m;
    c4[2] += (double)r2 * m;
    c4[3] += m;
  };
  const int n4 = LAY ? n >> 2 : 0;
  for (int q = lane; q < n4; q += 64) {
    float x[4], y[4], z[4];
    load_quad(a.R, f, q, LAY, x, y, z);
#pragma unroll
    for (int e = 0; e < 4; ++e) cen_add(4 * q + e, x[e], y[e], z[e]);
  }
  for (int k = 4 * n4 + lane; k < n; k += 64)
    cen_add(k, fcoord(a.R, f, k, 0), fcoord(a.R, f, k, 1), fcoord(a.R, f, k, 2));
#pragma unroll
  for (int i = 0; i < 4; ++i) c4[i] = rmsd_dev::wave_sum(c4[i]);
  double cen[3] = {0, 0, 0};
  const double tot = a.rmass ? c4[3] : (double)n;
  if (tot != 0.0) {
    cen[0] = c4[0] / tot;
    cen[1] = c4[1] / tot;
    cen[2] = c4[2] / tot;
  }
  // target shifted by its first selected atom (an exact identity for the
  // centred moments, so the sums carry the molecule's extent only)
  const double sh[3] = {(double)fcoord(a.T, f, 0, 0), (double)fcoord(a.T, f, 0, 1),
                        (double)fcoord(a.T, f, 0, 2)};
  // v[0..12]: target moments (w x, w x y^T, w |x|^2); v[13..16]: w y, w |y|^2
  double v[17];
#pragma unroll
  for (int i = 0; i < 17; ++i) v[i] = 0.0;
  auto mom_add = [&](int k, const float (&r)[3], const float (&t)[3]) {
    const double y0 = (double)r[0] - cen[0];
    const double y1 = (double)r[1] - cen[1];
    const double y2 = (double)r[2] - cen[2];
    const double w = WEIGHTED ? a.tmass[k] : 1.0;
    double m13[13];
#pragma unroll
    for (int i = 0; i < 13; ++i) m13[i] = v[i];
    accum<WEIGHTED>(m13, (double)t[0] - sh[0], (double)t[1] - sh[1], (double)t[2] - sh[2], w,
                    y0, y1, y2);
#pragma unroll
    for (int i = 0; i < 13; ++i) v[i] = m13[i];
    v[13] += w * y0;
    v[14] += w * y1;
    v[15] += w * y2;
    v[16] += w * (y0 * y0 + y1 * y1 + y2 * y2);
  };
  for (int q = lane; q < n4; q += 64) {
    float rx[4], ry[4], rz[4], tx[4], ty[4], tz[4];
    load_quad(a.R, f, q, LAY, rx, ry, rz);
    load_quad(a.T, f, q, LAY, tx, ty, tz);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float r[3] = {rx[e], ry[e], rz[e]}, t[3] = {tx[e], ty[e], tz[e]};
      mom_add(4 * q + e, r, t);
    }
  }
  for (int k = 4 * n4 + lane; k < n; k += 64) {
    float r[3], t[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = fcoord(a.R, f, k, c);
      t[c] = fcoord(a.T, f, k, c);
    }
    mom_add(k, r, t);
  }
#pragma unroll
  for (int i = 0; i < 17; ++i) v[i] = rmsd_dev::wave_sum(v[i]);
  if (lane != 0) return;
  double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double t1[3] = {0, 0, 0};
  double r;
  if (W < RMSD_SMALL) {
    r = -1.0;
  } else {
    const double c0 = v[0] / W, c1 = v[1] / W, c2 = v[2] / W;
    double R[9];
    R[0] = v[3] - c0 * v[13]; R[1] = v[4] - c0 * v[14]; R[2] = v[5] - c0 * v[15];
    R[3] = v[6] - c1 * v[13]; R[4] = v[7] - c1 * v[14]; R[5] = v[8] - c1 * v[15];
    R[6] = v[9] - c2 * v[13]; R[7] = v[10] - c2 * v[14]; R[8] = v[11] - c2 * v[15];
    const double sx = v[12] - W * (c0 * c0 + c1 * c1 + c2 * c2);
    const double E0 = 0.5 * (sx + v[16]);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(R);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    r = rmsd_dev::rmsd_from_lambda(E0, lam, W, a.clamp);
    rmsd_dev::qcp_rotation(R, lam, U);
    t1[0] = -(c0 + sh[0]);
    t1[1] = -(c1 + sh[1]);
    t1[2] = -(c2 + sh[2]);
  }
  a.rmsd[f] = r;
  if (a.U)
    for (int k = 0; k < 9; ++k) a.U[9 * f + k] = U[k];
  if (a.tvec)
    for (int k = 0; k < 3; ++k) {
      a.tvec[6 * f + k] = t1[k];
      a.tvec[6 * f + 3 + k] = cen[k];
    }
  if (a.xform) {
    double* x = a.xform + 15 * f;
    for (int k = 0; k < 9; ++k) x[k] = U[k];
    for (int k = 0; k < 3; ++k) {
      x[9 + k] = t1[k];
      x[12 + k] = cen[k];
    }
  }
}

// ---- no-fit ------------------------------------------------------------------
// sum w |y - x|^2 accumulated atom by atom per thread in the reference's
// operation order (no contraction), then reduced over the block.
template <bool WEIGHTED>
__global__ __launch_bounds__(FIT_THREADS) void k_nofit_frames(NofitArgs a, int ldr,
                                                              const double* consts, double* rmsd) {
#pragma clang fp contract(off)
  __shared__ double red[FIT_THREADS / 64];
  const int64_t f = blockIdx.x;
  double v[1] = {0.0};
  for (int k = threadIdx.x; k < a.fs.nsel; k += blockDim.x) {
    float p[3];
    load_atom(a.fs, f, k, p);
    const double dx = a.ref[k] - (double)p[0];
    const double dy = a.ref[ldr + k] - (double)p[1];
    const double dz = a.ref[2 * ldr + k] - (double)p[2];
    const double w = WEIGHTED ? a.mass[k] : 1.0;
    v[0] += (w * (dx * dx + dy * dy + dz * dz));
  }
  rmsd_dev::block_sum<1>(v, red);
  if (threadIdx.x == 0) {
    const double W = consts[4];
    double r;
    if (W < RMSD_SMALL) r = -1.0;
    else if (v[0] < 0) r = 0.0;
    else r = sqrt(v[0] / W);
    rmsd[f] = r;
  }
}

// ---- coordinate update -----------------------------------------------------------
// mode 1: x' = U (x + t1) + t2 (Frame::Trans_Rot_Trans); mode 2: x' = x + (t1 + t2).
// The transform comes from xform (15 per frame) or from U (9) + tvec (6); the
// atoms are all natom or the list atoms[0..natoms-1] (Trans_Rot_Trans with an
// AtomMask, as Action_Align's 'move' uses it).
__global__ __launch_bounds__(FIT_THREADS) void k_apply(ApplyArgs a, int mode) {
#pragma clang fp contract(off)
  const int64_t f = blockIdx.x;
  double X[15];
  if (a.xform) {
    for (int k = 0; k < 15; ++k) X[k] = a.xform[15 * f + k];
  } else {
    for (int k = 0; k < 9; ++k) X[k] = a.U ? a.U[9 * f + k] : (k % 4 == 0 ? 1.0 : 0.0);
    for (int k = 0; k < 6; ++k) X[9 + k] = a.tvec[6 * f + k];
  }
  float* fr = a.xyz + f * a.stride;  // all stored frames, in order
  const int nmove = a.atoms ? a.natoms : a.natom;
  for (int i = threadIdx.x; i < nmove; i += blockDim.x) {
    const int k = a.atoms ? a.atoms[i] : i;
    float* px = a.ld ? fr + k : fr + 3 * k;
    float* py = a.ld ? fr + a.ld + k : fr + 3 * k + 1;
    float* pz = a.ld ? fr + 2 * a.ld + k : fr + 3 * k + 2;
    if (mode == 1) {
      const double x = (double)*px + X[9];
      const double y = (double)*py + X[10];
      const double z = (double)*pz + X[11];
      *px = (float)(x * X[0] + y * X[1] + z * X[2] + X[12]);
      *py = (float)(x * X[3] + y * X[4] + z * X[5] + X[13]);
      *pz = (float)(x * X[6] + y * X[7] + z * X[8] + X[14]);
    } else {
      *px = (float)((double)*px + (X[9] + X[12]));
      *py = (float)((double)*py + (X[10] + X[13]));
      *pz = (float)((double)*pz + (X[11] + X[14]));
    }
  }
}

}  // namespace

hipError_t rmsd_launch_ref_prep(const FrameSetDev& fs, int f, const double* rmass,
                                const double* tmass, int center, double* ref_out,
                                double* ref_trans, double* consts, hipStream_t s) {
  const int ldr = (fs.nsel + 3) & ~3;
  hipLaunchKernelGGL(k_ref_prep, dim3(1), dim3(FIT_THREADS), 0, s, fs, f, rmass, tmass, center, ldr,
                     ref_out, ref_trans, consts);
  return hipGetLastError();
}

// moments must hold 13 doubles per frame followed by 3 per frame of shift.
// Box::RotateUcell after the rotate + translate update (src/Action_Rmsd.cpp:
// 381, src/Box.cpp:732-745): the three unit-cell vectors of frame f (rows of
// its 3x3 cell) rotated by the frame's Kabsch U, row k' = U row k, in the
// reference's operation order.
__global__ __launch_bounds__(256) void k_rotate_ucell(const double* U, int64_t nframes,
                                                      double* ucell) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nframes) return;
  double u[9], b[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    u[k] = U[9 * f + k];
    b[k] = ucell[9 * f + k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      ucell[9 * f + 3 * k + c] = __dadd_rn(__dadd_rn(__dmul_rn(u[3 * c], b[3 * k]),
                                                     __dmul_rn(u[3 * c + 1], b[3 * k + 1])),
                                           __dmul_rn(u[3 * c + 2], b[3 * k + 2]));
}

hipError_t rmsd_launch_rotate_ucell(const double* U, int64_t nframes, double* ucell,
                                    hipStream_t s) {
  if (nframes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rotate_ucell, dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, s, U,
                     nframes, ucell);
  return hipGetLastError();
}

hipError_t rmsd_launch_fit_moments(const FitArgs& a, hipStream_t s) {
  const int ldr = (a.fs.nsel + 3) & ~3;
  const int nf = a.fs.nframes;
  double* shift = a.moments + 13 * (int64_t)nf;
  if (nf == 0) return hipSuccess;
  const bool vec = a.fs.sel == nullptr && a.fs.ld > 0 && (a.fs.ld % 4) == 0 &&
                   (a.fs.stride % 4) == 0 && ((uintptr_t)a.fs.xyz % 16) == 0;
  const bool aos4 = a.fs.sel == nullptr && a.fs.ld == 0 && (a.fs.stride % 4) == 0 &&
                    ((uintptr_t)a.fs.xyz % 16) == 0;
  const bool w = a.mass != nullptr;
  // Wave-per-frame path whenever the reference fits in LDS (up to 5120
  // weighted / 6826 unweighted atoms); one-block-per-frame kernels otherwise.
  const size_t lds = wave_fit_lds(ldr, w);
  if (lds && a.variant != 1) {
    int dev = 0, ncu = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t waves = WAVE_FIT_THREADS / 64;
    const int64_t want = (nf + waves - 1) / waves;
    // Blocks loop over frames, so the grid must not exceed what is resident.
#define RMSD_FIT_WAVE(V, W)                                                                   \
  {                                                                                          \
    const void* fn = reinterpret_cast<const void*>(&k_fit_moments_wave<V, W>);               \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                           \
    int bpc = 1;                                                                             \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, WAVE_FIT_THREADS, lds);       \
    if (e != hipSuccess) return e;                                                           \
    if (bpc < 1) bpc = 1;                                                                    \
    const int grid = (int)(want < (int64_t)ncu * bpc ? want : (int64_t)ncu * bpc);           \
    hipLaunchKernelGGL((k_fit_moments_wave<V, W>), dim3(grid), dim3(WAVE_FIT_THREADS), lds, s, a, \
                       ldr, shift);                                                          \
  }
#define RMSD_FIT_WAVE2(W)                                                                     \
  {                                                                                          \
    const void* fn = reinterpret_cast<const void*>(&k_fit_moments_wave2<W>);               \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                           \
    int bpc = 1;                                                                             \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, WAVE2_FIT_THREADS, lds);      \
    if (e != hipSuccess) return e;                                                           \
    if (bpc < 1) bpc = 1;                                                                    \
    const int64_t want2 = (nf + 2 * (WAVE2_FIT_THREADS / 64) - 1) / (2 * (WAVE2_FIT_THREADS / 64)); \
    const int grid = (int)(want2 < (int64_t)ncu * bpc ? want2 : (int64_t)ncu * bpc);         \
    hipLaunchKernelGGL((k_fit_moments_wave2<W>), dim3(grid), dim3(WAVE2_FIT_THREADS), lds, s, a, \
                       ldr, shift);                                                          \
  }
#define RMSD_FIT_AOS4(W)                                                                      \
  {                                                                                          \
    const void* fn = reinterpret_cast<const void*>(&k_fit_moments_aos4<W>);                \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                           \
    int bpc = 1;                                                                             \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, WAVE_FIT_THREADS, lds);       \
    if (e != hipSuccess) return e;                                                           \
    if (bpc < 1) bpc = 1;                                                                    \
    const int grid = (int)(want < (int64_t)ncu * bpc ? want : (int64_t)ncu * bpc);           \
    hipLaunchKernelGGL((k_fit_moments_aos4<W>), dim3(grid), dim3(WAVE_FIT_THREADS), lds, s, a, \
                       ldr, shift);                                                          \
  }
    if (vec && kFitFramesPerWave == 2) {
      if (w) RMSD_FIT_WAVE2(true) else RMSD_FIT_WAVE2(false)
    } else if (aos4) {
      if (w) RMSD_FIT_AOS4(true) else RMSD_FIT_AOS4(false)
    } else if (vec) {
      if (w) RMSD_FIT_WAVE(true, true) else RMSD_FIT_WAVE(true, false)
    } else {
      if (w) RMSD_FIT_WAVE(false, true) else RMSD_FIT_WAVE(false, false)
    }
#undef RMSD_FIT_WAVE
#undef RMSD_FIT_WAVE2
#undef RMSD_FIT_AOS4
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_fit_shift, dim3((nf + 255) / 256), dim3(256), 0, s, a.fs, shift);
  if (vec) {
    if (w) hipLaunchKernelGGL(k_fit_moments_soa4<true>, dim3(nf), dim3(FIT_THREADS), 0, s, a, ldr);
    else hipLaunchKernelGGL(k_fit_moments_soa4<false>, dim3(nf), dim3(FIT_THREADS), 0, s, a, ldr);
  } else {
    if (w) hipLaunchKernelGGL(k_fit_moments_any<true>, dim3(nf), dim3(FIT_THREADS), 0, s, a, ldr);
    else hipLaunchKernelGGL(k_fit_moments_any<false>, dim3(nf), dim3(FIT_THREADS), 0, s, a, ldr);
  }
  return hipGetLastError();
}

hipError_t rmsd_launch_fit_solve(const SolveArgs& a, hipStream_t s) {
  const double* shift = a.moments + 13 * (int64_t)a.nframes;
  hipLaunchKernelGGL(k_fit_solve, dim3((a.nframes + 127) / 128), dim3(128), 0, s, a, shift);
  return hipGetLastError();
}

hipError_t rmsd_launch_nofit_frames(const NofitArgs& a, const double* consts, double* rmsd,
                                    hipStream_t s) {
  const int ldr = (a.fs.nsel + 3) & ~3;
  if (a.mass)
    hipLaunchKernelGGL(k_nofit_frames<true>, dim3(a.fs.nframes), dim3(FIT_THREADS), 0, s, a, ldr,
                       consts, rmsd);
  else
    hipLaunchKernelGGL(k_nofit_frames<false>, dim3(a.fs.nframes), dim3(FIT_THREADS), 0, s, a, ldr,
                       consts, rmsd);
  return hipGetLastError();
}

hipError_t rmsd_launch_mass_total(const double* mass, int nsel, double* W, hipStream_t s) {
  hipLaunchKernelGGL(k_mass_total, dim3(1), dim3(64), 0, s, mass, nsel, W);
  return hipGetLastError();
}

hipError_t rmsd_launch_fit_pairs(const PairFitArgs& a, hipStream_t s) {
  if (a.T.nframes <= 0) return hipSuccess;
  if (a.R.nframes < a.T.nframes || a.R.nsel != a.T.nsel || a.T.nsel <= 0)
    return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((a.T.nframes + FIT_THREADS / 64 - 1) / (FIT_THREADS / 64));
  const int lay = a.fit && quad_layout(a.T) == quad_layout(a.R) ? quad_layout(a.T) : 0;
#define RMSD_FIT_PAIRS(W, L) \
  hipLaunchKernelGGL((k_fit_pairs<W, L>), dim3(grid), dim3(FIT_THREADS), 0, s, a)
  if (a.tmass) {
    if (lay == 1) RMSD_FIT_PAIRS(true, 1);
    else if (lay == 2) RMSD_FIT_PAIRS(true, 2);
    else RMSD_FIT_PAIRS(true, 0);
  } else {
    if (lay == 1) RMSD_FIT_PAIRS(false, 1);
    else if (lay == 2) RMSD_FIT_PAIRS(false, 2);
    else RMSD_FIT_PAIRS(false, 0);
  }
#undef RMSD_FIT_PAIRS
  return hipGetLastError();
}

hipError_t rmsd_launch_apply(const ApplyArgs& a, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3(a.nframes), dim3(FIT_THREADS), 0, s, a, mode);
  return hipGetLastError();
}
