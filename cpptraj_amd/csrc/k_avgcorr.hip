// k_avgcorr.hip -- RMS average correlation ('rmsavgcorr').
//
//   k_avgcorr  Analysis_RmsAvgCorr::Analyze (src/Analysis_RmsAvgCorr.cpp:
//              119-276): for one window size per wavefront, the running sum
//              of the selected atoms over the last `w` frames (Frame
//              operator+= / operator-=, FP64, in the reference's order),
//              divided by w (Frame::Divide), fitted to the reference
//              (Frame::RMSD_CenteredRef); mean and standard deviation of the
//              RMSD over the F - w + 1 averaged frames.  The reference is the
//              centred reference frame, or ('first') the window's first
//              averaged frame, centred.
//
// The running sum lives in LDS ([3][nsel] FP64), and so does the reference
// in 'first' mode (a fixed reference is read from global memory: the same
// for every window, it stays in cache).  The fit is the one of the
// frame-to-reference path: single-pass moments of the averaged frame shifted
// by its first atom against the centred reference, one wave reduction, then
// FP64 QCP.  The QCP solves leave the frame loop's critical path: lane
// (step mod 64) keeps that step's reduced moments, and every 64 steps all
// lanes solve their own step at once.  The mean and SD then sum the RMSDs in
// lane order instead of frame order (a difference in the last bits).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

constexpr int AVG_THREADS = 64;

__device__ __forceinline__ float coord(const FrameSetDev& fs, int64_t f, int k, int c) {
  const int a = fs.sel ? fs.sel[k] : k;
  const float* fr = fs.xyz + frame_of(fs, f) * fs.stride;
  return fs.ld ? fr[c * fs.ld + a] : fr[3 * a + c];
}

// Value j of the 13 moments, summed over the wave, on every lane (read from
// lane 4 j into scalar registers).
__device__ __forceinline__ void wave_moments(const double (&v)[13], double (&out)[13]) {
  double p[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) p[i] = i < 13 ? v[i] : 0.0;
  const double r = rmsd_dev::wave_reduce16(p);  // lane l: total of value l >> 2
  const unsigned long long bits = (unsigned long long)__double_as_longlong(r);
  const unsigned lo = (unsigned)bits, hi = (unsigned)(bits >> 32);
#pragma unroll
  for (int i = 0; i < 13; ++i) {
    const unsigned long long b =
        ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hi, 4 * i) << 32) |
        (unsigned)__builtin_amdgcn_readlane((int)lo, 4 * i);
    out[i] = __longlong_as_double((long long)b);
  }
}

template <bool FIRST, bool WEIGHTED>
__global__ __launch_bounds__(AVG_THREADS) void k_avgcorr(AvgCorrArgs a) {
  extern __shared__ double lds[];
  const int n = a.fs.nsel;
  const int lane = threadIdx.x;
  const int wi = blockIdx.x;
  const int w = a.windows[wi];
  const double dw = (double)w;
  double* S = lds;  // running sum [3][n]
  // centred reference: LDS in 'first' mode, else the prepared reference
  double* Yl = lds + 3 * n;
  const double* Y = FIRST ? Yl : a.ref;
  const int ly = FIRST ? n : a.ldr;
  for (int k = lane; k < n; k += AVG_THREADS)
    for (int c = 0; c < 3; ++c) S[c * n + k] = 0.0;
  // Reference constants: sum w y (3), sum w |y|^2, W (sequential, as the
  // reference's running total).
  double K[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (!FIRST)
    for (int i = 0; i < 5; ++i) K[i] = a.consts[i];
  double W = 0.0;
  if (WEIGHTED) {
    for (int k = 0; k < n; ++k) W += a.mass[k];
  } else {
    W = (double)n;
  }
  __syncthreads();
  bool need_ref = FIRST;
  double sum_r = 0.0, sum_r2 = 0.0;
  // Deferred solves: lane j holds the moments of pending step j (and the
  // reference constants they go with).
  double pm[13], pK[4];
  int npend = 0;
  auto solve_pending = [&]() {
    if (lane < npend) {
      double r;
      if (W < RMSD_SMALL) {
        r = -1.0;
      } else {
        const double c0 = pm[0] / W, c1 = pm[1] / W, c2 = pm[2] / W;
        double R[9];
        R[0] = pm[3] - c0 * pK[0]; R[1] = pm[4] - c0 * pK[1]; R[2] = pm[5] - c0 * pK[2];
        R[3] = pm[6] - c1 * pK[0]; R[4] = pm[7] - c1 * pK[1]; R[5] = pm[8] - c1 * pK[2];
        R[6] = pm[9] - c2 * pK[0]; R[7] = pm[10] - c2 * pK[1]; R[8] = pm[11] - c2 * pK[2];
        const double sx = pm[12] - W * (c0 * c0 + c1 * c1 + c2 * c2);
        const double E0 = 0.5 * (sx + pK[3]);
        r = rmsd_dev::qcp_rmsd(R, E0, W, 1);
      }
      sum_r += r;
      sum_r2 += r * r;
    }
    npend = 0;
  };
  int64_t sub = 0;
  for (int64_t t = 0; t < a.fs.nframes; ++t) {
    for (int k = lane; k < n; k += AVG_THREADS)
      for (int c = 0; c < 3; ++c) S[c * n + k] += (double)coord(a.fs, t, k, c);
    __syncthreads();
    if (t > w - 2) {
      if (FIRST && need_ref) {
        // The first averaged frame, centred (Frame::CenterOnOrigin): the
        // window's reference.
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = lane; k < n; k += AVG_THREADS) {
          const double m = WEIGHTED ? a.mass[k] : 1.0;
          for (int c = 0; c < 3; ++c) v[c] += WEIGHTED ? (S[c * n + k] / dw) * m : S[c * n + k] / dw;
          v[3] += m;
        }
        for (int i = 0; i < 4; ++i) v[i] = rmsd_dev::wave_sum(v[i]);
        double cen[3] = {0.0, 0.0, 0.0};
        const double tot = WEIGHTED ? v[3] : (double)n;
        if (tot != 0.0)
          for (int c = 0; c < 3; ++c) cen[c] = v[c] / tot;
        double h[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = lane; k < n; k += AVG_THREADS) {
          double y[3];
          for (int c = 0; c < 3; ++c) {
            y[c] = S[c * n + k] / dw - cen[c];
            Yl[c * n + k] = y[c];
          }
          const double m = WEIGHTED ? a.mass[k] : 1.0;
          h[0] += m * y[0];
          h[1] += m * y[1];
          h[2] += m * y[2];
          h[3] += m * (y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        }
        for (int i = 0; i < 4; ++i) K[i] = rmsd_dev::wave_sum(h[i]);
        K[4] = W;
        need_ref = false;
        __syncthreads();
      }
      // Moments of the averaged frame (shifted by its first atom) against the
      // reference, then the fit of the frame-to-reference path.
      const double s0 = S[0] / dw, s1 = S[n] / dw, s2 = S[2 * n] / dw;
      double v[13];
#pragma unroll
      for (int i = 0; i < 13; ++i) v[i] = 0.0;
      for (int k = lane; k < n; k += AVG_THREADS) {
        const double x0 = S[k] / dw - s0, x1 = S[n + k] / dw - s1, x2 = S[2 * n + k] / dw - s2;
        const double y0 = Y[k], y1 = Y[ly + k], y2 = Y[2 * ly + k];
        double wx0 = x0, wx1 = x1, wx2 = x2;
        if (WEIGHTED) {
          const double m = a.mass[k];
          wx0 = m * x0;
          wx1 = m * x1;
          wx2 = m * x2;
        }
        v[0] += wx0;
        v[1] += wx1;
        v[2] += wx2;
        v[3] += wx0 * y0; v[4] += wx0 * y1; v[5] += wx0 * y2;
        v[6] += wx1 * y0; v[7] += wx1 * y1; v[8] += wx1 * y2;
        v[9] += wx2 * y0; v[10] += wx2 * y1; v[11] += wx2 * y2;
        v[12] += wx0 * x0 + wx1 * x1 + wx2 * x2;
      }
      double m[13];
      wave_moments(v, m);
      if (lane == npend) {
#pragma unroll
        for (int i = 0; i < 13; ++i) pm[i] = m[i];
#pragma unroll
        for (int i = 0; i < 4; ++i) pK[i] = K[i];
      }
      if (++npend == AVG_THREADS) solve_pending();
      // Drop the oldest frame of the window.
      for (int k = lane; k < n; k += AVG_THREADS)
        for (int c = 0; c < 3; ++c) S[c * n + k] -= (double)coord(a.fs, sub, k, c);
      ++sub;
      __syncthreads();
    }
  }
  solve_pending();
  sum_r = rmsd_dev::wave_sum(sum_r);
  sum_r2 = rmsd_dev::wave_sum(sum_r2);
  if (lane == 0) {
    const double d = 1.0 / ((double)a.fs.nframes - dw + 1.0);
    const double avg = sum_r * d;
    double sd = sum_r2 * d;
    sd -= avg * avg;
    a.ct[wi] = avg;
    a.sd[wi] = sd > 0.0 ? sqrt(sd) : 0.0;
  }
}

}  // namespace

hipError_t rmsd_launch_avgcorr(const AvgCorrArgs& a, bool first, hipStream_t s) {
  if (a.nwin == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (first ? 6 : 3) * (size_t)a.fs.nsel;
  if (lds > AVGCORR_LDS_MAX) return hipErrorInvalidValue;
  const bool wt = a.mass != nullptr;
  const void* fn = first ? (wt ? reinterpret_cast<const void*>(&k_avgcorr<true, true>)
                               : reinterpret_cast<const void*>(&k_avgcorr<true, false>))
                         : (wt ? reinterpret_cast<const void*>(&k_avgcorr<false, true>)
                               : reinterpret_cast<const void*>(&k_avgcorr<false, false>));
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void* args[] = {const_cast<AvgCorrArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(a.nwin), dim3(AVG_THREADS), args, lds, s);
}
