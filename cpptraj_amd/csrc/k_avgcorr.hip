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

// ---- prefix-sum form (fixed reference) --------------------------------------------
// With the selected coordinates x_f and their per-coordinate mean mu, the
// prefix sums P_i = sum_{f < i} (x_f - mu) (i = 0..F, P_0 = 0) give the
// window-w average ending at frame t (i = t + 1, j = i - w) as
// a = (P_i - P_j) / w + mu.  With d = P_i - P_j, v = mu - mu_c (mu_c the
// weighted centroid of mu) and the fit centroid c = (sum m d) / (w W) + mu_c:
//   sum m a y^T       = (Q_i - Q_j) / w + sum m mu y^T
//   sum m |a - c|^2   = sum m |d|^2 / w^2 + 2 (L_i - L_j) / w + sum m |v|^2
//                       - W |(C_i - C_j) / (w W)|^2
// with per-prefix moments C_i = sum m P_i, Q_i = sum m P_i y^T and
// L_i = sum m v . P_i.  Only sum m |d|^2 needs all coordinates of a pair:
// it is a squared distance between two prefix rows, computed from the
// difference itself (no |P_i|^2 + |P_j|^2 - 2 P_i . P_j cancellation), in
// 64 x 64 tiles of (i, j) pairs.  Each tile fits the pairs whose i - j is a
// window size and sums their RMSDs per diagonal; a last pass adds the
// per-tile sums of every window in a fixed order.  Subtracting mu first
// keeps the prefix sums, and the digits their differences lose, small.
namespace {

constexpr int AC_CHUNK = 128;  // frames per scan chunk
constexpr int AC_THREADS = 256;
constexpr int AC_TILE = 64;    // prefix rows per tile side
constexpr int AC_KC = 32;      // coordinates per LDS stage
constexpr int AC_LD = AC_TILE + 1;

__global__ __launch_bounds__(AC_THREADS) void k_ac_chunksum(AvgPrefixArgs a) {
  const int j = blockIdx.x * AC_THREADS + threadIdx.x;
  const int q = blockIdx.y;
  if (j >= a.k) return;
  const int kk = j / 3, c = j - 3 * kk;
  const int64_t f0 = (int64_t)q * AC_CHUNK;
  const int64_t f1 = f0 + AC_CHUNK < a.fs.nframes ? f0 + AC_CHUNK : a.fs.nframes;
  double s = 0.0;
  for (int64_t f = f0; f < f1; ++f) s += (double)coord(a.fs, f, kk, c);
  a.chunk[(int64_t)q * a.k + j] = s;
}

// Coordinate means, and the chunk sums turned into exclusive prefix sums of
// the mean-subtracted coordinates.
__global__ __launch_bounds__(AC_THREADS) void k_ac_scan(AvgPrefixArgs a) {
  const int j = blockIdx.x * AC_THREADS + threadIdx.x;
  if (j >= a.k) return;
  const int nq = (int)((a.fs.nframes + AC_CHUNK - 1) / AC_CHUNK);
  double tot = 0.0;
  for (int q = 0; q < nq; ++q) tot += a.chunk[(int64_t)q * a.k + j];
  const double mu = tot / (double)a.fs.nframes;
  a.mu[j] = mu;
  double off = 0.0;
  for (int q = 0; q < nq; ++q) {
    const int64_t f0 = (int64_t)q * AC_CHUNK;
    const int64_t cnt = f0 + AC_CHUNK < a.fs.nframes ? AC_CHUNK : a.fs.nframes - f0;
    const double s = a.chunk[(int64_t)q * a.k + j];
    a.chunk[(int64_t)q * a.k + j] = off;
    off += s - (double)cnt * mu;
  }
  a.P[j] = 0.0;  // P_0
}

// P rows 1..F, each coordinate times sqrt(mass) when weighted.
__global__ __launch_bounds__(AC_THREADS) void k_ac_prefix(AvgPrefixArgs a) {
  const int j = blockIdx.x * AC_THREADS + threadIdx.x;
  const int q = blockIdx.y;
  if (j >= a.k) return;
  const int kk = j / 3, c = j - 3 * kk;
  const double mu = a.mu[j];
  const double sw = a.mass ? sqrt(a.mass[kk]) : 1.0;
  double run = a.chunk[(int64_t)q * a.k + j];
  const int64_t f0 = (int64_t)q * AC_CHUNK;
  const int64_t f1 = f0 + AC_CHUNK < a.fs.nframes ? f0 + AC_CHUNK : a.fs.nframes;
  for (int64_t f = f0; f < f1; ++f) {
    run += (double)coord(a.fs, f, kk, c) - mu;
    a.P[(f + 1) * a.k + j] = run * sw;
  }
}

// Constants (row F + 1 of rows): sum m mu (3), sum m mu y^T (9),
// sum m |v|^2 (1).  One wave.
__global__ __launch_bounds__(64) void k_ac_consts(AvgPrefixArgs a) {
  const int n = a.fs.nsel;
  const double W = a.consts[4];
  double s[3] = {0.0, 0.0, 0.0};
  for (int kk = threadIdx.x; kk < n; kk += 64) {
    const double m = a.mass ? a.mass[kk] : 1.0;
    for (int c = 0; c < 3; ++c) s[c] += m * a.mu[3 * kk + c];
  }
  double mc[3];
  for (int c = 0; c < 3; ++c) {
    s[c] = rmsd_dev::wave_sum(s[c]);
    mc[c] = W > 0.0 ? s[c] / W : 0.0;
  }
  double v[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) v[q] = 0.0;
  for (int kk = threadIdx.x; kk < n; kk += 64) {
    const double m = a.mass ? a.mass[kk] : 1.0;
    const double y0 = a.ref ? a.ref[kk] : 0.0, y1 = a.ref ? a.ref[a.ldr + kk] : 0.0,
                 y2 = a.ref ? a.ref[2 * a.ldr + kk] : 0.0;
    for (int c = 0; c < 3; ++c) {
      const double mm = m * a.mu[3 * kk + c];
      const double vc = a.mu[3 * kk + c] - mc[c];
      v[3 + 3 * c] += mm * y0;
      v[4 + 3 * c] += mm * y1;
      v[5 + 3 * c] += mm * y2;
      v[12] += m * vc * vc;
    }
  }
  double r[13];
  wave_moments(v, r);
  if (threadIdx.x == 0) {
    double* o = a.rows + (a.fs.nframes + 1) * 13;
    for (int c = 0; c < 3; ++c) o[c] = s[c];
    for (int q = 3; q < 13; ++q) o[q] = r[q];
  }
}

// Per prefix row i (0..F): C_i (3), Q_i (9), L_i (1).  One wave per row.
__global__ __launch_bounds__(64) void k_ac_rows(AvgPrefixArgs a) {
  const int64_t i = blockIdx.x;
  const int n = a.fs.nsel;
  const double W = a.consts[4];
  const double* cr = a.rows + (a.fs.nframes + 1) * 13;
  double mc[3];
  for (int c = 0; c < 3; ++c) mc[c] = W > 0.0 ? cr[c] / W : 0.0;
  const double* Pi = a.P + i * a.k;
  double v[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) v[q] = 0.0;
  for (int kk = threadIdx.x; kk < n; kk += 64) {
    const double sw = a.mass ? sqrt(a.mass[kk]) : 1.0;
    const double y0 = a.ref ? a.ref[kk] : 0.0, y1 = a.ref ? a.ref[a.ldr + kk] : 0.0,
                 y2 = a.ref ? a.ref[2 * a.ldr + kk] : 0.0;
    for (int c = 0; c < 3; ++c) {
      const double mp = sw * Pi[3 * kk + c];  // m P
      v[c] += mp;
      v[3 + 3 * c] += mp * y0;
      v[4 + 3 * c] += mp * y1;
      v[5 + 3 * c] += mp * y2;
      v[12] += mp * (a.mu[3 * kk + c] - mc[c]);
    }
  }
  double r[13];
  wave_moments(v, r);
  if (threadIdx.x == 0)
    for (int q = 0; q < 13; ++q) a.rows[i * 13 + q] = r[q];
}

// One 64 x 64 tile of (i, j) prefix-row pairs.  Thread (ty, tx) = (t / 16,
// t % 16) owns rows i0 + ty + 16 a and j0 + tx + 16 b (a, b < 4).
__global__ __launch_bounds__(AC_THREADS) void k_ac_tiles(AvgPrefixArgs a) {
  // [2][KC][65] staged coordinates (odd row stride: conflict-free transposed
  // stores), later the RMSDs [64][65].
  __shared__ double sm[2 * AC_KC * AC_LD];
  static_assert(2 * AC_KC * AC_LD >= AC_TILE * AC_LD, "RMSD tile does not fit");
  double* si = sm;
  double* sj = sm + AC_KC * AC_LD;
  const int2 tile = a.tiles[blockIdx.x];
  const int ti = tile.x, tj = tile.y;
  const int64_t i0 = (int64_t)ti * AC_TILE, j0 = (int64_t)tj * AC_TILE;
  const int64_t nrow = a.fs.nframes + 1;
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  double acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
  // Staging: thread t loads coordinate (t % 32) of rows (t / 32) + 8 r.
  const int lc = t & (AC_KC - 1), lr = t >> 5;
  for (int k0 = 0; k0 < a.k; k0 += AC_KC) {
    const int kc = k0 + lc;
#pragma unroll
    for (int r = 0; r < AC_TILE / 8; ++r) {
      const int row = lr + 8 * r;
      const int64_t gi = i0 + row, gj = j0 + row;
      si[lc * AC_LD + row] = (kc < a.k && gi < nrow) ? a.P[gi * a.k + kc] : 0.0;
      sj[lc * AC_LD + row] = (kc < a.k && gj < nrow) ? a.P[gj * a.k + kc] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < AC_KC; ++c) {
      double pi[4], pj[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        pi[p] = si[c * AC_LD + ty + 16 * p];
        pj[p] = sj[c * AC_LD + tx + 16 * p];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double d = pi[p] - pj[q];
          acc[p][q] = fma(d, d, acc[p][q]);
        }
    }
    __syncthreads();
  }
  // Fits of the pairs on window diagonals.
  const double* K = a.consts;  // sum w y (3), sum w |y|^2, W
  const double W = K[4];
  const double* cm = a.rows + nrow * 13;
  double* rr = sm;
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) {
      const int il = ty + 16 * p, jl = tx + 16 * q;
      const int64_t i = i0 + il, j = j0 + jl;
      const int64_t w = i - j;
      double r = 0.0;
      if (w >= 1 && i < nrow && a.wmap[w] >= 0) {
        if (W < RMSD_SMALL) {
          r = -1.0;
        } else {
          const double dw = (double)w;
          const double* ri = a.rows + i * 13;
          const double* rj = a.rows + j * 13;
          double dc[3], c[3];
          for (int e = 0; e < 3; ++e) {
            dc[e] = (ri[e] - rj[e]) / (dw * W);
            c[e] = dc[e] + cm[e] / W;
          }
          double R[9];
          for (int e = 0; e < 9; ++e)
            R[e] = (ri[3 + e] - rj[3 + e]) / dw + cm[3 + e] - c[e / 3] * K[e % 3];
          const double sx = acc[p][q] / (dw * dw) + 2.0 * (ri[12] - rj[12]) / dw + cm[12] -
                            W * (dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
          const double E0 = 0.5 * (sx + K[3]);
          r = rmsd_dev::qcp_rmsd(R, E0, W, 1);
        }
      }
      rr[il * AC_LD + jl] = r;
    }
  __syncthreads();
  // Diagonal sums: thread dd + 63 takes local diagonal dd = il - jl.
  if (t < 2 * AC_TILE - 1) {
    const int dd = t - (AC_TILE - 1);
    const int64_t w = (i0 - j0) + dd;
    if (w >= 1 && w < nrow && a.wmap[w] >= 0) {
      double s1 = 0.0, s2 = 0.0;
      const int il0 = dd > 0 ? dd : 0, il1 = dd < 0 ? AC_TILE + dd : AC_TILE;
      for (int il = il0; il < il1; ++il) {
        const double r = rr[il * AC_LD + il - dd];
        s1 += r;
        s2 += r * r;
      }
      // Slot of this tile among the two j tiles that diagonal w crosses in
      // row tile ti.
      const int half = tj - (ti - (int)((w + AC_TILE - 1) / AC_TILE));
      double* o = a.part + (((int64_t)a.wmap[w] * a.ntile + ti) * 2 + half) * 2;
      o[0] = s1;
      o[1] = s2;
    }
  }
}

// Mean and standard deviation per window from the per-tile sums, in a fixed
// order.  One wave per window.
__global__ __launch_bounds__(64) void k_ac_final(AvgPrefixArgs a) {
  const int wi = blockIdx.x;
  const int w = a.windows[wi];
  double s1 = 0.0, s2 = 0.0;
  for (int slot = threadIdx.x; slot < 2 * a.ntile; slot += 64) {
    const int ti = slot >> 1, half = slot & 1;
    // The tile (ti, tj) that holds this slot; written when it exists and
    // diagonal w crosses it (a multiple of 64 crosses one tile per row tile).
    const int T = a.tile;
    const int tj = ti - (w + T - 1) / T + half;
    if (tj < 0 || tj > ti) continue;
    const int off = w - T * (ti - tj);
    if (off < 1 - T || off > T - 1) continue;
    const double* o = a.part + (((int64_t)wi * a.ntile + ti) * 2 + half) * 2;
    s1 += o[0];
    s2 += o[1];
  }
  s1 = rmsd_dev::wave_sum(s1);
  s2 = rmsd_dev::wave_sum(s2);
  if (threadIdx.x == 0) {
    const double dw = (double)w;
    const double d = 1.0 / ((double)a.fs.nframes - dw + 1.0);
    const double avg = s1 * d;
    double sd = s2 * d;
    sd -= avg * avg;
    a.ct[wi] = avg;
    a.sd[wi] = sd > 0.0 ? sqrt(sd) : 0.0;
  }
}

// ---- prefix form, 'first' reference --------------------------------------------------
// Window w's reference is its first averaged frame, centred:
// y_w = P_w / w + mu - c_w.  An averaged frame's cross moment is then
//   sum m a y_w^T = (1/w) sum m (P_i - P_j) y_w^T + sum m mu y_w^T,
// linear in the prefix difference (its rounding stays relative to the
// difference, as for the fixed reference).  k_ac_firstref writes sqrt(m) y_w
// per window and its constants (sum m y_w, sum m |y_w|^2, sum m mu y_w^T);
// k_ac_tiles_first runs 32 x 32 tiles of prefix-row pairs with the window
// references of the tile's 63 diagonals staged beside the rows, 2 x 2 pairs
// per thread (a squared distance and nine cross sums each).
constexpr int ACF_TILE = 32;
constexpr int ACF_KA = 8;                // atoms per LDS stage
constexpr int ACF_KC = 3 * ACF_KA;       // coordinates per stage
constexpr int ACF_LD = ACF_TILE + 1;
constexpr int ACF_NDIAG = 2 * ACF_TILE - 1;

__global__ __launch_bounds__(64) void k_ac_firstref(AvgPrefixArgs a) {
  const int wi = blockIdx.x;
  const int w = a.windows[wi];
  const double dw = (double)w;
  const int n = a.fs.nsel;
  const double W = a.consts[4];
  const double* cm = a.rows + (a.fs.nframes + 1) * 13;
  const double* Cw = a.rows + (int64_t)w * 13;
  double cw[3];
  for (int e = 0; e < 3; ++e) cw[e] = W > 0.0 ? (Cw[e] / dw + cm[e]) / W : 0.0;
  const double* Pw = a.P + (int64_t)w * a.k;
  double* ys = a.yref + (int64_t)wi * a.k;
  double v[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) v[q] = 0.0;
  for (int kk = threadIdx.x; kk < n; kk += 64) {
    const double m = a.mass ? a.mass[kk] : 1.0;
    const double sw = a.mass ? sqrt(m) : 1.0;
    double y[3];
    for (int e = 0; e < 3; ++e) {
      const double pw = sw > 0.0 ? Pw[3 * kk + e] / sw : 0.0;
      y[e] = pw / dw + a.mu[3 * kk + e] - cw[e];
      ys[3 * kk + e] = sw * y[e];
      v[e] += m * y[e];
      v[3] += m * y[e] * y[e];
    }
    for (int c = 0; c < 3; ++c) {
      const double mm = m * a.mu[3 * kk + c];
      for (int e = 0; e < 3; ++e) v[4 + 3 * c + e] += mm * y[e];
    }
  }
  double r[13];
  wave_moments(v, r);
  if (threadIdx.x == 0)
    for (int q = 0; q < 13; ++q) a.wconst[(int64_t)wi * 13 + q] = r[q];
}

__global__ __launch_bounds__(AC_THREADS) void k_ac_tiles_first(AvgPrefixArgs a) {
  // [2][KC][33] staged rows (later the RMSDs [32][33]) and [63][KC] window
  // references.
  __shared__ double sij[2 * ACF_KC * ACF_LD];
  __shared__ double sy[ACF_NDIAG * ACF_KC];
  static_assert(2 * ACF_KC * ACF_LD >= ACF_TILE * ACF_LD, "RMSD tile does not fit");
  double* si = sij;
  double* sj = sij + ACF_KC * ACF_LD;
  const int2 tile = a.tiles[blockIdx.x];
  const int ti = tile.x, tj = tile.y;
  const int64_t i0 = (int64_t)ti * ACF_TILE, j0 = (int64_t)tj * ACF_TILE;
  const int64_t base = i0 - j0;
  const int64_t nrow = a.fs.nframes + 1;
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  double D[2][2], X[2][2][9];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      D[p][q] = 0.0;
#pragma unroll
      for (int e = 0; e < 9; ++e) X[p][q][e] = 0.0;
    }
  // Diagonal of each of the thread's pairs (index into sy).
  int dg[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) dg[p][q] = (ty + 16 * p) - (tx + 16 * q) + (ACF_TILE - 1);
  for (int k0 = 0; k0 < a.k; k0 += ACF_KC) {
    // Rows: 32 x 24 coordinates each for i and j.
    for (int e = t; e < ACF_TILE * ACF_KC; e += AC_THREADS) {
      const int row = e / ACF_KC, kc = e - row * ACF_KC;
      const int64_t gi = i0 + row, gj = j0 + row;
      const int kg = k0 + kc;
      si[kc * ACF_LD + row] = (kg < a.k && gi < nrow) ? a.P[gi * a.k + kg] : 0.0;
      sj[kc * ACF_LD + row] = (kg < a.k && gj < nrow) ? a.P[gj * a.k + kg] : 0.0;
    }
    // Window references of the 63 diagonals.
    for (int e = t; e < ACF_NDIAG * ACF_KC; e += AC_THREADS) {
      const int d = e / ACF_KC, kc = e - d * ACF_KC;
      const int64_t w = base + d - (ACF_TILE - 1);
      const int kg = k0 + kc;
      double v = 0.0;
      if (w >= 1 && w < nrow && kg < a.k) {
        const int wi = a.wmap[w];
        if (wi >= 0) v = a.yref[(int64_t)wi * a.k + kg];
      }
      sy[d * ACF_KC + kc] = v;
    }
    __syncthreads();
#pragma unroll
    for (int at = 0; at < ACF_KA; ++at) {
      double pi[2][3], pj[2][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          pi[p][c] = si[(3 * at + c) * ACF_LD + ty + 16 * p];
          pj[p][c] = sj[(3 * at + c) * ACF_LD + tx + 16 * p];
        }
      // The pairs (0, 0) and (1, 1) lie on one diagonal: three references.
      double yv[3][3];
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const double* y = sy + (dg[0][0] + 16 * (u - 1)) * ACF_KC + 3 * at;
        yv[u][0] = y[0];
        yv[u][1] = y[1];
        yv[u][2] = y[2];
      }
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int u = p - q + 1;
          const double y0 = yv[u][0], y1 = yv[u][1], y2 = yv[u][2];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double d = pi[p][c] - pj[q][c];
            D[p][q] = fma(d, d, D[p][q]);
            X[p][q][3 * c] = fma(d, y0, X[p][q][3 * c]);
            X[p][q][3 * c + 1] = fma(d, y1, X[p][q][3 * c + 1]);
            X[p][q][3 * c + 2] = fma(d, y2, X[p][q][3 * c + 2]);
          }
        }
    }
    __syncthreads();
  }
  const double W = a.consts[4];
  const double* cm = a.rows + nrow * 13;
  double* rr = sij;  // [32][33] RMSDs, over the staging
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q) {
      const int il = ty + 16 * p, jl = tx + 16 * q;
      const int64_t i = i0 + il, j = j0 + jl;
      const int64_t w = i - j;
      double r = 0.0;
      if (w >= 1 && i < nrow && a.wmap[w] >= 0) {
        if (W < RMSD_SMALL) {
          r = -1.0;
        } else {
          const int wi = a.wmap[w];
          const double dw = (double)w;
          const double* ri = a.rows + i * 13;
          const double* rj = a.rows + j * 13;
          const double* K = a.wconst + (int64_t)wi * 13;  // sum m y (3), sum m |y|^2, M (9)
          double dc[3], c[3];
          for (int e = 0; e < 3; ++e) {
            dc[e] = (ri[e] - rj[e]) / (dw * W);
            c[e] = dc[e] + cm[e] / W;
          }
          double R[9];
          for (int e = 0; e < 9; ++e) R[e] = X[p][q][e] / dw + K[4 + e] - c[e / 3] * K[e % 3];
          const double sx = D[p][q] / (dw * dw) + 2.0 * (ri[12] - rj[12]) / dw + cm[12] -
                            W * (dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
          const double E0 = 0.5 * (sx + K[3]);
          r = rmsd_dev::qcp_rmsd(R, E0, W, 1);
        }
      }
      rr[il * ACF_LD + jl] = r;
    }
  __syncthreads();
  if (t < ACF_NDIAG) {
    const int dd = t - (ACF_TILE - 1);
    const int64_t w = base + dd;
    if (w >= 1 && w < nrow && a.wmap[w] >= 0) {
      double s1 = 0.0, s2 = 0.0;
      const int il0 = dd > 0 ? dd : 0, il1 = dd < 0 ? ACF_TILE + dd : ACF_TILE;
      for (int il = il0; il < il1; ++il) {
        const double r = rr[il * ACF_LD + il - dd];
        s1 += r;
        s2 += r * r;
      }
      const int half = tj - (ti - (int)((w + ACF_TILE - 1) / ACF_TILE));
      double* o = a.part + (((int64_t)a.wmap[w] * a.ntile + ti) * 2 + half) * 2;
      o[0] = s1;
      o[1] = s2;
    }
  }
}

// Total weight for the 'first' form (consts[4]; the rest unused): one wave.
__global__ __launch_bounds__(64) void k_ac_wsum(AvgPrefixArgs a, double* consts) {
  double s = 0.0;
  for (int kk = threadIdx.x; kk < a.fs.nsel; kk += 64) s += a.mass ? a.mass[kk] : 1.0;
  s = rmsd_dev::wave_sum(s);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) consts[q] = 0.0;
    consts[4] = s;
  }
}

}  // namespace

int rmsd_avgcorr_chunks(int64_t nframes) { return (int)((nframes + AC_CHUNK - 1) / AC_CHUNK); }
int rmsd_avgcorr_tile(bool first) { return first ? ACF_TILE : AC_TILE; }

hipError_t rmsd_launch_avgcorr_prefix(const AvgPrefixArgs& a, hipStream_t s) {
  if (a.nwin == 0 || a.ntiles == 0) return hipSuccess;
  const int nq = rmsd_avgcorr_chunks(a.fs.nframes);
  const dim3 gj((a.k + AC_THREADS - 1) / AC_THREADS, nq);
  hipLaunchKernelGGL(k_ac_chunksum, gj, dim3(AC_THREADS), 0, s, a);
  hipLaunchKernelGGL(k_ac_scan, dim3(gj.x), dim3(AC_THREADS), 0, s, a);
  hipLaunchKernelGGL(k_ac_prefix, gj, dim3(AC_THREADS), 0, s, a);
  hipLaunchKernelGGL(k_ac_consts, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_ac_rows, dim3((unsigned)(a.fs.nframes + 1)), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t rmsd_launch_avgcorr_tiles(const AvgPrefixArgs& a, hipStream_t s) {
  if (a.nwin == 0 || a.ntiles == 0) return hipSuccess;
  if (a.yref) {
    hipLaunchKernelGGL(k_ac_firstref, dim3(a.nwin), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_ac_tiles_first, dim3(a.ntiles), dim3(AC_THREADS), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_ac_tiles, dim3(a.ntiles), dim3(AC_THREADS), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t rmsd_launch_avgcorr_wsum(const AvgPrefixArgs& a, double* consts, hipStream_t s) {
  hipLaunchKernelGGL(k_ac_wsum, dim3(1), dim3(64), 0, s, a, consts);
  return hipGetLastError();
}

hipError_t rmsd_launch_avgcorr_final(const AvgPrefixArgs& a, hipStream_t s) {
  if (a.nwin == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ac_final, dim3(a.nwin), dim3(64), 0, s, a);
  return hipGetLastError();
}
