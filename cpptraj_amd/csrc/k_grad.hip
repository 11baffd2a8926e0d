// k_grad.hip -- gradients of best-fit RMSD with respect to the coordinates.
//
// The reference has no derivative of RMSD; this is the envelope-theorem result
// for the optimal superposition.  With x~ the centred target, y~ the centred
// reference, U the rotation mapping x~ onto y~ (Kabsch convention), weights w
// and M = sum w:
//   dRMSD/dx_k = w_k (x~_k - U^T y~_k) / (M RMSD)          (target atoms)
//   dRMSD/dy_k = w_k (y~_k - U x~_k) / (M RMSD)            (reference atoms)
// The centring terms vanish because sum_k w_k (U x~_k - y~_k) = 0 when both
// frames are centred with the covariance weights.  RMSD = 0 is a cusp; the
// gradient is reported as 0 there.  Gradients are FP64, per selected atom,
// AoS [frame][atom][3].
//
//   k_fit_grad    streaming: frame-to-reference fit (rotations recomputed by
//                 the moments + solve kernels of k_fit.hip, not stored)
//   k_center      per frame FP64 centred coordinates and E0 halves
//   k_pair_rot    per 16 x 16 block of pairs (i, j): covariances on FP64
//                 MFMA, then per pair QCP and rotation; stores c_ij = G_ij /
//                 RMSD_ij and c_ij U_ij (recomputed in backward)
//   k_grad_sum    the partner sums of the block gradient (rows, columns) on
//                 FP64 MFMA
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

// xform: 15 doubles per frame (U, t1 = -centroid, t2) from k_fit_solve.
__global__ __launch_bounds__(256) void k_fit_grad(FrameSetDev fs, const double* mass, const double* ref,
                                                  int ldr, const double* xform, const double* rmsd,
                                                  const double* gout, const double* consts,
                                                  double* grad) {
  const int64_t f = blockIdx.y;
  const double W = consts[4];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= fs.nsel) return;
  const double* X = xform + 15 * f;
  const double r = rmsd[f];
  double* g = grad + (f * fs.nsel + k) * 3;
  if (!(r > 1e-12) || W < RMSD_SMALL) {
    g[0] = g[1] = g[2] = 0.0;
    return;
  }
  const double s = gout[f] * (mass ? mass[k] : 1.0) / (W * r);
  float p[3];
  load_atom(fs, f, k, p);
  const double x0 = (double)p[0] + X[9];
  const double x1 = (double)p[1] + X[10];
  const double x2 = (double)p[2] + X[11];
  const double y0 = ref[k], y1 = ref[ldr + k], y2 = ref[2 * ldr + k];
  // U^T y
  const double u0 = X[0] * y0 + X[3] * y1 + X[6] * y2;
  const double u1 = X[1] * y0 + X[4] * y1 + X[7] * y2;
  const double u2 = X[2] * y0 + X[5] * y1 + X[8] * y2;
  g[0] = s * (x0 - u0);
  g[1] = s * (x1 - u1);
  g[2] = s * (x2 - u2);
}

// Centred FP64 coordinates, SoA [F][3][n], and sum w |x~|^2 per frame.
__global__ __launch_bounds__(256) void k_center(FrameSetDev fs, const double* w, double* xc,
                                                double* g) {
  __shared__ double red[4 * 4];
  const int64_t f = blockIdx.x;
  const int n = fs.nsel;
  double v[4] = {0, 0, 0, 0};
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double m = w ? w[k] : 1.0;
    float p[3];
    load_atom(fs, f, k, p);
    v[0] += m * (double)p[0];
    v[1] += m * (double)p[1];
    v[2] += m * (double)p[2];
    v[3] += m;
  }
  rmsd_dev::block_sum<4>(v, red);
  const double tot = w ? v[3] : (double)n;
  const double inv = tot != 0.0 ? 1.0 / tot : 0.0;
  const double c0 = v[0] * inv, c1 = v[1] * inv, c2 = v[2] * inv;
  double e[1] = {0.0};
  double* out = xc + f * 3 * n;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    float p[3];
    load_atom(fs, f, k, p);
    const double x0 = (double)p[0] - c0;
    const double x1 = (double)p[1] - c1;
    const double x2 = (double)p[2] - c2;
    out[k] = x0;
    out[n + k] = x1;
    out[2 * n + k] = x2;
    e[0] += (w ? w[k] : 1.0) * (x0 * x0 + x1 * x1 + x2 * x2);
  }
  rmsd_dev::block_sum<1>(e, red);
  if (threadIdx.x == 0) g[f] = e[0];
}

// One workgroup per 16 x 16 block of pairs (i target, j reference);
// cw[(i*FB + j)*10]: c_ij and c_ij * U_ij (row-major), c_ij = G_ij / RMSD_ij (0
// at RMSD = 0).  The covariances S_ij = sum_k w_k x~_i,k y~_j,k^T of the block
// run on v_mfma_f64_16x16x4f64, nine accumulator blocks (x coordinate c by y
// coordinate d): in atom steps of 16, lane l feeds atoms k0 + 4 (l / 16) + s,
// s = 0..3, of row frame i0 + l % 16 (A, 16 x 4) and of column frame j0 + l % 16
// (B, 4 x 16), four consecutive atoms of a coordinate row (two 16-byte loads
// when n is even).  The four waves take a quarter of the atoms each; their
// partial sums are added in a fixed order through LDS, and wave 0 solves the
// block's pairs: FP64 QCP, rotation, c_ij.  (One thread per pair with a serial
// atom loop took 6.2 ms for 512 x 512 pairs of 2,000 atoms; one wave per block
// over all atoms 0.58 ms.)
typedef __attribute__((ext_vector_type(4))) double f64x4_g;
typedef __attribute__((ext_vector_type(2))) double f64x2_g;
constexpr int ROT_SPLIT = 4;
__global__ __launch_bounds__(64 * ROT_SPLIT) __attribute__((amdgpu_waves_per_eu(2))) void k_pair_rot(
    const double* XA, const double* XB, const double* gA, const double* gB, const double* w, int n,
    int FA, int FB, const double* Mp, const double* G, double* cw) {
  __shared__ double part[2][36][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nbj = (FB + 15) / 16;
  const int64_t blk = blockIdx.x;  // the grid is the block count: no workgroup exits early
  const int64_t i0 = (blk / nbj) * 16, j0 = (blk % nbj) * 16;
  const int64_t fi = i0 + (lane & 15), fj = j0 + (lane & 15);
  const bool iok = fi < FA, jok = fj < FB;
  const double* x = XA + (iok ? fi : 0) * 3 * (int64_t)n;
  const double* y = XB + (jok ? fj : 0) * 3 * (int64_t)n;
  f64x4_g acc[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[c][d] = f64x4_g{0.0, 0.0, 0.0, 0.0};
  const int per = ((n + 15) / 16 + ROT_SPLIT - 1) / ROT_SPLIT * 16;  // atoms per wave
  const int kbeg = wave * per, kend = n < kbeg + per ? n : kbeg + per;
  const int kq = 4 * (lane >> 4);
  const bool pairs = (n & 1) == 0;  // rows of an even length: 16-byte aligned pairs of atoms
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    const int k = k0 + kq;
    double av[4][3], bv[4][3];
    if (pairs && k + 3 < kend) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        f64x2_g a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, b0 = {0.0, 0.0}, b1 = {0.0, 0.0};
        if (iok) {
          a0 = *reinterpret_cast<const f64x2_g*>(x + c * n + k);
          a1 = *reinterpret_cast<const f64x2_g*>(x + c * n + k + 2);
        }
        if (jok) {
          b0 = *reinterpret_cast<const f64x2_g*>(y + c * n + k);
          b1 = *reinterpret_cast<const f64x2_g*>(y + c * n + k + 2);
        }
        av[0][c] = a0.x; av[1][c] = a0.y; av[2][c] = a1.x; av[3][c] = a1.y;
        bv[0][c] = b0.x; bv[1][c] = b0.y; bv[2][c] = b1.x; bv[3][c] = b1.y;
      }
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const double m = w ? w[k + s4] : 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) av[s4][c] *= m;
      }
    } else {
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int ks = k + s4;
#pragma unroll
        for (int c = 0; c < 3; ++c) av[s4][c] = bv[s4][c] = 0.0;
        if (ks < kend) {
          const double m = w ? w[ks] : 1.0;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (iok) av[s4][c] = m * x[c * n + ks];
            if (jok) bv[s4][c] = y[c * n + ks];
          }
        }
      }
    }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          acc[c][d] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s4][c], bv[s4][d], acc[c][d], 0, 0, 0);
  }
  // (acc0 + acc2) + (acc1 + acc3), in that order
  if (wave >= 2)
#pragma unroll
    for (int e = 0; e < 36; ++e) part[wave - 2][e][lane] = acc[e / 12][(e / 4) % 3][e % 4];
  __syncthreads();
  if (wave < 2)
#pragma unroll
    for (int e = 0; e < 36; ++e) acc[e / 12][(e / 4) % 3][e % 4] += part[wave][e][lane];
  __syncthreads();
  if (wave == 1)
#pragma unroll
    for (int e = 0; e < 36; ++e) part[0][e][lane] = acc[e / 12][(e / 4) % 3][e % 4];
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e / 12][(e / 4) % 3][e % 4] += part[0][e][lane];
  const double M = Mp[0];
  // The accumulators hold row (lane / 16) + 4 v, column lane % 16, in item v.
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t i = i0 + (lane >> 4) + 4 * v, j = fj;
    if (i >= FA || !jok) continue;
    double S[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) S[3 * c + d] = acc[c][d][v];
    const int64_t p = i * FB + j;
    const double E0 = 0.5 * (gA[i] + gB[j]);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(S);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    const double r = rmsd_dev::rmsd_from_lambda(E0, lam, M, 1);
    double U[9];
    rmsd_dev::qcp_rotation(S, lam, U);
    const double cc = (r > 1e-12) ? G[p] / r : 0.0;
    double* o = cw + p * 10;
    o[0] = cc;
    for (int k = 0; k < 9; ++k) o[1 + k] = cc * U[k];
  }
}

// The partner sums of the block gradient on FP64 MFMA.  For own frame o, atom
// k and coordinate c:
//   ROWS (own i, partners j):     O[i][c][k] = sum_j sum_b (c U)_ij[b][c] y~_j[b][k]
//   columns (own j, partners i):  O[j][c][k] = sum_i sum_a (c U)_ij[c][a] x~_i[a][k]
//   grad[o][k][c] = w_k / M (cs_o own~_o[c][k] - O[o][c][k]),  cs_o = sum_p c_op.
// One wave per 16 own frames x 32 atoms.  Per step of 4 partners, lane l holds
// the pair record (c, c U) of own frame o0 + l % 16 and partner p0 + l / 16
// (five 16-byte loads): the A operands (16 x 4) of the nine (own coordinate,
// partner coordinate) products; B (4 x 16) is the partner's coordinate row at
// atoms k0 + 16 g + l % 16.  (A thread per atom with the coefficients of one
// partner per LDS stage: 0.58 ms per side for 512 x 512 pairs of 2,000 atoms.)
constexpr int GKT = 2;  // 16-atom tiles per wave (4: 0.97 ms, 1: 0.90 ms for the whole call)
template <bool ROWS>
__global__ __launch_bounds__(256) void k_grad_sum(const double* XA, const double* XB,
                                                  const double* cw, const double* w, int n, int FA,
                                                  int FB, const double* Mp, double* grad) {
  const int lane = threadIdx.x & 63;
  const int NO = ROWS ? FA : FB, NP = ROWS ? FB : FA;  // own frames, partners
  const double* XO = ROWS ? XA : XB;
  const double* XP = ROWS ? XB : XA;
  const int nkt = (n + 16 * GKT - 1) / (16 * GKT);
  const int64_t wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wv >= (int64_t)((NO + 15) / 16) * nkt) return;  // whole waves; no barriers here
  const int o0 = (int)(wv / nkt) * 16, k0 = (int)(wv % nkt) * 16 * GKT;
  const int ol = o0 + (lane & 15);
  const bool ook = ol < NO;
  f64x4_g acc[GKT][3];
#pragma unroll
  for (int g = 0; g < GKT; ++g)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[g][c] = f64x4_g{0.0, 0.0, 0.0, 0.0};
  double csum = 0.0;
  for (int p0 = 0; p0 < NP; p0 += 4) {
    const int pl = p0 + (lane >> 4);
    const bool pok = pl < NP;
    double rec[10];
#pragma unroll
    for (int t = 0; t < 10; ++t) rec[t] = 0.0;
    if (ook && pok) {
      const int64_t pair = ROWS ? (int64_t)ol * FB + pl : (int64_t)pl * FB + ol;
      const f64x2_g* r2 = reinterpret_cast<const f64x2_g*>(cw + pair * 10);
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        const f64x2_g v = r2[t];
        rec[2 * t] = v.x;
        rec[2 * t + 1] = v.y;
      }
    }
    csum += rec[0];
    const double* yp = XP + (int64_t)(pok ? pl : 0) * 3 * n;
#pragma unroll
    for (int g = 0; g < GKT; ++g) {
      const int k = k0 + 16 * g + (lane & 15);
      double bv[3] = {0.0, 0.0, 0.0};
      if (pok && k < n)
#pragma unroll
        for (int b = 0; b < 3; ++b) bv[b] = yp[b * n + k];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          acc[g][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ROWS ? rec[1 + 3 * b + c] : rec[1 + 3 * c + b],
                                                           bv[b], acc[g][c], 0, 0, 0);
    }
  }
  // cs of own frame o0 + r: lanes r, r + 16, r + 32, r + 48
  csum += __shfl_xor(csum, 16, 64);
  csum += __shfl_xor(csum, 32, 64);
  const double M = Mp[0];
  // The accumulators hold own frame o0 + (lane / 16) + 4 v, atom lane % 16, in item v.
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double cs = __shfl(csum, (lane >> 4) + 4 * v, 64);
    const int o = o0 + (lane >> 4) + 4 * v;
    if (o >= NO) continue;
    const double* xo = XO + (int64_t)o * 3 * n;
#pragma unroll
    for (int g = 0; g < GKT; ++g) {
      const int k = k0 + 16 * g + (lane & 15);
      if (k >= n) continue;
      const double s = (w ? w[k] : 1.0) / M;
      double* gp_ = grad + ((int64_t)o * n + k) * 3;
#pragma unroll
      for (int c = 0; c < 3; ++c) gp_[c] = s * (cs * xo[c * n + k] - acc[g][c][v]);
    }
  }
}

}  // namespace

hipError_t rmsd_launch_fit_grad(const FrameSetDev& fs, const double* mass, const double* ref,
                                const double* xform, const double* rmsd, const double* gout,
                                const double* consts, double* grad, hipStream_t s) {
  const int ldr = (fs.nsel + 3) & ~3;
  dim3 grid((fs.nsel + 255) / 256, fs.nframes);
  hipLaunchKernelGGL(k_fit_grad, grid, dim3(256), 0, s, fs, mass, ref, ldr, xform, rmsd, gout,
                     consts, grad);
  return hipGetLastError();
}

hipError_t rmsd_launch_center(const FrameSetDev& fs, const double* w, double* xc, double* g,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_center, dim3(fs.nframes), dim3(256), 0, s, fs, w, xc, g);
  return hipGetLastError();
}

hipError_t rmsd_launch_pair_grad(const double* XA, const double* XB, const double* gA,
                                 const double* gB, const double* w, int n, int FA, int FB,
                                 const double* M, const double* G, double* cw, double* gradA,
                                 double* gradB, hipStream_t s) {
  const int64_t nblk = (int64_t)((FA + 15) / 16) * ((FB + 15) / 16);
  if (nblk > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_rot, dim3((unsigned)nblk), dim3(64 * ROT_SPLIT), 0, s, XA, XB, gA, gB,
                     w, n, FA, FB, M, G, cw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nkt = (n + 16 * GKT - 1) / (16 * GKT);
  if (gradA) {
    const int64_t waves = (int64_t)((FA + 15) / 16) * nkt;
    hipLaunchKernelGGL(k_grad_sum<true>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, XA, XB,
                       cw, w, n, FA, FB, M, gradA);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (gradB) {
    const int64_t waves = (int64_t)((FB + 15) / 16) * nkt;
    hipLaunchKernelGGL(k_grad_sum<false>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, XA, XB,
                       cw, w, n, FA, FB, M, gradB);
    e = hipGetLastError();
  }
  return e;
}
