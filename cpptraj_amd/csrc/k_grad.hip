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
//   k_pair_rot    per pair (i, j): covariance, QCP, rotation; stores
//                 c_ij = G_ij / RMSD_ij and c_ij U_ij (recomputed in backward)
//   k_grad_rows / k_grad_cols   the partner sums of the block gradient
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

// One thread per pair (i target, j reference).  cw[(i*FB + j)*10]: c_ij and
// c_ij * U_ij (row-major), c_ij = G_ij / RMSD_ij (0 at RMSD = 0).
__global__ __launch_bounds__(256) void k_pair_rot(const double* XA, const double* XB,
                                                  const double* gA, const double* gB,
                                                  const double* w, int n, int FA, int FB,
                                                  const double* Mp, const double* G, double* cw) {
  const double M = Mp[0];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)FA * FB) return;
  const int64_t i = p / FB, j = p % FB;
  const double* x = XA + i * 3 * n;
  const double* y = XB + j * 3 * n;
  double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < n; ++k) {
    const double m = w ? w[k] : 1.0;
    const double t0 = m * x[k], t1 = m * x[n + k], t2 = m * x[2 * n + k];
    const double r0 = y[k], r1 = y[n + k], r2 = y[2 * n + k];
    S[0] += t0 * r0; S[1] += t0 * r1; S[2] += t0 * r2;
    S[3] += t1 * r0; S[4] += t1 * r1; S[5] += t1 * r2;
    S[6] += t2 * r0; S[7] += t2 * r1; S[8] += t2 * r2;
  }
  const double E0 = 0.5 * (gA[i] + gB[j]);
  const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(S);
  const double lam = rmsd_dev::qcp_lambda(co, E0);
  const double r = rmsd_dev::rmsd_from_lambda(E0, lam, M, 1);
  double U[9];
  rmsd_dev::qcp_rotation(S, lam, U);
  const double c = (r > 1e-12) ? G[p] / r : 0.0;
  double* o = cw + p * 10;
  o[0] = c;
  for (int k = 0; k < 9; ++k) o[1 + k] = c * U[k];
}

// grad_A[i][k] = w_k / M * (sum_j c_ij x~_i,k - sum_j (c_ij U_ij)^T y~_j,k).
// Block = (8 rows i, 256 atoms); the partner loop streams y~_j once per block.
constexpr int GR = 8;
__global__ __launch_bounds__(256) void k_grad_rows(const double* XA, const double* XB,
                                                   const double* cw, const double* w, int n, int FA,
                                                   int FB, const double* Mp, double* grad) {
  __shared__ double sc[GR][10];
  const double M = Mp[0];
  const int i0 = blockIdx.y * GR;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[GR][3], cs[GR];
#pragma unroll
  for (int a = 0; a < GR; ++a) acc[a][0] = acc[a][1] = acc[a][2] = cs[a] = 0.0;
  for (int j = 0; j < FB; ++j) {
    __syncthreads();
    if (threadIdx.x < GR * 10) {
      const int a = threadIdx.x / 10, e = threadIdx.x % 10;
      sc[a][e] = (i0 + a < FA) ? cw[((int64_t)(i0 + a) * FB + j) * 10 + e] : 0.0;
    }
    __syncthreads();
    if (k >= n) continue;
    const double* y = XB + (int64_t)j * 3 * n;
    const double y0 = y[k], y1 = y[n + k], y2 = y[2 * n + k];
#pragma unroll
    for (int a = 0; a < GR; ++a) {
      const double* W = &sc[a][1];  // c U, row-major; (c U)^T y
      acc[a][0] += W[0] * y0 + W[3] * y1 + W[6] * y2;
      acc[a][1] += W[1] * y0 + W[4] * y1 + W[7] * y2;
      acc[a][2] += W[2] * y0 + W[5] * y1 + W[8] * y2;
      cs[a] += sc[a][0];
    }
  }
  if (k >= n) return;
  const double s = (w ? w[k] : 1.0) / M;
#pragma unroll
  for (int a = 0; a < GR; ++a) {
    const int i = i0 + a;
    if (i >= FA) break;
    const double* x = XA + (int64_t)i * 3 * n;
    double* g = grad + ((int64_t)i * n + k) * 3;
    g[0] = s * (cs[a] * x[k] - acc[a][0]);
    g[1] = s * (cs[a] * x[n + k] - acc[a][1]);
    g[2] = s * (cs[a] * x[2 * n + k] - acc[a][2]);
  }
}

// grad_B[j][k] = w_k / M * (sum_i c_ij y~_j,k - sum_i (c_ij U_ij) x~_i,k).
__global__ __launch_bounds__(256) void k_grad_cols(const double* XA, const double* XB,
                                                   const double* cw, const double* w, int n, int FA,
                                                   int FB, const double* Mp, double* grad) {
  __shared__ double sc[GR][10];
  const double M = Mp[0];
  const int j0 = blockIdx.y * GR;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[GR][3], cs[GR];
#pragma unroll
  for (int a = 0; a < GR; ++a) acc[a][0] = acc[a][1] = acc[a][2] = cs[a] = 0.0;
  for (int i = 0; i < FA; ++i) {
    __syncthreads();
    if (threadIdx.x < GR * 10) {
      const int a = threadIdx.x / 10, e = threadIdx.x % 10;
      sc[a][e] = (j0 + a < FB) ? cw[((int64_t)i * FB + (j0 + a)) * 10 + e] : 0.0;
    }
    __syncthreads();
    if (k >= n) continue;
    const double* x = XA + (int64_t)i * 3 * n;
    const double x0 = x[k], x1 = x[n + k], x2 = x[2 * n + k];
#pragma unroll
    for (int a = 0; a < GR; ++a) {
      const double* W = &sc[a][1];  // (c U) x
      acc[a][0] += W[0] * x0 + W[1] * x1 + W[2] * x2;
      acc[a][1] += W[3] * x0 + W[4] * x1 + W[5] * x2;
      acc[a][2] += W[6] * x0 + W[7] * x1 + W[8] * x2;
      cs[a] += sc[a][0];
    }
  }
  if (k >= n) return;
  const double s = (w ? w[k] : 1.0) / M;
#pragma unroll
  for (int a = 0; a < GR; ++a) {
    const int j = j0 + a;
    if (j >= FB) break;
    const double* y = XB + (int64_t)j * 3 * n;
    double* g = grad + ((int64_t)j * n + k) * 3;
    g[0] = s * (cs[a] * y[k] - acc[a][0]);
    g[1] = s * (cs[a] * y[n + k] - acc[a][1]);
    g[2] = s * (cs[a] * y[2 * n + k] - acc[a][2]);
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
  const int64_t np = (int64_t)FA * FB;
  hipLaunchKernelGGL(k_pair_rot, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, XA, XB, gA, gB,
                     w, n, FA, FB, M, G, cw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (gradA) {
    hipLaunchKernelGGL(k_grad_rows, dim3((n + 255) / 256, (FA + GR - 1) / GR), dim3(256), 0, s, XA,
                       XB, cw, w, n, FA, FB, M, gradA);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (gradB) {
    hipLaunchKernelGGL(k_grad_cols, dim3((n + 255) / 256, (FB + GR - 1) / GR), dim3(256), 0, s, XA,
                       XB, cw, w, n, FA, FB, M, gradB);
    e = hipGetLastError();
  }
  return e;
}
