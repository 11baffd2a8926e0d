// rmsd_device.hpp -- device-side FP64 math shared by the RMSD kernels (gfx950).
//
// * qcp_lambda / qcp_rmsd: largest eigenvalue of the QCP key matrix by Newton
//   iteration from E0, following the quaternion characteristic polynomial of
//   Theobald (Acta Cryst. A61:478, 2005) and Liu, Agrafiotis & Theobald
//   (J. Comput. Chem. 31:1561, 2010), with the same coefficient formulas,
//   Newton start, stopping rule (|step| < 1e-11 lambda, <= 50 steps) and
//   adjugate-column eigenvector fallbacks as the reference's
//   Cpptraj::QCPRot::FastCalcRMSDAndRotation (src/qcprot.cpp:167-343).
// * qcp_rotation: the optimal rotation from the same eigenvector, returned in
//   the Kabsch convention of Frame::RMSD_CenteredRef (src/Frame.cpp:1247-1258):
//   x' = U x maps the centred target onto the centred reference.
// * wave / block reductions for doubles on 64-wide wavefronts.
//
// The QCP routines below restate the published algorithm of qcprot (version 1.5),
// which carries this notice:
//
//   Copyright (c) 2009-2016 Pu Liu and Douglas L. Theobald
//   All rights reserved.
//
//   Redistribution and use in source and binary forms, with or without
//   modification, are permitted provided that the following conditions are met:
//
//   * Redistributions of source code must retain the above copyright notice, this
//     list of conditions and the following disclaimer.
//   * Redistributions in binary form must reproduce the above copyright notice,
//     this list of conditions and the following disclaimer in the documentation
//     and/or other materials provided with the distribution.
//   * Neither the name of the <ORGANIZATION> nor the names of its contributors may
//     be used to endorse or promote products derived from this software without
//     specific prior written permission.
//
//   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
//   AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
//   IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE ARE
//   DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER OR CONTRIBUTORS BE LIABLE
//   FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL
//   DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR
//   SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER
//   CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY,
//   OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE
//   OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RMSD_SMALL 1.0e-14  // Constants::SMALL, src/Constants.h:35

namespace rmsd_dev {

struct QcpCoef {
  double c0, c1, c2;
};

// Quartic coefficients of the QCP key matrix for covariance S (row-major,
// S_ij = sum_k w_k t_ki r_kj).
__device__ __forceinline__ QcpCoef qcp_coefficients(const double* S) {
  const double Sxx = S[0], Sxy = S[1], Sxz = S[2];
  const double Syx = S[3], Syy = S[4], Syz = S[5];
  const double Szx = S[6], Szy = S[7], Szz = S[8];
  const double Sxx2 = Sxx * Sxx, Syy2 = Syy * Syy, Szz2 = Szz * Szz;
  const double Sxy2 = Sxy * Sxy, Syz2 = Syz * Syz, Sxz2 = Sxz * Sxz;
  const double Syx2 = Syx * Syx, Szy2 = Szy * Szy, Szx2 = Szx * Szx;
  const double pYZ = 2.0 * (Syz * Szy - Syy * Szz);
  const double qD = Syy2 + Szz2 - Sxx2 + Syz2 + Szy2;
  QcpCoef k;
  k.c2 = -2.0 * (Sxx2 + Syy2 + Szz2 + Sxy2 + Syx2 + Sxz2 + Szx2 + Syz2 + Szy2);
  k.c1 = 8.0 * (Sxx * Syz * Szy + Syy * Szx * Sxz + Szz * Sxy * Syx -
                Sxx * Syy * Szz - Syz * Szx * Sxy - Szy * Syx * Sxz);
  const double xzP = Sxz + Szx, yzP = Syz + Szy, xyP = Sxy + Syx;
  const double yzM = Syz - Szy, xzM = Sxz - Szx, xyM = Sxy - Syx;
  const double xxPyy = Sxx + Syy, xxMyy = Sxx - Syy;
  const double od = Sxy2 + Sxz2 - Syx2 - Szx2;
  k.c0 = od * od + (qD + pYZ) * (qD - pYZ) +
         (-(xzP) * (yzM) + (xyM) * (xxMyy - Szz)) * (-(xzM) * (yzP) + (xyM) * (xxMyy + Szz)) +
         (-(xzP) * (yzP) - (xyP) * (xxPyy - Szz)) * (-(xzM) * (yzM) - (xyP) * (xxPyy + Szz)) +
         (+(xyP) * (yzP) + (xzP) * (xxMyy + Szz)) * (-(xyM) * (yzM) + (xzP) * (xxPyy + Szz)) +
         (+(xyP) * (yzM) + (xzM) * (xxMyy - Szz)) * (-(xyM) * (yzP) + (xzM) * (xxPyy - Szz));
  return k;
}

// Newton iteration for the largest root, started at E0.
__device__ __forceinline__ double qcp_lambda(const QcpCoef& k, double E0) {
  double lam = E0;
  for (int it = 0; it < 50; ++it) {
    const double prev = lam;
    const double l2 = lam * lam;
    const double b = (l2 + k.c2) * lam;
    const double a = b + k.c1;
    lam -= (a * lam + k.c0) / (2.0 * l2 * lam + b + a);
    if (fabs(lam - prev) < fabs(1e-11 * lam)) break;
  }
  return lam;
}

// RMSD from E0 - lambda.  clamp != 0 reproduces the Kabsch path's
// "negative -> 0" (src/Frame.cpp:1262-1267), which also yields 0 where every
// atom sits on the centroid (one atom: E0 = 0 and Newton's 0/0); otherwise the
// QCP path's fabs (src/qcprot.cpp:246), NaN included.
__device__ __forceinline__ double rmsd_from_lambda(double E0, double lam, double M,
                                                   int clamp) {
  const double e = E0 - lam;
  if (clamp) return e > 0.0 ? sqrt((2.0 * e) / M) : 0.0;
  return sqrt(fabs(2.0 * e / M));
}

__device__ __forceinline__ double qcp_rmsd(const double* S, double E0, double M,
                                           int clamp) {
  const QcpCoef k = qcp_coefficients(S);
  return rmsd_from_lambda(E0, qcp_lambda(k, E0), M, clamp);
}

// The same solve with fewer FP64 instructions, for the MFMA-reduced pairwise
// forms, whose results carry a stated error bound against FP64 anyway: each
// Newton quotient is the numerator times the hardware reciprocal of the
// denominator after one Newton-Raphson refinement (relative error a few units
// of 2^-53, against the IEEE division's half unit; Newton's iteration corrects
// an inexact step, and the stopping rule is unchanged), and 1 / M is taken
// once per launch.  The result moves by a few units in the last place of the
// FP64 solve, which the bound's FP64 term (64 roundings of E0) covers.
__device__ __forceinline__ double fast_quot(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return n * r;
}

__device__ __forceinline__ double qcp_lambda_fast(const QcpCoef& k, double E0) {
  double lam = E0;
  for (int it = 0; it < 50; ++it) {
    const double prev = lam;
    const double l2 = lam * lam;
    const double b = (l2 + k.c2) * lam;
    const double a = b + k.c1;
    lam -= fast_quot(a * lam + k.c0, 2.0 * l2 * lam + b + a);
    if (fabs(lam - prev) < fabs(1e-11 * lam)) break;
  }
  return lam;
}

__device__ __forceinline__ double qcp_rmsd_fast(const double* S, double E0, double invM,
                                                int clamp) {
  const QcpCoef k = qcp_coefficients(S);
  const double e = E0 - qcp_lambda_fast(k, E0);
  if (clamp) return e > 0.0 ? sqrt(2.0 * e * invM) : 0.0;
  return sqrt(fabs(2.0 * e * invM));
}

// Error bound of one pair of the MFMA-reduced forms against the FP64
// computation.  The model (coefficients k*, per-frame norms ||D||, b, l from
// k_prep) bounds the covariance error
//   ||dS||_F <= kdrop l_i l_j + kmix (l_i ||D_j|| + ||D_i|| l_j)
//             + krep b_i b_j + kacc ||D_i|| ||D_j||
//           <= u_i . u_j,  u_f = (sqrt(kdrop) l, sqrt(krep) b,
//                                 sqrt(kacc) ||D|| + kmix / sqrt(kacc) l)
// (the last components' product adds kmix^2 / kacc l_i l_j >= 0), and by
// Cauchy-Schwarz ||dS||_F <= c_i c_j with c_f = |u_f|: one number per frame;
// the u_f of different frames are nearly parallel, so this loses little.
// dS moves the largest QCP eigenvalue by at most sqrt(3) ||dS||_F
// (lambda = max over rotations U of <U, S>, ||U||_F = sqrt(3)), so
// r^2 = 2 (E0 - lambda) / M moves by at most
//   eps = cs_i cs_j + ke0 E0 / M,   cs_f = c_f sqrt(2 sqrt(3) / M)
// (k_prep stores cs_f; ke0 E0 covers the FP64 reassembly and solve).  For the
// stored float r with r^2 > eps the FP64 result lies within
// eps / (r + sqrt(r^2 - eps)) <= eps r / (2 r^2 - eps) = b (as
// sqrt(r^2 - eps) >= r - eps / r), and the float roundings of both outputs add
// at most (r + b) 2^-23, so
//   B = b (1 + 2^-23) + r 2^-23
// is the stated bound, valid for any float r with r^2 > eps; for any r the
// FP64 result also lies in [0, sqrt(r^2 + eps)], so within sqrt(r^2 + eps) of r.
//
// Keep test (keep_limits): the host splits tol once per launch (KeepSplit).
// For r <= h1 the rounding term r 2^-23 is at most h1 2^-23, so with
// tol1 = tol - h1 2^-23 and a = 1 + 2^-23
//   B <= tol  <=  eps a r <= tol1 (2 r^2 - eps)  <=>  r >= R0,
//   R0 = (a eps + sqrt(a^2 eps^2 + 8 tol1^2 eps)) / (4 tol1),
// the larger root of that quadratic.  b decreases above sqrt(eps / 2), so when
// eps a h1 / (2 h1^2 - eps) <= tol - h2 2^-23 (eps <= e16) every r in [h1, h2]
// is kept as well.  keep_limits evaluates R0 and sqrt(eps) in float with every
// rounding (and the hardware square root's 1 ulp) covered by a 2^-18 margin
// upwards, so every float r in [lo, hi] has r^2 > eps and B <= tol: one eps
// serves many pairs with two float compares each.  Report only: lo =
// sqrt(eps)+, hi = infinity.
struct KeepSplit {
  float inv4t;  // 1 / (4 tol1), rounded up (0: report only)
  float c8;     // 8 tol1^2, rounded up
  float h1, h2; // kept ranges end at h1, or h2 when eps <= e16
  float e16;
  float tol;    // tol rounded up (infinity: report only)
};

__device__ __forceinline__ void keep_limits(float e, const KeepSplit& k, float* lo, float* hi) {
  const float t1 = e * (1.0f + 0x1p-22f);  // >= a e after rounding
  const float s = __builtin_amdgcn_sqrtf(fmaf(t1, t1, k.c8 * e));
  const float r0 = (t1 + s) * k.inv4t * (1.0f + 0x1p-18f);
  const float rs = __builtin_amdgcn_sqrtf(e) * (1.0f + 0x1p-18f);
  *lo = fmaxf(r0, rs);  // NaN eps: lo NaN, every compare fails (recomputed)
  *hi = e <= k.e16 ? k.h2 : k.h1;
}

// Upper bounds on B, in float with margins: kept pairs of one eps (e >= eps)
// with RMSDs in [kmin, kmax] (kmin >= sqrt(e)+: b is largest at kmin, the
// rounding term at kmax), and pairs up to r = rmax by the sqrt(r^2 + eps) form.
__device__ __forceinline__ float bound_kept(float e, float kmin, float kmax) {
  const float den = fmaf(2.0f * kmin, kmin, -e) * (1.0f - 0x1p-20f);
  const float b = e * (1.0f + 0x1p-22f) * kmin * __builtin_amdgcn_rcpf(den) * (1.0f + 0x1p-18f);
  return b + kmax * (0x1p-23f * (1.0f + 0x1p-20f));
}

__device__ __forceinline__ float bound_any(float e, float rmax) {
  const float b = __builtin_amdgcn_sqrtf(fmaf(rmax, rmax, e)) * (1.0f + 0x1p-18f);
  return (b + (rmax + b) * 0x1p-23f) * (1.0f + 0x1p-20f);
}

// Largest value over the 64 lanes of a wave (the same in every lane): DPP
// steps within quads, half rows and rows, then row broadcasts collect the
// four rows in lane 63.  For non-negative floats pass their bits (same order).
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false));
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false));
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false));
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false));
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
  v = max(v, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
  return (unsigned int)__builtin_amdgcn_readlane((int)v, 63);
}

// Optimal rotation for covariance S at eigenvalue lam, written as the Kabsch
// U (x' = U x for the centred target).  Returns 0 when every adjugate column
// is degenerate and the identity is returned (QCP fallback), 1 otherwise.
__device__ __forceinline__ int qcp_rotation(const double* S, double lam, double* U) {
  const double Sxx = S[0], Sxy = S[1], Sxz = S[2];
  const double Syx = S[3], Syy = S[4], Syz = S[5];
  const double Szx = S[6], Szy = S[7], Szz = S[8];
  const double xzP = Sxz + Szx, yzP = Syz + Szy, xyP = Sxy + Syx;
  const double yzM = Syz - Szy, xzM = Sxz - Szx, xyM = Sxy - Syx;
  const double xxPyy = Sxx + Syy, xxMyy = Sxx - Syy;
  const double a11 = xxPyy + Szz - lam, a12 = yzM, a13 = -xzM, a14 = xyM;
  const double a21 = yzM, a22 = xxMyy - Szz - lam, a23 = xyP, a24 = xzP;
  const double a31 = a13, a32 = a23, a33 = Syy - Sxx - Szz - lam, a34 = yzP;
  const double a41 = a14, a42 = a24, a43 = a34, a44 = Szz - xxPyy - lam;
  const double m3344 = a33 * a44 - a43 * a34, m3244 = a32 * a44 - a42 * a34;
  const double m3243 = a32 * a43 - a42 * a33, m3143 = a31 * a43 - a41 * a33;
  const double m3144 = a31 * a44 - a41 * a34, m3142 = a31 * a42 - a41 * a32;
  double q1 = a22 * m3344 - a23 * m3244 + a24 * m3243;
  double q2 = -a21 * m3344 + a23 * m3144 - a24 * m3143;
  double q3 = a21 * m3244 - a22 * m3144 + a24 * m3142;
  double q4 = -a21 * m3243 + a22 * m3143 - a23 * m3142;
  double qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
  const double eps = 1e-6;
  if (qq < eps) {
    q1 = a12 * m3344 - a13 * m3244 + a14 * m3243;
    q2 = -a11 * m3344 + a13 * m3144 - a14 * m3143;
    q3 = a11 * m3244 - a12 * m3144 + a14 * m3142;
    q4 = -a11 * m3243 + a12 * m3143 - a13 * m3142;
    qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
    if (qq < eps) {
      const double m1324 = a13 * a24 - a14 * a23, m1224 = a12 * a24 - a14 * a22;
      const double m1223 = a12 * a23 - a13 * a22, m1124 = a11 * a24 - a14 * a21;
      const double m1123 = a11 * a23 - a13 * a21, m1122 = a11 * a22 - a12 * a21;
      q1 = a42 * m1324 - a43 * m1224 + a44 * m1223;
      q2 = -a41 * m1324 + a43 * m1124 - a44 * m1123;
      q3 = a41 * m1224 - a42 * m1124 + a44 * m1122;
      q4 = -a41 * m1223 + a42 * m1123 - a43 * m1122;
      qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
      if (qq < eps) {
        q1 = a32 * m1324 - a33 * m1224 + a34 * m1223;
        q2 = -a31 * m1324 + a33 * m1124 - a34 * m1123;
        q3 = a31 * m1224 - a32 * m1124 + a34 * m1122;
        q4 = -a31 * m1223 + a32 * m1123 - a33 * m1122;
        qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
        if (qq < eps) {
          U[0] = U[4] = U[8] = 1.0;
          U[1] = U[2] = U[3] = U[5] = U[6] = U[7] = 0.0;
          return 0;
        }
      }
    }
  }
  const double nq = sqrt(qq);
  q1 /= nq;
  q2 /= nq;
  q3 /= nq;
  q4 /= nq;
  const double w2 = q1 * q1, x2 = q2 * q2, y2 = q3 * q3, z2 = q4 * q4;
  const double xy = q2 * q3, wz = q1 * q4, zx = q4 * q2, wy = q1 * q3, yz = q3 * q4,
               wx = q1 * q2;
  // QCP's matrix (rows) is the transpose of the Kabsch U; store U directly.
  U[0] = w2 + x2 - y2 - z2;
  U[3] = 2 * (xy + wz);
  U[6] = 2 * (zx - wy);
  U[1] = 2 * (xy - wz);
  U[4] = w2 - x2 + y2 - z2;
  U[7] = 2 * (yz + wx);
  U[2] = 2 * (zx + wy);
  U[5] = 2 * (yz - wx);
  U[8] = w2 - x2 - y2 + z2;
  return 1;
}

// ---- reductions -----------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Butterfly reduce-scatter of 16 doubles per lane over the wave: each step
// sends the half of the remaining values that the partner keeps, so the wave
// moves 8+4+2+1+1+1 values per lane instead of 16 x 6.  On return lane l holds
// the wave total of value (l >> 2); the four lanes 4m..4m+3 hold the same one.
__device__ __forceinline__ double wave_reduce16(const double (&v)[16]) {
  const int lane = threadIdx.x & 63;
  double a8[8], a4[4], a2[2];
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double keep = b5 ? v[i + 8] : v[i];
    const double send = b5 ? v[i] : v[i + 8];
    a8[i] = keep + __shfl_xor(send, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b4 ? a8[i + 4] : a8[i];
    const double send = b4 ? a8[i] : a8[i + 4];
    a4[i] = keep + __shfl_xor(send, 16, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b3 ? a4[i + 2] : a4[i];
    const double send = b3 ? a4[i] : a4[i + 2];
    a2[i] = keep + __shfl_xor(send, 8, 64);
  }
  double r = (b2 ? a2[1] : a2[0]) + __shfl_xor(b2 ? a2[0] : a2[1], 4, 64);
  r += __shfl_xor(r, 2, 64);
  r += __shfl_xor(r, 1, 64);
  return r;
}

// Sums NV doubles per thread over the whole block.  Every thread receives the
// totals.  red must hold NV * (blockDim.x / 64) doubles of shared memory (one
// wave: red is not touched, and there is no barrier).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (nw == 1) return;  // one wave: every lane already holds the totals
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wave * NV + i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w * NV + i];
    v[i] = s;
  }
  __syncthreads();
}

// Packed upper-triangle index without the diagonal (Matrix<T>::calcTriIndex,
// src/Matrix.h:110-122), for i < j.
__device__ __forceinline__ int64_t tri_index(int64_t n, int64_t i, int64_t j) {
  return n * i - ((i + 1) * i) / 2 + j - i - 1;
}

}  // namespace rmsd_dev
