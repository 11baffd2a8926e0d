/*
 * rmsd_oracle.c -- FP64 CPU oracle for cpptraj's best-fit RMSD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C and in FP64, the
 * arithmetic of the reference functions cited beside each routine.  It is the
 * checker for the HIP product path: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links or
 * calls it.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - every routine below is compared with the compiled reference sources
 *     (oracle/_ref, built by oracle/Makefile) on seeded random inputs and on
 *     degenerate inputs (identical, mirrored, collinear, zero-mass frames);
 *   - rms2d / fit routines are compared with the reference's own regression
 *     outputs (test/Test_2DRMS/*.save, test/Test_RMSD/*.save) through the
 *     fixtures under tests/golden/;
 *   - best-fit RMSD is checked against an SVD Kabsch (numpy.linalg.svd) and
 *     against closed forms (exact rigid motion -> 0, rotation recovered);
 *   - the analytic gradient is checked against central finite differences.
 *
 * Conventions shared with the reference (Frame.h / Matrix_3x3.h):
 *   coordinates are AoS doubles x0 y0 z0 x1 ...; 3x3 matrices are row-major;
 *   U is the Kabsch rotation with x' = U x applied to the centred target.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_SMALL 0.00000000000001 /* Constants::SMALL, src/Constants.h:35 */
#define ORC_JACOBI_MAX_SWEEPS 50   /* Matrix_3x3::MAX_ITERATIONS_, src/Matrix_3x3.cpp:97 */

enum { ORC_MODE_FIT = 0, ORC_MODE_NOFIT = 1, ORC_MODE_QUAT = 2 };

/* ------------------------------------------------------------------------ */
/* Centre helpers.  Follows Frame::CenterOnOrigin (src/Frame.cpp:1043-1055),
 * Frame::VCenterOfMass / VGeometricCenter (src/Frame.h:373-437) and
 * Frame::NegTranslate (src/Frame.h:500-506).  Returns the centre (the
 * translation origin -> original centre) and moves the atoms to the origin.   */
void orc_center_on_origin(double *xyz, const double *mass, int n, int use_mass,
                          double center[3])
{
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (use_mass) {
    double total = 0.0;
    for (int a = 0; a < n; ++a) {
      double m = mass[a];
      sx += xyz[3 * a] * m;
      sy += xyz[3 * a + 1] * m;
      sz += xyz[3 * a + 2] * m;
      total += m;
    }
    if (total == 0.0) {
      center[0] = center[1] = center[2] = 0.0;
    } else {
      center[0] = sx / total;
      center[1] = sy / total;
      center[2] = sz / total;
    }
  } else {
    for (int a = 0; a < n; ++a) {
      sx += xyz[3 * a];
      sy += xyz[3 * a + 1];
      sz += xyz[3 * a + 2];
    }
    double dn = (double)n;
    center[0] = sx / dn;
    center[1] = sy / dn;
    center[2] = sz / dn;
  }
  for (int a = 0; a < n; ++a) {
    xyz[3 * a] -= center[0];
    xyz[3 * a + 1] -= center[1];
    xyz[3 * a + 2] -= center[2];
  }
}

/* ------------------------------------------------------------------------ */
/* Cyclic Jacobi for a symmetric 3x3 matrix.  Restates Matrix_3x3::Diagonalize
 * (src/Matrix_3x3.cpp:110-211): at most 50 sweeps; threshold 0.2*sm/9 in the
 * first three sweeps, 0 afterwards; after sweep 4 an off-diagonal element that
 * is negligible against both diagonal entries is zeroed; the diagonal is
 * refreshed from the accumulated updates at the end of each sweep.
 * On return V holds the eigenvectors in its columns and d the eigenvalues.
 * Returns 0 on convergence, 1 when the sweep cap is hit.                      */
static void jrot(double *m, int r1, int c1, int r2, int c2, double s, double tau)
{
  double g = m[3 * r1 + c1];
  double h = m[3 * r2 + c2];
  m[3 * r1 + c1] = g - s * (h + g * tau);
  m[3 * r2 + c2] = h + s * (g - h * tau);
}

static int jacobi3(const double in[9], double V[9], double d[3])
{
  double a[9];
  memcpy(a, in, sizeof(a));
  for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
  double b[3], z[3] = {0.0, 0.0, 0.0};
  b[0] = d[0] = a[0];
  b[1] = d[1] = a[4];
  b[2] = d[2] = a[8];
  for (int sweep = 0; sweep < ORC_JACOBI_MAX_SWEEPS; ++sweep) {
    double off = fabs(a[1]) + fabs(a[2]) + fabs(a[5]);
    if (off == 0) return 0;
    double thr = (sweep < 3) ? 0.2 * off / 9 : 0;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        int pq = 3 * p + q;
        double big = 100.0 * fabs(a[pq]);
        if (sweep > 3 && fabs(d[p]) + big == fabs(d[p]) &&
            fabs(d[q]) + big == fabs(d[q])) {
          a[pq] = 0;
        } else if (fabs(a[pq]) > thr) {
          double diff = d[q] - d[p];
          double t;
          if (fabs(diff) + big == fabs(diff)) {
            t = a[pq] / diff;
          } else {
            double theta = 0.5 * diff / a[pq];
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          double c = 1.0 / sqrt(1 + t * t);
          double s = t * c;
          double tau = s / (1.0 + c);
          double h = t * a[pq];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[pq] = 0;
          for (int j = 0; j <= p - 1; ++j) jrot(a, j, p, j, q, s, tau);
          for (int j = p + 1; j <= q - 1; ++j) jrot(a, p, j, j, q, s, tau);
          for (int j = q + 1; j < 3; ++j) jrot(a, p, j, q, j, s, tau);
          for (int j = 0; j < 3; ++j) jrot(V, j, p, j, q, s, tau);
        }
      }
    }
    for (int k = 0; k < 3; ++k) {
      b[k] += z[k];
      d[k] = b[k];
      z[k] = 0;
    }
  }
  return 1;
}

/* Descending eigen-sort with eigenvectors moved into rows.  Restates
 * Matrix_3x3::Diagonalize_Sort (src/Matrix_3x3.cpp:219-268), including its
 * strict '>' tie handling.                                                    */
static int jacobi3_sorted(const double in[9], double E[9], double ev[3])
{
  double V[9], d[3];
  if (jacobi3(in, V, d)) return 1;
  int o0, o1, o2;
  if (d[0] > d[1] && d[0] > d[2]) {
    if (d[1] > d[2]) { o0 = 0; o1 = 1; o2 = 2; } else { o0 = 0; o1 = 2; o2 = 1; }
  } else if (d[1] > d[0] && d[1] > d[2]) {
    if (d[0] > d[2]) { o0 = 1; o1 = 0; o2 = 2; } else { o0 = 1; o1 = 2; o2 = 0; }
  } else if (d[0] > d[1]) {
    o0 = 2; o1 = 0; o2 = 1;
  } else {
    o0 = 2; o1 = 1; o2 = 0;
  }
  int ord[3] = {o0, o1, o2};
  for (int r = 0; r < 3; ++r) {
    E[3 * r + 0] = V[ord[r]];
    E[3 * r + 1] = V[ord[r] + 3];
    E[3 * r + 2] = V[ord[r] + 6];
    ev[r] = d[ord[r]];
  }
  return 0;
}

int orc_jacobi3_sorted(const double in[9], double E[9], double ev[3])
{
  return jacobi3_sorted(in, E, ev);
}

static void unit3(double *v)
{
  double inv = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= inv;
  v[1] *= inv;
  v[2] *= inv;
}

/* Kabsch solve from a covariance R (R_ij = sum m t_i r_j) and E0.
 * Steps of Frame::RMSD_CenteredRef after the covariance loop
 * (src/Frame.cpp:1213-1268): RRt = R R^T (Matrix_3x3::TransposeMult,
 * src/Matrix_3x3.cpp:421-433); sorted Jacobi; a3 = a1 x a2; b_k = R^T-projected
 * a_k, normalised; sig3 = sign((b1 x b2) . b3); b3 := b1 x b2;
 * U = sum_k a_k b_k^T laid out as in the reference; RMSD from the square roots
 * of the eigenvalues, clamped at 0.
 * Returns the RMSD, or 0 when Jacobi does not converge (reference behaviour). */
double orc_kabsch_from_cov(const double R[9], double E0, double total_mass,
                           double U[9])
{
  double RR[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      RR[3 * i + j] = R[3 * i] * R[3 * j] + R[3 * i + 1] * R[3 * j + 1] +
                      R[3 * i + 2] * R[3 * j + 2];
  double E[9], ev[3];
  if (jacobi3_sorted(RR, E, ev)) return 0;
  E[6] = (E[1] * E[5]) - (E[2] * E[4]);
  E[7] = (E[2] * E[3]) - (E[0] * E[5]);
  E[8] = (E[0] * E[4]) - (E[1] * E[3]);
  double b[9];
  for (int k = 0; k < 3; ++k) {
    const double *a = E + 3 * k;
    double *bk = b + 3 * k;
    bk[0] = a[0] * R[0] + a[1] * R[3] + a[2] * R[6];
    bk[1] = a[0] * R[1] + a[1] * R[4] + a[2] * R[7];
    bk[2] = a[0] * R[2] + a[1] * R[5] + a[2] * R[8];
    unit3(bk);
  }
  double c0 = (b[1] * b[5]) - (b[2] * b[4]);
  double c1 = (b[2] * b[3]) - (b[0] * b[5]);
  double c2 = (b[0] * b[4]) - (b[1] * b[3]);
  double sig3 = ((c0 * b[6] + c1 * b[7] + c2 * b[8]) < 0.0) ? -1.0 : 1.0;
  b[6] = c0;
  b[7] = c1;
  b[8] = c2;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      U[3 * r + c] = (E[c] * b[r]) + (E[3 + c] * b[3 + r]) + (E[6 + c] * b[6 + r]);
  double e = E0 - sqrt(fabs(ev[0])) - sqrt(fabs(ev[1])) - (sig3 * sqrt(fabs(ev[2])));
  if (e < 0) return 0.0;
  return sqrt((2.0 * e) / total_mass);
}

/* Centre the target and accumulate the covariance and E0.
 * Follows the first half of Frame::RMSD_CenteredRef (src/Frame.cpp:1141-1209):
 * target centroid with target masses (or atom count), Sigma m < SMALL -> error,
 * target translated to the origin, then R_ij = sum m t_i r_j and
 * E0 = 0.5 sum m (|t|^2 + |r|^2) with the target masses.
 * Returns the total mass, or -1 on the divide-by-zero guard.                 */
static double centre_and_cov(double *tgt, const double *ref, const double *mass,
                             int n, int use_mass, double R[9], double *E0,
                             double trans[3])
{
  double total;
  double tx = 0.0, ty = 0.0, tz = 0.0;
  if (use_mass) {
    total = 0.0;
    for (int a = 0; a < n; ++a) {
      total += mass[a];
      tx += tgt[3 * a] * mass[a];
      ty += tgt[3 * a + 1] * mass[a];
      tz += tgt[3 * a + 2] * mass[a];
    }
  } else {
    total = (double)n;
    for (int a = 0; a < n; ++a) {
      tx += tgt[3 * a];
      ty += tgt[3 * a + 1];
      tz += tgt[3 * a + 2];
    }
  }
  if (total < ORC_SMALL) {
    trans[0] = tx; trans[1] = ty; trans[2] = tz;
    return -1.0;
  }
  trans[0] = -(tx / total);
  trans[1] = -(ty / total);
  trans[2] = -(tz / total);
  for (int a = 0; a < n; ++a) {
    tgt[3 * a] += trans[0];
    tgt[3 * a + 1] += trans[1];
    tgt[3 * a + 2] += trans[2];
  }
  double ss = 0.0;
  for (int k = 0; k < 9; ++k) R[k] = 0.0;
  double w = 1.0;
  for (int a = 0; a < n; ++a) {
    double xt = tgt[3 * a], yt = tgt[3 * a + 1], zt = tgt[3 * a + 2];
    double xr = ref[3 * a], yr = ref[3 * a + 1], zr = ref[3 * a + 2];
    if (use_mass) w = mass[a];
    ss += w * ((xt * xt) + (yt * yt) + (zt * zt) + (xr * xr) + (yr * yr) + (zr * zr));
    R[0] += w * xt * xr;
    R[1] += w * xt * yr;
    R[2] += w * xt * zr;
    R[3] += w * yt * xr;
    R[4] += w * yt * yr;
    R[5] += w * yt * zr;
    R[6] += w * zt * xr;
    R[7] += w * zt * yr;
    R[8] += w * zt * zr;
  }
  *E0 = ss * 0.5;
  return total;
}

/* Frame::RMSD_CenteredRef(Ref, U, Trans, useMass), src/Frame.cpp:1137-1272.
 * tgt is moved to the origin (as in the reference).  trans receives -centroid.*/
double orc_rmsd_centered_ref(double *tgt, const double *ref, const double *mass,
                             int n, int use_mass, double U[9], double trans[3])
{
  double R[9], E0;
  double total = centre_and_cov(tgt, ref, mass, n, use_mass, R, &E0, trans);
  if (total < 0.0) {
    trans[0] = trans[1] = trans[2] = 0.0;
    return -1.0;
  }
  return orc_kabsch_from_cov(R, E0, total, U);
}

/* ------------------------------------------------------------------------ */
/* QCP (Theobald 2005; Liu, Agrafiotis & Theobald 2010).  Restates
 * Cpptraj::QCPRot::FastCalcRMSDAndRotation (src/qcprot.cpp:167-343): quartic
 * coefficients of the key 4x4 matrix from A, Newton from lambda = E0 (at most 50
 * steps, |step| < 1e-11 lambda), rmsd = sqrt(|2(E0-lambda)/len|), then the
 * eigenvector from adjugate columns with the reference's fallbacks (norm^2 <
 * 1e-6) down to the identity.  rot (if not NULL) receives the rotation in the
 * reference's QCP layout, which is the transpose of the Kabsch U.
 * Returns -1 when rot is NULL, 0 for the identity fallback, 1 otherwise.
 *
 * qcprot (version 1.5) is distributed under this notice, which applies to the
 * restated algorithm:
 *   Copyright (c) 2009-2016 Pu Liu and Douglas L. Theobald.  All rights reserved.
 *   Redistribution and use in source and binary forms, with or without
 *   modification, are permitted provided that the following conditions are met:
 *   * Redistributions of source code must retain the above copyright notice,
 *     this list of conditions and the following disclaimer.
 *   * Redistributions in binary form must reproduce the above copyright notice,
 *     this list of conditions and the following disclaimer in the documentation
 *     and/or other materials provided with the distribution.
 *   * Neither the name of the <ORGANIZATION> nor the names of its contributors
 *     may be used to endorse or promote products derived from this software
 *     without specific prior written permission.
 *   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
 *   AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
 *   IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE
 *   ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER OR CONTRIBUTORS BE
 *   LIABLE FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR
 *   CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF
 *   SUBSTITUTE GOODS OR SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS
 *   INTERRUPTION) HOWEVER CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN
 *   CONTRACT, STRICT LIABILITY, OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE)
 *   ARISING IN ANY WAY OUT OF THE USE OF THIS SOFTWARE, EVEN IF ADVISED OF THE
 *   POSSIBILITY OF SUCH DAMAGE.                                                */
int orc_qcp(double *rot, const double A[9], double *rmsd, double E0, double len)
{
  const double Sxx = A[0], Sxy = A[1], Sxz = A[2];
  const double Syx = A[3], Syy = A[4], Syz = A[5];
  const double Szx = A[6], Szy = A[7], Szz = A[8];
  const double Sxx2 = Sxx * Sxx, Syy2 = Syy * Syy, Szz2 = Szz * Szz;
  const double Sxy2 = Sxy * Sxy, Syz2 = Syz * Syz, Sxz2 = Sxz * Sxz;
  const double Syx2 = Syx * Syx, Szy2 = Szy * Szy, Szx2 = Szx * Szx;

  const double pYZ = 2.0 * (Syz * Szy - Syy * Szz);
  const double qDiag = Syy2 + Szz2 - Sxx2 + Syz2 + Szy2;
  double c2 = -2.0 * (Sxx2 + Syy2 + Szz2 + Sxy2 + Syx2 + Sxz2 + Szx2 + Syz2 + Szy2);
  double c1 = 8.0 * (Sxx * Syz * Szy + Syy * Szx * Sxz + Szz * Sxy * Syx -
                     Sxx * Syy * Szz - Syz * Szx * Sxy - Szy * Syx * Sxz);
  const double xzP = Sxz + Szx, yzP = Syz + Szy, xyP = Sxy + Syx;
  const double yzM = Syz - Szy, xzM = Sxz - Szx, xyM = Sxy - Syx;
  const double xxPyy = Sxx + Syy, xxMyy = Sxx - Syy;
  const double offd = Sxy2 + Sxz2 - Syx2 - Szx2;
  double c0 = offd * offd
      + (qDiag + pYZ) * (qDiag - pYZ)
      + (-(xzP) * (yzM) + (xyM) * (xxMyy - Szz)) * (-(xzM) * (yzP) + (xyM) * (xxMyy + Szz))
      + (-(xzP) * (yzP) - (xyP) * (xxPyy - Szz)) * (-(xzM) * (yzM) - (xyP) * (xxPyy + Szz))
      + (+(xyP) * (yzP) + (xzP) * (xxMyy + Szz)) * (-(xyM) * (yzM) + (xzP) * (xxPyy + Szz))
      + (+(xyP) * (yzM) + (xzM) * (xxMyy - Szz)) * (-(xyM) * (yzP) + (xzM) * (xxPyy - Szz));

  double lam = E0;
  for (int it = 0; it < 50; ++it) {
    double prev = lam;
    double l2 = lam * lam;
    double b = (l2 + c2) * lam;
    double a = b + c1;
    double step = ((a * lam + c0) / (2.0 * l2 * lam + b + a));
    lam -= step;
    if (fabs(lam - prev) < fabs(1e-11 * lam)) break;
  }
  *rmsd = sqrt(fabs(2.0 * (E0 - lam) / len));
  if (rot == 0) return -1;

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
  const double evecprec = 1e-6;
  if (qq < evecprec) {
    q1 = a12 * m3344 - a13 * m3244 + a14 * m3243;
    q2 = -a11 * m3344 + a13 * m3144 - a14 * m3143;
    q3 = a11 * m3244 - a12 * m3144 + a14 * m3142;
    q4 = -a11 * m3243 + a12 * m3143 - a13 * m3142;
    qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
    if (qq < evecprec) {
      const double m1324 = a13 * a24 - a14 * a23, m1224 = a12 * a24 - a14 * a22;
      const double m1223 = a12 * a23 - a13 * a22, m1124 = a11 * a24 - a14 * a21;
      const double m1123 = a11 * a23 - a13 * a21, m1122 = a11 * a22 - a12 * a21;
      q1 = a42 * m1324 - a43 * m1224 + a44 * m1223;
      q2 = -a41 * m1324 + a43 * m1124 - a44 * m1123;
      q3 = a41 * m1224 - a42 * m1124 + a44 * m1122;
      q4 = -a41 * m1223 + a42 * m1123 - a43 * m1122;
      qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
      if (qq < evecprec) {
        q1 = a32 * m1324 - a33 * m1224 + a34 * m1223;
        q2 = -a31 * m1324 + a33 * m1124 - a34 * m1123;
        q3 = a31 * m1224 - a32 * m1124 + a34 * m1122;
        q4 = -a31 * m1223 + a32 * m1123 - a33 * m1122;
        qq = q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4;
        if (qq < evecprec) {
          rot[0] = rot[4] = rot[8] = 1.0;
          rot[1] = rot[2] = rot[3] = rot[5] = rot[6] = rot[7] = 0.0;
          return 0;
        }
      }
    }
  }
  double nq = sqrt(qq);
  q1 /= nq;
  q2 /= nq;
  q3 /= nq;
  q4 /= nq;
  double w2 = q1 * q1, x2 = q2 * q2, y2 = q3 * q3, z2 = q4 * q4;
  double xy = q2 * q3, wz = q1 * q4, zx = q4 * q2, wy = q1 * q3, yz = q3 * q4, wx = q1 * q2;
  rot[0] = w2 + x2 - y2 - z2;
  rot[1] = 2 * (xy + wz);
  rot[2] = 2 * (zx - wy);
  rot[3] = 2 * (xy - wz);
  rot[4] = w2 - x2 + y2 - z2;
  rot[5] = 2 * (yz + wx);
  rot[6] = 2 * (zx + wy);
  rot[7] = 2 * (yz - wx);
  rot[8] = w2 - x2 - y2 + z2;
  return 1;
}

/* do_quaternion_rmsd (src/QuaternionRMSD.cpp:25-120): centre Tgt (target
 * masses; unweighted total = atom count), covariance A and E0 as above, QCP.
 * rot may be NULL.                                                           */
double orc_qrmsd_centered_ref(const double *ref, double *tgt, const double *mass,
                              int n, int use_mass, double *rot, double trans[3])
{
  double R[9], E0;
  double total = centre_and_cov(tgt, ref, mass, n, use_mass, R, &E0, trans);
  if (total < 0.0) {
    trans[0] = trans[1] = trans[2] = 0.0;
    return -1.0;
  }
  double r;
  orc_qcp(rot, R, &r, E0, total);
  return r;
}

/* Frame::RMSD_NoFit (src/Frame.cpp:1279-1307).                               */
double orc_rmsd_nofit(const double *tgt, const double *ref, const double *mass,
                      int n, int use_mass)
{
  double acc = 0.0, total = 0.0, w = 1.0;
  for (int a = 0; a < n; ++a) {
    double dx = ref[3 * a] - tgt[3 * a];
    double dy = ref[3 * a + 1] - tgt[3 * a + 1];
    double dz = ref[3 * a + 2] - tgt[3 * a + 2];
    if (use_mass) w = mass[a];
    total += w;
    acc += (w * (dx * dx + dy * dy + dz * dz));
  }
  if (total < ORC_SMALL) return -1.0;
  if (acc < 0) return 0.0;
  return sqrt(acc / total);
}

/* Masked float -> double gather (CompactFrameArray::GetToMaskDblPtr,
 * src/CompactFrameArray.cpp:244-262).  sel == NULL selects atoms 0..nsel-1.   */
void orc_gather(const float *frame, const int *sel, int nsel, double *out)
{
  for (int j = 0; j < nsel; ++j) {
    int a = sel ? sel[j] : j;
    out[3 * j] = (double)frame[3 * a];
    out[3 * j + 1] = (double)frame[3 * a + 1];
    out[3 * j + 2] = (double)frame[3 * a + 2];
  }
}

/* Packed matrix indices (Matrix<T>::calcFullIndex / calcTriIndex,
 * src/Matrix.h:94-122).                                                       */
int64_t orc_full_index(int64_t nx, int64_t x, int64_t y) { return y * nx + x; }

int64_t orc_tri_index(int64_t nx, int64_t x, int64_t y)
{
  int64_t i, j;
  if (y > x) { i = x; j = y; }
  else if (x > y) { i = y; j = x; }
  else return -1;
  int64_t i1 = i + 1;
  return ((nx * i) - ((i1 * i) / 2)) + j - i1;
}

/* ------------------------------------------------------------------------ */
/* Analysis_Rms2d::Calculate_2D (src/Analysis_Rms2d.cpp:196-295).
 * Frames are cpptraj CRD frames: AoS floats, frame f starts at
 * frames + f*stride.  When ref_frames == NULL the reference set is the target
 * set.  full != 0 selects the full Fref x Ftgt layout (index ntgt*nref_total +
 * nref); otherwise the packed upper triangle.  tmass / rmass are the masses of
 * the selected target / reference atoms.  Loops over reference rows with a
 * dynamic schedule, gathers the target per pair, stores (float)RMSD.          */
int orc_rms2d(const float *tgt_frames, int64_t tgt_stride, int ntgt,
              const float *ref_frames, int64_t ref_stride, int nref,
              const int *tsel, const int *rsel, int nsel,
              const double *tmass, const double *rmass,
              int mode, int use_mass, int full, float *out, int nthreads)
{
  if (ref_frames == 0) {
    ref_frames = tgt_frames;
    ref_stride = tgt_stride;
    nref = ntgt;
  }
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *R = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    double *T = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    if (!R || !T) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
    }
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int r = 0; r < nref; ++r) {
      if (!R || !T) continue;
      orc_gather(ref_frames + (int64_t)r * ref_stride, rsel, nsel, R);
      if (mode == ORC_MODE_FIT || mode == ORC_MODE_QUAT) {
        double c[3];
        orc_center_on_origin(R, rmass, nsel, use_mass, c);
      }
      int t0 = full ? 0 : r + 1;
      for (int t = t0; t < ntgt; ++t) {
        orc_gather(tgt_frames + (int64_t)t * tgt_stride, tsel, nsel, T);
        double v, U[9], tr[3];
        if (mode == ORC_MODE_FIT)
          v = orc_rmsd_centered_ref(T, R, tmass, nsel, use_mass, U, tr);
        else if (mode == ORC_MODE_NOFIT)
          v = orc_rmsd_nofit(T, R, tmass, nsel, use_mass);
        else
          v = orc_qrmsd_centered_ref(R, T, tmass, nsel, use_mass, 0, tr);
        int64_t idx = full ? orc_full_index(nref, r, t) : orc_tri_index(ntgt, r, t);
        out[idx] = (float)v;
      }
    }
    free(R);
    free(T);
  }
  return err;
}

/* Cluster pairwise cache fill: MetricArray::calcFrameDistances
 * (src/Cluster/MetricArray.cpp:766-801) with one RMS metric
 * (Metric_RMS::FrameDist, src/Cluster/Metric_RMS.cpp:44-63) whose distance is
 * Frame::RMSD (src/Frame.cpp:1084-1109: reference frame re-centred, then
 * RMSD_CenteredRef) or RMSD_NoFit, scaled by the metric weight
 * (MetricArray::Dist_Manhattan, :539-547).  frame_idx lists the frames to
 * cache (sieved set); the output is the packed triangle over that list.      */
int orc_cluster_cache(const float *frames, int64_t stride, const int *frame_idx,
                      int ncache, const int *sel, int nsel, const double *mass,
                      int nofit, int use_mass, double weight, float *out,
                      int nthreads)
{
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *A = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    double *B = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    if (!A || !B) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
    }
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int f1 = 0; f1 < ncache - 1; ++f1) {
      if (!A || !B) continue;
      for (int f2 = f1 + 1; f2 < ncache; ++f2) {
        int g1 = frame_idx ? frame_idx[f1] : f1;
        int g2 = frame_idx ? frame_idx[f2] : f2;
        orc_gather(frames + (int64_t)g1 * stride, sel, nsel, A);
        orc_gather(frames + (int64_t)g2 * stride, sel, nsel, B);
        double v;
        if (nofit) {
          v = orc_rmsd_nofit(A, B, mass, nsel, use_mass);
        } else {
          double c[3], U[9], tr[3];
          orc_center_on_origin(B, mass, nsel, use_mass, c);
          v = orc_rmsd_centered_ref(A, B, mass, nsel, use_mass, U, tr);
        }
        out[orc_tri_index(ncache, f1, f2)] = (float)(weight * v);
      }
    }
    free(A);
    free(B);
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* Action_Rmsd::DoAction over a set of frames against one fixed reference
 * (src/Action_Rmsd.cpp:361-417, reference modes 'first' / 'ref' of
 * ReferenceAction, src/ReferenceAction.h:73-84).  ref_sel is the selected
 * reference (nsel x 3 doubles, already centred when fitting) and ref_trans its
 * original centre.  modify: 0 = nomod, 1 = rotate+translate (Trans_Rot_Trans,
 * src/Frame.h:572-581), 2 = translate only.  U and tvec (Trans, refTrans as
 * 6 doubles per frame) may be NULL.  frames_out (may be NULL) receives the
 * modified frames (all natom atoms, float, same stride).                     */
int orc_fit_frames(const float *frames, int64_t stride, int nframes, int natom,
                   const int *sel, int nsel, const double *mass,
                   const double *ref_sel, const double ref_trans[3],
                   int fit, int use_mass, int modify, double *rmsd, double *U,
                   double *tvec, float *frames_out, int nthreads)
{
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *T = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    if (!T) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
    }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int f = 0; f < nframes; ++f) {
      if (!T) continue;
      const float *fr = frames + (int64_t)f * stride;
      orc_gather(fr, sel, nsel, T);
      double Uf[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tr[3] = {0, 0, 0};
      double v;
      if (!fit) {
        v = orc_rmsd_nofit(T, ref_sel, mass, nsel, use_mass);
      } else {
        v = orc_rmsd_centered_ref(T, ref_sel, mass, nsel, use_mass, Uf, tr);
      }
      rmsd[f] = v;
      if (fit && U) memcpy(U + 9 * (int64_t)f, Uf, sizeof(Uf));
      if (fit && tvec) {
        double *tv = tvec + 6 * (int64_t)f;
        tv[0] = tr[0]; tv[1] = tr[1]; tv[2] = tr[2];
        tv[3] = ref_trans[0]; tv[4] = ref_trans[1]; tv[5] = ref_trans[2];
      }
      if (frames_out) {
        float *fo = frames_out + (int64_t)f * stride;
        for (int a = 0; a < natom; ++a) {
          double x = fr[3 * a], y = fr[3 * a + 1], z = fr[3 * a + 2];
          if (fit && modify == 1) {
            /* Frame::Trans_Rot_Trans: translate, rotate with U, translate. */
            x += tr[0]; y += tr[1]; z += tr[2];
            double rx = Uf[0] * x + Uf[1] * y + Uf[2] * z;
            double ry = Uf[3] * x + Uf[4] * y + Uf[5] * z;
            double rz = Uf[6] * x + Uf[7] * y + Uf[8] * z;
            x = rx + ref_trans[0];
            y = ry + ref_trans[1];
            z = rz + ref_trans[2];
          } else if (fit && modify == 2) {
            /* TRANS_ONLY: translate by Trans + refTrans. */
            x += tr[0] + ref_trans[0];
            y += tr[1] + ref_trans[1];
            z += tr[2] + ref_trans[2];
          }
          fo[3 * a] = (float)x;
          fo[3 * a + 1] = (float)y;
          fo[3 * a + 2] = (float)z;
        }
      }
    }
    free(T);
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* Analytic gradient of the best-fit RMSD with respect to the target atoms.
 * The reference has no gradient; this is the envelope-theorem result
 *   dRMSD/dx_k = m_k (x~_k - U^T y~_k) / (M * RMSD)
 * for target atoms x (centred: x~), centred reference y~ and the Kabsch U
 * that maps the centred target onto the reference.  Written from the
 * mathematics; pinned by central finite differences in the tests.
 * grad (nsel x 3) receives the gradient; returns the RMSD (0 -> grad 0).      */
double orc_rmsd_grad_target(const double *tgt_in, const double *ref_centered,
                            const double *mass, int n, int use_mass,
                            double *grad)
{
  double *T = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  if (!T) return -1.0;
  memcpy(T, tgt_in, sizeof(double) * 3 * (size_t)n);
  double U[9], tr[3];
  double r = orc_rmsd_centered_ref(T, ref_centered, mass, n, use_mass, U, tr);
  double M = 0.0;
  for (int a = 0; a < n; ++a) M += use_mass ? mass[a] : 1.0;
  for (int a = 0; a < n; ++a) {
    if (!(r > 0.0)) {
      grad[3 * a] = grad[3 * a + 1] = grad[3 * a + 2] = 0.0;
      continue;
    }
    double m = use_mass ? mass[a] : 1.0;
    const double *y = ref_centered + 3 * a;
    double ry0 = U[0] * y[0] + U[3] * y[1] + U[6] * y[2];
    double ry1 = U[1] * y[0] + U[4] * y[1] + U[7] * y[2];
    double ry2 = U[2] * y[0] + U[5] * y[1] + U[8] * y[2];
    double s = m / (M * r);
    grad[3 * a] = s * (T[3 * a] - ry0);
    grad[3 * a + 1] = s * (T[3 * a + 1] - ry1);
    grad[3 * a + 2] = s * (T[3 * a + 2] - ry2);
  }
  free(T);
  return r;
}

/* Selected rows of the triangle of orc_rms2d (same frames and mask as target
 * and reference): for each listed reference row r, the pairs (r, t > r) are
 * written consecutively into out (row after row, in list order).  Used to
 * check sampled rows of matrices too large for the whole oracle.            */
int orc_rms2d_rows(const float *frames, int64_t stride, int nframes, const int *sel,
                   int nsel, const double *mass, int mode, int use_mass,
                   const int *rows, int nrows, float *out, int nthreads)
{
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrows + 1));
  if (!off) return 1;
  off[0] = 0;
  for (int q = 0; q < nrows; ++q) off[q + 1] = off[q] + (nframes - 1 - rows[q]);
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *R = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    double *T = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    if (!R || !T) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
    }
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int q = 0; q < nrows; ++q) {
      if (!R || !T) continue;
      const int r = rows[q];
      orc_gather(frames + (int64_t)r * stride, sel, nsel, R);
      if (mode == ORC_MODE_FIT || mode == ORC_MODE_QUAT) {
        double c[3];
        orc_center_on_origin(R, mass, nsel, use_mass, c);
      }
      for (int t = r + 1; t < nframes; ++t) {
        orc_gather(frames + (int64_t)t * stride, sel, nsel, T);
        double v, U[9], tr[3];
        if (mode == ORC_MODE_FIT)
          v = orc_rmsd_centered_ref(T, R, mass, nsel, use_mass, U, tr);
        else if (mode == ORC_MODE_NOFIT)
          v = orc_rmsd_nofit(T, R, mass, nsel, use_mass);
        else
          v = orc_qrmsd_centered_ref(R, T, mass, nsel, use_mass, 0, tr);
        out[off[q] + (t - r - 1)] = (float)v;
      }
    }
    free(R);
    free(T);
  }
  free(off);
  return err;
}

/* ------------------------------------------------------------------------ */
/* Per-residue no-fit RMSD of Action_Rmsd (src/Action_Rmsd.cpp:398-413, masks
 * from perResSetup :197-315).  For frame f and residue r the target atoms
 * tgt_atoms[res_off[r] .. res_off[r+1]) of the frame, as the fit's coordinate
 * update leaves it in double precision (modify 1: Frame::Trans_Rot_Trans,
 * src/Frame.h:572-581, with U and tvec of the fit; modify 2: translate by
 * Trans + refTrans; 0 or U == NULL: the frame as stored), are compared with
 * the reference atoms ref_atoms[...] of the uncentred reference frame
 * (ReferenceAction::CurrentReference, src/ReferenceAction.h:36): both sides
 * optionally centred (perrescenter: Frame::CenterOnOrigin with the target's
 * and the reference's own masses), then Frame::RMSD_NoFit with the target
 * masses.  out[f * nres + r].                                                 */
int orc_perres(const float *frames, int64_t stride, int nframes, const float *ref,
               int nres, const int *res_off, const int *tgt_atoms, const int *ref_atoms,
               const double *tmass, const double *rmass, int use_mass, int center,
               const double *U, const double *tvec, int modify, double *out)
{
  int maxn = 0;
  for (int r = 0; r < nres; ++r)
    if (res_off[r + 1] - res_off[r] > maxn) maxn = res_off[r + 1] - res_off[r];
  double *T = (double *)malloc(sizeof(double) * 3 * (size_t)(maxn > 0 ? maxn : 1));
  double *R = (double *)malloc(sizeof(double) * 3 * (size_t)(maxn > 0 ? maxn : 1));
  if (!T || !R) {
    free(T);
    free(R);
    return 1;
  }
  for (int f = 0; f < nframes; ++f) {
    const float *fr = frames + (int64_t)f * stride;
    const double *Uf = U ? U + 9 * (int64_t)f : NULL;
    const double *tv = tvec ? tvec + 6 * (int64_t)f : NULL;
    for (int r = 0; r < nres; ++r) {
      const int b = res_off[r], n = res_off[r + 1] - b;
      for (int k = 0; k < n; ++k) {
        const int a = tgt_atoms[b + k], q = ref_atoms[b + k];
        double x = fr[3 * a], y = fr[3 * a + 1], z = fr[3 * a + 2];
        if (Uf && tv && modify == 1) {
          x += tv[0]; y += tv[1]; z += tv[2];
          double rx = Uf[0] * x + Uf[1] * y + Uf[2] * z;
          double ry = Uf[3] * x + Uf[4] * y + Uf[5] * z;
          double rz = Uf[6] * x + Uf[7] * y + Uf[8] * z;
          x = rx + tv[3];
          y = ry + tv[4];
          z = rz + tv[5];
        } else if (tv && modify == 2) {
          x += tv[0] + tv[3];
          y += tv[1] + tv[4];
          z += tv[2] + tv[5];
        }
        T[3 * k] = x; T[3 * k + 1] = y; T[3 * k + 2] = z;
        R[3 * k] = ref[3 * q]; R[3 * k + 1] = ref[3 * q + 1]; R[3 * k + 2] = ref[3 * q + 2];
      }
      if (center) {
        double c[3];
        orc_center_on_origin(T, tmass ? tmass + b : NULL, n, use_mass, c);
        orc_center_on_origin(R, rmass ? rmass + b : NULL, n, use_mass, c);
      }
      out[(int64_t)f * nres + r] =
          n > 0 ? orc_rmsd_nofit(T, R, tmass ? tmass + b : NULL, n, use_mass) : -1.0;
    }
  }
  free(T);
  free(R);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Window sizes of Analysis_RmsAvgCorr::Analyze (src/Analysis_RmsAvgCorr.cpp:
 * 154-209): 1, then 1 + offset, 1 + 2 offset, ... below WindowMax (nframes,
 * or stop + 1 capped at nframes).  Writes up to cap sizes into w and returns
 * their number, or -1 when there are not enough frames.                      */
int orc_avgcorr_windows(int nframes, int maxwindow, int offset, int *w, int cap)
{
  if (offset < 1) offset = 1;
  int wmax = nframes;
  if (maxwindow != -1) {
    wmax = maxwindow + 1;
    if (wmax > nframes) wmax = nframes;
  }
  const int start = 1 + offset;
  int total = (wmax - start) / offset;
  if (total % offset > 0) ++total;
  if (total < 0) return -1;
  int n = 0;
  if (n < cap && w) w[n] = 1;
  ++n;
  for (int ws = start; ws < wmax; ws += offset) {
    if (n < cap && w) w[n] = ws;
    ++n;
  }
  return n;
}

/* Analysis_RmsAvgCorr::Analyze (src/Analysis_RmsAvgCorr.cpp:119-276): for
 * each window size, the running sum of the selected atoms (Frame operator+=
 * / operator-=, FP64), divided by the window (Frame::Divide), fitted to the
 * reference (Frame::RMSD_CenteredRef): the centred reference ref_sel, or
 * with ref_sel == NULL ('first') the window's first averaged frame, centred.
 * ct / sd receive the mean and standard deviation over the averaged frames,
 * one per window size (orc_avgcorr_windows).  Returns the number of window
 * sizes or -1.                                                                */
int orc_avgcorr(const float *frames, int64_t stride, int nframes, const int *sel, int nsel,
                const double *mass, const double *ref_sel, int use_mass, int maxwindow,
                int offset, double *ct, double *sd)
{
  const int nw = orc_avgcorr_windows(nframes, maxwindow, offset, NULL, 0);
  if (nw < 1) return -1;
  int *ws = (int *)malloc(sizeof(int) * (size_t)nw);
  double *sum = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
  double *tgt = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
  double *ref = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
  double *x = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
  if (!ws || !sum || !tgt || !ref || !x) {
    free(ws); free(sum); free(tgt); free(ref); free(x);
    return -1;
  }
  orc_avgcorr_windows(nframes, maxwindow, offset, ws, nw);
  for (int wi = 0; wi < nw; ++wi) {
    const int w = ws[wi];
    const double dw = (double)w;
    int first = ref_sel == NULL;
    if (ref_sel) memcpy(ref, ref_sel, sizeof(double) * 3 * (size_t)nsel);
    for (int k = 0; k < 3 * nsel; ++k) sum[k] = 0.0;
    double avg = 0.0, sq = 0.0;
    int sub = 0;
    for (int f = 0; f < nframes; ++f) {
      orc_gather(frames + (int64_t)f * stride, sel, nsel, x);
      for (int k = 0; k < 3 * nsel; ++k) sum[k] += x[k];
      if (f > w - 2) {
        for (int k = 0; k < 3 * nsel; ++k) tgt[k] = sum[k] / dw;
        if (first) {
          double c[3];
          memcpy(ref, tgt, sizeof(double) * 3 * (size_t)nsel);
          orc_center_on_origin(ref, mass, nsel, use_mass, c);
          first = 0;
        }
        double U[9], tr[3];
        const double r = orc_rmsd_centered_ref(tgt, ref, mass, nsel, use_mass, U, tr);
        avg += r;
        sq += r * r;
        orc_gather(frames + (int64_t)sub * stride, sel, nsel, x);
        for (int k = 0; k < 3 * nsel; ++k) sum[k] -= x[k];
        ++sub;
      }
    }
    const double d = 1.0 / ((double)nframes - dw + 1.0);
    avg *= d;
    sq *= d;
    sq -= avg * avg;
    ct[wi] = avg;
    sd[wi] = sq > 0.0 ? sqrt(sq) : 0.0;
  }
  free(ws); free(sum); free(tgt); free(ref); free(x);
  return nw;
}

/* ------------------------------------------------------------------------ */
/* Hungarian::Optimize (src/Hungarian.cpp:21-66) with AssignRowsToColumns
 * (:75-159), CoverZeroElements (:240-337) and UpdateMatrix (:165-190): the
 * reference's own heuristic form of the assignment, step for step (row then
 * column reduction; greedy zero assignment from the row or column with the
 * fewest free zeros, rows preferred on ties; greedy line cover by the most
 * zeros, later by zeros minus non-zeros, rows preferred on ties; add the
 * smallest uncovered value where lines cross and subtract it where none runs),
 * at most n*n rounds.  M (n x n, row-major, row = target atom, column =
 * reference atom) is modified.  map[col] receives the row assigned to column
 * col (-1 where none).  Returns the number of assignments.                   */
int orc_hungarian(double *M, int n, int *map)
{
  int *a_cr = (int *)malloc(sizeof(int) * (size_t)n);          /* row -> col */
  int *cnt = (int *)malloc(sizeof(int) * 4 * (size_t)n);
  unsigned char *lr = (unsigned char *)malloc((size_t)2 * n);
  if (!a_cr || !cnt || !lr) { free(a_cr); free(cnt); free(lr); return -1; }
  unsigned char *lc = lr + n;
  int *rz = cnt, *cz = cnt + n, *rnz = cnt + 2 * n, *cnz = cnt + 3 * n;
  for (int r = 0; r < n; ++r) {
    double mn = 1.7976931348623157e308;
    for (int c = 0; c < n; ++c) if (M[r * n + c] < mn) mn = M[r * n + c];
    for (int c = 0; c < n; ++c) M[r * n + c] -= mn;
  }
  for (int c = 0; c < n; ++c) {
    double mn = 1.7976931348623157e308;
    for (int r = 0; r < n; ++r) if (M[r * n + c] < mn) mn = M[r * n + c];
    for (int r = 0; r < n; ++r) M[r * n + c] -= mn;
  }
  int nassigned = 0;
  for (int it = 0; it < n * n; ++it) {
    /* greedy assignment of zeros */
    for (int k = 0; k < n; ++k) { map[k] = -1; a_cr[k] = -1; }
    nassigned = 0;
    for (;;) {
      int min_row = -1, min_row_z = n + 1;
      for (int r = 0; r < n; ++r) {
        if (a_cr[r] != -1) continue;
        int z = 0;
        for (int c = 0; c < n; ++c) if (map[c] == -1 && M[r * n + c] < ORC_SMALL) ++z;
        if (z != 0 && z < min_row_z) { min_row_z = z; min_row = r; }
      }
      int min_col = -1, min_col_z = n + 1;
      for (int c = 0; c < n; ++c) {
        if (map[c] != -1) continue;
        int z = 0;
        for (int r = 0; r < n; ++r) if (a_cr[r] == -1 && M[r * n + c] < ORC_SMALL) ++z;
        if (z != 0 && z < min_col_z) { min_col_z = z; min_col = c; }
      }
      if (min_row == -1 && min_col == -1) break;
      int done = 0;
      if (min_col_z < min_row_z) {
        for (int r = 0; r < n; ++r)
          if (M[r * n + min_col] < ORC_SMALL && a_cr[r] == -1) {
            map[min_col] = r; a_cr[r] = min_col; done = 1; break;
          }
      } else {
        for (int c = 0; c < n; ++c)
          if (M[min_row * n + c] < ORC_SMALL && map[c] == -1) {
            map[c] = min_row; a_cr[min_row] = c; done = 1; break;
          }
      }
      if (!done) break;
      ++nassigned;
    }
    if (nassigned == n) break;
    /* greedy line cover of the zeros */
    for (int k = 0; k < n; ++k) { lr[k] = 0; lc[k] = 0; rz[k] = cz[k] = rnz[k] = cnz[k] = 0; }
    int total = 0;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        if (M[r * n + c] < ORC_SMALL) { rz[r]++; cz[c]++; ++total; }
        else { rnz[r]++; cnz[c]++; }
      }
    for (int step = 0; total > 0; ++step) {
      int max_diff = -1, max_zero = -1, max_idx = -1, is_row = 1;
      for (int r = 0; r < n; ++r) {
        if (lr[r]) continue;
        const int d = step > 0 ? rz[r] - rnz[r] : rz[r];
        if (max_zero == -1 || d > max_diff) { max_diff = d; max_zero = rz[r]; max_idx = r; }
      }
      for (int c = 0; c < n; ++c) {
        if (lc[c]) continue;
        const int d = step > 0 ? cz[c] - cnz[c] : cz[c];
        if (d > max_diff) { max_diff = d; max_zero = cz[c]; max_idx = c; is_row = 0; }
      }
      if (max_idx < 0) break; /* every line drawn */
      if (is_row) {
        lr[max_idx] = 1;
        for (int c = 0; c < n; ++c) {
          if (M[max_idx * n + c] < ORC_SMALL) cz[c]--; else cnz[c]--;
        }
      } else {
        lc[max_idx] = 1;
        for (int r = 0; r < n; ++r) {
          if (M[r * n + max_idx] < ORC_SMALL) rz[r]--; else rnz[r]--;
        }
      }
      total -= max_zero;
    }
    /* update by the smallest uncovered element */
    double mu = 1.7976931348623157e308;
    for (int r = 0; r < n; ++r) {
      if (lr[r]) continue;
      for (int c = 0; c < n; ++c) if (!lc[c] && M[r * n + c] < mu) mu = M[r * n + c];
    }
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        if (lr[r] && lc[c]) M[r * n + c] += mu;
        else if (!lr[r] && !lc[c]) M[r * n + c] -= mu;
      }
  }
  free(a_cr); free(cnt); free(lr);
  return nassigned;
}

/* SymmetricRmsdCalc::SymmRMSD_CenteredRef (src/SymmetricRmsdCalc.cpp:125-193)
 * for one frame.  tgt: the selected target atoms (n x 3); ref: the selected
 * reference (centred when fitting, as ReferenceAction leaves it); mass: the
 * selected target masses.  Groups: atoms gatoms[goff[g] .. goff[g+1]) (indices
 * into the selection), as SymmetricRmsdCalc::SetupSymmRMSD leaves them.  With
 * fit: best fit of the target as given, rotated onto the reference (Frame::
 * Rotate, src/Frame.h:508-517); per group the squared distances
 * DIST2_NoImage(ref[ra], tgt'[ta]) (src/DistRoutines.cpp:411-419) into a
 * Hungarian cost matrix (row ta, column ra); amap[ra] = the target atom of
 * the assigned row; the target remapped (coordinates and masses,
 * Frame::SetCoordinatesByMap, src/Frame.cpp:765-797) and fitted again (or
 * RMSD_NoFit).  amap (n), U and trans (the final fit's, may be NULL) out.
 * A column the heuristic leaves unassigned keeps its own atom (the reference
 * would index the group at -1).                                               */
double orc_symm_rmsd(const double *tgt, const double *ref, const double *mass, int n,
                     int use_mass, int fit, int ngroups, const int *goff, const int *gatoms,
                     int *amap, double *U, double *trans)
{
  int gmax = 0;
  for (int g = 0; g < ngroups; ++g)
    if (goff[g + 1] - goff[g] > gmax) gmax = goff[g + 1] - goff[g];
  double *T = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *mr = (double *)malloc(sizeof(double) * (size_t)n);
  double *M = (double *)malloc(sizeof(double) * ((size_t)gmax * gmax + 1));
  int *map = (int *)malloc(sizeof(int) * ((size_t)gmax + 1));
  if (!T || !mr || !M || !map) { free(T); free(mr); free(M); free(map); return -2.0; }
  for (int i = 0; i < n; ++i) amap[i] = i;
  memcpy(T, tgt, sizeof(double) * 3 * (size_t)n);
  double Uw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tw[3] = {0, 0, 0};
  if (fit) {
    orc_rmsd_centered_ref(T, ref, mass, n, use_mass, Uw, tw);
    for (int i = 0; i < n; ++i) {
      const double x = T[3 * i], y = T[3 * i + 1], z = T[3 * i + 2];
      T[3 * i] = (x * Uw[0]) + (y * Uw[1]) + (z * Uw[2]);
      T[3 * i + 1] = (x * Uw[3]) + (y * Uw[4]) + (z * Uw[5]);
      T[3 * i + 2] = (x * Uw[6]) + (y * Uw[7]) + (z * Uw[8]);
    }
  }
  for (int g = 0; g < ngroups; ++g) {
    const int *ga = gatoms + goff[g];
    const int m = goff[g + 1] - goff[g];
    int e = 0;
    for (int ta = 0; ta < m; ++ta)
      for (int ra = 0; ra < m; ++ra) {
        const double *a1 = ref + 3 * ga[ra], *a2 = T + 3 * ga[ta];
        const double x = a1[0] - a2[0], y = a1[1] - a2[1], z = a1[2] - a2[2];
        M[e++] = (x * x + y * y + z * z);
      }
    orc_hungarian(M, m, map);
    for (int c = 0; c < m; ++c) amap[ga[c]] = map[c] >= 0 ? ga[map[c]] : ga[c];
  }
  for (int i = 0; i < n; ++i) {
    memcpy(T + 3 * i, tgt + 3 * amap[i], sizeof(double) * 3);
    mr[i] = mass ? mass[amap[i]] : 1.0;
  }
  double r;
  if (fit) {
    r = orc_rmsd_centered_ref(T, ref, mr, n, use_mass, Uw, tw);
  } else {
    r = orc_rmsd_nofit(T, ref, mr, n, use_mass);
  }
  if (U) memcpy(U, Uw, sizeof(Uw));
  if (trans) memcpy(trans, tw, sizeof(tw));
  free(T); free(mr); free(M); free(map);
  return r;
}

/* Action_SymmetricRmsd::DoAction (src/Action_SymmetricRmsd.cpp:88-112) over a
 * frame set against one prepared reference: gather, symmetric RMSD, and per
 * frame the atom map (amap, F x nsel, may be NULL), U (F x 9) and the
 * target translation (F x 3), both may be NULL.                              */
int orc_symm_frames(const float *frames, int64_t stride, int nframes, const int *sel, int nsel,
                    const double *mass, const double *ref_sel, int use_mass, int fit,
                    int ngroups, const int *goff, const int *gatoms, double *rmsd, int *amap,
                    double *U, double *trans)
{
  int err = 0;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *T = (double *)malloc(sizeof(double) * 3 * (size_t)nsel);
    int *am = (int *)malloc(sizeof(int) * (size_t)nsel);
    if (!T || !am) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
    }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int f = 0; f < nframes; ++f) {
      if (!T || !am) continue;
      orc_gather(frames + (int64_t)f * stride, sel, nsel, T);
      double Uf[9], tf[3];
      rmsd[f] = orc_symm_rmsd(T, ref_sel, mass, nsel, use_mass, fit, ngroups, goff, gatoms, am,
                              Uf, tf);
      if (amap) memcpy(amap + (int64_t)f * nsel, am, sizeof(int) * (size_t)nsel);
      if (U) memcpy(U + 9 * (int64_t)f, Uf, sizeof(Uf));
      if (trans) memcpy(trans + 3 * (int64_t)f, tf, sizeof(tf));
    }
    free(T);
    free(am);
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* Box::RotateUcell (src/Box.cpp:732-745), applied after the rotate +
 * translate update of Action_Rmsd::DoAction (src/Action_Rmsd.cpp:381): each
 * unit-cell vector (a row of ucell) is rotated by the Kabsch U, row k' = U row k.
 * The reference recomputes the fractional cell from the result; that is the
 * inverse of ucell' and left to the caller.                                 */
void orc_rotate_ucell(const double U[9], const double ucell[9], double out[9])
{
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < 3; ++c)
      out[3 * k + c] = U[3 * c] * ucell[3 * k] + U[3 * c + 1] * ucell[3 * k + 1] +
                       U[3 * c + 2] * ucell[3 * k + 2];
}
