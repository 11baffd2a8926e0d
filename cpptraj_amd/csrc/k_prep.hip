// k_prep.hip -- per-frame stage of the pairwise RMSD: centroid, E0 term and the
// fragment-packed operands of the MFMA contraction.
//
// Replaces the per-pair work the reference repeats inside its loop nest
// (Analysis_Rms2d::Calculate_2D, src/Analysis_Rms2d.cpp:262-275): the masked
// float->double gather (CompactFrameArray::GetToMaskDblPtr,
// src/CompactFrameArray.cpp:244-262), centring (Frame::CenterOnOrigin,
// src/Frame.cpp:1043-1055; the target centroid in Frame::RMSD_CenteredRef,
// src/Frame.cpp:1141-1171) and the E0 half-sums (src/Frame.cpp:1194).  These
// depend only on the frame, so they run once per frame here.
//
// One wave per frame (64-thread workgroups: 14 % faster than 256 on the
// 50k-frame matrix).  All sums are FP64.  With weights the coordinates
// are stored pre-multiplied by sqrt(w), so sum_k w_k t_k r_k^T becomes a plain
// dot product of the stored rows.
//
// For the MFMA-reduced precisions each frame is additionally rotated (FP64
// QCP) onto a common template T and only the deviation D = U x~ - T is
// stored at reduced precision.  The best-fit RMSD of a pair is unchanged by
// rigid motions of either frame, and D is a few Angstrom where x~ is tens, so
// the reduced-precision contraction only sees small numbers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

constexpr int PREP_THREADS = 64;

template <int PREC>
__device__ __forceinline__ int64_t packed_elem(int64_t f, int k, int c, int p, int KB) {
  constexpr int CH = chunk_atoms(PREC);
  constexpr int PER = CH / 4;
  constexpr int PARTS = chunk_parts(PREC);
  const int64_t fb = f >> 4;
  const int r = (int)(f & 15);
  const int kb = k / CH, kk = k % CH;
  const int lane = r + 16 * (kk / PER);
  const int e = kk % PER;
  const int64_t piece = ((((fb * KB + kb) * 3 + c) * PARTS + p) << 6) + lane;
  return piece * PER + e;
}

// Weighted centroid of frame f: returns (sum w x, sum w).
__device__ __forceinline__ void centroid(const FrameSetDev& fs, int64_t f, const double* w,
                                         double* red, double cen[3]) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < fs.nsel; k += blockDim.x) {
    const double m = w ? gp(w)[k] : 1.0;
    float p[3];
    load_atom(fs, f, k, p);
    v[0] += m * (double)p[0];
    v[1] += m * (double)p[1];
    v[2] += m * (double)p[2];
    v[3] += m;
  }
  rmsd_dev::block_sum<4>(v, red);
  const double tot = w ? v[3] : (double)fs.nsel;
  const double inv = tot != 0.0 ? 1.0 / tot : 0.0;
  cen[0] = v[0] * inv;
  cen[1] = v[1] * inv;
  cen[2] = v[2] * inv;
}

__device__ __forceinline__ void centred(const FrameSetDev& fs, int64_t f, int k, const double cen[3],
                                        const double* sw, double x[3]) {
  if (k < fs.nsel) {
    const double s = sw ? sqrt(gp(sw)[k]) : 1.0;
    float p[3];
    load_atom(fs, f, k, p);
    x[0] = ((double)p[0] - cen[0]) * s;
    x[1] = ((double)p[1] - cen[1]) * s;
    x[2] = ((double)p[2] - cen[2]) * s;
  } else {
    x[0] = x[1] = x[2] = 0.0;
  }
}

// Atoms 4q ... 4q + 3 of frame f: one quad of 16-byte loads when the layout
// allows it (quad_layout), else atom by atom; zeros beyond the selection.
__device__ __forceinline__ void atoms_quad(const FrameSetDev& fs, int64_t f, int q, int lay,
                                           float (&x)[4], float (&y)[4], float (&z)[4]) {
  if (lay && 4 * q + 3 < fs.nsel) {
    load_quad(fs, f, q, lay, x, y, z);
    return;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (4 * q + e < fs.nsel) load_atom(fs, f, 4 * q + e, p);
    x[e] = p[0];
    y[e] = p[1];
    z[e] = p[2];
  }
}

// The template's atoms 4q ... 4q + 3 (rows of T, Kp a multiple of 4 and the
// rows 32-byte aligned): two 16-byte loads per row.
__device__ __forceinline__ void template_quad(const double* T, int Kp, int q, double (&t)[4][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    typedef __attribute__((ext_vector_type(2))) double f64x2;
    const auto row = gp(reinterpret_cast<const f64x2*>(T + (int64_t)c * Kp + 4 * q));
    const f64x2 a = row[0], b = row[1];
    t[0][c] = a.x;
    t[1][c] = a.y;
    t[2][c] = b.x;
    t[3][c] = b.y;
  }
}

// ---- template / total weight -----------------------------------------------
__device__ __forceinline__ void setup_body(const SetupArgs& a) {
  __shared__ double red[16 * (PREP_THREADS / 64)];
  if (threadIdx.x == 0) {
    // Sequential running total, rounded as the reference's per-pair sum; the
    // weights are loaded 16 at a time so that their loads overlap.
    double tw = 0.0;
    if (a.smass) {
      const auto m = gp(a.smass);
      const int n = a.fs.nsel;
      int k = 0;
      for (; k + 16 <= n; k += 16) {
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = m[k + e];
#pragma unroll
        for (int e = 0; e < 16; ++e) tw += v[e];
      }
      for (; k < n; ++k) tw += m[k];
    } else {
      tw = (double)a.fs.nsel;
    }
    a.total_mass[0] = tw;
  }
  if (!a.T) return;
  double cen[3];
  centroid(a.fs, a.f0, a.cmass, red, cen);
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < a.Kp; k += blockDim.x) {
    double x[3];
    centred(a.fs, a.f0, k, cen, a.smass, x);
    a.T[k] = x[0];
    a.T[a.Kp + k] = x[1];
    a.T[2 * a.Kp + k] = x[2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) v[3 * c + d] += x[c] * x[d];
    v[9] += x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  }
  rmsd_dev::block_sum<10>(v, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 9; ++k) a.P[k] = v[k];
    a.gT[0] = v[9];
  }
}

__global__ __launch_bounds__(PREP_THREADS) void k_setup(SetupArgs a) { setup_body(a); }

// ---- mean-structure template ---------------------------------------------------
// The MFMA-reduced forms contract deviations D = U x~ - T, and their error
// grows with ||D_i|| ||D_j||: a template near the middle of the trajectory
// keeps D small.  Sample s (evenly spaced frame of the set) is rotated onto the
// first template (FP64 QCP, as in k_prep) and written to part[s][3][Kp];
// k_tmpl_mean averages the samples in a fixed order (deterministic) and
// k_tmpl_consts recomputes P = T^T T and sum |T|^2.
__device__ __forceinline__ void tmpl_sample_body(const SetupArgs& a, int s, int nsamp,
                                                 double* part) {
  __shared__ double red[16 * (PREP_THREADS / 64)];
  const int64_t f = ((int64_t)s * a.fs.nframes) / nsamp;
  const int Kp = a.Kp;
  double cen[3];
  centroid(a.fs, f, a.cmass, red, cen);
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < a.fs.nsel; k += blockDim.x) {
    double x[3];
    centred(a.fs, f, k, cen, a.smass, x);
    const double t0 = a.T[k], t1 = a.T[Kp + k], t2 = a.T[2 * Kp + k];
    v[0] += x[0] * t0; v[1] += x[0] * t1; v[2] += x[0] * t2;
    v[3] += x[1] * t0; v[4] += x[1] * t1; v[5] += x[1] * t2;
    v[6] += x[2] * t0; v[7] += x[2] * t1; v[8] += x[2] * t2;
    v[9] += x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  }
  rmsd_dev::block_sum<10>(v, red);
  double U[9];
  {
    const double E0 = 0.5 * (v[9] + a.gT[0]);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(v);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    rmsd_dev::qcp_rotation(v, lam, U);
  }
  double* out = part + (int64_t)s * 3 * Kp;
  for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
    double x[3];
    centred(a.fs, f, k, cen, a.smass, x);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[c * Kp + k] = U[3 * c] * x[0] + U[3 * c + 1] * x[1] + U[3 * c + 2] * x[2];
  }
}

__global__ __launch_bounds__(PREP_THREADS) void k_tmpl_sample(SetupArgs a, int nsamp,
                                                              double* part) {
  tmpl_sample_body(a, blockIdx.x, nsamp, part);
}

// Batched: workgroup b is sample b - sbase[k] of set k, sbase[k] <= b < sbase[k + 1].
__global__ __launch_bounds__(PREP_THREADS) void k_tmpl_sample_batch(const SetupArgs* sets,
                                                                    const int* sbase, int nsets) {
  const int b = blockIdx.x;
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sbase[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const SetupArgs a = sets[lo];
  tmpl_sample_body(a, b - sbase[lo], a.nsamp, a.part);
}

__device__ __forceinline__ void tmpl_mean_elem(const double* part, int nsamp, int n, int e,
                                               double* T) {
  if (e >= n) return;
  double acc = 0.0;
  for (int s = 0; s < nsamp; ++s) acc += part[(int64_t)s * n + e];
  T[e] = acc / (double)nsamp;
}

__global__ __launch_bounds__(256) void k_tmpl_mean(const double* part, int nsamp, int n,
                                                   double* T) {
  tmpl_mean_elem(part, nsamp, n, blockIdx.x * blockDim.x + threadIdx.x, T);
}

// Batched: blockIdx.y is the set.
__global__ __launch_bounds__(256) void k_tmpl_mean_batch(const SetupArgs* sets) {
  const SetupArgs& a = sets[blockIdx.y];
  tmpl_mean_elem(a.part, a.nsamp, 3 * a.Kp, blockIdx.x * blockDim.x + threadIdx.x, a.T);
}

__device__ __forceinline__ void tmpl_consts_body(const SetupArgs& a) {
  __shared__ double red[16 * (PREP_THREADS / 64)];
  const int Kp = a.Kp;
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
    const double x[3] = {a.T[k], a.T[Kp + k], a.T[2 * Kp + k]};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) v[3 * c + d] += x[c] * x[d];
    v[9] += x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  }
  rmsd_dev::block_sum<10>(v, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 9; ++k) a.P[k] = v[k];
    a.gT[0] = v[9];
  }
}

__global__ __launch_bounds__(PREP_THREADS) void k_tmpl_consts(SetupArgs a) { tmpl_consts_body(a); }

__global__ __launch_bounds__(PREP_THREADS) void k_tmpl_consts_batch(const SetupArgs* sets) {
  const SetupArgs a = sets[blockIdx.x];
  tmpl_consts_body(a);
}

// One workgroup per set (batched matrices).
__global__ __launch_bounds__(PREP_THREADS) void k_setup_batch(const SetupArgs* sets) {
  const SetupArgs a = sets[blockIdx.x];
  setup_body(a);
}

// ---- per-frame preparation -----------------------------------------------------
// QUADS = false compiles the atom-by-atom path only (fewer registers and no
// branch to the quad code: the batched launch for sets without quads).
template <int PREC, bool QUADS = true>
__device__ __forceinline__ void prep_body(const PrepArgs& a, int64_t f) {
  __shared__ double red[24 * (PREP_THREADS / 64)];
  const int Kp = a.Kp;

  if (f >= a.fs.nframes) {
    // Padding frames: zero operands so that padded rows and columns of a tile
    // contribute nothing.
    if (a.nrm && threadIdx.x < 4) gp(a.nrm)[4 * f + threadIdx.x] = 0.0;
    if (a.cen && threadIdx.x < 4) gp(a.cen)[4 * f + threadIdx.x] = 0.0;
    for (int k = threadIdx.x; k < Kp; k += blockDim.x)
      for (int c = 0; c < 3; ++c)
        for (int p = 0; p < chunk_parts(PREC); ++p) {
          const int64_t e = packed_elem<PREC>(f, k, c, p, a.KB);
          if (PREC == PREC_F64) gp(reinterpret_cast<double*>(a.packed))[e] = 0.0;
          else if (PREC == PREC_F32) gp(reinterpret_cast<float*>(a.packed))[e] = 0.0f;
          else gp(reinterpret_cast<__bf16*>(a.packed))[e] = (__bf16)0.0f;
        }
    return;
  }

  // MFMA forms on frames that allow 16-byte loads of four atoms: each lane
  // takes quads of atoms (four per step; the 50,000 x 1,200 AoS stage: 1.05 ->
  // 0.86 ms).  Otherwise lanes take consecutive atoms (coalesced 4-byte loads;
  // quads of such atoms would spread each load over four times the lines).
  const int lay = PREC == PREC_F64 || !QUADS ? 0 : quad_layout(a.fs);
  double cen[3];
  if (PREC == PREC_F64) {
    centroid(a.fs, f, a.cmass, red, cen);
    if (a.cen && threadIdx.x == 0) {
      gp(a.cen)[4 * f] = cen[0];
      gp(a.cen)[4 * f + 1] = cen[1];
      gp(a.cen)[4 * f + 2] = cen[2];
      gp(a.cen)[4 * f + 3] = 0.0;
    }
    double g[1] = {0.0};
    const auto out = gp(reinterpret_cast<double*>(a.packed));
    for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
      double x[3];
      centred(a.fs, f, k, cen, a.smass, x);
      out[packed_elem<PREC>(f, k, 0, 0, a.KB)] = x[0];
      out[packed_elem<PREC>(f, k, 1, 0, a.KB)] = x[1];
      out[packed_elem<PREC>(f, k, 2, 0, a.KB)] = x[2];
      g[0] += x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    }
    rmsd_dev::block_sum<1>(g, red);
    if (threadIdx.x == 0) gp(a.g)[f] = g[0];
    return;
  }

  // MFMA-reduced forms: align to the template in FP64.  One pass over the
  // frame gives the centroid and the moments about the origin: with s = sqrt(w)
  // the fit weights and c the centroid,
  //   M = sum s (x - c) t^T = sum s x t^T - c (sum s t)^T,
  // (sum s t is zero up to rounding for a template of centred frames), and the
  // QCP's Newton start E0 = (G + |T|^2) / 2 takes
  //   G = sum w |x - c|^2 = sum w |x|^2 - 2 c . sum w x + |c|^2 sum w.
  // The G stored for the pairs is summed from the centred coordinates below.
  // v: sum wc x (3), sum wc, sum s x_a t_b (9), sum s t (3), sum w x (3),
  // sum w, sum w |x|^2 (wc: centroid weights, w: fit weights).
  double v[21];
#pragma unroll
  for (int i = 0; i < 21; ++i) v[i] = 0.0;
  auto origin_moments = [&](int k, float px, float py, float pz, const double t[3]) {
    const double wc = a.cmass ? gp(a.cmass)[k] : 1.0;
    const double w = a.smass ? gp(a.smass)[k] : 1.0;
    const double sq = a.smass ? sqrt(w) : 1.0;
    const double x[3] = {(double)px, (double)py, (double)pz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] += wc * x[c];
      const double sx = sq * x[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) v[4 + 3 * c + d] += sx * t[d];
      v[13 + c] += sq * t[c];
      v[16 + c] += w * x[c];
    }
    v[3] += wc;
    v[19] += w;
    v[20] += w * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  };
  const int nq = (a.fs.nsel + 3) >> 2;
  if (lay) {
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
      float px[4], py[4], pz[4];
      double t[4][3];
      atoms_quad(a.fs, f, q, lay, px, py, pz);
      template_quad(a.T, Kp, q, t);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * q + e < a.fs.nsel) origin_moments(4 * q + e, px[e], py[e], pz[e], t[e]);
    }
  } else {
    for (int k = threadIdx.x; k < a.fs.nsel; k += blockDim.x) {
      float p[3];
      load_atom(a.fs, f, k, p);
      const double t[3] = {gp(a.T)[k], gp(a.T)[Kp + k], gp(a.T)[2 * Kp + k]};
      origin_moments(k, p[0], p[1], p[2], t);
    }
  }
  rmsd_dev::block_sum<21>(v, red);
  {
    const double tot = a.cmass ? v[3] : (double)a.fs.nsel;
    const double inv = tot != 0.0 ? 1.0 / tot : 0.0;
    cen[0] = v[0] * inv;
    cen[1] = v[1] * inv;
    cen[2] = v[2] * inv;
  }
  if (a.cen && threadIdx.x == 0) {
    gp(a.cen)[4 * f] = cen[0];
    gp(a.cen)[4 * f + 1] = cen[1];
    gp(a.cen)[4 * f + 2] = cen[2];
    gp(a.cen)[4 * f + 3] = 0.0;
  }
  // Every thread holds the block totals and solves the same FP64 QCP (same
  // inputs, same operations), so the rotation needs no broadcast or barrier.
  double U[9];
  {
    double Mc[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) Mc[3 * c + d] = v[4 + 3 * c + d] - cen[c] * v[13 + d];
    const double G = v[20] - 2.0 * (cen[0] * v[16] + cen[1] * v[17] + cen[2] * v[18]) +
                     (cen[0] * cen[0] + cen[1] * cen[1] + cen[2] * cen[2]) * v[19];
    const double E0 = 0.5 * (G + gp(a.gT)[0]);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(Mc);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    rmsd_dev::qcp_rotation(Mc, lam, U);
  }
  // centred, weighted coordinates of atoms 4q ... 4q + 3 (zeros past nsel)
  auto centred_quad = [&](int q, double (&x)[4][3]) {
    float px[4], py[4], pz[4];
    atoms_quad(a.fs, f, q, lay, px, py, pz);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * q + e;
      if (k < a.fs.nsel) {
        const double s = a.smass ? sqrt(gp(a.smass)[k]) : 1.0;
        x[e][0] = ((double)px[e] - cen[0]) * s;
        x[e][1] = ((double)py[e] - cen[1]) * s;
        x[e][2] = ((double)pz[e] - cen[2]) * s;
      } else {
        x[e][0] = x[e][1] = x[e][2] = 0.0;
      }
    }
  };

  // u = T^T D (9), then the sums behind the error bound's norms: ||D||^2,
  // sum_k D_c^4 per coordinate, and ||lo||^2 (the split's low parts; for FP32
  // the rounding residual D - float(D)).
  // u[14]: G = sum |x~|^2, the frame's E0 term.
  double u[15] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // one atom: its D (accumulated into u) rounded to the stored parts
  auto emit = [&](const double x[3], const double t[3], float (&hi)[3], float (&lo)[3]) {
    u[14] += x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    double D[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      D[c] = U[3 * c] * x[0] + U[3 * c + 1] * x[1] + U[3 * c + 2] * x[2] - t[c];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) u[3 * c + d] += t[c] * D[d];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double d2 = D[c] * D[c];
      u[9] += d2;
      u[10 + c] += d2 * d2;
      if (PREC == PREC_F32) {
        const float df = (float)D[c];
        hi[c] = df;
        const double r = D[c] - (double)df;
        u[13] += r * r;
      } else {
        const __bf16 h = (__bf16)(float)D[c];
        const __bf16 l = (__bf16)(float)(D[c] - (double)(float)h);
        hi[c] = (float)h;
        lo[c] = (float)l;
        const double ld = (double)(float)l;
        u[13] += ld * ld;
      }
    }
  };
  // Kp is a multiple of the chunk, so a quad never straddles two chunks and its
  // four values of one coordinate and part are consecutive in a 16-byte slot:
  // one 8-byte (split-BF16) or 16-byte (FP32) store each.
  for (int q = threadIdx.x; lay && q < (Kp >> 2); q += blockDim.x) {
    double x[4][3], t[4][3];
    centred_quad(q, x);
    template_quad(a.T, Kp, q, t);
    float hi[4][3], lo[4][3];
#pragma unroll
    for (int e = 0; e < 4; ++e) emit(x[e], t[e], hi[e], lo[e]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (PREC == PREC_F32) {
        f32x4_t v4 = {hi[0][c], hi[1][c], hi[2][c], hi[3][c]};
        *gp(reinterpret_cast<f32x4_t*>(reinterpret_cast<float*>(a.packed) +
                                       packed_elem<PREC>(f, 4 * q, c, 0, a.KB))) = v4;
      } else {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const float (&src)[4][3] = p ? lo : hi;
          // bfloat16 bits: the top half of each float (exact: the values are bfloat16)
          typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;
          const u32x2 w2 = {(__float_as_uint(src[0][c]) >> 16) | (__float_as_uint(src[1][c]) & 0xffff0000u),
                            (__float_as_uint(src[2][c]) >> 16) | (__float_as_uint(src[3][c]) & 0xffff0000u)};
          *gp(reinterpret_cast<u32x2*>(reinterpret_cast<__bf16*>(a.packed) +
                                       packed_elem<PREC>(f, 4 * q, c, p, a.KB))) = w2;
        }
      }
    }
  }
  for (int k = threadIdx.x; !lay && k < Kp; k += blockDim.x) {
    double x[3], t[3];
    centred(a.fs, f, k, cen, a.smass, x);
    t[0] = gp(a.T)[k];
    t[1] = gp(a.T)[Kp + k];
    t[2] = gp(a.T)[2 * Kp + k];
    float hi[3], lo[3];
    emit(x, t, hi, lo);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (PREC == PREC_F32) {
        gp(reinterpret_cast<float*>(a.packed))[packed_elem<PREC>(f, k, c, 0, a.KB)] = hi[c];
      } else {
        const auto out = gp(reinterpret_cast<__bf16*>(a.packed));
        out[packed_elem<PREC>(f, k, c, 0, a.KB)] = (__bf16)hi[c];
        out[packed_elem<PREC>(f, k, c, 1, a.KB)] = (__bf16)lo[c];
      }
    }
  }
  rmsd_dev::block_sum<15>(u, red);
  if (threadIdx.x == 0) {
    gp(a.g)[f] = u[14];
    for (int k = 0; k < 9; ++k) gp(a.u)[9 * f + k] = u[k];
    if (a.nrm) {
      // The error bound's per-frame coefficient (rmsd_dev::bound_keep).
      const double d2 = u[9];                                  // ||D||^2
      const double b2 = sqrt(u[10]) + sqrt(u[11]) + sqrt(u[12]);  // sum_c ||D_c||_4^2
      const double l2 = u[13];                                 // ||l||^2
      const BoundCoef& k = a.bound;
      // c^2 = kdrop l^2 + krep b^2 + (sqrt(kacc) ||D|| + kmix / sqrt(kacc) l)^2
      // (rmsd_dev::bound_keep; kacc > 0 for every MFMA form).
      const double sa = sqrt(k.acc);
      const double t = sa * sqrt(d2) + (k.mix / sa) * sqrt(l2);
      const double c2 = k.drop * l2 + k.rep * b2 + t * t;
      const double M = a.total_mass ? gp(a.total_mass)[0] : 0.0;
      gp(a.nrm)[4 * f] = M >= RMSD_SMALL ? sqrt(c2 * (2.0 * 1.7320508075688772 / M)) : 0.0;
      gp(a.nrm)[4 * f + 1] = sqrt(d2);
      gp(a.nrm)[4 * f + 2] = sqrt(b2);
      gp(a.nrm)[4 * f + 3] = sqrt(l2);
    }
  }
}

// Workgroup -> frame.  The hardware deals consecutive workgroups round-robin
// over the 8 XCDs, and the 16 frames of a packed frame block write interleaved
// 16-byte slots of the same 1 KiB pieces; over 8 L2s each 128-byte line went
// out in partial writes from 8 places.  Here the 16 frames of a block run on
// one XCD, whose L2 merges their slots before the line is written back
// (k_prep_batch 7.8 -> 5.9 ms on the 64-trajectory batch).  The map is a
// bijection of [0, prep_grid(n)) onto itself.
constexpr int PREP_XCDS = 8;
__device__ __forceinline__ int64_t prep_frame(int64_t b) {
  const int64_t xcd = b % PREP_XCDS, i = b / PREP_XCDS;
  return ((i / 16) * PREP_XCDS + xcd) * 16 + (i % 16);
}

int64_t prep_grid(int64_t n) {
  return (n + 16 * PREP_XCDS - 1) / (16 * PREP_XCDS) * 16 * PREP_XCDS;
}

template <int PREC>
__global__ __launch_bounds__(PREP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_prep(PrepArgs a) {
  const int64_t f = prep_frame(blockIdx.x);
  if (f < a.Fp) prep_body<PREC>(a, f);
}

// Every padded frame of every set of a batch in one launch: workgroup b is
// frame b - fbase[s] of the set s with fbase[s] <= b < fbase[s + 1].
// Frames fbase[0] ... fbase[nsets] - 1 of sets[0 .. nsets) (fbase: prefix of
// the padded frame counts, which may start above zero).
template <int PREC, bool QUADS>
__global__ __launch_bounds__(PREP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_prep_batch(
    const PrepArgs* sets, const int* fbase, int nsets) {
  // every set's padded frame count is a multiple of 16, so no frame block
  // straddles two sets
  const int b = fbase[0] + (int)prep_frame(blockIdx.x);
  if (b >= fbase[nsets]) return;
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fbase[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const PrepArgs a = sets[lo];
  prep_body<PREC, QUADS>(a, b - fbase[lo]);
}

}  // namespace

hipError_t rmsd_launch_setup(const SetupArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_setup, dim3(1), dim3(PREP_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t rmsd_launch_template_mean(const SetupArgs& a, int nsamp, double* part,
                                     hipStream_t s) {
  if (!a.T || nsamp < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tmpl_sample, dim3(nsamp), dim3(PREP_THREADS), 0, s, a, nsamp, part);
  const int n = 3 * a.Kp;
  hipLaunchKernelGGL(k_tmpl_mean, dim3((n + 255) / 256), dim3(256), 0, s, part, nsamp, n, a.T);
  hipLaunchKernelGGL(k_tmpl_consts, dim3(1), dim3(PREP_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t rmsd_launch_template_mean_batch(const SetupArgs* sets, const int* sbase, int nsets,
                                           int total_samples, int maxk, hipStream_t s) {
  if (nsets <= 0 || total_samples <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tmpl_sample_batch, dim3(total_samples), dim3(PREP_THREADS), 0, s, sets,
                     sbase, nsets);
  hipLaunchKernelGGL(k_tmpl_mean_batch, dim3((maxk + 255) / 256, nsets), dim3(256), 0, s, sets);
  hipLaunchKernelGGL(k_tmpl_consts_batch, dim3(nsets), dim3(PREP_THREADS), 0, s, sets);
  return hipGetLastError();
}

hipError_t rmsd_launch_setup_batch(const SetupArgs* sets, int nsets, hipStream_t s) {
  if (nsets <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_batch, dim3(nsets), dim3(PREP_THREADS), 0, s, sets);
  return hipGetLastError();
}

hipError_t rmsd_launch_prep_batch(int prec, const PrepArgs* sets, const int* fbase, int nsets,
                                  int nquad, int quad_frames, int total_frames, hipStream_t s) {
  if (nsets <= 0 || total_frames <= 0) return hipSuccess;
  if (nquad < 0 || nquad > nsets || quad_frames < 0 || quad_frames > total_frames)
    return hipErrorInvalidValue;
  const int rest = total_frames - quad_frames;
  switch (prec) {
    case PREC_F64:
      hipLaunchKernelGGL((k_prep_batch<PREC_F64, false>), dim3(prep_grid(total_frames)),
                         dim3(PREP_THREADS), 0, s, sets, fbase, nsets);
      break;
    case PREC_F32:
      if (nquad > 0)
        hipLaunchKernelGGL((k_prep_batch<PREC_F32, true>), dim3(prep_grid(quad_frames)),
                           dim3(PREP_THREADS), 0, s, sets, fbase, nquad);
      if (nquad < nsets)
        hipLaunchKernelGGL((k_prep_batch<PREC_F32, false>), dim3(prep_grid(rest)), dim3(PREP_THREADS),
                           0, s, sets + nquad, fbase + nquad, nsets - nquad);
      break;
    case PREC_BF16X3:
      if (nquad > 0)
        hipLaunchKernelGGL((k_prep_batch<PREC_BF16X3, true>), dim3(prep_grid(quad_frames)),
                           dim3(PREP_THREADS), 0, s, sets, fbase, nquad);
      if (nquad < nsets)
        hipLaunchKernelGGL((k_prep_batch<PREC_BF16X3, false>), dim3(prep_grid(rest)),
                           dim3(PREP_THREADS), 0, s, sets + nquad, fbase + nquad, nsets - nquad);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t rmsd_launch_prep(const PrepArgs& a, hipStream_t s) {
  switch (a.prec) {
    case PREC_F64:
      hipLaunchKernelGGL(k_prep<PREC_F64>, dim3(prep_grid(a.Fp)), dim3(PREP_THREADS), 0, s, a);
      break;
    case PREC_F32:
      hipLaunchKernelGGL(k_prep<PREC_F32>, dim3(prep_grid(a.Fp)), dim3(PREP_THREADS), 0, s, a);
      break;
    case PREC_BF16X3:
      hipLaunchKernelGGL(k_prep<PREC_BF16X3>, dim3(prep_grid(a.Fp)), dim3(PREP_THREADS), 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
