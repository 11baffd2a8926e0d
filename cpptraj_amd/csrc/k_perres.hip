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
// Layout.  The residue list is cut by the host into pieces: runs of whole
// residues whose LDS image (per entry: mass, reference atom, offset; per
// residue: reference centre; one frame's target atoms) fits PERRES_LDS_PIECE
// bytes (a single larger residue gets a piece of its own, up to
// PERRES_LDS_MAX).  A piece stages either the contiguous span of target atoms
// it touches (coalesced copy) or, when its atoms are scattered (a span more
// than twice its atom count), just its own atoms (gather).  A workgroup
// owns one piece and walks over frame batches.  It keeps the piece's
// reference atoms (FP32, exact), target masses and atom offsets in LDS for its
// lifetime, and stages the atom span of `nfb` frames at a time into LDS with
// coalesced loads; then one thread computes one (frame, residue) from LDS,
// summing atom by atom in the reference's order.  HBM traffic is the frames'
// atoms once (12 B per atom and frame) plus 8 B per result.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_kernels.hpp"

namespace {

constexpr double kSmall = 0.00000000000001;  // Constants::SMALL, src/Constants.h:35

__device__ __forceinline__ double3 moved(const float* p, const double* X, int modify) {
#pragma clang fp contract(off)
  double x = (double)p[0], y = (double)p[1], z = (double)p[2];
  if (modify == 1) {
    x += X[9];
    y += X[10];
    z += X[11];
    return make_double3(x * X[0] + y * X[1] + z * X[2] + X[12],
                        x * X[3] + y * X[4] + z * X[5] + X[13],
                        x * X[6] + y * X[7] + z * X[8] + X[14]);
  }
  if (modify == 2) return make_double3(x + (X[9] + X[12]), y + (X[10] + X[13]), z + (X[11] + X[14]));
  return make_double3(x, y, z);
}

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
  // LDS: tmass[ne] double | cref[3 nr] double | rref[3 ne] float | toff[ne] int |
  //      frames[nfb][3 span] float
  double* tm = (double*)lds;
  double* cref = tm + ne;
  float* rref = (float*)(cref + 3 * nr);
  int* toff = (int*)(rref + 3 * ne);
  float* fx = (float*)(toff + ne);

  const float* rfr = a.ref.xyz + frame_of(a.ref, a.ref_frame) * a.ref.stride;
  for (int k = threadIdx.x; k < ne; k += blockDim.x) {
    const int q = a.ref_atoms[e0 + k];
    for (int c = 0; c < 3; ++c) rref[3 * k + c] = a.ref.ld ? rfr[c * a.ref.ld + q] : rfr[3 * q + c];
    tm[k] = a.use_mass ? a.tmass[e0 + k] : 1.0;
    toff[k] = gather ? k : a.tgt_atoms[e0 + k] - lo;
  }
  __syncthreads();
  // Reference residue centres (CenterOnOrigin with the reference masses).
  for (int r = threadIdx.x; r < nr; r += blockDim.x) {
    double c[3] = {0.0, 0.0, 0.0};
    if (a.center) {
      const int b = a.res_off[r0 + r] - e0, n = a.res_off[r0 + r + 1] - e0 - b;
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

  const int nfb = a.nfb;
  const int fspan = 3 * span;
  for (int64_t f0 = (int64_t)blockIdx.x * nfb; f0 < a.nframes; f0 += (int64_t)gridDim.x * nfb) {
    const int nf = (int)((a.nframes - f0) < nfb ? (a.nframes - f0) : nfb);
    __syncthreads();  // previous batch done with fx (and the set-up above)
    for (int fs = 0; fs < nf; ++fs) {
      const int64_t f = f0 + fs;
      const float* fr = a.xyz + (a.fidx ? (int64_t)a.fidx[f] : f) * a.stride;
      float* dst = fx + fs * fspan;
      if (gather) {
        for (int k = threadIdx.x; k < ne; k += blockDim.x) {
          const int at = a.tgt_atoms[e0 + k];
          for (int c = 0; c < 3; ++c) dst[3 * k + c] = a.ld ? fr[c * a.ld + at] : fr[3 * (int64_t)at + c];
        }
      } else if (a.ld) {
        for (int c = 0; c < 3; ++c)
          for (int at = threadIdx.x; at < span; at += blockDim.x)
            dst[3 * at + c] = fr[c * a.ld + lo + at];
      } else {
        const float* src = fr + 3 * (int64_t)lo;
        for (int w = threadIdx.x; w < fspan; w += blockDim.x) dst[w] = src[w];
      }
    }
    __syncthreads();
    for (int it = threadIdx.x; it < nf * nr; it += blockDim.x) {
      const int fs = it / nr, r = it - fs * nr;
      const int64_t f = f0 + fs;
      double X[15];
      if (a.modify) {
        for (int k = 0; k < 9; ++k) X[k] = a.U ? a.U[9 * f + k] : (k % 4 == 0 ? 1.0 : 0.0);
        for (int k = 0; k < 6; ++k) X[9 + k] = a.tvec[6 * f + k];
      }
      const float* frm = fx + fs * fspan;
      const int b = a.res_off[r0 + r] - e0, n = a.res_off[r0 + r + 1] - e0 - b;
      double ct[3] = {0.0, 0.0, 0.0};
      if (a.center) {
        double s[3] = {0.0, 0.0, 0.0}, total = 0.0;
        for (int k = b; k < b + n; ++k) {
          const double3 t = moved(frm + 3 * toff[k], X, a.modify);
          if (a.use_mass) {
            s[0] += t.x * tm[k];
            s[1] += t.y * tm[k];
            s[2] += t.z * tm[k];
            total += tm[k];
          } else {
            s[0] += t.x;
            s[1] += t.y;
            s[2] += t.z;
          }
        }
        if (a.use_mass) {
          if (total != 0.0)
            for (int d = 0; d < 3; ++d) ct[d] = s[d] / total;
        } else {
          for (int d = 0; d < 3; ++d) ct[d] = s[d] / (double)n;
        }
      }
      const double* cr = cref + 3 * r;
      double acc = 0.0, total = 0.0;
      for (int k = b; k < b + n; ++k) {
        const double3 t = moved(frm + 3 * toff[k], X, a.modify);
        const double dx = ((double)rref[3 * k] - cr[0]) - (t.x - ct[0]);
        const double dy = ((double)rref[3 * k + 1] - cr[1]) - (t.y - ct[1]);
        const double dz = ((double)rref[3 * k + 2] - cr[2]) - (t.z - ct[2]);
        total += tm[k];
        acc += tm[k] * (dx * dx + dy * dy + dz * dz);
      }
      double v;
      if (total < kSmall) v = -1.0;
      else if (acc < 0.0) v = 0.0;
      else v = sqrt(acc / total);
      a.out[f * a.nres + r0 + r] = v;
    }
  }
}

}  // namespace

size_t rmsd_perres_lds_bytes(int nres_piece, int nentries, int span, int nfb) {
  return sizeof(double) * ((size_t)nentries + 3 * (size_t)nres_piece) +
         sizeof(float) * 3 * (size_t)nentries + sizeof(int) * (size_t)nentries +
         sizeof(float) * 3 * (size_t)span * (size_t)nfb;
}

hipError_t rmsd_launch_perres(const PerresArgs& a, int npieces, size_t lds_bytes, hipStream_t s) {
  if (a.nframes == 0 || npieces == 0) return hipSuccess;
  if (lds_bytes > PERRES_LDS_MAX) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_perres),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  const int64_t batches = (a.nframes + a.nfb - 1) / a.nfb;
  // Enough workgroups to fill the chip a few times over; each walks batches.
  const int64_t want = (2048 + npieces - 1) / npieces;
  const unsigned gx = (unsigned)(batches < want ? batches : want);
  hipLaunchKernelGGL(k_perres, dim3(gx, npieces), dim3(PERRES_THREADS), lds_bytes, s, a);
  return hipGetLastError();
}
