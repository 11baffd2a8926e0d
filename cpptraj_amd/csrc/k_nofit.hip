// k_nofit.hip -- all-pairs RMSD without fitting (rms2d 'nofit', cluster 'nofit').
//
// Frame::RMSD_NoFit (src/Frame.cpp:1279-1307) for every pair of the
// Analysis_Rms2d::Calculate_2D loop nest (src/Analysis_Rms2d.cpp:247-287):
// sqrt(sum_k w_k |r_k - t_k|^2 / sum_k w_k).  The per-pair sum runs atom by
// atom in the reference's operation order with FP64 arithmetic and no
// contraction, so results match the reference bit for bit.
//
// A 256-thread workgroup owns a 64 x 64 block of frame pairs (4 x 4 pairs per
// thread) and stages 16 atoms of its 64 row and 64 column frames in LDS per
// step (48 KiB).  This path is FP64 vector work; it is not the headline path.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

constexpr int NF_TILE = 64;
constexpr int NF_KS = 16;

__global__ __launch_bounds__(256) void k_pair_nofit(FrameSetDev A, FrameSetDev B,
                                                    const double* mass, const double* total_mass,
                                                    const int2* tiles, int ntiles, int per_xcd,
                                                    int nxcd, int tri, float* out) {
#pragma clang fp contract(off)
  __shared__ double sA[NF_KS][3][NF_TILE];
  __shared__ double sB[NF_KS][3][NF_TILE];
  __shared__ double sW[NF_KS];
  const int w = blockIdx.x;
  const int t = (w % nxcd) * per_xcd + (w / nxcd);
  if (t >= ntiles) return;
  const int2 tile = tiles[t];
  const int64_t i0 = (int64_t)tile.x * NF_TILE, j0 = (int64_t)tile.y * NF_TILE;
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int n = A.nsel;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;

  for (int k0 = 0; k0 < n; k0 += NF_KS) {
    __syncthreads();
    for (int e = threadIdx.x; e < NF_KS * NF_TILE; e += blockDim.x) {
      const int kk = e / NF_TILE, fr = e % NF_TILE;
      const int k = k0 + kk;
      const int64_t fi = i0 + fr, fj = j0 + fr;
      const bool kin = k < n;
      float pa[3] = {0.0f, 0.0f, 0.0f}, pb[3] = {0.0f, 0.0f, 0.0f};
      if (kin && fi < A.nframes) load_atom(A, fi, k, pa);
      if (kin && fj < B.nframes) load_atom(B, fj, k, pb);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sA[kk][c][fr] = (double)pa[c];
        sB[kk][c][fr] = (double)pb[c];
      }
    }
    if (threadIdx.x < NF_KS) {
      const int k = k0 + threadIdx.x;
      sW[threadIdx.x] = k < n ? (mass ? mass[k] : 1.0) : 0.0;
    }
    __syncthreads();
    const int kend = (n - k0) < NF_KS ? (n - k0) : NF_KS;
    for (int kk = 0; kk < kend; ++kk) {
      const double wk = sW[kk];
      double ra[4][3], cb[4][3];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) ra[a][c] = sA[kk][c][ty + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) cb[b][c] = sB[kk][c][tx + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          // Rows are the reference frames in the triangle layout and the
          // target frames in the full layout; (r - t)^2 is symmetric.
          const double dx = ra[a][0] - cb[b][0];
          const double dy = ra[a][1] - cb[b][1];
          const double dz = ra[a][2] - cb[b][2];
          if (mass) acc[a][b] += (wk * (dx * dx + dy * dy + dz * dz));
          else acc[a][b] += (dx * dx + dy * dy + dz * dz);
        }
    }
  }
  const double M = total_mass[0];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
      if (i >= A.nframes || j >= B.nframes || (tri && i >= j)) continue;
      double r;
      if (M < RMSD_SMALL) r = -1.0;
      else if (acc[a][b] < 0) r = 0.0;
      else r = sqrt(acc[a][b] / M);
      const int64_t idx = tri ? rmsd_dev::tri_index(A.nframes, i, j) : i * B.nframes + j;
      out[idx] = (float)r;
    }
}

}  // namespace

hipError_t rmsd_launch_pair_nofit_tiles(const FrameSetDev& A, const FrameSetDev& B,
                                        const double* mass, const double* total_mass,
                                        const int2* tiles, int ntiles, int per_xcd, int nxcd,
                                        int tri, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_nofit, dim3(per_xcd * nxcd), dim3(256), 0, s, A, B, mass, total_mass,
                     tiles, ntiles, per_xcd, nxcd, tri, out);
  return hipGetLastError();
}
