// k_matrix.hip -- reductions over a finished packed-triangle RMSD matrix.
//
//   k_autocorr_part / k_autocorr_sum   Analysis_Rms2d::CalcAutoCorr
//       (src/Analysis_Rms2d.cpp:301-319): ct[lag] = mean_j exp(-R(j, j+lag)),
//       ct[0] = 1, in FP64 from the FP32 matrix.
//
// Row j of the triangle holds R(j, j+1 .. n-1) contiguously, so for a fixed
// row the lags are consecutive addresses: a thread owns one lag, a block 256
// consecutive lags, and a block walks a chunk of rows (coalesced 1 KiB reads
// per wave).  Per-chunk partial sums go to scratch and a second kernel adds the
// chunks in a fixed order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rmsd_kernels.hpp"

namespace {

constexpr int AC_THREADS = 256;

__global__ __launch_bounds__(AC_THREADS) void k_autocorr_part(const float* tri, int n, int rows,
                                                              double* part) {
  const int lag = blockIdx.x * AC_THREADS + threadIdx.x + 1;  // 1 .. n-1
  const int64_t j0 = (int64_t)blockIdx.y * rows;
  const int64_t j1 = j0 + rows < n ? j0 + rows : n;
  double s = 0.0;
  if (lag < n) {
    for (int64_t j = j0; j < j1 && j + lag < n; ++j) {
      // index of (j, j+1) in the packed triangle, then + (lag - 1)
      const int64_t base = (int64_t)n * j - (j * (j + 1)) / 2;
      s += exp(-(double)tri[base + lag - 1]);
    }
  }
  if (lag < n) part[(int64_t)blockIdx.y * n + lag] = s;
}

__global__ __launch_bounds__(AC_THREADS) void k_autocorr_sum(const double* part, int n, int nchunks,
                                                             double* ct) {
  const int lag = blockIdx.x * AC_THREADS + threadIdx.x;
  if (lag >= n) return;
  if (lag == 0) {
    ct[0] = 1.0;
    return;
  }
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += part[(int64_t)c * n + lag];
  ct[lag] = s / (double)(n - lag);
}

}  // namespace

int rmsd_autocorr_chunks(int n, int* rows) {
  *rows = 512;
  return (n + *rows - 1) / *rows;
}

hipError_t rmsd_launch_autocorr(const float* tri, int n, double* part, double* ct, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int rows = 0;
  const int nchunks = rmsd_autocorr_chunks(n, &rows);
  if (n > 1) {
    dim3 grid((n - 1 + AC_THREADS - 1) / AC_THREADS, nchunks);
    hipLaunchKernelGGL(k_autocorr_part, grid, dim3(AC_THREADS), 0, s, tri, n, rows, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_autocorr_sum, dim3((n + AC_THREADS - 1) / AC_THREADS), dim3(AC_THREADS), 0,
                     s, part, n, nchunks, ct);
  return hipGetLastError();
}
