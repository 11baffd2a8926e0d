// k_symm.hip -- symmetry-corrected RMSD ('symmrmsd').
//
//   k_symm_assign  SymmetricRmsdCalc::SymmRMSD_CenteredRef
//                  (src/SymmetricRmsdCalc.cpp:125-176): per (frame, group of
//                  symmetric atoms) the squared distances between the
//                  reference atoms and the target atoms as the initial fit
//                  leaves them (translated by Trans, rotated by U), and the
//                  reference's Hungarian heuristic on them
//                  (Hungarian::Optimize, src/Hungarian.cpp:21-337); the
//                  assignment goes to the frame's map.
//   k_symm_final_fit / k_symm_final_nofit
//                  the rest of SymmRMSD_CenteredRef (:178-192): the target
//                  remapped (coordinates and masses, Frame::
//                  SetCoordinatesByMap) and fitted again (FP64 moments, QCP,
//                  rotation) or compared without fit.
//   k_symm_amap    the per-frame map of every selected atom, when asked for.
//
// The initial fit is the frame-to-reference fit (k_fit.hip) run first.  One
// lane per (frame, group) for the assignment: groups are small (2-3 atoms in
// proteins) and the heuristic is sequential and branchy; the groups are
// ordered by size and lanes of one wave take consecutive groups of one frame
// (one size per wave but at size boundaries, loads within one frame).  Groups of up to
// SYMM_GROUP_REG atoms run a compile-time-size form with every array in
// registers (no LDS: occupancy is set by registers alone); larger groups keep
// their arrays in a slab of global scratch per (frame, group).  The final pass is one wavefront per frame: with a fit it
// corrects the initial fit's moments for the atoms whose pairing changed;
// without, it sums over every atom.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <stdint.h>
#include "rmsd_device.hpp"
#include "rmsd_kernels.hpp"

namespace {

constexpr int SYM_THREADS = 64;

// All three coordinates of selected atom k of frame f.  An AoS atom is read
// as one 12-byte load: three 4-byte gathers would each look up the same
// cache lines again.
typedef float xyz3 __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ void satom(const FrameSetDev& fs, int64_t f, int k, double& x,
                                      double& y, double& z) {
  const int a = fs.sel ? fs.sel[k] : k;
  const float* fr = fs.xyz + frame_of(fs, f) * fs.stride;
  if (fs.ld) {
    x = fr[a];
    y = fr[fs.ld + a];
    z = fr[2 * fs.ld + a];
  } else {
    const xyz3 p = *reinterpret_cast<const xyz3*>(fr + 3 * a);
    x = p.x;
    y = p.y;
    z = p.z;
  }
}

// Hungarian::Optimize: row and column reduction, then at most n * n rounds of
// greedy zero assignment, greedy line cover and update by the smallest
// uncovered element.  M is n x n row-major (row = target atom, column =
// reference atom); map[col] receives the assigned row or -1.  ws holds 5 n
// entries.  Lines are bit masks (n <= 32).  Element k of every array sits at
// [k * S] (S = 1: a private slab).
template <int S, typename IT>
struct Strided {
  IT* p;
  __device__ IT& operator[](int k) const { return p[k * S]; }
};

template <int S, typename IT>
__device__ void hungarian(Strided<S, double> M, int n, Strided<S, IT> map, IT* wsp) {
#pragma clang fp contract(off)
  const Strided<S, IT> a_cr{wsp};  // row -> column
  const Strided<S, IT> rz{wsp + n * S};   // zeros per row / column, non-zeros per row / column
  const Strided<S, IT> cz{wsp + 2 * n * S};
  const Strided<S, IT> rnz{wsp + 3 * n * S};
  const Strided<S, IT> cnz{wsp + 4 * n * S};
  for (int r = 0; r < n; ++r) {
    double mn = DBL_MAX;
    for (int c = 0; c < n; ++c) mn = M[r * n + c] < mn ? M[r * n + c] : mn;
    for (int c = 0; c < n; ++c) M[r * n + c] -= mn;
  }
  for (int c = 0; c < n; ++c) {
    double mn = DBL_MAX;
    for (int r = 0; r < n; ++r) mn = M[r * n + c] < mn ? M[r * n + c] : mn;
    for (int r = 0; r < n; ++r) M[r * n + c] -= mn;
  }
  for (int it = 0; it < n * n; ++it) {
    for (int k = 0; k < n; ++k) {
      map[k] = -1;
      a_cr[k] = -1;
    }
    int nassigned = 0;
    for (;;) {
      int min_row = -1, min_row_z = n + 1;
      for (int r = 0; r < n; ++r) {
        if (a_cr[r] != -1) continue;
        int z = 0;
        for (int c = 0; c < n; ++c) z += (map[c] == -1 && M[r * n + c] < RMSD_SMALL) ? 1 : 0;
        if (z != 0 && z < min_row_z) {
          min_row_z = z;
          min_row = r;
        }
      }
      int min_col = -1, min_col_z = n + 1;
      for (int c = 0; c < n; ++c) {
        if (map[c] != -1) continue;
        int z = 0;
        for (int r = 0; r < n; ++r) z += (a_cr[r] == -1 && M[r * n + c] < RMSD_SMALL) ? 1 : 0;
        if (z != 0 && z < min_col_z) {
          min_col_z = z;
          min_col = c;
        }
      }
      if (min_row == -1 && min_col == -1) break;
      bool done = false;
      if (min_col_z < min_row_z) {
        for (int r = 0; r < n && !done; ++r)
          if (M[r * n + min_col] < RMSD_SMALL && a_cr[r] == -1) {
            map[min_col] = (IT)r;
            a_cr[r] = (IT)min_col;
            done = true;
          }
      } else {
        for (int c = 0; c < n && !done; ++c)
          if (M[min_row * n + c] < RMSD_SMALL && map[c] == -1) {
            map[c] = (IT)min_row;
            a_cr[min_row] = (IT)c;
            done = true;
          }
      }
      if (!done) break;
      ++nassigned;
    }
    if (nassigned == n) break;
    // Cover the zeros: first by most zeros, then by zeros minus non-zeros;
    // rows win ties.
    uint32_t lr = 0, lc = 0;
    int total = 0;
    for (int k = 0; k < n; ++k) rz[k] = cz[k] = rnz[k] = cnz[k] = 0;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        if (M[r * n + c] < RMSD_SMALL) {
          ++rz[r];
          ++cz[c];
          ++total;
        } else {
          ++rnz[r];
          ++cnz[c];
        }
      }
    for (int step = 0; total > 0 && step < 2 * n; ++step) {
      int max_diff = -1, max_zero = -1, max_idx = -1;
      bool is_row = true;
      for (int r = 0; r < n; ++r) {
        if (lr >> r & 1u) continue;
        const int d = step > 0 ? rz[r] - rnz[r] : rz[r];
        if (max_zero == -1 || d > max_diff) {
          max_diff = d;
          max_zero = rz[r];
          max_idx = r;
        }
      }
      for (int c = 0; c < n; ++c) {
        if (lc >> c & 1u) continue;
        const int d = step > 0 ? cz[c] - cnz[c] : cz[c];
        if (d > max_diff) {
          max_diff = d;
          max_zero = cz[c];
          max_idx = c;
          is_row = false;
        }
      }
      if (max_idx < 0) break;
      if (is_row) {
        lr |= 1u << max_idx;
        for (int c = 0; c < n; ++c) {
          if (M[max_idx * n + c] < RMSD_SMALL)
            --cz[c];
          else
            --cnz[c];
        }
      } else {
        lc |= 1u << max_idx;
        for (int r = 0; r < n; ++r) {
          if (M[r * n + max_idx] < RMSD_SMALL)
            --rz[r];
          else
            --rnz[r];
        }
      }
      total -= max_zero;
    }
    double mu = DBL_MAX;
    for (int r = 0; r < n; ++r) {
      if (lr >> r & 1u) continue;
      for (int c = 0; c < n; ++c)
        if (!(lc >> c & 1u) && M[r * n + c] < mu) mu = M[r * n + c];
    }
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        const bool br = lr >> r & 1u, bc = lc >> c & 1u;
        if (br && bc)
          M[r * n + c] += mu;
        else if (!br && !bc)
          M[r * n + c] -= mu;
      }
  }
}

// The same heuristic for a compile-time size N, every array in registers:
// all loops unroll, and the few data-dependent indices (the chosen row or
// column) are resolved by selects over the N candidates.
template <int N>
__device__ __forceinline__ double pick_col(const double (&M)[N * N], int r, int c_dyn) {
  double v = 0.0;
#pragma unroll
  for (int c = 0; c < N; ++c) v = c == c_dyn ? M[r * N + c] : v;
  return v;
}

template <int N>
__device__ __forceinline__ double pick_row(const double (&M)[N * N], int r_dyn, int c) {
  double v = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) v = r == r_dyn ? M[r * N + c] : v;
  return v;
}

template <int N>
__device__ void hungarian_n(double (&M)[N * N], int (&map)[N]) {
#pragma clang fp contract(off)
  int a_cr[N], rz[N], cz[N], rnz[N], cnz[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double mn = DBL_MAX;
#pragma unroll
    for (int c = 0; c < N; ++c) mn = M[r * N + c] < mn ? M[r * N + c] : mn;
#pragma unroll
    for (int c = 0; c < N; ++c) M[r * N + c] -= mn;
  }
#pragma unroll
  for (int c = 0; c < N; ++c) {
    double mn = DBL_MAX;
#pragma unroll
    for (int r = 0; r < N; ++r) mn = M[r * N + c] < mn ? M[r * N + c] : mn;
#pragma unroll
    for (int r = 0; r < N; ++r) M[r * N + c] -= mn;
  }
  for (int it = 0; it < N * N; ++it) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      map[k] = -1;
      a_cr[k] = -1;
    }
    int nassigned = 0;
    for (int round = 0; round < N; ++round) {
      int min_row = -1, min_row_z = N + 1;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        int z = 0;
#pragma unroll
        for (int c = 0; c < N; ++c) z += (map[c] == -1 && M[r * N + c] < RMSD_SMALL) ? 1 : 0;
        if (a_cr[r] == -1 && z != 0 && z < min_row_z) {
          min_row_z = z;
          min_row = r;
        }
      }
      int min_col = -1, min_col_z = N + 1;
#pragma unroll
      for (int c = 0; c < N; ++c) {
        int z = 0;
#pragma unroll
        for (int r = 0; r < N; ++r) z += (a_cr[r] == -1 && M[r * N + c] < RMSD_SMALL) ? 1 : 0;
        if (map[c] == -1 && z != 0 && z < min_col_z) {
          min_col_z = z;
          min_col = c;
        }
      }
      if (min_row == -1 && min_col == -1) break;
      int ar = -1, ac = -1;  // the assignment made this round
      if (min_col_z < min_row_z) {
#pragma unroll
        for (int r = N - 1; r >= 0; --r)
          if (a_cr[r] == -1 && pick_col<N>(M, r, min_col) < RMSD_SMALL) ar = r;
        ac = min_col;
      } else {
#pragma unroll
        for (int c = N - 1; c >= 0; --c)
          if (map[c] == -1 && pick_row<N>(M, min_row, c) < RMSD_SMALL) ac = c;
        ar = min_row;
      }
      if (ar < 0 || ac < 0) break;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        if (k == ac) map[k] = ar;
        if (k == ar) a_cr[k] = ac;
      }
      ++nassigned;
    }
    if (nassigned == N) break;
    uint32_t lr = 0, lc = 0;
    int total = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) rz[k] = cz[k] = rnz[k] = cnz[k] = 0;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c < N; ++c) {
        const bool z = M[r * N + c] < RMSD_SMALL;
        rz[r] += z ? 1 : 0;
        cz[c] += z ? 1 : 0;
        rnz[r] += z ? 0 : 1;
        cnz[c] += z ? 0 : 1;
        total += z ? 1 : 0;
      }
    for (int step = 0; total > 0 && step < 2 * N; ++step) {
      int max_diff = -1, max_zero = -1, max_idx = -1;
      bool is_row = true;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        const int d = step > 0 ? rz[r] - rnz[r] : rz[r];
        if (!(lr >> r & 1u) && (max_zero == -1 || d > max_diff)) {
          max_diff = d;
          max_zero = rz[r];
          max_idx = r;
        }
      }
#pragma unroll
      for (int c = 0; c < N; ++c) {
        const int d = step > 0 ? cz[c] - cnz[c] : cz[c];
        if (!(lc >> c & 1u) && d > max_diff) {
          max_diff = d;
          max_zero = cz[c];
          max_idx = c;
          is_row = false;
        }
      }
      if (max_idx < 0) break;
      if (is_row) {
        lr |= 1u << max_idx;
#pragma unroll
        for (int c = 0; c < N; ++c) {
          const bool z = pick_row<N>(M, max_idx, c) < RMSD_SMALL;
          cz[c] -= z ? 1 : 0;
          cnz[c] -= z ? 0 : 1;
        }
      } else {
        lc |= 1u << max_idx;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const bool z = pick_col<N>(M, r, max_idx) < RMSD_SMALL;
          rz[r] -= z ? 1 : 0;
          rnz[r] -= z ? 0 : 1;
        }
      }
      total -= max_zero;
    }
    double mu = DBL_MAX;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c < N; ++c)
        if (!(lr >> r & 1u) && !(lc >> c & 1u) && M[r * N + c] < mu) mu = M[r * N + c];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c < N; ++c) {
        const bool br = lr >> r & 1u, bc = lc >> c & 1u;
        if (br && bc)
          M[r * N + c] += mu;
        else if (!br && !bc)
          M[r * N + c] -= mu;
      }
  }
}

// Cost matrix of one group: target atoms (rows) as the initial fit leaves
// them (translated by Trans, Frame::Rotate by U) against the reference atoms
// (columns), DIST2_NoImage.
template <int S>
__device__ __forceinline__ void group_costs(const SymmArgs& a, int64_t f, const int* ga, int n,
                                            const double (&U)[9], const double (&t)[3],
                                            Strided<S, double> M) {
#pragma clang fp contract(off)
  for (int ta = 0; ta < n; ++ta) {
    double x, y, z;
    satom(a.fs, f, ga[ta], x, y, z);
    if (a.U0) {
      x += t[0];
      y += t[1];
      z += t[2];
      const double rx = (x * U[0]) + (y * U[1]) + (z * U[2]);
      const double ry = (x * U[3]) + (y * U[4]) + (z * U[5]);
      const double rz = (x * U[6]) + (y * U[7]) + (z * U[8]);
      x = rx;
      y = ry;
      z = rz;
    }
    for (int ra = 0; ra < n; ++ra) {
      const int k = ga[ra];
      const double dx = a.ref[k] - x, dy = a.ref[a.ldr + k] - y, dz = a.ref[2 * a.ldr + k] - z;
      M[ta * n + ra] = (dx * dx + dy * dy + dz * dz);
    }
  }
}

template <int N>
__device__ __forceinline__ void assign_reg(const SymmArgs& a, int64_t f, const int* ga,
                                           const double (&U)[9], const double (&t)[3], int* out) {
  double M[N * N];
  int map[N];
  group_costs(a, f, ga, N, U, t, Strided<1, double>{M});
  hungarian_n<N>(M, map);
#pragma unroll
  for (int c = 0; c < N; ++c) out[c] = ga[map[c] >= 0 ? map[c] : c];
}

// Lanes: consecutive groups of one frame (pair = frame * G + group); the
// groups are ordered by size, so a wave's groups share one size (but at the
// few size boundaries), its loads stay within one frame and the frame's
// initial fit is one broadcast load.
__global__ __launch_bounds__(SYM_THREADS) void k_symm_assign(SymmArgs a) {
  const int64_t pair = (int64_t)blockIdx.x * SYM_THREADS + threadIdx.x;
  if (pair >= (a.f1 - a.f0) * a.ngroups) return;
  const int64_t fl = pair / a.ngroups;  // frame within this launch
  const int64_t f = a.f0 + fl;
  const int g = (int)(pair - fl * a.ngroups);
  const int g0 = a.goff[g];
  const int n = a.goff[g + 1] - g0;
  const int* ga = a.gatoms + g0;
  int* out = a.gmap + f * a.ngatoms + g0;
  if (n == 1) {
    out[0] = ga[0];
    return;
  }
  double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t[3] = {0, 0, 0};
  if (a.U0) {
    for (int k = 0; k < 9; ++k) U[k] = a.U0[9 * f + k];
    for (int k = 0; k < 3; ++k) t[k] = a.t0[6 * f + k];
  }
  if (n == 2) {
    assign_reg<2>(a, f, ga, U, t, out);
  } else if (n == 3) {
    assign_reg<3>(a, f, ga, U, t, out);
  } else if (n == 4) {
    assign_reg<4>(a, f, ga, U, t, out);
  } else {
    double* slab = a.big + fl * a.bigstride + a.bigoff[g];
    const Strided<1, double> M{slab};
    signed char* ws = reinterpret_cast<signed char*>(slab + n * n);
    const Strided<1, signed char> map{ws + 5 * n};
    group_costs(a, f, ga, n, U, t, M);
    hungarian(M, n, map, ws);
    for (int c = 0; c < n; ++c) out[c] = ga[map[c] >= 0 ? map[c] : c];
  }
}

// The remapped fit.  A map permutes atoms inside groups, so the remapped
// frame's moments are the initial fit's (k_fit.hip: 13 per frame against the
// same reference and shift) plus, for each atom i whose target t(i) != i, the
// difference of its terms: w_t (x_t - s) - w_i (x_i - s) in the sums of
// d, d y_i^T and |d|^2, (w_t - w_i) y_i and (w_t - w_i) |y_i|^2 in the
// reference sums and w_t - w_i in the total weight.  Every term is corrected,
// so the result is the remapped frame's for any map.  Only the frame's group
// entries and the atoms whose pairing changed are read.
template <bool WEIGHTED>
__global__ __launch_bounds__(SYM_THREADS) void k_symm_final_fit(SymmArgs a) {
#pragma clang fp contract(off)
  const int64_t f = blockIdx.x;
  const int64_t F = a.fs.nframes;
  const int lane = threadIdx.x;
  const int* gm = a.gmap + f * a.ngatoms;
  const double* sh = a.moments + 13 * F + 3 * f;
  const double s0 = sh[0], s1 = sh[1], s2 = sh[2];
  double v[18];
#pragma unroll
  for (int q = 0; q < 18; ++q) v[q] = 0.0;
  for (int p = lane; p < a.ngatoms; p += SYM_THREADS) {
    const int i = a.gatoms[p];
    const int t = gm[p];
    if (t == i) continue;
    const double wt = WEIGHTED ? a.mass[t] : 1.0;
    const double wi = WEIGHTED ? a.mass[i] : 1.0;
    const double y0 = a.ref[i], y1 = a.ref[a.ldr + i], y2 = a.ref[2 * a.ldr + i];
    double xt0, xt1, xt2, xi0, xi1, xi2;
    satom(a.fs, f, t, xt0, xt1, xt2);
    satom(a.fs, f, i, xi0, xi1, xi2);
    const double t0 = xt0 - s0, t1 = xt1 - s1, t2 = xt2 - s2;
    const double i0 = xi0 - s0, i1 = xi1 - s1, i2 = xi2 - s2;
    const double e0 = wt * t0 - wi * i0, e1 = wt * t1 - wi * i1, e2 = wt * t2 - wi * i2;
    v[0] += e0;
    v[1] += e1;
    v[2] += e2;
    v[3] += e0 * y0; v[4] += e0 * y1; v[5] += e0 * y2;
    v[6] += e1 * y0; v[7] += e1 * y1; v[8] += e1 * y2;
    v[9] += e2 * y0; v[10] += e2 * y1; v[11] += e2 * y2;
    v[12] += (wt * t0 * t0 + wt * t1 * t1 + wt * t2 * t2) -
             (wi * i0 * i0 + wi * i1 * i1 + wi * i2 * i2);
    if (WEIGHTED) {
      const double dw = wt - wi;
      v[13] += dw * y0;
      v[14] += dw * y1;
      v[15] += dw * y2;
      v[16] += dw * (y0 * y0 + y1 * y1 + y2 * y2);
      v[17] += dw;
    }
  }
#pragma unroll
  for (int q = 0; q < 18; ++q) v[q] = rmsd_dev::wave_sum(v[q]);
  if (lane != 0) return;
  const double* mo = a.moments + 13 * f;
  const double* K = a.consts;
  const double W = K[4] + v[17];
  double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double tr[3] = {0, 0, 0};
  double r;
  if (W < RMSD_SMALL) {
    r = -1.0;
  } else {
    double m[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) m[q] = mo[q] + v[q];
    const double K0 = K[0] + v[13], K1 = K[1] + v[14], K2 = K[2] + v[15], K3 = K[3] + v[16];
    const double c0 = m[0] / W, c1 = m[1] / W, c2 = m[2] / W;
    double R[9];
    R[0] = m[3] - c0 * K0; R[1] = m[4] - c0 * K1; R[2] = m[5] - c0 * K2;
    R[3] = m[6] - c1 * K0; R[4] = m[7] - c1 * K1; R[5] = m[8] - c1 * K2;
    R[6] = m[9] - c2 * K0; R[7] = m[10] - c2 * K1; R[8] = m[11] - c2 * K2;
    const double sx = m[12] - W * (c0 * c0 + c1 * c1 + c2 * c2);
    const double E0 = 0.5 * (sx + K3);
    const rmsd_dev::QcpCoef co = rmsd_dev::qcp_coefficients(R);
    const double lam = rmsd_dev::qcp_lambda(co, E0);
    r = rmsd_dev::rmsd_from_lambda(E0, lam, W, 1);
    rmsd_dev::qcp_rotation(R, lam, U);
    tr[0] = -(c0 + s0);
    tr[1] = -(c1 + s1);
    tr[2] = -(c2 + s2);
  }
  a.rmsd[f] = r;
  if (a.U)
    for (int k = 0; k < 9; ++k) a.U[9 * f + k] = U[k];
  if (a.tvec)
    for (int k = 0; k < 3; ++k) {
      a.tvec[6 * f + k] = tr[k];
      a.tvec[6 * f + 3 + k] = a.ref_trans[k];
    }
}

// Without fit: the remapped frame against the reference as stored,
// RMSD_NoFit over every atom with the remapped masses.
template <bool WEIGHTED>
__global__ __launch_bounds__(SYM_THREADS) void k_symm_final_nofit(SymmArgs a) {
#pragma clang fp contract(off)
  const int64_t f = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = a.fs.nsel;
  const int* gm = a.gmap + f * a.ngatoms;
  double acc = 0.0, W = 0.0;
#pragma unroll 4
  for (int i = lane; i < n; i += SYM_THREADS) {
    const int p = a.pos[i];
    const int t = p >= 0 ? gm[p] : i;
    const double w = WEIGHTED ? a.mass[t] : 1.0;
    double xt0, xt1, xt2;
    satom(a.fs, f, t, xt0, xt1, xt2);
    const double dx = a.ref[i] - xt0;
    const double dy = a.ref[a.ldr + i] - xt1;
    const double dz = a.ref[2 * a.ldr + i] - xt2;
    W += w;
    acc += (w * (dx * dx + dy * dy + dz * dz));
  }
  acc = rmsd_dev::wave_sum(acc);
  W = rmsd_dev::wave_sum(W);
  if (lane == 0) a.rmsd[f] = W < RMSD_SMALL ? -1.0 : (acc < 0.0 ? 0.0 : sqrt(acc / W));
}

// The per-frame map of every selected atom (SymmetricRmsdCalc::AMap).
// Grid: x over frames (AMAP_FRAMES per block), y over atoms (256 per block).
constexpr int AMAP_FRAMES = 8;
__global__ __launch_bounds__(256) void k_symm_amap(SymmArgs a) {
  const int n = a.fs.nsel;
  const int i = blockIdx.y * 256 + threadIdx.x;
  if (i >= n) return;
  const int p = a.pos[i];
  const int64_t f0 = (int64_t)blockIdx.x * AMAP_FRAMES;
#pragma unroll
  for (int k = 0; k < AMAP_FRAMES; ++k) {
    const int64_t f = f0 + k;
    if (f >= a.fs.nframes) break;
    a.amap[f * n + i] = p >= 0 ? a.gmap[f * a.ngatoms + p] : i;
  }
}

}  // namespace

hipError_t rmsd_launch_symm(const SymmArgs& a, hipStream_t s) {
  const int64_t F = a.fs.nframes;
  if (F == 0) return hipSuccess;
  if (a.ngroups > 0) {
    // Frames in chunks whose big-group scratch fits SYMM_BIG_MAX (one chunk
    // when no group needs it); launches on one stream reuse it in order.
    int64_t chunk = F;
    if (a.bigstride > 0)
      chunk = std::max<int64_t>(1, (int64_t)(SYMM_BIG_MAX / (sizeof(double) * a.bigstride)));
    for (int64_t f0 = 0; f0 < F; f0 += chunk) {
      SymmArgs b = a;
      b.f0 = f0;
      b.f1 = std::min<int64_t>(F, f0 + chunk);
      const int64_t pairs = (b.f1 - b.f0) * a.ngroups;
      const int64_t blocks = (pairs + SYM_THREADS - 1) / SYM_THREADS;
      if (blocks > 0x7fffffff) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_symm_assign, dim3((unsigned)blocks), dim3(SYM_THREADS), 0, s, b);
    }
  }
  const dim3 grid((unsigned)F), block(SYM_THREADS);
  if (a.fit && a.mass)
    hipLaunchKernelGGL((k_symm_final_fit<true>), grid, block, 0, s, a);
  else if (a.fit)
    hipLaunchKernelGGL((k_symm_final_fit<false>), grid, block, 0, s, a);
  else if (a.mass)
    hipLaunchKernelGGL((k_symm_final_nofit<true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_symm_final_nofit<false>), grid, block, 0, s, a);
  if (a.amap) {
    const int64_t gx = (F + AMAP_FRAMES - 1) / AMAP_FRAMES;
    const int gy = (a.fs.nsel + 255) / 256;
    if (gy > 65535) return hipErrorInvalidValue;
    const dim3 g((unsigned)gx, (unsigned)gy);
    hipLaunchKernelGGL(k_symm_amap, g, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}
