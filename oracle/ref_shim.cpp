// ref_shim.cpp -- C entry points onto the compiled reference sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own hot-path translation units (Frame.cpp, Matrix_3x3.cpp,
// qcprot.cpp, QuaternionRMSD.cpp and the few files they pull in) straight from
// the reference tree into oracle/_ref/libcpptraj_ref.so.  Nothing from the
// reference is copied into this repository; the library is built here and is
// used only (a) to pin the C oracle in rmsd_oracle.c and (b) as the CPU
// baseline timed beside the GPU.  Each wrapper names the reference function it
// calls.
#include <vector>
#include <cstring>
#include <cmath>
#include <cstdint>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "Frame.h"
#include "Matrix_3x3.h"
#include "Vec3.h"
#include "QuaternionRMSD.h"
#include "qcprot.h"
#include "Hungarian.h"
#include "DistRoutines.h"
#include "Box.h"

namespace {

Frame make_frame(const double* xyz, const double* mass, int n) {
  std::vector<double> X(xyz, xyz + 3 * (size_t)n);
  std::vector<double> M((size_t)n, 1.0);
  if (mass) M.assign(mass, mass + n);
  Frame f;
  f.SetupFrameXM(X, M);
  return f;
}

void frame_from_floats(Frame& f, const float* fr, const int* sel, int nsel) {
  double* X = f.xAddress();
  for (int j = 0; j < nsel; ++j) {
    int a = sel ? sel[j] : j;
    X[3 * j] = (double)fr[3 * a];
    X[3 * j + 1] = (double)fr[3 * a + 1];
    X[3 * j + 2] = (double)fr[3 * a + 2];
  }
}

int64_t tri_index(int64_t nx, int64_t x, int64_t y) {
  int64_t i = x < y ? x : y, j = x < y ? y : x;
  return ((nx * i) - (((i + 1) * i) / 2)) + j - (i + 1);
}

}  // namespace

extern "C" {

// Frame::RMSD_CenteredRef(Frame const&, Matrix_3x3&, Vec3&, bool)
double cref_rmsd_centered_ref(double* tgt, const double* ref, const double* tmass,
                              const double* rmass, int n, int use_mass,
                              double* U, double* trans) {
  Frame T = make_frame(tgt, tmass, n);
  Frame R = make_frame(ref, rmass, n);
  Matrix_3x3 u;
  Vec3 tr;
  double r = T.RMSD_CenteredRef(R, u, tr, use_mass != 0);
  for (int k = 0; k < 9; ++k) U[k] = u[k];
  for (int k = 0; k < 3; ++k) trans[k] = tr[k];
  std::memcpy(tgt, T.xAddress(), sizeof(double) * 3 * (size_t)n);
  return r;
}

// Frame::RMSD(Frame&, Matrix_3x3&, Vec3&, Vec3&, bool)
double cref_rmsd(double* tgt, double* ref, const double* tmass, const double* rmass,
                 int n, int use_mass, double* U, double* trans, double* ref_trans) {
  Frame T = make_frame(tgt, tmass, n);
  Frame R = make_frame(ref, rmass, n);
  Matrix_3x3 u;
  Vec3 tr, rtr;
  double r = T.RMSD(R, u, tr, rtr, use_mass != 0);
  for (int k = 0; k < 9; ++k) U[k] = u[k];
  for (int k = 0; k < 3; ++k) { trans[k] = tr[k]; ref_trans[k] = rtr[k]; }
  std::memcpy(tgt, T.xAddress(), sizeof(double) * 3 * (size_t)n);
  std::memcpy(ref, R.xAddress(), sizeof(double) * 3 * (size_t)n);
  return r;
}

// Frame::RMSD_NoFit(Frame const&, bool)
double cref_rmsd_nofit(const double* tgt, const double* ref, const double* tmass,
                       int n, int use_mass) {
  Frame T = make_frame(tgt, tmass, n);
  Frame R = make_frame(ref, 0, n);
  return T.RMSD_NoFit(R, use_mass != 0);
}

// Frame::CenterOnOrigin(bool)
void cref_center_on_origin(double* xyz, const double* mass, int n, int use_mass,
                           double* center) {
  Frame F = make_frame(xyz, mass, n);
  Vec3 c = F.CenterOnOrigin(use_mass != 0);
  for (int k = 0; k < 3; ++k) center[k] = c[k];
  std::memcpy(xyz, F.xAddress(), sizeof(double) * 3 * (size_t)n);
}

// QuaternionRMSD_CenteredRef(Frame const&, Frame&, Matrix_3x3&, Vec3&, bool)
double cref_qrmsd_centered_ref(const double* ref, double* tgt, const double* tmass,
                               int n, int use_mass, double* rot, double* trans) {
  Frame R = make_frame(ref, 0, n);
  Frame T = make_frame(tgt, tmass, n);
  double r;
  Vec3 tr;
  if (rot) {
    Matrix_3x3 u;
    r = QuaternionRMSD_CenteredRef(R, T, u, tr, use_mass != 0);
    for (int k = 0; k < 9; ++k) rot[k] = u[k];
  } else {
    r = QuaternionRMSD_CenteredRef(R, T, use_mass != 0);
  }
  for (int k = 0; k < 3; ++k) trans[k] = tr[k];
  std::memcpy(tgt, T.xAddress(), sizeof(double) * 3 * (size_t)n);
  return r;
}

// Cpptraj::QCPRot::FastCalcRMSDAndRotation
int cref_qcp(double* rot, const double* A, double* rmsd, double E0, double len) {
  double a[9];
  std::memcpy(a, A, sizeof(a));
  return Cpptraj::QCPRot::FastCalcRMSDAndRotation(rot, a, rmsd, E0, len);
}

// Matrix_3x3::Diagonalize_Sort(Vec3&)
int cref_diagonalize_sort(const double* in, double* E, double* ev) {
  Matrix_3x3 m(in);
  Vec3 e;
  int err = m.Diagonalize_Sort(e);
  for (int k = 0; k < 9; ++k) E[k] = m[k];
  for (int k = 0; k < 3; ++k) ev[k] = e[k];
  return err;
}

// The loop nest of Analysis_Rms2d::Calculate_2D (src/Analysis_Rms2d.cpp:
// 247-287) driving the reference's Frame routines: OpenMP over reference rows
// with a dynamic schedule, reference centred once per row, target gathered per
// pair, (float) store in the packed triangle or full layout.
// mode: 0 fit, 1 nofit, 2 quaternion.
int cref_rms2d(const float* tgt_frames, int64_t tgt_stride, int ntgt,
               const float* ref_frames, int64_t ref_stride, int nref,
               const int* tsel, const int* rsel, int nsel,
               const double* tmass, const double* rmass,
               int mode, int use_mass, int full, float* out, int nthreads,
               int row_begin, int row_end) {
  if (ref_frames == 0) { ref_frames = tgt_frames; ref_stride = tgt_stride; nref = ntgt; }
  if (row_end < 0 || row_end > nref) row_end = nref;
  std::vector<double> zero(3 * (size_t)nsel, 0.0);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    Frame SelRef = make_frame(zero.data(), rmass, nsel);
    Frame SelTgt = make_frame(zero.data(), tmass, nsel);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int r = row_begin; r < row_end; ++r) {
      frame_from_floats(SelRef, ref_frames + (int64_t)r * ref_stride, rsel, nsel);
      if (mode == 0 || mode == 2) SelRef.CenterOnOrigin(use_mass != 0);
      int t0 = full ? 0 : r + 1;
      for (int t = t0; t < ntgt; ++t) {
        frame_from_floats(SelTgt, tgt_frames + (int64_t)t * tgt_stride, tsel, nsel);
        float v;
        if (mode == 0) v = (float)SelTgt.RMSD_CenteredRef(SelRef, use_mass != 0);
        else if (mode == 1) v = (float)SelTgt.RMSD_NoFit(SelRef, use_mass != 0);
        else v = (float)QuaternionRMSD_CenteredRef(SelRef, SelTgt, use_mass != 0);
        int64_t idx = full ? (int64_t)t * nref + r : tri_index(ntgt, r, t);
        if (out) out[idx] = v;
      }
    }
  }
  return 0;
}

// Listed reference rows of the same loop nest (triangle case), writing each
// row's pairs (r, t > r) consecutively; out may be NULL (timing only).
int cref_rms2d_rows(const float* frames, int64_t stride, int nframes, const int* sel, int nsel,
                    const double* mass, int mode, int use_mass, const int* rows, int nrows,
                    float* out, int nthreads) {
  std::vector<int64_t> off((size_t)nrows + 1, 0);
  for (int q = 0; q < nrows; ++q) off[q + 1] = off[q] + (nframes - 1 - rows[q]);
  std::vector<double> zero(3 * (size_t)nsel, 0.0);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    Frame SelRef = make_frame(zero.data(), mass, nsel);
    Frame SelTgt = make_frame(zero.data(), mass, nsel);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int q = 0; q < nrows; ++q) {
      const int r = rows[q];
      frame_from_floats(SelRef, frames + (int64_t)r * stride, sel, nsel);
      if (mode == 0 || mode == 2) SelRef.CenterOnOrigin(use_mass != 0);
      for (int t = r + 1; t < nframes; ++t) {
        frame_from_floats(SelTgt, frames + (int64_t)t * stride, sel, nsel);
        float v;
        if (mode == 0) v = (float)SelTgt.RMSD_CenteredRef(SelRef, use_mass != 0);
        else if (mode == 1) v = (float)SelTgt.RMSD_NoFit(SelRef, use_mass != 0);
        else v = (float)QuaternionRMSD_CenteredRef(SelRef, SelTgt, use_mass != 0);
        if (out) out[off[q] + (t - r - 1)] = v;
      }
    }
  }
  return 0;
}

// Per-frame body of Action_Rmsd::DoAction (src/Action_Rmsd.cpp:361-417) for a
// fixed, already centred reference: RMSD_CenteredRef (or RMSD_NoFit), then the
// coordinate update of the chosen mode (1 = Trans_Rot_Trans, 2 = translate).
int cref_fit_frames(const float* frames, int64_t stride, int nframes, int natom,
                    const int* sel, int nsel, const double* mass,
                    const double* ref_sel, const double* ref_trans,
                    int fit, int use_mass, int modify, double* rmsd, double* U,
                    double* tvec, float* frames_out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    Frame Ref = make_frame(ref_sel, 0, nsel);
    std::vector<double> zero(3 * (size_t)nsel, 0.0);
    Frame Tgt = make_frame(zero.data(), mass, nsel);
    std::vector<double> all(3 * (size_t)natom, 0.0);
    Frame Full = make_frame(all.data(), 0, natom);
    Vec3 refT(ref_trans[0], ref_trans[1], ref_trans[2]);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int f = 0; f < nframes; ++f) {
      const float* fr = frames + (int64_t)f * stride;
      frame_from_floats(Tgt, fr, sel, nsel);
      Matrix_3x3 u(1.0);
      u.Zero(); u[0] = u[4] = u[8] = 1.0;
      Vec3 tr(0.0, 0.0, 0.0);
      double v = fit ? Tgt.RMSD_CenteredRef(Ref, u, tr, use_mass != 0)
                     : Tgt.RMSD_NoFit(Ref, use_mass != 0);
      rmsd[f] = v;
      if (fit && U) for (int k = 0; k < 9; ++k) U[9 * (int64_t)f + k] = u[k];
      if (fit && tvec) {
        for (int k = 0; k < 3; ++k) {
          tvec[6 * (int64_t)f + k] = tr[k];
          tvec[6 * (int64_t)f + 3 + k] = refT[k];
        }
      }
      if (frames_out) {
        frame_from_floats(Full, fr, 0, natom);
        if (fit && modify == 1) Full.Trans_Rot_Trans(tr, u, refT);
        else if (fit && modify == 2) Full.Translate(tr + refT);
        const double* X = Full.xAddress();
        float* fo = frames_out + (int64_t)f * stride;
        for (int k = 0; k < 3 * natom; ++k) fo[k] = (float)X[k];
      }
    }
  }
  return 0;
}

// 'reftraj' mode of Action_Rmsd::DoAction: for every pair the reference
// frame's selected atoms centred with their masses (ReferenceAction::
// SelectRefAtoms, src/ReferenceAction.cpp:155-169), then
// Frame::RMSD_CenteredRef of the target frame (nomod).
int cref_fit_pairs(const float* tgt, const float* ref, int64_t stride, int nframes,
                   const int* sel, int nsel, const double* mass, int use_mass, double* rmsd,
                   int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    std::vector<double> zr(3 * (size_t)nsel, 0.0), zt(3 * (size_t)nsel, 0.0);
    Frame Ref = make_frame(zr.data(), mass, nsel);
    Frame Tgt = make_frame(zt.data(), mass, nsel);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int f = 0; f < nframes; ++f) {
      frame_from_floats(Ref, ref + (int64_t)f * stride, sel, nsel);
      Ref.CenterOnOrigin(use_mass != 0);
      frame_from_floats(Tgt, tgt + (int64_t)f * stride, sel, nsel);
      Matrix_3x3 u(1.0);
      Vec3 tr(0.0, 0.0, 0.0);
      rmsd[f] = Tgt.RMSD_CenteredRef(Ref, u, tr, use_mass != 0);
    }
  }
  return 0;
}

// Per-residue loop of Action_Rmsd::DoAction (src/Action_Rmsd.cpp:398-413):
// the frame after the fit's update (Frame::Trans_Rot_Trans / Translate in
// double), each residue copied out with its own masses, optionally centred
// (Frame::CenterOnOrigin on both), then Frame::RMSD_NoFit.  out[f*nres + r].
int cref_perres(const float* frames, int64_t stride, int nframes, int natom, const float* ref,
                int ref_natom, int nres, const int* res_off, const int* tgt_atoms,
                const int* ref_atoms, const double* tmass, const double* rmass, int use_mass,
                int center, const double* U, const double* tvec, int modify, double* out) {
  std::vector<double> all(3 * (size_t)natom, 0.0);
  Frame Full = make_frame(all.data(), 0, natom);
  std::vector<double> rall(3 * (size_t)ref_natom, 0.0);
  Frame RefFull = make_frame(rall.data(), 0, ref_natom);
  frame_from_floats(RefFull, ref, 0, ref_natom);
  for (int f = 0; f < nframes; ++f) {
    frame_from_floats(Full, frames + (int64_t)f * stride, 0, natom);
    if (modify && tvec) {
      const double* tv = tvec + 6 * (int64_t)f;
      Vec3 t1(tv[0], tv[1], tv[2]), t2(tv[3], tv[4], tv[5]);
      if (modify == 1) {
        Matrix_3x3 u(U + 9 * (int64_t)f);
        Full.Trans_Rot_Trans(t1, u, t2);
      } else {
        t1 += t2;
        Full.Translate(t1);
      }
    }
    for (int r = 0; r < nres; ++r) {
      const int b = res_off[r], n = res_off[r + 1] - b;
      std::vector<double> tx(3 * (size_t)n), rx(3 * (size_t)n);
      for (int k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) {
          tx[3 * k + c] = Full.xAddress()[3 * tgt_atoms[b + k] + c];
          rx[3 * k + c] = RefFull.xAddress()[3 * ref_atoms[b + k] + c];
        }
      Frame T = make_frame(tx.data(), tmass ? tmass + b : 0, n);
      Frame R = make_frame(rx.data(), rmass ? rmass + b : 0, n);
      if (center) {
        T.CenterOnOrigin(use_mass != 0);
        R.CenterOnOrigin(use_mass != 0);
      }
      out[(int64_t)f * nres + r] = T.RMSD_NoFit(R, use_mass != 0);
    }
  }
  return 0;
}

// Window loop of Analysis_RmsAvgCorr::Analyze (src/Analysis_RmsAvgCorr.cpp:
// 208-257) on the reference's Frame: operator+= / operator-= for the running
// sum, Divide(Frame, double), CenterOnOrigin for the 'first' reference and
// RMSD_CenteredRef.  ws: the window sizes; ref_sel (centred) or NULL.  As in
// the reference's OpenMP build (:222-233), threads take window sizes with a
// dynamic schedule, each with private frames (nthreads 0: the runtime's
// default; the results do not depend on it).
int cref_avgcorr(const float* frames, int64_t stride, int nframes, const int* sel, int nsel,
                 const double* mass, const double* ref_sel, int use_mass, const int* ws, int nw,
                 double* ct, double* sd, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
  std::vector<double> zero(3 * (size_t)nsel, 0.0);
  Frame Tgt = make_frame(zero.data(), mass, nsel);
  Frame Sum = make_frame(zero.data(), mass, nsel);
  Frame Ref = make_frame(ref_sel ? ref_sel : zero.data(), mass, nsel);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
  for (int wi = 0; wi < nw; ++wi) {
    const int w = ws[wi];
    const double dw = (double)w;
    Sum.ZeroCoords();
    bool first = ref_sel == 0;
    double avg = 0.0, sq = 0.0;
    int sub = 0;
    for (int f = 0; f < nframes; ++f) {
      frame_from_floats(Tgt, frames + (int64_t)f * stride, sel, nsel);
      Sum += Tgt;
      if (f > w - 2) {
        Tgt.Divide(Sum, dw);
        if (first) {
          Ref.SetCoordinates(Tgt);
          Ref.CenterOnOrigin(use_mass != 0);
          first = false;
        }
        const double r = Tgt.RMSD_CenteredRef(Ref, use_mass != 0);
        avg += r;
        sq += r * r;
        frame_from_floats(Tgt, frames + (int64_t)sub * stride, sel, nsel);
        Sum -= Tgt;
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
  }
  return 0;
}


// Hungarian::Initialize / AddElement / Optimize (src/Hungarian.cpp:9-66) on an
// n x n matrix given row by row; map[col] = row, as Optimize returns it.
int cref_hungarian(const double* M, int n, int* map) {
  Hungarian h;
  if (h.Initialize((size_t)n)) return -1;
  for (int k = 0; k < n * n; ++k) h.AddElement(M[k]);
  std::vector<int> r = h.Optimize();
  for (int c = 0; c < n; ++c) map[c] = r[c];
  return 0;
}

// The steps of SymmetricRmsdCalc::SymmRMSD_CenteredRef (src/SymmetricRmsdCalc.
// cpp:125-193) on the reference's Frame, Hungarian and DIST2_NoImage, for each
// frame: SetCoordinates, RMSD_CenteredRef + Rotate (fit), cost matrices per
// group, Optimize, SetCoordinatesByMap, RMSD_CenteredRef or RMSD_NoFit.  The
// groups are given (SymmetricRmsdCalc::SetupSymmRMSD derives them from a
// topology).  amap (F x nsel), U (F x 9), trans (F x 3) may be NULL.
int cref_symm_frames(const float* frames, int64_t stride, int nframes, const int* sel, int nsel,
                     const double* mass, const double* ref_sel, int use_mass, int fit,
                     int ngroups, const int* goff, const int* gatoms, double* rmsd, int* amap,
                     double* U, double* trans) {
  std::vector<double> zero(3 * (size_t)nsel, 0.0);
  Frame Sel = make_frame(zero.data(), mass, nsel);
  Frame Remap = make_frame(zero.data(), mass, nsel);
  Frame Ref = make_frame(ref_sel, 0, nsel);
  std::vector<int> AMap((size_t)nsel);
  Matrix_3x3 rot;
  Vec3 tt;
  Hungarian cost;
  for (int f = 0; f < nframes; ++f) {
    const float* fr = frames + (int64_t)f * stride;
    double* x = Sel.xAddress();
    for (int j = 0; j < nsel; ++j) {
      const int a = sel ? sel[j] : j;
      for (int c = 0; c < 3; ++c) x[3 * j + c] = (double)fr[3 * a + c];
    }
    for (int i = 0; i < nsel; ++i) AMap[i] = i;
    Remap.SetCoordinates(Sel);
    if (fit) {
      Remap.RMSD_CenteredRef(Ref, rot, tt, use_mass != 0);
      Remap.Rotate(rot);
    }
    for (int g = 0; g < ngroups; ++g) {
      const int* ga = gatoms + goff[g];
      const int m = goff[g + 1] - goff[g];
      cost.Initialize((size_t)m);
      for (int ta = 0; ta < m; ++ta)
        for (int ra = 0; ra < m; ++ra)
          cost.AddElement(DIST2_NoImage(Ref.XYZ(ga[ra]), Remap.XYZ(ga[ta])));
      std::vector<int> rm = cost.Optimize();
      for (int c = 0; c < m; ++c) AMap[ga[c]] = rm[c] >= 0 ? ga[rm[c]] : ga[c];
    }
    Remap.SetCoordinatesByMap(Sel, AMap);
    double r;
    if (fit)
      r = Remap.RMSD_CenteredRef(Ref, rot, tt, use_mass != 0);
    else
      r = Remap.RMSD_NoFit(Ref, use_mass != 0);
    rmsd[f] = r;
    if (amap) std::memcpy(amap + (int64_t)f * nsel, AMap.data(), sizeof(int) * (size_t)nsel);
    if (U) for (int k = 0; k < 9; ++k) U[9 * (int64_t)f + k] = fit ? rot[k] : (k % 4 == 0);
    if (trans) for (int k = 0; k < 3; ++k) trans[3 * (int64_t)f + k] = fit ? tt[k] : 0.0;
  }
  return 0;
}


// Box::RotateUcell (src/Box.cpp:732-745) as Action_Rmsd::DoAction calls it
// after the rotate + translate update (src/Action_Rmsd.cpp:381): the box is
// set up from the unit cell (rows = cell vectors), rotated by the Kabsch U,
// and its unit cell and recomputed fractional cell are returned.
int cref_rotate_ucell(const double* ucell, const double* U, double* ucell_out, double* frac_out) {
  Box b;
  if (b.SetupFromUcell(ucell)) return 1;
  b.RotateUcell(Matrix_3x3(U));
  const double* u = b.UnitCell().Dptr();
  const double* f = b.FracCell().Dptr();
  for (int k = 0; k < 9; ++k) {
    ucell_out[k] = u[k];
    if (frac_out) frac_out[k] = f[k];
  }
  return 0;
}

// Box::SetupFromXyzAbg: unit and fractional cell from lengths and angles.
int cref_ucell_from_xyzabg(const double* xyzabg, double* ucell_out, double* frac_out) {
  Box b;
  if (b.SetupFromXyzAbg(xyzabg)) return 1;
  for (int k = 0; k < 9; ++k) {
    ucell_out[k] = b.UnitCell().Dptr()[k];
    frac_out[k] = b.FracCell().Dptr()[k];
  }
  return 0;
}

// DIST2_ImageNonOrtho (src/DistRoutines.cpp): squared minimum-image distance
// in a general cell, as the distance action computes it with imaging.
double cref_dist2_image_nonortho(const double* a, const double* b, const double* ucell,
                                 const double* frac) {
  return DIST2_ImageNonOrtho(a, b, Matrix_3x3(ucell), Matrix_3x3(frac));
}

}  // extern "C"
