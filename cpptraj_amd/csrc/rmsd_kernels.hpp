// rmsd_kernels.hpp -- argument blocks and launchers shared by the HIP kernels
// and the C ABI implementation (capi.hip).  Internal: not part of the public
// header include/cpptraj_rmsd.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rmsd_device.hpp"

enum { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2 };

// Atoms per fragment chunk for each precision: one 16-byte piece per lane
// holds 2 doubles, 4 floats or 8 bfloat16 (x 4 lane groups).
inline constexpr int chunk_atoms(int prec) {
  return prec == PREC_F64 ? 8 : (prec == PREC_F32 ? 16 : 32);
}
inline constexpr int chunk_parts(int prec) { return prec == PREC_BF16X3 ? 2 : 1; }

// A set of frames as cpptraj stores them.  AoS (ld == 0): coordinate c of atom
// a in frame f at xyz[f*stride + 3a + c] (CompactFrameArray / Frame layout).
// SoA (ld > 0): at xyz[f*stride + c*ld + a].  sel (may be null) lists the
// selected atoms; nsel atoms are used.
struct FrameSetDev {
  const float* xyz;
  int64_t stride;
  int64_t ld;
  int nframes;
  const int* sel;
  int nsel;
  const int* fidx;  // frame f of the set is stored frame fidx[f] (null: f)
};

// p as a global-address-space pointer.  Kernels that read their arguments from a
// device array (the batched ones) would otherwise access memory through flat
// instructions (the compiler cannot tell where such a pointer points; flat
// accesses also wait on the LDS counter and cannot use a scalar base); every
// pointer of the argument sets addresses device memory.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gp(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

__host__ __device__ inline int64_t frame_of(const FrameSetDev& fs, int64_t f) {
  return fs.fidx ? (int64_t)fs.fidx[f] : f;
}

// All three coordinates of selected atom k of frame f.  An AoS atom is one
// 12-byte load: three 4-byte gathers would look its cache lines up three times.
typedef float f32x3_a4 __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ void load_atom(const FrameSetDev& fs, int64_t f, int k, float (&p)[3]) {
  const int a = fs.sel ? gp(fs.sel)[k] : k;
  const int64_t fo = fs.fidx ? (int64_t)gp(fs.fidx)[f] : f;
  const auto fr = gp(fs.xyz + fo * fs.stride);
  if (fs.ld) {
    p[0] = fr[a];
    p[1] = fr[fs.ld + a];
    p[2] = fr[2 * fs.ld + a];
  } else {
    const f32x3_a4 q = *gp(reinterpret_cast<const f32x3_a4*>(fs.xyz + fo * fs.stride + 3 * a));
    p[0] = q.x;
    p[1] = q.y;
    p[2] = q.z;
  }
}

// Four consecutive selected atoms from 16-byte loads: 1 for SoA rows (ld and
// stride multiples of 4), 2 for AoS frames (stride a multiple of 4), both with
// a 16-byte aligned base and no selection; 0 when neither applies.  One wave
// per frame is bound by the number of its load instructions: the pair fit on
// 100,000 pairs of 3,000 SoA atoms takes 1.54 ms with these loads and 4.14 ms
// with one 4-byte load per atom and coordinate.
__host__ __device__ inline int quad_layout(const FrameSetDev& fs) {
  if (fs.sel || ((uintptr_t)fs.xyz % 16) != 0 || (fs.stride % 4) != 0) return 0;
  if (fs.ld > 0) return (fs.ld % 4) == 0 ? 1 : 0;
  return 2;
}

typedef __attribute__((ext_vector_type(4))) float f32x4_t;

// Atoms 4q ... 4q + 3 of frame f (quad_layout(fs) == lay != 0, 4q + 3 < nsel).
__device__ __forceinline__ void load_quad(const FrameSetDev& fs, int64_t f, int q, int lay,
                                          float (&x)[4], float (&y)[4], float (&z)[4]) {
  const int64_t fo = fs.fidx ? (int64_t)gp(fs.fidx)[f] : f;
  const float* fr = fs.xyz + fo * fs.stride;
  if (lay == 1) {
    const f32x4_t a = *gp(reinterpret_cast<const f32x4_t*>(fr + 4 * q));
    const f32x4_t b = *gp(reinterpret_cast<const f32x4_t*>(fr + fs.ld + 4 * q));
    const f32x4_t c = *gp(reinterpret_cast<const f32x4_t*>(fr + 2 * fs.ld + 4 * q));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      x[e] = a[e];
      y[e] = b[e];
      z[e] = c[e];
    }
  } else {
    const auto V = gp(reinterpret_cast<const f32x4_t*>(fr + 12 * q));
    const f32x4_t p0 = V[0], p1 = V[1], p2 = V[2];
    x[0] = p0.x; y[0] = p0.y; z[0] = p0.z;
    x[1] = p0.w; y[1] = p1.x; z[1] = p1.y;
    x[2] = p1.z; y[2] = p1.w; z[2] = p2.x;
    x[3] = p2.y; y[3] = p2.z; z[3] = p2.w;
  }
}

// Coefficients of the per-pair error bound of the MFMA-reduced forms (the
// model and its constants are set on the host: rmsd_bound_coef).
struct BoundCoef {
  double drop;  // times l_i l_j: the split's dropped lo*lo products
  double mix;   // times (l_i ||D_j|| + ||D_i|| l_j): FP32 operand rounding
  double rep;   // times b_i b_j: residual of the two-term split
  double acc;   // times ||D_i|| ||D_j||: FP32 accumulation in the MFMAs
  double e0;    // times E0: FP64 reassembly and QCP
};

// Per-frame preparation for the pairwise contraction (k_prep.hip).
struct PrepArgs {
  FrameSetDev fs;
  const double* cmass;  // centroid weights (null: geometric centre)
  const double* smass;  // covariance weights (null: unit)
  int prec;
  int KB;               // chunks along the atoms
  int Fp;               // padded frame count (multiple of 64)
  void* packed;         // fragment-packed output
  double* g;            // per frame: sum w |x~|^2
  double* u;            // per frame (MFMA forms): T^T D, 9 doubles
  const double* T;      // template, SoA [3][Kp] FP64 (MFMA forms)
  int Kp;
  const double* gT;     // sum |T|^2 (device scalar)
  // MFMA forms, per frame (4 doubles each, may be null): the error bound's
  // {cs, ||D||_2, (sum_c ||D_c||_4^2)^(1/2), ||l||_2} (cs: the per-frame
  // coefficient of rmsd_dev::bound_keep, from `bound` and the total weight),
  // and the centroid {x, y, z, 0} that the FP64 recomputation of a pair needs.
  double* nrm = nullptr;
  double* cen = nullptr;
  const double* total_mass = nullptr;
  BoundCoef bound{};
};


struct SetupArgs {
  FrameSetDev fs;       // template source set
  int f0;               // template frame
  const double* cmass;
  const double* smass;
  int Kp;
  double* T;            // out: [3][Kp] (may be null when no template is needed)
  double* P;            // out: T^T T (9)
  double* gT;           // out: sum |T|^2
  double* total_mass;   // out: sum w (or nsel)
  // Mean-structure template (batched form): samples and their scratch.
  int nsamp = 0;
  double* part = nullptr;  // nsamp * 3 * Kp doubles
};

struct PairArgs {
  const void* A;        // packed row frames
  const void* B;        // packed column frames
  int KB;
  const double* gA;
  const double* gB;
  const double* uA;
  const double* uB;
  const double* P;
  const double* total_mass;
  const int2* tiles;    // (row tile, column tile), grouped order
  int ntiles;
  int per_xcd;
  int nxcd;
  int FA, FB;           // real frame counts
  int tri;              // packed upper triangle (A == B); else full FA x FB
  int clamp;            // 1: Kabsch-style clamp at 0; 0: QCP-style fabs
  float* out;
  // Error bound of the MFMA-reduced forms (off when wlist is null): per-frame
  // norms (PrepArgs::nrm, first entry cs), the FP64 term's coefficient, the
  // tolerance split.  Wave w of tile t (w = t * waves + wave) with a pair whose
  // bound exceeds tol stores each lane's pairs over it as a bit mask,
  // masks[w * 64 + lane] (bit (q WR + r) 4 + v: its accumulator block (r, q),
  // item v), and appends w to wlist (nflag[0] entries, nflag[0] zeroed before
  // the launch); rmsd_launch_pair_fallback recomputes those pairs in FP64.
  // stats[0] collects the largest bound of the pairs that stay (as the bits of
  // a non-negative double).
  const double* nrmA = nullptr;
  const double* nrmB = nullptr;
  double ke0 = 0.0;
  rmsd_dev::KeepSplit keep{};
  int* wlist = nullptr;
  unsigned int* nflag = nullptr;
  unsigned int* masks = nullptr;
  unsigned long long* stats = nullptr;
};

// Wave layout of a contraction variant: each wave owns (16 WR) x (16 WC) frame
// pairs of an NWR x NWC grid of waves.
struct WaveLayout {
  int WR, WC, NWR, NWC;
};

// FP64 recomputation of the pairs whose bound exceeds the tolerance
// (k_pairwise.hip, k_pair_fallback): one wave per 16 x 16 block of a wave the
// contraction listed (PairArgs::wlist, masks), FP64 MFMA over the atoms of
// the frames as stored.
struct FallbackArgs {
  FrameSetDev A, B;
  const double* smass;   // covariance weights (null: unit)
  const double* cenA;    // per-frame centroids (PrepArgs::cen)
  const double* cenB;
  const double* gA;
  const double* gB;
  const double* total_mass;
  const int2* tiles;
  int ntiles;
  const int* wlist;  // the contraction's flagged waves and their lane masks
  const unsigned int* masks;
  int FA, FB;
  int tri;
  int clamp;
  float* out;
  unsigned long long* stats;  // [1]: pairs recomputed
};

// Tiles of many matrices (k_pairwise.hip, batched kernels).
struct BatchArgs {
  const PairArgs* sets;  // one per matrix (device array)
  const int4* tiles;     // (set, row tile, column tile, 0), queue after queue
  const int* qbeg;       // nq + 1 queue offsets into tiles
  int nq;                // queues (one per XCD)
  int maxq;              // longest queue
  int* counters;         // nq position counters (persistent form)
};

// Streaming fit of every frame to one reference (k_fit.hip).
struct FitArgs {
  FrameSetDev fs;
  const double* mass;    // covariance / centroid weights of the target (null: unit)
  const double* ref;     // centred reference, SoA [3][nsel] FP64
  const double* consts;  // [0..2] sum w y, [3] sum w |y|^2, [4] total weight
  double* moments;       // out: 13 per frame, then 3 per frame of shift
  int variant = 0;       // 0: fastest fitting kernel, 1: one block per frame
};

struct SolveArgs {
  const double* moments;
  const double* consts;
  const double* ref_trans;  // origin -> reference centre (3)
  int nframes;
  int fit;                  // 0: no-fit RMSD
  int clamp;
  double* rmsd;
  double* U;                // 9 per frame or null
  double* tvec;             // 6 per frame or null
  double* xform;            // 15 per frame (U, Trans, refTrans) for the coordinate update, or null
};

// Fit of target frame f of T to reference frame f of R, every f (k_fit.hip,
// k_fit_pairs): the 'reftraj' and 'previous' reference modes of
// ReferenceAction (src/ReferenceAction.h:73-88) as one launch.
struct PairFitArgs {
  FrameSetDev T, R;
  const double* tmass;  // weights of the covariance and E0 (null: unit)
  const double* rmass;  // centring masses of the reference (null: unit)
  const double* W;      // sum of tmass, summed in atom order (device scalar)
  int fit;              // 0: RMSD_NoFit
  int clamp;
  double* rmsd;
  double* U;            // 9 per frame or null
  double* tvec;         // 6 per frame or null
  double* xform;        // 15 per frame for the coordinate update, or null
};

struct NofitArgs {
  FrameSetDev fs;
  const double* mass;
  const double* ref;     // SoA [3][ldr] FP64, not centred
};

struct ApplyArgs {
  float* xyz;             // frames to rewrite in place (all natom atoms)
  int64_t stride;
  int64_t ld;
  int natom;
  int nframes;
  const double* xform;    // 15 per frame: U, t1, t2 with x' = U (x + t1) + t2
  const int32_t* atoms = nullptr;  // atoms to move (null: all natom)
  int natoms = 0;
  const double* U = nullptr;       // instead of xform: U (9 per frame) ...
  const double* tvec = nullptr;    // ... and t1, t2 (6 per frame)
};

// Per-residue no-fit RMSD (k_perres.hip).  Residue r owns entries
// res_off[r] .. res_off[r+1] of tgt_atoms / ref_atoms / tmass / rmass.  pieces:
// {first residue, end residue, lowest target atom, highest target atom + 1},
// or {first, end, -1, -1} for a piece that gathers its own atoms; a workgroup
// stages `nfb` frames' atoms at a time.
struct PerresArgs {
  const float* xyz;
  int64_t stride, ld;
  const int32_t* fidx;
  int64_t nframes;
  FrameSetDev ref;        // reference set (xyz, stride, ld, fidx used) ...
  int ref_frame;          // ... and its frame
  int nres;
  const int32_t* res_off;
  const int32_t* tgt_atoms;
  const int32_t* ref_atoms;
  const double* tmass;
  const double* rmass;
  int use_mass, center;
  const double* U;        // 9 per frame (modify 1)
  const double* tvec;     // 6 per frame (modify 1, 2)
  int modify;             // 0: frames as stored; 1: x' = U (x + t) + r; 2: x + (t + r)
  const int4* pieces;
  int nfb;
  double* out;            // out[f * nres + r]
};
constexpr int PERRES_THREADS = 64;
constexpr int PERRES_RES_PIECE = 16;                   // residues per piece
constexpr size_t PERRES_LDS_PIECE = size_t(12) << 10;   // piece budget (LDS bytes at nfb = 1)
constexpr size_t PERRES_LDS_BATCH = size_t(24) << 10;  // budget when batching frames
constexpr size_t PERRES_LDS_MAX = size_t(160) << 10;   // one oversized residue alone
size_t rmsd_perres_lds_bytes(int nres_piece, int nentries, int span, int nfb);
hipError_t rmsd_launch_perres(const PerresArgs& a, int npieces, size_t lds_bytes, hipStream_t s);

// RMS average correlation (k_avgcorr.hip): one wavefront per window size.
struct AvgCorrArgs {
  FrameSetDev fs;         // frames (nsel selected atoms), in order
  const double* mass;     // nsel target masses, or null (unweighted)
  const double* ref;      // centred reference SoA [3][ldr] (null: 'first')
  const double* consts;   // reference constants of rmsd_launch_ref_prep
  int ldr;
  const int* windows;     // nwin window sizes
  int nwin;
  double* ct;             // nwin mean RMSD
  double* sd;             // nwin standard deviation
};
constexpr size_t AVGCORR_LDS_MAX = size_t(160) << 10;  // running sum + reference
hipError_t rmsd_launch_avgcorr(const AvgCorrArgs& a, bool first, hipStream_t s);
// Prefix-sum form of the RMS average correlation with a fixed reference
// (k_avgcorr.hip): prefix sums of the mean-subtracted coordinates and their
// per-row moments, then 64 x 64 tiles of prefix-row pairs (i, j) whose
// difference i - j is a window size, fitted and summed per diagonal.
struct AvgPrefixArgs {
  FrameSetDev fs;
  const double* mass;     // nsel or null
  const double* ref;      // centred reference SoA [3][ldr], null for 'first'
  const double* consts;   // reference constants (sum w y, sum w |y|^2, W); [4] = W
  int ldr;
  int k;                  // 3 nsel
  double* chunk;          // [nchunks][k] scan scratch
  double* mu;             // [k] coordinate means
  double* P;              // [F+1][k] prefix sums (row 0 = 0), times sqrt(mass) when weighted
  double* rows;           // [F+2][13] per-row moments; row F+1 = constants
  const int* wmap;        // [F+1] window index of a window size, else -1
  const int2* tiles;      // (row tile, column tile) pairs to run
  int ntiles;
  int ntile;              // row tiles: ceil((F+1) / 64)
  double* part;           // [nwin][ntile][2][2] per-tile sums of RMSD and RMSD^2
  int tile;               // frames per tile side (64 fixed reference, 32 'first')
  double* yref;           // 'first': [nwin][k] sqrt(m) y_w per window, else null
  double* wconst;         // 'first': [nwin][13] sum m y_w, sum m |y_w|^2, sum m mu y_w^T
  const int* windows;
  int nwin;
  double* ct;
  double* sd;
};
int rmsd_avgcorr_chunks(int64_t nframes);
int rmsd_avgcorr_tile(bool first);
hipError_t rmsd_launch_avgcorr_wsum(const AvgPrefixArgs& a, double* consts, hipStream_t s);
hipError_t rmsd_launch_avgcorr_prefix(const AvgPrefixArgs& a, hipStream_t s);
hipError_t rmsd_launch_avgcorr_tiles(const AvgPrefixArgs& a, hipStream_t s);
hipError_t rmsd_launch_avgcorr_final(const AvgPrefixArgs& a, hipStream_t s);
// Device bytes the prefix form may take (P, moments and per-tile sums).
constexpr size_t AVGCORR_PREFIX_MAX = size_t(16) << 30;

hipError_t rmsd_launch_setup(const SetupArgs& a, hipStream_t s);
hipError_t rmsd_launch_prep(const PrepArgs& a, hipStream_t s);
// Batched forms: one workgroup per set / per padded frame of every set.  The
// prep takes the sets that allow 16-byte loads of four atoms first (nquad sets,
// quad_frames padded frames: fbase[nquad]) and launches them and the rest
// separately.
hipError_t rmsd_launch_setup_batch(const SetupArgs* sets, int nsets, hipStream_t s);
hipError_t rmsd_launch_prep_batch(int prec, const PrepArgs* sets, const int* fbase, int nsets,
                                  int nquad, int quad_frames, int total_frames, hipStream_t s);
hipError_t rmsd_launch_pairwise(int prec, const PairArgs& a, hipStream_t s);
// Wave layout of rmsd_launch_pairwise / rmsd_launch_pairwise_batch variants
// (wlist and masks hold NWR * NWC waves per tile).
WaveLayout rmsd_pairwise_layout(int prec);
WaveLayout rmsd_batch_layout(int prec, int variant);
// nflag: the contraction's count of listed waves, then a work counter (zeroed
// with it).
hipError_t rmsd_launch_pair_fallback(const FallbackArgs& a, WaveLayout w, unsigned int* nflag,
                                     hipStream_t s);
hipError_t rmsd_launch_fill_int(int* p, int n, int v, hipStream_t s);
// Mean-structure template (k_prep.hip): the first template frame's T (from
// rmsd_launch_setup) is replaced by the mean of `nsamp` evenly spaced frames
// of the set, each rotated onto it (FP64 QCP); P and gT are recomputed.
// part: nsamp * 3 * Kp doubles of scratch.
hipError_t rmsd_launch_template_mean(const SetupArgs& a, int nsamp, double* part, hipStream_t s);
// Batched form over every set (SetupArgs::nsamp / part per set); sbase: prefix
// of the sample counts (nsets + 1); maxk: the largest 3 * Kp.
hipError_t rmsd_launch_template_mean_batch(const SetupArgs* sets, const int* sbase, int nsets,
                                           int total_samples, int maxk, hipStream_t s);
// FP64 recomputation over a batch's tile list: tile t = (set, row tile,
// column tile, t), fsets one per set (wlist / nflag / masks are the plan's).
hipError_t rmsd_launch_pair_fallback_batch(const FallbackArgs* fsets, const int4* tiles,
                                           int ntiles, WaveLayout w, const int* wlist,
                                           unsigned int* nflag, const unsigned int* masks,
                                           hipStream_t s);
// Column frames per pair tile of rmsd_launch_pairwise (rows: 64).
int rmsd_pairwise_tile_cols(int prec);
int rmsd_autocorr_chunks(int n, int* rows);
hipError_t rmsd_launch_autocorr(const float* tri, int n, double* part, double* ct, hipStream_t s);
int rmsd_batch_variant_count();
void rmsd_batch_tile_shape(int prec, int variant, int* rows, int* cols);
hipError_t rmsd_launch_batch_reset(const BatchArgs& b, hipStream_t s);
hipError_t rmsd_launch_pairwise_batch(int prec, int variant, const BatchArgs& b, hipStream_t s);
hipError_t rmsd_launch_pair_nofit_tiles(const FrameSetDev& A, const FrameSetDev& B,
                                        const double* mass, const double* total_mass,
                                        const int2* tiles, int ntiles, int per_xcd, int nxcd,
                                        int tri, float* out, hipStream_t s);
hipError_t rmsd_launch_fit_moments(const FitArgs& a, hipStream_t s);
hipError_t rmsd_launch_fit_solve(const SolveArgs& a, hipStream_t s);
// W = sum of the target masses (nsel when null), summed in atom order.
hipError_t rmsd_launch_mass_total(const double* mass, int nsel, double* W, hipStream_t s);
hipError_t rmsd_launch_fit_pairs(const PairFitArgs& a, hipStream_t s);
hipError_t rmsd_launch_nofit_frames(const NofitArgs& a, const double* consts, double* rmsd,
                                    hipStream_t s);
hipError_t rmsd_launch_apply(const ApplyArgs& a, int mode, hipStream_t s);
hipError_t rmsd_launch_rotate_ucell(const double* U, int64_t nframes, double* ucell,
                                    hipStream_t s);
hipError_t rmsd_launch_ref_prep(const FrameSetDev& fs, int f, const double* rmass,
                                const double* tmass, int center, double* ref_out,
                                double* ref_trans, double* consts, hipStream_t s);

// Gradients (k_grad.hip).
hipError_t rmsd_launch_fit_grad(const FrameSetDev& fs, const double* mass, const double* ref,
                                const double* xform, const double* rmsd, const double* gout,
                                const double* consts, double* grad, hipStream_t s);
hipError_t rmsd_launch_center(const FrameSetDev& fs, const double* w, double* xc, double* g,
                              hipStream_t s);
hipError_t rmsd_launch_pair_grad(const double* XA, const double* XB, const double* gA,
                                 const double* gB, const double* w, int n, int FA, int FB,
                                 const double* M, const double* G, double* cw, double* gradA,
                                 double* gradB, hipStream_t s);

// Symmetry-corrected RMSD (k_symm.hip).  Group g owns entries goff[g] ..
// goff[g+1] of gatoms (selection indices); pos[i] is atom i's entry or -1.
struct SymmArgs {
  FrameSetDev fs;
  const double* mass;     // selected target masses (null: unit)
  const double* ref;      // reference SoA [3][ldr]: centred (fit) or as stored
  int ldr;
  const double* ref_trans;
  int fit;
  const double* U0;       // initial fit: U (9 per frame) and Trans (tvec layout,
  const double* t0;       // 6 per frame); null without fit
  const double* moments;  // initial fit's moments (13 per frame, then 3 of shift)
  const double* consts;   // reference constants of rmsd_launch_ref_prep
  int ngroups;
  const int* goff;
  const int* gatoms;
  int ngatoms;            // entries of gatoms
  const int* pos;         // nsel
  const int* bigoff;      // per group: offset (doubles) of its working matrix
  int64_t bigstride;      // doubles per frame of big-group scratch
  double* big;            // groups over SYMM_GROUP_REG atoms, frames f0 .. f1
  int64_t f0, f1;         // frames of this assignment launch
  int* gmap;              // F x ngatoms: target atom for each group entry
  double* rmsd;
  int32_t* amap;          // F x nsel or null
  double* U;              // 9 per frame or null
  double* tvec;           // 6 per frame or null
};
constexpr int SYMM_GROUP_REG = 4;    // largest group kept in registers
constexpr int SYMM_GROUP_MAX = 32;   // largest group (bit-mask lines, byte indices)
constexpr size_t SYMM_BIG_MAX = size_t(1) << 30;  // bytes of big-group scratch per launch
hipError_t rmsd_launch_symm(const SymmArgs& a, hipStream_t s);
