/*
 * cpptraj_rmsd.h -- C ABI of libcpptraj_rmsd.so, the MI355X (gfx950) best-fit
 * RMSD library.
 *
 * Drop-in boundary for cpptraj's best-fit RMSD path.  Each entry point names
 * the reference loop or function it replaces (paths relative to the cpptraj
 * tree).  The caller keeps its own option handling (masks, 'mass', 'nofit',
 * 'qrmsd', reference modes) and passes the resolved atom selection, masses and
 * frame storage; the library owns the loop nest and the device work.
 *
 * Conventions
 *   - Every array argument is a device pointer (HBM) unless noted.  The caller
 *     owns all input and output buffers; the library owns the scratch memory of
 *     its context.  Nothing is allocated or freed across the boundary.
 *   - Frames are FP32, as cpptraj's COORDS sets store them
 *     (CompactFrameArray, src/CompactFrameArray.h:75-76).  Internally every
 *     sum, centroid, E0 and the QCP step are FP64.
 *   - 3x3 matrices are row-major doubles.  U is the Kabsch rotation of
 *     Frame::RMSD_CenteredRef (src/Frame.cpp:1247-1258): x' = U x maps the
 *     centred target onto the centred reference.
 *   - Output matrices use the packed layouts of Matrix<float>
 *     (src/Matrix.h:94-122): TRIANGLE index n*i - (i+1)*i/2 + j - i - 1 for
 *     i < j, FULL index ntgt*nref_total + nref.
 *   - Functions return RMSD_OK (0) or an error code; rmsd_last_error() gives
 *     the message.  Numerical sentinels follow the reference: -1 when the total
 *     mass is below 1e-14 (src/Frame.cpp:1160-1163), 0 for a negative
 *     residual in the fit mode (src/Frame.cpp:1264-1267).
 *   - A context is used by one host thread at a time; different contexts are
 *     independent (one HIP stream each) and may be used concurrently.
 */
#ifndef CPPTRAJ_RMSD_H
#define CPPTRAJ_RMSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMSD_OK 0
#define RMSD_ERR_ARG 1
#define RMSD_ERR_DEVICE 2
#define RMSD_ERR_ALLOC 3

/* Distance metric.  Analysis_Rms2d modes RMS_FIT, RMS_NOFIT and QUAT
 * (src/Analysis_Rms2d.cpp:33-35, :275-281); the cluster metric's
 * fit / nofit (src/Cluster/Metric_RMS.cpp:44-63). */
#define RMSD_MODE_FIT 0
#define RMSD_MODE_NOFIT 1
#define RMSD_MODE_QUAT 2

/* Arithmetic of the pairwise covariance contraction.  FP64 is the plain
 * kernel; FP32 and BF16X3 (bfloat16 hi+lo split, three products, FP32
 * accumulation) run on MFMA over per-frame pre-aligned deviations.  E0,
 * centroids and QCP are FP64 in every case. */
#define RMSD_PREC_FP64 0
#define RMSD_PREC_FP32 1
#define RMSD_PREC_BF16X3 2

/* Coordinate update of the fit (Action_Rmsd::DoAction, src/Action_Rmsd.cpp:
 * 378-389): none ('nomod'), rotate + translate (default), translate only
 * ('norotate'). */
#define RMSD_MODIFY_NONE 0
#define RMSD_MODIFY_ROT_TRANS 1
#define RMSD_MODIFY_TRANS 2

/* A set of frames in device memory.
 *   AoS (ld == 0): coordinate c of atom a of stored frame f at
 *                  xyz[f*stride + 3*a + c]   (cpptraj CRD frames)
 *   SoA (ld >  0): at xyz[f*stride + c*ld + a]
 * sel     nsel selected atom indices (AtomMask::Selected(), src/AtomMask.h:32),
 *         or NULL for atoms 0..nsel-1.
 * mass    nsel masses of the selected atoms (Frame::SetupFrameFromMask,
 *         src/Frame.cpp:502-512), or NULL; used only with use_mass.
 * fidx    nframes stored-frame indices (a sieved or reordered frame list, as
 *         in Cluster::Cframes), or NULL for frames 0..nframes-1. */
typedef struct rmsd_frames {
  const float* xyz;
  int64_t stride;
  int64_t ld;
  int32_t nframes;
  int32_t natom;
  const int32_t* sel;
  int32_t nsel;
  const double* mass;
  const int32_t* fidx;
} rmsd_frames;

typedef struct rmsd_ctx rmsd_ctx;

/* Context on HIP device `device` using stream `stream` (a hipStream_t, or NULL
 * for the null stream).  No reference counterpart: holds the stream, scratch
 * buffers and tile schedules that cpptraj would re-create per call. */
int rmsd_ctx_create(rmsd_ctx** ctx, int device, void* stream);
void rmsd_ctx_destroy(rmsd_ctx* ctx);
const char* rmsd_last_error(const rmsd_ctx* ctx);
/* Blocks until the context's stream is idle. */
int rmsd_ctx_synchronize(rmsd_ctx* ctx);
/* Bytes of device scratch currently held by the context. */
int64_t rmsd_ctx_scratch_bytes(const rmsd_ctx* ctx);

/* Stage timing (the counterpart of cpptraj's Timer around an analysis,
 * src/Timer.h:10-20, at kernel granularity).  With profiling on, every call
 * records HIP events on the context's stream; rmsd_ctx_stage_ms waits for the
 * last call and returns three durations in milliseconds:
 *   ms[0] per-frame stage (setup, gather/centre/pack; reference preparation),
 *   ms[1] the main kernel (pair contraction + QCP; streaming moments),
 *   ms[2] per-frame epilogue (fit solve / coordinate update; 0 for pairwise). */
int rmsd_ctx_set_profiling(rmsd_ctx* ctx, int enable);
int rmsd_ctx_stage_ms(rmsd_ctx* ctx, double* ms);

/* All-pairs RMSD matrix.
 * Replaces the nref/ntgt loop nest of Analysis_Rms2d::Calculate_2D
 * (src/Analysis_Rms2d.cpp:196-295):
 *   ref == NULL  -> same frames and mask: packed upper triangle of
 *                   tgt->nframes*(tgt->nframes-1)/2 floats (AllocateTriangle);
 *   ref != NULL  -> full matrix of tgt->nframes rows x ref->nframes columns,
 *                   element (ntgt, nref) at ntgt*ref->nframes + nref
 *                   (Allocate2D(totalref, totaltgt) + SetElement(nref, ntgt)).
 * The reference frames are centred with their own masses, the targets with
 * theirs, and the target masses weight the covariance and E0
 * (Frame::RMSD_CenteredRef / QuaternionRMSD_CenteredRef).  ref->nsel must
 * equal tgt->nsel.  precision is ignored for RMSD_MODE_NOFIT (FP64 direct
 * sum, bit-exact with Frame::RMSD_NoFit). */
int rmsd_pairwise(rmsd_ctx* ctx, const rmsd_frames* tgt, const rmsd_frames* ref, int mode,
                  int use_mass, int precision, float* out);

/* Error bound of the mixed-precision forms (RMSD_PREC_FP32, RMSD_PREC_BF16X3)
 * of rmsd_pairwise and rmsd_cluster_cache.  No reference counterpart: the
 * reference computes every pair in FP64.  Each pair's result carries a stated
 * bound on its distance from the FP64 computation (Angstrom), from per-frame
 * norms and a rounding model of the contraction; a pair whose bound exceeds the
 * context's tolerance is recomputed in FP64 in the same call, so every result
 * is within the tolerance of FP64.
 *   tol > 0   recompute above tol (default 1e-4 Angstrom);
 *   tol == 0  no bound (pure mixed precision; the report says -1);
 *   tol < 0   bound reported, nothing recomputed.
 * rmsd_ctx_error_report waits for the context's stream and returns, for the
 * last rmsd_pairwise / rmsd_cluster_cache call, the largest bound among the
 * results kept in mixed precision (0 for FP64 and no-fit calls) and how many
 * pairs were recomputed in FP64. */
int rmsd_ctx_set_tolerance(rmsd_ctx* ctx, double tol);
int rmsd_ctx_error_report(rmsd_ctx* ctx, double* max_bound, int64_t* n_fp64);

/* RMSD auto-correlation of a packed triangle (rms2d 'corr').
 * Replaces Analysis_Rms2d::CalcAutoCorr (src/Analysis_Rms2d.cpp:301-319):
 * ct[0] = 1 and ct[lag] = sum_j exp(-R(j, j+lag)) / (n - lag) for
 * lag = 1 .. n-1, FP64, over the n*(n-1)/2 floats of `tri` (device). */
int rmsd_matrix_autocorr(rmsd_ctx* ctx, const float* tri, int n, double* ct);

/* Cluster pairwise-distance cache fill.
 * Replaces Cluster::MetricArray::calcFrameDistances (src/Cluster/
 * MetricArray.cpp:766-801) for one RMS metric (Metric_RMS::FrameDist,
 * src/Cluster/Metric_RMS.cpp:44-63) with weight `weight` (Manhattan, one
 * metric: d = weight * RMSD).  frames->fidx lists the frames to cache
 * (framesToCache); out receives the packed upper triangle over that list
 * (DataSet_PairwiseCache_MEM::SetElement order). */
int rmsd_cluster_cache(rmsd_ctx* ctx, const rmsd_frames* frames, int nofit, int use_mass,
                       int precision, double weight, float* out);

/* Fit every frame of `frames` to one reference frame.
 * Replaces Action_Rmsd::DoAction (src/Action_Rmsd.cpp:361-417) over a batch of
 * frames for the 'first' / 'ref' / 'refindex' reference modes
 * (ReferenceAction, src/ReferenceAction.h:73-84):
 *   ref, ref_frame  the reference set and frame; its selection and masses
 *                   centre the reference (ReferenceAction::SelectRefAtoms);
 *   fit             0 for 'nofit' (Frame::RMSD_NoFit);
 *   modify          RMSD_MODIFY_*: frames->xyz is rewritten in place (all
 *                   natom atoms of every frame; the buffer must be writable);
 *   rmsd            nframes doubles (DataSet_double of the action);
 *   U               9 doubles per frame or NULL ('savematrices');
 *   tvec            6 doubles per frame or NULL: Trans (target -> origin) and
 *                   refTrans (origin -> reference), 'savevectors separate';
 *                   'combined' is their sum. */
int rmsd_fit_to_ref(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                    int ref_frame, int fit, int use_mass, int modify, double* rmsd, double* U,
                    double* tvec);

/* Fit of every target frame to its own reference frame: target f of `frames`
 * against frame f of `refs` (refs->nframes >= frames->nframes; either set may
 * use fidx).  Replaces the per-frame DoAction of Action_Rmsd
 * (src/Action_Rmsd.cpp:361-417) in the 'reftraj' and 'previous' reference
 * modes of ReferenceAction (src/ReferenceAction.h:73-88), where every frame has
 * a different reference: one call per batch instead of one per frame.  The
 * arguments and outputs mean what they mean for rmsd_fit_to_ref (the reference
 * frame is centred with refs->mass, the target's masses weight the covariance;
 * refTrans in tvec is frame f's reference centre).  With modify, the targets
 * are rewritten after every pair is solved, so refs may share the targets'
 * storage (refs->xyz = frames->xyz - frames->stride gives 'previous' on the
 * frames as stored).  For 'previous' with a coordinate update the reference
 * is the previous frame after its update: fit the frames as stored (modify
 * NONE), then compose on the host U_f = L_(f-1) U_f and refTrans_f =
 * L_(f-1) (refTrans_f + Trans_(f-1)) + refTrans'_(f-1), where L_(f-1) is the
 * rotation applied to frame f-1 (its composed U for RMSD_MODIFY_ROT_TRANS, the
 * identity for RMSD_MODIFY_TRANS) and refTrans'_(f-1) its composed refTrans
 * (RMSD is unchanged by a rigid update of the reference), and apply them with
 * rmsd_transform; the host layer does this. */
int rmsd_fit_pairs(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* refs, int fit,
                   int use_mass, int modify, double* rmsd, double* U, double* tvec);

/* Allocation-free variant of rmsd_fit_to_ref: same arguments, same kernels,
 * same results bit for bit.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * per-frame loops such as Action::DoAction in a serving path, harnesses that
 * forbid hipMalloc).  Its working memory is a fixed per-GPU device array of
 * RMSD_NOALLOC_SCRATCH_BYTES that the library holds until the process ends;
 * one call needs about 32 nsel + 512 bytes plus 128 bytes per frame (248 with
 * a coordinate update): larger calls are rejected with RMSD_ERR_ARG (split the
 * frames; 64 MiB hold about 500,000 frames, 270,000 with the update).  All calls on one GPU share the array,
 * so calls that overlap on one GPU (other streams or threads) race. */
#define RMSD_NOALLOC_SCRATCH_BYTES (64ll << 20)
int rmsd_fit_to_ref_noalloc(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                            int ref_frame, int fit, int use_mass, int modify, double* rmsd,
                            double* U, double* tvec);

/* Unit-cell rotation of the fit's coordinate update.
 * Replaces Box::RotateUcell (src/Box.cpp:732-745) as Action_Rmsd::DoAction
 * applies it after Trans_Rot_Trans in the default rotate + translate mode
 * (src/Action_Rmsd.cpp:378-382): frame f's unit cell (ucell[9f .. 9f+8], rows
 * are the three cell vectors, device memory) becomes row k' = U_f row k, with
 * U_f the fit's Kabsch rotation (rmsd_fit_to_ref's U output).  Lengths,
 * angles and volume are unchanged; the fractional cell is the inverse of the
 * result.  Frames without a box are simply not passed. */
int rmsd_rotate_ucell(rmsd_ctx* ctx, const double* U, int nframes, double* ucell);

/* Per-frame rigid update of selected atoms.
 * Replaces Frame::Trans_Rot_Trans(AtomMask, ...) as Action_Align::DoAction
 * uses it for 'align ... move <mask>' (src/Action_Align.cpp): for every stored
 * frame f of `frames` (xyz rewritten in place; fidx, sel and mass ignored),
 * atoms atoms[0..natoms-1] (every natom atom when atoms is NULL) become
 *   RMSD_MODIFY_ROT_TRANS  x' = U_f (x + t_f) + r_f
 *   RMSD_MODIFY_TRANS      x' = x + (t_f + r_f)
 * with U (9 doubles per frame) and tvec (6 per frame: t = Trans, r = refTrans)
 * as rmsd_fit_to_ref returns them. */
int rmsd_transform(rmsd_ctx* ctx, const rmsd_frames* frames, const int32_t* atoms, int natoms,
                   const double* U, const double* tvec, int mode);

/* Per-residue no-fit RMSD ('rmsd ... perres').
 * Replaces the per-residue loop of Action_Rmsd::DoAction (src/Action_Rmsd.cpp:
 * 398-413) over the residue masks of Action_Rmsd::perResSetup (:197-315),
 * which the caller resolves: residue r compares the target atoms
 * tgt_atoms[res_off[r] .. res_off[r+1]) of every frame f of `frames` (fidx
 * honoured; sel and mass ignored) with the reference atoms ref_atoms[...] of
 * frame ref_frame of `ref`, uncentred (ReferenceAction::CurrentReference,
 * src/ReferenceAction.h:36), by Frame::RMSD_NoFit (src/Frame.cpp:1279-1307)
 * with the target masses.
 *   res_off, tgt_atoms, ref_atoms, tmass, rmass   HOST arrays (nres+1 ints;
 *            res_off[nres] ints or doubles), res_off[0] = 0; masses are needed
 *            with use_mass only.  Copied to the device, once per new table.
 *   center   'perrescenter': target and reference residue each moved to its
 *            own centre first (Frame::CenterOnOrigin, src/Frame.cpp:1043-1055)
 *   modify   the target as the fit's coordinate update leaves it, in FP64,
 *            before any rounding to FP32: RMSD_MODIFY_ROT_TRANS with U and tvec
 *            of rmsd_fit_to_ref (Frame::Trans_Rot_Trans), RMSD_MODIFY_TRANS
 *            with tvec, RMSD_MODIFY_NONE for the frames as stored.  A caller
 *            fits with RMSD_MODIFY_NONE, calls this, then rmsd_transform.
 *   out      device, nframes * nres doubles: out[f*nres + r]; -1 for a residue
 *            of zero total mass (Constants::SMALL). */
/* Window sizes of the RMS average correlation (Analysis_RmsAvgCorr::Analyze,
 * src/Analysis_RmsAvgCorr.cpp:154-209): 1, then 1 + offset, 1 + 2 offset, ...
 * below min(maxwindow + 1, nframes) (maxwindow -1: nframes).  Host function:
 * writes up to cap sizes into windows (may be NULL) and returns their number,
 * or -1 when there are not enough frames. */
int rmsd_avgcorr_windows(int nframes, int maxwindow, int offset, int32_t* windows, int cap);

/* RMS average correlation ('rmsavgcorr').
 * Replaces Analysis_RmsAvgCorr::Analyze (src/Analysis_RmsAvgCorr.cpp:119-276):
 * for each window size w (rmsd_avgcorr_windows), the running average of the
 * selected atoms over w consecutive frames (FP64 running sum, as the
 * reference's operator+= / operator-= / Divide), best-fit RMSD of every
 * averaged frame (Frame::RMSD_CenteredRef) to the reference, and the mean and
 * standard deviation of those F - w + 1 values.
 *   ref      reference set and frame ('ref' / 'reference': its selected atoms,
 *            centred with its masses), or NULL for 'first' (the window's first
 *            averaged frame, centred)
 *   maxwindow, offset   'stop' (-1: every frame) and 'offset' (< 1 -> 1)
 *   ct, sd   device, one double per window size (DataSet Ct and Ct[SD])
 * The running-sum method takes at most 6826 selected atoms with a reference,
 * 3413 with 'first' (the running sum, and the 'first' reference, are kept in
 * LDS); see rmsd_avgcorr_set_method. */
int rmsd_avgcorr(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref, int ref_frame,
                 int use_mass, int maxwindow, int offset, double* ct, double* sd);

/* How rmsd_avgcorr computes; no counterpart in the reference (its running sum
 * is the RUNNING method).
 *   RMSD_AVGCORR_AUTO     PREFIX when its device memory (about
 *                         8 (F+1) 3 nsel + 32 nwindows ceil((F+1)/T) bytes, T = 64
 *                         with a reference, 32 with 'first', which also keeps
 *                         8 (3 nsel + 13) bytes per window) stays within
 *                         16 GiB, else RUNNING
 *   RMSD_AVGCORR_RUNNING  one wavefront per window size, FP64 running sum
 *                         (atom limits above)
 *   RMSD_AVGCORR_PREFIX   prefix sums of the mean-subtracted coordinates: every
 *                         averaged frame of every window is fitted
 *                         independently, T x T (frame, frame) tiles at a time;
 *                         agrees with RUNNING to rounding (the window sums are
 *                         differences of prefix sums, not running sums); no
 *                         atom limit; an error over the memory limit.
 * The setting stays with the context. */
#define RMSD_AVGCORR_AUTO 0
#define RMSD_AVGCORR_RUNNING 1
#define RMSD_AVGCORR_PREFIX 2
int rmsd_avgcorr_set_method(rmsd_ctx* ctx, int method);

int rmsd_perres_nofit(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                      int ref_frame, int nres, const int32_t* res_off, const int32_t* tgt_atoms,
                      const int32_t* ref_atoms, const double* tmass, const double* rmass,
                      int use_mass, int center, int modify, const double* U, const double* tvec,
                      double* out);

/* Symmetry-corrected RMSD ('symmrmsd').
 * Replaces SymmetricRmsdCalc::SymmRMSD_CenteredRef (src/SymmetricRmsdCalc.cpp:
 * 125-193) in Action_SymmetricRmsd::DoAction (src/Action_SymmetricRmsd.cpp:
 * 88-112), for every frame: with fit, the frame's best fit to the reference
 * (translated, rotated); per group of symmetric atoms, the squared distances
 * between its reference and target atoms and the reference's Hungarian
 * heuristic (Hungarian::Optimize, src/Hungarian.cpp:21-337) to pair them; the
 * target remapped (coordinates and masses) and fitted again, or compared
 * without fit (Frame::RMSD_NoFit).
 *   ref, ref_frame   the reference set and frame (selected atoms): centred with
 *            its masses for a fit, as stored without (ReferenceAction)
 *   ngroups, group_off, group_atoms   HOST arrays: group g is group_atoms
 *            [group_off[g] .. group_off[g+1]), selection indices, 1..32 atoms,
 *            disjoint -- SymmetricRmsdCalc::SetupSymmRMSD's groups (derived by
 *            the caller from the topology, src/SymmetricRmsdCalc.cpp:28-119)
 *   fit, use_mass    'nofit', 'mass'
 *   rmsd     device, one double per frame (-1 for zero total mass)
 *   amap     device, nframes * nsel int32 or NULL: amap[f*nsel + i] = the
 *            selected target atom that reference atom i is paired with
 *            (SymmetricRmsdCalc::AMap)
 *   U, tvec  device or NULL (fit only): the final fit's rotation and
 *            translations in rmsd_fit_to_ref's layout, for rmsd_transform of
 *            the remapped frame (the action's Trans_Rot_Trans)
 * A column the heuristic leaves unassigned keeps its own atom (the reference
 * would index the group at -1).  Groups of up to 4 atoms are solved in
 * registers; a larger group of n atoms takes n*n + ceil(6n/8) doubles of
 * context scratch per frame, and the frames are assigned in chunks that keep
 * that scratch within 1 GiB. */
int rmsd_symm(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref, int ref_frame,
              int ngroups, const int32_t* group_off, const int32_t* group_atoms, int fit,
              int use_mass, double* rmsd, int32_t* amap, double* U, double* tvec);

/* Gradient of the fit: no counterpart in the reference (cpptraj has no RMSD
 * derivative).  For loss = sum_f grad_rmsd[f] * RMSD_f of rmsd_fit_to_ref
 * (fit, same frames / reference / masses), writes d loss / d x for the selected
 * atoms of every frame: grad_xyz[(f*nsel + k)*3 + c], FP64.  The rotations are
 * recomputed from the frames, not taken from the forward call.  Envelope
 * theorem: w_k (x~_k - U^T y~_k) / (M RMSD_f) per atom; 0 where RMSD_f = 0. */
int rmsd_fit_backward(rmsd_ctx* ctx, const rmsd_frames* frames, const rmsd_frames* ref,
                      int ref_frame, int use_mass, const double* grad_rmsd, double* grad_xyz);

/* Gradient of a full pairwise block (rows = A as targets, columns = B as
 * references, as in rmsd_pairwise with ref != NULL, fit mode): for
 * loss = sum_ij grad_out[i*B->nframes + j] * RMSD_ij writes d loss / d A
 * (grad_a, A->nframes x nsel x 3) and d loss / d B (grad_b, may be NULL),
 * FP64.  Both sets are centred and weighted with A's masses (same selection
 * and weights on both sides).  Every pair's rotation is recomputed in FP64. */
int rmsd_pairwise_backward(rmsd_ctx* ctx, const rmsd_frames* A, const rmsd_frames* B,
                           int use_mass, const double* grad_out, double* grad_a,
                           double* grad_b);

/* Batched all-pairs matrices: many trajectories or masks, of uneven frame and
 * atom counts, in one launch of the contraction.  No single reference
 * counterpart: replaces one Analysis_Rms2d::Calculate_2D pass
 * (src/Analysis_Rms2d.cpp:196-295) per COORDS set, for the same-mask case
 * (packed upper triangle of each set into outs[s], nframes*(nframes-1)/2
 * floats, as rmsd_pairwise with ref == NULL).  Sets may differ in nframes,
 * natom, nsel, selection and masses.
 *
 * A plan binds the sets and outputs (device pointers) and owns one device
 * allocation with every set's operands and the tile queues: every set's
 * triangle tiles, sorted by cost and cut into one cost-balanced queue per XCD.
 * rmsd_batch_run recomputes all matrices from the current contents of the
 * bound frame buffers (per-set preparation kernels, then one batched kernel).
 * Flags:
 *   RMSD_BATCH_GRAPH     capture the pass as a hipGraph at creation; runs
 *                        replay it (ordered after prior work on the context's
 *                        stream, and later work waits for it);
 *   RMSD_BATCH_AUTOTUNE  time every kernel variant (tile side 64 or 32, grid or
 *                        persistent work-stealing form) on the plan's own data
 *                        at creation and keep the fastest.
 * RMSD_MODE_NOFIT runs the FP64 direct-sum kernel set by set (no batching). */
#define RMSD_BATCH_GRAPH 1
#define RMSD_BATCH_AUTOTUNE 2
typedef struct rmsd_batch rmsd_batch;
int rmsd_batch_create(rmsd_ctx* ctx, const rmsd_frames* sets, int nsets, int mode, int use_mass,
                      int precision, float* const* outs, int flags, rmsd_batch** plan);
int rmsd_batch_run(rmsd_batch* plan);
/* Selects kernel variant 0..3 (bit 0: persistent form, bit 1: 32-frame tile
 * side) and re-captures the graph. */
int rmsd_batch_set_variant(rmsd_batch* plan, int variant);
/* info (6 values): variant, tile side, tiles, graph captured (0/1), queues,
 * longest queue.  tune_ms (4 values): the tuner's time per variant, -1 when
 * not tuned. */
int rmsd_batch_info(const rmsd_batch* plan, int64_t* info, double* tune_ms);
/* Waits for the plan's work, then frees it.  rmsd_ctx_destroy frees every
 * plan still bound to the context; such a plan must not be used afterwards. */
void rmsd_batch_destroy(rmsd_batch* plan);
/* One-shot form: create (no graph, no tuning), run, destroy. */
int rmsd_pairwise_batch(rmsd_ctx* ctx, const rmsd_frames* sets, int nsets, int mode, int use_mass,
                        int precision, float* const* outs);

#ifdef __cplusplus
}
#endif

#endif /* CPPTRAJ_RMSD_H */
