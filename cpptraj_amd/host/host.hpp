// host.hpp -- C++ host layer of the best-fit RMSD path.
//
// The reference's command classes for this path, with the reference's names,
// options and outputs, built on the C ABI of libcpptraj_rmsd.so:
//   Analysis_Rms2d   rms2d / 2drms            (src/Analysis_Rms2d.cpp)
//   Action_Rmsd      rms / rmsd, align        (src/Action_Rmsd.cpp, src/Action_Align.cpp,
//                                              src/ReferenceAction.cpp)
//   Metric_RMS       cluster pairwise cache   (src/Cluster/Metric_RMS.cpp,
//                                              src/Cluster/MetricArray.cpp:722-801)
// Each owns its loop nest and calls the library once per matrix, batch of
// frames or cache (the reference calls its per-pair routine inside OpenMP
// loops).  Atom masks arrive resolved (the mask language stays with the
// caller, as Topology::SetupIntegerMask does in the reference); coordinates
// arrive as host FP32 COORDS sets, as DataSet_Coords_CRD holds them.
#pragma once

#include <stdint.h>

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cpptraj_rmsd.h"

namespace cpptraj_amd {

enum RetType { OK = 0, ERR = 1 };

// Command arguments with the consumption rules of the reference's ArgList
// (src/ArgList.cpp:273-436): every lookup marks the tokens it uses; a mask is
// the next unmarked token whose first character (after '!' / '(') is one of
// ':' '@' '^' '*' '=' (StrIsMask, src/StringRoutines.cpp:331-350).
class ArgList {
 public:
  ArgList() = default;
  explicit ArgList(const std::string& line);
  const std::string& Command() const;
  bool hasKey(const char* key);
  bool Contains(const char* key) const;
  std::string GetStringKey(const char* key);
  std::string GetStringNext();
  std::string GetMaskNext();
  int getKeyInt(const char* key, int def);
  std::vector<std::string> Unmarked() const;
  static bool StrIsMask(const std::string& s);

 private:
  std::vector<std::string> args_;
  std::vector<bool> marked_;
};

// Page-locked host memory (hipHostMalloc): COORDS sets live in it so that
// their copies to and from the device run asynchronously at full link speed.
void* pinned_alloc(size_t bytes);
void pinned_free(void* p) noexcept;
template <class T>
struct PinnedAllocator {
  using value_type = T;
  PinnedAllocator() = default;
  template <class U>
  PinnedAllocator(const PinnedAllocator<U>&) {}
  T* allocate(size_t n) { return static_cast<T*>(pinned_alloc(n * sizeof(T))); }
  void deallocate(T* p, size_t) noexcept { pinned_free(p); }
  template <class U>
  bool operator==(const PinnedAllocator<U>&) const { return true; }
  template <class U>
  bool operator!=(const PinnedAllocator<U>&) const { return false; }
};
using PinnedFloats = std::vector<float, PinnedAllocator<float>>;
using PinnedDoubles = std::vector<double, PinnedAllocator<double>>;

// Topology facts the path reads: atom count, masses and residues.
struct Topology {
  int natom = 0;
  std::vector<double> mass;
  std::vector<int> resFirst;         // first atom of each residue (0-based)
  std::vector<std::string> resName;  // residue names, trailing blanks removed
  std::vector<int> soluteRes;        // Topology::SoluteResidues() (0-based); empty: all
  // Groups of potentially symmetric atoms (atom indices), residue by residue,
  // as AtomMap::SymmetricAtoms (src/AtomMap.cpp:374-439) finds them from the
  // bonds and elements; given by the caller (the host layer reads no bonds).
  std::vector<std::vector<int>> symmGroups;
  int Nres() const { return (int)resFirst.size(); }
  int ResEnd(int r) const { return r + 1 < Nres() ? resFirst[r + 1] : natom; }
};

// A COORDS set: nframes FP32 frames of `stride` floats each, the 3*natom
// coordinates first (box values, if any, after them).
struct CoordsSet {
  std::string name;
  Topology top;
  int64_t stride = 0;
  PinnedFloats xyz;
  // Unit cell of every frame (9 doubles, rows = cell vectors, Box::UnitCell),
  // or empty when the set has no box.
  PinnedDoubles box;
  int nframes() const { return stride ? (int)(xyz.size() / (size_t)stride) : 0; }
  const float* frame(int f) const { return xyz.data() + (int64_t)f * stride; }
  float* frame(int f) { return xyz.data() + (int64_t)f * stride; }
};

// Resolved atom selection (AtomMask::Selected()).
using AtomMask = std::vector<int>;
// Mask expression -> selection for a set's topology.  Returns false when the
// expression cannot be resolved.
using MaskResolver =
    std::function<bool(const std::string& expr, const CoordsSet& set, AtomMask* sel)>;

// DataSet_MatrixFlt (src/DataSet_MatrixFlt.h, src/Matrix.h:94-122): TRI is the
// packed upper triangle of an n x n symmetric matrix (diagonal 0), FULL is
// ncols x nrows with element (x, y) at y * ncols + x.
class MatrixFlt {
 public:
  enum Kind { FULL = 0, TRI = 1 };
  void AllocateTriangle(size_t n);
  void Allocate2D(size_t ncols, size_t nrows);
  size_t Nrows() const { return nrows_; }
  size_t Ncols() const { return ncols_; }
  Kind kind() const { return kind_; }
  double GetElement(size_t x, size_t y) const;
  float* data() { return v_.data(); }
  const float* data() const { return v_.data(); }
  size_t size() const { return v_.size(); }
  std::string name;
  // Pairwise-cache bookkeeping (DataSet_PairwiseCache, src/DataSet_PairwiseCache.h):
  // set on a cluster cache; `frames` are the cached frame numbers (row order)
  // of a run over `ntotal` frames with sieve `sieve`.
  bool pwcache = false;
  size_t ntotal = 0;
  int sieve = 1;
  std::vector<int> frames;
  // DataSet_PairwiseCache::CachedFramesMatch (src/DataSet_PairwiseCache.cpp:39-48):
  // every frame of `want` is below ntotal and cached.
  bool CachedFramesMatch(const std::vector<int>& want) const;

 private:
  Kind kind_ = FULL;
  size_t ncols_ = 0, nrows_ = 0;
  std::vector<float> v_;
};

// A per-frame double series (DataSet_double, DataSet_Mat3x3, DataSet_Vector):
// ncomp values per frame, written with width w and precision p.
struct Series {
  std::string name;
  std::string legend;  // plot legend (Grace); the name when empty
  double xmin = 1.0, xstep = 1.0;  // x of value i: xmin + i * xstep (DataSet dimension)
  int ncomp = 1;
  int width = 12, prec = 4;
  std::vector<double> v;
  size_t nframes() const { return ncomp ? v.size() / (size_t)ncomp : 0; }
};

// Named data of a session (the part of DataSetList the path uses).
class DataSetList {
 public:
  CoordsSet* AddCoords(const std::string& name);
  CoordsSet* FindCoordsSet(const std::string& name) const;  // "" -> the default (first) set
  CoordsSet* AddReference(const std::string& name);         // reference frames ('ref <name>')
  CoordsSet* ReferenceAt(int idx) const;                    // 'refindex <#>'
  CoordsSet* FindReference(const std::string& name) const;  // "" -> the first reference
  MatrixFlt* AddMatrix(const std::string& name);
  Series* AddSeries(const std::string& name);
  const MatrixFlt* FindMatrix(const std::string& name) const;
  MatrixFlt* FindPairCache(const std::string& name);  // a matrix with pwcache set
  const Series* FindSeries(const std::string& name) const;
  std::string GenerateDefaultName(const std::string& base);

 private:
  std::vector<std::unique_ptr<CoordsSet>> coords_, refs_;
  std::vector<std::unique_ptr<MatrixFlt>> mats_;
  std::vector<std::unique_ptr<Series>> series_;
  std::map<std::string, int> counters_;
};

// Text output of DataIO_Std (src/DataIO_Std.cpp:1556-1620 for matrices, the
// 1D writer for series).
namespace dataio {
int WriteMatrix(const std::string& path, const MatrixFlt& m, bool square2d);
int WriteSeries(const std::string& path, const std::vector<const Series*>& sets);
// Grace xy sets (DataIO_Grace::WriteDataNormal / WriteDataInverted,
// src/DataIO_Grace.cpp:179-224, :276-313), x = frame number.
int WriteGrace(const std::string& path, const std::vector<const Series*>& sets, bool invert);
// XY mesh sets sharing one x column (DataSet_Mesh through DataIO_Std):
// "#<xlabel>" header, x as %8.3f, then each set's values.
int WriteMesh(const std::string& path, const std::string& xlabel, const std::vector<double>& x,
              const std::vector<const Series*>& sets);
}  // namespace dataio

// The library context plus device staging for host COORDS sets.
class GpuSession {
 public:
  explicit GpuSession(int device = 0);
  ~GpuSession();
  GpuSession(const GpuSession&) = delete;
  GpuSession& operator=(const GpuSession&) = delete;
  bool ok() const { return ctx_ != nullptr; }
  rmsd_ctx* ctx() const { return ctx_; }
  const std::string& error() const { return err_; }
  // Grow-only named device buffer.
  void* Buffer(const std::string& name, size_t bytes);
  int Upload(void* dst, const void* src, size_t bytes);
  int Download(void* dst, const void* src, size_t bytes);
  int Sync();
  // Device copy of a COORDS set (all frames) with a selection and its masses.
  int FramesOnDevice(const std::string& tag, const CoordsSet& set, const AtomMask& sel,
                     rmsd_frames* out, const float* xyz_override = nullptr, int nframes = -1);
  // Records an error from a library return code.
  int Check(int rc, const char* what);
  void SetError(const std::string& e) { err_ = e; }
  // Transfer slot k (0 or 1): a non-blocking stream with its own library
  // context, for double-buffered host <-> device streaming.
  int Slot(int k, rmsd_ctx** ctx, void** stream);

 private:
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  int device_ = 0;
  rmsd_ctx* ctx_ = nullptr;
  rmsd_ctx* slot_ctx_[2] = {nullptr, nullptr};
  void* slot_stream_[2] = {nullptr, nullptr};
  std::map<std::string, Buf> bufs_;
  std::string err_;
};

// rms2d / 2drms.
class Analysis_Rms2d {
 public:
  enum ModeType { RMS_FIT = 0, RMS_NOFIT, DME, SRMSD, QUAT };
  RetType Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  RetType Analyze();
  const MatrixFlt* Matrix() const { return rmsdataset_; }
  const Series* Corr() const { return Ct_; }

 private:
  ModeType mode_ = RMS_FIT;
  bool useMass_ = false;
  int precision_ = RMSD_PREC_FP64;
  bool square2d_ = true;
  std::string outfile_, corrfile_, tgtMaskExpr_, refMaskExpr_;
  CoordsSet* TgtTraj_ = nullptr;
  CoordsSet* RefTraj_ = nullptr;
  bool useReferenceTraj_ = false;
  MatrixFlt* rmsdataset_ = nullptr;
  Series* Ct_ = nullptr;
  MaskResolver resolve_;
  GpuSession* gpu_ = nullptr;
};

// rms / rmsd over the frames of a COORDS set (the action applied to every
// frame in order, as a trajectory run would).  Reference modes: first,
// previous, reftraj <set>, ref <name> / reference.
class Action_Rmsd {
 public:
  enum ModeType { ROT_AND_TRANS = 0, TRANS_ONLY, NONE };
  enum TvecType { NO_TVEC = 0, COMBINED, SEPARATE };
  enum RefModeType { FIRST = 0, PREVIOUS, TRAJ, FRAME };
  RetType Init(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  // 'align' (src/Action_Align.cpp): the same fit, always applied, to every atom
  // or to the atoms of 'move <mask>' only.
  RetType InitAlign(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  // Runs the action over every frame of `crd` (rewriting its coordinates
  // unless 'nofit' or 'nomod').  Frames go to the library in batches: against
  // one fixed reference (first / ref), or each against its own (reftraj /
  // previous, rmsd_fit_pairs); one at a time only for 'previous' / 'reftraj'
  // with per-residue values or a move mask.
  RetType DoFrames(CoordsSet& crd);
  // Output files; with 'perres', also Action_Rmsd::Print (src/Action_Rmsd.cpp:
  // 423-470): the per-residue series (perresout, Grace) and their average and
  // standard deviation (perresavg).
  int WriteOutput();
  const Series* Rmsd() const { return rmsd_; }

 private:
  RetType InitRef(ArgList& args, DataSetList& dsl, GpuSession& gpu);
  RetType RunBatch(CoordsSet& crd, const AtomMask& tsel, int f0, int nf, const float* refxyz,
                   const Topology& reftop, const std::string& refname);
  // 'reftraj' (refset = the reference trajectory: frame f against its frame f)
  // or 'previous' (refset = crd: frame f against frame f - 1 after this
  // action's update of it; frame 0 against itself), in chunks of frames.
  RetType RunPairs(CoordsSet& crd, const AtomMask& tsel, CoordsSet& refset, bool previous);
  // Action_Rmsd::perResSetup (src/Action_Rmsd.cpp:197-315): residue masks of
  // the target and reference topologies, one data set per residue.
  RetType PerResSetup(const Topology& tgt, const CoordsSet& tgtset, const Topology& ref,
                      const CoordsSet& refset);
  // Per-residue table (library layout) of the active residues.
  struct PerRes {
    int tgtRes = 0;  // 1-based residue number (the data set's index)
    Series* data = nullptr;
    bool active = false;
  };
  bool perres_ = false, perresCenter_ = false, perresInvert_ = false, perresReady_ = false;
  std::string perresOut_, perresAvg_, perresMask_, tgtRange_, refRange_;
  std::vector<PerRes> resRMS_;
  std::vector<int32_t> prOff_, prTgt_, prRef_;  // active residues only
  std::vector<double> prTmass_, prRmass_;
  std::vector<int> prSeries_;                   // active residue -> resRMS_ index
  bool fit_ = true, useMass_ = false;
  ModeType mode_ = ROT_AND_TRANS;
  TvecType tvecType_ = NO_TVEC;
  RefModeType refMode_ = FIRST;
  int precision_ = RMSD_PREC_FP64;
  int chunkFrames_ = 0;
  std::string outfile_, matricesout_, vecsout_, tgtMaskExpr_, refMaskExpr_, moveMaskExpr_;
  CoordsSet* refCrd_ = nullptr;
  Series* rmsd_ = nullptr;
  Series* rmatrices_ = nullptr;
  Series* tvecs_ = nullptr;
  MaskResolver resolve_;
  GpuSession* gpu_ = nullptr;
  DataSetList* dsl_ = nullptr;
};

// Action_SymmetricRmsd (src/Action_SymmetricRmsd.cpp): 'symmrmsd [<name>]
// <mask> [<refmask>] [out <file>] [nofit] [mass] [remap]' with the reference
// modes of rms (first, previous, reftraj <set>, ref <name> / reference /
// refindex <#>).  Symmetric groups: the target topology's symmGroups, kept for
// residues up to the last selected one and reduced to selected atoms
// (SymmetricRmsdCalc::SetupSymmRMSD, src/SymmetricRmsdCalc.cpp:28-119).  With
// a fixed reference every frame goes to the library in one call; 'remap' and
// the fit's coordinate update (Trans_Rot_Trans) are then applied to the
// frames of the set.
class Action_SymmetricRmsd {
 public:
  RetType Init(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  RetType DoFrames(CoordsSet& crd);
  int WriteOutput() const;
  const Series* Rmsd() const { return rmsd_; }

 private:
  RetType Run(CoordsSet& crd, const AtomMask& tsel, int f0, int nf, const float* refxyz,
              const Topology& reftop, const std::string& refname);
  bool fit_ = true, useMass_ = false, remap_ = false;
  Action_Rmsd::RefModeType refMode_ = Action_Rmsd::FIRST;
  std::string outfile_, tgtMaskExpr_, refMaskExpr_;
  CoordsSet* refCrd_ = nullptr;
  Series* rmsd_ = nullptr;
  std::vector<int32_t> goff_, gatoms_;
  MaskResolver resolve_;
  GpuSession* gpu_ = nullptr;
};

// Analysis_RmsAvgCorr (src/Analysis_RmsAvgCorr.cpp): 'rmsavgcorr [crdset <set>]
// [<name>] [<mask>] [out <file>] [output <file>] [mass] [stop <maxwindow>]
// [offset <offset>] {ref <name> | reference | refindex <#> [<refmask>] | first}'.
// One library call for every window size.
class Analysis_RmsAvgCorr {
 public:
  RetType Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  RetType Analyze();
  int WriteOutput() const;

 private:
  CoordsSet* coords_ = nullptr;
  CoordsSet* ref_ = nullptr;  // null: 'first'
  bool useMass_ = false;
  int maxwindow_ = -1, lagOffset_ = 1;
  std::string outfile_, separate_, tgtMaskExpr_, refMaskExpr_;
  Series* Ct_ = nullptr;
  Series* Csd_ = nullptr;
  std::vector<int32_t> windows_;
  MaskResolver resolve_;
  GpuSession* gpu_ = nullptr;
};

// Cluster::Metric_RMS with the pairwise cache of Cluster::MetricArray:
// 'cluster' arguments rms [<mask>] [mass] [nofit] [sieve <#>]
// [pairdist <name> [pairdistfile <file>]] [loadpairdist] [savepairdist]
// [pwrecalc]; loads, builds and optionally saves the cache (Cmatrix binary,
// src/Cluster/Cmatrix_Binary.cpp) with MetricArray::setupPairwiseCache's rules
// (src/Cluster/MetricArray.cpp:104-272): with 'loadpairdist', or 'pairdist'
// without 'savepairdist', an existing cache set of that name is used, else the
// file is read (a loaded file is not saved again unless 'pwrecalc'), else the
// distances are computed and saved.  A cache whose frames do not cover the
// frames to cluster is an error, or is recomputed with 'pwrecalc'
// (MetricArray::CacheDistances, :725-760).  'srmsd [<mask>] [mass]' caches
// Cluster::Metric_SRMSD's distances instead (symmetric groups from the set's
// topology).  The clustering algorithms that consume the cache are not part of
// this path.
class Metric_RMS {
 public:
  RetType Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve, GpuSession& gpu);
  RetType CacheDistances();
  // FrameDist(f1, f2) of the metric for one pair (one library call).
  double FrameDist(int f1, int f2);
  const MatrixFlt* Cache() const { return cache_; }
  const std::vector<int>& FramesToCache() const { return frames_; }

 private:
  RetType SaveCache();
  bool useMass_ = false, nofit_ = false, save_ = false, symm_ = false;
  bool mismatchFatal_ = true;  // no 'pwrecalc'
  std::vector<int32_t> goff_, gatoms_;  // symmetric groups (selection indices)
  int sieve_ = 1;
  int precision_ = RMSD_PREC_FP64;
  std::string maskExpr_, pairdistFile_;
  CoordsSet* crd_ = nullptr;
  AtomMask sel_;
  std::vector<int> frames_;
  MatrixFlt* cache_ = nullptr;
  MaskResolver resolve_;
  GpuSession* gpu_ = nullptr;
};

// Cmatrix binary ('CTM' v2, src/Cluster/Cmatrix_Binary.cpp).
namespace cmatrix {
// sieve: the sieve value stored in the header; frames: the cached frame list
// (size nrows) of a run over `actual_nrows` frames.
int Write(const std::string& path, const MatrixFlt& tri, size_t actual_nrows, int sieve,
          const std::vector<int>& frames);
int Read(const std::string& path, MatrixFlt* tri, size_t* actual_nrows, int* sieve,
         std::vector<int>* frames);
}  // namespace cmatrix

// Dispatches one command line ("rms2d ...", "2drms ...", "rms ...",
// "rmsd ...", "align ...", "symmrmsd ...", "rmsavgcorr ...", "cluster ...")
// on a session's data.
class Session {
 public:
  explicit Session(int device = 0) : gpu(device) {}
  int Run(const std::string& line);
  DataSetList dsl;
  GpuSession gpu;
  // Mask table: "<set>|<expression>" (that set's topology) or "<expression>"
  // (any set) -> selection.
  std::map<std::string, AtomMask> masks;
  std::string err;
};

}  // namespace cpptraj_amd
