// host_rms.cpp -- Analysis_Rms2d, Action_Rmsd and Metric_RMS over the C ABI,
// and the command dispatch of a Session.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "host.hpp"

namespace cpptraj_amd {

namespace {

// Optional precision keyword of the GPU path ('prec fp64|fp32|bf16x3');
// absent means FP64, the reference's arithmetic.
bool parse_prec(ArgList& args, int* prec, std::string* err) {
  const std::string p = args.GetStringKey("prec");
  if (p.empty() || p == "fp64") *prec = RMSD_PREC_FP64;
  else if (p == "fp32") *prec = RMSD_PREC_FP32;
  else if (p == "bf16x3") *prec = RMSD_PREC_BF16X3;
  else {
    *err = "unknown precision '" + p + "' (fp64, fp32, bf16x3)";
    return false;
  }
  return true;
}

// The whole topology when the expression is empty (an unset AtomMask selects
// every atom, src/ReferenceAction.cpp:141-144).
bool resolve_or_all(const MaskResolver& r, const std::string& expr, const CoordsSet& set,
                    AtomMask* sel) {
  if (expr.empty() || expr == "*") {
    sel->resize(set.top.natom);
    for (int i = 0; i < set.top.natom; ++i) (*sel)[i] = i;
    return true;
  }
  return r && r(expr, set, sel);
}

// Range::SetRange (src/Range.cpp:49-103): comma-separated numbers or
// "a-b" spans (inclusive), sorted, duplicates removed; empty is an error.
bool parse_range(const std::string& arg, std::vector<int>* out, std::string* err) {
  out->clear();
  if (arg.find_first_of(":@*") != std::string::npos) {
    *err = "Using a mask expression for range (" + arg + ")";
    return false;
  }
  size_t p = 0;
  while (p <= arg.size()) {
    size_t q = arg.find(',', p);
    if (q == std::string::npos) q = arg.size();
    const std::string tok = arg.substr(p, q - p);
    p = q + 1;
    if (tok.empty()) continue;
    const size_t dash = tok.find('-', 1);
    char* end = nullptr;
    const long a = strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str()) {
      *err = "Range::SetRange(" + arg + "): bad number in " + tok;
      return false;
    }
    long b = a;
    if (dash != std::string::npos) b = strtol(tok.c_str() + dash + 1, nullptr, 10);
    for (long v = a; v <= b; ++v) out->push_back((int)v);
  }
  std::sort(out->begin(), out->end());
  out->erase(std::unique(out->begin(), out->end()), out->end());
  if (out->empty()) {
    *err = "empty range '" + arg + "'";
    return false;
  }
  return true;
}

}  // namespace

// ---- symmetric groups and rows of symmetry-corrected RMSDs -----------------------------
// SymmetricRmsdCalc::SetupSymmRMSD (src/SymmetricRmsdCalc.cpp:28-119): the
// topology's groups for residues up to the last selected atom's residue, as
// selection indices; groups with no selected atom are dropped.
static bool symm_groups_for(const Topology& top, const AtomMask& tsel, std::vector<int32_t>* off,
                            std::vector<int32_t>* atoms, std::string* err) {
  std::vector<int> selIdx((size_t)top.natom, -1);
  for (size_t k = 0; k < tsel.size(); ++k) selIdx[tsel[k]] = (int)k;
  int lastRes = top.Nres();
  if (top.Nres() > 0 && !tsel.empty()) {
    const int lastAtom = *std::max_element(tsel.begin(), tsel.end());
    lastRes = (int)(std::upper_bound(top.resFirst.begin(), top.resFirst.end(), lastAtom) -
                    top.resFirst.begin());
  }
  off->assign(1, 0);
  atoms->clear();
  for (const std::vector<int>& grp : top.symmGroups) {
    if (grp.empty()) continue;
    if (top.Nres() > 0) {
      const int r = (int)(std::upper_bound(top.resFirst.begin(), top.resFirst.end(), grp[0]) -
                          top.resFirst.begin()) - 1;
      if (r >= lastRes) continue;
    }
    const size_t before = atoms->size();
    for (int a : grp) {
      if (a < 0 || a >= top.natom) {
        *err = "symmetric group atom out of range of the topology";
        return false;
      }
      if (selIdx[a] >= 0) atoms->push_back(selIdx[a]);
    }
    if (atoms->size() > before) off->push_back((int32_t)atoms->size());
  }
  return true;
}

// A matrix of symmetry-corrected RMSDs, one rmsd_symm call per reference
// frame (the SRMSD case of Analysis_Rms2d::Calculate_2D, src/Analysis_Rms2d.cpp:
// 244-287, and of the cluster cache's symmetric metric).  tri: references are
// the frames of T, each against the frames after it (packed triangle over T);
// otherwise each frame of R against every frame of T (full, x = reference).
// ref_is_column (triangle only): element (i, j), i < j, is frame i against
// frame j as the reference (the cluster cache's FrameDist(f1, f2)) instead of
// frame j against frame i (rms2d).  T may carry a frame list (fidx); a device
// copy of it is indexed by row.
static RetType symm_rows(GpuSession& g, const rmsd_frames& T, const rmsd_frames* R, bool tri,
                         int use_mass, const std::vector<int32_t>& goff,
                         const std::vector<int32_t>& gatoms, MatrixFlt* out,
                         bool ref_is_column = false) {
  const int nt = T.nframes;
  const int nr = tri ? nt : R->nframes;
  if (tri) out->AllocateTriangle(nt);
  else out->Allocate2D(nr, nt);
  if (nt == 0 || nr == 0) return OK;
  // Frame list of T on the device (identity when T has none), so that each
  // row's targets are a contiguous slice of it.
  std::vector<int32_t> ids((size_t)nt);
  for (int k = 0; k < nt; ++k) ids[k] = k;
  int32_t* dids = (int32_t*)g.Buffer("symm.rows.ids", sizeof(int32_t) * nt);
  double* drow = (double*)g.Buffer("symm.rows.out", sizeof(double) * nt);
  if (!dids || !drow) return ERR;
  if (T.fidx) {
    if (hipMemcpy(dids, T.fidx, sizeof(int32_t) * nt, hipMemcpyDeviceToDevice) != hipSuccess) {
      g.SetError("symm rows: frame list copy failed");
      return ERR;
    }
  } else if (g.Upload(dids, ids.data(), sizeof(int32_t) * nt)) {
    return ERR;
  }
  std::vector<int32_t> hostList((size_t)nt);
  if (g.Download(hostList.data(), dids, sizeof(int32_t) * nt)) return ERR;
  const int stored = *std::max_element(hostList.begin(), hostList.end()) + 1;
  std::vector<double> row((size_t)nt);
  const int ng = (int)goff.size() - 1;
  const int32_t* ga = gatoms.empty() ? nullptr : gatoms.data();
  float* o = out->data();
  for (int r = 0; r < nr; ++r) {
    rmsd_frames tg = T;
    tg.fidx = dids;
    rmsd_frames rf = tri ? T : *R;
    int ref_frame = r;
    if (tri) {
      if (ref_is_column ? r == 0 : r == nt - 1) continue;
      tg.fidx = ref_is_column ? dids : dids + r + 1;
      tg.nframes = ref_is_column ? r : nt - r - 1;
      rf.fidx = nullptr;  // the reference frame by its storage index
      rf.nframes = stored;
      ref_frame = hostList[r];
    }
    if (g.Check(rmsd_symm(g.ctx(), &tg, &rf, ref_frame, ng, goff.data(), ga, 1, use_mass, drow,
                          nullptr, nullptr, nullptr),
                "rmsd_symm"))
      return ERR;
    if (g.Download(row.data(), drow, sizeof(double) * tg.nframes)) return ERR;
    if (tri && ref_is_column) {
      for (int i = 0; i < r; ++i)  // element (i, r)
        o[(size_t)nt * i - ((size_t)i * (i + 1)) / 2 + r - i - 1] = (float)row[i];
    } else if (tri) {
      const size_t base = (size_t)nt * r - ((size_t)r * (r + 1)) / 2;  // element (r, r+1)
      for (int k = 0; k < tg.nframes; ++k) o[base + k] = (float)row[k];
    } else {
      for (int t = 0; t < nt; ++t) o[(size_t)t * nr + r] = (float)row[t];
    }
  }
  return OK;
}

// ---- Analysis_Rms2d ------------------------------------------------------------
// Setup follows src/Analysis_Rms2d.cpp:38-148; 'dme' (another metric) is
// rejected here; 'srmsd' uses the target topology's symmetric groups.
RetType Analysis_Rms2d::Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                              GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  const std::string setname = args.GetStringKey("crdset");
  TgtTraj_ = dsl.FindCoordsSet(setname);
  if (!TgtTraj_) {
    gpu.SetError("rms2d: could not locate COORDS set corresponding to " + setname);
    return ERR;
  }
  if (args.hasKey("nofit")) mode_ = RMS_NOFIT;
  else if (args.hasKey("dme")) mode_ = DME;
  else if (args.hasKey("srmsd")) mode_ = SRMSD;
  else if (args.hasKey("qrmsd")) mode_ = QUAT;
  else mode_ = RMS_FIT;
  if (mode_ == DME) {
    gpu.SetError("rms2d: 'dme' is not on the GPU RMSD path");
    return ERR;
  }
  useMass_ = args.hasKey("mass");
  std::string perr;
  if (!parse_prec(args, &precision_, &perr)) {
    gpu.SetError("rms2d: " + perr);
    return ERR;
  }
  outfile_ = args.GetStringKey("out");
  if (outfile_.empty()) outfile_ = args.GetStringKey("rmsout");
  if (args.hasKey("nosquare2d")) square2d_ = false;
  else if (args.hasKey("square2d")) square2d_ = true;
  const std::string reftraj = args.GetStringKey("reftraj");
  useReferenceTraj_ = !reftraj.empty();
  corrfile_ = args.GetStringKey("corr");
  tgtMaskExpr_ = args.GetMaskNext();
  refMaskExpr_ = args.GetMaskNext();
  if (refMaskExpr_.empty()) refMaskExpr_ = tgtMaskExpr_;
  if (!corrfile_.empty()) {
    if (useReferenceTraj_) {
      gpu.SetError("Keyword 'corr' not supported with 'reftraj'");
      return ERR;
    }
    if (tgtMaskExpr_ != refMaskExpr_) {
      gpu.SetError("Keyword 'corr' not supported when masks differ if not using 'reftraj'");
      return ERR;
    }
  }
  if (useReferenceTraj_) {
    RefTraj_ = dsl.FindCoordsSet(reftraj);
    if (!RefTraj_) {
      gpu.SetError("rms2d: reference trajectory '" + reftraj + "' is not a loaded COORDS set");
      return ERR;
    }
  }
  std::string name = args.GetStringNext();
  if (name.empty()) name = dsl.GenerateDefaultName("Rms2d");
  rmsdataset_ = dsl.AddMatrix(name);
  if (!corrfile_.empty()) {
    Ct_ = dsl.AddSeries(name + "[Corr]");
    Ct_->width = 8;
    Ct_->prec = 3;
  }
  return OK;
}

// Analyze + Calculate_2D (src/Analysis_Rms2d.cpp:151-295): the triangle when
// the masks are equal and there is no reference trajectory, else the full
// matrix (columns = reference frames, rows = targets); one library call.
RetType Analysis_Rms2d::Analyze() {
  GpuSession& g = *gpu_;
  AtomMask tsel, rsel;
  if (!resolve_or_all(resolve_, tgtMaskExpr_, *TgtTraj_, &tsel) || tsel.empty()) {
    g.SetError("No atoms selected for [" + tgtMaskExpr_ + "]");
    return ERR;
  }
  CoordsSet* ref = useReferenceTraj_ ? RefTraj_ : TgtTraj_;
  if (!resolve_or_all(resolve_, refMaskExpr_, *ref, &rsel) || rsel.empty()) {
    g.SetError("Could not set up reference mask [" + refMaskExpr_ + "]");
    return ERR;
  }
  if (rsel.size() != tsel.size()) {
    g.SetError("# Selected atoms in '" + tgtMaskExpr_ + "' not equal to selected # atoms in '" +
               refMaskExpr_ + "'");
    return ERR;
  }
  const int mode = mode_ == RMS_NOFIT ? RMSD_MODE_NOFIT
                                      : (mode_ == QUAT ? RMSD_MODE_QUAT : RMSD_MODE_FIT);
  const bool tri = !useReferenceTraj_ && tgtMaskExpr_ == refMaskExpr_;
  rmsd_frames T, R;
  if (g.FramesOnDevice("rms2d.tgt", *TgtTraj_, tsel, &T)) return ERR;
  const int ntgt = TgtTraj_->nframes();
  if (mode_ == SRMSD) {
    if (precision_ != RMSD_PREC_FP64) {
      g.SetError("rms2d: 'srmsd' is FP64 only");
      return ERR;
    }
    std::vector<int32_t> goff, gatoms;
    std::string err;
    if (!symm_groups_for(TgtTraj_->top, tsel, &goff, &gatoms, &err)) {
      g.SetError("rms2d: " + err);
      return ERR;
    }
    if (!tri && g.FramesOnDevice("rms2d.ref", *ref, rsel, &R)) return ERR;
    if (symm_rows(g, T, tri ? nullptr : &R, tri, useMass_ ? 1 : 0, goff, gatoms,
                  rmsdataset_) != OK)
      return ERR;
    if (tri && Ct_) {
      float* d = (float*)g.Buffer("rms2d.out", sizeof(float) * std::max<size_t>(1, rmsdataset_->size()));
      double* ct = (double*)g.Buffer("rms2d.corr", sizeof(double) * std::max(ntgt, 1));
      if (!d || !ct || g.Upload(d, rmsdataset_->data(), sizeof(float) * rmsdataset_->size()))
        return ERR;
      if (g.Check(rmsd_matrix_autocorr(g.ctx(), d, ntgt, ct), "rms2d corr")) return ERR;
      Ct_->v.resize(ntgt);
      if (g.Download(Ct_->v.data(), ct, sizeof(double) * ntgt)) return ERR;
    }
  } else if (tri) {
    rmsdataset_->AllocateTriangle(ntgt);
    float* d = (float*)g.Buffer("rms2d.out", sizeof(float) * std::max<size_t>(1, rmsdataset_->size()));
    if (!d) return ERR;
    if (g.Check(rmsd_pairwise(g.ctx(), &T, nullptr, mode, useMass_, precision_, d), "rms2d"))
      return ERR;
    if (Ct_) {
      double* ct = (double*)g.Buffer("rms2d.corr", sizeof(double) * std::max(ntgt, 1));
      if (!ct) return ERR;
      if (g.Check(rmsd_matrix_autocorr(g.ctx(), d, ntgt, ct), "rms2d corr")) return ERR;
      Ct_->v.resize(ntgt);
      if (g.Download(Ct_->v.data(), ct, sizeof(double) * ntgt)) return ERR;
    }
    if (g.Download(rmsdataset_->data(), d, sizeof(float) * rmsdataset_->size())) return ERR;
  } else {
    if (g.FramesOnDevice("rms2d.ref", *ref, rsel, &R)) return ERR;
    const int nref = ref->nframes();
    rmsdataset_->Allocate2D(nref, ntgt);
    float* d = (float*)g.Buffer("rms2d.out", sizeof(float) * std::max<size_t>(1, rmsdataset_->size()));
    if (!d) return ERR;
    if (g.Check(rmsd_pairwise(g.ctx(), &T, &R, mode, useMass_, precision_, d), "rms2d"))
      return ERR;
    if (g.Download(rmsdataset_->data(), d, sizeof(float) * rmsdataset_->size())) return ERR;
  }
  if (!outfile_.empty() && dataio::WriteMatrix(outfile_, *rmsdataset_, square2d_)) {
    g.SetError("cannot write " + outfile_);
    return ERR;
  }
  if (Ct_ && dataio::WriteSeries(corrfile_, {Ct_})) {
    g.SetError("cannot write " + corrfile_);
    return ERR;
  }
  return OK;
}

// ---- Action_Rmsd ------------------------------------------------------------------
// Init follows src/Action_Rmsd.cpp:39-189 and ReferenceAction::InitRef
// (src/ReferenceAction.cpp:72-130); 'perres' is another analysis and is
// rejected here.
RetType Action_Rmsd::Init(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                          GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  dsl_ = &dsl;
  // Per-residue options first (src/Action_Rmsd.cpp:35-50), so that their
  // values are not taken for masks.
  perres_ = args.hasKey("perres");
  if (perres_) {
    perresOut_ = args.GetStringKey("perresout");
    perresInvert_ = args.hasKey("perresinvert");
    tgtRange_ = args.GetStringKey("range");
    refRange_ = args.GetStringKey("refrange");
    perresMask_ = args.GetStringKey("perresmask");
    if (!perresMask_.empty() && perresMask_[0] == '&') perresMask_.erase(0, 1);
    perresCenter_ = args.hasKey("perrescenter");
    perresAvg_ = args.GetStringKey("perresavg");
  }
  fit_ = !args.hasKey("nofit");
  if (fit_) {
    if (args.hasKey("norotate")) mode_ = TRANS_ONLY;
    else if (args.hasKey("nomod")) mode_ = NONE;
  }
  useMass_ = args.hasKey("mass");
  std::string perr;
  if (!parse_prec(args, &precision_, &perr)) {
    gpu.SetError("rms: " + perr);
    return ERR;
  }
  if (precision_ != RMSD_PREC_FP64) {
    gpu.SetError("rms: the frame-to-reference fit is FP64 only");
    return ERR;
  }
  outfile_ = args.GetStringKey("out");
  // GPU path option: frames per streamed chunk (default: 256 MB of frames).
  chunkFrames_ = args.getKeyInt("chunk", 0);
  const bool saveMatrices = args.hasKey("savematrices");
  if (saveMatrices) matricesout_ = args.GetStringKey("matricesout");
  const std::string sv = args.GetStringKey("savevectors");
  if (sv.empty()) tvecType_ = NO_TVEC;
  else if (sv == "combined") tvecType_ = COMBINED;
  else if (sv == "separate") tvecType_ = SEPARATE;
  else {
    gpu.SetError("Expected 'combined' or 'separate' for 'savevectors'");
    return ERR;
  }
  if (tvecType_ != NO_TVEC) vecsout_ = args.GetStringKey("vecsout");
  if (InitRef(args, dsl, gpu) != OK) return ERR;
  tgtMaskExpr_ = args.GetMaskNext();
  refMaskExpr_ = args.GetMaskNext();
  if (refMaskExpr_.empty()) refMaskExpr_ = tgtMaskExpr_;
  std::string name = args.GetStringNext();
  if (name.empty()) name = dsl.GenerateDefaultName("RMSD");
  rmsd_ = dsl.AddSeries(name);
  rmsd_->width = 12;
  rmsd_->prec = 4;
  if (saveMatrices) {
    if (!fit_) {
      gpu.SetError("Must be fitting in order to save rotation matrices.");
      return ERR;
    }
    rmatrices_ = dsl.AddSeries(name + "[RM]");
    rmatrices_->ncomp = 9;
    rmatrices_->width = 12;
    rmatrices_->prec = 9;
  }
  if (tvecType_ != NO_TVEC) {
    if (!fit_) {
      gpu.SetError("Must be fitting in order to save translation vectors.");
      return ERR;
    }
    tvecs_ = dsl.AddSeries(name + "[TV]");
    tvecs_->ncomp = 6;
    tvecs_->width = 8;
    tvecs_->prec = 4;
  }
  return OK;
}

// ReferenceAction::InitRef (src/ReferenceAction.cpp:72-130): first (default),
// previous, reftraj <COORDS set>, ref <name> / reference.
// ReferenceAction::InitRef's modes (src/ReferenceAction.cpp:73-130): first
// (default), previous, reftraj <set>, ref <name> / reference, refindex <#>.
static RetType init_reference(ArgList& args, DataSetList& dsl, GpuSession& gpu,
                              Action_Rmsd::RefModeType* mode, CoordsSet** refCrd) {
  *mode = Action_Rmsd::FIRST;
  if (args.hasKey("previous")) {
    *mode = Action_Rmsd::PREVIOUS;
  } else if (!args.hasKey("first")) {
    if (args.Contains("reftraj")) {
      const std::string rt = args.GetStringKey("reftraj");
      *refCrd = dsl.FindCoordsSet(rt);
      if (!*refCrd) {
        gpu.SetError("reference trajectory '" + rt + "' is not a loaded COORDS set");
        return ERR;
      }
      *mode = Action_Rmsd::TRAJ;
    } else {
      std::string rn;
      bool want = false;
      if (args.Contains("ref")) {
        rn = args.GetStringKey("ref");
        want = true;
      } else if (args.hasKey("reference")) {
        want = true;
      }
      if (want) {
        *refCrd = dsl.FindReference(rn);
        if (!*refCrd) {
          gpu.SetError("reference '" + rn + "' not found");
          return ERR;
        }
        *mode = Action_Rmsd::FRAME;
      } else if (args.Contains("refindex")) {
        // 'refindex <#>': the reference frame at that position
        // (DataSetList::GetReferenceSet, src/DataSetList.cpp:1011-1014).
        const int idx = args.getKeyInt("refindex", -1);
        *refCrd = dsl.ReferenceAt(idx);
        if (!*refCrd) {
          gpu.SetError("no reference at index " + std::to_string(idx));
          return ERR;
        }
        *mode = Action_Rmsd::FRAME;
      }
    }
  }
  return OK;
}

RetType Action_Rmsd::InitRef(ArgList& args, DataSetList& dsl, GpuSession& gpu) {
  return init_reference(args, dsl, gpu, &refMode_, &refCrd_);
}

// Action_Rmsd::perResSetup (src/Action_Rmsd.cpp:197-315).  Target residues:
// 'range' or every solute residue; reference residues: 'refrange' or the same
// numbers.  Residue masks are ":<res>" intersected with 'perresmask'.  A
// residue out of range is skipped; one whose masks select nothing or differ in
// size keeps an (empty) data set but is not computed.
RetType Action_Rmsd::PerResSetup(const Topology& tgt, const CoordsSet& tgtset,
                                 const Topology& ref, const CoordsSet& refset) {
  GpuSession& g = *gpu_;
  std::string err;
  if (tgt.Nres() == 0 || ref.Nres() == 0) {
    g.SetError("perres: the topology has no residue information");
    return ERR;
  }
  std::vector<int> tr, rr;
  if (tgtRange_.empty()) {
    if (tgt.soluteRes.empty())
      for (int r = 0; r < tgt.Nres(); ++r) tr.push_back(r + 1);
    else
      for (int r : tgt.soluteRes) tr.push_back(r + 1);  // user numbering starts at 1
  } else if (!parse_range(tgtRange_, &tr, &err)) {
    g.SetError(err);
    return ERR;
  }
  if (refRange_.empty()) rr = tr;
  else if (!parse_range(refRange_, &rr, &err)) {
    g.SetError(err);
    return ERR;
  }
  if (tr.size() != rr.size()) {
    g.SetError("Number of target residues " + std::to_string(tr.size()) +
               " does not match number of reference residues " + std::to_string(rr.size()));
    return ERR;
  }
  // 'perresmask': one selection per topology, intersected with each residue.
  std::vector<char> tpick(tgt.natom, 1), rpick(ref.natom, 1);
  if (!perresMask_.empty()) {
    AtomMask ts, rs;
    if (!resolve_ || !resolve_(perresMask_, tgtset, &ts) || !resolve_(perresMask_, refset, &rs)) {
      g.SetError("perresmask [" + perresMask_ + "] cannot be resolved");
      return ERR;
    }
    std::fill(tpick.begin(), tpick.end(), 0);
    std::fill(rpick.begin(), rpick.end(), 0);
    for (int a : ts) tpick[a] = 1;
    for (int a : rs) rpick[a] = 1;
  }
  prOff_.assign(1, 0);
  prTgt_.clear();
  prRef_.clear();
  prTmass_.clear();
  prRmass_.clear();
  prSeries_.clear();
  int maxNatom = 0;
  for (size_t i = 0; i < tr.size(); ++i) {
    const int tres = tr[i], rres = rr[i];
    if (tres < 1 || tres > tgt.Nres() || rres < 1 || rres > ref.Nres()) continue;
    size_t p = 0;
    while (p < resRMS_.size() && resRMS_[p].tgtRes != tres) ++p;
    if (p == resRMS_.size()) {
      PerRes pr;
      pr.tgtRes = tres;
      pr.data = dsl_->AddSeries(rmsd_->name + "[res]:" + std::to_string(tres));
      const std::string& rn = (size_t)(tres - 1) < tgt.resName.size() ? tgt.resName[tres - 1] : "";
      pr.data->legend = rn + ":" + std::to_string(tres);
      pr.data->width = 12;
      pr.data->prec = 4;
      resRMS_.push_back(pr);
    }
    resRMS_[p].active = false;
    AtomMask ta, ra;
    for (int a = tgt.resFirst[tres - 1]; a < tgt.ResEnd(tres - 1); ++a)
      if (tpick[a]) ta.push_back(a);
    for (int a = ref.resFirst[rres - 1]; a < ref.ResEnd(rres - 1); ++a)
      if (rpick[a]) ra.push_back(a);
    if (ta.empty() || ra.empty() || ta.size() != ra.size()) continue;
    maxNatom = std::max(maxNatom, (int)ta.size());
    resRMS_[p].active = true;
    for (size_t k = 0; k < ta.size(); ++k) {
      prTgt_.push_back(ta[k]);
      prRef_.push_back(ra[k]);
      prTmass_.push_back((size_t)ta[k] < tgt.mass.size() ? tgt.mass[ta[k]] : 1.0);
      prRmass_.push_back((size_t)ra[k] < ref.mass.size() ? ref.mass[ra[k]] : 1.0);
    }
    prOff_.push_back((int32_t)prTgt_.size());
    prSeries_.push_back((int)p);
  }
  if (maxNatom == 0) {
    g.SetError("No residues selected for per-residue calculation.");
    return ERR;
  }
  perresReady_ = true;
  return OK;
}

// Frames f0 .. f0+nf-1 of crd against one reference frame (refxyz, whose
// selected atoms are the reference mask of reftop).  The frames stream
// through the device in chunks of at most kStreamBytes on two transfer slots
// (stream + library context each): while one chunk is fitted, the next one
// is uploaded and the previous one's results (and rewritten coordinates) come
// back.  One library call per chunk.
static const size_t kStreamBytes = size_t(256) << 20;

RetType Action_Rmsd::RunBatch(CoordsSet& crd, const AtomMask& tsel, int f0, int nf,
                              const float* refxyz, const Topology& reftop,
                              const std::string& refname) {
  GpuSession& g = *gpu_;
  AtomMask rsel;
  CoordsSet refset;
  refset.name = refname;
  refset.top = reftop;
  refset.stride = 3 * (int64_t)reftop.natom;
  refset.xyz.assign(refxyz, refxyz + refset.stride);
  if (!resolve_or_all(resolve_, refMaskExpr_, refset, &rsel) || rsel.empty()) {
    g.SetError("No reference atoms selected for [" + refMaskExpr_ + "]");
    return ERR;
  }
  if (rsel.size() != tsel.size()) {
    g.SetError("Number of atoms in target mask does not equal number of atoms in reference mask");
    return ERR;
  }
  // Reference, target selection and masses: uploaded once (synchronously).
  rmsd_frames R, T0;
  if (g.FramesOnDevice("rms.ref", refset, rsel, &R)) return ERR;
  if (g.FramesOnDevice("rms.tsel", crd, tsel, &T0, crd.frame(f0), 0)) return ERR;
  const int modify = !fit_ ? RMSD_MODIFY_NONE
                           : (mode_ == ROT_AND_TRANS ? RMSD_MODIFY_ROT_TRANS
                                                     : (mode_ == TRANS_ONLY ? RMSD_MODIFY_TRANS
                                                                            : RMSD_MODIFY_NONE));
  // 'align ... move <mask>': fit without update, then move the listed atoms.
  const bool moving = !moveMaskExpr_.empty();
  int32_t* dmove = nullptr;
  int nmove = 0;
  if (moving) {
    AtomMask mv;
    if (!resolve_or_all(resolve_, moveMaskExpr_, crd, &mv) || mv.empty()) {
      g.SetError("No atoms selected for move mask [" + moveMaskExpr_ + "]");
      return ERR;
    }
    for (int a : mv)
      if (a < 0 || a >= crd.top.natom) {
        g.SetError("move mask atom out of range");
        return ERR;
      }
    std::vector<int32_t> m32(mv.begin(), mv.end());
    nmove = (int)m32.size();
    dmove = (int32_t*)g.Buffer("rms.move", sizeof(int32_t) * nmove);
    if (!dmove || g.Upload(dmove, m32.data(), sizeof(int32_t) * nmove)) return ERR;
  }
  const size_t fbytes = sizeof(float) * (size_t)crd.stride;
  int chunk = (int)std::max<size_t>(1, std::min<size_t>(nf, kStreamBytes / fbytes));
  if (chunkFrames_ > 0) chunk = std::min(chunk, chunkFrames_);
  const int nchunks = (nf + chunk - 1) / chunk;
  const size_t base_r = rmsd_->v.size();
  rmsd_->v.resize(base_r + nf);
  const size_t base_u = rmatrices_ ? rmatrices_->v.size() : 0;
  if (rmatrices_) rmatrices_->v.resize(base_u + 9 * (size_t)nf);
  const size_t base_t = tvecs_ ? tvecs_->v.size() : 0;
  if (tvecs_) tvecs_->v.resize(base_t + 6 * (size_t)nf);

  // Per-residue values: the fit runs without its coordinate update, the
  // residues are measured on the frames as that update leaves them (FP64),
  // then the update is applied.
  const bool pr = perresReady_;
  const int prn = pr ? (int)prSeries_.size() : 0;
  // Box::RotateUcell after the rotate + translate update (src/Action_Rmsd.cpp:
  // 381; Action_Align::DoAction, src/Action_Align.cpp:98).
  const bool rotbox = fit_ && (modify == RMSD_MODIFY_ROT_TRANS || moving) &&
                      crd.box.size() == 9 * (size_t)crd.nframes() && !crd.box.empty();
  const bool needUT = rmatrices_ || moving || (pr && fit_) || rotbox;
  if (pr)
    for (int j = 0; j < prn; ++j) resRMS_[prSeries_[j]].data->v.resize(base_r + nf);
  struct SlotBufs {
    rmsd_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    float* dx = nullptr;
    double *dr = nullptr, *dU = nullptr, *dT = nullptr, *dpr = nullptr, *dbox = nullptr;
    PinnedDoubles hr, hU, hT, hpr;
    int c = -1;  // chunk in flight
  } sl[2];
  for (int k = 0; k < 2 && k < nchunks; ++k) {
    void* st = nullptr;
    if (g.Slot(k, &sl[k].ctx, &st)) return ERR;
    sl[k].st = (hipStream_t)st;
    const std::string t = "rms.slot" + std::to_string(k);
    sl[k].dx = (float*)g.Buffer(t + ".xyz", fbytes * chunk);
    sl[k].dr = (double*)g.Buffer(t + ".rmsd", sizeof(double) * chunk);
    if (needUT) sl[k].dU = (double*)g.Buffer(t + ".U", sizeof(double) * 9 * chunk);
    if (tvecs_ || needUT) sl[k].dT = (double*)g.Buffer(t + ".T", sizeof(double) * 6 * chunk);
    if (pr) sl[k].dpr = (double*)g.Buffer(t + ".perres", sizeof(double) * (size_t)prn * chunk);
    if (rotbox) sl[k].dbox = (double*)g.Buffer(t + ".box", sizeof(double) * 9 * chunk);
    if (!sl[k].dx || !sl[k].dr || (needUT && !sl[k].dU) || ((tvecs_ || needUT) && !sl[k].dT) ||
        (pr && !sl[k].dpr) || (rotbox && !sl[k].dbox))
      return ERR;
    sl[k].hr.resize(chunk);
    if (pr) sl[k].hpr.resize((size_t)prn * chunk);
    if (rmatrices_) sl[k].hU.resize(9 * (size_t)chunk);
    if (tvecs_) sl[k].hT.resize(6 * (size_t)chunk);
  }
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    g.SetError(std::string(what) + ": " + hipGetErrorString(e));
    return false;
  };
  // Waits for a slot's chunk and moves its results into the data sets.
  auto harvest = [&](SlotBufs& b) -> bool {
    if (b.c < 0) return true;
    if (!hip_ok(hipStreamSynchronize(b.st), "rms stream")) return false;
    const int c0 = b.c * chunk, n = std::min(chunk, nf - c0);
    std::copy(b.hr.begin(), b.hr.begin() + n, rmsd_->v.begin() + base_r + c0);
    for (int j = 0; j < prn; ++j) {
      std::vector<double>& v = resRMS_[prSeries_[j]].data->v;
      for (int f = 0; f < n; ++f) v[base_r + c0 + f] = b.hpr[(size_t)f * prn + j];
    }
    if (rmatrices_)
      std::copy(b.hU.begin(), b.hU.begin() + 9 * (size_t)n,
                rmatrices_->v.begin() + base_u + 9 * (size_t)c0);
    if (tvecs_)
      for (int f = 0; f < n; ++f) {
        double* o = tvecs_->v.data() + base_t + 6 * (size_t)(c0 + f);
        for (int c = 0; c < 3; ++c) {
          const double tr = b.hT[6 * f + c], rf = b.hT[6 * f + 3 + c];
          if (tvecType_ == COMBINED) {
            // Vxyz = Trans + refTrans; the origin columns stay 0.
            o[c] = tr + rf;
            o[3 + c] = 0.0;
          } else {
            o[c] = tr;
            o[3 + c] = rf;
          }
        }
      }
    b.c = -1;
    return true;
  };
  for (int c = 0; c < nchunks; ++c) {
    SlotBufs& b = sl[c & 1];
    if (!harvest(b)) return ERR;
    const int c0 = c * chunk, n = std::min(chunk, nf - c0);
    float* host = crd.frame(f0 + c0);
    if (!hip_ok(hipMemcpyAsync(b.dx, host, fbytes * n, hipMemcpyHostToDevice, b.st), "upload"))
      return ERR;
    rmsd_frames F = T0;
    F.xyz = b.dx;
    F.nframes = n;
    if (g.Check(rmsd_fit_to_ref(b.ctx, &F, &R, 0, fit_, useMass_,
                                (moving || pr) ? RMSD_MODIFY_NONE : modify, b.dr, b.dU, b.dT),
                "rms"))
      return ERR;
    if (pr) {
      if (g.Check(rmsd_perres_nofit(b.ctx, &F, &R, 0, prn, prOff_.data(), prTgt_.data(),
                                    prRef_.data(), prTmass_.data(), prRmass_.data(), useMass_,
                                    perresCenter_, modify, b.dU, b.dT, b.dpr),
                  "rms perres"))
        return ERR;
      if (modify != RMSD_MODIFY_NONE &&
          g.Check(rmsd_transform(b.ctx, &F, nullptr, 0, b.dU, b.dT, modify), "rms"))
        return ERR;
    }
    if (moving && g.Check(rmsd_transform(b.ctx, &F, dmove, nmove, b.dU, b.dT,
                                         RMSD_MODIFY_ROT_TRANS),
                          "align"))
      return ERR;
    if (rotbox) {
      double* hbox = crd.box.data() + 9 * (size_t)(f0 + c0);
      if (!hip_ok(hipMemcpyAsync(b.dbox, hbox, sizeof(double) * 9 * n, hipMemcpyHostToDevice,
                                 b.st),
                  "upload") ||
          g.Check(rmsd_rotate_ucell(b.ctx, b.dU, n, b.dbox), "rms box") ||
          !hip_ok(hipMemcpyAsync(hbox, b.dbox, sizeof(double) * 9 * n, hipMemcpyDeviceToHost,
                                 b.st),
                  "download"))
        return ERR;
    }
    if (pr && !hip_ok(hipMemcpyAsync(b.hpr.data(), b.dpr, sizeof(double) * (size_t)prn * n,
                                     hipMemcpyDeviceToHost, b.st),
                      "download"))
      return ERR;
    if (!hip_ok(hipMemcpyAsync(b.hr.data(), b.dr, sizeof(double) * n, hipMemcpyDeviceToHost, b.st),
                "download") ||
        (rmatrices_ && !hip_ok(hipMemcpyAsync(b.hU.data(), b.dU, sizeof(double) * 9 * n,
                                              hipMemcpyDeviceToHost, b.st),
                               "download")) ||
        (tvecs_ && !hip_ok(hipMemcpyAsync(b.hT.data(), b.dT, sizeof(double) * 6 * n,
                                          hipMemcpyDeviceToHost, b.st),
                           "download")) ||
        ((modify != RMSD_MODIFY_NONE || moving) &&
         !hip_ok(hipMemcpyAsync(host, b.dx, fbytes * n, hipMemcpyDeviceToHost, b.st), "download")))
      return ERR;
    b.c = c;
  }
  return harvest(sl[0]) && harvest(sl[1]) ? OK : ERR;
}

// Every frame against its own reference (ReferenceAction::ActionRef /
// PreviousRef, src/ReferenceAction.h:73-88) in chunks: one rmsd_fit_pairs
// call per chunk.  'previous' fits the frames as stored (frame f against f - 1)
// and composes the updates on the host: the reference of frame f is frame
// f - 1 after its update G_(f-1) (a rigid motion, which leaves the RMSD
// unchanged), so the fit onto it has U_f = L_(f-1) U_f and refTrans_f =
// L_(f-1) (refTrans_f + Trans_(f-1)) + refTrans'_(f-1), L the applied
// rotation (identity for the translate-only update).
RetType Action_Rmsd::RunPairs(CoordsSet& crd, const AtomMask& tsel, CoordsSet& refset,
                              bool previous) {
  GpuSession& g = *gpu_;
  const int nf = crd.nframes();
  AtomMask rsel;
  if (!resolve_or_all(resolve_, refMaskExpr_, refset, &rsel) || rsel.empty()) {
    g.SetError("No reference atoms selected for [" + refMaskExpr_ + "]");
    return ERR;
  }
  if (rsel.size() != tsel.size()) {
    g.SetError("Number of atoms in target mask does not equal number of atoms in reference mask");
    return ERR;
  }
  if (!previous && refset.nframes() < nf) {
    g.SetError("reference trajectory has fewer frames than the target");
    return ERR;
  }
  const int modify = !fit_ ? RMSD_MODIFY_NONE
                           : (mode_ == ROT_AND_TRANS ? RMSD_MODIFY_ROT_TRANS
                                                     : (mode_ == TRANS_ONLY ? RMSD_MODIFY_TRANS
                                                                            : RMSD_MODIFY_NONE));
  const bool compose = previous && modify != RMSD_MODIFY_NONE;
  const bool rotbox = fit_ && modify == RMSD_MODIFY_ROT_TRANS &&
                      crd.box.size() == 9 * (size_t)crd.nframes() && !crd.box.empty();
  const bool needU = fit_ && (rmatrices_ || compose || rotbox);
  const bool needT = fit_ && (tvecs_ || compose);
  const size_t fbytes = sizeof(float) * (size_t)crd.stride;
  int chunk = (int)std::max<size_t>(1, std::min<size_t>(nf, kStreamBytes / fbytes));
  if (chunkFrames_ > 0) chunk = std::min(chunk, chunkFrames_);
  const size_t base_r = rmsd_->v.size();
  rmsd_->v.resize(base_r + nf);
  const size_t base_u = rmatrices_ ? rmatrices_->v.size() : 0;
  if (rmatrices_) rmatrices_->v.resize(base_u + 9 * (size_t)nf);
  const size_t base_t = tvecs_ ? tvecs_->v.size() : 0;
  if (tvecs_) tvecs_->v.resize(base_t + 6 * (size_t)nf);
  double* dr = (double*)g.Buffer("pairs.rmsd", sizeof(double) * chunk);
  double* dU = needU ? (double*)g.Buffer("pairs.U", sizeof(double) * 9 * chunk) : nullptr;
  double* dT = needT ? (double*)g.Buffer("pairs.T", sizeof(double) * 6 * chunk) : nullptr;
  double* dbox = rotbox ? (double*)g.Buffer("pairs.box", sizeof(double) * 9 * chunk) : nullptr;
  if (!dr || (needU && !dU) || (needT && !dT) || (rotbox && !dbox)) return ERR;
  std::vector<double> hr(chunk), hU(needU ? 9 * (size_t)chunk : 0),
      hT(needT ? 6 * (size_t)chunk : 0);
  // 'previous': frame f0 - 1 as stored, before this action moved it, ahead of
  // the chunk's frames in one upload (for the first chunk frame 0 itself, its
  // own reference).
  std::vector<float> stage;
  std::vector<float> prevOrig(crd.frame(0), crd.frame(0) + crd.stride);
  double Lp[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tp[3] = {0, 0, 0}, rp[3] = {0, 0, 0};
  for (int f0 = 0; f0 < nf; f0 += chunk) {
    const int n = std::min(chunk, nf - f0);
    rmsd_frames T, R;
    if (previous) {
      stage.resize((size_t)(n + 1) * crd.stride);
      std::copy(prevOrig.begin(), prevOrig.end(), stage.begin());
      std::copy(crd.frame(f0), crd.frame(f0) + (size_t)n * crd.stride,
                stage.begin() + crd.stride);
      if (g.FramesOnDevice("pairs.prev", crd, rsel, &R, stage.data(), n + 1)) return ERR;
      std::copy(crd.frame(f0 + n - 1), crd.frame(f0 + n - 1) + crd.stride, prevOrig.begin());
      // targets: the same upload, one frame on (their own selection / masses)
      if (g.FramesOnDevice("pairs.tsel", crd, tsel, &T, crd.frame(f0), 0)) return ERR;
      T.xyz = R.xyz + crd.stride;
      T.nframes = n;
      R.nframes = n;
    } else {
      if (g.FramesOnDevice("pairs.tgt", crd, tsel, &T, crd.frame(f0), n) ||
          g.FramesOnDevice("pairs.ref", refset, rsel, &R, refset.frame(f0), n))
        return ERR;
    }
    if (g.Check(rmsd_fit_pairs(g.ctx(), &T, &R, fit_, useMass_,
                               compose ? RMSD_MODIFY_NONE : modify, dr, dU, dT),
                "rms"))
      return ERR;
    if (g.Download(hr.data(), dr, sizeof(double) * n) ||
        (needU && g.Download(hU.data(), dU, sizeof(double) * 9 * n)) ||
        (needT && g.Download(hT.data(), dT, sizeof(double) * 6 * n)))
      return ERR;
    if (compose) {
      for (int f = 0; f < n; ++f) {
        double* U = hU.data() + 9 * (size_t)f;
        double* t = hT.data() + 6 * (size_t)f;
        if (f0 + f > 0) {
          double Uc[9], rc[3];
          const double* L = modify == RMSD_MODIFY_ROT_TRANS ? Lp : nullptr;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              Uc[3 * i + j] = L ? L[3 * i] * U[j] + L[3 * i + 1] * U[3 + j] + L[3 * i + 2] * U[6 + j]
                                : U[3 * i + j];
          const double v[3] = {t[3] + tp[0], t[4] + tp[1], t[5] + tp[2]};
          for (int i = 0; i < 3; ++i)
            rc[i] = (L ? L[3 * i] * v[0] + L[3 * i + 1] * v[1] + L[3 * i + 2] * v[2] : v[i]) +
                    rp[i];
          std::copy(Uc, Uc + 9, U);
          std::copy(rc, rc + 3, t + 3);
        }
        std::copy(U, U + 9, Lp);
        std::copy(t, t + 3, tp);
        std::copy(t + 3, t + 6, rp);
      }
      rmsd_frames A = T;
      A.fidx = nullptr;
      if (g.Upload(dU, hU.data(), sizeof(double) * 9 * n) ||
          g.Upload(dT, hT.data(), sizeof(double) * 6 * n) ||
          g.Check(rmsd_transform(g.ctx(), &A, nullptr, 0, dU, dT, modify), "rms"))
        return ERR;
    }
    if (rotbox) {
      double* hbox = crd.box.data() + 9 * (size_t)f0;
      if (g.Upload(dbox, hbox, sizeof(double) * 9 * n) ||
          g.Check(rmsd_rotate_ucell(g.ctx(), dU, n, dbox), "rms box") ||
          g.Download(hbox, dbox, sizeof(double) * 9 * n))
        return ERR;
    }
    if (modify != RMSD_MODIFY_NONE &&
        g.Download(crd.frame(f0), T.xyz, fbytes * n))
      return ERR;
    std::copy(hr.begin(), hr.begin() + n, rmsd_->v.begin() + base_r + f0);
    if (rmatrices_)
      std::copy(hU.begin(), hU.begin() + 9 * (size_t)n,
                rmatrices_->v.begin() + base_u + 9 * (size_t)f0);
    if (tvecs_)
      for (int f = 0; f < n; ++f) {
        double* o = tvecs_->v.data() + base_t + 6 * (size_t)(f0 + f);
        for (int c = 0; c < 3; ++c) {
          const double tr = hT[6 * f + c], rf = hT[6 * f + 3 + c];
          if (tvecType_ == COMBINED) {
            o[c] = tr + rf;
            o[3 + c] = 0.0;
          } else {
            o[c] = tr;
            o[3 + c] = rf;
          }
        }
      }
  }
  return OK;
}

// Action_Align::Init (src/Action_Align.cpp): 'align <mask> [<refmask>]
// [move <mask>] [mass]' plus the reference keywords; always fits and moves
// (all atoms, or those of the move mask).  Its RMSD values are kept in a data
// set as a by-product.
RetType Action_Rmsd::InitAlign(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                               GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  fit_ = true;
  mode_ = ROT_AND_TRANS;
  useMass_ = args.hasKey("mass");
  chunkFrames_ = args.getKeyInt("chunk", 0);
  if (InitRef(args, dsl, gpu) != OK) return ERR;
  moveMaskExpr_ = args.GetStringKey("move");
  tgtMaskExpr_ = args.GetMaskNext();
  refMaskExpr_ = args.GetMaskNext();
  if (refMaskExpr_.empty()) refMaskExpr_ = tgtMaskExpr_;
  std::string name = args.GetStringNext();
  if (name.empty()) name = dsl.GenerateDefaultName("Align");
  rmsd_ = dsl.AddSeries(name);
  return OK;
}

RetType Action_Rmsd::DoFrames(CoordsSet& crd) {
  GpuSession& g = *gpu_;
  AtomMask tsel;
  if (!resolve_or_all(resolve_, tgtMaskExpr_, crd, &tsel) || tsel.empty()) {
    g.SetError("No atoms selected for [" + tgtMaskExpr_ + "]");
    return ERR;
  }
  if (perres_ && !perresReady_) {
    const CoordsSet& rs = (refMode_ == FRAME || refMode_ == TRAJ) ? *refCrd_ : crd;
    if (PerResSetup(crd.top, crd, rs.top, rs) != OK) return ERR;
  }
  const int nf = crd.nframes();
  if (nf == 0) return OK;
  switch (refMode_) {
    case FIRST: {
      // Reference = the first frame as read, before this action moves it.
      std::vector<float> first(crd.frame(0), crd.frame(0) + crd.stride);
      return RunBatch(crd, tsel, 0, nf, first.data(), crd.top, crd.name);
    }
    case FRAME:
      return RunBatch(crd, tsel, 0, nf, refCrd_->frame(0), refCrd_->top, refCrd_->name);
    case TRAJ:
      // Frame i against frame i of the reference trajectory
      // (ReferenceAction::ActionRef, src/ReferenceAction.h:72-84).
      if (refCrd_->nframes() < nf) {
        g.SetError("reference trajectory has fewer frames than the target");
        return ERR;
      }
      if (!perresReady_ && moveMaskExpr_.empty()) return RunPairs(crd, tsel, *refCrd_, false);
      for (int f = 0; f < nf; ++f)
        if (RunBatch(crd, tsel, f, 1, refCrd_->frame(f), refCrd_->top, refCrd_->name) != OK)
          return ERR;
      return OK;
    case PREVIOUS: {
      // Reference = the previous frame after this action's update of it
      // (ReferenceAction::PreviousRef, src/ReferenceAction.h:86-88).
      if (!perresReady_ && moveMaskExpr_.empty()) return RunPairs(crd, tsel, crd, true);
      std::vector<float> prev(crd.frame(0), crd.frame(0) + crd.stride);
      for (int f = 0; f < nf; ++f) {
        if (RunBatch(crd, tsel, f, 1, prev.data(), crd.top, crd.name) != OK) return ERR;
        std::copy(crd.frame(f), crd.frame(f) + crd.stride, prev.begin());
      }
      return OK;
    }
  }
  return ERR;
}

int Action_Rmsd::WriteOutput() {
  if (!outfile_.empty() && dataio::WriteSeries(outfile_, {rmsd_})) return 1;
  if (rmatrices_ && !matricesout_.empty() && dataio::WriteSeries(matricesout_, {rmatrices_}))
    return 1;
  if (tvecs_ && !vecsout_.empty() && dataio::WriteSeries(vecsout_, {tvecs_})) return 1;
  if (!perres_ || resRMS_.empty()) return 0;
  std::vector<const Series*> sets;
  for (const PerRes& p : resRMS_) sets.push_back(p.data);
  if (!perresOut_.empty() && dataio::WriteGrace(perresOut_, sets, perresInvert_)) return 1;
  if (!perresAvg_.empty()) {
    // DataSet_1D::Avg (src/DataSet_1D.cpp:13-61): mean and population
    // standard deviation; x = the residue number.
    Series* avg = dsl_->AddSeries(rmsd_->name + "[Avg]");
    Series* sd = dsl_->AddSeries(rmsd_->name + "[Stdev]");
    std::vector<double> x;
    for (const PerRes& p : resRMS_) {
      const std::vector<double>& v = p.data->v;
      double a = 0.0, d = 0.0;
      if (!v.empty()) {
        for (double y : v) a += y;
        a /= (double)v.size();
        for (double y : v) d += (a - y) * (a - y);
        d = std::sqrt(d / (double)v.size());
      }
      x.push_back((double)p.tgtRes);
      avg->v.push_back(a);
      sd->v.push_back(d);
    }
    if (dataio::WriteMesh(perresAvg_, "Residue", x, {avg, sd})) return 1;
  }
  return 0;
}

// ---- Action_SymmetricRmsd ------------------------------------------------------------
// Init follows src/Action_SymmetricRmsd.cpp:17-58, DoAction :88-112.
RetType Action_SymmetricRmsd::Init(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                                   GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  fit_ = !args.hasKey("nofit");
  useMass_ = args.hasKey("mass");
  outfile_ = args.GetStringKey("out");
  remap_ = args.hasKey("remap");
  if (init_reference(args, dsl, gpu, &refMode_, &refCrd_) != OK) return ERR;
  tgtMaskExpr_ = args.GetMaskNext();
  refMaskExpr_ = args.GetMaskNext();
  if (refMaskExpr_.empty()) refMaskExpr_ = tgtMaskExpr_;
  std::string name = args.GetStringNext();
  if (name.empty()) name = dsl.GenerateDefaultName("RMSD");
  rmsd_ = dsl.AddSeries(name);
  rmsd_->width = 12;
  rmsd_->prec = 4;
  return OK;
}

RetType Action_SymmetricRmsd::DoFrames(CoordsSet& crd) {
  GpuSession& g = *gpu_;
  AtomMask tsel;
  if (!resolve_or_all(resolve_, tgtMaskExpr_, crd, &tsel) || tsel.empty()) {
    g.SetError("No atoms selected for [" + tgtMaskExpr_ + "]");
    return ERR;
  }
  // SymmetricRmsdCalc::SetupSymmRMSD: groups of residues up to the last
  // selected atom's residue, as selection indices; groups with no selected
  // atom are dropped.
  const Topology& top = crd.top;
  std::vector<int> selIdx((size_t)top.natom, -1);
  for (size_t k = 0; k < tsel.size(); ++k) selIdx[tsel[k]] = (int)k;
  int lastRes = top.Nres();
  if (top.Nres() > 0) {
    const int lastAtom = *std::max_element(tsel.begin(), tsel.end());
    lastRes = (int)(std::upper_bound(top.resFirst.begin(), top.resFirst.end(), lastAtom) -
                    top.resFirst.begin());
  }
  goff_.assign(1, 0);
  gatoms_.clear();
  for (const std::vector<int>& grp : top.symmGroups) {
    if (grp.empty()) continue;
    if (top.Nres() > 0) {
      const int r = (int)(std::upper_bound(top.resFirst.begin(), top.resFirst.end(), grp[0]) -
                          top.resFirst.begin()) - 1;
      if (r >= lastRes) continue;
    }
    size_t before = gatoms_.size();
    for (int a : grp) {
      if (a < 0 || a >= top.natom) {
        g.SetError("symmetric group atom out of range of the topology");
        return ERR;
      }
      if (selIdx[a] >= 0) gatoms_.push_back(selIdx[a]);
    }
    if (gatoms_.size() > before) goff_.push_back((int32_t)gatoms_.size());
  }
  const int nf = crd.nframes();
  if (nf == 0) return OK;
  switch (refMode_) {
    case Action_Rmsd::FIRST: {
      std::vector<float> first(crd.frame(0), crd.frame(0) + crd.stride);
      return Run(crd, tsel, 0, nf, first.data(), crd.top, crd.name);
    }
    case Action_Rmsd::FRAME:
      return Run(crd, tsel, 0, nf, refCrd_->frame(0), refCrd_->top, refCrd_->name);
    case Action_Rmsd::TRAJ:
      if (refCrd_->nframes() < nf) {
        g.SetError("reference trajectory has fewer frames than the target");
        return ERR;
      }
      for (int f = 0; f < nf; ++f)
        if (Run(crd, tsel, f, 1, refCrd_->frame(f), refCrd_->top, refCrd_->name) != OK)
          return ERR;
      return OK;
    case Action_Rmsd::PREVIOUS: {
      std::vector<float> prev(crd.frame(0), crd.frame(0) + crd.stride);
      for (int f = 0; f < nf; ++f) {
        if (Run(crd, tsel, f, 1, prev.data(), crd.top, crd.name) != OK) return ERR;
        std::copy(crd.frame(f), crd.frame(f) + crd.stride, prev.begin());
      }
      return OK;
    }
  }
  return ERR;
}

RetType Action_SymmetricRmsd::Run(CoordsSet& crd, const AtomMask& tsel, int f0, int nf,
                                  const float* refxyz, const Topology& reftop,
                                  const std::string& refname) {
  GpuSession& g = *gpu_;
  AtomMask rsel;
  CoordsSet refset;
  refset.name = refname;
  refset.top = reftop;
  refset.stride = 3 * (int64_t)reftop.natom;
  refset.xyz.assign(refxyz, refxyz + refset.stride);
  if (!resolve_or_all(resolve_, refMaskExpr_, refset, &rsel) || rsel.empty()) {
    g.SetError("No reference atoms selected for [" + refMaskExpr_ + "]");
    return ERR;
  }
  if (rsel.size() != tsel.size()) {
    g.SetError("Number of atoms in target mask does not equal number of atoms in reference mask");
    return ERR;
  }
  rmsd_frames R, T;
  if (g.FramesOnDevice("symm.ref", refset, rsel, &R)) return ERR;
  if (g.FramesOnDevice("symm.tgt", crd, tsel, &T, crd.frame(f0), nf)) return ERR;
  const int n = (int)tsel.size();
  double* dr = (double*)g.Buffer("symm.rmsd", sizeof(double) * nf);
  int32_t* dm = remap_ ? (int32_t*)g.Buffer("symm.amap", sizeof(int32_t) * (size_t)nf * n) : nullptr;
  double* dU = fit_ ? (double*)g.Buffer("symm.U", sizeof(double) * 9 * nf) : nullptr;
  double* dT = fit_ ? (double*)g.Buffer("symm.T", sizeof(double) * 6 * nf) : nullptr;
  if (!dr || (remap_ && !dm) || (fit_ && (!dU || !dT))) return ERR;
  const int ng = (int)goff_.size() - 1;
  if (g.Check(rmsd_symm(g.ctx(), &T, &R, 0, ng, goff_.data(), gatoms_.empty() ? nullptr
                                                                                : gatoms_.data(),
                        fit_ ? 1 : 0, useMass_ ? 1 : 0, dr, dm, dU, dT),
              "rmsd_symm"))
    return ERR;
  std::vector<double> hr(nf), hU(fit_ ? 9 * (size_t)nf : 0), hT(fit_ ? 6 * (size_t)nf : 0);
  std::vector<int32_t> hm(remap_ ? (size_t)nf * n : 0);
  if (g.Download(hr.data(), dr, sizeof(double) * nf) ||
      (remap_ && g.Download(hm.data(), dm, sizeof(int32_t) * hm.size())) ||
      (fit_ && (g.Download(hU.data(), dU, sizeof(double) * hU.size()) ||
                g.Download(hT.data(), dT, sizeof(double) * hT.size()))))
    return ERR;
  rmsd_->v.insert(rmsd_->v.end(), hr.begin(), hr.end());
  if (!remap_ && !fit_) return OK;
  // The frame as the action leaves it: remapped (targetMap over all atoms,
  // Frame::SetCoordinatesByMap), then Trans_Rot_Trans(Trans, U, refTrans)
  // in FP64 (src/Action_SymmetricRmsd.cpp:96-108).
  const int natom = crd.top.natom;
  std::vector<double> x(3 * (size_t)natom);
  std::vector<int> tmap((size_t)natom);
  for (int f = 0; f < nf; ++f) {
    float* fr = crd.frame(f0 + f);
    for (int a = 0; a < natom; ++a) tmap[a] = a;
    if (remap_)
      for (int i = 0; i < n; ++i) tmap[tsel[i]] = tsel[hm[(size_t)f * n + i]];
    for (int a = 0; a < natom; ++a)
      for (int c = 0; c < 3; ++c) x[3 * a + c] = (double)fr[3 * tmap[a] + c];
    if (fit_) {
      const double* u = &hU[9 * (size_t)f];
      const double* t = &hT[6 * (size_t)f];
      for (int a = 0; a < natom; ++a) {
        const double px = x[3 * a] + t[0], py = x[3 * a + 1] + t[1], pz = x[3 * a + 2] + t[2];
        x[3 * a] = (px * u[0] + py * u[1] + pz * u[2]) + t[3];
        x[3 * a + 1] = (px * u[3] + py * u[4] + pz * u[5]) + t[4];
        x[3 * a + 2] = (px * u[6] + py * u[7] + pz * u[8]) + t[5];
      }
    }
    for (size_t k = 0; k < x.size(); ++k) fr[k] = (float)x[k];
  }
  return OK;
}

int Action_SymmetricRmsd::WriteOutput() const {
  if (!outfile_.empty() && dataio::WriteSeries(outfile_, {rmsd_})) return 1;
  return 0;
}

// ---- Analysis_RmsAvgCorr ---------------------------------------------------------------
// Setup follows src/Analysis_RmsAvgCorr.cpp:33-116.
RetType Analysis_RmsAvgCorr::Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                                   GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  const std::string setname = args.GetStringKey("crdset");
  coords_ = dsl.FindCoordsSet(setname);
  if (!coords_) {
    gpu.SetError("rmsavgcorr: Could not locate COORDS set corresponding to " + setname);
    return ERR;
  }
  lagOffset_ = std::max(1, args.getKeyInt("offset", 1));
  outfile_ = args.GetStringKey("out");
  separate_ = args.GetStringKey("output");
  useMass_ = args.hasKey("mass");
  maxwindow_ = args.getKeyInt("stop", -1);
  bool useFirst = args.hasKey("first");
  // DataSetList::GetReferenceFrame: 'ref <name>', 'reference', 'refindex <#>'.
  if (args.Contains("ref")) {
    const std::string rn = args.GetStringKey("ref");
    ref_ = dsl.FindReference(rn);
    if (!ref_) {
      gpu.SetError("reference '" + rn + "' not found");
      return ERR;
    }
  } else if (args.hasKey("reference")) {
    ref_ = dsl.FindReference("");
    if (!ref_) {
      gpu.SetError("no reference frame loaded");
      return ERR;
    }
  } else if (args.Contains("refindex")) {
    const int idx = args.getKeyInt("refindex", -1);
    ref_ = dsl.ReferenceAt(idx);
    if (!ref_) {
      gpu.SetError("no reference at index " + std::to_string(idx));
      return ERR;
    }
  }
  // Without a reference: the first running-averaged frame of each window;
  // 'first' cannot be combined with a reference (the reference wins).
  if (ref_) useFirst = false;
  (void)useFirst;
  tgtMaskExpr_ = args.GetMaskNext();
  if (ref_) {
    refMaskExpr_ = args.GetMaskNext();
    if (refMaskExpr_.empty()) refMaskExpr_ = tgtMaskExpr_;
  }
  std::string name = args.GetStringNext();
  if (name.empty()) name = dsl.GenerateDefaultName("RACorr");
  Ct_ = dsl.AddSeries(name);
  Csd_ = dsl.AddSeries(name + "[SD]");
  for (Series* s : {Ct_, Csd_}) {
    s->xmin = 1.0;
    s->xstep = (double)lagOffset_;
    s->width = 12;
    s->prec = 4;
  }
  return OK;
}

// Analyze (src/Analysis_RmsAvgCorr.cpp:119-276): every window size in one
// library call; window 1 (no averaging) first, then 1 + offset, ...
RetType Analysis_RmsAvgCorr::Analyze() {
  GpuSession& g = *gpu_;
  AtomMask tsel, rsel;
  if (!resolve_or_all(resolve_, tgtMaskExpr_, *coords_, &tsel) || tsel.empty()) {
    g.SetError("No atoms selected for [" + tgtMaskExpr_ + "]");
    return ERR;
  }
  rmsd_frames T, R;
  if (g.FramesOnDevice("avgcorr.tgt", *coords_, tsel, &T)) return ERR;
  if (ref_) {
    if (!resolve_or_all(resolve_, refMaskExpr_, *ref_, &rsel) || rsel.empty()) {
      g.SetError("Could not set up reference mask [" + refMaskExpr_ + "]");
      return ERR;
    }
    if (rsel.size() != tsel.size()) {
      g.SetError("Target mask " + tgtMaskExpr_ + " (" + std::to_string(tsel.size()) +
                 ") does not correspond to reference mask (" + std::to_string(rsel.size()) + ")");
      return ERR;
    }
    if (g.FramesOnDevice("avgcorr.ref", *ref_, rsel, &R, ref_->frame(0), 1)) return ERR;
  }
  const int nw = rmsd_avgcorr_windows(T.nframes, maxwindow_, lagOffset_, nullptr, 0);
  if (nw < 1) {
    g.SetError("Not enough frames to perform calculation.");
    return ERR;
  }
  windows_.resize(nw);
  rmsd_avgcorr_windows(T.nframes, maxwindow_, lagOffset_, windows_.data(), nw);
  double* d = (double*)g.Buffer("avgcorr.out", sizeof(double) * 2 * nw);
  if (!d) return ERR;
  if (g.Check(rmsd_avgcorr(g.ctx(), &T, ref_ ? &R : nullptr, 0, useMass_, maxwindow_, lagOffset_,
                           d, d + nw),
              "rmsavgcorr"))
    return ERR;
  std::vector<double> h(2 * (size_t)nw);
  if (g.Download(h.data(), d, sizeof(double) * h.size())) return ERR;
  Ct_->v.assign(h.begin(), h.begin() + nw);
  Csd_->v.assign(h.begin() + nw, h.end());
  return WriteOutput() ? ERR : OK;
}

int Analysis_RmsAvgCorr::WriteOutput() const {
  if (!outfile_.empty() && dataio::WriteSeries(outfile_, {Ct_, Csd_})) return 1;
  if (!separate_.empty()) {
    // The 'output' file: window size, mean, SD (src/Analysis_RmsAvgCorr.cpp:
    // 189-190, 259-260).
    FILE* f = fopen(separate_.c_str(), "w");
    if (!f) return 1;
    for (size_t i = 0; i < windows_.size(); ++i)
      fprintf(f, "%8i %f %f\n", windows_[i], Ct_->v[i], Csd_->v[i]);
    if (fclose(f) != 0) return 1;
  }
  return 0;
}

// ---- Metric_RMS + pairwise cache -------------------------------------------------------
RetType Metric_RMS::Setup(ArgList& args, DataSetList& dsl, MaskResolver resolve,
                          GpuSession& gpu) {
  gpu_ = &gpu;
  resolve_ = resolve;
  crd_ = dsl.FindCoordsSet(args.GetStringKey("crdset"));
  if (!crd_) {
    gpu.SetError("cluster: no COORDS set");
    return ERR;
  }
  if (args.hasKey("dme") || args.hasKey("data")) {
    gpu.SetError("cluster: only the 'rms' and 'srmsd' metrics are on the GPU RMSD path");
    return ERR;
  }
  symm_ = args.hasKey("srmsd");
  (void)args.hasKey("rms");
  useMass_ = args.hasKey("mass");
  nofit_ = args.hasKey("nofit");
  if (symm_ && nofit_) {
    // Metric_SRMSD without fit centres the reference frame but not the target
    // (SymmetricRmsdCalc::SymmRMSD), a combination rmsd_symm does not offer.
    gpu.SetError("cluster: 'srmsd' with 'nofit' is not on the GPU RMSD path");
    return ERR;
  }
  sieve_ = args.getKeyInt("sieve", 1);
  if (sieve_ < 1) {
    gpu.SetError("cluster: only regular sieves (sieve >= 1) are supported");
    return ERR;
  }
  if (args.hasKey("random")) {
    gpu.SetError("cluster: random sieving is not supported");
    return ERR;
  }
  std::string perr;
  if (!parse_prec(args, &precision_, &perr)) {
    gpu.SetError("cluster: " + perr);
    return ERR;
  }
  std::string pdSet = args.GetStringKey("pairdist");
  pairdistFile_ = args.GetStringKey("pairdistfile");
  const bool load = args.hasKey("loadpairdist");
  save_ = args.hasKey("savepairdist");
  mismatchFatal_ = !args.hasKey("pwrecalc");
  maskExpr_ = args.GetMaskNext();
  if (!resolve_or_all(resolve_, maskExpr_, *crd_, &sel_) || sel_.empty()) {
    gpu.SetError("cluster: no atoms selected for [" + maskExpr_ + "]");
    return ERR;
  }
  if (symm_) {
    std::string err;
    if (!symm_groups_for(crd_->top, sel_, &goff_, &gatoms_, &err)) {
      gpu.SetError("cluster: " + err);
      return ERR;
    }
  }
  const std::string name = args.GetStringNext();
  // MetricArray::setupPairwiseCache (src/Cluster/MetricArray.cpp:124-190).
  static const char* kDefault = "CpptrajPairDist";
  if (!pdSet.empty() && pairdistFile_.empty()) pairdistFile_ = pdSet;
  cache_ = nullptr;
  if (load || (!save_ && !pdSet.empty())) {
    if (pdSet.empty()) pdSet = kDefault;
    if (pairdistFile_.empty()) pairdistFile_ = kDefault;
    if ((cache_ = dsl.FindPairCache(pdSet)) == nullptr) {
      FILE* probe = fopen(pairdistFile_.c_str(), "rb");
      if (probe) {
        fclose(probe);
        MatrixFlt tri;
        size_t ntotal = 0;
        int sv = 1;
        std::vector<int> fr;
        if (cmatrix::Read(pairdistFile_, &tri, &ntotal, &sv, &fr)) {
          gpu.SetError("cluster: cannot read pairwise cache file " + pairdistFile_);
          return ERR;
        }
        cache_ = dsl.AddMatrix(pdSet);
        *cache_ = std::move(tri);
        cache_->name = pdSet;
        cache_->pwcache = true;
        cache_->ntotal = ntotal;
        cache_->sieve = sv;
        cache_->frames = std::move(fr);
        // a cache read from its file is not written back unless 'pwrecalc'
        if (save_ && mismatchFatal_) save_ = false;
      }
    }
    // neither a set nor a file: compute the distances and save them
    if (!cache_) save_ = true;
  }
  if (!cache_) {
    cache_ = dsl.AddMatrix(!pdSet.empty() ? pdSet
                                          : !name.empty() ? name : dsl.GenerateDefaultName(kDefault));
  }
  if (save_ && pairdistFile_.empty()) pairdistFile_ = kDefault;
  // Regular sieve (Cluster::Sieve::SetFramesToCluster, src/Cluster/Sieve.cpp:42-55).
  frames_.clear();
  for (int f = 0; f < crd_->nframes(); f += sieve_) frames_.push_back(f);
  return OK;
}

// MetricArray::CacheDistances + calcFrameDistances (src/Cluster/
// MetricArray.cpp:722-801): the triangle over framesToCache in one call.
RetType Metric_RMS::CacheDistances() {
  GpuSession& g = *gpu_;
  if (frames_.empty()) return OK;
  if (cache_->pwcache && cache_->Nrows() > 0) {
    // an existing cache: used when it holds every frame to cluster
    // (MetricArray::CacheDistances, src/Cluster/MetricArray.cpp:729-747)
    if (cache_->CachedFramesMatch(frames_)) return SaveCache();
    if (mismatchFatal_) {
      g.SetError("cluster: frames to cluster do not match those in existing cache '" +
                 cache_->name + "' (specify 'pwrecalc' to recalculate)");
      return ERR;
    }
  }
  rmsd_frames F;
  if (g.FramesOnDevice("cluster.crd", *crd_, sel_, &F)) return ERR;
  const int n = (int)frames_.size();
  int32_t* fidx = (int32_t*)g.Buffer("cluster.fidx", sizeof(int32_t) * std::max(n, 1));
  if (!fidx) return ERR;
  std::vector<int32_t> f32(frames_.begin(), frames_.end());
  if (g.Upload(fidx, f32.data(), sizeof(int32_t) * n)) return ERR;
  F.fidx = fidx;
  F.nframes = n;
  if (symm_) {
    // Metric_SRMSD::FrameDist(f1, f2): frame f1 against frame f2 as the
    // reference (src/Cluster/Metric_SRMSD.cpp:34-38).
    if (precision_ != RMSD_PREC_FP64) {
      g.SetError("cluster: 'srmsd' is FP64 only");
      return ERR;
    }
    if (symm_rows(g, F, nullptr, true, useMass_ ? 1 : 0, goff_, gatoms_, cache_, true) != OK)
      return ERR;
  } else {
    cache_->AllocateTriangle(n);
    float* d = (float*)g.Buffer("cluster.out", sizeof(float) * std::max<size_t>(1, cache_->size()));
    if (!d) return ERR;
    if (g.Check(rmsd_cluster_cache(g.ctx(), &F, nofit_, useMass_, precision_, 1.0, d), "cluster"))
      return ERR;
    if (g.Download(cache_->data(), d, sizeof(float) * cache_->size())) return ERR;
  }
  // DataSet_PairwiseCache::SetupCache's bookkeeping (src/DataSet_PairwiseCache.cpp)
  cache_->pwcache = true;
  cache_->ntotal = (size_t)crd_->nframes();
  cache_->sieve = sieve_;
  cache_->frames = frames_;
  return SaveCache();
}

// 'savepairdist' (or a 'pairdist' file that did not exist yet): the cache in
// the Cmatrix binary format, over the cache's own frames.
RetType Metric_RMS::SaveCache() {
  if (!save_) return OK;
  if (cmatrix::Write(pairdistFile_, *cache_, cache_->ntotal, cache_->sieve, cache_->frames)) {
    gpu_->SetError("cannot write " + pairdistFile_);
    return ERR;
  }
  return OK;
}

double Metric_RMS::FrameDist(int f1, int f2) {
  GpuSession& g = *gpu_;
  rmsd_frames A, B;
  if (g.FramesOnDevice("metric.a", *crd_, sel_, &A, crd_->frame(f1), 1) ||
      g.FramesOnDevice("metric.b", *crd_, sel_, &B, crd_->frame(f2), 1))
    return -1.0;
  if (symm_) {
    // Metric_SRMSD::FrameDist: f1 against f2 as the (centred) reference.
    double* dr = (double*)g.Buffer("metric.symm", sizeof(double));
    if (!dr || g.Check(rmsd_symm(g.ctx(), &A, &B, 0, (int)goff_.size() - 1, goff_.data(),
                                 gatoms_.empty() ? nullptr : gatoms_.data(), 1, useMass_ ? 1 : 0,
                                 dr, nullptr, nullptr, nullptr),
                       "FrameDist"))
      return -1.0;
    double v = 0.0;
    if (g.Download(&v, dr, sizeof(double))) return -1.0;
    return v;
  }
  float* d = (float*)g.Buffer("metric.out", sizeof(float));
  if (!d) return -1.0;
  // Full 1 x 1 block: row = f1 as the target, column = f2 as the reference
  // (Frame::RMSD(Ref) centres the reference, src/Cluster/Metric_RMS.cpp:44-63).
  if (g.Check(rmsd_pairwise(g.ctx(), &A, &B, nofit_ ? RMSD_MODE_NOFIT : RMSD_MODE_FIT, useMass_,
                            RMSD_PREC_FP64, d),
              "FrameDist"))
    return -1.0;
  float v = 0.f;
  if (g.Download(&v, d, sizeof(float))) return -1.0;
  return v;
}

// ---- Session ------------------------------------------------------------------------------
int Session::Run(const std::string& line) {
  err.clear();
  if (!gpu.ok()) {
    err = gpu.error();
    return 1;
  }
  ArgList args(line);
  const std::string& cmd = args.Command();
  MaskResolver resolve = [this](const std::string& expr, const CoordsSet& set, AtomMask* sel) {
    auto it = masks.find(set.name + "|" + expr);  // set-specific entry first
    if (it == masks.end()) it = masks.find(expr);
    if (it == masks.end()) return false;
    for (int a : it->second)
      if (a < 0 || a >= set.top.natom) return false;
    *sel = it->second;
    return true;
  };
  int rc = 1;
  if (cmd == "rms2d" || cmd == "2drms") {
    Analysis_Rms2d a;
    rc = a.Setup(args, dsl, resolve, gpu) == OK && a.Analyze() == OK ? 0 : 1;
  } else if (cmd == "rms" || cmd == "rmsd") {
    Action_Rmsd a;
    const std::string crdset = args.GetStringKey("crdset");
    rc = 1;
    if (a.Init(args, dsl, resolve, gpu) == OK) {
      CoordsSet* crd = dsl.FindCoordsSet(crdset);
      if (!crd) gpu.SetError("rms: no COORDS set");
      else if (a.DoFrames(*crd) == OK && a.WriteOutput() == 0) rc = 0;
      else if (gpu.error().empty()) gpu.SetError("rms: cannot write output");
    }
  } else if (cmd == "align") {
    Action_Rmsd a;
    const std::string crdset = args.GetStringKey("crdset");
    rc = 1;
    if (a.InitAlign(args, dsl, resolve, gpu) == OK) {
      CoordsSet* crd = dsl.FindCoordsSet(crdset);
      if (!crd) gpu.SetError("align: no COORDS set");
      else if (a.DoFrames(*crd) == OK) rc = 0;
    }
  } else if (cmd == "symmrmsd") {
    Action_SymmetricRmsd a;
    const std::string crdset = args.GetStringKey("crdset");
    rc = 1;
    if (a.Init(args, dsl, resolve, gpu) == OK) {
      CoordsSet* crd = dsl.FindCoordsSet(crdset);
      if (!crd) gpu.SetError("symmrmsd: no COORDS set");
      else if (a.DoFrames(*crd) == OK && a.WriteOutput() == 0) rc = 0;
      else if (gpu.error().empty()) gpu.SetError("symmrmsd: cannot write output");
    }
  } else if (cmd == "rmsavgcorr") {
    Analysis_RmsAvgCorr a;
    rc = a.Setup(args, dsl, resolve, gpu) == OK && a.Analyze() == OK ? 0 : 1;
  } else if (cmd == "cluster") {
    Metric_RMS m;
    rc = m.Setup(args, dsl, resolve, gpu) == OK && m.CacheDistances() == OK ? 0 : 1;
  } else {
    err = "unknown command '" + cmd + "'";
    return 1;
  }
  if (rc) err = gpu.error().empty() ? "command failed" : gpu.error();
  return rc;
}

}  // namespace cpptraj_amd
