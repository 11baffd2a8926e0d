// host_core.cpp -- ArgList, data containers, text and Cmatrix output, and the
// device session of the host layer.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <sstream>

#include "host.hpp"

namespace cpptraj_amd {

// ---- ArgList -----------------------------------------------------------------
ArgList::ArgList(const std::string& line) {
  std::istringstream is(line);
  std::string tok;
  while (is >> tok) args_.push_back(tok);
  marked_.assign(args_.size(), false);
  if (!marked_.empty()) marked_[0] = true;  // the command itself
}

const std::string& ArgList::Command() const {
  static const std::string empty;
  return args_.empty() ? empty : args_[0];
}

bool ArgList::hasKey(const char* key) {
  for (size_t i = 0; i < args_.size(); ++i)
    if (!marked_[i] && args_[i] == key) {
      marked_[i] = true;
      return true;
    }
  return false;
}

bool ArgList::Contains(const char* key) const {
  for (size_t i = 0; i < args_.size(); ++i)
    if (!marked_[i] && args_[i] == key) return true;
  return false;
}

std::string ArgList::GetStringKey(const char* key) {
  for (size_t i = 0; i + 1 < args_.size(); ++i)
    if (!marked_[i] && args_[i] == key) {
      marked_[i] = marked_[i + 1] = true;
      return args_[i + 1];
    }
  return std::string();
}

std::string ArgList::GetStringNext() {
  for (size_t i = 0; i < args_.size(); ++i)
    if (!marked_[i]) {
      marked_[i] = true;
      return args_[i];
    }
  return std::string();
}

bool ArgList::StrIsMask(const std::string& s) {
  size_t p = 0;
  while (p < s.size() && (s[p] == '!' || s[p] == '(')) ++p;
  if (p >= s.size()) return false;
  const char c = s[p];
  return c == '@' || c == ':' || c == '^' || c == '*' || c == '=';
}

std::string ArgList::GetMaskNext() {
  for (size_t i = 0; i < args_.size(); ++i)
    if (!marked_[i] && StrIsMask(args_[i])) {
      marked_[i] = true;
      return args_[i];
    }
  return std::string();
}

int ArgList::getKeyInt(const char* key, int def) {
  std::string v = GetStringKey(key);
  if (v.empty()) return def;
  return atoi(v.c_str());
}

std::vector<std::string> ArgList::Unmarked() const {
  std::vector<std::string> r;
  for (size_t i = 0; i < args_.size(); ++i)
    if (!marked_[i]) r.push_back(args_[i]);
  return r;
}

// ---- MatrixFlt ------------------------------------------------------------------
void MatrixFlt::AllocateTriangle(size_t n) {
  kind_ = TRI;
  ncols_ = nrows_ = n;
  v_.assign(n > 1 ? n * (n - 1) / 2 : 0, 0.0f);
}

void MatrixFlt::Allocate2D(size_t ncols, size_t nrows) {
  kind_ = FULL;
  ncols_ = ncols;
  nrows_ = nrows;
  v_.assign(ncols * nrows, 0.0f);
}

double MatrixFlt::GetElement(size_t x, size_t y) const {
  if (kind_ == FULL) return (double)v_[y * ncols_ + x];
  if (x == y) return 0.0;
  const size_t i = std::min(x, y), j = std::max(x, y);
  return (double)v_[ncols_ * i - (i * (i + 1)) / 2 + j - i - 1];
}

// ---- DataSetList ------------------------------------------------------------------
CoordsSet* DataSetList::AddCoords(const std::string& name) {
  coords_.emplace_back(new CoordsSet());
  coords_.back()->name = name;
  return coords_.back().get();
}

CoordsSet* DataSetList::FindCoordsSet(const std::string& name) const {
  if (name.empty()) return coords_.empty() ? nullptr : coords_.front().get();
  for (const auto& c : coords_)
    if (c->name == name) return c.get();
  return nullptr;
}

CoordsSet* DataSetList::AddReference(const std::string& name) {
  refs_.emplace_back(new CoordsSet());
  refs_.back()->name = name;
  return refs_.back().get();
}

CoordsSet* DataSetList::FindReference(const std::string& name) const {
  if (name.empty()) return refs_.empty() ? nullptr : refs_.front().get();
  for (const auto& c : refs_)
    if (c->name == name) return c.get();
  return nullptr;
}

CoordsSet* DataSetList::ReferenceAt(int idx) const {
  return idx >= 0 && idx < (int)refs_.size() ? refs_[idx].get() : nullptr;
}

MatrixFlt* DataSetList::AddMatrix(const std::string& name) {
  mats_.emplace_back(new MatrixFlt());
  mats_.back()->name = name;
  return mats_.back().get();
}

Series* DataSetList::AddSeries(const std::string& name) {
  series_.emplace_back(new Series());
  series_.back()->name = name;
  return series_.back().get();
}

const MatrixFlt* DataSetList::FindMatrix(const std::string& name) const {
  for (auto it = mats_.rbegin(); it != mats_.rend(); ++it)
    if ((*it)->name == name) return it->get();
  return nullptr;
}

MatrixFlt* DataSetList::FindPairCache(const std::string& name) {
  for (auto it = mats_.rbegin(); it != mats_.rend(); ++it)
    if ((*it)->pwcache && (*it)->name == name) return it->get();
  return nullptr;
}

bool MatrixFlt::CachedFramesMatch(const std::vector<int>& want) const {
  std::vector<char> present(ntotal, 0);
  for (int f : frames)
    if (f >= 0 && (size_t)f < ntotal) present[f] = 1;
  for (int f : want)
    if (f < 0 || (size_t)f >= ntotal || !present[f]) return false;
  return true;
}

const Series* DataSetList::FindSeries(const std::string& name) const {
  for (auto it = series_.rbegin(); it != series_.rend(); ++it)
    if ((*it)->name == name) return it->get();
  return nullptr;
}

std::string DataSetList::GenerateDefaultName(const std::string& base) {
  char buf[64];
  snprintf(buf, sizeof(buf), "%s_%05d", base.c_str(), counters_[base]++);
  return buf;
}

// ---- text output ------------------------------------------------------------------
namespace dataio {

// Width of a frame-number column (TextFormat::SetCoordFormat,
// src/TextFormat.cpp:52-82): the digits of the largest coordinate
// (step * n + min), widened for the precision, at least the default width.
static int coord_width(size_t n, int def_width, int prec, long minc = 1, double step = 1.0) {
  long maxc = (long)(step * (double)n + (double)minc);
  int w = 1;
  while (maxc >= 10) {
    maxc /= 10;
    ++w;
  }
  if (prec != 0 && w + prec + 1 > def_width) w = w + prec + 1;
  return std::max(w, def_width);
}

int WriteMatrix(const std::string& path, const MatrixFlt& m, bool square2d) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  const size_t nr = m.Nrows(), nc = m.Ncols();
  if (square2d) {
    // "#Frame" left-aligned in 8 columns, then the column numbers in the
    // value columns' width (9 = leading space + %8.3f); rows: frame number
    // %8.0f, then " %8.3f" per value.
    const int yw = coord_width(nr, 8, 0);
    const int xw = coord_width(nc, 9, 0);
    fprintf(f, "%-*s", 8, "#Frame");
    for (size_t x = 0; x < nc; ++x) fprintf(f, "%*.0f", xw, (double)(x + 1));
    fprintf(f, "\n");
    for (size_t y = 0; y < nr; ++y) {
      fprintf(f, "%*.0f", yw, (double)(y + 1));
      for (size_t x = 0; x < nc; ++x) fprintf(f, " %8.3f", m.GetElement(x, y));
      fprintf(f, "\n");
    }
  } else {
    // "#<xlabel> <ylabel> <legend>" (empty labels), then "x y value" lines
    // with the coordinates as %8.3f.
    fprintf(f, "#  %s\n", m.name.c_str());
    const int xw = coord_width(nc, 8, 3), yw = coord_width(nr, 8, 3);
    for (size_t y = 0; y < nr; ++y)
      for (size_t x = 0; x < nc; ++x)
        fprintf(f, "%*.3f %*.3f  %8.3f\n", xw, (double)(x + 1), yw, (double)(y + 1),
                m.GetElement(x, y));
  }
  return fclose(f) != 0;
}

int WriteSeries(const std::string& path, const std::vector<const Series*>& sets) {
  if (sets.empty()) return 0;
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  size_t n = 0;
  for (const Series* s : sets) n = std::max(n, s->nframes());
  // x column from the first set's dimension (min, step), as integers
  const double x0 = sets[0]->xmin, dx = sets[0]->xstep;
  const int fw = coord_width(n, 8, 0, (long)x0, dx);
  // Header: "#Frame", then each set's legend right-aligned over its columns
  // (ncomp x (1 + width)); spaces in names become underscores.
  fprintf(f, "%-*s", fw, "#Frame");
  for (const Series* s : sets) {
    std::string nm = s->name;
    std::replace(nm.begin(), nm.end(), ' ', '_');
    const int cw = s->ncomp * (s->width + 1);
    fprintf(f, "%*s", cw, nm.c_str());
  }
  fprintf(f, "\n");
  for (size_t i = 0; i < n; ++i) {
    fprintf(f, "%*.0f", fw, x0 + (double)i * dx);
    for (const Series* s : sets)
      for (int c = 0; c < s->ncomp; ++c) {
        const double v = i < s->nframes() ? s->v[i * s->ncomp + c] : 0.0;
        fprintf(f, " %*.*f", s->width, s->prec, v);
      }
    fprintf(f, "\n");
  }
  return fclose(f) != 0;
}

static const char* legend_of(const Series* s) {
  return s->legend.empty() ? s->name.c_str() : s->legend.c_str();
}

int WriteGrace(const std::string& path, const std::vector<const Series*>& sets, bool invert) {
  if (sets.empty()) return 1;
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  // Header: axis labels from the first set's dimension ("Frame"), swapped
  // when inverted.
  fprintf(f, "@with g0\n@  xaxis label \"%s\"\n@  yaxis label \"%s\"\n"
             "@  legend 0.2, 0.995\n@  legend char size 0.60\n",
          invert ? "" : "Frame", invert ? "Frame" : "");
  if (!invert) {
    // One xy set per data set: x = frame number (1-based).
    for (size_t k = 0; k < sets.size(); ++k) {
      const Series* s = sets[k];
      fprintf(f, "@  s%zu legend \"%s\"\n@target G0.S%zu\n@type xy\n", k, legend_of(s), k);
      const size_t n = s->nframes();
      const int xw = coord_width(n, 8, 3);
      for (size_t i = 0; i < n; ++i) {
        fprintf(f, "%*.3f", xw, (double)(i + 1));
        for (int c = 0; c < s->ncomp; ++c)
          fprintf(f, " %*.*f", s->width, s->prec, s->v[i * s->ncomp + c]);
        fprintf(f, "\n");
      }
    }
  } else {
    // One xy set per frame: x = data set number (0-based), then the legend.
    size_t n = 0;
    for (const Series* s : sets) n = std::max(n, s->nframes());
    const int xw = coord_width(sets.size(), 8, 3, 0);
    for (size_t i = 0; i < n; ++i) {
      fprintf(f, "@target G0.S%zu\n@type xy\n", i);
      for (size_t k = 0; k < sets.size(); ++k) {
        const Series* s = sets[k];
        fprintf(f, "%*.3f", xw, (double)k);
        for (int c = 0; c < s->ncomp; ++c)
          fprintf(f, " %*.*f", s->width, s->prec,
                  i < s->nframes() ? s->v[i * s->ncomp + c] : 0.0);
        fprintf(f, " \"%s\"\n", legend_of(s));
      }
    }
  }
  return fclose(f) != 0;
}

int WriteMesh(const std::string& path, const std::string& xlabel, const std::vector<double>& x,
              const std::vector<const Series*>& sets) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  // x column: %8.3f, widened for coordinates of more than four digits.
  double xmax = 0.0;
  for (double v : x) xmax = std::max(xmax, std::fabs(v));
  const int xw = coord_width(0, 8, 3, (long)xmax);
  fprintf(f, "%-*s", xw, ("#" + xlabel).c_str());
  for (const Series* s : sets) {
    std::string nm = s->name;
    std::replace(nm.begin(), nm.end(), ' ', '_');
    fprintf(f, "%*s", s->width + 1, nm.c_str());
  }
  fprintf(f, "\n");
  for (size_t i = 0; i < x.size(); ++i) {
    fprintf(f, "%*.3f", xw, x[i]);
    for (const Series* s : sets)
      fprintf(f, " %*.*f", s->width, s->prec, i < s->v.size() ? s->v[i] : 0.0);
    fprintf(f, "\n");
  }
  return fclose(f) != 0;
}

}  // namespace dataio

// ---- Cmatrix binary -----------------------------------------------------------------
namespace cmatrix {

static const unsigned char kMagic[4] = {'C', 'T', 'M', 2};

int Write(const std::string& path, const MatrixFlt& tri, size_t actual_nrows, int sieve,
          const std::vector<int>& frames) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return 1;
  const uint64_t original = actual_nrows;
  const uint64_t nrows = tri.Nrows();
  const int64_t sv = sieve;
  bool ok = fwrite(kMagic, 1, 4, f) == 4 && fwrite(&original, 8, 1, f) == 1 &&
            fwrite(&nrows, 8, 1, f) == 1 && fwrite(&sv, 8, 1, f) == 1;
  if (ok && tri.size()) ok = fwrite(tri.data(), sizeof(float), tri.size(), f) == tri.size();
  if (ok && sieve != 1) {
    // 'T' for a frame that was sieved out, 'F' for a cached one.
    std::vector<char> present(original, 'T');
    for (int fr : frames)
      if (fr >= 0 && (size_t)fr < original) present[fr] = 'F';
    ok = fwrite(present.data(), 1, present.size(), f) == present.size();
  }
  return (fclose(f) != 0 || !ok) ? 1 : 0;
}

int Read(const std::string& path, MatrixFlt* tri, size_t* actual_nrows, int* sieve,
         std::vector<int>* frames) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return 1;
  unsigned char magic[4];
  uint64_t original = 0, nrows = 0;
  int64_t sv = 1;
  bool ok = fread(magic, 1, 4, f) == 4 && memcmp(magic, kMagic, 4) == 0 &&
            fread(&original, 8, 1, f) == 1 && fread(&nrows, 8, 1, f) == 1 &&
            fread(&sv, 8, 1, f) == 1;
  if (ok) {
    tri->AllocateTriangle(nrows);
    if (tri->size()) ok = fread(tri->data(), sizeof(float), tri->size(), f) == tri->size();
  }
  frames->clear();
  if (ok && sv != 1) {
    std::vector<char> present(original);
    ok = fread(present.data(), 1, original, f) == original;
    for (size_t i = 0; ok && i < original; ++i)
      if (present[i] == 'F') frames->push_back((int)i);
  } else if (ok) {
    // sieve 1: every frame of the run is present
    // (DataIO_Cmatrix_Binary::ReadCmatrix, src/DataIO_Cmatrix_Binary.cpp:53-59)
    for (uint64_t i = 0; i < original; ++i) frames->push_back((int)i);
  }
  fclose(f);
  if (!ok) return 1;
  *actual_nrows = original;
  *sieve = (int)sv;
  return 0;
}

}  // namespace cmatrix

// ---- pinned host memory ---------------------------------------------------------------
void* pinned_alloc(size_t bytes) {
  if (bytes == 0) bytes = 16;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) throw std::bad_alloc();
  return p;
}

void pinned_free(void* p) noexcept {
  if (p) (void)hipHostFree(p);
}

// ---- GpuSession -------------------------------------------------------------------
GpuSession::GpuSession(int device) : device_(device) {
  if (rmsd_ctx_create(&ctx_, device, nullptr) != RMSD_OK) {
    ctx_ = nullptr;
    err_ = "cannot create the RMSD context on device " + std::to_string(device);
  }
}

GpuSession::~GpuSession() {
  for (int k = 0; k < 2; ++k) {
    if (slot_ctx_[k]) {
      rmsd_ctx_synchronize(slot_ctx_[k]);
      rmsd_ctx_destroy(slot_ctx_[k]);
    }
    if (slot_stream_[k]) (void)hipStreamDestroy((hipStream_t)slot_stream_[k]);
  }
  if (ctx_) {
    rmsd_ctx_synchronize(ctx_);
    rmsd_ctx_destroy(ctx_);
  }
  (void)hipSetDevice(device_);
  for (auto& kv : bufs_)
    if (kv.second.p) (void)hipFree(kv.second.p);
}

void* GpuSession::Buffer(const std::string& name, size_t bytes) {
  Buf& b = bufs_[name];
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return b.p;
  (void)hipSetDevice(device_);
  if (b.p) {
    Sync();
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    err_ = "cannot allocate " + std::to_string(bytes) + " device bytes for " + name;
    return nullptr;
  }
  b.bytes = bytes;
  return b.p;
}

int GpuSession::Upload(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  (void)hipSetDevice(device_);
  // The context's stream is the null stream: a synchronous copy is ordered
  // with the library's work.
  if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    err_ = "host to device copy failed";
    return 1;
  }
  return 0;
}

int GpuSession::Download(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  (void)hipSetDevice(device_);
  if (hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    err_ = "device to host copy failed";
    return 1;
  }
  return 0;
}

int GpuSession::Sync() {
  if (!ctx_) return 1;
  return Check(rmsd_ctx_synchronize(ctx_), "synchronize");
}

int GpuSession::Slot(int k, rmsd_ctx** ctx, void** stream) {
  if (k < 0 || k > 1) return 1;
  if (!slot_ctx_[k]) {
    (void)hipSetDevice(device_);
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      err_ = "cannot create a transfer stream";
      return 1;
    }
    if (rmsd_ctx_create(&slot_ctx_[k], device_, st) != RMSD_OK) {
      (void)hipStreamDestroy(st);
      slot_ctx_[k] = nullptr;
      err_ = "cannot create a transfer context";
      return 1;
    }
    slot_stream_[k] = st;
  }
  *ctx = slot_ctx_[k];
  *stream = slot_stream_[k];
  return 0;
}

int GpuSession::Check(int rc, const char* what) {
  if (rc == RMSD_OK) return 0;
  err_ = std::string(what) + ": " + (ctx_ ? rmsd_last_error(ctx_) : "no context");
  return 1;
}

int GpuSession::FramesOnDevice(const std::string& tag, const CoordsSet& set, const AtomMask& sel,
                               rmsd_frames* out, const float* xyz_override, int nframes) {
  const int nf = nframes < 0 ? set.nframes() : nframes;
  const float* src = xyz_override ? xyz_override : set.xyz.data();
  const size_t fbytes = sizeof(float) * (size_t)set.stride * (size_t)nf;
  float* dx = (float*)Buffer(tag + ".xyz", fbytes);
  int32_t* ds = (int32_t*)Buffer(tag + ".sel", sizeof(int32_t) * sel.size());
  double* dm = (double*)Buffer(tag + ".mass", sizeof(double) * sel.size());
  if (!dx || !ds || !dm) return 1;
  std::vector<int32_t> s32(sel.begin(), sel.end());
  std::vector<double> m(sel.size(), 1.0);
  for (size_t k = 0; k < sel.size(); ++k) {
    if (sel[k] < 0 || sel[k] >= set.top.natom) {
      err_ = "selection index out of range for set " + set.name;
      return 1;
    }
    if (!set.top.mass.empty()) m[k] = set.top.mass[sel[k]];
  }
  if (Upload(dx, src, fbytes) || Upload(ds, s32.data(), sizeof(int32_t) * s32.size()) ||
      Upload(dm, m.data(), sizeof(double) * m.size()))
    return 1;
  // Every atom in order: no selection array, which the library reads as the
  // same atoms and which lets it use its contiguous-frame loads.
  bool identity = (int)sel.size() == set.top.natom;
  for (size_t k = 0; identity && k < sel.size(); ++k) identity = sel[k] == (int)k;
  memset(out, 0, sizeof(*out));
  out->xyz = dx;
  out->stride = set.stride;
  out->ld = 0;
  out->nframes = nf;
  out->natom = set.top.natom;
  out->sel = identity ? nullptr : ds;
  out->nsel = (int32_t)sel.size();
  out->mass = dm;
  out->fidx = nullptr;
  return 0;
}

}  // namespace cpptraj_amd
