// host_capi.cpp -- C entry points over the host layer's Session.
#include <string.h>

#include <string>

#include "host.hpp"
#include "host_capi.h"

using namespace cpptraj_amd;

struct chost_session {
  explicit chost_session(int device) : s(device) {}
  Session s;
  std::string err;
};

extern "C" {

int chost_create(int device, chost_session** out) {
  if (!out) return 1;
  *out = new chost_session(device);
  if (!(*out)->s.gpu.ok()) {
    (*out)->err = (*out)->s.gpu.error();
    return 2;
  }
  return 0;
}

void chost_destroy(chost_session* s) { delete s; }

const char* chost_last_error(const chost_session* s) { return s ? s->err.c_str() : "no session"; }

int chost_add_coords(chost_session* s, const char* name, const float* xyz, int64_t nframes,
                     int64_t stride, int natom, const double* mass) {
  if (!s || !name || natom < 1 || stride < 3 * (int64_t)natom || nframes < 0 ||
      (nframes > 0 && !xyz)) {
    if (s) s->err = "chost_add_coords: bad arguments";
    return 1;
  }
  CoordsSet* c = s->s.dsl.AddCoords(name);
  c->top.natom = natom;
  c->top.mass.assign(natom, 1.0);
  if (mass) c->top.mass.assign(mass, mass + natom);
  c->stride = stride;
  c->xyz.assign(xyz, xyz + nframes * stride);
  return 0;
}

int chost_set_box(chost_session* s, const char* name, const double* ucell, int64_t nframes) {
  if (!s || !name || nframes < 0 || (nframes > 0 && !ucell)) return 1;
  CoordsSet* c = s->s.dsl.FindCoordsSet(name);
  if (!c) {
    s->err = std::string("no COORDS set '") + name + "'";
    return 1;
  }
  if (nframes != 0 && nframes != c->nframes()) {
    s->err = "chost_set_box: one unit cell per frame expected";
    return 1;
  }
  c->box.assign(ucell, ucell + 9 * nframes);
  return 0;
}

int chost_get_box(chost_session* s, const char* name, double* out, int64_t* nframes) {
  if (!s || !name) return 1;
  const CoordsSet* c = s->s.dsl.FindCoordsSet(name);
  if (!c) {
    s->err = std::string("no COORDS set '") + name + "'";
    return 1;
  }
  if (nframes) *nframes = (int64_t)(c->box.size() / 9);
  if (out && !c->box.empty()) memcpy(out, c->box.data(), sizeof(double) * c->box.size());
  return 0;
}

int chost_add_reference(chost_session* s, const char* name, const float* xyz, int natom,
                        const double* mass) {
  if (!s || !name || !xyz || natom < 1) {
    if (s) s->err = "chost_add_reference: bad arguments";
    return 1;
  }
  CoordsSet* c = s->s.dsl.AddReference(name);
  c->top.natom = natom;
  c->top.mass.assign(natom, 1.0);
  if (mass) c->top.mass.assign(mass, mass + natom);
  c->stride = 3 * (int64_t)natom;
  c->xyz.assign(xyz, xyz + c->stride);
  return 0;
}

int chost_set_residues(chost_session* s, const char* set, int nres, const int* first,
                       const char* const* names, int nsolute, const int* solute) {
  if (!s || !set || nres < 0 || (nres > 0 && !first) || nsolute < 0 ||
      (nsolute > 0 && !solute))
    return 1;
  CoordsSet* c = s->s.dsl.FindCoordsSet(set);
  if (!c) c = s->s.dsl.FindReference(set);
  if (!c || (nres > 0 && *set == 0)) {
    s->err = std::string("chost_set_residues: no set '") + set + "'";
    return 1;
  }
  for (int r = 0; r < nres; ++r)
    if (first[r] < 0 || first[r] >= c->top.natom || (r > 0 && first[r] <= first[r - 1])) {
      s->err = "chost_set_residues: residue first atoms must increase within the topology";
      return 1;
    }
  for (int k = 0; k < nsolute; ++k)
    if (solute[k] < 0 || solute[k] >= nres) {
      s->err = "chost_set_residues: solute residue out of range";
      return 1;
    }
  c->top.resFirst.assign(first, first + nres);
  c->top.resName.clear();
  for (int r = 0; r < nres; ++r) {
    std::string n = names && names[r] ? names[r] : "";
    while (!n.empty() && n.back() == ' ') n.pop_back();
    c->top.resName.push_back(n);
  }
  c->top.soluteRes.assign(solute, solute + nsolute);
  return 0;
}

int chost_set_symmetric_groups(chost_session* s, const char* set, int ngroups, const int* off,
                               const int* atoms) {
  if (!s || !set || ngroups < 0 || (ngroups > 0 && (!off || !atoms))) return 1;
  CoordsSet* c = s->s.dsl.FindCoordsSet(set);
  if (!c) c = s->s.dsl.FindReference(set);
  if (!c) {
    s->err = std::string("chost_set_symmetric_groups: no set '") + set + "'";
    return 1;
  }
  if (ngroups > 0 && off[0] != 0) {
    s->err = "chost_set_symmetric_groups: off[0] must be 0";
    return 1;
  }
  std::vector<std::vector<int>> g((size_t)ngroups);
  for (int k = 0; k < ngroups; ++k) {
    if (off[k + 1] < off[k]) {
      s->err = "chost_set_symmetric_groups: decreasing offsets";
      return 1;
    }
    for (int j = off[k]; j < off[k + 1]; ++j) {
      if (atoms[j] < 0 || atoms[j] >= c->top.natom) {
        s->err = "chost_set_symmetric_groups: atom out of range";
        return 1;
      }
      g[k].push_back(atoms[j]);
    }
  }
  c->top.symmGroups = std::move(g);
  return 0;
}

int chost_add_mask(chost_session* s, const char* set, const char* expr, const int* sel, int nsel) {
  if (!s || !expr || nsel < 0 || (nsel > 0 && !sel)) return 1;
  const std::string key = (set && *set) ? std::string(set) + "|" + expr : std::string(expr);
  s->s.masks[key] = AtomMask(sel, sel + nsel);
  return 0;
}

int chost_run(chost_session* s, const char* line) {
  if (!s || !line) return 1;
  const int rc = s->s.Run(line);
  s->err = s->s.err;
  return rc;
}

int chost_get_matrix(chost_session* s, const char* name, float* out, int64_t* ncols,
                     int64_t* nrows, int* is_tri) {
  if (!s || !name) return 1;
  const MatrixFlt* m = s->s.dsl.FindMatrix(name);
  if (!m) {
    s->err = std::string("no matrix '") + name + "'";
    return 1;
  }
  if (ncols) *ncols = (int64_t)m->Ncols();
  if (nrows) *nrows = (int64_t)m->Nrows();
  if (is_tri) *is_tri = m->kind() == MatrixFlt::TRI;
  if (out && m->size()) memcpy(out, m->data(), sizeof(float) * m->size());
  return 0;
}

int chost_get_series(chost_session* s, const char* name, double* out, int64_t* nframes,
                     int* ncomp) {
  if (!s || !name) return 1;
  const Series* d = s->s.dsl.FindSeries(name);
  if (!d) {
    s->err = std::string("no data set '") + name + "'";
    return 1;
  }
  if (nframes) *nframes = (int64_t)d->nframes();
  if (ncomp) *ncomp = d->ncomp;
  if (out && !d->v.empty()) memcpy(out, d->v.data(), sizeof(double) * d->v.size());
  return 0;
}

int chost_get_coords(chost_session* s, const char* name, float* out, int64_t* nfloats) {
  if (!s || !name) return 1;
  const CoordsSet* c = s->s.dsl.FindCoordsSet(name);
  if (!c) {
    s->err = std::string("no COORDS set '") + name + "'";
    return 1;
  }
  if (nfloats) *nfloats = (int64_t)c->xyz.size();
  if (out && !c->xyz.empty()) memcpy(out, c->xyz.data(), sizeof(float) * c->xyz.size());
  return 0;
}

int chost_write_matrix(const char* path, const float* data, int64_t ncols, int64_t nrows,
                       int is_tri, int square2d, const char* name) {
  if (!path || ncols < 0 || nrows < 0) return 1;
  MatrixFlt m;
  if (is_tri) m.AllocateTriangle((size_t)ncols);
  else m.Allocate2D((size_t)ncols, (size_t)nrows);
  if (m.size() && !data) return 1;
  if (m.size()) memcpy(m.data(), data, sizeof(float) * m.size());
  m.name = name ? name : "";
  return dataio::WriteMatrix(path, m, square2d != 0);
}

int chost_write_series(const char* path, int nsets, const char* const* names, const int* ncomp,
                       const int* width, const int* prec, const double* const* values,
                       int64_t nframes) {
  if (!path || nsets < 0 || nframes < 0) return 1;
  std::vector<Series> sets(nsets);
  std::vector<const Series*> ptrs;
  for (int k = 0; k < nsets; ++k) {
    Series& s = sets[k];
    s.name = names[k];
    s.ncomp = ncomp[k];
    s.width = width[k];
    s.prec = prec[k];
    s.v.assign(values[k], values[k] + nframes * s.ncomp);
    ptrs.push_back(&s);
  }
  return dataio::WriteSeries(path, ptrs);
}

int chost_write_cmatrix(const char* path, const float* tri, int64_t nrows, int64_t actual_nrows,
                        int sieve, const int* frames) {
  if (!path || nrows < 0 || (nrows > 1 && !tri) || (nrows > 0 && !frames)) return 1;
  MatrixFlt m;
  m.AllocateTriangle((size_t)nrows);
  if (m.size()) memcpy(m.data(), tri, sizeof(float) * m.size());
  std::vector<int> fr(frames, frames + nrows);
  return cmatrix::Write(path, m, (size_t)actual_nrows, sieve, fr);
}

int chost_read_cmatrix(const char* path, float* tri, int64_t* nrows, int64_t* actual_nrows,
                       int* sieve, int* frames) {
  MatrixFlt m;
  size_t actual = 0;
  int sv = 1;
  std::vector<int> fr;
  if (!path || cmatrix::Read(path, &m, &actual, &sv, &fr)) return 1;
  if (nrows) *nrows = (int64_t)m.Nrows();
  if (actual_nrows) *actual_nrows = (int64_t)actual;
  if (sieve) *sieve = sv;
  if (tri && m.size()) memcpy(tri, m.data(), sizeof(float) * m.size());
  if (frames && !fr.empty()) memcpy(frames, fr.data(), sizeof(int) * fr.size());
  return 0;
}

int chost_write_grace(const char* path, int nsets, const char* const* legends,
                      const double* const* values, int64_t nframes, int invert) {
  if (!path || nsets < 1 || nframes < 0 || !legends || (nframes > 0 && !values)) return 1;
  std::vector<Series> sets(nsets);
  std::vector<const Series*> ptrs;
  for (int k = 0; k < nsets; ++k) {
    Series& s = sets[k];
    s.legend = legends[k] ? legends[k] : "";
    s.name = s.legend;
    if (nframes > 0) s.v.assign(values[k], values[k] + nframes);
    ptrs.push_back(&s);
  }
  return dataio::WriteGrace(path, ptrs, invert != 0);
}

int chost_write_mesh(const char* path, const char* xlabel, const double* x, int64_t n, int nsets,
                     const char* const* names, const double* const* values) {
  if (!path || !xlabel || n < 0 || nsets < 0 || (n > 0 && (!x || (nsets > 0 && !values))))
    return 1;
  std::vector<Series> sets(nsets);
  std::vector<const Series*> ptrs;
  for (int k = 0; k < nsets; ++k) {
    sets[k].name = names && names[k] ? names[k] : "";
    if (n > 0) sets[k].v.assign(values[k], values[k] + n);
    ptrs.push_back(&sets[k]);
  }
  return dataio::WriteMesh(path, xlabel, std::vector<double>(x, x + n), ptrs);
}

}  // extern "C"
