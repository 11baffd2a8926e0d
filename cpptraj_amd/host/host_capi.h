/*
 * host_capi.h -- C entry points of the host layer (libcpptraj_rmsd_host.so),
 * for bindings that cannot use the C++ classes of host.hpp directly (the
 * Python tests).  A session holds named COORDS sets, reference frames, a mask
 * table (expression -> resolved selection) and the data sets that commands
 * produce; chost_run executes one command line of the reference's syntax
 * ("rms2d ...", "2drms ...", "rms ...", "rmsd ...", "align ...", "symmrmsd ...",
 * "rmsavgcorr ...", "cluster ...").
 */
#ifndef CPPTRAJ_RMSD_HOST_CAPI_H
#define CPPTRAJ_RMSD_HOST_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chost_session chost_session;

int chost_create(int device, chost_session** s);
void chost_destroy(chost_session* s);
const char* chost_last_error(const chost_session* s);
/* COORDS set: nframes frames of `stride` floats (3*natom coordinates first),
 * natom masses (NULL: all 1). */
int chost_add_coords(chost_session* s, const char* name, const float* xyz, int64_t nframes,
                     int64_t stride, int natom, const double* mass);
/* Unit cells of a COORDS set's frames: 9 doubles per frame, the rows being
 * the three cell vectors (Box::UnitCell).  Actions that rotate the frames
 * rotate the cells with them (Box::RotateUcell).  chost_get_box: NULL out
 * gives the frame count (0 without a box). */
int chost_set_box(chost_session* s, const char* name, const double* ucell, int64_t nframes);
int chost_get_box(chost_session* s, const char* name, double* out, int64_t* nframes);
/* Reference frame ('ref <name>' / 'reference'): one frame of 3*natom floats. */
int chost_add_reference(chost_session* s, const char* name, const float* xyz, int natom,
                        const double* mass);
/* Residues of a set's topology (COORDS set or reference): nres first atoms
 * (0-based, increasing), names, and the solute residues (0-based; NULL / 0:
 * every residue) that Topology::SoluteResidues would list. */
int chost_set_residues(chost_session* s, const char* set, int nres, const int* first,
                       const char* const* names, int nsolute, const int* solute);
/* Symmetric atom groups of a set's topology: group k is atoms[off[k] ..
 * off[k+1]) (0-based atom indices), as AtomMap::SymmetricAtoms finds them
 * residue by residue ('symmrmsd' reduces them to its selection). */
int chost_set_symmetric_groups(chost_session* s, const char* set, int ngroups, const int* off,
                               const int* atoms);
/* Mask expression -> selected atom indices, for the set (COORDS or reference)
 * named `set`, or for every set when `set` is NULL or empty. */
int chost_add_mask(chost_session* s, const char* set, const char* expr, const int* sel, int nsel);
int chost_run(chost_session* s, const char* line);
/* Data set queries.  Pass NULL buffers to get the sizes first. */
int chost_get_matrix(chost_session* s, const char* name, float* out, int64_t* ncols,
                     int64_t* nrows, int* is_tri);
int chost_get_series(chost_session* s, const char* name, double* out, int64_t* nframes,
                     int* ncomp);
int chost_get_coords(chost_session* s, const char* name, float* out, int64_t* nfloats);
/* Writers of the host layer, usable without a GPU: a float matrix (packed
 * triangle of n x n when is_tri, else nrows x ncols row-major) in DataIO_Std's
 * square2d or nosquare2d text; per-frame series (nsets sets of ncomp[s] values
 * per frame, widths / precisions per set, names) in its 1D text; a Cmatrix
 * file (triangle over `frames` (nrows entries) of a run over actual_nrows
 * frames with the given sieve). */
int chost_write_matrix(const char* path, const float* data, int64_t ncols, int64_t nrows,
                       int is_tri, int square2d, const char* name);
int chost_write_series(const char* path, int nsets, const char* const* names, const int* ncomp,
                       const int* width, const int* prec, const double* const* values,
                       int64_t nframes);
int chost_write_cmatrix(const char* path, const float* tri, int64_t nrows, int64_t actual_nrows,
                        int sieve, const int* frames);
/* Grace xy file of nsets per-frame series (one value per frame, %12.4f),
 * normal or inverted (DataIO_Grace). */
int chost_write_grace(const char* path, int nsets, const char* const* legends,
                      const double* const* values, int64_t nframes, int invert);
/* XY mesh text (x %8.3f, then each set %12.4f) with header "#<xlabel>" and
 * the set names. */
int chost_write_mesh(const char* path, const char* xlabel, const double* x, int64_t n, int nsets,
                     const char* const* names, const double* const* values);
/* Cmatrix file reader: fills sizes, then (second call) the triangle and the
 * cached frame list. */
int chost_read_cmatrix(const char* path, float* tri, int64_t* nrows, int64_t* actual_nrows,
                       int* sieve, int* frames);

#ifdef __cplusplus
}
#endif

#endif
