/* pffrg_h5.h -- HDF5 files in the layout PFFRGSolver.jl writes through HDF5.jl: the
 * observable file (Σ and χ per Λ, the reduced-lattice description) and the checkpoint file
 * (meshes, step size and action per Λ).  Host-side C ABI (libpffrg_h5.so, no GPU); the layout
 * functions that compose the dataset names are in pffrg/h5io.py.
 *
 * Arrays are passed with their Julia (column-major) dimensions and bytes; the file stores the
 * dimensions reversed, as HDF5.jl does, so Julia reads them back unchanged.  Paths such as
 * "a/0.5/Γ/spin/ch_s/q3" create their intermediate groups.  Every call returns 0 on success
 * (pffrg_h5_exists: 1 or 0) and PFFRG_H5_ERR on failure, with pffrg_h5_last_error. */
#ifndef PFFRG_H5_H
#define PFFRG_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFFRG_H5_ERR -1

/* open modes: "r", "r+" and HDF5.jl's "cw" (read/write, create if missing) */
#define PFFRG_H5_READ 0
#define PFFRG_H5_WRITE 1
#define PFFRG_H5_CREATE_WRITE 2

/* dataset kinds reported by pffrg_h5_shape */
#define PFFRG_H5_F64 0
#define PFFRG_H5_I64 1
#define PFFRG_H5_BOOL 2
#define PFFRG_H5_STR 3

typedef struct pffrg_h5 pffrg_h5;

/* h5open(path, mode)  (Launcher.jl:20-21) */
int pffrg_h5_open(const char *path, int32_t mode, pffrg_h5 **out);
void pffrg_h5_close(pffrg_h5 *f);
const char *pffrg_h5_last_error(const pffrg_h5 *f);

/* haskey(file, path)  (Launcher.jl:28, checkpoint_su2.jl:24-31) */
int pffrg_h5_exists(pffrg_h5 *f, const char *path);

/* file[path] = value  (checkpoint_su2.jl:11-36, Action/disk.jl:8-11, Observable/disk.jl:12-24,
 * Lattice/disk.jl:9-16); rank 0 writes a scalar.  An existing dataset is an error. */
int pffrg_h5_write_f64(pffrg_h5 *f, const char *path, int32_t rank, const int64_t *dims, const double *data);
int pffrg_h5_write_i64(pffrg_h5 *f, const char *path, int32_t rank, const int64_t *dims, const int64_t *data);
int pffrg_h5_write_bool(pffrg_h5 *f, const char *path, int32_t value);
int pffrg_h5_write_str(pffrg_h5 *f, const char *path, const char *s);

/* read(file, path)  (checkpoint_su2.jl:55-71): shape in Julia order (dims[8]) and kind, then
 * the data (n = number of elements) */
int pffrg_h5_shape(pffrg_h5 *f, const char *path, int32_t *rank, int64_t *dims, int32_t *kind);
int pffrg_h5_read_f64(pffrg_h5 *f, const char *path, int64_t n, double *data);
int pffrg_h5_read_i64(pffrg_h5 *f, const char *path, int64_t n, int64_t *data);
int pffrg_h5_read_bool(pffrg_h5 *f, const char *path, int32_t *value);
int pffrg_h5_read_str(pffrg_h5 *f, const char *path, char *buf, int64_t cap);

/* keys(file[group])  (checkpoint_su2.jl:48-51): link names in increasing name order, separated
 * by '\n'; *needed = bytes including the terminating 0 (buf may be NULL to query it) */
int pffrg_h5_keys(pffrg_h5 *f, const char *group, char *buf, int64_t cap, int64_t *needed);

#ifdef __cplusplus
}
#endif

#endif
