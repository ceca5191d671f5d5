/*
 * lcs_io.h -- C ABI of the restart and statistics files of the MI355X
 * backend: the on-disk format of LCS.jl's LCSIO (src/LCSIO/save.jl,
 * src/LCSIO/load.jl, src/LCSIO/hdf5.jl), written and read through the HDF5
 * C library.
 *
 * Built as its own library, liblcs_io.so, linked against libhdf5 and against
 * liblcs_mi355x.so (whose public calls move fields and particles between
 * the device and the host), so the compute library has no HDF5 dependency.
 *
 * Files hold what the reference's files hold, one process's view:
 *   restart.h5.<step>:  flow/{U,V,W,P}                 global nx*ny*nz arrays,
 *                                                      no halo, x fastest
 *                       particles/<i>/counts           Int64[1] = nvalid
 *                       particles/<i>/{id,x,y,z,u,v,w,diam}  nvalid elements
 *   stat.h5:            <step>/{state,stat,log}/...    nested groups of scalars
 *                                                      and vectors
 * Array shapes follow HDF5.jl, which stores a Julia (column-major) array of
 * size (nx, ny, nz) under the HDF5 dataspace (nz, ny, nx): the bytes on disk
 * are the column-major array.  Field datasets have the box's float type;
 * reads convert whatever float type the file holds.
 *
 * Every function returns LCS_OK (0) or a negative error class of
 * lcs_mi355x.h and sets a thread-local message readable with
 * lcs_io_last_error().  Files are opened and closed inside each call.
 */
#ifndef LCS_IO_H
#define LCS_IO_H

#include "lcs_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* element types of stat entries and of lcs_h5_read */
#define LCS_H5_I64 0
#define LCS_H5_F64 1
#define LCS_H5_F32 2
#define LCS_H5_STR 3 /* fixed-length UTF-8 string */

/* what lcs_h5_info finds at a path */
#define LCS_H5_NONE 0
#define LCS_H5_GROUP 1
#define LCS_H5_DATASET 2

/* One dataset to write: a leaf of Utils.deepwrite (src/utils/hdf5.jl:5-19)
 * or one array of a restart file, at a '/'-separated relative path (e.g.
 * "stat/flow/ke", "flow/U"); missing groups on the path are created. */
typedef struct {
    const char *path;
    int type;        /* LCS_H5_* */
    int rank;        /* 0 (scalar) to 3 */
    int64_t dims[3]; /* size in the reference's (Julia) order, dims[0] fastest;
                        for a LCS_H5_STR scalar dims[0] is the string's bytes */
    const void *data; /* column-major (dims[0] fastest), the host's byte order */
} lcs_h5_entry;

/* LCSIO.saverestart(file, buf, state, config, topo) (src/LCSIO/save.jl:40-56):
 * creates or opens file, writes flow/{U,V,W,P} from the box interior and,
 * per particle profile i (1-based group names), particles/<i>/counts and the
 * first nvalid elements of id, x, y, z, u, v, w, diam (no property dataset
 * when nvalid is 0, hdf5.jl:36).  A file that already holds "flow" is an
 * error, as create_group is in the reference.  Synchronises the box. */
int lcs_save_restart(lcs_box *box, const char *file);

/* LCSIO.load_flow_fields! (src/LCSIO/load.jl:23-42): reads flow/{U,V,W,P}
 * into the box interior (each dataset must be nx*ny*nz, hdf5.jl:28), then
 * fills every ghost cell of U, V, W, P periodically (FullHalo). */
int lcs_load_flow_fields(lcs_box *box, const char *file);

/* LCSIO.load_particle_counts! + load_particle_{id,position,velocity,size}!
 * (src/LCSIO/load.jl:44-118) for profile (0-based; group particles/<profile+1>).
 * counts must have one entry (one process, load.jl:60) and it must equal the
 * profile's nvalid: a profile's population is fixed when it is added. */
int lcs_load_particles(lcs_box *box, const char *file, int profile);

/* LCSIO.savestat(file, state, topo; stat, log) (src/LCSIO/save.jl:14-33):
 * creates or opens file, replaces group "<step>" and writes every entry
 * below it.  Entries with no elements are skipped, as deepwrite skips empty
 * arrays (src/utils/hdf5.jl:10-12). */
int lcs_save_stat(const char *file, int64_t step, const lcs_h5_entry *entries, int n);

/* Writes entries into file (created when absent) from host memory, e.g. a
 * restart file of host arrays: flow/U as F64 or F32 of dims (nx, ny, nz),
 * particles/1/counts as I64 of dims (1,), ...  An existing dataset at an
 * entry's path is an error.  Entries with no elements are skipped, as
 * write_property skips empty properties (src/LCSIO/hdf5.jl:36). */
int lcs_h5_write(const char *file, const lcs_h5_entry *entries, int n);

/* h[path] lookups for readers (HDF5.jl haskey/size/eltype): *kind is
 * LCS_H5_NONE, _GROUP or _DATASET; for a dataset *type is its LCS_H5_* class
 * (-1 if none applies), *rank its rank (at most 8) and dims[0..rank-1] its
 * size in the reference's (Julia) order, dims[0] fastest.  type, rank and
 * dims may be NULL. */
int lcs_h5_info(const char *file, const char *path, int *kind, int *type, int *rank, int64_t *dims);

/* read(h[path]) for a numeric dataset, converted to type (LCS_H5_I64, _F64,
 * _F32), into out, which holds capacity elements; or, for LCS_H5_STR, a
 * scalar string dataset into out as a NUL-terminated string of at most
 * capacity bytes. */
int lcs_h5_read(const char *file, const char *path, int type, void *out, int64_t capacity);

/* keys(h[path]) in name order, as NUL-terminated names packed into names
 * (len bytes); *n receives the number of keys.  LCS_ERR_CAPACITY when they
 * do not fit (then *n is still set and *needed, if not NULL, the bytes). */
int lcs_h5_keys(const char *file, const char *path, char *names, size_t len, int *n, size_t *needed);

const char *lcs_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LCS_IO_H */
