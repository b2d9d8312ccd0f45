/*
 * luna_h5.h -- C ABI of the HDF5 writer/reader behind HDF5Output (libluna_h5.so).
 *
 * The reference writes its output files through HDF5.jl (src/Output.jl:146-487,
 * Utils.save_dict_h5).  This library is the same file-level vocabulary on top of
 * the HDF5 C library (1.10): groups, whole datasets, chunked datasets that grow
 * along their last dimension, and reads.  The Output semantics themselves
 * (GridCondition saves, statistics, the continuation cache) live in the host
 * package (output.py), as they live in Output.jl.
 *
 * Conventions (those of HDF5.jl, so files are interchangeable):
 *   - dims are given in the caller's column-major order (Julia's size(A));
 *     the file stores them reversed, as HDF5.jl does;
 *   - complex values are the compound type {double r; double i;};
 *   - a dataset with ndim == 0 has a scalar dataspace;
 *   - strings are fixed-length UTF-8, null-terminated (scalar), or
 *     variable-length UTF-8 (1-D arrays of strings).
 * Paths are '/'-separated from the file root ("meta/cache/t").
 * Every call returns LH_H5_OK (0) or a negative LH_H5_E_* code.
 */
#ifndef LUNA_H5_H
#define LUNA_H5_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LH_H5_OK = 0,
    LH_H5_E_IO = -1,      /* the HDF5 library reported an error              */
    LH_H5_E_ARG = -2,     /* invalid argument                                */
    LH_H5_E_MISSING = -3, /* no object at that path                          */
    LH_H5_E_TYPE = -4,    /* element type or shape does not match            */
    LH_H5_E_SPACE = -5    /* the caller's buffer is too small                */
};

enum {
    LH_H5_F64 = 0,   /* Float64                                             */
    LH_H5_C128 = 1,  /* ComplexF64 as {r, i}                                */
    LH_H5_I64 = 2,   /* Int64                                               */
    LH_H5_U64 = 3,   /* UInt64 (cachehash)                                  */
    LH_H5_U8 = 4,    /* Bool / UInt8                                        */
    LH_H5_STR = 5    /* String: ndim 0 -> const char*, ndim 1 -> const char**/
};

/* mode 0: open read-only; 1: open read-write; 2: create (truncating).
 * HDF5.h5open(fpath, "r" | "r+" | "w") */
int lh_h5_open(const char *path, int mode, int64_t *file);
int lh_h5_close(int64_t file);

/* 1 if an object exists at path (every intermediate group too), 0 if not.
 * HDF5.haskey */
int lh_h5_exists(int64_t file, const char *path);
/* HDF5.create_group; intermediate groups are created as needed. */
int lh_h5_create_group(int64_t file, const char *path);
/* HDF5.delete_object */
int lh_h5_delete(int64_t file, const char *path);

/* parent[name] = value (a new contiguous dataset), or write(parent[name], value)
 * in place when the dataset exists with the same type and shape.
 * Missing parent groups are created. */
int lh_h5_write(int64_t file, const char *path, int type, int ndim, const int64_t *dims, const void *data);

/* HDF5.create_dataset(parent, name, datatype, (dims, maxdims), chunk=chunk)
 * with maxdims = dims except the last, which is unlimited (-1). */
int lh_h5_create_extendable(int64_t file, const char *path, int type, int ndim, const int64_t *dims,
                            const int64_t *chunk);
/* dataset[:, ..., :, index + 1] = slice (index 0-based along the last
 * dimension), growing the dataset by set_extent_dims when index is past its
 * end.  slice holds prod(dims[0:ndim-1]) elements. */
int lh_h5_write_slab(int64_t file, const char *path, int64_t index, const void *slice);

/* Element type, rank and dims (column-major order) of the dataset at path;
 * dims has room for 8 entries. */
int lh_h5_info(int64_t file, const char *path, int *type, int *ndim, int64_t *dims);
/* read(dataset) into data (nbytes available).  Strings: a scalar string is
 * copied null-terminated; a string array as consecutive null-terminated
 * strings. */
int lh_h5_read(int64_t file, const char *path, void *data, int64_t nbytes);
/* Names of the members of the group at path, each null-terminated, one after
 * another, in the library's name order; *count receives how many.  keys(group) */
int lh_h5_list(int64_t file, const char *path, char *names, int64_t nbytes, int *count);

#ifdef __cplusplus
}
#endif
#endif /* LUNA_H5_H */
