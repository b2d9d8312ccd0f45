/*
 * lh_h5.c -- the HDF5 file layer behind HDF5Output (declared in include/luna_h5.h).
 *
 * Host code on the HDF5 1.10 C library.  It provides what Output.jl asks of
 * HDF5.jl (src/Output.jl:146-487): groups, whole datasets written or
 * overwritten in place, chunked datasets that grow along their last
 * (column-major) dimension one slice at a time, and reads.  Dimensions cross
 * the interface in column-major order and are stored reversed, complex numbers
 * are the compound {r, i}: the layout HDF5.jl gives a Julia array.
 */
#include "../../include/luna_h5.h"

#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

#define MAXDIM 8

static void quiet(void) {
    static int done = 0;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, NULL, NULL);  /* errors become return codes */
        done = 1;
    }
}

/* Memory/file type of an element kind; the caller closes the result. */
static hid_t make_type(int type, int ndim) {
    hid_t t = -1;
    switch (type) {
        case LH_H5_F64: t = H5Tcopy(H5T_NATIVE_DOUBLE); break;
        case LH_H5_I64: t = H5Tcopy(H5T_NATIVE_INT64); break;
        case LH_H5_U64: t = H5Tcopy(H5T_NATIVE_UINT64); break;
        case LH_H5_U8: t = H5Tcopy(H5T_NATIVE_UINT8); break;
        case LH_H5_C128:
            t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
            if (t >= 0) {
                H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE);
                H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE);
            }
            break;
        case LH_H5_STR:
            t = H5Tcopy(H5T_C_S1);
            if (t >= 0) {
                H5Tset_cset(t, H5T_CSET_UTF8);
                if (ndim > 0) H5Tset_size(t, H5T_VARIABLE);
            }
            break;
        default: break;
    }
    return t;
}

static size_t elem_size(int type) {
    switch (type) {
        case LH_H5_F64: case LH_H5_I64: case LH_H5_U64: return 8;
        case LH_H5_U8: return 1;
        case LH_H5_C128: return 16;
        default: return 0;
    }
}

/* Element kind of a dataset's stored type, or -1. */
static int kind_of(hid_t dt) {
    H5T_class_t c = H5Tget_class(dt);
    size_t sz = H5Tget_size(dt);
    if (c == H5T_FLOAT && sz == 8) return LH_H5_F64;
    if (c == H5T_INTEGER) {
        H5T_sign_t sg = H5Tget_sign(dt);
        if (sz == 8) return sg == H5T_SGN_NONE ? LH_H5_U64 : LH_H5_I64;
        if (sz == 1) return LH_H5_U8;
        return -1;
    }
    if (c == H5T_COMPOUND && sz == 16 && H5Tget_nmembers(dt) == 2) {
        char* a = H5Tget_member_name(dt, 0);
        char* b = H5Tget_member_name(dt, 1);
        int ok = a && b && strcmp(a, "r") == 0 && strcmp(b, "i") == 0;
        H5free_memory(a);
        H5free_memory(b);
        return ok ? LH_H5_C128 : -1;
    }
    if (c == H5T_STRING) return LH_H5_STR;
    return -1;
}

/* Create the missing groups on the way to path's last component. */
static int ensure_parents(hid_t f, const char* path) {
    char buf[1024];
    size_t n = strlen(path);
    if (n == 0 || n >= sizeof(buf)) return LH_H5_E_ARG;
    memcpy(buf, path, n + 1);
    for (size_t i = 1; i < n; i++) {
        if (buf[i] != '/') continue;
        buf[i] = '\0';
        htri_t e = H5Lexists(f, buf, H5P_DEFAULT);
        if (e < 0) return LH_H5_E_IO;
        if (e == 0) {
            hid_t g = H5Gcreate2(f, buf, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (g < 0) return LH_H5_E_IO;
            H5Gclose(g);
        }
        buf[i] = '/';
    }
    return LH_H5_OK;
}

int lh_h5_open(const char* path, int mode, int64_t* file) {
    if (!path || !file || mode < 0 || mode > 2) return LH_H5_E_ARG;
    quiet();
    hid_t f;
    if (mode == 2)
        f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    else
        f = H5Fopen(path, mode == 1 ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return LH_H5_E_IO;
    *file = (int64_t)f;
    return LH_H5_OK;
}

int lh_h5_close(int64_t file) {
    return H5Fclose((hid_t)file) < 0 ? LH_H5_E_IO : LH_H5_OK;
}

int lh_h5_exists(int64_t file, const char* path) {
    if (!path || !path[0]) return LH_H5_E_ARG;
    char buf[1024];
    size_t n = strlen(path);
    if (n >= sizeof(buf)) return LH_H5_E_ARG;
    memcpy(buf, path, n + 1);
    for (size_t i = 1; i <= n; i++) {
        if (buf[i] != '/' && buf[i] != '\0') continue;
        char keep = buf[i];
        buf[i] = '\0';
        htri_t e = H5Lexists((hid_t)file, buf, H5P_DEFAULT);
        buf[i] = keep;
        if (e < 0) return LH_H5_E_IO;
        if (e == 0) return 0;
    }
    return 1;
}

int lh_h5_create_group(int64_t file, const char* path) {
    if (!path) return LH_H5_E_ARG;
    int st = ensure_parents((hid_t)file, path);
    if (st) return st;
    hid_t g = H5Gcreate2((hid_t)file, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return LH_H5_E_IO;
    H5Gclose(g);
    return LH_H5_OK;
}

int lh_h5_delete(int64_t file, const char* path) {
    if (!path) return LH_H5_E_ARG;
    if (lh_h5_exists(file, path) != 1) return LH_H5_E_MISSING;
    return H5Ldelete((hid_t)file, path, H5P_DEFAULT) < 0 ? LH_H5_E_IO : LH_H5_OK;
}

static void reversed(int ndim, const int64_t* dims, hsize_t* out) {
    for (int i = 0; i < ndim; i++) out[i] = (hsize_t)dims[ndim - 1 - i];
}

/* Scalar string: fixed-length type sized to the text (at least 1). */
static int write_string(hid_t f, const char* path, const char* text) {
    size_t len = strlen(text);
    hid_t t = make_type(LH_H5_STR, 0);
    H5Tset_size(t, len > 0 ? len + 1 : 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    hid_t sp = H5Screate(H5S_SCALAR);
    int st = LH_H5_OK;
    if (H5Lexists(f, path, H5P_DEFAULT) > 0 && H5Ldelete(f, path, H5P_DEFAULT) < 0) st = LH_H5_E_IO;
    hid_t d = st ? -1 : H5Dcreate2(f, path, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, text) < 0) st = LH_H5_E_IO;
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    H5Tclose(t);
    return st;
}

int lh_h5_write(int64_t file, const char* path, int type, int ndim, const int64_t* dims, const void* data) {
    if (!path || !data || ndim < 0 || ndim > MAXDIM || (ndim > 0 && !dims)) return LH_H5_E_ARG;
    if (type < 0 || type > LH_H5_STR || (type == LH_H5_STR && ndim > 1)) return LH_H5_E_ARG;
    hid_t f = (hid_t)file;
    int st = ensure_parents(f, path);
    if (st) return st;
    if (type == LH_H5_STR && ndim == 0) return write_string(f, path, (const char*)data);
    hsize_t hd[MAXDIM];
    reversed(ndim, dims, hd);
    hid_t t = make_type(type, ndim);
    if (t < 0) return LH_H5_E_IO;
    hid_t d = -1;
    htri_t ex = H5Lexists(f, path, H5P_DEFAULT);
    if (ex > 0) {
        /* overwrite in place when type and shape agree, else replace */
        d = H5Dopen2(f, path, H5P_DEFAULT);
        int same = 0;
        if (d >= 0) {
            hid_t dt = H5Dget_type(d), sp = H5Dget_space(d);
            int nd = H5Sget_simple_extent_ndims(sp);
            hsize_t cur[MAXDIM];
            if (kind_of(dt) == type && nd == ndim && type != LH_H5_STR) {
                H5Sget_simple_extent_dims(sp, cur, NULL);
                same = 1;
                for (int i = 0; i < nd; i++) same &= (cur[i] == hd[i]);
            }
            H5Tclose(dt);
            H5Sclose(sp);
            if (!same) {
                H5Dclose(d);
                d = -1;
                if (H5Ldelete(f, path, H5P_DEFAULT) < 0) st = LH_H5_E_IO;
            }
        }
    }
    if (!st && d < 0) {
        hid_t sp = ndim == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(ndim, hd, NULL);
        d = H5Dcreate2(f, path, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5Sclose(sp);
    }
    if (!st && (d < 0 || H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)) st = LH_H5_E_IO;
    if (d >= 0) H5Dclose(d);
    H5Tclose(t);
    return st;
}

int lh_h5_create_extendable(int64_t file, const char* path, int type, int ndim, const int64_t* dims,
                            const int64_t* chunk) {
    if (!path || !dims || !chunk || ndim < 1 || ndim > MAXDIM || type == LH_H5_STR || type < 0 || type > LH_H5_U8)
        return LH_H5_E_ARG;
    hid_t f = (hid_t)file;
    int st = ensure_parents(f, path);
    if (st) return st;
    hsize_t hd[MAXDIM], hm[MAXDIM], hc[MAXDIM];
    reversed(ndim, dims, hd);
    reversed(ndim, chunk, hc);
    for (int i = 0; i < ndim; i++) hm[i] = hd[i];
    hm[0] = H5S_UNLIMITED;  /* the last column-major dimension */
    hid_t sp = H5Screate_simple(ndim, hd, hm);
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, ndim, hc);
    hid_t t = make_type(type, ndim);
    hid_t d = H5Dcreate2(f, path, t, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    if (d < 0) st = LH_H5_E_IO;
    else H5Dclose(d);
    H5Tclose(t);
    H5Pclose(pl);
    H5Sclose(sp);
    return st;
}

int lh_h5_write_slab(int64_t file, const char* path, int64_t index, const void* slice) {
    if (!path || !slice || index < 0) return LH_H5_E_ARG;
    hid_t f = (hid_t)file;
    if (lh_h5_exists(file, path) != 1) return LH_H5_E_MISSING;
    hid_t d = H5Dopen2(f, path, H5P_DEFAULT);
    if (d < 0) return LH_H5_E_IO;
    int st = LH_H5_OK;
    hid_t dt = H5Dget_type(d);
    int kind = kind_of(dt);
    H5Tclose(dt);
    hid_t sp = H5Dget_space(d);
    int nd = H5Sget_simple_extent_ndims(sp);
    hsize_t cur[MAXDIM], mx[MAXDIM];
    H5Sget_simple_extent_dims(sp, cur, mx);
    H5Sclose(sp);
    if (kind < 0 || kind == LH_H5_STR || nd < 1 || nd > MAXDIM) st = LH_H5_E_TYPE;
    if (!st && (hsize_t)index >= cur[0]) {
        if (mx[0] != H5S_UNLIMITED && (hsize_t)index >= mx[0]) st = LH_H5_E_SPACE;
        cur[0] = (hsize_t)index + 1;
        if (!st && H5Dset_extent(d, cur) < 0) st = LH_H5_E_IO;
    }
    if (!st) {
        hid_t fs = H5Dget_space(d);
        hsize_t start[MAXDIM] = {0}, count[MAXDIM];
        start[0] = (hsize_t)index;
        count[0] = 1;
        for (int i = 1; i < nd; i++) count[i] = cur[i];
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, count, NULL);
        hid_t ms = H5Screate_simple(nd, count, NULL);
        hid_t t = make_type(kind, nd);
        if (H5Dwrite(d, t, ms, fs, H5P_DEFAULT, slice) < 0) st = LH_H5_E_IO;
        H5Tclose(t);
        H5Sclose(ms);
        H5Sclose(fs);
    }
    H5Dclose(d);
    return st;
}

int lh_h5_info(int64_t file, const char* path, int* type, int* ndim, int64_t* dims) {
    if (!path || !type || !ndim || !dims) return LH_H5_E_ARG;
    if (lh_h5_exists(file, path) != 1) return LH_H5_E_MISSING;
    H5O_info_t oi;
    if (H5Oget_info_by_name2((hid_t)file, path, &oi, H5O_INFO_BASIC, H5P_DEFAULT) < 0) return LH_H5_E_IO;
    if (oi.type != H5O_TYPE_DATASET) return LH_H5_E_TYPE;
    hid_t d = H5Dopen2((hid_t)file, path, H5P_DEFAULT);
    if (d < 0) return LH_H5_E_IO;
    hid_t dt = H5Dget_type(d), sp = H5Dget_space(d);
    int nd = H5Sget_simple_extent_ndims(sp);
    hsize_t cur[MAXDIM];
    int st = LH_H5_OK;
    if (nd < 0 || nd > MAXDIM) st = LH_H5_E_TYPE;
    else {
        H5Sget_simple_extent_dims(sp, cur, NULL);
        *ndim = nd;
        for (int i = 0; i < nd; i++) dims[i] = (int64_t)cur[nd - 1 - i];
        *type = kind_of(dt);
        if (*type < 0) st = LH_H5_E_TYPE;
    }
    H5Tclose(dt);
    H5Sclose(sp);
    H5Dclose(d);
    return st;
}

int lh_h5_read(int64_t file, const char* path, void* data, int64_t nbytes) {
    if (!path || !data || nbytes < 0) return LH_H5_E_ARG;
    int type, nd;
    int64_t dims[MAXDIM];
    int st = lh_h5_info(file, path, &type, &nd, dims);
    if (st) return st;
    int64_t n = 1;
    for (int i = 0; i < nd; i++) n *= dims[i];
    hid_t d = H5Dopen2((hid_t)file, path, H5P_DEFAULT);
    if (d < 0) return LH_H5_E_IO;
    if (type == LH_H5_STR) {
        hid_t dt = H5Dget_type(d);
        if (H5Tis_variable_str(dt) > 0) {
            char** v = (char**)calloc((size_t)(n > 0 ? n : 1), sizeof(char*));
            hid_t mt = make_type(LH_H5_STR, 1);
            if (!v || H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) st = LH_H5_E_IO;
            int64_t used = 0;
            for (int64_t i = 0; i < n && !st; i++) {
                size_t len = v[i] ? strlen(v[i]) : 0;
                if (used + (int64_t)len + 1 > nbytes) {
                    st = LH_H5_E_SPACE;
                    break;
                }
                if (len) memcpy((char*)data + used, v[i], len);
                ((char*)data)[used + len] = '\0';
                used += (int64_t)len + 1;
            }
            if (v && st != LH_H5_E_IO) {
                hid_t sp = H5Dget_space(d);
                H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, v);
                H5Sclose(sp);
            }
            free(v);
            H5Tclose(mt);
        } else {
            size_t sz = H5Tget_size(dt);
            if ((int64_t)(sz * (size_t)n + 1) > nbytes) st = LH_H5_E_SPACE;
            else {
                hid_t mt = H5Tcopy(dt);
                char* tmp = (char*)calloc((size_t)n * sz + 1, 1);
                if (!tmp || H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp) < 0) st = LH_H5_E_IO;
                else {
                    /* n fixed-length strings, each padded to sz */
                    int64_t used = 0;
                    for (int64_t i = 0; i < n; i++) {
                        size_t len = strnlen(tmp + i * sz, sz);
                        memcpy((char*)data + used, tmp + i * sz, len);
                        ((char*)data)[used + len] = '\0';
                        used += (int64_t)len + 1;
                    }
                }
                free(tmp);
                H5Tclose(mt);
            }
        }
        H5Tclose(dt);
    } else {
        if ((int64_t)elem_size(type) * n > nbytes) st = LH_H5_E_SPACE;
        else {
            hid_t mt = make_type(type, nd);
            if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) st = LH_H5_E_IO;
            H5Tclose(mt);
        }
    }
    H5Dclose(d);
    return st;
}

struct list_acc {
    char* out;
    int64_t cap, used;
    int count, overflow;
};

static herr_t list_one(hid_t g, const char* name, const H5L_info_t* info, void* op) {
    (void)g;
    (void)info;
    struct list_acc* a = (struct list_acc*)op;
    size_t len = strlen(name);
    if (a->used + (int64_t)len + 1 > a->cap) {
        a->overflow = 1;
        return 1;
    }
    memcpy(a->out + a->used, name, len + 1);
    a->used += (int64_t)len + 1;
    a->count++;
    return 0;
}

int lh_h5_list(int64_t file, const char* path, char* names, int64_t nbytes, int* count) {
    if (!path || !names || !count || nbytes < 0) return LH_H5_E_ARG;
    if (!(path[0] == '/' && path[1] == '\0') && lh_h5_exists(file, path) != 1) return LH_H5_E_MISSING;
    hid_t g = H5Gopen2((hid_t)file, path, H5P_DEFAULT);
    if (g < 0) return LH_H5_E_TYPE;
    struct list_acc a = {names, nbytes, 0, 0, 0};
    hsize_t idx = 0;
    herr_t r = H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, list_one, &a);
    H5Gclose(g);
    if (a.overflow) return LH_H5_E_SPACE;
    if (r < 0) return LH_H5_E_IO;
    *count = a.count;
    return LH_H5_OK;
}
