// h5io.cpp -- HDF5 files with the conventions of HDF5.jl, in which PFFRGSolver.jl writes its
// observable and checkpoint files (Launcher.jl:2-75, checkpoint_su2.jl:2-79, Action/disk.jl,
// Observable/disk.jl:2-29, Lattice/disk.jl:2-19).  Host code only (no GPU), over the HDF5 C API.
//
// Conventions of HDF5.jl that the reference's files carry:
//  - an array is stored with its Julia dimensions reversed, so that the bytes are the Julia
//    (column-major) array's; callers pass the Julia dimensions;
//  - a scalar is a scalar dataspace; Float64 is IEEE F64LE, Int64 STD I64LE, Bool an 8-bit
//    bitfield, String a fixed-length UTF-8 string of its byte length (null-terminated padding);
//  - writing "a/b/c" creates the intermediate groups;
//  - keys(group) lists the links in increasing name order.
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pffrg_h5.h"

struct pffrg_h5 {
    hid_t file = -1;
    std::string err;
};

namespace {

constexpr int MAX_RANK = 8;

int fail(pffrg_h5 *f, const std::string &msg) {
    if (f) f->err = msg;
    return PFFRG_H5_ERR;
}

// silence HDF5's own error printing; every failure is reported through the return code
struct Quiet {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// 1 if every component of the path exists, 0 if not, < 0 on error
int path_exists(hid_t file, const char *path) {
    std::string p(path), prefix;
    size_t pos = 0;
    while (pos <= p.size()) {
        size_t next = p.find('/', pos);
        if (next == std::string::npos) next = p.size();
        if (next > pos) {
            prefix += (prefix.empty() ? "" : "/") + p.substr(pos, next - pos);
            const htri_t e = H5Lexists(file, prefix.c_str(), H5P_DEFAULT);
            if (e <= 0) return e < 0 ? -1 : 0;
        }
        pos = next + 1;
    }
    return 1;
}

hid_t space_of(int rank, const int64_t *dims) {
    if (rank == 0) return H5Screate(H5S_SCALAR);
    hsize_t d[MAX_RANK];
    for (int k = 0; k < rank; ++k) d[k] = (hsize_t)dims[rank - 1 - k];  // Julia dims reversed
    return H5Screate_simple(rank, d, nullptr);
}

int write_dataset(pffrg_h5 *f, const char *path, hid_t ftype, hid_t mtype, hid_t space, const void *data) {
    if (path_exists(f->file, path) == 1) return fail(f, std::string("dataset exists: ") + path);
    const hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    const hid_t ds = H5Dcreate2(f->file, path, ftype, space, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    H5Pclose(lcpl);
    if (ds < 0) return fail(f, std::string("cannot create ") + path);
    const herr_t w = H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    if (w < 0) return fail(f, std::string("cannot write ") + path);
    return 0;
}

int write_numeric(pffrg_h5 *f, const char *path, int rank, const int64_t *dims, hid_t ftype, hid_t mtype,
                  const void *data) {
    if (!f || f->file < 0 || !path || !data || rank < 0 || rank > MAX_RANK || (rank > 0 && !dims))
        return fail(f, "invalid arguments");
    Quiet q;
    const hid_t sp = space_of(rank, dims);
    if (sp < 0) return fail(f, "invalid dimensions");
    const int rc = write_dataset(f, path, ftype, mtype, sp, data);
    H5Sclose(sp);
    return rc;
}

int read_numeric(pffrg_h5 *f, const char *path, hid_t mtype, int64_t n, void *data) {
    if (!f || f->file < 0 || !path || !data) return fail(f, "invalid arguments");
    Quiet q;
    const hid_t ds = H5Dopen2(f->file, path, H5P_DEFAULT);
    if (ds < 0) return fail(f, std::string("no dataset ") + path);
    const hid_t sp = H5Dget_space(ds);
    const hssize_t np = H5Sget_simple_extent_npoints(sp);
    H5Sclose(sp);
    int rc = 0;
    if (np != n) rc = fail(f, std::string("size mismatch reading ") + path);
    else if (H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) rc = fail(f, std::string("cannot read ") + path);
    H5Dclose(ds);
    return rc;
}

struct KeyList {
    std::vector<std::string> names;
};

herr_t collect(hid_t, const char *name, const H5L_info_t *, void *op) {
    static_cast<KeyList *>(op)->names.emplace_back(name);
    return 0;
}

}  // namespace

extern "C" {

int pffrg_h5_open(const char *path, int32_t mode, pffrg_h5 **out) {
    if (!path || !out || mode < 0 || mode > 2) return PFFRG_H5_ERR;
    *out = nullptr;
    Quiet q;
    hid_t id = -1;
    if (mode == PFFRG_H5_READ) {
        id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    } else {
        id = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
        if (id < 0 && mode == PFFRG_H5_CREATE_WRITE) id = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    }
    if (id < 0) return PFFRG_H5_ERR;
    pffrg_h5 *f = new pffrg_h5();
    f->file = id;
    *out = f;
    return 0;
}

void pffrg_h5_close(pffrg_h5 *f) {
    if (!f) return;
    if (f->file >= 0) H5Fclose(f->file);
    delete f;
}

const char *pffrg_h5_last_error(const pffrg_h5 *f) { return f ? f->err.c_str() : "null file"; }

int pffrg_h5_exists(pffrg_h5 *f, const char *path) {
    if (!f || f->file < 0 || !path) return fail(f, "invalid arguments");
    Quiet q;
    const int e = path_exists(f->file, path);
    return e < 0 ? fail(f, std::string("cannot look up ") + path) : e;
}

int pffrg_h5_write_f64(pffrg_h5 *f, const char *path, int32_t rank, const int64_t *dims, const double *data) {
    return write_numeric(f, path, rank, dims, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, data);
}

int pffrg_h5_write_i64(pffrg_h5 *f, const char *path, int32_t rank, const int64_t *dims, const int64_t *data) {
    return write_numeric(f, path, rank, dims, H5T_STD_I64LE, H5T_NATIVE_INT64, data);
}

int pffrg_h5_write_bool(pffrg_h5 *f, const char *path, int32_t value) {
    const uint8_t v = value ? 1 : 0;
    return write_numeric(f, path, 0, nullptr, H5T_NATIVE_B8, H5T_NATIVE_B8, &v);
}

int pffrg_h5_write_str(pffrg_h5 *f, const char *path, const char *s) {
    if (!f || f->file < 0 || !path || !s) return fail(f, "invalid arguments");
    Quiet q;
    const size_t len = std::strlen(s);
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, len > 0 ? len : 1);
    H5Tset_cset(t, H5T_CSET_UTF8);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    const hid_t sp = H5Screate(H5S_SCALAR);
    std::vector<char> buf(len > 0 ? len : 1, '\0');
    std::memcpy(buf.data(), s, len);
    const int rc = write_dataset(f, path, t, t, sp, buf.data());
    H5Sclose(sp);
    H5Tclose(t);
    return rc;
}

int pffrg_h5_shape(pffrg_h5 *f, const char *path, int32_t *rank, int64_t *dims, int32_t *kind) {
    if (!f || f->file < 0 || !path || !rank || !dims || !kind) return fail(f, "invalid arguments");
    Quiet q;
    const hid_t ds = H5Dopen2(f->file, path, H5P_DEFAULT);
    if (ds < 0) return fail(f, std::string("no dataset ") + path);
    const hid_t sp = H5Dget_space(ds), t = H5Dget_type(ds);
    const int r = H5Sget_simple_extent_ndims(sp);
    hsize_t d[MAX_RANK];
    int rc = 0;
    if (r < 0 || r > MAX_RANK) {
        rc = fail(f, std::string("unsupported rank in ") + path);
    } else {
        H5Sget_simple_extent_dims(sp, d, nullptr);
        *rank = r;
        for (int k = 0; k < r; ++k) dims[k] = (int64_t)d[r - 1 - k];  // back to Julia order
        switch (H5Tget_class(t)) {
            case H5T_FLOAT: *kind = PFFRG_H5_F64; break;
            case H5T_INTEGER: *kind = PFFRG_H5_I64; break;
            case H5T_BITFIELD: *kind = PFFRG_H5_BOOL; break;
            case H5T_STRING: *kind = PFFRG_H5_STR; break;
            default: rc = fail(f, std::string("unsupported type in ") + path);
        }
    }
    H5Tclose(t);
    H5Sclose(sp);
    H5Dclose(ds);
    return rc;
}

int pffrg_h5_read_f64(pffrg_h5 *f, const char *path, int64_t n, double *data) {
    return read_numeric(f, path, H5T_NATIVE_DOUBLE, n, data);
}

int pffrg_h5_read_i64(pffrg_h5 *f, const char *path, int64_t n, int64_t *data) {
    return read_numeric(f, path, H5T_NATIVE_INT64, n, data);
}

int pffrg_h5_read_bool(pffrg_h5 *f, const char *path, int32_t *value) {
    uint8_t v = 0;
    const int rc = read_numeric(f, path, H5T_NATIVE_B8, 1, &v);
    if (rc == 0 && value) *value = v ? 1 : 0;
    return rc;
}

int pffrg_h5_read_str(pffrg_h5 *f, const char *path, char *buf, int64_t cap) {
    if (!f || f->file < 0 || !path || !buf || cap < 1) return fail(f, "invalid arguments");
    Quiet q;
    const hid_t ds = H5Dopen2(f->file, path, H5P_DEFAULT);
    if (ds < 0) return fail(f, std::string("no dataset ") + path);
    const hid_t t = H5Dget_type(ds);
    int rc = 0;
    if (H5Tget_class(t) != H5T_STRING) {
        rc = fail(f, std::string("not a string: ") + path);
    } else if (H5Tis_variable_str(t) > 0) {
        const hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(t));
        char *s = nullptr;
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s) < 0 || !s) {
            rc = fail(f, std::string("cannot read ") + path);
        } else {
            const size_t len = std::strlen(s);
            if ((int64_t)len + 1 > cap) rc = fail(f, "buffer too small");
            else std::memcpy(buf, s, len + 1);
            H5free_memory(s);
        }
        H5Tclose(mt);
    } else {
        const size_t size = H5Tget_size(t);
        std::vector<char> tmp(size + 1, '\0');
        const hid_t mt = H5Tcopy(t);  // same size and character set; read as stored
        H5Tset_strpad(mt, H5T_STR_NULLPAD);
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()) < 0) {
            rc = fail(f, std::string("cannot read ") + path);
        } else {
            const size_t len = std::strlen(tmp.data());
            if ((int64_t)len + 1 > cap) rc = fail(f, "buffer too small");
            else std::memcpy(buf, tmp.data(), len + 1);
        }
        H5Tclose(mt);
    }
    H5Tclose(t);
    H5Dclose(ds);
    return rc;
}

int pffrg_h5_keys(pffrg_h5 *f, const char *group, char *buf, int64_t cap, int64_t *needed) {
    if (!f || f->file < 0 || !group || !needed) return fail(f, "invalid arguments");
    Quiet q;
    const hid_t g = H5Gopen2(f->file, group, H5P_DEFAULT);
    if (g < 0) return fail(f, std::string("no group ") + group);
    KeyList kl;
    hsize_t idx = 0;
    const herr_t it = H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, collect, &kl);
    H5Gclose(g);
    if (it < 0) return fail(f, std::string("cannot list ") + group);
    std::string all;
    for (size_t k = 0; k < kl.names.size(); ++k) all += (k ? "\n" : "") + kl.names[k];
    *needed = (int64_t)all.size() + 1;
    if (buf && cap >= *needed) std::memcpy(buf, all.c_str(), all.size() + 1);
    return 0;
}

}  // extern "C"
