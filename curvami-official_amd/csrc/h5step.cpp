// Step container: the HDF5 file of Exporter::writeBinaryStep
// (plate/Exporter.h:203-327) through the HDF5 C library; the ABI and layout
// are described in include/curvami_h5.h.
#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/curvami_h5.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& what) {
    g_err = what;
    return 2;
}

// Closes an HDF5 identifier of any kind at scope exit.
struct Id {
    hid_t id = -1;
    explicit Id(hid_t v = -1) : id(v) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id() {
        if (id < 0) return;
        switch (H5Iget_type(id)) {
            case H5I_FILE: H5Fclose(id); break;
            case H5I_DATASET: H5Dclose(id); break;
            case H5I_DATASPACE: H5Sclose(id); break;
            case H5I_DATATYPE: H5Tclose(id); break;
            case H5I_ATTR: H5Aclose(id); break;
            case H5I_GENPROP_LST: H5Pclose(id); break;
            default: break;
        }
    }
    bool ok() const { return id >= 0; }
};

bool file_exists(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// rows x cols doubles, chunked (min(rows, 1000), cols), deflate 6 (Exporter.h:222-232);
// a table without rows is stored contiguous (a chunk needs at least one row)
bool write_table(hid_t file, const char* name, const double* v, hsize_t rows, hsize_t cols, hid_t* out) {
    const hsize_t dims[2] = {rows, cols};
    Id space(H5Screate_simple(2, dims, nullptr));
    Id plist(H5Pcreate(H5P_DATASET_CREATE));
    const hsize_t chunk[2] = {std::min<hsize_t>(rows, 1000), cols};
    if (!space.ok() || !plist.ok()) return false;
    if (rows > 0 && (H5Pset_chunk(plist.id, 2, chunk) < 0 || H5Pset_deflate(plist.id, 6) < 0)) return false;
    const hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, plist.id, H5P_DEFAULT);
    if (ds < 0) return false;
    if (rows > 0 && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) {
        H5Dclose(ds);
        return false;
    }
    *out = ds;
    return true;
}

bool string_attr(hid_t obj, const char* name, const char* value) {
    Id type(H5Tcopy(H5T_C_S1));
    if (!type.ok() || H5Tset_size(type.id, std::strlen(value) + 1) < 0) return false;
    Id space(H5Screate(H5S_SCALAR));
    Id a(H5Acreate2(obj, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT));
    return a.ok() && H5Awrite(a.id, type.id, value) >= 0;
}

bool array_attr(hid_t obj, const char* name, hid_t base, const void* v, hsize_t n) {
    Id type(H5Tarray_create2(base, 1, &n));
    Id space(H5Screate(H5S_SCALAR));
    Id a(H5Acreate2(obj, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT));
    return type.ok() && a.ok() && H5Awrite(a.id, type.id, v) >= 0;
}

bool scalar_attr(hid_t obj, const char* name, double v) {
    Id space(H5Screate(H5S_SCALAR));
    Id a(H5Acreate2(obj, name, H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT));
    return a.ok() && H5Awrite(a.id, H5T_NATIVE_DOUBLE, &v) >= 0;
}

}  // namespace

extern "C" {

const char* cvm_h5_last_error(void) { return g_err.c_str(); }

int cvm_h5_write_step(const char* path, int step, const double* data, int64_t num_gps, int num_cols,
                      const double* creases, int64_t crease_rows, const double* angles, int nangles,
                      const int* is_driver, int ndriver, const double* scalars8, const char* git_commit) {
    if (!path || step < 0 || num_gps < 0 || num_cols < 7 || (num_gps > 0 && !data) || crease_rows < 0 ||
        (crease_rows > 0 && !creases) || nangles < 0 || ndriver < 0 || (nangles > 0 && !angles) ||
        (ndriver > 0 && !is_driver) || !scalars8) {
        g_err = "invalid argument";
        return 1;
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);     // errors come back as status codes
    const bool exists = file_exists(path);
    Id file(exists ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                   : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT));
    if (!file.ok()) return fail(std::string("cannot ") + (exists ? "open " : "create ") + path);
    char name[64];
    std::snprintf(name, sizeof(name), "/step_%05d", step);
    hid_t dsid = -1;
    if (!write_table(file.id, name, data, static_cast<hsize_t>(num_gps), static_cast<hsize_t>(num_cols), &dsid))
        return fail(std::string("writing ") + name + " failed (does the step exist?)");
    Id ds(dsid);
    if (crease_rows > 0) {
        char cname[80];
        std::snprintf(cname, sizeof(cname), "/step_%05d_creases", step);
        hid_t cid = -1;
        if (!write_table(file.id, cname, creases, static_cast<hsize_t>(crease_rows), 7, &cid))
            return fail(std::string("writing ") + cname + " failed");
        Id cds(cid);
        if (!string_attr(cds.id, "columns", "fold_index,crease_tag,X_ref,Y_ref,s_arc,theta_deg,weight"))
            return fail("writing the crease columns attribute failed");
    }
    const double zero = 0.0;
    const int izero = 0;
    bool ok = string_attr(ds.id, "git_commit", git_commit ? git_commit : "unknown");
    ok = ok && array_attr(ds.id, "angles", H5T_NATIVE_DOUBLE, nangles ? angles : &zero,
                          static_cast<hsize_t>(std::max(nangles, 1)));
    ok = ok && array_attr(ds.id, "is_driver", H5T_NATIVE_INT, ndriver ? is_driver : &izero,
                          static_cast<hsize_t>(std::max(ndriver, 1)));
    const char* names[8] = {"disp", "W_ext", "E_mem", "E_ben", "E_hin", "planar_rms", "planar_max", "ref_diag"};
    for (int k = 0; k < 8 && ok; ++k) ok = scalar_attr(ds.id, names[k], scalars8[k]);
    if (!ok) return fail(std::string("writing the attributes of ") + name + " failed");
    return 0;
}

int cvm_h5_read_dataset(const char* path, const char* name, int64_t* dims, double* out, int64_t cap) {
    if (!path || !name || !dims) { g_err = "invalid argument"; return 1; }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    Id ds(H5Dopen2(file.id, name, H5P_DEFAULT));
    if (!ds.ok()) return fail(std::string("no dataset ") + name);
    Id space(H5Dget_space(ds.id));
    hsize_t d[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(space.id) != 2) return fail(std::string(name) + " is not 2-D");
    H5Sget_simple_extent_dims(space.id, d, nullptr);
    dims[0] = static_cast<int64_t>(d[0]);
    dims[1] = static_cast<int64_t>(d[1]);
    if (!out) return 0;
    if (cap < dims[0] * dims[1]) { g_err = "output too small"; return 1; }
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(std::string("reading ") + name + " failed");
    return 0;
}

int cvm_h5_read_attr(const char* path, const char* dataset, const char* attr, double* out, int cap, int* n) {
    if (!path || !dataset || !attr || !n) { g_err = "invalid argument"; return 1; }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    Id ds(H5Dopen2(file.id, dataset, H5P_DEFAULT));
    if (!ds.ok()) return fail(std::string("no dataset ") + dataset);
    Id a(H5Aopen(ds.id, attr, H5P_DEFAULT));
    if (!a.ok()) return fail(std::string("no attribute ") + attr);
    Id type(H5Aget_type(a.id));
    hsize_t count = 1;
    if (H5Tget_class(type.id) == H5T_ARRAY) {
        Id base(H5Tget_super(type.id));
        hsize_t d = 0;
        H5Tget_array_dims2(type.id, &d);
        count = d;
        *n = static_cast<int>(count);
        if (!out) return 0;
        if (cap < *n) { g_err = "output too small"; return 1; }
        Id mem(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, &count));
        if (H5Aread(a.id, mem.id, out) < 0) return fail(std::string("reading ") + attr + " failed");
        return 0;
    }
    *n = 1;
    if (!out) return 0;
    if (cap < 1) { g_err = "output too small"; return 1; }
    if (H5Aread(a.id, H5T_NATIVE_DOUBLE, out) < 0) return fail(std::string("reading ") + attr + " failed");
    return 0;
}

int cvm_h5_list(const char* path, char* buf, int cap, int* needed) {
    if (!path || !needed || (cap > 0 && !buf)) { g_err = "invalid argument"; return 1; }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    std::string names;
    auto visit = [](hid_t, const char* name, const H5L_info_t*, void* op) -> herr_t {
        std::string& s = *static_cast<std::string*>(op);
        if (!s.empty()) s += ',';
        s += name;
        return 0;
    };
    if (H5Literate(file.id, H5_INDEX_NAME, H5_ITER_INC, nullptr, visit, &names) < 0)
        return fail("listing the datasets failed");
    *needed = static_cast<int>(names.size()) + 1;
    if (cap > 0) std::snprintf(buf, static_cast<size_t>(cap), "%s", names.c_str());
    return 0;
}

int cvm_h5_read_attr_string(const char* path, const char* dataset, const char* attr, char* buf, int cap) {
    if (!path || !dataset || !attr || !buf || cap <= 0) { g_err = "invalid argument"; return 1; }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    Id ds(H5Dopen2(file.id, dataset, H5P_DEFAULT));
    if (!ds.ok()) return fail(std::string("no dataset ") + dataset);
    Id a(H5Aopen(ds.id, attr, H5P_DEFAULT));
    if (!a.ok()) return fail(std::string("no attribute ") + attr);
    Id type(H5Aget_type(a.id));
    const size_t sz = H5Tget_size(type.id);
    std::vector<char> tmp(sz + 1, 0);
    if (H5Aread(a.id, type.id, tmp.data()) < 0) return fail(std::string("reading ") + attr + " failed");
    std::snprintf(buf, static_cast<size_t>(cap), "%s", tmp.data());
    return 0;
}

}  // extern "C"
