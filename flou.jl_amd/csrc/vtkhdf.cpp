// Host-side VTKHDF writer (libflouvtkhdf.so): lays Flou's mesh and point data out in an
// HDF5 file the way FlouBiz.open_for_write / add_*data! / close_file! do through HDF5.jl
// (src/FlouBiz/FlouBiz.jl:38-113, src/FlouSpatial/IO.jl:16-76).  Plain C++ over the HDF5
// C library; no GPU code.  The group and attribute layout, the dataset names and their
// element types follow the reference's files because VTK readers and Flou's own tools
// read exactly those.
#include <hdf5.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/flouvtkhdf.h"

struct flou_vtkhdf_file {
    hid_t file = -1;
    hid_t root = -1;  // the /VTKHDF group
    hid_t lcpl = -1;  // link creation: intermediate groups, UTF-8 names (as HDF5.jl)
};

namespace {

thread_local std::string g_err;

herr_t collect(unsigned n, const H5E_error2_t* e, void* data) {
    auto* s = static_cast<std::string*>(data);
    if (n < 4 && e && e->desc) {
        if (!s->empty()) *s += "; ";
        *s += e->func_name ? e->func_name : "?";
        *s += ": ";
        *s += e->desc;
    }
    return 0;
}

int fail(int code, const std::string& what) {
    g_err = what;
    return code;
}

// the HDF5 error stack of the failed call, innermost first, appended to `what`
int fail_hdf5(const std::string& what) {
    std::string stack;
    H5Ewalk2(H5E_DEFAULT, H5E_WALK_DOWNWARD, collect, &stack);
    H5Eclear2(H5E_DEFAULT);
    return fail(FLOU_VTKHDF_EHDF5, what + (stack.empty() ? "" : " (" + stack + ")"));
}

void quiet() {
    // failures come back as return codes with the message in flou_vtkhdf_last_error
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

hid_t native(int dtype) {
    switch (dtype) {
        case FLOU_VTKHDF_F64: return H5T_NATIVE_DOUBLE;
        case FLOU_VTKHDF_F32: return H5T_NATIVE_FLOAT;
        case FLOU_VTKHDF_I64: return H5T_NATIVE_INT64;
        case FLOU_VTKHDF_U8: return H5T_NATIVE_UINT8;
        default: return -1;
    }
}

// one contiguous dataset at loc/path with the file type equal to the memory type, as
// HDF5.jl writes Julia arrays
int write_dataset(hid_t loc, hid_t lcpl, const char* path, hid_t type, int rank, const hsize_t* dims,
                  const void* data) {
    static const char dummy = 0;
    hid_t space = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, nullptr);
    if (space < 0) return fail_hdf5(std::string("dataspace of ") + path);
    hid_t ds = H5Dcreate2(loc, path, type, space, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(space);
        return fail_hdf5(std::string("creating dataset ") + path);
    }
    herr_t st = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data ? data : &dummy);
    H5Dclose(ds);
    H5Sclose(space);
    if (st < 0) return fail_hdf5(std::string("writing dataset ") + path);
    return FLOU_VTKHDF_OK;
}

int write_count(hid_t loc, hid_t lcpl, const char* name, int64_t v) {
    hsize_t one = 1;
    return write_dataset(loc, lcpl, name, H5T_NATIVE_INT64, 1, &one, &v);
}

void release(flou_vtkhdf_file* f) {
    if (!f) return;
    if (f->lcpl >= 0) H5Pclose(f->lcpl);
    if (f->root >= 0) H5Gclose(f->root);
    if (f->file >= 0) H5Fclose(f->file);
    delete f;
}

int check_mesh(int64_t npoints, const void* points, int points_dtype, int64_t nconn, const int64_t* conn,
               int64_t ncells, const uint8_t* types, const int64_t* offsets, const int64_t* regions) {
    if (npoints < 0 || nconn < 0 || ncells < 0) return fail(FLOU_VTKHDF_EINVAL, "negative size");
    if (points_dtype != FLOU_VTKHDF_F64 && points_dtype != FLOU_VTKHDF_F32)
        return fail(FLOU_VTKHDF_EINVAL, "points must be float64 or float32");
    if ((npoints && !points) || (nconn && !conn) || (ncells && (!types || !regions)) || !offsets)
        return fail(FLOU_VTKHDF_EINVAL, "null array");
    if (offsets[0] != 0 || offsets[ncells] != nconn)
        return fail(FLOU_VTKHDF_EINVAL, "offsets must start at 0 and end at the connectivity length");
    for (int64_t c = 0; c < ncells; ++c)
        if (offsets[c + 1] < offsets[c]) return fail(FLOU_VTKHDF_EINVAL, "offsets must not decrease");
    for (int64_t k = 0; k < nconn; ++k)
        if (conn[k] < 0 || conn[k] >= npoints)
            return fail(FLOU_VTKHDF_EINVAL, "connectivity id " + std::to_string(conn[k]) + " at " +
                                                std::to_string(k) + " is not a point of the file");
    return FLOU_VTKHDF_OK;
}

}  // namespace

extern "C" {

const char* flou_vtkhdf_last_error(void) { return g_err.c_str(); }

const char* flou_vtkhdf_hdf5_version(void) {
    static char buf[32];
    unsigned a = 0, b = 0, c = 0;
    H5get_libversion(&a, &b, &c);
    std::snprintf(buf, sizeof buf, "%u.%u.%u", a, b, c);
    return buf;
}

int flou_vtkhdf_open(const char* filename, int64_t npoints, const void* points, int points_dtype, int64_t nconn,
                     const int64_t* connectivity, int64_t ncells, const uint8_t* types, const int64_t* offsets,
                     const int64_t* regions, flou_vtkhdf_file** out) {
    g_err.clear();
    if (!out) return fail(FLOU_VTKHDF_EINVAL, "null output handle");
    *out = nullptr;
    if (!filename || !*filename) return fail(FLOU_VTKHDF_EINVAL, "empty file name");
    if (int rc = check_mesh(npoints, points, points_dtype, nconn, connectivity, ncells, types, offsets, regions))
        return rc;
    quiet();
    auto* f = new (std::nothrow) flou_vtkhdf_file;
    if (!f) return fail(FLOU_VTKHDF_EINVAL, "out of host memory");

    // h5open(filename, "w"): create or truncate; closing the file closes what is open in it
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    if (fapl < 0 || H5Pset_fclose_degree(fapl, H5F_CLOSE_STRONG) < 0) {
        if (fapl >= 0) H5Pclose(fapl);
        release(f);
        return fail_hdf5("file access properties");
    }
    f->file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
    H5Pclose(fapl);
    if (f->file < 0) {
        release(f);
        return fail_hdf5(std::string("creating ") + filename);
    }
    f->lcpl = H5Pcreate(H5P_LINK_CREATE);
    if (f->lcpl < 0 || H5Pset_create_intermediate_group(f->lcpl, 1) < 0 ||
        H5Pset_char_encoding(f->lcpl, H5T_CSET_UTF8) < 0) {
        release(f);
        return fail_hdf5("link creation properties");
    }
    f->root = H5Gcreate2(f->file, "/VTKHDF", f->lcpl, H5P_DEFAULT, H5P_DEFAULT);
    if (f->root < 0) {
        release(f);
        return fail_hdf5("creating /VTKHDF");
    }

    // attribute Version = [1, 0] (int64)
    {
        const int64_t version[2] = {1, 0};
        hsize_t two = 2;
        hid_t sp = H5Screate_simple(1, &two, nullptr);
        hid_t at = sp < 0 ? -1 : H5Acreate2(f->root, "Version", H5T_NATIVE_INT64, sp, H5P_DEFAULT, H5P_DEFAULT);
        herr_t st = at < 0 ? -1 : H5Awrite(at, H5T_NATIVE_INT64, version);
        if (at >= 0) H5Aclose(at);
        if (sp >= 0) H5Sclose(sp);
        if (st < 0) {
            release(f);
            return fail_hdf5("attribute Version");
        }
    }
    // attribute Type = "UnstructuredGrid": fixed-length C string of the text's length,
    // ASCII, scalar dataspace
    {
        static const char kind[] = "UnstructuredGrid";
        hid_t ty = H5Tcopy(H5T_C_S1);
        herr_t st = ty < 0 ? -1 : H5Tset_size(ty, sizeof(kind) - 1);
        if (st >= 0) st = H5Tset_cset(ty, H5T_CSET_ASCII);
        hid_t sp = st < 0 ? -1 : H5Screate(H5S_SCALAR);
        hid_t at = sp < 0 ? -1 : H5Acreate2(f->root, "Type", ty, sp, H5P_DEFAULT, H5P_DEFAULT);
        st = at < 0 ? -1 : H5Awrite(at, ty, kind);
        if (at >= 0) H5Aclose(at);
        if (sp >= 0) H5Sclose(sp);
        if (ty >= 0) H5Tclose(ty);
        if (st < 0) {
            release(f);
            return fail_hdf5("attribute Type");
        }
    }

    int rc = FLOU_VTKHDF_OK;
    const hsize_t pdims[2] = {(hsize_t)npoints, 3};
    const hsize_t cdims = (hsize_t)nconn, ndims = (hsize_t)ncells, odims = (hsize_t)ncells + 1;
    if (!rc) rc = write_count(f->root, f->lcpl, "NumberOfPoints", npoints);
    if (!rc) rc = write_dataset(f->root, f->lcpl, "Points", native(points_dtype), 2, pdims, points);
    if (!rc) rc = write_count(f->root, f->lcpl, "NumberOfConnectivityIds", nconn);
    if (!rc) rc = write_dataset(f->root, f->lcpl, "Connectivity", H5T_NATIVE_INT64, 1, &cdims, connectivity);
    if (!rc) rc = write_count(f->root, f->lcpl, "NumberOfCells", ncells);
    if (!rc) rc = write_dataset(f->root, f->lcpl, "Types", H5T_NATIVE_UINT8, 1, &ndims, types);
    if (!rc) rc = write_dataset(f->root, f->lcpl, "Offsets", H5T_NATIVE_INT64, 1, &odims, offsets);
    if (!rc) rc = write_dataset(f->root, f->lcpl, "CellData/Region", H5T_NATIVE_INT64, 1, &ndims, regions);
    if (rc) {
        release(f);
        return rc;
    }
    *out = f;
    return FLOU_VTKHDF_OK;
}

int flou_vtkhdf_write(flou_vtkhdf_file* f, const char* path, int dtype, int rank, const int64_t* dims,
                      const void* data) {
    g_err.clear();
    if (!f || f->root < 0) return fail(FLOU_VTKHDF_EINVAL, "file is not open");
    if (!path || !*path || path[0] == '/') return fail(FLOU_VTKHDF_EINVAL, "path must be relative to /VTKHDF");
    hid_t type = native(dtype);
    if (type < 0) return fail(FLOU_VTKHDF_EINVAL, "unknown element type");
    if (rank < 0 || rank > 4 || (rank && !dims)) return fail(FLOU_VTKHDF_EINVAL, "rank must be 0..4");
    hsize_t hd[4];
    int64_t count = 1;
    for (int k = 0; k < rank; ++k) {
        if (dims[k] < 0) return fail(FLOU_VTKHDF_EINVAL, "negative dimension");
        hd[k] = (hsize_t)dims[k];
        count *= dims[k];
    }
    if (count && !data) return fail(FLOU_VTKHDF_EINVAL, "null data");
    quiet();
    htri_t there = H5Lexists(f->root, path, H5P_DEFAULT);
    if (there > 0) return fail(FLOU_VTKHDF_EINVAL, std::string("/VTKHDF/") + path + " already exists");
    H5Eclear2(H5E_DEFAULT);  // H5Lexists fails when a group along the path is missing
    return write_dataset(f->root, f->lcpl, path, type, rank, hd, data);
}

int flou_vtkhdf_close(flou_vtkhdf_file* f) {
    g_err.clear();
    if (!f) return fail(FLOU_VTKHDF_EINVAL, "null handle");
    quiet();
    int rc = FLOU_VTKHDF_OK;
    if (f->lcpl >= 0) H5Pclose(f->lcpl);
    if (f->root >= 0) H5Gclose(f->root);
    f->lcpl = f->root = -1;
    if (f->file >= 0 && H5Fclose(f->file) < 0) rc = fail_hdf5("closing the file");
    f->file = -1;
    delete f;
    return rc;
}

}  // extern "C"
