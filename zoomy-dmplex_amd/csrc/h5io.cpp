// HDF5 float64 datasets for initial-condition files (IOManager::LoadSolution,
// IOManager.hpp:122-192, which reads the "state" Vec through PETSc's HDF5 viewer).
//
// libhdf5 is opened at run time, so the library builds and runs where it is absent (the
// calls then fail with a message naming what was searched).  Only the part of the HDF5 C API
// that reading and writing one dataset needs is declared here, for HDF5 >= 1.10 (64-bit
// hid_t); the library's version is checked before use.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <string>

#include "zdm_mesh.h"

void zdm_mesh_set_error(const std::string& msg);   // mesh.cpp

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

constexpr hid_t kDefault = 0;                  // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr unsigned kAccRdonly = 0u, kAccTrunc = 2u;
constexpr int kClassInteger = 0, kClassFloat = 1;

struct Hdf5 {
    void* lib = nullptr;
    std::string err;
    herr_t (*open)(void) = nullptr;
    herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*) = nullptr;
    herr_t (*eset_auto)(hid_t, void*, void*) = nullptr;
    hid_t (*fopen)(const char*, unsigned, hid_t) = nullptr;
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
    herr_t (*fclose)(hid_t) = nullptr;
    hid_t (*dopen)(hid_t, const char*, hid_t) = nullptr;
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    hid_t (*dget_space)(hid_t) = nullptr;
    hid_t (*dget_type)(hid_t) = nullptr;
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*) = nullptr;
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
    herr_t (*dclose)(hid_t) = nullptr;
    int (*sndims)(hid_t) = nullptr;
    int (*sdims)(hid_t, hsize_t*, hsize_t*) = nullptr;
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*sclose)(hid_t) = nullptr;
    int (*tget_class)(hid_t) = nullptr;
    herr_t (*tclose)(hid_t) = nullptr;
    hid_t* native_double = nullptr;   // H5T_NATIVE_DOUBLE_g
    hid_t* ieee_f64le = nullptr;      // H5T_IEEE_F64LE_g
};

template <typename F>
bool sym(Hdf5& h, F& f, const char* name) {
    f = reinterpret_cast<F>(dlsym(h.lib, name));
    if (!f) h.err = std::string("libhdf5 lacks ") + name;
    return f != nullptr;
}

Hdf5* load() {
    static Hdf5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        std::string tried;
        // $ZDM_HDF5_LIB alone when set, else the usual names
        const char* env = std::getenv("ZDM_HDF5_LIB");
        const bool given = env && *env;
        const char* names[] = {given ? env : "libhdf5.so", given ? nullptr : "libhdf5.so.103",
                               given ? nullptr : "libhdf5.so.200", given ? nullptr : "libhdf5.so.310",
                               given ? nullptr : "/opt/conda/lib/libhdf5.so"};
        for (const char* n : names) {
            if (!n) continue;
            h.lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
            if (h.lib) break;
            tried += std::string(tried.empty() ? "" : ", ") + n;
        }
        if (!h.lib) {
            h.err = "HDF5 initial conditions need libhdf5 (set ZDM_HDF5_LIB); none found among: " + tried;
            return;
        }
        bool ok = sym(h, h.open, "H5open") && sym(h, h.get_libversion, "H5get_libversion") &&
                  sym(h, h.eset_auto, "H5Eset_auto2") && sym(h, h.fopen, "H5Fopen") &&
                  sym(h, h.fcreate, "H5Fcreate") && sym(h, h.fclose, "H5Fclose") && sym(h, h.dopen, "H5Dopen2") &&
                  sym(h, h.dcreate, "H5Dcreate2") && sym(h, h.dget_space, "H5Dget_space") &&
                  sym(h, h.dget_type, "H5Dget_type") && sym(h, h.dread, "H5Dread") &&
                  sym(h, h.dwrite, "H5Dwrite") && sym(h, h.dclose, "H5Dclose") &&
                  sym(h, h.sndims, "H5Sget_simple_extent_ndims") &&
                  sym(h, h.sdims, "H5Sget_simple_extent_dims") && sym(h, h.screate_simple, "H5Screate_simple") &&
                  sym(h, h.sclose, "H5Sclose") && sym(h, h.tget_class, "H5Tget_class") &&
                  sym(h, h.tclose, "H5Tclose") && sym(h, h.native_double, "H5T_NATIVE_DOUBLE_g") &&
                  sym(h, h.ieee_f64le, "H5T_IEEE_F64LE_g");
        unsigned maj = 0, min = 0, rel = 0;
        if (ok && (h.open() < 0 || h.get_libversion(&maj, &min, &rel) < 0)) {
            h.err = "libhdf5 failed to initialise";
            ok = false;
        }
        if (ok && (maj != 1 || min < 10)) {
            h.err = "libhdf5 " + std::to_string(maj) + "." + std::to_string(min) +
                    " has a 32-bit hid_t; HDF5 >= 1.10 is needed";
            ok = false;
        }
        if (!ok) {
            dlclose(h.lib);
            h.lib = nullptr;
            return;
        }
        h.eset_auto(kDefault, nullptr, nullptr);   // errors are reported here, not printed
    });
    if (!h.lib) {
        zdm_mesh_set_error(h.err);
        return nullptr;
    }
    return &h;
}

std::mutex g_h5;   // the library may be built without thread safety

int fail(const std::string& msg) {
    zdm_mesh_set_error(msg);
    return -1;
}

}  // namespace

extern "C" {

int zdm_h5_read(const char* path, const char* dataset, double* data, uint64_t capacity, int* ndims,
                uint64_t* dims) {
    if (!path || !dataset || !ndims || !dims) return fail("zdm_h5_read: null argument");
    Hdf5* h = load();
    if (!h) return -1;
    std::lock_guard<std::mutex> lk(g_h5);
    const hid_t f = h->fopen(path, kAccRdonly, kDefault);
    if (f < 0) return fail(std::string("cannot open HDF5 file ") + path);
    const hid_t d = h->dopen(f, dataset, kDefault);
    if (d < 0) {
        h->fclose(f);
        return fail(std::string(path) + ": no dataset \"" + dataset + "\"");
    }
    int rc = 0;
    const hid_t sp = h->dget_space(d), ty = h->dget_type(d);
    const int nd = sp >= 0 ? h->sndims(sp) : -1;
    const int cls = ty >= 0 ? h->tget_class(ty) : -1;
    hsize_t dd[4] = {0, 0, 0, 0};
    if (nd < 1 || nd > 4 || h->sdims(sp, dd, nullptr) < 0) {
        rc = fail(std::string(path) + ": dataset \"" + dataset + "\" must have 1 to 4 dimensions");
    } else if (cls != kClassFloat && cls != kClassInteger) {
        rc = fail(std::string(path) + ": dataset \"" + dataset + "\" is not numeric");
    } else {
        uint64_t n = 1;
        *ndims = nd;
        for (int i = 0; i < nd; ++i) {
            dims[i] = dd[i];
            n *= dd[i];
        }
        if (data) {
            if (capacity < n)
                rc = fail(std::string(path) + ": dataset holds " + std::to_string(n) + " values, room for " +
                          std::to_string(capacity));
            else if (n > 0 && h->dread(d, *h->native_double, kDefault, kDefault, kDefault, data) < 0)
                rc = fail(std::string(path) + ": reading \"" + dataset + "\" failed");
        }
    }
    if (ty >= 0) h->tclose(ty);
    if (sp >= 0) h->sclose(sp);
    h->dclose(d);
    h->fclose(f);
    return rc;
}

int zdm_h5_write(const char* path, const char* dataset, const double* data, int ndims, const uint64_t* dims) {
    if (!path || !dataset || !dims || ndims < 1 || ndims > 4) return fail("zdm_h5_write: bad arguments");
    Hdf5* h = load();
    if (!h) return -1;
    std::lock_guard<std::mutex> lk(g_h5);
    hsize_t dd[4];
    uint64_t n = 1;
    for (int i = 0; i < ndims; ++i) {
        dd[i] = dims[i];
        n *= dims[i];
    }
    if (n > 0 && !data) return fail("zdm_h5_write: null data");
    const hid_t f = h->fcreate(path, kAccTrunc, kDefault, kDefault);
    if (f < 0) return fail(std::string("cannot create HDF5 file ") + path);
    int rc = 0;
    const hid_t sp = h->screate_simple(ndims, dd, nullptr);
    const hid_t d = sp >= 0 ? h->dcreate(f, dataset, *h->ieee_f64le, sp, kDefault, kDefault, kDefault) : -1;
    if (d < 0)
        rc = fail(std::string(path) + ": cannot create dataset \"" + dataset + "\"");
    else if (n > 0 && h->dwrite(d, *h->native_double, kDefault, kDefault, kDefault, data) < 0)
        rc = fail(std::string(path) + ": writing \"" + dataset + "\" failed");
    if (d >= 0) h->dclose(d);
    if (sp >= 0) h->sclose(sp);
    if (h->fclose(f) < 0 && rc == 0) rc = fail(std::string(path) + ": closing failed");
    return rc;
}

}  // extern "C"
