// Writes an initial-condition file the way the reference's case tools do
// (make_bump_case.py:127-160, generate_ic.py:120-127: h5py, a flat cell-major float64 dataset
// "state" of ncells * ncomp values with integer attributes num_cells and num_components),
// through libhdf5's C API opened at run time, independently of the library's reader.
//
//   h5_ic_write <file.h5> <ncells> <ncomp>   < values (whitespace separated, cell-major)
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

template <typename F>
static F get(void* lib, const char* name) {
    F f = reinterpret_cast<F>(dlsym(lib, name));
    if (!f) {
        std::fprintf(stderr, "libhdf5 lacks %s\n", name);
        std::exit(2);
    }
    return f;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s file.h5 ncells ncomp < values\n", argv[0]);
        return 2;
    }
    const long nc = std::atol(argv[2]), nq = std::atol(argv[3]);
    std::vector<double> v((size_t)nc * nq);
    for (double& x : v)
        if (std::scanf("%lf", &x) != 1) {
            std::fprintf(stderr, "expected %ld values\n", nc * nq);
            return 2;
        }
    const char* env = std::getenv("ZDM_HDF5_LIB");
    void* lib = dlopen(env && *env ? env : "libhdf5.so", RTLD_NOW);
    if (!lib) lib = dlopen("/opt/conda/lib/libhdf5.so", RTLD_NOW);
    if (!lib) {
        std::fprintf(stderr, "no libhdf5\n");
        return 2;
    }
    auto H5open = get<herr_t (*)()>(lib, "H5open");
    auto H5Fcreate = get<hid_t (*)(const char*, unsigned, hid_t, hid_t)>(lib, "H5Fcreate");
    auto H5Fclose = get<herr_t (*)(hid_t)>(lib, "H5Fclose");
    auto H5Screate_simple = get<hid_t (*)(int, const hsize_t*, const hsize_t*)>(lib, "H5Screate_simple");
    auto H5Screate = get<hid_t (*)(int)>(lib, "H5Screate");
    auto H5Sclose = get<herr_t (*)(hid_t)>(lib, "H5Sclose");
    auto H5Dcreate2 = get<hid_t (*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t)>(lib, "H5Dcreate2");
    auto H5Dwrite = get<herr_t (*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*)>(lib, "H5Dwrite");
    auto H5Dclose = get<herr_t (*)(hid_t)>(lib, "H5Dclose");
    auto H5Acreate2 = get<hid_t (*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t)>(lib, "H5Acreate2");
    auto H5Awrite = get<herr_t (*)(hid_t, hid_t, const void*)>(lib, "H5Awrite");
    auto H5Aclose = get<herr_t (*)(hid_t)>(lib, "H5Aclose");
    if (H5open() < 0) return 3;
    const hid_t f64 = *get<hid_t*>(lib, "H5T_IEEE_F64LE_g"), nd = *get<hid_t*>(lib, "H5T_NATIVE_DOUBLE_g");
    const hid_t i64 = *get<hid_t*>(lib, "H5T_STD_I64LE_g"), ni = *get<hid_t*>(lib, "H5T_NATIVE_LLONG_g");
    const hid_t file = H5Fcreate(argv[1], 2u /* truncate */, 0, 0);
    if (file < 0) return 3;
    const hsize_t n = (hsize_t)v.size();
    const hid_t sp = H5Screate_simple(1, &n, nullptr);
    const hid_t ds = H5Dcreate2(file, "state", f64, sp, 0, 0, 0);
    if (ds < 0 || H5Dwrite(ds, nd, 0, 0, 0, v.data()) < 0) return 3;
    const long long attrs[2] = {nc, nq};
    const char* names[2] = {"num_cells", "num_components"};
    for (int i = 0; i < 2; ++i) {
        const hid_t as = H5Screate(0 /* scalar */);
        const hid_t a = H5Acreate2(ds, names[i], i64, as, 0, 0);
        if (a < 0 || H5Awrite(a, ni, &attrs[i]) < 0) return 3;
        H5Aclose(a);
        H5Sclose(as);
    }
    H5Dclose(ds);
    H5Sclose(sp);
    return H5Fclose(file) < 0 ? 3 : 0;
}
