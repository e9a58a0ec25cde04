// The solver-state restart record (RestartManager::save / load,
// core/include/curvami/RestartManager.h:11-117), byte for byte: magic
// "CURVRES\0", int format version 1, int k, doubles lambda, ds, dl_old, W_ext,
// int ns, then x[ns] and dU_old[ns] (native byte order). Loading verifies the
// whole record before it writes the caller's outputs, and refuses a foreign
// magic, another version, a truncated file, a non-positive or unexpected DOF
// count and non-finite values, as the reference does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/curvami_gpu.h"

namespace {

thread_local std::string g_restart_err;
constexpr char kMagic[8] = {'C', 'U', 'R', 'V', 'R', 'E', 'S', '\0'};
constexpr int kVersion = 1;

int fail(int code, const std::string& msg) {
    g_restart_err = msg;
    return code;
}

}  // namespace

extern "C" {

const char* cvm_restart_last_error(void) { return g_restart_err.c_str(); }

int cvm_restart_save(const char* path, int k, double lambda, double ds, double dl_old, double W_ext, int ns,
                     const double* x, const double* dU_old) {
    if (!path || ns < 0 || (ns > 0 && (!x || !dU_old))) return fail(CVM_E_INVALID, "restart save: null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CVM_E_INVALID, std::string("restart save: cannot open ") + path);
    const int version = kVersion;
    bool ok = std::fwrite(kMagic, 1, sizeof(kMagic), f) == sizeof(kMagic);
    ok = ok && std::fwrite(&version, sizeof(int), 1, f) == 1;
    ok = ok && std::fwrite(&k, sizeof(int), 1, f) == 1;
    for (double v : {lambda, ds, dl_old, W_ext}) ok = ok && std::fwrite(&v, sizeof(double), 1, f) == 1;
    ok = ok && std::fwrite(&ns, sizeof(int), 1, f) == 1;
    if (ns > 0) {
        ok = ok && std::fwrite(x, sizeof(double), static_cast<size_t>(ns), f) == static_cast<size_t>(ns);
        ok = ok && std::fwrite(dU_old, sizeof(double), static_cast<size_t>(ns), f) == static_cast<size_t>(ns);
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? CVM_OK : fail(CVM_E_INVALID, std::string("restart save: write failed for ") + path);
}

int cvm_restart_load(const char* path, int expected_ns, int* k, double* scalars4, int* ns, double* x,
                     double* dU_old, int cap) {
    if (!path || !k || !scalars4 || !ns) return fail(CVM_E_INVALID, "restart load: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CVM_E_INVALID, std::string("restart load: cannot open ") + path);
    struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
    char magic[sizeof(kMagic)] = {};
    int version = 0;
    if (std::fread(magic, 1, sizeof(magic), f) != sizeof(magic) || std::memcmp(magic, kMagic, sizeof(kMagic)) != 0 ||
        std::fread(&version, sizeof(int), 1, f) != 1)
        return fail(CVM_E_INVALID, std::string("restart file ") + path + " is not a Curvami restart record");
    if (version != kVersion)
        return fail(CVM_E_INVALID, std::string("restart file ") + path + " has format version " +
                                       std::to_string(version) + ", expected " + std::to_string(kVersion));
    int k_l = 0, ns_l = 0;
    double sc[4] = {0, 0, 0, 0};
    bool ok = std::fread(&k_l, sizeof(int), 1, f) == 1;
    ok = ok && std::fread(sc, sizeof(double), 4, f) == 4;
    ok = ok && std::fread(&ns_l, sizeof(int), 1, f) == 1;
    if (!ok) return fail(CVM_E_INVALID, std::string("restart file ") + path + " is truncated in its header");
    if (ns_l <= 0)
        return fail(CVM_E_INVALID, std::string("restart file ") + path + " reports a non-positive DOF count (" +
                                       std::to_string(ns_l) + ")");
    if (expected_ns > 0 && ns_l != expected_ns)
        return fail(CVM_E_INVALID, std::string("restart file ") + path + " holds " + std::to_string(ns_l) +
                                       " DOFs but this problem has " + std::to_string(expected_ns));
    if (!x && !dU_old) {                         // size query
        *ns = ns_l;
        return CVM_OK;
    }
    if (!x || !dU_old || cap < ns_l) return fail(CVM_E_INVALID, "restart load: output arrays too small");
    std::vector<double> xl(static_cast<size_t>(ns_l)), dl(static_cast<size_t>(ns_l));
    if (std::fread(xl.data(), sizeof(double), xl.size(), f) != xl.size() ||
        std::fread(dl.data(), sizeof(double), dl.size(), f) != dl.size())
        return fail(CVM_E_INVALID, std::string("restart file ") + path + " is truncated in its state arrays");
    for (size_t i = 0; i < xl.size(); ++i)
        if (!std::isfinite(xl[i]) || !std::isfinite(dl[i]))
            return fail(CVM_E_INVALID, std::string("restart file ") + path + " contains non-finite values");
    *k = k_l;
    std::memcpy(scalars4, sc, sizeof(sc));
    *ns = ns_l;
    std::memcpy(x, xl.data(), xl.size() * sizeof(double));
    std::memcpy(dU_old, dl.data(), dl.size() * sizeof(double));
    return CVM_OK;
}

}  // extern "C"
