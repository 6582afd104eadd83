// Checkpoints of a run as the reference's restart files (host only): the keys, dtypes and rules of
// Simulation::checkpoint_header / save_checkpoint / load_checkpoint (cpp/src/SimulationRestart.cpp:61-85, 162-190,
// 254-360) for a uniform mesh, written and read with the .npz container of fvg_npz.cpp.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fvg.h"
#include "fvg_npz.hpp"

namespace {

constexpr int kFormat = 1;   // format_version the reference writes and the newest it reads

int fail(char* err, size_t n, int code, const std::string& m) {
    if (err && n) std::snprintf(err, n, "%s", m.c_str());
    return code;
}

}   // namespace

extern "C" {

int fvg_checkpoint_save(const fvg_config* cfg, const double* U, size_t ncells, int ncomp, double t, int step, double lagged_speed,
                        const char* path, char* err, size_t err_len) {
    if (!cfg || !U || !path) return fail(err, err_len, FVG_EINVAL, "null argument");
    const size_t nc = size_t(cfg->nx) * size_t(cfg->ny) * size_t(cfg->nz);
    const int want = cfg->nz > 1 ? 5 : 4;   // the reference writes rho w only in 3-D
    if (ncells != nc) return fail(err, err_len, FVG_EINVAL, "ncells differs from the mesh");
    if (ncomp != want) return fail(err, err_len, FVG_EINVAL, "ncomp must be " + std::to_string(want) + " for this mesh");
    try {
        using namespace fvg::npz;
        std::map<std::string, Array> a;
        if (lagged_speed > 0.0) a["S"] = scalar_f8(lagged_speed);
        a["format_version"] = scalar_i8(kFormat);
        a["t"] = scalar_f8(t);
        a["step"] = scalar_i8(step);
        a["nx"] = scalar_i8(cfg->nx);
        a["ny"] = scalar_i8(cfg->ny);
        a["nz"] = scalar_i8(cfg->nz);
        a["Lx"] = scalar_f8(cfg->Lx);
        a["Ly"] = scalar_f8(cfg->Ly);
        a["Lz"] = scalar_f8(cfg->Lz);
        a["collection_time"] = vector_f8({});
        a["collection_file"] = strings({});
        a["U"] = view_f8(U, {int64_t(nc), int64_t(ncomp)});
        const char* z64 = std::getenv("FVG_NPZ_ZIP64_FROM");   // test knob: write zip64 fields from this many bytes
        write(path, a, z64 ? std::strtoull(z64, nullptr, 10) : 0xFFFFFFFFull);
    } catch (const std::exception& e) {
        return fail(err, err_len, FVG_EIO, e.what());
    }
    return FVG_OK;
}

int fvg_checkpoint_load(const fvg_config* cfg, const char* path, double* U, size_t ncells, int ncomp, double* t, int* step,
                        double* lagged_speed, char* err, size_t err_len) {
    if (!cfg || !path || !U || !t || !step) return fail(err, err_len, FVG_EINVAL, "null argument");
    const size_t nc = size_t(cfg->nx) * size_t(cfg->ny) * size_t(cfg->nz);
    if (ncells != nc) return fail(err, err_len, FVG_EINVAL, "ncells differs from the mesh");
    if (ncomp != 4 && ncomp != 5) return fail(err, err_len, FVG_EINVAL, "ncomp must be 4 or 5");
    try {
        using namespace fvg::npz;
        const std::map<std::string, Array> a = read(path);
        auto has = [&](const char* k) { return a.count(k) != 0; };
        if (has("format_version") && as_double(a.at("format_version")) > kFormat)
            return fail(err, err_len, FVG_EIO, std::string("the checkpoint ") + path + " has a newer format_version than this library reads");
        if (has("leaf_level")) return fail(err, err_len, FVG_EINVAL, std::string("the checkpoint ") + path + " is from an adaptive (AMR) run");
        if (!has("U") || !has("t") || !has("step")) return fail(err, err_len, FVG_EIO, std::string(path) + " is not a checkpoint (U, t, step)");
        const int mesh[3] = {cfg->nx, cfg->ny, cfg->nz};
        const char* mk[3] = {"nx", "ny", "nz"};
        for (int d = 0; d < 3; ++d)
            if (has(mk[d]) && int(as_double(a.at(mk[d]))) != mesh[d])
                return fail(err, err_len, FVG_EINVAL, std::string("the checkpoint's ") + mk[d] + " differs from the mesh");
        const Array& u = a.at("U");
        if (u.descr != "<f8" || u.shape.size() != 2 || (u.shape[1] != 4 && u.shape[1] != 5))
            return fail(err, err_len, FVG_EIO, "expected an (n, 4) or (n, 5) <f8 state array");
        if (uint64_t(u.shape[0]) != nc)
            return fail(err, err_len, FVG_EINVAL, "the checkpoint holds " + std::to_string(u.shape[0]) + " cells, the mesh " + std::to_string(nc));
        const int fc = int(u.shape[1]);
        const double* src = reinterpret_cast<const double*>(u.data);
        for (size_t c = 0; c < nc; ++c)
            for (int q = 0; q < ncomp; ++q) U[c * ncomp + q] = q < fc ? src[c * fc + q] : 0.0;
        *t = as_double(a.at("t"));
        *step = int(as_double(a.at("step")));
        if (lagged_speed) *lagged_speed = has("S") ? as_double(a.at("S")) : 0.0;
    } catch (const std::exception& e) {
        return fail(err, err_len, FVG_EIO, e.what());
    }
    return FVG_OK;
}

}   // extern "C"
