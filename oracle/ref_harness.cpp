// Harness over the reference itself (test infrastructure only): a C ABI around the reference's own
// Simulation / TimeScheme / RiemannSolver, compiled from the sources under /root/reference by oracle/Makefile
// into oracle/_ref/libfenice_ref.so. It sets up a run exactly as the reference's tests do
// (cpp/tests/test_common.hpp:83-90: baseline defaults, the case's defaults, then key = value overrides) and
// exposes residual / step / initial field / Riemann flux, so the oracle and the device path can be compared
// with the reference on the same seeded inputs. Nothing here restates an algorithm; it only calls it.
#include <cstdio>
#include <cstring>
#include <exception>
#include <filesystem>
#include <memory>
#include <string>

#include "fv/Params.hpp"
#include "fv/Schemes.hpp"
#include "fv/Simulation.hpp"
#include "fv/TestCases.hpp"
#include "fv/TimeScheme.hpp"

namespace {
struct Handle {
    fv::Simulation sim;
    fv::ResidualFn res;
};
thread_local std::string g_err;
}   // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// A run of test case `tc` with n overrides; null on an error (ref_last_error).
void* ref_create(const char* tc, int n, const char** keys, const char** vals) {
    try {
        fv::Params p = fv::baseline_defaults();
        p.merge(fv::get_test_case(tc).defaults);
        p.set("test_case", tc);
        p.set("output_root", "/tmp/fenice_ref_out");
        for (int k = 0; k < n; ++k) p.set(keys[k], vals[k]);
        auto h = std::make_unique<Handle>();
        h->sim.setup(fv::SimParams::from(p), false);
        Handle* raw = h.get();
        h->res = [raw](const fv::Field& U, fv::Field& R) { return raw->sim.residual(U, R); };
        return h.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_destroy(void* hp) { delete static_cast<Handle*>(hp); }

long long ref_num_cells(void* hp) { return static_cast<Handle*>(hp)->sim.mesh.num_cells(); }
double ref_dt_length(void* hp) { return static_cast<Handle*>(hp)->sim.dt_length; }

// The case's initial field (Simulation::initialize_field) into U (n x 5).
int ref_initial_field(void* hp, double* U) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        h->sim.initialize_field();
        std::memcpy(U, h->sim.U.data(), sizeof(double) * 5 * h->sim.U.size());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int ref_set_state(void* hp, const double* U) {
    Handle* h = static_cast<Handle*>(hp);
    h->sim.U.resize(static_cast<size_t>(h->sim.mesh.num_cells()));
    std::memcpy(h->sim.U.data(), U, sizeof(double) * 5 * h->sim.U.size());
    return 0;
}
int ref_get_state(void* hp, double* U) {
    Handle* h = static_cast<Handle*>(hp);
    std::memcpy(U, h->sim.U.data(), sizeof(double) * 5 * h->sim.U.size());
    return 0;
}

// R = residual(U) with the time scheme's predictor half step set to half_dt (0: plain); returns 0 and S.
int ref_residual(void* hp, const double* U, double half_dt, double* R, double* S) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        fv::Field Uf(static_cast<size_t>(h->sim.mesh.num_cells())), Rf;
        std::memcpy(Uf.data(), U, sizeof(double) * 5 * Uf.size());
        h->sim.scheme->predictor_half_dt = half_dt;
        *S = h->sim.residual(Uf, Rf);
        h->sim.scheme->predictor_half_dt = 0.0;
        std::memcpy(R, Rf.data(), sizeof(double) * 5 * Rf.size());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// One TimeScheme::step of the resident state (the solver's step: min_delta = dt_length).
int ref_step(void* hp, double t, double t_max, double dt_max, double* dt, double* S) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        const fv::StepResult r = h->sim.scheme->step(h->sim.U, h->res, h->sim.dt_length, t, t_max, dt_max);
        *dt = r.dt;
        *S = r.s_max;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Steps while t < t_end (and at most max_steps > 0), as Simulation::run_loop; returns the step count.
int ref_run(void* hp, double* t, double t_end, int max_steps) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        int steps = 0;
        while (*t < t_end) {
            const fv::StepResult r = h->sim.scheme->step(h->sim.U, h->res, h->sim.dt_length, *t, t_end, 0.0);
            ++steps;
            if (!(std::isfinite(r.s_max) && std::isfinite(r.dt) && *t + r.dt > *t)) {
                g_err = "diverged";
                return -steps;
            }
            *t += r.dt;
            if (max_steps > 0 && steps >= max_steps) break;
        }
        return steps;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1000000000;
    }
}

// The run's restart file (Simulation::checkpoint_path) into buf (cap bytes); its length, or -1 when it does not fit.
int ref_checkpoint_path(void* hp, char* buf, int cap) {
    const std::string p = static_cast<Handle*>(hp)->sim.checkpoint_path();
    if ((int)p.size() + 1 > cap) return -1;
    std::memcpy(buf, p.c_str(), p.size() + 1);
    return (int)p.size();
}

// Simulation::save_checkpoint of the current state at time t as step `step` (into checkpoint_path()).
int ref_save_checkpoint(void* hp, double t, int step) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        std::filesystem::create_directories(std::filesystem::path(h->sim.checkpoint_path()).parent_path());
        h->sim.t = t;
        h->sim.save_checkpoint(step);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Simulation::load_checkpoint from checkpoint_path(): state, t and lagged speed; returns the step, -1 on an error.
int ref_load_checkpoint(void* hp, double* t) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        const int step = h->sim.load_checkpoint();
        *t = h->sim.t;
        return step;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Simulation::volume_averages of U (n x 5) over the run's n cells: mass, kinetic energy and total energy per cell.
int ref_volume_averages(void* hp, const double* U, double* out) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        fv::Field Uf(static_cast<size_t>(h->sim.mesh.num_cells()));
        std::memcpy(Uf.data(), U, sizeof(double) * 5 * Uf.size());
        const std::array<double, 3> v = h->sim.volume_averages(Uf, static_cast<long long>(Uf.size()));
        for (int q = 0; q < 3; ++q) out[q] = v[q];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

double ref_get_lagged_speed(void* hp) { return static_cast<Handle*>(hp)->sim.scheme->lagged_speed(); }
void ref_set_lagged_speed(void* hp, double s) { static_cast<Handle*>(hp)->sim.scheme->set_lagged_speed(s); }

// RiemannSolver::face_flux of n faces (5 doubles per state).
int ref_face_flux(double gamma, int stiffened, double p_inf, int lowmach, double blend, const double* UL, const double* UR,
                  const int* axis, long long n, double* F, double* smax) {
    const fv::EOS eos = stiffened ? fv::EOS::stiffened(gamma, p_inf) : fv::EOS::ideal(gamma);
    const fv::RiemannSolver rs(eos, lowmach ? fv::RiemannType::LowMachHLLC : fv::RiemannType::HLLC, blend);
    for (long long k = 0; k < n; ++k) {
        fv::State ul, ur;
        std::memcpy(ul.data(), UL + 5 * k, sizeof ul);
        std::memcpy(ur.data(), UR + 5 * k, sizeof ur);
        const fv::State f = rs.face_flux(ul, ur, axis[k], smax[k]);
        std::memcpy(F + 5 * k, f.data(), sizeof f);
    }
    return 0;
}

}   // extern "C"
