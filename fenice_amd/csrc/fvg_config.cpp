// Run keys of the uniform-grid path: defaults, key = value parsing with the reference's labels and aliases,
// and the `auto` time-scheme rule. Host code only.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/fvg.h"

namespace {

std::string lower(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}
std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}
int fail(char* err, size_t n, const std::string& msg) {
    if (err && n) std::snprintf(err, n, "%s", msg.c_str());
    return FVG_EINVAL;
}
bool parse_num(const std::string& v, double& out) {
    const std::string t = trim(v);
    if (t.empty()) return false;
    char* end = nullptr;
    out = std::strtod(t.c_str(), &end);
    return end && *end == '\0';
}
// Params::integer: "100" or "100.0"
bool parse_int(const std::string& v, int& out) {
    double d;
    if (!parse_num(v, d) || d != std::floor(d)) return false;
    out = static_cast<int>(d);
    return true;
}
// Params::flag: true/false, 1/0, yes/no, on/off
bool parse_flag(const std::string& v, int& out) {
    const std::string s = lower(trim(v));
    if (s == "true" || s == "1" || s == "yes" || s == "on") return (out = 1), true;
    if (s == "false" || s == "0" || s == "no" || s == "off") return (out = 0), true;
    return false;
}

}   // namespace

extern "C" {

void fvg_config_defaults(fvg_config* c) {
    std::memset(c, 0, sizeof *c);
    c->nx = 0;
    c->ny = 0;
    c->nz = 1;
    c->Lx = c->Ly = c->Lz = 1.0;
    c->eos_type = FVG_EOS_IDEAL;
    c->gamma = 1.4;
    c->p_inf = 0.0;
    // BoundaryManager's sides before any case sets them (Boundaries.hpp:128-129)
    const int bc0[6] = {FVG_BC_TRANSMISSIVE, FVG_BC_TRANSMISSIVE, FVG_BC_WALL, FVG_BC_WALL, FVG_BC_PERIODIC, FVG_BC_PERIODIC};
    for (int s = 0; s < 6; ++s) c->bc[s] = bc0[s];
    c->u_in = 0.0;
    c->v_in = 0.0;
    c->rho_in = 1.0;
    c->p_out = 1.0;
    c->riemann = FVG_HLLC;
    c->shock_blend = 0.2;
    c->reconstruction = FVG_RECON_NONE;
    c->limiter = FVG_LIM_VAN_LEER;
    c->weno_epsilon = 1e-6;
    c->boundary_order = 2;
    c->time_scheme = FVG_TS_AUTO;
    c->cfl = 0.5;
    c->use_viscous = 0;
    c->mu = 1e-3;
    c->pr = 0.72;
    c->t_end = 0.005;
    c->dt_max = 0.0;
    c->max_steps = 0;
    c->flux_precision = FVG_FLUX_FP64;
    c->indicator = FVG_INDICATOR_NONE;
    c->indicator_audit = 50;
    c->indicator_label = 0.05;
    c->indicator_fn_limit = 0.01;
    c->indicator_fn_floor = 1e-4;
    c->indicator_label_source = FVG_LABEL_WENO5;
    c->flux_quadrature = FVG_QUAD_MIDPOINT;
    c->nrbc_sigma = 0.28;
    c->nrbc_beta = -1.0;
}

int fvg_config_set(fvg_config* c, const char* key_c, const char* value_c, char* err, size_t n) {
    if (!c || !key_c || !value_c) return fail(err, n, "null argument");
    const std::string key = trim(key_c), v = value_c, lv = lower(trim(value_c));
    double d;
    int k;
    auto num = [&](double& dst) {
        if (!parse_num(v, d)) return fail(err, n, "'" + key + "': not a number: '" + v + "'");
        dst = d;
        return int(FVG_OK);
    };
    auto integer = [&](int& dst) {
        if (!parse_int(v, k)) return fail(err, n, "'" + key + "': not an integer: '" + v + "'");
        dst = k;
        return int(FVG_OK);
    };
    if (key == "nx") return integer(c->nx);
    if (key == "ny") return integer(c->ny);
    if (key == "nz") return integer(c->nz);
    if (key == "Lx") return num(c->Lx);
    if (key == "Ly") return num(c->Ly);
    if (key == "Lz") return num(c->Lz);
    if (key == "gamma") return num(c->gamma);
    if (key == "p_inf") return num(c->p_inf);
    if (key == "u_in") return num(c->u_in);
    if (key == "v_in") return num(c->v_in);
    if (key == "rho_in") return num(c->rho_in);
    if (key == "p_out") return num(c->p_out);
    if (key == "shock_blend") return num(c->shock_blend);
    if (key == "weno_epsilon") return num(c->weno_epsilon);
    if (key == "boundary_order") return integer(c->boundary_order);
    if (key == "cfl") return num(c->cfl);
    if (key == "mu") return num(c->mu);
    if (key == "pr") return num(c->pr);
    if (key == "t_end") return num(c->t_end);
    if (key == "dt_max") return num(c->dt_max);
    if (key == "max_steps") return integer(c->max_steps);
    if (key == "use_viscous") {
        if (!parse_flag(v, k)) return fail(err, n, "'use_viscous': not a flag: '" + v + "'");
        c->use_viscous = k;
        return FVG_OK;
    }
    if (key == "eos_type") {   // SimParams::from accepts these two labels only (Params.cpp:476-478)
        if (lv == "ideal gas") return (c->eos_type = FVG_EOS_IDEAL), FVG_OK;
        if (lv == "stiffened gas") return (c->eos_type = FVG_EOS_STIFFENED), FVG_OK;
        return fail(err, n, "Unknown eos_type: '" + v + "' (use Ideal Gas or Stiffened Gas)");
    }
    if (key == "riemann_solver") {   // parse_riemann (Schemes.cpp:77-85)
        if (lv == "hllc") return (c->riemann = FVG_HLLC), FVG_OK;
        if (lv == "hllc (low-mach thornber)" || lv == "lowmach" || lv == "lowmachhllc" || lv == "low-mach")
            return (c->riemann = FVG_LOWMACH_HLLC), FVG_OK;
        if (lv == "hllc (preconditioned)" || lv == "preconditioned" || lv == "preconditionedhllc")
            return fail(err, n, "The preconditioned Riemann solver was removed: use 'HLLC (Low-Mach Thornber)' (or 'HLLC')");
        return fail(err, n, "Unknown Riemann solver: '" + v + "'");
    }
    if (key == "muscl_limiter") {   // parse_limiter (Schemes.cpp:87-94)
        if (lv == "minmod") return (c->limiter = FVG_LIM_MINMOD), FVG_OK;
        if (lv == "van leer" || lv == "vanleer") return (c->limiter = FVG_LIM_VAN_LEER), FVG_OK;
        if (lv == "mc") return (c->limiter = FVG_LIM_MC), FVG_OK;
        if (lv == "koren") return (c->limiter = FVG_LIM_KOREN), FVG_OK;
        return fail(err, n, "Unknown MUSCL limiter: '" + v + "' (use Minmod, Van Leer, MC or Koren)");
    }
    if (key == "reconstruction") {   // parse_reconstruction (Schemes.cpp:96-102) plus "none"
        if (lv == "none") return (c->reconstruction = FVG_RECON_NONE), FVG_OK;
        if (lv == "muscl") return (c->reconstruction = FVG_RECON_MUSCL), FVG_OK;
        if (lv == "weno3") return (c->reconstruction = FVG_RECON_WENO3), FVG_OK;
        if (lv == "weno5") return (c->reconstruction = FVG_RECON_WENO5), FVG_OK;
        return fail(err, n, "Unknown reconstruction: '" + v + "' (use muscl, weno3 or weno5)");
    }
    if (key == "time_scheme") {   // parse_time_scheme (TimeScheme.cpp:78-85) plus "auto"
        if (lv == "auto") return (c->time_scheme = FVG_TS_AUTO), FVG_OK;
        if (lv == "explicit euler" || lv == "euler" || lv == "explicit_euler") return (c->time_scheme = FVG_TS_EULER), FVG_OK;
        if (lv == "ssp runge-kutta 2" || lv == "ssprk2" || lv == "rk2") return (c->time_scheme = FVG_TS_SSPRK2), FVG_OK;
        if (lv == "ssp runge-kutta 3" || lv == "ssprk3" || lv == "rk3") return (c->time_scheme = FVG_TS_SSPRK3), FVG_OK;
        if (lv == "muscl-hancock" || lv == "hancock") return (c->time_scheme = FVG_TS_HANCOCK), FVG_OK;
        return fail(err, n, "Unknown time scheme: '" + v + "'");
    }
    if (key == "flux_precision") {
        if (lv == "fp64" || lv == "double") return (c->flux_precision = FVG_FLUX_FP64), FVG_OK;
        if (lv == "fp32" || lv == "float") return (c->flux_precision = FVG_FLUX_FP32), FVG_OK;
        return fail(err, n, "Unknown flux precision: '" + v + "' (use fp64 or fp32)");
    }
    if (key == "indicator") {
        if (lv == "none" || lv == "weno") return (c->indicator = FVG_INDICATOR_NONE), FVG_OK;
        if (lv == "learned") return (c->indicator = FVG_INDICATOR_LEARNED), FVG_OK;
        return fail(err, n, "Unknown indicator: '" + v + "' (use none or learned)");
    }
    if (key == "indicator_audit") return integer(c->indicator_audit);
    if (key == "indicator_label") return num(c->indicator_label);
    if (key == "indicator_fn_limit") return num(c->indicator_fn_limit);
    if (key == "indicator_fn_floor") return num(c->indicator_fn_floor);
    if (key == "indicator_label_source") {
        if (lv == "weno5" || lv == "weno") return (c->indicator_label_source = FVG_LABEL_WENO5), FVG_OK;
        if (lv == "muscl") return (c->indicator_label_source = FVG_LABEL_MUSCL), FVG_OK;
        if (lv == "both") return (c->indicator_label_source = FVG_LABEL_BOTH), FVG_OK;
        return fail(err, n, "Unknown indicator_label_source: '" + v + "' (use weno5, muscl or both)");
    }
    if (key == "nrbc_sigma") return num(c->nrbc_sigma);   // Params.cpp:274-275
    if (key == "nrbc_beta") return num(c->nrbc_beta);
    if (key == "flux_quadrature") {   // Params.cpp:348, 448-449
        if (lv == "midpoint") return (c->flux_quadrature = FVG_QUAD_MIDPOINT), FVG_OK;
        if (lv == "fourth") return (c->flux_quadrature = FVG_QUAD_FOURTH), FVG_OK;
        return fail(err, n, "flux_quadrature must be midpoint or fourth; got '" + lv + "'");
    }
    // The reference's other keys (Params.cpp:98-223). The run control of the host driver (output, monitor, restart,
    // threads) is accepted and has no effect here; the parameters of the immersed body and of AMR are inert while those
    // features are off; a feature outside the uniform-grid path is refused when it is switched on.
    static const char* host_keys[] = {"test_case", "save_interval", "write_interval_time", "monitor_interval", "resume",
                                      "dump_state", "vtk_format", "n_cpus"};
    for (const char* h : host_keys)
        if (key == h) return FVG_OK;
    auto refused = [&](const std::string& what) {
        if (err && n) std::snprintf(err, n, "%s", (key + " = " + v + ": " + what + " is outside the uniform-grid device path").c_str());
        return int(FVG_EUNSUPPORTED);
    };
    static const char* switches[][2] = {{"use_bodyforce", "a body force"}, {"use_ibm", "the immersed boundary"}, {"use_amr", "AMR"},
                                        {"use_mpi", "MPI decomposition"}, {"positivity_fallback", "the positivity fallback"}};
    for (const auto& sw : switches) {
        if (key != sw[0]) continue;
        if (!parse_flag(v, k)) return fail(err, n, "'" + key + "': not a flag: '" + v + "'");
        return k ? refused(sw[1]) : int(FVG_OK);
    }
    if (key == "source_rho" || key == "source_mx" || key == "source_my" || key == "source_mz" || key == "source_E")
        // a source term is on for an expression or a non-zero number (Params.cpp:282-286)
        return parse_num(v, d) && d == 0.0 ? int(FVG_OK) : refused("a source term");
    if (key == "gx" || key == "gy" || key == "gz" || key == "reynolds") {
        if (!parse_num(v, d)) return fail(err, n, "'" + key + "': not a number: '" + v + "'");
        if (key == "reynolds" && d != 0.0) return refused("mu from the Reynolds number of an immersed body");
        return FVG_OK;   // gravity: used only with use_bodyforce
    }
    if (key == "u_in_table" || key == "v_in_table" || key == "rho_in_table" || key == "p_out_table")
        return trim(v).empty() ? int(FVG_OK) : refused("a boundary value table");
    if (key.compare(0, 4, "ibm_") == 0 || key.compare(0, 4, "amr_") == 0) return FVG_OK;
    static const char* bc_keys[6] = {"bc_xmin", "bc_xmax", "bc_ymin", "bc_ymax", "bc_zmin", "bc_zmax"};
    for (int s = 0; s < 6; ++s) {
        if (key != bc_keys[s]) continue;
        // parse_bc (Boundaries.cpp:15-27)
        int b = -1;
        if (lv == "wall" || lv == "inviscid wall (slip)") b = FVG_BC_WALL;
        else if (lv == "no_slip" || lv == "viscous wall (no-slip)") b = FVG_BC_NO_SLIP;
        else if (lv == "moving_wall" || lv == "moving wall") b = FVG_BC_MOVING_WALL;
        else if (lv == "vel_inlet" || lv == "velocity inlet") b = FVG_BC_VEL_INLET;
        else if (lv == "massflow_inlet" || lv == "mass flow inlet") b = FVG_BC_MASSFLOW_INLET;
        else if (lv == "outlet" || lv == "pressure outlet") b = FVG_BC_OUTLET;
        else if (lv == "transmissive") b = FVG_BC_TRANSMISSIVE;
        else if (lv == "periodic") b = FVG_BC_PERIODIC;
        else if (lv == "nonreflecting" || lv == "non_reflecting" || lv == "non-reflecting outlet") b = FVG_BC_NONREFLECTING;
        if (b < 0) return fail(err, n, "Unknown boundary condition type: '" + v + "'");
        c->bc[s] = b;
        return FVG_OK;
    }
    return fail(err, n, "Unknown key: '" + key + "'");
}

int fvg_config_finalize(fvg_config* c, char* err, size_t n) {
    if (c->time_scheme == FVG_TS_AUTO)   // SimParams::from (Params.cpp:360-363)
        c->time_scheme = c->reconstruction != FVG_RECON_MUSCL ? FVG_TS_EULER : (c->use_viscous ? FVG_TS_SSPRK2 : FVG_TS_HANCOCK);
    if (c->nx <= 0 || c->ny <= 0 || c->nz <= 0) return fail(err, n, "nx, ny and nz must be positive");
    if (!(c->Lx > 0) || !(c->Ly > 0) || !(c->Lz > 0)) return fail(err, n, "Lx, Ly and Lz must be positive");
    for (int a = 0; a < 3; ++a) {
        const bool lo = c->bc[2 * a] == FVG_BC_PERIODIC, hi = c->bc[2 * a + 1] == FVG_BC_PERIODIC;
        if (a < 2 || c->nz > 1)
            if (lo != hi) return fail(err, n, "Periodic boundaries come in pairs");
    }
    // the checks of SimParams::from on the keys of the path (Params.cpp:429-460), with its messages
    if (!(c->weno_epsilon >= 0.0)) return fail(err, n, "weno_epsilon must be >= 0");
    if ((c->reconstruction == FVG_RECON_WENO3 || c->reconstruction == FVG_RECON_WENO5) && c->time_scheme == FVG_TS_SSPRK2)
        return fail(err, n, std::string("reconstruction = ") + (c->reconstruction == FVG_RECON_WENO3 ? "weno3" : "weno5") +
                                " is not used with SSP Runge-Kutta 2 (use SSP Runge-Kutta 3)");
    if (c->time_scheme == FVG_TS_HANCOCK && c->reconstruction != FVG_RECON_MUSCL)
        return fail(err, n, "time_scheme = MUSCL-Hancock needs reconstruction = muscl");
    if (c->flux_quadrature != FVG_QUAD_MIDPOINT && c->flux_quadrature != FVG_QUAD_FOURTH)
        return fail(err, n, "flux_quadrature must be midpoint or fourth");
    if (c->flux_quadrature == FVG_QUAD_FOURTH && c->reconstruction != FVG_RECON_WENO3 && c->reconstruction != FVG_RECON_WENO5)
        return fail(err, n, "flux_quadrature = fourth needs reconstruction = weno3 or weno5");
    if (c->boundary_order != 1 && c->boundary_order != 2) return fail(err, n, "boundary_order must be 1 or 2");
    if (!(c->nrbc_sigma >= 0.0)) return fail(err, n, "nrbc_sigma must be >= 0");   // Params.cpp:489-490
    if (!(c->nrbc_beta <= 1.0)) return fail(err, n, "nrbc_beta must be <= 1 (< 0: local Mach number)");
    if (!(c->cfl > 0.0)) return fail(err, n, "cfl must be positive");   // Params.cpp:491-496
    if (!(c->t_end > 0.0)) return fail(err, n, "t_end must be positive");
    if (!(c->dt_max >= 0.0) || !std::isfinite(c->dt_max)) return fail(err, n, "dt_max must be finite and >= 0 (0: off)");
    if (!(c->gamma > 1.0)) return fail(err, n, "gamma must be greater than 1");
    if (c->indicator == FVG_INDICATOR_LEARNED && (c->reconstruction != FVG_RECON_WENO5 || c->time_scheme == FVG_TS_HANCOCK))
        return fail(err, n, "indicator = learned needs reconstruction = weno5 and a Runge-Kutta scheme");
    if (c->indicator == FVG_INDICATOR_LEARNED && c->flux_quadrature != FVG_QUAD_MIDPOINT)
        return fail(err, n, "indicator = learned needs flux_quadrature = midpoint");
    if (c->indicator_audit < 0 || !(c->indicator_label > 0.0) || !(c->indicator_fn_limit >= 0.0) || !(c->indicator_fn_floor >= 0.0))
        return fail(err, n, "indicator_audit must be >= 0, indicator_label > 0 and indicator_fn_limit, indicator_fn_floor >= 0");
    // a first-order scheme has no reconstruction to extend to the boundary faces (Params.cpp:460)
    if (c->boundary_order == 2 && c->reconstruction == FVG_RECON_NONE) c->boundary_order = 1;
    return FVG_OK;
}

double fvg_dt_length(const fvg_config* c) {
    const double dx = c->Lx / c->nx, dy = c->Ly / c->ny;
    if (c->nz > 1) {
        const double dz = c->Lz / c->nz;
        return 2.0 / (1.0 / dx + 1.0 / dy + 1.0 / dz);
    }
    return (dy < dx) ? dy : dx;   // std::min(dx, dy)
}

}   // extern "C"
