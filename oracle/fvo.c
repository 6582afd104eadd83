/* The oracle (test infrastructure only; see fvo.h): a plain-C FP64 restatement of the reference's
 * uniform-grid residual and time step. Each function names the reference lines it follows. Compiled with
 * -ffp-contract=off -fno-fast-math: no multiply-add is fused, as in the reference (cpp/CMakeLists.txt:85).
 * min/max are the reference's compare-and-select forms: std::min(a, b) = (b < a) ? b : a and
 * std::max(a, b) = (a < b) ? b : a (NaN in the second operand is dropped, in the first it is kept).
 */
#include "fvo.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NV 5
static const int kMom[3] = {1, 2, 4};
static const double kEosEps = 1e-10;    /* Core.hpp:169 */
static const double kDenomEps = 1e-14;  /* Schemes.cpp:63 */
static const double kMachEps = 1e-12;   /* Schemes.cpp:65 */
static const double kDtEps = 1e-10;     /* TimeScheme.cpp:20 */

static int g_threads = 0;
void fvo_set_threads(int n) { g_threads = n; }

/* 1: every WENO5-Z face value takes the linear weights (upwind-5), the smooth-face value of the learned
 * indicator; 0 (default): the reference's WENO5-Z. */
static int g_weno_linear = 0;
void fvo_set_weno_linear(int on) { g_weno_linear = on; }
int fvo_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define NT fvo_get_threads()

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------------------------------ */
/* Mesh (Core.hpp:68-156)                                                                           */
/* ------------------------------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz, is3d;
    double dx, dy, dz;
    long long nc, nxf, nyf, nzf, nf;
} Mesh;

static Mesh make_mesh(const fvg_config* c) {
    Mesh m;
    m.nx = c->nx;
    m.ny = c->ny;
    m.nz = c->nz;
    m.is3d = c->nz > 1;
    m.dx = c->Lx / c->nx;
    m.dy = c->Ly / c->ny;
    m.dz = c->Lz / c->nz;   /* generate_cartesian_mesh_3d; unused by the 2-D formulas */
    m.nc = (long long)m.nx * m.ny * m.nz;
    m.nxf = (long long)(m.nx + 1) * m.ny * m.nz;
    m.nyf = (long long)m.nx * (m.ny + 1) * m.nz;
    m.nzf = m.is3d ? (long long)m.nx * m.ny * (m.nz + 1) : 0;
    m.nf = m.nxf + m.nyf + m.nzf;
    return m;
}
static inline long long CELL(const Mesh* m, int i, int j, int k) { return ((long long)k * m->ny + j) * m->nx + i; }
static inline long long XF(const Mesh* m, int i, int j, int k) { return ((long long)k * m->ny + j) * (m->nx + 1) + i; }
static inline long long YF(const Mesh* m, int i, int j, int k) { return m->nxf + ((long long)k * (m->ny + 1) + j) * m->nx + i; }
static inline long long ZF(const Mesh* m, int i, int j, int k) { return m->nxf + m->nyf + ((long long)k * m->ny + j) * m->nx + i; }
/* face_area (Core.hpp:96-99) */
static double face_area(const Mesh* m, int axis) {
    if (!m->is3d) return axis == 0 ? m->dy : m->dx;
    return axis == 0 ? m->dy * m->dz : (axis == 1 ? m->dx * m->dz : m->dx * m->dy);
}
static double cell_volume(const Mesh* m) { return m->is3d ? m->dx * m->dy * m->dz : m->dx * m->dy; }

double fvo_dt_length(const fvg_config* c) {
    /* Simulation.cpp:225-227 with ViscousBound.hpp:28-29 */
    const double dx = c->Lx / c->nx, dy = c->Ly / c->ny;
    if (c->nz > 1) {
        const double dz = c->Lz / c->nz;
        return 2.0 / (1.0 / dx + 1.0 / dy + 1.0 / dz);
    }
    return smin(dx, dy);
}

/* ------------------------------------------------------------------------------------------------ */
/* EOS (Core.hpp:173-238)                                                                           */
/* ------------------------------------------------------------------------------------------------ */
typedef struct {
    int stiff;
    double gamma, p_inf;
} Eos;

static inline void p2c(const Eos* e, const double* W, double* U) {
    const double rho = W[0], u = W[1], v = W[2], p = W[3], w = W[4];
    const double ke = 0.5 * rho * (u * u + v * v + w * w);
    const double en = !e->stiff ? p / (e->gamma - 1.0) : (p + e->gamma * e->p_inf) / (e->gamma - 1.0);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = en + ke;
    U[4] = rho * w;
}
static inline void c2p(const Eos* e, const double* U, double* W, double* inv_rho) {
    const double rho = smax(U[0], kEosEps);
    const double inv = 1.0 / rho;
    if (inv_rho) *inv_rho = inv;
    const double u = U[1] * inv, v = U[2] * inv, w = U[4] * inv;
    const double e_kin = 0.5 * rho * (u * u + v * v + w * w);
    const double rho_e = U[3] - e_kin;
    const double p = !e->stiff ? (e->gamma - 1.0) * rho_e : (e->gamma - 1.0) * rho_e - e->gamma * e->p_inf;
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = smax(p, kEosEps);
    W[4] = w;
}
static inline double sound_speed(const Eos* e, double p, double rho) {
    if (!e->stiff) return sqrt(e->gamma * smax(p, kEosEps) / smax(rho, kEosEps));
    return sqrt(e->gamma * (p + e->p_inf) / smax(rho, kEosEps));
}
static inline double sound_speed_inv(const Eos* e, double p, double inv_rho) {
    if (!e->stiff) return sqrt(e->gamma * smax(p, kEosEps) * inv_rho);
    return sqrt(e->gamma * (p + e->p_inf) * inv_rho);
}
static inline void roundtrip(const Eos* e, double* U) {
    double W[NV];
    c2p(e, U, W, NULL);
    p2c(e, W, U);
}

/* ------------------------------------------------------------------------------------------------ */
/* Riemann solvers (Schemes.cpp:104-229)                                                            */
/* ------------------------------------------------------------------------------------------------ */
typedef struct {
    Eos eos;
    int lowmach;
    double blend;
} Riemann;

static inline double safe_denom(double x) { return copysign(smax(fabs(x), kDenomEps), x); }

static void hllc_core(const Riemann* rs, const double* WL, const double* WR, const double* UL, const double* UR, double inv_L,
                      double inv_R, double c_L, double c_R, int axis, double* F, double* smax_out) {
    const double rho_L = WL[0], p_L = WL[3];
    const double rho_R = WR[0], p_R = WR[3];
    const double vel_L[3] = {WL[1], WL[2], WL[4]};
    const double vel_R[3] = {WR[1], WR[2], WR[4]};
    const double E_L = UL[3] * inv_L, E_R = UR[3] * inv_R;
    const double Vn_L = vel_L[axis], Vn_R = vel_R[axis];
    const int mn = kMom[axis];
    double FL[NV], FR[NV];
    FL[0] = rho_L * Vn_L;
    FR[0] = rho_R * Vn_R;
    for (int a = 0; a < 3; ++a) {
        FL[kMom[a]] = rho_L * vel_L[a] * Vn_L;
        FR[kMom[a]] = rho_R * vel_R[a] * Vn_R;
    }
    FL[mn] = FL[mn] + p_L;
    FR[mn] = FR[mn] + p_R;
    FL[3] = Vn_L * (rho_L * E_L + p_L);
    FR[3] = Vn_R * (rho_R * E_R + p_R);
    const double S_L = smin(Vn_L - c_L, Vn_R - c_R);
    const double S_R = smax(Vn_L + c_L, Vn_R + c_R);
    const double num = rho_R * Vn_R * (S_R - Vn_R) - rho_L * Vn_L * (S_L - Vn_L) + p_L - p_R;
    const double den = rho_R * (S_R - Vn_R) - rho_L * (S_L - Vn_L);
    const double S_M = num / safe_denom(den);
    const double den_star_L = safe_denom(S_L - S_M);
    const double den_star_R = safe_denom(S_R - S_M);
    const double den_pL = safe_denom(S_L - Vn_L);
    const double den_pR = safe_denom(S_R - Vn_R);
    const double fL = rho_L * (S_L - Vn_L) / den_star_L;
    const double fR = rho_R * (S_R - Vn_R) / den_star_R;
    double UsL[NV], UsR[NV];
    UsL[0] = fL;
    UsR[0] = fR;
    for (int a = 0; a < 3; ++a) {
        UsL[kMom[a]] = fL * (a == axis ? S_M : vel_L[a]);
        UsR[kMom[a]] = fR * (a == axis ? S_M : vel_R[a]);
    }
    UsL[3] = fL * (E_L + (S_M - Vn_L) * (S_M + p_L / (rho_L * den_pL)));
    UsR[3] = fR * (E_R + (S_M - Vn_R) * (S_M + p_R / (rho_R * den_pR)));
    double Fo[NV] = {0, 0, 0, 0, 0};
    if (S_L >= 0)
        for (int q = 0; q < NV; ++q) Fo[q] = FL[q];
    if (S_L < 0 && S_M >= 0)
        for (int q = 0; q < NV; ++q) Fo[q] = FL[q] + S_L * (UsL[q] - UL[q]);
    if (S_M < 0 && S_R >= 0)
        for (int q = 0; q < NV; ++q) Fo[q] = FR[q] + S_R * (UsR[q] - UR[q]);
    if (S_R < 0)
        for (int q = 0; q < NV; ++q) Fo[q] = FR[q];
    /* shock-face blend toward HLL, Schemes.cpp:180-196 */
    if (rs->blend > 0.0) {
        const double s = fabs(p_R - p_L) / smin(p_L, p_R);
        if (s > rs->blend) {
            const double phi = smin((s - rs->blend) / rs->blend, 1.0);
            double Fh[NV];
            if (S_L >= 0) {
                for (int q = 0; q < NV; ++q) Fh[q] = FL[q];
            } else if (S_R <= 0) {
                for (int q = 0; q < NV; ++q) Fh[q] = FR[q];
            } else {
                const double inv = 1.0 / (S_R - S_L);
                for (int q = 0; q < NV; ++q) Fh[q] = (S_R * FL[q] - S_L * FR[q] + S_L * S_R * (UR[q] - UL[q])) * inv;
            }
            for (int q = 0; q < NV; ++q) Fo[q] = (1.0 - phi) * Fo[q] + phi * Fh[q];
        }
    }
    *smax_out = smax(fabs(S_L), fabs(S_R));
    for (int q = 0; q < NV; ++q) F[q] = Fo[q];
}

/* RiemannSolver::face_flux and hllc (Schemes.cpp:108-113, 202-208) and lowmach_flux (:213-229) */
static void face_flux(const Riemann* rs, const double* UL, const double* UR, int axis, double* F, double* sm) {
    double inv_L, inv_R, WL[NV], WR[NV];
    c2p(&rs->eos, UL, WL, &inv_L);
    c2p(&rs->eos, UR, WR, &inv_R);
    if (!rs->lowmach) {
        hllc_core(rs, WL, WR, UL, UR, inv_L, inv_R, sound_speed_inv(&rs->eos, WL[3], inv_L), sound_speed_inv(&rs->eos, WR[3], inv_R),
                  axis, F, sm);
        return;
    }
    const double c_L = sound_speed_inv(&rs->eos, WL[3], inv_L);
    const double c_R = sound_speed_inv(&rs->eos, WR[3], inv_R);
    const double M_L = sqrt(WL[1] * WL[1] + WL[2] * WL[2] + WL[4] * WL[4]) / (c_L + kMachEps);
    const double M_R = sqrt(WR[1] * WR[1] + WR[2] * WR[2] + WR[4] * WR[4]) / (c_R + kMachEps);
    const double z = smin(smax(smax(M_L, M_R), 1e-4), 1.0);
    for (int a = 0; a < 3; ++a) {
        const int m = kMom[a];
        const double avg = 0.5 * (WL[m] + WR[m]);
        const double half = 0.5 * (WL[m] - WR[m]);
        WL[m] = avg + z * half;
        WR[m] = avg - z * half;
    }
    double UsL[NV], UsR[NV];
    p2c(&rs->eos, WL, UsL);
    p2c(&rs->eos, WR, UsR);
    hllc_core(rs, WL, WR, UsL, UsR, inv_L, inv_R, c_L, c_R, axis, F, sm);
}

static Eos make_eos(const fvg_config* c) {
    Eos e;
    e.stiff = c->eos_type == FVG_EOS_STIFFENED;
    e.gamma = c->gamma;
    e.p_inf = e.stiff ? c->p_inf : 0.0;
    return e;
}
static Riemann make_riemann(const fvg_config* c) {
    Riemann r;
    r.eos = make_eos(c);
    r.lowmach = c->riemann == FVG_LOWMACH_HLLC;
    r.blend = c->shock_blend;
    return r;
}

void fvo_face_flux(const fvg_config* cfg, const double* UL, const double* UR, int axis, double* F, double* sm) {
    const Riemann rs = make_riemann(cfg);
    face_flux(&rs, UL, UR, axis, F, sm);
}
void fvo_face_flux_batch(const fvg_config* cfg, const double* UL, const double* UR, const int* axis, size_t n, double* F,
                         double* sm) {
    const Riemann rs = make_riemann(cfg);
#pragma omp parallel for schedule(static) num_threads(NT)
    for (long long k = 0; k < (long long)n; ++k) face_flux(&rs, UL + NV * k, UR + NV * k, axis[k], F + NV * k, sm + k);
}

/* ------------------------------------------------------------------------------------------------ */
/* Limiters (Schemes.hpp:90-117) and WENO-Z (Schemes.cpp:323-353)                                  */
/* ------------------------------------------------------------------------------------------------ */
static inline double sgn(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }
static inline double lim_apply(int l, double a, double b) {
    switch (l) {
    case FVG_LIM_MINMOD: return 0.5 * (sgn(a) + sgn(b)) * smin(fabs(a), fabs(b));
    case FVG_LIM_VAN_LEER: {
        const double ab = a * b;
        return (ab > 0) ? 2.0 * ab / (a + b) : 0.0;
    }
    case FVG_LIM_MC:
        if (!(a * b > 0)) return 0.0;
        return smin(smin(2.0 * fabs(a), 0.5 * fabs(a + b)), 2.0 * fabs(b)) * sgn(a);
    default:
        if (!(a * b > 0)) return 0.0;
        return fmin(fmin(2.0 * fabs(a), fabs(2.0 * a + b) / 3.0), 2.0 * fabs(b)) * sgn(a);
    }
}
static inline double weno3z(double c0, double c1, double c2, double eps_rel) {
    const double eps = eps_rel * (c0 * c0 + c1 * c1 + c2 * c2) * (1.0 / 3.0) + 1e-300;
    const double v0 = -0.5 * c0 + 1.5 * c1, v1 = 0.5 * c1 + 0.5 * c2;
    const double b0 = (c1 - c0) * (c1 - c0), b1 = (c2 - c1) * (c2 - c1);
    const double tau = fabs(b0 - b1);
    const double a0 = (1.0 / 3.0) * (1.0 + tau / (b0 + eps)), a1 = (2.0 / 3.0) * (1.0 + tau / (b1 + eps));
    return (a0 * v0 + a1 * v1) / (a0 + a1);
}
static inline double weno5z(double c0, double c1, double c2, double c3, double c4, double eps_rel) {
    const double eps = eps_rel * (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4) * 0.2 + 1e-300;
    const double p0 = (2.0 * c0 - 7.0 * c1 + 11.0 * c2) / 6.0;
    const double p1 = (-c1 + 5.0 * c2 + 2.0 * c3) / 6.0;
    const double p2 = (2.0 * c2 + 5.0 * c3 - c4) / 6.0;
    const double s0 = c0 - 2.0 * c1 + c2, t0 = c0 - 4.0 * c1 + 3.0 * c2;
    const double s1 = c1 - 2.0 * c2 + c3, t1 = c1 - c3;
    const double s2 = c2 - 2.0 * c3 + c4, t2 = 3.0 * c2 - 4.0 * c3 + c4;
    const double b0 = (13.0 / 12.0) * s0 * s0 + 0.25 * t0 * t0;
    const double b1 = (13.0 / 12.0) * s1 * s1 + 0.25 * t1 * t1;
    const double b2 = (13.0 / 12.0) * s2 * s2 + 0.25 * t2 * t2;
    const double tau = fabs(b0 - b2);
    const double a0 = 0.1 * (1.0 + tau / (b0 + eps)), a1 = 0.6 * (1.0 + tau / (b1 + eps)), a2 = 0.3 * (1.0 + tau / (b2 + eps));
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}

/* ------------------------------------------------------------------------------------------------ */
/* Boundary conditions (Boundaries.cpp)                                                             */
/* ------------------------------------------------------------------------------------------------ */
static int side_axis(int side) { return side / 2; }
static int side_high(int side) { return side == FVG_RIGHT || side == FVG_TOP || side == FVG_BACK; }

/* ghost_state (Boundaries.cpp:106-153); u_target: the mass-flow inlet's velocity of this side (apply_side). A
 * periodic or non-reflecting side has no single-cell ghost (-1). */
static int ghost_state(const fvg_config* c, const Eos* e, int side, const double* Ui, double* Ug, double u_target) {
    const int bc = c->bc[side];
    if (bc == FVG_BC_TRANSMISSIVE) {
        for (int q = 0; q < NV; ++q) Ug[q] = Ui[q];
        return 0;
    }
    if (bc == FVG_BC_PERIODIC || bc == FVG_BC_NONREFLECTING) return -1;
    const int axis = side_axis(side);
    const double nxn = axis == 0 ? 1.0 : 0.0, nyn = axis == 1 ? 1.0 : 0.0, nzn = axis == 2 ? 1.0 : 0.0;
    double Wi[NV], Wg[NV];
    c2p(e, Ui, Wi, NULL);
    for (int q = 0; q < NV; ++q) Wg[q] = Wi[q];
    switch (bc) {
    case FVG_BC_WALL: {
        const double Vn = Wi[1] * nxn + Wi[2] * nyn + Wi[4] * nzn;
        Wg[1] = Wi[1] - 2.0 * Vn * nxn;
        Wg[2] = Wi[2] - 2.0 * Vn * nyn;
        Wg[4] = Wi[4] - 2.0 * Vn * nzn;
        break;
    }
    case FVG_BC_NO_SLIP:
        Wg[1] = -Wi[1];
        Wg[2] = -Wi[2];
        Wg[4] = -Wi[4];
        break;
    case FVG_BC_MOVING_WALL:
    case FVG_BC_VEL_INLET:
        Wg[1] = 2.0 * c->u_in - Wi[1];
        Wg[2] = 2.0 * c->v_in - Wi[2];
        Wg[4] = -Wi[4];
        break;
    case FVG_BC_MASSFLOW_INLET:
        Wg[1] = 2.0 * u_target - Wi[1];
        Wg[2] = 2.0 * c->v_in - Wi[2];
        Wg[4] = -Wi[4];
        break;
    case FVG_BC_OUTLET: {
        const double Vn = Wi[1] * nxn + Wi[2] * nyn + Wi[4] * nzn;
        const int outflow = Vn >= 0;
        Wg[1] = outflow ? Wi[1] : 0.0;
        Wg[2] = outflow ? Wi[2] : 0.0;
        Wg[4] = outflow ? Wi[4] : 0.0;
        Wg[3] = smax(c->p_out, 1e-5);
        break;
    }
    default: break;
    }
    p2c(e, Wg, Ug);
    return 0;
}

/* Number of boundary faces and their face / adjacent cell (Core.hpp:121-151). */
static long long boundary_count(const Mesh* m, int side) {
    switch (side) {
    case FVG_LEFT:
    case FVG_RIGHT: return (long long)m->ny * m->nz;
    case FVG_BOTTOM:
    case FVG_TOP: return (long long)m->nx * m->nz;
    default: return m->is3d ? (long long)m->nx * m->ny : 0;
    }
}
static long long boundary_face(const Mesh* m, int side, long long k) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    switch (side) {
    case FVG_LEFT: return k * (nx + 1);
    case FVG_RIGHT: return k * (nx + 1) + nx;
    case FVG_BOTTOM: return m->nxf + (k / nx) * (ny + 1) * nx + k % nx;
    case FVG_TOP: return m->nxf + ((k / nx) * (ny + 1) + ny) * nx + k % nx;
    case FVG_FRONT: return m->nxf + m->nyf + k;
    default: return m->nxf + m->nyf + (long long)nz * ny * nx + k;
    }
}
static long long boundary_cell(const Mesh* m, int side, long long k) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    switch (side) {
    case FVG_LEFT: return k * nx;
    case FVG_RIGHT: return k * nx + nx - 1;
    case FVG_BOTTOM: return (k / nx) * ny * nx + k % nx;
    case FVG_TOP: return ((k / nx) * ny + ny - 1) * nx + k % nx;
    case FVG_FRONT: return k;
    default: return (long long)(nz - 1) * ny * nx + k;
    }
}

/* The mass-flow inlet velocity of every side (apply_side, Boundaries.cpp:285-294): rho_in u_in over the mean
 * density of the side's boundary cells (summed in boundary order), floored at 1e-5; 0 on the other sides. */
static void massflow_targets(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, double ut[6]) {
    for (int side = 0; side < 6; ++side) {
        ut[side] = 0.0;
        if (c->bc[side] != FVG_BC_MASSFLOW_INLET || (side >= 4 && !m->is3d)) continue;
        const long long n = boundary_count(m, side);
        double sum = 0.0;
        const double count = (double)n;
        for (long long k = 0; k < n; ++k) {
            double W[NV];
            c2p(e, U + NV * boundary_cell(m, side, k), W, NULL);
            sum += W[0];
        }
        const double rho_avg = smax(sum / count, 1e-5);
        ut[side] = (c->rho_in * c->u_in) / rho_avg;
    }
}

/* nonreflecting_ghost (Boundaries.cpp:155-198) with the tangential neighbours of apply_side (Boundaries.cpp:314-360):
 * a characteristic outlet whose ghost differs from the cell only in the incoming acoustic wave, relaxed towards
 * p_out (nrbc_sigma, the domain length along the axis) with the transverse terms weighted by 1 - beta
 * (beta = nrbc_beta, or |M| when nrbc_beta < 0). Tangential derivatives are central differences of the primitives,
 * wrapped across a periodic axis, one-sided at the ends of the others. */
static void nonreflecting_ghost(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, int side, long long cell,
                                double* Ug) {
    const int isR = side_high(side), axis = side_axis(side);
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const int n3[3] = {nx, ny, nz};
    const int ijk[3] = {(int)(cell % nx), (int)((cell / nx) % ny), (int)(cell / ((long long)nx * ny))};
    int use[3] = {0, 0, 0};
    double Wlo[3][NV], Whi[3][NV], span[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < 3; ++t) {
        if (t == axis || n3[t] == 1) continue;
        const int wrap = c->bc[2 * t] == FVG_BC_PERIODIC && c->bc[2 * t + 1] == FVG_BC_PERIODIC;
        int lo = ijk[t] - 1, hi = ijk[t] + 1;
        double sp = 2.0;
        if (wrap) {
            lo = (lo + n3[t]) % n3[t];
            hi = hi % n3[t];
        } else {
            if (lo < 0) {
                lo = 0;
                sp -= 1.0;
            }
            if (hi >= n3[t]) {
                hi = n3[t] - 1;
                sp -= 1.0;
            }
            if (sp <= 0.0) continue;
        }
        int il[3] = {ijk[0], ijk[1], ijk[2]}, ih[3] = {ijk[0], ijk[1], ijk[2]};
        il[t] = lo;
        ih[t] = hi;
        use[t] = 1;
        c2p(e, U + NV * CELL(m, il[0], il[1], il[2]), Wlo[t], NULL);
        c2p(e, U + NV * CELL(m, ih[0], ih[1], ih[2]), Whi[t], NULL);
        span[t] = sp;
    }
    const double hs[3] = {m->dx, m->dy, m->dz};
    const double length[3] = {c->Lx, c->Ly, c->Lz};
    double Wi[NV], Wg[NV];
    c2p(e, U + NV * cell, Wi, NULL);
    for (int q = 0; q < NV; ++q) Wg[q] = Wi[q];
    const double s = isR ? 1.0 : -1.0;
    const int qv = kMom[axis];
    const double u_out = s * Wi[qv];
    const double cs = sound_speed(e, Wi[3], Wi[0]);
    const double M = smin(smax(u_out / cs, -1.0), 1.0);
    if (M >= 1.0) {
        p2c(e, Wg, Ug);
        return;
    }
    const double h = hs[axis];
    const double K = c->nrbc_sigma * cs * (1.0 - M * M) / length[axis];
    const double beta = c->nrbc_beta >= 0.0 ? c->nrbc_beta : fabs(M);
    double T1 = 0.0;
    if (beta < 1.0)
        for (int t = 0; t < 3; ++t) {
            if (!use[t]) continue;
            const double* Wl = Wlo[t];
            const double* Wh = Whi[t];
            const double ht = span[t] * hs[t];
            const int qt = kMom[t];
            T1 += Wi[qt] * (Wh[3] - Wl[3]) / ht + Wi[0] * cs * cs * (Wh[qt] - Wl[qt]) / ht -
                  Wi[0] * cs * s * Wi[qt] * (Wh[qv] - Wl[qv]) / ht;
        }
    const double dw1 = -h * (K * (Wi[3] - c->p_out) - (1.0 - beta) * T1) / (cs - u_out);
    Wg[3] = Wi[3] + 0.5 * dw1;
    Wg[qv] = Wi[qv] - s * dw1 / (2.0 * Wi[0] * cs);
    Wg[0] = Wi[0] + 0.5 * dw1 / (cs * cs);
    p2c(e, Wg, Ug);
}

/* apply_boundary_conditions + apply_side (Boundaries.cpp:54-104, 259-371): periodic, ghost_state and non-reflecting
 * sides; ut: the mass-flow targets of massflow_targets. */
static int apply_boundary_conditions(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, double* UL, double* UR,
                                     int interior, const double ut[6]) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    if (interior) {
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int kj = 0; kj < ny * nz; ++kj) {
            const int j = kj % ny, k = kj / ny;
            for (int i = 1; i < nx; ++i) {
                const long long f = XF(m, i, j, k);
                memcpy(UL + NV * f, U + NV * CELL(m, i - 1, j, k), sizeof(double) * NV);
                memcpy(UR + NV * f, U + NV * CELL(m, i, j, k), sizeof(double) * NV);
            }
        }
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int kj = 0; kj < (ny - 1) * nz; ++kj) {
            const int j = kj % (ny - 1) + 1, k = kj / (ny - 1);
            for (int i = 0; i < nx; ++i) {
                const long long f = YF(m, i, j, k);
                memcpy(UL + NV * f, U + NV * CELL(m, i, j - 1, k), sizeof(double) * NV);
                memcpy(UR + NV * f, U + NV * CELL(m, i, j, k), sizeof(double) * NV);
            }
        }
        if (m->is3d) {
#pragma omp parallel for schedule(static) num_threads(NT)
            for (int kj = 0; kj < ny * (nz - 1); ++kj) {
                const int j = kj % ny, k = kj / ny + 1;
                for (int i = 0; i < nx; ++i) {
                    const long long f = ZF(m, i, j, k);
                    memcpy(UL + NV * f, U + NV * CELL(m, i, j, k - 1), sizeof(double) * NV);
                    memcpy(UR + NV * f, U + NV * CELL(m, i, j, k), sizeof(double) * NV);
                }
            }
        }
    }
    static const int order[6] = {FVG_RIGHT, FVG_TOP, FVG_LEFT, FVG_BOTTOM, FVG_BACK, FVG_FRONT};
    for (int s = 0; s < (m->is3d ? 6 : 4); ++s) {
        const int side = order[s];
        const long long n = boundary_count(m, side);
        const int bc = c->bc[side];
        const int isR = side_high(side);
        for (long long k = 0; k < n; ++k) {
            const long long f = boundary_face(m, side, k), cell = boundary_cell(m, side, k);
            const double* Ui = U + NV * cell;
            double Ug[NV];
            if (bc == FVG_BC_PERIODIC) {
                long long opp;
                switch (side) {
                case FVG_LEFT: opp = cell + (nx - 1); break;
                case FVG_RIGHT: opp = cell - (nx - 1); break;
                case FVG_BOTTOM: opp = cell + (long long)(ny - 1) * nx; break;
                case FVG_TOP: opp = cell - (long long)(ny - 1) * nx; break;
                case FVG_FRONT: opp = cell + (long long)(nz - 1) * ny * nx; break;
                default: opp = cell - (long long)(nz - 1) * ny * nx; break;
                }
                memcpy(Ug, U + NV * opp, sizeof Ug);
            } else if (bc == FVG_BC_NONREFLECTING) {
                nonreflecting_ghost(c, m, e, U, side, cell, Ug);
            } else if (ghost_state(c, e, side, Ui, Ug, ut[side]) != 0) {
                return -1;
            }
            if (isR) {
                memcpy(UL + NV * f, Ui, sizeof Ug);
                memcpy(UR + NV * f, Ug, sizeof Ug);
            } else {
                memcpy(UL + NV * f, Ug, sizeof Ug);
                memcpy(UR + NV * f, Ui, sizeof Ug);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* Reconstruction on the padded primitive state (Schemes.cpp:389-700, 915-969)                       */
/* ------------------------------------------------------------------------------------------------ */
/* Which sides have a halo (Simulation.cpp:863-905 2-D, 1109-1170 3-D): periodic wrap when both sides of an
 * axis are periodic and the axis has at least Dh cells; else, with boundary_order = 2, a ghost slab on a
 * side that is neither periodic nor non-reflecting and has at least Dh cells (fill_ghost_slab). */
static void halo_sides(const fvg_config* c, const Mesh* m, int Dh, int has[6], int wrap[3]) {
    const int n3[3] = {m->nx, m->ny, m->nz};
    const int naxes = m->is3d ? 3 : 2;
    for (int s = 0; s < 6; ++s) has[s] = 0;
    for (int a = 0; a < 3; ++a) wrap[a] = 0;
    for (int a = 0; a < naxes; ++a) {
        const int per = c->bc[2 * a] == FVG_BC_PERIODIC && c->bc[2 * a + 1] == FVG_BC_PERIODIC;
        if (per && n3[a] >= Dh) {
            wrap[a] = 1;
            has[2 * a] = has[2 * a + 1] = 1;
            continue;
        }
        if (c->boundary_order == 2)
            for (int hi = 0; hi < 2; ++hi) {
                const int bt = c->bc[2 * a + hi];
                if (bt == FVG_BC_PERIODIC || bt == FVG_BC_NONREFLECTING) continue;
                if (n3[a] >= Dh) has[2 * a + hi] = 2;
            }
    }
}

typedef struct {
    int pad, zp, PX, PY, PZ;
    size_t plane;
    double* W;   /* 5 planes */
} Padded;

/* The conservative state of the halo cell `l` cells beyond side `side` (l = 0 next to it) in the transverse
 * position (i, j, k) of the padded cell: periodic wrap or the ghost slab layer (fill_ghost_slab,
 * Boundaries.cpp:202-228: layer l is the ghost of the interior cell l cells from the boundary). */
static int halo_state(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, int side, int wrap, int i, int j, int k,
                      double* out, const double ut[6]) {
    const int axis = side_axis(side), hi = side_high(side);
    const int n3[3] = {m->nx, m->ny, m->nz};
    int idx[3] = {i, j, k};
    const int t = idx[axis];   /* outside: t < 0 (low) or t >= n (high) */
    const int l = hi ? t - n3[axis] : -1 - t;
    if (wrap) {
        idx[axis] = hi ? l : n3[axis] - 1 - l;
        memcpy(out, U + NV * CELL(m, idx[0], idx[1], idx[2]), sizeof(double) * NV);
        return 0;
    }
    idx[axis] = hi ? n3[axis] - 1 - l : l;
    return ghost_state(c, e, side, U + NV * CELL(m, idx[0], idx[1], idx[2]), out, ut[side]);
}

static inline int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* pad_primitives (Schemes.cpp:915-969): clamp-replicated primitive state (prim = 0: the conservative state as it
 * is), with the halo slabs in place. */
static int pad_primitives(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, int pad, const int has[6],
                          const int wrap[3], Padded* P, int prim, const double ut[6]) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    P->pad = pad;
    P->zp = m->is3d ? pad : 0;
    P->PX = nx + 2 * pad;
    P->PY = ny + 2 * pad;
    P->PZ = nz + 2 * P->zp;
    P->plane = (size_t)P->PX * P->PY * P->PZ;
    P->W = (double*)malloc(sizeof(double) * NV * P->plane);
    if (!P->W) return -2;
    int bad = 0;
#pragma omp parallel for schedule(static) num_threads(NT) reduction(| : bad)
    for (int kj = 0; kj < P->PZ * P->PY; ++kj) {
        const int kk = kj / P->PY, jj = kj % P->PY;
        const int k = kk - P->zp, j = jj - pad;
        for (int ii = 0; ii < P->PX; ++ii) {
            const int i = ii - pad;
            const int oi = i < 0 || i >= nx, oj = j < 0 || j >= ny, ok = k < 0 || k >= nz;
            double Uc[NV], W[NV];
            const double* src = U + NV * CELL(m, clampi(i, nx), clampi(j, ny), clampi(k, nz));
            if (oi + oj + ok == 1) {
                const int axis = oi ? 0 : (oj ? 1 : 2);
                const int t = axis == 0 ? i : (axis == 1 ? j : k);
                const int side = 2 * axis + (t >= 0);
                if (has[side]) {
                    if (halo_state(c, m, e, U, side, wrap[axis], i, j, k, Uc, ut) != 0) bad = 1;
                    src = Uc;
                }
            }
            if (prim) c2p(e, src, W, NULL);
            else memcpy(W, src, sizeof W);
            const size_t p = ((size_t)kk * P->PY + jj) * P->PX + ii;
            for (int q = 0; q < NV; ++q) P->W[q * P->plane + p] = W[q];
        }
    }
    return bad ? -1 : 0;
}

/* Face states of every face (reconstruct / reconstruct_3d / reconstruct_weno, Schemes.cpp:389-749):
 * faces with a full stencil are reconstructed from the padded primitives (MUSCL on 4 cells, WENO3-Z on 4,
 * WENO5-Z on 6) and converted back; the other (side) faces keep their first-order states through the
 * conservative-primitive round trip. */
static void reconstruct_faces(const fvg_config* c, const Mesh* m, const Eos* e, const Padded* P, const int has[6], double* UL,
                              double* UR) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const int scheme = c->reconstruction, lim = c->limiter;
    const double we = c->weno_epsilon;
    const int pad = P->pad;
    for (int axis = 0; axis < (m->is3d ? 3 : 2); ++axis) {
        const int n_a = axis == 0 ? nx : (axis == 1 ? ny : nz);
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)P->PX : (size_t)P->PX * P->PY);
        const int ni = axis == 0 ? nx + 1 : nx, nj = axis == 1 ? ny + 1 : ny, nk = axis == 2 ? nz + 1 : nz;
        const int lo = has[2 * axis] != 0, hi = has[2 * axis + 1] != 0;
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int kj = 0; kj < nk * nj; ++kj) {
            const int j = kj % nj, k = kj / nj;
            for (int i = 0; i < ni; ++i) {
                const int t = axis == 0 ? i : (axis == 1 ? j : k);
                const long long f = axis == 0 ? XF(m, i, j, k) : (axis == 1 ? YF(m, i, j, k) : ZF(m, i, j, k));
                const int upgrade = (t > 0 && t < n_a) || (t == 0 && lo) || (t == n_a && hi);
                if (!upgrade) {
                    roundtrip(e, UL + NV * f);
                    roundtrip(e, UR + NV * f);
                    continue;
                }
                /* padded index of the first stencil cell: cell t - pad along the axis */
                const int pi = axis == 0 ? t : i + pad, pj = axis == 1 ? t : j + pad, pk = axis == 2 ? t : k + P->zp;
                const size_t base = ((size_t)pk * P->PY + pj) * P->PX + pi;
                double wl[NV], wr[NV];
                for (int q = 0; q < NV; ++q) {
                    const double* cq = P->W + q * P->plane + base;
                    if (scheme == FVG_RECON_MUSCL) {
                        /* pad 2: stencil t-2 .. t+1 */
                        const double c0 = cq[0], c1 = cq[stride], c2 = cq[2 * stride], c3 = cq[3 * stride];
                        const double b = c2 - c1;
                        wl[q] = c1 + 0.5 * lim_apply(lim, c1 - c0, b);
                        wr[q] = c2 - 0.5 * lim_apply(lim, c3 - c2, b);
                    } else if (scheme == FVG_RECON_WENO3) {
                        wl[q] = weno3z(cq[0], cq[stride], cq[2 * stride], we);
                        wr[q] = weno3z(cq[3 * stride], cq[2 * stride], cq[stride], we);
                    } else {
                        const double c0 = cq[0], c1 = cq[stride], c2 = cq[2 * stride], c3 = cq[3 * stride], c4 = cq[4 * stride],
                                     c5 = cq[5 * stride];
                        if (g_weno_linear) {   /* the weighted sum times the double nearest 1/60 */
                            wl[q] = (2.0 * c0 - 13.0 * c1 + 47.0 * c2 + 27.0 * c3 - 3.0 * c4) * (1.0 / 60.0);
                            wr[q] = (2.0 * c5 - 13.0 * c4 + 47.0 * c3 + 27.0 * c2 - 3.0 * c1) * (1.0 / 60.0);
                        } else {
                            wl[q] = weno5z(c0, c1, c2, c3, c4, we);
                            wr[q] = weno5z(c5, c4, c3, c2, c1, we);
                        }
                    }
                }
                p2c(e, wl, UL + NV * f);
                p2c(e, wr, UR + NV * f);
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* flux_quadrature = fourth (Schemes.cpp:1210-1349, 1954-2134)                                      */
/* ------------------------------------------------------------------------------------------------ */
/* p > 0 and rho > 0 without a division: g1 (E rho - |m|^2 / 2) > gp rho (Schemes.cpp:1241-1243, 1967-1968) */
static inline double cons_m2(const double* u) { return u[1] * u[1] + u[2] * u[2] + u[4] * u[4]; }
static inline int cons_positive(double g1, double gp, const double* u) {
    return u[0] > 0.0 && g1 * (u[3] * u[0] - 0.5 * cons_m2(u)) > gp * u[0];
}
static inline double cons_pressure(double g1, double gp, const double* u) { return g1 * (u[3] - 0.5 * cons_m2(u) / u[0]) - gp; }

/* Branch counts of the fourth-order fluxes, collected when g_q4_collect is set (fvo_fourth_stats, tests only):
 * faces reconstructed, admissibility fallbacks, good, pt, full, pt without full, good without pt. */
static long long g_q4_stats[7];
static int g_q4_collect = 0;

/* MusclReconstruction::reconstruct_conserved / ConservedLines::line (Schemes.cpp:1210-1349): WENO3-Z (four cells)
 * or WENO5-Z (six cells) of the conservative averages at every face with a full stencil; a face whose two states
 * are not positive, or whose density or pressure leaves the range of its stencil cells widened by the range's own
 * width, takes the two cells next to it. The other faces keep their first-order states as they are. */
static void reconstruct_conserved(const fvg_config* c, const Mesh* m, const Eos* e, const Padded* P, const int has[6], double* UL,
                                  double* UR) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const int five = c->reconstruction == FVG_RECON_WENO5;
    const double we = c->weno_epsilon;
    const double g1 = e->gamma - 1.0, gp = e->stiff ? e->gamma * e->p_inf : 0.0;
    const int pad = P->pad, left = five ? 2 : 1, ncell = five ? 6 : 4;
    long long n_up = 0, n_fb = 0;
    for (int axis = 0; axis < (m->is3d ? 3 : 2); ++axis) {
        const int n_a = axis == 0 ? nx : (axis == 1 ? ny : nz);
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)P->PX : (size_t)P->PX * P->PY);
        const int ni = axis == 0 ? nx + 1 : nx, nj = axis == 1 ? ny + 1 : ny, nk = axis == 2 ? nz + 1 : nz;
        const int lo = has[2 * axis] != 0, hi = has[2 * axis + 1] != 0;
#pragma omp parallel for schedule(static) num_threads(NT) reduction(+ : n_up, n_fb)
        for (int kj = 0; kj < nk * nj; ++kj) {
            const int j = kj % nj, k = kj / nj;
            for (int i = 0; i < ni; ++i) {
                const int t = axis == 0 ? i : (axis == 1 ? j : k);
                if (!((t > 0 && t < n_a) || (t == 0 && lo) || (t == n_a && hi))) continue;
                ++n_up;
                const long long f = axis == 0 ? XF(m, i, j, k) : (axis == 1 ? YF(m, i, j, k) : ZF(m, i, j, k));
                const int pi = axis == 0 ? t : i + pad, pj = axis == 1 ? t : j + pad, pk = axis == 2 ? t : k + P->zp;
                const size_t base = ((size_t)pk * P->PY + pj) * P->PX + pi;
                double* ul = UL + NV * f;
                double* ur = UR + NV * f;
                for (int q = 0; q < NV; ++q) {
                    const double* cq = P->W + q * P->plane + base;
                    if (five) {
                        ul[q] = weno5z(cq[0], cq[stride], cq[2 * stride], cq[3 * stride], cq[4 * stride], we);
                        ur[q] = weno5z(cq[5 * stride], cq[4 * stride], cq[3 * stride], cq[2 * stride], cq[stride], we);
                    } else {
                        ul[q] = weno3z(cq[0], cq[stride], cq[2 * stride], we);
                        ur[q] = weno3z(cq[3 * stride], cq[2 * stride], cq[stride], we);
                    }
                }
                int ok = cons_positive(g1, gp, ul) && cons_positive(g1, gp, ur);
                if (ok) {
                    double rlo = 1e300, rhi = -1e300, plo = 1e300, phi = -1e300;
                    for (int s = 0; s < ncell; ++s) {
                        double u[NV];
                        for (int q = 0; q < NV; ++q) u[q] = P->W[q * P->plane + base + s * stride];
                        const double pc = cons_pressure(g1, gp, u);
                        rlo = smin(rlo, u[0]);
                        rhi = smax(rhi, u[0]);
                        plo = smin(plo, pc);
                        phi = smax(phi, pc);
                    }
                    const double rw = rhi - rlo, pw = phi - plo;
                    rlo -= rw;
                    rhi += rw;
                    plo -= pw;
                    phi += pw;
                    const double* sts[2] = {ul, ur};
                    for (int h = 0; h < 2; ++h) {
                        const double pf = cons_pressure(g1, gp, sts[h]);
                        ok = ok && sts[h][0] >= rlo && sts[h][0] <= rhi && pf >= plo && pf <= phi;
                    }
                }
                if (ok) continue;
                ++n_fb;
                for (int q = 0; q < NV; ++q) {
                    ul[q] = P->W[q * P->plane + base + left * stride];
                    ur[q] = P->W[q * P->plane + base + (left + 1) * stride];
                }
            }
        }
    }
    if (g_q4_collect) {
        g_q4_stats[0] += n_up;
        g_q4_stats[1] += n_fb;
    }
}

/* fourth_order_fluxes (Schemes.cpp:1954-2134). Per face family a (faces normal to axis a), on the lattice
 * c = (i, j, k) of the family:
 *   good: a full stencil, both states positive, |pR - pL| <= shock_blend * min(pL, pR) (when shock_blend > 0);
 *   point values A = U - (0 + sum over transverse axes of (U[+1] - 2 U + U[-1])) / 24 where the face and its
 *     transverse neighbours are good and A is positive (pt); the neighbours wrap across a periodic axis;
 *   full: pt and every transverse neighbour pt;
 *   Fp = the Riemann flux of the point values (of the averages where there are none); its largest wave speed is
 *     the residual's;
 *   F = Fp + (0 + sum of (Fp[+1] - 2 Fp + Fp[-1])) / 24 where full, the flux of the averages where pt only, else Fp. */
static int fourth_order_fluxes(const fvg_config* c, const Mesh* m, const Riemann* rs, const double* UL, const double* UR, double* F,
                               double* S_out) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const int dims[3] = {nx, ny, nz};
    const int nax = m->is3d ? 3 : 2;
    const long long nf = m->nf;
    const double sb = rs->blend;
    const double g1 = rs->eos.gamma - 1.0, gp = rs->eos.stiff ? rs->eos.gamma * rs->eos.p_inf : 0.0;
    int periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = a < nax && c->bc[2 * a] == FVG_BC_PERIODIC && c->bc[2 * a + 1] == FVG_BC_PERIODIC;
    double* AL = (double*)malloc(sizeof(double) * NV * (size_t)nf);
    double* AR = (double*)malloc(sizeof(double) * NV * (size_t)nf);
    double* Fp = (double*)malloc(sizeof(double) * NV * (size_t)nf);
    unsigned char* good = (unsigned char*)calloc((size_t)nf, 1);
    unsigned char* pt = (unsigned char*)calloc((size_t)nf, 1);
    unsigned char* full = (unsigned char*)calloc((size_t)nf, 1);
    if (!AL || !AR || !Fp || !good || !pt || !full) {
        free(AL); free(AR); free(Fp); free(good); free(pt); free(full);
        return -2;
    }
    int has[6], wrap[3];
    halo_sides(c, m, c->reconstruction == FVG_RECON_WENO5 ? 3 : 2, has, wrap);
    const long long fb[3] = {0, m->nxf, m->nxf + m->nyf};
    /* pass 1: good */
    for (int a = 0; a < nax; ++a) {
        const int e0 = dims[0] + (a == 0), e1 = dims[1] + (a == 1), e2 = dims[2] + (a == 2);
        const int lo_a = has[2 * a] != 0, hi_a = has[2 * a + 1] != 0;
#pragma omp parallel for schedule(static) num_threads(NT)
        for (long long line = 0; line < (long long)e1 * e2; ++line) {
            int cc[3];
            cc[1] = (int)(line % e1);
            cc[2] = (int)(line / e1);
            for (int i = 0; i < e0; ++i) {
                cc[0] = i;
                const long long f = fb[a] + line * e0 + i;
                const int ca = cc[a], n_a = dims[a];
                if (!((ca > 0 && ca < n_a) || (ca == 0 && lo_a) || (ca == n_a && hi_a))) continue;
                const double* ul = UL + NV * f;
                const double* ur = UR + NV * f;
                if (!cons_positive(g1, gp, ul) || !cons_positive(g1, gp, ur)) continue;
                if (sb > 0.0) {
                    const double pL = cons_pressure(g1, gp, ul), pR = cons_pressure(g1, gp, ur);
                    if (fabs(pR - pL) > sb * smin(pL, pR)) continue;
                }
                good[f] = 1;
            }
        }
    }
    /* the transverse neighbour s = -1 / +1 of face f along t (-1: none) */
#define NBR(f, cc, t, s, st)                                                                                               \
    ((cc)[t] + (s) >= 0 && (cc)[t] + (s) < dims[t] ? (f) + (s) * (st)[t]                                               \
                                                   : (!periodic[t] ? -1 : (f) - (long long)(s) * (dims[t] - 1) * (st)[t]))
    /* pass 2: point values */
    for (int a = 0; a < nax; ++a) {
        const int e0 = dims[0] + (a == 0), e1 = dims[1] + (a == 1), e2 = dims[2] + (a == 2);
        const long long st[3] = {1, e0, (long long)e0 * e1};
#pragma omp parallel for schedule(static) num_threads(NT)
        for (long long line = 0; line < (long long)e1 * e2; ++line) {
            int cc[3];
            cc[1] = (int)(line % e1);
            cc[2] = (int)(line / e1);
            for (int i = 0; i < e0; ++i) {
                const long long f = fb[a] + line * e0 + i;
                memcpy(AL + NV * f, UL + NV * f, sizeof(double) * NV);
                memcpy(AR + NV * f, UR + NV * f, sizeof(double) * NV);
                if (!good[f]) continue;
                cc[0] = i;
                long long nb[2][2];
                int cnt = 0, ok = 1;
                for (int t = 0; t < nax && ok; ++t) {
                    if (t == a) continue;
                    nb[cnt][0] = NBR(f, cc, t, -1, st);
                    nb[cnt][1] = NBR(f, cc, t, 1, st);
                    ok = nb[cnt][0] >= 0 && nb[cnt][1] >= 0 && good[nb[cnt][0]] && good[nb[cnt][1]];
                    ++cnt;
                }
                if (!ok) continue;
                double pl[NV], pr[NV];
                for (int q = 0; q < NV; ++q) {
                    double dl = 0.0, dr = 0.0;
                    for (int u = 0; u < cnt; ++u) {
                        dl += UL[NV * nb[u][1] + q] - 2.0 * UL[NV * f + q] + UL[NV * nb[u][0] + q];
                        dr += UR[NV * nb[u][1] + q] - 2.0 * UR[NV * f + q] + UR[NV * nb[u][0] + q];
                    }
                    pl[q] = UL[NV * f + q] - dl / 24.0;
                    pr[q] = UR[NV * f + q] - dr / 24.0;
                }
                if (!cons_positive(g1, gp, pl) || !cons_positive(g1, gp, pr)) continue;
                memcpy(AL + NV * f, pl, sizeof pl);
                memcpy(AR + NV * f, pr, sizeof pr);
                pt[f] = 1;
            }
        }
    }
    /* pass 3: full */
    for (int a = 0; a < nax; ++a) {
        const int e0 = dims[0] + (a == 0), e1 = dims[1] + (a == 1), e2 = dims[2] + (a == 2);
        const long long st[3] = {1, e0, (long long)e0 * e1};
#pragma omp parallel for schedule(static) num_threads(NT)
        for (long long line = 0; line < (long long)e1 * e2; ++line) {
            int cc[3];
            cc[1] = (int)(line % e1);
            cc[2] = (int)(line / e1);
            for (int i = 0; i < e0; ++i) {
                const long long f = fb[a] + line * e0 + i;
                if (!pt[f]) continue;
                cc[0] = i;
                int ok = 1;
                for (int t = 0; t < nax && ok; ++t) {
                    if (t == a) continue;
                    const long long g0 = NBR(f, cc, t, -1, st), g1n = NBR(f, cc, t, 1, st);
                    ok = g0 >= 0 && g1n >= 0 && pt[g0] && pt[g1n];
                }
                full[f] = (unsigned char)ok;
            }
        }
    }
    /* the Riemann fluxes of the point values (compute_flux, Schemes.cpp:294-318) */
    double S = 0.0;
    const long long nxf = m->nxf, nxyf = m->nxf + m->nyf;
#pragma omp parallel num_threads(NT)
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (long long f = 0; f < nf; ++f) {
            double sm;
            face_flux(rs, AL + NV * f, AR + NV * f, f < nxf ? 0 : (f < nxyf ? 1 : 2), Fp + NV * f, &sm);
            if (sm > local) local = sm;
        }
#pragma omp critical(fvo_flux4_max)
        if (local > S) S = local;
    }
    /* pass 4: the corrected fluxes */
    for (int a = 0; a < nax; ++a) {
        const int e0 = dims[0] + (a == 0), e1 = dims[1] + (a == 1), e2 = dims[2] + (a == 2);
        const long long st[3] = {1, e0, (long long)e0 * e1};
#pragma omp parallel for schedule(static) num_threads(NT)
        for (long long line = 0; line < (long long)e1 * e2; ++line) {
            int cc[3];
            cc[1] = (int)(line % e1);
            cc[2] = (int)(line / e1);
            for (int i = 0; i < e0; ++i) {
                const long long f = fb[a] + line * e0 + i;
                cc[0] = i;
                if (full[f]) {
                    long long nb[2][2];
                    int cnt = 0;
                    for (int t = 0; t < nax; ++t) {
                        if (t == a) continue;
                        nb[cnt][0] = NBR(f, cc, t, -1, st);
                        nb[cnt][1] = NBR(f, cc, t, 1, st);
                        ++cnt;
                    }
                    for (int q = 0; q < NV; ++q) {
                        double d = 0.0;
                        for (int u = 0; u < cnt; ++u) d += Fp[NV * nb[u][1] + q] - 2.0 * Fp[NV * f + q] + Fp[NV * nb[u][0] + q];
                        F[NV * f + q] = Fp[NV * f + q] + d / 24.0;
                    }
                } else if (pt[f]) {
                    double sm;
                    face_flux(rs, UL + NV * f, UR + NV * f, a, F + NV * f, &sm);
                } else {
                    memcpy(F + NV * f, Fp + NV * f, sizeof(double) * NV);
                }
            }
        }
    }
#undef NBR
    if (g_q4_collect)
        for (long long f = 0; f < nf; ++f) {
            g_q4_stats[2] += good[f];
            g_q4_stats[3] += pt[f];
            g_q4_stats[4] += full[f];
            g_q4_stats[5] += pt[f] && !full[f];
            g_q4_stats[6] += good[f] && !pt[f];
        }
    free(AL); free(AR); free(Fp); free(good); free(pt); free(full);
    *S_out = S;
    return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* MUSCL-Hancock predictor (Schemes.cpp:2135-2199, 2907-2964, 2965-2988, 2989-3060)                  */
/* ------------------------------------------------------------------------------------------------ */
static void euler_flux_axis(const Eos* e, const double* U, int a, int three, double* F) {
    double W[NV];
    c2p(e, U, W, NULL);
    const double un = a == 0 ? W[1] : (a == 1 ? W[2] : W[4]), p = W[3];
    F[0] = U[0] * un;
    F[1] = U[1] * un + (a == 0 ? p : 0.0);
    F[2] = U[2] * un + (a == 1 ? p : 0.0);
    F[3] = (U[3] + p) * un;
    F[4] = three ? U[4] * un + (a == 2 ? p : 0.0) : U[4] * un;
}
static int admissible_state(const Eos* e, const double* U) {
    if (!(U[0] > kEosEps)) return 0;
    return 0.5 * (U[1] * U[1] + U[2] * U[2] + U[4] * U[4]) < U[0] * (U[3] - e->p_inf);
}
static long long hancock_predict(const fvg_config* c, const Mesh* m, const Eos* e, double half_dt, const int wrap[3], double* UL,
                                 double* UR) {
    const int nx = m->nx, ny = m->ny, nz = m->nz, three = m->is3d;
    const double hx = half_dt / m->dx, hy = half_dt / m->dy, hz = three ? half_dt / m->dz : 0.0;
    long long reverted = 0;
    (void)c;
#pragma omp parallel for schedule(static) num_threads(NT) reduction(+ : reverted)
    for (int kj = 0; kj < ny * nz; ++kj) {
        const int j = kj % ny, k = kj / ny;
        for (int i = 0; i < nx; ++i) {
            double* st[6];
            st[0] = UR + NV * XF(m, i, j, k);
            st[1] = UL + NV * XF(m, i + 1, j, k);
            st[2] = UR + NV * YF(m, i, j, k);
            st[3] = UL + NV * YF(m, i, j + 1, k);
            if (three) {
                st[4] = UR + NV * ZF(m, i, j, k);
                st[5] = UL + NV * ZF(m, i, j, k + 1);
            }
            double fl[6][NV];
            const int ns = three ? 6 : 4;
            for (int s = 0; s < ns; ++s) euler_flux_axis(e, st[s], s / 2, three, fl[s]);
            double d[NV];
            for (int q = 0; q < NV; ++q) {
                if (three)
                    d[q] = (hx * (fl[1][q] - fl[0][q]) + hy * (fl[3][q] - fl[2][q])) + hz * (fl[5][q] - fl[4][q]);
                else
                    d[q] = hx * (fl[1][q] - fl[0][q]) + hy * (fl[3][q] - fl[2][q]);
            }
            for (int s = 0; s < ns; ++s) {
                double n[NV];
                for (int q = 0; q < NV; ++q) n[q] = st[s][q] - d[q];
                if (admissible_state(e, n))
                    memcpy(st[s], n, sizeof n);
                else
                    ++reverted;
            }
        }
    }
    /* the two copies of a periodic seam face (Schemes.cpp:2934-2943, 3017-3034) */
    if (wrap[0])
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                memcpy(UL + NV * XF(m, 0, j, k), UL + NV * XF(m, nx, j, k), sizeof(double) * NV);
                memcpy(UR + NV * XF(m, nx, j, k), UR + NV * XF(m, 0, j, k), sizeof(double) * NV);
            }
    if (wrap[1])
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                memcpy(UL + NV * YF(m, i, 0, k), UL + NV * YF(m, i, ny, k), sizeof(double) * NV);
                memcpy(UR + NV * YF(m, i, ny, k), UR + NV * YF(m, i, 0, k), sizeof(double) * NV);
            }
    if (three && wrap[2])
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                memcpy(UL + NV * ZF(m, i, j, 0), UL + NV * ZF(m, i, j, nz), sizeof(double) * NV);
                memcpy(UR + NV * ZF(m, i, j, nz), UR + NV * ZF(m, i, j, 0), sizeof(double) * NV);
            }
    return reverted;
}
/* hancock_wall_ghosts (Schemes.cpp:2965-2988) with wall_side = wall / no-slip / moving wall (Simulation.cpp:921-930) */
static void hancock_wall_ghosts(const fvg_config* c, const Mesh* m, const Eos* e, double* UL, double* UR) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    int wall[6];
    for (int s = 0; s < 6; ++s) {
        const int bt = c->bc[s];
        wall[s] = bt == FVG_BC_WALL || bt == FVG_BC_NO_SLIP || bt == FVG_BC_MOVING_WALL;
    }
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            if (wall[0]) ghost_state(c, e, 0, UR + NV * XF(m, 0, j, k), UL + NV * XF(m, 0, j, k), 0.0);
            if (wall[1]) ghost_state(c, e, 1, UL + NV * XF(m, nx, j, k), UR + NV * XF(m, nx, j, k), 0.0);
        }
        for (int i = 0; i < nx; ++i) {
            if (wall[2]) ghost_state(c, e, 2, UR + NV * YF(m, i, 0, k), UL + NV * YF(m, i, 0, k), 0.0);
            if (wall[3]) ghost_state(c, e, 3, UL + NV * YF(m, i, ny, k), UR + NV * YF(m, i, ny, k), 0.0);
        }
    }
    if (m->is3d)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (wall[4]) ghost_state(c, e, 4, UR + NV * ZF(m, i, j, 0), UL + NV * ZF(m, i, j, 0), 0.0);
                if (wall[5]) ghost_state(c, e, 5, UL + NV * ZF(m, i, j, nz), UR + NV * ZF(m, i, j, nz), 0.0);
            }
}

/* ------------------------------------------------------------------------------------------------ */
/* Viscous padded state (Boundaries.cpp:379-700)                                                    */
/* ------------------------------------------------------------------------------------------------ */
static void match_kinetic_energy(double* g, const double* in) {
    g[3] = in[3] + 0.5 * (g[1] * g[1] + g[2] * g[2] + g[4] * g[4] - in[1] * in[1] - in[2] * in[2] - in[4] * in[4]) / in[0];
}
/* viscous_ghost_3d (Boundaries.cpp:253-269) */
static void viscous_ghost_3d(const fvg_config* c, int side, const double* cin, double* gh) {
    const double uw[3] = {c->u_in, c->v_in, 0.0};
    const int d = side_axis(side);
    for (int q = 0; q < NV; ++q) gh[q] = cin[q];
    switch (c->bc[side]) {
    case FVG_BC_NO_SLIP:
        for (int a = 0; a < 3; ++a) gh[kMom[a]] = -cin[kMom[a]];
        break;
    case FVG_BC_WALL: gh[kMom[d]] = -cin[kMom[d]]; break;
    case FVG_BC_MOVING_WALL:
        for (int t = 0; t < 3; ++t) gh[kMom[t]] = (t == d) ? -cin[kMom[t]] : 2.0 * cin[0] * uw[t] - cin[kMom[t]];
        if (c->boundary_order == 2) match_kinetic_energy(gh, cin);
        break;
    default: break;
    }
}
/* 3-D padded state: padded_3d + fix_padded_corners (Boundaries.cpp:606-700, 520-560) */
static double* padded_3d(const fvg_config* c, const Mesh* m, const double* U) {
    const int nx = m->nx, ny = m->ny, nz = m->nz, PX = nx + 2, PY = ny + 2;
    const size_t tot = (size_t)(nz + 2) * PY * PX;
    double* P = (double*)malloc(sizeof(double) * NV * tot);
    if (!P) return NULL;
#define AT3(ii, jj, kk) (P + NV * (((size_t)(kk) * PY + (jj)) * PX + (ii)))
#pragma omp parallel for schedule(static) num_threads(NT)
    for (int kk = 0; kk < nz + 2; ++kk) {
        const int k = clampi(kk - 1, nz);
        for (int jj = 0; jj < PY; ++jj) {
            const int j = clampi(jj - 1, ny);
            for (int ii = 0; ii < PX; ++ii) memcpy(AT3(ii, jj, kk), U + NV * CELL(m, clampi(ii - 1, nx), j, k), sizeof(double) * NV);
        }
    }
    static const int sides[6] = {FVG_BOTTOM, FVG_TOP, FVG_LEFT, FVG_RIGHT, FVG_FRONT, FVG_BACK};
    for (int s = 0; s < 6; ++s) {
        const int side = sides[s], bc = c->bc[side];
        if (bc != FVG_BC_NO_SLIP && bc != FVG_BC_WALL && bc != FVG_BC_MOVING_WALL && bc != FVG_BC_PERIODIC) continue;
        const int d = side_axis(side), hi = side_high(side);
        const int n_d = d == 0 ? nx : (d == 1 ? ny : nz);
        const int g = hi ? n_d + 1 : 0, in = hi ? n_d - 1 : 0, op = hi ? 0 : n_d - 1;
        const int na = d == 0 ? ny : nx, nb = d == 2 ? ny : nz;
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < na; ++a) {
                int i, j, k, ii, jj, kk, io, jo, ko;
                if (d == 0) {
                    i = in; j = a; k = b; ii = g; jj = j + 1; kk = k + 1; io = op; jo = j; ko = k;
                } else if (d == 1) {
                    i = a; j = in; k = b; ii = i + 1; jj = g; kk = k + 1; io = i; jo = op; ko = k;
                } else {
                    i = a; j = b; k = in; ii = i + 1; jj = j + 1; kk = g; io = i; jo = j; ko = op;
                }
                double* gh = AT3(ii, jj, kk);
                if (bc == FVG_BC_PERIODIC)
                    memcpy(gh, U + NV * CELL(m, io, jo, ko), sizeof(double) * NV);
                else
                    viscous_ghost_3d(c, side, U + NV * CELL(m, i, j, k), gh);
            }
    }
    if (c->boundary_order == 2) {
        const int n3[3] = {nx, ny, nz};
        for (int kk = 0; kk <= nz + 1; ++kk)
            for (int jj = 0; jj <= ny + 1; ++jj)
                for (int ii = 0; ii <= nx + 1; ++ii) {
                    int p[3] = {ii, jj, kk};
                    int ng = 0, ga[2] = {-1, -1};
                    for (int a = 0; a < 3; ++a)
                        if (p[a] == 0 || p[a] == n3[a] + 1) {
                            if (ng < 2) ga[ng] = a;
                            ++ng;
                        }
                    if (ng != 2) continue;
                    const int d = ga[0];
                    const int hi = p[d] != 0;
                    int q3[3] = {p[0], p[1], p[2]};
                    q3[d] = hi ? n3[d] : 1;
                    double cc[NV], g[NV];
                    memcpy(cc, AT3(q3[0], q3[1], q3[2]), sizeof cc);
                    const int side = 2 * d + (hi ? 1 : 0);
                    if (c->bc[side] == FVG_BC_PERIODIC) {
                        q3[d] = hi ? 1 : n3[d];
                        memcpy(g, AT3(q3[0], q3[1], q3[2]), sizeof g);
                    } else {
                        viscous_ghost_3d(c, side, cc, g);
                    }
                    memcpy(AT3(p[0], p[1], p[2]), g, sizeof g);
                }
    }
#undef AT3
    return P;
}

/* 2-D padded state: get_padded_U via get_padded_frame + fix_frame_corners (Boundaries.cpp:379-520, 562-604) */
static double* padded_2d(const fvg_config* c, const Mesh* m, const double* U) {
    const int nx = m->nx, ny = m->ny, W = nx + 2;
    double* P = (double*)malloc(sizeof(double) * NV * (size_t)(ny + 2) * W);
    if (!P) return NULL;
#define AT(jj, ii) (P + NV * ((size_t)(jj) * W + (ii)))
    for (int jj = 0; jj < ny + 2; ++jj)
        for (int ii = 0; ii < W; ++ii) memcpy(AT(jj, ii), U + NV * CELL(m, clampi(ii - 1, nx), clampi(jj - 1, ny), 0), sizeof(double) * NV);
    const double u_in = c->u_in, v_in = c->v_in;
    const int e2 = c->boundary_order == 2;
    /* bottom */
    switch (c->bc[FVG_BOTTOM]) {
    case FVG_BC_NO_SLIP:
        for (int i = 0; i < nx; ++i) { const double* in = U + NV * i; AT(0, i + 1)[1] = -in[1]; AT(0, i + 1)[2] = -in[2]; }
        break;
    case FVG_BC_MOVING_WALL:
        for (int i = 0; i < nx; ++i) {
            const double* in = U + NV * i;
            AT(0, i + 1)[1] = 2.0 * in[0] * u_in - in[1];
            AT(0, i + 1)[2] = -in[2];
            if (e2) match_kinetic_energy(AT(0, i + 1), in);
        }
        break;
    case FVG_BC_WALL:
        for (int i = 0; i < nx; ++i) AT(0, i + 1)[2] = -U[NV * i + 2];
        break;
    case FVG_BC_PERIODIC:
        for (int i = 0; i < nx; ++i) memcpy(AT(0, i + 1), U + NV * ((size_t)(ny - 1) * nx + i), sizeof(double) * NV);
        break;
    default: break;
    }
    /* top */
    switch (c->bc[FVG_TOP]) {
    case FVG_BC_NO_SLIP:
        for (int i = 0; i < nx; ++i) {
            const double* in = U + NV * ((size_t)(ny - 1) * nx + i);
            AT(ny + 1, i + 1)[1] = -in[1];
            AT(ny + 1, i + 1)[2] = -in[2];
        }
        break;
    case FVG_BC_WALL:
        for (int i = 0; i < nx; ++i) AT(ny + 1, i + 1)[2] = -U[NV * ((size_t)(ny - 1) * nx + i) + 2];
        break;
    case FVG_BC_PERIODIC:
        for (int i = 0; i < nx; ++i) memcpy(AT(ny + 1, i + 1), U + NV * i, sizeof(double) * NV);
        break;
    case FVG_BC_MOVING_WALL:
        for (int i = 0; i < nx; ++i) {
            const double* in = U + NV * ((size_t)(ny - 1) * nx + i);
            AT(ny + 1, i + 1)[1] = 2.0 * in[0] * u_in - in[1];
            AT(ny + 1, i + 1)[2] = -in[2];
            if (e2) match_kinetic_energy(AT(ny + 1, i + 1), in);
        }
        break;
    default: break;
    }
    /* left */
    switch (c->bc[FVG_LEFT]) {
    case FVG_BC_NO_SLIP:
        for (int j = 0; j < ny; ++j) { const double* in = U + NV * ((size_t)j * nx); AT(j + 1, 0)[1] = -in[1]; AT(j + 1, 0)[2] = -in[2]; }
        break;
    case FVG_BC_WALL:
        for (int j = 0; j < ny; ++j) AT(j + 1, 0)[1] = -U[NV * ((size_t)j * nx) + 1];
        break;
    case FVG_BC_PERIODIC:
        for (int j = 0; j < ny; ++j) memcpy(AT(j + 1, 0), U + NV * ((size_t)j * nx + nx - 1), sizeof(double) * NV);
        break;
    case FVG_BC_MOVING_WALL:
        for (int j = 0; j < ny; ++j) {
            const double* in = U + NV * ((size_t)j * nx);
            AT(j + 1, 0)[1] = -in[1];
            AT(j + 1, 0)[2] = 2.0 * in[0] * v_in - in[2];
            if (e2) match_kinetic_energy(AT(j + 1, 0), in);
        }
        break;
    default: break;
    }
    /* right */
    switch (c->bc[FVG_RIGHT]) {
    case FVG_BC_NO_SLIP:
        for (int j = 0; j < ny; ++j) {
            const double* in = U + NV * ((size_t)j * nx + nx - 1);
            AT(j + 1, nx + 1)[1] = -in[1];
            AT(j + 1, nx + 1)[2] = -in[2];
        }
        break;
    case FVG_BC_WALL:
        for (int j = 0; j < ny; ++j) AT(j + 1, nx + 1)[1] = -U[NV * ((size_t)j * nx + nx - 1) + 1];
        break;
    case FVG_BC_PERIODIC:
        for (int j = 0; j < ny; ++j) memcpy(AT(j + 1, nx + 1), U + NV * ((size_t)j * nx), sizeof(double) * NV);
        break;
    case FVG_BC_MOVING_WALL:
        for (int j = 0; j < ny; ++j) {
            const double* in = U + NV * ((size_t)j * nx + nx - 1);
            AT(j + 1, nx + 1)[1] = -in[1];
            AT(j + 1, nx + 1)[2] = 2.0 * in[0] * v_in - in[2];
            if (e2) match_kinetic_energy(AT(j + 1, nx + 1), in);
        }
        break;
    default: break;
    }
    if (e2) {
        /* fix_frame_corners without MPI (Boundaries.cpp:562-604): the corner takes the left / right ghost of the
         * entry next to it in the ghost row */
        for (int r = 0; r < 2; ++r) {
            const int jj = r == 0 ? 0 : ny + 1;
            double res[2][NV];
            for (int right = 0; right < 2; ++right) {
                double in[NV], g[NV];
                memcpy(in, AT(jj, right ? nx : 1), sizeof in);
                memcpy(g, in, sizeof g);
                switch (c->bc[right ? FVG_RIGHT : FVG_LEFT]) {
                case FVG_BC_NO_SLIP: g[1] = -in[1]; g[2] = -in[2]; break;
                case FVG_BC_WALL: g[1] = -in[1]; break;
                case FVG_BC_MOVING_WALL:
                    g[1] = -in[1];
                    g[2] = 2.0 * in[0] * c->v_in - in[2];
                    match_kinetic_energy(g, in);
                    break;
                case FVG_BC_PERIODIC: memcpy(g, AT(jj, right ? 1 : nx), sizeof g); break;
                default: break;
                }
                memcpy(res[right], g, sizeof g);
            }
            memcpy(AT(jj, 0), res[0], sizeof(double) * NV);
            memcpy(AT(jj, nx + 1), res[1], sizeof(double) * NV);
        }
    }
#undef AT
    return P;
}

/* ------------------------------------------------------------------------------------------------ */
/* Navier-Stokes viscous residual (Physics.cpp:44-221 2-D, 484-826 3-D; ViscousBound.hpp)           */
/* ------------------------------------------------------------------------------------------------ */
static const double MU_DIAG = 4.0 / 3.0, MU_OFF = 2.0 / 3.0;

static inline double nu_eff(double mu, double k_cond, double gamma, double rho) {
    return smax((4.0 / 3.0) * (mu / rho), (k_cond * (gamma - 1.0)) / rho) + 1e-12;
}

/* 2-D: the gradient version (viscous_residual), bit-identical to the row sweep the solver runs
 * (cpp/tests/test_core.cpp:1578-1667). Adds to R; returns the speed bound. */
static int viscous_2d(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, double* R, double* s_out) {
    const int nx = m->nx, ny = m->ny, n = nx * ny, PW = nx + 2;
    const double dx = m->dx, dy = m->dy, gamma = e->gamma, mu = c->mu;
    const double k_cond = mu * gamma / (c->pr * (gamma - 1.0));
    const int bface2 = c->boundary_order == 2;
    const int px = c->bc[0] == FVG_BC_PERIODIC && c->bc[1] == FVG_BC_PERIODIC;   /* CrossGhosts::periodic_x (Simulation.cpp:974-976) */
    const int py = c->bc[2] == FVG_BC_PERIODIC && c->bc[3] == FVG_BC_PERIODIC;
    double* P = padded_2d(c, m, U);
    if (!P) return -2;
    const size_t np = (size_t)(ny + 2) * PW;
    double* Wp = (double*)malloc(sizeof(double) * NV * np);
    double* scr = (double*)malloc(sizeof(double) * 7 * (size_t)n);
    double (*F)[NV] = (double (*)[NV])malloc(sizeof(double) * NV * ((size_t)(nx + 1) * ny + (size_t)nx * (ny + 1)));
    if (!Wp || !scr || !F) {
        free(P); free(Wp); free(scr); free(F);
        return -2;
    }
    for (size_t p = 0; p < np; ++p) c2p(e, P + NV * p, Wp + NV * p, NULL);
#define WPX(jj, ii) (Wp + NV * ((size_t)(jj) * PW + (ii)))
#define PP(jj, ii) (P + NV * ((size_t)(jj) * PW + (ii)))
    double *su = scr, *sv = su + n, *sT = sv + n, *dudx = sT + n, *dudy = dudx + n, *dvdx = dudy + n, *dvdy = dvdx + n;
#pragma omp parallel for schedule(static) num_threads(NT)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int cc = j * nx + i;
            const double* w = WPX(j + 1, i + 1);
            const double rho = w[0], u = w[1], v = w[2];
            su[cc] = u;
            sv[cc] = v;
            sT[cc] = w[3] / rho;
            double gx[3], gy[3];
            for (int q = 0; q < 3; ++q) {
                gx[q] = (PP(j + 1, i + 2)[q] - PP(j + 1, i)[q]) / (2.0 * dx);
                gy[q] = (PP(j + 2, i + 1)[q] - PP(j, i + 1)[q]) / (2.0 * dy);
            }
            const double drx = gx[0], dry = gy[0];
            dudx[cc] = (gx[1] - u * drx) / rho;
            dudy[cc] = (gy[1] - u * dry) / rho;
            dvdx[cc] = (gx[2] - v * drx) / rho;
            dvdy[cc] = (gy[2] - v * dry) / rho;
        }
    const long long nxf = (long long)(nx + 1) * ny;
    /* x faces */
#pragma omp parallel for schedule(static) num_threads(NT)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            const int lo = (i == 0), hi = (i == nx);
            const int cL = j * nx + (lo ? 0 : i - 1), cR = j * nx + (hi ? nx - 1 : i);
            const double* gL = WPX(j + 1, 0);
            const double* gR = WPX(j + 1, nx + 1);
            const double uL = lo ? gL[1] : su[cL], uR = hi ? gR[1] : su[cR];
            const double vL = lo ? gL[2] : sv[cL], vR = hi ? gR[2] : sv[cR];
            const double TL = lo ? gL[3] / gL[0] : sT[cL], TR = hi ? gR[3] / gR[0] : sT[cR];
            const double u_f = 0.5 * (uL + uR), v_f = 0.5 * (vL + vR);
            const double du_dx = (uR - uL) / dx, dv_dx = (vR - vL) / dx, dT_dx = (TR - TL) / dx;
            const int xL = (lo && px) ? j * nx + nx - 1 : cL, xR = (hi && px) ? j * nx : cR;
            double du_dy = 0.5 * (dudy[xL] + dudy[xR]);
            double dv_dy = 0.5 * (dvdy[xL] + dvdy[xR]);
            if (bface2 && ((lo && !px) || (hi && !px))) {
                const int ig = lo ? 0 : nx + 1, ic = lo ? 1 : nx;
                const double f2u = 0.5 * (WPX(j + 2, ig)[1] + WPX(j + 2, ic)[1]), f0u = 0.5 * (WPX(j, ig)[1] + WPX(j, ic)[1]);
                const double f2v = 0.5 * (WPX(j + 2, ig)[2] + WPX(j + 2, ic)[2]), f0v = 0.5 * (WPX(j, ig)[2] + WPX(j, ic)[2]);
                du_dy = (f2u - f0u) / (2.0 * dy);
                dv_dy = (f2v - f0v) / (2.0 * dy);
            }
            const double tau_xx = mu * (MU_DIAG * du_dx - MU_OFF * dv_dy);
            const double tau_xy = mu * (du_dy + dv_dx);
            const double q_x = -k_cond * dT_dx;
            double* o = F[(long long)j * (nx + 1) + i];
            o[0] = 0.0;
            o[1] = tau_xx * dy;
            o[2] = tau_xy * dy;
            o[3] = (u_f * tau_xx + v_f * tau_xy - q_x) * dy;
            o[4] = 0.0;
        }
    /* y faces */
#pragma omp parallel for schedule(static) num_threads(NT)
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int lo = (j == 0), hi = (j == ny);
            const int cB = (lo ? 0 : j - 1) * nx + i, cT = (hi ? ny - 1 : j) * nx + i;
            const double* gB = WPX(0, i + 1);
            const double* gT = WPX(ny + 1, i + 1);
            const double uB = lo ? gB[1] : su[cB], uT = hi ? gT[1] : su[cT];
            const double vB = lo ? gB[2] : sv[cB], vT = hi ? gT[2] : sv[cT];
            const double TB = lo ? gB[3] / gB[0] : sT[cB], TT = hi ? gT[3] / gT[0] : sT[cT];
            const double u_f = 0.5 * (uB + uT), v_f = 0.5 * (vB + vT);
            const double dv_dy = (vT - vB) / dy, du_dy = (uT - uB) / dy, dT_dy = (TT - TB) / dy;
            const int yB = (lo && py) ? (ny - 1) * nx + i : cB, yT = (hi && py) ? i : cT;
            double du_dx = 0.5 * (dudx[yB] + dudx[yT]);
            double dv_dx = 0.5 * (dvdx[yB] + dvdx[yT]);
            if (bface2 && ((lo && !py) || (hi && !py))) {
                const int jg = lo ? 0 : ny + 1, jc = lo ? 1 : ny;
                const double f2u = 0.5 * (WPX(jg, i + 2)[1] + WPX(jc, i + 2)[1]), f0u = 0.5 * (WPX(jg, i)[1] + WPX(jc, i)[1]);
                const double f2v = 0.5 * (WPX(jg, i + 2)[2] + WPX(jc, i + 2)[2]), f0v = 0.5 * (WPX(jg, i)[2] + WPX(jc, i)[2]);
                du_dx = (f2u - f0u) / (2.0 * dx);
                dv_dx = (f2v - f0v) / (2.0 * dx);
            }
            const double tau_yy = mu * (MU_DIAG * dv_dy - MU_OFF * du_dx);
            const double tau_yx = mu * (du_dy + dv_dx);
            const double q_y = -k_cond * dT_dy;
            double* o = F[nxf + (long long)j * nx + i];
            o[0] = 0.0;
            o[1] = tau_yx * dx;
            o[2] = tau_yy * dx;
            o[3] = (u_f * tau_yx + v_f * tau_yy - q_y) * dx;
            o[4] = 0.0;
        }
    const double inv_area = 1.0 / (dx * dy);
    const double L = smin(dx, dy), inv_h2 = 1.0 / (dx * dx) + 1.0 / (dy * dy);
    double s_all = 0.0;
#pragma omp parallel num_threads(NT)
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double* Fl = F[(long long)j * (nx + 1) + i];
                const double* Fr = F[(long long)j * (nx + 1) + i + 1];
                const double* Gb = F[nxf + (long long)j * nx + i];
                const double* Gt = F[nxf + (long long)(j + 1) * nx + i];
                double* r = R + NV * ((size_t)j * nx + i);
                for (int q = 0; q < NV; ++q) r[q] += ((0.0 + (Fr[q] - Fl[q])) + (Gt[q] - Gb[q])) * inv_area;
                const double* w = WPX(j + 1, i + 1);
                const double rho = w[0];
                const double ne = nu_eff(mu, k_cond, gamma, rho);
                const double s = smax(fabs(w[1]), fabs(w[2])) + sound_speed(e, w[3], rho) + 2.0 * L * inv_h2 * ne;
                local = smax(local, s);
            }
#pragma omp critical(fvo_visc2d)
        s_all = smax(s_all, local);
    }
#undef WPX
#undef PP
    free(P);
    free(Wp);
    free(scr);
    free(F);
    *s_out = s_all;
    return 0;
}

/* 3-D: viscous_residual_3d with the face-array passes (Physics.cpp:784-826), bit-identical to its plane sweep
 * (cpp/tests/test_core.cpp:1669-1777). Adds to R; returns the speed bound with L = dt_length. */
static int viscous_3d(const fvg_config* c, const Mesh* m, const Eos* e, const double* U, double length, double* R, double* s_out) {
    const int nx = m->nx, ny = m->ny, nz = m->nz, PX = nx + 2, PY = ny + 2;
    const long long n = m->nc;
    const double h[3] = {m->dx, m->dy, m->dz};
    const double gamma = e->gamma, mu = c->mu;
    const double k_cond = mu * gamma / (c->pr * (gamma - 1.0));
    const int bface2 = c->boundary_order == 2;
    int periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = c->bc[2 * a] == FVG_BC_PERIODIC && c->bc[2 * a + 1] == FVG_BC_PERIODIC;
    double* P = padded_3d(c, m, U);
    if (!P) return -2;
    const size_t np = (size_t)(nz + 2) * PY * PX;
    double* Wp = (double*)malloc(sizeof(double) * NV * np);
    double* scr = (double*)malloc(sizeof(double) * 13 * (size_t)n);
    double* Fx = (double*)malloc(sizeof(double) * NV * (size_t)m->nf);
    if (!Wp || !scr || !Fx) {
        free(P); free(Wp); free(scr); free(Fx);
        return -2;
    }
#pragma omp parallel for schedule(static) num_threads(NT)
    for (long long p = 0; p < (long long)np; ++p) c2p(e, P + NV * p, Wp + NV * p, NULL);
#define PADW(ii, jj, kk) (Wp + NV * (((size_t)(kk) * PY + (jj)) * PX + (ii)))
#define PADU(ii, jj, kk) (P + NV * (((size_t)(kk) * PY + (jj)) * PX + (ii)))
    double* vel[3] = {scr, scr + n, scr + 2 * n};
    double* sT = scr + 3 * n;
    double* G[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) G[a][b] = scr + (4 + 3 * a + b) * (size_t)n;
#pragma omp parallel for schedule(static) num_threads(NT)
    for (int kj = 0; kj < ny * nz; ++kj) {
        const int j = kj % ny, k = kj / ny;
        for (int i = 0; i < nx; ++i) {
            const long long cc = CELL(m, i, j, k);
            const double* w = PADW(i + 1, j + 1, k + 1);
            const double rho = w[0];
            sT[cc] = w[3] / rho;
            const double* nb[3][2] = {{PADU(i, j + 1, k + 1), PADU(i + 2, j + 1, k + 1)},
                                      {PADU(i + 1, j, k + 1), PADU(i + 1, j + 2, k + 1)},
                                      {PADU(i + 1, j + 1, k), PADU(i + 1, j + 1, k + 2)}};
            double grad[3][NV];
            static const int qs[4] = {0, 1, 2, 4};
            for (int b = 0; b < 3; ++b)
                for (int qi = 0; qi < 4; ++qi) {
                    const int q = qs[qi];
                    grad[b][q] = (nb[b][1][q] - nb[b][0][q]) / (2.0 * h[b]);
                }
            for (int a = 0; a < 3; ++a) {
                const double ua = w[kMom[a]];
                vel[a][cc] = ua;
                for (int b = 0; b < 3; ++b) G[a][b][cc] = (grad[b][kMom[a]] - ua * grad[b][0]) / rho;
            }
        }
    }
    /* faces of each axis, stored in the mesh face order */
    for (int d = 0; d < 3; ++d) {
        const int nj = ny + (d == 1), nk = nz + (d == 2), ni = nx + (d == 0);
        const int n_d = d == 0 ? nx : (d == 1 ? ny : nz);
        const double area = face_area(m, d);
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int kj = 0; kj < nj * nk; ++kj) {
            const int j = kj % nj, k = kj / nj;
            for (int i = 0; i < ni; ++i) {
                const int idx[3] = {i, j, k};
                const int t = idx[d];
                const int lo = (t == 0), hi = (t == n_d);
                int c3[3] = {i, j, k};
                c3[d] = lo ? 0 : t - 1;
                const long long cL = CELL(m, c3[0], c3[1], c3[2]);
                c3[d] = hi ? n_d - 1 : t;
                const long long cR = CELL(m, c3[0], c3[1], c3[2]);
                double vL[3], vR[3], TL, TR;
                if (lo) {
                    int p3[3] = {i + 1, j + 1, k + 1};
                    p3[d] = 0;
                    const double* g = PADW(p3[0], p3[1], p3[2]);
                    for (int a = 0; a < 3; ++a) vL[a] = g[kMom[a]];
                    TL = g[3] / g[0];
                } else {
                    for (int a = 0; a < 3; ++a) vL[a] = vel[a][cL];
                    TL = sT[cL];
                }
                if (hi) {
                    int p3[3] = {i + 1, j + 1, k + 1};
                    p3[d] = n_d + 1;
                    const double* g = PADW(p3[0], p3[1], p3[2]);
                    for (int a = 0; a < 3; ++a) vR[a] = g[kMom[a]];
                    TR = g[3] / g[0];
                } else {
                    for (int a = 0; a < 3; ++a) vR[a] = vel[a][cR];
                    TR = sT[cR];
                }
                c3[0] = i; c3[1] = j; c3[2] = k;
                c3[d] = n_d - 1;
                const long long xL = (lo && periodic[d]) ? CELL(m, c3[0], c3[1], c3[2]) : cL;
                c3[d] = 0;
                const long long xR = (hi && periodic[d]) ? CELL(m, c3[0], c3[1], c3[2]) : cR;
                const int phys = bface2 && !periodic[d] && (lo || hi);
#define CROSS(a, b)                                                                                                                \
    (phys ? ({                                                                                                                     \
        int pg[3] = {i + 1, j + 1, k + 1}, pc[3] = {i + 1, j + 1, k + 1};                                                          \
        pg[d] = lo ? 0 : n_d + 1;                                                                                                  \
        pc[d] = lo ? 1 : n_d;                                                                                                      \
        int g1[3] = {pg[0], pg[1], pg[2]}, c1[3] = {pc[0], pc[1], pc[2]}, g0[3] = {pg[0], pg[1], pg[2]}, c0[3] = {pc[0], pc[1], pc[2]}; \
        g1[b] += 1; c1[b] += 1; g0[b] -= 1; c0[b] -= 1;                                                                            \
        const double f1 = 0.5 * (PADW(g1[0], g1[1], g1[2])[kMom[a]] + PADW(c1[0], c1[1], c1[2])[kMom[a]]);                        \
        const double f0 = 0.5 * (PADW(g0[0], g0[1], g0[2])[kMom[a]] + PADW(c0[0], c0[1], c0[2])[kMom[a]]);                        \
        (f1 - f0) / (2.0 * h[b]);                                                                                                  \
    })                                                                                                                             \
          : 0.5 * (G[a][b][xL] + G[a][b][xR]))
                double Gn[3], vf[3];
                for (int a = 0; a < 3; ++a) {
                    Gn[a] = (vR[a] - vL[a]) / h[d];
                    vf[a] = 0.5 * (vL[a] + vR[a]);
                }
                const double dT = (TR - TL) / h[d];
                double Gc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                if (d == 0) {
                    Gc[0][1] = CROSS(0, 1);
                    Gc[1][1] = CROSS(1, 1);
                    Gc[2][2] = CROSS(2, 2);
                    Gc[0][2] = CROSS(0, 2);
                } else if (d == 1) {
                    Gc[0][0] = CROSS(0, 0);
                    Gc[1][0] = CROSS(1, 0);
                    Gc[2][2] = CROSS(2, 2);
                    Gc[1][2] = CROSS(1, 2);
                } else {
                    Gc[0][0] = CROSS(0, 0);
                    Gc[1][1] = CROSS(1, 1);
                    Gc[2][0] = CROSS(2, 0);
                    Gc[2][1] = CROSS(2, 1);
                }
#undef CROSS
                double tau[3];
                for (int a = 0; a < 3; ++a) {
                    if (a == d) {
                        double others = 0.0;
                        for (int o = 0; o < 3; ++o)
                            if (o != d) others += Gc[o][o];
                        tau[a] = mu * (MU_DIAG * Gn[d] - MU_OFF * others);
                    } else {
                        tau[a] = mu * (Gn[a] + Gc[d][a]);
                    }
                }
                const double q_d = -k_cond * dT;
                const long long f = d == 0 ? XF(m, i, j, k) : (d == 1 ? YF(m, i, j, k) : ZF(m, i, j, k));
                double* o = Fx + NV * f;
                o[0] = 0.0;
                o[1] = tau[0] * area;
                o[2] = tau[1] * area;
                o[3] = (vf[0] * tau[0] + vf[1] * tau[1] + vf[2] * tau[2] - q_d) * area;
                o[4] = tau[2] * area;
            }
        }
    }
    const double inv_vol = 1.0 / cell_volume(m);
    const double inv_h2 = 1.0 / (m->dx * m->dx) + 1.0 / (m->dy * m->dy) + 1.0 / (m->dz * m->dz);
    const double L = length > 0.0 ? length : 2.0 / (1.0 / m->dx + 1.0 / m->dy + 1.0 / m->dz);
    double s_all = 0.0;
#pragma omp parallel num_threads(NT)
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (int kj = 0; kj < ny * nz; ++kj) {
            const int j = kj % ny, k = kj / ny;
            for (int i = 0; i < nx; ++i) {
                const double* fl = Fx + NV * XF(m, i, j, k);
                const double* fr = Fx + NV * XF(m, i + 1, j, k);
                const double* fb = Fx + NV * YF(m, i, j, k);
                const double* ft = Fx + NV * YF(m, i, j + 1, k);
                const double* ff = Fx + NV * ZF(m, i, j, k);
                const double* fk = Fx + NV * ZF(m, i, j, k + 1);
                double* r = R + NV * CELL(m, i, j, k);
                for (int q = 0; q < NV; ++q) r[q] += (((0.0 + (fr[q] - fl[q])) + (ft[q] - fb[q])) + (fk[q] - ff[q])) * inv_vol;
                const double* w = PADW(i + 1, j + 1, k + 1);
                const double rho = w[0];
                const double ne = nu_eff(mu, k_cond, gamma, rho);
                double mv = fabs(w[1]);
                if (mv < fabs(w[2])) mv = fabs(w[2]);
                if (mv < fabs(w[4])) mv = fabs(w[4]);
                const double s = mv + sound_speed(e, w[3], rho) + 2.0 * L * inv_h2 * ne;
                local = smax(local, s);
            }
        }
#pragma omp critical(fvo_visc3d)
        s_all = smax(s_all, local);
    }
#undef PADW
#undef PADU
    free(P);
    free(Wp);
    free(scr);
    free(Fx);
    *s_out = s_all;
    return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* Residual (Simulation::assemble / assemble_3d, Simulation.cpp:811-1323)                           */
/* ------------------------------------------------------------------------------------------------ */
int fvo_residual(const fvg_config* c, const double* U, double* R, double half_dt, double* smax_out, long long* reverted) {
    const Mesh m = make_mesh(c);
    const Riemann rs = make_riemann(c);
    const Eos* e = &rs.eos;
    const int recon = c->reconstruction != FVG_RECON_NONE;
    const int fourth = c->flux_quadrature == FVG_QUAD_FOURTH;
    if (fourth && c->reconstruction != FVG_RECON_WENO3 && c->reconstruction != FVG_RECON_WENO5) return -3;
    double* UL = (double*)malloc(sizeof(double) * NV * (size_t)m.nf);
    double* UR = (double*)malloc(sizeof(double) * NV * (size_t)m.nf);
    double* F = (double*)malloc(sizeof(double) * NV * (size_t)m.nf);
    if (!UL || !UR || !F) {
        free(UL); free(UR); free(F);
        return -2;
    }
    double ut[6];
    massflow_targets(c, &m, e, U, ut);
    int rc = apply_boundary_conditions(c, &m, e, U, UL, UR, !recon, ut);
    long long rev = 0;
    if (rc == 0 && recon) {
        const int Dh = c->reconstruction == FVG_RECON_WENO5 ? 3 : 2;
        int has[6], wrap[3];
        halo_sides(c, &m, Dh, has, wrap);
        Padded P;
        rc = pad_primitives(c, &m, e, U, Dh, has, wrap, &P, !fourth, ut);
        if (rc == 0) {
            if (fourth) reconstruct_conserved(c, &m, e, &P, has, UL, UR);
            else reconstruct_faces(c, &m, e, &P, has, UL, UR);
        }
        free(P.W);
        if (rc == 0 && half_dt > 0.0) {
            rev = hancock_predict(c, &m, e, half_dt, wrap, UL, UR);
            hancock_wall_ghosts(c, &m, e, UL, UR);
        }
    }
    if (rc != 0) {
        free(UL); free(UR); free(F);
        return rc;
    }
    /* RiemannSolver::compute_flux (Schemes.cpp:294-318): every face; the largest wave speed, NaN dropped */
    double S = 0.0;
    const long long nxf = m.nxf, nxyf = m.nxf + m.nyf;
    if (fourth) {
        if (fourth_order_fluxes(c, &m, &rs, UL, UR, F, &S) != 0) {
            free(UL); free(UR); free(F);
            return -2;
        }
    } else
#pragma omp parallel num_threads(NT)
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (long long f = 0; f < m.nf; ++f) {
            double sm;
            face_flux(&rs, UL + NV * f, UR + NV * f, f < nxf ? 0 : (f < nxyf ? 1 : 2), F + NV * f, &sm);
            if (sm > local) local = sm;
        }
#pragma omp critical(fvo_flux_max)
        if (local > S) S = local;
    }
    /* divergence (Simulation.cpp:1022-1037 2-D, 1293-1310 3-D) */
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    if (!m.is3d) {
        const double inv_area = 1.0 / (m.dx * m.dy), dxl = m.dx, dyl = m.dy;
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double* fl = F + NV * XF(&m, i, j, 0);
                const double* fr = F + NV * XF(&m, i + 1, j, 0);
                const double* fb = F + NV * YF(&m, i, j, 0);
                const double* ft = F + NV * YF(&m, i, j + 1, 0);
                double* r = R + NV * CELL(&m, i, j, 0);
                for (int q = 0; q < NV; ++q) r[q] = ((0.0 - (fr[q] * dyl - fl[q] * dyl)) - (ft[q] * dxl - fb[q] * dxl)) * inv_area;
            }
    } else {
        const double Ax = face_area(&m, 0), Ay = face_area(&m, 1), Az = face_area(&m, 2), inv_vol = 1.0 / cell_volume(&m);
#pragma omp parallel for schedule(static) num_threads(NT)
        for (int kj = 0; kj < ny * nz; ++kj) {
            const int j = kj % ny, k = kj / ny;
            for (int i = 0; i < nx; ++i) {
                const double* fl = F + NV * XF(&m, i, j, k);
                const double* fr = F + NV * XF(&m, i + 1, j, k);
                const double* fb = F + NV * YF(&m, i, j, k);
                const double* ft = F + NV * YF(&m, i, j + 1, k);
                const double* ff = F + NV * ZF(&m, i, j, k);
                const double* fk = F + NV * ZF(&m, i, j, k + 1);
                double* r = R + NV * CELL(&m, i, j, k);
                for (int q = 0; q < NV; ++q)
                    r[q] = (((0.0 - (fr[q] * Ax - fl[q] * Ax)) - (ft[q] * Ay - fb[q] * Ay)) - (fk[q] * Az - ff[q] * Az)) * inv_vol;
            }
        }
    }
    free(UL);
    free(UR);
    free(F);
    if (c->use_viscous) {
        double sv = 0.0;
        rc = m.is3d ? viscous_3d(c, &m, e, U, fvo_dt_length(c), R, &sv) : viscous_2d(c, &m, e, U, R, &sv);
        if (rc != 0) return rc;
        S = smax(S, sv);
    }
    if (reverted) *reverted = rev;
    *smax_out = S;
    return 0;
}

int fvo_fourth_stats(const fvg_config* c, const double* U, long long out[7]) {
    const Mesh m = make_mesh(c);
    double* R = (double*)malloc(sizeof(double) * NV * (size_t)m.nc);
    if (!R) return -2;
    memset(g_q4_stats, 0, sizeof g_q4_stats);
    g_q4_collect = 1;
    double S;
    const int rc = fvo_residual(c, U, R, 0.0, &S, NULL);
    g_q4_collect = 0;
    free(R);
    memcpy(out, g_q4_stats, sizeof g_q4_stats);
    return rc;
}

/* ------------------------------------------------------------------------------------------------ */
/* Time schemes (TimeScheme.cpp:110-319) and the divergence check (SimulationInternal.hpp:259-267)   */
/* ------------------------------------------------------------------------------------------------ */
static double clamp_dt(double dt, double t, double t_max, double dt_max) {
    dt = smin(dt, t_max - t);
    if (dt_max > 0.0) dt = smin(dt, dt_max);
    return dt;
}

int fvo_step(const fvg_config* c, double* U, double t, double t_max, double dt_max, double* lagged_s, double* dt_out,
             double* smax_out) {
    const Mesh m = make_mesh(c);
    const size_t n = (size_t)m.nc * NV;
    const double L = fvo_dt_length(c);
    double* R = (double*)malloc(sizeof(double) * n);
    double* tmp = (double*)malloc(sizeof(double) * n);
    if (!R || !tmp) {
        free(R); free(tmp);
        return -2;
    }
    double dt = 0.0, S = 0.0;
    int rc = 0;
    if (c->time_scheme == FVG_TS_HANCOCK) {
        double sp = lagged_s ? *lagged_s : 0.0;
        if (!(sp > 0.0)) rc = fvo_residual(c, U, R, 0.0, &sp, NULL);
        if (rc == 0) {
            dt = clamp_dt(c->cfl * L / (sp + kDtEps), t, t_max, dt_max);
            rc = fvo_residual(c, U, R, 0.5 * dt, &S, NULL);
        }
        if (rc == 0) {
#pragma omp parallel for schedule(static) num_threads(NT)
            for (long long x = 0; x < (long long)n; ++x) U[x] = U[x] + dt * R[x];
            if (lagged_s) *lagged_s = S;
        }
    } else {
        double s1 = 0.0;
        rc = fvo_residual(c, U, R, 0.0, &s1, NULL);
        if (rc == 0) {
            dt = clamp_dt(c->cfl * L / (s1 + kDtEps), t, t_max, dt_max);
            S = s1;
            if (c->time_scheme == FVG_TS_EULER) {
#pragma omp parallel for schedule(static) num_threads(NT)
                for (long long x = 0; x < (long long)n; ++x) U[x] = U[x] + dt * R[x];
            } else {
#pragma omp parallel for schedule(static) num_threads(NT)
                for (long long x = 0; x < (long long)n; ++x) tmp[x] = U[x] + dt * R[x];
                double s2 = 0.0;
                rc = fvo_residual(c, tmp, R, 0.0, &s2, NULL);
                if (rc == 0 && c->time_scheme == FVG_TS_SSPRK2) {
#pragma omp parallel for schedule(static) num_threads(NT)
                    for (long long x = 0; x < (long long)n; ++x) U[x] = 0.5 * U[x] + 0.5 * (tmp[x] + dt * R[x]);
                    S = smax(s1, s2);
                } else if (rc == 0) {
#pragma omp parallel for schedule(static) num_threads(NT)
                    for (long long x = 0; x < (long long)n; ++x) tmp[x] = 0.75 * U[x] + 0.25 * (tmp[x] + dt * R[x]);
                    double s3 = 0.0;
                    rc = fvo_residual(c, tmp, R, 0.0, &s3, NULL);
                    if (rc == 0) {
#pragma omp parallel for schedule(static) num_threads(NT)
                        for (long long x = 0; x < (long long)n; ++x) U[x] = 1.0 / 3.0 * U[x] + 2.0 / 3.0 * (tmp[x] + dt * R[x]);
                        S = smax(s1, smax(s2, s3));
                    }
                }
            }
        }
    }
    free(R);
    free(tmp);
    if (rc != 0) return rc;
    *dt_out = dt;
    *smax_out = S;
    if (isfinite(S) && isfinite(dt) && t + dt > t) return 0;
    return -1;
}

void fvo_volume_averages(const fvg_config* c, const double* U, double out[3]) {
    const Mesh m = make_mesh(c);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (long long x = 0; x < m.nc; ++x) {
        const double* u = U + NV * x;
        s0 += u[0];
        s1 += 0.5 * (u[1] * u[1] + u[2] * u[2] + u[4] * u[4]) / u[0];
        s2 += u[3];
    }
    out[0] = s0 / (double)m.nc;
    out[1] = s1 / (double)m.nc;
    out[2] = s2 / (double)m.nc;
}
