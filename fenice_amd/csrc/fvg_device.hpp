// Device arithmetic of the uniform-grid path: equation of state, HLLC / low-Mach HLLC with the shock blend,
// TVD limiters, WENO3-Z / WENO5-Z, convective and viscous ghost states.
//
// Bit-for-bit rule: every expression keeps the operation order of the reference, the file is compiled with
// -ffp-contract=off (no multiply-add is fused) and min / max are the reference's compare-and-select forms
// (std::min(a, b) = (b < a) ? b : a, std::max(a, b) = (a < b) ? b : a). Division and square root are the
// IEEE-rounded ones of the compiler's default lowering. With these, FP64 results are the reference's bits.
//
// NV is the number of conserved variables carried: 5 in 3-D, 4 in 2-D. A 2-D field has rho*w = +0
// everywhere in the reference, and with w = 0 the first four components do not depend on it (w*w adds +0 to a
// non-negative sum), so the 2-D kernels drop the fifth plane and its arithmetic.
#pragma once

#include <hip/hip_runtime.h>

namespace fvg {

constexpr double kEosEps = 1e-10;    // Core.hpp:169
constexpr double kDenomEps = 1e-14;  // Schemes.cpp:63
constexpr double kMachEps = 1e-12;   // Schemes.cpp:65
constexpr double kDtEps = 1e-10;     // TimeScheme.cpp:20

// index of the momentum component along axis a in [rho, rho u, rho v, rho E, rho w]
__host__ __device__ constexpr int mom(int a) { return a == 0 ? 1 : (a == 1 ? 2 : 4); }

template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

// Division and square root of the flux arithmetic. FP64: the IEEE-rounded operations (the reference's bits).
// FP32 (flux_precision = fp32): the hardware reciprocal and square root (about 1 ulp), two instructions instead of
// the ten of a correctly rounded sequence; the error of the FP32 mode is measured against the FP64 path.
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sq(double a) { return sqrt(a); }
__device__ __forceinline__ float sq(float a) { return __builtin_amdgcn_sqrtf(a); }

// x / 6 correctly rounded in three operations, for x = 0 or 2^-960 <= |x| <= 2^905 (Markstein's correction step with
// a = RN(1/6) = (1/6)(1 - 2^-54)): x a is within ulp(x/6)/2 of x/6, so q0 = RN(x a) is within 1 ulp, r = 6 q0 - x is
// an integer multiple of 2 ulp(x/6) smaller than 3 of them and exact in one FMA, and q0 - r a equals
// x/6 + (q0 - x/6) 2^-54 before its one rounding. x/6 = (M/3) 2^e for an integer significand M is either
// representable or M/3 leaves a remainder of 1/3 or 2/3, which puts x/6 at least 1/6 ulp from every midpoint; an
// error below 2^-54 ulp cannot cross one, so the result is the IEEE quotient. Zeros keep their sign (r = +0, and
// -r a + q0 = q0). Below the range the quotient may be subnormal and a midpoint (x = 9 2^-1074 gives 2^-1074, the
// division 2 2^-1074); above it, or for non-finite x, the FMA can overflow or give NaN: callers keep the IEEE
// division there (range_fits_quot6).
__host__ __device__ __forceinline__ double quot6(double x) {
    constexpr double a = 1.0 / 6.0;
    const double q0 = x * a;
    const double r = __builtin_fma(q0, 6.0, -x);
    return __builtin_fma(-r, a, q0);
}
// v is 0 or 2^-900 <= |v| <= 2^900: the numerators of WENO5-Z's candidates (sums of 2, 7 and 11 times such values)
// are then 0 or between 2^-952 and 2^905 in magnitude, inside the range of quot6. NaN is outside.
__host__ __device__ __forceinline__ bool range_fits_quot6(double v) {
    const double m = v < 0.0 ? -v : v;
    return (v == 0.0) | ((m >= 0x1p-900) & (m <= 0x1p+900));   // no short-circuit branches
}
// the absolute floor of the WENO regularisation: 1e-300 in FP64 (Schemes.cpp:331, 343); the smallest power of ten
// above FLT_MIN in FP32, where 1e-300 would round to zero and 0/0 would reach the weights of an all-zero stencil
template <typename T> __device__ __forceinline__ T weno_floor() { return T(1e-300); }
template <> __device__ __forceinline__ float weno_floor<float>() { return 1e-37f; }

struct Eos {
    int stiff;
    double gamma, p_inf;
};

// primitive [rho, u, v, p, w] -> conservative (Core.hpp:209-215)
template <int NV, typename T>
__device__ __forceinline__ void p2c(const Eos& e, const T* W, T* U) {
    const T g = T(e.gamma);
    const T rho = W[0], u = W[1], v = W[2], p = W[3];
    T ke;
    if constexpr (NV == 5) {
        const T w = W[4];
        ke = T(0.5) * rho * (u * u + v * v + w * w);
        U[4] = rho * w;
    } else {
        ke = T(0.5) * rho * (u * u + v * v);
    }
    const T en = !e.stiff ? dv(p, g - T(1.0)) : dv(p + g * T(e.p_inf), g - T(1.0));
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = en + ke;
}

// conservative -> primitive with the floors of Core.hpp:219-228; *inv = 1 / rho
template <int NV, typename T>
__device__ __forceinline__ void c2p(const Eos& e, const T* U, T* W, T* inv_out) {
    const T g = T(e.gamma);
    const T rho = smax(U[0], T(kEosEps));
    const T inv = dv(T(1.0), rho);
    if (inv_out) *inv_out = inv;
    const T u = U[1] * inv, v = U[2] * inv;
    T e_kin;
    if constexpr (NV == 5) {
        const T w = U[4] * inv;
        e_kin = T(0.5) * rho * (u * u + v * v + w * w);
        W[4] = w;
    } else {
        e_kin = T(0.5) * rho * (u * u + v * v);
    }
    const T rho_e = U[3] - e_kin;
    const T p = !e.stiff ? (g - T(1.0)) * rho_e : (g - T(1.0)) * rho_e - g * T(e.p_inf);
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = smax(p, T(kEosEps));
}

template <typename T>
__device__ __forceinline__ T sound_speed(const Eos& e, T p, T rho) {
    if (!e.stiff) return sq(dv(T(e.gamma) * smax(p, T(kEosEps)), smax(rho, T(kEosEps))));
    return sq(dv(T(e.gamma) * (p + T(e.p_inf)), smax(rho, T(kEosEps))));
}
template <typename T>
__device__ __forceinline__ T sound_speed_inv(const Eos& e, T p, T inv_rho) {
    if (!e.stiff) return sq(T(e.gamma) * smax(p, T(kEosEps)) * inv_rho);
    return sq(T(e.gamma) * (p + T(e.p_inf)) * inv_rho);
}

template <int NV, typename T>
__device__ __forceinline__ void roundtrip(const Eos& e, T* U) {
    T W[NV];
    c2p<NV>(e, U, W, (T*)nullptr);
    p2c<NV>(e, W, U);
}

template <typename T> __device__ __forceinline__ T safe_denom(T x) { return copysign(smax(fabs(x), T(kDenomEps)), x); }

struct RiemannParams {
    Eos eos;
    int lowmach;
    double blend;
};

// HLLC with Einfeldt bounds, the contact speed of Toro (10.37) and the shock blend toward HLL
// (Schemes.cpp:116-200). vel index a = 0, 1, 2 maps to W[1], W[2], W[4].
template <int NV, typename T>
__device__ __forceinline__ void hllc_core(const RiemannParams& rs, const T* WL, const T* WR, const T* UL, const T* UR, T inv_L,
                                          T inv_R, T c_L, T c_R, int axis, T* F, T& smax_out) {
    const T rho_L = WL[0], p_L = WL[3];
    const T rho_R = WR[0], p_R = WR[3];
    const T E_L = UL[3] * inv_L, E_R = UR[3] * inv_R;
    const T Vn_L = WL[mom(axis)], Vn_R = WR[mom(axis)];
    const int mn = mom(axis);
    T FL[NV], FR[NV];
    FL[0] = rho_L * Vn_L;
    FR[0] = rho_R * Vn_R;
#pragma unroll
    for (int a = 0; a < NV - 2; ++a) {
        FL[mom(a)] = rho_L * WL[mom(a)] * Vn_L;
        FR[mom(a)] = rho_R * WR[mom(a)] * Vn_R;
    }
    FL[mn] = FL[mn] + p_L;
    FR[mn] = FR[mn] + p_R;
    FL[3] = Vn_L * (rho_L * E_L + p_L);
    FR[3] = Vn_R * (rho_R * E_R + p_R);
    const T S_L = smin(Vn_L - c_L, Vn_R - c_R);
    const T S_R = smax(Vn_L + c_L, Vn_R + c_R);
    const T num = rho_R * Vn_R * (S_R - Vn_R) - rho_L * Vn_L * (S_L - Vn_L) + p_L - p_R;
    const T den = rho_R * (S_R - Vn_R) - rho_L * (S_L - Vn_L);
    const T S_M = dv(num, safe_denom(den));
    const T den_star_L = safe_denom(S_L - S_M);
    const T den_star_R = safe_denom(S_R - S_M);
    const T den_pL = safe_denom(S_L - Vn_L);
    const T den_pR = safe_denom(S_R - Vn_R);
    const T fL = dv(rho_L * (S_L - Vn_L), den_star_L);
    const T fR = dv(rho_R * (S_R - Vn_R), den_star_R);
    T UsL[NV], UsR[NV];
    UsL[0] = fL;
    UsR[0] = fR;
#pragma unroll
    for (int a = 0; a < NV - 2; ++a) {
        UsL[mom(a)] = fL * (a == axis ? S_M : WL[mom(a)]);
        UsR[mom(a)] = fR * (a == axis ? S_M : WR[mom(a)]);
    }
    UsL[3] = fL * (E_L + (S_M - Vn_L) * (S_M + dv(p_L, rho_L * den_pL)));
    UsR[3] = fR * (E_R + (S_M - Vn_R) * (S_M + dv(p_R, rho_R * den_pR)));
    T Fo[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) Fo[q] = T(0);
    if (S_L >= T(0))
#pragma unroll
        for (int q = 0; q < NV; ++q) Fo[q] = FL[q];
    if (S_L < T(0) && S_M >= T(0))
#pragma unroll
        for (int q = 0; q < NV; ++q) Fo[q] = FL[q] + S_L * (UsL[q] - UL[q]);
    if (S_M < T(0) && S_R >= T(0))
#pragma unroll
        for (int q = 0; q < NV; ++q) Fo[q] = FR[q] + S_R * (UsR[q] - UR[q]);
    if (S_R < T(0))
#pragma unroll
        for (int q = 0; q < NV; ++q) Fo[q] = FR[q];
    if (rs.blend > 0.0) {
        const T blend = T(rs.blend);
        const T s = dv(fabs(p_R - p_L), smin(p_L, p_R));
        if (s > blend) {
            const T phi = smin(dv(s - blend, blend), T(1.0));
            T Fh[NV];
            if (S_L >= T(0)) {
#pragma unroll
                for (int q = 0; q < NV; ++q) Fh[q] = FL[q];
            } else if (S_R <= T(0)) {
#pragma unroll
                for (int q = 0; q < NV; ++q) Fh[q] = FR[q];
            } else {
                const T inv = dv(T(1.0), S_R - S_L);
#pragma unroll
                for (int q = 0; q < NV; ++q) Fh[q] = (S_R * FL[q] - S_L * FR[q] + S_L * S_R * (UR[q] - UL[q])) * inv;
            }
#pragma unroll
            for (int q = 0; q < NV; ++q) Fo[q] = (T(1.0) - phi) * Fo[q] + phi * Fh[q];
        }
    }
    smax_out = smax(fabs(S_L), fabs(S_R));
#pragma unroll
    for (int q = 0; q < NV; ++q) F[q] = Fo[q];
}

// RiemannSolver::face_flux (Schemes.cpp:108-113, 202-229)
template <int NV, typename T>
__device__ __forceinline__ void face_flux(const RiemannParams& rs, const T* UL, const T* UR, int axis, T* F, T& sm) {
    T WL[NV], WR[NV], inv_L, inv_R;
    c2p<NV>(rs.eos, UL, WL, &inv_L);
    c2p<NV>(rs.eos, UR, WR, &inv_R);
    if (!rs.lowmach) {
        hllc_core<NV>(rs, WL, WR, UL, UR, inv_L, inv_R, sound_speed_inv(rs.eos, WL[3], inv_L), sound_speed_inv(rs.eos, WR[3], inv_R),
                      axis, F, sm);
        return;
    }
    const T c_L = sound_speed_inv(rs.eos, WL[3], inv_L);
    const T c_R = sound_speed_inv(rs.eos, WR[3], inv_R);
    T ML2, MR2;
    if constexpr (NV == 5) {
        ML2 = WL[1] * WL[1] + WL[2] * WL[2] + WL[4] * WL[4];
        MR2 = WR[1] * WR[1] + WR[2] * WR[2] + WR[4] * WR[4];
    } else {
        ML2 = WL[1] * WL[1] + WL[2] * WL[2];
        MR2 = WR[1] * WR[1] + WR[2] * WR[2];
    }
    const T M_L = dv(sq(ML2), c_L + T(kMachEps));
    const T M_R = dv(sq(MR2), c_R + T(kMachEps));
    const T z = smin(smax(smax(M_L, M_R), T(1e-4)), T(1.0));
#pragma unroll
    for (int a = 0; a < NV - 2; ++a) {
        const int m = mom(a);
        const T avg = T(0.5) * (WL[m] + WR[m]);
        const T half = T(0.5) * (WL[m] - WR[m]);
        WL[m] = avg + z * half;
        WR[m] = avg - z * half;
    }
    T UsL[NV], UsR[NV];
    p2c<NV>(rs.eos, WL, UsL);
    p2c<NV>(rs.eos, WR, UsR);
    hllc_core<NV>(rs, WL, WR, UsL, UsR, inv_L, inv_R, c_L, c_R, axis, F, sm);
}

// ---- limiters (Schemes.hpp:90-117) and WENO-Z (Schemes.cpp:323-353) ----
template <typename T> __device__ __forceinline__ T sgn(T x) { return (x > T(0)) ? T(1) : ((x < T(0)) ? T(-1) : T(0)); }

template <typename T>
__device__ __forceinline__ T limiter(int l, T a, T b) {
    switch (l) {
    case 0: return T(0.5) * (sgn(a) + sgn(b)) * smin(fabs(a), fabs(b));
    case 1: {
        const T ab = a * b;
        return (ab > T(0)) ? dv(T(2.0) * ab, a + b) : T(0.0);
    }
    case 2:
        if (!(a * b > T(0))) return T(0.0);
        return smin(smin(T(2.0) * fabs(a), T(0.5) * fabs(a + b)), T(2.0) * fabs(b)) * sgn(a);
    default:
        if (!(a * b > T(0))) return T(0.0);
        return fmin(fmin(T(2.0) * fabs(a), dv(fabs(T(2.0) * a + b), T(3.0))), T(2.0) * fabs(b)) * sgn(a);
    }
}

template <typename T>
__device__ __forceinline__ T weno3z(T c0, T c1, T c2, T eps_rel) {
    const T eps = eps_rel * (c0 * c0 + c1 * c1 + c2 * c2) * T(1.0 / 3.0) + weno_floor<T>();
    const T v0 = T(-0.5) * c0 + T(1.5) * c1, v1 = T(0.5) * c1 + T(0.5) * c2;
    const T b0 = (c1 - c0) * (c1 - c0), b1 = (c2 - c1) * (c2 - c1);
    const T tau = fabs(b0 - b1);
    const T a0 = T(1.0 / 3.0) * (T(1.0) + dv(tau, b0 + eps)), a1 = T(2.0 / 3.0) * (T(1.0) + dv(tau, b1 + eps));
    return dv(a0 * v0 + a1 * v1, a0 + a1);
}

// the candidates' quotients by 6: the IEEE division, or quot6 (the same bits) when every stencil value is known to be
// in range_fits_quot6 (Q6)
template <bool Q6, typename T> __device__ __forceinline__ T by6(T x) {
    if constexpr (Q6 && std::is_same<T, double>::value) return quot6(x);
    else return dv(x, T(6.0));
}

template <typename T, bool Q6 = false>
__device__ __forceinline__ T weno5z(T c0, T c1, T c2, T c3, T c4, T eps_rel) {
    const T eps = eps_rel * (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4) * T(0.2) + weno_floor<T>();
    const T p0 = by6<Q6>(T(2.0) * c0 - T(7.0) * c1 + T(11.0) * c2);
    const T p1 = by6<Q6>(-c1 + T(5.0) * c2 + T(2.0) * c3);
    const T p2 = by6<Q6>(T(2.0) * c2 + T(5.0) * c3 - c4);
    const T s0 = c0 - T(2.0) * c1 + c2, t0 = c0 - T(4.0) * c1 + T(3.0) * c2;
    const T s1 = c1 - T(2.0) * c2 + c3, t1 = c1 - c3;
    const T s2 = c2 - T(2.0) * c3 + c4, t2 = T(3.0) * c2 - T(4.0) * c3 + c4;
    const T b0 = T(13.0 / 12.0) * s0 * s0 + T(0.25) * t0 * t0;
    const T b1 = T(13.0 / 12.0) * s1 * s1 + T(0.25) * t1 * t1;
    const T b2 = T(13.0 / 12.0) * s2 * s2 + T(0.25) * t2 * t2;
    const T tau = fabs(b0 - b2);
    const T a0 = T(0.1) * (T(1.0) + dv(tau, b0 + eps)), a1 = T(0.6) * (T(1.0) + dv(tau, b1 + eps)),
            a2 = T(0.3) * (T(1.0) + dv(tau, b2 + eps));
    return dv(a0 * p0 + a1 * p1 + a2 * p2, a0 + a1 + a2);
}

// ---- convective ghost state (Boundaries.cpp:106-153) ----
struct BcParams {
    int bc[6];
    double u_in, v_in, p_out;
    double rho_in;                   // mass-flow inlet density
    double nrbc_sigma, nrbc_beta;    // non-reflecting outlet: relaxation towards p_out, transverse weight
    double length[3];                // domain size per axis (the non-reflecting relaxation length)
    const double* ut;                // mass-flow inlet velocity per side, set on the device each residual (or null)
};

// The boundary type of a side, without indexing the parameter array by a run-time value (that would put the
// kernel arguments in scratch memory).
__device__ __forceinline__ int bc_of(const BcParams& b, int side) {
    switch (side) {
    case 0: return b.bc[0];
    case 1: return b.bc[1];
    case 2: return b.bc[2];
    case 3: return b.bc[3];
    case 4: return b.bc[4];
    default: return b.bc[5];
    }
}

// side: 0 left, 1 right, 2 bottom, 3 top, 4 front, 5 back. Periodic and non-reflecting sides never reach here; a
// mass-flow side reads its target velocity from b.ut (set by k_massflow before the residual's first pass).
template <int NV>
__device__ __forceinline__ void ghost_state(const Eos& e, const BcParams& b, int side, const double* Ui, double* Ug) {
    const int bc = bc_of(b, side);
    if (bc == 6 /* transmissive */) {
#pragma unroll
        for (int q = 0; q < NV; ++q) Ug[q] = Ui[q];
        return;
    }
    const int axis = side / 2;
    const double nxn = axis == 0 ? 1.0 : 0.0, nyn = axis == 1 ? 1.0 : 0.0, nzn = axis == 2 ? 1.0 : 0.0;
    double Wi[NV], Wg[NV];
    c2p<NV>(e, Ui, Wi, (double*)nullptr);
#pragma unroll
    for (int q = 0; q < NV; ++q) Wg[q] = Wi[q];
    const double wi4 = NV == 5 ? Wi[NV == 5 ? 4 : 0] : 0.0;
    if (bc == 0 /* wall */) {
        const double Vn = Wi[1] * nxn + Wi[2] * nyn + wi4 * nzn;
        Wg[1] = Wi[1] - 2.0 * Vn * nxn;
        Wg[2] = Wi[2] - 2.0 * Vn * nyn;
        if constexpr (NV == 5) Wg[4] = Wi[4] - 2.0 * Vn * nzn;
    } else if (bc == 1 /* no-slip */) {
        Wg[1] = -Wi[1];
        Wg[2] = -Wi[2];
        if constexpr (NV == 5) Wg[4] = -Wi[4];
    } else if (bc == 2 || bc == 3 /* moving wall, velocity inlet */) {
        Wg[1] = 2.0 * b.u_in - Wi[1];
        Wg[2] = 2.0 * b.v_in - Wi[2];
        if constexpr (NV == 5) Wg[4] = -Wi[4];
    } else if (bc == 4 /* mass-flow inlet: the side's target velocity */) {
        Wg[1] = 2.0 * b.ut[side] - Wi[1];
        Wg[2] = 2.0 * b.v_in - Wi[2];
        if constexpr (NV == 5) Wg[4] = -Wi[4];
    } else if (bc == 5 /* outlet */) {
        const double Vn = Wi[1] * nxn + Wi[2] * nyn + wi4 * nzn;
        const bool outflow = Vn >= 0;
        Wg[1] = outflow ? Wi[1] : 0.0;
        Wg[2] = outflow ? Wi[2] : 0.0;
        if constexpr (NV == 5) Wg[4] = outflow ? Wi[4] : 0.0;
        Wg[3] = smax(b.p_out, 1e-5);
    }
    p2c<NV>(e, Wg, Ug);
}

}   // namespace fvg
