// Learned troubled-face indicator of the WENO5-Z reconstruction (no reference counterpart; its labels come from the
// weights of weno5z, Schemes.cpp:338-353).
//
// Features of a face (15): for density, pressure and the velocity component with the largest total variation over
// the face's six-cell stencil c0..c5, the four second differences and the largest first difference, divided by
// sqrt(sum of the squared first differences + eps_rel * mean(c^2) + 1e-30) (eps_rel: WENO's relative
// regularisation). Smooth data gives small second differences; a jump or grid-scale noise gives O(1) ones.
// Network: 15 -> 32 (ReLU) -> 32 (ReLU) -> 1 logit; a face is troubled when the logit exceeds the threshold.
// Inference runs on the matrix cores, one wavefront of 64 faces at a time: two column blocks of 32 faces, each
// through v_mfma_f32_32x32x16_bf16 (hidden units as rows, faces as columns): one product for layer 1 (K = 16),
// two for layer 2 (K = 32). Lane l of a 32x32 accumulator holds rows 8j + 4h + i (h = l/32, j < 4, i < 4) of face
// column l % 32; its elements 8c .. 8c+7 are layer 2's B operand for K chunk c once W2's columns are permuted to
// match (K index 16c + 8h + s <- hidden row 16c + 8(s/4) + 4h + s%4), so the hidden layer never leaves the
// registers. x, W1 (bias b1 folded into column 15 against x15 = 1) and the first hidden layer are bfloat16 (round
// to nearest even); accumulation, b2, w3 and b3 are float.
#pragma once

#include "fvg_device.hpp"

namespace fvg {

constexpr int MLP_NIN = 15, MLP_NH = 32;
constexpr int MLP_NPARAM = MLP_NH * MLP_NIN + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1;   // 1601

// The network as the lanes of a wavefront read it.
struct Frag8 {
    unsigned long long lo, hi;   // eight bfloat16 values
};
struct MlpDev {
    Frag8 w1[64];        // W1 (b1 in column 15) for lane l: row l % 32, columns 8h .. 8h+7 (h = l / 32)
    Frag8 w2[2][64];     // W2 for K chunk c, lane l: row l % 32, hidden rows 16c + 8(s/4) + 4h + s%4, s < 8
    float b2h[2][16];    // b2 at the accumulator rows of lane half h: 8(e/4) + 4h + e%4
    float w3h[2][16];    // w3 at the same rows
    float b3, thr;
};

// Audit counters and the indicator's own switch, kept for the life of a context.
struct IndState {
    unsigned long long last[4];    // TP, FP, FN, TN of the last audit (positive = troubled)
    unsigned long long total[4];   // over all audits
    long long steps;               // steps completed by the context
    int off;                       // 1: the audit found too many missed troubled faces; WENO5-Z everywhere
    int audits;
    int pending;                   // the current step is being taken
    // per-block audit counts land in one of kAuditSlots slot rows (block index mod kAuditSlots), so that the atomics
    // of the blocks of one audit spread over many addresses; the control kernel sums the rows into last[]
    unsigned long long part[64][4];
};
constexpr int kAuditSlots = 64;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> two bfloat16 in one dword (a in the low half), rounded to nearest even (gfx950's
// v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned pack2(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}
__device__ __forceinline__ unsigned long long pack4(float a, float b, float c, float d) {
    return (unsigned long long)pack2(a, b) | ((unsigned long long)pack2(c, d) << 32);
}
__device__ __forceinline__ bf16x8 as_frag(unsigned long long lo, unsigned long long hi) {
    Frag8 f{lo, hi};
    return __builtin_bit_cast(bf16x8, f);
}
// the value of lane l ^ 32 (gfx950's v_permlane32_swap: of the two copies of v, the first gets lanes 0-31 of the
// second in its lanes 32-63, the second gets lanes 32-63 of the first in its lanes 0-31)
__device__ __forceinline__ unsigned xor32(unsigned v) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
}
__device__ __forceinline__ unsigned long long xor32(unsigned long long v) {
    return (unsigned long long)xor32((unsigned)v) | ((unsigned long long)xor32((unsigned)(v >> 32)) << 32);
}
__device__ __forceinline__ float xor32(float v) { return __uint_as_float(xor32(__float_as_uint(v))); }

// linear upwind-5 face value (weno5z with the linear weights)
template <typename T> __device__ __forceinline__ T upwind5(T c0, T c1, T c2, T c3, T c4) {
    return (T(2.0) * c0 - T(13.0) * c1 + T(47.0) * c2 + T(27.0) * c3 - T(3.0) * c4) * T(1.0 / 60.0);
}

// max_k |w_k - d_k| of the WENO5-Z weights of one face value, with the regularisation eps_rel * m + floor, where m
// is the mean square of the five values (density, pressure) or the face's kinetic-plus-acoustic scale (velocities)
template <typename T> __device__ __forceinline__ T weno5z_departure(T c0, T c1, T c2, T c3, T c4, T eps_rel, T m) {
    const T eps = eps_rel * m + weno_floor<T>();
    const T s0 = c0 - T(2.0) * c1 + c2, t0 = c0 - T(4.0) * c1 + T(3.0) * c2;
    const T s1 = c1 - T(2.0) * c2 + c3, t1 = c1 - c3;
    const T s2 = c2 - T(2.0) * c3 + c4, t2 = T(3.0) * c2 - T(4.0) * c3 + c4;
    const T b0 = T(13.0 / 12.0) * s0 * s0 + T(0.25) * t0 * t0;
    const T b1 = T(13.0 / 12.0) * s1 * s1 + T(0.25) * t1 * t1;
    const T b2 = T(13.0 / 12.0) * s2 * s2 + T(0.25) * t2 * t2;
    const T tau = fabs(b0 - b2);
    const T a0 = T(0.1) * (T(1.0) + dv(tau, b0 + eps)), a1 = T(0.6) * (T(1.0) + dv(tau, b1 + eps)),
            a2 = T(0.3) * (T(1.0) + dv(tau, b2 + eps));
    const T inv = dv(T(1.0), a0 + a1 + a2);
    return smax(smax(fabs(a0 * inv - T(0.1)), fabs(a1 * inv - T(0.6))), fabs(a2 * inv - T(0.3)));
}

template <typename T> __device__ __forceinline__ T mean_square5(T c0, T c1, T c2, T c3, T c4) {
    return (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4) * T(0.2);
}

// label value of a face: the largest departure over the variables and both face states. A velocity component is
// regularised by the face's scale mean(|v|^2 + p / rho) over its six cells rather than by its own mean square, so
// that a component near zero (where rounding of the much larger pressure flux leaves noise of ~1e-7) does not count
// as a discontinuity; density and pressure keep their own mean square, as WENO5-Z itself regularises them.
template <int NV, typename T> __device__ __forceinline__ T velocity_scale(const T (*c)[NV]) {
    T vs = T(0);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        T v2 = c[s][1] * c[s][1] + c[s][2] * c[s][2];
        if constexpr (NV == 5) v2 += c[s][4] * c[s][4];
        vs += v2 + dv(c[s][3], c[s][0]);
    }
    return vs * T(1.0 / 6.0);
}

template <int NV, typename T> __device__ __forceinline__ T face_departure(const T (*c)[NV], T eps_rel, T vs) {
    T d = T(0);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const bool vel = q == 1 || q == 2 || q == 4;
        const T ml = vel ? vs : mean_square5(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q]);
        const T mr = vel ? vs : mean_square5(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q]);
        d = smax(d, weno5z_departure(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], eps_rel, ml));
        d = smax(d, weno5z_departure(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], eps_rel, mr));
    }
    return d;
}

// How far the MUSCL limiter moves the slope of one face value off the central slope: |phi(a, b) - (a + b)/2| over
// sqrt((a^2 + b^2)/2 + eps_rel * m + 1e-30), with a the upwind and b the central difference as the reconstruction
// takes them (apply_limiter, Schemes.hpp:110-117; wl = c1 + phi(c1 - c0, c2 - c1) / 2, Schemes.cpp:261). Zero on
// linear data for every limiter (phi(a, a) = a).
template <typename T> __device__ __forceinline__ T muscl_clip(int lim, T a, T b, T eps_rel, T m) {
    const T phi = limiter(lim, a, b);
    const T mid = T(0.5) * (a + b);
    return dv(fabs(phi - mid), sqrt(T(0.5) * (a * a + b * b) + eps_rel * m + T(1e-30)));
}

template <typename T> __device__ __forceinline__ T mean_square4(T c0, T c1, T c2, T c3) {
    return (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3) * T(0.25);
}

// clip label value of a face: the largest clip over the variables and both face values of the MUSCL stencil c1..c4
// of the six-cell stencil (cells t-2 .. t+1 of the face between t-1 and t); the regularisation scale m as in
// face_departure (a velocity component takes the face's kinetic-plus-acoustic scale, over the six cells)
template <int NV, typename T> __device__ __forceinline__ T face_clip(const T (*c)[NV], T eps_rel, int lim, T vs) {
    T d = T(0);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const bool vel = q == 1 || q == 2 || q == 4;
        const T m = vel ? vs : mean_square4(c[1][q], c[2][q], c[3][q], c[4][q]);
        const T b = c[3][q] - c[2][q];
        d = smax(d, muscl_clip(lim, c[2][q] - c[1][q], b, eps_rel, m));   // left value: wl = c2 + phi(c2 - c1, c3 - c2) / 2
        d = smax(d, muscl_clip(lim, c[4][q] - c[3][q], b, eps_rel, m));   // right value: wr = c3 - phi(c4 - c3, c3 - c2) / 2
    }
    return d;
}

// label sources (enum fvg_label_source of include/fvg.h)
constexpr int LABEL_WENO5 = 0, LABEL_MUSCL = 1, LABEL_BOTH = 2;

// label value of a face from its six-cell stencil: the WENO5-Z weight departure, the MUSCL limiter clip of the
// configured limiter, or the larger of the two
template <int NV, typename T> __device__ __forceinline__ T face_label(const T (*c)[NV], T eps_rel, int src, int lim) {
    const T vs = velocity_scale<NV, T>(c);
    T d = T(0);
    if (src != LABEL_MUSCL) d = face_departure<NV, T>(c, eps_rel, vs);
    if (src != LABEL_WENO5) d = smax(d, face_clip<NV, T>(c, eps_rel, lim, vs));
    return d;
}

// five features of one variable over the six-cell stencil: its four second differences and its largest first
// difference, divided by N = sqrt(sum of the squared first differences + eps_rel * mean(c^2) + 1e-30)
__device__ __forceinline__ void feat5(float v0, float v1, float v2, float v3, float v4, float v5, float eps_rel, float* x) {
    const float d0 = v1 - v0, d1 = v2 - v1, d2 = v3 - v2, d3 = v4 - v3, d4 = v5 - v4;
    const float s = d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3 + d4 * d4;
    const float m = (v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3 + v4 * v4 + v5 * v5) * (1.0f / 6.0f);
    const float r = __builtin_amdgcn_rsqf(s + eps_rel * m + 1e-30f);
    x[0] = (d1 - d0) * r;
    x[1] = (d2 - d1) * r;
    x[2] = (d3 - d2) * r;
    x[3] = (d4 - d3) * r;
    x[4] = fmaxf(fmaxf(fmaxf(fabsf(d0), fabsf(d1)), fmaxf(fabsf(d2), fabsf(d3))), fabsf(d4)) * r;
}

// the 15 features of a face (x[15] = 1 carries the first layer's bias) from its six-cell stencil c; an FP64
// stencil is rounded to float first, so the network sees the same inputs in both flux precisions
template <int NV, typename T> __device__ __forceinline__ void face_features(const T (*cs)[NV], float eps_rel, float* x) {
    float c[6][NV];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int q = 0; q < NV; ++q) c[s][q] = float(cs[s][q]);
    feat5(c[0][0], c[1][0], c[2][0], c[3][0], c[4][0], c[5][0], eps_rel, x);
    feat5(c[0][3], c[1][3], c[2][3], c[3][3], c[4][3], c[5][3], eps_rel, x + 5);
    // the velocity slot with the largest total variation (the first on ties), picked by selects
    float v[6] = {c[0][1], c[1][1], c[2][1], c[3][1], c[4][1], c[5][1]}, best = -1.0f;
#pragma unroll
    for (int a = 0; a < NV - 2; ++a) {
        const int q = mom(a);
        float tv = 0.0f;
#pragma unroll
        for (int j = 0; j < 5; ++j) tv += fabsf(c[j + 1][q] - c[j][q]);
        const bool take = tv > best;
        best = take ? tv : best;
#pragma unroll
        for (int j = 0; j < 6; ++j) v[j] = take ? c[j][q] : v[j];
    }
    feat5(v[0], v[1], v[2], v[3], v[4], v[5], eps_rel, x + 10);
    x[15] = 1.0f;
}

// Logit of this lane's face; every lane of the wavefront must call it (the matrix instructions use all 64).
__device__ __forceinline__ float mlp_logit_wave(const MlpDev* __restrict__ m, const float* x) {
    const int lane = threadIdx.x & 63, half = lane >> 5;
    unsigned long long g[4], o[4];   // features 4q .. 4q+3 of the own face and of the partner lane's (l ^ 32)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        g[q] = pack4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
        o[q] = xor32(g[q]);
    }
    const Frag8 w1 = m->w1[lane], w2a = m->w2[0][lane], w2b = m->w2[1][lane];
    float out = 0.0f;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        // B operand: features 8h .. 8h+7 of face 32 cb + l % 32, owned by this lane when cb == h, else by the partner
        const bool own = cb == half;
        const unsigned long long b0 = own ? (half ? g[2] : g[0]) : (half ? o[2] : o[0]);
        const unsigned long long b1 = own ? (half ? g[3] : g[1]) : (half ? o[3] : o[1]);
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_frag(w1.lo, w1.hi), as_frag(b0, b1), acc, 0, 0, 0);
        f32x16 acc2;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc2[e] = m->b2h[half][e];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned long long h0 = pack4(fmaxf(acc[8 * c], 0.0f), fmaxf(acc[8 * c + 1], 0.0f), fmaxf(acc[8 * c + 2], 0.0f),
                                                fmaxf(acc[8 * c + 3], 0.0f));
            const unsigned long long h1 = pack4(fmaxf(acc[8 * c + 4], 0.0f), fmaxf(acc[8 * c + 5], 0.0f), fmaxf(acc[8 * c + 6], 0.0f),
                                                fmaxf(acc[8 * c + 7], 0.0f));
            const Frag8 w = c == 0 ? w2a : w2b;
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_frag(w.lo, w.hi), as_frag(h0, h1), acc2, 0, 0, 0);
        }
        float part = 0.0f;
#pragma unroll
        for (int e = 0; e < 16; ++e) part += m->w3h[half][e] * fmaxf(acc2[e], 0.0f);
        const float z = part + xor32(part) + m->b3;
        if (own) out = z;
    }
    return out;
}

}   // namespace fvg
