// Device code of the uniform-grid residual and time step: the helpers and the kernel templates. See
// fvg_kernels.hpp for the pass structure. Reference lines are cited per kernel; the arithmetic itself is in
// fvg_device.hpp. The templates are instantiated in fvg_kernels.hip (control, learned indicator, totals),
// fvg_kernels_f64.hip and fvg_kernels_f32.hip, compiled in parallel.
#pragma once

#include <type_traits>

#include "fvg_kernels.hpp"

namespace fvg {

constexpr double MU_DIAG = 4.0 / 3.0;   // Physics.cpp:488
constexpr double MU_OFF = 2.0 / 3.0;

__device__ __forceinline__ long long cidx(const KP& p, int i, int j, int k) { return ((long long)k * p.ny + j) * p.nx + i; }
__device__ __forceinline__ int ncell(const KP& p, int d) { return d == 0 ? p.nx : (d == 1 ? p.ny : p.nz); }
__device__ __forceinline__ double hstep(const KP& p, int d) { return d == 0 ? p.dx : (d == 1 ? p.dy : p.dz); }
// the coordinate along d of (i, j, k), and (i, j, k) with that coordinate replaced by t
__device__ __forceinline__ int coord(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }
__device__ __forceinline__ long long cidx_d(const KP& p, int d, int i, int j, int k, int t) {
    return cidx(p, d == 0 ? t : i, d == 1 ? t : j, d == 2 ? t : k);
}
// Accessors of the per-side / per-axis parameters without run-time array indexing (which would move the
// kernel arguments to scratch memory).
__device__ __forceinline__ int halo_of(const KP& p, int s) {
    switch (s) {
    case 0: return p.halo[0];
    case 1: return p.halo[1];
    case 2: return p.halo[2];
    case 3: return p.halo[3];
    case 4: return p.halo[4];
    default: return p.halo[5];
    }
}
__device__ __forceinline__ int vper_of(const KP& p, int a) { return a == 0 ? p.vper[0] : (a == 1 ? p.vper[1] : p.vper[2]); }
__device__ __forceinline__ int wrap_of(const KP& p, int a) { return a == 0 ? p.wrap[0] : (a == 1 ? p.wrap[1] : p.wrap[2]); }

template <int NV> __device__ __forceinline__ void load_cons(const KP& p, long long c, double* u) {
#pragma unroll
    for (int q = 0; q < NV; ++q) u[q] = p.U[q * p.nc + c];
}
// element of cell (i, j, k) in the padded primitive planes (each coordinate from -PH)
__device__ __forceinline__ long long pidx(const KP& p, int i, int j, int k) {
    return ((long long)(k + p.phz) * p.py + (j + PH)) * p.px + (i + PX0);
}
__device__ __forceinline__ long long pidx_d(const KP& p, int d, int i, int j, int k, int t) {
    return pidx(p, d == 0 ? t : i, d == 1 ? t : j, d == 2 ? t : k);
}
// the primitives of padded element pc
template <int NV, typename T = double> __device__ __forceinline__ void load_prim(const KP& p, long long pc, T* w) {
    const T* W = (const T*)p.W;
#pragma unroll
    for (int q = 0; q < NV; ++q) w[q] = W[q * p.npad + pc];
}

template <int NV, typename T> __device__ __forceinline__ void to_t(const double* a, T* b) {
#pragma unroll
    for (int q = 0; q < NV; ++q) b[q] = T(a[q]);
}

// Blocks are dispatched round-robin over the 8 XCDs of the chip, each with its own L2. The logical block of a
// dispatched block gives every XCD a contiguous range of logical blocks (rows and planes of the mesh), so that the
// stencil rows a block shares with the next one are found in that XCD's L2. The tail that does not divide by 8
// keeps its order.
__device__ __forceinline__ long long xcd_block() {
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb >> 3;
    if (b >= per * 8u) return b;
    return (long long)(b & 7u) * per + (b >> 3);
}

__device__ __forceinline__ unsigned mdiv(unsigned n, Magic g) { return (__umulhi(n, g.m) + n) >> g.l; }

// The row and chunk of a row block (a grid of nplanes * ny * nchunk blocks of rbt threads; the row of
// (j, k) holds nx elements). Blocks are dispatched round-robin over the 8 XCDs; when ny divides by 8, XCD x takes
// the rows j in [x ny/8, (x+1) ny/8) of every plane, plane after plane, so the few planes a z-stencil spans of its
// slab stay in that XCD's L2. Otherwise the rows go in mesh order.
struct RowBlock {
    int i, j, k;   // this thread's x position (may be >= nx in the last chunk), row and plane
};
__device__ __forceinline__ RowBlock row_block(const KP& p) {
    const unsigned b = blockIdx.x;
    unsigned m = b, j0 = 0;
    Magic rows = p.mg_ny;
    unsigned nrow = p.ny;
    if (p.slabs) {
        m = b >> 3;
        nrow = p.ny >> 3;
        j0 = (b & 7u) * nrow;
        rows = p.mg_slab;
    }
    const unsigned r = mdiv(m, p.mg_chunk), chunk = m - r * p.nchunk;
    const unsigned k = mdiv(r, rows), jj = r - k * nrow;
    return {int(chunk * p.rbt + threadIdx.x), int(j0 + jj), int(k)};
}

// Wave-level then block-level maximum of non-negative, non-NaN doubles, then one atomicMax of the bit pattern.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* dst) {
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double other = __shfl_xor(v, o, 64);
        v = (other > v) ? other : v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) m = (red[w] > m) ? red[w] : m;
        if (m > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(m));
    }
}
// the block maximum into this block's slot of quantity q (SPD_CONV, SPD_VISC)
__device__ __forceinline__ void block_max_slot(const KP& p, double v, int q) {
    block_max_to(v, p.spd + ((size_t)q * kSpeedSlots + blockIdx.x % kSpeedSlots) * kSlotStride);
}

// ------------------------------------------------------------------------------------------------------------
// Stencil access
// ------------------------------------------------------------------------------------------------------------

// The padded primitive state of the reconstruction at the cell whose coordinate along d is t (the other two
// from (i, j, k)): the cell itself, or outside the domain the periodic wrap, the ghost-slab layer
// (fill_ghost_slab, Boundaries.cpp:202-228) or the edge cell (pad_primitives, Schemes.cpp:915-969).
// The primitives a reconstruction stencil along d takes at coordinate t (-PH <= t < n + PH) of the line through
// (i, j, k): the cell itself inside the mesh, else the halo value k_cells stored there (push_halo).
template <int NV, typename T = double>
__device__ __forceinline__ void stencil_prim(const KP& p, int d, int i, int j, int k, int t, T* w) {
    load_prim<NV, T>(p, pidx_d(p, d, i, j, k, t), w);
}

// The non-reflecting outlet's ghost of boundary cell (i, j, k) on side SIDE (nonreflecting_ghost and the tangential
// neighbours of apply_side, Boundaries.cpp:155-198, 314-360): a characteristic outlet whose ghost differs from the
// cell only in the incoming acoustic wave, relaxed towards p_out with the transverse terms weighted by 1 - beta.
// Tangential derivatives are central differences of the primitives, wrapped across a periodic axis, one-sided at
// the ends of the others. k_nrbc computes it once per stage into the side's plane of KP::nrbc, which the face
// kernels read.
template <int NV, int SIDE>
__device__ __forceinline__ void nrbc_ghost(const KP& p, int i, int j, int k, double* ug) {
    constexpr int axis = SIDE >> 1;
    constexpr double s = (SIDE & 1) ? 1.0 : -1.0;
    constexpr int qv = mom(axis);
    const BcParams& b = p.bp;
    double Wi[NV], Wg[NV];
    load_cons<NV>(p, cidx(p, i, j, k), ug);
    c2p<NV>(p.rs.eos, ug, Wi, (double*)nullptr);
#pragma unroll
    for (int q = 0; q < NV; ++q) Wg[q] = Wi[q];
    const double u_out = s * Wi[qv];
    const double c = sound_speed(p.rs.eos, Wi[3], Wi[0]);
    const double M = smin(smax(u_out / c, -1.0), 1.0);
    if (M >= 1.0) {   // supersonic outflow: nothing enters
        p2c<NV>(p.rs.eos, Wg, ug);
        return;
    }
    const double h = hstep(p, axis);
    const double len = axis == 0 ? b.length[0] : (axis == 1 ? b.length[1] : b.length[2]);
    const double K = b.nrbc_sigma * c * (1.0 - M * M) / len;
    const double beta = b.nrbc_beta >= 0.0 ? b.nrbc_beta : fabs(M);
    double T1 = 0.0;
    if (beta < 1.0) {
#pragma unroll
        for (int t = 0; t < NV - 2; ++t) {
            const int n = ncell(p, t);
            if (t == axis || n == 1) continue;
            const int ct = coord(t, i, j, k);
            int lo = ct - 1, hi = ct + 1;
            double span = 2.0;
            if (vper_of(p, t)) {
                lo = (lo + n) % n;
                hi = hi % n;
            } else {
                if (lo < 0) {
                    lo = 0;
                    span -= 1.0;
                }
                if (hi >= n) {
                    hi = n - 1;
                    span -= 1.0;
                }
                if (span <= 0.0) continue;
            }
            double u[NV], Wl[NV], Wh[NV];
            load_cons<NV>(p, cidx_d(p, t, i, j, k, lo), u);
            c2p<NV>(p.rs.eos, u, Wl, (double*)nullptr);
            load_cons<NV>(p, cidx_d(p, t, i, j, k, hi), u);
            c2p<NV>(p.rs.eos, u, Wh, (double*)nullptr);
            const double ht = span * hstep(p, t);
            const int qt = mom(t);
            T1 += Wi[qt] * (Wh[3] - Wl[3]) / ht + Wi[0] * c * c * (Wh[qt] - Wl[qt]) / ht -
                  Wi[0] * c * s * Wi[qt] * (Wh[qv] - Wl[qv]) / ht;
        }
    }
    const double dw1 = -h * (K * (Wi[3] - b.p_out) - (1.0 - beta) * T1) / (c - u_out);
    Wg[3] = Wi[3] + 0.5 * dw1;
    Wg[qv] = Wi[qv] - s * dw1 / (2.0 * Wi[0] * c);
    Wg[0] = Wi[0] + 0.5 * dw1 / (c * c);
    p2c<NV>(p.rs.eos, Wg, ug);
}

// boundary cells of a side (Core.hpp:121-151) and the index of cell (i, j, k) among them
__device__ __forceinline__ long long side_cells(const KP& p, int side) {
    return side < 2 ? (long long)p.ny * p.nz : (side < 4 ? (long long)p.nx * p.nz : (long long)p.nx * p.ny);
}
__device__ __forceinline__ long long side_index(const KP& p, int side, int i, int j, int k) {
    return side < 2 ? (long long)k * p.ny + j : (side < 4 ? (long long)k * p.nx + i : (long long)j * p.nx + i);
}
__device__ __forceinline__ long long nrbc_off_of(const KP& p, int s) {
    switch (s) {
    case 0: return p.nrbc_off[0];
    case 1: return p.nrbc_off[1];
    case 2: return p.nrbc_off[2];
    case 3: return p.nrbc_off[3];
    case 4: return p.nrbc_off[4];
    default: return p.nrbc_off[5];
    }
}

// First-order state outside side `side` (along axis d) of the boundary cell (i, j, k) (apply_side,
// Boundaries.cpp:259-371): the opposite cell on a periodic side, the non-reflecting ghost, else the ghost state of
// the cell.
template <int NV>
__device__ __forceinline__ void outside_first_order(const KP& p, int d, int side, int i, int j, int k, double* ug) {
    const int bc = bc_of(p.bp, side);
    if (bc == 7) {
        load_cons<NV>(p, cidx_d(p, d, i, j, k, (side & 1) ? 0 : ncell(p, d) - 1), ug);
        return;
    }
    if (bc == 8) {   // the non-reflecting ghost k_nrbc stored for this boundary cell
        const long long n = side_cells(p, side), kb = side_index(p, side, i, j, k);
        const double* g = p.nrbc + nrbc_off_of(p, side);
#pragma unroll
        for (int q = 0; q < NV; ++q) ug[q] = g[q * n + kb];
        return;
    }
    double ui[NV];
    load_cons<NV>(p, cidx(p, i, j, k), ui);
    ghost_state<NV>(p.rs.eos, p.bp, side, ui, ug);
}

// MUSCL face values of one variable (Schemes.cpp:627-635): wl from (c0, c1, c2), wr from (c1, c2, c3)
__device__ __forceinline__ double muscl_wl(int lim, double c0, double c1, double c2) { return c1 + 0.5 * limiter(lim, c1 - c0, c2 - c1); }
__device__ __forceinline__ double muscl_wr(int lim, double c1, double c2, double c3) { return c2 - 0.5 * limiter(lim, c3 - c2, c2 - c1); }

// ------------------------------------------------------------------------------------------------------------
// Viscous padded state (Boundaries.cpp:379-700): cell coordinates may lie one cell outside the domain.
// ------------------------------------------------------------------------------------------------------------

// viscous_ghost_3d (Boundaries.cpp:253-269); the 2-D frame rules are the same on the first four components.
template <int NV>
__device__ __forceinline__ void viscous_ghost(const KP& p, int side, const double* c, double* g) {
#pragma unroll
    for (int q = 0; q < NV; ++q) g[q] = c[q];
    const int d = side >> 1, bc = bc_of(p.bp, side);
    if (bc == 1) {   // no-slip
#pragma unroll
        for (int a = 0; a < NV - 2; ++a) g[mom(a)] = -c[mom(a)];
    } else if (bc == 0) {   // slip wall
#pragma unroll
        for (int a = 0; a < NV - 2; ++a)
            if (a == d) g[mom(a)] = -c[mom(a)];
    } else if (bc == 2) {   // moving wall
#pragma unroll
        for (int t = 0; t < NV - 2; ++t) {
            const double uw = t == 0 ? p.bp.u_in : (t == 1 ? p.bp.v_in : 0.0);
            g[mom(t)] = (t == d) ? -c[mom(t)] : 2.0 * c[0] * uw - c[mom(t)];
        }
        if (p.bo2) {
            // match_kinetic_energy (Boundaries.cpp:194-199); in 2-D the w terms are +0 and are kept for the same rounding
            const double g4 = NV == 5 ? g[NV - 1] : 0.0, c4 = NV == 5 ? c[NV - 1] : 0.0;
            g[3] = c[3] + 0.5 * (g[1] * g[1] + g[2] * g[2] + g4 * g4 - c[1] * c[1] - c[2] * c[2] - c4 * c4) / c[0];
        }
    }
}

// The face-ghost entry of the viscous padded state at cell coordinates (i, j, k) outside the domain along axis
// a only (padded_3d / get_padded_frame side rules): the periodic opposite cell, the wall rules, else the edge copy.
template <int NV>
__device__ __forceinline__ void visc_face_ghost(const KP& p, int i, int j, int k, int a, double* out) {
    const int n = ncell(p, a);
    const int hi = coord(a, i, j, k) >= n, side = 2 * a + hi;
    const int bc = bc_of(p.bp, side);
    if (bc == 7) {
        load_cons<NV>(p, cidx_d(p, a, i, j, k, hi ? 0 : n - 1), out);
        return;
    }
    double in[NV];
    load_cons<NV>(p, cidx_d(p, a, i, j, k, hi ? n - 1 : 0), in);
    if (bc == 0 || bc == 1 || bc == 2) {
        viscous_ghost<NV>(p, side, in, out);
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = in[q];
    }
}

// Conservative padded state at cell coordinates (i, j, k), each in [-1, n]. One axis outside: the side rule;
// two axes outside with boundary_order = 2: the rule of the lower axis applied to the face-ghost entry of the
// higher one (fix_padded_corners, Boundaries.cpp:520-560; fix_frame_corners, :562-604); otherwise the edge copy.
template <int NV>
__device__ __forceinline__ void visc_pad_cons(const KP& p, int i, int j, int k, double* out) {
    const int ox = i < 0 || i >= p.nx, oy = j < 0 || j >= p.ny, oz = k < 0 || k >= p.nz;
    const int nout = ox + oy + oz;
    if (nout == 0) {
        load_cons<NV>(p, cidx(p, i, j, k), out);
        return;
    }
    if (nout == 1) {
        visc_face_ghost<NV>(p, i, j, k, ox ? 0 : (oy ? 1 : 2), out);
        return;
    }
    if (nout == 2 && p.bo2) {
        const int d = ox ? 0 : 1, second = (ox && oy) ? 1 : 2;
        const int n = ncell(p, d);
        const int hi = coord(d, i, j, k) >= n, side = 2 * d + hi;
        if (bc_of(p.bp, side) == 7) {
            const int t = hi ? 0 : n - 1;
            visc_face_ghost<NV>(p, d == 0 ? t : i, d == 1 ? t : j, k, second, out);
        } else {
            const int t = hi ? n - 1 : 0;
            double cin[NV];
            visc_face_ghost<NV>(p, d == 0 ? t : i, d == 1 ? t : j, k, second, cin);
            viscous_ghost<NV>(p, side, cin, out);
        }
        return;
    }
    const int ci = i < 0 ? 0 : (i >= p.nx ? p.nx - 1 : i);
    const int cj = j < 0 ? 0 : (j >= p.ny ? p.ny - 1 : j);
    const int ck = k < 0 ? 0 : (k >= p.nz ? p.nz - 1 : k);
    load_cons<NV>(p, cidx(p, ci, cj, ck), out);
}

template <int NV>
__device__ __forceinline__ void visc_pad_prim(const KP& p, int i, int j, int k, double* w) {
    double u[NV];
    visc_pad_cons<NV>(p, i, j, k, u);
    c2p<NV>(p.rs.eos, u, w, (double*)nullptr);
}

// ------------------------------------------------------------------------------------------------------------
// k_cells: primitives; with viscosity the cell velocity gradients and the viscous speed bound
// (Physics.cpp:44-76 2-D, 520-555 and 733-741 3-D; ViscousBound.hpp:16-24)
// ------------------------------------------------------------------------------------------------------------
// Viscous cell gradients G (velocity gradients from central differences of the conservative padded state) and the
// cell's viscous speed bound (Physics.cpp:44-76 2-D, 520-555 and 733-741 3-D; ViscousBound.hpp:16-24), FP64 in the
// reference's operation order.
template <int NV>
__device__ __forceinline__ double cell_gradients_f64(const KP& p, long long c, int i, int j, int k, const double* w) {
    const double rho = w[0];
    constexpr int ND = NV - 2;
    double grad[ND][ND + 1];   // grad[b][slot]: slots 0 (rho) and the momenta 1..ND
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        double lo[NV], hi[NV];
        const int o0[3] = {b == 0 ? -1 : 0, b == 1 ? -1 : 0, b == 2 ? -1 : 0};
        visc_pad_cons<NV>(p, i + o0[0], j + o0[1], k + o0[2], lo);
        visc_pad_cons<NV>(p, i - o0[0], j - o0[1], k - o0[2], hi);
        const double h2 = 2.0 * hstep(p, b);
        grad[b][0] = (hi[0] - lo[0]) / h2;
#pragma unroll
        for (int a = 0; a < ND; ++a) grad[b][a + 1] = (hi[mom(a)] - lo[mom(a)]) / h2;
    }
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const double ua = w[mom(a)];
#pragma unroll
        for (int b = 0; b < ND; ++b) ((double*)p.G)[(a * ND + b) * p.nc + c] = (grad[b][a + 1] - ua * grad[b][0]) / rho;
    }
    const double g = p.rs.eos.gamma;
    const double nu = smax((4.0 / 3.0) * (p.mu / rho), (p.k_cond * (g - 1.0)) / rho) + 1e-12;
    double mv = fabs(w[1]);
    if (mv < fabs(w[2])) mv = fabs(w[2]);
    if constexpr (NV == 5)
        if (mv < fabs(w[4])) mv = fabs(w[4]);
    return mv + sound_speed(p.rs.eos, w[3], rho) + 2.0 * p.Ldt * p.inv_h2 * nu;
}

// The same in float for the FP32-flux mode: products with the reciprocals of 2h and rho in place of divisions.
template <int NV>
__device__ __forceinline__ double cell_gradients_f32(const KP& p, long long c, int i, int j, int k, const float* w) {
    const float rho = w[0], inv_rho = __builtin_amdgcn_rcpf(rho);
    constexpr int ND = NV - 2;
    float grad[ND][ND + 1];
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        double lo[NV], hi[NV];
        const int o0[3] = {b == 0 ? -1 : 0, b == 1 ? -1 : 0, b == 2 ? -1 : 0};
        visc_pad_cons<NV>(p, i + o0[0], j + o0[1], k + o0[2], lo);
        visc_pad_cons<NV>(p, i - o0[0], j - o0[1], k - o0[2], hi);
        const float inv2h = float(0.5 / hstep(p, b));
        grad[b][0] = float(hi[0] - lo[0]) * inv2h;
#pragma unroll
        for (int a = 0; a < ND; ++a) grad[b][a + 1] = float(hi[mom(a)] - lo[mom(a)]) * inv2h;
    }
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float ua = w[mom(a)];
#pragma unroll
        for (int b = 0; b < ND; ++b) ((float*)p.G)[(a * ND + b) * p.nc + c] = (grad[b][a + 1] - ua * grad[b][0]) * inv_rho;
    }
    const double g = p.rs.eos.gamma;
    const double rd = (double)rho;
    const double nu = smax((4.0 / 3.0) * (p.mu / rd), (p.k_cond * (g - 1.0)) / rd) + 1e-12;
    float mv = fabsf(w[1]);
    if (mv < fabsf(w[2])) mv = fabsf(w[2]);
    if constexpr (NV == 5)
        if (mv < fabsf(w[4])) mv = fabsf(w[4]);
    return double(mv) + double(sound_speed(p.rs.eos, w[3], rho)) + 2.0 * p.Ldt * p.inv_h2 * nu;
}

// The halo of the padded primitive planes, pushed by the cells it copies (the padded state of Simulation.cpp:863-905,
// 1109-1170): along an axis with a periodic halo a cell within PH of one side writes its primitives into the halo
// beyond the other side; with a ghost slab (boundary_order = 2) it writes the primitives of its ghost state at its
// mirror position; a side with neither has its halo layers copy the edge cell. u: the cell's state, w: its
// primitives as stored. CONS: the planes Wo hold the conservative state (w = u, and a ghost is stored as it is).
// wide: set when a ghost value stored is outside range_fits_quot6 (the copies of w are the caller's to check).
template <int NV, typename T, bool CONS = false>
__device__ __forceinline__ void push_halo(const KP& p, T* Wo, int i, int j, int k, const double* u, const T* w, bool& wide) {
    auto put = [&](int d, int t, const T* v) {
        const long long pc = pidx_d(p, d, i, j, k, t);
#pragma unroll
        for (int q = 0; q < NV; ++q) Wo[q * p.npad + pc] = v[q];
    };
    auto put_ghost = [&](int d, int side, int t) {
        double ug[NV], wg[NV];
        T v[NV];
        ghost_state<NV>(p.rs.eos, p.bp, side, u, ug);
        if constexpr (CONS) {
            to_t<NV>(ug, v);
        } else {
            c2p<NV>(p.rs.eos, ug, wg, (double*)nullptr);
            to_t<NV>(wg, v);
        }
        if constexpr (std::is_same<T, double>::value) {   // FP64 primitives, or the conservative ghosts
#pragma unroll
            for (int q = 0; q < NV; ++q) wide |= !range_fits_quot6(v[q]);
        }
        put(d, t, v);
    };
#pragma unroll
    for (int d = 0; d < NV - 2; ++d) {
        const int t = coord(d, i, j, k), n = ncell(p, d);
        if (t >= PH && t < n - PH) continue;   // no halo value comes from this cell along d
        const int hl = halo_of(p, 2 * d), hh = halo_of(p, 2 * d + 1);
        if (hl == 1) {   // periodic: both sides wrap
            if (t >= n - PH) put(d, t - n, w);
            if (t < PH) put(d, t + n, w);
            continue;
        }
        if (hl == 2) {
            if (t < PH) put_ghost(d, 2 * d, -1 - t);
        } else if (t == 0) {
            for (int l = 1; l <= PH; ++l) put(d, -l, w);
        }
        if (hh == 2) {
            if (t >= n - PH) put_ghost(d, 2 * d + 1, 2 * n - 1 - t);
        } else if (t == n - 1) {
            for (int l = 0; l < PH; ++l) put(d, n + l, w);
        }
    }
}
template <int NV, typename T = double>
__global__ __launch_bounds__(256) void k_cells(KP p) {
    const RowBlock rb = row_block(p);
    const int i = rb.i, j = rb.j, k = rb.k;
    const long long c = ((long long)k * p.ny + j) * p.nx + i;
    double local = 0.0;
    bool wide = false;   // a primitive stored in the FP64 planes is outside range_fits_quot6
    if (i < p.nx && !p.st->done) {
        const long long pc = pidx(p, i, j, k);
        if constexpr (std::is_same<T, float>::value) {
            // FP32 fluxes: the primitives are the FP64 ones rounded to float (the viscous ghosts of a periodic side
            // round the same FP64 primitives, so the two faces of a periodic seam stay identical); the viscous cell
            // gradients are computed in float
            double u[NV], wd[NV];
            float w[NV];
            load_cons<NV>(p, c, u);
            c2p<NV>(p.rs.eos, u, wd, (double*)nullptr);
            to_t<NV>(wd, w);
            float* Wo = (float*)p.W;
#pragma unroll
            for (int q = 0; q < NV; ++q) Wo[q * p.npad + pc] = w[q];
            push_halo<NV, float>(p, Wo, i, j, k, u, w, wide);
            if (p.viscous) local = smax(local, cell_gradients_f32<NV>(p, c, i, j, k, w));
        } else {
            double u[NV], w[NV];
            load_cons<NV>(p, c, u);
            c2p<NV>(p.rs.eos, u, w, (double*)nullptr);
            if (p.quad4) {   // fourth-order fluxes: the padded conservative state; the primitives only for the viscous term
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    p.Wc[q * p.npad + pc] = u[q];
                    wide |= !range_fits_quot6(u[q]);
                }
                push_halo<NV, double, true>(p, p.Wc, i, j, k, u, u, wide);
            }
            if (!p.quad4 || p.viscous) {
                double* Wo = (double*)p.W;
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    Wo[q * p.npad + pc] = w[q];
                    wide |= !range_fits_quot6(w[q]);
                }
                push_halo<NV, double>(p, Wo, i, j, k, u, w, wide);
            }
            if (p.viscous) local = smax(local, cell_gradients_f64<NV>(p, c, i, j, k, w));
        }
    }
    if (wide) p.st->wide = 1;   // a plain store of the same value from every lane that found one
    if (p.viscous) block_max_slot(p, local, SPD_VISC);
}

// k_nrbc: the non-reflecting ghosts of side SIDE's boundary cells (thread per cell, in the side's boundary order),
// from the stage input, before the passes that read them
template <int NV, int SIDE>
__global__ __launch_bounds__(256) void k_nrbc(KP p) {
    const long long n = side_cells(p, SIDE), kb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (kb >= n || p.st->done) return;
    int i, j, k;
    if (SIDE < 2) {
        i = SIDE == 0 ? 0 : p.nx - 1;
        j = (int)(kb % p.ny);
        k = (int)(kb / p.ny);
    } else if (SIDE < 4) {
        i = (int)(kb % p.nx);
        j = SIDE == 2 ? 0 : p.ny - 1;
        k = (int)(kb / p.nx);
    } else {
        i = (int)(kb % p.nx);
        j = (int)(kb / p.nx);
        k = SIDE == 4 ? 0 : p.nz - 1;
    }
    double g[NV];
    nrbc_ghost<NV, SIDE>(p, i, j, k, g);
    double* out = p.nrbc + p.nrbc_off[SIDE];
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q * n + kb] = g[q];
}

// ------------------------------------------------------------------------------------------------------------
// k_hancock: the limited states of every cell advanced by half a step with the cell's own flux difference;
// an inadmissible state stays as it was (Schemes.cpp:2135-2199, 2907-2933, 2989-3016).
// ------------------------------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void euler_flux_axis(const Eos& e, const double* U, int a, double* F) {
    double W[NV];
    c2p<NV>(e, U, W, (double*)nullptr);
    const double un = W[mom(a)], pr = W[3];
    F[0] = U[0] * un;
    F[1] = U[1] * un + (a == 0 ? pr : 0.0);
    F[2] = U[2] * un + (a == 1 ? pr : 0.0);
    F[3] = (U[3] + pr) * un;
    if constexpr (NV == 5) F[4] = U[4] * un + (a == 2 ? pr : 0.0);
}

template <int NV>
__global__ __launch_bounds__(256) void k_hancock(KP p) {
    constexpr int ND = NV - 2, NS = 2 * ND;
    const long long c = xcd_block() * blockDim.x + threadIdx.x;
    unsigned long long reverted = 0;
    if (c < p.nc && !p.st->done) {
        const int i = (int)(c % p.nx);
        const long long r = c / p.nx;
        const int j = (int)(r % p.ny), k = (int)(r / p.ny);
        double st[NS][NV];
        double w0[NV];
        load_prim<NV, double>(p, pidx(p, i, j, k), w0);
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const int t = coord(d, i, j, k), n = ncell(p, d);
            const bool up_lo = t > 0 || halo_of(p, 2 * d);          // the cell's low face is reconstructed
            const bool up_hi = t + 1 < n || halo_of(p, 2 * d + 1);  // its high face
            double wm[NV], wp[NV];
            stencil_prim<NV>(p, d, i, j, k, t - 1, wm);
            stencil_prim<NV>(p, d, i, j, k, t + 1, wp);
            double lo[NV], hi[NV], uc[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                lo[q] = muscl_wr(p.lim, wm[q], w0[q], wp[q]);
                hi[q] = muscl_wl(p.lim, wm[q], w0[q], wp[q]);
            }
            load_cons<NV>(p, c, uc);
            roundtrip<NV>(p.rs.eos, uc);   // the first-order state of a side face without a halo
            double a0[NV], a1[NV];
            p2c<NV>(p.rs.eos, lo, a0);
            p2c<NV>(p.rs.eos, hi, a1);
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                st[2 * d][q] = up_lo ? a0[q] : uc[q];
                st[2 * d + 1][q] = up_hi ? a1[q] : uc[q];
            }
        }
        const double half_dt = p.st->half_dt;
        const double hx = half_dt / p.dx, hy = half_dt / p.dy;
        double f[NS][NV];
#pragma unroll
        for (int s = 0; s < NS; ++s) euler_flux_axis<NV>(p.rs.eos, st[s], s / 2, f[s]);
        double dd[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            if constexpr (NV == 5) {
                const double hz = half_dt / p.dz;
                dd[q] = (hx * (f[1][q] - f[0][q]) + hy * (f[3][q] - f[2][q])) + hz * (f[5][q] - f[4][q]);
            } else {
                dd[q] = hx * (f[1][q] - f[0][q]) + hy * (f[3][q] - f[2][q]);
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double nw[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) nw[q] = st[s][q] - dd[q];
            const double u4 = NV == 5 ? nw[NV - 1] : 0.0;
            // admissible_state (Schemes.cpp:2145-2148)
            const bool ok = (nw[0] > kEosEps) && (0.5 * (nw[1] * nw[1] + nw[2] * nw[2] + u4 * u4) < nw[0] * (nw[3] - p.rs.eos.p_inf));
            reverted += ok ? 0 : 1;
#pragma unroll
            for (int q = 0; q < NV; ++q) p.HS[((long long)s * NV + q) * p.nc + c] = ok ? nw[q] : st[s][q];
        }
    }
    if (reverted) atomicAdd((unsigned long long*)&p.st->reverted, reverted);
}

// ------------------------------------------------------------------------------------------------------------
// k_faces<axis>: the faces normal to one axis, one thread per face: face states (reconstruction, first order,
// or the MUSCL-Hancock predicted states), the Riemann flux, and the viscous flux of the face.
// Schemes.cpp:389-700 (states), 842-911 (fused line), 231-240 (Riemann); Physics.cpp:100-176 (2-D),
// 567-707 (3-D) for the viscous flux.
// ------------------------------------------------------------------------------------------------------------
// Cross derivative d u_a / d x_b on a face normal to D: the average of the cell gradients of the two cells xL, xR
// (the opposite cell across a periodic seam), or on a physical boundary face with boundary_order = 2 the
// tangential difference of the face values, ghost and cell layer averaged (Physics.cpp:614-628).
template <int NV, int D, typename T>
__device__ __forceinline__ T cross_term(const KP& p, bool phys, int i, int j, int k, int tg, int tc, long long xL, long long xR, int a,
                                             int b) {
    constexpr int ND = NV - 2;
    if (phys) {
        double f[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int off = s == 0 ? 1 : -1;
            const int oi = b == 0 ? off : 0, oj = b == 1 ? off : 0, ok = b == 2 ? off : 0;
            double wg[NV], wc[NV];
            visc_pad_prim<NV>(p, (D == 0 ? tg : i) + oi, (D == 1 ? tg : j) + oj, (D == 2 ? tg : k) + ok, wg);
            visc_pad_prim<NV>(p, (D == 0 ? tc : i) + oi, (D == 1 ? tc : j) + oj, (D == 2 ? tc : k) + ok, wc);
            f[s] = 0.5 * (wg[mom(a)] + wc[mom(a)]);
        }
        return T((f[0] - f[1]) / (2.0 * hstep(p, b)));
    }
    const T* Gp = (const T*)p.G + (long long)(a * ND + b) * p.nc;
    return T(0.5) * (Gp[xL] + Gp[xR]);
}

template <int NV, int D, typename T>
__device__ __forceinline__ void viscous_face(const KP& p, int i, int j, int k, int t, int n, T* out) {
    constexpr int ND = NV - 2;
    const bool lo = t == 0, hi = t == n;
    const long long cL = cidx_d(p, D, i, j, k, lo ? 0 : t - 1), cR = cidx_d(p, D, i, j, k, hi ? n - 1 : t);
    T vL[ND], vR[ND], TL, TR;
    {
        T w[NV];
        if (lo) {
            double wd[NV];
            visc_pad_prim<NV>(p, D == 0 ? -1 : i, D == 1 ? -1 : j, D == 2 ? -1 : k, wd);
#pragma unroll
            for (int q = 0; q < NV; ++q) w[q] = T(wd[q]);
        } else {
            load_prim<NV, T>(p, pidx_d(p, D, i, j, k, t - 1), w);
        }
#pragma unroll
        for (int a = 0; a < ND; ++a) vL[a] = w[mom(a)];
        TL = dv(w[3], w[0]);
        if (hi) {
            double wd[NV];
            visc_pad_prim<NV>(p, D == 0 ? n : i, D == 1 ? n : j, D == 2 ? n : k, wd);
#pragma unroll
            for (int q = 0; q < NV; ++q) w[q] = T(wd[q]);
        } else {
            load_prim<NV, T>(p, pidx_d(p, D, i, j, k, t), w);
        }
#pragma unroll
        for (int a = 0; a < ND; ++a) vR[a] = w[mom(a)];
        TR = dv(w[3], w[0]);
    }
    const long long xL = (lo && vper_of(p, D)) ? cidx_d(p, D, i, j, k, n - 1) : cL;
    const long long xR = (hi && vper_of(p, D)) ? cidx_d(p, D, i, j, k, 0) : cR;
    const bool phys = p.bo2 && !vper_of(p, D) && (lo || hi);
    const T h = T(hstep(p, D));
    const int tg = lo ? -1 : n, tc = lo ? 0 : n - 1;   // the ghost and the cell layer of a boundary face
#define CROSS(a, b) cross_term<NV, D, T>(p, phys, i, j, k, tg, tc, xL, xR, a, b)
    const T mu = T(p.mu), kc = T(p.k_cond), MD = T(MU_DIAG), MO = T(MU_OFF), half = T(0.5);
    if constexpr (ND == 2) {
        const T u_f = half * (vL[0] + vR[0]), v_f = half * (vL[1] + vR[1]);
        if constexpr (D == 0) {
            const T du_dx = dv(vR[0] - vL[0], h), dv_dx = dv(vR[1] - vL[1], h), dT_dx = dv(TR - TL, h);
            const T du_dy = CROSS(0, 1), dv_dy = CROSS(1, 1);
            const T tau_xx = mu * (MD * du_dx - MO * dv_dy);
            const T tau_xy = mu * (du_dy + dv_dx);
            const T q_x = -kc * dT_dx;
            const T ar = T(p.dy);
            out[0] = tau_xx * ar;
            out[1] = tau_xy * ar;
            out[2] = (u_f * tau_xx + v_f * tau_xy - q_x) * ar;
        } else {
            const T dv_dy = dv(vR[1] - vL[1], h), du_dy = dv(vR[0] - vL[0], h), dT_dy = dv(TR - TL, h);
            const T du_dx = CROSS(0, 0), dv_dx = CROSS(1, 0);
            const T tau_yy = mu * (MD * dv_dy - MO * du_dx);
            const T tau_yx = mu * (du_dy + dv_dx);
            const T q_y = -kc * dT_dy;
            const T ar = T(p.dx);
            out[0] = tau_yx * ar;
            out[1] = tau_yy * ar;
            out[2] = (u_f * tau_yx + v_f * tau_yy - q_y) * ar;
        }
    } else {
        T Gn[3], vf[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Gn[a] = dv(vR[a] - vL[a], h);
            vf[a] = half * (vL[a] + vR[a]);
        }
        const T dT = dv(TR - TL, h);
        T tau[3];
        if constexpr (D == 0) {
            const T g01 = CROSS(0, 1), g11 = CROSS(1, 1), g22 = CROSS(2, 2), g02 = CROSS(0, 2);
            tau[0] = mu * (MD * Gn[0] - MO * ((T(0.0) + g11) + g22));
            tau[1] = mu * (Gn[1] + g01);
            tau[2] = mu * (Gn[2] + g02);
        } else if constexpr (D == 1) {
            const T g00 = CROSS(0, 0), g10 = CROSS(1, 0), g22 = CROSS(2, 2), g12 = CROSS(1, 2);
            tau[0] = mu * (Gn[0] + g10);
            tau[1] = mu * (MD * Gn[1] - MO * ((T(0.0) + g00) + g22));
            tau[2] = mu * (Gn[2] + g12);
        } else {
            const T g00 = CROSS(0, 0), g11 = CROSS(1, 1), g20 = CROSS(2, 0), g21 = CROSS(2, 1);
            tau[0] = mu * (Gn[0] + g20);
            tau[1] = mu * (Gn[1] + g21);
            tau[2] = mu * (MD * Gn[2] - MO * ((T(0.0) + g00) + g11));
        }
        const T q_d = -kc * dT;
        const T area = T(D == 0 ? p.Ax : (D == 1 ? p.Ay : p.Az));
        out[0] = tau[0] * area;
        out[1] = tau[1] * area;
        out[2] = (vf[0] * tau[0] + vf[1] * tau[1] + vf[2] * tau[2] - q_d) * area;
        out[3] = tau[2] * area;
    }
}

#undef CROSS

// The MUSCL-Hancock state outside a boundary face (side `side`, boundary cell (i, j, k) whose predicted state
// on the face is `in`): the seam copy of the cell across a periodic wrap, the reconstructed state of the ghost
// layers, or the first-order state; on a wall side the ghost of the predicted state (hancock_wall_ghosts).
template <int NV, int D>
__device__ __forceinline__ void hancock_outside(const KP& p, int side, int i, int j, int k, int t, const double* in, double* out) {
    const int n = ncell(p, D), hi = side & 1;
    if (wrap_of(p, D)) {
        const long long cc = cidx_d(p, D, i, j, k, hi ? 0 : n - 1);
        const int s = hi ? 2 * D : 2 * D + 1;
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = p.HS[((long long)s * NV + q) * p.nc + cc];
    } else if (halo_of(p, side)) {
        double w[4][NV], v[NV];
#pragma unroll
        for (int s = 0; s < 4; ++s) stencil_prim<NV>(p, D, i, j, k, t - 2 + s, w[s]);
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = hi ? muscl_wr(p.lim, w[1][q], w[2][q], w[3][q]) : muscl_wl(p.lim, w[0][q], w[1][q], w[2][q]);
        p2c<NV>(p.rs.eos, v, out);
    } else {
        outside_first_order<NV>(p, D, side, i, j, k, out);
        roundtrip<NV>(p.rs.eos, out);
    }
    const int bc = bc_of(p.bp, side);
    if (bc == 0 || bc == 1 || bc == 2) ghost_state<NV>(p.rs.eos, p.bp, side, in, out);
}

// double-precision boundary states into the flux arithmetic type

__device__ __forceinline__ double area_of(const KP& p, int d) { return d == 0 ? p.Ax : (d == 1 ? p.Ay : p.Az); }

// The fluxes of face f normal to D. FP64: the convective flux and (with viscosity) the viscous flux in their own
// planes, combined by k_update in the reference's order. FP32 fluxes: one combined plane set, convective flux
// times the face area minus the viscous flux, so the update reads each face once.
template <int NV, int D, typename T>
__device__ __forceinline__ void store_face(const KP& p, long long f, int i, int j, int k, int t, int n, const T* F) {
    T* Fc = (T*)p.Fc[D];
    const long long ps = p.nfa[D];
    if constexpr (std::is_same<T, float>::value) {
        const float A = float(area_of(p, D));
        float G[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) G[q] = F[q] * A;
        if (p.viscous) {
            float v[NV - 1];
            viscous_face<NV, D, float>(p, i, j, k, t, n, v);
#pragma unroll
            for (int q = 1; q < NV; ++q) G[q] -= v[q - 1];
        }
#pragma unroll
        for (int q = 0; q < NV; ++q) Fc[q * ps + f] = G[q];
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) Fc[q * ps + f] = F[q];
        if (p.viscous) {
            T v[NV - 1];
            viscous_face<NV, D, T>(p, i, j, k, t, n, v);
            T* Fv = (T*)p.Fv[D];
#pragma unroll
            for (int q = 0; q < NV - 1; ++q) Fv[q * ps + f] = v[q];
        }
    }
}

// The two states of face (i, j, k) normal to D (t its coordinate along D, n the cells along D) that the Riemann
// solver takes: reconstructed from the padded primitives, first-order at a side without a halo, or the
// MUSCL-Hancock predicted states (Schemes.cpp:389-749, 2934-2988). Q6: WENO5-Z divides by 6 with quot6 (every
// stencil value in range_fits_quot6; the same bits).
template <int NV, int D, int MODE, typename T, bool Q6 = false>
__device__ __forceinline__ void face_states(const KP& p, int i, int j, int k, int t, int n, T* UL, T* UR) {
    const bool lo = t == 0, hi = t == n;
    // the cells left and right of the face (their coordinate along D is t - 1 and t)
    const int il = D == 0 ? t - 1 : i, jl = D == 1 ? t - 1 : j, kl = D == 2 ? t - 1 : k;
    if constexpr (MODE == MODE_FIRST_ORDER) {
        double ul[NV], ur[NV];
        if (!lo) load_cons<NV>(p, cidx(p, il, jl, kl), ul);
        else outside_first_order<NV>(p, D, 2 * D, i, j, k, ul);
        if (!hi) load_cons<NV>(p, cidx(p, i, j, k), ur);
        else outside_first_order<NV>(p, D, 2 * D + 1, il, jl, kl, ur);
        to_t<NV>(ul, UL);
        to_t<NV>(ur, UR);
    } else if constexpr (MODE == MODE_PLAIN) {
        const bool upgrade = (t > 0 && t < n) || (lo && halo_of(p, 2 * D)) || (hi && halo_of(p, 2 * D + 1));
        if (upgrade) {
            T wl[NV], wr[NV];
            const T we = T(p.weno_eps);
            if (p.recon == 3) {   // WENO5-Z on six cells t-3 .. t+2
                T c[6][NV];
#pragma unroll
                for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    wl[q] = weno5z<T, Q6>(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], we);
                    wr[q] = weno5z<T, Q6>(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], we);
                }
            } else {   // four cells t-2 .. t+1
                T c[4][NV];
#pragma unroll
                for (int s = 0; s < 4; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 2 + s, c[s]);
                if (p.recon == 2) {
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        wl[q] = weno3z(c[0][q], c[1][q], c[2][q], we);
                        wr[q] = weno3z(c[3][q], c[2][q], c[1][q], we);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        wl[q] = muscl_wl(p.lim, c[0][q], c[1][q], c[2][q]);
                        wr[q] = muscl_wr(p.lim, c[1][q], c[2][q], c[3][q]);
                    }
                }
            }
            p2c<NV>(p.rs.eos, wl, UL);
            p2c<NV>(p.rs.eos, wr, UR);
        } else {
            double ul[NV], ur[NV];
            if (!lo) load_cons<NV>(p, cidx(p, il, jl, kl), ul);
            else outside_first_order<NV>(p, D, 2 * D, i, j, k, ul);
            if (!hi) load_cons<NV>(p, cidx(p, i, j, k), ur);
            else outside_first_order<NV>(p, D, 2 * D + 1, il, jl, kl, ur);
            roundtrip<NV>(p.rs.eos, ul);
            roundtrip<NV>(p.rs.eos, ur);
            to_t<NV>(ul, UL);
            to_t<NV>(ur, UR);
        }
    } else {   // MUSCL-Hancock: the predicted states of the cells on both sides (Schemes.cpp:2934-2988)
        if (!lo) {
            const long long cc = cidx(p, il, jl, kl);
#pragma unroll
            for (int q = 0; q < NV; ++q) UL[q] = p.HS[((long long)(2 * D + 1) * NV + q) * p.nc + cc];
        }
        if (!hi) {
            const long long cc = cidx(p, i, j, k);
#pragma unroll
            for (int q = 0; q < NV; ++q) UR[q] = p.HS[((long long)(2 * D) * NV + q) * p.nc + cc];
        }
        if (lo) hancock_outside<NV, D>(p, 2 * D, i, j, k, t, UR, UL);
        if (hi) hancock_outside<NV, D>(p, 2 * D + 1, il, jl, kl, t, UL, UR);
    }
}

// (i, j, k) and the coordinate t along D of face f (< 2^31) normal to D, in that axis' face order
template <int NV, int D>
__device__ __forceinline__ void face_coords(const KP& p, long long f, int& i, int& j, int& k, int& t) {
    const unsigned fu = (unsigned)f, ni = p.nx + (D == 0), nj = p.ny + (D == 1);
    const unsigned r = mdiv(fu, p.mg_fi[D]), kk = mdiv(r, p.mg_fj[D]);
    i = (int)(fu - r * ni);
    j = (int)(r - kk * nj);
    k = (int)kk;
    t = coord(D, i, j, k);
}

template <int NV, int D, int MODE, typename T = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_faces(KP p) {
    const int ni = p.nx + (D == 0), nj = p.ny + (D == 1), nk = p.nz + (D == 2);
    const long long nf = (long long)ni * nj * nk;
    // z faces: row blocks (rows (j, k) of nx faces, k up to nz); x and y faces: one face per thread in face order
    const RowBlock rb = D == 2 ? row_block(p) : RowBlock{0, 0, 0};
    const long long f = D == 2 ? ((long long)rb.k * p.ny + rb.j) * p.nx + rb.i : xcd_block() * blockDim.x + threadIdx.x;
    double local = 0.0;
    if ((D == 2 ? rb.i < p.nx : f < nf) && !p.st->done) {
        int i = rb.i, j = rb.j, k = rb.k, t = rb.k;
        if (D != 2) face_coords<NV, D>(p, f, i, j, k, t);
        const int n = ncell(p, D);
        T UL[NV], UR[NV];
        if constexpr (MODE == MODE_PLAIN && std::is_same<T, double>::value) {
            // uniform over the launch: quot6 unless k_cells stored a primitive outside its range
            if (__builtin_amdgcn_readfirstlane(p.st->wide) == 0) face_states<NV, D, MODE, T, true>(p, i, j, k, t, n, UL, UR);
            else face_states<NV, D, MODE, T, false>(p, i, j, k, t, n, UL, UR);
        } else {
            face_states<NV, D, MODE, T>(p, i, j, k, t, n, UL, UR);
        }
        T F[NV], sm;
        face_flux<NV>(p.rs, UL, UR, D, F, sm);
        if (double(sm) > local) local = double(sm);
        store_face<NV, D, T>(p, f, i, j, k, t, n, F);
    }
    block_max_slot(p, local, SPD_CONV);
}

// ------------------------------------------------------------------------------------------------------------
// Learned troubled-face indicator (WENO5-Z; fvg_learned.hpp), in either flux precision T. Faces the network calls
// smooth take the linear upwind-5 value; the others, and every face once the indicator is off, take WENO5-Z (in T:
// FP64 WENO5-Z with FP64 fluxes, the reference's bits). Faces without a full stencil keep their first-order states as
// in k_faces. The network itself always runs in bfloat16 / float on the features of the float-rounded stencil.
// ------------------------------------------------------------------------------------------------------------
template <int NV, int D>
__device__ __forceinline__ bool face_upgrade(const KP& p, int t) {
    const int n = ncell(p, D);
    return (t > 0 && t < n) || (t == 0 && halo_of(p, 2 * D)) || (t == n && halo_of(p, 2 * D + 1));
}

// States and flux of face (i, j, k) normal to D through the learned indicator. Every lane of the wavefront calls
// it (the network runs on the matrix cores); lanes with active = false only take part in the network.
// USE: the network decides and its decision is kept for the step's later stages (LF_STORE, the first Runge-Kutta
// stage), or the kept decision of face f is used (LF_KEPT, stages 2 and 3, also in the fused last-axis kernel)
template <int NV, int D> __device__ __forceinline__ unsigned char* kept_of(const KP& p) {
    return D == 0 ? p.ind_kept[0] : (D == 1 ? p.ind_kept[1] : p.ind_kept[2]);
}
// Q6: the troubled faces' WENO5-Z divides by 6 with quot6 (every stencil value in range_fits_quot6; the same bits)
template <int NV, int D, int USE, typename T, bool Q6 = false>
__device__ __forceinline__ void learned_face(const KP& p, bool active, long long f, int i, int j, int k, int t, int n, T* F,
                                             T& sm) {
    const bool upgrade = active && face_upgrade<NV, D>(p, t);
    // once the audit has switched the indicator off every face takes WENO5-Z and the network is not run (the switch
    // changes only between steps; read once per wavefront, so the branch is uniform)
    const bool off = __builtin_amdgcn_readfirstlane(p.ind->off) != 0;
    T c[6][NV];
    bool troubled = true;
    if constexpr (USE == LF_KEPT) {
        if (upgrade) {
#pragma unroll
            for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
            troubled = off || kept_of<NV, D>(p)[f] != 0;
        }
    } else if (off) {   // LF_STORE, indicator off
        if (upgrade)
#pragma unroll
            for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
    } else {   // LF_STORE
        float x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = 0.0f;
        if (upgrade) {   // the stencil for the features only: it is loaded again below, so that it is not held in
                         // registers across the network (with FP64 stencils that spilled)
            T cf[6][NV];
#pragma unroll
            for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, cf[s]);
            face_features<NV, T>(cf, float(p.weno_eps), x);
        }
        const float z = mlp_logit_wave(p.mlp, x);   // every lane of the wavefront takes part
        troubled = z > p.mlp->thr;
        if (upgrade) {
            kept_of<NV, D>(p)[f] = troubled ? 1 : 0;
#pragma unroll
            for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
        }
    }
    if (!active) return;
    const bool lo = t == 0, hi = t == n;
    const int il = D == 0 ? t - 1 : i, jl = D == 1 ? t - 1 : j, kl = D == 2 ? t - 1 : k;
    T UL[NV], UR[NV];
    if (upgrade) {
        const bool smooth = !troubled && !off;
        T wl[NV], wr[NV];
        const T we = T(p.weno_eps);
        if (smooth) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                wl[q] = upwind5(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q]);
                wr[q] = upwind5(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                wl[q] = weno5z<T, Q6>(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], we);
                wr[q] = weno5z<T, Q6>(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], we);
            }
        }
        p2c<NV>(p.rs.eos, wl, UL);
        p2c<NV>(p.rs.eos, wr, UR);
    } else {
        double ul[NV], ur[NV];
        if (!lo) load_cons<NV>(p, cidx(p, il, jl, kl), ul);
        else outside_first_order<NV>(p, D, 2 * D, i, j, k, ul);
        if (!hi) load_cons<NV>(p, cidx(p, i, j, k), ur);
        else outside_first_order<NV>(p, D, 2 * D + 1, il, jl, kl, ur);
        roundtrip<NV>(p.rs.eos, ul);
        roundtrip<NV>(p.rs.eos, ur);
        to_t<NV>(ul, UL);
        to_t<NV>(ur, UR);
    }
    face_flux<NV>(p.rs, UL, UR, D, F, sm);
}

template <int NV, int D, int USE, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_faces_learned(KP p) {
    if (p.st->done) return;   // uniform: the whole grid leaves
    const int ni = p.nx + (D == 0), nj = p.ny + (D == 1), nk = p.nz + (D == 2);
    const long long nf = (long long)ni * nj * nk;
    // z faces: row blocks (as k_faces)
    const RowBlock rb = D == 2 ? row_block(p) : RowBlock{0, 0, 0};
    const long long f = D == 2 ? ((long long)rb.k * p.ny + rb.j) * p.nx + rb.i : xcd_block() * blockDim.x + threadIdx.x;
    const bool active = D == 2 ? rb.i < p.nx : f < nf;
    int i = 0, j = 0, k = 0, t = 0;
    if (D == 2) {
        i = rb.i;
        j = rb.j;
        k = t = rb.k;
    } else if (active) {
        face_coords<NV, D>(p, f, i, j, k, t);
    }
    const int n = ncell(p, D);
    T F[NV], sm = T(0);
    // uniform over the launch: quot6 for FP64 unless k_cells stored a primitive outside its range
    if (std::is_same<T, double>::value && __builtin_amdgcn_readfirstlane(p.st->wide) == 0)
        learned_face<NV, D, USE, T, std::is_same<T, double>::value>(p, active, f, i, j, k, t, n, F, sm);
    else
        learned_face<NV, D, USE, T>(p, active, f, i, j, k, t, n, F, sm);
    double local = 0.0;
    if (active) {
        if (double(sm) > local) local = double(sm);
        store_face<NV, D, T>(p, f, i, j, k, t, n, F);
    }
    block_max_slot(p, local, SPD_CONV);
}

// Audit of the indicator on the stage input of an audit step: the network's decision against the label (the label
// value of the configured source above ind_label) on every face with a full stencil; counts TP, FP, FN, TN.
template <int NV, int D, typename T>
__global__ __launch_bounds__(256) void k_audit(KP p) {
    if (p.st->done || p.ind_audit <= 0 || (p.ind->steps % p.ind_audit) != 0) return;   // uniform
    const int ni = p.nx + (D == 0), nj = p.ny + (D == 1), nk = p.nz + (D == 2);
    const long long nf = (long long)ni * nj * nk;
    const long long f = xcd_block() * blockDim.x + threadIdx.x;
    int i = 0, j = 0, k = 0, t = 0;
    if (f < nf) face_coords<NV, D>(p, f, i, j, k, t);
    const bool upgrade = f < nf && face_upgrade<NV, D>(p, t);
    T c[6][NV];
    float x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = 0.0f;
    bool label = false;
    if (upgrade) {
#pragma unroll
        for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
        face_features<NV, T>(c, float(p.weno_eps), x);
        label = face_label<NV, T>(c, T(p.weno_eps), p.ind_src, p.lim) > T(p.ind_label);
    }
    const float z = mlp_logit_wave(p.mlp, x);
    const bool pred = z > p.mlp->thr;
    // counts of the wavefront (ballots), of the block (LDS), then one atomic per block and class into the block's
    // slot row
    __shared__ unsigned cnt[4];
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long bal_tp = __ballot(upgrade && label && pred), bal_fp = __ballot(upgrade && !label && pred),
                             bal_fn = __ballot(upgrade && label && !pred), bal_tn = __ballot(upgrade && !label && !pred);
    if ((threadIdx.x & 63) == 0) {
        if (bal_tp) atomicAdd(&cnt[0], (unsigned)__popcll(bal_tp));
        if (bal_fp) atomicAdd(&cnt[1], (unsigned)__popcll(bal_fp));
        if (bal_fn) atomicAdd(&cnt[2], (unsigned)__popcll(bal_fn));
        if (bal_tn) atomicAdd(&cnt[3], (unsigned)__popcll(bal_tn));
    }
    __syncthreads();
    if (threadIdx.x < 4 && cnt[threadIdx.x])
        atomicAdd(&p.ind->part[blockIdx.x % kAuditSlots][threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// Training samples of one axis: the features (16 floats per face, the 16th = 1) and the label value of every face
// of the current primitives; faces without a full stencil get label -1.
template <int NV, int D, typename T>
__global__ __launch_bounds__(256) void k_samples(KP p, float* __restrict__ X, float* __restrict__ dep) {
    const int ni = p.nx + (D == 0), nj = p.ny + (D == 1), nk = p.nz + (D == 2);
    const long long nf = (long long)ni * nj * nk;
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    int i, j, k, t;
    face_coords<NV, D>(p, f, i, j, k, t);
    float x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = 0.0f;
    float d = -1.0f;
    if (face_upgrade<NV, D>(p, t)) {
        T c[6][NV];
#pragma unroll
        for (int s = 0; s < 6; ++s) stencil_prim<NV, T>(p, D, i, j, k, t - 3 + s, c[s]);
        face_features<NV, T>(c, float(p.weno_eps), x);
        d = float(face_label<NV, T>(c, T(p.weno_eps), p.ind_src, p.lim));
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) X[f * 16 + q] = x[q];
    dep[f] = d;
}

// ------------------------------------------------------------------------------------------------------------
// k_update: divergence of the face fluxes (Simulation.cpp:1022-1037, 1293-1310), the viscous balance added
// (Physics.cpp:196-204, 746-752) and the stage combination of the time scheme (TimeScheme.cpp:27-59, 150-156,
// 213-219), one cell per thread.
// ------------------------------------------------------------------------------------------------------------
// The stage combination of the time scheme (TimeScheme.cpp:27-59, 150-156, 213-219) for cell c.
template <int NV, int UPD>
__device__ __forceinline__ void combine(const KP& p, long long c, const double* R, const double* __restrict__ U0,
                                        const double* __restrict__ Ut, double* __restrict__ out) {
    const double dt = p.st->dt;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const long long x = q * p.nc + c;
        double o;
        if constexpr (UPD == UPD_RESIDUAL) o = R[q];
        else if constexpr (UPD == UPD_AXPY) o = U0[x] + dt * R[q];
        else if constexpr (UPD == UPD_RK2_2) o = 0.5 * U0[x] + 0.5 * (Ut[x] + dt * R[q]);
        else if constexpr (UPD == UPD_RK3_2) o = 0.75 * U0[x] + 0.25 * (Ut[x] + dt * R[q]);
        else o = (1.0 / 3.0) * U0[x] + (2.0 / 3.0) * (Ut[x] + dt * R[q]);
        out[x] = o;
    }
}

template <int NV, int UPD, typename T = double>
__global__ __launch_bounds__(256) void k_update(KP p, const double* __restrict__ U0, const double* __restrict__ Ut, double* __restrict__ out) {
    const RowBlock rb = row_block(p);
    const int i = rb.i, j = rb.j, k = rb.k;
    // the residual's face kernels have read the range flag of k_cells: clear it for the next residual
    if (blockIdx.x == 0 && threadIdx.x == 0) p.st->wide = 0;
    if (i >= p.nx || p.st->done) return;
    const long long c = ((long long)k * p.ny + j) * p.nx + i;
    const long long fx = ((long long)k * p.ny + j) * (p.nx + 1) + i;
    const long long fy = ((long long)k * (p.ny + 1) + j) * p.nx + i;
    const long long fz = c;
    const long long fyt = fy + p.nx, fzk = fz + (long long)p.nx * p.ny;
    double R[NV];
    if constexpr (std::is_same<T, float>::value) {
        // combined face fluxes (store_face); the differences of two floats are exact in double, so the balance
        // still telescopes to FP64 round-off
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const float* Fx = (const float*)p.Fc[0] + q * p.nfa[0];
            const float* Fy = (const float*)p.Fc[1] + q * p.nfa[1];
            double d = (double(Fx[fx + 1]) - double(Fx[fx])) + (double(Fy[fyt]) - double(Fy[fy]));
            if constexpr (NV == 5) {
                const float* Fz = (const float*)p.Fc[2] + q * p.nfa[2];
                d += double(Fz[fzk]) - double(Fz[fz]);
            }
            R[q] = -d * p.inv_vol;
        }
    } else {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const T* Fx = (const T*)p.Fc[0] + q * p.nfa[0];
        const T* Fy = (const T*)p.Fc[1] + q * p.nfa[1];
        const double fl = Fx[fx], fr = Fx[fx + 1], fb = Fy[fy], ft = Fy[fyt];
        if constexpr (NV == 5) {
            const T* Fz = (const T*)p.Fc[2] + q * p.nfa[2];
            const double ff = Fz[fz], fk = Fz[fzk];
            R[q] = (((0.0 - (fr * p.Ax - fl * p.Ax)) - (ft * p.Ay - fb * p.Ay)) - (fk * p.Az - ff * p.Az)) * p.inv_vol;
        } else {
            R[q] = ((0.0 - (fr * p.Ax - fl * p.Ax)) - (ft * p.Ay - fb * p.Ay)) * p.inv_vol;
        }
    }
    if (p.viscous) {
        R[0] = R[0] + 0.0;   // the zero viscous flux of the density slot
#pragma unroll
        for (int q = 1; q < NV; ++q) {
            const T* Fx = (const T*)p.Fv[0] + (q - 1) * p.nfa[0];
            const T* Fy = (const T*)p.Fv[1] + (q - 1) * p.nfa[1];
            const double fl = Fx[fx], fr = Fx[fx + 1], fb = Fy[fy], ft = Fy[fyt];
            double v;
            if constexpr (NV == 5) {
                const T* Fz = (const T*)p.Fv[2] + (q - 1) * p.nfa[2];
                v = (((0.0 + (fr - fl)) + (ft - fb)) + (double(Fz[fzk]) - double(Fz[fz]))) * p.inv_vol;
            } else {
                v = ((0.0 + (fr - fl)) + (ft - fb)) * p.inv_vol;
            }
            R[q] = R[q] + v;
        }
    }
    }
    combine<NV, UPD>(p, c, R, U0, Ut, out);
}

// ------------------------------------------------------------------------------------------------------------
// Fused residual passes (the default kernel sequence): the face fluxes never go to HBM. Every face is computed
// once (up to one face per line segment or wavefront) and its difference with the next face of the same axis is
// added at once to the cell between them, in the reference's order of the divergence (Simulation.cpp:1022-1037,
// 1293-1310; viscous balance Physics.cpp:196-204, 746-752):
//   FP64: a = ((0 - (fr Ax - fl Ax)) - (ft Ay - fb Ay)) - (fk Az - ff Az), R = a / V;
//         v = ((0 + (vr - vl)) + (vt - vb)) + (vk - vf), R = R + v / V (R[0] + 0);
//   FP32 fluxes (combined F A - v): d = ((fr - fl) + (ft - fb)) + (fk - ff) in double, R = -d / V.
// The partial sums after the first axes live in the accumulator planes A (FP64: NV convective, NV - 1 viscous);
// the last axis' pass finishes R and either stores it (UPD_RESIDUAL) or applies the stage combination at once
// (UPD_RK2_2, UPD_RK3_2, UPD_RK3_3: dt is known). The stage output never aliases the stage input, so no face reads
// a cell another thread has already updated.
//   x: k_xacc, one face per lane in face order; a wavefront holds 64 consecutive faces and updates the 63 cells
//      between them (neighbour face by lane shuffle), so wavefronts overlap by one face;
//   y (3-D), then the last axis: k_lacc, one lane per line segment of `seg` cells along the axis (64 lanes on
//      consecutive x), looping over its seg + 1 faces and keeping the last face's fluxes.
// ------------------------------------------------------------------------------------------------------------

// The fluxes of face (i, j, k) normal to D (t its coordinate along D) as the separate face pass stores them
// (store_face): FP64 the convective flux Fc and, with viscosity, the viscous flux Fv; FP32 the combined flux in Fc.
// Every lane of a wavefront calls it in the learned mode (the network runs on the matrix cores); lanes with
// active = false only take part in the network. The wave speed of the face goes into local.
template <int NV, int D, int MODE, int USE, typename T>
__device__ __forceinline__ void face_all(const KP& p, bool active, long long f, int i, int j, int k, int t, T* Fc, T* Fv, double& local) {
    const int n = ncell(p, D);
    T sm = T(0);
    if constexpr (MODE == MODE_LEARNED) {
        learned_face<NV, D, USE, T>(p, active, f, i, j, k, t, n, Fc, sm);
    } else {
        if (!active) return;
        T UL[NV], UR[NV];
        if constexpr (MODE == MODE_PLAIN && std::is_same<T, double>::value) {
            if (__builtin_amdgcn_readfirstlane(p.st->wide) == 0) face_states<NV, D, MODE, T, true>(p, i, j, k, t, n, UL, UR);
            else face_states<NV, D, MODE, T, false>(p, i, j, k, t, n, UL, UR);
        } else {
            face_states<NV, D, MODE, T>(p, i, j, k, t, n, UL, UR);
        }
        face_flux<NV>(p.rs, UL, UR, D, Fc, sm);
    }
    if (!active) return;
    if (double(sm) > local) local = double(sm);
    if (p.viscous) viscous_face<NV, D, T>(p, i, j, k, t, n, Fv);
    if constexpr (std::is_same<T, float>::value) {   // the combined flux of store_face
        const float A = float(area_of(p, D));
#pragma unroll
        for (int q = 0; q < NV; ++q) Fc[q] = Fc[q] * A;
        if (p.viscous)
#pragma unroll
            for (int q = 1; q < NV; ++q) Fc[q] -= Fv[q - 1];
    }
}

template <int NV, int MODE, int USE, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_xacc(KP p) {
    if (p.st->done) return;   // uniform: the whole grid leaves
    const long long nf = (long long)(p.nx + 1) * p.ny * p.nz;
    const int lane = threadIdx.x & 63;
    const long long f = (xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 63 + lane;
    const bool active = f < nf;
    int i = 0, j = 0, k = 0, t = 0;
    if (active) face_coords<NV, 0>(p, f, i, j, k, t);
    T Fc[NV], Fv[NV - 1];
#pragma unroll
    for (int q = 0; q < NV; ++q) Fc[q] = T(0);
#pragma unroll
    for (int q = 0; q < NV - 1; ++q) Fv[q] = T(0);
    double local = 0.0;
    face_all<NV, 0, MODE, USE, T>(p, active, f, i, j, k, t, Fc, Fv, local);
    // the next face along x is the next lane's (every lane takes part in the exchange)
    T Gc[NV], Gv[NV - 1];
#pragma unroll
    for (int q = 0; q < NV; ++q) Gc[q] = __shfl_down(Fc[q], 1, 64);
    constexpr bool VISC64 = std::is_same<T, double>::value;
    if (VISC64 && p.viscous)
#pragma unroll
        for (int q = 0; q < NV - 1; ++q) Gv[q] = __shfl_down(Fv[q], 1, 64);
    if (active && lane < 63 && i < p.nx) {   // the cell right of this face (its right face is face f + 1)
        const long long c = cidx(p, i, j, k);
        double* A = p.A;
        if constexpr (VISC64) {
#pragma unroll
            for (int q = 0; q < NV; ++q) A[q * p.nc + c] = 0.0 - (double(Gc[q]) * p.Ax - double(Fc[q]) * p.Ax);
            if (p.viscous)
#pragma unroll
                for (int q = 0; q < NV - 1; ++q) A[(NV + q) * p.nc + c] = 0.0 + (double(Gv[q]) - double(Fv[q]));
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) A[q * p.nc + c] = double(Gc[q]) - double(Fc[q]);
        }
    }
    block_max_slot(p, local, SPD_CONV);
}

// y faces (3-D, D = 1, UPD = -1: add to A) or the last axis (z in 3-D, y in 2-D; UPD: finish R, then store it or
// combine). A block is 4 wavefronts on 64 consecutive x positions of one line segment of seg cells along D; the
// wavefronts take 4 consecutive faces of the segment at a time (wavefront w face tb + w), so that they share their
// stencil rows in the caches, and step along the segment together. The face below a wavefront's cell is the
// previous wavefront's (wavefront 0: the last wavefront's of the previous step), passed through LDS between two
// barriers. Block b of the grid: 64-lane chunk along x, segment, and the other transverse coordinate o (z for D = 1,
// y for D = 2), in the order (o, segment, chunk) over the XCDs' contiguous ranges.
template <int NV, int D, int MODE, int USE, int UPD, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4))) void k_lacc(KP p, const double* __restrict__ U0,
                                                                                                        const double* __restrict__ Ut,
                                                                                                        double* __restrict__ out) {
    if (p.st->done) return;   // uniform
    const unsigned b = (unsigned)xcd_block();
    const unsigned r = mdiv(b, p.mg_lchunk), chunk = b - r * p.nlchunk;
    const unsigned o = mdiv(r, p.mg_nseg[D]), sidx = r - o * p.nseg[D];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = chunk * 64 + lane;
    const bool lane_ok = i < p.nx;
    const int n = ncell(p, D);
    const int t0 = (int)sidx * p.seg, t1 = t0 + p.seg < n ? t0 + p.seg : n;   // cells t0 .. t1 - 1, faces t0 .. t1
    constexpr bool VISC64 = std::is_same<T, double>::value;
    constexpr int NA = VISC64 ? 2 * NV - 1 : NV;
    const double area = D == 1 ? p.Ay : p.Az;
    double local = 0.0;
    __shared__ T sF[4][NA][64];   // the fluxes of each wavefront's face of the step
    for (int tb = t0; tb <= t1; tb += 4) {   // uniform over the block
        // the parameters (read in place in the kernel-argument segment, where p is the first argument) and the
        // coordinates pass through an empty asm each step, so that the compiler re-reads them instead of hoisting
        // the face code's address arithmetic out of the loop (which spills)
        typedef const __attribute__((address_space(4))) KP* KPconst;
        KPconst pp = (KPconst)__builtin_amdgcn_kernarg_segment_ptr();
        int ii = i;
        asm volatile("" : "+s"(pp));
        asm volatile("" : "+v"(ii));
        const KP& q = *(const KP*)pp;
        const int t = tb + w;   // this wavefront's face
        const bool face = lane_ok && t <= t1;
        const int jj = D == 1 ? t : (int)o, kk = D == 2 ? t : (D == 1 ? (int)o : 0);
        const long long f = D == 1 ? ((long long)kk * (p.ny + 1) + jj) * p.nx + i : ((long long)kk * p.ny + jj) * p.nx + i;
        // the cell this wavefront finishes (between face t - 1 below and face t above): its partial sums and, for a
        // stage combination, its stage states are loaded before the face, so that their latency hides behind it
        const bool cell = face && t > t0;
        const long long c = cidx(p, i, D == 1 ? t - 1 : jj, D == 2 ? t - 1 : kk);
        const bool visc_acc = VISC64 && p.viscous;
        double a[NA], u0[NV], ut[NV];
        if (cell) {
#pragma unroll
            for (int m = 0; m < NA; ++m) a[m] = (m < NV || visc_acc) ? p.A[m * p.nc + c] : 0.0;
            if constexpr (UPD == UPD_RK2_2 || UPD == UPD_RK3_2 || UPD == UPD_RK3_3)
#pragma unroll
                for (int m = 0; m < NV; ++m) {
                    u0[m] = U0[m * p.nc + c];
                    ut[m] = Ut[m * p.nc + c];
                }
        }
        T Fc[NV], Fv[NV - 1];
#pragma unroll
        for (int m = 0; m < NV; ++m) Fc[m] = T(0);
#pragma unroll
        for (int m = 0; m < NV - 1; ++m) Fv[m] = T(0);
        face_all<NV, D, MODE, USE, T>(q, face, f, ii, jj, kk, t, Fc, Fv, local);   // every lane of the wavefront
        if (w < 3) {
#pragma unroll
            for (int m = 0; m < NV; ++m) sF[w][m][lane] = Fc[m];
            if constexpr (VISC64)
#pragma unroll
                for (int m = 0; m < NV - 1; ++m) sF[w][NV + m][lane] = Fv[m];
        }
        __syncthreads();
        if (cell) {
            const int pw = w > 0 ? w - 1 : 3;   // the wavefront that holds face t - 1
            double* A = p.A;
            if constexpr (UPD < 0) {   // y faces of a 3-D mesh: add to the partial sums
                if constexpr (VISC64) {
#pragma unroll
                    for (int m = 0; m < NV; ++m) A[m * p.nc + c] = a[m] - (double(Fc[m]) * area - double(sF[pw][m][lane]) * area);
                    if (p.viscous)
#pragma unroll
                        for (int m = 0; m < NV - 1; ++m) A[(NV + m) * p.nc + c] = a[NV + m] + (double(Fv[m]) - double(sF[pw][NV + m][lane]));
                } else {
#pragma unroll
                    for (int m = 0; m < NV; ++m) A[m * p.nc + c] = a[m] + (double(Fc[m]) - double(sF[pw][m][lane]));
                }
            } else {   // the last axis: R of the cell, then the stage output
                double R[NV];
                if constexpr (VISC64) {
#pragma unroll
                    for (int m = 0; m < NV; ++m) R[m] = (a[m] - (double(Fc[m]) * area - double(sF[pw][m][lane]) * area)) * p.inv_vol;
                    if (p.viscous) {
                        R[0] = R[0] + 0.0;   // the zero viscous flux of the density slot
#pragma unroll
                        for (int m = 1; m < NV; ++m) {
                            const double v = (a[NV + m - 1] + (double(Fv[m - 1]) - double(sF[pw][NV + m - 1][lane]))) * p.inv_vol;
                            R[m] = R[m] + v;
                        }
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < NV; ++m) R[m] = -(a[m] + (double(Fc[m]) - double(sF[pw][m][lane]))) * p.inv_vol;
                }
                if constexpr (UPD == UPD_RESIDUAL) {
#pragma unroll
                    for (int m = 0; m < NV; ++m) out[m * p.nc + c] = R[m];
                } else {
                    const double dt = p.st->dt;
#pragma unroll
                    for (int m = 0; m < NV; ++m) {
                        double v;
                        if constexpr (UPD == UPD_RK2_2) v = 0.5 * u0[m] + 0.5 * (ut[m] + dt * R[m]);
                        else if constexpr (UPD == UPD_RK3_2) v = 0.75 * u0[m] + 0.25 * (ut[m] + dt * R[m]);
                        else v = (1.0 / 3.0) * u0[m] + (2.0 / 3.0) * (ut[m] + dt * R[m]);
                        out[m * p.nc + c] = v;
                    }
                }
            }
        }
        __syncthreads();
        if (w == 3) {   // the face below wavefront 0's cell of the next step
#pragma unroll
            for (int m = 0; m < NV; ++m) sF[3][m][lane] = Fc[m];
            if constexpr (VISC64)
#pragma unroll
                for (int m = 0; m < NV - 1; ++m) sF[3][NV + m][lane] = Fv[m];
        }
    }
    block_max_slot(p, local, SPD_CONV);
}

// The first-stage update out = U0 + dt R of a residual the fused passes stored (TimeScheme.cpp:27-59), one cell per
// thread; out may be U0.
template <int NV>
__global__ __launch_bounds__(256) void k_axpy(KP p, const double* U0, const double* __restrict__ R, double* out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nc || p.st->done) return;
    const double dt = p.st->dt;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const long long x = q * p.nc + c;
        out[x] = U0[x] + dt * R[x];
    }
}

}   // namespace fvg
