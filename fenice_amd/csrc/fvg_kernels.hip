// Kernels of the uniform-grid residual and time step that are not templates (time-step control, indicator control,
// conserved totals, the Riemann batch), and the instantiations of the learned-indicator, MUSCL-Hancock and totals
// kernels. The FP64 and FP32 residual kernels are instantiated in fvg_kernels_f64.hip and fvg_kernels_f32.hip.
#include "fvg_kernels_impl.hpp"

namespace fvg {

// ------------------------------------------------------------------------------------------------------------
// k_control: the scalar part of the time scheme on the device (TimeScheme.cpp:110-319; the divergence check of
// SimulationInternal.hpp:259-267; the loop condition of SimulationRun.cpp:167).
// ------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double stage_speed(DevState* s, int viscous) {
    const double sc = __longlong_as_double((long long)s->s_conv);
    if (!viscous) return sc;
    const double sv = __longlong_as_double((long long)s->s_visc);
    return smax(sc, sv);
}
__device__ __forceinline__ double clamp_dt(const DevState* s, double dt) {
    dt = smin(dt, s->t_max - s->t);
    if (s->dt_max > 0.0) dt = smin(dt, s->dt_max);
    return dt;
}
__device__ __forceinline__ void end_step(DevState* s) {
    if (isfinite(s->s_step) && isfinite(s->dt) && s->t + s->dt > s->t) {
        s->t += s->dt;
        s->steps += 1;
    } else {
        s->diverged = 1;
        s->done = 1;
        s->steps += 1;
    }
}

template __global__ void k_nrbc<4, 0>(KP);
template __global__ void k_nrbc<4, 1>(KP);
template __global__ void k_nrbc<4, 2>(KP);
template __global__ void k_nrbc<4, 3>(KP);
template __global__ void k_nrbc<5, 0>(KP);
template __global__ void k_nrbc<5, 1>(KP);
template __global__ void k_nrbc<5, 2>(KP);
template __global__ void k_nrbc<5, 3>(KP);
template __global__ void k_nrbc<5, 4>(KP);
template __global__ void k_nrbc<5, 5>(KP);

// The mass-flow inlet velocity of every mass-flow side (apply_side, Boundaries.cpp:285-294): rho_in u_in over the
// mean density of the side's boundary cells, floored at 1e-5. One block of kMassflowThreads per side; the densities
// are summed in the reference's boundary order (Core.hpp:121-151) by one thread, a chunk of kMassflowChunk values
// loaded into LDS by the whole block at a time, so that the sum has the reference's bits. The result goes to
// DevState::ut, which the ghost states of the residual read.
constexpr int kMassflowThreads = 1024, kMassflowChunk = 4096;
__global__ __launch_bounds__(kMassflowThreads) void k_massflow(KP p) {
    __shared__ double rho[kMassflowChunk];
    const int side = blockIdx.x;
    if (bc_of(p.bp, side) != 4 || (side >= 4 && !p.is3d) || p.st->done) return;   // uniform over the block
    const long long nx = p.nx, ny = p.ny, nz = p.nz;
    const long long n = side < 2 ? ny * nz : (side < 4 ? nx * nz : nx * ny);
    double sum = 0.0;
    for (long long base = 0; base < n; base += kMassflowChunk) {
        for (int l = threadIdx.x; l < kMassflowChunk; l += blockDim.x) {
            const long long k = base + l;
            if (k >= n) break;
            long long c;
            switch (side) {
            case 0: c = k * nx; break;
            case 1: c = k * nx + nx - 1; break;
            case 2: c = (k / nx) * ny * nx + k % nx; break;
            case 3: c = ((k / nx) * ny + ny - 1) * nx + k % nx; break;
            case 4: c = k; break;
            default: c = (nz - 1) * ny * nx + k; break;
            }
            rho[l] = smax(p.U[c], kEosEps);   // the density of cons_to_prim
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = n - base < kMassflowChunk ? (int)(n - base) : kMassflowChunk;
#pragma unroll 16
            for (int l = 0; l < m; ++l) sum += rho[l];   // in order: the loads run ahead of the chain of adds
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rho_avg = smax(sum / (double)n, 1e-5);
        p.st->ut[side] = (p.bp.rho_in * p.bp.u_in) / rho_avg;
    }
}

// op: the CTL_* step; arg: the time scheme (CTL_STAGE1_DT, CTL_END_RK) or the stage number (CTL_STAGE_S).
// Launched with kSpeedSlots threads: each takes one slot of each speed maximum, clears it, and the wavefront's
// maxima become s_conv / s_visc; the rest runs on one thread.
__global__ void k_control(KP p, int op, int arg) {
    const int scheme = arg;
    DevState* s = p.st;
    // every control launch comes after a residual's face kernels or between steps, never between a k_cells and its
    // faces: clear the quotient range flag for the next k_cells (the fused passes have no k_update to clear it)
    if (threadIdx.x == 0) s->wide = 0;
    if (!s->done) {
        const int l = threadIdx.x;
        unsigned long long* cs = p.spd + (size_t)(SPD_CONV * kSpeedSlots + l) * kSlotStride;
        unsigned long long* vs = p.spd + (size_t)(SPD_VISC * kSpeedSlots + l) * kSlotStride;
        unsigned long long c = *cs, v = *vs;
        *cs = 0;
        *vs = 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {   // bit patterns of non-negative doubles order as integers
            const unsigned long long oc = __shfl_xor(c, o, 64), ov = __shfl_xor(v, o, 64);
            c = oc > c ? oc : c;
            v = ov > v ? ov : v;
        }
        __syncthreads();   // every lane has read s->done before thread 0 may set it
        if (l == 0) {
            s->s_conv = c;
            s->s_visc = v;
        }
    }
    if (threadIdx.x != 0) return;
    // MUSCL-Hancock steps chained in one launch: the end of a step and the start of the next (the same operations, in
    // the same order, as the separate CTL_END_HANCOCK, CTL_BEGIN_STEP and CTL_HANCOCK_DT launches)
    if (op == CTL_END_BEGIN_HANCOCK) {
        if (s->done) return;
        const double S = stage_speed(s, p.viscous);
        s->s_conv = 0;
        s->s_visc = 0;
        s->lagged = S;
        s->s_step = S;
        s->half_dt = 0.0;
        end_step(s);
        op = CTL_BEGIN_HANCOCK;
    }
    if (op == CTL_BEGIN_HANCOCK) {
        if (s->done) return;
        if (!(s->t < s->t_end) || (s->max_steps > 0 && s->steps >= s->max_steps)) {
            s->done = 1;
            return;
        }
        s->t_max = s->t_end < s->t_max ? s->t_end : s->t_max;
        s->s_conv = 0;
        s->s_visc = 0;
        s->dt = clamp_dt(s, p.cfl * p.Ldt / (s->lagged + kDtEps));
        s->half_dt = 0.5 * s->dt;
        return;
    }
    if (op == CTL_BEGIN_STEP) {
        if (s->done) return;
        if (!(s->t < s->t_end) || (s->max_steps > 0 && s->steps >= s->max_steps)) {
            s->done = 1;
            return;
        }
        s->t_max = s->t_end < s->t_max ? s->t_end : s->t_max;
        s->s_conv = 0;
        s->s_visc = 0;
        return;
    }
    if (s->done) return;
    const double S = stage_speed(s, p.viscous);
    s->s_conv = 0;
    s->s_visc = 0;
    switch (op) {
    case CTL_STAGE1_DT:
        s->s1 = S;
        s->dt = clamp_dt(s, p.cfl * p.Ldt / (S + kDtEps));
        if (scheme == 0) {   // explicit Euler: the step is one stage
            s->s_step = S;
        }
        break;
    case CTL_STAGE_S:
        if (arg == 2) s->s2 = S;
        else s->s3 = S;
        break;
    case CTL_END_RK:
        if (scheme == 1) s->s_step = smax(s->s1, s->s2);
        else if (scheme == 2) s->s_step = smax(s->s1, smax(s->s2, s->s3));
        end_step(s);
        break;
    case CTL_SEED: s->lagged = S; break;
    case CTL_HANCOCK_DT:
        s->dt = clamp_dt(s, p.cfl * p.Ldt / (s->lagged + kDtEps));
        s->half_dt = 0.5 * s->dt;
        break;
    case CTL_END_HANCOCK:
        s->lagged = S;
        s->s_step = S;
        s->half_dt = 0.0;
        end_step(s);
        break;
    case CTL_RESIDUAL_S: s->s_step = S; break;
    default: break;
    }
}

// Indicator bookkeeping (one thread). Before a step's first stage: mark the step as taken (unless the run has
// ended) and clear the audit counters of an audit step. After its last stage: add the counters to the totals,
// switch the indicator off when it missed too many troubled faces, and count the step.
__global__ void k_indicator_control(KP p, int op) {
    IndState* d = p.ind;
    const bool audit_step = p.ind_audit > 0 && (d->steps % p.ind_audit) == 0;
    if (op == CTL_AUDIT_BEGIN) {
        if (p.st->done) return;
        d->pending = 1;
        if (audit_step)
            for (int r = 0; r < kAuditSlots; ++r)
                for (int q = 0; q < 4; ++q) d->part[r][q] = 0;
        return;
    }
    if (!d->pending) return;
    d->pending = 0;
    if (audit_step) {
        for (int q = 0; q < 4; ++q) {
            unsigned long long s = 0;
            for (int r = 0; r < kAuditSlots; ++r) s += d->part[r][q];
            d->last[q] = s;
            d->total[q] += s;
        }
        d->audits += 1;
        const double troubled = double(d->last[0] + d->last[2]), missed = double(d->last[2]);
        const double faces = double(d->last[0] + d->last[1] + d->last[2] + d->last[3]);
        if (troubled > 0.0 && missed > p.ind_fn_limit * troubled && missed > p.ind_fn_floor * faces) d->off = 1;
    }
    d->steps += 1;
}

// ------------------------------------------------------------------------------------------------------------
// Conserved totals and residual norms: a fixed-order FP64 tree (per-block partial sums, then one block over the
// partials). WHAT 0: mass, kinetic energy and total energy of a state (volume_averages, SimulationRun.cpp:45-64);
// WHAT 1: the sums of squares of a residual per quantity, rho, momentum and energy (add_residual,
// SimulationRun.cpp:121-126).
// ------------------------------------------------------------------------------------------------------------
template <int NV, int WHAT>
__global__ __launch_bounds__(256) void k_totals_partial(const double* __restrict__ U, long long nc, double* __restrict__ part) {
    __shared__ double sh[3][256];
    const long long per = (nc + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * per, b1 = b0 + per < nc ? b0 + per : nc;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (long long c = b0 + threadIdx.x; c < b1; c += blockDim.x) {
        const double r = U[c], m1 = U[nc + c], m2 = U[2 * nc + c], E = U[3 * nc + c];
        const double m4 = NV == 5 ? U[(NV - 1) * nc + c] : 0.0;
        if constexpr (WHAT == 0) {
            a0 += r;
            a1 += 0.5 * (m1 * m1 + m2 * m2 + m4 * m4) / r;
            a2 += E;
        } else {
            a0 += r * r * 1.0;
            a1 += (m1 * m1 + m2 * m2 + m4 * m4) * 1.0;
            a2 += E * E * 1.0;
        }
    }
    sh[0][threadIdx.x] = a0;
    sh[1][threadIdx.x] = a1;
    sh[2][threadIdx.x] = a2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[3 * blockIdx.x + q] = sh[q][0];
}
// means (norm 0), or root mean squares (norm 1: residual_norms, SimulationRun.cpp:128-131, every weight 1)
__global__ void k_totals_final(const double* __restrict__ part, int nb, long long nc, int norm, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < nb; ++b)
        for (int q = 0; q < 3; ++q) s[q] += part[3 * b + q];
    for (int q = 0; q < 3; ++q) out[q] = norm ? (nc > 0 ? sqrt(s[q] / (double)nc) : 0.0) : s[q] / (double)nc;
}

// ------------------------------------------------------------------------------------------------------------
// Known-answer kernel: the Riemann flux of independent faces (AoS, 5 doubles per state)
// ------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_riemann_batch(RiemannParams rs, const double* __restrict__ UL, const double* __restrict__ UR,
                                                       const int* __restrict__ axis, long long n, double* __restrict__ F,
                                                       double* __restrict__ smax_out) {
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double ul[5], ur[5], fo[5], sm;
    for (int q = 0; q < 5; ++q) {
        ul[q] = UL[5 * f + q];
        ur[q] = UR[5 * f + q];
    }
    face_flux<5>(rs, ul, ur, axis[f], fo, sm);
    for (int q = 0; q < 5; ++q) F[5 * f + q] = fo[q];
    smax_out[f] = sm;
}

// ------------------------------------------------------------------------------------------------------------
// Explicit instantiations used by the host
// ------------------------------------------------------------------------------------------------------------
template __global__ void k_hancock<4>(KP);
template __global__ void k_hancock<5>(KP);
template __global__ void k_totals_partial<4, 0>(const double*, long long, double*);
template __global__ void k_totals_partial<5, 0>(const double*, long long, double*);
template __global__ void k_totals_partial<4, 1>(const double*, long long, double*);
template __global__ void k_totals_partial<5, 1>(const double*, long long, double*);

}   // namespace fvg
