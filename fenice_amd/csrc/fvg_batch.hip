// An ensemble of independent 1-D runs, or of 2-D runs (fvg_batch_create_2d; see k_ensemble2), in one persistent
// kernel (include/fvg.h, fvg_batch_*).
//
// Instance i is the reference's 2-D run on an N_i x 1 mesh (periodic y, one row): nx = N_i, ny = 1, the scheme
// and the x boundaries of the shared configuration. One workgroup owns one instance for its whole run: it loops
// over the steps and stages of the time scheme itself (Simulation::run_loop, SimulationRun.cpp:161-191;
// TimeScheme.cpp:110-245), with workgroup barriers between the passes of a stage and the stage's wave speed
// reduced in LDS, so the whole ensemble is one launch. Instances may differ in size, time step and step count.
//
// Exactness against the reference's N x 1 run. The y-momentum of such a run is +0 in every cell at every step:
// the ghost and face states of the x sides allowed here keep v = +-0, v enters the other components only through
// v * v or products with zero, and an update adds +-0 to +0. So only (rho, rho u, E) are stored, and v = +0 is
// put back before the same device arithmetic as the 2-D path (NV = 4) runs. The two y faces of a cell see the
// cell itself on both sides (periodic, one row, no halo: the first-order branch of the face pass), so their
// fluxes are equal and cancel exactly in the balance (x - (+0) = x), but their wave speed enters the stage's
// maximum: it is computed per cell from the same face_flux call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fvg_host.hpp"

namespace fvg {

struct BInst {
    long long c0, f0;        // first cell / first face of the instance in the planes
    int n;                   // cells
    int halo_lo, halo_hi;    // reconstruction halo of the x sides (0 none, 1 periodic wrap, 2 ghost slab)
    int pad;
    double Ax, inv_vol, Ldt;   // mesh scalars of the N x 1 mesh
};

struct BOut {
    double t, s_last, dt_last;
    int steps, diverged;
};

// A team: m workgroups that run one instance together, each over its own chunk of cells (rank r: cells
// [n r / m, n (r + 1) / m)), exchanging the primitives of the KH cells next to their chunk and their wave-speed
// maxima through a TeamSync in global memory, with two team barriers per stage.
constexpr int KH = 3;   // stencil reach beyond a chunk: the WENO5-Z faces of a chunk read 3 cells on each side
struct BTeam {
    int inst, wg0, m, pad;   // instance, first workgroup, workgroups
    long long xoff;          // 2-D instances: the team's exchange rows in BP2::xch, from this element
};
struct TeamSync {
    unsigned long long count;    // barrier arrivals so far; barrier k of the run completes at (k + 1) m
    int broken;                  // a member gave up waiting (bounded spin): every member leaves the instance
    int pad;
    double smax[16];             // per rank: the wave-speed maximum of its chunk (the barrier after the faces)
    double halo[16][2][KH * 3];  // per rank, low and high end of its chunk: (rho, u, p) of its KH end cells
};
constexpr int kMaxTeam = 16;

struct BP {
    RiemannParams rs;
    BcParams bp;
    int recon, lim, scheme, max_steps;   // scheme: 0 explicit Euler, 1 SSP-RK2, 2 SSP-RK3
    double weno_eps, cfl, t_end, dt_max;
    long long nc, nf;   // plane strides: cells and faces of the whole ensemble
    double* U;          // 3 planes (rho, rho u, E): the state
    double* Ut;         // 3 planes: the stage state
    double* W;          // 3 planes (rho, u, p): primitives of the stage input
    double* F;          // 3 planes: x-face fluxes
    const BInst* inst;
    const int* order;   // workgroup -> instance (longest estimated run first) when queue is null
    BOut* out;
    // teams and queue (a cooperative launch of resident workgroups; null queue: one workgroup per instance)
    const BTeam* teams;
    int nteams, team_wgs;     // workgroups [0, team_wgs) belong to the teams, in team order
    const int* queue;         // the other instances, longest first, taken one at a time by any workgroup
    int nqueue, pad;
    unsigned* qnext;          // next queue entry (zeroed before the launch)
    TeamSync* sync;           // one per team (zeroed before the launch)
};

// the state of cell i of the instance as the 2-D path holds it: (rho, rho u, rho v = +0, E)
__device__ __forceinline__ void bload_cons(const BP& b, const BInst& in, const double* Us, int i, double* u) {
    const long long c = in.c0 + i;
    u[0] = Us[c];
    u[1] = Us[b.nc + c];
    u[2] = 0.0;
    u[3] = Us[2 * b.nc + c];
}
__device__ __forceinline__ void bload_prim(const BP& b, const BInst& in, int i, double* w) {
    const long long c = in.c0 + i;
    w[0] = b.W[c];
    w[1] = b.W[b.nc + c];
    w[2] = 0.0;
    w[3] = b.W[2 * b.nc + c];
}

// the padded primitive state at x index t (stencil_prim of the 2-D path, one row)
__device__ __forceinline__ void bstencil(const BP& b, const BInst& in, const double* Us, int t, double* w) {
    const int n = in.n;
    if (t >= 0 && t < n) {
        bload_prim(b, in, t, w);
        return;
    }
    const int hi = t >= n;
    const int h = hi ? in.halo_hi : in.halo_lo;
    if (h == 1) {
        bload_prim(b, in, hi ? t - n : t + n, w);
    } else if (h == 2) {
        const int l = hi ? t - n : -1 - t;
        double ui[4], ug[4];
        bload_cons(b, in, Us, hi ? n - 1 - l : l, ui);
        ghost_state<4>(b.rs.eos, b.bp, hi, ui, ug);
        c2p<4>(b.rs.eos, ug, w, (double*)nullptr);
    } else {
        bload_prim(b, in, t < 0 ? 0 : n - 1, w);
    }
}

// The cells of a workgroup's chunk [lo, hi) of the instance (the whole instance for a workgroup of its own), and
// the primitives (rho, u, p) of the KH cells below and above it that other members of its team hold (has_lo /
// has_hi; across the ends of a periodic instance the cells of the other end).
struct Chunk {
    int lo, hi;
    bool has_lo, has_hi;
    const double* hl;   // cells lo - KH .. lo - 1
    const double* hh;   // cells hi .. hi + KH - 1
};

// the padded primitive state at x index t as the stencils of the chunk's faces read it
__device__ __forceinline__ void cstencil(const BP& b, const BInst& in, const double* Us, const Chunk& ch, int t, double* w) {
    if (t >= ch.lo && t < ch.hi) {
        bload_prim(b, in, t, w);
        return;
    }
    const double* h = nullptr;
    if (t < ch.lo && ch.has_lo) h = ch.hl + (t - ch.lo + KH) * 3;
    else if (t >= ch.hi && ch.has_hi) h = ch.hh + (t - ch.hi) * 3;
    if (h) {
        w[0] = h[0];
        w[1] = h[1];
        w[2] = 0.0;
        w[3] = h[2];
        return;
    }
    bstencil(b, in, Us, t, w);
}

// the first-order state beyond x side `side` of its boundary cell (outside_first_order)
__device__ __forceinline__ void boutside(const BP& b, const BInst& in, const double* Us, int side, double* ug) {
    if (bc_of(b.bp, side) == 7) {
        bload_cons(b, in, Us, side ? 0 : in.n - 1, ug);
        return;
    }
    double ui[4];
    bload_cons(b, in, Us, side ? in.n - 1 : 0, ui);
    ghost_state<4>(b.rs.eos, b.bp, side, ui, ug);
}

// face f (between cells f - 1 and f) of the instance: the reconstruction (RECON = fvg_recon) and Riemann solve
// of the face pass
template <int RECON>
__device__ __forceinline__ void bface(const BP& b, const BInst& in, const double* Us, const Chunk& ch, int f, double* F, double& sm) {
    const int n = in.n;
    const bool lo = f == 0, hi = f == n;
    double UL[4], UR[4];
    if constexpr (RECON == 0) {   // first order (MODE_FIRST_ORDER): cell states, no round trip
        if (!lo) bload_cons(b, in, Us, f - 1, UL);
        else boutside(b, in, Us, 0, UL);
        if (!hi) bload_cons(b, in, Us, f, UR);
        else boutside(b, in, Us, 1, UR);
    } else if ((f > 0 && f < n) || (lo && in.halo_lo) || (hi && in.halo_hi)) {
        double wl[4], wr[4];
        const double we = b.weno_eps;
        if constexpr (RECON == 3) {
            double c[6][4];
#pragma unroll
            for (int s = 0; s < 6; ++s) cstencil(b, in, Us, ch, f - 3 + s, c[s]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q == 2) continue;   // v: see below
                wl[q] = weno5z(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], we);
                wr[q] = weno5z(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], we);
            }
        } else {
            double c[4][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) cstencil(b, in, Us, ch, f - 2 + s, c[s]);
            if constexpr (RECON == 2) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q == 2) continue;
                    wl[q] = weno3z(c[0][q], c[1][q], c[2][q], we);
                    wr[q] = weno3z(c[3][q], c[2][q], c[1][q], we);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q == 2) continue;
                    wl[q] = c[1][q] + 0.5 * limiter(b.lim, c[1][q] - c[0][q], c[2][q] - c[1][q]);
                    wr[q] = c[2][q] - 0.5 * limiter(b.lim, c[3][q] - c[2][q], c[2][q] - c[1][q]);
                }
            }
        }
        // v is +-0 in every stencil value (a no-slip ghost gives -0), so its reconstruction is a zero of some sign;
        // v enters the Riemann solver only squared (kinetic energy, the low-Mach Mach number) or in the rho v flux,
        // which the ensemble does not keep, so +0 gives the same (rho, rho u, E) fluxes and the reconstruction of v
        // is skipped
        wl[2] = 0.0;
        wr[2] = 0.0;
        p2c<4>(b.rs.eos, wl, UL);
        p2c<4>(b.rs.eos, wr, UR);
    } else {   // boundary face without a halo: first-order states through the primitive round trip
        if (!lo) bload_cons(b, in, Us, f - 1, UL);
        else boutside(b, in, Us, 0, UL);
        if (!hi) bload_cons(b, in, Us, f, UR);
        else boutside(b, in, Us, 1, UR);
        roundtrip<4>(b.rs.eos, UL);
        roundtrip<4>(b.rs.eos, UR);
    }
    face_flux<4>(b.rs, UL, UR, 0, F, sm);
}

// wave speed of the two (equal) y faces of cell i: the speed face_flux(u, u, y) returns for the cell's face state u,
// without its flux. Both states are u and the normal velocity is +0 (rho v = +0), so the Einfeldt bounds are
// S_L = 0 - c = -c and S_R = c, with c the sound speed of c2p(u), and max(|S_L|, |S_R|) = c; the low-Mach variant
// takes c before its velocity blend, which leaves equal states unchanged.
// w and inv: c2p of the cell's state, as the primitive pass computes them (the round trip's first half)
template <int RECON>
__device__ __forceinline__ double by_speed(const BP& b, const double* w, double inv) {
    if constexpr (RECON != 0) {   // the face state is the round trip p2c(c2p(u))
        double u2[4], w2[4], inv2;
        p2c<4>(b.rs.eos, w, u2);
        c2p<4>(b.rs.eos, u2, w2, &inv2);
        return sound_speed_inv(b.rs.eos, w2[3], inv2);
    }
    return sound_speed_inv(b.rs.eos, w[3], inv);
}

// maximum over the workgroup of non-negative, non-NaN doubles; every thread gets it
template <int BS>
__device__ __forceinline__ double bmax_all(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double other = __shfl_xor(v, o, 64);
        v = (other > v) ? other : v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int w = 1; w < BS / 64; ++w) m = (red[w] > m) ? red[w] : m;
    __syncthreads();
    return m;
}

// Barrier of the m workgroups of a team (thread 0 of each arrives, the workgroup waits on its own barrier around
// it). Arrivals only ever increase ts->count; barrier k of the run completes when it reaches (k + 1) m. The wait is
// bounded: a member that waits longer than kTeamSpin polls, or sees another member give up, marks the team broken
// and returns false, and the whole team leaves the instance (fvg_batch_run reports the failure).
//
// Ordering without agent-scope fences. The members share no plain memory: each reads and writes only the cells and
// faces of its own chunk, and everything they exchange (end-cell primitives, wave-speed maxima) goes through
// agent-scope atomic stores and loads (team_put / team_get: global_store / global_load with sc1, coherent across the
// XCDs like the counter itself), all stored by the first wavefront, which also arrives. Its s_waitcnt before the
// arrival (on gfx950 the vector memory counter covers stores) makes the published values complete before the
// counter moves, and the readers issue their loads after the workgroup barrier that follows the last arrival seen.
// A release / acquire fence at agent scope would add buffer_wbl2 / buffer_inv: a write-back and invalidation of the
// whole L2 of the member's XCD at every barrier, which every other workgroup of that XCD pays for (the 1024-instance
// run took 14.6 s with them, 12.3 s without; same bits).
constexpr unsigned kTeamSpin = 1u << 22;
// a double exchanged between the members of a team: device-scope atomic store and load of its bits
__device__ __forceinline__ void team_put(double* p, double v) {
    __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double team_get(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ bool team_barrier(TeamSync* ts, int m, unsigned long long& k, int* flag) {
    __syncthreads();   // the workgroup's writes of this phase precede thread 0's arrival
    if (threadIdx.x == 0) {
        const unsigned long long target = (k + 1) * (unsigned long long)m;
        int ok = 1;
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's published values are complete
        __hip_atomic_fetch_add(&ts->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(&ts->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&ts->broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > kTeamSpin) {
                __hip_atomic_store(&ts->broken, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *flag = ok;
    }
    ++k;
    __syncthreads();
    return *flag != 0;
}

// The run of instance id by rank r of a team of m workgroups (m = 1: the workgroup alone, ts unused): steps and
// stages of the time scheme over the rank's chunk of cells. With m > 1 each stage has two team barriers: after the
// primitives (the ranks publish the primitives of their KH end cells and read their neighbours'), and after the
// faces (the ranks publish their chunk's wave-speed maximum and take the team's). The maximum of the same values in
// another order is the same double, and every rank computes the same dt, time and step count, so the ranks leave
// the loop together and the instance takes the bits of a single workgroup's run. A chunk's faces are lo .. hi; face
// hi, which the next rank owns, is computed again for the chunk's last cell and kept in LDS.
template <int BS, int RECON>
__device__ void run_instance(const BP& b, int id, int r, int m, TeamSync* ts, double* red, double* hl, double* hh, double* fx,
                             int* flag) {
    const BInst in = b.inst[id];
    const int n = in.n;
    const int nst = b.scheme + 1;
    const bool team = m > 1;
    const bool per = in.halo_lo == 1;   // periodic: the first rank's low neighbour is the last rank
    Chunk ch;
    ch.lo = (int)((long long)n * r / m);
    ch.hi = (int)((long long)n * (r + 1) / m);
    ch.has_lo = team && (r > 0 || per);
    ch.has_hi = team && (r < m - 1 || per);
    ch.hl = hl;
    ch.hh = hh;
    const bool extra = team && r < m - 1;   // face hi belongs to the next rank
    const int r_lo = (r + m - 1) % m, r_hi = (r + 1) % m;
    unsigned long long nbar = 0;
    double t = 0.0, dt = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s_step = 0.0;
    int steps = 0, diverged = 0;
    while (t < b.t_end && steps < b.max_steps) {
        for (int st = 1; st <= nst; ++st) {
            const double* Us = st == 1 ? b.U : b.Ut;
            // primitives of the stage input (k_cells) and the wave speed of each cell's y faces
            double local = 0.0;
            for (int i = ch.lo + (int)threadIdx.x; i < ch.hi; i += BS) {
                double u[4], w[4], inv;
                bload_cons(b, in, Us, i, u);
                c2p<4>(b.rs.eos, u, w, &inv);
                const long long c = in.c0 + i;
                b.W[c] = w[0];
                b.W[b.nc + c] = w[1];
                b.W[2 * b.nc + c] = w[3];
                const double sy = by_speed<RECON>(b, w, inv);
                if (sy > local) local = sy;
            }
            if (team) {
                // lanes of the first wavefront publish the chunk's end cells (KH cells x (rho, u, p)) one value
                // each; the barrier's workgroup barrier and thread 0's release order them before its arrival. After
                // it the same lanes read the neighbours' into LDS, all loads in flight at once: the low halo is the
                // low neighbour's high end, the high halo the high neighbour's low end.
                __syncthreads();
                if (threadIdx.x < 2 * KH * 3) {
                    const int side = threadIdx.x / (KH * 3), e = threadIdx.x % (KH * 3);
                    const long long c = in.c0 + (side ? ch.hi - KH : ch.lo) + e / 3;
                    team_put(&ts->halo[r][side][e], b.W[(long long)(e % 3) * b.nc + c]);
                }
                if (!team_barrier(ts, m, nbar, flag)) {
                    diverged = 2;
                    break;
                }
                if (threadIdx.x < 2 * KH * 3) {
                    const int side = threadIdx.x / (KH * 3), e = threadIdx.x % (KH * 3);
                    (side ? hh : hl)[e] = team_get(&ts->halo[side ? r_hi : r_lo][side ? 0 : 1][e]);
                }
            }
            __syncthreads();
            // x faces (k_faces)
            for (int f = ch.lo + (int)threadIdx.x; f <= ch.hi; f += BS) {
                double F[4], sm;
                bface<RECON>(b, in, Us, ch, f, F, sm);
                if (sm > local) local = sm;
                if (extra && f == ch.hi) {
                    fx[0] = F[0];
                    fx[1] = F[1];
                    fx[2] = F[3];
                } else {
                    const long long x = in.f0 + f;
                    b.F[x] = F[0];
                    b.F[b.nf + x] = F[1];
                    b.F[2 * b.nf + x] = F[3];
                }
            }
            double S = bmax_all<BS>(local, red);   // its barriers also publish the face fluxes
            if (team) {   // the team's maximum
                if (threadIdx.x == 0) team_put(&ts->smax[r], S);
                if (!team_barrier(ts, m, nbar, flag)) {
                    diverged = 2;
                    break;
                }
                if (threadIdx.x < 64) {   // lane k reads rank k's maximum; a NaN of any rank is kept, so that the
                                          // team reports the divergence
                    double v = threadIdx.x < (unsigned)m ? team_get(&ts->smax[threadIdx.x]) : 0.0;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        const double w = __shfl_xor(v, o, 64);
                        v = (v < w || w != w) ? w : v;
                    }
                    if (threadIdx.x == 0) red[0] = v;
                }
                __syncthreads();
                S = red[0];
                __syncthreads();
            }
            // the scalar part of the scheme (k_control)
            if (st == 1) {
                s1 = S;
                dt = smin(b.cfl * in.Ldt / (S + kDtEps), b.t_end - t);
                if (b.dt_max > 0.0) dt = smin(dt, b.dt_max);
            } else if (st == 2) {
                s2 = S;
            } else {
                s3 = S;
            }
            // balance and stage combination (k_update)
            for (int i = ch.lo + (int)threadIdx.x; i < ch.hi; i += BS) {
                const long long c = in.c0 + i, x = in.f0 + i;
                const bool last = extra && i == ch.hi - 1;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double* Fq = b.F + q * b.nf;
                    const double fl = Fq[x], fr = last ? fx[q] : Fq[x + 1];
                    const double R = (0.0 - (fr * in.Ax - fl * in.Ax)) * in.inv_vol;
                    const long long y = q * b.nc + c;
                    if (b.scheme == 0) {
                        b.U[y] = b.U[y] + dt * R;
                    } else if (st == 1) {
                        b.Ut[y] = b.U[y] + dt * R;
                    } else if (b.scheme == 1) {
                        b.U[y] = 0.5 * b.U[y] + 0.5 * (b.Ut[y] + dt * R);
                    } else if (st == 2) {
                        b.Ut[y] = 0.75 * b.U[y] + 0.25 * (b.Ut[y] + dt * R);
                    } else {
                        b.U[y] = (1.0 / 3.0) * b.U[y] + (2.0 / 3.0) * (b.Ut[y] + dt * R);
                    }
                }
            }
            __syncthreads();
        }
        if (diverged == 2) break;
        s_step = b.scheme == 0 ? s1 : (b.scheme == 1 ? smax(s1, s2) : smax(s1, smax(s2, s3)));
        steps += 1;
        // check_progress (SimulationInternal.hpp:259-267)
        if (!(isfinite(s_step) && isfinite(dt) && t + dt > t)) {
            diverged = 1;
            break;
        }
        t += dt;
    }
    if (r == 0 && threadIdx.x == 0) {
        BOut o;
        o.t = t;
        o.s_last = s_step;
        o.dt_last = dt;
        o.steps = steps;
        o.diverged = diverged;
        b.out[id] = o;
    }
    __syncthreads();
}

// Workgroups [0, team_wgs) run the teams' instances, then every workgroup takes the queued instances one at a time
// (longest first) until none is left. Without a queue, workgroup k runs instance order[k] alone.
// 512 and 768 threads: at most 168 VGPRs (three wavefronts per SIMD; with 768 threads a CU runs 12 wavefronts of one
// workgroup), measured faster than the compiler's 210 at 512 threads (14.4 against 15.7 s on the 1024-instance run);
// the other sizes keep their default limits
template <int BS> constexpr int kEnsembleWaves = (BS == 512 || BS == 768) ? 3 : 1;
template <int BS, int RECON>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kEnsembleWaves<BS>))) void k_ensemble(BP b) {
    __shared__ double red[BS / 64];
    __shared__ double hl[KH * 3], hh[KH * 3], fx[3];
    __shared__ int flag, slot;
    const int wg = blockIdx.x;
    // 0: the launch's own instance (no queue), 1: the workgroup's team, 2: the queue
    int phase = !b.queue ? 0 : (wg < b.team_wgs ? 1 : 2);
    for (;;) {
        int id, r = 0, m = 1;
        TeamSync* ts = nullptr;
        if (phase == 0) {
            id = b.order[wg];
        } else if (phase == 1) {
            int k = 0;
            while (wg >= b.teams[k].wg0 + b.teams[k].m) ++k;
            id = b.teams[k].inst;
            r = wg - b.teams[k].wg0;
            m = b.teams[k].m;
            ts = b.sync + k;
        } else {
            if (threadIdx.x == 0) slot = (int)atomicAdd(b.qnext, 1u);
            __syncthreads();
            const int k = slot;
            __syncthreads();
            if (k >= b.nqueue) break;
            id = b.queue[k];
        }
        run_instance<BS, RECON>(b, id, r, m, ts, red, hl, hh, fx, &flag);
        if (phase == 0) break;
        phase = 2;
    }
}


// ------------------------------------------------------------------------------------------------------------
// 2-D instances (fvg_batch_create_2d): instance i is the reference's run on an nx_i x ny_i mesh with the four sides,
// Lx and Ly of the shared configuration; the same persistent layout (one workgroup per instance, its whole run),
// with the 2-D path's arithmetic (NV = 4): primitives, x and y faces, wave-speed maximum, dt, balance and stage
// combination per stage. Halos as the padded planes of the 2-D path hold them (push_halo): periodic wrap, ghost slab
// (boundary_order = 2: the ghost state of the mirrored cell), or the edge cell.
// ------------------------------------------------------------------------------------------------------------
struct BInst2 {
    long long c0, fx0, fy0;   // first cell, first x face and first y face of the instance in the planes
    int nx, ny;
    int halo[4];              // per side: 0 none, 1 periodic wrap, 2 ghost slab
    double Ax, Ay, inv_vol, Ldt;
};

struct BP2 {
    RiemannParams rs;
    BcParams bp;
    int recon, lim, scheme, max_steps;
    double weno_eps, cfl, t_end, dt_max;
    long long nc, nf;   // plane strides: cells and faces (x and y) of the whole ensemble
    double* U;          // 4 planes (rho, rho u, rho v, E)
    double* Ut;
    double* W;          // 4 planes (rho, u, v, p)
    double* F;          // 4 planes: the x faces then the y faces of each instance
    const BInst2* inst;
    const int* order;
    BOut* out;
    // teams and queue as for 1-D instances (BP); a team's ranks take row slabs of its instance, and exchange the
    // primitives of their KH end rows through xch (per team from BTeam::xoff: [rank][side][q][KH][nx] doubles, then
    // [rank][q][nx], the y-face fluxes of a rank's last face row, which the next rank owns)
    const BTeam* teams;
    int nteams, team_wgs;
    const int* queue;
    int nqueue, pad;
    unsigned* qnext;
    TeamSync* sync;
    double* xch;
};

__device__ __forceinline__ void b2load(const double* P, long long stride, long long c, double* v) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = P[q * stride + c];
}

// the padded primitive state at coordinate t along axis d of the line through (i, j) (stencil_prim of the 2-D path)
__device__ __forceinline__ void b2stencil(const BP2& b, const BInst2& in, const double* Us, int d, int i, int j, int t, double* w) {
    const int n = d == 0 ? in.nx : in.ny;
    auto cell = [&](int tt) { return in.c0 + (long long)(d == 1 ? tt : j) * in.nx + (d == 0 ? tt : i); };
    if (t >= 0 && t < n) {
        b2load(b.W, b.nc, cell(t), w);
        return;
    }
    const int hi = t >= n;
    const int h = in.halo[2 * d + hi];
    if (h == 1) {
        b2load(b.W, b.nc, cell(hi ? t - n : t + n), w);
    } else if (h == 2) {
        const int l = hi ? t - n : -1 - t;
        double ui[4], ug[4];
        b2load(Us, b.nc, cell(hi ? n - 1 - l : l), ui);
        ghost_state<4>(b.rs.eos, b.bp, 2 * d + hi, ui, ug);
        c2p<4>(b.rs.eos, ug, w, (double*)nullptr);
    } else {
        b2load(b.W, b.nc, cell(t < 0 ? 0 : n - 1), w);
    }
}

// The rows [lo, hi) of a 2-D instance that a team's rank owns, and the exchange rows of its team: the primitives of
// the KH rows below lo (has_lo) and from hi (has_hi) are the neighbours' end rows there (across the ends of a
// periodic y the rows of the other end).
struct Chunk2 {
    int lo, hi;
    bool has_lo, has_hi;
    const double* xl;   // low neighbour's high end rows: [q][KH][nx]
    const double* xh;   // high neighbour's low end rows
};

// the padded primitive state at row t of column i (y stencils) as the chunk's faces read it
__device__ __forceinline__ void c2stencil(const BP2& b, const BInst2& in, const double* Us, const Chunk2& ch, int i, int t, double* w) {
    const double* x = nullptr;
    int k = 0;
    if (t < ch.lo && ch.has_lo) {
        x = ch.xl;
        k = t - ch.lo + KH;
    } else if (t >= ch.hi && ch.has_hi) {
        x = ch.xh;
        k = t - ch.hi;
    }
    if (x) {
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = team_get(x + ((long long)q * KH + k) * in.nx + i);
        return;
    }
    b2stencil(b, in, Us, 1, i, 0, t, w);
}

// the first-order state beyond side 2d + hi of the boundary cell of the line through (i, j) (outside_first_order)
__device__ __forceinline__ void b2outside(const BP2& b, const BInst2& in, const double* Us, int d, int hi, int i, int j, double* ug) {
    const int n = d == 0 ? in.nx : in.ny, side = 2 * d + hi;
    auto cell = [&](int tt) { return in.c0 + (long long)(d == 1 ? tt : j) * in.nx + (d == 0 ? tt : i); };
    if (bc_of(b.bp, side) == 7) {
        b2load(Us, b.nc, cell(hi ? 0 : n - 1), ug);
        return;
    }
    double ui[4];
    b2load(Us, b.nc, cell(hi ? n - 1 : 0), ui);
    ghost_state<4>(b.rs.eos, b.bp, side, ui, ug);
}

// face t along axis d of the line through (i, j): the states of face_states (MODE_PLAIN / MODE_FIRST_ORDER) and the
// Riemann flux
template <int RECON>
__device__ __forceinline__ void b2face(const BP2& b, const BInst2& in, const double* Us, const Chunk2& ch, int d, int i, int j, int t, double* F,
                                       double& sm) {
    const int n = d == 0 ? in.nx : in.ny;
    const bool lo = t == 0, hi = t == n;
    auto cell = [&](int tt) { return in.c0 + (long long)(d == 1 ? tt : j) * in.nx + (d == 0 ? tt : i); };
    double UL[4], UR[4];
    const bool upgrade = RECON != 0 && ((t > 0 && t < n) || (lo && in.halo[2 * d]) || (hi && in.halo[2 * d + 1]));
    if (upgrade) {
        double wl[4], wr[4];
        const double we = b.weno_eps;
        if constexpr (RECON == 3) {
            double c[6][4];
#pragma unroll
            for (int s = 0; s < 6; ++s) d ? c2stencil(b, in, Us, ch, i, t - 3 + s, c[s]) : b2stencil(b, in, Us, 0, i, j, t - 3 + s, c[s]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                wl[q] = weno5z(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], we);
                wr[q] = weno5z(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], we);
            }
        } else {
            double c[4][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) d ? c2stencil(b, in, Us, ch, i, t - 2 + s, c[s]) : b2stencil(b, in, Us, 0, i, j, t - 2 + s, c[s]);
            if constexpr (RECON == 2) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    wl[q] = weno3z(c[0][q], c[1][q], c[2][q], we);
                    wr[q] = weno3z(c[3][q], c[2][q], c[1][q], we);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    wl[q] = c[1][q] + 0.5 * limiter(b.lim, c[1][q] - c[0][q], c[2][q] - c[1][q]);
                    wr[q] = c[2][q] - 0.5 * limiter(b.lim, c[3][q] - c[2][q], c[2][q] - c[1][q]);
                }
            }
        }
        p2c<4>(b.rs.eos, wl, UL);
        p2c<4>(b.rs.eos, wr, UR);
    } else {
        if (!lo) b2load(Us, b.nc, cell(t - 1), UL);
        else b2outside(b, in, Us, d, 0, i, j, UL);
        if (!hi) b2load(Us, b.nc, cell(t), UR);
        else b2outside(b, in, Us, d, 1, i, j, UR);
        if constexpr (RECON != 0) {   // a boundary face without a halo: the primitive round trip
            roundtrip<4>(b.rs.eos, UL);
            roundtrip<4>(b.rs.eos, UR);
        }
    }
    face_flux<4>(b.rs, UL, UR, d, F, sm);
}

// The run of 2-D instance id by rank r of a team of m workgroups (m = 1: the workgroup alone), over the rows
// [lo, hi) of its slab, with the barriers and the exchange of run_instance: after the primitives the ranks publish
// their KH end rows (every wavefront stores, then waits for its stores, before the workgroup barrier that precedes
// the arrival) and the y stencils read the neighbours' from xch; after the faces the team takes the maximum of the
// wave speeds. The y faces of a slab are rows lo .. hi; row hi, which the next rank owns, is computed again for the
// slab's last row and kept in the rank's own part of xch.
template <int BS, int RECON>
__device__ void run_instance2(const BP2& b, int id, int r, int m, TeamSync* ts, double* xch, double* red, int* flag) {
    const BInst2 in = b.inst[id];
    const int nx = in.nx, ny = in.ny;
    const int nst = b.scheme + 1;
    const bool team = m > 1;
    const bool per = in.halo[2] == 1;   // periodic y: the first rank's low neighbour is the last rank
    const long long rows_x = 2LL * 4 * KH * nx;   // exchange rows of one rank
    Chunk2 ch;
    ch.lo = (int)((long long)ny * r / m);
    ch.hi = (int)((long long)ny * (r + 1) / m);
    ch.has_lo = team && (r > 0 || per);
    ch.has_hi = team && (r < m - 1 || per);
    const int r_lo = (r + m - 1) % m, r_hi = (r + 1) % m;
    ch.xl = team ? xch + r_lo * rows_x + 4LL * KH * nx : nullptr;   // the low neighbour's high end
    ch.xh = team ? xch + r_hi * rows_x : nullptr;                    // the high neighbour's low end
    double* ytop = team ? xch + m * rows_x + (long long)r * 4 * nx : nullptr;   // this rank's copy of face row hi
    const bool extra = team && r < m - 1;
    const int c0 = ch.lo * nx, c1 = ch.hi * nx;   // the slab's cells
    const int nfx = (nx + 1) * (ch.hi - ch.lo), nfy = nx * (ch.hi - ch.lo + 1);
    unsigned long long nbar = 0;
    double t = 0.0, dt = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s_step = 0.0;
    int steps = 0, diverged = 0;
    while (t < b.t_end && steps < b.max_steps) {
        for (int st = 1; st <= nst; ++st) {
            const double* Us = st == 1 ? b.U : b.Ut;
            for (int c = c0 + (int)threadIdx.x; c < c1; c += BS) {   // primitives of the stage input (k_cells)
                double u[4], w[4];
                b2load(Us, b.nc, in.c0 + c, u);
                c2p<4>(b.rs.eos, u, w, (double*)nullptr);
#pragma unroll
                for (int q = 0; q < 4; ++q) b.W[q * b.nc + in.c0 + c] = w[q];
            }
            if (team) {   // the slab's KH end rows, side 0 the low end and side 1 the high end
                __syncthreads();
                for (int e = threadIdx.x; e < 2 * 4 * KH * nx; e += BS) {
                    const int side = e / (4 * KH * nx), q = (e / (KH * nx)) % 4, k = (e / nx) % KH, i = e % nx;
                    const int row = side ? ch.hi - KH + k : ch.lo + k;
                    team_put(xch + r * rows_x + e, b.W[q * b.nc + in.c0 + (long long)row * nx + i]);
                }
                __builtin_amdgcn_s_waitcnt(0);   // every wavefront's published rows are complete
                if (!team_barrier(ts, m, nbar, flag)) {
                    diverged = 2;
                    break;
                }
            }
            __syncthreads();
            double local = 0.0;
            for (int f = threadIdx.x; f < nfx + nfy; f += BS) {   // x faces, then y faces (k_faces)
                const int d = f < nfx ? 0 : 1;
                const int g = d == 0 ? f : f - nfx;
                const int w = d == 0 ? nx + 1 : nx;
                const int jj = g / w, ii = g - jj * w;   // x faces: face ii of slab row jj; y faces: face row jj of column ii
                const int j = ch.lo + jj;
                double F[4], sm;
                b2face<RECON>(b, in, Us, ch, d, d == 0 ? 0 : ii, d == 0 ? j : 0, d == 0 ? ii : j, F, sm);
                if (sm > local) local = sm;
                if (d == 1 && extra && j == ch.hi) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) ytop[q * nx + ii] = F[q];
                } else {
                    const long long x = d == 0 ? in.fx0 + (long long)j * (nx + 1) + ii : in.fy0 + (long long)j * nx + ii;
#pragma unroll
                    for (int q = 0; q < 4; ++q) b.F[q * b.nf + x] = F[q];
                }
            }
            double S = bmax_all<BS>(local, red);   // its barriers also publish the face fluxes
            if (team) {   // the team's maximum
                if (threadIdx.x == 0) team_put(&ts->smax[r], S);
                if (!team_barrier(ts, m, nbar, flag)) {
                    diverged = 2;
                    break;
                }
                if (threadIdx.x < 64) {   // as in run_instance: a NaN of any rank is kept
                    double v = threadIdx.x < (unsigned)m ? team_get(&ts->smax[threadIdx.x]) : 0.0;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        const double w = __shfl_xor(v, o, 64);
                        v = (v < w || w != w) ? w : v;
                    }
                    if (threadIdx.x == 0) red[0] = v;
                }
                __syncthreads();
                S = red[0];
                __syncthreads();
            }
            if (st == 1) {   // k_control
                s1 = S;
                dt = smin(b.cfl * in.Ldt / (S + kDtEps), b.t_end - t);
                if (b.dt_max > 0.0) dt = smin(dt, b.dt_max);
            } else if (st == 2) {
                s2 = S;
            } else {
                s3 = S;
            }
            for (int c = c0 + (int)threadIdx.x; c < c1; c += BS) {   // balance and stage combination (k_update)
                const int j = c / nx, i = c - j * nx;
                const long long fx = in.fx0 + (long long)j * (nx + 1) + i, fy = in.fy0 + (long long)j * nx + i;
                const bool top = extra && j == ch.hi - 1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double* Fq = b.F + q * b.nf;
                    const double fl = Fq[fx], fr = Fq[fx + 1], fb = Fq[fy], ft = top ? ytop[q * nx + i] : Fq[fy + nx];
                    const double R = ((0.0 - (fr * in.Ax - fl * in.Ax)) - (ft * in.Ay - fb * in.Ay)) * in.inv_vol;
                    const long long y = q * b.nc + in.c0 + c;
                    if (b.scheme == 0) {
                        b.U[y] = b.U[y] + dt * R;
                    } else if (st == 1) {
                        b.Ut[y] = b.U[y] + dt * R;
                    } else if (b.scheme == 1) {
                        b.U[y] = 0.5 * b.U[y] + 0.5 * (b.Ut[y] + dt * R);
                    } else if (st == 2) {
                        b.Ut[y] = 0.75 * b.U[y] + 0.25 * (b.Ut[y] + dt * R);
                    } else {
                        b.U[y] = (1.0 / 3.0) * b.U[y] + (2.0 / 3.0) * (b.Ut[y] + dt * R);
                    }
                }
            }
            __syncthreads();
        }
        if (diverged == 2) break;
        s_step = b.scheme == 0 ? s1 : (b.scheme == 1 ? smax(s1, s2) : smax(s1, smax(s2, s3)));
        steps += 1;
        if (!(isfinite(s_step) && isfinite(dt) && t + dt > t)) {   // check_progress (SimulationInternal.hpp:259-267)
            diverged = 1;
            break;
        }
        t += dt;
    }
    if (r == 0 && threadIdx.x == 0) {
        BOut o;
        o.t = t;
        o.s_last = s_step;
        o.dt_last = dt;
        o.steps = steps;
        o.diverged = diverged;
        b.out[id] = o;
    }
    __syncthreads();
}

// as k_ensemble: workgroups [0, team_wgs) run the teams' instances, then every workgroup takes queued instances;
// without a queue, workgroup k runs instance order[k] alone
template <int BS, int RECON>
__global__ __launch_bounds__(BS) void k_ensemble2(BP2 b) {
    __shared__ double red[BS / 64];
    __shared__ int flag, slot;
    const int wg = blockIdx.x;
    int phase = !b.queue ? 0 : (wg < b.team_wgs ? 1 : 2);
    for (;;) {
        int id, r = 0, m = 1;
        TeamSync* ts = nullptr;
        double* xch = nullptr;
        if (phase == 0) {
            id = b.order[wg];
        } else if (phase == 1) {
            int k = 0;
            while (wg >= b.teams[k].wg0 + b.teams[k].m) ++k;
            id = b.teams[k].inst;
            r = wg - b.teams[k].wg0;
            m = b.teams[k].m;
            ts = b.sync + k;
            xch = b.xch + b.teams[k].xoff;
        } else {
            if (threadIdx.x == 0) slot = (int)atomicAdd(b.qnext, 1u);
            __syncthreads();
            const int k = slot;
            __syncthreads();
            if (k >= b.nqueue) break;
            id = b.queue[k];
        }
        run_instance2<BS, RECON>(b, id, r, m, ts, xch, red, &flag);
        if (phase == 0) break;
        phase = 2;
    }
}


}   // namespace fvg

using namespace fvg;

namespace {
std::string g_batch_create_error;
const int kBlocks[] = {64, 128, 256, 512, 768, 1024};
}   // namespace

struct fvg_batch {
    fvg_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    int ninst = 0;
    long long nc = 0, nf = 0;
    int dim = 1;                  // 1: N x 1 instances (3 planes), 2: nx x ny instances (4 planes)
    int nplanes = 3;
    std::vector<int> n;           // cells per instance
    std::vector<BInst> inst;
    std::vector<BInst2> inst2;
    std::vector<int> order;
    BInst* d_inst = nullptr;
    BInst2* d_inst2 = nullptr;
    BP2 base2{};
    int* d_order = nullptr;
    BOut* d_out = nullptr;
    double *U = nullptr, *Ut = nullptr, *W = nullptr, *F = nullptr, *Usave = nullptr;
    BP base{};
    int block = 256;
    bool uploaded = false;
    double last_ms = 0.0;
    std::string err;
    // teams (1-D instances): the largest instances split over several workgroups of a cooperative launch
    int max_team = 8, min_chunk = 512, force_teams = 0;
    bool coop = false;                   // the device supports cooperative launches
    std::vector<double> cost;            // estimated work per instance (N^2 x initial wave speed)
    BTeam* d_teams = nullptr;
    int* d_queue = nullptr;
    unsigned* d_qnext = nullptr;
    TeamSync* d_sync = nullptr;
    double* d_xch = nullptr;             // 2-D teams: the exchange rows (BP2::xch), grown as plans need
    long long xch_cap = 0;               // its size in doubles
    int plan_teams = 0, plan_team_wgs = 0, plan_grid = 0;   // the last launch's plan
};

namespace {

int bset(fvg_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    else g_batch_create_error = m;
    return code;
}

#define BCHK(b, call)                                                                                      \
    do {                                                                                                   \
        const hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return bset(b, FVG_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void free_batch(fvg_batch* b) {
    if (!b) return;
    hipFree(b->d_inst);
    hipFree(b->d_inst2);
    hipFree(b->d_order);
    hipFree(b->d_out);
    hipFree(b->U);
    hipFree(b->Ut);
    hipFree(b->W);
    hipFree(b->F);
    hipFree(b->Usave);
    hipFree(b->d_teams);
    hipFree(b->d_queue);
    hipFree(b->d_qnext);
    hipFree(b->d_sync);
    hipFree(b->d_xch);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

template <int BS>
void launch_bs(fvg_batch* b, const BP& p) {
    const dim3 g(b->ninst), t(BS);
    if (b->dim == 2) {   // the run's time, dt_max and step limit go with the 2-D parameters
        BP2 q = b->base2;
        q.t_end = p.t_end;
        q.dt_max = p.dt_max;
        q.max_steps = p.max_steps;
        switch (q.recon) {
        case 0: hipLaunchKernelGGL((k_ensemble2<BS, 0>), g, t, 0, b->stream, q); break;
        case 1: hipLaunchKernelGGL((k_ensemble2<BS, 1>), g, t, 0, b->stream, q); break;
        case 2: hipLaunchKernelGGL((k_ensemble2<BS, 2>), g, t, 0, b->stream, q); break;
        default: hipLaunchKernelGGL((k_ensemble2<BS, 3>), g, t, 0, b->stream, q); break;
        }
        return;
    }
    switch (p.recon) {
    case 0: hipLaunchKernelGGL((k_ensemble<BS, 0>), g, t, 0, b->stream, p); break;
    case 1: hipLaunchKernelGGL((k_ensemble<BS, 1>), g, t, 0, b->stream, p); break;
    case 2: hipLaunchKernelGGL((k_ensemble<BS, 2>), g, t, 0, b->stream, p); break;
    default: hipLaunchKernelGGL((k_ensemble<BS, 3>), g, t, 0, b->stream, p); break;
    }
}

// The teams of a launch of `grid` resident workgroups of `bs` threads. With force_teams every instance of at least
// 2 min_chunk cells gets min(max_team, N / min_chunk) workgroups while they last. Otherwise the plan shortens the run
// greedily: an instance's run is modelled as steps x (a + b N / m + c [m > 1]) microseconds, with steps proportional
// to its estimated work over N (work = N^2 x initial wave speed), a the fixed latency of a stage, b the time per cell
// of a chunk and c the cost of the two team barriers of a stage (about 9, 0.0065 and 6 us measured on MI355X with
// 1,024 threads; b grows as the square root of 1024 / bs). The longest modelled run takes one more workgroup while
// that shortens it without raising the bound max(longest run, total workgroup time / grid), each chunk keeps at least min_chunk cells, a
// team at most max_team workgroups and the teams at most three quarters of the grid. The other instances form the
// queue, longest first.
void plan_teams(const fvg_batch* b, int grid, int bs, std::vector<BTeam>& teams, std::vector<int>& queue) {
    teams.clear();
    queue.clear();
    const int ninst = (int)b->order.size();
    std::vector<int> m(b->n.size(), 1);
    auto cap = [&](int id) {   // most workgroups instance id may take
        const int n = b->n[id];
        if (b->max_team < 2 || b->min_chunk < KH || n < 2 * b->min_chunk) return 1;
        int k = std::min(b->max_team, std::min(kMaxTeam, n / b->min_chunk));
        if (b->dim == 2) {   // slabs of whole rows, at least KH of them and min_chunk cells
            const int nx = b->inst2[id].nx, ny = b->inst2[id].ny;
            const int rows = std::max(KH, (b->min_chunk + nx - 1) / nx);
            k = std::min(k, ny / rows);
        }
        return std::max(1, k);
    };
    int wgs = 0;
    if (b->force_teams) {
        for (int id : b->order) {
            const int k = cap(id);
            if (k >= 2 && wgs + k <= grid) {
                m[id] = k;
                wgs += k;
            }
        }
    } else if (b->max_team > 1 && grid > 1) {
        const double A = 9.0, Bc = 0.0065 * std::sqrt(1024.0 / bs), C = 6.0;
        auto run = [&](int id, int k) {
            const double n = b->n[id], steps = b->cost[id] / std::max(1.0, n);
            return steps * (A + Bc * n / k + (k > 1 ? C : 0.0));
        };
        double total = 0.0;
        std::vector<std::pair<double, int>> heap;   // (modelled run, instance), longest on top
        for (int id = 0; id < ninst; ++id) {
            const double t = run(id, 1);
            total += t;
            heap.push_back({t, id});
        }
        std::make_heap(heap.begin(), heap.end());
        const int max_wgs = grid * 3 / 4;
        while (!heap.empty()) {
            std::pop_heap(heap.begin(), heap.end());
            const auto [t, id] = heap.back();
            heap.pop_back();
            const int k = m[id] + 1, extra = (m[id] == 1 ? 2 : 1);
            if (k > cap(id) || wgs + extra > max_wgs) break;
            const double t2 = run(id, k);
            const double total2 = total - m[id] * t + k * t2;
            const double next = heap.empty() ? 0.0 : heap.front().first;
            // taken while the bound does not rise (equal runs: the others follow) and this run gets shorter
            if (!(t2 < t) || std::max(std::max(next, t2), total2 / grid) > std::max(t, total / grid)) break;
            wgs += extra;
            m[id] = k;
            total = total2;
            heap.push_back({t2, id});
            std::push_heap(heap.begin(), heap.end());
        }
    }
    wgs = 0;
    for (int id : b->order) {
        if (m[id] >= 2) {
            teams.push_back({id, wgs, m[id], 0});
            wgs += m[id];
        } else {
            queue.push_back(id);
        }
    }
}

// The teamed launch of the 1-D kernel: workgroups resident at once (a cooperative launch), teams then the queue.
template <int BS, int RECON>
int launch_teams(fvg_batch* b, BP p) {
    int per_cu = 0, dev_cus = 0;
    BCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ensemble<BS, RECON>, BS, 0));
    BCHK(b, hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const int resident = per_cu * dev_cus;
    if (resident < 1) return bset(b, FVG_EHIP, "the ensemble kernel cannot be resident");
    std::vector<BTeam> teams;
    std::vector<int> queue;
    plan_teams(b, resident, BS, teams, queue);
    int wgs = 0;
    for (const BTeam& t : teams) wgs += t.m;
    const int grid = std::min(resident, wgs + (int)queue.size());
    if (!teams.empty()) BCHK(b, hipMemcpy(b->d_teams, teams.data(), sizeof(BTeam) * teams.size(), hipMemcpyHostToDevice));
    if (!queue.empty()) BCHK(b, hipMemcpy(b->d_queue, queue.data(), sizeof(int) * queue.size(), hipMemcpyHostToDevice));
    BCHK(b, hipMemsetAsync(b->d_qnext, 0, sizeof(unsigned), b->stream));
    if (!teams.empty()) BCHK(b, hipMemsetAsync(b->d_sync, 0, sizeof(TeamSync) * teams.size(), b->stream));
    p.teams = b->d_teams;
    p.nteams = (int)teams.size();
    p.team_wgs = wgs;
    p.queue = b->d_queue;
    p.nqueue = (int)queue.size();
    p.qnext = b->d_qnext;
    p.sync = b->d_sync;
    void* args[] = {&p};
    BCHK(b, hipLaunchCooperativeKernel((const void*)k_ensemble<BS, RECON>, dim3(grid), dim3(BS), args, 0, b->stream));
    b->plan_teams = (int)teams.size();
    b->plan_team_wgs = wgs;
    b->plan_grid = grid;
    return FVG_OK;
}

template <int BS>
int launch_teams_bs(fvg_batch* b, const BP& p) {
    switch (p.recon) {
    case 1: return launch_teams<BS, 1>(b, p);
    case 2: return launch_teams<BS, 2>(b, p);
    default: return launch_teams<BS, 3>(b, p);
    }
}

// The teamed launch of the 2-D kernel: as launch_teams, with each team's exchange rows in d_xch.
template <int BS, int RECON>
int launch_teams2(fvg_batch* b, const BP& p) {
    int per_cu = 0, dev_cus = 0;
    BCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ensemble2<BS, RECON>, BS, 0));
    BCHK(b, hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const int resident = per_cu * dev_cus;
    if (resident < 1) return bset(b, FVG_EHIP, "the ensemble kernel cannot be resident");
    std::vector<BTeam> teams;
    std::vector<int> queue;
    plan_teams(b, resident, BS, teams, queue);
    int wgs = 0;
    long long xn = 0;
    for (BTeam& t : teams) {
        wgs += t.m;
        t.xoff = xn;
        xn += (long long)t.m * (2 * 4 * KH + 4) * b->inst2[t.inst].nx;
    }
    if (xn > b->xch_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        if (b->d_xch) BCHK(b, hipFree(b->d_xch));
        b->d_xch = nullptr;
        b->xch_cap = 0;
        BCHK(b, hipMalloc(&b->d_xch, sizeof(double) * (size_t)xn));
        b->xch_cap = xn;
    }
    const int grid = std::min(resident, wgs + (int)queue.size());
    if (!teams.empty()) BCHK(b, hipMemcpy(b->d_teams, teams.data(), sizeof(BTeam) * teams.size(), hipMemcpyHostToDevice));
    if (!queue.empty()) BCHK(b, hipMemcpy(b->d_queue, queue.data(), sizeof(int) * queue.size(), hipMemcpyHostToDevice));
    BCHK(b, hipMemsetAsync(b->d_qnext, 0, sizeof(unsigned), b->stream));
    if (!teams.empty()) BCHK(b, hipMemsetAsync(b->d_sync, 0, sizeof(TeamSync) * teams.size(), b->stream));
    BP2 q = b->base2;
    q.t_end = p.t_end;
    q.dt_max = p.dt_max;
    q.max_steps = p.max_steps;
    q.teams = b->d_teams;
    q.nteams = (int)teams.size();
    q.team_wgs = wgs;
    q.queue = b->d_queue;
    q.nqueue = (int)queue.size();
    q.qnext = b->d_qnext;
    q.sync = b->d_sync;
    q.xch = b->d_xch;
    void* args[] = {&q};
    BCHK(b, hipLaunchCooperativeKernel((const void*)k_ensemble2<BS, RECON>, dim3(grid), dim3(BS), args, 0, b->stream));
    b->plan_teams = (int)teams.size();
    b->plan_team_wgs = wgs;
    b->plan_grid = grid;
    return FVG_OK;
}

template <int BS>
int launch_teams2_bs(fvg_batch* b, const BP& p) {
    switch (b->base2.recon) {
    case 1: return launch_teams2<BS, 1>(b, p);
    case 2: return launch_teams2<BS, 2>(b, p);
    default: return launch_teams2<BS, 3>(b, p);
    }
}

// whether a launch takes teams: cooperative launches available, a reconstruction (first-order faces read the
// conservative state of the next chunk, plain memory of another member, which the fence-free team barrier does not
// order) and max_team > 1
bool teamed(const fvg_batch* b, const BP& p) {
    return b->coop && b->max_team > 1 && (b->dim == 2 ? b->base2.recon : p.recon) != 0;
}

int launch_ensemble(fvg_batch* b, const BP& p, int block) {
    if (teamed(b, p) && b->dim == 2) {
        switch (block) {
        case 64: return launch_teams2_bs<64>(b, p);
        case 128: return launch_teams2_bs<128>(b, p);
        case 256: return launch_teams2_bs<256>(b, p);
        case 512: return launch_teams2_bs<512>(b, p);
        case 768: return launch_teams2_bs<768>(b, p);
        default: return launch_teams2_bs<1024>(b, p);
        }
    }
    if (teamed(b, p)) {
        switch (block) {
        case 64: return launch_teams_bs<64>(b, p);
        case 128: return launch_teams_bs<128>(b, p);
        case 256: return launch_teams_bs<256>(b, p);
        case 512: return launch_teams_bs<512>(b, p);
        case 768: return launch_teams_bs<768>(b, p);
        default: return launch_teams_bs<1024>(b, p);
        }
    }
    b->plan_teams = 0;
    b->plan_team_wgs = 0;
    b->plan_grid = b->ninst;
    switch (block) {
    case 64: launch_bs<64>(b, p); break;
    case 128: launch_bs<128>(b, p); break;
    case 256: launch_bs<256>(b, p); break;
    case 512: launch_bs<512>(b, p); break;
    case 768: launch_bs<768>(b, p); break;
    default: launch_bs<1024>(b, p); break;
    }
    BCHK(b, hipGetLastError());
    return FVG_OK;
}

// One timed launch over the uploaded state.
int run_once(fvg_batch* b, double t_end, double dt_max, int max_steps, int block, double* ms) {
    BP p = b->base;
    p.t_end = t_end;
    p.dt_max = dt_max;
    p.max_steps = max_steps;
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return bset(b, FVG_EHIP, "hipEventCreate failed");
    }
    int rc = FVG_OK;
    if (hipEventRecord(e0, b->stream) != hipSuccess) rc = bset(b, FVG_EHIP, "hipEventRecord failed");
    if (!rc) rc = launch_ensemble(b, p, block);
    if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = bset(b, FVG_EHIP, "hipEventRecord failed");
    if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = bset(b, FVG_EHIP, "ensemble kernel failed");
    float f = 0.0f;
    if (!rc && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = bset(b, FVG_EHIP, "hipEventElapsedTime failed");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ms) *ms = f;
    return rc;
}

std::string tune_key(const fvg_batch* b) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) arch = prop.gcnArchName;
    unsigned long long h = 1469598103934665603ull;   // FNV-1a over the instance sizes
    for (int v : b->n) {
        h ^= (unsigned long long)v;
        h *= 1099511628211ull;
    }
    for (const BInst2& in : b->inst2) {   // 2-D instances: their shapes
        h ^= (unsigned long long)in.nx * 1000003ull + (unsigned long long)in.ny;
        h *= 1099511628211ull;
    }
    std::ostringstream s;
    s << arch << "|" << b->ninst << "|" << b->nc << "|" << std::hex << h << std::dec << "|" << b->cfg.reconstruction << "|"
      << b->cfg.limiter << "|" << b->cfg.riemann << "|" << b->cfg.time_scheme;
    return s.str();
}

}   // namespace

extern "C" {

const char* fvg_batch_last_error(const fvg_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int fvg_batch_create(const fvg_config* cfg_in, int ninst, const int* ncells, int device, fvg_batch** out) {
    if (!cfg_in || !ncells || !out) return bset(nullptr, FVG_EINVAL, "null argument");
    *out = nullptr;
    if (ninst <= 0) return bset(nullptr, FVG_EINVAL, "ninst must be positive");
    fvg_config cfg = *cfg_in;
    cfg.nx = ncells[0];
    cfg.ny = 1;
    cfg.nz = 1;
    char msg[256] = {0};
    if (fvg_config_finalize(&cfg, msg, sizeof msg) != FVG_OK) return bset(nullptr, FVG_EINVAL, msg);
    if (cfg.bc[2] != FVG_BC_PERIODIC || cfg.bc[3] != FVG_BC_PERIODIC)
        return bset(nullptr, FVG_EUNSUPPORTED, "an ensemble instance is one periodic row: bc_ymin and bc_ymax must be periodic");
    for (int s = 0; s < 2; ++s) {
        const int bt = cfg.bc[s];
        if (!(bt == FVG_BC_WALL || bt == FVG_BC_NO_SLIP || bt == FVG_BC_TRANSMISSIVE || bt == FVG_BC_OUTLET || bt == FVG_BC_PERIODIC))
            return bset(nullptr, FVG_EUNSUPPORTED, "x boundary of an ensemble must be wall, no_slip, transmissive, outlet or periodic");
    }
    if (cfg.use_viscous) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble runs the Euler equations (use_viscous = false)");
    if (cfg.time_scheme == FVG_TS_HANCOCK) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble runs explicit Euler or SSP Runge-Kutta schemes");
    if (cfg.flux_precision != FVG_FLUX_FP64) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble computes in FP64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bset(nullptr, FVG_EHIP, "no HIP device");
    if (device < 0 || device >= ndev) return bset(nullptr, FVG_EINVAL, "device index out of range");

    fvg_batch* b = new fvg_batch();
    b->cfg = cfg;
    b->device = device;
    b->ninst = ninst;
    b->n.assign(ncells, ncells + ninst);
    b->inst.resize(ninst);
    long long c0 = 0;
    for (int i = 0; i < ninst; ++i) {
        if (ncells[i] < 1 || ncells[i] > (1 << 26)) {
            free_batch(b);
            return bset(nullptr, FVG_EINVAL, "instance " + std::to_string(i) + ": cell count out of range");
        }
        fvg_config ci = cfg;
        ci.nx = ncells[i];
        KP kp{};
        fill_kernel_params(ci, kp);
        BInst& in = b->inst[i];
        in.c0 = c0;
        in.f0 = c0 + i;
        in.n = ncells[i];
        in.halo_lo = kp.halo[0];
        in.halo_hi = kp.halo[1];
        in.pad = 0;
        in.Ax = kp.Ax;
        in.inv_vol = kp.inv_vol;
        in.Ldt = kp.Ldt;
        c0 += ncells[i];
    }
    b->nc = c0;
    b->nf = c0 + ninst;
    b->order.resize(ninst);
    for (int i = 0; i < ninst; ++i) b->order[i] = i;
    auto fail = [&](int code, const std::string& m) {
        free_batch(b);
        return bset(nullptr, code, m);
    };
    if (hipSetDevice(device) != hipSuccess) return fail(FVG_EHIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(FVG_EHIP, "hipStreamCreate failed");
    const size_t cb = sizeof(double) * 3 * (size_t)b->nc, fb = sizeof(double) * 3 * (size_t)b->nf;
    if (hipMalloc(&b->U, cb) != hipSuccess || hipMalloc(&b->Ut, cb) != hipSuccess || hipMalloc(&b->W, cb) != hipSuccess ||
        hipMalloc(&b->Usave, cb) != hipSuccess || hipMalloc(&b->F, fb) != hipSuccess ||
        hipMalloc(&b->d_inst, sizeof(BInst) * ninst) != hipSuccess || hipMalloc(&b->d_order, sizeof(int) * ninst) != hipSuccess ||
        hipMalloc(&b->d_out, sizeof(BOut) * ninst) != hipSuccess || hipMalloc(&b->d_teams, sizeof(BTeam) * ninst) != hipSuccess ||
        hipMalloc(&b->d_queue, sizeof(int) * ninst) != hipSuccess || hipMalloc(&b->d_qnext, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&b->d_sync, sizeof(TeamSync) * ninst) != hipSuccess)
        return fail(FVG_ENOMEM, "hipMalloc of the ensemble buffers");
    int coop = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) == hipSuccess) b->coop = coop != 0;
    b->cost.assign(ninst, 0.0);
    if (hipMemcpy(b->d_inst, b->inst.data(), sizeof(BInst) * ninst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_order, b->order.data(), sizeof(int) * ninst, hipMemcpyHostToDevice) != hipSuccess)
        return fail(FVG_EHIP, "hipMemcpy of the instance table");

    KP kp{};
    fill_kernel_params(cfg, kp);
    BP& p = b->base;
    p.rs = kp.rs;
    p.bp = kp.bp;
    p.recon = cfg.reconstruction;
    p.lim = cfg.limiter;
    p.scheme = cfg.time_scheme == FVG_TS_EULER ? 0 : (cfg.time_scheme == FVG_TS_SSPRK2 ? 1 : 2);
    p.weno_eps = cfg.weno_epsilon;
    p.cfl = cfg.cfl;
    p.nc = b->nc;
    p.nf = b->nf;
    p.U = b->U;
    p.Ut = b->Ut;
    p.W = b->W;
    p.F = b->F;
    p.inst = b->d_inst;
    p.order = b->d_order;
    p.out = b->d_out;
    *out = b;
    return FVG_OK;
}

int fvg_batch_destroy(fvg_batch* b) {
    free_batch(b);
    return FVG_OK;
}

int fvg_batch_create_2d(const fvg_config* cfg_in, int ninst, const int* nx, const int* ny, int device, fvg_batch** out) {
    if (!cfg_in || !nx || !ny || !out) return bset(nullptr, FVG_EINVAL, "null argument");
    *out = nullptr;
    if (ninst <= 0) return bset(nullptr, FVG_EINVAL, "ninst must be positive");
    fvg_config cfg = *cfg_in;
    cfg.nx = nx[0];
    cfg.ny = ny[0];
    cfg.nz = 1;
    char msg[256] = {0};
    if (fvg_config_finalize(&cfg, msg, sizeof msg) != FVG_OK) return bset(nullptr, FVG_EINVAL, msg);
    for (int s = 0; s < 4; ++s) {   // the side types whose ghost states need no other cell of the side
        const int bt = cfg.bc[s];
        if (bt == FVG_BC_MASSFLOW_INLET || bt == FVG_BC_NONREFLECTING)
            return bset(nullptr, FVG_EUNSUPPORTED, "mass-flow inlet and non-reflecting sides are not in the ensemble");
    }
    if (cfg.use_viscous) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble runs the Euler equations (use_viscous = false)");
    if (cfg.time_scheme == FVG_TS_HANCOCK) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble runs explicit Euler or SSP Runge-Kutta schemes");
    if (cfg.flux_precision != FVG_FLUX_FP64) return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble computes in FP64");
    if (cfg.flux_quadrature != FVG_QUAD_MIDPOINT || cfg.indicator != FVG_INDICATOR_NONE)
        return bset(nullptr, FVG_EUNSUPPORTED, "the ensemble takes the midpoint quadrature and no indicator");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bset(nullptr, FVG_EHIP, "no HIP device");
    if (device < 0 || device >= ndev) return bset(nullptr, FVG_EINVAL, "device index out of range");

    fvg_batch* b = new fvg_batch();
    b->cfg = cfg;
    b->device = device;
    b->ninst = ninst;
    b->dim = 2;
    b->nplanes = 4;
    b->n.resize(ninst);
    b->inst2.resize(ninst);
    long long c0 = 0, f0 = 0;
    for (int i = 0; i < ninst; ++i) {
        if (nx[i] < 1 || ny[i] < 1 || (long long)nx[i] * ny[i] > (1 << 24)) {
            free_batch(b);
            return bset(nullptr, FVG_EINVAL, "instance " + std::to_string(i) + ": mesh size out of range");
        }
        fvg_config ci = cfg;
        ci.nx = nx[i];
        ci.ny = ny[i];
        KP kp{};
        fill_kernel_params(ci, kp);
        BInst2& in = b->inst2[i];
        in.c0 = c0;
        in.fx0 = f0;
        in.fy0 = f0 + (long long)(nx[i] + 1) * ny[i];
        in.nx = nx[i];
        in.ny = ny[i];
        for (int s = 0; s < 4; ++s) in.halo[s] = kp.halo[s];
        in.Ax = kp.Ax;
        in.Ay = kp.Ay;
        in.inv_vol = kp.inv_vol;
        in.Ldt = kp.Ldt;
        b->n[i] = nx[i] * ny[i];
        c0 += (long long)nx[i] * ny[i];
        f0 += (long long)(nx[i] + 1) * ny[i] + (long long)nx[i] * (ny[i] + 1);
    }
    b->nc = c0;
    b->nf = f0;
    b->order.resize(ninst);
    for (int i = 0; i < ninst; ++i) b->order[i] = i;
    auto fail = [&](int code, const std::string& m) {
        free_batch(b);
        return bset(nullptr, code, m);
    };
    if (hipSetDevice(device) != hipSuccess) return fail(FVG_EHIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(FVG_EHIP, "hipStreamCreate failed");
    const size_t cb = sizeof(double) * 4 * (size_t)b->nc, fb = sizeof(double) * 4 * (size_t)b->nf;
    if (hipMalloc(&b->U, cb) != hipSuccess || hipMalloc(&b->Ut, cb) != hipSuccess || hipMalloc(&b->W, cb) != hipSuccess ||
        hipMalloc(&b->Usave, cb) != hipSuccess || hipMalloc(&b->F, fb) != hipSuccess ||
        hipMalloc(&b->d_inst2, sizeof(BInst2) * ninst) != hipSuccess || hipMalloc(&b->d_order, sizeof(int) * ninst) != hipSuccess ||
        hipMalloc(&b->d_out, sizeof(BOut) * ninst) != hipSuccess || hipMalloc(&b->d_teams, sizeof(BTeam) * ninst) != hipSuccess ||
        hipMalloc(&b->d_queue, sizeof(int) * ninst) != hipSuccess || hipMalloc(&b->d_qnext, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&b->d_sync, sizeof(TeamSync) * ninst) != hipSuccess)
        return fail(FVG_ENOMEM, "hipMalloc of the ensemble buffers");
    int coop = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) == hipSuccess) b->coop = coop != 0;
    // work per instance, first estimated as cells x the larger side (the steps grow with the resolution); longest first
    b->cost.resize(ninst);
    for (int i = 0; i < ninst; ++i) b->cost[i] = (double)b->n[i] * std::max(nx[i], ny[i]);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return b->cost[a] > b->cost[c]; });
    if (hipMemcpy(b->d_inst2, b->inst2.data(), sizeof(BInst2) * ninst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_order, b->order.data(), sizeof(int) * ninst, hipMemcpyHostToDevice) != hipSuccess)
        return fail(FVG_EHIP, "hipMemcpy of the instance table");
    KP kp{};
    fill_kernel_params(cfg, kp);
    BP2& p = b->base2;
    p.rs = kp.rs;
    p.bp = kp.bp;
    p.recon = cfg.reconstruction;
    p.lim = cfg.limiter;
    p.scheme = cfg.time_scheme == FVG_TS_EULER ? 0 : (cfg.time_scheme == FVG_TS_SSPRK2 ? 1 : 2);
    p.weno_eps = cfg.weno_epsilon;
    p.cfl = cfg.cfl;
    p.nc = b->nc;
    p.nf = b->nf;
    p.U = b->U;
    p.Ut = b->Ut;
    p.W = b->W;
    p.F = b->F;
    p.inst = b->d_inst2;
    p.order = b->d_order;
    p.out = b->d_out;
    *out = b;
    return FVG_OK;
}

int fvg_batch_upload(fvg_batch* b, const double* U, long long ncells, int ncomp) {
    if (!b || !U) return bset(b, FVG_EINVAL, "null argument");
    if (ncells != b->nc) return bset(b, FVG_EINVAL, "cell count differs from the ensemble's total");
    if (b->dim == 2 ? (ncomp != 4 && ncomp != 5) : (ncomp != 3 && ncomp != 4 && ncomp != 5))
        return bset(b, FVG_EINVAL, b->dim == 2 ? "ncomp must be 4 or 5" : "ncomp must be 3, 4 or 5");
    // the planes ((rho, rho u, E), or (rho, rho u, rho v, E) for 2-D instances); the estimated cost of each instance
    // (cells x the cells along its longer axis x the largest initial wave speed) orders the workgroups, longest first
    const int np = b->nplanes, e = ncomp == 3 ? 2 : 3;
    std::vector<double> h((size_t)np * b->nc);
    const double g = b->cfg.gamma, pinf = b->cfg.eos_type == FVG_EOS_STIFFENED ? b->cfg.p_inf : 0.0;
    std::vector<double> cost(b->ninst);
    for (int i = 0; i < b->ninst; ++i) {
        const long long c0 = b->dim == 2 ? b->inst2[i].c0 : b->inst[i].c0;
        const int nc_i = b->n[i];
        const int lmax = b->dim == 2 ? std::max(b->inst2[i].nx, b->inst2[i].ny) : nc_i;
        double smax0 = 0.0;
        for (int k = 0; k < nc_i; ++k) {
            const long long c = c0 + k;
            const double* s = U + (size_t)ncomp * c;
            h[c] = s[0];
            h[b->nc + c] = s[1];
            if (np == 4) h[2 * b->nc + c] = s[2];
            h[(np - 1) * b->nc + c] = s[e];
            const double rho = std::max(s[0], 1e-10), u = s[1] / rho, v = np == 4 ? s[2] / rho : 0.0;
            const double p = (g - 1.0) * (s[e] - 0.5 * rho * (u * u + v * v)) - (pinf != 0.0 ? g * pinf : 0.0);
            const double c2 = g * (std::max(p, 1e-10) + pinf) / rho;
            const double sp = std::sqrt(u * u + v * v) + std::sqrt(std::max(c2, 0.0));
            if (std::isfinite(sp) && sp > smax0) smax0 = sp;
        }
        cost[i] = (double)nc_i * lmax * (smax0 + 1e-3);
    }
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    b->cost = cost;
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipMemcpy(b->U, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_order, b->order.data(), sizeof(int) * b->ninst, hipMemcpyHostToDevice));
    b->uploaded = true;
    return FVG_OK;
}

int fvg_batch_download(fvg_batch* b, double* U, long long ncells, int ncomp) {
    if (!b || !U) return bset(b, FVG_EINVAL, "null argument");
    if (ncells != b->nc) return bset(b, FVG_EINVAL, "cell count differs from the ensemble's total");
    if (b->dim == 2 ? (ncomp != 4 && ncomp != 5) : (ncomp != 3 && ncomp != 4 && ncomp != 5))
        return bset(b, FVG_EINVAL, b->dim == 2 ? "ncomp must be 4 or 5" : "ncomp must be 3, 4 or 5");
    std::vector<double> h((size_t)b->nplanes * b->nc);
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipMemcpy(h.data(), b->U, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (long long c = 0; c < b->nc; ++c) {
        double* s = U + (size_t)ncomp * c;
        s[0] = h[c];
        s[1] = h[b->nc + c];
        if (b->nplanes == 4) {
            s[2] = h[2 * b->nc + c];
            s[3] = h[3 * b->nc + c];
            if (ncomp == 5) s[4] = 0.0;
        } else if (ncomp == 3) {
            s[2] = h[2 * b->nc + c];
        } else {
            s[2] = 0.0;
            s[3] = h[2 * b->nc + c];
            if (ncomp == 5) s[4] = 0.0;
        }
    }
    return FVG_OK;
}

int fvg_batch_run(fvg_batch* b, double t_end, double dt_max, int max_steps, double* t, int* steps, int* diverged, double* s_last,
                  double* ms) {
    if (!b) return bset(b, FVG_EINVAL, "null argument");
    if (!b->uploaded) return bset(b, FVG_ELOGIC, "no state uploaded");
    if (!(max_steps > 0)) return bset(b, FVG_EINVAL, "max_steps must be positive (it bounds the kernel's step loop)");
    if (!std::isfinite(t_end) || !(t_end >= 0.0)) return bset(b, FVG_EINVAL, "t_end must be finite and non-negative");
    BCHK(b, hipSetDevice(b->device));
    double m = 0.0;
    int rc = run_once(b, t_end, dt_max, max_steps, b->block, &m);
    if (rc) return rc;
    b->last_ms = m;
    if (ms) *ms = m;
    std::vector<BOut> o(b->ninst);
    BCHK(b, hipMemcpy(o.data(), b->d_out, sizeof(BOut) * b->ninst, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->ninst; ++i)
        if (o[i].diverged == 2)
            return bset(b, FVG_EHIP, "instance " + std::to_string(i) + ": a workgroup of its team timed out at a team barrier");
    for (int i = 0; i < b->ninst; ++i) {
        if (t) t[i] = o[i].t;
        if (steps) steps[i] = o[i].steps;
        if (diverged) diverged[i] = o[i].diverged;
        if (s_last) s_last[i] = o[i].s_last;
    }
    return FVG_OK;
}

int fvg_batch_set_block(fvg_batch* b, int block) {
    if (!b) return bset(b, FVG_EINVAL, "null argument");
    for (int k : kBlocks)
        if (k == block) {
            b->block = block;
            return FVG_OK;
        }
    return bset(b, FVG_EINVAL, "block must be 64, 128, 256, 512, 768 or 1024");
}

int fvg_batch_set_teams(fvg_batch* b, int max_team, int min_chunk, int force) {
    if (!b) return bset(b, FVG_EINVAL, "null argument");
    if (max_team < 1 || max_team > kMaxTeam) return bset(b, FVG_EINVAL, "max_team must be in 1 .. 16");
    if (min_chunk < KH) return bset(b, FVG_EINVAL, "min_chunk must be at least 3 cells");
    b->max_team = max_team;
    b->min_chunk = min_chunk;
    b->force_teams = force ? 1 : 0;
    return FVG_OK;
}

int fvg_batch_get_plan(const fvg_batch* b, int* nteams, int* team_wgs, int* grid) {
    if (!b) return FVG_EINVAL;
    if (nteams) *nteams = b->plan_teams;
    if (team_wgs) *team_wgs = b->plan_team_wgs;
    if (grid) *grid = b->plan_grid;
    return FVG_OK;
}

int fvg_batch_get_block(const fvg_batch* b, int* block) {
    if (!b || !block) return FVG_EINVAL;
    *block = b->block;
    return FVG_OK;
}

int fvg_batch_tune(fvg_batch* b, double t_probe, int max_steps, const char* cache_path, int* block, int* from_cache) {
    if (!b) return bset(b, FVG_EINVAL, "null argument");
    if (!b->uploaded) return bset(b, FVG_ELOGIC, "no state uploaded");
    if (!(max_steps > 0) || !std::isfinite(t_probe) || !(t_probe > 0.0)) return bset(b, FVG_EINVAL, "t_probe and max_steps must be positive");
    if (from_cache) *from_cache = 0;
    const std::string key = tune_key(b);
    if (cache_path && *cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream s(line);
            std::string k;
            int blk = 0;
            double tms = 0.0;
            if ((s >> k >> blk >> tms) && k == key && fvg_batch_set_block(b, blk) == FVG_OK) {
                if (block) *block = blk;
                if (from_cache) *from_cache = 1;
                return FVG_OK;
            }
        }
    }
    BCHK(b, hipSetDevice(b->device));
    const size_t cb = sizeof(double) * b->nplanes * (size_t)b->nc;
    BCHK(b, hipMemcpyAsync(b->Usave, b->U, cb, hipMemcpyDeviceToDevice, b->stream));
    int best = b->block;
    double best_ms = 1e300;
    for (int k : kBlocks) {
        BCHK(b, hipMemcpyAsync(b->U, b->Usave, cb, hipMemcpyDeviceToDevice, b->stream));
        double m = 0.0;
        const int rc = run_once(b, t_probe, 0.0, max_steps, k, &m);
        if (rc) {   // leave the ensemble as it was uploaded, as on success
            if (hipMemcpyAsync(b->U, b->Usave, cb, hipMemcpyDeviceToDevice, b->stream) == hipSuccess)
                (void)hipStreamSynchronize(b->stream);
            return rc;
        }
        if (m < best_ms) {
            best_ms = m;
            best = k;
        }
    }
    // the probe's step counts measure each instance's work better than the initial wave speed does (its waves form
    // within the first steps): the queue order and the teams of later runs take N x probe steps as the work
    std::vector<BOut> o(b->ninst);
    BCHK(b, hipMemcpy(o.data(), b->d_out, sizeof(BOut) * b->ninst, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->ninst; ++i) b->cost[i] = (double)b->n[i] * std::max(1, o[i].steps);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return b->cost[a] > b->cost[c]; });
    BCHK(b, hipMemcpy(b->d_order, b->order.data(), sizeof(int) * b->ninst, hipMemcpyHostToDevice));
    BCHK(b, hipMemcpyAsync(b->U, b->Usave, cb, hipMemcpyDeviceToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->block = best;
    if (block) *block = best;
    if (cache_path && *cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        if (f) f << key << " " << best << " " << best_ms << "\n";
    }
    return FVG_OK;
}

}   // extern "C"
