// flux_quadrature = fourth on the device (FP64, the reference's bits): fourth_order_fluxes and the conserved
// reconstruction it takes (Schemes.cpp:1210-1349, 1954-2134), as three passes over the faces of one axis at a time.
//
//   k4_states<D>  WENO3-Z / WENO5-Z of the padded conservative state Wc at every face with a full stencil, the
//                 admissibility check (positive states inside the widened range of the stencil cells, else the two
//                 cells next to the face); a face without a full stencil takes its first-order boundary state as
//                 it is. Flags the faces that are good: a full stencil, both states positive, the shock sensor quiet.
//   k4_points<D>  point values A = U - (0 + sum over the transverse axes of (U[+1] - 2 U + U[-1])) / 24 where the
//                 face and its transverse neighbours are good and A stays positive (flag pt), then the Riemann flux
//                 Fp of the point values (of the averages elsewhere); its block maxima give the wave speed.
//   k4_fluxes<D>  F = Fp + (0 + sum of (Fp[+1] - 2 Fp + Fp[-1])) / 24 where the face and its transverse
//                 neighbours have point values, the flux of the averaged states where only the face has, else Fp;
//                 stored with the viscous face flux for k_update, as k_faces stores its fluxes.
//
// Transverse neighbours are faces of the same axis one lattice step away; across a periodic axis they wrap to the
// other side (the reference's `periodic`: both sides of the axis periodic, whatever its halo), else the face has no
// correction. The three passes exchange the face states, flags and fluxes through q4_* buffers sized for the
// largest axis and reused axis after axis.
#include "fvg_kernels_impl.hpp"

namespace fvg {

namespace {

// p > 0 and rho > 0 without a division, and the pressure (Schemes.cpp:1241-1243, 1287, 1966-1969). With NV = 4
// the absent rho w adds +0 to |m|^2, which leaves it as it is.
template <int NV> __device__ __forceinline__ double q4_m2(const double* u) {
    double m = u[1] * u[1] + u[2] * u[2];
    if constexpr (NV == 5) m = m + u[4] * u[4];
    return m;
}
template <int NV> __device__ __forceinline__ bool q4_positive(double g1, double gp, const double* u) {
    return u[0] > 0.0 && g1 * (u[3] * u[0] - 0.5 * q4_m2<NV>(u)) > gp * u[0];
}
template <int NV> __device__ __forceinline__ double q4_pressure(double g1, double gp, const double* u) {
    return g1 * (u[3] - 0.5 * q4_m2<NV>(u) / u[0]) - gp;
}

// the face one lattice step s (-1, +1) from face f along transverse axis t, -1 when there is none
// (fourth_order_fluxes' nbr, Schemes.cpp:2022-2027)
template <int D>
__device__ __forceinline__ long long q4_nbr(const KP& p, long long f, int ct, int t, int s) {
    const long long ni = p.nx + (D == 0), nj = p.ny + (D == 1);
    const long long st = t == 0 ? 1 : (t == 1 ? ni : ni * nj);
    const int n = ncell(p, t);
    const int c = ct + s;
    if (c >= 0 && c < n) return f + s * st;
    if (!vper_of(p, t)) return -1;
    return f - s * (long long)(n - 1) * st;
}

// the transverse axes of the faces normal to D, in increasing order
template <int NV, int D> struct Trans {
    static constexpr int n = NV - 3;   // 1 in 2-D, 2 in 3-D
    static constexpr int t0 = D == 0 ? 1 : 0;
    static constexpr int t1 = D == 2 ? 1 : 2;
};

template <int NV>
__device__ __forceinline__ void q4_load(const double* planes, long long n, long long f, double* u) {
#pragma unroll
    for (int q = 0; q < NV; ++q) u[q] = planes[q * n + f];
}

}   // namespace

// R: the reconstruction, 2 WENO3-Z (four cells t-2 .. t+1), 3 WENO5-Z (six cells t-3 .. t+2)
// Q6: WENO5-Z's quotients by 6 in three operations (quot6; every value of Wc in range_fits_quot6, the same bits)
template <int NV, int D, int R, bool Q6>
__device__ __forceinline__ void k4_states_face(const KP& p, long long f) {
    int i, j, k, t;
    face_coords<NV, D>(p, f, i, j, k, t);
    const int n = ncell(p, D);
    const bool lo = t == 0, hi = t == n;
    const bool upgrade = (t > 0 && t < n) || (lo && halo_of(p, 2 * D)) || (hi && halo_of(p, 2 * D + 1));
    const double g1 = p.rs.eos.gamma - 1.0, gp = p.rs.eos.stiff ? p.rs.eos.gamma * p.rs.eos.p_inf : 0.0;
    double ul[NV], ur[NV];
    if (upgrade) {
        const double we = p.weno_eps;
        constexpr int NS = R == 3 ? 6 : 4;
        double c[NS][NV];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const long long pc = pidx_d(p, D, i, j, k, t - NS / 2 + s);
#pragma unroll
            for (int q = 0; q < NV; ++q) c[s][q] = p.Wc[q * p.npad + pc];
        }
        if constexpr (R == 3) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                ul[q] = weno5z<double, Q6>(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], we);
                ur[q] = weno5z<double, Q6>(c[5][q], c[4][q], c[3][q], c[2][q], c[1][q], we);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                ul[q] = weno3z(c[0][q], c[1][q], c[2][q], we);
                ur[q] = weno3z(c[3][q], c[2][q], c[1][q], we);
            }
        }
        // admissible: positive, and the density and pressure inside the range of the stencil cells widened by the
        // range's own width (ConservedLines::line, Schemes.cpp:1288-1318)
        bool ok = q4_positive<NV>(g1, gp, ul) && q4_positive<NV>(g1, gp, ur);
        if (ok) {
            double rlo = 1e300, rhi = -1e300, plo = 1e300, phi = -1e300;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double pc = q4_pressure<NV>(g1, gp, c[s]);
                rlo = smin(rlo, c[s][0]);
                rhi = smax(rhi, c[s][0]);
                plo = smin(plo, pc);
                phi = smax(phi, pc);
            }
            const double rw = rhi - rlo, pw = phi - plo;
            rlo -= rw;
            rhi += rw;
            plo -= pw;
            phi += pw;
            const double pl = q4_pressure<NV>(g1, gp, ul), pr = q4_pressure<NV>(g1, gp, ur);
            ok = ul[0] >= rlo && ul[0] <= rhi && pl >= plo && pl <= phi;
            ok = ok && ur[0] >= rlo && ur[0] <= rhi && pr >= plo && pr <= phi;
        }
        if (!ok) {   // the cells on both sides: stencil positions 2, 3 of the six, 1, 2 of the four
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                ul[q] = c[NS / 2 - 1][q];
                ur[q] = c[NS / 2][q];
            }
        }
    } else {   // a side without a halo: the first-order boundary states, not converted
        const int il = D == 0 ? t - 1 : i, jl = D == 1 ? t - 1 : j, kl = D == 2 ? t - 1 : k;
        if (!lo) load_cons<NV>(p, cidx(p, il, jl, kl), ul);
        else outside_first_order<NV>(p, D, 2 * D, i, j, k, ul);
        if (!hi) load_cons<NV>(p, cidx(p, i, j, k), ur);
        else outside_first_order<NV>(p, D, 2 * D + 1, il, jl, kl, ur);
    }
    // good: a full stencil, positive states, the shock sensor quiet (Schemes.cpp:1987-2005)
    bool good = upgrade && q4_positive<NV>(g1, gp, ul) && q4_positive<NV>(g1, gp, ur);
    if (good && p.rs.blend > 0.0) {
        const double pL = q4_pressure<NV>(g1, gp, ul), pR = q4_pressure<NV>(g1, gp, ur);
        if (fabs(pR - pL) > p.rs.blend * smin(pL, pR)) good = false;
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        p.q4_UL[q * p.q4_n + f] = ul[q];
        p.q4_UR[q * p.q4_n + f] = ur[q];
    }
    p.q4_good[f] = good ? 1 : 0;
}

template <int NV, int D, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k4_states(KP p) {
    const long long f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= p.nfa[D] || p.st->done) return;
    if constexpr (R == 3) {   // uniform over the launch: quot6 unless a value of Wc is outside its range
        if (__builtin_amdgcn_readfirstlane(p.st->wide) == 0) k4_states_face<NV, D, R, true>(p, f);
        else k4_states_face<NV, D, R, false>(p, f);
    } else {
        k4_states_face<NV, D, R, false>(p, f);
    }
}

template <int NV, int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k4_points(KP p) {
    using TR = Trans<NV, D>;
    const long long f = xcd_block() * blockDim.x + threadIdx.x;
    double local = 0.0;
    if (f < p.nfa[D] && !p.st->done) {
        int i, j, k, t;
        face_coords<NV, D>(p, f, i, j, k, t);
        double al[NV], ar[NV];
        q4_load<NV>(p.q4_UL, p.q4_n, f, al);
        q4_load<NV>(p.q4_UR, p.q4_n, f, ar);
        bool pt = false;
        if (p.q4_good[f]) {
            long long nb[2][2];
            const int ct0 = coord(TR::t0, i, j, k), ct1 = coord(TR::t1, i, j, k);
            nb[0][0] = q4_nbr<D>(p, f, ct0, TR::t0, -1);
            nb[0][1] = q4_nbr<D>(p, f, ct0, TR::t0, 1);
            bool ok = nb[0][0] >= 0 && nb[0][1] >= 0 && p.q4_good[nb[0][0]] && p.q4_good[nb[0][1]];
            if constexpr (TR::n == 2) {
                nb[1][0] = q4_nbr<D>(p, f, ct1, TR::t1, -1);
                nb[1][1] = q4_nbr<D>(p, f, ct1, TR::t1, 1);
                ok = ok && nb[1][0] >= 0 && nb[1][1] >= 0 && p.q4_good[nb[1][0]] && p.q4_good[nb[1][1]];
            }
            if (ok) {
                double pl[NV], pr[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const double* L = p.q4_UL + q * p.q4_n;
                    const double* R = p.q4_UR + q * p.q4_n;
                    double dl = 0.0, dr = 0.0;
#pragma unroll
                    for (int u = 0; u < TR::n; ++u) {
                        dl += L[nb[u][1]] - 2.0 * al[q] + L[nb[u][0]];
                        dr += R[nb[u][1]] - 2.0 * ar[q] + R[nb[u][0]];
                    }
                    pl[q] = al[q] - dl / 24.0;
                    pr[q] = ar[q] - dr / 24.0;
                }
                const double g1 = p.rs.eos.gamma - 1.0, gp = p.rs.eos.stiff ? p.rs.eos.gamma * p.rs.eos.p_inf : 0.0;
                if (q4_positive<NV>(g1, gp, pl) && q4_positive<NV>(g1, gp, pr)) {
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        al[q] = pl[q];
                        ar[q] = pr[q];
                    }
                    pt = true;
                }
            }
        }
        double F[NV], sm;
        face_flux<NV>(p.rs, al, ar, D, F, sm);
        if (sm > local) local = sm;
#pragma unroll
        for (int q = 0; q < NV; ++q) p.q4_Fp[q * p.q4_n + f] = F[q];
        p.q4_pt[f] = pt ? 1 : 0;
    }
    block_max_slot(p, local, SPD_CONV);
}

template <int NV, int D>
__global__ __launch_bounds__(256) void k4_fluxes(KP p) {
    using TR = Trans<NV, D>;
    const long long f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= p.nfa[D] || p.st->done) return;
    int i, j, k, t;
    face_coords<NV, D>(p, f, i, j, k, t);
    double F[NV];
    q4_load<NV>(p.q4_Fp, p.q4_n, f, F);
    if (p.q4_pt[f]) {
        long long nb[2][2];
        const int ct0 = coord(TR::t0, i, j, k), ct1 = coord(TR::t1, i, j, k);
        nb[0][0] = q4_nbr<D>(p, f, ct0, TR::t0, -1);
        nb[0][1] = q4_nbr<D>(p, f, ct0, TR::t0, 1);
        bool full = nb[0][0] >= 0 && nb[0][1] >= 0 && p.q4_pt[nb[0][0]] && p.q4_pt[nb[0][1]];
        if constexpr (TR::n == 2) {
            nb[1][0] = q4_nbr<D>(p, f, ct1, TR::t1, -1);
            nb[1][1] = q4_nbr<D>(p, f, ct1, TR::t1, 1);
            full = full && nb[1][0] >= 0 && nb[1][1] >= 0 && p.q4_pt[nb[1][0]] && p.q4_pt[nb[1][1]];
        }
        if (full) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const double* P = p.q4_Fp + q * p.q4_n;
                double d = 0.0;
#pragma unroll
                for (int u = 0; u < TR::n; ++u) d += P[nb[u][1]] - 2.0 * F[q] + P[nb[u][0]];
                F[q] = F[q] + d / 24.0;
            }
        } else {   // a point value without the full correction: the flux of the averaged states
            double ul[NV], ur[NV], sm;
            q4_load<NV>(p.q4_UL, p.q4_n, f, ul);
            q4_load<NV>(p.q4_UR, p.q4_n, f, ur);
            face_flux<NV>(p.rs, ul, ur, D, F, sm);
        }
    }
    store_face<NV, D, double>(p, f, i, j, k, t, ncell(p, D), F);
}

#define FVG_INST_Q4(NV, D)                              \
    template __global__ void k4_states<NV, D, 2>(KP); \
    template __global__ void k4_states<NV, D, 3>(KP); \
    template __global__ void k4_points<NV, D>(KP);    \
    template __global__ void k4_fluxes<NV, D>(KP);
FVG_INST_Q4(4, 0)
FVG_INST_Q4(4, 1)
FVG_INST_Q4(5, 0)
FVG_INST_Q4(5, 1)
FVG_INST_Q4(5, 2)

}   // namespace fvg
