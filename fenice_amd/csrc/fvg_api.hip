// Host side of the C ABI (include/fvg.h): a context owns one device, one stream and the device buffers of one
// mesh; residual / step / advance enqueue the kernels of fvg_kernels_impl.hpp and keep time, dt and the divergence flag
// on the device. Every HIP call is checked; failures come back as status codes with a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include <cstddef>

#include "../../include/fvg.h"
#include "fvg_host.hpp"

namespace fvg {
template <int NV, typename T> __global__ void k_cells(KP);
template <int NV> __global__ void k_hancock(KP);
template <int NV, int D, int MODE, typename T> __global__ void k_faces(KP);
template <int NV, int UPD, typename T> __global__ void k_update(KP, const double*, const double*, double*);
template <int NV, int WHAT> __global__ void k_totals_partial(const double*, long long, double*);
template <int NV, int D, int USE, typename T> __global__ void k_faces_learned(KP);
template <int NV, int MODE, int USE, typename T> __global__ void k_xacc(KP);
template <int NV, int D, int MODE, int USE, int UPD, typename T> __global__ void k_lacc(KP, const double*, const double*, double*);
template <int NV> __global__ void k_axpy(KP, const double*, const double*, double*);
template <int NV, int D, typename T> __global__ void k_audit(KP);
template <int NV, int D, typename T> __global__ void k_samples(KP, float*, float*);
template <int NV, int D, int R> __global__ void k4_states(KP);
template <int NV, int D> __global__ void k4_points(KP);
template <int NV, int D> __global__ void k4_fluxes(KP);
__global__ void k_indicator_control(KP, int);
__global__ void k_totals_final(const double*, int, long long, int, double*);
__global__ void k_control(KP, int, int);
__global__ void k_massflow(KP);
template <int NV, int SIDE> __global__ void k_nrbc(KP);
__global__ void k_riemann_batch(RiemannParams, const double*, const double*, const int*, long long, double*, double*);
}   // namespace fvg

using namespace fvg;

namespace {

constexpr int kBlock = 256;
static_assert(kBlock >= kRowBlock, "block sizes of the kernels' launch bounds");
constexpr int kGraphSteps = 4;   // steps per captured graph of fvg_advance
thread_local std::string g_create_error;

enum Family { FAM_CELLS, FAM_HANCOCK, FAM_FACES_X, FAM_FACES_Y, FAM_FACES_Z, FAM_UPDATE, FAM_CONTROL, FAM_AXPY, FAM_AUDIT,
              FAM_Q4_STATES, FAM_Q4_POINTS, FAM_Q4_FLUXES, FAM_N };
const char* kFamilyNames[FAM_N] = {"k_cells", "k_hancock", "k_faces_x", "k_faces_y", "k_faces_z", "k_update", "k_control", "k_axpy",
                                   "k_audit", "k4_states", "k4_points", "k4_fluxes"};

struct Err {
    int code;
    std::string msg;
};

}   // namespace

struct fvg_ctx {
    fvg_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    int NV = 5, ND = 3;
    long long nc = 0, nfa[3] = {0, 0, 0};
    double *U = nullptr, *Ut = nullptr, *Rb = nullptr, *HS = nullptr;
    void *W = nullptr, *G = nullptr;                                          // double, or float with fp32 fluxes
    void *Fc[3] = {nullptr, nullptr, nullptr}, *Fv[3] = {nullptr, nullptr, nullptr};
    bool fp32 = false;
    double *tot_part = nullptr, *tot_out = nullptr;
    DevState* st = nullptr;
    unsigned long long* spd = nullptr;   // wave-speed slots (KP::spd)
    DevState* st_host = nullptr;   // pinned mirror for the read-back
    KP base{};
    bool lagged_positive = false;
    std::string err;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    struct Rec {
        int fam;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    double fam_ms[FAM_N] = {};
    long long fam_n[FAM_N] = {};
    bool learned = false;              // learned troubled-face indicator
    unsigned char* ind_kept[3] = {nullptr, nullptr, nullptr};   // KP::ind_kept
    MlpDev* mlp = nullptr;
    IndState* ind = nullptr;
    bool graphs = true;                // fvg_advance replays a captured graph of steps
    bool fused = false;                // fused residual passes (k_xacc, k_lacc; no face fluxes in HBM, opt-in), or
                                       // separate face and update passes (k_faces, k_update; the default)
    double* A = nullptr;               // KP::A
    hipGraphExec_t graph = nullptr;    // kGraphSteps steps of the scheme
    bool quad4 = false;                // flux_quadrature = fourth
    bool massflow = false;             // a mass-flow inlet side: k_massflow runs before every k_cells
    int nrbc_sides = 0;                // bit s: side s is non-reflecting (k_nrbc runs for it before every k_cells)
    double* nrbc = nullptr;            // KP::nrbc
    double *Wc = nullptr, *q4_UL = nullptr, *q4_UR = nullptr, *q4_Fp = nullptr;
    unsigned char *q4_good = nullptr, *q4_pt = nullptr;
};

namespace {

int set_err(fvg_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    else g_create_error = m;
    return code;
}

#define HIPCHK(ctx, call)                                                                                         \
    do {                                                                                                          \
        hipError_t e_ = (call);                                                                                   \
        if (e_ != hipSuccess) return set_err(ctx, FVG_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// elements of one padded row of the primitive planes: the interior from PX0, PH halo cells on either side, rounded
// up to a multiple of PXA
int padded_row(int nx) { return (PX0 + nx + PH + PXA - 1) / PXA * PXA; }
// blocks of a row-block kernel (k_cells, k_update, z faces) over nplanes planes of ny rows
unsigned row_grid(const fvg_ctx* c, int nplanes) { return (unsigned)((long long)nplanes * c->base.ny * c->base.nchunk); }
unsigned cells_grid(const fvg_ctx* c) { return row_grid(c, c->base.nz); }
unsigned faces_grid(const fvg_ctx* c, int d) { return d == 2 ? row_grid(c, c->base.nz + 1) : grid_for(c->nfa[d]); }
unsigned cells_block(const fvg_ctx* c) { return (unsigned)c->base.rbt; }
unsigned faces_block(const fvg_ctx* c, int d) { return d == 2 ? (unsigned)c->base.rbt : (unsigned)kBlock; }

hipEvent_t take_event(fvg_ctx* c) {
    hipEvent_t e;
    if (!c->ev_pool.empty()) {
        e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch with optional event timing of the kernel family.
template <typename F>
int launch(fvg_ctx* c, int fam, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c->timing) {
        a = take_event(c);
        b = take_event(c);
        if (a) hipEventRecord(a, c->stream);
    }
    f();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, FVG_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (c->timing && a && b) {
        hipEventRecord(b, c->stream);
        c->recs.push_back({fam, a, b});
    }
    return FVG_OK;
}

int harvest_timing(fvg_ctx* c) {
    for (auto& r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->fam_ms[r.fam] += ms;
            c->fam_n[r.fam] += 1;
        }
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->recs.clear();
    return FVG_OK;
}

KP make_kp(fvg_ctx* c, const double* Uin) {
    KP p = c->base;
    p.U = Uin;
    return p;
}

// the mass-flow targets and the non-reflecting ghosts of the stage input, which the ghost states of k_cells and the
// face passes read
int enqueue_massflow(fvg_ctx* c, const KP& p) {
    int rc = FVG_OK;
    if (c->massflow && (rc = launch(c, FAM_CONTROL, [&] { hipLaunchKernelGGL(k_massflow, dim3(6), dim3(1024), 0, c->stream, p); })))
        return rc;
    for (int s = 0; s < 6 && !rc; ++s) {
        if (!(c->nrbc_sides >> s & 1)) continue;
        const long long n = s < 2 ? (long long)c->cfg.ny * c->cfg.nz : (s < 4 ? (long long)c->cfg.nx * c->cfg.nz : (long long)c->cfg.nx * c->cfg.ny);
        const dim3 g(grid_for(n)), b(kBlock);
        rc = launch(c, FAM_CONTROL, [&] {
            if (c->NV == 4) {
                switch (s) {
                case 0: hipLaunchKernelGGL((k_nrbc<4, 0>), g, b, 0, c->stream, p); break;
                case 1: hipLaunchKernelGGL((k_nrbc<4, 1>), g, b, 0, c->stream, p); break;
                case 2: hipLaunchKernelGGL((k_nrbc<4, 2>), g, b, 0, c->stream, p); break;
                default: hipLaunchKernelGGL((k_nrbc<4, 3>), g, b, 0, c->stream, p); break;
                }
            } else {
                switch (s) {
                case 0: hipLaunchKernelGGL((k_nrbc<5, 0>), g, b, 0, c->stream, p); break;
                case 1: hipLaunchKernelGGL((k_nrbc<5, 1>), g, b, 0, c->stream, p); break;
                case 2: hipLaunchKernelGGL((k_nrbc<5, 2>), g, b, 0, c->stream, p); break;
                case 3: hipLaunchKernelGGL((k_nrbc<5, 3>), g, b, 0, c->stream, p); break;
                case 4: hipLaunchKernelGGL((k_nrbc<5, 4>), g, b, 0, c->stream, p); break;
                default: hipLaunchKernelGGL((k_nrbc<5, 5>), g, b, 0, c->stream, p); break;
                }
            }
        });
    }
    return rc;
}

// reuse (learned indicator, Runge-Kutta stages 2 and 3): the faces take the decisions the network made on the step's
// first stage instead of running it again
template <int NV, typename T>
int enqueue_residual_nv(fvg_ctx* c, const double* Uin, int mode, bool audit, bool reuse) {
    const KP p = make_kp(c, Uin);
    int rc;
    if ((rc = enqueue_massflow(c, p))) return rc;
    if ((rc = launch(c, FAM_CELLS, [&] { hipLaunchKernelGGL((k_cells<NV, T>), dim3(cells_grid(c)), dim3(cells_block(c)), 0, c->stream, p); })))
        return rc;
    if (mode == MODE_LEARNED) {
#define LFACES(KERNEL, FAM, D, GRID, BLOCK) \
    launch(c, FAM, [&] { hipLaunchKernelGGL((KERNEL<NV, D, T>), dim3(GRID), dim3(BLOCK), 0, c->stream, p); })
#define LEARNED(FAM, D)                                                                                                            \
    launch(c, FAM, [&] {                                                                                                         \
        if (reuse) hipLaunchKernelGGL((k_faces_learned<NV, D, LF_KEPT, T>), dim3(faces_grid(c, D)), dim3(faces_block(c, D)), 0, c->stream, p); \
        else hipLaunchKernelGGL((k_faces_learned<NV, D, LF_STORE, T>), dim3(faces_grid(c, D)), dim3(faces_block(c, D)), 0, c->stream, p);     \
    })
        if (audit) {
            if ((rc = LFACES(k_audit, FAM_AUDIT, 0, grid_for(c->nfa[0]), kBlock)) || (rc = LFACES(k_audit, FAM_AUDIT, 1, grid_for(c->nfa[1]), kBlock)))
                return rc;
            if constexpr (NV == 5)
                if ((rc = LFACES(k_audit, FAM_AUDIT, 2, grid_for(c->nfa[2]), kBlock))) return rc;
        }
        if ((rc = LEARNED(FAM_FACES_X, 0)) || (rc = LEARNED(FAM_FACES_Y, 1))) return rc;
        if constexpr (NV == 5)
            if ((rc = LEARNED(FAM_FACES_Z, 2))) return rc;
#undef LEARNED
#undef LFACES
        return FVG_OK;
    }
    if constexpr (std::is_same<T, double>::value) {
        if (mode == MODE_FOURTH) {   // per axis: face states and flags, point values and their fluxes, corrected fluxes
            const bool w5 = c->cfg.reconstruction == FVG_RECON_WENO5;
#define Q4(KERNEL, FAM, D) \
    launch(c, FAM, [&] { hipLaunchKernelGGL((KERNEL<NV, D>), dim3(grid_for(c->nfa[D])), dim3(kBlock), 0, c->stream, p); })
#define Q4_STATES(D)                                                                                              \
    launch(c, FAM_Q4_STATES, [&] {                                                                                \
        if (w5) hipLaunchKernelGGL((k4_states<NV, D, 3>), dim3(grid_for(c->nfa[D])), dim3(kBlock), 0, c->stream, p); \
        else hipLaunchKernelGGL((k4_states<NV, D, 2>), dim3(grid_for(c->nfa[D])), dim3(kBlock), 0, c->stream, p);    \
    })
#define Q4_AXIS(D) \
    if ((rc = Q4_STATES(D)) || (rc = Q4(k4_points, FAM_Q4_POINTS, D)) || (rc = Q4(k4_fluxes, FAM_Q4_FLUXES, D))) return rc;
            Q4_AXIS(0)
            Q4_AXIS(1)
            if constexpr (NV == 5) {
                Q4_AXIS(2)
            }
#undef Q4_AXIS
#undef Q4_STATES
#undef Q4
            return FVG_OK;
        }
    }
    if (mode == MODE_HANCOCK)
        if ((rc = launch(c, FAM_HANCOCK, [&] { hipLaunchKernelGGL(k_hancock<NV>, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, p); })))
            return rc;
#define FACES(D, M)                                                                                                                       \
    launch(c, FAM_FACES_X + D, [&] { hipLaunchKernelGGL((k_faces<NV, D, M, T>), dim3(faces_grid(c, D)), dim3(faces_block(c, D)), 0, c->stream, p); })
#define AXIS(D) (mode == MODE_PLAIN ? FACES(D, MODE_PLAIN) : FACES(D, MODE_FIRST_ORDER))
    if ((rc = AXIS(0))) return rc;
    if ((rc = AXIS(1))) return rc;
    if constexpr (NV == 5)
        if ((rc = AXIS(2))) return rc;
#undef AXIS
#undef FACES
    return FVG_OK;
}

// MUSCL-Hancock faces (FP64 only)
template <int NV>
int enqueue_hancock_faces(fvg_ctx* c, const KP& p) {
    int rc;
#define HFACES(D)                                                                                                                         \
    launch(c, FAM_FACES_X + D, [&] { hipLaunchKernelGGL((k_faces<NV, D, MODE_HANCOCK, double>), dim3(faces_grid(c, D)), dim3(faces_block(c, D)), 0, \
                                                        c->stream, p); })
    if ((rc = HFACES(0))) return rc;
    if ((rc = HFACES(1))) return rc;
    if constexpr (NV == 5)
        if ((rc = HFACES(2))) return rc;
#undef HFACES
    return FVG_OK;
}

template <int NV>
int enqueue_hancock_residual(fvg_ctx* c, const double* Uin) {
    const KP p = make_kp(c, Uin);
    int rc;
    if ((rc = enqueue_massflow(c, p))) return rc;
    if ((rc = launch(c, FAM_CELLS, [&] { hipLaunchKernelGGL((k_cells<NV, double>), dim3(cells_grid(c)), dim3(cells_block(c)), 0, c->stream, p); })))
        return rc;
    if ((rc = launch(c, FAM_HANCOCK, [&] { hipLaunchKernelGGL(k_hancock<NV>, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, p); })))
        return rc;
    return enqueue_hancock_faces<NV>(c, p);
}

int enqueue_residual(fvg_ctx* c, const double* Uin, int mode, bool audit = false, bool reuse = false) {
    if (mode == MODE_HANCOCK)
        return c->NV == 5 ? enqueue_hancock_residual<5>(c, Uin) : enqueue_hancock_residual<4>(c, Uin);
    if (c->fp32)
        return c->NV == 5 ? enqueue_residual_nv<5, float>(c, Uin, mode, audit, reuse) : enqueue_residual_nv<4, float>(c, Uin, mode, audit, reuse);
    return c->NV == 5 ? enqueue_residual_nv<5, double>(c, Uin, mode, audit, reuse) : enqueue_residual_nv<4, double>(c, Uin, mode, audit, reuse);
}

int enqueue_indicator(fvg_ctx* c, int op) {
    const KP p = c->base;
    return launch(c, FAM_CONTROL, [&] { hipLaunchKernelGGL(k_indicator_control, dim3(1), dim3(1), 0, c->stream, p, op); });
}

int enqueue_control(fvg_ctx* c, int op, int arg) {
    const KP p = c->base;
    return launch(c, FAM_CONTROL, [&] { hipLaunchKernelGGL(k_control, dim3(1), dim3(kSpeedSlots), 0, c->stream, p, op, arg); });
}

template <int NV, typename T>
int enqueue_update_nv(fvg_ctx* c, int upd, const double* U0, const double* Ut, double* out) {
    const KP p = c->base;
    const dim3 g(cells_grid(c)), b(cells_block(c));
    return launch(c, FAM_UPDATE, [&] {
        switch (upd) {
        case UPD_RESIDUAL: hipLaunchKernelGGL((k_update<NV, UPD_RESIDUAL, T>), g, b, 0, c->stream, p, U0, Ut, out); break;
        case UPD_AXPY: hipLaunchKernelGGL((k_update<NV, UPD_AXPY, T>), g, b, 0, c->stream, p, U0, Ut, out); break;
        case UPD_RK2_2: hipLaunchKernelGGL((k_update<NV, UPD_RK2_2, T>), g, b, 0, c->stream, p, U0, Ut, out); break;
        case UPD_RK3_2: hipLaunchKernelGGL((k_update<NV, UPD_RK3_2, T>), g, b, 0, c->stream, p, U0, Ut, out); break;
        default: hipLaunchKernelGGL((k_update<NV, UPD_RK3_3, T>), g, b, 0, c->stream, p, U0, Ut, out); break;
        }
    });
}
int enqueue_update(fvg_ctx* c, int upd, const double* U0, const double* Ut, double* out) {
    if (c->fp32) return c->NV == 5 ? enqueue_update_nv<5, float>(c, upd, U0, Ut, out) : enqueue_update_nv<4, float>(c, upd, U0, Ut, out);
    return c->NV == 5 ? enqueue_update_nv<5, double>(c, upd, U0, Ut, out) : enqueue_update_nv<4, double>(c, upd, U0, Ut, out);
}

// ---- fused residual passes (KP::A) ----
// the (mode, use, upd) combinations the device units instantiate: MUSCL-Hancock and the learned network's first
// stage only store R (UPD_RESIDUAL; the update is k_axpy), the kept decisions of the learned indicator only serve
// the stages whose dt is known; upd < 0 is the y pass of a 3-D mesh
template <int MODE, int USE, int UPD> constexpr bool fused_combo() {
    if (MODE == MODE_HANCOCK) return USE == 0 && UPD <= UPD_RESIDUAL;
    if (MODE == MODE_LEARNED) return USE == LF_STORE ? UPD <= UPD_RESIDUAL : UPD != UPD_RESIDUAL;
    return USE == 0;
}

unsigned xacc_grid(const fvg_ctx* c) {   // 4 wavefronts per block, 63 cells per wavefront
    const long long waves = (c->nfa[0] - 1 + 62) / 63;
    return (unsigned)((waves + 3) / 4);
}
unsigned lacc_grid(const fvg_ctx* c, int d) {   // 4 wavefronts per block on 64 x positions of one line segment
    return (unsigned)((long long)c->base.nlchunk * c->base.nseg[d] * (d == 1 ? c->base.nz : c->base.ny));
}

template <int NV, int D, int MODE, int USE, int UPD, typename T>
void launch_lacc(fvg_ctx* c, const KP& p, const double* U0, const double* Ut, double* out) {
    if constexpr (fused_combo<MODE, USE, UPD>())
        hipLaunchKernelGGL((k_lacc<NV, D, MODE, USE, UPD, T>), dim3(lacc_grid(c, D)), dim3(kBlock), 0, c->stream, p, U0, Ut, out);
}

// x faces into the accumulators, y faces added (3-D), then the last axis finishes R: stored into out (UPD_RESIDUAL)
// or combined into the stage output (dt known)
template <int NV, int MODE, int USE, typename T>
int enqueue_fused_faces(fvg_ctx* c, const KP& p, int upd, const double* U0, const double* Ut, double* out) {
    constexpr int DL = NV == 5 ? 2 : 1;
    if (!((upd == UPD_RESIDUAL && fused_combo<MODE, USE, UPD_RESIDUAL>()) || (upd != UPD_RESIDUAL && fused_combo<MODE, USE, UPD_RK3_3>())))
        return set_err(c, FVG_ELOGIC, "no fused pass for this mode and stage");
    int rc;
    if ((rc = launch(c, FAM_FACES_X, [&] {
             hipLaunchKernelGGL((k_xacc<NV, MODE, USE, T>), dim3(xacc_grid(c)), dim3(kBlock), 0, c->stream, p);
         })))
        return rc;
    if constexpr (NV == 5)
        if ((rc = launch(c, FAM_FACES_Y, [&] { launch_lacc<NV, 1, MODE, USE, -1, T>(c, p, nullptr, nullptr, nullptr); }))) return rc;
    return launch(c, DL == 2 ? FAM_FACES_Z : FAM_FACES_Y, [&] {
        switch (upd) {
        case UPD_RESIDUAL: launch_lacc<NV, DL, MODE, USE, UPD_RESIDUAL, T>(c, p, U0, Ut, out); break;
        case UPD_RK2_2: launch_lacc<NV, DL, MODE, USE, UPD_RK2_2, T>(c, p, U0, Ut, out); break;
        case UPD_RK3_2: launch_lacc<NV, DL, MODE, USE, UPD_RK3_2, T>(c, p, U0, Ut, out); break;
        default: launch_lacc<NV, DL, MODE, USE, UPD_RK3_3, T>(c, p, U0, Ut, out); break;
        }
    });
}

// One fused residual of Uin: cells (and the MUSCL-Hancock predictor, the indicator's audit), then the face passes.
// reuse: the learned indicator's faces take the decisions of the step's first stage.
template <int NV, typename T>
int enqueue_fused_nv(fvg_ctx* c, const double* Uin, int mode, bool audit, bool reuse, int upd, const double* U0, const double* Ut,
                     double* out) {
    const KP p = make_kp(c, Uin);
    int rc;
    if ((rc = enqueue_massflow(c, p))) return rc;
    if ((rc = launch(c, FAM_CELLS, [&] { hipLaunchKernelGGL((k_cells<NV, T>), dim3(cells_grid(c)), dim3(cells_block(c)), 0, c->stream, p); })))
        return rc;
    if constexpr (std::is_same<T, double>::value) {
        if (mode == MODE_HANCOCK) {
            if ((rc = launch(c, FAM_HANCOCK, [&] { hipLaunchKernelGGL(k_hancock<NV>, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, p); })))
                return rc;
            return enqueue_fused_faces<NV, MODE_HANCOCK, 0, T>(c, p, upd, U0, Ut, out);
        }
    }
    if (mode == MODE_LEARNED) {
        if (audit) {
#define AUDIT(D) launch(c, FAM_AUDIT, [&] { hipLaunchKernelGGL((k_audit<NV, D, T>), dim3(grid_for(c->nfa[D])), dim3(kBlock), 0, c->stream, p); })
            if ((rc = AUDIT(0)) || (rc = AUDIT(1))) return rc;
            if constexpr (NV == 5)
                if ((rc = AUDIT(2))) return rc;
#undef AUDIT
        }
        return reuse ? enqueue_fused_faces<NV, MODE_LEARNED, LF_KEPT, T>(c, p, upd, U0, Ut, out)
                     : enqueue_fused_faces<NV, MODE_LEARNED, LF_STORE, T>(c, p, upd, U0, Ut, out);
    }
    if (mode == MODE_PLAIN) return enqueue_fused_faces<NV, MODE_PLAIN, 0, T>(c, p, upd, U0, Ut, out);
    if (mode == MODE_FIRST_ORDER) return enqueue_fused_faces<NV, MODE_FIRST_ORDER, 0, T>(c, p, upd, U0, Ut, out);
    return set_err(c, FVG_ELOGIC, "no fused pass for this mode");
}

int enqueue_fused(fvg_ctx* c, const double* Uin, int mode, bool audit, bool reuse, int upd, const double* U0, const double* Ut,
                  double* out) {
    if (c->fp32)
        return c->NV == 5 ? enqueue_fused_nv<5, float>(c, Uin, mode, audit, reuse, upd, U0, Ut, out)
                          : enqueue_fused_nv<4, float>(c, Uin, mode, audit, reuse, upd, U0, Ut, out);
    return c->NV == 5 ? enqueue_fused_nv<5, double>(c, Uin, mode, audit, reuse, upd, U0, Ut, out)
                      : enqueue_fused_nv<4, double>(c, Uin, mode, audit, reuse, upd, U0, Ut, out);
}

int enqueue_axpy(fvg_ctx* c, const double* U0, const double* R, double* out) {
    const KP p = c->base;
    return launch(c, FAM_AXPY, [&] {
        if (c->NV == 5) hipLaunchKernelGGL((k_axpy<5>), dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, p, U0, R, out);
        else hipLaunchKernelGGL((k_axpy<4>), dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, p, U0, R, out);
    });
}

// whether the residuals of this context take the fused passes (the fourth-order quadrature keeps its own passes)
bool fused_path(const fvg_ctx* c) { return c->fused && !c->quad4; }

// The residual of Uin stored into R (Rb for fvg_residual): fused passes, or separate face and update passes
int residual_into(fvg_ctx* c, const double* Uin, int mode, bool audit, double* R) {
    if (fused_path(c)) return enqueue_fused(c, Uin, mode, audit, false, UPD_RESIDUAL, nullptr, nullptr, R);
    int rc = enqueue_residual(c, Uin, mode, audit);
    return rc ? rc : enqueue_update(c, UPD_RESIDUAL, nullptr, nullptr, R);
}

int plain_mode(const fvg_ctx* c) {
    if (c->learned) return MODE_LEARNED;
    if (c->quad4) return MODE_FOURTH;
    return c->cfg.reconstruction == FVG_RECON_NONE ? MODE_FIRST_ORDER : MODE_PLAIN;
}

// One step of the configured scheme, enqueued (TimeScheme.cpp:110-319).
// link (MUSCL-Hancock with a lagged speed only, for steps enqueued back to back): LINK_AFTER, the previous step's end
// was left to this step's first control launch; LINK_BEFORE, leave this step's end to the next step's.
enum { LINK_AFTER = 1, LINK_BEFORE = 2 };
int enqueue_step(fvg_ctx* c, int link = 0) {
    int rc;
    const int ts = c->cfg.time_scheme;
    const bool fused = fused_path(c);
    if (ts == FVG_TS_HANCOCK) {
        const int mode = c->cfg.reconstruction == FVG_RECON_NONE ? MODE_FIRST_ORDER : MODE_HANCOCK;
        if (!c->lagged_positive) {   // no wave speed yet: one plain residual gives it (TimeScheme.cpp:298-302)
            link = 0;
            if ((rc = enqueue_control(c, CTL_BEGIN_STEP, ts))) return rc;
            if ((rc = fused ? residual_into(c, c->U, plain_mode(c), false, c->Rb) : enqueue_residual(c, c->U, plain_mode(c)))) return rc;
            if ((rc = enqueue_control(c, CTL_SEED, ts))) return rc;
            if ((rc = enqueue_control(c, CTL_HANCOCK_DT, ts))) return rc;
        } else if ((rc = enqueue_control(c, (link & LINK_AFTER) ? CTL_END_BEGIN_HANCOCK : CTL_BEGIN_HANCOCK, ts))) {
            return rc;
        }
        if (fused) {   // the residual into Rb, then U += dt Rb (the faces read U, so the update waits for them)
            if ((rc = residual_into(c, c->U, mode, false, c->Rb)) || (rc = enqueue_axpy(c, c->U, c->Rb, c->U))) return rc;
        } else {
            if ((rc = enqueue_residual(c, c->U, mode))) return rc;
            if ((rc = enqueue_update(c, UPD_AXPY, c->U, nullptr, c->U))) return rc;
        }
        if (!(link & LINK_BEFORE) && (rc = enqueue_control(c, CTL_END_HANCOCK, ts))) return rc;
        c->lagged_positive = true;
        return FVG_OK;
    }
    if ((rc = enqueue_control(c, CTL_BEGIN_STEP, ts))) return rc;
    const int mode = plain_mode(c);
    if (c->learned && (rc = enqueue_indicator(c, CTL_AUDIT_BEGIN))) return rc;
    if (fused) {
        // stage 1: R into Rb, dt from its wave speed, then the first update; stages 2-3 finish their combination in
        // the last face pass (the SSP-RK3 stage-2 output goes to Rb, which stage 1's R no longer needs)
        if ((rc = residual_into(c, c->U, mode, c->learned, c->Rb))) return rc;
        if ((rc = enqueue_control(c, CTL_STAGE1_DT, ts))) return rc;
        if ((rc = enqueue_axpy(c, c->U, c->Rb, ts == FVG_TS_EULER ? c->U : c->Ut))) return rc;
        if (ts == FVG_TS_SSPRK2) {
            if ((rc = enqueue_fused(c, c->Ut, mode, false, c->learned, UPD_RK2_2, c->U, c->Ut, c->U))) return rc;
            if ((rc = enqueue_control(c, CTL_STAGE_S, 2))) return rc;
        } else if (ts == FVG_TS_SSPRK3) {
            if ((rc = enqueue_fused(c, c->Ut, mode, false, c->learned, UPD_RK3_2, c->U, c->Ut, c->Rb))) return rc;
            if ((rc = enqueue_control(c, CTL_STAGE_S, 2))) return rc;
            if ((rc = enqueue_fused(c, c->Rb, mode, false, c->learned, UPD_RK3_3, c->U, c->Rb, c->U))) return rc;
            if ((rc = enqueue_control(c, CTL_STAGE_S, 3))) return rc;
        }
        if ((rc = enqueue_control(c, CTL_END_RK, ts))) return rc;
        return c->learned ? enqueue_indicator(c, CTL_AUDIT_END) : FVG_OK;
    }
    if ((rc = enqueue_residual(c, c->U, mode, c->learned))) return rc;
    if ((rc = enqueue_control(c, CTL_STAGE1_DT, ts))) return rc;
    if (ts == FVG_TS_EULER) {
        if ((rc = enqueue_update(c, UPD_AXPY, c->U, nullptr, c->U))) return rc;
        if ((rc = enqueue_control(c, CTL_END_RK, ts))) return rc;
        return c->learned ? enqueue_indicator(c, CTL_AUDIT_END) : FVG_OK;
    }
    if ((rc = enqueue_update(c, UPD_AXPY, c->U, nullptr, c->Ut))) return rc;
    if ((rc = enqueue_residual(c, c->Ut, mode, false, c->learned))) return rc;
    if ((rc = enqueue_control(c, CTL_STAGE_S, 2))) return rc;
    if (ts == FVG_TS_SSPRK2) {
        if ((rc = enqueue_update(c, UPD_RK2_2, c->U, c->Ut, c->U))) return rc;
        if ((rc = enqueue_control(c, CTL_END_RK, ts))) return rc;
        return c->learned ? enqueue_indicator(c, CTL_AUDIT_END) : FVG_OK;
    }
    if ((rc = enqueue_update(c, UPD_RK3_2, c->U, c->Ut, c->Ut))) return rc;
    if ((rc = enqueue_residual(c, c->Ut, mode, false, c->learned))) return rc;
    if ((rc = enqueue_control(c, CTL_STAGE_S, 3))) return rc;
    if ((rc = enqueue_update(c, UPD_RK3_3, c->U, c->Ut, c->U))) return rc;
    if ((rc = enqueue_control(c, CTL_END_RK, ts))) return rc;
    return c->learned ? enqueue_indicator(c, CTL_AUDIT_END) : FVG_OK;
}

// kGraphSteps steps of the scheme captured into one graph, once per context: every kernel reads time, dt and the
// done flag from the device state, so the same graph serves every later call, and steps after the run's end are
// no-ops. The MUSCL-Hancock seed residual is never in the graph (the first step after a reset of the lagged speed
// runs eagerly).
int ensure_graph(fvg_ctx* c) {
    if (c->graph) return FVG_OK;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = FVG_OK;
    // MUSCL-Hancock: the end of each step and the start of the next share one control launch
    const bool hancock = c->cfg.time_scheme == FVG_TS_HANCOCK && c->lagged_positive;
    for (int k = 0; k < kGraphSteps && !rc; ++k)
        rc = enqueue_step(c, hancock ? (k > 0 ? LINK_AFTER : 0) | (k + 1 < kGraphSteps ? LINK_BEFORE : 0) : 0);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc ? rc : set_err(c, FVG_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    const hipError_t e2 = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        c->graph = nullptr;
        return set_err(c, FVG_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    }
    return FVG_OK;
}

int write_state(fvg_ctx* c, const DevState& s) {
    *c->st_host = s;
    HIPCHK(c, hipMemcpyAsync(c->st, c->st_host, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    return FVG_OK;
}
int read_state(fvg_ctx* c, DevState& s) {
    HIPCHK(c, hipMemcpyAsync(c->st_host, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s = *c->st_host;
    harvest_timing(c);
    return FVG_OK;
}

DevState fresh_state(const fvg_ctx* c) {
    DevState s{};
    s.t = 0.0;
    s.t_end = std::numeric_limits<double>::infinity();
    s.t_max = std::numeric_limits<double>::infinity();
    s.lagged = c->st_host ? c->st_host->lagged : 0.0;
    return s;
}

void free_ctx(fvg_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->graph) hipGraphExecDestroy(c->graph);
    hipFree(c->mlp);
    hipFree(c->ind);
    void* bufs[] = {c->U,     c->Ut,    c->Rb,    c->W,        c->G,         c->HS,  c->Fc[0], c->Fc[1], c->Fc[2], c->A,
                    c->Fv[0], c->Fv[1], c->Fv[2], c->tot_part, c->tot_out, c->spd, c->Wc,    c->q4_UL, c->q4_UR,
                    c->q4_Fp, c->q4_good, c->q4_pt, c->nrbc, c->ind_kept[0], c->ind_kept[1], c->ind_kept[2]};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (c->st) hipFree(c->st);
    if (c->st_host) hipHostFree(c->st_host);
    for (auto& r : c->recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

template <typename P>
int checked_alloc(fvg_ctx* c, P** p, size_t n, size_t elem = sizeof(double)) {
    if (n == 0) return FVG_OK;
    hipError_t e = hipMalloc((void**)p, n * elem);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, FVG_ENOMEM, "hipMalloc of " + std::to_string(n * elem) + " bytes: " + hipGetErrorString(e));
    }
    return FVG_OK;
}

// the face-flux planes of the separate face and update passes: per axis NV convective planes (FP32 fluxes: the
// combined flux) and, with FP64 fluxes and viscosity, NV - 1 viscous planes
int alloc_face_fluxes(fvg_ctx* c, const fvg_config& cfg) {
    const size_t fe = c->fp32 ? sizeof(float) : sizeof(double);
    int rc = FVG_OK;
    for (int d = 0; d < c->ND; ++d) {
        if (!c->Fc[d]) rc = rc ? rc : checked_alloc(c, &c->Fc[d], (size_t)c->NV * c->nfa[d], fe);
        if (cfg.use_viscous && !c->fp32 && !c->Fv[d]) rc = rc ? rc : checked_alloc(c, &c->Fv[d], (size_t)(c->NV - 1) * c->nfa[d], fe);
    }
    return rc;
}

}   // namespace

namespace fvg {

static unsigned bf16_host(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// The network's parameter vector (W1 32x15, b1, W2 32x32, b2, w3, b3) as the lanes of a wavefront read it
// (fvg_learned.hpp): b1 folded into column 15 of W1, the weight matrices as bfloat16 MFMA fragments, W2's columns
// in the order the layer-1 accumulator hands its rows to layer 2, b2 and w3 at each lane half's rows.
MlpDev pack_mlp(const float* P, float thr) {
    MlpDev m{};
    const float* W1 = P;
    const float* b1 = P + 480;
    const float* W2 = P + 512;
    auto w1 = [&](int h, int k) { return k < 15 ? W1[h * 15 + k] : b1[h]; };
    auto frag = [](const float* v) {
        Frag8 f{0, 0};
        for (int s = 0; s < 4; ++s) f.lo |= (unsigned long long)bf16_host(v[s]) << (16 * s);
        for (int s = 0; s < 4; ++s) f.hi |= (unsigned long long)bf16_host(v[4 + s]) << (16 * s);
        return f;
    };
    for (int l = 0; l < 64; ++l) {
        const int row = l % 32, h = l / 32;
        float v[8];
        for (int s = 0; s < 8; ++s) v[s] = w1(row, 8 * h + s);
        m.w1[l] = frag(v);
        for (int c = 0; c < 2; ++c) {
            for (int s = 0; s < 8; ++s) v[s] = W2[row * 32 + 16 * c + 8 * (s / 4) + 4 * h + s % 4];
            m.w2[c][l] = frag(v);
        }
    }
    for (int h = 0; h < 2; ++h)
        for (int e = 0; e < 16; ++e) {
            const int r = 8 * (e / 4) + 4 * h + e % 4;
            m.b2h[h][e] = P[1536 + r];
            m.w3h[h][e] = P[1568 + r];
        }
    m.b3 = P[1600];
    m.thr = thr;
    return m;
}

// Mesh, numerics and viscous scalars of the kernels for a finalized configuration (buffers are set by the caller).
// divisor d (1 <= d < 2^31) as a Magic pair for 0 <= n < 2^31: l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1
Magic magic_for(unsigned d) {
    int l = 0;
    while ((1ull << l) < d) ++l;
    const unsigned long long m = (((1ull << 32) * ((1ull << l) - d)) / d) + 1;
    return Magic{(unsigned)m, l};
}

void fill_kernel_params(const fvg_config& cfg, KP& p) {
    const bool three = cfg.nz > 1;
    const bool visc = cfg.use_viscous != 0;
    p.nx = cfg.nx;
    p.ny = cfg.ny;
    p.nz = cfg.nz;
    p.is3d = three;
    p.nc = (long long)cfg.nx * cfg.ny * cfg.nz;
    p.dx = cfg.Lx / cfg.nx;
    p.dy = cfg.Ly / cfg.ny;
    p.dz = cfg.Lz / cfg.nz;
    if (three) {   // Mesh::face_area / cell_volume (Core.hpp:94-99)
        p.Ax = p.dy * p.dz;
        p.Ay = p.dx * p.dz;
        p.Az = p.dx * p.dy;
        p.inv_vol = 1.0 / (p.dx * p.dy * p.dz);
    } else {
        p.Ax = p.dy;
        p.Ay = p.dx;
        p.Az = 0.0;
        p.inv_vol = 1.0 / (p.dx * p.dy);
    }
    p.rs.eos.stiff = cfg.eos_type == FVG_EOS_STIFFENED;
    p.rs.eos.gamma = cfg.gamma;
    p.rs.eos.p_inf = p.rs.eos.stiff ? cfg.p_inf : 0.0;
    p.rs.lowmach = cfg.riemann == FVG_LOWMACH_HLLC;
    p.rs.blend = cfg.shock_blend;
    for (int s = 0; s < 6; ++s) p.bp.bc[s] = cfg.bc[s];
    p.bp.u_in = cfg.u_in;
    p.bp.v_in = cfg.v_in;
    p.bp.p_out = cfg.p_out;
    p.bp.rho_in = cfg.rho_in;
    p.bp.nrbc_sigma = cfg.nrbc_sigma;
    p.bp.nrbc_beta = cfg.nrbc_beta;
    p.bp.length[0] = cfg.Lx;
    p.bp.length[1] = cfg.Ly;
    p.bp.length[2] = cfg.Lz;
    p.bp.ut = nullptr;   // the device state's ut, set by fvg_create
    p.recon = cfg.reconstruction;
    p.lim = cfg.limiter;
    p.Dh = cfg.reconstruction == FVG_RECON_WENO5 ? 3 : 2;
    p.bo2 = cfg.boundary_order == 2;
    p.weno_eps = cfg.weno_epsilon;
    // halos of the reconstruction (Simulation.cpp:863-905, 1109-1170)
    const int n3[3] = {cfg.nx, cfg.ny, cfg.nz};
    for (int a = 0; a < 3; ++a) {
        p.wrap[a] = 0;
        p.halo[2 * a] = p.halo[2 * a + 1] = 0;
        if (a == 2 && !three) continue;
        if (cfg.reconstruction == FVG_RECON_NONE) continue;
        const bool per = cfg.bc[2 * a] == FVG_BC_PERIODIC && cfg.bc[2 * a + 1] == FVG_BC_PERIODIC;
        if (per && n3[a] >= p.Dh) {
            p.wrap[a] = 1;
            p.halo[2 * a] = p.halo[2 * a + 1] = 1;
            continue;
        }
        if (cfg.boundary_order == 2)
            for (int hi = 0; hi < 2; ++hi) {
                const int bt = cfg.bc[2 * a + hi];
                if (bt != FVG_BC_PERIODIC && bt != FVG_BC_NONREFLECTING && n3[a] >= p.Dh) p.halo[2 * a + hi] = 2;
            }
    }
    // padded primitive planes
    p.px = padded_row(cfg.nx);
    p.py = cfg.ny + 2 * PH;
    p.phz = three ? PH : 0;
    p.rbt = std::min(kRowBlock, (cfg.nx + 63) / 64 * 64);
    p.nchunk = (cfg.nx + p.rbt - 1) / p.rbt;
    p.slabs = cfg.ny % 8 == 0;
    p.mg_chunk = magic_for((unsigned)p.nchunk);
    p.mg_ny = magic_for((unsigned)cfg.ny);
    p.mg_slab = magic_for(p.slabs ? (unsigned)(cfg.ny / 8) : 1u);
    for (int d = 0; d < 3; ++d) {
        p.mg_fi[d] = magic_for((unsigned)(cfg.nx + (d == 0)));
        p.mg_fj[d] = magic_for((unsigned)(cfg.ny + (d == 1)));
    }
    p.npad = (long long)p.px * p.py * (cfg.nz + 2 * p.phz);
    // line segments of the fused passes along y and z: 4m - 1 cells (4m faces: whole steps of the block's 4
    // wavefronts), the longest of 127, 63, .., 3 that still gives the last axis' pass 1536 blocks (two rounds of
    // three blocks per CU); a segment of m cells computes m + 1 faces. FVG_LINE_SEGMENT overrides it (tuning).
    p.nlchunk = (cfg.nx + 63) / 64;
    p.mg_lchunk = magic_for((unsigned)p.nlchunk);
    {
        const int nl = three ? cfg.nz : cfg.ny;
        const long long lines = (long long)p.nlchunk * (three ? cfg.ny : 1);
        int seg = 127;
        while (seg > 3 && lines * ((nl + seg - 1) / seg) < 1536) seg = (seg + 1) / 2 - 1;
        if (const char* e = std::getenv("FVG_LINE_SEGMENT")) {
            const int v = std::atoi(e);
            if (v >= 1 && v <= 4096) seg = v;
        }
        p.seg = seg;
        for (int a = 0; a < 3; ++a) {
            p.nseg[a] = (n3[a] + seg - 1) / seg;
            p.mg_nseg[a] = magic_for((unsigned)p.nseg[a]);
        }
    }
    p.cfl = cfg.cfl;
    p.Ldt = fvg_dt_length(&cfg);
    p.viscous = visc;
    p.mu = cfg.mu;
    p.k_cond = cfg.mu * cfg.gamma / (cfg.pr * (cfg.gamma - 1.0));   // EOS::TransportLaw (Core.hpp:196)
    p.inv_h2 = three ? 1.0 / (p.dx * p.dx) + 1.0 / (p.dy * p.dy) + 1.0 / (p.dz * p.dz) : 1.0 / (p.dx * p.dx) + 1.0 / (p.dy * p.dy);
    for (int a = 0; a < 3; ++a) p.vper[a] = cfg.bc[2 * a] == FVG_BC_PERIODIC && cfg.bc[2 * a + 1] == FVG_BC_PERIODIC;
}

}   // namespace fvg

extern "C" {

const char* fvg_last_error(const fvg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int fvg_create(const fvg_config* cfg_in, int device, fvg_ctx** out) {
    if (!cfg_in || !out) return set_err(nullptr, FVG_EINVAL, "null argument");
    *out = nullptr;
    fvg_config cfg = *cfg_in;
    char msg[256] = {0};
    if (fvg_config_finalize(&cfg, msg, sizeof msg) != FVG_OK) return set_err(nullptr, FVG_EINVAL, msg);
    if (cfg.time_scheme == FVG_TS_HANCOCK && (cfg.reconstruction == FVG_RECON_WENO3 || cfg.reconstruction == FVG_RECON_WENO5))
        return set_err(nullptr, FVG_EUNSUPPORTED, "MUSCL-Hancock with a WENO reconstruction is not on the device path");
    if (cfg.flux_precision == FVG_FLUX_FP32 && cfg.time_scheme == FVG_TS_HANCOCK && cfg.reconstruction != FVG_RECON_NONE)
        return set_err(nullptr, FVG_EUNSUPPORTED, "flux_precision = fp32 with MUSCL-Hancock is not on the device path");
    if (cfg.flux_quadrature == FVG_QUAD_FOURTH && cfg.flux_precision == FVG_FLUX_FP32)
        return set_err(nullptr, FVG_EUNSUPPORTED, "flux_quadrature = fourth is on the FP64 device path only (flux_precision = fp64)");
    // face and row-block indices are 32-bit: every axis' face count stays below 2^31
    const long long fmax = std::max({(long long)(cfg.nx + 1) * cfg.ny * cfg.nz, (long long)cfg.nx * (cfg.ny + 1) * cfg.nz,
                                     (long long)cfg.nx * cfg.ny * (cfg.nz + 1)});
    if (fmax >= (1LL << 31))
        return set_err(nullptr, FVG_EUNSUPPORTED, "mesh too large: the faces of an axis must number below 2^31");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(nullptr, FVG_EHIP, "no HIP device");
    if (device < 0 || device >= ndev) return set_err(nullptr, FVG_EINVAL, "device index out of range");
    fvg_ctx* c = new fvg_ctx();
    c->cfg = cfg;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        free_ctx(c);
        return set_err(nullptr, FVG_EHIP, "hipSetDevice failed");
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        free_ctx(c);
        return set_err(nullptr, FVG_EHIP, "hipStreamCreate failed");
    }
    const bool three = cfg.nz > 1;
    c->NV = three ? 5 : 4;
    c->ND = three ? 3 : 2;
    c->nc = (long long)cfg.nx * cfg.ny * cfg.nz;
    c->nfa[0] = (long long)(cfg.nx + 1) * cfg.ny * cfg.nz;
    c->nfa[1] = (long long)cfg.nx * (cfg.ny + 1) * cfg.nz;
    c->nfa[2] = three ? (long long)cfg.nx * cfg.ny * (cfg.nz + 1) : 0;
    const int NV = c->NV, ND = c->ND;
    const bool visc = cfg.use_viscous != 0, hancock = cfg.time_scheme == FVG_TS_HANCOCK && cfg.reconstruction != FVG_RECON_NONE;
    int rc = FVG_OK;
    rc = rc ? rc : checked_alloc(c, &c->U, (size_t)NV * c->nc);
    rc = rc ? rc : checked_alloc(c, &c->Ut, (size_t)NV * c->nc);
    rc = rc ? rc : checked_alloc(c, &c->Rb, (size_t)NV * c->nc);
    c->fp32 = cfg.flux_precision == FVG_FLUX_FP32;
    const size_t fe = c->fp32 ? sizeof(float) : sizeof(double);
    const long long npad = (long long)padded_row(cfg.nx) * (cfg.ny + 2 * PH) * (cfg.nz + (three ? 2 * PH : 0));
    rc = rc ? rc : checked_alloc(c, &c->W, (size_t)NV * npad, fe);
    if (visc) rc = rc ? rc : checked_alloc(c, &c->G, (size_t)ND * ND * c->nc, fe);
    if (hancock) rc = rc ? rc : checked_alloc(c, &c->HS, (size_t)2 * ND * NV * c->nc);
    c->quad4 = cfg.flux_quadrature == FVG_QUAD_FOURTH;
    // the face-flux planes of the separate passes (the default, and the fourth-order quadrature's); the accumulators
    // of the fused passes are allocated when fvg_set_fused switches to them
    rc = rc ? rc : alloc_face_fluxes(c, cfg);
    const long long q4_n = std::max({c->nfa[0], c->nfa[1], c->nfa[2]});
    if (c->quad4) {   // the padded conservative state and one axis' face states, flags and point-value fluxes
        rc = rc ? rc : checked_alloc(c, &c->Wc, (size_t)NV * npad);
        rc = rc ? rc : checked_alloc(c, &c->q4_UL, (size_t)NV * q4_n);
        rc = rc ? rc : checked_alloc(c, &c->q4_UR, (size_t)NV * q4_n);
        rc = rc ? rc : checked_alloc(c, &c->q4_Fp, (size_t)NV * q4_n);
        rc = rc ? rc : checked_alloc(c, &c->q4_good, (size_t)q4_n, 1);
        rc = rc ? rc : checked_alloc(c, &c->q4_pt, (size_t)q4_n, 1);
    }
    long long nrbc_off[6] = {0, 0, 0, 0, 0, 0}, nrbc_total = 0;
    for (int s = 0; s < (three ? 6 : 4); ++s) {
        if (cfg.bc[s] != FVG_BC_NONREFLECTING) continue;
        c->nrbc_sides |= 1 << s;
        nrbc_off[s] = nrbc_total;
        nrbc_total += (long long)NV * (s < 2 ? (long long)cfg.ny * cfg.nz : (s < 4 ? (long long)cfg.nx * cfg.nz : (long long)cfg.nx * cfg.ny));
    }
    rc = rc ? rc : checked_alloc(c, &c->nrbc, (size_t)nrbc_total);
    rc = rc ? rc : checked_alloc(c, &c->tot_part, (size_t)3 * kTotalsBlocks);
    rc = rc ? rc : checked_alloc(c, &c->tot_out, 3);
    c->learned = cfg.indicator == FVG_INDICATOR_LEARNED;
    if (c->learned) {
        rc = rc ? rc : checked_alloc(c, &c->mlp, 1, sizeof(MlpDev));
        rc = rc ? rc : checked_alloc(c, &c->ind, 1, sizeof(IndState));
        for (int d = 0; d < ND; ++d) rc = rc ? rc : checked_alloc(c, &c->ind_kept[d], (size_t)c->nfa[d], 1);
        if (!rc) {   // until weights are set every face is troubled (WENO5-Z everywhere)
            MlpDev m{};
            m.thr = -std::numeric_limits<float>::max();
            IndState z{};
            if (hipMemcpy(c->mlp, &m, sizeof m, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->ind, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess)
                rc = set_err(c, FVG_EHIP, "hipMemcpy of the indicator state");
        }
    }
    if (!rc && (hipMalloc(&c->st, sizeof(DevState)) != hipSuccess || hipMemset(c->st, 0, sizeof(DevState)) != hipSuccess))
        rc = set_err(c, FVG_ENOMEM, "hipMalloc of the device state");
    const size_t spd_bytes = sizeof(unsigned long long) * 2 * kSpeedSlots * kSlotStride;
    if (!rc && (hipMalloc(&c->spd, spd_bytes) != hipSuccess || hipMemset(c->spd, 0, spd_bytes) != hipSuccess))
        rc = set_err(c, FVG_ENOMEM, "hipMalloc of the wave-speed slots");
    if (!rc && hipHostMalloc(&c->st_host, sizeof(DevState)) != hipSuccess) rc = set_err(c, FVG_ENOMEM, "hipHostMalloc of the state mirror");
    if (rc) {
        const std::string m = c->err;
        free_ctx(c);
        return set_err(nullptr, rc, m);
    }
    std::memset(c->st_host, 0, sizeof(DevState));

    // kernel parameters that do not change during the run
    KP& p = c->base;
    fill_kernel_params(cfg, p);
    p.U = c->U;
    p.W = c->W;
    p.G = c->G;
    p.HS = c->HS;
    for (int d = 0; d < 3; ++d) {
        p.Fc[d] = c->Fc[d];
        p.Fv[d] = c->Fv[d];
        p.nfa[d] = c->nfa[d];
    }
    p.A = c->A;
    p.st = c->st;
    p.spd = c->spd;
    p.bp.ut = reinterpret_cast<const double*>(reinterpret_cast<const char*>(c->st) + offsetof(DevState, ut));
    p.nrbc = c->nrbc;
    for (int s = 0; s < 6; ++s) p.nrbc_off[s] = nrbc_off[s];
    for (int s = 0; s < (three ? 6 : 4); ++s) c->massflow = c->massflow || cfg.bc[s] == FVG_BC_MASSFLOW_INLET;
    p.quad4 = c->quad4 ? 1 : 0;
    p.Wc = c->Wc;
    p.q4_UL = c->q4_UL;
    p.q4_UR = c->q4_UR;
    p.q4_Fp = c->q4_Fp;
    p.q4_good = c->q4_good;
    p.q4_pt = c->q4_pt;
    p.q4_n = q4_n;
    p.mlp = c->mlp;
    for (int d = 0; d < 3; ++d) p.ind_kept[d] = c->ind_kept[d];
    p.ind = c->ind;
    p.ind_audit = cfg.indicator_audit;
    p.ind_label = (float)cfg.indicator_label;
    p.ind_src = cfg.indicator_label_source;
    p.ind_fn_limit = cfg.indicator_fn_limit;
    p.ind_fn_floor = cfg.indicator_fn_floor;
    DevState s = fresh_state(c);
    if ((rc = write_state(c, s)) || (rc = (hipStreamSynchronize(c->stream) == hipSuccess ? 0 : FVG_EHIP))) {
        const std::string m = c->err;
        free_ctx(c);
        return set_err(nullptr, rc, m);
    }
    *out = c;
    return FVG_OK;
}

int fvg_destroy(fvg_ctx* c) {
    free_ctx(c);
    return FVG_OK;
}

int fvg_upload(fvg_ctx* c, const double* U, size_t ncells, int ncomp) {
    if (!c || !U) return set_err(c, FVG_EINVAL, "null argument");
    if ((long long)ncells != c->nc) return set_err(c, FVG_EINVAL, "ncells does not match the mesh");
    if (ncomp != 5 && !(ncomp == 4 && c->NV == 4)) return set_err(c, FVG_EINVAL, "ncomp must be 5 (or 4 on a 2-D mesh)");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> soa((size_t)c->NV * c->nc);
    for (long long x = 0; x < c->nc; ++x)
        for (int q = 0; q < c->NV; ++q) soa[(size_t)q * c->nc + x] = U[(size_t)x * ncomp + q];
    HIPCHK(c, hipMemcpyAsync(c->U, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FVG_OK;
}

static int download_planes(fvg_ctx* c, const double* src, double* U, int ncomp) {
    std::vector<double> soa((size_t)c->NV * c->nc);
    HIPCHK(c, hipMemcpyAsync(soa.data(), src, soa.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (long long x = 0; x < c->nc; ++x) {
        for (int q = 0; q < c->NV; ++q) U[(size_t)x * ncomp + q] = soa[(size_t)q * c->nc + x];
        if (c->NV == 4 && ncomp == 5) U[(size_t)x * 5 + 4] = 0.0;   // rho*w of a 2-D field
    }
    return FVG_OK;
}

int fvg_download(fvg_ctx* c, double* U, size_t ncells, int ncomp) {
    if (!c || !U) return set_err(c, FVG_EINVAL, "null argument");
    if ((long long)ncells != c->nc) return set_err(c, FVG_EINVAL, "ncells does not match the mesh");
    if (ncomp != 5 && !(ncomp == 4 && c->NV == 4)) return set_err(c, FVG_EINVAL, "ncomp must be 5 (or 4 on a 2-D mesh)");
    HIPCHK(c, hipSetDevice(c->device));
    return download_planes(c, c->U, U, ncomp);
}

int fvg_residual(fvg_ctx* c, double half_dt, double* R, int ncomp, double* s_max) {
    if (!c) return set_err(c, FVG_EINVAL, "null context");
    if (R && ncomp != 5 && !(ncomp == 4 && c->NV == 4)) return set_err(c, FVG_EINVAL, "ncomp must be 5 (or 4 on a 2-D mesh)");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    DevState s = fresh_state(c);
    // with fourth-order fluxes the reference's residual takes no predictor (Simulation.cpp:991-1007)
    const bool hancock = half_dt > 0.0 && c->cfg.reconstruction != FVG_RECON_NONE && !c->quad4;
    if (hancock && c->fp32)   // the predictor kernels are FP64 only, and the primitive planes hold floats in this mode
        return set_err(c, FVG_EUNSUPPORTED, "the MUSCL-Hancock predictor is not on the fp32-flux path");
    if (hancock && !c->HS) {
        if ((rc = checked_alloc(c, &c->HS, (size_t)2 * c->ND * c->NV * c->nc))) return rc;
        c->base.HS = c->HS;
    }
    s.half_dt = hancock ? half_dt : 0.0;
    if ((rc = write_state(c, s))) return rc;
    if ((rc = enqueue_control(c, CTL_BEGIN_STEP, 0))) return rc;
    if ((rc = residual_into(c, c->U, hancock ? MODE_HANCOCK : plain_mode(c), false, c->Rb))) return rc;
    if ((rc = enqueue_control(c, CTL_RESIDUAL_S, 0))) return rc;
    DevState o;
    if ((rc = read_state(c, o))) return rc;
    if (s_max) *s_max = o.s_step;
    if (R) return download_planes(c, c->Rb, R, ncomp);
    return FVG_OK;
}

int fvg_step(fvg_ctx* c, double t, double t_max, double dt_max, double* dt, double* s_max) {
    if (!c) return set_err(c, FVG_EINVAL, "null context");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    DevState s = fresh_state(c);
    s.t = t;
    s.t_max = t_max;
    s.dt_max = dt_max;
    s.max_steps = 1;
    if ((rc = write_state(c, s))) return rc;
    if ((rc = enqueue_step(c))) return rc;
    DevState o;
    if ((rc = read_state(c, o))) return rc;
    if (dt) *dt = o.dt;
    if (s_max) *s_max = o.s_step;
    if (o.diverged) {
        char m[256];
        std::snprintf(m, sizeof m, "the solution diverged (t = %.6g, dt = %.3g, largest wave speed = %.3g)", t, o.dt, o.s_step);
        return set_err(c, FVG_EDIVERGED, m);
    }
    return FVG_OK;
}

int fvg_advance(fvg_ctx* c, int max_steps, double* t, double t_end, double dt_max, int* steps, double* s_last) {
    if (!c || !t) return set_err(c, FVG_EINVAL, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    DevState s = fresh_state(c);
    s.t = *t;
    s.t_end = t_end;
    s.t_max = t_end;
    s.dt_max = dt_max;
    s.max_steps = max_steps > 0 ? max_steps : 0;
    if ((rc = write_state(c, s))) return rc;
    DevState o{};
    long long enqueued = 0;
    const int chunk = 16;
    const bool graphed = c->graphs && !c->timing;
    for (;;) {
        const long long left = max_steps > 0 ? max_steps - enqueued : chunk;
        int n = (int)std::min<long long>(chunk, left);
        if (graphed && c->cfg.time_scheme == FVG_TS_HANCOCK && !c->lagged_positive) {
            if ((rc = enqueue_step(c))) return rc;
            enqueued += 1;
            n -= 1;
        }
        if (graphed) {
            if ((rc = ensure_graph(c))) return rc;
            for (int k = 0; k < n; k += kGraphSteps) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
        } else {
            for (int k = 0; k < n; ++k)
                if ((rc = enqueue_step(c))) return rc;
        }
        enqueued += n;
        if ((rc = read_state(c, o))) return rc;
        if (o.done || o.diverged || (max_steps > 0 && enqueued >= max_steps)) break;
    }
    *t = o.t;
    if (steps) *steps = o.steps;
    if (s_last) *s_last = o.s_step;
    if (o.diverged) {
        char m[256];
        std::snprintf(m, sizeof m, "the solution diverged at step %d (t = %.6g, dt = %.3g, largest wave speed = %.3g)", o.steps, o.t,
                      o.dt, o.s_step);
        return set_err(c, FVG_EDIVERGED, m);
    }
    return FVG_OK;
}

int fvg_get_lagged_speed(const fvg_ctx* c, double* s) {
    if (!c || !s) return FVG_EINVAL;
    *s = c->st_host->lagged;
    return FVG_OK;
}

int fvg_set_lagged_speed(fvg_ctx* c, double sp) {
    if (!c) return FVG_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    c->st_host->lagged = sp;
    DevState s = fresh_state(c);
    s.lagged = sp;
    int rc = write_state(c, s);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lagged_positive = sp > 0.0;
    return FVG_OK;
}

int fvg_checkpoint_write(fvg_ctx* c, const char* path, double t, int step) {
    if (!c || !path) return set_err(c, FVG_EINVAL, "null argument");
    const int nq = c->NV;   // the reference's columns: 4 in 2-D, 5 in 3-D
    std::vector<double> U((size_t)nq * c->nc);
    int rc = fvg_download(c, U.data(), (size_t)c->nc, nq);
    if (rc) return rc;
    const double s = c->cfg.time_scheme == FVG_TS_HANCOCK ? c->st_host->lagged : 0.0;
    char err[512] = {0};
    rc = fvg_checkpoint_save(&c->cfg, U.data(), (size_t)c->nc, nq, t, step, s, path, err, sizeof err);
    return rc ? set_err(c, rc, err) : FVG_OK;
}

int fvg_checkpoint_read(fvg_ctx* c, const char* path, double* t, int* step) {
    if (!c || !path || !t || !step) return set_err(c, FVG_EINVAL, "null argument");
    const int nq = c->NV;
    std::vector<double> U((size_t)nq * c->nc);
    double s = 0.0;
    char err[512] = {0};
    int rc = fvg_checkpoint_load(&c->cfg, path, U.data(), (size_t)c->nc, nq, t, step, &s, err, sizeof err);
    if (rc) return set_err(c, rc, err);
    if ((rc = fvg_upload(c, U.data(), (size_t)c->nc, nq))) return rc;
    // a file without S leaves the lagged speed unset: the next MUSCL-Hancock step seeds it (SimulationRestart.cpp:343-344)
    if (c->cfg.time_scheme == FVG_TS_HANCOCK) return fvg_set_lagged_speed(c, s);
    return FVG_OK;
}

// the three sums of a plane set (what 0: conserved totals of a state, 1: residual sums of squares) into out[3]
static int plane_sums(fvg_ctx* c, int what, const double* planes, double out[3]) {
    const int nb = kTotalsBlocks;
    const dim3 g(nb), b(256);
    if (c->NV == 5 && what == 0) hipLaunchKernelGGL((k_totals_partial<5, 0>), g, b, 0, c->stream, planes, c->nc, c->tot_part);
    else if (c->NV == 5) hipLaunchKernelGGL((k_totals_partial<5, 1>), g, b, 0, c->stream, planes, c->nc, c->tot_part);
    else if (what == 0) hipLaunchKernelGGL((k_totals_partial<4, 0>), g, b, 0, c->stream, planes, c->nc, c->tot_part);
    else hipLaunchKernelGGL((k_totals_partial<4, 1>), g, b, 0, c->stream, planes, c->nc, c->tot_part);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1), 0, c->stream, c->tot_part, nb, c->nc, what, c->tot_out);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, c->tot_out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FVG_OK;
}

int fvg_conserved_totals(fvg_ctx* c, double out[3]) {
    if (!c || !out) return set_err(c, FVG_EINVAL, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    return plane_sums(c, 0, c->U, out);
}

int fvg_monitor(fvg_ctx* c, double out[6]) {
    if (!c || !out) return set_err(c, FVG_EINVAL, "null argument");
    int rc;
    if ((rc = fvg_residual(c, 0.0, nullptr, 5, nullptr))) return rc;   // the plain residual of the state, into Rb
    if ((rc = plane_sums(c, 0, c->U, out))) return rc;
    return plane_sums(c, 1, c->Rb, out + 3);
}

int fvg_indicator_set(fvg_ctx* c, const float* params, int nparam, float threshold) {
    if (!c || !params) return set_err(c, FVG_EINVAL, "null argument");
    if (!c->learned) return set_err(c, FVG_ELOGIC, "the context was not created with indicator = learned");
    if (nparam != MLP_NPARAM) return set_err(c, FVG_EINVAL, "the network has " + std::to_string(MLP_NPARAM) + " parameters");
    for (int k = 0; k < nparam; ++k)
        if (!std::isfinite(params[k])) return set_err(c, FVG_EINVAL, "non-finite network parameter");
    if (std::isnan(threshold)) return set_err(c, FVG_EINVAL, "threshold is NaN");
    HIPCHK(c, hipSetDevice(c->device));
    const MlpDev m = pack_mlp(params, threshold);
    IndState z{};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->mlp, &m, sizeof m, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->ind, &z, sizeof z, hipMemcpyHostToDevice));
    return FVG_OK;
}

int fvg_indicator_stats(fvg_ctx* c, unsigned long long counts[8], int* off, int* audits) {
    if (!c) return FVG_EINVAL;
    if (!c->learned) return set_err(c, FVG_ELOGIC, "the context was not created with indicator = learned");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    IndState h{};
    HIPCHK(c, hipMemcpy(&h, c->ind, sizeof h, hipMemcpyDeviceToHost));
    if (counts)
        for (int q = 0; q < 4; ++q) {
            counts[q] = h.last[q];
            counts[4 + q] = h.total[q];
        }
    if (off) *off = h.off;
    if (audits) *audits = h.audits;
    return FVG_OK;
}

int fvg_indicator_samples(fvg_ctx* c, int axis, float* X, float* dep, long long nfaces) {
    if (!c || !X || !dep) return set_err(c, FVG_EINVAL, "null argument");
    if (axis < 0 || axis >= c->ND) return set_err(c, FVG_EINVAL, "axis out of range");
    if (nfaces != c->nfa[axis]) return set_err(c, FVG_EINVAL, "nfaces differs from the axis' face count");
    HIPCHK(c, hipSetDevice(c->device));
    float *dX = nullptr, *dd = nullptr;
    if (hipMalloc(&dX, sizeof(float) * 16 * (size_t)nfaces) != hipSuccess) return set_err(c, FVG_ENOMEM, "hipMalloc of the samples");
    if (hipMalloc(&dd, sizeof(float) * (size_t)nfaces) != hipSuccess) {
        hipFree(dX);
        return set_err(c, FVG_ENOMEM, "hipMalloc of the samples");
    }
    KP p = make_kp(c, c->U);
    p.viscous = 0;   // only the primitives are needed: no gradients, and no speed maxima left behind in the slots
    const dim3 gc(cells_grid(c)), bc(cells_block(c)), gf((unsigned)((nfaces + kBlock - 1) / kBlock)), b(kBlock);
    if (int rc = enqueue_massflow(c, p)) {
        hipFree(dX);
        hipFree(dd);
        return rc;
    }
    auto samples = [&](auto tag) {   // the primitives in the context's flux precision, then the axis' samples
        using T = decltype(tag);
        if (c->NV == 5) {
            hipLaunchKernelGGL((k_cells<5, T>), gc, bc, 0, c->stream, p);
            if (axis == 0) hipLaunchKernelGGL((k_samples<5, 0, T>), gf, b, 0, c->stream, p, dX, dd);
            else if (axis == 1) hipLaunchKernelGGL((k_samples<5, 1, T>), gf, b, 0, c->stream, p, dX, dd);
            else hipLaunchKernelGGL((k_samples<5, 2, T>), gf, b, 0, c->stream, p, dX, dd);
        } else {
            hipLaunchKernelGGL((k_cells<4, T>), gc, bc, 0, c->stream, p);
            if (axis == 0) hipLaunchKernelGGL((k_samples<4, 0, T>), gf, b, 0, c->stream, p, dX, dd);
            else hipLaunchKernelGGL((k_samples<4, 1, T>), gf, b, 0, c->stream, p, dX, dd);
        }
    };
    if (c->fp32) samples(0.0f);
    else samples(0.0);
    int rc = FVG_OK;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, FVG_EHIP, "sample kernels failed");
    if (!rc && (hipMemcpy(X, dX, sizeof(float) * 16 * (size_t)nfaces, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(dep, dd, sizeof(float) * (size_t)nfaces, hipMemcpyDeviceToHost) != hipSuccess))
        rc = set_err(c, FVG_EHIP, "hipMemcpy of the samples");
    hipFree(dX);
    hipFree(dd);
    return rc;
}

int fvg_set_fused(fvg_ctx* c, int on) {
    if (!c) return FVG_EINVAL;
    if (on && !c->quad4 && !c->A) {   // fused passes: the accumulator planes
        HIPCHK(c, hipSetDevice(c->device));
        const bool visc = c->cfg.use_viscous != 0;
        if (int rc = checked_alloc(c, &c->A, (size_t)(visc && !c->fp32 ? 2 * c->NV - 1 : c->NV) * c->nc)) return rc;
        c->base.A = c->A;
    }
    c->fused = on != 0;
    if (c->graph) {   // the captured graph holds the old kernel sequence
        hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return FVG_OK;
}

int fvg_set_graphs(fvg_ctx* c, int on) {
    if (!c) return FVG_EINVAL;
    c->graphs = on != 0;
    return FVG_OK;
}

int fvg_set_timing(fvg_ctx* c, int on) {
    if (!c) return FVG_EINVAL;
    c->timing = on != 0;
    for (int f = 0; f < FAM_N; ++f) {
        c->fam_ms[f] = 0.0;
        c->fam_n[f] = 0;
    }
    return FVG_OK;
}

int fvg_kernel_times(const fvg_ctx* c, const char** names, double* ms, long long* launches, int cap) {
    if (!c) return 0;
    int k = 0;
    for (int f = 0; f < FAM_N && k < cap; ++f) {
        if (names) names[k] = kFamilyNames[f];
        if (ms) ms[k] = c->fam_ms[f];
        if (launches) launches[k] = c->fam_n[f];
        ++k;
    }
    return k;
}

int fvg_riemann_batch(const fvg_config* cfg, int device, const double* UL, const double* UR, const int* axis, size_t n, double* F,
                      double* smax) {
    if (!cfg || !UL || !UR || !axis || !F || !smax) return set_err(nullptr, FVG_EINVAL, "null argument");
    if (n == 0) return FVG_OK;
    for (size_t k = 0; k < n; ++k)
        if (axis[k] < 0 || axis[k] > 2) return set_err(nullptr, FVG_EINVAL, "axis must be 0, 1 or 2");
    if (hipSetDevice(device) != hipSuccess) return set_err(nullptr, FVG_EHIP, "hipSetDevice failed");
    RiemannParams rs;
    rs.eos.stiff = cfg->eos_type == FVG_EOS_STIFFENED;
    rs.eos.gamma = cfg->gamma;
    rs.eos.p_inf = rs.eos.stiff ? cfg->p_inf : 0.0;
    rs.lowmach = cfg->riemann == FVG_LOWMACH_HLLC;
    rs.blend = cfg->shock_blend;
    double *dUL = nullptr, *dUR = nullptr, *dF = nullptr, *dS = nullptr;
    int* dA = nullptr;
    int rc = FVG_OK;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return set_err(nullptr, FVG_EHIP, "hipStreamCreate failed");
    if (hipMalloc(&dUL, 5 * n * sizeof(double)) != hipSuccess || hipMalloc(&dUR, 5 * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dF, 5 * n * sizeof(double)) != hipSuccess || hipMalloc(&dS, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dA, n * sizeof(int)) != hipSuccess) {
        rc = set_err(nullptr, FVG_ENOMEM, "hipMalloc failed");
    }
    if (!rc) {
        hipMemcpyAsync(dUL, UL, 5 * n * sizeof(double), hipMemcpyHostToDevice, st);
        hipMemcpyAsync(dUR, UR, 5 * n * sizeof(double), hipMemcpyHostToDevice, st);
        hipMemcpyAsync(dA, axis, n * sizeof(int), hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(k_riemann_batch, dim3(grid_for((long long)n)), dim3(kBlock), 0, st, rs, dUL, dUR, dA, (long long)n, dF, dS);
        hipMemcpyAsync(F, dF, 5 * n * sizeof(double), hipMemcpyDeviceToHost, st);
        hipMemcpyAsync(smax, dS, n * sizeof(double), hipMemcpyDeviceToHost, st);
        hipError_t e = hipStreamSynchronize(st);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) rc = set_err(nullptr, FVG_EHIP, std::string("riemann batch: ") + hipGetErrorString(e));
    }
    if (dUL) hipFree(dUL);
    if (dUR) hipFree(dUR);
    if (dF) hipFree(dF);
    if (dS) hipFree(dS);
    if (dA) hipFree(dA);
    hipStreamDestroy(st);
    return rc;
}

}   // extern "C"
