// The FP32-flux instantiations of the residual kernels (float reconstruction and fluxes, FP64 state).
#include "fvg_kernels_impl.hpp"

namespace fvg {

#define FVG_INST_T(NV, T)                                                                                            \
    template __global__ void k_cells<NV, T>(KP);                                                                     \
    template __global__ void k_faces<NV, 0, MODE_PLAIN, T>(KP);                                                      \
    template __global__ void k_faces<NV, 1, MODE_PLAIN, T>(KP);                                                      \
    template __global__ void k_faces<NV, 0, MODE_FIRST_ORDER, T>(KP);                                                \
    template __global__ void k_faces<NV, 1, MODE_FIRST_ORDER, T>(KP);                                                \
    template __global__ void k_update<NV, UPD_RESIDUAL, T>(KP, const double*, const double*, double*);              \
    template __global__ void k_update<NV, UPD_AXPY, T>(KP, const double*, const double*, double*);                  \
    template __global__ void k_update<NV, UPD_RK2_2, T>(KP, const double*, const double*, double*);                 \
    template __global__ void k_update<NV, UPD_RK3_2, T>(KP, const double*, const double*, double*);                 \
    template __global__ void k_update<NV, UPD_RK3_3, T>(KP, const double*, const double*, double*);
FVG_INST_T(4, float)
FVG_INST_T(5, float)
template __global__ void k_faces<5, 2, MODE_PLAIN, float>(KP);
template __global__ void k_faces<5, 2, MODE_FIRST_ORDER, float>(KP);
#define FVG_LEARNED(T)                                                                \
    template __global__ void k_faces_learned<4, 0, LF_STORE, T>(KP);                  \
    template __global__ void k_faces_learned<4, 1, LF_STORE, T>(KP);                  \
    template __global__ void k_faces_learned<5, 0, LF_STORE, T>(KP);                  \
    template __global__ void k_faces_learned<5, 1, LF_STORE, T>(KP);                  \
    template __global__ void k_faces_learned<5, 2, LF_STORE, T>(KP);                  \
    template __global__ void k_faces_learned<4, 0, LF_KEPT, T>(KP);                   \
    template __global__ void k_faces_learned<4, 1, LF_KEPT, T>(KP);                   \
    template __global__ void k_faces_learned<5, 0, LF_KEPT, T>(KP);                   \
    template __global__ void k_faces_learned<5, 1, LF_KEPT, T>(KP);                   \
    template __global__ void k_faces_learned<5, 2, LF_KEPT, T>(KP);                   \
    template __global__ void k_audit<4, 0, T>(KP);                                    \
    template __global__ void k_audit<4, 1, T>(KP);                                    \
    template __global__ void k_audit<5, 0, T>(KP);                                    \
    template __global__ void k_audit<5, 1, T>(KP);                                    \
    template __global__ void k_audit<5, 2, T>(KP);                                    \
    template __global__ void k_samples<4, 0, T>(KP, float*, float*);                  \
    template __global__ void k_samples<4, 1, T>(KP, float*, float*);                  \
    template __global__ void k_samples<5, 0, T>(KP, float*, float*);                  \
    template __global__ void k_samples<5, 1, T>(KP, float*, float*);                  \
    template __global__ void k_samples<5, 2, T>(KP, float*, float*);
FVG_LEARNED(float)
// the fused residual passes (k_xacc, k_lacc) of every mode and stage the host launches
#define FVG_FUSED_MODE(NV, M, USE, T) template __global__ void k_xacc<NV, M, USE, T>(KP);
#define FVG_YPASS(M, USE, T) template __global__ void k_lacc<5, 1, M, USE, -1, T>(KP, const double*, const double*, double*);
#define FVG_LAST(NV, M, USE, UPD, T) \
    template __global__ void k_lacc<NV, (NV == 5 ? 2 : 1), M, USE, UPD, T>(KP, const double*, const double*, double*);
#define FVG_LAST_ALL(NV, M, T)                  \
    FVG_LAST(NV, M, 0, UPD_RESIDUAL, T)         \
    FVG_LAST(NV, M, 0, UPD_RK2_2, T)            \
    FVG_LAST(NV, M, 0, UPD_RK3_2, T)            \
    FVG_LAST(NV, M, 0, UPD_RK3_3, T)
#define FVG_FUSED(NV, T)                                      \
    FVG_FUSED_MODE(NV, MODE_PLAIN, 0, T)                      \
    FVG_FUSED_MODE(NV, MODE_FIRST_ORDER, 0, T)                \
    FVG_FUSED_MODE(NV, MODE_LEARNED, LF_STORE, T)             \
    FVG_FUSED_MODE(NV, MODE_LEARNED, LF_KEPT, T)              \
    FVG_LAST_ALL(NV, MODE_PLAIN, T)                           \
    FVG_LAST_ALL(NV, MODE_FIRST_ORDER, T)                     \
    FVG_LAST(NV, MODE_LEARNED, LF_STORE, UPD_RESIDUAL, T)     \
    FVG_LAST(NV, MODE_LEARNED, LF_KEPT, UPD_RK2_2, T)         \
    FVG_LAST(NV, MODE_LEARNED, LF_KEPT, UPD_RK3_2, T)         \
    FVG_LAST(NV, MODE_LEARNED, LF_KEPT, UPD_RK3_3, T)
FVG_FUSED(4, float)
FVG_FUSED(5, float)
FVG_YPASS(MODE_PLAIN, 0, float)
FVG_YPASS(MODE_FIRST_ORDER, 0, float)
FVG_YPASS(MODE_LEARNED, LF_STORE, float)
FVG_YPASS(MODE_LEARNED, LF_KEPT, float)

}   // namespace fvg
