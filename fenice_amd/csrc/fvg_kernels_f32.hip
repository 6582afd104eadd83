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
#define FVG_ZUPD(NV, M, T)                                                                                 \
    template __global__ void k_zupdate<NV, M, UPD_RK2_2, T>(KP, const double*, const double*, double*); \
    template __global__ void k_zupdate<NV, M, UPD_RK3_2, T>(KP, const double*, const double*, double*); \
    template __global__ void k_zupdate<NV, M, UPD_RK3_3, T>(KP, const double*, const double*, double*);
FVG_INST_T(4, float)
FVG_INST_T(5, float)
template __global__ void k_faces<5, 2, MODE_PLAIN, float>(KP);
template __global__ void k_faces<5, 2, MODE_FIRST_ORDER, float>(KP);
FVG_ZUPD(5, MODE_PLAIN, float)
FVG_ZUPD(5, MODE_FIRST_ORDER, float)
FVG_ZUPD(5, MODE_LEARNED, float)
FVG_ZUPD(4, MODE_PLAIN, float)
FVG_ZUPD(4, MODE_FIRST_ORDER, float)
FVG_ZUPD(4, MODE_LEARNED, float)
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

}   // namespace fvg
