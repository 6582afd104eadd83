// Kernels of the uniform-grid residual and time step (gfx950, FP64, structure-of-arrays planes).
//
// Layout: a set of planes is one allocation, plane q at base + q * stride. Cells are (k*ny + j)*nx + i.
// Faces normal to axis d are stored per axis in the reference's face order of that axis
// (x: ((k*ny + j)*(nx+1) + i), y: ((k*(ny+1) + j)*nx + i), z: ((k*ny + j)*nx + i) with k <= nz).
//
// One residual = k_cells (primitives, viscous cell gradients and the viscous speed bound) ->
// [k_hancock (MUSCL-Hancock only: the predicted face states of every cell)] -> k_faces<axis> per axis
// (stencil, reconstruction, Riemann solve, viscous face flux; block maximum of the wave speed then one 64-bit
// atomicMax) -> k_control (time step on the device) -> k_update (flux divergence, viscous balance and the
// stage combination of the time scheme, fused).
#pragma once

#include "fvg_learned.hpp"

namespace fvg {

// Time-loop state kept on the device, so that a run of steps needs no host round trip.
struct DevState {
    unsigned long long s_conv, s_visc;   // bit patterns of the stage's maxima (non-negative doubles order as integers)
    double t, dt, half_dt, t_end, t_max, dt_max;
    double s1, s2, s3, s_step, lagged;
    long long reverted;
    int done, diverged, steps, max_steps;
    double ut[6];   // mass-flow inlet velocity per side of the current residual (k_massflow)
    // nonzero: a value k_cells stored for the residual's WENO5-Z stencils (a primitive of the FP64 planes or, with
    // the fourth-order quadrature, a value of the padded conservative planes) is outside range_fits_quot6, so the
    // kernels take the IEEE division for WENO5-Z's quotients by 6. k_cells only sets it; k_update (the end of every
    // residual of the separate passes) and k_control (after a residual's faces or between steps) clear it, never
    // between a k_cells and the kernels that read its planes.
    int wide;
};

// n / d for 0 <= n < 2^31 as (umulhi(n, m) + n) >> l, with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1
// (the round-up method of Granlund and Montgomery, 1994); the sum stays below 2^32 for n < 2^31
struct Magic {
    unsigned m;
    int l;
};

struct KP {
    // mesh
    int nx, ny, nz, is3d;
    long long nc;
    double dx, dy, dz;
    double Ax, Ay, Az, inv_vol;   // face areas and 1 / volume (2-D: dy, dx, -, 1 / (dx dy))
    // numerics
    RiemannParams rs;
    BcParams bp;
    int recon, lim, Dh, bo2;
    double weno_eps;
    int halo[6];    // 0 none, 1 periodic wrap, 2 ghost slab (boundary_order = 2)
    int wrap[3];
    double cfl, Ldt;   // CFL number and dt_length
    // viscous
    int viscous;
    double mu, k_cond, inv_h2;
    int vper[3];    // periodic axes of the viscous stencil
    // padded primitive planes: W carries PH halo layers on each side of every axis (none along z in 2-D), filled
    // by k_cells with the value a reconstruction stencil takes there (periodic wrap, ghost slab or edge copy), so
    // stencils read W without boundary logic. Cell (i, j, k), coordinates from -PH, is element
    // ((k + phz) * py + j + PH) * px + i + PX0 of a plane of npad elements; rows are PXA elements apart in
    // multiples and start their interior at PX0, so that every row's interior starts on a cache-line boundary.
    int px, py, phz;
    long long npad;
    // row blocks (k_cells, k_update, z faces): blocks of rbt threads (nx rounded up to a multiple of 64, at most
    // kRowBlock) over chunks of a row of nx elements; divisors by multiply-high (Magic): chunks per row, rows per
    // XCD slab (ny / 8), rows per plane (ny)
    int rbt, nchunk, slabs;   // slabs: 1 when ny divides by 8 (each XCD takes a y-slab), else 0 (rows in mesh order)
    Magic mg_chunk, mg_slab, mg_ny;
    Magic mg_fi[3], mg_fj[3];   // faces per row and rows per plane of the faces normal to each axis
    // fused residual passes (k_xacc, k_lacc): accumulator planes A of the partial sums of the divergence (FP64: NV
    // convective, then NV - 1 viscous with FP64 fluxes and viscosity), and the line segments of the passes along y
    // and z: seg cells each, nseg[d] segments per line along axis d, blocks of 64 x positions
    double* A;
    int seg, nseg[3], nlchunk;   // nlchunk: 64-lane chunks of an x-row
    Magic mg_nseg[3], mg_lchunk;
    // buffers (plane sets)
    // W, G, Fc and Fv hold doubles in the FP64 mode and floats in the FP32-flux mode (flux_precision)
    const double* U;    // stage input state (always FP64)
    void* W;            // primitives of U
    void* G;            // viscous cell gradients (3-D: 9 planes G[a][b] at 3a+b; 2-D: dudx, dudy, dvdx, dvdy)
    double* HS;         // MUSCL-Hancock predicted states: (2*ndim) * NV planes, state s = 2*axis + high (FP64 only)
    void* Fc[3];        // convective face flux per axis, NV planes each
    void* Fv[3];        // viscous face flux per axis (3-D: slots 1..4; 2-D: slots 1..3)
    long long nfa[3];   // faces per axis (plane stride of Fc / Fv)
    DevState* st;
    // the block maxima of the wave speeds land in kSpeedSlots slots per quantity (convective, viscous), chosen by
    // block index and kSlotStride elements apart, so that the atomics of a pass spread over L2 channels;
    // k_control reduces and clears them
    unsigned long long* spd;
    // flux_quadrature = fourth (FP64): Wc holds the conservative state padded like W (the stencils of the
    // conserved reconstruction); the face states, flags and point-value fluxes of one axis at a time are q4_UL,
    // q4_UR, q4_Fp (NV planes of q4_n faces each) and q4_good, q4_pt
    int quad4;
    double* Wc;
    double *q4_UL, *q4_UR, *q4_Fp;
    unsigned char *q4_good, *q4_pt;
    long long q4_n;
    // non-reflecting sides: the ghost state of each boundary cell of the stage input (k_nrbc), NV planes per side
    // of the side's boundary cells, side s from element nrbc_off[s] (null without such a side)
    double* nrbc;
    long long nrbc_off[6];
    // learned troubled-face indicator (WENO5-Z, either flux precision)
    unsigned char* ind_kept[3];   // per axis, one byte per face: the first stage's decision (1 troubled), kept for the step's later stages
    const MlpDev* mlp;
    IndState* ind;
    int ind_audit;            // steps between audits (0: none)
    float ind_label;          // label value above which a face is troubled
    int ind_src;              // label source (enum fvg_label_source): WENO5-Z weight departure, MUSCL limiter clip, both
    double ind_fn_limit;      // missed fraction of troubled faces that switches the indicator off
    double ind_fn_floor;      // ... when the missed faces are also above this fraction of all audited faces
};

constexpr int kSpeedSlots = 64, kSlotStride = 32;   // 64 slots per quantity, 256 B apart
constexpr int SPD_CONV = 0, SPD_VISC = 1;
constexpr int kRowBlock = 256;   // most threads of a row block
constexpr int PH = 3;     // halo layers of the padded primitive planes (the WENO5-Z stencil reaches 3 cells out)
constexpr int PXA = 32;   // row alignment of the padded planes in elements (128 B in FP32, 256 B in FP64)
constexpr int PX0 = 32;   // element of a padded row that holds i = 0 (the low x halo sits just before it)

enum { MODE_PLAIN = 0, MODE_HANCOCK = 1, MODE_FIRST_ORDER = 2, MODE_LEARNED = 3, MODE_FOURTH = 4 };
enum { LF_STORE = 0, LF_KEPT = 1 };   // use of the learned indicator's decisions (learned_face)
enum { UPD_RESIDUAL = 0, UPD_AXPY = 1, UPD_RK2_2 = 2, UPD_RK3_2 = 3, UPD_RK3_3 = 4 };
enum { CTL_BEGIN_STEP = 0, CTL_STAGE1_DT = 1, CTL_STAGE_S = 2, CTL_END_RK = 3, CTL_SEED = 4, CTL_HANCOCK_DT = 5, CTL_END_HANCOCK = 6,
       CTL_RESIDUAL_S = 7, CTL_AUDIT_BEGIN = 8, CTL_AUDIT_END = 9,
       CTL_BEGIN_HANCOCK = 10,       // CTL_BEGIN_STEP then CTL_HANCOCK_DT, in one launch
       CTL_END_BEGIN_HANCOCK = 11 }; // CTL_END_HANCOCK of the previous step, then CTL_BEGIN_HANCOCK

}   // namespace fvg
