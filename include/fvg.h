/* fvg — MI355X (gfx950) residual and time step of Fenice's uniform Cartesian grid, behind a C ABI.
 *
 * State layout across the boundary: the reference's `Field` (cpp/include/fv/Core.hpp:20-24), i.e. an
 * array-of-structures of cells, cell (i, j, k) at index (k*ny + j)*nx + i, each cell holding
 * ncomp doubles [rho, rho*u, rho*v, rho*E(, rho*w)]. ncomp is 5 (the reference's State) or 4 (a 2-D
 * field without its rho*w slot, the layout of the reference's npz dumps, IO.hpp:153-158). On the device
 * the state lives as structure-of-arrays FP64 planes; the layout change happens only in upload/download.
 *
 * Every entry point returns an fvg_status; the message of the last failure of a context is
 * fvg_last_error(ctx). The reference raises exceptions instead: std::invalid_argument for bad names and
 * values (cpp/src/Schemes.cpp:84,93,101, cpp/src/Boundaries.cpp:27), std::logic_error for broken
 * contracts (cpp/src/Schemes.cpp:398-399), std::runtime_error when a run diverges
 * (cpp/src/SimulationInternal.hpp:259-267). FVG_EINVAL, FVG_ELOGIC and FVG_EDIVERGED carry those three.
 *
 * A context is one device plus one stream. It is not thread-safe: calls on one context must not overlap
 * (the reference's Simulation is not re-entrant either, cpp/include/fv/Physics.hpp:118-119).
 */
#ifndef FVG_H
#define FVG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fvg_status {
    FVG_OK = 0,
    FVG_EINVAL = 1,     /* bad name, value or size (std::invalid_argument in the reference) */
    FVG_ELOGIC = 2,     /* contract violation, e.g. a halo too shallow for the stencil (std::logic_error) */
    FVG_EDIVERGED = 3,  /* non-finite wave speed or time step, or t + dt == t (check_progress, std::runtime_error) */
    FVG_EHIP = 4,       /* a HIP runtime call failed */
    FVG_ENOMEM = 5,     /* device memory could not be allocated */
    FVG_EUNSUPPORTED = 6, /* a configuration the device path does not implement (the message names it) */
    FVG_EIO = 7          /* a file could not be read or written, or is not a checkpoint (std::runtime_error) */
} fvg_status;

/* Boundary-condition types, in the order of the reference's enum BCType (cpp/include/fv/Boundaries.hpp:15). */
enum fvg_bc {
    FVG_BC_WALL = 0,
    FVG_BC_NO_SLIP = 1,
    FVG_BC_MOVING_WALL = 2,
    FVG_BC_VEL_INLET = 3,
    FVG_BC_MASSFLOW_INLET = 4,
    FVG_BC_OUTLET = 5,
    FVG_BC_TRANSMISSIVE = 6,
    FVG_BC_PERIODIC = 7,
    FVG_BC_NONREFLECTING = 8
};
/* Sides, in the order of the reference's enum Face (cpp/include/fv/Core.hpp:44). */
enum fvg_side { FVG_LEFT = 0, FVG_RIGHT = 1, FVG_BOTTOM = 2, FVG_TOP = 3, FVG_FRONT = 4, FVG_BACK = 5 };
enum fvg_riemann { FVG_HLLC = 0, FVG_LOWMACH_HLLC = 1 };                 /* Schemes.hpp:16 */
enum fvg_recon { FVG_RECON_NONE = 0, FVG_RECON_MUSCL = 1, FVG_RECON_WENO3 = 2, FVG_RECON_WENO5 = 3 };
enum fvg_limiter { FVG_LIM_MINMOD = 0, FVG_LIM_VAN_LEER = 1, FVG_LIM_MC = 2, FVG_LIM_KOREN = 3 }; /* Schemes.hpp:82 */
enum fvg_time_scheme { FVG_TS_AUTO = -1, FVG_TS_EULER = 0, FVG_TS_SSPRK2 = 1, FVG_TS_SSPRK3 = 2, FVG_TS_HANCOCK = 3 };
enum fvg_eos { FVG_EOS_IDEAL = 0, FVG_EOS_STIFFENED = 1 };
/* Arithmetic of the reconstruction and the fluxes. FP64 is the reference's. FP32 evaluates reconstruction,
 * Riemann solver and viscous fluxes in float, while the state, the stage combinations, the time step and the
 * conservation sums stay FP64. */
enum fvg_flux_precision { FVG_FLUX_FP64 = 0, FVG_FLUX_FP32 = 1 };

/* Troubled-face indicator of the WENO5-Z reconstruction (no reference key). LEARNED (reconstruction = weno5, a
 * Runge-Kutta scheme, the midpoint quadrature, either flux precision): a small network classifies every face from 15
 * features of its stencil; faces it calls smooth take the linear upwind-5 value instead of WENO5-Z (in the flux
 * precision: FP64 WENO5-Z is the reference's). An audit every indicator_audit steps compares its decisions with the
 * label (the label value of indicator_label_source > indicator_label) and switches the indicator off
 * for the rest of the context's life when the missed troubled faces exceed both indicator_fn_limit of the troubled
 * faces and indicator_fn_floor of all the audited faces (the second keeps a handful of isolated label flips in a
 * smooth flow, e.g. where a velocity component crosses zero, from switching it off). */
enum fvg_indicator { FVG_INDICATOR_NONE = 0, FVG_INDICATOR_LEARNED = 1 };
/* Label sources of the indicator (key indicator_label_source): WENO5, the departure of the WENO5-Z weights from the
 * linear weights (max over the variables and both face states of max_k |w_k - d_k|; weno5z, Schemes.cpp:338-353);
 * MUSCL, how far the MUSCL limiter of the configured muscl_limiter (apply_limiter, Schemes.hpp:110-117, as the
 * reconstruction applies it at Schemes.cpp:261-262) moves a face value's slope off the central slope,
 * |phi(a, b) - (a + b) / 2| / sqrt((a^2 + b^2) / 2 + weno_epsilon * m + 1e-30) with a, b the one-sided differences
 * and m the regularisation scale of the departure; BOTH, the larger of the two. A face is troubled when its label
 * value exceeds indicator_label. */
enum fvg_label_source { FVG_LABEL_WENO5 = 0, FVG_LABEL_MUSCL = 1, FVG_LABEL_BOTH = 2 };

/* Flux quadrature over a face (flux_quadrature, Params.cpp:148, 448-455). FOURTH: WENO3-Z / WENO5-Z of the conserved
 * averages, point values and a fourth-order transverse correction of the face fluxes (Schemes.cpp:1954-2134); needs
 * reconstruction = weno3 or weno5. */
enum fvg_flux_quadrature { FVG_QUAD_MIDPOINT = 0, FVG_QUAD_FOURTH = 1 };

/* The keys of the path (cpp/src/Params.cpp:98-223 and SimParams::from), typed. Defaults: fvg_config_defaults. */
typedef struct fvg_config {
    int nx, ny, nz;              /* cells; nz = 1 is a 2-D mesh */
    double Lx, Ly, Lz;           /* extent */
    int eos_type;                /* enum fvg_eos */
    double gamma, p_inf;
    int bc[6];                   /* enum fvg_bc per side, enum fvg_side order (bc_xmin, bc_xmax, bc_ymin, bc_ymax, bc_zmin, bc_zmax) */
    double u_in, v_in, rho_in, p_out; /* BC parameters (moving wall, inlets, outlet) */
    int riemann;                 /* enum fvg_riemann */
    double shock_blend;          /* default 0.2 */
    int reconstruction;          /* enum fvg_recon */
    int limiter;                 /* enum fvg_limiter */
    double weno_epsilon;         /* default 1e-6 */
    int boundary_order;          /* 1 or 2 (default 2) */
    int time_scheme;             /* enum fvg_time_scheme; FVG_TS_AUTO is resolved by fvg_config_finalize */
    double cfl;                  /* default 0.5 */
    int use_viscous;
    double mu, pr;               /* defaults 1e-3, 0.72 */
    double t_end, dt_max;        /* run keys */
    int max_steps;
    int flux_precision;          /* enum fvg_flux_precision (no reference key; default FP64) */
    int indicator;               /* enum fvg_indicator (no reference key; default NONE) */
    int indicator_audit;         /* steps between audits (default 50; 0: no audit) */
    double indicator_label;      /* weight departure above which a face is troubled (default 0.05) */
    double indicator_fn_limit;   /* fallback limit of the missed fraction of troubled faces (default 0.01) */
    double indicator_fn_floor;   /* fallback limit of the missed fraction of all audited faces (default 1e-4) */
    int flux_quadrature;         /* enum fvg_flux_quadrature (default MIDPOINT) */
    double nrbc_sigma;           /* non-reflecting outlet: relaxation of p towards p_out (default 0.28; 0: none) */
    double nrbc_beta;            /* non-reflecting outlet: transverse weight, <= 1 (default -1: the local Mach number) */
    int indicator_label_source;  /* enum fvg_label_source: what makes a face troubled for the audit and the training
                                    samples (no reference key; default WENO5) */
} fvg_config;

/* The reference's defaults (Params.cpp parameter_table): a 2-D 1x1 domain is not meaningful, so nx, ny,
 * nz, Lx, Ly, Lz are set to 0 / 0.0 / 1 / 1.0 and must be given. */
void fvg_config_defaults(fvg_config* cfg);
/* Set one key from its text value, with the reference's key names and case-insensitive labels and aliases
 * ("HLLC (Low-Mach Thornber)", "lowmach", "Van Leer", "SSP Runge-Kutta 3", "periodic", ...). Replaces
 * SimParams::from + parse_riemann / parse_limiter / parse_reconstruction / parse_time_scheme / parse_bc
 * (Params.cpp, Schemes.cpp:77-102, TimeScheme.cpp:78-85, Boundaries.cpp:15-27).
 * The reference's other keys are accepted: run control of its driver (test_case, save_interval, monitor_interval,
 * resume, dump_state, ...) has no effect, immersed-body and AMR parameters are inert; switching on a feature outside
 * this path (use_ibm, use_amr, use_mpi, use_bodyforce, positivity_fallback, a source term, reynolds, a boundary
 * value table) returns FVG_EUNSUPPORTED. Returns FVG_EINVAL with a message in err (may be NULL) for an unknown key
 * or label. */
int fvg_config_set(fvg_config* cfg, const char* key, const char* value, char* err, size_t err_len);
/* Resolve time_scheme = auto (Params.cpp: MUSCL-Hancock for inviscid MUSCL, SSP-RK2 for viscous MUSCL,
 * Explicit Euler otherwise) and check sizes. */
int fvg_config_finalize(fvg_config* cfg, char* err, size_t err_len);
/* The CFL length L of dt = cfl L / (S + 1e-10): min(dx, dy) in 2-D, 2 / (1/dx + 1/dy + 1/dz) in 3-D
 * (Simulation.cpp:223-227, ViscousBound.hpp:28-29). */
double fvg_dt_length(const fvg_config* cfg);

typedef struct fvg_ctx fvg_ctx;

/* A context on HIP device `device` for the mesh and numerics of cfg: allocates the device state and work
 * buffers. Replaces Simulation::setup for the uniform path (Simulation.cpp:188). */
int fvg_create(const fvg_config* cfg, int device, fvg_ctx** out);
int fvg_destroy(fvg_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed fvg_create when ctx is NULL). */
const char* fvg_last_error(const fvg_ctx* ctx);

/* Host AoS state in / out (ncomp 4 or 5, ncells = nx*ny*nz). Synchronous. */
int fvg_upload(fvg_ctx* ctx, const double* U, size_t ncells, int ncomp);
int fvg_download(fvg_ctx* ctx, double* U, size_t ncells, int ncomp);

/* R = residual(U) of the resident state and the largest wave speed it saw: Simulation::residual
 * (Simulation.cpp:1351) as the ResidualFn of TimeScheme.hpp:14. predictor_half_dt > 0 runs the
 * MUSCL-Hancock predictor inside it (TimeScheme.hpp:92). R (may be NULL) receives the host AoS residual
 * with ncomp components. */
int fvg_residual(fvg_ctx* ctx, double predictor_half_dt, double* R, int ncomp, double* s_max);

/* One step of the configured scheme on the resident state: TimeScheme::step(U, residual, dt_length, t,
 * t_max, dt_max) -> {dt, s_max} (TimeScheme.hpp:17-20,48), followed by the reference's divergence check
 * (check_progress). Returns FVG_EDIVERGED (state left as the step made it) when the check fails. */
int fvg_step(fvg_ctx* ctx, double t, double t_max, double dt_max, double* dt, double* s_max);

/* Up to max_steps steps from time *t toward t_end without a host round trip per step (time, dt and the
 * divergence flag stay on the device; the run loop condition is `t < t_end`, SimulationRun.cpp:167).
 * max_steps <= 0: no step limit. On return *t is the new time and *steps the number of steps taken; s_last
 * (may be NULL) gets the last step's wave speed. */
int fvg_advance(fvg_ctx* ctx, int max_steps, double* t, double t_end, double dt_max, int* steps, double* s_last);

/* The MUSCL-Hancock lagged wave speed (TimeScheme.hpp:96-97; checkpoint key "S"). */
int fvg_get_lagged_speed(const fvg_ctx* ctx, double* s);
int fvg_set_lagged_speed(fvg_ctx* ctx, double s);

/* ---- Checkpoints (the reference's restart files) -------------------------------------------------------------
 * A checkpoint is the reference's uniform-mesh restart file: a .npz of stored .npy entries with U (ncells x 4 in
 * 2-D, x 5 in 3-D; FP64, the AoS Field), t, step, format_version = 1, nx, ny, nz, Lx, Ly, Lz, the output index
 * (collection_time, collection_file; empty here) and S, the MUSCL-Hancock lagged speed, when positive
 * (Simulation::checkpoint_header / save_checkpoint / load_checkpoint, SimulationRestart.cpp:61-85, 162-190, 254-360).
 * A run continued from a checkpoint takes the same steps, bit for bit, as the run that wrote it, on either side.
 * Errors: FVG_EIO (file, format), FVG_EINVAL (a mesh or size that differs from cfg / the context's). Entries and
 * offsets of 4 GiB and more carry zip64 fields; the environment variable FVG_NPZ_ZIP64_FROM (bytes) lowers that
 * size, so that tests can exercise them on small files. */
/* Writes U (ncells x ncomp, ncomp 4 or 5) with t, step and lagged_speed (written as S when > 0) as `path`,
 * through a temporary file renamed over it. Host only; err (may be NULL) gets the message. */
int fvg_checkpoint_save(const fvg_config* cfg, const double* U, size_t ncells, int ncomp, double t, int step, double lagged_speed,
                        const char* path, char* err, size_t err_len);
/* Reads a checkpoint of a cfg mesh into U (ncells x ncomp; a file with 4 columns gives rho w = 0 in 3-D), t, step
 * and lagged_speed (0 when the file has no S: the next MUSCL-Hancock step seeds it). Host only. */
int fvg_checkpoint_load(const fvg_config* cfg, const char* path, double* U, size_t ncells, int ncomp, double* t, int* step,
                        double* lagged_speed, char* err, size_t err_len);
/* The context's state and lagged speed to / from a checkpoint (fvg_download / fvg_upload around the two above). */
int fvg_checkpoint_write(fvg_ctx* ctx, const char* path, double t, int step);
int fvg_checkpoint_read(fvg_ctx* ctx, const char* path, double* t, int* step);

/* Volume averages of mass, kinetic energy and total energy (Simulation::volume_averages,
 * SimulationRun.cpp:45-64). The device sum is a fixed-order FP64 tree, so it matches the reference's
 * sequential sum to round-off, not bit for bit. */
int fvg_conserved_totals(fvg_ctx* ctx, double out[3]);

/* The volume averages of fvg_conserved_totals for a state the caller holds on the device, enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) without allocating anything: no hipMalloc, no hipFree, no context. It has no
 * counterpart in the reference; it is for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). U: ncomp (4 or 5) structure-of-arrays FP64 planes of ncells values each,
 * plane q at U + q * ncells, in the component order of a Field; out: 3 doubles of device memory (mass, kinetic
 * energy, total energy), written when the stream reaches the call. Same kernels and block count as
 * fvg_conserved_totals, so the same bits. Its partial sums live in one fixed device array per GPU of
 * FVG_NOALLOC_TOTALS_SCRATCH bytes that every call on that GPU shares: calls that overlap on one GPU (other
 * threads, other streams) race. 1 <= ncells <= FVG_NOALLOC_MAX_CELLS, else FVG_EINVAL. The first call on a device
 * looks up that device's array (hipGetSymbolAddress) and makes the device current meanwhile; make it outside a
 * capture. Returns without waiting for the stream. */
#define FVG_NOALLOC_TOTALS_SCRATCH 24576
#define FVG_NOALLOC_MAX_CELLS 2147483647LL
int fvg_conserved_totals_noalloc(int device, const double* U, long long ncells, int ncomp, double* out, void* stream);

/* The numbers of a row of the reference's monitor file (`<prefix>_monitor.csv`: step, t, mass, kinetic_energy,
 * total_energy, res_rho, res_mom, res_E; Simulation::monitor_row and its uniform-mesh caller,
 * SimulationRun.cpp:76-131, 413-424): out[0..2] the volume averages of the state (fvg_conserved_totals), out[3..5] the
 * root mean squares over the cells of the plain residual of the state for density, momentum (all components) and
 * energy. The residual is the reference's bits; its sums are a fixed-order tree, so the norms match the reference's
 * sequential sums to round-off. */
int fvg_monitor(fvg_ctx* ctx, double out[6]);

/* Timing of the kernels of the last fvg_advance / fvg_step calls, measured with HIP events on the
 * context's stream: name[k] (static strings), total milliseconds and launch count per kernel family.
 * Returns the number of families written (at most cap). */
int fvg_kernel_times(const fvg_ctx* ctx, const char** names, double* ms, long long* launches, int cap);
/* Switch the per-kernel event timing on (1) or off (0, the default: no events are recorded). */
int fvg_set_timing(fvg_ctx* ctx, int on);
/* Launch mode of fvg_advance (no counterpart in the reference): 1 (default) replays a HIP graph captured once per
 * context that holds four steps of the scheme; 0 enqueues every kernel. Results are the same bits either way;
 * with event timing on, kernels are always enqueued one by one. */
int fvg_set_graphs(fvg_ctx* ctx, int on);
/* Kernel sequence of the residual (no counterpart in the reference): 1, fused passes, in which no face flux goes to
 * device memory: the x faces' differences go straight into per-cell partial sums, the y faces' are added to them,
 * and the last axis' pass finishes the residual and applies the stage combination; 0 (default), separate face
 * passes storing the fluxes and one update pass. The same bits either way; on MI355X the separate passes measured
 * as fast for FP64 fluxes and faster for FP32 fluxes and the learned indicator. The fourth-order quadrature always
 * takes its own passes. */
int fvg_set_fused(fvg_ctx* ctx, int on);

/* Known-answer entry: the Riemann flux of n faces with host AoS conservative states (5 doubles per face),
 * axis per face, through the device kernel of the solver configured in cfg (riemann, shock_blend, EOS).
 * F receives 5 doubles per face and smax one. Replaces RiemannSolver::face_flux_batch (Schemes.hpp:43). */
int fvg_riemann_batch(const fvg_config* cfg, int device, const double* UL, const double* UR, const int* axis, size_t n,
                      double* F, double* smax);

/* ---- Learned troubled-face indicator (no reference counterpart; see enum fvg_indicator) -----------------------
 * The network: 15 features -> 32 (ReLU) -> 32 (ReLU) -> 1 logit; parameters as one float vector of
 * FVG_MLP_NPARAM values: W1 (32 x 15, row-major), b1 (32), W2 (32 x 32), b2 (32), w3 (32), b3. A face is troubled
 * when its logit exceeds the threshold. Features: for density, pressure and the velocity component with the largest
 * total variation over the face's six-cell stencil, the five differences of neighbouring cells divided by
 * sqrt(sum of their squares + weno_epsilon * mean square of the six values + 1e-30). */
#define FVG_MLP_NPARAM 1601
/* Weights of a context created with indicator = learned (until set, every face is troubled); resets the audit
 * counters and switches the indicator back on. */
int fvg_indicator_set(fvg_ctx* ctx, const float* params, int nparam, float threshold);
/* counts: TP, FP, FN, TN of the last audit, then the same over all audits (positive = troubled);
 * off: 1 once an audit switched the indicator off; audits: audits so far. Any output may be NULL. */
int fvg_indicator_stats(fvg_ctx* ctx, unsigned long long counts[8], int* off, int* audits);
/* Training samples of the resident state (a context with flux_precision = fp32): for each face normal to axis,
 * its 16 feature slots (the 16th is 1) in X[16 * nfaces] and its label value (the largest WENO5-Z weight
 * departure over the variables and both face states) in dep[nfaces]; -1 for faces without a full stencil. */
int fvg_indicator_samples(fvg_ctx* ctx, int axis, float* X, float* dep, long long nfaces);

/* Training on the device: weighted binary cross-entropy with logits (troubled samples weighted pos_weight), Adam.
 * X holds 16 floats per sample (the first 15 are the features), y 0 or 1. */
typedef struct fvg_trainer fvg_trainer;
int fvg_trainer_create(int device, long long capacity, fvg_trainer** out);
int fvg_trainer_destroy(fvg_trainer* t);
const char* fvg_trainer_last_error(const fvg_trainer* t);
int fvg_trainer_set_data(fvg_trainer* t, const float* X, const float* y, long long n);
/* Mean loss and its gradient over all samples (FVG_MLP_NPARAM floats), the same bits on every call. */
int fvg_trainer_grad(fvg_trainer* t, const float* params, float pos_weight, float* grad, double* loss);
/* epochs of shuffled mini-batches from params (updated in place); loss (may be NULL): the final mean loss. */
int fvg_trainer_fit(fvg_trainer* t, float* params, int epochs, int batch, float lr, float pos_weight, unsigned long long seed,
                    double* loss);
/* Logits of n feature rows (16 floats each) through the face kernels' inference routine (bfloat16 matrix cores). */
int fvg_mlp_logits(int device, const float* params, float threshold, const float* X, long long n, float* z);

/* ---- Ensemble of independent 1-D runs -------------------------------------------------------------------
 * ninst runs of the reference's solver, instance i on an ncells[i] x 1 mesh (nx = ncells[i], ny = 1, periodic
 * y), all with the scheme, x boundaries, Lx and Ly of cfg: each is Simulation::run_loop
 * (SimulationRun.cpp:161-191) over TimeScheme::step (TimeScheme.cpp:110-245) with check_progress
 * (SimulationInternal.hpp:259-267), run to its own t_end; the reference would run them one process each. The
 * whole ensemble is one kernel launch (one workgroup per instance). Instances keep (rho, rho u, E); their
 * rho v is +0 throughout, as in the reference's run. Explicit Euler / SSP-RK2 / SSP-RK3, Euler equations,
 * FP64; x sides wall, no_slip, transmissive, outlet or periodic. */
typedef struct fvg_batch fvg_batch;

int fvg_batch_create(const fvg_config* cfg, int ninst, const int* ncells, int device, fvg_batch** out);
/* An ensemble of 2-D runs: instance i on an nx[i] x ny[i] mesh with the scheme, the four sides (any but mass-flow
 * inlets and non-reflecting outlets), Lx and Ly of cfg, e.g. isentropic vortices of uneven resolution. The same
 * persistent kernel layout (one workgroup per instance); the instances keep (rho, rho u, rho v, E), and upload and
 * download take ncomp 4 or 5 (rho w written as +0). The other fvg_batch_* entries apply unchanged. */
int fvg_batch_create_2d(const fvg_config* cfg, int ninst, const int* nx, const int* ny, int device, fvg_batch** out);
int fvg_batch_destroy(fvg_batch* b);
/* Message of the last failure of b (of the last failed fvg_batch_create when b is NULL). */
const char* fvg_batch_last_error(const fvg_batch* b);
/* The instances' states, concatenated in instance order: ncells = sum of ncells[i]; ncomp 3 (rho, rho u, E),
 * 4 (the reference's 2-D State: rho, rho u, rho v, E) or 5. Download writes rho v (and rho w) as +0. */
int fvg_batch_upload(fvg_batch* b, const double* U, long long ncells, int ncomp);
int fvg_batch_download(fvg_batch* b, double* U, long long ncells, int ncomp);
/* Every instance from t = 0 while t < t_end and fewer than max_steps steps (max_steps > 0 is required: it bounds
 * the kernel). Per instance: final time, steps taken, diverged flag (1: check_progress failed on the last step,
 * whose state is left as the step made it), last step's wave speed; ms (may be NULL) the kernel time. Any
 * output pointer may be NULL. Divergence of an instance is reported through its flag, not the status. */
int fvg_batch_run(fvg_batch* b, double t_end, double dt_max, int max_steps, double* t, int* steps, int* diverged, double* s_last,
                  double* ms);
/* Workgroup size of the ensemble kernel: 64, 128, 256 (default), 512, 768 or 1024. Results do not depend on it. */
int fvg_batch_set_block(fvg_batch* b, int block);
int fvg_batch_get_block(const fvg_batch* b, int* block);
/* Autotuner: time every workgroup size on a probe run of the uploaded state (to t_probe, at most max_steps),
 * keep the fastest and restore the state. With cache_path, a line "key block ms" per tuned workload (device
 * architecture, instance sizes, scheme) is appended, and a later call for the same key reads it instead of
 * measuring (*from_cache = 1). */
int fvg_batch_tune(fvg_batch* b, double t_probe, int max_steps, const char* cache_path, int* block, int* from_cache);
/* Teams of 1-D and 2-D instances (no reference counterpart: the reference runs one process per instance). The
 * largest instances run on several workgroups of one cooperative launch, each over its own chunk of at least
 * min_chunk cells, with two barriers among a team's workgroups per stage; the other instances are taken one at a time, longest
 * first, by any workgroup. An instance gets a team when its estimated work is at least twice the average per resident
 * workgroup (as many workgroups as averages, at most max_team; teams use at most half of the workgroups); with
 * force, every instance of at least 2 min_chunk cells gets min(max_team, N / min_chunk). The chunks of a 2-D
 * instance are slabs of whole rows, at least 3 rows and min_chunk cells each. Results do not depend on the teams. Waits at a team barrier are bounded: a team that times out makes fvg_batch_run return FVG_EHIP.
 * max_team = 1, first-order faces or a device without cooperative launches: one workgroup per instance.
 * Defaults: max_team 8, min_chunk 512, force 0. max_team <= 16, min_chunk >= 3. */
int fvg_batch_set_teams(fvg_batch* b, int max_team, int min_chunk, int force);
/* The last fvg_batch_run's launch: teams, workgroups in teams, workgroups launched. */
int fvg_batch_get_plan(const fvg_batch* b, int* nteams, int* team_wgs, int* grid);

#ifdef __cplusplus
}
#endif

#endif /* FVG_H */
