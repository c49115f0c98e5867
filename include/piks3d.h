/*
 * piks3d.h -- C ABI of the MI355X DVM solver (libpiks3d.so).
 *
 * The reference (iPACT-Platform/PIKS3D) has no library interface: its hot path
 * is `subroutine iterate` (src/solver.F90:111) and `subroutine chkConverge`
 * (src/output.F90:60), called from the main loop in src/main.F90:95-128 and
 * operating on module globals set up by src/main.F90:75-89.  This header is the
 * boundary a host program (the Fortran host in piks3d_amd/fortran, the Python
 * package, or any ISO_C_BINDING / ctypes caller) binds instead.  Plain C types
 * only; no exceptions cross it.
 *
 * Conventions
 *  - Every call returns 0 on success or a negative PIKS_E* code; the message
 *    of the last failure on the calling thread is in piks_last_error().
 *  - Arrays returned to the host use the reference's layouts so its writers
 *    and checks can consume them unchanged: cell arrays have Ntotal =
 *    (Nx+4)(Ny+4)(Nz+4) entries with index l = (i+2) + (j+1)(Nx+4) +
 *    (k+1)(Nx+4)(Ny+4) - 1 (0-based) for i, j, k in [-1, N+2]
 *    (src/physicalGrid.F90:704, with Nx already including the fluid layers);
 *    f is f(Ntotal, Nc8, 8) in Fortran order (src/flow.F90:69,85).
 *  - A context is not re-entrant: call it from one host thread at a time
 *    (the reference is MPI_THREAD_FUNNELED, src/main.F90:70).
 */
#ifndef PIKS3D_H
#define PIKS3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIKS_OK 0
#define PIKS_E_ARG (-1)      /* invalid parameter or input */
#define PIKS_E_IMAGE (-2)    /* image byte not 0/1, or too short */
#define PIKS_E_VGRID (-3)    /* velocity grid not provided */
#define PIKS_E_GEOM (-4)     /* geometry the reference cannot run */
#define PIKS_E_HIP (-5)      /* HIP runtime failure */
#define PIKS_E_MEM (-6)      /* device or host memory exhausted / over budget */
#define PIKS_E_STATE (-7)    /* call not valid in this state */

#define PIKS_PREC_FP64 0       /* f stored and swept in double */
#define PIKS_PREC_FP32_REFINED 1 /* FP64 defect correction: FP32 correction iterations between
                                    FP64 residual iterations (see piks_set_precision) */

/* para.in contents (src/parameters.F90:78-82), after defaults.  nx is the
 * image width before the fluid layers are added (src/parameters.F90:121). */
typedef struct piks_params {
    int32_t nx, ny, nz;            /* &physicalNml Nx, Ny, Nz */
    int32_t wall_ext_order;        /* &physicalNml wallExtOrder (1, 2 or 3) */
    int32_t fluid_layer;           /* &physicalNml fluidLayer */
    double  ref_l;                 /* &physicalNml Ref_L */
    int32_t nc_fundamental;        /* &velocityNml Nc_fundamental */
    int32_t half_range;            /* &velocityNml halfRange (0/1) */
    int32_t chk_converge_step;     /* &solverNml chkConvergeStep */
    double  kn;                    /* &flowNml Kn */
    double  press_drop;            /* &flowNml pressDrop */
    double  accom;                 /* &flowNml accom */
    /* additions of this library (optional &gpuNml) */
    int32_t precision;             /* PIKS_PREC_* */
    int32_t device;                /* HIP device ordinal */
    int32_t exact_quadrature;      /* 1: exact Gauss-Hermite rules instead of the
                                      reference's tables (needed for Nc_fundamental=3) */
    int32_t exact_order;           /* 1: wall-flux and moment sums run serially in the
                                      reference's order, so every iteration is bit-for-bit
                                      the reference's (slower; for verification) */
    int32_t sweep;                 /* sweep schedule: 0 the library's default, 1 one launch per
                                      hyperplane, 2 tiles (one launch per tile diagonal; grids
                                      of at most 64 velocities per octant, FP64 contexts) */
    int32_t tile;                  /* sweep = 2: tx | ty << 8 | tz << 16 (tx * ty one of
                                      4x4, 4x2, 8x2, 2x2; 0: 2x2x2) */
    int32_t reserved[2];
} piks_params;

typedef struct piks_info {
    int32_t nx, ny, nz;            /* domain after fluid layers */
    int32_t nxt, nyt, nzt;         /* with two ghost layers per side */
    int64_t ntotal;                /* nxt * nyt * nzt */
    int32_t nc8;                   /* velocities per octant */
    int32_t nwall, ncorner;        /* wall cells; 3-fold corners */
    int64_t nfluid;                /* fluid cells (|mapF|) */
    int64_t updates_per_iter;      /* nfluid * 8 * nc8 (one update = one cell-velocity pair) */
    double  ds, mu, diff_flux;     /* grid spacing, sqrt(pi/2)/Kn, 4 sum cz w */
    double  mass;                  /* last mass-flow value (1 before the first check) */
    int64_t device_bytes;          /* device memory held by the context */
    int32_t coupled_walls;         /* walls whose extrapolation reads another wall */
    int32_t precision;
    int64_t steps_fp32;            /* FP32 correction iterations run (PIKS_PREC_FP32_REFINED) */
    int32_t dc_block;              /* current defect-correction block length (1: plain FP64) */
    int32_t sweep_path;            /* FP64 sweep: 0 one launch per hyperplane, 1 tiles, 2 the
                                      whole sweep in one launch (no counterpart in the reference) */
} piks_info;

/* Set-up: replaces setupVelocityGrid + setupPhysicalGrid + setupFlow +
 * allocateBuf (src/main.F90:82-89; src/velocityGrid.F90:90-192,
 * src/physicalGrid.F90:217-1625, src/flow.F90:76-163).  raw_image holds the
 * image file's bytes (nx*ny*nz, x fastest; fluid layers are added here).
 * Fails with PIKS_E_IMAGE on a byte other than 0/1 (src/physicalGrid.F90:268-275),
 * PIKS_E_VGRID for an unsupported Nc_fundamental (src/velocityGrid.F90:149-153),
 * PIKS_E_GEOM for a 3-fold corner within two x-layers of the inlet/outlet
 * (the reference's x self-message then overruns its receive and MPI aborts). */
int piks_create(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                void** ctx_out);

/* nsteps source iterations; replaces `call iterate` (src/solver.F90:111-1940,
 * called from src/main.F90:100). */
int piks_iterate(void* ctx, int32_t nsteps);

/* Precision mode of the next iterations, for a context created with
 * PIKS_PREC_FP32_REFINED (a PIKS_PREC_FP64 context accepts only FP64).
 * PIKS_PREC_FP32_REFINED is an FP64 defect correction of the reference's
 * iteration, an affine map f <- T(f) (src/solver.F90:111-1940, the source
 * being the pressure drop of src/solver.F90:1832-1862): piks_iterate runs
 * blocks of m iterations, one FP64 iteration of the state S that yields the
 * residual r = T(S) - S, then m - 1 FP32 iterations of the correction
 * e <- A(e + r) with the linear part A (pressure drop 0); S + e + r then
 * equals the plain iteration's state in exact arithmetic, so FP32 rounds
 * only the correction and every check tracks the FP64 solve (within 1e-8 on
 * the permeability).  m follows each piks_chk_converge (1 while the relative
 * change per iteration is above 1e-2, then about sqrt(0.1 / change), at
 * most 200; PIKS_DC_BLOCK in the environment fixes it).  Moments, wall
 * densities and the mass-flow check stay FP64.  Switching to PIKS_PREC_FP64
 * folds the correction into the FP64 state.  No counterpart in the
 * reference, which is FP64 only. */
int piks_set_precision(void* ctx, int32_t precision);

/* Defect-correction block length of a PIKS_PREC_FP32_REFINED context: m >= 1
 * fixes it from the next block on (1: plain FP64 iterations), m = 0 returns
 * to the length that follows the checks.  No counterpart in the reference. */
int piks_set_correction_block(void* ctx, int32_t m);

/* Convergence check; replaces chkConverge (src/output.F90:60-164):
 * out[0] = mass flow, out[1] = apparent permeability, out[2] = error, and the
 * stored mass becomes out[0], exactly like the reference. */
int piks_chk_converge(void* ctx, double out[3]);

/* Mass flow through the plane x = 2 without changing the stored mass
 * (massLocal of src/output.F90:140-142). */
int piks_mass_flow(void* ctx, double* mass);

/* piks_mass_flow for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): no counterpart in the reference.  The same
 * value, bit for bit, written to the device double mass_dev; enqueued on the
 * context's stream without synchronising and without allocating anything
 * (no hipMalloc, no host allocation): the plane is staged in a fixed device
 * array of the library of 4 MiB per device (at most 524288 cells, Ny * Nz),
 * larger planes fail with PIKS_E_ARG.  The array is shared by every call on
 * a device and keeps the last call's plane: calls that overlap on one device
 * (several threads or streams) race. */
int piks_mass_flow_noalloc(void* ctx, double* mass_dev);

/* Macroscopic fields Rho, Ux, Uy, Uz (src/flow.F90:72), Ntotal doubles each,
 * as the reference's writers read them (src/output.F90:176-316).  NULL
 * pointers are skipped. */
int piks_get_macro(void* ctx, double* rho, double* ux, double* uy, double* uz);

/* f (src/flow.F90:69) as doubles in the reference layout f(Ntotal, Nc8, 8);
 * fw (src/physicalGrid.F90:192) as fw(nCorner, Nc8, 8). */
int piks_get_f(void* ctx, double* f);
int piks_get_fw(void* ctx, double* fw);

/* Cell labels (`image`, src/physicalGrid.F90:124-134), Ntotal int32. */
int piks_get_labels(void* ctx, int32_t* labels);

int piks_get_info(void* ctx, piks_info* info);

/* Host-only image preprocessing (no device needed): the labelled image of
 * src/physicalGrid.F90:251-671 (Ntotal int32 into labels, may be NULL) and
 * its counts in info (nx..ntotal, nc8 = 0, nwall, ncorner, nfluid,
 * coupled_walls).  Same error behaviour as piks_create. */
int piks_preprocess(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                    int32_t* labels, piks_info* info);

/* Several right-hand sides on one geometry: nrhs solves of the same image,
 * velocity grid and parameters that differ only in Kn (kn[0..nrhs-1]; p->kn is
 * ignored), advanced together by piks_iterate with one sweep launch per
 * hyperplane for all of them (each task's geometry record is read once for
 * all right-hand sides).  No counterpart in the reference, which solves one Kn
 * per run; each right-hand side's iterates are those of piks_create with its
 * Kn.  piks_chk_converge_multi fills out[3*r + 0..2] = mass, permeability,
 * error of right-hand side r (src/output.F90:151-155 for each); the other
 * calls (piks_chk_converge, piks_get_macro, piks_get_f, ...) see right-hand
 * side 0. */
int piks_create_multi(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                      int32_t nrhs, const double* kn, void** ctx_out);
int piks_chk_converge_multi(void* ctx, double* out);
int piks_get_macro_rhs(void* ctx, int32_t rhs, double* rho, double* ux, double* uy, double* uz);

/* Device time of the last piks_iterate call, summed over its steps (the first
 * 4096) and split by phase (milliseconds): [0] sweeps, [1] switch commit +
 * exchange + wall BC + inlet/outlet + symmetry, [2] moments, [3] first step's
 * start to last step's end.  HIP events on the context's stream. */
int piks_last_timing(void* ctx, double ms[4]);

int piks_destroy(void* ctx);

/* Message of the last failure on this thread ("" if none). */
const char* piks_last_error(void);

/* ------------------------------------------------------------------ request server
 * Continuous batching of independent solves on one GPU.  No counterpart in
 * the reference, which runs one solve per program (src/main.F90); each
 * request runs the reference's loop (src/main.F90:95-121): iterations up to
 * the next check, chkConverge, stop at error <= eps or max_step.  Requests
 * are preprocessed on host threads, admitted in arrival order whenever a
 * slot is free and their device footprint fits the memory budget, and run
 * concurrently, each on its own stream with its iteration captured as a HIP
 * graph; a finished request releases its slot and memory to the next one.
 * Results are bit-for-bit those of piks_create + piks_iterate +
 * piks_chk_converge on the same request.  All calls are thread-safe. */
typedef struct piks_serve_params {
    int32_t device;            /* HIP device ordinal */
    int32_t max_slots;         /* solves on the device at once (<= 64; 0: 8) */
    int32_t prep_threads;      /* host preprocessing threads (0: 4) */
    int32_t use_graphs;        /* 1: each iteration is one HIP graph launch */
    int64_t memory_budget;     /* device bytes for admitted solves (0: 90% of free memory) */
    int32_t reserved[4];
} piks_serve_params;

#define PIKS_REQ_QUEUED 0      /* waiting for preprocessing or admission */
#define PIKS_REQ_RUNNING 1
#define PIKS_REQ_DONE 2
#define PIKS_REQ_FAILED 3

typedef struct piks_serve_result {
    int32_t status;            /* PIKS_REQ_* */
    int32_t error_code;        /* PIKS_E_* when FAILED */
    int32_t steps;             /* iterations run */
    int32_t steps32;           /* of which FP32 correction iterations (PIKS_PREC_FP32_REFINED) */
    int32_t converged;         /* 1: error <= eps was reached */
    int32_t reserved0;
    double  permeability;      /* last check's K (src/output.F90:155) */
    double  mass;
    double  error;
    double  queue_ms;          /* submit -> admitted (includes host preprocessing) */
    double  run_ms;            /* admitted -> done */
    double  total_ms;          /* submit -> done */
    int64_t device_bytes;      /* device memory the request held */
} piks_serve_result;

int piks_serve_create(const piks_serve_params* sp, void** server_out);
/* Queue one solve; p and the image as for piks_create.  refine_switch is
 * ignored (kept for binary compatibility; FP32_REFINED requests run the
 * defect correction of piks_set_precision). */
int piks_serve_submit(void* server, const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                      int32_t max_step, double eps, double refine_switch, int64_t* ticket_out);
int piks_serve_poll(void* server, int64_t ticket, piks_serve_result* out);   /* no wait */
int piks_serve_wait(void* server, int64_t ticket, piks_serve_result* out);   /* until done or failed */
int piks_serve_drain(void* server);                                          /* until all are */
int piks_serve_destroy(void* server);

#ifdef __cplusplus
}
#endif
#endif
