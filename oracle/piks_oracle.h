/*
 * piks_oracle.h -- CPU restatement of PIKS3D's DVM iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline.  The product (piks3d_amd, include/piks3d.h) never
 * links, loads or calls it.
 *
 * It follows the reference single-rank, single-thread semantics exactly (same
 * arithmetic in the same order, compiled with -ffp-contract=off), keeps the
 * reference's array layouts (f(Ntotal,Nc8,8) column-major, image with two
 * ghost layers) and is checked bit for bit against dumps of the reference
 * built from its own sources (oracle/Makefile, tests/golden/).
 */
#ifndef PIKS_ORACLE_H
#define PIKS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct po_params {
    int32_t nx, ny, nz;          /* physicalNml Nx, Ny, Nz (Nx before fluid layers) */
    int32_t wall_ext_order;      /* physicalNml wallExtOrder */
    int32_t fluid_layer;         /* physicalNml fluidLayer */
    double  ref_l;               /* physicalNml Ref_L */
    int32_t nc_fundamental;      /* velocityNml Nc_fundamental */
    int32_t half_range;          /* velocityNml halfRange */
    int32_t chk_converge_step;   /* solverNml chkConvergeStep */
    double  kn, press_drop, accom; /* flowNml */
    int32_t exact_quadrature;    /* 0: reference tables; 1: exact Gauss-Hermite rules */
} po_params;

typedef struct po_ctx po_ctx;

/* Returns NULL on error and writes a message into err (if non-NULL). */
po_ctx* po_create(const po_params* p, const uint8_t* raw_image, int64_t image_bytes,
                  char* err, int errlen);
void po_destroy(po_ctx* c);

/* nsteps calls of the reference's iterate (src/solver.F90:111-1940). */
void po_iterate(po_ctx* c, int nsteps, int nthreads);

/* The reference's chkConverge (src/output.F90:60-164): out = {mass, permeability,
 * error}; updates the stored mass like the reference does. */
void po_chk_converge(po_ctx* c, double out[3]);

/* Integer facts: [0]Nx [1]Ny [2]Nz (after fluid layers) [3]Nxtotal [4]Nytotal
 * [5]Nztotal [6]Ntotal [7]Nc8 [8]nWall [9]Nfluid [10]nCorner [11..18]Nstencil1..8 */
void po_info(const po_ctx* c, int64_t out[19]);
/* Double facts: [0]ds [1]mu [2]DiffFlux [3]mass */
void po_info_d(const po_ctx* c, double out[4]);

/* Views of the state (owned by the context; valid until po_destroy). */
double*  po_f(po_ctx* c);      /* Ntotal*Nc8*8, reference layout */
double*  po_fw(po_ctx* c);     /* nCorner*Nc8*8 */
double*  po_rho(po_ctx* c);    /* Ntotal each */
double*  po_ux(po_ctx* c);
double*  po_uy(po_ctx* c);
double*  po_uz(po_ctx* c);
int32_t* po_image(po_ctx* c);  /* Ntotal labels */
int32_t* po_which_corner(po_ctx* c);
int32_t* po_vecwall(po_ctx* c);  /* nWall, 1-based local ids like the reference */
double*  po_extcoef(po_ctx* c);  /* nWall*2, column-major like the reference */
int32_t* po_mapf(po_ctx* c);     /* Nfluid, 1-based */
double*  po_cx(po_ctx* c);
double*  po_cy(po_ctx* c);
double*  po_cz(po_ctx* c);
double*  po_w(po_ctx* c);

#ifdef __cplusplus
}
#endif
#endif
