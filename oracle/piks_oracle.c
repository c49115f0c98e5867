/*
 * piks_oracle.c -- CPU restatement of PIKS3D's DVM iteration (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * file, as the checker or as the timed CPU baseline.  It restates, single rank,
 * the reference's algorithm for the hot path with the same floating-point
 * operations in the same order (build with -ffp-contract=off), in the
 * reference's layouts, so it can be pinned bit for bit against dumps of the
 * reference built from its own sources (oracle/Makefile).  Each function names
 * the reference lines it follows (paths relative to the reference checkout).
 *
 * Single-rank facts used throughout (src/mpiParams.F90:159-201 with
 * mpi_{x,y,z}dim = 1): xl = 1, xu = Nx, xlg = -1, xug = Nx + 2 and likewise in
 * y and z; x is periodic through a self send/receive; y and z neighbours are
 * MPI_PROC_NULL so their receive buffers stay zero (src/flow.F90:151-156).
 */
#include "piks_oracle.h"
#include "gauss_hermite_tables.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* labels, src/physicalGrid.F90:127-134 */
enum { FLUID = 0, SOLID = 1, GHOST = 4 };
enum {
    L_E = 20, L_W = 21, L_N = 22, L_S = 23, L_F = 24, L_B = 25,
    L_EN = 30, L_WN = 31, L_ES = 32, L_WS = 33,
    L_NF = 40, L_NB = 41, L_SF = 42, L_SB = 43,
    L_EF = 50, L_EB = 51, L_WF = 52, L_WB = 53,
    L_ENF = 60, L_WNF = 61, L_WSF = 62, L_ESF = 63,
    L_ENB = 70, L_WNB = 71, L_WSB = 72, L_ESB = 73
};

/* octants 1..8 as 0..7; signs of (cx, cy, cz), solver.F90:174,220,265,309,353,397,441,485 */
static const int SX[8] = {+1, -1, -1, +1, +1, -1, -1, +1};
static const int SY[8] = {+1, +1, -1, -1, +1, +1, -1, -1};
static const int SZ[8] = {+1, +1, +1, +1, -1, -1, -1, -1};

struct po_ctx {
    po_params p;
    int Nx, Ny, Nz, Nxt, Nyt, Nzt;
    int64_t Nxyt, Ntotal;
    int Nc8;
    double ds, mu, diff_flux, mass;
    double *cx, *cy, *cz, *w;
    int32_t *image, *which_corner;
    int nwall, nfluid, ncorner;
    int32_t *vecwall;   /* 1-based local ids */
    double *extcoef;    /* extCoef(nWall,2) column-major */
    int32_t *mapf;      /* 1-based */
    int nst[8];
    int64_t *dir[8];    /* 0-based cell of each stencil entry */
    double *coef[8];    /* coefO(Nstencil,9) stored row per entry: [i*9 + m] */
    double *f, *fw, *rho, *ux, *uy, *uz;
    /* x-direction self-exchange buffers, src/flow.F90:124-127 (corner part empty) */
    int64_t xbuf;
    double *snd_w, *snd_e, *rcv_w, *rcv_e;
};

static inline int64_t CELL(const po_ctx* c, int i, int j, int k)
{
    return (int64_t)(i + 1) + (int64_t)(j + 1) * c->Nxt + (int64_t)(k + 1) * c->Nxyt;
}
static inline double* FP(po_ctx* c, int64_t cell, int l, int o)
{
    return &c->f[((int64_t)o * c->Nc8 + l) * c->Ntotal + cell];
}
static inline double* FWP(po_ctx* c, int corner /*1-based*/, int l, int o)
{
    return &c->fw[((int64_t)o * c->Nc8 + l) * c->ncorner + (corner - 1)];
}

static int has_letter(int label, char letter)
{
    switch (letter) {
    case 'E': return label == L_E || label == L_ES || label == L_EN || label == L_EF || label == L_EB ||
                     label == L_ENF || label == L_ENB || label == L_ESF || label == L_ESB;
    case 'W': return label == L_W || label == L_WS || label == L_WN || label == L_WF || label == L_WB ||
                     label == L_WNF || label == L_WNB || label == L_WSF || label == L_WSB;
    case 'N': return label == L_N || label == L_EN || label == L_WN || label == L_NF || label == L_NB ||
                     label == L_ENF || label == L_ENB || label == L_WNF || label == L_WNB;
    case 'S': return label == L_S || label == L_ES || label == L_WS || label == L_SF || label == L_SB ||
                     label == L_ESF || label == L_ESB || label == L_WSF || label == L_WSB;
    case 'F': return label == L_F || label == L_NF || label == L_SF || label == L_EF || label == L_WF ||
                     label == L_ENF || label == L_WNF || label == L_ESF || label == L_WSF;
    case 'B': return label == L_B || label == L_NB || label == L_SB || label == L_EB || label == L_WB ||
                     label == L_ENB || label == L_WNB || label == L_ESB || label == L_WSB;
    }
    return 0;
}
static int is_wall(int label) { return label >= 20; }
static int is_3corner(int label) { return label >= 60; }

/* neighbour-product -> label, src/physicalGrid.F90:554-667 */
static int label_of_product(int p)
{
    switch (p) {
    case 2: return L_E;    case 3: return L_W;    case 5: return L_N;    case 7: return L_S;
    case 9: return L_F;    case 11: return L_B;
    case 10: return L_EN;  case 15: return L_WN;  case 14: return L_ES;  case 21: return L_WS;
    case 18: return L_EF;  case 22: return L_EB;  case 27: return L_WF;  case 33: return L_WB;
    case 45: return L_NF;  case 55: return L_NB;  case 63: return L_SF;  case 77: return L_SB;
    case 90: return L_ENF; case 110: return L_ENB; case 135: return L_WNF; case 165: return L_WNB;
    case 126: return L_ESF; case 154: return L_ESB; case 189: return L_WSF; case 231: return L_WSB;
    }
    return -1;
}

/* ---------------------------------------------------------------- setup */

/* src/velocityGrid.F90:90-192 */
static int setup_velocity(po_ctx* c, char* err, int errlen)
{
    const double *xi0, *wi0;
    int n = c->p.nc_fundamental;
    if (piks_gh_table(c->p.half_range, n, c->p.exact_quadrature, &xi0, &wi0) != 0 ||
        (!c->p.exact_quadrature && !piks_gh_in_reference(n) && n != 3)) {
        snprintf(err, errlen, "Nc_fundamental xi/wi for %d has not been provided", n);
        return -1;
    }
    double xi[16];
    for (int l = 0; l < n; l++) xi[l] = xi0[l] * sqrt(2.0);
    int nc8 = n * n * n;
    c->Nc8 = nc8;
    c->cx = calloc(nc8, sizeof(double)); c->cy = calloc(nc8, sizeof(double));
    c->cz = calloc(nc8, sizeof(double)); c->w = calloc(nc8, sizeof(double));
    for (int l = 0; l < n; l++)
        for (int m = 0; m < n; m++)
            for (int q = 0; q < n; q++) {
                int k = q + m * n + l * n * n;
                c->cx[k] = xi[q]; c->cy[k] = xi[m]; c->cz[k] = xi[l];
                c->w[k] = wi0[q] * wi0[m] * wi0[l];
            }
    double df = 0.0;
    for (int l = 0; l < nc8; l++) df = df + c->cz[l] * c->w[l];
    c->diff_flux = df * 4;
    return 0;
}

/* one in-place drilling sweep, src/physicalGrid.F90:352-377 (and :427-452) */
static void drill(po_ctx* c, int32_t* A)
{
    int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    for (int k = 1; k <= Nz; k++)
        for (int j = 1; j <= Ny; j++)
            for (int i = 2; i <= Nx - 1; i++) {
                if (A[CELL(c, i, j, k)] != FLUID) continue;
                int thin =
                    (A[CELL(c, i + 1, j, k)] != FLUID && A[CELL(c, i - 1, j, k)] != FLUID) ||
                    (A[CELL(c, i, j + 1, k)] != FLUID && A[CELL(c, i, j - 1, k)] != FLUID) ||
                    (A[CELL(c, i, j, k + 1)] != FLUID && A[CELL(c, i, j, k - 1)] != FLUID);
                if (!thin) continue;
                int j0 = (j == Ny) ? j - 1 : j;
                int k0 = (k == Nz) ? k - 1 : k;
                for (int kk = k0; kk <= k0 + 1; kk++)
                    for (int jj = j0; jj <= j0 + 1; jj++)
                        for (int ii = i; ii <= i + 1; ii++) A[CELL(c, ii, jj, kk)] = FLUID;
            }
}

static void set_ghost_ring(po_ctx* c, int32_t* A, int value)
{
    for (int k = -1; k <= c->Nz + 2; k++)
        for (int j = -1; j <= c->Ny + 2; j++)
            for (int i = -1; i <= c->Nx + 2; i++)
                if (k < 1 || k > c->Nz || j < 1 || j > c->Ny || i < 1 || i > c->Nx)
                    A[CELL(c, i, j, k)] = value;
}

/* in-place removal of one-voxel-thick solid, src/physicalGrid.F90:396-408 (and :471-483) */
static void remove_thin(po_ctx* c, int32_t* A)
{
    for (int k = 1; k <= c->Nz; k++)
        for (int j = 1; j <= c->Ny; j++)
            for (int i = 1; i <= c->Nx; i++) {
                if (A[CELL(c, i, j, k)] != SOLID) continue;
                if ((A[CELL(c, i + 1, j, k)] == FLUID && A[CELL(c, i - 1, j, k)] == FLUID) ||
                    (A[CELL(c, i, j + 1, k)] == FLUID && A[CELL(c, i, j - 1, k)] == FLUID) ||
                    (A[CELL(c, i, j, k + 1)] == FLUID && A[CELL(c, i, j, k - 1)] == FLUID))
                    A[CELL(c, i, j, k)] = FLUID;
            }
}

/* src/physicalGrid.F90:217-1625 for a single rank */
static int setup_grid(po_ctx* c, const uint8_t* raw, int64_t nbytes, char* err, int errlen)
{
    int fl = c->p.fluid_layer;
    int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    c->ds = 1.0 / (Nx - 1 - 2 * fl);
    int64_t need = (int64_t)(Nx - 2 * fl) * Ny * Nz;
    if (nbytes < need) {
        snprintf(err, errlen, "image has %lld bytes, need %lld", (long long)nbytes, (long long)need);
        return -1;
    }
    int32_t* A = calloc(c->Ntotal, sizeof(int32_t)); /* array3Dg */
    /* read, :251-280 (fluid layers stay fluid) */
    int64_t pos = 0;
    for (int k = 1; k <= Nz; k++)
        for (int j = 1; j <= Ny; j++)
            for (int i = fl + 1; i <= Nx - fl; i++) {
                uint8_t v = raw[pos++];
                if (v == 0) A[CELL(c, i, j, k)] = FLUID;
                else if (v == 1) A[CELL(c, i, j, k)] = SOLID;
                else {
                    snprintf(err, errlen, "invalid image data file");
                    free(A);
                    return -1;
                }
            }
    set_ghost_ring(c, A, GHOST);                  /* :303-314 */
    drill(c, A);                                   /* :352-377 */
    set_ghost_ring(c, A, FLUID);                  /* :382-391 */
    remove_thin(c, A);                             /* :396-408 */
    set_ghost_ring(c, A, GHOST);                  /* :413-422 */
    drill(c, A);                                   /* :427-452 */
    set_ghost_ring(c, A, FLUID);                  /* :457-466 */
    remove_thin(c, A);                             /* :471-483 */
    set_ghost_ring(c, A, GHOST);                  /* :488-497 */

    /* local image = array3Dg, ghosts = ghost, :511-543 */
    c->image = malloc(c->Ntotal * sizeof(int32_t));
    memcpy(c->image, A, c->Ntotal * sizeof(int32_t));
    set_ghost_ring(c, c->image, GHOST);

    /* wall labels, :546-671 */
    for (int k = 1; k <= Nz; k++)
        for (int j = 1; j <= Ny; j++)
            for (int i = 1; i <= Nx; i++) {
                if (A[CELL(c, i, j, k)] != SOLID) continue;
                int prod = 1;
                if (A[CELL(c, i + 1, j, k)] == FLUID) prod *= 2;
                if (A[CELL(c, i - 1, j, k)] == FLUID) prod *= 3;
                if (A[CELL(c, i, j + 1, k)] == FLUID) prod *= 5;
                if (A[CELL(c, i, j - 1, k)] == FLUID) prod *= 7;
                if (A[CELL(c, i, j, k + 1)] == FLUID) prod *= 9;
                if (A[CELL(c, i, j, k - 1)] == FLUID) prod *= 11;
                int lab = label_of_product(prod);
                if (lab > 0) c->image[CELL(c, i, j, k)] = lab;
            }
    free(A);

    /* vecWall and which_corner, :701-748 */
    c->which_corner = calloc(c->Ntotal, sizeof(int32_t));
    int nw = 0, ncor = 0;
    int w_snd = 0, e_snd = 0;
    for (int64_t q = 0; q < c->Ntotal; q++) {
        if (is_wall(c->image[q])) nw++;
    }
    c->vecwall = malloc((nw > 0 ? nw : 1) * sizeof(int32_t));
    nw = 0;
    for (int k = -1; k <= Nz + 2; k++)
        for (int j = -1; j <= Ny + 2; j++)
            for (int i = -1; i <= Nx + 2; i++) {
                int64_t q = CELL(c, i, j, k);
                if (!is_wall(c->image[q])) continue;
                c->vecwall[nw++] = (int32_t)(q + 1);
                if (is_3corner(c->image[q])) {
                    c->which_corner[q] = ++ncor;
                    if (i <= 2) w_snd++;            /* i in [xl, xl+1] */
                    else if (i >= Nx - 1) e_snd++;  /* i in [xu-1, xu] */
                }
            }
    c->nwall = nw;
    c->ncorner = ncor;
    /* A 3-fold corner in the first or last two x-layers makes the reference's
     * x self-message longer than the matching receive (src/flow.F90:124-127
     * against the recount at src/physicalGrid.F90:835-885), which MPI aborts. */
    if (w_snd > 0 || e_snd > 0) {
        snprintf(err, errlen,
                 "3-fold corner within two layers of the x boundary (the reference aborts in MPI)");
        return -1;
    }

    /* extCoef, :891-948 (including its index decode, which lands one cell east) */
    c->extcoef = malloc(2 * (nw > 0 ? nw : 1) * sizeof(double));
    for (int l = 0; l < nw; l++) { c->extcoef[l] = 2.0; c->extcoef[nw + l] = -1.0; }
    {
        int xlg = -1, ylg = -1, zlg = -1;
        int xl = 1, yl = 1, zl = 1, xu = Nx, yu = Ny, zu = Nz;
        for (int l = 0; l < nw; l++) {
            int localid = c->vecwall[l];
            int k = zlg + localid / (int)c->Nxyt;
            int j = ylg + (localid - (k - zlg) * (int)c->Nxyt) / c->Nxt;
            int i = xlg + localid - (k - zlg) * (int)c->Nxyt - (j - ylg) * c->Nxt;
            int lab = c->image[CELL(c, i, j, k)];
            int first = 0;
            if (has_letter(lab, 'W') && i == xl + 1 && c->image[CELL(c, xl, j, k)] == FLUID) first = 1;
            if (has_letter(lab, 'E') && i == xu - 1 && c->image[CELL(c, xu, j, k)] == FLUID) first = 1;
            if (has_letter(lab, 'S') && j == yl + 1 && c->image[CELL(c, i, yl, k)] == FLUID) first = 1;
            if (has_letter(lab, 'N') && j == yu - 1 && c->image[CELL(c, i, yu, k)] == FLUID) first = 1;
            if (has_letter(lab, 'B') && k == zl + 1 && c->image[CELL(c, i, j, zl)] == FLUID) first = 1;
            if (has_letter(lab, 'F') && k == zu - 1 && c->image[CELL(c, i, j, zu)] == FLUID) first = 1;
            if (first) { c->extcoef[l] = 1.0; c->extcoef[nw + l] = 0.0; }
        }
        if (c->p.wall_ext_order == 2) {
            for (int l = 0; l < nw; l++) { c->extcoef[l] = 2.0; c->extcoef[nw + l] = -1.0; }
        } else if (c->p.wall_ext_order == 1) {
            for (int l = 0; l < nw; l++) { c->extcoef[l] = 1.0; c->extcoef[nw + l] = 0.0; }
        } else if (c->p.wall_ext_order != 3) {
            fprintf(stderr, " Error: wallExtOrder wroond shoud be [1|2|3]\n");
        }
    }

    /* mapF, :950-974 */
    int nf = 0;
    for (int k = 1; k <= Nz; k++)
        for (int j = 1; j <= Ny; j++)
            for (int i = 1; i <= Nx; i++)
                if (c->image[CELL(c, i, j, k)] == FLUID) nf++;
    c->mapf = malloc((nf > 0 ? nf : 1) * sizeof(int32_t));
    nf = 0;
    for (int k = 1; k <= Nz; k++)
        for (int j = 1; j <= Ny; j++)
            for (int i = 1; i <= Nx; i++)
                if (c->image[CELL(c, i, j, k)] == FLUID) c->mapf[nf++] = (int32_t)(CELL(c, i, j, k) + 1);
    c->nfluid = nf;

    /* sweep lists dir1..8 (:976-1246) and upwind coefficients coef1..8 (:1249-1622) */
    const double c0 = 1.5 / c->ds, c1 = 2.0 / c->ds, c2 = -0.5 / c->ds, d1 = 1.0 / c->ds;
    for (int o = 0; o < 8; o++) {
        int xlo = SX[o] > 0 ? 2 : 1, xhi = SX[o] > 0 ? Nx : Nx - 1;
        int ylo = SY[o] > 0 ? 2 : 1, yhi = SY[o] > 0 ? Ny : Ny - 1;
        int zlo = SZ[o] > 0 ? 2 : 1, zhi = SZ[o] > 0 ? Nz : Nz - 1;
        int cnt = 0;
        for (int k = zlo; k <= zhi; k++)
            for (int j = ylo; j <= yhi; j++)
                for (int i = xlo; i <= xhi; i++)
                    if (c->image[CELL(c, i, j, k)] == FLUID) cnt++;
        c->nst[o] = cnt;
        c->dir[o] = malloc((cnt > 0 ? cnt : 1) * sizeof(int64_t));
        c->coef[o] = malloc((cnt > 0 ? cnt : 1) * 9 * sizeof(double));
        int n = 0;
        for (int kk = 0; kk <= zhi - zlo; kk++) {
            int k = SZ[o] > 0 ? zlo + kk : zhi - kk;
            for (int jj = 0; jj <= yhi - ylo; jj++) {
                int j = SY[o] > 0 ? ylo + jj : yhi - jj;
                for (int ii = 0; ii <= xhi - xlo; ii++) {
                    int i = SX[o] > 0 ? xlo + ii : xhi - ii;
                    if (c->image[CELL(c, i, j, k)] != FLUID) continue;
                    c->dir[o][n] = CELL(c, i, j, k);
                    double* cf = &c->coef[o][(int64_t)n * 9];
                    int sgn[3] = {SX[o], SY[o], SZ[o]};
                    int first[3];
                    first[0] = c->image[CELL(c, i - 2 * SX[o], j, k)] == GHOST ||
                               has_letter(c->image[CELL(c, i - SX[o], j, k)], SX[o] > 0 ? 'E' : 'W');
                    first[1] = c->image[CELL(c, i, j - 2 * SY[o], k)] == GHOST ||
                               has_letter(c->image[CELL(c, i, j - SY[o], k)], SY[o] > 0 ? 'N' : 'S');
                    first[2] = c->image[CELL(c, i, j, k - 2 * SZ[o])] == GHOST ||
                               has_letter(c->image[CELL(c, i, j, k - SZ[o])], SZ[o] > 0 ? 'F' : 'B');
                    for (int a = 0; a < 3; a++) {
                        double v0 = first[a] ? d1 : c0, v1 = first[a] ? d1 : c1, v2 = first[a] ? 0.0 : c2;
                        cf[3 * a + 0] = sgn[a] > 0 ? v0 : -v0;
                        cf[3 * a + 1] = sgn[a] > 0 ? v1 : -v1;
                        cf[3 * a + 2] = sgn[a] > 0 ? v2 : (v2 == 0.0 ? 0.0 : -v2);
                    }
                    n++;
                }
            }
        }
    }
    return 0;
}

po_ctx* po_create(const po_params* p, const uint8_t* raw, int64_t nbytes, char* err, int errlen)
{
    char dummy[8];
    if (!err) { err = dummy; errlen = sizeof dummy; }
    err[0] = 0;
    po_ctx* c = calloc(1, sizeof(po_ctx));
    c->p = *p;
    /* src/parameters.F90:121 */
    c->Nx = p->nx + 2 * p->fluid_layer;
    c->Ny = p->ny;
    c->Nz = p->nz;
    c->Nxt = c->Nx + 4; c->Nyt = c->Ny + 4; c->Nzt = c->Nz + 4;
    c->Nxyt = (int64_t)c->Nxt * c->Nyt;
    c->Ntotal = c->Nxyt * c->Nzt;
    if (setup_velocity(c, err, errlen) != 0 || setup_grid(c, raw, nbytes, err, errlen) != 0) {
        po_destroy(c);
        return NULL;
    }
    /* src/flow.F90:85-112 */
    int64_t nf = c->Ntotal * c->Nc8 * 8;
    c->f = calloc(nf, sizeof(double));
    c->fw = calloc((int64_t)(c->ncorner > 0 ? c->ncorner : 1) * c->Nc8 * 8, sizeof(double));
    c->rho = calloc(c->Ntotal, sizeof(double));
    c->ux = calloc(c->Ntotal, sizeof(double));
    c->uy = calloc(c->Ntotal, sizeof(double));
    c->uz = calloc(c->Ntotal, sizeof(double));
    c->mu = sqrt((atan(1.0) * 4.0) / 2.0) / p->kn;
    c->mass = 1.0;
    c->xbuf = (int64_t)c->Nyt * c->Nzt * 2 * 8 * c->Nc8;
    c->snd_w = calloc(c->xbuf, sizeof(double));
    c->snd_e = calloc(c->xbuf, sizeof(double));
    c->rcv_w = calloc(c->xbuf, sizeof(double));
    c->rcv_e = calloc(c->xbuf, sizeof(double));
    return c;
}

void po_destroy(po_ctx* c)
{
    if (!c) return;
    free(c->cx); free(c->cy); free(c->cz); free(c->w);
    free(c->image); free(c->which_corner); free(c->vecwall); free(c->extcoef); free(c->mapf);
    for (int o = 0; o < 8; o++) { free(c->dir[o]); free(c->coef[o]); }
    free(c->f); free(c->fw); free(c->rho); free(c->ux); free(c->uy); free(c->uz);
    free(c->snd_w); free(c->snd_e); free(c->rcv_w); free(c->rcv_e);
    free(c);
}

/* ---------------------------------------------------------------- sweep */

/* In-sweep wall switches, src/solver.F90:184-204 and analogues.  src > 0 is a
 * slot number, src == 0 means fw(which_corner, l, o). */
typedef struct { int label, src; } sw_t;
static const sw_t YSW[8][3] = {
    {{L_EN, 3}, {L_ENF, 0}, {L_ENB, 0}}, {{L_WN, 4}, {L_WNF, 0}, {L_WNB, 0}},
    {{L_WS, 1}, {L_WSF, 0}, {L_WSB, 0}}, {{L_ES, 2}, {L_ESF, 0}, {L_ESB, 0}},
    {{L_EN, 7}, {L_ENF, 0}, {L_ENB, 0}}, {{L_WN, 8}, {L_WNF, 0}, {L_WNB, 0}},
    {{L_WS, 8}, {L_WSF, 0}, {L_WSB, 0}}, {{L_ES, 6}, {L_ESF, 0}, {L_ESB, 0}},
};
static const sw_t ZSW[8][5] = {
    {{L_NF, 8}, {L_EF, 6}, {L_WNF, 0}, {L_ESF, 0}, {L_ENF, 7}},
    {{L_NF, 7}, {L_WF, 5}, {L_ENF, 0}, {L_WSF, 0}, {L_WNF, 8}},
    {{L_SF, 6}, {L_WF, 8}, {L_WNF, 0}, {L_ESF, 0}, {L_WSF, 5}},
    {{L_SF, 5}, {L_EF, 7}, {L_ENF, 0}, {L_WSF, 0}, {L_ESF, 6}},
    {{L_NB, 4}, {L_EB, 2}, {L_ESB, 0}, {L_WNB, 0}, {L_ENB, 3}},
    {{L_NB, 3}, {L_WB, 1}, {L_ENB, 0}, {L_WSB, 0}, {L_WNB, 4}},
    {{L_SB, 2}, {L_WB, 4}, {L_WNB, 0}, {L_ESB, 0}, {L_WSB, 1}},
    {{L_SB, 1}, {L_EB, 3}, {L_ENB, 0}, {L_WSB, 0}, {L_ESB, 2}},
};

static void apply_switch(po_ctx* c, const sw_t* tab, int ntab, int64_t wcell, int l, int o)
{
    int lab = c->image[wcell];
    for (int t = 0; t < ntab; t++) {
        if (tab[t].label != lab) continue;
        double v = tab[t].src > 0 ? *FP(c, wcell, l, tab[t].src - 1)
                                  : *FWP(c, c->which_corner[wcell], l, o);
        *FP(c, wcell, l, o) = v;
        return;
    }
}

/* one octant sweep, src/solver.F90:176-216 (octant 1) and the seven analogues */
static void sweep_octant(po_ctx* c, int o, int nthreads)
{
    const int sx = SX[o], sy = SY[o], sz = SZ[o];
    const int64_t dx = sx, dy = (int64_t)sy * c->Nxt, dz = (int64_t)sz * c->Nxyt;
    const double mu = c->mu;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int l = 0; l < c->Nc8; l++) {
        const double cxl = c->cx[l], cyl = c->cy[l], czl = c->cz[l], wl = c->w[l];
        for (int n = 0; n < c->nst[o]; n++) {
            const int64_t k = c->dir[o][n];
            const double* cf = &c->coef[o][(int64_t)n * 9];
            apply_switch(c, YSW[o], 3, k - dy, l, o);
            apply_switch(c, ZSW[o], 5, k - dz, l, o);
            double t = sx > 0 ? cxl * c->ux[k] : -(cxl * c->ux[k]);
            t = sy > 0 ? t + cyl * c->uy[k] : t - cyl * c->uy[k];
            t = sz > 0 ? t + czl * c->uz[k] : t - czl * c->uz[k];
            const double feq = wl * (c->rho[k] + 1.0 * t);
            double* fk = FP(c, k, l, o);
            double num = mu * (feq - 0.5 * (*fk));
            double den = 0.5 * mu;
            const double fx1 = *FP(c, k - dx, l, o), fx2 = *FP(c, k - 2 * dx, l, o);
            const double fy1 = *FP(c, k - dy, l, o), fy2 = *FP(c, k - 2 * dy, l, o);
            const double fz1 = *FP(c, k - dz, l, o), fz2 = *FP(c, k - 2 * dz, l, o);
            if (sx > 0) { num = num + cxl * cf[1] * fx1; num = num + cxl * cf[2] * fx2; }
            else        { num = num - cxl * cf[1] * fx1; num = num - cxl * cf[2] * fx2; }
            if (sy > 0) { num = num + cyl * cf[4] * fy1; num = num + cyl * cf[5] * fy2; }
            else        { num = num - cyl * cf[4] * fy1; num = num - cyl * cf[5] * fy2; }
            if (sz > 0) { num = num + czl * cf[7] * fz1; num = num + czl * cf[8] * fz2; }
            else        { num = num - czl * cf[7] * fz1; num = num - czl * cf[8] * fz2; }
            den = sx > 0 ? den + cxl * cf[0] : den - cxl * cf[0];
            den = sy > 0 ? den + cyl * cf[3] : den - cyl * cf[3];
            den = sz > 0 ? den + czl * cf[6] : den - czl * cf[6];
            *fk = num / den;
        }
    }
}

/* ---------------------------------------------------------------- wall BC */

/* One wall label's boundary program, src/solver.F90:772-1823.
 * Value slots: 0..7 = f(W, l, 1..8), 8..15 = fw(corner, l, 1..8), 16..23 = fwZ(l, 1..8).
 * face: axis 0/1/2, dir (+1 fluid side towards +axis), store base (0, 8 or 16),
 *       incoming octants in the order the reference sums them.
 * write: dst slot <- accom*w*R[rho] + (1-accom)*src slot, in this order. */
typedef struct { int axis, dir, store, in[4]; } face_t;
typedef struct { int dst, rho, src; } wr_t;
typedef struct { int label, nface; face_t face[3]; int nwr; wr_t wr[12]; } wallprog_t;

#define F_(n) ((n) - 1)
#define W_(n) (8 + (n) - 1)
#define Z_(n) (16 + (n) - 1)
#define FX_P {0, +1, 0, {2, 3, 6, 7}}
#define FX_M {0, -1, 0, {1, 4, 5, 8}}
#define WX_P0 {F_(1), 0, F_(2)}, {F_(4), 0, F_(3)}, {F_(5), 0, F_(6)}, {F_(8), 0, F_(7)}
#define WX_M0 {F_(2), 0, F_(1)}, {F_(3), 0, F_(4)}, {F_(6), 0, F_(5)}, {F_(7), 0, F_(8)}

static const wallprog_t WALLS[26] = {
    {L_E, 1, {FX_P}, 4, {WX_P0}},
    {L_W, 1, {FX_M}, 4, {WX_M0}},
    {L_N, 1, {{1, +1, 0, {3, 4, 7, 8}}}, 4, {{F_(1), 0, F_(4)}, {F_(2), 0, F_(3)}, {F_(5), 0, F_(8)}, {F_(6), 0, F_(7)}}},
    {L_S, 1, {{1, -1, 0, {1, 2, 5, 6}}}, 4, {{F_(3), 0, F_(2)}, {F_(4), 0, F_(1)}, {F_(7), 0, F_(6)}, {F_(8), 0, F_(5)}}},
    {L_F, 1, {{2, +1, 0, {5, 6, 7, 8}}}, 4, {{F_(1), 0, F_(5)}, {F_(2), 0, F_(6)}, {F_(3), 0, F_(7)}, {F_(4), 0, F_(8)}}},
    {L_B, 1, {{2, -1, 0, {1, 2, 3, 4}}}, 4, {{F_(5), 0, F_(1)}, {F_(6), 0, F_(2)}, {F_(7), 0, F_(3)}, {F_(8), 0, F_(4)}}},
    {L_EN, 2, {FX_P, {1, +1, 16, {3, 4, 7, 8}}}, 8,
     {WX_P0, {F_(2), 1, Z_(3)}, {F_(6), 1, Z_(7)}, {F_(3), 1, Z_(4)}, {F_(7), 1, Z_(8)}}},
    {L_WN, 2, {FX_M, {1, +1, 16, {3, 4, 7, 8}}}, 8,
     {WX_M0, {F_(1), 1, Z_(4)}, {F_(5), 1, Z_(8)}, {F_(4), 1, Z_(3)}, {F_(8), 1, Z_(7)}}},
    {L_ES, 2, {FX_P, {1, -1, 16, {1, 2, 5, 6}}}, 8,
     {WX_P0, {F_(3), 1, Z_(2)}, {F_(7), 1, Z_(6)}, {F_(2), 1, Z_(1)}, {F_(6), 1, Z_(5)}}},
    {L_WS, 2, {FX_M, {1, -1, 16, {1, 2, 5, 6}}}, 8,
     {WX_M0, {F_(4), 1, Z_(1)}, {F_(8), 1, Z_(5)}, {F_(1), 1, Z_(2)}, {F_(5), 1, Z_(6)}}},
    {L_NF, 2, {{1, +1, 0, {3, 4, 7, 8}}, {2, +1, 16, {5, 6, 7, 8}}}, 8,
     {{F_(1), 0, F_(4)}, {F_(2), 0, F_(3)}, {F_(5), 0, F_(8)}, {F_(6), 0, F_(7)},
      {F_(3), 1, Z_(7)}, {F_(4), 1, Z_(8)}, {F_(7), 1, Z_(6)}, {F_(8), 1, Z_(5)}}},
    {L_NB, 2, {{1, +1, 0, {3, 4, 7, 8}}, {2, -1, 16, {1, 2, 3, 4}}}, 8,
     {{F_(1), 0, F_(4)}, {F_(2), 0, F_(3)}, {F_(5), 0, F_(8)}, {F_(6), 0, F_(7)},
      {F_(7), 1, Z_(3)}, {F_(8), 1, Z_(4)}, {F_(4), 1, Z_(1)}, {F_(3), 1, Z_(2)}}},
    {L_SF, 2, {{1, -1, 0, {1, 2, 5, 6}}, {2, +1, 16, {5, 6, 7, 8}}}, 8,
     {{F_(3), 0, F_(2)}, {F_(4), 0, F_(1)}, {F_(7), 0, F_(6)}, {F_(8), 0, F_(5)},
      {F_(1), 1, Z_(5)}, {F_(2), 1, Z_(6)}, {F_(6), 1, Z_(7)}, {F_(5), 1, Z_(8)}}},
    {L_SB, 2, {{1, -1, 0, {1, 2, 5, 6}}, {2, -1, 16, {1, 2, 3, 4}}}, 8,
     {{F_(3), 0, F_(2)}, {F_(4), 0, F_(1)}, {F_(7), 0, F_(6)}, {F_(8), 0, F_(5)},
      {F_(5), 1, Z_(1)}, {F_(6), 1, Z_(2)}, {F_(2), 1, Z_(3)}, {F_(1), 1, Z_(4)}}},
    {L_EF, 2, {FX_P, {2, +1, 16, {5, 6, 7, 8}}}, 8,
     {WX_P0, {F_(2), 1, Z_(6)}, {F_(3), 1, Z_(7)}, {F_(6), 1, Z_(5)}, {F_(7), 1, Z_(8)}}},
    {L_EB, 2, {FX_P, {2, -1, 16, {1, 2, 3, 4}}}, 8,
     {WX_P0, {F_(6), 1, Z_(2)}, {F_(7), 1, Z_(3)}, {F_(2), 1, Z_(1)}, {F_(3), 1, Z_(4)}}},
    {L_WF, 2, {FX_M, {2, +1, 16, {5, 6, 7, 8}}}, 8,
     {WX_M0, {F_(1), 1, Z_(5)}, {F_(4), 1, Z_(8)}, {F_(5), 1, Z_(6)}, {F_(8), 1, Z_(7)}}},
    {L_WB, 2, {FX_M, {2, -1, 16, {1, 2, 3, 4}}}, 8,
     {WX_M0, {F_(5), 1, Z_(1)}, {F_(8), 1, Z_(4)}, {F_(1), 1, Z_(2)}, {F_(4), 1, Z_(3)}}},
    {L_ENF, 3, {FX_P, {1, +1, 8, {3, 4, 7, 8}}, {2, +1, 16, {5, 6, 7, 8}}}, 12,
     {WX_P0, {F_(2), 1, W_(3)}, {F_(6), 1, W_(7)}, {F_(3), 2, Z_(7)},
      {W_(1), 1, W_(4)}, {W_(5), 1, W_(8)}, {W_(2), 2, Z_(6)}, {W_(4), 2, Z_(8)}, {F_(7), 2, Z_(5)}}},
    {L_WNF, 3, {FX_M, {1, +1, 8, {3, 4, 7, 8}}, {2, +1, 16, {5, 6, 7, 8}}}, 12,
     {WX_M0, {F_(1), 1, W_(4)}, {F_(5), 1, W_(8)}, {F_(4), 2, Z_(8)},
      {W_(2), 1, W_(3)}, {W_(6), 1, W_(7)}, {W_(1), 2, Z_(5)}, {W_(3), 2, Z_(7)}, {F_(8), 2, Z_(6)}}},
    {L_WSF, 3, {FX_M, {1, -1, 8, {1, 2, 5, 6}}, {2, +1, 16, {5, 6, 7, 8}}}, 12,
     {WX_M0, {F_(4), 1, W_(1)}, {F_(8), 1, W_(5)}, {F_(1), 2, Z_(5)},
      {W_(3), 1, W_(2)}, {W_(7), 1, W_(6)}, {W_(2), 2, Z_(6)}, {W_(4), 2, Z_(8)}, {F_(5), 2, Z_(7)}}},
    {L_ESF, 3, {FX_P, {1, -1, 8, {1, 2, 5, 6}}, {2, +1, 16, {5, 6, 7, 8}}}, 12,
     {WX_P0, {F_(3), 1, W_(2)}, {F_(7), 1, W_(6)}, {F_(2), 2, Z_(6)},
      {W_(4), 1, W_(1)}, {W_(8), 1, W_(5)}, {W_(1), 2, Z_(5)}, {W_(3), 2, Z_(7)}, {F_(6), 2, Z_(8)}}},
    {L_ENB, 3, {FX_P, {1, +1, 8, {3, 4, 7, 8}}, {2, -1, 16, {1, 2, 3, 4}}}, 12,
     {WX_P0, {F_(2), 1, W_(3)}, {F_(6), 1, W_(7)}, {F_(7), 2, Z_(3)},
      {W_(1), 1, W_(4)}, {W_(5), 1, W_(8)}, {W_(6), 2, Z_(2)}, {W_(8), 2, Z_(4)}, {F_(3), 2, Z_(1)}}},
    {L_WNB, 3, {FX_M, {1, +1, 8, {3, 4, 7, 8}}, {2, -1, 16, {1, 2, 3, 4}}}, 12,
     {WX_M0, {F_(1), 1, W_(4)}, {F_(5), 1, W_(8)}, {F_(8), 2, Z_(4)},
      {W_(2), 1, W_(3)}, {W_(6), 1, W_(7)}, {W_(5), 2, Z_(1)}, {W_(7), 2, Z_(3)}, {F_(4), 2, Z_(2)}}},
    {L_WSB, 3, {FX_M, {1, -1, 8, {1, 2, 5, 6}}, {2, -1, 16, {1, 2, 3, 4}}}, 12,
     {WX_M0, {F_(4), 1, W_(1)}, {F_(8), 1, W_(5)}, {F_(5), 2, Z_(1)},
      {W_(3), 1, W_(2)}, {W_(7), 1, W_(6)}, {W_(6), 2, Z_(2)}, {W_(8), 2, Z_(4)}, {F_(1), 2, Z_(3)}}},
    {L_ESB, 3, {FX_P, {1, -1, 8, {1, 2, 5, 6}}, {2, -1, 16, {1, 2, 3, 4}}}, 12,
     {WX_P0, {F_(3), 1, W_(2)}, {F_(7), 1, W_(6)}, {F_(6), 2, Z_(2)},
      {W_(4), 1, W_(1)}, {W_(8), 1, W_(5)}, {W_(5), 2, Z_(1)}, {W_(7), 2, Z_(3)}, {F_(2), 2, Z_(4)}}},
};

static const wallprog_t* wallprog(int label)
{
    for (int t = 0; t < 26; t++)
        if (WALLS[t].label == label) return &WALLS[t];
    return NULL;
}

/* value slot -> storage location for wall cell `cell`, velocity l */
static double* slot_ptr(po_ctx* c, int64_t cell, int corner, double* fwz, int l, int s)
{
    if (s < 8) return FP(c, cell, l, s);
    if (s < 16) return FWP(c, corner, l, s - 8);
    return &fwz[(int64_t)(s - 16) * c->Nc8 + l];
}

/* src/solver.F90:761-1826, walls in vecWall order */
static void wall_bc(po_ctx* c, double* fwz)
{
    const double accom = c->p.accom;
    const double* cvel[3] = {c->cx, c->cy, c->cz};
    for (int iw = 0; iw < c->nwall; iw++) {
        const int64_t cell = (int64_t)c->vecwall[iw] - 1;
        const int corner = c->which_corner[cell];
        const wallprog_t* P = wallprog(c->image[cell]);
        const double e1 = c->extcoef[iw], e2 = c->extcoef[c->nwall + iw];
        double R[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < c->Nc8; l++) {
            for (int a = 0; a < P->nface; a++) {
                const face_t* fc = &P->face[a];
                int64_t step = fc->axis == 0 ? 1 : (fc->axis == 1 ? (int64_t)c->Nxt : c->Nxyt);
                step *= fc->dir;
                double v[4];
                for (int q = 0; q < 4; q++) {
                    int o = fc->in[q] - 1;
                    double val = e1 * (*FP(c, cell + step, l, o)) + e2 * (*FP(c, cell + 2 * step, l, o));
                    *slot_ptr(c, cell, corner, fwz, l, fc->store + o) = val;
                    v[q] = val;
                }
                R[a] = R[a] + cvel[fc->axis][l] * (v[0] + v[1] + v[2] + v[3]);
            }
        }
        for (int a = 0; a < P->nface; a++) R[a] = R[a] / c->diff_flux;
        for (int l = 0; l < c->Nc8; l++) {
            for (int t = 0; t < P->nwr; t++) {
                const wr_t* wr = &P->wr[t];
                double src = *slot_ptr(c, cell, corner, fwz, l, wr->src);
                *slot_ptr(c, cell, corner, fwz, l, wr->dst) =
                    accom * c->w[l] * R[wr->rho] + (1.0 - accom) * src;
            }
        }
    }
}

/* ---------------------------------------------------------------- iterate */

void po_iterate(po_ctx* c, int nsteps, int nthreads)
{
    if (nthreads <= 0) nthreads = 1;
    const int Nc8 = c->Nc8;
    double* fwz = malloc((size_t)Nc8 * 8 * sizeof(double));
    for (int step = 0; step < nsteps; step++) {
        /* self send/receive posted at the top of iterate carries the buffers
         * packed by the previous call (src/solver.F90:146-170, 530-532) */
        memcpy(c->rcv_e, c->snd_w, c->xbuf * sizeof(double));
        memcpy(c->rcv_w, c->snd_e, c->xbuf * sizeof(double));

        for (int o = 0; o < 8; o++) sweep_octant(c, o, nthreads);

        /* pack/unpack x, src/solver.F90:537-559 */
        for (int k = 0; k < c->Nzt; k++)
            for (int j = 0; j < c->Nyt; j++)
                for (int i = 0; i < 2; i++) {
                    int64_t row = (int64_t)k * c->Nxyt + (int64_t)j * c->Nxt;
                    int64_t msnd = row + i + 2, mrcv = row + i;
                    int64_t psnd = row + i + c->Nx, prcv = row + i + 2 + c->Nx;
                    int64_t packid = (((int64_t)k * c->Nyt + j) * 2 + i) * 8 * Nc8;
                    for (int o = 0; o < 8; o++)
                        for (int v = 0; v < Nc8; v++) {
                            int64_t b = packid + (int64_t)o * Nc8 + v;
                            c->snd_w[b] = *FP(c, msnd, v, o);
                            *FP(c, mrcv, v, o) = c->rcv_w[b];
                            c->snd_e[b] = *FP(c, psnd, v, o);
                            *FP(c, prcv, v, o) = c->rcv_e[b];
                        }
                }
        /* unpack y and z from MPI_PROC_NULL neighbours: zeros, :564-612 */
        for (int k = 0; k < c->Nzt; k++)
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < c->Nxt; i++) {
                    int64_t a = (int64_t)k * c->Nxyt + (int64_t)j * c->Nxt + i;
                    int64_t b = (int64_t)k * c->Nxyt + (int64_t)(j + c->Ny + 2) * c->Nxt + i;
                    for (int o = 0; o < 8; o++)
                        for (int v = 0; v < Nc8; v++) { *FP(c, a, v, o) = 0.0; *FP(c, b, v, o) = 0.0; }
                }
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < c->Nyt; j++)
                for (int i = 0; i < c->Nxt; i++) {
                    int64_t a = (int64_t)k * c->Nxyt + (int64_t)j * c->Nxt + i;
                    int64_t b = (int64_t)(k + c->Nz + 2) * c->Nxyt + (int64_t)j * c->Nxt + i;
                    for (int o = 0; o < 8; o++)
                        for (int v = 0; v < Nc8; v++) { *FP(c, a, v, o) = 0.0; *FP(c, b, v, o) = 0.0; }
                }

        wall_bc(c, fwz);

        /* inlet/outlet, src/solver.F90:1832-1862 */
        const double pd = c->p.press_drop;
        for (int k = -1; k <= c->Nz + 2; k++)
            for (int j = -1; j <= c->Ny + 2; j++) {
                int64_t in = CELL(c, 1, j, k), out = CELL(c, c->Nx, j, k);
                static const int INO[4] = {0, 3, 4, 7}, OUTO[4] = {1, 2, 5, 6};
                for (int t = 0; t < 4; t++)
                    for (int v = 0; v < Nc8; v++)
                        *FP(c, in, v, INO[t]) = *FP(c, in - 1, v, INO[t]) + c->w[v] * pd;
                for (int t = 0; t < 4; t++)
                    for (int v = 0; v < Nc8; v++)
                        *FP(c, out, v, OUTO[t]) = *FP(c, out + 1, v, OUTO[t]) - c->w[v] * pd;
            }

        /* symmetry planes S, N, B, F, src/solver.F90:1867-1918 */
        for (int k = 1; k <= c->Nz; k++)
            for (int i = 1; i <= c->Nx; i++) {
                int64_t q = CELL(c, i, 1, k);
                for (int v = 0; v < Nc8; v++) {
                    *FP(c, q, v, 1) = *FP(c, q, v, 2); *FP(c, q, v, 0) = *FP(c, q, v, 3);
                    *FP(c, q, v, 5) = *FP(c, q, v, 6); *FP(c, q, v, 4) = *FP(c, q, v, 7);
                }
            }
        for (int k = 1; k <= c->Nz; k++)
            for (int i = 1; i <= c->Nx; i++) {
                int64_t q = CELL(c, i, c->Ny, k);
                for (int v = 0; v < Nc8; v++) {
                    *FP(c, q, v, 2) = *FP(c, q, v, 1); *FP(c, q, v, 3) = *FP(c, q, v, 0);
                    *FP(c, q, v, 6) = *FP(c, q, v, 5); *FP(c, q, v, 7) = *FP(c, q, v, 4);
                }
            }
        for (int j = 1; j <= c->Ny; j++)
            for (int i = 1; i <= c->Nx; i++) {
                int64_t q = CELL(c, i, j, 1);
                for (int v = 0; v < Nc8; v++) {
                    *FP(c, q, v, 0) = *FP(c, q, v, 4); *FP(c, q, v, 1) = *FP(c, q, v, 5);
                    *FP(c, q, v, 2) = *FP(c, q, v, 6); *FP(c, q, v, 3) = *FP(c, q, v, 7);
                }
            }
        for (int j = 1; j <= c->Ny; j++)
            for (int i = 1; i <= c->Nx; i++) {
                int64_t q = CELL(c, i, j, c->Nz);
                for (int v = 0; v < Nc8; v++) {
                    *FP(c, q, v, 4) = *FP(c, q, v, 0); *FP(c, q, v, 5) = *FP(c, q, v, 1);
                    *FP(c, q, v, 6) = *FP(c, q, v, 2); *FP(c, q, v, 7) = *FP(c, q, v, 3);
                }
            }

        /* moments over mapF, src/solver.F90:1925-1937 */
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int n = 0; n < c->nfluid; n++) {
            const int64_t k = (int64_t)c->mapf[n] - 1;
            double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
            for (int l = 0; l < Nc8; l++) {
                const double f1 = *FP(c, k, l, 0), f2 = *FP(c, k, l, 1), f3 = *FP(c, k, l, 2),
                             f4 = *FP(c, k, l, 3), f5 = *FP(c, k, l, 4), f6 = *FP(c, k, l, 5),
                             f7 = *FP(c, k, l, 6), f8 = *FP(c, k, l, 7);
                r = r + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
                x = x + c->cx[l] * (f1 - f2 - f3 + f4 + f5 - f6 - f7 + f8);
                y = y + c->cy[l] * (f1 + f2 - f3 - f4 + f5 + f6 - f7 - f8);
                z = z + c->cz[l] * (f1 + f2 + f3 + f4 - f5 - f6 - f7 - f8);
            }
            c->rho[k] = r; c->ux[k] = x; c->uy[k] = y; c->uz[k] = z;
        }
    }
    free(fwz);
}

/* src/output.F90:60-164 for one rank */
void po_chk_converge(po_ctx* c, double out[3])
{
    const int Ny = c->Ny, Nz = c->Nz, col = 2; /* column = 2, src/physicalGrid.F90:120 */
    const double ds = c->ds;
    double inner = 0.0, noth = 0.0, suth = 0.0, frnt = 0.0, back = 0.0;
    for (int k = 2; k <= Nz - 1; k++)
        for (int j = 2; j <= Ny - 1; j++) inner = inner + c->ux[CELL(c, col, j, k)] * ds * ds;
    for (int k = 2; k <= Nz - 1; k++) suth = suth + c->ux[CELL(c, col, 1, k)] * ds * ds * 0.5;
    for (int k = 2; k <= Nz - 1; k++) noth = noth + c->ux[CELL(c, col, Ny, k)] * ds * ds * 0.5;
    for (int j = 2; j <= Ny - 1; j++) back = back + c->ux[CELL(c, col, j, 1)] * ds * ds * 0.5;
    for (int j = 2; j <= Ny - 1; j++) frnt = frnt + c->ux[CELL(c, col, j, Nz)] * ds * ds * 0.5;
    const double sb = c->ux[CELL(c, col, 1, 1)] * ds * ds * 0.25;
    const double sf = c->ux[CELL(c, col, 1, Nz)] * ds * ds * 0.25;
    const double nb = c->ux[CELL(c, col, Ny, 1)] * ds * ds * 0.25;
    const double nf = c->ux[CELL(c, col, Ny, Nz)] * ds * ds * 0.25;
    const double rl = 1.0 / c->p.ref_l;
    const double m = (inner + suth + noth + frnt + back + sb + sf + nb + nf) * sqrt(1.0 / 2.0) /
                     c->p.press_drop / (rl * rl);
    const double error = fabs(1.0 - m / c->mass) / (double)c->p.chk_converge_step;
    c->mass = m;
    const double pi = atan(1.0) * 4.0;
    out[0] = m;
    out[1] = m * c->p.kn * sqrt(4.0 / pi);
    out[2] = error;
}

void po_info(const po_ctx* c, int64_t out[19])
{
    out[0] = c->Nx; out[1] = c->Ny; out[2] = c->Nz;
    out[3] = c->Nxt; out[4] = c->Nyt; out[5] = c->Nzt; out[6] = c->Ntotal;
    out[7] = c->Nc8; out[8] = c->nwall; out[9] = c->nfluid; out[10] = c->ncorner;
    for (int o = 0; o < 8; o++) out[11 + o] = c->nst[o];
}
void po_info_d(const po_ctx* c, double out[4])
{
    out[0] = c->ds; out[1] = c->mu; out[2] = c->diff_flux; out[3] = c->mass;
}
double* po_f(po_ctx* c) { return c->f; }
double* po_fw(po_ctx* c) { return c->fw; }
double* po_rho(po_ctx* c) { return c->rho; }
double* po_ux(po_ctx* c) { return c->ux; }
double* po_uy(po_ctx* c) { return c->uy; }
double* po_uz(po_ctx* c) { return c->uz; }
int32_t* po_image(po_ctx* c) { return c->image; }
int32_t* po_which_corner(po_ctx* c) { return c->which_corner; }
int32_t* po_vecwall(po_ctx* c) { return c->vecwall; }
double* po_extcoef(po_ctx* c) { return c->extcoef; }
int32_t* po_mapf(po_ctx* c) { return c->mapf; }
double* po_cx(po_ctx* c) { return c->cx; }
double* po_cy(po_ctx* c) { return c->cy; }
double* po_cz(po_ctx* c) { return c->cz; }
double* po_w(po_ctx* c) { return c->w; }
