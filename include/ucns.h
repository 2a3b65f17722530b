/*
 * ucns.h -- C-ABI of the MI355X residual + SSP-RK3 library (libucnsgpu.so).
 *
 * This is the drop-in boundary for UCNS3D's explicit spatial residual and
 * time update.  The reference has no FFI layer: the seams replaced here are
 * plain Fortran calls on module state, namely
 *   - the residual body of CALL_FLUX_SUBROUTINES_3D   (src/time_p.f90:2253-2497)
 *     = ARBITRARY_ORDER (src/reconstruct_p.f90:2954-3015)
 *     + CALCULATE_FLUXESHI_CONVECTIVE (src/flux_p.f90:478-1030)
 *   - the three stage loops of RUNGE_KUTTA3             (src/time_p.f90:885-1027)
 *   - the dt loop of CALCULATE_CFL + its MPI min        (src/time_p.f90:27-191, 4131)
 * A Fortran host binds these with ISO_C_BINDING; Python
 * binds them with ctypes (ucns3d_amd/native.py).
 *
 * Conventions
 *   - every array argument is caller-owned and only read/written during the
 *     call; the library copies what it needs into device memory it owns;
 *   - per-face arrays are [ncell][nface_max] row-major (face fastest);
 *   - index arrays carry an explicit index_base (1 for Fortran arrays);
 *   - matrices keep the reference's column-major element order, e.g.
 *     INVMAT_STENCILT(ideg,imax,s) is stored with ideg fastest;
 *   - every call returns 0 or a negative UCNS_E* code; ucns_last_error()
 *     gives the message.  Nothing throws across the ABI, nothing exits;
 *   - a context is not re-entrant: call it from one host thread (the
 *     reference's !$OMP MASTER discipline, src/communications.f90:1137-1242).
 */
#ifndef UCNS_H
#define UCNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UCNS_ABI_VERSION 2

/* return codes */
#define UCNS_OK             0
#define UCNS_EBADARG       -1
#define UCNS_EUNSUPPORTED  -2
#define UCNS_ENOMEM        -3
#define UCNS_EHIP          -4
#define UCNS_ENONFINITE    -5
#define UCNS_ESTATE        -6

/* precision modes */
#define UCNS_FP64   0   /* everything in FP64 */
#define UCNS_MIXED  1   /* reconstruction, WENO weights, HLLC in FP32; dU
                           formation, residual accumulation, RK update and
                           conserved->primitive recovery in FP64 */

/* state layouts accepted by set/get */
#define UCNS_LAYOUT_AOS  0   /* u[i*5+v]     = U_C(i)%VAL(1,v)   */
#define UCNS_LAYOUT_SOA  1   /* u[v*ncell+i] = U_C(i)%VAL(1,v)   */

/* cell order inside the library (ucns_scheme_v2.cell_order).  Every array
 * crossing the ABI stays in the caller's order; the library permutes on the
 * way in and out. */
#define UCNS_ORDER_CALLER  0   /* the caller's order as given */
#define UCNS_ORDER_RCM     1   /* reverse Cuthill-McKee on the face graph */
#define UCNS_ORDER_MORTON  2   /* Morton (Z-order) curve of the cell centroids (needs ucns_mesh_v2.centroid) */

/* reconstruction kernel family (ucns_scheme_v2.kernels) */
#define UCNS_KERNELS_AUTO      0   /* row-stream kernels where the stencil shape allows */
#define UCNS_KERNELS_REGISTER  1   /* register-prefetch kernels only */

/* Mesh topology and geometry, one partition (the whole mesh).
 * Mirrors IELEM(N,i)%{IFCA,INEIGH,INEIGHN,IBOUNDS->ICODE,TYPES_FACES,
 * FACEANGLEX,FACEANGLEY,SURF,TOTVOLUME,MINEDGE,XXC/YYC/ZZC}
 * (src/declarations.f90:644-712). */
typedef struct ucns_mesh_v2 {
    int32_t abi_version;        /* UCNS_ABI_VERSION */
    int32_t ncell;              /* KMAXE */
    int32_t index_base;         /* 0 or 1 */
    int32_t nface_max;          /* leading dimension of the per-face arrays */
    const int32_t *ifca;        /* [ncell] faces per cell (4 tet, 5 prism/pyramid, 6 hex) */
    const int32_t *ineigh;      /* [ncell*nface_max] cell across face L (interior or periodic; ignored on boundary faces) */
    const int32_t *ineighn;     /* [ncell*nface_max] face slot of L inside that neighbour */
    const int32_t *bcode;       /* [ncell*nface_max] 0 interior, 5 or 50 periodic, 1/2/3/4/6 boundary (ucns_boundary_v1) */
    const int32_t *types_faces; /* [ncell*nface_max] 5 quadrilateral, 6 triangle */
    const double *faceanglex;   /* [ncell*nface_max] outward normal angle a1 */
    const double *faceangley;   /* [ncell*nface_max] outward normal angle a2 */
    const double *surf;         /* [ncell*nface_max] face area */
    const double *totvolume;    /* [ncell] */
    const double *minedge;      /* [ncell] 3V/sum(S), src/grid_t.f90:1122 */
    const double *centroid;     /* [ncell][3] cell centroids IELEM%XXC,YYC,ZZC; may be NULL
                                   unless cell_order is UCNS_ORDER_MORTON */
} ucns_mesh_v2;

/* Scheme switches (the UCNS3D.DAT fields on the path, src/parameters.f90:215-287),
 * then the library's own options, which have no UCNS3D.DAT counterpart. */
typedef struct ucns_scheme_v2 {
    int32_t abi_version;
    int32_t precision;          /* UCNS_FP64 | UCNS_MIXED */
    int32_t iweno;              /* 1 = WENO (spatiladiscret 3), -1 = MUSCL (2) */
    int32_t iorder;             /* IORDER = spatialorder-1 */
    int32_t idegfree;           /* coefficients without the mean */
    int32_t inumneighbours;     /* stencil size including the cell */
    int32_t typesten;           /* stencil slots per cell (1 MUSCL, 7 WENO 3D) */
    int32_t ees;                /* 0 (WENO) or 5 (CWENO: the central polynomial reconstituted from the
                                   sector polynomials, src/reconstruct_p.f90:955-966) */
    int32_t wenoz;              /* 0, or 1 with ees = 5 (CWENO-Z weights, src/reconstruct_p.f90:1019-1029) */
    int32_t wenwrt;             /* 1 conserved-variable reconstruction */
    int32_t limiter;            /* 1 Barth-Jespersen */
    int32_t riemann;            /* 1 HLLC */
    int32_t qp_quad;            /* Gauss points per quadrilateral face */
    int32_t qp_triangle;        /* Gauss points per triangular face */
    double gamma;
    double lwci1;               /* wenocentralweight */
    double weno_eps;            /* 10E-12 = 1e-11 in the reference */
    const double *weights_quad; /* [qp_quad], sum 1 */
    const double *weights_tri;  /* [qp_triangle], sum 1 */
    /* library options */
    int32_t cell_order;         /* UCNS_ORDER_CALLER | UCNS_ORDER_RCM | UCNS_ORDER_MORTON */
    int32_t kernels;            /* UCNS_KERNELS_AUTO | UCNS_KERNELS_REGISTER */
    int32_t adj_keep_stages;    /* ucns_rk3_adjoint keeps the forward reconstruction of at most
                                   this many of a march's last stages; 0: no count limit (the
                                   byte limit alone); < 0: none (every reversed stage runs its
                                   reconstruction again) */
    int32_t reserved;           /* 0 */
    double adj_keep_gib;        /* and of at most this many GiB of them (never more than the free
                                   device memory less 4 GiB); <= 0 selects 64 */
} ucns_scheme_v2;

/* Per-cell precomputed reconstruction data (PRESTORE_RECONSTRUCTION3 and
 * MEMORY_FAST outputs, src/prestore.f90:167-815, src/reconstruct_p.f90:45-302).
 * nqp_max = max(qp_quad, qp_triangle); imax = inumneighbours-1. */
typedef struct ucns_prestore_v1 {
    const int32_t *admis;       /* [ncell] IELEM%ADMIS (complete stencils) */
    const int32_t *full;        /* [ncell] IELEM%FULL */
    const int32_t *ihexl;       /* [ncell][typesten][inumneighbours] IHEXL(s,1:nnb), first = cell itself */
    const double *invmat;       /* [ncell][typesten][imax][idegfree] INVMAT_STENCILT(ideg,imax,s) */
    const double *indicator;    /* [ncell][idegfree][idegfree] INDICATOR (WENO only, may be NULL for MUSCL) */
    const double *integ_basis;  /* [ncell][idegfree] INTEG_BASIS%VALUE */
    const double *volume11;     /* [ncell] ILOCAL_RECON3%VOLUME(1,1) */
    const double *qpoints;      /* [ncell][nface_max][nqp_max][3] QPOINTS(L,NGP,1:3), local frame */
} ucns_prestore_v1;

/* CWENO data (EES=5): the sector stencils s >= 2 (1-based) of a cell take a
 * polynomial of degree IORDER2 (idegfree2 coefficients: 3 for IORDER <= 3, 9
 * above) over their first inumneighbours2 members (12 or 20, the cell first),
 * with the means of INTEG_BASIS%VALUEC (src/grid_p.f90:6151-6163,
 * src/prestore.f90:244-263, 296-337, 520-535).  imax2 = inumneighbours2-1. */
typedef struct ucns_cweno_v1 {
    int32_t abi_version;
    int32_t idegfree2;
    int32_t inumneighbours2;
    int32_t reserved;
    const double *invmat_c;      /* [ncell][typesten][imax2][idegfree2] INVMAT_STENCILTC(ideg2,imax2,s);
                                    the slot of stencil 1 is not read */
    const double *indicator_c;   /* [ncell][idegfree2][idegfree2] INDICATORC */
    const double *integ_basis_c; /* [ncell][idegfree2] INTEG_BASIS%VALUEC */
} ucns_cweno_v1;

/* Accumulated device-side timings and traffic model (ucns_stats). */
typedef struct ucns_stats_v1 {
    int64_t n_recon, n_flux, n_update, n_dt;      /* launches */
    double ms_recon, ms_flux, ms_update, ms_dt;   /* summed kernel time (hipEvents on the library stream) */
    double bytes_recon_per_cell;                  /* algorithmic bytes per cell of the reconstruction kernel */
    int64_t n_reduced_last;                       /* cells sent to MUSCL by CHECKSOL in the last residual */
    int64_t n_firstorder_last;                    /* cells sent to first order by CHECKSOLX */
    int64_t n_recon_vjp, n_flux_vjp;              /* launches of the residual-product kernels */
    double ms_recon_vjp, ms_flux_vjp;             /* their summed kernel time */
    int64_t n_recon_adj;                          /* reconstructions run by the adjoint (with the scatter) */
    double ms_recon_adj;
} ucns_stats_v1;

/* Ghost states of boundary faces (BOUNDARYS, src/flow_operations.f90:2860-3375),
 * Euler 3D, no rotating frame.  Face codes (bcode): 1 inflow, 2 outflow,
 * 3 symmetry, 4 wall (slip, the inviscid branch), 6 farfield.  inflow/outflow
 * are the conserved states INFLOW() and OUTFLOW() return; this version takes
 * them uniform over the boundary (a free stream). */
typedef struct ucns_boundary_v1 {
    int32_t abi_version;
    int32_t boundtype;          /* BOUNDTYPE: 0 supersonic (inflow = free stream, outflow = copy),
                                   1 characteristic (subsonic/supersonic from the interior Mach number) */
    double inflow[5];           /* conserved */
    double outflow[5];          /* conserved */
} ucns_boundary_v1;

/* Cotangents of an objective with respect to the data a march reads from
 * the mesh and PRESTORE arrays (ucns_rk3_adjoint_geometry), in the layouts of
 * ucns_prestore_v1 and ucns_mesh_v2 and the caller's cell order.  Any pointer
 * may be NULL (not returned).  A host chains them with the derivatives of
 * its own geometry (PRESTORE_RECONSTRUCTION3, LOCALISE_STEN2, the face
 * angles and areas of GEOMETRY_CALC) to reach mesh-shape parameters. */
typedef struct ucns_geometry_bar_v1 {
    int32_t abi_version;
    int32_t reserved;
    double *invmat;       /* [ncell][typesten][imax][idegfree] */
    double *indicator;    /* [ncell][idegfree][idegfree]: entries k <= l (the upper triangle the
                             library reads; the matrix is symmetric), zero below */
    double *integ_basis;  /* [ncell][idegfree] */
    double *volume11;     /* [ncell] */
    double *qpoints;      /* [ncell][nface_max][nqp_max][3] */
    double *surf;         /* [ncell][nface_max]: on each face's owner record (the lower cell index,
                             or the only one on boundary faces), zero on the other */
    double *normal;       /* [ncell][nface_max][3]: the unit outward normal
                             (cos a1 sin a2, sin a1 sin a2, cos a2) of FACEANGLEX/Y = (a1, a2), on the
                             owner record; tangential to the unit sphere (the normal's own direction
                             has no effect) */
    double *totvolume;    /* [ncell] */
} ucns_geometry_bar_v1;

typedef struct ucns_ctx ucns_ctx;

/* Set up one partition on the current HIP device (topology, faces, scheme). */
int  ucns_create(const ucns_mesh_v2 *mesh, const ucns_scheme_v2 *scheme, ucns_ctx **out);
/* Upload PRESTORE data (pseudo-inverses, indicators, qp coordinates ...).
 * With ees = 5 (CWENO) the CWENO data must have been given first. */
int  ucns_upload_prestore(ucns_ctx *ctx, const ucns_prestore_v1 *pre);
/* CWENO data for a context created with ees = 5 (CP_RECONSTRUCTION's EES=5
 * branch, src/reconstruct_p.f90:955-1014), before ucns_upload_prestore. */
int  ucns_upload_cweno(ucns_ctx *ctx, const ucns_cweno_v1 *cw);
/* Free-stream data of the boundary codes 1, 2 and 6 (needed before the first
 * residual when the mesh has such faces; codes 3 and 4 need none). */
int  ucns_set_boundary(ucns_ctx *ctx, const ucns_boundary_v1 *bc);
/* Position-dependent INFLOW() / OUTFLOW() (src/flow_operations.f90:1161-1302,
 * evaluated by the host at each boundary face's centre, the mean of its nodes
 * as GET_STATES_BOUNDS takes it, src/dg_functions.f90:1818-1829): conserved
 * states per face record, [ncell][nface_max][5] in the mesh's cell order,
 * read at the faces of codes 1, 2 and 6 (the inflow array for 1 and 6, the
 * outflow array for 2 and 6).  A null array keeps the uniform state of
 * ucns_set_boundary for its side; both null return to uniform states.  Call
 * after ucns_set_boundary, which still gives BOUNDTYPE.  The adjoint entry
 * points refuse a context that holds per-face states. */
int  ucns_set_boundary_faces(ucns_ctx *ctx, const double *inflow, const double *outflow);
/* Cell averages U_C(i)%VAL(1,1:5) in / out. */
int  ucns_set_state(ucns_ctx *ctx, const double *u, int32_t layout);
int  ucns_get_state(ucns_ctx *ctx, double *u, int32_t layout);
/* Same, with u a device pointer (no host staging). */
int  ucns_set_state_device(ucns_ctx *ctx, const double *u_dev, int32_t layout);
int  ucns_get_state_device(ucns_ctx *ctx, double *u_dev, int32_t layout);
/* CALCULATE_CFL: dt = (CFL/3) min_i MINEDGE_i / (max(|u|,|v|,|w|)+c). */
int  ucns_compute_dt(ucns_ctx *ctx, double cfl, double *dt_out);
/* CALL_FLUX_SUBROUTINES_3D on the current state: RHS(i)%VAL(1:5) (face-flux
 * sums, not divided by the volume).  rhs_out may be NULL (result stays on device). */
int  ucns_residual(ucns_ctx *ctx, double *rhs_out, int32_t layout);
/* Same, with rhs_dev a device pointer (no host staging). */
int  ucns_residual_device(ucns_ctx *ctx, double *rhs_dev, int32_t layout);
/* ucns_residual_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (the face fluxes and troubled-cell counters) is a static device
 * array of 256 MiB per GPU, held from the library's first use on that GPU
 * until the process ends; meshes with more than about 6.7 million faces are
 * rejected with UCNS_EUNSUPPORTED.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (from several threads or contexts) race. */
int  ucns_residual_noalloc(ucns_ctx *ctx, double *rhs_dev, int32_t layout);
/* ARBITRARY_ORDER only: ILOCAL_RECON3(i)%ULEFT(1:5,L,NGP) as [ncell][nface_max][nqp_max][5]. */
int  ucns_reconstruct(ucns_ctx *ctx, double *uleft_out);
/* TIME_MARCHING's RK3 loop for nsteps steps.  If dt_fixed > 0 it is used for
 * every step, else dt is recomputed each step from cfl.  *t_io is advanced. */
int  ucns_rk3_steps(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, double *t_io);
/* Adjoint of the residual (no counterpart in the reference, which has no
 * adjoint): u_bar = (dR/dU)^T w at the current state, R as ucns_residual
 * returns it.  bc_bar (10 doubles, or NULL) receives
 * (dR/d(inflow[0..4], outflow[0..4]))^T w for the boundary data of
 * ucns_set_boundary.  Decisions that are discrete in the primal (CHECKSOL /
 * CHECKSOLX flags, HLLC and boundary regimes, the Barth limiter's active face
 * point and bound) are those of the current state. */
int  ucns_residual_vjp(ucns_ctx *ctx, const double *w, double *u_bar, int32_t layout, double *bc_bar);
/* Same with device pointers, AOS layout (bc_bar_dev: 10 doubles or NULL). */
int  ucns_residual_vjp_device(ucns_ctx *ctx, const double *w_dev, double *u_bar_dev, double *bc_bar_dev);
/* Adjoint of nsteps SSP-RK3 steps (no counterpart in the reference): runs
 * the steps from the current state as ucns_rk3_steps does, storing every
 * stage state on the device (3 * nsteps * ncell * 40 bytes), then sweeps back:
 * u0_bar = (dU_N/dU_0)^T uN_bar and bc_bar (10 doubles, or NULL) =
 * (dU_N/d(inflow, outflow))^T uN_bar.  The step sizes are those of the
 * forward sweep, held fixed in the derivative.  The state is left at U_N and
 * *t_io (or NULL) is advanced as by ucns_rk3_steps.  The last stages of the
 * forward sweep also keep their reconstruction for the reverse sweep, as
 * many as the scheme's adj_keep_gib and adj_keep_stages allow; the context
 * holds those buffers until ucns_destroy or ucns_upload_prestore. */
int  ucns_rk3_adjoint(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, const double *uN_bar,
                      double *u0_bar, int32_t layout, double *bc_bar, double *t_io);
/* Same with device pointers, AOS layout (bc_bar_dev: 10 doubles or NULL). */
int  ucns_rk3_adjoint_device(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, const double *uN_bar_dev,
                             double *u0_bar_dev, double *bc_bar_dev, double *t_io);
/* ucns_rk3_adjoint with a fixed step size that also returns, in *gbar, the
 * cotangents of the data the march reads from the mesh and PRESTORE arrays
 * (no counterpart in the reference, which has no adjoint): the stencil
 * pseudo-inverses, indicators, basis integrals, VOLUME(1,1), face Gauss
 * points, face areas and normals and the cell volumes, summed over every
 * stage of the march.  FP64 contexts on the register kernels only
 * (UCNS_EUNSUPPORTED otherwise); the same discrete decisions are held as in
 * ucns_residual_vjp.  gbar->abi_version must be UCNS_ABI_VERSION. */
int  ucns_rk3_adjoint_geometry(ucns_ctx *ctx, int32_t nsteps, double dt_fixed, const double *uN_bar, double *u0_bar,
                               int32_t layout, double *bc_bar, const ucns_geometry_bar_v1 *gbar, double *t_io);
int  ucns_stats(ucns_ctx *ctx, ucns_stats_v1 *out);
/* Fault injection for tests of the error paths (no counterpart in the
 * reference): the nth device allocation the context makes from now on
 * fails with UCNS_ENOMEM (0 disarms). */
int  ucns_debug_fail_alloc(ucns_ctx *ctx, int32_t nth);
int  ucns_reset_stats(ucns_ctx *ctx);
const char *ucns_last_error(const ucns_ctx *ctx);
void ucns_destroy(ucns_ctx *ctx);
/* ABI version compiled into the library. */
int  ucns_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* UCNS_H */
