/*
 * ucns_oracle.c -- CPU restatement of UCNS3D's explicit residual and SSP-RK3
 * update, in FP64, following the reference's own loop structure.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product path (ucns3d_amd/, libucnsgpu.so) never calls it.
 *
 * It reads the same ucns_*_v1 structs as the GPU library (include/ucns.h)
 * and keeps the reference's arithmetic form where it decides rounding:
 *   - one LSQ contraction per stencil over dU = U_j - U_i     (gradients.f90:253-497)
 *   - per-variable indicators B*g, then g.(B*g)                (reconstruct_p.f90:995-1004)
 *   - omega~ = lambda / (eps+beta)^4, normalised                (reconstruct_p.f90:1031-1049)
 *   - ULEFT = sum_s omega_s (U_i + CONSMATRIX*G_s), per stencil (reconstruct_p.f90:1126-1199, 1556-1583)
 *   - CHECKSOL 0.85 / MUSCL+Barth / CHECKSOLX 0.95              (reconstruct_p.f90:3126-3406, 2044-2352)
 *   - cell loop over faces, every interior face solved twice    (flux_p.f90:526-1019)
 *   - ROTATEF / HLLC / ROTATEB with cos/sin rebuilt per call     (grid_t.f90:4217-4305, riemann.f90:221-426)
 *   - Shu-Osher SSP-RK3 with the U_C slots 1..3                 (time_p.f90:885-1027)
 *   - dt = (CFL/3) min MINEDGE/(max|u_k| + c)                    (time_p.f90:27-116)
 *
 * Parity status of each function is listed in DESIGN.md (oracle section).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ucns.h"

#define NV 5

/* CONS2PRIM2 for one state, Euler 3D (flow_operations.f90:447-467) */
static void cons2prim(double gamma, const double *u, double *w) {
    double oo = 1.0 / u[0];
    w[0] = u[0];
    w[1] = u[1] * oo;
    w[2] = u[2] * oo;
    w[3] = u[3] * oo;
    w[4] = (gamma - 1.0) * (u[4] - 0.5 * u[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}

/* ROTATEF (grid_t.f90:4217-4258): rot = TRI * vec */
void oracle_rotatef(double *rot, const double *vec, double a1, double a2) {
    double coa1 = cos(a1), sia1 = sin(a1), coa2 = cos(a2), sia2 = sin(a2);
    rot[0] = vec[0];
    rot[1] = coa1 * sia2 * vec[1] + sia1 * sia2 * vec[2] + coa2 * vec[3];
    rot[2] = coa1 * coa2 * vec[1] + sia1 * coa2 * vec[2] + (-sia2) * vec[3];
    rot[3] = (-sia1) * vec[1] + coa1 * vec[2];
    rot[4] = vec[4];
}

/* ROTATEB (grid_t.f90:4261-4305): vec = INVTRI * rot */
void oracle_rotateb(double *vec, const double *rot, double a1, double a2) {
    double coa1 = cos(a1), sia1 = sin(a1), coa2 = cos(a2), sia2 = sin(a2);
    vec[0] = rot[0];
    vec[1] = coa1 * sia2 * rot[1] + coa1 * coa2 * rot[2] + (-sia1) * rot[3];
    vec[2] = sia1 * sia2 * rot[1] + sia1 * coa2 * rot[2] + coa1 * rot[3];
    vec[3] = coa2 * rot[1] + (-sia2) * rot[2];
    vec[4] = rot[4];
}

/* Free-stream data of the boundary codes (set by oracle_set_boundary). */
static ucns_boundary_v1 g_bc = {UCNS_ABI_VERSION, 0, {1, 0, 0, 0, 2.5}, {1, 0, 0, 0, 2.5}};

void oracle_set_boundary(const ucns_boundary_v1 *bc) { g_bc = *bc; }

#define TOLSMALL 1.0e-13 /* grid_p.f90:18 */

/* PRIM2CONS of (rho, u, v, w, p) with the reference's energy form
 * rho*(p/((gamma-1) rho)) + rho*(|u|^2/2) (flow_operations.f90:2929-2935) */
static void bc_prim2cons(double gamma, const double *s3, double *cr) {
    cr[0] = s3[0];
    cr[1] = s3[1] * s3[0];
    cr[2] = s3[2] * s3[0];
    cr[3] = s3[3] * s3[0];
    double skins = 0.5 * ((s3[1] * s3[1]) + (s3[2] * s3[2]) + (s3[3] * s3[3]));
    double ikins = s3[4] / ((gamma - 1.0) * (s3[0]));
    cr[4] = (s3[0] * (ikins)) + (s3[0] * skins);
}

/* subsonic inflow: pressure from the outgoing characteristic, the rest from
 * the free stream (flow_operations.f90:2907-2915); s1 free stream, s2 interior */
static void bc_sub_in(double gamma, const double *s1, const double *s2, double sps, double nx, double ny, double nz,
                      double *cr) {
    double s3[NV];
    s3[4] = 0.5 * ((s1[4]) + (s2[4]) - (s2[0] * sps * ((nx * (s1[1] - s2[1])) + (ny * (s1[2] - s2[2])) + (nz * (s1[3] - s2[3])))));
    s3[0] = s1[0] + (s3[4] - s1[4]) / (sps * sps);
    s3[1] = s1[1] - (nx * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s1[2] - (ny * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s1[3] - (nz * (s1[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

/* subsonic outflow: free-stream pressure, the rest from the interior
 * (flow_operations.f90:2995-3010); s1 free stream, s2 interior */
static void bc_sub_out(double gamma, const double *s1, const double *s2, double sps, double nx, double ny, double nz,
                       double *cr) {
    double s3[NV];
    s3[4] = s1[4];
    s3[0] = s2[0] + (s3[4] - s2[4]) / (sps * sps);
    s3[1] = s2[1] + (nx * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s2[2] + (ny * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s2[3] + (nz * (s2[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

/* BOUNDARYS (flow_operations.f90:2860-3375), Euler 3D without a rotating
 * frame: ghost state cr of a boundary face from the face-point state cl */
void oracle_ghost(int code, double gamma, const double *cl, double a1, double a2, double *cr) {
    const double nx = (cos(a1) * sin(a2)), ny = (sin(a1) * sin(a2)), nz = (cos(a2));
    double s1[NV], s2[NV];
    switch (code) {
    case 1: /* inflow (2880-2925) */
        if (g_bc.boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = g_bc.inflow[v];
        } else {
            cons2prim(gamma, g_bc.inflow, s1);
            cons2prim(gamma, cl, s2);
            double sps = sqrt((gamma * s2[4]) / (s2[0]));
            double vel = sqrt(s2[1] * s2[1] + s2[2] * s2[2] + s2[3] * s2[3]);
            if (vel / (sps + TOLSMALL) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = g_bc.inflow[v];
            } else {
                bc_sub_in(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 2: /* outflow (2965-3013) */
        if (g_bc.boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = cl[v];
        } else {
            cons2prim(gamma, g_bc.outflow, s1);
            cons2prim(gamma, cl, s2);
            double sps = sqrt((gamma * s2[4]) / (s2[0]));
            double vel = sqrt(s2[1] * s2[1] + s2[2] * s2[2] + s2[3] * s2[3]);
            if (vel / (sps + TOLSMALL) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                bc_sub_out(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 3: /* symmetry (3086-3131) */
    case 4: /* wall, inviscid branch (3132-3160) */
    {
        double rot[NV], rr[NV];
        oracle_rotatef(rot, cl, a1, a2);
        for (int v = 0; v < NV; ++v) rr[v] = rot[v];
        rr[1] = -rot[1];
        oracle_rotateb(cr, rr, a1, a2);
        break;
    }
    case 6: /* farfield (3203-3296): direction from the rotated normal momentum */
    {
        double rot[NV];
        oracle_rotatef(rot, cl, a1, a2);
        const double vnb = rot[1];
        cons2prim(gamma, cl, s2);
        double sps = sqrt((gamma * s2[4]) / (s2[0]));
        if (vnb <= 0.0) {
            if (fabs(vnb) >= sps) {
                for (int v = 0; v < NV; ++v) cr[v] = g_bc.inflow[v];
            } else {
                cons2prim(gamma, g_bc.inflow, s1);
                bc_sub_in(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        } else {
            if (fabs(vnb) >= sps) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                cons2prim(gamma, g_bc.outflow, s1);
                bc_sub_out(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        }
        break;
    }
    default:
        for (int v = 0; v < NV; ++v) cr[v] = cl[v];
    }
}

/* HLLC_RIEMANN_SOLVER, Euler, no MRF (riemann.f90:221-426) */
void oracle_hllc(double gamma, const double *cl, const double *cr, double *flux) {
    double wl[NV], wr[NV];
    cons2prim(gamma, cl, wl);
    cons2prim(gamma, cr, wr);
    double rl = wl[0], ul = wl[1], vl = wl[2], zl = wl[3], pl = wl[4], el = cl[4];
    double rr = wr[0], ur = wr[1], vr = wr[2], zr = wr[3], pr = wr[4], er = cr[4];
    double ccl = sqrt(gamma * pl / rl), ccr = sqrt(gamma * pr / rr);
    double srl = sqrt(rl), srr = sqrt(rr);
    double cc2 = sqrt((((ccl * ccl) * srl) + ((ccr * ccr) * srr)) / (srl + srr) +
                      (0.5 * ((srl * srr) / ((srl + srr) * (srl + srr))) * ((ur - ul) * (ur - ul))));
    double uu2 = ((ul * srl) + (ur * srr)) / (srl + srr);
    double sl = fmin(ul - ccl, uu2 - cc2), sr = fmax(ur + ccr, uu2 + cc2);
    sl = fmin(sl, 0.0);
    sr = fmax(sr, 0.0);
    double sm = (pr - pl + (rl * ul * (sl - ul)) - (rr * ur * (sr - ur))) / ((rl * (sl - ul)) - (rr * (sr - ur)));
    double fl[NV] = {rl * ul, (rl * (ul * ul)) + pl, rl * ul * vl, rl * ul * zl, ul * (el + pl)};
    double fr[NV] = {rr * ur, (rr * (ur * ur)) + pr, rr * ur * vr, rr * ur * zr, ur * (er + pr)};
    double mul = rl * ((sl - ul) / (sl - sm));
    double mur = rr * ((sr - ur) / (sr - sm));
    double lastl = (el / rl) + ((sm - ul) * (sm + ((pl) / (rl * (sl - ul)))));
    double lastr = (er / rr) + ((sm - ur) * (sm + ((pr) / (rr * (sr - ur)))));
    double usl[NV] = {mul, mul * sm, mul * vl, mul * zl, mul * lastl};
    double usr[NV] = {mur, mur * sm, mur * vr, mur * zr, mur * lastr};
    double s = copysign(1.0, sm);
    for (int k = 0; k < NV; ++k)
        flux[k] = (((1.0 + s) / 2.0) * (fl[k] + sl * (usl[k] - cl[k]))) +
                  (((1.0 - s) / 2.0) * (fr[k] + sr * (usr[k] - cr[k])));
}

/* mean-subtracted monomial basis at a local-frame point (BASIS_REC, basis.f90:11-102, 1720-1734) */
static const int MONO[34][3] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
    {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2},
    {3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}, {2, 0, 1}, {1, 1, 1}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {0, 0, 3},
    {4, 0, 0}, {3, 1, 0}, {2, 2, 0}, {1, 3, 0}, {0, 4, 0}, {3, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 3, 1},
    {2, 0, 2}, {1, 1, 2}, {0, 2, 2}, {1, 0, 3}, {0, 1, 3}, {0, 0, 4}};

static void basis_rec(const double *x, int ideg, const double *integ, double vol11, double *phi) {
    double px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = 1.0;
    for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * x[0];
        py[k] = py[k - 1] * x[1];
        pz[k] = pz[k - 1] * x[2];
    }
    double oov = 1.0 / vol11;
    for (int d = 0; d < ideg; ++d)
        phi[d] = px[MONO[d][0]] * py[MONO[d][1]] * pz[MONO[d][2]] - integ[d] * oov;
}

typedef struct {
    const ucns_mesh_v1 *m;
    const ucns_scheme_v1 *s;
    const ucns_prestore_v1 *p;
    int nf, nqm, ideg, nnb, imax, ts, base;
} ctx_t;

static int nqp_of(const ctx_t *c, int64_t i, int f) {
    return c->m->types_faces[i * c->nf + f] == 5 ? c->s->qp_quad : c->s->qp_triangle;
}

static double *uleft_at(const ctx_t *c, double *ul, int64_t i, int f, int q) {
    return ul + (((size_t)i * c->nf + f) * c->nqm + q) * NV;
}

/* omega_s = lambda_s / (eps + beta_s)^4, normalised; lambda = lwci1 for the
 * central stencil, 1 for the directional ones (reconstruct_p.f90:1031-1049) */
static void weno_weights(int admis, const double *beta, double lwci1, double eps, double *omega) {
    double wt[16], sum = 0.0;
    for (int ll = 0; ll < admis; ++ll) {
        double lam = (ll == 0) ? lwci1 : 1.0;
        double x = eps + beta[ll];
        double x2 = x * x;
        wt[ll] = lam / (x2 * x2);
    }
    for (int ll = 0; ll < admis; ++ll) sum += wt[ll];
    for (int ll = 0; ll < admis; ++ll) omega[ll] = wt[ll] / sum;
}

/* one cell of ARBITRARY_ORDER: LEAST_SQUARES + WENOWEIGHTS (CP_RECONSTRUCTION) */
static void weno_cell(const ctx_t *c, const double *u, int64_t i, double *ul, double *grad /*ts*ideg*5*/,
                      double *work) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb, ts = c->ts;
    const int admis = c->p->admis[i];
    const double *ui = u + i * NV;
    double *mat1 = work;                  /* imax*5 */
    double *phi = work + imax * NV;       /* ideg */
    double *bg = phi + ideg;              /* ideg */
    double omega[8][NV];
    /* LSQ (COMPUTE_GRADIENTS_MEAN_LSQ, LOCAL=1 branch) */
    for (int ll = 0; ll < admis; ++ll) {
        const int32_t *ih = c->p->ihexl + ((size_t)i * ts + ll) * nnb;
        const double *s1 = u + (size_t)(ih[0] - c->base) * NV;
        for (int iq = 0; iq < imax; ++iq) {
            const double *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
            for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
        }
        const double *P = c->p->invmat + ((size_t)i * ts + ll) * imax * ideg;
        for (int k = 0; k < ideg; ++k)
            for (int v = 0; v < NV; ++v) {
                double acc = 0.0;
                for (int iq = 0; iq < imax; ++iq) acc += P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
                grad[((size_t)ll * ideg + k) * NV + v] = acc;
            }
    }
    /* smoothness indicators and weights, per variable */
    const double *B = c->p->indicator + (size_t)i * ideg * ideg;
    const double eps = c->s->weno_eps;
    for (int v = 0; v < NV; ++v) {
        double beta[8], wt[8];
        for (int ll = 0; ll < admis; ++ll) {
            for (int a = 0; a < ideg; ++a) {
                double acc = 0.0;
                for (int b = 0; b < ideg; ++b) acc += B[(size_t)b * ideg + a] * grad[((size_t)ll * ideg + b) * NV + v];
                bg[a] = acc;
            }
            double dot = 0.0;
            for (int a = 0; a < ideg; ++a) dot += grad[((size_t)ll * ideg + a) * NV + v] * bg[a];
            beta[ll] = dot;
        }
        weno_weights(admis, beta, c->s->lwci1, eps, wt);
        for (int ll = 0; ll < admis; ++ll) omega[ll][v] = wt[ll];
    }
    /* extrapolation to the face Gauss points, stencil by stencil */
    const double *integ = c->p->integ_basis + (size_t)i * ideg;
    const double vol11 = c->p->volume11[i];
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            double *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) o[v] = 0.0;
        }
    for (int ll = 0; ll < admis; ++ll)
        for (int f = 0; f < c->m->ifca[i]; ++f)
            for (int q = 0; q < nqp_of(c, i, f); ++q) {
                const double *x = c->p->qpoints + (((size_t)i * c->nf + f) * c->nqm + q) * 3;
                basis_rec(x, ideg, integ, vol11, phi);
                double *o = uleft_at(c, ul, i, f, q);
                for (int v = 0; v < NV; ++v) {
                    double res = 0.0;
                    for (int k = 0; k < ideg; ++k) res += phi[k] * grad[((size_t)ll * ideg + k) * NV + v];
                    o[v] = o[v] + (ui[v] + res) * omega[ll][v];
                }
            }
}

/* CHECKSOL / CHECKSOLX face-point test (reconstruct_p.f90:3183-3196, 3294-3310) */
static int trouble_test(const ctx_t *c, const double *u, const double *ul, int64_t i, double jump) {
    double wi[NV], wq[NV];
    cons2prim(c->s->gamma, u + i * NV, wi);
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            cons2prim(c->s->gamma, uleft_at(c, (double *)ul, i, f, q), wq);
            if (fabs(wq[0] - wi[0]) >= jump * wi[0]) return 1;
            if (fabs(wq[4] - wi[4]) >= jump * wi[4]) return 1;
            if (wq[4] < 0.0 || wq[0] < 0.0) return 1;
        }
    return 0;
}

/* Barth-Jespersen psi of one face point value usol (reconstruct_p.f90:3439-3453) */
static double barth_psi(double ui, double umin, double umax, double usol) {
    double d2 = usol - ui;
    if (fabs(d2) <= 0.0) return 1.0;
    if (d2 > 0.0) return fmin(1.0, (umax - ui) / d2);
    return fmin(1.0, (umin - ui) / d2);
}

/* MUSCL on one cell: FIND_BOUNDS (strict) + unlimited central polynomial +
 * Barth-Jespersen (reconstruct_p.f90:1821-2039, 2044-2280, 3421-3453) */
static void muscl_cell(const ctx_t *c, const double *u, int64_t i, double *ul, const double *grad, double *work) {
    const int ideg = c->ideg, nf = c->nf;
    const double *ui = u + i * NV;
    double umin[NV], umax[NV], slope[NV];
    for (int v = 0; v < NV; ++v) umin[v] = umax[v] = ui[v];
    for (int f = 0; f < c->m->ifca[i]; ++f) {
        int bc = c->m->bcode[i * nf + f];
        if (bc != 0 && bc != 5 && bc != 50) continue;  /* non-periodic boundary faces add no neighbour */
        const double *un = u + (size_t)(c->m->ineigh[i * nf + f] - c->base) * NV;
        for (int v = 0; v < NV; ++v) {
            umin[v] = fmin(umin[v], un[v]);
            umax[v] = fmax(umax[v], un[v]);
        }
    }
    double *phi = work;
    const double *integ = c->p->integ_basis + (size_t)i * ideg;
    const double vol11 = c->p->volume11[i];
    for (int v = 0; v < NV; ++v) slope[v] = 1.0e16; /* TOLBIG */
    /* pass 1: USOL and the limiter */
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            const double *x = c->p->qpoints + (((size_t)i * nf + f) * c->nqm + q) * 3;
            basis_rec(x, ideg, integ, vol11, phi);
            for (int v = 0; v < NV; ++v) {
                double res = 0.0;
                for (int k = 0; k < ideg; ++k) res += phi[k] * grad[(size_t)k * NV + v];
                double usol = ui[v] + res;
                slope[v] = fmin(slope[v], barth_psi(ui[v], umin[v], umax[v], usol));
            }
        }
    /* pass 2: ULEFT = U_i + (USOL - U_i) * SLOPE */
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            const double *x = c->p->qpoints + (((size_t)i * nf + f) * c->nqm + q) * 3;
            basis_rec(x, ideg, integ, vol11, phi);
            double *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) {
                double res = 0.0;
                for (int k = 0; k < ideg; ++k) res += phi[k] * grad[(size_t)k * NV + v];
                double usol = ui[v] + res;
                o[v] = 0.0 + (ui[v] + ((usol - ui[v]) * slope[v]));
            }
        }
}

static void lsq_central(const ctx_t *c, const double *u, int64_t i, double *grad, double *work) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb;
    const int32_t *ih = c->p->ihexl + (size_t)i * c->ts * nnb;
    const double *s1 = u + (size_t)(ih[0] - c->base) * NV;
    double *mat1 = work;
    for (int iq = 0; iq < imax; ++iq) {
        const double *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
        for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
    }
    const double *P = c->p->invmat + (size_t)i * c->ts * imax * ideg;
    for (int k = 0; k < ideg; ++k)
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
            for (int iq = 0; iq < imax; ++iq) acc += P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
            grad[(size_t)k * NV + v] = acc;
        }
}

static void init_ctx(ctx_t *c, const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const ucns_prestore_v1 *p) {
    c->m = m;
    c->s = s;
    c->p = p;
    c->nf = m->nface_max;
    c->nqm = s->qp_quad > s->qp_triangle ? s->qp_quad : s->qp_triangle;
    c->ideg = s->idegfree;
    c->nnb = s->inumneighbours;
    c->imax = s->inumneighbours - 1;
    c->ts = s->typesten;
    c->base = m->index_base;
}

/* ARBITRARY_ORDER (reconstruct_p.f90:2954-3015) for all cells -> ULEFT.
 * counters (optional): [0] cells sent to MUSCL, [1] cells set to first order. */
int oracle_reconstruct(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const ucns_prestore_v1 *p, const double *u,
                       double *ul, int64_t *counters) {
    ctx_t c;
    init_ctx(&c, m, s, p);
    const int64_t nc = m->ncell;
    int64_t n_red = 0, n_first = 0;
    int *reduce = (int *)calloc((size_t)nc, sizeof(int));
    if (!reduce) return UCNS_ENOMEM;
#pragma omp parallel
    {
        double *grad = (double *)malloc(sizeof(double) * (size_t)c.ts * c.ideg * NV);
        double *work = (double *)malloc(sizeof(double) * ((size_t)c.imax * NV + 4 * c.ideg + 16));
        /* WENOWEIGHTS + CHECKSOL, or nothing for MUSCL (IWENO=-1) */
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            if (s->iweno == 1) {
                weno_cell(&c, u, i, ul, grad, work);
                int red = 0;
                if (p->full[i] == 0) red = 1;
                else red = trouble_test(&c, u, ul, i, 0.85);
                reduce[i] = red;
            } else {
                reduce[i] = 0;
            }
        }
        /* MUSCL for (TROUBLED & REDUCE) | (TROUBLED & ~FULL); TROUBLED=1 for every FV cell */
#pragma omp for schedule(static) reduction(+ : n_red)
        for (int64_t i = 0; i < nc; ++i) {
            if (reduce[i] == 1 || p->full[i] == 0) {
                lsq_central(&c, u, i, grad, work);
                muscl_cell(&c, u, i, ul, grad, work);
                n_red += 1;
            }
        }
        /* CHECKSOLX: first order where the limited values still jump */
#pragma omp for schedule(static) reduction(+ : n_first)
        for (int64_t i = 0; i < nc; ++i) {
            if (trouble_test(&c, u, ul, i, 0.95)) {
                for (int f = 0; f < m->ifca[i]; ++f)
                    for (int q = 0; q < c.nqm; ++q) {
                        double *o = uleft_at(&c, ul, i, f, q);
                        for (int v = 0; v < NV; ++v) o[v] = u[i * NV + v];
                    }
                n_first += 1;
            }
        }
        free(grad);
        free(work);
    }
    free(reduce);
    if (counters) {
        counters[0] = n_red;
        counters[1] = n_first;
    }
    return UCNS_OK;
}

/* CALCULATE_FLUXESHI_CONVECTIVE, cell based, every face solved from both sides */
int oracle_fluxes(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const ucns_prestore_v1 *p, const double *ul,
                  double *rhs) {
    ctx_t c;
    init_ctx(&c, m, s, p);
    const int64_t nc = m->ncell;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        double r[NV] = {0, 0, 0, 0, 0};
        for (int f = 0; f < m->ifca[i]; ++f) {
            const int64_t slot = i * c.nf + f;
            const double a1 = m->faceanglex[slot], a2 = m->faceangley[slot];
            const int iqp = nqp_of(&c, i, f);
            const double *wts = m->types_faces[slot] == 5 ? s->weights_quad : s->weights_tri;
            const int bcode = m->bcode[slot];
            const int bnd = bcode != 0 && bcode != 5 && bcode != 50;
            const int64_t j = bnd ? 0 : m->ineigh[slot] - c.base;
            const int fj = bnd ? 0 : m->ineighn[slot] - c.base;
            double god[NV] = {0, 0, 0, 0, 0};
            for (int q = 0; q < iqp; ++q) {
                const double *cl = uleft_at(&c, (double *)ul, i, f, q);
                double ghost[NV];
                const double *cr;
                if (bnd) { /* GET_STATES_BOUNDS -> BOUNDARYS (dg_functions.f90:1808-1828) */
                    oracle_ghost(bcode, s->gamma, cl, a1, a2, ghost);
                    cr = ghost;
                } else {
                    cr = uleft_at(&c, (double *)ul, j, fj, q);
                }
                double clr[NV], crr[NV], h[NV], rh[NV];
                oracle_rotatef(crr, cr, a1, a2);
                oracle_rotatef(clr, cl, a1, a2);
                oracle_hllc(s->gamma, clr, crr, h);
                oracle_rotateb(rh, h, a1, a2);
                for (int v = 0; v < NV; ++v) god[v] = god[v] + (rh[v] * (wts[q] * m->surf[slot]));
            }
            for (int v = 0; v < NV; ++v) r[v] = r[v] + god[v];
        }
        for (int v = 0; v < NV; ++v) rhs[i * NV + v] = r[v];
    }
    return UCNS_OK;
}

static size_t uleft_size(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s) {
    int nqm = s->qp_quad > s->qp_triangle ? s->qp_quad : s->qp_triangle;
    return (size_t)m->ncell * m->nface_max * nqm * NV;
}

int oracle_residual(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const ucns_prestore_v1 *p, const double *u,
                    double *rhs, int64_t *counters) {
    double *ul = (double *)calloc(uleft_size(m, s), sizeof(double));
    if (!ul) return UCNS_ENOMEM;
    int rc = oracle_reconstruct(m, s, p, u, ul, counters);
    if (rc == 0) rc = oracle_fluxes(m, s, p, ul, rhs);
    free(ul);
    return rc;
}

/* CALCULATE_CFL, Euler branch (time_p.f90:46, 67-116) */
int oracle_dt(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const double *u, double cfl, double *dt_out) {
    const double ccfl = cfl / 3.0;
    double dt = 1.0e16;
    const int64_t nc = m->ncell;
#pragma omp parallel for reduction(min : dt) schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        double w[NV];
        cons2prim(s->gamma, u + i * NV, w);
        double agrt = sqrt(w[4] * s->gamma / w[0]);
        double veln = fmax(fmax(fabs(w[1]), fabs(w[2])), fabs(w[3])) + agrt;
        double d = ccfl * ((m->minedge[i]) / (fabs(veln)));
        dt = fmin(dt, d);
    }
    *dt_out = dt;
    return UCNS_OK;
}

/* RUNGE_KUTTA3 (time_p.f90:885-1027) for nsteps steps; u is slot 1 (in/out). */
int oracle_rk3_steps(const ucns_mesh_v1 *m, const ucns_scheme_v1 *s, const ucns_prestore_v1 *p, double *u,
                     int32_t nsteps, double cfl, double dt_fixed, double *t_io) {
    const int64_t nc = m->ncell;
    const size_t n = (size_t)nc * NV;
    double *u2 = (double *)malloc(sizeof(double) * n);
    double *u3 = (double *)malloc(sizeof(double) * n);
    double *rhs = (double *)malloc(sizeof(double) * n);
    double *ul = (double *)calloc(uleft_size(m, s), sizeof(double));
    if (!u2 || !u3 || !rhs || !ul) {
        free(u2); free(u3); free(rhs); free(ul);
        return UCNS_ENOMEM;
    }
    const double to4 = 3.0 / 4.0, oo4 = 1.0 / 4.0, to3 = 2.0 / 3.0, oo3 = 1.0 / 3.0;
    double t = t_io ? *t_io : 0.0;
    int rc = UCNS_OK;
    for (int it = 0; it < nsteps && rc == UCNS_OK; ++it) {
        double dt = dt_fixed;
        if (!(dt_fixed > 0.0)) oracle_dt(m, s, u, cfl, &dt);
        /* stage 1 */
        rc = oracle_reconstruct(m, s, p, u, ul, NULL);
        if (rc) break;
        oracle_fluxes(m, s, p, ul, rhs);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            double oov = 1.0 / m->totvolume[i];
            for (int v = 0; v < NV; ++v) {
                u2[i * NV + v] = u[i * NV + v];
                u[i * NV + v] = u2[i * NV + v] - (dt * (rhs[i * NV + v] * oov));
            }
        }
        /* stage 2 */
        rc = oracle_reconstruct(m, s, p, u, ul, NULL);
        if (rc) break;
        oracle_fluxes(m, s, p, ul, rhs);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            double oov = 1.0 / m->totvolume[i];
            for (int v = 0; v < NV; ++v) {
                u3[i * NV + v] = u[i * NV + v];
                u[i * NV + v] = (to4 * u2[i * NV + v]) + (oo4 * u3[i * NV + v]) -
                                (((oo4)) * ((dt) * ((rhs[i * NV + v]) * (oov))));
            }
        }
        /* stage 3 */
        rc = oracle_reconstruct(m, s, p, u, ul, NULL);
        if (rc) break;
        oracle_fluxes(m, s, p, ul, rhs);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            double oov = 1.0 / m->totvolume[i];
            for (int v = 0; v < NV; ++v)
                u[i * NV + v] = ((oo3) * u2[i * NV + v]) + ((to3) * u[i * NV + v]) -
                                (((to3)) * ((dt) * ((rhs[i * NV + v]) * (oov))));
        }
        t += dt;
    }
    if (t_io) *t_io = t;
    free(u2); free(u3); free(rhs); free(ul);
    return rc;
}

/* WENO weights of one cell and variable from given indicators (the closed-form
 * tests of reconstruct_p.f90:1031-1049 call this through the same helper the
 * reconstruction uses) */
void oracle_weno_weights(int admis, const double *beta, double lwci1, double eps, double *omega) {
    weno_weights(admis, beta, lwci1, eps, omega);
}

/* Barth-Jespersen psi of one face point (reconstruct_p.f90:3439-3453), the
 * helper the MUSCL reconstruction uses */
double oracle_barth(double ui, double umin, double umax, double usol) { return barth_psi(ui, umin, umax, usol); }
