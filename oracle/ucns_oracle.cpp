/*
 * ucns_oracle.cpp -- CPU restatement of UCNS3D's explicit residual and SSP-RK3
 * update, in FP64, following the reference's own loop structure, plus its
 * tangent-linear (forward-mode) derivative.
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
 * Every function on the residual path is a template on the scalar type T.
 * T = double is the restatement above (the same expressions, evaluated in
 * the same order, as the plain-C form it replaces).  T = Dual carries one
 * directional derivative next to each value: the same code then evaluates
 * the tangent-linear model J(u) du exactly (to rounding), with every branch
 * (HLLC wave-speed clamps and sign(S_M), WENO, CHECKSOL/CHECKSOLX, Barth
 * min/max, boundary regimes) decided by the values, as the primal decides it.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/ucns.h"

#define NV 5

/* ---- forward-mode dual numbers ---- */
struct Dual {
    double v, d;
    Dual() = default;
    Dual(double x, double y) : v(x), d(y) {}
    explicit Dual(double x) : v(x), d(0.0) {}
};
static inline double val(double x) { return x; }
static inline double val(const Dual &x) { return x.v; }
static inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
static inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
static inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
static inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
static inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
static inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
static inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
static inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
static inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
static inline Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
static inline Dual sqrt(Dual a) {
    const double r = std::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
static inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
static inline Dual fmin(Dual a, Dual b) { return (b.v < a.v) ? b : a; }
static inline Dual fmax(Dual a, Dual b) { return (b.v > a.v) ? b : a; }
static inline Dual fmin(double a, Dual b) { return (b.v < a) ? b : Dual{a, 0.0}; }
static inline Dual fmin(Dual a, double b) { return (b < a.v) ? Dual{b, 0.0} : a; }
static inline Dual fmax(Dual a, double b) { return (b > a.v) ? Dual{b, 0.0} : a; }
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;

/* CONS2PRIM2 for one state, Euler 3D (flow_operations.f90:447-467) */
template <class T>
static void cons2prim(double gamma, const T *u, T *w) {
    T oo = 1.0 / u[0];
    w[0] = u[0];
    w[1] = u[1] * oo;
    w[2] = u[2] * oo;
    w[3] = u[3] * oo;
    w[4] = (gamma - 1.0) * (u[4] - 0.5 * u[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}

/* ROTATEF (grid_t.f90:4217-4258): rot = TRI * vec */
template <class T>
static void rotatef(T *rot, const T *vec, double a1, double a2) {
    double coa1 = cos(a1), sia1 = sin(a1), coa2 = cos(a2), sia2 = sin(a2);
    rot[0] = vec[0];
    rot[1] = coa1 * sia2 * vec[1] + sia1 * sia2 * vec[2] + coa2 * vec[3];
    rot[2] = coa1 * coa2 * vec[1] + sia1 * coa2 * vec[2] + (-sia2) * vec[3];
    rot[3] = (-sia1) * vec[1] + coa1 * vec[2];
    rot[4] = vec[4];
}

/* ROTATEB (grid_t.f90:4261-4305): vec = INVTRI * rot */
template <class T>
static void rotateb(T *vec, const T *rot, double a1, double a2) {
    double coa1 = cos(a1), sia1 = sin(a1), coa2 = cos(a2), sia2 = sin(a2);
    vec[0] = rot[0];
    vec[1] = coa1 * sia2 * rot[1] + coa1 * coa2 * rot[2] + (-sia1) * rot[3];
    vec[2] = sia1 * sia2 * rot[1] + sia1 * coa2 * rot[2] + coa1 * rot[3];
    vec[3] = coa2 * rot[1] + (-sia2) * rot[2];
    vec[4] = rot[4];
}

/* Free-stream data of the boundary codes (set by oracle_set_boundary), and
 * the tangent of the inflow[5], outflow[5] data for the Dual evaluation. */
static ucns_boundary_v1 g_bc = {UCNS_ABI_VERSION, 0, {1, 0, 0, 0, 2.5}, {1, 0, 0, 0, 2.5}};
static double g_dbc[2 * NV];

template <class T>
static T bc_at(const double *x, int k);
template <>
double bc_at<double>(const double *x, int k) {
    return x[k];
}
template <>
Dual bc_at<Dual>(const double *x, int k) {
    return {x[k], x == g_bc.inflow ? g_dbc[k] : g_dbc[NV + k]};
}

#define TOLSMALL 1.0e-13 /* grid_p.f90:18 */

/* PRIM2CONS of (rho, u, v, w, p) with the reference's energy form
 * rho*(p/((gamma-1) rho)) + rho*(|u|^2/2) (flow_operations.f90:2929-2935) */
template <class T>
static void bc_prim2cons(double gamma, const T *s3, T *cr) {
    cr[0] = s3[0];
    cr[1] = s3[1] * s3[0];
    cr[2] = s3[2] * s3[0];
    cr[3] = s3[3] * s3[0];
    T skins = 0.5 * ((s3[1] * s3[1]) + (s3[2] * s3[2]) + (s3[3] * s3[3]));
    T ikins = s3[4] / ((gamma - 1.0) * (s3[0]));
    cr[4] = (s3[0] * (ikins)) + (s3[0] * skins);
}

/* subsonic inflow: pressure from the outgoing characteristic, the rest from
 * the free stream (flow_operations.f90:2907-2915); s1 free stream, s2 interior */
template <class T>
static void bc_sub_in(double gamma, const T *s1, const T *s2, T sps, double nx, double ny, double nz, T *cr) {
    T s3[NV];
    s3[4] = 0.5 * ((s1[4]) + (s2[4]) - (s2[0] * sps * ((nx * (s1[1] - s2[1])) + (ny * (s1[2] - s2[2])) + (nz * (s1[3] - s2[3])))));
    s3[0] = s1[0] + (s3[4] - s1[4]) / (sps * sps);
    s3[1] = s1[1] - (nx * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s1[2] - (ny * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s1[3] - (nz * (s1[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

/* subsonic outflow: free-stream pressure, the rest from the interior
 * (flow_operations.f90:2995-3010); s1 free stream, s2 interior */
template <class T>
static void bc_sub_out(double gamma, const T *s1, const T *s2, T sps, double nx, double ny, double nz, T *cr) {
    T s3[NV];
    s3[4] = s1[4];
    s3[0] = s2[0] + (s3[4] - s2[4]) / (sps * sps);
    s3[1] = s2[1] + (nx * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s2[2] + (ny * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s2[3] + (nz * (s2[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

/* BOUNDARYS (flow_operations.f90:2860-3375), Euler 3D without a rotating
 * frame: ghost state cr of a boundary face from the face-point state cl */
/* per-face INFLOW() / OUTFLOW() states, [ncell][nf][5] (oracle_set_boundary_faces; null: g_bc's) */
static const double *g_bfin = nullptr, *g_bfout = nullptr;

template <class T>
static void ghost(int code, double gamma, const T *cl, double a1, double a2, T *cr, int64_t slot = -1) {
    const double nx = (cos(a1) * sin(a2)), ny = (sin(a1) * sin(a2)), nz = (cos(a2));
    T s1[NV], s2[NV], fin[NV], fout[NV];
    for (int v = 0; v < NV; ++v) {
        fin[v] = slot >= 0 && g_bfin ? T(g_bfin[slot * NV + v]) : bc_at<T>(g_bc.inflow, v);
        fout[v] = slot >= 0 && g_bfout ? T(g_bfout[slot * NV + v]) : bc_at<T>(g_bc.outflow, v);
    }
    switch (code) {
    case 1: /* inflow (2880-2925) */
        if (g_bc.boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = fin[v];
        } else {
            cons2prim(gamma, fin, s1);
            cons2prim(gamma, cl, s2);
            T sps = sqrt((gamma * s2[4]) / (s2[0]));
            T vel = sqrt(s2[1] * s2[1] + s2[2] * s2[2] + s2[3] * s2[3]);
            if (val(vel / (sps + TOLSMALL)) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = fin[v];
            } else {
                bc_sub_in(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 2: /* outflow (2965-3013) */
        if (g_bc.boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = cl[v];
        } else {
            cons2prim(gamma, fout, s1);
            cons2prim(gamma, cl, s2);
            T sps = sqrt((gamma * s2[4]) / (s2[0]));
            T vel = sqrt(s2[1] * s2[1] + s2[2] * s2[2] + s2[3] * s2[3]);
            if (val(vel / (sps + TOLSMALL)) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                bc_sub_out(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 3: /* symmetry (3086-3131) */
    case 4: /* wall, inviscid branch (3132-3160) */
    {
        T rot[NV], rr[NV];
        rotatef(rot, cl, a1, a2);
        for (int v = 0; v < NV; ++v) rr[v] = rot[v];
        rr[1] = -rot[1];
        rotateb(cr, rr, a1, a2);
        break;
    }
    case 6: /* farfield (3203-3296): direction from the rotated normal momentum */
    {
        T rot[NV];
        rotatef(rot, cl, a1, a2);
        const T vnb = rot[1];
        cons2prim(gamma, cl, s2);
        T sps = sqrt((gamma * s2[4]) / (s2[0]));
        if (val(vnb) <= 0.0) {
            if (val(fabs(vnb)) >= val(sps)) {
                for (int v = 0; v < NV; ++v) cr[v] = fin[v];
            } else {
                cons2prim(gamma, fin, s1);
                bc_sub_in(gamma, s1, s2, sps, nx, ny, nz, cr);
            }
        } else {
            if (val(fabs(vnb)) >= val(sps)) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                cons2prim(gamma, fout, s1);
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
template <class T>
static void hllc(double gamma, const T *cl, const T *cr, T *flux) {
    T wl[NV], wr[NV];
    cons2prim(gamma, cl, wl);
    cons2prim(gamma, cr, wr);
    T rl = wl[0], ul = wl[1], vl = wl[2], zl = wl[3], pl = wl[4], el = cl[4];
    T rr = wr[0], ur = wr[1], vr = wr[2], zr = wr[3], pr = wr[4], er = cr[4];
    T ccl = sqrt(gamma * pl / rl), ccr = sqrt(gamma * pr / rr);
    T srl = sqrt(rl), srr = sqrt(rr);
    T cc2 = sqrt((((ccl * ccl) * srl) + ((ccr * ccr) * srr)) / (srl + srr) +
                 (0.5 * ((srl * srr) / ((srl + srr) * (srl + srr))) * ((ur - ul) * (ur - ul))));
    T uu2 = ((ul * srl) + (ur * srr)) / (srl + srr);
    T sl = fmin(ul - ccl, uu2 - cc2), sr = fmax(ur + ccr, uu2 + cc2);
    sl = fmin(sl, 0.0);
    sr = fmax(sr, 0.0);
    T sm = (pr - pl + (rl * ul * (sl - ul)) - (rr * ur * (sr - ur))) / ((rl * (sl - ul)) - (rr * (sr - ur)));
    T fl[NV] = {rl * ul, (rl * (ul * ul)) + pl, rl * ul * vl, rl * ul * zl, ul * (el + pl)};
    T fr[NV] = {rr * ur, (rr * (ur * ur)) + pr, rr * ur * vr, rr * ur * zr, ur * (er + pr)};
    T mul = rl * ((sl - ul) / (sl - sm));
    T mur = rr * ((sr - ur) / (sr - sm));
    T lastl = (el / rl) + ((sm - ul) * (sm + ((pl) / (rl * (sl - ul)))));
    T lastr = (er / rr) + ((sm - ur) * (sm + ((pr) / (rr * (sr - ur)))));
    T usl[NV] = {mul, mul * sm, mul * vl, mul * zl, mul * lastl};
    T usr[NV] = {mur, mur * sm, mur * vr, mur * zr, mur * lastr};
    double s = copysign(1.0, val(sm));
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

struct ctx_t {
    const ucns_mesh_v2 *m;
    const ucns_scheme_v2 *s;
    const ucns_prestore_v1 *p;
    int nf, nqm, ideg, nnb, imax, ts, base;
};

static int nqp_of(const ctx_t *c, int64_t i, int f) {
    return c->m->types_faces[i * c->nf + f] == 5 ? c->s->qp_quad : c->s->qp_triangle;
}

/* Form each face value as U_i + sum_s omega_s res_s instead of the
 * reference's sum_s omega_s (U_i + res_s) (oracle_set_uleft_form): equal in
 * exact arithmetic since the weights sum to 1, rounded differently.  Test
 * hook for measuring the reference arithmetic's own rounding sensitivity. */
static int g_uleft_mean_first = 0;

template <class T>
static T *uleft_at(const ctx_t *c, T *ul, int64_t i, int f, int q) {
    return ul + (((size_t)i * c->nf + f) * c->nqm + q) * NV;
}

/* omega_s = lambda_s / (eps + beta_s)^4, normalised; lambda = lwci1 for the
 * central stencil, 1 for the directional ones (reconstruct_p.f90:1031-1049) */
template <class T>
static void weno_weights(int admis, const T *beta, double lwci1, double eps, T *omega) {
    T wt[16], sum = T(0.0);
    for (int ll = 0; ll < admis; ++ll) {
        double lam = (ll == 0) ? lwci1 : 1.0;
        T x = eps + beta[ll];
        T x2 = x * x;
        wt[ll] = lam / (x2 * x2);
    }
    for (int ll = 0; ll < admis; ++ll) sum = sum + wt[ll];
    for (int ll = 0; ll < admis; ++ll) omega[ll] = wt[ll] / sum;
}

/* CWENO data of the scheme's ees = 5 (oracle_set_cweno; NULL: none) */
static const ucns_cweno_v1 *g_cw = nullptr;

/* CWENO weights (reconstruct_p.f90:955-958, 1010-1049): lambda_1 = 1 - 1/LWCI1, lambda_l =
 * (1 - lambda_1)/(ADMIS-1); WENO: lambda/(eps+beta)^4; WENO-Z: lambda (1 + (tau/(eps+beta))^4) with
 * tau = sum_l |beta_1 - beta_l| / (ADMIS-1); normalised */
template <class T>
static void cweno_weights(int admis, const T *beta, double lwci1, double eps, int wenoz, T *omega) {
    const double lam1 = 1.0 - (1.0 / lwci1);
    const double laml = admis > 1 ? (1.0 - lam1) / (admis - 1) : 0.0;
    T wt[8], sum = T(0.0), tau = T(0.0);
    if (wenoz == 1 && admis > 1) {
        for (int ll = 0; ll < admis; ++ll) tau = tau + fabs(beta[0] - beta[ll]);
        tau = tau / (double)(admis - 1);
    }
    for (int ll = 0; ll < admis; ++ll) {
        const double lam = ll == 0 ? lam1 : laml;
        const T x = eps + beta[ll];
        if (wenoz == 1) {
            const T r = tau / x;
            const T r2 = r * r;
            wt[ll] = lam * (1.0 + r2 * r2);
        } else {
            const T x2 = x * x;
            wt[ll] = lam / (x2 * x2);
        }
    }
    for (int ll = 0; ll < admis; ++ll) sum = sum + wt[ll];
    for (int ll = 0; ll < admis; ++ll) omega[ll] = wt[ll] / sum;
}

/* per-thread scratch of the cell routines */
template <class T>
struct Work {
    std::vector<T> grad, mat1, bg;
    std::vector<double> phi;
    Work(const ctx_t &c)
        : grad((size_t)c.ts * c.ideg * NV), mat1((size_t)c.imax * NV), bg(c.ideg), phi(c.ideg) {}
};

/* one cell of ARBITRARY_ORDER with EES=5 (CP_RECONSTRUCTION, reconstruct_p.f90:955-1014, 1062-1199):
 * central LSQ of degree IORDER, sector LSQ of degree IORDER2 over the first numneighbours2 members
 * (INVMAT_STENCILTC, gradients.f90:362), the zero polynomial (G_1 - sum_l lambda_l G_l) / lambda_1
 * with lambda_1 = 1 - 1/LWCI1 and lambda_l = (1 - lambda_1)/(ADMIS-1), its indicator with INDICATOR
 * and the sectors' with INDICATORC, WENO or WENO-Z weights (power 4), and the face values
 * sum_l omega_l (U + res_l), res_1 on the high-order basis, res_l on the low-order one (VALUEC). */
template <class T>
static void cweno_cell(const ctx_t *c, const T *u, int64_t i, T *ul, Work<T> &w) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb, ts = c->ts;
    const int id2 = g_cw->idegfree2, im2 = g_cw->inumneighbours2 - 1;
    const int admis = c->p->admis[i];
    const T *ui = u + i * NV;
    T *mat1 = w.mat1.data();
    double *phi = w.phi.data();
    std::vector<T> g0((size_t)ideg * NV), gc((size_t)ts * id2 * NV), zp((size_t)ideg * NV);
    T omega[8][NV];
    /* LSQ: the central stencil (degree IORDER) */
    {
        const int32_t *ih = c->p->ihexl + (size_t)i * ts * nnb;
        const T *s1 = u + (size_t)(ih[0] - c->base) * NV;
        for (int iq = 0; iq < imax; ++iq) {
            const T *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
            for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
        }
        const double *P = c->p->invmat + (size_t)i * ts * imax * ideg;
        for (int k = 0; k < ideg; ++k)
            for (int v = 0; v < NV; ++v) {
                T acc = T(0.0);
                for (int iq = 0; iq < imax; ++iq) acc = acc + P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
                g0[(size_t)k * NV + v] = acc;
            }
    }
    /* the sectors (degree IORDER2, first numneighbours2 members) */
    for (int ll = 1; ll < admis; ++ll) {
        const int32_t *ih = c->p->ihexl + ((size_t)i * ts + ll) * nnb;
        const T *s1 = u + (size_t)(ih[0] - c->base) * NV;
        for (int iq = 0; iq < im2; ++iq) {
            const T *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
            for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
        }
        const double *P = g_cw->invmat_c + ((size_t)i * ts + ll) * im2 * id2;
        for (int k = 0; k < id2; ++k)
            for (int v = 0; v < NV; ++v) {
                T acc = T(0.0);
                for (int iq = 0; iq < im2; ++iq) acc = acc + P[(size_t)iq * id2 + k] * mat1[iq * NV + v];
                gc[((size_t)ll * id2 + k) * NV + v] = acc;
            }
    }
    const double lwc = c->s->lwci1;
    const double lam1 = 1.0 - (1.0 / lwc);
    const double laml = admis > 1 ? (1.0 - lam1) / (admis - 1) : 0.0;
    const double *B = c->p->indicator + (size_t)i * ideg * ideg;
    const double *Bc = g_cw->indicator_c + (size_t)i * id2 * id2;
    const double eps = c->s->weno_eps;
    for (int v = 0; v < NV; ++v) {
        /* GRAD3AL = sum of the sector polynomials, then the zero polynomial GRAD1AL */
        std::vector<T> g3((size_t)ideg, T(0.0));
        for (int ll = 1; ll < admis; ++ll)
            for (int k = 0; k < id2; ++k) g3[k] = g3[k] + (laml * gc[((size_t)ll * id2 + k) * NV + v]);
        for (int k = 0; k < ideg; ++k) zp[(size_t)k * NV + v] = (1.0 / lam1) * (g0[(size_t)k * NV + v] - g3[k]);
        T beta[8], wt[8];
        for (int ll = 0; ll < admis; ++ll) {
            const int n = ll == 0 ? ideg : id2;
            const double *BB = ll == 0 ? B : Bc;
            T *bg = w.bg.data();
            for (int a = 0; a < n; ++a) {
                T acc = T(0.0);
                for (int b = 0; b < n; ++b) {
                    const T gb = ll == 0 ? zp[(size_t)b * NV + v] : gc[((size_t)ll * id2 + b) * NV + v];
                    acc = acc + BB[(size_t)b * n + a] * gb;
                }
                bg[a] = acc;
            }
            T dot = T(0.0);
            for (int a = 0; a < n; ++a) {
                const T ga = ll == 0 ? zp[(size_t)a * NV + v] : gc[((size_t)ll * id2 + a) * NV + v];
                dot = dot + ga * bg[a];
            }
            beta[ll] = dot;
        }
        cweno_weights(admis, beta, lwc, eps, c->s->wenoz, wt);
        for (int ll = 0; ll < admis; ++ll) omega[ll][v] = wt[ll];
    }
    /* extrapolation, stencil by stencil */
    const double *integ = c->p->integ_basis + (size_t)i * ideg;
    const double *integc = g_cw->integ_basis_c + (size_t)i * id2;
    const double vol11 = c->p->volume11[i];
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            T *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) o[v] = T(0.0);
        }
    for (int ll = 0; ll < admis; ++ll)
        for (int f = 0; f < c->m->ifca[i]; ++f)
            for (int q = 0; q < nqp_of(c, i, f); ++q) {
                const double *x = c->p->qpoints + (((size_t)i * c->nf + f) * c->nqm + q) * 3;
                const int n = ll == 0 ? ideg : id2;
                basis_rec(x, n, ll == 0 ? integ : integc, vol11, phi);
                T *o = uleft_at(c, ul, i, f, q);
                for (int v = 0; v < NV; ++v) {
                    T res = T(0.0);
                    for (int k = 0; k < n; ++k)
                        res = res + phi[k] * (ll == 0 ? zp[(size_t)k * NV + v] : gc[((size_t)ll * id2 + k) * NV + v]);
                    if (g_uleft_mean_first)
                        o[v] = o[v] + res * omega[ll][v];
                    else
                        o[v] = o[v] + (ui[v] + res) * omega[ll][v];
                }
            }
    if (g_uleft_mean_first)
        for (int f = 0; f < c->m->ifca[i]; ++f)
            for (int q = 0; q < nqp_of(c, i, f); ++q) {
                T *o = uleft_at(c, ul, i, f, q);
                for (int v = 0; v < NV; ++v) o[v] = ui[v] + o[v];
            }
}

/* one cell of ARBITRARY_ORDER: LEAST_SQUARES + WENOWEIGHTS (CP_RECONSTRUCTION) */
template <class T>
static void weno_cell(const ctx_t *c, const T *u, int64_t i, T *ul, Work<T> &w) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb, ts = c->ts;
    const int admis = c->p->admis[i];
    const T *ui = u + i * NV;
    T *grad = w.grad.data(), *mat1 = w.mat1.data(), *bg = w.bg.data();
    double *phi = w.phi.data();
    T omega[8][NV];
    /* LSQ (COMPUTE_GRADIENTS_MEAN_LSQ, LOCAL=1 branch) */
    for (int ll = 0; ll < admis; ++ll) {
        const int32_t *ih = c->p->ihexl + ((size_t)i * ts + ll) * nnb;
        const T *s1 = u + (size_t)(ih[0] - c->base) * NV;
        for (int iq = 0; iq < imax; ++iq) {
            const T *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
            for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
        }
        const double *P = c->p->invmat + ((size_t)i * ts + ll) * imax * ideg;
        for (int k = 0; k < ideg; ++k)
            for (int v = 0; v < NV; ++v) {
                T acc = T(0.0);
                for (int iq = 0; iq < imax; ++iq) acc = acc + P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
                grad[((size_t)ll * ideg + k) * NV + v] = acc;
            }
    }
    /* smoothness indicators and weights, per variable */
    const double *B = c->p->indicator + (size_t)i * ideg * ideg;
    const double eps = c->s->weno_eps;
    for (int v = 0; v < NV; ++v) {
        T beta[8], wt[8];
        for (int ll = 0; ll < admis; ++ll) {
            for (int a = 0; a < ideg; ++a) {
                T acc = T(0.0);
                for (int b = 0; b < ideg; ++b) acc = acc + B[(size_t)b * ideg + a] * grad[((size_t)ll * ideg + b) * NV + v];
                bg[a] = acc;
            }
            T dot = T(0.0);
            for (int a = 0; a < ideg; ++a) dot = dot + grad[((size_t)ll * ideg + a) * NV + v] * bg[a];
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
            T *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) o[v] = T(0.0);
        }
    for (int ll = 0; ll < admis; ++ll)
        for (int f = 0; f < c->m->ifca[i]; ++f)
            for (int q = 0; q < nqp_of(c, i, f); ++q) {
                const double *x = c->p->qpoints + (((size_t)i * c->nf + f) * c->nqm + q) * 3;
                basis_rec(x, ideg, integ, vol11, phi);
                T *o = uleft_at(c, ul, i, f, q);
                for (int v = 0; v < NV; ++v) {
                    T res = T(0.0);
                    for (int k = 0; k < ideg; ++k) res = res + phi[k] * grad[((size_t)ll * ideg + k) * NV + v];
                    if (g_uleft_mean_first)
                        o[v] = o[v] + res * omega[ll][v];
                    else
                        o[v] = o[v] + (ui[v] + res) * omega[ll][v];
                }
            }
    if (g_uleft_mean_first)
        for (int f = 0; f < c->m->ifca[i]; ++f)
            for (int q = 0; q < nqp_of(c, i, f); ++q) {
                T *o = uleft_at(c, ul, i, f, q);
                for (int v = 0; v < NV; ++v) o[v] = ui[v] + o[v];
            }
}

/* Eigenvectors of the x-split Euler flux Jacobian at the face's averaged
 * state (COMPUTE_EIGENVECTORS, flow_operations.f90:3934-3995): density,
 * velocity and total energy averaged arithmetically between the two rotated
 * states rl, rr, pressure and speed of sound from those; R columns (u-a, u,
 * v-shear, w-shear, u+a), L = R^-1 in the textbook form (Toro, Riemann
 * Solvers, 3rd ed., sec. 3.2.2) with b = (gamma-1)/a^2.  R and L are [5][5]
 * row-major. */
template <class T>
static void char_eigen(double gamma, const T *rl, const T *rr, T *R, T *L) {
    const T rs = 0.5 * (rl[0] + rr[0]);
    const T us = 0.5 * (rl[1] / rl[0] + rr[1] / rr[0]);
    const T vs = 0.5 * (rl[2] / rl[0] + rr[2] / rr[0]);
    const T ws = 0.5 * (rl[3] / rl[0] + rr[3] / rr[0]);
    const T es = 0.5 * (rl[4] + rr[4]);
    const T q2 = us * us + vs * vs + ws * ws;
    const T ps = (gamma - 1.0) * (es - 0.5 * rs * q2);
    const T as = sqrt(gamma * ps / rs);
    const T hs = 0.5 * q2 + as * as / (gamma - 1.0);
    const T Rm[5][5] = {{T(1.0), T(1.0), T(0.0), T(0.0), T(1.0)},
                        {us - as, us, T(0.0), T(0.0), us + as},
                        {vs, vs, T(1.0), T(0.0), vs},
                        {ws, ws, T(0.0), T(1.0), ws},
                        {hs - us * as, 0.5 * q2, vs, ws, hs + us * as}};
    const T b = (gamma - 1.0) / (as * as), bq = 0.5 * b * q2, oa = 1.0 / as;
    const T Lm[5][5] = {{0.5 * (bq + us * oa), -0.5 * (b * us + oa), -0.5 * b * vs, -0.5 * b * ws, 0.5 * b},
                        {1.0 - bq, b * us, b * vs, b * ws, -b},
                        {-vs, T(0.0), T(1.0), T(0.0), T(0.0)},
                        {-ws, T(0.0), T(0.0), T(1.0), T(0.0)},
                        {0.5 * (bq - us * oa), -0.5 * (b * us - oa), -0.5 * b * vs, -0.5 * b * ws, 0.5 * b}};
    for (int a = 0; a < 5; ++a)
        for (int c = 0; c < 5; ++c) {
            R[a * 5 + c] = Rm[a][c];
            L[a * 5 + c] = Lm[a][c];
        }
}

/* Skip the troubled-cell logic (CHECKSOL, MUSCL, CHECKSOLX) in reconstruct
 * (oracle_set_raw): the WENO face values alone.  Test hook. */
static int g_raw = 0;

/* One cell of CHARACTERISTIC_RECONSTRUCTION (reconstruct_p.f90:497-758, EES=0):
 * per face, the neighbour's cell average (WENO_NEIGHBOUR, 761-917: the
 * boundary ghost of the cell's own average on boundary faces), both rotated
 * into the face frame, the eigenvectors of their average (char_eigen), every
 * stencil's coefficients and the cell average projected with L
 * (compute_gradcharv_smoothindicator, 1690-1818), indicators per
 * characteristic variable with INDICATOR (diag_At_B_A, 1608-1688), WENO
 * weights per characteristic variable, the weighted characteristic
 * coefficients projected back with R, and that face's values
 * FINDW_0 + sum_k phi_k FINDW_k. */
template <class T>
static void char_cell(const ctx_t *c, const T *u, int64_t i, T *ul, Work<T> &w) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb, ts = c->ts;
    const int admis = c->p->admis[i];
    const T *ui = u + i * NV;
    T *grad = w.grad.data(), *mat1 = w.mat1.data();
    double *phi = w.phi.data();
    const double eps = c->s->weno_eps;
    /* LSQ (COMPUTE_GRADIENTS_MEAN_LSQ) */
    for (int ll = 0; ll < admis; ++ll) {
        const int32_t *ih = c->p->ihexl + ((size_t)i * ts + ll) * nnb;
        const T *s1 = u + (size_t)(ih[0] - c->base) * NV;
        for (int iq = 0; iq < imax; ++iq) {
            const T *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
            for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
        }
        const double *P = c->p->invmat + ((size_t)i * ts + ll) * imax * ideg;
        for (int k = 0; k < ideg; ++k)
            for (int v = 0; v < NV; ++v) {
                T acc = T(0.0);
                for (int iq = 0; iq < imax; ++iq) acc = acc + P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
                grad[((size_t)ll * ideg + k) * NV + v] = acc;
            }
    }
    const double *B = c->p->indicator + (size_t)i * ideg * ideg;
    const double *integ = c->p->integ_basis + (size_t)i * ideg;
    const double vol11 = c->p->volume11[i];
    std::vector<T> W((size_t)admis * (ideg + 1) * NV), F((size_t)(ideg + 1) * NV), ba(ideg);
    for (int f = 0; f < c->m->ifca[i]; ++f) {
        const int64_t slot = i * c->nf + f;
        const double a1 = c->m->faceanglex[slot], a2 = c->m->faceangley[slot];
        const int bcode = c->m->bcode[slot];
        const int bnd = bcode != 0 && bcode != 5 && bcode != 50;
        T ur[NV], rl[NV], rr[NV], R[25], L[25];
        if (bnd)
            ghost(bcode, c->s->gamma, ui, a1, a2, ur, slot);
        else
            for (int v = 0; v < NV; ++v) ur[v] = u[(size_t)(c->m->ineigh[slot] - c->base) * NV + v];
        rotatef(rl, ui, a1, a2);
        rotatef(rr, ur, a1, a2);
        char_eigen(c->s->gamma, rl, rr, R, L);
        /* characteristic coefficients W[ll][k][ch], k = 0 the cell average */
        for (int ll = 0; ll < admis; ++ll)
            for (int k = 0; k <= ideg; ++k)
                for (int ch = 0; ch < NV; ++ch) {
                    T acc = T(0.0);
                    for (int v = 0; v < NV; ++v)
                        acc = acc + (k == 0 ? ui[v] : grad[((size_t)ll * ideg + k - 1) * NV + v]) * L[ch * 5 + v];
                    W[((size_t)ll * (ideg + 1) + k) * NV + ch] = acc;
                }
        /* indicators and weights per characteristic variable */
        T omega[8][NV];
        for (int ch = 0; ch < NV; ++ch) {
            T beta[8], wt[8];
            for (int ll = 0; ll < admis; ++ll) {
                for (int a = 0; a < ideg; ++a) {
                    T acc = T(0.0);
                    for (int b = 0; b < ideg; ++b)
                        acc = acc + B[(size_t)a * ideg + b] * W[((size_t)ll * (ideg + 1) + b + 1) * NV + ch];
                    ba[a] = acc;
                }
                T dot = T(0.0);
                for (int a = 0; a < ideg; ++a) dot = dot + W[((size_t)ll * (ideg + 1) + a + 1) * NV + ch] * ba[a];
                beta[ll] = dot;
            }
            weno_weights(admis, beta, c->s->lwci1, eps, wt);
            for (int ll = 0; ll < admis; ++ll) omega[ll][ch] = wt[ll];
        }
        /* weighted characteristic coefficients, back to conserved variables */
        for (int k = 0; k <= ideg; ++k) {
            T lim[NV];
            for (int ch = 0; ch < NV; ++ch) {
                T acc = T(0.0);
                for (int ll = 0; ll < admis; ++ll) acc = acc + W[((size_t)ll * (ideg + 1) + k) * NV + ch] * omega[ll][ch];
                lim[ch] = acc;
            }
            for (int v = 0; v < NV; ++v) {
                T acc = T(0.0);
                for (int ch = 0; ch < NV; ++ch) acc = acc + R[v * 5 + ch] * lim[ch];
                F[(size_t)k * NV + v] = acc;
            }
        }
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            const double *x = c->p->qpoints + (((size_t)i * c->nf + f) * c->nqm + q) * 3;
            basis_rec(x, ideg, integ, vol11, phi);
            T *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) {
                T res = T(0.0);
                for (int k = 0; k < ideg; ++k) res = res + phi[k] * F[(size_t)(k + 1) * NV + v];
                o[v] = (g_uleft_mean_first ? ui[v] : F[v]) + res;
            }
        }
    }
}

/* CHECKSOL / CHECKSOLX face-point test (reconstruct_p.f90:3183-3196, 3294-3310) */
template <class T>
static int trouble_test(const ctx_t *c, const T *u, const T *ul, int64_t i, double jump) {
    T wi[NV], wq[NV];
    cons2prim(c->s->gamma, u + i * NV, wi);
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            cons2prim(c->s->gamma, uleft_at(c, (T *)ul, i, f, q), wq);
            if (fabs(val(wq[0]) - val(wi[0])) >= jump * val(wi[0])) return 1;
            if (fabs(val(wq[4]) - val(wi[4])) >= jump * val(wi[4])) return 1;
            if (val(wq[4]) < 0.0 || val(wq[0]) < 0.0) return 1;
        }
    return 0;
}

/* Barth-Jespersen psi of one face point value usol (reconstruct_p.f90:3439-3453) */
template <class T>
static T barth_psi(T ui, T umin, T umax, T usol) {
    T d2 = usol - ui;
    if (fabs(val(d2)) <= 0.0) return T(1.0);
    if (val(d2) > 0.0) return fmin(1.0, (umax - ui) / d2);
    return fmin(1.0, (umin - ui) / d2);
}

/* MUSCL on one cell: FIND_BOUNDS (strict) + unlimited central polynomial +
 * Barth-Jespersen (reconstruct_p.f90:1821-2039, 2044-2280, 3421-3453) */
template <class T>
static void muscl_cell(const ctx_t *c, const T *u, int64_t i, T *ul, Work<T> &w) {
    const int ideg = c->ideg, nf = c->nf;
    const T *grad = w.grad.data();
    const T *ui = u + i * NV;
    T umin[NV], umax[NV], slope[NV];
    for (int v = 0; v < NV; ++v) umin[v] = umax[v] = ui[v];
    for (int f = 0; f < c->m->ifca[i]; ++f) {
        /* strict bounds (EXTENDED_BOUNDS=0, FIND_BOUNDS src/reconstruct_p.f90:1847-1885): every face
         * neighbour of an interior cell; of a cell with boundary faces, the interior neighbours and
         * the periodic ones with code 5 (1871), not those of code 50 nor the other boundary codes */
        int bc = c->m->bcode[i * nf + f];
        if (bc != 0 && bc != 5) continue;
        const T *un = u + (size_t)(c->m->ineigh[i * nf + f] - c->base) * NV;
        for (int v = 0; v < NV; ++v) {
            umin[v] = fmin(umin[v], un[v]);
            umax[v] = fmax(umax[v], un[v]);
        }
    }
    double *phi = w.phi.data();
    const double *integ = c->p->integ_basis + (size_t)i * ideg;
    const double vol11 = c->p->volume11[i];
    for (int v = 0; v < NV; ++v) slope[v] = T(1.0e16); /* TOLBIG */
    /* pass 1: USOL and the limiter */
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            const double *x = c->p->qpoints + (((size_t)i * nf + f) * c->nqm + q) * 3;
            basis_rec(x, ideg, integ, vol11, phi);
            for (int v = 0; v < NV; ++v) {
                T res = T(0.0);
                for (int k = 0; k < ideg; ++k) res = res + phi[k] * grad[(size_t)k * NV + v];
                T usol = ui[v] + res;
                slope[v] = fmin(slope[v], barth_psi(ui[v], umin[v], umax[v], usol));
            }
        }
    /* pass 2: ULEFT = U_i + (USOL - U_i) * SLOPE */
    for (int f = 0; f < c->m->ifca[i]; ++f)
        for (int q = 0; q < nqp_of(c, i, f); ++q) {
            const double *x = c->p->qpoints + (((size_t)i * nf + f) * c->nqm + q) * 3;
            basis_rec(x, ideg, integ, vol11, phi);
            T *o = uleft_at(c, ul, i, f, q);
            for (int v = 0; v < NV; ++v) {
                T res = T(0.0);
                for (int k = 0; k < ideg; ++k) res = res + phi[k] * grad[(size_t)k * NV + v];
                T usol = ui[v] + res;
                o[v] = 0.0 + (ui[v] + ((usol - ui[v]) * slope[v]));
            }
        }
}

template <class T>
static void lsq_central(const ctx_t *c, const T *u, int64_t i, Work<T> &w) {
    const int ideg = c->ideg, imax = c->imax, nnb = c->nnb;
    const int32_t *ih = c->p->ihexl + (size_t)i * c->ts * nnb;
    const T *s1 = u + (size_t)(ih[0] - c->base) * NV;
    T *mat1 = w.mat1.data(), *grad = w.grad.data();
    for (int iq = 0; iq < imax; ++iq) {
        const T *s2 = u + (size_t)(ih[iq + 1] - c->base) * NV;
        for (int v = 0; v < NV; ++v) mat1[iq * NV + v] = s2[v] - s1[v];
    }
    const double *P = c->p->invmat + (size_t)i * c->ts * imax * ideg;
    for (int k = 0; k < ideg; ++k)
        for (int v = 0; v < NV; ++v) {
            T acc = T(0.0);
            for (int iq = 0; iq < imax; ++iq) acc = acc + P[(size_t)iq * ideg + k] * mat1[iq * NV + v];
            grad[(size_t)k * NV + v] = acc;
        }
}

static void init_ctx(ctx_t *c, const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p) {
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
template <class T>
static int reconstruct(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const T *u, T *ul,
                       int64_t *counters) {
    ctx_t c;
    init_ctx(&c, m, s, p);
    const int64_t nc = m->ncell;
    int64_t n_red = 0, n_first = 0;
    std::vector<int> reduce((size_t)nc, 0);
#pragma omp parallel
    {
        Work<T> w(c);
        /* WENOWEIGHTS + CHECKSOL, or nothing for MUSCL (IWENO=-1) */
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            if (s->iweno == 1) {
                if (s->wenwrt == 2)
                    char_cell(&c, u, i, ul, w);
                else if (s->ees == 5)
                    cweno_cell(&c, u, i, ul, w);
                else
                    weno_cell(&c, u, i, ul, w);
                int red = 0;
                if (g_raw) red = 0;
                else if (p->full[i] == 0) red = 1;
                else red = trouble_test(&c, u, ul, i, 0.85);
                reduce[i] = red;
            } else {
                reduce[i] = 0;
            }
        }
        /* MUSCL for (TROUBLED & REDUCE) | (TROUBLED & ~FULL); TROUBLED=1 for every FV cell */
#pragma omp for schedule(static) reduction(+ : n_red)
        for (int64_t i = 0; i < nc; ++i) {
            if (g_raw) continue;
            if (reduce[i] == 1 || p->full[i] == 0) {
                lsq_central(&c, u, i, w);
                muscl_cell(&c, u, i, ul, w);
                n_red += 1;
            }
        }
        /* CHECKSOLX: first order where the limited values still jump */
#pragma omp for schedule(static) reduction(+ : n_first)
        for (int64_t i = 0; i < nc; ++i) {
            if (!g_raw && trouble_test(&c, u, ul, i, 0.95)) {
                for (int f = 0; f < m->ifca[i]; ++f)
                    for (int q = 0; q < c.nqm; ++q) {
                        T *o = uleft_at(&c, ul, i, f, q);
                        for (int v = 0; v < NV; ++v) o[v] = u[i * NV + v];
                    }
                n_first += 1;
            }
        }
    }
    if (counters) {
        counters[0] = n_red;
        counters[1] = n_first;
    }
    return UCNS_OK;
}

/* Sum each cell's faces in reverse slot order (oracle_set_face_order): the
 * same residual in exact arithmetic with another rounding of each cell's
 * face sum, for measuring how far the reference's own arithmetic moves a
 * result under a reordering of one sum. */
static int g_faces_reversed = 0;
/* Solve each interior face once, from its owner (the lower cell id; equal
 * ids: the lower slot), and give the neighbour the negated flux
 * (oracle_set_face_once), instead of solving it from both sides with each
 * side's own face angles and area as the reference does: equal in exact
 * arithmetic, rounded differently. */
static int g_face_once = 0;

/* CALCULATE_FLUXESHI_CONVECTIVE, cell based, every face solved from both sides */
template <class T>
static int fluxes(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const T *ul, T *rhs) {
    ctx_t c;
    init_ctx(&c, m, s, p);
    const int64_t nc = m->ncell;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        T r[NV] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
        for (int ff = 0; ff < m->ifca[i]; ++ff) {
            const int f = g_faces_reversed ? m->ifca[i] - 1 - ff : ff;
            const int64_t slot = i * c.nf + f;
            const double a1 = m->faceanglex[slot], a2 = m->faceangley[slot];
            const int iqp = nqp_of(&c, i, f);
            const double *wts = m->types_faces[slot] == 5 ? s->weights_quad : s->weights_tri;
            const int bcode = m->bcode[slot];
            const int bnd = bcode != 0 && bcode != 5 && bcode != 50;
            const int64_t j = bnd ? 0 : m->ineigh[slot] - c.base;
            const int fj = bnd ? 0 : m->ineighn[slot] - c.base;
            T god[NV] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
            if (g_face_once && !bnd && !(i < j || (i == j && f < fj))) {
                /* the owner's flux through this face, negated */
                const int64_t so = j * c.nf + fj;
                const double b1 = m->faceanglex[so], b2 = m->faceangley[so];
                const double *wo = m->types_faces[so] == 5 ? s->weights_quad : s->weights_tri;
                T go[NV] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
                for (int q = 0; q < iqp; ++q) {
                    T clr[NV], crr[NV], h[NV], rh[NV];
                    rotatef(crr, uleft_at(&c, (T *)ul, i, f, q), b1, b2);
                    rotatef(clr, uleft_at(&c, (T *)ul, j, fj, q), b1, b2);
                    hllc(s->gamma, clr, crr, h);
                    rotateb(rh, h, b1, b2);
                    for (int v = 0; v < NV; ++v) go[v] = go[v] + (rh[v] * (wo[q] * m->surf[so]));
                }
                for (int v = 0; v < NV; ++v) r[v] = r[v] - go[v];
                continue;
            }
            for (int q = 0; q < iqp; ++q) {
                const T *cl = uleft_at(&c, (T *)ul, i, f, q);
                T gst[NV];
                const T *cr;
                if (bnd) { /* GET_STATES_BOUNDS -> BOUNDARYS (dg_functions.f90:1808-1828) */
                    ghost(bcode, s->gamma, cl, a1, a2, gst, slot);
                    cr = gst;
                } else {
                    cr = uleft_at(&c, (T *)ul, j, fj, q);
                }
                T clr[NV], crr[NV], h[NV], rh[NV];
                rotatef(crr, cr, a1, a2);
                rotatef(clr, cl, a1, a2);
                hllc(s->gamma, clr, crr, h);
                rotateb(rh, h, a1, a2);
                for (int v = 0; v < NV; ++v) god[v] = god[v] + (rh[v] * (wts[q] * m->surf[slot]));
            }
            for (int v = 0; v < NV; ++v) r[v] = r[v] + god[v];
        }
        for (int v = 0; v < NV; ++v) rhs[i * NV + v] = r[v];
    }
    return UCNS_OK;
}

static size_t uleft_size(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s) {
    int nqm = s->qp_quad > s->qp_triangle ? s->qp_quad : s->qp_triangle;
    return (size_t)m->ncell * m->nface_max * nqm * NV;
}

template <class T>
static int residual(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const T *u, T *rhs,
                    int64_t *counters) {
    std::vector<T> ul(uleft_size(m, s), T(0.0));
    int rc = reconstruct(m, s, p, u, ul.data(), counters);
    if (rc == 0) rc = fluxes(m, s, p, ul.data(), rhs);
    return rc;
}

/* CALCULATE_CFL, Euler branch (time_p.f90:46, 67-116) */
static void cfl_dt(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const double *u, double cfl, double *dt_out) {
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
}

static void values(const double *u, double *out, size_t n) { std::memcpy(out, u, n * sizeof(double)); }
static void values(const Dual *u, double *out, size_t n) {
    for (size_t e = 0; e < n; ++e) out[e] = u[e].v;
}

/* RUNGE_KUTTA3 (time_p.f90:885-1027) for nsteps steps; u is slot 1 (in/out).
 * dt_fixed <= 0 takes dt from the CFL condition of each step's state (its
 * derivative is not carried: the tangent runs use dt_fixed). */
template <class T>
static int rk3_steps(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, T *u, int32_t nsteps,
                     double cfl, double dt_fixed, double *t_io) {
    const int64_t nc = m->ncell;
    const size_t n = (size_t)nc * NV;
    std::vector<T> u2(n), u3(n), rhs(n), ul(uleft_size(m, s), T(0.0));
    std::vector<double> uv(dt_fixed > 0.0 ? 0 : n);
    const double to4 = 3.0 / 4.0, oo4 = 1.0 / 4.0, to3 = 2.0 / 3.0, oo3 = 1.0 / 3.0;
    double t = t_io ? *t_io : 0.0;
    int rc = UCNS_OK;
    for (int it = 0; it < nsteps && rc == UCNS_OK; ++it) {
        double dt = dt_fixed;
        if (!(dt_fixed > 0.0)) {
            values(u, uv.data(), n);
            cfl_dt(m, s, uv.data(), cfl, &dt);
        }
        /* stage 1 */
        rc = reconstruct(m, s, p, u, ul.data(), (int64_t *)nullptr);
        if (rc) break;
        fluxes(m, s, p, ul.data(), rhs.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            double oov = 1.0 / m->totvolume[i];
            for (int v = 0; v < NV; ++v) {
                u2[i * NV + v] = u[i * NV + v];
                u[i * NV + v] = u2[i * NV + v] - (dt * (rhs[i * NV + v] * oov));
            }
        }
        /* stage 2 */
        rc = reconstruct(m, s, p, u, ul.data(), (int64_t *)nullptr);
        if (rc) break;
        fluxes(m, s, p, ul.data(), rhs.data());
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
        rc = reconstruct(m, s, p, u, ul.data(), (int64_t *)nullptr);
        if (rc) break;
        fluxes(m, s, p, ul.data(), rhs.data());
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
    return rc;
}

static std::vector<Dual> make_dual(const double *u, const double *du, size_t n) {
    std::vector<Dual> x(n);
    for (size_t e = 0; e < n; ++e) x[e] = {u[e], du ? du[e] : 0.0};
    return x;
}

static void set_dbc(const double *dbc) {
    for (int k = 0; k < 2 * NV; ++k) g_dbc[k] = dbc ? dbc[k] : 0.0;
}

extern "C" {

void oracle_rotatef(double *rot, const double *vec, double a1, double a2) { rotatef(rot, vec, a1, a2); }
void oracle_rotateb(double *vec, const double *rot, double a1, double a2) { rotateb(vec, rot, a1, a2); }
void oracle_set_boundary(const ucns_boundary_v1 *bc) { g_bc = *bc; }
void oracle_set_face_order(int reversed) { g_faces_reversed = reversed ? 1 : 0; }
void oracle_set_uleft_form(int mean_first) { g_uleft_mean_first = mean_first ? 1 : 0; }
void oracle_set_face_once(int once) { g_face_once = once ? 1 : 0; }
/* the CWENO data the scheme's ees = 5 reads (NULL: none); the struct and its
 * arrays must outlive the calls that use them */
void oracle_set_cweno(const ucns_cweno_v1 *cw) { g_cw = cw; }
void oracle_set_raw(int on) { g_raw = on; }
void oracle_set_boundary_faces(const double *inflow, const double *outflow) {
    g_bfin = inflow;
    g_bfout = outflow;
}
void oracle_char_eigen(double gamma, const double *rl, const double *rr, double *R, double *L) {
    char_eigen(gamma, rl, rr, R, L);
}
void oracle_cweno_weights(int admis, const double *beta, double lwci1, double eps, int wenoz, double *omega) {
    cweno_weights(admis, beta, lwci1, eps, wenoz, omega);
}
void oracle_ghost(int code, double gamma, const double *cl, double a1, double a2, double *cr) {
    ghost(code, gamma, cl, a1, a2, cr);
}
void oracle_hllc(double gamma, const double *cl, const double *cr, double *flux) { hllc(gamma, cl, cr, flux); }

int oracle_reconstruct(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const double *u,
                       double *ul, int64_t *counters) {
    return reconstruct(m, s, p, u, ul, counters);
}

int oracle_fluxes(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const double *ul,
                  double *rhs) {
    return fluxes(m, s, p, ul, rhs);
}

int oracle_residual(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const double *u,
                    double *rhs, int64_t *counters) {
    return residual(m, s, p, u, rhs, counters);
}

int oracle_dt(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const double *u, double cfl, double *dt_out) {
    cfl_dt(m, s, u, cfl, dt_out);
    return UCNS_OK;
}

int oracle_rk3_steps(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, double *u,
                     int32_t nsteps, double cfl, double dt_fixed, double *t_io) {
    return rk3_steps(m, s, p, u, nsteps, cfl, dt_fixed, t_io);
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

/* ---- tangent-linear model: the Dual instantiation of the same code ----
 * du: [ncell][5] state tangent; dbc (or null): tangent of the boundary
 * data inflow[0..4], outflow[0..4] set by oracle_set_boundary. */
int oracle_residual_tl(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, const double *u,
                       const double *du, const double *dbc, double *rhs, double *drhs) {
    const size_t n = (size_t)m->ncell * NV;
    set_dbc(dbc);
    std::vector<Dual> x = make_dual(u, du, n), r(n);
    int rc = residual(m, s, p, x.data(), r.data(), (int64_t *)nullptr);
    for (size_t e = 0; e < n; ++e) {
        if (rhs) rhs[e] = r[e].v;
        drhs[e] = r[e].d;
    }
    set_dbc(nullptr);
    return rc;
}

/* nsteps SSP-RK3 steps of fixed dt with the tangent carried along:
 * u, du in/out */
int oracle_rk3_tl(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, const ucns_prestore_v1 *p, double *u, double *du,
                  const double *dbc, int32_t nsteps, double dt_fixed) {
    if (!(dt_fixed > 0.0)) return UCNS_EBADARG;
    const size_t n = (size_t)m->ncell * NV;
    set_dbc(dbc);
    std::vector<Dual> x = make_dual(u, du, n);
    int rc = rk3_steps(m, s, p, x.data(), nsteps, 0.0, dt_fixed, (double *)nullptr);
    for (size_t e = 0; e < n; ++e) {
        u[e] = x[e].v;
        du[e] = x[e].d;
    }
    set_dbc(nullptr);
    return rc;
}

/* Jacobian-vector products of the point routines, for their own tests:
 * HLLC of (cl + t dcl, cr + t dcr) and the ghost state of cl + t dcl */
void oracle_hllc_tl(double gamma, const double *cl, const double *dcl, const double *cr, const double *dcr,
                    double *flux, double *dflux) {
    Dual a[NV], b[NV], f[NV];
    for (int v = 0; v < NV; ++v) {
        a[v] = {cl[v], dcl[v]};
        b[v] = {cr[v], dcr[v]};
    }
    hllc(gamma, a, b, f);
    for (int v = 0; v < NV; ++v) {
        flux[v] = f[v].v;
        dflux[v] = f[v].d;
    }
}

void oracle_ghost_tl(int code, double gamma, const double *cl, const double *dcl, const double *dbc, double a1,
                     double a2, double *cr, double *dcr) {
    Dual a[NV], g[NV];
    for (int v = 0; v < NV; ++v) a[v] = {cl[v], dcl[v]};
    set_dbc(dbc);
    ghost(code, gamma, a, a1, a2, g);
    set_dbc(nullptr);
    for (int v = 0; v < NV; ++v) {
        cr[v] = g[v].v;
        dcr[v] = g[v].d;
    }
}

}  // extern "C"
