// Host-side FP64 preprocessor for the residual path: synthetic periodic
// meshes, connectivity, geometry, stencils and PRESTORE arrays.
// Every stage cites the reference routine it restates (src/... file:line).
// All floating point work is FP64; loops over cells run under OpenMP.

#include "ucns_prep.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <array>
#include <map>
#include <memory>
#include <vector>
#include <omp.h>
#include <sys/mman.h>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ----------------------------------------------------------------------------
// Reference element tables (0-based local node numbers)
// Face node lists: src/grid_p.f90:1950-2063 (hex, tet), 1896-1925 (prism).
const int HEX_FACES[6][4] = {{1, 0, 3, 2}, {4, 5, 6, 7}, {5, 1, 2, 6},
                             {7, 3, 0, 4}, {6, 2, 3, 7}, {4, 0, 1, 5}};
const int TET_FACES[4][3] = {{1, 3, 0}, {1, 0, 2}, {1, 2, 3}, {2, 0, 3}};
// prism: faces 1-2 triangles (top, bottom), 3-5 quadrilaterals
const int PRISM_FACES[5][4] = {{4, 5, 3, -1}, {1, 0, 2, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}};
// pyramid (src/grid_p.f90:2100-2124): face 1 the quadrilateral base, 2-5 triangles to the apex (node 5)
const int PYR_FACES[5][4] = {{3, 2, 1, 0}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {3, 0, 4, -1}};
// Sub-tetrahedra of DECOMPOSE3, src/grid_t.f90:1027-1074.
const int HEX_TETS[6][4] = {{0, 5, 7, 4}, {0, 1, 7, 5}, {1, 6, 7, 5},
                            {0, 7, 2, 3}, {0, 7, 1, 2}, {1, 7, 6, 2}};
const int TET_TETS[1][4] = {{0, 1, 2, 3}};
// prism (DECOMPOSE3 element type 4, src/grid_t.f90:1065-1069)
const int PRISM_TETS[3][4] = {{0, 1, 2, 5}, {0, 1, 5, 4}, {0, 4, 5, 3}};
// pyramid (DECOMPOSE3 element type 3, src/grid_t.f90:1058-1060)
const int PYR_TETS[2][4] = {{0, 1, 2, 4}, {0, 2, 3, 4}};

// hexes of the channel's interface layer (pyramids and tetrahedra), an
// element kind of the generator only: its cells are UCNSP_PYRAMID and UCNSP_TET
constexpr int UCNSP_INTERFACE = 100;
// hexes of the blast box's shell around its hexahedra: a pyramid on each face
// towards a quadrilateral-faced neighbour (hexahedra, shell hexes, the box
// side), two tetrahedra on each face towards a triangle-faced one (Kuhn
// tetrahedra, prisms), all with the hex centre as apex
constexpr int UCNSP_SHELL = 101;
// the diagonal of each hex face (HEX_FACES order) that the Kuhn split and the
// prisms' split take: from the face's lowest corner to its highest, as local
// hex nodes (0 = (i,j,k), 6 = (i+1,j+1,k+1))
const int FACE_DIAG[6][2] = {{0, 2}, {4, 6}, {1, 6}, {0, 7}, {3, 6}, {0, 5}};
// prism layers of the channel: 10 (the configuration's wall region), fewer
// on coarse boxes so that tetrahedra fill at least two thirds of y
inline int channel_layers(int ny) { return std::max(1, std::min(10, ny / 3)); }

inline int nfaces_of(int shape) { return shape == UCNSP_HEX ? 6 : shape == UCNSP_TET ? 4 : 5; }
inline int nnodes_of(int shape) {
    return shape == UCNSP_HEX ? 8 : shape == UCNSP_PRISM ? 6 : shape == UCNSP_PYRAMID ? 5 : 4;
}
inline int face_nv(int shape, int f) {
    return shape == UCNSP_HEX ? 4 : shape == UCNSP_PRISM ? (f < 2 ? 3 : 4) : shape == UCNSP_PYRAMID ? (f == 0 ? 4 : 3) : 3;
}
inline int face_node(int shape, int f, int k) {
    return shape == UCNSP_HEX     ? HEX_FACES[f][k]
           : shape == UCNSP_PRISM ? PRISM_FACES[f][k]
           : shape == UCNSP_PYRAMID ? PYR_FACES[f][k]
                                    : TET_FACES[f][k];
}
inline int nsubtets(int shape) {
    return shape == UCNSP_HEX ? 6 : shape == UCNSP_PRISM ? 3 : shape == UCNSP_PYRAMID ? 2 : 1;
}
inline const int *subtet(int shape, int t) {
    return shape == UCNSP_HEX     ? HEX_TETS[t]
           : shape == UCNSP_PRISM ? PRISM_TETS[t]
           : shape == UCNSP_PYRAMID ? PYR_TETS[t]
                                    : TET_TETS[t];
}

// ----------------------------------------------------------------------------
// Monomial basis of BASIS_REC (POLY=1), src/basis.f90:11-102, up to degree 4.
const int MONO[34][3] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
    {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2},
    {3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}, {2, 0, 1}, {1, 1, 1}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {0, 0, 3},
    {4, 0, 0}, {3, 1, 0}, {2, 2, 0}, {1, 3, 0}, {0, 4, 0}, {3, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 3, 1},
    {2, 0, 2}, {1, 1, 2}, {0, 2, 2}, {1, 0, 3}, {0, 1, 3}, {0, 0, 4}};

inline int ndeg_of(int iorder) { return (iorder + 1) * (iorder + 2) * (iorder + 3) / 6 - 1; }

inline void monomials(double x, double y, double z, int ideg, double *m) {
    double px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = 1.0;
    for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * x;
        py[k] = py[k - 1] * y;
        pz[k] = pz[k - 1] * z;
    }
    for (int d = 0; d < ideg; ++d) m[d] = px[MONO[d][0]] * py[MONO[d][1]] * pz[MONO[d][2]];
}

// ----------------------------------------------------------------------------
// Quadrature
// Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double PI = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double t = std::cos(PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = t; p0 = 1.0; }
            double dp = n * (t * p1 - p0) / (t * t - 1.0);
            double dt = p1 / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-17) break;
        }
        double p0 = 1.0, p1 = t;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = (n == 1) ? 1.0 : n * (t * p1 - p0) / (t * t - 1.0);
        x[i] = t;
        w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
    if (n == 1) { x[0] = 0.0; w[0] = 2.0; }
    // sort ascending
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return x[a] < x[b]; });
    std::vector<double> xs(n), ws(n);
    for (int i = 0; i < n; ++i) { xs[i] = x[idx[i]]; ws[i] = w[idx[i]]; }
    x = xs;
    w = ws;
}

// Gauss-Legendre tables for n = 1..8, solved once.
struct GLTable {
    std::vector<double> x[9], w[9];
    GLTable() {
        for (int n = 1; n <= 8; ++n) gauss_legendre(n, x[n], w[n]);
    }
};
const GLTable &gl_table() {
    static const GLTable t;
    return t;
}

// Quadrilateral face rule in the reference's point order
// (QUADRATUREQUAD3D, src/grid_t.f90:2834-3035): 1-D nodes ordered
// n=1 {0}; n=2 {-g,+g}; n=3 {0,-g,+g}; n=4 {-g1,+g1,-g2,+g2}; tensor loop
// with R outer, S inner; weights scaled by 1/4 so they sum to 1.
int quad_rule(int n, std::vector<double> &rs, std::vector<double> &w) {
    if (n < 1 || n > 4) return -1;
    std::vector<double> x, wx;
    gauss_legendre(n, x, wx);  // ascending
    std::vector<double> nx, nw;
    if (n == 1) { nx = {0.0}; nw = {2.0}; }
    if (n == 2) { nx = {x[0], x[1]}; nw = {wx[0], wx[1]}; }
    if (n == 3) { nx = {0.0, x[0], x[2]}; nw = {wx[1], wx[0], wx[2]}; }
    if (n == 4) { nx = {x[1], x[2], x[0], x[3]}; nw = {wx[1], wx[2], wx[0], wx[3]}; }
    rs.clear();
    w.clear();
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            rs.push_back(nx[a]);
            rs.push_back(nx[b]);
            w.push_back(nw[a] * nw[b] * 0.25);
        }
    return n * n;
}

// Newton on a small square system f(v) = 0 (forward-difference Jacobian,
// Gaussian elimination with partial pivoting), for the orbit parameters of
// the symmetric triangle rules below.
template <int N, class Fn>
void newton_solve(double (&v)[N], Fn F) {
    for (int it = 0; it < 60; ++it) {
        double f0[N];
        F(v, f0);
        double M[N][N + 1];
        for (int j = 0; j < N; ++j) {
            double vp[N];
            for (int k = 0; k < N; ++k) vp[k] = v[k];
            const double h = 1e-7 * std::max(1.0, std::fabs(v[j]));
            vp[j] += h;
            double f1[N];
            F(vp, f1);
            for (int i = 0; i < N; ++i) M[i][j] = (f1[i] - f0[i]) / h;
        }
        for (int i = 0; i < N; ++i) M[i][N] = -f0[i];
        for (int c = 0; c < N; ++c) {
            int p = c;
            for (int r = c + 1; r < N; ++r)
                if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
            for (int j = 0; j <= N; ++j) std::swap(M[c][j], M[p][j]);
            for (int r = c + 1; r < N; ++r) {
                const double m = M[r][c] / M[c][c];
                for (int j = c; j <= N; ++j) M[r][j] -= m * M[c][j];
            }
        }
        double dx[N], nrm = 0.0;
        for (int r = N - 1; r >= 0; --r) {
            double s = M[r][N];
            for (int j = r + 1; j < N; ++j) s -= M[r][j] * dx[j];
            dx[r] = s / M[r][r];
        }
        for (int j = 0; j < N; ++j) {
            v[j] += dx[j];
            nrm += std::fabs(dx[j]);
        }
        if (nrm < 1e-17) break;
    }
}

// S3-invariants of a barycentric point: e2 = l1 l2 + l2 l3 + l3 l1, e3 = l1 l2 l3
inline void tri_invariants(double l1, double l2, double l3, double &e2, double &e3) {
    e2 = l1 * l2 + l2 * l3 + l3 * l1;
    e3 = l1 * l2 * l3;
}

// Triangle face rule (QUADRATURETRIANG, src/grid_t.f90:2054-2348): the
// symmetric rules with 1, 3, 6 and 10 points (degrees 1, 2, 4, 5) in the
// reference's point order.  Orbit parameters are solved to full double
// precision from the moment equations of the S3-invariants over the triangle
// (means: 1 -> 1, e2 -> 1/4, e3 -> 1/60, e2^2 -> 1/15, e2 e3 -> 1/210, from
// E[l1^a l2^b l3^c] = 2 a! b! c! / (a+b+c+2)!) instead of copying 15-digit
// constants.  The 6-point rule is fixed by its degree; the 10-point rule
// (orbits (a,b,b), the 6 permutations of (p,q,r), the centroid) is a
// one-parameter family at degree 5, and takes its vertex-orbit coordinate
// a = 0.888871894660413 from the reference (grid_t.f90:2099) with p, q and
// the weights solved from the moments.
int tri_rule(int n, std::vector<double> &bary, std::vector<double> &w) {
    bary.clear();
    w.clear();
    if (n == 1) {
        bary = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
        w = {1.0};
        return 1;
    }
    if (n == 2) {
        const double a = 2.0 / 3.0, b = 1.0 / 6.0;
        bary = {a, b, b, b, a, b, b, b, a};
        w = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
        return 3;
    }
    if (n == 3) {
        // unknowns: a (small orbit), b (large orbit), wa, wb (per point)
        double v[4] = {0.0915762135, 0.4459484909, 0.1099517437, 0.2233815897};
        auto F = [](const double *p, double *f) {
            double a = p[0], b = p[1], wa = p[2], wb = p[3];
            double e2a = 2 * a - 3 * a * a, e2b = 2 * b - 3 * b * b;
            double e3a = a * a * (1 - 2 * a), e3b = b * b * (1 - 2 * b);
            f[0] = 3 * wa + 3 * wb - 1.0;
            f[1] = 3 * wa * e2a + 3 * wb * e2b - 0.25;
            f[2] = 3 * wa * e3a + 3 * wb * e3b - 1.0 / 60.0;
            f[3] = 3 * wa * e2a * e2a + 3 * wb * e2b * e2b - 1.0 / 15.0;
        };
        for (int it = 0; it < 60; ++it) {
            double f0[4];
            F(v, f0);
            double J[4][4];
            for (int j = 0; j < 4; ++j) {
                double vp[4] = {v[0], v[1], v[2], v[3]};
                double h = 1e-7 * std::max(1.0, std::fabs(v[j]));
                vp[j] += h;
                double f1[4];
                F(vp, f1);
                for (int i = 0; i < 4; ++i) J[i][j] = (f1[i] - f0[i]) / h;
            }
            // solve J dx = -f0 (Gaussian elimination with pivoting)
            double M[4][5];
            for (int i = 0; i < 4; ++i) {
                for (int j = 0; j < 4; ++j) M[i][j] = J[i][j];
                M[i][4] = -f0[i];
            }
            for (int c = 0; c < 4; ++c) {
                int p = c;
                for (int r = c + 1; r < 4; ++r)
                    if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
                for (int j = 0; j < 5; ++j) std::swap(M[c][j], M[p][j]);
                for (int r = c + 1; r < 4; ++r) {
                    double m = M[r][c] / M[c][c];
                    for (int j = c; j < 5; ++j) M[r][j] -= m * M[c][j];
                }
            }
            double dx[4];
            for (int r = 3; r >= 0; --r) {
                double s = M[r][4];
                for (int j = r + 1; j < 4; ++j) s -= M[r][j] * dx[j];
                dx[r] = s / M[r][r];
            }
            double nrm = 0;
            for (int j = 0; j < 4; ++j) { v[j] += dx[j]; nrm += std::fabs(dx[j]); }
            if (nrm < 1e-17) break;
        }
        const double a = v[0], b = v[1], A = 1.0 - 2.0 * a, B = 1.0 - 2.0 * b;
        bary = {A, a, a, a, A, a, a, a, A, b, b, B, b, B, b, B, b, b};
        w = {v[2], v[2], v[2], v[3], v[3], v[3]};
        return 6;
    }
    if (n == 4) {
        const double a = 0.888871894660413, b = 0.5 * (1.0 - a);
        double ea2, ea3, ec2, ec3;
        tri_invariants(a, b, b, ea2, ea3);
        tri_invariants(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, ec2, ec3);
        // unknowns: p, q (r = 1 - p - q), wa (vertex orbit), wp (6-orbit); the
        // centroid takes 1 - 3 wa - 6 wp
        double v[4] = {0.2955, 0.6342, 0.042, 0.112};
        newton_solve(v, [&](const double *x, double *f) {
            double e2, e3;
            tri_invariants(x[0], x[1], 1.0 - x[0] - x[1], e2, e3);
            const double wa = 3.0 * x[2], wp = 6.0 * x[3], wc = 1.0 - wa - wp;
            f[0] = wa * ea2 + wp * e2 + wc * ec2 - 0.25;
            f[1] = wa * ea3 + wp * e3 + wc * ec3 - 1.0 / 60.0;
            f[2] = wa * ea2 * ea2 + wp * e2 * e2 + wc * ec2 * ec2 - 1.0 / 15.0;
            f[3] = wa * ea2 * ea3 + wp * e2 * e3 + wc * ec2 * ec3 - 1.0 / 210.0;
        });
        const double p = v[0], q = v[1], r = 1.0 - p - q, c = 1.0 / 3.0;
        bary = {a, b, b, b, a, b, b, b, a, p, q, r, p, r, q, r, p, q, q, p, r, q, r, p, r, q, p, c, c, c};
        w = {v[2], v[2], v[2], v[3], v[3], v[3], v[3], v[3], v[3], 1.0 - 3.0 * v[2] - 6.0 * v[3]};
        return 10;
    }
    return -1;
}

// Volume rules.  Stroud T3:3-1 (degree 3, 5 points) on a tetrahedron in
// barycentric form; collapsed (Duffy) Gauss-Legendre product rule of any
// degree; tensor Gauss-Legendre on a parallelepiped.  All return points in
// physical coordinates and weights that include the volume.
struct QPts {
    std::vector<double> x;  // n*3
    std::vector<double> w;
};

void tet_stroud3(const double v[4][3], QPts &q) {
    double e1[3], e2[3], e3[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = v[1][d] - v[0][d];
        e2[d] = v[2][d] - v[0][d];
        e3[d] = v[3][d] - v[0][d];
    }
    double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                 e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    double vol = std::fabs(det) / 6.0;
    const double L[5][4] = {{0.25, 0.25, 0.25, 0.25},
                            {0.5, 1.0 / 6, 1.0 / 6, 1.0 / 6},
                            {1.0 / 6, 0.5, 1.0 / 6, 1.0 / 6},
                            {1.0 / 6, 1.0 / 6, 0.5, 1.0 / 6},
                            {1.0 / 6, 1.0 / 6, 1.0 / 6, 0.5}};
    const double W[5] = {-0.8, 0.45, 0.45, 0.45, 0.45};
    for (int p = 0; p < 5; ++p) {
        for (int d = 0; d < 3; ++d)
            q.x.push_back(L[p][0] * v[0][d] + L[p][1] * v[1][d] + L[p][2] * v[2][d] + L[p][3] * v[3][d]);
        q.w.push_back(W[p] * vol);
    }
}

void tet_collapsed(const double v[4][3], int n, QPts &q) {
    const std::vector<double> &g = gl_table().x[n], &gw = gl_table().w[n];
    double e1[3], e2[3], e3[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = v[1][d] - v[0][d];
        e2[d] = v[2][d] - v[0][d];
        e3[d] = v[3][d] - v[0][d];
    }
    double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                 e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    double ad = std::fabs(det);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            for (int c = 0; c < n; ++c) {
                double u = 0.5 * (g[a] + 1), s = 0.5 * (g[b] + 1), t = 0.5 * (g[c] + 1);
                // reference simplex point: (xi,eta,zeta) with xi=u, eta=(1-u)s, zeta=(1-u)(1-s)t
                double xi = u, eta = (1 - u) * s, zeta = (1 - u) * (1 - s) * t;
                double jac = (1 - u) * (1 - u) * (1 - s);
                double wt = 0.125 * gw[a] * gw[b] * gw[c] * jac * ad;
                for (int d = 0; d < 3; ++d) q.x.push_back(v[0][d] + xi * e1[d] + eta * e2[d] + zeta * e3[d]);
                q.w.push_back(wt);
            }
}

void hex_tensor(const double v[8][3], int n, QPts &q) {
    // parallelepiped spanned at node 1 by (2-1), (4-1), (5-1)
    const std::vector<double> &g = gl_table().x[n], &gw = gl_table().w[n];
    double e1[3], e2[3], e3[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = v[1][d] - v[0][d];
        e2[d] = v[3][d] - v[0][d];
        e3[d] = v[4][d] - v[0][d];
    }
    double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                 e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    double ad = std::fabs(det);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            for (int c = 0; c < n; ++c) {
                double r = 0.5 * (g[a] + 1), s = 0.5 * (g[b] + 1), t = 0.5 * (g[c] + 1);
                for (int d = 0; d < 3; ++d) q.x.push_back(v[0][d] + r * e1[d] + s * e2[d] + t * e3[d]);
                q.w.push_back(0.125 * gw[a] * gw[b] * gw[c] * ad);
            }
}

bool hex_is_affine(const double v[8][3]) {
    double scale = 0;
    for (int d = 0; d < 3; ++d) scale = std::max(scale, std::fabs(v[6][d] - v[0][d]));
    const double tol = 1e-12 * std::max(scale, 1e-300);
    // every node equals v0 + combination of e1,e2,e3
    const int comb[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < 8; ++k)
        for (int d = 0; d < 3; ++d) {
            double p = v[0][d] + comb[k][0] * (v[1][d] - v[0][d]) + comb[k][1] * (v[3][d] - v[0][d]) +
                       comb[k][2] * (v[4][d] - v[0][d]);
            if (std::fabs(p - v[k][d]) > tol) return false;
        }
    return true;
}

// Integration points of a whole cell, exact for polynomials of total degree
// <= deg (affine hex: tensor GL; otherwise the DECOMPOSE3 sub-tetrahedra).
void cell_rule(int shape, const double v[8][3], int deg, QPts &q) {
    q.x.clear();
    q.w.clear();
    if (shape == UCNSP_HEX && hex_is_affine(v)) {
        int n = std::max(1, (deg + 2) / 2);
        hex_tensor(v, n, q);
        return;
    }
    for (int t = 0; t < nsubtets(shape); ++t) {
        const int *tt = subtet(shape, t);
        double tv[4][3];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) tv[k][d] = v[tt[k]][d];
        if (deg <= 3)
            tet_stroud3(tv, q);
        else
            tet_collapsed(tv, (deg + 4) / 2, q);
    }
}

double tet_volume(const double a[3], const double b[3], const double c[3], const double d[3]) {
    double e1[3], e2[3], e3[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = b[k] - a[k];
        e2[k] = c[k] - a[k];
        e3[k] = d[k] - a[k];
    }
    double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                 e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    return std::fabs(det) / 6.0;
}

// ----------------------------------------------------------------------------
uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double u01(uint64_t &s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

template <class T>
struct Buf {  // uninitialised heap array on 2 MiB pages, filled in parallel (first touch)
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        const size_t align = (size_t)2 << 20;
        size_t bytes = ((sizeof(T) * count + align - 1) / align) * align;
        p = static_cast<T *>(std::aligned_alloc(align, bytes));
        if (p) madvise(p, bytes, MADV_HUGEPAGE);
    }
    void release() {
        std::free(p);
        p = nullptr;
        n = 0;
    }
    ~Buf() { release(); }
};

}  // namespace

// ----------------------------------------------------------------------------
struct ucnsp_case {
    ucnsp_opts o{};
    int shape = 0;
    int64_t ncell = 0, nnode = 0;
    static const int NF = 6;  // nface_max
    double per[3] = {0, 0, 0};
    std::vector<double> xyz;
    std::vector<double> xyz_box;  // node coordinates before the bump
    std::vector<int32_t> canon;
    std::vector<int32_t> cnodes;  // ncell*8
    std::vector<int32_t> cshape;
    // connectivity / geometry (per cell-face: ncell*NF)
    std::vector<int32_t> ifca, ineigh, ineighn, bcode, types_faces, interior, reorient;
    std::vector<int32_t> fnodes;  // ncell*NF*4 face vertices after CHECKGRADS adoption
    std::vector<double> angx, angy, surf, normal, vol, minedge, centre;
    // scheme sizes
    int iorder = 1, ideg = 3, nnb = 9, imax = 8, typesten = 1, iselem = 0, igq = 1;
    int qp_quad = 1, qp_tri = 1, nqpmax = 1;
    std::vector<double> quad_rs, quad_w, tri_bary, tri_w;
    // stencils and prestore
    std::vector<int32_t> admis, full, ihexl;
    Buf<double> invmat;
    std::vector<double> indicator, integ_basis, volume11, qpoints, frame;
    std::vector<double> framedet;
    int64_t n_solved = 0;  // stencils whose pseudo-inverse was solved (others copied from a congruent one)
    // CWENO (EES=5): the sector stencils s >= 1 take a polynomial of degree
    // iorder2 (ideg2 coefficients) over their first imax2 members
    // (src/grid_p.f90:6151-6163, src/prestore.f90:244-263, 296-337, 520-535)
    int iorder2 = 0, ideg2 = 0, imax2 = 0;
    std::vector<double> integ_basis_c, indicator_c;
    Buf<double> invmat_c;
    bool congruence_cache = true;  // copy data between congruent cells (off once nodes have been moved)
    // a mesh read from GRID.cel/.vrt/.bnd: the boundary codes by the sorted
    // node ids of each GRID.bnd face (0-based; triangles end in -1)
    bool from_file = false;
    std::map<std::array<int32_t, 4>, int32_t> bnd_code;
};

namespace {

// Face matching (NEIGHBOURSS, src/grid_p.f90:1850): faces with the same set
// of periodic-canonical node ids are neighbours; a pair whose raw node ids
// differ is a periodic pair (boundary code 5).
void build_connectivity(ucnsp_case &c) {
    const int64_t nc = c.ncell;
    const int NF = ucnsp_case::NF;
    struct Key {
        int32_t k[4];
        int64_t slot;
    };
    std::vector<Key> keys;
    keys.reserve(nc * nfaces_of(c.shape));
    c.ifca.assign(nc, 0);
    for (int64_t i = 0; i < nc; ++i) {
        int sh = c.cshape[i];
        c.ifca[i] = nfaces_of(sh);
        for (int f = 0; f < nfaces_of(sh); ++f) {
            Key K;
            int nv = face_nv(sh, f);
            for (int k = 0; k < 4; ++k) K.k[k] = (k < nv) ? c.canon[c.cnodes[i * 8 + face_node(sh, f, k)]] : -1;
            std::sort(K.k, K.k + 4);
            K.slot = i * NF + f;
            keys.push_back(K);
        }
    }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        for (int k = 0; k < 4; ++k)
            if (a.k[k] != b.k[k]) return a.k[k] < b.k[k];
        return a.slot < b.slot;
    });
    c.ineigh.assign(nc * NF, -1);
    c.ineighn.assign(nc * NF, -1);
    c.bcode.assign(nc * NF, 0);
    c.types_faces.assign(nc * NF, 0);
    c.interior.assign(nc, 0);
    for (size_t a = 0; a + 1 < keys.size(); ++a) {
        const Key &A = keys[a], &B = keys[a + 1];
        if (std::memcmp(A.k, B.k, sizeof(A.k)) != 0) continue;
        int64_t i = A.slot / NF, fi = A.slot % NF, j = B.slot / NF, fj = B.slot % NF;
        c.ineigh[A.slot] = (int32_t)j;
        c.ineighn[A.slot] = (int32_t)fj;
        c.ineigh[B.slot] = (int32_t)i;
        c.ineighn[B.slot] = (int32_t)fi;
        // periodic if raw node sets differ
        int sh_i = c.cshape[i], sh_j = c.cshape[j];
        int32_t ri[4] = {-1, -1, -1, -1}, rj[4] = {-1, -1, -1, -1};
        for (int k = 0; k < face_nv(sh_i, (int)fi); ++k) ri[k] = c.cnodes[i * 8 + face_node(sh_i, (int)fi, k)];
        for (int k = 0; k < face_nv(sh_j, (int)fj); ++k) rj[k] = c.cnodes[j * 8 + face_node(sh_j, (int)fj, k)];
        std::sort(ri, ri + 4);
        std::sort(rj, rj + 4);
        if (std::memcmp(ri, rj, sizeof(ri)) != 0) {
            c.bcode[A.slot] = 5;
            c.bcode[B.slot] = 5;
            c.interior[i] = 1;
            c.interior[j] = 1;
        }
        ++a;
    }
    for (int64_t i = 0; i < nc; ++i)
        for (int f = 0; f < c.ifca[i]; ++f) c.types_faces[i * NF + f] = face_nv(c.cshape[i], f) == 4 ? 5 : 6;
    if (c.from_file) {
        // GRID.bnd: codes of the unmatched faces and of the periodic pairs (5 or 50)
        for (int64_t i = 0; i < nc; ++i)
            for (int f = 0; f < c.ifca[i]; ++f) {
                const int64_t sl = i * NF + f;
                if (c.ineigh[sl] >= 0 && c.bcode[sl] == 0) continue;
                std::array<int32_t, 4> k = {-1, -1, -1, -1};
                const int sh = c.cshape[i], nv = face_nv(sh, f);
                for (int q = 0; q < nv; ++q) k[q] = c.cnodes[i * 8 + face_node(sh, f, q)];
                std::sort(k.begin(), k.begin() + nv);
                const auto it = c.bnd_code.find(k);
                if (it != c.bnd_code.end())
                    c.bcode[sl] = it->second;
                else if (c.ineigh[sl] < 0)
                    c.bcode[sl] = -1;  // a boundary face GRID.bnd does not list
            }
        return;
    }
    // unmatched faces lie on a bounded side of the box: code of that side
    const double L[3] = {c.o.lx, c.o.ly, c.o.lz};
    for (int64_t i = 0; i < nc; ++i)
        for (int f = 0; f < c.ifca[i]; ++f) {
            if (c.ineigh[i * NF + f] >= 0) continue;
            int sh = c.cshape[i], nv = face_nv(sh, f), side = -1;
            for (int d = 0; d < 3 && side < 0; ++d)
                for (int e = 0; e < 2 && side < 0; ++e) {
                    bool on = true;
                    for (int k = 0; k < nv; ++k) {
                        const double x = c.xyz[(size_t)c.cnodes[i * 8 + face_node(sh, f, k)] * 3 + d];
                        if (std::fabs(x - e * L[d]) > 1e-9 * L[d]) on = false;
                    }
                    if (on) side = 2 * d + e;
                }
            c.bcode[i * NF + f] = side >= 0 ? c.o.bc[side] : -1;
        }
}

inline void node_xyz(const ucnsp_case &c, int32_t node, double *p) {
    p[0] = c.xyz[node * 3 + 0];
    p[1] = c.xyz[node * 3 + 1];
    p[2] = c.xyz[node * 3 + 2];
}

// Shift a point by whole periods so it lies within half a period of ref
// (strict test as in LOCALISE_STEN2, src/local_pt.f90:956-981).
inline void unwrap(const ucnsp_case &c, const double *ref, double *p) {
    for (int d = 0; d < 3; ++d) {
        if (c.per[d] <= 0) continue;
        if (std::fabs(p[d] - ref[d]) > 0.5 * c.per[d]) p[d] += c.per[d] * ((ref[d] - 0.5 * c.per[d]) >= 0 ? 1.0 : -1.0);
    }
}

// Angles of the outward normal (ANGLEX/ANGLEY, src/grid_t.f90:4441-4508).
void normal_angles(double nx, double ny, double nz, double &a1, double &a2) {
    const double PI = 3.14159265358979323846;
    a1 = 0.0;
    if (nx != 0.0 && ny != 0.0) {
        if (nx > 0 && ny > 0) a1 = std::atan(ny / nx);
        if (nx < 0 && ny > 0) a1 = PI + std::atan(ny / nx);
        if (nx < 0 && ny < 0) a1 = PI + std::atan(ny / nx);
        if (nx > 0 && ny < 0) a1 = 2.0 * PI + std::atan(ny / nx);
    } else if (nx == 0.0 && ny != 0.0) {
        a1 = ny > 0 ? PI / 2.0 : 3.0 * (PI / 2.0);
    } else if (nx != 0.0 && ny == 0.0) {
        a1 = nx > 0 ? 0.0 : PI;
    }
    a2 = std::acos(nz == 0.0 ? 0.0 : nz);
}

// Size a vector and fill it, keeping its storage when the size is unchanged
// (arrays handed out by ucnsp_array / ucnsp_fill stay valid across
// ucnsp_move_nodes).
template <class T>
void keep_assign(std::vector<T> &v, size_t n, T x) {
    if (v.size() == n)
        std::fill(v.begin(), v.end(), x);
    else
        v.assign(n, x);
}

// GEOMETRY_CALC: volumes (DECOMPOSE3 sums), centres (vertex average),
// face areas (QUADAREA / TRIANGLEAREA), Newell normals and their angles
// (FIND_ROT_ANGLES), MINEDGE = 3V/sum(S) (EDGE_CALCULATOR3d).
void build_geometry(ucnsp_case &c) {
    const int64_t nc = c.ncell;
    const int NF = ucnsp_case::NF;
    keep_assign(c.vol, nc, 0.0);
    keep_assign(c.centre, nc * 3, 0.0);
    keep_assign(c.minedge, nc, 0.0);
    keep_assign(c.surf, nc * NF, 0.0);
    keep_assign(c.angx, nc * NF, 0.0);
    keep_assign(c.angy, nc * NF, 0.0);
    keep_assign(c.normal, nc * NF * 3, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        int sh = c.cshape[i];
        int nn = nnodes_of(sh);
        double v[8][3];
        for (int k = 0; k < nn; ++k) node_xyz(c, c.cnodes[i * 8 + k], v[k]);
        double cen[3] = {0, 0, 0};
        for (int k = 0; k < nn; ++k)
            for (int d = 0; d < 3; ++d) cen[d] += v[k][d];
        for (int d = 0; d < 3; ++d) c.centre[i * 3 + d] = cen[d] / nn;
        double vol = 0.0;
        for (int t = 0; t < nsubtets(sh); ++t) {
            const int *tt = subtet(sh, t);
            vol += tet_volume(v[tt[0]], v[tt[1]], v[tt[2]], v[tt[3]]);
        }
        c.vol[i] = vol;
        double ssum = 0.0;
        for (int f = 0; f < c.ifca[i]; ++f) {
            int nv = face_nv(sh, f);
            double fv[4][3];
            for (int k = 0; k < nv; ++k)
                for (int d = 0; d < 3; ++d) fv[k][d] = v[face_node(sh, f, k)][d];
            double area;
            if (nv == 4) {
                double e[3], dd[3];
                for (int d = 0; d < 3; ++d) {
                    e[d] = fv[3][d] - fv[1][d];
                    dd[d] = fv[2][d] - fv[0][d];
                }
                double cx = e[1] * dd[2] - e[2] * dd[1], cy = e[2] * dd[0] - e[0] * dd[2],
                       cz = e[0] * dd[1] - e[1] * dd[0];
                area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
                // two-triangle check of SURFACE_CALCULATOR3 (src/grid_t.f90:1530-1556)
                auto tri_area = [](const double *a, const double *b, const double *cc) {
                    double u[3], w[3];
                    for (int d = 0; d < 3; ++d) {
                        u[d] = b[d] - a[d];
                        w[d] = cc[d] - a[d];
                    }
                    double x = u[1] * w[2] - u[2] * w[1], y = u[2] * w[0] - u[0] * w[2], z = u[0] * w[1] - u[1] * w[0];
                    return 0.5 * std::sqrt(x * x + y * y + z * z);
                };
                double a2 = tri_area(fv[0], fv[1], fv[2]) + tri_area(fv[0], fv[2], fv[3]);
                if (std::fabs((a2 - area) / area) * 100.0 > 10.0) area = a2;
            } else {
                double u[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    u[d] = fv[1][d] - fv[0][d];
                    w[d] = fv[2][d] - fv[0][d];
                }
                double x = u[1] * w[2] - u[2] * w[1], y = u[2] * w[0] - u[0] * w[2], z = u[0] * w[1] - u[1] * w[0];
                area = 0.5 * std::sqrt(x * x + y * y + z * z);
            }
            c.surf[i * NF + f] = area;
            ssum += area;
            // Newell normal over the (outward ordered) polygon
            double nx = 0, ny = 0, nz = 0;
            for (int k = 0; k < nv; ++k) {
                const double *a = fv[k], *b = fv[(k + 1) % nv];
                nx += (a[1] - b[1]) * (a[2] + b[2]);
                ny += (a[2] - b[2]) * (a[0] + b[0]);
                nz += (a[0] - b[0]) * (a[1] + b[1]);
            }
            double r = std::sqrt(nx * nx + ny * ny + nz * nz);
            nx /= r;
            ny /= r;
            nz /= r;
            double a1, a2;
            normal_angles(nx, ny, nz, a1, a2);
            c.angx[i * NF + f] = a1;
            c.angy[i * NF + f] = a2;
            c.normal[(i * NF + f) * 3 + 0] = nx;
            c.normal[(i * NF + f) * 3 + 1] = ny;
            c.normal[(i * NF + f) * 3 + 2] = nz;
        }
        c.minedge[i] = 3.0 * vol / ssum;
    }
    // CHECKGRADS face-vertex adoption (src/local_pt.f90:2235-2268): on a
    // shared face the lower-numbered cell takes the higher-numbered cell's
    // vertex order, so both sides build their Gauss points in the same order.
    keep_assign(c.fnodes, nc * NF * 4, -1);
    keep_assign(c.reorient, nc * NF, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        int sh = c.cshape[i];
        for (int f = 0; f < c.ifca[i]; ++f) {
            int64_t j = c.ineigh[i * NF + f];
            if (j >= 0 && j > i) {
                int fj = c.ineighn[i * NF + f];
                int shj = c.cshape[j];
                for (int k = 0; k < face_nv(shj, fj); ++k)
                    c.fnodes[(i * NF + f) * 4 + k] = c.cnodes[j * 8 + face_node(shj, fj, k)];
                c.reorient[i * NF + f] = 1;
            } else {
                for (int k = 0; k < face_nv(sh, f); ++k)
                    c.fnodes[(i * NF + f) * 4 + k] = c.cnodes[i * 8 + face_node(sh, f, k)];
            }
        }
    }
}

// Trihedral cone of CHECK_CONDITION (src/grid_p.f90:4582-4811): a point p
// is inside when VVA1 (p - o) >= 0 componentwise, VVA1 the inverse of
// [a-o, b-o, c-o] as COMPUTEJACOBIANS builds it (adjugate divided by the
// determinant, src/grid_t.f90:581-616).  The inverse is formed once per
// cone; the test keeps the reference's operation order (ties on cone faces
// are decided by this rounding).
struct Cone {
    double I[3][3];
    double o[3];
};

void make_cone(const double *o, const double *a, const double *b, const double *cc, Cone &k) {
    double A[3][3];
    for (int d = 0; d < 3; ++d) {
        A[d][0] = a[d] - o[d];
        A[d][1] = b[d] - o[d];
        A[d][2] = cc[d] - o[d];
        k.o[d] = o[d];
    }
    double det = A[0][0] * (A[2][2] * A[1][1] - A[2][1] * A[1][2]) - A[1][0] * (A[2][2] * A[0][1] - A[2][1] * A[0][2]) +
                 A[2][0] * (A[1][2] * A[0][1] - A[1][1] * A[0][2]);
    double (&I)[3][3] = k.I;
    I[0][0] = A[2][2] * A[1][1] - A[2][1] * A[1][2];
    I[0][1] = -(A[2][2] * A[0][1] - A[2][1] * A[0][2]);
    I[0][2] = A[1][2] * A[0][1] - A[1][1] * A[0][2];
    I[1][0] = -(A[2][2] * A[1][0] - A[2][0] * A[1][2]);
    I[1][1] = A[2][2] * A[0][0] - A[2][0] * A[0][2];
    I[1][2] = -(A[1][2] * A[0][0] - A[1][0] * A[0][2]);
    I[2][0] = A[2][1] * A[1][0] - A[2][0] * A[1][1];
    I[2][1] = -(A[2][1] * A[0][0] - A[2][0] * A[0][1]);
    I[2][2] = A[1][1] * A[0][0] - A[1][0] * A[0][1];
    for (int r = 0; r < 3; ++r)
        for (int c2 = 0; c2 < 3; ++c2) I[r][c2] = I[r][c2] / det;
}

inline bool in_cone(const Cone &k, const double *p) {
    const double q[3] = {p[0] - k.o[0], p[1] - k.o[1], p[2] - k.o[2]};
    for (int r = 0; r < 3; ++r) {
        double x = k.I[r][0] * q[0];
        x = x + k.I[r][1] * q[1];
        x = x + k.I[r][2] * q[2];
        if (!(x >= 0.0)) return false;
    }
    return true;
}

// Stencils: BFS over face neighbours (ALLSf, src/grid_p.f90:4411-4551) with
// the candidate pool truncated at ISELEM; central stencil = first nnb BFS
// cells (ICOMPACT=0: no distance sort); one directional stencil per face
// with a neighbour, filled with BFS candidates inside the face cone
// (STENCIILS, src/grid_p.f90:5822-5960); complete stencils are compacted in
// order (LOCALISE_STENCIL, src/local_pt.f90:653-760) -> ADMIS, FULL
// (STENCILS, src/grid_p.f90:6452-6520).
void build_stencils(ucnsp_case &c) {
    const int64_t nc = c.ncell;
    const int NF = ucnsp_case::NF;
    const int TS = c.typesten, NNB = c.nnb;
    c.admis.assign(nc, 0);
    c.full.assign(nc, 0);
    c.ihexl.assign((size_t)nc * TS * NNB, -1);
    const int nthr = omp_get_max_threads();
    std::vector<std::vector<int32_t>> stamp(nthr, std::vector<int32_t>(nc, -1));
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        std::vector<int32_t> &mark = stamp[tid];
        std::vector<int32_t> list;
        list.reserve(c.iselem);
        std::vector<int32_t> st((size_t)TS * NNB);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < nc; ++i) {
            list.clear();
            list.push_back((int32_t)i);
            mark[i] = (int32_t)i;
            size_t head = 0;
            while ((int)list.size() < c.iselem && head < list.size()) {
                int64_t cur = list[head++];
                for (int f = 0; f < c.ifca[cur] && (int)list.size() < c.iselem; ++f) {
                    int32_t nb = c.ineigh[cur * NF + f];
                    if (nb < 0 || mark[nb] == (int32_t)i) continue;
                    mark[nb] = (int32_t)i;
                    list.push_back(nb);
                }
            }
            std::fill(st.begin(), st.end(), -1);
            // stencil 1: central
            for (int k = 0; k < NNB && k < (int)list.size(); ++k) st[k] = list[k];
            // stencils 2..: directional
            if (TS > 1) {
                const double *bc = &c.centre[i * 3];
                int sh = c.cshape[i];
                for (int f = 0; f < c.ifca[i] && f + 1 < TS; ++f) {
                    if (c.ineigh[i * NF + f] < 0) continue;
                    int32_t *row = &st[(size_t)(f + 1) * NNB];
                    row[0] = (int32_t)i;
                    int cnt = 1;
                    int nv = face_nv(sh, f);
                    double fv[4][3];
                    for (int k = 0; k < nv; ++k) node_xyz(c, c.cnodes[i * 8 + face_node(sh, f, k)], fv[k]);
                    // sub-cones: triangle (n1,n3,n2); quadrilateral the four
                    // corner triangles (n2,n1,n3) (n2,n4,n3) (n1,n4,n3) (n1,n4,n2)
                    Cone cones[4];
                    int ncone;
                    if (nv == 3) {
                        make_cone(bc, fv[0], fv[2], fv[1], cones[0]);
                        ncone = 1;
                    } else {
                        make_cone(bc, fv[1], fv[0], fv[2], cones[0]);
                        make_cone(bc, fv[1], fv[3], fv[2], cones[1]);
                        make_cone(bc, fv[0], fv[3], fv[2], cones[2]);
                        make_cone(bc, fv[0], fv[3], fv[1], cones[3]);
                        ncone = 4;
                    }
                    for (size_t J = 1; J < list.size() && cnt < NNB; ++J) {
                        int32_t cand = list[J];
                        double vg[3] = {c.centre[cand * 3], c.centre[cand * 3 + 1], c.centre[cand * 3 + 2]};
                        for (int d = 0; d < 3; ++d)
                            if (c.per[d] > 0 && std::fabs(vg[d] - bc[d]) >= 0.5 * c.per[d])
                                vg[d] += c.per[d] * ((bc[d] - 0.5 * c.per[d]) >= 0 ? 1.0 : -1.0);
                        bool ok = false;
                        for (int kc = 0; kc < ncone && !ok; ++kc) ok = in_cone(cones[kc], vg);
                        if (ok) row[cnt++] = cand;
                    }
                }
            }
            // compact complete stencils
            int ad = 0;
            for (int s = 0; s < TS; ++s) {
                bool complete = true;
                for (int k = 0; k < NNB; ++k)
                    if (st[(size_t)s * NNB + k] < 0) { complete = false; break; }
                if (!complete) continue;
                std::memcpy(&c.ihexl[((size_t)i * TS + ad) * NNB], &st[(size_t)s * NNB], sizeof(int32_t) * NNB);
                ++ad;
            }
            c.admis[i] = ad;
            c.full[i] = (c.o.iweno == 1 && ad >= c.ifca[i] + 1) ? 1 : 0;
        }
    }
}

// Householder QR of a square matrix (QRDECOMPOSITION, src/matrix.f90:93-146)
// with the reflector applied in place instead of forming P_l, then the
// triangular inverse of INVERT (src/prestore.f90:4047-4069);
// returns invmat = R^-1 Q^T.
void qr_inverse(int n, const double *M, double *invmat, double *work) {
    double *R = work, *Q = work + n * n, *v = work + 2 * n * n, *Ri = work + 2 * n * n + n;
    for (int a = 0; a < n * n; ++a) R[a] = M[a];
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) Q[a * n + b] = (a == b) ? 1.0 : 0.0;
    for (int l = 0; l < n; ++l) {
        double nrm = 0.0;
        for (int r = l; r < n; ++r) nrm += R[r * n + l] * R[r * n + l];
        nrm = std::sqrt(nrm);
        for (int r = l; r < n; ++r) v[r] = R[r * n + l];
        v[l] += (R[l * n + l] >= 0.0 ? 1.0 : -1.0) * nrm;
        double vn = 0.0;
        for (int r = l; r < n; ++r) vn += v[r] * v[r];
        vn = std::sqrt(vn);
        for (int r = l; r < n; ++r) v[r] /= vn;
        for (int col = 0; col < n; ++col) {  // R = (I - 2 v v^T) R
            double s = 0.0;
            for (int r = l; r < n; ++r) s += v[r] * R[r * n + col];
            for (int r = l; r < n; ++r) R[r * n + col] -= 2.0 * v[r] * s;
        }
        for (int row = 0; row < n; ++row) {  // Q = Q (I - 2 v v^T)
            double s = 0.0;
            for (int r = l; r < n; ++r) s += Q[row * n + r] * v[r];
            for (int r = l; r < n; ++r) Q[row * n + r] -= 2.0 * s * v[r];
        }
    }
    for (int a = 0; a < n * n; ++a) Ri[a] = 0.0;
    for (int i = n - 1; i >= 0; --i) {
        Ri[i * n + i] = 1.0 / R[i * n + i];
        for (int j = i + 1; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < j; ++k) s -= R[k * n + j] * Ri[i * n + k];
            Ri[i * n + j] = s / R[j * n + j];
        }
    }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += Ri[a * n + k] * Q[b * n + k];
            invmat[a * n + b] = s;
        }
}

// Local frame of LOCALISE_STEN2 (src/local_pt.f90:1037-1101): the first
// DECOMPOSE3 sub-tetrahedron; x_local = VVA1 (x - v1), VVA1 = inverse of
// [v2-v1, v3-v1, v4-v1] as built by COMPUTEJACOBIANS (src/grid_t.f90:581-616).
void cell_frame(const ucnsp_case &c, int64_t i, double *J9, double *o3, double &detjc) {
    int sh = c.cshape[i];
    const int *tt = subtet(sh, 0);
    double v[4][3];
    for (int k = 0; k < 4; ++k) node_xyz(c, c.cnodes[i * 8 + tt[k]], v[k]);
    double A[3][3];
    for (int d = 0; d < 3; ++d) {
        A[d][0] = v[1][d] - v[0][d];
        A[d][1] = v[2][d] - v[0][d];
        A[d][2] = v[3][d] - v[0][d];
    }
    double det = A[0][0] * ((A[2][2] * A[1][1]) - (A[2][1] * A[1][2])) - A[1][0] * ((A[2][2] * A[0][1]) - (A[2][1] * A[0][2])) +
                 A[2][0] * ((A[1][2] * A[0][1]) - (A[1][1] * A[0][2]));
    double I[3][3];
    I[0][0] = A[2][2] * A[1][1] - A[2][1] * A[1][2];
    I[0][1] = -(A[2][2] * A[0][1] - A[2][1] * A[0][2]);
    I[0][2] = A[1][2] * A[0][1] - A[1][1] * A[0][2];
    I[1][0] = -(A[2][2] * A[1][0] - A[2][0] * A[1][2]);
    I[1][1] = A[2][2] * A[0][0] - A[2][0] * A[0][2];
    I[1][2] = -(A[1][2] * A[0][0] - A[1][0] * A[0][2]);
    I[2][0] = A[2][1] * A[1][0] - A[2][0] * A[1][1];
    I[2][1] = -(A[2][1] * A[0][0] - A[2][0] * A[0][1]);
    I[2][2] = A[1][1] * A[0][0] - A[1][0] * A[0][1];
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) J9[r * 3 + s] = I[r][s] / det;
    for (int d = 0; d < 3; ++d) o3[d] = v[0][d];
    detjc = det;
}

inline void to_local(const double *J9, const double *o, const double *x, double *y) {
    double q[3] = {x[0] - o[0], x[1] - o[1], x[2] - o[2]};
    for (int r = 0; r < 3; ++r) y[r] = J9[r * 3 + 0] * q[0] + J9[r * 3 + 1] * q[1] + J9[r * 3 + 2] * q[2];
}

// Node coordinates of stencil member j, unwrapped around cell i's centre
// and mapped into cell i's local frame.
void member_local_nodes(const ucnsp_case &c, int64_t i, int64_t j, const double *J9, const double *o, double v[8][3]) {
    const double *ci = &c.centre[i * 3];
    double cj[3] = {c.centre[j * 3], c.centre[j * 3 + 1], c.centre[j * 3 + 2]};
    double shift[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
        if (c.per[d] > 0 && std::fabs(cj[d] - ci[d]) > 0.5 * c.per[d])
            shift[d] = c.per[d] * ((ci[d] - 0.5 * c.per[d]) >= 0 ? 1.0 : -1.0);
    int nn = nnodes_of(c.cshape[j]);
    for (int k = 0; k < nn; ++k) {
        double p[3];
        node_xyz(c, c.cnodes[j * 8 + k], p);
        for (int d = 0; d < 3; ++d) p[d] += shift[d];
        to_local(J9, o, p, v[k]);
    }
}

// number of derivative multi-indices 1<=|a|<=p and D^a of monomial d at x
struct DerivTable {
    std::vector<int> alpha;  // na*3
    int na = 0;
};
DerivTable make_derivs(int p) {
    DerivTable t;
    for (int s = 1; s <= p; ++s)
        for (int a = s; a >= 0; --a)
            for (int b = s - a; b >= 0; --b) {
                int cz = s - a - b;
                t.alpha.push_back(a);
                t.alpha.push_back(b);
                t.alpha.push_back(cz);
                t.na++;
            }
    return t;
}
inline double falling(int e, int a) {  // e!/(e-a)!
    double r = 1.0;
    for (int k = 0; k < a; ++k) r *= (e - k);
    return r;
}

// Geometry signatures.  On unjittered meshes many cells are translates of
// each other; their PRESTORE data then agree to rounding, so the
// preprocessor solves one representative and copies it.  Two levels:
//   base signature    the cell's own node offsets from its centre and the
//                     adopted face-vertex order -> frame-relative data
//                     (INTEG_BASIS, VOLUME(1,1), INDICATOR, QPOINTS)
//   stencil signature base signature + the stencil's members' node offsets
//                     -> that stencil's pseudo-inverse
// Offsets are quantised at 1e-9 of the box size; equal hashes are verified
// element by element.  Never used on jittered meshes.
struct SigWriter {
    double q;
    uint64_t h = 1469598103934665603ULL;
    std::vector<int64_t> *out;
    void mix(int64_t v) {
        if (out) out->push_back(v);
        h ^= (uint64_t)v;
        h *= 1099511628211ULL;
    }
};

void sig_cell_nodes(const ucnsp_case &c, int64_t i, int64_t j, SigWriter &w) {
    const double *ci = &c.centre[i * 3];
    double shift[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
        if (c.per[d] > 0 && std::fabs(c.centre[j * 3 + d] - ci[d]) > 0.5 * c.per[d])
            shift[d] = c.per[d] * ((ci[d] - 0.5 * c.per[d]) >= 0 ? 1.0 : -1.0);
    w.mix(c.cshape[j]);
    for (int k = 0; k < nnodes_of(c.cshape[j]); ++k)
        for (int d = 0; d < 3; ++d) w.mix(std::llround((c.xyz[c.cnodes[j * 8 + k] * 3 + d] + shift[d] - ci[d]) / w.q));
}

uint64_t base_signature(const ucnsp_case &c, int64_t i, std::vector<int64_t> *out) {
    SigWriter w{1e-9 * std::max(c.o.lx, std::max(c.o.ly, c.o.lz)), 1469598103934665603ULL, out};
    sig_cell_nodes(c, i, i, w);
    const double *ci = &c.centre[i * 3];
    for (int f = 0; f < c.ifca[i]; ++f)
        for (int k = 0; k < 4; ++k) {
            int32_t n = c.fnodes[(i * ucnsp_case::NF + f) * 4 + k];
            if (n < 0) continue;
            double p[3] = {c.xyz[n * 3], c.xyz[n * 3 + 1], c.xyz[n * 3 + 2]};
            unwrap(c, ci, p);
            for (int d = 0; d < 3; ++d) w.mix(std::llround((p[d] - ci[d]) / w.q));
        }
    return w.h;
}

uint64_t stencil_signature(const ucnsp_case &c, int64_t i, int s, const std::vector<int64_t> &brep,
                           std::vector<int64_t> *out) {
    // cells are described by their verified geometry class (node offsets
    // from the centre) and members also by their centre offset
    SigWriter w{1e-9 * std::max(c.o.lx, std::max(c.o.ly, c.o.lz)), 1469598103934665603ULL, out};
    w.mix(brep[i]);
    const double *ci = &c.centre[i * 3];
    const int32_t *mem = &c.ihexl[((size_t)i * c.typesten + s) * c.nnb];
    for (int k = 1; k < c.nnb; ++k) {
        const int64_t j = mem[k];
        double cj[3] = {c.centre[j * 3], c.centre[j * 3 + 1], c.centre[j * 3 + 2]};
        unwrap(c, ci, cj);
        w.mix(brep[j]);
        for (int d = 0; d < 3; ++d) w.mix(std::llround((cj[d] - ci[d]) / w.q));
    }
    return w.h;
}

struct PrestoreScratch {
    std::vector<double> A, M, inv, work, mono, intg, mbar, volj, nzv, memint;
    std::vector<int> nzk;
    QPts q;
    PrestoreScratch(int ID, int IM) {
        A.resize((size_t)IM * ID);
        M.resize((size_t)ID * ID);
        inv.resize((size_t)ID * ID);
        work.resize(3 * ID * ID + ID + 8);
        mono.resize(ID);
        intg.resize(ID);
        mbar.resize(ID);
        volj.resize(IM);
        nzv.resize(ID);
        nzk.resize(ID);
        memint.resize((size_t)IM * (ID + 1));
    }
};

// Frame-relative data of one cell: INTEG_BASIS, VOLUME(1,1), INDICATOR
// (INDICATORMATRIX + PH*, src/prestore.f90:1529-2025) and the face Gauss
// points in the local frame (MEMORY_FAST, src/reconstruct_p.f90:178-244).
void base_cell(ucnsp_case &c, int64_t i, const DerivTable &dt, PrestoreScratch &S) {
    const int NF = ucnsp_case::NF;
    const int ID = c.ideg, NQ = c.nqpmax;
    QPts &q = S.q;
    const double *J9 = &c.frame[i * 12], *o = &c.frame[i * 12 + 9];
    const double adet = std::fabs(c.framedet[i]);
    double vi[8][3];
    member_local_nodes(c, i, i, J9, o, vi);
    // INTEG_BASIS: integral of the raw monomials (CALINTBASIS runs before
    // INTEG_BASIS is set, src/prestore.f90:222-225)
    cell_rule(c.cshape[i], vi, c.iorder, q);
    std::fill(S.intg.begin(), S.intg.end(), 0.0);
    for (size_t p = 0; p < q.w.size(); ++p) {
        monomials(q.x[p * 3], q.x[p * 3 + 1], q.x[p * 3 + 2], ID, S.mono.data());
        for (int d = 0; d < ID; ++d) S.intg[d] += q.w[p] * S.mono[d];
    }
    c.volume11[i] = c.vol[i] / adet;
    for (int d = 0; d < ID; ++d) c.integ_basis[i * ID + d] = S.intg[d];
    if (c.o.iweno == 1) {
        cell_rule(c.cshape[i], vi, 2 * (c.iorder - 1), q);
        double *B = &c.indicator[(size_t)i * ID * ID];
        std::fill(B, B + (size_t)ID * ID, 0.0);
        for (size_t p = 0; p < q.w.size(); ++p) {
            double x = q.x[p * 3], y = q.x[p * 3 + 1], z = q.x[p * 3 + 2];
            double px[5], py[5], pz[5];
            px[0] = py[0] = pz[0] = 1.0;
            for (int k = 1; k < 5; ++k) {
                px[k] = px[k - 1] * x;
                py[k] = py[k - 1] * y;
                pz[k] = pz[k - 1] * z;
            }
            for (int a = 0; a < dt.na; ++a) {
                int ax = dt.alpha[a * 3], ay = dt.alpha[a * 3 + 1], az = dt.alpha[a * 3 + 2];
                // D^a m_d is non-zero only for monomials that contain x^ax y^ay z^az
                int nz = 0;
                for (int d = 0; d < ID; ++d) {
                    int ex = MONO[d][0], ey = MONO[d][1], ez = MONO[d][2];
                    if (ex < ax || ey < ay || ez < az) continue;
                    S.nzk[nz] = d;
                    S.nzv[nz] = falling(ex, ax) * falling(ey, ay) * falling(ez, az) * px[ex - ax] * py[ey - ay] *
                                pz[ez - az];
                    ++nz;
                }
                const double wq = q.w[p];
                for (int u = 0; u < nz; ++u) {
                    const double wu = wq * S.nzv[u];
                    double *Brow = B + (size_t)S.nzk[u] * ID;
                    for (int t = 0; t < nz; ++t) Brow[S.nzk[t]] += wu * S.nzv[t];
                }
            }
        }
    }
    const double *ci = &c.centre[i * 3];
    for (int f = 0; f < c.ifca[i]; ++f) {
        int nv = c.types_faces[i * NF + f] == 5 ? 4 : 3;
        double fv[4][3], lv[4][3];
        for (int k = 0; k < nv; ++k) {
            node_xyz(c, c.fnodes[(i * NF + f) * 4 + k], fv[k]);
            unwrap(c, ci, fv[k]);
            to_local(J9, o, fv[k], lv[k]);
        }
        double *Q = &c.qpoints[((size_t)i * NF + f) * NQ * 3];
        if (nv == 4) {
            for (int g = 0; g < c.qp_quad; ++g) {
                double R = c.quad_rs[g * 2], Ss = c.quad_rs[g * 2 + 1];
                double N[4] = {0.25 * (1 - R) * (1 - Ss), 0.25 * (1 + R) * (1 - Ss), 0.25 * (1 + R) * (1 + Ss),
                               0.25 * (1 - R) * (1 + Ss)};
                for (int d = 0; d < 3; ++d)
                    Q[g * 3 + d] = N[0] * lv[0][d] + N[1] * lv[1][d] + N[2] * lv[2][d] + N[3] * lv[3][d];
            }
        } else {
            for (int g = 0; g < c.qp_tri; ++g) {
                const double *b = &c.tri_bary[g * 3];
                for (int d = 0; d < 3; ++d) Q[g * 3 + d] = b[0] * lv[0][d] + b[1] * lv[1][d] + b[2] * lv[2][d];
            }
        }
    }
}

// Pseudo-inverse of stencil s of cell i (PRESTORE_RECONSTRUCTION3,
// src/prestore.f90:285-529): A_jk = w_j * integral over V_j of (m_k - mbar_k),
// M = A^T A, Householder QR, invmat = R^-1 Q^T, P = invmat A^T with column j
// scaled by |V_j| w_j.  Returns false if the result has a NaN.
bool stencil_p(ucnsp_case &c, int64_t i, int s, PrestoreScratch &S) {
    const int IM = c.imax, ID = c.ideg, NNB = c.nnb, TS = c.typesten;
    QPts &q = S.q;
    const double *J9 = &c.frame[i * 12], *o = &c.frame[i * 12 + 9];
    const double adet = std::fabs(c.framedet[i]);
    const double vol11 = c.volume11[i];
    for (int d = 0; d < ID; ++d) S.mbar[d] = c.integ_basis[i * ID + d] / vol11;
    const int32_t *mem = &c.ihexl[((size_t)i * TS + s) * NNB];
    double *A = S.A.data();
    for (int k = 0; k < IM; ++k) {
        int64_t j = mem[k + 1];
        double wl = 1.0;
        if (c.o.weight_lsqr == 1) {
            double ci[3] = {c.centre[i * 3], c.centre[i * 3 + 1], c.centre[i * 3 + 2]};
            double cj[3] = {c.centre[j * 3], c.centre[j * 3 + 1], c.centre[j * 3 + 2]};
            unwrap(c, ci, cj);
            double dd = 0;
            for (int d = 0; d < 3; ++d) dd += (cj[d] - ci[d]) * (cj[d] - ci[d]);
            wl = 1.0 / std::sqrt(dd);
        }
        double vj[8][3];
        member_local_nodes(c, i, j, J9, o, vj);
        cell_rule(c.cshape[j], vj, c.iorder, q);
        double vq = 0.0;
        std::fill(S.intg.begin(), S.intg.end(), 0.0);
        for (size_t p = 0; p < q.w.size(); ++p) {
            monomials(q.x[p * 3], q.x[p * 3 + 1], q.x[p * 3 + 2], ID, S.mono.data());
            for (int d = 0; d < ID; ++d) S.intg[d] += q.w[p] * S.mono[d];
            vq += q.w[p];
        }
        // STENCILS(LL,K,:) = w * integral of (m - mbar) over V_j (COMPBASEL)
        for (int d = 0; d < ID; ++d) A[(size_t)k * ID + d] = wl * (S.intg[d] - S.mbar[d] * vq);
        S.volj[k] = (c.vol[j] / adet) * wl;  // ILOX_VOLUME * WLSQR
    }
    // LSCQM = A^T A (src/prestore.f90:437-476)
    double *M = S.M.data();
    for (int a2 = 0; a2 < ID; ++a2)
        for (int b2 = a2; b2 < ID; ++b2) {
            double s2 = 0.0;
            for (int k = 0; k < IM; ++k) s2 += A[(size_t)k * ID + a2] * A[(size_t)k * ID + b2];
            M[a2 * ID + b2] = s2;
            M[b2 * ID + a2] = s2;
        }
    qr_inverse(ID, M, S.inv.data(), S.work.data());
    bool bad = false;
    const double *inv = S.inv.data();
    double *P = c.invmat.p + ((size_t)i * TS + s) * IM * ID;
    for (int k = 0; k < IM; ++k)
        for (int a2 = 0; a2 < ID; ++a2) {
            double s2 = 0.0;
            for (int b2 = 0; b2 < ID; ++b2) s2 += inv[a2 * ID + b2] * A[(size_t)k * ID + b2];
            double val = s2 * S.volj[k];
            P[(size_t)k * ID + a2] = val;
            if (val != val) bad = true;
        }
    return !bad;
}

// representative of every item: rep[k] = smallest index with an equal
// (verified) signature
template <class Hash, class Full>
void classes(int64_t n, std::vector<int64_t> &rep, Hash hash, Full full) {
    std::vector<std::pair<uint64_t, int64_t>> sig(n);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) sig[k] = {hash(k), k};
    std::sort(sig.begin(), sig.end());
    rep.resize(n);
    for (int64_t k = 0; k < n; ++k) rep[k] = k;
    for (size_t a = 1; a < sig.size(); ++a)
        if (sig[a].first == sig[a - 1].first) rep[sig[a].second] = rep[sig[a - 1].second];
#pragma omp parallel
    {
        std::vector<int64_t> va, vb;
#pragma omp for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            if (rep[k] == k) continue;
            va.clear();
            vb.clear();
            full(k, &va);
            full(rep[k], &vb);
            if (va != vb) rep[k] = k;
        }
    }
}

// CWENO data of one mesh: INTEG_BASIS%VALUEC and INDICATORC (the low-order
// polynomial's moments and indicator in the cell's frame, CALINTBASIS with
// ICOMPWRT=1 and INDICATORMATRIX2, src/prestore.f90:244-263, 1700-1800), and
// INVMAT_STENCILTC of every sector stencil s >= 1 (its first imax2 members,
// the low-order basis with the VALUEC means; src/prestore.f90:296-337,
// 520-535).  Stencil 0 keeps the high-order data; its slot here is zero.
void build_cweno(ucnsp_case &c) {
    const int64_t nc = c.ncell;
    const int ID2 = c.ideg2, IM2 = c.imax2, TS = c.typesten, NNB = c.nnb;
    keep_assign(c.integ_basis_c, (size_t)nc * ID2, 0.0);
    keep_assign(c.indicator_c, (size_t)nc * ID2 * ID2, 0.0);
    if (c.invmat_c.n != (size_t)nc * TS * IM2 * ID2) c.invmat_c.alloc((size_t)nc * TS * IM2 * ID2);
    const DerivTable dt = make_derivs(c.iorder2);
#pragma omp parallel
    {
        QPts q;
        std::vector<double> mono(ID2), intg(ID2), mbar(ID2), A((size_t)IM2 * ID2), M((size_t)ID2 * ID2),
            inv((size_t)ID2 * ID2), work(3 * ID2 * ID2 + ID2 + 8), volj(IM2);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nc; ++i) {
            const double *J9 = &c.frame[i * 12], *o = &c.frame[i * 12 + 9];
            const double adet = std::fabs(c.framedet[i]);
            double vi[8][3];
            member_local_nodes(c, i, i, J9, o, vi);
            cell_rule(c.cshape[i], vi, c.iorder2, q);
            std::fill(intg.begin(), intg.end(), 0.0);
            for (size_t p = 0; p < q.w.size(); ++p) {
                monomials(q.x[p * 3], q.x[p * 3 + 1], q.x[p * 3 + 2], ID2, mono.data());
                for (int d = 0; d < ID2; ++d) intg[d] += q.w[p] * mono[d];
            }
            for (int d = 0; d < ID2; ++d) c.integ_basis_c[i * ID2 + d] = intg[d];
            // INDICATORC: sum over 1 <= |a| <= iorder2 of the Gram matrix of D^a m
            cell_rule(c.cshape[i], vi, 2 * (c.iorder2 - 1), q);
            double *B = &c.indicator_c[(size_t)i * ID2 * ID2];
            for (size_t p = 0; p < q.w.size(); ++p)
                for (int a = 0; a < dt.na; ++a) {
                    const int ax = dt.alpha[a * 3], ay = dt.alpha[a * 3 + 1], az = dt.alpha[a * 3 + 2];
                    double dv[34];
                    for (int d = 0; d < ID2; ++d) {
                        const int ex = MONO[d][0], ey = MONO[d][1], ez = MONO[d][2];
                        if (ex < ax || ey < ay || ez < az) {
                            dv[d] = 0.0;
                            continue;
                        }
                        dv[d] = falling(ex, ax) * falling(ey, ay) * falling(ez, az) *
                                std::pow(q.x[p * 3], ex - ax) * std::pow(q.x[p * 3 + 1], ey - ay) *
                                std::pow(q.x[p * 3 + 2], ez - az);
                    }
                    for (int u = 0; u < ID2; ++u)
                        for (int t = 0; t < ID2; ++t) B[u * ID2 + t] += q.w[p] * dv[u] * dv[t];
                }
            const double vol11 = c.volume11[i];
            for (int d = 0; d < ID2; ++d) mbar[d] = intg[d] / vol11;
            double *Pz = c.invmat_c.p + (size_t)i * TS * IM2 * ID2;
            std::fill(Pz, Pz + (size_t)IM2 * ID2, 0.0);  // stencil 0
            for (int st = 1; st < TS; ++st) {
                double *P = Pz + (size_t)st * IM2 * ID2;
                if (st >= c.admis[i]) {
                    std::fill(P, P + (size_t)IM2 * ID2, 0.0);
                    continue;
                }
                const int32_t *mem = &c.ihexl[((size_t)i * TS + st) * NNB];
                for (int k = 0; k < IM2; ++k) {
                    const int64_t j = mem[k + 1];
                    double vj[8][3];
                    member_local_nodes(c, i, j, J9, o, vj);
                    cell_rule(c.cshape[j], vj, c.iorder2, q);
                    double vq = 0.0;
                    std::fill(intg.begin(), intg.end(), 0.0);
                    for (size_t p = 0; p < q.w.size(); ++p) {
                        monomials(q.x[p * 3], q.x[p * 3 + 1], q.x[p * 3 + 2], ID2, mono.data());
                        for (int d = 0; d < ID2; ++d) intg[d] += q.w[p] * mono[d];
                        vq += q.w[p];
                    }
                    for (int d = 0; d < ID2; ++d) A[(size_t)k * ID2 + d] = intg[d] - mbar[d] * vq;
                    volj[k] = c.vol[j] / adet;
                }
                for (int a2 = 0; a2 < ID2; ++a2)
                    for (int b2 = a2; b2 < ID2; ++b2) {
                        double s2 = 0.0;
                        for (int k = 0; k < IM2; ++k) s2 += A[(size_t)k * ID2 + a2] * A[(size_t)k * ID2 + b2];
                        M[a2 * ID2 + b2] = s2;
                        M[b2 * ID2 + a2] = s2;
                    }
                qr_inverse(ID2, M.data(), inv.data(), work.data());
                bool bad = false;
                for (int k = 0; k < IM2; ++k)
                    for (int a2 = 0; a2 < ID2; ++a2) {
                        double s2 = 0.0;
                        for (int b2 = 0; b2 < ID2; ++b2) s2 += inv[a2 * ID2 + b2] * A[(size_t)k * ID2 + b2];
                        P[(size_t)k * ID2 + a2] = s2 * volj[k];
                        if (P[(size_t)k * ID2 + a2] != P[(size_t)k * ID2 + a2]) bad = true;
                    }
                if (bad) c.full[i] = 0;  // NaN stencils disable FULL (src/prestore.f90:541-544)
            }
        }
    }
}

void build_prestore(ucnsp_case &c) {
    const int64_t nc = c.ncell;
    const int NF = ucnsp_case::NF;
    const int TS = c.typesten, IM = c.imax, ID = c.ideg, NQ = c.nqpmax;
    if (c.invmat.n != (size_t)nc * TS * IM * ID) c.invmat.alloc((size_t)nc * TS * IM * ID);
    keep_assign(c.indicator, c.o.iweno == 1 ? (size_t)nc * ID * ID : 0, 0.0);
    keep_assign(c.integ_basis, (size_t)nc * ID, 0.0);
    keep_assign(c.volume11, nc, 0.0);
    keep_assign(c.qpoints, (size_t)nc * NF * NQ * 3, 0.0);
    keep_assign(c.frame, (size_t)nc * 12, 0.0);
    keep_assign(c.framedet, nc, 0.0);
    const DerivTable dt = make_derivs(c.iorder);
    const bool tim = std::getenv("UCNSP_TIMING") != nullptr;
    double t0 = omp_get_wtime();
    auto lap = [&](const char *what) {
        if (!tim) return;
        double t1 = omp_get_wtime();
        std::fprintf(stderr, "[ucnsp]   %-18s %8.3f s\n", what, t1 - t0);
        t0 = t1;
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        double J9[9], o[3], det;
        cell_frame(c, i, J9, o, det);
        for (int k = 0; k < 9; ++k) c.frame[i * 12 + k] = J9[k];
        for (int k = 0; k < 3; ++k) c.frame[i * 12 + 9 + k] = o[k];
        c.framedet[i] = det;
    }
    const bool cache = c.congruence_cache && c.o.jitter == 0.0;
    lap("frames");
    // 1. frame-relative per-cell data
    std::vector<int64_t> brep;
    std::vector<uint64_t> bhash(nc, 0);
    if (cache) {
        classes(nc, brep, [&](int64_t i) { return bhash[i] = base_signature(c, i, nullptr); },
                [&](int64_t i, std::vector<int64_t> *v) { base_signature(c, i, v); });
    } else {
        brep.resize(nc);
        for (int64_t i = 0; i < nc; ++i) brep[i] = i;
    }
    lap("base classes");
#pragma omp parallel
    {
        PrestoreScratch S(ID, IM);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nc; ++i)
            if (brep[i] == i) base_cell(c, i, dt, S);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            const int64_t r = brep[i];
            if (r == i) continue;
            if (c.o.iweno == 1)
                std::memcpy(&c.indicator[(size_t)i * ID * ID], &c.indicator[(size_t)r * ID * ID], sizeof(double) * ID * ID);
            std::memcpy(&c.integ_basis[(size_t)i * ID], &c.integ_basis[(size_t)r * ID], sizeof(double) * ID);
            c.volume11[i] = c.volume11[r];
            const size_t nq = (size_t)NF * NQ * 3;
            std::memcpy(&c.qpoints[(size_t)i * nq], &c.qpoints[(size_t)r * nq], sizeof(double) * nq);
        }
    }
    lap("base data");
    // 2. pseudo-inverses, one per stencil class
    const int64_t ns = nc * TS;
    std::vector<int64_t> srep, grep;
    if (cache) {
        const double qg = 1e-9 * std::max(c.o.lx, std::max(c.o.ly, c.o.lz));
        classes(nc, grep,
                [&](int64_t i) {
                    SigWriter w{qg, 1469598103934665603ULL, nullptr};
                    sig_cell_nodes(c, i, i, w);
                    return w.h;
                },
                [&](int64_t i, std::vector<int64_t> *v) {
                    SigWriter w{qg, 1469598103934665603ULL, v};
                    sig_cell_nodes(c, i, i, w);
                });
        classes(ns, srep,
                [&](int64_t k) {
                    const int64_t i = k / TS;
                    const int s = (int)(k % TS);
                    if (s >= c.admis[i]) return (uint64_t)0;
                    return stencil_signature(c, i, s, grep, nullptr);
                },
                [&](int64_t k, std::vector<int64_t> *v) {
                    const int64_t i = k / TS;
                    const int s = (int)(k % TS);
                    if (s >= c.admis[i]) {
                        v->push_back(-1);
                        return;
                    }
                    stencil_signature(c, i, s, grep, v);
                });
    } else {
        srep.resize(ns);
        for (int64_t k = 0; k < ns; ++k) srep[k] = k;
    }
    lap("stencil classes");
    std::vector<char> ok(ns, 1);
    int64_t solved = 0;
#pragma omp parallel reduction(+ : solved)
    {
        PrestoreScratch S(ID, IM);
#pragma omp for schedule(dynamic, 64)
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t i = k / TS;
            const int s = (int)(k % TS);
            // CWENO: the sector stencils' high-order pseudo-inverses are not
            // formed (their low-order ones are, by build_cweno)
            if (s >= c.admis[i] || (c.ideg2 > 0 && s > 0)) {
                double *P = c.invmat.p + (size_t)k * IM * ID;
                std::fill(P, P + (size_t)IM * ID, 0.0);
                continue;
            }
            if (srep[k] != k) continue;
            ok[k] = stencil_p(c, i, s, S) ? 1 : 0;
            ++solved;
        }
#pragma omp for schedule(static)
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t r = srep[k];
            const int64_t i = k / TS;
            if (r == k || (int)(k % TS) >= c.admis[i] || (c.ideg2 > 0 && k % TS > 0)) continue;
            std::memcpy(c.invmat.p + (size_t)k * IM * ID, c.invmat.p + (size_t)r * IM * ID, sizeof(double) * IM * ID);
            ok[k] = ok[r];
        }
    }
    lap("pseudo-inverses");
    c.n_solved = solved;
    if (c.ideg2 > 0) build_cweno(c);
    for (int64_t k = 0; k < ns; ++k)
        if (!ok[k]) c.full[k / TS] = 0;  // NaN stencils disable FULL (src/prestore.f90:526-529)
}

// The lower-wall bump of the channel configuration, faded out towards y = ly;
// applied after the box sides are classified (on the undeformed box).
void apply_bump(ucnsp_case &c) {
    const ucnsp_opts &o = c.o;
    bool bumped = false;
    for (int q = 0; q < 8; ++q) bumped = bumped || o.bump[q] != 0.0;
    if (!bumped) return;
    for (int64_t n = 0; n < c.nnode; ++n) {
        double *p = &c.xyz[n * 3];
        const double xh = 4.0 * p[0] / o.lx, zh = p[2] / o.lz;
        if (xh < 1.0 || xh > 3.0) continue;
        double w = 0.0;
        for (int q = 0; q < 8; ++q) w += o.bump[q] * std::sin((q + 1) * M_PI * (xh - 1.0) / 2.0);
        p[1] += o.ly * (1.0 - p[1] / o.ly) * w * std::sin(M_PI * zh);
    }
}

// signed volumes (x6) of every cell's DECOMPOSE3 sub-tetrahedra, [cell][6]
std::vector<double> subtet_orientation(const ucnsp_case &c) {
    std::vector<double> o(c.ncell * 6, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < c.ncell; ++i) {
        const int sh = c.cshape[i];
        double v[8][3];
        for (int k = 0; k < nnodes_of(sh); ++k) node_xyz(c, c.cnodes[i * 8 + k], v[k]);
        for (int t = 0; t < nsubtets(sh); ++t) {
            const int *tt = subtet(sh, t);
            const double *a = v[tt[0]], *b = v[tt[1]], *cc = v[tt[2]], *d = v[tt[3]];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {cc[0] - a[0], cc[1] - a[1], cc[2] - a[2]};
            const double e3[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
            o[i * 6 + t] = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                           e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
        }
    }
    return o;
}

// the first cell one of whose sub-tetrahedra the bump turned inside out or
// flattened (its signed volume changed sign or vanished), or -1
int64_t first_inverted_cell(const ucnsp_case &c, const std::vector<double> &before) {
    const std::vector<double> after = subtet_orientation(c);
    for (int64_t i = 0; i < c.ncell; ++i)
        for (int t = 0; t < nsubtets(c.cshape[i]); ++t)
            if (!(before[i * 6 + t] * after[i * 6 + t] > 0.0)) return i;
    return -1;
}

void gen_box(ucnsp_case &c) {
    const ucnsp_opts &o = c.o;
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const double hx = o.lx / nx, hy = o.ly / ny, hz = o.lz / nz;
    const bool pax[3] = {o.bc[0] == 0, o.bc[2] == 0, o.bc[4] == 0};  // periodic axes
    c.per[0] = pax[0] ? o.lx : 0.0;
    c.per[1] = pax[1] ? o.ly : 0.0;
    c.per[2] = pax[2] ? o.lz : 0.0;
    c.nnode = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    c.xyz.resize(c.nnode * 3);
    c.canon.resize(c.nnode);
    auto nid = [&](int i, int j, int k) { return (int64_t)i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k); };
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int64_t n = nid(i, j, k);
                double p[3] = {i * hx, j * hy, k * hz};
                if (o.jitter > 0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
                    uint64_t s = o.seed * 0x100000001B3ULL + (uint64_t)n * 3;
                    p[0] += o.jitter * hx * (2.0 * u01(s) - 1.0);
                    p[1] += o.jitter * hy * (2.0 * u01(s) - 1.0);
                    p[2] += o.jitter * hz * (2.0 * u01(s) - 1.0);
                }
                c.xyz[n * 3] = p[0];
                c.xyz[n * 3 + 1] = p[1];
                c.xyz[n * 3 + 2] = p[2];
                // nodes on opposite sides of a periodic axis share a canonical id
                const int ci = pax[0] ? i % nx : i, cj = pax[1] ? j % ny : j, ck = pax[2] ? k % nz : k;
                c.canon[n] = (int32_t)(ci + (nx + 1) * (cj + (int64_t)(ny + 1) * ck));
            }
    const int64_t nhex = (int64_t)nx * ny * nz;
    const int64_t ngrid = c.nnode;
    // element type of hex (i, j): one type; for the hybrid box hexahedra,
    // pyramids and prisms in thirds of x (their shared faces are the hexes'
    // quadrilaterals, so the mesh is conforming); for the channel prisms in
    // the wall layers j < channel_layers(ny), the interface layer of
    // pyramids and tetrahedra (UCNSP_INTERFACE) and Kuhn tetrahedra above
    const int layers = channel_layers(ny);
    // blast box: hexahedra in [c0, c1) on every axis, the shell one hex
    // around them, prism layers at the z ends
    const int blast_pl = std::max(1, nz / 10);
    const int bc0[3] = {nx / 4, ny / 4, nz / 4}, bc1[3] = {nx - nx / 4, ny - ny / 4, nz - nz / 4};
    auto etype = [&](int i, int j, int k) {
        if (o.shape == UCNSP_BLAST) {
            const int q[3] = {i, j, k};
            bool core = true, near = true;
            for (int d = 0; d < 3; ++d) {
                core = core && q[d] >= bc0[d] && q[d] < bc1[d];
                near = near && q[d] >= bc0[d] - 1 && q[d] <= bc1[d];
            }
            if (core) return UCNSP_HEX;
            if (near) return UCNSP_SHELL;
            return k < blast_pl || k >= nz - blast_pl ? UCNSP_PRISM : UCNSP_TET;
        }
        if (o.shape == UCNSP_CHANNEL) return j < layers ? UCNSP_PRISM : j == layers ? UCNSP_INTERFACE : UCNSP_TET;
        return o.shape != UCNSP_HYBRID ? o.shape
               : i < nx / 3            ? UCNSP_HEX
               : i < 2 * nx / 3        ? UCNSP_PYRAMID
                                       : UCNSP_PRISM;
    };
    // shell hexes: which faces (HEX_FACES order) meet triangle-faced neighbours
    auto shell_tri = [&](int i, int j, int k) {
        const int d[6][3] = {{0, 0, -1}, {0, 0, 1}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
        const int n3[3] = {nx, ny, nz};
        int mask = 0;
        for (int f = 0; f < 6; ++f) {
            int q[3] = {i + d[f][0], j + d[f][1], k + d[f][2]};
            bool out = false;
            for (int a = 0; a < 3; ++a) {
                if (q[a] >= 0 && q[a] < n3[a]) continue;
                if (pax[a]) q[a] = (q[a] + n3[a]) % n3[a];
                else out = true;
            }
            if (out) continue;
            const int t = etype(q[0], q[1], q[2]);
            if (t == UCNSP_TET || t == UCNSP_PRISM) mask |= 1 << f;
        }
        return mask;
    };
    auto per_hex = [&](int t, int i, int j, int k) {
        if (t == UCNSP_SHELL) return 6 + __builtin_popcount(shell_tri(i, j, k));
        return t == UCNSP_HEX ? 1 : t == UCNSP_PRISM ? 2 : t == UCNSP_INTERFACE ? 7 : 6;
    };
    auto has_apex = [&](int i, int j, int k) {
        const int t = etype(i, j, k);
        return t == UCNSP_PYRAMID || t == UCNSP_INTERFACE || t == UCNSP_SHELL;
    };
    if (o.shape == UCNSP_PYRAMID || o.shape == UCNSP_HYBRID || o.shape == UCNSP_CHANNEL || o.shape == UCNSP_BLAST) {
        // one apex node per hex that holds pyramids (the node slot of every
        // hex is reserved), at the mean of its corners
        c.nnode += nhex;
        c.xyz.resize(c.nnode * 3);
        c.canon.resize(c.nnode);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t h = i + (int64_t)nx * (j + (int64_t)ny * k), n = ngrid + h;
                    c.canon[n] = (int32_t)n;
                    if (!has_apex(i, j, k)) {
                        for (int d = 0; d < 3; ++d) c.xyz[n * 3 + d] = 0.0;
                        continue;
                    }
                    double m[3] = {0, 0, 0};
                    for (int b = 0; b < 8; ++b) {
                        const int64_t q = nid(i + (b & 1), j + ((b >> 1) & 1), k + (b >> 2));
                        for (int d = 0; d < 3; ++d) m[d] += 0.125 * c.xyz[q * 3 + d];
                    }
                    for (int d = 0; d < 3; ++d) c.xyz[n * 3 + d] = m[d];
                }
    }
    c.shape = o.shape;
    c.ncell = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) c.ncell += per_hex(etype(i, j, k), i, j, k);
    c.cnodes.assign(c.ncell * 8, -1);
    c.cshape.assign(c.ncell, o.shape);
    int64_t cell0 = 0;  // first cell of the current hex
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int64_t h = i + (int64_t)nx * (j + (int64_t)ny * k);
                int64_t v[8] = {nid(i, j, k),     nid(i + 1, j, k),     nid(i + 1, j + 1, k),     nid(i, j + 1, k),
                                nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
                const int et = etype(i, j, k);
                const int nper = per_hex(et, i, j, k);
                for (int q = 0; q < nper; ++q) c.cshape[cell0 + q] = et;
                // a tetrahedron from four nodes, numbered with positive orientation
                auto add_tet = [&](int64_t cell, int64_t a, int64_t b, int64_t cc, int64_t d) {
                    double x[4][3];
                    const int64_t tv0[4] = {a, b, cc, d};
                    for (int q = 0; q < 4; ++q)
                        for (int e = 0; e < 3; ++e) x[q][e] = c.xyz[tv0[q] * 3 + e];
                    double e1[3], e2[3], e3[3];
                    for (int e = 0; e < 3; ++e) {
                        e1[e] = x[1][e] - x[0][e];
                        e2[e] = x[2][e] - x[0][e];
                        e3[e] = x[3][e] - x[0][e];
                    }
                    const double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                       e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
                    int64_t tv[4] = {a, b, cc, d};
                    if (det < 0) std::swap(tv[1], tv[2]);
                    c.cshape[cell] = UCNSP_TET;
                    for (int q = 0; q < 4; ++q) c.cnodes[cell * 8 + q] = (int32_t)tv[q];
                };
                if (et == UCNSP_SHELL) {
                    // a pyramid per quadrilateral-faced neighbour, two
                    // tetrahedra along the Kuhn diagonal per triangle-faced one
                    const int mask = shell_tri(i, j, k);
                    int64_t cell = cell0;
                    for (int f = 0; f < 6; ++f) {
                        if (!(mask & (1 << f))) {
                            c.cshape[cell] = UCNSP_PYRAMID;
                            for (int q = 0; q < 4; ++q) c.cnodes[cell * 8 + q] = (int32_t)v[HEX_FACES[f][3 - q]];
                            c.cnodes[cell * 8 + 4] = (int32_t)(ngrid + h);
                            ++cell;
                            continue;
                        }
                        // the face's corners in cyclic order, the diagonal's ends a and c
                        const int *fc = HEX_FACES[f];
                        int ia = 0;
                        while (fc[ia] != FACE_DIAG[f][0]) ++ia;
                        const int a = fc[ia], b = fc[(ia + 1) % 4], cc = fc[(ia + 2) % 4], d = fc[(ia + 3) % 4];
                        add_tet(cell++, v[a], v[b], v[cc], ngrid + h);
                        add_tet(cell++, v[a], v[cc], v[d], ngrid + h);
                    }
                } else if (et == UCNSP_INTERFACE) {
                    // between the prism layers (quadrilateral faces at y = j)
                    // and the Kuhn tetrahedra (triangles at y = j + 1): five
                    // pyramids on the hex faces other than y = j + 1, whose
                    // quadrilateral is split along the diagonal 3-6 (the one
                    // the Kuhn split of the hex above takes, through its
                    // corner (0,0,0)) into two tetrahedra with the same apex
                    const int pyr_faces[5] = {5, 0, 1, 2, 3};  // y = j, z = k, z = k+1, x = i+1, x = i
                    for (int t = 0; t < 5; ++t) {
                        const int64_t cell = cell0 + t;
                        c.cshape[cell] = UCNSP_PYRAMID;
                        for (int q = 0; q < 4; ++q) c.cnodes[cell * 8 + q] = (int32_t)v[HEX_FACES[pyr_faces[t]][3 - q]];
                        c.cnodes[cell * 8 + 4] = (int32_t)(ngrid + h);
                    }
                    add_tet(cell0 + 5, v[3], v[2], v[6], ngrid + h);
                    add_tet(cell0 + 6, v[3], v[6], v[7], ngrid + h);
                } else if (et == UCNSP_HEX) {
                    for (int q = 0; q < 8; ++q) c.cnodes[cell0 * 8 + q] = (int32_t)v[q];
                } else if (et == UCNSP_PYRAMID) {
                    // six pyramids per hex: each hex face is a base (reversed, so
                    // that its normal points to the apex), the hex centre the apex
                    for (int f = 0; f < 6; ++f) {
                        const int64_t cell = cell0 + f;
                        for (int q = 0; q < 4; ++q) c.cnodes[cell * 8 + q] = (int32_t)v[HEX_FACES[f][3 - q]];
                        c.cnodes[cell * 8 + 4] = (int32_t)(ngrid + h);
                    }
                } else if (et == UCNSP_PRISM) {
                    // two prisms per hex, split along the bottom face's 0-2 diagonal;
                    // bottom triangle counter-clockwise seen from +z, top above it
                    const int pr[2][6] = {{0, 1, 2, 4, 5, 6}, {0, 2, 3, 4, 6, 7}};
                    for (int t = 0; t < 2; ++t)
                        for (int q = 0; q < 6; ++q) c.cnodes[(cell0 + t) * 8 + q] = (int32_t)v[pr[t][q]];
                } else {
                    // Kuhn split: 6 tets along the 000-111 diagonal, one per axis order
                    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                    auto corner = [&](int bx, int by, int bz) {
                        const int idx[2][2][2] = {{{0, 4}, {3, 7}}, {{1, 5}, {2, 6}}};
                        return v[idx[bx][by][bz]];
                    };
                    for (int t = 0; t < 6; ++t) {
                        int b[3] = {0, 0, 0};
                        int64_t tv[4];
                        tv[0] = corner(0, 0, 0);
                        b[perm[t][0]] = 1;
                        tv[1] = corner(b[0], b[1], b[2]);
                        b[perm[t][1]] = 1;
                        tv[2] = corner(b[0], b[1], b[2]);
                        tv[3] = corner(1, 1, 1);
                        // positive orientation (odd axis orders swap nodes 2,3)
                        double a0[3], a1[3], a2[3], a3[3];
                        for (int d = 0; d < 3; ++d) {
                            a0[d] = c.xyz[tv[0] * 3 + d];
                            a1[d] = c.xyz[tv[1] * 3 + d];
                            a2[d] = c.xyz[tv[2] * 3 + d];
                            a3[d] = c.xyz[tv[3] * 3 + d];
                        }
                        double e1[3], e2[3], e3[3];
                        for (int d = 0; d < 3; ++d) {
                            e1[d] = a1[d] - a0[d];
                            e2[d] = a2[d] - a0[d];
                            e3[d] = a3[d] - a0[d];
                        }
                        double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                     e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
                        if (det < 0) std::swap(tv[1], tv[2]);
                        int64_t cell = cell0 + t;
                        for (int q = 0; q < 4; ++q) c.cnodes[cell * 8 + q] = (int32_t)tv[q];
                    }
                }
                cell0 += nper;
            }
}

}  // namespace

// Everything after the mesh: connectivity, the bump, geometry, scheme sizes,
// stencils and PRESTORE.  Deletes c on failure.
static int build_after_mesh(ucnsp_case *c, bool bumped) {
    const ucnsp_opts &o = c->o;
    const bool tim = std::getenv("UCNSP_TIMING") != nullptr;
    double t0 = omp_get_wtime();
    auto lap = [&](const char *what) {
        if (!tim) return;
        double t1 = omp_get_wtime();
        std::fprintf(stderr, "[ucnsp] %-14s %8.3f s\n", what, t1 - t0);
        t0 = t1;
    };
    build_connectivity(*c);
    lap("connectivity");
    for (int64_t i = 0; i < c->ncell; ++i)
        for (int f = 0; f < c->ifca[i]; ++f)
            if (c->ineigh[i * ucnsp_case::NF + f] < 0 && c->bcode[i * ucnsp_case::NF + f] <= 0) {
                delete c;
                return fail(UCNS_EBADARG, c->from_file ? "a boundary face of the mesh is missing from GRID.bnd"
                                                       : "unmatched face off the box sides");
            }
    std::vector<double> orient;
    if (bumped) orient = subtet_orientation(*c);
    c->xyz_box = c->xyz;
    apply_bump(*c);
    if (bumped) {
        const int64_t bad = first_inverted_cell(*c, orient);
        if (bad >= 0) {
            delete c;
            return fail(UCNS_EBADARG, "the lower-wall bump inverts cell " + std::to_string(bad) +
                                          " (a sub-tetrahedron of non-positive volume): smaller a_k or more cells");
        }
    }
    build_geometry(*c);
    lap("geometry");
    // scheme sizes (DETERMINE_SIZE, src/grid_p.f90:6097-6288; parameters.f90:1270-1394)
    c->iorder = o.iorder;
    c->ideg = ndeg_of(o.iorder);
    const int ilx = c->ideg + 1;
    int iextend;
    if (o.iweno == 1) {
        c->nnb = ilx * o.extf;
        c->typesten = 7;
        iextend = 12;
        int64_t pool = (int64_t)ilx * o.extf * iextend;
        c->iselem = (int)std::min<int64_t>(pool, c->ncell - 1);
    } else {
        c->nnb = 9;
        c->typesten = 1;
        iextend = 2;
        c->iselem = ilx * o.extf * iextend;
    }
    if (c->iselem < c->nnb) {
        delete c;
        return fail(UCNS_EBADARG, "mesh too small for the stencil size");
    }
    c->imax = c->nnb - 1;
    if (o.ees == 5) {  // CWENO sectors (DETERMINE_SIZE, src/grid_p.f90:6151-6163)
        if (o.iweno != 1) {
            delete c;
            return fail(UCNS_EBADARG, "CWENO (ees 5) needs the WENO path");
        }
        c->iorder2 = o.iorder <= 3 ? 1 : 2;
        c->ideg2 = ndeg_of(c->iorder2);
        c->imax2 = (c->iorder2 == 1 ? 12 : 20) - 1;
        if (c->imax2 > c->imax) {
            delete c;
            return fail(UCNS_EUNSUPPORTED, "the sector stencils are shorter than the low-order CWENO stencils");
        }
    } else if (o.ees != 0) {
        delete c;
        return fail(UCNS_EUNSUPPORTED, "ees must be 0 (WENO) or 5 (CWENO)");
    }
    c->igq = std::min(o.iorder, 6);
    c->qp_quad = quad_rule(c->igq, c->quad_rs, c->quad_w);
    c->qp_tri = tri_rule(c->igq, c->tri_bary, c->tri_w);
    if (c->qp_quad < 0 || c->qp_tri < 0) {
        delete c;
        return fail(UCNS_EUNSUPPORTED, "face quadrature rule not available");
    }
    c->nqpmax = std::max(c->qp_quad, c->qp_tri);
    build_stencils(*c);
    // leading stencil dimension = the largest ADMIS present (tets use 5 of
    // the 7 slots of TYPESTEN); complete stencils are already compacted
    {
        int ts = 1;
        for (int64_t i = 0; i < c->ncell; ++i) ts = std::max(ts, (int)c->admis[i]);
        if (ts < c->typesten) {
            std::vector<int32_t> h((size_t)c->ncell * ts * c->nnb);
            for (int64_t i = 0; i < c->ncell; ++i)
                std::memcpy(&h[(size_t)i * ts * c->nnb], &c->ihexl[(size_t)i * c->typesten * c->nnb],
                            sizeof(int32_t) * ts * c->nnb);
            c->ihexl.swap(h);
            c->typesten = ts;
        }
    }
    lap("stencils");
    build_prestore(*c);
    lap("prestore");
    return UCNS_OK;
}

namespace {

// Fortran unformatted sequential records (BINIO = 1): a 4-byte length before
// and after each record's payload
void put_rec(FILE *f, const void *p, int32_t n) {
    std::fwrite(&n, 4, 1, f);
    std::fwrite(p, 1, (size_t)n, f);
    std::fwrite(&n, 4, 1, f);
}
bool get_rec(FILE *f, void *p, int32_t n) {
    int32_t a = 0, b = 0;
    if (std::fread(&a, 4, 1, f) != 1 || a != n) return false;
    if (std::fread(p, 1, (size_t)n, f) != (size_t)n) return false;
    return std::fread(&b, 4, 1, f) == 1 && b == n;
}

// GRID.cel node lists: eight ids with repeats that say the shape (the rule
// of src/grid_p.f90:1670-1684), and the shape's own node list in this
// preprocessor's 8-slot order
int cel_shape(const int32_t *I) {
    if (I[4] != I[5] && I[5] != I[6] && I[6] != I[7] && I[2] != I[3]) return UCNSP_HEX;
    if (I[2] == I[3] && I[4] == I[5] && I[5] == I[6] && I[6] == I[7]) return UCNSP_TET;
    if (I[2] != I[3] && I[4] == I[5] && I[5] == I[6] && I[6] == I[7]) return UCNSP_PYRAMID;
    if (I[2] == I[3] && I[4] != I[5] && I[6] == I[7]) return UCNSP_PRISM;
    return 0;
}
// slots of I1..I8 that hold the shape's nodes, in node order
const int CEL_SLOTS[5][8] = {{0}, {0, 1, 2, 3, 4, 5, 6, 7}, {0, 1, 2, 4}, {0, 1, 2, 4, 5, 6}, {0, 1, 2, 3, 4}};
// I1..I8 from a shape's nodes n[]: tetrahedron n0 n1 n2 n2 n3 n3 n3 n3,
// pyramid n0..n3 n4 n4 n4 n4, prism n0 n1 n2 n2 n3 n4 n5 n5
void cel_line(int shape, const int32_t *n, int32_t *I) {
    switch (shape) {
    case UCNSP_HEX:
        for (int k = 0; k < 8; ++k) I[k] = n[k];
        break;
    case UCNSP_TET: {
        const int32_t t[8] = {n[0], n[1], n[2], n[2], n[3], n[3], n[3], n[3]};
        std::memcpy(I, t, sizeof(t));
        break;
    }
    case UCNSP_PYRAMID: {
        const int32_t t[8] = {n[0], n[1], n[2], n[3], n[4], n[4], n[4], n[4]};
        std::memcpy(I, t, sizeof(t));
        break;
    }
    default: {
        const int32_t t[8] = {n[0], n[1], n[2], n[2], n[3], n[4], n[5], n[5]};
        std::memcpy(I, t, sizeof(t));
    }
    }
}

struct FileCloser {
    FILE *f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int ucnsp_write_grid(const ucnsp_case *c, const char *dir, int32_t binary) {
    if (!c || !dir) return fail(UCNS_EBADARG, "null argument");
    const std::string d(dir);
    const char *mode = binary ? "wb" : "w";
    FileCloser cel{std::fopen((d + "/GRID.cel").c_str(), mode)}, vrt{std::fopen((d + "/GRID.vrt").c_str(), mode)},
        bnd{std::fopen((d + "/GRID.bnd").c_str(), mode)};
    if (!cel.f || !vrt.f || !bnd.f) return fail(UCNS_EBADARG, "cannot create GRID.cel/.vrt/.bnd in " + d);
    // nodes (1-based ids; %.17g keeps every double exactly)
    for (int64_t n = 0; n < c->nnode; ++n) {
        const int32_t id = (int32_t)(n + 1);
        const double *x = &c->xyz[(size_t)n * 3];
        if (binary) {
            char buf[28];
            std::memcpy(buf, &id, 4);
            std::memcpy(buf + 4, x, 24);
            put_rec(vrt.f, buf, 28);
        } else {
            std::fprintf(vrt.f, "%d %.17g %.17g %.17g\n", id, x[0], x[1], x[2]);
        }
    }
    // cells
    for (int64_t i = 0; i < c->ncell; ++i) {
        int32_t n[8], I[9];
        for (int k = 0; k < 8; ++k) n[k] = c->cnodes[i * 8 + k] + 1;
        I[0] = (int32_t)(i + 1);
        cel_line(c->cshape[i], n, I + 1);
        if (binary)
            put_rec(cel.f, I, 36);
        else
            std::fprintf(cel.f, "%d %d %d %d %d %d %d %d %d\n", I[0], I[1], I[2], I[3], I[4], I[5], I[6], I[7], I[8]);
    }
    // boundary faces (codes 1, 2, 3, 4, 6) and both sides of periodic pairs (5, 50)
    int32_t id = 0;
    for (int64_t i = 0; i < c->ncell; ++i)
        for (int f = 0; f < c->ifca[i]; ++f) {
            const int64_t sl = i * ucnsp_case::NF + f;
            const int code = c->bcode[sl];
            if (code == 0) continue;
            const int sh = c->cshape[i], nv = face_nv(sh, f);
            int32_t B[6];
            B[0] = ++id;
            for (int q = 0; q < 4; ++q) B[1 + q] = c->cnodes[i * 8 + face_node(sh, f, q < nv ? q : nv - 1)] + 1;
            B[5] = code;
            if (binary)
                put_rec(bnd.f, B, 24);
            else
                std::fprintf(bnd.f, "%d %d %d %d %d %d\n", B[0], B[1], B[2], B[3], B[4], B[5]);
        }
    return UCNS_OK;
}

int ucnsp_read_grid(const char *dir, int32_t binary, const ucnsp_opts *opts, ucnsp_case **out) {
    if (!dir || !opts || !out) return fail(UCNS_EBADARG, "null argument");
    const ucnsp_opts &o = *opts;
    if (o.iorder < 1 || o.iorder > 4) return fail(UCNS_EUNSUPPORTED, "iorder must be 1..4");
    if (o.iweno != 1 && o.iweno != -1) return fail(UCNS_EBADARG, "iweno must be 1 or -1");
    if (o.iweno == -1 && o.iorder != 1) return fail(UCNS_EUNSUPPORTED, "MUSCL path supports iorder 1");
    if (o.iweno == 1 && o.iorder < 2) return fail(UCNS_EUNSUPPORTED, "WENO needs iorder >= 2");
    const std::string d(dir);
    const char *mode = binary ? "rb" : "r";
    FileCloser cel{std::fopen((d + "/GRID.cel").c_str(), mode)}, vrt{std::fopen((d + "/GRID.vrt").c_str(), mode)},
        bnd{std::fopen((d + "/GRID.bnd").c_str(), mode)};
    if (!cel.f || !vrt.f || !bnd.f) return fail(UCNS_EBADARG, "cannot open GRID.cel/.vrt/.bnd in " + d);
    if (o.nthreads > 0) omp_set_num_threads(o.nthreads);
    std::unique_ptr<ucnsp_case> c(new ucnsp_case());
    c->o = o;
    for (int q = 0; q < 8; ++q) c->o.bump[q] = 0.0;
    c->o.jitter = 0.0;
    c->from_file = true;
    c->shape = UCNSP_HYBRID;
    // nodes: ids 1..N in order
    for (;;) {
        int32_t id;
        double x[3];
        if (binary) {
            char buf[28];
            if (!get_rec(vrt.f, buf, 28)) break;
            std::memcpy(&id, buf, 4);
            std::memcpy(x, buf + 4, 24);
        } else if (std::fscanf(vrt.f, "%d %lf %lf %lf", &id, &x[0], &x[1], &x[2]) != 4) {
            break;
        }
        if (id != (int32_t)(c->nnode + 1)) return fail(UCNS_EBADARG, "GRID.vrt: node ids must run 1, 2, 3, ...");
        c->xyz.insert(c->xyz.end(), x, x + 3);
        ++c->nnode;
    }
    if (c->nnode < 4) return fail(UCNS_EBADARG, "GRID.vrt: no nodes");
    // cells
    for (;;) {
        int32_t I[9];
        if (binary) {
            if (!get_rec(cel.f, I, 36)) break;
        } else if (std::fscanf(cel.f, "%d %d %d %d %d %d %d %d %d", &I[0], &I[1], &I[2], &I[3], &I[4], &I[5], &I[6],
                               &I[7], &I[8]) != 9) {
            break;
        }
        if (I[0] != (int32_t)(c->ncell + 1)) return fail(UCNS_EBADARG, "GRID.cel: cell ids must run 1, 2, 3, ...");
        const int sh = cel_shape(I + 1);
        if (!sh) return fail(UCNS_EBADARG, "GRID.cel: cell " + std::to_string(I[0]) + " has no element shape");
        int32_t n[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
        for (int k = 0; k < nnodes_of(sh); ++k) {
            const int32_t v = I[1 + CEL_SLOTS[sh][k]];
            if (v < 1 || v > c->nnode) return fail(UCNS_EBADARG, "GRID.cel: node id out of range");
            n[k] = v - 1;
        }
        c->cnodes.insert(c->cnodes.end(), n, n + 8);
        c->cshape.push_back(sh);
        ++c->ncell;
    }
    if (c->ncell < 1) return fail(UCNS_EBADARG, "GRID.cel: no cells");
    // boundary faces
    std::vector<std::array<int32_t, 5>> bf;
    for (;;) {
        int32_t B[6];
        if (binary) {
            if (!get_rec(bnd.f, B, 24)) break;
        } else if (std::fscanf(bnd.f, "%d %d %d %d %d %d", &B[0], &B[1], &B[2], &B[3], &B[4], &B[5]) != 6) {
            break;
        }
        const int code = B[5];
        if (code != 1 && code != 2 && code != 3 && code != 4 && code != 5 && code != 6 && code != 50)
            return fail(UCNS_EUNSUPPORTED, "GRID.bnd: codes 1, 2, 3, 4, 5, 6 and 50");
        std::array<int32_t, 4> k;
        const int nv = B[3] == B[4] ? 3 : 4;  // triangle when IB3 = IB4 (src/bc_p.f90:100-110)
        for (int q = 0; q < 4; ++q) {
            if (q < nv && (B[1 + q] < 1 || B[1 + q] > c->nnode)) return fail(UCNS_EBADARG, "GRID.bnd: node id out of range");
            k[q] = q < nv ? B[1 + q] - 1 : -1;
        }
        std::sort(k.begin(), k.begin() + nv);
        c->bnd_code[k] = code;
        bf.push_back({k[0], k[1], k[2], nv == 4 ? k[3] : -1, code});
    }
    // the box and its periodic axes: a periodic face lies on a side of the
    // bounding box; its axis's period is the box's extent
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t n = 0; n < c->nnode; ++n)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], c->xyz[n * 3 + a]);
            hi[a] = std::max(hi[a], c->xyz[n * 3 + a]);
        }
    c->o.lx = hi[0] - lo[0];
    c->o.ly = hi[1] - lo[1];
    c->o.lz = hi[2] - lo[2];
    bool pax[3] = {false, false, false};
    for (const auto &F : bf) {
        if (F[4] != 5 && F[4] != 50) continue;
        const int nv = F[3] < 0 ? 3 : 4;
        int axis = -1;
        for (int a = 0; a < 3 && axis < 0; ++a) {
            const double ext = hi[a] - lo[a], tol = 1e-9 * std::max(ext, 1e-300);
            for (const double side : {lo[a], hi[a]}) {
                bool on = true;
                for (int q = 0; q < nv; ++q) on = on && std::fabs(c->xyz[(size_t)F[q] * 3 + a] - side) <= tol;
                if (on) axis = a;
            }
        }
        if (axis < 0) return fail(UCNS_EUNSUPPORTED, "GRID.bnd: a periodic face off the bounding box's sides");
        pax[axis] = true;
    }
    c->canon.resize(c->nnode);
    for (int64_t n = 0; n < c->nnode; ++n) c->canon[n] = (int32_t)n;
    for (int a = 0; a < 3; ++a) {
        if (!pax[a]) continue;
        c->per[a] = hi[a] - lo[a];
        const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        const double scale = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
        const double tol = 1e-9 * scale;
        auto key = [&](int64_t n) {
            return std::make_pair((int64_t)std::llround(c->xyz[n * 3 + b1] / tol), (int64_t)std::llround(c->xyz[n * 3 + b2] / tol));
        };
        std::map<std::pair<int64_t, int64_t>, int32_t> lower;
        for (int64_t n = 0; n < c->nnode; ++n)
            if (std::fabs(c->xyz[n * 3 + a] - lo[a]) <= tol) lower[key(n)] = (int32_t)n;
        for (int64_t n = 0; n < c->nnode; ++n) {
            if (std::fabs(c->xyz[n * 3 + a] - hi[a]) > tol) continue;
            const auto it = lower.find(key(n));
            if (it == lower.end()) return fail(UCNS_EBADARG, "periodic node without a partner on the opposite side");
            c->canon[n] = c->canon[it->second];
        }
    }
    ucnsp_case *raw = c.release();
    const int rc = build_after_mesh(raw, false);
    if (rc != UCNS_OK) return rc;
    *out = raw;
    return UCNS_OK;
}

/* RESTART.dat of an explicit march (CHECKPOINT, src/io.f90:13457-13639): the
 * step count (int32), the time (float64) and the cell averages U_C(i)%VAL(1,1:5)
 * in global cell order (float64), native byte order, no record markers */
int ucnsp_write_restart(const char *path, int32_t it, double t, const double *u, int64_t ncell) {
    if (!path || !u || ncell < 1) return fail(UCNS_EBADARG, "null argument");
    FileCloser f{std::fopen(path, "wb")};
    if (!f.f) return fail(UCNS_EBADARG, std::string("cannot create ") + path);
    if (std::fwrite(&it, 4, 1, f.f) != 1 || std::fwrite(&t, 8, 1, f.f) != 1 ||
        std::fwrite(u, 8, (size_t)ncell * 5, f.f) != (size_t)ncell * 5)
        return fail(UCNS_EBADARG, std::string("short write to ") + path);
    return UCNS_OK;
}

int ucnsp_read_restart(const char *path, int64_t ncell, int32_t *it, double *t, double *u) {
    if (!path || !it || !t || !u || ncell < 1) return fail(UCNS_EBADARG, "null argument");
    FileCloser f{std::fopen(path, "rb")};
    if (!f.f) return fail(UCNS_EBADARG, std::string("cannot open ") + path);
    std::fseek(f.f, 0, SEEK_END);
    const long size = std::ftell(f.f);
    std::fseek(f.f, 0, SEEK_SET);
    if (size != 12 + 40 * (long)ncell)
        return fail(UCNS_EBADARG, "RESTART.dat holds " + std::to_string(size) + " bytes, not 12 + 40 x " +
                                      std::to_string(ncell) + " (step, time, 5 doubles per cell)");
    if (std::fread(it, 4, 1, f.f) != 1 || std::fread(t, 8, 1, f.f) != 1 ||
        std::fread(u, 8, (size_t)ncell * 5, f.f) != (size_t)ncell * 5)
        return fail(UCNS_EBADARG, std::string("short read from ") + path);
    return UCNS_OK;
}

}  // extern "C"

extern "C" {

const char *ucnsp_last_error(void) { return g_err.c_str(); }

int ucnsp_shape_table(int32_t shape, int32_t *out) {
    if (!out) return fail(UCNS_EBADARG, "null argument");
    if (shape != UCNSP_HEX && shape != UCNSP_TET && shape != UCNSP_PRISM && shape != UCNSP_PYRAMID)
        return fail(UCNS_EBADARG, "shape must be a cell shape (hex, tet, prism, pyramid)");
    for (int k = 0; k < 64; ++k) out[k] = -1;
    out[0] = nnodes_of(shape);
    out[1] = nfaces_of(shape);
    for (int f = 0; f < nfaces_of(shape); ++f) {
        out[2 + f] = face_nv(shape, f);
        for (int k = 0; k < face_nv(shape, f); ++k) out[8 + f * 4 + k] = face_node(shape, f, k);
    }
    out[32] = nsubtets(shape);
    for (int t = 0; t < nsubtets(shape); ++t)
        for (int k = 0; k < 4; ++k) out[33 + t * 4 + k] = subtet(shape, t)[k];
    return UCNS_OK;
}

int ucnsp_face_rule(int32_t igq, int32_t tri, double *pts, double *w, int32_t *n) {
    std::vector<double> a, b;
    int m = tri ? tri_rule(igq, a, b) : quad_rule(igq, a, b);
    if (m < 0) return fail(UCNS_EUNSUPPORTED, "face rule not available");
    for (int g = 0; g < m; ++g) {
        if (tri) {
            pts[g * 3] = a[g * 3];
            pts[g * 3 + 1] = a[g * 3 + 1];
            pts[g * 3 + 2] = a[g * 3 + 2];
        } else {
            pts[g * 3] = a[g * 2];
            pts[g * 3 + 1] = a[g * 2 + 1];
            pts[g * 3 + 2] = 0.0;
        }
        w[g] = b[g];
    }
    *n = m;
    return UCNS_OK;
}

int ucnsp_build(const ucnsp_opts *opts, ucnsp_case **out) {
    if (!opts || !out) return fail(UCNS_EBADARG, "null argument");
    const ucnsp_opts &o = *opts;
    if (o.shape < UCNSP_HEX || o.shape > UCNSP_BLAST)
        return fail(UCNS_EUNSUPPORTED,
                    "shape must be hex(1), tet(2), prism(3), pyramid(4), hybrid(5), channel(6) or blast(7)");
    if (o.shape == UCNSP_BLAST && (o.nx < 8 || o.ny < 8 || o.nz < 8))
        return fail(UCNS_EBADARG, "the blast box needs at least 8 hexes per axis (core, shell, tetrahedra, prisms)");
    if (o.nx < 3 || o.ny < 3 || o.nz < 3) return fail(UCNS_EBADARG, "the box needs at least 3 cells per direction");
    for (int d = 0; d < 3; ++d) {
        const int a = o.bc[2 * d], b = o.bc[2 * d + 1];
        if ((a == 0) != (b == 0)) return fail(UCNS_EBADARG, "a periodic axis needs code 0 on both sides");
        for (int v : {a, b})
            if (v != 0 && v != 1 && v != 2 && v != 3 && v != 4 && v != 6)
                return fail(UCNS_EUNSUPPORTED, "side codes: 0 periodic, 1 inflow, 2 outflow, 3 symmetry, 4 wall, 6 farfield");
    }
    if (o.iorder < 1 || o.iorder > 4) return fail(UCNS_EUNSUPPORTED, "iorder must be 1..4");
    if (o.iweno != 1 && o.iweno != -1) return fail(UCNS_EBADARG, "iweno must be 1 or -1");
    if (o.iweno == -1 && o.iorder != 1) return fail(UCNS_EUNSUPPORTED, "MUSCL path supports iorder 1");
    if (o.iweno == 1 && o.iorder < 2) return fail(UCNS_EUNSUPPORTED, "WENO needs iorder >= 2");
    bool bumped = false;
    for (int q = 0; q < 8; ++q) bumped = bumped || o.bump[q] != 0.0;
    if (bumped && o.bc[2] == 0)
        return fail(UCNS_EBADARG, "a lower-wall bump needs a bounded y axis (the y sides cannot be periodic)");
    if (o.nthreads > 0) omp_set_num_threads(o.nthreads);
    ucnsp_case *c = new ucnsp_case();
    c->o = o;
    const bool tim = std::getenv("UCNSP_TIMING") != nullptr;
    double t0 = omp_get_wtime();
    auto lap = [&](const char *what) {
        if (!tim) return;
        double t1 = omp_get_wtime();
        std::fprintf(stderr, "[ucnsp] %-14s %8.3f s\n", what, t1 - t0);
        t0 = t1;
    };
    gen_box(*c);
    lap("mesh");
    const int rc = build_after_mesh(c, bumped);
    if (rc != UCNS_OK) return rc;
    *out = c;
    return UCNS_OK;
}

int ucnsp_fill(const ucnsp_case *c, int32_t precision, double gamma, double lwci1, ucns_mesh_v2 *m, ucns_scheme_v2 *s,
               ucns_prestore_v1 *p) {
    if (!c || !m || !s || !p) return fail(UCNS_EBADARG, "null argument");
    std::memset(m, 0, sizeof(*m));
    std::memset(s, 0, sizeof(*s));
    std::memset(p, 0, sizeof(*p));
    m->abi_version = UCNS_ABI_VERSION;
    m->ncell = (int32_t)c->ncell;
    m->index_base = 0;
    m->nface_max = ucnsp_case::NF;
    m->ifca = c->ifca.data();
    m->ineigh = c->ineigh.data();
    m->ineighn = c->ineighn.data();
    m->bcode = c->bcode.data();
    m->types_faces = c->types_faces.data();
    m->faceanglex = c->angx.data();
    m->faceangley = c->angy.data();
    m->surf = c->surf.data();
    m->totvolume = c->vol.data();
    m->minedge = c->minedge.data();
    m->centroid = c->centre.data();
    s->abi_version = UCNS_ABI_VERSION;
    s->precision = precision;
    s->iweno = c->o.iweno;
    s->iorder = c->iorder;
    s->idegfree = c->ideg;
    s->inumneighbours = c->nnb;
    s->typesten = c->typesten;
    s->ees = c->ideg2 > 0 ? 5 : 0;
    s->wenoz = 0;
    s->wenwrt = 1;
    s->limiter = 1;
    s->riemann = 1;
    s->qp_quad = c->qp_quad;
    s->qp_triangle = c->qp_tri;
    s->gamma = gamma;
    s->lwci1 = lwci1;
    s->weno_eps = 1e-11;
    s->weights_quad = c->quad_w.data();
    s->weights_tri = c->tri_w.data();
    p->admis = c->admis.data();
    p->full = c->full.data();
    p->ihexl = c->ihexl.data();
    p->invmat = c->invmat.p;
    p->indicator = c->indicator.empty() ? nullptr : c->indicator.data();
    p->integ_basis = c->integ_basis.data();
    p->volume11 = c->volume11.data();
    p->qpoints = c->qpoints.data();
    return UCNS_OK;
}

int ucnsp_info(const ucnsp_case *c, int64_t *ints, double *dbls) {
    if (!c) return fail(UCNS_EBADARG, "null case");
    if (ints) {
        int64_t v[16] = {c->ncell, c->nnode, ucnsp_case::NF, c->iorder, c->ideg, c->nnb, c->imax, c->typesten,
                         c->iselem, c->igq, c->qp_quad, c->qp_tri, c->nqpmax, c->shape, c->o.iweno, c->n_solved};
        std::memcpy(ints, v, sizeof(v));
    }
    if (dbls) {
        double v[8] = {c->per[0], c->per[1], c->per[2], 0, 0, 0, 0, 0};
        std::memcpy(dbls, v, sizeof(v));
    }
    return UCNS_OK;
}

int ucnsp_array(const ucnsp_case *c, const char *name, void **ptr, int64_t *count, int32_t *dtype) {
    if (!c || !name || !ptr || !count || !dtype) return fail(UCNS_EBADARG, "null argument");
    std::string n(name);
    auto I = [&](const std::vector<int32_t> &v) {
        *ptr = (void *)v.data();
        *count = (int64_t)v.size();
        *dtype = 0;
        return UCNS_OK;
    };
    auto D = [&](const std::vector<double> &v) {
        *ptr = (void *)v.data();
        *count = (int64_t)v.size();
        *dtype = 1;
        return UCNS_OK;
    };
    if (n == "xyz") return D(c->xyz);
    if (n == "xyz_box") return D(c->xyz_box);
    if (n == "anglex") return D(c->angx);
    if (n == "angley") return D(c->angy);
    if (n == "reorient") return I(c->reorient);
    if (n == "quad_rs") return D(c->quad_rs);
    if (n == "tri_bary") return D(c->tri_bary);
    if (n == "cell_nodes") return I(c->cnodes);
    if (n == "cell_shape") return I(c->cshape);
    if (n == "centre") return D(c->centre);
    if (n == "volume") return D(c->vol);
    if (n == "minedge") return D(c->minedge);
    if (n == "surf") return D(c->surf);
    if (n == "normal") return D(c->normal);
    if (n == "ineigh") return I(c->ineigh);
    if (n == "ineighn") return I(c->ineighn);
    if (n == "bcode") return I(c->bcode);
    if (n == "face_nodes") return I(c->fnodes);
    if (n == "admis") return I(c->admis);
    if (n == "full") return I(c->full);
    if (n == "ihexl") return I(c->ihexl);
    if (n == "indicator") return D(c->indicator);
    if (n == "integ_basis") return D(c->integ_basis);
    if (n == "volume11") return D(c->volume11);
    if (n == "qpoints") return D(c->qpoints);
    if (n == "frame") return D(c->frame);
    if (n == "quad_w") return D(c->quad_w);
    if (n == "tri_w") return D(c->tri_w);
    if (n == "integ_basis_c") return D(c->integ_basis_c);
    if (n == "indicator_c") return D(c->indicator_c);
    if (n == "invmat_c") {
        *ptr = (void *)c->invmat_c.p;
        *count = (int64_t)c->invmat_c.n;
        *dtype = 1;
        return UCNS_OK;
    }
    if (n == "invmat") {
        *ptr = (void *)c->invmat.p;
        *count = (int64_t)c->invmat.n;
        *dtype = 1;
        return UCNS_OK;
    }
    return fail(UCNS_EBADARG, "unknown array " + n);
}

int ucnsp_init_vortex(const ucnsp_case *c, double gamma, double noise_amp, uint64_t seed, double *u) {
    if (!c || !u) return fail(UCNS_EBADARG, "null argument");
    const int64_t nc = c->ncell;
#pragma omp parallel
    {
        QPts q;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            double v[8][3];
            int nn = nnodes_of(c->cshape[i]);
            for (int k = 0; k < nn; ++k) node_xyz(*c, c->cnodes[i * 8 + k], v[k]);
            cell_rule(c->cshape[i], v, 5, q);
            double rho = 1.0;
            if (noise_amp != 0.0) {
                uint64_t s = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i;
                double u1 = std::max(u01(s), 1e-300), u2 = u01(s);
                rho += noise_amp * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
            }
            double vol = 0, su = 0, sv = 0, sw = 0, sp = 0, sk = 0;
            for (size_t p = 0; p < q.w.size(); ++p) {
                double x = q.x[p * 3], y = q.x[p * 3 + 1], z = q.x[p * 3 + 2], w = q.w[p];
                double uu = std::sin(x) * std::cos(y) * std::cos(z);
                double vv = -std::cos(x) * std::sin(y) * std::cos(z);
                double ww = 0.0;
                double pp = 100.0 / gamma + (rho / 16.0) * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
                vol += w;
                su += w * uu;
                sv += w * vv;
                sw += w * ww;
                sp += w * pp;
                sk += w * 0.5 * rho * (uu * uu + vv * vv + ww * ww);
            }
            u[i * 5 + 0] = rho;
            u[i * 5 + 1] = rho * su / vol;
            u[i * 5 + 2] = rho * sv / vol;
            u[i * 5 + 3] = rho * sw / vol;
            u[i * 5 + 4] = sp / vol / (gamma - 1.0) + sk / vol;
        }
    }
    return UCNS_OK;
}

int ucnsp_fill_cweno(const ucnsp_case *c, ucns_cweno_v1 *cw) {
    if (!c || !cw) return fail(UCNS_EBADARG, "null argument");
    if (c->ideg2 == 0) return fail(UCNS_EBADARG, "the case was built without CWENO (ees 5)");
    cw->abi_version = UCNS_ABI_VERSION;
    cw->idegfree2 = c->ideg2;
    cw->inumneighbours2 = c->imax2 + 1;
    cw->reserved = 0;
    cw->invmat_c = c->invmat_c.p;
    cw->indicator_c = c->indicator_c.data();
    cw->integ_basis_c = c->integ_basis_c.data();
    return UCNS_OK;
}

int ucnsp_move_nodes(ucnsp_case *c, const double *xyz) {
    if (!c || !xyz) return fail(UCNS_EBADARG, "null argument");
    for (int64_t k = 0; k < c->nnode * 3; ++k)
        if (!std::isfinite(xyz[k])) return fail(UCNS_EBADARG, "non-finite node coordinate");
    std::copy(xyz, xyz + c->nnode * 3, c->xyz.begin());
    c->congruence_cache = false;
    build_geometry(*c);
    for (int64_t i = 0; i < c->ncell; ++i)
        if (!(c->vol[i] > 0)) return fail(UCNS_EBADARG, "moved nodes give cell " + std::to_string(i) + " no volume");
    build_prestore(*c);
    return UCNS_OK;
}

void ucnsp_free(ucnsp_case *c) { delete c; }

}  // extern "C"
