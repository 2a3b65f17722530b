// libucnsgpu.so -- MI355X (gfx950) residual + SSP-RK3 path behind include/ucns.h.
//
// Per RK stage three kernels run on the library stream:
//   k_recon   one wavefront per cell: stencil LSQ contraction (P_s dU_s),
//             smoothness indicators and WENO weights, the weighted
//             polynomial at every face Gauss point, CHECKSOL -> MUSCL/Barth
//             -> CHECKSOLX, ULEFT written once per cell
//             (ARBITRARY_ORDER, src/reconstruct_p.f90:2954-3015)
//   k_flux    one thread per unique face: rotate, HLLC, rotate back at each
//             Gauss point; the area-weighted sum is written once per face
//             (CALCULATE_FLUXESHI_CONVECTIVE, src/flux_p.f90:478-1030, which
//             solves every face twice; here each face is solved once and
//             scattered atomic-free by the gather below)
//   k_update  one thread per cell: RHS = signed sum of its faces' fluxes, the
//             Shu-Osher stage update (RUNGE_KUTTA3, src/time_p.f90:885-1027),
//             and after the last stage the next dt as a min-reduction
//             (CALCULATE_CFL, src/time_p.f90:27-116)
// Device layouts are chosen for coalescing and differ from the ABI layouts;
// ucns_upload_prestore converts.  FP64 everywhere in UCNS_FP64; in UCNS_MIXED
// the stencil matrices, indicators, Gauss points, reconstruction, weights and
// HLLC run in FP32 while dU formation, conserved->primitive recovery, flux
// accumulation, the RK update and dt stay FP64.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ucns.h"

namespace {

constexpr int NV = 5;
constexpr int WAVE = 64;
constexpr int NXCD = 8;
// distinct stencil members of a tile held in LDS by the residual product's scatter
#define UCNS_TILE_UQ 512
#define UCNS_NOSLOT 0xFFFFu
// rows of a tile's member table: 8 neighbouring cells share about 343
// distinct members with WENO-4 on tets (5 stencils), 373 with WENO-3 on
// hexes, 660 with WENO-4 on hexes (7 stencils of 39 members)
constexpr int tile_uq(int ts, int imax) { return ts > 5 && imax > 19 ? 768 : UCNS_TILE_UQ; }

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return UCNS_EHIP;                                                               \
        }                                                                                   \
    } while (0)

// monomial exponents of BASIS_REC (src/basis.f90:11-102), cubic max here
__constant__ int8_t c_mono[34][3] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
    {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2},
    {3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}, {2, 0, 1}, {1, 1, 1}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {0, 0, 3},
    {4, 0, 0}, {3, 1, 0}, {2, 2, 0}, {1, 3, 0}, {0, 4, 0}, {3, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 3, 1},
    {2, 0, 2}, {1, 1, 2}, {0, 2, 2}, {1, 0, 3}, {0, 1, 3}, {0, 0, 4}};


// per-cell info word
//   bits 0-2 ifca, 3-5 admis, 6 full, 8-13 face k is a triangle
__host__ __device__ inline int ci_ifca(int w) { return w & 7; }
__host__ __device__ inline int ci_admis(int w) { return (w >> 3) & 7; }
__host__ __device__ inline int ci_full(int w) { return (w >> 6) & 1; }
__host__ __device__ inline int ci_tri(int w, int f) { return (w >> (8 + f)) & 1; }

template <typename R>
struct ReconParams {
    int ncell, nf, nqm, ts, per_xcd;
    int iweno, qpq, qpt;
    double gamma, lwci1, eps;
    const double *U;       // [ncell][5]
    const int32_t *nbr;    // [ncell][ts][imax]
    const R *P;            // [ncell][imax][ts][ideg]
    const R *B;            // [ncell][ideg*(ideg+1)/2] packed upper triangle by rows
    const R *mbar;         // [ncell][ideg] INTEG_BASIS/VOLUME(1,1)
    const R *qp;           // [ncell][nf][nqm][3]
    const int32_t *cinfo;  // [ncell]
    const int32_t *cnb;    // [ncell][nf] face neighbours
    R *uleft;              // [ncell][nf][nqm][5]
    int *counters;         // [0] MUSCL cells, [1] first-order cells
    // CWENO (ees 5) in the per-cell kernel: P holds the central stencil only
    // ([ncell][imax][ideg]); the sectors' low-order pseudo-inverses, their
    // indicators (packed) and mbar - mbar_c over their coefficients
    const R *Pc;           // [ncell][ts-1][imax2][ideg2]
    const R *Bc;           // [ncell][ideg2*(ideg2+1)/2]
    const R *dmb;          // [ncell][ideg2]
    int wenoz;
    // characteristic WENO (wenwrt 2): per cell face the neighbour (negative:
    // the boundary code) and the cell's own face angles (cos a1, sin a1, cos
    // a2, sin a2); the free-stream data of the boundary ghosts
    const int32_t *chnb;   // [ncell][nf]
    const double *chang;   // [ncell][nf][4]
    ucns_boundary_v1 bc;
    const int32_t *cface;  // [ncell][nf] +(face+1) where the cell owns the face
    const double *bface;   // [nface][10] per-face INFLOW(), OUTFLOW() states, or null: bc's
};

__device__ inline void cons2prim_d(double gamma, const double *u, double *w) {
    double oo = 1.0 / u[0];
    w[0] = u[0];
    w[1] = u[1] * oo;
    w[2] = u[2] * oo;
    w[3] = u[3] * oo;
    w[4] = (gamma - 1.0) * (u[4] - 0.5 * u[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}

// ---------------------------------------------------------------------------
// Forward-mode dual numbers for the adjoint: the same scalar code as the
// primal, run on (value, one directional derivative).  Branches follow the
// value, so the derivative is the one of the branch the primal takes.
struct Dual {
    double v, d;
};
// the Dual overloads below would otherwise hide the global math functions
using ::fabs;
using ::fmax;
using ::fmin;
using ::sqrt;
__device__ inline double val(double x) { return x; }
__device__ inline double val(const Dual &x) { return x.v; }
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ inline Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ inline Dual sqrt(Dual a) {
    const double r = ::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
__device__ inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
__device__ inline Dual fmin(Dual a, Dual b) { return (b.v < a.v) ? b : a; }
__device__ inline Dual fmax(Dual a, Dual b) { return (b.v > a.v) ? b : a; }
__device__ inline Dual fmin(Dual a, double b) { return (b < a.v) ? Dual{b, 0.0} : a; }
__device__ inline Dual fmax(Dual a, double b) { return (b > a.v) ? Dual{b, 0.0} : a; }

template <class T>
__device__ inline void cons2prim_t(double gamma, const T *u, T *w) {
    T oo = 1.0 / u[0];
    w[0] = u[0];
    w[1] = u[1] * oo;
    w[2] = u[2] * oo;
    w[3] = u[3] * oo;
    w[4] = (gamma - 1.0) * (u[4] - 0.5 * u[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}

// CHECKSOL / CHECKSOLX test of one face point against the cell average
// (src/reconstruct_p.f90:3183-3196): primitive recovery in FP64.
__device__ inline bool jump_test(double gamma, const double *wi, const double *uq, double jump) {
    double wq[NV];
    cons2prim_d(gamma, uq, wq);
    if (fabs(wq[0] - wi[0]) >= jump * wi[0]) return true;
    if (fabs(wq[4] - wi[4]) >= jump * wi[4]) return true;
    if (wq[4] < 0.0 || wq[0] < 0.0) return true;
    return false;
}

template <typename R, int IDEG>
__device__ inline void basis_at(const R *x, const R *mbar, R *phi) {
    R px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = R(1);
#pragma unroll
    for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * x[0];
        py[k] = py[k - 1] * x[1];
        pz[k] = pz[k - 1] * x[2];
    }
#pragma unroll
    for (int d = 0; d < IDEG; ++d) phi[d] = px[c_mono[d][0]] * py[c_mono[d][1]] * pz[c_mono[d][2]] - mbar[d];
}

__device__ inline double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
    return v;
}

// face rotation and boundary ghost states (defined with the flux kernels)
template <class T>
__device__ inline void rotf(const T *a, double c1, double s1, double c2, double s2, T *r);
template <class T, class F>
__device__ inline void ghost_state(int code, int boundtype, double gamma, const T *cl, double c1, double s1, double c2,
                                   double s2, const F *fin, const F *fout, T *cr);

// Eigenvectors of the x-split Euler flux Jacobian at the average of two
// rotated states (COMPUTE_EIGENVECTORS, src/flow_operations.f90:3934-3995):
// density, velocity and total energy averaged arithmetically; R columns
// (u-a, u, v-shear, w-shear, u+a), L = R^-1 in the textbook form with
// b = (gamma-1)/a^2.  Row-major [5][5].
__device__ inline void char_eigen_d(double gamma, const double *rl, const double *rr, double *R, double *L) {
    const double rs = 0.5 * (rl[0] + rr[0]);
    const double us = 0.5 * (rl[1] / rl[0] + rr[1] / rr[0]);
    const double vs = 0.5 * (rl[2] / rl[0] + rr[2] / rr[0]);
    const double ws = 0.5 * (rl[3] / rl[0] + rr[3] / rr[0]);
    const double es = 0.5 * (rl[4] + rr[4]);
    const double q2 = us * us + vs * vs + ws * ws;
    const double ps = (gamma - 1.0) * (es - 0.5 * rs * q2);
    const double as = sqrt(gamma * ps / rs);
    const double hs = 0.5 * q2 + as * as / (gamma - 1.0);
    const double b = (gamma - 1.0) / (as * as), bq = 0.5 * b * q2, oa = 1.0 / as;
    const double Rm[25] = {1.0, 1.0, 0.0, 0.0, 1.0,
                           us - as, us, 0.0, 0.0, us + as,
                           vs, vs, 1.0, 0.0, vs,
                           ws, ws, 0.0, 1.0, ws,
                           hs - us * as, 0.5 * q2, vs, ws, hs + us * as};
    const double Lm[25] = {0.5 * (bq + us * oa), -0.5 * (b * us + oa), -0.5 * b * vs, -0.5 * b * ws, 0.5 * b,
                           1.0 - bq, b * us, b * vs, b * ws, -b,
                           -vs, 0.0, 1.0, 0.0, 0.0,
                           -ws, 0.0, 0.0, 1.0, 0.0,
                           0.5 * (bq - us * oa), -0.5 * (b * us - oa), -0.5 * b * vs, -0.5 * b * ws, 0.5 * b};
    for (int e = 0; e < 25; ++e) {
        R[e] = Rm[e];
        L[e] = Lm[e];
    }
}

// CWS: the CWENO layout, the sectors' rows compact (IM2 members, ID2
// coefficients each) and the indicator packed (upper triangle by rows)
template <typename R, int IDEG, int IMAX, int TSMAX, bool CWS = false, int ID2 = 9, int IM2 = 19>
struct ReconShared {
    static constexpr int NQMAX = 6 * 16;
    static constexpr int DU = CWS ? IMAX + (TSMAX - 1) * IM2 : TSMAX * IMAX;
    static constexpr int GN = CWS ? IDEG + (TSMAX - 1) * ID2 : TSMAX * IDEG;
    static constexpr int BN = CWS ? IDEG * (IDEG + 1) / 2 : IDEG * IDEG;
    union {
        R du[DU * NV];  // phase 1
        struct {
            R gh[IDEG * NV];         // combined polynomial
            R beta[TSMAX * NV];
            R omega[TSMAX * NV];
            R ul[NQMAX * NV];        // face-point values
        } late;
    } u;
    R g[GN * NV];                    // per-stencil coefficients
    R b[BN];                         // indicator, unpacked (CWS: packed)
    R mb[IDEG];
    double ui[NV];
    // CWENO: the sectors' indicator (unpacked), the indicators in FP64 and
    // the face-value shift of the sectors' low-order means
    R bc[9 * 9];
    double betad[TSMAX * NV];
    R corr[NV];
};

// Fused reconstruction, one wavefront per cell.  CW: CWENO with sectors of
// ID2 coefficients over their first IM2 members (src/reconstruct_p.f90:955-1049).
// CH: characteristic-variable WENO, face by face (src/reconstruct_p.f90:497-758).
template <typename R, int IDEG, int IMAX, int TSMAX, bool CW = false, int ID2 = 9, int IM2 = 19, bool CH = false>
__global__ __launch_bounds__(WAVE, CW ? 2 : 1) void k_recon(ReconParams<R> p) {
    static_assert(!CW || (ID2 <= 9 && IM2 <= IMAX), "CWENO sectors: at most 9 coefficients");
    constexpr int NROW = (TSMAX * IDEG + WAVE - 1) / WAVE;
    __shared__ ReconShared<R, IDEG, IMAX, TSMAX, CW, ID2, IM2> sh;
    // CWENO: row of sector s >= 1's coefficient k in sh.g
    auto crow = [](int s, int k) { return IDEG + (s - 1) * ID2 + k; };
    const int lane = threadIdx.x;
    // XCD-aware order: consecutive blocks go to different XCDs, so give
    // each XCD a contiguous range of cells (its L2 then holds the
    // neighbourhood of its cells)
    const int b = blockIdx.x;
    const int cell = (b % NXCD) * p.per_xcd + b / NXCD;
    if (cell >= p.ncell) return;

    const int info = p.cinfo[cell];
    const int ifca = ci_ifca(info), admis = ci_admis(info), full = ci_full(info);
    const int ts = p.ts;
    const double *Ui = p.U + (size_t)cell * NV;
    double ui[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) ui[v] = Ui[v];
    if (lane < NV) sh.ui[lane] = ui[lane];

    // ---- phase 0: dU of every stencil member (FP64, then rounded to R) ----
    const int32_t *nb = p.nbr + (size_t)cell * ts * IMAX;
    if constexpr (CW) {
        // the central stencil's IMAX members, then each sector's first IM2
        for (int e = lane; e < IMAX + (admis - 1) * IM2; e += WAVE) {
            const int m = e < IMAX ? e : ((e - IMAX) / IM2 + 1) * IMAX + (e - IMAX) % IM2;
            const double *Un = p.U + (size_t)nb[m] * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) sh.u.du[e * NV + v] = (R)(Un[v] - ui[v]);
        }
        const R *bcp = p.Bc + (size_t)cell * (ID2 * (ID2 + 1) / 2);
        for (int e = lane; e < ID2 * (ID2 + 1) / 2; e += WAVE) {
            int k = 0, base = 0;
            while (e >= base + (ID2 - k)) {
                base += ID2 - k;
                ++k;
            }
            const int l = k + (e - base);
            sh.bc[k * ID2 + l] = bcp[e];
            sh.bc[l * ID2 + k] = bcp[e];
        }
    } else {
        for (int e = lane; e < admis * IMAX; e += WAVE) {
            const double *Un = p.U + (size_t)nb[e] * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) sh.u.du[e * NV + v] = (R)(Un[v] - ui[v]);
        }
    }
    if (CW) {
        const R *Bc = p.B + (size_t)cell * (IDEG * (IDEG + 1) / 2);
        for (int e = lane; e < IDEG * (IDEG + 1) / 2; e += WAVE) sh.b[e] = Bc[e];
    } else if (p.iweno == 1) {
        const R *Bc = p.B + (size_t)cell * (IDEG * (IDEG + 1) / 2);
        for (int e = lane; e < IDEG * (IDEG + 1) / 2; e += WAVE) {
            // packed upper triangle by rows: row k holds l = k..IDEG-1
            int k = 0, base = 0;
            while (e >= base + (IDEG - k)) {
                base += IDEG - k;
                ++k;
            }
            int l = k + (e - base);
            R val = Bc[e];
            sh.b[k * IDEG + l] = val;
            sh.b[l * IDEG + k] = val;
        }
    }
    for (int e = lane; e < IDEG; e += WAVE) sh.mb[e] = p.mbar[(size_t)cell * IDEG + e];
    __syncthreads();

    // ---- phase 1: G_s = P_s dU_s, one (stencil,coefficient) row per lane ----
    if constexpr (CW) {
        // rows 0..IDEG-1: the central stencil; then ID2 rows per sector
        constexpr int NROWC = (IDEG + (TSMAX - 1) * ID2 + WAVE - 1) / WAVE;
        const int nrows = IDEG + (admis - 1) * ID2;
        const R *Pcen = p.P + (size_t)cell * IMAX * IDEG;
        const R *Psec = p.Pc + (size_t)cell * (ts - 1) * IM2 * ID2;
#pragma unroll
        for (int t = 0; t < NROWC; ++t) {
            const int r = lane + t * WAVE;
            if (r >= nrows) continue;
            R acc[NV] = {R(0), R(0), R(0), R(0), R(0)};
            if (r < IDEG) {
                for (int j = 0; j < IMAX; ++j) {
                    const R pv = Pcen[(size_t)j * IDEG + r];
                    const R *du = &sh.u.du[j * NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] += pv * du[v];
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) sh.g[r * NV + v] = acc[v];
            } else {
                const int sec = (r - IDEG) / ID2, k = (r - IDEG) % ID2;
                const R *Ps = Psec + (size_t)sec * IM2 * ID2;
                for (int j = 0; j < IM2; ++j) {
                    const R pv = Ps[(size_t)j * ID2 + k];
                    const R *du = &sh.u.du[(IMAX + sec * IM2 + j) * NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] += pv * du[v];
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) sh.g[crow(sec + 1, k) * NV + v] = acc[v];
            }
        }
    } else {
        R acc[NROW][NV];
        int srow[NROW];
#pragma unroll
        for (int t = 0; t < NROW; ++t) {
            srow[t] = (lane + t * WAVE) / IDEG;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[t][v] = R(0);
        }
        const int nrows = admis * IDEG;
        const R *Pc = p.P + (size_t)cell * IMAX * ts * IDEG;
#pragma unroll 4
        for (int j = 0; j < IMAX; ++j) {
            const R *Pj = Pc + (size_t)j * ts * IDEG;
#pragma unroll
            for (int t = 0; t < NROW; ++t) {
                const int r = lane + t * WAVE;
                if (r < nrows) {
                    const R pv = Pj[r];
                    const R *du = &sh.u.du[(srow[t] * IMAX + j) * NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[t][v] += pv * du[v];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NROW; ++t) {
            const int r = lane + t * WAVE;
            if (r < nrows) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sh.g[r * NV + v] = acc[t][v];
            }
        }
    }
    __syncthreads();

    bool reduce = (full == 0);
    const int nslot = ifca * p.nqm;
    if constexpr (CW) {
        // ---- CWENO: the zero polynomial (G_1 - sum_l lambda_l G_l) / lambda_1,
        // lambda_1 = 1 - 1/LWCI1, lambda_l = (1 - lambda_1)/(ADMIS-1) ----
        const double lam0 = 1.0 - (1.0 / p.lwci1);
        const double laml = admis > 1 ? (1.0 - lam0) / (admis - 1) : 0.0;
        for (int e = lane; e < IDEG * NV; e += WAVE) {
            const int k = e / NV, v = e % NV;
            R sl = R(0);
            if (k < ID2)
                for (int s = 1; s < admis; ++s) sl += (R)laml * sh.g[crow(s, k) * NV + v];
            sh.u.late.gh[e] = (R)(1.0 / lam0) * (sh.g[e] - sl);
        }
        __syncthreads();
        // indicators: the zero polynomial's with INDICATOR, a sector's with INDICATORC
        if (lane < admis * NV) {
            const int s = lane / NV, v = lane % NV;
            double beta = 0.0;
            if (s == 0) {
                // over the packed triangle: sum_k z_k (B_kk z_k + 2 sum_{l>k} B_kl z_l)
                int e = 0;
                for (int k = 0; k < IDEG; ++k) {
                    const double zk = (double)sh.u.late.gh[k * NV + v];
                    double bg = 0.0;
                    for (int l = k + 1; l < IDEG; ++l) bg += (double)sh.b[e + l - k] * (double)sh.u.late.gh[l * NV + v];
                    beta += zk * ((double)sh.b[e] * zk + 2.0 * bg);
                    e += IDEG - k;
                }
            } else {
                const R *gs = &sh.g[crow(s, 0) * NV + v];
#pragma unroll
                for (int k = 0; k < ID2; ++k) {
                    double bg = 0.0;
#pragma unroll
                    for (int l = 0; l < ID2; ++l) bg += (double)sh.bc[k * ID2 + l] * (double)gs[l * NV];
                    beta += (double)gs[k * NV] * bg;
                }
            }
            sh.betad[lane] = beta;
        }
        __syncthreads();
        // weights in FP64: WENO (power 4) or WENO-Z, tau = sum_l |beta_1 - beta_l| / (ADMIS-1)
        if (lane < admis * NV) {
            const int s = lane / NV, v = lane % NV;
            double tau = 0.0;
            if (p.wenoz && admis > 1) {
                for (int t = 1; t < admis; ++t) tau += fabs(sh.betad[v] - sh.betad[t * NV + v]);
                tau /= (double)(admis - 1);
            }
            double sum = 0.0, mine = 0.0;
            for (int t = 0; t < admis; ++t) {
                const double lam = t == 0 ? lam0 : laml;
                const double x = p.eps + sh.betad[t * NV + v];
                double w;
                if (p.wenoz) {
                    const double r = tau / x, r2 = r * r;
                    w = lam * (1.0 + r2 * r2);
                } else {
                    const double x2 = x * x;
                    w = lam / (x2 * x2);
                }
                sum += w;
                if (t == s) mine = w;
            }
            sh.u.late.omega[lane] = (R)(mine / sum);
        }
        __syncthreads();
        // weighted polynomial on the high-order basis; the sectors' polynomials
        // use the means of VALUEC: their weighted coefficients times
        // (mbar - mbar_c) shift every face value alike
        for (int e = lane; e < IDEG * NV; e += WAVE) {
            const int k = e / NV, v = e % NV;
            R acc = sh.u.late.omega[v] * sh.u.late.gh[e];
            if (k < ID2)
                for (int s = 1; s < admis; ++s) acc += sh.u.late.omega[s * NV + v] * sh.g[crow(s, k) * NV + v];
            sh.u.late.gh[e] = acc;
        }
        if (lane < NV) {
            const R *dm = p.dmb + (size_t)cell * ID2;
            R c = R(0);
#pragma unroll
            for (int k = 0; k < ID2; ++k) {
                R wg = R(0);
                for (int s = 1; s < admis; ++s) wg += sh.u.late.omega[s * NV + lane] * sh.g[crow(s, k) * NV + lane];
                c += dm[k] * wg;
            }
            sh.corr[lane] = c;
        }
        __syncthreads();
    }
    if constexpr (CH) {
        // ---- characteristic WENO: per face, the eigenvectors of the averaged
        // cell / neighbour state in the face frame (WENO_NEIGHBOUR: the ghost of
        // the cell's own average on boundary faces), every stencil's
        // coefficients projected with L, indicators and weights per
        // characteristic variable, the weighted coefficients projected back
        // with R, and that face's values (compute_gradcharv_smoothindicator,
        // diag_At_B_A, src/reconstruct_p.f90:1608-1818) ----
        __shared__ double chR[25], chL[25];
        __shared__ double cbeta[TSMAX * NV];
        __shared__ R wch[TSMAX * IDEG * NV];
        double wi[NV];
        cons2prim_d(p.gamma, ui, wi);
        int trouble = 0;
        const R *qpc = p.qp + (size_t)cell * p.nf * p.nqm * 3;
        for (int f = 0; f < ifca; ++f) {
            if (lane == 0) {
                const int nbf = p.chnb[(size_t)cell * p.nf + f];
                const double *an = p.chang + ((size_t)cell * p.nf + f) * 4;
                double ur[NV], rl[NV], rr[NV];
                if (nbf < 0) {
                    const double *bf = p.bface ? p.bface + (size_t)(p.cface[(size_t)cell * p.nf + f] - 1) * 10 : nullptr;
                    ghost_state(-nbf, p.bc.boundtype, p.gamma, ui, an[0], an[1], an[2], an[3],
                                bf ? bf : p.bc.inflow, bf ? bf + 5 : p.bc.outflow, ur);
                }
                else
#pragma unroll
                    for (int v = 0; v < NV; ++v) ur[v] = p.U[(size_t)nbf * NV + v];
                rotf(ui, an[0], an[1], an[2], an[3], rl);
                rotf(ur, an[0], an[1], an[2], an[3], rr);
                char_eigen_d(p.gamma, rl, rr, chR, chL);
            }
            __syncthreads();
            for (int e = lane; e < admis * IDEG * NV; e += WAVE) {
                const int c = e % NV, sk = e / NV;
                double acc = 0.0;
#pragma unroll
                for (int v = 0; v < NV; ++v) acc += (double)sh.g[sk * NV + v] * chL[c * 5 + v];
                wch[e] = (R)acc;
            }
            __syncthreads();
            if (lane < admis * NV) {
                const int s = lane / NV, c = lane % NV;
                const R *ws = &wch[s * IDEG * NV + c];
                double beta = 0.0;
                for (int a = 0; a < IDEG; ++a) {
                    double ba = 0.0;
                    for (int b2 = 0; b2 < IDEG; ++b2) ba += (double)sh.b[a * IDEG + b2] * (double)ws[b2 * NV];
                    beta += (double)ws[a * NV] * ba;
                }
                cbeta[lane] = beta;
            }
            __syncthreads();
            if (lane < admis * NV) {
                // lambda/(eps+beta)^4 per characteristic variable, in FP64 in both precisions
                const int s = lane / NV, c = lane % NV;
                double sum = 0.0, mine = 0.0;
                for (int t = 0; t < admis; ++t) {
                    const double x = p.eps + cbeta[t * NV + c];
                    const double x2 = x * x;
                    const double w = (t == 0 ? p.lwci1 : 1.0) / (x2 * x2);
                    sum += w;
                    if (t == s) mine = w;
                }
                sh.u.late.omega[lane] = (R)(mine / sum);
            }
            __syncthreads();
            for (int e = lane; e < IDEG * NV; e += WAVE) {
                const int k = e / NV, v = e % NV;
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    double lim = 0.0;
                    for (int s = 0; s < admis; ++s)
                        lim += (double)wch[(s * IDEG + k) * NV + c] * (double)sh.u.late.omega[s * NV + c];
                    acc += chR[v * 5 + c] * lim;
                }
                sh.u.late.gh[e] = (R)acc;
            }
            __syncthreads();
            const int nq = ci_tri(info, f) ? p.qpt : p.qpq;
#pragma unroll 1
            for (int q = lane; q < nq; q += WAVE) {
                asm volatile("" ::: "memory");
                const int e = f * p.nqm + q;
                R x[3] = {qpc[e * 3], qpc[e * 3 + 1], qpc[e * 3 + 2]};
                R phi[IDEG];
                basis_at<R, IDEG>(x, sh.mb, phi);
                double uq[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    R res = R(0);
                    for (int k = 0; k < IDEG; ++k) res += phi[k] * sh.u.late.gh[k * NV + v];
                    uq[v] = ui[v] + (double)res;
                    sh.u.late.ul[e * NV + v] = res;
                }
                if (full && jump_test(p.gamma, wi, uq, 0.85)) trouble = 1;
            }
            __syncthreads();
        }
        reduce = reduce || (__syncthreads_or(trouble) != 0);
    } else if (p.iweno == 1) {
        if constexpr (!CW) {
        // ---- phase 2: beta_{s,v} = g^T (B g) (src/reconstruct_p.f90:995-1004) ----
        if (lane < admis * NV) {
            const int s = lane / NV, v = lane % NV;
            const R *gs = &sh.g[s * IDEG * NV + v];
            R beta = R(0);
            for (int k = 0; k < IDEG; ++k) {
                R bg = R(0);
#pragma unroll
                for (int l = 0; l < IDEG; ++l) bg += sh.b[k * IDEG + l] * gs[l * NV];
                beta += gs[k * NV] * bg;
            }
            sh.u.late.beta[lane] = beta;
        }
        __syncthreads();
        // ---- phase 3: weights (src/reconstruct_p.f90:1031-1049) ----
        if (lane < admis * NV) {
            const int s = lane / NV, v = lane % NV;
            const R eps = (R)p.eps;
            R om;
            if (sizeof(R) == 8) {
                // lambda/(eps+beta)^4 as in the reference
                R sum = R(0), mine = R(0);
                for (int t = 0; t < admis; ++t) {
                    R x = eps + sh.u.late.beta[t * NV + v];
                    R x2 = x * x;
                    R w = (t == 0 ? (R)p.lwci1 : R(1)) / (x2 * x2);
                    sum += w;
                    if (t == s) mine = w;
                }
                om = mine / sum;
            } else {
                // FP32: (eps+beta)^-4 underflows, so scale every weight by
                // (eps+beta_min)^4 first; the normalised weights are unchanged
                R xmin = eps + sh.u.late.beta[v];
                for (int t = 1; t < admis; ++t) xmin = fminf(xmin, eps + sh.u.late.beta[t * NV + v]);
                R sum = R(0), mine = R(0);
                for (int t = 0; t < admis; ++t) {
                    R r = xmin / (eps + sh.u.late.beta[t * NV + v]);
                    R r2 = r * r;
                    R w = (t == 0 ? (R)p.lwci1 : R(1)) * (r2 * r2);
                    sum += w;
                    if (t == s) mine = w;
                }
                om = mine / sum;
            }
            sh.u.late.omega[lane] = om;
        }
        __syncthreads();
        // ---- phase 4: combined coefficients sum_s omega_s G_s ----
        for (int e = lane; e < IDEG * NV; e += WAVE) {
            const int v = e % NV;
            R acc = R(0);
            for (int s = 0; s < admis; ++s) acc += sh.u.late.omega[s * NV + v] * sh.g[s * IDEG * NV + e];
            sh.u.late.gh[e] = acc;
        }
        __syncthreads();
        }
        // ---- phase 5: ULEFT at the face points + CHECKSOL (0.85) ----
        double wi[NV];
        cons2prim_d(p.gamma, ui, wi);
        int trouble = 0;
        const R *qpc = p.qp + (size_t)cell * p.nf * p.nqm * 3;
#pragma unroll 1
        for (int e = lane; e < nslot; e += WAVE) {
            // keep the coefficient loads inside the point loop (hoisted, they
            // would hold IDEG x 5 values in registers across it)
            asm volatile("" ::: "memory");
            const int f = e / p.nqm, q = e % p.nqm;
            const int nq = ci_tri(info, f) ? p.qpt : p.qpq;
            if (q >= nq) continue;
            R x[3] = {qpc[e * 3], qpc[e * 3 + 1], qpc[e * 3 + 2]};
            R phi[IDEG];
            basis_at<R, IDEG>(x, sh.mb, phi);
            double uq[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R res = R(0);
#pragma unroll
                for (int k = 0; k < IDEG; ++k) res += phi[k] * sh.u.late.gh[k * NV + v];
                if constexpr (CW) res += sh.corr[v];
                // ULEFT is kept as the deviation from the FP64 cell average
                uq[v] = ui[v] + (double)res;
                sh.u.late.ul[e * NV + v] = res;
            }
            if (full && jump_test(p.gamma, wi, uq, 0.85)) trouble = 1;
        }
        reduce = reduce || (__syncthreads_or(trouble) != 0);
    } else {
        reduce = true;  // MUSCL scheme: every cell is non-FULL
        __syncthreads();
    }

    // ---- phase 6: MUSCL + Barth-Jespersen (src/reconstruct_p.f90:2044-2280) ----
    if (reduce) {
        double umin[NV], umax[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) umin[v] = umax[v] = ui[v];
        for (int f = 0; f < ifca; ++f) {
            const int n = p.cnb[(size_t)cell * p.nf + f];
            if (n < 0) continue;
            const double *Un = p.U + (size_t)n * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                umin[v] = fmin(umin[v], Un[v]);
                umax[v] = fmax(umax[v], Un[v]);
            }
        }
        const R *qpc = p.qp + (size_t)cell * p.nf * p.nqm * 3;
        double slope[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) slope[v] = 1.0e16;
#pragma unroll 1
        for (int e = lane; e < nslot; e += WAVE) {
            // keep the coefficient loads inside the point loop (hoisted, they
            // would hold IDEG x 5 values in registers across it)
            asm volatile("" ::: "memory");
            const int f = e / p.nqm, q = e % p.nqm;
            const int nq = ci_tri(info, f) ? p.qpt : p.qpq;
            if (q >= nq) continue;
            R x[3] = {qpc[e * 3], qpc[e * 3 + 1], qpc[e * 3 + 2]};
            R phi[IDEG];
            basis_at<R, IDEG>(x, sh.mb, phi);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R res = R(0);
#pragma unroll
                for (int k = 0; k < IDEG; ++k) res += phi[k] * sh.g[k * NV + v];
                const double d2 = (double)res;
                double psi;
                if (fabs(d2) <= 0.0)
                    psi = 1.0;
                else if (d2 > 0.0)
                    psi = fmin(1.0, (umax[v] - ui[v]) / d2);
                else
                    psi = fmin(1.0, (umin[v] - ui[v]) / d2);
                slope[v] = fmin(slope[v], psi);
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) slope[v] = wave_min(slope[v]);
#pragma unroll 1
        for (int e = lane; e < nslot; e += WAVE) {
            // keep the coefficient loads inside the point loop (hoisted, they
            // would hold IDEG x 5 values in registers across it)
            asm volatile("" ::: "memory");
            const int f = e / p.nqm, q = e % p.nqm;
            const int nq = ci_tri(info, f) ? p.qpt : p.qpq;
            if (q >= nq) continue;
            R x[3] = {qpc[e * 3], qpc[e * 3 + 1], qpc[e * 3 + 2]};
            R phi[IDEG];
            basis_at<R, IDEG>(x, sh.mb, phi);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R res = R(0);
#pragma unroll
                for (int k = 0; k < IDEG; ++k) res += phi[k] * sh.g[k * NV + v];
                sh.u.late.ul[e * NV + v] = (R)((double)res * slope[v]);
            }
        }
        if (lane == 0) atomicAdd(&p.counters[0], 1);
        __syncthreads();
    }

    // ---- phase 7: CHECKSOLX (0.95) -> first order (src/reconstruct_p.f90:3245-3406) ----
    {
        double wi[NV];
        cons2prim_d(p.gamma, ui, wi);
        int trouble = 0;
#pragma unroll 1
        for (int e = lane; e < nslot; e += WAVE) {
            // keep the coefficient loads inside the point loop (hoisted, they
            // would hold IDEG x 5 values in registers across it)
            asm volatile("" ::: "memory");
            const int f = e / p.nqm, q = e % p.nqm;
            const int nq = ci_tri(info, f) ? p.qpt : p.qpq;
            if (q >= nq) continue;
            double uq[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) uq[v] = ui[v] + (double)sh.u.late.ul[e * NV + v];
            if (jump_test(p.gamma, wi, uq, 0.95)) trouble = 1;
        }
        const bool first = __syncthreads_or(trouble) != 0;
        if (first) {
            for (int e = lane; e < nslot; e += WAVE)
#pragma unroll
                for (int v = 0; v < NV; ++v) sh.u.late.ul[e * NV + v] = R(0);
            if (lane == 0) atomicAdd(&p.counters[1], 1);
            __syncthreads();
        }
    }

    // ---- phase 8: ULEFT out, one contiguous block per cell ----
    R *out = p.uleft + (size_t)cell * p.nf * p.nqm * NV;
    for (int e = lane; e < nslot * NV; e += WAVE) out[e] = sh.u.late.ul[e];
}

// ---------------------------------------------------------------------------
// WENO reconstruction, lane per (cell, stencil).  A wavefront owns a tile of
// 8 consecutive cells; lane = 8*local_cell + stencil.  Each lane streams its
// stencil's pseudo-inverse from the tiled layout [tile][j][k][8*ts] (one
// contiguous block per load instruction), forms dU_j on the fly from the
// FP64 states and keeps its ideg x 5 coefficients in registers.  The
// indicators, weights and the weighted sum over stencils use shuffles
// inside the 8-lane group of a cell; face points are split over the group.
template <typename R>
struct ReconWParams {
    int ncell, ntile, tiles_per_xcd, ts, nf, nqm, qpq, qpt;
    double gamma, lwci1, eps;
    const double *U;
    const double *U64;     // [ncell][8] the state in 64-byte records (row-stream gathers)
    const int32_t *nbr;    // [ntile][imax][8*ts]
    const R *P;            // [ntile][imax][ideg][8*ts]
    const R *B;            // [ncell][ideg*(ideg+1)/2]
    const R *mbar;         // [ncell][ideg]
    const R *qp;           // [ncell][nf][nqm][3]
    const int32_t *cinfo;
    const int32_t *cnb;    // [ncell][nf]
    R *uleft;              // [ncell][nf][nqm][5] deviations from the cell average
    int *counters;
    const unsigned char *stream;  // per-tile row stream of the DMA kernels (StreamShape)
    long long tile_bytes;
    R *gstore;             // optional [ntile][ideg*5 + 10][8*ts]: every stencil's coefficients, then
                           // its indicators beta[5] and weights omega[5], for the adjoint
    int *fstore;           // with gstore: per cell, 1 = MUSCL (CHECKSOL), 2 = first order (CHECKSOLX)
    // scatter of a residual product riding on the row stream (k_recon_sv*):
    // member j of lane position pos receives P_j^T gbs[.][.][pos]
    const double *gbs;        // [ntile][ideg*5][8*ts] cotangents of the stencil coefficients
    double *ubar;             // [ncell][5] accumulated
    const uint16_t *slot;     // [ntile][imax][8*ts] row of the tile's member table (UCNS_NOSLOT: HBM)
    int uq_cap;               // rows of a member table at most (tile_uq(ts, imax))
    const int32_t *uq, *ucount;
    const int64_t *toff;      // first row of each tile's member table in tab
    double *tab;              // [rows][5] the member tables, summed per cell by k_adj_gather
    // CWENO (ees 5): the sectors' low-order indicators (packed, [ncell][6]) and
    // the differences of the basis means mbar - mbar_c over their 3
    // coefficients ([ncell][3]); the sectors' P rows hold their low-order
    // pseudo-inverses (zero past 3 coefficients and imax2 members)
    const R *Bc;
    const R *dmb;
    int wenoz;
};

template <typename T>
__device__ inline T lane_bcast(T v, int src) {
    return __shfl(v, src, WAVE);
}

template <typename T>
__device__ inline T grp_sum8(T v) {
    v += __shfl_xor(v, 1, WAVE);
    v += __shfl_xor(v, 2, WAVE);
    v += __shfl_xor(v, 4, WAVE);
    return v;
}
template <typename T>
__device__ inline T grp_min8(T v) {
    v = fmin(v, __shfl_xor(v, 1, WAVE));
    v = fmin(v, __shfl_xor(v, 2, WAVE));
    v = fmin(v, __shfl_xor(v, 4, WAVE));
    return v;
}
__device__ inline int grp_or8(int v) {
    v |= __shfl_xor(v, 1, WAVE);
    v |= __shfl_xor(v, 2, WAVE);
    v |= __shfl_xor(v, 4, WAVE);
    return v;
}

// index of B(k,l) in the packed upper triangle stored by rows
__host__ __device__ constexpr int pidx(int ideg, int k, int l) {
    return k <= l ? k * ideg - k * (k - 1) / 2 + (l - k) : l * ideg - l * (l - 1) / 2 + (k - l);
}

// Direct global->LDS loads (global_load_lds_dword / _dwordx4): lane l's 4 or
// 16 bytes from its own global address land at LDS byte (M0 + 4 l) or
// (M0 + 16 l).  Issued as inline asm so that the compiler neither counts nor
// waits for them; every reader waits with an explicit s_waitcnt.  The nop
// covers the M0-write hazard ahead of the load.
__device__ __forceinline__ unsigned lds_addr(const void *p) {
    return (unsigned)(size_t)(const __attribute__((address_space(3))) char *)p;
}
__device__ __forceinline__ void lds_load16(const void *g, unsigned lds_base) {
    asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "{m0}"(lds_base) : "memory");
}

// one 1 KiB piece: lane l moves bytes [16 l, 16 l + 16) of the chunk when
// they lie below `valid`
__device__ __forceinline__ void dma_chunk(const char *src, char *lds_dst, int lane, int valid) {
    if (lane * 16 < valid) lds_load16(src + lane * 16, lds_addr(lds_dst));
}

// STORE: also write every stencil's coefficients to p.gstore (the adjoint's forward pass).
// qpc / mb: the cell's face points and basis means, in HBM or staged in LDS
// by loads still in flight (QWAIT: wait for them before the face points).
// STG: the face values pass through stg (LDS, 8 x 8 x NV values, free once
// the indicators are formed) and leave as coalesced words.
template <typename R, int IDEG, int NQL, bool STORE = false, bool CW = false, bool QWAIT = false, bool STG = false>
__device__ __forceinline__ void recon_late(const ReconWParams<R> &p, R (&g)[IDEG][NV], const double (&ui)[NV],
                                           const R (*sB)[IDEG * (IDEG + 1) / 2], R (*sG0)[IDEG * NV], int lc, int s,
                                           int cell, bool cvalid, int info, bool act, const R *qpc, const R *mb,
                                           R *stg = nullptr) {
    const int ifca = ci_ifca(info), full = ci_full(info);
    if (STORE && act) {
        // coalesced: for each coefficient the 8 ts stencils of the tile are contiguous
        const int TP = 8 * p.ts;
        R *gs = p.gstore + (size_t)(cell >> 3) * (IDEG + 2) * NV * TP + (lc * p.ts + s);
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) gs[(size_t)(k * NV + v) * TP] = g[k][v];
    }
    // central-stencil coefficients kept for the MUSCL fallback
    if (s == 0) {
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) sG0[lc][k * NV + v] = g[k][v];
    }
    __syncthreads();

    R om[NV];
    R corr[NV] = {R(0), R(0), R(0), R(0), R(0)};  // CWENO: the sectors' low-order means
    if constexpr (CW) {
        // ---- CWENO (src/reconstruct_p.f90:955-1029): lambda_1 = 1 - 1/LWCI1 for
        // the central stencil, (1 - lambda_1)/(ADMIS-1) for each sector; lane 0
        // turns its central polynomial into the zero polynomial
        // (G_1 - sum_l lambda_l G_l) / lambda_1 (the sectors' coefficients are
        // zero past the first three) ----
        const int admis = ci_admis(info);
        const double lam0 = 1.0 - (1.0 / p.lwci1);
        const double laml = admis > 1 ? (1.0 - lam0) / (admis - 1) : 0.0;
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R sl = grp_sum8((s >= 1 && act) ? (R)laml * g[k][v] : R(0));
                if (s == 0) g[k][v] = (R)(1.0 / lam0) * (g[k][v] - sl);
            }
        // indicators: the zero polynomial's with INDICATOR, a sector's with INDICATORC
        double bd[NV] = {0, 0, 0, 0, 0};
        if (act && s == 0) {
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                R bg[NV] = {R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
                for (int l = k + 1; l < IDEG; ++l) {
                    const R bkl = sB[lc][pidx(IDEG, k, l)];
#pragma unroll
                    for (int v = 0; v < NV; ++v) bg[v] += bkl * g[l][v];
                }
                const R bkk = sB[lc][pidx(IDEG, k, k)];
#pragma unroll
                for (int v = 0; v < NV; ++v) bd[v] += (double)(g[k][v] * (bkk * g[k][v] + R(2) * bg[v]));
            }
        } else if (act) {
            const R *bc = p.Bc + (size_t)cell * 6;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R bg[NV] = {R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
                for (int l = k + 1; l < 3; ++l) {
                    const R bkl = bc[pidx(3, k, l)];
#pragma unroll
                    for (int v = 0; v < NV; ++v) bg[v] += bkl * g[l][v];
                }
                const R bkk = bc[pidx(3, k, k)];
#pragma unroll
                for (int v = 0; v < NV; ++v) bd[v] += (double)(g[k][v] * (bkk * g[k][v] + R(2) * bg[v]));
            }
        }
        // weights in FP64 in both precisions: WENO (power 4) or WENO-Z with
        // tau = sum_l |beta_1 - beta_l| / (ADMIS-1)
        const double lam = s == 0 ? lam0 : laml;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double b0 = lane_bcast(bd[v], lc * 8);
            double w = 0.0;
            if (p.wenoz) {
                const double tau = admis > 1 ? grp_sum8((s >= 1 && act) ? fabs(b0 - bd[v]) : 0.0) / (admis - 1) : 0.0;
                const double r = tau / (p.eps + bd[v]);
                const double r2 = r * r;
                w = lam * (1.0 + r2 * r2);
            } else {
                const double x = p.eps + bd[v];
                const double x2 = x * x;
                w = lam / (x2 * x2);
            }
            w = act ? w : 0.0;
            const double sum = grp_sum8(w);
            om[v] = act ? (R)(w / sum) : R(0);
        }
        // the sectors' polynomials use the means of VALUEC: their weighted
        // coefficients times (mbar - mbar_c) shift every face value alike
        const R *dm = p.dmb + (size_t)(cvalid ? cell : 0) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) corr[v] += dm[k] * grp_sum8((s >= 1 && act) ? om[v] * g[k][v] : R(0));
    } else {
    // ---- smoothness indicators beta = g^T (B g) (src/reconstruct_p.f90:995-1004),
    // over the packed triangle of the symmetric B:
    //   beta = sum_k g_k (B_kk g_k + 2 sum_{l>k} B_kl g_l) ----
    R beta[NV] = {R(0), R(0), R(0), R(0), R(0)};
    if (act) {
#pragma unroll
        for (int k = 0; k < IDEG; ++k) {
            R bg[NV] = {R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
            for (int l = k + 1; l < IDEG; ++l) {
                const R bkl = sB[lc][pidx(IDEG, k, l)];
#pragma unroll
                for (int v = 0; v < NV; ++v) bg[v] += bkl * g[l][v];
            }
            const R bkk = sB[lc][pidx(IDEG, k, k)];
#pragma unroll
            for (int v = 0; v < NV; ++v) beta[v] += g[k][v] * (bkk * g[k][v] + R(2) * bg[v]);
        }
    }
    // ---- weights (src/reconstruct_p.f90:1031-1049) and the weighted sum ----
    const R lam = (s == 0) ? (R)p.lwci1 : R(1);
    const R eps = (R)p.eps;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        R w;
        if (sizeof(R) == 8) {
            const R x = eps + beta[v];
            const R x2 = x * x;
            w = act ? lam / (x2 * x2) : R(0);
        } else {
            // FP32: scale by (eps+beta_min)^4 so (eps+beta)^-4 cannot underflow
            const R x = eps + beta[v];
            const R xmin = grp_min8(act ? x : (R)INFINITY);
            const R r = xmin / x;
            const R r2 = r * r;
            w = act ? lam * (r2 * r2) : R(0);
        }
        const R sum = grp_sum8(w);
        om[v] = act ? w / sum : R(0);
    }
    if (STORE && act) {
        const int TP = 8 * p.ts;
        R *gs = p.gstore + ((size_t)(cell >> 3) * (IDEG + 2) + IDEG) * NV * TP + (lc * p.ts + s);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gs[(size_t)v * TP] = beta[v];
            gs[(size_t)(NV + v) * TP] = om[v];
        }
    }
    }
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[k][v] = grp_sum8(om[v] * g[k][v]);

    // ---- face points: deviation from the average + CHECKSOL (0.85) ----
    // Values are stored as they are formed.  A cell that passes CHECKSOL
    // cannot fail CHECKSOLX (its 0.95 jumps imply the 0.85 ones), so only
    // the MUSCL branch below re-tests and overwrites its face points.
    const int nslot = ifca * p.nqm;
    if constexpr (QWAIT) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    R *out = p.uleft + (size_t)(cvalid ? cell : 0) * p.nf * p.nqm * NV;
    double wi[NV];
    cons2prim_d(p.gamma, ui, wi);
    int trouble = 0;
#pragma unroll 1
    for (int q = s; q < NQL * 8; q += 8) {
        const int f = q / p.nqm, qq = q - f * p.nqm;
        const bool ok = cvalid && q < nslot && qq < (ci_tri(info, f) ? p.qpt : p.qpq);
        if (ok) {
            R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
            R phi[IDEG];
            basis_at<R, IDEG>(x, mb, phi);
            double uq[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R res = corr[v];
#pragma unroll
                for (int k = 0; k < IDEG; ++k) res += phi[k] * g[k][v];
                if constexpr (STG)
                    stg[(lc * 8 + s) * NV + v] = res;
                else
                    out[q * NV + v] = res;
                uq[v] = ui[v] + (double)res;
            }
            if (full && jump_test(p.gamma, wi, uq, 0.85)) trouble = 1;
        } else if constexpr (STG) {
#pragma unroll
            for (int v = 0; v < NV; ++v) stg[(lc * 8 + s) * NV + v] = R(0);
        } else if (cvalid && q < nslot) {
#pragma unroll
            for (int v = 0; v < NV; ++v) out[q * NV + v] = R(0);
        }
        if constexpr (STG) {
            // The tile's points q0..q0+7 go out from LDS as whole 8-byte
            // words, consecutive lanes on consecutive words: each cell's
            // 8 points are one contiguous run of its block.  Words past a
            // cell's block (its last points when nf*nqm < NQL*8) stay unwritten.
            asm volatile("" ::: "memory");
            constexpr int WPC = 8 * NV * (int)sizeof(R) / 8;  // words per cell and pass
            const int q0 = q - s;
            const int cbytes = p.nf * p.nqm * NV * (int)sizeof(R);
            const int seg = q0 * NV * (int)sizeof(R);
            const int cell0 = cell - lc;
#pragma unroll
            for (int e = lc * 8 + s; e < 8 * WPC; e += WAVE) {
                const int c = e / WPC, j = e - c * WPC;
                if (cell0 + c < p.ncell && seg + (j + 1) * 8 <= cbytes) {
                    const uint64_t w = ((const uint64_t *)(stg + c * 8 * NV))[j];
                    *(uint64_t *)((char *)p.uleft + (size_t)(cell0 + c) * cbytes + seg + j * 8) = w;
                }
            }
            asm volatile("" ::: "memory");
        }
    }
    const int any_trouble = grp_or8(trouble);  // every lane takes part in the shuffle
    const bool reduce = cvalid && (!full || any_trouble);
    bool first = false;

    // ---- MUSCL + Barth-Jespersen on the central polynomial (src/reconstruct_p.f90:2044-2280) ----
    // then CHECKSOLX (0.95) -> first order (src/reconstruct_p.f90:3245-3406)
    if (reduce) {
        // the staged words of these points are written before they are overwritten
        if constexpr (STG) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        double umin[NV], umax[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) umin[v] = umax[v] = ui[v];
        for (int f = 0; f < ifca; ++f) {
            const int n = p.cnb[(size_t)cell * p.nf + f];
            if (n < 0) continue;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double un = p.U[(size_t)n * NV + v];
                umin[v] = fmin(umin[v], un);
                umax[v] = fmax(umax[v], un);
            }
        }
        double slope[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) slope[v] = 1.0e16;
        R res0[NQL][NV];
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            const int f = q / p.nqm, qq = q - f * p.nqm;
            const bool ok = q < nslot && qq < (ci_tri(info, f) ? p.qpt : p.qpq);
#pragma unroll
            for (int v = 0; v < NV; ++v) res0[r][v] = R(0);
            if (ok) {
                R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
                R phi[IDEG];
                basis_at<R, IDEG>(x, mb, phi);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    R res = R(0);
#pragma unroll
                    for (int k = 0; k < IDEG; ++k) res += phi[k] * sG0[lc][k * NV + v];
                    res0[r][v] = res;
                    const double d2 = (double)res;
                    double psi;
                    if (fabs(d2) <= 0.0)
                        psi = 1.0;
                    else if (d2 > 0.0)
                        psi = fmin(1.0, (umax[v] - ui[v]) / d2);
                    else
                        psi = fmin(1.0, (umin[v] - ui[v]) / d2);
                    slope[v] = fmin(slope[v], psi);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) slope[v] = grp_min8(slope[v]);
        int tr2 = 0;
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            const int f = q / p.nqm, qq = q - f * p.nqm;
            const bool ok = q < nslot && qq < (ci_tri(info, f) ? p.qpt : p.qpq);
#pragma unroll
            for (int v = 0; v < NV; ++v) res0[r][v] = (R)((double)res0[r][v] * slope[v]);
            if (ok) {
                double uq[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) uq[v] = ui[v] + (double)res0[r][v];
                if (jump_test(p.gamma, wi, uq, 0.95)) tr2 = 1;
            }
        }
        first = grp_or8(tr2) != 0;
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            if (q < nslot) {
#pragma unroll
                for (int v = 0; v < NV; ++v) out[q * NV + v] = first ? R(0) : res0[r][v];
            }
        }
        if (s == 0) {
            atomicAdd(&p.counters[0], 1);
            if (first) atomicAdd(&p.counters[1], 1);
        }
    }
    if (STORE && s == 0 && cvalid) p.fstore[cell] = (reduce ? 1 : 0) | (first ? 2 : 0);
}

template <typename R, int IDEG, int IMAX, int NQL, int D, bool CW = false>
__device__ __forceinline__ void recon_w_body(const ReconWParams<R> &p) {
    constexpr int NP = IDEG * (IDEG + 1) / 2;
    __shared__ R sB[8][NP];
    __shared__ R sG0[8][IDEG * NV];
    __shared__ int sN[IMAX * 64];  // stencil members of the tile, [j][pos]
    const int lane = threadIdx.x, lc = lane >> 3, s = lane & 7;
    const int b = blockIdx.x;
    const int tile = (b % NXCD) * p.tiles_per_xcd + b / NXCD;
    if (tile >= p.ntile) return;
    const int cell = tile * 8 + lc;
    const bool cvalid = cell < p.ncell;
    const int info = cvalid ? p.cinfo[cell] : 0;
    const int ifca = ci_ifca(info), admis = ci_admis(info), full = ci_full(info);
    const bool act = cvalid && s < admis;
    const int TP = 8 * p.ts, pos = lc * p.ts + s;
    double ui[NV] = {1.0, 0.0, 0.0, 0.0, 1.0};
    if (cvalid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) ui[v] = p.U[(size_t)cell * NV + v];
    }
    const int nvc = min(8, p.ncell - tile * 8);
    R g[IDEG][NV];
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[k][v] = R(0);
    {
        // indicators and stencil members of the tile -> LDS (contiguous blocks)
        {
            const R *Bt = p.B + (size_t)tile * 8 * NP;
            for (int e = lane; e < nvc * NP; e += WAVE) (&sB[0][0])[e] = Bt[e];
            const int32_t *nt = p.nbr + (size_t)tile * IMAX * TP;
            for (int e = lane; e < IMAX * TP; e += WAVE) sN[e] = nt[e];
        }
        __syncthreads();
        // ---- LSQ: g_s = P_s dU_s (src/gradients.f90:253-497) ----
        // D members are in flight (their states and pseudo-inverse columns in
        // registers) while one is accumulated.
        if (act) {
            const R *Pt = p.P + (size_t)tile * IMAX * IDEG * TP + pos;
            R pb[D][IDEG];
            double ub[D][NV];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int n = sN[d * TP + pos];
#pragma unroll
                for (int v = 0; v < NV; ++v) ub[d][v] = p.U[(size_t)n * NV + v];
                const R *Pj = Pt + (size_t)d * IDEG * TP;
#pragma unroll
                for (int k = 0; k < IDEG; ++k) pb[d][k] = Pj[k * TP];
            }
#pragma unroll 1
            for (int j0 = 0; j0 < IMAX; j0 += D) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const int j = j0 + d;
                    if (j < IMAX) {
                        R du[NV], pc[IDEG];
#pragma unroll
                        for (int v = 0; v < NV; ++v) du[v] = (R)(ub[d][v] - ui[v]);
#pragma unroll
                        for (int k = 0; k < IDEG; ++k) pc[k] = pb[d][k];
                        const int jn = j + D;
                        if (jn < IMAX) {
                            const int n = sN[jn * TP + pos];
#pragma unroll
                            for (int v = 0; v < NV; ++v) ub[d][v] = p.U[(size_t)n * NV + v];
                            const R *Pj = Pt + (size_t)jn * IDEG * TP;
#pragma unroll
                            for (int k = 0; k < IDEG; ++k) pb[d][k] = Pj[k * TP];
                        }
#pragma unroll
                        for (int k = 0; k < IDEG; ++k)
#pragma unroll
                            for (int v = 0; v < NV; ++v) g[k][v] += pc[k] * du[v];
                    }
                }
            }
        }
    }
    const size_t cq = cvalid ? cell : 0;
    recon_late<R, IDEG, NQL, false, CW>(p, g, ui, sB, sG0, lc, s, cell, cvalid, info, act,
                                        p.qp + cq * p.nf * p.nqm * 3, p.mbar + cq * IDEG);
}

// ---------------------------------------------------------------------------
// Row-stream reconstruction: every byte the LSQ loop needs arrives by direct
// global->LDS loads (global_load_lds), so no register load is outstanding
// while rows are in flight.  Vector-memory counters complete in order, and
// the compiler does not count these loads, so the loop waits with explicit
// s_waitcnt vmcnt(N) where N is a lower bound on the loads issued after the
// row it needs.
//
// HBM stream of one tile: header = member indices of rows 0..DEPTH-1, then
// per member row j: the pseudo-inverse block [k][8 ts] (R) followed by the
// member indices of row j + DEPTH ([8 ts] int32), so the indices of a row
// are in LDS one ring turn before its states are fetched.
//
// LDS slot of a row: the row block as streamed, then the member states of
// its TP columns as 48-byte records (5 doubles and a pad), gathered in
// 16-byte pieces from a copy of the state padded to 64-byte records, so a
// record is one cache line; consecutive lanes take consecutive pieces and a
// row is NSI gathers.  After the loop the slots hold the tile's indicators and the
// central coefficients.
constexpr int stream_depth(bool mixed, int ts) { return ts == 5 ? 4 : (mixed ? 2 : 3); }

template <typename R, int IDEG, int IMAX, int TS, int DEPTH>
struct StreamShape {
    static constexpr int NP = IDEG * (IDEG + 1) / 2;
    static constexpr int TP = 8 * TS;
    static constexpr int PB = IDEG * TP * (int)sizeof(R);  // pseudo-inverse bytes of a row
    static constexpr int RD = PB + TP * 4;                 // row bytes in HBM
    static constexpr int HD = DEPTH * TP * 4;              // tile header bytes
    static constexpr int NCH = (RD + 1023) / 1024;         // 1 KiB loads per row
    // the row's member states: the TP columns' 48-byte records, three
    // 16-byte pieces each, gathered by NSI loads of 64 pieces; they start
    // right after the row block (its last 1 KiB load writes only to RD)
    static constexpr int NSI = (3 * TP + 63) / 64;
    static constexpr int UA = (RD + 15) / 16 * 16;
    static constexpr int SB = UA + NSI * 1024;             // slot bytes
    static constexpr int PER = NCH + NSI;                  // loads per row
    static constexpr int BBYTES = 8 * NP * (int)sizeof(R);
    static constexpr int NBC = (BBYTES + 1023) / 1024;
    static constexpr int NB = (BBYTES + SB - 1) / SB;
    static constexpr int S0 = (IMAX - DEPTH) % DEPTH;      // slot of the first tail row
    static constexpr bool TAILB = S0 + NB <= DEPTH && NB < DEPTH;
    static constexpr int TB = NB - 1;                      // tail step after which they are free
    static constexpr int BSLOT = TAILB ? S0 : 0;
    static constexpr int GSLOT = TAILB ? (S0 + NB) % DEPTH : NB;
    static constexpr long long TILE = HD + (long long)IMAX * RD;
    static_assert(RD % 16 == 0 && HD % 16 == 0, "16-byte pieces");
    static_assert(HD <= 1024, "header is one load");
    static_assert(GSLOT < DEPTH && NBC * 1024 <= NB * SB, "late data must fit the ring");
    static_assert(8 * IDEG * NV * (int)sizeof(R) <= SB, "central coefficients exceed a slot");
    static_assert((DEPTH - 1) * PER + NBC <= 63, "vmcnt immediate out of range");
    // after the stream the ring holds B ([B0, B1): every chunk issue_b
    // writes) and the central coefficients ([G0, G1)); the rest is free
    static constexpr int B0 = BSLOT * SB, B1 = B0 + NBC * 1024;
    static constexpr int G0 = GSLOT * SB, G1 = G0 + 8 * IDEG * NV * (int)sizeof(R);
};

// first 1 KiB-aligned offset in [0, ring) where `need` bytes avoid the three
// given ranges; -1 if there is none
constexpr int free_lds(int ring, int need, int a0, int a1, int b0, int b1, int c0, int c1) {
    for (int o = 0; o + need <= ring; o += 1024) {
        const int e = o + need;
        if ((e > a0 && o < a1) || (e > b0 && o < b1) || (e > c0 && o < c1)) continue;
        return o;
    }
    return -1;
}

template <typename R, int IDEG, int IMAX, int NQL, int TS, int DEPTH, bool STORE = false>
__device__ __forceinline__ void recon_s_body(const ReconWParams<R> &p) {
    using SS = StreamShape<R, IDEG, IMAX, TS, DEPTH>;
    constexpr int NP = SS::NP, TP = SS::TP;
    // ring slots, then the header's index rows
    __shared__ __attribute__((aligned(1024))) char ring[DEPTH * SS::SB + SS::HD];
    const int *hdr = (const int *)(ring + DEPTH * SS::SB);
    const int lane = threadIdx.x, lc = lane >> 3, s = lane & 7;
    const int b = blockIdx.x;
    const int tile = (b % NXCD) * p.tiles_per_xcd + b / NXCD;
    if (tile >= p.ntile) return;
    const int cell = tile * 8 + lc;
    const bool cvalid = cell < p.ncell;
    int info = cvalid ? p.cinfo[cell] : 0;
    double ui[NV] = {1.0, 0.0, 0.0, 0.0, 1.0};
    if (cvalid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) ui[v] = p.U[(size_t)cell * NV + v];
    }
    // the cell's own data is in registers before the first row load
    asm volatile("s_waitcnt vmcnt(0)"
                 : "+v"(info), "+v"(ui[0]), "+v"(ui[1]), "+v"(ui[2]), "+v"(ui[3]), "+v"(ui[4])
                 :
                 : "memory");
    const int admis = ci_admis(info);
    const bool act = cvalid && s < admis;
    const int rl = act ? lc * TS + s : 0;
    const int nvc = min(8, p.ncell - tile * 8);
    const char *T = (const char *)p.stream + (size_t)tile * SS::TILE;

    auto issue_row = [&](int j) {
        const char *src = T + SS::HD + (size_t)j * SS::RD;
        char *dst = ring + (j % DEPTH) * SS::SB;
#pragma unroll
        for (int c = 0; c < SS::NCH; ++c) dma_chunk(src + c * 1024, dst + c * 1024, lane, SS::RD - c * 1024);
    };
    // piece q = lane + 64 c of a row is bytes [16 (q % 3), +16) of column
    // q / 3's record; pieces past the TP columns re-read cell 0.  Every lane
    // issues every gather, so each row adds exactly NSI loads to the count.
    auto piece_cells = [&](const int *cols, int (&ix)[SS::NSI]) {
#pragma unroll
        for (int c = 0; c < SS::NSI; ++c) {
            const int q = lane + 64 * c;
            ix[c] = q < 3 * TP ? cols[q / 3] : 0;
        }
    };
    auto issue_states = [&](int j, const int (&ix)[SS::NSI]) {
        char *dst = ring + (j % DEPTH) * SS::SB + SS::UA;
#pragma unroll
        for (int c = 0; c < SS::NSI; ++c) {
            const int q = lane + 64 * c;
            const int part = q < 3 * TP ? q % 3 : 0;
            lds_load16(p.U64 + (size_t)ix[c] * 8 + 2 * part, lds_addr(dst + c * 1024));
        }
    };
    // every chunk is issued, also past the last cell's block (lane 0 then
    // re-reads the block's first bytes), so the loads counted below exist
    auto issue_b = [&]() {
        const char *src = (const char *)(p.B + (size_t)tile * 8 * NP);
        char *dst = ring + SS::BSLOT * SS::SB;
        const int bb = nvc * NP * (int)sizeof(R);
#pragma unroll
        for (int c = 0; c < SS::NBC; ++c) {
            const int valid = bb - c * 1024;
            dma_chunk(valid > 0 ? src + c * 1024 : src, dst + c * 1024, lane, valid > 0 ? valid : 16);
        }
    };


    // prologue: header, then (row, states) pairs for rows 0..DEPTH-1
    dma_chunk(T, ring + DEPTH * SS::SB, lane, SS::HD);
    issue_row(0);
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(SS::NCH) : "memory");  // header landed
#pragma unroll
    for (int r = 0; r < DEPTH; ++r) {
        if (r > 0) issue_row(r);
        int ix[SS::NSI];
        piece_cells(hdr + r * TP, ix);
        issue_states(r, ix);
    }

    // ---- LSQ: g_s = P_s dU_s (src/gradients.f90:253-497) ----
    R g[IDEG][NV];
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[k][v] = R(0);
    auto consume = [&](int i) {
        if (act) {
            const char *slot = ring + (i % DEPTH) * SS::SB;
            const double *st = (const double *)(slot + SS::UA + rl * 48);
            const double2 a = *(const double2 *)st;
            const double2 c = *(const double2 *)(st + 2);
            const double u4 = st[4];
            const R du[NV] = {(R)(a.x - ui[0]), (R)(a.y - ui[1]), (R)(c.x - ui[2]), (R)(c.y - ui[3]),
                              (R)(u4 - ui[4])};
            const R *row = (const R *)slot + rl;
            R pc[IDEG];
#pragma unroll
            for (int k = 0; k < IDEG; ++k) pc[k] = row[k * TP];
#pragma unroll
            for (int k = 0; k < IDEG; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) g[k][v] += pc[k] * du[v];
        }
    };
#pragma unroll 1
    for (int i = 0; i < IMAX - DEPTH; ++i) {
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"((DEPTH - 1) * SS::PER) : "memory");
        consume(i);
        // members of row i+DEPTH, read before that row overwrites this slot
        const int *cols = (const int *)(ring + (i % DEPTH) * SS::SB + SS::PB);
        int ix[SS::NSI];
        piece_cells(cols, ix);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        issue_row(i + DEPTH);
        issue_states(i + DEPTH, ix);
    }
    // tail: the last DEPTH rows; the indicators follow into the first free slots
    [&]<int... TT>(std::integer_sequence<int, TT...>) {
        (
            [&] {
                constexpr int t = TT;
                constexpr int pend = (DEPTH - 1 - t) * SS::PER + ((SS::TAILB && t > SS::TB) ? SS::NBC : 0);
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"(pend) : "memory");
                consume(IMAX - DEPTH + t);
                if constexpr (SS::TAILB && t == SS::TB) {
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                    issue_b();
                }
            }(),
            ...);
    }(std::make_integer_sequence<int, DEPTH>{});
    if constexpr (!SS::TAILB) {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        issue_b();
    }
    // The tile's face points and basis means follow into free LDS, so that
    // the face-point loop reads them there instead of waiting on HBM once
    // per point; the indicators wait for B only.  Where they do not fit
    // (hexahedra in FP32) the loop reads them from HBM.
    constexpr int RING = DEPTH * SS::SB;
    constexpr int NQC = (8 * NQL * 8 * 3 * (int)sizeof(R) + 1023) / 1024;
    constexpr int NMC = (8 * IDEG * (int)sizeof(R) + 1023) / 1024;
    constexpr int QO = free_lds(RING, NQC * 1024, SS::B0, SS::B1, SS::G0, SS::G1, 0, 0);
    constexpr int MO = QO < 0 ? -1 : free_lds(RING, NMC * 1024, SS::B0, SS::B1, SS::G0, SS::G1, QO, QO + NQC * 1024);
    const size_t cq = cvalid ? cell : 0;
    const R *qpc = p.qp + cq * p.nf * p.nqm * 3;
    const R *mb = p.mbar + cq * IDEG;
    if constexpr (QO >= 0) {
        static_assert(NQC + NMC <= 63, "vmcnt immediate out of range");
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // the tail's reads of these slots are done
        auto issue_blk = [&](const R *src, int bytes, int off, int nch) {
#pragma unroll 1
            for (int c = 0; c < nch; ++c) {
                const int valid = bytes - c * 1024;
                dma_chunk(valid > 0 ? (const char *)src + c * 1024 : (const char *)src, ring + off + c * 1024, lane,
                          valid > 0 ? valid : 16);
            }
        };
        const int per = p.nf * p.nqm * 3;
        issue_blk(p.qp + (size_t)tile * 8 * per, nvc * per * (int)sizeof(R), QO, NQC);
        qpc = (const R *)(ring + QO) + lc * per;
        if constexpr (MO >= 0) {
            issue_blk(p.mbar + (size_t)tile * 8 * IDEG, nvc * IDEG * (int)sizeof(R), MO, NMC);
            mb = (const R *)(ring + MO) + lc * IDEG;
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(NQC + NMC) : "memory");  // indicators landed
        } else {
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(NQC) : "memory");
        }
    } else {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    const R (*sB)[NP] = (const R (*)[NP])(ring + SS::BSLOT * SS::SB);
    R (*sG0)[IDEG * NV] = (R (*)[IDEG * NV])(ring + SS::GSLOT * SS::SB);
    // FP64 face values leave through LDS as coalesced words (128^3 hexahedra
    // 20.8 -> 20.1 ms, 64^3 x 6 tetrahedra 10.86 -> 10.71); in FP32 the
    // direct stores measured as fast or faster (hexahedra 10.8 ms direct,
    // 11.6 staged)
    constexpr bool STG = sizeof(R) == 8;
    static_assert(!STG || 8 * 8 * NV * (int)sizeof(R) <= SS::B1 - SS::B0, "face-value staging exceeds the B block");
    recon_late<R, IDEG, NQL, STORE, false, (QO >= 0), STG>(p, g, ui, sB, sG0, lc, s, cell, cvalid, info, act, qpc, mb,
                                                           (R *)(ring + SS::B0));
}

template <int IDEG, int IMAX, int NQL, int TS>
__global__ __launch_bounds__(WAVE) void k_recon_s64(ReconWParams<double> p) {
    recon_s_body<double, IDEG, IMAX, NQL, TS, stream_depth(false, TS)>(p);
}
// Scatter of a residual product on the row stream, without the
// reconstruction: lane (cell, stencil) holds its g_bar (p.gbs) and streams
// the tile's pseudo-inverse rows through the same LDS ring and layout as the
// reconstruction (no neighbour states); dU_bar_j = P_j^T g_bar goes to the
// tile's member table (then p.tab, summed per cell by k_adj_gather) and
// minus the same to the cell.
template <typename R, int IDEG, int IMAX, int TS, int DEPTH>
__device__ __forceinline__ void scatter_s_body(const ReconWParams<R> &p) {
    using SS = StreamShape<R, IDEG, IMAX, TS, DEPTH>;
    constexpr int TP = SS::TP, NCH = SS::NCH, SL = NCH * 1024;
    constexpr int UQ = tile_uq(TS, IMAX);  // LDS table rows
    __shared__ __attribute__((aligned(1024))) char ring[DEPTH * SL + SS::HD];
    __shared__ double sAcc[UQ][NV];
    __shared__ uint16_t sSlot[IMAX * TP];
    const int *hdr = (const int *)(ring + DEPTH * SL);
    const int lane = threadIdx.x, lc = lane >> 3, s = lane & 7;
    const int b = blockIdx.x;
    const int tile = (b % NXCD) * p.tiles_per_xcd + b / NXCD;
    if (tile >= p.ntile) return;
    const int cell = tile * 8 + lc;
    const bool cvalid = cell < p.ncell;
    int info = cvalid ? p.cinfo[cell] : 0;
    double gb[IDEG][NV];
    {
        const double *gs = p.gbs + (size_t)tile * IDEG * NV * TP + (lc * TS + s);
        const bool in = cvalid && s < TS;
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) gb[k][v] = in ? gs[(size_t)(k * NV + v) * TP] : 0.0;
    }
    const uint16_t *st = p.slot + (size_t)tile * IMAX * TP;
    for (int e = lane; e < IMAX * TP; e += WAVE) sSlot[e] = st[e];
    const int nuq = min(p.ucount[tile], p.uq_cap);
    for (int e = lane; e < min(nuq, UQ) * NV; e += WAVE) (&sAcc[0][0])[e] = 0.0;
    // every cotangent is used before the row stream starts, so the compiler
    // waits for their loads here and not inside the stream
    double gsum = 0.0;
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) gsum += gb[k][v];
    asm volatile("s_waitcnt vmcnt(0)" : "+v"(info), "+v"(gsum) : : "memory");
    __syncthreads();  // sSlot, sAcc
    const bool act = cvalid && s < ci_admis(info);
    const int rl = act ? lc * TS + s : 0;
    const char *T = (const char *)p.stream + (size_t)tile * SS::TILE;
    auto issue_row = [&](int j) {
        const char *src = T + SS::HD + (size_t)j * SS::RD;
        char *dst = ring + (j % DEPTH) * SL;
#pragma unroll
        for (int c = 0; c < NCH; ++c) dma_chunk(src + c * 1024, dst + c * 1024, lane, SS::RD - c * 1024);
    };
    int qn[DEPTH];  // members of the next DEPTH rows, oldest first
    double acc[NV] = {0, 0, 0, 0, 0};
    dma_chunk(T, ring + DEPTH * SL, lane, SS::HD);
    issue_row(0);
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(NCH) : "memory");  // header landed
#pragma unroll
    for (int r = 0; r < DEPTH; ++r) qn[r] = hdr[r * TP + rl];
#pragma unroll
    for (int r = 1; r < DEPTH; ++r) issue_row(r);
    auto consume = [&](int i) {
        if (act) {
            const R *row = (const R *)(ring + (i % DEPTH) * SL) + rl;
            double d[NV] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                const double pk = (double)row[k * TP];
#pragma unroll
                for (int v = 0; v < NV; ++v) d[v] += pk * gb[k][v];
            }
            const unsigned sl = sSlot[i * TP + rl];
            if (sl < (unsigned)UQ) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    __hip_atomic_fetch_add(&sAcc[sl][v], d[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) unsafeAtomicAdd(&p.ubar[(size_t)qn[0] * NV + v], d[v]);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] -= d[v];
        }
#pragma unroll
        for (int r = 0; r + 1 < DEPTH; ++r) qn[r] = qn[r + 1];
    };
#pragma unroll 1
    for (int i = 0; i < IMAX - DEPTH; ++i) {
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"((DEPTH - 1) * NCH) : "memory");
        consume(i);
        const int nn = *(const int *)(ring + (i % DEPTH) * SL + SS::PB + rl * 4);  // members of row i+DEPTH
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        issue_row(i + DEPTH);
        qn[DEPTH - 1] = nn;
    }
    [&]<int... TT>(std::integer_sequence<int, TT...>) {
        (
            [&] {
                constexpr int t = TT;
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"((DEPTH - 1 - t) * NCH) : "memory");
                consume(IMAX - DEPTH + t);
            }(),
            ...);
    }(std::make_integer_sequence<int, DEPTH>{});
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double t = grp_sum8(acc[v]);
        if (s == 0 && cvalid) unsafeAtomicAdd(&p.ubar[(size_t)cell * NV + v], t);
    }
    __syncthreads();
    double *tt = p.tab + (size_t)p.toff[tile] * NV;
    for (int e = lane; e < nuq * NV; e += WAVE) tt[e] = e < UQ * NV ? (&sAcc[0][0])[e] : 0.0;
}
template <int IDEG, int IMAX, int TS>
__global__ __launch_bounds__(WAVE) void k_scatter_s64(ReconWParams<double> p) {
    scatter_s_body<double, IDEG, IMAX, TS, stream_depth(false, TS)>(p);
}
template <int IDEG, int IMAX, int TS>
__global__ __launch_bounds__(WAVE) void k_scatter_s32(ReconWParams<float> p) {
    scatter_s_body<float, IDEG, IMAX, TS, stream_depth(true, TS)>(p);
}

// the reconstruction storing every stencil's coefficients for a residual product
template <int IDEG, int IMAX, int NQL, int TS>
__global__ __launch_bounds__(WAVE) void k_recon_sg64(ReconWParams<double> p) {
    recon_s_body<double, IDEG, IMAX, NQL, TS, stream_depth(false, TS), true>(p);
}
template <int IDEG, int IMAX, int NQL, int TS>
__global__ __launch_bounds__(WAVE, 2) void k_recon_sg32(ReconWParams<float> p) {
    recon_s_body<float, IDEG, IMAX, NQL, TS, stream_depth(true, TS), true>(p);
}
template <int IDEG, int IMAX, int NQL, int TS>
__global__ __launch_bounds__(WAVE, 2) void k_recon_s32(ReconWParams<float> p) {
    recon_s_body<float, IDEG, IMAX, NQL, TS, stream_depth(true, TS)>(p);
}

// upload: pseudo-inverses of cells cell0 .. cell0+cn-1 ([cell][tsa][imax][ideg]
// FP64, the ABI order) into the row stream
template <typename R>
__global__ void k_stream_p(const double *src, unsigned char *stream, int cell0, int cn, int tsa, int ts, int imax,
                           int ideg, int hd, int rd, long long tile_bytes) {
    const size_t n = (size_t)cn * ts * imax * ideg;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % ideg);
        const int j = (int)((e / ideg) % imax);
        const int st = (int)((e / ((size_t)ideg * imax)) % ts);
        const int cl = (int)(e / ((size_t)ideg * imax * ts));
        const int i = cell0 + cl;
        const double v = src[(((size_t)cl * tsa + st) * imax + j) * ideg + k];
        R *row = (R *)(stream + (size_t)(i / 8) * tile_bytes + hd + (size_t)j * rd);
        row[(size_t)k * 8 * ts + (i % 8) * ts + st] = (R)v;
    }
}

// upload: member-index rows [ntile][depth+imax][8 ts] (row r = members j=r,
// zero past imax) into the header (r < depth) and behind the pseudo-inverse
// block of row r - depth
__global__ void k_stream_idx(const int32_t *idx, unsigned char *stream, int ntile, int depth, int imax, int tp, int hd,
                             int pb, int rd, long long tile_bytes) {
    const size_t n = (size_t)ntile * (depth + imax) * tp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(e % tp);
        const int r = (int)((e / tp) % (depth + imax));
        const size_t t = e / ((size_t)tp * (depth + imax));
        unsigned char *base = stream + t * tile_bytes;
        int32_t *dst = r < depth ? (int32_t *)(base + (size_t)r * tp * 4) + q
                                 : (int32_t *)(base + hd + (size_t)(r - depth) * rd + pb) + q;
        *dst = idx[e];
    }
}

// FP64: registers for the 19x5 coefficients dominate, one wave per SIMD;
// FP32: the compiler is asked for WAVES32 waves per SIMD.
#ifndef UCNS_WAVES32
#define UCNS_WAVES32 1
#endif
#ifndef UCNS_DEPTH64
#define UCNS_DEPTH64 1
#endif
#ifndef UCNS_DEPTH32
#define UCNS_DEPTH32 2
#endif
template <int IDEG, int IMAX, int NQL>
__global__ __launch_bounds__(WAVE) void k_recon_w64(ReconWParams<double> p) {
    recon_w_body<double, IDEG, IMAX, NQL, UCNS_DEPTH64>(p);
}
template <int IDEG, int IMAX, int NQL>
__global__ __launch_bounds__(WAVE) void k_recon_cw64(ReconWParams<double> p) {
    recon_w_body<double, IDEG, IMAX, NQL, UCNS_DEPTH64, true>(p);
}
template <int IDEG, int IMAX, int NQL>
__global__ __launch_bounds__(WAVE, UCNS_WAVES32) void k_recon_cw32(ReconWParams<float> p) {
    recon_w_body<float, IDEG, IMAX, NQL, UCNS_DEPTH32, true>(p);
}
template <int IDEG, int IMAX, int NQL>
__global__ __launch_bounds__(WAVE, UCNS_WAVES32) void k_recon_w32(ReconWParams<float> p) {
    recon_w_body<float, IDEG, IMAX, NQL, UCNS_DEPTH32>(p);
}
// [cells][ts_in][imax][ideg] (FP64, ABI) -> tiled [tile][imax][ideg][8*ts] (R)
template <typename R>
__global__ void k_tile_p(const double *src, R *dst, int cell0, int ncells_chunk, int ncell, int ts_in, int ts,
                         int imax, int ideg) {
    // dst covers the tiles of cells [cell0, cell0+ncells_chunk), cell0 % 8 == 0
    const int TP = 8 * ts;
    const size_t per_tile = (size_t)imax * ideg * TP;
    const size_t ntile = ((size_t)ncells_chunk + 7) / 8;
    const size_t total = per_tile * ntile;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t t = e / per_tile;
        size_t r = e % per_tile;
        const int pos = (int)(r % TP);
        r /= TP;
        const int k = (int)(r % ideg);
        const int j = (int)(r / ideg);
        const int lc = pos / ts, s = pos % ts;
        const size_t cl = t * 8 + lc;  // cell within the chunk
        R val = R(0);
        if (cl < (size_t)ncells_chunk && cell0 + cl < (size_t)ncell && s < ts_in)
            val = (R)src[((cl * ts_in + s) * imax + j) * ideg + k];
        dst[e] = val;
    }
}

// ---------------------------------------------------------------------------
// HLLC (src/riemann.f90:221-426) on rotated FP64 states.  Primitive recovery
// is FP64.  R=double keeps the reference's arithmetic.  R=float evaluates the
// wave speeds and the star-state correction in FP32 and each side's
// physical flux F_K in FP64:
//   F = F_K + S_K (U*_K - U_K),  K = L if S_M >= 0 else R,
// with U*_K - U_K in its factored form, every component carrying
// (S_M - u_K) / (S_K - S_M):
//   mass rho d, normal momentum rho S_K d, tangential rho v_t d,
//   energy E d + rho chi (S_M - u_K)(S_M + p/(rho (S_K - u_K))),
//   d = (S_M - u_K)/(S_K - S_M), chi = (S_K - u_K)/(S_K - S_M),
// so the FP32 rounding scales with the jump across the face and not with
// the size of the state (the pressure, the energy) as it would if U* and
// U_K were rounded separately and subtracted.
template <class T>
__device__ inline void hllc_t(double gamma, const T *cl, const T *cr, T *flux) {
    T wl[NV], wr[NV];
    cons2prim_t(gamma, cl, wl);
    cons2prim_t(gamma, cr, wr);
    const T rl = wl[0], uL = wl[1], vl = wl[2], zl = wl[3], pl = wl[4], el = cl[4];
    const T rr = wr[0], uR = wr[1], vr = wr[2], zr = wr[3], pr = wr[4], er = cr[4];
    const T ccl = sqrt(gamma * pl / rl), ccr = sqrt(gamma * pr / rr);
    const T srl = sqrt(rl), srr = sqrt(rr);
    const T cc2 = sqrt((((ccl * ccl) * srl) + ((ccr * ccr) * srr)) / (srl + srr) +
                       (0.5 * ((srl * srr) / ((srl + srr) * (srl + srr))) * ((uR - uL) * (uR - uL))));
    const T uu2 = ((uL * srl) + (uR * srr)) / (srl + srr);
    T sl = fmin(uL - ccl, uu2 - cc2), sr = fmax(uR + ccr, uu2 + cc2);
    sl = fmin(sl, 0.0);
    sr = fmax(sr, 0.0);
    const T sm = (pr - pl + (rl * uL * (sl - uL)) - (rr * uR * (sr - uR))) / ((rl * (sl - uL)) - (rr * (sr - uR)));
    const T fl[NV] = {rl * uL, (rl * (uL * uL)) + pl, rl * uL * vl, rl * uL * zl, uL * (el + pl)};
    const T fr[NV] = {rr * uR, (rr * (uR * uR)) + pr, rr * uR * vr, rr * uR * zr, uR * (er + pr)};
    const T mul = rl * ((sl - uL) / (sl - sm));
    const T mur = rr * ((sr - uR) / (sr - sm));
    const T lastl = (el / rl) + ((sm - uL) * (sm + ((pl) / (rl * (sl - uL)))));
    const T lastr = (er / rr) + ((sm - uR) * (sm + ((pr) / (rr * (sr - uR)))));
    const T usl[NV] = {mul, mul * sm, mul * vl, mul * zl, mul * lastl};
    const T usr[NV] = {mur, mur * sm, mur * vr, mur * zr, mur * lastr};
    const double s = copysign(1.0, val(sm));
#pragma unroll
    for (int k = 0; k < NV; ++k)
        flux[k] = (((1.0 + s) / 2.0) * (fl[k] + sl * (usl[k] - cl[k]))) +
                  (((1.0 - s) / 2.0) * (fr[k] + sr * (usr[k] - cr[k])));
}
__device__ inline void hllc64(double gamma, const double *cl, const double *cr, double *flux) {
    hllc_t<double>(gamma, cl, cr, flux);
}

// Reverse mode of hllc_t (FP64): given the flux cotangent hb, the cotangents
// of both rotated states.  The forward values repeat hllc_t; the branches
// (Einfeldt min/max, the clamps of S_L/S_R at 0, the side sign(S_M) picks)
// are taken as hllc_t takes them, ties included, so this is the transpose
// of the derivative the Dual passes of hllc_t compute.
__device__ inline void hllc_vjp(double g, const double *cl, const double *cr, const double *hb, double *clb,
                                double *crb) {
    const double gm = g - 1.0;
    // ---- forward ----
    const double rl = cl[0], ool = 1.0 / rl, uL = cl[1] * ool, vl = cl[2] * ool, zl = cl[3] * ool, el = cl[4];
    const double rr = cr[0], oor = 1.0 / rr, uR = cr[1] * oor, vr = cr[2] * oor, zr = cr[3] * oor, er = cr[4];
    const double ql = uL * uL + vl * vl + zl * zl, qr = uR * uR + vr * vr + zr * zr;
    const double pl = gm * (el - 0.5 * rl * ql), pr = gm * (er - 0.5 * rr * qr);
    const double ccl = sqrt(g * pl / rl), ccr = sqrt(g * pr / rr);
    const double srl = sqrt(rl), srr = sqrt(rr), S = srl + srr;
    const double du = uR - uL;
    const double T1 = ((ccl * ccl) * srl + (ccr * ccr) * srr) / S;
    const double cc2 = sqrt(T1 + (0.5 * ((srl * srr) / (S * S)) * (du * du)));
    const double uu2 = ((uL * srl) + (uR * srr)) / S;
    const double a1 = uL - ccl, a2 = uu2 - cc2, b1 = uR + ccr, b2 = uu2 + cc2;
    const bool selL = a2 < a1, selR = b2 > b1;  // the second argument wins
    double sl = selL ? a2 : a1, sr = selR ? b2 : b1;
    const bool clL = 0.0 < sl, clR = 0.0 > sr;  // clamped to 0
    if (clL) sl = 0.0;
    if (clR) sr = 0.0;
    const double num = pr - pl + (rl * uL * (sl - uL)) - (rr * uR * (sr - uR));
    const double den = (rl * (sl - uL)) - (rr * (sr - uR));
    const double sm = num / den;
    const bool left = copysign(1.0, sm) > 0.0;
    // ---- reverse through the chosen side K: F = F_K + S_K (U*_K - U_K) ----
    const double rK = left ? rl : rr, uK = left ? uL : uR, vK = left ? vl : vr, zK = left ? zl : zr;
    const double pK = left ? pl : pr, eK = left ? el : er, sK = left ? sl : sr;
    const double *cK = left ? cl : cr;
    const double dK = sK - uK, Q = sK - sm, mu = rK * (dK / Q);
    const double D = rK * dK, A = sm - uK, Bq = sm + pK / D;
    const double lastK = (eK / rK) + A * Bq;
    const double us[NV] = {mu, mu * sm, mu * vK, mu * zK, mu * lastK};
    double sKb = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) sKb += hb[k] * (us[k] - cK[k]);
    double usb[NV], cKb[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        usb[k] = sK * hb[k];
        cKb[k] = -sK * hb[k];
    }
    const double mub = usb[0] + usb[1] * sm + usb[2] * vK + usb[3] * zK + usb[4] * lastK;
    double smb = usb[1] * mu;
    double vKb = usb[2] * mu, zKb = usb[3] * mu;
    const double lastb = usb[4] * mu;
    double eKb = lastb / rK, rKb = -lastb * eK / (rK * rK);
    const double Ab = lastb * Bq, Bqb = lastb * A;
    smb += Ab + Bqb;
    double uKb = -Ab;
    double pKb = Bqb / D;
    const double Db = -Bqb * pK / (D * D);
    rKb += Db * dK;
    double dKb = Db * rK;
    rKb += mub * dK / Q;
    dKb += mub * rK / Q;
    const double Qb = -mub * mu / Q;
    sKb += Qb;
    smb -= Qb;
    sKb += dKb;
    uKb -= dKb;
    // F_K = [rK uK, rK uK^2 + pK, rK uK vK, rK uK zK, uK (eK + pK)]
    rKb += hb[0] * uK + hb[1] * uK * uK + hb[2] * uK * vK + hb[3] * uK * zK;
    uKb += hb[0] * rK + hb[1] * 2.0 * rK * uK + hb[2] * rK * vK + hb[3] * rK * zK + hb[4] * (eK + pK);
    pKb += hb[1] + hb[4] * uK;
    vKb += hb[2] * rK * uK;
    zKb += hb[3] * rK * uK;
    eKb += hb[4] * uK;
    // primitive cotangents of both sides
    double rlb = 0, ulb = 0, vlb = 0, zlb = 0, plb = 0, elb = 0, slb = 0;
    double rrb = 0, urb = 0, vrb = 0, zrb = 0, prb = 0, erb = 0, srb = 0;
    if (left) {
        rlb = rKb; ulb = uKb; vlb = vKb; zlb = zKb; plb = pKb; elb = eKb; slb = sKb;
    } else {
        rrb = rKb; urb = uKb; vrb = vKb; zrb = zKb; prb = pKb; erb = eKb; srb = sKb;
    }
    // S_M = num / den
    const double numb = smb / den, denb = -smb * sm / den;
    prb += numb;
    plb -= numb;
    rlb += numb * uL * (sl - uL);
    ulb += numb * rl * (sl - 2.0 * uL);
    slb += numb * rl * uL;
    rrb -= numb * uR * (sr - uR);
    urb -= numb * rr * (sr - 2.0 * uR);
    srb -= numb * rr * uR;
    rlb += denb * (sl - uL);
    slb += denb * rl;
    ulb -= denb * rl;
    rrb -= denb * (sr - uR);
    srb -= denb * rr;
    urb += denb * rr;
    // S_L, S_R: clamps and Einfeldt min / max
    double cclb = 0, ccrb = 0, cc2b = 0, uu2b = 0;
    if (!clL) {
        if (selL) {
            uu2b += slb;
            cc2b -= slb;
        } else {
            ulb += slb;
            cclb -= slb;
        }
    }
    if (!clR) {
        if (selR) {
            uu2b += srb;
            cc2b += srb;
        } else {
            urb += srb;
            ccrb += srb;
        }
    }
    ulb += uu2b * srl / S;
    urb += uu2b * srr / S;
    double srlb = uu2b * (uL - uu2) / S, srrb = uu2b * (uR - uu2) / S;
    const double Xb = cc2b / (2.0 * cc2);
    cclb += Xb * 2.0 * ccl * srl / S;
    ccrb += Xb * 2.0 * ccr * srr / S;
    srlb += Xb * (ccl * ccl - T1) / S;
    srrb += Xb * (ccr * ccr - T1) / S;
    const double S3 = S * S * S, hdu2 = 0.5 * du * du;
    srlb += Xb * hdu2 * srr * (srr - srl) / S3;
    srrb += Xb * hdu2 * srl * (srl - srr) / S3;
    const double dub = Xb * ((srl * srr) / (S * S)) * du;
    urb += dub;
    ulb -= dub;
    rlb += srlb / (2.0 * srl);
    rrb += srrb / (2.0 * srr);
    const double ylb = cclb / (2.0 * ccl), yrb = ccrb / (2.0 * ccr);
    plb += ylb * g / rl;
    rlb -= ylb * g * pl / (rl * rl);
    prb += yrb * g / rr;
    rrb -= yrb * g * pr / (rr * rr);
    // conserved -> primitive: rho = c0, u = c1/c0 ..., p = (g-1)(c4 - c0 q/2)
    auto back = [&](const double *c, double r_b, double u_b, double v_b, double z_b, double p_b, double e_b,
                    double u, double v, double z, double q, double *cb) {
        const double qb = -gm * 0.5 * c[0] * p_b;
        const double ub = u_b + 2.0 * u * qb, vb = v_b + 2.0 * v * qb, zb = z_b + 2.0 * z * qb;
        const double oo = 1.0 / c[0];
        cb[0] += r_b - gm * 0.5 * q * p_b - (ub * u + vb * v + zb * z) * oo;
        cb[1] += ub * oo;
        cb[2] += vb * oo;
        cb[3] += zb * oo;
        cb[4] += gm * p_b + e_b;
    };
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        clb[k] = left ? cKb[k] : 0.0;
        crb[k] = left ? 0.0 : cKb[k];
    }
    back(cl, rlb, ulb, vlb, zlb, plb, elb, uL, vl, zl, ql, clb);
    back(cr, rrb, urb, vrb, zrb, prb, erb, uR, vr, zr, qr, crb);
}

__device__ inline void hllc32(double gamma, const double *cl, const double *cr, double *flux) {
    double wl[NV], wr[NV];
    cons2prim_d(gamma, cl, wl);
    cons2prim_d(gamma, cr, wr);
    const float g = (float)gamma;
    const float rl = (float)wl[0], uL = (float)wl[1], pl = (float)wl[4];
    const float rr = (float)wr[0], uR = (float)wr[1], pr = (float)wr[4];
    const float ccl = sqrtf(g * pl / rl), ccr = sqrtf(g * pr / rr);
    const float srl = sqrtf(rl), srr = sqrtf(rr);
    const float cc2 = sqrtf((((ccl * ccl) * srl) + ((ccr * ccr) * srr)) / (srl + srr) +
                            (0.5f * ((srl * srr) / ((srl + srr) * (srl + srr))) * ((uR - uL) * (uR - uL))));
    const float uu2 = ((uL * srl) + (uR * srr)) / (srl + srr);
    float sl = fminf(uL - ccl, uu2 - cc2), sr = fmaxf(uR + ccr, uu2 + cc2);
    sl = fminf(sl, 0.0f);
    sr = fmaxf(sr, 0.0f);
    const float dp = (float)(wr[4] - wl[4]);
    const float sm = (dp + (rl * uL * (sl - uL)) - (rr * uR * (sr - uR))) / ((rl * (sl - uL)) - (rr * (sr - uR)));
    const bool left = !(copysignf(1.0f, sm) < 0.0f);
    const double *w = left ? wl : wr;
    const double *c = left ? cl : cr;
    const float sk = left ? sl : sr;
    const float rk = (float)w[0], uk = (float)w[1], pk = (float)w[4];
    // physical flux of the upwind side, FP64
    const double fk[NV] = {w[0] * w[1], (w[0] * (w[1] * w[1])) + w[4], w[0] * w[1] * w[2], w[0] * w[1] * w[3],
                           w[1] * (c[4] + w[4])};
    // star-state correction, FP32, factored
    const float den = sk - sm;
    const float d = (sm - uk) / den;
    const float chi = (sk - uk) / den;
    const float du[NV] = {rk * d, rk * sk * d, rk * (float)w[2] * d, rk * (float)w[3] * d,
                          (float)c[4] * d + rk * chi * (sm - uk) * (sm + pk / (rk * (sk - uk)))};
#pragma unroll
    for (int k = 0; k < NV; ++k) flux[k] = fk[k] + (double)(sk * du[k]);
}

template <typename R>
struct FluxParams {
    int nface, nf, nqm, qpq, qpt;
    double gamma;
    const int32_t *fcell;   // [nface][2] left (owner), right
    const int32_t *fslot;   // [nface][2] face slot in left, right; bit 8 of slot 0: triangle
    const double *frot;     // [nface][4] cos a1, sin a1, cos a2, sin a2 of the owner's outward normal
    const double *farea;    // [nface]
    const double *U;        // [ncell][5] cell averages
    const R *uleft;         // face-point deviations from the cell average
    double *fflux;          // [nface][5] outward from the owner
    const double *wq;       // [qpq] quadrilateral-face weights of this context, sum 1
    const double *wt;       // [qpt] triangular-face weights, sum 1
    ucns_boundary_v1 bc;    // free-stream data of the boundary codes
    const double *bface;    // [nface][10] per-face INFLOW(), OUTFLOW() states, or null: bc's
};

// ---------------------------------------------------------------------------
// Ghost states of boundary faces: BOUNDARYS (src/flow_operations.f90:2860-3375),
// Euler 3D without a rotating frame, in FP64 in both precisions.
#define UCNS_TOLSMALL 1.0e-13  // src/grid_p.f90:18

template <class T>
__device__ inline void bc_prim2cons(double gamma, const T *s3, T *cr) {
    cr[0] = s3[0];
    cr[1] = s3[1] * s3[0];
    cr[2] = s3[2] * s3[0];
    cr[3] = s3[3] * s3[0];
    const T skins = 0.5 * ((s3[1] * s3[1]) + (s3[2] * s3[2]) + (s3[3] * s3[3]));
    const T ikins = s3[4] / ((gamma - 1.0) * (s3[0]));
    cr[4] = (s3[0] * (ikins)) + (s3[0] * skins);
}

// subsonic inflow (src/flow_operations.f90:2907-2915): s1 free stream, s2 interior
template <class T, class F, class N>
__device__ inline void bc_sub_in(double gamma, const F *s1, const T *s2, T sps, N nx, N ny, N nz, T *cr) {
    T s3[NV];
    s3[4] = 0.5 * ((s1[4]) + (s2[4]) - (s2[0] * sps * ((nx * (s1[1] - s2[1])) + (ny * (s1[2] - s2[2])) + (nz * (s1[3] - s2[3])))));
    s3[0] = s1[0] + (s3[4] - s1[4]) / (sps * sps);
    s3[1] = s1[1] - (nx * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s1[2] - (ny * (s1[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s1[3] - (nz * (s1[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

// subsonic outflow (src/flow_operations.f90:2995-3010): s1 free stream, s2 interior
template <class T, class F, class N>
__device__ inline void bc_sub_out(double gamma, const F *s1, const T *s2, T sps, N nx, N ny, N nz, T *cr) {
    T s3[NV];
    s3[4] = T{s1[4]};
    s3[0] = s2[0] + (s3[4] - s2[4]) / (sps * sps);
    s3[1] = s2[1] + (nx * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[2] = s2[2] + (ny * (s2[4] - s3[4])) / (sps * s2[0]);
    s3[3] = s2[3] + (nz * (s2[4] - s3[4])) / (sps * s2[0]);
    bc_prim2cons(gamma, s3, cr);
}

// ROTATEF / ROTATEB with precomputed cos/sin of the face angles
template <class T>
__device__ inline void rotf(const T *a, double c1, double s1, double c2, double s2, T *r) {
    r[0] = a[0];
    r[1] = c1 * s2 * a[1] + s1 * s2 * a[2] + c2 * a[3];
    r[2] = c1 * c2 * a[1] + s1 * c2 * a[2] + (-s2) * a[3];
    r[3] = (-s1) * a[1] + c1 * a[2];
    r[4] = a[4];
}
template <class T>
__device__ inline void rotb(const T *h, double c1, double s1, double c2, double s2, T *r) {
    r[0] = h[0];
    r[1] = c1 * s2 * h[1] + c1 * c2 * h[2] + (-s1) * h[3];
    r[2] = s1 * s2 * h[1] + s1 * c2 * h[2] + c1 * h[3];
    r[3] = c2 * h[1] + (-s2) * h[2];
    r[4] = h[4];
}

// The face frame of the rotations: the angles' cosines and sines as the
// primal kernels use them, or the frame vectors themselves (normal n and
// tangents t1, t2, rows of the rotation) in a scalar type C that may carry a
// tangent.  Both give the same rotation: n = (c1 s2, s1 s2, c2),
// t1 = (c1 c2, s1 c2, -s2), t2 = (-s1, c1, 0).
struct AngleFrame {
    double c1, s1, c2, s2;
    template <class T>
    __device__ void fwd(const T *a, T *r) const { rotf(a, c1, s1, c2, s2, r); }
    template <class T>
    __device__ void back(const T *h, T *r) const { rotb(h, c1, s1, c2, s2, r); }
    __device__ double nx() const { return c1 * s2; }
    __device__ double ny() const { return s1 * s2; }
    __device__ double nz() const { return c2; }
};
template <class C>
struct VectorFrame {
    C n[3], t1[3], t2[3];
    template <class T>
    __device__ void fwd(const T *a, T *r) const {
        r[0] = a[0];
        r[1] = n[0] * a[1] + n[1] * a[2] + n[2] * a[3];
        r[2] = t1[0] * a[1] + t1[1] * a[2] + t1[2] * a[3];
        r[3] = t2[0] * a[1] + t2[1] * a[2] + t2[2] * a[3];
        r[4] = a[4];
    }
    template <class T>
    __device__ void back(const T *h, T *r) const {
        r[0] = h[0];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[1 + d] = n[d] * h[1] + t1[d] * h[2] + t2[d] * h[3];
        r[4] = h[4];
    }
    __device__ C nx() const { return n[0]; }
    __device__ C ny() const { return n[1]; }
    __device__ C nz() const { return n[2]; }
};

// fin/fout: the conserved free-stream states (inflow, outflow) of the
// boundary data, of the scalar type F (Dual when their tangent is carried)
template <class T, class F, class Frame>
__device__ inline void ghost_state_fr(int code, int boundtype, double gamma, const T *cl, const Frame &fr,
                                      const F *fin, const F *fout, T *cr) {
    const auto nx = fr.nx(), ny = fr.ny(), nz = fr.nz();
    F f[NV];
    T w[NV];
    switch (code) {
    case 1:  // inflow (src/flow_operations.f90:2880-2925)
        if (boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = T{fin[v]};
        } else {
            cons2prim_t(gamma, fin, f);
            cons2prim_t(gamma, cl, w);
            const T sps = sqrt((gamma * w[4]) / (w[0]));
            const T vel = sqrt(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
            if (val(vel / (sps + UCNS_TOLSMALL)) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = T{fin[v]};
            } else {
                bc_sub_in(gamma, f, w, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 2:  // outflow (src/flow_operations.f90:2965-3013)
        if (boundtype == 0) {
            for (int v = 0; v < NV; ++v) cr[v] = cl[v];
        } else {
            cons2prim_t(gamma, fout, f);
            cons2prim_t(gamma, cl, w);
            const T sps = sqrt((gamma * w[4]) / (w[0]));
            const T vel = sqrt(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
            if (val(vel / (sps + UCNS_TOLSMALL)) > 1.0) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                bc_sub_out(gamma, f, w, sps, nx, ny, nz, cr);
            }
        }
        break;
    case 3:  // symmetry (src/flow_operations.f90:3086-3131)
    case 4:  // wall, inviscid branch (src/flow_operations.f90:3132-3160)
    {
        T rot[NV];
        fr.fwd(cl, rot);
        rot[1] = -rot[1];
        fr.back(rot, cr);
        break;
    }
    default:  // 6, farfield (src/flow_operations.f90:3203-3296)
    {
        T rot[NV];
        fr.fwd(cl, rot);
        const T vnb = rot[1];
        cons2prim_t(gamma, cl, w);
        const T sps = sqrt((gamma * w[4]) / (w[0]));
        if (val(vnb) <= 0.0) {
            if (val(fabs(vnb)) >= val(sps)) {
                for (int v = 0; v < NV; ++v) cr[v] = T{fin[v]};
            } else {
                cons2prim_t(gamma, fin, f);
                bc_sub_in(gamma, f, w, sps, nx, ny, nz, cr);
            }
        } else {
            if (val(fabs(vnb)) >= val(sps)) {
                for (int v = 0; v < NV; ++v) cr[v] = cl[v];
            } else {
                cons2prim_t(gamma, fout, f);
                bc_sub_out(gamma, f, w, sps, nx, ny, nz, cr);
            }
        }
    }
    }
}

template <class T, class F>
__device__ inline void ghost_state(int code, int boundtype, double gamma, const T *cl, double c1, double s1, double c2,
                                   double s2, const F *fin, const F *fout, T *cr) {
    ghost_state_fr(code, boundtype, gamma, cl, AngleFrame{c1, s1, c2, s2}, fin, fout, cr);
}

template <typename R>
__global__ __launch_bounds__(256) void k_flux(FluxParams<R> p) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= p.nface) return;
    const int cL = p.fcell[2 * f], cR = p.fcell[2 * f + 1];
    const int s0 = p.fslot[2 * f], sR = p.fslot[2 * f + 1];
    const int sL = s0 & 255;
    const bool tri = (s0 >> 8) & 1;
    const int nq = tri ? p.qpt : p.qpq;
    const double *w = tri ? p.wt : p.wq;
    const double c1 = p.frot[4 * f], s1 = p.frot[4 * f + 1], c2 = p.frot[4 * f + 2], s2 = p.frot[4 * f + 3];
    const double area = p.farea[f];
    const bool bnd = cR < 0;  // boundary face: -code
    const double *fin = p.bface ? p.bface + (size_t)f * 10 : p.bc.inflow;
    const double *fout = p.bface ? p.bface + (size_t)f * 10 + 5 : p.bc.outflow;
    double uL[NV], uR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        uL[v] = p.U[(size_t)cL * NV + v];
        uR[v] = bnd ? 0.0 : p.U[(size_t)cR * NV + v];
    }
    const R *ulL = p.uleft + ((size_t)cL * p.nf + sL) * p.nqm * NV;
    const R *ulR = p.uleft + ((size_t)(bnd ? cL : cR) * p.nf + (bnd ? sL : sR)) * p.nqm * NV;
    double acc[NV] = {0, 0, 0, 0, 0};
    // QB points' deviations are loaded before any is used (indices clamped
    // to the face's nq points; the surplus is not evaluated)
    constexpr int QB = 3;
    R dl[QB][NV], dr[QB][NV];
    for (int q = 0; q < nq; ++q) {
        const int qb = q % QB;
        if (qb == 0) {
#pragma unroll
            for (int u = 0; u < QB; ++u) {
                const int qq = min(q + u, nq - 1);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    dl[u][v] = ulL[qq * NV + v];
                    dr[u][v] = ulR[qq * NV + v];
                }
            }
        }
        double a[NV], bq[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R xl = dl[0][v], xr = dr[0][v];
#pragma unroll
            for (int u = 1; u < QB; ++u)
                if (qb == u) {
                    xl = dl[u][v];
                    xr = dr[u][v];
                }
            a[v] = uL[v] + (double)xl;
            bq[v] = uR[v] + (double)xr;
        }
        if (bnd) ghost_state(-cR, p.bc.boundtype, p.gamma, a, c1, s1, c2, s2, fin, fout, bq);  // GET_STATES_BOUNDS
        // ROTATEF (src/grid_t.f90:4217-4258)
        const double la[NV] = {a[0], c1 * s2 * a[1] + s1 * s2 * a[2] + c2 * a[3],
                               c1 * c2 * a[1] + s1 * c2 * a[2] + (-s2) * a[3], (-s1) * a[1] + c1 * a[2], a[4]};
        const double ra[NV] = {bq[0], c1 * s2 * bq[1] + s1 * s2 * bq[2] + c2 * bq[3],
                               c1 * c2 * bq[1] + s1 * c2 * bq[2] + (-s2) * bq[3], (-s1) * bq[1] + c1 * bq[2], bq[4]};
        double hd[NV];
        if (sizeof(R) == 8)
            hllc64(p.gamma, la, ra, hd);
        else
            hllc32(p.gamma, la, ra, hd);
        // ROTATEB (src/grid_t.f90:4261-4305)
        const double rb[NV] = {hd[0], c1 * s2 * hd[1] + c1 * c2 * hd[2] + (-s1) * hd[3],
                               s1 * s2 * hd[1] + s1 * c2 * hd[2] + c1 * hd[3], c2 * hd[1] + (-s2) * hd[2], hd[4]};
        const double wa = w[q] * area;
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = acc[v] + (rb[v] * wa);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) p.fflux[(size_t)f * NV + v] = acc[v];
}

struct UpdateParams {
    int ncell, nf, stage;
    double cfl, gamma;
    const int32_t *cface;    // [ncell][nf] +(face+1) owner, -(face+1) neighbour, 0 none
    const double *fflux;
    const double *ovol;      // [ncell] 1/TOTVOLUME
    const double *minedge;
    double *U, *U2;
    double *rhs_out;         // optional [ncell][5]
    const double *dt;        // this step's dt
    unsigned long long *dt_next;  // atomic-min target (bits of a positive double), stage 3 only
    double *t;               // advanced by dt in stage 3
};

__device__ inline void atomic_min_pos(unsigned long long *addr, double v) {
    atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

// stage 0: gather only (rhs_out); stage 1..3: Shu-Osher update
__global__ __launch_bounds__(256) void k_update(UpdateParams p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dtl = 1.0e300;
    if (i < p.ncell) {
        // every face's flux is loaded before the first is summed (face
        // slots past nf, and slots with no face, read face 0 and add nothing)
        constexpr int NFM = 6;
        int cf[NFM];
#pragma unroll
        for (int f = 0; f < NFM; ++f) cf[f] = f < p.nf ? p.cface[(size_t)i * p.nf + f] : 0;
        double F[NFM][NV];
#pragma unroll
        for (int f = 0; f < NFM; ++f) {
            const size_t id = cf[f] != 0 ? (size_t)(abs(cf[f]) - 1) : 0;
#pragma unroll
            for (int v = 0; v < NV; ++v) F[f][v] = p.fflux[id * NV + v];
        }
        double r[NV] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int f = 0; f < NFM; ++f) {
            if (cf[f] > 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) r[v] = r[v] + F[f][v];
            } else if (cf[f] < 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) r[v] = r[v] - F[f][v];
            }
        }
        if (p.rhs_out) {
#pragma unroll
            for (int v = 0; v < NV; ++v) p.rhs_out[(size_t)i * NV + v] = r[v];
        }
        if (p.stage > 0) {
            const double dt = *p.dt, oov = p.ovol[i];
            double *u = p.U + (size_t)i * NV, *u2 = p.U2 + (size_t)i * NV;
            const double to4 = 3.0 / 4.0, oo4 = 1.0 / 4.0, to3 = 2.0 / 3.0, oo3 = 1.0 / 3.0;
            double un[NV];
            if (p.stage == 1) {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    u2[v] = u[v];
                    un[v] = u[v] - (dt * (r[v] * oov));
                }
            } else if (p.stage == 2) {
#pragma unroll
                for (int v = 0; v < NV; ++v) un[v] = (to4 * u2[v]) + (oo4 * u[v]) - (oo4 * (dt * (r[v] * oov)));
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) un[v] = (oo3 * u2[v]) + (to3 * u[v]) - (to3 * (dt * (r[v] * oov)));
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) u[v] = un[v];
            if (p.stage == 3 && p.dt_next) {
                double w[NV];
                cons2prim_d(p.gamma, un, w);
                const double c = sqrt(w[4] * p.gamma / w[0]);
                const double veln = fmax(fmax(fabs(w[1]), fabs(w[2])), fabs(w[3])) + c;
                dtl = (p.cfl / 3.0) * (p.minedge[i] / fabs(veln));
            }
        }
    }
    if (p.stage == 3) {
        if (p.dt_next) {
            dtl = wave_min(dtl);
            if ((threadIdx.x & (WAVE - 1)) == 0 && dtl < 1.0e300) atomic_min_pos(p.dt_next, dtl);
        }
        if (i == 0 && p.t) *p.t += *p.dt;
    }
}

struct DtParams {
    int ncell;
    double cfl, gamma;
    const double *U, *minedge;
    unsigned long long *dt_out;
};

__global__ __launch_bounds__(256) void k_dt(DtParams p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 1.0e300;
    if (i < p.ncell) {
        double w[NV];
        cons2prim_d(p.gamma, p.U + (size_t)i * NV, w);
        const double c = sqrt(w[4] * p.gamma / w[0]);
        const double veln = fmax(fmax(fabs(w[1]), fabs(w[2])), fabs(w[3])) + c;
        d = (p.cfl / 3.0) * (p.minedge[i] / fabs(veln));
    }
    d = wave_min(d);
    if ((threadIdx.x & (WAVE - 1)) == 0 && d < 1.0e300) atomic_min_pos(p.dt_out, d);
}

// ---------------------------------------------------------------------------
// Adjoint of the residual: vector-Jacobian products w -> (dR/dU)^T w.
//
// The residual is R_i = sum over the faces of cell i of +-F_f, F_f the
// quadrature of ROTATEB(HLLC(ROTATEF(a_q), ROTATEF(b_q))) w_q |f|, with face
// states a_q = U_L + ULEFT_L(q) and b_q = U_R + ULEFT_R(q) (or the ghost
// state of a_q).  The reverse sweep runs in two kernels:
//   k_flux_vjp  face-parallel: F_bar = w_owner - w_neighbour, and through
//               each quadrature point the cotangents of a_q and b_q, written
//               to the face-point slots of the two cells (each slot belongs
//               to exactly one face, so no atomics);
//   k_recon_vjp tile of 8 cells per wave, lane = 8 * cell + stencil: repeats
//               the forward decisions of the reconstruction bit for bit
//               (same arithmetic and order), then pulls the face-point
//               cotangents back through the extrapolation, the WENO weights
//               and indicators or the Barth limiter, and scatters
//               P_s^T g_bar_s to the stencil members with FP64 atomics.
// Discrete decisions (CHECKSOL / CHECKSOLX flags, HLLC and boundary
// regimes, the limiter's arg-min face point and arg-max/min neighbour) are
// those of the state the product is taken at, so the result is the exact
// derivative of the branch the primal takes.

// pseudo-inverse rows and stencil members in whichever device layout the
// forward kernels use: 0 per cell [cell][j][s][k] + [cell][s][j],
// 1 tiled [tile][j][k][8 ts] + [tile][j][8 ts], 2 the row stream
template <typename R>
struct LsqView {
    int mode, ts, tp, imax, ideg, depth, hd, pb, rd;
    long long tile_bytes;
    const int32_t *nbr;
    const R *P;
    const unsigned char *rows;
    __device__ int member(int tile, int cell, int s, int j, int pos) const {
        if (mode == 0) return nbr[((size_t)cell * ts + s) * imax + j];
        if (mode == 1) return nbr[((size_t)tile * imax + j) * tp + pos];
        const unsigned char *b = rows + (size_t)tile * tile_bytes;
        return j < depth ? ((const int32_t *)b)[j * tp + pos]
                         : ((const int32_t *)(b + hd + (size_t)(j - depth) * rd + pb))[pos];
    }
    // coefficient k of member j is row[k * kst]
    __device__ const R *row(int tile, int cell, int s, int j, int pos, int &kst) const {
        if (mode == 0) {
            kst = 1;
            return P + (((size_t)cell * imax + j) * ts + s) * ideg;
        }
        kst = tp;
        if (mode == 1) return P + ((size_t)tile * imax + j) * ideg * tp + pos;
        return (const R *)(rows + (size_t)tile * tile_bytes + hd + (size_t)j * rd) + pos;
    }
};

template <typename R>
struct VjpParams {
    int ncell, ntile, tiles_per_xcd, nf, nqm, qpq, qpt, iweno;
    double gamma, lwci1, eps;
    const double *U;
    LsqView<R> lsq;
    const R *B;            // [ncell][ideg*(ideg+1)/2]
    const R *mbar;         // [ncell][ideg]
    const R *qp;           // [ncell][nf][nqm][3]
    const int32_t *cinfo;
    const int32_t *cnb;    // [ncell][nf]
    const double *ubf;     // [ncell][nf][nqm][5] cotangents of the face-point states
    double *ubar;          // [ncell][5] accumulated
    const R *gload;        // row stream: the forward's stencil coefficients, indicators and
                           // weights [ntile][ideg*5 + 10][8 ts]
    const int *fload;      // row stream: the forward's per-cell flags (1 MUSCL, 2 first order)
    // the tile's distinct stencil members: member (j, pos) of tile t scatters
    // to table row slot[t][j][pos] (UCNS_NOSLOT: straight to HBM), and row r
    // of tile t belongs to cell uq[t][r], r < min(ucount[t], uq_cap); rows
    // past a kernel's LDS table also go straight to HBM and are written as 0
    const uint16_t *slot;
    const int32_t *uq, *ucount;
    int uq_cap;
    double *gbout;         // row stream: each stencil's g_bar [ntile][ideg*5][8 ts], for the scatter
    const int64_t *toff;   // first row of each tile's member table in tab
    double *tab;           // [rows][5] the member tables, summed per cell by k_adj_gather
    // geometry cotangents (k_recon_vjp with GEO, FP64), accumulated: pseudo-
    // inverses [ncell][lsq ts][imax][ideg], indicators [ncell][ideg][ideg]
    // (upper triangle, the part the library reads), basis means [ncell][ideg]
    // and face points [ncell][nf][nqm][3]
    double *gP, *gB, *gM, *gQ;
};

__device__ inline void atomic_add_f64(double *a, double v) { unsafeAtomicAdd(a, v); }


// Distinct members per tile: a wave's scatter goes through an LDS table of
// the tile's distinct member cells (a tile of 8 neighbouring cells shares
// most of its stencil members), flushed with one HBM atomic per member and
// variable instead of one per (stencil, member, variable).
constexpr int HT = 4096;  // hash slots, >= 2 * the 8 * 8 * 39 member entries of a tile

template <typename R, int IMAX>
__global__ __launch_bounds__(WAVE) void k_tile_members(LsqView<R> L, int ntile, int ncell, const int32_t *cinfo,
                                                       int cap, int32_t *uq, int32_t *ucount, uint16_t *slot) {
    __shared__ int32_t keys[HT];
    __shared__ uint16_t cid[HT];
    __shared__ uint16_t hs[IMAX * 64];
    const int lane = threadIdx.x;
    const int tile = blockIdx.x;
    if (tile >= ntile) return;
    const int ts = L.ts, tp = L.tp;
    for (int h = lane; h < HT; h += WAVE) keys[h] = -1;
    __syncthreads();
    // insert every (member row j, lane position) entry; linear probing
    for (int e = lane; e < IMAX * tp; e += WAVE) {
        const int j = e / tp, pos = e - j * tp;
        const int lc = pos / ts, s = pos - lc * ts;
        const int cell = tile * 8 + lc;
        int h = 0xFFFF;
        if (cell < ncell && s < ci_admis(cinfo[cell])) {
            const int n = L.member(tile, cell, s, j, pos);
            h = (int)(((unsigned)n * 2654435761u) >> 20) & (HT - 1);
            for (int probe = 0; probe < HT; ++probe) {
                const int old = atomicCAS(&keys[h], -1, n);
                if (old == -1 || old == n) break;
                h = (h + 1) & (HT - 1);
            }
        }
        hs[e] = (uint16_t)h;
    }
    __syncthreads();
    // compact: lane l owns hash slots [64 l, 64 l + 64)
    constexpr int PER = HT / WAVE;
    int cnt = 0;
    for (int k = 0; k < PER; ++k) cnt += keys[lane * PER + k] >= 0;
    int incl = cnt;
    for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
    }
    int id = incl - cnt;
    for (int k = 0; k < PER; ++k) {
        const int h = lane * PER + k;
        const int key = keys[h];
        if (key >= 0) {
            cid[h] = id < cap ? (uint16_t)id : (uint16_t)UCNS_NOSLOT;
            if (id < cap) uq[(size_t)tile * cap + id] = key;
            ++id;
        }
    }
    if (lane == WAVE - 1) ucount[tile] = incl;
    __syncthreads();
    for (int e = lane; e < IMAX * tp; e += WAVE) {
        const int h = hs[e];
        slot[(size_t)tile * IMAX * tp + e] = h == 0xFFFF ? (uint16_t)UCNS_NOSLOT : cid[h];
    }
}

// pw(x, x^2, x^3, e) = x^e for e = 0..3
template <typename R>
__device__ inline R pw(R x, R x2, R x3, int e) {
    return e == 0 ? R(1) : e == 1 ? x : e == 2 ? x2 : x3;
}

// Reconstruction product on the tiled and per-cell layouts: the forward is
// repeated with its own arithmetic (the same bits, so the same decisions),
// then g_bar of each stencil is scattered through its pseudo-inverse rows.
template <typename R, int IDEG, int IMAX, int NQL, bool GEO = false>
__global__ __launch_bounds__(WAVE) void k_recon_vjp(VjpParams<R> p) {
    constexpr int NP = IDEG * (IDEG + 1) / 2;
    constexpr int KL = (IDEG + 7) / 8;    // coefficients per lane in the face-point pass
    __shared__ double sGb[8][IDEG * NV];  // cotangent of the combined coefficients
    __shared__ R sGh[8][IDEG * NV];       // combined coefficients sum_s omega_s G_s
    __shared__ R sG0[8][IDEG * NV];       // central-stencil coefficients
    __shared__ R sB[8][NP];               // the tile's indicator matrices
    __shared__ double sAcc[UCNS_TILE_UQ][NV];  // scatter to the tile's distinct members
    const int lane = threadIdx.x, lc = lane >> 3, s = lane & 7;
    const int b = blockIdx.x;
    const int tile = (b % NXCD) * p.tiles_per_xcd + b / NXCD;
    if (tile >= p.ntile) return;
    const int nuq = p.slot ? min(p.ucount[tile], p.uq_cap) : 0;
    for (int e = lane; e < min(nuq, UCNS_TILE_UQ) * NV; e += WAVE) (&sAcc[0][0])[e] = 0.0;
    if (p.iweno == 1) {
        const int nvc = min(8, p.ncell - tile * 8);
        const R *Bt = p.B + (size_t)tile * 8 * NP;
        for (int e = lane; e < nvc * NP; e += WAVE) (&sB[0][0])[e] = Bt[e];
    }
    const int cell = tile * 8 + lc;
    const bool cvalid = cell < p.ncell;
    const int cs = cvalid ? cell : 0;
    const int info = cvalid ? p.cinfo[cell] : 0;
    const int ifca = ci_ifca(info), admis = ci_admis(info), full = ci_full(info);
    const bool act = cvalid && s < admis;
    const int pos = lc * p.lsq.ts + s;
    double ui[NV] = {1.0, 0.0, 0.0, 0.0, 1.0};
    if (cvalid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) ui[v] = p.U[(size_t)cell * NV + v];
    }
    const int nslot = ifca * p.nqm;
    const R *qpc = p.qp + (size_t)cs * p.nf * p.nqm * 3;
    const R *mb = p.mbar + (size_t)cs * IDEG;
    const double *ubc = p.ubf + (size_t)cs * p.nf * p.nqm * NV;
    auto qvalid = [&](int q) {
        const int f = q / p.nqm, qq = q - f * p.nqm;
        return cvalid && q < nslot && qq < (ci_tri(info, f) ? p.qpt : p.qpq);
    };

    // ---- A: cotangent of the combined coefficients, sum_q phi_q ubar_q
    // (lane s of a cell takes coefficients s, s+8, s+16 over all its face
    // points), and of the cell average through a_q = U_i + ULEFT(q) ----
    double usum[NV] = {0, 0, 0, 0, 0};
    {
        double ab[KL][NV];
        R mk[KL];
        int ex[KL][3];
#pragma unroll
        for (int t = 0; t < KL; ++t) {
            const int k = s + 8 * t;
            mk[t] = k < IDEG ? mb[k] : R(0);
#pragma unroll
            for (int d = 0; d < 3; ++d) ex[t][d] = k < IDEG ? c_mono[k][d] : 0;
#pragma unroll
            for (int v = 0; v < NV; ++v) ab[t][v] = 0.0;
        }
#pragma unroll 1
        for (int q = 0; q < nslot; ++q) {
            if (!qvalid(q)) continue;
            const R x0 = qpc[q * 3], x1 = qpc[q * 3 + 1], x2 = qpc[q * 3 + 2];
            const R x0s = x0 * x0, x1s = x1 * x1, x2s = x2 * x2;
            const R x0c = x0s * x0, x1c = x1s * x1, x2c = x2s * x2;
            double ub[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                ub[v] = ubc[q * NV + v];
                usum[v] += ub[v];
            }
#pragma unroll
            for (int t = 0; t < KL; ++t) {
                const double ph = (double)(pw(x0, x0s, x0c, ex[t][0]) * pw(x1, x1s, x1c, ex[t][1]) *
                                               pw(x2, x2s, x2c, ex[t][2]) - mk[t]);
#pragma unroll
                for (int v = 0; v < NV; ++v) ab[t][v] += ph * ub[v];
            }
        }
#pragma unroll
        for (int t = 0; t < KL; ++t) {
            const int k = s + 8 * t;
            if (k < IDEG) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sGb[lc][k * NV + v] = ab[t][v];
            }
        }
    }
    __syncthreads();  // sGb, sB

    // ---- B: the stencil coefficients G_s = P_s dU_s, as the forward ----
    R g[IDEG][NV];
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[k][v] = R(0);
    if (act) {
#pragma unroll 1
        for (int j = 0; j < IMAX; ++j) {
            const int n = p.lsq.member(tile, cell, s, j, pos);
            R du[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) du[v] = (R)(p.U[(size_t)n * NV + v] - ui[v]);
            int kst;
            const R *row = p.lsq.row(tile, cell, s, j, pos, kst);
            R pc[IDEG];
#pragma unroll
            for (int k = 0; k < IDEG; ++k) pc[k] = row[(size_t)k * kst];
#pragma unroll
            for (int k = 0; k < IDEG; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) g[k][v] += pc[k] * du[v];
        }
    }
    if (s == 0) {
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) sG0[lc][k * NV + v] = g[k][v];
    }

    // ---- C: indicators, weights and combined coefficients (forward, with
    // its arithmetic: the same bits) ----
    R beta[NV] = {R(0), R(0), R(0), R(0), R(0)};
    R om[NV] = {R(0), R(0), R(0), R(0), R(0)};
    double obar[NV] = {0, 0, 0, 0, 0};
    double wi[NV];
    cons2prim_d(p.gamma, ui, wi);
    bool reduce = cvalid;
    if (p.iweno == 1) {
        if (act) {
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                R bg[NV] = {R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
                for (int l = 0; l < IDEG; ++l) {
                    const R bkl = sB[lc][pidx(IDEG, k, l)];
#pragma unroll
                    for (int v = 0; v < NV; ++v) bg[v] += bkl * g[l][v];
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) beta[v] += g[k][v] * bg[v];
            }
        }
        const R lam = (s == 0) ? (R)p.lwci1 : R(1);
        const R eps = (R)p.eps;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R w;
            if (sizeof(R) == 8) {
                const R x = eps + beta[v];
                const R x2 = x * x;
                w = act ? lam / (x2 * x2) : R(0);
            } else {
                const R x = eps + beta[v];
                const R xmin = grp_min8(act ? x : (R)INFINITY);
                const R r = xmin / x;
                const R r2 = r * r;
                w = act ? lam * (r2 * r2) : R(0);
            }
            const R sum = grp_sum8(w);
            om[v] = act ? w / sum : R(0);
        }
#pragma unroll
        for (int k = 0; k < IDEG; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R gh = grp_sum8(om[v] * g[k][v]);
                if (s == 0) sGh[lc][k * NV + v] = gh;
                obar[v] += sGb[lc][k * NV + v] * (double)g[k][v];
            }
    }
    __syncthreads();  // sGb, sGh, sG0

    // ---- D: CHECKSOL (0.85) on the WENO face values (forward) ----
    if (p.iweno == 1) {
        int trouble = 0;
#pragma unroll 1
        for (int q = s; q < NQL * 8; q += 8) {
            if (!qvalid(q)) continue;
            R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
            R phi[IDEG];
            basis_at<R, IDEG>(x, mb, phi);
            double uq[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R res = R(0);
#pragma unroll
                for (int k = 0; k < IDEG; ++k) res += phi[k] * sGh[lc][k * NV + v];
                uq[v] = ui[v] + (double)res;
            }
            if (full && jump_test(p.gamma, wi, uq, 0.85)) trouble = 1;
        }
        const int any_trouble = grp_or8(trouble);
        reduce = cvalid && (!full || any_trouble);
    }
    __syncthreads();  // sG0

    // ---- E: MUSCL + Barth (forward), with the limiter's active arguments ----
    bool first = false;
    double slope[NV], d2s[NV];
    int qs[NV], ext[NV];
    R res0[NQL][NV];
    if (reduce) {
        double umin[NV], umax[NV];
        int imin[NV], imax_[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            umin[v] = umax[v] = ui[v];
            imin[v] = imax_[v] = cell;
        }
        for (int f = 0; f < ifca; ++f) {
            const int n = p.cnb[(size_t)cell * p.nf + f];
            if (n < 0) continue;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double un = p.U[(size_t)n * NV + v];
                if (un < umin[v]) imin[v] = n;
                if (un > umax[v]) imax_[v] = n;
                umin[v] = fmin(umin[v], un);
                umax[v] = fmax(umax[v], un);
            }
        }
        double lmin[NV];
        int lq[NV], lw[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lmin[v] = 1.0e16;
            lq[v] = -1;
            lw[v] = 0;
        }
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
#pragma unroll
            for (int v = 0; v < NV; ++v) res0[r][v] = R(0);
            if (qvalid(q)) {
                R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
                R phi[IDEG];
                basis_at<R, IDEG>(x, mb, phi);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    R res = R(0);
#pragma unroll
                    for (int k = 0; k < IDEG; ++k) res += phi[k] * sG0[lc][k * NV + v];
                    res0[r][v] = res;
                    const double d2 = (double)res;
                    double psi;
                    int wsel = 0;
                    if (fabs(d2) <= 0.0)
                        psi = 1.0;
                    else if (d2 > 0.0)
                        psi = fmin(1.0, (umax[v] - ui[v]) / d2);
                    else {
                        psi = fmin(1.0, (umin[v] - ui[v]) / d2);
                        wsel = 1;
                    }
                    if (psi < lmin[v]) {
                        lmin[v] = psi;
                        lq[v] = q;
                        lw[v] = wsel;
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            slope[v] = grp_min8(lmin[v]);
            // lowest lane of the group holding the minimum
            int cand = (lmin[v] == slope[v]) ? s : 8;
            cand = min(cand, __shfl_xor(cand, 1, WAVE));
            cand = min(cand, __shfl_xor(cand, 2, WAVE));
            cand = min(cand, __shfl_xor(cand, 4, WAVE));
            const int src = lc * 8 + cand;
            qs[v] = lane_bcast(lq[v], src);
            const int wsel = lane_bcast(lw[v], src);
            ext[v] = wsel ? imin[v] : imax_[v];
            d2s[v] = 0.0;
        }
        // the unlimited value at the arg-min point, for the active limiters
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double mine = 0.0;
#pragma unroll
            for (int r = 0; r < NQL; ++r)
                if (s + 8 * r == qs[v]) mine = (double)res0[r][v];
            d2s[v] = grp_sum8(mine);
        }
        int tr2 = 0;
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            if (qvalid(q)) {
                double uq[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) uq[v] = ui[v] + (double)(R)((double)res0[r][v] * slope[v]);
                if (jump_test(p.gamma, wi, uq, 0.95)) tr2 = 1;
            }
        }
        first = grp_or8(tr2) != 0;
    }

    // ---- F: cotangent of each lane's stencil coefficients ----
    double gb[IDEG][NV];
#pragma unroll
    for (int k = 0; k < IDEG; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) gb[k][v] = 0.0;
    auto put = [&](int k, int v, double x) { gb[k][v] = x; };
    double ubi[NV];  // extra cotangent of U_i from the limiter
    double cq[NV], bb[NV];  // limiter correction at the arg-min point; indicator cotangents
    bool scatter = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) ubi[v] = cq[v] = bb[v] = 0.0;
    if (cvalid && !first && reduce) {
        // ULEFT(q) = slope * d2_q: slope_bar = sum_q ubar_q d2_q; through an
        // active limiter slope = (u_ext - u_i) / d2_{q*}
        double sb[NV] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            if (qvalid(q)) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sb[v] += ubc[q * NV + v] * (double)res0[r][v];
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sb[v] = grp_sum8(sb[v]);
            const bool on = slope[v] < 1.0 && qs[v] >= 0;
            const double e = on ? sb[v] / d2s[v] : 0.0;
            cq[v] = on ? e * slope[v] : 0.0;
            ubi[v] = -e;
            if (s == 0 && on) atomic_add_f64(&p.ubar[(size_t)ext[v] * NV + v], e);
        }
        scatter = s == 0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R phi[IDEG];
            const bool corr = s == 0 && cq[v] != 0.0;
            if (corr) {
                const int q = qs[v];
                R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
                basis_at<R, IDEG>(x, mb, phi);
            }
#pragma unroll
            for (int k = 0; k < IDEG; ++k)
                put(k, v, s == 0 ? slope[v] * sGb[lc][k * NV + v] - (corr ? cq[v] * (double)phi[k] : 0.0) : 0.0);
        }
    } else if (cvalid && !first) {
        // omega_s = lam_s (eps+beta_s)^-4 / sum_t (...):
        // beta_bar_s = -4 omega_s/(eps+beta_s) (obar_s - sum_t omega_t obar_t)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double so = grp_sum8(act ? (double)om[v] * obar[v] : 0.0);
            bb[v] = act ? -4.0 * (double)om[v] / ((double)p.eps + (double)beta[v]) * (obar[v] - so) : 0.0;
        }
        scatter = act;
        // g_bar_s = omega_s gh_bar + 2 beta_bar_s B G_s
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                double hg = 0.0;
#pragma unroll
                for (int l = 0; l < IDEG; ++l) hg += (double)sB[lc][pidx(IDEG, k, l)] * (double)g[l][v];
                put(k, v, act ? (double)om[v] * sGb[lc][k * NV + v] + 2.0 * bb[v] * hg : 0.0);
            }
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int k = 0; k < IDEG; ++k) put(k, v, 0.0);
    }

    if constexpr (GEO) {
        // ---- geometry: face points and basis means.  A face value is
        // U_i + sum_k phi_k(x_q) c_k with phi_k = m_k - mbar_k and c the
        // combined WENO coefficients (sGh) or, on the MUSCL path, slope times
        // the central ones (sG0, with the limiter's own dependence on the
        // arg-min point's value); phi_bar_k(q) = sum_v r_bar_v(q) c_kv ----
        const bool weno = cvalid && !first && !reduce && p.iweno == 1;
        const bool muscl = cvalid && !first && reduce;
        double mba[IDEG];
#pragma unroll
        for (int k = 0; k < IDEG; ++k) mba[k] = 0.0;
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            if (!(weno || muscl) || !qvalid(q)) continue;
            const double x[3] = {(double)qpc[q * 3], (double)qpc[q * 3 + 1], (double)qpc[q * 3 + 2]};
            double rb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double ub = ubc[q * NV + v];
                rb[v] = weno ? ub : slope[v] * ub - (q == qs[v] ? cq[v] : 0.0);
            }
            double xb[3] = {0.0, 0.0, 0.0};
            double pw3[3][4];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                pw3[d][0] = 1.0;
#pragma unroll
                for (int e = 1; e < 4; ++e) pw3[d][e] = pw3[d][e - 1] * x[d];
            }
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                double pb = 0.0;
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    pb += rb[v] * (double)(weno ? sGh[lc][k * NV + v] : sG0[lc][k * NV + v]);
                mba[k] -= pb;
                const int e0 = c_mono[k][0], e1 = c_mono[k][1], e2 = c_mono[k][2];
                if (e0 > 0) xb[0] += pb * e0 * pw3[0][e0 - 1] * pw3[1][e1] * pw3[2][e2];
                if (e1 > 0) xb[1] += pb * e1 * pw3[0][e0] * pw3[1][e1 - 1] * pw3[2][e2];
                if (e2 > 0) xb[2] += pb * e2 * pw3[0][e0] * pw3[1][e1] * pw3[2][e2 - 1];
            }
            double *gq = p.gQ + ((size_t)cell * p.nf * p.nqm + q) * 3;
#pragma unroll
            for (int d = 0; d < 3; ++d) gq[d] += xb[d];
        }
#pragma unroll
        for (int k = 0; k < IDEG; ++k) {
            const double t = grp_sum8(mba[k]);
            if (s == 0 && cvalid) p.gM[(size_t)cell * IDEG + k] += t;
        }
        // ---- indicators: beta_s,v = sum_kl B_kl g_k g_l over the full square
        // of the packed upper triangle ----
        if (p.iweno == 1) {
            const bool on = weno && act;
#pragma unroll
            for (int k = 0; k < IDEG; ++k)
#pragma unroll
                for (int l = k; l < IDEG; ++l) {
                    double c = 0.0;
                    if (on) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) c += bb[v] * (double)g[k][v] * (double)g[l][v];
                    }
                    const double t = grp_sum8(k == l ? c : 2.0 * c);
                    if (s == 0 && cvalid) p.gB[((size_t)cell * IDEG + k) * IDEG + l] += t;
                }
        }
    }

    // ---- G: dU_bar_j = P_j^T g_bar to the members, minus the same to U_i ----
    double acc[NV] = {0, 0, 0, 0, 0};
    if (scatter) {
        // member j+1's row and index are in flight while member j is scattered
        int kst;
        int n = p.lsq.member(tile, cell, s, 0, pos);
        const R *row = p.lsq.row(tile, cell, s, 0, pos, kst);
        R pc[IDEG];
#pragma unroll
        for (int k = 0; k < IDEG; ++k) pc[k] = row[(size_t)k * kst];
#pragma unroll 1
        for (int j = 0; j < IMAX; ++j) {
            R pn[IDEG];
            int nn = 0;
            if (j + 1 < IMAX) {
                nn = p.lsq.member(tile, cell, s, j + 1, pos);
                const R *rn = p.lsq.row(tile, cell, s, j + 1, pos, kst);
#pragma unroll
                for (int k = 0; k < IDEG; ++k) pn[k] = rn[(size_t)k * kst];
            }
            double d[NV] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                const double pk = (double)pc[k];
#pragma unroll
                for (int v = 0; v < NV; ++v) d[v] += pk * gb[k][v];
            }
            if constexpr (GEO) {
                // G_s = P_s dU: P_bar(j, k) = sum_v g_bar(k, v) dU_j(v)
                double du[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) du[v] = p.U[(size_t)n * NV + v] - ui[v];
                double *gp = p.gP + (((size_t)cell * p.lsq.ts + s) * IMAX + j) * IDEG;
#pragma unroll
                for (int k = 0; k < IDEG; ++k) {
                    double t = 0.0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) t += gb[k][v] * du[v];
                    gp[k] += t;
                }
            }
            const unsigned sl = p.slot ? p.slot[((size_t)tile * IMAX + j) * p.lsq.tp + pos] : UCNS_NOSLOT;
            if (sl < UCNS_TILE_UQ) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    __hip_atomic_fetch_add(&sAcc[sl][v], d[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) atomic_add_f64(&p.ubar[(size_t)n * NV + v], d[v]);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] -= d[v];
            n = nn;
#pragma unroll
            for (int k = 0; k < IDEG; ++k) pc[k] = pn[k];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double t = grp_sum8(acc[v]);
        if (s == 0 && cvalid) atomic_add_f64(&p.ubar[(size_t)cell * NV + v], t + usum[v] + ubi[v]);
    }
    if (nuq > 0) {
        __syncthreads();
        double *tt = p.tab + (size_t)p.toff[tile] * NV;
        for (int e = lane; e < nuq * NV; e += WAVE) tt[e] = e < UCNS_TILE_UQ * NV ? (&sAcc[0][0])[e] : 0.0;
    }
}

// Reconstruction product on the row-stream layouts (TS stencils per cell,
// TP = 8 TS columns per tile row).  The forward kept each stencil's
// coefficients G_s, indicators and weights ([tile][ideg*5 + 10][TP], gload)
// and each cell's flags (fload: 1 MUSCL, 2 first order), so nothing of the
// forward is repeated: the flags pick one of three paths per cell, and g_bar
// of every stencil goes to p.gbout ([tile][ideg*5][TP]) for the row-stream
// scatter.  On the WENO path every step is per variable (indicators and
// weights are), so a lane holds one variable's column of G_s at a time; B G_s
// runs over the packed triangle of B once per variable.
template <typename R, int IDEG, int NQL, int TS>
__global__ __launch_bounds__(WAVE) void k_recon_vjp_s(VjpParams<R> p) {
    constexpr int NP = IDEG * (IDEG + 1) / 2;
    constexpr int KL = (IDEG + 3) / 4;  // coefficients per lane in the face-point pass
    constexpr int QU = 3;               // face points in flight per lane (and group)
    constexpr int TP = 8 * TS;
    constexpr unsigned KS = NV * TP;    // stride of one coefficient row (elements)
    __shared__ double sGb[8][IDEG * NV];  // cotangent of the combined coefficients
    __shared__ R sB[8][NP];               // the tile's indicator matrices
    const int lane = threadIdx.x, lc = lane >> 3, s = lane & 7;
    const int b = blockIdx.x;
    const int tile = (b % NXCD) * p.tiles_per_xcd + b / NXCD;
    if (tile >= p.ntile) return;
    {
        const int nvc = min(8, p.ncell - tile * 8);
        const R *Bt = p.B + (size_t)tile * 8 * NP;
        for (int e = lane; e < nvc * NP; e += WAVE) (&sB[0][0])[e] = Bt[e];
    }
    const int cell = tile * 8 + lc;
    const bool cvalid = cell < p.ncell;
    const int cs = cvalid ? cell : 0;
    const int info = cvalid ? p.cinfo[cell] : 0;
    const int fl = cvalid ? p.fload[cell] : 0;
    const int ifca = ci_ifca(info), admis = ci_admis(info);
    const bool act = cvalid && s < admis;
    const int nslot = cvalid ? ifca * p.nqm : 0;
    const R *qpc = p.qp + (size_t)cs * p.nf * p.nqm * 3;
    const R *mb = p.mbar + (size_t)cs * IDEG;
    const double *ubc = p.ubf + (size_t)cs * p.nf * p.nqm * NV;
    auto qvalid = [&](int q) {
        const int f = q / p.nqm, qq = q - f * p.nqm;
        return q < nslot && qq < (ci_tri(info, f) ? p.qpt : p.qpq);
    };

    // ---- cotangent of the combined coefficients, sum_q phi_q ubar_q, and of
    // the cell average through a_q = U_i + ULEFT(q).  The 8 lanes of a cell
    // are two groups of 4: group g takes the face points of parity g, lane i
    // of a group the coefficients i, i+4, i+8, ...; the groups' sums meet by
    // one exchange ----
    double usum[NV] = {0, 0, 0, 0, 0};
    {
        const int g = s >> 2, il = s & 3;
        double ab[KL][NV];
        R mk[KL];
        int ex[KL][3];
#pragma unroll
        for (int t = 0; t < KL; ++t) {
            const int k = il + 4 * t;
            mk[t] = k < IDEG ? mb[k] : R(0);
#pragma unroll
            for (int d = 0; d < 3; ++d) ex[t][d] = k < IDEG ? c_mono[k][d] : 0;
#pragma unroll
            for (int v = 0; v < NV; ++v) ab[t][v] = 0.0;
        }
#pragma unroll 1
        for (int q0 = 0; q0 < nslot; q0 += 2 * QU) {
            // QU points' loads are issued before any is used (indices stay in
            // the cell's block: q < nslot <= nf nqm)
            R xq[QU][3];
            double ub[QU][NV];
#pragma unroll
            for (int u = 0; u < QU; ++u) {
                const int q = min(q0 + 2 * u + g, nslot - 1);
#pragma unroll
                for (int d = 0; d < 3; ++d) xq[u][d] = qpc[q * 3 + d];
#pragma unroll
                for (int v = 0; v < NV; ++v) ub[u][v] = ubc[q * NV + v];
            }
#pragma unroll
            for (int u = 0; u < QU; ++u) {
                const int q = q0 + 2 * u + g;
                const bool ok = q < nslot && qvalid(q);
                const R x0 = xq[u][0], x1 = xq[u][1], x2 = xq[u][2];
                const R x0s = x0 * x0, x1s = x1 * x1, x2s = x2 * x2;
                const R x0c = x0s * x0, x1c = x1s * x1, x2c = x2s * x2;
                double uv[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    uv[v] = ok ? ub[u][v] : 0.0;
                    usum[v] += uv[v];
                }
#pragma unroll
                for (int t = 0; t < KL; ++t) {
                    const double ph = (double)(pw(x0, x0s, x0c, ex[t][0]) * pw(x1, x1s, x1c, ex[t][1]) *
                                                   pw(x2, x2s, x2c, ex[t][2]) - mk[t]);
#pragma unroll
                    for (int v = 0; v < NV; ++v) ab[t][v] += ph * uv[v];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            usum[v] += __shfl_xor(usum[v], 4, WAVE);
#pragma unroll
            for (int t = 0; t < KL; ++t) ab[t][v] += __shfl_xor(ab[t][v], 4, WAVE);
        }
        if (g == 0) {
#pragma unroll
            for (int t = 0; t < KL; ++t) {
                const int k = il + 4 * t;
                if (k < IDEG) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) sGb[lc][k * NV + v] = ab[t][v];
                }
            }
        }
    }
    __syncthreads();  // sGb, sB

    const bool reduce = cvalid && (fl & 1), first = (fl & 2) != 0;
    // the tile's blocks (uniform) and this lane's column in them; lanes past
    // the cell's stencils read stencil 0's column and write nothing
    const R *gt = p.gload + (size_t)tile * (IDEG + 2) * KS;
    double *gbt = p.gbout + (size_t)tile * IDEG * KS;
    const unsigned c0 = lc * TS, col = c0 + (s < TS ? s : 0);
    const bool wr = cvalid && s < TS;
    // a MUSCL cell has no use for its indicator matrix: its LDS row takes the
    // central stencil's coefficients
    static_assert(IDEG * NV <= NP, "central coefficients exceed the indicator row");
    R *g0 = &sB[lc][0];
    if (reduce && !first)
        for (int e = s; e < IDEG * NV; e += 8) g0[e] = gt[c0 + (unsigned)e * TP];
    __syncthreads();  // g0
    double ubi[NV] = {0, 0, 0, 0, 0};  // extra cotangent of U_i from the limiter
    if (cvalid && !reduce) {
        // WENO: omega_s = lam_s (eps+beta_s)^-4 / sum_t (...), so with
        // obar_s = gh_bar . G_s:  beta_bar_s = -4 omega_s/(eps+beta_s)
        // (obar_s - sum_t omega_t obar_t),  g_bar_s = omega_s gh_bar + 2 beta_bar_s B G_s
#pragma unroll 1
        for (int v = 0; v < NV; ++v) {
            // B is re-read from LDS per variable: kept in registers across the
            // loop it would take NP of them
            asm volatile("" ::: "memory");
            const unsigned cv = col + (unsigned)v * TP;
            R gv[IDEG];
#pragma unroll
            for (int k = 0; k < IDEG; ++k) gv[k] = act ? gt[cv + k * KS] : R(0);
            const double beta = act ? (double)gt[cv + IDEG * KS] : 0.0;
            const double om = act ? (double)gt[cv + IDEG * KS + NV * TP] : 0.0;
            double obar = 0.0;
#pragma unroll
            for (int k = 0; k < IDEG; ++k) obar += sGb[lc][k * NV + v] * (double)gv[k];
            const double so = grp_sum8(act ? om * obar : 0.0);
            const double bb = act ? -4.0 * om / ((double)p.eps + beta) * (obar - so) : 0.0;
            double hg[IDEG];
#pragma unroll
            for (int k = 0; k < IDEG; ++k) hg[k] = (double)sB[lc][pidx(IDEG, k, k)] * (double)gv[k];
            // one row of B in registers at a time: each row's products are
            // pinned before the next row is read (else the compiler defers
            // them to the stores and keeps all of B live)
#pragma unroll
            for (int k = 0; k < IDEG; ++k) {
                asm volatile("" ::: "memory");
#pragma unroll
                for (int l = k + 1; l < IDEG; ++l) {
                    const double bkl = (double)sB[lc][pidx(IDEG, k, l)];
                    hg[k] += bkl * (double)gv[l];
                    hg[l] += bkl * (double)gv[k];
                }
#pragma unroll
                for (int l = k; l < IDEG; ++l) asm volatile("" : "+v"(hg[l]));
            }
            if (wr) {
#pragma unroll
                for (int k = 0; k < IDEG; ++k)
                    gbt[cv + k * KS] = act ? om * sGb[lc][k * NV + v] + 2.0 * bb * hg[k] : 0.0;
            }
        }
    } else if (reduce && !first) {
        // MUSCL + Barth with the forward's arithmetic, for the limiter's
        // active arguments: its arg-min face point and bound neighbour
        double ui[NV], umin[NV], umax[NV];
        int imin[NV], imax_[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            ui[v] = p.U[(size_t)cell * NV + v];
            umin[v] = umax[v] = ui[v];
            imin[v] = imax_[v] = cell;
        }
        for (int f = 0; f < ifca; ++f) {
            const int n = p.cnb[(size_t)cell * p.nf + f];
            if (n < 0) continue;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double un = p.U[(size_t)n * NV + v];
                if (un < umin[v]) imin[v] = n;
                if (un > umax[v]) imax_[v] = n;
                umin[v] = fmin(umin[v], un);
                umax[v] = fmax(umax[v], un);
            }
        }
        double lmin[NV];
        int lq[NV], lw[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lmin[v] = 1.0e16;
            lq[v] = -1;
            lw[v] = 0;
        }
        R res0[NQL][NV];
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
#pragma unroll
            for (int v = 0; v < NV; ++v) res0[r][v] = R(0);
            if (qvalid(q)) {
                R x[3] = {qpc[q * 3], qpc[q * 3 + 1], qpc[q * 3 + 2]};
                R phi[IDEG];
                basis_at<R, IDEG>(x, mb, phi);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    R res = R(0);
#pragma unroll
                    for (int k = 0; k < IDEG; ++k) res += phi[k] * g0[k * NV + v];
                    res0[r][v] = res;
                    const double d2 = (double)res;
                    double psi;
                    int wsel = 0;
                    if (fabs(d2) <= 0.0)
                        psi = 1.0;
                    else if (d2 > 0.0)
                        psi = fmin(1.0, (umax[v] - ui[v]) / d2);
                    else {
                        psi = fmin(1.0, (umin[v] - ui[v]) / d2);
                        wsel = 1;
                    }
                    if (psi < lmin[v]) {
                        lmin[v] = psi;
                        lq[v] = q;
                        lw[v] = wsel;
                    }
                }
            }
        }
        // ULEFT(q) = slope * d2_q: slope_bar = sum_q ubar_q d2_q; through an
        // active limiter slope = (u_ext - u_i) / d2_{q*}
        double sb[NV] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < NQL; ++r) {
            const int q = s + 8 * r;
            if (qvalid(q)) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sb[v] += ubc[q * NV + v] * (double)res0[r][v];
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double slope = grp_min8(lmin[v]);
            // lowest lane of the group holding the minimum
            int cand = (lmin[v] == slope) ? s : 8;
            cand = min(cand, __shfl_xor(cand, 1, WAVE));
            cand = min(cand, __shfl_xor(cand, 2, WAVE));
            cand = min(cand, __shfl_xor(cand, 4, WAVE));
            const int src = lc * 8 + cand;
            const int qs = lane_bcast(lq[v], src);
            const int wsel = lane_bcast(lw[v], src);
            const int ext = wsel ? imin[v] : imax_[v];
            // the unlimited value at the arg-min point
            double mine = 0.0;
#pragma unroll
            for (int r = 0; r < NQL; ++r)
                if (s + 8 * r == qs) mine = (double)res0[r][v];
            const double d2s = grp_sum8(mine);
            const double sbv = grp_sum8(sb[v]);
            const bool on = slope < 1.0 && qs >= 0;
            const double e = on ? sbv / d2s : 0.0;
            const double cq = on ? e * slope : 0.0;
            ubi[v] = -e;
            if (s == 0 && on) atomic_add_f64(&p.ubar[(size_t)ext * NV + v], e);
            if (wr) {
                // only the central stencil (lane 0) carries a coefficient cotangent
                R phi[IDEG];
                const bool corr = s == 0 && cq != 0.0;
                if (corr) {
                    R x[3] = {qpc[qs * 3], qpc[qs * 3 + 1], qpc[qs * 3 + 2]};
                    basis_at<R, IDEG>(x, mb, phi);
                }
                const unsigned cv = col + (unsigned)v * TP;
#pragma unroll
                for (int k = 0; k < IDEG; ++k)
                    gbt[cv + k * KS] = s == 0 ? slope * sGb[lc][k * NV + v] - (corr ? cq * (double)phi[k] : 0.0)
                                              : 0.0;
            }
        }
    } else if (wr) {
        // first-order cells: no coefficient cotangent
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int k = 0; k < IDEG; ++k) gbt[col + (unsigned)v * TP + k * KS] = 0.0;
    }
    if (s == 0 && cvalid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) atomic_add_f64(&p.ubar[(size_t)cell * NV + v], usum[v] + ubi[v]);
    }
}

// ubar[n] += the rows of every tile's member table that belong to cell n
// (CSR: rows ent[off[n]] .. ent[off[n+1]-1]); each cell is one thread's, so
// no atomics
__global__ __launch_bounds__(256) void k_adj_gather(int ncell, const int64_t *off, const int32_t *ent,
                                                    const double *tab, double *ubar) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ncell) return;
    double a[NV] = {0, 0, 0, 0, 0};
    // GU rows in flight: their indices, then the rows, are loaded before any
    // is added (the sum keeps the entries' order)
    constexpr int GU = 4;
    const int64_t e0 = off[n], e1 = off[n + 1];
    for (int64_t e = e0; e < e1; e += GU) {
        int32_t rr[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u) rr[u] = ent[min(e + u, e1 - 1)];
        double x[GU][NV];
#pragma unroll
        for (int u = 0; u < GU; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v) x[u][v] = tab[(size_t)rr[u] * NV + v];
#pragma unroll
        for (int u = 0; u < GU; ++u)
            if (e + u < e1) {
#pragma unroll
                for (int v = 0; v < NV; ++v) a[v] += x[u][v];
            }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) ubar[(size_t)n * NV + v] += a[v];
}

// CSR of the member tables by cell: counts, then positions
__global__ void k_adj_csr(int ntile, int cap, const int32_t *uq, const int32_t *ucount, const int64_t *toff,
                          int64_t *cnt, const int64_t *off, int32_t *ent) {
    const int t = blockIdx.x;
    if (t >= ntile) return;
    const int nu = min(ucount[t], cap);
    for (int r = threadIdx.x; r < nu; r += blockDim.x) {
        const int n = uq[(size_t)t * cap + r];
        const unsigned long long k = atomicAdd((unsigned long long *)&cnt[n], 1ull);
        if (off) ent[off[n] + (int64_t)k] = (int32_t)(toff[t] + r);
    }
}

// Face-parallel adjoint of the HLLC flux quadrature: HLLC by its reverse
// mode (hllc_vjp); the ghost state's Jacobian (5 inputs, plus the 10
// free-stream values) from forward-mode passes of the same templated code the
// primal runs, one pass per input; F_bar is pulled back through ROTATEB
// (whose transpose is ROTATEF) and ROTATEF (transpose ROTATEB).
// BND: the launch over boundary faces (ghost-state Jacobians by forward
// mode); the other launch takes the interior and periodic faces, whose
// kernel then carries none of the boundary code's registers
template <typename R, bool BND>
__global__ __launch_bounds__(256) void k_flux_vjp(FluxParams<R> p, const double *lam, double *ubf, double *bcbar) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= p.nface) return;
    const int cL = p.fcell[2 * f], cR = p.fcell[2 * f + 1];
    if ((cR < 0) != BND) return;
    const int s0 = p.fslot[2 * f], sR = p.fslot[2 * f + 1];
    const int sL = s0 & 255;
    const bool tri = (s0 >> 8) & 1;
    const int nq = tri ? p.qpt : p.qpq;
    const double *w = tri ? p.wt : p.wq;
    const double c1 = p.frot[4 * f], s1 = p.frot[4 * f + 1], c2 = p.frot[4 * f + 2], s2 = p.frot[4 * f + 3];
    const double area = p.farea[f];
    constexpr bool bnd = BND;
    double uL[NV], uR[NV], Fb[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        uL[v] = p.U[(size_t)cL * NV + v];
        uR[v] = bnd ? 0.0 : p.U[(size_t)cR * NV + v];
        Fb[v] = lam[(size_t)cL * NV + v] - (bnd ? 0.0 : lam[(size_t)cR * NV + v]);
    }
    double hb0[NV];
    rotf(Fb, c1, s1, c2, s2, hb0);
    const R *ulL = p.uleft + ((size_t)cL * p.nf + sL) * p.nqm * NV;
    const R *ulR = p.uleft + ((size_t)(bnd ? cL : cR) * p.nf + (bnd ? sL : sR)) * p.nqm * NV;
    double *obL = ubf + ((size_t)cL * p.nf + sL) * p.nqm * NV;
    double *obR = ubf + ((size_t)(bnd ? cL : cR) * p.nf + (bnd ? sL : sR)) * p.nqm * NV;
    double bcacc[2 * NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // QB points' deviations are loaded before any is used, as in k_flux
    constexpr int QB = 3;
    R dl[QB][NV], dr[QB][NV];
    for (int q = 0; q < nq; ++q) {
        const int qb = q % QB;
        if (qb == 0) {
#pragma unroll
            for (int u = 0; u < QB; ++u) {
                const int qq = min(q + u, nq - 1);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    dl[u][v] = ulL[qq * NV + v];
                    dr[u][v] = ulR[qq * NV + v];
                }
            }
        }
        double a[NV], bq[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R xl = dl[0][v], xr = dr[0][v];
#pragma unroll
            for (int u = 1; u < QB; ++u)
                if (qb == u) {
                    xl = dl[u][v];
                    xr = dr[u][v];
                }
            a[v] = uL[v] + (double)xl;
            bq[v] = uR[v] + (double)xr;
        }
        if (bnd) ghost_state(-cR, p.bc.boundtype, p.gamma, a, c1, s1, c2, s2, p.bc.inflow, p.bc.outflow, bq);
        double la[NV], ra[NV];
        rotf(a, c1, s1, c2, s2, la);
        rotf(bq, c1, s1, c2, s2, ra);
        const double wa = w[q] * area;
        double hb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) hb[v] = wa * hb0[v];
        double lab[NV], rab[NV];
        hllc_vjp(p.gamma, la, ra, hb, lab, rab);
        double ab[NV], bb[NV];
        rotb(lab, c1, s1, c2, s2, ab);
        rotb(rab, c1, s1, c2, s2, bb);
        if (bnd) {
            const int code = -cR;
#pragma unroll 1
            for (int m = 0; m < NV; ++m) {
                Dual A[NV], G[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) A[v] = Dual{a[v], m == v ? 1.0 : 0.0};
                ghost_state(code, p.bc.boundtype, p.gamma, A, c1, s1, c2, s2, p.bc.inflow, p.bc.outflow, G);
                double d = 0.0;
#pragma unroll
                for (int v = 0; v < NV; ++v) d += bb[v] * G[v].d;
                ab[m] += d;
            }
            if (bcbar && (code == 1 || code == 2 || code == 6)) {
#pragma unroll 1
                for (int m = 0; m < 2 * NV; ++m) {
                    Dual A[NV], G[NV], fi[NV], fo[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        A[v] = Dual{a[v], 0.0};
                        fi[v] = Dual{p.bc.inflow[v], m == v ? 1.0 : 0.0};
                        fo[v] = Dual{p.bc.outflow[v], m == NV + v ? 1.0 : 0.0};
                    }
                    ghost_state(code, p.bc.boundtype, p.gamma, A, c1, s1, c2, s2, fi, fo, G);
                    double d = 0.0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) d += bb[v] * G[v].d;
                    bcacc[m] += d;
                }
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) obR[q * NV + v] = bb[v];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) obL[q * NV + v] = ab[v];
    }
    if (bcbar && bnd) {
#pragma unroll
        for (int m = 0; m < 2 * NV; ++m)
            if (bcacc[m] != 0.0) atomic_add_f64(&bcbar[m], bcacc[m]);
    }
}

// Geometry cotangents of one residual product lam . R(U) on the face side
// (FP64): per face the area (garea), the unit outward normal (gnrm) and the
// cells' 1/TOTVOLUME (govol, through the update term dt R / V that pairs
// with lam).  The face sum S = area sum_q w_q <rotf(Fb), H_q> is linear in
// the area.  It depends on the rotation only through the normal (HLLC and
// the ghost states are invariant under turning the tangents), so its
// normal derivative is taken along two rotations of the whole frame: about
// t1 (n -> n + phi t2, t2 -> t2 - phi n) and about t2 (n -> n + psi t1,
// t1 -> t1 - psi n), each with a dual number; gnrm = area (dS/dphi t2 +
// dS/dpsi t1), the tangential gradient, free of the angles' singularity at
// n = +-z.  Every output is accumulated (+=) over the calls.
template <bool BND>
__global__ __launch_bounds__(256) void k_flux_geom(FluxParams<double> p, const double *lam, const double *ovol,
                                                  double *garea, double *gnrm, double *govol) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= p.nface) return;
    const int cL = p.fcell[2 * f], cR = p.fcell[2 * f + 1];
    if ((cR < 0) != BND) return;
    const int s0 = p.fslot[2 * f], sR = p.fslot[2 * f + 1];
    const int sL = s0 & 255;
    const bool tri = (s0 >> 8) & 1;
    const int nq = tri ? p.qpt : p.qpq;
    const double *w = tri ? p.wt : p.wq;
    const double c1 = p.frot[4 * f], s1 = p.frot[4 * f + 1], c2 = p.frot[4 * f + 2], s2 = p.frot[4 * f + 3];
    const double area = p.farea[f];
    const double nv[3] = {c1 * s2, s1 * s2, c2}, t1[3] = {c1 * c2, s1 * c2, -s2}, t2[3] = {-s1, c1, 0.0};
    double uL[NV], uR[NV], Fb[NV], fbr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        uL[v] = p.U[(size_t)cL * NV + v];
        uR[v] = BND ? 0.0 : p.U[(size_t)cR * NV + v];
        Fb[v] = lam[(size_t)cL * NV + v] - (BND ? 0.0 : lam[(size_t)cR * NV + v]);
    }
    rotf(Fb, c1, s1, c2, s2, fbr);
    const double *ulL = p.uleft + ((size_t)cL * p.nf + sL) * p.nqm * NV;
    const double *ulR = p.uleft + ((size_t)(BND ? cL : cR) * p.nf + (BND ? sL : sR)) * p.nqm * NV;
    double S = 0.0, dphi = 0.0, dpsi = 0.0, F[NV] = {0, 0, 0, 0, 0};
    for (int q = 0; q < nq; ++q) {
        double a[NV], b[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            a[v] = uL[v] + ulL[q * NV + v];
            b[v] = uR[v] + ulR[q * NV + v];
        }
        if (BND) ghost_state(-cR, p.bc.boundtype, p.gamma, a, c1, s1, c2, s2, p.bc.inflow, p.bc.outflow, b);
        double la[NV], ra[NV], h[NV], hb[NV];
        rotf(a, c1, s1, c2, s2, la);
        rotf(b, c1, s1, c2, s2, ra);
        hllc64(p.gamma, la, ra, h);
        rotb(h, c1, s1, c2, s2, hb);
        double sq = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sq += fbr[v] * h[v];
            F[v] += w[q] * hb[v];
        }
        S += w[q] * sq;
#pragma unroll 1
        for (int seed = 0; seed < 2; ++seed) {
            // seed 0: about t1 (n moves along t2); seed 1: about t2 (n moves along t1)
            VectorFrame<Dual> fr;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                fr.n[d] = Dual{nv[d], seed == 0 ? t2[d] : t1[d]};
                fr.t1[d] = Dual{t1[d], seed == 0 ? 0.0 : -nv[d]};
                fr.t2[d] = Dual{t2[d], seed == 0 ? -nv[d] : 0.0};
            }
            Dual A[NV], B[NV], FB[NV], LA[NV], RA[NV], H[NV], FR[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                A[v] = Dual{a[v], 0.0};
                B[v] = Dual{b[v], 0.0};
                FB[v] = Dual{Fb[v], 0.0};
            }
            if (BND) ghost_state_fr(-cR, p.bc.boundtype, p.gamma, A, fr, p.bc.inflow, p.bc.outflow, B);
            fr.fwd(A, LA);
            fr.fwd(B, RA);
            hllc_t<Dual>(p.gamma, LA, RA, H);
            fr.fwd(FB, FR);
            double d = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) d += (FR[v] * H[v]).d;
            (seed == 0 ? dphi : dpsi) += w[q] * d;
        }
    }
    garea[f] += S;
#pragma unroll
    for (int d = 0; d < 3; ++d) gnrm[(size_t)f * 3 + d] += area * (dphi * t2[d] + dpsi * t1[d]);
    // R_L holds +area F, R_R -area F
    double eL = 0.0, eR = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        eL += lam[(size_t)cL * NV + v] * F[v];
        if (!BND) eR += lam[(size_t)cR * NV + v] * F[v];
    }
    atomic_add_f64(&govol[cL], area * eL / ovol[cL]);
    if (!BND) atomic_add_f64(&govol[cR], -area * eR / ovol[cR]);
}

// the state [ncell][5] as 64-byte records [ncell][8] (the last three zero)
__global__ __launch_bounds__(256) void k_pack64(size_t n64, const double *U, double *U64) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n64) return;
    const size_t i = e / 8, v = e - i * 8;
    U64[e] = v < NV ? U[i * NV + v] : 0.0;
}

__global__ void k_zero_f64(double *p, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

// Reverse of one Shu-Osher stage (the three stage updates are affine in
// the stage state and linear in R):
//   stage 3: u = 1/3 u0 + 2/3 u2 - 2/3 dt R(u2)/V   -> u2_bar, u0_bar
//   stage 2: u2 = 3/4 u0 + 1/4 u1 - 1/4 dt R(u1)/V  -> u1_bar, u0_bar
//   stage 1: u1 = u0 - dt R(u0)/V                    -> u0_bar
// k_rk3_adj_w forms w = -c dt ybar / V (the cotangent handed to the residual
// product) for the stage whose output cotangent is ybar; k_rk3_adj_acc then
// adds a ybar and the product to the cotangents of the stage's inputs.
__global__ void k_rk3_adj_w(int ncell, const double *ybar, const double *ovol, const double *dt, double c, double *w) {
    const size_t n = (size_t)ncell * NV;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        w[e] = -(c * *dt) * ybar[e] * ovol[e / NV];
}
// the cotangents of one stage's inputs from its output cotangent y and the
// residual product jt = (dR/du)^T w:
//   stage 3: x = 2/3 y + jt (x = u2_bar), u0_bar = 1/3 y
//   stage 2: x = 1/4 y + jt (x = u1_bar), u0_bar += 3/4 y   (x may alias y)
//   stage 1: u0_bar += y + jt
__global__ void k_rk3_adj_stage(int ncell, int stage, const double *y, const double *jt, double *x, double *u0) {
    const size_t n = (size_t)ncell * NV;
    const double to4 = 3.0 / 4.0, oo4 = 1.0 / 4.0, to3 = 2.0 / 3.0, oo3 = 1.0 / 3.0;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const double ye = y[e], je = jt[e];
        if (stage == 3) {
            x[e] = to3 * ye + je;
            u0[e] = oo3 * ye;
        } else if (stage == 2) {
            u0[e] += to4 * ye;
            x[e] = oo4 * ye + je;
        } else {
            u0[e] += ye + je;
        }
    }
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_zero_i32(int *p, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = 0;
}

// Per-cell 5-vectors across the ABI: the caller's cell o = perm[c] (o = c
// without a permutation) in the caller's layout <-> the library's cell c in
// AoS.  src and dst never alias.
__global__ __launch_bounds__(256) void k_cells_in(int ncell, const int32_t *perm, const double *src, int layout,
                                                  double *dst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const size_t o = perm ? (size_t)perm[c] : (size_t)c;
#pragma unroll
    for (int v = 0; v < NV; ++v)
        dst[(size_t)c * NV + v] = layout == UCNS_LAYOUT_AOS ? src[o * NV + v] : src[(size_t)v * ncell + o];
}
__global__ __launch_bounds__(256) void k_cells_out(int ncell, const int32_t *perm, const double *src, int layout,
                                                   double *dst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const size_t o = perm ? (size_t)perm[c] : (size_t)c;
#pragma unroll
    for (int v = 0; v < NV; ++v)
        dst[layout == UCNS_LAYOUT_AOS ? o * NV + v : (size_t)v * ncell + o] = src[(size_t)c * NV + v];
}

__global__ void k_nonfinite(const double *u, size_t n, int *flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(u[i])) bad = 1;
    if (bad) atomicOr(flag, 1);
}

// [cells][ts_in][imax][ideg] (FP64, ABI) -> [cells][imax][ts_out][ideg] (R)
template <typename R>
__global__ void k_relayout_p(const double *src, R *dst, int ncells, int ts_in, int ts_out, int imax, int ideg) {
    const size_t per_out = (size_t)imax * ts_out * ideg;
    const size_t total = per_out * ncells;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / per_out;
        size_t r = e % per_out;
        const int k = (int)(r % ideg);
        r /= ideg;
        const int s = (int)(r % ts_out);
        const int j = (int)(r / ts_out);
        dst[e] = (R)src[((c * ts_in + s) * imax + j) * ideg + k];
    }
}

template <typename T>
int dalloc(ucns_ctx *ctx, T **p, size_t n);

// Scratch of ucns_residual_noalloc: its troubled-cell counters (the first
// 256 bytes) and face fluxes, in a file-scope device array that every GPU the
// library runs on holds from the library's first use on it until the process
// ends, so that the call allocates nothing.  In host code the array's name is
// a stand-in, not a device address: scratch_on() takes each device's own
// address once, with hipGetSymbolAddress, and only that address is used.
constexpr size_t UCNS_SCRATCH_BYTES = (size_t)256 << 20;  // 256 MiB: 6.7 M faces
constexpr size_t UCNS_SCRATCH_HEAD = 256;                  // counters

}  // namespace

static __device__ __align__(256) unsigned char scratch[UCNS_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[256] = {};

// the address of `scratch` on the current device (taken once per device)
int scratch_on(int device, void **out, std::string &err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= 256) {
        err = "device ordinal out of range for the scratch table";
        return UCNS_EUNSUPPORTED;
    }
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return UCNS_EHIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return UCNS_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
struct ucns_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    // sizes
    int ncell = 0, nf = 0, nqm = 0, nface = 0;
    int precision = UCNS_FP64, iweno = 1, ideg = 0, nnb = 0, imax = 0, ts_abi = 0, ts = 0;
    int qpq = 0, qpt = 0, nf_abi = 0, nqm_abi = 0, index_base = 0;
    double gamma = 1.4, lwci1 = 1.0, eps = 1e-11;
    bool have_prestore = false, have_state = false;
    bool tiled = false;  // WENO-4 lane-per-stencil kernel with the tiled layout
    int smode = 0;      // stencils per cell of the row-stream kernel in use (0: none)
    bool needs_bc = false, have_bc = false;  // boundary codes 1/2/6 present; ucns_set_boundary called
    double *bface = nullptr;  // [nface][10] per-face INFLOW(), OUTFLOW() states (ucns_set_boundary_faces)
    bool has_bnd_faces = false;              // any face with a ghost state (codes 1/2/3/4/6)
    ucns_boundary_v1 bc{};
    int ntile = 0;
    // mesh (device)
    int32_t *cinfo = nullptr, *cnb = nullptr, *cface = nullptr, *fcell = nullptr, *fslot = nullptr;
    double *frot = nullptr, *farea = nullptr, *ovol = nullptr, *minedge = nullptr;
    double *wquad = nullptr;  // [128] face weights: [0, 64) quadrilateral, [64, 128) triangle
    std::vector<int32_t> h_cinfo_base;  // ifca + face types (admis/full added at upload)
    // cell order: library cell c is the caller's cell h_perm[c] (both empty:
    // the caller's order); h_iperm is the inverse
    int cell_order = UCNS_ORDER_CALLER;
    std::vector<int32_t> h_perm, h_iperm;
    int32_t *perm = nullptr;    // device copy of h_perm, or null
    double *stage = nullptr;    // [ncell][5] staging buffer of host transfers
    void *scratch = nullptr;    // this device's address of the static scratch array
    // CWENO (ees 5): scheme switches, the caller's low-order data by library
    // cell (ucns_upload_cweno), and on the device the sectors' packed
    // indicators and the means' differences mbar - mbar_c
    int ees = 0, wenoz = 0, ideg2 = 0, imax2 = 0;
    // characteristic WENO (wenwrt 2): per cell face the neighbour or -(boundary
    // code), and the cell's own face angles' cosines and sines
    int wenwrt = 1;
    int32_t *chnb = nullptr;
    double *chang = nullptr;
    bool have_cweno = false;
    std::vector<double> h_cwP, h_cwB, h_cwIB;
    void *Bc = nullptr, *dmb = nullptr;
    void *Pc = nullptr;     // per-cell kernel: the sectors' low-order pseudo-inverses [ncell][ts-1][imax2][ideg2]
    bool cw_cell = false;   // CWENO on the per-cell kernel (WENO-5)
    // library options of the scheme
    int kernels = UCNS_KERNELS_AUTO;
    int adj_keep_stages = 0;
    double adj_keep_gib = 64.0;
    // prestore (device)
    int32_t *nbr = nullptr;
    void *P = nullptr, *B = nullptr, *mbar = nullptr, *qp = nullptr, *uleft = nullptr;
    unsigned char *rows = nullptr;  // row stream of the DMA reconstruction kernels
    double *u64 = nullptr;          // [ncell][8] the state in 64-byte records, for their gathers
    long long tile_bytes = 0;
    // state (device)
    double *U = nullptr, *U2 = nullptr, *fflux = nullptr, *rhs = nullptr, *t = nullptr, *dtfix = nullptr;
    unsigned long long *dtv = nullptr;  // [2] ping-pong dt slots
    int *counters = nullptr, *flag = nullptr;
    // adjoint (allocated on first use): face-point cotangents, state-sized
    // work arrays, the free-stream cotangent, stage checkpoints and step dts
    double *adj_ubf = nullptr, *adj_lam = nullptr, *adj_ubar = nullptr, *adj_x = nullptr, *adj_a = nullptr;
    double *adj_u0 = nullptr, *adj_bc = nullptr, *ckpt = nullptr, *dts = nullptr;
    void *adj_g = nullptr;  // stencil coefficients, indicators, weights of the row-stream reconstruction (R)
    double *adj_gb = nullptr;  // stencil-coefficient cotangents handed to the row stream, [ntile][ideg*5][8 ts]
    uint16_t *adj_slot = nullptr;                      // [ntile][imax][8 ts] LDS row of each member entry
    int32_t *adj_uq = nullptr, *adj_ucount = nullptr;  // [ntile][uq_cap], [ntile]
    int uq_cap = UCNS_TILE_UQ;                         // member-table rows per tile (tile_uq(ts, imax))
    int64_t *adj_toff = nullptr, *adj_off = nullptr;  // table rows per tile, CSR offsets by cell
    int32_t *adj_ent = nullptr;                       // CSR entries: table rows (fewer than 2^31)
    double *adj_tab = nullptr;                          // [rows][5] member tables
    bool adj_g_on = false;  // the next tiled reconstruction stores them
    int *adj_flags = nullptr;  // [ncell] its MUSCL / first-order flags
    bool adj_ready = false;  // every adjoint buffer and member table is in place
    int fail_alloc_in = 0;   // > 0: the fail_alloc_in-th device allocation from now fails (ucns_debug_fail_alloc)
    size_t ckpt_steps = 0;
    // the last stages of a march keep their forward reconstruction for the
    // reverse sweep: face values, and in row-stream mode the stencil
    // coefficients, indicators, weights (as adj_g) and flags
    std::vector<void *> keep_ul, keep_g;
    std::vector<int *> keep_f;
    // geometry cotangents of ucns_rk3_adjoint_geometry (device, FP64, live
    // during that call only): pseudo-inverses, indicators, basis means, face
    // points (k_recon_vjp), face areas, face normals, 1/TOTVOLUME (k_flux_geom)
    double *geo_P = nullptr, *geo_B = nullptr, *geo_M = nullptr, *geo_Q = nullptr;
    double *geo_area = nullptr, *geo_nrm = nullptr, *geo_ovol = nullptr;
    bool geo_on = false;
    // host copies for mapping them back to the caller's arrays: the faces'
    // owner cell and slot, 1/TOTVOLUME and VOLUME(1,1) by library cell
    std::vector<int32_t> h_fcell, h_fslot;
    std::vector<double> h_ovol, h_vol11;
    // stats
    bool timing = false;
    hipEvent_t ev[8] = {};
    ucns_stats_v1 st{};
};

namespace {

template <typename T>
int dalloc(ucns_ctx *ctx, T **p, size_t n) {
    if (ctx->fail_alloc_in > 0 && --ctx->fail_alloc_in == 0) {
        *p = nullptr;
        ctx->err = "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed (injected)";
        return UCNS_ENOMEM;
    }
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " bytes): " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? UCNS_ENOMEM : UCNS_EHIP;
    }
    return UCNS_OK;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int per_xcd(int ncell) { return (ncell + NXCD - 1) / NXCD; }

// kernel selection by (precision, ideg, imax)
template <typename R>
int launch_recon_t(ucns_ctx *ctx) {
    ReconParams<R> p{};
    p.ncell = ctx->ncell;
    p.nf = ctx->nf;
    p.nqm = ctx->nqm;
    p.ts = ctx->ts;
    p.per_xcd = per_xcd(ctx->ncell);
    p.iweno = ctx->iweno;
    p.qpq = ctx->qpq;
    p.qpt = ctx->qpt;
    p.gamma = ctx->gamma;
    p.lwci1 = ctx->lwci1;
    p.eps = ctx->eps;
    p.U = ctx->U;
    p.nbr = ctx->nbr;
    p.P = (const R *)ctx->P;
    p.B = (const R *)ctx->B;
    p.mbar = (const R *)ctx->mbar;
    p.qp = (const R *)ctx->qp;
    p.cinfo = ctx->cinfo;
    p.cnb = ctx->cnb;
    p.uleft = (R *)ctx->uleft;
    p.counters = ctx->counters;
    p.Pc = (const R *)ctx->Pc;
    p.Bc = (const R *)ctx->Bc;
    p.dmb = (const R *)ctx->dmb;
    p.wenoz = ctx->wenoz;
    p.chnb = ctx->chnb;
    p.chang = ctx->chang;
    p.bc = ctx->bc;
    p.cface = ctx->cface;
    p.bface = ctx->bface;
    if (ctx->tiled) {
        ReconWParams<R> w{};
        w.ncell = ctx->ncell;
        w.ntile = ctx->ntile;
        w.tiles_per_xcd = (ctx->ntile + NXCD - 1) / NXCD;
        w.ts = ctx->ts;
        w.nf = ctx->nf;
        w.nqm = ctx->nqm;
        w.qpq = ctx->qpq;
        w.qpt = ctx->qpt;
        w.gamma = ctx->gamma;
        w.lwci1 = ctx->lwci1;
        w.eps = ctx->eps;
        w.U = ctx->U;
        w.nbr = ctx->nbr;
        w.P = (const R *)ctx->P;
        w.B = (const R *)ctx->B;
        w.mbar = (const R *)ctx->mbar;
        w.qp = (const R *)ctx->qp;
        w.cinfo = ctx->cinfo;
        w.cnb = ctx->cnb;
        w.uleft = (R *)ctx->uleft;
        w.counters = ctx->counters;
        const dim3 grid(w.tiles_per_xcd * NXCD), block(WAVE);
        w.stream = ctx->rows;
        w.tile_bytes = ctx->tile_bytes;
        if (ctx->smode) {
            // the row-stream kernels gather the member states from 64-byte records
            const size_t n64 = (size_t)ctx->ncell * 8;
            hipLaunchKernelGGL(k_pack64, dim3((unsigned)((n64 + 255) / 256)), dim3(256), 0, ctx->stream, n64, ctx->U,
                               ctx->u64);
            w.U64 = ctx->u64;
        }
        w.gstore = ctx->adj_g_on ? (R *)ctx->adj_g : nullptr;
        w.fstore = ctx->adj_g_on ? ctx->adj_flags : nullptr;
        if (ctx->smode && ctx->adj_g_on) {
            if constexpr (sizeof(R) == 8) {
                if (ctx->ideg == 19)
                    if (ctx->smode == 5)
                        hipLaunchKernelGGL((k_recon_sg64<19, 39, 3, 5>), grid, block, 0, ctx->stream, w);
                    else
                        hipLaunchKernelGGL((k_recon_sg64<19, 39, 7, 7>), grid, block, 0, ctx->stream, w);
                else if (ctx->smode == 5)
                    hipLaunchKernelGGL((k_recon_sg64<9, 19, 3, 5>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_sg64<9, 19, 3, 7>), grid, block, 0, ctx->stream, w);
            } else {
                if (ctx->ideg == 19)
                    if (ctx->smode == 5)
                        hipLaunchKernelGGL((k_recon_sg32<19, 39, 3, 5>), grid, block, 0, ctx->stream, w);
                    else
                        hipLaunchKernelGGL((k_recon_sg32<19, 39, 7, 7>), grid, block, 0, ctx->stream, w);
                else if (ctx->smode == 5)
                    hipLaunchKernelGGL((k_recon_sg32<9, 19, 3, 5>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_sg32<9, 19, 3, 7>), grid, block, 0, ctx->stream, w);
            }
            return UCNS_OK;
        }
        if (ctx->smode) {
            // (ideg, imax, face-point slots, stencils): WENO-4 tets / hexes, WENO-3 tets / hexes
            if constexpr (sizeof(R) == 8) {
                if (ctx->ideg == 19)
                    if (ctx->smode == 5)
                        hipLaunchKernelGGL((k_recon_s64<19, 39, 3, 5>), grid, block, 0, ctx->stream, w);
                    else
                        hipLaunchKernelGGL((k_recon_s64<19, 39, 7, 7>), grid, block, 0, ctx->stream, w);
                else if (ctx->smode == 5)
                    hipLaunchKernelGGL((k_recon_s64<9, 19, 3, 5>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_s64<9, 19, 3, 7>), grid, block, 0, ctx->stream, w);
            } else {
                if (ctx->ideg == 19)
                    if (ctx->smode == 5)
                        hipLaunchKernelGGL((k_recon_s32<19, 39, 3, 5>), grid, block, 0, ctx->stream, w);
                    else
                        hipLaunchKernelGGL((k_recon_s32<19, 39, 7, 7>), grid, block, 0, ctx->stream, w);
                else if (ctx->smode == 5)
                    hipLaunchKernelGGL((k_recon_s32<9, 19, 3, 5>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_s32<9, 19, 3, 7>), grid, block, 0, ctx->stream, w);
            }
            return UCNS_OK;
        }
        if (ctx->ees == 5) {
            // CWENO on the register kernels
            w.Bc = (const R *)ctx->Bc;
            w.dmb = (const R *)ctx->dmb;
            w.wenoz = ctx->wenoz;
            if constexpr (sizeof(R) == 8) {
                if (ctx->ideg == 9)
                    hipLaunchKernelGGL((k_recon_cw64<9, 19, 3>), grid, block, 0, ctx->stream, w);
                else if (ctx->nf * ctx->nqm <= 24)
                    hipLaunchKernelGGL((k_recon_cw64<19, 39, 3>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_cw64<19, 39, 7>), grid, block, 0, ctx->stream, w);
            } else {
                if (ctx->ideg == 9)
                    hipLaunchKernelGGL((k_recon_cw32<9, 19, 3>), grid, block, 0, ctx->stream, w);
                else if (ctx->nf * ctx->nqm <= 24)
                    hipLaunchKernelGGL((k_recon_cw32<19, 39, 3>), grid, block, 0, ctx->stream, w);
                else
                    hipLaunchKernelGGL((k_recon_cw32<19, 39, 7>), grid, block, 0, ctx->stream, w);
            }
            return UCNS_OK;
        }
        if constexpr (sizeof(R) == 8) {
            if (ctx->ideg == 9)
                hipLaunchKernelGGL((k_recon_w64<9, 19, 3>), grid, block, 0, ctx->stream, w);
            else if (ctx->nf * ctx->nqm <= 24)
                hipLaunchKernelGGL((k_recon_w64<19, 39, 3>), grid, block, 0, ctx->stream, w);
            else
                hipLaunchKernelGGL((k_recon_w64<19, 39, 7>), grid, block, 0, ctx->stream, w);
        } else {
            if (ctx->ideg == 9)
                hipLaunchKernelGGL((k_recon_w32<9, 19, 3>), grid, block, 0, ctx->stream, w);
            else if (ctx->nf * ctx->nqm <= 24)
                hipLaunchKernelGGL((k_recon_w32<19, 39, 3>), grid, block, 0, ctx->stream, w);
            else
                hipLaunchKernelGGL((k_recon_w32<19, 39, 7>), grid, block, 0, ctx->stream, w);
        }
        return UCNS_OK;
    }
    const dim3 grid(p.per_xcd * NXCD), block(WAVE);
    if (ctx->wenwrt == 2) {
        if (ctx->ideg == 19 && ctx->imax == 39)
            hipLaunchKernelGGL((k_recon<R, 19, 39, 7, false, 9, 19, true>), grid, block, 0, ctx->stream, p);
        else if (ctx->ideg == 9 && ctx->imax == 19)
            hipLaunchKernelGGL((k_recon<R, 9, 19, 7, false, 9, 19, true>), grid, block, 0, ctx->stream, p);
        else if (ctx->ideg == 34 && ctx->imax == 69)
            hipLaunchKernelGGL((k_recon<R, 34, 69, 7, false, 9, 19, true>), grid, block, 0, ctx->stream, p);
        else {
            ctx->err = "no characteristic reconstruction kernel for idegfree=" + std::to_string(ctx->ideg);
            return UCNS_EUNSUPPORTED;
        }
        return UCNS_OK;
    }
    if (ctx->cw_cell) {
        if (ctx->ideg == 34 && ctx->imax == 69 && ctx->ideg2 == 9 && ctx->imax2 == 19) {
            hipLaunchKernelGGL((k_recon<R, 34, 69, 7, true, 9, 19>), grid, block, 0, ctx->stream, p);
            return UCNS_OK;
        }
        ctx->err = "no CWENO kernel for idegfree=" + std::to_string(ctx->ideg) + " idegfree2=" +
                   std::to_string(ctx->ideg2) + " imax2=" + std::to_string(ctx->imax2);
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->ideg == 19 && ctx->imax == 39)
        hipLaunchKernelGGL((k_recon<R, 19, 39, 7>), grid, block, 0, ctx->stream, p);
    else if (ctx->ideg == 34 && ctx->imax == 69)
        hipLaunchKernelGGL((k_recon<R, 34, 69, 7>), grid, block, 0, ctx->stream, p);
    else if (ctx->ideg == 9 && ctx->imax == 19)
        hipLaunchKernelGGL((k_recon<R, 9, 19, 7>), grid, block, 0, ctx->stream, p);
    else if (ctx->ideg == 3 && ctx->imax == 8)
        hipLaunchKernelGGL((k_recon<R, 3, 8, 1>), grid, block, 0, ctx->stream, p);
    else {
        ctx->err = "no reconstruction kernel for idegfree=" + std::to_string(ctx->ideg) +
                   " imax=" + std::to_string(ctx->imax);
        return UCNS_EUNSUPPORTED;
    }
    return UCNS_OK;
}

template <typename R>
void launch_flux_t(ucns_ctx *ctx) {
    FluxParams<R> p{};
    p.nface = ctx->nface;
    p.nf = ctx->nf;
    p.nqm = ctx->nqm;
    p.qpq = ctx->qpq;
    p.qpt = ctx->qpt;
    p.gamma = ctx->gamma;
    p.fcell = ctx->fcell;
    p.fslot = ctx->fslot;
    p.frot = ctx->frot;
    p.farea = ctx->farea;
    p.U = ctx->U;
    p.uleft = (const R *)ctx->uleft;
    p.fflux = ctx->fflux;
    p.wq = ctx->wquad;
    p.wt = ctx->wquad + 64;
    p.bc = ctx->bc;
    p.bface = ctx->bface;
    hipLaunchKernelGGL((k_flux<R>), dim3((ctx->nface + 255) / 256), dim3(256), 0, ctx->stream, p);
}

void tstart(ucns_ctx *ctx, int k) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev[2 * k], ctx->stream);
}
void tstop(ucns_ctx *ctx, int k) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev[2 * k + 1], ctx->stream);
}
void tcollect(ucns_ctx *ctx, int k, double *acc, int64_t *n) {
    if (!ctx->timing) return;
    float ms = 0;
    (void)hipEventSynchronize(ctx->ev[2 * k + 1]);
    if (hipEventElapsedTime(&ms, ctx->ev[2 * k], ctx->ev[2 * k + 1]) == hipSuccess) {
        *acc += ms;
        *n += 1;
    }
}

// reconstruction + face fluxes of the current state (ULEFT and face fluxes on device)
int residual_kernels(ucns_ctx *ctx) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 2);
    tstart(ctx, 0);
    int rc = ctx->precision == UCNS_MIXED ? launch_recon_t<float>(ctx) : launch_recon_t<double>(ctx);
    if (rc) return rc;
    tstop(ctx, 0);
    tcollect(ctx, 0, &ctx->st.ms_recon, &ctx->st.n_recon);
    tstart(ctx, 1);
    if (ctx->precision == UCNS_MIXED)
        launch_flux_t<float>(ctx);
    else
        launch_flux_t<double>(ctx);
    tstop(ctx, 1);
    tcollect(ctx, 1, &ctx->st.ms_flux, &ctx->st.n_flux);
    return UCNS_OK;
}

int update_kernel(ucns_ctx *ctx, int stage, const double *dt, unsigned long long *dt_next, double cfl,
                  double *rhs_out) {
    UpdateParams p{};
    p.ncell = ctx->ncell;
    p.nf = ctx->nf;
    p.stage = stage;
    p.cfl = cfl;
    p.gamma = ctx->gamma;
    p.cface = ctx->cface;
    p.fflux = ctx->fflux;
    p.ovol = ctx->ovol;
    p.minedge = ctx->minedge;
    p.U = ctx->U;
    p.U2 = ctx->U2;
    p.rhs_out = rhs_out;
    p.dt = dt;
    p.dt_next = dt_next;
    p.t = ctx->t;
    tstart(ctx, 2);
    hipLaunchKernelGGL(k_update, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, p);
    tstop(ctx, 2);
    tcollect(ctx, 2, &ctx->st.ms_update, &ctx->st.n_update);
    return UCNS_OK;
}

int dt_kernel(ucns_ctx *ctx, double cfl, unsigned long long *out) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, ctx->stream, out,
                       (unsigned long long)0x7FF0000000000000ULL);
    DtParams p{};
    p.ncell = ctx->ncell;
    p.cfl = cfl;
    p.gamma = ctx->gamma;
    p.U = ctx->U;
    p.minedge = ctx->minedge;
    p.dt_out = out;
    tstart(ctx, 3);
    hipLaunchKernelGGL(k_dt, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, p);
    tstop(ctx, 3);
    tcollect(ctx, 3, &ctx->st.ms_dt, &ctx->st.n_dt);
    return UCNS_OK;
}

// ---- adjoint -----------------------------------------------------------
void release_adjoint(ucns_ctx *ctx) {
    dfree(ctx->adj_ubf);
    dfree(ctx->adj_lam);
    dfree(ctx->adj_ubar);
    dfree(ctx->adj_x);
    dfree(ctx->adj_a);
    dfree(ctx->adj_u0);
    dfree(ctx->adj_bc);
    dfree(ctx->ckpt);
    dfree(ctx->dts);
    if (ctx->adj_g) (void)hipFree(ctx->adj_g);
    ctx->adj_g = nullptr;
    dfree(ctx->adj_flags);
    dfree(ctx->adj_gb);
    dfree(ctx->adj_slot);
    dfree(ctx->adj_uq);
    dfree(ctx->adj_ucount);
    dfree(ctx->adj_toff);
    dfree(ctx->adj_off);
    dfree(ctx->adj_ent);
    dfree(ctx->adj_tab);
    ctx->ckpt_steps = 0;
    ctx->adj_ready = false;
    for (void *q : ctx->keep_ul) (void)hipFree(q);
    for (void *q : ctx->keep_g) (void)hipFree(q);
    for (int *q : ctx->keep_f) (void)hipFree(q);
    ctx->keep_ul.clear();
    ctx->keep_g.clear();
    ctx->keep_f.clear();
}

template <typename R>
LsqView<R> lsq_view(const ucns_ctx *ctx) {
    LsqView<R> L{};
    L.mode = ctx->smode ? 2 : ctx->tiled ? 1 : 0;
    L.ts = ctx->ts;
    L.tp = 8 * ctx->ts;
    L.imax = ctx->imax;
    L.ideg = ctx->ideg;
    L.depth = stream_depth(sizeof(R) == 4, ctx->ts);
    L.pb = ctx->ideg * L.tp * (int)sizeof(R);
    L.rd = L.pb + L.tp * 4;
    L.hd = L.depth * L.tp * 4;
    L.tile_bytes = ctx->tile_bytes;
    L.nbr = ctx->nbr;
    L.P = (const R *)ctx->P;
    L.rows = ctx->rows;
    return L;
}

template <typename R>
void launch_tile_members(ucns_ctx *ctx) {
    const LsqView<R> L = lsq_view<R>(ctx);
    const dim3 grid(ctx->ntile), block(WAVE);
    if (ctx->imax == 39)
        hipLaunchKernelGGL((k_tile_members<R, 39>), grid, block, 0, ctx->stream, L, ctx->ntile, ctx->ncell,
                           ctx->cinfo, ctx->uq_cap, ctx->adj_uq, ctx->adj_ucount, ctx->adj_slot);
    else if (ctx->imax == 19)
        hipLaunchKernelGGL((k_tile_members<R, 19>), grid, block, 0, ctx->stream, L, ctx->ntile, ctx->ncell,
                           ctx->cinfo, ctx->uq_cap, ctx->adj_uq, ctx->adj_ucount, ctx->adj_slot);
    else
        hipLaunchKernelGGL((k_tile_members<R, 8>), grid, block, 0, ctx->stream, L, ctx->ntile, ctx->ncell,
                           ctx->cinfo, ctx->uq_cap, ctx->adj_uq, ctx->adj_ucount, ctx->adj_slot);
}

// frees a temporary device buffer on every exit of the scope
template <typename T>
struct DevTemp {
    T *p = nullptr;
    ~DevTemp() { dfree(p); }
};

int ensure_adjoint_steps(ucns_ctx *ctx) {
    const size_t n = (size_t)ctx->ncell * NV;
    int rc;
    if ((rc = dalloc(ctx, &ctx->adj_ubf, (size_t)ctx->ncell * ctx->nf * ctx->nqm * NV))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_lam, n))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_ubar, n))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_x, n))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_a, n))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_u0, n))) return rc;
    if ((rc = dalloc(ctx, &ctx->adj_bc, 2 * NV))) return rc;
    if (ctx->smode) {
        const size_t ng = (size_t)ctx->ntile * (ctx->ideg + 2) * NV * 8 * ctx->ts;
        if ((rc = dalloc(ctx, &ctx->adj_flags, (size_t)ctx->ncell))) return rc;
        if (ctx->precision == UCNS_MIXED)
            rc = dalloc(ctx, (float **)&ctx->adj_g, ng);
        else
            rc = dalloc(ctx, (double **)&ctx->adj_g, ng);
        if (rc) return rc;
    }
    if (ctx->smode) {
        if ((rc = dalloc(ctx, &ctx->adj_gb, (size_t)ctx->ntile * ctx->ideg * NV * 8 * ctx->ts))) return rc;
    }
    // the tiles' distinct stencil members (once per upload)
    if (ctx->ts <= 8 && (ctx->imax == 39 || ctx->imax == 19 || ctx->imax == 8)) {
        if ((rc = dalloc(ctx, &ctx->adj_slot, (size_t)ctx->ntile * ctx->imax * 8 * ctx->ts))) return rc;
        ctx->uq_cap = tile_uq(ctx->ts, ctx->imax);
        if ((rc = dalloc(ctx, &ctx->adj_uq, (size_t)ctx->ntile * ctx->uq_cap))) return rc;
        if ((rc = dalloc(ctx, &ctx->adj_ucount, (size_t)ctx->ntile))) return rc;
        if (ctx->precision == UCNS_MIXED)
            launch_tile_members<float>(ctx);
        else
            launch_tile_members<double>(ctx);
        HIP_TRY(hipGetLastError());
        // table rows per tile (prefix sums on the host: set-up only).  The
        // library stream is non-blocking: every host read of what its kernels
        // wrote waits for the stream first, and every clear is queued on it.
        const int nt = ctx->ntile, nc = ctx->ncell;
        std::vector<int32_t> uc(nt);
        HIP_TRY(hipMemcpyAsync(uc.data(), ctx->adj_ucount, nt * sizeof(int32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<int64_t> toff(nt + 1, 0);
        for (int t = 0; t < nt; ++t) toff[t + 1] = toff[t] + std::min(uc[t], ctx->uq_cap);
        const int64_t rows = toff[nt];
        if (rows >= ((int64_t)1 << 31)) {
            ctx->err = "member tables exceed 2^31 rows";
            return UCNS_EUNSUPPORTED;
        }
        if ((rc = dalloc(ctx, &ctx->adj_toff, (size_t)nt + 1))) return rc;
        if ((rc = dalloc(ctx, &ctx->adj_tab, (size_t)rows * NV))) return rc;
        if ((rc = dalloc(ctx, &ctx->adj_off, (size_t)nc + 1))) return rc;
        if ((rc = dalloc(ctx, &ctx->adj_ent, (size_t)std::max<int64_t>(rows, 1)))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->adj_toff, toff.data(), (nt + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                               ctx->stream));
        // CSR by cell: count, scan, fill
        DevTemp<int64_t> cnt_buf;
        if ((rc = dalloc(ctx, &cnt_buf.p, (size_t)nc))) return rc;
        int64_t *cnt = cnt_buf.p;
        HIP_TRY(hipMemsetAsync(cnt, 0, nc * sizeof(int64_t), ctx->stream));
        hipLaunchKernelGGL(k_adj_csr, dim3(nt), dim3(256), 0, ctx->stream, nt, ctx->uq_cap, ctx->adj_uq,
                           ctx->adj_ucount, ctx->adj_toff, cnt, (const int64_t *)nullptr, (int32_t *)nullptr);
        std::vector<int64_t> hc(nc), off(nc + 1, 0);
        HIP_TRY(hipMemcpyAsync(hc.data(), cnt, nc * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < nc; ++i) off[i + 1] = off[i] + hc[i];
        if (off[nc] != rows) {  // every table row belongs to exactly one cell
            ctx->err = "member-table index: " + std::to_string(off[nc]) + " entries for " + std::to_string(rows) +
                       " rows";
            return UCNS_EHIP;
        }
        HIP_TRY(hipMemcpyAsync(ctx->adj_off, off.data(), (nc + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                               ctx->stream));
        HIP_TRY(hipMemsetAsync(cnt, 0, nc * sizeof(int64_t), ctx->stream));
        hipLaunchKernelGGL(k_adj_csr, dim3(nt), dim3(256), 0, ctx->stream, nt, ctx->uq_cap, ctx->adj_uq,
                           ctx->adj_ucount, ctx->adj_toff, cnt, (const int64_t *)ctx->adj_off, ctx->adj_ent);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return UCNS_OK;
}

// The adjoint's buffers and member tables, set up on first use.  A context
// counts as set up only once every step has succeeded; a failed step frees
// what the earlier ones allocated, so the next call starts over and returns
// the error again instead of running kernels on missing buffers.
int ensure_adjoint(ucns_ctx *ctx) {
    if (ctx->ees == 5) {
        ctx->err = "the adjoint does not cover CWENO (ees 5)";
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->ideg == 34) {
        ctx->err = "the adjoint covers idegfree 3, 9 and 19 (not WENO-5, idegfree 34)";
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->wenwrt == 2) {
        ctx->err = "the adjoint does not cover characteristic reconstruction (wenwrt 2)";
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->bface) {
        ctx->err = "the adjoint takes uniform boundary states (ucns_set_boundary), not per-face ones";
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->adj_ready) return UCNS_OK;
    release_adjoint(ctx);
    const int rc = ensure_adjoint_steps(ctx);
    if (rc != UCNS_OK) {
        release_adjoint(ctx);
        return rc;
    }
    ctx->adj_ready = true;
    return UCNS_OK;
}

// the member tables of the last scatter into ubar
void adj_gather(ucns_ctx *ctx, double *ubar) {
    hipLaunchKernelGGL(k_adj_gather, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, ctx->ncell,
                       ctx->adj_off, ctx->adj_ent, ctx->adj_tab, ubar);
}

void zero_dev(ucns_ctx *ctx, double *p, size_t n) {
    const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_zero_f64, dim3(std::max(blocks, 1)), dim3(256), 0, ctx->stream, p, n);
}

// P^T g_bar of the last products (adj_gb) onto the row stream, into ubar
template <typename R>
void launch_scatter_t(ucns_ctx *ctx, double *ubar) {
    ReconWParams<R> w{};
    w.ncell = ctx->ncell;
    w.ntile = ctx->ntile;
    w.tiles_per_xcd = (ctx->ntile + NXCD - 1) / NXCD;
    w.ts = ctx->ts;
    w.cinfo = ctx->cinfo;
    w.stream = ctx->rows;
    w.tile_bytes = ctx->tile_bytes;
    w.gbs = ctx->adj_gb;
    w.ubar = ubar;
    w.slot = ctx->adj_slot;
    w.uq_cap = ctx->uq_cap;
    w.ucount = ctx->adj_ucount;
    w.toff = ctx->adj_toff;
    w.tab = ctx->adj_tab;
    const dim3 grid(w.tiles_per_xcd * NXCD), block(WAVE);
    if constexpr (sizeof(R) == 8) {
        if (ctx->ideg == 19)
            if (ctx->smode == 5)
                hipLaunchKernelGGL((k_scatter_s64<19, 39, 5>), grid, block, 0, ctx->stream, w);
            else
                hipLaunchKernelGGL((k_scatter_s64<19, 39, 7>), grid, block, 0, ctx->stream, w);
        else if (ctx->smode == 5)
            hipLaunchKernelGGL((k_scatter_s64<9, 19, 5>), grid, block, 0, ctx->stream, w);
        else
            hipLaunchKernelGGL((k_scatter_s64<9, 19, 7>), grid, block, 0, ctx->stream, w);
    } else {
        if (ctx->ideg == 19)
            if (ctx->smode == 5)
                hipLaunchKernelGGL((k_scatter_s32<19, 39, 5>), grid, block, 0, ctx->stream, w);
            else
                hipLaunchKernelGGL((k_scatter_s32<19, 39, 7>), grid, block, 0, ctx->stream, w);
        else if (ctx->smode == 5)
            hipLaunchKernelGGL((k_scatter_s32<9, 19, 5>), grid, block, 0, ctx->stream, w);
        else
            hipLaunchKernelGGL((k_scatter_s32<9, 19, 7>), grid, block, 0, ctx->stream, w);
    }
}

// the forward reconstruction of the current state for a residual product:
// ULEFT, and in row-stream mode every stencil's coefficients, indicators,
// weights and the cell flags
template <typename R>
int vjp_recon_t(ucns_ctx *ctx) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 2);
    ctx->adj_g_on = ctx->smode && ctx->adj_g;
    tstart(ctx, 3);
    const int rc = launch_recon_t<R>(ctx);
    tstop(ctx, 3);
    tcollect(ctx, 3, &ctx->st.ms_recon_adj, &ctx->st.n_recon_adj);
    ctx->adj_g_on = false;
    return rc;
}

// the products of one residual at the state of the last vjp_recon_t: ubar
// (zeroed here) = the flux product, the reconstruction product and, in
// row-stream mode, the scatter of P^T g_bar (adj_gb) through the member tables
template <typename R>
int vjp_products_t(ucns_ctx *ctx, const double *lam, double *ubar, double *bcbar) {
    const int nfq = ctx->nf * ctx->nqm;
    if (ctx->ts > 8 || nfq > 56) {
        ctx->err = "the residual product needs at most 8 stencils and 56 face points per cell";
        return UCNS_EUNSUPPORTED;
    }
    zero_dev(ctx, ubar, (size_t)ctx->ncell * NV);
    FluxParams<R> fp{};
    fp.nface = ctx->nface;
    fp.nf = ctx->nf;
    fp.nqm = ctx->nqm;
    fp.qpq = ctx->qpq;
    fp.qpt = ctx->qpt;
    fp.gamma = ctx->gamma;
    fp.fcell = ctx->fcell;
    fp.fslot = ctx->fslot;
    fp.frot = ctx->frot;
    fp.farea = ctx->farea;
    fp.U = ctx->U;
    fp.uleft = (const R *)ctx->uleft;
    fp.fflux = nullptr;
    fp.wq = ctx->wquad;
    fp.wt = ctx->wquad + 64;
    fp.bc = ctx->bc;
    tstart(ctx, 1);
    hipLaunchKernelGGL((k_flux_vjp<R, false>), dim3((ctx->nface + 255) / 256), dim3(256), 0, ctx->stream, fp, lam,
                       ctx->adj_ubf, bcbar);
    if (ctx->has_bnd_faces)
        hipLaunchKernelGGL((k_flux_vjp<R, true>), dim3((ctx->nface + 255) / 256), dim3(256), 0, ctx->stream, fp, lam,
                       ctx->adj_ubf, bcbar);
    tstop(ctx, 1);
    tcollect(ctx, 1, &ctx->st.ms_flux_vjp, &ctx->st.n_flux_vjp);
    if constexpr (sizeof(R) == 8) {
        if (ctx->geo_on) {
            hipLaunchKernelGGL(k_flux_geom<false>, dim3((ctx->nface + 255) / 256), dim3(256), 0, ctx->stream, fp, lam,
                               ctx->ovol, ctx->geo_area, ctx->geo_nrm, ctx->geo_ovol);
            if (ctx->has_bnd_faces)
                hipLaunchKernelGGL(k_flux_geom<true>, dim3((ctx->nface + 255) / 256), dim3(256), 0, ctx->stream, fp,
                                   lam, ctx->ovol, ctx->geo_area, ctx->geo_nrm, ctx->geo_ovol);
        }
    }
    VjpParams<R> vp{};
    vp.ncell = ctx->ncell;
    vp.ntile = ctx->ntile;
    vp.tiles_per_xcd = (ctx->ntile + NXCD - 1) / NXCD;
    vp.nf = ctx->nf;
    vp.nqm = ctx->nqm;
    vp.qpq = ctx->qpq;
    vp.qpt = ctx->qpt;
    vp.iweno = ctx->iweno;
    vp.gamma = ctx->gamma;
    vp.lwci1 = ctx->lwci1;
    vp.eps = ctx->eps;
    vp.U = ctx->U;
    vp.lsq = lsq_view<R>(ctx);
    vp.B = (const R *)ctx->B;
    vp.mbar = (const R *)ctx->mbar;
    vp.qp = (const R *)ctx->qp;
    vp.cinfo = ctx->cinfo;
    vp.cnb = ctx->cnb;
    vp.ubf = ctx->adj_ubf;
    vp.ubar = ubar;
    vp.gload = (ctx->smode && ctx->adj_g) ? (const R *)ctx->adj_g : nullptr;
    vp.fload = ctx->adj_flags;
    vp.slot = ctx->adj_slot;
    vp.uq_cap = ctx->uq_cap;
    vp.uq = ctx->adj_uq;
    vp.ucount = ctx->adj_ucount;
    vp.gbout = ctx->smode ? ctx->adj_gb : nullptr;
    vp.toff = ctx->adj_toff;
    vp.tab = ctx->adj_tab;
    vp.gP = ctx->geo_P;
    vp.gB = ctx->geo_B;
    vp.gM = ctx->geo_M;
    vp.gQ = ctx->geo_Q;
    const dim3 grid(vp.tiles_per_xcd * NXCD), block(WAVE);
    const bool q3 = nfq <= 24;
    tstart(ctx, 0);
    bool geo_done = false;
    if constexpr (sizeof(R) == 8) {
        if (ctx->geo_on) {
            geo_done = true;
            if (ctx->ideg == 19 && ctx->imax == 39) {
                if (q3)
                    hipLaunchKernelGGL((k_recon_vjp<R, 19, 39, 3, true>), grid, block, 0, ctx->stream, vp);
                else
                    hipLaunchKernelGGL((k_recon_vjp<R, 19, 39, 7, true>), grid, block, 0, ctx->stream, vp);
            } else if (ctx->ideg == 9 && ctx->imax == 19) {
                if (q3)
                    hipLaunchKernelGGL((k_recon_vjp<R, 9, 19, 3, true>), grid, block, 0, ctx->stream, vp);
                else
                    hipLaunchKernelGGL((k_recon_vjp<R, 9, 19, 7, true>), grid, block, 0, ctx->stream, vp);
            } else if (ctx->ideg == 3 && ctx->imax == 8) {
                if (q3)
                    hipLaunchKernelGGL((k_recon_vjp<R, 3, 8, 3, true>), grid, block, 0, ctx->stream, vp);
                else
                    hipLaunchKernelGGL((k_recon_vjp<R, 3, 8, 7, true>), grid, block, 0, ctx->stream, vp);
            } else {
                ctx->err = "no geometry product for idegfree=" + std::to_string(ctx->ideg);
                return UCNS_EUNSUPPORTED;
            }
        }
    }
    if (geo_done) {
    } else if (ctx->smode) {
        // row-stream shapes: from the forward's stored coefficients; the
        // scatter follows on the row stream
        if (!vp.gload || !vp.fload) {
            ctx->err = "the row-stream product needs the forward's stored coefficients";
            return UCNS_EUNSUPPORTED;
        }
        if (ctx->ideg == 19 && ctx->smode == 5)
            hipLaunchKernelGGL((k_recon_vjp_s<R, 19, 3, 5>), grid, block, 0, ctx->stream, vp);
        else if (ctx->ideg == 19)
            hipLaunchKernelGGL((k_recon_vjp_s<R, 19, 7, 7>), grid, block, 0, ctx->stream, vp);
        else if (ctx->smode == 5)
            hipLaunchKernelGGL((k_recon_vjp_s<R, 9, 3, 5>), grid, block, 0, ctx->stream, vp);
        else
            hipLaunchKernelGGL((k_recon_vjp_s<R, 9, 3, 7>), grid, block, 0, ctx->stream, vp);
    } else if (ctx->ideg == 19 && ctx->imax == 39) {
        if (q3)
            hipLaunchKernelGGL((k_recon_vjp<R, 19, 39, 3>), grid, block, 0, ctx->stream, vp);
        else
            hipLaunchKernelGGL((k_recon_vjp<R, 19, 39, 7>), grid, block, 0, ctx->stream, vp);
    } else if (ctx->ideg == 9 && ctx->imax == 19) {
        if (q3)
            hipLaunchKernelGGL((k_recon_vjp<R, 9, 19, 3>), grid, block, 0, ctx->stream, vp);
        else
            hipLaunchKernelGGL((k_recon_vjp<R, 9, 19, 7>), grid, block, 0, ctx->stream, vp);
    } else if (ctx->ideg == 3 && ctx->imax == 8) {
        if (q3)
            hipLaunchKernelGGL((k_recon_vjp<R, 3, 8, 3>), grid, block, 0, ctx->stream, vp);
        else
            hipLaunchKernelGGL((k_recon_vjp<R, 3, 8, 7>), grid, block, 0, ctx->stream, vp);
    } else {
        ctx->err = "no residual product for idegfree=" + std::to_string(ctx->ideg);
        return UCNS_EUNSUPPORTED;
    }
    if (ctx->smode) launch_scatter_t<R>(ctx, ubar);
    if (ctx->adj_slot) adj_gather(ctx, ubar);
    tstop(ctx, 0);
    tcollect(ctx, 0, &ctx->st.ms_recon_vjp, &ctx->st.n_recon_vjp);
    return UCNS_OK;
}

// (dR/dU)^T lam at the current state into ubar; bcbar (device, 10) += the
// free-stream cotangent when non-null
template <typename R>
int residual_vjp_t(ucns_ctx *ctx, const double *lam, double *ubar, double *bcbar) {
    int rc;
    if ((rc = vjp_recon_t<R>(ctx))) return rc;
    return vjp_products_t<R>(ctx, lam, ubar, bcbar);
}

// (dR/dU)^T lam at the current state into ubar; bcbar (device, 10) += the
// free-stream cotangent when non-null
int residual_vjp_dev(ucns_ctx *ctx, const double *lam, double *ubar, double *bcbar) {
    int rc = ensure_adjoint(ctx);
    if (rc) return rc;
    return ctx->precision == UCNS_MIXED ? residual_vjp_t<float>(ctx, lam, ubar, bcbar)
                                        : residual_vjp_t<double>(ctx, lam, ubar, bcbar);
}

// How many of a march's last stages keep their forward reconstruction (at
// most nstage): the scheme's adj_keep_stages caps the count (0: no cap,
// < 0: none), adj_keep_gib the bytes (never more than the free device memory
// less 4 GiB).
// Buffers only grow; a failed allocation leaves the count at what was had.
int adj_keep(ucns_ctx *ctx, int nstage) {
    const size_t esz = ctx->precision == UCNS_MIXED ? 4 : 8;
    const size_t ul = (size_t)ctx->ncell * ctx->nf * ctx->nqm * NV * esz;
    const size_t g = ctx->smode ? (size_t)ctx->ntile * (ctx->ideg + 2) * NV * 8 * ctx->ts * esz : 0;
    const size_t fl = ctx->smode ? (size_t)ctx->ncell * sizeof(int) : 0;
    const size_t per = ul + g + fl;
    long want = nstage;
    if (ctx->adj_keep_stages < 0) return 0;
    if (ctx->adj_keep_stages > 0) want = std::min<long>(want, ctx->adj_keep_stages);
    size_t budget = (size_t)(ctx->adj_keep_gib * (double)(1ull << 30));
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        const size_t have = fr + ctx->keep_ul.size() * per, margin = 4ull << 30;
        budget = std::min(budget, have > margin ? have - margin : 0);
    }
    want = std::min<long>(want, (long)(budget / per));
    while ((long)ctx->keep_ul.size() < want) {
        void *a = nullptr, *b = nullptr;
        int *c = nullptr;
        if (hipMalloc(&a, ul) != hipSuccess) break;
        if (ctx->smode && (hipMalloc(&b, g) != hipSuccess || hipMalloc((void **)&c, fl) != hipSuccess)) {
            (void)hipFree(a);
            if (b) (void)hipFree(b);
            break;
        }
        ctx->keep_ul.push_back(a);
        ctx->keep_g.push_back(b);
        ctx->keep_f.push_back(c);
    }
    (void)hipGetLastError();  // a refused allocation only lowers the count
    return (int)std::min<long>(want, (long)ctx->keep_ul.size());
}

// Point the reconstruction's outputs (face values; stencil coefficients and
// flags in row-stream mode) at kept stage b, and back to the context's own
// buffers on destruction.
struct KeptStage {
    ucns_ctx *ctx;
    void *ul, *g;
    int *f;
    KeptStage(ucns_ctx *c, int b) : ctx(c), ul(c->uleft), g(c->adj_g), f(c->adj_flags) {
        ctx->uleft = ctx->keep_ul[b];
        if (ctx->smode) {
            ctx->adj_g = ctx->keep_g[b];
            ctx->adj_flags = ctx->keep_f[b];
        }
    }
    ~KeptStage() {
        ctx->uleft = ul;
        ctx->adj_g = g;
        ctx->adj_flags = f;
    }
};

// one SSP-RK3 step forward from ctx->U with the stage states stored; stages
// 3 it + stage - 1 >= keep0 also keep their reconstruction
int rk3_step_ckpt(ucns_ctx *ctx, int it, const double *dt, unsigned long long *next, double cfl, int keep0) {
    const size_t n = (size_t)ctx->ncell * NV;
    for (int stage = 1; stage <= 3; ++stage) {
        const int q = 3 * it + stage - 1;
        HIP_TRY(hipMemcpyAsync(ctx->ckpt + (size_t)q * n, ctx->U, n * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
        int rc;
        if (q >= keep0) {
            KeptStage k(ctx, q - keep0);
            ctx->adj_g_on = ctx->smode != 0;
            rc = residual_kernels(ctx);
            ctx->adj_g_on = false;
        } else {
            rc = residual_kernels(ctx);
        }
        if (rc) return rc;
        if (stage == 3 && next)
            hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, ctx->stream, next,
                               (unsigned long long)0x7FF0000000000000ULL);
        update_kernel(ctx, stage, dt, stage == 3 ? next : nullptr, cfl, nullptr);
    }
    return UCNS_OK;
}

// The reverse sweep over nsteps stored steps: abar holds the cotangent of
// the final state on entry and of the initial state on return.  Stages are
// reversed in the order (step n: 3, 2, 1), each at its stored state, with
//   w = -c dt ybar / V, jt = (dR/du)^T w, and k_rk3_adj_stage combining ybar
//   and jt into the cotangents of the stage's inputs.
// In row-stream mode the scatter of a stage's reconstruction product rides on
// the forward reconstruction of the next stage reversed, so each reversed
// stage streams the pseudo-inverses once.
//
// Stages q = 3 it + stage - 1 >= keep0 kept their forward reconstruction, so
// their products start from it and only the other stages run the
// reconstruction again.
template <typename R>
int rk3_reverse_t(ucns_ctx *ctx, int nsteps, double *abar, double *bcbar, int keep0) {
    const size_t n = (size_t)ctx->ncell * NV;
    const int blocks = (int)std::min<size_t>(8192, (n + 255) / 256);
    double *const ustate = ctx->U;
    double *a = abar, *u0 = ctx->adj_u0, *x = ctx->adj_x, *w = ctx->adj_lam, *jt = ctx->adj_ubar;
    const double cst[4] = {0.0, 1.0, 1.0 / 4.0, 2.0 / 3.0};  // -w / (dt ybar / V) per stage
    int rc = UCNS_OK;
    for (int it = nsteps - 1; it >= 0 && rc == UCNS_OK; --it) {
        const double *dt = ctx->dts + it;
        for (int stage = 3; stage >= 1 && rc == UCNS_OK; --stage) {
            ctx->U = ctx->ckpt + ((size_t)3 * it + stage - 1) * n;
            const double *y = stage == 3 ? a : x;
            hipLaunchKernelGGL(k_rk3_adj_w, dim3(blocks), dim3(256), 0, ctx->stream, ctx->ncell, y, ctx->ovol, dt,
                               cst[stage], w);
            const int q = 3 * it + stage - 1;
            if (q >= keep0) {
                KeptStage k(ctx, q - keep0);
                if ((rc = vjp_products_t<R>(ctx, w, jt, bcbar))) break;
            } else {
                if ((rc = vjp_recon_t<R>(ctx))) break;
                if ((rc = vjp_products_t<R>(ctx, w, jt, bcbar))) break;
            }
            // this stage's product is complete: fold it into its input cotangents
            hipLaunchKernelGGL(k_rk3_adj_stage, dim3(blocks), dim3(256), 0, ctx->stream, ctx->ncell, stage, y, jt, x,
                               u0);
            if (stage == 1) (void)hipMemcpyAsync(a, u0, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
        }
    }
    ctx->U = ustate;
    if (rc == UCNS_OK) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            ctx->err = std::string("rk3 reverse: ") + hipGetErrorString(e);
            rc = UCNS_EHIP;
        }
    }
    return rc;
}

int rk3_reverse(ucns_ctx *ctx, int nsteps, double *abar, double *bcbar, int keep0) {
    return ctx->precision == UCNS_MIXED ? rk3_reverse_t<float>(ctx, nsteps, abar, bcbar, keep0)
                                        : rk3_reverse_t<double>(ctx, nsteps, abar, bcbar, keep0);
}

template <typename R>
int upload_real(ucns_ctx *ctx, void **dst, const std::vector<double> &src) {
    std::vector<R> tmp(src.size());
    for (size_t e = 0; e < src.size(); ++e) tmp[e] = (R)src[e];
    int rc = dalloc(ctx, (R **)dst, tmp.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(R), hipMemcpyHostToDevice));
    return UCNS_OK;
}

int upload_real_any(ucns_ctx *ctx, void **dst, const std::vector<double> &src) {
    return ctx->precision == UCNS_MIXED ? upload_real<float>(ctx, dst, src) : upload_real<double>(ctx, dst, src);
}

void release_prestore(ucns_ctx *ctx) {
    dfree(ctx->nbr);
    dfree(ctx->P);
    dfree(ctx->rows);
    dfree(ctx->u64);
    dfree(ctx->B);
    dfree(ctx->mbar);
    if (ctx->Bc) (void)hipFree(ctx->Bc);
    if (ctx->dmb) (void)hipFree(ctx->dmb);
    if (ctx->Pc) (void)hipFree(ctx->Pc);
    ctx->Bc = ctx->dmb = ctx->Pc = nullptr;
    dfree(ctx->qp);
    ctx->have_prestore = false;
}


// ---- cell order -------------------------------------------------------------
// The library may hold the cells in an order of its own, so that a tile of 8
// consecutive cells is a compact cluster whatever order the caller's
// partitioner produced.  perm[c] = the caller's cell of library cell c.

// reverse Cuthill-McKee on the face graph (interior and periodic faces):
// per connected component, breadth-first from a pseudo-peripheral cell of
// least degree, neighbours in increasing degree, then the whole order reversed
void order_rcm(const ucns_mesh_v2 *m, std::vector<int32_t> &perm) {
    const int nc = m->ncell, nfa = m->nface_max, base = m->index_base;
    std::vector<int32_t> deg(nc, 0);
    auto nbr = [&](int i, int f) -> int {
        if (f >= m->ifca[i]) return -1;
        const int bc = m->bcode[(size_t)i * nfa + f];
        if (bc != 0 && bc != 5 && bc != 50) return -1;
        const int j = m->ineigh[(size_t)i * nfa + f] - base;
        return j == i ? -1 : j;
    };
    for (int i = 0; i < nc; ++i)
        for (int f = 0; f < nfa; ++f) deg[i] += nbr(i, f) >= 0;
    std::vector<int32_t> level(nc, -1), q;
    q.reserve(nc);
    // breadth-first levels from r over the cells not yet ordered; returns the
    // cell of least degree in the last level and the number of levels
    std::vector<char> done(nc, 0);
    std::vector<int32_t> touched;
    auto bfs_far = [&](int r, int &nlev) {
        touched.clear();
        std::vector<int32_t> frontier{r}, next;
        level[r] = 0;
        touched.push_back(r);
        nlev = 1;
        int far = r;
        while (!frontier.empty()) {
            next.clear();
            for (int i : frontier)
                for (int f = 0; f < nfa; ++f) {
                    const int j = nbr(i, f);
                    if (j >= 0 && !done[j] && level[j] < 0) {
                        level[j] = level[i] + 1;
                        touched.push_back(j);
                        next.push_back(j);
                    }
                }
            if (next.empty()) {
                far = frontier[0];
                for (int i : frontier)
                    if (deg[i] < deg[far] || (deg[i] == deg[far] && i < far)) far = i;
            } else {
                ++nlev;
            }
            frontier.swap(next);
        }
        for (int i : touched) level[i] = -1;
        return far;
    };
    perm.clear();
    perm.reserve(nc);
    std::vector<int32_t> by_deg(nc);
    for (int i = 0; i < nc; ++i) by_deg[i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int32_t> cand;
    for (int seed : by_deg) {
        if (done[seed]) continue;
        int r = seed, nlev = 0;
        int far = bfs_far(r, nlev);
        for (int it = 0; it < 8; ++it) {  // pseudo-peripheral cell (George and Liu)
            int nl2 = 0;
            const int far2 = bfs_far(far, nl2);
            if (nl2 <= nlev) break;
            r = far;
            nlev = nl2;
            far = far2;
        }
        size_t head = perm.size();
        perm.push_back(r);
        done[r] = 1;
        while (head < perm.size()) {
            const int i = perm[head++];
            cand.clear();
            for (int f = 0; f < nfa; ++f) {
                const int j = nbr(i, f);
                if (j >= 0 && !done[j]) {
                    done[j] = 1;
                    cand.push_back(j);
                }
            }
            std::sort(cand.begin(), cand.end(),
                      [&](int a, int b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); });
            perm.insert(perm.end(), cand.begin(), cand.end());
        }
    }
    std::reverse(perm.begin(), perm.end());
}

// Morton (Z-order) curve of the centroids: 21 bits per axis over the
// bounding box, ties by the caller's index
void order_morton(const ucns_mesh_v2 *m, std::vector<int32_t> &perm) {
    const int nc = m->ncell;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < nc; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], m->centroid[(size_t)3 * i + d]);
            hi[d] = std::max(hi[d], m->centroid[(size_t)3 * i + d]);
        }
    auto spread = [](uint64_t x) {  // 21 bits -> every third bit of 63
        x &= 0x1FFFFFull;
        x = (x | x << 32) & 0x1F00000000FFFFull;
        x = (x | x << 16) & 0x1F0000FF0000FFull;
        x = (x | x << 8) & 0x100F00F00F00F00Full;
        x = (x | x << 4) & 0x10C30C30C30C30C3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    std::vector<uint64_t> key(nc);
    for (int i = 0; i < nc; ++i) {
        uint64_t k = 0;
        for (int d = 0; d < 3; ++d) {
            const double ext = hi[d] - lo[d];
            const double t = ext > 0 ? (m->centroid[(size_t)3 * i + d] - lo[d]) / ext : 0.0;
            const uint64_t g = (uint64_t)std::min(2097151.0, std::max(0.0, std::floor(t * 2097152.0)));
            k |= spread(g) << (2 - d);
        }
        key[i] = k;
    }
    perm.resize(nc);
    for (int i = 0; i < nc; ++i) perm[i] = i;
    std::sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
}

// the caller's cell of library cell i, and the library cell of caller cell o
inline size_t src_cell(const ucns_ctx *ctx, int i) { return ctx->h_perm.empty() ? (size_t)i : (size_t)ctx->h_perm[i]; }
inline int lib_cell(const ucns_ctx *ctx, int o) { return ctx->h_iperm.empty() ? o : ctx->h_iperm[o]; }

// caller-order 5-vectors (host or device memory, either layout) -> the
// library's AoS device array dst
int cells_in(ucns_ctx *ctx, const double *src, bool host, int32_t layout, double *dst) {
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) {
        ctx->err = "unknown layout";
        return UCNS_EBADARG;
    }
    const size_t n = (size_t)ctx->ncell * NV;
    if (host) {
        HIP_TRY(hipMemcpyAsync(ctx->stage, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        src = ctx->stage;
    }
    hipLaunchKernelGGL(k_cells_in, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, ctx->ncell, ctx->perm,
                       src, layout, dst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UCNS_OK;
}

// the library's AoS device array src -> caller-order 5-vectors
int cells_out(ucns_ctx *ctx, const double *src, int32_t layout, double *dst, bool host) {
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) {
        ctx->err = "unknown layout";
        return UCNS_EBADARG;
    }
    const size_t n = (size_t)ctx->ncell * NV;
    hipLaunchKernelGGL(k_cells_out, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, ctx->ncell, ctx->perm,
                       src, layout, host ? ctx->stage : dst);
    HIP_TRY(hipGetLastError());
    if (host) HIP_TRY(hipMemcpyAsync(dst, ctx->stage, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UCNS_OK;
}

// The caller's mesh with its cells in the library's order (per-cell rows
// moved, neighbour ids renumbered); the vectors own the storage the struct
// points into.
struct PermutedMesh {
    ucns_mesh_v2 m{};
    std::vector<int32_t> ifca, ineigh, ineighn, bcode, types;
    std::vector<double> angx, angy, surf, vol, minedge;
    void build(const ucns_mesh_v2 *src, const std::vector<int32_t> &perm, const std::vector<int32_t> &iperm) {
        const int nc = src->ncell, nfa = src->nface_max, base = src->index_base;
        const size_t nfc = (size_t)nc * nfa;
        ifca.resize(nc);
        vol.resize(nc);
        minedge.resize(nc);
        ineigh.resize(nfc);
        ineighn.resize(nfc);
        bcode.resize(nfc);
        types.resize(nfc);
        angx.resize(nfc);
        angy.resize(nfc);
        surf.resize(nfc);
        for (int c = 0; c < nc; ++c) {
            const size_t o = perm[c];
            ifca[c] = src->ifca[o];
            vol[c] = src->totvolume[o];
            minedge[c] = src->minedge[o];
            for (int f = 0; f < nfa; ++f) {
                const size_t d = (size_t)c * nfa + f, sl = o * nfa + f;
                const int bc = src->bcode[sl];
                const int j = src->ineigh[sl] - base;
                const bool link = f < src->ifca[o] && (bc == 0 || bc == 5 || bc == 50) && j >= 0 && j < nc;
                ineigh[d] = link ? iperm[j] + base : src->ineigh[sl];
                ineighn[d] = src->ineighn[sl];
                bcode[d] = bc;
                types[d] = src->types_faces[sl];
                angx[d] = src->faceanglex[sl];
                angy[d] = src->faceangley[sl];
                surf[d] = src->surf[sl];
            }
        }
        m = *src;
        m.ifca = ifca.data();
        m.ineigh = ineigh.data();
        m.ineighn = ineighn.data();
        m.bcode = bcode.data();
        m.types_faces = types.data();
        m.faceanglex = angx.data();
        m.faceangley = angy.data();
        m.surf = surf.data();
        m.totvolume = vol.data();
        m.minedge = minedge.data();
        m.centroid = nullptr;
    }
};

}  // namespace

extern "C" {

int ucns_abi_version(void) { return UCNS_ABI_VERSION; }

const char *ucns_last_error(const ucns_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ucns_create(const ucns_mesh_v2 *m, const ucns_scheme_v2 *s, ucns_ctx **out) {
    if (!out) return UCNS_EBADARG;
    *out = nullptr;
    if (!m || !s) return UCNS_EBADARG;
    PermutedMesh pm;
    ucns_ctx *ctx = new ucns_ctx();
    *out = ctx;
    auto bad = [&](int code, const std::string &msg) {
        ctx->err = msg;
        return code;
    };
    if (m->abi_version != UCNS_ABI_VERSION || s->abi_version != UCNS_ABI_VERSION)
        return bad(UCNS_EBADARG, "abi_version mismatch");
    if (m->ncell <= 0 || m->nface_max <= 0 || m->nface_max > 6) return bad(UCNS_EBADARG, "bad ncell/nface_max");
    if (!m->ifca || !m->ineigh || !m->ineighn || !m->bcode || !m->types_faces || !m->faceanglex ||
        !m->faceangley || !m->surf || !m->totvolume || !m->minedge)
        return bad(UCNS_EBADARG, "null mesh array");
    if (s->precision != UCNS_FP64 && s->precision != UCNS_MIXED) return bad(UCNS_EBADARG, "bad precision");
    if (s->iweno != 1 && s->iweno != -1) return bad(UCNS_EUNSUPPORTED, "iweno must be 1 (WENO) or -1 (MUSCL)");
    if (s->ees != 0 && s->ees != 5) return bad(UCNS_EUNSUPPORTED, "ees must be 0 (WENO) or 5 (CWENO)");
    if (s->ees == 5 && s->iweno != 1) return bad(UCNS_EBADARG, "CWENO (ees 5) needs iweno 1");
    if (s->wenoz != 0 && s->wenoz != 1) return bad(UCNS_EBADARG, "wenoz must be 0 or 1");
    if (s->wenoz == 1 && s->ees != 5) return bad(UCNS_EUNSUPPORTED, "wenoz 1 is a CWENO option (ees 5)");
    if (s->wenwrt != 1 && s->wenwrt != 2)
        return bad(UCNS_EUNSUPPORTED, "wenwrt must be 1 (conserved) or 2 (characteristic variables)");
    if (s->wenwrt == 2 && (s->iweno != 1 || s->ees != 0))
        return bad(UCNS_EUNSUPPORTED, "characteristic reconstruction (wenwrt 2) is built for WENO with ees 0");
    if (s->limiter != 1) return bad(UCNS_EUNSUPPORTED, "only the Barth-Jespersen limiter (1)");
    if (s->riemann != 1) return bad(UCNS_EUNSUPPORTED, "only the HLLC solver (1)");
    if (s->typesten < 1 || s->typesten > 7) return bad(UCNS_EUNSUPPORTED, "typesten must be 1..7");
    if (s->qp_quad < 1 || s->qp_quad > 64 || s->qp_triangle < 1 || s->qp_triangle > 64 || !s->weights_quad ||
        !s->weights_tri)
        return bad(UCNS_EBADARG, "bad face quadrature");
    if (s->cell_order != UCNS_ORDER_CALLER && s->cell_order != UCNS_ORDER_RCM && s->cell_order != UCNS_ORDER_MORTON)
        return bad(UCNS_EBADARG, "cell_order must be UCNS_ORDER_CALLER, _RCM or _MORTON");
    if (s->cell_order == UCNS_ORDER_MORTON && !m->centroid)
        return bad(UCNS_EBADARG, "UCNS_ORDER_MORTON needs the mesh centroids");
    if (s->kernels != UCNS_KERNELS_AUTO && s->kernels != UCNS_KERNELS_REGISTER)
        return bad(UCNS_EBADARG, "kernels must be UCNS_KERNELS_AUTO or UCNS_KERNELS_REGISTER");
    if (s->reserved != 0) return bad(UCNS_EBADARG, "reserved must be 0");
    if (!std::isfinite(s->adj_keep_gib)) return bad(UCNS_EBADARG, "adj_keep_gib must be finite");
    ctx->cell_order = s->cell_order;
    ctx->kernels = s->kernels;
    ctx->adj_keep_stages = s->adj_keep_stages;
    ctx->adj_keep_gib = s->adj_keep_gib > 0 ? s->adj_keep_gib : 64.0;
    ctx->ncell = m->ncell;
    ctx->nf_abi = m->nface_max;
    ctx->index_base = m->index_base;
    ctx->precision = s->precision;
    ctx->iweno = s->iweno;
    ctx->ees = s->ees;
    ctx->wenwrt = s->wenwrt;
    ctx->wenoz = s->wenoz;
    ctx->ideg = s->idegfree;
    ctx->nnb = s->inumneighbours;
    ctx->imax = s->inumneighbours - 1;
    ctx->ts_abi = s->typesten;
    ctx->qpq = s->qp_quad;
    ctx->qpt = s->qp_triangle;
    ctx->nqm_abi = std::max(s->qp_quad, s->qp_triangle);
    ctx->gamma = s->gamma;
    ctx->lwci1 = s->lwci1;
    ctx->eps = s->weno_eps;
    if (!((ctx->ideg == 34 && ctx->imax == 69) || (ctx->ideg == 19 && ctx->imax == 39) ||
          (ctx->ideg == 9 && ctx->imax == 19) || (ctx->ideg == 3 && ctx->imax == 8)))
        return bad(UCNS_EUNSUPPORTED, "supported (idegfree, inumneighbours): (34, 70) WENO-5, (19, 40) WENO-4, "
                                      "(9, 20) WENO-3 and (3, 9) MUSCL");
    if (ctx->iweno == -1 && ctx->ideg != 3) return bad(UCNS_EUNSUPPORTED, "MUSCL path is linear (idegfree 3)");
    if (m->ncell > 0 && m->ifca)
        for (int i = 0; i < m->ncell; ++i)
            if (m->ifca[i] < 4 || m->ifca[i] > m->nface_max)
                return bad(UCNS_EBADARG, "ifca out of range at cell " + std::to_string(i));
    // the caller's face links, checked before anything reads through them:
    // codes, neighbour ranges and the symmetry of (cell, slot) pairs
    {
        const int nc0 = m->ncell, nfa0 = m->nface_max, b0 = m->index_base;
        for (int i = 0; i < nc0; ++i)
            for (int f = 0; f < m->ifca[i]; ++f) {
                const size_t sl = (size_t)i * nfa0 + f;
                const int bc = m->bcode[sl];
                if (bc == 1 || bc == 2 || bc == 3 || bc == 4 || bc == 6) continue;
                if (bc != 0 && bc != 5 && bc != 50)
                    return bad(UCNS_EUNSUPPORTED, "face codes: 0 interior, 5/50 periodic, 1/2/3/4/6 boundary");
                const int j = m->ineigh[sl] - b0, fj = m->ineighn[sl] - b0;
                if (j < 0 || j >= nc0 || fj < 0 || fj >= m->ifca[j])
                    return bad(UCNS_EBADARG, "face neighbour out of range at cell " + std::to_string(i));
                const size_t slj = (size_t)j * nfa0 + fj;
                if (m->ineigh[slj] - b0 != i || m->ineighn[slj] - b0 != f)
                    return bad(UCNS_EBADARG, "face neighbours are not symmetric at cell " + std::to_string(i));
            }
    }
    if (ctx->cell_order != UCNS_ORDER_CALLER) {
        if (ctx->cell_order == UCNS_ORDER_RCM)
            order_rcm(m, ctx->h_perm);
        else
            order_morton(m, ctx->h_perm);
        ctx->h_iperm.assign(m->ncell, 0);
        for (int c = 0; c < m->ncell; ++c) ctx->h_iperm[ctx->h_perm[c]] = c;
        pm.build(m, ctx->h_perm, ctx->h_iperm);
        m = &pm.m;
    }

    const int nc = m->ncell, nfa = m->nface_max, base = m->index_base;
    // faces in use, Gauss points per face type present
    int nf = 0, nqm = 0;
    bool any_quad = false, any_tri = false;
    for (int i = 0; i < nc; ++i) {
        const int ifca = m->ifca[i];
        if (ifca < 4 || ifca > nfa) return bad(UCNS_EBADARG, "ifca out of range at cell " + std::to_string(i));
        nf = std::max(nf, ifca);
        for (int f = 0; f < ifca; ++f) {
            const int t = m->types_faces[(size_t)i * nfa + f];
            if (t == 5) any_quad = true;
            else if (t == 6) any_tri = true;
            else return bad(UCNS_EBADARG, "types_faces must be 5 or 6");
        }
    }
    if (any_quad) nqm = std::max(nqm, s->qp_quad);
    if (any_tri) nqm = std::max(nqm, s->qp_triangle);
    if (nf > 6) return bad(UCNS_EUNSUPPORTED, "at most 6 faces per cell");
    ctx->nf = nf;
    ctx->nqm = nqm;
    if (nf * nqm > 6 * 16) return bad(UCNS_EUNSUPPORTED, "too many face Gauss points per cell");
    // unique faces: owner = lower cell id (equal ids: lower slot)
    std::vector<int32_t> cinfo(nc, 0), cnb((size_t)nc * nf, -1), cface((size_t)nc * nf, 0);
    std::vector<int32_t> chnb(ctx->wenwrt == 2 ? (size_t)nc * nf : 0, 0);
    std::vector<double> chang(ctx->wenwrt == 2 ? (size_t)nc * nf * 4 : 0, 0.0);
    std::vector<int32_t> fcell, fslot;
    std::vector<double> frot, farea;
    for (int i = 0; i < nc; ++i) {
        const int ifca = m->ifca[i];
        int w = ifca;
        for (int f = 0; f < ifca; ++f)
            if (m->types_faces[(size_t)i * nfa + f] == 6) w |= 1 << (8 + f);
        cinfo[i] = w;
        for (int f = 0; f < ifca; ++f) {
            const size_t sl = (size_t)i * nfa + f;
            const int bc = m->bcode[sl];
            if (!chnb.empty()) {
                const bool bnd = bc == 1 || bc == 2 || bc == 3 || bc == 4 || bc == 6;
                chnb[(size_t)i * nf + f] = bnd ? -bc : m->ineigh[sl] - base;
                const double a1 = m->faceanglex[sl], a2 = m->faceangley[sl];
                double *an = &chang[((size_t)i * nf + f) * 4];
                an[0] = std::cos(a1);
                an[1] = std::sin(a1);
                an[2] = std::cos(a2);
                an[3] = std::sin(a2);
            }
            if (bc == 1 || bc == 2 || bc == 3 || bc == 4 || bc == 6) {
                // boundary face: owner only, the right state is a ghost (BOUNDARYS)
                if (bc != 3 && bc != 4) ctx->needs_bc = true;
                ctx->has_bnd_faces = true;
                cnb[(size_t)i * nf + f] = -1;
                const int id = (int)farea.size();
                fcell.push_back(i);
                fcell.push_back(-bc);
                fslot.push_back(f | (m->types_faces[sl] == 6 ? 256 : 0));
                fslot.push_back(0);
                const double a1 = m->faceanglex[sl], a2 = m->faceangley[sl];
                frot.push_back(std::cos(a1));
                frot.push_back(std::sin(a1));
                frot.push_back(std::cos(a2));
                frot.push_back(std::sin(a2));
                farea.push_back(m->surf[sl]);
                cface[(size_t)i * nf + f] = id + 1;
                continue;
            }
            if (bc != 0 && bc != 5 && bc != 50)  // periodic: 5 or 50 (src/dg_functions.f90:1786)
                return bad(UCNS_EUNSUPPORTED, "face codes: 0 interior, 5/50 periodic, 1/2/3/4/6 boundary");
            const int j = m->ineigh[sl] - base, fj = m->ineighn[sl] - base;
            if (j < 0 || j >= nc || fj < 0 || fj >= m->ifca[j])
                return bad(UCNS_EBADARG, "face neighbour out of range at cell " + std::to_string(i));
            const size_t slj = (size_t)j * nfa + fj;
            if (m->ineigh[slj] - base != i || m->ineighn[slj] - base != f)
                return bad(UCNS_EBADARG, "face neighbours are not symmetric at cell " + std::to_string(i));
            // MUSCL bounds take the neighbours of interior faces and of periodic
            // faces of code 5, not of code 50 (FIND_BOUNDS, src/reconstruct_p.f90:1871)
            cnb[(size_t)i * nf + f] = bc == 50 ? -1 : j;
            const bool owner = (i < j) || (i == j && f < fj);
            if (!owner) continue;
            const int id = (int)farea.size();
            fcell.push_back(i);
            fcell.push_back(j);
            fslot.push_back(f | (m->types_faces[sl] == 6 ? 256 : 0));
            fslot.push_back(fj);
            const double a1 = m->faceanglex[sl], a2 = m->faceangley[sl];
            frot.push_back(std::cos(a1));
            frot.push_back(std::sin(a1));
            frot.push_back(std::cos(a2));
            frot.push_back(std::sin(a2));
            farea.push_back(m->surf[sl]);
            cface[(size_t)i * nf + f] = id + 1;
            cface[(size_t)j * nf + fj] = -(id + 1);
        }
    }
    ctx->nface = (int)farea.size();
    ctx->h_cinfo_base = cinfo;
    std::vector<double> ovol(nc), minedge(nc);
    for (int i = 0; i < nc; ++i) {
        if (!(m->totvolume[i] > 0)) return bad(UCNS_EBADARG, "non-positive volume");
        ovol[i] = 1.0 / m->totvolume[i];
        minedge[i] = m->minedge[i];
    }
    ctx->h_fcell = fcell;
    ctx->h_fslot = fslot;
    ctx->h_ovol = ovol;
    // the device resources, once the host data has been checked
    HIP_TRY(hipGetDevice(&ctx->device));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto &e : ctx->ev) HIP_TRY(hipEventCreate(&e));
    {
        const int rc0 = scratch_on(ctx->device, &ctx->scratch, ctx->err);  // with the device current
        if (rc0) return rc0;
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->cinfo, nc)) || (rc = dalloc(ctx, &ctx->cnb, cnb.size())) ||
        (rc = dalloc(ctx, &ctx->cface, cface.size())) || (rc = dalloc(ctx, &ctx->fcell, fcell.size())) ||
        (rc = dalloc(ctx, &ctx->fslot, fslot.size())) || (rc = dalloc(ctx, &ctx->frot, frot.size())) ||
        (rc = dalloc(ctx, &ctx->farea, farea.size())) || (rc = dalloc(ctx, &ctx->ovol, nc)) ||
        (rc = dalloc(ctx, &ctx->minedge, nc)) || (rc = dalloc(ctx, &ctx->U, (size_t)nc * NV)) ||
        (rc = dalloc(ctx, &ctx->U2, (size_t)nc * NV)) || (rc = dalloc(ctx, &ctx->rhs, (size_t)nc * NV)) ||
        (rc = dalloc(ctx, &ctx->fflux, (size_t)ctx->nface * NV)) || (rc = dalloc(ctx, &ctx->t, 1)) ||
        (rc = dalloc(ctx, &ctx->dtfix, 1)) || (rc = dalloc(ctx, &ctx->dtv, 2)) ||
        (rc = dalloc(ctx, &ctx->counters, 2)) || (rc = dalloc(ctx, &ctx->flag, 1)) ||
        (rc = dalloc(ctx, &ctx->stage, (size_t)nc * NV)))
        return rc;
    if (!ctx->h_perm.empty()) {
        if ((rc = dalloc(ctx, &ctx->perm, nc))) return rc;
        HIP_TRY(hipMemcpy(ctx->perm, ctx->h_perm.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    const size_t ul_elems = (size_t)nc * nf * nqm * NV;
    rc = ctx->precision == UCNS_MIXED ? dalloc(ctx, (float **)&ctx->uleft, ul_elems)
                                      : dalloc(ctx, (double **)&ctx->uleft, ul_elems);
    if (rc) return rc;
    HIP_TRY(hipMemset(ctx->uleft, 0, ul_elems * (ctx->precision == UCNS_MIXED ? 4 : 8)));
    HIP_TRY(hipMemcpy(ctx->cinfo, cinfo.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->cnb, cnb.data(), cnb.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!chnb.empty()) {
        if ((rc = dalloc(ctx, &ctx->chnb, chnb.size())) || (rc = dalloc(ctx, &ctx->chang, chang.size()))) return rc;
        HIP_TRY(hipMemcpy(ctx->chnb, chnb.data(), chnb.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->chang, chang.data(), chang.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(ctx->cface, cface.data(), cface.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->fcell, fcell.data(), fcell.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->fslot, fslot.data(), fslot.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->frot, frot.data(), frot.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->farea, farea.data(), farea.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->ovol, ovol.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->minedge, minedge.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    double wq[64] = {0}, wt[64] = {0};
    for (int q = 0; q < s->qp_quad; ++q) wq[q] = s->weights_quad[q];
    for (int q = 0; q < s->qp_triangle; ++q) wt[q] = s->weights_tri[q];
    if ((rc = dalloc(ctx, &ctx->wquad, 128))) return rc;
    HIP_TRY(hipMemcpy(ctx->wquad, wq, sizeof(wq), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->wquad + 64, wt, sizeof(wt), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(ctx->t, 0, sizeof(double)));
    return UCNS_OK;
}

int ucns_upload_prestore(ucns_ctx *ctx, const ucns_prestore_v1 *pre) {
    if (!ctx) return UCNS_EBADARG;
    if (!pre || !pre->admis || !pre->full || !pre->ihexl || !pre->invmat || !pre->integ_basis || !pre->volume11 ||
        !pre->qpoints) {
        ctx->err = "null prestore array";
        return UCNS_EBADARG;
    }
    if (ctx->iweno == 1 && !pre->indicator) {
        ctx->err = "WENO needs the indicator matrices";
        return UCNS_EBADARG;
    }
    if (ctx->ees == 5 && !ctx->have_cweno) {
        ctx->err = "CWENO (ees 5): call ucns_upload_cweno before ucns_upload_prestore";
        return UCNS_ESTATE;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    release_prestore(ctx);
    release_adjoint(ctx);  // its buffers are sized by the stencil count
    const int nc = ctx->ncell, ID = ctx->ideg, IM = ctx->imax, NNB = ctx->nnb, TSA = ctx->ts_abi;
    const int base = ctx->index_base;
    // admissible stencils, flags
    int ts = 1;
    std::vector<int32_t> cinfo = ctx->h_cinfo_base;
    for (int i = 0; i < nc; ++i) {
        const size_t o = src_cell(ctx, i);
        const int ad = pre->admis[o];
        if (ad < 1 || ad > TSA) {
            ctx->err = "admis out of range at cell " + std::to_string(o);
            return UCNS_EBADARG;
        }
        ts = std::max(ts, ad);
        cinfo[i] |= (ad << 3) | ((pre->full[o] ? 1 : 0) << 6);
    }
    if (ctx->iweno == -1 && ts != 1) {
        ctx->err = "MUSCL uses the central stencil only (admis must be 1)";
        return UCNS_EUNSUPPORTED;
    }
    ctx->ts = ts;
    const bool mixed = ctx->precision == UCNS_MIXED;
    ctx->tiled = (ctx->iweno == 1 && ctx->wenwrt == 1 && ((ID == 19 && IM == 39) || (ID == 9 && IM == 19)) &&
                  ts <= 8 && ctx->nf * ctx->nqm <= 56);
    ctx->cw_cell = ctx->ees == 5 && !ctx->tiled;
    if (ctx->cw_cell && ID != 34) {
        ctx->err = "CWENO (ees 5) runs on the WENO-3 / WENO-4 register kernels (idegfree 9 or 19) or the WENO-5 "
                   "per-cell kernel";
        return UCNS_EUNSUPPORTED;
    }
    {
        // row-stream kernels: tetrahedra (5 stencils, 24 face points) and
        // hexahedra (7 stencils, 54 face points), unless the scheme asks for
        // the register-prefetch kernels
        const int nfq = ctx->nf * ctx->nqm;
        const bool shape_ok = ID == 19 ? ((ts == 5 && nfq <= 24) || (ts == 7 && nfq > 24))
                                       : ((ts == 5 || ts == 7) && nfq <= 24);
        const bool stream_ok = ctx->kernels == UCNS_KERNELS_AUTO && ctx->tiled && shape_ok && ctx->ees != 5;
        ctx->smode = stream_ok ? ts : 0;
    }
    ctx->ntile = (nc + 7) / 8;
    // stencil members without the cell itself
    //   tiled: [tile][imax][8*ts], lane position 8-cell-tile local
    //   plain: [cell][ts][imax]
    std::vector<int32_t> nbr(ctx->smode ? (size_t)0 : ctx->tiled ? (size_t)ctx->ntile * IM * 8 * ts : (size_t)nc * ts * IM, 0);
    for (int i = 0; i < nc; ++i) {
        const size_t o = src_cell(ctx, i);
        for (int s = 0; s < pre->admis[o]; ++s) {
            const int32_t *row = pre->ihexl + (o * TSA + s) * NNB;
            if ((size_t)(row[0] - base) != o) {
                ctx->err = "IHEXL(s,1) must be the cell itself (cell " + std::to_string(o) + ")";
                return UCNS_EBADARG;
            }
            for (int k = 0; k < IM; ++k) {
                const int jo = row[k + 1] - base;
                if (jo < 0 || jo >= nc) {
                    ctx->err = "stencil member out of range at cell " + std::to_string(o);
                    return UCNS_EBADARG;
                }
                const int j = lib_cell(ctx, jo);
                if (ctx->smode) {
                    continue;  // carried by the row stream
                } else if (ctx->tiled) {
                    const size_t TP = 8 * (size_t)ts;
                    nbr[((size_t)(i / 8) * IM + k) * TP + (size_t)(i % 8) * ts + s] = j;
                } else {
                    nbr[((size_t)i * ts + s) * IM + k] = j;
                }
            }
        }
    }
    int rc = dalloc(ctx, &ctx->nbr, std::max<size_t>(nbr.size(), 1));
    if (rc) return rc;
    if (!nbr.empty())
        HIP_TRY(hipMemcpy(ctx->nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->cinfo, cinfo.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    // pseudo-inverses: staged in chunks, re-laid out on the device (CWENO on
    // the per-cell kernel: the central stencil's only)
    const int tsP = ctx->cw_cell ? 1 : ts;
    const size_t per_cell_abi = (size_t)TSA * IM * ID, per_cell_dev = (size_t)IM * tsP * ID;
    const size_t dev_elems = ctx->smode ? 1 : ctx->tiled ? (size_t)ctx->ntile * 8 * per_cell_dev : per_cell_dev * nc;
    rc = mixed ? dalloc(ctx, (float **)&ctx->P, dev_elems) : dalloc(ctx, (double **)&ctx->P, dev_elems);
    if (rc) return rc;
    const int sTP = 8 * ts, sD = stream_depth(mixed, ts), sPB = ID * sTP * (mixed ? 4 : 8);
    const int sRD = sPB + sTP * 4, sHD = sD * sTP * 4;
    if (ctx->smode) {
        ctx->tile_bytes = sHD + (long long)IM * sRD;
        if ((rc = dalloc(ctx, &ctx->rows, (size_t)ctx->ntile * ctx->tile_bytes))) return rc;
        if ((rc = dalloc(ctx, &ctx->u64, (size_t)nc * 8))) return rc;
        // member-index rows: row r holds members j = r (zero past imax)
        const size_t rows = (size_t)sD + IM;
        std::vector<int32_t> idx((size_t)ctx->ntile * rows * sTP, 0);
        for (int i = 0; i < nc; ++i) {
            const size_t o = src_cell(ctx, i);
            for (int s = 0; s < pre->admis[o]; ++s) {
                const int32_t *row = pre->ihexl + (o * TSA + s) * NNB;
                for (int r = 0; r < IM; ++r)
                    idx[((size_t)(i / 8) * rows + r) * sTP + (size_t)(i % 8) * ts + s] = lib_cell(ctx, row[r + 1] - base);
            }
        }
        int32_t *didx = nullptr;
        if ((rc = dalloc(ctx, &didx, idx.size()))) return rc;
        HIP_TRY(hipMemcpy(didx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_stream_idx, dim3(4096), dim3(256), 0, ctx->stream, didx, ctx->rows, ctx->ntile, sD, IM,
                           sTP, sHD, sPB, sRD, ctx->tile_bytes);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        dfree(didx);
    }
    {
        size_t chunk_cells = std::max<size_t>(8, (size_t)(256u << 20) / (per_cell_abi * sizeof(double)));
        chunk_cells -= chunk_cells % 8;
        double *stage = nullptr, *host_pin = nullptr;
        if ((rc = dalloc(ctx, &stage, chunk_cells * per_cell_abi))) return rc;
        if (hipHostMalloc((void **)&host_pin, chunk_cells * per_cell_abi * sizeof(double), hipHostMallocDefault) !=
            hipSuccess)
            host_pin = nullptr;
        std::vector<double> host_buf;  // pageable staging when pinned memory is refused and cells move
        if (!host_pin && (!ctx->h_perm.empty() || ctx->ees == 5)) host_buf.resize(chunk_cells * per_cell_abi);
        double *gather = host_pin ? host_pin : host_buf.data();
        for (size_t c0 = 0; c0 < (size_t)nc; c0 += chunk_cells) {
            const size_t cn = std::min(chunk_cells, (size_t)nc - c0);
            const size_t bytes = cn * per_cell_abi * sizeof(double);
            const double *src = pre->invmat + c0 * per_cell_abi;
            if (!ctx->h_perm.empty() || ctx->ees == 5) {  // the chunk's cells from their places in the caller's order
                for (size_t c = 0; c < cn; ++c)
                    std::memcpy(gather + c * per_cell_abi, pre->invmat + src_cell(ctx, (int)(c0 + c)) * per_cell_abi,
                                per_cell_abi * sizeof(double));
                if (ctx->ees == 5 && ctx->tiled) {
                    // CWENO: a sector's rows are its low-order pseudo-inverse, zero
                    // past idegfree2 coefficients and imax2 members
                    const int ID2 = ctx->ideg2, IM2 = ctx->imax2;
                    for (size_t c = 0; c < cn; ++c)
                        for (int st = 1; st < TSA; ++st) {
                            double *Pd = gather + c * per_cell_abi + (size_t)st * IM * ID;
                            const double *Pc = ctx->h_cwP.data() + (((c0 + c) * TSA + st) * IM2) * ID2;
                            std::fill(Pd, Pd + (size_t)IM * ID, 0.0);
                            for (int j = 0; j < IM2; ++j)
                                for (int k = 0; k < ID2; ++k) Pd[(size_t)j * ID + k] = Pc[(size_t)j * ID2 + k];
                        }
                }
                src = gather;
            } else if (host_pin) {
                std::memcpy(host_pin, src, bytes);
                src = host_pin;
            }
            HIP_TRY(hipMemcpyAsync(stage, src, bytes, hipMemcpyHostToDevice, ctx->stream));
            if (ctx->smode) {
                const int blocks = (int)std::min<size_t>(4096, (cn * per_cell_dev + 255) / 256);
                if (mixed)
                    hipLaunchKernelGGL(k_stream_p<float>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       ctx->rows, (int)c0, (int)cn, TSA, ts, IM, ID, sHD, sRD, ctx->tile_bytes);
                else
                    hipLaunchKernelGGL(k_stream_p<double>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       ctx->rows, (int)c0, (int)cn, TSA, ts, IM, ID, sHD, sRD, ctx->tile_bytes);
            } else if (ctx->tiled) {
                const size_t out_elems = ((cn + 7) / 8) * 8 * per_cell_dev;
                const int blocks = (int)std::min<size_t>(4096, (out_elems + 255) / 256);
                if (mixed)
                    hipLaunchKernelGGL(k_tile_p<float>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       (float *)ctx->P + c0 * per_cell_dev, (int)c0, (int)cn, nc, TSA, ts, IM, ID);
                else
                    hipLaunchKernelGGL(k_tile_p<double>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       (double *)ctx->P + c0 * per_cell_dev, (int)c0, (int)cn, nc, TSA, ts, IM, ID);
            } else {
                const int blocks = (int)std::min<size_t>(4096, (cn * per_cell_dev + 255) / 256);
                if (mixed)
                    hipLaunchKernelGGL(k_relayout_p<float>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       (float *)ctx->P + c0 * per_cell_dev, (int)cn, TSA, tsP, IM, ID);
                else
                    hipLaunchKernelGGL(k_relayout_p<double>, dim3(blocks), dim3(256), 0, ctx->stream, stage,
                                       (double *)ctx->P + c0 * per_cell_dev, (int)cn, TSA, tsP, IM, ID);
            }
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
        dfree(stage);
        if (host_pin) (void)hipHostFree(host_pin);
    }
    // indicators, packed upper triangle
    if (ctx->iweno == 1) {
        const size_t np = (size_t)ID * (ID + 1) / 2;
        std::vector<double> bp((size_t)nc * np);
        for (int i = 0; i < nc; ++i) {
            const double *Bi = pre->indicator + src_cell(ctx, i) * ID * ID;
            size_t e = (size_t)i * np;
            for (int k = 0; k < ID; ++k)
                for (int l = k; l < ID; ++l) bp[e++] = Bi[(size_t)k * ID + l];
        }
        // zero pad: the tile loads move whole 16-byte pieces, the last of
        // which may reach past the final cell's block
        bp.resize(bp.size() + 512, 0.0);
        if ((rc = upload_real_any(ctx, &ctx->B, bp))) return rc;
    } else {
        std::vector<double> bp(1, 0.0);
        if ((rc = upload_real_any(ctx, &ctx->B, bp))) return rc;
    }
    // basis means
    {
        std::vector<double> mb((size_t)nc * ID);
        ctx->h_vol11.resize(nc);
        for (int i = 0; i < nc; ++i) {
            const size_t o = src_cell(ctx, i);
            ctx->h_vol11[i] = pre->volume11[o];
            const double oov = 1.0 / pre->volume11[o];
            for (int d = 0; d < ID; ++d) mb[(size_t)i * ID + d] = pre->integ_basis[o * ID + d] * oov;
        }
        // zero pad, as B: the row-stream kernels stage a tile's means in LDS
        // with whole 16-byte pieces
        mb.resize(mb.size() + 512, 0.0);
        if ((rc = upload_real_any(ctx, &ctx->mbar, mb))) return rc;
        if (ctx->ees == 5) {
            // the sectors' indicators (packed) and mbar - mbar_c over their coefficients
            const int ID2 = ctx->ideg2, NB2 = ID2 * (ID2 + 1) / 2;
            std::vector<double> bc((size_t)nc * NB2), dm((size_t)nc * ID2);
            for (int i = 0; i < nc; ++i) {
                const double *Bi = ctx->h_cwB.data() + (size_t)i * ID2 * ID2;
                size_t e = (size_t)i * NB2;
                for (int k = 0; k < ID2; ++k)
                    for (int l = k; l < ID2; ++l) bc[e++] = Bi[(size_t)k * ID2 + l];
                const size_t o = src_cell(ctx, i);
                const double oov = 1.0 / pre->volume11[o];
                for (int d = 0; d < ID2; ++d)
                    dm[(size_t)i * ID2 + d] = mb[(size_t)i * ID + d] - ctx->h_cwIB[(size_t)i * ID2 + d] * oov;
            }
            if ((rc = upload_real_any(ctx, &ctx->Bc, bc)) || (rc = upload_real_any(ctx, &ctx->dmb, dm))) return rc;
            if (ctx->cw_cell) {
                // the sectors' low-order pseudo-inverses, stencils 2..ts
                const int IM2 = ctx->imax2;
                const size_t pc = (size_t)IM2 * ID2;
                std::vector<double> P2((size_t)nc * std::max(ts - 1, 1) * pc, 0.0);
                for (int i = 0; i < nc; ++i)
                    for (int st = 1; st < ts; ++st)
                        std::memcpy(P2.data() + ((size_t)i * (ts - 1) + st - 1) * pc,
                                    ctx->h_cwP.data() + ((size_t)i * TSA + st) * pc, pc * sizeof(double));
                if ((rc = upload_real_any(ctx, &ctx->Pc, P2))) return rc;
            }
        }
    }
    // Gauss points in the local frames: [cell][nf][nqm][3]
    {
        const int nf = ctx->nf, nqm = ctx->nqm, nfa = ctx->nf_abi, nqa = ctx->nqm_abi;
        std::vector<double> q((size_t)nc * nf * nqm * 3 + 512, 0.0);  // zero pad, as the means
        for (int i = 0; i < nc; ++i)
            for (int f = 0; f < nf; ++f)
                for (int g = 0; g < nqm && g < nqa; ++g)
                    for (int d = 0; d < 3; ++d)
                        q[(((size_t)i * nf + f) * nqm + g) * 3 + d] =
                            pre->qpoints[((src_cell(ctx, i) * nfa + f) * nqa + g) * 3 + d];
        if ((rc = upload_real_any(ctx, &ctx->qp, q))) return rc;
    }
    ctx->have_prestore = true;
    ctx->st.bytes_recon_per_cell =
        (double)((size_t)IM * ts * ID * (mixed ? 4 : 8)) + (ctx->iweno == 1 ? (double)ID * (ID + 1) / 2 * (mixed ? 4 : 8) : 0) +
        (double)ts * IM * 4 + (double)ctx->nf * ctx->nqm * 3 * (mixed ? 4 : 8) + (double)ID * (mixed ? 4 : 8);
    if (ctx->cw_cell) {
        // central P, the sectors' low-order P, both indicators, the members read, mbar - mbar_c
        const int ID2 = ctx->ideg2, IM2 = ctx->imax2, r = mixed ? 4 : 8;
        ctx->st.bytes_recon_per_cell = (double)IM * ID * r + (double)(ts - 1) * IM2 * ID2 * r +
                                       (double)ID * (ID + 1) / 2 * r + (double)ID2 * (ID2 + 1) / 2 * r +
                                       (double)(IM + (ts - 1) * IM2) * 4 + (double)ctx->nf * ctx->nqm * 3 * r +
                                       (double)(ID + ID2) * r;
    }
    return UCNS_OK;
}

int ucns_upload_cweno(ucns_ctx *ctx, const ucns_cweno_v1 *cw) {
    if (!ctx || !cw) return UCNS_EBADARG;
    if (cw->abi_version != UCNS_ABI_VERSION) {
        ctx->err = "ucns_cweno_v1: abi_version mismatch";
        return UCNS_EBADARG;
    }
    if (ctx->ees != 5) {
        ctx->err = "ucns_upload_cweno: the scheme is not CWENO (ees 5)";
        return UCNS_ESTATE;
    }
    if (!cw->invmat_c || !cw->indicator_c || !cw->integ_basis_c) {
        ctx->err = "null CWENO array";
        return UCNS_EBADARG;
    }
    // IORDER <= 3 (idegfree 9 and 19) take linear sectors, IORDER 4 (idegfree
    // 34) quadratic ones (DETERMINE_SIZE, src/grid_p.f90:6151-6161)
    const int want2 = ctx->ideg == 34 ? 9 : 3;
    if (cw->idegfree2 != want2 || cw->inumneighbours2 < want2 + 1 || cw->inumneighbours2 > ctx->nnb) {
        ctx->err = "CWENO sectors: idegfree2 must be " + std::to_string(want2) +
                   " for this idegfree and inumneighbours2 in [idegfree2 + 1, inumneighbours]";
        return UCNS_EUNSUPPORTED;
    }
    const int nc = ctx->ncell, TSA = ctx->ts_abi, ID2 = cw->idegfree2, IM2 = cw->inumneighbours2 - 1;
    const size_t pc = (size_t)TSA * IM2 * ID2;
    ctx->h_cwP.resize((size_t)nc * pc);
    ctx->h_cwB.resize((size_t)nc * ID2 * ID2);
    ctx->h_cwIB.resize((size_t)nc * ID2);
    for (int i = 0; i < nc; ++i) {
        const size_t o = src_cell(ctx, i);
        std::memcpy(ctx->h_cwP.data() + (size_t)i * pc, cw->invmat_c + o * pc, pc * sizeof(double));
        std::memcpy(ctx->h_cwB.data() + (size_t)i * ID2 * ID2, cw->indicator_c + o * ID2 * ID2,
                    (size_t)ID2 * ID2 * sizeof(double));
        std::memcpy(ctx->h_cwIB.data() + (size_t)i * ID2, cw->integ_basis_c + o * ID2, (size_t)ID2 * sizeof(double));
    }
    ctx->ideg2 = ID2;
    ctx->imax2 = IM2;
    ctx->have_cweno = true;
    ctx->have_prestore = false;  // the pseudo-inverses are laid out again by ucns_upload_prestore
    return UCNS_OK;
}

int ucns_set_boundary(ucns_ctx *ctx, const ucns_boundary_v1 *bc) {
    if (!ctx || !bc) return UCNS_EBADARG;
    if (bc->abi_version != UCNS_ABI_VERSION) {
        ctx->err = "abi_version mismatch";
        return UCNS_EBADARG;
    }
    if (bc->boundtype != 0 && bc->boundtype != 1) {
        ctx->err = "boundtype must be 0 or 1";
        return UCNS_EBADARG;
    }
    for (int v = 0; v < NV; ++v)
        if (!std::isfinite(bc->inflow[v]) || !std::isfinite(bc->outflow[v])) {
            ctx->err = "non-finite free-stream state";
            return UCNS_EBADARG;
        }
    if (!(bc->inflow[0] > 0) || !(bc->outflow[0] > 0)) {
        ctx->err = "free-stream density must be positive";
        return UCNS_EBADARG;
    }
    ctx->bc = *bc;
    ctx->have_bc = true;
    return UCNS_OK;
}

int ucns_set_boundary_faces(ucns_ctx *ctx, const double *inflow, const double *outflow) {
    if (!ctx) return UCNS_EBADARG;
    if (!ctx->have_bc) {
        ctx->err = "call ucns_set_boundary first (BOUNDTYPE and the states of the faces not given here)";
        return UCNS_ESTATE;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    dfree(ctx->bface);
    ctx->bface = nullptr;
    if (!inflow && !outflow) return UCNS_OK;
    const int nfa = ctx->nf_abi;
    std::vector<double> h((size_t)ctx->nface * 10);
    for (int f = 0; f < ctx->nface; ++f) {
        const int code = -ctx->h_fcell[2 * f + 1];
        const int ci = ctx->h_fcell[2 * f], sl = ctx->h_fslot[2 * f] & 255;
        const size_t rec = (src_cell(ctx, ci) * nfa + sl) * NV;
        for (int v = 0; v < NV; ++v) {
            h[(size_t)f * 10 + v] = ctx->bc.inflow[v];
            h[(size_t)f * 10 + 5 + v] = ctx->bc.outflow[v];
        }
        if (code <= 0) continue;  // interior and periodic faces: no ghost
        for (int v = 0; v < NV; ++v) {
            if (inflow) h[(size_t)f * 10 + v] = inflow[rec + v];
            if (outflow) h[(size_t)f * 10 + 5 + v] = outflow[rec + v];
        }
        for (int v = 0; v < 10; ++v)
            if (!std::isfinite(h[(size_t)f * 10 + v])) {
                ctx->err = "non-finite boundary state at cell " + std::to_string(src_cell(ctx, ci)) + " face " +
                           std::to_string(sl);
                return UCNS_EBADARG;
            }
        if (!(h[(size_t)f * 10] > 0) || !(h[(size_t)f * 10 + 5] > 0)) {
            ctx->err = "boundary density must be positive (cell " + std::to_string(src_cell(ctx, ci)) + " face " +
                       std::to_string(sl) + ")";
            return UCNS_EBADARG;
        }
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->bface, h.size()))) return rc;
    HIP_TRY(hipMemcpy(ctx->bface, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return UCNS_OK;
}

int ucns_set_state(ucns_ctx *ctx, const double *u, int32_t layout) {
    if (!ctx || !u) return UCNS_EBADARG;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = cells_in(ctx, u, true, layout, ctx->U);
    if (rc == 0) ctx->have_state = true;
    return rc;
}

int ucns_get_state(ucns_ctx *ctx, double *u, int32_t layout) {
    if (!ctx || !u) return UCNS_EBADARG;
    HIP_TRY(hipSetDevice(ctx->device));
    return cells_out(ctx, ctx->U, layout, u, true);
}

int ucns_set_state_device(ucns_ctx *ctx, const double *u, int32_t layout) {
    if (!ctx || !u) return UCNS_EBADARG;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = cells_in(ctx, u, false, layout, ctx->U);
    if (rc == 0) ctx->have_state = true;
    return rc;
}

int ucns_get_state_device(ucns_ctx *ctx, double *u, int32_t layout) {
    if (!ctx || !u) return UCNS_EBADARG;
    HIP_TRY(hipSetDevice(ctx->device));
    return cells_out(ctx, ctx->U, layout, u, false);
}

static int ready(ucns_ctx *ctx) {
    if (ctx->needs_bc && !ctx->have_bc) {
        ctx->err = "the mesh has inflow/outflow/farfield faces: call ucns_set_boundary first";
        return UCNS_ESTATE;
    }
    if (!ctx->have_prestore) {
        ctx->err = "ucns_upload_prestore has not been called";
        return UCNS_ESTATE;
    }
    if (!ctx->have_state) {
        ctx->err = "ucns_set_state has not been called";
        return UCNS_ESTATE;
    }
    return UCNS_OK;
}

int ucns_compute_dt(ucns_ctx *ctx, double cfl, double *dt_out) {
    if (!ctx || !dt_out || !(cfl > 0)) return UCNS_EBADARG;
    if (!ctx->have_state) {
        ctx->err = "ucns_set_state has not been called";
        return UCNS_ESTATE;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    dt_kernel(ctx, cfl, &ctx->dtv[0]);
    HIP_TRY(hipMemcpyAsync(dt_out, &ctx->dtv[0], sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UCNS_OK;
}

int ucns_reconstruct(ucns_ctx *ctx, double *uleft_out) {
    if (!ctx || !uleft_out) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 2);
    rc = ctx->precision == UCNS_MIXED ? launch_recon_t<float>(ctx) : launch_recon_t<double>(ctx);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    const size_t n = (size_t)ctx->ncell * ctx->nf * ctx->nqm * NV;
    std::vector<double> d(n);
    if (ctx->precision == UCNS_MIXED) {
        std::vector<float> f(n);
        HIP_TRY(hipMemcpyAsync(f.data(), ctx->uleft, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (size_t e = 0; e < n; ++e) d[e] = f[e];
    } else {
        HIP_TRY(hipMemcpyAsync(d.data(), ctx->uleft, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    // the device keeps deviations from the cell average; hand out states
    std::vector<double> u((size_t)ctx->ncell * NV);
    HIP_TRY(hipMemcpy(u.data(), ctx->U, u.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int nfa = ctx->nf_abi, nqa = ctx->nqm_abi;
    std::memset(uleft_out, 0, sizeof(double) * (size_t)ctx->ncell * nfa * nqa * NV);
    for (int i = 0; i < ctx->ncell; ++i) {
        const size_t o = src_cell(ctx, i);
        for (int f = 0; f < ctx->nf; ++f)
            for (int q = 0; q < ctx->nqm; ++q)
                for (int v = 0; v < NV; ++v)
                    uleft_out[((o * nfa + f) * nqa + q) * NV + v] =
                        u[(size_t)i * NV + v] + d[(((size_t)i * ctx->nf + f) * ctx->nqm + q) * NV + v];
    }
    int cnt[2];
    HIP_TRY(hipMemcpy(cnt, ctx->counters, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->st.n_reduced_last = cnt[0];
    ctx->st.n_firstorder_last = cnt[1];
    return UCNS_OK;
}

static int residual_to(ucns_ctx *ctx, double *rhs_out, int32_t layout, bool host) {
    if (!ctx) return UCNS_EBADARG;
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = residual_kernels(ctx))) return rc;
    update_kernel(ctx, 0, nullptr, nullptr, 0.0, ctx->rhs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int cnt[2];
    HIP_TRY(hipMemcpy(cnt, ctx->counters, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->st.n_reduced_last = cnt[0];
    ctx->st.n_firstorder_last = cnt[1];
    if (rhs_out) return cells_out(ctx, ctx->rhs, layout, rhs_out, host);
    return UCNS_OK;
}

int ucns_residual(ucns_ctx *ctx, double *rhs_out, int32_t layout) { return residual_to(ctx, rhs_out, layout, true); }

int ucns_residual_noalloc(ucns_ctx *ctx, double *rhs_dev, int32_t layout) {
    if (!ctx || !rhs_dev) return UCNS_EBADARG;
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    const size_t need = UCNS_SCRATCH_HEAD + (size_t)ctx->nface * NV * sizeof(double);
    if (need > UCNS_SCRATCH_BYTES) {
        ctx->err = "ucns_residual_noalloc: " + std::to_string(ctx->nface) + " faces need " + std::to_string(need) +
                   " bytes of scratch, more than the library's " + std::to_string(UCNS_SCRATCH_BYTES) +
                   " (ucns_residual_device has no such limit)";
        return UCNS_EUNSUPPORTED;
    }
    if (!ctx->scratch) {
        ctx->err = "ucns_residual_noalloc: no scratch address for this device";
        return UCNS_ESTATE;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // the same kernels as ucns_residual_device, with their counters and face
    // fluxes in the static scratch (the counters are cleared by a kernel at
    // the start of every reconstruction; every face flux is written before
    // it is read)
    int *const own_counters = ctx->counters;
    double *const own_fflux = ctx->fflux;
    ctx->counters = (int *)ctx->scratch;
    ctx->fflux = (double *)((unsigned char *)ctx->scratch + UCNS_SCRATCH_HEAD);
    rc = residual_kernels(ctx);
    if (rc == UCNS_OK) update_kernel(ctx, 0, nullptr, nullptr, 0.0, ctx->rhs);
    int cnt[2] = {0, 0};
    if (rc == UCNS_OK && hipMemcpyAsync(cnt, ctx->counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = UCNS_EHIP;
    if (rc == UCNS_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = UCNS_EHIP;
    ctx->counters = own_counters;
    ctx->fflux = own_fflux;
    if (rc == UCNS_EHIP) {
        ctx->err = std::string("ucns_residual_noalloc: ") + hipGetErrorString(hipGetLastError());
        return rc;
    }
    if (rc) return rc;
    ctx->st.n_reduced_last = cnt[0];
    ctx->st.n_firstorder_last = cnt[1];
    return cells_out(ctx, ctx->rhs, layout, rhs_dev, false);
}

int ucns_residual_device(ucns_ctx *ctx, double *rhs_dev, int32_t layout) {
    if (!rhs_dev) return UCNS_EBADARG;
    return residual_to(ctx, rhs_dev, layout, false);
}

int ucns_rk3_steps(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, double *t_io) {
    if (!ctx || nsteps < 0) return UCNS_EBADARG;
    if (!(dt_fixed > 0) && !(cfl > 0)) {
        ctx->err = "need cfl > 0 or dt_fixed > 0";
        return UCNS_EBADARG;
    }
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const double t0 = t_io ? *t_io : 0.0;
    HIP_TRY(hipMemcpyAsync(ctx->t, &t0, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const bool fixed = dt_fixed > 0;
    if (fixed) HIP_TRY(hipMemcpyAsync(ctx->dtfix, &dt_fixed, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int cur = 0;
    if (!fixed && nsteps > 0) dt_kernel(ctx, cfl, &ctx->dtv[cur]);
    for (int it = 0; it < nsteps; ++it) {
        const double *dt = fixed ? ctx->dtfix : (const double *)&ctx->dtv[cur];
        unsigned long long *next = fixed ? nullptr : &ctx->dtv[1 - cur];
        for (int stage = 1; stage <= 3; ++stage) {
            if ((rc = residual_kernels(ctx))) return rc;
            if (stage == 3 && next)
                hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, ctx->stream, next,
                                   (unsigned long long)0x7FF0000000000000ULL);
            update_kernel(ctx, stage, dt, stage == 3 ? next : nullptr, cfl, nullptr);
        }
        HIP_TRY(hipGetLastError());
        if (!fixed) cur = 1 - cur;
    }
    HIP_TRY(hipMemsetAsync(ctx->flag, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_nonfinite, dim3(1024), dim3(256), 0, ctx->stream, ctx->U, (size_t)ctx->ncell * NV,
                       ctx->flag);
    int flag = 0;
    double t1 = t0;
    HIP_TRY(hipMemcpyAsync(&flag, ctx->flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&t1, ctx->t, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t_io) *t_io = t1;
    int cnt[2];
    HIP_TRY(hipMemcpy(cnt, ctx->counters, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->st.n_reduced_last = cnt[0];
    ctx->st.n_firstorder_last = cnt[1];
    if (flag) {
        ctx->err = "non-finite state after the RK3 steps";
        return UCNS_ENONFINITE;
    }
    return UCNS_OK;
}

int ucns_residual_vjp_device(ucns_ctx *ctx, const double *w_dev, double *u_bar_dev, double *bc_bar_dev) {
    if (!ctx || !w_dev || !u_bar_dev) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ensure_adjoint(ctx))) return rc;
    if ((rc = cells_in(ctx, w_dev, false, UCNS_LAYOUT_AOS, ctx->adj_a))) return rc;
    zero_dev(ctx, ctx->adj_bc, 2 * NV);
    if ((rc = residual_vjp_dev(ctx, ctx->adj_a, ctx->adj_x, ctx->adj_bc))) return rc;
    HIP_TRY(hipGetLastError());
    if (bc_bar_dev)
        HIP_TRY(hipMemcpyAsync(bc_bar_dev, ctx->adj_bc, 2 * NV * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
    return cells_out(ctx, ctx->adj_x, UCNS_LAYOUT_AOS, u_bar_dev, false);
}

int ucns_residual_vjp(ucns_ctx *ctx, const double *w, double *u_bar, int32_t layout, double *bc_bar) {
    if (!ctx || !w || !u_bar) return UCNS_EBADARG;
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ensure_adjoint(ctx))) return rc;
    if ((rc = cells_in(ctx, w, true, layout, ctx->adj_a))) return rc;
    zero_dev(ctx, ctx->adj_bc, 2 * NV);
    if ((rc = residual_vjp_dev(ctx, ctx->adj_a, ctx->adj_x, ctx->adj_bc))) return rc;
    HIP_TRY(hipGetLastError());
    double bc[2 * NV];
    HIP_TRY(hipMemcpyAsync(bc, ctx->adj_bc, sizeof(bc), hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = cells_out(ctx, ctx->adj_x, layout, u_bar, true))) return rc;
    if (bc_bar) std::memcpy(bc_bar, bc, sizeof(bc));
    return UCNS_OK;
}

// forward sweep with stage checkpoints + reverse sweep; a_dev holds uN_bar
// on entry and u0_bar on return
static int rk3_adjoint_impl(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, double *a_dev, double *t_io) {
    if (nsteps < 1) {
        ctx->err = "nsteps must be >= 1";
        return UCNS_EBADARG;
    }
    if (!(dt_fixed > 0) && !(cfl > 0)) {
        ctx->err = "need cfl > 0 or dt_fixed > 0";
        return UCNS_EBADARG;
    }
    int rc = ensure_adjoint(ctx);
    if (rc) return rc;
    const size_t n = (size_t)ctx->ncell * NV;
    if (ctx->ckpt_steps < (size_t)nsteps) {
        dfree(ctx->ckpt);
        dfree(ctx->dts);
        ctx->ckpt_steps = 0;
        if ((rc = dalloc(ctx, &ctx->ckpt, (size_t)3 * nsteps * n))) return rc;
        if ((rc = dalloc(ctx, &ctx->dts, (size_t)nsteps))) return rc;
        ctx->ckpt_steps = nsteps;
    }
    const int keep0 = 3 * nsteps - adj_keep(ctx, 3 * nsteps);
    const double t0 = t_io ? *t_io : 0.0;
    HIP_TRY(hipMemcpyAsync(ctx->t, &t0, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const bool fixed = dt_fixed > 0;
    if (fixed) HIP_TRY(hipMemcpyAsync(ctx->dtfix, &dt_fixed, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int cur = 0;
    if (!fixed) dt_kernel(ctx, cfl, &ctx->dtv[cur]);
    for (int it = 0; it < nsteps; ++it) {
        const double *dt = fixed ? ctx->dtfix : (const double *)&ctx->dtv[cur];
        unsigned long long *next = fixed ? nullptr : &ctx->dtv[1 - cur];
        HIP_TRY(hipMemcpyAsync(ctx->dts + it, dt, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        if ((rc = rk3_step_ckpt(ctx, it, dt, next, cfl, keep0))) return rc;
        HIP_TRY(hipGetLastError());
        if (!fixed) cur = 1 - cur;
    }
    zero_dev(ctx, ctx->adj_bc, 2 * NV);
    if ((rc = rk3_reverse(ctx, nsteps, a_dev, ctx->adj_bc, keep0))) return rc;
    HIP_TRY(hipGetLastError());
    double t1 = t0;
    HIP_TRY(hipMemcpyAsync(&t1, ctx->t, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t_io) *t_io = t1;
    return UCNS_OK;
}

int ucns_rk3_adjoint_device(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, const double *uN_bar_dev,
                            double *u0_bar_dev, double *bc_bar_dev, double *t_io) {
    if (!ctx || !uN_bar_dev || !u0_bar_dev) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ensure_adjoint(ctx))) return rc;
    if ((rc = cells_in(ctx, uN_bar_dev, false, UCNS_LAYOUT_AOS, ctx->adj_a))) return rc;
    if ((rc = rk3_adjoint_impl(ctx, nsteps, cfl, dt_fixed, ctx->adj_a, t_io))) return rc;
    if (bc_bar_dev)
        HIP_TRY(hipMemcpyAsync(bc_bar_dev, ctx->adj_bc, 2 * NV * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
    return cells_out(ctx, ctx->adj_a, UCNS_LAYOUT_AOS, u0_bar_dev, false);
}

int ucns_rk3_adjoint(ucns_ctx *ctx, int32_t nsteps, double cfl, double dt_fixed, const double *uN_bar,
                     double *u0_bar, int32_t layout, double *bc_bar, double *t_io) {
    if (!ctx || !uN_bar || !u0_bar) return UCNS_EBADARG;
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) return UCNS_EBADARG;
    int rc = ready(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ensure_adjoint(ctx))) return rc;
    if ((rc = cells_in(ctx, uN_bar, true, layout, ctx->adj_a))) return rc;
    if ((rc = rk3_adjoint_impl(ctx, nsteps, cfl, dt_fixed, ctx->adj_a, t_io))) return rc;
    double bc[2 * NV];
    HIP_TRY(hipMemcpyAsync(bc, ctx->adj_bc, sizeof(bc), hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = cells_out(ctx, ctx->adj_a, layout, u0_bar, true))) return rc;
    if (bc_bar) std::memcpy(bc_bar, bc, sizeof(bc));
    return UCNS_OK;
}

namespace {
void release_geometry(ucns_ctx *ctx) {
    dfree(ctx->geo_P);
    dfree(ctx->geo_B);
    dfree(ctx->geo_M);
    dfree(ctx->geo_Q);
    dfree(ctx->geo_area);
    dfree(ctx->geo_nrm);
    dfree(ctx->geo_ovol);
    ctx->geo_on = false;
}

int geo_download(ucns_ctx *ctx, const double *dev, std::vector<double> &host, size_t n) {
    host.resize(n);
    HIP_TRY(hipMemcpyAsync(host.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return UCNS_OK;
}

// the accumulated device cotangents in the caller's arrays and cell order
int geo_to_caller(ucns_ctx *ctx, const ucns_geometry_bar_v1 *g) {
    const int nc = ctx->ncell, ID = ctx->ideg, IM = ctx->imax, ts = ctx->ts, TSA = ctx->ts_abi;
    const int nf = ctx->nf, nqm = ctx->nqm, nfa = ctx->nf_abi, nqa = ctx->nqm_abi;
    std::vector<double> P, B, M, Q, A, N, V, mb;
    int rc;
    if ((g->invmat && (rc = geo_download(ctx, ctx->geo_P, P, (size_t)nc * ts * IM * ID))) ||
        (g->indicator && ctx->iweno == 1 && (rc = geo_download(ctx, ctx->geo_B, B, (size_t)nc * ID * ID))) ||
        ((g->integ_basis || g->volume11) && (rc = geo_download(ctx, ctx->geo_M, M, (size_t)nc * ID))) ||
        (g->volume11 && (rc = geo_download(ctx, (const double *)ctx->mbar, mb, (size_t)nc * ID))) ||
        (g->qpoints && (rc = geo_download(ctx, ctx->geo_Q, Q, (size_t)nc * nf * nqm * 3))) ||
        (g->surf && (rc = geo_download(ctx, ctx->geo_area, A, (size_t)ctx->nface))) ||
        (g->normal && (rc = geo_download(ctx, ctx->geo_nrm, N, (size_t)ctx->nface * 3))) ||
        (g->totvolume && (rc = geo_download(ctx, ctx->geo_ovol, V, (size_t)nc))))
        return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (g->invmat) std::memset(g->invmat, 0, sizeof(double) * (size_t)nc * TSA * IM * ID);
    if (g->indicator) std::memset(g->indicator, 0, sizeof(double) * (size_t)nc * ID * ID);
    if (g->qpoints) std::memset(g->qpoints, 0, sizeof(double) * (size_t)nc * nfa * nqa * 3);
    if (g->surf) std::memset(g->surf, 0, sizeof(double) * (size_t)nc * nfa);
    if (g->normal) std::memset(g->normal, 0, sizeof(double) * (size_t)nc * nfa * 3);
    for (int i = 0; i < nc; ++i) {
        const size_t o = src_cell(ctx, i);
        if (g->invmat)
            for (int st = 0; st < std::min(ts, TSA); ++st)
                std::memcpy(g->invmat + ((o * TSA + st) * IM) * ID, P.data() + (((size_t)i * ts + st) * IM) * ID,
                            sizeof(double) * IM * ID);
        if (g->indicator && !B.empty())
            std::memcpy(g->indicator + o * ID * ID, B.data() + (size_t)i * ID * ID, sizeof(double) * ID * ID);
        // mbar = INTEG_BASIS / VOLUME(1,1)
        const double v11 = ctx->h_vol11[i];
        double vb = 0.0;
        for (int d = 0; d < ID; ++d) {
            const double m = M.empty() ? 0.0 : M[(size_t)i * ID + d];
            if (g->integ_basis) g->integ_basis[o * ID + d] = m / v11;
            if (g->volume11) vb -= m * mb[(size_t)i * ID + d] / v11;
        }
        if (g->volume11) g->volume11[o] = vb;
        if (g->qpoints)
            for (int f = 0; f < std::min(nf, nfa); ++f)
                for (int q = 0; q < std::min(nqm, nqa); ++q)
                    for (int d = 0; d < 3; ++d)
                        g->qpoints[((o * nfa + f) * nqa + q) * 3 + d] = Q[(((size_t)i * nf + f) * nqm + q) * 3 + d];
        // 1/TOTVOLUME
        if (g->totvolume) g->totvolume[o] = -V[i] * ctx->h_ovol[i] * ctx->h_ovol[i];
    }
    for (int f = 0; f < ctx->nface; ++f) {
        const size_t o = src_cell(ctx, ctx->h_fcell[2 * f]);
        const int sl = ctx->h_fslot[2 * f] & 255;
        if (g->surf) g->surf[o * nfa + sl] = A[f];
        if (g->normal)
            for (int d = 0; d < 3; ++d) g->normal[(o * nfa + sl) * 3 + d] = N[(size_t)f * 3 + d];
    }
    return UCNS_OK;
}
}  // namespace

int ucns_rk3_adjoint_geometry(ucns_ctx *ctx, int32_t nsteps, double dt_fixed, const double *uN_bar, double *u0_bar,
                              int32_t layout, double *bc_bar, const ucns_geometry_bar_v1 *gbar, double *t_io) {
    if (!ctx || !uN_bar || !u0_bar || !gbar) return UCNS_EBADARG;
    if (layout != UCNS_LAYOUT_AOS && layout != UCNS_LAYOUT_SOA) return UCNS_EBADARG;
    if (gbar->abi_version != UCNS_ABI_VERSION) {
        ctx->err = "ucns_geometry_bar_v1: abi_version mismatch";
        return UCNS_EBADARG;
    }
    if (!(dt_fixed > 0)) {
        ctx->err = "ucns_rk3_adjoint_geometry: needs dt_fixed > 0 (the step size is held fixed under the geometry)";
        return UCNS_EBADARG;
    }
    int rc = ready(ctx);
    if (rc) return rc;
    if (ctx->precision != UCNS_FP64 || ctx->smode) {
        ctx->err = "ucns_rk3_adjoint_geometry: needs an FP64 context on the register kernels "
                   "(precision UCNS_FP64; kernels UCNS_KERNELS_REGISTER where the row-stream kernels would run)";
        return UCNS_EUNSUPPORTED;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ensure_adjoint(ctx))) return rc;
    const size_t nc = ctx->ncell;
    release_geometry(ctx);
    if ((rc = dalloc(ctx, &ctx->geo_P, nc * ctx->ts * ctx->imax * ctx->ideg)) ||
        (rc = dalloc(ctx, &ctx->geo_B, nc * ctx->ideg * ctx->ideg)) || (rc = dalloc(ctx, &ctx->geo_M, nc * ctx->ideg)) ||
        (rc = dalloc(ctx, &ctx->geo_Q, nc * ctx->nf * ctx->nqm * 3)) ||
        (rc = dalloc(ctx, &ctx->geo_area, (size_t)ctx->nface)) ||
        (rc = dalloc(ctx, &ctx->geo_nrm, (size_t)ctx->nface * 3)) || (rc = dalloc(ctx, &ctx->geo_ovol, nc))) {
        release_geometry(ctx);
        return rc;
    }
    zero_dev(ctx, ctx->geo_P, nc * ctx->ts * ctx->imax * ctx->ideg);
    zero_dev(ctx, ctx->geo_B, nc * ctx->ideg * ctx->ideg);
    zero_dev(ctx, ctx->geo_M, nc * ctx->ideg);
    zero_dev(ctx, ctx->geo_Q, nc * ctx->nf * ctx->nqm * 3);
    zero_dev(ctx, ctx->geo_area, (size_t)ctx->nface);
    zero_dev(ctx, ctx->geo_nrm, (size_t)ctx->nface * 3);
    zero_dev(ctx, ctx->geo_ovol, nc);
    if ((rc = cells_in(ctx, uN_bar, true, layout, ctx->adj_a)) == UCNS_OK) {
        ctx->geo_on = true;
        rc = rk3_adjoint_impl(ctx, nsteps, 0.0, dt_fixed, ctx->adj_a, t_io);
        ctx->geo_on = false;
    }
    double bc[2 * NV];
    if (rc == UCNS_OK && hipMemcpyAsync(bc, ctx->adj_bc, sizeof(bc), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = UCNS_EHIP;
    if (rc == UCNS_OK) rc = cells_out(ctx, ctx->adj_a, layout, u0_bar, true);
    if (rc == UCNS_OK) rc = geo_to_caller(ctx, gbar);
    release_geometry(ctx);
    if (rc == UCNS_EHIP && ctx->err.empty()) ctx->err = "ucns_rk3_adjoint_geometry: HIP error";
    if (rc) return rc;
    if (bc_bar) std::memcpy(bc_bar, bc, sizeof(bc));
    return UCNS_OK;
}

int ucns_stats(ucns_ctx *ctx, ucns_stats_v1 *out) {
    if (!ctx || !out) return UCNS_EBADARG;
    *out = ctx->st;
    return UCNS_OK;
}

int ucns_debug_fail_alloc(ucns_ctx *ctx, int32_t nth) {
    if (!ctx || nth < 0) return UCNS_EBADARG;
    ctx->fail_alloc_in = nth;
    return UCNS_OK;
}

int ucns_reset_stats(ucns_ctx *ctx) {
    if (!ctx) return UCNS_EBADARG;
    const double b = ctx->st.bytes_recon_per_cell;
    ctx->st = ucns_stats_v1{};
    ctx->st.bytes_recon_per_cell = b;
    ctx->timing = true;
    return UCNS_OK;
}

void ucns_destroy(ucns_ctx *ctx) {
    if (!ctx) return;
    if (!ctx->stream) {  // rejected before any device resource was created
        delete ctx;
        return;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_prestore(ctx);
    release_adjoint(ctx);
    dfree(ctx->cinfo);
    dfree(ctx->cnb);
    dfree(ctx->chnb);
    dfree(ctx->chang);
    dfree(ctx->bface);
    dfree(ctx->cface);
    dfree(ctx->fcell);
    dfree(ctx->fslot);
    dfree(ctx->frot);
    dfree(ctx->farea);
    dfree(ctx->ovol);
    dfree(ctx->minedge);
    dfree(ctx->wquad);
    dfree(ctx->stage);
    dfree(ctx->perm);
    dfree(ctx->U);
    dfree(ctx->U2);
    dfree(ctx->rhs);
    dfree(ctx->fflux);
    dfree(ctx->t);
    dfree(ctx->dtfix);
    dfree(ctx->dtv);
    dfree(ctx->counters);
    dfree(ctx->flag);
    dfree(ctx->uleft);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"
