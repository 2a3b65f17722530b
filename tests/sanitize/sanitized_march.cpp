// The host preprocessor and the oracle run under AddressSanitizer and
// UndefinedBehaviorSanitizer: each mesh construction (hexahedra,
// tetrahedra, prisms, pyramids, the hybrid, the bumped channel, the blast
// box), each scheme (MUSCL, WENO-3/4/5, CWENO, CWENO-Z, characteristic WENO)
// and each boundary kind at a small size, through PRESTORE, the
// reconstruction, the residual, RK3 steps, the tangent-linear residual, a
// node move and the mesh and restart files.  Exit status 0 when every call
// returned UCNS_OK and every state stayed finite; a sanitizer finding aborts.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ucns.h"
#include "ucns_prep.h"

extern "C" {
void oracle_set_boundary(const ucns_boundary_v1 *bc);
void oracle_set_cweno(const ucns_cweno_v1 *cw);
int oracle_reconstruct(const ucns_mesh_v2 *, const ucns_scheme_v2 *, const ucns_prestore_v1 *, const double *u,
                       double *ul, int64_t *counters);
int oracle_residual(const ucns_mesh_v2 *, const ucns_scheme_v2 *, const ucns_prestore_v1 *, const double *u,
                    double *rhs, int64_t *counters);
int oracle_rk3_steps(const ucns_mesh_v2 *, const ucns_scheme_v2 *, const ucns_prestore_v1 *, double *u,
                     int32_t nsteps, double cfl, double dt_fixed, double *t_io);
int oracle_residual_tl(const ucns_mesh_v2 *, const ucns_scheme_v2 *, const ucns_prestore_v1 *, const double *u,
                       const double *du, const double *dbc, double *rhs, double *drhs);
}

namespace {

struct Run {
    const char *name;
    int32_t shape, n, iorder, ees, wenoz, wenwrt;
    int32_t bc[6];
    double bump0;
};

bool finite(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int fail(const char *name, const char *what, int rc) {
    std::fprintf(stderr, "%s: %s failed (%d): %s\n", name, what, rc, ucnsp_last_error());
    return 1;
}

int run(const Run &r, const std::string &dir) {
    ucnsp_opts o;
    std::memset(&o, 0, sizeof(o));
    o.shape = r.shape;
    o.nx = r.shape == UCNSP_CHANNEL ? 4 * r.n : r.n;
    o.ny = o.nz = r.n;
    o.lx = r.shape == UCNSP_CHANNEL ? 4.0 : 1.0;
    o.ly = o.lz = 1.0;
    o.jitter = r.shape == UCNSP_CHANNEL || r.shape == UCNSP_BLAST ? 0.0 : 0.1;
    o.seed = 7;
    o.iorder = r.iorder;
    o.iweno = r.iorder == 1 ? -1 : 1;
    o.extf = r.iorder == 1 ? 3 : 2;
    o.nthreads = 2;
    std::memcpy(o.bc, r.bc, sizeof(o.bc));
    o.bump[0] = r.bump0;
    o.ees = r.ees;
    ucnsp_case *c = nullptr;
    int rc = ucnsp_build(&o, &c);
    if (rc) return fail(r.name, "ucnsp_build", rc);
    int64_t in[16];
    double db[8];
    ucnsp_info(c, in, db);
    const int64_t ncell = in[0], nf = in[2], nq = in[12];
    ucns_mesh_v2 m;
    ucns_scheme_v2 s;
    ucns_prestore_v1 p;
    if ((rc = ucnsp_fill(c, UCNS_FP64, 1.4, 10000.0, &m, &s, &p))) return fail(r.name, "ucnsp_fill", rc);
    s.wenoz = r.wenoz;
    s.wenwrt = r.wenwrt;
    ucns_cweno_v1 cw;
    if (r.ees == 5) {
        if ((rc = ucnsp_fill_cweno(c, &cw))) return fail(r.name, "ucnsp_fill_cweno", rc);
        oracle_set_cweno(&cw);
    } else {
        oracle_set_cweno(nullptr);
    }
    ucns_boundary_v1 bc;
    std::memset(&bc, 0, sizeof(bc));
    bc.abi_version = UCNS_ABI_VERSION;
    bc.boundtype = 1;
    const double fs[5] = {1.0, 0.6 * std::sqrt(1.4), 0.0, 0.0, 1.0 / 0.4 + 0.5 * 0.36 * 1.4};
    std::memcpy(bc.inflow, fs, sizeof(fs));
    std::memcpy(bc.outflow, fs, sizeof(fs));
    oracle_set_boundary(&bc);

    std::vector<double> u(ncell * 5), rhs(ncell * 5), ul(ncell * nf * nq * 5), du(ncell * 5), drhs(ncell * 5);
    if ((rc = ucnsp_init_vortex(c, 1.4, 0.01, 3, u.data()))) return fail(r.name, "ucnsp_init_vortex", rc);
    for (size_t e = 0; e < du.size(); ++e) du[e] = std::sin(0.37 * (double)e);
    int64_t cnt[2] = {0, 0};
    if ((rc = oracle_reconstruct(&m, &s, &p, u.data(), ul.data(), cnt))) return fail(r.name, "reconstruct", rc);
    if ((rc = oracle_residual(&m, &s, &p, u.data(), rhs.data(), cnt))) return fail(r.name, "residual", rc);
    if ((rc = oracle_residual_tl(&m, &s, &p, u.data(), du.data(), nullptr, rhs.data(), drhs.data())))
        return fail(r.name, "residual_tl", rc);
    double t = 0.0;
    if ((rc = oracle_rk3_steps(&m, &s, &p, u.data(), 2, 0.5, 0.0, &t))) return fail(r.name, "rk3", rc);
    if (!finite(rhs) || !finite(drhs) || !finite(u)) {
        std::fprintf(stderr, "%s: non-finite result\n", r.name);
        return 1;
    }
    // the mesh and restart files, read back into a second case
    const std::string d = dir + "/" + r.name;
    if (std::system(("mkdir -p '" + d + "'").c_str()) != 0) return fail(r.name, "mkdir", -1);
    for (int binary = 0; binary < 2; ++binary) {
        if ((rc = ucnsp_write_grid(c, d.c_str(), binary))) return fail(r.name, "ucnsp_write_grid", rc);
        ucnsp_case *b = nullptr;
        if ((rc = ucnsp_read_grid(d.c_str(), binary, &o, &b))) return fail(r.name, "ucnsp_read_grid", rc);
        int64_t ib[16];
        ucnsp_info(b, ib, nullptr);
        if (ib[0] != ncell) return fail(r.name, "cell count of the read mesh", -1);
        ucnsp_free(b);
    }
    const std::string rp = d + "/RESTART.dat";
    std::vector<double> v(u.size());
    int32_t it = 0;
    if ((rc = ucnsp_write_restart(rp.c_str(), 2, t, u.data(), ncell))) return fail(r.name, "write_restart", rc);
    if ((rc = ucnsp_read_restart(rp.c_str(), ncell, &it, &t, v.data()))) return fail(r.name, "read_restart", rc);
    if (it != 2 || v != u) return fail(r.name, "restart round trip", -1);
    // nodes moved onto themselves: geometry and PRESTORE rebuilt in place
    void *ptr = nullptr;
    int64_t count = 0;
    int32_t dtype = 0;
    if ((rc = ucnsp_array(c, "xyz", &ptr, &count, &dtype))) return fail(r.name, "ucnsp_array", rc);
    std::vector<double> xyz((const double *)ptr, (const double *)ptr + count);
    if ((rc = ucnsp_move_nodes(c, xyz.data()))) return fail(r.name, "ucnsp_move_nodes", rc);
    if ((rc = oracle_residual(&m, &s, &p, v.data(), rhs.data(), cnt))) return fail(r.name, "residual after move", rc);
    ucnsp_free(c);
    std::printf("%-22s %6lld cells ok\n", r.name, (long long)ncell);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    const Run runs[] = {
        {"hex-muscl-walls", UCNSP_HEX, 6, 1, 0, 0, 1, {1, 2, 4, 4, 3, 3}, 0.0},
        {"tet-weno4", UCNSP_TET, 4, 3, 0, 0, 1, {0, 0, 0, 0, 0, 0}, 0.0},
        {"prism-weno3", UCNSP_PRISM, 5, 2, 0, 0, 1, {0, 0, 0, 0, 0, 0}, 0.0},
        {"pyramid-weno3", UCNSP_PYRAMID, 4, 2, 0, 0, 1, {0, 0, 6, 6, 0, 0}, 0.0},
        {"hybrid-weno5-cweno-z", UCNSP_HYBRID, 6, 4, 5, 1, 1, {0, 0, 0, 0, 0, 0}, 0.0},
        {"hex-weno4-char", UCNSP_HEX, 6, 3, 0, 0, 2, {1, 2, 4, 4, 3, 3}, 0.0},
        {"channel-weno3", UCNSP_CHANNEL, 3, 2, 0, 0, 1, {1, 2, 4, 4, 4, 4}, 0.05},
        {"blast-weno3-cweno", UCNSP_BLAST, 8, 2, 5, 0, 1, {2, 2, 2, 2, 2, 2}, 0.0},
    };
    int bad = 0;
    for (const Run &r : runs) bad += run(r, argv[1]);
    return bad ? 1 : 0;
}
