"""Geometry cotangents of the RK3 adjoint (ucns_rk3_adjoint_geometry): the
derivative of J = <c, U_N> after N steps of fixed dt with respect to each
array of mesh/PRESTORE data the march reads, checked array by array against
central differences of the oracle's march with that array perturbed in a
random direction (the same discrete decisions on both sides at h small).

Face data (areas, normals) are perturbed on both records of a face alike, so
that the face keeps one geometry; the library reads one record per face and
returns its cotangent there.  The march starts from a smooth non-uniform
state: from a uniform one, an area changed alone opens the cells (the face
normals times areas no longer sum to zero), every MUSCL cell then limits a
ratio of two differences of the size of the change, and J is not
differentiable there (its one-sided differences disagree at every h).  A normal is moved tangentially and its angles
(FACEANGLEX/Y) recomputed as the preprocessor forms them."""
import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, TET, Case, CaseSpec, reduced
from ucns3d_amd.native import UCNS_FP64, UCNS_KERNELS_REGISTER
from ucns3d_amd.solver import Solver

pytestmark = pytest.mark.gpu

# (spec, steps, initial state: "smooth" perturbed or the case's own)
CASES = {
    # prisms, pyramids and tetrahedra; inflow, outflow and slip walls; WENO-3
    # with MUSCL cells at the walls
    "c4-reduced": (lambda: reduced("c4", 3), 4, "smooth"),
    # jittered periodic tetrahedra, WENO-4
    "tet-weno4": (lambda: CaseSpec("tet-weno4", TET, 4, "weno4", jitter=0.1, seed=7), 4, "smooth"),
    # periodic hexahedra, MUSCL + Barth (the limiter's active points)
    "hex-muscl": (lambda: CaseSpec("hex-muscl", HEX, 6, "muscl", seed=5), 4, "smooth"),
}


def normal_angles(n):
    """FACEANGLEX/Y of unit normals [..., 3] as the preprocessor forms them
    (src/grid_t.f90:4441-4508)"""
    nx, ny, nz = n[..., 0], n[..., 1], n[..., 2]
    a1 = np.zeros(nx.shape)
    q = np.where((nx != 0) & (ny != 0), np.arctan(np.divide(ny, nx, out=np.zeros_like(ny), where=nx != 0)), 0.0)
    a1 = np.where((nx > 0) & (ny > 0), q, a1)
    a1 = np.where((nx < 0) & (ny != 0), np.pi + q, a1)
    a1 = np.where((nx > 0) & (ny < 0), 2 * np.pi + q, a1)
    a1 = np.where((nx == 0) & (ny > 0), np.pi / 2, a1)
    a1 = np.where((nx == 0) & (ny < 0), 3 * np.pi / 2, a1)
    a1 = np.where((nx < 0) & (ny == 0), np.pi, a1)
    a2 = np.arccos(np.where(nz == 0, 0.0, nz))
    return a1, a2


def smooth_state(case):
    """the case's initial state with a smooth, asymmetric perturbation of
    density, velocity and pressure (a few percent of the sound speed)"""
    g = case.spec.gamma
    u = case.init_state()
    x = case.array("centre").reshape(-1, 3)
    L = np.array(case.spec.dims[3:]) if case.spec.box else np.full(3, case.spec.length)
    k = 2 * np.pi / L
    ph = x * k
    rho = u[:, 0]
    vel = u[:, 1:4] / rho[:, None]
    p = (g - 1) * (u[:, 4] - 0.5 * rho * (vel ** 2).sum(1))
    c = np.sqrt(g * p / rho)
    rho2 = rho * (1 + 0.02 * np.sin(ph[:, 0] + 0.7 * ph[:, 1] + 0.3 * ph[:, 2] + 0.4))
    dv = np.stack([np.sin(0.5 * ph[:, 0] + ph[:, 1] + 0.2), np.cos(ph[:, 1] + 0.6 * ph[:, 2] + 0.9),
                   np.sin(0.8 * ph[:, 2] + 0.3 * ph[:, 0] + 1.3)], 1)
    vel2 = vel + 0.03 * c[:, None] * dv
    p2 = p * (1 + 0.02 * np.cos(0.6 * ph[:, 0] + 0.9 * ph[:, 2] + 0.1))
    out = np.empty_like(u)
    out[:, 0] = rho2
    out[:, 1:4] = rho2[:, None] * vel2
    out[:, 4] = p2 / (g - 1) + 0.5 * rho2 * (vel2 ** 2).sum(1)
    return out


class Setup:
    def __init__(self, key):
        make, self.nsteps, state = CASES[key]
        spec = make()
        self.case = case = Case(spec)
        self.fs = spec.freestream() if spec.mach > 0 else None
        if self.fs is not None:
            oracle.set_boundary(1, self.fs, self.fs)
        rng = np.random.default_rng(11)
        # a smooth state without the initial data's symmetries and uniform
        # regions: the Barth limiter is not differentiable where a cell's
        # neighbours all hold its own value (a ratio of vanishing differences)
        # or where two face points tie for its minimum, and a random field
        # would put such switches everywhere
        self.u0 = smooth_state(case) if state == "smooth" else case.init_state()
        self.orc = Oracle(case)
        self.dt = 0.5 * self.orc.dt(self.u0, spec.cfl)
        self.c = rng.standard_normal((case.ncell, 5))
        self.rng = rng

    def J(self):
        u, _ = self.orc.rk3(self.u0, self.nsteps, dt_fixed=self.dt)
        return float((self.c * u).sum())

    def adjoint(self):
        s = Solver(self.case, UCNS_FP64, kernels=UCNS_KERNELS_REGISTER)
        if self.fs is not None:
            s.set_boundary(1, self.fs, self.fs)
        s.set_state(self.u0)
        _, _, g, _ = s.rk3_adjoint_geometry(self.nsteps, self.c, self.dt)
        s.close()
        return g


_setups = {}


def setup(key):
    if key not in _setups:
        _setups[key] = Setup(key)
        _setups[key].g = _setups[key].adjoint()
    return _setups[key]


def central(S, arrays, apply, h):
    """(J(x + h d) - J(x - h d)) / 2h: apply(t) moves the case's arrays (in
    place) to x + t d; they are restored afterwards"""
    base = [a.copy() for a in arrays]
    out = []
    try:
        for t in (h, -h):
            for a, b in zip(arrays, base):
                a[:] = b
            apply(t)
            out.append(S.J())
    finally:
        for a, b in zip(arrays, base):
            a[:] = b
    return (out[0] - out[1]) / (2 * h)


def check(name, fd, an, scale, tol):
    """|fd - an| <= tol * scale, scale the directional derivative's own size
    (the sum of |terms|)"""
    err = abs(fd - an) / scale
    print(f"  {name:24s} fd {fd: .12e} adjoint {an: .12e} err/scale {err:.2e}", flush=True)
    assert err <= tol, (name, fd, an, scale)


def directions(S, gbar, a, mask=None):
    """two directions for one array: along the adjoint's own gradient
    (scaled by the entries, so the largest change is h times its entry) and
    a random one of entry-relative size"""
    w = np.abs(a) if mask is None else np.where(mask, np.abs(a), 0.0)
    grad = gbar * w * w
    grad /= max(np.abs(gbar * w).max(), 1e-300)
    rnd = S.rng.standard_normal(a.shape) * w
    return {"gradient": grad, "random": rnd}


# bars: along the gradient the derivative is the squared norm and the
# differences resolve it to 1e-6; a random direction's derivative is a sum of
# terms of both signs that cancel, so it is held to 1e-5 of their sum of
# magnitudes (central differences at h = 1e-7 in the oracle's rounding)
TOL = {"gradient": 1e-6, "random": 1e-5}


@pytest.mark.parametrize("key", list(CASES))
def test_cell_data_cotangents(key):
    S = setup(key)
    case, g = S.case, S.g
    print(f"[geometry cotangents] {key} ncell={case.ncell} steps={S.nsteps}", flush=True)
    names = ["invmat", "integ_basis", "volume11", "qpoints", "volume"] + (["indicator"] if case.iweno == 1 else [])
    for name in names:
        a = case.array(name)
        gb = g["totvolume" if name == "volume" else name].reshape(-1)
        for kind, d in directions(S, gb, a).items():
            if name == "indicator":  # symmetric directions: the matrix stays symmetric
                dd = d.reshape(case.ncell, case.ideg, case.ideg)
                d = (dd + dd.transpose(0, 2, 1)).reshape(-1) * 0.5
            base = a.copy()
            fd = central(S, [a], lambda t: a.__setitem__(slice(None), base + t * d), 1e-7)
            check(f"{name} ({kind})", fd, float((gb * d).sum()), float(np.abs(gb * d).sum()), TOL[kind])


@pytest.mark.parametrize("key", list(CASES))
def test_face_data_cotangents(key):
    S = setup(key)
    case, g = S.case, S.g
    nf = case.nface_max
    surf = case.array("surf")
    ax, ay = case.array("anglex"), case.array("angley")
    normal = case.array("normal").reshape(case.ncell, nf, 3)
    ineigh = case.array("ineigh").reshape(case.ncell, nf)
    ineighn = case.array("ineighn").reshape(case.ncell, nf)
    nfc = (case.array("face_nodes").reshape(case.ncell, nf, 4)[:, :, 0] >= 0).sum(1)
    # the record of each face the library reads: the lower cell index (the
    # lower slot when a cell meets itself across the period), or the only one
    i = np.repeat(np.arange(case.ncell), nf).reshape(case.ncell, nf)
    f = np.tile(np.arange(nf), case.ncell).reshape(case.ncell, nf)
    live = f < nfc[:, None]
    own = live & ((ineigh < 0) | (i < ineigh) | ((i == ineigh) & (f < ineighn)))
    assert np.all(g["surf"][live & ~own] == 0.0) and np.all(g["normal"][live & ~own] == 0.0)
    cells, slots = np.nonzero(own)
    pj, pf = ineigh[cells, slots], ineighn[cells, slots]
    q = pj >= 0
    print(f"[geometry cotangents] {key}: {len(cells)} faces, {int((~q).sum())} on the boundary", flush=True)
    s2 = surf.reshape(case.ncell, nf)
    gs = g["surf"][cells, slots]
    for kind, d in directions(S, gs, s2[cells, slots]).items():
        def set_area(t):
            s2[cells, slots] += t * d
            s2[pj[q], pf[q]] += t * d[q]

        fd = central(S, [surf], set_area, 1e-7)
        check(f"surf ({kind})", fd, float((gs * d).sum()), float(np.abs(gs * d).sum()), TOL[kind])
    # normals: tangential directions, the angles of both records recomputed;
    # not on faces near n = +-z, where the angles' own rounding (a2 from
    # acos(n_z)) swamps a change of 1e-7
    n = normal[cells, slots]
    gn = g["normal"][cells, slots]
    # the normal cotangent is tangential to the unit sphere
    assert np.abs((gn * n).sum(1)).max() <= 1e-9 * np.abs(gn).max()
    ax2, ay2 = ax.reshape(case.ncell, nf), ay.reshape(case.ncell, nf)
    far = (np.abs(n[:, 2]) <= 0.9)[:, None]
    for kind, d in directions(S, gn, np.ones_like(n), np.broadcast_to(far, n.shape)).items():
        d = d - (d * n).sum(1, keepdims=True) * n

        def set_normal(t):
            m = n + t * d
            m /= np.linalg.norm(m, axis=1, keepdims=True)
            ax2[cells, slots], ay2[cells, slots] = normal_angles(m)
            ax2[pj[q], pf[q]], ay2[pj[q], pf[q]] = normal_angles(-m[q])

        fd = central(S, [ax, ay], set_normal, 1e-7)
        check(f"normal ({kind})", fd, float((gn * d).sum()), float(np.abs(gn * d).sum()), TOL[kind])
