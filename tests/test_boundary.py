"""Boundary ghost states of the oracle (BOUNDARYS, src/flow_operations.f90:2860-3375)
pinned against what the mathematics fixes: mirror states at symmetry/slip
walls, exact copies for transmissive and supersonic faces, the textbook
characteristic subsonic inflow/outflow states (Blazek, 'Computational Fluid
Dynamics: Principles and Applications', the subsonic inflow and outflow
conditions of the boundary-condition chapter), and mesh-level conservation
and free-stream preservation on bounded boxes."""
import math

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, TET, Case, CaseSpec

from test_oracle import G, phys_flux, prim2cons, random_state


def normal(a1, a2):
    return np.array([math.cos(a1) * math.sin(a2), math.sin(a1) * math.sin(a2), math.cos(a2)])


def prim(U):
    r = U[0]
    u = U[1:4] / r
    p = (G - 1) * (U[4] - 0.5 * r * u @ u)
    return r, u, p


def blazek_inflow(far, interior, n):
    """subsonic inflow: p_b = (p_a + p_d - rho_0 c_0 n.(u_a - u_d))/2,
    rho_b = rho_a + (p_b - p_a)/c_0^2, u_b = u_a - n (p_a - p_b)/(rho_0 c_0);
    a = free stream, d = interior, 0 = interior reference state"""
    ra, ua, pa = prim(far)
    rd, ud, pd = prim(interior)
    c0 = math.sqrt(G * pd / rd)
    pb = 0.5 * (pa + pd - rd * c0 * n @ (ua - ud))
    rb = ra + (pb - pa) / c0 ** 2
    ub = ua - n * (pa - pb) / (rd * c0)
    return prim2cons(rb, *ub, pb)


def blazek_outflow(far, interior, n):
    """subsonic outflow: p_b = p_a, rho_b = rho_d + (p_b - p_d)/c_0^2,
    u_b = u_d + n (p_d - p_b)/(rho_0 c_0)"""
    ra, ua, pa = prim(far)
    rd, ud, pd = prim(interior)
    c0 = math.sqrt(G * pd / rd)
    pb = pa
    rb = rd + (pb - pd) / c0 ** 2
    ub = ud + n * (pd - pb) / (rd * c0)
    return prim2cons(rb, *ub, pb)


@pytest.fixture(autouse=True)
def _reset_boundary():
    oracle.set_boundary(0, prim2cons(1, 0, 0, 0, 1), prim2cons(1, 0, 0, 0, 1))
    yield


def angles(rng):
    return rng.uniform(-math.pi, math.pi), rng.uniform(0.05, math.pi - 0.05)


@pytest.mark.parametrize("code", [3, 4])
def test_symmetry_and_wall_mirror_the_state(code):
    rng = np.random.default_rng(20)
    for _ in range(100):
        a1, a2 = angles(rng)
        n = normal(a1, a2)
        U = random_state(rng, 1.5)
        g = oracle.ghost(code, G, U, a1, a2)
        m = U[1:4]
        np.testing.assert_allclose(g, np.r_[U[0], m - 2 * (m @ n) * n, U[4]], rtol=1e-13, atol=1e-13)
        # the face flux carries no mass and no energy, and pushes along n only
        rl, rr = oracle.rotatef(U, a1, a2), oracle.rotatef(g, a1, a2)
        F = oracle.rotateb(oracle.hllc(G, rl, rr), a1, a2)
        scale = abs(F[1:4]).max() + 1.0
        assert abs(F[0]) < 1e-12 * scale and abs(F[4]) < 1e-12 * scale
        tang = F[1:4] - (F[1:4] @ n) * n
        assert np.abs(tang).max() < 1e-12 * scale


def test_supersonic_type_copies_states():
    rng = np.random.default_rng(21)
    far = prim2cons(1.1, 2.5, 0.1, -0.2, 1.0)
    oracle.set_boundary(0, far, prim2cons(0.9, 0.1, 0, 0, 0.8))
    for _ in range(20):
        a1, a2 = angles(rng)
        U = random_state(rng, 0.5)
        assert np.array_equal(oracle.ghost(1, G, U, a1, a2), far)
        assert np.array_equal(oracle.ghost(2, G, U, a1, a2), U)


def test_characteristic_inflow_and_outflow_match_the_textbook():
    rng = np.random.default_rng(22)
    for _ in range(200):
        far = random_state(rng, 0.6)
        out = random_state(rng, 0.6)
        oracle.set_boundary(1, far, out)
        a1, a2 = angles(rng)
        n = normal(a1, a2)
        U = random_state(rng, 0.5)  # |u| <= 0.5 sqrt(3) c: subsonic interior
        np.testing.assert_allclose(oracle.ghost(1, G, U, a1, a2), blazek_inflow(far, U, n), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(oracle.ghost(2, G, U, a1, a2), blazek_outflow(out, U, n), rtol=1e-12, atol=1e-12)
        # supersonic interior: inflow takes the free stream, outflow the interior
        r, u, p = prim(U)
        c = math.sqrt(G * p / r)
        Us = prim2cons(r, *(u / np.linalg.norm(u) * 1.5 * c), p)
        assert np.array_equal(oracle.ghost(1, G, Us, a1, a2), far)
        assert np.array_equal(oracle.ghost(2, G, Us, a1, a2), Us)


def test_characteristic_states_have_the_free_stream_as_fixed_point():
    rng = np.random.default_rng(23)
    for _ in range(50):
        far = random_state(rng, 0.6)
        oracle.set_boundary(1, far, far)
        a1, a2 = angles(rng)
        for code in (1, 2, 6):
            np.testing.assert_allclose(oracle.ghost(code, G, far, a1, a2), far, rtol=1e-13, atol=1e-13)


def test_farfield_picks_inflow_or_outflow_from_the_normal_momentum():
    """code 6 (src/flow_operations.f90:3203-3296): inflow when the rotated
    normal momentum rho*u.n <= 0, outflow otherwise; supersonic when
    |rho u.n| >= c (the reference compares the momentum, not the velocity,
    with the speed of sound)"""
    rng = np.random.default_rng(24)
    far = prim2cons(1.0, 0.2, 0.0, 0.1, 1.0)
    out = prim2cons(1.0, 0.2, 0.0, 0.1, 0.9)
    oracle.set_boundary(1, far, out)
    for _ in range(100):
        a1, a2 = angles(rng)
        n = normal(a1, a2)
        r = rng.uniform(0.5, 2.0)
        p = rng.uniform(0.5, 2.0)
        c = math.sqrt(G * p / r)
        for un, kind in ((-0.3 * c, "sub_in"), (0.3 * c, "sub_out"), (-2.0 * c / r - 0.1, "sup_in"),
                         (2.0 * c / r + 0.1, "sup_out")):
            t = np.cross(n, [0.3, -0.5, 0.81])
            U = prim2cons(r, *(un * n + 0.2 * t), p)
            g = oracle.ghost(6, G, U, a1, a2)
            want = {"sub_in": lambda: blazek_inflow(far, U, n), "sub_out": lambda: blazek_outflow(out, U, n),
                    "sup_in": lambda: far, "sup_out": lambda: U}[kind]()
            np.testing.assert_allclose(g, want, rtol=1e-12, atol=1e-12, err_msg=kind)


BOXES = {
    "hex-weno4-walls": lambda: CaseSpec("walls", HEX, 12, "weno4", seed=3, bc=(4, 4, 4, 4, 3, 3)),
    "tet-weno4-walls": lambda: CaseSpec("walls", TET, 5, "weno4", jitter=0.1, seed=3, bc=(4, 4, 3, 3, 4, 4)),
    "hex-muscl-walls": lambda: CaseSpec("walls", HEX, 6, "muscl", seed=3, bc=(4, 4, 4, 4, 4, 4)),
}


@pytest.mark.parametrize("name", list(BOXES))
def test_closed_box_conserves_mass_and_energy(name):
    case = Case(BOXES[name]())
    orc = Oracle(case)
    u = case.init_state()
    vol = case.array("volume")[:, None]
    u1, _ = orc.rk3(u, 3, cfl=0.4)
    tot0, tot1 = (u * vol).sum(0), (u1 * vol).sum(0)
    for v in (0, 4):
        assert abs(tot1[v] - tot0[v]) <= 1e-13 * abs(tot0[v]), (v, tot0[v], tot1[v])
    assert np.all(np.isfinite(u1))


@pytest.mark.parametrize("shape,scheme,bc,boundtype", [
    (HEX, "weno4", (2, 2, 2, 2, 2, 2), 0),
    (TET, "weno4", (1, 2, 2, 2, 3, 4), 0),
    (HEX, "weno4", (6, 6, 6, 6, 6, 6), 1),
    (TET, "muscl", (1, 2, 6, 6, 3, 3), 1),
])
def test_uniform_flow_is_preserved(shape, scheme, bc, boundtype):
    """a uniform state equal to the free stream, moving along x (so that the
    symmetry/wall sides, normal to y or z, see no normal velocity), is
    steady with transmissive, supersonic, characteristic and farfield faces"""
    case = Case(CaseSpec("uniform", shape, 8 if shape == HEX else 4, scheme, jitter=0.1 if shape == TET else 0,
                         seed=5, bc=bc))
    orc = Oracle(case)
    mach = 2.0 if boundtype == 0 else 0.5
    U0 = prim2cons(1.0, mach * math.sqrt(G), 0.0, 0.0, 1.0)
    oracle.set_boundary(boundtype, U0, U0)
    u = np.tile(U0, (case.ncell, 1))
    rhs, _ = orc.residual(u)
    scale = np.abs(phys_flux(U0, np.array([1.0, 0, 0]))).max() * case.array("surf").max()
    assert np.abs(rhs).max() <= 1e-12 * scale


def code50_structs(case):
    """the case's structs with the periodic faces whose outward normal is +x
    carrying code 50 (their partners keep 5), as a host numbering periodic
    pairs that way passes them; returns (holder, structs)"""
    m, s, p = case.structs()
    nf = case.nface_max
    bcode = case.array("bcode").copy().reshape(-1, nf)
    normal = case.array("normal").reshape(-1, nf, 3)
    plus_x = (bcode == 5) & (normal[:, :, 0] > 0.9)
    bcode[plus_x] = 50
    bcode = np.ascontiguousarray(bcode.reshape(-1)).astype(np.int32)
    from ucns3d_amd import native
    m.bcode = native.i32p(bcode)
    return (bcode, m, s, p), (m, s, p), int(plus_x.sum())


def test_code50_neighbours_stay_out_of_the_muscl_bounds():
    """FIND_BOUNDS with strict bounds (src/reconstruct_p.f90:1847-1885)
    takes, for a cell with boundary faces, the interior neighbours and the
    periodic neighbours of code 5 (1871), not those of code 50.  On a
    periodic MUSCL box whose x+ periodic faces carry 50: the limited face
    values of those cells stay within the bounds without the x+ neighbour
    (and some leave the bounds that include it), the fluxes still pair
    every face (sum of residuals 0), and the other cells are unchanged."""
    case = Case(CaseSpec("muscl-50", HEX, 6, "muscl", seed=5))
    orc5 = Oracle(case)
    holder, (m, s, p), n50 = code50_structs(case)
    assert n50 == 36
    orc50 = Oracle(case)
    orc50.m = m
    rng = np.random.default_rng(4)
    u = np.stack([prim2cons(1 + 0.3 * rng.uniform(), 0.1, 0.2, -0.1, 1 + 0.3 * rng.uniform())
                  for _ in range(case.ncell)])
    ul5, _ = orc5.reconstruct(u)
    ul50, _ = orc50.reconstruct(u)
    nf = 6
    bcode = holder[0].reshape(-1, nf)
    ineigh = case.array("ineigh").reshape(-1, nf)
    has50 = (bcode == 50).any(1)
    np.testing.assert_array_equal(ul5[~has50], ul50[~has50])
    q = ul50[:, :nf, :case.qp_quad, :]
    keep = (bcode == 0) | (bcode == 5)
    nb = np.where(keep, ineigh, np.arange(case.ncell)[:, None])
    lo = np.minimum(u[:, None, :], u[nb]).min(1)
    hi = np.maximum(u[:, None, :], u[nb]).max(1)
    tol = 1e-12 * np.abs(u).max(0)
    assert np.all(q >= lo[:, None, None, :] - tol) and np.all(q <= hi[:, None, None, :] + tol)
    assert not np.array_equal(ul5[has50], ul50[has50])
    r50, _ = orc50.residual(u)
    assert np.all(np.abs(r50.sum(0)) <= 1e-12 * np.abs(r50).sum(0))


def face_centres(case):
    """mean of each face record's nodes (CORDINATES3, src/grid_t.f90:953-967),
    [ncell][nface_max][3]; NaN where a cell has no such face"""
    fn = case.array("face_nodes").reshape(case.ncell, case.nface_max, 4)
    xyz = case.array("xyz").reshape(-1, 3)
    pts = np.where((fn >= 0)[..., None], xyz[np.clip(fn, 0, None)], 0.0)
    cnt = (fn >= 0).sum(2)[..., None]
    return np.where(cnt > 0, pts.sum(2) / np.maximum(cnt, 1), np.nan)


@pytest.mark.parametrize("code", [1, 2])
def test_per_face_boundary_states_reach_their_own_faces(code):
    """per-face INFLOW()/OUTFLOW() states (position-dependent, evaluated at the
    face centres): every cell's residual equals the one computed with the
    uniform state of its own boundary faces' region, bit for bit, and cells
    without such faces are untouched"""
    case = Case(CaseSpec("jet", HEX, 8, "weno3", jitter=0.1, seed=6, bc=(1, 2, 4, 4, 3, 3)))
    orc = Oracle(case)
    u = case.init_state()
    SA = prim2cons(1.0, 0.5, 0.0, 0.0, 100.0 / 1.4)
    SB = prim2cons(1.3, 0.8, 0.1, -0.1, 90.0 / 1.4)
    c = face_centres(case)
    inside = np.linalg.norm(c[..., 1:] - 0.5 * case.spec.length, axis=-1) < 0.3 * case.spec.length
    per_face = np.where(inside[..., None], SB, SA)
    bcode = case.array("bcode").reshape(case.ncell, case.nface_max)
    on = bcode == code
    assert on[inside].sum() > 0 and on[~inside & (bcode >= 0)].sum() > 0
    oracle.set_boundary(1, SA, SA)
    r_a, _ = orc.residual(u)
    oracle.set_boundary(1, SB, SB)
    r_b, _ = orc.residual(u)
    oracle.set_boundary(1, SA, SA)
    try:
        if code == 1:
            oracle.set_boundary_faces(per_face, None)
        else:
            oracle.set_boundary_faces(None, per_face)
        r, _ = orc.residual(u)
    finally:
        oracle.set_boundary_faces(None, None)
    in_a = (on & ~inside).any(1)
    in_b = (on & inside).any(1)
    assert np.array_equal(r[in_b & ~in_a], r_b[in_b & ~in_a])
    assert np.array_equal(r[~in_b], r_a[~in_b])
    assert (in_b & ~in_a).sum() > 0 and (in_a & ~in_b).sum() > 0
    assert not np.array_equal(r_a[in_b], r_b[in_b])  # the regions' states matter
