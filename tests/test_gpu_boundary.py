"""GPU parity on bounded boxes: boundary faces with ghost states (BOUNDARYS,
src/flow_operations.f90:2860-3375) through the C ABI against the CPU oracle.
Same bar as the periodic cases: FP64 relative L2 per conserved variable
<= 1e-12 after one SSP-RK3 step, <= 1e-9 after 20 steps; mixed <= 1e-5."""
import math

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec
from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED
from ucns3d_amd.solver import Solver

pytestmark = pytest.mark.gpu

G = 1.4


def prim2cons(r, u, v, w, p):
    return np.array([r, r * u, r * v, r * w, p / (G - 1) + 0.5 * r * (u * u + v * v + w * w)])


P0 = 100.0 / 1.4  # pressure level of the vortex field
BOXES = {
    # subsonic channel: characteristic inflow/outflow, slip walls, symmetry
    "hex-weno4-channel": (CaseSpec("channel", HEX, 12, "weno4", seed=3, bc=(1, 2, 4, 4, 3, 3)), 1,
                          prim2cons(1.0, 0.5, 0.0, 0.0, P0), prim2cons(1.0, 0.5, 0.0, 0.0, P0)),
    # farfield on every side
    "tet-weno4-farfield": (CaseSpec("farfield", TET, 6, "weno4", jitter=0.1, seed=4, bc=(6,) * 6), 1,
                           prim2cons(1.0, 0.3, 0.1, 0.0, P0), prim2cons(1.0, 0.3, 0.1, 0.0, P0)),
    # supersonic inflow, transmissive exits, walls
    "tet-weno4-supersonic": (CaseSpec("supersonic", TET, 5, "weno4", jitter=0.1, seed=5, bc=(1, 2, 4, 4, 2, 2)), 0,
                             prim2cons(1.0, 20.0, 0.0, 0.0, P0), prim2cons(1.0, 20.0, 0.0, 0.0, P0)),
    # closed box, MUSCL
    "hex-muscl-walls": (CaseSpec("walls", HEX, 8, "muscl", seed=6, bc=(4,) * 6), 0,
                        prim2cons(1, 0, 0, 0, P0), prim2cons(1, 0, 0, 0, P0)),
    # the adjoint channel's construction: hexahedra, pyramids and prisms, a lower-wall bump
    # (a_k ~ U(0, 0.01)), WENO-3, subsonic inflow/outflow and slip walls
    "hybrid-weno3-bump": (CaseSpec("bump", HYBRID, 6, "weno3", seed=7, bc=(1, 2, 4, 4, 3, 3),
                                   bump=tuple(np.random.default_rng(22).uniform(0, 0.01, 8))), 1,
                          prim2cons(1.0, 0.6 * math.sqrt(G * P0), 0.0, 0.0, P0),
                          prim2cons(1.0, 0.6 * math.sqrt(G * P0), 0.0, 0.0, P0)),
}
_cache = {}


def get(name):
    if name not in _cache:
        spec, bt, fin, fout = BOXES[name]
        case = Case(spec)
        u = case.init_state()
        if name == "tet-weno4-supersonic":
            u = u.copy()
            u[:, 4] += 0.5 * 400.0 * u[:, 0]  # add x-velocity 20 to the vortex
            u[:, 1] += 20.0 * u[:, 0]
        _cache[name] = (case, Oracle(case), u, bt, fin, fout)
    return _cache[name]


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1, 5)
    b = np.asarray(b).reshape(-1, 5)
    den = np.sqrt((b ** 2).sum(axis=0))
    return np.sqrt(((a - b) ** 2).sum(axis=0)) / np.where(den > 0, den, 1.0)


def solver(name, precision):
    case, orc, u, bt, fin, fout = get(name)
    s = Solver(case, precision)
    s.set_boundary(bt, fin, fout)
    s.set_state(u)
    oracle.set_boundary(bt, fin, fout)
    return s


def test_boundary_data_is_required():
    case, orc, u, *_ = get("hex-weno4-channel")
    s = Solver(case, UCNS_FP64)
    s.set_state(u)
    with pytest.raises(RuntimeError, match="ucns_set_boundary"):
        s.residual()
    s.close()


@pytest.mark.parametrize("name", list(BOXES))
def test_residual_fp64(name):
    case, orc, u, *_ = get(name)
    print(f"[bnd residual] {name} ncell={case.ncell}", flush=True)
    s = solver(name, UCNS_FP64)
    r = s.residual()
    st = s.stats()
    r_ref, cnt = orc.residual(u)
    err = rel_l2(r, r_ref)
    print("  rel L2", err, "muscl cells", st["n_reduced_last"], cnt[0])
    assert st["n_reduced_last"] == cnt[0] and st["n_firstorder_last"] == cnt[1]
    assert np.all(err <= 1e-11), err
    s.close()


@pytest.mark.parametrize("name", list(BOXES))
def test_steps_fp64(name):
    case, orc, u, *_ = get(name)
    print(f"[bnd rk3] {name}", flush=True)
    s = solver(name, UCNS_FP64)
    s.rk3_steps(1, cfl=0.4)
    u1 = s.get_state()
    u1_ref, _ = orc.rk3(u, 1, cfl=0.4)
    e1 = rel_l2(u1, u1_ref)
    s.rk3_steps(19, cfl=0.4)
    u20 = s.get_state()
    u20_ref, _ = orc.rk3(u1_ref, 19, cfl=0.4)
    e20 = rel_l2(u20, u20_ref)
    print("  rel L2 step1", e1, "step20", e20)
    assert np.all(e1 <= 1e-12), e1
    assert np.all(e20 <= 1e-9), e20
    s.close()


@pytest.mark.parametrize("name", ["hex-weno4-channel", "tet-weno4-farfield"])
def test_mixed_vs_fp64(name):
    case, orc, u, *_ = get(name)
    print(f"[bnd mixed] {name}", flush=True)
    s = solver(name, UCNS_MIXED)
    s.rk3_steps(1, cfl=0.4)
    u1 = s.get_state()
    u1_ref, _ = orc.rk3(u, 1, cfl=0.4)
    err = rel_l2(u1, u1_ref)
    print("  rel L2 mixed vs FP64 oracle", err)
    assert np.all(err <= 1e-5), err
    s.close()


@pytest.mark.parametrize("scheme", ["muscl", "weno4"])
def test_code50_periodic_faces(scheme):
    """periodic faces numbered 5 on one side and 50 on the other (the x+
    side here): code-50 neighbours stay out of the MUSCL bounds
    (src/reconstruct_p.f90:1871) while both codes pair their fluxes; GPU
    against the oracle given the same structs, one step <= 1e-12"""
    from test_boundary import code50_structs
    case = Case(CaseSpec(f"{scheme}-50", HEX, 8 if scheme == "muscl" else 12, scheme, seed=5))
    holder, st, n50 = code50_structs(case)
    print(f"[code 50] {scheme} ncell={case.ncell} faces={n50}", flush=True)
    orc = Oracle(case)
    orc.m = st[0]
    u = case.init_state()
    x = case.array("centre").reshape(-1, 3)
    u = u * (1 + 0.05 * np.sin(0.37 * x[:, :1] + 0.71 * x[:, 1:2] + 0.53 * x[:, 2:3] + 0.2))
    s = Solver(case, UCNS_FP64, structs=st)
    s.set_state(u)
    ul = s.reconstruct()
    ul_ref, cnt = orc.reconstruct(u)
    st_gpu = s.stats()
    assert st_gpu["n_reduced_last"] == cnt[0]
    assert scheme != "muscl" or cnt[0] == case.ncell
    s.rk3_steps(1, cfl=0.5)
    u1_ref, _ = orc.rk3(u, 1, cfl=0.5)
    err = np.sqrt(((s.get_state() - u1_ref) ** 2).sum(0)) / np.sqrt((u1_ref ** 2).sum(0))
    print("  rel L2", err)
    assert np.all(err <= 1e-12)
    m = np.isfinite(ul_ref)
    assert np.abs(ul[m] - ul_ref[m]).max() <= 1e-12 * np.abs(ul_ref[m]).max()
    s.close()


def jet_case(wenwrt=1):
    """subsonic box with a jet in the inflow side: the INFLOW() state differs on
    the faces whose centres lie within 0.3 L of the side's centre"""
    from test_boundary import face_centres
    case = Case(CaseSpec("jet", HEX, 12, "weno4", jitter=0.1, seed=9, bc=(1, 2, 4, 4, 3, 3), wenwrt=wenwrt))
    SA = prim2cons(1.0, 0.5, 0.0, 0.0, P0)
    SB = prim2cons(1.4, 2.0, 0.2, -0.1, 0.9 * P0)
    c = face_centres(case)
    inside = np.linalg.norm(c[..., 1:] - 0.5 * case.spec.length, axis=-1) < 0.3 * case.spec.length
    return case, SA, np.where(inside[..., None], SB, SA)


def test_uniform_per_face_states_give_the_same_bits():
    case, SA, _ = jet_case()
    u = case.init_state()
    print(f"[per-face boundary] uniform arrays ncell={case.ncell}", flush=True)
    s = Solver(case, UCNS_FP64)
    s.set_boundary(1, SA, SA)
    s.set_state(u)
    r0 = s.residual()
    s.set_boundary_faces(np.broadcast_to(SA, (case.ncell, case.nface_max, 5)),
                         np.broadcast_to(SA, (case.ncell, case.nface_max, 5)))
    r1 = s.residual()
    s.close()
    assert np.array_equal(r0, r1)


@pytest.mark.parametrize("wenwrt", [1, 2])
def test_jet_inflow_parity(wenwrt):
    case, SA, per_face = jet_case(wenwrt)
    u = case.init_state()
    print(f"[per-face boundary] jet wenwrt={wenwrt} ncell={case.ncell}", flush=True)
    orc = Oracle(case)
    oracle.set_boundary(1, SA, SA)
    oracle.set_boundary_faces(per_face, None)
    try:
        s = Solver(case, UCNS_FP64)
        s.set_boundary(1, SA, SA)
        s.set_boundary_faces(per_face, None)
        s.set_state(u)
        r = s.residual()
        r_ref, _ = orc.residual(u)
        err = rel_l2(r, r_ref)
        s.rk3_steps(10, cfl=0.5)
        u10 = s.get_state()
        s.close()
        u10_ref, _ = orc.rk3(u, 10, cfl=0.5)
        err10 = rel_l2(u10, u10_ref)
        # the jet changed the flow: against the uniform inflow the states differ
        oracle.set_boundary_faces(None, None)
        u10_uni, _ = orc.rk3(u, 10, cfl=0.5)
    finally:
        oracle.set_boundary_faces(None, None)
    print("  residual rel L2", err, "10 steps", err10, "jet effect", rel_l2(u10_uni, u10_ref))
    assert np.all(err <= 1e-11), err
    assert np.all(err10 <= 1e-9), err10
    assert rel_l2(u10_uni, u10_ref).max() > 1e-4
