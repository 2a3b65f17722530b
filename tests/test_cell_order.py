"""Cell order across the ABI.

A host hands its cells over in whatever order its partitioner left them
(src/main.f90:103-137); the library may keep its own order inside
(UCNS_ORDER_RCM, UCNS_ORDER_MORTON) and permutes on the way in and out.

CPU: the shuffle helper is pinned by the oracle, whose cell loop makes each
cell's residual independent of the numbering, so a shuffled caller order
gives bitwise the permuted residual.
GPU: the library with a shuffled caller order and each internal order
matches the unshuffled run (every entry point that takes or returns cell
data), and the internal order restores the tile locality a random order
destroys."""
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec, reduced
from ucns3d_amd.native import (UCNS_FP64, UCNS_LAYOUT_SOA, UCNS_MIXED, UCNS_ORDER_CALLER, UCNS_ORDER_MORTON,
                               UCNS_ORDER_RCM)

from shuffle import shuffled_structs
from test_tangent import prim2cons

CASES = {
    "tet-weno4": lambda: reduced("c2", 5),
    "hex-weno4": lambda: reduced("c3", 10),
    "hex-muscl": lambda: reduced("c1", 8),
    "channel-weno3": lambda: CaseSpec("channel", HYBRID, 6, "weno3", seed=7, bc=(1, 2, 4, 4, 3, 3),
                                      bump=tuple(np.random.default_rng(22).uniform(0, 0.01, 8))),
}
# Mach 0.6 at the vortex's mean pressure, characteristic inflow and outflow
FREESTREAM = (1, prim2cons(1.0, 6.0, 0.0, 0.0, 100 / 1.4), prim2cons(1.0, 6.0, 0.0, 0.0, 100 / 1.4))
_cache = {}


def get(name):
    """the case, a state without the vortex's mirror symmetries and a random
    caller order.  Mirrored cells hold equal values, so the Barth limiter's
    bound and active face point are ties there; states that differ in the
    last bit (as the two runs' stage states do, since each face's owner and
    so its rounding depend on the order) would take different branches of
    the tie in the adjoint"""
    if name not in _cache:
        case = Case(CASES[name]())
        rng = np.random.default_rng(3)
        sh = rng.permutation(case.ncell).astype(np.int64)
        x = case.array("centre").reshape(-1, 3)
        u = case.init_state()
        u = u * (1 + 0.02 * np.sin(0.37 * x[:, :1] + 0.71 * x[:, 1:2] + 0.53 * x[:, 2:3] + 0.2))
        u = u * (1 + 1e-3 * rng.uniform(-1, 1, u.shape))
        _cache[name] = (case, u, sh)
    return _cache[name]


def rel(a, b):
    return np.sqrt(((a - b) ** 2).sum(0)) / np.maximum(np.sqrt((b ** 2).sum(0)), 1e-300)


@pytest.mark.parametrize("name", ["tet-weno4", "hex-muscl", "channel-weno3"])
def test_oracle_on_a_shuffled_caller_order_is_the_permuted_result(name):
    from oracle import oracle
    case, u, sh = get(name)
    orc = Oracle(case)
    if case.spec.bc[0]:
        oracle.set_boundary(*FREESTREAM)
    r0, c0 = orc.residual(u)
    keep, (m, s, p) = shuffled_structs(case, sh)
    orc.m, orc.s, orc.p = m, s, p
    r1, c1 = orc.residual(u[sh])
    assert np.array_equal(r1, r0[sh])
    assert list(c0) == list(c1)
    u0, _ = Oracle(case).rk3(u, 2, cfl=0.5)
    u1, _ = orc.rk3(u[sh], 2, cfl=0.5)
    assert np.array_equal(u1, u0[sh])


# ---------------------------------------------------------------- GPU

def solver(case, structs=None, order=UCNS_ORDER_CALLER, precision=UCNS_FP64):
    from ucns3d_amd.solver import Solver
    s = Solver(case, precision, cell_order=order, structs=structs)
    if case.spec.bc[0]:
        s.set_boundary(*FREESTREAM)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("order", [UCNS_ORDER_RCM, UCNS_ORDER_MORTON])
def test_library_order_matches_the_caller_order(name, order):
    """residual, reconstruction, one RK3 step, the residual product and the
    RK3 adjoint through a shuffled caller order with the library's own
    order, against the unshuffled caller order without one.  Face ownership
    and so the rounding of each face sum change with the order: the
    residuals, sums of cancelling face fluxes, agree to 1e-11 (the parity
    bar), the states after two steps to 1e-12, the RK3 cotangents to 1e-9
    (the stage states differ in their last bits, and on the bumped channel,
    where most cells take the Barth limiter, the cotangent moves with the
    limiter's near-ties)"""
    import torch
    case, u, sh = get(name)
    print(f"[order] {name} order={order} ncell={case.ncell}", flush=True)
    ref = solver(case)
    ref.set_state(u)
    r0 = ref.residual()
    ul0 = ref.reconstruct()
    dt = 0.5 * ref.compute_dt(0.5)
    w = np.random.default_rng(5).normal(size=u.shape)
    g0, gb0 = ref.residual_vjp(w)
    a0, ab0, _ = ref.rk3_adjoint(2, w, dt_fixed=dt)
    u2 = ref.get_state()
    ref.close()

    keep, st = shuffled_structs(case, sh)
    s = solver(case, st, order)
    s.set_state(u[sh])
    assert np.array_equal(s.get_state(), u[sh])
    r1 = s.residual()
    print("  residual", rel(r1, r0[sh]))
    assert np.all(rel(r1, r0[sh]) <= 1e-11)
    ul1 = s.reconstruct()
    assert np.allclose(ul1, ul0[sh], rtol=0, atol=1e-12 * np.abs(ul0).max())
    g1, gb1 = s.residual_vjp(w[sh])
    print("  vjp", rel(g1, g0[sh]))
    assert np.all(rel(g1, g0[sh]) <= 1e-11)
    assert np.allclose(gb1, gb0, rtol=1e-11, atol=1e-11 * (np.abs(gb0).max() + 1e-300))
    s.set_state(u[sh])
    a1, ab1, _ = s.rk3_adjoint(2, w[sh], dt_fixed=dt)
    print("  rk3 adjoint", rel(a1, a0[sh]), "state", rel(s.get_state(), u2[sh]))
    assert np.all(rel(a1, a0[sh]) <= 1e-9)
    assert np.all(rel(s.get_state(), u2[sh]) <= 1e-12)
    # device entry points and the SoA layout go through the same permutation
    t = torch.tensor(u[sh], dtype=torch.float64, device="cuda")
    s.set_state_device(t.data_ptr())
    out = torch.empty_like(t)
    s.residual_device(out.data_ptr())
    assert np.all(rel(out.cpu().numpy(), r0[sh]) <= 1e-11)
    soa = np.empty((5, case.ncell))
    from ucns3d_amd import native
    assert s.L.ucns_get_state(s._h, native.f64p(soa), UCNS_LAYOUT_SOA) == 0
    assert np.array_equal(soa.T, u[sh])
    s.close()


@pytest.mark.gpu
def test_library_order_restores_throughput_of_a_shuffled_mesh():
    """32^3 x 6 tetrahedra (196,608 cells), WENO-4, mixed precision: a
    random caller order with the library's Morton order runs at the speed of
    the generator's lexicographic order (within 10%), and both orders give
    the same state after 3 steps (1e-5, the mixed path's bar: the FP32 face
    solves round differently when the owner of a face changes)"""
    import torch
    case = Case(CaseSpec("tet-32", TET, 32, "weno4", jitter=0.1, seed=2))
    u = case.init_state()
    sh = np.random.default_rng(9).permutation(case.ncell).astype(np.int64)
    print(f"[order throughput] ncell={case.ncell}", flush=True)

    def timed(s, u0):
        s.set_state(u0)
        dt = s.compute_dt(0.5)
        s.rk3_steps(1, dt_fixed=dt)
        s.set_state(u0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.rk3_steps(3, dt_fixed=dt)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, s.get_state()

    s = solver(case, precision=UCNS_MIXED)
    t_lex, x_lex = timed(s, u)
    s.close()
    keep, st = shuffled_structs(case, sh, UCNS_MIXED)
    s = solver(case, st, UCNS_ORDER_MORTON, precision=UCNS_MIXED)
    t_mor, x_mor = timed(s, u[sh])
    s.close()
    print(f"  lexicographic {t_lex * 1e3 / 3:.2f} ms/step, shuffled + Morton {t_mor * 1e3 / 3:.2f} ms/step")
    assert np.all(rel(x_mor, x_lex[sh]) <= 1e-5)
    assert t_mor <= 1.10 * t_lex
