"""Contexts side by side, and the library's error paths, on the GPU.

- Live contexts keep their own face quadrature: a MUSCL context (one Gauss
  point per face, config 1's scheme) and WENO-4 contexts (six per triangle
  on tetrahedra, nine per quadrilateral on hexahedra) interleave their calls, each checked
  against its oracle (the weights are w_q * SURF per scheme,
  src/flux_p.f90:669).
- A state that is not finite after an RK3 march is reported as
  UCNS_ENONFINITE (the march ran: the NaN reached the update through the
  ordinary kernels), and the context steps a finite state afterwards exactly
  as a fresh one does.
- A failed adjoint set-up (an allocation refused by ucns_debug_fail_alloc)
  returns its error again on the next call instead of running kernels on
  missing buffers, and a later call with the allocation allowed gives the
  result of a fresh context.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from ucns3d_amd.case import Case, reduced
from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED
from ucns3d_amd.solver import Solver, UcnsError

pytestmark = pytest.mark.gpu

UCNS_ENOMEM = -3
UCNS_ENONFINITE = -5


def rel(a, b):
    return np.sqrt(((a - b) ** 2).sum(0)) / np.maximum(np.sqrt((b ** 2).sum(0)), 1e-300)


def test_two_contexts_keep_their_face_rules():
    cm = Case(reduced("c1", 8))   # hexes, MUSCL: 1 point per quadrilateral
    cw = Case(reduced("c2", 4))   # tets, WENO-4: 6 points per triangle
    ch = Case(reduced("c3", 8))   # hexes, WENO-4: 9 points per quadrilateral
    assert (cm.qp_quad, cw.qp_tri, ch.qp_quad) == (1, 6, 9)
    print(f"[two contexts] {cm.ncell} MUSCL hexes, {cw.ncell} WENO-4 tets, {ch.ncell} WENO-4 hexes", flush=True)
    cases = [cm, cw, ch]
    us = [c.init_state() for c in cases]
    refs = [Oracle(c).residual(u)[0] for c, u in zip(cases, us)]
    solvers = []
    for c, u in zip(cases, us):   # created in turn: each creation once overwrote the others' weights
        s = Solver(c, UCNS_FP64)
        s.set_state(u)
        solvers.append(s)
    for rnd in range(2):
        for k in (2, 0, 1, 0, 2):
            err = rel(solvers[k].residual(), refs[k])
            print(f"  round {rnd} context {k}: rel L2 {err.max():.2e}")
            assert np.all(err <= 1e-11), (k, err)
    steps = [Oracle(c).rk3(u, 1, cfl=0.5)[0] for c, u in zip(cases, us)]
    for k in (1, 0, 2):
        solvers[k].set_state(us[k])
        solvers[k].rk3_steps(1, cfl=0.5)
    for k in range(3):
        assert np.all(rel(solvers[k].get_state(), steps[k]) <= 1e-12)
        solvers[k].close()


@pytest.mark.parametrize("name", ["c2", "c1"])
def test_failed_adjoint_setup_reports_again_and_recovers(name):
    """tet WENO-4 (row stream: member tables, CSR, stored coefficients) and
    hex MUSCL (per-cell layout): the k-th allocation of the adjoint's
    set-up fails, for every k up to the number it makes"""
    case = Case(reduced(name, 4))
    u = case.init_state()
    w = np.random.default_rng(2).normal(size=u.shape)
    fresh = Solver(case, UCNS_FP64)
    fresh.set_state(u)
    ref, _ = fresh.residual_vjp(w)
    fresh.close()
    print(f"[adjoint set-up failure] {name} ncell={case.ncell}", flush=True)
    k = 1
    while True:
        s = Solver(case, UCNS_FP64)
        s.set_state(u)
        s.debug_fail_alloc(k)
        try:
            s.residual_vjp(w)
            failed = False
        except UcnsError as e:
            failed = True
            assert e.code == UCNS_ENOMEM, e
            assert "injected" in str(e)
        if not failed:   # the set-up makes fewer than k allocations: all were covered
            s.close()
            break
        # the next call must not find a half-built adjoint: the same failure
        # reports again, and with nothing injected the set-up runs in full
        # and gives the fresh context's product
        s.debug_fail_alloc(k)
        with pytest.raises(UcnsError) as ei:
            s.residual_vjp(w)
        assert ei.value.code == UCNS_ENOMEM
        s.debug_fail_alloc(0)
        g, _ = s.residual_vjp(w)
        assert np.all(rel(g, ref) <= 1e-14), (k, rel(g, ref))
        s.close()
        k += 1
    print(f"  {k - 1} allocation points failed and recovered")
    assert k > 5


@pytest.mark.parametrize("precision", [UCNS_FP64, UCNS_MIXED])
def test_non_finite_state_is_reported(precision):
    case = Case(reduced("c2", 4))
    u = case.init_state()
    fresh = Solver(case, precision)
    fresh.set_state(u)
    fresh.rk3_steps(1, dt_fixed=1e-3)
    ref = fresh.get_state()
    fresh.close()
    bad = u.copy()
    bad[case.ncell // 3, 4] = np.nan
    print(f"[non-finite] c2 reduced ncell={case.ncell} precision={precision}", flush=True)
    s = Solver(case, precision)
    for kw in ({"dt_fixed": 1e-3}, {"cfl": 0.5}):
        s.set_state(bad)
        with pytest.raises(UcnsError) as ei:
            s.rk3_steps(1, **kw)
        assert ei.value.code == UCNS_ENONFINITE, ei.value
        assert "non-finite" in str(ei.value)
        assert not np.all(np.isfinite(s.get_state()))
    s.set_state(u)
    s.rk3_steps(1, dt_fixed=1e-3)
    assert np.array_equal(s.get_state(), ref)
    s.close()
