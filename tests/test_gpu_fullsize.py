"""Parity at the benchmark sizes: C2 (64^3 x 6 jittered tetrahedra,
1,572,864 cells) and C3 (128^3 hexahedra with density noise, 2,097,152
cells), WENO-4 HLLC, one SSP-RK3 step of the CFL step size, FP64 and mixed,
against the CPU oracle run on the same mesh and state.

Two choices of the GPU path are exact in real arithmetic and round
differently from the reference:
  - the reference forms each face value as sum_s w_s (U_i + res_s)
    (EXTRAPOLATE_BOUND, src/reconstruct_p.f90:1577), the GPU as
    U_i + sum_s w_s res_s (the weights sum to 1);
  - the reference solves every interior face from both sides, each with
    its own face angles and area (src/flux_p.f90:526-749), the GPU once,
    from the owner, negated for the neighbour.
A few ulps of E (through HLLC's acoustic upwinding) and of the rotated
pressure force then move the mass and z-momentum residuals, small
differences of large face terms for this nearly divergence-free vortex, by
parts in 1e11.  So the oracle runs both ways, and once more with each
cell's face sum reversed (a reordering of a single sum, the smallest
change of rounding there is):
  - against the oracle making the GPU's two choices, the FP64 path holds
    1e-12 per variable, or twice what reversing one sum does to the oracle
    where that is larger (rho*w: 2e-12 on C2 and C3; the GPU's remaining
    differences are of that kind: contracted multiply-adds, the order of
    the least-squares and indicator sums);
  - against the oracle in the reference's arithmetic, it stays within twice
    the oracle's own difference between the two ways, or 1e-12.
Mixed: <= 1e-5 against the FP64 oracle per variable, except rho*w, a field
1e-5 the size of rho*u after one step (w = 0 in the initial vortex): its
error is held to 1e-5 of the momentum vector's norm and 2e-5 of its own."""
import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import PRESETS, Case
from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED
from ucns3d_amd.solver import Solver

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))


@pytest.mark.parametrize("key", ["c2", "c3"])
def test_one_step_at_full_size(key):
    spec = PRESETS[key]
    case = Case(spec)
    u = case.init_state()
    print(f"[full size] {spec.name}: {case.ncell} cells", flush=True)
    orc = Oracle(case)
    dt = orc.dt(u, spec.cfl)
    ref, _ = orc.rk3(u, 1, dt_fixed=dt)
    oracle.set_uleft_form(True)
    oracle.set_face_once(True)
    try:
        ref_mf, _ = orc.rk3(u, 1, dt_fixed=dt)
    finally:
        oracle.set_uleft_form(False)
        oracle.set_face_once(False)
    oracle.set_face_order(True)
    try:
        ref_rev, _ = orc.rk3(u, 1, dt_fixed=dt)
    finally:
        oracle.set_face_order(False)
    sens = rel(ref_mf, ref)
    sens_sum = rel(ref_rev, ref)
    print("  oracle, the GPU's arithmetic choices vs the reference's:", sens)
    print("  oracle, faces summed in reverse vs in order:            ", sens_sum, flush=True)
    out = {}
    for prec in (UCNS_FP64, UCNS_MIXED):
        s = Solver(case, prec)
        s.set_state(u)
        assert abs(s.compute_dt(spec.cfl) - dt) <= 1e-15 * dt
        s.rk3_steps(1, dt_fixed=dt)
        out[prec] = s.get_state()
        s.close()
    e64, e64_mf, emx = rel(out[UCNS_FP64], ref), rel(out[UCNS_FP64], ref_mf), rel(out[UCNS_MIXED], ref)
    print("  fp64 vs oracle (reference's choices):", e64)
    print("  fp64 vs oracle (GPU's choices):      ", e64_mf)
    print("  mixed vs oracle:                     ", emx, flush=True)
    assert np.all(e64_mf <= np.maximum(1e-12, 2.0 * sens_sum)), (e64_mf, sens_sum)
    assert np.all(e64 <= np.maximum(1e-12, 2.0 * sens)), (e64, sens)
    assert np.all(emx[[0, 1, 2, 4]] <= 1e-5), emx
    mom = np.sqrt((ref[:, 1:4] ** 2).sum())
    assert np.linalg.norm(out[UCNS_MIXED][:, 3] - ref[:, 3]) <= 1e-5 * mom and emx[3] <= 2e-5, emx
