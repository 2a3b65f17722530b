"""GPU parity of CWENO (EES=5) and CWENO-Z: the register kernels' central +
linear-sector reconstruction (IORDER 2, 3) and the per-cell kernel's quartic
central + quadratic-sector one (IORDER 4) (src/reconstruct_p.f90:955-1049, 1062-1199)
through the C-ABI against the oracle.

Bars as for WENO: FP64 reconstruction, residual and one SSP-RK3 step <= 1e-12
relative L2 per conserved variable, 1e-9 after 100 steps; the mixed path <=
1e-5 against the FP64 oracle after one step.  The periodic cases start from
the vortex with cell noise (every stencil sees non-smooth data, so the weights
take all values between the linear ones and a single sector); the channel
from a smooth perturbation of its free stream, with inflow, outflow, slip
walls and MUSCL cells at the walls."""
import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec, reduced
from ucns3d_amd.native import UCNS_FP64, UCNS_KERNELS_REGISTER, UCNS_MIXED
from ucns3d_amd.solver import Solver, UcnsError

pytestmark = pytest.mark.gpu

CASES = {
    "tet-weno4": lambda: CaseSpec("cw-tet4", TET, 5, "weno4", jitter=0.1, seed=3, ees=5),
    "hex-weno4": lambda: CaseSpec("cw-hex4", HEX, 10, "weno4", jitter=0.1, seed=5, ees=5),
    "hex-weno4-z": lambda: CaseSpec("cw-hex4z", HEX, 10, "weno4", jitter=0.1, seed=5, ees=5, wenoz=1),
    "hybrid-weno3": lambda: CaseSpec("cw-hyb3", HYBRID, 6, "weno3", jitter=0.05, seed=2, ees=5),
    "tet-weno3-z": lambda: CaseSpec("cw-tet3z", TET, 5, "weno3", jitter=0.1, seed=4, ees=5, wenoz=1),
    "channel-weno3": lambda: CaseSpec(**{**reduced("c4", 5).__dict__, "name": "cw-channel", "ees": 5}),
    "channel-weno3-z": lambda: CaseSpec(**{**reduced("c4", 5).__dict__, "name": "cw-channel-z", "ees": 5,
                                           "wenoz": 1}),
    # IORDER 4 (WENO-5 central, quadratic sectors on 20 members): the per-cell kernel
    "hex-weno5": lambda: CaseSpec("cw-hex5", HEX, 12, "weno5", seed=5, ees=5),
    "tet-weno5-z": lambda: CaseSpec("cw-tet5z", TET, 6, "weno5", jitter=0.1, seed=3, ees=5, wenoz=1),
    "hybrid-weno5": lambda: CaseSpec("cw-hyb5", HYBRID, 9, "weno5", jitter=0.1, seed=6, ees=5),
}

_cache = {}


def get(name):
    if name not in _cache:
        from test_gpu_geometry import smooth_state
        case = Case(CASES[name]())
        fs = case.spec.freestream() if case.spec.mach > 0 else None
        u = smooth_state(case) if fs is not None else case.init_state()
        _cache[name] = (case, Oracle(case), u, fs)
    case, orc, u, fs = _cache[name]
    if fs is not None:
        oracle.set_boundary(1, fs, fs)
    return _cache[name]


def solver(name, precision=UCNS_FP64, **kw):
    case, _, u, fs = get(name)
    s = Solver(case, precision, **kw)
    if fs is not None:
        s.set_boundary(1, fs, fs)
    s.set_state(u)
    return s


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1, 5)
    b = np.asarray(b).reshape(-1, 5)
    num = np.sqrt(((a - b) ** 2).sum(axis=0))
    den = np.sqrt((b ** 2).sum(axis=0))
    return num / np.where(den > 0, den, 1.0)


@pytest.mark.parametrize("name", list(CASES))
def test_reconstruction_fp64(name):
    from test_gpu_parity import valid_uleft_mask
    case, orc, u, _ = get(name)
    print(f"[cweno recon] {name} ncell={case.ncell} full={(case.array('full') == 1).sum()}", flush=True)
    s = solver(name)
    ul_gpu = s.reconstruct()
    st = s.stats()
    s.close()
    ul_ref, cnt = orc.reconstruct(u)
    m = valid_uleft_mask(case)
    err = rel_l2(ul_gpu[m], ul_ref[m])
    print("  rel L2", err, "muscl cells gpu/oracle", st["n_reduced_last"], cnt[0])
    assert st["n_reduced_last"] == cnt[0] and st["n_firstorder_last"] == cnt[1]
    assert np.all(err <= 1e-12), err


@pytest.mark.parametrize("name", list(CASES))
def test_residual_fp64(name):
    """against the oracle as written (1e-11: a sum of opposite face fluxes)
    and in the GPU path's arithmetic (face values as U + sum_l w_l res_l,
    each face solved once): 1e-12"""
    case, orc, u, _ = get(name)
    print(f"[cweno residual] {name}", flush=True)
    s = solver(name)
    r_gpu = s.residual()
    s.close()
    r_ref, _ = orc.residual(u)
    err = rel_l2(r_gpu, r_ref)
    oracle.set_uleft_form(True)
    oracle.set_face_once(True)
    try:
        r_ref2, _ = orc.residual(u)
    finally:
        oracle.set_uleft_form(False)
        oracle.set_face_once(False)
    err2 = rel_l2(r_gpu, r_ref2)
    print("  rel L2", err, "in the GPU arithmetic", err2)
    assert np.all(err <= 1e-11), err
    assert np.all(err2 <= 1e-12), err2


@pytest.mark.parametrize("name", list(CASES))
def test_one_step_fp64(name):
    case, orc, u, _ = get(name)
    print(f"[cweno rk3 x1] {name}", flush=True)
    s = solver(name)
    t = s.rk3_steps(1, cfl=case.spec.cfl)
    u1 = s.get_state()
    s.close()
    u1_ref, t_ref = orc.rk3(u, 1, cfl=case.spec.cfl)
    assert abs(t - t_ref) <= 1e-14 * t_ref
    err = rel_l2(u1, u1_ref)
    print("  rel L2", err)
    assert np.all(err <= 1e-12), err


@pytest.mark.parametrize("name", ["hex-weno4", "channel-weno3-z", "hex-weno5"])
def test_100_steps_fp64(name):
    case, orc, u, _ = get(name)
    print(f"[cweno rk3 x100] {name}", flush=True)
    s = solver(name)
    t = s.rk3_steps(100, cfl=case.spec.cfl)
    u100 = s.get_state()
    s.close()
    u100_ref, t_ref = orc.rk3(u, 100, cfl=case.spec.cfl)
    err = rel_l2(u100, u100_ref)
    print("  rel L2", err, "t", t, t_ref)
    assert abs(t - t_ref) <= 1e-12 * t_ref
    assert np.all(err <= 1e-9), err


@pytest.mark.parametrize("name", ["tet-weno4", "hex-weno4-z", "hybrid-weno3", "channel-weno3", "tet-weno5-z",
                                  "hybrid-weno5"])
def test_mixed_vs_fp64(name):
    case, orc, u, _ = get(name)
    print(f"[cweno mixed] {name}", flush=True)
    s = solver(name, UCNS_MIXED)
    s.rk3_steps(1, cfl=case.spec.cfl)
    u1 = s.get_state()
    s.close()
    u1_ref, _ = orc.rk3(u, 1, cfl=case.spec.cfl)
    err = rel_l2(u1, u1_ref)
    print("  rel L2 mixed vs FP64 oracle", err)
    assert np.all(err <= 1e-5), err


def test_register_kernels_requested_and_automatic_agree():
    """CWENO runs on the register kernels whichever the caller asks for: the
    automatic choice (which would take the row-stream kernels for WENO on
    hexahedra) gives the same bits"""
    a = solver("hex-weno4")
    b = solver("hex-weno4", kernels=UCNS_KERNELS_REGISTER)
    ra, rb = a.residual(), b.residual()
    a.close()
    b.close()
    assert np.array_equal(ra, rb)


def test_conservation_and_freestream():
    case, _, u, _ = get("tet-weno4")
    s = solver("tet-weno4")
    r = s.residual()
    scale = np.abs(r).sum(axis=0)
    assert np.all(np.abs(r.sum(axis=0)) <= 1e-12 * scale)
    uf = np.tile(np.array([1.2, 0.3, -0.2, 0.1, 3.0]), (case.ncell, 1))
    s.set_state(uf)
    assert np.abs(s.residual()).max() <= 1e-12 * 3.0
    s.close()


def test_adjoint_is_refused():
    """the adjoint does not cover CWENO: a clear error, not a wrong gradient"""
    case, _, u, _ = get("tet-weno4")
    s = solver("tet-weno4")
    with pytest.raises(UcnsError, match="CWENO"):
        s.rk3_adjoint(2, np.ones((case.ncell, 5)), dt_fixed=1e-4)
    s.close()
