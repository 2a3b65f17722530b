"""Adjoint of the residual and of SSP-RK3 on the GPU, through the C ABI.

Pinned three ways:
  - dot-product test against the oracle's tangent-linear model (the Dual
    instantiation of the FP64 restatement, itself pinned by central
    differences in test_tangent.py):  <w, J du> = <J^T w, du>, for the state
    and for the free-stream data;
  - the same identity over several RK3 steps with stage checkpoints;
  - central finite differences of the GPU forward itself (h = 1e-6).
Every row layout of the reconstruction is covered: the row stream (WENO-4
tetrahedra and hexahedra, WENO-3), the tiled layout (UCNS_KERNELS_REGISTER)
and the per-cell layout (MUSCL), on periodic and bounded boxes."""
import numpy as np
import pytest

from oracle import oracle
from ucns3d_amd.native import UCNS_FP64, UCNS_KERNELS_AUTO, UCNS_KERNELS_REGISTER, UCNS_MIXED
from ucns3d_amd.solver import Solver

from test_tangent import CASES, setup

pytestmark = pytest.mark.gpu


def solver_for(name, precision=UCNS_FP64, stream=True, **opts):
    case, orc, u, du, bc = setup(name)
    s = Solver(case, precision, kernels=UCNS_KERNELS_AUTO if stream else UCNS_KERNELS_REGISTER, **opts)
    if bc is not None:
        s.set_boundary(*bc)
    s.set_state(u)
    return s


def dot_gap(a, b, scale):
    return abs(a - b) / scale


VARIANTS = [(n, True) for n in CASES] + [("tet-weno4", False), ("hex-weno4-channel", False)]


@pytest.mark.parametrize("name,stream", VARIANTS)
def test_residual_vjp_dot_product(name, stream):
    case, orc, u, du, bc = setup(name)
    print(f"[vjp dot] {name} stream={stream} ncell={case.ncell}", flush=True)
    s = solver_for(name, stream=stream)
    rng = np.random.default_rng(11)
    w = rng.normal(size=u.shape)
    ub, bcb = s.residual_vjp(w)
    _, jdu = orc.residual_tl(u, du)
    lhs, rhs = float((w * jdu).sum()), float((ub * du).sum())
    scale = np.linalg.norm(w) * np.linalg.norm(jdu)
    print("  <w,J du>", lhs, "<J^T w,du>", rhs, "gap", dot_gap(lhs, rhs, scale))
    assert dot_gap(lhs, rhs, scale) < 1e-12
    # the state passed through: the forward of the product's state is the state's
    np.testing.assert_allclose(s.get_state(), u, rtol=0, atol=0)
    if bc is not None:
        dbc = rng.normal(size=10) * np.abs(np.r_[bc[1], bc[2]])
        _, jb = orc.residual_tl(u, np.zeros_like(u), dbc)
        lhs, rhs = float((w * jb).sum()), float(bcb @ dbc)
        scale = np.linalg.norm(w) * np.linalg.norm(jb)
        print("  boundary data: gap", dot_gap(lhs, rhs, scale))
        assert dot_gap(lhs, rhs, scale) < 1e-12
    else:
        assert np.all(bcb == 0)
    s.close()


@pytest.mark.parametrize("name", ["tet-weno4", "hex-muscl", "hex-weno4-channel", "tet-weno3-farfield"])
def test_rk3_adjoint_dot_product(name):
    case, orc, u, du, bc = setup(name)
    print(f"[rk3 adjoint dot] {name}", flush=True)
    s = solver_for(name)
    dt = 0.5 * s.compute_dt(0.5)
    rng = np.random.default_rng(12)
    a = rng.normal(size=u.shape)
    u0b, bcb, t = s.rk3_adjoint(3, a, dt_fixed=dt)
    assert abs(t - 3 * dt) < 1e-15
    uN, duN = orc.rk3_tl(u, du, 3, dt)
    # the state is left at U_N
    np.testing.assert_allclose(s.get_state(), uN, rtol=1e-12, atol=1e-12 * np.abs(uN).max())
    lhs, rhs = float((a * duN).sum()), float((u0b * du).sum())
    scale = np.linalg.norm(a) * np.linalg.norm(duN)
    print("  gap", dot_gap(lhs, rhs, scale))
    assert dot_gap(lhs, rhs, scale) < 1e-11
    if bc is not None:
        dbc = rng.normal(size=10) * np.abs(np.r_[bc[1], bc[2]])
        _, jb = orc.rk3_tl(u, np.zeros_like(u), 3, dt, dbc)
        lhs, rhs = float((a * jb).sum()), float(bcb @ dbc)
        scale = np.linalg.norm(a) * np.linalg.norm(jb)
        print("  boundary data: gap", dot_gap(lhs, rhs, scale))
        assert dot_gap(lhs, rhs, scale) < 1e-11
    s.close()


@pytest.mark.parametrize("name,stream,precision", [("tet-weno4", True, UCNS_FP64), ("tet-weno4", False, UCNS_FP64),
                                                    ("hex-muscl", True, UCNS_FP64),
                                                    ("hex-weno4-channel", True, UCNS_FP64),
                                                    ("tet-weno4", True, UCNS_MIXED)])
def test_rk3_adjoint_kept_stages_agree_with_recomputed(name, stream, precision):
    """The reverse sweep starts a stage's products from the reconstruction the
    forward sweep kept, or runs the reconstruction again (the scheme's
    adj_keep_stages caps the kept stages, adj_keep_gib their bytes): all,
    some (4 of 9) and none give the same cotangents"""
    case, orc, u, du, bc = setup(name)
    print(f"[rk3 adjoint keep] {name} stream={stream} precision={precision}", flush=True)
    a = np.random.default_rng(15).normal(size=u.shape)
    out = {}
    settings = {"9": dict(adj_keep_stages=9), "4": dict(adj_keep_stages=4), "0": dict(adj_keep_stages=-1),
                "bytes": dict(adj_keep_gib=1e-9)}
    dt = None
    for key, opts in settings.items():
        s = solver_for(name, precision=precision, stream=stream, **opts)
        if dt is None:
            dt = 0.5 * s.compute_dt(0.5)
        out[key] = s.rk3_adjoint(3, a, dt_fixed=dt)
        s.close()
    ref, rb, _ = out["9"]
    for keep in ("4", "0", "bytes"):
        g, gb, _ = out[keep]
        rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        print(f"  keep {keep}: rel {rel:.3e}")
        assert rel <= 1e-12
        np.testing.assert_allclose(gb, rb, rtol=1e-12, atol=1e-12 * (np.abs(rb).max() + 1e-300))
    s.close()


@pytest.mark.parametrize("name", ["tet-weno4", "hex-weno4-channel"])
def test_rk3_adjoint_matches_differences_of_the_gpu_forward(name):
    """J(U_0) = <c, U_N> over 5 steps; dJ/dU_0 . d against central
    differences of the GPU's own RK3 steps"""
    case, orc, u, du, bc = setup(name)
    print(f"[rk3 adjoint fd] {name}", flush=True)
    s = solver_for(name)
    dt = 0.5 * s.compute_dt(0.5)
    rng = np.random.default_rng(13)
    c = rng.normal(size=u.shape)
    g, _, _ = s.rk3_adjoint(5, c, dt_fixed=dt)

    def J(x):
        s.set_state(x)
        s.rk3_steps(5, dt_fixed=dt)
        return float((c * s.get_state()).sum())
    h = 1e-6
    fd = (J(u + h * du) - J(u - h * du)) / (2 * h)
    an = float((g * du).sum())
    print("  adjoint", an, "fd", fd, "rel", abs(an - fd) / abs(fd))
    assert abs(an - fd) <= 1e-6 * abs(fd)
    s.close()


@pytest.mark.parametrize("name", ["tet-weno4", "hex-weno4"])
def test_mixed_precision_vjp_is_close_to_fp64(name):
    case, orc, u, du, bc = setup(name)
    print(f"[vjp mixed] {name}", flush=True)
    rng = np.random.default_rng(14)
    w = rng.normal(size=u.shape)
    s64 = solver_for(name, UCNS_FP64)
    u64, _ = s64.residual_vjp(w)
    s64.close()
    s32 = solver_for(name, UCNS_MIXED)
    u32, _ = s32.residual_vjp(w)
    s32.close()
    err = np.linalg.norm(u32 - u64, axis=0) / np.linalg.norm(u64, axis=0)
    print("  rel L2 per variable", err)
    assert np.all(np.isfinite(u32))
    assert np.all(err[[0, 1, 2, 4]] < 1e-4), err


# ---------------------------------------------------------------- torch harness

def test_autograd_gradcheck_of_the_march():
    """torch.autograd.gradcheck of RK3March (forward ucns_rk3_steps, backward
    ucns_rk3_adjoint_device) on 512 WENO-3 hexahedra, every cell FULL (no
    MUSCL cell, whose limiter a one-hot difference could switch between face
    points): every column of the Jacobian by differences against every row
    by the adjoint.  The faces are axis-aligned and the vortex is carried
    with a uniform velocity larger than its own, so no face sits near
    S_M = 0, where HLLC switches star states with a kink (on tetrahedra,
    faces nearly parallel to the flow put such kinks within 1e-7 of the
    state, inside the difference stencil)"""
    import torch
    from ucns3d_amd.autograd import RK3March
    from ucns3d_amd.case import HEX, Case, CaseSpec
    from test_tangent import G
    print("[gradcheck] hex weno3 n=8", flush=True)
    case = Case(CaseSpec("h3", HEX, 8, "weno3", seed=2))
    assert case.array("full").min() == 1
    u = case.init_state()
    r = u[:, :1]
    vel = u[:, 1:4] / r + np.array([1.5, 1.2, 1.3])
    p = (G - 1) * (u[:, 4] - 0.5 * (u[:, 1:4] ** 2).sum(1) / u[:, 0])
    u = np.c_[r, r * vel, p / (G - 1) + 0.5 * r[:, 0] * (vel ** 2).sum(1)]
    # non-uniform density: with rho uniform its indicators are 0 and a 1e-6
    # difference moves them to ~eps, where the weights are far from linear.
    # 1% cell noise: in a smooth field the two arguments of the Einfeldt
    # estimate S_L = min(u_L - c_L, u~ - c~) differ only by the face jump of
    # the reconstruction, and a one-hot 1e-6 difference would flip the min
    x = case.array("centre").reshape(-1, 3)
    u = u * (1 + 0.02 * np.sin(0.37 * x[:, :1] + 0.71 * x[:, 1:2] + 0.53 * x[:, 2:3] + 0.2))
    u = u * (1 + 1e-2 * np.random.default_rng(15).uniform(-1, 1, u.shape))
    s = Solver(case, UCNS_FP64)
    s.set_state(u)
    dt = 0.5 * s.compute_dt(0.5)
    u0 = torch.tensor(u, dtype=torch.float64, device="cuda", requires_grad=True)
    ok = torch.autograd.gradcheck(lambda x: RK3March.apply(x, None, s, 2, dt), (u0,), eps=1e-6, atol=1e-6,
                                  rtol=1e-4, nondet_tol=1e-12, fast_mode=False)
    assert ok
    s.close()


def test_wall_pressure_sensitivity_to_the_inflow_mach_number():
    """J = integrated pressure on the lower wall of a channel (inflow x-,
    characteristic outflow x+, slip walls y, symmetry z) after 40 steps from
    a free stream at Mach M carrying a pressure/density blob next to the
    wall.  dJ/dM through the initial state and the inflow/outflow data
    (torch autograd over RK3March, i.e. ucns_rk3_adjoint_device) against
      - the FP64 tangent-linear reference: the oracle's Dual instantiation
        run forward with dU_0/dM and d(inflow, outflow)/dM (exact for the
        branch the primal takes), to 1e-9;
      - central differences of the GPU forward, to 1e-6 at h = 1e-8 (at
        h = 1e-6 they are 2.7e-5 off: on this mesh 94% of the cells are
        MUSCL cells next to the boundaries, and steps of 1e-6 in M straddle
        switches of the Barth limiter's active face point and of the
        Einfeldt min, where J has kinks)"""
    import torch
    from ucns3d_amd.autograd import RK3March
    from ucns3d_amd.case import HEX, Case, CaseSpec
    from oracle.oracle import Oracle
    print("[dJ/dMach] hex weno3 channel n=8", flush=True)
    case = Case(CaseSpec("channel", HEX, 8, "weno3", seed=3, bc=(1, 2, 4, 4, 3, 3)))
    g = 1.4
    c = g ** 0.5  # free-stream speed of sound (rho = p = 1)
    dev = "cuda"
    xc = case.array("centre").reshape(-1, 3)
    x = torch.tensor(xc, dtype=torch.float64, device=dev)
    nf = case.nface_max
    bcode = case.array("bcode").reshape(-1, nf)
    normal = case.array("normal").reshape(-1, nf, 3)
    surf = case.array("surf").reshape(-1, nf)
    wall = (bcode == 4) & (normal[:, :, 1] < -0.5)
    area_np = (surf * wall).sum(1)
    area = torch.tensor(area_np, dtype=torch.float64, device=dev)
    assert area_np.sum() > 0
    s = Solver(case, UCNS_FP64)

    def state(M):
        # a blob next to the lower wall (y = 0), upstream of the middle
        blob = torch.exp(-2.0 * ((x - torch.tensor([1.5, 0.8, 3.0], dtype=torch.float64, device=dev)) ** 2).sum(1))
        rho = 1.0 + 0.2 * blob
        pr = 1.0 + 0.3 * blob
        u = M * c * torch.ones_like(rho)
        U0 = torch.stack([rho, rho * u, 0 * rho, 0 * rho, pr / (g - 1) + 0.5 * rho * u * u], 1)
        fs = torch.stack([torch.ones_like(M), M * c, 0 * M, 0 * M, 1.0 / (g - 1) + 0.5 * (M * c) ** 2])
        return U0, torch.cat([fs, fs])

    def J(M, nsteps, dt):
        U0, bc = state(M)
        UN = RK3March.apply(U0, bc, s, nsteps, dt, 1)
        p = (g - 1) * (UN[:, 4] - 0.5 * (UN[:, 1:4] ** 2).sum(1) / UN[:, 0])
        return (p * area).sum()

    nsteps = 40
    M = torch.tensor(0.6, dtype=torch.float64, device=dev, requires_grad=True)
    U0, bc = state(M.detach())
    s.set_boundary(1, bc[:5].cpu().numpy(), bc[5:].cpu().numpy())
    s.set_state(U0.cpu().numpy())
    dt = 0.5 * s.compute_dt(0.5)
    Jm = J(M, nsteps, dt)
    Jm.backward()
    an = M.grad.item()
    # tangent-linear reference on the CPU
    Mt = torch.tensor(0.6, dtype=torch.float64, device=dev)
    U0t, bct = state(Mt)
    dU0 = torch.autograd.functional.jacobian(lambda m: state(m)[0], Mt).cpu().numpy()  # (ncell, 5)
    dbc = torch.autograd.functional.jacobian(lambda m: state(m)[1], Mt).cpu().numpy()  # (10,)
    orc = Oracle(case)
    b = bct.cpu().numpy()
    oracle.set_boundary(1, b[:5], b[5:])
    uN, duN = orc.rk3_tl(U0t.cpu().numpy(), dU0, nsteps, dt, dbc)
    r, m = uN[:, 0], uN[:, 1:4]
    vel = m / r[:, None]
    dpdU = (g - 1) * np.c_[0.5 * (vel ** 2).sum(1), -vel, np.ones_like(r)]
    tl = float(((dpdU * duN).sum(1) * area_np).sum())
    fds = {}
    with torch.no_grad():
        for h in (1e-6, 1e-7, 1e-8):
            fds[h] = (J(M.detach() + h, nsteps, dt) - J(M.detach() - h, nsteps, dt)).item() / (2 * h)
    print("  J", Jm.item(), "dJ/dM adjoint", an, "tangent-linear", tl, "rel(tl)", abs(an - tl) / abs(tl))
    for h, fd in fds.items():
        print(f"  fd h={h:.0e} {fd:.15e} rel {abs(an - fd) / abs(fd):.2e}")
    assert abs(an - tl) <= 1e-9 * abs(tl)
    assert abs(an - fds[1e-8]) <= 1e-6 * abs(fds[1e-8])
    s.close()
