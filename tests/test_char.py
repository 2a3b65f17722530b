"""Characteristic-variable WENO (WENWRT=2) in the oracle: face by face, the
stencils' coefficients projected with the left eigenvectors of the face's
averaged state, weighted per characteristic variable and projected back
(src/reconstruct_p.f90:497-758, 1608-1818; src/flow_operations.f90:3934-3995).

Pinned by:
  - the eigenvectors: L R = I and R diag(u-a, u, u, u, u+a) L = the x-flux
    Jacobian at the averaged state (complex-step derivative of the flux);
  - the identity that defines the scheme: on every face, the characteristic
    face values equal R times the component-wise WENO face values of the
    frozen characteristic variables L U (the component-wise path is pinned
    separately), with the troubled-cell logic off on both sides;
  - polynomials of the scheme's degree reconstructed exactly;
  - convergence at the scheme's order on a smooth wave."""
import math

import numpy as np
import pytest

from meshcheck import Cubic, Mesh, monomial_exponents
from oracle import oracle
from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec

G = 1.4


def xflux(U, g=G):
    r, mu, mv, mw, E = U
    u, v, w = mu / r, mv / r, mw / r
    p = (g - 1) * (E - 0.5 * r * (u * u + v * v + w * w))
    return np.array([mu, mu * u + p, mv * u, mw * u, (E + p) * u])


def prim2cons(r, u, v, w, p, g=G):
    return np.array([r, r * u, r * v, r * w, p / (g - 1) + 0.5 * r * (u * u + v * v + w * w)])


def test_eigenvectors_diagonalise_the_flux_jacobian():
    rng = np.random.default_rng(1)
    for _ in range(100):
        a = prim2cons(*rng.uniform(0.5, 2, 1), *rng.uniform(-1, 1, 3), *rng.uniform(0.5, 2, 1))
        b = prim2cons(*rng.uniform(0.5, 2, 1), *rng.uniform(-1, 1, 3), *rng.uniform(0.5, 2, 1))
        R, L = oracle.char_eigen(G, a, b)
        np.testing.assert_allclose(L @ R, np.eye(5), atol=1e-13)
        # the averaged state: density, velocity, total energy averaged
        rs = 0.5 * (a[0] + b[0])
        vel = 0.5 * (a[1:4] / a[0] + b[1:4] / b[0])
        es = 0.5 * (a[4] + b[4])
        us = np.r_[rs, rs * vel, es]
        A = np.array([np.imag(xflux(us + 1e-30j * np.eye(5)[j])) / 1e-30 for j in range(5)]).T
        p = (G - 1) * (es - 0.5 * rs * vel @ vel)
        c = math.sqrt(G * p / rs)
        lam = np.diag([vel[0] - c, vel[0], vel[0], vel[0], vel[0] + c])
        np.testing.assert_allclose(R @ lam @ L, A, rtol=1e-12, atol=1e-12 * np.abs(A).max())


SPECS = {
    "tet-weno4": lambda: CaseSpec("ch-tet4", TET, 4, "weno4", jitter=0.1, seed=3, wenwrt=2),
    "hex-weno4-walls": lambda: CaseSpec("ch-hex4", HEX, 8, "weno4", jitter=0.1, seed=5, wenwrt=2,
                                        bc=(1, 2, 4, 4, 3, 3), mach=0.5),
    "hybrid-weno3": lambda: CaseSpec("ch-hyb3", HYBRID, 6, "weno3", jitter=0.1, seed=2, wenwrt=2),
}


def state_of(case):
    u = case.init_state()
    if case.spec.mach > 0:
        # a smooth perturbation, so that every stencil sees non-trivial data
        x = case.array("centre").reshape(-1, 3)
        u = u * (1 + 0.05 * np.sin(2.1 * x[:, :1] + 1.3 * x[:, 1:2] + 0.7 * x[:, 2:3] + np.arange(5)))
        oracle.set_boundary(1, case.spec.freestream(), case.spec.freestream())
    return u


@pytest.mark.parametrize("name", list(SPECS))
def test_characteristic_values_are_componentwise_weno_of_frozen_characteristic_variables(name):
    case = Case(SPECS[name]())
    m = Mesh(case)
    u = state_of(case)
    comp = Case(CaseSpec(**{**case.spec.__dict__, "wenwrt": 1}))
    oracle.set_raw(True)
    try:
        ul, _ = Oracle(case).reconstruct(u)
        nf = case.nface_max
        a1 = case.array("anglex").reshape(-1, nf)
        a2 = case.array("angley").reshape(-1, nf)
        bcode = case.array("bcode").reshape(-1, nf)
        ineigh = case.array("ineigh").reshape(-1, nf)
        rng = np.random.default_rng(7)
        worst, n = 0.0, 0
        for i in rng.choice(case.ncell, 12, replace=False):
            for f in range(m.nfaces(i)):
                bnd = bcode[i, f] not in (0, 5, 50)
                ur = oracle.ghost(int(bcode[i, f]), G, u[i], a1[i, f], a2[i, f]) if bnd else u[ineigh[i, f]]
                R, L = oracle.char_eigen(G, oracle.rotatef(u[i], a1[i, f], a2[i, f]),
                                         oracle.rotatef(ur, a1[i, f], a2[i, f]))
                w = u @ L.T  # the characteristic variables of every cell, frozen at this face
                ulw, _ = Oracle(comp).reconstruct(w)
                nq = case.qp_quad if (m.fnodes[i, f] >= 0).sum() == 4 else case.qp_tri
                got = ul[i, f, :nq]
                ref = ulw[i, f, :nq] @ R.T
                worst = max(worst, np.abs(got - ref).max() / np.abs(u[i]).max())
                n += 1
    finally:
        oracle.set_raw(False)
    assert n >= 40
    assert worst <= 1e-13, worst


POLY = {
    "tet-weno4": (lambda: CaseSpec("ch-tet4-6", TET, 6, "weno4", jitter=0.1, seed=5, wenwrt=2), 3),
    "hybrid-weno3": (lambda: CaseSpec("ch-hyb3-9", HYBRID, 9, "weno3", jitter=0.1, seed=21, wenwrt=2), 2),
}


@pytest.mark.parametrize("name", list(POLY))
def test_polynomials_are_reconstructed_exactly(name):
    from test_oracle import interior_cells
    make, deg = POLY[name]
    case = Case(make())
    m = Mesh(case)
    rng = np.random.default_rng(11)
    base = prim2cons(1.0, 0.2, -0.1, 0.3, 1.0)
    size = np.array(case.spec.dims[3:])
    polys = [Cubic(rng, deg, 0.5 * size, size.max()) for _ in range(5)]
    amp = 2e-3 * np.abs(base).clip(0.2)
    u = np.empty((case.ncell, 5))
    for i in range(case.ncell):
        x, w = m.cell_rule(i, 4)
        for v in range(5):
            u[i, v] = base[v] + amp[v] * (polys[v](x) * w).sum() / w.sum()
    ul, cnt = Oracle(case).reconstruct(u)
    assert cnt[1] == 0
    full = case.array("full")
    cells, _ = interior_cells(case, m, monomial_exponents(case.ideg), np.nonzero(full == 1)[0])
    assert len(cells) >= 8
    qp = case.array("qpoints").reshape(case.ncell, case.nface_max, case.nqpmax, 3)
    worst = 0.0
    for i in cells:
        for f in range(m.nfaces(i)):
            nq = case.qp_quad if (m.fnodes[i, f] >= 0).sum() == 4 else case.qp_tri
            xg = m.to_global(i, qp[i, f, :nq])
            for v in range(5):
                exact = base[v] + amp[v] * polys[v](xg)
                worst = max(worst, np.abs(ul[i, f, :nq, v] - exact).max() / amp[v])
    assert worst < 1e-9, worst


@pytest.mark.parametrize("shape,scheme,ns,order", [(HEX, "weno4", (12, 24), 3.0), (TET, "weno4", (4, 8), 2.7)])
def test_residual_converges_at_the_scheme_order(shape, scheme, ns, order):
    from test_oracle import wave_averages
    errs = []
    for n in ns:
        case = Case(CaseSpec(f"ch-wave-{n}", shape, n, scheme, seed=1, wenwrt=2))
        m = Mesh(case)
        U, dU = wave_averages(m, np.array([1.0, 1.0, 1.0]), np.array([0.7, -0.4, 0.3]))
        rhs, cnt = Oracle(case).residual(U)
        assert cnt[1] == 0
        got = -rhs / m.vol[:, None]
        errs.append(np.abs(got[:, 0] - dU[:, 0]).mean() / np.abs(dU[:, 0]).mean())
    rate = math.log(errs[0] / errs[1]) / math.log(ns[1] / ns[0])
    assert rate >= order, (errs, rate)
