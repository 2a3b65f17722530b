"""The torch restatement of the preprocessor's geometry (ucns3d_amd.shape),
on the CPU: it reproduces the C++ preprocessor's arrays on tetrahedral,
hexahedral, hybrid and channel meshes, and its vector-Jacobian product
matches central differences of the C++ preprocessor itself under node
moves (ucnsp_move_nodes), array by array."""
import numpy as np
import pytest
import torch

from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec, reduced
from ucns3d_amd.shape import ShapeChain, bump_basis

NAMES = {"invmat": "invmat", "indicator": "indicator", "integ_basis": "integ_basis", "volume11": "volume11",
         "qpoints": "qpoints", "surf": "surf", "normal": "normal", "totvolume": "volume"}

CASES = {
    "channel": lambda: reduced("c4", 3),
    "tet-weno4": lambda: CaseSpec("tet-weno4", TET, 4, "weno4", jitter=0.1, seed=7),
    "hex-weno4-jitter": lambda: CaseSpec("hex-weno4-jitter", HEX, 4, "weno4", jitter=0.1, seed=3),
    "hex-muscl": lambda: CaseSpec("hex-muscl", HEX, 4, "muscl", seed=5),
    "hybrid-weno3": lambda: CaseSpec("hybrid-weno3", HYBRID, 6, "weno3", jitter=0.05, seed=2),
}


def moved(key):
    case = Case(CASES[key]())
    x0 = case.array("xyz").reshape(-1, 3).copy()
    case.move_nodes(x0)
    return case, x0


@pytest.mark.parametrize("key", list(CASES))
def test_restatement_reproduces_the_preprocessor(key):
    case, x0 = moved(key)
    d = ShapeChain(case).data(torch.tensor(x0))
    for a, b in NAMES.items():
        if a == "indicator" and case.iweno != 1:
            continue
        ref = case.array(b).reshape(d[a].shape)
        x = d[a].detach().numpy()
        if a == "invmat":
            # the normal equations' solve rounds to about eps cond(A)^2 per
            # stencil (Householder QR in the preprocessor, LU here)
            # the stencils the residual reads (a FULL cell's, the first of the others)
            P = ref.reshape(-1, case.imax, case.ideg)
            Q = x.reshape(P.shape)
            live = (np.abs(P).max((1, 2)) > 0) & ShapeChain(case).solvable.reshape(-1)
            assert np.all(Q[~live] == 0)
            sv = np.linalg.svd(P[live], compute_uv=False)
            cond = sv[:, 0] / sv[:, -1]
            err = np.abs(Q[live] - P[live]).max((1, 2)) / np.abs(P[live]).max((1, 2))
            print(f"  {key} invmat: worst {err.max():.2e}, worst err/(eps cond^2) {(err / (2.2e-16 * cond ** 2)).max():.2e}")
            assert np.all(err <= np.maximum(1e-12, 100 * 2.2e-16 * cond ** 2))
        else:
            err = np.abs(x - ref).max() / np.abs(ref).max()
            print(f"  {key} {a}: {err:.2e}")
            assert err <= 1e-13, a


@pytest.mark.parametrize("key", ["channel", "tet-weno4", "hybrid-weno3"])
def test_vjp_matches_differences_of_the_preprocessor(key):
    """<vjp_nodes(gbar), dx> against (<gbar, data(x0 + h dx)> - <gbar,
    data(x0 - h dx)>) / 2h with the C++ preprocessor's data, per array.

    The preprocessor's pseudo-inverses come from the normal equations and
    round to about eps cond(A)^2, which central differences divide by h: they
    are weighted on stencils with cond(A) < 300 only and differenced at
    h = 1e-5 (1e-6 elsewhere), and held to 1e-5 (1e-6 elsewhere).  Arrays a
    move leaves unchanged (a tetrahedron's data in its own frame) are held
    to the differences' rounding (1e-9 sum |gbar data| / h) instead of
    their vanishing derivative."""
    case, x0 = moved(key)
    rng = np.random.default_rng(3)
    chain = ShapeChain(case)
    shapes = {k: v.shape for k, v in chain.data(torch.tensor(x0), np.arange(1)).items()}
    if key == "channel":
        dx = bump_basis(case)[2]
    else:  # interior nodes moved at random (boundary and periodic nodes kept)
        lo, hi = x0.min(0), x0.max(0)
        inner = np.all((x0 > lo + 1e-9) & (x0 < hi - 1e-9), 1)
        dx = rng.standard_normal(x0.shape) * inner[:, None] * (hi - lo).max() / 10
    for a, b in NAMES.items():
        if a == "indicator" and case.iweno != 1:
            continue
        ref = case.array(b).copy()
        g = rng.standard_normal(ref.shape)
        h, bar = 1e-6, 1e-6
        if a == "invmat":
            P = ref.reshape(-1, case.imax, case.ideg)
            live = (np.abs(P).max((1, 2)) > 0) & chain.solvable.reshape(-1)
            cond = np.full(len(P), np.inf)
            sv = np.linalg.svd(P[live], compute_uv=False)
            cond[live] = sv[:, 0] / sv[:, -1]
            g = (g.reshape(P.shape) * (cond < 300)[:, None, None]).reshape(-1)
            h, bar = 1e-5, 1e-5
        vals = []
        for t in (h, -h):
            case.move_nodes(x0 + t * dx)
            vals.append(float((g * case.array(b)).sum()))
        case.move_nodes(x0)
        fd = (vals[0] - vals[1]) / (2 * h)
        gb = {a: g.reshape((case.ncell,) + tuple(shapes[a][1:]))}
        an = float((chain.vjp_nodes(gb, x0) * dx).sum())
        # floor: the differences' rounding, about eps sum |gbar data| / h
        err = abs(an - fd) / max(abs(fd), 1e-9 * float(np.abs(g * ref).sum()) / h)
        print(f"  {key} {a:12s} fd {fd: .10e} vjp {an: .10e} err {err:.2e}")
        assert err <= bar, a
