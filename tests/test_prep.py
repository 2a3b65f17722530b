"""Host preprocessor (ucns3d_amd/csrc/prep) against independent geometry:
volumes, face areas and normals, neighbour pairing, the face quadrature
rules, Gauss points, stencils, the local frame data and the least-squares
pseudo-inverses.  Everything here is recomputed in numpy from the mesh
vertices (tests/meshcheck.py), not from the preprocessor's own formulas."""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle
from ucns3d_amd.case import CHANNEL, HEX, HYBRID, PRISM, PYRAMID, TET, Case, CaseSpec, face_rule, reduced

from meshcheck import Mesh, factorial_ratio_tri, monomial_exponents, monomials

CASES = {
    "hex-muscl": lambda: reduced("c1", 5),
    "hex-weno4": lambda: reduced("c3", 8),
    "tet-weno4": lambda: reduced("c2", 4),
    "tet-muscl": lambda: CaseSpec("tet-muscl", TET, 4, "muscl", jitter=0.15, seed=7),
    "tet-weno3": lambda: CaseSpec("tet-weno3", TET, 4, "weno3", jitter=0.1, seed=8),
    "hex-weno4-box": lambda: CaseSpec("hex-box", HEX, 8, "weno4", seed=9, bc=(1, 2, 4, 4, 3, 3)),
    # prisms (2 triangles + 3 quadrilaterals), pyramids (1 + 4), the three
    # in one mesh, and the channel: prism wall layers, an interface layer of
    # pyramids and tetrahedra, tetrahedra above; bounded, with and without
    # its lower-wall bump (whose wall quadrilaterals are not planar)
    # (jittered: on the unjittered pyramid mesh the symmetric member centres
    # make every central stencil's quadratic fit singular)
    "prism-weno3": lambda: CaseSpec("prism", PRISM, 5, "weno3", jitter=0.1, seed=16),
    "pyramid-weno3": lambda: CaseSpec("pyramid", PYRAMID, 4, "weno3", jitter=0.1, seed=17),
    "hybrid-weno4": lambda: CaseSpec("hybrid", HYBRID, 6, "weno4", jitter=0.1, seed=21),
    "channel-weno3": lambda: CaseSpec("channel", CHANNEL, 3, "weno3", seed=4, bc=(1, 2, 4, 4, 4, 4),
                                      box=(12, 3, 3, 4.0, 1.0, 1.0)),
    "channel-bump-weno3": lambda: reduced("c4", 3),
    # quartic bases: 70-cell stencils, 16-point quadrilateral and 10-point
    # triangle face rules, derivatives up to fourth order in the indicator
    "hex-weno5": lambda: CaseSpec("hex-weno5", HEX, 8, "weno5", jitter=0.1, seed=9),
    "tet-weno5": lambda: CaseSpec("tet-weno5", TET, 4, "weno5", jitter=0.1, seed=10),
    # the blast box: hexahedra at the centre, a shell of pyramids and
    # tetrahedron pairs, Kuhn tetrahedra, prism layers at the z ends; outflow sides
    "blast-weno3": lambda: dataclasses.replace(reduced("c5", 8), scheme="weno3", ees=0),
}
_cache = {}


def get(name):
    if name not in _cache:
        case = Case(CASES[name]())
        _cache[name] = (case, Mesh(case))
    return _cache[name]


@pytest.fixture(params=list(CASES))
def cm(request):
    return get(request.param)


def box(case):
    nx, ny, nz, lx, ly, lz = case.spec.dims
    return np.array([lx, ly, lz])


def planar(case):
    """every face planar: no bump (a bumped wall bends the quadrilaterals
    next to it), and no jitter unless every face is a triangle"""
    return not any(case.spec.bump) and (case.spec.jitter == 0 or case.spec.shape == TET)


def test_volumes_tile_the_box(cm):
    case, m = cm
    L = box(case).prod()
    ref = np.array([m.cell_rule(i, 2)[1].sum() for i in range(m.nc)])
    assert np.all(m.vol > 0)
    np.testing.assert_allclose(m.vol, ref, rtol=1e-12)
    if planar(case):
        assert abs(m.vol.sum() - L) <= 1e-12 * L
    elif not any(case.spec.bump):
        # jittered quadrilaterals are bent, and two cells may split a shared
        # one along different diagonals: the cells tile the box to the
        # second order in the bend
        assert abs(m.vol.sum() - L) <= 1e-3 * L, (m.vol.sum(), L)
    else:
        # the bump takes the volume under the wall out of the box: the
        # trapezoidal sum of the wall heights over the wall's node grid (exact
        # for a bilinear wall; the split of each bent quadrilateral moves it
        # by its twist)
        nx, ny, nz, lx, ly, lz = case.spec.dims
        xs, zs = np.linspace(0, lx, nx + 1), np.linspace(0, lz, nz + 1)
        xh = 4 * xs / lx
        wx = sum(a * np.sin((k + 1) * math.pi * (xh - 1) / 2) for k, a in enumerate(case.spec.bump))
        wx = np.where((xh >= 1) & (xh <= 3), wx, 0.0)
        h = ly * wx[:, None] * np.sin(math.pi * zs / lz)[None, :]
        lost = (lx / nx) * (lz / nz) * 0.25 * (h[:-1, :-1] + h[1:, :-1] + h[:-1, 1:] + h[1:, 1:]).sum()
        assert lost > 0
        assert abs(m.vol.sum() - (L - lost)) <= 0.02 * lost, (m.vol.sum(), L - lost)


def test_other_split_of_planar_prisms_and_pyramids_agrees(cm):
    """the DECOMPOSE3 sub-tetrahedra and a different split give the same
    volume and first moments on cells with planar faces"""
    case, m = cm
    if not planar(case):
        return
    cells = [i for i in range(m.nc) if int(m.shape[i]) in (PRISM, PYRAMID)]
    for i in cells[::max(1, len(cells) // 50)]:
        x0, w0 = m.cell_rule(i, 2)
        x1, w1 = m.cell_rule(i, 2, alt=True)
        assert abs(w0.sum() - w1.sum()) <= 1e-13 * w0.sum()
        np.testing.assert_allclose((w0[:, None] * x0).sum(0), (w1[:, None] * x1).sum(0), rtol=1e-12,
                                   atol=1e-13 * w0.sum())


def test_faces_close_each_cell_and_point_outward(cm):
    case, m = cm
    for i in range(m.nc):
        nf = m.nfaces(i)
        s = (m.surf[i, :nf, None] * m.normal[i, :nf]).sum(0)
        # closed cells: the area vectors sum to zero (bent quadrilaterals:
        # to the bend, here < 1e-4 of the surface)
        assert np.abs(s).max() <= (1e-12 if planar(case) else 1e-4) * m.surf[i, :nf].sum()
        for f in range(nf):
            v = m.face_verts(i, f)
            n = m.normal[i, f]
            assert abs(np.linalg.norm(n) - 1) < 1e-14
            assert np.dot(v.mean(0) - m.centre[i], n) > 0
            if not planar(case):
                continue
            # planar faces: all vertices on the plane
            assert np.abs((v - v[0]) @ n).max() < 1e-12 * math.sqrt(m.surf[i, f])
            if len(v) == 3:
                area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
            else:
                area = 0.5 * np.linalg.norm(np.cross(v[2] - v[0], v[3] - v[1]))
            assert abs(area - m.surf[i, f]) <= 1e-12 * area


def test_face_angles_encode_the_normal(cm):
    case, m = cm
    ms, _, _ = case.structs()
    nf = case.nface_max
    a1 = np.ctypeslib.as_array(ms.faceanglex, shape=(m.nc * nf,)).reshape(m.nc, nf)
    a2 = np.ctypeslib.as_array(ms.faceangley, shape=(m.nc * nf,)).reshape(m.nc, nf)
    for i in range(0, m.nc, 7):
        for f in range(m.nfaces(i)):
            n = np.array([math.cos(a1[i, f]) * math.sin(a2[i, f]), math.sin(a1[i, f]) * math.sin(a2[i, f]),
                          math.cos(a2[i, f])])
            np.testing.assert_allclose(n, m.normal[i, f], atol=1e-12)
            # ROTATEF takes the face normal to the first axis
            r = oracle.rotatef(np.r_[0.0, m.normal[i, f], 0.0], a1[i, f], a2[i, f])
            # a2 = acos(n_z) loses digits as 1/sin(a2) near n = +-z (jittered
            # hexahedra have such faces)
            np.testing.assert_allclose(r, [0, 1, 0, 0, 0], atol=1e-12 * max(1.0, 0.1 / max(abs(math.sin(a2[i, f])), 1e-3)))


def test_neighbours_pair_up_across_shared_faces(cm):
    case, m = cm
    for i in range(m.nc):
        for f in range(m.nfaces(i)):
            j, fj = m.ineigh[i, f], m.ineighn[i, f]
            code = m.bcode[i, f]
            if code in (1, 2, 3, 4, 6):
                # a boundary face lies on the side of the box its code names
                v = m.face_verts(i, f)
                side = case.spec.bc.index(code) if case.spec.bc.count(code) == 1 else None
                assert j < 0
                L = box(case)
                lower_wall = not planar(case) and np.all(v[:, 1] < 0.5 * L[1])  # the bump moves it in y
                assert lower_wall or any(np.allclose(v[:, d], e * L[d]) for d in range(3) for e in (0, 1))
                if side is not None:
                    assert np.allclose(v[:, side // 2], (side % 2) * L[side // 2])
                continue
            assert j >= 0, "interior and periodic faces have a neighbour"
            assert m.ineigh[j, fj] == i and m.ineighn[j, fj] == f
            a = m.face_verts(i, f)
            b = m.face_verts(j, fj)
            b = b + m.unwrap_shift(a.mean(0), b.mean(0))
            key = lambda p: sorted(map(tuple, np.round(p, 10)))
            assert key(a) == key(b)
            np.testing.assert_allclose(m.normal[i, f], -m.normal[j, fj], atol=1e-13 if planar(case) else 1e-11)
            assert abs(m.surf[i, f] - m.surf[j, fj]) <= 1e-13 * m.surf[i, f]


def test_minedge_is_three_volumes_over_the_surface(cm):
    case, m = cm
    ssum = np.array([m.surf[i, :m.nfaces(i)].sum() for i in range(m.nc)])
    np.testing.assert_allclose(m.minedge, 3 * m.vol / ssum, rtol=1e-14)
    if case.spec.shape == HEX and case.spec.jitter == 0 and not case.spec.box:
        np.testing.assert_allclose(m.minedge, 0.5 * case.spec.length / case.spec.n, rtol=1e-14)


@pytest.mark.parametrize("igq,tri,deg", [(1, False, 1), (2, False, 3), (3, False, 5), (4, False, 7),
                                         (1, True, 1), (2, True, 2), (3, True, 4), (4, True, 5)])
def test_face_rules_integrate_polynomials_exactly(igq, tri, deg):
    pts, w = face_rule(igq, tri)
    assert abs(w.sum() - 1) < 1e-14
    if tri:
        assert np.allclose(pts.sum(1), 1.0)
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                for c in range(deg + 1 - a - b):
                    got = (w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum()
                    assert abs(got - factorial_ratio_tri(a, b, c)) < 1e-14, (a, b, c)
        # and not beyond
        worst = max(abs((w * pts[:, 0] ** a * pts[:, 1] ** (deg + 1 - a)).sum() - factorial_ratio_tri(a, deg + 1 - a, 0))
                    for a in range(deg + 2))
        assert worst > 1e-8
    else:
        mean = lambda k: 0.0 if k % 2 else 1.0 / (k + 1)  # mean of x^k on [-1,1]
        for a in range(deg + 1):
            for b in range(deg + 1):
                got = (w * pts[:, 0] ** a * pts[:, 1] ** b).sum()
                assert abs(got - mean(a) * mean(b)) < 1e-14, (a, b)
        assert abs((w * pts[:, 0] ** (deg + 1)).sum() - mean(deg + 1)) > 1e-8


def test_ten_point_triangle_rule_is_the_references_orbit_family():
    """QUADRATURETRIANG case 4 (src/grid_t.f90:2097-2108): orbits (a,b,b) x3,
    the 6 permutations of (p,q,r), the centroid, in that order; degree 5 leaves
    one free parameter, the vertex-orbit coordinate a = 0.888871894660413"""
    pts, w = face_rule(4, True)
    a = 0.888871894660413
    b = 0.5 * (1 - a)
    np.testing.assert_allclose(pts[:3], [[a, b, b], [b, a, b], [b, b, a]], rtol=0, atol=1e-15)
    np.testing.assert_allclose(pts[9], [1 / 3] * 3, rtol=0, atol=1e-16)
    p, q, r = pts[3]
    np.testing.assert_allclose(pts[3:9], [[p, q, r], [p, r, q], [r, p, q], [q, p, r], [q, r, p], [r, q, p]], atol=0)
    assert np.all(w > 0) and np.all(w[:3] == w[0]) and np.all(w[3:9] == w[3])
    assert len({round(p, 6), round(q, 6), round(r, 6)}) == 3


def face_param(v, x):
    """barycentric coordinates (triangle) or bilinear parameters (planar quad) of x"""
    if len(v) == 3:
        A = np.stack([v[1] - v[0], v[2] - v[0]], axis=1)
        lam, *_ = np.linalg.lstsq(A, x - v[0], rcond=None)
        return np.r_[1 - lam.sum(), lam]
    # parallelogram faces (all faces of the tested hex meshes)
    A = np.stack([v[1] - v[0], v[3] - v[0]], axis=1)
    st, *_ = np.linalg.lstsq(A, x - v[0], rcond=None)
    return st


def test_gauss_points_lie_on_faces_and_match_across_them(cm):
    case, m = cm
    if not planar(case):
        return  # bent quadrilaterals: the Gauss points are off the vertex plane
    nqm = case.nqpmax
    qp = case.array("qpoints").reshape(m.nc, 6, nqm, 3)
    qw = case.array("quad_w")
    tw = case.array("tri_w")
    for i in range(0, m.nc, 3):
        for f in range(m.nfaces(i)):
            v = m.face_verts(i, f)
            nq = case.qp_tri if len(v) == 3 else case.qp_quad
            w = tw if len(v) == 3 else qw
            xg = m.to_global(i, qp[i, f, :nq])
            for x in xg:
                p = face_param(v, x)
                assert np.all(p > -1e-12) and np.all(p < 1 + 1e-12), p
                assert abs(np.dot(x - v[0], m.normal[i, f])) < 1e-12
            # the rule integrates linear functions: weighted mean = face centroid
            np.testing.assert_allclose((w[:, None] * xg).sum(0), v.mean(0), atol=1e-12)
            # the neighbour sees the same physical points in the same order
            j, fj = m.ineigh[i, f], m.ineighn[i, f]
            if j < 0:
                continue
            xj = m.to_global(j, qp[j, fj, :nq])
            xj = xj + m.unwrap_shift(xg.mean(0), xj.mean(0))
            np.testing.assert_allclose(xj, xg, atol=1e-12)


def bfs(m, i, limit):
    out = [i]
    seen = {i}
    head = 0
    while len(out) < limit and head < len(out):
        c = out[head]
        head += 1
        for f in range(m.nfaces(c)):
            nb = m.ineigh[c, f]
            if nb < 0 or nb in seen or len(out) >= limit:
                continue
            seen.add(nb)
            out.append(nb)
    return out


def cone_test(o, face, pts):
    """coordinates of pts in every sub-cone from o through the face: a
    triangle gives one cone, a quadrilateral the cones of its four vertex
    triples (the point set of a cone does not depend on vertex order).
    Returns the per-point max over sub-cones of the smallest coordinate."""
    tris = [face] if len(face) == 3 else [face[[0, 1, 2]], face[[1, 2, 3]], face[[0, 2, 3]], face[[0, 1, 3]]]
    best = np.full(len(pts), -np.inf)
    for t in tris:
        A = np.stack([t[0] - o, t[1] - o, t[2] - o], axis=1)
        lam = np.linalg.solve(A, (pts - o).T).T
        best = np.maximum(best, lam.min(axis=1) / np.abs(lam).max(axis=1).clip(1e-300))
    return best


def test_stencils_are_bfs_and_face_cones(cm):
    """central stencil = the first nnb cells of the face-neighbour BFS (no
    distance sort); directional stencil of face f = the cell, then the BFS
    candidates in BFS order whose centres lie in the cone from the cell
    centre through face f; complete stencils are kept in face order.
    Candidates on a cone face (ties decided by rounding) may go either way."""
    case, m = cm
    nnb, ts = case.nnb, case.typesten
    ih = case.array("ihexl").reshape(m.nc, ts, nnb)
    admis = case.array("admis")
    full = case.array("full")
    tol = 1e-9
    checked = 0
    for i in range(0, m.nc, max(1, m.nc // 40)):
        pool = bfs(m, i, case.iselem)
        assert list(ih[i, 0]) == pool[:nnb]
        if ts == 1:
            assert admis[i] == 1 and full[i] == 0
            continue
        rank = {c: k for k, c in enumerate(pool)}
        o = m.centre[i]
        cand = np.array(pool[1:])
        # candidate images for the cone test (STENCIILS, src/grid_p.f90:5822-5960):
        # a centre at least half a period away moves by one period towards
        # the half of the box that holds the cell
        pc = m.centre[cand].copy()
        far = np.abs(pc - o) >= 0.5 * m.L
        pc += far * np.where(o - 0.5 * m.L >= 0, 1.0, -1.0) * m.L
        k = 1
        for f in range(m.nfaces(i)):
            score = cone_test(o, m.face_verts(i, f), pc)
            loose = set(cand[score >= -tol].tolist())
            strict = [c for c in cand[score > tol].tolist()]
            if len(loose) < nnb - 1:
                continue  # cannot be complete
            ok = False
            if k < admis[i]:
                row = ih[i, k]
                ranks = [rank[c] for c in row[1:]]
                ok = (row[0] == i and set(row[1:].tolist()) <= loose and ranks == sorted(ranks)
                      and all(c in set(row[1:].tolist()) for c in strict if rank[c] <= ranks[-1]))
            if ok:
                k += 1
                checked += 1
            else:
                assert len(strict) < nnb - 1, (i, f, "complete stencil missing")
        assert admis[i] == k
        assert full[i] == (1 if admis[i] >= m.nfaces(i) + 1 else 0)
    if ts > 1:
        assert checked > 0


def test_local_frame_moments(cm):
    """INTEG_BASIS / VOLUME(1,1) = local-frame mean of every basis monomial
    over the cell; VOLUME(1,1) = local volume."""
    case, m = cm
    exps = monomial_exponents(case.ideg)
    ib = case.array("integ_basis").reshape(m.nc, case.ideg)
    v11 = case.array("volume11")
    for i in range(0, m.nc, max(1, m.nc // 25)):
        x, w = m.cell_rule(i, 4)
        detJ = abs(np.linalg.det(m.frame[i, :9].reshape(3, 3)))
        assert abs(v11[i] - m.vol[i] * detJ) <= 1e-12 * v11[i]
        mon = monomials(m.to_local(i, x), exps)
        np.testing.assert_allclose(ib[i], (mon * w[:, None]).sum(0) * detJ, rtol=1e-11, atol=1e-13)


def test_pseudo_inverse_is_the_volume_weighted_least_squares_solution(cm):
    """INVMAT_STENCILT (src/prestore.f90:437-560) maps the member-average
    differences dU_j to the coefficients g minimising
    sum_j (V_j (A g - dU)_j)^2, A_jk = mean over member j of (m_k - mbar_k):
    P = pinv(diag(V) A) diag(V), and P A = I.

    On unjittered hexahedra the cone test settles ties (centres on cone
    faces) by rounding, and some directional stencils then have no unique
    cubic fit (A rank-deficient): their P is whatever the QR inverse of the
    singular normal matrix gives, identical in every consumer; the test
    requires it finite and pins the others."""
    case, m = cm
    exps = monomial_exponents(case.ideg)
    ts, nnb, imax, ideg = case.typesten, case.nnb, case.imax, case.ideg
    ih = case.array("ihexl").reshape(m.nc, ts, nnb)
    admis = case.array("admis")
    P = case.array("invmat").reshape(m.nc, ts, imax, ideg)
    detJs = np.abs(np.linalg.det(m.frame[:, :9].reshape(-1, 3, 3)))
    pinned = degenerate = 0
    for i in range(0, m.nc, max(1, m.nc // 6)):
        for s in range(admis[i]):
            mem = ih[i, s, 1:]
            A, V = m.design_matrix(i, mem, exps)
            V = V * detJs[i]
            got = P[i, s].T  # (ideg, imax)
            sv = np.linalg.svd(V[:, None] * A, compute_uv=False)
            if sv[-1] < 1e-10 * sv[0]:
                assert s > 0 and np.all(np.isfinite(got)), (i, s)
                degenerate += 1
                continue
            ref = np.linalg.pinv(V[:, None] * A) * V[None, :]
            # the preprocessor solves the normal equations: error ~ eps cond^2
            tol = max(1e-8, 1e-15 * (sv[0] / sv[-1]) ** 2)
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (i, s)
            assert np.abs(got @ A - np.eye(ideg)).max() < tol
            pinned += 1
    assert pinned > 4 * degenerate
    if case.spec.jitter > 0:
        assert degenerate == 0


def test_indicator_is_the_derivative_gram_matrix(cm):
    """INDICATOR_kl = sum over 1<=|alpha|<=p of the local-frame integral of
    D^alpha m_k D^alpha m_l over the cell (src/prestore.f90:1529-2025)."""
    case, m = cm
    if case.iweno != 1:
        return
    ideg = case.ideg
    exps = monomial_exponents(ideg)
    p = case.iorder
    alphas = [(a, b, c) for s in range(1, p + 1) for a in range(s + 1) for b in range(s + 1 - a) for c in [s - a - b]]
    B = case.array("indicator").reshape(m.nc, ideg, ideg)

    def falling(e, a):
        return float(np.prod([e - k for k in range(a)])) if a <= e else 0.0

    for i in range(0, m.nc, max(1, m.nc // 8)):
        # products of derivatives reach degree 2(p-1); the collapsed rule
        # with n points per axis is exact to degree 2n-3
        x, w = m.cell_rule(i, max(4, p + 2))
        detJ = abs(np.linalg.det(m.frame[i, :9].reshape(3, 3)))
        y = m.to_local(i, x)
        G = np.zeros((ideg, ideg))
        for al in alphas:
            D = np.zeros((len(y), ideg))
            for k, e in enumerate(exps):
                if all(e[d] >= al[d] for d in range(3)):
                    D[:, k] = np.prod([falling(e[d], al[d]) * y[:, d] ** (e[d] - al[d]) for d in range(3)], axis=0)
            G += (D * (w * detJ)[:, None]).T @ D
        np.testing.assert_allclose(B[i], G, rtol=1e-10, atol=1e-12 * np.abs(G).max())
