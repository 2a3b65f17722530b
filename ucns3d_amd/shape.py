"""Mesh-shape sensitivities: the host side of the geometry chain.

ucns_rk3_adjoint_geometry returns the cotangents of a march's objective with
respect to the data it reads from the mesh and PRESTORE arrays (pseudo-
inverses, indicators, basis integrals, VOLUME(1,1), face Gauss points, face
areas and normals, cell volumes).  This module restates, in torch FP64, how
the preprocessor forms those arrays from the node coordinates for one mesh's
fixed topology, stencils and discrete choices, so that the cotangents chain
back to the nodes and on to shape parameters:

    GEOMETRY_CALC volumes, areas, normals      src/grid_t.f90:1192-1242, 4441-4508
    LOCALISE_STEN2 cell frames                 src/local_pt.f90:1037-1101
    PRESTORE_RECONSTRUCTION3 pseudo-inverses   src/prestore.f90:285-529
    INDICATORMATRIX smoothness indicators      src/prestore.f90:1529-2025
    MEMORY_FAST face Gauss points              src/reconstruct_p.f90:178-244

The restatement follows the C++ preprocessor (ucns3d_amd/csrc/prep) step by
step: the cell frame of the first DECOMPOSE3 sub-tetrahedron, integrals of
the raw monomials by the sub-tetrahedra (Stroud T3:3-1 up to degree 3, the
collapsed Gauss-Legendre rule above) or the tensor rule of an affine
hexahedron, A_jk = integral over member j of (m_k - mbar_k), P = (A^T A)^-1
A^T with column j scaled by |V_j| / |det J_i|, Newell normals and the
two-triangle area check of quadrilaterals.  What is discrete stays as the
preprocessor chose it on the given mesh: stencil members, periodic shifts,
the affine-hexahedron test and the area check's branch.

The dependence of the data on the nodes is local (a cell's own nodes and its
stencil members' nodes), so the chain runs over chunks of cells and needs no
global graph.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native

# monomials of the basis (BASIS_REC, POLY=1, src/basis.f90:11-102): exponents
# of x, y, z in the basis order, up to degree 4
MONO = np.array([
    (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2),
    (3, 0, 0), (2, 1, 0), (1, 2, 0), (0, 3, 0), (2, 0, 1), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2), (0, 0, 3),
    (4, 0, 0), (3, 1, 0), (2, 2, 0), (1, 3, 0), (0, 4, 0), (3, 0, 1), (2, 1, 1), (1, 2, 1), (0, 3, 1),
    (2, 0, 2), (1, 1, 2), (0, 2, 2), (1, 0, 3), (0, 1, 3), (0, 0, 4)], dtype=np.int64)

HEX, TET, PRISM, PYRAMID = 1, 2, 3, 4

# Stroud T3:3-1: barycentric points and weights (times the volume)
_STROUD_L = np.array([[0.25, 0.25, 0.25, 0.25],
                      [0.5, 1 / 6, 1 / 6, 1 / 6], [1 / 6, 0.5, 1 / 6, 1 / 6],
                      [1 / 6, 1 / 6, 0.5, 1 / 6], [1 / 6, 1 / 6, 1 / 6, 0.5]])
_STROUD_W = np.array([-0.8, 0.45, 0.45, 0.45, 0.45])


def shape_table(shape: int) -> dict:
    """element tables of the preprocessor (ucnsp_shape_table)"""
    out = (C.c_int32 * 64)()
    lib = native.prep_lib()
    if lib.ucnsp_shape_table(shape, out) != 0:
        raise ValueError(lib.ucnsp_last_error().decode())
    o = list(out)
    nf = o[1]
    faces = [[o[8 + 4 * f + k] for k in range(o[2 + f])] for f in range(nf)]
    tets = [[o[33 + 4 * t + k] for k in range(4)] for t in range(o[32])]
    return {"nnodes": o[0], "faces": faces, "subtets": tets}


def _gauss_legendre(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def _derivs(p: int):
    """multi-indices 1 <= |a| <= p of the indicator's derivatives"""
    out = []
    for s in range(1, p + 1):
        for a in range(s, -1, -1):
            for b in range(s - a, -1, -1):
                out.append((a, b, s - a - b))
    return np.array(out, dtype=np.int64)


class ShapeChain:
    """The preprocessor's data of one Case as torch functions of the node
    coordinates, and their vector-Jacobian product."""

    def __init__(self, case, device: str = "cpu"):
        import torch
        self.torch = torch
        self.case = case
        self.dev = torch.device(device)
        self.nc = case.ncell
        self.NF = case.nface_max
        self.ID, self.IM, self.TS = case.ideg, case.imax, case.typesten
        self.iorder = case.iorder
        self.iweno = case.iweno
        self.shape = case.array("cell_shape").astype(np.int64).copy()
        self.cnodes = case.array("cell_nodes").reshape(self.nc, 8).astype(np.int64).copy()
        self.fnodes = case.array("face_nodes").reshape(self.nc, self.NF, 4).astype(np.int64).copy()
        self.admis = case.array("admis").astype(np.int64).copy()
        self.ihexl = case.array("ihexl").reshape(self.nc, self.TS, case.nnb).astype(np.int64).copy()
        # stencils the residual reads: all admissible ones of a FULL cell, the
        # first of the others (their reconstruction is MUSCL on the central
        # stencil; the sector stencils of such cells may be singular, with a
        # non-finite or meaningless pseudo-inverse) -- the rest are left out
        P = case.array("invmat").reshape(self.nc, self.TS, -1)
        full = case.array("full").astype(bool)
        self.solvable = np.isfinite(P).all(2) & (full[:, None] | (np.arange(self.TS) == 0)[None, :])
        self.per = np.array(case.period, dtype=np.float64)
        self.tables = {s: shape_table(s) for s in (HEX, TET, PRISM, PYRAMID)}
        self.nqm = max(case.qp_quad, case.qp_tri)
        self.quad_rs = case.array("quad_rs").reshape(-1, 2).copy()
        self.tri_bary = case.array("tri_bary").reshape(-1, 3).copy()
        xyz = case.array("xyz").reshape(-1, 3).copy()
        # discrete choices on the given mesh
        centre = case.array("centre").reshape(-1, 3).copy()
        self.centre0 = centre
        self.affine = np.zeros(self.nc, dtype=bool)
        for i in np.nonzero(self.shape == HEX)[0]:
            self.affine[i] = self._hex_affine(xyz[self.cnodes[i]])
        surf0 = case.array("surf").reshape(self.nc, self.NF)
        self.two_tri = np.zeros((self.nc, self.NF), dtype=bool)
        for sh, tab in self.tables.items():
            cells = np.nonzero(self.shape == sh)[0]
            for f, fn in enumerate(tab["faces"]):
                if len(fn) != 4 or len(cells) == 0:
                    continue
                v = xyz[self.cnodes[cells][:, fn]]
                a4 = 0.5 * np.linalg.norm(np.cross(v[:, 3] - v[:, 1], v[:, 2] - v[:, 0]), axis=1)
                self.two_tri[cells, f] = np.abs(surf0[cells, f] - a4) > 1e-14 * np.maximum(a4, 1e-300)

    @staticmethod
    def _hex_affine(v):
        scale = np.abs(v[6] - v[0]).max()
        tol = 1e-12 * max(scale, 1e-300)
        comb = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        for k, (a, b, c) in enumerate(comb):
            p = v[0] + a * (v[1] - v[0]) + b * (v[3] - v[0]) + c * (v[4] - v[0])
            if np.abs(p - v[k]).max() > tol:
                return False
        return True

    # -- building blocks -------------------------------------------------------
    def _shift(self, ref, pts):
        """whole periods that bring pts within half a period of ref (numpy,
        on the given mesh)"""
        s = np.zeros(pts.shape)
        for d in range(3):
            if self.per[d] <= 0:
                continue
            far = np.abs(pts[..., d] - ref[..., d]) > 0.5 * self.per[d]
            sg = np.where(ref[..., d] - 0.5 * self.per[d] >= 0, 1.0, -1.0)
            s[..., d] = np.where(far, self.per[d] * sg, 0.0)
        return s

    def _monos(self, y, nd):
        """raw monomials [.., nd] at points y [.., 3]"""
        torch = self.torch
        cols = []
        e = MONO[:nd]
        p = [[None] * 5 for _ in range(3)]
        for d in range(3):
            p[d][0] = torch.ones_like(y[..., d])
            for k in range(1, 5):
                p[d][k] = p[d][k - 1] * y[..., d]
        for k in range(nd):
            cols.append(p[0][e[k, 0]] * p[1][e[k, 1]] * p[2][e[k, 2]])
        return torch.stack(cols, -1)

    def _rule(self, shape, v, deg, affine=None):
        """integration points [n, np, 3] and weights [n, np] of cells of one
        shape with nodes v [n, nn, 3] (cell_rule: affine hexahedra by the
        tensor rule, others by their DECOMPOSE3 sub-tetrahedra)"""
        torch = self.torch
        tab = self.tables[shape]
        xs, ws = [], []
        if shape == HEX and affine is not None and affine.any():
            if not affine.all():
                raise ValueError("a batch mixes affine and non-affine hexahedra")
            n = max(1, (deg + 2) // 2)
            g, gw = _gauss_legendre(n)
            r = 0.5 * (g + 1)
            e1, e2, e3 = v[:, 1] - v[:, 0], v[:, 3] - v[:, 0], v[:, 4] - v[:, 0]
            det = (e1[:, 0] * (e2[:, 1] * e3[:, 2] - e2[:, 2] * e3[:, 1]) -
                   e1[:, 1] * (e2[:, 0] * e3[:, 2] - e2[:, 2] * e3[:, 0]) +
                   e1[:, 2] * (e2[:, 0] * e3[:, 1] - e2[:, 1] * e3[:, 0]))
            ad = det.abs()
            for a in range(n):
                for b in range(n):
                    for c in range(n):
                        xs.append(v[:, 0] + r[a] * e1 + r[b] * e2 + r[c] * e3)
                        ws.append(0.125 * gw[a] * gw[b] * gw[c] * ad)
            return torch.stack(xs, 1), torch.stack(ws, 1)
        for t in tab["subtets"]:
            tv = v[:, t]  # [n, 4, 3]
            e1, e2, e3 = tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0], tv[:, 3] - tv[:, 0]
            det = (e1[:, 0] * (e2[:, 1] * e3[:, 2] - e2[:, 2] * e3[:, 1]) -
                   e1[:, 1] * (e2[:, 0] * e3[:, 2] - e2[:, 2] * e3[:, 0]) +
                   e1[:, 2] * (e2[:, 0] * e3[:, 1] - e2[:, 1] * e3[:, 0]))
            if deg <= 3:
                vol = det.abs() / 6.0
                for p in range(5):
                    L = _STROUD_L[p]
                    xs.append(L[0] * tv[:, 0] + L[1] * tv[:, 1] + L[2] * tv[:, 2] + L[3] * tv[:, 3])
                    ws.append(_STROUD_W[p] * vol)
            else:
                n = (deg + 4) // 2
                g, gw = _gauss_legendre(n)
                ad = det.abs()
                for a in range(n):
                    for b in range(n):
                        for c in range(n):
                            u, s, tt = 0.5 * (g[a] + 1), 0.5 * (g[b] + 1), 0.5 * (g[c] + 1)
                            xi, eta, zeta = u, (1 - u) * s, (1 - u) * (1 - s) * tt
                            jac = (1 - u) * (1 - u) * (1 - s)
                            xs.append(tv[:, 0] + xi * e1 + eta * e2 + zeta * e3)
                            ws.append(0.125 * gw[a] * gw[b] * gw[c] * jac * ad)
        return torch.stack(xs, 1), torch.stack(ws, 1)

    @staticmethod
    def _frame(v):
        """LOCALISE_STEN2 frame of the first sub-tetrahedron v [n, 4, 3]:
        J = inverse of [v2-v1, v3-v1, v4-v1] (adjugate / determinant), origin
        v1, determinant"""
        import torch
        A = torch.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], -1)  # columns
        a = lambda r, c: A[:, r, c]  # noqa: E731
        det = (a(0, 0) * ((a(2, 2) * a(1, 1)) - (a(2, 1) * a(1, 2))) -
               a(1, 0) * ((a(2, 2) * a(0, 1)) - (a(2, 1) * a(0, 2))) +
               a(2, 0) * ((a(1, 2) * a(0, 1)) - (a(1, 1) * a(0, 2))))
        I = torch.stack([
            torch.stack([a(2, 2) * a(1, 1) - a(2, 1) * a(1, 2), -(a(2, 2) * a(0, 1) - a(2, 1) * a(0, 2)),
                         a(1, 2) * a(0, 1) - a(1, 1) * a(0, 2)], -1),
            torch.stack([-(a(2, 2) * a(1, 0) - a(2, 0) * a(1, 2)), a(2, 2) * a(0, 0) - a(2, 0) * a(0, 2),
                         -(a(1, 2) * a(0, 0) - a(1, 0) * a(0, 2))], -1),
            torch.stack([a(2, 1) * a(1, 0) - a(2, 0) * a(1, 1), -(a(2, 1) * a(0, 0) - a(2, 0) * a(0, 1)),
                         a(1, 1) * a(0, 0) - a(1, 0) * a(0, 1)], -1)], -2)
        return I / det[:, None, None], v[:, 0], det

    # -- the data ---------------------------------------------------------------
    def data(self, xyz, cells=None) -> dict:
        """the preprocessor's arrays of `cells` (default: all) as functions of
        the node coordinates xyz [nnode, 3] (a torch tensor)"""
        torch = self.torch
        if cells is None:
            cells = np.arange(self.nc)
        cells = np.asarray(cells, dtype=np.int64)
        n = len(cells)
        ID, IM, TS, NF, NQ = self.ID, self.IM, self.TS, self.NF, self.nqm
        dev = xyz.device
        T = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
        xyz_np = xyz.detach().cpu().numpy()
        out = {
            "invmat": torch.zeros(n, TS, IM, ID, dtype=xyz.dtype, device=dev),
            "indicator": torch.zeros(n, ID, ID, dtype=xyz.dtype, device=dev),
            "integ_basis": torch.zeros(n, ID, dtype=xyz.dtype, device=dev),
            "volume11": torch.zeros(n, dtype=xyz.dtype, device=dev),
            "qpoints": torch.zeros(n, NF, NQ, 3, dtype=xyz.dtype, device=dev),
            "surf": torch.zeros(n, NF, dtype=xyz.dtype, device=dev),
            "normal": torch.zeros(n, NF, 3, dtype=xyz.dtype, device=dev),
            "totvolume": torch.zeros(n, dtype=xyz.dtype, device=dev),
        }
        # volumes of every cell the chunk touches (its cells and their members)
        need = np.unique(np.concatenate([cells, self.ihexl[cells].reshape(-1)]))
        need = need[need >= 0]
        vol_of = {}
        vol_t = torch.zeros(self.nc, dtype=xyz.dtype, device=dev)
        for sh, tab in self.tables.items():
            sel = need[self.shape[need] == sh]
            if len(sel) == 0:
                continue
            v = xyz[T(self.cnodes[sel][:, :tab["nnodes"]])]
            vol = 0.0
            for t in tab["subtets"]:
                tv = v[:, t]
                e1, e2, e3 = tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0], tv[:, 3] - tv[:, 0]
                det = (e1[:, 0] * (e2[:, 1] * e3[:, 2] - e2[:, 2] * e3[:, 1]) -
                       e1[:, 1] * (e2[:, 0] * e3[:, 2] - e2[:, 2] * e3[:, 0]) +
                       e1[:, 2] * (e2[:, 0] * e3[:, 1] - e2[:, 1] * e3[:, 0]))
                vol = vol + det.abs() / 6.0
            vol_t = vol_t.index_put((T(sel),), vol)
            vol_of[sh] = sel
        out["totvolume"] = vol_t[T(cells)]
        # frames of the chunk's cells
        J = torch.zeros(n, 3, 3, dtype=xyz.dtype, device=dev)
        O = torch.zeros(n, 3, dtype=xyz.dtype, device=dev)
        D = torch.zeros(n, dtype=xyz.dtype, device=dev)
        for sh, tab in self.tables.items():
            loc = np.nonzero(self.shape[cells] == sh)[0]
            if len(loc) == 0:
                continue
            t0 = tab["subtets"][0]
            v = xyz[T(self.cnodes[cells[loc]][:, t0])]
            Ji, Oi, Di = self._frame(v)
            J = J.index_put((T(loc),), Ji)
            O = O.index_put((T(loc),), Oi)
            D = D.index_put((T(loc),), Di)
        adet = D.abs()
        out["volume11"] = out["totvolume"] / adet
        tolocal = lambda Jc, Oc, x: torch.einsum("nrs,n...s->n...r", Jc, x - Oc.reshape(  # noqa: E731
            Oc.shape[0], *([1] * (x.dim() - 2)), 3))
        c0 = self.centre0
        # own-cell data: INTEG_BASIS, INDICATOR, face points, areas, normals
        dts = _derivs(self.iorder)
        for sh, tab in self.tables.items():
            loc = np.nonzero(self.shape[cells] == sh)[0]
            if len(loc) == 0:
                continue
            ci = cells[loc]
            nn = tab["nnodes"]
            v = xyz[T(self.cnodes[ci][:, :nn])]
            vl = tolocal(J[T(loc)], O[T(loc)], v)
            aff = self.affine[ci] if sh == HEX else None
            for part in ([aff, ~aff] if aff is not None and aff.any() and not aff.all() else [None]):
                sub = np.arange(len(loc)) if part is None else np.nonzero(part)[0]
                pa = None if aff is None else aff[sub]
                x, w = self._rule(sh, vl[T(sub)], self.iorder, pa)
                ib = (w[..., None] * self._monos(x, ID)).sum(1)
                out["integ_basis"] = out["integ_basis"].index_put((T(loc[sub]),), ib)
                if self.iweno == 1:
                    x, w = self._rule(sh, vl[T(sub)], 2 * (self.iorder - 1), pa)
                    B = 0.0
                    for al in dts:
                        dm = self._dmonos(x, ID, al)  # [m, np, ID]
                        B = B + torch.einsum("mp,mpk,mpl->mkl", w, dm, dm)
                    out["indicator"] = out["indicator"].index_put((T(loc[sub]),), B)
            for f, fn in enumerate(tab["faces"]):
                fv = v[:, fn]  # own orientation
                if len(fn) == 4:
                    cr = torch.cross(fv[:, 3] - fv[:, 1], fv[:, 2] - fv[:, 0], dim=1)
                    area = 0.5 * torch.linalg.norm(cr, dim=1)
                    tt = self.two_tri[ci, f]
                    if tt.any():
                        a2 = (0.5 * torch.linalg.norm(torch.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0], dim=1), dim=1)
                              + 0.5 * torch.linalg.norm(torch.cross(fv[:, 2] - fv[:, 0], fv[:, 3] - fv[:, 0], dim=1),
                                                        dim=1))
                        area = torch.where(T(tt), a2, area)
                else:
                    area = 0.5 * torch.linalg.norm(torch.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0], dim=1), dim=1)
                nv = len(fn)
                nx = sum((fv[:, k, 1] - fv[:, (k + 1) % nv, 1]) * (fv[:, k, 2] + fv[:, (k + 1) % nv, 2])
                         for k in range(nv))
                ny = sum((fv[:, k, 2] - fv[:, (k + 1) % nv, 2]) * (fv[:, k, 0] + fv[:, (k + 1) % nv, 0])
                         for k in range(nv))
                nz = sum((fv[:, k, 0] - fv[:, (k + 1) % nv, 0]) * (fv[:, k, 1] + fv[:, (k + 1) % nv, 1])
                         for k in range(nv))
                nrm = torch.stack([nx, ny, nz], 1)
                nrm = nrm / torch.linalg.norm(nrm, dim=1, keepdim=True)
                out["surf"] = out["surf"].index_put((T(loc), T(np.full(len(loc), f))), area)
                out["normal"] = out["normal"].index_put((T(loc), T(np.full(len(loc), f))), nrm)
                # Gauss points from the adopted face vertices, unwrapped around the centre
                fidx = self.fnodes[ci, f, :nv]
                shift = self._shift(c0[ci][:, None, :], xyz_np[fidx])
                lv = tolocal(J[T(loc)], O[T(loc)], xyz[T(fidx)] + T(shift))
                if nv == 4:
                    R, S = self.quad_rs[:, 0], self.quad_rs[:, 1]
                    N = np.stack([0.25 * (1 - R) * (1 - S), 0.25 * (1 + R) * (1 - S), 0.25 * (1 + R) * (1 + S),
                                  0.25 * (1 - R) * (1 + S)], 1)  # [nq, 4]
                else:
                    N = self.tri_bary
                Q = torch.einsum("qk,nkd->nqd", T(N), lv)
                qf = torch.cat([Q, torch.zeros(len(loc), NQ - N.shape[0], 3, dtype=xyz.dtype, device=dev)], 1)
                out["qpoints"] = out["qpoints"].index_put((T(loc), T(np.full(len(loc), f))), qf)
        # pseudo-inverses: members' monomial integrals in each cell's frame
        mbar = out["integ_basis"] / out["volume11"][:, None]
        A = torch.zeros(n, TS, IM, ID, dtype=xyz.dtype, device=dev)
        mem = self.ihexl[cells][:, :, 1:]  # [n, TS, IM]
        act = (np.arange(TS)[None, :] < self.admis[cells][:, None]) & self.solvable[cells]  # [n, TS]
        ii, ss, kk = np.nonzero(np.broadcast_to(act[:, :, None], mem.shape))
        jj = mem[ii, ss, kk]
        for sh, tab in self.tables.items():
            sel = np.nonzero(self.shape[jj] == sh)[0]
            if len(sel) == 0:
                continue
            nn = tab["nnodes"]
            i_, j_ = ii[sel], jj[sel]
            shift = self._shift(c0[cells[i_]], c0[j_])  # by the members' centres
            v = xyz[T(self.cnodes[j_][:, :nn])] + T(shift)[:, None, :]
            vl = tolocal(J[T(i_)], O[T(i_)], v)
            aff = self.affine[j_] if sh == HEX else None
            parts = [aff, ~aff] if aff is not None and aff.any() and not aff.all() else [None]
            for part in parts:
                sub = np.arange(len(sel)) if part is None else np.nonzero(part)[0]
                pa = None if aff is None else aff[sub]
                x, w = self._rule(sh, vl[T(sub)], self.iorder, pa)
                intg = (w[..., None] * self._monos(x, ID)).sum(1)
                vq = w.sum(1)
                row = intg - mbar[T(i_[sub])] * vq[:, None]
                A = A.index_put((T(i_[sub]), T(ss[sel][sub]), T(kk[sel][sub])), row)
        volj = torch.zeros(n, TS, IM, dtype=xyz.dtype, device=dev)
        volj = volj.index_put((T(ii), T(ss), T(kk)), vol_t[T(jj)] / adet[T(ii)])
        M = torch.einsum("nskd,nske->nsde", A, A)
        eye = torch.eye(ID, dtype=xyz.dtype, device=dev)
        M = torch.where(T(act)[:, :, None, None], M, eye)
        X = torch.linalg.solve(M, A.transpose(-1, -2))  # [n, TS, ID, IM] = M^-1 A^T
        P = X.transpose(-1, -2) * volj[..., None]
        out["invmat"] = torch.where(T(act)[:, :, None, None], P, torch.zeros_like(P))
        return out

    def _dmonos(self, y, nd, al):
        """D^al of the raw monomials at points y"""
        torch = self.torch
        cols = []
        p = [[None] * 5 for _ in range(3)]
        for d in range(3):
            p[d][0] = torch.ones_like(y[..., d])
            for k in range(1, 5):
                p[d][k] = p[d][k - 1] * y[..., d]
        for k in range(nd):
            e = MONO[k]
            if (e < al).any():
                cols.append(torch.zeros_like(y[..., 0]))
                continue
            coef = 1.0
            for d in range(3):
                coef *= math.perm(int(e[d]), int(al[d]))
            cols.append(coef * p[0][e[0] - al[0]] * p[1][e[1] - al[1]] * p[2][e[2] - al[2]])
        return torch.stack(cols, -1)

    def vjp_nodes(self, gbar: dict, xyz, chunk: int = 4096):
        """sum over the arrays of <gbar, d data / d xyz>: the cotangent of the
        node coordinates [nnode, 3] (numpy), chunk cells at a time"""
        torch = self.torch
        x = torch.as_tensor(xyz, dtype=torch.float64, device=self.dev).clone().requires_grad_(True)
        total = torch.zeros_like(x)
        keys = ["invmat", "indicator", "integ_basis", "volume11", "qpoints", "surf", "normal", "totvolume"]
        for c0 in range(0, self.nc, chunk):
            cells = np.arange(c0, min(self.nc, c0 + chunk))
            d = self.data(x, cells)
            s = 0.0
            for k in keys:
                if k not in gbar or (k == "indicator" and self.iweno != 1):
                    continue
                g = torch.as_tensor(np.asarray(gbar[k])[cells], dtype=torch.float64, device=self.dev)
                s = s + (g * d[k]).sum()
            (gx,) = torch.autograd.grad(s, x)
            total += gx
        return total.cpu().numpy()


def bump_basis(case) -> np.ndarray:
    """d xyz / d a_k of the channel's lower-wall bump [8, nnode, 3]: with
    x^ = 4x/lx, z^ = z/lz the nodes move by dy = ly (1 - y/ly) sum_k a_k
    sin(k pi (x^-1)/2) sin(pi z^) for 1 <= x^ <= 3 (y, x, z before the bump;
    the map is linear in a)"""
    lx, ly, lz = case.spec.dims[3:]
    p = case.array("xyz_box").reshape(-1, 3)
    xh, zh = 4.0 * p[:, 0] / lx, p[:, 2] / lz
    inside = (xh >= 1.0) & (xh <= 3.0)
    out = np.zeros((8, len(p), 3))
    for q in range(8):
        out[q, :, 1] = np.where(inside, ly * (1.0 - p[:, 1] / ly) * np.sin((q + 1) * np.pi * (xh - 1.0) / 2.0)
                                * np.sin(np.pi * zh), 0.0)
    return out


def shape_gradient(case, gbar: dict, device: str = "cpu", chunk: int = 4096) -> np.ndarray:
    """dJ/da_k (k = 1..8) of the channel bump from the geometry cotangents of
    ucns_rk3_adjoint_geometry (plus any explicit dependence of J on the data,
    added to gbar by the caller)"""
    chain = ShapeChain(case, device)
    gx = chain.vjp_nodes(gbar, case.array("xyz").reshape(-1, 3), chunk)
    return np.einsum("knd,nd->k", bump_basis(case), gx)
