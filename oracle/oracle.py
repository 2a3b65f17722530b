"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return os.path.join(HERE, "liboracle.so")


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        vp = C.c_void_p
        f64 = C.POINTER(C.c_double)
        i64 = C.POINTER(C.c_int64)
        L.oracle_residual.argtypes = [vp, vp, vp, f64, f64, i64]
        L.oracle_reconstruct.argtypes = [vp, vp, vp, f64, f64, i64]
        L.oracle_fluxes.argtypes = [vp, vp, vp, f64, f64]
        L.oracle_dt.argtypes = [vp, vp, f64, C.c_double, f64]
        L.oracle_rk3_steps.argtypes = [vp, vp, vp, f64, C.c_int32, C.c_double, C.c_double, f64]
        L.oracle_hllc.argtypes = [C.c_double, f64, f64, f64]
        L.oracle_hllc.restype = None
        L.oracle_rotatef.argtypes = [f64, f64, C.c_double, C.c_double]
        L.oracle_rotatef.restype = None
        L.oracle_rotateb.argtypes = [f64, f64, C.c_double, C.c_double]
        L.oracle_rotateb.restype = None
        L.oracle_weno_weights.argtypes = [C.c_int, f64, C.c_double, C.c_double, f64]
        L.oracle_weno_weights.restype = None
        L.oracle_barth.argtypes = [C.c_double] * 4
        L.oracle_barth.restype = C.c_double
        L.oracle_set_boundary.argtypes = [vp]
        L.oracle_set_boundary.restype = None
        L.oracle_set_cweno.argtypes = [vp]
        L.oracle_set_cweno.restype = None
        L.oracle_cweno_weights.argtypes = [C.c_int, f64, C.c_double, C.c_double, C.c_int, f64]
        L.oracle_cweno_weights.restype = None
        L.oracle_set_face_order.argtypes = [C.c_int]
        L.oracle_set_face_order.restype = None
        L.oracle_set_uleft_form.argtypes = [C.c_int]
        L.oracle_set_uleft_form.restype = None
        L.oracle_set_face_once.argtypes = [C.c_int]
        L.oracle_set_face_once.restype = None
        L.oracle_set_raw.argtypes = [C.c_int]
        L.oracle_set_raw.restype = None
        L.oracle_set_boundary_faces.argtypes = [vp, vp]
        L.oracle_set_boundary_faces.restype = None
        L.oracle_char_eigen.argtypes = [C.c_double, f64, f64, f64, f64]
        L.oracle_char_eigen.restype = None
        L.oracle_ghost.argtypes = [C.c_int, C.c_double, f64, C.c_double, C.c_double, f64]
        L.oracle_ghost.restype = None
        L.oracle_residual_tl.argtypes = [vp, vp, vp, f64, f64, f64, f64, f64]
        L.oracle_rk3_tl.argtypes = [vp, vp, vp, f64, f64, f64, C.c_int32, C.c_double]
        L.oracle_hllc_tl.argtypes = [C.c_double, f64, f64, f64, f64, f64, f64]
        L.oracle_hllc_tl.restype = None
        L.oracle_ghost_tl.argtypes = [C.c_int, C.c_double, f64, f64, f64, C.c_double, C.c_double, f64, f64]
        L.oracle_ghost_tl.restype = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Oracle bound to one case's ucns_*_v1 structs."""

    def __init__(self, case, precision=0):
        self.case = case
        self.m, self.s, self.p = case.structs(precision)
        self.cw = case.cweno() if hasattr(case, "cweno") else None
        self.L = lib()
        nqm = max(case.qp_quad, case.qp_tri)
        self.uleft_shape = (case.ncell, case.nface_max, nqm, 5)

    def _refs(self):
        # the CWENO data of this case (or none) for the calls that follow
        self.L.oracle_set_cweno(C.byref(self.cw) if self.cw is not None else None)
        return C.byref(self.m), C.byref(self.s), C.byref(self.p)

    def reconstruct(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        ul = np.zeros(self.uleft_shape)
        cnt = np.zeros(2, dtype=np.int64)
        m, s, p = self._refs()
        rc = self.L.oracle_reconstruct(m, s, p, _p(u), _p(ul), cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        assert rc == 0
        return ul, cnt

    def residual(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        rhs = np.zeros_like(u)
        cnt = np.zeros(2, dtype=np.int64)
        m, s, p = self._refs()
        rc = self.L.oracle_residual(m, s, p, _p(u), _p(rhs), cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        assert rc == 0
        return rhs, cnt

    def dt(self, u, cfl):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = C.c_double()
        m, s, _ = self._refs()
        self.L.oracle_dt(m, s, _p(u), cfl, C.byref(out))
        return out.value

    def rk3(self, u, nsteps, cfl=0.5, dt_fixed=0.0, t0=0.0):
        u = np.array(u, dtype=np.float64, copy=True)
        t = C.c_double(t0)
        m, s, p = self._refs()
        rc = self.L.oracle_rk3_steps(m, s, p, _p(u), nsteps, cfl, dt_fixed, C.byref(t))
        assert rc == 0
        return u, t.value

    def residual_tl(self, u, du, dbc=None):
        """tangent-linear residual: (R(u), J(u) du + dR/dbc dbc); dbc is the
        tangent of the boundary data (inflow[5], outflow[5])"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.ascontiguousarray(du, dtype=np.float64)
        rhs, drhs = np.zeros_like(u), np.zeros_like(u)
        m, s, p = self._refs()
        db = None if dbc is None else np.ascontiguousarray(dbc, dtype=np.float64)
        rc = self.L.oracle_residual_tl(m, s, p, _p(u), _p(du), None if db is None else _p(db), _p(rhs), _p(drhs))
        assert rc == 0
        return rhs, drhs

    def rk3_tl(self, u, du, nsteps, dt, dbc=None):
        """nsteps SSP-RK3 steps of fixed dt with the tangent: (u_N, du_N)"""
        u = np.array(u, dtype=np.float64, copy=True)
        du = np.array(du, dtype=np.float64, copy=True)
        m, s, p = self._refs()
        db = None if dbc is None else np.ascontiguousarray(dbc, dtype=np.float64)
        rc = self.L.oracle_rk3_tl(m, s, p, _p(u), _p(du), None if db is None else _p(db), nsteps, dt)
        assert rc == 0
        return u, du


def hllc(gamma, cl, cr):
    cl = np.ascontiguousarray(cl, dtype=np.float64)
    cr = np.ascontiguousarray(cr, dtype=np.float64)
    out = np.zeros(5)
    lib().oracle_hllc(gamma, _p(cl), _p(cr), _p(out))
    return out


def rotatef(vec, a1, a2):
    vec = np.ascontiguousarray(vec, dtype=np.float64)
    out = np.zeros(5)
    lib().oracle_rotatef(_p(out), _p(vec), a1, a2)
    return out


def rotateb(vec, a1, a2):
    vec = np.ascontiguousarray(vec, dtype=np.float64)
    out = np.zeros(5)
    lib().oracle_rotateb(_p(out), _p(vec), a1, a2)
    return out


def weno_weights(beta, lwci1, eps=1e-11):
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    out = np.zeros(len(beta))
    lib().oracle_weno_weights(len(beta), _p(beta), lwci1, eps, _p(out))
    return out


def barth(ui, umin, umax, usol):
    return lib().oracle_barth(ui, umin, umax, usol)


def cweno_weights(beta, lwci1, eps=1e-11, wenoz=0):
    """CWENO weights of one variable for the stencils' indicators beta (central first)"""
    b = np.ascontiguousarray(beta, dtype=np.float64)
    out = np.zeros_like(b)
    lib().oracle_cweno_weights(len(b), _p(b), lwci1, eps, wenoz, _p(out))
    return out


def set_boundary(boundtype, inflow, outflow):
    """free-stream data the oracle's boundary faces use (process-wide)"""
    from ucns3d_amd import native
    b = native.BoundaryV1(abi_version=native.UCNS_ABI_VERSION, boundtype=boundtype,
                          inflow=(C.c_double * 5)(*inflow), outflow=(C.c_double * 5)(*outflow))
    lib().oracle_set_boundary(C.byref(b))


def set_face_order(reversed_: bool) -> None:
    """sum each cell's faces in reverse slot order (process-wide): the same
    residual in exact arithmetic, rounded differently"""
    lib().oracle_set_face_order(1 if reversed_ else 0)


def set_uleft_form(mean_first: bool) -> None:
    """form each WENO face value as U_i + sum_s w_s res_s instead of the
    reference's sum_s w_s (U_i + res_s) (process-wide): equal in exact
    arithmetic (the weights sum to 1), rounded differently"""
    lib().oracle_set_uleft_form(1 if mean_first else 0)


def set_face_once(once: bool) -> None:
    """solve each interior face once from its owner and negate it for the
    neighbour, instead of from both sides with each side's own face angles
    and area (process-wide): equal in exact arithmetic, rounded differently"""
    lib().oracle_set_face_once(1 if once else 0)


_BFACES = None


def set_boundary_faces(inflow=None, outflow=None) -> None:
    """per-face INFLOW() / OUTFLOW() states [ncell][nface_max][5] for the
    ghosts of codes 1, 2, 6 (process-wide; the arrays are held until the next
    call); None: the uniform state of set_boundary for that side"""
    global _BFACES
    fi = None if inflow is None else np.ascontiguousarray(inflow, dtype=np.float64)
    fo = None if outflow is None else np.ascontiguousarray(outflow, dtype=np.float64)
    _BFACES = (fi, fo)
    lib().oracle_set_boundary_faces(None if fi is None else fi.ctypes.data, None if fo is None else fo.ctypes.data)


def set_raw(on: bool) -> None:
    """reconstruct without the troubled-cell logic (CHECKSOL, MUSCL,
    CHECKSOLX): the WENO face values alone (process-wide)"""
    lib().oracle_set_raw(1 if on else 0)


def char_eigen(gamma, rl, rr):
    """(R, L): the eigenvectors of characteristic WENO at the average of the
    two rotated states rl, rr"""
    rl = np.ascontiguousarray(rl, dtype=np.float64)
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    R, L = np.zeros(25), np.zeros(25)
    lib().oracle_char_eigen(gamma, _p(rl), _p(rr), _p(R), _p(L))
    return R.reshape(5, 5), L.reshape(5, 5)


def ghost(code, gamma, cl, a1, a2):
    cl = np.ascontiguousarray(cl, dtype=np.float64)
    out = np.zeros(5)
    lib().oracle_ghost(code, gamma, _p(cl), a1, a2, _p(out))
    return out


def hllc_tl(gamma, cl, dcl, cr, dcr):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cl, dcl, cr, dcr)]
    f, df = np.zeros(5), np.zeros(5)
    lib().oracle_hllc_tl(gamma, *[_p(x) for x in a], _p(f), _p(df))
    return f, df


def ghost_tl(code, gamma, cl, dcl, a1, a2, dbc=None):
    cl = np.ascontiguousarray(cl, dtype=np.float64)
    dcl = np.ascontiguousarray(dcl, dtype=np.float64)
    db = None if dbc is None else np.ascontiguousarray(dbc, dtype=np.float64)
    g, dg = np.zeros(5), np.zeros(5)
    lib().oracle_ghost_tl(code, gamma, _p(cl), _p(dcl), None if db is None else _p(db), a1, a2, _p(g), _p(dg))
    return g, dg
