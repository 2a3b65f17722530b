"""Host-side mirror of the reference's time-marching interface over the C-ABI.

    RUNGE_KUTTA3 / TIME_MARCHING  (src/time_p.f90:885-1027, 4026-4580) -> Solver.rk3_steps
    CALL_FLUX_SUBROUTINES_3D       (src/time_p.f90:2253-2497)            -> Solver.residual
    ARBITRARY_ORDER                (src/reconstruct_p.f90:2954-3015)     -> Solver.reconstruct
    CALCULATE_CFL                  (src/time_p.f90:27-191)               -> Solver.compute_dt

and the adjoint of the path (no counterpart in the reference):
    (dR/dU)^T w                                                          -> Solver.residual_vjp
    (dU_N/dU_0)^T uN_bar over RK3 steps with stage checkpoints           -> Solver.rk3_adjoint

Every call goes through libucnsgpu.so (include/ucns.h); there is no CPU
fallback: constructing a Solver fails if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .native import UCNS_FP64, UCNS_LAYOUT_AOS, UCNS_MIXED  # noqa: F401


class UcnsError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


class Solver:
    """One library context over a Case.

    cell_order, kernels, adj_keep_stages and adj_keep_gib are the library
    options of ucns_scheme_v2 (UCNS_ORDER_*, UCNS_KERNELS_*; 0 = defaults);
    `structs` replaces the Case's (mesh, scheme, prestore) structs, e.g. with
    a caller order of the cells of one's own."""

    def __init__(self, case, precision: int = UCNS_FP64, lib_path: str | None = None, cell_order: int = 0,
                 kernels: int = 0, adj_keep_stages: int = 0, adj_keep_gib: float = 0.0, structs=None,
                 cweno=None):
        self.case = case
        self.precision = precision
        self.L = native.gpu_lib(lib_path)
        self._m, self._s, self._p = structs if structs is not None else case.structs(precision)
        self._s.cell_order = cell_order
        self._s.kernels = kernels
        self._s.adj_keep_stages = adj_keep_stages
        self._s.adj_keep_gib = adj_keep_gib
        h = C.c_void_p()
        rc = self.L.ucns_create(C.byref(self._m), C.byref(self._s), C.byref(h))
        self._h = h
        self._check("ucns_create", rc)
        self._cw = cweno if cweno is not None else (case.cweno() if structs is None else None)
        if self._cw is not None:
            self._check("ucns_upload_cweno", self.L.ucns_upload_cweno(self._h, C.byref(self._cw)))
        rc = self.L.ucns_upload_prestore(self._h, C.byref(self._p))
        self._check("ucns_upload_prestore", rc)
        self.ncell = case.ncell
        self.uleft_shape = (case.ncell, case.nface_max, max(case.qp_quad, case.qp_tri), 5)

    def _check(self, fn: str, rc: int) -> None:
        if rc != 0:
            msg = self.L.ucns_last_error(self._h).decode() if self._h else ""
            raise UcnsError(fn, rc, msg)

    # -- boundaries --------------------------------------------------------------
    def set_boundary(self, boundtype: int, inflow, outflow) -> None:
        """free-stream data of the inflow/outflow/farfield codes (BOUNDARYS,
        src/flow_operations.f90:2860-3375); conserved 5-vectors"""
        b = native.BoundaryV1(abi_version=native.UCNS_ABI_VERSION, boundtype=boundtype,
                              inflow=(C.c_double * 5)(*inflow), outflow=(C.c_double * 5)(*outflow))
        self._check("ucns_set_boundary", self.L.ucns_set_boundary(self._h, C.byref(b)))

    def set_boundary_faces(self, inflow=None, outflow=None) -> None:
        """position-dependent INFLOW() / OUTFLOW() states per face record
        [ncell][nface_max][5] (read at the faces of codes 1, 2, 6); None keeps
        the uniform state of that side"""
        def arr(a):
            if a is None:
                return None
            return np.ascontiguousarray(a, dtype=np.float64).reshape(self.ncell, self.case.nface_max, 5)
        fi, fo = arr(inflow), arr(outflow)
        self._check("ucns_set_boundary_faces",
                    self.L.ucns_set_boundary_faces(self._h, native.f64p(fi) if fi is not None else None,
                                                   native.f64p(fo) if fo is not None else None))

    # -- state -----------------------------------------------------------------
    def set_state(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.ncell, 5)
        self._check("ucns_set_state", self.L.ucns_set_state(self._h, native.f64p(u), UCNS_LAYOUT_AOS))

    def get_state(self) -> np.ndarray:
        u = np.empty((self.ncell, 5))
        self._check("ucns_get_state", self.L.ucns_get_state(self._h, native.f64p(u), UCNS_LAYOUT_AOS))
        return u

    def set_state_device(self, ptr: int) -> None:
        self._check("ucns_set_state_device", self.L.ucns_set_state_device(self._h, C.c_void_p(ptr), UCNS_LAYOUT_AOS))

    def get_state_device(self, ptr: int) -> None:
        self._check("ucns_get_state_device", self.L.ucns_get_state_device(self._h, C.c_void_p(ptr), UCNS_LAYOUT_AOS))

    # -- path ------------------------------------------------------------------
    def residual(self) -> np.ndarray:
        rhs = np.empty((self.ncell, 5))
        self._check("ucns_residual", self.L.ucns_residual(self._h, native.f64p(rhs), UCNS_LAYOUT_AOS))
        return rhs

    def residual_device(self, ptr: int) -> None:
        self._check("ucns_residual_device", self.L.ucns_residual_device(self._h, C.c_void_p(ptr), UCNS_LAYOUT_AOS))

    def debug_fail_alloc(self, nth: int) -> None:
        self._check("ucns_debug_fail_alloc", self.L.ucns_debug_fail_alloc(self._h, nth))

    def reconstruct(self) -> np.ndarray:
        ul = np.empty(self.uleft_shape)
        self._check("ucns_reconstruct", self.L.ucns_reconstruct(self._h, native.f64p(ul)))
        return ul

    def compute_dt(self, cfl: float) -> float:
        dt = C.c_double()
        self._check("ucns_compute_dt", self.L.ucns_compute_dt(self._h, cfl, C.byref(dt)))
        return dt.value

    def rk3_steps(self, nsteps: int, cfl: float = 0.5, dt_fixed: float = 0.0, t0: float = 0.0) -> float:
        t = C.c_double(t0)
        self._check("ucns_rk3_steps", self.L.ucns_rk3_steps(self._h, nsteps, cfl, dt_fixed, C.byref(t)))
        return t.value

    # -- adjoint ---------------------------------------------------------------
    def residual_vjp(self, w: np.ndarray):
        """(dR/dU)^T w at the current state, and (dR/d(inflow, outflow))^T w"""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ncell, 5)
        ub = np.empty((self.ncell, 5))
        bc = np.zeros(10)
        self._check("ucns_residual_vjp", self.L.ucns_residual_vjp(self._h, native.f64p(w), native.f64p(ub),
                                                                   UCNS_LAYOUT_AOS, native.f64p(bc)))
        return ub, bc

    def residual_vjp_device(self, w_ptr: int, ubar_ptr: int, bc_ptr: int | None = None) -> None:
        self._check("ucns_residual_vjp_device",
                    self.L.ucns_residual_vjp_device(self._h, C.c_void_p(w_ptr), C.c_void_p(ubar_ptr),
                                                    C.c_void_p(bc_ptr) if bc_ptr else None))

    def rk3_adjoint(self, nsteps: int, uN_bar: np.ndarray, cfl: float = 0.5, dt_fixed: float = 0.0,
                    t0: float = 0.0):
        """nsteps RK3 steps from the current state (left at U_N) and the
        cotangents (dU_N/dU_0)^T uN_bar and (dU_N/d(inflow, outflow))^T uN_bar"""
        a = np.ascontiguousarray(uN_bar, dtype=np.float64).reshape(self.ncell, 5)
        u0b = np.empty((self.ncell, 5))
        bc = np.zeros(10)
        t = C.c_double(t0)
        self._check("ucns_rk3_adjoint", self.L.ucns_rk3_adjoint(self._h, nsteps, cfl, dt_fixed, native.f64p(a),
                                                                 native.f64p(u0b), UCNS_LAYOUT_AOS,
                                                                 native.f64p(bc), C.byref(t)))
        return u0b, bc, t.value

    # shapes of the geometry cotangents, in the layouts of ucns_prestore_v1 / ucns_mesh_v2
    def geometry_shapes(self) -> dict:
        c = self.case
        nqm = max(c.qp_quad, c.qp_tri)
        return {"invmat": (c.ncell, c.typesten, c.imax, c.ideg), "indicator": (c.ncell, c.ideg, c.ideg),
                "integ_basis": (c.ncell, c.ideg), "volume11": (c.ncell,),
                "qpoints": (c.ncell, c.nface_max, nqm, 3), "surf": (c.ncell, c.nface_max),
                "normal": (c.ncell, c.nface_max, 3), "totvolume": (c.ncell,)}

    def rk3_adjoint_geometry(self, nsteps: int, uN_bar: np.ndarray, dt_fixed: float, t0: float = 0.0):
        """ucns_rk3_adjoint_geometry: nsteps RK3 steps of dt_fixed from the
        current state (left at U_N) and the cotangents of uN_bar . U_N with
        respect to U_0, the boundary data and the mesh/PRESTORE data, the
        last as a dict of arrays named after ucns_geometry_bar_v1's fields"""
        a = np.ascontiguousarray(uN_bar, dtype=np.float64).reshape(self.ncell, 5)
        u0b = np.empty((self.ncell, 5))
        bc = np.zeros(10)
        out = {k: np.empty(shp) for k, shp in self.geometry_shapes().items()}
        g = native.GeometryBarV1(abi_version=native.UCNS_ABI_VERSION, reserved=0,
                                 **{k: native.f64p(v) for k, v in out.items()})
        t = C.c_double(t0)
        self._check("ucns_rk3_adjoint_geometry",
                    self.L.ucns_rk3_adjoint_geometry(self._h, nsteps, dt_fixed, native.f64p(a), native.f64p(u0b),
                                                     UCNS_LAYOUT_AOS, native.f64p(bc), C.byref(g), C.byref(t)))
        return u0b, bc, out, t.value

    def rk3_adjoint_device(self, nsteps: int, uN_bar_ptr: int, u0_bar_ptr: int, bc_ptr: int | None = None,
                           cfl: float = 0.5, dt_fixed: float = 0.0, t0: float = 0.0) -> float:
        t = C.c_double(t0)
        self._check("ucns_rk3_adjoint_device",
                    self.L.ucns_rk3_adjoint_device(self._h, nsteps, cfl, dt_fixed, C.c_void_p(uN_bar_ptr),
                                                   C.c_void_p(u0_bar_ptr), C.c_void_p(bc_ptr) if bc_ptr else None,
                                                   C.byref(t)))
        return t.value

    def reset_stats(self) -> None:
        self._check("ucns_reset_stats", self.L.ucns_reset_stats(self._h))

    def stats(self) -> dict:
        st = native.StatsV1()
        self._check("ucns_stats", self.L.ucns_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in native.StatsV1._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.L.ucns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
