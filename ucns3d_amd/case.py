"""Synthetic cases: mesh + scheme + PRESTORE arrays built by the host preprocessor.

A Case owns the native preprocessor object (libucnsprep.so) and hands out
the ucns_*_v1 structs of include/ucns.h that point into its storage, so the
GPU library (and any other consumer of the C-ABI) receives exactly what a
Fortran host would pass.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from dataclasses import dataclass

import numpy as np

from . import native

HEX = 1
TET = 2
PRISM = 3
PYRAMID = 4
HYBRID = 5  # hexes, pyramids and prisms in thirds of x
CHANNEL = 6  # prism wall layers, a pyramid/tetrahedron interface layer, tetrahedra above
BLAST = 7  # a central box of hexahedra, a pyramid/tetrahedron shell, tetrahedra, prism layers at the z ends


@dataclass(frozen=True)
class CaseSpec:
    """Mesh and scheme choices of one configuration."""
    name: str
    shape: int            # HEX, TET (6 Kuhn tets per hex), PRISM (2 per hex), PYRAMID (6 per hex) or HYBRID
    n: int                # hexes per direction
    scheme: str           # "muscl" (2nd order, Barth), "weno3" (IORDER 2), "weno4" (IORDER 3) or "weno5" (IORDER 4)
    jitter: float = 0.0   # interior vertex jitter, fraction of the spacing
    seed: int = 1
    noise: float = 0.0    # N(0,1) density perturbation amplitude per cell
    cfl: float = 0.5
    steps: int = 100
    length: float = 2.0 * math.pi
    lwci1: float = 10000.0   # wenocentralweight of tests/taylor_green_vortex/UCNS3D.DAT
    gamma: float = 1.4
    bc: tuple = (0, 0, 0, 0, 0, 0)  # codes of the x-, x+, y-, y+, z-, z+ sides; 0 = periodic axis
    bump: tuple = ()      # lower-wall bump coefficients a_1..a_8 (channel configuration)
    box: tuple = ()       # (nx, ny, nz, lx, ly, lz) for a box other than n^3 cubes of side `length`
    mach: float = 0.0     # > 0: the initial state is a uniform flow along x at this Mach number, rho = p = 1
    ees: int = 0          # 5: CWENO (central polynomial reconstituted from low-order sector polynomials)
    wenoz: int = 0        # 1: CWENO-Z weights (with ees = 5)
    wenwrt: int = 1       # 2: WENO in characteristic variables, face by face (WENWRT 2)
    init: str = ""        # "blast": rho = 8, p = 10 for cell centres within 0.2 of the box centre, rho = p = 1
                          # elsewhere, at rest; otherwise the vortex (or the free stream when mach > 0)

    @property
    def iorder(self) -> int:
        return {"muscl": 1, "weno3": 2, "weno4": 3, "weno5": 4}[self.scheme]

    @property
    def iweno(self) -> int:
        return -1 if self.scheme == "muscl" else 1

    @property
    def extf(self) -> int:
        # CODE_PROFILE 0 (extf=3) for the MUSCL case, profile 8 (extf=2) for WENO
        return 3 if self.scheme == "muscl" else 2

    @property
    def dims(self):
        """(nx, ny, nz, lx, ly, lz)"""
        return tuple(self.box) if self.box else (self.n, self.n, self.n, self.length, self.length, self.length)

    @property
    def ncell(self) -> int:
        nx, ny, nz = self.dims[:3]
        if self.shape == HYBRID:
            return ny * nz * sum(1 if i < nx // 3 else 6 if i < 2 * nx // 3 else 2 for i in range(nx))
        if self.shape == CHANNEL:
            layers = max(1, min(10, ny // 3))
            return nx * nz * (2 * layers + 7 + 6 * (ny - layers - 1))
        if self.shape == BLAST:
            core = [n - 2 * (n // 4) for n in (nx, ny, nz)]
            box = [c + 2 for c in core]
            ncore, nbox = math.prod(core), math.prod(box)
            # shell hexes: 6 pyramids, one more cell per face on the shell box's outside
            shell = 6 * (nbox - ncore) + 2 * (box[0] * box[1] + box[1] * box[2] + box[0] * box[2])
            pl = max(1, nz // 10)
            nprism = nx * ny * 2 * pl
            ntet = nx * ny * nz - nbox - nx * ny * 2 * pl
            return ncore + shell + 2 * nprism + 6 * ntet
        return nx * ny * nz * {HEX: 1, TET: 6, PRISM: 2, PYRAMID: 6}[self.shape]

    def freestream(self):
        """conserved state of the uniform flow along x at Mach `mach`, rho = p = 1"""
        u = self.mach * math.sqrt(self.gamma)
        return np.array([1.0, u, 0.0, 0.0, 1.0 / (self.gamma - 1.0) + 0.5 * u * u])


# The benchmark configurations (sizes as specified) and reduced versions of
# the same construction used by the parity tests.
PRESETS = {
    "c1": CaseSpec("c1: 24^3 hex, MUSCL+Barth, HLLC, SSP-RK3", HEX, 24, "muscl", seed=1, steps=100),
    "c2": CaseSpec("c2: 64^3x6 jittered tets, WENO-4, HLLC, SSP-RK3, FP64", TET, 64, "weno4", jitter=0.1, seed=2,
                   steps=200),
    "c3": CaseSpec("c3: 128^3 hex + 1e-3 density noise, WENO-4, HLLC, SSP-RK3, mixed", HEX, 128, "weno4", seed=3,
                   noise=1e-3, steps=500),
    # [0,4]x[0,1]x[0,1], 120 x 30 x 30 hexes split into prisms (10 wall
    # layers), pyramids and tetrahedra (one interface layer) and tetrahedra:
    # 507,600 cells; lower-wall bump a_k ~ U(0, 0.01) (seed 4); inflow x-,
    # characteristic outflow x+, slip walls elsewhere; uniform M 0.6, p = rho = 1
    "c4": CaseSpec("c4: 120x30x30 bump channel (prism wall layers, pyramid/tet interface, tets), WENO-3, M0.6",
                   CHANNEL, 30, "weno3", seed=4, steps=50, bc=(1, 2, 4, 4, 4, 4),
                   bump=tuple(float(a) for a in np.random.default_rng(4).uniform(0.0, 0.01, 8)),
                   box=(120, 30, 30, 4.0, 1.0, 1.0), mach=0.6),
    # the unit cube, 121^3 hexes: a 61^3 box of hexahedra at the centre, a
    # one-hex shell of pyramids and tetrahedra around it, prism layers (12
    # hexes deep) at z = 0 and z = 1, Kuhn tetrahedra elsewhere; 5% vertex
    # jitter; 8,112,739 cells.  Blast at the centre, outflow (copy) on every
    # side; WENO-5 central polynomial with quadratic CWENO sectors
    "c5": CaseSpec("c5: 121^3 blast box (hex core, pyramid/tet shell, tets, prism layers), WENO-5 CWENO, mixed",
                   BLAST, 121, "weno5", jitter=0.05, seed=5, steps=300, bc=(2, 2, 2, 2, 2, 2),
                   box=(121, 121, 121, 1.0, 1.0, 1.0), ees=5, init="blast"),
}


def reduced(key: str, n: int) -> CaseSpec:
    s = PRESETS[key]
    box = ()
    if s.box:  # the same aspect (c4: 4n x n x n cells; c5: n^3)
        box = tuple(n * b // s.n for b in s.box[:3]) + tuple(s.box[3:])
    return dataclasses.replace(s, name=f"{key}-reduced-{n}", n=n, box=box)


class Case:
    def __init__(self, spec: CaseSpec, nthreads: int = 0, grid_dir: str | None = None, binary: bool = False):
        """the synthetic mesh of spec, or (grid_dir) the mesh of the GRID.cel /
        GRID.vrt / GRID.bnd files there with spec's scheme options"""
        self.spec = spec
        lib = native.prep_lib()
        nx, ny, nz, lx, ly, lz = spec.dims
        o = native.PrepOpts(shape=spec.shape, nx=nx, ny=ny, nz=nz,
                            lx=lx, ly=ly, lz=lz, jitter=spec.jitter, seed=spec.seed,
                            iorder=spec.iorder, iweno=spec.iweno, extf=spec.extf, weight_lsqr=0, nthreads=nthreads,
                            bc=(C.c_int32 * 6)(*spec.bc),
                            bump=(C.c_double * 8)(*(tuple(spec.bump) + (0.0,) * (8 - len(spec.bump)))),
                            ees=spec.ees)
        h = C.c_void_p()
        if grid_dir is None:
            rc = lib.ucnsp_build(C.byref(o), C.byref(h))
        else:
            rc = lib.ucnsp_read_grid(grid_dir.encode(), 1 if binary else 0, C.byref(o), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"ucnsp_{'build' if grid_dir is None else 'read_grid'} failed ({rc}): "
                               f"{lib.ucnsp_last_error().decode()}")
        self._h = h
        self._lib = lib
        ints = (C.c_int64 * 16)()
        dbls = (C.c_double * 8)()
        lib.ucnsp_info(h, ints, dbls)
        (self.ncell, self.nnode, self.nface_max, self.iorder, self.ideg, self.nnb, self.imax, self.typesten,
         self.iselem, self.igq, self.qp_quad, self.qp_tri, self.nqpmax, self.shape, self.iweno, _) = list(ints)
        self.period = (dbls[0], dbls[1], dbls[2])

    # -- raw arrays --------------------------------------------------------
    def array(self, name: str) -> np.ndarray:
        ptr = C.c_void_p()
        cnt = C.c_int64()
        dt = C.c_int32()
        rc = self._lib.ucnsp_array(self._h, name.encode(), C.byref(ptr), C.byref(cnt), C.byref(dt))
        if rc != 0:
            raise KeyError(self._lib.ucnsp_last_error().decode())
        if cnt.value == 0:
            return np.zeros(0, dtype=np.int32 if dt.value == 0 else np.float64)
        ctype = C.c_int32 if dt.value == 0 else C.c_double
        buf = (ctype * cnt.value).from_address(ptr.value)
        return np.ctypeslib.as_array(buf)

    def structs(self, precision: int = native.UCNS_FP64):
        m, s, p = native.MeshV2(), native.SchemeV2(), native.PrestoreV1()
        rc = self._lib.ucnsp_fill(self._h, precision, self.spec.gamma, self.spec.lwci1, C.byref(m), C.byref(s),
                                  C.byref(p))
        if rc != 0:
            raise RuntimeError(self._lib.ucnsp_last_error().decode())
        s.wenoz = self.spec.wenoz
        s.wenwrt = self.spec.wenwrt
        return m, s, p

    def cweno(self):
        """the ucns_cweno_v1 struct of a CWENO case (ees 5), else None"""
        if self.spec.ees != 5:
            return None
        cw = native.CwenoV1()
        rc = self._lib.ucnsp_fill_cweno(self._h, C.byref(cw))
        if rc != 0:
            raise RuntimeError(self._lib.ucnsp_last_error().decode())
        return cw

    def write_grid(self, grid_dir: str, binary: bool = False) -> None:
        """the mesh as GRID.cel / GRID.vrt / GRID.bnd in grid_dir (text, or
        Fortran unformatted records)"""
        rc = self._lib.ucnsp_write_grid(self._h, grid_dir.encode(), 1 if binary else 0)
        if rc != 0:
            raise RuntimeError(f"ucnsp_write_grid failed ({rc}): {self._lib.ucnsp_last_error().decode()}")

    def move_nodes(self, xyz: np.ndarray) -> None:
        """the same mesh with its nodes at xyz [nnode, 3]: geometry and
        PRESTORE data rebuilt, topology and stencils kept (ucnsp_move_nodes);
        the arrays and structs handed out before now hold the new data"""
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.nnode, 3)
        rc = self._lib.ucnsp_move_nodes(self._h, native.f64p(x))
        if rc != 0:
            raise RuntimeError(f"ucnsp_move_nodes failed ({rc}): {self._lib.ucnsp_last_error().decode()}")

    def init_state(self) -> np.ndarray:
        if self.spec.mach > 0:
            return np.tile(self.spec.freestream(), (self.ncell, 1))
        if self.spec.init == "blast":
            x = self.array("centre").reshape(-1, 3)
            c = 0.5 * np.array(self.spec.dims[3:])
            inside = np.linalg.norm(x - c, axis=1) < 0.2
            u = np.zeros((self.ncell, 5))
            u[:, 0] = np.where(inside, 8.0, 1.0)
            u[:, 4] = np.where(inside, 10.0, 1.0) / (self.spec.gamma - 1.0)
            return u
        u = np.empty((self.ncell, 5), dtype=np.float64)
        rc = self._lib.ucnsp_init_vortex(self._h, self.spec.gamma, self.spec.noise, self.spec.seed,
                                         native.f64p(u))
        if rc != 0:
            raise RuntimeError(self._lib.ucnsp_last_error().decode())
        return u

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ucnsp_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def face_rule(igq: int, tri: bool):
    lib = native.prep_lib()
    pts = np.zeros((64, 3))
    w = np.zeros(64)
    n = C.c_int32()
    rc = lib.ucnsp_face_rule(igq, 1 if tri else 0, native.f64p(pts), native.f64p(w), C.byref(n))
    if rc != 0:
        raise RuntimeError(lib.ucnsp_last_error().decode())
    return pts[: n.value].copy(), w[: n.value].copy()


def write_restart(path: str, it: int, t: float, u: np.ndarray) -> None:
    """RESTART.dat of an explicit march: step, time and the cell averages [ncell][5]"""
    lib = native.prep_lib()
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 5)
    rc = lib.ucnsp_write_restart(path.encode(), it, t, native.f64p(u), len(u))
    if rc != 0:
        raise RuntimeError(f"ucnsp_write_restart failed ({rc}): {lib.ucnsp_last_error().decode()}")


def read_restart(path: str, ncell: int):
    """(step, time, cell averages [ncell][5]) of a RESTART.dat"""
    lib = native.prep_lib()
    u = np.empty((ncell, 5))
    it, t = C.c_int32(), C.c_double()
    rc = lib.ucnsp_read_restart(path.encode(), ncell, C.byref(it), C.byref(t), native.f64p(u))
    if rc != 0:
        raise RuntimeError(f"ucnsp_read_restart failed ({rc}): {lib.ucnsp_last_error().decode()}")
    return it.value, t.value, u
