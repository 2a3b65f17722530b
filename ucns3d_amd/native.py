"""ctypes bindings of the C-ABI in include/ucns.h and of the host preprocessor.

The structs below mirror ucns_mesh_v2 / ucns_scheme_v2 / ucns_prestore_v1
field for field.  Loading fails loudly when a library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

UCNS_ABI_VERSION = 2
UCNS_OK = 0
UCNS_FP64 = 0
UCNS_MIXED = 1
UCNS_LAYOUT_AOS = 0
UCNS_LAYOUT_SOA = 1
UCNS_ORDER_CALLER = 0
UCNS_ORDER_RCM = 1
UCNS_ORDER_MORTON = 2
UCNS_KERNELS_AUTO = 0
UCNS_KERNELS_REGISTER = 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class MeshV2(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("ncell", C.c_int32), ("index_base", C.c_int32), ("nface_max", C.c_int32),
        ("ifca", _i32p), ("ineigh", _i32p), ("ineighn", _i32p), ("bcode", _i32p), ("types_faces", _i32p),
        ("faceanglex", _f64p), ("faceangley", _f64p), ("surf", _f64p), ("totvolume", _f64p), ("minedge", _f64p),
        ("centroid", _f64p),
    ]


class SchemeV2(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("precision", C.c_int32), ("iweno", C.c_int32), ("iorder", C.c_int32),
        ("idegfree", C.c_int32), ("inumneighbours", C.c_int32), ("typesten", C.c_int32), ("ees", C.c_int32),
        ("wenoz", C.c_int32), ("wenwrt", C.c_int32), ("limiter", C.c_int32), ("riemann", C.c_int32),
        ("qp_quad", C.c_int32), ("qp_triangle", C.c_int32),
        ("gamma", C.c_double), ("lwci1", C.c_double), ("weno_eps", C.c_double),
        ("weights_quad", _f64p), ("weights_tri", _f64p),
        ("cell_order", C.c_int32), ("kernels", C.c_int32), ("adj_keep_stages", C.c_int32), ("reserved", C.c_int32),
        ("adj_keep_gib", C.c_double),
    ]


class PrestoreV1(C.Structure):
    _fields_ = [
        ("admis", _i32p), ("full", _i32p), ("ihexl", _i32p), ("invmat", _f64p), ("indicator", _f64p),
        ("integ_basis", _f64p), ("volume11", _f64p), ("qpoints", _f64p),
    ]


class BoundaryV1(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("boundtype", C.c_int32),
        ("inflow", C.c_double * 5), ("outflow", C.c_double * 5),
    ]


class CwenoV1(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("idegfree2", C.c_int32), ("inumneighbours2", C.c_int32),
        ("reserved", C.c_int32), ("invmat_c", _f64p), ("indicator_c", _f64p), ("integ_basis_c", _f64p),
    ]


class GeometryBarV1(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("reserved", C.c_int32),
        ("invmat", _f64p), ("indicator", _f64p), ("integ_basis", _f64p), ("volume11", _f64p),
        ("qpoints", _f64p), ("surf", _f64p), ("normal", _f64p), ("totvolume", _f64p),
    ]


class StatsV1(C.Structure):
    _fields_ = [
        ("n_recon", C.c_int64), ("n_flux", C.c_int64), ("n_update", C.c_int64), ("n_dt", C.c_int64),
        ("ms_recon", C.c_double), ("ms_flux", C.c_double), ("ms_update", C.c_double), ("ms_dt", C.c_double),
        ("bytes_recon_per_cell", C.c_double),
        ("n_reduced_last", C.c_int64), ("n_firstorder_last", C.c_int64),
        ("n_recon_vjp", C.c_int64), ("n_flux_vjp", C.c_int64),
        ("ms_recon_vjp", C.c_double), ("ms_flux_vjp", C.c_double),
        ("n_recon_adj", C.c_int64), ("ms_recon_adj", C.c_double),
    ]


class PrepOpts(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double), ("jitter", C.c_double),
        ("seed", C.c_uint64), ("iorder", C.c_int32), ("iweno", C.c_int32), ("extf", C.c_int32),
        ("weight_lsqr", C.c_int32), ("nthreads", C.c_int32), ("bc", C.c_int32 * 6), ("bump", C.c_double * 8),
        ("ees", C.c_int32),
    ]


# every entry point declared in include/ucns.h: name -> (restype, argtypes)
GPU_API = {
    "ucns_create": (C.c_int, [C.POINTER(MeshV2), C.POINTER(SchemeV2), C.POINTER(C.c_void_p)]),
    "ucns_upload_prestore": (C.c_int, [C.c_void_p, C.POINTER(PrestoreV1)]),
    "ucns_upload_cweno": (C.c_int, [C.c_void_p, C.POINTER(CwenoV1)]),
    "ucns_set_boundary": (C.c_int, [C.c_void_p, C.POINTER(BoundaryV1)]),
    "ucns_set_boundary_faces": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "ucns_set_state": (C.c_int, [C.c_void_p, _f64p, C.c_int32]),
    "ucns_get_state": (C.c_int, [C.c_void_p, _f64p, C.c_int32]),
    "ucns_set_state_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "ucns_get_state_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "ucns_compute_dt": (C.c_int, [C.c_void_p, C.c_double, _f64p]),
    "ucns_residual": (C.c_int, [C.c_void_p, _f64p, C.c_int32]),
    "ucns_residual_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "ucns_residual_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "ucns_reconstruct": (C.c_int, [C.c_void_p, _f64p]),
    "ucns_rk3_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, _f64p]),
    "ucns_residual_vjp": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                                    C.POINTER(C.c_double)]),
    "ucns_residual_vjp_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ucns_rk3_adjoint": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "ucns_rk3_adjoint_device": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.POINTER(C.c_double)]),
    "ucns_rk3_adjoint_geometry": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_double),
                                            C.POINTER(GeometryBarV1), C.POINTER(C.c_double)]),
    "ucns_stats": (C.c_int, [C.c_void_p, C.POINTER(StatsV1)]),
    "ucns_reset_stats": (C.c_int, [C.c_void_p]),
    "ucns_debug_fail_alloc": (C.c_int, [C.c_void_p, C.c_int32]),
    "ucns_last_error": (C.c_char_p, [C.c_void_p]),
    "ucns_destroy": (None, [C.c_void_p]),
    "ucns_abi_version": (C.c_int, []),
}

PREP_API = {
    "ucnsp_build": (C.c_int, [C.POINTER(PrepOpts), C.POINTER(C.c_void_p)]),
    "ucnsp_write_grid": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "ucnsp_read_grid": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(PrepOpts), C.POINTER(C.c_void_p)]),
    "ucnsp_write_restart": (C.c_int, [C.c_char_p, C.c_int32, C.c_double, C.POINTER(C.c_double), C.c_int64]),
    "ucnsp_read_restart": (C.c_int, [C.c_char_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "ucnsp_fill": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double,
                             C.POINTER(MeshV2), C.POINTER(SchemeV2), C.POINTER(PrestoreV1)]),
    "ucnsp_array": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                              C.POINTER(C.c_int32)]),
    "ucnsp_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "ucnsp_init_vortex": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_uint64, _f64p]),
    "ucnsp_face_rule": (C.c_int, [C.c_int32, C.c_int32, _f64p, _f64p, C.POINTER(C.c_int32)]),
    "ucnsp_shape_table": (C.c_int, [C.c_int32, C.POINTER(C.c_int32)]),
    "ucnsp_fill_cweno": (C.c_int, [C.c_void_p, C.POINTER(CwenoV1)]),
    "ucnsp_move_nodes": (C.c_int, [C.c_void_p, _f64p]),
    "ucnsp_free": (None, [C.c_void_p]),
    "ucnsp_last_error": (C.c_char_p, []),
}

_libs: dict[str, C.CDLL] = {}


def _load(name: str, api: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = name if os.path.isabs(name) else os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python -m ucns3d_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    for fn, (res, args) in api.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def prep_lib() -> C.CDLL:
    return _load("libucnsprep.so", PREP_API)


def gpu_lib(path: str | None = None) -> C.CDLL:
    """The HIP library; `path` selects another build of it (an absolute path).

    torch (when installed) is imported first: its wheel carries its own HIP
    runtime under the same soname, and a process that loads the system
    runtime first cannot initialise torch's devices afterwards.  With torch
    loaded, the library binds to the runtime already in the process."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return _load(path or "libucnsgpu.so", GPU_API)


def gpu_lib_path() -> str:
    return os.path.join(LIBDIR, "libucnsgpu.so")


def f64p(arr) -> C.POINTER(C.c_double):
    return arr.ctypes.data_as(_f64p)


def i32p(arr) -> C.POINTER(C.c_int32):
    return arr.ctypes.data_as(_i32p)
