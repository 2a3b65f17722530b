"""The C-ABI boundary without a GPU: every symbol include/ucns.h declares is
exported by libucnsgpu.so, the ctypes table covers it, and argument
validation rejects bad input with the documented codes before any device
call is made.  The preprocessor and oracle libraries are checked the same
way."""
import ctypes as C
import os
import re

import pytest

from ucns3d_amd import native
from ucns3d_amd.case import Case, reduced

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "ucns.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ucns_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def header_macros():
    text = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"#define\s+(UCNS_[A-Z_0-9]+)\s+(-?\d+)", text)}


@pytest.fixture(scope="module")
def glib():
    if not os.path.exists(native.gpu_lib_path()):
        pytest.fail("libucnsgpu.so is not built: run __graft_entry__.build()")
    return native.gpu_lib()


@pytest.fixture(scope="module")
def small_case():
    return Case(reduced("c1", 4))


def test_header_lists_the_path():
    fns = header_functions()
    for fn in ["ucns_create", "ucns_upload_prestore", "ucns_set_state", "ucns_get_state", "ucns_residual",
               "ucns_reconstruct", "ucns_rk3_steps", "ucns_compute_dt", "ucns_destroy", "ucns_last_error",
               "ucns_residual_vjp", "ucns_residual_vjp_device", "ucns_rk3_adjoint", "ucns_rk3_adjoint_device"]:
        assert fn in fns
    assert len(fns) == 26


def test_every_declared_symbol_is_exported(glib):
    fns = header_functions()
    missing = [f for f in fns if not hasattr(glib, f)]
    assert not missing, missing
    assert set(fns) == set(native.GPU_API), set(fns) ^ set(native.GPU_API)


def test_constants_match_bindings(glib):
    m = header_macros()
    assert m["UCNS_ABI_VERSION"] == native.UCNS_ABI_VERSION == glib.ucns_abi_version()
    for k in ("UCNS_ORDER_CALLER", "UCNS_ORDER_RCM", "UCNS_ORDER_MORTON", "UCNS_KERNELS_AUTO", "UCNS_KERNELS_REGISTER"):
        assert m[k] == getattr(native, k), k
    assert m["UCNS_FP64"] == native.UCNS_FP64 and m["UCNS_MIXED"] == native.UCNS_MIXED
    assert m["UCNS_LAYOUT_AOS"] == native.UCNS_LAYOUT_AOS and m["UCNS_LAYOUT_SOA"] == native.UCNS_LAYOUT_SOA
    codes = [m[k] for k in ("UCNS_EBADARG", "UCNS_EUNSUPPORTED", "UCNS_ENOMEM", "UCNS_EHIP", "UCNS_ENONFINITE",
                            "UCNS_ESTATE")]
    assert len(set(codes)) == len(codes) and all(c < 0 for c in codes)


def test_struct_layouts_match_header():
    """ctypes mirrors: field count and order of each struct in the header."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for cname, py in [("ucns_mesh_v2", native.MeshV2), ("ucns_scheme_v2", native.SchemeV2),
                      ("ucns_prestore_v1", native.PrestoreV1), ("ucns_stats_v1", native.StatsV1),
                      ("ucns_boundary_v1", native.BoundaryV1)]:
        body = re.search(r"typedef struct %s \{(.*?)\}" % cname, text, flags=re.S).group(1)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
            fields += [n.strip().lstrip("*").strip().split("[")[0] for n in names.split(",")]
        assert fields == [f for f, _ in py._fields_], (cname, fields)


def _check(glib, ctx_out, code, needle):
    assert code is not None
    msg = glib.ucns_last_error(ctx_out).decode()
    glib.ucns_destroy(ctx_out)
    assert needle in msg, msg


def test_create_rejects_bad_arguments(glib, small_case):
    h = C.c_void_p()
    assert glib.ucns_create(None, None, C.byref(h)) == -1  # UCNS_EBADARG
    assert not h.value
    assert glib.ucns_create(None, None, None) == -1
    m, s, _ = small_case.structs()

    def attempt(mut_m=None, mut_s=None):
        m2, s2, _ = small_case.structs()
        if mut_m:
            mut_m(m2)
        if mut_s:
            mut_s(s2)
        h = C.c_void_p()
        rc = glib.ucns_create(C.byref(m2), C.byref(s2), C.byref(h))
        return rc, h

    cases = [
        (lambda m: setattr(m, "abi_version", 99), None, -1, "abi_version"),
        (lambda m: setattr(m, "ncell", 0), None, -1, "ncell"),
        (lambda m: setattr(m, "nface_max", 7), None, -1, "nface_max"),
        (lambda m: setattr(m, "surf", None), None, -1, "null mesh array"),
        (None, lambda s: setattr(s, "precision", 5), -1, "precision"),
        (None, lambda s: setattr(s, "iweno", 0), -2, "iweno"),
        (None, lambda s: setattr(s, "ees", 3), -2, "ees must be 0"),
        (None, lambda s: setattr(s, "wenoz", 1), -2, "wenoz 1 is a CWENO option"),
        (None, lambda s: setattr(s, "wenoz", 2), -1, "wenoz must be"),
        (None, lambda s: setattr(s, "wenwrt", 3), -2, "wenwrt must be 1"),
        # (the small case is MUSCL: characteristic reconstruction is a WENO option)
        (None, lambda s: setattr(s, "wenwrt", 2), -2, "characteristic reconstruction"),
        (None, lambda s: setattr(s, "riemann", 2), -2, "HLLC"),
        (None, lambda s: setattr(s, "limiter", 2), -2, "Barth"),
        (None, lambda s: setattr(s, "qp_quad", 0), -1, "quadrature"),
        (None, lambda s: setattr(s, "idegfree", 9), -2, "idegfree"),
        (None, lambda s: setattr(s, "cell_order", 7), -1, "cell_order"),
        (lambda m: setattr(m, "centroid", None), lambda s: setattr(s, "cell_order", native.UCNS_ORDER_MORTON), -1,
         "centroids"),
        (None, lambda s: setattr(s, "kernels", 3), -1, "kernels"),
        (None, lambda s: setattr(s, "reserved", 1), -1, "reserved"),
        (None, lambda s: setattr(s, "adj_keep_gib", float("nan")), -1, "adj_keep_gib"),
    ]
    for mm, ms, want, needle in cases:
        rc, h = attempt(mm, ms)
        assert rc == want, (needle, rc)
        assert h.value, "a context carrying the message is returned"
        _check(glib, h, rc, needle)


def test_null_context_is_rejected(glib):
    d = C.c_double()
    st = native.StatsV1()
    assert glib.ucns_upload_prestore(None, None) == -1
    assert glib.ucns_set_boundary(None, None) == -1
    assert glib.ucns_set_state(None, None, 0) == -1
    assert glib.ucns_get_state(None, None, 0) == -1
    assert glib.ucns_set_state_device(None, None, 0) == -1
    assert glib.ucns_get_state_device(None, None, 0) == -1
    assert glib.ucns_compute_dt(None, 0.5, C.byref(d)) == -1
    assert glib.ucns_residual(None, None, 0) == -1
    assert glib.ucns_residual_device(None, None, 0) == -1
    assert glib.ucns_debug_fail_alloc(None, 1) == -1
    assert glib.ucns_reconstruct(None, None) == -1
    assert glib.ucns_rk3_steps(None, 1, 0.5, 0.0, None) == -1
    assert glib.ucns_stats(None, C.byref(st)) == -1
    assert glib.ucns_reset_stats(None) == -1
    assert glib.ucns_last_error(None) == b"null context"
    glib.ucns_destroy(None)


def test_prep_library_exports_and_errors():
    lib = native.prep_lib()
    for fn in native.PREP_API:
        assert hasattr(lib, fn), fn
    o = native.PrepOpts(shape=9, nx=2, ny=2, nz=2, lx=1, ly=1, lz=1, jitter=0, seed=1, iorder=3, iweno=1, extf=2,
                        weight_lsqr=0, nthreads=1)
    h = C.c_void_p()
    assert lib.ucnsp_build(C.byref(o), C.byref(h)) != 0
    assert lib.ucnsp_last_error().decode()


def test_prep_rejects_a_bump_it_cannot_apply():
    """a lower-wall bump needs bounded y sides (periodic partners would not
    move together), and one that folds the cells next to the wall is refused
    with the cell named"""
    lib = native.prep_lib()

    def build(bc, a):
        o = native.PrepOpts(shape=1, nx=8, ny=4, nz=4, lx=4, ly=1, lz=1, jitter=0, seed=1, iorder=1, iweno=-1,
                            extf=3, weight_lsqr=0, nthreads=1, bc=(C.c_int32 * 6)(*bc),
                            bump=(C.c_double * 8)(*([a] + [0.0] * 7)))
        h = C.c_void_p()
        rc = lib.ucnsp_build(C.byref(o), C.byref(h))
        if rc == 0:
            lib.ucnsp_free(h)
        return rc, lib.ucnsp_last_error().decode()

    rc, msg = build((1, 2, 0, 0, 3, 3), 0.01)
    assert rc == -1 and "bounded y" in msg
    rc, msg = build((1, 2, 4, 4, 3, 3), 0.01)
    assert rc == 0, msg
    # slope ~0.9 along x over cells 0.5 long and 0.25 high: the wall nodes of
    # a cell differ in height by more than the cell, folding sub-tetrahedra
    rc, msg = build((1, 2, 4, 4, 3, 3), 0.6)
    assert rc == -1 and "inverts cell" in msg, msg


def test_oracle_library_exports():
    from oracle import oracle
    L = oracle.lib()
    for fn in ["oracle_residual", "oracle_reconstruct", "oracle_fluxes", "oracle_dt", "oracle_rk3_steps",
               "oracle_hllc", "oracle_rotatef", "oracle_rotateb", "oracle_weno_weights", "oracle_barth"]:
        assert hasattr(L, fn), fn
