"""The Fortran binding (ucns3d_amd/fortran/ucns_gpu.f90) compiled with
amdflang into a host program (tests/fortran/host_march.f90) that packs its
own Fortran arrays with 1-based indices and drives the library as UCNS3D's
TIME_MARCHING would: per step ucns_compute_dt (CALCULATE_CFL), then one
ucns_rk3_steps step with that dt (RUNGE_KUTTA3), T += DT.

CPU: the module's ABI version matches the library's, and ucns_create's
error paths return their codes and messages through the binding.
GPU: the Fortran host's march equals the ctypes path's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from ucns3d_amd import native

HOST = os.path.join(native.LIBDIR, "ucns_fortran_host")


def host():
    if not os.path.exists(HOST):
        pytest.fail("ucns_fortran_host is not built: run __graft_entry__.build()")
    return HOST


def test_fortran_binding_error_paths():
    r = subprocess.run([host(), "errors"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FORTRAN_ERRORS_OK" in r.stdout
    assert "library ABI version %d" % native.UCNS_ABI_VERSION in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("shape,n,steps", [("tet", 4, 5), ("hex", 8, 5)])
def test_fortran_host_march_matches_the_ctypes_path(shape, n, steps, tmp_path):
    """the same case (reduced C2: WENO-4 tets; reduced C1: MUSCL hexes), the
    same steps: the Fortran host, 1-based indices and Fortran-owned arrays,
    against the Solver on the preprocessor's 0-based arrays"""
    from ucns3d_amd.case import Case, reduced
    from ucns3d_amd.native import UCNS_FP64
    from ucns3d_amd.solver import Solver
    out = tmp_path / "u.bin"
    print(f"[fortran host] {shape} n={n} steps={steps}", flush=True)
    r = subprocess.run([host(), "march", shape, str(n), str(steps), str(out)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FORTRAN_MARCH_OK" in r.stdout
    raw = out.read_bytes()
    nc = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    t_f = float(np.frombuffer(raw[4:12], dtype=np.float64)[0])
    u_f = np.frombuffer(raw[12:], dtype=np.float64).reshape(nc, 5)
    case = Case(reduced("c2" if shape == "tet" else "c1", n))
    assert case.ncell == nc
    s = Solver(case, UCNS_FP64)
    s.set_state(case.init_state())
    t = 0.0
    for _ in range(steps):
        dt = s.compute_dt(0.5)
        t = s.rk3_steps(1, dt_fixed=dt, t0=t)
    u = s.get_state()
    s.close()
    assert t == t_f
    assert np.array_equal(u, u_f)
