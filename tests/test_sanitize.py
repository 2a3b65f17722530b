"""The host preprocessor and the oracle under AddressSanitizer (with leak
checks) and UndefinedBehaviorSanitizer: tests/sanitize/sanitized_march.cpp
drives every mesh construction, scheme and boundary kind at a small size
through PRESTORE, the reconstruction, the residual and its tangent, RK3
steps, a node move and the mesh and restart files; any finding aborts the
run with a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = ["tests/sanitize/sanitized_march.cpp", "ucns3d_amd/csrc/prep/ucns_prep.cpp", "oracle/ucns_oracle.cpp"]


@pytest.mark.timeout(900)
def test_preprocessor_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "march")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fopenmp", "-ffp-contract=off", "-Iinclude", "-Iucns3d_amd/csrc/prep",
           *SOURCES, "-o", exe, "-lm"]
    b = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(tmp_path / "files")], cwd=str(tmp_path), capture_output=True, text=True, env=env,
                       timeout=800)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert r.stdout.count(" ok") == 8, r.stdout
