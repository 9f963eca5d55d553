"""C callers of the boundary, built on the CPU.

* ``tests/c_abi/wf_abi_check.c`` includes ``include/wf_po.h`` and links
  ``libwfpo.so`` the way a reference driver would; its ``errors`` mode checks
  argument validation and error reporting without device work (built with
  AddressSanitizer and UndefinedBehaviorSanitizer on the caller's side).
* ``tests/c_abi/oracle_sanitize.c`` drives every function of the C oracle in
  a build with AddressSanitizer and UndefinedBehaviorSanitizer.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "wavefiber3d_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")


def _build(tmp_path, name, srcs, extra):
    exe = os.path.join(tmp_path, name)
    cmd = ["gcc", "-std=c11", "-O1", "-g", *srcs, "-o", exe, *extra, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_san",
                 [os.path.join(ROOT, "tests", "c_abi", "oracle_sanitize.c"),
                  os.path.join(ROOT, "oracle", "po_oracle.c")],
                 [*SAN, "-ffp-contract=off", "-fopenmp", "-I" + os.path.join(ROOT, "oracle")])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


def test_c_caller_error_paths(tmp_path):
    lib = os.path.join(LIBDIR, "libwfpo.so")
    if not os.path.exists(lib):
        pytest.skip("libwfpo.so not built")
    exe = _build(tmp_path, "abi_check", [os.path.join(ROOT, "tests", "c_abi", "wf_abi_check.c")],
                 [*SAN, "-I" + os.path.join(ROOT, "include"), "-L" + LIBDIR, "-lwfpo",
                  "-Wl,-rpath," + LIBDIR])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")  # the HIP runtime's own allocations
    r = subprocess.run([exe, "errors"], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors: ok" in r.stdout


def test_analytic_normal_derivative_of_currents(tmp_path):
    """The reverse pass's dL/dn uses the analytic derivative of the Fresnel
    currents (wavefiber3d_amd/csrc/wf_currents.h, current_point_dn); on the
    host it matches central differences of the currents to 1e-6 over random
    lit points, normals and incidences (tests/c_abi/currents_dn_check.cpp)."""
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not found")
    exe = os.path.join(tmp_path, "cdn")
    r = subprocess.run([hipcc, "-std=c++17", "-O2",
                        "-I" + os.path.join(ROOT, "wavefiber3d_amd", "csrc"),
                        os.path.join(ROOT, "tests", "c_abi", "currents_dn_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
