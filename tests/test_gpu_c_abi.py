"""A C program that includes include/wf_po.h and links libwfpo.so runs one
tree solve on the GPU and checks it against the FP64 direct sum (relative L2
of intensity <= 1e-3), through the C-ABI only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "wavefiber3d_amd")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")]


def test_c_caller_solve(tmp_path):
    exe = os.path.join(tmp_path, "abi_check")
    cmd = ["gcc", "-std=c11", "-O2", os.path.join(ROOT, "tests", "c_abi", "wf_abi_check.c"),
           "-o", exe, "-I" + os.path.join(ROOT, "include"), "-L" + LIBDIR, "-lwfpo",
           "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("case: C caller, 3000-point sphere, depth 2", flush=True)
    r = subprocess.run([exe, "solve"], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "solve: ok" in r.stdout
