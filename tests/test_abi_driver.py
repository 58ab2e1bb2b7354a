"""The C ABI from C (tests/abi_driver.c): it compiles against
include/prnu_api.h and links to libprnu_hip.so with no other header or
binding (CPU test), and on the GPU its per-image / per-pair drop-in calls
agree with the batched Gram and the FP64 definition (GPU test)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "rocket_amd")


def build_driver(out_dir):
    exe = os.path.join(out_dir, "abi_driver")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "abi_driver.c"), "-L", PKG, "-lprnu_hip", f"-Wl,-rpath,{PKG}",
           "-Wl,-rpath-link,/opt/rocm/lib", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_driver_builds_against_the_header(tmp_path):
    assert os.path.exists(build_driver(str(tmp_path)))


@pytest.mark.gpu
def test_driver_runs_on_the_gpu(tmp_path):
    exe = build_driver(str(tmp_path))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "abi driver:" in r.stdout
