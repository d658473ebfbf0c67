"""A C++ consumer of include/exadg_amd.hpp (the binding an ExaDG build would compile), built
with g++ against libexadg_amd.so. The CPU part runs here; the GPU part solves a manufactured
problem on the device."""
import os
import subprocess

import pytest

import exadg_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "consumer.cpp")


def _build(tmp_path):
    exe = str(tmp_path / "consumer")
    libdir = os.path.dirname(exadg_amd.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include", SRC, "-o", exe,
           "-L" + libdir, "-lexadg_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + libdir + ":/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_consumer_cpu(tmp_path):
    assert os.path.exists(exadg_amd.LIB_PATH), "build with python -m exadg_amd.build"
    r = subprocess.run([_build(tmp_path), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "cpu ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_consumer_gpu_solve(tmp_path):
    r = subprocess.run([_build(tmp_path), "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "gpu ok" in r.stdout, r.stdout + r.stderr
