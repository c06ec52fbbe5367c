"""The C++ functor interface (include/gplum_fdps.hpp) driven like FDPS drives the reference's
calcForceEPEPWithSearch / calcForceEPSP and the MultiWalk dispatch/retrieve pair."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "_build", "fdps_functor_demo")


def build_demo():
    """Compiles tests/fdps_functor_demo.cpp (a C++ program that drives include/gplum_fdps.hpp
    as FDPS would) against the built library; called by __graft_entry__.build()."""
    root = os.path.dirname(HERE)
    lib_dir = os.path.join(root, "gplum_amd", "lib")
    os.makedirs(os.path.dirname(DEMO), exist_ok=True)
    cmd = ["g++", "-std=c++14", "-O2", "-I", os.path.join(root, "include"),
           os.path.join(HERE, "fdps_functor_demo.cpp"), "-L", lib_dir, "-lgplum_gpu",
           "-Wl,-rpath," + os.path.relpath(lib_dir, os.path.dirname(DEMO)).join(["$ORIGIN/", ""]), "-o", DEMO]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for the functor demo:\n%s" % r.stderr[-4000:])


def test_fdps_functor_program():
    assert os.path.exists(DEMO), "build first (__graft_entry__.build())"
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "fdps functors: OK" in r.stdout
