"""The C++ functor interface (include/gplum_fdps.hpp) driven like FDPS drives the reference's
calcForceEPEPWithSearch / calcForceEPSP and the MultiWalk dispatch/retrieve pair."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gplum_amd", "lib",
                    "fdps_functor_demo")


def test_fdps_functor_program():
    assert os.path.exists(DEMO), "build first (python -m gplum_amd.build)"
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "fdps functors: OK" in r.stdout
