import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def init3000(tmp_path_factory):
    """The reference's sample snapshot (sample/INIT3000.dat), as (header, particles)."""
    from gplum_amd import ic
    path = tmp_path_factory.mktemp("golden") / "INIT3000.dat"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "INIT3000.dat.gz"), "rt") as fh:
        path.write_text(fh.read())
    return ic.read_snapshot(str(path))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle
