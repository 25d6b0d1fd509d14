import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "e1039-gpu_amd")
ORACLE = os.path.join(ROOT, "oracle")
sys.path.insert(0, PKG)
sys.path.insert(0, ORACLE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP chain)")


def _ensure_built():
    if not os.path.exists(os.path.join(PKG, "libe1039or.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(PKG, "csrc")], check=True)
    if not os.path.exists(os.path.join(ORACLE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def oc():
    import orchain
    return orchain


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def geo(oc):
    return oc.Geometry.load()


@pytest.fixture(scope="session")
def ogeo(oc, orc):
    return orc.OracleGeometry(oc.DEFAULT_GEOMETRY)


@pytest.fixture(scope="session")
def session(oc, geo):
    return oc.Session(geo, oc.default_params(max_hits_per_event=2048), device=0)
