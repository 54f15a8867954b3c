"""Test configuration: markers, import path, and in-tree builds when missing."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU renders, run with -m slow")


def _ensure_built():
    from vviewer_amd import build as B
    if not os.path.exists(B.ORACLE_LIB):
        B.build_oracle()
    if not os.path.exists(B.LIB):
        B.build_product()


@pytest.fixture(scope="session", autouse=True)
def built_libraries():
    _ensure_built()


@pytest.fixture(scope="session")
def has_reference():
    return os.path.isdir(os.path.join(REFERENCE, "assets", "unittests"))


@pytest.fixture(scope="session")
def proc_meshes():
    """Procedural stand-ins for the reference's uvsphere/plane/cube assets (identical on every host)."""
    from vviewer_amd import meshes as M
    return {"uvsphere": M.uv_sphere(), "plane": M.plane(), "cube": M.cube()}
