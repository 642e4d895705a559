import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_londonhp100():
    """test/data/londonhp100coords.csv + londonhp100data.csv (copied as fixtures)."""
    c = np.loadtxt(os.path.join(GOLDEN, "londonhp100coords.csv"), delimiter=",", skiprows=1)
    d = np.loadtxt(os.path.join(GOLDEN, "londonhp100data.csv"), delimiter=",", skiprows=1)
    return c, d


@pytest.fixture(scope="session")
def londonhp100():
    c, d = load_londonhp100()
    y = d[:, 0]
    X = np.column_stack([np.ones(len(y)), d[:, 1:4]])          # [1, FLOORSZ, UNEMPLOY, PROF]
    Xm = np.column_stack([np.ones(len(y)), d[:, 1], d[:, 3]])  # [1, FLOORSZ, PROF]
    return dict(coords=c, y=y, X=X, Xm=Xm)


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from libgwmodel_amd import capi
    ctx = capi.Context(0)
    yield ctx
    ctx.close()
