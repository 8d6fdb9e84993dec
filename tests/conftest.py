import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def mesh_from(d):
    from mixsolver_amd.meshgen import Mesh
    return Mesh(x=d["x"], fptr=d["fptr"], fnodes=d["fnodes"], c0=d["c0"], c1=d["c1"],
                zone=d["zone"], ncells=int(d["ncells"]), ctype=int(d["ctype"]))


@pytest.fixture
def golden():
    return load_golden
