import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden_index():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as handle:
        return json.load(handle)


def load_golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN_DIR, f"{name}.npz")) as data:
        return {key: data[key] for key in data.files}


FULL_CASES = ["one_element", "scrambled_none", "scrambled_rcm", "extended_loads", "grid6_jacobi", "grid6_ic0",
              "grid6_block1", "grid16_stress", "grid16_rcm_coloring", "grid16_rcm_ic0", "grid16_rcm_block1",
              "grid16x12_strain_loge"]

PRECOND_OF = {"jacobi": 0, "block": 1, "ic0": 2}


def case_precond(meta: dict) -> int:
    return PRECOND_OF[meta["preconditioner"]]
