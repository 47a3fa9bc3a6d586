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
              "grid16x12_strain_loge", "grid6_bsr_jacobi", "grid6_bsr_ic0", "grid6_bsr_block_rcb2",
              "grid6_bsr_block_chunk4", "grid16_rcm_block64", "grid16_block_chunk8"]

PRECOND_OF = {"jacobi": 0, "block": 1, "ic0": 2}


def case_precond(meta: dict) -> int:
    """Oracle preconditioner code; block cases with more than one node per subdomain
    use the general subdomain form (code 3) on the case's dof -> subdomain map."""
    if meta["preconditioner"] == "block" and meta.get("subdomain_nodes", 1) != 1:
        return 3
    return PRECOND_OF[meta["preconditioner"]]


def case_storage(meta: dict) -> int:
    return 1 if meta.get("storage") == "bsr" else 0
