import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


@pytest.fixture(scope="session")
def ref_runs():
    with np.load(os.path.join(GOLDEN, "ref_runs.npz")) as z:
        data = {k: z[k] for k in z.files}
    meta = json.load(open(os.path.join(GOLDEN, "ref_runs.json")))
    return data, meta


@pytest.fixture(scope="session")
def examples():
    return json.load(open(os.path.join(GOLDEN, "examples.json")))


def build_deck(meta_entry):
    from piafs_amd import decks
    return getattr(decks, meta_entry["builder"])(**meta_entry["kwargs"])
