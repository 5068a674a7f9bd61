import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU case")


@pytest.fixture(scope="session")
def gmp_fixtures():
    path = os.path.join(GOLDEN, "gmp_residues.json")
    with open(path) as f:
        data = json.load(f)
    out = {}
    for p, rows in data.items():
        for r in rows:
            out[(int(p), r["iterations"])] = r
    return out


def announce(msg):
    """Name the case before a GPU step, so a fault is attributable."""
    print("[case] " + msg, flush=True)
