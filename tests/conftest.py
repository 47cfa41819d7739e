import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_NAMES = ["water_1500", "water_6k", "chain_3k", "chain_lmd", "ions_3k", "water_54", "chain_bonded",
                "chain_lmd_bonded", "chain_tors", "chain_lmd_tors"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_golden(name):
    """(System rebuilt from the fixture's recipe, fixture arrays).  The
    generator must reproduce the exact inputs md_run saw (fingerprint)."""
    from make_golden import build_system, fingerprint
    f = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    rec = json.loads(str(f["recipe"]))
    s = build_system(rec["kind"], rec["args"])
    assert fingerprint(s) == str(f["fingerprint"]), f"generator no longer reproduces {name}; regenerate fixtures"
    return s, f


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_golden(name)
        return cache[name]
    return get


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
