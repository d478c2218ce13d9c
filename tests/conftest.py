import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmcint.so")


def _ensure_oracle():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def orc():
    return _ensure_oracle()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "libcint_testsuite.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def fixture_basis(golden):
    """atm/bas/env of libcint's testsuite/test_cint.py (8 shells + exponent-0 s shell)."""
    return (np.array(golden["atm"], dtype=np.int32), np.array(golden["bas"], dtype=np.int32),
            np.array(golden["env"], dtype=np.float64))


def libcint_quartets(nb=8):
    """(i<=j, k<=l) quartets in the loop order of test_cint.py:235-256."""
    out = []
    for l in range(nb):
        for k in range(l + 1):
            for j in range(nb):
                for i in range(j + 1):
                    out.append((i, j, k, l))
    return np.array(out, dtype=np.int32)


def close_place(v, ref, count, place):
    """libcint's checksum criterion (testsuite/test_cint.py:148-149)."""
    return round(abs(v - ref) / count ** 0.5, place) == 0
