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


def fixture_quartets(name):
    """Shell tuples of libcint's test-suite loops for one integral."""
    if name.startswith("int3c2e"):
        return np.array([(i, j, k, 0) for k in range(4) for j in range(4) for i in range(4)],
                        np.int32)
    return libcint_quartets()


def class_fingerprints(bas, q, out, off, name):
    """Per class (l_i, l_j, l_k[, l_l]): signed sum_k x_k cos(k) of each block,
    and sum |x| (the tolerance scale)."""
    three = name.startswith("int3c2e")
    fps, mag = {}, {}
    for n, s in enumerate(q):
        blk = out[off[n]:off[n + 1]]
        key = ",".join(str(int(bas[s[a], 1])) for a in range(3 if three else 4))
        fps[key] = fps.get(key, 0.0) + float(blk.dot(np.cos(np.arange(blk.size))))
        mag[key] = mag.get(key, 0.0) + float(np.abs(blk).sum()) + 1e-2 * blk.size
    return fps, mag


def check_class_fingerprints(golden_fps, fps, mag, what):
    """Each class's signed fingerprint within 1e-10 of its scale."""
    assert sorted(golden_fps) == sorted(fps), what
    bad = [(k, fps[k], golden_fps[k]) for k in fps
           if abs(fps[k] - golden_fps[k]) > 1e-10 * mag[k]]
    assert not bad, (what, bad[:5])
