"""The committed MI355X tuning table (libcint_amd/tuning/mi355x.json) is
well formed and every entry is accepted by the library's host-side tuning
calls (no GPU): team sizes among the tuner's candidates, classes within the
register kernels' range, wave-kernel choices only for classes and team sizes
that have a wave kernel."""
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "libcint_amd", "tuning", "mi355x.json")


@pytest.fixture(scope="module")
def mc():
    import libcint_amd
    if not os.path.exists(libcint_amd.LIB_PATH):
        from libcint_amd import build
        build.build()
    libcint_amd.lib()
    return libcint_amd


def _cls(key):
    c = [int(v) for v in key.split(",")]
    assert len(c) == 4 and all(0 <= v <= 4 for v in c), key
    return c


def test_tuning_table_is_well_formed(mc):
    from libcint_amd import autotune
    with open(TABLE) as fh:
        data = json.load(fh)
    cands = data["candidates"]
    tables = data["tables"]
    assert {"int2e_sph", "int2e_ip1_sph", "int3c2e_sph", "wave"} <= set(tables)
    for name, table in tables.items():
        if name == "wave":
            continue
        assert name in autotune.INTOR_CODES, name
        for key, ppl in table.items():
            _cls(key)
            assert ppl in cands, (name, key, ppl)
    for key, mode in tables["wave"].items():
        cls, _, team = key.partition(":")
        c = _cls(cls)
        t = int(team)
        assert mode in (1, 2) and 1 <= t <= 64 and t & (t - 1) == 0, key
        if mode == 2:
            assert mc.wave_kernel_exists(c), key
    # the library accepts every entry (host-side calls); applying twice is
    # what loading the library already did, so the state is unchanged
    autotune.apply(autotune.load(TABLE))
