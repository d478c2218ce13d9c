"""The C-ABI library loads and exports every symbol include/mcint.h declares
(no GPU needed: nothing here calls into HIP)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "mcint.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:CACHE_SIZE_T|FINT|int|void|double|const char \*)\s*\*?\s*(\w+)\s*\(",
                       txt, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def libpath():
    import libcint_amd
    if not os.path.exists(libcint_amd.LIB_PATH):
        from libcint_amd import build
        build.build()
    return libcint_amd.LIB_PATH


def test_header_declares_reference_entry_points():
    names = _declared_functions()
    for n in ("int2e_sph", "int2e_cart", "int2e_ip1_sph", "int2e_ip1_cart", "int3c2e_sph",
              "int2e_optimizer", "int2e_ip1_optimizer", "int3c2e_optimizer", "CINTdel_optimizer",
              "CINTdel_2e_optimizer", "cint2e_sph", "cint2e_cart", "cint2e_sph_optimizer",
              "cint2e_ip1_sph", "CINTgto_norm", "CINTcgto_spheric", "CINTcgto_cart",
              "CINTtot_cgto_spheric", "CINTshells_spheric_offset", "CINTrys_roots", "CINTsr_rys_roots",
              "int2e_sph_batch", "int2e_ip1_sph_batch", "int3c2e_sph_batch", "int2e_ip2_sph",
              "int2e_ip2_optimizer", "int2e_ipip1_sph", "int2e_ipip1_cart", "int2e_ipip1_optimizer",
              "int2e_ip1ip2_sph", "int2e_ip1ip2_cart", "int2e_ip1ip2_optimizer", "int3c2e_ip1_sph",
              "int3c2e_ip2_sph", "int2c2e_sph"):
        assert n in names, n


def test_library_exports_every_declared_symbol(libpath):
    L = ctypes.CDLL(libpath)
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert not [n for n in _declared_functions() if n not in exported]


def test_host_helpers_match_libcint_conventions(libpath):
    import numpy as np
    import libcint_amd
    from libcint_amd.synthetic import gto_norm
    L = libcint_amd.lib()
    for l in range(5):
        for a in (0.1, 1.3, 27.0):
            assert abs(L.CINTgto_norm(l, a) - gto_norm(l, a)) <= 1e-14 * gto_norm(l, a)
    bas = np.array([[0, 2, 1, 2, 0, 20, 21, 0], [0, 1, 1, 1, 0, 20, 21, 0]], dtype=np.int32)
    Lb = L
    Lb.CINTtot_cgto_spheric.restype = ctypes.c_int
    assert Lb.CINTtot_cgto_spheric(bas.ctypes.data_as(ctypes.c_void_p), 2) == 10 + 3
    ao = np.zeros(2, dtype=np.int32)
    Lb.CINTshells_spheric_offset(ao.ctypes.data_as(ctypes.c_void_p), bas.ctypes.data_as(ctypes.c_void_p), 2)
    assert list(ao) == [0, 10]


def test_python_package_fails_loudly_without_library(monkeypatch, tmp_path):
    import libcint_amd
    monkeypatch.setattr(libcint_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(libcint_amd, "_lib", None)
    with pytest.raises(libcint_amd.McintError):
        libcint_amd.lib()


def test_wave_kernel_choice_is_validated_on_the_host(libpath):
    """mcint_wave_kernel_exists names the classes with a wave kernel, and
    mcint_tune_class_wave refuses the wave kernel for a class without one
    (host-side calls, no GPU)."""
    import libcint_amd as mc
    mc.lib()
    assert mc.wave_kernel_exists((3, 2, 2, 1))       # (fd|dp): 46 rows x 16 columns
    assert not mc.wave_kernel_exists((1, 0, 0, 0))   # single-chunk class
    assert not mc.wave_kernel_exists((3, 3, 3, 3))   # 2 x 74 accumulators per lane: no
    with pytest.raises(mc.McintError):
        mc.tune_class_wave((1, 0, 0, 0), 2)
    old = mc.tune_class_wave((3, 2, 2, 1), 2, team=8)
    assert mc.tune_class_wave((3, 2, 2, 1), old, team=8) == 2
    with pytest.raises(mc.McintError):
        mc.tune_class_wave((3, 2, 2, 1), 3)
    with pytest.raises(mc.McintError):
        mc.tune_class_wave((3, 2, 2, 1), 2, team=3)


def test_short_range_rule_arguments_are_checked_on_the_host(libpath):
    """mcint_sr_rys_roots_batch and CINTsr_rys_roots refuse bad arguments
    before any device call (nroots outside 1..9, x < 0 or not finite, lower
    outside [0, 1)); CINTsr_rys_roots then leaves u and w untouched.  An
    empty batch succeeds without touching the device."""
    import ctypes

    import numpy as np

    import libcint_amd as mc
    L = mc.lib()
    for n, x, lo in [(0, 1.0, 0.2), (10, 1.0, 0.2), (3, -1.0, 0.2), (3, float("inf"), 0.2),
                     (3, float("nan"), 0.2), (3, 1.0, 1.0), (3, 1.0, -0.1)]:
        with pytest.raises(mc.McintError):
            mc.sr_rys_roots(n, x, lo)
        u = np.full(max(n, 1), 7.0)
        w = np.full(max(n, 1), 7.0)
        dp = ctypes.POINTER(ctypes.c_double)
        L.CINTsr_rys_roots(n, x, lo, u.ctypes.data_as(dp), w.ctypes.data_as(dp))
        assert (u == 7.0).all() and (w == 7.0).all()
        assert b"mcint_sr_rys_roots_batch" in L.mcint_last_error()
    u, w = mc.sr_rys_roots(4, np.zeros(0), np.zeros(0))
    assert u.shape == (0, 4) and w.shape == (0, 4)
