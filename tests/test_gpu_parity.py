"""GPU parity: libmcint (HIP, through the C ABI) against the CPU oracle.

Bar: per element |gpu - ref| <= 1e-12 + 1e-10 |ref| (FP64 path), the
checksums of libcint's own test suite, and device Rys roots within 1e-13
relative of the exact quad-precision rule.
"""
import os
import sys
import zlib

import numpy as np
import pytest

from conftest import (ROOT, check_class_fingerprints, class_fingerprints, close_place,
                      fixture_quartets, libcint_quartets)

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-12, 1e-10


def _say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mc():
    import libcint_amd
    libcint_amd.lib()
    return libcint_amd


def _cmp(got, ref, what):
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    err = np.abs(got - ref)
    bad = err > ATOL + RTOL * np.abs(ref)
    if bad.any():
        i = int(np.argmax(err - (ATOL + RTOL * np.abs(ref))))
        raise AssertionError(f"{what}: {int(bad.sum())} elements off, worst at {i}: "
                             f"got {got.ravel()[i]!r} ref {ref.ravel()[i]!r}")


def test_device_rys_roots(mc, orc):
    rng = np.random.default_rng(2)
    xs = np.concatenate([[0.0, 1e-9], rng.uniform(0, 2, 60), rng.uniform(0, 110, 200),
                         np.arange(0.0, 100.0, 1.0), [150.0, 1e4]])
    for n in range(1, 10):
        _say("roots n =", n)
        u, w = mc.rys_roots(n, xs)
        for i, x in enumerate(xs):
            u0, w0 = orc.rys_roots(n, float(x), exact=True)
            np.testing.assert_allclose(u[i], u0, rtol=1e-13, atol=0, err_msg=f"u n={n} x={x}")
            np.testing.assert_allclose(w[i], w0, rtol=1e-13, atol=0, err_msg=f"w n={n} x={x}")


def test_device_short_range_rule(mc, orc):
    """CINTsr_rys_roots on the device (weight exp(-x t^2) on [lower, 1]) against
    the oracle's rule (itself pinned by its moments against mpmath), against
    the device full-range rule at lower = 0, and by its own moments
    sum_i w_i s_i^k = int_lower^1 t^2k exp(-x t^2) dt, k < 2n (closed form in
    mpmath).
    Cases with x lower^2 > 600 are left out: their weights are below
    1e-260 and lose relative precision in double."""
    import mpmath as mp

    xs = np.array([0.0, 1e-6, 0.3, 2.0, 17.0, 60.0, 250.0, 1000.0])
    lows = np.array([0.0, 0.05, 0.3, 0.7, 0.95])
    X, L = np.meshgrid(xs, lows, indexing="ij")
    keep = X * L * L <= 600
    X, L = X[keep], L[keep]
    worst_u = worst_w = 0.0
    for n in range(1, 10):
        u, w = mc.sr_rys_roots(n, X, L)
        for i in range(X.size):
            u0, w0 = orc.sr_rys_roots(n, float(X[i]), float(L[i]))
            # roots compared in s = t^2 = u / (1 + u), weights relative to the largest
            s, s0 = u[i] / (1 + u[i]), u0 / (1 + u0)
            eu = np.max(np.abs(s - s0))
            ew = np.max(np.abs(w[i] - w0)) / np.max(np.abs(w0))
            worst_u, worst_w = max(worst_u, eu), max(worst_w, ew)
            assert eu < 1e-13 and ew < 1e-11, (n, X[i], L[i], eu, ew)
        uf, wf = mc.rys_roots(n, xs)
        u0, w0 = mc.sr_rys_roots(n, xs, 0.0)
        np.testing.assert_allclose(u0, uf, rtol=1e-11, atol=0, err_msg=f"lower=0 u n={n}")
        np.testing.assert_allclose(w0, wf, rtol=1e-11, atol=0, err_msg=f"lower=0 w n={n}")
    _say(f"short-range rule vs oracle: max |ds| {worst_u:.2e}, max rel |dw| {worst_w:.2e}")
    # moments in closed form: I_0 = sqrt(pi)/(2 sqrt x) (erfc(sqrt x lower) -
    # erfc(sqrt x)), I_k = (lower^(2k-1) e^(-x lower^2) - e^(-x) + (2k-1) I_(k-1)) / 2x
    mp.mp.dps = 60
    for n, x, lo in [(9, 2.0, 0.3), (6, 60.0, 0.05), (9, 250.0, 0.7), (4, 0.3, 0.95),
                     (9, 250.0, 0.0)]:
        u, w = mc.sr_rys_roots(n, x, lo)
        s = u[0] / (1 + u[0])
        xm, lo_m = mp.mpf(x), mp.mpf(lo)
        mom = mp.sqrt(mp.pi) / (2 * mp.sqrt(xm)) * (mp.erfc(mp.sqrt(xm) * lo_m) - mp.erfc(mp.sqrt(xm)))
        for k in range(2 * n):
            if k:
                mom = (lo_m ** (2 * k - 1) * mp.exp(-xm * lo_m ** 2) - mp.exp(-xm)
                       + (2 * k - 1) * mom) / (2 * xm)
            got = float(np.sum(w[0] * s ** k))
            assert abs(got - float(mom)) <= 1e-12 * float(mom), (n, x, lo, k, got, float(mom))


@pytest.mark.parametrize("name,key", [("int2e_sph", "cint2e_sph"), ("int2e_cart", None),
                                      ("int2e_ip1_sph", "cint2e_ip1_sph"),
                                      ("int2e_ip1_cart", None), ("int2e_ip2_sph", None),
                                      ("int2e_ip2_cart", None), ("int2e_ipip1_sph", None),
                                      ("int2e_ipip1_cart", None), ("int2e_ip1ip2_sph", None),
                                      ("int2e_ip1ip2_cart", None)])
def test_libcint_fixture(mc, orc, golden, fixture_basis, name, key):
    """All (i<=j, k<=l) quartets of testsuite/test_cint.py's fixture: l <= 4,
    a 2-primitive 2-contraction d shell, up to 9 Rys roots (the second
    derivatives skip (gg|gg), whose 10 roots the device tables do not hold)."""
    _say("fixture", name)
    q = libcint_quartets()
    if "ipip1" in name or "ip1ip2" in name:
        ls = fixture_basis[1][:, 1]
        q = q[(ls[q].sum(axis=1) + 2) // 2 + 1 <= 9]
    B = mc.Basis(*fixture_basis)
    got, off, not0 = mc.batch(name, q, B)
    ref, off_r, not0_r = orc.batch(name, q, orc.Basis(*fixture_basis))
    np.testing.assert_array_equal(off, off_r)
    _cmp(got, ref, name)
    np.testing.assert_array_equal(not0, not0_r)
    if key:
        v = float(np.abs(got).sum())
        gref, place = golden["golden"][key]
        assert close_place(v, gref, got.size, place), (v, gref)


@pytest.mark.parametrize("name", ["int2e_sph", "int2e_cart", "int2e_ip1_sph", "int3c2e_sph"])
def test_signed_class_fingerprints(mc, fixture_basis, name):
    """Device blocks over libcint's fixture against the committed signed
    per-class fingerprints sum_k x_k cos(k) (computed by the oracle, whose
    spherical transform is pinned to libcint's table): a sign flip in any
    spherical row, Cartesian component or derivative direction of any class
    changes its class's fingerprint."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "signed_fingerprints.json")) as fh:
        gold = json.load(fh)[name]
    _say("signed fingerprints", name)
    q = fixture_quartets(name)
    out, off, _ = mc.batch(name, q, mc.Basis(*fixture_basis))
    fps, mag = class_fingerprints(fixture_basis[1], q, out, off, name)
    check_class_fingerprints(gold, fps, mag, name)


def test_libcint_fixture_3c2e(mc, orc, golden, fixture_basis):
    _say("fixture int3c2e_sph")
    q = np.array([(i, j, k, 0) for k in range(4) for j in range(4) for i in range(4)], np.int32)
    got, _, _ = mc.batch("int3c2e_sph", q, mc.Basis(*fixture_basis))
    ref, _, _ = orc.batch("int3c2e_sph", q, orc.Basis(*fixture_basis))
    _cmp(got, ref, "int3c2e_sph")
    v = float(np.abs(got).sum())
    gref, place = golden["golden"]["cint3c2e_sph"]
    assert close_place(v, gref, got.size, place), (v, gref)


def test_libcint_fixture_3c2e_cart(mc, orc, fixture_basis):
    _say("fixture int3c2e_cart")
    q = np.array([(i, j, k, 0) for k in range(8) for j in range(8) for i in range(8)], np.int32)
    got, _, nz = mc.batch("int3c2e_cart", q, mc.Basis(*fixture_basis))
    ref, _, nz_r = orc.batch("int3c2e_cart", q, orc.Basis(*fixture_basis))
    _cmp(got, ref, "int3c2e_cart")
    np.testing.assert_array_equal(nz, nz_r)


@pytest.mark.parametrize("name,ref", [("int3c2e_ip1_sph", 2242.052249221302),
                                      ("int3c2e_ip2_sph", 1970.982483824248),
                                      ("int2c2e_sph", 782.3104849606677)])
def test_fixture_3c2e_derivatives_and_2c2e(mc, orc, golden, fixture_basis, name, ref):
    """int3c2e_ip1 / int3c2e_ip2 / int2c2e over testsuite/test_3c2e.py's loops
    (i, j, k over the four regular shells; (i, k) for int2c2e): element by
    element against the oracle, non-zero flags, and the reference's
    checksums (test_3c2e.py:303-319, 10 places); the Cartesian twins over all
    eight shells against the oracle."""
    _say("fixture", name)
    B = mc.Basis(*fixture_basis)
    if name.startswith("int2c2e"):
        q = np.array([(i, k, 0, 0) for k in range(4) for i in range(4)], np.int32)
        qc = np.array([(i, k, 0, 0) for k in range(8) for i in range(8)], np.int32)
    else:
        q = np.array([(i, j, k, 0) for k in range(4) for j in range(4) for i in range(4)], np.int32)
        qc = np.array([(i, j, k, 0) for k in range(8) for j in range(8) for i in range(8)], np.int32)
    got, off, nz = mc.batch(name, q, B)
    ref_o, off_r, nz_r = orc.batch(name, q, orc.Basis(*fixture_basis))
    np.testing.assert_array_equal(off, off_r)
    _cmp(got, ref_o, name)
    np.testing.assert_array_equal(nz, nz_r)
    v = float(np.abs(got).sum())
    assert close_place(v, ref, got.size, 10), (name, v, ref)
    cart = name.replace("_sph", "_cart")
    got, _, nz = mc.batch(cart, qc, B)
    ref_o, _, nz_r = orc.batch(cart, qc, orc.Basis(*fixture_basis))
    _cmp(got, ref_o, cart)
    np.testing.assert_array_equal(nz, nz_r)


@pytest.mark.parametrize("omega", [0.0, 0.3, -0.3])
def test_c4_3c2e_derivative_and_2c2e_samples(mc, orc, omega):
    """The 40-atom cluster's orbital and auxiliary shells: int3c2e_ip1 /
    int3c2e_ip2 on (i, j | P) triples and int2c2e on (P|Q) pairs of the
    auxiliary basis, plain, long-range erf and short-range erfc."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    env = m.env.copy()
    env[8] = omega
    rng = np.random.default_rng(44)
    n = 300
    q = np.zeros((n, 4), np.int32)
    q[:, 0] = rng.integers(0, m.aux_start, n)
    q[:, 1] = rng.integers(0, m.aux_start, n)
    q[:, 2] = rng.integers(m.aux_start, m.nbas, n)
    q2 = np.zeros((n, 4), np.int32)
    q2[:, 0] = rng.integers(m.aux_start, m.nbas, n)
    q2[:, 1] = rng.integers(m.aux_start, m.nbas, n)
    _say("c4 3c derivative / 2c samples, omega", omega)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        for name, qq in (("int3c2e_ip1_sph", q), ("int3c2e_ip2_sph", q), ("int2c2e_sph", q2),
                         ("int3c2e_ip2_cart", q[:100]), ("int2c2e_cart", q2[:100])):
            got, _, nz = mc.batch(name, qq, mc.Basis(m.atm, m.bas, env))
            ref, _, nz_r = orc.batch(name, qq, orc.Basis(m.atm, m.bas, env))
            _cmp(got, ref, f"c4 {name} omega {omega}")
            np.testing.assert_array_equal(nz, nz_r)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)


def test_batch_with_caller_offsets(mc, fixture_basis):
    """int2e_sph_batch with out_offsets (blocks placed with gaps, reverse
    order) gives byte-identical blocks to the packed call; the gaps stay
    untouched."""
    import ctypes
    _say("batch with caller offsets")
    L = mc.lib()
    q = libcint_quartets()[:200]
    B = mc.Basis(*fixture_basis)
    packed, off, _ = mc.batch("int2e_sph", q, B)
    sizes = np.diff(off)
    # reverse order with a 3-double gap before each block
    pos = np.zeros(q.shape[0], dtype=np.uint64)
    run = 0
    for n in reversed(range(q.shape[0])):
        run += 3
        pos[n] = run
        run += int(sizes[n])
    out = np.full(run + 3, 7.25)
    not0 = np.zeros(q.shape[0], dtype=np.int32)
    shls = np.ascontiguousarray(q, dtype=np.int32)
    rc = L.int2e_sph_batch(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           pos.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)),
                           shls.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), q.shape[0],
                           B.atm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), B.natm,
                           B.bas.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), B.nbas,
                           B.env.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None,
                           not0.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None)
    assert rc == 0, mc.last_error()
    touched = np.zeros(out.size, bool)
    for n in range(q.shape[0]):
        a = int(pos[n])
        np.testing.assert_array_equal(out[a:a + sizes[n]], packed[off[n]:off[n + 1]])
        touched[a:a + sizes[n]] = True
    assert np.all(out[~touched] == 7.25)


def _sample_quartets(nbas, n, seed):
    rng = np.random.default_rng(seed)
    q = rng.integers(0, nbas, size=(n, 4)).astype(np.int32)
    return q


@pytest.mark.parametrize("cfg,n", [("c0", 3000), ("c1", 1200), ("c3", 800)])
def test_synthetic_samples(mc, orc, cfg, n):
    from libcint_amd.synthetic import make_molecule
    m = make_molecule(cfg)
    nb = m.nbas if m.aux_start < 0 else m.aux_start
    q = _sample_quartets(nb, n, seed=zlib.crc32(cfg.encode()) % 1000)
    _say(cfg, "sample of", n, "quartets")
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        for name in ("int2e_sph", "int2e_cart", "int2e_ip1_sph", "int2e_ip1_cart",
                     "int2e_ip2_sph", "int2e_ip2_cart", "int2e_ipip1_sph", "int2e_ip1ip2_sph",
                     "int2e_ip1ip2_cart"):
            qq = {"int2e_sph": q, "int2e_cart": q[: n // 2], "int2e_ip1_sph": q[: n // 4],
                  "int2e_ip1_cart": q[: n // 8], "int2e_ip2_sph": q[n // 4: n // 2],
                  "int2e_ip2_cart": q[n // 2: n // 2 + n // 8],
                  "int2e_ipip1_sph": q[n // 2: n // 2 + n // 8],
                  "int2e_ip1ip2_sph": q[3 * n // 4: 3 * n // 4 + n // 8],
                  "int2e_ip1ip2_cart": q[7 * n // 8:]}[name]
            got, _, nz = mc.batch(name, qq, mc.Basis(m.atm, m.bas, m.env))
            ref, _, nz_r = orc.batch(name, qq, orc.Basis(m.atm, m.bas, m.env))
            _cmp(got, ref, f"{cfg} {name}")
            np.testing.assert_array_equal(nz, nz_r)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)


def test_c4_3c2e_sample(mc, orc):
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    rng = np.random.default_rng(4)
    n = 600
    q = np.zeros((n, 4), np.int32)
    q[:, 0] = rng.integers(0, m.aux_start, n)
    q[:, 1] = rng.integers(0, m.aux_start, n)
    q[:, 2] = rng.integers(m.aux_start, m.nbas, n)
    _say("c4 int3c2e sample")
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        for name in ("int3c2e_sph", "int3c2e_cart"):
            got, _, nz = mc.batch(name, q, mc.Basis(m.atm, m.bas, m.env))
            ref, _, nz_r = orc.batch(name, q, orc.Basis(m.atm, m.bas, m.env))
            _cmp(got, ref, f"c4 {name}")
            np.testing.assert_array_equal(nz, nz_r)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)


def test_per_call_with_dims(mc, orc, fixture_basis):
    """int2e_sph(out, dims, ...) writes into a larger block (src/cint2e.c:853-856)."""
    import ctypes
    _say("per-call dims")
    L = mc.lib()
    atm, bas, env = fixture_basis
    B = mc.Basis(atm, bas, env)
    shls = np.array([1, 2, 0, 5], np.int32)
    d = [B.shell_dim(int(s)) for s in shls]
    dims = np.array([d[0] + 2, d[1] + 1, d[2] + 3, d[3]], np.int32)
    out = np.full(int(np.prod(dims)), 7.0)
    rc = L.int2e_sph(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                     dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     shls.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     B.atm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), B.natm,
                     B.bas.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), B.nbas,
                     B.env.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None)
    ref, rc_r = orc.intor("int2e_sph", tuple(shls), orc.Basis(atm, bas, env))
    assert rc == rc_r == 1
    o = out.reshape(dims, order="F")
    _cmp(o[: d[0], : d[1], : d[2], : d[3]], ref, "dims block")
    assert np.all(o[d[0]:] == 7.0) and np.all(o[:, d[1]:] == 7.0)


def test_per_call_threads_share_optimizer(mc, fixture_basis):
    """libcint's documented usage (examples/time_c2h6.c:719-745): one
    int2e_optimizer shared by 8 host threads, each calling int2e_sph /
    int2e_ip1_sph per shell tuple.  Every block equals the batch entry
    point's; the optimizer holds the device state (basis, all pair data), so
    a call with it only launches and copies.  Per-call latency with and
    without the optimizer is printed."""
    import ctypes
    import time
    from concurrent.futures import ThreadPoolExecutor
    L = mc.lib()
    B = mc.Basis(*fixture_basis)
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    args = (B.atm.ctypes.data_as(ip), B.natm, B.bas.ctypes.data_as(ip), B.nbas,
            B.env.ctypes.data_as(dp))
    for name in ("int2e_sph", "int2e_ip1_sph", "int2e_ip2_sph", "int2e_ip1ip2_sph"):
        q = libcint_quartets()
        if name == "int2e_ip1ip2_sph":
            ls = fixture_basis[1][:, 1]
            q = q[(ls[q].sum(axis=1) + 2) // 2 + 1 <= 9]
        ref, off, nz_ref = mc.batch(name, q, B)
        opt = ctypes.c_void_p()
        getattr(L, name.replace("_sph", "") + "_optimizer")(ctypes.byref(opt), *args)
        assert L.mcint_optimizer_on_device(opt) == 1
        got = np.full(ref.size, np.nan)
        st = np.full(q.shape[0], -1, dtype=np.int32)
        f = getattr(L, name)
        shl = np.ascontiguousarray(q, dtype=np.int32)

        def work(rows):
            for n in rows:
                dst = got[off[n]:off[n + 1]]
                st[n] = f(dst.ctypes.data_as(dp), None, shl[n].ctypes.data_as(ip), args[0], args[1],
                          args[2], args[3], args[4], opt, None)

        _say(name, "per call from 8 threads sharing one optimizer,", q.shape[0], "tuples")
        with ThreadPoolExecutor(8) as ex:
            list(ex.map(work, [range(t, q.shape[0], 8) for t in range(8)]))
        np.testing.assert_array_equal(st, nz_ref)
        _cmp(got, ref, f"{name} per call with a shared optimizer")
        # latency, one thread: with the optimizer, and without (upload + plan)
        rows = range(0, q.shape[0], max(1, q.shape[0] // 200))
        t0 = time.perf_counter()
        work(rows)
        t_opt = (time.perf_counter() - t0) / len(rows)
        t0 = time.perf_counter()
        for n in rows:
            dst = got[off[n]:off[n + 1]]
            f(dst.ctypes.data_as(dp), None, shl[n].ctypes.data_as(ip), args[0], args[1], args[2],
              args[3], args[4], None, None)
        t_none = (time.perf_counter() - t0) / len(rows)
        _say(f"{name} per-call latency: {t_opt * 1e6:.0f} us with the optimizer, "
             f"{t_none * 1e6:.0f} us without")
        _cmp(got, ref, f"{name} per call without an optimizer")
        L.CINTdel_optimizer(ctypes.byref(opt))
        assert not opt.value


def test_screened_quartet_is_zero_filled(mc, orc):
    """Far-apart tight primitives: every primitive pair screened -> zeros, not0 = 0."""
    env = [0.0] * 20
    atm, bas = [], []
    for r in ((0, 0, 0), (40.0, 0, 0)):
        atm.append([1, len(env), 1, 0, 0, 0])
        env.extend(r)
    for ia in (0, 1):
        bas.append([ia, 1, 1, 1, 0, len(env), len(env) + 1, 0])
        env.extend([3.0, 1.0])
    B = (np.array(atm, np.int32), np.array(bas, np.int32), np.array(env))
    q = np.array([[0, 1, 0, 1], [0, 0, 1, 1]], np.int32)
    got, _, nz = mc.batch("int2e_sph", q, mc.Basis(*B))
    ref, _, nz_r = orc.batch("int2e_sph", q, orc.Basis(*B))
    np.testing.assert_array_equal(nz, nz_r)
    assert nz[0] == 0 and nz[1] == 1
    _cmp(got, ref, "screened")
    assert np.all(got[:81] == 0)


def test_plan_graph_replay_matches_direct(mc):
    """Plan outputs via HIP-graph replays (with eager work and syncs between)
    equal the directly launched plan, on every replay, with the buckets in
    one graph branch and forked over several concurrent branches."""
    import torch
    from libcint_amd.synthetic import make_molecule, unique_quartets
    m = make_molecule("c0")
    q = unique_quartets(m.nbas)[::97].copy()
    _say("plan graph replay", len(q), "quartets")
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    plan = mc.Plan(db, "int2e_sph", q)
    dev = torch.device("cuda:0")
    ref = torch.zeros(plan.out_size, dtype=torch.float64, device=dev)
    plan.execute(ref.data_ptr(), graph=False)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    out = torch.empty(plan.out_size, dtype=torch.float64, device=dev)
    # one branch, then the buckets forked over 4 and 3 concurrent branches
    # (each capture replayed three times)
    for nbr in (1, 4, 3):
        old = mc.tune(mc.TUNE_GRAPH_BRANCHES, nbr)
        try:
            for rep in range(3):
                out.fill_(float("nan"))
                torch.cuda.synchronize()
                plan.execute(out.data_ptr(), stream_ptr=s.cuda_stream, graph=True)
                s.synchronize()
                junk = torch.randn(1000, device=dev).sum()  # eager work between replays
                torch.cuda.synchronize()
                assert torch.equal(out, ref), f"{nbr} branches, replay {rep}"
        finally:
            mc.tune(mc.TUNE_GRAPH_BRANCHES, old)
    del junk


def test_schwarz_bounds(mc, orc):
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c0")
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    _say("schwarz")
    qb = db.schwarz()
    B = orc.Basis(m.atm, m.bas, m.env)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        for i, j in ((0, 0), (5, 3), (47, 12), (30, 30), (47, 46)):
            v, _ = orc.intor("int2e_sph", (i, j, i, j), B)
            ref = np.sqrt(np.abs(v).max())
            assert abs(qb[i, j] - ref) <= 1e-12 * ref
            assert qb[j, i] == qb[i, j]
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)


def _with_omega(basis_arrays, omega):
    atm, bas, env = basis_arrays
    env = np.array(env, dtype=np.float64, copy=True)
    env[8] = omega  # PTR_RANGE_OMEGA
    return atm, bas, env


@pytest.mark.parametrize("omega", [0.5, -0.5, -0.2])
def test_range_separated_fixture(mc, orc, fixture_basis, omega):
    """Long-range (omega > 0) and short-range (omega < 0) int2e_sph on the
    libcint fixture quartets against the oracle, which uses libcint's own
    short-range rule on [sqrt(theta), 1] for rys_order > 3 (the device takes
    full minus long range on doubled roots for every order)."""
    _say("fixture int2e_sph omega", omega)
    q = libcint_quartets()
    arr = _with_omega(fixture_basis, omega)
    got, off, not0 = mc.batch("int2e_sph", q, mc.Basis(*arr))
    ref, off_r, not0_r = orc.batch("int2e_sph", q, orc.Basis(*arr))
    _cmp(got, ref, f"int2e_sph omega={omega}")
    np.testing.assert_array_equal(not0, not0_r)


def test_erf_plus_erfc_is_coulomb(mc, fixture_basis):
    """testsuite/test_cint.py:258-302 on the device: LR + SR = full."""
    q = libcint_quartets()
    full, _, _ = mc.batch("int2e_sph", q, mc.Basis(*fixture_basis))
    for omega in (0.2, 0.5, 0.8):
        _say("erf + erfc omega", omega)
        lr, _, _ = mc.batch("int2e_sph", q, mc.Basis(*_with_omega(fixture_basis, omega)))
        sr, _, _ = mc.batch("int2e_sph", q, mc.Basis(*_with_omega(fixture_basis, -omega)))
        assert np.round(np.abs(lr + sr - full).max(), 9) == 0


@pytest.mark.parametrize("omega", [-0.3, 0.3])
def test_c4_3c2e_range_separated_sample(mc, orc, omega):
    """The 40-atom C/H cluster's (ij|P) with the erfc (omega = -0.3) and erf
    (omega = 0.3) operators against the oracle; aux shells up to g, so the
    short-range rys_order reaches 6.  The erfc case runs on the three-centre
    register kernels (single-kernel and chunked classes), erf on the generic
    kernel."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    rng = np.random.default_rng(7)
    n = 1500
    q = np.zeros((n, 4), np.int32)
    q[:, 0] = rng.integers(0, m.aux_start, n)
    q[:, 1] = rng.integers(0, m.aux_start, n)
    q[:, 2] = rng.integers(m.aux_start, m.nbas, n)
    arr = _with_omega((m.atm, m.bas, m.env), omega)
    _say("c4 int3c2e omega", omega)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        got, _, nz = mc.batch("int3c2e_sph", q, mc.Basis(*arr))
        ref, _, nz_r = orc.batch("int3c2e_sph", q, orc.Basis(*arr))
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)
    _cmp(got, ref, f"c4 int3c2e_sph omega={omega}")
    np.testing.assert_array_equal(nz, nz_r)


def test_c4_3c2e_fp32_tier(mc):
    """FP32 tier of the three-centre path on a c4 sample: every quartet
    flagged; the single-kernel classes run coefficients, recurrence and
    contraction in FP32 (roots, weights and the result stay FP64).  Stated
    error: max |FP32 - FP64| <= 1e-5 max|FP64| over the sample."""
    import torch
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    rng = np.random.default_rng(11)
    n = 4000
    q = np.zeros((n, 4), np.int32)
    q[:, 0] = rng.integers(0, m.aux_start, n)
    q[:, 1] = rng.integers(0, m.aux_start, n)
    q[:, 2] = rng.integers(m.aux_start, m.nbas, n)
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    p64 = mc.Plan(db, "int3c2e_sph", q)
    p32 = mc.Plan(db, "int3c2e_sph", q, fp32_flags=np.ones(n, np.uint8))
    o64 = torch.zeros(p64.out_size, dtype=torch.float64, device="cuda")
    o32 = torch.zeros(p32.out_size, dtype=torch.float64, device="cuda")
    p64.execute(o64.data_ptr(), graph=False)
    p32.execute(o32.data_ptr(), graph=False)
    torch.cuda.synchronize()
    st = p32.stats()
    err = float((o32 - o64).abs().max())
    scale = float(o64.abs().max())
    _say("c4 3c2e FP32 tier:", st["fp32_tasks"], "of", st["tasks"], "tasks FP32; max|d|", err,
         "max|ref|", scale)
    assert st["fp32_tasks"] > 0
    assert err <= 1e-5 * scale
    p64.close()
    p32.close()
    db.close()


def _ip1_translation_residual(batch, q, bas):
    """Per quartet, max |d/dA + d/dB + d/dC + d/dD| of int2e_ip1_sph, the four
    derivatives taken from the ip1 blocks of (ij|kl), (ji|kl), (kl|ij),
    (lk|ij); zero up to rounding by translational invariance."""
    perm = np.concatenate([q[:, [0, 1, 2, 3]], q[:, [1, 0, 2, 3]], q[:, [2, 3, 0, 1]],
                           q[:, [3, 2, 0, 1]]])
    out, off, _ = batch("int2e_ip1_sph", perm)
    d = 2 * bas[:, 1] + 1
    n = q.shape[0]
    res = np.zeros(n)
    for t in range(n):
        i, j, k, l = (int(d[s]) for s in q[t])
        blk = [out[off[t + m * n]:off[t + m * n + 1]] for m in range(4)]
        tot = (blk[0].reshape(3, l, k, j, i)
               + blk[1].reshape(3, l, k, i, j).transpose(0, 1, 2, 4, 3)
               + blk[2].reshape(3, j, i, l, k).transpose(0, 3, 4, 1, 2)
               + blk[3].reshape(3, j, i, k, l).transpose(0, 4, 3, 1, 2))
        res[t] = np.abs(tot).max()
    return res


# Share of the quartets of one ip1 comparison that may use the allowance
# of _cmp_ip1 (at least two quartets).
IP1_ALLOWANCE_FRAC = 0.02


def _cmp_ip1(mc, orc, m, q, got, ref, off, what):
    """ip1 parity: every element within the bar, except in quartets where the
    reference's own result is off by more than that (its translational
    invariance residual exceeds the bar); there the GPU result must satisfy
    invariance much better than the reference and differ from it by no more
    than the reference's residual.  At most IP1_ALLOWANCE_FRAC of the
    quartets (and at least two) may take the allowance."""
    err = np.abs(got - ref)
    bad = err > ATOL + RTOL * np.abs(ref)
    if not bad.any():
        _say(what, "ip1: every element within the bar")
        return
    qi = np.unique(np.searchsorted(off, np.nonzero(bad)[0], side="right") - 1)
    _say(what, "ip1 allowance taken by", len(qi), "of", len(q), "quartets")
    assert len(qi) <= max(2, IP1_ALLOWANCE_FRAC * len(q)), (what, len(qi), len(q))
    qb = q[qi]
    r_ref = _ip1_translation_residual(
        lambda n, qq: orc.batch(n, qq, orc.Basis(m.atm, m.bas, m.env)), qb, m.bas)
    r_gpu = _ip1_translation_residual(
        lambda n, qq: mc.batch(n, qq, mc.Basis(m.atm, m.bas, m.env)), qb, m.bas)
    for t, n in enumerate(qi):
        e = err[off[n]:off[n + 1]].max()
        _say(what, "quartet", qb[t], "off by", e, "reference residual", r_ref[t],
             "gpu residual", r_gpu[t])
        assert r_ref[t] > ATOL and e <= 2 * r_ref[t] + ATOL and r_gpu[t] <= 0.1 * r_ref[t], \
            (what, qb[t], e, r_ref[t], r_gpu[t])


def test_ip1_reference_rounding_on_tight_primitives(mc, orc):
    """(p f|d f) with p primitives up to exponent 1968 and |A-B| = 4 bohr:
    libcint's transfer of the 2a-weighted raised integrals, done per
    primitive quartet, loses about 1e-10 relative on the derivative block of
    the p shell (its translational-invariance residual is 1.5e-11), while the
    GPU path transfers after contraction and keeps the residual below 1e-12.
    The two differ by the reference's own rounding error."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c1")
    q = np.array([[44, 79, 48, 79]], np.int32)
    orc.set_root_mode(orc.ROOTS_EXACT)
    B = orc.Basis(m.atm, m.bas, m.env)
    r_ref = _ip1_translation_residual(lambda n, qq: orc.batch(n, qq, B), q, m.bas)[0]
    r_gpu = _ip1_translation_residual(
        lambda n, qq: mc.batch(n, qq, mc.Basis(m.atm, m.bas, m.env)), q, m.bas)[0]
    got, off, _ = mc.batch("int2e_ip1_sph", q, mc.Basis(m.atm, m.bas, m.env))
    ref, _, _ = orc.batch("int2e_ip1_sph", q, B)
    _say("tight-primitive ip1: reference residual", r_ref, "gpu residual", r_gpu,
         "difference", np.abs(got - ref).max())
    assert r_ref > 1e-12 and r_gpu < 1e-12
    assert np.abs(got - ref).max() <= 2 * r_ref


@pytest.mark.parametrize("ppl", [1, 4096])
def test_team_shapes_and_screening_walk(mc, orc, ppl):
    """The same contracted quartets computed by teams of 64 lanes (1 primitive
    quartet per lane: the screening staircase in LDS) and by single lanes
    (4096 per lane: rows beyond the LDS row budget, every primitive quartet
    tested) agree with the oracle, values and non-zero flags."""
    from libcint_amd import autotune
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c1")
    q = _sample_quartets(m.nbas, 1500, seed=77)
    _say("c1 sample, primitive quartets per lane", ppl)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        for name in ("int2e_sph", "int2e_ip1_sph"):
            autotune.reset(name)
        mc.tune(mc.TUNE_PRIMS_PER_LANE, ppl)
        for name in ("int2e_sph", "int2e_ip1_sph"):
            qq = q if name == "int2e_sph" else q[:400]
            got, off, nz = mc.batch(name, qq, mc.Basis(m.atm, m.bas, m.env))
            ref, _, nz_r = orc.batch(name, qq, orc.Basis(m.atm, m.bas, m.env))
            if name == "int2e_sph":
                _cmp(got, ref, f"ppl {ppl} {name}")
            else:
                _cmp_ip1(mc, orc, m, qq, got, ref, off, f"ppl {ppl} {name}")
            np.testing.assert_array_equal(nz, nz_r)
    finally:
        mc.tune(mc.TUNE_PRIMS_PER_LANE, 32)
        autotune.apply(autotune.load())
        orc.set_root_mode(orc.ROOTS_EXACT)


def test_edge_cases(mc, fixture_basis):
    """Empty quartet lists and out-of-range shell indices through the C ABI:
    an empty result, and a clear error before anything is written or
    launched, for every integral kind; the library stays usable."""
    import ctypes
    L = mc.lib()
    B = mc.Basis(*fixture_basis)
    P = ctypes.POINTER
    _say("edge cases: empty lists, bad shell indices")

    def call(name, q):
        q = np.ascontiguousarray(q, dtype=np.int32).reshape(-1, 4)
        out = np.full(4096, 7.25)
        not0 = np.full(max(q.shape[0], 1), -3, np.int32)
        rc = getattr(L, name + "_batch")(
            out.ctypes.data_as(P(ctypes.c_double)), None, q.ctypes.data_as(P(ctypes.c_int)),
            q.shape[0], B.atm.ctypes.data_as(P(ctypes.c_int)), B.natm,
            B.bas.ctypes.data_as(P(ctypes.c_int)), B.nbas, B.env.ctypes.data_as(P(ctypes.c_double)),
            None, not0.ctypes.data_as(P(ctypes.c_int)), None)
        return rc, out, not0

    for name in ("int2e_sph", "int2e_cart", "int2e_ip1_sph", "int3c2e_sph"):
        rc, out, _ = call(name, np.zeros((0, 4)))
        assert rc == 0, mc.last_error()
        assert np.all(out == 7.25)
        for bad in ([0, 0, B.nbas, 0], [0, -1, 0, 0]):
            rc, out, not0 = call(name, np.array([[0, 0, 0, 0], bad]))
            assert rc != 0 and "out of range" in mc.last_error(), (name, bad, rc, mc.last_error())
            assert np.all(out == 7.25) and np.all(not0 == -3)
    got, _, _ = mc.batch("int2e_sph", np.array([[1, 0, 1, 0]], np.int32), B)
    assert np.isfinite(got).all() and np.abs(got).max() > 0


@pytest.mark.parametrize("name,omega", [("int2e_sph", 0.3), ("int2e_ip1_sph", 0.3),
                                        ("int3c2e_sph", 0.3), ("int2e_sph", -0.3)])
def test_range_separated_sample_on_register_kernels(mc, orc, name, omega):
    """erf(0.3 r)/r (omega > 0) and erfc(0.3 r)/r (omega < 0) on a c1 sample
    (contracted s/p shells, screening) for the plain, derivative and
    three-centre register kernels against the oracle."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c1")
    env = m.env.copy()
    env[8] = omega  # PTR_RANGE_OMEGA
    q = _sample_quartets(m.nbas, 900 if name == "int2e_sph" else 300, seed=31)
    _say("range-separated sample", name, omega)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        got, off, nz = mc.batch(name, q, mc.Basis(m.atm, m.bas, env))
        ref, _, nz_r = orc.batch(name, q, orc.Basis(m.atm, m.bas, env))
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)
    if name == "int2e_ip1_sph":
        _cmp_ip1(mc, orc, type("M", (), {"atm": m.atm, "bas": m.bas, "env": env})(), q, got, ref,
                 off, f"omega {omega} {name}")
    else:
        _cmp(got, ref, f"omega {omega} {name}")
    np.testing.assert_array_equal(nz, nz_r)
