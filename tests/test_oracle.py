"""Pins the CPU oracle (oracle/cint_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake (a dropped term, a wrong
sign, index or normalisation, a transposed operand) fails at least one:
  * libcint's own golden values: checksums of int2e_sph / int2e_ip1_sph /
    int3c2e_sph over the test-suite fixture, Boys constants, root
    fingerprints (tests/golden/libcint_testsuite.json);
  * mathematics: Gauss-rule moment identities against mpmath's incomplete
    gamma, orthonormality of the solid harmonics, the (ss|ss) closed form,
    permutational symmetry, finite differences for the i-derivative, and
    the (ij|k) = (ij|k s0) identity of the 3-centre integral;
  * libcint's Cartesian -> spherical table (src/cart2sph.c:21-806) read as
    data (tests/golden/libcint_cart2sph.json), entry by entry and with
    signs, for the oracle and for the device's generated table.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import (ROOT, check_class_fingerprints, class_fingerprints, close_place,
                      fixture_quartets, libcint_quartets)


def fp(a):
    a = np.asarray(a)
    return a.ravel().dot(np.cos(np.arange(a.size)))


def mp_boys(k, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * k + 1)
    return mp.gammainc(k + mp.mpf(1) / 2, 0, x) / (2 * x ** (k + mp.mpf(1) / 2))


def test_boys_constants(orc, golden):
    # testsuite/test_rys_roots.py:64-65
    np.testing.assert_allclose(orc.boys(2, 1.0), golden["golden"]["boys_x1"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(orc.boys(2, 7.0), golden["golden"]["boys_x7"], rtol=0, atol=1e-15)


def test_boys_vs_mpmath(orc):
    mp.mp.dps = 30
    for x in (0.0, 1e-12, 0.3, 5.0, 29.9, 60.0, 150.0, 250.0):
        f = orc.boys(18, x)
        for k in range(19):
            ref = float(mp_boys(k, x))
            assert abs(f[k] - ref) <= 2e-16 * ref, (x, k)


def test_rys_fingerprints(orc, golden):
    # testsuite/test_rys_roots.py:72-85 (the reference generated these with
    # a 15-digit mpmath run; the n=6 pair carries that run's ~1e-10 error)
    for n, x, ref in golden["golden"]["fp_rys"]:
        u, w = orc.rys_roots(n, x)
        assert abs(fp(np.array([u, w])) - ref) < 1e-12 * max(1.0, abs(ref)) * 10, (n, x)
    x, fr, fw = golden["golden"]["fp_rys6"]
    u, w = orc.rys_roots(6, x)
    assert abs(fp(u) - fr) < 1e-9
    assert abs(fp(w) - fw) < 1e-9


@pytest.mark.parametrize("exact", [True, False])
def test_rys_moments(orc, exact):
    """n-point Gauss rule of exp(-x t^2) on [0,1]: sum w s^k = F_k(x), k < 2n."""
    mp.mp.dps = 30
    rng = np.random.default_rng(7)
    xs = [0.0, 1e-7, 0.37, 2.5, 11.2, 33.3, 47.9, 70.1, 95.0, 300.0] + list(rng.uniform(0, 100, 6))
    for n in range(1, 10):
        for x in xs:
            u, w = orc.rys_roots(n, x, exact=exact)
            s = u / (1 + u)
            assert np.all(w > 0) and np.all(np.diff(s) > 0)
            for k in range(2 * n):
                ref = float(mp_boys(k, x))
                got = float(np.sum(w * s ** k))
                assert abs(got - ref) <= 2e-13 * ref, (n, x, k, got, ref)


def test_rys_table_vs_exact(orc):
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(0, 1, 40), rng.uniform(0, 100, 160), np.arange(0, 95, 1.0)])
    for n in range(1, 10):
        for x in xs:
            u0, w0 = orc.rys_roots(n, x, exact=True)
            u1, w1 = orc.rys_roots(n, x, exact=False)
            np.testing.assert_allclose(u1, u0, rtol=2e-14, atol=0)
            np.testing.assert_allclose(w1, w0, rtol=2e-14, atol=0)


def _sphere_monomial(a, b, c):
    if a % 2 or b % 2 or c % 2:
        return 0.0
    g = math.gamma
    return 2 * g((a + 1) / 2) * g((b + 1) / 2) * g((c + 1) / 2) / g((a + b + c + 3) / 2)


def test_c2s_orthonormal(orc):
    for l in range(2, 7):
        C = orc.c2s_matrix(l)
        comps = [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]
        nc = len(comps)
        S = np.array([[_sphere_monomial(*(np.add(comps[i], comps[j]))) for j in range(nc)]
                      for i in range(nc)])
        np.testing.assert_allclose(C @ S @ C.T, np.eye(2 * l + 1), atol=1e-13)
    # d shell values of libcint's table (src/cart2sph.c:49-80)
    d = orc.c2s_matrix(2)
    assert abs(d[0, 1] - 1.092548430592079070) < 1e-15
    assert abs(d[2, 5] - 0.630783130505040012) < 1e-15
    assert abs(d[2, 0] + 0.315391565252520002) < 1e-15
    assert abs(d[4, 3] + 0.546274215296039535) < 1e-15
    np.testing.assert_array_equal(orc.c2s_matrix(1), np.eye(3))


def _libcint_c2s():
    """libcint's g_trans_cart2sph values (src/cart2sph.c:21-806) for l <= 6,
    committed as data by tests/golden/make_c2s_fixture.py."""
    with open(os.path.join(ROOT, "tests", "golden", "libcint_cart2sph.json")) as fh:
        d = json.load(fh)
    return {int(l): np.array(m) for l, m in d["l"].items()}


def _device_c2s_tables():
    """The device's generated table (libcint_amd/csrc/c2s_tables.inc), parsed."""
    import re
    with open(os.path.join(ROOT, "libcint_amd", "csrc", "c2s_tables.inc")) as fh:
        txt = fh.read()
    off = [int(x) for x in re.search(r"c2s_offset\[\d+\] = \{([^}]*)\}", txt).group(1).split(",")]
    coef = np.array([float(x) for x in
                     re.search(r"c2s_coef\[\d+\] = \{([^}]*)\}", txt).group(1).split(",")
                     if x.strip()])
    out = {}
    for l in range(len(off)):
        nc = (l + 1) * (l + 2) // 2
        out[l] = coef[off[l]:off[l] + (2 * l + 1) * nc].reshape(2 * l + 1, nc)
    return out


def _c2s_mismatch(a, b):
    """Rows of a and b that differ beyond rounding (sign flips included)."""
    tol = 1e-14 + 1e-13 * np.abs(b)
    return sorted(set(np.nonzero(np.abs(a - b) > tol)[0].tolist()))


def test_c2s_matches_libcint_table(orc):
    """Oracle and device transform equal libcint's table entry by entry for
    l = 0..6 (every sign of every row, p as px, py, pz)."""
    ref = _libcint_c2s()
    dev = _device_c2s_tables()
    assert sorted(ref) == list(range(7))
    for l in range(7):
        assert _c2s_mismatch(orc.c2s_matrix(l), ref[l]) == [], ("oracle", l)
        assert _c2s_mismatch(dev[l], ref[l]) == [], ("device table", l)


def test_c2s_check_sees_a_flipped_row():
    """The comparison above fails when one row of an f or g block changes
    sign (the Condon-Shortley / m<0 convention mistake it guards against)."""
    ref = _libcint_c2s()
    for l, row in ((3, 1), (3, 6), (4, 0), (2, 3)):
        bad = ref[l].copy()
        bad[row] *= -1
        assert _c2s_mismatch(bad, ref[l]) == [row]


@pytest.mark.parametrize("name", ["int2e_sph", "int2e_cart", "int2e_ip1_sph", "int3c2e_sph"])
def test_signed_class_fingerprints(orc, fixture_basis, name):
    """The oracle reproduces the committed signed per-class fingerprints
    (tests/golden/signed_fingerprints.json; regression pin of the oracle,
    which the GPU parity suite checks the device against)."""
    with open(os.path.join(ROOT, "tests", "golden", "signed_fingerprints.json")) as fh:
        gold = json.load(fh)[name]
    q = fixture_quartets(name)
    orc.set_root_mode(orc.ROOTS_EXACT)
    out, off, _ = orc.batch(name, q, orc.Basis(*fixture_basis))
    fps, mag = class_fingerprints(fixture_basis[1], q, out, off, name)
    check_class_fingerprints(gold, fps, mag, name)


@pytest.mark.parametrize("mode", ["table", "exact"])
def test_int2e_checksums(orc, golden, fixture_basis, mode):
    """testsuite/test_cint.py:478-481 over all 8 shells (l <= 4, nroots <= 9)."""
    orc.set_root_mode(orc.ROOTS_EXACT if mode == "exact" else orc.ROOTS_TABLE)
    try:
        B = orc.Basis(*fixture_basis)
        q = libcint_quartets()
        for name, key in (("int2e_sph", "cint2e_sph"), ("int2e_ip1_sph", "cint2e_ip1_sph")):
            out, off, not0 = orc.batch(name, q, B)
            v = float(np.abs(out).sum())
            ref, place = golden["golden"][key]
            assert close_place(v, ref, out.size, place), (name, v, ref)
            assert abs(v - ref) < 1e-12 * ref, (name, v, ref)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)


def test_int3c2e_checksum_and_s0_identity(orc, golden, fixture_basis):
    """testsuite/test_3c2e.py:170-202 and :303-307."""
    atm, bas, env = fixture_basis
    bas = bas.copy()
    v, cnt = 0.0, 0
    for k in range(4):
        bas[8, 0] = bas[k, 0]  # s0 shell on the atom of k
        B = orc.Basis(atm, bas, env)
        for j in range(4):
            for i in range(4):
                a, _ = orc.intor("int3c2e_sph", (i, j, k), B)
                ref, _ = orc.intor("int2e_sph", (i, j, k, 8), B)
                tol = 1e-12 * max(1.0, np.abs(ref).max())
                np.testing.assert_allclose(a.ravel(order="F"), ref.ravel(order="F"), rtol=1e-11,
                                           atol=tol)
                v += float(np.abs(a).sum())
                cnt += a.size
    gref, place = golden["golden"]["cint3c2e_sph"]
    assert close_place(v, gref, cnt, place), (v, gref)


def _single_s_basis(exps, coefs, centres):
    atm, bas, env = [], [], [0.0] * 20
    for a, c, r in zip(exps, coefs, centres):
        atm.append([1, len(env), 1, 0, 0, 0])
        env.extend(r)
        bas.append([len(atm) - 1, 0, 1, 1, 0, len(env), len(env) + 1, 0])
        env.extend([a, c])
    return np.array(atm, np.int32), np.array(bas, np.int32), np.array(env)


def test_ssss_closed_form(orc):
    """(ss|ss) = 2 pi^(5/2) / (p q sqrt(p+q)) K_AB K_CD F_0(T) on raw Gaussians."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        ex = rng.uniform(0.1, 5.0, 4)
        cf = rng.uniform(0.5, 2.0, 4)
        cen = rng.uniform(-2, 2, (4, 3))
        B = orc.Basis(*_single_s_basis(ex, cf, cen))
        val, _ = orc.intor("int2e_sph", (0, 1, 2, 3), B)
        a, b, c, d = ex
        p, q = a + b, c + d
        P = (a * cen[0] + b * cen[1]) / p
        Q = (c * cen[2] + d * cen[3]) / q
        kab = math.exp(-a * b / p * np.sum((cen[0] - cen[1]) ** 2))
        kcd = math.exp(-c * d / q * np.sum((cen[2] - cen[3]) ** 2))
        T = p * q / (p + q) * np.sum((P - Q) ** 2)
        f0 = float(mp_boys(0, T))
        ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * kab * kcd * f0
        ref *= np.prod(cf) * (0.5 / math.sqrt(math.pi)) ** 4
        assert abs(val.item() - ref) <= 1e-14 * abs(ref), (val.item(), ref)


def test_permutational_symmetry(orc):
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c0")
    B = orc.Basis(m.atm, m.bas, m.env)
    rng = np.random.default_rng(5)
    for _ in range(25):
        i, j, k, l = rng.integers(0, m.nbas, 4)
        a, _ = orc.intor("int2e_sph", (i, j, k, l), B)
        b, _ = orc.intor("int2e_sph", (j, i, k, l), B)
        c, _ = orc.intor("int2e_sph", (k, l, i, j), B)
        d, _ = orc.intor("int2e_sph", (i, j, l, k), B)
        np.testing.assert_allclose(b.transpose(1, 0, 2, 3), a, rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(c.transpose(2, 3, 0, 1), a, rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(d.transpose(0, 1, 3, 2), a, rtol=1e-11, atol=1e-14)


def test_ip1_finite_difference(orc):
    """int2e_ip1 = <d/dr i> = -d/dA_i of int2e_cart (shell i alone on its atom)."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c3")
    # give shell i its own atom by appending a ghost atom at the same place
    rng = np.random.default_rng(9)
    h = 1e-4
    done = 0
    while done < 6:
        i, j, k, l = (int(v) for v in rng.integers(0, m.nbas, 4))
        if i in (j, k, l):
            continue  # only shell i may move
        done += 1
        atm = np.vstack([m.atm, m.atm[m.bas[i, 0]]]).astype(np.int32)
        env = np.concatenate([m.env, m.env[m.atm[m.bas[i, 0], 1]:m.atm[m.bas[i, 0], 1] + 3]])
        atm[-1, 1] = len(m.env)
        bas = m.bas.copy()
        bas[i, 0] = len(atm) - 1
        g, _ = orc.intor("int2e_ip1_cart", (i, j, k, l), orc.Basis(atm, bas, env))
        for d in range(3):
            ep, em = env.copy(), env.copy()
            ep[len(m.env) + d] += h
            em[len(m.env) + d] -= h
            bp = orc.Basis(atm, bas, ep)
            bm = orc.Basis(atm, bas, em)
            # shells j,k,l may share the original atom: they stay put
            vp, _ = orc.intor("int2e_cart", (i, j, k, l), bp)
            vm, _ = orc.intor("int2e_cart", (i, j, k, l), bm)
            fd = -(vp - vm) / (2 * h)
            np.testing.assert_allclose(g[..., d], fd, rtol=0, atol=2e-7 * max(1.0, np.abs(fd).max()))


def test_sr_rys_moments(orc):
    """Short-range rule (src/rys_roots.c:149-267): the n-point Gauss rule of
    exp(-x t^2) on [lower, 1], checked by its moments against mpmath's
    quadrature of the same integral, k < 2n."""
    mp.mp.dps = 30
    for n in (1, 2, 3, 4, 5, 6, 7):
        for x in (0.0, 0.3, 2.5, 11.0, 37.0):
            for lower in (0.05, 0.4, 0.8, 0.97):
                u, w = orc.sr_rys_roots(n, x, lower)
                s = u / (1 + u)
                assert np.all(w > 0) and np.all(s >= lower * lower * (1 - 1e-12))
                for k in range(2 * n):
                    ref = float(mp.quad(lambda t: t ** (2 * k) * mp.e ** (-x * t * t), [lower, 1]))
                    got = float(np.sum(w * s ** k))
                    assert abs(got - ref) <= 1e-13 * ref, (n, x, lower, k, got, ref)


@pytest.mark.parametrize("omega", [0.2, 0.5, 0.8])
def test_erf_plus_erfc_is_coulomb(orc, fixture_basis, omega):
    """testsuite/test_cint.py:258-302: erf(w r)/r + erfc(w r)/r = 1/r for
    every quartet (i<=j, k<=l) of the fixture to 9 places, with the oracle's
    short-range path (doubled roots for rys_order <= 3, the [sqrt(theta), 1]
    rule above) and long-range path; also for (ij|k) 3-centre integrals."""
    atm, bas, env = fixture_basis
    q = libcint_quartets()
    q3 = np.array([(i, j, k, 0) for k in range(8) for j in range(4) for i in range(4)], np.int32)
    res = {}
    for o in (0.0, omega, -omega):
        e2 = np.array(env, copy=True)
        e2[8] = o
        B = orc.Basis(atm, bas, e2)
        res[o] = (orc.batch("int2e_sph", q, B, nthreads=8)[0],
                  orc.batch("int3c2e_sph", q3, B, nthreads=8)[0])
    for t in range(2):
        d = res[omega][t] + res[-omega][t] - res[0.0][t]
        assert np.round(np.abs(d).max(), 9) == 0, (omega, t, np.abs(d).max())
        # the short-range part is a genuine fraction of the full integral
        assert np.abs(res[-omega][t]).max() > 1e-3 * np.abs(res[0.0][t]).max()


def _own_atom(m, shell):
    """atm/bas/env with `shell` moved to a new atom at the same place; returns
    (atm, bas, env, offset of the new atom's coordinates in env)."""
    a0 = m.bas[shell, 0]
    atm = np.vstack([m.atm, m.atm[a0]]).astype(np.int32)
    env = np.concatenate([m.env, m.env[m.atm[a0, 1]:m.atm[a0, 1] + 3]])
    atm[-1, 1] = len(m.env)
    bas = m.bas.copy()
    bas[shell, 0] = len(atm) - 1
    return atm, bas, env, len(m.env)


@pytest.mark.parametrize("name,slot", [("int3c2e_ip1_cart", 0), ("int3c2e_ip2_cart", 2)])
def test_3c2e_derivatives_finite_difference(orc, name, slot):
    """int3c2e_ip1 = <d/dr i> = -d/dA_i and int3c2e_ip2 = <d/dr k> = -d/dC_k of
    int3c2e_cart, by central differences (h = 1e-4) with the differentiated
    shell alone on its atom; the c4 cluster's orbital and auxiliary shells."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    rng = np.random.default_rng(31 + slot)
    h = 1e-4
    done = 0
    while done < 6:
        t = [int(rng.integers(0, m.aux_start)), int(rng.integers(0, m.aux_start)),
             int(rng.integers(m.aux_start, m.nbas))]
        if t[slot] in [t[s] for s in range(3) if s != slot]:
            continue
        done += 1
        atm, bas, env, pc = _own_atom(m, t[slot])
        q = (t[0], t[1], t[2], 0)
        g, _ = orc.intor(name, q, orc.Basis(atm, bas, env))
        for d in range(3):
            ep, em = env.copy(), env.copy()
            ep[pc + d] += h
            em[pc + d] -= h
            vp, _ = orc.intor("int3c2e_cart", q, orc.Basis(atm, bas, ep))
            vm, _ = orc.intor("int3c2e_cart", q, orc.Basis(atm, bas, em))
            fd = -(vp - vm) / (2 * h)
            np.testing.assert_allclose(g[..., d], fd, rtol=0,
                                       atol=2e-7 * max(1.0, np.abs(fd).max()), err_msg=str(q))


def test_3c2e_translational_invariance(orc, fixture_basis):
    """(d/dr i j|k) + (i d/dr j|k) + (i j|d/dr k) = 0 for every triple of the
    fixture: ties the signs and index orders of int3c2e_ip1 and int3c2e_ip2
    together (the second term is int3c2e_ip1 of (j, i, k), transposed)."""
    B = orc.Basis(*fixture_basis)
    worst = 0.0
    for i in range(4):
        for j in range(4):
            for k in range(4):
                a, _ = orc.intor("int3c2e_ip1_sph", (i, j, k, 0), B)
                b, _ = orc.intor("int3c2e_ip1_sph", (j, i, k, 0), B)
                c, _ = orc.intor("int3c2e_ip2_sph", (i, j, k, 0), B)
                tot = a + b.transpose(1, 0, 2, 3) + c
                worst = max(worst, float(np.abs(tot).max()))
                assert np.abs(tot).max() < 1e-11 * max(1.0, np.abs(a).max()), (i, j, k)
    assert worst < 1e-10


def test_int2c2e_checksum_symmetry_and_3c_identity(orc, golden, fixture_basis):
    """testsuite/test_3c2e.py:311-319: sum |int2c2e_sph(i, k)| over the four
    regular shells is 782.3104849606677 (10 places); (i|k) = (k|i)^T; and
    (i|k) equals int3c2e_sph(i, s0, k) with the fixture's exponent-0 shell."""
    B = orc.Basis(*fixture_basis)
    q = np.array([(i, k, 0, 0) for k in range(4) for i in range(4)], np.int32)
    out, off, nz = orc.batch("int2c2e_sph", q, B)
    assert np.all(nz == 1)
    v = float(np.abs(out).sum())
    assert close_place(v, 782.3104849606677, out.size, 10), v
    q3 = np.array([(i, 8, k, 0) for k in range(4) for i in range(4)], np.int32)
    o3, _, _ = orc.batch("int3c2e_sph", q3, B)
    np.testing.assert_allclose(out, o3, rtol=0, atol=1e-12)
    for i in range(4):
        for k in range(4):
            a, _ = orc.intor("int2c2e_sph", (i, k), B)
            b, _ = orc.intor("int2c2e_sph", (k, i), B)
            np.testing.assert_allclose(a, b.T, rtol=0, atol=1e-14 * max(1.0, np.abs(a).max()))


@pytest.mark.parametrize("name,ref", [("int3c2e_ip1_sph", 2242.052249221302),
                                      ("int3c2e_ip2_sph", 1970.982483824248)])
def test_int3c2e_derivative_checksums(orc, fixture_basis, name, ref):
    """testsuite/test_3c2e.py:303-307: sum |x| over all (i, j, k) of the four
    regular shells, 3 components, to 10 places."""
    q = np.array([(i, j, k, 0) for k in range(4) for j in range(4) for i in range(4)], np.int32)
    out, _, _ = orc.batch(name, q, orc.Basis(*fixture_basis))
    v = float(np.abs(out).sum())
    assert close_place(v, ref, out.size, 10), (name, v, ref)


def test_oracle_under_address_and_undefined_behaviour_sanitizers(golden, fixture_basis, tmp_path):
    """The oracle built with -fsanitize=address,undefined (findings abort)
    over libcint's test-suite loops for every kind, exact and tabulated
    roots: a clean exit and libcint's golden checksums
    (testsuite/test_cint.py:478-481, test_3c2e.py:303-320)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    atm, bas, env = fixture_basis
    path = tmp_path / "basis.txt"
    with open(path, "w") as fh:
        fh.write(f"{atm.shape[0]} {bas.shape[0]} {env.size}\n")
        fh.write(" ".join(map(str, atm.ravel())) + "\n" + " ".join(map(str, bas.ravel())) + "\n")
        fh.write(" ".join(repr(float(x)) for x in env) + "\n")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle_sanitize"), str(path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    sums = {}
    for line in r.stdout.splitlines():
        mode, name, v = line.split()
        sums[(mode, name)] = float(v)
    gold = {"int2e_sph": (golden["golden"]["cint2e_sph"][0], 7),
            "int2e_ip1_sph": (golden["golden"]["cint2e_ip1_sph"][0], 7),
            "int3c2e_sph": (golden["golden"]["cint3c2e_sph"][0], 9),
            "int3c2e_ip1_sph": (2242.052249221302, 9), "int3c2e_ip2_sph": (1970.982483824248, 9),
            "int2c2e_sph": (782.3104849606677, 9)}
    for mode in ("exact", "table"):
        for name, (ref, place) in gold.items():
            assert round(abs(sums[(mode, name)] - ref), place - 2) == 0, (mode, name, sums[(mode, name)], ref)


def test_ip2_finite_difference(orc):
    """int2e_ip2 = (i j|d/dr k l) = -d/dC_k of int2e_cart by central
    differences (h = 1e-4), shell k alone on its atom (src/autocode/grad2.c:101-170)."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c3")
    rng = np.random.default_rng(19)
    h = 1e-4
    done = 0
    while done < 6:
        q = tuple(int(v) for v in rng.integers(0, m.nbas, 4))
        if q[2] in (q[0], q[1], q[3]):
            continue
        done += 1
        atm, bas, env, pc = _own_atom(m, q[2])
        g, _ = orc.intor("int2e_ip2_cart", q, orc.Basis(atm, bas, env))
        for d in range(3):
            ep, em = env.copy(), env.copy()
            ep[pc + d] += h
            em[pc + d] -= h
            vp, _ = orc.intor("int2e_cart", q, orc.Basis(atm, bas, ep))
            vm, _ = orc.intor("int2e_cart", q, orc.Basis(atm, bas, em))
            fd = -(vp - vm) / (2 * h)
            np.testing.assert_allclose(g[..., d], fd, rtol=0,
                                       atol=2e-7 * max(1.0, np.abs(fd).max()), err_msg=str(q))


@pytest.mark.parametrize("name,first,slot", [("int2e_ipip1_cart", "int2e_ip1_cart", 0),
                                              ("int2e_ipip1_sph", "int2e_ip1_sph", 0),
                                              ("int2e_ip1ip2_cart", "int2e_ip1_cart", 2),
                                              ("int2e_ip1ip2_sph", "int2e_ip1_sph", 2)])
def test_second_derivatives_finite_difference(orc, name, first, slot):
    """Component 3p+q of int2e_ipip1 = (d2/dr_p dr_q i j|k l) is -d/dA_q of
    component p of int2e_ip1 (shell i moved), and of int2e_ip1ip2 =
    (d/dr_p i j|d/dr_q k l) it is -d/dC_q of the same (shell k moved); central
    differences (h = 1e-4) of the ip1 integrals, which the cint2e_ip1_sph
    checksum pins (src/autocode/hess.c:1125-1174, 1289-1338)."""
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c3")
    rng = np.random.default_rng(41 + slot + (name.endswith("sph")))
    h = 1e-4
    done = 0
    while done < 5:
        q = tuple(int(v) for v in rng.integers(0, m.nbas, 4))
        if q[slot] in [q[s] for s in range(4) if s != slot]:
            continue
        done += 1
        atm, bas, env, pc = _own_atom(m, q[slot])
        g, _ = orc.intor(name, q, orc.Basis(atm, bas, env))
        assert g.shape[-1] == 9
        for d in range(3):
            ep, em = env.copy(), env.copy()
            ep[pc + d] += h
            em[pc + d] -= h
            vp, _ = orc.intor(first, q, orc.Basis(atm, bas, ep))
            vm, _ = orc.intor(first, q, orc.Basis(atm, bas, em))
            fd = -(vp - vm) / (2 * h)
            for p_ in range(3):
                np.testing.assert_allclose(g[..., 3 * p_ + d], fd[..., p_], rtol=0,
                                           atol=2e-7 * max(1.0, np.abs(fd).max()),
                                           err_msg="%s %s p=%d q=%d" % (name, q, p_, d))


def test_second_derivative_symmetries(orc, fixture_basis):
    """Over quartets of libcint's fixture: int2e_ipip1 is symmetric in its two
    axes (component 3p+q = 3q+p), and (d_p i j|d_q k l) equals (d_q k l|d_p i j)
    with the shell axes transposed (int2e_ip1ip2 of (k, l, i, j), component
    3q+p), which fixes which operator the first index of a component names."""
    B = orc.Basis(*fixture_basis)
    rng = np.random.default_rng(8)
    for _ in range(25):
        i, j, k, l = (int(v) for v in rng.integers(0, 8, 4))
        a, _ = orc.intor("int2e_ipip1_sph", (i, j, k, l), B)
        scale = max(1.0, float(np.abs(a).max()))
        a9 = a.reshape(a.shape[:4] + (3, 3))
        np.testing.assert_allclose(a9, a9.swapaxes(4, 5), rtol=0, atol=1e-12 * scale)
        c, _ = orc.intor("int2e_ip1ip2_sph", (i, j, k, l), B)
        c_swap, _ = orc.intor("int2e_ip1ip2_sph", (k, l, i, j), B)
        scale = max(1.0, float(np.abs(c).max()))
        c9 = c.reshape(c.shape[:4] + (3, 3))
        s9 = c_swap.reshape(c_swap.shape[:4] + (3, 3)).transpose(2, 3, 0, 1, 5, 4)
        np.testing.assert_allclose(c9, s9, rtol=0, atol=1e-12 * scale)


def test_ip2_electron_swap_and_translational_invariance(orc, fixture_basis):
    """Over quartets of libcint's fixture: (ij|d/dr k l) equals (d/dr k l|ij)
    with the axes transposed (int2e_ip1 of (k, l, i, j), pinned by the
    cint2e_ip1_sph checksum), and the derivatives on the four centres sum to
    zero: (d i j|kl) + (i d j|kl) + (ij|d k l) + (ij|k d l) = 0."""
    B = orc.Basis(*fixture_basis)
    rng = np.random.default_rng(5)
    for _ in range(40):
        i, j, k, l = (int(v) for v in rng.integers(0, 8, 4))
        c, _ = orc.intor("int2e_ip2_sph", (i, j, k, l), B)
        c_swap, _ = orc.intor("int2e_ip1_sph", (k, l, i, j), B)
        scale = max(1.0, float(np.abs(c).max()))
        np.testing.assert_allclose(c, c_swap.transpose(2, 3, 0, 1, 4), rtol=0, atol=1e-12 * scale)
        a, _ = orc.intor("int2e_ip1_sph", (i, j, k, l), B)
        b, _ = orc.intor("int2e_ip1_sph", (j, i, k, l), B)
        d, _ = orc.intor("int2e_ip2_sph", (i, j, l, k), B)
        tot = a + b.transpose(1, 0, 2, 3, 4) + c + d.transpose(0, 1, 3, 2, 4)
        assert np.abs(tot).max() < 1e-10 * scale, ((i, j, k, l), float(np.abs(tot).max()))
