"""GPU J/K build, two-electron energy and nuclear gradient vs the CPU oracle.

J and K are checked against a dense contraction of the oracle's full ERI
tensor; the gradient against (1) the oracle's int2e_ip1 blocks contracted
with the density on the CPU and (2) central finite differences of the GPU
energy (step 1e-4 Bohr, every coordinate).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-jk]", *a, flush=True)
    sys.stdout.flush()


def _full_eri_oracle(orc, m, nshell):
    from libcint_amd.synthetic import unique_quartets
    B = orc.Basis(m.atm, m.bas, m.env)
    q = unique_quartets(nshell)
    out, off, not0 = orc.batch("int2e_sph", q, B, nthreads=16)
    ao = m.ao_loc_sph(nshell)
    nao = int(ao[-1])
    eri = np.zeros((nao, nao, nao, nao))
    for n, (i, j, k, l) in enumerate(q):
        d = [ao[s + 1] - ao[s] for s in (i, j, k, l)]
        blk = out[off[n]:off[n + 1]].reshape(d, order="F")
        si, sj, sk, sl = (slice(ao[s], ao[s + 1]) for s in (i, j, k, l))
        for a, b, c, e, t in ((si, sj, sk, sl, (0, 1, 2, 3)), (sj, si, sk, sl, (1, 0, 2, 3)),
                              (si, sj, sl, sk, (0, 1, 3, 2)), (sj, si, sl, sk, (1, 0, 3, 2)),
                              (sk, sl, si, sj, (2, 3, 0, 1)), (sl, sk, si, sj, (3, 2, 0, 1)),
                              (sk, sl, sj, si, (2, 3, 1, 0)), (sl, sk, sj, si, (3, 2, 1, 0))):
            eri[a, b, c, e] = blk.transpose(t)
    return eri, (q, out, off, not0)


@pytest.fixture(scope="module")
def c0_setup(orc):
    from libcint_amd.synthetic import make_molecule, random_density
    m = make_molecule("c0")
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        eri, raw = _full_eri_oracle(orc, m, m.nbas)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)
    dm = random_density(m.nao_sph(), seed=7)
    return m, eri, dm, raw


def test_jk_and_energy_vs_oracle(c0_setup):
    import libcint_amd as mc
    from libcint_amd import jk
    m, eri, dm, _ = c0_setup
    _say("c0 J/K")
    vj, vk = jk.get_jk(mc.Basis(m.atm, m.bas, m.env), dm)
    vj, vk = vj.cpu().numpy(), vk.cpu().numpy()
    J = np.einsum("ijkl,kl->ij", eri, dm)
    K = np.einsum("ijkl,jl->ik", eri, dm)
    np.testing.assert_allclose(vj, J, rtol=0, atol=1e-11 * max(1, np.abs(J).max()))
    np.testing.assert_allclose(vk, K, rtol=0, atol=1e-11 * max(1, np.abs(K).max()))
    e_ref = 0.5 * np.sum(dm * J) - 0.25 * np.sum(dm * K)
    e = 0.5 * np.sum(dm * vj) - 0.25 * np.sum(dm * vk)
    assert abs(e - e_ref) < 1e-10 * max(1.0, abs(e_ref))


def test_fused_jk_paths(c0_setup):
    """J/K contracted inside the integral kernels (no ERI buffer) on every
    kernel path: the register kernels (default), the generic kernel alone,
    and the stored-buffer digest, each against the oracle's dense
    contraction; J alone and K alone leave the other matrix untouched."""
    import torch
    import libcint_amd as mc
    from libcint_amd import jk
    m, eri, dm, _ = c0_setup
    B = mc.Basis(m.atm, m.bas, m.env)
    J = np.einsum("ijkl,kl->ij", eri, dm)
    K = np.einsum("ijkl,jl->ik", eri, dm)
    tol_j = 1e-11 * max(1, np.abs(J).max())
    tol_k = 1e-11 * max(1, np.abs(K).max())
    for label, kw, generic in (("stored", {"fused": False}, 0), ("fused", {}, 0),
                               ("fused generic", {}, 1)):
        _say("c0 J/K", label)
        old = mc.tune(mc.TUNE_GENERIC_ONLY, generic)
        try:
            vj, vk = jk.get_jk(B, dm, **kw)
        finally:
            mc.tune(mc.TUNE_GENERIC_ONLY, old)
        np.testing.assert_allclose(vj.cpu().numpy(), J, rtol=0, atol=tol_j, err_msg=label)
        np.testing.assert_allclose(vk.cpu().numpy(), K, rtol=0, atol=tol_k, err_msg=label)
    jb = jk.JKBuilder(B)
    try:
        vj, _ = jb.fused_jk(dm, with_k=False)
        _, vk = jb.fused_jk(dm, with_j=False)
        torch.cuda.synchronize()
    finally:
        jb.close()
    np.testing.assert_allclose(vj.cpu().numpy(), J, rtol=0, atol=tol_j)
    np.testing.assert_allclose(vk.cpu().numpy(), K, rtol=0, atol=tol_k)


def test_fused_jk_long_range_and_short_range_refusal():
    """Long range (omega = 0.3): the fused J/K equals the stored-buffer digest
    of the erf-attenuated integrals (c3); short range (omega < 0) has no fused
    digest and is refused with a status, not computed wrongly."""
    import libcint_amd as mc
    from libcint_amd import jk
    from libcint_amd.synthetic import make_molecule, random_density
    m = make_molecule("c3")
    dm = random_density(m.nao_sph(), seed=6)
    env = m.env.copy()
    env[8] = 0.3
    B = mc.Basis(m.atm, m.bas, env)
    vj_s, vk_s = jk.get_jk(B, dm, fused=False)
    vj_f, vk_f = jk.get_jk(B, dm)
    dj = float((vj_f - vj_s).abs().max())
    dk = float((vk_f - vk_s).abs().max())
    _say(f"c3 long range fused vs stored: max|dJ| {dj:.2e} max|dK| {dk:.2e}")
    assert dj < 1e-12 * max(1.0, float(vj_s.abs().max()))
    assert dk < 1e-12 * max(1.0, float(vk_s.abs().max()))
    env[8] = -0.3
    with pytest.raises(mc.McintError):
        jk.get_jk(mc.Basis(m.atm, m.bas, env), dm)


def test_fused_jk_matches_stored_on_c3():
    """c3 (20 atoms, d shells): the fused J/K equals the stored-buffer
    digest, in FP64 and with the FP32 tier (the same blocks, summed in
    another order)."""
    import libcint_amd as mc
    from libcint_amd import jk
    from libcint_amd.synthetic import make_molecule, random_density
    m = make_molecule("c3")
    B = mc.Basis(m.atm, m.bas, m.env)
    dm = random_density(m.nao_sph(), seed=5)
    for thr in (0.0, 1e-7):
        _say("c3 fused vs stored J/K, fp32 threshold", thr)
        vj_s, vk_s = jk.get_jk(B, dm, fp32_threshold=thr, fused=False)
        vj_f, vk_f = jk.get_jk(B, dm, fp32_threshold=thr)
        sj = max(1.0, float(vj_s.abs().max()))
        sk = max(1.0, float(vk_s.abs().max()))
        dj = float((vj_f - vj_s).abs().max())
        dk = float((vk_f - vk_s).abs().max())
        _say(f"max|dJ| {dj:.2e} max|dK| {dk:.2e}")
        assert dj < 1e-12 * sj and dk < 1e-12 * sk


def _small(natoms=5):
    from libcint_amd.synthetic import make_molecule, random_density, subset
    m = subset(make_molecule("c3"), natoms)
    dm = random_density(m.nao_sph(), seed=3)
    return m, dm


def test_gradient_vs_oracle_ip1(orc):
    import libcint_amd as mc
    from libcint_amd import jk
    m, dm = _small(4)
    _say("gradient vs oracle ip1", m.nbas, "shells")
    g = jk.nuclear_gradient(mc.Basis(m.atm, m.bas, m.env), dm, m.natm).cpu().numpy()
    q = jk.gradient_quartets(m.nbas)
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        out, off, _ = orc.batch("int2e_ip1_sph", q, orc.Basis(m.atm, m.bas, m.env), nthreads=16)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)
    ao = m.ao_loc_sph()
    ref = np.zeros((m.natm, 3))
    for n, (i, j, k, l) in enumerate(q):
        d = [ao[s + 1] - ao[s] for s in (i, j, k, l)]
        blk = out[off[n]:off[n + 1]].reshape(d + [3], order="F")
        Dij = dm[ao[i]:ao[i + 1], ao[j]:ao[j + 1]]
        Dkl = dm[ao[k]:ao[k + 1], ao[l]:ao[l + 1]]
        Dik = dm[ao[i]:ao[i + 1], ao[k]:ao[k + 1]]
        Djl = dm[ao[j]:ao[j + 1], ao[l]:ao[l + 1]]
        Dil = dm[ao[i]:ao[i + 1], ao[l]:ao[l + 1]]
        Djk = dm[ao[j]:ao[j + 1], ao[k]:ao[k + 1]]
        w = -2 * np.einsum("ab,cd->abcd", Dij, Dkl) + 0.5 * (np.einsum("ac,bd->abcd", Dik, Djl)
                                                           + np.einsum("ad,bc->abcd", Dil, Djk))
        if k != l:
            w = 2 * w
        ref[m.bas[i, 0]] += np.einsum("abcdx,abcd->x", blk, w)
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-10 * max(1, np.abs(ref).max()))


def test_gradient_finite_difference():
    import libcint_amd as mc
    from libcint_amd import jk
    m, dm = _small(5)
    _say("gradient vs finite differences", m.natm, "atoms")
    g = jk.nuclear_gradient(mc.Basis(m.atm, m.bas, m.env), dm, m.natm).cpu().numpy()
    h = 1e-4

    def e_at(coords):
        mm = m.with_coords(coords)
        vj, vk = jk.get_jk(mc.Basis(mm.atm, mm.bas, mm.env), dm)
        return float(jk.energy(vj.new_tensor(dm), vj, vk))

    fd = np.zeros_like(g)
    for a in range(m.natm):
        for x in range(3):
            cp = m.coords.copy()
            cm = m.coords.copy()
            cp[a, x] += h
            cm[a, x] -= h
            fd[a, x] = (e_at(cp) - e_at(cm)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-7)
    # translational invariance of the analytic gradient
    assert np.abs(g.sum(axis=0)).max() < 1e-10 * max(1, np.abs(g).max())


def test_autograd_function_matches_digest():
    import torch
    import libcint_amd as mc
    from libcint_amd import jk
    m, dm = _small(4)
    _say("autograd")
    coords = torch.tensor(m.coords, dtype=torch.float64, requires_grad=True)
    dmt = torch.tensor(dm, dtype=torch.float64, device="cuda", requires_grad=True)
    e = jk.TwoElectronEnergy.apply(coords, dmt, m)
    e.backward()
    g = jk.nuclear_gradient(mc.Basis(m.atm, m.bas, m.env), dm, m.natm).cpu().numpy()
    np.testing.assert_allclose(coords.grad.numpy(), g, rtol=0, atol=1e-12)
    vj, vk = jk.get_jk(mc.Basis(m.atm, m.bas, m.env), dm)
    np.testing.assert_allclose(dmt.grad.cpu().numpy(), (vj - 0.5 * vk).cpu().numpy(), atol=1e-12)


def test_c0_every_unique_quartet_vs_oracle(c0_setup):
    """All 692,076 unique quartets of the 12-atom cluster: GPU blocks and
    non-zero flags element by element against the oracle's."""
    import libcint_amd as mc
    m, _, _, (q, ref, off, not0) = c0_setup
    _say("c0 all unique quartets", len(q))
    got, off_g, nz = mc.batch("int2e_sph", q, mc.Basis(m.atm, m.bas, m.env))
    np.testing.assert_array_equal(off_g, off)
    np.testing.assert_array_equal(nz, not0)
    err = np.abs(got - ref)
    bad = err > 1e-12 + 1e-10 * np.abs(ref)
    _say("c0 max |err|", float(err.max()), "elements", got.size)
    assert not bad.any(), (int(bad.sum()), float(err.max()))


def test_mixed_precision_jk(c0_setup):
    """FP32 tier for Schwarz*max|D| < 1e-7: J, K and energy against the
    oracle's dense contraction (errors stated) and against the all-FP64
    GPU result."""
    import libcint_amd as mc
    from libcint_amd import jk
    m, eri, dm, _ = c0_setup
    B = mc.Basis(m.atm, m.bas, m.env)
    vj, vk = jk.get_jk(B, dm)
    st = {}
    vj32, vk32 = jk.get_jk(B, dm, fp32_threshold=1e-7, stats=st)
    _say("mixed J/K", st)
    assert st["fp32_tasks"] > 0
    J = np.einsum("ijkl,kl->ij", eri, dm)
    K = np.einsum("ijkl,jl->ik", eri, dm)
    e_ref = 0.5 * np.sum(dm * J) - 0.25 * np.sum(dm * K)
    j32, k32 = vj32.cpu().numpy(), vk32.cpu().numpy()
    e32 = 0.5 * np.sum(dm * j32) - 0.25 * np.sum(dm * k32)
    dj_o = float(np.abs(j32 - J).max())
    dk_o = float(np.abs(k32 - K).max())
    _say(f"mixed vs oracle: max|dJ| {dj_o:.2e} max|dK| {dk_o:.2e} dE {e32 - e_ref:.2e}")
    assert dj_o < 1e-8 and dk_o < 1e-8
    assert abs(e32 - e_ref) < 1e-9
    dj = float((vj32 - vj).abs().max())
    dk = float((vk32 - vk).abs().max())
    de = float(jk.energy(vj.new_tensor(dm), vj32, vk32) - jk.energy(vj.new_tensor(dm), vj, vk))
    _say(f"mixed vs FP64: max|dJ| {dj:.2e} max|dK| {dk:.2e} dE {de:.2e}")
    assert dj < 1e-8 and dk < 1e-8
    assert abs(de) < 1e-9


def test_mixed_precision_gradient():
    """FP32 tier of the gradient digest (Schwarz x max|D|^2 < 1e-6): within
    1e-8 Ha/Bohr of the FP64 gradient."""
    import libcint_amd as mc
    from libcint_amd import jk
    m, dm = _small(6)
    B = mc.Basis(m.atm, m.bas, m.env)
    g64 = jk.nuclear_gradient(B, dm, m.natm).cpu().numpy()
    gp = jk.GradientPlan(B, m.natm, fp32_threshold=1e-6, dmax=float(np.abs(dm).max()))
    st = gp.stats()
    g32 = gp.run(dm).cpu().numpy()
    gp.close()
    err = float(np.abs(g32 - g64).max())
    _say("mixed gradient", st["fp32_tasks"], "of", st["tasks"], "FP32 tasks; max|dg|", err)
    assert st["fp32_tasks"] > 0
    assert err < 1e-8
