"""Full-size GPU checks on the synthetic workloads.

* Every bucket of the real plans: at least four quartets (tasks picked at
  random positions) from each (class, team size, kernel path) bucket and
  each original angular-momentum order (axis permutation) inside it of the c1 int2e_sph plan (205M Schwarz-surviving quartets of
  the 30-atom cluster), of the c3 ip1 gradient list, the c3 int2e_cart list
  and the c4 int3c2e list, checked element by element against the oracle.
* c2: the mixed-precision J/K build on the 30-atom cluster, FP32 tier for
  quartets with Schwarz bound x max|D| < 1e-7, against the all-FP64 build
  (max |dJ|, max |dK|, energy error stated and bounded).
* c3: the nuclear gradient of the 20-atom cluster against central finite
  differences of the GPU energy on all 60 coordinates (step 1e-4 Bohr) and
  against the oracle's int2e_ip1_sph contracted with the density on the CPU
  for the whole cluster (tests/golden/c3_gradient.json).
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import ROOT
from test_gpu_parity import _cmp, _cmp_ip1

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-full]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mc():
    import libcint_amd
    libcint_amd.lib()
    return libcint_amd


def _bucket_keys(plan):
    return sorted({(b["cls"], b["team"], b["path"]) for b in plan.buckets()})


def _order_code(bas, q):
    """The angular momenta of each tuple in its original order, packed: the
    axis map that canonicalises a tuple (the per-task permutation inside a
    bucket) is a function of this order."""
    code = np.zeros(len(q), np.int64)
    for a in range(4):
        code = code * 8 + bas[q[:, a], 1]
    return code


def _stratified_sample(plan, q, bas, per_group, seed):
    """per_group tasks at random positions of every (bucket, original
    angular-momentum order) group (all of a smaller group), as rows of the
    plan's list q; also the set of groups."""
    rng = np.random.default_rng(seed)
    picks, groups = [], set()
    for i, b in enumerate(plan.buckets()):
        ids = plan.bucket_tasks(i, 0, int(b["tasks"])).astype(np.int64)
        code = _order_code(bas, q[ids])
        for c in np.unique(code):
            sel = ids[code == c]
            groups.add((b["cls"], b["team"], b["path"], int(c)))
            picks.append(rng.choice(sel, size=min(per_group, sel.size), replace=False))
    return q[np.unique(np.concatenate(picks))], groups


def _check_stratified(mc, orc, m, name, q, per_bucket=4, seed=0):
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    plan = mc.Plan(db, name, q)
    keys = _bucket_keys(plan)
    sample, groups = _stratified_sample(plan, q, m.bas, per_bucket, seed)
    plan.close()
    splan = mc.Plan(db, name, sample)
    skeys = _bucket_keys(splan)
    sgroups = {(b["cls"], b["team"], b["path"], int(c))
               for i, b in enumerate(splan.buckets())
               for c in np.unique(_order_code(m.bas, sample[splan.bucket_tasks(i, 0, int(b["tasks"]))]))}
    splan.close()
    db.close()
    _say(m.name, name, len(q), "quartets,", len(keys), "buckets,", len(groups),
         "(bucket, axis order) groups; sample of", len(sample))
    assert skeys == keys, "the sample's plan must hold every bucket of the full plan"
    assert sgroups == groups, "the sample must hold every axis order of every bucket"
    got, off, nz = mc.batch(name, sample, mc.Basis(m.atm, m.bas, m.env))
    orc.set_root_mode(orc.ROOTS_TABLE)
    try:
        ref, off_r, nz_r = orc.batch(name, sample, orc.Basis(m.atm, m.bas, m.env), nthreads=16)
    finally:
        orc.set_root_mode(orc.ROOTS_EXACT)
    np.testing.assert_array_equal(off, off_r)
    if "ip1" in name:
        _cmp_ip1(mc, orc, m, sample, got, ref, off, f"{m.name} {name} stratified")
    else:
        _cmp(got, ref, f"{m.name} {name} stratified")
    np.testing.assert_array_equal(nz, nz_r)


def test_stratified_c1_int2e_sph(mc, orc):
    from libcint_amd.synthetic import make_molecule, screened_quartets
    m = make_molecule("c1")
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    q, _ = screened_quartets(db.schwarz(m.nbas), m.nbas, 1e-12)
    db.close()
    _check_stratified(mc, orc, m, "int2e_sph", q, seed=1)


def test_wave_kernels_match_chunk_kernels_c1(mc):
    """Every wave kernel (16, 32 or 64 lanes per quartet) against the
    component-chunk kernels of the same buckets: up to 20,000 quartets of
    each bucket (every team size: one to several candidate rounds and
    batches of primitive quartets) of c1's chunked classes give the same
    blocks on both paths, to the parity bar of the oracle comparisons
    (1e-12 + 1e-10 |x|; the two paths sum primitives in different orders),
    plain and long-range operator."""
    import torch
    from libcint_amd.synthetic import make_molecule, screened_quartets
    m = make_molecule("c1")
    B = mc.Basis(m.atm, m.bas, m.env)
    db = mc.DeviceBasis(B)
    q, _ = screened_quartets(db.schwarz(m.nbas), m.nbas, 1e-12)
    plan = mc.Plan(db, "int2e_sph", q)
    rng = np.random.default_rng(5)
    picks = []
    for i, b in enumerate(plan.buckets()):
        if b["path"] == "chunked" and mc.wave_kernel_exists(b["cls"]):
            ids = plan.bucket_tasks(i, 0, int(b["tasks"])).astype(np.int64)
            picks.append(np.sort(rng.choice(ids, size=min(20000, ids.size), replace=False)))
    plan.close()
    db.close()
    sub = q[np.unique(np.concatenate(picks))]
    for omega in (0.0, 0.4):
        env = m.env.copy()
        env[8] = omega
        dbo = mc.DeviceBasis(mc.Basis(m.atm, m.bas, env))
        pl = mc.Plan(dbo, "int2e_sph", sub)
        outs = []
        for wave in (2, 0):
            old = mc.tune(mc.TUNE_WAVE_KERNELS, wave)
            try:
                o = torch.full((pl.out_size,), float("nan"), dtype=torch.float64, device="cuda")
                pl.execute(o.data_ptr(), graph=False)
                torch.cuda.synchronize()
                outs.append(o.cpu().numpy())
            finally:
                mc.tune(mc.TUNE_WAVE_KERNELS, old)
        pl.close()
        dbo.close()
        _say("wave vs chunk kernels, omega", omega, ":", len(sub), "quartets,", outs[0].size, "ERIs")
        assert np.isfinite(outs[0]).all() and np.isfinite(outs[1]).all()
        err = np.abs(outs[0] - outs[1])
        _say("max |difference|", float(err.max()), "max relative",
             float((err / np.maximum(np.abs(outs[1]), 1e-300)).max()))
        assert (err <= 1e-12 + 1e-10 * np.abs(outs[1])).all(), float(err.max())


def test_fused_jk_through_wave_kernels_c1(mc):
    """J/K of c1 quartets from every chunked bucket (up to 5,000 each, every
    team size) with the digest fused into the kernels, the wave kernel taken
    wherever it exists, equals the stored-buffer digest of the same list."""
    from libcint_amd import jk
    from libcint_amd.synthetic import make_molecule, random_density, screened_quartets
    m = make_molecule("c1")
    B = mc.Basis(m.atm, m.bas, m.env)
    db = mc.DeviceBasis(B)
    q, _ = screened_quartets(db.schwarz(m.nbas), m.nbas, 1e-12)
    plan = mc.Plan(db, "int2e_sph", q)
    rng = np.random.default_rng(9)
    picks = []
    for i, b in enumerate(plan.buckets()):
        if b["path"] == "chunked":
            ids = plan.bucket_tasks(i, 0, int(b["tasks"])).astype(np.int64)
            picks.append(rng.choice(ids, size=min(5000, ids.size), replace=False))
    plan.close()
    db.close()
    sub = q[np.unique(np.concatenate(picks))]
    dm = random_density(m.nao_sph(), seed=21)
    old = mc.tune(mc.TUNE_WAVE_KERNELS, 2)
    try:
        vj_f, vk_f = jk.get_jk(B, dm, quartets=sub)
    finally:
        mc.tune(mc.TUNE_WAVE_KERNELS, old)
    vj_s, vk_s = jk.get_jk(B, dm, quartets=sub, fused=False)
    dj = float((vj_f - vj_s).abs().max())
    dk = float((vk_f - vk_s).abs().max())
    _say("c1 fused J/K through the wave kernels vs stored digest:", len(sub), "quartets,",
         f"max|dJ| {dj:.2e} max|dK| {dk:.2e}")
    assert dj < 1e-12 * max(1.0, float(vj_s.abs().max()))
    assert dk < 1e-12 * max(1.0, float(vk_s.abs().max()))


def test_stratified_c3_ip1_and_cart(mc, orc):
    from libcint_amd.jk import gradient_quartets
    from libcint_amd.synthetic import make_molecule, unique_quartets
    m = make_molecule("c3")
    _check_stratified(mc, orc, m, "int2e_ip1_sph", gradient_quartets(m.nbas), seed=3)
    _check_stratified(mc, orc, m, "int2e_cart", unique_quartets(m.nbas), seed=4)


def test_stratified_c4_int3c2e(mc, orc):
    from libcint_amd.synthetic import make_molecule
    m = make_molecule("c4")
    ns = m.aux_start
    ii, jj = np.tril_indices(ns)
    aux = np.arange(ns, m.nbas, dtype=np.int32)
    q = np.zeros((ii.size * aux.size, 4), np.int32)
    q[:, 0] = np.repeat(ii, aux.size)
    q[:, 1] = np.repeat(jj, aux.size)
    q[:, 2] = np.tile(aux, ii.size)
    _check_stratified(mc, orc, m, "int3c2e_sph", q, seed=5)


def test_c2_mixed_jk_30_atoms(mc):
    """Mixed-precision J/K on the 30-atom C10N3O2H15/TZ cluster (every unique
    quartet with Schwarz bound >= 1e-12): quartets with Schwarz bound x
    max|D| < 1e-7 run recurrences and contraction in FP32.  Errors against
    the all-FP64 build (whose blocks the stratified c1 test pins to the
    oracle): max|dJ|, max|dK| <= 1e-8, energy error <= 1e-9 Ha."""
    import torch
    from libcint_amd import jk
    from libcint_amd.synthetic import make_molecule, random_density, screened_quartets
    m = make_molecule("c2")
    B = mc.Basis(m.atm, m.bas, m.env)
    db = mc.DeviceBasis(B)
    q, _ = screened_quartets(db.schwarz(m.nbas), m.nbas, 1e-12)
    db.close()
    dm = random_density(m.nao_sph(), seed=11)
    dmt = torch.as_tensor(dm, device="cuda")
    res = {}
    for tier, thr in (("fp64", 0.0), ("mixed", 1e-7)):
        jb = jk.JKBuilder(B, q, fp32_threshold=thr, dmax=float(np.abs(dm).max()))
        st = jb.stats()
        vj, vk = jb.run(dm)
        torch.cuda.synchronize()
        res[tier] = (vj.cpu().numpy(), vk.cpu().numpy(), float(jk.energy(dmt, vj, vk)),
                     st["fp32_tasks"], st["tasks"])
        jb.close()
        del vj, vk, jb
        torch.cuda.empty_cache()
    j64, k64, e64, _, ntask = res["fp64"]
    j32, k32, e32, n32, _ = res["mixed"]
    dj = float(np.abs(j32 - j64).max())
    dk = float(np.abs(k32 - k64).max())
    _say(f"c2: {n32} of {ntask} quartets FP32; max|dJ| {dj:.2e} max|dK| {dk:.2e} "
         f"dE {e32 - e64:.2e} (E {e64:.6f})")
    assert n32 > 0
    assert dj <= 1e-8 and dk <= 1e-8
    assert abs(e32 - e64) <= 1e-9


def _c3_gradient_inputs():
    from libcint_amd.synthetic import make_molecule, random_density
    with open(os.path.join(ROOT, "tests", "golden", "c3_gradient.json")) as fh:
        gold = json.load(fh)
    m = make_molecule("c3")
    assert gold["nao"] == m.nao_sph() and gold["natm"] == m.natm
    dm = random_density(m.nao_sph(), seed=5)
    return m, dm, np.array(gold["grad"])


def test_c3_gradient_vs_oracle_full_cluster(mc):
    """GPU gradient digest of the 20-atom cluster against the oracle's
    int2e_ip1_sph blocks of every (i, j, k >= l) contracted with D on the
    CPU (tests/golden/c3_gradient.json, tests/golden/make_c3_gradient.py)."""
    from libcint_amd import jk
    m, dm, ref = _c3_gradient_inputs()
    g = jk.nuclear_gradient(mc.Basis(m.atm, m.bas, m.env), dm, m.natm).cpu().numpy()
    err = float(np.abs(g - ref).max())
    _say("c3 gradient vs oracle: max |dg|", err, "max |g|", float(np.abs(ref).max()))
    assert err <= 1e-10 * max(1.0, float(np.abs(ref).max()))
    assert np.abs(g.sum(axis=0)).max() < 1e-10


@pytest.mark.timeout(900)
def test_c3_gradient_finite_difference_all_coordinates(mc):
    """Central differences (h = 1e-4 Bohr) of the GPU energy
    E = 1/2 D J - 1/4 D K on all 60 coordinates of the 20-atom cluster
    against the analytic gradient digest: within 1e-7 Ha/Bohr."""
    import torch
    from libcint_amd import jk
    from libcint_amd.synthetic import unique_quartets
    m, dm, _ = _c3_gradient_inputs()
    g = jk.nuclear_gradient(mc.Basis(m.atm, m.bas, m.env), dm, m.natm).cpu().numpy()
    q = unique_quartets(m.nbas)
    dmt = torch.as_tensor(dm, device="cuda")
    h = 1e-4

    def e_at(coords):
        mm = m.with_coords(coords)
        jb = jk.JKBuilder(mc.Basis(mm.atm, mm.bas, mm.env), q)
        try:
            vj, vk = jb.run(dm)
            return float(jk.energy(dmt, vj, vk))
        finally:
            jb.close()

    fd = np.zeros_like(g)
    for a in range(m.natm):
        for x in range(3):
            cp = m.coords.copy()
            cm = m.coords.copy()
            cp[a, x] += h
            cm[a, x] -= h
            fd[a, x] = (e_at(cp) - e_at(cm)) / (2 * h)
        _say("fd atom", a, "max |fd - g|", float(np.abs(fd[a] - g[a]).max()))
    err = float(np.abs(fd - g).max())
    _say("c3 FD, 60 coordinates: max |fd - g|", err)
    assert err <= 1e-7
