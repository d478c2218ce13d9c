"""mcint_plan_execute_noalloc (per-task blocks in the library's fixed device
array, no device allocation) against mcint_plan_execute_direct (the plan's
own scratch) through the C ABI: two calls in a row, the second on a smaller
list, element for element, plus the not-all-zero flags.  The lists hold
single-kernel, chunked and generic-kernel classes, and the first one a
chunked bucket too large for the fixed array at once (run in slices)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[no-alloc]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mc():
    import libcint_amd
    libcint_amd.lib()
    return libcint_amd


def _random_quartets(nbas, n, seed):
    """n random quartets with i >= j, k >= l."""
    q = np.random.default_rng(seed).integers(0, nbas, size=(n, 4))
    q[:, [0, 1]] = np.sort(q[:, [0, 1]], axis=1)[:, ::-1]
    q[:, [2, 3]] = np.sort(q[:, [2, 3]], axis=1)[:, ::-1]
    return q.astype(np.int32)


def _f_quartets(m, n, seed):
    """n quartets (i >= j, k >= l) of f shells of the molecule (the (ff|ff) chunked
    class, 5476 doubles of block per task)."""
    f = np.flatnonzero(m.bas[:m.nbas, 1] == 3)
    rng = np.random.default_rng(seed)
    q = rng.choice(f, size=(4 * n, 4))
    q[:, [0, 1]] = np.sort(q[:, [0, 1]], axis=1)[:, ::-1]
    q[:, [2, 3]] = np.sort(q[:, [2, 3]], axis=1)[:, ::-1]
    return q[:n].astype(np.int32)


def _run_both(mc, db, name, q, torch):
    plan = mc.Plan(db, name, q)
    dev = torch.device("cuda:0")
    ref = torch.full((plan.out_size,), float("nan"), dtype=torch.float64, device=dev)
    nz_ref = torch.full((len(q),), -1, dtype=torch.int32, device=dev)
    plan.execute(ref.data_ptr(), nz_ref.data_ptr(), graph=False)
    got = torch.full((plan.out_size,), float("nan"), dtype=torch.float64, device=dev)
    nz = torch.full((len(q),), -1, dtype=torch.int32, device=dev)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    plan.execute_noalloc(got.data_ptr(), nz.data_ptr(), stream_ptr=s.cuda_stream)
    s.synchronize()
    torch.cuda.synchronize()
    buckets = plan.buckets()
    plan.close()
    assert not torch.isnan(ref).any()
    assert torch.equal(got, ref), f"{name}: max |diff| {float((got - ref).abs().max()):.3e}"
    assert torch.equal(nz, nz_ref)
    return buckets


def test_noalloc_matches_direct_two_calls(mc):
    import torch
    from libcint_amd.synthetic import make_molecule

    assert mc.lib().mcint_noalloc_scratch_doubles() == 8 << 20
    m = make_molecule("c1")
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    # call 1: a spread of classes (single-kernel and chunked) plus 3000 (ff|ff)
    # quartets: 3000 x 5476 doubles exceed the 7M-double block area
    q1 = np.concatenate([_random_quartets(m.nbas, 20000, 11), _f_quartets(m, 3000, 1)])
    _say("call 1:", len(q1), "quartets (int2e_sph)")
    b1 = _run_both(mc, db, "int2e_sph", q1, torch)
    paths = {b["path"] for b in b1}
    assert "chunked" in paths and "fast" in paths, paths
    # call 2: a smaller list; the f-shell derivative classes run on the
    # generic kernel (per-task blocks in the same array)
    q2 = np.concatenate([_random_quartets(m.nbas, 3000, 12), _f_quartets(m, 200, 2)])
    _say("call 2:", len(q2), "quartets (int2e_ip1_sph)")
    b2 = _run_both(mc, db, "int2e_ip1_sph", q2, torch)
    assert "generic" in {b["path"] for b in b2}
    db.close()


def test_noalloc_rejects_null_output(mc):
    import torch
    from libcint_amd.synthetic import make_molecule, unique_quartets
    m = make_molecule("c0")
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    plan = mc.Plan(db, "int2e_sph", unique_quartets(m.nbas)[:10])
    with pytest.raises(mc.McintError):
        plan.execute_noalloc(0)
    plan.close()
    db.close()
    torch.cuda.synchronize()
