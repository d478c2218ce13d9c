#!/usr/bin/env python3
"""Benchmark: batched plain FP64 int2e_sph on one MI355X.

Workload (default): the 30-atom C10N3O2H15 cluster with the TZ-like basis
(l <= 3, contracted s/p shells of 8-10 primitives; libcint_amd.synthetic
'c1'), all unique shell quartets (i>=j, k>=l, ij>=kl) whose Schwarz bound
sqrt((ij|ij)) sqrt((kl|kl)) is at least 1e-12.  The quartets are sorted into
class buckets and run from a captured HIP graph.  One step computes every
ERI block of that list into a device buffer (inputs and outputs resident in
HBM).  Reported value: spherical ERIs per second (sum of di dj dk dl over the
computed quartets / time), the definition of examples/time_c60.c:188,230.

Also reported: the dominant bucket's algorithmic FLOP rate against MI355X's
FP64 peak (HIP events on the plan's stream), the CPU oracle (this
repository's C restatement of libcint, OpenMP over quartets) on a bounded
random sample of the same quartets, ("gradient") the J/K build and the
ip1 nuclear-gradient digest on the 20-atom C8H12 cluster timed against the
plain FP64 int2e_sph pass over the same molecule, and ("three_centre") the
int3c2e batch of the 40-atom C16H24 cluster with its aux basis, FP64, FP32
tier and short range.

    python bench.py [--steps K] [--warmup W] [--config c1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "shell-quartet ERIs/sec (int2e_sph) + ip1 gradient time vs plain FP64, err≤1e-10"
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_quartets(db, nshell, thresh):
    from libcint_amd.synthetic import screened_quartets
    return screened_quartets(db.schwarz(nshell), nshell, thresh)


def _cpu_info() -> dict:
    """CPU model, logical CPUs, SMT and the CPUs this process may run on."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        with open("/proc/cpuinfo") as fh:
            txt = fh.read()
        for line in txt.splitlines():
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
        sib = [int(x.split(":")[1]) for x in txt.splitlines() if x.startswith("siblings")]
        cor = [int(x.split(":")[1]) for x in txt.splitlines() if x.startswith("cpu cores")]
        if sib and cor:
            info["smt_threads_per_core"] = sib[0] // max(1, cor[0])
    except OSError:
        pass
    return info


def cpu_baseline(basis_np, quartets, target_s=12.0, max_threads=16):
    """Time the CPU oracle (the restatement of libcint's path, tabulated
    roots, OpenMP dynamic schedule over quartets) on a random sample of the
    benchmark's list: with every CPU this process may use (at most 16, one
    GPU's share of the host) and on one core."""
    import oracle
    oracle.build()
    oracle.set_root_mode(oracle.ROOTS_TABLE)
    B = oracle.Basis(basis_np.atm, basis_np.bas, basis_np.env)
    info = _cpu_info()
    nthreads = max(1, min(max_threads, info["affinity"]))

    def sample_rate(threads, budget):
        rng = np.random.default_rng(123)
        n = 256
        while True:
            sel = quartets[rng.choice(quartets.shape[0], size=min(n, quartets.shape[0]),
                                      replace=False)]
            t0 = time.perf_counter()
            _, off, _ = oracle.batch("int2e_sph", sel, B, nthreads=threads)
            dt = time.perf_counter() - t0
            if dt >= budget * 0.5 or n >= quartets.shape[0]:
                return int(off[-1]) / dt, sel.shape[0], int(off[-1]), dt
            n = int(n * min(8.0, max(2.0, budget / max(dt, 1e-3))))

    rate, nq, eris, dt = sample_rate(nthreads, target_s)
    rate1, nq1, eris1, dt1 = sample_rate(1, target_s / 2)
    return dict(value=rate, unit="ERIs/s", cores=nthreads, kind="port",
                single_core_value=rate1, **info,
                sample=f"{nq} random quartets of the same list ({eris} ERIs) in {dt:.1f}s on "
                       f"{nthreads} threads; single core: {nq1} quartets ({eris1} ERIs) in "
                       f"{dt1:.1f}s; oracle/cint_oracle.c (libcint's algorithm restated in C) "
                       f"with the GPU's tabulated Rys roots, OpenMP dynamic schedule")


def sampled_check(db, B, plan, quart, out, per_bucket=2, seed=7) -> dict:
    """Blocks of per_bucket tasks at random positions of every bucket, read
    from the timed run's output, against the same quartets recomputed by the
    generic kernel (a separate implementation: runtime class, recurrences in
    LDS; the parity tests pin both to the oracle)."""
    import torch
    import libcint_amd as mc
    rng = np.random.default_rng(seed)
    picks = []
    for i, b in enumerate(plan.buckets()):
        for p0 in rng.choice(int(b["tasks"]), size=min(per_bucket, int(b["tasks"])), replace=False):
            picks.append(int(plan.bucket_tasks(i, int(p0), 1)[0]))
    idx = np.unique(np.array(picks, dtype=np.int64))
    sizes = B.block_sizes("int2e_sph", quart)
    off = np.zeros(sizes.size + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    pos = np.concatenate([np.arange(off[q], off[q + 1]) for q in idx])
    got = out[torch.as_tensor(pos, device=out.device)].cpu().numpy()
    old = mc.tune(mc.TUNE_GENERIC_ONLY, 1)
    try:
        gp = mc.Plan(db, "int2e_sph", quart[idx])
    finally:
        mc.tune(mc.TUNE_GENERIC_ONLY, old)
    ref_t = torch.empty(gp.out_size, dtype=torch.float64, device=out.device)
    gp.execute(ref_t.data_ptr(), graph=False)
    torch.cuda.synchronize()
    ref = ref_t.cpu().numpy()
    gp.close()
    err = np.abs(got - ref)
    return {"quartets": int(idx.size), "buckets": len(plan.buckets()),
            "reference": "generic kernel on the same quartets",
            "max_abs_err": float(err.max()),
            "max_rel_err": float((err / np.maximum(np.abs(ref), 1e-300)).max()),
            "within_1e-12_abs_1e-10_rel": bool(np.all(err <= 1e-12 + 1e-10 * np.abs(ref)))}


def gradient_section(reps: int = 3) -> dict:
    """The derivative and J/K legs on the 20-atom C8H12 cluster (DZ basis
    with d on H): plain FP64 int2e_sph and int2e_cart over all unique
    quartets, the J/K
    build (ERIs + grouped digest, FP64 and with the FP32 tier), and the
    nuclear-gradient digest over the ip1 list (i, j, k >= l), FP64 and with
    the FP32 tier (Schwarz x max|D|^2 < 1e-7), each timed on device-resident
    data after one untimed run."""
    import torch
    import libcint_amd as mc
    from libcint_amd import jk
    from libcint_amd.synthetic import make_molecule, random_density, unique_quartets

    m = make_molecule("c3")
    B = mc.Basis(m.atm, m.bas, m.env)
    dm = random_density(m.nao_sph(), seed=5)

    def timed(fn, reps_=None):
        n = reps if reps_ is None else reps_
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n, r

    db = mc.DeviceBasis(B)
    q = unique_quartets(m.nbas)
    plan = mc.Plan(db, "int2e_sph", q)
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    t_plain, _ = timed(lambda: plan.execute(out.data_ptr(), graph=False))
    plan.close()
    del out
    plan = mc.Plan(db, "int2e_cart", q)
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    t_cart, _ = timed(lambda: plan.execute(out.data_ptr(), graph=False))
    cart_eris = plan.out_size
    plan.close()
    del out
    jb = jk.JKBuilder(B)
    t_jk, (vj, vk) = timed(lambda: jb.run(dm))
    dmt = torch.as_tensor(dm, device="cuda")
    e64 = float(jk.energy(dmt, vj, vk))
    jb.close()
    jbs = jk.JKBuilder(B, fused=False)
    t_jk_stored, (vjs, vks) = timed(lambda: jbs.run(dm))
    fused_vs_stored = max(float((vjs - vj).abs().max()), float((vks - vk).abs().max()))
    jbs.close()
    del vjs, vks
    jb32 = jk.JKBuilder(B, fp32_threshold=1e-7, dmax=float(np.abs(dm).max()))
    fp32_tasks = jb32.stats()["fp32_tasks"]
    t_jk32, (vj32, vk32) = timed(lambda: jb32.run(dm))
    e32 = float(jk.energy(dmt, vj32, vk32))
    dj = float((vj32 - vj).abs().max())
    dk = float((vk32 - vk).abs().max())
    jb32.close()
    t0 = time.perf_counter()
    gp = jk.GradientPlan(B, m.natm)
    t_gplan = time.perf_counter() - t0
    t_grad, g = timed(lambda: gp.run(dm))
    gsum = float(g.sum(0).abs().max())
    nq_grad = gp.nquartets
    gp.close()
    gp32 = jk.GradientPlan(B, m.natm, fp32_threshold=1e-7, dmax=float(np.abs(dm).max()))
    g32_tasks = gp32.stats()["fp32_tasks"]
    t_grad32, g32 = timed(lambda: gp32.run(dm))
    g32_err = float((g32 - g).abs().max())
    gp32.close()
    # the plain FP64 GPU path for the same work: every quartet on the
    # generic kernel (one workgroup per quartet, runtime class, no register
    # kernels, no FP32 tier); one run each
    old = mc.tune(mc.TUNE_GENERIC_ONLY, 1)
    try:
        plan = mc.Plan(db, "int2e_sph", q)
        out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
        t_plain_gen, _ = timed(lambda: plan.execute(out.data_ptr(), graph=False), reps_=1)
        plan.close()
        del out
        jbg = jk.JKBuilder(B)
        t_jk_gen, (vjg, vkg) = timed(lambda: jbg.run(dm), reps_=1)
        jbg.close()
        gpg = jk.GradientPlan(B, m.natm)
        t_grad_gen, gg = timed(lambda: gpg.run(dm), reps_=1)
        gpg.close()
    finally:
        mc.tune(mc.TUNE_GENERIC_ONLY, old)
    gen_err = max(float((vjg - vj).abs().max()), float((vkg - vk).abs().max()),
                  float((gg - g).abs().max()))
    db.close()
    return {
        "workload": f"c3: {m.name}, {m.natm} atoms, {m.nbas} shells, {m.nao_sph()} AOs; "
                    f"random density D = A A^T / N",
        "plain_int2e_sph_ms": t_plain * 1e3,
        "plain_int2e_cart_ms": t_cart * 1e3,
        "plain_int2e_cart_eris": int(cart_eris),
        "plain_unique_quartets": int(q.shape[0]),
        "ip1_gradient_ms": t_grad * 1e3,
        "ip1_gradient_quartets": int(nq_grad),
        "ip1_gradient_vs_plain": t_grad / t_plain,
        "gradient_plan_create_s": t_gplan,
        "ip1_gradient_mixed_ms": t_grad32 * 1e3,
        "ip1_gradient_mixed_fp32_quartets": int(g32_tasks),
        "ip1_gradient_mixed_max_abs_error": g32_err,
        "gradient_translational_residual": gsum,
        "jk_fp64_ms": t_jk * 1e3,
        "jk_fp64_stored_buffer_ms": t_jk_stored * 1e3,
        "jk_fused_vs_stored_max_abs_diff": fused_vs_stored,
        "jk_mixed_ms": t_jk32 * 1e3,
        "jk_mixed_fp32_quartets": int(fp32_tasks),
        "jk_mixed_max_abs_dJ": dj,
        "jk_mixed_max_abs_dK": dk,
        "jk_mixed_energy_error": e32 - e64,
        # north-star figures on this molecule, all for the same work: one
        # int2e_sph pass over the unique quartets + one J/K build + one
        # nuclear-gradient digest
        "north_star": {
            "batched_mixed_ms": (t_plain + t_jk32 + t_grad32) * 1e3,
            "batched_fp64_ms": (t_plain + t_jk + t_grad) * 1e3,
            "plain_fp64_generic_ms": (t_plain_gen + t_jk_gen + t_grad_gen) * 1e3,
            "speedup_batched_mixed_vs_plain_fp64": (t_plain_gen + t_jk_gen + t_grad_gen)
                                                   / (t_plain + t_jk32 + t_grad32),
            "speedup_mixed_vs_fp64_batched": (t_plain + t_jk + t_grad) / (t_plain + t_jk32 + t_grad32),
            "plain_fp64_generic_max_abs_diff_jk_grad": gen_err,
            "definition": "plain FP64 GPU path = every quartet on the generic kernel (one "
                          "workgroup per quartet, runtime class, FP64); batched mixed = class "
                          "buckets on the register kernels in concurrent graph branches with "
                          "the FP32 tier for Schwarz x max|D|^2 < 1e-7",
        },
    }


def fused_jk_section(plan, m, reps: int, t_step: float) -> dict:
    """J and K of the timed workload (the same plan over the same unique
    quartets) with the density contraction inside the integral kernels: no
    ERI buffer; against the time of the ERI step that stores every block."""
    import torch
    from libcint_amd.synthetic import random_density
    nao = m.nao_sph()
    dm = torch.as_tensor(random_density(nao, seed=13), device="cuda")
    vj = torch.empty((nao, nao), dtype=torch.float64, device="cuda")
    vk = torch.empty_like(vj)
    plan.execute_jk(dm.data_ptr(), nao, vj.data_ptr(), vk.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute_jk(dm.data_ptr(), nao, vj.data_ptr(), vk.data_ptr())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    e = float(0.5 * (dm * vj).sum() - 0.25 * (dm * vk).sum())
    return {"workload": f"{m.name}: J and K for a random density D = A A^T / N, {nao} AOs",
            "ms": dt * 1e3, "eri_step_ms": t_step * 1e3, "vs_eri_step": dt / t_step,
            "eri_buffer_bytes": 0, "energy": e,
            "device_memory_beyond_plan_bytes": 2 * nao * nao * 8}


def c2_section(reps: int = 2) -> dict:
    """Mixed-precision J and K on the 30-atom C10N3O2H15/TZ cluster (every
    unique quartet with Schwarz bound >= 1e-12): quartets whose Schwarz
    bound x max|D| is below 1e-7 run their recurrences and contraction in
    FP32; roots, weights and the J/K accumulation stay FP64.  Errors are
    stated against the all-FP64 build on the same list; each build (ERIs +
    grouped digest) is timed on device-resident data after one untimed run."""
    import torch
    import libcint_amd as mc
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
        t0 = time.perf_counter()
        for _ in range(reps):
            vj, vk = jb.run(dm)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[tier] = (vj.cpu().numpy(), vk.cpu().numpy(), float(jk.energy(dmt, vj, vk)), dt,
                     int(st["fp32_tasks"]))
        jb.close()
        del vj, vk, jb
        torch.cuda.empty_cache()
    j64, k64, e64, t64, _ = res["fp64"]
    j32, k32, e32, t32, n32 = res["mixed"]
    return {
        "workload": f"c2: {m.name}, {m.natm} atoms, {m.nbas} shells, {m.nao_sph()} AOs, "
                    f"{q.shape[0]} quartets (Schwarz >= 1e-12); D = A A^T / N, seed 11",
        "jk_fp64_ms": t64 * 1e3,
        "jk_mixed_ms": t32 * 1e3,
        "fp32_quartets": n32,
        "fp32_threshold": 1e-7,
        "max_abs_dJ": float(np.abs(j32 - j64).max()),
        "max_abs_dK": float(np.abs(k32 - k64).max()),
        "energy_fp64": e64,
        "energy_error": e32 - e64,
    }


def three_centre_section(reps: int = 3) -> dict:
    """int3c2e_sph over every (i >= j | P) triple of the 40-atom C16H24
    cluster (TZ-like orbital basis, aux shells up to g): FP64, the FP32 tier
    for triples whose orbital pair has sqrt((ij|ij)) < 1e-4 (error against
    FP64 stated), and the short-range erfc variant with omega = 0.3."""
    import torch
    import libcint_amd as mc
    from libcint_amd.synthetic import make_molecule

    m = make_molecule("c4")
    ns = m.aux_start
    ii, jj = np.tril_indices(ns)
    aux = np.arange(ns, m.nbas, dtype=np.int32)
    q = np.zeros((ii.size * aux.size, 4), np.int32)
    q[:, 0] = np.repeat(ii, aux.size)
    q[:, 1] = np.repeat(jj, aux.size)
    q[:, 2] = np.tile(aux, ii.size)

    def run(env, flags=None):
        db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, env))
        plan = mc.Plan(db, "int3c2e_sph", q, fp32_flags=flags)
        out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.execute(out.data_ptr())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        st = plan.stats()
        plan.close()
        db.close()
        return dt, out, st

    t64, o64, st = run(m.env)
    db = mc.DeviceBasis(mc.Basis(m.atm, m.bas, m.env))
    qs = db.schwarz(ns)
    db.close()
    flags = (qs[q[:, 0], q[:, 1]] < 1e-4).astype(np.uint8)
    t32, o32, st32 = run(m.env, flags)
    err32 = float((o32 - o64).abs().max())
    del o32
    env_sr = m.env.copy()
    env_sr[8] = -0.3
    tsr, osr, _ = run(env_sr)
    del osr
    return {
        "workload": f"c4: {m.name}, {ns} orbital + {m.nbas - ns} aux shells, every (i>=j|P)",
        "triples": int(q.shape[0]),
        "eris": int(st["out_size"]),
        "fp64_ms": t64 * 1e3,
        "fp64_eris_per_s": st["out_size"] / t64,
        "fp32_tier_ms": t32 * 1e3,
        "fp32_tier_triples": int(st32["fp32_tasks"]),
        "fp32_tier_max_abs_error": err32,
        "fp64_max_abs_value": float(o64.abs().max()),
        "erfc_omega_0.3_ms": tsr * 1e3,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--schwarz", type=float, default=1e-12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the J/K, nuclear-gradient and three-centre legs")
    ap.add_argument("--bucket-table", default="",
                    help="write every bucket's class, path, tasks, time and FLOP rate (JSON)")
    ap.add_argument("--only-buckets", default="",
                    help="run only the buckets of these classes, e.g. '3221:1,3321:1' "
                         "(la lb lc ld : team), --steps times each, then exit (for counter runs)")
    ap.add_argument("--wave-kernels", type=int, default=1, choices=(0, 1, 2),
                    help="chunked buckets on the wave kernel: 0 never, 1 where measured "
                         "faster (default: the tuned table), 2 every bucket that has one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); a multi-GPU run "
              "would split the quartet list by shell pair, one shard per GPU, and all-reduce "
              "J, K and the gradient", file=sys.stderr)
        sys.exit(2)

    import torch
    import libcint_amd as mc
    from libcint_amd.flops import bucket_flops
    from libcint_amd.synthetic import make_molecule

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    mc.tune(mc.TUNE_WAVE_KERNELS, args.wave_kernels)
    m = make_molecule(args.config)
    nshell = m.nbas if m.aux_start < 0 else m.aux_start
    B = mc.Basis(m.atm, m.bas, m.env)
    db = mc.DeviceBasis(B)
    t0 = time.perf_counter()
    quart, nuniq = build_quartets(db, nshell, args.schwarz)
    log(f"{m.name}: {nshell} shells, {nuniq} unique quartets, {quart.shape[0]} pass Schwarz "
        f">= {args.schwarz:g} ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    plan = mc.Plan(db, "int2e_sph", quart)
    st = plan.stats()
    log(f"plan: {st} ({time.perf_counter() - t0:.1f}s)")
    eris = plan.out_size
    out = torch.empty(eris, dtype=torch.float64, device=dev)
    log(f"output {eris * 8 / 1e9:.1f} GB")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    if args.only_buckets:
        want = {(tuple(int(c) for c in k.split(":")[0]), int(k.split(":")[1]))
                for k in args.only_buckets.split(",")}
        bl = plan.buckets()
        nprim = plan.count_primitives()
        for i, b in enumerate(bl):
            if (tuple(b["cls"]), b["team"]) not in want:
                continue
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            plan.execute_bucket(i, out.data_ptr(), stream_ptr=sp)
            e0.record(stream)
            for _ in range(args.steps):
                plan.execute_bucket(i, out.data_ptr(), stream_ptr=sp)
            e1.record(stream)
            e1.synchronize()
            t = e0.elapsed_time(e1) * 1e-3 / args.steps
            fl = bucket_flops(b["cls"], nprim[i], b["tasks"])
            print(json.dumps(dict(bucket=i, cls=list(b["cls"]), team=b["team"], path=b["path"],
                                  tasks=b["tasks"], prim_quartets=int(nprim[i]), ms=t * 1e3,
                                  model_flops=fl, frac=fl / t / 1e12 / FP64_PEAK_TFLOPS)), flush=True)
        return

    for _ in range(args.warmup):
        plan.execute(out.data_ptr(), stream_ptr=sp)
        stream.synchronize()
    log("warmup done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        plan.execute(out.data_ptr(), stream_ptr=sp)
    stream.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    log(f"timed {args.steps} steps: {ms:.1f} ms/step")

    # per-bucket times (HIP events on the plan's stream) and the model FLOPs
    buckets = plan.buckets()
    nprim = plan.count_primitives()
    btime = []
    for i in range(len(buckets)):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.execute_bucket(i, out.data_ptr(), stream_ptr=sp)
        e1.record(stream)
        e1.synchronize()
        btime.append(e0.elapsed_time(e1) * 1e-3)
    dom = int(np.argmax(btime))
    reps = max(3, args.steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.execute_bucket(dom, out.data_ptr(), stream_ptr=sp)
    e1.record(stream)
    e1.synchronize()
    dom_t = e0.elapsed_time(e1) * 1e-3 / reps
    bflops = [bucket_flops(b["cls"], nprim[i], b["tasks"]) for i, b in enumerate(buckets)]
    dom_flops = bflops[dom]
    achieved = dom_flops / dom_t / 1e12
    total_flops = float(sum(bflops))
    # HBM traffic of the dominant launch (all kernels of the bucket), when a
    # committed counter profile (rocprofv3 FETCH_SIZE/WRITE_SIZE) matches this
    # bucket's class, path, team size and task count
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "dominant_traffic.json")) as fh:
            tr = json.load(fh)
        b = buckets[dom]
        if (tuple(tr["cls"]) == tuple(b["cls"]) and tr["path"] == b["path"]
                and tr["team"] == b["team"] and tr["tasks"] == b["tasks"]):
            traffic = tr["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        traffic = None
    if args.bucket_table:
        rows = [dict(bucket=i, cls=list(b["cls"]), team=b["team"], path=b["path"],
                     tasks=b["tasks"], prim_quartets=int(nprim[i]), ms=btime[i] * 1e3,
                     model_flops=bflops[i], tflops=bflops[i] / btime[i] / 1e12,
                     frac=bflops[i] / btime[i] / 1e12 / FP64_PEAK_TFLOPS)
                for i, b in enumerate(buckets)]
        with open(args.bucket_table, "w") as fh:
            json.dump({"config": args.config, "step_ms": ms, "buckets": rows}, fh)
    top = sorted(range(len(buckets)), key=lambda i: -btime[i])[:8]
    for i in top:
        b = buckets[i]
        log(f"bucket {i}: cls={b['cls']} team={b['team']} {b['path']} tasks={b['tasks']} "
            f"prims={nprim[i]} t={btime[i] * 1e3:.2f} ms  {bflops[i] / btime[i] / 1e12:.2f} TFLOP/s")

    result = {
        "metric": METRIC,
        "value": eris * args.steps / dt,
        "unit": "ERIs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: batched int2e_sph over the Schwarz-surviving unique shell "
                        f"quartets of {m.name}",
            "quartets": int(quart.shape[0]),
            "unique_quartets": nuniq,
            "schwarz_threshold": args.schwarz,
            "eris_per_step": int(eris),
            "buckets": len(buckets),
            "fast_tasks": st["fast_tasks"],
            "generic_tasks": st["generic_tasks"],
            "surviving_primitive_quartets": int(nprim.sum()),
        },
        "quartets_per_s": quart.shape[0] * args.steps / dt,
        "model_tflops_per_s": total_flops / (dt / args.steps) / 1e12,
        "roofline": {
            # vector FP64 FMA (no MFMA instruction is issued): bounded by the
            # FP64 VALU peak, 78.6 TFLOP/s on MI355X
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": f"class {buckets[dom]['cls']} team {buckets[dom]['team']} "
                      f"({buckets[dom]['path']}"
                      f"{': component-chunk kernels + LDS epilogue, timed together' if buckets[dom]['path'] == 'chunked' else ''}), "
                      f"{buckets[dom]['tasks']} quartets, {int(nprim[dom])} primitive quartets",
            "launch_ms": dom_t * 1e3,
            "flops_per_launch": dom_flops,
        },
    }
    log("sampled check of the timed output ...")
    result["sampled_check"] = sampled_check(db, B, plan, quart, out)
    log(f"sampled_check: {result['sampled_check']}")
    del out
    torch.cuda.empty_cache()
    if not args.no_gradient:
        log("J/K fused into the integral kernels on the same plan (no ERI buffer) ...")
        result["fused_jk"] = fused_jk_section(plan, m, args.steps, t_step=dt / args.steps)
        log(f"fused_jk: {result['fused_jk']}")
    plan.close()
    if not args.no_gradient:
        log("gradient / J/K legs ...")
        result["gradient"] = gradient_section()
        log(f"gradient: {result['gradient']}")
        log("c2 leg (mixed J/K, 30 atoms) ...")
        result["c2_mixed_jk"] = c2_section()
        log(f"c2_mixed_jk: {result['c2_mixed_jk']}")
        log("three-centre leg ...")
        result["three_centre"] = three_centre_section()
        log(f"three_centre: {result['three_centre']}")
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(B, quart, target_s=args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
