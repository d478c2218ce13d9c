"""Per-class autotuner: team size of the register kernels, and the kernel
of the team-of-one buckets.

A team of lanes shares one contracted quartet; its size follows from the
number of primitive quartets each lane should handle (``mcint_tune`` /
``mcint_tune_class``).  Few primitive quartets per lane spread a quartet
over many lanes (short loops, but the team reduction and the epilogue are
paid by every lane); many keep lanes busy longer per quartet.  The best
value depends on the class and on the contraction depth of the workload, so
it is measured: for each candidate the plan is rebuilt and every bucket is
timed alone with HIP events, the times are summed per class, and each class
keeps its fastest candidate.

Second dimension (plain int2e): a chunked class's bucket of each team size
runs either its component-chunk kernels or the wave kernel (16-64 lanes per
quartet, ``mcint_tune_class_wave``); both are timed per bucket and each
(class, team size) keeps the faster one (``"wave"`` in the table, keyed
"la,lb,lc,ld:team": 2 wave kernel, 1 chunk kernels).

    python -m libcint_amd.autotune out.json           # c1 / c3 / c4 workloads
    python -m libcint_amd.autotune --wave-only out.json   # kernel choice only,
                                                      # on top of the committed table

The committed table (``libcint_amd/tuning/mi355x.json``) was produced this
way on an MI355X and is applied when the library is loaded
(``libcint_amd.lib()``); ``apply(load(path))`` installs another one.
"""
from __future__ import annotations

import collections
import json
import os
import sys

import numpy as np

from . import (INTOR_CODES, TUNE_PRIMS_PER_LANE, TUNE_WAVE_KERNELS, DeviceBasis, Plan, tune,
               tune_class, tune_class_wave, wave_kernel_exists)

CANDIDATES = (8, 16, 32, 64, 128)
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x.json")


def bucket_times(plan: Plan, out_ptr: int, reps: int = 2) -> dict:
    """Seconds per (class, path), each bucket timed alone (best of reps)."""
    import torch
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    agg: dict = collections.defaultdict(float)
    for i, b in enumerate(plan.buckets()):
        best = float("inf")
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.execute_bucket(i, out_ptr, stream_ptr=sp)
            e1.record(stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        agg[(b["cls"], b["path"])] += best
    return agg


def autotune(dbasis: DeviceBasis, name: str, quartets, candidates=CANDIDATES, reps: int = 2,
             log=print) -> dict:
    """Best primitive quartets per lane for every register-kernel class of
    the workload (generic-path classes are left out)."""
    import torch
    reset(name)
    times: dict = {}
    out = None
    for ppl in candidates:
        tune(TUNE_PRIMS_PER_LANE, ppl)
        plan = Plan(dbasis, name, quartets)
        if out is None:
            out = torch.empty(max(plan.out_size, 1), dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr(), graph=False)
        torch.cuda.synchronize()
        for key, t in bucket_times(plan, out.data_ptr(), reps).items():
            times.setdefault(key, {})[ppl] = t
        plan.close()
        log(f"[autotune] {name} prims/lane {ppl}: {sum(v[ppl] for v in times.values() if ppl in v) * 1e3:.2f} ms")
    tune(TUNE_PRIMS_PER_LANE, 32)
    table = {}
    for (cls, path), by in sorted(times.items()):
        if path == "generic":
            continue
        best = min(by, key=by.get)
        table[",".join(str(c) for c in cls)] = best
    total_default = sum(by.get(32, 0.0) for by in times.values())
    total_best = sum(min(by.values()) for by in times.values())
    log(f"[autotune] {name}: sum of bucket times {total_default * 1e3:.2f} ms at 32 -> "
        f"{total_best * 1e3:.2f} ms tuned")
    return table


def autotune_wave(dbasis: DeviceBasis, quartets, reps: int = 2, margin: float = 0.03,
                  log=print) -> dict:
    """Kernel of every bucket of the chunked int2e_sph classes: the wave
    kernel where it is at least `margin` faster than the chunk kernels, else
    the chunk kernels ({"la,lb,lc,ld:team": 2 or 1})."""
    import torch
    plan = Plan(dbasis, "int2e_sph", quartets)
    out = torch.empty(max(plan.out_size, 1), dtype=torch.float64, device="cuda")
    one = [(i, b) for i, b in enumerate(plan.buckets())
           if b["path"] == "chunked" and wave_kernel_exists(b["cls"])]
    times = {}
    for mode in (1, 2):
        for _, b in one:
            tune_class_wave(b["cls"], mode, b["team"])
        plan.execute(out.data_ptr(), graph=False)
        torch.cuda.synchronize()
        for i, b in one:
            best = float("inf")
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                plan.execute_bucket(i, out.data_ptr())
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1) * 1e-3)
            times.setdefault((tuple(b["cls"]), b["team"]), {})[mode] = best
    plan.close()
    table = {}
    for (cls, team), by in sorted(times.items()):
        choice = 2 if by[2] < (1.0 - margin) * by[1] else 1
        tune_class_wave(cls, 0, team)
        table[",".join(str(c) for c in cls) + ":%d" % team] = choice
        log(f"[autotune] wave {cls} team {team}: chunk kernels {by[1] * 1e3:.3f} ms, wave kernel "
            f"{by[2] * 1e3:.3f} ms -> {'wave' if choice == 2 else 'chunk'}")
    return table


def reset(name: str) -> None:
    """Drop every per-class override of one integral kind."""
    for cls in np.ndindex(7, 7, 7, 7):
        tune_class(name, cls, 0)


def apply(tables: dict) -> None:
    """Install {intor name: {"la,lb,lc,ld": prims per lane}} overrides and,
    under "wave", {"la,lb,lc,ld": 2 wave kernel | 1 chunk kernels}."""
    for name, table in tables.items():
        if name == "wave":
            for key, mode in table.items():
                cls, _, team = key.partition(":")
                c = [int(v) for v in cls.split(",")]
                if int(mode) != 2 or wave_kernel_exists(c):
                    tune_class_wave(c, int(mode), int(team or 1))
            continue
        if name not in INTOR_CODES:
            continue
        for cls, ppl in table.items():
            tune_class(name, [int(c) for c in cls.split(",")], int(ppl))


def load(path: str = DEFAULT_TABLE) -> dict:
    with open(path) as fh:
        data = json.load(fh)
    return data.get("tables", {})


def _workloads():
    """The tuning workloads: c1 int2e_sph (Schwarz 1e-12), c3 int2e_ip1_sph
    (gradient list) and c4 int3c2e_sph (every (i>=j|P))."""
    from . import Basis
    from .jk import gradient_quartets
    from .synthetic import make_molecule, unique_quartets

    m1 = make_molecule("c1")
    db1 = DeviceBasis(Basis(m1.atm, m1.bas, m1.env))
    qs = db1.schwarz(m1.nbas)
    ii, jj = np.tril_indices(m1.nbas)
    qp = qs[ii, jj]
    p, r = np.tril_indices(ii.size)
    q1 = unique_quartets(m1.nbas)[qp[p] * qp[r] >= 1e-12]
    yield "int2e_sph", db1, q1
    m3 = make_molecule("c3")
    yield "int2e_ip1_sph", DeviceBasis(Basis(m3.atm, m3.bas, m3.env)), gradient_quartets(m3.nbas)
    m4 = make_molecule("c4")
    ns = m4.aux_start
    ii, jj = np.tril_indices(ns)
    aux = np.arange(ns, m4.nbas, dtype=np.int32)
    q4 = np.zeros((ii.size * aux.size, 4), np.int32)
    q4[:, 0] = np.repeat(ii, aux.size)
    q4[:, 1] = np.repeat(jj, aux.size)
    q4[:, 2] = np.tile(aux, ii.size)
    yield "int3c2e_sph", DeviceBasis(Basis(m4.atm, m4.bas, m4.env)), q4


def main(argv):
    wave_only = "--wave-only" in argv
    args = [a for a in argv[1:] if a != "--wave-only"]
    out_path = args[0] if args else DEFAULT_TABLE
    tables = load() if wave_only else {}
    tune(TUNE_WAVE_KERNELS, 1)
    for name, db, q in _workloads():
        if name == "int2e_sph":
            if not wave_only:
                print(f"[autotune] {name}: {q.shape[0]} quartets", flush=True)
                tables[name] = autotune(db, name, q, log=lambda s: print(s, flush=True))
            apply({name: tables.get(name, {})})
            tables["wave"] = autotune_wave(db, q, log=lambda s: print(s, flush=True))
        elif not wave_only:
            print(f"[autotune] {name}: {q.shape[0]} quartets", flush=True)
            tables[name] = autotune(db, name, q, log=lambda s: print(s, flush=True))
        db.close()
    with open(out_path, "w") as fh:
        json.dump({"device": "MI355X (gfx950)", "candidates": list(CANDIDATES),
                   "tables": tables}, fh, indent=1, sort_keys=True)
    print(f"[autotune] wrote {out_path}", flush=True)


if __name__ == "__main__":
    main(sys.argv)
