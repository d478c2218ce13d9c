"""libcint_amd: libcint's two-electron Rys integrals on an AMD MI355X.

Python mirror of the C ABI in include/mcint.h (ctypes).  The product is the
shared library libcint_amd/libmcint.so, built by ``python -m
libcint_amd.build``; importing this package never falls back to a CPU
implementation -- if the library is missing, ``lib()`` raises.

Names follow libcint: ``intor('int2e_sph', shls, basis)`` is one call of
libcint's ``int2e_sph`` (src/cint2e.c:1186-1193); ``batch`` is the drop-in for
a loop over such calls; ``Plan`` is the device-resident throughput path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcint.so")

INTOR_NAMES = ("int2e_sph", "int2e_cart", "int2e_ip1_sph", "int2e_ip1_cart", "int3c2e_sph",
               "int3c2e_cart", "int3c2e_ip1_sph", "int3c2e_ip1_cart", "int3c2e_ip2_sph",
               "int3c2e_ip2_cart", "int2c2e_sph", "int2c2e_cart", "int2e_ip2_sph", "int2e_ip2_cart",
               "int2e_ipip1_sph", "int2e_ipip1_cart", "int2e_ip1ip2_sph", "int2e_ip1ip2_cart")
INTOR_CODES = {n: i for i, n in enumerate(INTOR_NAMES)}
(INT2E_SPH, INT2E_CART, INT2E_IP1_SPH, INT2E_IP1_CART, INT3C2E_SPH, INT3C2E_CART, INT3C2E_IP1_SPH,
 INT3C2E_IP1_CART, INT3C2E_IP2_SPH, INT3C2E_IP2_CART, INT2C2E_SPH, INT2C2E_CART, INT2E_IP2_SPH,
 INT2E_IP2_CART) = range(14)


def intor_shape(name: str):
    """(slots of a 4-int shell tuple the integral uses, derivative components)."""
    ncomp = 9 if ("_ipip" in name or "_ip1ip2" in name) else (3 if "_ip" in name else 1)
    if name.startswith("int2c2e"):
        return (0, 1), ncomp
    if name.startswith("int3c2e"):
        return (0, 1, 2), ncomp
    return (0, 1, 2, 3), ncomp

_lib = None


class McintError(RuntimeError):
    pass


def lib():
    """Load libmcint.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise McintError(f"{LIB_PATH} is missing: run `python -m libcint_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname, libamdhip64.so.7).  Loading torch first makes libmcint bind to
    # that copy instead of pulling in a second runtime from /opt/rocm.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    for name in INTOR_NAMES:
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [dp, ip, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp, vp, dp]
    for name in INTOR_NAMES:
        f = getattr(L, name + "_batch")
        f.restype = ctypes.c_int
        f.argtypes = [dp, ctypes.POINTER(ctypes.c_size_t), ip, ctypes.c_size_t, ip, ctypes.c_int,
                      ip, ctypes.c_int, dp, vp, ip, vp]
    L.mcint_basis_create.restype = ctypes.c_int
    L.mcint_basis_create.argtypes = [ctypes.POINTER(vp), ip, ctypes.c_int, ip, ctypes.c_int, dp]
    L.mcint_basis_destroy.argtypes = [vp]
    L.mcint_plan_create.restype = ctypes.c_int
    L.mcint_plan_create.argtypes = [ctypes.POINTER(vp), vp, ctypes.c_int, ip, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.mcint_plan_create_mixed.restype = ctypes.c_int
    L.mcint_plan_create_mixed.argtypes = [ctypes.POINTER(vp), vp, ctypes.c_int, ip, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_int64)]
    L.mcint_plan_fp32_tasks.restype = ctypes.c_int64
    L.mcint_plan_fp32_tasks.argtypes = [vp]
    L.mcint_plan_execute.restype = ctypes.c_int
    L.mcint_plan_execute.argtypes = [vp, vp, vp, vp]
    L.mcint_plan_execute_direct.restype = ctypes.c_int
    L.mcint_plan_execute_direct.argtypes = [vp, vp, vp, vp]
    L.mcint_plan_execute_noalloc.restype = ctypes.c_int
    L.mcint_plan_execute_noalloc.argtypes = [vp, vp, vp, vp]
    L.mcint_noalloc_scratch_doubles.restype = ctypes.c_int64
    L.mcint_noalloc_scratch_doubles.argtypes = []
    L.mcint_plan_destroy.argtypes = [vp]
    L.mcint_plan_stats.restype = ctypes.c_int
    L.mcint_plan_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mcint_plan_nbuckets.restype = ctypes.c_int
    L.mcint_plan_nbuckets.argtypes = [vp]
    L.mcint_plan_bucket_info.restype = ctypes.c_int
    L.mcint_plan_bucket_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.mcint_plan_bucket_tasks.restype = ctypes.c_int
    L.mcint_plan_bucket_tasks.argtypes = [vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ip]
    L.mcint_plan_execute_bucket.restype = ctypes.c_int
    L.mcint_plan_execute_bucket.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    L.mcint_plan_count_primitives.restype = ctypes.c_int
    L.mcint_plan_count_primitives.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mcint_plan_execute_grad.restype = ctypes.c_int
    L.mcint_plan_execute_grad.argtypes = [vp, vp, ctypes.c_int, vp, vp]
    L.mcint_plan_execute_jk.restype = ctypes.c_int
    L.mcint_plan_execute_jk.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp]
    L.mcint_jk_digest.restype = ctypes.c_int
    L.mcint_jk_digest.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp, vp, vp, vp]
    L.mcint_jk_plan_create.restype = ctypes.c_int
    L.mcint_jk_plan_create.argtypes = [ctypes.POINTER(vp), vp, ip, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_int64)]
    L.mcint_jk_plan_execute.restype = ctypes.c_int
    L.mcint_jk_plan_execute.argtypes = [vp, vp, vp, vp, vp, vp]
    L.mcint_jk_plan_destroy.argtypes = [vp]
    L.mcint_schwarz.restype = ctypes.c_int
    L.mcint_schwarz.argtypes = [vp, ctypes.c_int, dp]
    L.mcint_tune.restype = ctypes.c_int
    L.mcint_tune.argtypes = [ctypes.c_int, ctypes.c_int]
    L.mcint_tune_class.restype = ctypes.c_int
    L.mcint_tune_class.argtypes = [ctypes.c_int] * 6
    L.mcint_tune_class_wave.restype = ctypes.c_int
    L.mcint_tune_class_wave.argtypes = [ctypes.c_int] * 6
    L.mcint_wave_kernel_exists.restype = ctypes.c_int
    L.mcint_wave_kernel_exists.argtypes = [ctypes.c_int] * 4
    L.mcint_last_error.restype = ctypes.c_char_p
    for name in ("int2e_optimizer", "int2e_ip1_optimizer", "int3c2e_optimizer",
                 "int3c2e_ip1_optimizer", "int3c2e_ip2_optimizer", "int2c2e_optimizer",
                 "int2e_ip2_optimizer", "int2e_ipip1_optimizer", "int2e_ip1ip2_optimizer"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.POINTER(vp), ip, ctypes.c_int, ip, ctypes.c_int, dp]
    L.CINTdel_optimizer.restype = None
    L.CINTdel_optimizer.argtypes = [ctypes.POINTER(vp)]
    L.mcint_optimizer_on_device.restype = ctypes.c_int
    L.mcint_optimizer_on_device.argtypes = [vp]
    L.CINTrys_roots.restype = None
    L.CINTrys_roots.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
    L.mcint_rys_roots_batch.restype = ctypes.c_int
    L.mcint_rys_roots_batch.argtypes = [ctypes.c_int, dp, ctypes.c_size_t, dp, dp]
    L.CINTsr_rys_roots.restype = None
    L.CINTsr_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
    L.mcint_sr_rys_roots_batch.restype = ctypes.c_int
    L.mcint_sr_rys_roots_batch.argtypes = [ctypes.c_int, dp, dp, ctypes.c_size_t, dp, dp]
    L.CINTgto_norm.restype = ctypes.c_double
    L.CINTgto_norm.argtypes = [ctypes.c_int, ctypes.c_double]
    _lib = L
    # per-class team sizes measured by the autotuner on MI355X
    from . import autotune
    if os.path.exists(autotune.DEFAULT_TABLE):
        autotune.apply(autotune.load())
    return L


TUNE_PRIMS_PER_LANE = 0
TUNE_GENERIC_ONLY = 1
TUNE_GRAPH_BRANCHES = 2
TUNE_WAVE_KERNELS = 3


def tune(key: int, value: int) -> int:
    """Set a plan-construction tuning value (mcint_tune); returns the old one."""
    rc = lib().mcint_tune(key, value)
    if rc < 0:
        raise McintError(f"mcint_tune failed ({rc}): {last_error()}")
    return rc


def tune_class(name: str, cls, prims_per_lane: int) -> int:
    """Per-class primitive quartets per lane (mcint_tune_class); 0 resets."""
    rc = lib().mcint_tune_class(INTOR_CODES[name], *[int(c) for c in cls], int(prims_per_lane))
    if rc < 0:
        raise McintError(f"mcint_tune_class failed ({rc}): {last_error()}")
    return rc


def tune_class_wave(cls, mode: int, team: int = 1) -> int:
    """Buckets of a chunked plain int2e class with teams of `team` lanes:
    2 the wave kernel, 1 the chunk kernels, 0 the built-in rule
    (mcint_tune_class_wave)."""
    rc = lib().mcint_tune_class_wave(*[int(c) for c in cls], int(team), int(mode))
    if rc < 0:
        raise McintError(f"mcint_tune_class_wave failed ({rc}): {last_error()}")
    return rc


def wave_kernel_exists(cls) -> bool:
    return bool(lib().mcint_wave_kernel_exists(*[int(c) for c in cls]))


def last_error() -> str:
    return lib().mcint_last_error().decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise McintError(f"{what} failed ({rc}): {last_error()}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Basis:
    """atm/bas/env arrays (host), in libcint's layout."""

    def __init__(self, atm, bas, env):
        self.atm = np.ascontiguousarray(atm, dtype=np.int32)
        self.bas = np.ascontiguousarray(bas, dtype=np.int32)
        self.env = np.ascontiguousarray(env, dtype=np.float64)
        self.natm = int(self.atm.shape[0])
        self.nbas = int(self.bas.shape[0])

    def shell_dim(self, ish: int, sph: bool = True) -> int:
        l = int(self.bas[ish, 1])
        return (2 * l + 1 if sph else (l + 1) * (l + 2) // 2) * int(self.bas[ish, 3])

    def block_sizes(self, name: str, shls: np.ndarray) -> np.ndarray:
        sph = not name.endswith("cart")
        slots, ncomp = intor_shape(name)
        shls = np.asarray(shls).reshape(-1, 4)
        size = np.full(shls.shape[0], ncomp, dtype=np.int64)
        for a in slots:
            l = self.bas[shls[:, a], 1].astype(np.int64)
            nc = self.bas[shls[:, a], 3].astype(np.int64)
            size *= ((2 * l + 1) if sph else (l + 1) * (l + 2) // 2) * nc
        return size


def intor(name: str, shls, basis: Basis, dims=None):
    """One libcint-style call; returns (Fortran-ordered array, return value)."""
    L = lib()
    sph = not name.endswith("cart")
    slots, ncomp = intor_shape(name)
    d = [basis.shell_dim(int(shls[a]), sph) for a in slots]
    out = np.zeros(int(np.prod(d)) * ncomp)
    sh = np.asarray(list(shls) + [0] * (4 - len(shls)), dtype=np.int32)
    rc = getattr(L, name)(_dp(out), None, _ip(sh), _ip(basis.atm), basis.natm, _ip(basis.bas),
                          basis.nbas, _dp(basis.env), None, None)
    shape = tuple(d) + ((ncomp,) if ncomp > 1 else ())
    return out.reshape(shape, order="F"), rc


def batch(name: str, shls, basis: Basis):
    """Host batch: packed outputs of a quartet list; returns (out, offsets, not0)."""
    L = lib()
    shls = np.ascontiguousarray(shls, dtype=np.int32).reshape(-1, 4)
    nq = shls.shape[0]
    sizes = basis.block_sizes(name, shls)
    off = np.zeros(nq + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    out = np.zeros(max(int(off[-1]), 1))
    not0 = np.zeros(nq, dtype=np.int32)
    f = getattr(L, name + "_batch")
    rc = f(_dp(out), None, _ip(shls), nq, _ip(basis.atm), basis.natm, _ip(basis.bas), basis.nbas,
           _dp(basis.env), None, _ip(not0), None)
    _check(rc, name + "_batch")
    return out[: int(off[-1])], off, not0


def rys_roots(nroots: int, x):
    """Device Rys roots/weights for an array of x: returns (u[n, nroots], w[n, nroots])."""
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    u = np.zeros((x.size, nroots))
    w = np.zeros((x.size, nroots))
    _check(lib().mcint_rys_roots_batch(nroots, _dp(x), x.size, _dp(u), _dp(w)),
           "mcint_rys_roots_batch")
    return u, w


def sr_rys_roots(nroots: int, x, lower):
    """Device short-range rule (weight exp(-x t^2) on [lower, 1]) for arrays of
    x and lower (broadcast together): returns (u[n, nroots], w[n, nroots])."""
    x, lower = np.broadcast_arrays(np.atleast_1d(np.asarray(x, dtype=np.float64)),
                                   np.atleast_1d(np.asarray(lower, dtype=np.float64)))
    x = np.ascontiguousarray(x.ravel())
    lower = np.ascontiguousarray(lower.ravel())
    u = np.zeros((x.size, nroots))
    w = np.zeros((x.size, nroots))
    _check(lib().mcint_sr_rys_roots_batch(nroots, _dp(x), _dp(lower), x.size, _dp(u), _dp(w)),
           "mcint_sr_rys_roots_batch")
    return u, w


class DeviceBasis:
    """mcint_basis: atm/bas/env uploaded once to the current device."""

    def __init__(self, basis: Basis):
        self.basis = basis
        h = ctypes.c_void_p()
        _check(lib().mcint_basis_create(ctypes.byref(h), _ip(basis.atm), basis.natm,
                                        _ip(basis.bas), basis.nbas, _dp(basis.env)),
               "mcint_basis_create")
        self.handle = h

    def schwarz(self, nshell: int = 0) -> np.ndarray:
        q = np.zeros((self.basis.nbas, self.basis.nbas))
        _check(lib().mcint_schwarz(self.handle, nshell, _dp(q)), "mcint_schwarz")
        return q

    def close(self):
        if self.handle:
            lib().mcint_basis_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """mcint_plan over a host list of shell tuples; outputs live on the device."""

    def __init__(self, dbasis: DeviceBasis, name: str, shls, out_offsets=None, fp32_flags=None):
        self.dbasis = dbasis
        self.name = name
        shls = np.ascontiguousarray(shls, dtype=np.int32).reshape(-1, 4)
        self.nq = shls.shape[0]
        h = ctypes.c_void_p()
        osz = ctypes.c_int64()
        offp = None
        if out_offsets is not None:
            self._off = np.ascontiguousarray(out_offsets, dtype=np.int64)
            offp = self._off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        if fp32_flags is None:
            _check(lib().mcint_plan_create(ctypes.byref(h), dbasis.handle, INTOR_CODES[name],
                                           _ip(shls), self.nq, offp, ctypes.byref(osz)),
                   "mcint_plan_create")
        else:
            self._flags = np.ascontiguousarray(fp32_flags, dtype=np.uint8)
            assert self._flags.size == self.nq
            _check(lib().mcint_plan_create_mixed(ctypes.byref(h), dbasis.handle, INTOR_CODES[name],
                                                 _ip(shls), self.nq, offp,
                                                 self._flags.ctypes.data_as(ctypes.c_void_p),
                                                 ctypes.byref(osz)), "mcint_plan_create_mixed")
        self.handle = h
        self.out_size = int(osz.value)

    def stats(self) -> dict:
        n = (ctypes.c_int64 * 6)()
        _check(lib().mcint_plan_stats(self.handle, n), "mcint_plan_stats")
        return dict(tasks=n[0], buckets=n[1], fast_tasks=n[2], generic_tasks=n[3],
                    prim_quartets=n[4], out_size=n[5],
                    fp32_tasks=int(lib().mcint_plan_fp32_tasks(self.handle)))

    def buckets(self) -> list:
        """Per bucket: dict(cls=(la,lb,lc,ld), perm (-1: every axis order of the class), team,
        path, tasks)."""
        L = lib()
        out = []
        info = (ctypes.c_int64 * 8)()
        for i in range(L.mcint_plan_nbuckets(self.handle)):
            _check(L.mcint_plan_bucket_info(self.handle, i, info), "mcint_plan_bucket_info")
            out.append(dict(cls=tuple(info[:4]), perm=info[4], team=info[5],
                            path=("generic", "fast", "chunked")[info[6]], tasks=info[7]))
        return out

    def bucket_tasks(self, i: int, first: int = 0, n: int | None = None) -> np.ndarray:
        """List positions of tasks first..first+n-1 of bucket i (launch order)."""
        if n is None:
            n = self.buckets()[i]["tasks"] - first
        out = np.zeros(max(n, 1), dtype=np.int32)
        _check(lib().mcint_plan_bucket_tasks(self.handle, i, first, n, _ip(out)),
               "mcint_plan_bucket_tasks")
        return out[:n]

    def count_primitives(self) -> np.ndarray:
        n = lib().mcint_plan_nbuckets(self.handle)
        c = np.zeros(max(n, 1), dtype=np.int64)
        _check(lib().mcint_plan_count_primitives(
            self.handle, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))), "count_primitives")
        return c[:n]

    def execute_bucket(self, i: int, out_ptr: int, stream_ptr: int = 0):
        _check(lib().mcint_plan_execute_bucket(self.handle, i, ctypes.c_void_p(out_ptr), None,
                                               ctypes.c_void_p(stream_ptr or None)),
               "mcint_plan_execute_bucket")

    def execute(self, out_ptr: int, not0_ptr: int = 0, stream_ptr: int = 0, graph: bool = True):
        """Run on device pointers (e.g. torch tensor .data_ptr(), stream.cuda_stream)."""
        f = lib().mcint_plan_execute if graph else lib().mcint_plan_execute_direct
        _check(f(self.handle, ctypes.c_void_p(out_ptr), ctypes.c_void_p(not0_ptr or None),
                 ctypes.c_void_p(stream_ptr or None)), "mcint_plan_execute")

    def execute_noalloc(self, out_ptr: int, not0_ptr: int = 0, stream_ptr: int = 0):
        """As execute(graph=False), allocating no device memory: per-task
        blocks live in the library's fixed device array (calls that overlap
        on one GPU race)."""
        _check(lib().mcint_plan_execute_noalloc(self.handle, ctypes.c_void_p(out_ptr),
                                                ctypes.c_void_p(not0_ptr or None),
                                                ctypes.c_void_p(stream_ptr or None)),
               "mcint_plan_execute_noalloc")

    def execute_grad(self, dm_ptr: int, nao: int, grad_ptr: int, stream_ptr: int = 0):
        """Gradient digest of an int2e_ip1_sph plan: grad (natm x 3, device)
        += contraction of the derivative integrals with the density dm."""
        _check(lib().mcint_plan_execute_grad(self.handle, ctypes.c_void_p(dm_ptr), int(nao),
                                             ctypes.c_void_p(grad_ptr),
                                             ctypes.c_void_p(stream_ptr or None)),
               "mcint_plan_execute_grad")

    def execute_jk(self, dm_ptr: int, nao: int, vj_ptr: int = 0, vk_ptr: int = 0,
                   stream_ptr: int = 0):
        """J and K (nao x nao, device, overwritten; either pointer 0 to skip)
        contracted inside the integral kernels of an int2e_sph plan over
        unique quartets; no ERI buffer."""
        _check(lib().mcint_plan_execute_jk(self.handle, ctypes.c_void_p(dm_ptr), int(nao),
                                           ctypes.c_void_p(vj_ptr or None),
                                           ctypes.c_void_p(vk_ptr or None),
                                           ctypes.c_void_p(stream_ptr or None)),
               "mcint_plan_execute_jk")

    def close(self):
        if self.handle:
            lib().mcint_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
