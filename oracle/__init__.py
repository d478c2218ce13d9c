"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product library
(libcint_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

ROOTS_EXACT = 0
ROOTS_TABLE = 1

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        for name in ("orc_int2e_sph", "orc_int2e_cart", "orc_int2e_ip1_sph",
                     "orc_int2e_ip1_cart", "orc_int3c2e_sph", "orc_int3c2e_cart",
                     "orc_int3c2e_ip1_sph", "orc_int3c2e_ip1_cart", "orc_int3c2e_ip2_sph",
                     "orc_int3c2e_ip2_cart", "orc_int2c2e_sph", "orc_int2c2e_cart",
                     "orc_int2e_ip2_sph", "orc_int2e_ip2_cart", "orc_int2e_ipip1_sph",
                     "orc_int2e_ipip1_cart", "orc_int2e_ip1ip2_sph", "orc_int2e_ip1ip2_cart"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [dp, ip, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp]
        L.orc_batch.restype = ctypes.c_int
        L.orc_batch.argtypes = [ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int64), ip,
                                ctypes.c_int64, ip, ctypes.c_int, ip, ctypes.c_int, dp, ip,
                                ctypes.c_int]
        L.orc_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.orc_rys_roots_exact.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_rys_roots_table.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_sr_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_set_root_mode.argtypes = [ctypes.c_int]
        L.orc_c2s_matrix.argtypes = [ctypes.c_int, dp]
        L.orc_grad_ip1.restype = ctypes.c_int
        L.orc_grad_ip1.argtypes = [ip, ctypes.c_int, ip, ctypes.c_int, dp, dp, ctypes.c_int, dp,
                                   ctypes.c_int]
        L.orc_gto_norm.restype = ctypes.c_double
        L.orc_gto_norm.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def set_root_mode(mode: int) -> None:
    lib().orc_set_root_mode(mode)


def boys(mmax: int, x: float) -> np.ndarray:
    out = np.zeros(mmax + 1)
    lib().orc_boys(mmax, x, _dp(out))
    return out


def rys_roots(n: int, x: float, exact: bool = True):
    u = np.zeros(n)
    w = np.zeros(n)
    f = lib().orc_rys_roots_exact if exact else lib().orc_rys_roots_table
    f(n, x, _dp(u), _dp(w))
    return u, w


def sr_rys_roots(n: int, x: float, lower: float):
    """Short-range rule (weight exp(-x t^2) on [lower, 1]): (u, w)."""
    u = np.zeros(n)
    w = np.zeros(n)
    lib().orc_sr_rys_roots(n, x, lower, _dp(u), _dp(w))
    return u, w


def c2s_matrix(l: int) -> np.ndarray:
    nc = (l + 1) * (l + 2) // 2
    out = np.zeros((2 * l + 1, nc))
    lib().orc_c2s_matrix(l, _dp(out))
    return out


class Basis:
    """Holds contiguous int32/float64 copies of atm/bas/env."""

    def __init__(self, atm, bas, env):
        self.atm = np.ascontiguousarray(atm, dtype=np.int32)
        self.bas = np.ascontiguousarray(bas, dtype=np.int32)
        self.env = np.ascontiguousarray(env, dtype=np.float64)
        self.natm = self.atm.shape[0]
        self.nbas = self.bas.shape[0]

    def shell_dim(self, ish: int, sph: bool = True) -> int:
        l = int(self.bas[ish, 1])
        n = 2 * l + 1 if sph else (l + 1) * (l + 2) // 2
        return n * int(self.bas[ish, 3])


_KIND = {"int2e_sph": 0, "int2e_cart": 1, "int2e_ip1_sph": 2, "int2e_ip1_cart": 3,
         "int3c2e_sph": 4, "int3c2e_cart": 5, "int3c2e_ip1_sph": 6, "int3c2e_ip1_cart": 7,
         "int3c2e_ip2_sph": 8, "int3c2e_ip2_cart": 9, "int2c2e_sph": 10, "int2c2e_cart": 11,
         "int2e_ip2_sph": 12, "int2e_ip2_cart": 13, "int2e_ipip1_sph": 14,
         "int2e_ipip1_cart": 15, "int2e_ip1ip2_sph": 16, "int2e_ip1ip2_cart": 17}


def _shape(name):
    """(shell slots of shls used, derivative components) of an integral."""
    ncomp = 9 if ("_ipip" in name or "_ip1ip2" in name) else (3 if "_ip" in name else 1)
    if name.startswith("int2c2e"):
        return (0, 1), ncomp
    if name.startswith("int3c2e"):
        return (0, 1, 2), ncomp
    return (0, 1, 2, 3), ncomp


def intor(name: str, shls, basis: Basis, dims=None):
    """One call of a libcint-named integral; returns (array Fortran-shaped, not0)."""
    sph = not name.endswith("cart")
    slots, ncomp = _shape(name)
    d = [basis.shell_dim(int(shls[s]), sph) for s in slots]
    n = int(np.prod(d)) * ncomp
    out = np.zeros(n)
    sh = np.asarray(list(shls) + [0] * (4 - len(shls)), dtype=np.int32)
    f = getattr(lib(), "orc_" + name)
    st = f(_dp(out), None, _ip(sh), _ip(basis.atm), basis.natm, _ip(basis.bas), basis.nbas,
           _dp(basis.env))
    shape = tuple(d) + ((ncomp,) if ncomp > 1 else ())
    return out.reshape(shape, order="F"), st


def grad_ip1(basis: Basis, dm: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """dE/dR (natm x 3) of E = 1/2 D J - 1/4 D K from int2e_ip1_sph blocks of
    every (i, j, k >= l) contracted with the density on the CPU."""
    dm = np.ascontiguousarray(dm, dtype=np.float64)
    g = np.zeros((basis.natm, 3))
    rc = lib().orc_grad_ip1(_ip(basis.atm), basis.natm, _ip(basis.bas), basis.nbas,
                            _dp(basis.env), _dp(dm), dm.shape[0], _dp(g), int(nthreads))
    if rc != 0:
        raise ValueError("orc_grad_ip1: density dimension does not match the basis")
    return g


def batch(name: str, shls: np.ndarray, basis: Basis, nthreads: int = 0):
    """Packed outputs for a list of quartets; returns (out, offsets, not0)."""
    sph = not name.endswith("cart")
    slots, ncomp = _shape(name)
    shls = np.ascontiguousarray(shls, dtype=np.int32).reshape(-1, 4)
    nq = shls.shape[0]
    dims = np.ones((nq, 4), dtype=np.int64)
    for s in slots:
        l = basis.bas[shls[:, s], 1].astype(np.int64)
        nc = basis.bas[shls[:, s], 3].astype(np.int64)
        dims[:, s] = ((2 * l + 1) if sph else (l + 1) * (l + 2) // 2) * nc
    sizes = dims.prod(axis=1) * ncomp
    off = np.zeros(nq + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    out = np.zeros(int(off[-1]))
    not0 = np.zeros(nq, dtype=np.int32)
    lib().orc_batch(_KIND[name], _dp(out), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                    _ip(shls), nq, _ip(basis.atm), basis.natm, _ip(basis.bas), basis.nbas,
                    _dp(basis.env), _ip(not0), int(nthreads))
    return out, off, not0
