"""Coulomb/exchange matrices, the two-electron energy and its nuclear gradient.

The caller-side layer of the integral path (not part of libcint; PySCF's
``_vhf`` plays this role for the reference): on the device,

    J_pq = sum_rs (pq|rs) D_rs,     K_pr = sum_qs (pq|rs) D_qs,
    E    = 1/2 sum D_pq J_pq - 1/4 sum D_pr K_pr,
    dE/dA_x = sum_{i on A} sum_jkl (d/dr_x i j|kl) [-2 D_ij D_kl + D_ik D_jl],

where (d/dr_x i j|kl) is libcint's int2e_ip1 (src/autocode/grad2.c:19-68).
The gradient is a digest of the derivative integrals computed on the fly
(they are never stored).  ``TwoElectronEnergy`` exposes E(coords, D) as a
torch.autograd.Function whose backward runs that digest.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Basis, DeviceBasis, Plan, _check, lib
from .synthetic import Molecule, unique_quartets


def _ptr(t: torch.Tensor):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def nao_sph(basis: Basis) -> int:
    return int(((2 * basis.bas[:, 1] + 1) * basis.bas[:, 3]).sum())


def fp32_quartet_flags(db: DeviceBasis, quartets: np.ndarray, dm, threshold: float) -> np.ndarray:
    """1 where sqrt((ij|ij)) sqrt((kl|kl)) max|D| < threshold (FP32 tier);
    dm may be the density or any array whose largest |entry| bounds it."""
    qs = db.schwarz()
    dmax = float(np.abs(np.asarray(dm)).max())
    bound = qs[quartets[:, 0], quartets[:, 1]] * qs[quartets[:, 2], quartets[:, 3]] * dmax
    return (bound < threshold).astype(np.uint8)


class JKBuilder:
    """J and K of a basis for any symmetric density, on the device.

    Holds the ERI plan of the unique quartets (optionally Schwarz-screened
    and split into an FP32 tier).  With fused=True (the default) ``run(dm)``
    contracts every block with the density inside the integral kernels and
    no ERI buffer exists; with fused=False it keeps the device ERI buffer
    and a grouped digest plan, and ``run(dm)`` recomputes the ERIs unless
    ``recompute=False`` and the buffer is current, then digests them.
    fp32_threshold > 0 selects the mixed-precision tier: quartets whose
    Schwarz bound times dmax (a bound on |D|) is below it run their
    recurrences in FP32.
    """

    def __init__(self, basis: Basis, quartets: np.ndarray | None = None, device="cuda",
                 fp32_threshold: float = 0.0, dmax: float = 1.0, fused: bool = True):
        import ctypes
        self.dev = torch.device(device)
        self.nao = nao_sph(basis)
        self.db = DeviceBasis(basis)
        q = unique_quartets(basis.nbas) if quartets is None else np.ascontiguousarray(quartets, np.int32)
        self.quartets = q
        flags = None
        if fp32_threshold > 0:
            flags = fp32_quartet_flags(self.db, q, np.array([dmax]), fp32_threshold)
        self.plan = Plan(self.db, "int2e_sph", q, fp32_flags=flags)
        self.fused = fused
        self.jk_handle = None
        self.fresh = False
        if fused:
            self.eri = None
            return
        self.eri = torch.empty(max(self.plan.out_size, 1), dtype=torch.float64, device=self.dev)
        h = ctypes.c_void_p()
        _check(lib().mcint_jk_plan_create(ctypes.byref(h), self.db.handle,
                                          q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          q.shape[0], None), "mcint_jk_plan_create")
        self.jk_handle = h

    def stats(self) -> dict:
        return self.plan.stats()

    def compute_eri(self, stream_ptr: int = 0, graph: bool = False):
        if self.fused:
            raise RuntimeError("JKBuilder(fused=True) keeps no ERI buffer")
        self.plan.execute(self.eri.data_ptr(), stream_ptr=stream_ptr, graph=graph)
        self.fresh = True

    def fused_jk(self, dm, with_j=True, with_k=True, stream_ptr: int = 0):
        """J and K with the contraction inside the integral kernels."""
        dm_t = torch.as_tensor(dm, dtype=torch.float64, device=self.dev).contiguous()
        vj = torch.empty((self.nao, self.nao), dtype=torch.float64, device=self.dev) if with_j else None
        vk = torch.empty((self.nao, self.nao), dtype=torch.float64, device=self.dev) if with_k else None
        self.plan.execute_jk(dm_t.data_ptr(), self.nao, vj.data_ptr() if with_j else 0,
                             vk.data_ptr() if with_k else 0, stream_ptr)
        return vj, vk

    def digest(self, dm, with_j=True, with_k=True, stream_ptr: int = 0):
        dm_t = torch.as_tensor(dm, dtype=torch.float64, device=self.dev).contiguous()
        vj = torch.empty((self.nao, self.nao), dtype=torch.float64, device=self.dev) if with_j else None
        vk = torch.empty((self.nao, self.nao), dtype=torch.float64, device=self.dev) if with_k else None
        _check(lib().mcint_jk_plan_execute(self.jk_handle, _ptr(self.eri), _ptr(dm_t),
                                           _ptr(vj) if with_j else None,
                                           _ptr(vk) if with_k else None,
                                           stream_ptr or None), "mcint_jk_plan_execute")
        return vj, vk

    def run(self, dm, recompute: bool = True, stream_ptr: int = 0):
        if self.fused:
            return self.fused_jk(dm, stream_ptr=stream_ptr)
        if recompute or not self.fresh:
            self.compute_eri(stream_ptr)
        return self.digest(dm, stream_ptr=stream_ptr)

    def close(self):
        if self.jk_handle:
            lib().mcint_jk_plan_destroy(self.jk_handle)
            self.jk_handle = None
        self.plan.close()
        self.db.close()


def get_jk(basis: Basis, dm, nshell: int | None = None, quartets: np.ndarray | None = None,
           device="cuda", fp32_threshold: float = 0.0, stats: dict | None = None,
           fused: bool = True):
    """J and K (torch float64 on the device) for a symmetric density dm.

    fp32_threshold > 0 selects the mixed-precision tier: quartets whose
    Schwarz bound times max|D| is below it run their recurrences in FP32.
    fused=False stores the ERIs and digests the buffer afterwards.
    """
    dm_np = np.asarray(dm, dtype=np.float64)
    if quartets is None and nshell is not None:
        quartets = unique_quartets(nshell)
    jb = JKBuilder(basis, quartets, device=device, fp32_threshold=fp32_threshold,
                   dmax=float(np.abs(dm_np).max()), fused=fused)
    try:
        if stats is not None:
            st = jb.stats()
            for k in ("tasks", "fp32_tasks"):
                stats[k] = stats.get(k, 0) + st[k]
        vj, vk = jb.run(dm_np)
        torch.cuda.synchronize(jb.dev)
    finally:
        jb.close()
    return vj, vk


def energy(dm: torch.Tensor, vj: torch.Tensor, vk: torch.Tensor) -> torch.Tensor:
    return 0.5 * torch.sum(dm * vj) - 0.25 * torch.sum(dm * vk)


def gradient_quartets(nshell: int) -> np.ndarray:
    """(i, j, k>=l) for all ordered i, j: the list the gradient digest needs."""
    ii, jj = np.meshgrid(np.arange(nshell), np.arange(nshell), indexing="ij")
    kk, ll = np.tril_indices(nshell)
    nij, nkl = ii.size, kk.size
    out = np.empty((nij * nkl, 4), dtype=np.int32)
    out[:, 0] = np.repeat(ii.ravel(), nkl)
    out[:, 1] = np.repeat(jj.ravel(), nkl)
    out[:, 2] = np.tile(kk, nij)
    out[:, 3] = np.tile(ll, nij)
    return out


class GradientPlan:
    """Device plans of the ip1 list (i, j, k >= l) of a basis, reusable for
    any density: ``run(dm)`` returns dE/dR (natm x 3) of E = 1/2 D J - 1/4 D K."""

    def __init__(self, basis: Basis, natm: int, chunk: int = 64_000_000, device="cuda",
                 fp32_threshold: float = 0.0, dmax: float = 1.0):
        self.dev = torch.device(device)
        self.natm = natm
        self.nao = nao_sph(basis)
        self.db = DeviceBasis(basis)
        q_all = gradient_quartets(basis.nbas)
        self.nquartets = q_all.shape[0]
        flags = None
        if fp32_threshold > 0:
            # the digest weighs each quartet by products of two density
            # elements: FP32 where sqrt((ij|ij)) sqrt((kl|kl)) dmax^2 is small
            flags = fp32_quartet_flags(self.db, q_all, np.array([dmax * dmax]), fp32_threshold)
        self.plans = [Plan(self.db, "int2e_ip1_sph", q_all[c0:c0 + chunk],
                           fp32_flags=None if flags is None else flags[c0:c0 + chunk])
                      for c0 in range(0, q_all.shape[0], chunk)]

    def stats(self) -> dict:
        out: dict = {}
        for p in self.plans:
            for k, v in p.stats().items():
                out[k] = out.get(k, 0) + v
        return out

    def run(self, dm, stream_ptr: int = 0) -> torch.Tensor:
        dm_t = torch.as_tensor(dm, dtype=torch.float64, device=self.dev).contiguous()
        assert dm_t.shape == (self.nao, self.nao)
        grad = torch.zeros((self.natm, 3), dtype=torch.float64, device=self.dev)
        for plan in self.plans:
            plan.execute_grad(dm_t.data_ptr(), self.nao, grad.data_ptr(), stream_ptr)
        return grad

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []
        self.db.close()


def nuclear_gradient(basis: Basis, dm, natm: int, device="cuda", fp32_threshold: float = 0.0):
    """dE/dR (natm x 3, torch float64) of E = 1/2 D J - 1/4 D K.

    fp32_threshold > 0: quartets with Schwarz bound x max|D|^2 below it run
    their derivative recurrences in FP32 (the digest stays FP64)."""
    dm = np.asarray(dm, dtype=np.float64)
    gp = GradientPlan(basis, natm, device=device, fp32_threshold=fp32_threshold,
                      dmax=float(np.abs(dm).max()))
    try:
        grad = gp.run(dm)
        torch.cuda.synchronize(gp.dev)
    finally:
        gp.close()
    return grad


class TwoElectronEnergy(torch.autograd.Function):
    """E(coords, D) = 1/2 sum D J - 1/4 sum D K for a fixed basis template.

    forward: J/K on the device; backward: the nuclear-gradient digest for
    coords and dE/dD = J - K/2 for the density.
    """

    @staticmethod
    def forward(ctx, coords: torch.Tensor, dm: torch.Tensor, mol: Molecule):
        m = mol.with_coords(coords.detach().cpu().numpy())
        basis = Basis(m.atm, m.bas, m.env)
        dm_d = dm.detach().to(dtype=torch.float64)
        vj, vk = get_jk(basis, dm_d.cpu().numpy(), device=dm_d.device if dm_d.is_cuda else "cuda")
        ctx.mol = m
        ctx.save_for_backward(coords, dm_d)
        ctx.vj, ctx.vk = vj, vk
        e = energy(dm_d.to(vj.device), vj, vk)
        return e.to(coords.device)

    @staticmethod
    def backward(ctx, g):
        coords, dm = ctx.saved_tensors
        m = ctx.mol
        gc = gd = None
        if ctx.needs_input_grad[0]:
            grad = nuclear_gradient(Basis(m.atm, m.bas, m.env), dm.cpu().numpy(), m.natm)
            gc = (g.to(grad.device) * grad).to(coords.device, coords.dtype)
        if ctx.needs_input_grad[1]:
            gd = (g.to(ctx.vj.device) * (ctx.vj - 0.5 * ctx.vk)).to(dm.device, dm.dtype)
        return gc, gd, None
