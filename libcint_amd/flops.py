"""Algorithmic FLOP model of libcint's Rys ERI path, from class metadata only.

Counts follow the reference's loop bounds, per surviving primitive quartet of
class (li, lj, lk, ll) with n = nroots and nf = product of Cartesian counts:
  coefficients  25 n                        (src/g2e.c:4518-4540)
  2D recurrence 3 n 5 (nmax+1)(mmax+1)      (src/g2e.c:272-421)
  2D -> 4D      3 n 2 N_hrr                 (src/g2e.c:428-693)
  gout          3 n nf                      (src/cint2e.c:969-1181)
and per contracted quartet the dense Cartesian->spherical transform of
src/cart2sph.c:5366-5370.  Root evaluation is not counted.  This is the
work libcint's algorithm does; the GPU kernels do the same mathematics with
fewer operations (transfers after contraction), so the FLOP rate quoted
against it is an algorithmic rate.
"""
from __future__ import annotations


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def n_hrr(li: int, lj: int, lk: int, ll: int) -> int:
    """Entries produced by the two transfers (per root, per direction)."""
    nmax, mmax = li + lj, lk + ll
    bra_target = min(li, lj)
    ket_target = min(lk, ll)
    bra = sum((nmax - t + 1) * (mmax + 1) for t in range(1, bra_target + 1))
    ket = (li + 1) * (lj + 1) * sum(mmax - t + 1 for t in range(1, ket_target + 1))
    return bra + ket


def prim_flops(li: int, lj: int, lk: int, ll: int) -> int:
    n = (li + lj + lk + ll) // 2 + 1
    nf = ncart(li) * ncart(lj) * ncart(lk) * ncart(ll)
    f_coef = 25 * n
    f_vrr = 3 * n * 5 * (li + lj + 1) * (lk + ll + 1)
    f_hrr = 3 * n * 2 * n_hrr(li, lj, lk, ll)
    f_gout = 3 * n * nf
    return f_coef + f_vrr + f_hrr + f_gout


def c2s_flops(li: int, lj: int, lk: int, ll: int) -> int:
    nfi, nfj, nfk, nfl = ncart(li), ncart(lj), ncart(lk), ncart(ll)
    di, dj, dk, dl = 2 * li + 1, 2 * lj + 1, 2 * lk + 1, 2 * ll + 1
    return 2 * (nfi * nfk * nfl * nfj * dj + nfi * nfk * dj * nfl * dl
                + nfi * dl * dj * nfk * dk + nfi * di * dk * dl * dj)


def bucket_flops(cls, n_prim_quartets: int, n_quartets: int) -> int:
    li, lj, lk, ll = cls
    return prim_flops(li, lj, lk, ll) * int(n_prim_quartets) + c2s_flops(li, lj, lk, ll) * int(n_quartets)
