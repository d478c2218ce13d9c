"""Seeded synthetic molecules and basis sets in libcint's atm/bas/env layout.

The data model is libcint's (include/cint.h.in:26-67): ``atm`` holds
ATM_SLOTS=6 ints per atom (charge, pointer to the coordinates in ``env``),
``bas`` holds BAS_SLOTS=8 ints per shell (atom, l, nprim, nctr, kappa,
pointer to exponents, pointer to coefficients), ``env`` holds doubles with the
user data starting at PTR_ENV_START=20.  Contraction coefficients stored in
``env`` already include the radial normalisation ``CINTgto_norm``
(src/misc.c:72-93), exactly as libcint's callers store them
(examples/c_call_spheric.c).

Five workloads are defined, each from a fixed seed:

* ``c0``: C4H8, DZ-like, l <= 2 (48 shells, 96 spherical AOs)
* ``c1``: C10N3O2H15, TZ-like, l <= 3 (210 shells, 540 spherical AOs)
* ``c2``: the c1 molecule (the 30-atom TZ cluster) plus a random symmetric
  density matrix, for the mixed-precision J/K build
* ``c3``: C8H12 with the c0 basis plus a d shell on H
* ``c4``: C16H24, c1-like orbital basis plus an even-tempered auxiliary basis
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# slots (include/cint.h.in:26-67)
PTR_EXPCUTOFF = 0
PTR_RANGE_OMEGA = 8
PTR_ENV_START = 20
CHARGE_OF, PTR_COORD, NUC_MOD_OF, PTR_ZETA = 0, 1, 2, 3
ATM_SLOTS = 6
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = range(7)
BAS_SLOTS = 8

CHARGES = {"H": 1, "C": 6, "N": 7, "O": 8}


def gto_norm(l: int, a: float) -> float:
    """1/sqrt(int_0^inf r^(2l+2) exp(-2 a r^2) dr)  (src/misc.c:72-93)."""
    n1 = l + 1.5
    integral = math.gamma(n1) / (2.0 * (2.0 * a) ** n1)
    return 1.0 / math.sqrt(integral)


def _radial_overlap(l: int, a: float, b: float) -> float:
    n1 = l + 1.5
    return math.gamma(n1) / (2.0 * (a + b) ** n1)


def normalized_coefficients(l: int, exps, coefs) -> np.ndarray:
    """Coefficients (times gto_norm) of a unit-norm contracted radial function."""
    exps = np.asarray(exps, dtype=np.float64)
    c = np.asarray(coefs, dtype=np.float64) * np.array([gto_norm(l, a) for a in exps])
    s = 0.0
    for p in range(len(exps)):
        for q in range(len(exps)):
            s += c[p] * c[q] * _radial_overlap(l, exps[p], exps[q])
    return c / math.sqrt(s)


@dataclasses.dataclass
class Molecule:
    """atm/bas/env arrays plus bookkeeping for one synthetic system."""

    name: str
    atm: np.ndarray  # int32 [natm, 6]
    bas: np.ndarray  # int32 [nbas, 8]
    env: np.ndarray  # float64
    symbols: list
    coords: np.ndarray  # [natm, 3] Bohr
    aux_start: int = -1  # first auxiliary shell (c4), -1 if none

    @property
    def natm(self) -> int:
        return int(self.atm.shape[0])

    @property
    def nbas(self) -> int:
        return int(self.bas.shape[0])

    def nao_sph(self, shells=None) -> int:
        b = self.bas if shells is None else self.bas[shells]
        return int(((2 * b[:, ANG_OF] + 1) * b[:, NCTR_OF]).sum())

    def ao_loc_sph(self, nshell=None) -> np.ndarray:
        b = self.bas[: (self.nbas if nshell is None else nshell)]
        d = (2 * b[:, ANG_OF] + 1) * b[:, NCTR_OF]
        return np.concatenate([[0], np.cumsum(d)]).astype(np.int64)

    def ao_loc_cart(self, nshell=None) -> np.ndarray:
        b = self.bas[: (self.nbas if nshell is None else nshell)]
        l = b[:, ANG_OF]
        d = (l + 1) * (l + 2) // 2 * b[:, NCTR_OF]
        return np.concatenate([[0], np.cumsum(d)]).astype(np.int64)

    def with_coords(self, coords: np.ndarray) -> "Molecule":
        """Copy with atoms moved (used for finite-difference gradients)."""
        env = self.env.copy()
        for ia in range(self.natm):
            p = self.atm[ia, PTR_COORD]
            env[p : p + 3] = coords[ia]
        return dataclasses.replace(self, env=env, coords=np.array(coords, dtype=np.float64))


class _Builder:
    def __init__(self):
        self.env = [0.0] * PTR_ENV_START
        self.atm = []
        self.bas = []

    def add_atom(self, symbol, xyz):
        ptr = len(self.env)
        self.env.extend(float(v) for v in xyz)
        self.atm.append([CHARGES[symbol], ptr, 1, 0, 0, 0])
        return len(self.atm) - 1

    def add_shell(self, iatm, l, exps, coefs, normalise=True):
        exps = np.asarray(exps, dtype=np.float64)
        if normalise:
            c = normalized_coefficients(l, exps, coefs)
        else:
            c = np.asarray(coefs, dtype=np.float64)
        pe = len(self.env)
        self.env.extend(exps.tolist())
        pc = len(self.env)
        self.env.extend(c.tolist())
        self.bas.append([iatm, l, len(exps), 1, 0, pe, pc, 0])

    def finish(self, name, symbols, coords, aux_start=-1):
        return Molecule(
            name=name,
            atm=np.array(self.atm, dtype=np.int32).reshape(-1, ATM_SLOTS),
            bas=np.array(self.bas, dtype=np.int32).reshape(-1, BAS_SLOTS),
            env=np.array(self.env, dtype=np.float64),
            symbols=list(symbols),
            coords=np.array(coords, dtype=np.float64),
            aux_start=aux_start,
        )


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _heavy_chain(rng, n, dmin, dmax):
    """Random chain: each atom bonded to a random earlier atom, no close contacts."""
    pos = [np.zeros(3)]
    while len(pos) < n:
        anchor = pos[rng.integers(len(pos))] if len(pos) > 1 else pos[0]
        if len(pos) > 1 and rng.random() < 0.6:
            anchor = pos[-1]
        p = anchor + rng.uniform(dmin, dmax) * _unit(rng)
        if min(np.linalg.norm(p - q) for q in pos) >= dmin - 1e-9:
            pos.append(p)
    return pos


def _hydrogens(rng, heavy, nh, r0, dr, dmin=1.9):
    pos = []
    k = 0
    while len(pos) < nh:
        c = heavy[k % len(heavy)] if nh <= 2 * len(heavy) else heavy[rng.integers(len(heavy))]
        p = c + rng.uniform(r0 - dr, r0 + dr) * _unit(rng)
        others = list(heavy) + pos
        near = min(np.linalg.norm(p - q) for q in others if q is not c)
        if near >= dmin:
            pos.append(p)
            k += 1
    return pos


def _even_tempered(a0, ratio, n):
    return [a0 * ratio**k for k in range(n)]


def _dz_shells(b, rng, iatm, heavy, d_on_h=False):
    """c0 basis: per C 3s+2p+1d, per H 2s+1p (+1d for c3).

    The first s and the first p shell of each atom are contracted over 6
    even-tempered primitives 0.15*3.2^k; the other shells are single
    primitives with exponents uniform in [0.1, 1.2]; coefficients uniform in
    (0.1, 1] before normalisation.
    """
    ctr = _even_tempered(0.15, 3.2, 6)

    def contracted(l):
        b.add_shell(iatm, l, ctr, 1.1 - rng.uniform(0.1, 1.0, size=6))

    def single(l):
        b.add_shell(iatm, l, [rng.uniform(0.1, 1.2)], [1.0])

    if heavy:
        contracted(0); single(0); single(0)
        contracted(1); single(1)
        single(2)
    else:
        contracted(0); single(0)
        contracted(1)
        if d_on_h:
            single(2)


def _tz_shells(b, rng, iatm, heavy):
    """c1 basis: per heavy atom 4s+3p+2d+1f, per H 3s+1p.

    The first s and p shells are contracted over 8-10 even-tempered
    primitives 0.1*3^k; other shells are single primitives in [0.1, 1.2].
    """
    def contracted(l):
        n = int(rng.integers(8, 11))
        b.add_shell(iatm, l, _even_tempered(0.1, 3.0, n), 1.1 - rng.uniform(0.1, 1.0, size=n))

    def single(l):
        b.add_shell(iatm, l, [rng.uniform(0.1, 1.2)], [1.0])

    if heavy:
        contracted(0); single(0); single(0); single(0)
        contracted(1); single(1); single(1)
        single(2); single(2)
        single(3)
    else:
        contracted(0); single(0); single(0)
        contracted(1)


def _aux_shells(b, iatm, heavy):
    """c4 auxiliary basis: even-tempered single primitives 0.08*2.2^k."""
    counts = (12, 8, 6, 4, 2) if heavy else (6, 3, 2)
    for l, n in enumerate(counts):
        for a in _even_tempered(0.08, 2.2, n):
            b.add_shell(iatm, l, [a], [1.0])


def make_molecule(config: str) -> Molecule:
    """Build one of the five synthetic workloads by name ('c0'..'c4')."""
    cid = int(config[1:])
    rng = np.random.default_rng(20260 + cid)
    if config == "c0":
        heavy_sym = ["C"] * 4
        nh = 8
        basis = "dz"
    elif config in ("c1", "c2"):
        rng = np.random.default_rng(20261)
        heavy_sym = ["C"] * 10 + ["N"] * 3 + ["O"] * 2
        nh = 15
        basis = "tz"
    elif config == "c3":
        heavy_sym = ["C"] * 8
        nh = 12
        basis = "dz_dh"
    elif config == "c4":
        heavy_sym = ["C"] * 16
        nh = 24
        basis = "tz_aux"
    else:
        raise ValueError(f"unknown config {config!r}")

    if basis.startswith("dz"):
        heavy = _heavy_chain(rng, len(heavy_sym), 2.6, 3.0)
    else:
        heavy = _heavy_chain(rng, len(heavy_sym), 2.5, 3.2)
    hyd = _hydrogens(rng, heavy, nh, 2.06, 0.05)
    symbols = heavy_sym + ["H"] * nh
    coords = np.array(heavy + hyd)

    b = _Builder()
    for s, xyz in zip(symbols, coords):
        b.add_atom(s, xyz)
    for ia, s in enumerate(symbols):
        if basis == "dz":
            _dz_shells(b, rng, ia, s != "H")
        elif basis == "dz_dh":
            _dz_shells(b, rng, ia, s != "H", d_on_h=True)
        else:
            _tz_shells(b, rng, ia, s != "H")
    aux_start = -1
    if basis == "tz_aux":
        aux_start = len(b.bas)
        for ia, s in enumerate(symbols):
            _aux_shells(b, ia, s != "H")
    name = {"c0": "C4H8/DZ", "c1": "C10N3O2H15/TZ", "c2": "C10N3O2H15/TZ",
            "c3": "C8H12/DZ(d on H)", "c4": "C16H24/TZ+aux"}[config]
    return b.finish(name, symbols, coords, aux_start)


def random_density(nao: int, seed: int) -> np.ndarray:
    """D = A A^T / N with A ~ N(0, 0.1^2): symmetric positive semidefinite."""
    rng = np.random.default_rng(seed)
    a = rng.normal(0.0, 0.1, size=(nao, nao))
    return a @ a.T / nao


def unique_quartets(nbas: int) -> np.ndarray:
    """All (i,j,k,l) with i>=j, k>=l, ij>=kl (8-fold symmetry), as int32 [n,4]."""
    ii, jj = np.tril_indices(nbas)
    npair = ii.size
    p, q = np.tril_indices(npair)
    out = np.empty((p.size, 4), dtype=np.int32)
    out[:, 0] = ii[p]
    out[:, 1] = jj[p]
    out[:, 2] = ii[q]
    out[:, 3] = jj[q]
    return out


def screened_quartets(q_schwarz: np.ndarray, nshell: int, thresh: float):
    """Unique quartets (i>=j, k>=l, ij>=kl) of the first nshell shells whose
    Schwarz bound sqrt((ij|ij)) sqrt((kl|kl)) (q_schwarz[i, j] = sqrt(max
    |(ij|ij)|)) is at least thresh, in the order of unique_quartets; also
    returns the number of unique quartets before screening."""
    ii, jj = np.tril_indices(nshell)
    qpair = q_schwarz[ii, jj]
    quart = unique_quartets(nshell)
    p, r = np.tril_indices(ii.size)
    keep = qpair[p] * qpair[r] >= thresh
    return quart[keep], int(quart.shape[0])


def subset(mol: Molecule, natoms: int) -> Molecule:
    """The first natoms atoms of a molecule with their shells (same env)."""
    keep = mol.bas[:, ATOM_OF] < natoms
    if mol.aux_start >= 0:
        keep[mol.aux_start:] = False
    return dataclasses.replace(mol, name=f"{mol.name}[:{natoms}]", atm=mol.atm[:natoms].copy(),
                               bas=mol.bas[keep].copy(), symbols=mol.symbols[:natoms],
                               coords=mol.coords[:natoms].copy(), aux_start=-1)
