"""Extended-Hueckel STO parameters and the per-orbital basis record.

The parameter rows are the reference table's values for the elements the
propagator workloads use (my_eht_parameters.dat:1 (H), :9-14 (C, N, O),
:43-45 (Ti)); the reader converts zeta from 1/a0 to 1/Angstrom and truncates
zeta and the contraction coefficients to 1e-4 exactly as
EHT_input.f:77-92 does.  Basis order per atom is s, p(m=-1,0,1 = y,z,x),
d(m=-2..2 = xy,yz,z2,xz,x2-y2) (structure.f:211-223, 228-324).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

A_BOHR = 0.52917720859  # constants_m.f:21
H_BAR = 6.58264e-4  # eV*ps, constants_m.f:23
K_WH = 1.75  # babel.f:120
QM_CUTOFF = 12.0  # Angstrom, constants_m.f:34

# symbol: (AtNo, Nvalen, [(n, l, Nzeta, IP, zeta1, zeta2, coef1, coef2), ...])
# zetas in 1/a0, exactly as the table stores them.
_RAW = {
    "H": (1, 1, [(1, 0, 1, -13.600, 1.3000, 0.0000, 1.0000, 0.0000)]),
    "C": (6, 4, [(2, 0, 1, -21.400, 1.6250, 0.0000, 1.0000, 0.0000),
                 (2, 1, 1, -11.400, 1.6250, 0.0000, 1.0000, 0.0000)]),
    "N": (7, 5, [(2, 0, 1, -26.000, 1.9500, 0.0000, 1.0000, 0.0000),
                 (2, 1, 1, -13.400, 1.9500, 0.0000, 1.0000, 0.0000)]),
    "O": (8, 6, [(2, 0, 1, -32.300, 2.2750, 0.0000, 1.0000, 0.0000),
                 (2, 1, 1, -14.800, 2.2750, 0.0000, 1.0000, 0.0000)]),
    "Ti": (22, 4, [(4, 0, 1, -8.970, 1.0750, 0.0000, 1.0000, 0.0000),
                   (4, 1, 1, -5.440, 1.0750, 0.0000, 1.0000, 0.0000),
                   (3, 2, 2, -10.810, 4.5500, 1.4000, 0.4206, 0.7839)]),
}


def _trunc4(x: float) -> float:
    """x -> int(x*1e4)*1e-4 with Fortran truncation toward zero (EHT_input.f:81-92)."""
    y = x * 1.0e4
    y = float(int(y))
    return y * 1.0e-4


@dataclass(frozen=True)
class Shell:
    n: int
    l: int
    nzeta: int
    ip: float
    zeta: tuple  # 1/Angstrom, truncated
    coef: tuple  # truncated


@dataclass(frozen=True)
class Element:
    symbol: str
    atno: int
    nvalen: int
    shells: tuple

    @property
    def dos(self) -> int:
        return sum(2 * s.l + 1 for s in self.shells)

    @property
    def ang_max(self) -> int:
        return max(s.l for s in self.shells)


def _build_table() -> dict:
    table = {}
    for sym, (atno, nval, rows) in _RAW.items():
        shells = []
        for (n, l, nz, ip, z1, z2, c1, c2) in rows:
            zz = tuple(_trunc4(z / A_BOHR) for z in (z1, z2))
            cc = tuple(_trunc4(c) for c in (c1, c2))
            shells.append(Shell(n, l, nz, ip, zz, cc))
        table[sym] = Element(sym, atno, nval, tuple(shells))
    return table


ELEMENTS = _build_table()


@dataclass
class Basis:
    """Struct-of-arrays form of the reference's STO_basis record (types_EHT.f:114-166)."""
    n_atoms: int
    first_orb: np.ndarray  # int32 [n_atoms+1]
    atom_lmax: np.ndarray  # int32 [n_atoms]
    n: np.ndarray  # int32 [N]
    l: np.ndarray
    m: np.ndarray
    nzeta: np.ndarray
    zeta: np.ndarray  # float64 [N, 2]
    coef: np.ndarray  # float64 [N, 2]
    ip: np.ndarray  # float64 [N]
    k_wh: np.ndarray
    v_shift: np.ndarray
    atom: np.ndarray  # int32 [N] owning atom
    fragment: np.ndarray  # int32 [N] fragment index of the owning atom

    @property
    def N(self) -> int:
        return int(self.n.shape[0])


def build_basis(symbols, fragments=None, k_wh: float = K_WH, v_shift=None) -> Basis:
    """Basis_Builder (structure.f:228-324): per atom, shells by l, m = -l..l."""
    n_atoms = len(symbols)
    if fragments is None:
        fragments = np.zeros(n_atoms, dtype=np.int32)
    if v_shift is None:
        v_shift = np.zeros(n_atoms)
    first = np.zeros(n_atoms + 1, dtype=np.int32)
    lmax = np.zeros(n_atoms, dtype=np.int32)
    rows = []
    for ia, sym in enumerate(symbols):
        el = ELEMENTS[sym]
        first[ia] = len(rows)
        lmax[ia] = el.ang_max
        for sh in el.shells:
            for m in range(-sh.l, sh.l + 1):
                rows.append((sh.n, sh.l, m, sh.nzeta, sh.zeta, sh.coef, sh.ip, ia))
    first[n_atoms] = len(rows)
    N = len(rows)
    b = Basis(
        n_atoms=n_atoms,
        first_orb=first,
        atom_lmax=lmax,
        n=np.array([r[0] for r in rows], dtype=np.int32),
        l=np.array([r[1] for r in rows], dtype=np.int32),
        m=np.array([r[2] for r in rows], dtype=np.int32),
        nzeta=np.array([r[3] for r in rows], dtype=np.int32),
        zeta=np.array([r[4] for r in rows], dtype=np.float64).reshape(N, 2),
        coef=np.array([r[5] for r in rows], dtype=np.float64).reshape(N, 2),
        ip=np.array([r[6] for r in rows], dtype=np.float64),
        k_wh=np.full(N, k_wh, dtype=np.float64),
        v_shift=np.array([v_shift[r[7]] for r in rows], dtype=np.float64),
        atom=np.array([r[7] for r in rows], dtype=np.int32),
        fragment=np.array([fragments[r[7]] for r in rows], dtype=np.int32),
    )
    return b


def n_valence(symbols) -> int:
    return sum(ELEMENTS[s].nvalen for s in symbols)
