"""Writes and reads the text decks of the Fortran replay harness (dh_harness).

Deck: `n_atoms N n_frag n_steps delta_t rebuild_every`, then per atom
`x y z fragment(1-based)`, per orbital `atom(1-based) n l m Nzeta zeta1 zeta2
coef1 coef2 IP k_WH V_shift`, then the electron and hole AO coefficients
(re im per line), then for every rebuild a per-atom displacement block.
"""
from __future__ import annotations

import numpy as np


def _g(x) -> str:
    """Shortest round-trip decimal form of a float (17 significant digits)."""
    return f"{float(x):.17g}"


def write_deck(path, st, basis, c_el, c_hl, n_steps, delta_t, rebuild_every=0, displacements=()):
    b = basis
    with open(path, "w") as f:
        f.write(f"{st.n_atoms} {b.N} {len(st.fragment_names)} {n_steps} {_g(delta_t)} {rebuild_every}\n")
        for x, fr in zip(st.coords, st.fragments):
            f.write(f"{_g(x[0])} {_g(x[1])} {_g(x[2])} {int(fr) + 1}\n")
        for i in range(b.N):
            f.write(f"{b.atom[i] + 1} {b.n[i]} {b.l[i]} {b.m[i]} {b.nzeta[i]} {_g(b.zeta[i, 0])} {_g(b.zeta[i, 1])} "
                    f"{_g(b.coef[i, 0])} {_g(b.coef[i, 1])} {_g(b.ip[i])} {_g(b.k_wh[i])} {_g(b.v_shift[i])}\n")
        for c in (c_el, c_hl):
            for v in np.asarray(c, dtype=np.complex128):
                f.write(f"{_g(v.real)} {_g(v.imag)}\n")
        for d in displacements:
            for row in d:
                f.write(f"{_g(row[0])} {_g(row[1])} {_g(row[2])}\n")


def read_output(path, N, n_frag, n_steps):
    with open(path) as f:
        rows = [np.array([float(v) for v in line.split()]) for line in f if line.strip()]
    nd = n_steps + 1
    dyn = np.stack(rows[:nd]).reshape(nd, 2, n_frag + 2)  # [it][particle][t, frags..., total]
    psi = np.stack(rows[nd:nd + 2 * N]).reshape(2, N, 4)
    ket = psi[:, :, 0] + 1j * psi[:, :, 1]
    bra = psi[:, :, 2] + 1j * psi[:, :, 3]
    erg = rows[nd + 2 * N]
    kern = rows[nd + 2 * N + 1:nd + 3 * N + 1]  # Ehrenfest kernel, column i on line i
    K = np.stack(kern).T if len(kern) == N else None
    return dyn, bra, ket, np.array([erg[0] + 1j * erg[1], erg[2] + 1j * erg[3]]), K
