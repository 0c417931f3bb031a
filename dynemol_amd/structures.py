"""Seeded synthetic geometries for the propagator workloads, plus PDB I/O.

The PDB writer emits the column layout the reference reader expects
(babel.f:184-262, format 115 at babel.f:290: name t12 a5, residue t18 a3,
residue number t23 i7, x/y/z t31/t39/t47 f8.3, element t77 a2), with a header
line, a CRYST1 line and a closing MASTER line.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .eht import ELEMENTS, build_basis, Basis


@dataclass
class Structure:
    symbols: list
    coords: np.ndarray  # [n_atoms, 3] Angstrom
    fragments: np.ndarray  # int32 [n_atoms]
    fragment_names: list = field(default_factory=lambda: ["#"])
    name: str = "system"

    @property
    def n_atoms(self) -> int:
        return len(self.symbols)

    def basis(self) -> Basis:
        return build_basis(self.symbols, self.fragments)

    @property
    def n_orbitals(self) -> int:
        return sum(ELEMENTS[s].dos for s in self.symbols)

    def jittered(self, rng: np.random.Generator, sigma: float) -> "Structure":
        return Structure(list(self.symbols), self.coords + rng.normal(0.0, sigma, self.coords.shape),
                         self.fragments.copy(), list(self.fragment_names), self.name)


def _place(rng, n, box, dmin, existing=None, center=None, max_tries=200000):
    """Uniform random points in a cube with a minimum pair distance (rejection)."""
    pts = [] if existing is None else list(existing)
    start = len(pts)
    lo = np.zeros(3) if center is None else np.asarray(center) - box / 2
    tries = 0
    while len(pts) - start < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms with the requested minimum distance")
        p = lo + rng.uniform(0.0, box, 3)
        if pts:
            d = np.linalg.norm(np.asarray(pts) - p, axis=1)
            if d.min() < dmin:
                continue
        pts.append(p)
    return np.asarray(pts[start:])


def _cap_hydrogens(rng, heavy, n_h, bond=1.09, dmin=1.0, max_tries=100000):
    """Cap n_h distinct heavy atoms with one H each, at `bond` Angstrom in a random direction."""
    hosts = rng.permutation(len(heavy))
    hs = []
    used = 0
    k = 0
    while used < n_h:
        if k >= len(hosts):
            raise RuntimeError("not enough heavy atoms to cap")
        host = hosts[k]
        k += 1
        for _ in range(max_tries):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            p = heavy[host] + bond * v
            dh = np.linalg.norm(heavy - p, axis=1)
            dh[host] = np.inf
            if dh.min() < 1.35:
                continue
            if hs and np.linalg.norm(np.asarray(hs) - p, axis=1).min() < dmin:
                continue
            hs.append(p)
            used += 1
            break
    return np.asarray(hs).reshape(-1, 3)


def ch_cluster(seed: int = 1, n_c: int = 96, n_h: int = 32, box: float = 14.0) -> Structure:
    """128-atom C/H cluster: 96 C in a 14 A box (C-C >= 1.35 A), 32 C capped by H at 1.09 A.

    N = 96*4 + 32 = 416 orbitals.  Fragments: 0 = C atoms, 1 = H atoms.
    """
    rng = np.random.default_rng(seed)
    c = _place(rng, n_c, box, 1.35)
    h = _cap_hydrogens(rng, c, n_h)
    coords = np.vstack([c, h])
    symbols = ["C"] * n_c + ["H"] * n_h
    frag = np.array([0] * n_c + [1] * n_h, dtype=np.int32)
    return Structure(symbols, coords, frag, ["C", "H"], name=f"CH{n_c}_{n_h}")


def heavy_cluster(seed: int, n_orbitals: int, elements=("C", "N", "O"), h_fraction: float = 0.2,
                  density: float = 0.035, rng=None) -> Structure:
    """Random C/H/N/O (or C/H/O) cluster with exactly `n_orbitals` orbitals.

    Heavy atoms carry 4 orbitals and H one, so 4*k + h = N with h ~ h_fraction*N
    and h <= k (each heavy atom is capped by at most one H).
    """
    rng = np.random.default_rng(seed) if rng is None else rng
    N = int(n_orbitals)
    h = int(round(h_fraction * N))
    h += (N - h) % 4
    k = (N - h) // 4
    while h > k:
        h -= 4
        k += 1
    box = (k / density) ** (1.0 / 3.0)
    heavy = _place(rng, k, box, 1.35)
    hs = _cap_hydrogens(rng, heavy, h)
    syms = list(rng.choice(list(elements), size=k))
    coords = np.vstack([heavy, hs])
    symbols = syms + ["H"] * h
    frag = np.array([0] * k + [1] * h, dtype=np.int32)
    s = Structure(symbols, coords, frag, ["X", "H"], name=f"cluster_N{N}")
    assert s.n_orbitals == N, (s.n_orbitals, N)
    return s


def _anatase_block(na, nb, nc):
    """Anatase TiO2 (I4_1/amd; a = 3.785, c = 9.514 A, u = 0.208) replicated na x nb x nc."""
    a, c, u = 3.785, 9.514, 0.208
    ti_frac = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [0, 0.5, 0.25], [0.5, 0, 0.75]])
    o_frac = []
    for t in ti_frac:
        o_frac.append(t + [0, 0, u])
        o_frac.append(t - [0, 0, u])
    o_frac = np.array(o_frac)
    cell = np.array([a, a, c])
    ti, o = [], []
    for i in range(na):
        for j in range(nb):
            for k in range(nc):
                shift = np.array([i, j, k])
                ti.extend((ti_frac + shift) * cell)
                o.extend((o_frac + shift) * cell)
    return np.array(ti), np.array(o)


def tio2_slab_adsorbate(seed: int = 2, n_tio2: int = 236) -> Structure:
    """Anatase-like TiO2 slab (236 TiO2 units) plus a C16H20 adsorbate: N = 236*17 + 84 = 4096.

    Fragments: 0 = slab (Ti, O), 1 = adsorbate (C, H).
    """
    rng = np.random.default_rng(seed)
    ti, o = _anatase_block(8, 8, 3)
    center = np.array([ti[:, 0].mean(), ti[:, 1].mean(), ti[:, 2].mean()])
    # flattened slab: weight z distance more so the cluster is plate-like
    w = np.array([1.0, 1.0, 2.5])
    d_ti = np.linalg.norm((ti - center) * w, axis=1)
    sel_ti = ti[np.argsort(d_ti, kind="stable")[:n_tio2]]
    d_o = np.min(np.linalg.norm(o[:, None, :] - sel_ti[None, :, :], axis=2), axis=1)
    d_o = d_o + 1e-3 * np.linalg.norm((o - center) * w, axis=1)
    sel_o = o[np.argsort(d_o, kind="stable")[:2 * n_tio2]]
    slab = np.vstack([sel_ti, sel_o])
    top = slab[:, 2].max()
    # adsorbate: 16-carbon zigzag chain 2.8 A above the slab top, 20 H caps
    c = []
    ang = np.deg2rad(120.0)
    bond = 1.40
    p = np.array([center[0] - 8.0, center[1], top + 2.8])
    dirs = [np.array([np.cos(ang / 2), np.sin(ang / 2), 0.0]), np.array([np.cos(ang / 2), -np.sin(ang / 2), 0.0])]
    for i in range(16):
        c.append(p.copy())
        p = p + bond * dirs[i % 2]
    c = np.array(c)
    c += rng.normal(0.0, 0.02, c.shape)
    hs = []
    for i in range(16):
        nh = 2 if i in (0, 15, 5, 10) else 1
        for _ in range(nh):
            for _t in range(100000):
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                q = c[i] + 1.09 * v
                if q[2] < top + 2.0:
                    continue
                dc = np.linalg.norm(c - q, axis=1)
                dc[i] = np.inf
                if dc.min() < 1.5:
                    continue
                if hs and np.linalg.norm(np.array(hs) - q, axis=1).min() < 1.4:
                    continue
                hs.append(q)
                break
    hs = np.array(hs)
    assert len(hs) == 20
    symbols = ["Ti"] * n_tio2 + ["O"] * (2 * n_tio2) + ["C"] * 16 + ["H"] * 20
    coords = np.vstack([slab, c, hs])
    frag = np.array([0] * (3 * n_tio2) + [1] * 36, dtype=np.int32)
    s = Structure(symbols, coords, frag, ["S", "M"], name="TiO2_slab_C16H20")
    assert s.n_orbitals == 4096
    return s


def _random_rotation(rng) -> np.ndarray:
    """Uniform random rotation matrix (QR of a Gaussian matrix, sign-fixed, det +1)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def tio2_cluster(seed: int = 6, n_tio2: int = 28, n_c: int = 6, n_n: int = 2, n_h: int = 6,
                 sigma: float = 0.05) -> Structure:
    """Small anatase TiO2 cluster in a random orientation with a C/N/H adsorbate layer.

    n_tio2 Ti nearest the block centre and the 2*n_tio2 O nearest them, rotated
    by a random rotation and jittered by sigma (Angstrom), so Ti-O, Ti-Ti and
    O-O pairs lie along general directions; C and N atoms (>= 1.35 A apart)
    sit 1.9-3.5 A above the cluster surface and n_h of them are capped by H.
    N = 17*n_tio2 + 4*(n_c + n_n) + n_h (28 units: 476 + 32 + 6 = 514).
    Fragments: 0 = Ti, 1 = O, 2 = C/N/H.
    """
    rng = np.random.default_rng(seed)
    ti, o = _anatase_block(4, 4, 3)
    center = ti.mean(axis=0)
    sel_ti = ti[np.argsort(np.linalg.norm(ti - center, axis=1), kind="stable")[:n_tio2]]
    d_o = np.min(np.linalg.norm(o[:, None, :] - sel_ti[None, :, :], axis=2), axis=1)
    d_o = d_o + 1e-3 * np.linalg.norm(o - center, axis=1)
    sel_o = o[np.argsort(d_o, kind="stable")[:2 * n_tio2]]
    core = np.vstack([sel_ti, sel_o]) - center
    core = core @ _random_rotation(rng).T
    core += rng.normal(0.0, sigma, core.shape)
    # adsorbate atoms on the cluster surface, at general positions
    heavy = []
    rmax = np.linalg.norm(core, axis=1).max()
    while len(heavy) < n_c + n_n:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        p = v * rng.uniform(0.5 * rmax, rmax + 3.0)
        dc = np.linalg.norm(core - p, axis=1).min()
        if not 1.9 <= dc <= 3.5:
            continue
        if heavy and np.linalg.norm(np.asarray(heavy) - p, axis=1).min() < 1.35:
            continue
        heavy.append(p)
    heavy = np.asarray(heavy)
    hs = []
    for i in range(n_h):
        for _ in range(100000):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            q = heavy[i] + 1.09 * v
            d = np.linalg.norm(np.vstack([core, np.delete(heavy, i, axis=0)]) - q, axis=1).min()
            if d < 1.5:
                continue
            if hs and np.linalg.norm(np.asarray(hs) - q, axis=1).min() < 1.4:
                continue
            hs.append(q)
            break
        else:
            raise RuntimeError("could not cap the adsorbate")
    symbols = ["Ti"] * n_tio2 + ["O"] * (2 * n_tio2) + ["C"] * n_c + ["N"] * n_n + ["H"] * n_h
    coords = np.vstack([core, heavy, np.asarray(hs).reshape(-1, 3)])
    frag = np.array([0] * n_tio2 + [1] * (2 * n_tio2) + [2] * (n_c + n_n + n_h), dtype=np.int32)
    return Structure(symbols, coords, frag, ["Ti", "O", "A"], name=f"TiO2_{n_tio2}_cluster")


def interface_cluster(seed: int = 4, n_heavy: int = 1931, n_h: int = 468) -> Structure:
    """C/H/O interface cluster with N = 4*1931 + 468 = 8192 orbitals (2399 atoms).

    2400 atoms cannot give N = 8192 exactly (4k + h = 8192 and k + h = 2400
    has no integer solution), so one H fewer is used.
    """
    rng = np.random.default_rng(seed)
    box = (n_heavy / 0.035) ** (1.0 / 3.0)
    heavy = _place(rng, n_heavy, box, 1.35)
    # two half-spaces: C-rich below, O-rich above (an interface)
    syms = []
    zmid = heavy[:, 2].mean()
    for p in heavy:
        if p[2] < zmid:
            syms.append("C" if rng.uniform() < 0.85 else "O")
        else:
            syms.append("O" if rng.uniform() < 0.55 else "C")
    hs = _cap_hydrogens(rng, heavy, n_h)
    symbols = syms + ["H"] * n_h
    frag = np.array([0 if p[2] < zmid else 1 for p in heavy] + [2] * n_h, dtype=np.int32)
    s = Structure(symbols, np.vstack([heavy, hs]), frag, ["L", "U", "H"], name="CHO_interface_N8192")
    assert s.n_orbitals == 8192
    return s


def log_uniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


# ----------------------------------------------------------------------------- PDB
def write_pdb(path, s: Structure, title="dynemol-hip synthetic system"):
    """Fixed columns of babel.f:290 (format 115) on HETATM records."""
    def put(buf, col, text):  # col is 1-based as in the Fortran T edit descriptor
        buf[col - 1:col - 1 + len(text)] = list(text)

    lines = [f"{title:<72.72}"]
    span = s.coords.max(axis=0) - s.coords.min(axis=0) + 10.0
    lines.append("CRYST1" + "".join(f"{v:9.3f}" for v in span) + "  90.00  90.00  90.00 P 1           1")
    for i, (sym, x) in enumerate(zip(s.symbols, s.coords)):
        res = s.fragment_names[s.fragments[i]][:3] if s.fragment_names else "UNK"
        buf = [" "] * 80
        put(buf, 1, "HETATM")
        put(buf, 7, f"{(i + 1) % 100000:5d}")
        put(buf, 12, f"{sym:<5.5}")
        put(buf, 18, f"{res:<3.3}")
        put(buf, 23, f"{1:7d}")
        put(buf, 31, f"{x[0]:8.3f}{x[1]:8.3f}{x[2]:8.3f}")
        put(buf, 55, "  1.00  0.00")
        put(buf, 77, f"{sym:>2}")
        lines.append("".join(buf).rstrip())
    lines.append("MASTER")
    lines.append("END")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_pdb(path) -> Structure:
    """Reads the fixed columns of babel.f:290 (format 115)."""
    with open(path) as f:
        rows = f.read().splitlines()
    symbols, coords = [], []
    it = iter(rows[1:])
    for line in it:
        if line[:6] == "CRYST1":
            continue
        if line[:6] == "MASTER":
            break
        coords.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
        symbols.append(line[76:78].strip())
    return Structure(symbols, np.array(coords), np.zeros(len(symbols), dtype=np.int32), ["#"], name=path)
