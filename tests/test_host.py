"""Host-side checks that need no GPU: C-ABI exports, EHT table, basis order, geometries, PDB I/O."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest
from scipy.spatial.distance import pdist

from dynemol_amd import eht, hip
from dynemol_amd.structures import (ch_cluster, heavy_cluster, interface_cluster, read_pdb, tio2_slab_adsorbate,
                                    write_pdb)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("dynemol_hip.h", "dynemol_hip_testing.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:int|long long|const char \*)\s*(dh_\w+)\s*\(", text, re.M)
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(hip.LIB_PATH), "run python -m dynemol_amd.build"
    L = ctypes.CDLL(hip.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # the binding covers exactly what the headers declare
    assert set(names) == set(hip.SYMBOLS) | set(hip.TEST_SYMBOLS)


def test_eht_truncation_and_units():
    # zeta: int(zeta/a0 * 1e4) * 1e-4 (EHT_input.f:81-92)
    assert eht.ELEMENTS["C"].shells[0].zeta[0] == 3.0708
    assert eht.ELEMENTS["H"].shells[0].zeta[0] == pytest.approx(2.4566, abs=1e-15)
    ti_d = eht.ELEMENTS["Ti"].shells[2]
    assert (ti_d.n, ti_d.l, ti_d.nzeta) == (3, 2, 2)
    assert ti_d.zeta == pytest.approx((8.5982, 2.6456), abs=1e-12)
    assert ti_d.coef == pytest.approx((0.4206, 0.7839), abs=1e-12)
    assert [eht.ELEMENTS[s].dos for s in ("H", "C", "N", "O", "Ti")] == [1, 4, 4, 4, 9]


def test_basis_order():
    b = eht.build_basis(["Ti", "H"])
    # s, p(y,z,x) = m -1,0,1, d(xy,yz,z2,xz,x2-y2) = m -2..2 (structure.f:211-223)
    assert b.l.tolist() == [0, 1, 1, 1, 2, 2, 2, 2, 2, 0]
    assert b.m.tolist() == [0, -1, 0, 1, -2, -1, 0, 1, 2, 0]
    assert b.first_orb.tolist() == [0, 9, 10]


@pytest.mark.parametrize("gen,N", [(ch_cluster, 416), (tio2_slab_adsorbate, 4096)])
def test_generators(gen, N):
    st = gen()
    assert st.n_orbitals == N
    assert pdist(st.coords).min() > 1.0
    heavy = np.array([s != "H" for s in st.symbols])
    if gen is ch_cluster:
        assert heavy.sum() == 96 and (~heavy).sum() == 32
        assert pdist(st.coords[heavy]).min() >= 1.35
        assert st.coords[heavy].max(axis=0).max() - st.coords[heavy].min(axis=0).min() <= 14.0 + 1e-9


def test_interface_cluster_size():
    st = interface_cluster()
    assert st.n_orbitals == 8192 and st.n_atoms == 2399


def test_heavy_cluster_exact_N():
    for n in (256, 777, 2048):
        assert heavy_cluster(n, n).n_orbitals == n


def test_pdb_roundtrip(tmp_path):
    st = ch_cluster()
    p = tmp_path / "input.pdb"
    write_pdb(str(p), st)
    back = read_pdb(str(p))
    assert back.symbols == st.symbols
    assert np.abs(back.coords - st.coords).max() <= 5e-4
    lines = p.read_text().splitlines()
    assert lines[1].startswith("CRYST1") and lines[-2] == "MASTER"
    # babel.f:290 columns: x at 31-38, element at 77-78
    assert lines[2][76:78].strip() == "C"


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_mixed_footprint_counts_fp32_copies():
    """Serving admission counts the FP32 operator copies of a mixed-precision
    system, and in every mode the multi-step series' accumulators (10 of
    ld x 2 n_wp complex), their coefficient rows (10 x 4098 complex) and the
    16-slot ring of terms (ld x 2 n_wp complex each)."""
    from dynemol_amd import hip
    ld = 448
    base = int(hip.lib().dh_system_bytes(400, 2))
    multi = (10 * (ld * 4 + 4098) + 16 * ld * 4) * 16
    assert hip.system_bytes(400, 2, hip.DH_PREC_MIXED) - base == 4 * ld * ld + 2 * ld * 4 * 2 * 4 + multi
    assert hip.system_bytes(400, 2, hip.DH_PREC_F32_STORAGE) - base == 4 * ld * ld + multi
    assert hip.system_bytes(400, 2) - base == multi
