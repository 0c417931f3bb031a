"""Multi-RHS matrix-core term kernel (FP64) and the mixed-precision (FP32 MFMA) mode."""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DT = 5e-4


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def c1():
    st = ch_cluster(seed=1)
    b = st.basis()
    S = O.overlap(st.coords, b)
    h = O.huckel(b, S)
    E, C = sl.eigh(h, S)
    return st, b, S, h, E, C


def _states(C, S, n, seed):
    rng = np.random.default_rng(seed)
    kets = []
    for w in range(n):
        if w == 0:
            k = C[:, 208] + 0j
        elif w == 1:
            k = C[:, 207] + 0j
        else:
            k = C[:, 190:230] @ (rng.normal(size=40) + 1j * rng.normal(size=40))
            k /= np.sqrt(np.real(np.conj(k) @ S @ k))
        kets.append(k)
    return kets


def test_gemm_term_many_wavepackets_exact(c1):
    st, b, S, h, E, C = c1
    n_wp = 20  # 40 complex RHS = 80 real columns: two 64-column tiles
    _say(f"FP64 GEMM term kernel, {n_wp} wavepackets")
    kets = _states(C, S, n_wp, 1)
    hip.force_matvec_variant(15)
    try:
        with hip.Context(0) as ctx:
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
            s.build(st.coords)
            for w, k in enumerate(kets):
                s.set_wavepacket(w, S @ k, k)
            s.propagate(0.0, DT)
            for w, k in enumerate(kets):
                bra, ket = s.wavepacket(w)
                ke = C @ (np.exp(-1j * E * DT / O.H_BAR) * (C.T @ (S @ k)))
                assert np.abs(ket - ke).max() <= 1e-10 * np.abs(ke).max()
            s.close()
    finally:
        hip.force_matvec_variant(-1)


def test_mixed_precision_tracks_fp64(c1):
    st, b, S, h, E, C = c1
    n_wp, nsteps = 4, 20
    _say(f"mixed vs FP64, {n_wp} wavepackets, {nsteps} steps")
    kets = _states(C, S, n_wp, 2)
    out = {}
    with hip.Context(0) as ctx:
        for mixed in (False, True):
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
            s.set_precision(mixed)
            s.build(st.coords)
            for w, k in enumerate(kets):
                s.set_wavepacket(w, S @ k, k)
            t = 0.0
            for _ in range(nsteps):
                s.propagate(t, t + DT)
                t += DT
            out[mixed] = [(s.wavepacket(w), s.observables(w)[1]) for w in range(n_wp)]
            s.close()
    for w in range(n_wp):
        (b64, k64), p64 = out[False][w]
        (b32, k32), p32 = out[True][w]
        assert np.abs(k32 - k64).max() < 1e-4 * np.abs(k64).max()
        assert np.abs(p32 - p64).max() < 1e-4
        assert np.abs(p32 - p64).max() > 0.0  # the mixed path really ran in FP32


@pytest.mark.parametrize("n_wp", [2, 20])
def test_fp32_storage_is_exact_for_rounded_operator(c1, n_wp):
    """Mode 2 reads H_orth rounded to FP32 and does FP64 arithmetic: it must be
    the exact propagator of that rounded operator to FP64 accuracy.  n_wp = 2
    (4 RHS) runs the symmetric half-storage kernel, n_wp = 20 (40 RHS) the
    split-K GEMM with widened FP32 operands."""
    st, b, S, h, E, C = c1
    nsteps = 3
    _say(f"FP32-stored operator, {n_wp} wavepackets, {nsteps} steps")
    kets = _states(C, S, n_wp, 3)
    with hip.Context(0) as ctx:
        s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
        s.set_precision(hip.DH_PREC_F32_STORAGE)
        s.build(st.coords)
        L = np.tril(s.matrix(hip.DH_MAT_L))  # stored mirrored
        Hq = s.matrix(hip.DH_MAT_HORTH).astype(np.float32).astype(np.float64)
        for w, k in enumerate(kets):
            s.set_wavepacket(w, S @ k, k)
        t = 0.0
        for _ in range(nsteps):
            s.propagate(t, t + DT)
            t += DT
        out = [s.wavepacket(w) for w in range(n_wp)]
        s.close()
    assert np.abs(L @ L.T - S).max() < 1e-12
    lam, V = np.linalg.eigh(Hq)
    U = V @ (np.exp(-1j * lam * t / O.H_BAR)[:, None] * V.T)  # exp(-i Hq t / hbar)
    Uex = C @ (np.exp(-1j * E * t / O.H_BAR)[:, None] * (C.T @ S))  # unrounded operator, AO basis
    for w, k in enumerate(kets):
        bra, ket = out[w]
        ke = sl.solve_triangular(L, U @ (L.T @ k), trans="T", lower=True)  # L^-T U L^T k
        be = L @ (U @ sl.solve_triangular(L, S @ k, lower=True))  # L U L^-1 S k
        assert np.abs(ket - ke).max() <= 1e-10 * np.abs(ke).max()
        assert np.abs(bra - be).max() <= 1e-10 * np.abs(be).max()
        # and it is not the FP64 operator: rounding H_orth moves the result
        assert np.abs(ket - Uex @ k).max() > 1e-12 * np.abs(ke).max()
