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


def test_mixed_precision_holds_the_bar_over_1000_steps(c1):
    """1000 steps of 0.5 fs, 4 wavepackets: the corrected mixed mode (FP32
    products + the bf16-stored FP64 residual of the FP32 operator) stays within
    the 1e-6 population bar of FP64; the plain mixed mode is reported beside it
    and must at least hold 1e-5.  (On this 416-orbital system the random FP32
    rounding, not the rounding of the operator, sets both errors, so the
    correction's gain shows on the 8192-orbital workload instead.)  The pre-renormalisation norm drift of one step
    stays at FP32 level (the renormalisation cannot hide a larger error)."""
    st, b, S, h, E, C = c1
    n_wp, nsteps = 4, 1000
    _say(f"mixed modes vs FP64, {n_wp} wavepackets, {nsteps} steps")
    kets = _states(C, S, n_wp, 2)
    out = {}
    with hip.Context(0) as ctx:
        for mode in (hip.DH_PREC_FP64, hip.DH_PREC_MIXED, hip.DH_PREC_MIXED_CORRECTED):
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
            s.set_precision(mode)
            s.build(st.coords)
            for w, k in enumerate(kets):
                s.set_wavepacket(w, S @ k, k)
            pops, erg, _ = s.propagate_steps(nsteps, DT)
            out[mode] = (pops, [s.wavepacket(w) for w in range(n_wp)], s.stats()["norm_drift"])
            s.close()
    p64 = out[hip.DH_PREC_FP64][0]
    err1 = np.abs(out[hip.DH_PREC_MIXED][0] - p64).max(axis=(1, 2))
    err3 = np.abs(out[hip.DH_PREC_MIXED_CORRECTED][0] - p64).max(axis=(1, 2))
    _say(f"  max population difference after 100/1000 steps: mixed {err1[99]:.2e}/{err1[-1]:.2e}, "
         f"corrected {err3[99]:.2e}/{err3[-1]:.2e}; last-step drift mixed {out[1][2]:.1e}, corrected {out[3][2]:.1e}")
    assert err3[-1] <= 1e-6
    assert err1[-1] <= 1e-5
    assert err3[-1] > 0.0 and err1[-1] > 0.0  # the FP32 paths really ran
    assert out[hip.DH_PREC_MIXED][2] < 1e-5 and out[hip.DH_PREC_MIXED_CORRECTED][2] < 1e-5
    # the final kets against the exact propagator: FP64 at its bar; the FP32
    # paths carry a phase error that grows with time (populations do not see it)
    for w, k in enumerate(kets):
        ke = C @ (np.exp(-1j * E * nsteps * DT / O.H_BAR) * (C.T @ (S @ k)))
        _, k64 = out[hip.DH_PREC_FP64][1][w]
        _, k3 = out[hip.DH_PREC_MIXED_CORRECTED][1][w]
        assert np.abs(k64 - ke).max() <= 1e-9 * np.abs(ke).max()
        e3 = np.abs(k3 - ke).max() / np.abs(ke).max()
        _say(f"  wavepacket {w}: corrected mixed |dPsi|/|Psi| = {e3:.1e} after {nsteps} steps")
        assert e3 <= 1e-3


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
