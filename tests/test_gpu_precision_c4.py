"""The precision workload's shape on the device: N = 8192 (the 2399-atom C/H/O
interface cluster), 64 wavepackets (128 right-hand sides) through the split-K
matrix-core term, a few steps.  FP64 against the exact propagator on sampled
columns, from all eigenpairs of the device operators (the divide-and-conquer
eigensolver, itself checked against LAPACK in test_gpu_eigen.py); both mixed
modes against FP64."""
import numpy as np
import pytest

from dynemol_amd import hip
from dynemol_amd.structures import interface_cluster

pytestmark = pytest.mark.gpu

DT = 5.0e-4


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def test_precision_workload_N8192_64_wavepackets():
    st = interface_cluster(seed=4)
    b = st.basis()
    N, n_wp, nsteps = b.N, 64, 3
    _say(f"N={N}, {n_wp} wavepackets, {nsteps} steps: FP64 vs exact, mixed modes vs FP64")
    rng = np.random.default_rng(4)
    out = {}
    with hip.Context(0) as ctx:
        for mode in (hip.DH_PREC_FP64, hip.DH_PREC_MIXED_CORRECTED, hip.DH_PREC_MIXED):
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=len(st.fragment_names))
            s.set_precision(mode)
            s.build(st.coords)
            if mode == hip.DH_PREC_FP64:
                S = s.matrix(hip.DH_MAT_S)
                K = rng.normal(size=(N, n_wp)) + 1j * rng.normal(size=(N, n_wp))
                SK = S @ np.ascontiguousarray(K.real) + 1j * (S @ np.ascontiguousarray(K.imag))
                nrm = np.sqrt(np.real(np.sum(np.conj(K) * SK, axis=0)))
                K /= nrm
                SK /= nrm
            for w in range(n_wp):
                s.set_wavepacket(w, SK[:, w], K[:, w])
            pops, _, k = s.propagate_steps(nsteps, DT)
            _say(f"  mode {mode}: {k} terms per step, variant {s.tuning()['term']}")
            kets = {w: s.wavepacket(w)[1] for w in (0, 37)}
            if mode == hip.DH_PREC_FP64:
                E, C = s.eigenpairs(1, N)
            out[mode] = (pops, kets)
            s.close()
    for w in (0, 37):
        ke = C @ (np.exp(-1j * E * nsteps * DT / hip.H_BAR) * (C.T @ SK[:, w]))
        k64 = out[hip.DH_PREC_FP64][1][w]
        err = np.abs(k64 - ke).max() / np.abs(ke).max()
        _say(f"  FP64 wavepacket {w}: |dPsi|/|Psi| = {err:.2e} from the exact propagator")
        assert err <= 1e-9
    p64 = out[hip.DH_PREC_FP64][0]
    e3 = np.abs(out[hip.DH_PREC_MIXED_CORRECTED][0] - p64).max()
    e1 = np.abs(out[hip.DH_PREC_MIXED][0] - p64).max()
    _say(f"  max population difference from FP64 after {nsteps} steps: corrected {e3:.2e}, mixed {e1:.2e}")
    # a few steps in: the random FP32 rounding of the products dominates both
    # modes (the operator's rounding, which the correction removes, grows
    # linearly and shows over hundreds of steps: the 1000-step study)
    assert e3 <= 1e-7 and e1 <= 1e-7
