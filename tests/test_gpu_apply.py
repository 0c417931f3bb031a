"""dh_chebyshev_apply: propagation of caller-owned device operands.

The operator is the orthogonalised H_orth of a built system (TiO2 cluster,
N = 514, and the 128-atom C/H cluster, N = 416), copied into a caller-owned
device buffer with the identity padding; the state is R complex columns.  Bar:
phi after one 0.5 fs step within 1e-9 (relative, max-norm) of the exact
exp(-i H_orth dt / hbar) phi from eigh(H_orth), FP64 parity as for the system
path; the same term count as dh_propagate on the same operator and state, and
its AO-basis result within the same bar.
"""
import numpy as np
import pytest

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster, tio2_cluster

pytestmark = pytest.mark.gpu

DT = 5.0e-4


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def _operator(ctx, st):
    b = st.basis()
    s = hip.System(ctx, b, n_wp=1, n_frag=len(st.fragment_names))
    s.build(st.coords)
    Ho = s.matrix(hip.DH_MAT_HORTH)
    bounds = s.bounds()[:2].copy()
    s.close()
    return Ho, bounds


def _padded(Ho):
    N = Ho.shape[0]
    ld = (N + 63) // 64 * 64
    Hp = np.eye(ld)
    Hp[:N, :N] = Ho
    return np.asfortranarray(Hp), ld


def _exact(Ho, phi, dt):
    w, V = np.linalg.eigh(Ho)
    return V @ (np.exp(-1j * w * dt / hip.H_BAR)[:, None] * (V.T @ phi))


@pytest.mark.parametrize("which,R", [("tio2", 4), ("tio2", 1), ("ch", 2), ("ch", 8)])
def test_apply_matches_exact_propagator(ctx, which, R):
    st = tio2_cluster(seed=6) if which == "tio2" else ch_cluster(seed=1)
    Ho, bounds = _operator(ctx, st)
    N = Ho.shape[0]
    Hp, ld = _padded(Ho)
    _say(f"apply {which} N={N} R={R}")
    rng = np.random.default_rng(R)
    phi = np.zeros((ld, R), dtype=np.complex128)
    phi[:N] = rng.normal(size=(N, R)) + 1j * rng.normal(size=(N, R))
    phi[:N] /= np.linalg.norm(phi[:N], axis=0)
    phi[N:] = 7.0  # padding rows are the library's to clear
    wb = hip.chebyshev_work_bytes(N, R)
    Hd, Pd, Wd = hip.DeviceBuffer(Hp.nbytes), hip.DeviceBuffer(phi.nbytes), hip.DeviceBuffer(wb)
    Hd.upload(Hp)
    Pd.upload(phi)
    k = hip.chebyshev_apply(ctx, N, Hd.ptr, ld, bounds, DT, R, Pd.ptr, Wd.ptr, wb)
    out = Pd.download((ld, R), np.complex128)
    ref = _exact(Ho, phi[:N], DT)
    err = np.abs(out[:N] - ref).max() / np.abs(ref).max()
    _say(f"  {k} terms, |phi - exact| / |phi| = {err:.2e}")
    assert err <= 1e-9
    assert np.all(out[N:] == 0)
    # too small a work buffer is refused
    with pytest.raises(hip.DHError):
        hip.chebyshev_apply(ctx, N, Hd.ptr, ld, bounds, DT, R, Pd.ptr, Wd.ptr, wb - 256)
    for buf in (Hd, Pd, Wd):
        buf.close()


def test_apply_equals_system_step(ctx):
    """Same operator, same state: the apply entry point and dh_propagate run the
    same term kernels and the same term count."""
    st = tio2_cluster(seed=6)
    b = st.basis()
    N = b.N
    s = hip.System(ctx, b, n_wp=2, n_frag=len(st.fragment_names))
    s.build(st.coords)
    Ho = s.matrix(hip.DH_MAT_HORTH)
    L = np.tril(s.matrix(hip.DH_MAT_L))
    Li = np.tril(s.matrix(hip.DH_MAT_LINV))
    bounds = s.bounds()[:2].copy()
    rng = np.random.default_rng(3)
    kets = rng.normal(size=(N, 2)) + 1j * rng.normal(size=(N, 2))
    S = s.matrix(hip.DH_MAT_S)
    bras = S @ kets
    for w in range(2):
        s.set_wavepacket(w, bras[:, w], kets[:, w])
    k_sys = s.propagate(0.0, DT)
    # the system's columns: [L^-1 bra_0, L^-1 bra_1, L^T ket_0, L^T ket_1]
    phi0 = np.concatenate([Li @ bras, L.T @ kets], axis=1)
    Hp, ld = _padded(Ho)
    phi = np.zeros((ld, 4), dtype=np.complex128)
    phi[:N] = phi0
    wb = hip.chebyshev_work_bytes(N, 4)
    Hd, Pd, Wd = hip.DeviceBuffer(Hp.nbytes), hip.DeviceBuffer(phi.nbytes), hip.DeviceBuffer(wb)
    Hd.upload(Hp)
    Pd.upload(phi)
    k = hip.chebyshev_apply(ctx, N, Hd.ptr, ld, bounds, DT, 4, Pd.ptr, Wd.ptr, wb)
    out = Pd.download((ld, 4), np.complex128)[:N]
    _say(f"apply vs dh_propagate N={N}: {k} vs {k_sys} terms")
    assert k == k_sys
    ref = _exact(Ho, phi0, DT)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()
    # back to the AO basis as the system does it: ket = L^-T phi_ket, bra = L phi_bra
    for w in range(2):
        bra, ket = s.wavepacket(w)
        assert np.abs(ket - Li.T @ out[:, 2 + w]).max() <= 1e-9 * np.abs(ket).max()
        assert np.abs(bra - L @ out[:, w]).max() <= 1e-9 * np.abs(bra).max()
    for buf in (Hd, Pd, Wd):
        buf.close()
    s.close()
