"""GPU parity of the FP64 drop-in path against the CPU oracle (config 1: 128-atom C/H, N = 416).

Tolerances: S and h bit-for-bit up to rounding (1e-13 absolute); populations
<= 1e-10 absolute; Psi <= 1e-9 relative (max-norm) -- the FP64 parity targets.
The oracle's Taylor propagator runs at tight tolerance (element 1e-13, norm
1e-11); its own distance to the exact propagator is checked in test_oracle.py.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 5.0e-4  # 0.5 fs in ps


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c1():
    st = ch_cluster(seed=1)
    b = st.basis()
    S = O.overlap(st.coords, b)
    h = O.huckel(b, S)
    E, C = sl.eigh(h, S)
    return st, b, S, h, E, C


def exact(E, C, S, psi0, t):
    return C @ (np.exp(-1j * E * t / O.H_BAR) * (C.T @ (S @ psi0)))


def test_build_SH_matches_oracle(ctx, c1):
    st, b, S, h, E, C = c1
    _say("build S,h c1 N=416")
    sysm = hip.System(ctx, b, n_wp=2, n_frag=2)
    sysm.build(st.coords)
    Sg = sysm.matrix(hip.DH_MAT_S)
    hg = sysm.matrix(hip.DH_MAT_H)
    assert np.abs(Sg - S).max() < 1e-13
    assert np.abs(hg - h).max() < 1e-12
    sysm.close()


def test_orthogonalisation(ctx, c1):
    st, b, S, h, E, C = c1
    _say("cholesky/H_orth c1")
    sysm = hip.System(ctx, b, n_wp=1, n_frag=2)
    sysm.build(st.coords)
    L = np.tril(sysm.matrix(hip.DH_MAT_L))
    Li = np.tril(sysm.matrix(hip.DH_MAT_LINV))
    Ho = sysm.matrix(hip.DH_MAT_HORTH)
    assert np.abs(L @ L.T - S).max() < 1e-13
    assert np.abs(Li @ L - np.eye(b.N)).max() < 1e-12
    assert np.abs(Ho - Ho.T).max() == 0.0
    w = np.linalg.eigvalsh(Ho)
    assert np.abs(w - E).max() < 1e-10 * max(1.0, np.abs(E).max())
    bnd = sysm.bounds()
    assert bnd[0] <= E.min() and bnd[1] >= E.max(), (bnd, E.min(), E.max())
    # the enclosure is not loose: within 10% of the width on each side
    width = E.max() - E.min()
    assert E.min() - bnd[0] < 0.1 * width and bnd[1] - E.max() < 0.1 * width
    sysm.close()


def test_one_step_vs_exact(ctx, c1):
    st, b, S, h, E, C = c1
    _say("one step c1 vs exact")
    sysm = hip.System(ctx, b, n_wp=2, n_frag=2)
    sysm.build(st.coords)
    lumo, homo = C[:, 208], C[:, 207]
    for w, c in enumerate((lumo, homo)):
        sysm.set_wavepacket(w, S @ c + 0j, c + 0j)
    k = sysm.propagate(0.0, DT)
    assert 5 < k < 200
    for w, c in enumerate((lumo, homo)):
        bra, ket = sysm.wavepacket(w)
        ke = exact(E, C, S, c + 0j, DT)
        be = S @ ke
        assert np.abs(ket - ke).max() <= 1e-9 * np.abs(ke).max()
        assert np.abs(bra - be).max() <= 1e-9 * np.abs(be).max()
        ao, pops, erg = sysm.observables(w)
        assert np.abs(ao - np.linalg.solve(S, bra)).max() <= 1e-9 * np.abs(ao).max()
        pe = O.populations(b.fragment, 2, np.conj(ke), be)
        assert np.abs(pops - pe).max() < 1e-10
        # <H> of an eigenstate is its eigenvalue
        assert abs(erg - E[208 if w == 0 else 207]) < 1e-9
    sysm.close()


def test_many_steps_vs_taylor_oracle(ctx, c1):
    st, b, S, h, E, C = c1
    nsteps = 20
    _say(f"{nsteps} steps c1 vs Taylor oracle (tight)")
    rng = np.random.default_rng(7)
    x = rng.normal(size=b.N) + 1j * rng.normal(size=b.N)
    x /= np.sqrt(np.real(np.conj(x) @ S @ x))
    kets0 = [C[:, 208] + 0j, x]
    sysm = hip.System(ctx, b, n_wp=2, n_frag=2)
    sysm.build(st.coords)
    for w, k0 in enumerate(kets0):
        sysm.set_wavepacket(w, S @ k0, k0)
    ref = O.ElHlReference(b, 2, np.stack([S @ k for k in kets0], 1), np.stack(kets0, 1), err_tol=1e-13,
                          norm_tol=1e-11)
    t = 0.0
    for it in range(2, nsteps + 2):
        sysm.propagate(t, DT * (it - 1))
        t, out = ref.step(st.coords, t, DT, it, S=S, h=h)
    for w in range(2):
        bra, ket = sysm.wavepacket(w)
        assert np.abs(ket - ref.psi_ket[:, w]).max() <= 1e-9 * np.abs(ket).max()
        assert np.abs(bra - ref.psi_bra[:, w]).max() <= 1e-9 * np.abs(bra).max()
        _, pops, erg = sysm.observables(w)
        assert np.abs(pops - out["pops"][:, w]).max() < 1e-10
        assert abs(erg - out["erg"][w]) < 1e-9
    sysm.close()


def test_dropin_propagationelhl(ctx, c1):
    st, b, S, h, E, C = c1
    _say("drop-in propagationelhl c1")
    c = C[:, 207] + 0j
    bra, ket, ao, Hp, tau, save_tau = ctx.propagation_elhl(S, h, S @ c, c, 0.0, DT)
    Hp_ref = np.linalg.solve(S, h)
    assert np.abs(Hp - Hp_ref).max() <= 1e-10 * np.abs(Hp_ref).max()
    ke = exact(E, C, S, c, DT)
    assert np.abs(ket - ke).max() <= 1e-9 * np.abs(ke).max()
    assert np.abs(bra - S @ ke).max() <= 1e-9 * np.abs(ke).max()
    assert np.abs(ao - ke).max() <= 1e-9 * np.abs(ke).max()  # S^-1 bra = ket here
    assert abs(tau - DT / O.H_BAR) < 1e-12 and abs(save_tau - DT / O.H_BAR) < 1e-12
