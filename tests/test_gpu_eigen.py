"""Initial-state eigenpairs of (h, S) on the GPU (config 1: 128-atom C/H cluster, N = 416).

Pinned by what the eigenproblem fixes, independent of each column's sign:
eigenvalues against scipy.linalg.eigh(h, S) (1e-10 eV); residual
|h C - S C E| <= 1e-10 |h|; C^T S C = I to 1e-10; vectors equal to +-scipy's
where the eigenvalue is separated from its neighbours by more than 1e-3 eV
(1e-8); a seeded eigenstate only gains the phase e^{-iE t/hbar} over a step
(1e-9) and its <H> is E.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 5.0e-4


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


@pytest.mark.parametrize("il,iu", [(200, 215), (1, 3), (414, 416), (209, 209)])
def test_eigenpairs_match_scipy(ctx, c1, il, iu):
    st, b, S, h, E, C = c1
    _say(f"eigenpairs {il}..{iu} c1 N=416")
    sysm = hip.System(ctx, b, n_wp=2, n_frag=2)
    sysm.build(st.coords)
    Eg, Cg = sysm.eigenpairs(il, iu)
    sysm.close()
    ref = E[il - 1:iu]
    assert np.abs(Eg - ref).max() < 1e-10
    assert np.abs(h @ Cg - S @ Cg * Eg).max() <= 1e-10 * np.abs(h).max()
    assert np.abs(Cg.T @ S @ Cg - np.eye(iu - il + 1)).max() < 1e-10
    gaps = np.minimum(np.diff(E, prepend=-np.inf), np.diff(E, append=np.inf))[il - 1:iu]
    for c, g in enumerate(gaps):
        if g > 1e-3:
            r = C[:, il - 1 + c]
            sgn = np.sign(r @ S @ Cg[:, c])
            assert np.abs(Cg[:, c] - sgn * r).max() < 1e-8


def test_eigenstate_seeds_a_stationary_wavepacket(ctx, c1):
    st, b, S, h, E, C = c1
    _say("set_eigenstate LUMO/HOMO c1, one step")
    sysm = hip.System(ctx, b, n_wp=2, n_frag=2)
    sysm.build(st.coords)
    e_lumo = sysm.set_eigenstate(0, 209)
    e_homo = sysm.set_eigenstate(1, 208)
    assert abs(e_lumo - E[208]) < 1e-10 and abs(e_homo - E[207]) < 1e-10
    start = [sysm.wavepacket(w)[1] for w in range(2)]
    sysm.propagate(0.0, DT)
    for w, e in enumerate((e_lumo, e_homo)):
        _, ket = sysm.wavepacket(w)
        assert np.abs(ket - np.exp(-1j * e * DT / hip.H_BAR) * start[w]).max() <= 1e-9 * np.abs(ket).max()
        assert abs(sysm.observables(w)[2] - e) < 1e-9
    sysm.close()


def test_eigenpairs_rejects_bad_range(ctx, c1):
    st, b, S, h, E, C = c1
    sysm = hip.System(ctx, b, n_wp=1, n_frag=2)
    with pytest.raises(hip.DHError):
        sysm.eigenpairs(1, 2)  # operators not built yet
    sysm.build(st.coords)
    with pytest.raises(hip.DHError):
        sysm.eigenpairs(3, 2)
    with pytest.raises(hip.DHError):
        sysm.eigenpairs(1, b.N + 1)
    sysm.close()
