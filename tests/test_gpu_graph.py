"""The captured hipGraph term series replays to the same bits as separate launches."""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.mark.parametrize("n,n_wp", [(300, 2), (700, 3)])
def test_graph_replays_match_eager_launches(n, n_wp):
    rng = np.random.default_rng(n)
    _say(f"graph vs eager N={n} n_wp={n_wp}")
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    A = rng.normal(size=(n, n))
    h = (A + A.T) * 2.0
    E, C = sl.eigh(h, S)
    kets = [C[:, n // 2 - 4:n // 2 + 4] @ (rng.normal(size=8) + 1j * rng.normal(size=8)) for _ in range(n_wp)]
    dt = 5e-4
    with hip.Context(0) as ctx:
        systems = []
        for graph in (False, True):
            s = hip.System(ctx, None, n_wp=n_wp, n_frag=1, n_orb=n)
            s.load_operators(S, h)
            s.set_graph(graph)
            for w in range(n_wp):
                s.set_wavepacket(w, S @ kets[w], kets[w])
            systems.append(s)
        eager, graphed = systems
        t = 0.0
        for step in range(8):
            # eager work and host synchronisation between replays
            eager.propagate(t, t + dt)
            graphed.propagate(t, t + dt)
            t += dt
            for w in range(n_wp):
                be, ke = eager.wavepacket(w)
                bg, kg = graphed.wavepacket(w)
                assert np.array_equal(ke, kg) and np.array_equal(be, bg), (step, w)
                _, pe, ee = eager.observables(w)
                _, pg, eg = graphed.observables(w)
                assert np.array_equal(pe, pg) and np.array_equal(ee, eg)
        # and both follow the exact propagator
        for w in range(n_wp):
            _, k = graphed.wavepacket(w)
            ke = C @ (np.exp(-1j * E * t / hip.H_BAR) * (C.T @ (S @ kets[w])))
            assert np.abs(k - ke).max() <= 1e-9 * np.abs(ke).max()
        for s in systems:
            s.close()
