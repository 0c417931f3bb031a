"""Multi-system batches: uneven-N systems advanced by grouped launches == the same systems alone."""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.structures import heavy_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DT = 5e-4


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _system_inputs(seed, N, n_wp=4):
    st = heavy_cluster(seed, N)
    b = st.basis()
    S = O.overlap(st.coords, b)
    h = O.huckel(b, S)
    E, C = sl.eigh(h, S)
    rng = np.random.default_rng(seed)
    nel = sum({"C": 4, "N": 5, "O": 6, "H": 1}[s] for s in st.symbols)
    homo = nel // 2 - 1
    kets = []
    for _ in range(n_wp):  # random unit-norm superpositions of 8 frontier orbitals
        cc = rng.normal(size=8) + 1j * rng.normal(size=8)
        k = C[:, homo - 3:homo + 5] @ cc
        kets.append(k / np.sqrt(np.real(np.conj(k) @ S @ k)))
    return st, b, S, E, C, kets


@pytest.mark.parametrize("n_wp", [4, 8])
def test_batch_matches_single_and_exact(n_wp):
    """4 wavepackets: one symmetric pass (R = 8); 8 wavepackets: two column halves (R = 16)."""
    sizes = [300, 517, 256, 800, 411, 1000]
    _say(f"batch of {len(sizes)} systems, N={sizes}, n_wp={n_wp}")
    data = [_system_inputs(100 + i, n, n_wp) for i, n in enumerate(sizes)]
    with hip.Context(0) as ctx:
        batched, single = [], []
        for st, b, S, E, C, kets in data:
            for group in (batched, single):
                s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
                s.build(st.coords)
                for w, k in enumerate(kets):
                    s.set_wavepacket(w, S @ k, k)
                group.append(s)
        batch = hip.Batch(ctx, batched)
        t = 0.0
        for _ in range(3):
            batch.propagate(t, t + DT)
            for s in single:
                s.propagate(t, t + DT)
            t += DT
        for (st, b, S, E, C, kets), sb, ss in zip(data, batched, single):
            for w in range(n_wp):
                bb, kb = sb.wavepacket(w)
                bs, ks = ss.wavepacket(w)
                assert np.abs(kb - ks).max() <= 1e-11 * np.abs(ks).max()
                assert np.abs(bb - bs).max() <= 1e-11 * np.abs(bs).max()
                ke = C @ (np.exp(-1j * E * t / O.H_BAR) * (C.T @ (S @ kets[w])))
                assert np.abs(kb - ke).max() <= 1e-9 * np.abs(ke).max()
                _, pb, eb = sb.observables(w)
                _, ps, es = ss.observables(w)
                assert np.abs(pb - ps).max() < 1e-12
                assert abs(eb - es) < 1e-10
        batch.close()


def test_batch_graph_replays_match_launches():
    """The captured term series (default) against individual launches, bitwise, over
    steps that replay one graph, with host reads between the replays, and a step of
    another length that captures a new one."""
    sizes = [300, 517, 256, 700]
    _say(f"batch graph vs launches, N={sizes}")
    data = [_system_inputs(200 + i, n) for i, n in enumerate(sizes)]
    with hip.Context(0) as ctx:
        groups = {True: [], False: []}
        for st, b, S, E, C, kets in data:
            for on in (True, False):
                s = hip.System(ctx, b, n_wp=4, n_frag=2)
                s.build(st.coords)
                for w, k in enumerate(kets):
                    s.set_wavepacket(w, S @ k, k)
                groups[on].append(s)
        batches = {on: hip.Batch(ctx, groups[on]) for on in (True, False)}
        batches[False].set_graph(False)
        t = 0.0
        for dt in (DT, DT, DT, 0.5 * DT, DT):
            for on in (True, False):
                batches[on].propagate(t, t + dt)
            t += dt
            for sg, sl_ in zip(groups[True], groups[False]):
                for w in range(4):
                    bg, kg = sg.wavepacket(w)
                    bl, kl = sl_.wavepacket(w)
                    assert np.array_equal(kg, kl) and np.array_equal(bg, bl)
                    assert np.array_equal(sg.observables(w)[1], sl_.observables(w)[1])
        for on in (True, False):
            batches[on].close()
