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


@pytest.mark.parametrize("family", [1, 0])
def test_batch_steps_match_single_steps(family):
    """dh_batch_propagate_steps (10 steps per recurrence for a symmetric-family
    batch; one at a time for the dense family) against each system stepped
    alone with dh_propagate, and against the exact propagator: 13 steps (the
    tuning step, a series of 10, a series of 2)."""
    sizes = [300, 517, 256, 800, 411, 1000]
    n_wp, steps = 4, 13
    _say(f"batch steps of {len(sizes)} systems, family {family}, {steps} steps")
    data = [_system_inputs(200 + i, n, n_wp) for i, n in enumerate(sizes)]
    hip.force_batch_family(2 * n_wp, family)
    try:
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
            kmax = batch.propagate_steps(steps, DT)
            assert batch.tuning() == ("symmetric" if family == 1 else "dense")
            for s in single:
                for _ in range(steps):
                    s.propagate(0.0, DT)
            _say(f"  most terms of one recurrence: {kmax}")
            for (st, b, S, E, C, kets), sb, ss in zip(data, batched, single):
                for w in range(n_wp):
                    bb, kb = sb.wavepacket(w)
                    bs, ks = ss.wavepacket(w)
                    assert np.abs(kb - ks).max() <= 1e-11 * np.abs(ks).max()
                    assert np.abs(bb - bs).max() <= 1e-11 * np.abs(bs).max()
                    ke = C @ (np.exp(-1j * E * steps * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
                    assert np.abs(kb - ke).max() <= 1e-9 * np.abs(ke).max()
                    _, pb, eb = sb.observables(w)
                    _, ps, es = ss.observables(w)
                    assert np.abs(pb - ps).max() < 1e-12
                    assert abs(eb - es) < 1e-10
            batch.close()
    finally:
        hip.force_batch_family(2 * n_wp, -1)


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


def test_batch_32_systems_up_to_2048_vs_exact():
    """32 C/H/N/O systems with N log-uniform in [256, 2048], 4 wavepackets each,
    10 steps in one batch (the many-system configuration at a tenth of its
    size); sampled systems against the exact propagator (FP64 bar)."""
    from dynemol_amd.structures import log_uniform_int
    rng = np.random.default_rng(33)
    sizes = [log_uniform_int(rng, 256, 2048) for _ in range(32)]
    sizes[0], sizes[1] = 2048, 256
    _say(f"batch of 32 systems, N in [{min(sizes)}, {max(sizes)}]")
    check = (0, 1, 7, 19)
    with hip.Context(0) as ctx:
        systems, kept = [], {}
        for i, N in enumerate(sizes):
            st = heavy_cluster(500 + i, N)
            b = st.basis()
            s = hip.System(ctx, b, n_wp=4, n_frag=2)
            s.build(st.coords)
            E, C = s.eigenpairs(N // 2 - 3, N // 2 + 4)
            cs = []
            for w in range(4):
                a = rng.normal(size=8) + 1j * rng.normal(size=8)
                c = C @ (a / np.linalg.norm(a))
                s.set_initial_state(w, c, c)
                cs.append(c)
            if i in check:
                kept[i] = (st, b, cs)
            systems.append(s)
        batch = hip.Batch(ctx, systems)
        for _ in range(10):
            batch.propagate(0.0, DT)
        _say(f"  batch kernel: {batch.tuning()}")
        for i, (st, b, cs) in kept.items():
            S = O.overlap(st.coords, b)
            h = O.huckel(b, S)
            E, C = sl.eigh(h, S)
            for w, c in enumerate(cs):
                ke = C @ (np.exp(-1j * E * 10 * DT / O.H_BAR) * (C.T @ (S @ c)))
                _, k = systems[i].wavepacket(w)
                assert np.abs(k - ke).max() <= 1e-9 * np.abs(ke).max(), (i, w)
                _, pops, _ = systems[i].observables(w)
                pe = O.populations(b.fragment, 2, np.conj(ke), S @ ke)
                assert np.abs(pops - pe).max() < 1e-10, (i, w)
        batch.close()
        for s in systems:
            s.close()


def test_batch_with_mixed_precision_member():
    """A mixed-precision member runs its own FP32 matrix-core terms inside the
    batch step: its state follows the same member stepped alone in the same
    mode, and stays within 1e-6 in populations of the FP64 exact propagator."""
    sizes = [300, 517, 640]
    data = [_system_inputs(700 + i, n, 2) for i, n in enumerate(sizes)]
    _say(f"batch with a mixed member, N={sizes}")
    with hip.Context(0) as ctx:
        batched, alone = [], None
        for j, (st, b, S, E, C, kets) in enumerate(data):
            s = hip.System(ctx, b, n_wp=2, n_frag=2)
            if j == 2:
                s.set_precision(hip.DH_PREC_MIXED)
            s.build(st.coords)
            for w, k in enumerate(kets):
                s.set_wavepacket(w, S @ k, k)
            batched.append(s)
        st, b, S, E, C, kets = data[2]
        alone = hip.System(ctx, b, n_wp=2, n_frag=2)
        alone.set_precision(hip.DH_PREC_MIXED)
        alone.build(st.coords)
        for w, k in enumerate(kets):
            alone.set_wavepacket(w, S @ k, k)
        batch = hip.Batch(ctx, batched)
        for _ in range(20):
            batch.propagate(0.0, DT)
            alone.propagate(0.0, DT)
        for w in range(2):
            bb, kb = batched[2].wavepacket(w)
            ba, ka = alone.wavepacket(w)
            assert np.abs(kb - ka).max() <= 1e-11 * np.abs(ka).max()
            ke = C @ (np.exp(-1j * E * 20 * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
            pe = O.populations(b.fragment, 2, np.conj(ke), S @ ke)
            _, pb, _ = batched[2].observables(w)
            assert np.abs(pb - pe).max() <= 1e-6
        # the FP64 members against their exact propagators
        for j in (0, 1):
            st, b, S, E, C, kets = data[j]
            for w in range(2):
                ke = C @ (np.exp(-1j * E * 20 * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
                _, kb = batched[j].wavepacket(w)
                assert np.abs(kb - ke).max() <= 1e-9 * np.abs(ke).max()
        batch.close()
