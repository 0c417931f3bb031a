"""dh_propagate_steps: the nuclear-step loop on the device (Chebyshev_driver.f:102-169).

With one recurrence per step (dh_set_steps_per_series(1)), several steps and
operator rebuilds in one call must give the same bits as the step-by-step
sequence dh_set_coords + dh_build_operators at the rebuild steps and
dh_propagate(0, dt) for every step: the same kernels in the same order, so the
final state, every step's populations and every step's energies agree exactly.
By default the steps of an interval share recurrences (each step's state from
the same terms with its own coefficients); those results agree with the
step-by-step ones to rounding, step for step.  Over a fixed geometry both are
also checked against the exact propagator (FP64 bar: 1e-9 relative in Psi,
1e-10 in populations).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster, tio2_cluster, tio2_slab_adsorbate
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


def _pair(ctx, st, kets, S):
    out = []
    for _ in range(2):
        s = hip.System(ctx, st.basis(), n_wp=len(kets), n_frag=len(st.fragment_names))
        s.build(st.coords)
        for w, k in enumerate(kets):
            s.set_wavepacket(w, S @ k, k)
        out.append(s)
    return out


@pytest.mark.parametrize("which,n_steps,every", [("ch", 12, 5), ("tio2", 12, 4), ("slab", 21, 10), ("ch", 7, 0)])
def test_steps_equal_step_by_step(ctx, which, n_steps, every):
    st = {"ch": lambda: ch_cluster(seed=1), "tio2": lambda: tio2_cluster(seed=6),
          "slab": lambda: tio2_slab_adsorbate(seed=2)}[which]()
    b = st.basis()
    S = O.overlap(st.coords, b)
    rng = np.random.default_rng(9)
    kets = []
    for _ in range(2):
        x = rng.normal(size=b.N) + 1j * rng.normal(size=b.N)
        kets.append(x / np.sqrt(np.real(np.conj(x) @ S @ x)))
    n_geo = (n_steps - 1) // every if every else 0
    geos = [st.coords + rng.normal(0.0, 0.02, st.coords.shape) for _ in range(n_geo)]
    _say(f"{which} N={b.N}: {n_steps} steps, rebuild every {every}, one call vs step by step")
    A, B = _pair(ctx, st, kets, S)
    A.set_steps_per_series(1)
    pops_a, erg_a, k_a = A.propagate_steps(n_steps, DT, every, geos)
    pops_b = np.zeros_like(pops_a)
    erg_b = np.zeros_like(erg_a)
    for j in range(n_steps):
        if every and j > 0 and j % every == 0:
            B.build(geos[j // every - 1])
        B.propagate(0.0, DT)
        for w in range(2):
            _, pops_b[j, w], erg_b[j, w] = B.observables(w)
    assert np.array_equal(pops_a, pops_b)
    assert np.array_equal(erg_a, erg_b)
    for w in range(2):
        ba, ka = A.wavepacket(w)
        bb, kb = B.wavepacket(w)
        assert np.array_equal(ba, bb) and np.array_equal(ka, kb)
    sa, sb = A.stats(), B.stats()
    assert sa["steps"] == sb["steps"] == n_steps and sa["rebuilds"] == sb["rebuilds"]
    _say(f"  identical: {k_a} terms per step, {sa['rebuilds']} builds")
    if not every:
        # fixed geometry: the exact propagator
        h = O.huckel(b, S)
        E, C = sl.eigh(h, S)
        for w in range(2):
            ke = C @ (np.exp(-1j * E * n_steps * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
            _, ka = A.wavepacket(w)
            assert np.abs(ka - ke).max() <= 1e-9 * np.abs(ke).max()
            pe = O.populations(b.fragment, len(st.fragment_names), np.conj(ke), S @ ke)
            assert np.abs(pops_a[-1, w] - pe).max() < 1e-10
    A.close()
    B.close()


# term kernels with the multi-step epilogue: 18 and 21 (symmetric half storage,
# term_epilogue_pre), 15 (split-K GEMM, term_epilogue); -1 = the autotuned one
@pytest.mark.parametrize("which,n_steps,every,variant", [("ch", 23, 0, 18), ("ch", 23, 0, 15), ("tio2", 12, 5, 21),
                                                         ("slab", 21, 10, -1)])
def test_multi_step_series_match_single_steps(ctx, which, n_steps, every, variant):
    st = {"ch": lambda: ch_cluster(seed=1), "tio2": lambda: tio2_cluster(seed=6),
          "slab": lambda: tio2_slab_adsorbate(seed=2)}[which]()
    b = st.basis()
    S = O.overlap(st.coords, b)
    rng = np.random.default_rng(11)
    kets = []
    for _ in range(2):
        x = rng.normal(size=b.N) + 1j * rng.normal(size=b.N)
        kets.append(x / np.sqrt(np.real(np.conj(x) @ S @ x)))
    n_geo = (n_steps - 1) // every if every else 0
    geos = [st.coords + rng.normal(0.0, 0.02, st.coords.shape) for _ in range(n_geo)]
    _say(f"{which} N={b.N}: {n_steps} steps, rebuild every {every}, variant {variant}, multi-step series vs one per step")
    hip.force_matvec_variant(variant)
    try:
        _multi_vs_single(ctx, st, b, S, kets, geos, n_steps, every)
    finally:
        hip.force_matvec_variant(-1)


def _multi_vs_single(ctx, st, b, S, kets, geos, n_steps, every):
    A, B = _pair(ctx, st, kets, S)
    A.set_timing(True)
    B.set_timing(True)
    B.set_steps_per_series(1)
    pops_a, erg_a, k_a = A.propagate_steps(n_steps, DT, every, geos)
    pops_b, erg_b, k_b = B.propagate_steps(n_steps, DT, every, geos)
    sa, sb = A.stats(), B.stats()
    _say(f"  terms: {sa['series_launches']:.0f} in shared series against {sb['series_launches']:.0f}")
    assert sa["series_launches"] < sb["series_launches"]
    # every step's outputs, and the final state
    assert np.abs(pops_a - pops_b).max() < 1e-12
    assert np.abs(erg_a - erg_b).max() < 1e-10 * np.abs(erg_b).max()
    for w in range(2):
        ba, ka = A.wavepacket(w)
        bb, kb = B.wavepacket(w)
        assert np.abs(ka - kb).max() <= 1e-11 * np.abs(kb).max()
        assert np.abs(ba - bb).max() <= 1e-11 * np.abs(bb).max()
    if not every:
        h = O.huckel(b, S)
        E, C = sl.eigh(h, S)
        for w in range(2):
            ke = C @ (np.exp(-1j * E * n_steps * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
            _, ka = A.wavepacket(w)
            assert np.abs(ka - ke).max() <= 1e-9 * np.abs(ke).max()
            pe = O.populations(b.fragment, len(st.fragment_names), np.conj(ke), S @ ke)
            assert np.abs(pops_a[-1, w] - pe).max() < 1e-10
            # an intermediate step too (step 13: second series, third accumulator)
            ke13 = C @ (np.exp(-1j * E * 13 * DT / O.H_BAR) * (C.T @ (S @ kets[w])))
            pe13 = O.populations(b.fragment, len(st.fragment_names), np.conj(ke13), S @ ke13)
            assert np.abs(pops_a[12, w] - pe13).max() < 1e-10
    A.close()
    B.close()


# a step so short that a recurrence has fewer terms than one panel of 16: the
# panel route (15) adds them all in the pass after the last term, the paired
# route (18) ends on a single or a paired update depending on the last term's
# parity; both must give the single steps' results
@pytest.mark.parametrize("variant", [18, 15])
def test_multi_step_short_series(ctx, variant):
    st = ch_cluster(seed=1)
    b = st.basis()
    S = O.overlap(st.coords, b)
    rng = np.random.default_rng(12)
    kets = []
    for _ in range(2):
        x = rng.normal(size=b.N) + 1j * rng.normal(size=b.N)
        kets.append(x / np.sqrt(np.real(np.conj(x) @ S @ x)))
    dt, n_steps = 1.0e-8, 10
    _say(f"ch N={b.N}: {n_steps} steps of {dt} ps, variant {variant}, short multi-step series vs one per step")
    hip.force_matvec_variant(variant)
    try:
        A, B = _pair(ctx, st, kets, S)
        B.set_steps_per_series(1)
        pops_a, _, k_a = A.propagate_steps(n_steps, dt)
        pops_b, _, _ = B.propagate_steps(n_steps, dt)
        _say(f"  {k_a} terms in the last shared series")
        assert k_a <= 16
        assert np.abs(pops_a - pops_b).max() < 1e-13
        for w in range(2):
            _, ka = A.wavepacket(w)
            _, kb = B.wavepacket(w)
            assert np.abs(ka - kb).max() <= 1e-12 * np.abs(kb).max()
        A.close()
        B.close()
    finally:
        hip.force_matvec_variant(-1)


def test_steps_argument_checks(ctx):
    st = ch_cluster(seed=1)
    s = hip.System(ctx, st.basis(), n_wp=1, n_frag=2)
    with pytest.raises(hip.DHError):
        s.propagate_steps(3, DT)  # operators not built yet
    s.build(st.coords)
    s.set_initial_state(0, np.ones(s.N) + 0j, np.ones(s.N) + 0j)
    with pytest.raises(hip.DHError):
        s.propagate_steps(0, DT)
    with pytest.raises(hip.DHError):
        s.propagate_steps(3, -DT)
    s.close()
