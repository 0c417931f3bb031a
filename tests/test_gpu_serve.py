"""Serving: continuous batching gives each request the same result alone and in the mix."""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd.serve import Request, Server, latency_stats
from dynemol_amd.structures import heavy_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def make_request(rid, N, n_wp, steps, arrival, seed, mixed=False):
    st = heavy_cluster(seed, N)
    b = st.basis()
    S = O.overlap(st.coords, b)
    h = O.huckel(b, S)
    E, C = sl.eigh(h, S)
    rng = np.random.default_rng(seed)
    kets = []
    for _ in range(n_wp):
        k = C[:, N // 2 - 6:N // 2 + 6] @ (rng.normal(size=12) + 1j * rng.normal(size=12))
        kets.append(k / np.sqrt(np.real(np.conj(k) @ S @ k)))
    return Request(rid=rid, basis=b, coords=st.coords, psi_bra=[S @ k for k in kets], psi_ket=kets, steps=steps,
                   arrival=arrival, n_frag=2, mixed=mixed)


def test_alone_equals_in_mix():
    specs = [(0, 300, 3, 6, 0.0), (1, 520, 1, 4, 0.0), (2, 260, 5, 8, 0.01), (3, 700, 2, 5, 0.02),
             (4, 410, 3, 7, 0.03), (5, 333, 16, 3, 0.05), (6, 290, 2, 6, 0.0)]
    _say(f"serve mix of {len(specs)} requests")
    with hip.Context(0) as ctx:
        mix = [make_request(i, n, w, k, a, 200 + i) for (i, n, w, k, a) in specs]
        mix.append(make_request(7, 350, 4, 5, 0.0, 300, mixed=True))
        # N > 400 steps on its own (tuned single-system kernels), N <= 400 in grouped batches
        srv = Server(ctx, hbm_budget_bytes=64 * 2**30, batch_max_n=400)
        done = {r.rid: r for r in srv.run(mix)}
        assert len(done) == len(mix)
        for probe in (0, 1, 2, 3, 5, 7):
            i, n, w, k, a = specs[probe] if probe < len(specs) else (7, 350, 4, 5, 0.0)
            alone_req = make_request(i, n, w, k, 0.0, 200 + i if probe < len(specs) else 300, mixed=(probe == 7))
            alone = Server(ctx, hbm_budget_bytes=64 * 2**30, batch_max_n=400).run([alone_req])[0]
            # Grouped kernels have a fixed shape, and a solo request runs the variant
            # the autotuner chose for its shape once per process: both bitwise.
            tol = 0.0
            for q in range(len(alone.result_ket)):
                assert np.abs(done[i].result_ket[q] - alone.result_ket[q]).max() <= tol
                assert np.abs(done[i].pops[q] - alone.pops[q]).max() <= tol
        stats = latency_stats(list(done.values()))
        assert stats["n"] == len(mix) and stats["p99_s"] >= stats["p50_s"] > 0


def test_admission_waits_for_memory():
    _say("serve admission under a tight budget")
    with hip.Context(0) as ctx:
        reqs = [make_request(i, 400, 2, 3, 0.0, 400 + i) for i in range(4)]
        one = hip.system_bytes(400, 2)
        srv = Server(ctx, hbm_budget_bytes=Server.workspace_bytes(400) + 2 * one + one // 2)
        done = srv.run(reqs)
        assert len(done) == 4
        # at most two requests were ever resident together
        spans = sorted((r.admitted, r.finished) for r in done)
        for t in [a for a, _ in spans]:
            assert sum(1 for a, f in spans if a <= t < f) <= 2
