"""Continuous-batching request loop for the propagator.

A request is (geometry + basis, wavepackets Psi_bra/Psi_ket, k nuclear steps of
dt): what an MD driver hands the propagator each nuclear step
(Chebyshev_driver.f:102-169 calls ElHl_Chebyshev once per step).  The server:

  * admits arrived requests first-come first-served by device footprint
    (dh_system_bytes_prec) against an HBM budget, and hands each to a lane;
  * runs its lanes concurrently: each lane is a library context (its own HIP
    stream and build workspace) driven by its own host thread, so the lanes'
    kernels overlap on the device.  Lane 0 takes FP64 requests of N <=
    batch_max_n and advances them together, grouped by padded wavepacket count
    into one grouped launch per Chebyshev term (small systems: launch overhead
    dominates).  The other lanes take the larger and the mixed-precision
    requests (FP32 matrix-core products with the FP64 residual correction),
    each stepping with the kernels the autotuner picked for its shape
    (symmetric half-storage term, matrix-core GEMM), several steps per call
    (dh_propagate_steps) round-robin over the lane's requests;
  * a lane builds S, h and the orthogonal operators of a request on the device
    when it takes it and keeps its state resident until its k steps end;
    requests that arrive meanwhile join their lane at its next iteration (no
    drain), and a finished request frees its slot and memory at once.

Every step of every request is a fixed step of dt from t = 0 (dh_propagate(0,
dt)), so a request's arithmetic does not depend on which requests it shares a
lane or a batch with: its result is the same bits alone and in the mix.

Latency = completion wall time - arrival wall time.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time
from collections import deque

import numpy as np

from . import hip

PAD_W = (1, 2, 4, 8, 16)


@dataclasses.dataclass
class Request:
    rid: int
    basis: object
    coords: np.ndarray
    psi_bra: list            # per wavepacket, N complex (unused with ao_init)
    psi_ket: list            # per wavepacket, N complex; AO coefficients C with ao_init
    steps: int
    dt: float = 5.0e-4
    n_frag: int = 1
    mixed: bool = False
    ao_init: bool = False    # psi_ket holds C: Psi_ket = C, Psi_bra = S C formed on the device
    arrival: float = 0.0     # seconds on the server clock
    # filled by the server
    admitted: float = -1.0
    finished: float = -1.0
    done_steps: int = 0
    lane: int = -1
    result_bra: list = None
    result_ket: list = None
    pops: list = None
    erg: list = None
    system: object = None
    footprint: int = 0


def padded_wp(n):
    for w in PAD_W:
        if n <= w:
            return w
    raise ValueError("at most 16 wavepackets per request")


class _Lane:
    """One context and the host thread that drives it."""

    def __init__(self, server, index, device, batched):
        self.server = server
        self.index = index
        self.batched = batched
        self.ctx = hip.Context(device)
        self.inbox: queue.Queue = queue.Queue()
        self.active: list = []
        self.load = 0.0           # resident N^2 * wavepackets, for lane choice
        self._batches = {}
        self._batch_keys = {}
        self.error = None
        self.thread = threading.Thread(target=self._run, name=f"dh-lane-{index}", daemon=True)

    # ---------------------------------------------------------------- admission
    def _take(self, req):
        w = padded_wp(len(req.psi_ket))
        s = hip.System(self.ctx, req.basis, n_wp=w, n_frag=req.n_frag)
        # lanes run concurrently: no stream capture, which must not overlap the
        # other lanes' allocations and frees
        s.set_graph(False)
        if req.mixed:
            s.set_precision(hip.DH_PREC_MIXED_CORRECTED)
        s.build(req.coords)
        N = req.basis.N
        for i in range(w):
            if i < len(req.psi_ket):
                if req.ao_init:
                    s.set_initial_state(i, req.psi_ket[i], req.psi_ket[i])
                else:
                    s.set_wavepacket(i, req.psi_bra[i], req.psi_ket[i])
            else:  # padding wavepackets are zero and stay zero
                z = np.zeros(N, dtype=np.complex128)
                s.set_wavepacket(i, z, z)
        req.system = s
        self.active.append(req)

    # ---------------------------------------------------------------- stepping
    def _step_batched(self):
        groups = {}
        for r in self.active:
            groups.setdefault((padded_wp(len(r.psi_ket)), r.dt), []).append(r)
        for key, reqs in groups.items():
            ids = tuple(r.rid for r in reqs)
            if self._batch_keys.get(key) != ids:
                old = self._batches.pop(key, None)
                if old is not None:
                    old.close()
                self._batches[key] = hip.Batch(self.ctx, [r.system for r in reqs])
                self._batches[key].set_graph(False)
                self._batch_keys[key] = ids
            self._batches[key].propagate(0.0, key[1])
            for r in reqs:
                r.done_steps += 1

    def _step_solo(self):
        chunk = self.server.solo_chunk
        for r in self.active:
            m = min(chunk, r.steps - r.done_steps)
            if m > 0:
                r.system.propagate_steps(m, r.dt)
                r.done_steps += m

    def _retire(self):
        still, gone = [], []
        for r in self.active:
            (gone if r.done_steps >= r.steps else still).append(r)
        if not gone:
            return
        for r in gone:
            n = len(r.psi_ket)
            r.result_bra, r.result_ket, r.pops, r.erg = [], [], [], []
            for i in range(n):
                b, k = r.system.wavepacket(i)
                _, p, e = r.system.observables(i)
                r.result_bra.append(b)
                r.result_ket.append(k)
                r.pops.append(p)
                r.erg.append(e)
        # batches referencing retired systems are closed before the systems go
        for b in self._batches.values():
            b.close()
        self._batches.clear()
        self._batch_keys.clear()
        for r in gone:
            r.system.close()
            r.system = None
        self.active = still
        self.server._finished(self, gone)

    def _run(self):
        try:
            while True:
                block = not self.active
                try:
                    while True:
                        item = self.inbox.get(block=block, timeout=0.05 if block else None)
                        if item is None:
                            return
                        self._take(item)
                        block = False
                except queue.Empty:
                    pass
                if not self.active:
                    continue
                if self.batched:
                    self._step_batched()
                else:
                    self._step_solo()
                self._retire()
        except Exception as e:  # surfaced by Server.run
            self.error = e
            self.server._wake()

    def close(self):
        for b in self._batches.values():
            b.close()
        self._batches.clear()
        for r in self.active:
            if r.system is not None:
                r.system.close()
        self.active = []
        self.ctx.close()


class Server:
    """n_solo_lanes + 1 lanes on one device; hbm_budget_bytes covers the resident
    systems and every lane's build workspace."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 96 << 30, n_solo_lanes: int = 2,
                 batch_max_n: int = 2048, max_active: int = 512, solo_chunk: int = 8):
        self.device = device
        self.batch_max_n = batch_max_n
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.n_solo_lanes = max(1, int(n_solo_lanes))
        self.solo_chunk = max(1, int(solo_chunk))
        self.used = 0
        self.n_active = 0
        self.queue: deque = deque()
        self.done: list = []
        self.rejected: list = []
        self.t0 = None
        self._cv = threading.Condition()

    def now(self):
        return time.perf_counter() - self.t0

    @staticmethod
    def footprint(req: Request) -> int:
        """Resident device bytes of the request's system (operators, state,
        buffers, and the FP32 operator copies of a mixed-precision request)."""
        return hip.system_bytes(req.basis.N, padded_wp(len(req.psi_ket)),
                                hip.DH_PREC_MIXED_CORRECTED if req.mixed else hip.DH_PREC_FP64)

    @staticmethod
    def workspace_bytes(n_max: int) -> int:
        """A context's operator-build workspace (S, h and two N x N work matrices) for the largest N."""
        ld = (n_max + 63) // 64 * 64
        return 4 * 8 * ld * ld

    def _solo(self, r):
        return r.mixed or r.basis.N > self.batch_max_n

    def _wake(self):
        with self._cv:
            self._cv.notify_all()

    def _finished(self, lane, reqs):
        with self._cv:
            for r in reqs:
                r.finished = self.now()
                self.used -= r.footprint
                self.n_active -= 1
                lane.load -= r.basis.N ** 2 * len(r.psi_ket)
                self.done.append(r)
            self._cv.notify_all()

    def _admit(self, lanes):
        """Hands every admissible request to a lane (called with the lock held)."""
        while self.queue and self.n_active < self.max_active:
            req = self.queue[0]
            if req.arrival > self.now() or self.used + req.footprint > self.budget:
                break  # first-come first-served: wait for the arrival or for memory
            self.queue.popleft()
            lane = lanes[0] if not self._solo(req) else min(lanes[1:], key=lambda l: l.load)
            req.lane = lane.index
            req.admitted = self.now()
            self.used += req.footprint
            self.n_active += 1
            lane.load += req.basis.N ** 2 * len(req.psi_ket)
            lane.inbox.put(req)

    def run(self, requests, log=None, max_seconds=None):
        """Serve `requests` (arrival times relative to the start); returns the finished list."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        lanes = [_Lane(self, 0, self.device, True)] + \
                [_Lane(self, 1 + i, self.device, False) for i in range(self.n_solo_lanes)]
        n_max_small = max([r.basis.N for r in reqs if not self._solo(r)], default=0)
        n_max_large = max([r.basis.N for r in reqs if self._solo(r)], default=0)
        budget = self.budget - self.workspace_bytes(n_max_small) - self.n_solo_lanes * self.workspace_bytes(n_max_large)
        self.budget = budget
        for r in reqs:
            r.footprint = self.footprint(r)
            if r.footprint > self.budget:
                r.finished = -2.0
                self.rejected.append(r)  # can never fit
            else:
                self.queue.append(r)
        total = len(self.queue)
        self.t0 = time.perf_counter()
        for lane in lanes:
            lane.thread.start()
        last = 0.0
        try:
            with self._cv:
                while len(self.done) < total:
                    for lane in lanes:
                        if lane.error is not None:
                            raise RuntimeError(f"serving lane {lane.index} failed") from lane.error
                    self._admit(lanes)
                    wait = 0.05
                    if self.queue and self.queue[0].arrival > self.now():
                        wait = min(wait, self.queue[0].arrival - self.now())
                    self._cv.wait(timeout=max(wait, 1e-4))
                    if log and self.now() - last > 5.0:
                        last = self.now()
                        log(f"t={last:.1f}s active={self.n_active} queued={len(self.queue)} done={len(self.done)} "
                            f"mem={self.used / 2**30:.1f}GiB lanes=" + ",".join(str(len(l.active)) for l in lanes))
                    if max_seconds is not None and self.now() > max_seconds:
                        break
        finally:
            for lane in lanes:
                lane.inbox.put(None)
            for lane in lanes:
                lane.thread.join()
            for lane in lanes:
                lane.close()
        return self.done


def latency_stats(done):
    lat = np.array([r.finished - r.arrival for r in done if r.finished >= 0])
    if lat.size == 0:
        return {}
    wp_steps = sum(len(r.psi_ket) * r.steps for r in done)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "n": int(lat.size), "wavepacket_steps": int(wp_steps)}
