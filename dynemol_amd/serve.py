"""Continuous-batching request loop for the propagator.

A request is (geometry + basis, wavepackets Psi_bra/Psi_ket, k nuclear steps of
dt): what an MD driver hands the propagator each nuclear step
(Chebyshev_driver.f:102-169 calls ElHl_Chebyshev once per step).  The server:

  * admits queued requests by device footprint (dh_system_bytes) against an HBM
    budget, building S, h and the orthogonal operators on the device at
    admission, then keeps the request's state resident until its k steps end;
  * advances every resident request one step per loop iteration, grouping FP64
    requests of N <= batch_max_n by (padded) wavepacket count into one grouped
    launch per Chebyshev term (small systems: launch overhead dominates);
    larger systems and mixed-precision requests step on their own with the
    kernels the autotuner picked for their shapes (symmetric half-storage
    matvec, matrix-core GEMM);
  * requests that arrive while others run join at the next iteration (no
    drain), and a finished request frees its slot and memory at once.

Latency = completion wall time - arrival wall time.
"""
from __future__ import annotations

import dataclasses
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
    t: float = 0.0
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


class Server:
    def __init__(self, ctx: hip.Context, hbm_budget_bytes: int, max_active: int = 512, batch_max_n: int = 2048):
        self.ctx = ctx
        self.batch_max_n = batch_max_n
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.used = 0
        self.queue: deque = deque()
        self.active: list = []
        self.done: list = []
        self.t0 = None
        self._batches = {}
        self._batch_keys = {}
        self.rejected: list = []

    def now(self):
        return time.perf_counter() - self.t0

    def footprint(self, req: Request) -> int:
        """Resident device bytes of the request's system (operators, state, buffers)."""
        return hip.system_bytes(req.basis.N, padded_wp(len(req.psi_ket)))

    @staticmethod
    def workspace_bytes(n_max: int) -> int:
        """The context's operator-build workspace (S, h and two N x N work matrices) for the largest N."""
        ld = (n_max + 63) // 64 * 64
        return 4 * 8 * ld * ld

    def submit(self, req: Request):
        req.footprint = self.footprint(req)
        if req.footprint > self.budget:
            req.finished = -2.0
            self.rejected.append(req)  # can never fit
            return
        self.queue.append(req)

    def _admit(self):
        admitted = False
        while self.queue and len(self.active) < self.max_active:
            req = self.queue[0]
            if req.arrival > self.now():
                break
            if self.used + req.footprint > self.budget:
                break  # first-come first-served: wait for memory to free up
            self.queue.popleft()
            w = padded_wp(len(req.psi_ket))
            s = hip.System(self.ctx, req.basis, n_wp=w, n_frag=req.n_frag)
            if req.mixed:
                s.set_precision(True)
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
            req.admitted = self.now()
            self.used += req.footprint
            self.active.append(req)
            admitted = True
        return admitted

    def _solo(self, r):
        return r.mixed or r.basis.N > self.batch_max_n

    def _groups(self):
        groups = {}
        for r in self.active:
            if self._solo(r):
                continue
            groups.setdefault((padded_wp(len(r.psi_ket)), r.dt), []).append(r)
        return groups

    def _step(self):
        groups = self._groups()
        for w, reqs in groups.items():
            key = tuple(r.rid for r in reqs)
            if self._batch_keys.get(w) != key:
                old = self._batches.pop(w, None)
                if old is not None:
                    old.close()
                self._batches[w] = hip.Batch(self.ctx, [r.system for r in reqs])
                self._batch_keys[w] = key
            dt = reqs[0].dt  # a group shares the step length
            t = reqs[0].t
            self._batches[w].propagate(t, t + dt)
            for r in reqs:
                r.t += dt
                r.done_steps += 1
        for r in self.active:
            if self._solo(r):
                r.system.propagate(r.t, r.t + r.dt)
                r.t += r.dt
                r.done_steps += 1

    def _retire(self):
        still = []
        for r in self.active:
            if r.done_steps >= r.steps:
                n = len(r.psi_ket)
                r.result_bra, r.result_ket, r.pops, r.erg = [], [], [], []
                for i in range(n):
                    b, k = r.system.wavepacket(i)
                    _, p, e = r.system.observables(i)
                    r.result_bra.append(b)
                    r.result_ket.append(k)
                    r.pops.append(p)
                    r.erg.append(e)
                r.finished = self.now()
                self.used -= r.footprint
                self.done.append(r)
            else:
                still.append(r)
        if len(still) != len(self.active):
            # batches referencing retired systems are rebuilt before the systems go
            for b in self._batches.values():
                b.close()
            self._batches.clear()
            self._batch_keys.clear()
            for r in self.active:
                if r.done_steps >= r.steps:
                    r.system.close()
                    r.system = None
        self.active = still

    def run(self, requests, log=None, max_seconds=None):
        """Serve `requests` (arrival times relative to the start); returns the finished list."""
        if requests:
            self.budget -= self.workspace_bytes(max(r.basis.N for r in requests))
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        self.t0 = time.perf_counter()
        last = 0.0
        while self.queue or self.active:
            self._admit()
            if not self.active:
                # idle until the next arrival
                nxt = self.queue[0].arrival if self.queue else None
                if nxt is not None:
                    time.sleep(max(0.0, min(0.01, nxt - self.now())))
                continue
            self._step()
            self._retire()
            if log and self.now() - last > 5.0:
                last = self.now()
                log(f"t={last:.1f}s active={len(self.active)} queued={len(self.queue)} done={len(self.done)} "
                    f"mem={self.used / 2**30:.1f}GiB")
            if max_seconds is not None and self.now() > max_seconds:
                break
        for b in self._batches.values():
            b.close()
        self._batches.clear()
        return self.done


def latency_stats(done):
    lat = np.array([r.finished - r.arrival for r in done if r.finished >= 0])
    if lat.size == 0:
        return {}
    wp_steps = sum(len(r.psi_ket) * r.steps for r in done)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "n": int(lat.size), "wavepacket_steps": int(wp_steps)}
