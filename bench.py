"""Benchmark: wavepacket-steps/s of the electron/hole propagator at N = 4096.

Workload (one GPU): anatase-like TiO2 slab (236 TiO2) + C16H20 adsorbate, N = 4096
STO orbitals, electron (LUMO) and hole (HOMO) wavepackets, steps of 0.5 fs; S
and h rebuilt on the device every 10 steps from coordinates jittered by a
Gaussian of sigma = 0.02 A.  A step advances both wavepackets (2
wavepacket-steps) and includes populations and <H> on the device.

Prints one JSON line (rank 0).  --gpus N > 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT_PS = 5.0e-4
REBUILD_EVERY = 10
JITTER = 0.02
HBM_PEAK_GBS = 8000.0
# HBM bytes per Chebyshev term (both kernels of the symmetric variant, N=4096,
# 4 RHS) from rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by what the same
# counter reads for the library's plain copy kernel over 256 MiB (FETCH_SIZE x 2.000,
# WRITE_SIZE x 1.000): profiles/cheb_term_sym_tio2_N4096_counters.md.
# Variants 17, 18 and 20 likewise: profiles/cheb_term_symm_tio2_N4096_counters.md,
# profiles/cheb_term_symm4_tio2_N4096_counters.md and profiles/cheb_term_symm3_tio2_N4096_counters.md;
# variant 21: profiles/cheb_term_symm_up_N4096_counters.md.
MEASURED_TRAFFIC_BYTES = {16: int(((35409.8125 + 5487.5) * 2.0 + (10924.0 + 256.0)) * 1024),
                          17: int(((34796.36781939338 + 5874.9525547445255) * 2.0
                                   + (10924.001321231617 + 508.26277372262774)) * 1024),
                          18: int(((34810.57783777574 + 5872.213768115942) * 2.0
                                   + (10924.024931066177 + 506.4347826086956)) * 1024),
                          20: int(((34719.24046415441 + 5869.5) * 2.0
                                   + (10925.4521484375 + 504.63309352517985)) * 1024),
                          21: int(((34452.90 + 5881.50) * 2.0 + (10924.01 + 512.00)) * 1024)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_S_h(st, b):
    """S and h as the library builds them on the device (untimed setup)."""
    from dynemol_amd import hip
    with hip.Context(0) as ctx:
        s = hip.System(ctx, b, n_wp=1, n_frag=len(st.fragment_names))
        s.build(st.coords)
        S, h = s.matrix(hip.DH_MAT_S), s.matrix(hip.DH_MAT_H)
        s.close()
    return S, h


def homo_index(st):
    """0-based HOMO index of the (h, S) pencil: valence electrons / 2 - 1."""
    nel = sum({"Ti": 4, "O": 6, "C": 4, "H": 1, "N": 5}[s] for s in st.symbols)
    return nel // 2 - 1


HOST_EIGH_S = None  # host time of the setup's LUMO/HOMO solve, reported beside the device one


def initial_states(st, b):
    """LUMO/HOMO of the (h, S) pencil (whole system flagged El and Hl, FMO.f:294-350)."""
    global HOST_EIGH_S
    import scipy.linalg as sl
    S, h = device_S_h(st, b)
    homo = homo_index(st)
    t0 = time.perf_counter()
    E, C = sl.eigh(h, S, subset_by_index=[homo, homo + 1])
    HOST_EIGH_S = time.perf_counter() - t0
    return S, h, C[:, 1], C[:, 0], E


def run_gpu(args, st, b, S, lumo, homo):
    """The drop-in's nuclear-step loop on the device (dh_propagate_steps): W
    warm-up steps, then the timed region: the rebuild at its first step from new
    coordinates and K steps with a rebuild every REBUILD_EVERY steps, each from
    coordinates jittered by JITTER (all geometries generated before the clock
    starts; the library copies them to the device once per call)."""
    from dynemol_amd import hip
    ctx = hip.Context(0)
    sysm = hip.System(ctx, b, n_wp=2, n_frag=len(st.fragment_names))
    rng = np.random.default_rng(2)
    sysm.build(st.coords)
    for w, c in enumerate((lumo, homo)):
        sysm.set_wavepacket(w, S @ c + 0j, c + 0j)

    def geometries(n):
        return np.stack([st.coords + rng.normal(0.0, JITTER, st.coords.shape) for _ in range(max(n, 0))]) \
            if n > 0 else None

    if args.warmup > 0:
        sysm.propagate_steps(args.warmup, DT_PS, REBUILD_EVERY, geometries((args.warmup - 1) // REBUILD_EVERY))
    first = st.coords + rng.normal(0.0, JITTER, st.coords.shape)
    geos = geometries((args.steps - 1) // REBUILD_EVERY)
    sysm.set_timing(True)
    st0 = sysm.stats()
    t_start = time.perf_counter()
    t0 = time.perf_counter()
    sysm.build(first)
    first_build_s = time.perf_counter() - t0
    pops_hist, erg_hist, k = sysm.propagate_steps(args.steps, DT_PS, REBUILD_EVERY, geos)
    elapsed = time.perf_counter() - t_start
    st1 = sysm.stats()
    n_rebuilds = st1["rebuilds"] - st0["rebuilds"]
    side_ms = (st1["rebuild_ms_in_steps"] - st0["rebuild_ms_in_steps"]) / max(1, n_rebuilds - 1)
    pops = [pops_hist[-1, w] for w in range(2)]
    # outside the timed region: the same steps without rebuilds, and rebuilds alone
    t0 = time.perf_counter()
    sysm.propagate_steps(20, DT_PS)
    prop_only = 2 * 20 / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(2):
        sysm.build(st.coords + rng.normal(0.0, JITTER, st.coords.shape))
    rebuild_s = (time.perf_counter() - t0) / 2
    bounds = sysm.bounds()
    tuning = sysm.tuning()
    ehrenfest = None if args.no_ehrenfest else time_ehrenfest(sysm)
    eigen = None if args.no_ehrenfest else time_eigenpairs(sysm, homo_index(st))
    sysm.close()
    ctx.close()
    series_ms = st1["series_ms"] - st0["series_ms"]
    launches = st1["series_launches"] - st0["series_launches"]
    return {
        "elapsed": elapsed, "rebuild_s": rebuild_s, "terms": [k], "series_ms": series_ms, "prop_only": prop_only,
        "side_build_ms": side_ms, "first_build_s": first_build_s,
        "series_launches": launches, "pops": pops, "bounds": bounds.tolist(), "rebuilds": n_rebuilds,
        "tuning": tuning, "ehrenfest": ehrenfest, "eigenpairs": eigen,
    }


def time_eigenpairs(sysm, homo, reps=3):
    """HOMO and LUMO of the final geometry's (h, S) on the device (dh_eigenpairs),
    beside the host scipy.linalg.eigh subset solve of the setup."""
    k = homo + 1  # 1-based HOMO
    sysm.eigenpairs(k, k + 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        E, C = sysm.eigenpairs(k, k + 1)
    dev_s = (time.perf_counter() - t0) / reps
    log(f"eigenpairs HOMO/LUMO N={sysm.N}: {1e3 * dev_s:.1f} ms on the device"
        + (f", setup's host eigh {1e3 * HOST_EIGH_S:.0f} ms" if HOST_EIGH_S else ""))
    return {"n_orbitals": sysm.N, "indices_1based": [k, k + 1], "device_ms": round(1e3 * dev_s, 2),
            "energies_eV": [round(float(x), 6) for x in E],
            "host_scipy_eigh_ms": round(1e3 * HOST_EIGH_S, 1) if HOST_EIGH_S else None}


def time_ehrenfest(sysm, reps=5):
    """Diabatic-Ehrenfest kernel K = X o A - H' A on the drop-in's final state
    (after the timed steps): device time per call and with K copied to the host,
    and the invariant <K, S>_F = 0 as a check."""
    from dynemol_amd import hip
    N = sysm.N
    t0 = time.perf_counter()
    K = sysm.ehrenfest_kernel(0, 1)  # forms H' (one N^3 triangular product) first
    first_s = time.perf_counter() - t0
    S = sysm.matrix(hip.DH_MAT_S)
    inv = float(abs(np.sum(K * S)) / np.sum(np.abs(K * S)))
    t0 = time.perf_counter()
    for _ in range(reps):
        sysm.ehrenfest_kernel(0, 1, to_host=False)
    dev_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    sysm.ehrenfest_kernel(0, 1)
    host_s = time.perf_counter() - t0
    # the reference's CPU branch is K = X*A - matmul(H', A) (diabatic-Ehren.f:112-116): the same
    # arithmetic with the host BLAS behind numpy, on matrices of the same size
    rng = np.random.default_rng(0)
    Hh, Ah, Xh = (np.asfortranarray(rng.normal(size=(N, N))) for _ in range(3))
    t0 = time.perf_counter()
    _ = Xh * Ah - Hh @ Ah
    cpu_s = time.perf_counter() - t0
    log(f"ehrenfest kernel N={N}: {1e3 * dev_s:.2f} ms on the device, {1e3 * host_s:.2f} ms with K to the host, "
        f"{1e3 * cpu_s:.0f} ms with numpy on the host")
    return {"n_orbitals": N, "device_ms": round(1e3 * dev_s, 3), "with_host_copy_ms": round(1e3 * host_s, 3),
            "first_call_ms": round(1e3 * first_s, 3), "gemm_tflops": round(2.0 * N ** 3 / dev_s / 1e12, 2),
            "fp64_mfma_peak_tflops": 78.6, "K_S_frobenius_rel": inv,
            "host_numpy_ms": round(1e3 * cpu_s, 1), "host_threads": os.environ.get("OMP_NUM_THREADS")}


# symmetric half-storage term kernels (vector FMAs; matrix-core 16x16x4, 4x4x4; super-tiles, column
# strips, segments loaded up front)
SYM_VARIANTS = (16, 17, 18, 19, 20, 21)


def term_bytes(N, R):
    """Algorithmic HBM bytes of one Chebyshev term: H read once for all R
    complex right-hand sides (8 N^2) plus 80 N R of vector traffic (read
    T_{k-1} and T_{k-2}, write T_k, read and write the accumulator; 16 B per
    complex)."""
    return 8 * N * N + 80 * N * R


def half_storage_bytes(N, R):
    """Bytes the symmetric half-storage kernels cannot avoid: the nb (nb + 1) / 2
    tiles of 64 x 64 on and below the diagonal of H, plus the same vectors."""
    nb = (N + 63) // 64
    return 8 * 64 * 64 * nb * (nb + 1) // 2 + 80 * N * R


def run_batched(st, b, S, n_wp, steps, seed=11):
    """n_wp wavepackets on the N = 4096 system through the multi-RHS term kernel:
    FP64, mixed (FP32 products) and mixed with the residual correction."""
    from dynemol_amd import hip
    rng = np.random.default_rng(seed)
    K = rng.normal(size=(b.N, n_wp)) + 1j * rng.normal(size=(b.N, n_wp))
    SK = S @ np.ascontiguousarray(K.real) + 1j * (S @ np.ascontiguousarray(K.imag))  # real BLAS GEMMs
    nrm = np.sqrt(np.real(np.sum(np.conj(K) * SK, axis=0)))
    kets = [K[:, w] / nrm[w] for w in range(n_wp)]
    bras = [SK[:, w] / nrm[w] for w in range(n_wp)]
    out = {}
    modes = (hip.DH_PREC_FP64, hip.DH_PREC_MIXED, hip.DH_PREC_MIXED_CORRECTED)
    with hip.Context(0) as ctx:
        for mode in modes:
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=len(st.fragment_names))
            s.set_precision(mode)
            s.build(st.coords)
            for w, k in enumerate(kets):
                s.set_wavepacket(w, bras[w], k)
            s.propagate_steps(1, DT_PS)  # warm-up step (also autotunes)
            t0 = time.perf_counter()
            s.propagate_steps(steps, DT_PS)
            el = time.perf_counter() - t0
            pops = np.array([s.observables(w)[1] for w in range(n_wp)])
            out[mode] = {"value": n_wp * steps / el, "pops": pops, "tuning": s.tuning()["term"],
                         "terms": s.stats()["terms"]}
            s.close()
    f = out[hip.DH_PREC_FP64]
    m1, m3 = out[hip.DH_PREC_MIXED], out[hip.DH_PREC_MIXED_CORRECTED]
    log(f"batched: fp64 {f['value']:.1f}, mixed {m1['value']:.1f}, corrected {m3['value']:.1f} wavepacket-steps/s")
    return {"n_orbitals": b.N, "wavepackets": n_wp, "steps": steps + 1,
            "fp64_value": round(f["value"], 3), "mixed_value": round(m1["value"], 3),
            "mixed_corrected_value": round(m3["value"], 3),
            "fp64_term_variant": f["tuning"], "chebyshev_terms": f["terms"],
            "mixed_pop_err_vs_fp64": float(np.abs(m1["pops"] - f["pops"]).max()),
            "mixed_corrected_pop_err_vs_fp64": float(np.abs(m3["pops"] - f["pops"]).max()),
            "unit": "wavepacket-steps/s"}


def run_serve(n_req, lam, seed=5, n_lo=256, n_hi=8192, budget_gib=96, check=True):
    """Poisson arrivals; N log-uniform, 1-16 wavepackets, 10-200 steps each;
    requests of 8 or more wavepackets run in the mixed precision mode (the
    multi-RHS FP32 matrix-core term) beside the FP64 ones.  After the stream
    (untimed), each FP64 request of N <= 2048 is checked against the exact
    propagator of its device-built S and h, and each mixed request of N <= 4096
    against the same request run alone in FP64."""
    from dynemol_amd import hip
    from dynemol_amd.serve import Request, Server, latency_stats
    from dynemol_amd.structures import heavy_cluster, log_uniform_int
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(n_req):
        N = log_uniform_int(rng, n_lo, n_hi)
        st = heavy_cluster(int(rng.integers(1 << 30)), N)
        b = st.basis()
        w = int(rng.integers(1, 17))
        # AO coefficients; the server forms Psi_ket = C, Psi_bra = S C on the device at admission
        ao = [(rng.normal(size=N) + 1j * rng.normal(size=N)) / np.sqrt(2.0 * N) for _ in range(w)]
        t += rng.exponential(1.0 / lam)
        reqs.append(Request(rid=i, basis=b, coords=st.coords, psi_bra=None, psi_ket=ao, ao_init=True,
                            steps=int(rng.integers(10, 201)), arrival=t, n_frag=2, mixed=w >= 8))
    log(f"serve: {n_req} requests generated, N in [{min(r.basis.N for r in reqs)}, {max(r.basis.N for r in reqs)}], "
        f"{sum(r.mixed for r in reqs)} mixed")
    srv = Server(0, hbm_budget_bytes=budget_gib * 2**30)
    t0 = time.perf_counter()
    done = srv.run(reqs, log=log)
    el = time.perf_counter() - t0
    st = latency_stats(done)
    st.update({"requests": n_req, "arrival_rate_per_s": lam, "throughput_wp_steps_per_s": st["wavepacket_steps"] / el,
               "wall_s": el, "mixed_requests": sum(r.mixed for r in reqs), "lanes": 1 + srv.n_solo_lanes,
               "max_concurrent": _max_concurrent(done)})
    if check:
        st.update(_serve_errors(done))
    return {k: (round(v, 4) if isinstance(v, float) and v > 1e-3 else v) for k, v in st.items()}


def _max_concurrent(done):
    ev = sorted([(r.admitted, 1) for r in done] + [(r.finished, -1) for r in done])
    cur = best = 0
    for _, d in ev:
        cur += d
        best = max(best, cur)
    return best


def _serve_errors(done, fp64_max_n=2048, mixed_max_n=4096):
    """Population error per request: FP64 against the exact propagator, mixed
    against the same request run alone in FP64."""
    import scipy.linalg as sl
    from dynemol_amd import hip
    from dynemol_amd.serve import Request, Server
    err64, errmx, unchecked = [], [], 0
    with hip.Context(0) as ctx:
        for r in done:
            if r.mixed and r.basis.N <= mixed_max_n:
                again = Request(rid=r.rid, basis=r.basis, coords=r.coords, psi_bra=None, psi_ket=r.psi_ket,
                                ao_init=True, steps=r.steps, n_frag=r.n_frag, mixed=False)
                ref = Server(0, hbm_budget_bytes=96 << 30).run([again])[0]
                errmx.append(max(float(np.abs(r.pops[q] - ref.pops[q]).max()) for q in range(len(r.psi_ket))))
            elif not r.mixed and r.basis.N <= fp64_max_n:
                s = hip.System(ctx, r.basis, n_wp=1, n_frag=r.n_frag)
                s.build(r.coords)
                S, h = s.matrix(hip.DH_MAT_S), s.matrix(hip.DH_MAT_H)
                s.close()
                E, C = sl.eigh(h, S)
                ph = np.exp(-1j * E * r.steps * r.dt / hip.H_BAR)
                e = 0.0
                for q, c0 in enumerate(r.psi_ket):
                    ke = C @ (ph * (C.T @ (S @ c0)))
                    z = np.conj(ke) * (S @ ke)
                    pe = np.array([z[r.basis.fragment == f].real.sum() for f in range(r.n_frag)] + [z.real.sum()])
                    e = max(e, float(np.abs(r.pops[q] - pe).max()))
                err64.append(e)
            else:
                unchecked += 1
    log(f"serve errors: fp64 vs exact max {max(err64, default=0):.2e} over {len(err64)}, "
        f"mixed vs fp64 max {max(errmx, default=0):.2e} over {len(errmx)}, {unchecked} unchecked")
    return {"pop_err_max": max(err64 + errmx, default=None), "pop_err_max_fp64_vs_exact": max(err64, default=None),
            "pop_err_max_mixed_vs_fp64": max(errmx, default=None), "checked_fp64": len(err64),
            "checked_mixed": len(errmx), "unchecked": unchecked}


def _exact(S, h, psi0, t):
    """Exact propagation C exp(-i E t / hbar) C^T S psi0 from the (h, S) eigenpairs."""
    import scipy.linalg as sl
    from dynemol_amd import hip
    E, C = sl.eigh(h, S)
    return C @ (np.exp(-1j * E * t / hip.H_BAR) * (C.T @ (S @ psi0)))


def run_many_systems(n_sys=256, steps=500, seed=3, n_wp=4, check=6):
    """Many systems in one batch: C/H/N/O clusters, N log-uniform in [256, 2048],
    4 wavepackets each (random S-normalised superpositions of 8 frontier orbitals
    from the device eigensolver), grouped launches per Chebyshev term, 10 steps
    per recurrence (dh_batch_propagate_steps); 11 untimed steps (the batch's
    tuning step, then one 10-step recurrence that captures its graph) before
    `steps` timed ones; sampled systems checked against the exact propagator of
    their device-built S and h (untimed)."""
    import scipy.linalg as sl
    from dynemol_amd import hip
    from dynemol_amd.structures import heavy_cluster, log_uniform_int
    rng = np.random.default_rng(seed)
    out = {"systems": n_sys, "wavepackets_per_system": n_wp, "steps": steps}
    with hip.Context(0) as ctx:
        systems, init, mats = [], [], []
        t0 = time.perf_counter()
        for i in range(n_sys):
            N = log_uniform_int(rng, 256, 2048)
            st = heavy_cluster(int(rng.integers(1 << 30)), N)
            b = st.basis()
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=2)
            s.build(st.coords)
            mid = N // 2
            E, C = s.eigenpairs(mid - 3, mid + 4)
            cs = []
            for w in range(n_wp):
                a = rng.normal(size=8) + 1j * rng.normal(size=8)
                c = C @ (a / np.linalg.norm(a))
                s.set_initial_state(w, c, c)
                cs.append(c)
            systems.append(s)
            init.append(cs)
            if (i + 1) % 64 == 0:
                log(f"many systems: {i + 1} systems built")
            mats.append((s.matrix(hip.DH_MAT_S), s.matrix(hip.DH_MAT_H)) if i < check else None)
        out["setup_s"] = round(time.perf_counter() - t0, 2)
        out["n_min"] = min(s.N for s in systems)
        out["n_max"] = max(s.N for s in systems)
        out["n_mean"] = round(float(np.mean([s.N for s in systems])), 1)
        batch = hip.Batch(ctx, systems)
        batch.propagate(0.0, DT_PS)  # first step: task tables, kernel choice
        batch.propagate_steps(10, DT_PS)  # a 10-step recurrence: its graph
        t0 = time.perf_counter()
        done, kmax = 0, 0
        while done < steps:
            m = min(100, steps - done)
            kmax = max(kmax, batch.propagate_steps(m, DT_PS))
            done += m
            if done % 100 == 0:
                log(f"many systems: step {done}")
        el = time.perf_counter() - t0
        total = steps + 11
        out["wavepacket_steps_per_s"] = round(n_sys * n_wp * steps / el, 1)
        out["ms_per_step"] = round(1e3 * el / steps, 3)
        out["batch_kernel"] = batch.tuning()
        out["most_terms_per_recurrence"] = kmax
        errs, perr = [], []
        for i in range(check):
            S, h = mats[i]
            E, C = sl.eigh(h, S)
            for w in range(n_wp):
                ke = C @ (np.exp(-1j * E * total * DT_PS / hip.H_BAR) * (C.T @ (S @ init[i][w])))
                _, k = systems[i].wavepacket(w)
                errs.append(float(np.abs(k - ke).max() / np.abs(ke).max()))
                _, pops, _ = systems[i].observables(w)
                z = np.conj(ke) * (S @ ke)
                frag = systems[i].basis.fragment
                pe = np.array([z[frag == f].real.sum() for f in range(2)] + [z.real.sum()])
                perr.append(float(np.abs(pops - pe).max()))
        out["checked_systems"] = check
        out["max_rel_psi_err_vs_exact"] = max(errs)
        out["max_pop_err_vs_exact"] = max(perr)
        batch.close()
        for s in systems:
            s.close()
    log(f"many systems: {out['wavepacket_steps_per_s']} wavepacket-steps/s, {out['ms_per_step']} ms per step, "
        f"max pop err {out['max_pop_err_vs_exact']:.1e}")
    return out


def run_precision_study(steps=1000, seed=4, n_wp=64, check_every=100, system="interface"):
    """n_wp random S-normalised wavepackets on one fixed system, FP64
    matrix-core terms against FP32 products (mixed) and against the FP32-stored
    operator with FP64 products: population, norm and energy differences after
    each block of steps.  system: "interface" = the N = 8192 C/H/O interface
    cluster; "tio2" = the N = 4096 TiO2 slab + C16H20 of the drop-in."""
    from dynemol_amd import hip
    from dynemol_amd.structures import interface_cluster, tio2_slab_adsorbate
    log(f"precision study: generating the {system} system")
    st = interface_cluster(seed=seed) if system == "interface" else tio2_slab_adsorbate(seed=2)
    b = st.basis()
    log(f"precision study: {st.n_atoms} atoms, N={b.N}")
    rng = np.random.default_rng(seed)
    out = {"system": system, "n_orbitals": b.N, "atoms": st.n_atoms, "wavepackets": n_wp, "rhs_columns": 2 * n_wp,
           "steps": steps}
    with hip.Context(0) as ctx:
        res = {}
        names = {hip.DH_PREC_FP64: "fp64", hip.DH_PREC_MIXED_CORRECTED: "mixed_corrected",
                 hip.DH_PREC_MIXED: "mixed", hip.DH_PREC_F32_STORAGE: "f32_storage"}
        for mode, name in names.items():
            s = hip.System(ctx, b, n_wp=n_wp, n_frag=len(st.fragment_names))
            s.set_precision(mode)
            s.build(st.coords)
            log(f"precision study {name}: operators built, bounds {s.bounds()[:2]}")
            if mode == hip.DH_PREC_FP64:
                S = s.matrix(hip.DH_MAT_S)
                K = rng.normal(size=(b.N, n_wp)) + 1j * rng.normal(size=(b.N, n_wp))
                SK = S @ np.ascontiguousarray(K.real) + 1j * (S @ np.ascontiguousarray(K.imag))  # real BLAS GEMMs
                nrm = np.sqrt(np.real(np.sum(np.conj(K) * SK, axis=0)))
                kets = [K[:, w] / nrm[w] for w in range(n_wp)]
                bras = [SK[:, w] / nrm[w] for w in range(n_wp)]
                del S, SK
                log("precision study: initial states ready")
            for w in range(n_wp):
                s.set_wavepacket(w, bras[w], kets[w])
            el, hist = 0.0, []
            for block in range(steps // check_every):
                t0 = time.perf_counter()
                s.propagate_steps(check_every, DT_PS)
                el += time.perf_counter() - t0
                log(f"precision study {name}: step {(block + 1) * check_every}, {s.stats()['terms']} terms")
                obs = [s.observables(w) for w in range(n_wp)]
                hist.append((np.array([o[1] for o in obs]), np.array([o[2] for o in obs])))
            res[name] = {"hist": hist, "value": n_wp * steps / el, "terms": s.stats()["terms"],
                         "bounds": s.bounds()[:2], "drift": s.stats()["norm_drift"]}
            s.close()
    f = res["fp64"]
    out["chebyshev_terms"] = f["terms"]
    out["fp64_wavepacket_steps_per_s"] = round(f["value"], 2)
    out["fp64_norm_drift_last_step"] = f["drift"]
    for name, key in (("mixed_corrected", "corr_"), ("mixed", ""), ("f32_storage", "f32s_")):
        m = res[name]
        out[f"{name}_wavepacket_steps_per_s"] = round(m["value"], 2)
        out[f"{key}pop_err_by_step"] = {str(check_every * (i + 1)): float(np.abs(pm - pf).max())
                                        for i, ((pf, _), (pm, _)) in enumerate(zip(f["hist"], m["hist"]))}
        pf, ef = f["hist"][-1]
        pm, em = m["hist"][-1]
        out[f"{key}final_pop_err"] = float(np.abs(pm - pf).max())
        out[f"{key}final_norm_err"] = float(np.abs(pm[:, -1] - pf[:, -1]).max())
        out[f"{key}final_energy_err_eV"] = float(np.abs(em - ef).max())
        out[f"{key}pop_err_bar_1e-6_met"] = out[f"{key}final_pop_err"] <= 1e-6
    return out


def _host_cpu():
    """CPU model (as lscpu prints it, from /proc/cpuinfo) and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def run_cpu_baseline(st, b, S, h, lumo, homo, steps):
    """Reference algorithm (Taylor.f Propagation, tolerance 1e-8, order 25) restated
    in C/OpenMP, timed on the host: `steps` whole 0.5 fs steps of electron + hole
    on all the host cores this process may use (OMP_NUM_THREADS), then the same
    GEMV on one thread, timed over a bounded sample of GEMVs and scaled by the
    GEMVs per wavepacket-step of the all-core run."""
    import scipy.linalg as sl
    from oracle import oracle as O
    N = b.N
    model, avail = _host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS", avail))
    Hp = np.asfortranarray(sl.solve(S, h, assume_a="pos"))
    brak = np.stack([S @ lumo, S @ homo], 1).astype(complex)
    kets = np.stack([lumo, homo], 1).astype(complex)
    tau_max = DT_PS / O.H_BAR
    save = [tau_max, tau_max]
    O.set_threads(cores)
    matvecs = 0
    t = 0.0
    t0 = time.perf_counter()
    for it in range(steps):
        for p in range(2):
            tau = tau_max if it == 0 else min(save[p] * 1.15, tau_max)
            brak[:, p], kets[:, p], _, save[p], stt = O.propagation(Hp, brak[:, p], kets[:, p], t, t + DT_PS, tau)
            matvecs += stt["matvec"]
        t += DT_PS
    el = time.perf_counter() - t0
    gemv_per_wp_step = matvecs / (2 * steps)
    host_gbs = matvecs * 8.0 * N * N / el / 1e9
    # one thread: the same GEMV (H' streamed once per call) over a bounded sample
    O.set_threads(1)
    x = kets[:, 0].copy()
    O.op_x_ket(Hp, x)
    n1 = 16
    t0 = time.perf_counter()
    for _ in range(n1):
        x = O.op_x_ket(Hp, x)
        x /= np.abs(x).max()
    el1 = time.perf_counter() - t0
    O.set_threads(cores)
    gemv1_s = el1 / n1
    log(f"cpu baseline: {cores} threads {2 * steps / el:.4f} wavepacket-steps/s ({matvecs} GEMVs, "
        f"{host_gbs:.1f} GB/s); 1 thread {1e3 * gemv1_s:.1f} ms per GEMV")
    return {"value": 2 * steps / el, "unit": "wavepacket-steps/s", "cores": cores, "kind": "port",
            "sample": f"Taylor propagation (reference algorithm restated in C/OpenMP, not the ifx/MKL build; "
                      f"tol 1e-8, order 25) of electron + hole for {steps} x 0.5 fs steps at N={N} on the TiO2 "
                      f"system, {cores} threads; S^-1 h setup untimed; {matvecs} GEMVs",
            "cpu_model": model, "cores_available": avail, "steps": steps, "wall_s": round(el, 3),
            "gemv_total": matvecs, "gemv_per_wp_step": round(gemv_per_wp_step, 1),
            "host_GBs": round(host_gbs, 2), "bytes_per_gemv": 8 * N * N,
            "single_thread": {"cores": 1, "gemv_ms": round(1e3 * gemv1_s, 3),
                              "host_GBs": round(8.0 * N * N / gemv1_s / 1e9, 2),
                              "value": 1.0 / (gemv1_s * gemv_per_wp_step), "unit": "wavepacket-steps/s",
                              "sample": f"{n1} GEMVs of the same path on one thread, scaled by the "
                                        f"{gemv_per_wp_step:.0f} GEMVs per wavepacket-step of the all-core run"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed nuclear steps (config 2 runs 2000)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ehrenfest", action="store_true",
                    help="skip the diabatic-Ehrenfest kernel timing after the drop-in steps")
    ap.add_argument("--batched-wp", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--many-steps", type=int, default=30,
                    help="steps of the 256-system batch leg of the default run (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=1.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-sweep", type=str, default="",
                    help="comma-separated arrival rates: run only the serving leg at each and print one JSON line per rate")
    ap.add_argument("--no-extra", action="store_true", help="only the FP64 drop-in leg")
    ap.add_argument("--term-variant", type=int, default=-1,
                    help="force one matvec variant instead of the autotuner's choice (for profiling: a "
                         "kernel tracer adds per-dispatch cost that skews the autotuner's timings)")
    ap.add_argument("--many-systems", type=int, default=0, metavar="STEPS",
                    help="only run the 256-system batch case for STEPS steps and print its JSON line")
    ap.add_argument("--precision-study", type=int, default=0, metavar="STEPS",
                    help="only run the N=8192, 64-wavepacket FP64 / mixed / FP32-stored precision case for STEPS "
                         "steps and print its JSON line")
    ap.add_argument("--precision-system", choices=("interface", "tio2"), default="interface",
                    help="system of --precision-study: the N=8192 interface cluster or the N=4096 TiO2 drop-in system")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the library's plain copy kernel over MIB MiB (5 launches), the "
                         "reference for scaling FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.many_systems:
        print(json.dumps(run_many_systems(steps=args.many_systems)), flush=True)
        return
    if args.precision_study:
        print(json.dumps(run_precision_study(steps=args.precision_study, system=args.precision_system)), flush=True)
        return
    if args.serve_sweep:
        for lam in (float(x) for x in args.serve_sweep.split(",")):
            log(f"serve sweep: rate {lam}/s")
            print(json.dumps(run_serve(args.serve_requests, lam, check=False)), flush=True)
        return
    if args.calibrate_copy > 0:
        from dynemol_amd import hip
        nbytes = args.calibrate_copy << 20
        ms = hip.run_copy(nbytes, 5)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 5,
                          "ms_per_copy": ms, "GBs": 2 * nbytes / ms / 1e6}))
        for mib in (32, 67, 134, 268, 1024):  # streaming-read ceiling at operator sizes
            rms = hip.run_read(mib * 1000 * 1000, 20)
            print(json.dumps({"read_bytes": mib * 1000 * 1000, "us_per_read": round(rms * 1e3, 2),
                              "GBs": round(mib * 1000 * 1000 / rms / 1e6, 1)}))
        return

    if args.term_variant >= 0:
        from dynemol_amd import hip
        hip.force_matvec_variant(args.term_variant)
    from dynemol_amd.structures import tio2_slab_adsorbate
    st = tio2_slab_adsorbate(seed=2)
    b = st.basis()
    N = b.N
    log(f"system {st.name}: {st.n_atoms} atoms, N={N}")
    S, h, lumo, homo, E = initial_states(st, b)
    log("initial states ready; GPU run")
    g = run_gpu(args, st, b, S, lumo, homo)

    R = 4  # complex columns: electron and hole, bra and ket each
    variant = g["tuning"]["term"]
    sym = variant in SYM_VARIANTS
    # algorithmic bytes of the kernel that ran: the symmetric variants read H's
    # lower-triangle tiles once (half storage), the dense ones all of H
    bytes_term = half_storage_bytes(N, R) if sym else term_bytes(N, R)
    avg_launch_s = g["series_ms"] / 1e3 / max(1, g["series_launches"])
    achieved = bytes_term / avg_launch_s / 1e9
    traffic = MEASURED_TRAFFIC_BYTES.get(variant) if N == 4096 else None
    value = 2 * args.steps / g["elapsed"]
    prop_only = g["prop_only"]
    # the same term alone (HIP events around repeated launches on a zero-filled
    # operator of the same shape, one accumulator, nothing beside it): the
    # figure a kernel trace's phase 1 + phase 2 averages compare with
    from dynemol_amd import hip
    alone_us = hip.time_matvec(N, R, variant, reps=50)
    out = {
        "metric": "wavepacket-steps/s at N=4096 (mixed vs FP64, pop.err<=1e-6) + serve p50/p99",
        "value": round(value, 3),
        "unit": "wavepacket-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * g["elapsed"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded TiO2 slab + C16H20 geometry, EHT parameters of the reference table)",
        "config": {"workload": "FP64 drop-in: TiO2 anatase slab + C16H20, N=4096, 2 wavepackets (electron LUMO, "
                               "hole HOMO), 0.5 fs steps, S/h/Cholesky/Lanczos rebuilt every 10 steps (jitter 0.02 A), "
                               "the step loop on the device (dh_propagate_steps)",
                   "n_orbitals": N, "wavepackets": 2, "rebuild_every": REBUILD_EVERY},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"Chebyshev term, FP64, 4 RHS (autotuned variant {variant}: "
                               f"{'lower-triangle tiles of symmetric H' if sym else 'dense H'})",
                     "bytes_per_launch": bytes_term, "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     # measured HBM bytes per term over the same time
                     "traffic_frac": round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     # the same time against a dense term's bytes (H read in full, 8 N^2 + 80 N R): the rate a
                     # dense-storage kernel would need to match this one, not an HBM rate
                     # the timed region's rebuilds run on a second stream beside the terms (dh_propagate_steps),
                     # so this average includes their interference
                     "dense_equivalent_bytes": term_bytes(N, R),
                     "dense_equivalent_frac": round(term_bytes(N, R) / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4),
                     "alone_us": round(alone_us, 3),
                     "alone_frac": round(bytes_term / (alone_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
        "propagation_only_value": round(prop_only, 3),
        # terms launched in the timed region per step (an interval's steps share recurrences)
        "chebyshev_terms_per_step": round(g["series_launches"] / args.steps, 1),
        "terms_last_recurrence": int(np.median(g["terms"])),
        # one operator build (S, h, Cholesky, L^-1, H_orth, bounds) on its own; in the timed region
        # the first build of the region runs so and the others beside the propagation
        "rebuild_ms": round(1e3 * g["rebuild_s"], 3),
        "rebuilds_in_timed_region": g["rebuilds"],
        "overlapped_rebuild_host_ms": round(g["side_build_ms"], 3),
        "spectral_bounds_eV": [round(x, 4) for x in g["bounds"][:2]],
        "ehrenfest": g["ehrenfest"],
        "eigenpairs": g["eigenpairs"],
        "batched": None,
        "many_systems": None,
        "serve": None,
    }
    if not args.no_extra:
        log("batched leg")
        out["batched"] = run_batched(st, b, S, args.batched_wp, args.batched_steps)
        out["batched"]["fp64_vs_dropin"] = round(out["batched"]["fp64_value"] / value, 3)
        out["batched"]["mixed_vs_dropin"] = round(out["batched"]["mixed_value"] / value, 3)
        out["batched"]["mixed_corrected_vs_dropin"] = round(out["batched"]["mixed_corrected_value"] / value, 3)
        if args.many_steps > 0:
            log("many-systems leg")
            out["many_systems"] = run_many_systems(steps=args.many_steps)
        if args.serve_requests > 0:
            log("serve leg")
            out["serve"] = run_serve(args.serve_requests, args.serve_rate)
    if not args.no_cpu and args.cpu_steps > 0:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline(st, b, S, h, lumo, homo, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
