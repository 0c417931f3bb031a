"""ctypes binding of libdynemol_hip.so (the C ABI declared in include/dynemol_hip.h).

There is no CPU fallback: if the library is missing or no GPU is visible, the
calls raise.  The classes mirror the reference's per-step sequence:
`System.build` ~ Overlap_Matrix + Build_Huckel (+ orthogonalisation),
`System.propagate` ~ Propagation for every wavepacket (Taylor.f:35-127),
`System.observables` ~ QuasiParticleEnergies + Populations.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DYNEMOL_HIP_LIB points the binding at another build of the same library
LIB_PATH = os.environ.get("DYNEMOL_HIP_LIB", os.path.join(HERE, "libdynemol_hip.so"))

DH_MAT_S, DH_MAT_H, DH_MAT_L, DH_MAT_LINV, DH_MAT_HORTH, DH_MAT_HPRIME = range(6)
DH_PREC_FP64, DH_PREC_MIXED, DH_PREC_F32_STORAGE, DH_PREC_MIXED_CORRECTED = range(4)
H_BAR = 6.58264e-4

_lib = None
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

SYMBOLS = [
    "dh_last_error", "dh_version", "dh_init", "dh_finalize", "dh_system_create", "dh_system_destroy",
    "dh_system_bytes", "dh_system_bytes_prec", "dh_set_coords", "dh_build_operators", "dh_load_operators", "dh_orthogonalize",
    "dh_get_matrix", "dh_get_bounds", "dh_set_lanczos", "dh_set_wavepacket", "dh_get_wavepacket", "dh_propagate",
    "dh_get_observables", "dh_get_stats", "dh_set_timing", "dh_set_steps_per_series", "dh_matvec_tuning", "dh_set_precision",
    "dh_propagation_elhl", "dh_batch_create", "dh_batch_destroy", "dh_batch_propagate", "dh_batch_propagate_steps", "dh_set_initial_state",
    "dh_set_graph", "dh_ehrenfest_kernel", "dh_eigenpairs", "dh_set_eigenstate", "dh_batch_set_graph",
    "dh_chebyshev_work_bytes", "dh_chebyshev_apply", "dh_chebyshev_apply_noalloc", "dh_propagate_steps",
    "dh_batch_tuning",
]
TEST_SYMBOLS = ["dh_test_gemm", "dh_test_cholesky", "dh_test_copy", "dh_test_force_matvec_variant",
                "dh_test_matvec_timing", "dh_test_read", "dh_test_device_alloc", "dh_test_device_free",
                "dh_test_memcpy", "dh_test_bessel_coefficients", "dh_test_force_batch_family"]

# Every live context and system is released before interpreter shutdown, while
# the HIP runtime is still up (a handle freed during finalisation aborts).
_LIVE = weakref.WeakSet()


@atexit.register
def _release_all():
    rank = {"Batch": 0, "System": 1}
    for obj in sorted(list(_LIVE), key=lambda o: rank.get(type(o).__name__, 2)):
        try:
            obj.close()
        except Exception:
            pass


class DHError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DHError(f"{LIB_PATH} is not built; run `python -m dynemol_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.dh_last_error.restype = ctypes.c_char_p
    L.dh_init.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.dh_finalize.argtypes = [_vp]
    L.dh_system_create.argtypes = [_vp, ctypes.c_int, _i, _i, ctypes.c_int, _i, _i, _i, _i, _d, _d, _d, _d, _d, _i,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.dh_system_destroy.argtypes = [_vp]
    L.dh_system_bytes.restype = ctypes.c_longlong
    L.dh_system_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.dh_system_bytes_prec.restype = ctypes.c_longlong
    L.dh_system_bytes_prec.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.dh_set_coords.argtypes = [_vp, _d]
    L.dh_build_operators.argtypes = [_vp, ctypes.c_double, ctypes.c_double]
    L.dh_load_operators.argtypes = [_vp, _d, _d, ctypes.c_int]
    L.dh_orthogonalize.argtypes = [_vp, ctypes.c_int]
    L.dh_get_matrix.argtypes = [_vp, ctypes.c_int, _d, ctypes.c_int]
    L.dh_get_bounds.argtypes = [_vp, _d]
    L.dh_set_lanczos.argtypes = [_vp, ctypes.c_int, ctypes.c_double]
    L.dh_set_wavepacket.argtypes = [_vp, ctypes.c_int, _d, _d]
    L.dh_get_wavepacket.argtypes = [_vp, ctypes.c_int, _d, _d]
    L.dh_propagate.argtypes = [_vp, ctypes.c_double, ctypes.c_double, _i]
    L.dh_get_observables.argtypes = [_vp, ctypes.c_int, _d, _d, _d]
    L.dh_propagate_steps.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _d, ctypes.c_double,
                                     ctypes.c_double, _d, _d, _i]
    L.dh_ehrenfest_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int]
    L.dh_eigenpairs.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _d, _d, ctypes.c_int]
    L.dh_batch_set_graph.argtypes = [_vp, ctypes.c_int]
    L.dh_set_eigenstate.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _d]
    L.dh_get_stats.argtypes = [_vp, _d]
    L.dh_propagation_elhl.argtypes = [_vp, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d]
    L.dh_test_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int, _d,
                               ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int]
    L.dh_test_cholesky.argtypes = [ctypes.c_int, _d, _d, _i]
    L.dh_test_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, _d]
    L.dh_test_read.argtypes = [ctypes.c_longlong, ctypes.c_int, _d]
    L.dh_test_bessel_coefficients.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _d,
                                              ctypes.POINTER(ctypes.c_int)]
    L.dh_set_timing.argtypes = [_vp, ctypes.c_int]
    L.dh_set_steps_per_series.argtypes = [_vp, ctypes.c_int]
    L.dh_set_graph.argtypes = [_vp, ctypes.c_int]
    L.dh_matvec_tuning.argtypes = [_vp, _i]
    L.dh_set_precision.argtypes = [_vp, ctypes.c_int]
    L.dh_set_initial_state.argtypes = [_vp, ctypes.c_int, _d, _d]
    L.dh_batch_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
    L.dh_batch_destroy.argtypes = [_vp]
    L.dh_batch_propagate.argtypes = [_vp, ctypes.c_double, ctypes.c_double, _i]
    L.dh_batch_propagate_steps.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _i]
    L.dh_batch_tuning.argtypes = [_vp, _i]
    L.dh_test_force_matvec_variant.argtypes = [ctypes.c_int]
    L.dh_test_matvec_timing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
    L.dh_chebyshev_work_bytes.restype = ctypes.c_longlong
    L.dh_chebyshev_work_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.dh_chebyshev_apply.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _d, ctypes.c_double, ctypes.c_int, _vp,
                                     _vp, ctypes.c_longlong, _i]
    L.dh_chebyshev_apply_noalloc.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _d, ctypes.c_double,
                                             ctypes.c_int, _vp, _i]
    L.dh_test_device_alloc.argtypes = [ctypes.c_longlong, ctypes.POINTER(_vp)]
    L.dh_test_device_free.argtypes = [_vp]
    L.dh_test_memcpy.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_int]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().dh_last_error().decode(errors="replace")
        raise DHError(f"{what} failed ({rc}): {msg}")


def _p(a, t=_d):
    return a.ctypes.data_as(t)


class Context:
    def __init__(self, device: int = 0):
        self.h = _vp()
        _check(lib().dh_init(device, ctypes.byref(self.h)), "dh_init")
        self.systems = weakref.WeakSet()
        _LIVE.add(self)

    def close(self):
        if self.h:
            for s in list(self.systems):
                s.close()
            lib().dh_finalize(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def propagation_elhl(self, S, h, psi_bra, psi_ket, t_init, t_max, want_hprime=True):
        """propagationelhl_gpucaller_ semantics; returns (psi_bra, psi_ket, ao_bra, H_prime, tau, save_tau)."""
        N = S.shape[0]
        S = np.asfortranarray(S, dtype=np.float64)
        h = np.asfortranarray(h, dtype=np.float64)
        bra = np.array(psi_bra, dtype=np.complex128, copy=True)
        ket = np.array(psi_ket, dtype=np.complex128, copy=True)
        ao = np.zeros(N, dtype=np.complex128)
        Hp = np.zeros((N, N), order="F") if want_hprime else None
        n = ctypes.c_int(N)
        ti, tm = ctypes.c_double(t_init), ctypes.c_double(t_max)
        tau, st = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(lib().dh_propagation_elhl(self.h, ctypes.byref(n), S.ctypes.data_as(_d), h.ctypes.data_as(_d),
                                         Hp.ctypes.data_as(_d) if Hp is not None else None,
                                         ao.ctypes.data_as(_d), None, bra.ctypes.data_as(_d), ket.ctypes.data_as(_d),
                                         ctypes.byref(ti), ctypes.byref(tm), ctypes.byref(tau), ctypes.byref(st)),
               "dh_propagation_elhl")
        return bra, ket, ao, Hp, tau.value, st.value


class System:
    """Device-resident operators and `n_wp` wavepackets of one molecular system."""

    def __init__(self, ctx: Context, basis=None, n_wp: int = 2, n_frag: int = 1, n_orb: int | None = None):
        self.ctx = ctx
        self.h = _vp()
        _LIVE.add(self)
        ctx.systems.add(self)
        self.n_wp = n_wp
        self.n_frag = n_frag
        if basis is None:
            assert n_orb is not None
            self.N = int(n_orb)
            self.n_atoms = 0
            _check(lib().dh_system_create(ctx.h, 0, None, None, self.N, None, None, None, None, None, None, None,
                                          None, None, None, n_frag, n_wp, ctypes.byref(self.h)), "dh_system_create")
            return
        b = basis
        self.basis = b
        self.N = b.N
        self.n_atoms = b.n_atoms
        self._keep = [np.ascontiguousarray(x) for x in (b.first_orb, b.atom_lmax, b.n, b.l, b.m, b.nzeta, b.zeta,
                                                         b.coef, b.ip, b.k_wh, b.v_shift, b.fragment)]
        fo, lm, n, l, m, nz, z, c, ip, kw, vs, fr = self._keep
        fr = fr.astype(np.int32)
        self._keep[-1] = fr
        _check(lib().dh_system_create(ctx.h, b.n_atoms, _p(fo, _i), _p(lm, _i), b.N, _p(n, _i), _p(l, _i), _p(m, _i),
                                      _p(nz, _i), _p(z), _p(c), _p(ip), _p(kw), _p(vs), _p(fr, _i), n_frag, n_wp,
                                      ctypes.byref(self.h)), "dh_system_create")

    def close(self):
        if self.h:
            lib().dh_system_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- operators
    def set_coords(self, coords):
        x = np.ascontiguousarray(coords, dtype=np.float64)
        assert x.shape == (self.n_atoms, 3)
        _check(lib().dh_set_coords(self.h, _p(x)), "dh_set_coords")

    def build(self, coords=None, cutoff=12.0, s_drop=0.0):
        if coords is not None:
            self.set_coords(coords)
        _check(lib().dh_build_operators(self.h, cutoff, s_drop), "dh_build_operators")

    def load_operators(self, S, h, want_hprime=False):
        S = np.asfortranarray(S, dtype=np.float64)
        h = np.asfortranarray(h, dtype=np.float64)
        _check(lib().dh_load_operators(self.h, S.ctypes.data_as(_d), h.ctypes.data_as(_d), S.shape[0]),
               "dh_load_operators")
        _check(lib().dh_orthogonalize(self.h, 1 if want_hprime else 0), "dh_orthogonalize")

    def matrix(self, which):
        out = np.zeros((self.N, self.N), order="F")
        _check(lib().dh_get_matrix(self.h, which, out.ctypes.data_as(_d), self.N), "dh_get_matrix")
        return out

    def bounds(self):
        out = np.zeros(6)
        _check(lib().dh_get_bounds(self.h, _p(out)), "dh_get_bounds")
        return out

    def set_lanczos(self, steps, rel_margin):
        _check(lib().dh_set_lanczos(self.h, steps, rel_margin), "dh_set_lanczos")

    # -- wavepackets
    def set_wavepacket(self, wp, psi_bra, psi_ket):
        b = np.ascontiguousarray(psi_bra, dtype=np.complex128)
        k = np.ascontiguousarray(psi_ket, dtype=np.complex128)
        assert b.shape == (self.N,) and k.shape == (self.N,)
        _check(lib().dh_set_wavepacket(self.h, wp, b.ctypes.data_as(_d), k.ctypes.data_as(_d)), "dh_set_wavepacket")

    def set_initial_state(self, wp, ao_bra, ao_ket):
        b = np.ascontiguousarray(ao_bra, dtype=np.complex128)
        k = np.ascontiguousarray(ao_ket, dtype=np.complex128)
        _check(lib().dh_set_initial_state(self.h, wp, b.ctypes.data_as(_d), k.ctypes.data_as(_d)),
               "dh_set_initial_state")

    def eigenpairs(self, il, iu, vectors=True):
        """Eigenpairs il..iu (1-based, ascending) of h C = S C E on the device: (E, C)."""
        n = iu - il + 1
        E = np.zeros(n)
        C = np.zeros((self.N, n), order="F") if vectors else None
        _check(lib().dh_eigenpairs(self.h, il, iu, _p(E), None if C is None else C.ctypes.data_as(_d), self.N),
               "dh_eigenpairs")
        return E, C

    def set_eigenstate(self, wp, k):
        """Seed wavepacket wp with eigenvector k (1-based) of (h, S); returns E_k."""
        e = ctypes.c_double(0.0)
        _check(lib().dh_set_eigenstate(self.h, wp, k, ctypes.byref(e)), "dh_set_eigenstate")
        return e.value

    def wavepacket(self, wp):
        b = np.zeros(self.N, dtype=np.complex128)
        k = np.zeros(self.N, dtype=np.complex128)
        _check(lib().dh_get_wavepacket(self.h, wp, b.ctypes.data_as(_d), k.ctypes.data_as(_d)), "dh_get_wavepacket")
        return b, k

    def propagate(self, t_init, t_max):
        n = ctypes.c_int(0)
        _check(lib().dh_propagate(self.h, t_init, t_max, ctypes.byref(n)), "dh_propagate")
        return n.value

    def propagate_steps(self, n_steps, dt, rebuild_every=0, geometries=None, cutoff=12.0, s_drop=0.0):
        """n_steps steps of dt (ps) on the device, rebuilding the operators from
        geometries[g] (n_atoms x 3) before step (g + 1) * rebuild_every.  Returns
        (pops [n_steps, n_wp, n_frag + 1], erg [n_steps, n_wp] complex, terms)."""
        n_geo = (n_steps - 1) // rebuild_every if rebuild_every > 0 else 0
        geo = None
        if n_geo > 0:
            geo = np.ascontiguousarray(np.asarray(geometries, dtype=np.float64)[:n_geo])
            assert geo.shape[0] == n_geo, f"need {n_geo} geometries"
        pops = np.zeros((n_steps, self.n_wp, self.n_frag + 1))
        erg = np.zeros((n_steps, self.n_wp, 2))
        n = ctypes.c_int(0)
        _check(lib().dh_propagate_steps(self.h, int(n_steps), float(dt), int(rebuild_every),
                                        None if geo is None else geo.ctypes.data_as(_d), cutoff, s_drop, _p(pops),
                                        _p(erg), ctypes.byref(n)), "dh_propagate_steps")
        return pops, erg[..., 0] + 1j * erg[..., 1], n.value

    def observables(self, wp):
        ao = np.zeros(self.N, dtype=np.complex128)
        pops = np.zeros(self.n_frag + 1)
        erg = np.zeros(2)
        _check(lib().dh_get_observables(self.h, wp, ao.ctypes.data_as(_d), _p(pops), _p(erg)), "dh_get_observables")
        return ao, pops, complex(erg[0], erg[1])

    def ehrenfest_kernel(self, wp_el=0, wp_hl=1, X=None, to_host=True):
        """Diabatic-Ehrenfest kernel K = X o A - H' A of the current state (N x N,
        Fortran order); X = None forms the Hueckel factors from the system's basis.
        to_host = False computes K on the device only and returns None."""
        if not to_host:
            xp = None if X is None else np.asfortranarray(X, dtype=np.float64)
            _check(lib().dh_ehrenfest_kernel(self.h, wp_el, wp_hl, None if xp is None else xp.ctypes.data_as(_d),
                                             self.N, None, self.N), "dh_ehrenfest_kernel")
            return None
        K = np.zeros((self.N, self.N), order="F")
        if X is None:
            xp, ldx = None, self.N
        else:
            X = np.asfortranarray(X, dtype=np.float64)
            assert X.shape == (self.N, self.N)
            xp, ldx = X.ctypes.data_as(_d), self.N
        _check(lib().dh_ehrenfest_kernel(self.h, wp_el, wp_hl, xp, ldx, K.ctypes.data_as(_d), self.N),
               "dh_ehrenfest_kernel")
        return K

    def stats(self):
        out = np.zeros(9)
        _check(lib().dh_get_stats(self.h, _p(out)), "dh_get_stats")
        return {"terms": int(out[0]), "steps": int(out[1]), "matvecs": int(out[2]), "retries": int(out[3]),
                "norm_drift": float(out[4]), "series_ms": float(out[5]), "series_launches": int(out[6]),
                "rebuilds": int(out[7]), "rebuild_ms_in_steps": float(out[8])}

    def tuning(self):
        out = np.zeros(7, dtype=np.int32)
        _check(lib().dh_matvec_tuning(self.h, _p(out, _i)), "dh_matvec_tuning")
        keys = ["term", "energy", "L_phi_bra", "LinvT_phi", "LT_ket", "Linv_bra", "lanczos"]
        return dict(zip(keys, out.tolist()))

    def set_precision(self, mode):
        """Term arithmetic: 0 / False = FP64, 1 / True = mixed (FP32 products),
        2 = FP32-stored H_orth with FP64 products, 3 = mixed with the FP64
        residual correction (DH_PREC_* constants)."""
        m = int(mode)
        _check(lib().dh_set_precision(self.h, m), "dh_set_precision")

    def set_timing(self, on=True):
        _check(lib().dh_set_timing(self.h, 1 if on else 0), "dh_set_timing")

    def set_steps_per_series(self, m):
        """Steps per Chebyshev recurrence in propagate_steps (0: automatic, 1: one per step)."""
        _check(lib().dh_set_steps_per_series(self.h, int(m)), "dh_set_steps_per_series")

    def set_graph(self, on=True):
        """Term series as one captured hipGraph per step (default) or as separate launches."""
        _check(lib().dh_set_graph(self.h, 1 if on else 0), "dh_set_graph")


def run_gemm(A, B, C, alpha=1.0, beta=0.0, transA=0, transB=0, kmode=0, lower_only=0):
    """Runs the library's FP64 MFMA GEMM kernel on host operands (column-major)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    M, N = C.shape
    K = A.shape[0] if transA else A.shape[1]
    _check(lib().dh_test_gemm(M, N, K, A.ctypes.data_as(_d), A.shape[0], B.ctypes.data_as(_d), B.shape[0],
                              C.ctypes.data_as(_d), M, alpha, beta, transA, transB, kmode, lower_only), "dh_test_gemm")
    return C


def run_cholesky(S):
    A = np.array(S, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    X = np.zeros_like(A)
    info = ctypes.c_int(0)
    _check(lib().dh_test_cholesky(n, A.ctypes.data_as(_d), X.ctypes.data_as(_d), ctypes.byref(info)),
           "dh_test_cholesky")
    return A, X, info.value


class Batch:
    """Many systems advanced together (one grouped launch per Chebyshev term)."""

    def __init__(self, ctx: Context, systems):
        self.ctx = ctx
        self.systems = list(systems)
        arr = (_vp * len(self.systems))(*[s.h for s in self.systems])
        self.h = _vp()
        _check(lib().dh_batch_create(ctx.h, len(self.systems), arr, ctypes.byref(self.h)), "dh_batch_create")
        _LIVE.add(self)

    def propagate(self, t_init, t_max):
        n = ctypes.c_int(0)
        _check(lib().dh_batch_propagate(self.h, t_init, t_max, ctypes.byref(n)), "dh_batch_propagate")
        return n.value

    def propagate_steps(self, n_steps, dt):
        """n_steps steps of dt (ps) for every member, several steps per recurrence;
        returns the most terms of one recurrence."""
        n = ctypes.c_int(0)
        _check(lib().dh_batch_propagate_steps(self.h, int(n_steps), float(dt), ctypes.byref(n)),
               "dh_batch_propagate_steps")
        return n.value

    def set_graph(self, on=True):
        """Term series as one captured hipGraph per step (default) or as separate launches."""
        _check(lib().dh_batch_set_graph(self.h, 1 if on else 0), "dh_batch_set_graph")

    def tuning(self):
        """Grouped term kernel the batch autotuner chose: 'symmetric' or 'dense'."""
        f = ctypes.c_int(0)
        _check(lib().dh_batch_tuning(self.h, ctypes.byref(f)), "dh_batch_tuning")
        return "symmetric" if f.value == 1 else "dense"

    def close(self):
        if self.h:
            lib().dh_batch_destroy(self.h)
            self.h = _vp()


def force_batch_family(R, family):
    _check(lib().dh_test_force_batch_family(int(R), int(family)), "dh_test_force_batch_family")


def force_matvec_variant(v):
    _check(lib().dh_test_force_matvec_variant(int(v)), "dh_test_force_matvec_variant")


def time_matvec(N, R, variant, reps=20):
    us = ctypes.c_double(0.0)
    _check(lib().dh_test_matvec_timing(N, R, variant, reps, ctypes.byref(us)), "dh_test_matvec_timing")
    return us.value


def bessel_coefficients(x, phase=0.0, tol=1e-16, kmax=4096):
    """The propagation's Chebyshev coefficients (host computation, no device
    work): complex array of length K + 2 (the last one 0) and K."""
    out = np.zeros(2 * (kmax + 2))
    K = ctypes.c_int(0)
    _check(lib().dh_test_bessel_coefficients(float(x), float(phase), float(tol), int(kmax),
                                             out.ctypes.data_as(_d), ctypes.byref(K)), "dh_test_bessel_coefficients")
    c = out[0::2] + 1j * out[1::2]
    return c[:K.value + 2], K.value


def run_read(nbytes, repeats=20):
    ms = ctypes.c_double(0.0)
    _check(lib().dh_test_read(int(nbytes), repeats, ctypes.byref(ms)), "dh_test_read")
    return ms.value


def run_copy(nbytes, repeats=20):
    ms = ctypes.c_double(0.0)
    _check(lib().dh_test_copy(int(nbytes), repeats, ctypes.byref(ms)), "dh_test_copy")
    return ms.value


def system_bytes(n_orb, n_wp, precision=DH_PREC_FP64):
    return int(lib().dh_system_bytes_prec(n_orb, n_wp, int(precision)))


def chebyshev_work_bytes(N, R):
    return int(lib().dh_chebyshev_work_bytes(int(N), int(R)))


def chebyshev_apply(ctx: Context, N, H_dev, ldh, bounds, dt, R, phi_dev, work_dev, work_bytes):
    """phi <- exp(-i H dt / hbar) phi on caller-owned device buffers (device
    addresses as ints); returns the Chebyshev term count."""
    b = np.ascontiguousarray(bounds[:2], dtype=np.float64)
    n = ctypes.c_int(0)
    _check(lib().dh_chebyshev_apply(ctx.h, int(N), _vp(H_dev), int(ldh), _p(b), float(dt), int(R), _vp(phi_dev),
                                    _vp(work_dev), int(work_bytes), ctypes.byref(n)), "dh_chebyshev_apply")
    return n.value


def chebyshev_apply_noalloc(ctx: Context, N, H_dev, ldh, bounds, dt, R, phi_dev):
    """chebyshev_apply with the library's static scratch instead of a work buffer."""
    b = np.ascontiguousarray(bounds[:2], dtype=np.float64)
    n = ctypes.c_int(0)
    _check(lib().dh_chebyshev_apply_noalloc(ctx.h, int(N), _vp(H_dev), int(ldh), _p(b), float(dt), int(R),
                                            _vp(phi_dev), ctypes.byref(n)), "dh_chebyshev_apply_noalloc")
    return n.value


class DeviceBuffer:
    """Device memory through the library's test hooks (for the entry points that
    take device pointers)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib().dh_test_device_alloc(self.nbytes, ctypes.byref(p)), "dh_test_device_alloc")
        self.ptr = p.value
        _LIVE.add(self)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().dh_test_memcpy(_vp(self.ptr), a.ctypes.data_as(_vp), a.nbytes, 1), "dh_test_memcpy")

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().dh_test_memcpy(out.ctypes.data_as(_vp), _vp(self.ptr), out.nbytes, 0), "dh_test_memcpy")
        return out

    def close(self):
        if self.ptr:
            lib().dh_test_device_free(_vp(self.ptr))
            self.ptr = None
