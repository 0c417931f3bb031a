"""ctypes binding of the CPU oracle (liboracle.so) and the reference step sequence.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
restatement, never by the product package.

`elhl_reference_step` restates one call of ElHl_Chebyshev (ElHl_Chebyshev.f:
148-291) for both particles on one process: S and h rebuild, S^-1, H' = S^-1 h,
Taylor propagation of the electron and the hole, DUAL/AO outputs,
quasiparticle energies and populations.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

H_BAR = 6.58264e-4

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_l = ctypes.POINTER(ctypes.c_long)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_solap.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double, _d]
        L.orc_anor.restype = ctypes.c_double
        L.orc_anor.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_double]
        L.orc_anor_div.restype = ctypes.c_double
        L.orc_anor_div.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_rotar.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.orc_rotation_overlap.argtypes = [_d, _d, ctypes.c_int, _d, _d]
        L.orc_overlap.restype = ctypes.c_int
        L.orc_overlap.argtypes = [ctypes.c_int, _d, _i, _i, ctypes.c_int, _i, _i, _i, _i, _d, _d, ctypes.c_double,
                                  ctypes.c_double, _d]
        L.orc_huckel.argtypes = [ctypes.c_int, _d, _d, _d, _d, _d]
        L.orc_cholesky.restype = ctypes.c_int
        L.orc_cholesky.argtypes = [ctypes.c_int, _d, _d]
        L.orc_syinvert.restype = ctypes.c_int
        L.orc_syinvert.argtypes = [ctypes.c_int, _d, _d]
        L.orc_dgemm_nn.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_propagation.restype = ctypes.c_int
        L.orc_propagation.argtypes = [ctypes.c_int, _d, _d, _d, ctypes.c_double, ctypes.c_double, _d, _d,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, _l]
        L.orc_op_x_ket.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_bra_x_op.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_populations.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d]
        L.orc_qp_energy.argtypes = [ctypes.c_int, _d, _d, _d, _d]
        L.orc_huckel_x.argtypes = [ctypes.c_int, _d, _d, _d, _d]
        L.orc_ehrenfest_kernel.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _d, _d, _d]
        L.orc_util_init()
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def solap(na, la, za, nb, lb, zb, R):
    out = np.zeros(6)
    lib().orc_solap(na, la, za, nb, lb, zb, R, _p(out))
    return out


def anor(expa, expb, na, nb, la, lb):
    aux = 1.0 / (np.math.factorial(2 * na) * np.math.factorial(2 * nb)) if hasattr(np, "math") else None
    import math
    aux = 1.0 / (math.factorial(2 * na) * math.factorial(2 * nb))
    return lib().orc_anor(expa, expb, na, nb, la, lb, aux)


def anor_div(la, lb, m):
    return lib().orc_anor_div(la, lb, m)


def rotation(xa, xb, lmax):
    """Returns (Rab, rl) with rl[l][mp+l, m+l] for l <= lmax (real-harmonic rotation)."""
    xa = np.ascontiguousarray(xa, dtype=np.float64)
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    raw = np.zeros(11 * 11 * 6)
    R = np.zeros(1)
    lib().orc_rotation_overlap(_p(xa), _p(xb), lmax, _p(R), _p(raw))
    rl = raw.reshape(6, 11, 11)  # [l][m][mp] (Fortran column-major (mp, m, l))
    out = []
    for l in range(lmax + 1):
        blk = np.zeros((2 * l + 1, 2 * l + 1))
        for mp in range(-l, l + 1):
            for m in range(-l, l + 1):
                blk[mp + l, m + l] = rl[l, m + 5, mp + 5]
        out.append(blk)
    return float(R[0]), out


def overlap(coords, basis, cutoff=12.0, drop_below=0.0):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    N = basis.N
    S = np.zeros((N, N), order="F")
    b = basis
    zeta = np.ascontiguousarray(b.zeta)
    coef = np.ascontiguousarray(b.coef)
    rc = lib().orc_overlap(b.n_atoms, _p(coords), _p(b.first_orb, _i), _p(b.atom_lmax, _i), N, _p(b.n, _i),
                           _p(b.l, _i), _p(b.m, _i), _p(b.nzeta, _i), _p(zeta), _p(coef), cutoff, drop_below,
                           S.ctypes.data_as(_d))
    assert rc == 0
    return S


def huckel(basis, S):
    N = basis.N
    S = np.asfortranarray(S)
    h = np.zeros((N, N), order="F")
    lib().orc_huckel(N, _p(basis.ip), _p(basis.k_wh), _p(basis.v_shift), S.ctypes.data_as(_d),
                     h.ctypes.data_as(_d))
    return h


def huckel_x(basis):
    """X_ij with h_ij = X_ij S_ij (hamiltonians.f:33-63; Huckel_stuff, diabatic-Ehren.f:250-276)."""
    N = basis.N
    X = np.zeros((N, N), order="F")
    lib().orc_huckel_x(N, _p(basis.ip), _p(basis.k_wh), _p(basis.v_shift), X.ctypes.data_as(_d))
    return X


def ehrenfest_kernel(Hp, X, ao_bra, ao_ket):
    """K = X o A - H' A and A = (rho + rho^T)/2 (diabatic-Ehren.f:100-116, 366-392).
    ao_bra, ao_ket: N x n_part complex (electron column 0, hole column 1), ao_bra
    already conjugated as the reference passes it.  Returns (K, A)."""
    N = Hp.shape[0]
    b = np.asfortranarray(np.asarray(ao_bra, dtype=np.complex128).reshape(N, -1))
    k = np.asfortranarray(np.asarray(ao_ket, dtype=np.complex128).reshape(N, -1))
    assert b.shape == k.shape and b.shape[1] in (1, 2)
    Hp = np.asfortranarray(Hp, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    K = np.zeros((N, N), order="F")
    A = np.zeros((N, N), order="F")
    lib().orc_ehrenfest_kernel(N, b.shape[1], Hp.ctypes.data_as(_d), X.ctypes.data_as(_d), b.ctypes.data_as(_d),
                               k.ctypes.data_as(_d), A.ctypes.data_as(_d), K.ctypes.data_as(_d))
    return K, A


def cholesky(S):
    N = S.shape[0]
    S = np.asfortranarray(S)
    L = np.zeros((N, N), order="F")
    info = lib().orc_cholesky(N, S.ctypes.data_as(_d), L.ctypes.data_as(_d))
    if info:
        raise np.linalg.LinAlgError(f"S not positive definite at column {info}")
    return L


def syinvert(S):
    N = S.shape[0]
    S = np.asfortranarray(S)
    Si = np.zeros((N, N), order="F")
    info = lib().orc_syinvert(N, S.ctypes.data_as(_d), Si.ctypes.data_as(_d))
    if info:
        raise np.linalg.LinAlgError(f"S not positive definite at column {info}")
    return Si


def dgemm(A, B):
    N = A.shape[0]
    A = np.asfortranarray(A)
    B = np.asfortranarray(B)
    C = np.zeros((N, N), order="F")
    lib().orc_dgemm_nn(N, A.ctypes.data_as(_d), B.ctypes.data_as(_d), C.ctypes.data_as(_d))
    return C


def propagation(Hp, bra, ket, t_init, t_max, tau, err_tol=1e-8, norm_tol=1e-8, order=25):
    """Taylor.f Propagation.  Returns (bra, ket, tau, save_tau, stats)."""
    N = Hp.shape[0]
    Hp = np.asfortranarray(Hp)
    bra = np.ascontiguousarray(bra, dtype=np.complex128).copy()
    ket = np.ascontiguousarray(ket, dtype=np.complex128).copy()
    t = np.array([tau], dtype=np.float64)
    st = np.array([tau], dtype=np.float64)
    stats = np.zeros(4, dtype=np.int64)
    rc = lib().orc_propagation(N, Hp.ctypes.data_as(_d), bra.ctypes.data_as(_d), ket.ctypes.data_as(_d),
                               t_init, t_max, _p(t), _p(st), err_tol, norm_tol, order, _p(stats, _l))
    if rc != 0:
        raise RuntimeError("Taylor propagation did not converge")
    return bra, ket, float(t[0]), float(st[0]), {"matvec": int(stats[0]), "substeps": int(stats[1]),
                                                  "rescale": int(stats[2]), "shrink": int(stats[3])}


def set_threads(n):
    """OpenMP threads of the following oracle calls."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def op_x_ket(Hp, ket):
    """One Taylor-path GEMV, H' ket (the dzgemv of Matrix_math.f:225-301)."""
    Hp = np.asfortranarray(Hp)
    ket = np.ascontiguousarray(ket, dtype=np.complex128)
    out = np.zeros_like(ket)
    lib().orc_op_x_ket(Hp.shape[0], Hp.ctypes.data_as(_d), ket.ctypes.data_as(_d), out.ctypes.data_as(_d))
    return out


def populations(frag, n_frag, dual_bra, dual_ket):
    out = np.zeros(n_frag + 1)
    frag = np.ascontiguousarray(frag, dtype=np.int32)
    a = np.ascontiguousarray(dual_bra, dtype=np.complex128)
    b = np.ascontiguousarray(dual_ket, dtype=np.complex128)
    lib().orc_populations(len(a), n_frag, _p(frag, _i), a.ctypes.data_as(_d), b.ctypes.data_as(_d), _p(out))
    return out


def qp_energy(ao_bra, H, ao_ket):
    out = np.zeros(2)
    a = np.ascontiguousarray(ao_bra, dtype=np.complex128)
    b = np.ascontiguousarray(ao_ket, dtype=np.complex128)
    H = np.asfortranarray(H)
    lib().orc_qp_energy(len(a), a.ctypes.data_as(_d), H.ctypes.data_as(_d), b.ctypes.data_as(_d), _p(out))
    return complex(out[0], out[1])


# --------------------------------------------------------------------------- the reference step
class ElHlReference:
    """One-process restatement of ElHl_Chebyshev's electron and hole propagation.

    State mirrors the reference module variables: Psi_t_bra/ket (N x 2),
    save_tau(2), first_call_ (ElHl_Chebyshev.f:35-40).
    """

    def __init__(self, basis, n_frag, psi_bra, psi_ket, err_tol=1e-8, norm_tol=1e-8, order=25):
        self.basis = basis
        self.n_frag = n_frag
        self.psi_bra = np.array(psi_bra, dtype=np.complex128)
        self.psi_ket = np.array(psi_ket, dtype=np.complex128)
        self.save_tau = np.zeros(2)
        self.first_call = True
        self.err_tol, self.norm_tol, self.order = err_tol, norm_tol, order
        self.stats = []
        self.times = {"build": 0.0, "invert": 0.0, "propagate": 0.0, "outputs": 0.0}

    def step(self, coords, t, delta_t, it, frame_step=1, S=None, h=None, drop_below=0.0):
        """ElHl_Chebyshev.f:148-291.  Returns (t_new, outputs dict)."""
        t0 = time.perf_counter()
        if S is None:
            S = overlap(coords, self.basis, drop_below=drop_below)
        if h is None:
            h = huckel(self.basis, S)
        t1 = time.perf_counter()
        Sinv = syinvert(S)
        Hp = dgemm(Sinv, h)
        t2 = time.perf_counter()
        t_init = t
        t_max = delta_t * frame_step * (it - 1)
        tau_max = delta_t / H_BAR
        out = {}
        for p in range(2):
            tau = tau_max if self.first_call else self.save_tau[p] * 1.15
            tau = tau_max if tau > tau_max else tau
            b, k, tau, st, stats = propagation(Hp, self.psi_bra[:, p], self.psi_ket[:, p], t_init, t_max, tau,
                                               self.err_tol, self.norm_tol, self.order)
            self.psi_bra[:, p] = b
            self.psi_ket[:, p] = k
            self.save_tau[p] = st
            self.stats.append(stats)
        t3 = time.perf_counter()
        t_new = t_init + delta_t * frame_step
        dual_bra = np.conj(self.psi_ket)
        dual_ket = self.psi_bra.copy()
        ao_bra = np.conj(Sinv @ self.psi_bra)
        ao_ket = self.psi_ket.copy()
        erg = [qp_energy(ao_bra[:, p], h, ao_ket[:, p]) for p in range(2)]
        pops = np.stack([populations(self.basis.fragment, self.n_frag, dual_bra[:, p], dual_ket[:, p])
                         for p in range(2)], axis=1)
        t4 = time.perf_counter()
        self.first_call = False
        self.times["build"] += t1 - t0
        self.times["invert"] += t2 - t1
        self.times["propagate"] += t3 - t2
        self.times["outputs"] += t4 - t3
        out.update(dual_bra=dual_bra, dual_ket=dual_ket, ao_bra=ao_bra, ao_ket=ao_ket, erg=erg, pops=pops,
                   Hprime=Hp, S=S, h=h)
        return t_new, out
