"""FP64 MFMA GEMM, blocked Cholesky and triangular inverse on the GPU against numpy."""
import numpy as np
import pytest

from dynemol_amd import hip

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.mark.parametrize("transA,transB", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_transposes(transA, transB):
    rng = np.random.default_rng(1)
    M, N, K = 96, 80, 70
    _say(f"gemm M={M} N={N} K={K} tA={transA} tB={transB}")
    A = rng.normal(size=(K, M) if transA else (M, K))
    B = rng.normal(size=(N, K) if transB else (K, N))
    C = rng.normal(size=(M, N))
    out = hip.run_gemm(A, B, C, alpha=0.5, beta=-2.0, transA=transA, transB=transB)
    ref = 0.5 * ((A.T if transA else A) @ (B.T if transB else B)) - 2.0 * C
    assert np.abs(out - ref).max() < 1e-12


def test_gemm_single_tile_layout():
    # one 64x64 tile with K = 4: checks the MFMA operand/accumulator lane layout directly
    _say("gemm layout 64x64x4")
    A = np.arange(64 * 4, dtype=np.float64).reshape(64, 4) / 100.0
    B = np.arange(4 * 64, dtype=np.float64).reshape(4, 64)[:, ::-1] / 50.0
    out = hip.run_gemm(A, B, np.zeros((64, 64)))
    assert np.abs(out - A @ B).max() < 1e-12


def test_gemm_lower_only_and_kmodes():
    rng = np.random.default_rng(2)
    n = 192
    _say("gemm kmodes")
    L = np.tril(rng.normal(size=(n, n)))
    B = rng.normal(size=(n, n))
    # A lower triangular: k < row tile end
    out = hip.run_gemm(L, B, np.zeros((n, n)), kmode=1)
    assert np.abs(out - L @ B).max() < 1e-11
    # B^T with B lower: opB(k, j) = L[j, k] nonzero for k <= j
    out = hip.run_gemm(B, L, np.zeros((n, n)), transB=1, kmode=2)
    assert np.abs(out - B @ L.T).max() < 1e-11
    # A^T with A lower: opA(i, k) = L[k, i] nonzero for k >= i
    out = hip.run_gemm(L, B, np.zeros((n, n)), transA=1, kmode=3)
    assert np.abs(out - L.T @ B).max() < 1e-11
    C0 = rng.normal(size=(n, n))
    out = hip.run_gemm(B, B, C0, alpha=-1.0, beta=1.0, transB=1, lower_only=1)
    ref = C0 - B @ B.T
    for bi in range(3):
        for bj in range(3):
            blk = (slice(64 * bi, 64 * bi + 64), slice(64 * bj, 64 * bj + 64))
            if bi >= bj:
                assert np.abs(out[blk] - ref[blk]).max() < 1e-11
            else:
                assert np.array_equal(out[blk], C0[blk])


@pytest.mark.parametrize("transA,transB", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_large_tiles(transA, transB):
    """M, N >= 256 take the 128 x 128 pipelined kernel; ragged edges and odd K."""
    rng = np.random.default_rng(3)
    M, N, K = 300, 280, 70
    _say(f"gemm128 M={M} N={N} K={K} tA={transA} tB={transB}")
    A = rng.normal(size=(K, M) if transA else (M, K))
    B = rng.normal(size=(N, K) if transB else (K, N))
    C = rng.normal(size=(M, N))
    out = hip.run_gemm(A, B, C, alpha=0.5, beta=-2.0, transA=transA, transB=transB)
    ref = 0.5 * ((A.T if transA else A) @ (B.T if transB else B)) - 2.0 * C
    assert np.abs(out - ref).max() < 1e-12


def test_gemm_large_tiles_kmodes_and_lower_only():
    rng = np.random.default_rng(4)
    n = 448
    _say("gemm128 kmodes")
    L = np.tril(rng.normal(size=(n, n)))
    B = rng.normal(size=(n, n))
    assert np.abs(hip.run_gemm(L, B, np.zeros((n, n)), kmode=1) - L @ B).max() < 1e-11
    assert np.abs(hip.run_gemm(B, L, np.zeros((n, n)), transB=1, kmode=2) - B @ L.T).max() < 1e-11
    assert np.abs(hip.run_gemm(L, B, np.zeros((n, n)), transA=1, kmode=3) - L.T @ B).max() < 1e-11
    # B lower: opB(k, j) = L[k, j] nonzero for k >= j
    assert np.abs(hip.run_gemm(B, L, np.zeros((n, n)), kmode=4) - B @ L).max() < 1e-11
    C0 = rng.normal(size=(n, n))
    out = hip.run_gemm(B, B, C0, alpha=-1.0, beta=1.0, transB=1, lower_only=1)
    ref = C0 - B @ B.T
    for bi in range(n // 64):
        for bj in range(n // 64):
            blk = (slice(64 * bi, 64 * bi + 64), slice(64 * bj, 64 * bj + 64))
            if bi >= bj:
                assert np.abs(out[blk] - ref[blk]).max() < 1e-11
            else:
                assert np.array_equal(out[blk], C0[blk])


@pytest.mark.parametrize("n", [64, 192, 448, 1024, 1216])
def test_cholesky_and_inverse(n):
    rng = np.random.default_rng(n)
    _say(f"cholesky n={n}")
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    L, Li, info = hip.run_cholesky(S)
    assert info == 0
    assert np.array_equal(np.triu(L, 1), np.zeros((n, n)))
    assert np.abs(L - np.linalg.cholesky(S)).max() < 1e-12
    assert np.abs(np.tril(Li) @ L - np.eye(n)).max() < 1e-12


def test_cholesky_reports_non_spd():
    n = 128
    _say("cholesky non-SPD")
    S = np.eye(n)
    S[70, 70] = -1.0
    _, _, info = hip.run_cholesky(S)
    assert info == 71


@pytest.mark.parametrize("variant", list(range(22)))
@pytest.mark.parametrize("n_wp", [1, 2, 3, 4])
def test_matvec_variants_propagate_exactly(variant, n_wp):
    """Every matvec kernel variant the autotuner may pick gives the exact propagator."""
    rng = np.random.default_rng(10 + n_wp)
    n = 300
    _say(f"variant {variant} n_wp {n_wp} N={n}")
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    A = rng.normal(size=(n, n))
    h = (A + A.T) * 2.0
    import scipy.linalg as sl
    E, C = sl.eigh(h, S)
    dt = 5e-4
    hip.force_matvec_variant(variant)
    try:
        ctx = hip.Context(0)
        s = hip.System(ctx, None, n_wp=n_wp, n_frag=1, n_orb=n)
        s.load_operators(S, h)
        kets = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(n_wp)]
        bras = [S @ k for k in kets]
        for w in range(n_wp):
            s.set_wavepacket(w, bras[w], kets[w])
        s.propagate(0.0, dt)
        for w in range(n_wp):
            b, k = s.wavepacket(w)
            ke = C @ (np.exp(-1j * E * dt / hip.H_BAR) * (C.T @ (S @ kets[w])))
            assert np.abs(k - ke).max() <= 1e-10 * np.abs(ke).max()
            assert np.abs(b - S @ ke).max() <= 1e-10 * np.abs(S @ ke).max()
        s.close()
        ctx.close()
    finally:
        hip.force_matvec_variant(-1)


@pytest.mark.parametrize("variant", [16, 17, 18, 19, 20, 21])
@pytest.mark.parametrize("n,n_wp", [(1100, 1), (1100, 2), (1100, 4), (1200, 8)])
def test_symmetric_half_storage_variant_large(n, n_wp, variant):
    """Variants 16 (vector FMAs), 17 to 21 (matrix-core tile products): lower-triangle
    tiles, two-phase, over many block rows and segments (also loaded up front); even and
    odd block counts (18 and 19 blocks of 64)."""
    rng = np.random.default_rng(n + n_wp)
    _say(f"symmetric variant {variant} N={n} n_wp={n_wp}")
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    A = rng.normal(size=(n, n))
    h = (A + A.T) * 2.0
    import scipy.linalg as sl
    E, C = sl.eigh(h, S)
    dt = 5e-4
    hip.force_matvec_variant(variant)
    try:
        ctx = hip.Context(0)
        s = hip.System(ctx, None, n_wp=n_wp, n_frag=1, n_orb=n)
        s.load_operators(S, h)
        kets = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(n_wp)]
        for w in range(n_wp):
            s.set_wavepacket(w, S @ kets[w], kets[w])
        s.propagate(0.0, dt)
        for w in range(n_wp):
            b, k = s.wavepacket(w)
            ke = C @ (np.exp(-1j * E * dt / hip.H_BAR) * (C.T @ (S @ kets[w])))
            assert np.abs(k - ke).max() <= 1e-10 * np.abs(ke).max()
            assert np.abs(b - S @ ke).max() <= 1e-10 * np.abs(S @ ke).max()
        s.close()
        ctx.close()
    finally:
        hip.force_matvec_variant(-1)
