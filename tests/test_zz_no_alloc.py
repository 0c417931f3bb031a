"""dh_chebyshev_apply_noalloc against dh_chebyshev_apply: the allocation-free
entry point runs on the library's static per-device scratch and must give the
workspace entry point's result (bit for bit: same kernels, same shape) on two
calls in a row, the second with a smaller system, and within 1e-9 of the exact
propagator (the FP64 bar).  Shapes beyond the static scratch are refused."""
import numpy as np
import pytest

from dynemol_amd import hip
from dynemol_amd.structures import ch_cluster, tio2_cluster

pytestmark = pytest.mark.gpu

DT = 5.0e-4


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _case(ctx, st, R, seed):
    b = st.basis()
    s = hip.System(ctx, b, n_wp=1, n_frag=len(st.fragment_names))
    s.build(st.coords)
    Ho = s.matrix(hip.DH_MAT_HORTH)
    bounds = s.bounds()[:2].copy()
    s.close()
    N = b.N
    ld = (N + 63) // 64 * 64
    Hp = np.eye(ld)
    Hp[:N, :N] = Ho
    rng = np.random.default_rng(seed)
    phi = np.zeros((ld, R), dtype=np.complex128)
    phi[:N] = rng.normal(size=(N, R)) + 1j * rng.normal(size=(N, R))
    phi[:N] /= np.linalg.norm(phi[:N], axis=0)
    return N, ld, np.asfortranarray(Hp), Ho, bounds, phi


def test_noalloc_matches_workspace_entry_point_twice():
    with hip.Context(0) as ctx:
        # larger system first, then a smaller one: the scratch keeps the first call's contents
        for st, R, seed in ((tio2_cluster(seed=6), 4, 1), (ch_cluster(seed=1), 2, 2)):
            N, ld, Hp, Ho, bounds, phi = _case(ctx, st, R, seed)
            _say(f"no-alloc vs workspace: N={N} R={R}")
            wb = hip.chebyshev_work_bytes(N, R)
            Hd, Pa, Pb, Wd = (hip.DeviceBuffer(n) for n in (Hp.nbytes, phi.nbytes, phi.nbytes, wb))
            Hd.upload(Hp)
            Pa.upload(phi)
            Pb.upload(phi)
            ka = hip.chebyshev_apply(ctx, N, Hd.ptr, ld, bounds, DT, R, Pa.ptr, Wd.ptr, wb)
            kb = hip.chebyshev_apply_noalloc(ctx, N, Hd.ptr, ld, bounds, DT, R, Pb.ptr)
            a = Pa.download((ld, R), np.complex128)
            bb = Pb.download((ld, R), np.complex128)
            assert ka == kb
            assert np.array_equal(a, bb)
            w, V = np.linalg.eigh(Ho)
            ref = V @ (np.exp(-1j * w * DT / hip.H_BAR)[:, None] * (V.T @ phi[:N]))
            err = np.abs(bb[:N] - ref).max() / np.abs(ref).max()
            _say(f"  {kb} terms, identical to the workspace call, {err:.2e} from exact")
            assert err <= 1e-9
            for buf in (Hd, Pa, Pb, Wd):
                buf.close()


def test_noalloc_refuses_shapes_beyond_its_scratch():
    with hip.Context(0) as ctx:
        N, R = 4096, 8  # 23 MB of work, beyond the 12 MiB array
        assert hip.chebyshev_work_bytes(N, R) > 12 << 20
        dummy = hip.DeviceBuffer(256)
        with pytest.raises(hip.DHError, match="beyond"):
            hip.chebyshev_apply_noalloc(ctx, N, dummy.ptr, N, np.array([-1.0, 1.0]), DT, R, dummy.ptr)
        dummy.close()
