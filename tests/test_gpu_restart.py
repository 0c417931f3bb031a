"""Device state through the reference's restart file (config 1, N = 416).

A run stopped after 3 steps, written with Security_Copy_Cheb's layout, read back
and restored into a fresh system continues as the uninterrupted run: states to
1e-10 relative and populations to 1e-12 after 2 more steps.  The snapshot's
arrays satisfy the definitions DUAL_bra = conj(Psi_ket), DUAL_ket = Psi_bra,
AO_bra = conj(S^-1 Psi_bra), AO_ket = Psi_ket (1e-10 relative).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from dynemol_amd import hip
from dynemol_amd import restart as RS
from dynemol_amd.structures import ch_cluster
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 5.0e-4


def test_restart_file_continues_the_run(tmp_path):
    print("[gpu] restart through Security_copy.dat, c1 N=416", flush=True)
    st = ch_cluster(seed=1)
    b = st.basis()
    S = O.overlap(st.coords, b)
    h = O.huckel(b, S)
    E, C = sl.eigh(h, S)
    rng = np.random.default_rng(21)
    with hip.Context(0) as ctx:
        a = hip.System(ctx, b, n_wp=2, n_frag=2)
        a.build(st.coords)
        for w, lo in enumerate((208, 196)):
            c = C[:, lo:lo + 10] @ (rng.normal(size=10) + 1j * rng.normal(size=10))
            a.set_wavepacket(w, S @ c, c)
        t = 0.0
        for _ in range(3):
            a.propagate(t, t + DT)
            t += DT
        dual_bra, dual_ket, ao_bra, ao_ket = RS.snapshot(a)
        for n in range(2):
            psi_bra, psi_ket = a.wavepacket(n)
            assert np.array_equal(dual_bra[:, n], np.conj(psi_ket))
            assert np.array_equal(dual_ket[:, n], psi_bra)
            ref = np.conj(np.linalg.solve(S, psi_bra))
            assert np.abs(ao_bra[:, n] - ref).max() <= 1e-10 * np.abs(ref).max()
        q = RS.net_charge(dual_bra, dual_ket, b.atom, st.n_atoms)
        path = tmp_path / "Security_copy.dat"
        RS.write_security_copy(str(path), dual_bra, dual_ket, ao_bra, ao_ket, t, 4, q)
        for _ in range(2):
            a.propagate(t, t + DT)
            t += DT
        want = [a.wavepacket(n) for n in range(2)]
        want_pops = [a.observables(n)[1] for n in range(2)]
        a.close()

        r = RS.read_restart_copy(str(path))
        assert r["it"] == 4 and r["eh_tag"] == ("el", "hl") and np.array_equal(r["net_charge"], q)
        bsys = hip.System(ctx, b, n_wp=2, n_frag=2)
        bsys.build(st.coords)
        RS.restore(bsys, r["dual_ket"], r["ao_ket"])
        t = r["t"]
        for _ in range(2):
            bsys.propagate(t, t + DT)
            t += DT
        for n in range(2):
            got = bsys.wavepacket(n)
            for g, w in zip(got, want[n]):
                assert np.abs(g - w).max() <= 1e-10 * np.abs(w).max()
            assert np.abs(bsys.observables(n)[1] - want_pops[n]).max() < 1e-12
        bsys.close()
