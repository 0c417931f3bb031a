"""The reference's output and restart file formats (host side, no GPU).

Expected text lines are written out from the Fortran edit descriptors of
Taylor.f:280-283 (Fw.d: right-justified in w columns with d decimals; Iw:
right-justified integer; Aw: right-justified string; a leading '/' writes an
empty record).  The unformatted snapshot is checked byte for byte against the
sequential-record layout (4-byte length, payload, 4-byte length).
"""
import struct

import numpy as np

from dynemol_amd import restart as RS


def test_survival_and_energy_lines():
    assert RS.survival_header(["C", "H"]) == [
        "#        1         2         3         4",
        "",
        "         #         C         H     total",
    ]
    assert RS.survival_record([0.0005, 0.25, 0.75, 1.0]) == "   0.000500   0.25000   0.75000   1.00000"
    assert RS.survival_record([12.5, -0.000004, 1.0]) == "  12.500000  -0.00000   1.00000"
    assert RS.energy_record(0.5, complex(-3.2, 0.001)) == "    0.500000   -3.200000    0.001000"
    assert RS._F(123456.0, 10, 5) == "*" * 10  # too wide for the field


def test_dump_qdyn_files(tmp_path):
    d = tmp_path / "dyn.trunk"
    for it, t in ((1, 0.0), (2, 0.0005)):
        RS.dump_qdyn(str(d), it, [[t, 0.9, 0.1, 1.0], [t, 0.2, 0.8, 1.0]], [complex(-1.5, 0), complex(-9.25, 1e-7)],
                     ["C", "H"])
    el = (d / "el_survival.dat").read_text().splitlines()
    assert el[:3] == RS.survival_header(["C", "H"])
    assert el[3:] == ["   0.000000   0.90000   0.10000   1.00000", "   0.000500   0.90000   0.10000   1.00000"]
    hl = (d / "hl_wp_energy.dat").read_text().splitlines()
    assert hl == ["    0.000000   -9.250000    0.000000", "    0.000500   -9.250000    0.000000"]


def test_security_copy_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    N = 3
    arrs = [rng.normal(size=(N, 2)) + 1j * rng.normal(size=(N, 2)) for _ in range(4)]
    q = rng.normal(size=5)
    path = tmp_path / "Security_copy.dat"
    RS.write_security_copy(str(path), *arrs, t=0.125, it=7, net_charge=q, frame=3)
    raw = path.read_bytes()
    payloads = [4, 4, 8, 4, 4, 4, 4, N * 32, N * 32, N * 32, N * 32, 5 * 5 * 8]
    assert len(raw) == sum(8 + p for p in payloads)
    # frame, it and t records, framed by their lengths
    assert raw[:12] == struct.pack("<iii", 4, 3, 4)
    assert raw[12:24] == struct.pack("<iii", 4, 7, 4)
    assert raw[24:40] == struct.pack("<idi", 8, 0.125, 8)
    # the tag record holds "elhl"; the first state record interleaves DUAL_bra_i, DUAL_ket_i
    off = sum(8 + p for p in payloads[:6])  # the tag record starts here
    assert raw[off:off + 12] == struct.pack("<i", 4) + b"elhl" + struct.pack("<i", 4)
    off += 12
    assert raw[off:off + 4] == struct.pack("<i", N * 32)
    first = np.frombuffer(raw[off + 4:off + 4 + N * 32], dtype="<c16")
    assert first[0] == arrs[0][0, 0] and first[1] == arrs[1][0, 0] and first[2] == arrs[0][1, 0]
    back = RS.read_restart_copy(str(path), with_frame=True)
    assert back["frame"] == 3 and back["it"] == 7 and back["t"] == 0.125 and back["eh_tag"] == ("el", "hl")
    for name, a in zip(("dual_bra", "dual_ket", "ao_bra", "ao_ket"), arrs):
        assert np.array_equal(back[name], a)
    assert np.array_equal(back["net_charge"], q)


def test_net_charge():
    # orbitals 0, 1 on atom 0 and orbital 2 on atom 1; electron on atom 0, hole split
    atom = np.array([0, 0, 1])
    bra = np.array([[0.6, 0.5], [0.8, 0.0], [0.0, 0.5]], dtype=complex)
    ket = np.array([[0.6, 1.0], [0.8, 0.0], [0.0, 1.0]], dtype=complex)
    q = RS.net_charge(bra, ket, atom, 2)
    assert np.allclose(q, [-(0.36 + 0.64) + 0.5, 0.5])


def _fortran_record(items):
    """One unformatted sequential record: the I/O list's values in order, framed by
    4-byte little-endian lengths."""
    body = b"".join(items)
    return struct.pack("<i", len(body)) + body + struct.pack("<i", len(body))


def test_security_copy_bytes_from_io_lists(tmp_path):
    """The whole file assembled item by item from Security_Copy_Cheb's I/O lists
    (backup.f:372-386): `( DUAL_bra(i,j) , DUAL_ket(i,j) , i=1,basis_size )` puts
    the two arrays out interleaved, and `( Net_Charge , i=1,size(Net_Charge) )`
    puts the whole Net_Charge array out once per i, n_atoms**2 reals."""
    rng = np.random.default_rng(5)
    N, n_part, n_atoms = 4, 2, 3
    db, dk, ab, ak = (rng.normal(size=(N, n_part)) + 1j * rng.normal(size=(N, n_part)) for _ in range(4))
    q = rng.normal(size=n_atoms)
    c16 = lambda z: struct.pack("<dd", z.real, z.imag)
    exp = _fortran_record([struct.pack("<i", 9)])
    exp += _fortran_record([struct.pack("<d", 0.0625)])
    exp += _fortran_record([struct.pack("<i", N)]) + _fortran_record([struct.pack("<i", n_part)])
    exp += _fortran_record([struct.pack("<i", 2)]) + _fortran_record([b"el", b"hl"])
    for j in range(n_part):
        exp += _fortran_record([c16(x) for i in range(N) for x in (db[i, j], dk[i, j])])
        exp += _fortran_record([c16(x) for i in range(N) for x in (ab[i, j], ak[i, j])])
    exp += _fortran_record([struct.pack("<d", v) for _i in range(n_atoms) for v in q])
    path = tmp_path / "Security_copy.dat"
    RS.write_security_copy(str(path), db, dk, ab, ak, t=0.0625, it=9, net_charge=q)
    assert path.read_bytes() == exp


def test_restart_reads_reference_layout_file(tmp_path):
    """A Restart_copy.dat written in the reference's layout (Net_Charge repeated
    n_atoms times; the copies need not agree) reads back as Restart_State_Cheb
    reads it: the last copy stays in Net_Charge (backup.f:436)."""
    N, n_atoms = 2, 3
    z = np.array([[1 + 2j], [3 - 1j]])
    copies = [np.array([0.1, 0.2, 0.3]), np.array([0.4, 0.5, 0.6]), np.array([-1.0, 0.25, 2.0])]
    c16 = lambda v: struct.pack("<dd", v.real, v.imag)
    raw = _fortran_record([struct.pack("<i", 4)]) + _fortran_record([struct.pack("<d", 1.5)])
    raw += _fortran_record([struct.pack("<i", N)]) + _fortran_record([struct.pack("<i", 1)])
    raw += _fortran_record([struct.pack("<i", 1)]) + _fortran_record([b"el"])
    raw += _fortran_record([c16(x) for i in range(N) for x in (z[i, 0], -z[i, 0])])
    raw += _fortran_record([c16(x) for i in range(N) for x in (2 * z[i, 0], 3 * z[i, 0])])
    raw += _fortran_record([struct.pack("<d", v) for c in copies for v in c])
    path = tmp_path / "Restart_copy.dat"
    path.write_bytes(raw)
    back = RS.read_restart_copy(str(path))
    assert back["it"] == 4 and back["t"] == 1.5 and back["eh_tag"] == ("el",)
    assert np.array_equal(back["dual_bra"], z) and np.array_equal(back["dual_ket"], -z)
    assert np.array_equal(back["ao_bra"], 2 * z) and np.array_equal(back["ao_ket"], 3 * z)
    assert np.array_equal(back["net_charge"], copies[-1])
    # and what this library writes is read back the same way
    RS.write_security_copy(str(tmp_path / "again.dat"), back["dual_bra"], back["dual_ket"], back["ao_bra"],
                           back["ao_ket"], t=1.5, it=4, net_charge=back["net_charge"], eh_tag=("el",))
    again = RS.read_restart_copy(str(tmp_path / "again.dat"))
    assert np.array_equal(again["net_charge"], copies[-1])
