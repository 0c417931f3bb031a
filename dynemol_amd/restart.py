"""The reference's propagator output and restart files, for the device-resident state.

* `dump_qdyn` writes one step of dump_Qdyn (Taylor.f:244-285): `<tag>_survival.dat`
  (time, population of every fragment, total; FORMAT(F11.6,14F10.5) under the
  two header lines FORMAT(A,I9,14I10) and FORMAT(/15A10)) and `<tag>_wp_energy.dat`
  (time, Re and Im of the wavepacket energy; FORMAT(3F12.6)) for tags "el" and
  "hl".  Step 1 replaces the files and writes the headers; later steps append.
* `write_security_copy` / `read_restart_copy` are Security_Copy_Cheb and
  Restart_State_Cheb (backup.f:329-445): a Fortran unformatted sequential file of
  records [frame], it, t, N, n_part, size(eh_tag), eh_tag(:), then per particle
  (DUAL_bra_i, DUAL_ket_i) and (AO_bra_i, AO_ket_i) interleaved over the
  orbitals, then Net_Charge, put out size(Net_Charge) times over by the implied
  DO of backup.f:386 (n_atoms**2 values in one record).  Records are framed by 4-byte little-endian
  lengths, the record layout of gfortran and of ifort (the reference's compiler)
  on x86-64.
* `net_charge` is Net_Charge of Populations_mtx (data_output.f:132-138).
* `snapshot` / `restore` move the reference's DUAL/AO arrays out of and into a
  `hip.System`: after a step DUAL_bra = conj(Psi_ket), DUAL_ket = Psi_bra,
  AO_bra = conj(S^-1 Psi_bra), AO_ket = Psi_ket (ElHl_Chebyshev.f:272-279); a
  restart sets Psi_bra = DUAL_ket and Psi_ket = AO_ket
  (preprocess_from_restart, ElHl_Chebyshev.f:391-395).
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

EH_TAG = ("el", "hl")
CHARGE_SIGN = (-1.0, 1.0)  # [el, hl] (data_output.f:97)


# ---------------------------------------------------------------- Fortran edit descriptors
def _F(x: float, w: int, d: int) -> str:
    s = f"{x:.{d}f}"
    return "*" * w if len(s) > w else s.rjust(w)


def _I(n: int, w: int) -> str:
    s = str(int(n))
    return "*" * w if len(s) > w else s.rjust(w)


def _A(s: str, w: int) -> str:
    return s.rjust(w) if len(s) <= w else s[:w]


def survival_header(fragments) -> list[str]:
    """The two header records of <tag>_survival.dat: FORMAT(A,I9,14I10) with the
    column numbers 1..nf+2, then FORMAT(/15A10) with '#', the fragment names and
    'total' (the slash writes an empty record first)."""
    nums = list(range(1, len(fragments) + 3))
    line1 = "#" + _I(nums[0], 9) + "".join(_I(v, 10) for v in nums[1:])
    line2 = "".join(_A(s, 10) for s in ["#", *fragments, "total"])
    return [line1, "", line2]


def survival_record(dyn_row) -> str:
    """FORMAT(F11.6,14F10.5): time, then the fragment populations and the total."""
    return _F(dyn_row[0], 11, 6) + "".join(_F(v, 10, 5) for v in dyn_row[1:])


def energy_record(t: float, erg: complex) -> str:
    """FORMAT(3F12.6): time, Re and Im of QM_wp_erg."""
    return _F(t, 12, 6) + _F(erg.real, 12, 6) + _F(erg.imag, 12, 6)


def dump_qdyn(directory, it: int, dyn_it, wp_erg, fragments, eh_tag=EH_TAG):
    """One call of dump_Qdyn for step `it` (1-based).  dyn_it[n] = [t, pops of
    each fragment..., total] of particle n; wp_erg[n] its complex energy.
    Particles tagged "XX" are skipped, as in the reference."""
    os.makedirs(directory, exist_ok=True)
    for n, tag in enumerate(eh_tag):
        if tag == "XX":
            continue
        mode = "w" if it == 1 else "a"
        with open(os.path.join(directory, f"{tag}_survival.dat"), mode) as f:
            if it == 1:
                f.write("\n".join(survival_header(fragments)) + "\n")
            f.write(survival_record(dyn_it[n]) + "\n")
        with open(os.path.join(directory, f"{tag}_wp_energy.dat"), mode) as f:
            f.write(energy_record(dyn_it[n][0], complex(wp_erg[n])) + "\n")


# ---------------------------------------------------------------- unformatted records
def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def _read_records(data: bytes) -> list[bytes]:
    out, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        body = data[pos + 4:pos + 4 + n]
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        if m != n or len(body) != n:
            raise ValueError(f"corrupt record at byte {pos}: head {n}, tail {m}")
        out.append(body)
        pos += 8 + n
    return out


def _interleave(a, b) -> bytes:
    """(a_i, b_i, i = 1..N) of complex*16 values, as one record's payload."""
    x = np.empty((len(a), 2), dtype=np.complex128)
    x[:, 0] = a
    x[:, 1] = b
    return x.astype("<c16").tobytes()


def write_security_copy(path, dual_bra, dual_ket, ao_bra, ao_ket, t: float, it: int, net_charge,
                        eh_tag=EH_TAG, frame: int | None = None):
    """Security_Copy_Cheb (backup.f:329-400).  The state arrays are N x n_part."""
    dual_bra, dual_ket, ao_bra, ao_ket = (np.asarray(a, dtype=np.complex128).reshape(len(a), -1)
                                          for a in (dual_bra, dual_ket, ao_bra, ao_ket))
    N, n_part = ao_bra.shape
    recs = []
    if frame is not None:
        recs.append(struct.pack("<i", frame))
    recs += [struct.pack("<i", it), struct.pack("<d", t), struct.pack("<i", N), struct.pack("<i", n_part),
             struct.pack("<i", len(eh_tag)), "".join(_A(s, 2)[:2].ljust(2) for s in eh_tag[:n_part]).encode()]
    for j in range(n_part):
        recs.append(_interleave(dual_bra[:, j], dual_ket[:, j]))
        recs.append(_interleave(ao_bra[:, j], ao_ket[:, j]))
    # write(33) ( Net_Charge , i=1,size(Net_Charge) ) (backup.f:386): the implied
    # DO puts the whole array out size(Net_Charge) times, n_atoms**2 values in one record
    q = np.asarray(net_charge, dtype="<f8").ravel()
    recs.append(np.tile(q, len(q)).tobytes())
    with open(path, "wb") as f:
        for r in recs:
            f.write(_record(r))


def read_restart_copy(path, with_frame: bool = False) -> dict:
    """Restart_State_Cheb (backup.f:405-445)."""
    with open(path, "rb") as f:
        recs = _read_records(f.read())
    k = 0
    out = {}
    if with_frame:
        out["frame"] = struct.unpack("<i", recs[0])[0]
        k = 1
    out["it"] = struct.unpack("<i", recs[k])[0]
    out["t"] = struct.unpack("<d", recs[k + 1])[0]
    N = struct.unpack("<i", recs[k + 2])[0]
    n_part = struct.unpack("<i", recs[k + 3])[0]
    tags = recs[k + 5].decode()  # the first n_part of size(eh_tag) tags, character(2) each
    out["eh_tag"] = tuple(tags[2 * i:2 * i + 2] for i in range(len(tags) // 2))
    arrays = {name: np.zeros((N, n_part), dtype=np.complex128) for name in ("dual_bra", "dual_ket", "ao_bra", "ao_ket")}
    for j in range(n_part):
        d = np.frombuffer(recs[k + 6 + 2 * j], dtype="<c16").reshape(N, 2)
        a = np.frombuffer(recs[k + 7 + 2 * j], dtype="<c16").reshape(N, 2)
        arrays["dual_bra"][:, j], arrays["dual_ket"][:, j] = d[:, 0], d[:, 1]
        arrays["ao_bra"][:, j], arrays["ao_ket"][:, j] = a[:, 0], a[:, 1]
    out.update(arrays)
    # read(33) ( Net_Charge , i=1,size(Net_Charge) ) (backup.f:436) reads the array
    # size(Net_Charge) times over; what stays is the last copy
    raw = np.frombuffer(recs[k + 6 + 2 * n_part], dtype="<f8")
    n_atoms = math.isqrt(len(raw))
    if n_atoms * n_atoms != len(raw):
        raise ValueError(f"Net_Charge record holds {len(raw)} values, not n_atoms**2")
    out["net_charge"] = raw[len(raw) - n_atoms:].copy()
    return out


def net_charge(dual_bra, dual_ket, atom_of_orbital, n_atoms):
    """Net_Charge(a) = sum_n ChargeSign(n) |sum_{i on a} DUAL_bra_i,n DUAL_ket_i,n|
    (data_output.f:132-138); atom_of_orbital is 0-based."""
    dual_bra = np.asarray(dual_bra).reshape(len(dual_bra), -1)
    dual_ket = np.asarray(dual_ket).reshape(len(dual_ket), -1)
    q = np.zeros(n_atoms)
    for n in range(dual_bra.shape[1]):
        z = dual_bra[:, n] * dual_ket[:, n]
        s = np.bincount(atom_of_orbital, weights=z.real, minlength=n_atoms) \
            + 1j * np.bincount(atom_of_orbital, weights=z.imag, minlength=n_atoms)
        q += CHARGE_SIGN[n] * np.abs(s)
    return q


# ---------------------------------------------------------------- device state
def snapshot(system, n_part: int = 2):
    """(DUAL_bra, DUAL_ket, AO_bra, AO_ket), N x n_part each, of a hip.System's
    current state, as ElHl_Chebyshev leaves them after a step."""
    N = system.N
    dual_bra, dual_ket, ao_bra, ao_ket = (np.zeros((N, n_part), dtype=np.complex128) for _ in range(4))
    for n in range(n_part):
        psi_bra, psi_ket = system.wavepacket(n)
        ao, _, _ = system.observables(n)
        dual_bra[:, n] = np.conj(psi_ket)
        dual_ket[:, n] = psi_bra
        ao_bra[:, n] = np.conj(ao)
        ao_ket[:, n] = psi_ket
    return dual_bra, dual_ket, ao_bra, ao_ket


def restore(system, dual_ket, ao_ket):
    """Psi_bra = DUAL_ket, Psi_ket = AO_ket for every particle (preprocess_from_restart)."""
    dual_ket = np.asarray(dual_ket).reshape(system.N, -1)
    ao_ket = np.asarray(ao_ket).reshape(system.N, -1)
    for n in range(dual_ket.shape[1]):
        system.set_wavepacket(n, dual_ket[:, n], ao_ket[:, n])
