"""SALMON's on-disk formats on either side of the RT path.

Input: the ground-state orbitals that ARTED writes and reads in
``read_write_gs_wfn_k`` (src/ARTED/GS/io_gs_wfn_k.f90:36-76):
``<dir>/gs_wfn_k/occupation`` holds one unformatted sequential record of
``occ(NB, NK)`` (real(8)), and ``<dir>/gs_wfn_k/wfn_gs_k%07d.wfn`` (k counted
from 1) one record of ``zu_GS(NL, NB)`` (complex(8)) each.  The RT run
takes the first ``NBoccmax`` bands of each (io_gs_wfn_k.f90:90).

Output: ``<SYS>_rt.data`` and ``<SYS>_rt_energy.data``
(src/ARTED/control/control_sc.f90:493-604): '#' header lines naming the
columns, then one row per iteration 0..niter written with the Fortran
format ``(F16.8,99(1X,E23.15E3))``; the fields are E(t) = -(A(t+dt) -
A(t-dt))/(2 dt) (control_sc.f90:228-230) and the current javt(t).

Unformatted sequential records are those of gfortran and flang: a 4-byte
little-endian length before and after each record; a record longer than
2**31 - 1 bytes is split into subrecords whose leading length is negative
on every subrecord but the last.

The arrays here are in this package's layout: ``zu`` is (NK, NB, NL) in C
order, which is the reference's ``zu(NL, NB, NK)`` in Fortran order byte for
byte, and ``occ`` is (NK, NB), the reference's ``occ(NB, NK)``.
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np

_MAX_SUB = 2**31 - 1

# unit names and factors from atomic units (src/io/inputoutput.f90:21-23,
# 1410-1530): 'au' and 'A_eV_fs'
_AU_TIME_FS = 0.02418884326505
_AU_ENERGY_EV = 27.21138505
_AU_LENGTH_AA = 0.52917721067


def units(unit_system="au"):
    """{quantity: (name, factor from a.u.)} for the rt.data columns."""
    if unit_system == "au":
        return dict(time=("a.u.", 1.0), ac=("a.u.", 1.0), elec=("a.u.", 1.0), current=("a.u.", 1.0),
                    energy=("a.u.", 1.0))
    if unit_system == "A_eV_fs":
        t, e, l = _AU_TIME_FS, _AU_ENERGY_EV, _AU_LENGTH_AA
        return dict(time=("fs", t), ac=("fs*V/Angstrom", t * e / l), elec=("V/Angstrom", 51.42206707),
                    current=("1/fs*Angstrom^2", (l / t) / l**3), energy=("eV", e))
    raise ValueError(f"unit_system {unit_system!r} is neither 'au' nor 'A_eV_fs'")


# --- unformatted sequential records ------------------------------------------

def read_record(f):
    """The payload of the next record of an open binary file."""
    parts = []
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise EOFError("no record marker")
        (n,) = struct.unpack("<i", head)
        body = f.read(abs(n))
        (tail,) = struct.unpack("<i", f.read(4))
        if len(body) != abs(n) or abs(tail) != abs(n):
            raise ValueError(f"damaged record: marker {n}, trailer {tail}, {len(body)} bytes read")
        parts.append(body)
        if n >= 0:
            return b"".join(parts)


def write_record(f, payload):
    """One record (subrecords beyond 2**31 - 1 bytes)."""
    payload = memoryview(payload).cast("B")
    n = len(payload)
    chunks = [(i, min(i + _MAX_SUB, n)) for i in range(0, n, _MAX_SUB)] or [(0, 0)]
    for j, (a, b) in enumerate(chunks):
        last, first = j == len(chunks) - 1, j == 0
        ln = b - a
        f.write(struct.pack("<i", ln if last else -ln))
        f.write(payload[a:b])
        f.write(struct.pack("<i", ln if first else -ln))


# --- gs_wfn_k -----------------------------------------------------------------

def _wfn_dir(directory):
    return os.path.join(directory, "gs_wfn_k")


def wfn_file(directory, ik):
    """Path of the orbitals of k point ik (counted from 0 here, from 1 in the name)."""
    return os.path.join(_wfn_dir(directory), "wfn_gs_k%07d.wfn" % (ik + 1))


def write_gs_wfn_k(directory, zu, occ):
    """Writes zu (NK, NB, NL) complex and occ (NK, NB) as ARTED's gs_wfn_k."""
    zu = np.ascontiguousarray(zu, dtype=np.complex128)
    occ = np.ascontiguousarray(occ, dtype=np.float64)
    if zu.ndim != 3 or occ.shape != zu.shape[:2]:
        raise ValueError(f"zu {zu.shape} and occ {occ.shape} do not match as (NK, NB, NL) and (NK, NB)")
    os.makedirs(_wfn_dir(directory), exist_ok=True)
    with open(os.path.join(_wfn_dir(directory), "occupation"), "wb") as f:
        write_record(f, occ.tobytes())
    for ik in range(zu.shape[0]):
        with open(wfn_file(directory, ik), "wb") as f:
            write_record(f, zu[ik].tobytes())


def read_gs_wfn_k(directory, NL, NB, NK, nb_occ=None):
    """(zu (NK, nb_occ, NL), occ (NK, NB)) from ARTED's gs_wfn_k; nb_occ
    (default NB) keeps the first bands, as the RT start does."""
    nb_occ = NB if nb_occ is None else int(nb_occ)
    if not 0 < nb_occ <= NB:
        raise ValueError(f"nb_occ {nb_occ} outside 1..{NB}")
    with open(os.path.join(_wfn_dir(directory), "occupation"), "rb") as f:
        raw = read_record(f)
    if len(raw) != 8 * NB * NK:
        raise ValueError(f"occupation holds {len(raw)} bytes, expected {8 * NB * NK} for NB={NB} NK={NK}")
    occ = np.frombuffer(raw, np.float64).reshape(NK, NB).copy()
    zu = np.empty((NK, nb_occ, NL), np.complex128)
    for ik in range(NK):
        with open(wfn_file(directory, ik), "rb") as f:
            raw = read_record(f)
        if len(raw) != 16 * NL * NB:
            raise ValueError(f"{wfn_file(directory, ik)} holds {len(raw)} bytes, expected {16 * NL * NB}")
        zu[ik] = np.frombuffer(raw, np.complex128).reshape(NB, NL)[:nb_occ]
    return zu, occ


# --- rt.data / rt_energy.data --------------------------------------------------

def fortran_e(x, w=23, d=15, e=3):
    """x as Fortran's Ew.dEe edit descriptor writes it (0.ddd...E+eee)."""
    x = float(x)
    if not np.isfinite(x):
        s = "NaN" if np.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w)
    if x == 0.0:
        body = "0." + "0" * d + "E+" + "0" * e
        return (("-" if np.signbit(x) else "") + body).rjust(w)
    m, ex = ("%.*e" % (d - 1, abs(x))).split("e")
    digits = m.replace(".", "")
    ex = int(ex) + 1
    body = "0." + digits + ("E+" if ex >= 0 else "E-") + str(abs(ex)).rjust(e, "0")
    return (("-" if x < 0 else "") + body).rjust(w)


def _row(values):
    t, rest = values[0], values[1:]
    return ("%16.8f" % t) + "".join(" " + fortran_e(v) for v in rest) + "\n"


def _header_columns(cols):
    return "#" + "".join(f" {i + 1}:{name}[{unit}]" for i, (name, unit) in enumerate(cols)) + "\n"


def field_from_ac(ac, dt):
    """E(t) = -(A(t+dt) - A(t-dt))/(2 dt) for t = 0..niter from A on -1..niter+1."""
    ac = np.asarray(ac, dtype=np.float64)
    return -(ac[2:] - ac[:-2]) / (2.0 * dt)


def write_rt_data(path, dt, ac_ext, ac_tot, javt, unit_system="au"):
    """``<SYS>_rt.data``: ac_ext and ac_tot are (niter+3, 3) on iterations
    -1..niter+1 (the reference's Ac arrays), javt (niter+1, 3) on 0..niter."""
    ac_ext = np.asarray(ac_ext, dtype=np.float64)
    ac_tot = np.asarray(ac_tot, dtype=np.float64)
    javt = np.asarray(javt, dtype=np.float64)
    niter = javt.shape[0] - 1
    if ac_ext.shape != (niter + 3, 3) or ac_tot.shape != (niter + 3, 3) or javt.shape != (niter + 1, 3):
        raise ValueError(f"shapes ac_ext {ac_ext.shape}, ac_tot {ac_tot.shape}, javt {javt.shape}: "
                         "expected (niter+3, 3), (niter+3, 3), (niter+1, 3)")
    u = units(unit_system)
    e_ext, e_tot = field_from_ac(ac_ext, dt), field_from_ac(ac_tot, dt)
    cols = [("Time", u["time"][0])]
    for name, q in (("Ac_ext", "ac"), ("E_ext", "elec"), ("Ac_tot", "ac"), ("E_tot", "elec"), ("Jm", "current")):
        cols += [(f"{name}_{a}", u[q][0]) for a in "xyz"]
    with open(path, "w") as f:
        f.write("# Real time calculation\n")
        for key, text in (("Ac_ext", "External vector potential field"), ("E_ext", "External electric field"),
                          ("Ac_tot", "Total vector potential field"), ("E_tot", "Total electric field"),
                          ("Jm", "Matter current density")):
            f.write(f"# {key}: {text}\n")
        f.write(_header_columns(cols))
        for it in range(niter + 1):
            f.write(_row([it * dt * u["time"][1],
                          *(ac_ext[it + 1] * u["ac"][1]), *(e_ext[it] * u["elec"][1]),
                          *(ac_tot[it + 1] * u["ac"][1]), *(e_tot[it] * u["elec"][1]),
                          *(javt[it] * u["current"][1])]))


def write_rt_energy_data(path, dt, eall, eall0, nstep_energy_calc=1, unit_system="au"):
    """``<SYS>_rt_energy.data``: eall (niter+1,) on iterations 0..niter, rows
    every nstep_energy_calc iterations."""
    eall = np.asarray(eall, dtype=np.float64)
    u = units(unit_system)
    en = u["energy"][0]
    with open(path, "w") as f:
        f.write("# Real time calculation\n")
        f.write("# Eall: Total energy\n")
        f.write("# Eall0: Initial energy\n")
        f.write(_header_columns([("Time", u["time"][0]), ("Eall", en), ("Eall-Eall0", en)]))
        for it in range(eall.shape[0]):
            if it % nstep_energy_calc:
                continue
            f.write(_row([it * dt * u["time"][1], eall[it] * u["energy"][1], (eall[it] - eall0) * u["energy"][1]]))


_COL = re.compile(r"(\d+):([^\[\s]+)\[([^\]]*)\]")


def read_data(path):
    """{column name: values} and {column name: unit} of a SALMON ``.data``
    file (the numbered '#' header line names the columns)."""
    names, unit_of = None, {}
    with open(path) as f:
        for line in f:
            if not line.startswith("#"):
                break
            cols = _COL.findall(line)
            if cols:
                names = [c[1] for c in sorted(cols, key=lambda c: int(c[0]))]
                unit_of = {c[1]: c[2] for c in cols}
    if names is None:
        raise ValueError(f"{path}: no numbered column header")
    data = np.loadtxt(path, comments="#", ndmin=2)
    if data.shape[1] != len(names):
        raise ValueError(f"{path}: {data.shape[1]} columns for {len(names)} names")
    return {n: data[:, i] for i, n in enumerate(names)}, unit_of
