"""SALMON's on-disk formats (salmon_amd/formats.py) against what a Fortran
compiler writes with the reference's own I/O statements (the probe program
tests/fortran/formats_probe.f90, built with the image's flang), and the
rt.data layout against the reference testsuite's reader
(testsuites/002_bulk_rt/verification: np.loadtxt, row = iteration,
column 10-1 Ac_tot_z, 13-1 E_tot_z, 16-1 Jm_z; rt_energy column 3-1)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from salmon_amd import formats

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler in this image")
    d = tmp_path_factory.mktemp("fprobe")
    exe = str(d / "formats_probe")
    subprocess.check_call([FLANG, "-O1", os.path.join(HERE, "fortran", "formats_probe.f90"), "-o", exe], cwd=d)
    return d, exe


def _values(rng, n_rows):
    """Rows of a time (F16.8) and four values spanning the double range, with
    the rounding and sign cases of Ew.dEe."""
    special = [0.0, 1.0, -1.0, 0.1, 0.99999999999999995, 0.999999999999999, 9.999999999999995e-300,
               -2.5e-308, 1.7976931348623157e308, 5e-324, 123456789012345.6, 0.5, -0.05, 1e-10, 3.3e200, 1e100]
    out = []
    for r in range(n_rows):
        t = r * 0.1 - 3.0 if r % 7 else -1e-9 * r
        vals = [special[(4 * r + j) % len(special)] if r < 8 else
                float(rng.choice([-1, 1]) * 10.0 ** rng.uniform(-300, 300) * rng.random()) for j in range(4)]
        out += [t] + vals
    return np.array(out)


def test_wfn_and_rows_match_fortran(probe):
    d, exe = probe
    rng = np.random.default_rng(11)
    NL, NB, NK = 7, 3, 2
    zu = rng.standard_normal((NK, NB, NL)) + 1j * rng.standard_normal((NK, NB, NL))
    occ = rng.random((NK, NB))
    v = _values(rng, 40)
    (d / "gs_wfn_k").mkdir(exist_ok=True)
    with open(d / "in.bin", "wb") as f:
        f.write(struct.pack("<4i", NL, NB, NK, v.size))
        f.write(zu.tobytes())  # (NK, NB, NL) C order = zu(NL, NB, NK) Fortran order
        f.write(occ.tobytes())
        f.write(v.tobytes())
    subprocess.check_call([exe], cwd=d)
    # reading what Fortran wrote
    z2, o2 = formats.read_gs_wfn_k(str(d), NL, NB, NK)
    assert np.array_equal(z2, zu) and np.array_equal(o2, occ)
    z3, _ = formats.read_gs_wfn_k(str(d), NL, NB, NK, nb_occ=2)
    assert np.array_equal(z3, zu[:, :2])
    # writing the same bytes
    mine = d / "mine"
    formats.write_gs_wfn_k(str(mine), zu, occ)
    for name in ["occupation"] + ["wfn_gs_k%07d.wfn" % (k + 1) for k in range(NK)]:
        assert (mine / "gs_wfn_k" / name).read_bytes() == (d / "gs_wfn_k" / name).read_bytes(), name
    # the row format, field for field
    rows = (d / "rows.txt").read_text().splitlines()
    assert len(rows) == v.size // 5
    for i, line in enumerate(rows):
        assert formats._row(v[5 * i: 5 * i + 5]).rstrip() == line.rstrip(), (i, v[5 * i: 5 * i + 5])


def test_wrong_sizes_and_damaged_records_raise(tmp_path):
    zu = np.ones((1, 2, 5), np.complex128)
    formats.write_gs_wfn_k(str(tmp_path), zu, np.ones((1, 2)))
    with pytest.raises(ValueError, match="expected"):
        formats.read_gs_wfn_k(str(tmp_path), 6, 2, 1)
    p = tmp_path / "gs_wfn_k" / "wfn_gs_k0000001.wfn"
    raw = bytearray(p.read_bytes())
    raw[-4:] = struct.pack("<i", 7)
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="damaged"):
        formats.read_gs_wfn_k(str(tmp_path), 5, 2, 1)
    with pytest.raises(ValueError, match="match"):
        formats.write_gs_wfn_k(str(tmp_path), zu, np.ones((2, 2)))


def test_subrecords_round_trip(tmp_path, monkeypatch):
    """Records beyond the 4-byte length limit (shrunk here to 16 bytes):
    subrecords with negative leading lengths, read back whole."""
    monkeypatch.setattr(formats, "_MAX_SUB", 16)
    payload = bytes(range(50))
    with open(tmp_path / "r.bin", "wb") as f:
        formats.write_record(f, payload)
        formats.write_record(f, b"")
        formats.write_record(f, b"abc")
    raw = (tmp_path / "r.bin").read_bytes()
    heads, pos = [], 0
    for _ in range(6):
        (n,) = struct.unpack_from("<i", raw, pos)
        heads.append(n)
        pos += 8 + abs(n)
    assert heads == [-16, -16, -16, 2, 0, 3] and pos == len(raw)
    with open(tmp_path / "r.bin", "rb") as f:
        assert formats.read_record(f) == payload
        assert formats.read_record(f) == b""
        assert formats.read_record(f) == b"abc"


def test_rt_data_layout_as_the_testsuite_reads_it(tmp_path):
    rng = np.random.default_rng(2)
    niter, dt = 30, 0.1
    ac_ext = rng.standard_normal((niter + 3, 3))
    ac_tot = ac_ext + 0.01 * rng.standard_normal((niter + 3, 3))
    javt = rng.standard_normal((niter + 1, 3)) * 1e-4
    p = tmp_path / "Si_rt.data"
    formats.write_rt_data(str(p), dt, ac_ext, ac_tot, javt)
    data = np.loadtxt(p)
    assert data.shape == (niter + 1, 16)
    it = 17
    assert data[it, 0] == pytest.approx(it * dt, abs=1e-8)
    assert data[it, 10 - 1] == pytest.approx(ac_tot[it + 1, 2], rel=1e-14)
    assert data[it, 13 - 1] == pytest.approx(-(ac_tot[it + 2, 2] - ac_tot[it, 2]) / (2 * dt), rel=1e-14)
    assert data[it, 16 - 1] == pytest.approx(javt[it, 2], rel=1e-14)
    assert data[it, 4 - 1] == pytest.approx(ac_ext[it + 1, 2], rel=1e-14)
    cols, unit_of = formats.read_data(str(p))
    assert list(cols)[:4] == ["Time", "Ac_ext_x", "Ac_ext_y", "Ac_ext_z"] and len(cols) == 16
    assert np.allclose(cols["Jm_y"], javt[:, 1], rtol=1e-14, atol=0)
    assert unit_of["Jm_z"] == "a.u."
    # rt_energy.data: every 10th iteration, column 3 is Eall - Eall0
    e = -10.0 + rng.standard_normal(niter + 1) * 1e-3
    pe = tmp_path / "Si_rt_energy.data"
    formats.write_rt_energy_data(str(pe), dt, e, e[0], nstep_energy_calc=10)
    de = np.loadtxt(pe)
    assert de.shape == (4, 3)
    assert de[2, 3 - 1] == pytest.approx(e[20] - e[0], rel=1e-14)
    # A_eV_fs: the testsuite's unit names
    formats.write_rt_data(str(p), dt, ac_ext, ac_tot, javt, unit_system="A_eV_fs")
    _, unit_of = formats.read_data(str(p))
    assert unit_of["Time"] == "fs" and unit_of["E_tot_z"] == "V/Angstrom"
    with pytest.raises(ValueError, match="niter"):
        formats.write_rt_data(str(p), dt, ac_ext[:-1], ac_tot, javt)
