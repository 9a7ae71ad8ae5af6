"""File formats on the CPU: the history file against the byte sequence the reference's own reader
(read_history.m) consumes, and the field-file round trip."""
import io
import os
import struct

import numpy as np

import dns_io


def _read_like_read_history_m(path):
    """The fread sequence of read_history.m (one file), restated: header, Ngal, then records while
    flag_ctpress stays 0 or 1."""
    f = open(path, "rb")
    rd = lambda fmt: struct.unpack("<" + fmt, f.read(struct.calcsize("<" + fmt)))
    rd("i")
    t_, Re, alp, bet, mpgx_ = rd("5d")
    (nband,) = rd("i")
    rd("2i")
    Ngal = np.array(rd("%di" % (4 * (nband + 2)))).reshape(nband + 2, 4).T
    rd("i")
    rd("i")
    raw = f.read(4)
    rows = []
    while len(raw) == 4:
        (flag,) = struct.unpack("<i", raw)
        names = dns_io.HIST_FIELDS[flag]
        (it,) = rd("i")
        vals = rd("%dd" % len(names))
        rd("i")                 # end of this record
        d = dict(zip(names, vals))
        d.update(flag_ctpress=flag, iter=it)
        rows.append(d)
        f.read(4)               # start of the next record (empty at the end of the file)
        raw = f.read(4)
    return dict(t=t_, Re=Re, alp=alp, bet=bet, mpgx=mpgx_, nband=nband), Ngal, rows


def test_hist_file_matches_reference_reader(tmp_path):
    N = np.zeros((4, 5), np.int32)
    N[0, 1:4] = 32
    N[0, 4] = -2
    N[1, 1:4] = 32
    N[2, 3] = 151
    N[3, 3] = 152
    head = dict(t=0.0, Re=3250.0, alp=1.0, bet=2.0, mpgx=-3.0675e-3, nband=3)
    for flag in (1, 0):
        names = dns_io.HIST_FIELDS[flag]
        recs = [dict(flag_ctpress=flag, iter=k, **{n: 0.5 * k + i * 1e-3 for i, n in enumerate(names)})
                for k in range(3)]
        path = os.path.join(str(tmp_path), "hist_%d.dat" % flag)
        dns_io.write_hist(path, head, N, recs[:2])
        dns_io.write_hist(path, head, N, recs[2:], append=True)
        h2, N2, rows = _read_like_read_history_m(path)
        assert h2 == head and np.array_equal(N2, N)
        assert rows == recs
        h3, N3, rows3 = dns_io.read_hist(path)
        assert h3 == head and np.array_equal(N3, N) and rows3 == recs


def test_field_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    Kx, Nz, nr = 9, 16, 5
    planes = rng.standard_normal((nr, Nz, Kx)) + 1j * rng.standard_normal((nr, Nz, Kx))
    N = np.zeros((4, 5), np.int32)
    N[:, 3] = [16, 16, 3, 4]
    y = np.linspace(-1, 1, nr)
    ff = dns_io.FieldFile(1.25, 3250.0, 1.0, 2.0, -3e-3, 3, 7, N, y, 1.0, y * 2, np.arange(nr), planes, y)
    path = os.path.join(str(tmp_path), "u1_x.dat")
    dns_io.write_field(path, ff)
    g = dns_io.read_field(path)
    assert (g.t, g.Re, g.iter, g.nband) == (1.25, 3250.0, 7, 3)
    assert np.array_equal(g.planes, planes) and np.array_equal(g.N, N) and np.array_equal(g.y, y)


def test_fold_spectrum_known_answer():
    """recvwrspec's folding (spectra.f90:142-165): +kz and -kz share a slot, i >= 1 and kz < Nz/2
    doubled, the kz = -Nz/2 row and i = 0 kept single."""
    Nx, Nz = 8, 6
    Kx = Nx // 2 + 1
    sp = np.zeros((Kx * Nz, 2))
    col = lambda i, kz: (kz if kz >= 0 else kz + Nz) * Kx + i
    sp[col(1, 1), 0] = 1.0
    sp[col(1, -1), 0] = 10.0
    sp[col(0, 0), 1] = 3.0
    sp[col(2, -3), 1] = 5.0           # kz = -Nz/2
    sp[col(0, 2), 1] = 7.0
    b = dns_io.fold_spectrum(sp, Nx, Nz)
    assert b.shape == (2, Nz // 2 + 1, Kx)
    assert b[0, 1, 1] == 22.0 and np.count_nonzero(b[0]) == 1
    assert b[1, 0, 0] == 3.0 and b[1, Nz // 2, 2] == 5.0 and b[1, 2, 0] == 7.0
    assert np.count_nonzero(b[1]) == 3
