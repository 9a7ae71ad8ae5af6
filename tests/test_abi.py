"""CPU tests of the boundary: the C-ABI library loads and exports every declared symbol, host-side
set-up is bit-exact, argument errors come back as status codes, and the file/IC helpers work.
No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import dns_ic
import dns_io
import dnship
import pyoracle
from conftest import ROOT


def declared_symbols():
    with open(os.path.join(ROOT, "include", "dns_hip.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(dns_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = dnship.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(dnship.EXPORTS)


def test_fortran_interface_binds_every_entry_point():
    src = os.path.join(ROOT, "final-single-block-_amd", "fortran", "dns_hip_mod.f90")
    with open(src) as f:
        txt = f.read().lower()
    for n in declared_symbols():
        assert "name='%s'" % n.lower() in txt or 'name="%s"' % n.lower() in txt, n


@pytest.mark.parametrize("ny", [13, 17, 33, 153, 257, 385])
def test_host_ygrid_bit_exact_vs_oracle(ny):
    g = dnship.ygrid(ny)
    s = pyoracle.Oracle(24, 24, ny).setup()
    for k in ("yu", "yv", "dthdyu", "dthdyv"):
        assert np.array_equal(g[k], s[k]), k


def test_host_ygrid_bit_exact_vs_reference_fixture(golden):
    g = golden("golden_grid_153.npz")
    h = dnship.ygrid(153)
    for k in ("yu", "yv", "dthdyu", "dthdyv"):
        assert np.array_equal(h[k], g[k]), k


@pytest.mark.parametrize("Mx,Mz,ny,msg", [
    (50, 48, 33, "multiples of 6"),        # not a multiple of 6
    (60, 48, 33, "prime factor"),          # 60 = 2^2 3 5: FFTPACK's fatal factor (rft_buff.f:308-310)
    (48, 48, 4, "too small"),
])
def test_create_rejects_bad_sizes(Mx, Mz, ny, msg):
    with pytest.raises(dnship.DnsError) as e:
        dnship.DNS(Mx, Mz, ny)
    assert msg in str(e.value)
    assert "status 1" in str(e.value)


def test_null_handle_is_an_error():
    L = dnship.lib()
    rc = L.dns_rk3_step(None, 1)
    assert rc == 1
    assert b"null handle" in L.dns_last_error()


def test_seeded_ic_divergence_free_and_hermitian():
    Mx = Mz = 96
    ny = 65
    u1, u2, u3, p = dns_ic.seeded_channel(Mx, Mz, ny, seed=3)
    d = dns_ic.divergence(u1, u2, u3, 64, 64, ny)
    assert np.abs(d).max() < 1e-13
    Kx, Nz = 33, 64
    c = u1.reshape(Nz, Kx, -1)
    for kz in range(1, Nz // 2):
        assert np.allclose(c[kz, 0], np.conj(c[Nz - kz, 0]))
    assert np.all(u2[:, 0] == 0) and np.all(u2[:, -1] == 0)


def test_field_file_roundtrip(tmp_path):
    Mx = Mz = 24
    ny = 17
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=2)
    g = dnship.ygrid(ny)
    Nx = Nz = 16
    N = np.zeros((4, 5), np.int32)
    N[:, 3] = [Nx, Nz, ny - 2, ny - 1]
    names = [("u1", st[0], 0, g["yu"], g["dthdyu"]), ("u2", st[1], 0, g["yv"], g["dthdyv"]),
             ("u3", st[2], 0, g["yu"], g["dthdyu"]), ("p", st[3], 1, g["yu"], g["dthdyu"])]
    for name, f, j0, y, dth in names:
        planes = dns_io.columns_to_planes(f, Nz)
        js = np.arange(j0, j0 + planes.shape[0])
        ff = dns_io.FieldFile(1.5, 3250.0, 1.0, 2.0, -3e-3, 3, 42, N, y, 1.0, dth, js, planes, y[js])
        dns_io.write_field(str(tmp_path / ("%s_test.dat" % name)), ff)
    u1, u2, u3, p, head = dns_io.load_state(str(tmp_path), "_test.dat", ny)
    assert head.t == 1.5 and head.iter == 42
    for a, b in zip((u1, u2, u3, p), st):
        assert np.array_equal(a, b)


@pytest.mark.skipif(not os.path.isdir("/root/reference/input"), reason="reference fixtures absent")
def test_reads_reference_fixture():
    u1, u2, u3, p, head = dns_io.load_state("/root/reference/input", "_0032x0032x0153_t00000.dat", 153)
    assert u1.shape == (17 * 32, 154) and u2.shape == (544, 153) and p.shape == (544, 152)
    assert head.iter == 273750 and abs(head.Re - 3250.0) < 1e-12
    assert np.all(u2[:, 0] == 0) and np.all(u2[:, -1] == 0)   # v = 0 at both walls


def test_read_input_in():
    cfg = dns_io.read_input(os.path.join(ROOT, "final-single-block-_amd", "fortran", "input.in"))
    assert cfg["Mx"] == 48 and cfg["ny"] == 153 and cfg["flag_ctpress"] == 1
    assert cfg["mpgx"] == -3.0675e-3 and cfg["CFL"] == 0.7


def test_fortran_wrappers_stop_out_of_order_calls():
    """The drop-in wrappers keep the fields on the GPU and advance them only in the reference's call order
    (littleharsh.f90:165-222); a substep that starts with RHS0_u2 stops with a message naming the call
    expected, before anything reaches the GPU."""
    import subprocess
    prog = os.path.join(ROOT, "final-single-block-_amd", "fortran", "lh_order_check")
    assert os.path.exists(prog), "build the Fortran programs first (__graft_entry__.build())"
    r = subprocess.run([prog], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "RHS0_u2 called out of the reference's order; expected RHS0_u1" in r.stderr
