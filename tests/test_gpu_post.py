"""The four_to_phys post-processor on the GPU: the Fortran program (four_to_phys_hip) against the
reference's own post-processors' output files, dns_four_to_phys at full size against the oracle, and the
driver's snapshot files converted end to end."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import dns_io
import dnship
import pyoracle
from conftest import PKG, ftp_compare

pytestmark = pytest.mark.gpu
PROG = os.path.join(PKG, "fortran", "four_to_phys_hip")
FIELDS = ("u1", "u2", "u3", "p")


def run_prog(d, nvar):
    assert os.path.exists(PROG), "build the Fortran programs first (__graft_entry__.build())"
    r = subprocess.run([PROG] + (["N"] if nvar else []), cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


@pytest.mark.parametrize("case,nvar", [("A", False), ("B", True)])
def test_four_to_phys_program_vs_reference_golden(tmp_path, golden, case, nvar):
    """Output files of four_to_phys (A) / four_to_phys_N (B), compiled from the reference's sources,
    on random 16x16x9 field files (golden_ftp_16x16x9.npz)."""
    g = golden("golden_ftp_16x16x9.npz")
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "output"))
    tag = "_0016x0016x0009_t00000.dat"
    for n in FIELDS:
        with open(os.path.join(d, "output", n + tag), "wb") as f:
            f.write(g["%s_in_%s" % (case, n)].tobytes())
    with open(os.path.join(d, "input_FtP.txt"), "w") as f:
        f.write("mgalx1=%10d\nmgalz1=%10d\nny=    %10d\nt=     %10s\n" % (16, 16, 9, "0.0d0"))
    print("case: four_to_phys program %s" % case, flush=True)
    run_prog(d, nvar)
    for n in FIELDS:
        with open(os.path.join(d, "output", n + "_phys" + tag), "rb") as f:
            ftp_compare(f.read(), g["%s_out_%s" % (case, n)].tobytes(), 1e-13)


@pytest.mark.parametrize("Nx,Nz,Mx,Mz", [(384, 384, 576, 576), (384, 384, 384, 384), (768, 384, 1152, 576),
                                         (768, 384, 768, 384)])
def test_four_to_phys_full_size_vs_oracle(Nx, Nz, Mx, Mz):
    """Spectral planes of the 384 x 384 and 768 x 384 workloads to the 3/2 grid and to the spectral grid,
    against the oracle's transform (numpy FFTs), every entry of the input set."""
    rng = np.random.default_rng(7)
    planes = rng.standard_normal((6, Nz, Nx + 2))
    print("case: dns_four_to_phys %dx%d -> %dx%d" % (Nx, Nz, Mx, Mz), flush=True)
    out, ms = dnship.four_to_phys(planes, Mx, Mz)
    assert ms > 0
    for k in range(planes.shape[0]):
        want = pyoracle.ftp_plane(planes[k], Mx, Mz)
        assert np.abs(out[k] - want).max() <= 1e-13 * np.abs(want).max(), k


def test_four_to_phys_rejects_bad_sizes():
    p = np.zeros((1, 16, 18))
    with pytest.raises(RuntimeError):
        dnship.four_to_phys(p, 12, 24)      # Mx < Nx
    with pytest.raises(RuntimeError):
        dnship.four_to_phys(p, 40, 24)      # Mx/2 = 20 is not an instantiated length


def test_driver_snapshot_to_physical(tmp_path):
    """The drop-in driver's snapshot (record_out's layout) through four_to_phys_hip, against the oracle's
    restatement of four_to_phys on the same files."""
    from test_gpu_fortran import DRIVER, write_case
    Mx, Mz, ny = 48, 48, 33
    d = str(tmp_path)
    write_case(d, Mx, Mz, ny, seed=5)
    print("case: driver snapshot -> four_to_phys_hip", flush=True)
    r = subprocess.run([DRIVER, "2", "fused"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    outdir = os.path.join(d, "output")
    last = max((f for f in os.listdir(outdir) if f.startswith("u1_")),
               key=lambda f: dns_io.read_field(os.path.join(outdir, f)).iter)
    ttag = last[-9:-4]                       # int(100 t) of record_out; four_to_phys names by int(10 t)
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    files = {}
    for n in FIELDS:
        src = os.path.join(outdir, "%s_%04dx%04dx%04d_t%s.dat" % (n, Nx, Nz, ny, ttag))
        with open(src, "rb") as f:
            files[n] = f.read()
    with open(os.path.join(d, "input_FtP.txt"), "w") as f:
        f.write("mgalx1=%10d\nmgalz1=%10d\nny=    %10d\nt=     %10s\n" % (Nx, Nz, ny, "%.5fd0" % (int(ttag) / 10 + 2e-5)))
    run_prog(d, False)
    want = pyoracle.four_to_phys_files(files, False)
    for n in FIELDS:
        with open(os.path.join(outdir, "%s_phys_%04dx%04dx%04d_t%s.dat" % (n, Nx, Nz, ny, ttag)), "rb") as f:
            ftp_compare(f.read(), want[n], 1e-13)


def test_n_variant_on_driver_snapshot(tmp_path):
    """four_to_phys_N's p ranges read nyu + 1 planes from a file with record_out's current header.  A
    one-rank record_out writes that many (limPL_incw(pgrid) = 1..nyu+1, start.f90:1187-1194; plane nyu+1
    of ppPL is zero), and so does the drop-in driver, so the N variant converts all four fields; the
    output files equal the oracle's restatement of four_to_phys_N on the same files."""
    from test_gpu_fortran import DRIVER, write_case
    Mx, Mz, ny = 48, 48, 33
    d = str(tmp_path)
    write_case(d, Mx, Mz, ny, seed=6)
    r = subprocess.run([DRIVER, "1", "fused"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    outdir = os.path.join(d, "output")
    files = {}
    for n in FIELDS:
        with open(os.path.join(outdir, "%s_%04dx%04dx%04d_t00000.dat" % (n, Nx, Nz, ny)), "rb") as f:
            files[n] = f.read()
    ff = dns_io.read_field(os.path.join(outdir, "p_%04dx%04dx%04d_t00000.dat" % (Nx, Nz, ny)))
    assert list(ff.j) == list(range(1, ny + 1)) and not np.any(ff.planes[-1])   # p planes 1..nyu+1
    with open(os.path.join(d, "input_FtP.txt"), "w") as f:
        f.write("mgalx1=%10d\nmgalz1=%10d\nny=    %10d\nt=     %10s\n" % (Nx, Nz, ny, "0.0d0"))
    print("case: four_to_phys_hip N on a driver snapshot", flush=True)
    r = subprocess.run([PROG, "N"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    want = pyoracle.four_to_phys_files(files, True)
    for n in FIELDS:
        with open(os.path.join(outdir, "%s_phys_%04dx%04dx%04d_t00000.dat" % (n, Nx, Nz, ny)), "rb") as f:
            ftp_compare(f.read(), want[n], 1e-13)
