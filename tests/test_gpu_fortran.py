"""The Fortran drop-in driver (reference subroutine names over ISO_C_BINDING) on the GPU: reads
reference-format input.in and field files, runs the reference's main loop, writes field files; the
result is compared with the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import dns_ic
import dns_io
import dnship
import pyoracle
from conftest import PKG, rel_l2, stats_errors

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(PKG, "fortran", "littleharsh_hip")


def write_case(d, Mx, Mz, ny, seed, flpress=1):
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=seed)
    g = dnship.ygrid(ny)
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    N = np.zeros((4, 5), np.int32)
    N[:, 3] = [Nx, Nz, ny - 2, ny - 1]
    os.makedirs(os.path.join(d, "input"))
    os.makedirs(os.path.join(d, "output"))
    tag = "_%04dx%04dx%04d_t00000.dat" % (Nx, Nz, ny)
    for name, f, j0, y, dth in (("u1", st[0], 0, g["yu"], g["dthdyu"]), ("u2", st[1], 0, g["yv"], g["dthdyv"]),
                                ("u3", st[2], 0, g["yu"], g["dthdyu"]), ("p", st[3], 1, g["yu"], g["dthdyu"])):
        planes = dns_io.columns_to_planes(f, Nz)
        js = np.arange(j0, j0 + planes.shape[0])
        ff = dns_io.FieldFile(0.0, 3250.0, 1.0, 2.0, -3.0675e-3, 3, 0, N, y, 1.0, dth, js, planes, y[js])
        dns_io.write_field(os.path.join(d, "input", name + tag), ff)
    with open(os.path.join(PKG, "fortran", "input.in")) as f:
        lines = f.read().split("\n")
    lines[14] = "Mx   = %10d" % Mx
    lines[16] = "Mz   = %10d" % Mz
    lines[18] = "ny   = %10d" % ny
    lines[27] = "fnmeb = u1" + tag
    lines[35] = "flpress    =%5d" % flpress
    with open(os.path.join(d, "input.in"), "w") as f:
        f.write("\n".join(lines))
    return st, tag


@pytest.mark.parametrize("mode,ctp", [("staged", 1), ("fused", 1), ("mirror", 1), ("mirror", 0), ("staged", 0)])
def test_fortran_driver_matches_oracle(tmp_path, mode, ctp):
    """mirror: the staged loop with DNS_HIP_MIRROR=1 (the wrappers copy every field they advance back into
    the caller's arrays after each call): the same results.  ctp = 0: constant flow rate (meanflow_ctU
    between the solves)."""
    assert os.path.exists(DRIVER), "build the Fortran driver first (__graft_entry__.build())"
    Mx, Mz, ny, n = 48, 48, 33, 5
    st, tag = write_case(str(tmp_path), Mx, Mz, ny, seed=12, flpress=ctp)
    args = [DRIVER, str(n)] + (["fused"] if mode == "fused" else [])
    env = dict(os.environ)
    env["DNS_HIP_MIRROR"] = "1" if mode == "mirror" else "0"
    print("case: fortran driver %s %dx%dx%d %d steps" % (mode, Mx, Mz, ny, n), flush=True)
    r = subprocess.run(args, cwd=str(tmp_path), capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr + r.stdout
    outdir = os.path.join(str(tmp_path), "output")
    # record_out at the start of every step (nwrite = 1) and once at the end; files are named by
    # int(100 t) as in the reference, so snapshots close in time share a file and the newest wins
    outs = sorted(f for f in os.listdir(outdir) if f.startswith("u1_"))
    heads = [dns_io.read_field(os.path.join(outdir, f)) for f in outs]
    last = max(heads, key=lambda hd: hd.iter)
    otag = [f for f, hd in zip(outs, heads) if hd.iter == n][0][2:]
    u1, u2, u3, p, head = dns_io.load_state(outdir, otag, ny)
    # with a constant flow rate the reference reads no mpgx (start.f90:198-202; the file header's is not
    # used, :1577-1579), so the run starts from mpgx = 0, as the driver does
    o = pyoracle.Oracle(Mx, Mz, ny, flag_ctpress=ctp, **({} if ctp == 1 else {"mpgx": 0.0}))
    o.set_state(*st)
    o.step(n)
    for a, b in zip((u1, u2, u3, p), o.get_state()):
        assert rel_l2(a, b) <= 1e-12
    assert head.iter == n == last.iter
    assert abs(head.t - o.scalars()["t"]) <= 1e-13
    hists = [f for f in os.listdir(outdir) if f.startswith("hist_")]
    assert hists == ["hist_%04dx%04dx%04d.dat" % (2 * Mx // 3, 2 * Mz // 3, ny)]
    hh, N, rows = dns_io.read_hist(os.path.join(outdir, hists[0]))
    assert [r["iter"] for r in rows] == list(range(n + 1))
    assert rows[0]["t"] == 0.0 and abs(rows[-1]["t"] - o.scalars()["t"]) <= 1e-13
    if ctp == 1:
        assert abs(rows[-1]["Qx"] - o.scalars()["Qx"]) <= 1e-13 * abs(o.scalars()["Qx"])
    assert all(r["flag_ctpress"] == ctp for r in rows) and hh["Re"] == 3250.0


def test_fortran_driver_stats_files_vs_oracle(tmp_path):
    """Statistics and spectra files of the driver (every nstat = min(20, nwrite) = 1 steps, written and
    reset at each write step) against the oracle's statistics of the state at the file's iteration."""
    Mx, Mz, ny, n = 48, 48, 33, 4
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    st, tag = write_case(str(tmp_path), Mx, Mz, ny, seed=13)
    print("case: fortran driver statistics files %dx%dx%d %d steps" % (Mx, Mz, ny, n), flush=True)
    r = subprocess.run([DRIVER, str(n), "fused"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    outdir = os.path.join(str(tmp_path), "output")
    files = sorted(f for f in os.listdir(outdir) if f.startswith("stats_"))
    assert files
    o = pyoracle.Oracle(Mx, Mz, ny)
    grid = o.setup()
    o.set_state(*st)
    done = 0
    seen = set()
    for f in sorted(files, key=lambda f: dns_io.read_stats(os.path.join(outdir, f))["iter"]):
        s = dns_io.read_stats(os.path.join(outdir, f))
        assert s["istat"] == (1 if s["iter"] >= 1 else 0)
        if s["istat"] == 0:
            assert all(not np.any(s[k]) for k in dns_io.STATS_FILE_ORDER)
            continue
        o.step(s["iter"] - done)
        done = s["iter"]
        seen.add(done)
        ref = pyoracle.plane_stats(Mx, Mz, ny, *o.get_state(), grid["yu"], grid["yv"], grid["dthdyv"], 2.0)
        nyu = ny - 1
        exp = {k: ref[k][:nyu + 1] if k in ("Vm", "V2m", "VWm", "wxm", "wx2m") else
               ref[k][1:nyu + 1] if k in ("Pm", "P2m") else ref[k] for k in dns_io.STATS_FILE_ORDER}
        err = stats_errors(s, exp, Mx * Mz, Nx * Nz)
        assert max(err.values()) <= 1e-11, (f, err)
        for name in ("spU", "spV", "spW", "spUV", "spP"):
            sp = dns_io.read_spectrum(os.path.join(outdir, name + f[5:]), Nx, Nz)
            want = dns_io.fold_spectrum(ref[name], Nx, Nz)
            assert sp["iter"] == done and sp["istat"] == 1
            assert np.abs(sp["sp"] - want).max() <= 1e-11 * np.abs(want).max(), name
    assert seen, files


def test_fortran_driver_stats_pressure_of_last_write(tmp_path):
    """nwrite = 40, so nstat = min(20, nwrite) = 20: statistics at steps 20 and 40, files at 0 and 40.  Pm
    and P2m sum ppPL, which the reference refreshes only at write steps, before that step's statistics
    (FOU3D.f90:149-152 precede stats at :340): step 20 sums the pressure of step 0, step 40 its own.
    Every other statistic sums the state of its own step."""
    Mx, Mz, ny, n = 48, 48, 33, 41
    st, tag = write_case(str(tmp_path), Mx, Mz, ny, seed=14)
    path = os.path.join(str(tmp_path), "input.in")
    with open(path) as f:
        lines = f.read().split("\n")
    lines[34] = "nwrt =%11d" % 40
    with open(path, "w") as f:
        f.write("\n".join(lines))
    print("case: fortran driver statistics nwrite=40 %dx%dx%d %d steps" % (Mx, Mz, ny, n), flush=True)
    r = subprocess.run([DRIVER, str(n), "fused"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    outdir = os.path.join(str(tmp_path), "output")
    files = [dns_io.read_stats(os.path.join(outdir, f)) for f in os.listdir(outdir) if f.startswith("stats_")]
    assert sorted(s["iter"] for s in files) == [0, 40]
    s = [s for s in files if s["iter"] == 40][0]
    assert s["istat"] == 2
    o = pyoracle.Oracle(Mx, Mz, ny)
    grid = o.setup()
    o.set_state(*st)

    def stats_now():
        return pyoracle.plane_stats(Mx, Mz, ny, *o.get_state(), grid["yu"], grid["yv"], grid["dthdyv"], 2.0)

    p0 = stats_now()
    o.step(20)
    s20 = stats_now()
    o.step(20)
    s40 = stats_now()
    nyu = ny - 1
    exp = {}
    for k in dns_io.STATS_FILE_ORDER:
        if k in ("Pm", "P2m"):
            exp[k] = (p0[k] + s40[k])[1:nyu + 1]
        elif k in ("Vm", "V2m", "VWm", "wxm", "wx2m"):
            exp[k] = (s20[k] + s40[k])[:nyu + 1]
        else:
            exp[k] = s20[k] + s40[k]
    err = stats_errors(s, exp, Mx * Mz, (2 * Mx // 3) * (2 * Mz // 3))
    assert max(err.values()) <= 1e-11, err
    # the pressure of step 20 itself differs: the test sees the difference it is about
    cur = (s20["Pm"] + s40["Pm"])[1:nyu + 1]
    assert np.abs(cur - exp["Pm"]).max() > 1e3 * 1e-11 * np.sqrt(np.abs(exp["P2m"]).max())


def test_fortran_driver_rejects_bad_size(tmp_path):
    write_case(str(tmp_path), 48, 48, 33, seed=1)
    with open(os.path.join(str(tmp_path), "input.in")) as f:
        lines = f.read().split("\n")
    lines[14] = "Mx   = %10d" % 60   # 60 = 2^2 3 5: FFTPACK's fatal factor
    with open(os.path.join(str(tmp_path), "input.in"), "w") as f:
        f.write("\n".join(lines))
    r = subprocess.run([DRIVER, "1"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "prime factor" in r.stderr
