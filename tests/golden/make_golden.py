"""Writes the golden fixtures under tests/golden/ from the reference's own compiled hot path.

Run in a container that has the reference sources (oracle/_ref/ref_harness built by
``make -C oracle ref``):  python tests/golden/make_golden.py

Fixtures (inputs and expected outputs only):
  golden_steps_24x24x17.npz : seeded state (seed 7) and the reference's fields after 10 RK3 steps,
                              the per-step (t, dt) history; flag_ctpress = 1 and = 0 runs.
  golden_stages_12x12x13.npz: one substep kRK = 1, 2, 3 with the stage-by-stage arrays
                              (RHS0, nonlinear, fluxes, solveU, meanflow, solveP, v_corr).
  golden_fft.npz            : random spectral planes through four_to_phys_du / phys_to_four_du.
  golden_grid_153.npz       : y grid stored in the reference's 32x32x153 fixture files.
  golden_fixture32_steps.npz: per-field norms and 8 columns of the reference run of 4 steps from its
                              own 32x32x153 fixture (input/*_0032x0032x0153_t00000.dat).
  golden_stats_24x24x17.npz : a state (the reference's fields 3 steps after the seeded IC, seed 21) and
                              the reference's statistics and spectra of it (stats.f90, spectra.f90,
                              omega_x of FOU3D.f90:307-338) from one kRK = 1 nonlinear call with the
                              statistics flag set, p in its planes as on a write step.

  golden_ftp_16x16x9.npz    : the reference's post-processors (four_to_phys, four_to_phys_N, compiled
                              from their sources into oracle/_ref) on small random field files: the
                              input and output files, byte for byte.  Case A: four_to_phys on files
                              with record_out's current header; case B: four_to_phys_N on files with
                              a three-band header like the reference's own input files.

  golden_traj_128x128x153.npz: norms and 10 columns of the reference run from the 128x128x153 seeded IC
                              after 10, 100 and 1000 steps, the per-step (t, dt) history, and the same
                              from a 1e-15-perturbed IC (the flow's own amplification, for scale).
  golden_fixed_dt_32x32x153.npz: 10 steps at fixed dt = 1e-3 from the 32x32x153 seeded IC: every column's
                              and every row's norm, 16 whole columns, the per-step history.
  golden_dg.npz             : the reference's pressure matrices DG (LU_buildP/LU_decP).
  fixture_32x32x153.npz     : the reference's own 32x32x153 initial fields (input/*.dat), as data.

``python make_golden.py NAME...`` writes only those (stats, ftp, traj128, fixed_dt, dg, fixture32).
"""
import subprocess
import tempfile
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "final-single-block-_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import dns_ic  # noqa: E402
import dns_io  # noqa: E402
import pyoracle  # noqa: E402

REF_INPUT = "/root/reference/input"


def stats_golden():
    Mx = Mz = 24
    ny = 17
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=21, amp=0.1, broadband=True)
    r = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=3)
    state = [r[k] for k in ("u1", "u2", "u3", "p")]
    s = pyoracle.run_reference(Mx, Mz, ny, *state, mode=5, t=r["t"], it=3)
    out = {"u1_in": state[0], "u2_in": state[1], "u3_in": state[2], "p_in": state[3], "istat": s["istat"]}
    for k in pyoracle.STATS_NAMES + ("spU", "spV", "spW", "spUV", "spP"):
        out[k] = s[k]
    np.savez_compressed(os.path.join(HERE, "golden_stats_24x24x17.npz"), **out)


FTP_NX, FTP_NZ, FTP_NY = 16, 16, 9
FTP_FIELDS = (("u1", 1), ("u2", 2), ("u3", 3), ("p", 4))


def ftp_case_files(case, seed):
    """Random field files (every entry set, including those the transforms ignore) for one case."""
    Nx, Nz, ny = FTP_NX, FTP_NZ, FTP_NY
    nyu, nyv = ny - 1, ny - 2
    N = np.zeros((4, 5), np.int32)
    if case == "A":      # record_out's header (FOU3D.f90:2024-2030)
        N[0, 1:4] = Nx
        N[0, 4] = -2
        N[1, 1:4] = Nz
        N[2, 3] = nyv
        N[3, 3] = nyu
    else:                # three bands, as in the reference's input/*.dat files
        N[:, 1] = [Nx, Nz, 2, 3]
        N[:, 2] = [Nx, Nz, 5, 5]
        N[:, 3] = [Nx, Nz, nyv, nyu]
        N[:, 4] = [-2, 0, 0, 0]
    rng = np.random.default_rng(seed)
    yu = np.sort(rng.uniform(-1, 1, nyu + 2))
    yv = np.sort(rng.uniform(-1, 1, nyv + 2))
    files = {}
    tag = "_%04dx%04dx%04d_t00000.dat" % (Nx, Nz, ny)
    for name, gid in FTP_FIELDS:
        y, js = (yv, np.arange(0, nyv + 2)) if name == "u2" else (yu, np.arange(1, nyu + 1) if name == "p"
                                                                   else np.arange(0, nyu + 2))
        planes = rng.standard_normal((len(js), Nz, Nx // 2 + 1)) + 1j * rng.standard_normal((len(js), Nz, Nx // 2 + 1))
        ff = dns_io.FieldFile(0.004, 3250.0, 1.0, 2.0, -3.0675e-3, 3, 17, N, y, 1.0, 0.5 + 0 * y, js, planes, y[js])
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "f.dat")
            dns_io.write_field(path, ff, grid_id=gid)
            with open(path, "rb") as f:
                files[name + tag] = f.read()
    return files


def ftp_input_text():
    return "mgalx1=%10d\nmgalz1=%10d\nny=    %10d\nt=     %10s\n" % (FTP_NX, FTP_NZ, FTP_NY, "0.0d0")


def ftp_golden():
    out = {}
    for case, prog, seed in (("A", "four_to_phys", 3), ("B", "four_to_phys_N", 4)):
        files = ftp_case_files(case, seed)
        with tempfile.TemporaryDirectory() as d:
            os.makedirs(os.path.join(d, "output"))
            # The programs re-OPEN unit 10 on the next field's file without a CLOSE.  amdflang's runtime
            # then keeps reading at the previous file's end offset instead of the new file's start, so
            # each input file is preceded by as many zero bytes as the files read before it (the
            # programs write unit 20 from the start of each new file, so the outputs need nothing).
            pad = 0
            for name, _ in FTP_FIELDS:
                fn = [k for k in files if k.startswith(name + "_")][0]
                with open(os.path.join(d, "output", fn), "wb") as f:
                    f.write(b"\0" * pad + files[fn])
                pad += len(files[fn])
            with open(os.path.join(d, "input_FtP.txt"), "w") as f:
                f.write(ftp_input_text())
            subprocess.run([os.path.join(ROOT, "oracle", "_ref", prog)], cwd=d, check=True,
                           capture_output=True, timeout=120)
            for fn, b in files.items():
                name = fn.split("_")[0]
                out["%s_in_%s" % (case, name)] = np.frombuffer(b, np.uint8)
                with open(os.path.join(d, "output", fn.replace(name + "_", name + "_phys_", 1)), "rb") as f:
                    out["%s_out_%s" % (case, name)] = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "golden_ftp_16x16x9.npz"), **out)


TRAJ_COLS = np.array([0, 1, 2, 64, 65, 130, 1000, 4159, 4160, 8319])
TRAJ_STOPS = (10, 100, 1000)


TRAJ_CHUNK = 100  # steps per reference run (each run must finish inside run_reference's time limit)


def traj128_run(pert: bool):
    """The reference's own code from the 128x128x153 seeded IC (seed 1, amplitude 0.1, lowest 8 modes;
    pert: every velocity value times 1 + 1e-15 N(0,1), rng seed 1), CFL steps, as chained runs of at most
    TRAJ_CHUNK steps (each run starts from the fields the last one wrote, which carries the whole state at
    a constant pressure gradient): per-field norms and the TRAJ_COLS columns at each of TRAJ_STOPS, and the
    per-step (t, dt) history.  After every run the state is checkpointed in the temporary directory, so an
    interrupted generation resumes where it stopped."""
    Mx, Mz, ny = 192, 192, 153
    ckpt = os.path.join(tempfile.gettempdir(), "traj128_%s.npz" % ("pert" if pert else "plain"))
    st = list(dns_ic.seeded_channel(Mx, Mz, ny, seed=1))
    if pert:
        rng = np.random.default_rng(1)
        st = [a * (1 + 1e-15 * rng.standard_normal(a.shape)) for a in st[:3]] + [st[3]]
    out, t, done, hist = {}, 0.0, 0, []
    if os.path.exists(ckpt):
        c = np.load(ckpt, allow_pickle=False)
        st = [c[k] for k in ("u1", "u2", "u3", "p")]
        t, done = float(c["t"]), int(c["done"])
        hist = [c["hist"]] if c["hist"].size else []
        out = {k[4:]: c[k] for k in c.files if k.startswith("out_")}
        print("traj128 %s: resumed at %d steps" % ("perturbed" if pert else "plain", done), flush=True)
    while done < TRAJ_STOPS[-1]:
        stop = min([x for x in TRAJ_STOPS if x > done] + [done + TRAJ_CHUNK])
        r = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=stop - done, t=t, it=done)
        st = [r[k] for k in ("u1", "u2", "u3", "p")]
        t, done = r["t"], stop
        hist.append(r["hist"])
        if stop in TRAJ_STOPS:
            for k, a in zip(("u1", "u2", "u3", "p"), st):
                out["norm_%s_%d" % (k, stop)] = np.linalg.norm(a)
                out["cols_%s_%d" % (k, stop)] = a[TRAJ_COLS]
        np.savez(ckpt, u1=st[0], u2=st[1], u3=st[2], p=st[3], t=t, done=done, hist=np.concatenate(hist),
                 **{"out_" + k: v for k, v in out.items()})
        print("traj128 %s: %d steps" % ("perturbed" if pert else "plain", stop), flush=True)
    out["hist"] = np.concatenate(hist)
    return out


def traj128_golden():
    """golden_traj_128x128x153.npz: the reference's trajectory from the seeded IC over 1000 steps, and for
    scale the trajectory from the 1e-15-perturbed IC (keys prefixed pert_); both runs in parallel."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(2) as ex:
        a, b = ex.map(traj128_run, (False, True))
    out = {"cols": TRAJ_COLS, "stops": np.array(TRAJ_STOPS)}
    out.update(a)
    out.update({"pert_" + k: v for k, v in b.items()})
    st = dns_ic.seeded_channel(192, 192, 153, seed=1)
    out.update({"ic_norm_" + k: np.linalg.norm(x) for k, x in zip(("u1", "u2", "u3", "p"), st)})
    np.savez_compressed(os.path.join(HERE, "golden_traj_128x128x153.npz"), **out)


def fixed_dt_golden():
    """golden_fixed_dt_32x32x153.npz: 32x32x153 (48x48 dealiased), Re 3250, the seeded IC (seed 1,
    amplitude 0.1, lowest 8 modes), 10 RK3 steps at the fixed dt = 1e-3 (the harness sets dt, dti and t after
    the reference's nonlinear at kRK = 1)."""
    Mx, Mz, ny = 48, 48, 153
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=1)
    r = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=10, dt=1e-3)
    cols = np.array([0, 1, 2, 8, 16, 17, 18, 33, 100, 136, 271, 272, 288, 400, 527, 543])
    out = {"hist": r["hist"], "cols": cols}
    out.update({"ic_norm_" + k: np.linalg.norm(x) for k, x in zip(("u1", "u2", "u3", "p"), st)})
    for k in ("u1", "u2", "u3", "p"):      # every column's and every row's norm, and 16 whole columns
        out["colnorm_" + k] = np.linalg.norm(r[k], axis=1)
        out["rownorm_" + k] = np.linalg.norm(r[k], axis=0)
        out["cols_" + k] = r[k][cols]
    np.savez_compressed(os.path.join(HERE, "golden_fixed_dt_32x32x153.npz"), **out)


def dg_golden():
    """golden_dg.npz: the pressure matrices DG(3, 1:nyu, column) that the reference's LU_buildP + LU_decP
    build (tridLU_3D.f90:224-297, 366-405; harness setup output), at 12x12x13 and 24x24x17."""
    out = {}
    for Mx, ny in ((12, 13), (24, 17)):
        st = dns_ic.seeded_channel(Mx, Mx, ny, seed=1)
        r = pyoracle.run_reference(Mx, Mx, ny, *st, mode=1, nsteps=0)
        out["DG_%dx%d" % (Mx, ny)] = r["setup"]["DG"]
    np.savez_compressed(os.path.join(HERE, "golden_dg.npz"), **out)


def fixture32_state():
    """fixture_32x32x153.npz: the reference's own 32x32x153 initial fields (input/*_0032x0032x0153_t00000.dat)
    as column fields, with the header's t, so that the GPU tests start from them."""
    u1, u2, u3, p, head = dns_io.load_state(REF_INPUT, "_0032x0032x0153_t00000.dat", 153)
    np.savez_compressed(os.path.join(HERE, "fixture_32x32x153.npz"), u1=u1, u2=u2, u3=u3, p=p, t=head.t,
                        iter=head.iter, Re=head.Re, mpgx=head.mpgx)


ONLY = {"stats": lambda: stats_golden(), "ftp": lambda: ftp_golden(), "traj128": lambda: traj128_golden(),
        "fixed_dt": lambda: fixed_dt_golden(), "dg": lambda: dg_golden(), "fixture32": lambda: fixture32_state()}


def main():
    if not pyoracle.have_reference():
        raise SystemExit("oracle/_ref/ref_harness is not built")
    if sys.argv[1:] and sys.argv[1] in ONLY:
        for name in sys.argv[1:]:
            ONLY[name]()
        return
    stats_golden()
    ftp_golden()
    fixed_dt_golden()
    dg_golden()
    Mx = Mz = 24
    ny = 17
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=7, amp=0.3, nmodes=4)
    out = {"u1_in": st[0], "u2_in": st[1], "u3_in": st[2], "p_in": st[3]}
    for ctp in (1, 0):
        r = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=10, flag_ctpress=ctp)
        for k in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3"):
            out["%s_ctp%d" % (k, ctp)] = r[k]
        out["hist_ctp%d" % ctp] = r["hist"]
    np.savez_compressed(os.path.join(HERE, "golden_steps_24x24x17.npz"), **out)

    # stage-by-stage for each kRK; kRK > 1 starts from the state after one full step, with the
    # nonlinear terms and dt of that step
    Mx = Mz = 12
    ny = 13
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=5, amp=0.3, nmodes=3)
    r1 = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=1)
    base = [r1["u1"], r1["u2"], r1["u3"], r1["p"]]
    Nu = [r1["Nu1"], r1["Nu2"], r1["Nu3"]]
    out = {"u1_in": base[0], "u2_in": base[1], "u3_in": base[2], "p_in": base[3],
           "Nu1_in": Nu[0], "Nu2_in": Nu[1], "Nu3_in": Nu[2], "dt_in": r1["dt"], "t_in": r1["t"]}
    for k in (1, 2, 3):
        for ctp in (1, 0):
            r = pyoracle.run_reference(Mx, Mz, ny, *base, mode=2, kRK=k, dt=r1["dt"], t=r1["t"], Nu=Nu,
                                       flag_ctpress=ctp)
            tag = "k%d_ctp%d" % (k, ctp)
            for q, n in enumerate(("du1", "du2", "du3")):
                out["rhs0_%s_%s" % (n, tag)] = r["rhs0"][q]
                out["dunl_%s_%s" % (n, tag)] = r["du_nl"][q]
            for q, n in enumerate(("Nu1", "Nu2", "Nu3")):
                out["nl_%s_%s" % (n, tag)] = r["Nu_new"][q]
            for q, n in enumerate(("uv_f", "vv_c", "wv_f")):
                out["flux_%s_%s" % (n, tag)] = r["flux"][q]
            for q, n in enumerate(("u1", "u2", "u3")):
                out["solveU_%s_%s" % (n, tag)] = r["u_solved"][q]
                out["vcorr_%s_%s" % (n, tag)] = r["u_corr"][q]
            out["meanflow_u1_%s" % tag] = r["u1_meanflow"]
            out["solveP_p_%s" % tag] = r["p_new"]
            out["solveP_psi_%s" % tag] = r["psi"]
            out["vcorr_div_%s" % tag] = r["div"]
            sc = r["stage_scalars"]
            out["scalars_%s" % tag] = np.array([sc["dt"], sc["t"], sc["mpgx"], sc["dgx"], sc["dtc"]])
    np.savez_compressed(os.path.join(HERE, "golden_stages_12x12x13.npz"), **out)

    # FFT vectors: random retained spectral coefficients (reference plane layout), zero padding
    out = {}
    rng = np.random.default_rng(11)
    for (mx, mz, n) in ((48, 48, 2), (24, 24, 2), (96, 96, 1)):
        nx, nz = 2 * mx // 3, 2 * mz // 3
        kx = nx // 2 + 1
        pl = np.zeros((n, mz, mx + 2))
        for s in range(n):
            for zr in list(range(nz // 2)) + list(range(mz - nz // 2, mz)):
                pl[s, zr, :2 * kx] = rng.standard_normal(2 * kx)
        a, b = pyoracle.run_reference(mx, mz, 9, None, None, None, None, mode=3, fft_planes=pl)
        out["in_%d" % mx], out["phys_%d" % mx], out["back_%d" % mx] = pl, a, b
    np.savez_compressed(os.path.join(HERE, "golden_fft.npz"), **out)

    if os.path.isdir(REF_INPUT):
        ff = dns_io.read_field(os.path.join(REF_INPUT, "u1_0032x0032x0153_t00000.dat"))
        fv = dns_io.read_field(os.path.join(REF_INPUT, "u2_0032x0032x0153_t00000.dat"))
        np.savez_compressed(os.path.join(HERE, "golden_grid_153.npz"), yu=ff.y, dthdyu=ff.dthdy, yv=fv.y,
                            dthdyv=fv.dthdy)
        u1, u2, u3, p, head = dns_io.load_state(REF_INPUT, "_0032x0032x0153_t00000.dat", 153)
        r = pyoracle.run_reference(48, 48, 153, u1, u2, u3, p, mode=1, nsteps=4, t=head.t)
        cols = np.array([0, 1, 5, 17, 100, 271, 400, 543])
        out = {"cols": cols, "hist": r["hist"]}
        for k in ("u1", "u2", "u3", "p"):
            out["norm_" + k] = np.linalg.norm(r[k])
            out["cols_" + k] = r[k][cols]
        np.savez_compressed(os.path.join(HERE, "golden_fixture32_steps.npz"), **out)
        fixture32_state()
    traj128_golden()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
