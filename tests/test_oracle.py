"""CPU tests that pin the oracle (oracle/dns_oracle.c) before anything is checked against it.

Pins, in order of strength:
  * the reference's own compiled Fortran hot path: committed golden fixtures written by
    tests/golden/make_golden.py from oracle/_ref/ref_harness, and live runs where that harness is built;
  * the reference's own fixture data: the y grid stored in input/*_0032x0032x0153_t00000.dat;
  * mathematics: the DFT definition (numpy.fft) for the padded/pruned transforms, dense solves of the
    assembled tridiagonal matrices, and the projection's zero divergence.
"""
import os

import numpy as np
import pytest

import dns_ic
import pyoracle
from conftest import ftp_compare, rel_l2, stats_errors

TOL = 1e-12


def seeded(Mx, Mz, ny, seed, **kw):
    return dns_ic.seeded_channel(Mx, Mz, ny, seed=seed, **kw)


# ---------------------------------------------------------------------------------------------
def test_grid_bit_exact_against_reference_fixture(golden):
    g = golden("golden_grid_153.npz")
    o = pyoracle.Oracle(48, 48, 153)
    s = o.setup()
    assert np.array_equal(s["yu"], g["yu"])
    assert np.array_equal(s["yv"], g["yv"])
    assert np.array_equal(s["dthdyu"], g["dthdyu"])
    assert np.array_equal(s["dthdyv"], g["dthdyv"])


def test_seeded_ic_matches_golden_input(golden):
    g = golden("golden_steps_24x24x17.npz")
    st = seeded(24, 24, 17, 7, amp=0.3, nmodes=4)
    for k, a in zip(("u1", "u2", "u3", "p"), st):
        assert np.array_equal(a, g[k + "_in"]), k


@pytest.mark.parametrize("ctp", [1, 0])
def test_oracle_10_steps_vs_reference_golden(golden, ctp):
    g = golden("golden_steps_24x24x17.npz")
    o = pyoracle.Oracle(24, 24, 17, flag_ctpress=ctp)
    o.set_state(g["u1_in"], g["u2_in"], g["u3_in"], g["p_in"])
    hist = o.step(10)
    for k, a in zip(("u1", "u2", "u3", "p"), o.get_state()):
        assert rel_l2(a, g["%s_ctp%d" % (k, ctp)]) <= TOL, k
    for k, a in zip(("Nu1", "Nu2", "Nu3"), o.get_nu()):
        assert rel_l2(a, g["%s_ctp%d" % (k, ctp)]) <= TOL, k
    ref_hist = g["hist_ctp%d" % ctp]
    np.testing.assert_allclose(hist[:, 0], ref_hist[:, 1], rtol=1e-14)   # t
    np.testing.assert_allclose(hist[:, 1], ref_hist[:, 2], rtol=1e-14)   # dt
    # mpgx changes when ctp = 0 by g = (QxT - Qx0)/Qx1, a difference of nearly equal flow rates
    np.testing.assert_allclose(hist[:, 3], ref_hist[:, 5], rtol=0, atol=1e-11)


@pytest.mark.parametrize("kRK", [1, 2, 3])
@pytest.mark.parametrize("ctp", [1, 0])
def test_oracle_stages_vs_reference_golden(golden, kRK, ctp):
    g = golden("golden_stages_12x12x13.npz")
    tag = "k%d_ctp%d" % (kRK, ctp)
    o = pyoracle.Oracle(12, 12, 13, flag_ctpress=ctp)
    o.set_state(g["u1_in"], g["u2_in"], g["u3_in"], g["p_in"], t=float(g["t_in"]))
    o.set_nu(g["Nu1_in"], g["Nu2_in"], g["Nu3_in"])
    if kRK > 1:
        o.set_dt(float(g["dt_in"]))
    o.rhs0(kRK)
    for n, a in zip(("du1", "du2", "du3"), o.get_du()):
        assert rel_l2(a, g["rhs0_%s_%s" % (n, tag)]) <= 1e-13, n
    o.nonlinear(kRK)
    for n, a in zip(("Nu1", "Nu2", "Nu3"), o.get_nu()):
        assert rel_l2(a, g["nl_%s_%s" % (n, tag)]) <= 1e-13, n
    for n, a in zip(("du1", "du2", "du3"), o.get_du()):
        assert rel_l2(a, g["dunl_%s_%s" % (n, tag)]) <= 1e-13, n
    for n, a in zip(("uv_f", "vv_c", "wv_f"), o.get_flux()):
        assert rel_l2(a, g["flux_%s_%s" % (n, tag)]) <= 1e-13, n
    sc = g["scalars_%s" % tag]
    np.testing.assert_allclose(o.scalars()["dt"], sc[0], rtol=1e-14)
    o.solveU(kRK)
    for n, a in zip(("u1", "u2", "u3"), o.get_state()[:3]):
        assert rel_l2(a, g["solveU_%s_%s" % (n, tag)]) <= 1e-13, n
    if ctp == 0:
        o.meanflow_ctU(kRK)
    assert rel_l2(o.get_state()[0], g["meanflow_u1_%s" % tag]) <= 1e-13
    o.solveP(kRK)
    psi, _ = o.get_psi()
    assert rel_l2(o.get_state()[3], g["solveP_p_%s" % tag]) <= 1e-12
    assert rel_l2(psi, g["solveP_psi_%s" % tag]) <= 1e-12
    o.vcorr(kRK)
    for n, a in zip(("u1", "u2", "u3"), o.get_state()[:3]):
        assert rel_l2(a, g["vcorr_%s_%s" % (n, tag)]) <= 1e-13, n
    np.testing.assert_allclose(o.scalars()["mpgx"], sc[2], rtol=0, atol=1e-11)


@pytest.mark.parametrize("M", [24, 48, 96])
def test_oracle_plane_fft_vs_reference_golden(golden, M):
    g = golden("golden_fft.npz")
    phys, back = pyoracle.plane_fft(M, M, g["in_%d" % M])
    scale = np.abs(g["phys_%d" % M]).max()
    assert np.abs(phys - g["phys_%d" % M]).max() <= 1e-13 * scale
    assert np.abs(back - g["back_%d" % M]).max() <= 1e-13 * np.abs(g["back_%d" % M]).max()


@pytest.mark.parametrize("M", [12, 24, 48, 96])
def test_oracle_plane_fft_matches_dft_definition(M):
    """four_to_phys_du / phys_to_four_du restated with numpy's DFT (rft_buff.f:26-48, cft_buff.f:3-13):
    u(x_m, z_l) = sum_{kx, kz} c(kx, kz) e^{+i(2 pi kx m/Mx + 2 pi kz l/Mz)} over the Hermitian x-extension,
    c = (1/(Mx Mz)) sum u e^{-i...}; Im c(0, .) ignored on the inverse; retained modes only."""
    rng = np.random.default_rng(M)
    Nx = Nz = 2 * M // 3
    Kx = Nx // 2 + 1
    pl = np.zeros((1, M, M + 2))
    for zr in list(range(Nz // 2)) + list(range(M - Nz // 2, M)):
        pl[0, zr, :2 * Kx] = rng.standard_normal(2 * Kx)
    phys, back = pyoracle.plane_fft(M, M, pl)
    c = pl[0, :, 0:2 * Kx:2] + 1j * pl[0, :, 1:2 * Kx:2]      # [z-mode row, kx]
    Cz = np.fft.ifft(c, axis=0) * M                             # z synthesis, unnormalised
    Cz[:, 0] = Cz[:, 0].real                                    # Im of kx = 0 ignored by rft
    full = np.zeros((M, M // 2 + 1), complex)
    full[:, :Kx] = Cz
    u = np.fft.irfft(full, n=M, axis=1) * M
    assert np.abs(phys[0, :, :M] - u).max() <= 1e-12 * np.abs(u).max()
    X = np.fft.rfft(u, axis=1) / M
    Xz = np.fft.fft(X[:, :Kx], axis=0) / M
    Xz[M // 2, :] = 0.0
    got = back[0, :, 0:2 * Kx:2] + 1j * back[0, :, 1:2 * Kx:2]
    assert np.abs(got - Xz).max() <= 1e-12 * np.abs(Xz).max()
    # retained modes come back up to rounding; the kx = 0 column comes back Hermitian-symmetrised,
    # because the real x-synthesis drops its imaginary part: (c(kz) + conj c(-kz)) / 2
    kept = list(range(Nz // 2)) + list(range(M - Nz // 2, M))
    ref = c.copy()
    neg = (M - np.arange(M)) % M
    ref[:, 0] = 0.5 * (c[:, 0] + np.conj(c[neg, 0]))
    assert np.abs(got[kept] - ref[kept]).max() <= 1e-12 * np.abs(ref).max()


def test_oracle_solveU_matches_dense_solve():
    """LU_build/LU_dec/LUsolU restated against a dense solve of the assembled matrix (tridLU_3D.f90:163-221)."""
    Mx, Mz, ny = 24, 24, 17
    st = seeded(Mx, Mz, ny, 3, amp=0.2, nmodes=3)
    o = pyoracle.Oracle(Mx, Mz, ny)
    o.set_state(*st)
    o.rhs0(1)
    o.nonlinear(1)
    du = o.get_du()
    sc = o.scalars()
    o.solveU(1)
    u1 = o.get_state()[0]
    s = o.setup()
    dt, Re = sc["dt"], 3250.0
    beta = (4.0 / 15.0) * dt / Re
    Kx, Nz = o.Kx, o.Nz
    n = ny + 1
    for col in (0, 3, Kx + 1, 5 * Kx + 2):
        i, k = col % Kx, col // Kx
        kz = k if k < Nz // 2 else k - Nz
        k2 = -(1.0 * i) ** 2 - (2.0 * kz) ** 2
        A = np.zeros((n, n))
        A[0, 0], A[0, 1] = 1.0, s["gw"][0]
        A[-1, -1], A[-1, -2] = 1.0, s["gw"][1]
        for j in range(1, n - 1):
            A[j, j - 1] = -beta * s["dyu2i"][j, 0]
            A[j, j] = 1.0 - beta * (s["dyu2i"][j, 1] + k2)
            A[j, j + 1] = -beta * s["dyu2i"][j, 2]
        rhs = np.zeros(n, complex)
        rhs[1:-1] = st[0][col, 1:-1] + dt * du[0][col]
        x = np.linalg.solve(A, rhs)
        assert np.abs(x - u1[col]).max() <= 1e-12 * max(np.abs(x).max(), 1e-30)


def test_oracle_projection_is_divergence_free():
    Mx, Mz, ny = 48, 48, 33
    st = seeded(Mx, Mz, ny, 4)
    o = pyoracle.Oracle(Mx, Mz, ny)
    o.set_state(*st)
    o.step(2)
    u1, u2, u3, _ = o.get_state()
    d = dns_ic.divergence(u1, u2, u3, o.Nx, o.Nz, ny)
    assert np.abs(d).max() <= 1e-10
    assert np.abs(u2[:, 0]).max() == 0.0 and np.abs(u2[:, -1]).max() == 0.0


def _plane_stats(Mx, Mz, ny, st, bet=2.0):
    s = pyoracle.Oracle(Mx, Mz, ny).setup()
    return pyoracle.plane_stats(Mx, Mz, ny, *st, s["yu"], s["yv"], s["dthdyv"], bet)


def test_oracle_plane_stats_vs_reference_golden(golden):
    """The sum-rule statistics and the spectra against stats.f90 / spectra.f90 run by the reference's
    own code (physical-space sums after its FFTs) on a state 3 steps after a seeded IC."""
    g = golden("golden_stats_24x24x17.npz")
    o = _plane_stats(24, 24, 17, (g["u1_in"], g["u2_in"], g["u3_in"], g["p_in"]))
    err = stats_errors(o, g, 24 * 24, 16 * 16)
    assert max(err.values()) < 1e-12, err
    for k in ("spU", "spV", "spW", "spUV", "spP"):
        assert np.abs(o[k] - g[k]).max() <= 1e-14 * np.abs(g[k]).max(), k


def test_plane_stats_planted_mistake_is_caught(golden):
    """A wrong spanwise wavenumber in omega_x must break the golden comparison of wxm/wx2m."""
    g = golden("golden_stats_24x24x17.npz")
    o = _plane_stats(24, 24, 17, (g["u1_in"], g["u2_in"], g["u3_in"], g["p_in"]), bet=2.0 * (1 + 1e-6))
    err = stats_errors(o, g, 24 * 24, 16 * 16)
    assert err["wx2m"] > 1e-9 and err["U2m"] < 1e-12


def test_plane_stats_mean_flow_closed_form():
    """A pure mean flow u1(y) on the interior planes: Um = Mx Mz U(y), U2m = Mx Mz U(y)^2, nothing else."""
    Mx, Mz, ny = 24, 24, 17
    st = [np.zeros_like(a) for a in seeded(Mx, Mz, ny, 3)]
    U = np.linspace(0.3, 1.1, ny + 1)
    st[0][0, :] = U
    o = _plane_stats(Mx, Mz, ny, st)
    j = np.arange(1, ny)
    assert np.allclose(o["Um"][j], Mx * Mz * U[j], rtol=1e-15)
    assert np.allclose(o["U2m"][j], Mx * Mz * U[j] ** 2, rtol=1e-14)
    assert np.all(o["UVm"] == 0) and np.all(o["W2m"] == 0) and np.all(o["wx2m"] == 0)


@pytest.mark.parametrize("case,nvar", [("A", False), ("B", True)])
def test_oracle_four_to_phys_vs_reference_golden(golden, case, nvar):
    """The post-processors four_to_phys (case A: record_out's header) and four_to_phys_N (case B: a
    three-band header), compiled from the reference's sources, on random field files: every output
    record (header records and plane headers byte for byte, planes to rounding)."""
    g = golden("golden_ftp_16x16x9.npz")
    files = {n: g["%s_in_%s" % (case, n)].tobytes() for n in ("u1", "u2", "u3", "p")}
    out = pyoracle.four_to_phys_files(files, nvar)
    for n in ("u1", "u2", "u3", "p"):
        ftp_compare(out[n], g["%s_out_%s" % (case, n)].tobytes(), 1e-14)


@pytest.mark.parametrize("Mx,Mz", [(24, 24), (16, 16), (24, 16)])
def test_ftp_plane_single_modes_closed_form(Mx, Mz):
    """One mode c at (kx, kz) gives 2 Re(c e^{2 pi i (kx x/Mx + kz z/Mz)}), and Re(c e^{2 pi i kz z/Mz})
    at kx = 0 (the real x-inverse takes Re of that column); the kz = -Nz/2 row is dropped on the 3/2 grid
    (Mz > Nz) and kept when Mz == Nz."""
    Nx = Nz = 16
    x = np.arange(Mx)[None, :]
    z = np.arange(Mz)[:, None]
    for kx, kz, c in ((1, 2, 1.0 + 0.5j), (0, 3, 0.7 - 0.2j), (5, -4, -0.3 + 1.1j), (2, -8, 1.0 + 0.0j)):
        b = np.zeros((Nz, Nx + 2))
        b[kz % Nz, 2 * kx] = c.real
        b[kz % Nz, 2 * kx + 1] = c.imag
        u = pyoracle.ftp_plane(b, Mx, Mz)
        ph = np.exp(2j * np.pi * (kx * x / Mx + kz * z / Mz))
        if kz == -Nz // 2 and Mz > Nz:
            want = np.zeros((Mz, Mx))
        elif kx == 0:
            want = (c * ph).real
        else:
            want = 2 * (c * ph).real
        assert np.allclose(u[:, :Mx], want, atol=1e-12), (kx, kz)
        assert not np.any(u[:, Mx:])


def test_planted_mistake_is_caught(golden):
    """A wrong RK coefficient in the oracle's inputs must break the golden comparison."""
    g = golden("golden_steps_24x24x17.npz")
    o = pyoracle.Oracle(24, 24, 17, Re=3250.0 * (1 + 1e-9))
    o.set_state(g["u1_in"], g["u2_in"], g["u3_in"], g["p_in"])
    o.step(10)
    assert rel_l2(o.get_state()[0], g["u1_ctp1"]) > TOL


# ---------------------------------------------------------------------------------------------
# live comparisons with the reference's compiled code (only where oracle/_ref is built)
needs_ref = pytest.mark.skipif(not pyoracle.have_reference(), reason="oracle/_ref not built")


@needs_ref
def test_oracle_vs_reference_live_seeded_32():
    Mx, Mz, ny = 48, 48, 153
    st = seeded(Mx, Mz, ny, 1)
    r = pyoracle.run_reference(Mx, Mz, ny, *st, mode=1, nsteps=10)
    o = pyoracle.Oracle(Mx, Mz, ny)
    o.set_state(*st)
    o.step(10)
    for k, a in zip(("u1", "u2", "u3", "p"), o.get_state()):
        assert rel_l2(a, r[k]) <= TOL, k


def test_oracle_vs_reference_fixture32_golden(golden):
    """The reference's own 32x32x153 initial fields (input/*_0032x0032x0153_t00000.dat, committed as
    fixture_32x32x153.npz): 4 oracle steps against the reference's run of the same 4 steps."""
    g = golden("golden_fixture32_steps.npz")
    fx = golden("fixture_32x32x153.npz")
    u1, u2, u3, p = (fx[k] for k in ("u1", "u2", "u3", "p"))
    o = pyoracle.Oracle(48, 48, 153)
    o.set_state(u1, u2, u3, p, t=float(fx["t"]))
    o.step(4)
    cols = g["cols"]
    for k, a in zip(("u1", "u2", "u3", "p"), o.get_state()):
        assert abs(np.linalg.norm(a) - g["norm_" + k]) <= TOL * g["norm_" + k], k
        assert rel_l2(a[cols], g["cols_" + k]) <= 1e-11, k


@pytest.mark.skipif(not os.path.isdir("/root/reference/input"), reason="reference fixtures absent")
def test_fixture32_npz_equals_reference_files(golden):
    """The committed fixture holds exactly the fields of the reference's input files."""
    import dns_io
    fx = golden("fixture_32x32x153.npz")
    u1, u2, u3, p, head = dns_io.load_state("/root/reference/input", "_0032x0032x0153_t00000.dat", 153)
    for k, a in zip(("u1", "u2", "u3", "p"), (u1, u2, u3, p)):
        assert np.array_equal(fx[k], a), k
    assert float(fx["t"]) == head.t


def test_oracle_DG_bit_exact_vs_reference(golden):
    """setup_DG (the pressure matrices of LU_buildP + LU_decP, tridLU_3D.f90:224-297, 366-405) equals the
    reference's own matrices bit for bit (golden_dg.npz, from the harness's set-up output)."""
    g = golden("golden_dg.npz")
    for Mx, ny in ((12, 13), (24, 17)):
        DG = pyoracle.Oracle(Mx, Mx, ny).setup()["DG"].ravel()
        assert np.array_equal(DG, g["DG_%dx%d" % (Mx, ny)]), (Mx, ny)


def fixed_dt_errors(state, g):
    """Largest relative error of the column norms, row norms and whole columns against the golden."""
    err = 0.0
    for k, a in zip(("u1", "u2", "u3", "p"), state):
        cn, rn = g["colnorm_" + k], g["rownorm_" + k]
        err = max(err, np.abs(np.linalg.norm(a, axis=1) - cn).max() / cn.max(),
                  np.abs(np.linalg.norm(a, axis=0) - rn).max() / rn.max(),
                  rel_l2(a[g["cols"]], g["cols_" + k]))
    return err


def test_oracle_fixed_dt_vs_reference_golden(golden):
    """A run at fixed dt = 1e-3 (32x32x153, 10 steps from the seeded IC) against the reference's own code run
    with its dt fixed the same way; a dt off by a relative 1e-9 must fail."""
    g = golden("golden_fixed_dt_32x32x153.npz")
    st = seeded(48, 48, 153, 1)
    for k, a in zip(("u1", "u2", "u3", "p"), st):
        assert abs(np.linalg.norm(a) - g["ic_norm_" + k]) <= 1e-15 * g["ic_norm_" + k], k
    o = pyoracle.Oracle(48, 48, 153, fixed_dt=1e-3)
    o.set_state(*st)
    hist = o.step(10)
    assert fixed_dt_errors(o.get_state(), g) <= TOL
    assert np.array_equal(hist[:, 1], g["hist"][:, 2]) and np.allclose(hist[:, 0], g["hist"][:, 1], rtol=1e-15)
    o = pyoracle.Oracle(48, 48, 153, fixed_dt=1e-3 * (1 + 1e-9))
    o.set_state(*st)
    o.step(10)
    assert fixed_dt_errors(o.get_state(), g) > TOL
