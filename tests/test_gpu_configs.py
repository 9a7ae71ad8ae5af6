"""GPU parity at the stated workloads: the HIP path through the C-ABI against the reference's own outputs
(committed golden fixtures) and the oracle.

* 32x32x153 at fixed dt = 1e-3, 10 steps (the reference's CPU-runnable case as a fixed-step run);
* the reference's own 32x32x153 initial fields (input/*_0032x0032x0153_t00000.dat, as a fixture);
* 128x128x153 over 1000 CFL steps against the reference's own trajectory, step by step through dt;
* the 48-member ensemble of three sizes (24 x 32^2x153, 16 x 64^2x97, 8 x 128^2x153), FP64 and mixed, one
  grouped launch per kernel class per substep, replayed from a hipGraph, against single channels;
* one full-size FP64 step at 384x384x257 and at 768x384x385 against the oracle, with the mixed path's
  error against the oracle.
FP64 bar: relative L2 <= 1e-12.
"""
import os

import numpy as np
import pytest

import dns_ic
import dnship
import pyoracle
from conftest import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def fixed_dt_errors(state, g):
    err = 0.0
    for k, a in zip(("u1", "u2", "u3", "p"), state):
        cn, rn = g["colnorm_" + k], g["rownorm_" + k]
        err = max(err, np.abs(np.linalg.norm(a, axis=1) - cn).max() / cn.max(),
                  np.abs(np.linalg.norm(a, axis=0) - rn).max() / rn.max(),
                  rel_l2(a[g["cols"]], g["cols_" + k]))
    return err


@pytest.mark.parametrize("path", [0, 1])
def test_fixed_dt_32x32x153_vs_reference(golden, path):
    """32x32x153 (48x48 dealiased), Re 3250, the seeded IC (seed 1, amplitude 0.1, lowest 8 modes), 10 steps
    at fixed dt = 1e-3: every column's and row's norm and 16 whole columns against the reference's own run,
    the whole fields against the oracle, and t after every step."""
    g = golden("golden_fixed_dt_32x32x153.npz")
    Mx, Mz, ny = 48, 48, 153
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=1)
    for k, a in zip(("u1", "u2", "u3", "p"), st):
        assert abs(np.linalg.norm(a) - g["ic_norm_" + k]) <= 1e-15 * g["ic_norm_" + k], k
    print("case: fixed dt 48x48x153, column path %d" % path, flush=True)
    h = dnship.DNS(Mx, Mz, ny, fixed_dt=1e-3, column_path=path)
    h.upload(*st)
    ts = []
    for _ in range(10):
        h.step(1)
        sc = h.scalars()
        assert sc["dt"] == 1e-3
        ts.append(sc["t"])
    out = h.download()
    h.close()
    assert fixed_dt_errors(out, g) <= TOL
    assert np.allclose(ts, g["hist"][:, 1], rtol=1e-15, atol=0)
    o = pyoracle.Oracle(Mx, Mz, ny, fixed_dt=1e-3, threads=THREADS)
    o.set_state(*st)
    o.step(10)
    for k, a, b in zip(("u1", "u2", "u3", "p"), out, o.get_state()):
        assert rel_l2(a, b) <= TOL, k


def test_reference_fixture_32x32x153(golden):
    """The reference's own turbulent 32x32x153 initial fields: 4 steps against the reference's run of the
    same 4 steps (norms and 8 columns), 10 steps against the oracle (whole fields)."""
    fx = golden("fixture_32x32x153.npz")
    g = golden("golden_fixture32_steps.npz")
    st = [fx[k] for k in ("u1", "u2", "u3", "p")]
    t0 = float(fx["t"])
    print("case: reference fixture 32x32x153", flush=True)
    h = dnship.DNS(48, 48, 153)
    h.upload(*st, t=t0)
    h.step(4)
    out = h.download()
    assert abs(h.scalars()["t"] - g["hist"][-1, 1]) <= 1e-13 * g["hist"][-1, 1]
    cols = g["cols"]
    for k, a in zip(("u1", "u2", "u3", "p"), out):
        assert abs(np.linalg.norm(a) - g["norm_" + k]) <= TOL * g["norm_" + k], k
        assert rel_l2(a[cols], g["cols_" + k]) <= 1e-11, k
    h.step(6)
    out = h.download()
    h.close()
    o = pyoracle.Oracle(48, 48, 153, threads=THREADS)
    o.set_state(*st, t=t0)
    o.step(10)
    for k, a, b in zip(("u1", "u2", "u3", "p"), out, o.get_state()):
        assert rel_l2(a, b) <= TOL, k


TRAJ = os.path.join(GOLDEN, "golden_traj_128x128x153.npz")


@pytest.mark.skipif(not os.path.exists(TRAJ), reason="golden_traj_128x128x153.npz not generated")
def test_trajectory_128x128x153_vs_reference(golden):
    """128x128x153 (192x192 dealiased) from the seeded IC, 1000 CFL steps, compared step by step with the
    reference's own run through dt (a function of the CFL maxima, so of the whole field), and at 10, 100
    and 1000 steps through the field norms and 10 whole columns.  Up to 10 steps the bar is 1e-12.  Beyond,
    the flow amplifies every rounding difference: the reference's own run from an IC perturbed by a
    relative 1e-15 drifts away from its unperturbed run, and the GPU's fields must stay within ten times
    that drift (never below 1e-12), its dt within 100 times (see below)."""
    g = golden("golden_traj_128x128x153.npz")
    Mx, Mz, ny = 192, 192, 153
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=1)
    for k, a in zip(("u1", "u2", "u3", "p"), st):
        assert abs(np.linalg.norm(a) - g["ic_norm_" + k]) <= 1e-15 * g["ic_norm_" + k], k
    print("case: 128x128x153, 1000 steps vs the reference's trajectory", flush=True)
    h = dnship.DNS(Mx, Mz, ny)
    h.upload(*st)
    dts, ts = [], []
    cols = g["cols"]
    worst = {}
    for stop in (int(s) for s in g["stops"]):
        while len(dts) < stop:
            h.step(1)
            sc = h.scalars()
            dts.append(sc["dt"])
            ts.append(sc["t"])
        out = h.download()
        for k, a in zip(("u1", "u2", "u3", "p"), out):
            nrm, pn = g["norm_%s_%d" % (k, stop)], g["pert_norm_%s_%d" % (k, stop)]
            ref, pert = g["cols_%s_%d" % (k, stop)], g["pert_cols_%s_%d" % (k, stop)]
            d = max(abs(np.linalg.norm(a) - nrm) / nrm, rel_l2(a[cols], ref))
            drift = max(abs(pn - nrm) / nrm, rel_l2(pert, ref))
            bound = TOL if stop <= 10 else max(TOL, 10 * drift)
            worst[(k, stop)] = (d, drift)
            print("%s after %d steps: GPU vs reference %.2e, 1e-15-perturbed reference vs reference %.2e"
                  % (k, stop, d, drift), flush=True)
            assert d <= bound, (k, stop, d, drift)
    h.close()
    ref_dt, pert_dt = g["hist"][:, 2], g["pert_hist"][:, 2]
    ddt = np.abs(np.array(dts) - ref_dt) / ref_dt
    drift_dt = np.abs(pert_dt - ref_dt) / ref_dt
    assert ddt[:10].max() <= TOL
    # step by step: against the perturbed run's largest dt drift up to that step (>= 1e-12).  The GPU run
    # rounds differently in every operation of every step, not once in its initial fields, so it starts
    # further from the reference than the 1e-15-perturbed run (5-10 times at 10 steps, above), and the flow
    # amplifies both alike; the perturbed run is one sample of that divergence, so the GPU's dt may lie up
    # to 100 times its envelope at a given step.
    run_max = np.maximum.accumulate(drift_dt)
    env = np.maximum(run_max * 100, TOL)
    ratio = ddt[10:] / np.maximum(run_max[10:], TOL / 100)
    print("dt after 10 steps: GPU vs reference at most %.1f times the perturbed run's drift so far (step %d)"
          % (ratio.max(), 10 + int(np.argmax(ratio))), flush=True)
    assert np.all(ddt <= env), int(np.argmax(ddt > env))
    assert abs(ts[-1] - g["hist"][-1, 1]) <= max(TOL, 100 * abs(g["pert_hist"][-1, 1] - g["hist"][-1, 1])) * ts[-1]


# The ensemble of three sizes: 24 x (32^2 x 153), 16 x (64^2 x 97), 8 x (128^2 x 153) modes; Re_tau per member
# uniform in [150, 400] (seed 7), bulk Re from Dean's correlation Re_tau = 0.09 (2 Re)^0.88, mpgx =
# -(Re_tau/Re)^2; member IC seed 100 + member index, broadband amplitude 0.1.
ENSEMBLE = [(24, 48, 48, 153), (16, 96, 96, 97), (8, 192, 192, 153)]


def ensemble_members():
    rng = np.random.default_rng(7)
    out, idx = [], 0
    for nm, Mx, Mz, ny in ENSEMBLE:
        grp = []
        for m in range(nm):
            ret = rng.uniform(150.0, 400.0)
            Re = 0.5 * (ret / 0.09) ** (1.0 / 0.88)
            grp.append((Re, -(ret / Re) ** 2, 100 + idx))
            idx += 1
        out.append(grp)
    return out


@pytest.mark.parametrize("launch", [0, 1])
@pytest.mark.parametrize("prec", [dnship.PREC_FP64, dnship.PREC_MIXED])
def test_ensemble_48_members_three_sizes(prec, launch):
    """The 48-member ensemble of three sizes, in one grouped launch per kernel class per substep with tile
    shapes from the ensemble autotuner (launch 0) or one persistent launch per substep over a task queue
    (launch 1), its RK3 step captured once in a
    hipGraph and replayed (3 steps): every member equals the same channel run alone (autotuned) bit for bit
    (own IC, Re, mpgx, CFL dt), and the first member of each size agrees with the oracle (FP64: <= 1e-12;
    mixed: its FP32 plane stage's error, <= 1e-6 after 3 steps)."""
    nsteps = 3
    members = ensemble_members()
    handles, ics = [], []
    for (nm, Mx, Mz, ny), grp in zip(ENSEMBLE, members):
        print("case: ensemble group %d x %dx%dx%d prec %d" % (nm, Mx, Mz, ny, prec), flush=True)
        e = dnship.DNS(Mx, Mz, ny, precision=prec, members=nm)
        sts = []
        for m, (Re, mp, sd) in enumerate(grp):
            st = dns_ic.seeded_channel(Mx, Mz, ny, seed=sd, amp=0.1, broadband=True)
            e.set_member_params(m, Re, mp)
            e.upload(*st, member=m)
            sts.append(st)
        handles.append(e)
        ics.append(sts)
    ens = dnship.Ensemble(handles)
    if launch:
        ens.set_launch(1)
        mode, nwg, ntask = ens.launch()
        print("persistent launch: %d workgroups, %s tasks per substep" % (nwg, ntask), flush=True)
        assert mode == 1 and nwg >= 256 and min(ntask) > 0
    else:
        # tile shapes from the ensemble's autotuner, which puts the fields back as it found them
        var, wg, ms = ens.autotune(iters=2)
        print("autotuned tiles: var %s wg %s" % (var, wg), flush=True)
        assert (var, wg) == ens.tiles()
        for c in range(5):
            k = dnship.Ensemble.PLANE_SHAPES.index((wg[c], var[c]))
            assert ms[c][k] == min(t for t in ms[c] if t >= 0)
        assert ms[5][var[5]] == min(ms[5])
    ens.step_graph(nsteps)
    ens.sync()  # raises if a task of the persistent launch waited too long
    ens.close()
    for (nm, Mx, Mz, ny), grp, sts, e in zip(ENSEMBLE, members, ics, handles):
        print("case: ensemble group %dx%dx%d members vs single channels" % (Mx, Mz, ny), flush=True)
        for m, ((Re, mp, _), st) in enumerate(zip(grp, sts)):
            got = e.download(member=m)
            s = dnship.DNS(Mx, Mz, ny, Re=Re, mpgx=mp, precision=prec)
            s.upload(*st)
            if m == 0:
                s.autotune(iters=2)
                s.upload(*st)
            s.step(nsteps)
            ref = s.download()
            sref = s.scalars()
            s.close()
            for x, y in zip(got, ref):
                assert np.array_equal(x, y), (Mx, m)
            sc = e.scalars(member=m)
            assert sc["t"] == sref["t"] and sc["dt"] == sref["dt"] and sc["iter"] == nsteps
        Re, mp, _ = grp[0]
        o = pyoracle.Oracle(Mx, Mz, ny, Re=Re, mpgx=mp, threads=THREADS)
        o.set_state(*sts[0])
        o.step(nsteps)
        got = e.download(member=0)
        for k, a, b in zip(("u1", "u2", "u3", "p"), got, o.get_state()):
            err = rel_l2(a, b)
            print("member 0 of %dx%dx%d: %s rel L2 vs oracle %.2e" % (Mx, Mz, ny, k, err), flush=True)
            if prec == dnship.PREC_FP64:
                assert err <= TOL, k
            elif k != "p":
                assert err <= 1e-6, k
        e.close()


@pytest.mark.parametrize("launch", [0, 1])
@pytest.mark.parametrize("prec", [dnship.PREC_FP64, dnship.PREC_MIXED])
def test_ensemble_grouped_eager_and_graph(prec, launch):
    """Grouped launches (launch 0) or the persistent launch (launch 1, one unit per task: the most tasks and
    waits) with a constant-flow-rate group (mean-flow correction after the column pass), a rectangular
    plane and a group of one: 2 eager steps then 2 graph-replayed steps equal every member's single-channel
    run bit for bit, t, dt, mpgx and iter included; the handles' own downloads wait for the ensemble's
    work."""
    groups = ENS_GROUPS
    handles, ics = [], []
    for Mx, Mz, ny, ctp, mem in groups:
        h = dnship.DNS(Mx, Mz, ny, precision=prec, members=len(mem), flag_ctpress=ctp)
        sts = []
        for m, (Re, mp, sd) in enumerate(mem):
            st = dns_ic.seeded_channel(Mx, Mz, ny, seed=sd, broadband=True)
            h.set_member_params(m, Re, mp)
            h.upload(*st, member=m)
            sts.append(st)
        handles.append(h)
        ics.append(sts)
    print("case: ensemble, 3 groups, prec %d, launch %d" % (prec, launch), flush=True)
    ens = dnship.Ensemble(handles)
    if launch:
        ens.set_launch(1, units={c: 1 for c in dnship.Ensemble.TASK_CLASSES})
    ens.step(2)
    ens.step_graph(2)
    ens.sync()
    for (Mx, Mz, ny, ctp, mem), sts, h in zip(groups, ics, handles):
        for m, ((Re, mp, _), st) in enumerate(zip(mem, sts)):
            got, sc = h.download(member=m), h.scalars(member=m)
            s = dnship.DNS(Mx, Mz, ny, Re=Re, mpgx=mp, precision=prec, flag_ctpress=ctp)
            s.upload(*st)
            s.step(4)
            ref, sref = s.download(), s.scalars()
            s.close()
            for x, y in zip(got, ref):
                assert np.array_equal(x, y), (Mx, Mz, ny, m)
            for k in ("t", "dt", "mpgx", "dgx", "Qx", "iter", "divmax"):
                assert sc[k] == sref[k], (Mx, Mz, ny, m, k)
    ens.close()
    for h in handles:
        h.close()


ENS_GROUPS = [(48, 48, 33, 0, [(3250.0, -3.0675e-3, 31), (2500.0, -3.5e-3, 32)]),
              (96, 192, 65, 1, [(4000.0, -2.8e-3, 33)]),
              (192, 96, 17, 1, [(3000.0, -3.2e-3, 34), (3500.0, -3.0e-3, 35), (2800.0, -3.3e-3, 36)])]


@pytest.mark.parametrize("prec", [dnship.PREC_FP64, dnship.PREC_MIXED])
def test_ensemble_every_tile_shape(prec):
    """Every tile shape of the grouped launches (the five plane shapes of every plane class, the three
    column-stage variants) gives every member the bits of its single-channel run (one step)."""
    handles, ics, refs = [], [], []
    for Mx, Mz, ny, ctp, mem in ENS_GROUPS:
        h = dnship.DNS(Mx, Mz, ny, precision=prec, members=len(mem), flag_ctpress=ctp)
        sts, rs = [], []
        for m, (Re, mp, sd) in enumerate(mem):
            st = dns_ic.seeded_channel(Mx, Mz, ny, seed=sd, broadband=True)
            h.set_member_params(m, Re, mp)
            s = dnship.DNS(Mx, Mz, ny, Re=Re, mpgx=mp, precision=prec, flag_ctpress=ctp)
            s.upload(*st)
            s.step(1)
            rs.append(s.download())
            s.close()
            sts.append(st)
        handles.append(h)
        ics.append(sts)
        refs.append(rs)
    ens = dnship.Ensemble(handles)
    assert ens.tiles() == ([1, 1, 1, 1, 1, 0], [128] * 5)    # the defaults
    for k, (wg, v) in enumerate(dnship.Ensemble.PLANE_SHAPES):
        var = [v] * 5 + [k % 3]
        print("case: ensemble tiles var %s wg %d prec %d" % (var, wg, prec), flush=True)
        ens.set_tiles(var, [wg] * 5)
        assert ens.tiles() == (var, [wg] * 5)
        for h, sts in zip(handles, ics):
            for m, st in enumerate(sts):
                h.upload(*st, member=m)
        ens.step(1)
        ens.sync()
        for (Mx, Mz, ny, _, mem), h, rs in zip(ENS_GROUPS, handles, refs):
            for m, ref in enumerate(rs):
                for x, y in zip(h.download(member=m), ref):
                    assert np.array_equal(x, y), (k, Mx, Mz, ny, m)
    ens.close()
    for h in handles:
        h.close()


def test_ensemble_rejects_unsupported_groups():
    a = dnship.DNS(48, 48, 33)
    b = dnship.DNS(576, 576, 17)                  # 576 points: not a grouped-launch length
    c = dnship.DNS(48, 48, 193)                   # 192 rows: beyond the grouped column pass
    d = dnship.DNS(48, 48, 33, precision=dnship.PREC_MIXED)
    for bad in ([a, b], [a, c], [a, d], [a, a], []):
        with pytest.raises(dnship.DnsError):
            dnship.Ensemble(bad)
    e = dnship.Ensemble([a])
    with pytest.raises(dnship.DnsError):          # no fields uploaded
        e.step(1)
    with pytest.raises(dnship.DnsError):          # no such launch mode
        e.set_launch(2)
    e.close()
    f = dnship.DNS(48, 48, 34)                    # ny - 1 = 33: z-inverse chunks of two planes are not whole
    e = dnship.Ensemble([a, f])
    with pytest.raises(dnship.DnsError):
        e.set_launch(1)
    assert e.launch()[0] == 0                     # still the grouped launches
    # several z-inverse bodies per workgroup need whole chunks of planes, too
    with pytest.raises(dnship.DnsError):
        e.set_tiles([0, 1, 1, 1, 1, 0], [256, 128, 128, 128, 128])
    e.set_tiles([1, 1, 1, 1, 1, 0], [128, 256, 128, 128, 128])
    for var, wg in (([2, 1, 1, 1, 1, 0], [128] * 5),   # 256-thread bodies in 128-thread workgroups
                    ([1, 3, 1, 1, 1, 0], [128] * 5),   # no variant 3
                    ([1, 1, 1, 1, 1, 3], [128] * 5),
                    ([1, 1, 1, 1, 1, 0], [512] * 5)):
        with pytest.raises(dnship.DnsError):
            e.set_tiles(var, wg)
    assert e.tiles() == ([1, 1, 1, 1, 1, 0], [128, 256, 128, 128, 128])
    e.close()
    for h in (a, b, c, d, f):
        h.close()


@pytest.mark.parametrize("Mx,Mz,ny,Re,retau,seed", [
    (576, 576, 257, 10000.0, 550.0, 3),      # 384 x 384 x 257, Re_tau ~ 550
    (1152, 576, 385, 19800.0, 1000.0, 5),    # 768 x 384 x 385, Re_tau ~ 1000
])
def test_full_size_step_vs_oracle(Mx, Mz, ny, Re, retau, seed):
    """One RK3 step at full size from the seeded broadband IC: FP64 against the oracle (OpenMP on the host
    cores) at 1e-12, and the mixed-precision step's error against the oracle (FP32 plane stage: <= 1e-6)."""
    mpgx = -(retau / Re) ** 2
    st = dns_ic.seeded_channel(Mx, Mz, ny, seed=seed, amp=0.1, broadband=True)
    res = {}
    for prec in (dnship.PREC_FP64, dnship.PREC_MIXED):
        print("case: full size %dx%dx%d prec %d, one step" % (Mx, Mz, ny, prec), flush=True)
        h = dnship.DNS(Mx, Mz, ny, Re=Re, mpgx=mpgx, precision=prec)
        h.upload(*st)
        h.step(1)
        res[prec] = (h.download(), h.download_nonlinear(), h.scalars())
        h.close()
    print("case: oracle %dx%dx%d on %d threads" % (Mx, Mz, ny, THREADS), flush=True)
    o = pyoracle.Oracle(Mx, Mz, ny, Re=Re, mpgx=mpgx, threads=THREADS)
    o.set_state(*st)
    o.step(1)
    ref, rnu, rsc = o.get_state(), o.get_nu(), o.scalars()
    del o
    out, nu, sc = res[dnship.PREC_FP64]
    for k, a, b in zip(("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3"), out + nu, ref + rnu):
        err = rel_l2(a, b)
        print("fp64 %s rel L2 vs oracle %.2e" % (k, err), flush=True)
        assert err <= TOL, k
    assert abs(sc["t"] - rsc["t"]) <= 1e-13 * rsc["t"] and sc["divmax"] < 1e-8
    mix = res[dnship.PREC_MIXED][0]
    for k, a, b in zip(("u1", "u2", "u3"), mix, ref):
        err = rel_l2(a, b)
        print("mixed %s rel L2 vs oracle %.2e" % (k, err), flush=True)
        assert 0.0 < err <= 1e-6, k
