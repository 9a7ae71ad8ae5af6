"""Benchmark: RK3 steps/s of the channel-flow DNS time step at 384 x 384 x 257 (576 x 576 dealiased
physical points) on one MI355X, FP64 and mixed precision, with the HBM roofline of the dominant kernel
and the reference's own CPU code timed on the host as a baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed]

Prints one JSON line.  value = FP64 steps/s (the fields, the RK3 accumulation and the solves all FP64);
the mixed-precision run (FP32 plane stage) is reported under "mixed" with its relative L2 velocity
error against the FP64 run after the same number of steps from the same initial state, and after
--error-steps (default 1000) steps beside the drift of an FP64 run from a 1e-10-perturbed state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "final-single-block-_amd")
sys.path.insert(0, PKG)

METRIC = "RK3 steps/s at 384x384x257: FP64 vs mixed-prec at <=1e-6 err; % HBM peak"
MX, MZ, NY = 576, 576, 257            # 384 x 384 Fourier modes x 257 y points
RE, RETAU = 10000.0, 550.0            # Re_tau = Re sqrt(-mpgx) under a constant pressure gradient
MPGX = -(RETAU / RE) ** 2
PLANE_KERNELS = ["zinv", "xpass", "zfwd"]
SWEEPS = ["col_up", "col_down", "col_final"]


def kernel_times(tim, path):
    """Per-kernel ms per substep; the one-pass column stage (path 0) is reported as col_pass."""
    k = {n: tim[n] for n in PLANE_KERNELS}
    if path == 0:
        k["col_pass"] = tim["col_up"]
    else:
        k.update({n: tim[n] for n in SWEEPS})
    return k


def gpu_leg(st, prec, steps, warmup, size=None, Re=RE, mpgx=MPGX):
    import dnship
    Mx, Mz, ny = size or (MX, MZ, NY)
    g = dnship.DNS(Mx, Mz, ny, Re=Re, mpgx=mpgx, precision=prec)
    g.upload(*st)
    chosen, ms = g.autotune(iters=3)   # tile variants of the plane kernels, timed on this workload
    g.upload(*st)
    g.step(warmup)
    g.sync()
    g.set_timing(1)
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    path = g.column_path()
    tim = kernel_times(g.timing(), path)
    sc = g.scalars()
    fields = g.download()
    g.close()
    tune = {"kernels": list(dnship.DNS.KERNELS), "chosen": [int(c) for c in chosen],
            "ms_per_variant": [[round(float(x), 4) for x in row] for row in ms],
            "column_path": "one pass" if path == 0 else "three sweeps"}
    return (t1 - t0) / steps, tim, sc, fields, tune


# Ensemble workload: 24 x (32^2 x 153), 16 x (64^2 x 97), 8 x (128^2 x 153) modes; Re_tau per member
# uniform in [150, 400] (seeded), bulk Re on the half height from Dean's correlation
# Re_tau = 0.09 (2 Re)^0.88, mpgx = -(Re_tau/Re)^2; member IC seed = 100 + member index.
ENSEMBLE = [(24, 48, 48, 153), (16, 96, 96, 97), (8, 192, 192, 153)]


def ensemble_leg(steps, warmup, prec, eager=False):
    """The 48-member ensemble (three sizes) with one grouped launch per kernel class per RK3 substep for
    all members, the step replayed from a captured hipGraph (value); beside it the same members as three
    handles on three streams, each handle's step autotuned and replayed from its own graph."""
    import numpy as np
    import dnship
    import dns_ic
    import dns_roofline
    rng = np.random.default_rng(7)
    handles, by, idx, ics = [], 0.0, 0, []
    for nm, Mx, Mz, ny in ENSEMBLE:
        g = dnship.DNS(Mx, Mz, ny, precision=prec, members=nm)
        sts = []
        for m in range(nm):
            ret = rng.uniform(150.0, 400.0)
            Re = 0.5 * (ret / 0.09) ** (1.0 / 0.88)
            g.set_member_params(m, Re, -(ret / Re) ** 2)
            st = dns_ic.seeded_channel(Mx, Mz, ny, seed=100 + idx, amp=0.1)
            g.upload(*st, member=m)
            sts.append(st)
            idx += 1
        handles.append(g)
        ics.append(sts)
        by += nm * dns_roofline.kernel_bytes(Mx, Mz, ny, prec == dnship.PREC_MIXED)["step"]

    def reupload():
        for g, sts in zip(handles, ics):
            for m, st in enumerate(sts):
                g.upload(*st, member=m)

    ens = dnship.Ensemble(handles)
    tvar, twg, tms = ens.autotune(iters=5)   # tile shapes of the grouped launches, timed on this workload
    ens_step = ens.step if eager else ens.step_graph
    ens_step(max(warmup, 1))
    ens.sync()
    t0 = time.perf_counter()
    ens_step(steps)
    ens.sync()
    sec = (time.perf_counter() - t0) / steps
    reupload()
    ens.set_timing(1)
    ens.step(2)
    ens.sync()
    tim = ens.timing()
    ens.set_timing(0)
    # the same members in one persistent launch per substep (task queue over every member and class)
    reupload()
    ens.set_launch(1)
    _, nwg, ntask = ens.launch()
    ens_step(max(warmup, 1))
    ens.sync()
    t0 = time.perf_counter()
    ens_step(steps)
    ens.sync()
    sec_persist = (time.perf_counter() - t0) / steps
    ens.close()
    # the same members as three handles on three streams, each autotuned, each step from its own graph
    for g in handles:
        g.autotune(iters=3)
    reupload()
    for g in handles:
        (g.step if eager else g.step_graph)(max(warmup, 1))
    for g in handles:
        g.sync()
    t0 = time.perf_counter()
    for g in handles:
        (g.step if eager else g.step_graph)(steps)
    for g in handles:
        g.sync()
    sec_streams = (time.perf_counter() - t0) / steps
    for g in handles:
        g.close()
    return {"value": 1.0 / sec, "unit": "ensemble steps/s (all 48 members advance one RK3 step)",
            "ms_per_step": sec * 1e3, "members": idx, "groups": [list(x) for x in ENSEMBLE],
            "launch": "one grouped launch per kernel class per substep for all members, tile shapes from the "
                      "ensemble autotuner, " + ("eager steps" if eager else "step replayed from a hipGraph"),
            "tiles": {"classes": list(dnship.Ensemble.TILE_CLASSES), "variant": tvar, "workgroup": twg,
                      "ms_per_shape": [[round(t, 4) for t in c] for c in tms]},
            "kernels_ms_per_substep": {k: v for k, v in tim.items() if k != "n"},
            "algorithmic_GB_per_step": by / 1e9, "frac_hbm": by / sec / 1e9 / dns_roofline.HBM_PEAK_GBS,
            "persistent": {"value": 1.0 / sec_persist, "ms_per_step": sec_persist * 1e3,
                           "launch": "one persistent launch per substep: %d workgroups over a queue of "
                                     "%d/%d/%d tasks (substeps 1/2/3)" % (nwg, *ntask)},
            "three_streams": {"value": 1.0 / sec_streams, "ms_per_step": sec_streams * 1e3,
                              "launch": "one handle per size on its own stream, autotuned, own hipGraph"}}


def cpu_leg(st):
    """The reference's own hot path (oracle/_ref, compiled from its sources) on one host core, timed on a
    bounded sample: one RK3 substep of the same 384x384x257 workload; falls back to the oracle's C
    restatement on all host cores when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if pyoracle.have_reference():
        t = pyoracle.time_reference_substeps(MX, MZ, NY, *st, nsub=1, Re=RE, mpgx=MPGX)
        return {"value": 1.0 / (3.0 * t), "unit": "steps/s", "cores": 1, "kind": "reference",
                "sample": "one RK3 substep (kRK=1) of the 384x384x257 workload, single MPI rank, %.1f s; "
                          "steps/s = 1/(3 x substep time)" % t}
    return cpu_all_cores_leg(st)


def cpu_all_cores_leg(st):
    """The oracle's C restatement of the step on every host core the process may use (OpenMP), one RK3
    step of the 384x384x257 workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    n = len(os.sched_getaffinity(0))
    o = pyoracle.Oracle(MX, MZ, NY, Re=RE, mpgx=MPGX, threads=n)
    o.set_state(*st)
    t0 = time.perf_counter()
    o.step(1)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "steps/s", "cores": n, "kind": "port",
            "sample": "one RK3 step of the 384x384x257 workload, OpenMP over %d threads, %.1f s" % (n, t)}


def class_fracs(kms, fp32):
    """HBM fraction of the plane stage (z-inverse + x-pass + z-forward) and of the column stage, each
    charged its kernels' own compulsory bytes (dns_roofline.kernel_bytes) over their event-timed ms per
    substep."""
    import dns_roofline
    by = dns_roofline.kernel_bytes(MX, MZ, NY, fp32)
    plane_ms = sum(kms[k] for k in PLANE_KERNELS)
    col = [k for k in kms if k not in PLANE_KERNELS]
    col_ms = sum(kms[k] for k in col)
    plane_b = sum(by[k] for k in PLANE_KERNELS)
    col_b = sum(by[k] for k in col)
    f = lambda b, ms: b / (ms * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS
    return {"plane_stage": f(plane_b, plane_ms), "column": f(col_b, col_ms),
            "plane_stage_GB_per_substep": plane_b / 1e9, "column_GB_per_substep": col_b / 1e9,
            "plane_stage_ms_per_substep": plane_ms, "column_stage_ms_per_substep": col_ms}


def traffic(kernel, prec_tag):
    path = os.path.join(ROOT, "profiles", "traffic_%s_384x384x257_%s.json" % (kernel, prec_tag))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


# The largest configuration: 768 x 384 Fourier modes x 385 y points (1152 x 576 dealiased), Re_tau ~ 1000
LARGE = (1152, 576, 385)
LARGE_RE = 19800.0
LARGE_MPGX = -(1000.0 / LARGE_RE) ** 2


def large_leg(steps, warmup):
    """FP64 and mixed steps/s at 768x384x385 (seeded broadband IC, seed 5), with the HBM fractions of the
    step's own and reference-design bytes and the mixed run's error against FP64."""
    import numpy as np
    import dnship
    import dns_ic
    import dns_roofline
    st = dns_ic.seeded_channel(*LARGE, seed=5, amp=0.1, broadband=True)
    out = {"workload": "RK3 step, 768x384 Fourier modes x 385 y points (1152x576 dealiased), Re=19800, "
                       "Re_tau~1000"}
    res = {}
    for tag, prec in (("fp64", dnship.PREC_FP64), ("mixed", dnship.PREC_MIXED)):
        sec, tim, sc, f, _ = gpu_leg(st, prec, steps, warmup, LARGE, LARGE_RE, LARGE_MPGX)
        fp32 = prec == dnship.PREC_MIXED
        by = dns_roofline.kernel_bytes(*LARGE, fp32)["step"]
        ref = dns_roofline.reference_design_step_bytes(*LARGE, fp32)
        out[tag] = {"value": 1.0 / sec, "ms_per_step": sec * 1e3, "kernels_ms": tim,
                    "step_frac_hbm": by / sec / 1e9 / dns_roofline.HBM_PEAK_GBS,
                    "step_frac_hbm_reference_design": ref / sec / 1e9 / dns_roofline.HBM_PEAK_GBS,
                    "max_div": sc["divmax"]}
        res[tag] = f
    out["mixed"]["speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    out["mixed"]["rel_l2_vs_fp64"] = {k: float(np.linalg.norm(a - b) / np.linalg.norm(b))
                                      for k, a, b in zip(("u1", "u2", "u3"), res["mixed"], res["fp64"])}
    return out


def error_leg(st, nsteps):
    """Relative L2 velocity error of the mixed-precision path against FP64 after nsteps steps from the same
    state (each path in its own handle, CFL time step; not timed), and, for scale, the difference that an
    FP64 run from the same state with every velocity value perturbed by a relative 1e-10 (Gaussian, seed
    11) has grown to by then."""
    import numpy as np
    import dnship
    rng = np.random.default_rng(11)
    pert = [a * (1 + 1e-10 * rng.standard_normal(a.shape)) for a in st[:3]] + [st[3]]
    out = {}
    for tag, prec, state in (("fp64", dnship.PREC_FP64, st), ("mixed", dnship.PREC_MIXED, st),
                             ("fp64_perturbed", dnship.PREC_FP64, pert)):
        g = dnship.DNS(MX, MZ, NY, Re=RE, mpgx=MPGX, precision=prec)
        g.upload(*state)
        done = 0
        while done < nsteps:          # in chunks, so a long run reports progress
            n = min(100, nsteps - done)
            g.step(n)
            g.sync()
            done += n
            print("error leg: %s %d/%d steps" % (tag, done, nsteps), file=sys.stderr, flush=True)
        out[tag] = (g.download(), g.scalars())
        g.close()
    (a, sa), (b, sb), (c, _) = out["fp64"], out["mixed"], out["fp64_perturbed"]
    rel = lambda u, v: {k: float(np.linalg.norm(y - x) / np.linalg.norm(x))
                        for k, x, y in zip(("u1", "u2", "u3"), u[:3], v[:3])}
    return {"steps": nsteps, "t_fp64": sa["t"], "t_mixed": sb["t"], "rel_l2": rel(a, b),
            "fp64_perturbed_1e-10_rel_l2": rel(a, c)}


# The smaller single-channel configurations: 32 x 32 x 153 (the reference's CPU-runnable case) and
# 128 x 128 x 153 modes, Re = 3250, Re_tau = 180 (mpgx = -3.0675e-3), seeded broadband IC (seed 1).
SMALL = [(48, 48, 153), (192, 192, 153)]


def small_leg(steps, warmup, eager=False):
    """FP64 and mixed steps/s of single channels at the smaller configurations, each RK3 step replayed from
    a captured hipGraph (at these sizes the launches, not HBM, set the pace without one)."""
    import dnship
    import dns_ic
    import dns_roofline
    out = {}
    for Mx, Mz, ny in SMALL:
        st = dns_ic.seeded_channel(Mx, Mz, ny, seed=1, amp=0.1, broadband=True)
        key = "%dx%dx%d" % (2 * Mx // 3, 2 * Mz // 3, ny)
        out[key] = {}
        for tag, prec in (("fp64", dnship.PREC_FP64), ("mixed", dnship.PREC_MIXED)):
            g = dnship.DNS(Mx, Mz, ny, Re=3250.0, mpgx=-3.0675e-3, precision=prec)
            g.upload(*st)
            g.autotune(iters=3)
            g.upload(*st)
            step = g.step if eager else g.step_graph
            step(max(warmup, 1))
            g.sync()
            t0 = time.perf_counter()
            step(steps)
            g.sync()
            sec = (time.perf_counter() - t0) / steps
            g.close()
            by = dns_roofline.reference_design_step_bytes(Mx, Mz, ny, prec == dnship.PREC_MIXED)
            out[key][tag] = {"value": 1.0 / sec, "ms_per_step": sec * 1e3,
                             "step_frac_hbm_reference_design": by / sec / 1e9 / dns_roofline.HBM_PEAK_GBS}
    return out


def stats_leg(st, reps=10):
    """dns_accumulate_stats on the 384x384x257 state (plane statistics and spectra of stats.f90 and
    spectra.f90, on the device): time per accumulation (host clock around synchronised calls) and its HBM
    rate for the algorithmic bytes (u1, u2, u3, p read once; the five spectra read and written once)."""
    import dnship
    import dns_roofline
    g = dnship.DNS(MX, MZ, NY, Re=RE, mpgx=MPGX)
    g.upload(*st)
    for _ in range(3):
        g.accumulate_stats()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.accumulate_stats()
    g.sync()
    sec = (time.perf_counter() - t0) / reps
    C = g.C
    nyu, nyv = NY - 1, NY - 2
    fields = (2 * (nyu + 2) + (nyv + 2) + nyu) * C * 16
    spectra = 2 * (3 * (nyu + 2) + (nyv + 2) + nyu) * C * 8
    g.close()
    by = fields + spectra
    return {"ms": sec * 1e3, "algorithmic_GB": by / 1e9, "GBs": by / sec / 1e9,
            "frac_hbm": by / sec / 1e9 / dns_roofline.HBM_PEAK_GBS,
            "what": "one accumulation of the 13 plane statistics and 5 spectra at 384x384x257 (FP64)"}


def post_leg(u1):
    """four_to_phys of the u1 field (384 x 384 modes, 258 planes) to the 576 x 576 grid on the GPU: kernel
    time of dns_four_to_phys (event-timed, host copies outside) and its algorithmic HBM rate (input and
    output planes once)."""
    import numpy as np
    import dnship
    import dns_io
    import dns_roofline
    cx = dns_io.columns_to_planes(u1, 2 * MZ // 3)            # [planes, Nz, Kx] complex
    planes = np.empty(cx.shape[:2] + (2 * cx.shape[2],))
    planes[:, :, 0::2] = cx.real
    planes[:, :, 1::2] = cx.imag
    dnship.four_to_phys(planes[:2], MX, MZ)                    # load the kernels
    _, ms = dnship.four_to_phys(planes, MX, MZ)
    n, nz, nx = planes.shape
    byts = n * (nx * nz + (MX + 2) * MZ) * 8.0
    mid = n * 2 * (nx // 2) * MZ * 16.0
    return {"planes": n, "ms": ms, "GBs": byts / (ms * 1e-3) / 1e9,
            "frac_hbm": byts / (ms * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS,
            "algorithmic_GB": byts / 1e9, "intermediate_GB": mid / 1e9,
            "what": "u1 snapshot (%d planes of %d x %d modes) to %d x %d physical planes" % (n, nx - 2, nz, MX, MZ)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-post", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="step the ensemble and small legs eagerly instead of replaying captured hipGraphs "
                         "(for kernel traces)")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="the mixed path's error against FP64 after this many steps, with the drift of a "
                         "1e-10-perturbed FP64 run beside it (0: leave out)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (every box here has one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dnship
    import dns_ic
    import dns_roofline

    st = dns_ic.seeded_channel(MX, MZ, NY, seed=3, amp=0.1, broadband=True)
    sec64, tim64, sc64, f64, tune64 = gpu_leg(st, dnship.PREC_FP64, args.steps, args.warmup)
    out = {"metric": METRIC, "value": 1.0 / sec64, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sec64 * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: laminar Poiseuille + seeded (seed 3) broadband divergence-free perturbation, "
                   "amplitude 0.1",
           "config": {"workload": "RK3 step, 384x384 Fourier modes x 257 y points (576x576 dealiased), "
                                  "box 2pi x 2 x pi, Re=10000, mpgx=-(550/Re)^2 (Re_tau~550), CFL=0.7",
                      "Mx": MX, "Mz": MZ, "ny": NY, "precision": "fp64"}}
    kms = tim64
    dom = max(kms, key=lambda k: kms[k])
    roof = dns_roofline.roofline(dom, MX, MZ, NY, False, kms[dom])
    roof["kernel"] = dom
    roof["avg_ms"] = kms[dom]
    roof["traffic"] = traffic(dom, "fp64")
    out["roofline"] = roof
    by = dns_roofline.kernel_bytes(MX, MZ, NY, False)
    out["kernels"] = {k: {"avg_ms": kms[k], "GBs": by[k] / (kms[k] * 1e-3) / 1e9,
                          "frac_hbm": by[k] / (kms[k] * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS}
                      for k in kms}
    out["step_GB"] = by["step"] / 1e9
    out["step_frac_hbm"] = by["step"] / sec64 / 1e9 / dns_roofline.HBM_PEAK_GBS
    ref64 = dns_roofline.reference_design_step_bytes(MX, MZ, NY, False)
    out["reference_design_step_GB"] = ref64 / 1e9
    out["step_frac_hbm_reference_design"] = ref64 / sec64 / 1e9 / dns_roofline.HBM_PEAK_GBS
    out["class_frac_hbm"] = class_fracs(kms, False)
    cls = dns_roofline.reference_design_substep_classes(MX, MZ, NY, False)
    plane_ms = sum(kms[k] for k in PLANE_KERNELS)
    col_ms = sum(v for k, v in kms.items() if k not in PLANE_KERNELS)
    out["class_frac_hbm_reference_design"] = {
        "note": "comparison only: the bytes the reference's own decomposition would move (its 12 transposes "
                "per substep included), not what these kernels must move",
        "fft_and_transposes": cls["fft_and_transposes"] / (plane_ms * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS,
        "column": cls["column"] / (col_ms * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS}
    out["max_div"] = sc64["divmax"]
    out["autotune"] = {"fp64": tune64}
    if not args.no_mixed:
        secmx, timmx, scmx, fmx, tunemx = gpu_leg(st, dnship.PREC_MIXED, args.steps, args.warmup)
        err = {k: float(np.linalg.norm(a - b) / np.linalg.norm(b)) for k, a, b in zip(("u1", "u2", "u3"), fmx, f64)}
        bym = dns_roofline.kernel_bytes(MX, MZ, NY, True)
        out["mixed"] = {"value": 1.0 / secmx, "ms_per_step": secmx * 1e3, "speedup_vs_fp64": sec64 / secmx,
                        "rel_l2_vs_fp64": err, "steps_compared": args.steps + args.warmup,
                        "kernels_ms": timmx, "step_GB": bym["step"] / 1e9,
                        "step_frac_hbm": bym["step"] / secmx / 1e9 / dns_roofline.HBM_PEAK_GBS,
                        "step_frac_hbm_reference_design":
                            dns_roofline.reference_design_step_bytes(MX, MZ, NY, True) / secmx / 1e9
                            / dns_roofline.HBM_PEAK_GBS}
        out["mixed"]["class_frac_hbm"] = class_fracs(timmx, True)
        clm = dns_roofline.reference_design_substep_classes(MX, MZ, NY, True)
        pm = sum(timmx[k] for k in PLANE_KERNELS)
        cm = sum(v for k, v in timmx.items() if k not in PLANE_KERNELS)
        out["mixed"]["class_frac_hbm_reference_design"] = {
            "fft_and_transposes": clm["fft_and_transposes"] / (pm * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS,
            "column": clm["column"] / (cm * 1e-3) / 1e9 / dns_roofline.HBM_PEAK_GBS}
        out["autotune"]["mixed"] = tunemx
        if args.error_steps > 0:
            out["mixed"]["long_run_error"] = error_leg(st, args.error_steps)
    if not args.no_ensemble:
        out["ensemble"] = {"fp64": ensemble_leg(args.steps, args.warmup, dnship.PREC_FP64, args.eager)}
        if not args.no_mixed:
            out["ensemble"]["mixed"] = ensemble_leg(args.steps, args.warmup, dnship.PREC_MIXED, args.eager)
    if not args.no_large:
        out["large"] = large_leg(max(1, min(args.steps, 5)), 1)
    if not args.no_small:
        out["small"] = small_leg(max(args.steps, 20) * 5, args.warmup, args.eager)
    if not args.no_post:
        out["four_to_phys"] = post_leg(f64[0])
        out["stats"] = stats_leg(st)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(st)
        if out["cpu_baseline"]["kind"] == "reference":
            out["cpu_port_all_cores"] = cpu_all_cores_leg(st)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
