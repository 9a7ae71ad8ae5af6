"""ctypes binding of the C-ABI in include/dns_hip.h (libdns_hip.so, built next to this file).

This is plumbing for Python callers (tests, bench): the product path is the HIP library; there is no
Python or CPU fallback.  Loading fails loudly when the library is missing.

Column fields cross the boundary in the reference layout x(j, column): numpy complex128 arrays of
shape [ncol, nrows] (C-contiguous, so j is fastest, as in Fortran x(j, column)).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libdns_hip.so")

PREC_FP64 = 0
PREC_MIXED = 1


class DnsError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("Mx", ctypes.c_int), ("Mz", ctypes.c_int), ("ny", ctypes.c_int), ("ppp", ctypes.c_int),
                ("flag_ctpress", ctypes.c_int), ("precision", ctypes.c_int),
                ("Re", ctypes.c_double), ("alp", ctypes.c_double), ("bet", ctypes.c_double),
                ("CFL", ctypes.c_double), ("mpgx", ctypes.c_double), ("dyq", ctypes.c_double),
                ("fixed_dt", ctypes.c_double), ("plane_batch", ctypes.c_int), ("members", ctypes.c_int),
                ("column_path", ctypes.c_int), ("reserved", ctypes.c_int * 5)]


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None

EXPORTS = ["dns_last_error", "dns_ygrid", "dns_create", "dns_destroy", "dns_sizes", "dns_setup_arrays",
           "dns_upload_fields", "dns_download_fields", "dns_download_nonlinear", "dns_upload_nonlinear",
           "dns_download_rhs", "dns_download_flux", "dns_download_pressure_correction", "dns_get_scalars",
           "dns_set_dt", "dns_rhs0", "dns_nonlinear", "dns_solveU", "dns_meanflow_ctU", "dns_solveP",
           "dns_vcorr", "dns_meanflow_ctP", "dns_rk3_substep", "dns_rk3_step", "dns_rk3_step_graph",
           "dns_plane_fft_roundtrip", "dns_modes_to_planes", "dns_planes_to_modes", "dns_sync",
           "dns_set_timing", "dns_get_timing", "dns_get_plane_variants", "dns_set_plane_variants",
           "dns_autotune", "dns_upload_member", "dns_download_member", "dns_set_member_params",
           "dns_get_member_scalars", "dns_set_column_path", "dns_get_column_path",
           "dns_history_record", "dns_accumulate_stats", "dns_get_stats", "dns_reset_stats",
           "dns_refresh_stats_pressure", "dns_four_to_phys", "dns_four_to_phys_noalloc", "dns_ensemble_create",
           "dns_ensemble_destroy",
           "dns_ensemble_step", "dns_ensemble_step_graph", "dns_ensemble_sync", "dns_ensemble_set_timing",
           "dns_ensemble_get_timing", "dns_ensemble_set_launch", "dns_ensemble_get_launch",
           "dns_ensemble_set_tiles", "dns_ensemble_get_tiles", "dns_ensemble_autotune"]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise DnsError("%s is missing: run __graft_entry__.build() (no CPU fallback exists)" % LIBPATH)
    L = ctypes.CDLL(LIBPATH)
    L.dns_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    L.dns_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.dns_destroy.argtypes = [vp]
    L.dns_sizes.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.dns_setup_arrays.argtypes = [vp] + [_dp] * 7
    L.dns_ygrid.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int] + [_dp] * 4
    L.dns_upload_fields.argtypes = [vp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.dns_upload_member.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double]
    L.dns_download_member.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.dns_set_member_params.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.dns_get_member_scalars.argtypes = [vp, ctypes.c_int, _dp]
    for n in ("dns_download_fields",):
        getattr(L, n).argtypes = [vp, _dp, _dp, _dp, _dp]
    for n in ("dns_download_nonlinear", "dns_upload_nonlinear", "dns_download_rhs", "dns_download_flux"):
        getattr(L, n).argtypes = [vp, _dp, _dp, _dp]
    L.dns_download_pressure_correction.argtypes = [vp, _dp, _dp]
    L.dns_get_scalars.argtypes = [vp, _dp]
    L.dns_set_dt.argtypes = [vp, ctypes.c_double]
    for n in ("dns_rhs0", "dns_nonlinear", "dns_solveU", "dns_meanflow_ctU", "dns_solveP", "dns_vcorr",
              "dns_rk3_substep", "dns_rk3_step", "dns_rk3_step_graph", "dns_set_timing"):
        getattr(L, n).argtypes = [vp, ctypes.c_int]
    L.dns_meanflow_ctP.argtypes = [vp]
    L.dns_sync.argtypes = [vp]
    L.dns_plane_fft_roundtrip.argtypes = [vp, vp, vp, vp, ctypes.c_int]
    L.dns_modes_to_planes.argtypes = [vp, vp, ctypes.c_int, vp]
    L.dns_planes_to_modes.argtypes = [vp, vp, ctypes.c_int, vp]
    L.dns_get_timing.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int)]
    _ip = ctypes.POINTER(ctypes.c_int)
    L.dns_get_plane_variants.argtypes = [vp, _ip, _ip]
    L.dns_set_plane_variants.argtypes = [vp, _ip]
    L.dns_autotune.argtypes = [vp, ctypes.c_int, _ip, _dp]
    L.dns_set_column_path.argtypes = [vp, ctypes.c_int]
    L.dns_history_record.argtypes = [vp, ctypes.c_int, _dp]
    L.dns_accumulate_stats.argtypes = [vp]
    L.dns_reset_stats.argtypes = [vp]
    L.dns_refresh_stats_pressure.argtypes = [vp]
    L.dns_get_stats.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    L.dns_four_to_phys.argtypes = [ctypes.c_int] * 5 + [_dp, _dp, _dp]
    L.dns_four_to_phys_noalloc.argtypes = [ctypes.c_int] * 5 + [vp, vp, vp]
    L.dns_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    for n in ("dns_ensemble_destroy", "dns_ensemble_sync"):
        getattr(L, n).argtypes = [vp]
    for n in ("dns_ensemble_step", "dns_ensemble_step_graph", "dns_ensemble_set_timing"):
        getattr(L, n).argtypes = [vp, ctypes.c_int]
    L.dns_ensemble_get_timing.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int)]
    L.dns_ensemble_set_launch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.dns_ensemble_get_launch.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]
    L.dns_ensemble_set_tiles.argtypes = [vp, _ip, _ip]
    L.dns_ensemble_get_tiles.argtypes = [vp, _ip, _ip]
    L.dns_ensemble_autotune.argtypes = [vp, ctypes.c_int, _ip, _ip, _dp]
    L.dns_get_column_path.argtypes = [vp, _ip]
    _lib = L
    return L


def _chk(rc):
    if rc != 0:
        raise DnsError("status %d: %s" % (rc, lib().dns_last_error().decode()))


def _p(a):
    return a.ctypes.data_as(_dp)


def ygrid(ny, dyq=10.0, ppp=5):
    """Host-only y grid (start.f90:665-806): dict(yu, yv, dthdyu, dthdyv)."""
    yu, yv, du, dv = np.zeros(ny + 1), np.zeros(ny), np.zeros(ny + 1), np.zeros(ny)
    _chk(lib().dns_ygrid(int(ny), float(dyq), int(ppp), _p(yu), _p(yv), _p(du), _p(dv)))
    return dict(yu=yu, yv=yv, dthdyu=du, dthdyv=dv)



def four_to_phys(planes, Mx, Mz):
    """Spectral planes of a field file [n, Nz, Nx+2] (float64, the file's buffSR1 layout) to physical planes
    [n, Mz, Mx+2] on the GPU (four_to_phys.f90 with Mx, Mz = 3/2 Nx, Nz; four_to_phys_N.f90 with Mx, Mz =
    Nx, Nz).  Returns (planes, device milliseconds of the transform kernels)."""
    planes = np.ascontiguousarray(planes, dtype=np.float64)
    n, Nz, nx = planes.shape
    out = np.empty((n, Mz, Mx + 2))
    ms = ctypes.c_double()
    _chk(lib().dns_four_to_phys(nx - 2, Nz, Mx, Mz, n, _p(planes), _p(out), ctypes.byref(ms)))
    return out, ms.value

def four_to_phys_noalloc(d_in, d_out, n, Nx, Nz, Mx, Mz, stream=None):
    """dns_four_to_phys_noalloc on device pointers (ints): n dense planes (Nx+2) x Nz at d_in to n dense
    physical planes (Mx+2) x Mz at d_out, with no device allocation (the library's static workspace)."""
    _chk(lib().dns_four_to_phys_noalloc(int(Nx), int(Nz), int(Mx), int(Mz), int(n), d_in, d_out, stream))


class DNS:
    """One channel on the GPU (one dns_handle)."""

    def __init__(self, Mx, Mz, ny, Re=3250.0, alp=1.0, bet=2.0, CFL=0.7, mpgx=-3.0675e-3, dyq=10.0, ppp=5,
                 flag_ctpress=1, precision=PREC_FP64, fixed_dt=0.0, plane_batch=0, members=1,
                 column_path=0):
        L = lib()
        self.prm = Params(Mx=Mx, Mz=Mz, ny=ny, ppp=ppp, flag_ctpress=flag_ctpress, precision=precision,
                          Re=Re, alp=alp, bet=bet, CFL=CFL, mpgx=mpgx, dyq=dyq, fixed_dt=fixed_dt,
                          plane_batch=plane_batch, members=members, column_path=column_path)
        self.members = max(1, members)
        h = ctypes.c_void_p()
        _chk(L.dns_create(ctypes.byref(self.prm), ctypes.byref(h)))
        self.h = h
        s = (ctypes.c_longlong * 10)()
        _chk(L.dns_sizes(self.h, s))
        self.Nx, self.Nz, self.Kx, self.C, self.nyu, self.nyv = (int(s[q]) for q in range(6))
        self.Mx, self.Mz, self.ny, self.mib = int(s[6]), int(s[7]), int(s[8]), int(s[9])

    def close(self):
        if getattr(self, "h", None):
            lib().dns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shapes in the reference layout
    def shape(self, name):
        C, ny = self.C, self.ny
        return {"u1": (C, ny + 1), "u3": (C, ny + 1), "u2": (C, ny), "p": (C, ny - 1),
                "Nu1": (C, ny - 1), "Nu3": (C, ny - 1), "Nu2": (C, ny - 2)}[name]

    def setup_arrays(self):
        ny = self.ny
        yu, yv, du, dv = np.zeros(ny + 1), np.zeros(ny), np.zeros(ny + 1), np.zeros(ny)
        d2u, d2v, w = np.zeros((ny + 1, 3)), np.zeros((ny, 3)), np.zeros(6)
        _chk(lib().dns_setup_arrays(self.h, *[_p(a) for a in (yu, yv, du, dv, d2u, d2v, w)]))
        return dict(yu=yu, yv=yv, dthdyu=du, dthdyv=dv, dyu2i=d2u, dyv2i=d2v, dtv=w[0], gw=w[1:3], gbc=w[3],
                    gi=w[4:6])

    def _fields(self, u1, u2, u3, p):
        arrs = []
        for name, a in (("u1", u1), ("u2", u2), ("u3", u3), ("p", p)):
            a = np.ascontiguousarray(a, dtype=np.complex128)
            if a.shape != self.shape(name):
                raise DnsError("%s has shape %s, expected %s" % (name, a.shape, self.shape(name)))
            arrs.append(a)
        return arrs

    def upload(self, u1, u2, u3, p, t=0.0, it=0, QxT=0.0, member=None):
        """Fields into every ensemble member (member=None) or into one member."""
        arrs = self._fields(u1, u2, u3, p)
        ptrs = [_p(a.view(np.float64)) for a in arrs]
        if member is None:
            _chk(lib().dns_upload_fields(self.h, *ptrs, float(t), int(it), float(QxT)))
        else:
            _chk(lib().dns_upload_member(self.h, int(member), *ptrs, float(t), int(it), float(QxT)))

    def download(self, member=0):
        out = [np.zeros(self.shape(n), np.complex128) for n in ("u1", "u2", "u3", "p")]
        _chk(lib().dns_download_member(self.h, int(member), *[_p(a.view(np.float64)) for a in out]))
        return out

    def set_member_params(self, member, Re, mpgx):
        _chk(lib().dns_set_member_params(self.h, int(member), float(Re), float(mpgx)))

    def download_nonlinear(self):
        out = [np.zeros(self.shape(n), np.complex128) for n in ("Nu1", "Nu2", "Nu3")]
        _chk(lib().dns_download_nonlinear(self.h, *[_p(a.view(np.float64)) for a in out]))
        return out

    def upload_nonlinear(self, Nu1, Nu2, Nu3):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (Nu1, Nu2, Nu3)]
        _chk(lib().dns_upload_nonlinear(self.h, *[_p(a.view(np.float64)) for a in arrs]))

    def download_rhs(self):
        out = [np.zeros(self.shape(n), np.complex128) for n in ("Nu1", "Nu2", "Nu3")]
        _chk(lib().dns_download_rhs(self.h, *[_p(a.view(np.float64)) for a in out]))
        return out

    def download_flux(self):
        out = [np.zeros(self.shape(n), np.complex128) for n in ("u2", "u1", "u2")]
        _chk(lib().dns_download_flux(self.h, *[_p(a.view(np.float64)) for a in out]))
        return out

    def download_psi_div(self):
        out = [np.zeros(self.shape("p"), np.complex128) for _ in range(2)]
        _chk(lib().dns_download_pressure_correction(self.h, *[_p(a.view(np.float64)) for a in out]))
        return out

    def scalars(self, member=0):
        s = np.zeros(10)
        _chk(lib().dns_get_member_scalars(self.h, int(member), _p(s)))
        return dict(zip(["t", "dt", "dti", "mpgx", "Qx", "dgx", "dtc", "QxT", "iter", "divmax"], s))

    def set_dt(self, dt):
        _chk(lib().dns_set_dt(self.h, float(dt)))

    # ---- staged entry points (one per reference subroutine)
    def rhs0(self, k): _chk(lib().dns_rhs0(self.h, k))
    def nonlinear(self, k): _chk(lib().dns_nonlinear(self.h, k))
    def solveU(self, k): _chk(lib().dns_solveU(self.h, k))
    def meanflow_ctU(self, k): _chk(lib().dns_meanflow_ctU(self.h, k))
    def solveP(self, k): _chk(lib().dns_solveP(self.h, k))
    def vcorr(self, k): _chk(lib().dns_vcorr(self.h, k))
    def meanflow_ctP(self): _chk(lib().dns_meanflow_ctP(self.h))

    # ---- fused path
    def substep(self, k): _chk(lib().dns_rk3_substep(self.h, k))
    def step(self, n=1): _chk(lib().dns_rk3_step(self.h, n))
    def step_graph(self, n=1): _chk(lib().dns_rk3_step_graph(self.h, n))
    def sync(self): _chk(lib().dns_sync(self.h))

    def set_timing(self, on): _chk(lib().dns_set_timing(self.h, 1 if on else 0))

    def timing(self):
        t = np.zeros(7)
        n = ctypes.c_int()
        _chk(lib().dns_get_timing(self.h, _p(t), ctypes.byref(n)))
        return dict(zinv=t[0], xpass=t[1], zfwd=t[2], col_up=t[3], col_down=t[4], col_final=t[5],
                    substep=t[6], n=n.value)

    # ---- plane-kernel tile variants and the autotuner
    KERNELS = ("zinv", "xpass_u", "xpass_v", "zfwd_u", "zfwd_v", "col_up", "col_down", "col_final", "col_pass")
    NKERN = 9

    def plane_variants(self):
        """(current variant per kernel, number of variants per kernel), int arrays of NKERN."""
        v = np.zeros(self.NKERN, np.int32)
        c = np.zeros(self.NKERN, np.int32)
        _ip = ctypes.POINTER(ctypes.c_int)
        _chk(lib().dns_get_plane_variants(self.h, v.ctypes.data_as(_ip), c.ctypes.data_as(_ip)))
        return v, c

    def set_plane_variants(self, var):
        v = np.ascontiguousarray(var, np.int32)
        assert v.shape == (self.NKERN,)
        _chk(lib().dns_set_plane_variants(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def autotune(self, iters=5):
        """Times every variant; returns (chosen[NKERN], ms[NKERN, 3]) with -1 where not timed."""
        ch = np.zeros(self.NKERN, np.int32)
        ms = np.zeros(3 * self.NKERN)
        _chk(lib().dns_autotune(self.h, iters, ch.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(ms)))
        return ch, ms.reshape(self.NKERN, 3)

    HIST_KEYS = ("flag_ctpress", "iter", "t", "dtv", "dtc", "dt", "err", "Qx", "QxT", "mpgx", "dgx", "Umax",
                 "Uslip")

    def history_record(self, member=0):
        """record_out's history record (FOU3D.f90:2169-2200) at the current state, as a dict."""
        r = np.zeros(13)
        _chk(lib().dns_history_record(self.h, int(member), _p(r)))
        d = dict(zip(self.HIST_KEYS, r))
        d["flag_ctpress"] = int(d["flag_ctpress"])
        d["iter"] = int(d["iter"])
        return d

    STATS = ("Um", "U2m", "Wm", "W2m", "UVm", "UWm", "Vm", "V2m", "VWm", "wxm", "wx2m", "Pm", "P2m")

    def accumulate_stats(self):
        """Adds the statistics and spectra of the current state (stats.f90, spectra.f90)."""
        _chk(lib().dns_accumulate_stats(self.h))

    def reset_stats(self):
        _chk(lib().dns_reset_stats(self.h))

    def refresh_stats_pressure(self):
        """ppPL = modes_to_planes_UVP(p) (FOU3D.f90:149-152): Pm, P2m sum this copy of the current p from
        now on, until the next call (the reference refreshes it at write steps)."""
        _chk(lib().dns_refresh_stats_pressure(self.h))

    def stats(self, member=0, spectra=False):
        """dict: istat, the 13 plane statistics (arrays of nyu+2; Pm, P2m at index j = 1..nyu+1) and, with
        spectra=True, spU, spV, spW, spUV, spP as [ncol, rows] (the reference's sp(j, column))."""
        nj = self.nyu + 2
        st = np.zeros(13 * nj)
        shapes = [(self.C, self.nyu + 2), (self.C, self.nyv + 2), (self.C, self.nyu + 2), (self.C, self.nyu + 2),
                  (self.C, self.nyu)]
        sp = np.zeros(sum(a * b for a, b in shapes)) if spectra else None
        n = ctypes.c_int()
        _chk(lib().dns_get_stats(self.h, int(member), _p(st), _p(sp) if spectra else None, ctypes.byref(n)))
        out = {"istat": n.value}
        for q, name in enumerate(self.STATS):
            out[name] = st[q * nj:(q + 1) * nj].copy()
        if spectra:
            off = 0
            for name, shp in zip(("spU", "spV", "spW", "spUV", "spP"), shapes):
                out[name] = sp[off:off + shp[0] * shp[1]].reshape(shp).copy()
                off += shp[0] * shp[1]
        return out

    def set_column_path(self, path):
        """0: one-pass column kernel where it applies, 1: three sweeps."""
        _chk(lib().dns_set_column_path(self.h, int(path)))

    def column_path(self):
        """Column path in use: 0 one pass, 1 three sweeps."""
        v = ctypes.c_int()
        _chk(lib().dns_get_column_path(self.h, ctypes.byref(v)))
        return v.value

    # ---- device-pointer utilities (the caller owns the device memory: any HIP allocation)
    def plane_fft_roundtrip(self, d_in, d_phys, d_back, n):
        _chk(lib().dns_plane_fft_roundtrip(self.h, d_in, d_phys, d_back, n))

    def modes_to_planes(self, d_cols, nrows, d_planes):
        _chk(lib().dns_modes_to_planes(self.h, d_cols, nrows, d_planes))

    def planes_to_modes(self, d_planes, nrows, d_cols):
        _chk(lib().dns_planes_to_modes(self.h, d_planes, nrows, d_cols))


class Ensemble:
    """Several DNS handles (one per size, each with its members) advanced together: one grouped launch per
    kernel class per RK3 substep for every member, or one persistent launch per substep (set_launch)
    (dns_ensemble_*).  The handles stay usable for uploads, downloads and scalars, and must stay open while
    the ensemble is."""

    def __init__(self, handles):
        self.handles = list(handles)
        arr = (ctypes.c_void_p * len(self.handles))(*[h.h.value for h in self.handles])
        e = ctypes.c_void_p()
        _chk(lib().dns_ensemble_create(arr, len(self.handles), ctypes.byref(e)))
        self.e = e

    def step(self, n=1): _chk(lib().dns_ensemble_step(self.e, n))
    def step_graph(self, n=1): _chk(lib().dns_ensemble_step_graph(self.e, n))
    def sync(self): _chk(lib().dns_ensemble_sync(self.e))
    def set_timing(self, on): _chk(lib().dns_ensemble_set_timing(self.e, 1 if on else 0))

    # task classes of the persistent launch, in queue order (units per task: dns_ensemble_set_launch)
    TASK_CLASSES = ("reset", "ghost", "zinv", "xpass_u", "xpass_v", "cfl_reduce", "zfwd_u", "zfwd_v",
                    "col_pass", "meanflow")

    def set_launch(self, mode, units=None):
        """mode 0: grouped launches (one per kernel class per substep); 1: one persistent launch per
        substep over a task queue.  units: {class name: units per task} (others default)."""
        arr = None
        if units:
            arr = (ctypes.c_int * len(self.TASK_CLASSES))(*[int(units.get(c, 0)) for c in self.TASK_CLASSES])
        _chk(lib().dns_ensemble_set_launch(self.e, int(mode), arr))

    def launch(self):
        """(mode, workgroups of a persistent launch, tasks of substeps 1..3)"""
        mode, nwg, nt = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 3)()
        _chk(lib().dns_ensemble_get_launch(self.e, ctypes.byref(mode), ctypes.byref(nwg), nt))
        return mode.value, nwg.value, list(nt)

    # kernel classes of the grouped launches with tile shapes (dns_ensemble_set_tiles); the plane classes'
    # shapes (threads per workgroup, variant) in the order of the autotuner's times
    TILE_CLASSES = ("zinv", "xpass_u", "xpass_v", "zfwd_u", "zfwd_v", "col_pass")
    PLANE_SHAPES = ((128, 0), (128, 1), (256, 0), (256, 1), (256, 2))
    TUNE_MS = 28

    def tiles(self):
        """(var, wg): var[6] variant of each grouped class (TILE_CLASSES), wg[5] threads per workgroup of
        the five plane classes."""
        var, wg = (ctypes.c_int * 6)(), (ctypes.c_int * 5)()
        _chk(lib().dns_ensemble_get_tiles(self.e, var, wg))
        return list(var), list(wg)

    def set_tiles(self, var, wg):
        _chk(lib().dns_ensemble_set_tiles(self.e, (ctypes.c_int * 6)(*[int(v) for v in var]),
                                          (ctypes.c_int * 5)(*[int(w) for w in wg])))

    def autotune(self, iters=3):
        """Times every tile shape of every grouped class on the whole launch and keeps the fastest; returns
        (var, wg, ms) with ms[class][k], k over PLANE_SHAPES for the plane classes and over the three
        variants for the column stage (-1: not available)."""
        ms = np.full(self.TUNE_MS, -1.0)
        _chk(lib().dns_ensemble_autotune(self.e, int(iters), None, None, _p(ms)))
        var, wg = self.tiles()
        return var, wg, [ms[5 * c:5 * c + 5].tolist() for c in range(5)] + [ms[25:28].tolist()]

    def timing(self):
        t = np.zeros(7)
        n = ctypes.c_int()
        _chk(lib().dns_ensemble_get_timing(self.e, _p(t), ctypes.byref(n)))
        return dict(zinv=t[0], xpass=t[1], zfwd=t[2], col_pass=t[3], substep=t[6], n=n.value)

    def close(self):
        if getattr(self, "e", None):
            lib().dns_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
