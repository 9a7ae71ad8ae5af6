"""ORACLE TEST INFRASTRUCTURE -- imported only by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py.  Never imported by the product.

* ``Oracle``: ctypes wrapper of liboracle.so (dns_oracle.c, the C restatement of the reference's
  RK3 step).
* ``run_reference``: runs the reference's own Fortran hot-path code (oracle/_ref/ref_harness, built by
  ``make -C oracle ref`` from the reference sources) on one state and returns its outputs.

Column fields are complex128 arrays [ncol, nrows] in the reference's x(j, column) order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [ctypes.c_int] * 2
        for name in ("orc_destroy", "orc_meanflow_ctP"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("orc_rhs0", "orc_nonlinear", "orc_solveU", "orc_meanflow_ctU", "orc_solveP",
                     "orc_vcorr", "orc_set_threads"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_nu.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_get_nu.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_get_du.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_get_flux.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_get_psi.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_get_scalars.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_fixed_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_get_setup.argtypes = [ctypes.c_void_p] + [_dp] * 8
        L.orc_plane_fft.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


class Oracle:
    def __init__(self, Mx, Mz, ny, Re=3250.0, alp=1.0, bet=2.0, CFL=0.7, mpgx=-3.0675e-3, dyq=10.0,
                 ppp=5, flag_ctpress=1, threads=None, fixed_dt=0.0):
        L = lib()
        self.h = L.orc_create(Mx, Mz, ny, Re, alp, bet, CFL, mpgx, dyq, ppp, flag_ctpress)
        if not self.h:
            raise ValueError("unsupported sizes %d x %d" % (Mx, Mz))
        if fixed_dt > 0:
            L.orc_set_fixed_dt(self.h, fixed_dt)
        s = (ctypes.c_int * 8)()
        L.orc_sizes(self.h, s)
        self.Nx, self.Nz, self.Kx, self.C, self.nyu, self.nyv = s[0], s[1], s[2], s[3], s[4], s[5]
        self.Mx, self.Mz, self.ny = Mx, Mz, ny
        if threads:
            L.orc_set_threads(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def shapes(self):
        C, ny = self.C, self.ny
        return {"u1": (C, ny + 1), "u2": (C, ny), "u3": (C, ny + 1), "p": (C, ny - 1),
                "Nu1": (C, ny - 1), "Nu2": (C, ny - 2), "Nu3": (C, ny - 1)}

    def set_state(self, u1, u2, u3, p, t=0.0, it=0, QxT=0.0):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (u1, u2, u3, p)]
        lib().orc_set_state(self.h, *[_ptr(a.view(np.float64)) for a in arrs], t, it, QxT)

    def get_state(self):
        sh = self.shapes()
        out = [np.zeros(sh[k], np.complex128) for k in ("u1", "u2", "u3", "p")]
        lib().orc_get_state(self.h, *[_ptr(a.view(np.float64)) for a in out])
        return out

    def set_nu(self, Nu1, Nu2, Nu3):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (Nu1, Nu2, Nu3)]
        lib().orc_set_nu(self.h, *[_ptr(a.view(np.float64)) for a in arrs])

    def get_nu(self):
        sh = self.shapes()
        out = [np.zeros(sh[k], np.complex128) for k in ("Nu1", "Nu2", "Nu3")]
        lib().orc_get_nu(self.h, *[_ptr(a.view(np.float64)) for a in out])
        return out

    def get_du(self):
        sh = self.shapes()
        out = [np.zeros(sh[k], np.complex128) for k in ("Nu1", "Nu2", "Nu3")]
        lib().orc_get_du(self.h, *[_ptr(a.view(np.float64)) for a in out])
        return out

    def get_flux(self):
        C, ny = self.C, self.ny
        out = [np.zeros((C, ny), np.complex128), np.zeros((C, ny + 1), np.complex128),
               np.zeros((C, ny), np.complex128)]
        lib().orc_get_flux(self.h, *[_ptr(a.view(np.float64)) for a in out])
        return out

    def get_psi(self):
        out = [np.zeros((self.C, self.ny - 1), np.complex128) for _ in range(2)]
        lib().orc_get_psi(self.h, *[_ptr(a.view(np.float64)) for a in out])
        return out

    def scalars(self):
        sc = np.zeros(8)
        lib().orc_get_scalars(self.h, _ptr(sc))
        return dict(zip(["t", "dt", "dti", "mpgx", "Qx", "dgx", "dtc", "QxT"], sc))

    def set_dt(self, dt):
        lib().orc_set_dt(self.h, dt)

    def setup(self):
        ny, C = self.ny, self.C
        yu, yv = np.zeros(ny + 1), np.zeros(ny)
        du, dv = np.zeros(ny + 1), np.zeros(ny)
        d2u, d2v = np.zeros((ny + 1, 3)), np.zeros((ny, 3))
        DG = np.zeros((C, ny - 1, 3))
        w = np.zeros(6)
        lib().orc_get_setup(self.h, *[_ptr(a) for a in (yu, yv, du, dv, d2u, d2v, DG, w)])
        return dict(yu=yu, yv=yv, dthdyu=du, dthdyv=dv, dyu2i=d2u, dyv2i=d2v, DG=DG, dtv=w[0],
                    gw=w[1:3], gbc=w[3], gi=w[4:6])

    def rhs0(self, k): lib().orc_rhs0(self.h, k)
    def nonlinear(self, k): lib().orc_nonlinear(self.h, k)
    def solveU(self, k): lib().orc_solveU(self.h, k)
    def meanflow_ctU(self, k): lib().orc_meanflow_ctU(self.h, k)
    def solveP(self, k): lib().orc_solveP(self.h, k)
    def vcorr(self, k): lib().orc_vcorr(self.h, k)

    def step(self, n=1):
        hist = np.zeros((n, 4))
        lib().orc_step(self.h, n, _ptr(hist))
        return hist


def plane_fft(Mx, Mz, planes):
    """four_to_phys_du then phys_to_four_du on real (Mx+2) x Mz planes ([n, Mz, Mx+2] float64)."""
    planes = np.ascontiguousarray(planes, dtype=np.float64)
    n = planes.shape[0]
    out = np.zeros((2 * n, Mz, Mx + 2))
    lib().orc_plane_fft(Mx, Mz, n, _ptr(planes), _ptr(out))
    return out[:n], out[n:]


# ------------------------------------------------------------------------------------------------
def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


def time_reference_substeps(Mx, Mz, ny, u1, u2, u3, p, nsub=1, timeout=1800, **kw):
    """Wall time (s) of ``nsub`` substeps of the reference's own hot path (mode 4 of the harness,
    MPI_Wtime around the loop only; set-up and file I/O are outside the timed region)."""
    return run_reference(Mx, Mz, ny, u1, u2, u3, p, mode=4, nsteps=nsub, timeout=timeout, **kw)


# plane statistics of stats.f90 in the order the harness writes them (mode 5): u-grid planes 0..nyu+1
# (U, W, UV, UW), v-grid planes 0..nyu+1 (V, VW, omega_x), p-grid planes 1..nyu+1 (P)
STATS_NAMES = ("Um", "U2m", "Wm", "W2m", "UVm", "UWm", "Vm", "V2m", "VWm", "wxm", "wx2m", "Pm", "P2m")


def run_reference(Mx, Mz, ny, u1, u2, u3, p, mode=1, nsteps=1, kRK=1, dt=0.0, t=0.0, it=0,
                  Re=3250.0, alp=1.0, bet=2.0, CFL=0.7, mpgx=-3.0675e-3, dyq=10.0, ppp=5,
                  flag_ctpress=1, QxT=0.0, Nu=None, fft_planes=None, timeout=3600):
    """Run the reference's compiled hot path (oracle/_ref/ref_harness) on one state.

    mode 1: ``nsteps`` full RK3 steps -> dict(u1,u2,u3,p,Nu1,Nu2,Nu3,t,dt,mpgx,Qx,dgx,hist)
    mode 2: one substep ``kRK`` with the stage-by-stage arrays (``dt`` is used when kRK > 1)
    mode 3: ``fft_planes`` [n, Mz, Mx+2] through four_to_phys_du / phys_to_four_du
    mode 5: statistics and spectra of the state (stats.f90, spectra.f90) -> dict of STATS_NAMES arrays,
            spU, spV, spW, spUV, spP [ncol, rows] and istat
    """
    if not have_reference():
        raise FileNotFoundError(REF_HARNESS)
    nsteps_field = nsteps if mode != 3 else fft_planes.shape[0]
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "params.txt"), "w") as f:
            f.write("%d\n%d %d %d\n" % (mode, Mx, Mz, ny))
            f.write("%.17e %.17e %.17e %.17e %.17e\n" % (Re, alp, bet, CFL, mpgx))
            f.write("%.17e %d %d\n" % (dyq, ppp, flag_ctpress))
            f.write("%.17e %d %d %d %.17e %.17e\n" % (t, it, nsteps_field, kRK, dt, QxT))
        if mode == 3:
            np.ascontiguousarray(fft_planes, dtype=np.float64).tofile(os.path.join(d, "fft_in.bin"))
        else:
            with open(os.path.join(d, "state_in.bin"), "wb") as f:
                for a in (u1, u2, u3, p):
                    f.write(np.ascontiguousarray(a, dtype=np.complex128).tobytes())
                if mode == 2:
                    for a in Nu:
                        f.write(np.ascontiguousarray(a, dtype=np.complex128).tobytes())
        env = dict(os.environ)
        env.setdefault("OMP_NUM_THREADS", "1")
        subprocess.run([REF_HARNESS], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, timeout=timeout, env=env)
        if mode == 4:
            with open(os.path.join(d, "timing.txt")) as f:
                return float(f.read().split()[1])
        if mode == 5:
            Nx, Nz = (2 * Mx) // 3, (2 * Mz) // 3
            C = (Nx // 2 + 1) * Nz
            nyu = ny - 1
            raw = np.fromfile(os.path.join(d, "stats_out.bin"), dtype=np.uint8)
            res = {"istat": int(raw[:4].view(np.int32)[0])}
            vals = raw[4:].view(np.float64)
            off = 0
            for name in STATS_NAMES:
                n = nyu + 1 if name in ("Pm", "P2m") else nyu + 2
                res[name] = vals[off:off + n].copy()
                off += n
            for name, nr in (("spU", nyu + 2), ("spV", nyu + 1), ("spW", nyu + 2), ("spUV", nyu + 2),
                             ("spP", nyu)):
                res[name] = vals[off:off + C * nr].reshape(C, nr).copy()
                off += C * nr
            return res
        if mode == 3:
            out = np.fromfile(os.path.join(d, "fft_out.bin"), dtype=np.float64)
            n = fft_planes.shape[0]
            out = out.reshape(2 * n, Mz, Mx + 2)
            return out[:n], out[n:]
        Nx, Nz = (2 * Mx) // 3, (2 * Mz) // 3
        C = (Nx // 2 + 1) * Nz
        nyv = ny - 2
        nyu = nyv + 1
        sizes = [("u1", nyu + 2), ("u2", nyv + 2), ("u3", nyu + 2), ("p", nyu), ("Nu1", nyu),
                 ("Nu2", nyv), ("Nu3", nyu)]
        raw = np.fromfile(os.path.join(d, "state_out.bin"), dtype=np.uint8)
        res = {}
        off = 0
        for name, nr in sizes:
            nbytes = C * nr * 16
            res[name] = raw[off:off + nbytes].view(np.complex128).reshape(C, nr).copy()
            off += nbytes
        sc = raw[off:off + 40].view(np.float64)
        res.update(dict(t=sc[0], dt=sc[1], mpgx=sc[2], Qx=sc[3], dgx=sc[4]))
        if mode == 1:
            res["hist"] = np.loadtxt(os.path.join(d, "hist.txt"), ndmin=2)
        if mode == 2:
            st = np.fromfile(os.path.join(d, "stages.bin"), dtype=np.uint8)
            off = 0

            def take(nr):
                nonlocal off
                nbytes = C * nr * 16
                a = st[off:off + nbytes].view(np.complex128).reshape(C, nr).copy()
                off += nbytes
                return a
            res["rhs0"] = [take(nyu), take(nyv), take(nyu)]
            res["Nu_new"] = [take(nyu), take(nyv), take(nyu)]
            res["du_nl"] = [take(nyu), take(nyv), take(nyu)]
            res["flux"] = [take(nyv + 2), take(nyu + 2), take(nyv + 2)]
            res["u_solved"] = [take(nyu + 2), take(nyv + 2), take(nyu + 2)]
            res["u1_meanflow"] = take(nyu + 2)
            res["p_new"], res["psi"] = take(nyu), take(nyu)
            res["u_corr"] = [take(nyu + 2), take(nyv + 2), take(nyu + 2)]
            res["div"] = take(nyu)
            sc2 = st[off:off + 40].view(np.float64)
            res["stage_scalars"] = dict(dt=sc2[0], t=sc2[1], mpgx=sc2[2], dgx=sc2[3], dtc=sc2[4])
        sb = np.fromfile(os.path.join(d, "setup_out.bin"), dtype=np.float64)
        o = 0
        setup = {}
        for name, n in (("yu", ny + 1), ("yv", ny), ("dthdyu", ny + 1), ("dthdyv", ny),
                        ("dyu2i", 3 * (ny + 1)), ("dyv2i", 3 * ny), ("DG", 3 * (ny - 1) * C)):
            setup[name] = sb[o:o + n].copy()
            o += n
        setup["dtv"] = sb[o]
        setup["gw"] = sb[o + 1:o + 3]
        setup["gi"] = sb[o + 3:o + 5]
        setup["gbc"] = sb[o + 5]
        res["setup"] = setup
        return res


# ------------------------------------------------------------------------------------------------
def plane_stats(Mx, Mz, ny, u1, u2, u3, p, yu, yv, dthdyv, bet):
    """One accumulation of stats.f90:1-97 (with omega_x of FOU3D.f90:307-338) from the spectral state,
    by the discrete sum rules of the reference's transforms: a physical-space plane sum of the 3/2
    grid (Mx x Mz points) is Mx Mz Re u(0, 0); a sum of products is Mx Mz Q(u, v) with
    Q = sum_{kx>=1} 2 Re(u v*) + sum_{kx=0} (Re(u_kz v_kz*) + Re(u_kz v_-kz)) / 2 (the real synthesis
    of rft_buff.f:76-84 ignores Im of the kx = 0 coefficient; kz = -Nz/2 has no partner on the padded
    grid).  omega_x is summed on the Nx x Nz grid of four_to_phys_N, where kx = Nx/2 also takes the
    half-terms and kz = -Nz/2 pairs with itself.  Planes the reference keeps in spectral layout (u
    planes 0 and nyu+1, v plane 0, p planes) sum the interleaved (Re, Im) entries.  Fields [C, rows]."""
    Nx, Nz = 2 * (Mx // 3), 2 * (Mz // 3)
    Kx = Nx // 2 + 1
    nyu, nyv = ny - 1, ny - 2
    C = Kx * Nz
    ii = np.arange(C) % Kx
    kk = np.arange(C) // Kx
    kz = np.where(kk < Nz // 2, kk, kk - Nz) * bet
    k0 = np.arange(Nz) * Kx                      # kx = 0 columns
    kn = k0 + Nx // 2                            # kx = Nx/2 columns
    pk = np.where(np.arange(Nz) == 0, 0, Nz - np.arange(Nz))
    pad_ok = np.arange(Nz) != Nz // 2

    def Qp(a, b):
        s = 2.0 * np.sum((a[ii >= 1] * np.conj(b[ii >= 1])).real)
        t = (a[k0] * np.conj(b[k0])).real + np.where(pad_ok, (a[k0] * b[pk * Kx]).real, 0.0)
        return Mx * Mz * (s + 0.5 * np.sum(t))

    def Qn(w):
        mid = (ii >= 1) & (ii < Nx // 2)
        s = 2.0 * np.sum(np.abs(w[mid]) ** 2)
        for cols in (k0, kn):
            s += 0.5 * np.sum(np.abs(w[cols]) ** 2 + (w[cols] * w[pk * Kx + (cols[0] % Kx)]).real)
        return Nx * Nz * s

    def ssum(a): return float(np.sum(a.real + a.imag))
    def sdot(a, b): return float(np.sum(a.real * b.real + a.imag * b.imag))

    gi1 = (yu[0] - yu[1]) / (yv[0] - yu[1])
    gi2 = (yu[nyu + 1] - yu[nyu]) / (yv[nyv + 1] - yu[nyu])
    u2i = np.zeros((C, nyu + 2), complex)          # interp_v (FOU3D.f90:396-418)
    for j in range(1, nyu + 1):
        u2i[:, j] = ((yu[j] - yv[j - 1]) * u2[:, j] + (yv[j] - yu[j]) * u2[:, j - 1]) / (yv[j] - yv[j - 1])
    u2i[:, 0] = gi1 * (u2[:, 0] - u2i[:, 1]) + u2i[:, 1]
    u2i[:, nyu + 1] = gi2 * (u2[:, nyv + 1] - u2i[:, nyu]) + u2i[:, nyu]
    u3i = np.zeros((C, nyv + 2), complex)          # interp_u (FOU3D.f90:373-392)
    for j in range(nyv + 2):
        u3i[:, j] = ((yv[j] - yu[j]) * u3[:, j + 1] + (yu[j + 1] - yv[j]) * u3[:, j]) / (yu[j + 1] - yu[j])
    out = {k: np.zeros(nyu + 2) for k in STATS_NAMES}
    N = Mx * Mz
    for j in range(nyu + 2):
        if 1 <= j <= nyu:
            out["Um"][j] = N * u1[0, j].real
            out["U2m"][j] = Qp(u1[:, j], u1[:, j])
            out["Wm"][j] = N * u3[0, j].real
            out["W2m"][j] = Qp(u3[:, j], u3[:, j])
            out["UVm"][j] = Qp(u1[:, j], u2i[:, j])
            out["UWm"][j] = Qp(u1[:, j], u3[:, j])
            out["Vm"][j] = N * u2[0, j].real
            out["V2m"][j] = Qp(u2[:, j], u2[:, j])
            out["VWm"][j] = Qp(u2[:, j], u3i[:, j])
            out["Pm"][j] = ssum(p[:, j - 1])
            out["P2m"][j] = sdot(p[:, j - 1], p[:, j - 1])
        else:
            out["Um"][j] = ssum(u1[:, j])
            out["U2m"][j] = sdot(u1[:, j], u1[:, j])
            out["Wm"][j] = ssum(u3[:, j])
            out["W2m"][j] = sdot(u3[:, j], u3[:, j])
            out["UVm"][j] = sdot(u1[:, j], u2i[:, j])
            out["UWm"][j] = sdot(u1[:, j], u3[:, j])
            if j == 0:
                out["Vm"][j] = ssum(u2[:, 0])
                out["V2m"][j] = sdot(u2[:, 0], u2[:, 0])
                out["VWm"][j] = sdot(u2[:, 0], u3i[:, 0])
        if j <= nyu:
            w = -(1j * kz) * u2[:, j] + (u3[:, j + 1] - u3[:, j]) * dthdyv[j]
            out["wxm"][j] = Nx * Nz * w[0].real
            out["wx2m"][j] = Qn(w)
    sp = {"spU": np.abs(u1) ** 2, "spV": np.abs(u2) ** 2, "spW": np.abs(u3) ** 2,
          "spUV": u1.real * u2i.real + u1.imag * u2i.imag, "spP": np.abs(p) ** 2}  # spectra.f90:29-60
    out.update(sp)
    return out


# ------------------------------------------------------------------------------------------------
def _frecords(buf):
    import struct
    off, out = 0, []
    while off < len(buf):
        (n,) = struct.unpack_from("<i", buf, off)
        out.append(buf[off + 4:off + 4 + n])
        off += n + 8
    return out


def _frec(payload):
    import struct
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def ftp_plane(b, Mx, Mz):
    """resize + cft + rft of four_to_phys.f90:76-79 (resize at :318-337, four_to_phys_N.f90's resize keeps
    the kz = -Nz/2 row, which is what Mz == Nz selects here) on one spectral plane b[z, x] (Nz x (Nx+2)):
    unnormalised inverse transforms, z then x; the real x-inverse takes Re of kx = 0 and of kx = Mx/2,
    and entries Mx, Mx+1 keep the kx = Mx/2 coefficient.  Returns the physical plane [Mz, Mx+2]."""
    Nz, nx = b.shape
    Nx = nx - 2
    KX = Nx // 2 + 1
    H = Mx // 2
    c = b[:, 0::2] + 1j * b[:, 1::2]
    full = np.zeros((Mz, KX), complex)
    full[:Nz // 2] = c[:Nz // 2]
    lo = Mz - Nz // 2 + (1 if Mz > Nz else 0)
    full[lo:] = c[lo + Nz - Mz:]
    zi = Mz * np.fft.ifft(full, axis=0)
    cx = np.zeros((Mz, H + 1), complex)
    cx[:, :KX] = zi
    out = np.zeros((Mz, Mx + 2))
    out[:, :Mx] = Mx * np.fft.irfft(cx, n=Mx, axis=1)
    out[:, Mx] = cx[:, H].real
    out[:, Mx + 1] = cx[:, H].imag
    return out


def four_to_phys_files(files, nvar=False):
    """four_to_phys.f90 (nvar: four_to_phys_N.f90) on the four field files {u1, u2, u3, p: bytes}: the
    output files' bytes, record for record (headers of u2, u3 and p from u1's; planes in the order of the
    band loops, :60-113 for u1, :116-186 for u2, :188-257 for u3, :260-305 for p)."""
    import struct
    out = {}
    r1 = _frecords(files["u1"])
    head = r1[0]
    nband = struct.unpack_from("<i", head, 40)[0]
    N = np.frombuffer(r1[1], "<i4").reshape(nband + 2, 4).T
    Ngal = N.copy()
    if not nvar:
        Ngal[0, 1:nband + 1] = N[0, 1:nband + 1] * 3 // 2
        Ngal[1, 1:nband + 1] = N[1, 1:nband + 1] * 3 // 2
    Nx, Nz = int(N[0, 1]), int(N[1, 1])
    Mx, Mz = int(Ngal[0, 1]), int(Ngal[1, 1])
    ylen_u = int(N[3, nband]) + 2 - int(N[3, 0])
    ryu = r1[2]
    yu = np.frombuffer(ryu, "<f8")[:ylen_u]
    dthetai_u1 = ryu[8 * ylen_u:8 * ylen_u + 8]
    r2 = _frecords(files["u2"])
    ylen_v = int(N[2, nband]) + 2 - int(N[2, 0])
    yv = np.frombuffer(r2[2], "<f8")[:ylen_v]
    dthetai_u2 = r2[2][8 * ylen_v:8 * ylen_v + 8]

    def planes(recs):
        for r in recs[3:]:
            jd, gid, nx, nz = struct.unpack_from("<4i", r, 0)
            (y,) = struct.unpack_from("<d", r, 16)
            yield jd, gid, y, np.frombuffer(r, "<f8", offset=24, count=(Nx + 2) * Nz).reshape(Nz, Nx + 2)

    def plane_rec(jd, gid, y, b):
        return _frec(struct.pack("<4id", jd, gid, Mx + 2, Mz, y) + ftp_plane(b, Mx, Mz).astype("<f8").tobytes())

    # u3 and p write u1's y record with the dthetai last read (u2's): the programs share one variable
    yrec_u3p = ryu[:8 * ylen_u] + dthetai_u2 + ryu[8 * ylen_u + 8:]
    for name, recs, yrec in (("u1", r1, ryu), ("u2", r2, r2[2]), ("u3", _frecords(files["u3"]), yrec_u3p)):
        it = planes(recs)
        row = 2 if name == "u2" else 3
        y = yv if name == "u2" else yu
        y0 = int(N[row, 0])
        o = [_frec(head), _frec(Ngal.T.astype("<i4").tobytes()), _frec(yrec)]
        last = None
        for iband in range(1, nband + 1):
            if iband == 1:
                last = next(it)
                o.append(plane_rec(*last))
            else:
                jj = int(N[row, iband - 1])
                if name == "u2":
                    o.append(plane_rec(last[0], last[1], y[jj - y0], last[3]))
                else:
                    o.append(plane_rec(jj, 1, y[jj - y0], last[3]))
            for _ in range(int(N[row, iband - 1]) + 1, int(N[row, iband]) + 1):
                last = next(it)
                o.append(plane_rec(*last))
            if iband == nband:
                last = next(it)
                o.append(plane_rec(*last))
        out[name] = b"".join(o)
    it = planes(_frecords(files["p"]))
    o = [_frec(head), _frec(Ngal.T.astype("<i4").tobytes()), _frec(yrec_u3p)]
    N4 = [int(v) for v in N[3]]
    for iband in range(1, nband + 1):
        if iband == 1:
            rng = range(N4[0] + 1, N4[1])
        elif iband == 3:
            rng = range(N4[2] + 2, N4[3] + 1) if nvar else range(N4[2] + 1, N4[3])
        else:
            rng = range(N4[1], N4[2] + 2) if nvar else range(N4[1], N4[2] + 1)
        for _ in rng:
            o.append(plane_rec(*next(it)))
    out["p"] = b"".join(o)
    return out
