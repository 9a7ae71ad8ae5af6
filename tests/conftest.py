import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "final-single-block-_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


def stats_errors(got, ref, npts_pad, npts_un):
    """Largest error of each statistic relative to its scale; a mean against sqrt(points x largest
    second moment), the size of its rounding in a physical-space sum."""
    means = {"Um": ("U2m", npts_pad), "Wm": ("W2m", npts_pad), "Vm": ("V2m", npts_pad),
             "wxm": ("wx2m", npts_un), "Pm": ("P2m", 1)}
    err = {}
    for name in ("Um", "U2m", "Wm", "W2m", "UVm", "UWm", "Vm", "V2m", "VWm", "wxm", "wx2m", "Pm", "P2m"):
        r = np.asarray(ref[name])
        g = np.asarray(got[name])
        if len(g) == len(r) + 1:
            g = g[1:]
        scale = np.abs(r).max()
        if name in means:
            sq, npts = means[name]
            scale = max(scale, np.sqrt(npts * np.abs(ref[sq]).max()))
        err[name] = np.abs(g - r).max() / max(scale, 1e-300)
    return err


def ftp_compare(got: bytes, ref: bytes, tol: float):
    """Two four_to_phys output files record by record: the header records and every plane record's
    (jd, id, nx, nz, y) byte for byte, the plane values within tol of the plane's largest value."""
    import struct

    def recs(buf):
        off, out = 0, []
        while off < len(buf):
            (n,) = struct.unpack_from("<i", buf, off)
            out.append(buf[off + 4:off + 4 + n])
            off += n + 8
        return out
    a, b = recs(got), recs(ref)
    assert len(a) == len(b), (len(a), len(b))
    worst = 0.0
    for k, (x, y) in enumerate(zip(a, b)):
        assert len(x) == len(y), k
        if k < 3:
            assert x == y, "header record %d" % k
            continue
        assert x[:24] == y[:24], (k, struct.unpack_from("<4id", x), struct.unpack_from("<4id", y))
        u, v = np.frombuffer(x, "<f8", offset=24), np.frombuffer(y, "<f8", offset=24)
        worst = max(worst, float(np.abs(u - v).max() / max(np.abs(v).max(), 1e-300)))
    assert worst <= tol, worst
    return worst

