"""The allocation-free entry point dns_four_to_phys_noalloc through the C-ABI: device input and output, the
library's static workspace instead of device allocations.  Its output equals dns_four_to_phys's on the
same planes bit for bit, on two calls in a row (the second with fewer, smaller planes, so the workspace
holds the first call's contents when the second starts), and sizes whose plane does not fit the
workspace or pointers that are not 16-byte aligned are rejected with an error."""
import numpy as np
import pytest

import dnship
from hipmem import DeviceBuffer

pytestmark = pytest.mark.gpu


def spectral_planes(n, Nx, Nz, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, Nz, Nx + 2))


@pytest.mark.parametrize("cases", [
    [(40, 384, 384, 576, 576), (3, 384, 384, 384, 384)],    # 40 planes take several workspace batches
    [(9, 768, 384, 1152, 576), (2, 64, 64, 96, 96)],
])
def test_noalloc_equals_allocating_entry(cases):
    for call, (n, Nx, Nz, Mx, Mz) in enumerate(cases):
        print("case: noalloc call %d, %d planes %dx%d -> %dx%d" % (call, n, Nx, Nz, Mx, Mz), flush=True)
        pl = spectral_planes(n, Nx, Nz, seed=call + 3)
        want, _ = dnship.four_to_phys(pl, Mx, Mz)
        d_in = DeviceBuffer.from_array(pl)
        d_out = DeviceBuffer(n * Mz * (Mx + 2) * 8)
        dnship.four_to_phys_noalloc(d_in.ptr, d_out.ptr, n, Nx, Nz, Mx, Mz)
        got = d_out.to_array((n, Mz, Mx + 2), np.float64)
        d_in.free()
        d_out.free()
        assert np.array_equal(got, want), call


def test_noalloc_rejects_what_does_not_fit():
    d = DeviceBuffer(1 << 20)
    with pytest.raises(dnship.DnsError):
        dnship.four_to_phys_noalloc(d.ptr + 8, d.ptr, 1, 16, 16, 24, 24)       # misaligned input
    with pytest.raises(dnship.DnsError):
        dnship.four_to_phys_noalloc(d.ptr, d.ptr, 1, 16, 16, 40, 24)           # Mx/2 = 20 not instantiated
    with pytest.raises(dnship.DnsError):
        dnship.four_to_phys_noalloc(d.ptr, d.ptr, 1, 2304, 1152, 2304, 1152)   # one plane > the workspace
    d.free()
