"""pvlbm_sample_kinds_noalloc, the entry point that allocates nothing, against pvlbm_sample_kinds.

Both go through the C ABI on one handle.  Two calls in a row, the second with a smaller sample
grid, must each give exactly the bytes of the allocating entry point (the second call must not see
anything the first left in the shared scratch array); a request larger than the scratch array is
rejected with PVLBM_EINVAL and leaves the handle usable.
"""
import ctypes as C

import numpy as np
import pytest

from pv_dec_original_amd import _native as N
from tests import gpu_cases as G

pytestmark = pytest.mark.gpu


def _call(fn, h, ix, iy, iz):
    ax = [np.ascontiguousarray(a, dtype=np.int32) for a in (ix, iy, iz)]
    p = [a.ctypes.data_as(C.POINTER(C.c_int32)) for a in ax]
    out = np.full(len(ax[0]) * len(ax[1]) * len(ax[2]), 0xEE, dtype=np.uint8)
    rc = fn(h, p[0], len(ax[0]), p[1], len(ax[1]), p[2], len(ax[2]), N.u8ptr(out))
    return rc, out.reshape(len(ax[2]), len(ax[1]), len(ax[0]))


def test_noalloc_sample_kinds_matches_the_allocating_entry_point():
    case = G.small_ybif_case(q=19)
    sd, sim, _ = G.build(case, normals=False, threads=1)
    L = N.lib()
    h = C.c_void_p(sd.handle)
    nx, ny, nz = case.geometry.dims
    rng = np.random.default_rng(17)
    kinds = sd.kinds()
    calls = [(rng.integers(0, nx, 96), rng.integers(0, ny, 48), rng.integers(0, nz, 40)),
             (rng.integers(0, nx, 9), rng.integers(0, ny, 5), rng.integers(0, nz, 3))]
    for i, (ix, iy, iz) in enumerate(calls):
        print(f"no-alloc call {i}: {len(ix)} x {len(iy)} x {len(iz)} samples", flush=True)
        rc_a, ref = _call(L.pvlbm_sample_kinds, h, ix, iy, iz)
        rc_n, got = _call(L.pvlbm_sample_kinds_noalloc, h, ix, iy, iz)
        assert rc_a == 0 and rc_n == 0, N.lib().pvlbm_last_error(h)
        assert np.array_equal(got, ref)
        assert np.array_equal(ref, kinds[np.ix_(iz, iy, ix)])
        assert len(np.unique(ref)) > 1
    # too large for the scratch array: rejected, and the handle still works
    big = np.zeros(256, dtype=np.int32)
    rc, _ = _call(L.pvlbm_sample_kinds_noalloc, h, big, big, big)
    assert rc == 1 and b"scratch" in L.pvlbm_last_error(h)
    rc, got = _call(L.pvlbm_sample_kinds_noalloc, h, *calls[1])
    assert rc == 0 and np.array_equal(got, kinds[np.ix_(calls[1][2], calls[1][1], calls[1][0])])
