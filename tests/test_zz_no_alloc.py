"""The no-allocation key sort (mhpc_sort_keys_noalloc) against the context's key sort
(mhpc_sort), through the C-ABI: two calls in a row, the second smaller and full of equal keys,
so that the static scratch's leftovers from the first call are exercised.  The caller's arrays
come from mhpc_device_alloc; the sort runs on the null stream."""
import ctypes as C

import numpy as np
import pytest

from miluphpc_amd import ics
from conftest import print_case

pytestmark = pytest.mark.gpu


def test_sort_keys_noalloc_matches_mhpc_sort(gpu_lib):
    lib = gpu_lib
    rng = np.random.default_rng(77)
    big = lib.NOALLOC_SORT_MAX
    cases = [("plummer", ics.plummer(big, seed=7))]
    # second call: fewer keys, many of them equal (a coarse lattice repeated)
    m = 300_001
    g = rng.integers(0, 8, (m, 3)).astype(np.float64) / 8.0 - 0.4
    cases.append(("ties", {"x": g[:, 0].copy(), "y": g[:, 1].copy(), "z": g[:, 2].copy(),
                           "mass": np.full(m, 1.0 / m)}))
    with lib.Context(big, lib.default_params(curve=lib.CURVE_HILBERT)) as ctx:
        dk = ctx.device_alloc(8 * big)
        dsk = ctx.device_alloc(8 * big)
        dsv = ctx.device_alloc(4 * big)
        marker = np.full(big, -1, dtype=np.int64)
        for name, s in cases:
            n = len(s["x"])
            print_case(f"no-alloc key sort, {name}, n={n}")
            # the context's own sort of the same keys
            ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
            ctx.bbox()
            ctx.keys()
            ctx.sort()
            key, perm = ctx.download("key"), ctx.download("sorted")
            ctx.device_upload(dk, key)
            ctx.device_upload(dsk, marker)
            ctx.device_upload(dsv, marker.astype(np.int32))
            st = lib.LIB.mhpc_sort_keys_noalloc(C.c_void_p(dk), n, C.c_void_p(dsk), C.c_void_p(dsv), None)
            assert st == lib.OK, lib.LIB.mhpc_status_string(st)
            got_keys = ctx.device_download(dsk, big, np.uint64)
            got_perm = ctx.device_download(dsv, big, np.int32)
            assert np.array_equal(got_perm[:n], perm)
            assert np.array_equal(got_keys[:n], key[perm])
            assert np.array_equal(perm, np.argsort(key, kind="stable"))
            # nothing written past n
            assert np.all(got_perm[n:] == -1) and np.all(got_keys[n:].view(np.int64) == -1)
        # beyond the scratch: refused before anything is queued
        st = lib.LIB.mhpc_sort_keys_noalloc(C.c_void_p(dk), big + 1, C.c_void_p(dsk), C.c_void_p(dsv), None)
        assert st == lib.E_SCRATCH
        assert lib.LIB.mhpc_sort_keys_noalloc(None, 5, None, None, None) == lib.E_ARG
        assert lib.LIB.mhpc_sort_keys_noalloc(None, 0, None, None, None) == lib.OK
        for p in (dk, dsk, dsv):
            ctx.device_free(p)
