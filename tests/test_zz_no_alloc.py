"""The no-allocation preprocess entry point against the workspace one.

gaffa_hip_detrend_normalise_batch_noalloc keeps its scratch in a
library-owned device array; it must give exactly what
gaffa_hip_detrend_normalise_batch gives with a caller workspace, on two calls
in a row (the second smaller, so stale scratch from the first would show),
and reject batches beyond its scratch before launching anything.
"""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class PreprocessArgs(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("nseries", ctypes.c_uint64),
                ("nsamples", ctypes.c_uint64), ("stride", ctypes.c_uint64),
                ("input_is_u32", ctypes.c_int32), ("detrend", ctypes.c_int32),
                ("scrunch_factor", ctypes.c_uint64), ("median_window", ctypes.c_uint64),
                ("normalise", ctypes.c_int32), ("reject_constant", ctypes.c_int32),
                ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_uint64),
                ("status", ctypes.c_void_p)]


def _lib():
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    lib.gaffa_hip_last_error.restype = ctypes.c_char_p
    lib.gaffa_hip_preprocess_workspace_bytes.restype = ctypes.c_uint64
    lib.gaffa_hip_preprocess_workspace_bytes.argtypes = [ctypes.c_uint64] * 3
    lib.gaffa_hip_noalloc_scratch_bytes.restype = ctypes.c_uint64
    return lib


def _run(lib, plane, *, factor, window, noalloc):
    """One preprocess call on a device copy of `plane` (uint32 sums); returns
    the float series (as raw bits) and the per-series status."""
    nseries, n = plane.shape
    ptrs = []

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert lib.gaffa_hip_malloc(ctypes.byref(p), ctypes.c_uint64(nbytes)) == 0
        ptrs.append(p)
        return p

    try:
        d = dmalloc(plane.nbytes)
        st = dmalloc(4 * nseries)
        assert lib.gaffa_hip_memcpy(d, plane.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_uint64(plane.nbytes), None) == 0
        assert lib.gaffa_hip_memset(st, 0, ctypes.c_uint64(4 * nseries), None) == 0
        ws, ws_bytes = None, 0
        if not noalloc:
            ws_bytes = lib.gaffa_hip_preprocess_workspace_bytes(nseries, n, factor)
            ws = dmalloc(ws_bytes)
        a = PreprocessArgs(d, nseries, n, n, 1, 1, factor, window, 1, 1, ws, ws_bytes, st)
        fn = lib.gaffa_hip_detrend_normalise_batch_noalloc if noalloc else \
            lib.gaffa_hip_detrend_normalise_batch
        rc = fn(ctypes.byref(a), None)
        assert rc == 0, lib.gaffa_hip_last_error()
        assert lib.gaffa_hip_stream_synchronize(None) == 0
        out = np.empty((nseries, n), np.uint32)
        status = np.empty(nseries, np.int32)
        assert lib.gaffa_hip_memcpy(out.ctypes.data_as(ctypes.c_void_p), d,
                                    ctypes.c_uint64(out.nbytes), None) == 0
        assert lib.gaffa_hip_memcpy(status.ctypes.data_as(ctypes.c_void_p), st,
                                    ctypes.c_uint64(status.nbytes), None) == 0
        return out, status
    finally:
        for p in ptrs:
            lib.gaffa_hip_free(p)


def test_noalloc_preprocess_equals_workspace_entry_point(gpu):
    lib = _lib()
    rng = np.random.default_rng(7)
    # two calls in a row, the second smaller; the third series of the first
    # batch is constant (rejected with a status, data left detrended)
    for nseries, n, factor, window in ((20, 100_000, 31, 101), (7, 41_000, 13, 51)):
        print(f"no-alloc preprocess: {nseries} series x {n} samples, scrunch {factor}, "
              f"window {window}", flush=True)
        sys.stdout.flush()
        plane = np.rint(rng.normal(131072.0, 512.0, size=(nseries, n))).astype(np.uint32)
        if nseries == 20:
            plane[2] = 5000
        want, want_status = _run(lib, plane, factor=factor, window=window, noalloc=False)
        got, got_status = _run(lib, plane, factor=factor, window=window, noalloc=True)
        np.testing.assert_array_equal(got_status, want_status)
        np.testing.assert_array_equal(got, want)  # float bits
        assert np.any(want_status != 0) == (nseries == 20)


def test_noalloc_preprocess_rejects_oversize_and_workspace(gpu):
    lib = _lib()
    cap = lib.gaffa_hip_noalloc_scratch_bytes()
    assert cap == 64 << 20
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    big = PreprocessArgs(dummy, 4096, 4_137_020, 4_137_020, 1, 1, 309, 101, 1, 1, None, 0,
                         dummy)
    assert lib.gaffa_hip_preprocess_workspace_bytes(4096, 4_137_020, 309) > cap
    assert lib.gaffa_hip_detrend_normalise_batch_noalloc(ctypes.byref(big), None) == 2
    assert b"split the batch" in lib.gaffa_hip_last_error()
    with_ws = PreprocessArgs(dummy, 4, 1000, 1000, 1, 1, 3, 101, 1, 1, dummy, 1 << 20, dummy)
    assert lib.gaffa_hip_detrend_normalise_batch_noalloc(ctypes.byref(with_ws), None) == 1
    assert b"workspace must be null" in lib.gaffa_hip_last_error()
