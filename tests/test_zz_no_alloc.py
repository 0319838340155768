"""GPU test of tdp_filter_noalloc, the filter entry point that allocates no
device memory: called through the C ABI, its device-resident rows and count
equal tdp_filter's on two calls in a row (the second one smaller, so stale
scratch contents from the first would show), then on a longer cycle with the
null stream, plus invalid arguments.

The test allocates the device buffers itself through the HIP runtime the
library is linked against (the entry point only takes them)."""
import ctypes

import numpy as np
import pytest

import tdp1375_amd as T
from tdp1375_amd import _lib
from tdp1375_amd.workloads import FILTER_ORIENTED_5_CYCLE

pytestmark = pytest.mark.gpu

CYCLE = FILTER_ORIENTED_5_CYCLE["cycle"]
TRIPLE = FILTER_ORIENTED_5_CYCLE["triple"]
HIP_MEMCPY_D2H = 2


def _hip():
    # the runtime libtdp1375_jni.so itself links (already loaded)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


class _DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(1, nbytes)) == 0
        self.nbytes = nbytes

    def read(self, nbytes):
        out = np.zeros(max(1, nbytes), np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, nbytes, HIP_MEMCPY_D2H) == 0
        return out[:nbytes]

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(lib, hip, stream, cycle, triple, n, skip, cap):
    rows = _DeviceBuffer(hip, cap * n)
    total = _DeviceBuffer(hip, 8)
    try:
        st = lib.tdp_filter_noalloc(_lib.u8(cycle), len(cycle), _lib.u8(triple), len(triple), n, int(skip),
                                    rows.ptr, cap, total.ptr, stream)
        _lib.check(st)
        assert hip.hipStreamSynchronize(stream) == 0
        count = int(total.read(8).view(np.uint64)[0])
        got = rows.read(min(count, cap) * n).reshape(-1, n)
        return got, count
    finally:
        rows.free()
        total.free()


@pytest.mark.parametrize("skip", [False, True])
def test_noalloc_filter_matches_tdp_filter(gpu_lib, skip):
    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # n = 13: 12! ranks in eight 2^26-rank passes, capped list; then
        # n = 9 on the same scratch array
        for n, cap in ((13, 1 << 20), (9, 5000)):
            want_rows, want_total = T.filter_pi_permutations(CYCLE, 5, TRIPLE, 3, n, skip, max_rows=cap)
            got_rows, got_total = _noalloc(gpu_lib, hip, stream, CYCLE, TRIPLE, n, skip, cap)
            assert got_total == want_total
            assert np.array_equal(got_rows, want_rows)
        # a longer cycle (more 2-move candidates) after the short one, null stream
        cyc, tri = [0, 5, 2, 7, 4, 1, 6], [1, 3, 8]
        want_rows, want_total = T.filter_pi_permutations(cyc, len(cyc), tri, 3, 10, skip, max_rows=1 << 16)
        got_rows, got_total = _noalloc(gpu_lib, hip, None, cyc, tri, 10, skip, 1 << 16)
        assert got_total == want_total
        assert np.array_equal(got_rows, want_rows)
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_filter_rejects_bad_arguments(gpu_lib):
    hip = _hip()
    total = _DeviceBuffer(hip, 8)
    try:
        st = gpu_lib.tdp_filter_noalloc(_lib.u8(CYCLE), 5, _lib.u8(TRIPLE), 3, 22, 0, None, 0, total.ptr, None)
        assert st == _lib.TDP_ERR_INVALID_ARGUMENT
        assert b"n too large" in gpu_lib.tdp_last_error()
        st = gpu_lib.tdp_filter_noalloc(_lib.u8(CYCLE), 5, _lib.u8(TRIPLE), 3, 9, 0, None, 0, None, None)
        assert st == _lib.TDP_ERR_INVALID_ARGUMENT
    finally:
        total.free()
