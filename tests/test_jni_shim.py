"""The JNI natives of GPUSortingSearch driven through a fake JNIEnv.

The fake environment is a function table (JNI specification indices) whose
entries are ctypes callbacks keeping Java objects as Python values, so the
shim's marshalling and exception mapping (jni_sorter.cu:16-150) are checked
without a JVM.  GPU round trips are marked gpu.
"""
import ctypes

import numpy as np
import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import _lib
from tdp1375_amd import workloads as W

VP = ctypes.c_void_p
JINT = ctypes.c_int32


class FakeJNIEnv:
    IDX = dict(FindClass=6, ThrowNew=14, ExceptionClear=17, DeleteLocalRef=23, GetArrayLength=171,
               NewObjectArray=172, SetObjectArrayElement=174, NewByteArray=176, NewIntArray=179,
               GetByteArrayRegion=200, GetIntArrayRegion=203, SetByteArrayRegion=208, SetIntArrayRegion=211,
               SetFloatArrayRegion=213, ExceptionCheck=228)

    def __init__(self):
        self.objects = {}
        self.next_handle = 16
        self.exception = None
        self.live_refs = 0
        self._keep = []
        self.table = (VP * 235)()
        cb = self._install
        cb("FindClass", ctypes.CFUNCTYPE(VP, VP, ctypes.c_char_p), lambda env, name: self._new(("class", name.decode())))
        cb("ThrowNew", ctypes.CFUNCTYPE(JINT, VP, VP, ctypes.c_char_p), self._throw_new)
        cb("ExceptionClear", ctypes.CFUNCTYPE(None, VP), lambda env: setattr(self, "exception", None))
        cb("ExceptionCheck", ctypes.CFUNCTYPE(ctypes.c_uint8, VP), lambda env: int(self.exception is not None))
        cb("DeleteLocalRef", ctypes.CFUNCTYPE(None, VP, VP), self._delete)
        cb("GetArrayLength", ctypes.CFUNCTYPE(JINT, VP, VP), lambda env, a: len(self.objects[a][1]))
        cb("NewObjectArray", ctypes.CFUNCTYPE(VP, VP, JINT, VP, VP),
           lambda env, n, cls, init: self._new(("objarray", [None] * n)))
        cb("SetObjectArrayElement", ctypes.CFUNCTYPE(None, VP, VP, JINT, VP), self._set_elem)
        cb("NewByteArray", ctypes.CFUNCTYPE(VP, VP, JINT), lambda env, n: self._new(("bytes", [0] * n)))
        cb("NewIntArray", ctypes.CFUNCTYPE(VP, VP, JINT), lambda env, n: self._new(("ints", [0] * n)))
        cb("GetByteArrayRegion", ctypes.CFUNCTYPE(None, VP, VP, JINT, JINT, ctypes.POINTER(ctypes.c_int8)),
           self._get_bytes)
        cb("SetByteArrayRegion", ctypes.CFUNCTYPE(None, VP, VP, JINT, JINT, ctypes.POINTER(ctypes.c_int8)),
           self._set_region)
        cb("SetIntArrayRegion", ctypes.CFUNCTYPE(None, VP, VP, JINT, JINT, ctypes.POINTER(JINT)),
           self._set_region)
        cb("GetIntArrayRegion", ctypes.CFUNCTYPE(None, VP, VP, JINT, JINT, ctypes.POINTER(JINT)), self._get_ints)
        cb("SetFloatArrayRegion", ctypes.CFUNCTYPE(None, VP, VP, JINT, JINT, ctypes.POINTER(ctypes.c_float)),
           self._set_region)
        self.env_struct = (VP * 1)(ctypes.cast(self.table, VP))
        self.env = ctypes.cast(self.env_struct, VP)

    def _install(self, name, proto, fn):
        f = proto(fn)
        self._keep.append(f)
        self.table[self.IDX[name]] = ctypes.cast(f, VP)

    def _new(self, obj):
        h = self.next_handle
        self.next_handle += 8
        self.objects[h] = obj
        self.live_refs += 1
        return h

    def _delete(self, env, h):
        self.live_refs -= 1

    def _throw_new(self, env, cls, msg):
        self.exception = (self.objects[cls][1], msg.decode())
        return 0

    def _set_elem(self, env, arr, i, v):
        self.objects[arr][1][i] = v

    def _get_bytes(self, env, arr, start, n, buf):
        data = self.objects[arr][1]
        for q in range(n):
            buf[q] = ctypes.c_int8(data[start + q]).value

    def _get_ints(self, env, arr, start, n, buf):
        data = self.objects[arr][1]
        for q in range(n):
            buf[q] = int(data[start + q])

    def _set_region(self, env, arr, start, n, buf):
        data = self.objects[arr][1]
        for q in range(n):
            data[start + q] = buf[q]

    def byte_array(self, values):
        return self._new(("bytes", [ctypes.c_int8(v & 0xFF).value for v in values]))

    def int_array(self, values):
        return self._new(("ints", [int(v) for v in values]))

    def float_array(self, n):
        return self._new(("floats", [0.0] * n))

    def value(self, h):
        kind, data = self.objects[h]
        if kind == "objarray":
            return [self.value(x) for x in data]
        if kind == "bytes":
            return [v & 0xFF for v in data]
        return list(data)


def natives():
    lib = T.load()
    pre = "Java_br_unb_cic_tdp_base_GPUSortingSearch_"
    init = getattr(lib, pre + "initSlots")
    init.argtypes = [VP, VP, JINT, JINT, ctypes.c_int64, JINT]
    init.restype = None
    destroy = getattr(lib, pre + "destroySlots")
    destroy.argtypes = [VP, VP]
    destroy.restype = None
    search = getattr(lib, pre + "searchForSortingSeq")
    search.argtypes = [VP, VP, VP, VP, JINT, ctypes.c_float, JINT, ctypes.c_uint8, JINT]
    search.restype = VP
    cancel = getattr(lib, pre + "cancelSort")
    cancel.argtypes = [VP, VP, JINT]
    cancel.restype = None
    filt = getattr(lib, pre + "filterPiPermutations")
    filt.argtypes = [VP, VP, VP, JINT, VP, JINT, JINT, ctypes.c_uint8]
    filt.restype = VP
    return init, destroy, search, cancel, filt


def batch_natives():
    lib = T.load()
    pre = "Java_br_unb_cic_tdp_base_GPUSortingSearch_"
    batch = getattr(lib, pre + "searchBatch")
    batch.argtypes = [VP, VP, VP, VP, VP, VP, ctypes.c_float, JINT, ctypes.c_uint8, JINT, JINT]
    batch.restype = VP
    canon = getattr(lib, pre + "canonicalBatch")
    canon.argtypes = [VP, VP, VP, VP, VP, VP, VP, VP, VP]
    canon.restype = JINT
    return batch, canon


def _rows(cases, attr):
    out = []
    for c in cases:
        row = list(getattr(c, attr)) + [0] * (64 - c.n)
        out.extend(row)
    return out


def test_errors_become_runtime_exceptions():
    init, destroy, search, cancel, filt = natives()
    env = FakeJNIEnv()
    # filterPiPermutations with n = 0 (pi_filter.cu:166)
    r = filt(env.env, None, env.byte_array([0, 3, 1, 4, 2]), 5, env.byte_array([0, 1, 2]), 3, 0, 0)
    assert r is None
    assert env.exception == ("java/lang/RuntimeException", "n must be at least 1")
    # searchForSortingSeq on a slot that does not exist
    env.exception = None
    pi = list(range(8))
    r = search(env.env, None, env.byte_array(pi), env.byte_array(pi), 8, 1.0, 4, 0, 99)
    assert r is None
    assert env.exception == ("java/lang/RuntimeException", "invalid slot")
    # spi shorter than pi
    env.exception = None
    r = search(env.env, None, env.byte_array(pi), env.byte_array(pi[:4]), 8, 1.0, 4, 0, 0)
    assert r is None and "inconsistent input sizes" in env.exception[1]
    # invalid maximumDepth (sorting_search.cu:715)
    env.exception = None
    r = search(env.env, None, env.byte_array(pi), env.byte_array(pi), 8, 1.0, 16, 0, 0)
    assert r is None and env.exception[1] == "performGpuSortingSearch: invalid maximumDepth"
    # initSlots with no slots
    env.exception = None
    init(env.env, None, 0, 1, 1 << 20, 1 << 10)
    assert env.exception == ("java/lang/RuntimeException", "totalSlots must be positive")
    # cancelSort never throws
    env.exception = None
    cancel(env.env, None, 3)
    assert env.exception is None


def test_bytes_are_read_unsigned():
    """Symbols >= 128 must arrive as n >= 128 and be rejected as n-range errors,
    not as negative values (jni_sorter.cu:63,68 mask with 0xFF)."""
    init, destroy, search, cancel, filt = natives()
    env = FakeJNIEnv()
    r = filt(env.env, None, env.byte_array([200]), 1, env.byte_array([0]), 1, 5, 1)
    assert r is None and env.exception[1] == "cycle symbol out of range"


@pytest.mark.gpu
def test_jni_round_trip_on_gpu(gpu_lib):
    init, destroy, search, cancel, filt = natives()
    env = FakeJNIEnv()
    init(env.env, None, 1, 1, 64 << 20, 1 << 20)
    assert env.exception is None
    try:
        wl = W.workload_random_n12()
        for case in wl.cases[:8]:
            r = search(env.env, None, env.byte_array(case.pi), env.byte_array(case.spi), case.even_cycles,
                       wl.min_rate, wl.max_depth, 0, 0)
            assert env.exception is None and r is not None
            positions = [tuple(m) for m in env.value(r)]
            want = oracle.search(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth)
            assert bool(positions) == want.found
            if positions:
                e, _, _ = oracle.replay_positions(case.pi, case.spi, positions)
                assert oracle.satisfies_rate(e, len(positions), case.n, case.even_cycles, wl.min_rate, False)
        r = filt(env.env, None, env.byte_array([0, 3, 1, 4, 2]), 5, env.byte_array([0, 1, 2]), 3, 8, 0)
        rows = env.value(r)
        _, want_rows = oracle.filter_pi([0, 3, 1, 4, 2], [0, 1, 2], 8, False, 1 << 20)
        assert np.array_equal(np.array(rows, np.uint8).reshape(-1, 8), want_rows)
    finally:
        destroy(env.env, None)
    assert env.exception is None


@pytest.mark.gpu
def test_jni_out_of_memory_becomes_runtime_exception(gpu_lib):
    """initSlots with a queue budget no device holds throws RuntimeException
    with the allocation message (jni_sorter.cu:96-103, cuda_memory.cuh:82-85);
    initSlots works again afterwards."""
    init, destroy, search, cancel, filt = natives()
    env = FakeJNIEnv()
    _lib.check(gpu_lib.tdp_set_alloc_timeout(0))
    try:
        init(env.env, None, 2, 1, 1 << 40, 1 << 16)
        assert env.exception == ("java/lang/RuntimeException",
                                 f"HIP out of memory while allocating {1 << 40} bytes for sort.nodes")
    finally:
        _lib.check(gpu_lib.tdp_set_alloc_timeout(60000))
    env.exception = None
    init(env.env, None, 1, 1, 64 << 20, 1 << 20)
    assert env.exception is None
    try:
        wl = W.workload_random_n12()
        case = wl.cases[1]
        r = search(env.env, None, env.byte_array(case.pi), env.byte_array(case.spi), case.even_cycles,
                   wl.min_rate, 6, 0, 0)
        assert env.exception is None and r is not None
    finally:
        destroy(env.env, None)


@pytest.mark.gpu
def test_jni_concurrent_callers_share_the_serving_session(gpu_lib, monkeypatch):
    """TDP_SERVE=1: initSlots starts one serving session holding the slots'
    queue memory, and searchForSortingSeq calls from eight threads (one slot
    and one JNIEnv each, as ForkJoin workers) share it.  Every result equals
    the oracle's label, in the mix and alone (one call at a time from one
    thread); per-call latency of the JNI path is printed."""
    from helpers import load_serving_pool
    import threading
    import time
    sat, unsat, rate, doc = load_serving_pool()
    pool = sat[:48] + unsat[:48]
    labels = [True] * 48 + [False] * 48
    monkeypatch.setenv("TDP_SERVE", "1")
    monkeypatch.setenv("TDP_SERVE_LIFETIME_S", "300")
    init, destroy, search, cancel, filt = natives()
    env0 = FakeJNIEnv()
    init(env0.env, None, 8, 1, 1 << 30, 1 << 22)
    assert env0.exception is None
    try:
        def call(env, c, slot):
            r = search(env.env, None, env.byte_array(c.pi), env.byte_array(c.spi), c.even_cycles, rate,
                       doc["max_depth"], 0, slot)
            assert env.exception is None and r is not None, env.exception
            return [tuple(m) for m in env.value(r)]

        results = [None] * len(pool)
        lat = [0.0] * len(pool)
        errors = []

        def worker(t):
            env = FakeJNIEnv()
            try:
                for q in range(t, len(pool), 8):
                    t0 = time.perf_counter()
                    results[q] = call(env, pool[q], t)
                    lat[q] = time.perf_counter() - t0
            except BaseException as e:  # noqa: BLE001
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(240)
        assert not errors, errors
        assert [bool(r) for r in results] == labels
        for c, pos in zip(pool, results):
            if pos:
                e, _, _ = oracle.replay_positions(c.pi, c.spi, pos)
                assert oracle.satisfies_rate(e, len(pos), c.n, c.even_cycles, rate, False)
        alone = [bool(call(env0, c, 0)) for c in pool[::3]]
        assert alone == labels[::3]
        ms = np.array(lat) * 1e3
        print(f"JNI path, 8 threads: p50 {np.percentile(ms, 50):.2f} ms, p99 {np.percentile(ms, 99):.2f} ms",
              flush=True)
    finally:
        destroy(env0.env, None)
    assert env0.exception is None


def test_batch_natives_reject_bad_arguments_without_a_gpu():
    """searchBatch / canonicalBatch: size mismatches, an invalid slot and a
    configuration with a fixed point become RuntimeExceptions (checked before
    any device work)."""
    batch, canon = batch_natives()
    env = FakeJNIEnv()
    wl = W.workload_random_n12()
    cases = wl.cases[:3]
    r = batch(env.env, None, env.byte_array(_rows(cases, "pi")), env.byte_array(_rows(cases, "spi")[:64]),
              env.int_array([12] * 3), env.int_array([c.even_cycles for c in cases]), wl.min_rate, 6, 0, 0, 0)
    assert r is None and env.exception[1] == "performGpuSortingSearch: inconsistent input sizes"
    env.exception = None
    r = batch(env.env, None, env.byte_array(_rows(cases, "pi")), env.byte_array(_rows(cases, "spi")),
              env.int_array([12] * 3), env.int_array([c.even_cycles for c in cases]), wl.min_rate, 6, 0, 0, 9)
    assert r is None and env.exception == ("java/lang/RuntimeException", "invalid slot")
    env.exception = None
    spi = list(range(5)) + [0] * 59     # identity: every symbol a fixed point
    pi = list(range(5)) + [0] * 59
    got = canon(env.env, None, env.byte_array(spi), env.byte_array(pi), env.int_array([5]), env.float_array(64),
                env.int_array([0]), env.int_array([0]), env.byte_array([0] * 64))
    assert got == 0 and env.exception == ("java/lang/RuntimeException", "spi has a fixed point")


@pytest.mark.gpu
def test_batch_natives_round_trip_on_gpu(gpu_lib):
    """searchBatch equals the oracle (deterministic mode: verdicts and
    sequences) and canonicalBatch reproduces ConfigurationTest's canonical
    signatures and hash codes."""
    import json
    import os
    from oracle import canon_ref as C
    batch, canon = batch_natives()
    init, destroy, search, cancel, filt = natives()
    env = FakeJNIEnv()
    init(env.env, None, 1, 1, 256 << 20, 1 << 20)
    try:
        wl = W.workload_random_n12()
        cases = wl.cases[:64]
        r = batch(env.env, None, env.byte_array(_rows(cases, "pi")), env.byte_array(_rows(cases, "spi")),
                  env.int_array([c.n for c in cases]), env.int_array([c.even_cycles for c in cases]), wl.min_rate,
                  wl.max_depth, 0, T.FLAG_DETERMINISTIC, 0)
        assert env.exception is None and r is not None
        out = np.array(env.value(r), np.int64).reshape(len(cases), 47)
        for c, case in enumerate(cases):
            want = oracle.search(case.pi, case.spi, case.even_cycles, wl.min_rate, wl.max_depth)
            assert bool(out[c, 0]) == want.found
            assert int(out[c, 1]) == want.depth
            assert [tuple(out[c, 2 + 3 * m: 5 + 3 * m]) for m in range(want.depth)] == want.moves
    finally:
        destroy(env.env, None)
    with open(os.path.join(os.path.dirname(__file__), "golden", "configuration_vectors.json")) as fh:
        gold = json.load(fh)["configurations"]
    spis, pis, ns = [], [], []
    for g in gold:
        cyc = C.parse_cycles(g["spi"])
        n = sum(len(c) for c in cyc)
        spi = W.one_line(cyc, n)
        spis += spi + [0] * (64 - n)
        pis += list(range(n)) + [0] * (64 - n)
        ns.append(n)
    sig, hashes, index, cs = env.float_array(64 * len(ns)), env.int_array([0] * len(ns)), \
        env.int_array([0] * len(ns)), env.byte_array([0] * 64 * len(ns))
    got = canon(env.env, None, env.byte_array(spis), env.byte_array(pis), env.int_array(ns), sig, hashes, index, cs)
    assert env.exception is None and got == len(ns)
    sigs = np.array(env.value(sig), np.float32).reshape(len(ns), 64)
    for q, g in enumerate(gold):
        assert env.value(hashes)[q] == g["hash"]
        assert np.array_equal(sigs[q, :ns[q]], np.asarray(g["canonical"], np.float32))
