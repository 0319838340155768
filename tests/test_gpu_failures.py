"""GPU failure paths forced from the host: queue overflow and cancellation."""
import threading
import time

import pytest

import oracle
import tdp1375_amd as T
from tdp1375_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_queue_overflow_is_reported(gpu_lib):
    """A queue budget below the root count raises the reference's overflow error."""
    T.init_slots(1, 1, 4096, 1 << 16)
    try:
        wl = W.workload_four_3cycles_n24(False)
        case = wl.cases[0]
        with pytest.raises(T.QueueOverflowError, match="GPU sorting search queue overflow"):
            T.search_for_sorting_seq(case.pi, case.spi, case.even_cycles, wl.min_rate, 9, False, 0,
                                     T.FLAG_SINGLE_WAVE)
    finally:
        T.destroy_slots()


def test_cancel_from_another_thread(gpu_lib):
    """cancelSort during a long single-wave search raises the cancelled error."""
    T.init_slots(1, 1, 512 << 20, 1 << 22)
    try:
        wl = W.workload_odd_cycle_mix_d11(8)
        case = wl.cases[7]
        box = {}

        def run():
            try:
                box["res"] = T.search_for_sorting_seq(case.pi, case.spi, case.even_cycles, wl.min_rate, 11, False,
                                                      0, T.FLAG_SINGLE_WAVE)
            except Exception as e:  # noqa: BLE001
                box["err"] = e

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.3)
        T.cancel_sort(0)
        th.join(60)
        assert not th.is_alive()
        assert isinstance(box.get("err"), T.CancelledError), box
        assert str(box["err"]) == "GPU sorting search cancelled"
        # the slot is usable again after a cancel
        small = W.workload_random_n12().cases[0]
        res = T.search_for_sorting_seq(small.pi, small.spi, small.even_cycles, wl.min_rate, 6, False, 0)
        want = oracle.search(small.pi, small.spi, small.even_cycles, wl.min_rate, 6)
        assert bool(res.positions) == want.found
    finally:
        T.destroy_slots()


def test_out_of_memory_is_reported_and_slots_recover(gpu_lib):
    """A queue budget no device can hold fails with the reference's allocation
    error (cuda_memory.cuh:82-85 with "HIP"), after the configured retry
    timeout (0 here: no retry); the failed initSlots leaves no slots, and
    slots can be initialised and used again afterwards."""
    budget = 1 << 40   # 1 TiB of queue memory
    T._lib.check(gpu_lib.tdp_set_alloc_timeout(0))
    try:
        with pytest.raises(T.TdpError) as ei:
            T.init_slots(1, 1, budget, 1 << 16)
        assert ei.value.status == T._lib.TDP_ERR_OUT_OF_MEMORY
        assert str(ei.value) == f"HIP out of memory while allocating {budget} bytes for sort.nodes"
        small = W.workload_random_n12().cases[0]
        with pytest.raises(T.TdpError, match="invalid slot"):
            T.search_for_sorting_seq(small.pi, small.spi, small.even_cycles, 1.0, 6, False, 0)
        # one retry (100 ms backoff) before giving up at a 50 ms timeout
        retries = gpu_lib.tdp_alloc_retries()
        T._lib.check(gpu_lib.tdp_set_alloc_timeout(50))
        with pytest.raises(T.TdpError, match="HIP out of memory while allocating"):
            T.init_slots(1, 1, budget, 1 << 16)
        assert gpu_lib.tdp_alloc_retries() >= retries + 1
    finally:
        T._lib.check(gpu_lib.tdp_set_alloc_timeout(60000))
    T.init_slots(1, 1, 64 << 20, 1 << 20)
    try:
        wl = W.workload_random_n12()
        small = wl.cases[0]
        res = T.search_for_sorting_seq(small.pi, small.spi, small.even_cycles, wl.min_rate, 6, False, 0)
        want = oracle.search(small.pi, small.spi, small.even_cycles, wl.min_rate, 6)
        assert bool(res.positions) == want.found
    finally:
        T.destroy_slots()
