"""MI355X-native (gfx950) proof-search path of tdp1375.

The hot path is HIP C++ in ``csrc/`` built into ``libtdp1375_jni.so``; this
package is the host-side mirror of the reference's GPUSortingSearch facade
over the library's C ABI, plus the seeded workload generators.
"""
from ._lib import (FLAG_DETERMINISTIC, FLAG_SINGLE_WAVE, CancelledError, QueueOverflowError, TdpError,
                   load)
from .gpu_sorting_search import (GPUSortingSearch, cancel_sort, destroy_slots, filter_pi_permutations,
                                 init_slots, last_launch_stats, positions_to_symbols, search_batch,
                                 search_for_sorting_seq, set_expansion_batch, set_tuning)

__all__ = [
    "GPUSortingSearch", "init_slots", "destroy_slots", "cancel_sort", "filter_pi_permutations",
    "search_for_sorting_seq", "search_batch", "set_tuning", "set_expansion_batch", "positions_to_symbols", "last_launch_stats", "load",
    "TdpError", "CancelledError", "QueueOverflowError", "FLAG_DETERMINISTIC", "FLAG_SINGLE_WAVE",
]
