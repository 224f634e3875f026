"""The autotuner's selection logic on the CPU: the fastest candidate wins, a candidate whose build
fails is skipped, results are cached per workload signature, and the block-size sweep leaves the
chosen size in the population's parameters."""
import types

import pytest

from neuwon_amd import autotune, synth


def _cells(n_cells=3, n=40, seed=0):
    return synth.uniform_population(num_cells=n_cells, n=n, seed=seed)


def test_choose_store_picks_the_fastest_and_skips_failures():
    times = {"chebyshev": 2.0, "rational": 1.5, "broken": None}

    def build(store):
        if times[store] is None:
            raise RuntimeError("cell of 6000 segments exceeds the store")
        return types.SimpleNamespace(store=store)

    choice, rep = autotune.choose_store(_cells(), candidates=("chebyshev", "rational", "broken"),
                                        build=build, timer=lambda p: times[p.store], dt=0.1)
    assert choice == "rational"
    assert rep["ms_per_step"] == {"chebyshev": 2.0, "rational": 1.5}
    assert "broken" in rep["skipped"] and not rep["cached"]
    times["rational"] = 3.0
    choice, _ = autotune.choose_store(_cells(), candidates=("chebyshev", "rational"),
                                      build=build, timer=lambda p: times[p.store], dt=0.5)
    assert choice == "chebyshev"


def test_choose_store_without_a_fitting_candidate_raises():
    def build(store):
        raise RuntimeError("too large")
    with pytest.raises(RuntimeError, match="no candidate store"):
        autotune.choose_store(_cells(), build=build, timer=lambda p: 1.0)


def test_cache_is_keyed_by_the_workload():
    cells = _cells(seed=3)
    key = autotune._signature(cells, {"dt": 0.1}, "store", autotune.MATRIX_FREE)
    autotune._CACHE[key] = {"choice": "rational", "ms_per_step": {}, "skipped": {},
                            "tuning_s": 0.0, "cached": False}
    try:
        choice, rep = autotune.choose_store(cells, dt=0.1)   # no build: served from the cache
        assert choice == "rational" and rep["cached"]
        other = autotune._signature(cells, {"dt": 0.5}, "store", autotune.MATRIX_FREE)
        assert other != key
        assert autotune._signature(_cells(n=41, seed=3), {"dt": 0.1}, "store",
                                   autotune.MATRIX_FREE) != key
    finally:
        autotune._CACHE.pop(key, None)


def test_choose_block_threads_sets_the_choice():
    pop = types.SimpleNamespace(params=types.SimpleNamespace(block_threads=0))
    cost = {128: 9.0, 256: 4.0, 512: 5.0, 1024: 6.0}
    seen = []

    def timer(p):
        seen.append(p.params.block_threads)
        return cost[p.params.block_threads]
    choice, rep = autotune.choose_block_threads(pop, timer=timer)
    assert choice == 256 and pop.params.block_threads == 256
    assert seen == [128, 256, 512, 1024] and rep["ms_per_step"] == cost
