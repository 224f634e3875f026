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


def test_choose_cheb_launch_sets_the_choice():
    pop = types.SimpleNamespace(params=types.SimpleNamespace(cheb_launch=0))
    cost = {1: 3.0, 2: 2.5, 3: 2.6}          # NW_CHEB_LAUNCH_HALF, _STEP, _TWO_STEPS
    seen = []

    def timer(p):
        seen.append(p.params.cheb_launch)
        return cost[p.params.cheb_launch]
    choice, rep = autotune.choose_cheb_launch(pop, timer=timer)
    assert choice == "step" and pop.params.cheb_launch == 2
    assert seen == [1, 2, 3]
    assert rep["ms_per_step"] == {"half": 3.0, "step": 2.5, "two_steps": 2.6}


def test_cheb_launch_names_match_the_header():
    """population.CHEB_LAUNCH and _lib's constants give the values include/neuwon_amd.h defines."""
    import os
    import re
    from neuwon_amd import _lib, population
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "neuwon_amd.h")).read()
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define NW_CHEB_LAUNCH_(\w+) (\d+)", header)}
    assert defs == {"DEFAULT": 0, "HALF": 1, "STEP": 2, "TWO_STEPS": 3}
    assert population.CHEB_LAUNCH == {None: defs["DEFAULT"], "half": defs["HALF"],
                                      "step": defs["STEP"], "two_steps": defs["TWO_STEPS"]}
    assert (_lib.NW_CHEB_LAUNCH_DEFAULT, _lib.NW_CHEB_LAUNCH_HALF, _lib.NW_CHEB_LAUNCH_STEP,
            _lib.NW_CHEB_LAUNCH_TWO_STEPS) == (0, 1, 2, 3)
    fields = [f for f, _ in _lib.NwHHParams._fields_]
    assert fields[-3:] == ["launch_mode", "cheb_launch", "reserved"]

