"""Autotuner of the HH population's launches: times candidates on the workload itself.

Two choices are measured rather than fixed:

* the matrix-free product (``store="auto"``): the Chebyshev series costs about sqrt(rho) terms
  per product, rho growing with the time step and the stiffness of the cells, while the rational
  store costs five tree solves whatever the step (one barrier per tree level); which is faster
  depends on dt, the cells' sizes and depths and the device;
* the workgroup size of the stored formats (``block_threads="auto"``): 128, 256, 512 or 1024
  threads per cell, whose best value depends on the rows per cell and the stored entries per row
  (LDS and register occupancy against HBM streaming);
* the launch form of the Chebyshev store (``cheb_launch="auto"``): one launch per half step, per
  step or per two steps.  Fewer launches pay fewer launch starts and tails and keep v in LDS, but
  the fused builds hold more registers (the one-row-group launches, at 64 VGPRs for four
  workgroups per CU, spill more as the form grows), so the best form depends on the cells.

Each candidate runs ``warmup`` untimed steps, then ``steps`` steps in one call timed with HIP
events on the launching stream; the fastest mean time per step wins.  Results are kept per
process, keyed by the workload's signature, so a second population of the same shape is not
tuned again.  The matrix-free stores apply the same unpruned exp(A dt) (to about 1e-10), so the
choice changes the time, not the results beyond rounding.
"""
import time

import numpy as np

MATRIX_FREE = ("chebyshev", "rational")
BLOCKS = (128, 256, 512, 1024)

_CACHE = {}


def time_steps(pop, steps=6, warmup=2):
    """Mean milliseconds per step of ``pop.advance(steps)`` after ``warmup`` steps (HIP events on
    the current stream; the population's state advances)."""
    torch = pop.torch
    if warmup:
        pop.advance(warmup)
    stream = torch.cuda.current_stream(pop.dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize(pop.dev)
    e0.record(stream)
    pop.advance(steps)
    e1.record(stream)
    torch.cuda.synchronize(pop.dev)
    return e0.elapsed_time(e1) / steps


def _signature(cells, kwargs, what, candidates):
    sizes = np.array([len(c) for c in cells], dtype=np.int64)
    keys = ("dt", "celsius", "gate_mode", "device", "prune", "store")
    stim = kwargs.get("stim") or {}
    return (what, tuple(candidates), len(cells), int(sizes.sum()), int(sizes.max()),
            int(np.sort(sizes)[len(sizes) // 2]), tuple(kwargs.get(k) for k in keys),
            stim.get("mode"))


def choose_store(cells, candidates=MATRIX_FREE, steps=6, warmup=2, build=None, timer=None,
                 **kwargs):
    """The fastest matrix-free store for an HHPopulation of ``cells`` built with ``kwargs``.

    build(store) -> population and timer(population) -> ms per step default to HHPopulation and
    time_steps.  A candidate whose build raises (a cell beyond its limits) is skipped.  Returns
    (store, report) with report = {"choice", "ms_per_step": {store: ms}, "skipped": {store: msg},
    "tuning_s", "cached"}."""
    key = _signature(cells, kwargs, "store", candidates)
    if build is None and timer is None and key in _CACHE:
        return _CACHE[key]["choice"], dict(_CACHE[key], cached=True)
    if build is None:
        from .population import HHPopulation
        build = lambda store: HHPopulation(cells, store=store, **kwargs)
    timer = timer or (lambda pop: time_steps(pop, steps, warmup))
    t0 = time.perf_counter()
    ms, skipped = {}, {}
    for store in candidates:
        try:
            pop = build(store)
        except (RuntimeError, ValueError) as e:
            skipped[store] = str(e)
            continue
        ms[store] = float(timer(pop))
        del pop
    if not ms:
        raise RuntimeError(f"no candidate store fits these cells: {skipped}")
    choice = min(ms, key=ms.get)
    report = {"choice": choice, "ms_per_step": ms, "skipped": skipped,
              "tuning_s": time.perf_counter() - t0, "cached": False}
    _CACHE[key] = report
    return choice, report


def choose_block_threads(pop, candidates=BLOCKS, steps=6, warmup=2, timer=None):
    """The fastest workgroup size of a population on a stored format, timed on the population
    itself (its state advances; the caller resets it).  Returns (block_threads, report)."""
    timer = timer or (lambda p: time_steps(p, steps, warmup))
    t0 = time.perf_counter()
    ms = {}
    for bt in candidates:
        pop.params.block_threads = int(bt)
        ms[int(bt)] = float(timer(pop))
    choice = min(ms, key=ms.get)
    pop.params.block_threads = choice
    return choice, {"choice": choice, "ms_per_step": ms, "tuning_s": time.perf_counter() - t0}


CHEB_LAUNCHES = ("half", "step", "two_steps")


def choose_cheb_launch(pop, candidates=CHEB_LAUNCHES, steps=6, warmup=2, timer=None):
    """The fastest launch form of a population on the Chebyshev store, timed on the population
    itself (its state advances; the caller resets it).  The forms give bitwise the same results.
    Returns (form, report)."""
    from .population import CHEB_LAUNCH
    timer = timer or (lambda p: time_steps(p, steps, warmup))
    t0 = time.perf_counter()
    ms = {}
    for form in candidates:
        pop.params.cheb_launch = CHEB_LAUNCH[form]
        ms[form] = float(timer(pop))
    choice = min(ms, key=ms.get)
    pop.params.cheb_launch = CHEB_LAUNCH[choice]
    return choice, {"choice": choice, "ms_per_step": ms, "tuning_s": time.perf_counter() - t0}
