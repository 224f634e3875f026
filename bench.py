"""Throughput of the per-timestep advance on one MI355X.

Workload (default): a population of independent random-tree cells, 200-5000 compartments each
(uniform), drawn until 10M compartments; leak + HH Na/K as two-state kinetic schemes on every
segment, per-cell gNa/gK scaled logU(0.5, 2); Poisson 5 Hz current pulses (0.5 nA x 1 ms) at every
root; FP32 state; dt = 0.1 ms.  The propagator exp(A dt) of every cell is applied without a stored matrix,
unpruned, as a Chebyshev series of its tree operator (--format chebyshev) or as five exact tree
solves of a rational approximation (--format rational) -- by default (--format auto) whichever
the autotuner times faster on this workload before the timed region -- or stored pruned at 1e-9
(--format symmetric / general, FP64 construction).

A step is one staggered time step of every compartment.  W warm-up steps run first, then exactly K
timed steps, bracketed by torch.cuda.synchronize().  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per compartment-step of the step kernel on this workload, per propagator format, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE counter runs scaled by what the same counters read for a
# plain 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); None where not measured
MEASURED_TRAFFIC_PER_CMPT_STEP = {"general": 6896.1, "symmetric": 3753.3, "chebyshev": 154.4,
                                  "rational": None}
# Kept entries per row of the propagator pruned at 1e-9 (pair rule), this seeded 10M workload:
# the canonical algorithmic count of the Chebyshev store, which keeps no matrix, uses it
NNZ_BAR_PRUNED_1E9 = 567.99


def algorithmic_bytes(nnz_bar):
    """Canonical algorithmic bytes per compartment-step of the fused step: two half steps of
    8 nnz + 28 B (FP32 value + i32 index per kept entry, row pointer, v/g/d/I/(1/C) reads, v write)
    and a 68-B HH mechanism pass.  nnz_bar = kept entries per row (both triangles).  The canonical
    count stays the same whatever the storage format, so a compressed format shows up as an
    effective bandwidth above what it moves."""
    return 16.0 * nnz_bar + 124.0


def format_bytes(fmt, nnz_bar):
    """Bytes per compartment-step the stored format itself needs: per half step the stored entries
    (FP32 value + u16 column; the symmetric format stores the diagonal and one triangle) and per row
    the offset (4 B) and FP64 row sum (8 B) plus, symmetric, the FP64 weight (8 B); per step 68 B of
    per-segment mechanism state and constants."""
    if fmt == "symmetric":
        return 2.0 * (6.0 * (nnz_bar + 1.0) / 2.0 + 20.0) + 68.0
    if fmt == "chebyshev":   # per half step the LDS place, 2 far places and 4 FP32 coefficients,
        return 2.0 * (2.0 + 4.0 + 16.0 + 16.0) + 68.0   # FP64 v in/out
    if fmt == "rational":    # per half step the level place, link, 2 FP32 edge coefficients, five
        return 2.0 * (2.0 + 4.0 + 8.0 + 5 * 16.0 + 16.0) + 68.0   # FP64 pivots, FP64 v in/out
    return 2.0 * (6.0 * nnz_bar + 12.0) + 68.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--compartments", type=float, default=1e7)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--dt", type=float, default=0.1, help="time step (ms) of the population")
    ap.add_argument("--launch-steps", type=int, default=0,
                    help="steps per kernel launch (0: all timed steps in one launch)")
    ap.add_argument("--block-threads", type=int, default=0)
    ap.add_argument("--format", choices=("symmetric", "general", "chebyshev", "rational", "auto"),
                    default="auto",
                    help="propagator store: symmetric half (S = W P W^-1), general CSR, "
                         "chebyshev (no stored matrix: series of the tree operator, unpruned), "
                         "rational (no stored matrix: five-pole rational approximation, exact "
                         "tree solves, unpruned) or auto (the faster matrix-free store on this "
                         "workload, timed by the autotuner before the timed region)")
    ap.add_argument("--cpu-sample-cells", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the gradients/s measurement of the reverse pass")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--accumulate", action="store_true",
                    help="only time the segmented accumulation (nw_accumulate) on --compartments "
                         "segments with three instances each (hh's na, k, l), one JSON line")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _sample_propagator(pop, sample, rule, prune):
    """FP64 propagator of the first `sample` cells from the device construction path, pruned at
    `prune` by the entry rule, the pair rule of the symmetric store, or not at all ("none")."""
    import scipy.sparse
    from neuwon_amd import population
    blocks = []
    for c in range(sample):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        args = (pop.parent_local[r0:r1], pop.R[r0:r1], pop.C[r0:r1], 0.5 * pop.dt)
        if rule == "none":
            blocks.append(population.cell_propagator_f64(*args, 0.0))
        elif rule == "pair":
            full = population.cell_propagator_f64(*args, 0.0).toarray()
            keep = np.maximum(np.abs(full), np.abs(full.T)) >= prune
            np.fill_diagonal(keep, True)
            blocks.append(scipy.sparse.csr_matrix(np.where(keep, full, 0.0)))
            del full, keep
        else:
            blocks.append(population.cell_propagator_f64(*args, prune))
    return scipy.sparse.block_diag(blocks, format="csr")


def cpu_baseline(pop, args, total_steps):
    """The oracle's C FP64 restatement of the same step (CSR SpMV, table gates, Poisson stimulus),
    fed the FP64 propagator the device built, pruned at 1e-9 as the store prunes it (the entry
    rule for the general and Chebyshev stores -- the latter keeps no matrix, and the baseline is
    the reference's operator of this workload -- the pair rule for the symmetric store), on a
    bounded sample of the population's cells: one thread (the reference's serial path) and every
    host core.  The all-cores run covers exactly the steps the GPU ran, so its voltages also give
    the precision of the GPU run on the sample; for the Chebyshev store, which applies exp(A dt)
    unpruned, a second all-cores run with the unpruned propagator on the first few cells does."""
    from oracle import neuwon_oracle as O
    cores = len(os.sched_getaffinity(0))
    sample = min(args.cpu_sample_cells, pop.num_cells)
    r_end = int(pop.cell_ptr[sample])
    rule = "pair" if pop.propagator.symmetric else "entry"
    P = _sample_propagator(pop, sample, rule, 1e-9 if pop.propagator.chebyshev else pop.propagator.prune)
    p = pop.params
    stim = ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
            int(p.pulse_ticks))

    def make(cells, P):
        r = int(pop.cell_ptr[cells])
        gbar = {k: pop.g_bar[k][:r] for k in ("na", "k", "l")}
        return O.CHHPopulation(pop.cell_ptr[:cells + 1], P, pop.C[:r], gbar, dt=pop.dt,
                               celsius=pop.celsius, gate_mode="table", stim=stim)
    rates = {}
    # one thread: a bounded run
    co = make(sample, P)
    co.advance(3, threads=1)   # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        co.advance(1, threads=1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds / 2 or steps >= 2000:
            break
    rates[1] = (r_end * steps / el, steps)
    # every core: exactly the GPU's steps (3 untimed warm-up steps first), recording every
    # cell's root voltage after each step
    co = make(sample, P)
    roots = pop.cell_ptr[:sample]
    trace = np.empty((sample, total_steps))
    warm = min(3, total_steps - 1)
    for k in range(warm):
        co.advance(1, threads=cores)
        trace[:, k] = co.v[roots]
    t0 = time.perf_counter()
    for k in range(warm, total_steps):
        co.advance(1, threads=cores)
        trace[:, k] = co.v[roots]
    el = time.perf_counter() - t0
    rates[cores] = (r_end * (total_steps - warm) / el, total_steps - warm)
    for t in (1, cores):
        log(f"cpu baseline {t} thread(s): {rates[t][0] / 1e6:.2f} M cmpt-steps/s")
    base = {"value": rates[cores][0], "unit": "compartment-steps/s", "cores": cores, "kind": "port",
            "value_1_thread": rates[1][0],
            "sample": f"first {sample} cells ({r_end} compartments, {P.nnz} FP64 propagator entries) "
                      f"of the same population; {rates[1][1]} steps on 1 thread, {rates[cores][1]} "
                      f"steps on {cores} threads; oracle C FP64 restatement (OpenMP over cells)"}
    if not pop.propagator.chebyshev:
        return base, co.v.copy(), trace, r_end
    few = min(6, sample)
    P0 = _sample_propagator(pop, few, "none", 0.0)
    co = make(few, P0)
    trace = np.empty((few, total_steps))
    for k in range(total_steps):
        co.advance(1, threads=cores)
        trace[:, k] = co.v[pop.cell_ptr[:few]]
    return base, co.v.copy(), trace, int(pop.cell_ptr[few])


def _trace_error(v, ref, dt, window_ms=2.0, threshold=20.0):
    """Errors of a voltage trace against the FP64 one: relative over samples more than window_ms
    from every upward threshold crossing of the reference, absolute (mV) within window_ms."""
    up = np.nonzero((ref[1:] >= threshold) & (ref[:-1] < threshold))[0] + 1
    t = np.arange(len(ref)) * dt
    near = np.zeros(len(ref), dtype=bool)
    for k in up:
        near |= np.abs(t - t[k]) <= window_ms
    d = np.abs(v - ref)
    rel = d / np.abs(ref)
    return (len(up), rel[~near].max() if (~near).any() else 0.0,
            d[near].max() if near.any() else 0.0, np.median(rel))


def precision(pop, cells, args, v_ref, trace_ref, r_end, total_steps):
    """Voltage error of the GPU against the FP64 oracle on the sampled cells: (1) the full state
    of the timed population after all its steps; (2) every cell's root voltage after every step,
    from the same cells advanced again as a population of their own in one launch (cells are
    independent; the rerun differs from the timed run only by the launch split — the general
    store moves its per-cell reference voltage, the Chebyshev store rounds the voltage to FP32
    between calls — and the order of the symmetric store's atomics: rerun_max_abs_diff reports by
    how much), split into subthreshold samples (relative) and samples within 2 ms of a spike
    (absolute)."""
    import torch
    from neuwon_amd import population
    v_gpu = pop.v.cpu().numpy().astype(np.float64)[:r_end]
    d = np.abs(v_gpu - v_ref)
    sample = len(trace_ref)
    sub = population.HHPopulation(cells[:sample], dt=pop.dt, celsius=pop.celsius,
                                  prune=pop.propagator.prune, gate_mode="table",
                                  stim=pop.stim, block_threads=args.block_threads,
                                  store=pop.propagator.store)
    tr = torch.zeros((sample, total_steps), dtype=torch.float32, device=sub.dev)
    sub.advance(total_steps, trace=tr)
    rerun = float(np.abs(sub.v.cpu().numpy().astype(np.float64) - v_gpu).max())
    tr = tr.cpu().numpy().astype(np.float64)
    errs = [_trace_error(tr[c], trace_ref[c], pop.dt) for c in range(sample)]
    return {"state_max_abs_v_error_mV": float(d.max()),
            "state_median_rel_v_error": float(np.median(d / np.abs(v_ref))),
            "trace_spikes": int(sum(e[0] for e in errs)),
            "trace_max_rel_v_error_subthreshold": float(max(e[1] for e in errs)),
            "trace_max_abs_v_error_near_spikes_mV": float(max(e[2] for e in errs)),
            "trace_median_rel_v_error": float(np.median([e[3] for e in errs])),
            "rerun_max_abs_diff_mV": rerun,
            "sample": f"first {sample} cells, {r_end} compartments, {total_steps} steps",
            "definition": "FP64 oracle with the store's propagator (pruned as stored; unpruned for "
                          "the Chebyshev store), tables and stimulus; "
                          "state: every compartment after all steps; trace: root voltage of every "
                          "sampled cell after every step, subthreshold = more than 2 ms from a "
                          "spike of the oracle trace"}


def adjoint_rate(seed=4, store=None):
    """Gradients per second of the trajectory loss through the fused step: 1024 cells x 500
    compartments, 2000 steps at dt 0.05 ms, checkpoints every 100 steps, targets from a forward run
    with hidden per-segment conductances (logU(0.5, 2) x the defaults).  One warm-up gradient, then
    the mean of two timed ones."""
    import torch
    from neuwon_amd import adjoint, synth
    T, amp = 2000, 0.5
    cells = synth.uniform_population(num_cells=1024, n=500, seed=seed)
    prob = adjoint.HHAdjoint(cells, dt=0.05, stim_amp_na=amp, checkpoint_every=100,
                             forward_store=store, reverse_store=store)
    pop = prob.pop
    rng = np.random.default_rng(seed)
    gb = {k: pop.g_bar[k] for k in ("na", "k", "l")}
    hidden = {k: gb[k] * np.exp(rng.uniform(np.log(0.5), np.log(2.0), len(gb[k]))) for k in gb}
    f32 = lambda a: torch.as_tensor(a.astype(np.float32), device=pop.dev)
    prob.forward(f32(hidden["na"]), f32(hidden["k"]), f32(hidden["l"]), np.ones(6),
                 torch.zeros((pop.num_cells, T), device=pop.dev))
    target = prob.trace.clone()
    args = (f32(gb["na"]), f32(gb["k"]), f32(gb["l"]), np.ones(6), target)
    prob.loss_and_grad(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        prob.loss_and_grad(*args)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 2
    log(f"adjoint: {el:.3f} s per gradient")
    # the same gradient with the record of all 2000 steps kept in HBM (20.5 GB): no recomputation
    prob.K = None
    prob.loss_and_grad(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        prob.loss_and_grad(*args)
    torch.cuda.synchronize()
    el_rec = (time.perf_counter() - t0) / 2
    log(f"adjoint, whole record in HBM: {el_rec:.3f} s per gradient")
    return {"gradients_per_s": 1.0 / el, "seconds_per_gradient": el,
            "gradients_per_s_whole_record": 1.0 / el_rec,
            "store": prob.forward_store,
            "config": "1024 cells x 500 compartments, 2000 steps at dt 0.05 ms, checkpoints every "
                      "100 steps; gradients of sum (v_soma - target)^2 w.r.t. per-segment gNa, gK, "
                      "gL and 6 HH rate multipliers"}


def copy_probe(args):
    """A plain device-to-device copy of known size: the calibration run for rocprof's
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = int(args.copy_probe_mib) * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 3}), flush=True)


def accumulate_probe(args):
    """The segmented accumulation alone (the model path's nw_accumulate): every segment carries
    three conductance sources (hh's na, k and l, instances sorted by segment, the conductances
    gathered through an index as the model's interleaved mechanism output is), FP32 sources,
    int64 offsets and indices, sums in FP64.  Algorithmic bytes: per source the conductance, the
    reversal potential and the index (4 + 4 + 8), per segment its offset (8) and the three outputs
    (12)."""
    import ctypes
    import torch
    from neuwon_amd import _lib
    N, per = int(args.compartments), 3
    M = N * per
    dev = "cuda"
    gen = torch.Generator(device=dev).manual_seed(args.seed)
    ptr = torch.arange(0, M + 1, per, dtype=torch.int64, device=dev)
    g = torch.rand(M, generator=gen, device=dev) * 1e-9
    idx = torch.arange(M, dtype=torch.int64, device=dev)
    e = torch.tensor([50.0, -77.0, -54.3], device=dev).repeat(N)
    out = [torch.empty(N, device=dev) for _ in range(3)]
    lib = _lib.load()
    P = _lib.ptr
    stream = torch.cuda.current_stream()

    def launch():
        _lib.check(lib.nw_accumulate(N, P(ptr), P(g), P(idx), P(e), None, P(out[0]), P(out[1]),
                                     P(out[2]), ctypes.c_void_p(stream.cuda_stream)))
    for _ in range(max(args.warmup, 1)):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # check a sample against a host sum
    k = min(N, 4096)
    gs = g[:k * per].double().view(k, per).cpu().numpy()
    es = e[:k * per].double().view(k, per).cpu().numpy()
    want_g = gs.sum(1)
    want_d = (gs * es).sum(1) / want_g
    err = max(float(np.max(np.abs(out[0][:k].double().cpu().numpy() - want_g) / want_g)),
              float(np.max(np.abs(out[1][:k].double().cpu().numpy() - want_d) / np.abs(want_d))))
    alg = M * 16.0 + N * 20.0
    gbs = alg / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "segmented accumulation (nw_accumulate) bandwidth", "value": gbs,
                      "unit": "GB/s", "segments": N, "sources_per_segment": per,
                      "avg_launch_ms": ms, "launches": args.steps,
                      "algorithmic_bytes_per_launch": alg,
                      "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS},
                      "max_rel_error_vs_host_fp64": err}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (cells are "
              "independent, a multi-GPU run would shard whole cells with no per-step exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args)
        return
    if args.accumulate:
        accumulate_probe(args)
        return
    import torch
    from neuwon_amd import population, synth
    t0 = time.time()
    cells = synth.population(total=int(args.compartments), seed=args.seed)
    log(f"generated {len(cells)} cells in {time.time() - t0:.1f}s")
    t0 = time.time()
    pop = population.HHPopulation(cells, dt=args.dt, celsius=6.3, prune=1e-9, gate_mode="table",
                                  stim={"mode": "poisson", "amp_na": 0.5, "rate_hz": 5.0,
                                        "duration_ms": 1.0, "seed": args.seed},
                                  block_threads=args.block_threads, store=args.format)
    st = pop.propagator.stats
    fmt = pop.propagator.store
    if pop.tuning.get("store"):
        log(f"autotuner: {pop.tuning['store']}")
    nnz_bar = NNZ_BAR_PRUNED_1E9 if pop.propagator.chebyshev else pop.propagator.mean_nnz_per_row
    log(f"population ready in {time.time() - t0:.1f}s: {pop.num_segments} compartments, "
        f"nnz/row {nnz_bar:.1f}, propagator build {st['build_seconds']:.1f}s, "
        f"{st['device_bytes'] / 1e9:.2f} GB on device")
    N = pop.num_segments
    K, W = args.steps, args.warmup
    L = args.launch_steps or K
    if W:
        pop.advance(W)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = []
    t_start = time.perf_counter()
    done = 0
    while done < K:
        k = min(L, K - done)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        pop.advance(k)
        e1.record(stream)
        ev.append((k, e0, e1))
        done += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kern_ms = sum(a.elapsed_time(b) for _, a, b in ev)
    # kernel launches: one per advance() call, two per step on the Chebyshev store
    launches = 2 * K if pop.propagator.chebyshev else len(ev)
    value = N * K / wall
    b_step = algorithmic_bytes(nnz_bar)
    avg_launch_s = kern_ms / 1e3 / launches
    bytes_per_launch = b_step * N * (K / launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    stored_bytes_step = (st["stored"] * 6.0 * 2) / N  # FP32 value + u16 column, two half steps
    f_step = format_bytes(fmt, nnz_bar)
    traffic = MEASURED_TRAFFIC_PER_CMPT_STEP[fmt]
    units_per_s = N * (K / launches) / avg_launch_s
    out = {
        "metric": "compartment-steps/s @10M cmpts: FP32 batched vs FP64 plain; adjoint grads/s",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-tree cells, Poisson stimulus)",
        "config": {"workload": ("batched HH population, FP32 state, exp(A dt) applied as a five-pole "
                                "rational approximation, each pole an exact solve over the cell's "
                                "tree (no stored matrix, unpruned)" if pop.propagator.rational else
                                "batched HH population, FP32 state, exp(A dt) applied as a Chebyshev "
                                "series of each cell's tree operator (no stored matrix, unpruned)"
                                if pop.propagator.chebyshev else
                                "batched HH population, FP32 state, FP64-built propagator pruned 1e-9"),
                   "propagator_format": fmt, "compartments": N, "cells": pop.num_cells, "dt_ms": args.dt,
                   "nnz_per_row": round(nnz_bar, 2), "steps_per_launch": L,
                   "propagator_build_s": round(st["build_seconds"], 2),
                   "autotune": pop.tuning.get("store")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic * N * (K / launches),
                     "kernel": ("nw::k_hh_half_cheb (two launches per step)"
                                if pop.propagator.chebyshev else "nw::k_hh_step"),
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_cmpt_step": b_step,
                     "format_bytes_per_cmpt_step": f_step,
                     "format_GBs": f_step * units_per_s / 1e9,
                     "stored_propagator_bytes_per_cmpt_step": stored_bytes_step,
                     "measured_traffic_GBs": None if traffic is None else traffic * units_per_s / 1e9},
    }
    if pop.propagator.rational:
        # per pole and row: leaves to root a child's q read (16 B, one child per row on average)
        # and q written; root to leaves the row's q and its parent's z read and z written
        lds_b = 2.0 * 5 * 5 * 16.0
        out["roofline"].update({"lds_bytes_per_cmpt_step": lds_b,
                                "lds_GBs": lds_b * N * K / (kern_ms / 1e3) / 1e9,
                                "lds_peak_GBs": 256 * 128 * 2.4,
                                "tree_levels_per_row": st["mean_levels"]})
    elif pop.propagator.chebyshev:
        # the series runs out of LDS: per row and term two far-neighbour slots read (an unused
        # slot reads a zero) and the new term written, 8 B each; the row's own last two terms and
        # its r +- 1 neighbours' terms come from registers
        terms = st["mean_terms"]
        lds_b = 2.0 * terms * 8.0 * 3.0
        out["roofline"].update({"lds_bytes_per_cmpt_step": lds_b,
                                "lds_GBs": lds_b * N * K / (kern_ms / 1e3) / 1e9,
                                "lds_peak_GBs": 256 * 128 * 2.4,
                                "series_terms_per_row": terms})
    log(f"timed {K} steps: {wall:.3f}s wall, kernel {kern_ms:.1f} ms, {value / 1e9:.3f} G cmpt-steps/s")
    if not args.no_cpu:
        out["cpu_baseline"], v_ref, trace_ref, r_end = cpu_baseline(pop, args, W + K)
        out["precision"] = precision(pop, cells, args, v_ref, trace_ref, r_end, W + K)
        log(f"precision: {out['precision']}")
    pop_store_free = pop.propagator.chebyshev
    if not args.no_adjoint:
        del pop
        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_rate(store=fmt if pop_store_free else None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
