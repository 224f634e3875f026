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
FP64_PEAK_TFLOPS = 78.6          # MI355X dense FP64 (vector and matrix)
LDS_PEAK_GBS = 256 * 128 * 2.4   # 128 B per clock per CU, 256 CUs, 2.4 GHz
# HBM bytes per compartment-step of the step kernels on this workload, per propagator store, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE counter runs scaled by what the same counters read for a
# plain 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); None where not measured.  The
# profile summary each figure comes from is named beside it.
MEASURED_TRAFFIC_PER_CMPT_STEP = {
    "general": (6896.1, "profiles/k_hh_step_10M_population.md"),
    "symmetric": (3753.3, "profiles/k_hh_step_sym_10M_population.md"),
    "chebyshev": (116.5, "profiles/k_hh_step2_cheb_10M_population.md"),
    "rational": (None, None)}
# Kept entries per row of the propagator pruned at 1e-9 (pair rule), this seeded 10M workload:
# the canonical algorithmic byte count of the pruned-matrix algorithm, reported beside the
# matrix-free stores' own roofline
NNZ_BAR_PRUNED_1E9 = 567.99
# FP64 flop per row and product of the matrix-free stores: a Chebyshev term is six FMAs per row
# (2 d T, three neighbours, - T_k-2, c_k T_k); a rational pole is a complex elimination sweep
# (one child's term, the pivot product) and a back substitution (the parent's term, the output)
CHEB_FLOP_PER_ROW_TERM = 12.0
RATIONAL_FLOP_PER_ROW_POLE = 24.0
RATIONAL_POLES = 5


def algorithmic_bytes(nnz_bar):
    """Canonical algorithmic bytes per compartment-step of the stored-matrix step: two half steps of
    8 nnz + 28 B (FP32 value + i32 index per kept entry, row pointer, v/g/d/I/(1/C) reads, v write)
    and a 68-B HH mechanism pass.  nnz_bar = kept entries per row (both triangles).  The canonical
    count stays the same whatever the storage format, so a compressed format shows up as an
    effective bandwidth above what it moves."""
    return 16.0 * nnz_bar + 124.0


def product_flop(pop):
    """FP64 flop per compartment-step of a matrix-free store's two products (half steps)."""
    st = pop.propagator.stats
    if pop.propagator.rational:
        return 2.0 * RATIONAL_POLES * RATIONAL_FLOP_PER_ROW_POLE
    return 2.0 * st["mean_terms"] * CHEB_FLOP_PER_ROW_TERM


def format_bytes(fmt, nnz_bar):
    """Bytes per compartment-step the stored format itself needs: per half step the stored entries
    (FP32 value + u16 column; the symmetric format stores the diagonal and one triangle) and per row
    the offset (4 B) and FP64 row sum (8 B) plus, symmetric, the FP64 weight (8 B); per step 68 B of
    per-segment mechanism state and constants."""
    if fmt == "symmetric":
        return 2.0 * (6.0 * (nnz_bar + 1.0) / 2.0 + 20.0) + 68.0
    if fmt == "chebyshev":   # per half step the LDS place, the port's place word and 4 FP32
        return 2.0 * (2.0 + 4.0 + 16.0 + 16.0) + 68.0   # coefficients, FP64 v in/out
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
    ap.add_argument("--cheb-launch", choices=("auto", "default", "half", "step", "two_steps"),
                    default="auto",
                    help="launch form of the Chebyshev store: one launch per half step, per step "
                         "or per two steps, the library's default, or auto (each timed by the "
                         "autotuner before the timed region; the forms give bitwise the same "
                         "results)")
    ap.add_argument("--cpu-sample-cells", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--precision-steps", type=int, default=1000,
                    help="steps of the precision traces (at least 1000)")
    ap.add_argument("--precision-cells", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the gradients/s measurement of the reverse pass")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--kinetic", action="store_true",
                    help="only the 13-state kinetic workload (--instances instances, one step per "
                         "launch), one JSON line")
    ap.add_argument("--instances", type=float, default=1e6)
    ap.add_argument("--species", action="store_true",
                    help="only the Ca2+ diffusion + buffer workload (--species-compartments), one "
                         "JSON line")
    ap.add_argument("--species-compartments", type=float, default=5e6)
    ap.add_argument("--accumulate", action="store_true",
                    help="only time the segmented accumulation (nw_accumulate) on --compartments "
                         "segments with three instances each (hh's na, k, l), one JSON line")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _sample_propagator(pop, sample, rule, prune):
    """FP64 propagator of the first `sample` cells of population `pop` from the device
    construction path, pruned at `prune` by the entry rule, the pair rule of the symmetric store,
    or not at all ("none")."""
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


def _oracle_rule(pop):
    """(rule, prune) of the FP64 operator the oracle runs to check store `pop.propagator`: the
    matrix-free stores apply exp(A dt) unpruned; the stored ones prune at their threshold by the
    entry rule (general) or the pair rule (symmetric)."""
    prop = pop.propagator
    if prop.chebyshev:
        return "none", 0.0
    return ("pair" if prop.symmetric else "entry"), prop.prune


def _oracle_population(pop, cells, P, stim):
    """The oracle's C FP64 restatement of the fused step for the first `cells` cells of `pop`."""
    from oracle import neuwon_oracle as O
    r = int(pop.cell_ptr[cells])
    gbar = {k: pop.g_bar[k][:r] for k in ("na", "k", "l")}
    return O.CHHPopulation(pop.cell_ptr[:cells + 1], P, pop.C[:r], gbar, dt=pop.dt,
                           celsius=pop.celsius, gate_mode=pop.gate_mode, stim=stim)


def _oracle_stim(pop):
    p = pop.params
    if p.stim_mode == 2:
        return ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
                int(p.pulse_ticks))
    if p.stim_mode == 1:
        return ("constant", float(p.stim_amp_a))
    return None


def cpu_baseline(pop, args, total_steps):
    """The oracle's C FP64 restatement of the same step (CSR SpMV, table gates, Poisson stimulus)
    fed the FP64 propagator the device built, pruned at 1e-9 (the reference's algorithm on this
    workload: a stored pruned matrix -- entry rule, or the pair rule for the symmetric store), on
    a bounded sample of the population's cells.  Timed on one thread (the reference's serial
    path) and on every host core, each the best of three repeats (3 untimed warm-up steps, then
    timed steps).  A further all-cores run covers exactly the steps the GPU ran and returns the
    final voltages (the timed run's state check; for the matrix-free stores, which apply
    exp(A dt) unpruned, that check runs the unpruned operator on the first few cells)."""
    cores = len(os.sched_getaffinity(0))
    sample = min(args.cpu_sample_cells, pop.num_cells)
    r_end = int(pop.cell_ptr[sample])
    rule = "pair" if pop.propagator.symmetric else "entry"
    P = _sample_propagator(pop, sample, rule, 1e-9 if pop.propagator.chebyshev
                           else pop.propagator.prune)
    stim = _oracle_stim(pop)
    rates = {}
    for threads, budget in ((1, args.cpu_seconds / 6.0), (cores, args.cpu_seconds / 6.0)):
        best, steps_done = 0.0, 0
        for _ in range(3):
            co = _oracle_population(pop, sample, P, stim)
            co.advance(3, threads=threads)   # warm-up
            steps, t0 = 0, time.perf_counter()
            while True:
                co.advance(1, threads=threads)
                steps += 1
                el = time.perf_counter() - t0
                if (el >= budget and steps >= 20) or steps >= 2000:
                    break
            if r_end * steps / el > best:
                best, steps_done = r_end * steps / el, steps
        rates[threads] = (best, steps_done)
        log(f"cpu baseline {threads} thread(s): {best / 1e6:.2f} M cmpt-steps/s "
            f"(best of 3, {steps_done} steps)")
    base = {"value": rates[cores][0], "unit": "compartment-steps/s", "cores": cores, "kind": "port",
            "value_1_thread": rates[1][0],
            "sample": f"first {sample} cells ({r_end} compartments, {P.nnz} FP64 propagator "
                      f"entries, pruned 1e-9) of the same population; best of 3 repeats of "
                      f">= 20 steps after 3 warm-up steps: {rates[1][1]} steps on 1 thread, "
                      f"{rates[cores][1]} steps on {cores} threads; oracle C FP64 restatement "
                      f"(OpenMP over cells)"}
    # the timed run's state: the GPU's exact step count on the store's own operator
    o_rule, o_prune = _oracle_rule(pop)
    few = sample if o_rule != "none" else min(6, sample)
    P0 = P if o_rule != "none" else _sample_propagator(pop, few, "none", 0.0)
    co = _oracle_population(pop, few, P0, stim)
    co.advance(total_steps, threads=cores)
    return base, co.v.copy(), int(pop.cell_ptr[few])


def _trace_error(v, ref, dt, window_ms=2.0, threshold=20.0, excursion=-20.0):
    """Errors of a voltage trace against the FP64 one: relative over the samples more than
    window_ms from every sample where the reference is at or above `excursion` mV (an action
    potential or a depolarisation approaching one: the voltage crosses 0 mV there and a relative
    error is not defined), absolute (mV) within; spikes are upward crossings of `threshold`."""
    up = np.nonzero((ref[1:] >= threshold) & (ref[:-1] < threshold))[0] + 1
    t = np.arange(len(ref)) * dt
    near = np.zeros(len(ref), dtype=bool)
    for k in np.nonzero(ref >= excursion)[0]:
        near |= np.abs(t - t[k]) <= window_ms
    d = np.abs(v - ref)
    rel = d / np.abs(ref)
    return (len(up), rel[~near].max() if (~near).any() else 0.0,
            d[near].max() if near.any() else 0.0, np.median(rel))


def _trace_precision(pop, cells, steps, stim, args, label):
    """Every root voltage after every step of `cells` (a population of their own, one launch,
    the timed population's store and settings) against the FP64 oracle run on the store's
    operator, beside the same oracle with FP32 storage emulated (the floor FP32 state sets:
    per-segment conductances, 1/C and gate tables in FP32, gates rounded after every update and,
    for the stored formats, voltages after every step); spikes counted on the oracle trace."""
    import torch
    from neuwon_amd import population
    cores = len(os.sched_getaffinity(0))
    sub = population.HHPopulation(cells, dt=pop.dt, celsius=pop.celsius,
                                  prune=pop.propagator.prune, gate_mode=pop.gate_mode,
                                  stim=stim, block_threads=args.block_threads,
                                  store=pop.propagator.store)
    tr = torch.zeros((sub.num_cells, steps), dtype=torch.float32, device=sub.dev)
    sub.advance(steps, trace=tr)
    tr = tr.cpu().numpy().astype(np.float64)
    v_gpu = sub.v.cpu().numpy().astype(np.float64)
    rule, prune = _oracle_rule(sub)
    P = _sample_propagator(sub, sub.num_cells, rule, prune)
    roots = sub.cell_ptr[:-1]
    runs = []
    for emulate in (False, True):
        from oracle import neuwon_oracle as O
        r = int(sub.cell_ptr[-1])
        co = O.CHHPopulation(sub.cell_ptr, P, sub.C[:r], {k: sub.g_bar[k][:r] for k in sub.g_bar},
                             dt=sub.dt, celsius=sub.celsius, gate_mode=sub.gate_mode,
                             stim=_oracle_stim(sub), fp32_storage=emulate,
                             round_v=emulate and not sub.propagator.chebyshev)
        ref = np.empty((sub.num_cells, steps))
        for k in range(steps):
            co.advance(1, threads=cores)
            ref[:, k] = co.v[roots]
        runs.append((ref, co.v.copy()))
    (ref, v_ref), (flo, v_flo) = runs
    errs = [_trace_error(tr[c], ref[c], sub.dt) for c in range(sub.num_cells)]
    ferr = [_trace_error(flo[c], ref[c], sub.dt) for c in range(sub.num_cells)]
    d, df = np.abs(v_gpu - v_ref), np.abs(v_flo - v_ref)
    sizes = np.diff(sub.cell_ptr)
    within = all(e[1] < max(1e-5, 3.0 * f[1]) and e[2] < max(2e-3, 3.0 * f[2])
                 for e, f in zip(errs, ferr))
    return {"sample": f"{label}: {sub.num_cells} cells of {int(sizes.min())}-{int(sizes.max())} "
                      f"compartments ({sub.num_segments} in all), {steps} steps",
            "spikes": int(sum(e[0] for e in errs)),
            "max_rel_v_error_subthreshold": float(max(e[1] for e in errs)),
            "max_abs_v_error_near_excursions_mV": float(max(e[2] for e in errs)),
            "median_rel_v_error": float(np.median([e[3] for e in errs])),
            "state_max_abs_v_error_mV": float(d.max()),
            "state_max_rel_v_error": float(np.max(d / np.abs(v_ref))),
            "per_cell_rel_subthreshold": [float("%.2e" % e[1]) for e in errs],
            "fp32_storage_floor": {
                "per_cell_rel_subthreshold": [float("%.2e" % f[1]) for f in ferr],
                "max_abs_v_error_near_excursions_mV": float(max(f[2] for f in ferr)),
                "state_max_abs_v_error_mV": float(df.max())},
            "within_1e-5_or_3x_floor_per_cell": bool(within)}


def precision(pop, cells, args, v_ref, r_end, total_steps):
    """Voltage error of the GPU against the FP64 oracle on the CPU:
    (1) the timed population's state on the sampled cells after all its steps;
    (2) the workload's first cells with its Poisson stimulus over >= 1000 steps, every root
        voltage after every step, split into subthreshold samples (relative error) and samples
        within 2 ms of a spike (absolute error, mV);
    (3) the same over >= 1000 steps on the smallest of the workload's first 200 cells driven by
        the single-neuron workload's 0.5 nA constant clamp at every root, so that every cell
        spikes repeatedly."""
    v_gpu = pop.v.cpu().numpy().astype(np.float64)[:r_end]
    d = np.abs(v_gpu - v_ref)
    steps = max(1000, args.precision_steps)
    few = min(args.precision_cells, pop.num_cells)
    out = {"timed_run": {"state_max_abs_v_error_mV": float(d.max()),
                         "state_median_rel_v_error": float(np.median(d / np.abs(v_ref))),
                         "sample": f"first {int(np.searchsorted(pop.cell_ptr, r_end))} cells, "
                                   f"{r_end} compartments, after {total_steps} steps"}}
    out["workload_poisson"] = _trace_precision(pop, cells[:few], steps, pop.stim, args,
                                               "first cells of the workload, Poisson stimulus")
    log(f"precision, workload cells: {out['workload_poisson']}")
    head = cells[:min(200, len(cells))]
    small = sorted(range(len(head)), key=lambda c: len(head[c]))[:few]
    out["clamp_0.5nA"] = _trace_precision(pop, [head[c] for c in sorted(small)], steps,
                                          {"mode": "constant", "amp_na": 0.5}, args,
                                          "smallest of the first 200 cells, 0.5 nA clamp")
    log(f"precision, clamped cells: {out['clamp_0.5nA']}")
    out["bar"] = {"subthreshold_rel": 1e-5,
                  "met_at_1e-5": all(out[k]["max_rel_v_error_subthreshold"] < 1e-5
                                     for k in ("workload_poisson", "clamp_0.5nA")),
                  "met_at_1e-5_or_3x_fp32_floor": all(out[k]["within_1e-5_or_3x_floor_per_cell"]
                                                      for k in ("workload_poisson", "clamp_0.5nA"))}
    out["definition"] = ("FP64 oracle (C restatement) with the store's operator: unpruned exp(A dt) "
                         "for the matrix-free stores, pruned as stored otherwise; the same tables and "
                         "stimulus; subthreshold = more than 2 ms from every sample where the "
                         "oracle's voltage is at or above -20 mV; fp32_storage_floor = the same "
                         "oracle with FP32 storage emulated, against the FP64 one")
    return out


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
    # FP64 roofline of the series work: per gradient the forward pass, the recomputation of
    # every checkpointed segment and the reverse pass each apply two products per step (the
    # whole-record gradient skips the recomputation); the gate and mechanism arithmetic around
    # the products is not counted
    out = {"gradients_per_s": 1.0 / el, "seconds_per_gradient": el,
           "gradients_per_s_whole_record": 1.0 / el_rec,
           "store": prob.forward_store}
    if pop.propagator.chebyshev:
        per_pass = pop.num_segments * T * product_flop(pop)
        tf = 3 * per_pass / el / 1e12
        tf_rec = 2 * per_pass / el_rec / 1e12
        out["roofline"] = {"bound": "fp64", "achieved": tf, "peak": FP64_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS, "traffic": None,
                           "flop_per_gradient": 3 * per_pass,
                           "definition": "FP64 flop of the products (three passes of two per step: "
                                         "forward, recomputation, reverse) per second of wall time "
                                         "per gradient",
                           "whole_record": {"achieved": tf_rec, "frac": tf_rec / FP64_PEAK_TFLOPS,
                                            "flop_per_gradient": 2 * per_pass}}
    out["config"] = ("1024 cells x 500 compartments, 2000 steps at dt 0.05 ms, checkpoints every "
                     "100 steps; gradients of sum (v_soma - target)^2 w.r.t. per-segment gNa, gK, "
                     "gL and 6 HH rate multipliers")
    return out


KINETIC_BYTES_PER_INSTANCE_STEP = 4.0 + 2 * 13 * 4.0   # v read, 13 FP32 states read and written
# HBM bytes per instance-step of k_lti_grouped<13, 1> on the coherent protocol, from FETCH_SIZE and
# WRITE_SIZE runs (scaled as for the step kernels)
# the model's call adds the segment index (4 B), gmax (4 B) and the conductance output (4 B)
KINETIC_MODEL_BYTES_PER_INSTANCE_STEP = KINETIC_BYTES_PER_INSTANCE_STEP + 12.0
# measured HBM bytes per instance-step of that call (k_lti_grouped<13, 1, true>)
MEASURED_KINETIC_MODEL_TRAFFIC = (131.0, "profiles/k_lti_grouped_1M_13state.md")
MEASURED_KINETIC_TRAFFIC = (113.4, "profiles/k_lti_grouped_1M_13state.md")


def _kinetic_protocol(torch, name, v0, steps, dev):
    """Voltages of every step [steps, M] (FP32, device): "coherent" -- every instance's own level
    plus a common 10 mV oscillation of period 20 ms (instances in order of their levels, as a
    model orders instances by segment); "random" -- every voltage drawn anew each step, uniform
    over -80..40 mV (no two neighbouring instances share a bucket: the worst case)."""
    M = v0.numel()
    if name == "coherent":
        k = torch.arange(steps, device=dev, dtype=torch.float64)
        wave = (10.0 * torch.sin(2.0 * math.pi * k * 0.1 / 20.0)).float()
        return v0[None, :] + wave[:, None], wave.double().cpu().numpy()
    gen = torch.Generator(device=dev).manual_seed(5)
    return torch.rand((steps, M), generator=gen, device=dev) * 120.0 - 80.0, None


def kinetic_bench(args):
    """The 13-state kinetic workload: 1M instances of a voltage-dependent Markov scheme on the
    AMPA13 graph (rates a exp(b v), seeded), 2001-node FP32 tables over -100..100 mV, linear
    interpolation, one step per launch (the path the model runs) with the voltages changing every
    step; states initialised to the steady state at the node below each instance's level."""
    import torch
    from neuwon_amd import channels, kinetic
    from oracle import neuwon_oracle as O
    dev = torch.device("cuda", 0)
    M, K, W = int(args.instances), args.steps, args.warmup
    sc = channels.random_ampa13_scheme(seed=2, time_step=0.1)
    node32 = sc.node_table().astype(np.float32)
    nodes = sc.grid.nodes()
    ss = np.stack([sc.initial_state(x) for x in nodes])
    rng = np.random.default_rng(2)
    v0 = np.sort(rng.uniform(-80.0, 40.0, M))
    b0 = np.clip(np.searchsorted(nodes, v0) - 1, 0, len(nodes) - 1)
    st0 = ss[b0]
    st0 = (st0 / st0.sum(axis=1, keepdims=True)).T.astype(np.float32)
    tab = kinetic.LTITable(torch.as_tensor(node32, device=dev), -100.0, 100.0)
    v0d = torch.as_tensor(v0.astype(np.float32), device=dev)
    stream = torch.cuda.current_stream()
    # the model's form of the call: instances on every other segment of a segment table twice
    # their number (read through a segment index), and the conductance output
    # cond_out[i] = gmax[i] * sum_s weights[s] state_s[i]
    seg = torch.arange(M, device=dev, dtype=torch.int32) * 2
    weights = torch.zeros(13, device=dev, dtype=torch.float32)
    weights[[11, 12]] = 1.0
    gmax = torch.full((M,), 1e-3, device=dev, dtype=torch.float32)
    cond = torch.empty(M, device=dev, dtype=torch.float32)
    results = {}
    for proto in ("coherent", "random", "model_path"):
        V, wave = _kinetic_protocol(torch, "random" if proto == "random" else "coherent", v0d,
                                    W + K, dev)
        kw = {"conserve_sum": 1.0}
        if proto == "model_path":
            Vs = torch.full((W + K, 2 * M), -65.0, device=dev, dtype=torch.float32)
            Vs[:, 0::2] = V
            V = Vs
            kw.update(seg_index=seg, weights=weights, gmax=gmax, cond_out=cond)
        st = torch.as_tensor(st0.copy(), device=dev)
        for k in range(W):
            tab.advance(st, V[k], **kw)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for k in range(W, W + K):
            tab.advance(st, V[k], **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1)
        # distinct buckets per wavefront of 64 instances, at the last step
        vl = V[W + K - 1][0::2] if proto == "model_path" else V[W + K - 1]
        bk = torch.clamp(((vl.double() + 100.0) * 10.0).trunc(), 0, 1999).long()
        bk = bk[:M // 64 * 64].view(-1, 64).cpu().numpy()
        distinct = 1 + (np.diff(np.sort(bk, axis=1), axis=1) != 0).sum(axis=1)
        results[proto] = {"wall_s": wall, "kernel_ms": ms, "distinct": distinct,
                          "state": st.cpu().numpy() if proto != "random" else None,
                          "cond": cond.cpu().numpy() if proto == "model_path" else None,
                          "wave": wave}
        log(f"kinetic {proto}: {K} launches, {ms / K * 1e3:.1f} us per launch, "
            f"{M * K / (ms * 1e-3) / 1e9:.2f} G instance-steps/s; distinct buckets per wavefront "
            f"mean {distinct.mean():.2f}, max {distinct.max()}")
        del V, st
        torch.cuda.empty_cache()
    co = results["coherent"]
    avg = co["kernel_ms"] / K * 1e-3
    achieved = KINETIC_BYTES_PER_INSTANCE_STEP * M / avg / 1e9
    traffic, traffic_src = MEASURED_KINETIC_TRAFFIC
    rnd = results["random"]
    out = {"metric": "instance-steps/s, 13-state kinetic scheme, 1M instances, one step per launch",
           "value": M * K / co["wall_s"], "unit": "instance-steps/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": co["wall_s"] * 1e3 / K, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 state, f64 arithmetic",
           "data": "synthetic (seeded rates and voltages)",
           "config": {"workload": "13-state voltage-dependent Markov scheme on the AMPA13 graph, "
                                  "2001-node FP32 table over -100..100 mV, linear interpolation; "
                                  "voltages: every instance's level U(-80, 40) mV (instances in "
                                  "order of level) plus a common 10 mV, 20 ms oscillation",
                      "instances": M, "states": 13, "table_nodes": len(nodes),
                      "kernel_launches_per_step": 1,
                      "distinct_buckets_per_wavefront_mean": float(co["distinct"].mean()),
                      "group_max": 2},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None if traffic is None else traffic * M,
                        "traffic_source": traffic_src,
                        "kernel": "nw::k_lti_grouped<13, 1>", "avg_launch_ms": avg * 1e3,
                        "algorithmic_bytes_per_instance_step": KINETIC_BYTES_PER_INSTANCE_STEP},
           "random_voltages": {
               "value": M * K / rnd["wall_s"], "avg_launch_ms": rnd["kernel_ms"] / K,
               "roofline_frac": KINETIC_BYTES_PER_INSTANCE_STEP * M / (rnd["kernel_ms"] / K * 1e-3)
                                / 1e9 / HBM_PEAK_GBS,
               "distinct_buckets_per_wavefront_mean": float(rnd["distinct"].mean()),
               "note": "every voltage redrawn each step: no two instances of a wavefront share a "
                       "bucket, every wavefront gathers its own node matrices"}}
    mp = results["model_path"]
    mp_avg = mp["kernel_ms"] / K * 1e-3
    # the same voltages as the coherent leg: the states must agree bit for bit; the conductance
    # output against gmax * (state_11 + state_12) formed in FP64 from the FP32 states
    st_mp = mp["state"]
    ref_cond = float(np.float32(1e-3)) * (st_mp[11].astype(np.float64)
                                          + st_mp[12].astype(np.float64))
    out["model_path"] = {
        "value": M * K / mp["wall_s"], "avg_launch_ms": mp["kernel_ms"] / K,
        "achieved_GBs": KINETIC_MODEL_BYTES_PER_INSTANCE_STEP * M / mp_avg / 1e9,
        "roofline_frac": KINETIC_MODEL_BYTES_PER_INSTANCE_STEP * M / mp_avg / 1e9 / HBM_PEAK_GBS,
        "algorithmic_bytes_per_instance_step": KINETIC_MODEL_BYTES_PER_INSTANCE_STEP,
        "traffic": MEASURED_KINETIC_MODEL_TRAFFIC[0] * M,
        "traffic_source": MEASURED_KINETIC_MODEL_TRAFFIC[1],
        "kernel": "nw::k_lti_grouped<13, 1, true>",
        "states_bitwise_equal_to_coherent": bool(np.array_equal(st_mp, co["state"])),
        "cond_out_max_rel_err": float(np.max(np.abs(mp["cond"] - ref_cond)
                                             / np.maximum(np.abs(ref_cond), 1e-30))),
        "note": "the call RxD_Model makes: voltages read through a segment index (instances on "
                "every other segment of a 2M-segment table), gmax and conductance output, the "
                "coherent voltages"}
    log(f"kinetic model path: {mp_avg * 1e6:.1f} us per launch, "
        f"frac {out['model_path']['roofline_frac']:.3f}, states equal to the coherent leg: "
        f"{out['model_path']['states_bitwise_equal_to_coherent']}")
    # parity on a sample of the coherent run against the C FP64 oracle fed the same FP32 table,
    # beside the oracle with its states rounded to FP32 after every step (the storage floor)
    poly = O.node_table_to_polynomial(node32.astype(np.float64))
    sample = np.sort(rng.choice(M, 2000, replace=False))
    vs = v0[sample].astype(np.float32).astype(np.float64)
    wave = co["wave"].astype(np.float32).astype(np.float64)
    ref = st0[:, sample].astype(np.float64)
    emu = ref.copy()
    for k in range(W + K):
        # the device adds the FP32 wave to the FP32 level
        vk = (vs.astype(np.float32) + np.float32(wave[k])).astype(np.float64)
        ref = O.c_lti_run(ref, vk, poly, conserve_sum=1.0)
        emu = O.c_lti_run(emu, vk, poly, conserve_sum=1.0).astype(np.float32).astype(np.float64)
    got = co["state"][:, sample].astype(np.float64)
    big = ref >= 1e-6
    out["precision"] = {
        "max_abs_state_error": float(np.abs(got - ref).max()),
        "max_rel_state_error_occupancy_ge_1e-6": float((np.abs(got - ref) / ref)[big].max()),
        "fp32_storage_floor_abs": float(np.abs(emu - ref).max()),
        "fp32_storage_floor_rel_occupancy_ge_1e-6": float((np.abs(emu - ref) / ref)[big].max()),
        "conservation_max_abs": float(np.abs(co["state"].astype(np.float64).sum(0) - 1.0).max()),
        "sample": f"2000 instances of the coherent run after {W + K} steps",
        "definition": "C FP64 oracle of the reference kernel (codegen.py:39-106) on the same FP32 "
                      "table and voltages; floor: the same oracle with its states rounded to FP32 "
                      "after every step"}
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        n_cpu = min(M, 100_000)
        rates = {}
        for threads in (1, cores):
            best = 0.0
            for _ in range(3):
                stc = st0[:, :n_cpu].astype(np.float64)
                steps = 2 if threads == 1 else 20
                t0 = time.perf_counter()
                O.c_lti_run(stc, v0[:n_cpu], poly, wave=co["wave"][:steps], conserve_sum=1.0,
                            threads=threads)
                best = max(best, n_cpu * steps / (time.perf_counter() - t0))
            rates[threads] = best
        out["cpu_baseline"] = {"value": rates[cores], "unit": "instance-steps/s", "cores": cores,
                               "kind": "port", "value_1_thread": rates[1],
                               "sample": f"{n_cpu} instances of the coherent run, 2 steps on 1 "
                                         f"thread and 20 on {cores}, best of 3; oracle C FP64 "
                                         f"restatement of the reference kernel (OpenMP over "
                                         f"instances)"}
    print(json.dumps(out), flush=True)


# HBM bytes per compartment-step of k_ca_step on the 5M workload (FETCH_SIZE x 2 + WRITE_SIZE)
MEASURED_SPECIES_TRAFFIC = (116.9, "profiles/k_ca_step_5M_species.md")


def species_bench(args):
    """The Ca2+ reaction-diffusion workload: 5M compartments of random-tree cells (seed 3),
    diffusion D = 0.6 um^2/ms with the exact-integration propagator of every cell (FP64 device
    expm, pruned 1e-9, FP32 values), a first-order buffer (kon 0.1 /uM/ms, koff 0.01 /ms, 100 uM),
    Ca 50 nM plus 1 uM pulses in 5 % of the compartments; dt 0.1 ms; all steps in one launch.
    The total Ca (free + bound) is checked in FP64 before and after."""
    import torch
    from neuwon_amd import diffusion, synth
    from oracle import neuwon_oracle as O
    t0 = time.time()
    cells = synth.population(total=int(args.species_compartments), seed=3)
    pop = diffusion.CaPopulation(cells, seed=3, prune=1e-9)
    st = pop.propagator.stats
    N, nnz = pop.num_segments, pop.propagator.mean_nnz_per_row
    log(f"species population ready in {time.time() - t0:.1f}s: {N} compartments, {pop.num_cells} "
        f"cells, nnz/row {nnz:.2f}, propagator build {st['build_seconds']:.1f}s")
    K, W = args.steps, args.warmup
    m0 = pop.mass()
    if W:
        pop.advance(W)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    e0.record(stream)
    pop.advance(K)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    kern_s = e0.elapsed_time(e1) / 1e3
    m1 = pop.mass()
    alg = 2.0 * (8.0 * nnz + 20.0)
    achieved = alg * N * K / kern_s / 1e9
    traffic, traffic_src = MEASURED_SPECIES_TRAFFIC
    out = {"metric": "compartment-steps/s, Ca2+ diffusion + buffer, 5M compartments",
           "value": N * K / wall, "unit": "compartment-steps/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": wall * 1e3 / K, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 state, f64 arithmetic",
           "data": "synthetic (seeded random-tree cells and Ca pulses)",
           "config": {"workload": "intracellular Ca2+ diffusion with a first-order buffer, exact "
                                  "integration propagator pruned 1e-9, all steps in one launch",
                      "compartments": N, "cells": pop.num_cells, "dt_ms": pop.dt,
                      "nnz_per_row": round(nnz, 3), "propagator_build_s": round(st["build_seconds"], 1),
                      "kernel_launches": 1},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None if traffic is None else traffic * N * K,
                        "traffic_source": traffic_src, "kernel": "nw::k_ca_step",
                        "avg_launch_ms": kern_s * 1e3,
                        "algorithmic_bytes_per_cmpt_step": alg,
                        "definition": "two products per step of 8 nnz + 20 B (FP32 value and i32 "
                                      "index per kept entry; row pointer; Ca and bound buffer read "
                                      "and written)"},
           "mass": {"fp64_total_before": m0, "fp64_total_after": m1,
                    "relative_drift": (m1 - m0) / m0}}
    # precision: the smallest of the first 50 cells against the FP64 oracle, the same steps
    ca = pop.ca.cpu().numpy().astype(np.float64)
    c = min(range(min(50, pop.num_cells)), key=lambda i: pop.cell_ptr[i + 1] - pop.cell_ptr[i])
    r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
    P = O.diffusion_propagator(pop.parent_local[r0:r1].astype(np.int64), pop.k[r0:r1],
                               pop.vol[r0:r1], 0.5 * pop.dt, prune=pop.propagator.prune)
    sim = O.CaSim(P, pop.vol[r0:r1], pop.ca0[r0:r1], pop.cab0[r0:r1], dt=pop.dt)
    for _ in range(W + K):
        sim.advance()
    out["precision"] = {"max_rel_ca_error": float(np.abs(ca[r0:r1] - sim.ca).max() / np.abs(sim.ca).max()),
                        "sample": f"cell {c} ({r1 - r0} compartments) after {W + K} steps",
                        "definition": "FP64 oracle (scipy CSR products) with the same pruned "
                                      "propagator; max |error| / max |Ca| over the cell"}
    if not args.no_cpu:
        import scipy.sparse
        ncell = min(50, pop.num_cells)
        R1 = int(pop.cell_ptr[ncell])
        blocks = [O.diffusion_propagator(pop.parent_local[int(pop.cell_ptr[i]):int(pop.cell_ptr[i + 1])].astype(np.int64),
                                         pop.k[int(pop.cell_ptr[i]):int(pop.cell_ptr[i + 1])],
                                         pop.vol[int(pop.cell_ptr[i]):int(pop.cell_ptr[i + 1])],
                                         0.5 * pop.dt, prune=pop.propagator.prune)
                  for i in range(ncell)]
        Ps = scipy.sparse.block_diag(blocks, format="csr")
        best, steps_done = 0.0, 0
        for _ in range(3):
            sim = O.CaSim(Ps, pop.vol[:R1], pop.ca0[:R1], pop.cab0[:R1], dt=pop.dt)
            for _ in range(3):
                sim.advance()
            steps, t2 = 0, time.perf_counter()
            while True:
                sim.advance()
                steps += 1
                el = time.perf_counter() - t2
                if (el >= args.cpu_seconds / 6 and steps >= 20) or steps >= 5000:
                    break
            if R1 * steps / el > best:
                best, steps_done = R1 * steps / el, steps
        out["cpu_baseline"] = {"value": best, "unit": "compartment-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first {ncell} cells ({R1} compartments), best of 3 runs "
                                         f"of {steps_done} steps after 3 warm-up steps; oracle FP64 "
                                         f"numpy/scipy restatement (scipy CSR products, one thread)"}
    print(json.dumps(out), flush=True)


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
    if args.kinetic:
        kinetic_bench(args)
        return
    if args.species:
        species_bench(args)
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
                                  block_threads=args.block_threads, store=args.format,
                                  cheb_launch=None if args.cheb_launch == "default" else args.cheb_launch)
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
    # the matrix-free stores replay a captured graph of every call's half-step launches: capture
    # the graph of the timed calls before the timed region
    graphs = pop.prepare(L) if pop.propagator.chebyshev else 0
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
    kern_s = kern_ms / 1e3
    value = N * K / wall
    free = pop.propagator.chebyshev
    # kernel launches, replayed from one graph per advance() call on the matrix-free stores: on
    # the Chebyshev store one per two steps when built for five row groups (a cell's workgroup
    # runs both steps; an odd call's last step alone), one per step when built for one; two per
    # step (one per half step) on the rational store; one per call (all its steps) on the stored
    # formats
    cheb = free and not pop.propagator.rational
    form = {0: "two_steps" if st["row_groups"] > 1 else "step", 1: "half", 2: "step",
            3: "two_steps"}[int(pop.params.cheb_launch)] if cheb else "half"
    fused = form == "two_steps"
    launches = ((sum((k + 1) // 2 for k, _, _ in ev) if fused else K if form == "step" else 2 * K)
                if free else len(ev))
    avg_launch_s = kern_s / launches
    units_per_launch = N * K / launches          # compartment-steps per launch
    b_step = algorithmic_bytes(nnz_bar)
    traffic, traffic_src = MEASURED_TRAFFIC_PER_CMPT_STEP[fmt]
    f_step = format_bytes(fmt, nnz_bar)
    rate = N * K / kern_s                        # compartment-steps per second of kernel time
    workload = ("batched HH population, FP32 state, exp(A dt) applied as a five-pole rational "
                "approximation, each pole an exact solve over the cell's tree (no stored matrix, "
                "unpruned)" if pop.propagator.rational else
                "batched HH population, FP32 state, exp(A dt) applied as a Chebyshev series of "
                "each cell's tree operator (no stored matrix, unpruned)" if free else
                "batched HH population, FP32 state, FP64-built propagator pruned 1e-9")
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
        "config": {"workload": workload, "propagator_format": fmt, "compartments": N,
                   "cells": pop.num_cells, "dt_ms": args.dt, "steps_per_call": L,
                   "calls": len(ev), "kernel_launches": launches,
                   "kernel_launches_per_step": launches / K,
                   "graph_launches_per_call": 1 if graphs else 0,
                   "propagator_build_s": round(st["build_seconds"], 2),
                   "autotune": pop.tuning.get("store"),
                   "launch_form": form, "launch_autotune": pop.tuning.get("cheb_launch")},
    }
    if free:
        # the products run out of LDS and FP64 registers: the binding resource is FP64 issue
        flop = product_flop(pop)
        tflops = flop * rate / 1e12
        roof = {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / FP64_PEAK_TFLOPS,
                "traffic": None if traffic is None else traffic * units_per_launch,
                "traffic_source": traffic_src,
                "kernel": ("nw::k_hh_step2_cheb (one launch per two steps)" if fused else
                           "nw::k_hh_step_cheb (one launch per step)" if form == "step" else
                           "nw::k_hh_half_cheb (one launch per half step)"),
                "avg_launch_ms": avg_launch_s * 1e3,
                "flop_per_cmpt_step": flop,
                "flop_per_launch": flop * units_per_launch}
        if pop.propagator.rational:
            roof["tree_levels_per_row"] = st["mean_levels"]
            # per pole and row: leaves to root a child's q read (16 B, one child per row on
            # average) and q written; root to leaves the row's q and its parent's z read, z written
            lds_b = 2.0 * RATIONAL_POLES * 5 * 16.0
        else:
            roof["series_terms_per_row"] = st["mean_terms"]
            # per row and term the slot's read port (a far neighbour's term, or the zero) and the
            # new term written, 8 B each, and the series coefficient (one broadcast read per wave
            # of 64 rows); the row's own last two terms and its r +- 1 neighbours' terms come from
            # registers or the neighbouring lane
            lds_b = 2.0 * st["mean_terms"] * (8.0 * 2.0 + 8.0 / 64.0)
        roof["lds"] = {"bytes_per_cmpt_step": lds_b, "GBs": lds_b * rate / 1e9,
                       "peak_GBs": LDS_PEAK_GBS, "frac": lds_b * rate / 1e9 / LDS_PEAK_GBS}
        roof["hbm"] = {"store_bytes_per_cmpt_step": f_step,
                       "measured_bytes_per_cmpt_step": traffic,
                       "GBs": None if traffic is None else traffic * rate / 1e9,
                       "frac": None if traffic is None else traffic * rate / 1e9 / HBM_PEAK_GBS}
        roof["canonical_pruned_matrix"] = {
            "nnz_per_row": nnz_bar, "bytes_per_cmpt_step": b_step,
            "effective_GBs": b_step * rate / 1e9,
            "note": "16 nnz + 124 B of the pruned-matrix algorithm at prune 1e-9 (the stored "
                    "formats' nnz of this seeded workload): an effective bandwidth, not HBM traffic"}
    else:
        achieved = b_step * rate / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": None if traffic is None else traffic * units_per_launch,
                "traffic_source": traffic_src,
                "kernel": "nw::k_hh_step (all steps of a call in one launch)",
                "avg_launch_ms": avg_launch_s * 1e3,
                "nnz_per_row": nnz_bar,
                "algorithmic_bytes_per_cmpt_step": b_step,
                "format_bytes_per_cmpt_step": f_step,
                "format_GBs": f_step * rate / 1e9,
                "measured_traffic_GBs": None if traffic is None else traffic * rate / 1e9}
    out["roofline"] = roof
    log(f"timed {K} steps: {wall:.3f}s wall, kernel {kern_ms:.1f} ms, {value / 1e9:.3f} G cmpt-steps/s")
    if not args.no_cpu:
        out["cpu_baseline"], v_ref, r_end = cpu_baseline(pop, args, W + K)
        out["precision"] = precision(pop, cells, args, v_ref, r_end, W + K)
    pop_store_free = pop.propagator.chebyshev
    if not args.no_adjoint:
        del pop
        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_rate(store=fmt if pop_store_free else None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
