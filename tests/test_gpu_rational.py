"""The rational store (no stored matrix: exp(A dt) x as a five-pole rational approximation, each
pole an exact solve over the cell's tree in level order, unpruned) against the FP64 oracle with
the unpruned propagator (the reference's default threshold, electric.py:114-115): a single
neuron, cells across the launch shapes, a Poisson-driven population beside the Chebyshev store,
and a model whose soma has six children."""
import os
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

import fp32_floor
from test_gpu_parity import _c0_cell, _oracle_cell, voltage_error

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-rational] {msg}", flush=True)
    sys.stdout.flush()


def _unpruned(pop):
    import scipy.sparse
    from neuwon_amd import population
    blocks = []
    for c in range(pop.num_cells):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        blocks.append(population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                                     pop.C[r0:r1], 0.5 * pop.dt, 0.0))
    return scipy.sparse.block_diag(blocks, format="csr")


def _oracle_population(pop, stim):
    return O.CHHPopulation(pop.cell_ptr, _unpruned(pop), pop.C, pop.g_bar, dt=pop.dt,
                           celsius=pop.celsius, gate_mode="table", stim=stim)


@pytest.mark.parametrize("gate_mode", ["table", "exact"])
def test_rational_single_neuron_trajectory(torch_gpu, gate_mode):
    """1000-compartment tree, HH, 0.5 nA step at the root, dt 0.1 ms x 1000 steps."""
    from neuwon_amd import population, channels
    torch = torch_gpu
    steps = 1000
    cells = _c0_cell()
    pop = population.HHPopulation(cells, gate_mode=gate_mode, store="rational",
                                  stim={"mode": "constant", "amp_na": 0.5})
    st = pop.propagator.stats
    assert st["rational"] == 1 and st["chebyshev"] == 1 and st["stored"] == 0
    trace = torch.zeros((1, steps), dtype=torch.float32, device=pop.dev)
    pop.advance(steps, trace=trace)
    torch.cuda.synchronize()
    v_gpu = pop.v.cpu().numpy().astype(np.float64)
    tr_gpu = trace.cpu().numpy()[0].astype(np.float64)
    g, R, C, P = _oracle_cell(cells[0])            # prune = machine epsilon
    gbar = {k: channels.HH_GBAR[k] * 1e-8 * g["surface_area"] for k in ("na", "k", "l")}
    stim = np.zeros(len(cells[0]))
    stim[0] = 0.5e-9
    sim = O.HHSim(P, C, gbar, v0=-70.0, dt=0.1, celsius=6.3, gate_mode=gate_mode,
                  stim=lambda t: stim)
    tr = np.empty(steps)
    for k in range(steps):
        sim.advance()
        tr[k] = sim.v[0]
    err = voltage_error(tr_gpu, tr, dt=0.1)
    rel_v = np.abs(v_gpu - sim.v) / np.abs(sim.v)
    _say(f"{gate_mode}: {st['max_levels']} levels, build {st['build_seconds'] * 1e3:.1f} ms; root "
         f"trace {err}, all segments at the end: max rel {rel_v.max():.3e}")
    assert err["spikes"] >= 1
    assert err["rel_subthreshold"] < 1e-5
    assert err["abs_spiking_mV"] < 2e-3
    assert rel_v.max() < 1e-5
    for gname, arr in (("m", pop.m), ("h", pop.h), ("n", pop.n)):
        assert np.abs(arr.cpu().numpy().astype(np.float64) - sim.gates[gname]).max() < 1e-5


@pytest.mark.parametrize("sizes,threads", [
    ([1, 2, 3, 63, 64, 65, 1023, 1024, 1025, 2049, 5000], 1024),
    ([1, 2, 63, 64, 65, 300, 512], 512)])
def test_rational_cell_sizes(torch_gpu, sizes, threads):
    """Cells from 1 to 5000 segments (one to five row groups per thread, levels wider than a
    wave) and, alone, cells up to 512 segments (512-thread workgroups); 0.5 nA constant current
    at every root, 200 steps: root traces and every segment at the end against the FP64
    oracle."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    rng = np.random.default_rng(31)
    cells = [synth.random_cell(n, rng) for n in sizes]
    pop = population.HHPopulation(cells, store="rational", stim={"mode": "constant", "amp_na": 0.5})
    st = pop.propagator.stats
    assert st["workgroup_threads"] == threads
    steps = 200
    trace = torch.zeros((pop.num_cells, steps), dtype=torch.float32, device=pop.dev)
    pop.advance(steps, trace=trace)
    torch.cuda.synchronize()
    tr = trace.cpu().numpy().astype(np.float64)
    v = pop.v.cpu().numpy().astype(np.float64)
    (ref, ref_v), (flo, flo_v) = fp32_floor.oracle_traces(pop, _unpruned(pop),
                                                          ("constant", 0.5e-9), steps)
    heads = np.full(pop.num_cells, pop.v0)
    errs = fp32_floor.check_cells(tr, ref, flo, 0.1, say=_say, heads=heads)
    _say(f"sizes {sizes} ({st['max_levels']} levels at most): end max |dv| "
         f"{np.abs(v - ref_v).max():.2e} mV (floor {np.abs(flo_v - ref_v).max():.2e})")
    assert np.array_equal(pop.ap_count.cpu().numpy(), [e["spikes"] for e, _ in errs])
    assert np.abs(v - ref_v).max() < max(2e-3, 3.0 * np.abs(flo_v - ref_v).max())


def test_rational_population_against_oracle_and_chebyshev(torch_gpu):
    """Poisson-driven population (20k compartments, cells of 200-5000 segments, 600 steps in three
    launches) against the FP64 C oracle with the unpruned propagator and beside the Chebyshev
    store: the same spikes; per cell, both stores' root traces within 1e-5 relative below
    excursions and 2e-3 mV near them, or three times the error FP32 storage alone gives the
    oracle on that cell."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    cells = synth.population(total=20_000, seed=11, size_range=(200, 5000))
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 5}
    runs = {}
    for store in ("rational", "chebyshev"):
        pop = population.HHPopulation(cells, stim=stim, store=store)
        parts = []
        for _ in range(3):
            trace = torch.zeros((pop.num_cells, 200), dtype=torch.float32, device=pop.dev)
            pop.advance(200, trace=trace)
            parts.append(trace)
        torch.cuda.synchronize()
        runs[store] = (torch.cat(parts, dim=1).cpu().numpy().astype(np.float64),
                       pop.ap_count.cpu().numpy().copy(), pop.v.cpu().numpy().copy(),
                       pop.propagator.stats)
    tr, ap, v_end, st = runs["rational"]
    p = pop.params
    stim_t = ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
              int(p.pulse_ticks))
    (ref, ref_v), (flo, flo_v) = fp32_floor.oracle_traces(pop, _unpruned(pop), stim_t, 600)
    _say(f"{pop.num_cells} cells, {st['mean_levels']:.1f} levels per row (max {st['max_levels']}), "
         f"build {st['build_seconds'] * 1e3:.0f} ms: spikes {ap.sum()}; rational vs Chebyshev at "
         f"the end {np.abs(v_end - runs['chebyshev'][2]).max():.2e} mV")
    errs = fp32_floor.check_cells(tr, ref, flo, 0.1, say=_say)
    fp32_floor.check_cells(runs["chebyshev"][0], ref, flo, 0.1)
    assert ap.sum() > 0 and ap.sum() == sum(e["spikes"] for e, _ in errs)
    assert np.array_equal(ap, runs["chebyshev"][1])
    assert np.abs(ref_v - v_end).max() < max(2e-3, 3.0 * np.abs(flo_v - ref_v).max())


def test_rational_model_with_a_many_branched_soma(torch_gpu):
    """RxD_Model(propagator_store='rational'): a soma with six dendrites (six children, beyond the
    two a row keeps in registers) and hh everywhere, 2 nA for 1 ms; the soma trace and every
    segment at the end against the general store built without pruning."""
    from neuwon_amd.model import RxD_Model, TimeSeries

    def run(store):
        m = RxD_Model(time_step=0.025, temperature=6.3, mechanisms=["hh"],
                      propagator_store=store, prune=0.0)
        root = m.Neuron([0, 0, 0], 14).root
        segs = [root]
        for k in range(6):
            a = 2 * np.pi * k / 6
            segs += root.add_section([120 * np.cos(a), 120 * np.sin(a), 0], 1.5,
                                     maximum_segment_length=20)
        for s in segs:
            s.insert("hh")
        probe = TimeSeries().record(root, "voltage")
        root.inject_current(2.0, 1.0)
        for _ in range(400):
            m.advance()
        return np.array(probe.get_data()), m.database.get_data("Segment.voltage").copy()
    t_rat, v_rat = run("rational")
    t_gen, v_gen = run("general")
    err = voltage_error(t_rat, t_gen, dt=0.025)
    _say(f"soma with 6 dendrites: peak {t_rat.max():.1f} mV, trace {err}, final segments max "
         f"|dv| {np.abs(v_rat - v_gen).max():.2e} mV")
    assert t_rat.max() > 20
    assert err["rel_subthreshold"] < 1e-5 and err["abs_spiking_mV"] < 2e-3
    assert np.abs(v_rat - v_gen).max() < 2e-3


def test_rational_refusals(torch_gpu):
    """Cells above 5120 segments and the Chebyshev coefficient query are refused with a message;
    the species step needs a stored matrix."""
    import ctypes
    from neuwon_amd import _lib, population, synth
    pop = population.HHPopulation(synth.single_cell(300, seed=2), store="rational")
    with pytest.raises(RuntimeError, match="Chebyshev store"):
        pop.propagator.cell_chebyshev(0)
    with pytest.raises(RuntimeError, match="rational store"):
        population.HHPopulation(synth.single_cell(6000, seed=2), store="rational")
    p = _lib.NwCaParams()
    p.dt_ms = 0.1
    rc = _lib.load().nw_ca_step(pop.propagator.handle, ctypes.byref(p), 1, _lib.ptr(pop.v),
                                _lib.ptr(pop.m), _lib.ptr(pop.h), None)
    assert rc != 0 and b"general format" in _lib.load().nw_last_error()


def test_rational_deep_cable_and_level_limit(torch_gpu):
    """An unbranched cable of 1,500 segments (1,500 levels, one row per level: every level step
    is a single row) with 0.5 nA at the root, 150 steps (a spike travels down the cable), every
    segment against the FP64 oracle beside the Chebyshev store: within 1e-5 relative or three
    times the Chebyshev store's own error (the FP32-storage floor at the spike's front), and
    within 2e-2 mV; a cable deeper than the 4,095 levels the store holds is refused."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    rng = np.random.default_rng(5)

    def cable(n):
        parent = np.arange(-1, n - 1, dtype=np.int32)
        return synth.CellSpec(parent, rng.uniform(10.0, 30.0, n),
                              np.exp(rng.uniform(np.log(0.5), np.log(4.0), n)))
    cells = [cable(1500)]
    steps = 150
    v = {}
    for store in ("rational", "chebyshev"):
        pop = population.HHPopulation(cells, store=store, stim={"mode": "constant", "amp_na": 0.5})
        if store == "rational":
            assert pop.propagator.stats["max_levels"] == 1500
        pop.advance(steps)
        torch.cuda.synchronize()
        v[store] = pop.v.cpu().numpy().astype(np.float64)
    co = _oracle_population(pop, ("constant", 0.5e-9))
    co.advance(steps, threads=len(os.sched_getaffinity(0)))
    rel = {k: np.abs(x - co.v) / np.abs(co.v) for k, x in v.items()}
    i = int(np.argmax(rel["rational"]))
    _say(f"cable of 1500 segments: max rel rational {rel['rational'].max():.2e} (segment {i}, "
         f"{co.v[i]:.2f} mV), Chebyshev {rel['chebyshev'].max():.2e}; median rational "
         f"{np.median(rel['rational']):.1e}; max |dv| {np.abs(v['rational'] - co.v).max():.2e} mV")
    assert rel["rational"].max() < max(1e-5, 3.0 * rel["chebyshev"].max())
    assert np.abs(v["rational"] - co.v).max() < 2e-2
    assert np.median(rel["rational"]) < 1e-6
    with pytest.raises(RuntimeError, match="levels"):
        population.HHPopulation([cable(4200)], store="rational")
