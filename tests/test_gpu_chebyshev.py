"""The Chebyshev propagator store (no stored matrix: products are Chebyshev series of each cell's
tree operator, unpruned) against the FP64 oracle with the unpruned propagator (the reference's
default threshold, electric.py:114-115): a single neuron, a Poisson-driven population, a model
whose soma has more neighbours than a row keeps in registers, and the entry points that refuse
the store."""
import os
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

import fp32_floor
from test_gpu_parity import _c0_cell, _oracle_cell, voltage_error

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-cheb] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("gate_mode", ["table", "exact"])
def test_chebyshev_single_neuron_trajectory(torch_gpu, gate_mode):
    """1000-compartment tree, HH, 0.5 nA step at the root, dt 0.1 ms x 1000 steps."""
    from neuwon_amd import population, channels
    torch = torch_gpu
    steps = 1000
    cells = _c0_cell()
    pop = population.HHPopulation(cells, gate_mode=gate_mode, store="chebyshev",
                                  stim={"mode": "constant", "amp_na": 0.5})
    st = pop.propagator.stats
    rho, coef = pop.propagator.cell_chebyshev(0)
    # rho is the spectral radius of B = A dt: the symmetric similar matrix's extreme eigenvalue
    import scipy.linalg
    g0, R0, C0, _ = _oracle_cell(cells[0])
    A = O.coefficient_matrix(np.asarray(cells[0].parent, dtype=np.int64), R0, C0, 0.05).toarray()
    w = np.sqrt(C0)
    S = (A * w[:, None]) / w[None, :]
    lam = -scipy.linalg.eigvalsh(0.5 * (S + S.T)).min()
    _say(f"rho {rho:.2f} (largest eigenvalue {lam:.2f}, Gershgorin {2 * (-np.diag(A)).max():.0f}), "
         f"{len(coef) - 1} terms, build {st['build_seconds'] * 1e3:.1f} ms")
    assert lam <= rho <= lam * (1.0 + 1e-6)
    assert st["chebyshev"] == 1 and st["stored"] == 0 and abs(coef.sum() - 1.0) < 1e-14
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
    _say(f"{gate_mode}: root trace {err}, all segments at the end: max rel {rel_v.max():.3e}")
    assert err["spikes"] >= 1
    assert err["rel_subthreshold"] < 1e-5
    assert err["abs_spiking_mV"] < 2e-3
    assert rel_v.max() < 1e-5
    for gname, arr in (("m", pop.m), ("h", pop.h), ("n", pop.n)):
        assert np.abs(arr.cpu().numpy().astype(np.float64) - sim.gates[gname]).max() < 1e-5


def test_chebyshev_population_against_oracle(torch_gpu):
    """Poisson-driven population (20k compartments, cells of 200-5000 segments, 600 steps in three
    launches) against the FP64 C oracle with the unpruned propagator, beside the general store
    built without pruning: the same spikes; per cell, the root's subthreshold voltages within
    1e-5 relative and within 2e-3 mV near excursions, or within three times the error that FP32
    storage alone gives the oracle on that cell (a spiking cell: 1.7e-5 relative)."""
    import scipy.sparse
    from neuwon_amd import population, synth
    torch = torch_gpu
    cells = synth.population(total=20_000, seed=11, size_range=(200, 5000))
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 5}
    runs = {}
    for store in ("chebyshev", "general"):
        pop = population.HHPopulation(cells, stim=stim, store=store, prune=0.0)
        parts = []
        for _ in range(3):
            trace = torch.zeros((pop.num_cells, 200), dtype=torch.float32, device=pop.dev)
            pop.advance(200, trace=trace)
            parts.append(trace)
        torch.cuda.synchronize()
        runs[store] = (torch.cat(parts, dim=1).cpu().numpy().astype(np.float64),
                       pop.ap_count.cpu().numpy().copy(), pop.v.cpu().numpy().copy(),
                       pop.propagator.stats)
    tr, ap, v_end, st = runs["chebyshev"]
    blocks = []
    for c in range(pop.num_cells):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        blocks.append(population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                                     pop.C[r0:r1], 0.05, 0.0))
    P = scipy.sparse.block_diag(blocks, format="csr")
    p = pop.params
    stim_t = ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
              int(p.pulse_ticks))
    (ref, ref_v), (flo, flo_v) = fp32_floor.oracle_traces(pop, P, stim_t, 600)
    _say(f"{pop.num_cells} cells, {st['mean_terms']:.1f} series terms per row and product, build "
         f"{st['build_seconds'] * 1e3:.0f} ms: spikes {ap.sum()} (oracle "
         f"{sum(voltage_error(tr[c], ref[c], dt=0.1)['spikes'] for c in range(pop.num_cells))})")
    errs = fp32_floor.check_cells(tr, ref, flo, 0.1, say=_say)
    gen = fp32_floor.check_cells(runs["general"][0], ref, flo, 0.1)
    assert ap.sum() > 0 and ap.sum() == sum(e["spikes"] for e, _ in errs)
    assert np.array_equal(ap, runs["general"][1])
    floor_end = np.abs(flo_v - ref_v).max()
    _say(f"every segment at the end: {np.abs(ref_v - v_end).max():.2e} mV (floor {floor_end:.2e})")
    assert np.abs(ref_v - v_end).max() < max(2e-3, 3.0 * floor_end)


@pytest.mark.parametrize("steps", [200, 201])
def test_chebyshev_cell_sizes_across_row_groups(torch_gpu, steps):
    """Cells of 1, 2, 3, 63, 64, 65, 1023, 1024, 1025 and 2049 segments in one population (one,
    two and three row groups per thread, rows on either side of the wave and workgroup
    boundaries, layouts with empty slots), 0.5 nA constant current at every root, 200 steps (two
    per launch) or 201 (the last one alone): every segment against the FP64 oracle with the
    unpruned propagator."""
    import scipy.sparse
    from neuwon_amd import population, synth
    torch = torch_gpu
    rng = np.random.default_rng(21)
    sizes = [1, 2, 3, 63, 64, 65, 1023, 1024, 1025, 2049]
    cells = [synth.random_cell(n, rng) for n in sizes]
    pop = population.HHPopulation(cells, store="chebyshev", stim={"mode": "constant", "amp_na": 0.5})
    pop.advance(steps)
    torch.cuda.synchronize()
    v = pop.v.cpu().numpy().astype(np.float64)
    blocks = []
    for c in range(pop.num_cells):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        blocks.append(population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                                     pop.C[r0:r1], 0.05, 0.0))
    P = scipy.sparse.block_diag(blocks, format="csr")
    co = O.CHHPopulation(pop.cell_ptr, P, pop.C, pop.g_bar, dt=pop.dt, celsius=pop.celsius,
                         gate_mode="table", stim=("constant", 0.5e-9))
    co.advance(steps, threads=len(os.sched_getaffinity(0)))
    rel = np.abs(v - co.v) / np.abs(co.v)
    per = [float("%.1e" % rel[int(pop.cell_ptr[c]):int(pop.cell_ptr[c + 1])].max())
           for c in range(pop.num_cells)]
    _say(f"sizes {sizes}: max rel per cell {per}, spikes {pop.ap_count.cpu().numpy().tolist()}")
    assert pop.propagator.stats["workgroup_threads"] == 1024
    assert pop.propagator.stats["row_groups"] == 5
    assert rel.max() < 1e-5


def test_chebyshev_small_cells_half_workgroups(torch_gpu):
    """A population whose cells all fit one row group of 512 slots runs 512-thread workgroups
    (four per CU): cells of 1, 2, 63, 64, 65, 127, 128, 300 and 500 segments, 0.5 nA constant
    current at every root, 200 steps.  The same cells beside a 600-segment cell run 1024-thread
    workgroups: every voltage bit for bit the same (each thread does the same arithmetic on the
    same LDS places).  Against the FP64 oracle with the unpruned propagator: spike counts, and per
    cell the root trace within 1e-5 relative below excursions and 2e-3 mV near them, or three
    times the error FP32 storage alone gives the oracle on that cell (strongly driven cells of
    one to three segments), every segment at the end likewise."""
    import scipy.sparse
    from neuwon_amd import population, synth
    torch = torch_gpu
    rng = np.random.default_rng(22)
    sizes = [1, 2, 63, 64, 65, 127, 128, 300, 500]
    cells = [synth.random_cell(n, rng) for n in sizes]
    steps = 200
    runs = []
    for extra, threads in (([], 512), ([synth.random_cell(600, rng)], 1024)):
        pop = population.HHPopulation(cells + extra, store="chebyshev",
                                      stim={"mode": "constant", "amp_na": 0.5})
        assert pop.propagator.stats["workgroup_threads"] == threads
        assert pop.propagator.stats["row_groups"] == 1
        trace = torch.zeros((pop.num_cells, steps), dtype=torch.float32, device=pop.dev)
        pop.advance(steps, trace=trace)
        torch.cuda.synchronize()
        runs.append((pop, trace.cpu().numpy().astype(np.float64), pop.v.cpu().numpy().astype(np.float64)))
    pop, tr, v = runs[0]
    n = int(pop.cell_ptr[-1])
    dv = np.abs(v - runs[1][2][:n]).max()
    blocks = []
    for c in range(pop.num_cells):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        blocks.append(population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                                     pop.C[r0:r1], 0.05, 0.0))
    P = scipy.sparse.block_diag(blocks, format="csr")
    (ref, ref_v), (flo, flo_v) = fp32_floor.oracle_traces(pop, P, ("constant", 0.5e-9), steps)
    # the initial voltage heads each trace, so a spike in the first step counts as a crossing
    heads = np.full(pop.num_cells, pop.v0)
    errs = fp32_floor.check_cells(tr, ref, flo, 0.1, say=_say, heads=heads)
    _say(f"sizes {sizes}: 512 vs 1024 threads max |dv| {dv:.2e} mV; spikes "
         f"{pop.ap_count.cpu().numpy().tolist()}, end max |dv| {np.abs(v - ref_v).max():.2e} mV "
         f"(floor {np.abs(flo_v - ref_v).max():.2e})")
    assert dv == 0.0
    assert np.array_equal(pop.ap_count.cpu().numpy(), [e["spikes"] for e, _ in errs])
    assert np.abs(v - ref_v).max() < max(2e-3, 3.0 * np.abs(flo_v - ref_v).max())


def test_chebyshev_large_cell(torch_gpu):
    """A 4600-segment cell (rows in the fifth register row of a 1024-thread block), 0.05 nA
    constant current, 300 steps, every segment against the FP64 oracle."""
    from neuwon_amd import population, channels
    torch = torch_gpu
    cells = _c0_cell(4600, seed=3)
    pop = population.HHPopulation(cells, store="chebyshev", stim={"mode": "constant", "amp_na": 0.05})
    pop.advance(300)
    torch.cuda.synchronize()
    g, R, C, P = _oracle_cell(cells[0])
    gbar = {k: channels.HH_GBAR[k] * 1e-8 * g["surface_area"] for k in ("na", "k", "l")}
    stim = np.zeros(len(cells[0]))
    stim[0] = 0.05e-9
    sim = O.HHSim(P, C, gbar, v0=-70.0, dt=0.1, celsius=6.3, gate_mode="table",
                  stim=lambda t: stim)
    for _ in range(300):
        sim.advance()
    v = pop.v.cpu().numpy().astype(np.float64)
    rel = np.abs(v - sim.v) / np.abs(sim.v)
    _say(f"4600 segments: max rel {rel.max():.2e}, rows >= 4096 {rel[4096:].max():.2e}")
    assert rel.max() < 1e-5


def test_chebyshev_model_with_a_many_branched_soma(torch_gpu):
    """RxD_Model(propagator_store='chebyshev'): a soma with six dendrites (seven neighbours, the
    overflow list) and hh everywhere, 2 nA for 1 ms; the soma trace and every segment at the end
    against the general store built without pruning (both the same exp(A dt) to ~5e-9)."""
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
    t_cheb, v_cheb = run("chebyshev")
    t_gen, v_gen = run("general")
    err = voltage_error(t_cheb, t_gen, dt=0.025)
    _say(f"soma with 6 dendrites: peak {t_cheb.max():.1f} mV, trace {err}, final segments max "
         f"|dv| {np.abs(v_cheb - v_gen).max():.2e} mV")
    assert t_cheb.max() > 20
    assert err["rel_subthreshold"] < 1e-5 and err["abs_spiking_mV"] < 2e-3
    assert np.abs(v_cheb - v_gen).max() < 2e-3


def test_chebyshev_refusals(torch_gpu):
    """The species step needs a stored matrix; cells above 5120 segments and the matrix export
    are refused with a message."""
    import ctypes
    from neuwon_amd import _lib, population, synth
    pop = population.HHPopulation(synth.single_cell(300, seed=2), store="chebyshev")
    with pytest.raises(RuntimeError, match="no matrix"):
        pop.propagator.export_cell_stored(0)
    with pytest.raises(RuntimeError, match="Chebyshev"):
        population.HHPopulation(synth.single_cell(6000, seed=2), store="chebyshev")
    p = _lib.NwCaParams()
    p.dt_ms = 0.1
    rc = _lib.load().nw_ca_step(pop.propagator.handle, ctypes.byref(p), 1, _lib.ptr(pop.v),
                                _lib.ptr(pop.m), _lib.ptr(pop.h), None)
    assert rc != 0 and b"general format" in _lib.load().nw_last_error()
