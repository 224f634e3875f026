"""Edge cases of the fused step: ragged cell sizes down to one compartment, zero steps, an empty
population, cells at and beyond the LDS budget, and an empty model."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-edges] {msg}", flush=True)
    sys.stdout.flush()


def _store_pruned_block(pop, c):
    """FP64 propagator of cell c, pruned as the population's store prunes it."""
    from neuwon_amd import population
    r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
    args = (pop.parent_local[r0:r1], pop.R[r0:r1], pop.C[r0:r1], 0.5 * pop.dt)
    prune = pop.propagator.prune
    if not pop.propagator.symmetric:
        return population.cell_propagator_f64(*args, prune)
    full = population.cell_propagator_f64(*args, 0.0).toarray()
    keep = np.maximum(np.abs(full), np.abs(full.T)) >= prune
    np.fill_diagonal(keep, True)
    return scipy.sparse.csr_matrix(np.where(keep, full, 0.0))


@pytest.mark.parametrize("symmetric", [False, True])
def test_ragged_population_parity(torch_gpu, symmetric):
    """Cells of 1, 2, 3, 17, 64, 65, 300 and 1500 compartments in one launch, Poisson-driven,
    300 steps, against the FP64 C oracle: every compartment at the end and every root over time."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    rng = np.random.default_rng(21)
    sizes = [1, 2, 3, 17, 64, 65, 300, 1500]
    cells = [synth.random_cell(n, rng) for n in sizes]
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 60.0, "duration_ms": 1.0, "seed": 9}
    pop = population.HHPopulation(cells, prune=1e-9, stim=stim, symmetric=symmetric)
    steps = 300
    trace = torch.zeros((len(cells), steps), dtype=torch.float32, device=pop.dev)
    pop.advance(steps, trace=trace)
    torch.cuda.synchronize()
    v = pop.v.cpu().numpy().astype(np.float64)
    tr = trace.cpu().numpy().astype(np.float64)
    P = scipy.sparse.block_diag([_store_pruned_block(pop, c) for c in range(len(cells))],
                                format="csr")
    p = pop.params
    co = O.CHHPopulation(pop.cell_ptr, P, pop.C, pop.g_bar, dt=pop.dt, celsius=pop.celsius,
                         gate_mode="table",
                         stim=("poisson", float(p.stim_amp_a), int(p.stim_seed),
                               int(p.poisson_threshold), int(p.pulse_ticks)))
    ref = np.empty((len(cells), steps))
    for k in range(steps):
        co.advance(1)
        ref[:, k] = co.v[pop.cell_ptr[:-1]]
    d = np.abs(tr - ref)
    quiet = np.abs(ref) >= 10.0
    _say(f"{'symmetric' if symmetric else 'general'}: spikes per cell "
         f"{[int(((ref[c, 1:] >= 20) & (ref[c, :-1] < 20)).sum()) for c in range(len(cells))]}, "
         f"root traces max abs {d.max():.2e} mV, max rel (|v| >= 10 mV) "
         f"{(d[quiet] / np.abs(ref[quiet])).max():.2e}; final state max abs "
         f"{np.abs(v - co.v).max():.2e} mV")
    assert ((ref[:, 1:] >= 20) & (ref[:, :-1] < 20)).sum() >= 3
    assert d.max() < 5e-3
    assert np.abs(v - co.v).max() < 5e-3
    assert np.array_equal(pop.ap_count.cpu().numpy(),
                          ((ref[:, 1:] >= 20) & (ref[:, :-1] < 20)).sum(axis=1) +
                          (ref[:, 0] >= 20))


def test_zero_steps_is_noop(torch_gpu):
    from neuwon_amd import population, synth
    pop = population.HHPopulation(synth.single_cell(120, seed=3), symmetric=True,
                                  stim={"mode": "constant", "amp_na": 0.5})
    pop.advance(5)
    before = [t.clone() for t in (pop.v, pop.m, pop.h, pop.n)]
    pop.advance(0)
    torch_gpu.cuda.synchronize()
    for a, b in zip(before, (pop.v, pop.m, pop.h, pop.n)):
        assert torch_gpu.equal(a, b)
    assert pop.steps == 5


def test_empty_population_raises(torch_gpu):
    from neuwon_amd import population
    with pytest.raises(ValueError):
        population.HHPopulation([])


def test_lds_budget(torch_gpu):
    """A 14,000-compartment cell: 112 KB of LDS in the general store (advances), 168 KB in the
    symmetric store, more than a workgroup has (a clean error, nothing launched)."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    cells = synth.single_cell(14_000, seed=5)
    gen = population.HHPopulation(cells, symmetric=False, stim={"mode": "constant", "amp_na": 1.0})
    gen.advance(20)
    torch.cuda.synchronize()
    v = gen.v.cpu().numpy()
    _say(f"14,000-compartment cell, general store: v in [{v.min():.2f}, {v.max():.2f}] mV, "
         f"{gen.propagator.mean_nnz_per_row:.1f} entries per row")
    assert np.all(np.isfinite(v)) and v.max() < 60.0 and v.min() > -100.0
    sym = population.HHPopulation(cells, symmetric=True)
    with pytest.raises(RuntimeError, match="LDS"):
        sym.advance(1)
    assert sym.steps == 0


def test_empty_model_advances(torch_gpu):
    """A model with no segments, and one with a neuron but no mechanisms, step without error."""
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(mechanisms=["hh"])
    for _ in range(3):
        m.advance()
    assert m.clock() == pytest.approx(0.3)
    m2 = RxD_Model()
    root = m2.Neuron([0, 0, 0], 12).root
    for _ in range(10):
        m2.advance()
    assert root.voltage == pytest.approx(-70.0, abs=1e-6)   # no conductance, no current
    m2.check()
