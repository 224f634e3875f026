"""Intracellular Ca2+ diffusion with a first-order buffer on the GPU against the FP64 oracle."""
import sys
import time

import numpy as np
import pytest

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-species] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_cell(pop, c):
    r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
    parent = pop.parent_local[r0:r1].astype(np.int64)
    k = pop.k[r0:r1]
    P = O.diffusion_propagator(parent, k, pop.vol[r0:r1], 0.5 * pop.dt, prune=pop.propagator.prune)
    return r0, r1, P


def test_diffusion_buffer_matches_oracle(torch_gpu):
    from neuwon_amd import diffusion, synth
    torch = torch_gpu
    cells = synth.population(total=2500, seed=3, size_range=(200, 900))
    pop = diffusion.CaPopulation(cells, seed=3, prune=1e-9)
    st = pop.propagator.stats
    _say(f"cells {pop.num_cells} nnz/row {pop.propagator.mean_nnz_per_row:.2f} max {st['max_row_nnz']}")
    m0 = pop.mass()
    steps = 300
    pop.advance(steps)
    torch.cuda.synchronize()
    ca = pop.ca.cpu().numpy().astype(np.float64)
    cab = pop.cab()
    worst = 0.0
    for c in range(pop.num_cells):
        r0, r1, P = _oracle_cell(pop, c)
        exported = pop.propagator.export_cell(c)
        assert (exported != 0).sum() == P.nnz
        sim = O.CaSim(P, pop.vol[r0:r1], pop.ca0[r0:r1], pop.cab0[r0:r1])
        for _ in range(steps):
            sim.advance()
        err_ca = np.abs(ca[r0:r1] - sim.ca).max() / np.abs(sim.ca).max()
        err_cab = np.abs(cab[r0:r1] - sim.cab).max() / np.abs(sim.cab).max()
        worst = max(worst, err_ca, err_cab)
    m1 = pop.mass()
    _say(f"max rel err vs oracle {worst:.2e}; mass drift {(m1 - m0) / m0:.2e}")
    assert worst < 1e-5
    assert abs(m1 - m0) / m0 < 1e-6


def test_ca_workload_5M(torch_gpu):
    """5M compartments, 5000 steps: FP64 mass check every 1000 steps, oracle parity on 3 cells."""
    from neuwon_amd import diffusion, synth
    torch = torch_gpu
    t0 = time.time()
    cells = synth.population(total=5_000_000, seed=3)
    pop = diffusion.CaPopulation(cells, seed=3, prune=1e-9)
    _say(f"5M: {pop.num_segments} compartments, {pop.num_cells} cells, nnz/row "
         f"{pop.propagator.mean_nnz_per_row:.2f}, build {pop.propagator.stats['build_seconds']:.1f}s, "
         f"setup {time.time() - t0:.1f}s")
    m0 = pop.mass()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    drift = []
    ms = 0.0
    for _ in range(5):
        e0.record()
        pop.advance(1000)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
        drift.append((pop.mass() - m0) / m0)
    rate = pop.num_segments * 5000 / (ms * 1e-3)
    _say(f"5M x 5000 steps: {ms:.1f} ms kernel, {rate / 1e9:.2f} G cmpt-steps/s, "
         f"mass drift every 1000 steps {[f'{d:.2e}' for d in drift]}")
    assert max(abs(d) for d in drift) < 1e-5
    ca = pop.ca.cpu().numpy().astype(np.float64)
    for c in (0, pop.num_cells // 2, pop.num_cells - 1):
        r0, r1, P = _oracle_cell(pop, c)
        if r1 - r0 > 2000:
            continue
        sim = O.CaSim(P, pop.vol[r0:r1], pop.ca0[r0:r1], pop.cab0[r0:r1])
        for _ in range(5000):
            sim.advance()
        err = np.abs(ca[r0:r1] - sim.ca).max() / np.abs(sim.ca).max()
        _say(f"cell {c} ({r1 - r0} cmpts): max rel err {err:.2e}")
        assert err < 1e-5
