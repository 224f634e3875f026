"""Symmetric half storage of the electric propagator (S = W P W^-1, w = sqrt(C), upper triangle):
the stored triangle against the FP64 propagator, the fused step in that format against the FP64
oracle and against the general format, and the species step, which refuses the format."""
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

import fp32_floor
from test_gpu_parity import _c0_cell, _oracle_cell, voltage_error

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-sym] {msg}", flush=True)
    sys.stdout.flush()


def _raw_rows(prop, cell, n):
    from neuwon_amd import _lib
    off = np.empty(n + 1, dtype=np.uint32)
    _lib.check(_lib.load().nw_propagator_cell_export(prop.handle, cell, _lib.ptr(off), None, None))
    cols = np.empty(int(off[-1]), dtype=np.uint16)
    vals = np.empty(int(off[-1]), dtype=np.float32)
    _lib.check(_lib.load().nw_propagator_cell_export(prop.handle, cell, None, _lib.ptr(cols),
                                                     _lib.ptr(vals)))
    return off, cols, vals


def _check_chunk_order(off, cols, n):
    """Within every 256-entry chunk of a row, the entries the lanes load q-th (slots 4l + q) are in
    increasing column order, padding (column = the row) last."""
    for r in range(n):
        b, e = int(off[r]), int(off[r + 1])
        assert (e - b) % 4 == 0
        for c0 in range(b, e, 256):
            m = min(256, e - c0) // 4
            logical = cols[c0:c0 + 4 * m].reshape(m, 4).T.ravel().astype(np.int64)
            real = logical[logical != r]
            assert np.all(np.diff(real) > 0)


def test_stored_chunk_order(torch_gpu):
    from neuwon_amd import synth, population
    cells = synth.population(total=6000, seed=5, size_range=(300, 3000))
    for sym in (False, True):
        pop = population.HHPopulation(cells, prune=1e-9, symmetric=sym)
        for c in range(pop.num_cells):
            n = int(pop.cell_ptr[c + 1] - pop.cell_ptr[c])
            off, cols, _ = _raw_rows(pop.propagator, c, n)
            _check_chunk_order(off, cols, n)


def test_symmetric_store_rebuilds_propagator(torch_gpu):
    from neuwon_amd import synth, population
    prune = 1e-9
    cells = synth.population(total=3000, seed=3, size_range=(200, 900))
    pop = population.HHPopulation(cells, prune=prune, symmetric=True)
    prop = pop.propagator
    st = prop.stats
    assert st["symmetric"] == 1
    _say(f"cells {pop.num_cells} nnz/row {prop.mean_nnz_per_row:.1f} stored/row "
         f"{st['stored'] / st['num_rows']:.1f}")
    assert st["stored"] < 0.55 * st["nnz"] + 4 * st["num_rows"]
    worst = 0.0
    for c in range(pop.num_cells):
        r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
        n = r1 - r0
        stored = prop.export_cell_stored(c)
        assert np.all(stored.indices >= np.repeat(np.arange(n), np.diff(stored.indptr)))  # upper
        w = prop.cell_weights(c)
        assert np.allclose(w, np.sqrt(pop.C[r0:r1] / pop.C[r0]), rtol=1e-15)
        P32 = prop.export_cell(c).toarray()
        full = population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                              pop.C[r0:r1], 0.5 * pop.dt, 0.0).toarray()
        # a pair is kept when either entry reaches the threshold
        keep = np.maximum(full, full.T) >= prune
        np.fill_diagonal(keep, True)
        assert np.array_equal(P32 != 0, keep & (full != 0))
        off = ~np.eye(n, dtype=bool) & keep
        # FP32 rounding of S_ij (w_j / w_i exact in FP64), plus the FP64 expm's absolute error
        # that averaging the two triangles spreads over small entries
        err = np.abs(P32[off] - full[off])
        worst = max(worst, (err / np.abs(full[off])).max())
        assert np.all(err <= 1.2e-7 * np.abs(full[off]) + 1e-15)
        # the FP32 diagonal restores the FP64 row sum of the kept entries up to its own rounding
        dropped = np.where(keep, 0.0, full).sum(axis=1)
        assert np.abs(P32.sum(axis=1) - (full.sum(axis=1) - dropped)).max() < 1e-7
    _say(f"max rel off-diagonal error {worst:.2e}")


@pytest.mark.parametrize("gate_mode", ["table", "exact"])
def test_symmetric_single_neuron_trajectory(torch_gpu, gate_mode):
    """1000-compartment tree, HH, 0.5 nA step at the root, dt 0.1 ms x 1000 steps, symmetric store."""
    from neuwon_amd import population, channels
    torch = torch_gpu
    steps = 1000
    cells = _c0_cell()
    pop = population.HHPopulation(cells, prune=O.EPSILON, gate_mode=gate_mode,
                                  stim={"mode": "constant", "amp_na": 0.5}, symmetric=True)
    trace = torch.zeros((1, steps), dtype=torch.float32, device=pop.dev)
    pop.advance(steps, trace=trace)
    torch.cuda.synchronize()
    v_gpu = pop.v.cpu().numpy().astype(np.float64)
    tr_gpu = trace.cpu().numpy()[0].astype(np.float64)
    g, R, C, P = _oracle_cell(cells[0])
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


def _pair_pruned(full, prune):
    """Pruning of the symmetric store: a pair is kept when either entry reaches the threshold."""
    keep = np.maximum(full, full.T) >= prune
    np.fill_diagonal(keep, True)
    return np.where(keep, full, 0.0)


@pytest.mark.parametrize("block_threads", [128, 512, 1024])
def test_symmetric_population_matches_general(torch_gpu, block_threads):
    """A Poisson-driven population (20k compartments, 600 steps in three launches) in both formats,
    each against the FP64 C oracle run with the same pruned propagator (entry rule for the general
    store, pair rule for the symmetric one): the same spikes, subthreshold voltages within 1e-5
    relative.  Near a spike the trajectories are compared in mV."""
    import os
    import scipy.sparse
    from neuwon_amd import population, synth
    torch = torch_gpu
    prune = 1e-9
    cells = synth.population(total=20_000, seed=11, size_range=(200, 2500))
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 5}
    threads = len(os.sched_getaffinity(0))
    out = {}
    for sym in (False, True):
        pop = population.HHPopulation(cells, prune=prune, stim=stim, symmetric=sym,
                                      block_threads=block_threads)
        parts = []
        for k in range(3):
            trace = torch.zeros((pop.num_cells, 200), dtype=torch.float32, device=pop.dev)
            pop.advance(200, trace=trace)
            parts.append(trace)
        torch.cuda.synchronize()
        tr = torch.cat(parts, dim=1).cpu().numpy().astype(np.float64)
        ap = pop.ap_count.cpu().numpy().copy()
        blocks = []
        for c in range(pop.num_cells):
            r0, r1 = int(pop.cell_ptr[c]), int(pop.cell_ptr[c + 1])
            full = population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1],
                                                  pop.C[r0:r1], 0.05, 0.0).toarray()
            Pc = _pair_pruned(full, prune) if sym else np.where(np.abs(full) >= prune, full, 0.0)
            blocks.append(scipy.sparse.csr_matrix(Pc))
        P = scipy.sparse.block_diag(blocks, format="csr")
        p = pop.params
        stim_t = ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
                  int(p.pulse_ticks))
        # the stored formats keep FP32 voltages between steps: the floor rounds them too
        (ref, _), (flo, _) = fp32_floor.oracle_traces(pop, P, stim_t, 600, round_v=True)
        name = "symmetric" if sym else "general"
        _say(f"block {block_threads} {name}: spikes {ap.sum()}")
        errs = fp32_floor.check_cells(tr, ref, flo, 0.1, say=_say)
        out[sym] = (tr, ap, [e for e, _ in errs])
    for sym in (False, True):
        tr, ap, errs = out[sym]
        assert ap.sum() > 0
        assert ap.sum() == sum(e["spikes"] for e in errs)
    assert np.array_equal(out[False][1], out[True][1])


def test_symmetric_refused_by_species_step(torch_gpu):
    import ctypes
    from neuwon_amd import _lib, population, synth
    torch = torch_gpu
    pop = population.HHPopulation(synth.single_cell(200, seed=2), symmetric=True)
    lib = _lib.load()
    z = torch.zeros(pop.num_segments, device=pop.dev)
    cp = _lib.NwCaParams()
    cp.dt_ms = 0.1
    rc = lib.nw_ca_step(pop.propagator.handle, ctypes.byref(cp), 1, _lib.ptr(z), _lib.ptr(z),
                        _lib.ptr(z), None)
    assert rc != 0 and b"symmetric" in lib.nw_last_error()
