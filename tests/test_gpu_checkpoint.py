"""State snapshots on the GPU: advancing from a restored snapshot reproduces the run that wrote
it.  Bit for bit with the general propagator store; the symmetric store adds its atomically
ordered sums in a different order each run, so there it agrees to FP32 rounding.  A GPU snapshot
restored into the FP64 oracle continues the same trajectory as the GPU resume, at the parity
bar."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-ckpt] {msg}", flush=True)
    sys.stdout.flush()


def _hh_state(pop):
    return {k: getattr(pop, k).cpu().numpy().copy() for k in pop._STATE}


def _compare(a, b, bitwise, what):
    for k in a:
        if bitwise or a[k].dtype.kind in "iu":
            assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"
        else:
            d = np.abs(a[k].astype(np.float64) - b[k]).max()
            assert d < 1e-3, f"{what}: {k} differs by {d}"


@pytest.mark.parametrize("store", ["general", "symmetric", "chebyshev"])
def test_hh_population_resume(torch_gpu, tmp_path, store):
    """2,000 steps of a Poisson-stimulated population; snapshot at step 1,000; a second population
    of the same cells loads it and runs the last 1,000 steps."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    cells = synth.population(total=20_000, seed=11, size_range=(200, 2000))
    stim = {"mode": "poisson", "amp_na": 0.6, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 5}
    symmetric = store == "symmetric"
    a = population.HHPopulation(cells, stim=stim, store=store)
    a.advance(1000)
    f = tmp_path / "hh.npz"
    a.save(f)
    a.advance(1000)
    torch.cuda.synchronize()
    want = _hh_state(a)
    b = population.HHPopulation(cells, store=store)
    b.load(f)
    assert b.steps == 1000 and b.stim == stim
    b.advance(1000)
    torch.cuda.synchronize()
    got = _hh_state(b)
    spikes = int(want["ap_count"].sum())
    _say(f"{store} store: {spikes} spikes over 2,000 steps, "
         f"max |dv| after resume {np.abs(want['v'] - got['v']).max():.2e} mV")
    assert spikes > 0
    _compare(want, got, not symmetric, "HHPopulation")
    with pytest.raises(ValueError):
        population.HHPopulation(cells[:3]).load(f)


def test_ca_population_resume(torch_gpu, tmp_path):
    from neuwon_amd import diffusion, synth
    torch = torch_gpu
    cells = synth.population(total=5000, seed=3, size_range=(200, 900))
    a = diffusion.CaPopulation(cells, seed=3, prune=1e-9)
    a.advance(200)
    f = tmp_path / "ca.npz"
    a.save(f)
    a.advance(300)
    torch.cuda.synchronize()
    want = {k: getattr(a, k).cpu().numpy().copy() for k in ("ca", "cab_offset", "dcdt")}
    b = diffusion.CaPopulation(cells, seed=3, prune=1e-9)
    b.load(f)
    b.advance(300)
    torch.cuda.synchronize()
    got = {k: getattr(b, k).cpu().numpy().copy() for k in ("ca", "cab_offset", "dcdt")}
    _say(f"Ca population: max |dca| after resume {np.abs(want['ca'] - got['ca']).max():.2e}")
    assert b.steps == a.steps == 500
    _compare(want, got, True, "CaPopulation")


def _model(symmetric, mechanisms):
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(time_step=0.025, temperature=6.3, mechanisms=mechanisms,
                  symmetric_propagator=symmetric)
    segs = []
    for k in range(3):
        root = m.Neuron([300.0 * k, 0, 0], 12).root
        dend = root.add_section([300.0 * k, 150, 0], 2, maximum_segment_length=15)
        segs += [root] + dend
        for s in [root] + dend[k % 2::2]:
            s.insert({"markov13": 1, "hh": 1})
    return m, segs


@pytest.mark.parametrize("symmetric", [False, True])
def test_model_resume(torch_gpu, tmp_path, symmetric):
    """RxD_Model with hh and a 13-state kinetic scheme on three neurons: a pulse during the first
    300 steps, snapshot, 300 more steps; a second model built the same way loads the snapshot and
    runs the same 300 steps.  Every device attribute of every class and the clocks agree."""
    from neuwon_amd import channels
    from neuwon_amd.model import LTIMechanism
    scheme = channels.random_ampa13_scheme(seed=2)
    mechs = lambda: [LTIMechanism(scheme, gmax_s_per_cm2=0.0005, reversal=0.0, name="markov13"),
                     "hh"]
    a, segs = _model(symmetric, mechs())
    segs[0].inject_current(0.3, 1.0)
    for _ in range(300):
        a.advance()
    f = tmp_path / "model.npz"
    a.save_state(f)
    for _ in range(300):
        a.advance()
    b, _ = _model(symmetric, mechs())
    b.load_state(f)
    assert b.clock.ticks == 300 and b.input_hook.ticks == 600
    for _ in range(300):
        b.advance()
    assert b.clock.ticks == a.clock.ticks and b.input_hook.ticks == a.input_hook.ticks
    va = a.database.get_data("Segment.voltage")
    vb = b.database.get_data("Segment.voltage")
    _say(f"model, {'symmetric' if symmetric else 'general'} store: peak {va.max():.1f} mV at the "
         f"end, max |dv| after resume {np.abs(va - vb).max():.2e} mV")
    for cname, c in a.database.classes.items():
        for col in c.columns.values():
            if not col.device:
                continue
            x = a.database.get_data(f"{cname}.{col.name}")
            y = b.database.get_data(f"{cname}.{col.name}")
            _compare({col.name: np.asarray(x)}, {col.name: np.asarray(y)}, not symmetric,
                     f"{cname}.{col.name}")
    # a model of another shape refuses the snapshot
    c, _ = _model(symmetric, mechs())
    c.Neuron([5000, 0, 0], 5)
    with pytest.raises(ValueError):
        c.load_state(f)


def _oracle_stim(pop):
    p = pop.params
    if p.stim_mode == 2:
        return ("poisson", float(p.stim_amp_a), int(p.stim_seed), int(p.poisson_threshold),
                int(p.pulse_ticks))
    if p.stim_mode == 1:
        return ("constant", float(p.stim_amp_a))
    return None


@pytest.mark.parametrize("store", ["general", "chebyshev"])
def test_hh_snapshot_resumed_in_the_oracle(torch_gpu, tmp_path, store):
    """A snapshot the GPU wrote after 1,000 Poisson-stimulated steps, restored into the FP64 oracle
    (neuwon/database/sql.py:78-128 is the persistence it replaces) and into a fresh GPU
    population: both run 1,000 more steps, every cell's root voltage after every step and every
    voltage at the end compared at the parity bar, beside the oracle with FP32 storage emulated
    from the same snapshot, on the FP32 matrix the device stores for the stored format (the
    floor)."""
    import scipy.sparse

    import fp32_floor
    from neuwon_amd import checkpoint, population, synth
    from oracle import neuwon_oracle as O
    torch = torch_gpu
    cells = synth.population(total=20_000, seed=11, size_range=(200, 2000))
    stim = {"mode": "poisson", "amp_na": 0.6, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 5}
    a = population.HHPopulation(cells, stim=stim, store=store)
    a.advance(1000)
    f = tmp_path / "hh.npz"
    a.save(f)
    b = population.HHPopulation(cells, store=store)
    b.load(f)
    steps = 1000
    tr = torch.zeros((b.num_cells, steps), dtype=torch.float32, device=b.dev)
    b.advance(steps, trace=tr)
    torch.cuda.synchronize()
    tr = tr.cpu().numpy().astype(np.float64)
    v_gpu = b.v.cpu().numpy().astype(np.float64)
    # the oracle from the snapshot file: the operator the store applies (pruned as stored for the
    # general store, exp(A dt) unpruned for the matrix-free one), the state, the step count
    arrays, meta = checkpoint.load(f, "HHPopulation")
    prune = b.propagator.prune if store == "general" else 0.0
    blocks = [population.cell_propagator_f64(b.parent_local[r0:r1], b.R[r0:r1], b.C[r0:r1],
                                             0.5 * b.dt, prune)
              for r0, r1 in zip(b.cell_ptr[:-1], b.cell_ptr[1:])]
    P = scipy.sparse.block_diag(blocks, format="csr")
    # the floor: FP32 storage of the state and, for the stored format, the FP32 matrix the device
    # holds (exported as stored: FP32 values, the diagonal that keeps the FP64 row sums)
    P_floor = P if store != "general" else scipy.sparse.block_diag(
        [b.propagator.export_cell(c).astype(np.float64) for c in range(b.num_cells)], format="csr")
    roots = b.cell_ptr[:-1]
    runs = []
    for emulate in (False, True):
        co = O.CHHPopulation(b.cell_ptr, P_floor if emulate else P, b.C, b.g_bar, dt=b.dt,
                             celsius=b.celsius,
                             gate_mode=b.gate_mode, stim=_oracle_stim(b), fp32_storage=emulate,
                             round_v=emulate and store == "general")
        co.v[:] = arrays["v"].astype(np.float64)
        for g in "mhn":
            co.gates[g][:] = arrays[g].astype(np.float64)
        m, h, n = (co.gates[g] for g in "mhn")
        g_na, g_k, g_l = co.g["na"] * m ** 3 * h, co.g["k"] * n ** 4, co.g["l"]
        co.sum_g[:] = g_na + g_k + g_l
        co.d_num[:] = g_na * b.reversal["na"] + g_k * b.reversal["k"] + g_l * b.reversal["l"]
        co.steps = int(meta["steps"])
        ref = np.empty((b.num_cells, steps))
        for k in range(steps):
            co.advance(1, threads=len(os.sched_getaffinity(0)))
            ref[:, k] = co.v[roots]
        runs.append((ref, co.v.copy()))
    (ref, v_ref), (flo, v_flo) = runs
    spikes = int(b.ap_count.cpu().numpy().sum() - arrays["ap_count"].sum())
    _say(f"{store} store, resumed at step {meta['steps']} in the oracle: {b.num_cells} cells, "
         f"{spikes} spikes in the last {steps} steps")
    fp32_floor.check_cells(tr, ref, flo, b.dt, say=_say)
    end, floor_end = np.abs(v_gpu - v_ref).max(), np.abs(v_flo - v_ref).max()
    _say(f"every segment at the end: {end:.2e} mV (floor {floor_end:.2e})")
    assert spikes > 0
    assert end < max(2e-3, 3.0 * floor_end)
