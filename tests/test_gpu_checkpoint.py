"""State snapshots on the GPU: advancing from a restored snapshot reproduces the run that wrote
it.  Bit for bit with the general propagator store; the symmetric store adds its atomically
ordered sums in a different order each run, so there it agrees to FP32 rounding."""
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
