"""The autotuner on the GPU: HHPopulation(store="auto") times both matrix-free stores and then
behaves exactly like a population built with the chosen store; block_threads="auto" on a stored
format times every workgroup size, resets the state and then matches a fresh population with the
chosen size bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-autotune] {msg}", flush=True)
    sys.stdout.flush()


def _run(pop, steps):
    pop.advance(steps)
    pop.torch.cuda.synchronize()
    return pop.v.cpu().numpy().copy(), pop.ap_count.cpu().numpy().copy()


def test_store_auto(torch_gpu):
    from neuwon_amd import population, synth
    cells = synth.population(total=30_000, seed=7, size_range=(200, 5000))
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 40.0, "duration_ms": 1.0, "seed": 3}
    pop = population.HHPopulation(cells, stim=stim, store="auto")
    rep = pop.tuning["store"]
    _say(f"{pop.num_cells} cells: {rep}")
    assert set(rep["ms_per_step"]) == {"chebyshev", "rational"}
    assert pop.propagator.store == rep["choice"] == min(rep["ms_per_step"], key=rep["ms_per_step"].get)
    assert pop.steps == 0
    ref = population.HHPopulation(cells, stim=stim, store=rep["choice"])
    v, ap = _run(pop, 60)
    v_ref, ap_ref = _run(ref, 60)
    assert np.array_equal(v, v_ref) and np.array_equal(ap, ap_ref)
    again = population.HHPopulation(cells, stim=stim, store="auto")
    assert again.tuning["store"]["cached"] and again.propagator.store == rep["choice"]


def test_block_threads_auto_on_the_general_store(torch_gpu):
    from neuwon_amd import population, synth
    cells = synth.population(total=8_000, seed=8, size_range=(200, 2000))
    stim = {"mode": "constant", "amp_na": 0.5}
    pop = population.HHPopulation(cells, stim=stim, store="general", block_threads="auto")
    rep = pop.tuning["block_threads"]
    _say(f"{pop.num_cells} cells, general store: {rep}")
    assert set(rep["ms_per_step"]) == {128, 256, 512, 1024}
    assert pop.block_threads == rep["choice"] == pop.params.block_threads
    assert pop.steps == 0 and float(pop.v.max()) == float(pop.v.min()) == -70.0
    ref = population.HHPopulation(cells, stim=stim, store="general", block_threads=rep["choice"])
    v, ap = _run(pop, 40)
    v_ref, ap_ref = _run(ref, 40)
    assert np.array_equal(v, v_ref) and np.array_equal(ap, ap_ref)
    cheb = population.HHPopulation(cells, stim=stim, store="chebyshev", block_threads="auto")
    assert cheb.tuning["block_threads"]["choice"] == 0


@pytest.mark.parametrize("sizes", [(200, 5000), (100, 500)])
def test_cheb_launch_forms_agree_and_auto(torch_gpu, sizes):
    """The Chebyshev store's launch forms (one launch per half step, per step, per two steps)
    give bitwise the same trajectory over 41 steps (an odd count: the two-step form ends with a
    step alone), on cells of up to five row groups and on cells of one row group (512-thread
    launches); cheb_launch="auto" times the three forms, resets the state and then matches."""
    from neuwon_amd import population, synth
    cells = synth.population(total=20_000, seed=9, size_range=sizes)
    stim = {"mode": "poisson", "amp_na": 0.5, "rate_hz": 200.0, "duration_ms": 1.0, "seed": 4}
    runs = {}
    for form in ("half", "step", "two_steps"):
        pop = population.HHPopulation(cells, stim=stim, store="chebyshev", cheb_launch=form)
        pop.advance(20)
        runs[form] = _run(pop, 21) + (pop.m.cpu().numpy().copy(),)
    for form in ("step", "two_steps"):
        assert all(np.array_equal(a, b) for a, b in zip(runs[form], runs["half"])), form
    pop = population.HHPopulation(cells, stim=stim, store="chebyshev", cheb_launch="auto")
    rep = pop.tuning["cheb_launch"]
    _say(f"{pop.num_cells} cells of {sizes}, row groups {pop.propagator.stats['row_groups']}: {rep}")
    assert set(rep["ms_per_step"]) == {"half", "step", "two_steps"}
    assert pop.steps == 0 and float(pop.v.max()) == float(pop.v.min()) == -70.0
    pop.advance(20)
    v, ap = _run(pop, 21)
    assert np.array_equal(v, runs["half"][0]) and np.array_equal(ap, runs["half"][1])
    with pytest.raises(ValueError):
        population.HHPopulation(cells, store="chebyshev", cheb_launch="three_steps")


def test_model_propagator_store_auto(torch_gpu):
    """RxD_Model(propagator_store="auto"): both matrix-free stores are timed on the model's half
    step (on a copy of the voltages) and the model then runs exactly as one built with the chosen
    store."""
    from neuwon_amd.model import RxD_Model, TimeSeries

    def run(store):
        m = RxD_Model(time_step=0.1, temperature=6.3, mechanisms=["hh"], propagator_store=store)
        root = m.Neuron([0, 0, 0], 12).root
        segs = [root]
        for k in range(3):
            a = 2 * np.pi * k / 3
            segs += root.add_section([200 * np.cos(a), 200 * np.sin(a), 0], 1.0,
                                     maximum_segment_length=10)
        for s in segs:
            s.insert("hh")
        probe = TimeSeries().record(root, "voltage")
        root.inject_current(1.0, 2.0)
        for _ in range(150):
            m.advance()
        return m, np.array(probe.get_data()), m.database.get_data("Segment.voltage").copy()
    m, trace, v = run("auto")
    rep = m.tuning["propagator_store"]
    _say(f"model: {rep}")
    assert set(rep["ms_per_half_step"]) == {"chebyshev", "rational"}
    _, trace_ref, v_ref = run(rep["choice"])
    assert trace.max() > 20
    assert np.array_equal(trace, trace_ref) and np.array_equal(v, v_ref)
