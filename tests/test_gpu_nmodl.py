"""Kinetic .mod files through the NMODL front end on the GPU (files written for these tests,
tests/golden/kinetic): the LTI kernel on the parsed tables, and a model carrying both channels
against the FP64 oracle step with the same tables."""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

from test_gpu_parity import voltage_error

pytestmark = pytest.mark.gpu
OWN = os.path.join(os.path.dirname(__file__), "golden", "kinetic")


def _say(msg):
    print(f"[gpu-nmodl] {msg}", flush=True)
    sys.stdout.flush()


def test_parsed_scheme_on_lti_kernel(torch_gpu):
    """200,000 instances of the 6-state na6 scheme at voltages held in [-90, 40] mV, 200 steps of
    0.025 ms in one launch, against the oracle's FP64 table kernel."""
    from neuwon_amd import _lib, nmodl
    torch = torch_gpu
    scheme = nmodl.load_kinetic(os.path.join(OWN, "na6.mod"), time_step=0.025, celsius=22.0)
    node = scheme.node_table()
    poly = O.node_table_to_polynomial(node)
    M, steps = 200_000, 200
    rng = np.random.default_rng(3)
    v = rng.uniform(-90.0, 40.0, M).astype(np.float32)
    base = np.stack([scheme.initial_state(x) for x in (-90.0, -60.0, -30.0, 0.0)], axis=1)
    st0 = base[:, rng.integers(0, 4, M)]
    tbl = torch.as_tensor(node.astype(np.float32), device="cuda")
    vs = torch.as_tensor(v, device="cuda")
    st = torch.as_tensor(np.ascontiguousarray(st0, dtype=np.float32), device="cuda")
    grid = scheme.grid
    _lib.check(_lib.load().nw_lti_advance(6, grid.num_buckets + 1, grid.minimum, grid.maximum,
                                          _lib.ptr(tbl), M, _lib.ptr(vs), None, _lib.ptr(st), M,
                                          1.0, steps, None, None, None, None))
    torch.cuda.synchronize()
    got = st.cpu().numpy().astype(np.float64)
    sample = rng.choice(M, 3000, replace=False)
    ref = st0[:, sample].astype(np.float32).astype(np.float64)
    vv = v[sample].astype(np.float64)
    for _ in range(steps):
        ref = O.lti_advance_vec(ref, vv, poly, conserve_sum=1.0)
    err = np.abs(got[:, sample] - ref).max()
    _say(f"na6 scheme, {M} instances x {steps} steps: max abs state error {err:.2e}")
    assert err < 1e-5
    assert np.abs(got.sum(axis=0) - 1.0).max() < 1e-5


@pytest.mark.parametrize("accuracy", [None, 1e-4])
def test_model_with_kinetic_mod_files(torch_gpu, accuracy):
    """RxD_Model(mechanisms=[na6.mod, k4.mod]) on a soma and a six-segment dendrite, 0.3 nA for
    1 ms from t = 1 ms, dt 0.025 ms x 400 steps at 22 degC: every segment's voltage and the
    channels' states against KineticSim (FP64) fed the same tables, geometry and stimulus; the
    tables on the fixed grids, or sized for the reference's 1e-4 (kinetic_table_accuracy)."""
    from neuwon_amd import channels, nmodl
    from neuwon_amd.model import RxD_Model, TimeSeries
    torch = torch_gpu
    dt, celsius, steps = 0.025, 22.0, 400
    files = [os.path.join(OWN, "na6.mod"), os.path.join(OWN, "k4.mod")]
    species = [{"name": "na", "reversal_potential": 50.0}, {"name": "k", "reversal_potential": -85.0}]
    m = RxD_Model(time_step=dt, temperature=celsius, initial_voltage=-65.0, mechanisms=files,
                  species=species, kinetic_table_accuracy=accuracy)
    if accuracy is not None:
        for name in ("na6", "k4"):
            scheme = m.mechanisms[name].spec.scheme
            assert scheme.inputs[0].num_buckets < 2000 and scheme.table_error <= accuracy
    root = m.Neuron([0, 0, 0], 15).root
    dend = root.add_section([120, 0, 0], 2, maximum_segment_length=20)
    segs = [root] + dend
    for s in segs:
        s.insert({"na6": 1, "k4": 1})
    probe = TimeSeries().record(root, "voltage")
    m.advance()
    first = {name: np.stack([m.database.get_data(f"{name}.{st}")
                             for st in nmodl.KineticModel(f, celsius=celsius).states])
             for name, f in (("na6", files[0]), ("k4", files[1]))}
    first_v = np.array([s.voltage for s in segs])
    first_g = m.database.get_data("Segment.sum_conductance").copy()
    for k in range(1, steps):
        if m.clock.is_now(1.0):
            root.inject_current(0.3, 1.0)
        m.advance()
    v_gpu = np.array([s.voltage for s in segs])
    trace = np.array(probe.get_data())

    # the oracle: same geometry (a straight section whose diameter tapers from the soma's to the
    # tip's, segment.py:75-104), tables, magnitudes and stimulus
    n = len(segs)
    parent = np.array([-1] + list(range(n - 1)))
    coords = np.array([np.asarray(s.coordinates, dtype=np.float64) for s in segs])
    g = O.geometry(parent, coords, np.array([s.diameter for s in segs]))
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    P = O.propagator(parent, R, C, 0.5 * dt)
    mechs = []
    for f, ion in zip(files, ("na", "k")):
        scheme = nmodl.load_kinetic(f, time_step=dt, celsius=celsius, accuracy=accuracy)
        km = scheme.model
        w = km.conductance_weights(f"e{ion}")
        init = scheme.initial_state(-65.0)
        mechs.append({"poly": O.node_table_to_polynomial(scheme.node_table()), "weights": w,
                      "gmax": g["surface_area"] * 1e-8, "E": 50.0 if ion == "na" else -85.0,
                      "states": np.repeat(init[:, None], n, axis=1), "conserve": 1.0})
    pulse = []

    def stim(tick):
        I = np.zeros(n)
        for t0, nt in pulse:
            if t0 <= tick < t0 + nt:
                I[0] += 0.3e-9
        return I
    sim = O.KineticSim(P, C, mechs, v0=-65.0, dt=dt, stim=stim)
    ref = [sim.v[0]]
    sim.advance()
    ref.append(sim.v[0])
    assert np.abs(first_v - sim.v).max() < 1e-4
    assert np.allclose(first_g, sim.sum_g, rtol=1e-5)
    for k, name in enumerate(("na6", "k4")):
        assert np.abs(first[name] - sim.mech[k]["states"]).max() < 1e-6
    for k in range(1, steps):
        if k == round(1.0 / dt):
            pulse.append((sim.input_ticks + 1, O.pulse_ticks(1.0, 0.5 * dt)))
        sim.advance()
        ref.append(sim.v[0])
    ref = np.array(ref)
    d = np.abs(trace - ref)
    up = lambda r: int(np.argmax(r > 0.0)) if (r > 0.0).any() else -1
    _say(f"model with na6 + k4: soma peak {trace.max():.1f} mV (oracle {ref.max():.1f}), spike "
         f"at step {up(trace)} (oracle {up(ref)}), max |dv| before the stimulus {d[:41].max():.2e}, "
         f"soma trace {d.max():.2e} mV, all segments at the end "
         f"{np.abs(v_gpu - sim.v).max():.2e} mV")
    assert trace.max() > 0.0                     # the channels fire a spike
    err = voltage_error(trace, ref, dt)
    _say(f"soma trace: subthreshold max rel {err['rel_subthreshold']:.2e}, within 2 ms of the "
         f"spike max abs {err['abs_spiking_mV']:.2e} mV")
    assert err["spikes"] == 1
    # before the stimulus the cell is at rest: the 1e-5 relative bar
    assert (d[:41] / np.abs(ref[:41])).max() < 1e-5
    # this cell rises slowly to threshold and repolarises over several ms, so the samples more
    # than 2 ms from the upstroke are still on the action potential.  The FP64 oracle with v,
    # states, conductances and tables rounded to FP32 every step is 9.5e-5 relative there and
    # 8.2e-3 mV at the upstroke from the plain FP64 oracle: the bars below are about 3x those
    assert err["rel_subthreshold"] < 3e-4
    assert err["abs_spiking_mV"] < 2.5e-2
    assert np.abs(v_gpu - sim.v).max() < 5e-3
    states = np.stack([m.database.get_data(f"na6.{s}") for s in mechs_state_names(files[0])])
    assert np.abs(states.sum(axis=0) - 1.0).max() < 1e-5


def mechs_state_names(path):
    from neuwon_amd import nmodl
    return nmodl.KineticModel(path).states


def _grid_dict(g):
    from neuwon_amd import channels
    log = isinstance(g, channels.LogarithmicInput)
    return dict(kind="log" if log else "linear", minimum=g.minimum, maximum=g.maximum,
                num_buckets=g.num_buckets, scale=g.scale if log else 0.0)


def _nd_inputs(scheme, arrays):
    from neuwon_amd import _lib, channels
    arr = (_lib.NwLtiInput * scheme.num_inputs)()
    for k, (g, a) in enumerate(zip(scheme.inputs, arrays)):
        log = isinstance(g, channels.LogarithmicInput)
        arr[k].values, arr[k].index = a.data_ptr(), None
        arr[k].mapping = _lib.NW_INPUT_LOG2 if log else _lib.NW_INPUT_LINEAR
        arr[k].num_nodes = g.num_buckets + 1
        arr[k].minimum, arr[k].maximum = g.minimum, g.maximum
        arr[k].scale = g.scale if log else 0.0
    return arr


@pytest.mark.parametrize("fname,steps", [("glu5.mod", 100), ("mgblock4.mod", 50)])
def test_input_driven_scheme_on_nd_kernel(torch_gpu, fname, steps):
    """200,000 instances of a transmitter-gated scheme (glu5: one logarithmic concentration input;
    mgblock4: concentration and voltage, bilinear tables) with inputs held over the launch,
    against the oracle's node-table kernel fed the same FP32 tables; concentrations span
    1e-7..3e3 mM (below and above the table) and include zeros."""
    from neuwon_amd import _lib, nmodl
    torch = torch_gpu
    scheme = nmodl.load_kinetic(os.path.join(OWN, fname), time_step=0.025, celsius=30.0)
    T32 = np.ascontiguousarray(scheme.node_table(), dtype=np.float32)
    S, M = scheme.num_states, 200_000
    rng = np.random.default_rng(11)
    xs = []
    for n in scheme.input_names:
        if n == "v":
            xs.append(rng.uniform(-110.0, 110.0, M).astype(np.float32))
        else:
            c = 10.0 ** rng.uniform(-7.0, 3.5, M)
            c[rng.random(M) < 0.02] = 0.0
            xs.append(c.astype(np.float32))
    st0 = rng.dirichlet(np.ones(S), M).T.astype(np.float32)
    weights = rng.uniform(0.0, 1.0, S).astype(np.float32)
    gmax = rng.uniform(0.5, 2.0, M).astype(np.float32)
    dev = [torch.as_tensor(x, device="cuda") for x in xs]
    st = torch.as_tensor(np.ascontiguousarray(st0), device="cuda")
    cond = torch.zeros(M, dtype=torch.float32, device="cuda")
    tbl = torch.as_tensor(T32, device="cuda")
    w_d, g_d = torch.as_tensor(weights, device="cuda"), torch.as_tensor(gmax, device="cuda")
    desc = _nd_inputs(scheme, dev)
    _lib.check(_lib.load().nw_lti_advance_nd(
        S, scheme.num_inputs, ctypes.addressof(desc), _lib.ptr(tbl), M, _lib.ptr(st), M, 1.0,
        steps, _lib.ptr(w_d), _lib.ptr(g_d), _lib.ptr(cond), None))
    torch.cuda.synchronize()
    got = st.cpu().numpy().astype(np.float64)
    sample = rng.choice(M, 4000, replace=False)
    ref = st0[:, sample].astype(np.float64)
    grids = [_grid_dict(g) for g in scheme.inputs]
    xin = [x[sample].astype(np.float64) for x in xs]
    for _ in range(steps):
        ref = O.lti_advance_nodes(ref, xin, T32, grids, 1.0)
    err = np.abs(got[:, sample] - ref).max()
    g_ref = gmax[sample].astype(np.float64) * (weights.astype(np.float64) @
                                                ref.astype(np.float32).astype(np.float64))
    g_err = np.abs(cond.cpu().numpy()[sample] - g_ref).max() / np.abs(g_ref).max()
    _say(f"{fname}: {M} instances x {steps} held steps, {scheme.num_inputs} input(s): max abs "
         f"state error {err:.2e}, conductance rel error {g_err:.2e}")
    assert err < 1e-6
    assert g_err < 1e-6
    assert np.abs(got.sum(axis=0) - 1.0).max() < 1e-5


def test_model_with_transmitter_receptor(torch_gpu):
    """RxD_Model(mechanisms=[k4.mod, glu5.mod]): k4 on a soma and dendrite, a GLU5 point process
    (magnitude 5) on the dendrite, 1 mM glutamate for 1 ms played into its 'gluo' attribute.  The
    soma voltage against KineticSim (FP64) fed the same tables, geometry and the recorded
    glutamate sequence."""
    from neuwon_amd import nmodl
    from neuwon_amd.model import RxD_Model, TimeSeries
    dt, celsius, steps = 0.025, 30.0, 400
    files = [os.path.join(OWN, "k4.mod"), os.path.join(OWN, "glu5.mod")]
    m = RxD_Model(time_step=dt, temperature=celsius, initial_voltage=-70.0, mechanisms=files,
                  species=[{"name": "k", "reversal_potential": -85.0}])
    root = m.Neuron([0, 0, 0], 12).root
    dend = root.add_section([150, 0, 0], 2, maximum_segment_length=25)
    segs = [root] + dend
    for s in segs:
        s.insert({"k4": 1})
    syn = dend[3].insert({"GLU5": 5})["GLU5"]
    probe = TimeSeries().record(root, "voltage")
    for _ in range(40):
        m.advance()
    pulse = TimeSeries().square_wave(1.0, 0.0, 2.0)          # 1 mM for 1 ms, then 0
    pulse.play(syn, "gluo", mode="=", immediate=False)
    seen = TimeSeries().record(syn, "gluo", immediate=False)
    for _ in range(40, steps):
        m.advance()
    trace = np.array(probe.get_data())
    glu_after_tick = np.zeros(steps + 1)
    glu_after_tick[41:41 + len(seen)] = seen.get_data()[:steps - 40]
    assert max(seen.get_data()) == 1.0 and len(seen.get_data()) == steps - 40

    n = len(segs)
    parent = np.array([-1] + list(range(n - 1)))
    coords = np.array([np.asarray(s.coordinates, dtype=np.float64) for s in segs])
    g = O.geometry(parent, coords, np.array([s.diameter for s in segs]))
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    P = O.propagator(parent, R, C, 0.5 * dt)
    k4 = nmodl.load_kinetic(files[0], time_step=dt, celsius=celsius)
    glu = nmodl.load_kinetic(files[1], time_step=dt, celsius=celsius)
    syn_row = 4
    gmax_syn = np.zeros(n)
    gmax_syn[syn_row] = 5.0
    mechs = [{"poly": O.node_table_to_polynomial(k4.node_table()),
              "weights": k4.model.conductance_weights("ek"), "gmax": g["surface_area"] * 1e-8,
              "E": -85.0, "states": np.repeat(k4.initial_state(-70.0)[:, None], n, axis=1),
              "conserve": 1.0},
             {"table": glu.node_table(), "grids": [_grid_dict(glu.inputs[0])],
              "inputs": lambda sim: [np.full(n, glu_after_tick[sim.steps])],
              "weights": glu.model.conductance_weights(None) * 1e-6, "gmax": gmax_syn, "E": 0.0,
              "states": np.repeat(glu.initial_state(0.0)[:, None], n, axis=1), "conserve": 1.0}]
    sim = O.KineticSim(P, C, mechs, v0=-70.0, dt=dt)
    ref = [sim.v[0]]
    for k in range(steps):
        sim.advance()
        sim.steps = k + 1
        ref.append(sim.v[0])
    ref = np.array(ref)
    d = np.abs(trace - ref)
    rise = trace.max() - trace[40]
    _say(f"GLU5 synapse: soma EPSP {rise:.2f} mV (oracle {ref.max() - ref[40]:.2f}), max |dv| "
         f"{d.max():.2e} mV, max rel {(d / np.abs(ref)).max():.2e}")
    assert rise > 1.0
    assert (d / np.abs(ref)).max() < 3e-5
