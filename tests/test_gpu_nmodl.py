"""Kinetic .mod files through the NMODL front end on the GPU (files written for these tests,
tests/golden/kinetic): the LTI kernel on the parsed tables, and a model carrying both channels
against the FP64 oracle step with the same tables."""
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


def test_model_with_kinetic_mod_files(torch_gpu):
    """RxD_Model(mechanisms=[na6.mod, k4.mod]) on a soma and a six-segment dendrite, 0.3 nA for
    1 ms from t = 1 ms, dt 0.025 ms x 400 steps at 22 degC: every segment's voltage and the
    channels' states against KineticSim (FP64) fed the same tables, geometry and stimulus."""
    from neuwon_amd import channels, nmodl
    from neuwon_amd.model import RxD_Model, TimeSeries
    torch = torch_gpu
    dt, celsius, steps = 0.025, 22.0, 400
    files = [os.path.join(OWN, "na6.mod"), os.path.join(OWN, "k4.mod")]
    species = [{"name": "na", "reversal_potential": 50.0}, {"name": "k", "reversal_potential": -85.0}]
    m = RxD_Model(time_step=dt, temperature=celsius, initial_voltage=-65.0, mechanisms=files,
                  species=species)
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
        scheme = nmodl.load_kinetic(f, time_step=dt, celsius=celsius)
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
