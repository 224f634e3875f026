"""The drop-in RxD_Model API on the GPU, following the reference's own tests
(neuwon/rxd/tests/test_electric.py, test_model.py, test_segment.py)."""
import bisect
import os
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(f"[gpu-model] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("store", [None, "symmetric", "general"])
def test_time_constant(torch_gpu, store):
    """test_electric.py:28-46 through Segment._advance_electric (the three propagator stores)."""
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(propagator_store=store)
    root = m.Neuron([0, 0, 0], 20).root
    m.advance()   # builds the model on the device
    root.voltage = 0
    root.driving_voltage = 1
    root.sum_conductance = 1e-10
    root.sum_current = 0
    tau = root.get_time_constant()
    nstep = 1000
    dt = tau / nstep * 1e3
    for _ in range(nstep):
        m.Segment._advance_electric(dt)
    _say(f"time constant: v(tau) = {root.voltage!r}")
    assert root.voltage == pytest.approx(1 - 1 / np.e, rel=1e-6)
    m.check()


@pytest.mark.parametrize("store", [None, "symmetric", "general"])
def test_inject_current(torch_gpu, store):
    """test_electric.py:132-150 (the three propagator stores)."""
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(time_step=.01, propagator_store=store)
    root = m.Neuron([0, 0, 0], 10).root
    init_v = root.voltage
    while m.clock() < 3:
        m.advance()
    assert init_v == pytest.approx(root.voltage)
    init_Q = 1e-3 * root.voltage * root.capacitance
    root.inject_current(1, 2)
    delta_Q = 1e-9 * 2e-3
    while m.clock() < 6:
        m.advance()
    m.check()
    new_Q = 1e-3 * root.voltage * root.capacitance
    _say(f"inject: dQ/expected - 1 = {(new_Q - init_Q) / delta_Q - 1:.2e}")
    assert init_Q + delta_Q == pytest.approx(new_Q, rel=1e-5)


@pytest.mark.parametrize("store", [None, "symmetric", "general"])
def test_length_constant(torch_gpu, store):
    """test_electric.py:48-94 (case 1): attenuation 1/e at one length constant."""
    from neuwon_amd.model import RxD_Model
    diam, rm, max_len, dt, ptol = 1, 4e11, 2, .01, .05
    m = RxD_Model(propagator_store=store)
    root = m.Neuron([0, 0, 0], diam).root
    section = [root] + root.add_section([3000, 0, 0], diam, max_len)
    m.advance()
    for seg in section:
        seg.voltage = 0
        seg.driving_voltage = 0
        seg.sum_conductance = seg.surface_area / rm
        seg.sum_current = 0
    probe = section[len(section) // 2]
    probe_x = probe.coordinates[0]
    lam = probe.get_length_constant()
    for _ in range(round(10 / dt)):
        probe.voltage = 1
        m.Segment._advance_electric(dt)
    xs = [seg.coordinates[0] for seg in section]
    test = section[bisect.bisect_left(xs, probe_x + lam)]
    ratio = test.voltage / probe.voltage
    _say(f"length constant {lam:.1f} um, attenuation {ratio:.4f}")
    assert ratio == pytest.approx(1 / np.e, ptol)


def _golden_model(mechanism, store=None):
    from neuwon_amd.model import RxD_Model, TimeSeries
    m = RxD_Model(time_step=.1, temperature=6.3, mechanisms=[mechanism],
                  propagator_store=store)
    hh = m.mechanisms[mechanism]
    root = tip = m.Neuron([-1, 0, 7], 5.7).root
    hh(root)
    for x in range(10):
        tip = m.Segment(tip, [x, 0, 7], 1)
        hh(tip)
    x = TimeSeries().record(tip, 'voltage')
    m.advance()
    m.check()
    while m.clock() < 20:
        for t in [5, 15]:
            if m.clock.is_now(t):
                root.inject_current(.1, 1)
        m.advance()
    m.check()
    return np.array(x.get_data())


@pytest.mark.parametrize("mechanism,store", [("hh", None), ("hh_kinetic", None),
                                            ("hh", "symmetric"), ("hh", "general")])
def test_model_hh_golden(torch_gpu, mechanism, store):
    """test_model.py:35-128 through the drop-in API, against the golden trace and the oracle."""
    golden = np.loadtxt(os.path.join(GOLDEN, "hh_trace_test_model.txt"))
    v = _golden_model(mechanism, store)
    assert len(v) == 201
    err = np.abs(v - golden)
    _say(f"{mechanism}: max |v - golden| = {err.max():.3f} mV")
    assert err.max() < 30.0
    assert err[:51].max() < 0.2
    # oracle of the same scheme
    parents = [-1] + list(range(10))
    coords = [np.array([-1.0, 0, 7])] + [np.array([float(x), 0, 7]) for x in range(10)]
    g = O.geometry(np.array(parents), np.array(coords), np.array([5.7] + [1.0] * 10))
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    P = O.propagator(np.array(parents), R, C, 0.05)
    mag = g["area_at_creation"]
    gbar = {"na": 0.12e-8 * mag, "k": 0.036e-8 * mag, "l": 0.0003e-8 * mag}
    pulses = []

    def stim(tick):
        I = np.zeros(11)
        for t0, nt in pulses:
            if t0 <= tick < t0 + nt:
                I[0] += 0.1e-9
        return I
    sim = O.HHSim(P, C, gbar, gate_mode="exact" if mechanism == "hh" else "table", stim=stim)
    ref = [sim.v[-1]]
    sim.advance()
    ref.append(sim.v[-1])
    ticks = 1
    while ticks * 0.1 < 20:
        for t in (5, 15):
            if round(t / 0.1) == ticks:
                pulses.append((sim.input_ticks + 1, O.pulse_ticks(1.0, 0.05)))
        sim.advance()
        ticks += 1
        ref.append(sim.v[-1])
    ref = np.array(ref)
    d = np.abs(v - ref)
    _say(f"{mechanism}: max |v - oracle| = {d.max():.2e} mV")
    assert d.max() < 5e-3


def test_insert_api(torch_gpu):
    """test_segment.py:51-69: insert by dict, name and list; duplicates allowed."""
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(mechanisms=['hh'])
    root = m.Neuron([0, 0, 0], 10).root
    segs = root.add_section([100, 0, 0], 3, maximum_segment_length=10)
    root.insert({})
    inst = root.insert({'hh': 1})
    segs[2].insert('hh')
    segs[3].insert(['hh'])
    segs[4].insert({'hh': 1})
    segs[4].insert({'hh': 1})
    assert inst['hh'].get_database_class().get_name() == 'hh'
    for _ in range(20):
        m.advance()
    m.check()
    g = m.database.get_data("Segment.sum_conductance")
    # two hh instances on segs[4]: twice the conductance per unit area of segs[3]
    ratio = (g[segs[4].get_unstable_index()] / segs[4].surface_area) / \
        (g[segs[3].get_unstable_index()] / segs[3].surface_area)
    assert 1.8 < ratio < 2.2, ratio


def test_lti_mechanism_api(torch_gpu):
    """A kinetic mechanism given as an LTI scheme runs through nw_lti_advance."""
    from neuwon_amd import channels
    from neuwon_amd.model import RxD_Model, LTIMechanism
    scheme = channels.random_ampa13_scheme(seed=2)
    mech = LTIMechanism(scheme, gmax_s_per_cm2=0.002, reversal=0.0, name="markov13")
    m = RxD_Model(time_step=0.1, temperature=6.3, mechanisms=[mech, 'hh'])
    root = m.Neuron([0, 0, 0], 10).root
    segs = root.add_section([200, 0, 0], 2, maximum_segment_length=20)
    for s in [root] + segs:
        s.insert({'hh': 1, 'markov13': 1})
    for _ in range(50):
        m.advance()
    m.check()
    states = np.stack([m.database.get_data(f"markov13.s{k}") for k in range(13)])
    assert np.abs(states.sum(axis=0) - 1).max() < 1e-5
    v = m.database.get_data("Segment.voltage")
    assert np.all(np.isfinite(v)) and v.max() < 60


def test_interacting_mechanisms(torch_gpu):
    """test_segment.py:69-113: no spontaneous AP at start-up; an AP after 0.1 nA x 0.5 ms; a second
    mechanism holding a pointer to the hh instance (local.mod: hh' = -hh / 100, solved exactly
    per step) decays its magnitude until the same stimulus no longer fires.  The plugin gets the hh
    instance through other_mechanisms() (mechanisms.py:33-40, segment.py:141-160)."""
    import math
    from neuwon_amd.model import RxD_Model, Mechanism, TimeSeries

    class Local(Mechanism):
        instances = []

        @classmethod
        def initialize(cls, model):
            cls.model = model
            cls.instances = []

        @classmethod
        def get_name(cls):
            return "local"

        @classmethod
        def other_mechanisms(cls):
            return ["hh"]

        def __init__(self, segment, outside, magnitude, hh):
            self.hh = hh
            type(self).instances.append(self)

        @classmethod
        def advance(cls):
            decay = math.exp(-cls.model.get_time_step() / 100.0)
            for inst in cls.instances:
                inst.hh.magnitude = inst.hh.magnitude * decay

    m = RxD_Model(time_step=0.1, temperature=6.3, mechanisms=["hh", Local],
                  species=[{"name": "na", "reversal_potential": +60},
                           {"name": "k", "reversal_potential": -88},
                           {"name": "l", "reversal_potential": -54.3}])
    root = m.Neuron([0, 0, 0], 10).root
    inst = root.insert({"hh": 1, "local": 1})
    assert inst["local"].hh is inst["hh"]
    mag0 = inst["hh"].magnitude
    probe = TimeSeries().record(root, "voltage", discard_after=30)
    while m.clock() < 20:
        m.advance()
    quiet = max(probe.get_data())
    m.clock.reset()
    root.inject_current(.1, .5)
    while m.clock() < 20:
        m.advance()
    fired = max(probe.get_data())
    while m.clock() < 500:
        m.advance()
    decayed = inst["hh"].magnitude / mag0
    m.clock.reset()
    root.inject_current(.1, .5)
    while m.clock() < 20:
        m.advance()
    after = max(probe.get_data())
    _say(f"interacting: quiet max {quiet:.1f} mV, stimulated max {fired:.1f} mV, magnitude "
         f"x{decayed:.4f}, stimulated after decay max {after:.1f} mV")
    assert quiet < -20
    assert fired > 20
    assert decayed == pytest.approx(math.exp(-5.2), rel=1e-3)   # 5,200 steps of 0.1 ms
    assert after < -20
    m.check()


def test_advance_from_worker_thread(torch_gpu):
    """The reference's GUI advances the model from a worker thread and reads arrays afterwards
    (gui/model_runner/model_thread.py:100-123): the same calls from a thread give the same
    voltages as on the main thread."""
    import threading
    from neuwon_amd.model import RxD_Model

    def run(steps):
        m = RxD_Model(time_step=0.1, temperature=6.3, mechanisms=["hh"])
        root = m.Neuron([0, 0, 0], 10).root
        for s in [root] + root.add_section([150, 0, 0], 2, maximum_segment_length=15):
            s.insert("hh")
        root.inject_current(.2, 2)
        for _ in range(steps):
            m.advance()
        return m.database.get_data("Segment.voltage").copy()

    main = run(150)
    out = {}
    t = threading.Thread(target=lambda: out.setdefault("v", run(150)))
    t.start()
    t.join(timeout=120)
    assert not t.is_alive()
    assert np.array_equal(out["v"], main)


@pytest.mark.parametrize("mechanism", ["hh", "hh_kinetic"])
def test_mechanism_on_a_subset_of_segments(torch_gpu, mechanism):
    """A mechanism instance reads the voltage of its own segment: a passive neuron created first
    (its rows come first) does not change the spike of an hh neuron created after it, and the
    passive neuron's own voltage follows only its own (absent) stimulus."""
    from neuwon_amd.model import RxD_Model, TimeSeries

    def run(with_passive):
        m = RxD_Model(time_step=0.025, temperature=6.3, mechanisms=[mechanism])
        if with_passive:
            other = m.Neuron([0, 0, 0], 10).root
            other.add_section([80, 0, 0], 2, maximum_segment_length=20)
        root = m.Neuron([500, 0, 0], 12).root
        dend = root.add_section([500, 100, 0], 2, maximum_segment_length=25)
        for s in [root] + dend[1::2]:
            s.insert(mechanism)
        probe = TimeSeries().record(root, "voltage")
        root.inject_current(0.2, 1.0)
        for _ in range(400):
            m.advance()
        v = m.database.get_data("Segment.voltage")
        passive = v[:len(v) - len(dend) - 1] if with_passive else np.zeros(0)
        return np.array(probe.get_data()), passive
    alone, _ = run(False)
    beside, passive = run(True)
    d = np.abs(alone - beside).max()
    _say(f"{mechanism} on a subset: peak {alone.max():.1f} mV alone, {beside.max():.1f} mV beside a "
         f"passive neuron (max |dv| {d:.2e} mV); passive neuron in [{passive.min():.3f}, "
         f"{passive.max():.3f}] mV")
    assert alone.max() > 20
    assert d < 1e-3
    assert np.abs(passive - (-70.0)).max() < 1e-3
