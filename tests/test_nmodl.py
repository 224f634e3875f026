"""The NMODL kinetic front end on the CPU: the reference's own known answers for its test files
(neuwon/rxd/lti_sim/tests/test_lti_sim.py:9-64), every file of the kinetic library parsed, and the
parsed rate matrices against the same rates computed by hand from the files' formulas."""
import math
import os

import numpy as np
import pytest
import scipy.linalg

from neuwon_amd import channels, nmodl

REF = "/root/reference"
LTI_TESTS = os.path.join(REF, "neuwon/rxd/lti_sim/tests")
LIBRARY = os.path.join(REF, "nmodl_library/kinetic")
OWN = os.path.join(os.path.dirname(__file__), "golden", "kinetic")

needs_reference = pytest.mark.skipif(not os.path.isdir(LTI_TESTS),
                                     reason="the reference's .mod files are not on this machine")


def _rate_matrix(km, *inputs):
    S = km.num_states
    return np.array([km.derivative(*inputs, *[float(i == j) for i in range(S)])
                     for j in range(S)]).T


def _steady_state(km, *inputs):
    """lti_model.py:67-78: expm(A * 1 hour), columns normalised, applied to the uniform state."""
    P = scipy.linalg.expm(_rate_matrix(km, *inputs) * 3600e3)
    P /= P.sum(axis=0, keepdims=True)
    return dict(zip(km.states, P.dot(np.full(km.num_states, km.conserve_sum / km.num_states))))


@needs_reference
def test_reference_nav11_known_answers():
    """test_lti_sim.py:9-25 (v over [-200, 200], initial value the minimum)."""
    km = nmodl.KineticModel(os.path.join(LTI_TESTS, "Nav11.mod"), celsius=37.0)
    assert km.name == "na11a"
    assert km.inputs == ["v"] and km.num_inputs == 1
    assert km.num_states == 6 and km.states == sorted(km.states)
    assert km.conserve_sum == 1.0
    assert km.parameters["C1C2b2"] == 18
    init = _steady_state(km, -200.0)
    assert init.pop("C1") == pytest.approx(1)
    for value in init.values():
        assert value == pytest.approx(0, abs=1e-6)


@needs_reference
def test_reference_ampa13_known_answers():
    """test_lti_sim.py:27-42 (the concentration input C, initial value 0)."""
    km = nmodl.KineticModel(os.path.join(LTI_TESTS, "ampa13.mod"), celsius=37.0, inputs=["C"])
    assert km.name == "AMPA13"
    assert km.num_states == 13 and km.states == sorted(km.states)
    assert km.conserve_sum == 1.0
    assert km.parameters["Rb1"] == 800
    init = _steady_state(km, 0.0)
    assert init.pop("C0") == pytest.approx(1)
    for value in init.values():
        assert value == pytest.approx(0, abs=1e-6)


@needs_reference
def test_reference_nmda_inputs():
    """test_lti_sim.py:44-60: two inputs, sorted (C, v)."""
    km = nmodl.KineticModel(os.path.join(LTI_TESTS, "NMDA.mod"), celsius=37.0)
    assert km.name == "NMDA_Mg"
    assert km.inputs == ["C", "v"]
    assert km.num_states == 10 and km.conserve_sum == 1.0
    assert km.parameters["Rb"] == 10e-3
    init = _steady_state(km, 0.0, -200.0)
    assert init.pop("UMg") == pytest.approx(1)
    for value in init.values():
        assert value == pytest.approx(0, abs=1e-6)
    # two inputs: a logarithmic concentration grid and a linear voltage grid (bilinear tables)
    sc = nmodl.load_kinetic(os.path.join(LTI_TESTS, "NMDA.mod"), time_step=0.1,
                            inputs=[channels.LogarithmicInput(0, 1000, 8, 1e-4),
                                    channels.LinearInput(-100, 100, 8)])
    assert sc.input_names == ["C", "v"] and sc.node_table().shape == (9, 9, 10, 10)


@needs_reference
def test_reference_nonlinear_detected():
    """test_lti_sim.py:62-65."""
    with pytest.raises(nmodl.NmodlError, match="non-linear"):
        nmodl.KineticModel(os.path.join(LTI_TESTS, "nonlinear.mod"))


@needs_reference
def test_library_parses():
    """Every file of nmodl_library/kinetic: a linear scheme, CONSERVE 1, rate matrices with zero
    column sums and non-negative off-diagonal rates at -80, 0 and +40 mV."""
    files = sorted(f for f in os.listdir(LIBRARY) if f.endswith(".mod"))
    assert len(files) == 34
    for f in files:
        km = nmodl.KineticModel(os.path.join(LIBRARY, f), celsius=37.0)
        assert km.conserve_sum == 1.0, f
        if km.inputs != ["v"]:
            assert km.point_process, f            # AMPA, NMDA: transmitter-gated synapses
            continue
        for v in (-80.0, 0.0, 40.0):
            A = _rate_matrix(km, v)
            assert np.abs(A.sum(axis=0)).max() < 1e-12 * max(1.0, np.abs(A).max()), f
            off = A - np.diag(np.diag(A))
            assert off.min() >= 0.0, f


@needs_reference
def test_nav11_rates_by_hand():
    """The Nav1.1 generator at 37 degC restated from the file's formulas (rates2 and the six
    reactions, nmodl_library/kinetic/Nav11.mod) against the parsed one."""
    km = nmodl.KineticModel(os.path.join(LIBRARY, "Nav11.mod"), celsius=37.0)
    q10 = 3.0 ** ((37.0 - 20.0) / 10.0)

    def r2(v, b, vv, k):
        return b / (1.0 + math.exp((v - vv) / k))
    for v in (-95.0, -40.0, 12.5):
        k = {
            ("C1", "C2"): q10 * r2(v, 18, -7, -10),
            ("C2", "C1"): q10 * (r2(v, 3, -37, 10) + r2(v, 18, -7, -10)),
            ("C2", "O1"): q10 * r2(v, 18, -7, -10),
            ("O1", "C2"): q10 * (r2(v, 3, -37, 10) + r2(v, 18, -7, -10)),
            ("C2", "O2"): q10 * r2(v, 0.08, -10, -15),
            ("O2", "C2"): q10 * (r2(v, 2, -50, 7) + r2(v, 0.2, -20, -10)),
            ("O1", "I1"): q10 * (r2(v, 8, -37, 13) + r2(v, 17, -7, -15)),
            ("I1", "O1"): q10 * r2(v, 0.00001, -37, 10),
            ("I1", "C1"): q10 * r2(v, 0.21, -61, 7),
            ("C1", "I1"): q10 * r2(v, 0.3, -61, -5.5),
            ("I1", "I2"): q10 * r2(v, 0.0015, -90, -5),
            ("I2", "I1"): q10 * r2(v, 0.0075, -90, 15),
        }
        A = np.zeros((6, 6))
        idx = {s: i for i, s in enumerate(km.states)}
        for (a, b), rate in k.items():
            A[idx[b], idx[a]] += rate
            A[idx[a], idx[a]] -= rate
        np.testing.assert_allclose(_rate_matrix(km, v), A, rtol=1e-13, atol=1e-15)


@needs_reference
def test_kv11_rates_by_hand():
    """Kv1.1 (nmodl_library/kinetic/Kv11.mod): INITIAL sets F and A, rates() uses them, and the
    reactions carry temperature factors; CS -> OS and OS -> CS restated at 30 degC."""
    celsius = 30.0
    km = nmodl.KineticModel(os.path.join(LIBRARY, "Kv11.mod"), celsius=celsius)
    F, A = 96.485, 8.134 * (celsius + 273.15)
    Zc, Vc, Hvc, Hkc, kc = 1.6756, -17.6122, 0.91934, 0.12166, 0.030179
    ki, Hki = 0.00025091, 0.44916
    for v in (-60.0, -10.0, 30.0):
        kco = kc * math.exp(Zc * F / A * (v - Vc - Hvc * (celsius - 25))) * math.exp(Hkc * (celsius - 25))
        koc = kc * math.exp(-Zc * F / A * (v - Vc - Hvc * (celsius - 25))) * math.exp(Hkc * (celsius - 25))
        M = _rate_matrix(km, v)
        i = {s: n for n, s in enumerate(km.states)}
        assert M[i["OS"], i["CS"]] == pytest.approx(kco, rel=1e-13)
        # OS leaves to CS (koc), to IC1 (ki e^{Hki (T-25)}) and to IN
        assert M[i["CS"], i["OS"]] == pytest.approx(koc, rel=1e-13)
        assert M[i["IC1"], i["OS"]] == pytest.approx(ki * math.exp(Hki * (celsius - 25)), rel=1e-13)
    # conductance: g = gbar OS with gbar 36 mS/cm2, ik = g (v - ek) 1e-3 -> 0.036 S/cm2 in OS
    w = km.conductance_weights("ek")
    assert w[km.states.index("OS")] == pytest.approx(0.036, rel=1e-14)
    assert np.count_nonzero(w) == 1


@needs_reference
def test_nav11_conductance_weights():
    km = nmodl.KineticModel(os.path.join(LIBRARY, "Nav11.mod"), celsius=37.0)
    w = km.conductance_weights("ena")
    want = {"O1": 0.1, "O2": 0.1}
    for s, x in zip(km.states, w):
        assert x == pytest.approx(want.get(s, 0.0), abs=1e-15)
    assert km.current_ion == "na"


def test_own_files_parse():
    """The kinetic files under tests/golden/kinetic (written for the GPU tests)."""
    na = nmodl.KineticModel(os.path.join(OWN, "na6.mod"), celsius=22.0)
    assert na.name == "na6" and na.num_states == 6 and na.inputs == ["v"]
    assert na.current_ion == "na"
    k = nmodl.KineticModel(os.path.join(OWN, "k4.mod"), celsius=22.0)
    assert k.name == "k4" and k.num_states == 4 and k.current_ion == "k"
    for km in (na, k):
        for v in (-90.0, -20.0, 35.0):
            A = _rate_matrix(km, v)
            assert np.abs(A.sum(axis=0)).max() < 1e-12 * np.abs(A).max()
            assert (A - np.diag(np.diag(A))).min() >= 0.0
    scheme = nmodl.load_kinetic(os.path.join(OWN, "k4.mod"), time_step=0.025, celsius=22.0)
    assert isinstance(scheme, channels.LTIScheme) and scheme.num_states == 4
    tbl = scheme.node_table()
    assert np.abs(tbl.sum(axis=1) - 1.0).max() < 1e-12        # columns of every node sum to one


def test_unsupported_constructs_raise(tmp_path):
    src = """NEURON { SUFFIX bad }
STATE { a b }
KINETIC k { ~ a <-> b (1, 2) }
BREAKPOINT { SOLVE k METHOD sparse }
NET_RECEIVE (w) { }
"""
    f = tmp_path / "bad.mod"
    f.write_text(src)
    with pytest.raises(nmodl.NmodlError, match="NET_RECEIVE"):
        nmodl.KineticModel(str(f))
    g = tmp_path / "deriv.mod"
    g.write_text("NEURON { SUFFIX d }\nSTATE { x }\nDERIVATIVE s { x' = -x }\n")
    with pytest.raises(nmodl.NmodlError):
        nmodl.KineticModel(str(g))


@needs_reference
def test_ampa_library_rates_by_hand():
    """nmodl_library/kinetic/AMPA.mod at 37 degC: the transmitter-gated generator restated from
    the file's reactions and Q10 factors; conductance weights of the four open states (uS)."""
    km = nmodl.KineticModel(os.path.join(LIBRARY, "AMPA.mod"), celsius=37.0)
    assert km.inputs == ["gluo"] and km.point_process and km.num_states == 13
    q = 2.4 ** ((37.0 - 22.0) / 10.0)
    Rb, Ru = [800, 600, 400, 200], [30, 40, 60, 80]
    Rd, Rr = [.25, .25, 1, 1], [0.05, 0.05, 0.022, 0.022]
    Ro, Rc = [3, 4, 4, 4], [1.5, 1, 1, 1.5]
    idx = {s: i for i, s in enumerate(km.states)}
    for glu in (0.0, 3e-4, 0.3, 10.0):
        A = np.zeros((13, 13))

        def add(a, b, kf, kr):
            A[idx[b], idx[a]] += kf
            A[idx[a], idx[a]] -= kf
            A[idx[a], idx[b]] += kr
            A[idx[b], idx[b]] -= kr
        for k in range(4):
            add(f"C{k}", f"C{k + 1}", Rb[k] * glu * q, Ru[k] * q)
            add(f"C{k + 1}", f"D{k + 1}", Rd[k] * q, Rr[k] * q)
            add(f"C{k + 1}", f"O{k + 1}", Ro[k] * q, Rc[k] * q)
        np.testing.assert_allclose(_rate_matrix(km, glu), A, rtol=1e-13, atol=1e-13)
    w = km.conductance_weights(None)
    want = {"O1": 0.25, "O2": 0.5, "O3": 0.75, "O4": 1.0}
    for s, x in zip(km.states, w):
        assert x == pytest.approx(1e-6 * 20 * want.get(s, 0.0), abs=1e-18)
    assert km.output_variable == "gzero" and km.output_ion == "zero"


@needs_reference
def test_library_transmitter_schemes_load():
    """AMPA (one concentration input) and NMDA (concentration and voltage) tabulate on the
    logarithmic / linear grids; the vectorised rate matrices equal the scalar ones."""
    from neuwon_amd import channels
    for f, inputs in (("AMPA.mod", ["gluo"]), ("NMDA.mod", ["C", "v"])):
        grids = [channels.LinearInput(-100, 100, 20) if n == "v" else
                 channels.LogarithmicInput(0, 1000, 30, 1e-4) for n in inputs]
        sc = nmodl.load_kinetic(os.path.join(LIBRARY, f), time_step=0.025, celsius=37.0,
                                inputs=grids)
        assert sc.input_names == inputs
        T = sc.node_table()
        assert T.shape == tuple(g.num_buckets + 1 for g in grids) + (sc.num_states,) * 2
        assert np.abs(T.sum(axis=-2) - 1.0).max() < 1e-12
        km = sc.model
        xs = [np.array([-30.0, 10.0]) if n == "v" else np.array([1e-3, 0.5]) for n in inputs]
        A = km.rate_matrices(*xs)
        for k in range(2):
            np.testing.assert_allclose(A[k], _rate_matrix(km, *[x[k] for x in xs]), rtol=1e-14,
                                       atol=1e-14)


def test_own_receptor_files():
    """tests/golden/kinetic/glu5.mod (transmitter read from outside) and mgblock4.mod (a POINTER
    concentration and the voltage): inputs, generators by hand, point-process weights and the
    reversal potentials their currents imply."""
    glu = nmodl.KineticModel(os.path.join(OWN, "glu5.mod"), celsius=30.0)
    assert glu.inputs == ["gluo"] and glu.point_process
    t = 2.0 ** ((30.0 - 22.0) / 10.0)
    idx = {s: i for i, s in enumerate(glu.states)}
    A = np.zeros((5, 5))

    def add(A, a, b, kf, kr):
        A[idx[b], idx[a]] += kf
        A[idx[a], idx[a]] -= kf
        A[idx[a], idx[b]] += kr
        A[idx[b], idx[b]] -= kr
    c = 0.7
    add(A, "R", "RA", 2 * 5 * c * t, 2 * t)
    add(A, "RA", "RA2", 5 * c * t, 2 * 2 * t)
    add(A, "RA2", "O", 6 * t, 1.2 * t)
    add(A, "RA2", "D", 0.4 * t, 0.03 * t)
    np.testing.assert_allclose(_rate_matrix(glu, c), A, rtol=1e-14, atol=1e-14)
    w = glu.conductance_weights(None)
    assert w[idx["O"]] == pytest.approx(0.002, rel=1e-14) and np.count_nonzero(w) == 1  # 2 nS
    assert glu.current_reversal() == pytest.approx(0.0, abs=1e-12)
    mg = nmodl.KineticModel(os.path.join(OWN, "mgblock4.mod"), celsius=30.0)
    assert mg.inputs == ["C", "v"] and mg.point_process
    ix = {s: i for i, s in enumerate(mg.states)}
    for C, v in ((0.0, -70.0), (0.2, 10.0)):
        B = np.zeros((4, 4))
        for a, b, kf, kr in (("U", "B", 8 * C, 0.05), ("B", "O", 0.3, 0.1),
                             ("O", "OB", 1.2 * math.exp(-v / 16), 0.4 * math.exp(v / 16))):
            B[ix[b], ix[a]] += kf
            B[ix[a], ix[a]] -= kf
            B[ix[a], ix[b]] += kr
            B[ix[b], ix[b]] -= kr
        np.testing.assert_allclose(_rate_matrix(mg, C, v), B, rtol=1e-14, atol=1e-14)
    assert mg.conductance_weights(None)[ix["O"]] == pytest.approx(8e-4, rel=1e-14)   # 800 pS
    assert mg.current_reversal() == pytest.approx(3.0, rel=1e-12)
    mech = nmodl.kinetic_mechanism(os.path.join(OWN, "glu5.mod"), 0.025, 30.0, {})
    assert mech.point_process and mech.input_names == ["gluo"] and mech.reversal == 0.0
    assert mech.scheme.weights[idx["O"]] == pytest.approx(2e-9, rel=1e-14)            # Siemens


def _kinetic_fixture(name):
    import json
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "kinetic_tables", f"{name}.npz"),
                allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    S = meta["num_states"]
    shape = tuple(meta["node_shape"])
    A = np.zeros((int(np.prod(shape)), S, S))
    A[:, z["pattern"][:, 0], z["pattern"][:, 1]] = z["values"]
    return A.reshape(shape + (S, S)), meta


@pytest.mark.parametrize("name", ["Nav11", "ampa13"])
def test_kinetic_table_fixtures(name):
    """The rate matrices the GPU tests tabulate: rate matrices (non-negative transitions, columns
    summing to zero), the grid of the front end, and -- where the reference is on this machine --
    equal to a fresh parse of its file."""
    A, meta = _kinetic_fixture(name)
    S = meta["num_states"]
    off = ~np.eye(S, dtype=bool)
    assert (A[..., off] >= 0.0).all()
    assert np.abs(A.sum(axis=-2)).max() <= 1e-9 * np.abs(A).max()
    g = meta["grids"][0]
    assert A.shape[:-2] == (g["num_buckets"] + 1,)
    assert meta["conserve_sum"] == 1.0
    path = os.path.join(REF, meta["source"])
    if os.path.exists(path):
        sc = nmodl.load_kinetic(path, time_step=meta["time_step"], celsius=meta["celsius"])
        assert sc.state_names == meta["states"] and sc.input_names == meta["inputs"]
        np.testing.assert_array_equal(sc.node_rates(), A)
