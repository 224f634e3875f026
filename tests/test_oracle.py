"""Pins of the FP64 oracle: reference closed-form tests, library routines, known answers.

Each check is chosen so that a plausible slip (dt vs dt/2, a unit factor, a transposed matrix,
a wrong sign in the relaxation) fails it.
"""
import bisect
import math
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse.linalg

from oracle import neuwon_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------------------------------------
# helpers: reference-style segment construction (segment.py:27-104)
# ------------------------------------------------------------------------------------------------
def add_section(parent_idx, coords, diam, parents, target, target_diam, max_len):
    """Segment.add_section (segment.py:75-104): indices of the new segments."""
    p_coords = np.array(coords[parent_idx], dtype=float)
    displace = p_coords - np.asarray(target, dtype=float)
    length = float(np.linalg.norm(displace))
    is_sphere = parents[parent_idx] == -1
    has_children = any(p == parent_idx for p in parents)
    if is_sphere or has_children:
        r = 0.5 * diam[parent_idx]
        if r < length - O.EPSILON:
            p_coords -= displace * (r / length)
            length -= r
    div = max(1, int(np.ceil(length / max_len)))
    xs = [np.linspace(p_coords[k], target[k], div + 1) for k in range(3)]
    ds = np.linspace(diam[parent_idx], target_diam, div + 1)
    out = []
    prev = parent_idx
    for i in range(1, div + 1):
        parents.append(prev)
        coords.append(np.array([xs[0][i], xs[1][i], xs[2][i]]))
        diam.append(ds[i])
        prev = len(parents) - 1
        out.append(prev)
    return out


def build(parents, coords, diam, Ra=100.0, Cm=1.0):
    parent = np.array(parents, dtype=np.int64)
    g = O.geometry(parent, np.array(coords), np.array(diam))
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"], Ra, Cm)
    return parent, g, R, C


# ------------------------------------------------------------------------------------------------
# electric closed forms (neuwon/rxd/tests/test_electric.py)
# ------------------------------------------------------------------------------------------------
def test_time_constant():
    """test_electric.py:28-46: one sphere relaxes to 1 - 1/e after one time constant."""
    parent, g, R, C = build([-1], [np.zeros(3)], [20.0])
    v = np.array([0.0])
    gsum = np.array([1e-10])
    d = np.array([1.0])
    tau = C[0] / gsum[0]
    nstep = 1000
    dt = tau / nstep * 1e3
    P = O.propagator(parent, R, C, dt)
    for _ in range(nstep):
        v, _ = O.half_step(v, gsum, d, np.zeros(1), C, P, dt)
    assert v[0] == pytest.approx(1 - 1 / np.e)


def test_time_constant_catches_unit_slip():
    """The same closed form with the exponent in ms instead of s misses by far."""
    parent, g, R, C = build([-1], [np.zeros(3)], [20.0])
    tau = C[0] / 1e-10
    dt = tau / 1000 * 1e3
    v = 0.0
    for _ in range(1000):
        v = 1.0 - (1.0 - v) * math.exp(-dt * 1e-10 / C[0])   # dt left in ms
    assert abs(v - (1 - 1 / np.e)) > 0.3


def _measure_length_constant(diam, rm, max_len, dt, ptol):
    """test_electric.py:48-85."""
    parents, coords, diams = [-1], [np.zeros(3)], [diam]
    section = [0] + add_section(0, coords, diams, parents, [3000, 0, 0], diam, max_len)
    parent, g, R, C = build(parents, coords, diams)
    n = len(parent)
    v = np.zeros(n)
    gsum = g["surface_area"] / rm
    d = np.zeros(n)
    probe = section[len(section) // 2]
    probe_x = coords[probe][0]
    rm_l = 1.0 / (gsum[probe] / g["length"][probe])
    ri = R[probe] / g["length"][probe]
    lam = (rm_l / ri) ** 0.5
    P = O.propagator(parent, R, C, dt)
    for _ in range(round(10 / dt)):
        v[probe] = 1.0
        v, _ = O.half_step(v, gsum, d, np.zeros(n), C, P, dt)
    xcoords = [coords[s][0] for s in section]
    test = section[bisect.bisect_left(xcoords, probe_x + lam)]
    ratio = v[test] / v[probe]
    assert ratio == pytest.approx(1 / np.e, ptol)
    assert coords[test][0] - probe_x == pytest.approx(lam, 0.1)
    assert np.sum(P.data) / P.shape[0] == pytest.approx(1)   # charge conserved (test_electric.py:84-85)


def test_length_constant_1():
    _measure_length_constant(diam=1, rm=4e11, max_len=2, dt=0.01, ptol=0.05)


def test_length_constant_2():
    _measure_length_constant(diam=0.9, rm=2e11, max_len=10, dt=0.025, ptol=0.15)


def test_length_constant_5():
    _measure_length_constant(diam=2, rm=3e11, max_len=20, dt=0.1, ptol=0.25)


def test_inject_current_charge():
    """test_electric.py:132-150: 1 nA for 2 ms adds exactly 2 pC to an isolated sphere."""
    parent, g, R, C = build([-1], [np.zeros(3)], [10.0])
    dt = 0.01
    half = 0.5 * dt
    P = O.propagator(parent, R, C, half)
    v = np.array([-70.0])
    zero = np.zeros(1)
    # no conductance: the state holds
    for _ in range(300):
        v, _ = O.half_step(v, zero, zero, zero, C, P, half)
        v, _ = O.half_step(v, zero, zero, zero, C, P, half)
    assert v[0] == pytest.approx(-70.0)
    q0 = 1e-3 * v[0] * C[0]
    ticks = O.pulse_ticks(2.0, half)
    I = np.array([1e-9])
    for t in range(1, 601):
        v, _ = O.half_step(v, zero, zero, I if t <= ticks else zero, C, P, half)
    q1 = 1e-3 * v[0] * C[0]
    assert q0 + 1e-9 * 2e-3 == pytest.approx(q1)


def test_section_counts():
    """test_segment.py:6-25: section lengths after subtracting the root radius."""
    parents, coords, diams = [-1], [np.zeros(3)], [4.0]
    s1 = add_section(0, coords, diams, parents, [10, 0, 0], 1, 3)
    s2 = add_section(0, coords, diams, parents, [10, 10, 0], 1, 3)
    assert len(s1) == 3 and len(s2) == 5


# ------------------------------------------------------------------------------------------------
# propagator invariants and the reference's own expm routine
# ------------------------------------------------------------------------------------------------
def _random_tree(n, seed):
    rng = np.random.default_rng(seed)
    parent = np.full(n, -1)
    for i in range(1, n):
        parent[i] = rng.integers(0, i)
    coords = np.zeros((n, 3))
    for i in range(1, n):
        coords[i] = coords[parent[i]] + rng.normal(size=3) * 8
    diam = np.exp(rng.uniform(np.log(0.5), np.log(4), n))
    return parent, coords, diam


def test_propagator_matches_scipy_sparse_expm():
    parent, coords, diam = _random_tree(60, 1)
    g = O.geometry(parent, coords, diam)
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    Pd = O.propagator(parent, R, C, 0.05, prune=0.0, method="dense").toarray()
    Ps = O.propagator(parent, R, C, 0.05, prune=0.0, method="sparse").toarray()
    assert np.abs(Pd - Ps).max() < 1e-12
    assert np.abs(Pd.sum(axis=1) - 1).max() < 1e-12           # zero row-sum generator
    assert np.abs(C @ Pd - C).max() / C.max() < 1e-12         # c^T P = c^T (charge)
    assert Pd.min() >= 0.0                                     # exp of a Metzler matrix
    # the generator is the transpose-sensitive part: A^T gives a different propagator
    A = O.coefficient_matrix(parent, R, C, 0.05).toarray()
    assert np.abs(scipy.linalg.expm(A.T) - Pd).max() > 1e-3
    # semigroup: P(dt)^2 = P(2 dt)
    P2 = O.propagator(parent, R, C, 0.1, prune=0.0).toarray()
    assert np.abs(Pd @ Pd - P2).max() < 1e-12


def test_propagator_symmetric_under_capacitance_scaling():
    """A = C^-1 L with L symmetric, so S = C^1/2 P C^-1/2 = expm(C^-1/2 L C^-1/2 dt) is symmetric:
    the identity the symmetric propagator store rests on.  A generator with the coupling
    transposed (A^T = L C^-1) does not have it."""
    parent, coords, diam = _random_tree(70, 3)
    g = O.geometry(parent, coords, diam)
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    assert C.max() / C.min() > 2                       # non-uniform, so the scaling matters
    w = np.sqrt(C / C[0])
    P = O.propagator(parent, R, C, 0.05, prune=0.0).toarray()
    S = w[:, None] * P / w[None, :]
    assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
    A = O.coefficient_matrix(parent, R, C, 0.05).toarray()
    St = w[:, None] * scipy.linalg.expm(A.T) / w[None, :]
    assert np.abs(St - St.T).max() > 1e-3 * np.abs(St).max()
    # the transpose product the adjoint uses: P^T y = W S W^-1 y = C P C^-1 y
    y = np.random.default_rng(0).normal(size=len(C))
    assert np.allclose(P.T @ y, w * (S @ (y / w)), rtol=1e-12, atol=1e-14)
    assert np.allclose(P.T @ y, C * (P @ (y / C)), rtol=1e-12, atol=1e-14)


def test_propagator_pruning_is_reference_rule():
    parent, coords, diam = _random_tree(80, 2)
    g = O.geometry(parent, coords, diam)
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    full = O.propagator(parent, R, C, 0.05, prune=0.0).toarray()
    pr = O.propagator(parent, R, C, 0.05, prune=1e-9).toarray()
    keep = np.abs(full) >= 1e-9
    assert np.array_equal(pr != 0, keep)
    assert np.array_equal(pr[keep], full[keep])


# ------------------------------------------------------------------------------------------------
# HH and LTI
# ------------------------------------------------------------------------------------------------
def test_vtrap_continuous_and_rates_positive():
    v = np.linspace(-100, 100, 20001)
    for a, b in O.hh_alpha_beta(v).values():
        assert np.all(a > 0) and np.all(b > 0)
    x = np.array([-1e-5, -1e-7, 0.0, 1e-7, 1e-5])
    vals = O.vtrap(x, 10.0)
    assert np.abs(np.diff(vals)).max() < 1e-5


def test_gate_exact_matches_closed_form_expm():
    for gate in "mhn":
        for v in (-90.0, -65.0, -40.0, 0.0, 30.0):
            a, b = (float(z) for z in O.hh_alpha_beta(v)[gate])
            P = scipy.linalg.expm(np.array([[-a, b], [a, -b]]) * 0.1)
            for x in (0.0, 0.3, 1.0):
                assert O.gate_exact(x, v, gate, 0.1) == pytest.approx(P[1, 0] * (1 - x) + P[1, 1] * x,
                                                                     rel=1e-12, abs=1e-15)


def test_two_state_table_is_closed_form():
    deriv = O.two_state_derivative(O.hh_gate_rate("n"))
    v = 12.3
    M = O.lti_matrix(deriv, [v], 2, 0.1)
    a, b = (float(z) for z in O.hh_alpha_beta(v)["n"])
    assert np.abs(M - O.two_state_closed_form(a, b, 0.1)).max() < 1e-14
    assert np.abs(M.sum(axis=0) - 1).max() < 1e-15


def test_bucket_location_known_answers():
    """lti_sim/tests/test_inputs.py:5-19."""
    b, f = O.bucket_location(4.0, 4.0, 7.0, 17)
    assert (b, f) == (0, 0.0)
    b, f = O.bucket_location(7 - 1e-12, 4.0, 7.0, 17)
    assert b == 16 and f == pytest.approx(1.0)
    b, f = O.bucket_location(7.5, 4.0, 7.0, 17)
    assert (b, f) == (16, 1.0)
    b, f = O.bucket_location(3.0, 4.0, 7.0, 17)     # more than one bucket below: clamped
    assert (b, f) == (0, 0.0)
    b, f = O.bucket_location(4.0 - 0.5 * 3 / 17, 4.0, 7.0, 17)   # truncation toward zero
    assert b == 0 and f == pytest.approx(-0.5)


def test_lti_polynomial_table_is_linear_interpolation():
    scheme_rates = [(0, 1, 0.3, 0.02), (1, 0, 0.1, -0.03), (1, 2, 0.5, 0.01), (2, 1, 0.2, 0.0)]

    def deriv(v, *s):
        d = [0.0, 0.0, 0.0]
        for i, j, a, bb in scheme_rates:
            f = a * math.exp(bb * v) * s[i]
            d[i] -= f
            d[j] += f
        return d
    node = O.lti_node_table(deriv, 3, 0.1, num_buckets=200)
    poly = O.node_table_to_polynomial(node)
    rng = np.random.default_rng(0)
    v = rng.uniform(-99, 99, 50)
    st = rng.dirichlet(np.ones(3), 50).T
    out_loop = O.lti_advance(st, v, poly, conserve_sum=1.0)
    out_vec = O.lti_advance_vec(st, v, poly, conserve_sum=1.0)
    assert np.abs(out_loop - out_vec).max() < 1e-15
    for k in range(50):
        b, f = O.bucket_location(v[k], -100, 100, 200)
        M = (1 - f) * node[b] + f * node[b + 1]
        ref = M @ st[:, k]
        assert np.abs(ref / ref.sum() - out_loop[:, k]).max() < 1e-14
    # at a node the table reproduces the exact propagator
    vn = np.array([-100 + 37 * 1.0])
    exact = O.lti_matrix(deriv, [vn[0]], 3, 0.1)
    assert np.abs(O.lti_advance_vec(st[:, :1], vn, poly) [:, 0] - exact @ st[:, 0]).max() < 1e-14
    # the C restatement (the kinetic workload's CPU baseline): the same arithmetic, one step
    # held and several steps at moving voltages, inputs beyond the table clamped
    v2 = np.concatenate([v, [-150.0, 150.0, -100.0, 100.0]])
    st2 = rng.dirichlet(np.ones(3), len(v2)).T
    assert np.abs(O.c_lti_run(st2, v2, poly, conserve_sum=1.0)
                  - O.lti_advance_vec(st2, v2, poly, conserve_sum=1.0)).max() < 1e-15
    wave = np.array([0.0, 3.7, -12.5, 40.0])
    ref = st2.copy()
    for w in wave:
        ref = O.lti_advance_vec(ref, v2 + w, poly, conserve_sum=1.0)
    got = O.c_lti_run(st2, v2, poly, wave=wave, conserve_sum=1.0, threads=3)
    assert np.abs(got - ref).max() < 1e-14
    # a planted slip (the previous bucket) is seen
    bad = O.lti_advance_vec(st2, v2 - 1.0, poly, conserve_sum=1.0)
    assert np.abs(O.c_lti_run(st2, v2, poly, conserve_sum=1.0) - bad).max() > 1e-6


def test_golden_hh_trace():
    """neuwon/rxd/tests/test_model.py:35-128 through the oracle (exact gates)."""
    golden = np.loadtxt(os.path.join(GOLDEN, "hh_trace_test_model.txt"))
    parents = [-1] + list(range(10))
    coords = [np.array([-1.0, 0, 7])] + [np.array([float(x), 0, 7]) for x in range(10)]
    parent, g, R, C = build(parents, coords, [5.7] + [1.0] * 10)
    P = O.propagator(parent, R, C, 0.05)
    mag = g["area_at_creation"]          # every hh instance inserted at creation
    gbar = {"na": 0.12e-8 * mag, "k": 0.036e-8 * mag, "l": 0.0003e-8 * mag}
    pulses = []

    def stim(tick):
        I = np.zeros(11)
        for t0, nt in pulses:
            if t0 <= tick < t0 + nt:
                I[0] += 0.1e-9
        return I
    sim = O.HHSim(P, C, gbar, v0=-70.0, dt=0.1, celsius=6.3, stim=stim)
    trace = [sim.v[-1]]
    sim.advance()
    trace.append(sim.v[-1])
    ticks = 1
    while ticks * 0.1 < 20:
        for t in (5, 15):
            if round(t / 0.1) == ticks:
                pulses.append((sim.input_ticks + 1, O.pulse_ticks(1.0, 0.05)))
        sim.advance()
        ticks += 1
        trace.append(sim.v[-1])
    err = np.abs(np.array(trace) - golden)
    assert len(trace) == 201
    assert err.max() < 30.0              # the reference test's bar
    assert err.max() < 3.5               # what the oracle achieves (spike timing shift < 0.02 ms)
    assert err[:51].max() < 0.2          # relaxation before the first stimulus


def test_accumulate_scatter_true_sums():
    rng = np.random.default_rng(3)
    seg = rng.integers(0, 20, 200)
    gi = rng.uniform(0, 1e-9, 200)
    E = rng.uniform(-90, 60, 200)
    sg, num = O.accumulate_scatter(20, seg, gi, E)
    for s in range(20):
        m = seg == s
        assert sg[s] == pytest.approx(gi[m].sum(), rel=1e-13, abs=0)
        assert num[s] == pytest.approx((gi[m] * E[m]).sum(), rel=1e-12, abs=1e-25)
    d = O.driving_voltage(np.array([0.0, 2.0]), np.array([0.0, -130.0]))
    assert d[0] == 0.0 and d[1] == -65.0


def test_stim_hash_bit_identities():
    # avalanche: one flipped input bit changes about half of the output bits
    flips = [bin(int(O.stim_hash(5, 7, t)) ^ int(O.stim_hash(5, 7, t ^ 1))).count("1")
             for t in range(0, 2000, 2)]
    assert 14 < np.mean(flips) < 18
    # onsets are Poisson at the requested rate
    thr = O.poisson_threshold(5.0, 0.05)
    hits = sum(int(O.stim_hash(1, c, t)) < thr for c in range(20) for t in range(20000))
    expect = 20 * 20000 * thr / 2 ** 32
    assert abs(hits - expect) < 4 * math.sqrt(expect)
    assert O.pulse_ticks(1.0, 0.05) == 20


def test_c_restatement_matches_numpy_oracle():
    """The C oracle (CPU baseline) and the numpy oracle run the same step to rounding."""
    import scipy.sparse
    from neuwon_amd import synth, geometry, channels
    cells = synth.population(total=800, seed=2, size_range=(100, 300))
    g = synth.concat(cells)
    geo = geometry.geometry(g["parent"], g["length"], g["diam"])
    R, C = geometry.passive(geo["length"], geo["cross_sectional_area"], geo["area_at_creation"])
    blocks = []
    for c in range(len(cells)):
        r0, r1 = g["cell_ptr"][c], g["cell_ptr"][c + 1]
        blocks.append(O.propagator(cells[c].parent.astype(np.int64), R[r0:r1], C[r0:r1], 0.05,
                                   prune=1e-9))
    P = scipy.sparse.block_diag(blocks, format="csr")
    gbar = {k: channels.HH_GBAR[k] * 1e-8 * geo["surface_area"] for k in ("na", "k", "l")}
    thr = O.poisson_threshold(200.0, 0.05)
    for mode in ("table", "exact"):
        co = O.CHHPopulation(g["cell_ptr"], P, C, gbar, gate_mode=mode,
                             stim=("poisson", 0.5e-9, 3, thr, 20))
        co.advance(60, threads=2)
        for c in range(len(cells)):
            r0, r1 = g["cell_ptr"][c], g["cell_ptr"][c + 1]

            def stim(t, c=c, n=r1 - r0):
                I = np.zeros(n)
                I[0] = 0.5e-9 * O.poisson_pulses_active(3, c, t, thr, 20)
                return I
            sim = O.HHSim(blocks[c], C[r0:r1], {k: v[r0:r1] for k, v in gbar.items()},
                          gate_mode=mode, stim=stim)
            for _ in range(60):
                sim.advance()
            assert np.abs(sim.v - co.v[r0:r1]).max() < 1e-11
            assert np.abs(sim.gates["n"] - co.gates["n"][r0:r1]).max() < 1e-13


# ------------------------------------------------------------------------------------------------
# species diffusion and buffer (the reference's diffusion is a stub, species.py:74-75)
# ------------------------------------------------------------------------------------------------
def test_diffusion_gaussian_spread():
    """A point release on a long uniform cable spreads with variance 2 D t."""
    n, L, d, D = 801, 1.0, 1.0, 0.6
    parent = np.arange(-1, n - 1)
    length = np.full(n, L)
    xarea = np.full(n, 0.25 * np.pi * d * d)
    vol = xarea * L
    k = O.diffusion_coupling(parent, length, xarea, D)
    P = O.diffusion_propagator(parent, k, vol, 0.05, prune=0.0)
    c = np.zeros(n)
    c[n // 2] = 1.0
    steps = 400                      # 20 ms
    for _ in range(steps):
        c = P.dot(c)
    x = (np.arange(n) - n // 2) * L
    var = np.sum(c * x * x) / np.sum(c)
    assert var == pytest.approx(2 * D * steps * 0.05, rel=0.02)
    assert np.sum(c) == pytest.approx(1.0, rel=1e-12)


def test_diffusion_conserves_mass_and_uniform():
    parent, coords, diam = _random_tree(70, 4)
    g = O.geometry(parent, coords, diam)
    k = O.diffusion_coupling(parent, g["length"], g["cross_sectional_area"], 0.6)
    P = O.diffusion_propagator(parent, k, g["volume"], 0.05, prune=0.0).toarray()
    assert np.abs(P.sum(axis=1) - 1).max() < 1e-13
    assert np.abs(g["volume"] @ P - g["volume"]).max() / g["volume"].max() < 1e-13
    assert P.min() >= 0


def test_buffer_equilibrium_and_mass():
    kon, koff, btot = 0.1, 0.01, 100.0
    ca, cab = np.array([1.0]), np.array([0.0])
    vol = np.array([3.0])
    m0 = np.sum(vol * (ca + cab))
    sim = O.CaSim(scipy.sparse.identity(1, format="csr"), vol, ca, cab, kon=kon, koff=koff, btot=btot)
    for _ in range(20000):
        sim.advance()
    assert sim.mass() == pytest.approx(m0, rel=1e-12)
    assert sim.cab[0] == pytest.approx(btot * sim.ca[0] / (sim.ca[0] + koff / kon), rel=1e-9)
    # the exact two-state update equals the closed-form expm of its rate matrix
    a, b = kon * 0.7, koff
    P = scipy.linalg.expm(np.array([[-a, b], [a, -b]]) * 0.1)
    got = O.buffer_exact(np.array([30.0]), 0.7, 0.1, kon, koff, btot)[0]
    assert got == pytest.approx(P[1, 0] * (btot - 30.0) + P[1, 1] * 30.0, rel=1e-13)


# ------------------------------------------------------------------------------------------------
# adjoint of the staggered step: the FP64 reverse pass against central finite differences
# ------------------------------------------------------------------------------------------------
def test_oracle_adjoint_matches_finite_differences():
    from neuwon_amd import synth, geometry, channels
    cell = synth.single_cell(40, seed=4)[0]
    parent = cell.parent.astype(np.int64)
    geo = geometry.geometry(parent, cell.length, cell.diam)
    R, C = geometry.passive(geo["length"], geo["cross_sectional_area"], geo["area_at_creation"])
    P = O.propagator(parent, R, C, 0.025, prune=1e-9)
    gb = {k: channels.HH_GBAR[k] * 1e-8 * geo["surface_area"] for k in ("na", "k", "l")}
    prob = O.HHAdjointProblem(P, C, 0.1e-9, dt=0.05)
    rng = np.random.default_rng(1)
    hidden = {k: gb[k] * np.exp(rng.uniform(np.log(0.5), np.log(2), len(C))) for k in gb}
    T = 250
    target, _ = prob.forward(hidden, {}, T)
    assert target.max() > 20.0            # the trajectory spikes
    loss, grad = prob.gradient(gb, {}, target)
    h = 1e-6
    for k, i in (("na", 3), ("k", 11), ("l", 0)):
        gp = {kk: vv.copy() for kk, vv in gb.items()}
        gm = {kk: vv.copy() for kk, vv in gb.items()}
        d = h * gb[k][i]
        gp[k][i] += d
        gm[k][i] -= d
        fd = (prob.forward(gp, {}, T, target)[1] - prob.forward(gm, {}, T, target)[1]) / (2 * d)
        assert grad[k][i] == pytest.approx(fd, rel=1e-4)
    for g in "mhn":
        for j in range(2):
            tp = {g: (1 + h, 1.0) if j == 0 else (1.0, 1 + h)}
            tm = {g: (1 - h, 1.0) if j == 0 else (1.0, 1 - h)}
            fd = (prob.forward(gb, tp, T, target)[1] - prob.forward(gb, tm, T, target)[1]) / (2 * h)
            assert grad["theta"][g][j] == pytest.approx(fd, rel=1e-4)


def test_theta_table_derivatives_are_expm_frechet():
    """The derivative columns of the rate-multiplier tables against scipy.linalg.expm_frechet of
    the 2-state generator Q = [[-alpha, beta], [alpha, -beta]] (state order closed, open):
    a = P_oc, b = P_oo - P_oc, and d/d(t_alpha) carries the factor alpha_0."""
    dt = 0.05
    tbl = O.hh_theta_tables(dt, 6.3, {"m": (1.1, 0.9), "h": (1.0, 1.2), "n": (0.8, 1.05)})
    v = -100.0 + np.arange(2001) * 0.1
    rates = O.hh_alpha_beta(v, 6.3)
    th = {"m": (1.1, 0.9), "h": (1.0, 1.2), "n": (0.8, 1.05)}
    dQa = np.array([[-1.0, 0.0], [1.0, 0.0]])
    dQb = np.array([[0.0, 1.0], [0.0, -1.0]])
    for g in "mhn":
        a0, b0 = rates[g]
        for k in range(0, 2001, 97):
            al, be = th[g][0] * a0[k], th[g][1] * b0[k]
            Q = np.array([[-al, be], [al, -be]]) * dt
            P, La = scipy.linalg.expm_frechet(Q, dQa * dt)
            _, Lb = scipy.linalg.expm_frechet(Q, dQb * dt)
            want = [P[1, 0], P[1, 1] - P[1, 0], a0[k] * La[1, 0], a0[k] * (La[1, 1] - La[1, 0]),
                    b0[k] * Lb[1, 0], b0[k] * (Lb[1, 1] - Lb[1, 0])]
            np.testing.assert_allclose(tbl[g][k], want, rtol=1e-9, atol=1e-15)


def test_theta_tables_match_oracle():
    from neuwon_amd import channels
    th = [1.1, 0.9, 1.0, 1.2, 0.8, 1.05]
    got = channels.hh_gate_table_theta(0.05, 6.3, th)
    ref = O.hh_theta_tables(0.05, 6.3, {"m": (1.1, 0.9), "h": (1.0, 1.2), "n": (0.8, 1.05)})
    for gi, g in enumerate("mhn"):
        np.testing.assert_allclose(got[gi], ref[g], rtol=1e-12, atol=1e-18)
    # derivative columns against differences of the value columns
    eps = 1e-6
    up = channels.hh_gate_table_theta(0.05, 6.3, [1.1 + eps, 0.9, 1, 1.2, 0.8, 1.05])
    dn = channels.hh_gate_table_theta(0.05, 6.3, [1.1 - eps, 0.9, 1, 1.2, 0.8, 1.05])
    fd = (up[0, :, :2] - dn[0, :, :2]) / (2 * eps)
    np.testing.assert_allclose(got[0, :, 2:4], fd, rtol=1e-5, atol=1e-12)


def test_kinetic_sim_constant_conductance_closed_form():
    """KineticSim with one mechanism frozen open (zero rates): a single compartment relaxes to
    the reversal potential exactly as exp(-t g / C), one half step late (the first half step
    sees no conductance)."""
    C, g, E, dt, v0 = np.array([2e-12]), 3e-10, -54.0, 0.1, -70.0
    zero = lambda v, a, b: [0.0, 0.0]
    poly = O.node_table_to_polynomial(O.lti_node_table(zero, 2, dt))
    mech = {"poly": poly, "weights": [0.0, 1.0], "gmax": np.array([g]), "E": E,
            "states": np.array([[0.0], [1.0]]), "conserve": 1.0}
    sim = O.KineticSim(scipy.sparse.csr_matrix(np.eye(1)), C, [mech], v0=v0, dt=dt)
    for k in range(1, 41):
        sim.advance()
        t_s = (k * dt - 0.5 * dt) * 1e-3
        want = E + (v0 - E) * np.exp(-t_s * g / C[0])
        assert sim.v[0] == pytest.approx(want, rel=1e-12)


# ---------------------------------------------------------------------------------------------
# Input mappings and node-table kinetics (one or two inputs)
# ---------------------------------------------------------------------------------------------
def _grid(g):
    from neuwon_amd import channels
    log = isinstance(g, channels.LogarithmicInput)
    return dict(kind="log" if log else "linear", minimum=g.minimum, maximum=g.maximum,
                num_buckets=g.num_buckets, scale=g.scale if log else 0.0)


def test_input_position_known_answers():
    """lti_sim/tests/test_inputs.py:5-68 against the oracle's mapping and the product's grids."""
    import random
    from neuwon_amd import channels
    lin = channels.LinearInput(4, 7, 17)
    log = channels.LogarithmicInput(0, 2e3, 37, scale=1e-3)
    for g, lo, hi in ((lin, 4, 7), (log, 0, 2e3)):
        for x, want_b, want_f in ((lo, 0, 0.0), (hi - 1e-12, g.num_buckets - 1, 1.0)):
            b, f = O.input_position(np.array([x]), _grid(g))
            assert int(b[0]) == want_b and float(f[0]) == pytest.approx(want_f)
            assert g.get_bucket_location(x)[0] == want_b
            assert g.get_bucket_location(x)[1] == pytest.approx(want_f)
    # bisect_inputs(5, 6): the midpoint in bucket space
    mid = log.get_input_value(0.5 * (log.get_bucket_value(5) + log.get_bucket_value(6)))
    assert 5.25 < mid < 5.49
    assert lin.get_input_value(0.5 * (lin.get_bucket_value(5) + lin.get_bucket_value(6))) == \
        pytest.approx(5.5)
    # inverse round trips, small concentrations included (test_inputs.py:38-68)
    rng = random.Random(5)
    lin = channels.LinearInput(11, 57, 107)
    log = channels.LogarithmicInput(0, 57, 107, scale=1e-3)
    for g in (lin, log):
        trials = [rng.uniform(g.minimum, g.maximum) for _ in range(100)] + [g.minimum, g.maximum]
        if g is log:
            trials += [7.0 ** -x for x in range(100)]
        for x in trials:
            u = g.get_bucket_value(x)
            assert g.get_input_value(u) == pytest.approx(x, abs=1e-12)
            assert 0.0 <= u <= g.num_buckets + 1e-9
            b, f = O.input_position(np.array([x]), _grid(g))
            assert int(b[0]) == min(int(u), g.num_buckets - 1)
    # below the range and NaN take node 0 (the reference has no lower clamp for concentrations)
    b, f = O.input_position(np.array([-1.0, np.nan]), _grid(log))
    assert list(b) == [0, 0] and list(f) == [0.0, 0.0]


def test_lti_advance_nodes_pins():
    """Voltage-only tables: the node-table step equals the polynomial-table kernel; at grid nodes
    the interpolated step is the exact expm of that node (one and two inputs); between nodes it
    stays within the measured interpolation error of the exact propagator."""
    import os
    from neuwon_amd import channels, nmodl
    own = os.path.join(os.path.dirname(__file__), "golden", "kinetic")
    rng = np.random.default_rng(8)
    # (a) voltage input
    sc = channels.random_ampa13_scheme(seed=3, time_step=0.1)
    T = sc.node_table()
    M = 500
    st = rng.dirichlet(np.ones(13), M).T
    v = rng.uniform(-120.0, 120.0, M)
    a = O.lti_advance_nodes(st, [v], T, [_grid(sc.grid)], 1.0)
    b = O.lti_advance_vec(st, v, O.node_table_to_polynomial(T), conserve_sum=1.0)
    assert np.abs(a - b).max() < 1e-14
    # (b) nodes of two inputs (log concentration x linear voltage), small grids
    path = os.path.join(own, "mgblock4.mod")
    grids = [channels.LogarithmicInput(0.0, 1000.0, 60, 1e-4), channels.LinearInput(-100.0, 100.0, 40)]
    sc2 = nmodl.load_kinetic(path, time_step=0.05, celsius=30.0, inputs=grids)
    T2 = sc2.node_table()
    G2 = [_grid(g) for g in grids]
    nodes = [g.nodes() for g in grids]
    for i, j in ((0, 0), (17, 9), (33, 21), (59, 39), (60, 40)):
        x = (nodes[0][i], nodes[1][j])
        P = sc2.make_matrix(*x)
        got = O.lti_advance_nodes(np.eye(4), [np.full(4, x[0]), np.full(4, x[1])], T2, G2)
        assert np.abs(got - P).max() < 1e-13, (i, j)
    # (c) between nodes: the measured error of the default grids
    for f, bound in (("glu5.mod", 5e-6), ("mgblock4.mod", 6e-5)):
        sc3 = nmodl.load_kinetic(os.path.join(own, f), time_step=0.025, celsius=30.0)
        T3 = sc3.node_table()
        G3 = [_grid(g) for g in sc3.inputs]
        worst = 0.0
        for _ in range(300):
            x = [rng.uniform(-90, 40) if n == "v" else 10 ** rng.uniform(-5, 1)
                 for n in sc3.input_names]
            P = sc3.make_matrix(*x)
            S = sc3.num_states
            got = O.lti_advance_nodes(np.eye(S), [np.full(S, xi) for xi in x], T3, G3)
            worst = max(worst, np.abs(got - P).max())
        assert worst < bound, (f, worst)
