"""Kinetic channel models and their FP64 propagator tables.

* Hodgkin-Huxley gates (nmodl_library/artificial/hh.mod) written as two-state kinetic schemes
  ``~ C <-> O (alpha(v), beta(v))``; the exact propagator of a two-state scheme is closed form.
* Generic linear-time-invariant schemes: ``LTIScheme`` builds the propagator at every node of a
  voltage grid with scipy.linalg.expm, as LTI_Model.make_matrix does
  (neuwon/rxd/lti_sim/lti_model.py:51-65: column j of A is the derivative at unit state e_j,
  P = expm(A dt), every column divided by its sum).
* ``LinearInput`` mirrors lti_sim/inputs.py:72-84 (bucket mapping of a linear input).

Tables are built in FP64 and stored on the device in FP32.  Between grid nodes the propagator is
linearly interpolated: this is the reference's per-bucket polynomial table with terms (1, frac)
(lti_sim/codegen.py:29-37, 68-94) whose coefficients are (T[b], T[b+1] - T[b]).
"""
import math

import numpy as np
import scipy.linalg

HH_REVERSAL = {"na": 60.0, "k": -88.0, "l": -54.3}        # hh.mod:32-34 (mV)
HH_GBAR = {"na": 0.12, "k": 0.036, "l": 0.0003}            # hh.mod:29-31 (S/cm^2)
CM2_TO_UM2 = (1e-6 * 1e-6) / (1e-2 * 1e-2)                 # nmodl/__init__.py:106


def q10_factor(celsius):
    """hh.mod:67."""
    return 3.0 ** ((float(celsius) - 6.3) / 10.0)


def _vtrap(x, y):
    """hh.mod:88-94."""
    x = np.asarray(x, dtype=np.float64)
    r = x / y
    out = np.empty_like(r)
    small = np.abs(r) < 1e-6
    out[small] = y * (1.0 - r[small] / 2.0)
    out[~small] = x[~small] / np.expm1(r[~small])
    return out


def hh_rates(v, celsius=6.3):
    """{gate: (alpha, beta)} in 1/ms including q10; hh.mod:62-86."""
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    q = q10_factor(celsius)
    return {
        "m": (q * 0.1 * _vtrap(-(v + 40.0), 10.0), q * 4.0 * np.exp(-(v + 65.0) / 18.0)),
        "h": (q * 0.07 * np.exp(-(v + 65.0) / 20.0), q * 1.0 / (np.exp(-(v + 35.0) / 10.0) + 1.0)),
        "n": (q * 0.01 * _vtrap(-(v + 55.0), 10.0), q * 0.125 * np.exp(-(v + 65.0) / 80.0)),
    }


def hh_steady_state(v, celsius=6.3):
    """Open fraction of every gate at rest (INITIAL block, hh.mod:48-53)."""
    r = hh_rates(v, celsius)
    return {g: float((a / (a + b))[0]) for g, (a, b) in r.items()}


class LinearInput:
    """Bucket mapping of a linear input; lti_sim/inputs.py:36-48, 72-84."""

    def __init__(self, minimum, maximum, num_buckets):
        self.minimum = float(minimum)
        self.maximum = float(maximum)
        assert self.minimum < self.maximum
        self.num_buckets = int(num_buckets)
        self.range = self.maximum - self.minimum
        self.bucket_frq = self.num_buckets / self.range
        self.bucket_width = self.range / self.num_buckets

    def get_bucket_value(self, value):
        return (value - self.minimum) * self.bucket_frq

    def get_input_value(self, bucket_value):
        return self.minimum + bucket_value * self.bucket_width

    def get_bucket_location(self, value):
        location = self.get_bucket_value(value)
        b = int(location)
        if b < 0:
            return 0, 0.0
        if b >= self.num_buckets:
            return self.num_buckets - 1, 1.0
        return b, location - b

    def nodes(self):
        return self.get_input_value(np.arange(self.num_buckets + 1, dtype=np.float64))


class LogarithmicInput:
    """Bucket mapping of a concentration: log2(x + scale) on linear buckets; lti_sim/inputs.py:86-109
    (minimum 0)."""

    def __init__(self, minimum, maximum, num_buckets, scale):
        self.minimum = float(minimum)
        self.maximum = float(maximum)
        assert self.minimum == 0.0, "logarithmic inputs start at 0"
        assert self.minimum < self.maximum
        self.num_buckets = int(num_buckets)
        self.scale = float(scale)
        self.range = self.maximum - self.minimum
        self.log2_minimum = np.log2(self.minimum + self.scale)
        self.log2_maximum = np.log2(self.maximum + self.scale)
        self.log2_range = self.log2_maximum - self.log2_minimum
        self.bucket_frq = self.num_buckets / self.log2_range
        self.bucket_width = self.log2_range / self.num_buckets

    def get_bucket_value(self, value):
        return (np.log2(value + self.scale) - self.log2_minimum) * self.bucket_frq

    def get_input_value(self, bucket_value):
        return 2.0 ** (self.log2_minimum + bucket_value * self.bucket_width) - self.scale

    def get_bucket_location(self, value):
        location = self.get_bucket_value(value)
        b = int(location) if np.isfinite(location) else -1
        if b < 0:
            return 0, 0.0
        if b >= self.num_buckets:
            return self.num_buckets - 1, 1.0
        return b, location - b

    def nodes(self):
        x = self.get_input_value(np.arange(self.num_buckets + 1, dtype=np.float64))
        x[0], x[-1] = self.minimum, self.maximum        # inputs.py:57-59
        return x


def two_state_propagator(alpha, beta, dt):
    """Exact expm([[-a, b], [a, -b]] dt), state order (C, O), vectorised over a and b.
    Returns (P_OC, P_OO): open' = P_OC * closed + P_OO * open."""
    s = alpha + beta
    e = np.exp(-s * dt)
    oinf = alpha / s
    return oinf * (1.0 - e), oinf + (1.0 - oinf) * e


def hh_gate_table(dt, celsius=6.3, grid=None):
    """FP64 table [3 gates (m, h, n)][node][2]: (a, b) with open' = a + b * open, where
    a = P_OC and b = P_OO - P_OC (the CONSERVE sum of a two-state scheme is 1)."""
    grid = grid or LinearInput(-100.0, 100.0, 2000)
    v = grid.nodes()
    rates = hh_rates(v, celsius)
    out = np.empty((3, len(v), 2))
    for gi, g in enumerate("mhn"):
        poc, poo = two_state_propagator(*rates[g], dt)
        out[gi, :, 0] = poc
        out[gi, :, 1] = poo - poc
    return out


class LTIScheme:
    """A kinetic scheme that is linear in its states, driven by one or two inputs (the membrane
    voltage and/or a concentration).

    ``derivative(*inputs, *states) -> list of d(state)/dt`` (1/ms); ``inputs``: the bucket
    mapping of each input (``LinearInput`` / ``LogarithmicInput``; default one voltage input over
    [-100, 100] mV, 2000 buckets), ``input_names`` their names; ``conserve_sum`` as the CONSERVE
    statement (None when absent); ``rate_matrices`` optionally evaluates the rate matrices of many
    input values at once.  Mirrors LTI_Model (lti_sim/lti_model.py:5-78)."""

    def __init__(self, name, state_names, derivative, conserve_sum=None, time_step=0.1,
                 grid=None, weights=None, inputs=None, input_names=None, rate_matrices=None):
        self.name = str(name)
        self.state_names = list(state_names)
        self.num_states = len(self.state_names)
        self.derivative = derivative
        self.conserve_sum = None if conserve_sum is None else float(conserve_sum)
        self.time_step = float(time_step)
        self.inputs = list(inputs) if inputs is not None else \
            [grid or LinearInput(-100.0, 100.0, 2000)]
        self.input_names = list(input_names) if input_names is not None else ["v"]
        assert len(self.inputs) == len(self.input_names) and 1 <= len(self.inputs) <= 2
        self.grid = self.inputs[0]
        self.weights = None if weights is None else np.asarray(weights, dtype=np.float64)
        self._rate_matrices = rate_matrices
        self._check_is_lti()

    @property
    def num_inputs(self):
        return len(self.inputs)

    def _check_is_lti(self):
        """lti_model.py:40-49."""
        rng = np.random.default_rng(12345)
        for _ in range(3):
            x = [rng.uniform(g.minimum, g.maximum) for g in self.inputs]
            s1 = rng.uniform(0.0, 1.0, self.num_states)
            d1 = np.asarray(self.derivative(*x, *s1), dtype=np.float64)
            d2 = np.asarray(self.derivative(*x, *(2.0 * s1)), dtype=np.float64)
            if not (np.all(np.isfinite(d1)) and
                    np.all(np.abs(d1 - d2 / 2.0) <= 1e-12 * np.maximum(1.0, np.abs(d1)))):
                raise ValueError(f"{self.name}: non-linear system detected")

    def rate_matrix(self, *x):
        S = self.num_states
        A = np.empty((S, S))
        for col in range(S):
            unit = [float(k == col) for k in range(S)]
            A[:, col] = self.derivative(*x, *unit)
        return A

    def _rates(self, *arrays):
        """Rate matrices [..., S, S] at arrays of input values (broadcast together)."""
        arrays = np.broadcast_arrays(*[np.asarray(a, dtype=np.float64) for a in arrays])
        if self._rate_matrices is not None:
            return self._rate_matrices(*arrays)
        flat = [a.ravel() for a in arrays]
        A = np.stack([self.rate_matrix(*[f[k] for f in flat]) for k in range(flat[0].size)])
        return A.reshape(arrays[0].shape + (self.num_states, self.num_states))

    def make_matrix(self, *x, time_step=None):
        """lti_model.py:51-65: expm(A dt), every column divided by its sum."""
        dt = self.time_step if time_step is None else float(time_step)
        P = scipy.linalg.expm(self.rate_matrix(*x) * dt)
        P /= P.sum(axis=0, keepdims=True)
        return P

    def node_rates(self):
        """Rate matrices [node_0](x [node_1])[S][S] at the grid nodes (input 0 slowest)."""
        grids = np.meshgrid(*[g.nodes() for g in self.inputs], indexing="ij")
        return self._rates(*grids)

    def node_table(self):
        """FP64 [node_0](x [node_1])[row][col] propagators on the grid nodes (input 0 slowest),
        on the host (scipy.linalg.expm); ``kinetic.device_node_table`` builds the same table on
        the device."""
        P = scipy.linalg.expm(self.node_rates() * self.time_step)
        P /= P.sum(axis=-2, keepdims=True)
        return P

    def initial_state(self, *x):
        """Steady state at the given input values: expm(A * 1 hour) applied to the uniform state
        (lti_model.py:67-78, evaluated at the given inputs)."""
        if self.conserve_sum is None:
            return np.zeros(self.num_states)
        P = self.make_matrix(*x, time_step=3600e3)
        return P.dot(np.full(self.num_states, self.conserve_sum / self.num_states))


def resized(grid, num_buckets):
    """The same input mapping with round(num_buckets) buckets."""
    n = max(1, int(round(num_buckets)))
    if isinstance(grid, LogarithmicInput):
        return LogarithmicInput(grid.minimum, grid.maximum, n, grid.scale)
    return LinearInput(grid.minimum, grid.maximum, n)


def interpolation_error(scheme, inputs=None, samples_per_bucket=2, seed=0):
    """Largest |P_table(x) - P(x)| of a table on `inputs` (default the scheme's), over random
    inputs in every bucket: P(x) = expm(A(x) dt) with every column divided by its sum (as
    ``make_matrix``), P_table the node propagators interpolated linearly (one input) or
    bilinearly (two) at the input's place in its bucket, as the device kernels apply them.  The
    reference measures its tables by the same statistic, the max abs difference over random
    samples in every bucket (lti_sim/approx.py:33-60 MatrixSamples.sample, 183-201
    measure_error).  Returns the maximum and the per-bucket maxima (shape: the bucket counts)."""
    inputs = list(scheme.inputs if inputs is None else inputs)
    if len(inputs) != scheme.num_inputs:
        raise ValueError(f"{scheme.name}: {len(inputs)} input grids for {scheme.num_inputs} inputs")
    if int(samples_per_bucket) < 1:
        raise ValueError("samples_per_bucket must be at least 1")
    rng = np.random.default_rng(seed)
    dt = scheme.time_step
    mesh = np.meshgrid(*[g.nodes() for g in inputs], indexing="ij")
    Pn = scipy.linalg.expm(scheme._rates(*mesh) * dt)
    Pn /= Pn.sum(axis=-2, keepdims=True)
    shape = tuple(g.num_buckets for g in inputs)
    nb = int(np.prod(shape))
    k = int(samples_per_bucket)
    flat = np.repeat(np.arange(nb), k)
    bidx = np.unravel_index(flat, shape)
    loc = [rng.uniform(size=flat.size) for _ in inputs]
    xs = [g.get_input_value(b + u) for g, b, u in zip(inputs, bidx, loc)]
    Px = scipy.linalg.expm(scheme._rates(*xs) * dt)
    Px /= Px.sum(axis=-2, keepdims=True)
    if len(inputs) == 1:
        b, u = bidx[0], loc[0][:, None, None]
        Pi = Pn[b] + u * (Pn[b + 1] - Pn[b])
    else:
        b0, b1 = bidx
        u0, u1 = loc[0][:, None, None], loc[1][:, None, None]
        lo = Pn[b0, b1] + u1 * (Pn[b0, b1 + 1] - Pn[b0, b1])
        hi = Pn[b0 + 1, b1] + u1 * (Pn[b0 + 1, b1 + 1] - Pn[b0 + 1, b1])
        Pi = lo + u0 * (hi - lo)
    per_bucket = np.abs(Pi - Px).max(axis=(-2, -1)).reshape(nb, k).max(axis=1).reshape(shape)
    return float(per_bucket.max()), per_bucket


def choose_grid(scheme, accuracy=1e-4, start=None, samples_per_bucket=None, seed=0,
                max_buckets=200_000):
    """Bucket counts of the scheme's inputs whose interpolated table meets `accuracy` (the max abs
    error of ``interpolation_error``) with few buckets, by the reference's search for the
    linear form the device kernels use (the reference also searches polynomial degrees by
    timing its generated kernel: lti_sim/optimizer.py:152-192 one input, 213-263 two inputs):
    grow the count by max(1.5, 1.7^log10(error / accuracy)) until the error is met (an error
    that does not fall is an error), then shrink it by 0.9 while the error stays met.  With two
    inputs each growth step tries both inputs x 1.5 and keeps the one with the lower error (both
    when they are within 25 % of each other); each shrink step tries both x 0.9 and keeps the
    lower error.  Returns (inputs, error)."""
    acc = float(accuracy)
    if not 0.0 < acc < 1.0:
        raise ValueError(f"accuracy {accuracy!r}: a max abs propagator error in (0, 1) expected")
    k = samples_per_bucket or (2 if scheme.num_inputs == 1 else 1)
    grids = list(scheme.inputs)

    def measure(counts):
        if int(np.prod([max(1, round(c)) for c in counts])) > max_buckets:
            raise RuntimeError(f"{scheme.name}: accuracy {acc:g} needs more than {max_buckets} "
                               "buckets")
        ins = [resized(g, c) for g, c in zip(grids, counts)]
        return ins, interpolation_error(scheme, ins, k, seed)[0]

    if scheme.num_inputs == 1:
        n = float(start or 10)
        ins, err = measure([n])
        floor = 1
        while err > acc:
            floor = ins[0].num_buckets
            n *= max(1.5, 1.7 ** math.log10(err / acc))
            new_ins, new_err = measure([n])
            if not new_err < err:
                raise RuntimeError(f"{scheme.name}: the error stopped falling at {err:g}")
            ins, err = new_ins, new_err
        while True:
            n *= 0.9
            if round(n) <= floor:
                break
            new_ins, new_err = measure([n])
            if new_err > acc:
                break
            ins, err = new_ins, new_err
        return ins, err
    n = [float(c) for c in (start or (10, 10))]
    ins, err = measure(n)
    while err > acc:
        a_ins, a_err = measure([n[0] * 1.5, n[1]])
        b_ins, b_err = measure([n[0], n[1] * 1.5])
        diff = 2.0 * (a_err - b_err) / (a_err + b_err)
        if diff < -0.25:
            cand = ([n[0] * 1.5, n[1]], a_ins, a_err)
        elif diff > 0.25:
            cand = ([n[0], n[1] * 1.5], b_ins, b_err)
        else:
            both = [n[0] * 1.5, n[1] * 1.5]
            cand = (both,) + measure(both)
        if not cand[2] < err:
            raise RuntimeError(f"{scheme.name}: the error stopped falling at {err:g}")
        n, ins, err = cand
    while True:
        trials = [[n[0] * 0.9, n[1]], [n[0], n[1] * 0.9]]
        # a trial must change a rounded count (and keep every count at one bucket or more)
        trials = [t for t in trials if min(round(t[0]), round(t[1])) >= 1 and
                  (round(t[0]), round(t[1])) != (round(n[0]), round(n[1]))]
        if not trials:
            break
        results = [(t,) + measure(t) for t in trials]
        t, t_ins, t_err = min(results, key=lambda r: r[2])
        if t_err > acc:
            break
        n, ins, err = t, t_ins, t_err
    return ins, err


def voltage_markov_scheme(edges, a, b, time_step=0.1, name="markov", conserve_sum=1.0,
                          num_states=None, weights=None):
    """Scheme with transitions i -> j at rate a_ij exp(b_ij v) (1/ms, v in mV).

    edges: list of (i, j) directed transitions; a, b: arrays of the same length."""
    edges = [(int(i), int(j)) for i, j in edges]
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    S = num_states or (max(max(e) for e in edges) + 1)

    def derivative(v, *states):
        d = [0.0] * S
        for (i, j), ai, bi in zip(edges, a, b):
            flux = ai * math.exp(bi * v) * states[i]
            d[i] -= flux
            d[j] += flux
        return d

    scheme = LTIScheme(name, [f"s{k}" for k in range(S)], derivative, conserve_sum, time_step,
                       weights=weights)
    scheme.edges, scheme.a, scheme.b = edges, a, b

    def rate_matrix(v):
        A = np.zeros((S, S))
        for (i, j), ai, bi in zip(edges, a, b):
            k = ai * math.exp(bi * v)
            A[i, i] -= k
            A[j, i] += k
        return A

    scheme.rate_matrix = rate_matrix
    return scheme


# AMPA13 topology (lti_sim/tests/ampa13.mod, KINETIC kstates): states
# C0..C4 (0..4), D1..D4 (5..8), O1..O4 (9..12); 12 reversible transitions.
AMPA13_STATES = ["C0", "C1", "C2", "C3", "C4", "D1", "D2", "D3", "D4", "O1", "O2", "O3", "O4"]
AMPA13_REACTIONS = [(0, 1), (1, 2), (2, 3), (3, 4),      # C0<->C1 ... C3<->C4
                    (1, 5), (2, 6), (3, 7), (4, 8),      # Ck<->Dk
                    (1, 9), (2, 10), (3, 11), (4, 12)]   # Ck<->Ok
AMPA13_WEIGHTS = [0.0] * 9 + [0.25, 0.5, 0.75, 1.0]      # g = gmax (O4 + .75 O3 + .5 O2 + .25 O1)


def random_ampa13_scheme(seed, time_step=0.1):
    """13-state voltage-dependent scheme on the AMPA13 graph: every transition (both directions)
    has rate a exp(b v), a ~ logU(1e-2, 1) /ms, b ~ U(-0.05, 0.05) /mV."""
    rng = np.random.default_rng(seed)
    edges = []
    for i, j in AMPA13_REACTIONS:
        edges += [(i, j), (j, i)]
    a = np.exp(rng.uniform(np.log(1e-2), np.log(1.0), len(edges)))
    b = rng.uniform(-0.05, 0.05, len(edges))
    return voltage_markov_scheme(edges, a, b, time_step=time_step, name="markov13",
                                 conserve_sum=1.0, num_states=13, weights=AMPA13_WEIGHTS)


def hh_gate_table_theta(dt, celsius=6.3, theta=None, grid=None):
    """Gate tables for rates scaled by theta = [ta_m, tb_m, ta_h, tb_h, ta_n, tb_n] (multipliers of
    alpha and beta): FP64 [3 gates][node][6] = (a, b, da/dta, db/dta, da/dtb, db/dtb) with
    a = P_OC, b = P_OO - P_OC = exp(-(alpha + beta) dt) of the two-state scheme."""
    grid = grid or LinearInput(-100.0, 100.0, 2000)
    theta = np.ones(6) if theta is None else np.asarray(theta, dtype=np.float64)
    v = grid.nodes()
    rates = hh_rates(v, celsius)
    out = np.empty((3, len(v), 6))
    for gi, g in enumerate("mhn"):
        a0, b0 = rates[g]
        ta, tb = theta[2 * gi], theta[2 * gi + 1]
        al, be = ta * a0, tb * b0
        s = al + be
        e = np.exp(-s * dt)
        oinf = al / s
        out[gi, :, 0] = oinf * (1.0 - e)
        out[gi, :, 1] = e
        out[gi, :, 2] = a0 * (be / s ** 2 * (1.0 - e) + oinf * dt * e)
        out[gi, :, 3] = a0 * (-dt * e)
        out[gi, :, 4] = b0 * (-al / s ** 2 * (1.0 - e) + oinf * dt * e)
        out[gi, :, 5] = b0 * (-dt * e)
    return out


def hh_rest_theta(v, celsius=6.3, theta=None):
    """Open fractions at rest for scaled rates, and their derivatives d/dta, d/dtb per gate."""
    theta = np.ones(6) if theta is None else np.asarray(theta, dtype=np.float64)
    rates = hh_rates(v, celsius)
    rest, drest = [], []
    for gi, g in enumerate("mhn"):
        a0, b0 = (float(x[0]) for x in rates[g])
        ta, tb = theta[2 * gi], theta[2 * gi + 1]
        s = ta * a0 + tb * b0
        rest.append(ta * a0 / s)
        drest.append((a0 * tb * b0 / s ** 2, -ta * a0 * b0 / s ** 2))
    return rest, drest
