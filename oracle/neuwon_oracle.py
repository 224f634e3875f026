"""FP64 CPU restatement of NEUWON's per-timestep advance -- TEST INFRASTRUCTURE ONLY.

This module is the oracle: tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it as the checker.  The product (``neuwon_amd``) never imports it.

Every function restates one piece of the reference algorithm in plain numpy/scipy FP64 and cites
the reference file:line it follows (paths relative to the reference repository root).  Loops are
written the way the reference reads, not for speed.

Pinning (what each function is checked against, see ``tests/test_oracle.py``):
  * geometry / passive R,C      -- closed-form areas, reference section counts, HH golden trace
  * propagator                  -- scipy.sparse.linalg.expm (the reference's own routine), P.1 = 1,
                                   c^T P = c^T, RC relaxation 1-1/e, injected charge dQ = I.t
  * half_step                   -- RC closed form, charge injection, length-constant attenuation
  * hh gates (exact)            -- HH golden trace (neuwon/rxd/tests/test_model.py:77-113)
  * input_position              -- bucket known answers of lti_sim/tests/test_inputs.py:5-68
                                   (linear and logarithmic mappings, ends, inverse round trips)
  * lti_advance_nodes           -- equals lti_advance_vec on voltage tables; at grid nodes the
                                   step is scipy.linalg.expm exactly; interpolation error bounds
  * LTI tables / kernel         -- 2-state closed-form expm (exact), scipy.linalg.expm, column sums,
                                   bucket-mapping known answers (lti_sim/tests/test_inputs.py:5-35)
  * accumulation                -- np.add.at brute force
  * stimulus hash               -- integer identities (bit-exact uint32 arithmetic)
"""
import math

import numpy as np
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

EPSILON = float(np.finfo(np.float64).eps)   # neuwon/database/dtypes.py:4
NULL = -1                                    # root marker (reference uses 0xFFFFFFFF, dtypes.py:6)


# ---------------------------------------------------------------------------------------------
# Geometry and passive electric properties
# ---------------------------------------------------------------------------------------------
def geometry(parent, coords, diam):
    """Per-segment geometry, segments given in creation order (parent index < child index).

    Follows neuwon/rxd/neuron/geometry.py:75-155 and segment.py:27-39:
      * root is a sphere: length = 2/3 d, area = pi d^2, volume = 4/3 pi r^3
      * other segments are cylinders; the first child of a cylinder is "primary", all other
        children (and every child of a sphere) are secondary branches (geometry.py:80-87)
      * secondary length subtracts the parent's radius unless that would envelop it (112-127)
      * surface area: lateral cylinder + tip disk for leaves - attachment disks of non-primary
        children (129-144).  The reference computes the *capacitance* once when the segment is
        created (electric.py:45-54 called from segment.py:37), when it has no children yet, so
        ``area_at_creation`` (no attachment disks, leaf tip included) is kept separately from the
        final ``surface_area`` (recomputed on the parent whenever a child is added, geometry.py:91).
    """
    parent = np.asarray(parent, dtype=np.int64)
    coords = np.asarray(coords, dtype=np.float64)
    diam = np.asarray(diam, dtype=np.float64)
    n = len(parent)
    length = np.empty(n)
    primary = np.zeros(n, dtype=bool)
    children = [[] for _ in range(n)]
    area_at_creation = np.empty(n)
    surface_area = np.empty(n)

    def is_sphere(i):
        return parent[i] == NULL

    def secondary_length(i):
        L = length[i]
        if not primary[i]:
            r = 0.5 * diam[parent[i]]
            if not (L < r + EPSILON):
                L -= r
        return L

    def area_of(i):
        d = diam[i]
        if is_sphere(i):
            a = math.pi * d * d
        else:
            a = math.pi * d * secondary_length(i)
            if len(children[i]) == 0:
                a += 0.25 * math.pi * d * d
        for c in children[i]:
            if not primary[c]:
                a -= 0.25 * math.pi * min(d, diam[c]) ** 2
        return a

    for i in range(n):
        p = parent[i]
        if p != NULL:
            children[p].append(i)
        if is_sphere(i):
            primary[i] = False
            length[i] = (2.0 / 3.0) * diam[i]
        else:
            primary[i] = False if is_sphere(p) else (len(children[p]) < 2)
            length[i] = float(np.linalg.norm(coords[i] - coords[p]))
        surface_area[i] = area_of(i)
        area_at_creation[i] = surface_area[i]
        if p != NULL:
            surface_area[p] = area_of(p)
    xarea = 0.25 * math.pi * diam ** 2
    volume = np.empty(n)
    for i in range(n):
        if is_sphere(i):
            volume[i] = (4.0 / 3.0) * math.pi * (0.5 * diam[i]) ** 3
        else:
            volume[i] = xarea[i] * secondary_length(i)
    return dict(length=length, primary=primary, surface_area=surface_area,
                area_at_creation=area_at_creation, cross_sectional_area=xarea, volume=volume)


def passive(length, xarea, area, cytoplasmic_resistance=100.0, membrane_capacitance=1.0):
    """Axial resistance (Ohm) and capacitance (F).  neuwon/rxd/neuron/electric.py:49-54."""
    Ra = cytoplasmic_resistance * 1e4
    Cm = membrane_capacitance * 1e-14
    return Ra * np.asarray(length) / np.asarray(xarea), Cm * np.asarray(area)


# ---------------------------------------------------------------------------------------------
# Electric propagator and half step
# ---------------------------------------------------------------------------------------------
def coefficient_matrix(parent, R, C, time_step_ms):
    """Sparse A*dt (dt in seconds) of dV/dt = A V.  electric.py:85-111.

    Per edge (child c, parent p): A[p,c]=+dt/(r_c C_p), A[c,c]=-dt/(r_c C_c),
    A[c,p]=+dt/(r_c C_c), A[p,p]=-dt/(r_c C_p); duplicates sum (COO -> CSC)."""
    dt = time_step_ms * 1e-3
    src, dst, coef = [], [], []
    for child, p in enumerate(parent):
        if p == NULL:
            continue
        r = R[child]
        src += [child, child, p, p]
        dst += [p, child, child, p]
        coef += [+dt / (r * C[p]), -dt / (r * C[child]), +dt / (r * C[child]), -dt / (r * C[p])]
    n = len(parent)
    return scipy.sparse.csc_matrix((coef, (dst, src)), shape=(n, n), dtype=np.float64)


def propagator(parent, R, C, time_step_ms, prune=EPSILON, method="dense"):
    """P = expm(A*dt), entries with |P| < prune dropped.  electric.py:79-117.

    ``method="sparse"`` calls scipy.sparse.linalg.expm exactly as the reference does
    (electric.py:112); ``"dense"`` uses scipy.linalg.expm (same result to ~1e-14, faster for the
    small trees the tests use).  Returns a scipy CSR matrix."""
    A = coefficient_matrix(parent, R, C, time_step_ms)
    if method == "sparse":
        P = scipy.sparse.linalg.expm(A).tocsr()
    else:
        P = scipy.sparse.csr_matrix(scipy.linalg.expm(A.toarray()))
    P.data[np.abs(P.data) < prune] = 0.0
    P.eliminate_zeros()
    return P


def half_step(v, g, d, I, C, P, time_step_ms):
    """One electric half step; ``time_step_ms`` is the half step (dt/2).  electric.py:57-76.

    dv = dt_ms*I/C ; exponent = -dt_s*g/C ; alpha = exp(exponent);
    v <- P (d - (d - v) alpha + dv); integral_v = dt_s d - exponent (d-v) alpha + 0.5 dt_s dv."""
    dt = time_step_ms * 1e-3
    dv_currents = time_step_ms * I / C
    exponent = -dt * g / C
    alpha = np.exp(exponent)
    diff_v = d - v
    v_new = P.dot(d - diff_v * alpha + dv_currents)
    integral_v = dt * d - exponent * diff_v * alpha + (0.5 * dt * dv_currents)
    return v_new, integral_v


# ---------------------------------------------------------------------------------------------
# Hodgkin-Huxley rates and the two gate integrators
# ---------------------------------------------------------------------------------------------
def vtrap(x, y):
    """nmodl_library/artificial/hh.mod:88-94."""
    x = np.asarray(x, dtype=np.float64)
    small = np.abs(x / y) < 1e-6
    with np.errstate(divide="ignore", invalid="ignore"):
        big = x / (np.exp(x / y) - 1.0)
    return np.where(small, y * (1.0 - x / y / 2.0), big)


def hh_alpha_beta(v, celsius=6.3):
    """Per-gate (alpha, beta) in /ms including q10.  hh.mod:62-86 (q10 at 67)."""
    v = np.asarray(v, dtype=np.float64)
    q10 = 3.0 ** ((celsius - 6.3) / 10.0)
    am = 0.1 * vtrap(-(v + 40.0), 10.0)
    bm = 4.0 * np.exp(-(v + 65.0) / 18.0)
    ah = 0.07 * np.exp(-(v + 65.0) / 20.0)
    bh = 1.0 / (np.exp(-(v + 35.0) / 10.0) + 1.0)
    an = 0.01 * vtrap(-(v + 55.0), 10.0)
    bn = 0.125 * np.exp(-(v + 65.0) / 80.0)
    return {"m": (q10 * am, q10 * bm), "h": (q10 * ah, q10 * bh), "n": (q10 * an, q10 * bn)}


def gate_inf(v, gate, celsius=6.3):
    a, b = hh_alpha_beta(v, celsius)[gate]
    return a / (a + b)


def gate_exact(x, v, gate, dt, celsius=6.3):
    """Exact linear-ODE update x' = (inf-x)/tau over dt.  solver.py:48-55 applied to hh.mod:55-60:
    x(t+dt) = -c0/c1 + (c0 + c1 x) exp(c1 dt)/c1 with c1 = -1/tau, c0 = inf/tau."""
    a, b = hh_alpha_beta(v, celsius)[gate]
    s = a + b
    tau = 1.0 / s
    inf = a / s
    c1 = -1.0 / tau
    c0 = inf / tau
    return -c0 / c1 + (c0 + c1 * x) * np.exp(c1 * dt) / c1


# ---------------------------------------------------------------------------------------------
# LTI kinetic tables and the table kernel
# ---------------------------------------------------------------------------------------------
def lti_matrix(derivative, inputs, num_states, time_step):
    """Exact propagator of an LTI system at fixed inputs.  lti_sim/lti_model.py:51-65:
    column j of A is derivative(inputs, e_j); P = expm(A dt); every column divided by its sum."""
    A = np.empty((num_states, num_states))
    for col in range(num_states):
        state = [float(x == col) for x in range(num_states)]
        A[:, col] = derivative(*inputs, *state)
    M = scipy.linalg.expm(A * time_step)
    for col in range(num_states):
        M[:, col] *= 1.0 / sum(M[:, col].flat)
    return M


def lti_node_table(derivative, num_states, time_step, vmin=-100.0, vmax=100.0, num_buckets=2000):
    """Propagator at every node of the voltage grid (num_buckets+1 nodes), shape [node][row][col]."""
    nodes = vmin + np.arange(num_buckets + 1) * ((vmax - vmin) / num_buckets)
    return np.stack([lti_matrix(derivative, [v], num_states, time_step) for v in nodes])


def node_table_to_polynomial(node_table):
    """Linear interpolation between grid nodes written in the reference's polynomial table layout
    [bucket][col][row][term] with terms (1, frac): c0 = T[b], c1 = T[b+1] - T[b].
    Layout: lti_sim/codegen.py:29-37 (transpose to column-major at 35)."""
    T = np.asarray(node_table)
    c0 = T[:-1]
    c1 = T[1:] - T[:-1]
    tbl = np.stack([c0, c1], axis=-1)          # [bucket][row][col][term]
    return np.ascontiguousarray(tbl.transpose(0, 2, 1, 3))   # [bucket][col][row][term]


def bucket_location(value, minimum, maximum, num_buckets):
    """Reference kernel's bucket lookup for a linear input.  codegen.py:44-59:
    input = (v - min) * nb/range; b = int(input) (truncates toward zero);
    b > nb-1 -> (nb-1, 1.0); b < 0 -> (0, 0.0); else frac = input - b."""
    bucket_frq = num_buckets / (maximum - minimum)
    u = (value - minimum) * bucket_frq
    b = int(u)
    if b > num_buckets - 1:
        return num_buckets - 1, 1.0
    if b < 0:
        return 0, 0.0
    return b, u - b


def lti_advance(states, v, poly_table, minimum=-100.0, maximum=100.0, conserve_sum=None):
    """The generated per-instance LTI kernel.  codegen.py:39-106.

    states: [S][M] array (SoA), v: [M] input voltages, poly_table: [bucket][col][row][term].
    Returns the new states [S][M]."""
    S = states.shape[0]
    nb = poly_table.shape[0]
    terms = poly_table.shape[3]
    out = np.empty_like(states)
    for inst in range(states.shape[1]):
        b, frac = bucket_location(float(v[inst]), minimum, maximum, nb)
        basis = [frac ** k for k in range(terms)]
        scratch = np.zeros(S)
        for col in range(S):
            s = states[col, inst]
            for row in range(S):
                poly = 0.0
                for k in range(terms):
                    poly += basis[k] * poly_table[b, col, row, k]
                scratch[row] += poly * s
        if conserve_sum is not None:
            scratch *= conserve_sum / scratch.sum()
        out[:, inst] = scratch
    return out


def lti_advance_vec(states, v, poly_table, minimum=-100.0, maximum=100.0, conserve_sum=None):
    """Vectorised form of ``lti_advance`` (same arithmetic per instance, summed in the same
    column-then-row order); checked against the loop form in tests/test_oracle.py."""
    S, M = states.shape
    nb = poly_table.shape[0]
    terms = poly_table.shape[3]
    bucket_frq = nb / (maximum - minimum)
    u = (np.asarray(v, dtype=np.float64) - minimum) * bucket_frq
    b = np.trunc(u).astype(np.int64)
    frac = u - b
    hi = b > nb - 1
    lo = b < 0
    b = np.where(hi, nb - 1, np.where(lo, 0, b))
    frac = np.where(hi, 1.0, np.where(lo, 0.0, frac))
    basis = np.stack([frac ** k for k in range(terms)])          # [term][M]
    scratch = np.zeros((S, M))
    for col in range(S):
        s = states[col]
        for row in range(S):
            poly = np.zeros(M)
            for k in range(terms):
                poly += basis[k] * poly_table[b, col, row, k]
            scratch[row] += poly * s
    if conserve_sum is not None:
        scratch *= conserve_sum / scratch.sum(axis=0)
    return scratch


def input_position(x, grid):
    """Fractional node position of input values x on a grid (dict: kind "linear" or "log",
    minimum, maximum, num_buckets, scale), with the reference's truncating bucket lookup and
    clamping (codegen.py:42-59): linear u = (x - min) * nb / range (inputs.py:72-84); log
    u = (log2(x + scale) - log2(min + scale)) * nb / log2range (inputs.py:86-109).  Returns
    (bucket, fraction) arrays; values below the range (and NaN) take (0, 0)."""
    x = np.asarray(x, dtype=np.float64)
    nb = int(grid["num_buckets"])
    if grid["kind"] == "log":
        lo = np.log2(grid["minimum"] + grid["scale"])
        hi = np.log2(grid["maximum"] + grid["scale"])
        with np.errstate(invalid="ignore", divide="ignore"):
            u = (np.log2(x + grid["scale"]) - lo) * (nb / (hi - lo))
    else:
        u = (x - grid["minimum"]) * (nb / (grid["maximum"] - grid["minimum"]))
    below = ~(u >= 0.0)
    above = u >= nb
    b = np.where(below | above, 0, np.trunc(np.where(below, 0.0, u))).astype(np.int64)
    f = np.where(below | above, 0.0, u - b)
    b = np.where(above, nb - 1, b)
    f = np.where(above, 1.0, f)
    return b, f


def lti_advance_nodes(states, xs, table, grids, conserve_sum=None):
    """One step of a kinetic scheme whose propagators are tabulated at grid nodes of one or two
    inputs: table [n0](x [n1])[row][col]; the propagator of each instance is interpolated
    linearly (one input) or bilinearly (two) between the surrounding nodes, then applied
    column by column as codegen.py:81-102 does (CONSERVE renormalisation included).
    states [S][M]; xs: list of input arrays [M]."""
    T = np.asarray(table, dtype=np.float64)
    S, M = states.shape
    pos = [input_position(x, g) for x, g in zip(xs, grids)]
    if len(pos) == 1:
        (b0, f0), = pos
        P = (1.0 - f0)[:, None, None] * T[b0] + f0[:, None, None] * T[b0 + 1]
    else:
        (b0, f0), (b1, f1) = pos
        P = (((1.0 - f0) * (1.0 - f1))[:, None, None] * T[b0, b1] +
             (f0 * (1.0 - f1))[:, None, None] * T[b0 + 1, b1] +
             ((1.0 - f0) * f1)[:, None, None] * T[b0, b1 + 1] +
             (f0 * f1)[:, None, None] * T[b0 + 1, b1 + 1])
    out = np.einsum("mrc,cm->rm", P, states)
    if conserve_sum is not None:
        out *= conserve_sum / out.sum(axis=0)
    return out


def two_state_derivative(rate_fn):
    """Derivative of the scheme ``~ C <-> O (alpha, beta)`` with state order (C, O)."""
    def deriv(v, C, O):
        a, b = rate_fn(v)
        return [-a * C + b * O, a * C - b * O]
    return deriv


def two_state_closed_form(alpha, beta, dt):
    """Closed-form expm of [[-a, b], [a, -b]] dt (state order C, O) -> 2x2 matrix."""
    s = alpha + beta
    e = math.exp(-s * dt)
    oinf = alpha / s
    cinf = beta / s
    return np.array([[cinf + oinf * e, cinf * (1 - e)],
                     [oinf * (1 - e), oinf + cinf * e]])


def hh_gate_rate(gate, celsius=6.3):
    return lambda v: tuple(float(x) for x in hh_alpha_beta(v, celsius)[gate])


# ---------------------------------------------------------------------------------------------
# Accumulation
# ---------------------------------------------------------------------------------------------
def accumulate_scatter(num_segments, segment_of_instance, g_instance, reversal):
    """Nonspecific conductance scatter: sum_g[seg] += g; d_num[seg] += g E.
    species.py:256-260 (numpy fancy-index += keeps only one write per duplicate index; the
    build uses true sums, as the numba pointer path in nmodl/__init__.py:396-406 does)."""
    sum_g = np.zeros(num_segments)
    d_num = np.zeros(num_segments)
    np.add.at(sum_g, segment_of_instance, g_instance)
    np.add.at(d_num, segment_of_instance, g_instance * reversal)
    return sum_g, d_num


def driving_voltage(sum_g, d_num):
    """driving_voltage /= sum_conductance; NaN -> 0.  rxd_model.py:105-109."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.nan_to_num(d_num / sum_g)


# ---------------------------------------------------------------------------------------------
# Stimulus: counter-based Poisson pulses (device-generated in the build)
# ---------------------------------------------------------------------------------------------
def stim_hash(seed, cell, tick):
    """32-bit hash of (seed, cell, tick); bit-exact restatement of the device generator."""
    M = 0xFFFFFFFF
    h = (np.uint64(seed) * np.uint64(0x9E3779B1)) & M
    h ^= (np.uint64(cell) * np.uint64(0x85EBCA77)) & M
    h ^= (np.uint64(tick) * np.uint64(0xC2B2AE3D)) & M
    h = np.uint64(h)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & M
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & M
    h ^= h >> np.uint64(16)
    return h


def poisson_threshold(rate_hz, half_dt_ms):
    return int(math.floor(rate_hz * half_dt_ms * 1e-3 * 4294967296.0))


def poisson_pulses_active(seed, cell, tick, threshold, pulse_ticks):
    """Number of pulses covering input tick ``tick``: onsets at ticks (tick-pulse_ticks, tick]."""
    count = 0
    for k in range(pulse_ticks):
        t = tick - k
        if t < 0:
            break
        if int(stim_hash(seed, cell, t)) < threshold:
            count += 1
    return count


def pulse_ticks(duration_ms, half_dt_ms):
    """Number of input ticks a constant_wave plays for.  time.py:271-273 with constant_wave 406-408."""
    return math.ceil(duration_ms / half_dt_ms) - math.floor(0.0 / half_dt_ms)


# ---------------------------------------------------------------------------------------------
# Staggered advance for cells carrying HH channels
# ---------------------------------------------------------------------------------------------
class HHSim:
    """Cells with leak + HH Na (m^3 h) + HH K (n^4) on every segment; the staggered step of
    rxd_model.py:63-123.

    gate_mode "exact"  : hh.mod DERIVATIVE block solved exactly (solver.py:48-55)
    gate_mode "table"  : each gate a two-state kinetic scheme advanced by the LTI table kernel
                         (codegen.py:39-106) with linear interpolation on a 0.1 mV grid.
    g_bar: dict of per-segment Siemens (gbar * magnitude), E: reversal potentials (mV).
    stim: callable(tick) -> per-segment current (A) for input tick ``tick`` (1-based, the value of
          input_hook.ticks after its increment, time.py:114-117)."""

    def __init__(self, P, C, g_bar, E=None, v0=-70.0, dt=0.1, celsius=6.3, gate_mode="exact",
                 stim=None, ap_segments=(), ap_threshold=20.0, tables=None):
        self.P, self.C = P, np.asarray(C, dtype=np.float64)
        self.n = len(self.C)
        self.g_bar = {k: np.asarray(v, dtype=np.float64) for k, v in g_bar.items()}
        self.E = E or {"na": 60.0, "k": -88.0, "l": -54.3}
        self.dt = float(dt)
        self.half = 0.5 * self.dt
        self.celsius = celsius
        self.mode = gate_mode
        self.v = np.full(self.n, float(v0))
        self.gates = {g: np.full(self.n, float(gate_inf(v0, g, celsius))) for g in "mhn"}
        self.sum_g = np.zeros(self.n)     # accumulators start at zero (data_components initial 0.0)
        self.d_num = np.zeros(self.n)
        self.stim = stim
        self.input_ticks = 0
        self.steps = 0
        self.ap_segments = list(ap_segments)
        self.ap_threshold = ap_threshold
        self.ap_state = np.zeros(len(self.ap_segments), dtype=bool)
        self.ap_detected = np.zeros(len(self.ap_segments), dtype=bool)
        self.tables = tables
        if gate_mode == "table" and tables is None:
            self.tables = {}
            for g in "mhn":
                deriv = two_state_derivative(hh_gate_rate(g, celsius))
                self.tables[g] = node_table_to_polynomial(lti_node_table(deriv, 2, self.dt))

    def _half(self):
        self.input_ticks += 1
        I = np.zeros(self.n)
        if self.stim is not None:
            I = I + self.stim(self.input_ticks)
        d = driving_voltage(self.sum_g, self.d_num)
        self.v, _ = half_step(self.v, self.sum_g, d, I, self.C, self.P, self.half)

    def _mechanisms(self):
        v = self.v
        for g in "mhn":
            x = self.gates[g]
            if self.mode == "exact":
                self.gates[g] = gate_exact(x, v, g, self.dt, self.celsius)
            else:
                st = np.stack([1.0 - x, x])
                self.gates[g] = lti_advance_vec(st, v, self.tables[g], conserve_sum=1.0)[1]
        m, h, n_ = self.gates["m"], self.gates["h"], self.gates["n"]
        gna = self.g_bar["na"] * m * m * m * h
        gk = self.g_bar["k"] * n_ * n_ * n_ * n_
        gl = self.g_bar["l"].copy()
        self.sum_g = gna + gk + gl
        self.d_num = gna * self.E["na"] + gk * self.E["k"] + gl * self.E["l"]

    def advance(self):
        self._half()
        self._mechanisms()
        self._half()
        if self.ap_segments:
            over = self.v[self.ap_segments] >= self.ap_threshold     # neuron.py:72-76
            self.ap_detected = over & ~self.ap_state
            self.ap_state = over
        self.steps += 1


class KineticSim:
    """Cells whose segments carry linear kinetic mechanisms, stepped like HHSim (the staggered
    step of rxd_model.py:63-123): every mechanism is advanced by the LTI table kernel
    (codegen.py:39-106) at v(t + dt/2) and contributes g = gmax * sum_s weights[s] state_s with
    its reversal potential (species.py:155-160, rxd_model.py:105-109).

    mechanisms: list of dicts with poly (polynomial table of node_table_to_polynomial),
    weights [S], gmax [n] (Siemens per unit of sum_s weights[s] state_s), E (mV), states [S, n],
    conserve (CONSERVE total or None).  A mechanism driven by other inputs instead holds table
    (node table of lti_advance_nodes), grids, and inputs: a function (sim) -> list of input
    arrays [n] in the table's input order (the voltage is sim.v).  stim as in HHSim."""

    def __init__(self, P, C, mechanisms, v0=-70.0, dt=0.1, stim=None):
        self.P, self.C = P, np.asarray(C, dtype=np.float64)
        self.n = len(self.C)
        self.mech = [dict(m, states=np.array(m["states"], dtype=np.float64)) for m in mechanisms]
        self.dt, self.half = float(dt), 0.5 * float(dt)
        self.v = np.full(self.n, float(v0))
        self.sum_g = np.zeros(self.n)     # the first half step sees no conductance
        self.d_num = np.zeros(self.n)
        self.stim = stim
        self.input_ticks = 0
        self.steps = 0

    def _half(self):
        self.input_ticks += 1
        I = np.zeros(self.n) if self.stim is None else np.asarray(self.stim(self.input_ticks))
        d = driving_voltage(self.sum_g, self.d_num)
        self.v, _ = half_step(self.v, self.sum_g, d, I, self.C, self.P, self.half)

    def _mechanisms(self):
        self.sum_g = np.zeros(self.n)
        self.d_num = np.zeros(self.n)
        for m in self.mech:
            if "table" in m:
                m["states"] = lti_advance_nodes(m["states"], m["inputs"](self), m["table"],
                                                m["grids"], m["conserve"])
            else:
                m["states"] = lti_advance_vec(m["states"], self.v, m["poly"],
                                              conserve_sum=m["conserve"])
            g = m["gmax"] * (np.asarray(m["weights"]) @ m["states"])
            self.sum_g += g
            self.d_num += g * m["E"]

    def advance(self):
        self._half()
        self._mechanisms()
        self._half()
        self.steps += 1


# ---------------------------------------------------------------------------------------------
# C restatement (oracle/neuwon_oracle_c.c) -- same step, FP64, OpenMP over cells
# ---------------------------------------------------------------------------------------------
def _c_lib():
    import ctypes
    import os
    # NEUWON_ORACLE_LIB: another build of the same source (the sanitizer build of the tests)
    path = os.environ.get("NEUWON_ORACLE_LIB") or \
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)

    class Params(ctypes.Structure):
        _fields_ = [("dt", ctypes.c_double), ("vmin", ctypes.c_double), ("vmax", ctypes.c_double),
                    ("num_buckets", ctypes.c_int32), ("e_na", ctypes.c_double),
                    ("e_k", ctypes.c_double), ("e_l", ctypes.c_double),
                    ("stim_mode", ctypes.c_int32), ("stim_amp", ctypes.c_double),
                    ("seed", ctypes.c_uint32), ("threshold", ctypes.c_uint32),
                    ("pulse_ticks", ctypes.c_int32), ("gate_mode", ctypes.c_int32),
                    ("celsius", ctypes.c_double), ("fp32_state", ctypes.c_int32)]
    lib.oracle_hh_run.restype = ctypes.c_int
    lib.oracle_lti_run.restype = ctypes.c_int
    lib.oracle_lti_run.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int32, ctypes.c_double, ctypes.c_int32]
    lib.oracle_stim_hash.restype = ctypes.c_uint32
    lib.oracle_stim_hash.argtypes = [ctypes.c_uint32] * 3
    lib.Params = Params
    return lib


def c_lti_run(states, v0, poly_table, wave=None, minimum=-100.0, maximum=100.0,
              conserve_sum=None, threads=0):
    """The C restatement of ``lti_advance_vec`` (lti_sim/codegen.py:39-106), OpenMP over instances:
    states [S][M] FP64 advanced in place by len(wave) steps (or one step with wave None), step k at
    the voltages v0 + wave[k]."""
    import ctypes
    lib = _c_lib()
    states_ = np.ascontiguousarray(states, dtype=np.float64)
    S, M = states_.shape
    nb, _, _, terms = poly_table.shape
    poly = np.ascontiguousarray(poly_table, dtype=np.float64)
    v0_ = np.ascontiguousarray(v0, dtype=np.float64)
    w = None if wave is None else np.ascontiguousarray(wave, dtype=np.float64)
    n = 1 if w is None else len(w)
    P = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.oracle_lti_run(S, nb, terms, minimum, maximum, P(poly), M, P(states_), P(v0_), P(w), n,
                            float(conserve_sum or 0.0), int(threads))
    assert rc == 0
    return states_


class CHHPopulation:
    """Cells with leak + HH Na/K advanced by the C restatement.

    cell_ptr: int64 [cells+1]; P: scipy CSR over all segments (block diagonal, global columns);
    C, g_bar (dict na/k/l): per segment; stim: None | ('constant', amp_A) |
    ('poisson', amp_A, seed, threshold, pulse_ticks), applied to segment 0 of every cell.

    fp32_storage: an emulation of FP32 storage, to measure the error floor it sets (it is not
    the reference's algorithm, which is FP64 throughout): the per-segment maximal conductances,
    1/C and the gate tables' node values are rounded to FP32 once, every gate after each update;
    round_v also rounds every voltage at the end of every step."""

    def __init__(self, cell_ptr, P, C, g_bar, v0=-70.0, dt=0.1, celsius=6.3, gate_mode="table",
                 stim=None, E=None, tables=None, gates0=None, fp32_storage=False, round_v=False):
        import ctypes
        self.lib = _c_lib()
        self.cell_ptr = np.ascontiguousarray(cell_ptr, dtype=np.int64)
        P = P.tocsr()
        self.rowptr = np.ascontiguousarray(P.indptr, dtype=np.int64)
        self.col = np.ascontiguousarray(P.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(P.data, dtype=np.float64)
        f32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if fp32_storage \
            else (lambda a: np.asarray(a, dtype=np.float64))
        C = np.asarray(C, dtype=np.float64)
        self.C = np.ascontiguousarray(1.0 / f32(1.0 / C) if fp32_storage else C)
        self.g = {k: np.ascontiguousarray(f32(v)) for k, v in g_bar.items()}
        N = len(self.C)
        self.v = np.full(N, float(v0))
        gates0 = gates0 or {g: float(gate_inf(v0, g, celsius)) for g in "mhn"}
        self.gates = {g: np.full(N, float(gates0[g])) for g in "mhn"}
        self.sum_g = np.zeros(N)
        self.d_num = np.zeros(N)
        self.steps = 0
        E = E or {"na": 60.0, "k": -88.0, "l": -54.3}
        p = self.lib.Params()
        p.dt, p.vmin, p.vmax, p.num_buckets = dt, -100.0, 100.0, 2000
        p.e_na, p.e_k, p.e_l = E["na"], E["k"], E["l"]
        p.gate_mode = {"table": 0, "exact": 1}[gate_mode]
        p.celsius = celsius
        if stim is None:
            p.stim_mode = 0
        elif stim[0] == "constant":
            p.stim_mode, p.stim_amp = 1, stim[1]
        else:
            p.stim_mode, p.stim_amp, p.seed, p.threshold, p.pulse_ticks = 2, stim[1], stim[2], stim[3], stim[4]
        p.fp32_state = (1 if fp32_storage else 0) | (2 if round_v else 0)
        self.p = p
        if tables is None:
            tables = {}
            for g in "mhn":
                deriv = two_state_derivative(hh_gate_rate(g, celsius))
                tables[g] = node_table_to_polynomial(f32(lti_node_table(deriv, 2, dt)))
        self.poly = np.ascontiguousarray(np.stack([tables[g] for g in "mhn"]), dtype=np.float64)

    def advance(self, n_steps=1, threads=0):
        import ctypes
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = self.lib.oracle_hh_run(
            ctypes.byref(self.p), ctypes.c_int64(len(self.cell_ptr) - 1), P(self.cell_ptr),
            P(self.rowptr), P(self.col), P(self.val), P(self.C), P(self.g["na"]), P(self.g["k"]),
            P(self.g["l"]), P(self.v), P(self.gates["m"]), P(self.gates["h"]), P(self.gates["n"]),
            P(self.sum_g), P(self.d_num), P(self.poly), ctypes.c_int64(self.steps),
            ctypes.c_int32(n_steps), ctypes.c_int32(threads))
        assert rc == 0
        self.steps += n_steps


# ---------------------------------------------------------------------------------------------
# Intracellular species: concentration update, diffusion, first-order buffer
# ---------------------------------------------------------------------------------------------
def diffusion_coupling(parent, length, xarea, D):
    """Per edge (child c, parent p) the diffusive conductance k = D xarea_c / L_c (um^3/ms): the flux
    through the child's cross-section over its length.  The reference's matrix is a stub
    (species.py:74-75, extracellular.py:52-53); the layout follows the electric matrix
    (electric.py:92-108) with 1/(r C) replaced by k / vol."""
    parent = np.asarray(parent)
    k = np.zeros(len(parent))
    child = parent != NULL
    k[child] = D * np.asarray(xarea)[child] / np.asarray(length)[child]
    return k


def diffusion_matrix(parent, k, vol, time_step_ms):
    """Sparse A_D dt with A[c,p] = k/vol_c, A[c,c] = -k/vol_c, A[p,c] = k/vol_p, A[p,p] -= k/vol_p."""
    src, dst, coef = [], [], []
    for c, p in enumerate(parent):
        if p == NULL:
            continue
        src += [c, c, p, p]
        dst += [p, c, c, p]
        coef += [time_step_ms * k[c] / vol[p], -time_step_ms * k[c] / vol[c],
                 time_step_ms * k[c] / vol[c], -time_step_ms * k[c] / vol[p]]
    n = len(parent)
    return scipy.sparse.csc_matrix((coef, (dst, src)), shape=(n, n), dtype=np.float64)


def diffusion_propagator(parent, k, vol, time_step_ms, prune=EPSILON):
    A = diffusion_matrix(parent, k, vol, time_step_ms)
    P = scipy.sparse.csr_matrix(scipy.linalg.expm(A.toarray()))
    P.data[np.abs(P.data) < prune] = 0.0
    P.eliminate_zeros()
    return P


def buffer_exact(cab, ca, dt, kon, koff, btot):
    """Two-state buffer B <-> CaB with forward rate kon*Ca (Ca held over the step) and backward rate
    koff, CONSERVE B + CaB = btot, solved exactly (solver.py:48-55 applied to
    dCaB/dt = kon Ca (btot - CaB) - koff CaB)."""
    a = kon * ca
    s = a + koff
    inf = btot * a / s
    return inf + (cab - inf) * np.exp(-s * dt)


class CaSim:
    """Ca2+ with a first-order buffer in every compartment and intracellular diffusion.

    The step follows the staggered order of rxd_model.py:63-123 with the species update of
    species.py:54-70 in each half step: c = max(0, c + dc/dt * dt/2); c = P_D c.  The buffer is a
    mechanism: it reads Ca at t + dt/2, advances CaB over dt and writes dc/dt = -(CaB' - CaB)/dt
    (the accumulator ``{ion}_derivative``, nmodl/__init__.py:204-207), zeroed before mechanisms run
    (species.py:49-52)."""

    def __init__(self, P_half, vol, ca0, cab0, dt=0.1, kon=0.1, koff=0.01, btot=100.0):
        self.P = P_half
        self.vol = np.asarray(vol, dtype=np.float64)
        self.ca = np.asarray(ca0, dtype=np.float64).copy()
        self.cab = np.asarray(cab0, dtype=np.float64).copy()
        self.dcdt = np.zeros_like(self.ca)
        self.dt, self.kon, self.koff, self.btot = dt, kon, koff, btot

    def _half(self):
        c = np.maximum(0.0, self.ca + self.dcdt * (0.5 * self.dt))
        self.ca = self.P.dot(c)

    def advance(self):
        self._half()
        new = buffer_exact(self.cab, self.ca, self.dt, self.kon, self.koff, self.btot)
        self.dcdt = -(new - self.cab) / self.dt
        self.cab = new
        self._half()

    def mass(self):
        """Conserved total: the last buffer change is paid back in Ca over two half steps, the
        second of them in the next step, hence the pending dcdt dt/2."""
        return float(np.sum(self.vol * (self.ca + self.cab + 0.5 * self.dt * self.dcdt)))


# ---------------------------------------------------------------------------------------------
# Adjoint of the staggered HH step (table gates): FP64 forward, reverse pass, finite differences
# ---------------------------------------------------------------------------------------------
def hh_theta_tables(dt, celsius=6.3, theta=None, vmin=-100.0, vmax=100.0, num_buckets=2000):
    """Node tables of every gate for rates scaled by theta[gate] = (t_alpha, t_beta):
    open' = a(v) + b(v) open with a = P_OC, b = P_OO - P_OC = exp(-(alpha+beta) dt), and the
    derivatives of a and b with respect to t_alpha and t_beta (closed form of the 2x2 expm,
    lti_model.py:51-65).  Returns dict gate -> array [node][6] = (a, b, da/dta, db/dta, da/dtb, db/dtb)."""
    theta = theta or {}
    v = vmin + np.arange(num_buckets + 1) * ((vmax - vmin) / num_buckets)
    rates = hh_alpha_beta(v, celsius)
    out = {}
    for g in "mhn":
        ta, tb = theta.get(g, (1.0, 1.0))
        a0, b0 = rates[g]
        al, be = ta * a0, tb * b0
        s = al + be
        e = np.exp(-s * dt)
        oinf = al / s
        a = oinf * (1.0 - e)
        da_dal = be / s ** 2 * (1.0 - e) + oinf * dt * e
        da_dbe = -al / s ** 2 * (1.0 - e) + oinf * dt * e
        db_dal = -dt * e
        db_dbe = -dt * e
        out[g] = np.stack([a, e, a0 * da_dal, a0 * db_dal, b0 * da_dbe, b0 * db_dbe], axis=1)
    return out


class HHAdjointProblem:
    """Cells with HH on every segment, constant current at segment 0, loss
    L = sum_steps (v_soma - target)^2 with v_soma the voltage of segment 0 after each step.

    Parameters: per-segment conductances g_na, g_k, g_l (S) and six rate multipliers
    theta[gate] = (t_alpha, t_beta).  Same step as HHSim in table mode, with the table
    interpolation written on node values (identical to the polynomial form)."""

    def __init__(self, P, C, stim_amp, dt=0.05, celsius=6.3, v0=-70.0, E=None):
        self.P = P.tocsr()
        self.PT = self.P.T.tocsr()
        self.C = np.asarray(C, dtype=np.float64)
        self.n = len(self.C)
        self.stim = np.zeros(self.n)
        self.stim[0] = stim_amp
        self.dt, self.celsius, self.v0 = dt, celsius, v0
        self.E = E or {"na": 60.0, "k": -88.0, "l": -54.3}
        self.nb = 2000
        self.frq = self.nb / 200.0

    def _lookup(self, tbl, v):
        u = (v + 100.0) * self.frq
        b = np.trunc(u).astype(np.int64)
        f = u - b
        hi, lo = b > self.nb - 1, b < 0
        inside = ~(hi | lo)
        b = np.where(hi, self.nb - 1, np.where(lo, 0, b))
        f = np.where(hi, 1.0, np.where(lo, 0.0, f))
        t0, t1 = tbl[b], tbl[b + 1]
        val = t0 + f[:, None] * (t1 - t0)
        slope = np.where(inside[:, None], (t1 - t0) * self.frq, 0.0)
        return val, slope

    def _g(self, gb, m, h, n):
        gna = gb["na"] * m ** 3 * h
        gk = gb["k"] * n ** 4
        gl = gb["l"]
        G = gna + gk + gl
        N = gna * self.E["na"] + gk * self.E["k"] + gl * self.E["l"]
        return gna, gk, G, N

    def forward(self, gb, theta, steps, target=None, record=False, fp32_storage=False):
        """fp32_storage: emulate FP32 storage (per-segment conductances, 1/C, the gate tables'
        nodes, every gate after its update, and the record's voltages rounded to FP32; the
        voltages themselves stay FP64 between steps, as on the matrix-free stores), to measure
        the floor that FP32 storage sets (not the reference's arithmetic)."""
        r32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if fp32_storage \
            else (lambda a: a)
        tables = hh_theta_tables(self.dt, self.celsius, theta)
        if fp32_storage:
            gb = {k: r32(gb[k]) for k in gb}
            for g in "mhn":
                tables[g] = tables[g].copy()
                tables[g][:, :2] = r32(tables[g][:, :2])
        self._C_run = 1.0 / r32(1.0 / self.C) if fp32_storage else self.C
        self._gb_run = gb
        hs = 0.5 * self.dt * 1e-3
        hdt = 0.5 * self.dt
        bI = hdt * self.stim / self._C_run
        v = np.full(self.n, self.v0)
        ss = {g: float(gate_inf(self.v0, g, self.celsius)) for g in "mhn"}
        # rest state of the scaled rates
        gates = {}
        for g in "mhn":
            a0, b0 = (float(x) for x in hh_alpha_beta(self.v0, self.celsius)[g])
            ta, tb = theta.get(g, (1.0, 1.0))
            gates[g] = np.full(self.n, ta * a0 / (ta * a0 + tb * b0))
        soma, rec = [], []
        for k in range(steps):
            m, h, n_ = gates["m"], gates["h"], gates["n"]
            if k == 0:
                x = v + bI                       # accumulators start at zero
            else:
                _, _, G, N = self._g(gb, m, h, n_)
                D = N / G
                al = np.exp(-hs * G / self._C_run)
                x = D * (1 - al) + v * al + bI
            v1 = self.P.dot(x)
            new = {}
            for g in "mhn":
                val, _ = self._lookup(tables[g], v1)
                new[g] = r32(val[:, 0] + val[:, 1] * gates[g])
            if record:
                rec.append((r32(v.copy()), dict(gates), r32(v1.copy())))
            _, _, G, N = self._g(gb, new["m"], new["h"], new["n"])
            D = N / G
            al = np.exp(-hs * G / self._C_run)
            v = self.P.dot(D * (1 - al) + v1 * al + bI)
            gates = new
            soma.append(v[0])
        soma = np.array(soma)
        loss = None if target is None else float(np.sum((soma - target) ** 2))
        if record:
            self._rec, self._final_gates, self._tables = rec, gates, tables
        return soma, loss

    def gradient(self, gb, theta, target, fp32_storage=False):
        """Reverse pass over the recorded forward; returns (loss, grads) with grads
        {'na','k','l': per segment, 'theta': {gate: (dL/dt_alpha, dL/dt_beta)}}.  fp32_storage:
        the reverse pass of the FP32-storage emulation of forward() (FP64 arithmetic)."""
        steps = len(target)
        soma, loss = self.forward(gb, theta, steps, target, record=True, fp32_storage=fp32_storage)
        gb, C = self._gb_run, self._C_run
        hs = 0.5 * self.dt * 1e-3
        E = self.E
        lv = np.zeros(self.n)
        lg = {g: np.zeros(self.n) for g in "mhn"}
        grad = {"na": np.zeros(self.n), "k": np.zeros(self.n), "l": np.zeros(self.n)}
        gth = {g: [0.0, 0.0] for g in "mhn"}
        tables = self._tables

        def g_adjoint(gates, lG, lN):
            """Adjoint through G, N of the given gates: gate adjoints and parameter grads."""
            m, h, n_ = gates["m"], gates["h"], gates["n"]
            l_gna = lG + E["na"] * lN
            l_gk = lG + E["k"] * lN
            l_gl = lG + E["l"] * lN
            grad["na"] += l_gna * m ** 3 * h
            grad["k"] += l_gk * n_ ** 4
            grad["l"] += l_gl
            return {"m": l_gna * gb["na"] * 3 * m ** 2 * h, "h": l_gna * gb["na"] * m ** 3,
                    "n": l_gk * gb["k"] * 4 * n_ ** 3}

        def half_adjoint(v_in, gates, lx, first):
            """x = D(1-al) + v al + b: returns lambda_v_in and gate adjoints of the half step."""
            if first:
                return lx, {g: np.zeros(self.n) for g in "mhn"}
            _, _, G, N = self._g(gb, gates["m"], gates["h"], gates["n"])
            D = N / G
            al = np.exp(-hs * G / C)
            lD = (1 - al) * lx
            lal = (v_in - D) * lx
            lG = lal * al * (-hs / C) - lD * D / G
            lN = lD / G
            return al * lx, g_adjoint(gates, lG, lN)

        next_gates = self._final_gates
        for k in range(steps - 1, -1, -1):
            v_k, gates_k, v1 = self._rec[k]
            lv[0] += 2.0 * (soma[k] - target[k])
            # half B: v2 = P x_B, gates after the mechanism = next_gates
            lxB = self.PT.dot(lv)
            lv1, lgB = half_adjoint(v1, next_gates, lxB, False)
            for g in "mhn":
                lg[g] = lg[g] + lgB[g]
            # mechanism: g' = a(v1) + b(v1) g
            for g in "mhn":
                val, slope = self._lookup(tables[g], v1)
                lv1 = lv1 + (slope[:, 0] + slope[:, 1] * gates_k[g]) * lg[g]
                gth[g][0] += float(np.sum((val[:, 2] + val[:, 3] * gates_k[g]) * lg[g]))
                gth[g][1] += float(np.sum((val[:, 4] + val[:, 5] * gates_k[g]) * lg[g]))
                lg[g] = val[:, 1] * lg[g]
            # half A: v1 = P x_A with the gates at the start of the step
            lxA = self.PT.dot(lv1)
            lv, lgA = half_adjoint(v_k, gates_k, lxA, k == 0)
            for g in "mhn":
                lg[g] = lg[g] + lgA[g]
            next_gates = gates_k
        # initial gates depend on theta through the rest state
        for g in "mhn":
            a0, b0 = (float(x) for x in hh_alpha_beta(self.v0, self.celsius)[g])
            ta, tb = theta.get(g, (1.0, 1.0))
            s = ta * a0 + tb * b0
            d_ta = a0 * (tb * b0) / s ** 2
            d_tb = -(ta * a0) * b0 / s ** 2
            gth[g][0] += float(np.sum(lg[g])) * d_ta
            gth[g][1] += float(np.sum(lg[g])) * d_tb
        grad["theta"] = {g: tuple(gth[g]) for g in "mhn"}
        return loss, grad


def theta_poly_tables(dt, celsius=6.3, theta=None):
    """Polynomial LTI tables (codegen layout) of the three gates for rates scaled by theta, and the
    rest state of the scaled rates at v0 = -70 mV."""
    t = hh_theta_tables(dt, celsius, theta)
    out = {}
    for g in "mhn":
        a, b = t[g][:, 0], t[g][:, 1]
        node = np.empty((len(a), 2, 2))
        node[:, 1, 0], node[:, 1, 1] = a, a + b
        node[:, 0, 0], node[:, 0, 1] = 1.0 - a, 1.0 - (a + b)
        out[g] = node_table_to_polynomial(node)
    return out


def rest_gates(v0, celsius=6.3, theta=None):
    theta = theta or {}
    out = {}
    for g in "mhn":
        a0, b0 = (float(x) for x in hh_alpha_beta(v0, celsius)[g])
        ta, tb = theta.get(g, (1.0, 1.0))
        out[g] = ta * a0 / (ta * a0 + tb * b0)
    return out
