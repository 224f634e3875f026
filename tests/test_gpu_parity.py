"""GPU parity: the HIP path against the FP64 oracle on the same seeded inputs."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_cell(cell, dt=0.1, prune=O.EPSILON):
    """Oracle geometry + propagator of one synth.CellSpec (independent restatement)."""
    n = len(cell)
    parent = np.asarray(cell.parent, dtype=np.int64)
    # coordinates that reproduce the drawn lengths exactly enough for the oracle geometry
    coords = np.zeros((n, 3))
    for i in range(1, n):
        coords[i] = coords[parent[i]] + np.array([cell.length[i], 0.0, 0.0])
    g = O.geometry(parent, coords, cell.diam)
    R, C = O.passive(g["length"], g["cross_sectional_area"], g["area_at_creation"])
    P = O.propagator(parent, R, C, 0.5 * dt, prune=prune)
    return g, R, C, P


def test_propagator_cell_build_matches_oracle(torch_gpu):
    from neuwon_amd import synth, population
    for n, prune in [(300, O.EPSILON), (1000, 1e-9)]:
        _say(f"propagator build n={n} prune={prune}")
        cell = synth.single_cell(n, seed=7)[0]
        g, R, C, P_ref = _oracle_cell(cell, prune=prune)
        P_gpu = population.cell_propagator_f64(cell.parent, R, C, 0.05, prune).toarray()
        P_ref = P_ref.toarray()
        diff = np.abs(P_gpu - P_ref)
        only_one = (P_gpu != 0) != (P_ref != 0)
        both = (P_gpu != 0) & (P_ref != 0)
        _say(f"  nnz gpu {np.count_nonzero(P_gpu)} oracle {np.count_nonzero(P_ref)}, "
             f"pattern differs at {only_one.sum()}, max abs diff {diff.max():.3e}")
        # entries kept by only one side sit at the threshold
        assert np.all(diff[only_one] < 2 * prune + 1e-15)
        rel = diff[both] / np.abs(P_ref[both])
        assert rel.max() < 1e-9, rel.max()
        assert only_one.sum() <= 1e-4 * both.sum()
        # rows sum to one, charge conserved (mass-weighted column sums)
        full = O.propagator(np.asarray(cell.parent), R, C, 0.05, prune=0.0)
        assert np.abs(np.asarray(full.sum(axis=1)).ravel() - 1).max() < 1e-12


def test_stored_fp32_propagator(torch_gpu):
    from neuwon_amd import synth, population
    cells = synth.population(total=3000, seed=3, size_range=(200, 900))
    pop = population.HHPopulation(cells, prune=1e-9)
    prop = pop.propagator
    _say(f"stored propagator cells={pop.num_cells} nnz/row={prop.mean_nnz_per_row:.1f}")
    for c in range(pop.num_cells):
        r0, r1 = pop.cell_ptr[c], pop.cell_ptr[c + 1]
        P32 = prop.export_cell(c)
        P64 = population.cell_propagator_f64(pop.parent_local[r0:r1], pop.R[r0:r1], pop.C[r0:r1],
                                             0.05, 1e-9)
        assert (P32 != 0).sum() == P64.nnz or np.array_equal(P32.indices, P64.indices)
        off = P32 - scipy.sparse.diags(P32.diagonal())
        off64 = P64 - scipy.sparse.diags(P64.diagonal())
        d = np.abs((off - off64).data)
        assert d.size == 0 or d.max() <= 6e-8 * abs(off64).max()
        rs32 = np.asarray(P32.sum(axis=1)).ravel()
        rs64 = np.asarray(P64.sum(axis=1)).ravel()
        assert np.abs(rs32 - rs64).max() < 1e-7


def voltage_error(v, ref, dt, window_ms=2.0, threshold=20.0, excursion=-20.0):
    """Errors of a voltage trace against the FP64 reference, split into subthreshold samples
    (relative error) and samples near an excursion (absolute error, mV): within window_ms of any
    sample where the reference is at or above `excursion` mV -- every action potential and every
    depolarisation that approaches one, where the voltage crosses 0 mV and a relative error is
    not defined.  Spikes: upward crossings of `threshold` (the AP detector's 20 mV)."""
    v, ref = np.asarray(v, np.float64), np.asarray(ref, np.float64)
    up = np.nonzero((ref[1:] >= threshold) & (ref[:-1] < threshold))[0] + 1
    t = np.arange(len(ref)) * dt
    near = np.zeros(len(ref), dtype=bool)
    for k in np.nonzero(ref >= excursion)[0]:
        near |= np.abs(t - t[k]) <= window_ms
    d = np.abs(v - ref)
    rel = d / np.abs(ref)
    return {"spikes": int(len(up)),
            "rel_subthreshold": float(rel[~near].max()) if (~near).any() else 0.0,
            "abs_spiking_mV": float(d[near].max()) if near.any() else 0.0,
            "rel_all": float(rel.max()), "abs_all_mV": float(d.max())}


def _c0_cell(n=1000, seed=0):
    from neuwon_amd import synth
    return synth.single_cell(n, seed=seed)


@pytest.mark.parametrize("gate_mode", ["table", "exact"])
def test_single_neuron_trajectory(torch_gpu, gate_mode):
    """1000-compartment tree, HH, 0.5 nA step at the root, dt 0.1 ms x 1000 steps."""
    from neuwon_amd import population, channels
    torch = torch_gpu
    steps = 1000
    _say(f"single neuron {gate_mode}")
    cells = _c0_cell()
    pop = population.HHPopulation(cells, prune=O.EPSILON, gate_mode=gate_mode,
                                  stim={"mode": "constant", "amp_na": 0.5})
    trace = torch.zeros((1, steps), dtype=torch.float32, device=pop.dev)
    pop.advance(steps, trace=trace)
    torch.cuda.synchronize()
    v_gpu = pop.v.cpu().numpy().astype(np.float64)
    tr_gpu = trace.cpu().numpy()[0].astype(np.float64)

    g, R, C, P = _oracle_cell(cells[0])
    area = g["surface_area"]
    gbar = {k: channels.HH_GBAR[k] * 1e-8 * area for k in ("na", "k", "l")}
    stim = np.zeros(len(cells[0]))
    stim[0] = 0.5e-9
    sim = O.HHSim(P, C, gbar, v0=-70.0, dt=0.1, celsius=6.3, gate_mode=gate_mode,
                  stim=lambda t: stim)
    tr = np.empty(steps)
    for k in range(steps):
        sim.advance()
        tr[k] = sim.v[0]
    err = voltage_error(tr_gpu, tr, dt=0.1)
    rel_v = np.abs(v_gpu - sim.v) / np.abs(sim.v)
    _say(f"{gate_mode}: root trace {err}, all segments at the end: max rel {rel_v.max():.3e}, "
         f"v_root range [{tr.min():.2f}, {tr.max():.2f}]")
    assert err["spikes"] >= 1
    assert err["rel_subthreshold"] < 1e-5
    assert err["abs_spiking_mV"] < 2e-3
    assert rel_v.max() < 1e-5
    for gname, arr in (("m", pop.m), ("h", pop.h), ("n", pop.n)):
        ga = arr.cpu().numpy().astype(np.float64)
        assert np.abs(ga - sim.gates[gname]).max() < 1e-5


def test_hh_golden_trace(torch_gpu):
    """Reference HH integration test (neuwon/rxd/tests/test_model.py:35-128): soma 5.7 um +
    ten 1 um segments, hh everywhere, 0.1 nA x 1 ms at t = 5 and 15 ms, dt 0.1, 6.3 degC."""
    from neuwon_amd import population, synth
    torch = torch_gpu
    golden = np.loadtxt(os.path.join(GOLDEN, "hh_trace_test_model.txt"))
    parent = np.array([-1] + list(range(10)), dtype=np.int32)
    length = np.array([0.0] + [1.0] * 10)
    diam = np.array([5.7] + [1.0] * 10)
    cell = synth.CellSpec(parent, length, diam)
    for mode in ("exact", "table"):
        _say(f"golden trace {mode}")
        pop = population.HHPopulation([cell], prune=O.EPSILON, gate_mode=mode,
                                      insert_area="creation")
        pop.ap_local[0] = 10
        trace = torch.zeros((1, 200), dtype=torch.float32, device=pop.dev)
        out = [-70.0]
        schedule = [(50, None), (10, 0.1), (90, None), (10, 0.1), (40, None)]
        for nsteps, amp in schedule:
            pop.set_stim({"mode": "constant", "amp_na": amp} if amp else None)
            pop.advance(nsteps, trace=trace)
            torch.cuda.synchronize()
            out += list(trace.cpu().numpy()[0, :nsteps])
        out = np.array(out)
        err = np.abs(out - golden)
        _say(f"golden {mode}: max abs diff {err.max():.3f} mV")
        assert err.max() < 30.0           # the reference test's own bar (test_model.py:128)
        assert err[:51].max() < 0.2       # pre-stimulus relaxation (first half-step offset only)
        assert int(pop.ap_count[0]) == 2  # two action potentials


def test_lti_13_state(torch_gpu):
    from neuwon_amd import channels, _lib
    torch = torch_gpu
    _say("lti 13-state")
    scheme = channels.random_ampa13_scheme(seed=2)
    node = scheme.node_table()
    poly = O.node_table_to_polynomial(node)
    M = 4096
    rng = np.random.default_rng(5)
    v = rng.uniform(-80, 40, M)
    st0 = np.stack([scheme.initial_state(x) for x in v[:64]] * (M // 64), axis=1)
    tbl = torch.as_tensor(node.astype(np.float32), device="cuda")
    vs = torch.as_tensor(v.astype(np.float32), device="cuda")
    st = torch.as_tensor(st0.astype(np.float32).copy(), device="cuda")
    steps = 20
    lib = _lib.load()
    _lib.check(lib.nw_lti_advance(13, node.shape[0], -100.0, 100.0, _lib.ptr(tbl), M, _lib.ptr(vs),
                                  None, _lib.ptr(st), M, 1.0, steps, None, None, None, None))
    torch.cuda.synchronize()
    ref = st0.copy()
    for _ in range(steps):
        ref = O.lti_advance_vec(ref, v.astype(np.float32).astype(np.float64), poly, conserve_sum=1.0)
    got = st.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max()
    _say(f"lti 13: max abs err {err:.3e}")
    assert err < 2e-6
    assert np.abs(got.sum(axis=0) - 1).max() < 1e-6


@pytest.mark.parametrize("layout", ["few", "crowded", "indexed"])
def test_accumulate(torch_gpu, layout):
    """0-3 sources per segment, one segment with 3,000 sources (many rounds of four slots), and
    the conductances gathered through a permuted index (no run of consecutive indices: one
    gather per slot)."""
    from neuwon_amd import _lib
    torch = torch_gpu
    rng = np.random.default_rng(9)
    nseg = 5000
    per = rng.integers(0, 4, nseg)
    if layout == "crowded":
        per[300] = 3000
    ptr = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(per, out=ptr[1:])
    M = int(ptr[-1])
    seg_of = np.repeat(np.arange(nseg), per)
    g = rng.uniform(0, 1e-9, M)
    E = rng.choice([60.0, -88.0, -54.3], M)
    Inon = rng.uniform(-1e-10, 1e-10, nseg)
    t = lambda a, dt=np.float32: torch.as_tensor(np.asarray(a, dtype=dt), device="cuda")
    d_ptr, d_e, d_i = t(ptr, np.int64), t(E), t(Inon)
    if layout == "indexed":      # source k's conductance sits at g_buf[perm[k]]
        perm = rng.permutation(M)
        g_buf = np.empty(M)
        g_buf[perm] = g
        d_g, d_idx = t(g_buf), t(perm, np.int64)
    else:
        d_g, d_idx = t(g), None
    sg, dv, si = (torch.zeros(nseg, device="cuda") for _ in range(3))
    _lib.check(_lib.load().nw_accumulate(nseg, _lib.ptr(d_ptr), _lib.ptr(d_g), _lib.ptr(d_idx),
                                         _lib.ptr(d_e), _lib.ptr(d_i), _lib.ptr(sg), _lib.ptr(dv),
                                         _lib.ptr(si), None))
    torch.cuda.synchronize()
    g = np.asarray(g, np.float32).astype(np.float64)
    E = np.asarray(E, np.float32).astype(np.float64)
    ref_g, ref_num = O.accumulate_scatter(nseg, seg_of, g, E)
    ref_d = O.driving_voltage(ref_g, ref_num)
    assert np.allclose(sg.cpu().numpy(), ref_g, rtol=1e-6, atol=1e-16)
    assert np.allclose(dv.cpu().numpy(), ref_d, rtol=1e-5, atol=1e-5)
    assert np.allclose(si.cpu().numpy(), Inon, rtol=1e-7, atol=0)
    assert np.all(dv.cpu().numpy()[per == 0] == 0)
