"""The bucket-grouped LTI kernel (nw_lti_table_advance) against the per-lane gather form of the same
kernel (bit for bit) and against the FP64 oracle of the reference kernel (lti_sim/codegen.py:39-106)
fed the same FP32 table."""
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-lti-grouped] {msg}", flush=True)
    sys.stdout.flush()


def _scheme(states=13):
    """The 13-state AMPA13-graph scheme or a 6-state chain."""
    from neuwon_amd import channels
    if states == 13:
        sc = channels.random_ampa13_scheme(seed=2, time_step=0.1)
    else:
        rng = np.random.default_rng(9)
        edges = [(i, i + 1) for i in range(states - 1)] + [(i + 1, i) for i in range(states - 1)]
        sc = channels.voltage_markov_scheme(edges, np.exp(rng.uniform(-4.0, 0.0, len(edges))),
                                            rng.uniform(-0.05, 0.05, len(edges)), time_step=0.1,
                                            num_states=states,
                                            weights=np.eye(states)[-1])
    node32 = sc.node_table().astype(np.float32)
    return sc, node32


def _steady(sc, v):
    """Steady states at the grid node below each voltage (a valid distribution per instance)."""
    nodes = sc.grid.nodes()
    ss = np.stack([sc.initial_state(x) for x in nodes])
    b = np.clip(np.searchsorted(nodes, v) - 1, 0, len(nodes) - 1)
    st = ss[b]
    return (st / st.sum(axis=1, keepdims=True)).T


def _table(torch, node32):
    from neuwon_amd import kinetic
    return kinetic.LTITable(torch.as_tensor(node32, device="cuda"), -100.0, 100.0)


@pytest.mark.parametrize("order", ["sorted", "random"])
@pytest.mark.parametrize("states", [13, 16, 8, 6, 5, 2])
def test_grouped_equals_gather_form(torch_gpu, order, states):
    """Sorted voltages (one or two buckets per wavefront: the grouped path) and random ones (up to
    64 buckets per wavefront): group_max 0 (every wavefront gathers), the default and 64 (every
    wavefront grouped) give the same states and conductances, bit for bit."""
    torch = torch_gpu
    sc, node32 = _scheme(states)
    tab = _table(torch, node32)
    M = 200_003                     # not a multiple of the wavefront or the workgroup
    rng = np.random.default_rng(7)
    v = rng.uniform(-80.0, 40.0, M)
    if order == "sorted":
        v.sort()
    st0 = _steady(sc, v).astype(np.float32)
    vs = torch.as_tensor(v.astype(np.float32), device="cuda")
    w = torch.as_tensor(np.asarray(sc.weights, dtype=np.float32), device="cuda")
    gmax = torch.as_tensor(rng.uniform(0.5, 2.0, M).astype(np.float32), device="cuda")
    outs = {}
    for gm in (0, -1, 64):
        st = torch.as_tensor(st0.copy(), device="cuda")
        cond = torch.zeros(M, device="cuda")
        for _ in range(3):
            tab.advance(st, vs, conserve_sum=1.0, weights=w, gmax=gmax, cond_out=cond,
                        group_max=gm)
        torch.cuda.synchronize()
        outs[gm] = (st.cpu().numpy(), cond.cpu().numpy())
    for gm in (-1, 64):
        assert np.array_equal(outs[gm][0], outs[0][0])
        assert np.array_equal(outs[gm][1], outs[0][1])
    # against the FP64 oracle on the same FP32 table
    poly = O.node_table_to_polynomial(node32.astype(np.float64))
    ref = st0.astype(np.float64)
    for _ in range(3):
        ref = O.lti_advance_vec(ref, v.astype(np.float32).astype(np.float64), poly,
                                conserve_sum=1.0)
    got = outs[0][0].astype(np.float64)
    err = np.abs(got - ref).max()
    rel = (np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)).max()
    wref = gmax.cpu().numpy().astype(np.float64) * (np.asarray(sc.weights)[:, None] * ref).sum(0)
    cerr = np.abs(outs[0][1] - wref).max() / np.abs(wref).max()
    _say(f"{order}: 3 steps, max abs state error {err:.2e}, max rel {rel:.2e}, conductance "
         f"rel {cerr:.2e}")
    assert err < 1e-6 and rel < 1e-5 and cerr < 1e-6


def test_out_of_range_and_indexed_voltages(torch_gpu):
    """Voltages below and above the table (clamped to the end nodes as the reference's lookup
    does), exactly on nodes, read through an index, and an empty launch."""
    torch = torch_gpu
    sc, node32 = _scheme()
    tab = _table(torch, node32)
    M = 4099
    rng = np.random.default_rng(3)
    seg_v = np.concatenate([[-300.0, -100.0, -100.05, 100.0, 100.05, 250.0, 0.0, -50.0],
                            rng.uniform(-120.0, 120.0, 992)])
    index = rng.integers(0, len(seg_v), M).astype(np.int32)
    index[:len(seg_v)] = np.arange(len(seg_v))
    v = seg_v[index]
    st0 = _steady(sc, np.clip(v, -100, 100)).astype(np.float32)
    st = torch.as_tensor(st0.copy(), device="cuda")
    tab.advance(st, torch.as_tensor(seg_v.astype(np.float32), device="cuda"),
                seg_index=torch.as_tensor(index, device="cuda"), conserve_sum=1.0)
    tab.advance(st[:, :0].contiguous(), torch.zeros(1, device="cuda"), conserve_sum=1.0)
    torch.cuda.synchronize()
    poly = O.node_table_to_polynomial(node32.astype(np.float64))
    ref = O.lti_advance_vec(st0.astype(np.float64), v.astype(np.float32).astype(np.float64), poly,
                            conserve_sum=1.0)
    err = np.abs(st.cpu().numpy().astype(np.float64) - ref).max()
    _say(f"out-of-range and indexed voltages: max abs state error {err:.2e}")
    assert err < 1e-6


def test_markov13_1M_evolving_voltages(torch_gpu):
    """1M instances of the 13-state scheme, one step per launch for 1000 steps while every
    voltage follows a common 10 mV oscillation around its own level (instances in order of
    their levels): the path the model runs.  Per-state relative error against the FP64 oracle on
    a sample, fed the same FP32 table; the bar is three times the error of the oracle itself with
    its states rounded to FP32 after every step (the floor FP32 state storage sets)."""
    torch = torch_gpu
    sc, node32 = _scheme()
    tab = _table(torch, node32)
    M, steps = 1_000_000, 1000
    rng = np.random.default_rng(2)
    v0 = np.sort(rng.uniform(-80.0, 40.0, M))
    st0 = _steady(sc, v0).astype(np.float32)
    st = torch.as_tensor(st0.copy(), device="cuda")
    v0d = torch.as_tensor(v0.astype(np.float32), device="cuda")
    vs = torch.empty_like(v0d)
    wave = 10.0 * np.sin(2.0 * np.pi * np.arange(steps) * 0.1 / 20.0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = 0.0
    for k in range(steps):
        torch.add(v0d, float(wave[k]), out=vs)
        e0.record()
        tab.advance(st, vs, conserve_sum=1.0)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
    rate = M * steps / (ms * 1e-3)
    _say(f"1M x 1000 steps, one launch per step: {ms / steps * 1e3:.1f} us per launch, "
         f"{rate / 1e9:.1f} G instance-steps/s ({108 * rate / 1e12:.2f} TB/s of the 108 B per "
         f"instance-step)")
    got = st.cpu().numpy().astype(np.float64)
    assert np.abs(got.sum(axis=0) - 1.0).max() < 1e-5
    sample = np.sort(rng.choice(M, 2000, replace=False))
    poly = O.node_table_to_polynomial(node32.astype(np.float64))
    ref = st0[:, sample].astype(np.float64)
    emu = ref.copy()
    for k in range(steps):
        vk = (v0[sample].astype(np.float32) + np.float32(wave[k])).astype(np.float64)
        ref = O.lti_advance_vec(ref, vk, poly, conserve_sum=1.0)
        emu = O.lti_advance_vec(emu, vk, poly, conserve_sum=1.0).astype(np.float32).astype(np.float64)
    d = np.abs(got[:, sample] - ref)
    rel = d / np.maximum(ref, 1e-30)
    floor = np.abs(emu - ref) / np.maximum(ref, 1e-30)
    big = ref >= 1e-6    # occupancies below 1e-6 are rounded states of FP32 magnitude noise
    _say(f"after {steps} steps: max abs state error {d.max():.2e}; per-state relative error "
         f"{rel[big].max():.2e} where the occupancy >= 1e-6 (FP32-storage floor "
         f"{floor[big].max():.2e}); smallest occupancy in the sample {ref.min():.1e}")
    assert d.max() < max(1e-6, 3.0 * np.abs(emu - ref).max())
    assert rel[big].max() < max(1e-5, 3.0 * floor[big].max())
