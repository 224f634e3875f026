"""13-state kinetic workload: 1M instances of a voltage-dependent Markov scheme on the AMPA13
graph, 2001-node FP32 tables over -100..100 mV, voltages held constant, 1000 steps."""
import sys
import time

import numpy as np
import pytest

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-kinetic] {msg}", flush=True)
    sys.stdout.flush()


def _initial_states(scheme, node_ss, v):
    """Steady state at v, linearly interpolated between the steady states of the grid nodes."""
    u = (v + 100.0) * 10.0
    b = np.clip(np.trunc(u).astype(np.int64), 0, 1999)
    f = np.clip(u - b, 0.0, 1.0)
    st = (1 - f)[:, None] * node_ss[b] + f[:, None] * node_ss[b + 1]
    return (st / st.sum(axis=1, keepdims=True)).T


def test_markov13_1M(torch_gpu):
    from neuwon_amd import channels, _lib
    torch = torch_gpu
    t0 = time.time()
    scheme = channels.random_ampa13_scheme(seed=2, time_step=0.1)
    node = scheme.node_table()
    node_ss = np.stack([scheme.initial_state(v) for v in scheme.grid.nodes()])
    M, steps = 1_000_000, 1000
    rng = np.random.default_rng(2)
    v = rng.uniform(-80.0, 40.0, M).astype(np.float32)
    st0 = _initial_states(scheme, node_ss, v.astype(np.float64)).astype(np.float32)
    _say(f"setup {time.time() - t0:.1f}s")
    tbl = torch.as_tensor(node.astype(np.float32), device="cuda")
    vs = torch.as_tensor(v, device="cuda")
    lib = _lib.load()

    def run(states, n_steps, launches):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(launches):
            _lib.check(lib.nw_lti_advance(13, node.shape[0], -100.0, 100.0, _lib.ptr(tbl), M,
                                          _lib.ptr(vs), None, _lib.ptr(states), M, 1.0, n_steps,
                                          None, None, None, None))
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    held = torch.as_tensor(st0.copy(), device="cuda")
    run(held.clone(), 2, 1)                       # warm-up
    ms_held = run(held, steps, 1)
    stepwise = torch.as_tensor(st0.copy(), device="cuda")
    ms_step = run(stepwise, 1, 100)
    rate_held = M * steps / (ms_held * 1e-3)
    rate_step = M * 100 / (ms_step * 1e-3)
    _say(f"1000 steps in one launch: {ms_held:.2f} ms = {rate_held / 1e9:.1f} G instance-steps/s; "
         f"one step per launch: {ms_step / 100:.3f} ms/step = {rate_step / 1e9:.2f} G instance-steps/s "
         f"({108 * rate_step / 1e12:.2f} TB/s of state traffic)")
    got = held.cpu().numpy().astype(np.float64)
    assert np.abs(got.sum(axis=0) - 1.0).max() < 1e-5
    # parity on a sample against the reference kernel's FP64 arithmetic (polynomial table form)
    # on the same FP32 table the GPU reads
    node32 = node.astype(np.float32).astype(np.float64)
    poly = O.node_table_to_polynomial(node32)
    sample = rng.choice(M, 2000, replace=False)
    ref = st0[:, sample].astype(np.float64)
    vv = v[sample].astype(np.float64)
    for _ in range(steps):
        ref = O.lti_advance_vec(ref, vv, poly, conserve_sum=1.0)
    # what the held kernel rounds: each instance's interpolated propagator is formed once in FP32
    # (registers), the states stay FP64 over the launch; its distance from the FP64 oracle is the
    # floor of this form
    u = (vv + 100.0) * 10.0
    b = np.clip(np.trunc(u).astype(np.int64), 0, 1999)
    f = np.where(np.trunc(u) > 1999, 1.0, u - b)
    Pm = (node32[b] + f[:, None, None] * (node32[b + 1] - node32[b])).astype(np.float32)
    emu = st0[:, sample].astype(np.float64).T[:, :, None]
    Pm = Pm.astype(np.float64)
    for _ in range(steps):
        emu = Pm @ emu
        emu /= emu.sum(axis=1, keepdims=True)
    emu = emu[:, :, 0].T
    big = ref >= 1e-6
    d = np.abs(got[:, sample] - ref)
    rel = (d / np.maximum(ref, 1e-30))[big].max()
    floor = (np.abs(emu - ref) / np.maximum(ref, 1e-30))[big].max()
    vs_emu = (np.abs(got[:, sample] - emu) / np.maximum(emu, 1e-30))[big].max()
    _say(f"after {steps} held steps: max abs state error vs FP64 oracle {d.max():.2e}; per-state "
         f"relative {rel:.2e} where the occupancy >= 1e-6 (the held form's FP32-matrix floor "
         f"{floor:.2e}); against an emulation of the held form {vs_emu:.2e}; smallest occupancy "
         f"{ref.min():.1e}")
    assert rel < max(1e-5, 3.0 * floor)
    assert d.max() < max(1e-6, 3.0 * np.abs(emu - ref).max())
    assert vs_emu < 1e-6
    # the one-step-per-launch path: FP64 interpolation each step, states rounded to FP32 between
    # launches (its floor: the oracle with the same rounding)
    got1 = stepwise.cpu().numpy().astype(np.float64)
    ref1 = st0[:, sample].astype(np.float64)
    emu1 = ref1.copy()
    for _ in range(100):
        ref1 = O.lti_advance_vec(ref1, vv, poly, conserve_sum=1.0)
        emu1 = O.lti_advance_vec(emu1, vv, poly, conserve_sum=1.0).astype(np.float32).astype(np.float64)
    big1 = ref1 >= 1e-6
    rel1 = (np.abs(got1[:, sample] - ref1) / np.maximum(ref1, 1e-30))[big1].max()
    floor1 = (np.abs(emu1 - ref1) / np.maximum(ref1, 1e-30))[big1].max()
    _say(f"100 steps one per launch: per-state relative {rel1:.2e} (FP32-state floor {floor1:.2e})")
    assert rel1 < max(1e-5, 3.0 * floor1)
    assert np.abs(got1[:, sample] - ref1).max() < max(1e-6, 3.0 * np.abs(emu1 - ref1).max())
