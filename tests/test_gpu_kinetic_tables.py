"""Kinetic tables built on the device (nw_lti_expm_batched, lti_sim/lti_model.py:51-65) against
scipy.linalg.expm, and the reference's own schemes (nmodl_library/kinetic/Nav11.mod,
neuwon/rxd/lti_sim/tests/ampa13.mod; rate matrices from tests/golden/kinetic_tables, written by
tests/golden/make_kinetic_tables.py) tabulated on the device and run on the kinetic kernels
against the FP64 oracle."""
import json
import os
import sys

import numpy as np
import pytest
import scipy.linalg

from oracle import neuwon_oracle as O

pytestmark = pytest.mark.gpu

TABLES = os.path.join(os.path.dirname(__file__), "golden", "kinetic_tables")


def _say(msg):
    print(f"[gpu-kinetic-tables] {msg}", flush=True)
    sys.stdout.flush()


def _fixture(name):
    z = np.load(os.path.join(TABLES, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    S = meta["num_states"]
    shape = tuple(meta["node_shape"])
    A = np.zeros((int(np.prod(shape)), S, S))
    p = z["pattern"]
    A[:, p[:, 0], p[:, 1]] = z["values"]
    return A.reshape(shape + (S, S)), meta


def _scipy_table(A, dt, normalise=True):
    P = scipy.linalg.expm(A * dt)
    if normalise:
        P /= P.sum(axis=-2, keepdims=True)
    return P


def _rel(P, ref):
    """Largest entry error relative to its column's largest entry."""
    scale = np.abs(ref).max(axis=-2, keepdims=True)
    return float((np.abs(P - ref) / np.maximum(scale, 1e-300)).max())


def test_expm_batched_random_rate_matrices(torch_gpu):
    """Rate matrices of 1 to 16 states with rates spanning 1e-3..1e3 /ms (up to ~12 squarings),
    a zero matrix and an unnormalised call, against scipy.linalg.expm."""
    from neuwon_amd import kinetic
    torch = torch_gpu
    rng = np.random.default_rng(4)
    for S in (1, 2, 3, 6, 10, 13, 16):
        n = 300
        A = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), (n, S, S)))
        A *= rng.random((n, S, S)) < 0.5                     # sparse schemes
        idx = np.arange(S)
        A[:, idx, idx] = 0.0
        A[:, idx, idx] = -A.sum(axis=1)                      # columns sum to zero
        A[0] = 0.0
        for normalise in (True, False):
            P = kinetic.expm_batched(torch.as_tensor(A, device="cuda"), 0.1, normalise).cpu().numpy()
            ref = _scipy_table(A, 0.1, normalise)
            err = _rel(P, ref)
            _say(f"S = {S:2d}, normalise {normalise}: max error relative to the column {err:.1e}")
            assert err < 1e-12
        assert np.abs(P[0] - np.eye(S)).max() == 0.0


def _exact_table(M, dt):
    """The column-normalised exponential to 50 digits (mpmath), rounded to FP64."""
    import mpmath
    mpmath.mp.dps = 50
    e = mpmath.expm(mpmath.matrix(M * dt))
    S = M.shape[0]
    E = np.array([[float(e[i, j]) for j in range(S)] for i in range(S)])
    return E / E.sum(axis=0, keepdims=True)


@pytest.mark.parametrize("name", ["Nav11", "ampa13"])
def test_reference_scheme_tables_on_the_device(torch_gpu, name):
    """Every node of the reference scheme's table: device expm (column-normalised) against
    scipy.linalg.expm, and on the stiffest nodes and a random sample against the exponential to
    50 digits (mpmath: at AMPA13's highest concentrations, |A dt| ~ 3e5, scipy itself is 1.5e-12
    from it); the nodes' columns sum to one."""
    from neuwon_amd import kinetic
    torch = torch_gpu
    A, meta = _fixture(name)
    dt = meta["time_step"]
    P = kinetic.expm_batched(torch.as_tensor(A, device="cuda"), dt).cpu().numpy()
    ref = _scipy_table(A, dt)
    err = _rel(P, ref)
    flat, pflat, sflat = A.reshape(-1, *A.shape[-2:]), P.reshape(-1, *P.shape[-2:]), \
        ref.reshape(-1, *ref.shape[-2:])
    stiff = np.argsort(-np.abs(np.diagonal(flat, axis1=1, axis2=2)).max(axis=1))[:6]
    pick = np.unique(np.r_[stiff, np.random.default_rng(1).choice(len(flat), 6, replace=False)])
    exact_dev = max(_rel(pflat[k], _exact_table(flat[k], dt)) for k in pick)
    exact_sp = max(_rel(sflat[k], _exact_table(flat[k], dt)) for k in pick)
    _say(f"{name} ({meta['source']}): {A.shape[:-2]} nodes of {meta['num_states']} states: "
         f"max error relative to the column against scipy {err:.1e}; on {len(pick)} nodes against "
         f"the 50-digit exponential: device {exact_dev:.1e}, scipy {exact_sp:.1e}")
    assert exact_dev < 1e-14
    assert err < max(1e-12, 2.0 * exact_sp)
    assert np.abs(P.sum(axis=-2) - 1.0).max() < 1e-13


def test_nav11_on_the_grouped_kernel(torch_gpu):
    """Nav1.1 (6 states, voltage input): 200k channels in order of their resting level with a
    common 15 mV oscillation, 300 steps, one per launch, on the device-built FP32 table, against
    the C FP64 oracle on the same table; bar tied to the oracle with FP32 states."""
    from neuwon_amd import kinetic
    torch = torch_gpu
    A, meta = _fixture("Nav11")
    g = meta["grids"][0]
    T64 = kinetic.expm_batched(torch.as_tensor(A, device="cuda"), meta["time_step"])
    T32 = T64.float().contiguous()
    tab = kinetic.LTITable(T32, g["minimum"], g["maximum"])
    S, M, steps = meta["num_states"], 200_000, 300
    rng = np.random.default_rng(6)
    v0 = np.sort(rng.uniform(-85.0, 30.0, M)).astype(np.float32)
    st0 = rng.dirichlet(np.ones(S), M).T.astype(np.float32)
    st = torch.as_tensor(st0.copy(), device="cuda")
    vd = torch.as_tensor(v0, device="cuda")
    vs = torch.empty_like(vd)
    wave = (15.0 * np.sin(2 * np.pi * np.arange(steps) * 0.1 / 10.0)).astype(np.float32)
    for k in range(steps):
        torch.add(vd, float(wave[k]), out=vs)
        tab.advance(st, vs, conserve_sum=meta["conserve_sum"])
    torch.cuda.synchronize()
    got = st.cpu().numpy().astype(np.float64)
    node32 = T32.cpu().numpy().astype(np.float64)
    poly = O.node_table_to_polynomial(node32)
    sample = np.sort(rng.choice(M, 3000, replace=False))
    ref = st0[:, sample].astype(np.float64)
    emu = ref.copy()
    for k in range(steps):
        vk = (v0[sample] + wave[k]).astype(np.float64)
        ref = O.c_lti_run(ref, vk, poly, minimum=g["minimum"], maximum=g["maximum"],
                          conserve_sum=meta["conserve_sum"])
        emu = O.c_lti_run(emu, vk, poly, minimum=g["minimum"], maximum=g["maximum"],
                          conserve_sum=meta["conserve_sum"]).astype(np.float32).astype(np.float64)
    d = np.abs(got[:, sample] - ref)
    floor = np.abs(emu - ref)
    big = ref >= 1e-6
    rel = (d / np.maximum(ref, 1e-30))[big].max()
    frel = (floor / np.maximum(ref, 1e-30))[big].max()
    _say(f"Nav11, {M} channels x {steps} steps: max abs state error {d.max():.2e} (FP32-state "
         f"floor {floor.max():.2e}), per-state relative {rel:.2e} (floor {frel:.2e})")
    assert d.max() < max(1e-6, 3.0 * floor.max())
    assert rel < max(1e-5, 3.0 * frel)


def test_ampa13_on_the_concentration_kernel(torch_gpu):
    """AMPA13 (13 states, glutamate input C on the logarithmic grid): 100k receptors at
    concentrations from 0 to 3 mM (and beyond the table), 50 steps held, on the device-built
    table through nw_lti_advance_nd, against the oracle's node-table kernel."""
    import ctypes
    from neuwon_amd import _lib, kinetic
    torch = torch_gpu
    A, meta = _fixture("ampa13")
    g = meta["grids"][0]
    T64 = kinetic.expm_batched(torch.as_tensor(A, device="cuda"), meta["time_step"])
    T32 = T64.float().contiguous()
    S, M, steps = meta["num_states"], 100_000, 50
    rng = np.random.default_rng(7)
    C = np.concatenate([[0.0, 1e-9, 1000.0, 2000.0], 10.0 ** rng.uniform(-7, np.log10(3.0), M - 4)])
    C = C.astype(np.float32)
    st0 = rng.dirichlet(np.ones(S), M).T.astype(np.float32)
    st = torch.as_tensor(st0.copy(), device="cuda")
    Cd = torch.as_tensor(C, device="cuda")
    inp = (_lib.NwLtiInput * 1)()
    inp[0].values = Cd.data_ptr()
    inp[0].index = None
    inp[0].mapping = _lib.NW_INPUT_LOG2
    inp[0].num_nodes = g["num_buckets"] + 1
    inp[0].minimum, inp[0].maximum, inp[0].scale = g["minimum"], g["maximum"], g["scale"]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().nw_lti_advance_nd(S, 1, ctypes.addressof(inp), _lib.ptr(T32), M,
                                             _lib.ptr(st), M, meta["conserve_sum"], steps, None,
                                             None, None, s))
    torch.cuda.synchronize()
    got = st.cpu().numpy().astype(np.float64)
    node32 = T32.cpu().numpy().astype(np.float64)
    ref = st0.astype(np.float64)
    for _ in range(steps):
        ref = O.lti_advance_nodes(ref, [C.astype(np.float64)], node32, [g],
                                  conserve_sum=meta["conserve_sum"])
    d = np.abs(got - ref)
    _say(f"AMPA13, {M} receptors x {steps} held steps: max abs state error {d.max():.2e}")
    assert d.max() < 2e-6
