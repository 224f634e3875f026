"""Host-side checks that need no GPU: the C-ABI library, product host logic against the oracle."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import neuwon_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "neuwon_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nw_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from neuwon_amd import _lib
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_argument_errors_without_gpu():
    """Argument checks run before any device call and report through nw_last_error."""
    from neuwon_amd import _lib
    lib = _lib.load()
    out = ctypes.c_void_p()
    cp = np.zeros(1, dtype=np.int64)
    rc = lib.nw_propagator_build(0, 0, _lib.ptr(cp), _lib.ptr(cp), _lib.ptr(cp), _lib.ptr(cp),
                                 0.05, 1e-9, ctypes.byref(out))
    assert rc == -1 and b"num_cells" in lib.nw_last_error()
    dummy = np.zeros(4, dtype=np.float32)
    rc = lib.nw_lti_advance(20, 11, -100.0, 100.0, _lib.ptr(dummy), 1, _lib.ptr(dummy), None,
                            _lib.ptr(dummy), 1, 1.0, 1, None, None, None, None)
    assert rc == -1 and b"state count" in lib.nw_last_error()
    rc = lib.nw_lti_advance(2, 1, -100.0, 100.0, _lib.ptr(dummy), 1, _lib.ptr(dummy), None,
                            _lib.ptr(dummy), 1, 1.0, 1, None, None, None, None)
    assert rc == -1 and b"grid" in lib.nw_last_error()
    with pytest.raises(RuntimeError, match="neuwon_amd error -1"):
        _lib.check(-1)


def test_stim_hash_library_matches_oracle():
    from neuwon_amd import _lib
    lib = _lib.load()
    for seed, cell, tick in [(0, 0, 0), (1, 2, 3), (123456, 9999, 4000000), (2**32 - 1, 7, 11)]:
        assert lib.nw_stim_hash(seed, cell, tick) == int(O.stim_hash(seed, cell, tick))


def test_geometry_matches_oracle():
    from neuwon_amd import geometry, synth
    for seed in range(3):
        cell = synth.random_cell(300, np.random.default_rng(seed))
        parent = cell.parent.astype(np.int64)
        coords = np.zeros((300, 3))
        rng = np.random.default_rng(99)
        for i in range(1, 300):
            u = rng.normal(size=3)
            coords[i] = coords[parent[i]] + cell.length[i] * u / np.linalg.norm(u)
        ref = O.geometry(parent, coords, cell.diam)
        for src in (coords, cell.length):
            got = geometry.geometry(parent, src, cell.diam)
            for k in ("length", "surface_area", "area_at_creation", "cross_sectional_area", "volume"):
                np.testing.assert_allclose(got[k], ref[k], rtol=1e-12)
            assert np.array_equal(got["primary"], ref["primary"])
        R, C = geometry.passive(got["length"], got["cross_sectional_area"], got["area_at_creation"])
        R0, C0 = O.passive(ref["length"], ref["cross_sectional_area"], ref["area_at_creation"])
        np.testing.assert_allclose(R, R0, rtol=1e-12)
        np.testing.assert_allclose(C, C0, rtol=1e-12)


def test_population_geometry_matches_per_cell():
    from neuwon_amd import geometry, synth
    cells = synth.population(total=2000, seed=5, size_range=(50, 400))
    g = synth.concat(cells)
    pop = geometry.geometry(g["parent"], g["length"], g["diam"])
    for c, cell in enumerate(cells):
        r0, r1 = g["cell_ptr"][c], g["cell_ptr"][c + 1]
        one = geometry.geometry(cell.parent.astype(np.int64), cell.length, cell.diam)
        np.testing.assert_allclose(pop["surface_area"][r0:r1], one["surface_area"], rtol=1e-15)


def test_hh_gate_table_matches_oracle_expm():
    from neuwon_amd import channels
    tbl = channels.hh_gate_table(0.1, 6.3)
    for gi, g in enumerate("mhn"):
        deriv = O.two_state_derivative(O.hh_gate_rate(g))
        node = O.lti_node_table(deriv, 2, 0.1)
        poc, poo = node[:, 1, 0], node[:, 1, 1]
        np.testing.assert_allclose(tbl[gi, :, 0], poc, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(tbl[gi, :, 1], poo - poc, rtol=1e-12, atol=1e-15)


def test_lti_scheme_matches_oracle():
    from neuwon_amd import channels
    scheme = channels.random_ampa13_scheme(seed=2)
    node = scheme.node_table()
    ref = O.lti_node_table(scheme.derivative, 13, 0.1)
    assert np.abs(node - ref).max() < 1e-13
    assert np.abs(node.sum(axis=1) - 1).max() < 1e-14     # column sums (lti_model.py:63-64)
    ss = scheme.initial_state(-40.0)
    P = scheme.make_matrix(-40.0)
    assert np.abs(P @ ss - ss).max() < 1e-10 and ss.sum() == pytest.approx(1.0)


def test_linear_input_known_answers():
    """lti_sim/tests/test_inputs.py:5-19 and 37-50."""
    from neuwon_amd import channels
    a = channels.LinearInput(4, 7, 17)
    assert a.get_bucket_location(4) == (0, 0.0)
    b, f = a.get_bucket_location(7 - 1e-12)
    assert b == 16 and f == pytest.approx(1.0)
    a = channels.LinearInput(11, 57, 107)
    rng = np.random.default_rng(0)
    for v in list(rng.uniform(11, 57, 100)) + [11.0, 57.0]:
        assert a.get_input_value(a.get_bucket_value(v)) == pytest.approx(v, abs=1e-12)


def test_synth_is_deterministic():
    from neuwon_amd import synth
    a = synth.population(total=5000, seed=7)
    b = synth.population(total=5000, seed=7)
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(x.parent, y.parent) and np.array_equal(x.diam, y.diam)
    assert sum(len(c) for c in a) >= 5000
    for c in a:
        assert 200 <= len(c) <= 5000
        counts = np.bincount(c.parent[1:], minlength=len(c))
        assert counts.max() <= 2                                 # binary tree
        assert np.all(c.parent[1:] < np.arange(1, len(c)))       # creation order


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
