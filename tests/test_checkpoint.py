"""The snapshot file format on the CPU: arrays and metadata round trip without pickling, and
files of another kind or format are refused."""
import json

import numpy as np
import pytest

from neuwon_amd import checkpoint


def test_round_trip(tmp_path):
    f = tmp_path / "s.npz"
    arrays = {"v": np.linspace(-70, 30, 11, dtype=np.float32),
              "state:hh.m": np.arange(6, dtype=np.float32).reshape(2, 3),
              "flags": np.array([0, 1, 1], dtype=np.uint8),
              "count": np.array([3, 0], dtype=np.int32)}
    checkpoint.save(f, arrays, {"kind": "HHPopulation", "steps": 7, "stim": {"mode": "constant",
                                                                              "amp_na": 0.5}})
    got, meta = checkpoint.load(f, "HHPopulation")
    assert set(got) == set(arrays)
    for k, a in arrays.items():
        assert got[k].dtype == a.dtype and np.array_equal(got[k], a)
    assert meta["steps"] == 7 and meta["stim"] == {"mode": "constant", "amp_na": 0.5}
    assert meta["format"] == checkpoint.FORMAT
    with np.load(f, allow_pickle=False) as z:      # plain arrays: readable without pickling
        assert json.loads(str(z["__meta__"]))["kind"] == "HHPopulation"


def test_refusals(tmp_path):
    f = tmp_path / "s.npz"
    with pytest.raises(TypeError):
        checkpoint.save(f, {"x": np.array([object()], dtype=object)}, {"kind": "A"})
    checkpoint.save(f, {"x": np.zeros(2)}, {"kind": "A"})
    with pytest.raises(ValueError, match="not B"):
        checkpoint.load(f, "B")
    g = tmp_path / "other.npz"
    np.savez(g, x=np.zeros(2), __meta__=np.array(json.dumps({"format": "something else"})))
    with pytest.raises(ValueError, match="not a snapshot"):
        checkpoint.load(g)


def test_copy_into_checks_shape():
    torch = pytest.importorskip("torch")
    t = torch.zeros(4)
    checkpoint.copy_into(t, np.arange(4, dtype=np.float64), "x")
    assert t.tolist() == [0.0, 1.0, 2.0, 3.0]
    with pytest.raises(ValueError, match="shape"):
        checkpoint.copy_into(t, np.zeros(5), "x")
