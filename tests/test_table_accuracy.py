"""Table sizing for an accuracy target (channels.interpolation_error / choose_grid): the error
statistic against a closed-form two-state scheme swept densely, its second-order convergence,
the search's result against its own stopping rule, and the NMODL loader's ``accuracy=``."""
import os

import numpy as np
import pytest

from neuwon_amd import channels, nmodl

OWN = os.path.join(os.path.dirname(__file__), "golden", "kinetic")


def _two_state(a0=0.5, b0=0.04, a1=0.2, b1=-0.03, dt=0.1):
    return channels.voltage_markov_scheme([(0, 1), (1, 0)], [a0, a1], [b0, b1], time_step=dt,
                                          name="two", conserve_sum=1.0)


def test_error_statistic_against_closed_form_sweep():
    """The sampled max error lies within 10 % below the error of a dense sweep of the closed
    form (two_state_propagator at the nodes and at 400 points per bucket, no expm)."""
    a0, b0, a1, b1, dt = 0.5, 0.04, 0.2, -0.03, 0.1
    scheme = _two_state(a0, b0, a1, b1, dt)
    grid = channels.LinearInput(-100.0, 100.0, 40)
    err, per_bucket = channels.interpolation_error(scheme, [grid], samples_per_bucket=20, seed=3)

    def P(v):
        poc, poo = channels.two_state_propagator(a0 * np.exp(b0 * v), a1 * np.exp(b1 * v), dt)
        return np.stack([np.stack([1.0 - poc, 1.0 - poo], -1), np.stack([poc, poo], -1)], -2)

    nodes = P(grid.nodes())
    u = (np.arange(400) + 0.5) / 400
    dense = np.empty(grid.num_buckets)
    for b in range(grid.num_buckets):
        exact = P(grid.get_input_value(b + u))
        interp = nodes[b] + u[:, None, None] * (nodes[b + 1] - nodes[b])
        dense[b] = np.abs(interp - exact).max()
    assert per_bucket.shape == (40,)
    # (the sweep's 400 points miss a bucket's peak by at most 1/800 of the bucket: a few 1e-6)
    assert np.all(per_bucket <= dense * (1 + 1e-4) + 1e-15)
    assert err >= 0.9 * dense.max()
    assert np.argmax(per_bucket) == np.argmax(dense)


def test_error_is_second_order():
    """Linear interpolation of a smooth propagator: twice the buckets, a quarter of the error."""
    scheme = channels.random_ampa13_scheme(7)
    e1 = channels.interpolation_error(scheme, [channels.LinearInput(-100, 100, 100)], 8, 1)[0]
    e2 = channels.interpolation_error(scheme, [channels.LinearInput(-100, 100, 200)], 8, 1)[0]
    assert 3.5 < e1 / e2 < 4.5


def test_choose_grid_one_input():
    """The chosen count meets the target (also on other samples, within sampling slack) and 0.8
    of it does not (the search stopped where 0.9 of it failed)."""
    scheme = channels.random_ampa13_scheme(7)
    acc = 1e-4
    (grid,), err = channels.choose_grid(scheme, acc)
    assert err <= acc and 10 < grid.num_buckets < 2000
    assert channels.interpolation_error(scheme, [grid], 8, 11)[0] <= 1.2 * acc
    fewer = channels.resized(grid, 0.8 * grid.num_buckets)
    assert channels.interpolation_error(scheme, [fewer], 2, 0)[0] > acc


def test_choose_grid_two_inputs():
    """A two-state receptor driven by the voltage and a concentration (log2 buckets)."""
    def derivative(v, c, s0, s1):
        on = 0.5 * c / (c + 0.2) * np.exp(0.02 * v)
        off = 0.3 * np.exp(-0.01 * v)
        return [off * s1 - on * s0, on * s0 - off * s1]

    def rates(v, c):
        on = 0.5 * c / (c + 0.2) * np.exp(0.02 * v)
        off = 0.3 * np.exp(-0.01 * v)
        A = np.zeros(np.shape(v) + (2, 2))
        A[..., 0, 0], A[..., 1, 0] = -on, on
        A[..., 0, 1], A[..., 1, 1] = off, -off
        return A

    inputs = [channels.LinearInput(-100, 100, 20), channels.LogarithmicInput(0, 1000, 20, 1e-4)]
    scheme = channels.LTIScheme("rec2", ["C", "O"], derivative, 1.0, 0.1, inputs=inputs,
                                input_names=["v", "c"], rate_matrices=rates)
    acc = 1e-3
    grids, err = channels.choose_grid(scheme, acc)
    assert err <= acc
    assert channels.interpolation_error(scheme, grids, 4, 5)[0] <= 1.2 * acc
    n = [g.num_buckets for g in grids]
    for k in range(2):
        fewer = list(grids)
        fewer[k] = channels.resized(grids[k], 0.8 * n[k])
        # shrinking either input further fails, or that input is already at the search floor
        assert channels.interpolation_error(scheme, fewer, 1, 0)[0] > acc or n[k] <= 12


def test_choose_grid_refuses_unreachable_targets():
    scheme = _two_state()
    with pytest.raises(RuntimeError):
        channels.choose_grid(scheme, 1e-12, max_buckets=500)
    for bad in (0.0, 1.0, -1e-4):
        with pytest.raises(ValueError):
            channels.choose_grid(scheme, bad)
    with pytest.raises(ValueError):
        channels.interpolation_error(scheme, [scheme.grid, scheme.grid])


def test_nmodl_loader_accuracy():
    """load_kinetic(accuracy=1e-4) sizes the file's voltage grid for that error: fewer buckets
    than the fixed grid, the node table on the chosen nodes."""
    fixed = nmodl.load_kinetic(os.path.join(OWN, "k4.mod"), time_step=0.025, celsius=22.0)
    sized = nmodl.load_kinetic(os.path.join(OWN, "k4.mod"), time_step=0.025, celsius=22.0,
                               accuracy=1e-4)
    g = sized.inputs[0]
    assert g.num_buckets < fixed.inputs[0].num_buckets
    assert sized.table_error <= 1e-4
    assert channels.interpolation_error(sized, None, 2, 0)[0] == pytest.approx(sized.table_error)
    table = sized.node_table()
    assert table.shape == (g.num_buckets + 1, sized.num_states, sized.num_states)
    assert np.allclose(table.sum(axis=-2), 1.0, atol=1e-12)
