"""nw_hh_step on the matrix-free stores as one HIP graph launch per call (NW_LAUNCH_GRAPH): the
replays must equal the launches issued one by one, bit for bit, across calls (the Poisson
stimulus and the trace depend on the step index, which a replay reads from device memory)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-graph] {msg}", flush=True)
    sys.stdout.flush()


def _pop(cells, store):
    from neuwon_amd import population
    return population.HHPopulation(cells, store=store,
                                   stim={"mode": "poisson", "amp_na": 0.5, "rate_hz": 500.0,
                                         "duration_ms": 1.0, "seed": 5})


@pytest.mark.parametrize("store", ["chebyshev", "rational"])
def test_graph_replays_equal_direct_launches(torch_gpu, store):
    from neuwon_amd import synth
    torch = torch_gpu
    cells = synth.population(total=12000, seed=9, size_range=(200, 3000))
    runs = {}
    for graph in (False, True):
        pop = _pop(cells, store)
        traces = []
        # calls of different lengths, the same length again (a replay), a call of one step (direct)
        for n in (40, 40, 1, 60, 40):
            tr = torch.zeros((pop.num_cells, n), device="cuda")
            pop.advance(n, trace=tr, graph=graph)
            traces.append(tr.cpu().numpy())
        torch.cuda.synchronize()
        runs[graph] = (pop.v.cpu().numpy(), pop.m.cpu().numpy(), pop.ap_count.cpu().numpy(),
                       np.concatenate(traces, axis=1))
        _say(f"{store}, graph={graph}: steps {pop.steps}, "
             f"AP counts {int(runs[graph][2].sum())}")
    for a, b in zip(runs[False], runs[True]):
        assert np.array_equal(a, b)
    assert runs[True][2].sum() > 0   # the stimulus made spikes: the step index reached the kernels


def test_graph_record_prepare_and_eviction(torch_gpu):
    """prepare() captures without running; a record written by a replay equals the direct one;
    more argument sets than the propagator keeps (4) are captured again when they come back."""
    from neuwon_amd import synth
    torch = torch_gpu
    cells = synth.population(total=6000, seed=2, size_range=(300, 2000))
    out = {}
    for graph in (False, True):
        pop = _pop(cells, "chebyshev")
        n = 5
        rec = torch.zeros((5 * n + 4) * pop.num_segments, device="cuda")
        if graph:
            v0 = pop.v.clone()
            assert pop.prepare(n, record=rec) == 1
            torch.cuda.synchronize()
            assert pop.steps == 0 and torch.equal(pop.v, v0)
        pop.advance(n, record=rec, graph=graph)
        # six different argument sets (trace buffers of different widths), then the first again
        bufs = [torch.zeros((pop.num_cells, n + k), device="cuda") for k in range(6)]
        for tr in bufs + bufs[:1]:
            pop.advance(n, trace=tr, graph=graph)
        torch.cuda.synchronize()
        out[graph] = (rec.cpu().numpy(), pop.v.cpu().numpy(),
                      [b.cpu().numpy() for b in bufs])
    assert np.array_equal(out[False][0], out[True][0])
    assert np.array_equal(out[False][1], out[True][1])
    for a, b in zip(out[False][2], out[True][2]):
        assert np.array_equal(a, b)


def test_graph_inside_an_outer_capture(torch_gpu):
    """A call made while the caller captures its own stream launches directly (into the caller's
    graph) rather than replaying a graph of its own."""
    from neuwon_amd import synth
    torch = torch_gpu
    cells = synth.population(total=3000, seed=4, size_range=(300, 1500))
    ref = _pop(cells, "chebyshev")
    ref.advance(4, graph=False)
    pop = _pop(cells, "chebyshev")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        pop.advance(4, graph=True)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(pop.v, ref.v)
