"""Device-side clock callbacks (nw_clock_tick, nw_trace_all) and the captured step graph, against
the clock-callback oracle (oracle/time_callbacks.py) driven with the same values, and against the
eager step."""
import sys

import numpy as np
import pytest

from oracle import time_callbacks as T

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-events] {msg}", flush=True)
    sys.stdout.flush()


def _f32(x):
    return float(np.float32(x))


def test_event_table_through_the_abi(torch_gpu):
    """A table written by hand: '=' and '+=' players on one element (in order), a looping player,
    a player that starts later, recorders of an FP32 and a uint8 column, an FP64 trace; a ring of
    8 ticks drained every 8 launches.  Every value bitwise against the oracle's callbacks."""
    import ctypes
    from neuwon_amd import _lib, events
    torch = torch_gpu
    dev = "cuda"
    col = torch.zeros(64, dtype=torch.float32, device=dev)
    col[10] = 5.0
    u8 = torch.zeros(8, dtype=torch.uint8, device=dev)
    state = torch.zeros(3, dtype=torch.float64, device=dev)
    clock = T.Clock(0.1)
    c3, c10, u5 = T.Cell(0.0, _f32), T.Cell(5.0, _f32), T.Cell(0, int)
    a = T.TimeSeries().set_data([0.0, 1.0, 3.0, 2.0], [0.0, 0.2, 0.4, 0.6])
    a.play(c3, clock, mode="=", immediate=False)
    b = T.TimeSeries().sine_wave(-1, 1, 0.5)
    b.play(c3, clock, mode="+=", loop=True, immediate=True)
    r3 = T.TimeSeries().record(c3, clock, immediate=False)
    tr = T.Trace(c3, clock, 2.0)
    ru = T.TimeSeries().record(u5, clock, record_duration=2.0, immediate=False)
    # the device starts from the oracle's state after the immediate callbacks
    col[3] = c3.get()
    state.copy_(torch.tensor([tr.mean, tr.var, tr.start], dtype=torch.float64))
    waves = [torch.as_tensor(np.asarray(x.play_data, np.float64), device=dev) for x in (a, b)]
    c = T.TimeSeries().set_data([1.0, -2.0, 0.5], [0.0, 0.1, 0.2])
    late = None
    n_ev, cap = 7, 8
    ring = torch.zeros((3, cap), dtype=torch.float32, device=dev)
    arr = events.c_events(n_ev)

    def put(i, kind, column, row, first, end, **kw):
        e = arr[i]
        e.kind, e.column, e.row, e.first_tick, e.end_tick = kind, column.data_ptr(), row, first, end
        e.dtype = _lib.NW_DTYPE_U8 if column.dtype == torch.uint8 else _lib.NW_DTYPE_F32
        for k, v in kw.items():
            setattr(e, k, v)
    big = events.UNBOUNDED
    put(0, _lib.NW_EVENT_SET, col, 3, 1, 1 + 6, wave=waves[0].data_ptr(), length=6, first_index=0)
    put(1, _lib.NW_EVENT_ADD, col, 3, 1, big, wave=waves[1].data_ptr(), length=len(b.play_data),
        first_index=1, loop=1)
    put(2, _lib.NW_EVENT_RECORD, col, 3, 1, big, slot=0)
    put(3, _lib.NW_EVENT_TRACE, col, 3, 1, big, state=state.data_ptr(), alpha=tr.alpha, beta=tr.beta)
    n_u = events.num_samples(2.0, 0.1)
    put(4, _lib.NW_EVENT_RECORD, u8, 5, 1, 1 + n_u, slot=1)
    # launch 4 on: a '+=' player on row 10 and its recorder (registered after three ticks)
    put(5, _lib.NW_EVENT_ADD, col, 10, 4, 4 + 2, length=3, first_index=1)
    put(6, _lib.NW_EVENT_RECORD, col, 10, 4, big, slot=2)
    assert len(a.play_data) == 6
    wave_c = None
    # run_end of every sampler run
    arr[2].run_end = arr[3].run_end = arr[4].run_end = 5
    arr[6].run_end = 7
    tick = torch.zeros(1, dtype=torch.int64, device=dev)
    count = torch.tensor([n_ev], dtype=torch.int32, device=dev)
    table = torch.empty(ctypes.sizeof(arr), dtype=torch.uint8, device=dev)
    rng = np.random.default_rng(0)
    got = {0: [], 1: [], 2: []}
    L = _lib.load()
    for m in range(1, 41):
        if m == 4:
            r10 = None
            c.play(c10, clock, mode="+=", immediate=True)      # writes sample 0 now
            col[10] = c10.get()
            wave_c = torch.as_tensor(np.asarray(c.play_data, np.float64), device=dev)
            arr[5].wave = wave_c.data_ptr()
            arr[5].length = len(c.play_data)
            arr[5].end_tick = 4 + len(c.play_data) - 1
            r10 = T.TimeSeries().record(c10, clock, immediate=False)
        table.copy_(torch.as_tensor(np.frombuffer(bytes(arr), dtype=np.uint8).copy()))
        x = int(rng.integers(0, 256))
        u5.set(x)
        u8[5] = x
        clock.tick()
        _lib.check(L.nw_clock_tick(_lib.ptr(tick), _lib.ptr(count), _lib.ptr(table),
                                   _lib.ptr(ring), cap, None))
        torch.cuda.synchronize()
        assert col[3].item() == c3.get(), m
        assert col[10].item() == c10.get(), m
        if m % cap == 0 or m == 40:
            h = ring.cpu().numpy()
            for t in range(m - ((m - 1) % cap), m + 1):
                got[0].append(float(h[0, t % cap]))
                if t <= n_u:
                    got[1].append(int(h[1, t % cap]))
                if t >= 4:
                    got[2].append(float(h[2, t % cap]))
    assert tick.item() == 40
    assert got[0] == list(r3.data)
    assert got[1] == list(ru.data)
    assert got[2] == list(r10.data)
    mean, var, start = state.cpu().tolist()
    assert (mean, var, start) == (tr.mean, tr.var, tr.start)
    _say(f"table: 40 ticks, {len(got[0])} + {len(got[1])} + {len(got[2])} samples and the trace "
         "bitwise equal to the oracle")


def _foo_model(n):
    """Objects with one FP32 device attribute each: single-segment neurons, attribute
    'driving_voltage' (written only by advance(), which these tests do not call)."""
    from neuwon_amd.model import RxD_Model
    m = RxD_Model(time_step=0.1)
    objs = [m.Neuron([10.0 * i, 0, 0], 5).root for i in range(n)]
    m._ensure_built()
    return m, objs


def test_time_series_buffers_on_device(torch_gpu):
    """test_time.py:31-54 on a device attribute: 51 samples, then play back with '='; samples,
    timestamps and played values equal to the oracle's."""
    from neuwon_amd.model import TimeSeries
    m, (f, f2) = _foo_model(2)
    clock = T.Clock(0.1)
    of, of2 = T.Cell(0, _f32), T.Cell(0, _f32)
    b, ob = TimeSeries(), T.TimeSeries()
    f.driving_voltage = 0.0
    b.record(f, "driving_voltage")
    ob.record(of, clock)
    for i in range(50):
        f.driving_voltage = (i / 10) % 2
        of.set((i / 10) % 2)
        m.clock.tick()
        clock.tick()
    b.stop()
    ob.stop()
    assert len(b) == 51
    assert b.get_data() == list(ob.data)
    assert b.get_timestamps() == list(ob.timestamps)
    b.play(f2, "driving_voltage", mode="=")
    ob.play(of2, clock, mode="=")
    for i in range(6):
        m.clock.tick()
        clock.tick()
        assert f2.driving_voltage == pytest.approx((i / 10) % 2, 1e-6)
        assert f2.driving_voltage == of2.get()
    b.stop()
    b.clear()
    assert not len(b)
    m.clock.tick()


def test_waveforms_on_device(torch_gpu):
    """test_time.py:56-126 on device attributes: square / triangle / sawtooth / sine played in a
    loop for 500 periods of 73 ticks with recordings (through ring drains every 4096 ticks); no
    drift, the right frequency, every sample equal to the oracle's."""
    from neuwon_amd.model import RxD_Model, TimeSeries
    m = RxD_Model(time_step=1.0)
    objs = [m.Neuron([10.0 * i, 0, 0], 5).root for i in range(4)]
    m._ensure_built()
    clock = T.Clock(1.0)
    cells = [T.Cell(0, _f32) for _ in range(4)]
    p, n_cycles = 73, 500
    makers = ("square_wave", "triangle_wave", "sawtooth_wave", "sine_wave")
    for o, c, mk in zip(objs, cells, makers):
        getattr(TimeSeries(), mk)(-1, 1, p).play(o, "driving_voltage", mode="=", loop=True)
        getattr(T.TimeSeries(), mk)(-1, 1, p).play(c, clock, mode="=", loop=True)
    bufs = [TimeSeries().record(o, "driving_voltage") for o in objs]
    obufs = [T.TimeSeries().record(c, clock) for c in cells]
    for _ in range(round(n_cycles * p / 1.0)):
        m.clock.tick()
        clock.tick()
    assert [o.driving_voltage for o in objs[:3]] == [-1, -1, -1]
    assert objs[3].driving_voltage == pytest.approx(0, abs=1e-6)
    for buf, ob in zip(bufs, obufs):
        data = buf.get_data()
        assert data == list(ob.data)
        assert buf.get_timestamps() == list(ob.timestamps)
        fft = np.abs(np.fft.fft(data))
        freq = np.fft.fftfreq(len(fft), 1.0)
        maxf = next(f for f in freq[np.argsort(-fft)] if f >= 0)
        assert maxf == pytest.approx(1.0 / p, 1e-4)
    _say(f"waveforms: {len(bufs[0].get_data())} samples per recording equal to the oracle")


def test_traces_on_device(torch_gpu):
    """test_time.py:128-176: Trace on one object and TraceAll on a whole attribute, host writes
    every tick; FP64 trace state bitwise, TraceAll's FP32 mean and variance bitwise against the
    oracle with FP32 storage, and the reference's statistical bars."""
    from neuwon_amd.model import Trace, TraceAll
    m, objs = _foo_model(32)
    f = objs[0]
    clock = T.Clock(0.1)
    of = T.Cell(42, _f32)
    f.driving_voltage = 42
    t = Trace(f, "driving_voltage", 100)
    ot = T.Trace(of, clock, 100)
    for _ in range(3):
        f.driving_voltage = 42
        m.clock.tick()
        clock.tick()
    assert t.get_mean() == pytest.approx(42) and t.get_standard_deviation() == pytest.approx(0)
    t.reset()
    ot.reset()
    f.driving_voltage = 64
    of.set(64)
    m.clock.tick()
    clock.tick()
    assert t.get_mean() == pytest.approx(64) and t.get_standard_deviation() == pytest.approx(0)
    rng = np.random.default_rng(1)
    for _ in range(3000):
        x = rng.normal(12, 5)
        f.driving_voltage = x
        of.set(x)
        m.clock.tick()
        clock.tick()
    assert t.get_mean() == pytest.approx(12, rel=.05)
    assert t.get_standard_deviation() == pytest.approx(5, rel=.05)
    assert (t.mean, t.var, t.start) == (ot.mean, ot.var, ot.start)
    t.reset()
    # TraceAll
    seg = m.database.get_class("Segment")
    n = seg.size
    order = seg.row_of                      # creation id -> device row
    box = {"x": rng.normal(-120, 0.3, n).astype(np.float32)}
    m.database.set_data("Segment.driving_voltage", box["x"][np.argsort(order)])
    ta = TraceAll(m.database.get("Segment.driving_voltage"), 100)
    ota = T.TraceAll(lambda: box["x"].astype(np.float64), clock, 100,
                     store=lambda a: np.asarray(a, np.float32))
    for i in range(5000):                   # test_time.py:165: five time constants
        box["x"] = rng.normal(-120, 0.3, n).astype(np.float32)
        m.database.set_data("Segment.driving_voltage", box["x"][np.argsort(order)])
        m.clock.tick()
        clock.tick()
    mean = ta.get_mean()[order]
    var = ta.get_variance()[order]
    assert np.array_equal(mean, ota.mean) and np.array_equal(var, ota.var)
    assert np.all(np.abs(mean + 120) <= 0.05 * 120)
    assert np.all(np.abs(np.sqrt(var) - 0.3) <= 0.10 * 0.3)
    assert "driving_voltage_mean" in seg.columns and "driving_voltage_variance" in seg.columns
    m.check()
    _say("Trace and TraceAll equal to the oracle")


def _hh_run(step_graph, steps, ring=None, symmetric=False):
    """Two hh neurons with dendrites: recordings of voltages and of AP_detected, a Trace and a
    TraceAll; host work between steps (reads, a magnitude change, a second pulse, a clock
    reset).  Returns the recordings and the final state."""
    from neuwon_amd import events
    from neuwon_amd.model import RxD_Model, TimeSeries, Trace, TraceAll
    if ring is not None:
        events.RING_TICKS = ring
    try:
        m = RxD_Model(time_step=0.025, temperature=6.3, mechanisms=["hh"], step_graph=step_graph,
                      symmetric_propagator=symmetric)
        neurons = [m.Neuron([300.0 * k, 0, 0], 12) for k in range(2)]
        segs = []
        for nrn in neurons:
            root = nrn.root
            d = root.add_section([root.coordinates[0], 120, 0], 2, maximum_segment_length=15)
            for s in [root] + d:
                s.insert("hh")
            segs += [root] + d
        neurons[0].set_AP_detector()
        recs = [TimeSeries().record(s, "voltage") for s in segs[::3]]
        recs.append(TimeSeries().record(neurons[0], "AP_detected"))
        recs.append(TimeSeries().record(segs[1], "sum_current", record_duration=2.0,
                                        discard_after=0.5))
        tr = Trace(segs[0], "voltage", 5.0)
        segs[0].inject_current(0.25, 1.0)
        reads = []
        for k in range(steps):
            if k == 40:
                reads.append(segs[2].voltage)                          # a host read (sync)
            if k == 60:
                inst = m.mechanisms["hh"]._db_class
                m.database.get_class("hh").columns["magnitude"].set_item(0, 0.7 * inst.columns["magnitude"].get_item(0))
            if k == 90:
                segs[len(segs) // 2].inject_current(0.3, 0.5)          # second neuron
            if k == 120:
                ta = TraceAll(m.database.get("Segment.voltage"), 2.0)
            if k == 150:
                m.clock.reset()
            m.advance()
        v = m.database.get_data("Segment.voltage").copy()
        out = {"recs": [(r.get_data(), r.get_timestamps()) for r in recs], "v": v, "reads": reads,
               "trace": (tr.get_mean(), tr.get_variance()), "ta": ta.get_mean().copy(),
               "ticks": (m.clock.ticks, m.input_hook.ticks),
               "graph": m._graph is not None}
        return out
    finally:
        events.RING_TICKS = 4096


def test_step_graph_matches_eager(torch_gpu):
    """300 steps with step_graph=False and True: every recorded sample, timestamp, trace value and
    the final voltages bitwise equal; a ring of 16 ticks in the graph run forces drains between
    replays."""
    eager = _hh_run(False, 300)
    graph = _hh_run(True, 300, ring=16)
    assert graph["graph"] and not eager["graph"]
    for (d0, t0), (d1, t1) in zip(eager["recs"], graph["recs"]):
        assert d0 == d1 and t0 == t1
    assert np.array_equal(eager["v"], graph["v"])
    assert eager["reads"] == graph["reads"]
    assert eager["trace"] == graph["trace"]
    assert np.array_equal(eager["ta"], graph["ta"])
    assert eager["ticks"] == graph["ticks"] == (150, 600)
    spikes = sum(eager["recs"][-2][0])
    v0 = eager["recs"][0][0]
    _say(f"step graph: 300 steps bitwise equal to eager ({spikes} AP detections, soma peak "
         f"{max(v0):.1f} mV, {len(v0)} samples per voltage recording)")
    assert spikes >= 1 and max(v0) > 20


def test_step_graph_speed(torch_gpu):
    """A 1,000-segment model with hh and two recordings: time per advance() eagerly and as graph
    replays (same results)."""
    import time
    import torch
    from neuwon_amd.model import RxD_Model, TimeSeries
    res = {}
    for graph in (False, True):
        m = RxD_Model(time_step=0.025, temperature=6.3, mechanisms=["hh"], step_graph=graph)
        for k in range(10):
            root = m.Neuron([300.0 * k, 0, 0], 12).root
            for s in [root] + root.add_section([300.0 * k, 500, 0], 2, maximum_segment_length=5.1):
                s.insert("hh")
        n = m.database.get_class("Segment").size
        rec = TimeSeries().record(root, "voltage")
        root.inject_current(0.2, 1.0)
        for _ in range(20):
            m.advance()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        steps = 400
        for _ in range(steps):
            m.advance()
        torch.cuda.synchronize()
        res[graph] = ((time.perf_counter() - t0) / steps, m.database.get_data("Segment.voltage").copy(),
                      rec.get_data(), n)
    (te, ve, re, n), (tg, vg, rg, _) = res[False], res[True]
    _say(f"{n} segments: {te * 1e6:.0f} us per advance() eagerly, {tg * 1e6:.0f} us as graph "
         f"replays ({te / tg:.1f}x)")
    assert np.array_equal(ve, vg) and re == rg
    assert tg < te
