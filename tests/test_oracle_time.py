"""The clock-callback oracle (oracle/time_callbacks.py) against the reference's own tests of
neuwon/database/time.py (neuwon/database/tests/test_time.py:7-176), on the CPU."""
import numpy as np
import pytest
from pytest import approx

from oracle import time_callbacks as T


def test_clock():
    """test_time.py:7-21."""
    c = T.Clock(1 / 40)
    assert c() == 0
    for _ in range(40):
        c.tick()
    assert c() == 1
    c.reset()
    assert c() == 0
    c.set_time(99)
    for _ in range(40):
        c.tick()
    assert c() == 100


def test_time_series_buffers():
    """test_time.py:31-54."""
    clock = T.Clock(0.1)
    f = T.Cell(0)
    b = T.TimeSeries().record(f, clock)
    for i in range(50):
        f.set((i / 10) % 2)
        clock.tick()
    b.stop()
    assert len(b.data) == 51
    f2 = T.Cell(0)
    b.play(f2, clock, mode="=")
    for i in range(6):
        clock.tick()
        assert f2.get() == approx((i / 10) % 2, 1e-9)


def test_waveforms_basic():
    """test_time.py:56-81."""
    clock = T.Clock(0.1)
    f = T.Cell(0)
    sqr = T.TimeSeries().square_wave(1, 0, 1)
    sqr.play(f, clock, mode="=")
    assert f.get() == 1
    for _ in range(7):
        clock.tick()
    assert f.get() == 0
    assert sum(sqr.play_data) == 5 and len(sqr.play_data) == 10
    tri = T.TimeSeries().triangle_wave(0, 1, 2)
    tri.play(f, clock, mode="=", loop=True)
    assert max(tri.play_data) == approx(1, .01)
    saw = T.TimeSeries().sawtooth_wave(0, 1, 100)
    f2 = T.Cell(0)
    saw.play(f2, clock, mode="=", loop=True)
    assert max(saw.play_data) == approx(1, .01)
    for _ in range(int(99.9 / clock.dt)):
        clock.tick()
    assert f2.get() == approx(1, .01)


def test_waveform_frequency():
    """test_time.py:83-126: 500 cycles of period 73 ticks, no drift, the right frequency."""
    clock = T.Clock(1)
    p, n_cycles = 73, 500
    cells = [T.Cell(0) for _ in range(4)]
    T.TimeSeries().square_wave(-1, 1, p).play(cells[0], clock, mode="=", loop=True)
    T.TimeSeries().triangle_wave(-1, 1, p).play(cells[1], clock, mode="=", loop=True)
    T.TimeSeries().sawtooth_wave(-1, 1, p).play(cells[2], clock, mode="=", loop=True)
    T.TimeSeries().sine_wave(-1, 1, p).play(cells[3], clock, mode="=", loop=True)
    bufs = [T.TimeSeries().record(c, clock) for c in cells]
    for _ in range(round(n_cycles * p / clock.dt)):
        clock.tick()
    assert [c.get() for c in cells[:3]] == [-1, -1, -1]
    assert cells[3].get() == approx(0, abs=1e-12)
    for buf in bufs:
        data = np.array(buf.data)
        fft = np.abs(np.fft.fft(data))
        freq = np.fft.fftfreq(len(fft), clock.dt)
        maxf = next(f for f in freq[np.argsort(-fft)] if f >= 0)
        assert maxf == approx(1.0 / p, 1e-4)


def test_obj_trace():
    """test_time.py:128-152."""
    clock = T.Clock(0.1)
    f = T.Cell(42)
    t = T.Trace(f, clock, 100)
    for _ in range(3):
        f.set(42)
        clock.tick()
    assert t.get_mean() == approx(42)
    assert t.get_variance() ** 0.5 == approx(0)
    t.reset()
    f.set(64)
    clock.tick()
    assert t.get_mean() == approx(64)
    assert t.get_variance() ** 0.5 == approx(0)
    rng = np.random.default_rng(1)
    for _ in range(3000):
        f.set(rng.normal(12, 5))
        clock.tick()
    assert t.get_mean() == approx(12, rel=.05)
    assert t.get_variance() ** 0.5 == approx(5, rel=.05)


def test_attr_trace():
    """test_time.py:154-176."""
    mean, std, n = -120, 0.3, 32
    clock = T.Clock(0.1)
    rng = np.random.default_rng(2)
    box = {"x": rng.normal(mean, std, n)}
    t = T.TraceAll(lambda: box["x"], clock, 100)
    for _ in range(5000):
        box["x"] = rng.normal(mean, std, n)
        clock.tick()
    assert np.all(np.abs(t.mean - mean) <= 0.05 * abs(mean))
    assert np.all(np.abs(np.sqrt(t.var) - std) <= 0.10 * std)


def test_discard_after_and_duration():
    """time.py:224-232: record_duration counts samples by repeated subtraction of dt, and the
    discard rule keeps the samples within discard_after of the newest."""
    clock = T.Clock(0.1)
    f = T.Cell(0)
    b = T.TimeSeries().record(f, clock, record_duration=1.0, discard_after=0.35)
    for i in range(30):
        f.set(i + 1)
        clock.tick()
    # 0.1 subtracted from 1.0 ten times leaves 1.4e-16 > 0: eleven samples (ticks 0..10)
    assert list(b.timestamps) == approx([0.7, 0.8, 0.9, 1.0])
    assert list(b.data) == [7, 8, 9, 10]
    assert not b.is_recording()
