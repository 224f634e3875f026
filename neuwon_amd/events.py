"""Device-side clock callbacks: play back, recording and exponentially weighted statistics.

The reference runs these as Python callbacks after every tick of a ``Clock``
(neuwon/database/time.py:114-117): ``TimeSeries._play_implementation`` writes one sample into an
attribute (287-302), ``TimeSeries._record_implementation`` appends the attribute's value (224-232),
``Trace._callback`` (509-515) and ``TraceAll._callback`` (627-637) update moving statistics.  Each
of those reads or writes one element of a device array, so on the GPU every callback would be a
host round trip per tick.  Here the callbacks of one clock form a table in device memory that one
launch per tick (``nw_clock_tick``) applies in registration order; recorded samples go to a ring
in device memory that the host drains when it fills or when the data are read.

Ticks are counted by the table itself (a monotonic count of launches, kept in device memory so a
captured step graph replays successive ticks), not by the clock's ``ticks``: a callback in the
reference acts on every tick after it was registered, also after ``Clock.reset`` or ``set_time``
move the clock.  The clock value of each launch (for timestamps) is kept on the host as a list of
epochs, one per jump of the clock.
"""
import math
from collections import deque

import numpy as np

from . import _lib

UNBOUNDED = (1 << 62)
RING_TICKS = 4096


def _dtype_code(tensor):
    name = str(tensor.dtype)
    if name == "torch.float32":
        return _lib.NW_DTYPE_F32
    if name == "torch.uint8":
        return _lib.NW_DTYPE_U8
    if name == "torch.int32":
        return _lib.NW_DTYPE_I32
    raise TypeError(f"no device callback for columns of {tensor.dtype}")


def device_column(db_object, component):
    """The device column behind ``db_object.component``, or None when the attribute has no device
    storage (host-only attributes, class attributes, objects of Python plugins)."""
    cls = getattr(type(db_object), "_db_class", None)
    if cls is None or getattr(type(db_object), "_model", None) is None:
        return None
    col = cls.columns.get(component)
    if col is None or not col.device or col.shape:
        return None
    return col


class _Entry:
    kind = None

    def __init__(self, events, db_object, component):
        self.events = events
        self.obj, self.component = db_object, component
        self.first = events.launched + 1          # first launch this entry acts on
        self.end = UNBOUNDED

    def target(self):
        col = device_column(self.obj, self.component)
        cls = type(self.obj)._db_class
        return col, int(cls.row_of[self.obj._cid])


class Player(_Entry):
    """TimeSeries.play: sample ``first_index + (m - first)`` on launch m."""

    def __init__(self, events, db_object, component, wave, first_index, loop, mode):
        super().__init__(events, db_object, component)
        self.wave = np.ascontiguousarray(wave, dtype=np.float64)
        self.first_index, self.loop = int(first_index), bool(loop)
        self.kind = {"=": _lib.NW_EVENT_SET, "+=": _lib.NW_EVENT_ADD}[mode]
        if not self.loop:
            self.end = self.first + (len(self.wave) - self.first_index)
        self.wave_dev = None

    def playing(self):
        """True until the launch that writes the last sample has run (time.py:296-302)."""
        return self.loop or self.events.launched < self.end - 1


class Recorder(_Entry):
    """TimeSeries.record: one sample per launch in [first, end)."""
    kind = _lib.NW_EVENT_RECORD

    def __init__(self, events, db_object, component, num_samples, sink):
        super().__init__(events, db_object, component)
        if num_samples != math.inf:
            self.end = self.first + int(num_samples)
        self.sink = sink          # receives (launch, value) batches when the ring drains
        self.slot = -1

    def recording(self):
        return self.events.launched < self.end - 1 if self.end != UNBOUNDED else True


class Tracer(_Entry):
    """Trace: FP64 mean, variance and start weight on the device."""
    kind = _lib.NW_EVENT_TRACE

    def __init__(self, events, db_object, component, state, alpha, beta):
        super().__init__(events, db_object, component)
        self.state, self.alpha, self.beta = state, float(alpha), float(beta)


class DeviceEvents:
    """The device-side callbacks of one clock."""

    def __init__(self, model, clock):
        self.model, self.clock = model, clock
        self.entries = []
        self.trace_all = []
        self.launched = 0                      # launches so far (host view of the device count)
        self.epochs = [(1, clock.ticks + 1)]   # (launch m, clock ticks at launch m)
        self.flushed = 0                       # launches whose samples are on the host
        self.dev_count = 0                     # what the device tick count holds
        self.dirty = True
        self.version = None
        self.torch = None
        self.tick = self.count = self.table = self.ring = None
        self.ring_rows = 0
        self.cap = RING_TICKS

    # ---------------------------------------------------------------- registration
    def active(self):
        return bool(self.entries) or bool(self.trace_all)

    def add(self, entry):
        self.entries.append(entry)
        self.dirty = True
        return entry

    def remove(self, entry):
        if entry in self.entries:
            self.flush()
            self.entries.remove(entry)
            self.dirty = True

    def clock_moved(self, next_ticks):
        """The clock was set (reset, set_time, a snapshot): the next launch sees next_ticks."""
        self.epochs.append((self.launched + 1, int(next_ticks)))

    def clock_ticks_at(self, m):
        """The clock's tick count when launch m ran."""
        for start, ticks in reversed(self.epochs):
            if m >= start:
                return ticks + (m - start)
        return self.epochs[0][1] + (m - self.epochs[0][0])

    # ---------------------------------------------------------------- device table
    def _ensure_buffers(self, torch, dev):
        if self.tick is None:
            self.torch = torch
            self.tick = torch.zeros(1, dtype=torch.int64, device=dev)
            self.count = torch.zeros(1, dtype=torch.int32, device=dev)
            self.tick.fill_(self.launched)
            self.dev_count = self.launched

    def _upload(self):
        torch = self.torch
        model = self.model
        dev = model._dev
        self.flush()
        # entries that can no longer act are dropped
        self.entries = [e for e in self.entries
                        if e.end > self.launched + 1 or isinstance(e, Tracer)]
        recs = [e for e in self.entries if isinstance(e, Recorder)]
        for k, e in enumerate(recs):
            e.slot = k
        if len(recs) > self.ring_rows:
            rows = max(len(recs), 2 * self.ring_rows, 4)
            self.ring = torch.zeros((rows, self.cap), dtype=torch.float32, device=dev)
            self.ring_rows = rows
        n = len(self.entries)
        arr = (_lib.NwEvent * max(n, 1))()
        keep = []
        for i, e in enumerate(self.entries):
            col, row = e.target()
            if col is None or col.dev is None:
                raise RuntimeError(f"{e.component} of {e.obj!r} has no device storage")
            ev = arr[i]
            ev.kind, ev.dtype = e.kind, _dtype_code(col.dev)
            ev.column, ev.row = col.dev.data_ptr(), row
            ev.first_tick, ev.end_tick = e.first, e.end
            keep.append(col.dev)
            if isinstance(e, Player):
                if ev.dtype != _lib.NW_DTYPE_F32:
                    raise TypeError(f"play back into {e.component}: FP32 attributes only")
                if e.wave_dev is None or e.wave_dev.device != dev:
                    e.wave_dev = torch.as_tensor(e.wave, device=dev)
                ev.wave, ev.length = e.wave_dev.data_ptr(), len(e.wave)
                ev.first_index, ev.loop = e.first_index, int(e.loop)
            elif isinstance(e, Recorder):
                ev.slot = e.slot
            else:
                ev.state, ev.alpha, ev.beta = e.state.data_ptr(), e.alpha, e.beta
        # runs of samplers: each knows where its run ends
        end = n
        for i in range(n - 1, -1, -1):
            if arr[i].kind in (_lib.NW_EVENT_SET, _lib.NW_EVENT_ADD):
                end = i
            else:
                arr[i].run_end = end
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        if self.table is None or self.table.numel() < raw.size:
            self.table = torch.empty(max(raw.size, 2 * (0 if self.table is None else
                                                        self.table.numel())),
                                     dtype=torch.uint8, device=dev)
        self.table[:raw.size].copy_(torch.as_tensor(raw.copy()), non_blocking=False)
        self.count.fill_(n)
        self._columns = keep
        self.dirty = False
        self.version = model._build_count

    def prepare(self, k):
        """Before k launches: the table is current and the ring has room for k more samples."""
        import torch
        model = self.model
        model._ensure_built()
        self._ensure_buffers(torch, model._dev)
        if self.dirty or self.version != model._build_count:
            self._upload()
        if self.dev_count != self.launched:     # ticks counted on the host while inactive
            self.tick.fill_(self.launched)
            self.dev_count = self.launched
        if self.ring is not None and self.launched + k - self.flushed > self.cap:
            self.flush()

    def launch_kernels(self, stream):
        """nw_clock_tick and the TraceAll updates of one tick (no host bookkeeping)."""
        P = _lib.ptr
        _lib.check(_lib.load().nw_clock_tick(
            P(self.tick), P(self.count), P(self.table),
            P(self.ring) if self.ring is not None else None, self.cap, stream))
        for ta in self.trace_all:
            ta._launch(stream)

    def launch(self):
        """One tick, eagerly (Clock.tick)."""
        self.prepare(1)
        self.launch_kernels(self.model._stream())
        self.launched += 1
        self.dev_count += 1

    def graph_key(self):
        """What a captured launch of this table depends on."""
        return (self.tick.data_ptr(), self.count.data_ptr(), self.table.data_ptr(),
                0 if self.ring is None else self.ring.data_ptr(),
                tuple(id(t) for t in self.trace_all))

    # ---------------------------------------------------------------- recording
    def flush(self):
        """Copy the recorded samples of launches flushed+1 .. launched to their TimeSeries."""
        if self.ring is None or self.flushed >= self.launched:
            self.flushed = self.launched
            return
        host = self.ring.cpu().numpy()
        lo, hi = self.flushed + 1, self.launched
        for e in self.entries:
            if not isinstance(e, Recorder) or e.slot < 0:
                continue
            a, b = max(lo, e.first), min(hi, e.end - 1)
            if a > b:
                continue
            ms = np.arange(a, b + 1)
            e.sink(ms, host[e.slot, ms % self.cap])
        self.flushed = self.launched


class SampleBuffer:
    """Host side of one recording: the reference's deque of samples and timestamps, with the
    discard rule applied after every appended sample (time.py:224-232)."""

    def __init__(self, series, events, discard_after, integer):
        self.series, self.events = series, events
        self.discard_after = float(discard_after)
        self.integer = integer

    def __call__(self, launches, values):
        dt = self.events.clock.dt
        data, ts = self.series.data, self.series.timestamps
        for m, x in zip(launches.tolist(), values.tolist()):
            data.append(int(x) if self.integer else x)
            ts.append(self.events.clock_ticks_at(m) * dt)
            while ts[-1] - ts[0] > self.discard_after:
                data.popleft()
                ts.popleft()


def num_samples(record_duration, dt):
    """How many samples a recording of record_duration takes (time.py:224-232: one sample per
    callback while the remaining duration, decremented by dt each sample, is positive)."""
    rd = float(record_duration)
    if rd == math.inf:
        return math.inf
    n = 0
    while rd > 0:
        n += 1
        rd -= dt
    return n


def as_deque(x):
    return x if isinstance(x, deque) else deque(x)


def c_events(n):
    """A zeroed ctypes array of n events (tests build tables by hand with it)."""
    return (_lib.NwEvent * n)()

