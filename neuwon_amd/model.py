"""Drop-in ``RxD_Model`` / ``Model`` whose ``advance()`` runs on the MI355X.

Mirrors the public surface of neuwon/rxd/rxd_model.py:8-123 and neuwon/model.py:6-22:

    m = RxD_Model(time_step=0.1, temperature=6.3, mechanisms=['hh'])
    root = m.Neuron([0, 0, 0], 10).root
    tip = root.add_segment([20, 0, 0], 1)
    tip.insert('hh')
    root.inject_current(0.5, 1.0)
    m.advance()                          # one staggered step on the GPU
    m.database.get_data('Segment.voltage')

Storage follows the reference's structure-of-arrays Database (database.py:11-259): every class
(Segment, Neuron, each mechanism) keeps one array per attribute.  While a model is being built the
arrays live on the host; the first ``advance()`` sorts segments by cell (the reference's
``sort_key='neuron'``, neuron.py:12), builds the per-cell propagators on the device and moves the
state to device memory.  ``get_data`` in the default "host" memory space returns a host copy whose
writes are uploaded before the next step (the reference returns a live view); inside
``database.using_memory_space('cuda')`` (alias 'hip') it returns the live device tensor.

Each half step runs nw_accumulate + nw_electric_half_step, the mechanisms run their own kernels
(nw_hh_mechanism, nw_lti_advance) or, for Python plugin mechanisms, their ``advance()``; there is
no CPU execution path.
"""
import collections
import contextlib
import ctypes
import math
import os

import numpy as np

from . import _lib, channels, events, geometry
from .population import CHEBYSHEV_MAX_CELL

EPSILON = float(np.finfo(np.float64).eps)
NULL = -1


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("neuwon_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


# ================================================================================================
# Clock and TimeSeries (neuwon/database/time.py)
# ================================================================================================
class Clock:
    """Tick counter with callbacks; time.py:54-121.  A clock of a model may also carry device
    callbacks (``events.DeviceEvents``), which run first on every tick."""

    def __init__(self, tick_period, units=""):
        self.dt = self.time_step = self.tick_period = float(tick_period)
        self.ticks = 0
        self.units = str(units)
        self._callbacks = []
        self._events = None

    def get_time(self):
        return self.ticks * self.dt

    __call__ = get_time

    def get_ticks(self):
        return self.ticks

    def get_time_step(self):
        return self.dt

    get_tick_period = get_time_step

    def get_units(self):
        return self.units

    def register_callback(self, function, period=1, weakref=False):
        """Callbacks run after every tick, in registration order, and are removed when they
        return True (time.py:90-103)."""
        if period != 1:
            raise NotImplementedError("period != 1")
        self._callbacks.append(function)

    def _run_callbacks(self):
        current, self._callbacks = self._callbacks, []
        keep = [fn for fn in current if not fn()]
        self._callbacks = keep + self._callbacks   # plus any registered while running

    def _set_ticks(self, ticks, callbacks=False):
        """Move the clock; with callbacks (reset, set_time) they run at the new time."""
        self.ticks = int(ticks)
        if self._events is not None:
            self._events.clock_moved(self.ticks if callbacks else self.ticks + 1)
        if callbacks:
            self._call()

    def reset(self):
        """time.py:105-108: zero, then the callbacks."""
        self._set_ticks(0, callbacks=True)

    def set_time(self, new_time):
        self._set_ticks(round(float(new_time) / self.dt), callbacks=True)

    def tick(self):
        self.ticks += 1
        self._call()

    def _call(self):
        ev = self._events
        if ev is not None:
            if ev.active():
                ev.launch()
            else:
                ev.launched += 1
        self._run_callbacks()

    def _count(self, k):
        """k ticks whose device work a replayed step graph has done (such a clock has no host
        callbacks)."""
        self.ticks += k
        self._events.launched += k
        self._events.dev_count += k

    def is_now(self, time):
        return round(float(time) / self.dt) == self.ticks


class TimeSeries:
    """Buffer of (timestamp, value) samples with play back and recording (time.py:123-408 subset).
    Play back into, and recording of, a device attribute of a model run on the device, one table
    launch per tick for all of a clock's series (``events.DeviceEvents``); other attributes use
    host callbacks."""

    def __init__(self, initial_data=None, timestamps=None):
        self.data = collections.deque(initial_data) if initial_data is not None else collections.deque()
        self.timestamps = collections.deque(timestamps) if timestamps is not None else collections.deque()
        self._playing = None
        self._recording = False
        self._entry = None
        self.play_data = None

    def _sync(self):
        if self._entry is not None and isinstance(self._entry, events.Recorder):
            self._entry.events.flush()

    def get_data(self):
        self._sync()
        return list(self.data)

    def get_timestamps(self):
        self._sync()
        return list(self.timestamps)

    def __len__(self):
        self._sync()
        return len(self.data)

    def set_data(self, data_samples, timestamps):
        assert self.is_stopped()
        self.data = collections.deque(data_samples)
        self.timestamps = collections.deque(timestamps)
        assert len(self.data) == len(self.timestamps)
        return self

    def clear(self):
        assert self.is_stopped()
        self.data, self.timestamps = collections.deque(), collections.deque()
        return self

    def constant_wave(self, value, duration):
        """time.py:406-408."""
        return self.set_data([value, value], [0, duration])

    def square_wave(self, extremum_1, extremum_2, period, duty_cycle=0.5):
        """time.py:410-420: one cycle."""
        a, b, period, duty_cycle = float(extremum_1), float(extremum_2), float(period), float(duty_cycle)
        assert 0.0 <= duty_cycle <= 1.0
        mid = period * duty_cycle
        return self.set_data([a, a, b, b], [0, mid, mid, period])

    def sine_wave(self, extremum_1, extremum_2, period):
        """time.py:422-432: one cycle, 1000 points."""
        a, b, period = float(extremum_1), float(extremum_2), float(period)
        amplitude = 0.5 * (a - b)
        offset = b + amplitude
        return self.set_data(offset + amplitude * np.sin(np.linspace(0.0, 2.0 * np.pi, 1000)),
                             np.linspace(0, period, 1000))

    def triangle_wave(self, extremum_1, extremum_2, period):
        """time.py:434-439."""
        a, b, period = float(extremum_1), float(extremum_2), float(period)
        return self.set_data([a, b, a], [0.0, 0.5 * period, period])

    def sawtooth_wave(self, extremum_1, extremum_2, period):
        """time.py:441-446."""
        a, b, period = float(extremum_1), float(extremum_2), float(period)
        return self.set_data([a, b, a], [0.0, period, period])

    def is_playing(self):
        if isinstance(self._entry, events.Player):
            return self._entry.playing()
        return self._playing is not None

    def is_recording(self):
        if isinstance(self._entry, events.Recorder):
            return self._entry.recording()
        return self._recording

    def is_stopped(self):
        return not self.is_playing() and not self.is_recording()

    def stop(self):
        if self._entry is not None:
            self._entry.events.remove(self._entry)
            self._entry = None
        self._playing = None
        self._recording = False
        return self

    def play(self, db_object, component, mode="+=", loop=False, clock=None, immediate=True):
        """Write one value per clock tick into ``db_object.component`` (time.py:245-302)."""
        assert self.is_stopped()
        self._entry = None
        model = getattr(type(db_object), "_model", None)
        clock = clock or db_object._model.clock
        dt = clock.get_tick_period()
        start = math.floor(self.timestamps[0] / dt)
        end = math.ceil(self.timestamps[-1] / dt)
        ts = dt * np.arange(start, end)
        import scipy.interpolate
        f = scipy.interpolate.interp1d(list(self.timestamps), list(self.data),
                                       fill_value=(self.data[0], self.data[-1]), bounds_error=False)
        values = np.vectorize(f)(ts)
        self.play_data = values
        if mode not in ("=", "+="):
            raise NotImplementedError(mode)
        self._playing = dict(obj=db_object, component=component, data=values, index=0,
                             loop=bool(loop), mode=mode)
        if events.device_column(db_object, component) is not None and model is not None:
            if immediate:
                self._play_tick()
                if self._playing is None:
                    return self
            ev = model._events_for(clock)
            self._entry = ev.add(events.Player(ev, db_object, component, values,
                                               self._playing["index"], loop, mode))
            self._playing = None
            return self
        clock.register_callback(self._play_tick)
        if immediate:
            self._play_tick()
        return self

    def _play_tick(self):
        st = self._playing
        if st is None:
            return True
        value = float(st["data"][st["index"]])
        obj, name = st["obj"], st["component"]
        if st["mode"] == "=":
            setattr(obj, name, value)
        else:
            setattr(obj, name, value + getattr(obj, name))
        st["index"] += 1
        if st["index"] >= len(st["data"]):
            if st["loop"]:
                st["index"] = 0
            else:
                self._playing = None
                return True
        return False

    def record(self, db_object, component, record_duration=np.inf, discard_after=np.inf,
               clock=None, immediate=True):
        """Sample ``db_object.component`` after every clock tick (time.py:189-232)."""
        assert self.is_stopped()
        self._entry = None
        model = getattr(type(db_object), "_model", None)
        self._clock = clock or db_object._model.clock
        self._rec = (db_object, component)
        self._recording = True
        self._remaining = float(record_duration)
        self._discard_after = float(discard_after)
        col = events.device_column(db_object, component)
        if col is not None and model is not None:
            n = events.num_samples(record_duration, self._clock.dt)
            if immediate and n:
                self._record_tick()
                n -= 1
            ev = model._events_for(self._clock)
            sink = events.SampleBuffer(self, ev, discard_after,
                                       integer=np.dtype(col.device_dtype).kind in "ui")
            self._entry = ev.add(events.Recorder(ev, db_object, component, n, sink))
            self._recording = False
            return self
        self._clock.register_callback(self._record_tick)
        if immediate:
            self._record_tick()
        return self

    def _record_tick(self):
        if not self._recording or self._remaining <= 0:
            self._recording = False
            return True
        obj, name = self._rec
        self.data.append(getattr(obj, name))
        self.timestamps.append(self._clock())
        while self.timestamps[-1] - self.timestamps[0] > self._discard_after:
            self.data.popleft()
            self.timestamps.popleft()
        self._remaining -= self._clock.dt
        return False


class Trace:
    """Exponentially weighted mean and standard deviation of one attribute of one object
    (time.py:461-537).  On a device attribute the update runs in the clock's device table in
    FP64 with the reference's roundings; otherwise as a host callback."""

    def __init__(self, db_object, component, period):
        self.db_object, self.component_name = db_object, str(component)
        self.period = float(period)
        assert self.period > 0.0
        model = type(db_object)._model
        self.clock = model.clock
        dt = self.clock.get_tick_period()
        self.alpha = float(np.exp(-dt / self.period))
        self.beta = 1.0 - self.alpha
        self.mean = self.var = 0.0
        self.start = 1.0
        self._entry = None
        self._callback()           # the current value is the first sample
        if events.device_column(db_object, component) is not None:
            torch = _torch()
            model._ensure_built()
            self._state = torch.tensor([self.mean, self.var, self.start], dtype=torch.float64,
                                       device=model._dev)
            ev = model._events_for(self.clock)
            self._entry = ev.add(events.Tracer(ev, db_object, component, self._state,
                                               self.alpha, self.beta))
        else:
            self.clock.register_callback(self._callback)

    def _pull(self):
        if self._entry is not None:
            self.mean, self.var, self.start = self._state.cpu().tolist()

    def reset(self):
        self.mean = self.var = 0.0
        self.start = 1.0
        if self._entry is not None:
            self._state.copy_(_torch().tensor([0.0, 0.0, 1.0], dtype=_torch().float64))

    def _callback(self):
        value = float(getattr(self.db_object, self.component_name))
        self.mean += self.beta * (value - self.mean)
        self.start *= self.alpha
        true_mean = self.mean / (1.0 - self.start)
        self.var += self.beta * (value - true_mean) ** 2
        self.var *= self.alpha

    def get_mean(self):
        self._pull()
        return self.mean / (1.0 - self.start)

    def get_variance(self):
        self._pull()
        return self.var / (1.0 - self.start)

    def get_standard_deviation(self):
        return self.get_variance() ** 0.5


class TraceAll:
    """Exponentially weighted mean and variance of every instance of an attribute, kept as new
    attributes of its class (time.py:539-647): ``nw_trace_all`` after every tick of the model's
    clock."""

    def __init__(self, db_component, period, mean=True, variance=True):
        self.component = db_component
        self.period = float(period)
        assert self.period > 0.0 and mean
        cls = db_component.db_class
        model = cls.database._model
        name = db_component.name
        if mean is True:
            mean = f"{name}_mean"
        if variance is True:
            variance = f"{name}_variance"
        model._ensure_built()
        self.model, self.cls = model, cls
        self.clock = model.clock
        dt = self.clock.get_tick_period()
        self.alpha = float(np.exp(-dt / self.period))
        self.beta = 1.0 - self.alpha
        self.mean = cls.add_attribute(mean, db_component.initial_value, units=db_component.units)
        self.var = cls.add_attribute(variance, 0.0, valid_range=(0.0, None)) if variance else None
        order = cls._order if getattr(cls, "_order", None) is not None else np.arange(cls.size)
        for col in (self.mean, self.var):
            if col is not None:
                col.to_device(order, model._dev)
        self.mean.dev.copy_(db_component.dev)
        self._launch(model._stream())           # the current value is the first sample
        model._events_for(self.clock).trace_all.append(self)

    def _launch(self, stream):
        P = _lib.ptr
        x = self.component.dev
        _lib.check(_lib.load().nw_trace_all(
            x.numel(), P(x), P(self.mean.dev), P(self.var.dev) if self.var is not None else None,
            self.alpha, self.beta, stream))

    def reset(self):
        self.mean.dev.copy_(self.component.dev)
        if self.var is not None:
            self.var.dev.zero_()

    def get_mean(self):
        return self.mean.host_array()

    def get_variance(self):
        assert self.var is not None
        return self.var.host_array()

    def get_standard_deviation(self):
        return self.get_variance() ** 0.5


# ================================================================================================
# Structure-of-arrays storage
# ================================================================================================
class Column:
    """One attribute of one class: host list while building, device tensor once the model runs."""

    def __init__(self, db_class, name, initial_value=None, dtype=np.float64, units=None,
                 valid_range=None, device_dtype="float32", shape=(), device=True, on_write=None):
        self.db_class, self.name = db_class, name
        self.initial_value, self.dtype, self.units = initial_value, np.dtype(dtype), units
        self.valid_range, self.device_dtype, self.shape = valid_range, device_dtype, tuple(shape)
        self.device = bool(device)   # False: host-only (references between objects, geometry input)
        self.on_write = on_write     # called after a host-only column changes
        self.host = []
        self.dev = None          # torch tensor, rows in device (sorted) order
        self.checked_out = None  # host copy handed out by get_data('host'), uploaded on next use

    def get_name(self):
        return self.name

    def get_units(self):
        return self.units

    def get_initial_value(self):
        return self.initial_value

    def append(self, value=None):
        if value is None:
            value = self.initial_value
        self.host.append(value)

    # ---- element access (rows in device order once finalized)
    def get_item(self, row):
        self._flush()
        if self.dev is not None:
            return self.dev[row].item()
        return self.host[row]

    def set_item(self, row, value):
        self._flush()
        if self.dev is not None:
            self.dev[row] = value
        else:
            self.host[row] = value
            if self.on_write is not None:
                self.on_write()

    # ---- bulk access
    def host_array(self):
        self._flush()
        if self.dev is not None:
            return self.dev.cpu().numpy()
        return np.asarray(self.host, dtype=self.dtype)

    def get_data(self, space):
        """Host space: a copy whose changes are written back at the next use of the column.
        Device space: the live device tensor."""
        self._flush()
        if space == "host":
            if self.dev is not None:
                arr = self.dev.cpu().numpy().copy()
            else:
                arr = np.array(self.host, dtype=self.dtype)
            self.checked_out = arr
            return arr
        if self.dev is None:
            raise RuntimeError(f"{self.db_class.name}.{self.name} is not on the device yet "
                               "(call advance() first)")
        return self.dev

    def set_data(self, value):
        value = np.asarray(value)
        if self.dev is not None:
            torch = _torch()
            self.dev.copy_(torch.as_tensor(value.astype(self.dev.cpu().numpy().dtype), device=self.dev.device))
        else:
            assert len(value) == len(self.host)
            self.host = list(value)

    def _flush(self):
        arr, self.checked_out = self.checked_out, None
        if arr is None:
            return
        if self.dev is not None:
            torch = _torch()
            src = torch.as_tensor(np.ascontiguousarray(arr), device=self.dev.device)
            self.dev.copy_(src.to(self.dev.dtype))
        else:
            self.host = list(arr)

    def to_device(self, order, device):
        if not self.device:
            return
        torch = _torch()
        arr = np.asarray(self.host, dtype=self.dtype)[order] if len(self.host) else \
            np.zeros((0,) + self.shape, dtype=self.dtype)
        self.dev = torch.as_tensor(np.ascontiguousarray(arr.astype(self.device_dtype)), device=device)

    def to_host(self, inverse_order):
        if self.dev is not None:
            arr = self.dev.cpu().numpy()
            self.host = list(arr[inverse_order].astype(self.dtype))
            self.dev = None
            self.checked_out = None

    def check(self):
        arr = self.host_array().astype(np.float64)
        if arr.size and np.isnan(arr).any():
            raise AssertionError(f"{self.db_class.name}.{self.name} contains NaN")
        if self.valid_range is not None and arr.size:
            lo, hi = self.valid_range
            if lo is not None and np.any(arr < lo):
                raise AssertionError(f"{self.db_class.name}.{self.name} is less than {lo}")
            if hi is not None and np.any(arr > hi):
                raise AssertionError(f"{self.db_class.name}.{self.name} is greater than {hi}")


class DBClass:
    """A class of objects with structure-of-arrays attributes (database.py:261-655 subset)."""

    def __init__(self, database, name):
        self.database, self.name = database, name
        self.columns = {}
        self.class_attributes = {}
        self.size = 0
        self._row_of = np.zeros(16, dtype=np.int64)  # creation id -> row, grown by doubling
        database.classes[name] = self

    @property
    def row_of(self):
        """Row of every object by creation id (the identity until the model is built)."""
        return self._row_of[:self.size]

    @row_of.setter
    def row_of(self, rows):
        self._row_of = np.asarray(rows, dtype=np.int64)

    def get_name(self):
        return self.name

    def __len__(self):
        return self.size

    def add_attribute(self, name, initial_value=None, **kw):
        col = Column(self, name, initial_value, **kw)
        for _ in range(self.size):
            col.append()
        self.columns[name] = col
        return col

    def add_class_attribute(self, name, initial_value, units=None, valid_range=None):
        self.class_attributes[name] = initial_value
        return initial_value

    def get(self, name):
        return self.columns[name]

    def get_data(self, name):
        return self.columns[name].get_data(self.database.memory_space)

    def set_data(self, name, value):
        self.columns[name].set_data(value)

    def get_all_components(self):
        return list(self.columns.values())

    def new_row(self):
        for col in self.columns.values():
            col.append()
        cid = self.size
        if cid >= len(self._row_of):  # amortised O(1): objects are created one at a time
            grown = np.empty(max(16, 2 * len(self._row_of)), dtype=np.int64)
            grown[:cid] = self._row_of[:cid]
            self._row_of = grown
        self._row_of[cid] = cid
        self.size += 1
        return cid


class Database:
    """get_data / set_data / using_memory_space / check (database.py:11-259 subset)."""

    def __init__(self):
        self.classes = {}
        self.memory_space = "host"
        self.clock = None

    def get(self, name):
        cls, _, attr = str(name).partition(".")
        c = self.classes[cls]
        return c.get(attr) if attr else c

    def get_class(self, name):
        return self.classes[str(name).partition(".")[0]]

    def get_data(self, name):
        cls, _, attr = str(name).partition(".")
        if self.memory_space == "device" and getattr(self, "_model", None) is not None:
            self._model._ensure_built()
        return self.classes[cls].get_data(attr)

    def set_data(self, name, value):
        cls, _, attr = str(name).partition(".")
        self.classes[cls].set_data(attr, value)

    def get_memory_space(self):
        return self.memory_space

    @contextlib.contextmanager
    def using_memory_space(self, memory_space):
        space = {"cuda": "device", "hip": "device", "device": "device", "host": "host"}[memory_space]
        prior = self.memory_space
        self.memory_space = space
        try:
            yield
        finally:
            self.memory_space = prior

    def add_clock(self, clock):
        assert self.clock is None
        self.clock = clock
        return clock

    def get_clock(self):
        return self.clock

    def check(self, name=None):
        errors = []
        classes = [self.get_class(name)] if name else list(self.classes.values())
        for c in classes:
            for col in c.columns.values():
                try:
                    col.check()
                except AssertionError as e:
                    errors.append(str(e))
        if errors:
            raise AssertionError(",\n".join(sorted(errors)) + ".")


class DBObject:
    """Handle to one row of a DBClass; attribute access reads/writes its columns."""
    __slots__ = ("_cid",)
    _db_class = None

    def __getattr__(self, name):
        cls = type(self)._db_class
        col = cls.columns.get(name)
        if col is None:
            if name in cls.class_attributes:
                return cls.class_attributes[name]
            raise AttributeError(name)
        cid = object.__getattribute__(self, "_cid")
        return col.get_item(int(cls.row_of[cid]) if col.dev is not None else cid)

    def __setattr__(self, name, value):
        if name in DBObject.__slots__:
            object.__setattr__(self, name, value)
            return
        cls = type(self)._db_class
        col = cls.columns.get(name)
        if col is not None:
            col.set_item(int(cls.row_of[self._cid]) if col.dev is not None else self._cid, value)
        else:
            object.__setattr__(self, name, value)

    def get_unstable_index(self):
        return int(type(self)._db_class.row_of[self._cid])

    def __eq__(self, other):
        return type(self) is type(other) and self._cid == other._cid

    def __hash__(self):
        return hash((type(self).__name__, self._cid))


# ================================================================================================
# Mechanisms (neuwon/rxd/mechanisms.py:4-54)
# ================================================================================================
class Mechanism:
    """Plugin ABI of the reference (mechanisms.py:4-54)."""
    omnipresent = False

    @classmethod
    def initialize(cls, model):
        pass

    @classmethod
    def get_name(cls):
        raise NotImplementedError(cls)

    @classmethod
    def other_mechanisms(cls):
        return []

    def __init__(self, segment, outside, magnitude, *other_mechanisms):
        raise NotImplementedError(type(self))

    @classmethod
    def advance(cls):
        raise NotImplementedError(cls)


class _DeviceMechanism(DBObject):
    """Built-in mechanism instance: one row of its class; instance i sits on segment[i]."""
    __slots__ = ()
    _model = None
    omnipresent = False

    @classmethod
    def get_name(cls):
        return cls._name

    @classmethod
    def other_mechanisms(cls):
        return []

    @classmethod
    def get_database_class(cls):
        return cls._db_class


def _make_hh_mechanism(model, name, gate_mode, gbar=None, reversal=None):
    """hh.mod (nmodl_library/artificial/hh.mod): leak + Na (m^3 h) + K (n^4), the ions written as
    nonspecific conductances (USEION ... WRITE gna, ena; nmodl/__init__.py:168-200).  gate_mode
    "exact" is the DERIVATIVE block solved exactly (solver.py:48-55); "table" the same gates as
    two-state kinetic schemes advanced by LTI tables (lti_sim/codegen.py:39-106)."""
    db = model.database
    cls_db = DBClass(db, name)
    gbar = dict(channels.HH_GBAR, **(gbar or {}))
    E = dict(channels.HH_REVERSAL, **(reversal or {}))
    ss = channels.hh_steady_state(model.initial_voltage, model.temperature)
    cls_db.add_attribute("segment", NULL, dtype=np.int64, device=False,
                         on_write=model._invalidate_mechanisms)
    # magnitude lives on the device: a change (another mechanism's pointer, set_data) rescales the
    # instance's maximal conductances at the next advance without rebuilding the model
    cls_db.add_attribute("magnitude", 1.0, valid_range=(0.0, None))
    for g in "mhn":
        cls_db.add_attribute(g, ss[g])
    for ion in ("na", "k", "l"):
        cls_db.add_class_attribute(f"{ion}_reversal_potential", E[ion])

    def __init__(self, segment, outside=None, magnitude=1.0, *others):
        cid = type(self)._db_class.new_row()
        object.__setattr__(self, "_cid", cid)
        model._invalidate_mechanisms()
        self.segment = segment._cid
        self.magnitude = float(magnitude) * segment.surface_area   # nmodl/__init__.py:499

    mech = type(name, (_DeviceMechanism,), {"__slots__": (), "__init__": __init__, "_name": name,
                                           "_db_class": cls_db, "_model": model,
                                           "gate_mode": gate_mode, "gbar": gbar, "E": E})

    def advance(cls):
        model._run_hh_mechanism(cls)
    mech.advance = classmethod(advance)
    mech.conductance_sources = ("na", "k", "l")
    return mech


class LTIMechanism:
    """Factory of a kinetic mechanism from an ``channels.LTIScheme`` whose conductance is
    gmax * magnitude * sum_s weights[s] state_s with reversal potential ``reversal`` (mV).  For a
    density mechanism gmax is in S/cm^2 and the magnitude is scaled by the segment's surface area
    (nmodl/__init__.py:97-107, 499); for a point process (point_process=True) gmax * weights are
    Siemens and the magnitude is the inserted one (nmodl/__init__.py:494-497).  Inputs of the
    scheme other than the voltage ('v') are attributes of each instance, initially the input's
    minimum (lti_sim/inputs.py:10-21)."""

    def __init__(self, scheme, gmax_s_per_cm2, reversal, name=None, point_process=False):
        self.scheme = scheme
        self.point_process = bool(point_process)
        self.gmax = float(gmax_s_per_cm2) * (1.0 if self.point_process else channels.CM2_TO_UM2)
        self.reversal = float(reversal)
        self.name = name or scheme.name
        self.input_names = list(getattr(scheme, "input_names", ["v"]))
        self.inputs = list(getattr(scheme, "inputs", [scheme.grid]))

    def initialize(self, model):
        db = model.database
        cls_db = DBClass(db, self.name)
        cls_db.add_attribute("segment", NULL, dtype=np.int64, device=False,
                             on_write=model._invalidate_mechanisms)
        cls_db.add_attribute("magnitude", 1.0, valid_range=(0.0, None))
        start = [model.initial_voltage if n == "v" else g.minimum
                 for n, g in zip(self.input_names, self.inputs)]
        for n, x0 in zip(self.input_names, start):
            if n != "v":
                cls_db.add_attribute(n, float(x0))
        init = self.scheme.initial_state(*start)
        for s, x0 in zip(self.scheme.state_names, init):
            cls_db.add_attribute(s, float(x0))
        spec = self
        point = self.point_process

        def __init__(self, segment, outside=None, magnitude=1.0, *others):
            cid = type(self)._db_class.new_row()
            object.__setattr__(self, "_cid", cid)
            model._invalidate_mechanisms()
            self.segment = segment._cid
            self.magnitude = float(magnitude) * (1.0 if point else segment.surface_area)

        mech = type(self.name, (_DeviceMechanism,), {"__slots__": (), "__init__": __init__,
                                                    "_name": self.name, "_db_class": cls_db,
                                                    "_model": model, "spec": spec})
        mech.advance = classmethod(lambda cls: model._run_lti_mechanism(cls))
        mech.conductance_sources = ("g",)
        return mech


_BUILTIN = {
    "hh": lambda model: _make_hh_mechanism(model, "hh", "exact"),
    "hh_kinetic": lambda model: _make_hh_mechanism(model, "hh_kinetic", "table"),
}


# ================================================================================================
# Neuron and Segment (neuwon/rxd/neuron/{neuron,segment,geometry,electric}.py)
# ================================================================================================
class _Segment(DBObject):
    __slots__ = ()
    _model = None

    def __init__(self, parent, coordinates, diameter, *, segment_type=None):
        model = type(self)._model
        model._invalidate_structure()
        cid = type(self)._db_class.new_row()
        object.__setattr__(self, "_cid", cid)
        self.parent = NULL if parent is None else parent._cid
        self.coordinates = tuple(float(x) for x in coordinates)
        self.diameter = float(diameter)
        self.neuron = NULL if parent is None else parent.neuron
        model._children.append([])
        if parent is not None:
            model._children[parent._cid].append(cid)

    @classmethod
    def get_database_class(cls):
        return cls._db_class

    def is_root(self):
        return self.parent == NULL

    is_sphere = is_root

    def is_cylinder(self):
        return not self.is_root()

    @property
    def children(self):
        return [type(self)._from_cid(c) for c in type(self)._model._children[self._cid]]

    @classmethod
    def _from_cid(cls, cid):
        obj = object.__new__(cls)
        object.__setattr__(obj, "_cid", int(cid))
        return obj

    def get_parent(self):
        return None if self.parent == NULL else type(self)._from_cid(self.parent)

    def add_segment(self, coordinates, diameter, segment_type=None):
        return type(self)(self, coordinates, diameter, segment_type=segment_type)

    def add_section(self, coordinates, diameter, maximum_segment_length=np.inf, segment_type=None):
        """segment.py:75-104."""
        maximum_segment_length = float(maximum_segment_length)
        assert maximum_segment_length > 0
        coords = [float(x) for x in coordinates]
        diameter = float(diameter)
        displace = np.subtract(self.coordinates, coords)
        length = float(np.linalg.norm(displace))
        p_coords = np.array(self.coordinates, dtype=float)
        if self.is_sphere() or len(type(self)._model._children[self._cid]) > 0:
            self_r = 0.5 * self.diameter
            if self_r < length - EPSILON:
                p_coords -= displace * (self_r / length)
                length -= self_r
        divisions = max(1, int(np.ceil(length / maximum_segment_length)))
        x = np.linspace(p_coords[0], coords[0], divisions + 1)
        y = np.linspace(p_coords[1], coords[1], divisions + 1)
        z = np.linspace(p_coords[2], coords[2], divisions + 1)
        d = np.linspace(self.diameter, diameter, divisions + 1)
        seg, section = self, []
        for k in range(1, divisions + 1):
            seg = seg.add_segment((x[k], y[k], z[k]), d[k], segment_type=segment_type)
            section.append(seg)
        return section

    def insert(self, mechanisms, outside=None):
        """segment.py:124-160."""
        from collections.abc import Mapping, Iterable
        if isinstance(mechanisms, Mapping):
            pass
        elif isinstance(mechanisms, str):
            mechanisms = {mechanisms: 1.0}
        elif isinstance(mechanisms, Iterable):
            mechanisms = {name: 1.0 for name in mechanisms}
        else:
            raise ValueError(f'Expected dictionary, not "{type(mechanisms)}"')
        model = type(self)._model
        all_mech = model.mechanisms
        instances = {}

        def insert_recursive(name):
            # a mechanism's other_mechanisms() are created first and handed to its constructor;
            # each name is instantiated once per call (segment.py:141-160)
            if name in instances:
                return instances[name]
            others = [insert_recursive(x) for x in model._local_dependencies.get(name, ())]
            inst = all_mech[name](self, outside, float(mechanisms[name]), *others)
            instances[name] = inst
            return inst

        for name in mechanisms:
            insert_recursive(name)
        return instances

    def inject_current(self, current=1.0, duration=1.0):
        """electric.py:119-125: current (nA) for duration (ms), from the next input tick."""
        current = float(current) * 1e-9
        duration = float(duration)
        assert duration >= 0
        TimeSeries().constant_wave(current, duration).play(
            self, "sum_current", clock=type(self)._model.input_hook, immediate=False)

    # derived geometry (computed when the model is built)
    def _geo(self, key):
        model = type(self)._model
        model._ensure_geometry()
        return float(model._geo[key][self._cid])

    @property
    def length(self):
        return self._geo("length")

    @property
    def surface_area(self):
        return self._geo("surface_area")

    @property
    def cross_sectional_area(self):
        return self._geo("cross_sectional_area")

    @property
    def volume(self):
        return self._geo("volume")

    @property
    def capacitance(self):
        return self._geo("capacitance")

    @property
    def axial_resistance(self):
        return self._geo("axial_resistance")

    def get_time_constant(self):
        return self.capacitance / self.sum_conductance

    def get_length_constant(self):
        rm = 1.0 / (self.sum_conductance / self.length)
        ri = self.axial_resistance / self.length
        return (rm / ri) ** 0.5

    @classmethod
    def _advance_electric(cls, time_step):
        """One electric step of ``time_step`` ms with the current arrays (electric.py:57-76)."""
        cls._model._electric_step(float(time_step))


class _Neuron(DBObject):
    __slots__ = ()
    _model = None

    def __init__(self, coordinates, diameter, *, neuron_type=None, segment_type=None):
        model = type(self)._model
        cid = type(self)._db_class.new_row()
        object.__setattr__(self, "_cid", cid)
        root = model.Segment(None, coordinates, diameter, segment_type=segment_type)
        root.neuron = cid
        self.root = root._cid
        self.AP_detector_segment = root._cid
        self.neuron_type = neuron_type

    def __getattr__(self, name):
        if name == "root":
            cid = DBObject.__getattr__(self, "root")
            return type(self)._model.Segment._from_cid(int(cid))
        return DBObject.__getattr__(self, name)

    def __setattr__(self, name, value):
        if name == "neuron_type":
            type(self)._model._neuron_types[self._cid] = value
            return
        DBObject.__setattr__(self, name, value)

    @property
    def neuron_type(self):
        return type(self)._model._neuron_types.get(self._cid)

    def set_AP_detector(self, segment=None, threshold=None):
        if segment is not None:
            self.AP_detector_segment = segment._cid
        if threshold is not None:
            type(self)._db_class.class_attributes["AP_detector_threshold"] = float(threshold)


# ================================================================================================
# The model
# ================================================================================================
class RxD_Model:
    """rxd_model.py:8-123 on the MI355X (see module docstring)."""

    def __init__(self, time_step=0.1, *, temperature=37, initial_voltage=-70,
                 cytoplasmic_resistance=100, membrane_capacitance=1, extracellular_tortuosity=1.55,
                 extracellular_max_distance=20e-6, species=(), mechanisms=(), device=0,
                 prune=EPSILON, symmetric_propagator=False, step_graph=True,
                 propagator_store=None, kinetic_table_accuracy=None):
        self.time_step = float(time_step)
        self.temperature = float(temperature)
        self.initial_voltage = float(initial_voltage)
        self.cytoplasmic_resistance = float(cytoplasmic_resistance)
        self.membrane_capacitance = float(membrane_capacitance)
        self.prune = float(prune)
        # symmetric_propagator: store sqrt(C) P sqrt(C)^-1, one triangle (half the bytes per half
        # step; pairs pruned together; sums in atomic order, so not bitwise reproducible)
        self.symmetric_propagator = bool(symmetric_propagator)
        # propagator_store: "general" | "symmetric" | "chebyshev" | "rational" (overrides
        # symmetric_propagator), or "auto": both matrix-free stores are built and timed on this
        # model's half step, the faster is kept (``tuning``).  Unset: the Chebyshev store
        # (exp(A dt) without a stored matrix) when nothing is pruned, the reference's default
        # threshold, and every cell fits it; otherwise the general store
        from .population import store_name
        self.tuning = {}
        if propagator_store == "auto":
            self.propagator_store = "auto"
        else:
            self.propagator_store = (None if propagator_store is None and not self.symmetric_propagator
                                     else store_name(propagator_store, self.symmetric_propagator))
        # step_graph: when every callback of a step runs on the device (device mechanisms only,
        # play back / recording / traces of device attributes), advance() replays a captured
        # graph of the step's kernels instead of launching them one by one from Python
        self.step_graph = bool(step_graph)
        # kinetic_table_accuracy: the max abs propagator error the tables of kinetic .mod files
        # are sized for (the reference sizes them for 1e-4, nmodl/__init__.py:304); unset: the
        # fixed grids of nmodl.input_grid
        self.kinetic_table_accuracy = (None if kinetic_table_accuracy is None
                                       else float(kinetic_table_accuracy))
        self.device = int(device)
        self.database = db = Database()
        db._model = self
        self.input_hook = Clock(0.5 * self.time_step, units="ms")
        self.advance_hook = self.clock = Clock(self.time_step, units="ms")
        db.add_clock(self.clock)
        for c in (self.input_hook, self.clock):
            c._events = events.DeviceEvents(self, c)
        self._graph = self._graph_key = self._graph_seen = None
        self._build_count = 0
        self._children = []
        self._neuron_types = {}
        self._built = False
        self._geo = None
        self._propagators = {}
        # Segment class and attributes (segment.py:11-25, geometry.py:42-73, electric.py:10-43)
        seg = DBClass(db, "Segment")
        seg.add_attribute("parent", NULL, dtype=np.int64, device=False)
        seg.add_attribute("neuron", NULL, dtype=np.int64, device=False)
        seg.add_attribute("coordinates", (0.0, 0.0, 0.0), dtype=np.float64, shape=(3,),
                          device=False, on_write=self._invalidate_structure)
        seg.add_attribute("diameter", 1.0, valid_range=(EPSILON, None), device=False,
                          on_write=self._invalidate_structure)
        seg.add_attribute("voltage", self.initial_voltage)
        seg.add_attribute("integral_voltage", 0.0)
        seg.add_attribute("nonspecific_current", 0.0)
        seg.add_attribute("sum_current", 0.0)
        seg.add_attribute("sum_conductance", 0.0, valid_range=(0.0, None))
        seg.add_attribute("driving_voltage", 0.0)
        seg.add_class_attribute("cytoplasmic_resistance", self.cytoplasmic_resistance)
        seg.add_class_attribute("membrane_capacitance", self.membrane_capacitance)
        self.Segment = type("Segment", (_Segment,), {"__slots__": (), "_db_class": seg, "_model": self})
        neu = DBClass(db, "Neuron")
        neu.add_attribute("root", NULL, dtype=np.int64, device=False)
        neu.add_attribute("AP_detector_segment", NULL, dtype=np.int64, device=False,
                          on_write=self._invalidate_references)
        neu.add_attribute("AP_detected", 0, dtype=np.uint8, device_dtype="uint8")
        neu.add_attribute("_AP_true_state", 0, dtype=np.uint8, device_dtype="uint8")
        neu.add_class_attribute("AP_detector_threshold", 20.0)
        self.Neuron = type("Neuron", (_Neuron,), {"__slots__": (), "_db_class": neu, "_model": self})
        # species: electric species with constant reversal potentials (species.py:85-142)
        self.species = {}
        for kw in species:
            name = kw["name"]
            rev = kw.get("reversal_potential")
            if rev is not None:
                seg.add_attribute(f"{name}_conductance", 0.0, valid_range=(0.0, None))
                seg.add_class_attribute(f"{name}_reversal_potential", float(rev))
            self.species[name] = dict(kw)
        # mechanisms (mechanisms.py:56-99)
        self.mechanisms = {}
        self._local_dependencies = {}
        for m in mechanisms:
            self._load_mechanism(m)

    # ------------------------------------------------------------------ accessors (rxd_model.py:37-45)
    def get_temperature(self):
        return self.temperature

    def get_clock(self):
        return self.clock

    def get_database(self):
        return self.database

    def get_mechanisms(self):
        return dict(self.mechanisms)

    def get_Neuron(self):
        return self.Neuron

    def get_Segment(self):
        return self.Segment

    def get_species(self):
        return dict(self.species)

    def get_time_step(self):
        return self.time_step

    def _events_for(self, clock):
        """The device callback table of a clock (created for clocks other than the model's)."""
        if clock._events is None:
            clock._events = events.DeviceEvents(self, clock)
        return clock._events

    def register_input_callback(self, function):
        self.input_hook.register_callback(function)

    def register_advance_callback(self, function):
        self.advance_hook.register_callback(function)

    def check(self):
        self._ensure_built()
        self.database.check()

    # ------------------------------------------------------------------ mechanisms
    def _load_mechanism(self, m):
        if isinstance(m, str):
            key = os.path.splitext(os.path.basename(m))[0]
            if key in _BUILTIN:
                mech = _BUILTIN[key](self)
            elif m.endswith(".mod") and os.path.exists(m):
                # a kinetic scheme file (nmodl_library/kinetic/*.mod): parsed, tabulated on its
                # input grids, advanced by nw_lti_advance / nw_lti_advance_nd (mechanisms.py:63-66)
                from . import nmodl
                rev = {name: kw["reversal_potential"] for name, kw in self.species.items()
                       if kw.get("reversal_potential") is not None}
                mech = nmodl.kinetic_mechanism(m, self.time_step, self.temperature, rev,
                                               accuracy=self.kinetic_table_accuracy).initialize(self)
            else:
                raise ValueError(f"unsupported mechanism {m!r} (built-ins: {sorted(_BUILTIN)}; "
                                 "or the path of a kinetic .mod file)")
        elif isinstance(m, LTIMechanism):
            mech = m.initialize(self)
        else:
            retval = m.initialize(self)
            mech = retval if retval is not None else m
        name = str(mech.get_name())
        assert name not in self.mechanisms
        if not getattr(mech, "omnipresent", False):   # mechanisms.py:84-96
            deps = mech.other_mechanisms()
            self._local_dependencies[name] = ((deps,) if isinstance(deps, str)
                                              else tuple(str(x) for x in deps))
        self.mechanisms[name] = mech

    # ------------------------------------------------------------------ structure changes
    def _invalidate_structure(self):
        if self._built:
            self._to_host()
        self._geo = None
        self._propagators = {}

    def _invalidate_mechanisms(self):
        if self._built:
            self._to_host()

    def _invalidate_references(self):
        self._refs_dirty = True

    def _to_host(self):
        for c in self.database.classes.values():
            order = getattr(c, "_order", None)
            if order is not None:
                inv = np.empty(len(order), dtype=np.int64)
                inv[order] = np.arange(len(order))
                for col in c.columns.values():
                    col.to_host(inv)
            c.row_of = np.arange(c.size)
            c._order = None
        self._built = False

    # ------------------------------------------------------------------ build
    def _ensure_geometry(self):
        if self._geo is not None:
            return
        seg = self.database.get_class("Segment")
        parent = np.asarray(seg.columns["parent"].host, dtype=np.int64)
        coords = np.asarray(seg.columns["coordinates"].host, dtype=np.float64).reshape(-1, 3)
        diam = np.asarray(seg.columns["diameter"].host, dtype=np.float64)
        geo = geometry.geometry(parent, coords, diam) if len(parent) else \
            {k: np.zeros(0) for k in ("length", "surface_area", "area_at_creation",
                                      "cross_sectional_area", "volume")}
        R, C = geometry.passive(geo["length"], geo["cross_sectional_area"], geo["area_at_creation"],
                                self.cytoplasmic_resistance, self.membrane_capacitance)
        geo["axial_resistance"], geo["capacitance"] = R, C
        self._geo = geo  # indexed by creation id

    def _ensure_built(self):
        if self._built:
            return
        torch = _torch()
        dev = torch.device("cuda", self.device)
        self._dev = dev
        self._ensure_geometry()
        db = self.database
        seg = db.get_class("Segment")
        N = seg.size
        parent = np.asarray(seg.columns["parent"].host, dtype=np.int64).reshape(-1)
        # cells: connected trees, ordered by the creation of their root (stable)
        root = np.arange(N)
        for i in range(N):
            if parent[i] != NULL:
                root[i] = root[parent[i]]
        order = np.argsort(root, kind="stable")
        row_of = np.empty(N, dtype=np.int64)
        row_of[order] = np.arange(N)
        sroot = root[order]
        starts = np.nonzero(np.r_[True, sroot[1:] != sroot[:-1]])[0] if N else np.zeros(0, int)
        cell_ptr = np.r_[starts, N].astype(np.int64)
        parent_local = np.where(parent[order] < 0, -1, row_of[np.maximum(parent[order], 0)] -
                                np.repeat(cell_ptr[:-1], np.diff(cell_ptr)))
        self._cell_ptr = cell_ptr
        self._parent_local = parent_local.astype(np.int32)
        self._R = self._geo["axial_resistance"][order]
        self._C = self._geo["capacitance"][order]
        seg._order, seg.row_of = order, row_of
        for col in seg.columns.values():
            col.to_device(order, dev)
        self._inv_c = torch.as_tensor((1.0 / self._C).astype(np.float32), device=dev)
        # neurons keep creation order; their AP segments are device rows
        neu = db.get_class("Neuron")
        neu._order = np.arange(neu.size)
        for col in neu.columns.values():
            col.to_device(neu._order, dev)
        self._upload_references()
        # mechanisms: instances sorted by the row of their segment
        self._sources = []
        gbuf_len = 0
        for mech in self.mechanisms.values():
            if not isinstance(mech, type) or not issubclass(mech, _DeviceMechanism):
                continue
            mc = mech._db_class
            segs = np.asarray(mc.columns["segment"].host, dtype=np.int64)
            rows = row_of[segs] if len(segs) else segs
            morder = np.argsort(rows, kind="stable")
            mc._order = morder
            mc.row_of = np.empty(mc.size, dtype=np.int64)
            mc.row_of[morder] = np.arange(mc.size)
            for col in mc.columns.values():
                col.to_device(morder, dev)
            mech._seg_dev = torch.as_tensor(rows[morder].astype(np.int32), device=dev)
            k = len(mech.conductance_sources)
            mech._g_offset = gbuf_len
            for j, src in enumerate(mech.conductance_sources):
                E = mech.E[src] if src in ("na", "k", "l") else mech.spec.reversal
                self._sources.append((rows[morder], gbuf_len + j + k * np.arange(mc.size), E))
            gbuf_len += k * mc.size
            self._prepare_mechanism(mech, dev)
        # species conductances: one source per segment each
        self._species_cols = []
        for name, kw in self.species.items():
            if kw.get("reversal_potential") is None:
                continue
            self._species_cols.append((name, gbuf_len))
            self._sources.append((np.arange(N), gbuf_len + np.arange(N),
                                  float(kw["reversal_potential"])))
            gbuf_len += N
        self._gbuf = torch.zeros(max(gbuf_len, 1), dtype=torch.float32, device=dev)
        if self._sources:
            srow = np.concatenate([s[0] for s in self._sources])
            sidx = np.concatenate([s[1] for s in self._sources])
            sE = np.concatenate([np.full(len(s[0]), s[2]) for s in self._sources])
        else:
            srow = sidx = np.zeros(0, dtype=np.int64)
            sE = np.zeros(0)
        so = np.argsort(srow, kind="stable")
        ptr = np.zeros(N + 1, dtype=np.int64)
        np.cumsum(np.bincount(srow, minlength=N), out=ptr[1:])
        self._acc_ptr = torch.as_tensor(ptr, device=dev)
        self._acc_idx = torch.as_tensor(sidx[so].astype(np.int64), device=dev)
        self._acc_E = torch.as_tensor(sE[so].astype(np.float32), device=dev)
        self._built = True
        self._build_count += 1

    def _upload_references(self):
        torch = _torch()
        neu = self.database.get_class("Neuron")
        seg = self.database.get_class("Segment")
        ap = np.asarray(neu.columns["AP_detector_segment"].host, dtype=np.int64)
        rows = np.where(ap < 0, -1, seg.row_of[np.maximum(ap, 0)]) if len(ap) else ap
        self._ap_seg = torch.as_tensor(rows.astype(np.int32), device=self._dev_or(torch))
        self._refs_dirty = False

    def _dev_or(self, torch):
        return getattr(self, "_dev", None) or torch.device("cuda", self.device)

    def _scaled_by_magnitude(self, mech):
        """Per-instance maximal conductances (magnitude x density), recomputed when the device
        magnitude column has been written since the last call."""
        col = mech._db_class.columns["magnitude"]
        col._flush()
        version = col.dev._version
        if getattr(mech, "_mag_version", None) != version:
            # written into the existing arrays, so a captured step graph keeps reading them
            mag = col.dev
            if hasattr(mech, "gate_mode"):
                new = {ion: (mag * (mech.gbar[ion] * channels.CM2_TO_UM2)).contiguous()
                       for ion in ("na", "k", "l")}
                old = getattr(mech, "_gbar_dev", None)
                if old is not None and all(old[i].shape == new[i].shape for i in new):
                    for i in new:
                        old[i].copy_(new[i])
                else:
                    mech._gbar_dev = new
            else:
                new = (mag * mech.spec.gmax).contiguous()
                old = getattr(mech, "_gmax_dev", None)
                if old is not None and old.shape == new.shape:
                    old.copy_(new)
                else:
                    mech._gmax_dev = new
            mech._mag_version = version

    def _prepare_mechanism(self, mech, dev):
        torch = _torch()
        mc = mech._db_class
        mech._mag_version = None
        if hasattr(mech, "gate_mode"):
            grid = channels.LinearInput(-100.0, 100.0, 2000)
            mech._grid = grid
            mech._table = torch.as_tensor(
                channels.hh_gate_table(self.time_step, self.temperature, grid).astype(np.float32),
                device=dev)
        else:
            spec = mech.spec
            spec.scheme.time_step = self.time_step
            mech._node_table = torch.as_tensor(
                np.ascontiguousarray(spec.scheme.node_table(), dtype=np.float32), device=dev)
            mech._weights = torch.as_tensor(np.asarray(spec.scheme.weights, dtype=np.float32), device=dev)
            S = spec.scheme.num_states
            mech._states = torch.stack([mc.columns[s].dev for s in spec.scheme.state_names]).contiguous()
            for k, s in enumerate(spec.scheme.state_names):
                mc.columns[s].dev = mech._states[k]
            # inputs other than the voltage alone: the general table kernel, inputs described once
            # per build (the arrays do not move until the next build)
            mech._nd_inputs = None
            mech._lti_table = None
            if spec.input_names == ["v"]:
                # the voltage alone: the bucket-grouped kernel (instances are in segment order, so
                # neighbouring instances share voltage buckets)
                from .kinetic import LTITable
                g = spec.scheme.grid
                mech._lti_table = LTITable(mech._node_table, g.minimum, g.maximum,
                                           device=dev.index or 0)
            if spec.input_names != ["v"]:
                arr = (_lib.NwLtiInput * len(spec.inputs))()
                for k, (n, g) in enumerate(zip(spec.input_names, spec.inputs)):
                    if n == "v":
                        arr[k].values = self._col("Segment", "voltage").data_ptr()
                        arr[k].index = mech._seg_dev.data_ptr()
                    else:
                        arr[k].values = mc.columns[n].dev.data_ptr()
                        arr[k].index = None
                    log = isinstance(g, channels.LogarithmicInput)
                    arr[k].mapping = _lib.NW_INPUT_LOG2 if log else _lib.NW_INPUT_LINEAR
                    arr[k].num_nodes = g.num_buckets + 1
                    arr[k].minimum, arr[k].maximum = g.minimum, g.maximum
                    arr[k].scale = g.scale if log else 0.0
                mech._nd_inputs = arr

    # ------------------------------------------------------------------ kernels
    def _stream(self):
        return _torch().cuda.current_stream(self._dev).cuda_stream

    def _col(self, cls, name):
        return self.database.get_class(cls).columns[name].dev

    def _run_hh_mechanism(self, mech):
        if mech._db_class.size:
            self._scaled_by_magnitude(mech)
            self._hh_kernel(mech)

    def _hh_kernel(self, mech):
        mc = mech._db_class
        M = mc.size
        if M == 0:
            return
        P = _lib.ptr
        out = self._gbuf[mech._g_offset:mech._g_offset + 3 * M]
        g = mech._grid
        _lib.check(_lib.load().nw_hh_mechanism(
            M, 1 if mech.gate_mode == "exact" else 0, self.time_step,
            channels.q10_factor(self.temperature), P(mech._table), g.num_buckets + 1, g.minimum,
            g.maximum, P(self._col("Segment", "voltage")), P(mech._seg_dev),
            P(mc.columns["m"].dev), P(mc.columns["h"].dev), P(mc.columns["n"].dev),
            P(mech._gbar_dev["na"]), P(mech._gbar_dev["k"]), P(mech._gbar_dev["l"]), P(out),
            self._stream()))

    def _run_lti_mechanism(self, mech):
        if mech._db_class.size:
            self._scaled_by_magnitude(mech)
            self._lti_kernel(mech)

    def _lti_kernel(self, mech):
        mc = mech._db_class
        M = mc.size
        if M == 0:
            return
        spec = mech.spec
        P = _lib.ptr
        out = self._gbuf[mech._g_offset:mech._g_offset + M]
        if mech._nd_inputs is not None:
            _lib.check(_lib.load().nw_lti_advance_nd(
                spec.scheme.num_states, len(mech._nd_inputs), ctypes.addressof(mech._nd_inputs),
                P(mech._node_table), M, P(mech._states), M, spec.scheme.conserve_sum or 0.0, 1,
                P(mech._weights), P(mech._gmax_dev), P(out), self._stream()))
            return
        _lib.check(_lib.load().nw_lti_table_advance(
            mech._lti_table.handle, M, P(self._col("Segment", "voltage")), P(mech._seg_dev),
            P(mech._states), M, spec.scheme.conserve_sum or 0.0, 1, P(mech._weights),
            P(mech._gmax_dev), P(out), -1, self._stream()))

    def _propagator(self, time_step):
        from .population import Propagator
        key = round(time_step, 12)
        prop = self._propagators.get(key)
        if prop is None:
            store = self.propagator_store
            if store is None:
                fits = int(np.diff(self._cell_ptr).max(initial=0)) <= CHEBYSHEV_MAX_CELL
                store = "chebyshev" if self.prune <= EPSILON and fits else "general"
            if store == "auto":
                prop = self._tuned_propagator(time_step)
            else:
                prop = Propagator(self._cell_ptr, self._parent_local, self._R, self._C, time_step,
                                  self.prune, self.device, store=store)
            self._propagators = {key: prop}
        return prop

    def _tuned_propagator(self, time_step, warmup=2, steps=6):
        """The faster matrix-free store for this model's half step: each candidate runs the half
        step on a copy of the voltages (the model's state is not touched), timed with HIP events;
        a candidate that does not fit the cells is skipped."""
        import time as _time
        from .population import Propagator
        torch = _torch()
        t0 = _time.perf_counter()
        v = self._col("Segment", "voltage")
        scratch = v.clone()
        args = [self._col("Segment", n) for n in ("sum_conductance", "driving_voltage", "sum_current")]
        ms, skipped, built = {}, {}, {}
        P = _lib.ptr
        for store in ("chebyshev", "rational"):
            try:
                prop = Propagator(self._cell_ptr, self._parent_local, self._R, self._C, time_step,
                                  self.prune, self.device, store=store)
            except (RuntimeError, ValueError) as e:
                skipped[store] = str(e)
                continue

            def run(k):
                for _ in range(k):
                    _lib.check(_lib.load().nw_electric_half_step(
                        prop.handle, time_step, P(scratch), P(args[0]), P(args[1]), P(args[2]),
                        P(self._inv_c), None, self._stream()))
            scratch.copy_(v)
            run(warmup)
            stream = torch.cuda.current_stream(self._dev)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run(steps)
            e1.record(stream)
            torch.cuda.synchronize(self._dev)
            ms[store] = e0.elapsed_time(e1) / steps
            built[store] = prop
        if not ms:
            raise RuntimeError(f"no matrix-free store fits these cells: {skipped}")
        choice = min(ms, key=ms.get)
        self.tuning["propagator_store"] = {"choice": choice, "ms_per_half_step": ms,
                                           "skipped": skipped,
                                           "tuning_s": _time.perf_counter() - t0}
        return built[choice]

    def _electric_step(self, time_step):
        self._ensure_built()
        if self.database.get_class("Segment").size == 0:
            return
        for col in self.database.get_class("Segment").columns.values():
            col._flush()
        self._half_step_kernel(self._propagator(time_step), time_step)

    def _half_step_kernel(self, prop, time_step):
        P = _lib.ptr
        _lib.check(_lib.load().nw_electric_half_step(
            prop.handle, time_step, P(self._col("Segment", "voltage")),
            P(self._col("Segment", "sum_conductance")), P(self._col("Segment", "driving_voltage")),
            P(self._col("Segment", "sum_current")), P(self._inv_c),
            P(self._col("Segment", "integral_voltage")), self._stream()))

    # ------------------------------------------------------------------ state snapshots
    def save_state(self, path):
        """Write a snapshot: every device-resident attribute of every class (voltages,
        accumulators, mechanism states and magnitudes) in creation order, the structure it belongs
        to (classes, sizes, parents, coordinates, diameters, mechanism segments), both clocks and
        the mechanisms' last conductances.  TimeSeries that are playing or recording are not part
        of it (sql.py:78-128 keeps less: a subset of the columns)."""
        from . import checkpoint
        self._ensure_built()
        arrays, sizes = {}, {}
        for cname, c in self.database.classes.items():
            sizes[cname] = c.size
            cid_rows = np.asarray(c.row_of[np.arange(c.size)], dtype=np.int64)
            for col in c.columns.values():
                arr = col.host_array()
                if col.dev is not None:
                    arr = arr[cid_rows]
                arrays[f"{'state' if col.device else 'structure'}:{cname}.{col.name}"] = arr
        arrays["model:conductances"] = self._gbuf
        checkpoint.save(path, arrays, {"kind": "RxD_Model", "sizes": sizes,
                                       "ticks": self.clock.ticks,
                                       "input_ticks": self.input_hook.ticks,
                                       "time_step": self.time_step})

    def load_state(self, path):
        """Restore a snapshot written by ``save_state`` into a model of the same structure."""
        from . import checkpoint
        arrays, meta = checkpoint.load(path, "RxD_Model")
        if meta["time_step"] != self.time_step:
            raise ValueError(f"{path}: time step {meta['time_step']} ms, this model {self.time_step}")
        self._ensure_built()
        sizes = {n: c.size for n, c in self.database.classes.items()}
        if sizes != meta["sizes"]:
            raise ValueError(f"{path}: classes and sizes {meta['sizes']} differ from {sizes}")
        for key, arr in arrays.items():
            kind, _, name = key.partition(":")
            if kind == "structure":
                cname, attr = name.split(".", 1)
                mine = self.database.get_class(cname).columns[attr].host_array()
                if not np.array_equal(np.asarray(mine), arr):
                    raise ValueError(f"{path}: {name} differs from this model's")
        for key, arr in arrays.items():
            kind, _, name = key.partition(":")
            if kind != "state":
                continue
            cname, attr = name.split(".", 1)
            c = self.database.get_class(cname)
            col = c.columns[attr]
            col._flush()
            dev_order = np.empty_like(arr)
            dev_order[np.asarray(c.row_of[np.arange(c.size)], dtype=np.int64)] = arr
            col.set_data(dev_order)
        from .checkpoint import copy_into
        copy_into(self._gbuf, arrays["model:conductances"], "conductances")
        self.clock._set_ticks(int(meta["ticks"]))
        self.input_hook._set_ticks(int(meta["input_ticks"]))

    # ------------------------------------------------------------------ the step (rxd_model.py:63-123)
    def advance(self):
        """Advance the model by one time step: half step, mechanisms, half step, AP detector."""
        self._ensure_built()
        if self._refs_dirty:
            self._upload_references()
        if self.step_graph and self._graphable():
            self._advance_graphed()
            return
        self._advance_species()
        self._advance_mechanisms()
        self._advance_species()
        self._advance_ap_detector()
        self.advance_hook.tick()

    # ------------------------------------------------------------------ the step as one graph
    def _device_mechanisms(self):
        return [m for m in self.mechanisms.values()
                if isinstance(m, type) and issubclass(m, _DeviceMechanism)]

    def _graphable(self):
        """Every callback of the step runs on the device: no host callbacks on either clock and
        no Python plugin mechanisms."""
        if self.clock._callbacks or self.input_hook._callbacks:
            return False
        if len(self._device_mechanisms()) != len(self.mechanisms):
            return False
        return self.database.get_class("Segment").size > 0

    def _launch_step(self, prop):
        """The kernels of one step (rxd_model.py:63-123) with no host work between them."""
        half = 0.5 * self.time_step
        inputs, clock = self.input_hook._events, self.clock._events
        for _ in range(2):
            self._accumulate_kernel()
            inputs.launch_kernels(self._stream())
            self._half_step_kernel(prop, half)
            if _ == 0:
                self._zero_accumulators()
                for m in self.mechanisms.values():
                    if hasattr(m, "gate_mode"):
                        self._hh_kernel(m)
                    else:
                        self._lti_kernel(m)
                self._copy_species_conductances()
        self._ap_kernel()
        clock.launch_kernels(self._stream())

    def _advance_graphed(self):
        """One step as a replay of a captured graph of ``_launch_step``.  The first step of each
        configuration runs the same kernels eagerly; the next one captures them.  Host writes,
        magnitude changes and the clocks' tables are brought up to date before every replay."""
        torch = _torch()
        for c in self.database.classes.values():
            for col in c.columns.values():
                col._flush()
        for m in self._device_mechanisms():
            if m._db_class.size:
                self._scaled_by_magnitude(m)
        self.input_hook._events.prepare(2)
        self.clock._events.prepare(1)
        prop = self._propagator(0.5 * self.time_step)
        key = (self._build_count, id(prop), self._gbuf.data_ptr(), self._ap_seg.data_ptr(),
               float(self.database.get_class("Neuron").class_attributes["AP_detector_threshold"]),
               tuple(id(m) for m in self.mechanisms.values()),
               self.input_hook._events.graph_key(), self.clock._events.graph_key())
        if key != self._graph_key:
            self._graph = None
            if self._graph_seen != key:
                self._graph_seen = key
                self._launch_step(prop)
                self.input_hook._count(2)
                self.clock._count(1)
                return
            g = torch.cuda.CUDAGraph()
            # no cyclic garbage collection while the stream is captured: finalisers of collected
            # objects (a dropped model's propagator frees device memory, a graph or an event is
            # destroyed) make HIP calls that are not allowed during a capture.  torch.cuda.graph
            # collects just before the capture begins.
            import gc
            was_enabled = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.device(self._dev), torch.cuda.graph(g):
                    self._launch_step(prop)
            finally:
                if was_enabled:
                    gc.enable()
            self._graph, self._graph_key = g, key
        self._graph.replay()
        self.input_hook._count(2)
        self.clock._count(1)

    # ------------------------------------------------------------------ the step, eagerly
    def _advance_species(self):
        """Accumulate conductances and currents, run input callbacks, one dt/2 electric step
        (rxd_model.py:91-113)."""
        seg = self.database.get_class("Segment")
        N = seg.size
        if N == 0:
            self.input_hook.tick()
            return
        for col in seg.columns.values():
            col._flush()
        self._accumulate_kernel()
        self.input_hook.tick()
        self._electric_step(0.5 * self.time_step)

    def _accumulate_kernel(self):
        seg = self.database.get_class("Segment")
        N = seg.size
        P = _lib.ptr
        _lib.check(_lib.load().nw_accumulate(
            N, P(self._acc_ptr), P(self._gbuf), P(self._acc_idx), P(self._acc_E),
            P(self._col("Segment", "nonspecific_current")), P(self._col("Segment", "sum_conductance")),
            P(self._col("Segment", "driving_voltage")), P(self._col("Segment", "sum_current")),
            self._stream()))

    def _advance_mechanisms(self):
        """Zero the accumulators and advance every mechanism (rxd_model.py:115-123)."""
        self._zero_accumulators()
        for name, m in self.mechanisms.items():
            try:
                m.advance()
            except Exception:
                print(f"ERROR in mechanism {name}.")
                raise
        self._copy_species_conductances()

    def _zero_accumulators(self):
        self._gbuf.zero_()
        for name, _ in self._species_cols:          # species._zero_accumulators (species.py:147-153)
            self._col("Segment", f"{name}_conductance").zero_()
        self._col("Segment", "nonspecific_current").zero_()

    def _copy_species_conductances(self):
        for name, off in self._species_cols:
            seg_col = self._col("Segment", f"{name}_conductance")
            self._gbuf[off:off + len(seg_col)].copy_(seg_col)

    def _advance_ap_detector(self):
        self._ap_kernel()

    def _ap_kernel(self):
        neu = self.database.get_class("Neuron")
        if neu.size == 0:
            return
        thr = neu.class_attributes["AP_detector_threshold"]
        P = _lib.ptr
        _lib.check(_lib.load().nw_ap_detect(
            neu.size, P(self._ap_seg), P(self._col("Segment", "voltage")),
            float(thr), P(self._col("Neuron", "_AP_true_state")), P(self._col("Neuron", "AP_detected")),
            self._stream()))


class Model(RxD_Model):
    """neuwon/model.py:6-22: ``Model(simulation={...}, species=..., mechanisms=...)``."""

    def __init__(self, simulation=None, *, species=(), mechanisms=(), regions=None, segments=None,
                 neurons=None, synapses=None):
        self.parameters = dict(simulation=simulation or {}, species=species, mechanisms=mechanisms,
                               regions=regions or {}, segments=segments or {},
                               neurons=neurons or {}, synapses=synapses or {})
        if regions or neurons or synapses:
            raise NotImplementedError("regions, neuron growth and synapses are outside the advance path")
        if isinstance(species, dict):
            species = [dict(v, name=k) for k, v in species.items()]
        if isinstance(mechanisms, dict):
            mechanisms = list(mechanisms.values())
        super().__init__(species=species, mechanisms=mechanisms, **(simulation or {}))

    def get_parameters(self):
        return self.parameters
