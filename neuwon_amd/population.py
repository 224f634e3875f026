"""Device-resident batches of independent HH cells advanced by the fused step kernel.

This is the throughput path: thousands of cells of uneven size share one block-diagonal
propagator and one launch advances all of them through any number of steps
(RxD_Model.advance, neuwon/rxd/rxd_model.py:63-123, once per step for every cell).
"""
import ctypes
import math

import numpy as np

from . import _lib, channels, geometry, synth

STIM_NONE, STIM_CONSTANT, STIM_POISSON = 0, 1, 2


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("neuwon_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


# cheb_launch names -> NW_CHEB_LAUNCH_* (include/neuwon_amd.h)
CHEB_LAUNCH = {None: 0, "half": 1, "step": 2, "two_steps": 3}
STORES = ("general", "symmetric", "chebyshev", "rational")


def store_name(store=None, symmetric=False):
    """The propagator store asked for by ``store`` (a name of STORES) or the older ``symmetric``
    flag."""
    if store is None:
        return "symmetric" if symmetric else "general"
    if store not in STORES:
        raise ValueError(f"propagator store {store!r}: one of {STORES}")
    return store


CHEBYSHEV_MAX_CELL = 5120   # segments per cell the Chebyshev store takes


class Propagator:
    """Owner of an ``nw_propagator`` handle.

    store="general": per-cell pruned CSR, FP32 values.  store="symmetric" (or symmetric=True)
    stores S = W P W^-1 (w = sqrt(C)) on and above the diagonal only; a pair (i, j) is then kept
    when max(P_ij, P_ji) >= prune.  store="chebyshev" stores no matrix: products are Chebyshev
    series of the cell's tree operator, unpruned (prune is ignored), for cells up to 5120
    segments.  store="rational" stores no matrix either: exp(B) x as a five-pole rational
    approximation, each pole an exact solve over the cell's tree in level order (unpruned; cells
    up to 5120 segments, at most 63 children per segment); it runs in the Chebyshev store's
    launches, so ``chebyshev`` is true for both matrix-free stores.  The fused HH step and the generic half step read every store, the adjoint the
    first two, the species step the general one.
    """

    def __init__(self, cell_ptr, parent_local, R, C, half_dt_ms, prune, device=0, symmetric=False,
                 store=None):
        lib = _lib.load()
        self.cell_ptr = np.ascontiguousarray(cell_ptr, dtype=np.int64)
        self._parent = np.ascontiguousarray(parent_local, dtype=np.int32)
        self._R = np.ascontiguousarray(R, dtype=np.float64)
        self._C = np.ascontiguousarray(C, dtype=np.float64)
        self.half_dt_ms, self.prune, self.device = float(half_dt_ms), float(prune), int(device)
        self.store = store_name(store, symmetric)
        self.symmetric = self.store == "symmetric"
        self.chebyshev = self.store in ("chebyshev", "rational")
        self.rational = self.store == "rational"
        flags = {"general": 0, "symmetric": _lib.NW_PROPAGATOR_SYMMETRIC,
                 "chebyshev": _lib.NW_PROPAGATOR_CHEBYSHEV,
                 "rational": _lib.NW_PROPAGATOR_RATIONAL}[self.store]
        h = ctypes.c_void_p()
        with _lib.lock():
            _lib.check(lib.nw_propagator_build_ex(self.device, len(self.cell_ptr) - 1,
                                                  _lib.ptr(self.cell_ptr), _lib.ptr(self._parent),
                                                  _lib.ptr(self._R), _lib.ptr(self._C),
                                                  self.half_dt_ms, self.prune, flags,
                                                  ctypes.byref(h)))
        self.handle = h
        st = _lib.NwPropagatorStats()
        _lib.check(lib.nw_propagator_stats_get(self.handle, ctypes.byref(st)))
        self.stats = {k: getattr(st, k) for k, _ in st._fields_}

    @property
    def mean_nnz_per_row(self):
        return self.stats["nnz"] / max(1, self.stats["num_rows"])

    def export_cell(self, cell):
        """The FP32 propagator block of one cell as a scipy matrix (FP64 values).  In the
        symmetric format P_ij = S_ij w_j / w_i is rebuilt from the stored triangle."""
        m = self.export_cell_stored(cell)
        if not self.symmetric:
            return m
        w = self.cell_weights(cell)
        full = (m + m.T - _diag_part(m)).tocsr()
        full = (full.multiply(w[None, :]).multiply(1.0 / w[:, None])).tocsr()
        full.eliminate_zeros()
        return full

    def cell_chebyshev(self, cell):
        """(rho, coefficients) of one cell's series (Chebyshev store)."""
        K, rho = ctypes.c_int32(), ctypes.c_double()
        lib = _lib.load()
        _lib.check(lib.nw_propagator_cell_chebyshev(self.handle, cell, ctypes.byref(K),
                                                    ctypes.byref(rho), None))
        c = np.empty(K.value + 1, dtype=np.float64)
        _lib.check(lib.nw_propagator_cell_chebyshev(self.handle, cell, ctypes.byref(K),
                                                    ctypes.byref(rho), _lib.ptr(c)))
        return rho.value, c

    def cell_weights(self, cell):
        """w of every segment of one cell (symmetric format)."""
        n = int(self.cell_ptr[cell + 1] - self.cell_ptr[cell])
        w = np.empty(n, dtype=np.float64)
        _lib.check(_lib.load().nw_propagator_cell_weights(self.handle, cell, _lib.ptr(w)))
        return w

    def export_cell_stored(self, cell):
        """Stored (FP32 value, u16 column) CSR block of one cell as a scipy matrix (FP64)."""
        import scipy.sparse
        n = int(self.cell_ptr[cell + 1] - self.cell_ptr[cell])
        off = np.empty(n + 1, dtype=np.uint32)
        _lib.check(_lib.load().nw_propagator_cell_export(self.handle, cell, _lib.ptr(off), None, None))
        cols = np.empty(int(off[-1]), dtype=np.uint16)
        vals = np.empty(int(off[-1]), dtype=np.float32)
        _lib.check(_lib.load().nw_propagator_cell_export(self.handle, cell, None, _lib.ptr(cols),
                                                         _lib.ptr(vals)))
        m = scipy.sparse.csr_matrix((vals.astype(np.float64), cols.astype(np.int64),
                                     off.astype(np.int64)), shape=(n, n))
        m.sum_duplicates()      # row padding: zero entries on the diagonal column
        m.eliminate_zeros()
        return m

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().nw_propagator_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _diag_part(m):
    import scipy.sparse
    return scipy.sparse.diags(m.diagonal(), format="csr")


def cell_propagator_f64(parent_local, R, C, half_dt_ms, prune, device=0):
    """Pruned FP64 propagator of one cell from the device construction path (scipy CSR)."""
    import scipy.sparse
    lib = _lib.load()
    parent_local = np.ascontiguousarray(parent_local, dtype=np.int32)
    R = np.ascontiguousarray(R, dtype=np.float64)
    C = np.ascontiguousarray(C, dtype=np.float64)
    n = len(parent_local)
    nnz = ctypes.c_int64()
    args = (device, n, _lib.ptr(parent_local), _lib.ptr(R), _lib.ptr(C), float(half_dt_ms),
            float(prune))
    _lib.check(lib.nw_propagator_cell_f64(*args, ctypes.byref(nnz), None, None, None))
    rp = np.empty(n + 1, dtype=np.int64)
    ci = np.empty(nnz.value, dtype=np.int32)
    va = np.empty(nnz.value, dtype=np.float64)
    _lib.check(lib.nw_propagator_cell_f64(*args, ctypes.byref(nnz), _lib.ptr(rp), _lib.ptr(ci),
                                          _lib.ptr(va)))
    return scipy.sparse.csr_matrix((va, ci, rp), shape=(n, n))


class HHPopulation:
    """Cells carrying leak + HH Na (m^3 h) + HH K (n^4) on every segment.

    cells: list of ``synth.CellSpec``.  Mechanisms are inserted after each cell is grown, so
    conductances use the final surface area (magnitude = area, nmodl/__init__.py:499) while the
    capacitance uses the area at creation (electric.py:45-54).

    stim: dict(mode='constant'|'poisson'|None, amp_na, rate_hz, duration_ms, seed) applied to the
    root (segment 0) of every cell.
    symmetric: store the propagator in the symmetric half format (half the bytes per step).
    store: "general", "symmetric", "chebyshev" or "rational" (no stored matrix; see Propagator),
    or "auto": the faster of the two matrix-free stores on these cells, timed by the autotuner.
    block_threads: workgroup size of the stored formats (128, 256, 512, 1024; 0: the default), or
    "auto": timed by the autotuner on this population, whose state is then reset.
    cheb_launch: how the Chebyshev store's steps map to launches ("half": one per half step,
    "step": one per step, "two_steps": one per two steps; None: two steps per launch when the
    store is built for five row groups, else one step), or "auto": each form timed by the
    autotuner on this population, whose state is then reset.  The forms give bitwise the same
    results.  The autotuner's measurements are kept in ``tuning``.
    """

    def __init__(self, cells, dt=0.1, celsius=6.3, v0=-70.0, cytoplasmic_resistance=100.0,
                 membrane_capacitance=1.0, gbar=None, reversal=None, prune=1e-9, gate_mode="table",
                 stim=None, ap_threshold=20.0, device=0, block_threads=0, insert_area="final",
                 symmetric=False, store=None, cheb_launch=None):
        if len(cells) == 0:
            raise ValueError("HHPopulation needs at least one cell")
        torch = _torch()
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.dt = float(dt)
        self.celsius = float(celsius)
        self.gate_mode = gate_mode
        self.tuning = {}
        if store == "auto":
            from . import autotune
            store, self.tuning["store"] = autotune.choose_store(
                cells, dt=dt, celsius=celsius, v0=v0, cytoplasmic_resistance=cytoplasmic_resistance,
                membrane_capacitance=membrane_capacitance, gbar=gbar, reversal=reversal,
                prune=prune, gate_mode=gate_mode, stim=stim, ap_threshold=ap_threshold,
                device=device, insert_area=insert_area)
        tune_blocks = block_threads == "auto"
        self.block_threads = 0 if tune_blocks else int(block_threads)
        tune_launch = cheb_launch == "auto"
        if not tune_launch and cheb_launch not in CHEB_LAUNCH:
            raise ValueError(f"cheb_launch {cheb_launch!r}: one of {sorted(CHEB_LAUNCH, key=str)} "
                             "or 'auto'")
        self.v0 = float(v0)
        g = synth.concat(cells)
        self.num_cells = len(cells)
        self.cell_ptr = g["cell_ptr"]
        self.num_segments = int(self.cell_ptr[-1])
        geo = geometry.geometry(g["parent"], g["length"], g["diam"])
        R, C = geometry.passive(geo["length"], geo["cross_sectional_area"], geo["area_at_creation"],
                                cytoplasmic_resistance, membrane_capacitance)
        self.R, self.C, self.geo, self.parent_local = R, C, geo, g["parent_local"]
        self.propagator = Propagator(self.cell_ptr, g["parent_local"], R, C, 0.5 * self.dt, prune,
                                     device, symmetric=symmetric, store=store)
        gbar = dict(channels.HH_GBAR, **(gbar or {}))
        self.reversal = dict(channels.HH_REVERSAL, **(reversal or {}))
        # magnitude = surface area when the mechanism is inserted: after the cell is grown
        # ("final") or right after each segment is created ("creation")
        area = geo["surface_area"] if insert_area == "final" else geo["area_at_creation"]
        sna = np.repeat([c.scale_na for c in cells], np.diff(self.cell_ptr))
        sk = np.repeat([c.scale_k for c in cells], np.diff(self.cell_ptr))
        self.g_bar = {"na": gbar["na"] * channels.CM2_TO_UM2 * area * sna,
                      "k": gbar["k"] * channels.CM2_TO_UM2 * area * sk,
                      "l": gbar["l"] * channels.CM2_TO_UM2 * area}
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=self.dev)
        N = self.num_segments
        self.v = torch.empty(N, dtype=torch.float32, device=self.dev)
        self.m, self.h, self.n = (torch.empty_like(self.v) for _ in range(3))
        self.g_na, self.g_k, self.g_l = f32(self.g_bar["na"]), f32(self.g_bar["k"]), f32(self.g_bar["l"])
        self.inv_c = f32(1.0 / C)
        self.grid = channels.LinearInput(-100.0, 100.0, 2000)
        self.table64 = channels.hh_gate_table(self.dt, celsius, self.grid)
        self.table = f32(self.table64)
        zeros_i32 = np.zeros(self.num_cells, dtype=np.int32)
        self.stim_local = torch.as_tensor(zeros_i32, device=self.dev)
        self.ap_local = torch.as_tensor(zeros_i32, device=self.dev)
        self.ap_state = torch.zeros(self.num_cells, dtype=torch.uint8, device=self.dev)
        self.ap_count = torch.zeros(self.num_cells, dtype=torch.int32, device=self.dev)
        self.reset_state()
        p = _lib.NwHHParams()
        p.dt_ms, p.vmin, p.vmax, p.num_nodes = self.dt, self.grid.minimum, self.grid.maximum, \
            self.grid.num_buckets + 1
        p.e_na, p.e_k, p.e_l = self.reversal["na"], self.reversal["k"], self.reversal["l"]
        p.ap_threshold = ap_threshold
        p.gate_mode = {"table": 0, "exact": 1}[gate_mode]
        p.q10 = channels.q10_factor(celsius)
        p.block_threads = self.block_threads
        p.cheb_launch = 0 if tune_launch else CHEB_LAUNCH[cheb_launch]
        self.params = p
        self.set_stim(stim)
        if tune_blocks:
            if self.propagator.chebyshev:
                self.tuning["block_threads"] = {
                    "choice": 0, "note": "matrix-free stores: the launch shape follows the cell sizes"}
            else:
                from . import autotune
                self.block_threads, self.tuning["block_threads"] = \
                    autotune.choose_block_threads(self)
                self.reset_state()
        if tune_launch:
            if self.propagator.chebyshev and not self.propagator.rational:
                from . import autotune
                _, self.tuning["cheb_launch"] = autotune.choose_cheb_launch(self)
                self.reset_state()
            else:
                self.tuning["cheb_launch"] = {
                    "choice": None, "note": "the launch form applies to the Chebyshev store only"}

    def reset_state(self):
        """Voltages at v0, gates at their steady state, AP detectors and the step count at 0."""
        ss = channels.hh_steady_state(self.v0, self.celsius)
        self.v.fill_(self.v0)
        self.m.fill_(float(ss["m"]))
        self.h.fill_(float(ss["h"]))
        self.n.fill_(float(ss["n"]))
        self.ap_state.zero_()
        self.ap_count.zero_()
        self.steps = 0

    def set_stim(self, stim):
        """Stimulus of the root of every cell from the next step on (None switches it off)."""
        p = self.params
        stim = stim or {}
        mode = stim.get("mode")
        self.stim = stim
        if mode == "constant":
            p.stim_mode = STIM_CONSTANT
            p.stim_amp_a = float(stim["amp_na"]) * 1e-9
        elif mode == "poisson":
            p.stim_mode = STIM_POISSON
            p.stim_amp_a = float(stim["amp_na"]) * 1e-9
            p.stim_seed = int(stim.get("seed", 0)) & 0xFFFFFFFF
            half = 0.5 * self.dt
            p.poisson_threshold = int(math.floor(float(stim["rate_hz"]) * half * 1e-3 * 4294967296.0))
            p.pulse_ticks = int(math.ceil(float(stim["duration_ms"]) / half))
        else:
            p.stim_mode = STIM_NONE

    def advance(self, n_steps=1, trace=None, stream=None, record=None, graph=True):
        """Advance every cell by n_steps time steps.  trace: optional float32 device tensor
        [num_cells, >= n_steps] receiving the AP-segment voltage after each step.  record:
        optional float32 device tensor of (5 n_steps + 4) * num_segments values receiving the
        state of every step for the reverse pass.
        The stored formats run all steps in one launch.  The matrix-free stores run one launch per
        half step; with graph=True (and n_steps >= 2) the 2 n_steps launches are replayed as one
        HIP graph, captured the first time these arguments (step count, trace, record) are seen
        and kept by the propagator (``prepare`` captures it ahead of time)."""
        self._launch(n_steps, trace, stream, record,
                     _lib.NW_LAUNCH_GRAPH if graph and n_steps >= 2 else _lib.NW_LAUNCH_DIRECT)

    def prepare(self, n_steps, trace=None, record=None, stream=None):
        """Capture the graph that advance(n_steps, trace, record) replays on a matrix-free store
        (nothing runs, the state is untouched); returns the number of graphs captured or found
        (0 on the stored formats, which need none)."""
        if not self.propagator.chebyshev or n_steps < 2:
            return 0
        steps = self.steps
        self._launch(n_steps, trace, stream, record, _lib.NW_LAUNCH_GRAPH_PREPARE)
        self.steps = steps
        return 1

    def _launch(self, n_steps, trace, stream, record, mode):
        torch = self.torch
        n_steps = int(n_steps)
        if n_steps < 0:
            raise ValueError(f"n_steps = {n_steps}: must be non-negative")
        self._check_output(trace, "trace", (self.num_cells, n_steps))
        self._check_output(record, "record", ((5 * n_steps + 4) * self.num_segments,))
        p = self.params
        p.step0 = self.steps
        p.conductances_valid = 1 if self.steps > 0 else 0
        p.launch_mode = mode
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        tr_ptr, tr_stride = None, 0
        if trace is not None:
            tr_ptr, tr_stride = _lib.ptr(trace), trace.shape[1]
        P = _lib.ptr
        with _lib.lock():
            _lib.check(_lib.load().nw_hh_step(
                self.propagator.handle, ctypes.byref(p), int(n_steps), P(self.table), P(self.v),
                P(self.m), P(self.h), P(self.n), P(self.g_na), P(self.g_k), P(self.g_l),
                P(self.inv_c), P(self.stim_local), P(self.ap_local), P(self.ap_state),
                P(self.ap_count), tr_ptr, tr_stride, P(record), ctypes.c_void_p(s.cuda_stream)))
        self.steps += n_steps

    def _check_output(self, t, name, shape):
        """An output array the kernels write: a contiguous float32 tensor on this population's
        device with at least ``shape`` (trace: [num_cells, >= n_steps]; record: a flat array of
        >= (5 n_steps + 4) num_segments values).  Anything else would be written out of bounds or
        through a host address, so it is refused before the launch."""
        if t is None:
            return
        torch = self.torch
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a torch tensor (got {type(t).__name__})")
        if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.dev:
            raise ValueError(f"{name} must be a contiguous float32 tensor on {self.dev} (got "
                             f"{t.dtype}, device {t.device}, contiguous {t.is_contiguous()})")
        if len(shape) == 1:
            if t.numel() < shape[0]:
                raise ValueError(f"{name} holds {t.numel()} values, {shape[0]} needed")
        elif t.dim() != 2 or t.shape[0] != shape[0] or t.shape[1] < shape[1]:
            raise ValueError(f"{name} of shape {tuple(t.shape)}: [{shape[0]}, >= {shape[1]}] needed")

    def time(self):
        return self.steps * self.dt

    # state snapshots ----------------------------------------------------------------------
    _STATE = ("v", "m", "h", "n", "g_na", "g_k", "g_l", "ap_state", "ap_count", "ap_local",
              "stim_local")

    def save(self, path):
        """Write every array the next step reads, and the step count (the Poisson stimulus is a
        function of (seed, cell, tick), so it needs no state of its own)."""
        from . import checkpoint
        checkpoint.save(path, {k: getattr(self, k) for k in self._STATE},
                        {"kind": "HHPopulation", "steps": self.steps, "dt": self.dt,
                         "num_segments": self.num_segments, "num_cells": self.num_cells,
                         "store": self.propagator.store, "stim": self.stim or None})

    def load(self, path):
        """Restore a snapshot written by ``save`` for the same cells."""
        from . import checkpoint
        arrays, meta = checkpoint.load(path, "HHPopulation")
        if (meta["num_segments"], meta["num_cells"]) != (self.num_segments, self.num_cells) or \
                meta["dt"] != self.dt:
            raise ValueError(f"{path}: snapshot of {meta['num_cells']} cells / "
                             f"{meta['num_segments']} segments at dt {meta['dt']}")
        for k in self._STATE:
            checkpoint.copy_into(getattr(self, k), arrays[k], k)
        self.steps = int(meta["steps"])
        self.set_stim(meta["stim"])
