"""Intracellular Ca2+ reaction-diffusion on batches of cells (species advance with diffusion).

The reference's species advance is c = max(0, c + dc/dt * dt); c = P_D c
(neuwon/rxd/species.py:54-70), run in every half step of the staggered step, but its diffusion matrix
is a stub (species.py:74-75).  Here the diffusion propagator of every cell is built like the electric
one (device FP64 expm, pruned CSR): per edge (child c, parent p) a diffusive conductance
k = D * xarea_c / L_c (um^3/ms) couples the two volumes, A[c,p] = k/vol_c, A[p,c] = k/vol_p.
A first-order buffer B <-> CaB (forward kon*Ca, backward koff, B + CaB = btot) is the mechanism; it
writes the Ca derivative accumulator.  Concentrations in uM, time in ms.
"""
import ctypes

import numpy as np

from . import _lib, geometry, synth
from .population import Propagator, _torch


def diffusion_edges(parent, length, xarea, D):
    """Per-segment diffusive conductance k of the edge to its parent (0 for roots)."""
    parent = np.asarray(parent)
    k = np.zeros(len(parent))
    child = parent >= 0
    k[child] = D * np.asarray(xarea)[child] / np.asarray(length)[child]
    return k


class CaPopulation:
    """Cells with diffusing Ca2+ and a first-order buffer in every compartment."""

    def __init__(self, cells, dt=0.1, D=0.6, kon=0.1, koff=0.01, btot=100.0, ca0=0.05,
                 pulse_fraction=0.05, pulse=1.0, seed=3, prune=1e-9, device=0, block_threads=0):
        torch = _torch()
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.dt, self.D, self.kon, self.koff, self.btot = float(dt), float(D), kon, koff, btot
        g = synth.concat(cells)
        self.cell_ptr = g["cell_ptr"]
        self.num_cells = len(cells)
        self.num_segments = N = int(self.cell_ptr[-1])
        geo = geometry.geometry(g["parent"], g["length"], g["diam"])
        self.geo, self.parent_local = geo, g["parent_local"]
        self.k = diffusion_edges(g["parent"], geo["length"], geo["cross_sectional_area"], D)
        self.vol = geo["volume"]
        # electric-form arguments: coefficient = time_step_ms * 1e-3 / (R * C) = dt k / vol
        with np.errstate(divide="ignore"):
            self.R_form = np.where(self.k > 0, 1e-3 / np.where(self.k > 0, self.k, 1.0), 0.0)
        self.propagator = Propagator(self.cell_ptr, g["parent_local"], self.R_form, self.vol,
                                     0.5 * self.dt, prune, device)
        rng = np.random.default_rng(seed)
        ca = np.full(N, float(ca0))
        ca[rng.random(N) < pulse_fraction] += float(pulse)
        self.cab_ref = btot * kon * ca0 / (kon * ca0 + koff)      # buffer at rest with ca0
        cab = np.full(N, self.cab_ref)
        self.ca0, self.cab0 = ca, cab
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=self.dev)
        # the bound buffer is kept as an offset from cab_ref
        self.ca, self.cab_offset, self.dcdt = f32(ca), f32(cab - self.cab_ref), f32(np.zeros(N))
        self.vol_dev = torch.as_tensor(self.vol.astype(np.float64), device=self.dev)
        self.steps = 0
        p = _lib.NwCaParams()
        p.dt_ms, p.kon, p.koff, p.btot, p.block_threads = self.dt, kon, koff, btot, int(block_threads)
        p.cab_ref = self.cab_ref
        self.params = p
        self._partials = torch.zeros(1024, dtype=torch.float64, device=self.dev)

    def advance(self, n_steps=1, stream=None):
        s = stream if stream is not None else self.torch.cuda.current_stream(self.dev)
        P = _lib.ptr
        _lib.check(_lib.load().nw_ca_step(self.propagator.handle, ctypes.byref(self.params),
                                          int(n_steps), P(self.ca), P(self.cab_offset), P(self.dcdt),
                                          ctypes.c_void_p(s.cuda_stream)))
        self.steps += n_steps

    def mass(self):
        """Conserved total Ca, sum vol (Ca + CaB + dcdt dt/2), in FP64 on the device: the last
        buffer change is compensated in Ca over two half steps, the second of them in the next step."""
        s = self.torch.cuda.current_stream(self.dev)
        P = _lib.ptr
        _lib.check(_lib.load().nw_weighted_sum(self.num_segments, P(self.vol_dev), P(self.ca),
                                               P(self.cab_offset), P(self.dcdt), 0.5 * self.dt,
                                               P(self._partials), 1024,
                                               ctypes.c_void_p(s.cuda_stream)))
        return float(self._partials.sum().item()) + self.cab_ref * float(self.vol.sum())

    def cab(self):
        """Bound buffer per compartment (host, FP64)."""
        return self.cab_ref + self.cab_offset.cpu().numpy().astype(np.float64)

    def save(self, path):
        """Snapshot of Ca, the bound buffer (offset from cab_ref), the pending dc/dt and the step."""
        from . import checkpoint
        checkpoint.save(path, {"ca": self.ca, "cab_offset": self.cab_offset, "dcdt": self.dcdt},
                        {"kind": "CaPopulation", "steps": self.steps, "dt": self.dt,
                         "num_segments": self.num_segments, "cab_ref": self.cab_ref})

    def load(self, path):
        from . import checkpoint
        arrays, meta = checkpoint.load(path, "CaPopulation")
        if meta["num_segments"] != self.num_segments or meta["cab_ref"] != self.cab_ref:
            raise ValueError(f"{path}: snapshot of another population")
        for k in ("ca", "cab_offset", "dcdt"):
            checkpoint.copy_into(getattr(self, k), arrays[k], k)
        self.steps = int(meta["steps"])
