"""Gradients of a voltage-trace loss through the fused step: a hand-written reverse pass with step
checkpointing, behind a torch.autograd.Function.

Loss: L = sum over cells and steps of (v_soma(step) - target)^2, v_soma the voltage of every
cell's root after each step.  Parameters: per-segment conductances g_na, g_k, g_l (Siemens) and six
rate multipliers theta = (ta_m, tb_m, ta_h, tb_h, ta_n, tb_n) scaling the HH alpha and beta rates
(hh.mod:62-86).  The forward pass keeps the state every `checkpoint_every` steps; the reverse pass
recomputes each segment with nw_hh_step writing a per-step record and runs nw_hh_adjoint over it.
The reference has no differentiation; this is the build's own extension of its step.

Both passes run on the Chebyshev store by default (no stored matrix, unpruned, the fastest step;
the reverse pass applies P^T lambda = C P (C^-1 lambda) with the same series); the symmetric and
general stores (pruned at `prune`) remain available for either pass.
"""
import ctypes

import numpy as np

from . import _lib, channels, population


class HHAdjoint:
    """Batched cells with HH on every segment and a constant current at every root.

    forward_store / reverse_store: propagator store of the forward passes (the loss and the
    recomputation with the record) and of the reverse pass: "chebyshev" (the default for both
    when every cell has at most 5120 segments), "symmetric" or "general" (see
    population.Propagator).  symmetric: the stored format chosen where a store is not named and
    the cells do not fit the Chebyshev store (symmetric: half the bytes).
    checkpoint_every: steps between the forward pass's checkpoints, each segment recomputed with
    its record in the reverse pass; None keeps the record of every step from the forward pass
    in device memory instead ((5 T + 4) FP32 values per segment) and recomputes nothing."""

    def __init__(self, cells, dt=0.05, celsius=6.3, stim_amp_na=0.5, prune=1e-9,
                 checkpoint_every=100, v0=-70.0, device=0, symmetric=True, forward_store=None,
                 reverse_store=None):
        fits = max(len(c) for c in cells) <= population.CHEBYSHEV_MAX_CELL
        default = "chebyshev" if fits else ("symmetric" if symmetric else "general")
        forward_store = forward_store or default
        reverse_store = reverse_store or default
        self.pop = population.HHPopulation(cells, dt=dt, celsius=celsius, v0=v0, prune=prune,
                                           gate_mode="table", device=device,
                                           stim={"mode": "constant", "amp_na": stim_amp_na},
                                           store=forward_store)
        pop = self.pop
        if forward_store == reverse_store:
            self.reverse = pop.propagator
        else:
            self.reverse = population.Propagator(pop.cell_ptr, pop.parent_local, pop.R, pop.C,
                                                 0.5 * pop.dt, prune, device=device,
                                                 store=reverse_store)
        self.forward_store = forward_store
        self.torch = self.pop.torch
        self.dev = self.pop.dev
        # checkpoint_every=None keeps the whole record in device memory (5 T + 4 floats per
        # segment: 20.5 GB for 512k segments x 2000 steps): no recomputation in the reverse pass
        self.K = None if checkpoint_every is None else int(checkpoint_every)
        self.v0 = float(v0)
        self.N = self.pop.num_segments
        self.num_cells = self.pop.num_cells
        self.soma = self.pop.ap_local          # roots (segment 0 of every cell)
        self._record = None

    # ------------------------------------------------------------------------------------------
    def _set_parameters(self, g_na, g_k, g_l, theta):
        pop, torch = self.pop, self.torch
        pop.g_na.copy_(g_na)
        pop.g_k.copy_(g_k)
        pop.g_l.copy_(g_l)
        theta = np.asarray(theta, dtype=np.float64)
        self.theta = theta
        t6 = channels.hh_gate_table_theta(pop.dt, pop.celsius, theta, pop.grid)
        self.table6 = torch.as_tensor(t6.astype(np.float32), device=self.dev)
        pop.table.copy_(torch.as_tensor(t6[:, :, :2].astype(np.float32).copy(), device=self.dev))
        self.rest, self.drest = channels.hh_rest_theta(self.v0, pop.celsius, theta)

    def _reset_state(self):
        pop = self.pop
        pop.v.fill_(self.v0)
        for arr, r in zip((pop.m, pop.h, pop.n), self.rest):
            arr.fill_(r)
        pop.steps = 0

    def forward(self, g_na, g_k, g_l, theta, target):
        """Run the trajectory, keep checkpoints, return the loss (FP64 float)."""
        torch = self.torch
        pop = self.pop
        target = torch.as_tensor(target)
        if target.dim() != 2 or target.shape[0] != self.num_cells or target.shape[1] < 1:
            raise ValueError(f"target of shape {tuple(target.shape)}: one row of at least one "
                             f"step per cell ({self.num_cells} rows) expected")
        self._set_parameters(g_na, g_k, g_l, theta)
        self._reset_state()
        T = int(target.shape[1])
        self.T = T
        # the reverse kernel reads FP32 rows of T values on the device: any other dtype, device
        # or layout is converted here
        self.target = target.to(device=self.dev, dtype=torch.float32).contiguous()
        trace = torch.zeros((self.num_cells, T), dtype=torch.float32, device=self.dev)
        self.ckpts = []
        if self.K is None:   # one pass that writes the whole record
            need = (5 * T + 4) * self.N
            if self._record is None or self._record.numel() < need:
                self._record = None
                self._record = torch.empty(need, dtype=torch.float32, device=self.dev)
            pop.advance(T, trace=trace, record=self._record)
            self.trace = trace
            return float(((trace.double() - self.target.double()) ** 2).sum().item())
        # one trace buffer for every segment of every gradient, so that every full segment
        # replays the same graph of the step's launches
        if getattr(self, "_seg_trace", None) is None:
            self._seg_trace = torch.zeros((self.num_cells, self.K), dtype=torch.float32,
                                          device=self.dev)
        seg_trace = self._seg_trace
        for s0 in range(0, T, self.K):
            n = min(self.K, T - s0)
            self.ckpts.append((s0, n, pop.v.clone(), pop.m.clone(), pop.h.clone(), pop.n.clone()))
            tr = seg_trace if n == self.K else \
                torch.zeros((self.num_cells, n), dtype=torch.float32, device=self.dev)
            pop.advance(n, trace=tr)
            trace[:, s0:s0 + n] = tr
        self.trace = trace
        return float(((trace.double() - self.target.double()) ** 2).sum().item())

    def backward(self):
        """Reverse pass; returns dict of gradients (torch FP64 device tensors for g_*, numpy theta)."""
        torch = self.torch
        pop = self.pop
        N = self.N
        z32 = lambda: torch.zeros(N, dtype=torch.float32, device=self.dev)
        z64 = lambda: torch.zeros(N, dtype=torch.float64, device=self.dev)
        lam = [z32() for _ in range(4)]
        grads = [z64() for _ in range(3)]
        gth = torch.zeros((self.num_cells, 6), dtype=torch.float64, device=self.dev)
        loss = torch.zeros(self.num_cells, dtype=torch.float64, device=self.dev)
        P = _lib.ptr
        prm = _lib.NwHHParams.from_buffer_copy(pop.params)
        if self.K is None:   # the forward pass kept the record of every step
            segments = [(0, self.T, None, None, None, None)]
        else:
            need = (5 * self.K + 4) * N
            if self._record is None or self._record.numel() < need:
                self._record = torch.empty(need, dtype=torch.float32, device=self.dev)
            segments = reversed(self.ckpts)
        for s0, n, v, m, h, nn in segments:
            if v is not None:   # recompute the segment from its checkpoint, writing the record
                pop.v.copy_(v)
                pop.m.copy_(m)
                pop.h.copy_(h)
                pop.n.copy_(nn)
                pop.steps = s0
                pop.advance(n, record=self._record)
            prm.step0 = s0
            prm.conductances_valid = 1 if s0 > 0 else 0
            s = torch.cuda.current_stream(self.dev)
            _lib.check(_lib.load().nw_hh_adjoint(
                self.reverse.handle, ctypes.byref(prm), n, P(self._record), P(self.table6),
                P(pop.g_na), P(pop.g_k), P(pop.g_l), P(pop.inv_c), P(self.soma), P(self.target),
                self.T, P(lam[0]), P(lam[1]), P(lam[2]), P(lam[3]), P(grads[0]), P(grads[1]),
                P(grads[2]), P(gth), P(loss), ctypes.c_void_p(s.cuda_stream)))
        theta_grad = gth.sum(dim=0).cpu().numpy()
        # the initial gates are the rest state of the scaled rates
        for gi in range(3):
            lsum = float(lam[1 + gi].double().sum().item())
            theta_grad[2 * gi] += lsum * self.drest[gi][0]
            theta_grad[2 * gi + 1] += lsum * self.drest[gi][1]
        self.last_loss = float(loss.sum().item())
        return {"na": grads[0], "k": grads[1], "l": grads[2], "theta": theta_grad}

    def loss_and_grad(self, g_na, g_k, g_l, theta, target):
        loss = self.forward(g_na, g_k, g_l, theta, target)
        return loss, self.backward()


def trajectory_loss(problem, g_na, g_k, g_l, theta, target):
    """torch.autograd entry: L(g_na, g_k, g_l, theta) with the hand-written reverse pass."""
    import torch

    class _TrajectoryLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, g_na, g_k, g_l, theta):
            loss = problem.forward(g_na.detach().float(), g_k.detach().float(),
                                   g_l.detach().float(), theta.detach().cpu().numpy(), target)
            return torch.tensor(loss, dtype=torch.float64, device=g_na.device)

        @staticmethod
        def backward(ctx, grad_out):
            gr = problem.backward()
            go = grad_out.to(torch.float64)
            th = torch.as_tensor(gr["theta"], dtype=torch.float64, device=grad_out.device)
            return (go * gr["na"]).to(g_na.dtype), (go * gr["k"]).to(g_k.dtype), \
                (go * gr["l"]).to(g_l.dtype), (go * th).to(theta.dtype)

    return _TrajectoryLoss.apply(g_na, g_k, g_l, theta)
