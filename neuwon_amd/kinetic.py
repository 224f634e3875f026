"""Device-resident LTI node tables and the bucket-grouped kinetic advance (nw_lti_table_*).

The reference's generated kernel (neuwon/rxd/lti_sim/codegen.py:39-106) interpolates every
instance's propagator between the two table nodes of its voltage bucket.  Here the instances of a
wavefront that share a bucket read the two node matrices once (scalar loads), so a mechanism whose
instances sit on neighbouring segments -- voltages close together -- reads its table once per
wavefront and bucket instead of once per instance.
"""
import ctypes

import numpy as np

from . import _lib

DEFAULT_GROUP_MAX = -1   # the library's default number of distinct buckets per grouped wavefront


class LTITable:
    """Owner of an ``nw_lti_table`` handle.

    node_table: float32 device tensor [num_nodes, S, S] (row-major propagators on the linear grid
    vmin..vmax); kept referenced, since the handle reads it for held-input launches."""

    def __init__(self, node_table, vmin, vmax, device=0):
        if node_table.dim() != 3 or node_table.shape[1] != node_table.shape[2]:
            raise ValueError(f"node table of shape {tuple(node_table.shape)}: [nodes, S, S] expected")
        self.node_table = node_table.contiguous()
        self.num_nodes, self.S = int(node_table.shape[0]), int(node_table.shape[1])
        self.vmin, self.vmax = float(vmin), float(vmax)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().nw_lti_table_create(int(device), self.S, self.num_nodes, self.vmin,
                                                   self.vmax, _lib.ptr(self.node_table),
                                                   ctypes.byref(h)))
        self.handle = h

    def advance(self, states, voltage, seg_index=None, conserve_sum=0.0, n_steps=1,
                weights=None, gmax=None, cond_out=None, group_max=DEFAULT_GROUP_MAX, stream=None):
        """states: float32 device tensor [S, M] (SoA; row stride may exceed M through
        ``state_stride``); voltage gathered through seg_index (int32, or None: instance i reads
        voltage[i])."""
        import torch
        if states.dim() != 2 or states.shape[0] != self.S:
            raise ValueError(f"states of shape {tuple(states.shape)}: [{self.S}, M] expected")
        M = int(states.shape[1])
        s = stream if stream is not None else torch.cuda.current_stream(states.device)
        P = _lib.ptr
        _lib.check(_lib.load().nw_lti_table_advance(
            self.handle, M, P(voltage), P(seg_index), P(states), M, float(conserve_sum),
            int(n_steps), P(weights), P(gmax), P(cond_out), int(group_max),
            ctypes.c_void_p(s.cuda_stream)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().nw_lti_table_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def expm_batched(A, time_step, normalise_columns=True, stream=None):
    """expm(A time_step) of every matrix of a float64 device tensor [..., S, S] (S <= 16) on the
    device (nw_lti_expm_batched), every column divided by its sum when normalise_columns (the
    table build of lti_sim/lti_model.py:51-65)."""
    import torch
    S = int(A.shape[-1])
    if A.dim() < 2 or A.shape[-2] != S or A.dtype != torch.float64 or not A.is_cuda:
        raise ValueError(f"expm_batched: float64 device tensor [..., S, S] expected (got "
                         f"{tuple(A.shape)} {A.dtype} on {A.device})")
    A = A.contiguous()
    P = torch.empty_like(A)
    count = A.numel() // (S * S)
    s = stream if stream is not None else torch.cuda.current_stream(A.device)
    _lib.check(_lib.load().nw_lti_expm_batched(count, S, _lib.ptr(A), float(time_step),
                                               1 if normalise_columns else 0, _lib.ptr(P),
                                               ctypes.c_void_p(s.cuda_stream)))
    return P


def device_node_table(scheme, device=0):
    """The scheme's node table [nodes...][S][S] (float64, on the device): its rate matrices at the
    grid nodes (host, from the scheme's kinetic equations) exponentiated on the device."""
    import torch
    A = torch.as_tensor(np.ascontiguousarray(scheme.node_rates(), dtype=np.float64),
                        device=torch.device("cuda", device))
    return expm_batched(A, scheme.time_step, normalise_columns=True)
