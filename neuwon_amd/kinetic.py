"""Device-resident LTI node tables and the bucket-grouped kinetic advance (nw_lti_table_*).

The reference's generated kernel (neuwon/rxd/lti_sim/codegen.py:39-106) interpolates every
instance's propagator between the two table nodes of its voltage bucket.  Here the instances of a
wavefront that share a bucket read the two node matrices once (scalar loads), so a mechanism whose
instances sit on neighbouring segments -- voltages close together -- reads its table once per
wavefront and bucket instead of once per instance.
"""
import ctypes

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
