"""State snapshots: every state array of a population or model in one ``.npz`` file.

The reference's only persistence is its SQL export (neuwon/database/sql.py:78-128), which keeps a
subset of the columns; a snapshot here holds everything the next step reads, so advancing from a
restored snapshot reproduces the original run (bit for bit with the general propagator store;
the symmetric store's atomically accumulated sums agree to FP64 rounding).  Files are plain
numpy archives, written and read without pickling: arrays plus one JSON string of metadata.
"""
import json

import numpy as np

FORMAT = "neuwon_amd-snapshot-1"


def save(path, arrays, meta):
    """Write ``arrays`` (name -> numpy array or tensor) and ``meta`` (JSON-serialisable dict)."""
    out = {}
    for name, a in arrays.items():
        if hasattr(a, "detach"):
            a = a.detach().cpu().numpy()
        a = np.asarray(a)
        if a.dtype == object:
            raise TypeError(f"{name}: object arrays cannot be stored without pickling")
        out[name] = a
    out["__meta__"] = np.array(json.dumps(dict(meta, format=FORMAT)))
    with open(path, "wb") as f:
        np.savez(f, **out)


def load(path, kind=None):
    """Read a snapshot: (arrays, meta).  ``kind`` checks the writer's kind."""
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files if k != "__meta__"}
        meta = json.loads(str(z["__meta__"]))
    if meta.get("format") != FORMAT:
        raise ValueError(f"{path}: not a snapshot of this library ({meta.get('format')!r})")
    if kind is not None and meta.get("kind") != kind:
        raise ValueError(f"{path}: a {meta.get('kind')} snapshot, not {kind}")
    return arrays, meta


def copy_into(tensor, array, name):
    """Overwrite a device tensor with a saved array of the same shape."""
    if tuple(tensor.shape) != tuple(array.shape):
        raise ValueError(f"{name}: snapshot shape {array.shape} does not match {tuple(tensor.shape)}")
    import torch
    tensor.copy_(torch.as_tensor(np.ascontiguousarray(array), device=tensor.device).to(tensor.dtype))
