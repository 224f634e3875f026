"""Writes tests/golden/kinetic_tables/*.npz: the rate matrices at the table nodes of two of the
reference's kinetic schemes, parsed in this repository's container by the build's own NMODL front
end (neuwon_amd.nmodl) from the reference's files

    nmodl_library/kinetic/Nav11.mod     (6 states, input v; celsius 37)
    neuwon/rxd/lti_sim/tests/ampa13.mod (13 states, input C; celsius 37)

The reference does not travel to the GPU box, so its schemes reach the GPU tests as these vectors:
per scheme the nonzero pattern of the rate matrices and their values at every node of the table's
grid (float64), the grid, the state names and the CONSERVE total.  Usage:

    python tests/golden/make_kinetic_tables.py /root/reference
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from neuwon_amd import channels, nmodl  # noqa: E402

FILES = {"Nav11": ("nmodl_library/kinetic/Nav11.mod", None),
         "ampa13": ("neuwon/rxd/lti_sim/tests/ampa13.mod", ["C"])}


def grid_json(g):
    if isinstance(g, channels.LogarithmicInput):
        return {"kind": "log", "minimum": g.minimum, "maximum": g.maximum,
                "num_buckets": g.num_buckets, "scale": g.scale}
    return {"kind": "linear", "minimum": g.minimum, "maximum": g.maximum,
            "num_buckets": g.num_buckets}


def main(ref):
    out_dir = os.path.join(HERE, "kinetic_tables")
    os.makedirs(out_dir, exist_ok=True)
    for name, (rel, inputs) in FILES.items():
        path = os.path.join(ref, rel)
        km = nmodl.KineticModel(path, celsius=37.0, inputs=inputs)
        sc = nmodl.load_kinetic(path, time_step=0.1, celsius=37.0)
        A = sc.node_rates()
        S = sc.num_states
        flat = A.reshape(-1, S, S)
        pattern = np.argwhere(np.any(flat != 0.0, axis=0)).astype(np.int32)
        values = flat[:, pattern[:, 0], pattern[:, 1]].astype(np.float64)
        meta = {"name": sc.name, "source": rel, "states": sc.state_names,
                "inputs": sc.input_names, "grids": [grid_json(g) for g in sc.inputs],
                "conserve_sum": sc.conserve_sum, "time_step": sc.time_step, "celsius": 37.0,
                "num_states": S, "node_shape": list(A.shape[:-2]),
                "weights": None if sc.weights is None else list(map(float, sc.weights))}
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), pattern=pattern, values=values,
                            meta=np.array(json.dumps(meta)))
        print(f"{name}: {S} states, inputs {sc.input_names}, nodes {A.shape[:-2]}, "
              f"{len(pattern)} nonzero rates; parsed as {km.name}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
