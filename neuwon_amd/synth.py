"""Seeded synthetic morphologies and populations.

A cell is a random binary tree grown one compartment at a time: each new compartment attaches to
a compartment drawn uniformly among those with fewer than two children.  Compartment 0 is the
root (a sphere, as every NEUWON root is, geometry.py:36-40); lengths are U(10, 30) um and
diameters logU(0.5, 4) um.  Every draw comes from numpy's PCG64 generator, so a seed fixes the
whole population.
"""
import numpy as np


class CellSpec:
    """One cell: parent (local indices, -1 for the root), lengths, diameters, channel scales."""
    __slots__ = ("parent", "length", "diam", "scale_na", "scale_k")

    def __init__(self, parent, length, diam, scale_na=1.0, scale_k=1.0):
        self.parent, self.length, self.diam = parent, length, diam
        self.scale_na, self.scale_k = float(scale_na), float(scale_k)

    def __len__(self):
        return len(self.parent)


def random_tree(n, rng):
    """Parent array (local, -1 root) of a random binary tree of n compartments."""
    parent = np.empty(n, dtype=np.int32)
    parent[0] = -1
    if n == 1:
        return parent
    u = rng.random(n)
    nchild = np.zeros(n, dtype=np.int8)
    open_ = np.empty(n, dtype=np.int32)
    open_[0] = 0
    n_open = 1
    for i in range(1, n):
        k = int(u[i] * n_open)
        p = int(open_[k])
        parent[i] = p
        nchild[p] += 1
        if nchild[p] == 2:
            n_open -= 1
            open_[k] = open_[n_open]
        open_[n_open] = i
        n_open += 1
    return parent


def random_cell(n, rng, length_range=(10.0, 30.0), diam_range=(0.5, 4.0), scale_range=None):
    parent = random_tree(n, rng)
    length = rng.uniform(length_range[0], length_range[1], n)
    diam = np.exp(rng.uniform(np.log(diam_range[0]), np.log(diam_range[1]), n))
    sna = sk = 1.0
    if scale_range is not None:
        lo, hi = np.log(scale_range[0]), np.log(scale_range[1])
        sna, sk = np.exp(rng.uniform(lo, hi, 2))
    return CellSpec(parent, length, diam, sna, sk)


def single_cell(n=1000, seed=0):
    """One random tree of n compartments (single-neuron workload)."""
    return [random_cell(n, np.random.default_rng(seed))]


def population(total=10_000_000, size_range=(200, 5000), seed=1, scale_range=(0.5, 2.0),
               max_cells=None):
    """Independent cells with sizes ~ U{size_range}, drawn until their sum reaches `total`
    compartments (or `max_cells` cells).  gNa and gK are scaled per cell by logU(scale_range)."""
    rng = np.random.default_rng(seed)
    cells, acc = [], 0
    while acc < total and (max_cells is None or len(cells) < max_cells):
        n = int(rng.integers(size_range[0], size_range[1] + 1))
        cells.append(random_cell(n, rng, scale_range=scale_range))
        acc += n
    return cells


def uniform_population(num_cells=1024, n=500, seed=4, scale_range=None):
    rng = np.random.default_rng(seed)
    return [random_cell(n, rng, scale_range=scale_range) for _ in range(num_cells)]


def concat(cells):
    """Global arrays of a list of cells: cell_ptr (int64), parent (global, -1 roots),
    parent_local (int32), length, diam, cell index of every segment."""
    sizes = np.array([len(c) for c in cells], dtype=np.int64)
    cell_ptr = np.zeros(len(cells) + 1, dtype=np.int64)
    np.cumsum(sizes, out=cell_ptr[1:])
    parent_local = np.concatenate([c.parent for c in cells]).astype(np.int32)
    offs = np.repeat(cell_ptr[:-1], sizes)
    parent = np.where(parent_local < 0, -1, parent_local.astype(np.int64) + offs)
    length = np.concatenate([c.length for c in cells])
    diam = np.concatenate([c.diam for c in cells])
    cell_of = np.repeat(np.arange(len(cells), dtype=np.int64), sizes)
    return dict(cell_ptr=cell_ptr, parent=parent, parent_local=parent_local, length=length,
                diam=diam, cell_of=cell_of)
