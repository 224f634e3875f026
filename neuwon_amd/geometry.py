"""Segment geometry and passive electric properties, vectorised over whole populations (FP64).

Semantics follow neuwon/rxd/neuron/geometry.py:75-155 and electric.py:49-54:
  * a root segment is a sphere (length 2/3 d, area pi d^2); all others are cylinders;
  * the first child of a cylinder is its "primary" extension, every other child -- and every child
    of a sphere -- is a secondary branch whose length loses the parent's radius (geometry.py:112-127);
  * the final surface area drops the attachment disk of every secondary child and adds a tip disk on
    leaves (geometry.py:129-144);
  * the capacitance is computed once, when the segment is created (electric.py:45-54 called from
    segment.py:37), so it uses the area the segment had then: no children yet, tip disk included.
    Mechanism magnitudes use the area at insertion time (nmodl/__init__.py:499).
"""
import numpy as np

EPSILON = float(np.finfo(np.float64).eps)


def geometry(parent, length_or_coords, diam):
    """Geometry of segments in creation order (every parent index precedes its children).

    parent: int array, -1 for roots (global indices).  length_or_coords: either per-segment
    (N, 3) coordinates (length = distance to the parent) or an (N,) array of lengths (ignored for
    roots).  Returns a dict of FP64 arrays."""
    parent = np.asarray(parent, dtype=np.int64)
    diam = np.asarray(diam, dtype=np.float64)
    n = len(parent)
    idx = np.arange(n)
    is_root = parent < 0
    p = np.where(is_root, 0, parent)
    x = np.asarray(length_or_coords, dtype=np.float64)
    if x.ndim == 2:
        seg_len = np.linalg.norm(x - x[p], axis=1)
    else:
        seg_len = x.copy()
    length = np.where(is_root, (2.0 / 3.0) * diam, seg_len)
    child = ~is_root
    first_child = np.full(n, n, dtype=np.int64)
    np.minimum.at(first_child, p[child], idx[child])
    parent_is_sphere = child & is_root[p]
    primary = child & ~parent_is_sphere & (first_child[p] == idx)
    r_par = 0.5 * diam[p]
    shorten = child & ~primary & ~(length < r_par + EPSILON)
    sec_len = np.where(shorten, length - r_par, length)
    xarea = 0.25 * np.pi * diam ** 2
    disk = 0.25 * np.pi * diam * diam
    area_creation = np.where(is_root, np.pi * diam * diam, np.pi * diam * sec_len + disk)
    nchild = np.bincount(p[child], minlength=n)
    leaf = nchild == 0
    surface = np.where(is_root, np.pi * diam * diam, np.pi * diam * sec_len + np.where(leaf, disk, 0.0))
    sec_child = child & ~primary
    attach = 0.25 * np.pi * np.minimum(diam[p], diam) ** 2
    np.subtract.at(surface, p[sec_child], attach[sec_child])
    volume = np.where(is_root, (4.0 / 3.0) * np.pi * (0.5 * diam) ** 3, xarea * sec_len)
    return dict(length=length, primary=primary, surface_area=surface,
                area_at_creation=area_creation, cross_sectional_area=xarea, volume=volume)


def passive(length, xarea, area_at_creation, cytoplasmic_resistance=100.0, membrane_capacitance=1.0):
    """Axial resistance (Ohm) and capacitance (F); electric.py:49-54."""
    Ra = float(cytoplasmic_resistance) * 1e4  # ohm-cm -> ohm-um
    Cm = float(membrane_capacitance) * 1e-14  # uF/cm^2 -> F/um^2
    return Ra * np.asarray(length) / np.asarray(xarea), Cm * np.asarray(area_at_creation)
