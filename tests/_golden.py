"""Helpers shared by the tests: golden fixtures and the normwise error metric."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRIDS = ("g1", "g2", "g3")


def load(grid):
    z = np.load(os.path.join(GOLDEN, f"ref_ops_{grid}.npz"))
    params = {k[len("params_"):]: z[k].item() for k in z.files if k.startswith("params_")}
    inp = {k[len("in_"):]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[len("out_"):]: z[k] for k in z.files if k.startswith("out_")}
    return params, inp, out


def normwise(a, ref, mask=None):
    """max|a-ref| / max|ref| over the selected points (the per-field metric)."""
    a = np.asarray(a); ref = np.asarray(ref)
    if mask is not None:
        a = a[mask]; ref = ref[mask]
    den = np.max(np.abs(ref))
    num = np.max(np.abs(a - ref))
    return num / den if den > 0 else num


def interior(d, shape=None):
    """Boolean mask of the interior cells of a ghosted 3-D field."""
    m = np.zeros((d.kcells, d.jcells, d.icells), dtype=bool)
    m[d.kstart:d.kend, d.jstart:d.jend, d.istart:d.iend] = True
    return m
