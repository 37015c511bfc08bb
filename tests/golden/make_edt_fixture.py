"""Writes edt_ties.npz: scipy.ndimage.distance_transform_edt(~active, return_indices=True) on small
seeded 3-D rasters with many equidistant active cells (the nearest-cell lookup of rb/rbsim.py:152).
The stored indices pin the tie rule sample_nodes reproduces on the GPU (smallest (k, j, i) among
the nearest cells) independently of the scipy installed where the tests run.

    python tests/golden/make_edt_fixture.py
"""
import os

import numpy as np
import scipy
from scipy import ndimage


def main():
    rng = np.random.default_rng(2024)
    out = {"scipy_version": np.array(scipy.__version__)}
    for r, (shape, dens) in enumerate((((8, 6, 7), 0.05), ((9, 7, 5), 0.15), ((6, 9, 8), 0.02))):
        act = rng.uniform(size=shape) < dens
        act[r, 0, 0] = True
        _, ind = ndimage.distance_transform_edt(~act, return_indices=True)
        out[f"active{r}"] = act
        out[f"indices{r}"] = ind.astype(np.int16)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "edt_ties.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
