"""Learned warm start of the voxel solve (no reference counterpart): a small 3-D U-Net trained and
evaluated with the project's BF16 matrix-core kernels (include/comech_learned.h), whose prediction,
projected onto its Galerkin coarse space, is the PCG initial guess (x0 of rb/gpusolve.py:229);
plain GMG-PCG is the fallback.

``WEIGHTS``: the network trained by ``python -m comech_amd.learned.train`` on 3,600 engine solves
of the hindfoot family at 96x64x64 voxels of 0.5 mm (450 seeded feet x 8 actions; 50 feet held
out).  ``load_default(device)`` returns it."""
import os

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights_96x64x64.npz")


def load_default(device):
    from .net import Net
    return Net.load(WEIGHTS, device)
