"""Learned warm start and coarse correction of the voxel solve (no reference counterpart): a small
3-D U-Net trained and evaluated with the project's BF16 matrix-core kernels (include/comech_learned.h),
whose prediction is the PCG initial guess (x0 of rb/gpusolve.py:229) and spans an additive coarse
correction (Z of rb/gpusolve.py:217-228); plain GMG-PCG is the fallback."""
