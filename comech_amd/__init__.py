"""comech_amd: the voxel finite-element solve path of CoMech, rebuilt for AMD Instinct MI355X.

Matrix-free 8-node hex elasticity, split nodes and springs, geometric-multigrid preconditioned CG
and von Mises recovery, as hand-written HIP kernels (gfx950) behind the reference's Python entry
points:

    comech_amd.fe.solver   -> fe/solver.py   (Grid, lame, solve, cell_stress, ...)
    comech_amd.fe.gmg      -> fe/gmg.py      (pcg_gmg)
    comech_amd.fe.simulate -> fe/simulate.py (simulate; simulate_raster for voxel rasters)
    comech_amd.rb.splitfe  -> rb/splitfe.py  (split_connectivity_fast, solve_cn, cell_stress_cn, ...)
    comech_amd.rb.gpusolve -> rb/gpusolve.py (pcg_gmg_gpu, pick_device, free_gpu)
    comech_amd.rb.replica3 -> rb/replica3.py (spring_matrix_cn)
    comech_amd.rb.rbsim    -> rb/rbsim.py    (sample_fe)

The HIP library (libcomech_hip.so) is required: there is no CPU fallback.
"""
__version__ = "0.1.0"
