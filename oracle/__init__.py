"""CPU FP64 oracle for the voxel finite-element solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``comech_amd``) imports, calls or
executes anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as the timed CPU
baseline), never as the thing that is measured or shipped.

The oracle is a plain numpy/scipy restatement of the reference algorithm (xtli12/CoMech):

* ``fe/solver.py``   -- grid numbering, Q1 hex element matrices, CSR assembly, springs, loads,
                         penalties, rigid modes, ``solve`` and ``cell_stress``;
* ``fe/gmg.py``      -- trilinear prolongation, Galerkin hierarchy, 3x3 block-Jacobi,
                         Chebyshev smoothing, V-cycle, ``pcg_gmg``;
* ``rb/splitfe.py``  -- split connectivity (loop and vectorised forms), SRI volumetric term,
                         ``assemble_cn``, ``solve_cn``, ``cell_stress_cn``, ``element_forces_cn``;
* ``rb/replica3.py`` -- ligament ``spring_matrix_cn``;
* ``rb/gpusolve.py`` -- the hand-rolled PCG with its residual-check cadence;
* ``rb/rbsim.py``    -- ``sample_fe`` resampling;
* ``net_oracle.py``  -- the learned warm-start network (no reference counterpart) in FP64, the
                         checker of the GPU training kernels, gradient-checked by central
                         differences as the reference's ml/test_ad.py:7-22 checks its autodiff.

Each function cites the reference file:line it follows.  How every function is pinned
(closed forms, invariants, brute force, library routines) is listed in ``oracle/fe_oracle.py``'s
module docstring.
"""
