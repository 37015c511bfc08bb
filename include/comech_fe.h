/* C ABI of the MI355X voxel finite-element solve path (libcomech_hip.so).
 *
 * The reference (xtli12/CoMech) is pure Python; its GPU solver plugs in by replacing the module
 * attribute fe.gmg.pcg_gmg with rb.gpusolve.pcg_gmg_gpu (rb/rbfe.py:20-27).  This library is what
 * the Python drop-ins of that surface (comech_amd.fe.solver, comech_amd.fe.gmg,
 * comech_amd.rb.splitfe, comech_amd.rb.gpusolve) bind through ctypes.  Each entry point below names
 * the reference interface it serves.
 *
 * Conventions
 *  - Return value 0 on success, -1 on failure; cm_last_error() gives the message (thread-local).
 *  - Host pointers are read during the call only.  "device" pointers are HIP device memory owned
 *    by the caller.
 *  - Node vectors are AoS with 3 components per node.  Level-0 node ids are the reference's grid
 *    ids i + (nx+1)(j + (ny+1)k) (fe/solver.py:107) followed by split-node copies nnode + r
 *    (rb/splitfe.py:268-280).  Cell ids passed in are the reference's z-fastest ids
 *    (i*ny + j)*nz + k (fe/solver.py:104-105,128).
 *  - Precision codes: 0 = FP64, 1 = FP32.
 */
#ifndef COMECH_FE_H
#define COMECH_FE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CM_ABI_VERSION 1

typedef struct cm_sys cm_sys;

/* Result of cm_sys_solve; mirrors the info dict of rb/gpusolve.py:259-260,285. */
typedef struct cm_solve_info {
  int iters;        /* PCG iterations (all inner iterations in mixed mode) */
  int flag;         /* 0 converged, else maxiter (rb/gpusolve.py:252-254); -2: the early-exit
                       rule of cm_sys_set_fallback stopped the solve */
  int outer;        /* refinement steps (mixed mode), recurrence starts otherwise */
  int levels;       /* multigrid levels */
  double relres;    /* ||b - A x|| / ||b|| (FP64) at exit */
  double t_solve_ms;
  double t_setup_ms;
} cm_solve_info;

const char* cm_last_error(void);
int cm_version(void);
int cm_device_count(int* n);

/* ---- system description (replaces the CSR assembly of fe/solver.py:137-162 and
 *      rb/splitfe.py:72-91 / rb/gpusolve.py:306-347: the GPU path is matrix-free) ------------- */

/* Grid of nx*ny*nz cells with edges xe[nx+1], ye[ny+1], ze[nz+1] (fe/solver.py:93-100).
 * sri = 1: volumetric term at the cell centre (rb/splitfe.py:58-69); 0: 2x2x2 Gauss
 * (fe/solver.py:77-90).  stream: hipStream_t for every launch (0 = default stream). */
int cm_sys_create(cm_sys** out, void* stream, int nx, int ny, int nz, const double* xe, const double* ye,
                  const double* ze, int sri);
int cm_sys_destroy(cm_sys* sys);

/* Materials as labels: lam_tab/mu_tab[nmat] (nmat <= 255), lab_ref[ncell] in reference cell order,
 * 255 = inactive cell.  Equivalent to lam[cells], mu[cells] of fe/solver.py:152. */
int cm_sys_set_labels(cm_sys* sys, int nmat, const double* lam_tab, const double* mu_tab, const uint8_t* lab_ref);

/* Materials per cell: lam_ref, mu_ref, active_ref[ncell] (fe/solver.py:131-152). */
int cm_sys_set_cellmat(cm_sys* sys, const double* lam_ref, const double* mu_ref, const uint8_t* active_ref);

/* Split connectivity (rb/splitfe.py:261-305): copy_orig[ncopy] = original grid node of copy r;
 * irregular cells (active cells whose cn row differs from the grid corners) with their 8 node
 * ids in local order l = a + 2b + 4c (fe/solver.py:11-12). */
int cm_sys_set_split(cm_sys* sys, int64_t ncopy, const int64_t* copy_orig, int64_t nirr, const int64_t* irr_cell_ref,
                     const int64_t* irr_nodes);

/* Extra stiffness in COO over level-0 node DOFs 3*node+c: diagonal springs (fe/solver.py:210-220,
 * rb/splitfe.py:139-148), rank-1 penalties (fe/solver.py:222-232), ligament springs
 * (rb/replica3.py:68-86) and interface springs.  Duplicates are summed. */
int cm_sys_set_extra(cm_sys* sys, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals);

/* Extra stiffness as rank-one terms sum_t k[t] w_t w_t^T, w_t = (dofs, w) entries ptr[t] ..
 * ptr[t+1]-1 over level-0 node DOFs: node-pair springs (interface: one term per component, w = e_a
 * - e_b), directional ligaments (w = +d at a, -d at b; rb/replica3.py:68-86 patches), penalties
 * (fe/solver.py:222-232).  Coarsened exactly as w -> P^T w. */
int cm_sys_set_rank1(cm_sys* sys, int64_t nterm, const int64_t* ptr, const int64_t* dofs, const double* w,
                     const double* k);

/* Mark nodes as unknowns in addition to the corners of active cells (nn = grid nodes + copies).
 * Used when the stiffness comes only from cm_sys_set_extra, as for the CSR input of
 * fe/gmg.py:120 pcg_gmg(A, b, used_nodes, ...). */
int cm_sys_set_used(cm_sys* sys, int64_t nn, const uint8_t* used);

/* Additive rigid-mode correction z += Z Einv Z^T r of fe/gmg.py:127-135 / rb/gpusolve.py:217-228:
 * Z in COO over level-0 DOFs (ncol columns), Einv dense ncol x ncol (row-major). */
int cm_sys_set_rigid(cm_sys* sys, int ncol, int64_t nnz, const int64_t* rows, const int64_t* cols,
                     const double* vals, const double* Einv);

/* Build the multigrid hierarchy (fe/gmg.py:79-102, rb/gpusolve.py:39-65): coarsen 2x (every other
 * node plus the last) until the level has <= min_dofs active DOFs or max_levels levels
 * (fixed_levels > 0 forces that count).  smoother 0 = Chebyshev of degree `degree` with 3x3
 * block Jacobi (fe/gmg.py:47-76), 1 = damped block Jacobi with npre/npost sweeps. */
int cm_sys_build(cm_sys* sys, int min_dofs, int max_levels, int fixed_levels, int smoother, int degree, double omega,
                 int npre, int npost);

/* Coarsening of the step to the coarsest level (call before cm_sys_build): bit mask of the
 * coarsened axes, x = 1, y = 2, z = 4.  7 is the reference rule (fe/gmg.py:91-96: every axis,
 * every level).  The default 3 coarsens x and y only on that last step when the resulting level
 * still has <= min_dofs DOFs (semi-coarsening; no reference counterpart): the coarsest grid keeps
 * the vertical resolution, which on the hindfoot raster cuts the PCG iterations from about 100 to
 * about 60 for the same tolerance.  Only the automatic level rule uses it (fixed_levels = 0). */
int cm_sys_set_last_coarsening(cm_sys* sys, int axes);

/* Level table: 6 int64 per level: nx, ny, nz, nodes (incl. copies), compact DOFs, extra rows. */
int cm_sys_levels(cm_sys* sys, int* nlev, int64_t* desc, int cap);

/* Operator structure of one level (diagnostics; no reference counterpart): up to 15 int64 --
 * cells, irregular cells, extra-stiffness CSR rows and entries, rank-one terms and their DOF
 * entries, coarse block-row entries and the longest block row, rank-one node-list entries and the
 * longest list of one node, then the z extent of a structured-march block for FP64 products, FP64
 * smoothing steps, FP32 products and FP32 smoothing steps, then (level 1) the number of cells whose
 * level-0 children carry several materials (applied by the heterogeneous-cell pass). */
int cm_sys_level_stats(cm_sys* sys, int level, int64_t* stats, int cap);

/* ---- compute (device pointers) ------------------------------------------------------------- */

/* y = A_level u (A64 @ p of rb/gpusolve.py:239; L.A @ x of :104). */
int cm_sys_apply(cm_sys* sys, int level, int prec, const void* u, void* y);

/* x = V-cycle(b) on level 0 (GPUMG.vcycle, rb/gpusolve.py:163-171). */
int cm_sys_vcycle(cm_sys* sys, int prec, const void* b, void* x);

/* Early exit of a solve started from a predicted initial guess (modes 0 and 1): at iteration
 * `iters` the FP64 true residual ||b - A x|| must be below ||b - A x0|| / drop (or already below
 * tol), else cm_sys_solve stops with info->flag = -2 and the caller solves again without the
 * prediction (plain GMG-PCG as the fallback).  iters = 0 turns the rule off (default).  No
 * reference counterpart: the reference's rb engine ignores U0 (fe/simulate.py:17). */
int cm_sys_set_fallback(cm_sys* sys, int iters, double drop);

/* Multigrid level internals (fe/gmg.py:29-76; tests and diagnostics).  Levels below the
 * coarsest only.  what 0: the Chebyshev bounds (lmax, lmin) of the level in precision prec, 2
 * host doubles at out.  what 1: the block-Jacobi inverses of fe/gmg.py:29-44 (blocks with
 * |det| < 1e-30 -> I), 6 values per level node (xx, xy, xz, yy, yz, zz of the symmetric 3x3
 * inverse), copied to device memory at out in precision prec.  Level 0 with split-copy smoother
 * blocks (cm_sys_set_copy_blocks) holds zero blocks for the group members.  what 2 (any level):
 * the level's used-node mask, one host byte per node (the kept DOFs of fe/gmg.py:95-96, / 3). */
int cm_sys_level_data(cm_sys* sys, int level, int prec, int what, void* out);

/* Chebyshev bounds of one level from 12 power steps of Dinv A started at x0 (device, the level's
 * node layout, precision prec): the Level constructor of fe/gmg.py:51-58 (there x0 is
 * numpy.random.default_rng(0).standard_normal on the level's compact DOFs).  lmax = 1.1 lam and
 * lmin = lmax / 30 replace the level's bounds; *lam receives the estimate. */
int cm_sys_power(cm_sys* sys, int level, int prec, const void* x0, double* lam);

/* One call of the level's smoother in place: Level.smooth(x, b) of fe/gmg.py:61-76 (Chebyshev)
 * or npre damped block-Jacobi sweeps x += omega Dinv (b - A x).  zero_start = 1: x = smooth(0, b)
 * (x is not read).  b, x: device, the level's node layout, precision prec. */
int cm_sys_smooth(cm_sys* sys, int level, int prec, const void* b, void* x, int zero_start);

/* PCG to ||b - A x|| <= tol ||b|| (rb/gpusolve.py:206-262, fe/gmg.py:120-137).  b, x: FP64
 * device, level-0 layout; x holds the initial guess (x0) on entry.
 * mode 0: FP64 PCG, FP64 V-cycle.  mode 1: FP64 PCG, FP32 V-cycle (mg_dtype="float32" of
 * rb/gpusolve.py:265).  mode 2: iterative refinement -- FP64 true residual b - A64 x, inner FP32
 * PCG (FP32 operator, V-cycle and vectors, FP64 alpha/beta reductions, FP64 coarse solve) to
 * relative tolerance inner_tol, x += d in FP64; convergence is decided on the FP64 residual.
 * The residual is checked at iterations < check_every and at multiples of check_every
 * (rb/gpusolve.py:244 uses 5); check_every < 1 checks every iteration.  Modes 0 and 1 confirm a converged recurrence with the FP64 true
 * residual and, if that is still above tol (stiff systems), restart the recurrence from it
 * (residual replacement) in flexible (Polak-Ribiere) form; the reference stops on the recurrence
 * alone.  info->outer counts
 * refinement steps (mode 2) or recurrence starts (modes 0, 1); info->relres is the true
 * residual.  A restart that twice fails to halve the true residual ends the solve with
 * flag = iterations done (not converged). */
int cm_sys_solve(cm_sys* sys, int mode, const double* b, double* x, double tol, int maxiter, int check_every,
                 double inner_tol, cm_solve_info* info);

/* ---- batched solves in lock-step ----------------------------------------------------------------
 * Replaces the reference's one-solve-per-process batch (rb/gpusolve.py:174-180,
 * exp/rb_run_engine_v4.py:113-146) on one GPU: systems in a batch are solved together by the PCG
 * of cm_sys_solve mode 1 (FP64 PCG, FP32 V-cycle), every launch of an iteration covering all of
 * them.  A system joins with cm_batch_add (after its build, on any stream; its stream waits for
 * the batch while it is in) and leaves when converged, stopped or at maxiter; cm_batch_poll hands
 * back its cm_solve_info.  Systems whose launch structure differs from the batch's (level count,
 * springs on a coarse level) or that carry a rigid-mode correction (cm_sys_set_rigid) are not
 * added: *slot = -1, solve them with cm_sys_solve.  The
 * residual is checked every check_every iterations of the batch. */
typedef struct cm_batch cm_batch;
/* Everything a mode-1 solve needs before its first launch (work vectors, FP32 V-cycle data, FP64
 * operator data, the march's block orders), done now on the system's stream; a system joins a
 * batch without a host wait after this (no reference counterpart: the reference builds its
 * hierarchy inside each solve, rb/gpusolve.py:229-236). */
int cm_sys_prepare_solve(cm_sys* sys);
int cm_batch_create(cm_batch** out, void* stream, int check_every);
int cm_batch_destroy(cm_batch* batch);
/* Starts PCG on (sys, b, x); x holds the initial guess.  *slot: the slot id, or -1 (not added). */
int cm_batch_add(cm_batch* batch, cm_sys* sys, const double* b, double* x, double tol, int maxiter, int* slot);
/* Runs iterations until a system has left the batch or max_iters ran; *nlive: systems still in. */
int cm_batch_step(cm_batch* batch, int max_iters, int* nlive);
/* One finished system: its slot and result (*slot = -1: none waiting).  The slot is free again. */
int cm_batch_poll(cm_batch* batch, int* slot, cm_solve_info* info);

/* Cell-centre strain/stress/von Mises (fe/solver.py:257-272, rb/splitfe.py:161-175) for n listed
 * cells.  hx/hy/hz: cell sizes per axis; cells_ref: reference cell ids; nodes8: 8 node ids per
 * listed cell or NULL for the grid's own corners; lam_ref/mu_ref indexed by reference cell id;
 * U: nodal displacement (3 per node, NaN -> 0); outputs sig[n*6], vm[n], eps[n*6].  All device. */
int cm_cell_stress(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                   int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* lam_ref,
                   const double* mu_ref, const double* U, double* sig, double* vm, double* eps);

/* cm_cell_stress for the z-orthotropic materials of fe/solver.py:43-74 (the coeffs of solve and
 * cell_stress, fe/solver.py:153-154,267-269): coeffs_ref[ncell*6] (device) in reference cell order,
 * stress sig = D eps with D = sum_k c_k P_k. */
int cm_cell_stress_ortho(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                         int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* coeffs_ref,
                         const double* U, double* sig, double* vm, double* eps);

/* Nodal element forces K_e u_e (rb/splitfe.py:178-193) of n listed cells, out[n*24] in element
 * DOF order 3*l + c; arguments as cm_cell_stress, sri as cm_sys_create.  All device. */
int cm_element_forces(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                      int sri, int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* lam_ref,
                      const double* mu_ref, const double* U, double* out);

/* Average kernel time of nrep launches of one level's operator, measured with HIP events on the
 * system's stream: structured_only = 0 the whole product (structured stencil, sparse pass,
 * fix-up), 1 the structured stencil kernel alone, 2 the structured kernel of one fused Chebyshev
 * smoothing step (r = b - A u, d = c1 u + c2 Dinv r, x += u + d on the level's work vectors; y is
 * not used). */
int cm_sys_time_apply(cm_sys* sys, int level, int prec, const void* u, void* y, int nrep, int structured_only,
                      double* avg_ms);

/* Tile autotuner: for every level, time the structured stencil kernel in precision `prec` for
 * z extents 1..32 (nrep launches each, HIP events) and keep the fastest, separately for products
 * (timed on K u) and fused smoothing steps (timed on one step; FP32 always, FP64 when the level
 * has its FP64 smoother, else the product's extent).  Results are cached process-wide per level
 * shape and kernel variant; a cached level is not timed again (nrep < 0: timed again with -nrep
 * launches, replacing the cached extent).  The GPU should be otherwise idle.
 * No reference counterpart (project autotuner). */
int cm_sys_autotune(cm_sys* sys, int prec, int nrep);

/* Set the z extent of one level's structured-march blocks (prec 0 = FP64, 1 = FP32; smooth = 1 for
 * fused smoothing steps, 0 for products and residuals; 0 = the heuristic), e.g. from a tuning table
 * the autotuner wrote earlier.  No reference counterpart. */
int cm_sys_set_tuned_kz(cm_sys* sys, int level, int prec, int smooth, int kz);

/* Minimum number of structured-march blocks per launch the z-extent heuristic aims for (default
 * 600; fewer, longer blocks suit a system that shares the GPU with concurrent solves).  Levels with
 * a tuned extent keep it.  No reference counterpart. */
int cm_sys_set_march_blocks(cm_sys* sys, int min_blocks);

/* Build stage times in ms: level-0 preparation, hierarchy structure, uploads, coarse moments,
 * diagonal blocks, coarsest inverse, total, then the structure stage's used-mask and
 * extra-stiffness coarsening parts (up to cap values, at most 9). */
int cm_sys_timings(cm_sys* sys, double* out, int cap);

/* Engine preparation on the GPU (all arrays device memory, reference cell order
 * c = (i ny + j) nz + k).
 *
 * Lattice materials: rb/rbfe.py:75-134 (materials: class fractions, mixing rule, joint cells,
 * activity), rb/replica3.py:17-29 (bone_share) and rb/rbfe.py:140-141 (part) in one pass.
 * L: [ncell][ns3] body labels (int16); lut_cls / lut_grp: [nlab] class and bone group of every
 * label (rb/rbgeom.py:45 classes, rb/replica2.py:34-46 groups); lam_k / mu_k: Lame parameters of
 * the 10 classes.  Outputs E, nu [ncell] (FP64), active, ser [ncell] (0/1), F [ncell][10] class
 * fractions, S [ncell][ngrp] bone shares (float), part [ncell]. */
int cm_lattice_materials(void* stream, int nx, int ny, int nz, int ns3, const int16_t* L, int nlab,
                         const int8_t* lut_cls, const int8_t* lut_grp, int ngrp, const double* lam_k,
                         const double* mu_k, double* E, double* nu, uint8_t* active, double* F, uint8_t* ser, float* S,
                         int32_t* part);

/* Body fractions of the lattice cells for several body sets at once (the bf(bodies) =
 * np.isin(L, bodies).mean(1) of rb/rbfe.py:138 and rb/rbsim.py:168, used by the region labels of
 * rb/rbfe.py:191-201 and the effect descriptors of rb/rbsim.py:164-203).  mask[b] (b < nlab): bit s
 * set when body b is in set s; out[s][c] = (sub-samples of cell c whose body is in set s) / ns3
 * (FP64, bit-exact to the numpy mean).  nset <= 16. */
int cm_body_fractions(void* stream, int64_t ncell, int ns3, const int16_t* L, int nlab, const uint32_t* mask, int nset,
                      double* out);

/* Node sets of the bone parts (rb/rbfe.py:166-169,178: np.unique of the nodes of the active cells
 * of each part): node_mask[n] gets bit p for every active cell of part p (1..63) that uses node n;
 * part_cells[p] = active cells of part p (64 entries).  cn: (ncell, 8) int64 in reference cell
 * order.  Device memory; the call clears both outputs first. */
int cm_part_nodes(void* stream, int64_t ncell, const int64_t* cn, const int32_t* part, const uint8_t* active,
                  int64_t nnode, uint64_t* node_mask, int32_t* part_cells);

/* Squared distance from each point to the nearest node of each bone (rb/rbfe.py:173-175, one KD
 * tree per bone there): d2min[e * nbone + b]; bone b's nodes are bnodes[bptr[b] .. bptr[b+1]),
 * placed at their original's grid position (orig, null = identity) on edges xe, ye, ze.  The
 * squares are summed as (dx^2 + dy^2) + dz^2 with each operation rounded once. */
int cm_nearest_bone(void* stream, int npt, const double* pts, int nbone, const int64_t* bptr, const int64_t* bnodes,
                    int nx, int ny, const double* xe, const double* ye, const double* ze, const int64_t* orig,
                    double* d2min);

/* Ligament patches (rb/replica3.py:68-86): for spring end e at pts[e] on bone bone[e], the bone's
 * used nodes (used null: all) within radius[e] = max(4, distance of the min(7, count - 1)-th nearest)
 * get flags[foff[e] + q] = 1 (q = the node's position in the bone's list); count[e] = used
 * candidates (0: radius -1, no flag set). */
int cm_spring_patches(void* stream, int nend, const double* pts, const int32_t* bone, const int64_t* bptr,
                      const int64_t* bnodes, const uint8_t* used, int nx, int ny, const double* xe, const double* ye,
                      const double* ze, const int64_t* orig, const int64_t* foff, uint8_t* flags, double* radius,
                      int32_t* count);

/* Split connectivity: rb/splitfe.py:261-305 (split_connectivity_fast).  active [ncell] (0/1),
 * part [ncell] (bone per cell, 0 = none), share [ncell][nb] (float).  Outputs cn [ncell][8]
 * (node ids, copies at nnode + rank), copy_orig [*ncopy] (original node of every copy; fails
 * with the needed size when it exceeds cap), *ncopy, *nsplit (host). */
int cm_split_connectivity(void* stream, int nx, int ny, int nz, const uint8_t* active, const int32_t* part,
                          const float* share, int nb, int64_t* cn, int64_t* copy_orig, int64_t cap, int64_t* ncopy,
                          int64_t* nsplit);

/* Bank record of one action (exp/bank.py:115-121): vm16 = FP16 von Mises [ncell], uc16 = FP16
 * mean relative displacement of every analysis cell over its 8 nodes [ncell][3] (u_nodes
 * [nnode][3] FP32, cell_nodes [ncell][8]), lab8 = int8 labels, frac8 = round(255 frac) as uint8
 * [ncell][ncls].  All device memory. */
int cm_bank_pack(void* stream, int64_t ncell, int ncls, const float* vm, const float* u_nodes, const int64_t* cell_nodes,
                 const int32_t* lab, const float* frac, void* vm16, void* uc16, int8_t* lab8, uint8_t* frac8);

/* Override the z extent of a structured block (0 = automatic). */
int cm_sys_set_kz(cm_sys* sys, int kz);

/* Level-0 smoother blocks over split-copy groups (on by default): a grid node and its split
 * copies, coupled by extra stiffness (interface springs), are smoothed with the inverse of their
 * joint diagonal block instead of one 3x3 block per node; 0 = the per-node 3x3 block Jacobi of
 * fe/gmg.py:29-44.  Same solution to the tolerance; stiff interfaces (fused joints) converge in
 * fewer iterations.  Call before cm_sys_build.  No reference counterpart. */
int cm_sys_set_copy_blocks(cm_sys* sys, int on);

/* Structured-march blocks whose cells have no stiffness run only the epilogue (on by default;
 * 0 = every block runs the full march).  Both give the same results; no reference counterpart. */
int cm_sys_set_block_skip(cm_sys* sys, int on);

/* V-cycle replay through HIP graphs (on by default): cm_sys_solve and cm_sys_vcycle run on the
 * system's own stream (ordered after the caller's stream by events) and replay the captured
 * V-cycle while its buffers stay the same.  on = 0 launches every kernel eagerly.  No reference
 * counterpart (the reference launches eagerly from Python, rb/gpusolve.py:163-171). */
int cm_sys_set_graphs(cm_sys* sys, int on);

int cm_sys_sync(cm_sys* sys);

/* Analysis-grid resampling of rb/rbsim.py:110-140 (sample_fe).  Per axis sample tables (FE cell
 * index, local coordinate, inside flag) for mx*sub, my*sub, mz*sub samples; FE grid nx,ny,nz;
 * active[ncell], cn[ncell*8] (reference cell order), VM[ncell], SIG[ncell*6] (float32),
 * Urel[nodes*3] (FP64).  Outputs per analysis cell (reference order).  All device pointers. */
int cm_sample_fe(void* stream, int mx, int my, int mz, int sub, int nx, int ny, int nz, const int* si,
                 const double* st, const uint8_t* sin_, const int* sj, const double* su, const uint8_t* sjn,
                 const int* sk, const double* sv, const uint8_t* skn, const uint8_t* active, const int64_t* cn,
                 const float* VM, const float* SIG, const double* Urel, float* vm_out, float* u_out, double* cover,
                 float* sig_out);

/* Node resampling of rb/rbsim.py:143-161 (sample_nodes): the relative displacement at the
 * (mx+1)(my+1)(mz+1) analysis nodes (fe/solver.py:120-123 order), trilinear in the nearest active
 * FE cell (scipy.ndimage.distance_transform_edt feature transform of the inactive mask in cell-index
 * units, rb/rbsim.py:152; ties: smallest (k, j, i)).  mxe/mye/mze: analysis node coordinates per
 * axis; ci/cj/ck: their FE cell per axis (clip(searchsorted(edges, x) - 1, 0, n - 1)); kq[nq]: the
 * distinct z cells of ck, qz[mz+1]: each z node's index into kq.  FE grid edges xe/ye/ze,
 * active[ncell], cn[ncell*8], Urel[nodes*3] (FP64).  Work arrays dz2, kbest: nq*nx*ny ints each.
 * out[(mx+1)(my+1)(mz+1)*3] float32.  At least one FE cell must be active.  All device pointers. */
int cm_sample_nodes(void* stream, int mx, int my, int mz, const double* mxe, const double* mye, const double* mze,
                    const int* ci, const int* cj, const int* ck, int nq, const int* kq, const int* qz, int nx, int ny,
                    int nz, const double* xe, const double* ye, const double* ze, const uint8_t* active,
                    const int64_t* cn, const double* Urel, int* dz2, int* kbest, float* out);

/* Allocation-free variant of cm_sample_nodes.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and result as cm_sample_nodes without the work arrays
 * dz2 and kbest: their tables live in a scratch array the library holds per device
 * (2 * CM_SAMPLE_NODES_MAX_COLUMNS ints, 32 MiB, kept until the process ends).  Requires
 * nq * nx * ny <= CM_SAMPLE_NODES_MAX_COLUMNS, else fails with a message; the current device must
 * be the one the stream and buffers belong to.  The call allocates nothing (no hipMalloc, hipFree
 * or allocating library routine).  Calls that overlap on one device share the scratch and race:
 * serialise them (one stream, or events). */
#define CM_SAMPLE_NODES_MAX_COLUMNS 4194304
int cm_sample_nodes_noalloc(void* stream, int mx, int my, int mz, const double* mxe, const double* mye,
                            const double* mze, const int* ci, const int* cj, const int* ck, int nq, const int* kq,
                            const int* qz, int nx, int ny, int nz, const double* xe, const double* ye, const double* ze,
                            const uint8_t* active, const int64_t* cn, const double* Urel, float* out);

#ifdef __cplusplus
}
#endif

#endif /* COMECH_FE_H */
