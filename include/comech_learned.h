/* C ABI of the learned warm start and coarse correction (libcomech_hip.so): the kernels of a small
 * 3-D convolutional network trained and evaluated on the GPU, its raster features, and the maps
 * between its per-cell output and the solver's node vectors.
 *
 * No counterpart in the reference (xtli12/CoMech has no learned component on the solve path).  The
 * network's prediction enters the solver through the reference's own hooks: the initial guess x0
 * of pcg_gmg / pcg_gmg_gpu (fe/gmg.py:120,136, rb/gpusolve.py:229) and the additive coarse
 * correction Z E^-1 Z^T of rb/gpusolve.py:217-228 (cm_sys_set_rigid), with plain GMG-PCG as the
 * fallback.  The host side is comech_amd.learned.
 *
 * Conventions as include/comech_fe.h: 0 on success, -1 on failure with cm_nn_last_error(); every
 * pointer is device memory unless stated; "stream" is a hipStream_t.  BF16 values are stored as
 * uint16 (the upper half of the FP32 pattern, round to nearest even).  Feature maps are
 * channels-last [B * V][C] with V the raster's cells in the reference's z-fastest order
 * (i * ny + j) * nz + k (fe/solver.py:104-105).  A 3x3x3 convolution keeps its BF16 weights as
 * Wt[co][kp], kp = tap * Cin + ci, tap = 9 (di + 1) + 3 (dj + 1) + (dk + 1), zero-padded to Kp
 * columns (a multiple of 32).
 */
#ifndef COMECH_LEARNED_H
#define COMECH_LEARNED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* cm_nn_last_error(void);

/* C[m][n] = alpha sum_k A[m][k] B[n][k] on the BF16 matrix cores with FP32 accumulation; A: M x K
 * (row stride lda), B: N x K (ldb); K, lda, ldb multiples of 32, 8, 8; 16-byte aligned operands.
 * mode 0: C FP32 = act(alpha acc + bias[n] + add[m][n]); mode 1: the same stored as BF16; mode 2:
 * C FP32 += alpha acc with atomics (ksplit > 1 splits K over blocks).  bias (FP32 [N]) and add
 * (BF16 [M][ldadd]) may be null; relu = 1 makes act = max(., 0), else the identity. */
int cm_nn_gemm_tn(void* stream, int M, int N, int K, const void* A, long long lda, const void* B, long long ldb, void* C,
                  long long ldc, const float* bias, const void* add, long long ldadd, int relu, int mode, int ksplit,
                  float alpha);

/* im2col of a batch of B rasters nx x ny x nz with C channels (C multiple of 8) for a 3x3x3
 * convolution with zero padding: col[m][kp] (Kp >= 27 C, multiple of 32), or transposed = 1:
 * colT[kp][m] with Mp >= B V rows per kp (multiple of 8; rows beyond B V are zero). */
int cm_nn_im2col(void* stream, int B, int nx, int ny, int nz, int C, int Kp, const void* X, void* col, int transposed,
                 long long Mp);

/* out[c][m] = in[m][c] as BF16 for m < M, zero for M <= m < Mp; the source is BF16 (in_bf16) or
 * FP32 (in_f32, rounded), exactly one of them non-null. */
int cm_nn_transpose(void* stream, long long M, long long Mp, int C, const void* in_bf16, const float* in_f32, void* out);

/* BF16 compute weights from FP32 master weights W[co][kp] (Coutp rows, Kp columns, zero beyond the
 * layer's Cout outputs and 27 Cin columns): Wt = W rounded to BF16, and (Wf non-null) the flipped
 * transpose for the input gradient, Wf[ci][kpf] = W[co][(26 - tap) Cin + ci] at kpf = tap Coutp +
 * co, Kpf >= 27 Coutp. */
int cm_nn_weights(void* stream, int Cin, int Coutp, int Kp, int Kpf, const float* W, void* Wt, void* Wf);

/* Adam step `step` (from 1) on n FP32 master weights w with moments m, v and gradient g. */
int cm_nn_adam(void* stream, long long n, float* w, float* m, float* v, const float* g, float lr, float b1, float b2,
               float eps, int step);

/* out[n] += sum over m < M of g[m][n] (g FP32 [M][N], N <= 256): the bias gradient. */
int cm_nn_colsum(void* stream, long long M, int N, const float* g, float* out);

/* g = dy where y > 0, else 0 (y: the BF16 post-ReLU activation, null for a linear layer);
 * written as BF16 (g_bf16) and/or FP32 (g_f32). */
int cm_nn_relu_bwd(void* stream, long long n, const float* dy, const void* y, void* g_bf16, float* g_f32);

/* 2x2x2 average pooling of BF16 maps (even sizes) and its adjoint accumulated into FP32 dx. */
int cm_nn_pool(void* stream, int B, int nx, int ny, int nz, int C, const void* x, void* y);
int cm_nn_pool_bwd(void* stream, int B, int nx, int ny, int nz, int C, const float* dy, float* dx);

/* Nearest upsampling by 2 of BF16 maps (coarse size cx, cy, cz) and its adjoint accumulated into
 * FP32 dx. */
int cm_nn_up(void* stream, int B, int cx, int cy, int cz, int C, const void* x, void* y);
int cm_nn_up_bwd(void* stream, int B, int cx, int cy, int cz, int C, const float* dy, float* dx);

/* Masked squared error of the first nc channels of out [M][ldo] against t [M][nc]: per 256-row
 * block the sum of mask (out - t)^2 into part[block] (FP64), and the gradient
 * d = 2 mask (out - t) inv_norm (FP32 [M][ldo], other channels zero). */
int cm_nn_mse(void* stream, long long M, int nc, int ldo, const float* out, const float* t, const uint8_t* mask,
              float inv_norm, float* d, double* part);

/* The network's 16 input channels per cell (BF16 [V][16]) from the solver's level-0 description:
 * active, the one-hot material class cls[label] (4 classes), (nu[label] - 0.3) / 0.15, split
 * cells x kif, and the sums over the cell's 8 corners of per-grid-node ground-spring weights, ligament
 * ends and downward load (x load_scale), then the normalised cell centre. lab: uint8 labels,
 * 255 inactive; cls: int8 per label; nu: FP32 per label; splitcell: uint8 per cell. */
int cm_nn_features(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int8_t* cls, const float* nu,
                   const uint8_t* splitcell, float kif, const float* ground, const float* lig, const float* load,
                   float load_scale, void* F);

/* Cell targets: the mean of the cell's 8 corner displacements (u: FP64, 3 per node; the corners
 * of cells with irr_of[v] >= 0 come from irr_nodes[8 irr_of[v] ..]) x scale, FP32 [V][3]; zero for
 * inactive cells. */
int cm_nn_node_to_cell(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int* irr_of,
                       const long long* irr_nodes, const double* u, float scale, float* t);

/* Node vector from per-cell predictions (first 3 of ldp channels): every node takes scale x the
 * mean over the active cells that use it, 0 where none does.  acc: FP32 work [nn][4]; x: FP64 [nn][3]. */
int cm_nn_cell_to_node(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int* irr_of,
                       const long long* irr_nodes, const float* pred, int ldp, long long nn, float* acc, double scale,
                       double* x);

/* Galerkin coarse space of a prediction (the learned initial guess x0 = Z (Z^T A Z)^-1 Z^T b, the
 * form of the additive correction of rb/gpusolve.py:217-228).  Node groups: group[n] = -1 unused, 0
 * soft tissue (nodes of active part-0 cells only), g >= 1 bone part g (part: uint8 per cell).  Z
 * has 3 + 6 ngroup columns: the prediction's x, y, z components on soft nodes (pred: FP64 [nn][3]),
 * then per bone g the rigid modes of fe/solver.py:177-199 (3 translations, 3 rotations of
 * r = (p - centre_g) / 20; p: the node's grid position, a split copy at its original node, from
 * the edges xe, ye, ze; copy_orig[nn - nnode]). */
int cm_nn_node_groups(void* stream, int nx, int ny, int nz, const uint8_t* lab, const uint8_t* part, const int* irr_of,
                      const long long* irr_nodes, long long nn, int* group);
/* acc[ngroup][4]: sums of the positions of each bone group's nodes and their count. */
int cm_nn_group_centres(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
                        const double* xe, const double* ye, const double* ze, int ngroup, double* acc);
/* out[3 + 6 ngroup] += Z^T v (v: FP64 [nn][3]); ngroup <= 40. */
int cm_nn_zt(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
             const double* xe, const double* ye, const double* ze, int ngroup, const double* centre, const double* pred,
             const double* v, double* out);
/* x = Z c (FP64 [nn][3]). */
int cm_nn_zc(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
             const double* xe, const double* ye, const double* ze, int ngroup, const double* centre, const double* pred,
             const double* c, double* x);

/* Small vector helpers: fill, FP32 <-> BF16 casts (to_bf16 = 1: f32 -> bf16), y += a x. */
int cm_nn_fill(void* stream, long long n, float* p, float v);
int cm_nn_cast(void* stream, long long n, void* f32, void* bf16, int to_bf16);
int cm_nn_axpy(void* stream, long long n, float a, const float* x, float* y);

#ifdef __cplusplus
}
#endif

#endif /* COMECH_LEARNED_H */
