// Learned warm start and coarse correction: the kernels of a small 3-D convolutional network
// (3x3x3 convolutions, 2x2x2 average pooling, nearest upsampling, additive skips) trained and
// evaluated on the GPU, plus the raster features it reads and the cell <-> node maps between its
// output and the solver's node vectors.  No counterpart in the reference; the prediction enters the
// solver where the reference takes an initial guess (x0 of rb/gpusolve.py:229, fe/gmg.py:120,136)
// and an additive coarse correction (Z of rb/gpusolve.py:217-228).
//
// Layouts.  Feature maps are channels-last: row m = b * V + v of a batch of B rasters, v the cell
// id (i * ny + j) * nz + k (the reference's z-fastest order, fe/solver.py:104-105), C contiguous
// channels per row; BF16 (stored as uint16) for activations, FP32 for gradients.  A convolution
// with Cin inputs and Cout outputs keeps its weights as Wt[co][kp] (BF16), kp = tap * Cin + ci with
// tap = 9 (di + 1) + 3 (dj + 1) + (dk + 1), zero-padded to Kp = 32 * ceil(27 Cin / 32) columns.
//
// Every matrix product is one kernel, k_gemm_tn: C[m][n] = sum_k A[m][k] B[n][k] with both
// operands K-contiguous (rows 16-byte aligned), on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation).  Forward: A = im2col(X) [M][Kp], B = Wt.  Backward to the input: the same
// convolution of the output gradient with the flipped, transposed weights.  Backward to the
// weights: A = g^T [Cout][M], B = im2col(X)^T [Kp][M], split over M with FP32 atomics.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/comech_learned.h"
#include "common.hpp"

namespace cm {
namespace nn {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// BF16 <-> FP32 (round to nearest even; NaN stays NaN)
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// ---- matrix product on the matrix cores ------------------------------------------------------
// Block: 4 waves in a 2 x 2 arrangement; a wave computes TM x TN tiles of 16 x 16 outputs.  Lane
// l of a wave feeds row (l % 16) of an A tile and column (l % 16) of a B tile with the 8
// consecutive k values 8 (l / 16) .. 8 (l / 16) + 7 of the current 32-wide K step (the same k for
// both operands), and holds outputs C[4 (l / 16) + r][l % 16], r = 0..3, of each tile.
// Epilogue: y = act(alpha acc (+ bias[n]) (+ add[m][n])); FP32 or BF16 store, or FP32 atomic
// accumulation (split K: blockIdx.z takes K range [z kc, (z + 1) kc)).
enum { EPI_F32 = 0, EPI_BF16 = 1, EPI_ATOMIC = 2 };

struct GemmArgs {
  int M, N, K;                 // K: multiple of 32
  const uint16_t* A; long long lda;
  const uint16_t* B; long long ldb;
  void* C; long long ldc;
  const float* bias;           // [N] or null
  const uint16_t* add; long long ldadd;   // BF16 [M][ldadd] added before the activation, or null
  int relu, mode, kc;          // kc: K per split (multiple of 32)
  float alpha;
};

template <int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm_tn(GemmArgs g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long long m0 = (long long)blockIdx.x * (32 * TM) + wm * (16 * TM);
  const int n0 = blockIdx.y * (32 * TN) + wn * (16 * TN);
  const int kb = blockIdx.z * g.kc, ke = min(g.K, kb + g.kc);
  const int r16 = lane & 15, kq = 8 * (lane >> 4);
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint16_t* arow[TM];
  const uint16_t* brow[TN];
  bool aok[TM], bok[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const long long m = m0 + 16 * i + r16;
    aok[i] = m < g.M;
    arow[i] = g.A + (aok[i] ? m : 0) * g.lda + kq;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int n = n0 + 16 * j + r16;
    bok[j] = n < g.N;
    brow[j] = g.B + (long long)(bok[j] ? n : 0) * g.ldb + kq;
  }
  const bf16x8 zero = {};
  for (int k = kb; k < ke; k += 32) {
    bf16x8 a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) a[i] = aok[i] ? *reinterpret_cast<const bf16x8*>(arow[i] + k) : zero;
#pragma unroll
    for (int j = 0; j < TN; ++j) b[j] = bok[j] ? *reinterpret_cast<const bf16x8*>(brow[j] + k) : zero;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + 16 * j + r16;
      if (n >= g.N) continue;
      const float bn = g.bias ? g.bias[n] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long m = m0 + 16 * i + 4 * (lane >> 4) + r;
        if (m >= g.M) continue;
        float v = g.alpha * acc[i][j][r];
        if (g.mode == EPI_ATOMIC) {
          atomicAdd(static_cast<float*>(g.C) + m * g.ldc + n, v);
          continue;
        }
        v += bn;
        if (g.add) v += bf2f(g.add[m * g.ldadd + n]);
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.mode == EPI_F32) static_cast<float*>(g.C)[m * g.ldc + n] = v;
        else static_cast<uint16_t*>(g.C)[m * g.ldc + n] = f2bf(v);
      }
    }
}

// ---- im2col ----------------------------------------------------------------------------------
// col[m][kp], kp = tap * C + c (8 channels per thread, 16-byte copies), zero outside the raster
// and in the padding columns 27 C <= kp < Kp.  transposed = 1 writes colT[kp][m] instead.
__global__ void k_im2col(int B, int nx, int ny, int nz, int C, int Kp, const uint16_t* __restrict__ X,
                         uint16_t* __restrict__ col) {
  const long long V = (long long)nx * ny * nz;
  const int g8 = Kp / 8;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= B * V * g8) return;
  const long long m = id / g8;
  const int q = (int)(id - m * g8), kp = 8 * q;
  uint4 val = make_uint4(0, 0, 0, 0);
  if (kp < 27 * C) {
    const int tap = kp / C, c = kp - tap * C;
    const long long b = m / V, v = m - b * V;
    const int i = (int)(v / ((long long)ny * nz)), j = (int)((v / nz) % ny), k = (int)(v % nz);
    const int ii = i + tap / 9 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap % 3 - 1;
    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz)
      val = *reinterpret_cast<const uint4*>(X + (b * V + ((long long)ii * ny + jj) * nz + kk) * C + c);
  }
  *reinterpret_cast<uint4*>(col + m * Kp + kp) = val;
}

// colT[kp][m] for m in [0, Mp) (Mp: M rounded up to 8; rows beyond M are zero), one thread per
// (kp, 8 consecutive m).
__global__ void k_im2col_t(int B, int nx, int ny, int nz, int C, int Kp, long long Mp, const uint16_t* __restrict__ X,
                           uint16_t* __restrict__ colT) {
  const long long V = (long long)nx * ny * nz, M = B * V;
  const long long g8 = Mp / 8;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)Kp * g8) return;
  const int kp = (int)(id / g8);
  const long long m0 = 8 * (id - (long long)kp * g8);
  uint16_t out[8];
  const bool real = kp < 27 * C;
  const int tap = real ? kp / C : 0, c = real ? kp - tap * C : 0;
  const int di = tap / 9 - 1, dj = (tap / 3) % 3 - 1, dk = tap % 3 - 1;
  for (int q = 0; q < 8; ++q) {
    const long long m = m0 + q;
    uint16_t x = 0;
    if (real && m < M) {
      const long long b = m / V, v = m - b * V;
      const int i = (int)(v / ((long long)ny * nz)) + di, j = (int)((v / nz) % ny) + dj, k = (int)(v % nz) + dk;
      if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz)
        x = X[(b * V + ((long long)i * ny + j) * nz + k) * C + c];
    }
    out[q] = x;
  }
  *reinterpret_cast<uint4*>(colT + (long long)kp * Mp + m0) = *reinterpret_cast<const uint4*>(out);
}

// out[c][m] = in[m][c] for m < M (BF16 source, BF16 result; FP32 source converted), rows
// padded with zeros to Mp
__global__ void k_transpose_bf(long long M, long long Mp, int C, const uint16_t* __restrict__ in,
                               const float* __restrict__ in32, uint16_t* __restrict__ out) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= Mp * C) return;
  const int c = (int)(id / Mp);
  const long long m = id - (long long)c * Mp;
  uint16_t v = 0;
  if (m < M) v = in ? in[m * C + c] : f2bf(in32[m * C + c]);
  out[id] = v;
}

// ---- weights ---------------------------------------------------------------------------------
// BF16 compute copies of the FP32 master weights W[co][kp] (Coutp rows, Kp columns, zero padding):
// Wt = W rounded, and the flipped transpose used for the input gradient, Wf[ci][kpf],
// kpf = tap' * Coutp + co, Wf[ci][(26 - tap) Coutp + co] = W[co][tap Cin + ci].
__global__ void k_weights_bf16(int Cin, int Coutp, int Kp, int Kpf, const float* __restrict__ W,
                               uint16_t* __restrict__ Wt, uint16_t* __restrict__ Wf) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nt = (long long)Coutp * Kp, nf = (long long)Cin * Kpf;
  if (id < nt) {
    Wt[id] = f2bf(W[id]);
  } else if (id < nt + nf && Wf) {
    const long long q = id - nt;
    const int ci = (int)(q / Kpf), kpf = (int)(q - (long long)ci * Kpf);
    uint16_t v = 0;
    if (kpf < 27 * Coutp) {
      const int tp = kpf / Coutp, co = kpf - tp * Coutp;
      v = f2bf(W[(long long)co * Kp + (26 - tp) * Cin + ci]);
    }
    Wf[q] = v;
  }
}

// Adam (Kingma and Ba 2015) on FP32 master weights: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
__global__ void k_adam(long long n, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// column sums (bias gradient): out[n] += sum_m g[m][n], g FP32 [M][N]; one block per 256 rows
__global__ void k_colsum(long long M, int N, const float* __restrict__ g, float* __restrict__ out) {
  const int n = threadIdx.x;
  if (n >= N) return;
  const long long m0 = (long long)blockIdx.x * 256, m1 = min(M, m0 + 256);
  float s = 0.f;
  for (long long m = m0; m < m1; ++m) s += g[m * N + n];
  atomicAdd(out + n, s);
}

// ---- elementwise -----------------------------------------------------------------------------
// ReLU backward to a BF16 GEMM operand: g = dy (y > 0), y the post-activation output
__global__ void k_relu_bwd(long long n, const float* __restrict__ dy, const uint16_t* __restrict__ y,
                           uint16_t* __restrict__ g, float* __restrict__ g32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = (!y || bf2f(y[i]) > 0.f) ? dy[i] : 0.f;
  if (g) g[i] = f2bf(v);
  if (g32) g32[i] = v;
}

// 2x2x2 average pooling (BF16) and its adjoint (FP32, accumulated into dx)
__global__ void k_pool(int B, int nx, int ny, int nz, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long Vc = (long long)cx * cy * cz;
  if (id >= B * Vc * C) return;
  const int c = (int)(id % C);
  const long long mc = id / C, b = mc / Vc, v = mc - b * Vc;
  const int i = (int)(v / ((long long)cy * cz)), j = (int)((v / cz) % cy), k = (int)(v % cz);
  const long long V = (long long)nx * ny * nz;
  float s = 0.f;
  for (int q = 0; q < 8; ++q) {
    const long long f = ((long long)(2 * i + (q & 1)) * ny + 2 * j + ((q >> 1) & 1)) * nz + 2 * k + (q >> 2);
    s += bf2f(x[(b * V + f) * C + c]);
  }
  y[id] = f2bf(0.125f * s);
}
__global__ void k_pool_bwd(int B, int nx, int ny, int nz, int C, const float* __restrict__ dy, float* __restrict__ dx) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long V = (long long)nx * ny * nz;
  if (id >= B * V * C) return;
  const int c = (int)(id % C);
  const long long m = id / C, b = m / V, v = m - b * V;
  const int i = (int)(v / ((long long)ny * nz)), j = (int)((v / nz) % ny), k = (int)(v % nz);
  const int cy = ny / 2, cz = nz / 2;
  const long long Vc = (long long)(nx / 2) * cy * cz;
  dx[id] += 0.125f * dy[(b * Vc + ((long long)(i / 2) * cy + j / 2) * cz + k / 2) * C + c];
}
// nearest upsampling by 2 (BF16) and its adjoint (FP32: the sum of the 8 fine gradients)
__global__ void k_up(int B, int cx, int cy, int cz, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int nx = 2 * cx, ny = 2 * cy, nz = 2 * cz;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long V = (long long)nx * ny * nz;
  if (id >= B * V * C) return;
  const int c = (int)(id % C);
  const long long m = id / C, b = m / V, v = m - b * V;
  const int i = (int)(v / ((long long)ny * nz)), j = (int)((v / nz) % ny), k = (int)(v % nz);
  const long long Vc = (long long)cx * cy * cz;
  y[id] = x[(b * Vc + ((long long)(i / 2) * cy + j / 2) * cz + k / 2) * C + c];
}
__global__ void k_up_bwd(int B, int cx, int cy, int cz, int C, const float* __restrict__ dy, float* __restrict__ dx) {
  const int ny = 2 * cy, nz = 2 * cz;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long Vc = (long long)cx * cy * cz;
  if (id >= B * Vc * C) return;
  const int c = (int)(id % C);
  const long long mc = id / C, b = mc / Vc, v = mc - b * Vc;
  const int i = (int)(v / ((long long)cy * cz)), j = (int)((v / cz) % cy), k = (int)(v % cz);
  const long long V = 8 * Vc;
  float s = 0.f;
  for (int q = 0; q < 8; ++q) {
    const long long f = ((long long)(2 * i + (q & 1)) * ny + 2 * j + ((q >> 1) & 1)) * nz + 2 * k + (q >> 2);
    s += dy[(b * V + f) * C + c];
  }
  dx[id] += s;
}

// masked mean-squared error over the first nc channels of out [M][ldo] against t [M][nc], rows
// weighted by mask [M] (0/1): partial sums of w (out - t)^2 into part[block], and the gradient
// d = 2 w (out - t) / norm in the same layout as out (other channels 0)
__global__ __launch_bounds__(256) void k_mse(long long M, int nc, int ldo, const float* __restrict__ out,
                                             const float* __restrict__ t, const uint8_t* __restrict__ mask, float inv_norm,
                                             float* __restrict__ d, double* __restrict__ part) {
  __shared__ double red[256];
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (m < M) {
    const float w = mask[m] ? 1.f : 0.f;
    for (int c = 0; c < ldo; ++c) {
      float gv = 0.f;
      if (c < nc) {
        const float e = out[m * ldo + c] - t[m * nc + c];
        s += (double)(w * e * e);
        gv = 2.f * w * e * inv_norm;
      }
      d[m * ldo + c] = gv;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---- raster features and cell <-> node maps ----------------------------------------------------
// Input channels per cell (16, BF16): 0 active; 1-4 one-hot material class of the cell's label
// (cls[label] in 0..3); 5 (nu - 0.3) / 0.15; 6 split cell x log10(k_iface) / 6; 7 ground-spring
// weight of the cell's corners; 8 ligament ends among the corners; 9 downward load on the corners
// (per node, scaled by load_scale); 10-12 cell centre coordinates / extent; 13-15 zero.  Node
// quantities (ground, lig, load) are per grid node; a split copy counts at its original node.
__global__ void k_features(int nx, int ny, int nz, const uint8_t* __restrict__ lab, const int8_t* __restrict__ cls,
                           const float* __restrict__ nu, const uint8_t* __restrict__ splitcell, float kif,
                           const float* __restrict__ ground, const float* __restrict__ lig, const float* __restrict__ load,
                           float load_scale, uint16_t* __restrict__ F) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long V = (long long)nx * ny * nz;
  if (v >= V) return;
  const int i = (int)(v / ((long long)ny * nz)), j = (int)((v / nz) % ny), k = (int)(v % nz);
  float f[16];
  for (int c = 0; c < 16; ++c) f[c] = 0.f;
  const int L = lab[v];
  if (L != 255) {
    f[0] = 1.f;
    const int q = cls[L];
    if (q >= 0 && q < 4) f[1 + q] = 1.f;
    f[5] = (nu[L] - 0.3f) / 0.15f;
    f[6] = splitcell[v] ? kif : 0.f;
    float gs = 0.f, ls = 0.f, fl = 0.f;
    for (int c = 0; c < 8; ++c) {
      const long long n = (i + (c & 1)) + (long long)(nx + 1) * ((j + ((c >> 1) & 1)) + (long long)(ny + 1) * (k + (c >> 2)));
      gs += ground[n];
      ls += lig[n];
      fl += load[n];
    }
    f[7] = gs;
    f[8] = ls;
    f[9] = fl * load_scale;
  }
  f[10] = (i + 0.5f) / nx;
  f[11] = (j + 0.5f) / ny;
  f[12] = (k + 0.5f) / nz;
  for (int c = 0; c < 16; ++c) F[v * 16 + c] = f2bf(f[c]);
}

// corner node ids of cell v: the grid's own, or the irregular cell's list
__device__ __forceinline__ long long cell_node(int nx, int ny, int nz, long long v, int c, const int* __restrict__ irr_of,
                                               const long long* __restrict__ irr_nodes) {
  const int q = irr_of[v];
  if (q >= 0) return irr_nodes[8LL * q + c];
  const int i = (int)(v / ((long long)ny * nz)), j = (int)((v / nz) % ny), k = (int)(v % nz);
  return (i + (c & 1)) + (long long)(nx + 1) * ((j + ((c >> 1) & 1)) + (long long)(ny + 1) * (k + (c >> 2)));
}

// cell targets: the mean of the 8 corner displacements (node vector u, 3 per node) x scale
__global__ void k_node_to_cell(int nx, int ny, int nz, const uint8_t* __restrict__ lab, const int* __restrict__ irr_of,
                               const long long* __restrict__ irr_nodes, const double* __restrict__ u, float scale,
                               float* __restrict__ t) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (long long)nx * ny * nz) return;
  double s[3] = {0.0, 0.0, 0.0};
  if (lab[v] != 255) {
    for (int c = 0; c < 8; ++c) {
      const long long n = cell_node(nx, ny, nz, v, c, irr_of, irr_nodes);
      for (int d = 0; d < 3; ++d) s[d] += u[3 * n + d];
    }
  }
  for (int d = 0; d < 3; ++d) t[3 * v + d] = (float)(s[d] * 0.125 * scale);
}

// node vector from cell predictions: every used node takes the mean of the active cells around it
// (pred [V][ldp], first 3 channels, x scale); acc [nn][4] work (3 sums + count), cleared first
__global__ void k_cell_to_node_acc(int nx, int ny, int nz, const uint8_t* __restrict__ lab, const int* __restrict__ irr_of,
                                   const long long* __restrict__ irr_nodes, const float* __restrict__ pred, int ldp,
                                   float* __restrict__ acc) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (long long)nx * ny * nz || lab[v] == 255) return;
  for (int c = 0; c < 8; ++c) {
    const long long n = cell_node(nx, ny, nz, v, c, irr_of, irr_nodes);
    for (int d = 0; d < 3; ++d) atomicAdd(acc + 4 * n + d, pred[v * ldp + d]);
    atomicAdd(acc + 4 * n + 3, 1.f);
  }
}
__global__ void k_cell_to_node_fin(long long nn, const float* __restrict__ acc, double scale, double* __restrict__ x) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const float cnt = acc[4 * n + 3];
  for (int d = 0; d < 3; ++d) x[3 * n + d] = cnt > 0.f ? scale * (double)(acc[4 * n + d] / cnt) : 0.0;
}

// ---- Galerkin coarse space of the prediction ------------------------------------------------
// Node groups: -1 unused, 0 soft tissue (a node of active part-0 cells only), g >= 1 bone part g
// (a node of a cell of that part; split copies belong to one bone).  Columns of Z: 0..2 the
// prediction's x, y, z components on soft nodes; 3 + 6 (g - 1) + q the rigid modes of bone g
// (translations x, y, z, then rotations about x, y, z of r = (p - centre_g) / 20, the modes of
// fe/solver.py:177-199).
__global__ void k_node_groups(int nx, int ny, int nz, const uint8_t* __restrict__ lab, const uint8_t* __restrict__ part,
                              const int* __restrict__ irr_of, const long long* __restrict__ irr_nodes,
                              int* __restrict__ group) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (long long)nx * ny * nz || lab[v] == 255) return;
  const int p = part[v];
  for (int c = 0; c < 8; ++c) atomicMax(group + cell_node(nx, ny, nz, v, c, irr_of, irr_nodes), p);
}

// grid position of level-0 node n (a split copy sits at its original node)
__device__ __forceinline__ void node_pos(long long n, long long nnode, const long long* __restrict__ copy_orig, int nx,
                                         int ny, const double* __restrict__ xe, const double* __restrict__ ye,
                                         const double* __restrict__ ze, double p[3]) {
  const long long g = n < nnode ? n : copy_orig[n - nnode];
  p[0] = xe[g % (nx + 1)];
  p[1] = ye[(g / (nx + 1)) % (ny + 1)];
  p[2] = ze[g / ((long long)(nx + 1) * (ny + 1))];
}

struct Groups {
  const int* group;                  // [nn]
  const long long* copy_orig;        // [nn - nnode]
  const double *xe, *ye, *ze;
  long long nn, nnode;
  int nx, ny, ngroup;                // ngroup: bone groups (g = 1..ngroup)
  const double* centre;              // [ngroup][3]
  const double* pred;                // [nn][3] FP64 prediction
};

__device__ __forceinline__ void zrow(const Groups& G, long long n, int& g, double z[6][3]) {
  g = G.group[n];
  if (g < 0) return;
  if (g == 0) {
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) z[c][d] = c == d ? G.pred[3 * n + d] : 0.0;
    return;
  }
  double p[3];
  node_pos(n, G.nnode, G.copy_orig, G.nx, G.ny, G.xe, G.ye, G.ze, p);
  const double* cg = G.centre + 3 * (g - 1);
  const double r0 = (p[0] - cg[0]) / 20.0, r1 = (p[1] - cg[1]) / 20.0, r2 = (p[2] - cg[2]) / 20.0;
  const double m[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, -r2, r1}, {r2, 0, -r0}, {-r1, r0, 0}};
  for (int q = 0; q < 6; ++q)
    for (int d = 0; d < 3; ++d) z[q][d] = m[q][d];
}
__device__ __forceinline__ int zcol0(int g) { return g == 0 ? 0 : 3 + 6 * (g - 1); }
__device__ __forceinline__ int zncol(int g) { return g == 0 ? 3 : 6; }

// group centres: sums of node positions and counts per bone group into acc[ngroup][4]
__global__ void k_group_centres(Groups G, double* __restrict__ acc) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.nn) return;
  const int g = G.group[n];
  if (g < 1) return;
  double p[3];
  node_pos(n, G.nnode, G.copy_orig, G.nx, G.ny, G.xe, G.ye, G.ze, p);
  for (int d = 0; d < 3; ++d) atomicAdd(acc + 4 * (g - 1) + d, p[d]);
  atomicAdd(acc + 4 * (g - 1) + 3, 1.0);
}

// out[k] += (Z^T v)[k] for ncol = 3 + 6 ngroup columns; block partial sums in LDS, then atomics
__global__ __launch_bounds__(256) void k_zt(Groups G, const double* __restrict__ v, double* __restrict__ out) {
  extern __shared__ double sh[];
  const int ncol = 3 + 6 * G.ngroup;
  for (int k = threadIdx.x; k < ncol; k += blockDim.x) sh[k] = 0.0;
  __syncthreads();
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < G.nn) {
    int g;
    double z[6][3];
    zrow(G, n, g, z);
    if (g >= 0) {
      const int c0 = zcol0(g), nc = zncol(g);
      for (int q = 0; q < nc; ++q) {
        const double s = z[q][0] * v[3 * n] + z[q][1] * v[3 * n + 1] + z[q][2] * v[3 * n + 2];
        if (s != 0.0) atomicAdd(sh + c0 + q, s);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ncol; k += blockDim.x)
    if (sh[k] != 0.0) atomicAdd(out + k, sh[k]);
}

// x = Z c (every node: its group's columns)
__global__ void k_zc(Groups G, const double* __restrict__ c, double* __restrict__ x) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.nn) return;
  int g;
  double z[6][3];
  zrow(G, n, g, z);
  double o[3] = {0.0, 0.0, 0.0};
  if (g >= 0) {
    const int c0 = zcol0(g), nc = zncol(g);
    for (int q = 0; q < nc; ++q)
      for (int d = 0; d < 3; ++d) o[d] += z[q][d] * c[c0 + q];
  }
  for (int d = 0; d < 3; ++d) x[3 * n + d] = o[d];
}

__global__ void k_fill_i32(long long n, int* __restrict__ p, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(long long n, double* __restrict__ p, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(long long n, float* __restrict__ p, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_cast_bf(long long n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}
__global__ void k_cast_f(long long n, const uint16_t* __restrict__ a, float* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = bf2f(a[i]);
}
__global__ void k_axpy_f(long long n, float a, const float* __restrict__ x, float* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

inline dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace nn
}  // namespace cm

using namespace cm::nn;

static thread_local std::string nn_error;

#define NN_BEGIN try {
#define NN_END                                  \
  return 0;                                     \
  }                                             \
  catch (const std::exception& e) {             \
    nn_error = e.what();                        \
    return -1;                                  \
  }

extern "C" {

const char* cm_nn_last_error(void) { return nn_error.c_str(); }

int cm_nn_gemm_tn(void* stream, int M, int N, int K, const void* A, long long lda, const void* B, long long ldb, void* C,
                  long long ldc, const float* bias, const void* add, long long ldadd, int relu, int mode, int ksplit,
                  float alpha) {
  NN_BEGIN
  CM_CHECK(M >= 0 && N > 0 && K > 0 && K % 32 == 0, "gemm: K must be a positive multiple of 32");
  CM_CHECK(lda % 8 == 0 && ldb % 8 == 0 && lda >= K && ldb >= K, "gemm: lda, ldb must be multiples of 8 and >= K");
  CM_CHECK(((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0, "gemm: operands must be 16-byte aligned");
  CM_CHECK(mode >= 0 && mode <= 2, "gemm: bad mode");
  CM_CHECK(ksplit >= 1, "gemm: ksplit >= 1");
  if (M == 0) return 0;
  GemmArgs g{M, N, K, (const uint16_t*)A, lda, (const uint16_t*)B, ldb, C, ldc, bias, (const uint16_t*)add, ldadd,
             relu, mode, K, alpha};
  const int steps = K / 32;
  const int per = (steps + ksplit - 1) / ksplit;
  g.kc = 32 * per;
  const int nz = (steps + per - 1) / per;
  CM_CHECK(nz == 1 || mode == EPI_ATOMIC, "gemm: split K needs the atomic mode");
  const hipStream_t st = (hipStream_t)stream;
  // wave tiles: TM x TN of 16 x 16; block = 2 x 2 waves
  if (N >= 128) {
    dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 127) / 128), nz);
    hipLaunchKernelGGL((k_gemm_tn<2, 4>), grid, dim3(256), 0, st, g);
  } else if (N >= 64) {
    dim3 grid((unsigned)((M + 127) / 128), (unsigned)((N + 63) / 64), nz);
    hipLaunchKernelGGL((k_gemm_tn<4, 2>), grid, dim3(256), 0, st, g);
  } else {
    dim3 grid((unsigned)((M + 127) / 128), (unsigned)((N + 31) / 32), nz);
    hipLaunchKernelGGL((k_gemm_tn<4, 1>), grid, dim3(256), 0, st, g);
  }
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_im2col(void* stream, int B, int nx, int ny, int nz, int C, int Kp, const void* X, void* col, int transposed,
                 long long Mp) {
  NN_BEGIN
  CM_CHECK(C % 8 == 0 && Kp % 32 == 0 && Kp >= 27 * C, "im2col: C multiple of 8, Kp >= 27 C, multiple of 32");
  const long long M = (long long)B * nx * ny * nz;
  const hipStream_t st = (hipStream_t)stream;
  if (transposed) {
    CM_CHECK(Mp >= M && Mp % 8 == 0, "im2col: Mp >= M, multiple of 8");
    hipLaunchKernelGGL(k_im2col_t, g1((long long)Kp * (Mp / 8)), dim3(256), 0, st, B, nx, ny, nz, C, Kp, Mp,
                       (const uint16_t*)X, (uint16_t*)col);
  } else {
    hipLaunchKernelGGL(k_im2col, g1(M * (Kp / 8)), dim3(256), 0, st, B, nx, ny, nz, C, Kp, (const uint16_t*)X,
                       (uint16_t*)col);
  }
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_transpose(void* stream, long long M, long long Mp, int C, const void* in_bf16, const float* in_f32, void* out) {
  NN_BEGIN
  CM_CHECK((in_bf16 != nullptr) != (in_f32 != nullptr), "transpose: exactly one source");
  CM_CHECK(Mp >= M, "transpose: Mp >= M");
  hipLaunchKernelGGL(k_transpose_bf, g1(Mp * C), dim3(256), 0, (hipStream_t)stream, M, Mp, C,
                     (const uint16_t*)in_bf16, in_f32, (uint16_t*)out);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_weights(void* stream, int Cin, int Coutp, int Kp, int Kpf, const float* W, void* Wt, void* Wf) {
  NN_BEGIN
  CM_CHECK(Kp >= 27 * Cin && (Wf == nullptr || Kpf >= 27 * Coutp), "weights: padded sizes too small");
  const long long n = (long long)Coutp * Kp + (Wf ? (long long)Cin * Kpf : 0);
  hipLaunchKernelGGL(k_weights_bf16, g1(n), dim3(256), 0, (hipStream_t)stream, Cin, Coutp, Kp, Kpf, W, (uint16_t*)Wt,
                     (uint16_t*)Wf);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_adam(void* stream, long long n, float* w, float* m, float* v, const float* g, float lr, float b1, float b2,
               float eps, int step) {
  NN_BEGIN
  CM_CHECK(step >= 1, "adam: step counts from 1");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, (hipStream_t)stream, n, w, m, v, g, lr, b1, b2, eps, c1, c2);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_colsum(void* stream, long long M, int N, const float* g, float* out) {
  NN_BEGIN
  CM_CHECK(N >= 1 && N <= 256, "colsum: N <= 256");
  if (M == 0) return 0;
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, (hipStream_t)stream, M, N, g, out);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_relu_bwd(void* stream, long long n, const float* dy, const void* y, void* g_bf16, float* g_f32) {
  NN_BEGIN
  hipLaunchKernelGGL(k_relu_bwd, g1(n), dim3(256), 0, (hipStream_t)stream, n, dy, (const uint16_t*)y,
                     (uint16_t*)g_bf16, g_f32);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_pool(void* stream, int B, int nx, int ny, int nz, int C, const void* x, void* y) {
  NN_BEGIN
  CM_CHECK(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "pool: even sizes");
  hipLaunchKernelGGL(k_pool, g1((long long)B * (nx / 2) * (ny / 2) * (nz / 2) * C), dim3(256), 0, (hipStream_t)stream,
                     B, nx, ny, nz, C, (const uint16_t*)x, (uint16_t*)y);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_pool_bwd(void* stream, int B, int nx, int ny, int nz, int C, const float* dy, float* dx) {
  NN_BEGIN
  CM_CHECK(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "pool: even sizes");
  hipLaunchKernelGGL(k_pool_bwd, g1((long long)B * nx * ny * nz * C), dim3(256), 0, (hipStream_t)stream, B, nx, ny, nz,
                     C, dy, dx);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_up(void* stream, int B, int cx, int cy, int cz, int C, const void* x, void* y) {
  NN_BEGIN
  hipLaunchKernelGGL(k_up, g1((long long)B * 8 * cx * cy * cz * C), dim3(256), 0, (hipStream_t)stream, B, cx, cy, cz,
                     C, (const uint16_t*)x, (uint16_t*)y);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_up_bwd(void* stream, int B, int cx, int cy, int cz, int C, const float* dy, float* dx) {
  NN_BEGIN
  hipLaunchKernelGGL(k_up_bwd, g1((long long)B * cx * cy * cz * C), dim3(256), 0, (hipStream_t)stream, B, cx, cy, cz,
                     C, dy, dx);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_mse(void* stream, long long M, int nc, int ldo, const float* out, const float* t, const uint8_t* mask,
              float inv_norm, float* d, double* part) {
  NN_BEGIN
  CM_CHECK(nc <= ldo, "mse: nc <= ldo");
  hipLaunchKernelGGL(k_mse, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, (hipStream_t)stream, M, nc, ldo, out, t,
                     mask, inv_norm, d, part);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_features(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int8_t* cls, const float* nu,
                   const uint8_t* splitcell, float kif, const float* ground, const float* lig, const float* load,
                   float load_scale, void* F) {
  NN_BEGIN
  hipLaunchKernelGGL(k_features, g1((long long)nx * ny * nz), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, lab, cls,
                     nu, splitcell, kif, ground, lig, load, load_scale, (uint16_t*)F);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_node_to_cell(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int* irr_of,
                       const long long* irr_nodes, const double* u, float scale, float* t) {
  NN_BEGIN
  hipLaunchKernelGGL(k_node_to_cell, g1((long long)nx * ny * nz), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, lab,
                     irr_of, irr_nodes, u, scale, t);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_cell_to_node(void* stream, int nx, int ny, int nz, const uint8_t* lab, const int* irr_of,
                       const long long* irr_nodes, const float* pred, int ldp, long long nn, float* acc, double scale,
                       double* x) {
  NN_BEGIN
  const hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_f32, g1(4 * nn), dim3(256), 0, st, 4 * nn, acc, 0.f);
  hipLaunchKernelGGL(k_cell_to_node_acc, g1((long long)nx * ny * nz), dim3(256), 0, st, nx, ny, nz, lab, irr_of,
                     irr_nodes, pred, ldp, acc);
  hipLaunchKernelGGL(k_cell_to_node_fin, g1(nn), dim3(256), 0, st, nn, acc, scale, x);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_node_groups(void* stream, int nx, int ny, int nz, const uint8_t* lab, const uint8_t* part, const int* irr_of,
                      const long long* irr_nodes, long long nn, int* group) {
  NN_BEGIN
  const hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_i32, g1(nn), dim3(256), 0, st, nn, group, -1);
  hipLaunchKernelGGL(k_node_groups, g1((long long)nx * ny * nz), dim3(256), 0, st, nx, ny, nz, lab, part, irr_of,
                     irr_nodes, group);
  CM_LAUNCH_CHECK();
  NN_END
}

static Groups groups_of(int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
                        const double* xe, const double* ye, const double* ze, int ngroup, const double* centre,
                        const double* pred) {
  return Groups{group, copy_orig, xe, ye, ze, nn, (long long)(nx + 1) * (ny + 1) * (nz + 1), nx, ny, ngroup, centre,
                pred};
}

int cm_nn_group_centres(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
                        const double* xe, const double* ye, const double* ze, int ngroup, double* acc) {
  NN_BEGIN
  const hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_f64, g1(4LL * ngroup), dim3(256), 0, st, 4LL * ngroup, acc, 0.0);
  hipLaunchKernelGGL(k_group_centres, g1(nn), dim3(256), 0, st,
                     groups_of(nx, ny, nz, nn, group, copy_orig, xe, ye, ze, ngroup, nullptr, nullptr), acc);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_zt(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
             const double* xe, const double* ye, const double* ze, int ngroup, const double* centre, const double* pred,
             const double* v, double* out) {
  NN_BEGIN
  CM_CHECK(ngroup >= 0 && ngroup <= 40, "zt: at most 40 bone groups");
  const hipStream_t st = (hipStream_t)stream;
  const int ncol = 3 + 6 * ngroup;
  hipLaunchKernelGGL(k_zt, g1(nn), dim3(256), ncol * sizeof(double), st,
                     groups_of(nx, ny, nz, nn, group, copy_orig, xe, ye, ze, ngroup, centre, pred), v, out);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_zc(void* stream, int nx, int ny, int nz, long long nn, const int* group, const long long* copy_orig,
             const double* xe, const double* ye, const double* ze, int ngroup, const double* centre, const double* pred,
             const double* c, double* x) {
  NN_BEGIN
  hipLaunchKernelGGL(k_zc, g1(nn), dim3(256), 0, (hipStream_t)stream,
                     groups_of(nx, ny, nz, nn, group, copy_orig, xe, ye, ze, ngroup, centre, pred), c, x);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_fill(void* stream, long long n, float* p, float v) {
  NN_BEGIN
  hipLaunchKernelGGL(k_fill_f32, g1(n), dim3(256), 0, (hipStream_t)stream, n, p, v);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_cast(void* stream, long long n, void* f32, void* bf16, int to_bf16) {
  NN_BEGIN
  if (to_bf16)
    hipLaunchKernelGGL(k_cast_bf, g1(n), dim3(256), 0, (hipStream_t)stream, n, (const float*)f32, (uint16_t*)bf16);
  else
    hipLaunchKernelGGL(k_cast_f, g1(n), dim3(256), 0, (hipStream_t)stream, n, (const uint16_t*)bf16, (float*)f32);
  CM_LAUNCH_CHECK();
  NN_END
}

int cm_nn_axpy(void* stream, long long n, float a, const float* x, float* y) {
  NN_BEGIN
  hipLaunchKernelGGL(k_axpy_f, g1(n), dim3(256), 0, (hipStream_t)stream, n, a, x, y);
  CM_LAUNCH_CHECK();
  NN_END
}

}  // extern "C"
