// Vector kernels of the smoothers and of PCG.  Dot products and norms accumulate in FP64 with
// wavefront shuffles, one partial per block, then a single-block finalize in fixed order
// (deterministic).  PCG scalars live in a small device array S so that no iteration needs a host
// round trip; the host only reads the residual norm at the check cadence.
#pragma once
#include "common.hpp"

namespace cm {

constexpr int RED_BLOCKS = 2048;
constexpr int RED_THREADS = 256;

// partial[block] = sum a[i] * b[i]
template <typename Ta, typename Tb>
__global__ __launch_bounds__(RED_THREADS) void k_dot(long long n, const Ta* __restrict__ a, const Tb* __restrict__ b,
                                                     double* __restrict__ partial) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (long long)gridDim.x * RED_THREADS)
    s += (double)a[i] * (double)b[i];
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// S[dst] = sum partial[0..nb)
__global__ __launch_bounds__(RED_THREADS) void k_finalize(int nb, const double* __restrict__ partial,
                                                          double* __restrict__ S, int dst) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RED_THREADS) s += partial[i];
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) S[dst] = s;
}

// ---- smoothers -------------------------------------------------------------------------
// Dinv holds the symmetric 3x3 block inverses as 6 values per node (00 01 02 11 12 22).
// First step of a smoother from x = 0: x = c * Dinv b.  Threads past nn take the split-copy
// groups (x = c Ginv b over each group); the per-node threads skip group members.
template <typename T>
__global__ void k_cheb0(long long nn, T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ Dinv, T c,
                        CopyGroups<T> ga) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) {
    if (n - nn < ga.ng) group_apply(ga, (int)(n - nn), b, x, c);
    return;
  }
  const T* m = Dinv + 6 * n;
  if (ga.ng && group_member(m)) return;
  const T r0 = b[3 * n], r1 = b[3 * n + 1], r2 = b[3 * n + 2];
  x[3 * n] = c * (m[0] * r0 + m[1] * r1 + m[2] * r2);
  x[3 * n + 1] = c * (m[1] * r0 + m[3] * r1 + m[4] * r2);
  x[3 * n + 2] = c * (m[2] * r0 + m[4] * r1 + m[5] * r2);
}

// y = M x (block-Jacobi apply, split-copy groups as in k_cheb0), for the power iteration
template <typename T>
__global__ void k_block_apply(long long nn, const T* __restrict__ Dinv, const T* __restrict__ x, T* __restrict__ y,
                              CopyGroups<T> ga) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) {
    if (n - nn < ga.ng) group_apply(ga, (int)(n - nn), x, y, T(1));
    return;
  }
  const T* m = Dinv + 6 * n;
  if (ga.ng && group_member(m)) return;
  const T x0 = x[3 * n], x1 = x[3 * n + 1], x2 = x[3 * n + 2];
  y[3 * n] = m[0] * x0 + m[1] * x1 + m[2] * x2;
  y[3 * n + 1] = m[1] * x0 + m[3] * x1 + m[4] * x2;
  y[3 * n + 2] = m[2] * x0 + m[4] * x1 + m[5] * x2;
}

template <typename T>
__global__ void k_scale(long long n, T* __restrict__ x, const double* __restrict__ S, int slot) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = T((double)x[i] / sqrt(S[slot]));
}

// deterministic pseudo-normal start vector of the power iteration (seed fixed), zero on unused
template <typename T>
__global__ void k_randn(long long nn, const uint8_t* __restrict__ used, unsigned long long seed, T* __restrict__ x) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  for (int c = 0; c < 3; ++c) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(3 * n + c + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double u1 = ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    unsigned long long w = z * 0x2545F4914F6CDD1DULL + 0x632BE59BD9B4E019ULL;
    w ^= w >> 29;
    const double u2 = ((w >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    x[3 * n + c] = (used == nullptr || used[n]) ? T(g) : T(0);
  }
}

// ---- PCG -------------------------------------------------------------------------------
// alpha = S[rz] / S[pap];  x += alpha p;  r -= alpha Ap;  partial ||r||^2;  optional cast of r
template <typename TX, typename T, typename TC>
__global__ __launch_bounds__(RED_THREADS) void k_pcg_xr(long long n, TX* __restrict__ x, T* __restrict__ r,
                                                        const T* __restrict__ p, const T* __restrict__ Ap,
                                                        const double* __restrict__ S, int s_rz, int s_pap,
                                                        double* __restrict__ partial, TC* __restrict__ rc) {
  const double alpha = S[s_rz] / S[s_pap];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (long long)gridDim.x * RED_THREADS) {
    x[i] = TX((double)x[i] + alpha * (double)p[i]);
    const T ri = T((double)r[i] - alpha * (double)Ap[i]);
    r[i] = ri;
    if (rc) rc[i] = TC(ri);
    s += (double)ri * (double)ri;
  }
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// beta = S[rz_new] / S[rz_old];  p = z + beta p
template <typename T, typename TZ = T>
__global__ void k_pcg_p(long long n, T* __restrict__ p, const TZ* __restrict__ z, const double* __restrict__ S,
                        int s_new, int s_old, int s_sub = -1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // beta = (r.z) / (r_old.z_old); flexible (Polak-Ribiere) form subtracts r.z_old (slot s_sub)
  const double beta = (S[s_new] - (s_sub >= 0 ? S[s_sub] : 0.0)) / S[s_old];
  p[i] = T((double)z[i] + beta * (double)p[i]);
}

// out = in (with conversion)
template <typename TO, typename TI>
__global__ void k_cast(long long n, TO* __restrict__ out, const TI* __restrict__ in) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = TO(in[i]);
}

// x += d (d converted)
template <typename TX, typename TD>
__global__ void k_add(long long n, TX* __restrict__ x, const TD* __restrict__ d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = TX((double)x[i] + (double)d[i]);
}

// r = b - t (FP64) with partial ||r||^2 and optional cast (iterative-refinement residual)
template <typename TC>
__global__ __launch_bounds__(RED_THREADS) void k_resid_cast(long long n, double* __restrict__ r,
                                                            const double* __restrict__ b, const double* __restrict__ t,
                                                            double* __restrict__ partial, TC* __restrict__ rc,
                                                            double scale) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (long long)gridDim.x * RED_THREADS) {
    const double ri = b[i] - t[i];
    r[i] = ri;
    if (rc) rc[i] = TC(ri * scale);
    s += ri * ri;
  }
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x += scale * d
template <typename TD>
__global__ void k_axpy_scaled(long long n, double* __restrict__ x, const TD* __restrict__ d, double scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += scale * (double)d[i];
}

// Rigid-mode (deflation) correction of rb/gpusolve.py:217-228 / fe/gmg.py:127-135:
// z += Z Einv Z^T r with Z given as CSR rows over DOFs (few columns).
// Step 1: partials of w = Z^T r -- one block per chunk of at most ZT_CHUNK entries of one column
// (CSC layout), so that the few long columns spread over many CUs.
constexpr int ZT_CHUNK = 2048;
__global__ __launch_bounds__(RED_THREADS) void k_zt_r(const int* __restrict__ chk, const int* __restrict__ rows,
                                                      const double* __restrict__ vals, const double* __restrict__ r,
                                                      double* __restrict__ part) {
  const int b = blockIdx.x;
  double s = 0.0;
  for (int q = chk[b] + threadIdx.x; q < chk[b + 1]; q += RED_THREADS) s += vals[q] * r[rows[q]];
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) part[b] = s;
}
// Step 2 (one block): w_j = the column's chunk partials in order, then v = Einv w
__global__ __launch_bounds__(RED_THREADS) void k_rigid_small(int ncol, const int* __restrict__ cchk,
                                                             const double* __restrict__ part,
                                                             const double* __restrict__ Einv, double* __restrict__ w,
                                                             double* __restrict__ v) {
  for (int j = threadIdx.x; j < ncol; j += RED_THREADS) {
    double s = 0.0;
    for (int c = cchk[j]; c < cchk[j + 1]; ++c) s += part[c];
    w[j] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ncol; i += RED_THREADS) {
    double s = 0.0;
    for (int j = 0; j < ncol; ++j) s += Einv[(long long)i * ncol + j] * w[j];
    v[i] = s;
  }
}
// Step 3: z[row] += sum_j Z[row, j] v[j] (CSR over the touched rows)
__global__ void k_z_v(int nrow, const int* __restrict__ rows, const int* __restrict__ ptr,
                      const int* __restrict__ cols, const double* __restrict__ vals, const double* __restrict__ v,
                      double* __restrict__ z) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  double s = 0.0;
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) s += vals[q] * v[cols[q]];
  z[rows[r]] += s;
}

}  // namespace cm
