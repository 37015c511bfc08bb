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
__device__ __forceinline__ void dot_body(const int bid, const int nblk, long long n, const Ta* __restrict__ a,
                                         const Tb* __restrict__ b, double* __restrict__ partial) {
  double s = 0.0;
  for (long long i = (long long)bid * RED_THREADS + threadIdx.x; i < n; i += (long long)nblk * RED_THREADS)
    s += (double)a[i] * (double)b[i];
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) partial[bid] = s;
}
template <typename Ta, typename Tb>
__global__ __launch_bounds__(RED_THREADS) void k_dot(long long n, const Ta* __restrict__ a, const Tb* __restrict__ b,
                                                     double* __restrict__ partial) {
  dot_body<Ta, Tb>(blockIdx.x, gridDim.x, n, a, b, partial);
}
template <typename Ta, typename Tb>
struct DotL {
  long long n;
  const Ta* a;
  const Tb* b;
  double* partial;
  __device__ __forceinline__ void run(int bid, int nblk) const { dot_body<Ta, Tb>(bid, nblk, n, a, b, partial); }
};

// S[dst] = sum partial[0..nb)
__device__ __forceinline__ void finalize_body(int nb, const double* __restrict__ partial, double* __restrict__ S,
                                              int dst) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RED_THREADS) s += partial[i];
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) S[dst] = s;
}
__global__ __launch_bounds__(RED_THREADS) void k_finalize(int nb, const double* __restrict__ partial,
                                                          double* __restrict__ S, int dst) {
  finalize_body(nb, partial, S, dst);
}
struct FinalizeL {
  int nb;
  const double* partial;
  double* S;
  int dst;
  __device__ __forceinline__ void run(int, int) const { finalize_body(nb, partial, S, dst); }
};
// S[slot] = v (one thread)
struct SlotSetL {
  double* S;
  int slot;
  double v;
  __device__ __forceinline__ void run(int, int) const {
    if (threadIdx.x == 0) S[slot] = v;
  }
};
// out = in, converted
template <typename TO, typename TI>
struct CastL {
  long long n;
  TO* out;
  const TI* in;
  __device__ __forceinline__ void run(int bid, int) const {
    const long long i = (long long)bid * blockDim.x + threadIdx.x;
    if (i < n) out[i] = TO(in[i]);
  }
};
// S[dst] = S[src] (one thread)
struct SlotCopyL {
  double* S;
  int dst, src;
  __device__ __forceinline__ void run(int, int) const {
    if (threadIdx.x == 0) S[dst] = S[src];
  }
};

// ---- smoothers -------------------------------------------------------------------------
// Dinv holds the symmetric 3x3 block inverses as 6 values per node (00 01 02 11 12 22).
// First step of a smoother from x = 0: x = c * Dinv b.  Threads past nn take the split-copy
// groups (x = c Ginv b over each group); the per-node threads skip group members.
template <typename T>
__device__ __forceinline__ void cheb0_body(const int bid, long long nn, T* __restrict__ x, const T* __restrict__ b,
                                           const T* __restrict__ Dinv, T c, const CopyGroups<T>& ga) {
  const long long n = (long long)bid * blockDim.x + threadIdx.x;
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
template <typename T>
__global__ void k_cheb0(long long nn, T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ Dinv, T c,
                        CopyGroups<T> ga) {
  cheb0_body<T>(blockIdx.x, nn, x, b, Dinv, c, ga);
}
template <typename T>
struct Cheb0L {
  long long nn;
  T* x;
  const T* b;
  const T* Dinv;
  T c;
  CopyGroups<T> ga;
  __device__ __forceinline__ void run(int bid, int) const { cheb0_body<T>(bid, nn, x, b, Dinv, c, ga); }
};

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
__device__ __forceinline__ void pcg_xr_body(const int bid, const int nblk, long long n, TX* __restrict__ x,
                                            T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ Ap,
                                            const double* __restrict__ S, int s_rz, int s_pap,
                                            double* __restrict__ partial, TC* __restrict__ rc) {
  const double alpha = S[s_rz] / S[s_pap];
  double s = 0.0;
  for (long long i = (long long)bid * RED_THREADS + threadIdx.x; i < n; i += (long long)nblk * RED_THREADS) {
    x[i] = TX((double)x[i] + alpha * (double)p[i]);
    const T ri = T((double)r[i] - alpha * (double)Ap[i]);
    r[i] = ri;
    if (rc) rc[i] = TC(ri);
    s += (double)ri * (double)ri;
  }
  s = block_sum<RED_THREADS>(s);
  if (threadIdx.x == 0) partial[bid] = s;
}
template <typename TX, typename T, typename TC>
__global__ __launch_bounds__(RED_THREADS) void k_pcg_xr(long long n, TX* __restrict__ x, T* __restrict__ r,
                                                        const T* __restrict__ p, const T* __restrict__ Ap,
                                                        const double* __restrict__ S, int s_rz, int s_pap,
                                                        double* __restrict__ partial, TC* __restrict__ rc) {
  pcg_xr_body<TX, T, TC>(blockIdx.x, gridDim.x, n, x, r, p, Ap, S, s_rz, s_pap, partial, rc);
}
template <typename TX, typename T, typename TC>
struct PcgXrL {
  long long n;
  TX* x;
  T* r;
  const T* p;
  const T* Ap;
  const double* S;
  int s_rz, s_pap;
  double* partial;
  TC* rc;
  __device__ __forceinline__ void run(int bid, int nblk) const {
    pcg_xr_body<TX, T, TC>(bid, nblk, n, x, r, p, Ap, S, s_rz, s_pap, partial, rc);
  }
};

// beta = S[rz_new] / S[rz_old];  p = z + beta p
template <typename T, typename TZ>
__device__ __forceinline__ void pcg_p_body(const int bid, long long n, T* __restrict__ p, const TZ* __restrict__ z,
                                           const double* __restrict__ S, int s_new, int s_old, int s_sub) {
  const long long i = (long long)bid * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // beta = (r.z) / (r_old.z_old); flexible (Polak-Ribiere) form subtracts r.z_old (slot s_sub)
  const double beta = (S[s_new] - (s_sub >= 0 ? S[s_sub] : 0.0)) / S[s_old];
  p[i] = T((double)z[i] + beta * (double)p[i]);
}
template <typename T, typename TZ = T>
__global__ void k_pcg_p(long long n, T* __restrict__ p, const TZ* __restrict__ z, const double* __restrict__ S,
                        int s_new, int s_old, int s_sub = -1) {
  pcg_p_body<T, TZ>(blockIdx.x, n, p, z, S, s_new, s_old, s_sub);
}
template <typename T, typename TZ>
struct PcgPL {
  long long n;
  T* p;
  const TZ* z;
  const double* S;
  int s_new, s_old, s_sub;
  __device__ __forceinline__ void run(int bid, int) const { pcg_p_body<T, TZ>(bid, n, p, z, S, s_new, s_old, s_sub); }
};

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
template <typename TX, typename TD>
struct AddL {
  long long n;
  TX* x;
  const TD* d;
  __device__ __forceinline__ void run(int bid, int) const {
    const long long i = (long long)bid * blockDim.x + threadIdx.x;
    if (i < n) x[i] = TX((double)x[i] + (double)d[i]);
  }
};

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
template <typename T>
__global__ void k_z_v(int nrow, const int* __restrict__ rows, const int* __restrict__ ptr,
                      const int* __restrict__ cols, const double* __restrict__ vals, const double* __restrict__ v,
                      T* __restrict__ z) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  double s = 0.0;
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) s += vals[q] * v[cols[q]];
  z[rows[r]] = T((double)z[rows[r]] + s);
}

// out[i] = S_i[slot] for the n slot arrays S (one block): the residual norms of a batch's systems
// read back with one copy
__global__ void k_gather_slot(int n, double* const* __restrict__ S, int slot, double* __restrict__ out) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = S[i][slot];
}

// ---- launch records --------------------------------------------------------------------
// A launch record L holds a kernel's arguments and runs block bid of nblk (256 threads).
// k_one launches one record; k_multi launches the records of several systems as one grid:
// blocks start[s] .. start[s+1]-1 run record s (start has nsys + 1 entries).
template <class L>
__global__ __launch_bounds__(256) void k_one(L a) {
  a.run(blockIdx.x, gridDim.x);
}
template <class L>
__global__ __launch_bounds__(256) void k_multi(const L* __restrict__ a, const int* __restrict__ start, int nsys) {
  const int b = blockIdx.x;
  int lo = 0, hi = nsys - 1;   // the last s with start[s] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  // s is the same on every lane: say so, so that the record's fields are read with scalar loads
  // into scalar registers (not once per lane into vector registers)
  const int s = __builtin_amdgcn_readfirstlane(lo);
  // the record read through the constant address space: its pointers then count as global
  // (loads and stores through them compile to global, not flat, instructions)
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) L CL;
  L la = *reinterpret_cast<CL*>(reinterpret_cast<uintptr_t>(a + s));
  la.run(b - start[s], start[s + 1] - start[s]);
#endif
}

}  // namespace cm
