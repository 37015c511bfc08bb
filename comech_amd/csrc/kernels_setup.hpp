// Hierarchy setup kernels: material moments of coarse cells, 3x3 diagonal blocks, the dense
// coarsest-level operator and its inverse.  Everything here runs once per system (FP64).
#pragma once
#include "common.hpp"
#include "element.hpp"

namespace cm {

// Per-axis description of how the cells of a fine level sit inside the cells of the next level:
// coarse cell C along this axis covers fine cells cs[C] .. cs[C]+cn[C]-1 (1 or 2 cells); for every
// fine cell, a (scale) and b (offset) map its coordinate into the parent's: xi_p = a xi_c + b.
struct AxisNest {
  const int* __restrict__ cs;
  const int* __restrict__ cnum;
  const double* __restrict__ a;   // per fine cell
  const double* __restrict__ b;   // per fine cell
};

// Level-1 moments from level-0 cells.  Fine cell materials come as (lam, mu) per fine cell in
// x-fastest order (inactive = 0).  sri: lam integrated at the fine cell centre.
__global__ void k_moments_from_fine(int ncx, int ncy, int ncz, int nfx, int nfy, AxisNest ax, AxisNest ay, AxisNest az,
                                    const double* __restrict__ hx, const double* __restrict__ hy,
                                    const double* __restrict__ hz, const double* __restrict__ lam,
                                    const double* __restrict__ mu, int sri, double* __restrict__ mom) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nc = (long long)ncx * ncy * ncz;
  if (id >= nc) return;
  const int I = id % ncx, J = (id / ncx) % ncy, K = id / ((long long)ncx * ncy);
  double mL[27], mM[27];
  for (int q = 0; q < 27; ++q) { mL[q] = 0.0; mM[q] = 0.0; }
  for (int dk = 0; dk < az.cnum[K]; ++dk) {
    const int k = az.cs[K] + dk;
    for (int dj = 0; dj < ay.cnum[J]; ++dj) {
      const int j = ay.cs[J] + dj;
      for (int di = 0; di < ax.cnum[I]; ++di) {
        const int i = ax.cs[I] + di;
        const long long f = i + (long long)nfx * (j + (long long)nfy * k);
        const double l = lam[f], m = mu[f];
        if (l == 0.0 && m == 0.0) continue;
        const double V = hx[i] * hy[j] * hz[k];
        const double A[3] = {ax.a[i], ay.a[j], az.a[k]};
        const double B[3] = {ax.b[i], ay.b[j], az.b[k]};
        double t[27];
        mom_const<double>(l, V, sri != 0, t);
        mom_shift_add<double>(t, A, B, mL);
        mom_const<double>(m, V, false, t);
        mom_shift_add<double>(t, A, B, mM);
      }
    }
  }
  for (int q = 0; q < 27; ++q) {
    mom[q * nc + id] = mL[q];
    mom[(27 + q) * nc + id] = mM[q];
  }
}

// Level l+1 moments from level-l moments (shift of the child moments into the parent).
__global__ void k_moments_from_moments(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz, AxisNest ax, AxisNest ay,
                                       AxisNest az, const double* __restrict__ momf, double* __restrict__ mom) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nc = (long long)ncx * ncy * ncz;
  if (id >= nc) return;
  const long long nf = (long long)nfx * nfy * nfz;
  const int I = id % ncx, J = (id / ncx) % ncy, K = id / ((long long)ncx * ncy);
  double mL[27], mM[27];
  for (int q = 0; q < 27; ++q) { mL[q] = 0.0; mM[q] = 0.0; }
  for (int dk = 0; dk < az.cnum[K]; ++dk) {
    const int k = az.cs[K] + dk;
    for (int dj = 0; dj < ay.cnum[J]; ++dj) {
      const int j = ay.cs[J] + dj;
      for (int di = 0; di < ax.cnum[I]; ++di) {
        const int i = ax.cs[I] + di;
        const long long f = i + (long long)nfx * (j + (long long)nfy * k);
        if (momf[f] == 0.0 && momf[27 * nf + f] == 0.0) continue;
        const double A[3] = {ax.a[i], ay.a[j], az.a[k]};
        const double B[3] = {ax.b[i], ay.b[j], az.b[k]};
        double t[27];
        for (int q = 0; q < 27; ++q) t[q] = momf[q * nf + f];
        mom_shift_add<double>(t, A, B, mL);
        for (int q = 0; q < 27; ++q) t[q] = momf[(27 + q) * nf + f];
        mom_shift_add<double>(t, A, B, mM);
      }
    }
  }
  for (int q = 0; q < 27; ++q) {
    mom[q * nc + id] = mL[q];
    mom[(27 + q) * nc + id] = mM[q];
  }
}

// 3x3 diagonal block of corner l of one cell's element matrix: D[r][c] = (K e_{l,c})[3l + r].
template <bool MOM, bool SRI>
__device__ void corner_block(int l, const double* mL, const double* mM, double lam, double mu, double hx, double hy,
                             double hz, double* D) {
  for (int c = 0; c < 3; ++c) {
    double u[24], f[24];
    for (int q = 0; q < 24; ++q) u[q] = 0.0;
    u[3 * l + c] = 1.0;
    if (MOM) {
      elem_apply_mom<double>(u, f, mL, mM, 0.25 / hx, 0.25 / hy, 0.25 / hz);
    } else {
      const double V = hx * hy * hz;
      CellW<double> w{0.25 / hx, 0.25 / hy, 0.25 / hz, lam * V, mu * V};
      elem_apply_const<double, SRI>(u, f, w);
    }
    for (int r = 0; r < 3; ++r) D[3 * r + c] = f[3 * l + r];
  }
}

// Diagonal blocks of the structured part at every grid node (gather over the 8 cells).
// Level 0: constant-material cells with (lam, mu) per cell (x-fastest, irregular cells zeroed).
// Coarse levels: moments.
template <bool MOM, bool SRI>
__global__ void k_diag_structured(int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
                                  const double* __restrict__ hz, const double* __restrict__ lam,
                                  const double* __restrict__ mu, const double* __restrict__ mom,
                                  double* __restrict__ D) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)(nx + 1) * (ny + 1) * (nz + 1);
  if (id >= nn) return;
  const long long ncell = (long long)nx * ny * nz;
  const int i = id % (nx + 1), j = (id / (nx + 1)) % (ny + 1), k = id / ((long long)(nx + 1) * (ny + 1));
  double acc[9];
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        // cell whose corner (a,b,c) is this node
        const int ci = i - a, cj = j - b, ck = k - c;
        if (ci < 0 || ci >= nx || cj < 0 || cj >= ny || ck < 0 || ck >= nz) continue;
        const long long cid = ci + (long long)nx * (cj + (long long)ny * ck);
        double blk[9];
        const int l = a + 2 * b + 4 * c;
        if (MOM) {
          double mL[27], mM[27];
          for (int q = 0; q < 27; ++q) { mL[q] = mom[q * ncell + cid]; mM[q] = mom[(27 + q) * ncell + cid]; }
          if (mL[0] == 0.0 && mM[0] == 0.0) continue;
          corner_block<true, SRI>(l, mL, mM, 0, 0, hx[ci], hy[cj], hz[ck], blk);
        } else {
          const double lv = lam[cid], mv = mu[cid];
          if (lv == 0.0 && mv == 0.0) continue;
          corner_block<false, SRI>(l, nullptr, nullptr, lv, mv, hx[ci], hy[cj], hz[ck], blk);
        }
        for (int q = 0; q < 9; ++q) acc[q] += blk[q];
      }
  for (int q = 0; q < 9; ++q) D[9 * id + q] = acc[q];
}

// Diagonal blocks of irregular cells (atomics into D).
template <bool SRI>
__global__ void k_diag_irregular(int nirr, const int* __restrict__ nodes, const int* __restrict__ ijk,
                                 const double* __restrict__ lam, const double* __restrict__ mu,
                                 const double* __restrict__ hx, const double* __restrict__ hy,
                                 const double* __restrict__ hz, double* __restrict__ D) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nirr) return;
  const int i = ijk[3 * e], j = ijk[3 * e + 1], k = ijk[3 * e + 2];
  for (int l = 0; l < 8; ++l) {
    double blk[9];
    corner_block<false, SRI>(l, nullptr, nullptr, lam[e], mu[e], hx[i], hy[j], hz[k], blk);
    const long long n = nodes[8 * e + l];
    for (int q = 0; q < 9; ++q) atomicAdd(&D[9 * n + q], blk[q]);
  }
}

// Diagonal-block entries of the extra stiffness (rows unique, entries with col/3 == row/3).
__global__ void k_diag_extra(int nrow, const int* __restrict__ rows, const int* __restrict__ ptr,
                             const int* __restrict__ cols, const double* __restrict__ vals, double* __restrict__ D) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  const int row = rows[r];
  const int n = row / 3, rc = row % 3;
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) {
    const int col = cols[q];
    if (col / 3 == n) atomicAdd(&D[9LL * n + 3 * rc + col % 3], vals[q]);
  }
}

// Diagonal blocks of rank-one terms: D[node] += k w_i w_j for entry pairs on the same node.
__global__ void k_diag_rank1(int nterm, const int* __restrict__ ptr, const int* __restrict__ dof,
                             const double* __restrict__ w, const double* __restrict__ k, double* __restrict__ D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterm) return;
  for (int a = ptr[t]; a < ptr[t + 1]; ++a)
    for (int b = ptr[t]; b < ptr[t + 1]; ++b) {
      const int na = dof[a] / 3, nb = dof[b] / 3;
      if (na != nb) continue;
      atomicAdd(&D[9LL * na + 3 * (dof[a] % 3) + dof[b] % 3], k[t] * w[a] * w[b]);
    }
}

// Rank-one terms into the dense coarsest matrix.
__global__ void k_dense_add_rank1(int nterm, const int* __restrict__ ptr, const int* __restrict__ dof,
                                  const double* __restrict__ w, const double* __restrict__ k, int n,
                                  double* __restrict__ A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterm) return;
  for (int a = ptr[t]; a < ptr[t + 1]; ++a)
    for (int b = ptr[t]; b < ptr[t + 1]; ++b) atomicAdd(&A[(long long)dof[a] * n + dof[b]], k[t] * w[a] * w[b]);
}

// Inverse of the 3x3 blocks; blocks with |det| < 1e-30 become the identity (fe/gmg.py:39-41).
template <typename T>
__global__ void k_block_inverse(long long nn, const double* __restrict__ D, T* __restrict__ Dinv) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double* m = D + 9 * n;
  T* o = Dinv + 9 * n;
  const double c0 = m[4] * m[8] - m[5] * m[7];
  const double c1 = m[5] * m[6] - m[3] * m[8];
  const double c2 = m[3] * m[7] - m[4] * m[6];
  const double det = m[0] * c0 + m[1] * c1 + m[2] * c2;
  if (fabs(det) < 1e-30) {
    for (int q = 0; q < 9; ++q) o[q] = (q % 4 == 0) ? T(1) : T(0);
    return;
  }
  const double id = 1.0 / det;
  o[0] = T(c0 * id);
  o[1] = T((m[2] * m[7] - m[1] * m[8]) * id);
  o[2] = T((m[1] * m[5] - m[2] * m[4]) * id);
  o[3] = T(c1 * id);
  o[4] = T((m[0] * m[8] - m[2] * m[6]) * id);
  o[5] = T((m[2] * m[3] - m[0] * m[5]) * id);
  o[6] = T(c2 * id);
  o[7] = T((m[1] * m[6] - m[0] * m[7]) * id);
  o[8] = T((m[0] * m[4] - m[1] * m[3]) * id);
}

// Dense coarsest operator: one block of 24 threads per coarse cell, thread = element column.
__global__ void k_dense_assemble(int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
                                 const double* __restrict__ hz, const double* __restrict__ mom, int n,
                                 double* __restrict__ A) {
  const long long cid = blockIdx.x;
  const long long ncell = (long long)nx * ny * nz;
  const int col = threadIdx.x;
  if (cid >= ncell || col >= 24) return;
  if (mom[cid] == 0.0 && mom[27 * ncell + cid] == 0.0) return;
  const int ci = cid % nx, cj = (cid / nx) % ny, ck = cid / ((long long)nx * ny);
  double mL[27], mM[27];
  for (int q = 0; q < 27; ++q) { mL[q] = mom[q * ncell + cid]; mM[q] = mom[(27 + q) * ncell + cid]; }
  double u[24], f[24];
  for (int q = 0; q < 24; ++q) u[q] = 0.0;
  u[col] = 1.0;
  elem_apply_mom<double>(u, f, mL, mM, 0.25 / hx[ci], 0.25 / hy[cj], 0.25 / hz[ck]);
  auto gdof = [&](int q) {
    const int l = q / 3, c = q % 3;
    const int a = l & 1, b = (l >> 1) & 1, cc = l >> 2;
    const long long node = (ci + a) + (long long)(nx + 1) * ((cj + b) + (long long)(ny + 1) * (ck + cc));
    return 3 * node + c;
  };
  const long long gc = gdof(col);
  for (int r = 0; r < 24; ++r) atomicAdd(&A[gdof(r) * n + gc], f[r]);
}

__global__ void k_dense_add_extra(int nrow, const int* __restrict__ rows, const int* __restrict__ ptr,
                                  const int* __restrict__ cols, const double* __restrict__ vals, int n,
                                  double* __restrict__ A) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) A[(long long)rows[r] * n + cols[q]] += vals[q];
}

// Rows that are not unknowns (all zero) get a unit diagonal.
__global__ void k_dense_fix_empty(int n, double* __restrict__ A) {
  const int r = blockIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) s += fabs(A[(long long)r * n + c]);
  s = block_sum<256>(s);
  if (threadIdx.x == 0 && s == 0.0) A[(long long)r * n + r] = 1.0;
}

// In-place blocked Gauss-Jordan inversion of an SPD matrix (no pivoting), pivot blocks of GJB
// rows.  For the pivot block K (all other rows/columns I, J):
//   Dk = A_KK^-1;  R_KJ = Dk A_KJ;  A_IJ -= A_IK R_KJ;  A_IK = -A_IK Dk;  A_KJ = R_KJ;  A_KK = Dk
// which is GJB consecutive steps of scalar Gauss-Jordan.
constexpr int GJB = 64;

// (1) Dk = inverse of the nb x nb diagonal block at k0 (one block of 256 threads, LDS).
__global__ __launch_bounds__(256) void k_gjb_diag(int n, int k0, int nb, const double* __restrict__ A,
                                                  double* __restrict__ Dk) {
  __shared__ double M[GJB][GJB + 1];
  __shared__ double col[GJB];
  for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
    const int i = t / GJB, j = t % GJB;
    M[i][j] = (i < nb && j < nb) ? A[(long long)(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    if (threadIdx.x < GJB) col[threadIdx.x] = M[threadIdx.x][k];
    __syncthreads();
    const double ip = 1.0 / col[k];
    for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
      const int i = t / GJB, j = t % GJB;
      if (i == k) {
        M[i][j] = (j == k) ? ip : M[i][j] * ip;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
      const int i = t / GJB, j = t % GJB;
      if (i != k) M[i][j] = (j == k) ? -col[i] * ip : M[i][j] - col[i] * M[k][j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
    const int i = t / nb, j = t % nb;
    Dk[i * GJB + j] = M[i][j];
  }
}

// (2) R[:, j] = Dk A[K, j] for every column j (R is GJB x n; columns inside K are skipped later).
__global__ __launch_bounds__(256) void k_gjb_rowpanel(int n, int k0, int nb, const double* __restrict__ A,
                                                      const double* __restrict__ Dk, double* __restrict__ R) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double a[GJB];
  for (int q = 0; q < nb; ++q) a[q] = A[(long long)(k0 + q) * n + j];
  for (int i = 0; i < nb; ++i) {
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += Dk[i * GJB + q] * a[q];
    R[(long long)i * n + j] = s;
  }
}

// (3) A_IJ -= A_IK R_KJ for I, J outside K: 64x64 output tiles, K-panel staged in LDS.
__global__ __launch_bounds__(256) void k_gjb_update(int n, int k0, int nb, double* __restrict__ A,
                                                    const double* __restrict__ R) {
  __shared__ double sa[64][GJB + 1];   // A[I-tile, K]
  __shared__ double sr[GJB][64 + 1];   // R[K, J-tile]
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * GJB; t += blockDim.x) {
    const int r = t / GJB, q = t % GJB;
    const int i = i0 + r;
    sa[r][q] = (i < n && q < nb) ? A[(long long)i * n + k0 + q] : 0.0;
    const int q2 = t / 64, c = t % 64;
    const int j = j0 + c;
    sr[q2][c] = (j < n && q2 < nb) ? R[(long long)q2 * n + j] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;   // 16x16 threads, 4x4 outputs each
  double acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int q = 0; q < nb; ++q) {
    double x[4], y[4];
    for (int a = 0; a < 4; ++a) x[a] = sa[tr + 16 * a][q];
    for (int b = 0; b < 4; ++b) y[b] = sr[q][tc + 16 * b];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + tr + 16 * a;
    if (i >= n || (i >= k0 && i < k0 + nb)) continue;
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tc + 16 * b;
      if (j >= n || (j >= k0 && j < k0 + nb)) continue;
      A[(long long)i * n + j] -= acc[a][b];
    }
  }
}

// (4) A_IK = -A_IK Dk for rows outside K; (5) A_KJ = R_KJ for columns outside K, A_KK = Dk.
__global__ __launch_bounds__(256) void k_gjb_colpanel(int n, int k0, int nb, double* __restrict__ A,
                                                      const double* __restrict__ Dk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (i >= k0 && i < k0 + nb)) return;
  double a[GJB];
  for (int q = 0; q < nb; ++q) a[q] = A[(long long)i * n + k0 + q];
  for (int j = 0; j < nb; ++j) {
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += a[q] * Dk[q * GJB + j];
    A[(long long)i * n + k0 + j] = -s;
  }
}
__global__ __launch_bounds__(256) void k_gjb_finish(int n, int k0, int nb, double* __restrict__ A,
                                                    const double* __restrict__ R, const double* __restrict__ Dk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const bool inK = (j >= k0 && j < k0 + nb);
  for (int q = 0; q < nb; ++q)
    A[(long long)(k0 + q) * n + j] = inK ? Dk[q * GJB + (j - k0)] : R[(long long)q * n + j];
}

// Coarse solve x = Ainv b (dense FP64), one wavefront per row.
template <typename T>
__global__ void k_dense_gemv(int n, const double* __restrict__ Ainv, const T* __restrict__ b, T* __restrict__ x) {
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wid >= n) return;
  const double* row = Ainv + (long long)wid * n;
  double s = 0.0;
  for (int c = lane; c < n; c += 64) s += row[c] * (double)b[c];
  s = wave_sum(s);
  if (lane == 0) x[wid] = T(s);
}

}  // namespace cm
