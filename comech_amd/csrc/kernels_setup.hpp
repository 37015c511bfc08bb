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
__global__ __launch_bounds__(256) void k_moments_from_fine(int ncx, int ncy, int ncz, int nfx, int nfy, AxisNest ax, AxisNest ay, AxisNest az,
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

// Level-1 child weights: lam V and mu V of the (up to) 2x2x2 level-0 children of every level-1
// cell, slot dx + 2 dy + 4 dz (lam 0..7, mu 8..15; a missing child or inactive cell is 0).  The
// same fine materials and volumes as k_moments_from_fine; the march expands them into moments.
__global__ __launch_bounds__(256) void k_child_weights(int ncx, int ncy, int ncz, int nfx, int nfy, AxisNest ax,
                                                       AxisNest ay, AxisNest az, const double* __restrict__ hx,
                                                       const double* __restrict__ hy, const double* __restrict__ hz,
                                                       const double* __restrict__ lam, const double* __restrict__ mu,
                                                       double* __restrict__ cw) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nc = (long long)ncx * ncy * ncz;
  if (id >= nc) return;
  const int I = id % ncx, J = (id / ncx) % ncy, K = id / ((long long)ncx * ncy);
  double w[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) w[q] = 0.0;
  for (int dk = 0; dk < az.cnum[K]; ++dk)
    for (int dj = 0; dj < ay.cnum[J]; ++dj)
      for (int di = 0; di < ax.cnum[I]; ++di) {
        const int i = ax.cs[I] + di, j = ay.cs[J] + dj, k = az.cs[K] + dk;
        const long long f = i + (long long)nfx * (j + (long long)nfy * k);
        const double V = hx[i] * hy[j] * hz[k];
        const int slot = di + 2 * dj + 4 * dk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (q == slot) {   // select chain: constant register indices
            w[q] = lam[f] * V;
            w[8 + q] = mu[f] * V;
          }
        }
      }
#pragma unroll
  for (int q = 0; q < 16; ++q) cw[q * nc + id] = w[q];
}

// Level-1 split (SplitMat): a coarse cell whose existing level-0 children all carry the same
// (lam, mu) and split every axis into equal halves (or not at all) gets cl = lam V, cm = mu V (the
// sums of its child weights); a cell with stiffness that is not such a cell is heterogeneous
// (het = 1, zero weights); a cell without stiffness gets zeros and het = 0.
__global__ __launch_bounds__(256) void k_level1_split(int ncx, int ncy, int ncz, int nfx, int nfy, AxisNest ax,
                                                      AxisNest ay, AxisNest az, const double* __restrict__ lam,
                                                      const double* __restrict__ mu, const double* __restrict__ cw,
                                                      double* __restrict__ cl, double* __restrict__ cm,
                                                      uint8_t* __restrict__ het) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nc = (long long)ncx * ncy * ncz;
  if (id >= nc) return;
  const int I = id % ncx, J = (id / ncx) % ncy, K = id / ((long long)ncx * ncy);
  auto halves = [](const AxisNest& a, int C) {
    if (a.cnum[C] != 2) return true;
    const double a0 = a.a[a.cs[C]], a1 = a.a[a.cs[C] + 1];
    return fabs(a0 - a1) <= 1e-12 * (a0 + a1);
  };
  bool same = halves(ax, I) && halves(ay, J) && halves(az, K);
  double l0 = 0.0, m0 = 0.0;
  bool first = true;
  for (int dk = 0; dk < az.cnum[K]; ++dk)
    for (int dj = 0; dj < ay.cnum[J]; ++dj)
      for (int di = 0; di < ax.cnum[I]; ++di) {
        const long long f = (ax.cs[I] + di) + (long long)nfx * ((ay.cs[J] + dj) + (long long)nfy * (az.cs[K] + dk));
        if (first) {
          l0 = lam[f];
          m0 = mu[f];
          first = false;
        } else {
          same = same && lam[f] == l0 && mu[f] == m0;
        }
      }
  double sl = 0.0, sm = 0.0;
  for (int q = 0; q < 8; ++q) {
    sl += cw[q * nc + id];
    sm += cw[(8 + q) * nc + id];
  }
  const bool stiff = sl != 0.0 || sm != 0.0;
  cl[id] = (stiff && same) ? sl : 0.0;
  cm[id] = (stiff && same) ? sm : 0.0;
  het[id] = (stiff && !same) ? 1 : 0;
}

// Level l+1 moments from level-l moments (shift of the child moments into the parent).
__global__ __launch_bounds__(256) void k_moments_from_moments(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz, AxisNest ax, AxisNest ay,
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
__device__ __forceinline__ void corner_block(int l, const double* mL, const double* mM, double lam, double mu,
                                             double hx, double hy, double hz, double* D) {
#pragma unroll
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
__global__ __launch_bounds__(256) void k_diag_structured(int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
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
  // fully unrolled: the corner index l is a compile-time constant in every copy of the body,
  // so the element vectors stay in registers
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
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
__global__ __launch_bounds__(256) void k_diag_irregular(int nirr, const int* __restrict__ nodes, const int* __restrict__ ijk,
                                 const double* __restrict__ lam, const double* __restrict__ mu,
                                 const double* __restrict__ hx, const double* __restrict__ hy,
                                 const double* __restrict__ hz, double* __restrict__ D) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nirr) return;
  const int i = ijk[3 * e], j = ijk[3 * e + 1], k = ijk[3 * e + 2];
#pragma unroll
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

// 3x3 node blocks of rank-one terms k w w^T: one wavefront per term.  The DOFs of a term are
// sorted and unique, so the DOFs of one node are at most two positions apart.
__global__ __launch_bounds__(256) void k_diag_rank1(int nterm, const int* __restrict__ ptr, const int* __restrict__ dof,
                                                    const double* __restrict__ w, const double* __restrict__ k,
                                                    double* __restrict__ D) {
  const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= nterm) return;
  const int q0 = ptr[t], q1 = ptr[t + 1];
  const double kt = k[t];
  for (int a = q0 + lane; a < q1; a += 64) {
    const int da = dof[a], na = da / 3;
    const double ka = kt * w[a];
    for (int b = max(q0, a - 2); b < min(q1, a + 3); ++b) {
      const int db = dof[b];
      if (db / 3 == na) atomicAdd(&D[9LL * na + 3 * (da % 3) + db % 3], ka * w[b]);
    }
  }
}

// Rank-one terms into the dense coarsest matrix: one block per term, threads over the (a, b) pairs.
__global__ __launch_bounds__(256) void k_dense_add_rank1(const int* __restrict__ ptr, const int* __restrict__ dof,
                                                         const double* __restrict__ w, const double* __restrict__ k,
                                                         int n, double* __restrict__ A) {
  const int t = blockIdx.x;
  const int q0 = ptr[t], s = ptr[t + 1] - q0;
  const double kt = k[t];
  for (long long p = threadIdx.x; p < (long long)s * s; p += blockDim.x) {
    const int a = q0 + (int)(p / s), b = q0 + (int)(p % s);
    atomicAdd(&A[(long long)dof[a] * n + dof[b]], kt * w[a] * w[b]);
  }
}

// ---- level-0 cell arrays from the reference-order description ------------------------------
// Reference cell id c = (i ny + j) nz + k (fe/solver.py:128) -> internal x-fastest id
// x = i + nx (j + ny k).  Label mode (lab_ref != nullptr): labels 255 = no stiffness, materials
// from the tables; value mode: per-cell (lam, mu) and activity.  Writes the real labels, FP64
// materials and activity in x-fastest order.
__global__ void k_prep_cells(int nx, int ny, int nz, const uint8_t* __restrict__ lab_ref,
                             const double* __restrict__ tabL, const double* __restrict__ tabM,
                             const double* __restrict__ lam_ref, const double* __restrict__ mu_ref,
                             const uint8_t* __restrict__ act_ref, uint8_t* __restrict__ lab_x,
                             double* __restrict__ matL, double* __restrict__ matM, uint8_t* __restrict__ act_x) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncell = (long long)nx * ny * nz;
  if (c >= ncell) return;
  const long long k = c % nz, j = (c / nz) % ny, i = c / ((long long)ny * nz);
  const long long x = i + (long long)nx * (j + (long long)ny * k);
  bool act;
  double l = 0.0, m = 0.0;
  if (lab_ref) {
    const uint8_t lb = lab_ref[c];
    act = lb != 255;
    if (act) { l = tabL[lb]; m = tabM[lb]; }
    lab_x[x] = lb;
  } else {
    act = act_ref[c] != 0;
    if (act) { l = lam_ref[c]; m = mu_ref[c]; }
  }
  matL[x] = l;
  matM[x] = m;
  act_x[x] = act ? 1 : 0;
}

// Irregular cells (x ids) leave the structured operator: no label, zero apply materials.
__global__ void k_prep_irregular(int nirr, const int* __restrict__ irr_x, uint8_t* __restrict__ lab_apply,
                                 double* __restrict__ applyL, double* __restrict__ applyM, uint8_t* __restrict__ irr) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nirr) return;
  const long long x = irr_x[e];
  if (lab_apply) lab_apply[x] = 255;
  applyL[x] = 0.0;
  applyM[x] = 0.0;
  irr[x] = 1;
}

// Used grid nodes (fe/solver.py:141, rb/splitfe.py:73): corners of active, non-irregular cells.
__global__ void k_used_grid(int nx, int ny, int nz, const uint8_t* __restrict__ act_x, const uint8_t* __restrict__ irr,
                            uint8_t* __restrict__ used) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nxn = nx + 1, nyn = ny + 1;
  if (n >= nxn * nyn * (nz + 1)) return;
  const int i = (int)(n % nxn), j = (int)((n / nxn) % nyn), k = (int)(n / (nxn * nyn));
  uint8_t u = 0;
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        const int ci = i - a, cj = j - b, ck = k - c;
        if (ci < 0 || ci >= nx || cj < 0 || cj >= ny || ck < 0 || ck >= nz) continue;
        const long long x = ci + (long long)nx * (cj + (long long)ny * ck);
        if (act_x[x] && !irr[x]) u = 1;
      }
  used[n] = u;
}

// used[nodes[q]] = 1 (the 8 nodes of every irregular cell; the same value from every writer)
__global__ void k_mark_nodes(long long n, const int* __restrict__ nodes, uint8_t* __restrict__ used) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) used[nodes[q]] = 1;
}

// Inverse of the 3x3 blocks (symmetric), stored as 6 values per node: 00 01 02 11 12 22 (the
// off-diagonal pairs averaged).  Blocks with |det| < 1e-30 become the identity (fe/gmg.py:39-41).
template <typename T>
__global__ void k_block_inverse(long long nn, const double* __restrict__ D, T* __restrict__ Dinv) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double* m = D + 9 * n;
  T* o = Dinv + 6 * n;
  const double c0 = m[4] * m[8] - m[5] * m[7];
  const double c1 = m[5] * m[6] - m[3] * m[8];
  const double c2 = m[3] * m[7] - m[4] * m[6];
  const double det = m[0] * c0 + m[1] * c1 + m[2] * c2;
  if (fabs(det) < 1e-30) {
    o[0] = T(1); o[1] = T(0); o[2] = T(0); o[3] = T(1); o[4] = T(0); o[5] = T(1);
    return;
  }
  const double id = 1.0 / det;
  const double i01 = (m[2] * m[7] - m[1] * m[8]) * id, i10 = c1 * id;
  const double i02 = (m[1] * m[5] - m[2] * m[4]) * id, i20 = c2 * id;
  const double i12 = (m[2] * m[3] - m[0] * m[5]) * id, i21 = (m[1] * m[6] - m[0] * m[7]) * id;
  o[0] = T(c0 * id);
  o[1] = T(0.5 * (i01 + i10));
  o[2] = T(0.5 * (i02 + i20));
  o[3] = T((m[0] * m[8] - m[2] * m[6]) * id);
  o[4] = T(0.5 * (i12 + i21));
  o[5] = T((m[0] * m[4] - m[1] * m[3]) * id);
}

// Split-copy group inverses, one thread per group: W = the group's principal submatrix of A
// (off-diagonal member blocks from the host in off, diagonal blocks from D), inverted in place by
// Gauss-Jordan without pivoting (W is SPD), symmetrised into ginv; the members' 3x3 slots of Dinv
// become zero.  A non-positive pivot keeps the members' own 3x3 inverses (block diagonal ginv).
template <typename T>
__global__ void k_group_inverse(int ng, const int* __restrict__ gptr, const int* __restrict__ gslot,
                                const int* __restrict__ tnode, const int* __restrict__ goff,
                                const double* __restrict__ off, const double* __restrict__ D,
                                double* __restrict__ W, T* __restrict__ Dinv, T* __restrict__ ginv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int b = gptr[g], m = gptr[g + 1] - b, N = 3 * m;
  double* w = W + goff[g];
  T* gi = ginv + goff[g];
  for (int k = 0; k < N * N; ++k) w[k] = off[goff[g] + k];
  for (int j = 0; j < m; ++j) {
    const double* dm = D + 9 * (long long)tnode[gslot[b + j]];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) w[(3 * j + p) * N + 3 * j + q] = dm[3 * p + q];
  }
  bool ok = true;
  for (int k = 0; k < N && ok; ++k) {
    const double piv = w[k * N + k];
    if (!(piv > 0.0) || !isfinite(piv)) { ok = false; break; }
    const double ip = 1.0 / piv;
    w[k * N + k] = 1.0;
    for (int j = 0; j < N; ++j) w[k * N + j] *= ip;
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const double f = w[i * N + k];
      if (f == 0.0) continue;
      w[i * N + k] = 0.0;
      for (int j = 0; j < N; ++j) w[i * N + j] -= f * w[k * N + j];
    }
  }
  if (ok) {
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) gi[i * N + j] = T(0.5 * (w[i * N + j] + w[j * N + i]));
  } else {
    for (int k = 0; k < N * N; ++k) gi[k] = T(0);
    for (int j = 0; j < m; ++j) {
      const T* o = Dinv + 6 * (long long)tnode[gslot[b + j]];
      const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) gi[(3 * j + p) * N + 3 * j + q] = o[idx[p][q]];
    }
  }
  for (int j = 0; j < m; ++j) {
    T* o = Dinv + 6 * (long long)tnode[gslot[b + j]];
    for (int k = 0; k < 6; ++k) o[k] = T(0);
  }
}

// Dense coarsest operator: one block of 24 threads per coarse cell, thread = element column.
__global__ __launch_bounds__(64) void k_dense_assemble(int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
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
#pragma unroll
  for (int q = 0; q < 24; ++q) u[q] = (q == col) ? 1.0 : 0.0;   // no runtime index: stays in registers
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
//   Dk = A_KK^-1;  A_KJ <- Dk A_KJ;  A_IJ -= A_IK A_KJ;  A_IK <- -A_IK Dk;  A_KK <- Dk
// which is GJB consecutive steps of scalar Gauss-Jordan.  Pivot blocks start at multiples of
// GJB, so the K rows/columns are exactly one 64-wide tile of every kernel below.
//
// Symmetry: starting from a symmetric A, after the pivot blocks P have been processed
// A_JI = s A_IJ^T with s = -1 when exactly one of I, J is in P (else +1), so only the tiles I >= J
// are kept.  With pivots taken in order (P = the blocks before K) and C_I = A_IK (old), the old row
// is A_KJ = t_J C_J^T with t_J = +1 for J > K, -1 for J < K, and with W_I = C_I Dk the step is
//   A_IJ -= t_J W_I C_J^T  (I >= J, both != K);   A_IK <- -W_I (I > K);   A_KI <- -W_I^T (I < K)
// -- half the trailing update of the unsymmetric form.  At the end every block is processed, A is
// the symmetric inverse, and the upper tiles are filled from the lower ones.
constexpr int GJB = 64;

// (1) Dk = inverse of the nb x nb diagonal block at k0, padded with the identity to 64 x 64
// (padding rows/columns stay the identity through the elimination).  256 threads: thread
// (g = t / 64, j = t % 64) keeps M[g + 4r][j], r = 0..15, in registers; each of the 64 steps
// publishes the pivot row and column through LDS (double-buffered, one barrier per step).  The
// step loop stays rolled (a one-pass kernel runs its code once, so a fully unrolled body is
// fetched cold); the pivot row's register is picked with a select chain, and the update of the
// 16 entries is branch-free (all three candidate values computed, then selected) so that the
// column reads are issued together instead of one wait per row.
__global__ __launch_bounds__(256) void k_gjb_diag(int n, int k0, int nb, const double* __restrict__ A,
                                                  double* __restrict__ Dk) {
  __shared__ double rowk[2][GJB], colk[2][GJB];
  const int j = threadIdx.x & 63, g = threadIdx.x >> 6;
  double m[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = g + 4 * r;
    m[r] = (i < nb && j < nb) ? A[(long long)(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
#pragma unroll 1
  for (int k = 0; k < GJB; ++k) {
    const int buf = k & 1, kr = k >> 2;
    if (g == (k & 3)) {
      double v = m[0];
#pragma unroll
      for (int r = 1; r < 16; ++r) v = (r == kr) ? m[r] : v;
      rowk[buf][j] = v;
    }
    if (j == k) {
#pragma unroll
      for (int r = 0; r < 16; ++r) colk[buf][g + 4 * r] = m[r];
    }
    __syncthreads();
    double cv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) cv[r] = colk[buf][g + 4 * r];
    const double ip = 1.0 / rowk[buf][k];
    const double rk = rowk[buf][j] * ip;       // new pivot row entry (j != k)
    const bool jk = (j == k);
    const double rowv = jk ? ip : rk;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const double upd = jk ? -cv[r] * ip : m[r] - cv[r] * rk;
      m[r] = (g + 4 * r == k) ? rowv : upd;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) Dk[(g + 4 * r) * GJB + j] = m[r];
}

// 64x64x64 product of two LDS tiles, 16x16 threads with 4x4 outputs each
// (rows tr + 16a, columns tc + 16b).
__device__ __forceinline__ void gjb_tile_mul(const double (*X)[GJB + 1], const double (*Y)[GJB + 1], int tr, int tc,
                                             double acc[4][4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
#pragma unroll 4
  for (int q = 0; q < GJB; ++q) {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = X[tr + 16 * a][q];
#pragma unroll
    for (int b = 0; b < 4; ++b) y[b] = Y[q][tc + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
}

// (2) Panels of pivot block K, one 64-row tile I per block: C_I = A_IK (from the kept tile (I, K),
// or -(A_KI)^T from the kept tile (K, I) when I < K) and W_I = C_I Dk, stored k-major
// (Ct[c][i] = C_I[i][c], Wt likewise, 64 x n) for the update; then the new K tiles of the kept
// triangle: A_IK = -W_I (I > K), A_KI = -W_I^T (I < K), A_KK = Dk.  Each block reads and writes
// only its own tile of the K row/column.
__global__ __launch_bounds__(256) void k_gjs_panel(int n, int k0, int nb, double* __restrict__ A,
                                                   const double* __restrict__ Dk, double* __restrict__ Ct,
                                                   double* __restrict__ Wt) {
  __shared__ double sc[GJB][GJB + 1], sd[GJB][GJB + 1];
  const int i0 = blockIdx.x * 64;
  if (i0 == k0) {
    for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
      const int q = t >> 6, c = t & 63;
      if (q < nb && c < nb) A[(long long)(k0 + q) * n + k0 + c] = Dk[q * GJB + c];
      if (k0 + c < n) { Ct[(long long)q * n + k0 + c] = 0.0; Wt[(long long)q * n + k0 + c] = 0.0; }
    }
    return;
  }
  const bool below = i0 > k0;
  for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
    const int r = t >> 6, c = t & 63;
    // below: row-major tile (I, K); above: tile (K, I) read transposed (c rows, r columns)
    double v = 0.0;
    if (i0 + r < n && c < nb)
      v = below ? A[(long long)(i0 + r) * n + k0 + c] : -A[(long long)(k0 + c) * n + i0 + r];
    sc[r][c] = v;
    sd[r][c] = (r < nb && c < nb) ? Dk[r * GJB + c] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4];
  gjb_tile_mul(sc, sd, tr, tc, acc);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int r = tr + 16 * a, i = i0 + r;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int c = tc + 16 * b;
      Ct[(long long)c * n + i] = sc[r][c];
      Wt[(long long)c * n + i] = acc[a][b];
      if (c < nb) {
        if (below) A[(long long)i * n + k0 + c] = -acc[a][b];
        else A[(long long)(k0 + c) * n + i] = -acc[a][b];
      }
    }
  }
}

// (3) A_IJ -= t_J W_I C_J^T for every kept 64 x 64 tile (I >= J) outside row and column K, on the
// FP64 matrix cores.  sa[k][i] = W_I[i][k], sb[k][j] = C_J[j][k] straight from the k-major panels;
// wave w computes the 32 x 32 quarter (w / 2, w % 2) as 2 x 2 products of
// v_mfma_f64_16x16x4f64 over the 64-deep K.  Operand lanes: A[l % 16][l / 16], B[l / 16][l % 16];
// result lane l, register r: D[l / 16 + 4 r][l % 16] (the layout of the builtin on gfx950).
typedef double gjb_d4 __attribute__((ext_vector_type(4)));
constexpr int GJP = GJB + 4;   // LDS row pitch (doubles)
__global__ __launch_bounds__(256) void k_gjs_update(int n, int k0, const double* __restrict__ Ct,
                                                    const double* __restrict__ Wt, double* __restrict__ A) {
  __shared__ double sa[GJB][GJP];
  __shared__ double sb[GJB][GJP];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  if (j0 > i0 || i0 == k0 || j0 == k0) return;
  for (int t = threadIdx.x; t < 64 * GJB; t += blockDim.x) {
    const int q = t >> 6, r = t & 63;
    sa[q][r] = (i0 + r < n) ? Wt[(long long)q * n + i0 + r] : 0.0;
    sb[q][r] = (j0 + r < n) ? Ct[(long long)q * n + j0 + r] : 0.0;
  }
  __syncthreads();
  const double tj = j0 > k0 ? 1.0 : -1.0;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = l & 15, lk = l >> 4;
  gjb_d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = gjb_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k4 = 0; k4 < GJB / 4; ++k4) {
    const int kk = 4 * k4 + lk;
    double av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) av[a] = sa[kk][wr + 16 * a + lr];
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = sb[kk][wc + 16 * b + lr];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = j0 + wc + 16 * b + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * a + lk + 4 * r;
        if (i < n && j < n) A[(long long)i * n + j] -= tj * acc[a][b][r];
      }
    }
}

// (4) The finished inverse is symmetric: copy every kept tile below the diagonal to its mirror.
// 64 x 64 tiles (I > J) through LDS, so that reads and writes are both row-contiguous.
__global__ __launch_bounds__(256) void k_gjs_mirror(int n, double* __restrict__ A) {
  __shared__ double st[GJB][GJB + 1];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  if (j0 >= i0) return;
  for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
    const int r = t >> 6, c = t & 63;
    st[r][c] = (i0 + r < n && j0 + c < n) ? A[(long long)(i0 + r) * n + j0 + c] : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
    const int r = t >> 6, c = t & 63;     // destination (j0 + r, i0 + c) = source (i0 + c, j0 + r)
    if (j0 + r < n && i0 + c < n) A[(long long)(j0 + r) * n + i0 + c] = st[c][r];
  }
}

// Coarse solve x = Ainv b (dense FP64), one 256-thread block per row: four waves share a row of a
// few thousand doubles, four times the loads in flight of one wavefront per row.
// The inverse is over the level's used DOFs only: map[i] = full DOF of compact DOF i (nullptr: the
// identity).  x at DOFs outside the map is not written (kept zero).
template <typename T>
__device__ __forceinline__ void dense_gemv_body(const int row, int n, const double* __restrict__ Ainv,
                                                const int* __restrict__ map, const T* __restrict__ b,
                                                T* __restrict__ x) {
  if (row >= n) return;   // block-uniform
  const double* r = Ainv + (long long)row * n;
  double s = 0.0;
  for (int c = threadIdx.x; c < n; c += 256) s += r[c] * (double)b[map ? map[c] : c];
  s = block_sum<256>(s);
  if (threadIdx.x == 0) x[map ? map[row] : row] = T(s);
}
template <typename T>
__global__ __launch_bounds__(256) void k_dense_gemv(int n, const double* __restrict__ Ainv,
                                                    const int* __restrict__ map, const T* __restrict__ b,
                                                    T* __restrict__ x) {
  dense_gemv_body<T>(blockIdx.x, n, Ainv, map, b, x);
}
template <typename T>
struct GemvL {
  int n;
  const double* Ainv;
  const int* map;
  const T* b;
  T* x;
  __device__ __forceinline__ void run(int bid, int) const { dense_gemv_body<T>(bid, n, Ainv, map, b, x); }
};

// Ac[i][j] = A[map[i]][map[j]] (m x m from n x n)
__global__ void k_dense_gather(int m, const int* __restrict__ map, int n, const double* __restrict__ A,
                               double* __restrict__ Ac) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)m * m) return;
  const int i = (int)(id / m), j = (int)(id % m);
  Ac[id] = A[(long long)map[i] * n + map[j]];
}

}  // namespace cm
