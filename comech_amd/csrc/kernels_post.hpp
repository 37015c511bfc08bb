// Post-processing kernels: cell-centre strain, stress and von Mises (fe/solver.py:257-272,
// rb/splitfe.py:161-175) and the analysis-grid resampling of rb/rbsim.py:110-161.
#pragma once
#include <hip/hip_fp16.h>

#include "common.hpp"
#include "element.hpp"

namespace cm {

// cells: reference cell ids (z-fastest, fe/solver.py:128); nodes8: explicit node ids per listed
// cell (split connectivity) or nullptr for the grid's own corners; lam/mu indexed by reference id.
__global__ void k_cell_stress(long long n, const long long* __restrict__ cells, const long long* __restrict__ nodes8,
                              int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
                              const double* __restrict__ hz, const double* __restrict__ lam,
                              const double* __restrict__ mu, const double* __restrict__ U, double* __restrict__ sig,
                              double* __restrict__ vm, double* __restrict__ eps,
                              const double* __restrict__ coeffs = nullptr) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long long c = cells[e];
  const int k = c % nz, j = (c / nz) % ny, i = c / ((long long)ny * nz);
  double ue[24];
  for (int l = 0; l < 8; ++l) {
    long long node;
    if (nodes8) {
      node = nodes8[8 * e + l];
    } else {
      const int a = l & 1, b = (l >> 1) & 1, cc = l >> 2;
      node = (i + a) + (long long)(nx + 1) * ((j + b) + (long long)(ny + 1) * (k + cc));
    }
    for (int q = 0; q < 3; ++q) {
      const double v = U[3 * node + q];
      ue[3 * l + q] = (v == v) ? v : 0.0;   // NaN (unused node) -> 0, rb/splitfe.py:162
    }
  }
  double ep[6], sg[6], v;
  cell_centre_stress<double>(ue, 0.25 / hx[i], 0.25 / hy[j], 0.25 / hz[k], coeffs ? 0.0 : lam[c], coeffs ? 0.0 : mu[c],
                             ep, sg, &v);
  if (coeffs) {
    // z-orthotropic D = sum_k c_k P_k (fe/solver.py:43-52, 267-269): normal block [[c0 c1 c2] [c1 c0 c2]
    // [c2 c2 c3]], shear c4 (xy), c5 (yz, xz); von Mises of fe/solver.py:270-271
    const double* cc = coeffs + 6 * c;
    sg[0] = cc[0] * ep[0] + cc[1] * ep[1] + cc[2] * ep[2];
    sg[1] = cc[1] * ep[0] + cc[0] * ep[1] + cc[2] * ep[2];
    sg[2] = cc[2] * ep[0] + cc[2] * ep[1] + cc[3] * ep[2];
    sg[3] = cc[4] * ep[3];
    sg[4] = cc[5] * ep[4];
    sg[5] = cc[5] * ep[5];
    const double dxy = sg[0] - sg[1], dyz = sg[1] - sg[2], dzx = sg[2] - sg[0];
    v = sqrt(0.5 * (dxy * dxy + dyz * dyz + dzx * dzx) + 3.0 * (sg[3] * sg[3] + sg[4] * sg[4] + sg[5] * sg[5]));
  }
  for (int q = 0; q < 6; ++q) { sig[6 * e + q] = sg[q]; eps[6 * e + q] = ep[q]; }
  vm[e] = v;
}

// Nodal element forces K_e u_e of listed cells (rb/splitfe.py:178-193), out[n][24].
template <bool SRI>
__global__ void k_element_forces(long long n, const long long* __restrict__ cells, const long long* __restrict__ nodes8,
                                 int nx, int ny, int nz, const double* __restrict__ hx, const double* __restrict__ hy,
                                 const double* __restrict__ hz, const double* __restrict__ lam,
                                 const double* __restrict__ mu, const double* __restrict__ U, double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long long c = cells[e];
  const int k = c % nz, j = (c / nz) % ny, i = c / ((long long)ny * nz);
  double ue[24], f[24];
  for (int l = 0; l < 8; ++l) {
    long long node;
    if (nodes8) {
      node = nodes8[8 * e + l];
    } else {
      const int a = l & 1, b = (l >> 1) & 1, cc = l >> 2;
      node = (i + a) + (long long)(nx + 1) * ((j + b) + (long long)(ny + 1) * (k + cc));
    }
    for (int q = 0; q < 3; ++q) {
      const double v = U[3 * node + q];
      ue[3 * l + q] = (v == v) ? v : 0.0;
    }
  }
  const double V = hx[i] * hy[j] * hz[k];
  CellW<double> w{0.25 / hx[i], 0.25 / hy[j], 0.25 / hz[k], lam[c] * V, mu[c] * V};
  elem_apply_const<double, SRI>(ue, f, w);
  for (int q = 0; q < 24; ++q) out[24 * e + q] = f[q];
}

// sample_fe (rb/rbsim.py:110-140): every analysis cell averages sub^3 sample points.  Per axis the
// host supplies, for every sample coordinate, the FE cell index, the local coordinate t in [0,1]
// and whether the sample lies inside the FE box.  One thread per analysis cell.
__global__ void k_sample_fe(int mx, int my, int mz, int sub, int nx, int ny, int nz, const int* __restrict__ si,
                            const double* __restrict__ st, const uint8_t* __restrict__ sin_, const int* __restrict__ sj,
                            const double* __restrict__ su, const uint8_t* __restrict__ sjn, const int* __restrict__ sk,
                            const double* __restrict__ sv, const uint8_t* __restrict__ skn,
                            const uint8_t* __restrict__ active, const long long* __restrict__ cn,
                            const float* __restrict__ VM, const float* __restrict__ SIG,
                            const double* __restrict__ Urel, float* __restrict__ vm_out, float* __restrict__ u_out,
                            double* __restrict__ cover, float* __restrict__ sig_out) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncm = (long long)mx * my * mz;
  if (id >= ncm) return;
  // analysis cell order (X, Y, Z) -> id = (X*my + Y)*mz + Z (fe/solver.py:128)
  const int Z = id % mz, Y = (id / mz) % my, X = id / ((long long)my * mz);
  double vms = 0.0, us[3] = {0, 0, 0}, ss[6] = {0, 0, 0, 0, 0, 0}, okc = 0.0;
  for (int a = 0; a < sub; ++a) {
    const int px = X * sub + a;
    for (int b = 0; b < sub; ++b) {
      const int py = Y * sub + b;
      for (int c = 0; c < sub; ++c) {
        const int pz = Z * sub + c;
        const int i = si[px], j = sj[py], k = sk[pz];
        const long long cell = ((long long)i * ny + j) * nz + k;
        const bool ok = sin_[px] && sjn[py] && skn[pz] && active[cell];
        if (!ok) continue;
        const double tx = st[px], ty = su[py], tz = sv[pz];
        okc += 1.0;
        vms += VM[cell];
        for (int q = 0; q < 6; ++q) ss[q] += SIG[6 * cell + q];
        for (int l = 0; l < 8; ++l) {
          const double w = ((l & 1) ? tx : 1 - tx) * (((l >> 1) & 1) ? ty : 1 - ty) * ((l >> 2) ? tz : 1 - tz);
          const long long nd = cn[8 * cell + l];
          for (int q = 0; q < 3; ++q) {
            const double v = Urel[3 * nd + q];
            us[q] += w * ((v == v) ? v : 0.0);
          }
        }
      }
    }
  }
  const double den = okc > 1.0 ? okc : 1.0;
  vm_out[id] = float(vms / den);
  for (int q = 0; q < 3; ++q) u_out[3 * id + q] = float(us[q] / den);
  for (int q = 0; q < 6; ++q) sig_out[6 * id + q] = float(ss[q] / den);
  cover[id] = okc / (double)(sub * sub * sub);
}

// Bank record of one action (exp/bank.py:117-121): von Mises and the cell-mean relative
// displacement (mean over the analysis cell's 8 nodes, exp/bank.py:116) in FP16, labels as int8,
// tissue fractions as round(255 f) in uint8.  One thread per analysis cell; u_nodes: [nnode][3],
// cell_nodes: [ncell][8], frac: [ncell][ncls].
__global__ __launch_bounds__(256) void k_bank_pack(long long ncell, int ncls, const float* __restrict__ vm,
                                                   const float* __restrict__ u_nodes,
                                                   const long long* __restrict__ cell_nodes,
                                                   const int* __restrict__ lab, const float* __restrict__ frac,
                                                   __half* __restrict__ vm16, __half* __restrict__ uc16,
                                                   int8_t* __restrict__ lab8, uint8_t* __restrict__ frac8) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  vm16[c] = __float2half(vm[c]);
  // FP32 mean of the 8 node values: numpy's mean over axis 1 of a (ncell, 8, 3) float32 array
  // accumulates the node values in index order, v0 + v1 + ... + v7, in float32, then divides by 8
  float s[3] = {0.f, 0.f, 0.f};
  for (int l = 0; l < 8; ++l) {
    const long long n = cell_nodes[8 * c + l];
    for (int q = 0; q < 3; ++q) s[q] = (l == 0) ? u_nodes[3 * n + q] : s[q] + u_nodes[3 * n + q];
  }
  for (int q = 0; q < 3; ++q) uc16[3 * c + q] = __float2half(s[q] / 8.0f);
  lab8[c] = (int8_t)lab[c];
  for (int k = 0; k < ncls; ++k) {
    // FP32 product, round half to even (numpy's round of a float32 array)
    frac8[c * ncls + k] = (uint8_t)rintf(255.0f * frac[c * ncls + k]);
  }
}

// ---------------------------------------------------------------------------------------------
// sample_nodes (rb/rbsim.py:143-161): the relative displacement at every node of the analysis grid,
// trilinear in the FE cell of the nearest active FE cell (the containing cell when it is active).
// "Nearest" is the Euclidean feature transform of the inactive mask in cell-index units
// (scipy.ndimage.distance_transform_edt(~active, return_indices=True) at rb/rbsim.py:152); among
// cells at the same squared distance it returns the one with the smallest (k, j, i) (last axis
// first), the rule scipy 1.15.3 follows on every tie of random 3-D rasters.
//
// Stage 1: for every FE cell column (i, j) and every z layer k the analysis nodes query, the
// nearest active cell of that column: squared distance (k - k')^2, or -1 for a column without one,
// and k' (the lower one on a tie).  Layout [query layer q][column i*ny + j].
__global__ void k_column_nearest(int nx, int ny, int nz, int nq, const int* __restrict__ kq,
                                 const uint8_t* __restrict__ active, int* __restrict__ dz2, int* __restrict__ kbest) {
  const long long ncol = (long long)nx * ny;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= ncol * nq) return;
  const int q = (int)(id / ncol);
  const long long col = id - q * ncol;
  const int k = kq[q];
  const uint8_t* a = active + col * nz;
  int best = -1, d2 = -1;
  const int dmax = max(k, nz - 1 - k);
  for (int d = 0; d <= dmax; ++d) {
    if (k - d >= 0 && a[k - d]) { best = k - d; d2 = d * d; break; }
    if (k + d < nz && a[k + d]) { best = k + d; d2 = d * d; break; }
  }
  dz2[id] = d2;
  kbest[id] = best;
}

// numpy.nan_to_num of a double: NaN -> 0, +-inf -> +-largest finite value
__device__ __forceinline__ double nan_to_num(double v) {
  if (v != v) return 0.0;
  if (v == __longlong_as_double(0x7ff0000000000000LL)) return 1.7976931348623157e308;
  if (v == -__longlong_as_double(0x7ff0000000000000LL)) return -1.7976931348623157e308;
  return v;
}

// Stage 2: one thread per analysis node n = a + (mx+1)(b + (my+1)c) (fe/solver.py:120-123).  ci..ck:
// FE cell of every analysis node coordinate per axis (clip(searchsorted(edges, x) - 1)), qz: the
// node layer c's index into the stage-1 query layers.  The interpolation repeats the reference's
// FP64 expressions operation by operation (no fused multiply-add), so the float32 result is the
// reference's bit for bit.
__global__ void k_sample_nodes(int mx1, int my1, int mz1, const double* __restrict__ mxe, const double* __restrict__ mye,
                               const double* __restrict__ mze, const int* __restrict__ ci, const int* __restrict__ cj,
                               const int* __restrict__ ck, const int* __restrict__ qz, int nx, int ny, int nz,
                               const double* __restrict__ xe, const double* __restrict__ ye,
                               const double* __restrict__ ze, const uint8_t* __restrict__ active,
                               const int* __restrict__ dz2, const int* __restrict__ kbest,
                               const long long* __restrict__ cn, const double* __restrict__ Urel,
                               float* __restrict__ out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long mnode = (long long)mx1 * my1 * mz1;
  if (n >= mnode) return;
  const int a = (int)(n % mx1), b = (int)((n / mx1) % my1), c = (int)(n / ((long long)mx1 * my1));
  const int i = ci[a], j = cj[b], k = ck[c];
  int bi = i, bj = j, bk = k;
  if (!active[((long long)i * ny + j) * nz + k]) {
    // smallest (d2, k', j', i') over all columns; a column farther in xy than the best d2 cannot win
    const long long ncol = (long long)nx * ny;
    const int* D = dz2 + (long long)qz[c] * ncol;
    const int* Kb = kbest + (long long)qz[c] * ncol;
    long long bd2 = 0x7fffffffffffffffLL;
    bi = bj = bk = -1;
    for (int jj = 0; jj < ny; ++jj) {
      const long long dj2 = (long long)(j - jj) * (j - jj);
      if (dj2 > bd2) continue;
      for (int ii = 0; ii < nx; ++ii) {
        const long long col = (long long)ii * ny + jj;
        const int dz = D[col];
        if (dz < 0) continue;
        const long long d2 = (long long)(i - ii) * (i - ii) + dj2 + dz;
        if (d2 > bd2) continue;
        const int kk = Kb[col];
        if (d2 < bd2 || kk < bk || (kk == bk && (jj < bj || (jj == bj && ii < bi)))) {
          bd2 = d2; bi = ii; bj = jj; bk = kk;
        }
      }
    }
    if (bi < 0) {   // no active cell at all (the host rejects that case before the launch)
      for (int q = 0; q < 3; ++q) out[3 * n + q] = 0.0f;
      return;
    }
  }
  // tx = clip((P - xe[ci]) / dx[ci], 0, 1) with dx = np.diff(xe)
  const double tx = fmin(fmax(__ddiv_rn(__dsub_rn(mxe[a], xe[bi]), __dsub_rn(xe[bi + 1], xe[bi])), 0.0), 1.0);
  const double ty = fmin(fmax(__ddiv_rn(__dsub_rn(mye[b], ye[bj]), __dsub_rn(ye[bj + 1], ye[bj])), 0.0), 1.0);
  const double tz = fmin(fmax(__ddiv_rn(__dsub_rn(mze[c], ze[bk]), __dsub_rn(ze[bk + 1], ze[bk])), 0.0), 1.0);
  const long long cell = ((long long)bi * ny + bj) * nz + bk;
  double u[3] = {0.0, 0.0, 0.0};
  for (int l = 0; l < 8; ++l) {
    const double wx = (l & 1) ? tx : __dsub_rn(1.0, tx);
    const double wy = ((l >> 1) & 1) ? ty : __dsub_rn(1.0, ty);
    const double wz = (l >> 2) ? tz : __dsub_rn(1.0, tz);
    const double w = __dmul_rn(__dmul_rn(wx, wy), wz);
    const long long nd = cn[8 * cell + l];
    for (int q = 0; q < 3; ++q) u[q] = __dadd_rn(u[q], __dmul_rn(w, nan_to_num(Urel[3 * nd + q])));
  }
  for (int q = 0; q < 3; ++q) out[3 * n + q] = (float)u[q];
}

}  // namespace cm
