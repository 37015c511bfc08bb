// Post-processing kernels: cell-centre strain, stress and von Mises (fe/solver.py:257-272,
// rb/splitfe.py:161-175) and the analysis-grid resampling of rb/rbsim.py:110-140.
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
                              double* __restrict__ vm, double* __restrict__ eps) {
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
  cell_centre_stress<double>(ue, 0.25 / hx[i], 0.25 / hy[j], 0.25 / hz[k], lam[c], mu[c], ep, sg, &v);
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
  // FP32 mean of the 8 node values, summed pairwise ((0+1)+(2+3))+((4+5)+(6+7)) as numpy sums
  // eight float32 values
  float v[8][3];
  for (int l = 0; l < 8; ++l) {
    const long long n = cell_nodes[8 * c + l];
    for (int q = 0; q < 3; ++q) v[l][q] = u_nodes[3 * n + q];
  }
  for (int q = 0; q < 3; ++q) {
    const float s = ((v[0][q] + v[1][q]) + (v[2][q] + v[3][q])) + ((v[4][q] + v[5][q]) + (v[6][q] + v[7][q]));
    uc16[3 * c + q] = __float2half(s / 8.0f);
  }
  lab8[c] = (int8_t)lab[c];
  for (int k = 0; k < ncls; ++k) {
    // FP32 product, round half to even (numpy's round of a float32 array)
    frac8[c * ncls + k] = (uint8_t)rintf(255.0f * frac[c * ncls + k]);
  }
}

}  // namespace cm
