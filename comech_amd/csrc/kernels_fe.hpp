// Matrix-free operator kernels for the voxel hex8 system on a rectilinear grid.
//
// Layout in HBM (all levels):
//   node vectors  AoS, 3 components per node, node id i + (nx+1)(j + (ny+1)k) -- the reference's
//                 grid numbering (fe/solver.py:107); level 0 appends the split-node copies after
//                 the grid nodes (ids nnode + rank, rb/splitfe.py:268-280)
//   cells         x-fastest internal order i + nx (j + ny k); the boundary converts from the
//                 reference's z-fastest cell ids (fe/solver.py:104-105,128)
//   level 0       one uint8 label per cell (255 = no stiffness) + a (lam, mu) table, or two
//                 per-cell values; irregular cells (touching split copies) are applied by a
//                 separate list kernel
//   level >= 1    27 lam-moments + 27 mu-moments per coarse cell (structure of arrays)
//
// The structured apply marches along z (2.5-D blocking): a block of 16x16 threads owns a 15x15
// column of output nodes; every z step stages one node layer of u (17x17 nodes) in LDS, computes
// the 16x16 cells of one cell layer once, carries the cells' top-corner forces to the next step
// in registers and gathers the four in-plane neighbours' corner forces through LDS.
#pragma once
#include "common.hpp"
#include "element.hpp"

namespace cm {

constexpr int TILE = 16;          // threads per axis in the xy tile (cells)
constexpr int TOUT = TILE - 1;    // output nodes per axis
constexpr int TN = TILE + 1;      // staged nodes per axis
constexpr int BLOCK = TILE * TILE;
constexpr uint8_t LAB_NONE = 255;

// -------------------------------------------------------------------------------------------
// Cell material policies
template <typename T, bool SRI>
struct FineLabelCells {
  const uint8_t* __restrict__ lab;   // x-fastest
  const T* __restrict__ tabL;        // lam per label
  const T* __restrict__ tabM;        // mu per label
  static constexpr bool kMoments = false;
};

template <typename T, bool SRI>
struct FineValueCells {
  const T* __restrict__ cl;          // lam per cell (x-fastest), 0 where no stiffness
  const T* __restrict__ cm;          // mu per cell
  static constexpr bool kMoments = false;
};

template <typename T>
struct MomentCells {
  const T* __restrict__ mom;         // [54][ncell]: 27 lam moments then 27 mu moments
  static constexpr bool kMoments = true;
};

// Epilogue modes of the structured apply
enum ApplyMode : int {
  MODE_APPLY = 0,   // y = A u
  MODE_RESID = 1,   // y = b - A u
};

template <typename T>
struct GridArgs {
  int nx, ny, nz;                    // cells
  const T* __restrict__ hx;          // cell sizes per axis
  const T* __restrict__ hy;
  const T* __restrict__ hz;
};

// Load the material of cell (i,j,k) into per-cell weights.  Returns false when the cell has no
// stiffness (outside, inactive, or irregular for the structured pass).
template <typename T, bool SRI>
__device__ __forceinline__ bool load_cell(const FineLabelCells<T, SRI>& c, const T* sL, const T* sM, long long cid,
                                          T& lam, T& mu) {
  const uint8_t l = c.lab[cid];
  if (l == LAB_NONE) return false;
  lam = sL[l];
  mu = sM[l];
  return true;
}

template <typename T, bool SRI>
__device__ __forceinline__ bool load_cell(const FineValueCells<T, SRI>& c, const T*, const T*, long long cid, T& lam,
                                          T& mu) {
  lam = c.cl[cid];
  mu = c.cm[cid];
  return (lam != T(0)) || (mu != T(0));
}

template <class C>
struct cell_traits;
template <typename T, bool SRI>
struct cell_traits<FineLabelCells<T, SRI>> {
  static constexpr bool sri = SRI;
  static constexpr bool labels = true;
};
template <typename T, bool SRI>
struct cell_traits<FineValueCells<T, SRI>> {
  static constexpr bool sri = SRI;
  static constexpr bool labels = false;
};

// -------------------------------------------------------------------------------------------
// Structured apply, constant-material (level 0) cells.
template <typename T, class Cells, int MODE>
__global__ __launch_bounds__(BLOCK) void k_apply_fine(GridArgs<T> g, Cells cells, const T* __restrict__ u,
                                                       T* y, const T* b, int kz) {
  constexpr bool SRI = cell_traits<Cells>::sri;
  __shared__ T s_u[2][TN * TN * 3];
  __shared__ T s_g[BLOCK * 9];
  __shared__ T s_tL[256], s_tM[256];
  const int tid = threadIdx.x;
  const int tx = tid % TILE, ty = tid / TILE;
  const int I0 = blockIdx.x * TOUT, J0 = blockIdx.y * TOUT, K0 = blockIdx.z * kz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long nxn = nx + 1, nyn = ny + 1;
  if constexpr (cell_traits<Cells>::labels) {
    s_tL[tid] = cells.tabL[tid];
    s_tM[tid] = cells.tabM[tid];
  }
  const int ci = I0 - 1 + tx, cj = J0 - 1 + ty;
  const bool cin = (ci >= 0 && ci < nx && cj >= 0 && cj < ny);
  T gx = T(0), gy = T(0), axy = T(0);
  if (cin) {
    const T hx = g.hx[ci], hy = g.hy[cj];
    gx = T(0.25) / hx;
    gy = T(0.25) / hy;
    axy = hx * hy;
  }
  auto stage = [&](int kk) {
    T* dst = s_u[kk & 1];
    const bool kin = (kk >= 0 && kk <= nz);
    for (int e = tid; e < TN * TN * 3; e += BLOCK) {
      const int r = e / (TN * 3);
      const int q = e - r * (TN * 3);
      const int gj = J0 - 1 + r;
      const int gi = I0 - 1 + q / 3;
      T v = T(0);
      if (kin && gj >= 0 && gj <= ny && gi >= 0 && gi <= nx)
        v = u[((kk * nyn + gj) * nxn + (I0 - 1)) * 3 + q];
      dst[e] = v;
    }
  };
  T carry[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) carry[q] = T(0);
  const int kend = min(K0 + kz, nz + 1);   // exclusive end of output node layers
  stage(K0 - 1);
  for (int k = K0 - 1; k < kend; ++k) {
    stage(k + 1);
    __syncthreads();
    T f[24];
    bool act = false;
    if (cin && k >= 0 && k < nz) {
      T lam, mu;
      act = load_cell(cells, s_tL, s_tM, (long long)ci + nx * ((long long)cj + (long long)ny * k), lam, mu);
      if (act) {
        T ue[24];
        const T* lo = s_u[k & 1];
        const T* hi = s_u[(k + 1) & 1];
#pragma unroll
        for (int l = 0; l < 8; ++l) {
          const int a = l & 1, bb = (l >> 1) & 1, c = l >> 2;
          const T* src = (c ? hi : lo) + ((ty + bb) * TN + tx + a) * 3;
          ue[3 * l] = src[0]; ue[3 * l + 1] = src[1]; ue[3 * l + 2] = src[2];
        }
        const T hz = g.hz[k];
        const T V = axy * hz;
        CellW<T> w{gx, gy, T(0.25) / hz, lam * V, mu * V};
        elem_apply_const<T, SRI>(ue, f, w);
      }
    }
    if (!act) {
#pragma unroll
      for (int q = 0; q < 24; ++q) f[q] = T(0);
    }
    T g0[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g0[c] = f[c] + carry[c];
#pragma unroll
    for (int q = 0; q < 9; ++q) s_g[tid * 9 + q] = f[3 + q] + carry[3 + q];
#pragma unroll
    for (int q = 0; q < 12; ++q) carry[q] = f[12 + q];
    __syncthreads();
    if (tx >= 1 && ty >= 1 && k >= K0 && ci <= nx && cj <= ny) {
      const T* p1 = s_g + (tid - 1) * 9;        // cell (tx-1, ty): corner l=1
      const T* p2 = s_g + (tid - TILE) * 9;     // cell (tx, ty-1): corner l=2
      const T* p3 = s_g + (tid - TILE - 1) * 9; // cell (tx-1, ty-1): corner l=3
      const long long n = ((long long)k * nyn + cj) * nxn + ci;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T val = g0[c] + p1[c] + p2[3 + c] + p3[6 + c];
        if (MODE == MODE_APPLY) y[3 * n + c] = val;
        else y[3 * n + c] = b[3 * n + c] - val;
      }
    }
  }
}

// Structured apply, coarse levels (moment cells).
template <typename T, int MODE>
__global__ __launch_bounds__(BLOCK) void k_apply_mom(GridArgs<T> g, MomentCells<T> cells, const T* __restrict__ u,
                                                      T* y, const T* b, int kz) {
  __shared__ T s_u[2][TN * TN * 3];
  __shared__ T s_g[BLOCK * 9];
  const int tid = threadIdx.x;
  const int tx = tid % TILE, ty = tid / TILE;
  const int I0 = blockIdx.x * TOUT, J0 = blockIdx.y * TOUT, K0 = blockIdx.z * kz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long nxn = nx + 1, nyn = ny + 1;
  const long long ncell = (long long)nx * ny * nz;
  const int ci = I0 - 1 + tx, cj = J0 - 1 + ty;
  const bool cin = (ci >= 0 && ci < nx && cj >= 0 && cj < ny);
  T gx = T(0), gy = T(0);
  if (cin) { gx = T(0.25) / g.hx[ci]; gy = T(0.25) / g.hy[cj]; }
  auto stage = [&](int kk) {
    T* dst = s_u[kk & 1];
    const bool kin = (kk >= 0 && kk <= nz);
    for (int e = tid; e < TN * TN * 3; e += BLOCK) {
      const int r = e / (TN * 3);
      const int q = e - r * (TN * 3);
      const int gj = J0 - 1 + r;
      const int gi = I0 - 1 + q / 3;
      T v = T(0);
      if (kin && gj >= 0 && gj <= ny && gi >= 0 && gi <= nx)
        v = u[((kk * nyn + gj) * nxn + (I0 - 1)) * 3 + q];
      dst[e] = v;
    }
  };
  T carry[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) carry[q] = T(0);
  const int kend = min(K0 + kz, nz + 1);
  stage(K0 - 1);
  for (int k = K0 - 1; k < kend; ++k) {
    stage(k + 1);
    __syncthreads();
    T f[24];
#pragma unroll
    for (int q = 0; q < 24; ++q) f[q] = T(0);
    if (cin && k >= 0 && k < nz) {
      const long long cid = (long long)ci + nx * ((long long)cj + (long long)ny * k);
      T mL[27], mM[27];
      bool nz_ = false;
#pragma unroll
      for (int q = 0; q < 27; ++q) {
        mL[q] = cells.mom[q * ncell + cid];
        mM[q] = cells.mom[(27 + q) * ncell + cid];
      }
      nz_ = (mL[0] != T(0)) || (mM[0] != T(0));
      if (nz_) {
        T ue[24];
        const T* lo = s_u[k & 1];
        const T* hi = s_u[(k + 1) & 1];
#pragma unroll
        for (int l = 0; l < 8; ++l) {
          const int a = l & 1, bb = (l >> 1) & 1, c = l >> 2;
          const T* src = (c ? hi : lo) + ((ty + bb) * TN + tx + a) * 3;
          ue[3 * l] = src[0]; ue[3 * l + 1] = src[1]; ue[3 * l + 2] = src[2];
        }
        elem_apply_mom<T>(ue, f, mL, mM, gx, gy, T(0.25) / g.hz[k]);
      }
    }
    T g0[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g0[c] = f[c] + carry[c];
#pragma unroll
    for (int q = 0; q < 9; ++q) s_g[tid * 9 + q] = f[3 + q] + carry[3 + q];
#pragma unroll
    for (int q = 0; q < 12; ++q) carry[q] = f[12 + q];
    __syncthreads();
    if (tx >= 1 && ty >= 1 && k >= K0 && ci <= nx && cj <= ny) {
      const T* p1 = s_g + (tid - 1) * 9;
      const T* p2 = s_g + (tid - TILE) * 9;
      const T* p3 = s_g + (tid - TILE - 1) * 9;
      const long long n = ((long long)k * nyn + cj) * nxn + ci;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T val = g0[c] + p1[c] + p2[3 + c] + p3[6 + c];
        if (MODE == MODE_APPLY) y[3 * n + c] = val;
        else y[3 * n + c] = b[3 * n + c] - val;
      }
    }
  }
}

// -------------------------------------------------------------------------------------------
// Irregular (split-copy) cells: explicit 8 node ids, scatter with atomics.  sign = +1 adds A u,
// -1 subtracts (residual mode).
template <typename T, bool SRI>
__global__ void k_apply_irregular(int nirr, const int* __restrict__ nodes, const int* __restrict__ ijk,
                                  const T* __restrict__ lam, const T* __restrict__ mu, GridArgs<T> g,
                                  const T* __restrict__ u, T* __restrict__ y, T sign) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nirr) return;
  const int i = ijk[3 * e], j = ijk[3 * e + 1], k = ijk[3 * e + 2];
  const T hx = g.hx[i], hy = g.hy[j], hz = g.hz[k];
  const T V = hx * hy * hz;
  CellW<T> w{T(0.25) / hx, T(0.25) / hy, T(0.25) / hz, lam[e] * V, mu[e] * V};
  T ue[24], f[24];
  int nd[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    nd[l] = nodes[8 * e + l];
    ue[3 * l] = u[3LL * nd[l]];
    ue[3 * l + 1] = u[3LL * nd[l] + 1];
    ue[3 * l + 2] = u[3LL * nd[l] + 2];
  }
  elem_apply_const<T, SRI>(ue, f, w);
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    atomicAdd(&y[3LL * nd[l]], sign * f[3 * l]);
    atomicAdd(&y[3LL * nd[l] + 1], sign * f[3 * l + 1]);
    atomicAdd(&y[3LL * nd[l] + 2], sign * f[3 * l + 2]);
  }
}

// Sparse extra stiffness (springs, penalties, ligaments; Galerkin-coarsened on coarse levels):
// y[row] += sign * sum_k val[k] u[col[k]] over the listed rows (rows unique).  W lanes share a row
// (strided over its entries, shuffle-reduced), W chosen per level from the mean row length.
template <typename T, int W>
__global__ void k_apply_extra(int nrow, const int* __restrict__ rows, const int* __restrict__ ptr,
                              const int* __restrict__ cols, const T* __restrict__ vals, const T* __restrict__ u,
                              T* y, T sign) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(gid / W), lane = (int)(gid % W);
  if (r >= nrow) return;   // the W lanes of a row leave together
  T s = T(0);
  for (int q = ptr[r] + lane; q < ptr[r + 1]; q += W) s += vals[q] * u[cols[q]];
  if (W > 1) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) s += __shfl_down(s, off, W);
  }
  if (lane == 0) y[rows[r]] += sign * s;
}

// Rank-one extra stiffness sum_t k_t w_t w_t^T (springs, penalties, ligaments and their exact
// Galerkin coarsening): y += sign * k_t (w_t . u) w_t.  One wavefront per term: lanes stride over
// the term's entries, the dot product is shuffle-reduced, each lane scatters its entries (atomics).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_rank1(int nterm, const int* __restrict__ ptr,
                                                     const int* __restrict__ dof, const T* __restrict__ w,
                                                     const T* __restrict__ k, const T* __restrict__ u, T* y,
                                                     T sign) {
  const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= nterm) return;     // the 64 lanes of a term leave together
  const int q0 = ptr[t], q1 = ptr[t + 1];
  T s = T(0);
  for (int q = q0 + lane; q < q1; q += 64) s += w[q] * u[dof[q]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  s *= sign * k[t];
  for (int q = q0 + lane; q < q1; q += 64) atomicAdd(&y[dof[q]], s * w[q]);
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, long long n, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Residual fix for the copies region when the structured kernel computed y = b - A_s u:
// copies have no structured part, so y = b there before the irregular/extra passes subtract.
template <typename T>
__global__ void k_copy_range(T* __restrict__ dst, const T* __restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// -------------------------------------------------------------------------------------------
// Grid transfers between level l (fine, nodes nf per axis) and l+1 (coarse).
// Prolongation per axis: fine node i -> coarse index ia[i] and weight w[i] of ia[i]+1
// (value = (1-w) e[ia] + w e[ia+1]; coarse-coincident nodes have w = 0), fe/gmg.py:7-26.
template <typename T>
struct AxisP {
  const int* __restrict__ ia;
  const T* __restrict__ w;
};
// Restriction per axis: coarse node I gathers fine nodes rs[I] .. rs[I]+2 with weights rw[3I..3I+2].
template <typename T>
struct AxisR {
  const int* __restrict__ rs;
  const T* __restrict__ rw;
};

// rc = P^T r over the grid nodes (structured part).
template <typename T>
__global__ void k_restrict(int nfx, int nfy, int nfz, int ncx, int ncy, int ncz, AxisR<T> rx, AxisR<T> ry,
                           AxisR<T> rz, const T* __restrict__ r, T* __restrict__ rc) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)ncx * ncy * ncz;
  if (id >= nn) return;
  const int I = id % ncx, J = (id / ncx) % ncy, K = id / ((long long)ncx * ncy);
  T s0 = T(0), s1 = T(0), s2 = T(0);
  const int i0 = rx.rs[I], j0 = ry.rs[J], k0 = rz.rs[K];
  for (int c = 0; c < 3; ++c) {
    const T wz = rz.rw[3 * K + c];
    const int kf = k0 + c;
    if (wz == T(0) || kf >= nfz) continue;
    for (int bq = 0; bq < 3; ++bq) {
      const T wy = ry.rw[3 * J + bq];
      const int jf = j0 + bq;
      if (wy == T(0) || jf >= nfy) continue;
      const T wyz = wy * wz;
      const long long rowbase = ((long long)kf * nfy + jf) * nfx;
      for (int a = 0; a < 3; ++a) {
        const T wx = rx.rw[3 * I + a];
        const int if_ = i0 + a;
        if (wx == T(0) || if_ >= nfx) continue;
        const T ww = wx * wyz;
        const T* p = r + 3 * (rowbase + if_);
        s0 += ww * p[0]; s1 += ww * p[1]; s2 += ww * p[2];
      }
    }
  }
  rc[3 * id] = s0; rc[3 * id + 1] = s1; rc[3 * id + 2] = s2;
}

// Interpolated coarse correction at fine grid node (i,j,k).
template <typename T>
__device__ __forceinline__ void interp_node(int i, int j, int k, int ncx, int ncy, AxisP<T> px, AxisP<T> py,
                                            AxisP<T> pz, const T* __restrict__ ec, T* out) {
  const int a = px.ia[i], b = py.ia[j], c = pz.ia[k];
  const T wx = px.w[i], wy = py.w[j], wz = pz.w[k];
  out[0] = out[1] = out[2] = T(0);
#pragma unroll
  for (int dz = 0; dz < 2; ++dz) {
    const T fz = dz ? wz : T(1) - wz;
    if (fz == T(0)) continue;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const T fy = dy ? wy : T(1) - wy;
      if (fy == T(0)) continue;
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const T fx = dx ? wx : T(1) - wx;
        if (fx == T(0)) continue;
        const T ww = fx * fy * fz;
        const long long n = ((long long)(c + dz) * ncy + (b + dy)) * ncx + (a + dx);
        out[0] += ww * ec[3 * n]; out[1] += ww * ec[3 * n + 1]; out[2] += ww * ec[3 * n + 2];
      }
    }
  }
}

// x += P ec on used fine grid nodes; unused nodes are set to zero (they are not unknowns).
template <typename T>
__global__ void k_prolong_add(int nfx, int nfy, int nfz, int ncx, int ncy, AxisP<T> px, AxisP<T> py, AxisP<T> pz,
                              const T* __restrict__ ec, const uint8_t* __restrict__ used, T* __restrict__ x) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)nfx * nfy * nfz;
  if (id >= nn) return;
  if (!used[id]) {
    x[3 * id] = x[3 * id + 1] = x[3 * id + 2] = T(0);
    return;
  }
  const int i = id % nfx, j = (id / nfx) % nfy, k = id / ((long long)nfx * nfy);
  T e[3];
  interp_node(i, j, k, ncx, ncy, px, py, pz, ec, e);
  x[3 * id] += e[0]; x[3 * id + 1] += e[1]; x[3 * id + 2] += e[2];
}

// Split copies: a copy is interpolated like its original grid node (rb/splitfe.py:153,
// used_nodes = orig[nused]).  Prolongation of copies:
template <typename T>
__global__ void k_prolong_copies(int ncopy, long long nng, const int* __restrict__ orig, int nfx, int nfy, int ncx,
                                 int ncy, AxisP<T> px, AxisP<T> py, AxisP<T> pz, const T* __restrict__ ec,
                                 T* __restrict__ x) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncopy) return;
  const int o = orig[q];
  const int i = o % nfx, j = (o / nfx) % nfy, k = o / (nfx * nfy);
  T e[3];
  interp_node(i, j, k, ncx, ncy, px, py, pz, ec, e);
  const long long n = nng + q;
  x[3 * n] += e[0]; x[3 * n + 1] += e[1]; x[3 * n + 2] += e[2];
}

// Restriction of copies: rc += P[orig]^T r[copy] (atomics; few copies).
template <typename T>
__global__ void k_restrict_copies(int ncopy, long long nng, const int* __restrict__ orig, int nfx, int nfy, int ncx,
                                  int ncy, AxisP<T> px, AxisP<T> py, AxisP<T> pz, const T* __restrict__ r,
                                  T* __restrict__ rc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncopy) return;
  const int o = orig[q];
  const int i = o % nfx, j = (o / nfx) % nfy, k = o / (nfx * nfy);
  const long long n = nng + q;
  const T v0 = r[3 * n], v1 = r[3 * n + 1], v2 = r[3 * n + 2];
  const int a = px.ia[i], b = py.ia[j], c = pz.ia[k];
  const T wx = px.w[i], wy = py.w[j], wz = pz.w[k];
  for (int dz = 0; dz < 2; ++dz) {
    const T fz = dz ? wz : T(1) - wz;
    if (fz == T(0)) continue;
    for (int dy = 0; dy < 2; ++dy) {
      const T fy = dy ? wy : T(1) - wy;
      if (fy == T(0)) continue;
      for (int dx = 0; dx < 2; ++dx) {
        const T fx = dx ? wx : T(1) - wx;
        if (fx == T(0)) continue;
        const T ww = fx * fy * fz;
        const long long m = ((long long)(c + dz) * ncy + (b + dy)) * ncx + (a + dx);
        atomicAdd(&rc[3 * m], ww * v0);
        atomicAdd(&rc[3 * m + 1], ww * v1);
        atomicAdd(&rc[3 * m + 2], ww * v2);
      }
    }
  }
}

}  // namespace cm
