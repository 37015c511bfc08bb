// Matrix-free operator kernels for the voxel hex8 system on a rectilinear grid.
//
// Layout in HBM (all levels):
//   node vectors  AoS, 3 components per node, node id i + (nx+1)(j + (ny+1)k) -- the reference's
//                 grid numbering (fe/solver.py:107); level 0 appends the split-node copies after
//                 the grid nodes (ids nnode + rank, rb/splitfe.py:268-280)
//   cells         x-fastest internal order i + nx (j + ny k); the boundary converts from the
//                 reference's z-fastest cell ids (fe/solver.py:104-105,128)
//   level 0       one uint8 label per cell (255 = no stiffness) + a (lam, mu) table, or two
//                 per-cell values; irregular cells (touching split copies) go to the sparse pass
//   level >= 1    27 lam-moments + 27 mu-moments per coarse cell (structure of arrays)
//
// One operator application = sparse pass (irregular cells, spring rows, rank-one terms into a
// compact vector over the few nodes they touch) + structured march over the grid (every cell
// computed once, epilogue fused: plain product, residual, or a whole smoothing step) + fix-up of
// the touched nodes.
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
// Cell material policies of the structured march
//   kind 0: level 0, one uint8 label per cell + (lam, mu) table
//   kind 1: level 0, per-cell (lam, mu) values
//   kind 2: coarse levels, 27 lam + 27 mu moments per cell
//   kind 4: level 1 split by cells: cells whose children share one material take the constant-
//           material operator, the others go through the heterogeneous-cell pass (het_body)
// ISO = the level-0 grid is made of identical cubes (h_x = h_y = h_z = h everywhere); the common
// factor V g^2 = h / 16 is then folded into the material weights (GridArgs::iso).
template <typename T, bool SRI_, bool ISO_>
struct LabelMat {
  const uint8_t* __restrict__ lab;   // x-fastest, LAB_NONE = no stiffness (or irregular)
  const T* __restrict__ tabL;        // lam per label
  const T* __restrict__ tabM;        // mu per label
  static constexpr int kind = 0;
  static constexpr bool sri = SRI_, iso = ISO_;
};

template <typename T, bool SRI_, bool ISO_>
struct ValueMat {
  const T* __restrict__ cl;          // lam per cell (x-fastest), 0 where no stiffness
  const T* __restrict__ cm;          // mu per cell
  static constexpr int kind = 1;
  static constexpr bool sri = SRI_, iso = ISO_;
};

// Coarse levels also carry extra stiffness (the Galerkin-coarsened springs as 3x3 block rows,
// rank-one terms k w w^T with a wide support): k_coarse_extra, launched just before the march,
// accumulates its product with s into ex (3 per node), and the march epilogue of every node adds
// ex and sets it back to zero (ex is all zero between applications).
template <typename T>
struct MomentMat {
  const T* __restrict__ mom;         // [54][ncell]: 27 lam moments then 27 mu moments
  T* ex;                             // nullptr: no extra stiffness
  static constexpr int kind = 2;
  static constexpr bool sri = false, iso = false;
};

// Level-1 cells are unions of (up to) 2x2x2 level-0 cells of constant material.  Along every
// axis a child occupies xi in its parent's coordinates as xi = a xi_c + b; the parent's moments
// int xi^g m dV are then sums over the children of w_c prod_axes phi(g), phi = (1, b, a^2/3 + b^2)
// (exact integration) or (1, b, b^2) (lam at the child's centre, SRI), w_c = m_c V_c: the child
// weights cw ([16][ncell]: lam weights of the children dx + 2dy + 4dz, then mu) and the per-axis
// factor rows ph (per coarse index: b0, q0, s0, b1, q1, s1 with q = a^2/3 + b^2, s = b^2).
// A level-1 cell whose existing level-0 children all carry one material and split every axis into
// equal halves (or not at all) has the Galerkin operator of a constant material whose lam term is
// integrated at the children's SRI points: elem_mid_q with the per-axis second moments q (1/4 for
// two halves, 0 for one child; 1/3 without SRI).  cl, cm hold lam V and mu V of such cells and 0 for
// every other cell.  The cells with several materials are applied by het_cell (extra blocks of the
// coarse-springs launch before the march) into hf, 24 forces per cell; the march epilogue of each of
// their nodes adds its corners' forces (hmap: the node's row of the slot lists, -1 for none; slots
// 8 * cell + corner, in increasing order) -- a fixed summation order and no launch of its own.
template <typename T, bool SRI_>
struct SplitMat {
  const T* __restrict__ cl;          // lam V per cell (x-fastest), 0 where the cell is not uniform
  const T* __restrict__ cm;          // mu V per cell
  const T* __restrict__ qx;          // lam second moment per cell column, row and layer
  const T* __restrict__ qy;
  const T* __restrict__ qz;
  T* ex;
  const int* __restrict__ hmap;      // per node: row of hptr, or -1
  const int* __restrict__ hptr;
  const int* __restrict__ hslot;
  const T* __restrict__ hf;
  static constexpr int kind = 4;
  static constexpr bool sri = SRI_, iso = false;
};

// Per-axis factor tables of a level-1 cell's children from its ph rows (b, a^2/3 + b^2, b^2 per
// child): f?L for the lam moments (SRI: centre points), f?M for the mu moments.
template <typename T, bool SRI>
__device__ __forceinline__ void child_factors(const T* __restrict__ p, T fL[2][2], T fM[2][2]) {
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    fM[d][0] = fL[d][0] = p[3 * d];
    fM[d][1] = p[3 * d + 1];
    fL[d][1] = SRI ? p[3 * d + 2] : p[3 * d + 1];
  }
}

// Expands child weights w[8] into the 27 moments (index gx + 3 gy + 9 gz) with per-axis factor
// tables f?[d][g] (g = 1, 2; g = 0 is 1).
template <typename T>
__device__ __forceinline__ void child_moments(const T w[8], const T fx[2][2], const T fy[2][2], const T fz[2][2],
                                              T m[27]) {
  T t[3][4];                  // [gx][dy + 2 dz]
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const T w0 = w[2 * p], w1 = w[2 * p + 1];
    t[0][p] = w0 + w1;
    t[1][p] = fx[0][0] * w0 + fx[1][0] * w1;
    t[2][p] = fx[0][1] * w0 + fx[1][1] * w1;
  }
  T u[3][3][2];               // [gx][gy][dz]
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int dz = 0; dz < 2; ++dz) {
      const T v0 = t[a][2 * dz], v1 = t[a][2 * dz + 1];
      u[a][0][dz] = v0 + v1;
      u[a][1][dz] = fy[0][0] * v0 + fy[1][0] * v1;
      u[a][2][dz] = fy[0][1] * v0 + fy[1][1] * v1;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const T v0 = u[a][b][0], v1 = u[a][b][1];
      m[a + 3 * b] = v0 + v1;
      m[a + 3 * b + 9] = fz[0][0] * v0 + fz[1][0] * v1;
      m[a + 3 * b + 18] = fz[0][1] * v0 + fz[1][1] * v1;
    }
}

template <typename T>
struct GridArgs {
  int nx, ny, nz;                    // cells
  const T* __restrict__ hx;          // cell sizes per axis
  const T* __restrict__ hy;
  const T* __restrict__ hz;
  T iso;                             // h / 16 when every cell is the same cube, else 0
};

// Block order of the structured march.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8); tile ids are handed out so that every XCD gets one contiguous run
// of tiles (x fastest, then y, then z), which keeps the tiles sharing halo nodes in one L2.
struct TileMap {
  int ntx, nty, ntz;
  int q, rem;                        // nblocks / 8, nblocks % 8
  __device__ __forceinline__ void decode(int b, int& bx, int& by, int& bz) const {
    const int xcd = b & 7, pos = b >> 3;
    const int id = xcd * q + min(xcd, rem) + pos;
    bx = id % ntx;
    const int r = id / ntx;
    by = r % nty;
    bz = r / nty;
  }
};

// -------------------------------------------------------------------------------------------
// Epilogues.  au = (A s) at one node from the structured part, sn = the node's value of s.
enum EpiMode : int {
  EPI_APPLY = 0,   // y = A s
  EPI_RESID = 1,   // y = rin - A s
  EPI_CHEB = 2,    // one smoothing step: r = rin - A s;  d = c1 sn + c2 Dinv r;  x += xa sn + d
};
// Names used by the entry points that speak of the plain operator.
constexpr int MODE_APPLY = EPI_APPLY;
constexpr int MODE_RESID = EPI_RESID;

template <typename T>
struct Epi {
  const T* rin;       // RESID / CHEB: right-hand side or running residual
  T* y;               // APPLY / RESID: output;  CHEB: updated residual r (nullptr: not kept)
  const T* Dinv;      // CHEB: symmetric 3x3 block inverses, 6 per node (00 01 02 11 12 22)
  T* dout;            // CHEB: new direction d (nullptr: not kept)
  T* x;               // CHEB: iterate, updated in place (nullptr: not updated)
  T c1, c2, xa;
  double* part;       // APPLY: per-block partial sums of s . (A s) (nullptr: none)
  const int* fold;    // RESID: original grid node of every split copy; the copy's residual is
                      // added to its original's (a copy restricts like its original), nullptr: none
};

template <typename T>
__host__ __device__ __forceinline__ void sym3_mul(const T* __restrict__ m, const T r[3], T z[3]) {
  const T a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
  z[0] = a * r[0] + b * r[1] + c * r[2];
  z[1] = b * r[0] + d * r[1] + e * r[2];
  z[2] = c * r[0] + e * r[1] + f * r[2];
}

// Epilogue inputs of one node (the residual or right-hand side, the block inverse, the iterate).
// Level 0 loads them one z step ahead, so that their latency overlaps the next step's work.
template <typename T>
struct EpiIn {
  T rin[3], dinv[6], x[3];
};

template <typename T, int EM>
__device__ __forceinline__ void epi_load(const Epi<T>& ep, long long n, EpiIn<T>& q) {
  if (EM == EPI_RESID || EM == EPI_CHEB) {
#pragma unroll
    for (int c = 0; c < 3; ++c) q.rin[c] = ep.rin[3 * n + c];
  }
  if (EM == EPI_CHEB) {
#pragma unroll
    for (int m = 0; m < 6; ++m) q.dinv[m] = ep.Dinv[6 * n + m];
    if (ep.x) {
#pragma unroll
      for (int c = 0; c < 3; ++c) q.x[c] = ep.x[3 * n + c];
    }
  }
}

template <typename T, int EM>
__device__ __forceinline__ void epilogue_in(const Epi<T>& ep, long long n, const T au[3], const T sn[3],
                                            const EpiIn<T>& q) {
  if (EM == EPI_APPLY) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = au[c];
  } else if (EM == EPI_RESID) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = q.rin[c] - au[c];
  } else {
    T r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = q.rin[c] - au[c];
    if (ep.y) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = r[c];
    }
    sym3_mul(q.dinv, r, z);
    T dn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dn[c] = ep.c1 * sn[c] + ep.c2 * z[c];
    if (ep.dout) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.dout[3 * n + c] = dn[c];
    }
    if (ep.x) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.x[3 * n + c] = q.x[c] + (ep.xa * sn[c] + dn[c]);
    }
  }
}

// The same epilogue with its inputs loaded where they are used.
template <typename T, int EM>
__device__ __forceinline__ void epilogue(const Epi<T>& ep, long long n, const T au[3], const T sn[3]) {
  if (EM == EPI_APPLY) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = au[c];
  } else if (EM == EPI_RESID) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = ep.rin[3 * n + c] - au[c];
  } else {
    T r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = ep.rin[3 * n + c] - au[c];
    if (ep.y) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.y[3 * n + c] = r[c];
    }
    sym3_mul(ep.Dinv + 6 * n, r, z);
    T dn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dn[c] = ep.c1 * sn[c] + ep.c2 * z[c];
    if (ep.dout) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.dout[3 * n + c] = dn[c];
    }
    if (ep.x) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.x[3 * n + c] += ep.xa * sn[c] + dn[c];
    }
  }
}

// -------------------------------------------------------------------------------------------
// Structured march (all levels).  A block of 16x16 threads owns the 15x15 output node columns of
// one xy tile and marches a z-range of kz node layers.  Thread (tx, ty) owns cell column
// (I0-1+tx, J0-1+ty).  Per z step:
//   * the next node layer of s (17x17 nodes, padded to 4 values) is written to LDS from registers
//     loaded one step earlier (the global loads of layer k+2 are in flight during step k);
//   * the thread transforms its cell's top face in xy (the bottom face's transform is the previous
//     step's top face, kept in registers), finishes the z transform, applies the cell operator in
//     transform space and splits the forces into bottom/top faces (still in xy transform space);
//   * the top-face forces are carried to the next step, the bottom-face forces plus the carry
//     are transformed back to the 4 bottom corners; three corners go to LDS for the in-plane
//     neighbours, the fourth is the thread's own output node.
// Three values of a 4-value LDS slot (16-byte aligned), read and written as whole slots: a
// 64-lane access of consecutive slots is then served without bank conflicts (a 3-value access is
// split into dwords 4 words apart, which meet on a quarter of the banks).
__device__ __forceinline__ void lds_get3(const float* p, float v[3]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  v[0] = q.x; v[1] = q.y; v[2] = q.z;
}
__device__ __forceinline__ void lds_get3(const double* p, double v[3]) {
  const double2 a = *reinterpret_cast<const double2*>(p), b = *reinterpret_cast<const double2*>(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x;
}
__device__ __forceinline__ void lds_put3(float* p, const float v[3]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], 0.0f);
}
__device__ __forceinline__ void lds_put3(double* p, const double v[3]) {
  *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], 0.0);
}

// Value of lane l - 1 of the same 16-lane DPP row (row_shr:1), 0 in a row's first lane.  The
// march lays out one tile row of 16 threads per DPP row, so this is the value of the thread at
// tx - 1 (lanes at tx = 0 are halo lanes).
__device__ __forceinline__ float row_prev(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ double row_prev(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), 0x111, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x111, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// LDS layout of n 3-vectors.  FP32: 16-byte slots as above.  FP64: one plane of n values per
// component; 8-byte accesses of consecutive lanes are conflict-free, and the planes need 3/4 of the
// LDS of 32-byte slots, which lets a fourth FP64 march block share a CU.
template <typename T>
struct V3Lds {
  static constexpr bool slots = sizeof(T) == 4;
  static constexpr int size(int n) { return slots ? 4 * n : 3 * n; }
  static __device__ __forceinline__ int idx(int n, int i, int c) { return slots ? 4 * i + c : c * n + i; }
  static __device__ __forceinline__ void get(const T* base, int n, int i, T v[3]) {
    if constexpr (slots) {
      lds_get3(base + 4 * i, v);
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = base[c * n + i];
    }
  }
  static __device__ __forceinline__ void put(T* base, int n, int i, const T v[3]) {
    if constexpr (slots) {
      lds_put3(base + 4 * i, v);
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) base[c * n + i] = v[c];
    }
  }
};

// Extra stiffness of a coarse node (kinds 2, 4): au += ex[n], ex[n] = 0; kind 4 also adds the
// forces of the node's corners of several-material cells.
template <typename T, class M>
__device__ __forceinline__ void add_extra(const M& mat, long long n, T au[3]) {
  if constexpr (M::kind == 4) {
    const int h = mat.hmap[n];
    if (h >= 0) {
      for (int q = mat.hptr[h]; q < mat.hptr[h + 1]; ++q) {
        const long long o = 3LL * mat.hslot[q];
#pragma unroll
        for (int c = 0; c < 3; ++c) au[c] += mat.hf[o + c];
      }
    }
  }
  if constexpr (M::kind >= 2) {
    if (mat.ex) {
      T e3[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) e3[c] = mat.ex[3 * n + c];
      if (e3[0] != T(0) || e3[1] != T(0) || e3[2] != T(0)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          au[c] += e3[c];
          mat.ex[3 * n + c] = T(0);
        }
      }
    }
  }
}

// Sparse pass: everything the structured march leaves out, accumulated into a compact vector tc
// over the level's "touched" nodes (3 values per listed node):
//   * irregular cells (level 0: cells touching split copies), 8 explicit node ids;
//   * extra stiffness rows (springs, penalties, ligaments, their Galerkin products), CSR;
//   * rank-one terms k w w^T with a support wider than one node neighbourhood (coarse levels).
// One launch, three block ranges.  Atomics: rows and cells overlap on the touched nodes.
// Extra stiffness of a coarse level, accumulated into ex (all zero on entry; the march that
// follows adds it to its nodes and clears it).  Two block ranges:
//   * rank-one terms k_t w_t w_t^T, one wavefront per term: d = k_t (w_t . s) reduced over the
//     wavefront, then ex[dof] += d w for every entry of the term;
//   * block entries (row node brow_e, column node bcol_e, 3x3 block B_e row-major), one thread each:
//     ex[brow_e] += B_e s[bcol_e].
template <typename T>
struct CoarseExtra {
  int nterm, nblk;
  int nb_term;                       // blocks of the first range
  int rw;                            // lanes per rank-one term (power of two <= 64)
  const int* __restrict__ r1ptr;
  const int* __restrict__ r1dof;
  const T* __restrict__ r1w;
  const T* __restrict__ r1k;
  const int* __restrict__ brow;
  const int* __restrict__ bcol;
  const T* __restrict__ bval;
};
// The pass runs as extra blocks of the level-0 march launch (it does not depend on the march, the
// fix-up that follows reads both), so it costs no launch of its own.
template <typename T>
struct SparseArgs {
  int nb;                            // blocks of the pass (0: none)
  int nirr, nrow, nterm;
  int nb_irr, nb_row;                // blocks of the first two ranges
  int xw;                            // lanes per CSR row (power of two <= 64)
  int rw;                            // lanes per rank-one term (power of two <= 64)
  const int* __restrict__ irr_nodes; // [nirr][8] node ids
  const int* __restrict__ irr_ci;    // [nirr][8] compact node index
  const int* __restrict__ irr_ijk;   // [nirr][3]
  const T* __restrict__ irrL;
  const T* __restrict__ irrM;
  const int* __restrict__ xrc;       // compact DOF of each CSR row
  const int* __restrict__ xptr;
  const int* __restrict__ xcol;      // full DOF
  const T* __restrict__ xval;
  const int* __restrict__ r1ptr;
  const int* __restrict__ r1dof;     // full DOF
  const int* __restrict__ r1c;       // compact DOF
  const T* __restrict__ r1w;
  const T* __restrict__ r1k;
  T* __restrict__ tc;                // accumulator over the touched nodes
};

template <typename T, bool SRI>
__device__ __forceinline__ void sparse_body(const SparseArgs<T>& a, const GridArgs<T>& g, const T* __restrict__ s,
                                            int b) {
  T* __restrict__ tc = a.tc;
  if (b < a.nb_irr) {
    const int e = b * 256 + threadIdx.x;
    if (e >= a.nirr) return;
    const int i = a.irr_ijk[3 * e], j = a.irr_ijk[3 * e + 1], k = a.irr_ijk[3 * e + 2];
    const T hx = g.hx[i], hy = g.hy[j], hz = g.hz[k];
    const T V = hx * hy * hz;
    CellW<T> w{T(0.25) / hx, T(0.25) / hy, T(0.25) / hz, a.irrL[e] * V, a.irrM[e] * V};
    T ue[24], f[24];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const long long nd = a.irr_nodes[8 * e + l];
      ue[3 * l] = s[3 * nd];
      ue[3 * l + 1] = s[3 * nd + 1];
      ue[3 * l + 2] = s[3 * nd + 2];
    }
    elem_apply_const<T, SRI>(ue, f, w);
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const int q = 3 * a.irr_ci[8 * e + l];
      atomicAdd(&tc[q], f[3 * l]);
      atomicAdd(&tc[q + 1], f[3 * l + 1]);
      atomicAdd(&tc[q + 2], f[3 * l + 2]);
    }
    return;
  }
  b -= a.nb_irr;
  if (b < a.nb_row) {
    const int W = a.xw;
    const long long gid = (long long)b * 256 + threadIdx.x;
    const int r = (int)(gid / W), lane = (int)(gid % W);
    if (r >= a.nrow) return;   // the W lanes of a row leave together
    T acc = T(0);
    for (int q = a.xptr[r] + lane; q < a.xptr[r + 1]; q += W) acc += a.xval[q] * s[a.xcol[q]];
    for (int off = W / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, W);
    if (lane == 0) atomicAdd(&tc[a.xrc[r]], acc);
    return;
  }
  b -= a.nb_row;
  const int rw = a.rw;
  const long long gid = (long long)b * 256 + threadIdx.x;
  const int t = (int)(gid / rw), lane = (int)(gid % rw);
  if (t >= a.nterm) return;   // the rw lanes of a term leave together
  const int q0 = a.r1ptr[t], q1 = a.r1ptr[t + 1];
  T acc = T(0);
  for (int q = q0 + lane; q < q1; q += rw) acc += a.r1w[q] * s[a.r1dof[q]];
  for (int off = rw / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, rw);
  acc *= a.r1k[t];
  for (int q = q0 + lane; q < q1; q += rw) atomicAdd(&tc[a.r1c[q]], acc * a.r1w[q]);
}

// Buffer resources (128-bit descriptors held in SGPRs) over one node or cell layer.  An access at
// a byte offset at or past the descriptor's length reads 0 and writes nothing; lanes without an
// entry use the offset OOB.  The march then runs no 64-bit address arithmetic and no select per
// access: a layer's base moves in scalar registers, every lane keeps its 32-bit offsets.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int RS_FLAGS = 0x00020000;   // raw (untyped) buffer, 32-bit data format
constexpr unsigned OOB = 0x80000000u;

// Descriptor of [base + off, base + off + len) bytes; ok = false gives an empty one.
__device__ __forceinline__ Rsrc layer_rs(const void* base, long long off, long long len, bool ok) {
  const int n = ok ? (int)min(len, 0x7fffffffLL) : 0;
  char* p = const_cast<char*>(static_cast<const char*>(base)) + (ok ? off : 0);
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, n, RS_FLAGS);
}
__device__ __forceinline__ void bld(Rsrc r, unsigned off, float& v) {
  v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bld(Rsrc r, unsigned off, double& v) {
  v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
// 3 consecutive values (one node's xyz).  FP32: one 96-bit access, its components taken from a
// float3 bit-cast of the whole vector (a bit-cast of each 32-bit component compiled into a single
// dword load, the other two components left undefined).  FP64: three 64-bit accesses.
typedef float cm_f3 __attribute__((ext_vector_type(3)));
typedef unsigned int cm_u3 __attribute__((ext_vector_type(3)));
typedef unsigned int cm_u2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void bld3(Rsrc r, unsigned off, float v[3]) {
  const cm_f3 q = __builtin_bit_cast(cm_f3, __builtin_amdgcn_raw_buffer_load_b96(r, (int)off, 0, 0));
  v[0] = q.x;
  v[1] = q.y;
  v[2] = q.z;
}
__device__ __forceinline__ void bld3(Rsrc r, unsigned off, double v[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) bld(r, off + 8u * c, v[c]);
}
__device__ __forceinline__ void bst3(Rsrc r, unsigned off, const float v[3]) {
  cm_f3 q;
  q.x = v[0];
  q.y = v[1];
  q.z = v[2];
  __builtin_amdgcn_raw_buffer_store_b96(__builtin_bit_cast(cm_u3, q), r, (int)off, 0, 0);
}
__device__ __forceinline__ void bst3(Rsrc r, unsigned off, const double v[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(cm_u2, v[c]), r, (int)(off + 8u * c), 0, 0);
}

// Epilogue arrays of one node layer as buffer descriptors.
template <typename T>
struct EpiRs {
  Rsrc rin, y, dinv, dout, x;
};
template <typename T, int EM>
__device__ __forceinline__ EpiRs<T> epi_rs(const Epi<T>& ep, long long k, long long lay, bool ok) {
  const long long b3 = 3 * lay * (long long)sizeof(T);
  EpiRs<T> r;
  r.y = layer_rs(ep.y, k * b3, b3, ok && ep.y != nullptr);
  if (EM == EPI_RESID || EM == EPI_CHEB) r.rin = layer_rs(ep.rin, k * b3, b3, ok);
  if (EM == EPI_CHEB) {
    r.dinv = layer_rs(ep.Dinv, 2 * k * b3, 2 * b3, ok);
    r.dout = layer_rs(ep.dout, k * b3, b3, ok && ep.dout != nullptr);
    r.x = layer_rs(ep.x, k * b3, b3, ok && ep.x != nullptr);
  }
  return r;
}
// Epilogue inputs of one node from buffer descriptors (off3: byte offset of the node's 3-vector).
template <typename T, int EM>
__device__ __forceinline__ void epi_load_rs(const EpiRs<T>& r, bool has_x, unsigned off3, EpiIn<T>& q) {
  if (EM == EPI_RESID || EM == EPI_CHEB) bld3(r.rin, off3, q.rin);
  if (EM == EPI_CHEB) {
    bld3(r.dinv, 2 * off3, q.dinv);
    bld3(r.dinv, 2 * off3 + 3 * sizeof(T), q.dinv + 3);
    bld3(r.x, off3, q.x);   // reads 0 without an iterate (empty descriptor)
  }
}
// Epilogue outputs of one node: y = A s (APPLY), rin - A s (RESID) or r (CHEB); d and x (CHEB).
template <typename T>
struct EpiOut {
  T y[3], d[3], x[3];
};
template <typename T, int EM>
__device__ __forceinline__ void epi_compute(const Epi<T>& ep, const T au[3], const T sn[3], const EpiIn<T>& q,
                                            EpiOut<T>& o) {
  if (EM == EPI_APPLY) {
#pragma unroll
    for (int c = 0; c < 3; ++c) o.y[c] = au[c];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) o.y[c] = q.rin[c] - au[c];
  }
  if (EM == EPI_CHEB) {
    T z[3];
    sym3_mul(q.dinv, o.y, z);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o.d[c] = ep.c1 * sn[c] + ep.c2 * z[c];
      o.x[c] = q.x[c] + (ep.xa * sn[c] + o.d[c]);
    }
  }
}
// Stores of the epilogue outputs; arrays the caller does not keep have empty descriptors.
template <typename T, int EM>
__device__ __forceinline__ void epi_store_rs(const EpiRs<T>& r, unsigned off3, const EpiOut<T>& o) {
  bst3(r.y, off3, o.y);
  if (EM == EPI_CHEB) {
    bst3(r.dout, off3, o.d);
    bst3(r.x, off3, o.x);
  }
}

template <int P>
struct Par {
  static constexpr int value = P;
};

// Prefetch depths of the structured march: node layers of s (MD) and epilogue inputs (ED) are
// loaded this many z steps before their use (1 or 2).
#ifndef CM_MARCH_MD
#define CM_MARCH_MD 1
#endif
#ifndef CM_MARCH_ED
#define CM_MARCH_ED 1
#endif
constexpr int MD = CM_MARCH_MD, ED = CM_MARCH_ED;
static_assert((MD == 1 || MD == 2) && (ED == 1 || ED == 2), "prefetch depth 1 or 2");

// order (nullptr: TileMap order): the tile of every block, ~tile for a tile none of whose cells has
// stiffness.  Such a block's structured part is zero at all its output nodes; it runs only the
// epilogue there.  The order puts the tiles with stiffness first, dealt to the XCDs in contiguous
// runs, so the VALU-heavy blocks spread evenly over all CUs.
// The march of block bid (the kernel's block index, or a block of one system in a launch that
// covers several systems, k_multi).
template <typename T, class M, int EM>
__device__ __forceinline__ void march_body(const int bid, GridArgs<T> g, M mat, const T* __restrict__ s, Epi<T> ep,
                                           int kz, TileMap tm, const int* __restrict__ order, SparseArgs<T> sp) {
  if constexpr (M::kind <= 1) {
    const int nbm = tm.ntx * tm.nty * tm.ntz;
    if (bid >= nbm) {   // block-uniform: the sparse pass rides on this launch
      sparse_body<T, M::sri>(sp, g, s, bid - nbm);
      return;
    }
  }
  constexpr int NST = TN * TN * 3;                 // scalars per staged node layer
  constexpr int NSTP = (NST + BLOCK - 1) / BLOCK;  // per thread
  using VL = V3Lds<T>;
  constexpr int NL = TN * TN;                      // nodes per staged layer
  // staged node layers: buffer (kk - (K0 - 1)) & 1 holds layer kk, so that the two z steps of an
  // unrolled pair address fixed buffers
  __shared__ __align__(16) T s_u[2][VL::size(NL)];
  // per thread: the forces its cell and the cell at tx - 1 pass to the node at ty + 1 (xyz, pad):
  // consecutive threads write and read consecutive 16-byte slots (no LDS bank conflicts)
  static_assert(TILE == 16, "one tile row per 16-lane DPP row");
  __shared__ __align__(16) T s_g[VL::size(BLOCK)];
  __shared__ T s_tL[M::kind == 0 ? 256 : 1], s_tM[M::kind == 0 ? 256 : 1];
  const int tid = threadIdx.x;
  const int tx = tid % TILE, ty = tid / TILE;
  int bx, by, bz;
  bool empty = false;
  if (order) {
    int id = order[bid];
    empty = id < 0;
    id = empty ? ~id : id;
    bx = id % tm.ntx;
    by = (id / tm.ntx) % tm.nty;
    bz = id / (tm.ntx * tm.nty);
  } else {
    tm.decode(bid, bx, by, bz);
  }
  const int I0 = bx * TOUT, J0 = by * TOUT, K0 = bz * kz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long nxn = nx + 1, nyn = ny + 1, lay = nxn * nyn;
  const long long ncl = (long long)nx * ny;
  const bool want_dot = (EM == EPI_APPLY) && ep.part != nullptr;
  double dsum = 0.0;     // s . (A s) over this block's output nodes (APPLY with part)
  const int ci = I0 - 1 + tx, cj = J0 - 1 + ty;
  // the thread's output node (cj, ci) on every layer, or none (halo threads, past the grid)
  const bool own = tx >= 1 && ty >= 1 && ci <= nx && cj <= ny;
  if (empty) {   // block-uniform
    if (own) {
      const int kend = min(K0 + kz, nz + 1);
      for (int k = K0; k < kend; ++k) {
        const long long n = (long long)k * lay + (long long)cj * nxn + ci;
        T au[3] = {T(0), T(0), T(0)}, sn[3] = {T(0), T(0), T(0)};
        if (EM == EPI_CHEB || want_dot) {
#pragma unroll
          for (int c = 0; c < 3; ++c) sn[c] = s[3 * n + c];
        }
        add_extra<T, M>(mat, n, au);
        if (want_dot) dsum += (double)sn[0] * au[0] + (double)sn[1] * au[1] + (double)sn[2] * au[2];
        epilogue<T, EM>(ep, n, au, sn);
      }
    }
    if (want_dot) {
      dsum = block_sum<BLOCK>(dsum);
      if (tid == 0) ep.part[bid] = dsum;
    }
    return;
  }
  if constexpr (M::kind == 0) {
    // label LAB_NONE (no stiffness, or an irregular cell of the sparse pass) reads zero weights,
    // so every cell runs the same branch-free operator
    const T sc = M::iso ? g.iso : T(1);
    s_tL[tid] = tid == LAB_NONE ? T(0) : mat.tabL[tid] * sc;
    s_tM[tid] = tid == LAB_NONE ? T(0) : mat.tabM[tid] * sc;
  }
  const bool cin = (ci >= 0 && ci < nx && cj >= 0 && cj < ny);
  const long long cxy = cin ? ci + (long long)nx * cj : 0;
  T gx = T(0), gy = T(0), axy = T(0);
  if (!M::iso && cin) {
    const T hx = g.hx[ci], hy = g.hy[cj];
    gx = T(0.25) / hx;
    gy = T(0.25) / hy;
    axy = hx * hy;
  }
  // kind 4: lam second moments of the thread's cell column and row
  T qxv = T(0), qyv = T(0);
  if constexpr (M::kind == 4) {
    if (cin) {
      qxv = mat.qx[ci];
      qyv = mat.qy[cj];
    }
  }
  // staging plan (fixed for all layers): scalar e = tid + m*BLOCK of a layer is node row r,
  // entry q of the row's 3*TN contiguous scalars.  A slot outside the grid reads 0 (offset OOB);
  // a slot past the layer (e >= NST, only in the last round) writes the unused 4th value of a
  // node's FP32 LDS slot, or nothing in FP64.
  static_assert((NSTP - 1) * BLOCK < NST && NST <= NSTP * BLOCK, "staging rounds");
  unsigned goff[NSTP];
  int loff[NSTP];
#pragma unroll
  for (int m = 0; m < NSTP; ++m) {
    const int e = tid + m * BLOCK;
    const int r = e / (TN * 3), q = e - r * (TN * 3);
    const int gj = J0 - 1 + r, gi = I0 - 1 + q / 3;
    const bool sv = e < NST && gj >= 0 && gj <= ny && gi >= 0 && gi <= nx;
    goff[m] = sv ? (unsigned)(((gj * nxn + gi) * 3 + q % 3) * sizeof(T)) : OOB;
    loff[m] = e < NST ? VL::idx(NL, r * TN + q / 3, q % 3) : (VL::slots ? (tid % NL) * 4 + 3 : -1);
  }
  const long long lbytes = 3 * lay * (long long)sizeof(T);   // one node layer of s
  // node layers in flight: layer kk is loaded MD z steps before it is staged in LDS, into
  // register set (kk - K0) % MD (MD = 1 or 2; the two z steps of an unrolled pair use fixed sets)
  T pre[MD][NSTP];
  auto gload = [&](int kk, int set) {
    const Rsrc r = layer_rs(s, kk * lbytes, lbytes, kk >= 0 && kk <= nz);
#pragma unroll
    for (int m = 0; m < NSTP; ++m) bld(r, goff[m], pre[set][m]);
  };
  auto lstore = [&](int buf, int set) {
    T* dst = s_u[buf];
#pragma unroll
    for (int m = 0; m < NSTP; ++m) {
      if (VL::slots || loff[m] >= 0) dst[loff[m]] = pre[set][m];
    }
  };
  // xy transform of the cell's face on the node layer staged in buffer buf: f[4*comp + (mx + 2 my)]
  auto face = [&](int buf, T* f) {
    const T* src = s_u[buf];
    T v[4][3];
#pragma unroll
    for (int ab = 0; ab < 4; ++ab) VL::get(src, NL, (ty + (ab >> 1)) * TN + tx + (ab & 1), v[ab]);
    face_fwd(v, f);
  };
  // material of cell layer kk, loaded one step ahead (kinds 0, 1 and 4), 0 for cells outside the
  // grid
  const unsigned coff = cin ? (unsigned)cxy : OOB;                   // byte offset of a label
  const unsigned coffT = cin ? (unsigned)(cxy * sizeof(T)) : OOB;     // of a T value
  const long long clb = ncl * (long long)sizeof(T);                    // one cell layer of T values
  // level 0 loads its materials MM = MD steps ahead (set (kk - K0 + 1) % MM), level 1 one step
  constexpr int MM = M::kind <= 1 ? MD : 1;
  uint8_t labn[MM];
  bool labok[MM];       // block-uniform: labn belongs to a cell layer of the grid
  T lamn[MM], mun[MM];
  auto mload = [&](int kk, int set) {
    const bool okz = kk >= 0 && kk < nz;
    if constexpr (M::kind == 0) {
      // a cell outside the grid or an empty layer reads 0, a real label: the step that uses labn
      // replaces it by LAB_NONE there (a select at the load would wait for it)
      labn[set] = __builtin_amdgcn_raw_buffer_load_b8(layer_rs(mat.lab, kk * ncl, ncl, okz), (int)coff, 0, 0);
      labok[set] = okz;
    }
    if constexpr (M::kind == 1 || M::kind == 4) {
      bld(layer_rs(mat.cl, kk * clb, clb, okz), coffT, lamn[set]);
      bld(layer_rs(mat.cm, kk * clb, clb, okz), coffT, mun[set]);
    }
  };

  const int kend = min(K0 + kz, nz + 1);   // exclusive end of the output node layers
  mload(K0 - 1, 0);
  if (MM == 2) mload(K0, 1);
  gload(K0 - 1, 0);
  lstore(0, 0);
  gload(K0, 0);
  if (MD == 2) gload(K0 + 1, MD - 1);
  __syncthreads();
  // epilogue inputs (residual and smoothing steps on level 0) of the thread's output node on layer
  // K0, then always one layer ahead
  constexpr bool PF = (M::kind <= 1) && (EM != EPI_APPLY);
  const unsigned off3 = own ? (unsigned)((cj * nxn + ci) * 3 * sizeof(T)) : OOB;
  // epilogue inputs of output layer k are loaded ED steps ahead into set (k - K0 + 1) % ED
  EpiIn<T> qn[ED];
  if constexpr (PF) {
    epi_load_rs<T, EM>(epi_rs<T, EM>(ep, K0, lay, K0 < kend), ep.x != nullptr, off3, qn[1 % ED]);
    if (ED == 2) epi_load_rs<T, EM>(epi_rs<T, EM>(ep, K0 + 1, lay, K0 + 1 < kend), ep.x != nullptr, off3, qn[0]);
  }
  // LDS slot of the forces the row below passes to this thread's node
  const int i2 = ty >= 1 ? tid - TILE : tid;
  T uA[12], uB[12], cA[12], cB[12];
  face(0, uA);
#pragma unroll
  for (int q = 0; q < 12; ++q) cA[q] = T(0);

  // one z step: cell layer k (bottom face transform ulo, top face computed into uhi), carry in car
  // (top-face forces of cell k-1), carry out carn (top-face forces of cell k), output node layer k
  auto step = [&](auto PAR, const int k, T(&ulo)[12], T(&uhi)[12], T(&car)[12], T(&carn)[12]) {
    constexpr int BLO = decltype(PAR)::value, BHI = 1 - BLO;   // buffers of layers k and k+1
    constexpr int SET = MD == 2 ? BLO : 0, QS = ED == 2 ? BLO : 0, MSET = MM == 2 ? BLO : 0;
    // Issue order of the memory operations of a step: the material MM layers ahead, node layer
    // k+1+MD, (after the cell work) the epilogue inputs ED layers ahead, then this step's stores.
    // Each prefetched value is then waited for behind loads only, never behind a store.
    // this step's cell material
    bool act = true;
    CellW<T> w{gx, gy, T(0), T(0), T(0)};
    if constexpr (M::kind == 0) {
      const uint8_t lb = (cin && labok[MSET]) ? labn[MSET] : LAB_NONE;
      w.wl = s_tL[lb];
      w.wm = s_tM[lb];
    } else if constexpr (M::kind == 1 || M::kind == 4) {
      const T sc = M::iso ? g.iso : T(1);
      w.wl = lamn[MSET] * sc;
      w.wm = mun[MSET] * sc;
    } else {
      act = cin && k >= 0 && k < nz;
    }
    lstore(BHI, SET);         // node layer k+1, loaded MD steps ago
    if constexpr (M::kind != 2) mload(k + MM, MSET);
    gload(k + 1 + MD, SET);   // past kend: an empty descriptor (no traffic)
    __syncthreads();
    face(BHI, uhi);
    T sn[3] = {T(0), T(0), T(0)};
    if (EM == EPI_CHEB || want_dot) VL::get(s_u[BLO], NL, ty * TN + tx, sn);
    T Flo[12];
    if constexpr (M::kind <= 1 && M::iso && M::sri) {
      // level 0, cubic cells, SRI: the face form of the cell operator (carry included)
      cell_faces_iso_sri<T>(ulo, uhi, car, Flo, carn, w.wl, w.wm);
    } else if constexpr (M::kind <= 1) {
      // level 0: every lane computes its cell (zero weights where there is no stiffness)
      T c3[3][8], F3[3][8];
      faces_to_cell(ulo, uhi, c3);
      if (!M::iso) {
        const T hz = g.hz[min(max(k, 0), nz - 1)];   // layers past the grid have zero weights
        const T V = axy * hz;
        w.gz = T(0.25) / hz;
        w.wl *= V;
        w.wm *= V;
      }
      elem_mid<T, M::sri, M::iso>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w);
      cell_to_faces(F3, Flo, carn);
    } else if constexpr (M::kind == 4) {
      // level 1, cells of one material (zero weights elsewhere: the heterogeneous-cell pass)
      T c3[3][8], F3[3][8];
      faces_to_cell(ulo, uhi, c3);
      const int kc = min(max(k, 0), nz - 1);   // layers past the grid have zero weights
      w.gz = T(0.25) / g.hz[kc];
      elem_mid_q<T>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w, qxv, qyv, mat.qz[kc]);
      cell_to_faces(F3, Flo, carn);
    } else {
#pragma unroll
      for (int q = 0; q < 12; ++q) {
        Flo[q] = T(0);
        carn[q] = T(0);
      }
      if (act) {
        T c3[3][8];
        faces_to_cell(ulo, uhi, c3);
        T F3[3][8];
        // kind 2: moments of the cell (zero for a cell without stiffness)
        const long long ncell = ncl * nz;
        const long long cid = cxy + ncl * k;
        T mL[27], mM[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) {
          mL[q] = mat.mom[q * ncell + cid];
          mM[q] = mat.mom[(27 + q) * ncell + cid];
        }
        const bool done = (mL[0] != T(0)) || (mM[0] != T(0));
        if (done) elem_mid_mom<T>(c3, F3, mL, mM, gx, gy, T(0.25) / g.hz[k]);
        if (done) cell_to_faces(F3, Flo, carn);
      }
    }
    if constexpr (!(M::kind <= 1 && M::iso && M::sri)) {
#pragma unroll
      for (int q = 0; q < 12; ++q) Flo[q] += car[q];
    }
    // back to the nodes in two butterflies (face_bwd split by axis): the x butterfly adds the
    // corners of the cell at tx - 1, read from the previous lane of the DPP row; R = corners 0 and
    // 1 that meet at this thread's node, S = corners 2 and 3 that meet at the node at ty + 1, passed
    // through LDS
    T R[3], S[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T t0 = Flo[4 * c] - Flo[4 * c + 1], t1 = Flo[4 * c] + Flo[4 * c + 1];
      const T t2 = Flo[4 * c + 2] - Flo[4 * c + 3], t3 = Flo[4 * c + 2] + Flo[4 * c + 3];
      const T P = t0 + row_prev(t1), Q = t2 + row_prev(t3);
      R[c] = P - Q;
      S[c] = P + Q;
    }
    VL::put(s_g, BLOCK, tid, S);
    __syncthreads();
    // Every step issues the same memory operations (those of the first step, which has no output
    // layer, and those of arrays the caller does not keep go to empty descriptors): with the same
    // count of memory operations on every path, the compiler's wait before each use of a prefetched
    // value covers only that value, instead of draining all loads in flight.
    {   // halo lanes compute and store nowhere (offset OOB)
      T p2[3];
      VL::get(s_g, BLOCK, i2, p2);   // cells (tx-1 .. tx, ty-1): their corners 3 and 2
      T au[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) au[c] = R[c] + p2[c];
      if constexpr (M::kind >= 2) {
        if (own && k >= K0) add_extra<T, M>(mat, (long long)k * lay + (long long)cj * nxn + ci, au);
      }
      if (want_dot && own && k >= K0) dsum += (double)sn[0] * au[0] + (double)sn[1] * au[1] + (double)sn[2] * au[2];
      const EpiRs<T> ro = epi_rs<T, EM>(ep, k, lay, k >= K0);
      EpiOut<T> o;
      if constexpr (PF) {
        epi_compute<T, EM>(ep, au, sn, qn[QS], o);
        epi_load_rs<T, EM>(epi_rs<T, EM>(ep, k + ED, lay, k + ED < kend), ep.x != nullptr, off3, qn[QS]);
      } else {
        EpiIn<T> q;
        epi_load_rs<T, EM>(ro, ep.x != nullptr, off3, q);
        epi_compute<T, EM>(ep, au, sn, q, o);
      }
      epi_store_rs<T, EM>(ro, off3, o);
    }
  };
  int k = K0 - 1;
  for (; k + 1 < kend; k += 2) {
    step(Par<0>{}, k, uA, uB, cA, cB);
    step(Par<1>{}, k + 1, uB, uA, cB, cA);
  }
  if (k < kend) step(Par<0>{}, k, uA, uB, cA, cB);
  if (want_dot) {
    dsum = block_sum<BLOCK>(dsum);
    if (tid == 0) ep.part[bid] = dsum;
  }
}

template <typename T, class M, int EM>
__global__ __launch_bounds__(BLOCK) void k_march(GridArgs<T> g, M mat, const T* __restrict__ s, Epi<T> ep, int kz,
                                                 TileMap tm, const int* __restrict__ order, SparseArgs<T> sp) {
  march_body<T, M, EM>(blockIdx.x, g, mat, s, ep, kz, tm, order, sp);
}
// launch record of one march (k_one / k_multi)
template <typename T, class M, int EM>
struct MarchL {
  GridArgs<T> g;
  M mat;
  const T* s;
  Epi<T> ep;
  int kz;
  TileMap tm;
  const int* order;
  SparseArgs<T> sp;
  __device__ __forceinline__ void run(int bid, int) const { march_body<T, M, EM>(bid, g, mat, s, ep, kz, tm, order, sp); }
};

// Several-material cells of a level-1 split march (SplitMat): one thread per listed cell forms the
// cell's moments from its child weights, applies the moment operator to s and writes the cell's 24
// forces to hf[cell][corner][comp] (the march epilogue gathers them per node).
template <typename T>
struct HetArgs {
  int n;                             // listed cells (0: none)
  int sri;
  const int* __restrict__ cells;     // x-fastest cell ids
  int nx, ny, nz;
  const T* __restrict__ hx;
  const T* __restrict__ hy;
  const T* __restrict__ hz;
  const T* __restrict__ cw;          // [16][ncell] child weights
  const T* __restrict__ phx;
  const T* __restrict__ phy;
  const T* __restrict__ phz;
  T* __restrict__ hf;
};
template <typename T, bool SRI>
__device__ __forceinline__ void het_cell(const HetArgs<T>& h, int t, const T* __restrict__ s) {
  const long long ncl = (long long)h.nx * h.ny, ncell = ncl * h.nz;
  const long long c = h.cells[t];
  const int i = (int)(c % h.nx), j = (int)((c / h.nx) % h.ny), k = (int)(c / ncl);
  const long long nxn = h.nx + 1, lay = nxn * (h.ny + 1);
  T cc[3][8];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const long long nd = (i + (l & 1)) + nxn * (j + ((l >> 1) & 1)) + lay * (k + (l >> 2));
#pragma unroll
    for (int q = 0; q < 3; ++q) cc[q][l] = s[3 * nd + q];
  }
  wht_fwd(cc[0]); wht_fwd(cc[1]); wht_fwd(cc[2]);
  T w[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) w[q] = h.cw[q * ncell + c];
  T fxL[2][2], fxM[2][2], fyL[2][2], fyM[2][2], fzL[2][2], fzM[2][2];
  child_factors<T, SRI>(h.phx + 6 * i, fxL, fxM);
  child_factors<T, SRI>(h.phy + 6 * j, fyL, fyM);
  child_factors<T, SRI>(h.phz + 6 * k, fzL, fzM);
  const T gx = T(0.25) / h.hx[i], gy = T(0.25) / h.hy[j], gz = T(0.25) / h.hz[k];
  T F[3][8];
  {
    T mL[27];
    child_moments<T>(w, fxL, fyL, fzL, mL);
    elem_mid_mom_lam<T>(cc, F, mL, gx, gy, gz);
  }
  {
    T mM[27];
    child_moments<T>(w + 8, fxM, fyM, fzM, mM);
    elem_mid_mom_mu<T>(cc, F, mM, gx, gy, gz);
  }
  wht_bwd(F[0]); wht_bwd(F[1]); wht_bwd(F[2]);
#pragma unroll
  for (int l = 0; l < 8; ++l)
#pragma unroll
    for (int q = 0; q < 3; ++q) h.hf[(8LL * t + l) * 3 + q] = F[q][l];
}

// Block flags of the structured march for z extent kz: one block per tile id (x fastest, then y,
// then z), flag = 1 when any cell the march block reads (cell columns I0-1 .. I0+14, rows
// J0-1 .. J0+14, layers K0-1 .. K0+kz-1) has stiffness.
template <typename T, class M>
__global__ __launch_bounds__(BLOCK) void k_block_flags(GridArgs<T> g, M mat, int kz, int ntx, int nty,
                                                       uint8_t* __restrict__ flag) {
  const int id = blockIdx.x;
  const int bx = id % ntx, by = (id / ntx) % nty, bz = id / (ntx * nty);
  const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
  const int ci = bx * TOUT - 1 + tx, cj = by * TOUT - 1 + ty;
  const long long ncl = (long long)g.nx * g.ny, ncell = ncl * g.nz;
  int any = 0;
  if (ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny) {
    const int k0 = max(bz * kz - 1, 0), k1 = min(bz * kz + kz, g.nz);
    for (int k = k0; k < k1 && !any; ++k) {
      const long long cid = ci + (long long)g.nx * cj + ncl * k;
      if constexpr (M::kind == 0) any = mat.lab[cid] != LAB_NONE;
      if constexpr (M::kind == 1) any = mat.cl[cid] != T(0) || mat.cm[cid] != T(0);
      if constexpr (M::kind == 2) any = mat.mom[cid] != T(0) || mat.mom[27 * ncell + cid] != T(0);
      if constexpr (M::kind == 4) any = mat.cl[cid] != T(0) || mat.cm[cid] != T(0);
    }
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) flag[id] = any ? 1 : 0;
}

// -------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void coarse_extra_body(const int bid, const CoarseExtra<T>& a, const T* __restrict__ s,
                                                  T* __restrict__ ex) {
  if (bid < a.nb_term) {
    const int rw = a.rw;
    const long long gid = (long long)bid * blockDim.x + threadIdx.x;
    const int t = (int)(gid / rw), lane = (int)(gid % rw);
    if (t >= a.nterm) return;   // the rw lanes of a term leave together
    const int q0 = a.r1ptr[t], q1 = a.r1ptr[t + 1];
    const T kt = a.r1k[t];
    // the first rw entries stay in registers between the dot and the scatter
    const int qf = q0 + lane;
    const bool has = qf < q1;
    const int df = has ? a.r1dof[qf] : 0;
    const T wf = has ? a.r1w[qf] : T(0);
    T acc = has ? wf * s[df] : T(0);
    for (int q = qf + rw; q < q1; q += rw) acc += a.r1w[q] * s[a.r1dof[q]];
    for (int off = rw / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, rw);
    const T dk = kt * acc;
    if (has) atomicAdd(ex + df, dk * wf);
    for (int q = qf + rw; q < q1; q += rw) atomicAdd(ex + a.r1dof[q], dk * a.r1w[q]);
    return;
  }
  const int e = (bid - a.nb_term) * blockDim.x + threadIdx.x;
  if (e >= a.nblk) return;
  const long long m = a.bcol[e], n = a.brow[e];
  const T* B = a.bval + 9LL * e;
  const T s0 = s[3 * m], s1 = s[3 * m + 1], s2 = s[3 * m + 2];
#pragma unroll
  for (int c = 0; c < 3; ++c) atomicAdd(ex + 3 * n + c, B[3 * c] * s0 + B[3 * c + 1] * s1 + B[3 * c + 2] * s2);
}
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_extra(CoarseExtra<T> a, const T* __restrict__ s, T* __restrict__ ex) {
  coarse_extra_body<T>(blockIdx.x, a, s, ex);
}
// launch record of the coarse springs (blocks 0 .. nb_extra-1) and, on level 1, the
// several-material cells (the blocks after them)
template <typename T>
struct CoarseExtraL {
  CoarseExtra<T> a;
  const T* s;
  T* ex;
  int nb_extra;
  HetArgs<T> h;
  __device__ __forceinline__ void run(int bid, int) const {
    if (bid < nb_extra) {
      coarse_extra_body<T>(bid, a, s, ex);
      return;
    }
    const int t = (bid - nb_extra) * (int)blockDim.x + (int)threadIdx.x;
    if (t >= h.n) return;
    if (h.sri) het_cell<T, true>(h, t, s);
    else het_cell<T, false>(h, t, s);
  }
};

// Fix-up after the structured march, one thread per touched node: adds the sparse part tc (and
// clears it for the next pass).  Grid nodes got the structured epilogue without tc; the epilogue
// is affine in A s, so the correction is -tc (RESID), +tc (APPLY), or for a smoothing step
// r -= t, d -= c2 Dinv t, x -= c2 Dinv t.  Split copies have no structured part: their epilogue
// runs here in full with A s = tc.
//
// Smoothing steps with split-copy groups (threads past ntouch, one per group): the members' 3x3
// blocks in Dinv are zero, so the march and the copies' step below leave d = c1 sn and
// x += xa sn + d; the group thread completes every member's residual r (ep.y, always kept when
// groups exist), then adds c2 Ginv r_group to d and x.
// MC > 0: a group of exactly MC members, every loop unrolled (the residual stays in registers and
// the rows of Ginv load together); MC = 0: any size up to GROUP_MAX.
template <typename T, int MC>
__device__ __forceinline__ void group_fixup_m(const CopyGroups<T>& ga, int g, long long nng, T* __restrict__ tc,
                                              const T* __restrict__ s, const Epi<T>& ep) {
  const int b = ga.gptr[g], m = MC > 0 ? MC : ga.gptr[g + 1] - b, N = 3 * m;
  // every index of rg is a compile-time constant (registers, no scratch): loops run to the bound
  // and stop at m
  constexpr int MB = MC > 0 ? MC : GROUP_MAX;
  T rg[3 * MB];
#pragma unroll
  for (int j = 0; j < MB; ++j) {
    if (MC == 0 && j >= m) break;
    const int slot = ga.gslot[b + j];
    const long long n = ga.tnode[slot];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T t = tc[3 * slot + q];
      tc[3 * slot + q] = T(0);
      T r;
      if (n >= nng) {   // split copy: its whole step, with a zero 3x3 block
        const T sn = s[3 * n + q];
        r = ep.rin[3 * n + q] - t;
        const T dn = ep.c1 * sn;
        if (ep.dout) ep.dout[3 * n + q] = dn;
        if (ep.x) ep.x[3 * n + q] += ep.xa * sn + dn;
      } else {          // grid node: the march wrote r without the sparse part t
        r = ep.y[3 * n + q] - t;
      }
      ep.y[3 * n + q] = r;
      rg[3 * j + q] = r;
    }
  }
  const T* G = ga.ginv + ga.goff[g];
#pragma unroll
  for (int j = 0; j < MB; ++j) {
    if (MC == 0 && j >= m) break;
    const long long n = ga.tnode[ga.gslot[b + j]];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      T acc = T(0);
      const T* row = G + (size_t)(3 * j + q) * N;
#pragma unroll
      for (int k = 0; k < 3 * MB; ++k) {
        if (MC == 0 && k >= N) break;
        acc += row[k] * rg[k];
      }
      const T dz = ep.c2 * acc;
      if (ep.dout) ep.dout[3 * n + q] += dz;
      if (ep.x) ep.x[3 * n + q] += dz;
    }
  }
}
template <typename T>
__device__ void group_fixup(const CopyGroups<T>& ga, int g, long long nng, T* __restrict__ tc, const T* __restrict__ s,
                            const Epi<T>& ep) {
  const int m = ga.gptr[g + 1] - ga.gptr[g];
  if (m == 2) group_fixup_m<T, 2>(ga, g, nng, tc, s, ep);
  else if (m == 3) group_fixup_m<T, 3>(ga, g, nng, tc, s, ep);
  else group_fixup_m<T, 0>(ga, g, nng, tc, s, ep);
}

template <typename T, int EM>
__device__ __forceinline__ void fixup_body(const int bid, int ntouch, const int* __restrict__ tnode, long long nng,
                                           T* __restrict__ tc, const T* __restrict__ s, const Epi<T>& ep,
                                           const CopyGroups<T>& ga) {
  const int i = bid * blockDim.x + threadIdx.x;
  if (EM == EPI_APPLY && ep.part) {
    // s . tc over the touched nodes: A s = structured part (its dot came from the march) + tc
    double d = 0.0;
    if (i < ntouch) {
      const long long n = tnode[i];
#pragma unroll
      for (int c = 0; c < 3; ++c) d += (double)s[3 * n + c] * (double)tc[3 * i + c];
    }
    d = block_sum<256>(d);
    if (threadIdx.x == 0) ep.part[bid] = d;
  }
  if (EM == EPI_CHEB && ga.ng) {
    if (i >= ntouch) {
      if (i - ntouch < ga.ng) group_fixup<T>(ga, i - ntouch, nng, tc, s, ep);
      return;
    }
    if (ga.gmem[i]) return;
  }
  if (i >= ntouch) return;
  const long long n = tnode[i];
  T t[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    t[c] = tc[3 * i + c];
    tc[3 * i + c] = T(0);
  }
  if (n >= nng) {
    T sn[3] = {T(0), T(0), T(0)};
    if (EM == EPI_CHEB) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sn[c] = s[3 * n + c];
    }
    epilogue<T, EM>(ep, n, t, sn);
    if (EM == EPI_RESID && ep.fold) {
      const long long o = ep.fold[n - nng];
#pragma unroll
      for (int c = 0; c < 3; ++c) atomicAdd(&ep.y[3 * o + c], ep.y[3 * n + c]);
    }
    return;
  }
  if (EM == EPI_APPLY) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ep.y[3 * n + c] += t[c];
  } else if (EM == EPI_RESID) {
    if (ep.fold) {   // copies of this node may add to it concurrently
#pragma unroll
      for (int c = 0; c < 3; ++c) atomicAdd(&ep.y[3 * n + c], -t[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.y[3 * n + c] -= t[c];
    }
  } else {
    if (ep.y) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.y[3 * n + c] -= t[c];
    }
    T z[3];
    sym3_mul(ep.Dinv + 6 * n, t, z);
    if (ep.dout) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.dout[3 * n + c] -= ep.c2 * z[c];
    }
    if (ep.x) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep.x[3 * n + c] -= ep.c2 * z[c];
    }
  }
}

template <typename T, int EM>
__global__ __launch_bounds__(256) void k_fixup(int ntouch, const int* __restrict__ tnode, long long nng,
                                               T* __restrict__ tc, const T* __restrict__ s, Epi<T> ep,
                                               CopyGroups<T> ga = CopyGroups<T>{}) {
  fixup_body<T, EM>(blockIdx.x, ntouch, tnode, nng, tc, s, ep, ga);
}
template <typename T, int EM>
struct FixupL {
  int ntouch;
  const int* tnode;
  long long nng;
  T* tc;
  const T* s;
  Epi<T> ep;
  CopyGroups<T> ga;
  __device__ __forceinline__ void run(int bid, int) const { fixup_body<T, EM>(bid, ntouch, tnode, nng, tc, s, ep, ga); }
};

template <typename T>
__global__ void k_fill(T* __restrict__ p, long long n, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
template <typename T>
struct FillL {
  T* p;
  long long n;
  T v;
  __device__ __forceinline__ void run(int bid, int) const {
    const long long i = (long long)bid * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
  }
};

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
__device__ __forceinline__ void restrict_body(const int bid, int nfx, int nfy, int nfz, int ncx, int ncy, int ncz,
                                              AxisR<T> rx, AxisR<T> ry, AxisR<T> rz, const T* __restrict__ r,
                                              T* __restrict__ rc, const T* __restrict__ dinv_c, T* __restrict__ xc,
                                              T c0) {
  const long long id = (long long)bid * blockDim.x + threadIdx.x;
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
  if (xc) {   // the coarse level's first smoothing step from zero: xc = c0 Dinv rc (as k_cheb0)
    const T* m = dinv_c + 6 * id;
    xc[3 * id] = c0 * (m[0] * s0 + m[1] * s1 + m[2] * s2);
    xc[3 * id + 1] = c0 * (m[1] * s0 + m[3] * s1 + m[4] * s2);
    xc[3 * id + 2] = c0 * (m[2] * s0 + m[4] * s1 + m[5] * s2);
  }
}
template <typename T>
__global__ void k_restrict(int nfx, int nfy, int nfz, int ncx, int ncy, int ncz, AxisR<T> rx, AxisR<T> ry,
                           AxisR<T> rz, const T* __restrict__ r, T* __restrict__ rc, const T* __restrict__ dinv_c,
                           T* __restrict__ xc, T c0) {
  restrict_body<T>(blockIdx.x, nfx, nfy, nfz, ncx, ncy, ncz, rx, ry, rz, r, rc, dinv_c, xc, c0);
}
template <typename T>
struct RestrictL {
  int nfx, nfy, nfz, ncx, ncy, ncz;
  AxisR<T> rx, ry, rz;
  const T* r;
  T* rc;
  const T* dinv_c;
  T* xc;
  T c0;
  __device__ __forceinline__ void run(int bid, int) const {
    restrict_body<T>(bid, nfx, nfy, nfz, ncx, ncy, ncz, rx, ry, rz, r, rc, dinv_c, xc, c0);
  }
};

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

// x += P ec on used fine grid nodes; unused nodes are set to zero (they are not unknowns).  The
// split copies (ncopy, interpolated like their originals, rb/splitfe.py:153) are the blocks after
// the grid nodes' nbg blocks of the same launch: both write disjoint nodes.
template <typename T>
__device__ __forceinline__ void prolong_body(const int bid, int nfx, int nfy, int nfz, int ncx, int ncy, AxisP<T> px,
                                             AxisP<T> py, AxisP<T> pz, const T* __restrict__ ec,
                                             const uint8_t* __restrict__ used, T* __restrict__ x, int nbg, int ncopy,
                                             const int* __restrict__ orig) {
  const long long nn = (long long)nfx * nfy * nfz;
  if (bid >= nbg) {   // block-uniform
    const int q = (bid - nbg) * blockDim.x + threadIdx.x;
    if (q >= ncopy) return;
    const int o = orig[q];
    const int i = o % nfx, j = (o / nfx) % nfy, k = o / (nfx * nfy);
    T e[3];
    interp_node(i, j, k, ncx, ncy, px, py, pz, ec, e);
    const long long n = nn + q;
    x[3 * n] += e[0]; x[3 * n + 1] += e[1]; x[3 * n + 2] += e[2];
    return;
  }
  const long long id = (long long)bid * blockDim.x + threadIdx.x;
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


template <typename T>
__global__ void k_prolong_add(int nfx, int nfy, int nfz, int ncx, int ncy, AxisP<T> px, AxisP<T> py, AxisP<T> pz,
                              const T* __restrict__ ec, const uint8_t* __restrict__ used, T* __restrict__ x,
                              int nbg, int ncopy, const int* __restrict__ orig) {
  prolong_body<T>(blockIdx.x, nfx, nfy, nfz, ncx, ncy, px, py, pz, ec, used, x, nbg, ncopy, orig);
}
template <typename T>
struct ProlongL {
  int nfx, nfy, nfz, ncx, ncy;
  AxisP<T> px, py, pz;
  const T* ec;
  const uint8_t* used;
  T* x;
  int nbg, ncopy;
  const int* orig;
  __device__ __forceinline__ void run(int bid, int) const {
    prolong_body<T>(bid, nfx, nfy, nfz, ncx, ncy, px, py, pz, ec, used, x, nbg, ncopy, orig);
  }
};

}  // namespace cm
