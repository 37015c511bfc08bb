// Engine preparation kernels: lattice material mixing and split connectivity.
//
// Lattice: every cell (reference order c = (i ny + j) nz + k) holds ns3 sub-sample body labels.
// Per cell: class fractions F (10 tissue classes), bone-group fractions S, the mixing rule of
// rb/rbfe.py:75-134 (Voigt average in plain cells; hard phases in series with joint layers,
// gaps and soft tissue in joint cells), joint detection (>= 2 bone groups within the 3x3x3
// neighbourhood, or a joint layer present), activity and the bone part (rb/rbfe.py:140-141).
//
// Split connectivity (rb/splitfe.py:261-305): a node touched by cells of >= 2 bones gets one
// id per bone; the lowest bone keeps the grid id, every further (node, bone) pair gets
// nnode + rank in (node, bone) order; a part-0 cell at a split node uses the bone with the
// largest share in the cell (first maximum), or, when the cell holds none, the bone with most
// cells at the node.
#pragma once
#include "common.hpp"

namespace cm {

constexpr int LAT_NCLS = 10;       // void cortical cancellous cartilage fascia soft plate fibrous fat gap
constexpr int LAT_MAX_NS3 = 64;    // sub-samples per cell

struct LatticeArgs {
  int nx, ny, nz, ns3;
  const int16_t* __restrict__ L;         // [ncell][ns3] body labels
  const int8_t* __restrict__ lut_cls;    // [nlab] class per label
  const int8_t* __restrict__ lut_grp;    // [nlab] bone group per label (0 = none)
  int nlab, ngrp;                        // ngrp = largest group + 1 (<= 64)
  const double* __restrict__ lam_k;      // [10] per class
  const double* __restrict__ mu_k;
  double* __restrict__ F;                // [ncell][10]
  float* __restrict__ S;                 // [ncell][ngrp]
  double* __restrict__ mix;              // [7][ncell] scratch: lamV muV lam_c mu_c lam_s mu_s occ
  unsigned long long* __restrict__ pres; // [ncell] bone groups present in the cell
  uint8_t* __restrict__ cart;            // [ncell] joint layer present
};

// Class and group fractions, presence mask and the three candidate (lam, mu) of every cell.
__global__ __launch_bounds__(256) void k_lattice_cells(LatticeArgs a) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncell = (long long)a.nx * a.ny * a.nz;
  if (c >= ncell) return;
  const int16_t* Lc = a.L + c * a.ns3;
  int cnt[LAT_NCLS];
#pragma unroll
  for (int k = 0; k < LAT_NCLS; ++k) cnt[k] = 0;
  unsigned long long pm = 0ull;
  for (int s = 0; s < a.ns3; ++s) {
    const int lab = Lc[s];
    const int cl = (lab >= 0 && lab < a.nlab) ? a.lut_cls[lab] : 0;
    const int gr = (lab >= 0 && lab < a.nlab) ? a.lut_grp[lab] : 0;
#pragma unroll
    for (int k = 0; k < LAT_NCLS; ++k) cnt[k] += (cl == k);
    if (gr > 0) pm |= 1ull << gr;
  }
  const double inv = (double)a.ns3;
  double F[LAT_NCLS];
#pragma unroll
  for (int k = 0; k < LAT_NCLS; ++k) {
    F[k] = (double)cnt[k] / inv;
    a.F[c * LAT_NCLS + k] = F[k];
  }
  // bone-group fractions (float, as the reference's share array)
  for (int g = 0; g < a.ngrp; ++g) {
    int n = 0;
    if (g > 0 && ((pm >> g) & 1ull))
      for (int s = 0; s < a.ns3; ++s) {
        const int lab = Lc[s];
        n += ((lab >= 0 && lab < a.nlab) ? a.lut_grp[lab] : 0) == g;
      }
    a.S[c * a.ngrp + g] = (float)((double)n / inv);
  }
  a.pres[c] = pm;
  const double* lk = a.lam_k;
  const double* mk = a.mu_k;
  double Mk[LAT_NCLS];
#pragma unroll
  for (int k = 0; k < LAT_NCLS; ++k) Mk[k] = lk[k] + 2.0 * mk[k];
  // phase sets: hard {1,2,4}, joint layers {3,7}, soft {5,8}, gap {9}; occupied = hard+layers+soft
  // (sums in the reference's order of the occupied set: 1 2 4 3 7 5 8)
  const int PH[7] = {1, 2, 4, 3, 7, 5, 8};
  double occ = 0.0, sl = 0.0, sm = 0.0;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    occ += F[PH[q]];
    sl += F[PH[q]] * lk[PH[q]];
    sm += F[PH[q]] * mk[PH[q]];
  }
  const double hard = F[1] + F[2] + F[4];
  const double lay = F[3] + F[7];
  const double fgap = F[9];
  const double lamV = sl / occ, muV = sm / occ;
  const double lam_h = (F[1] * lk[1] + F[2] * lk[2] + F[4] * lk[4]) / hard;
  const double mu_h = (F[1] * mk[1] + F[2] * mk[2] + F[4] * mk[4]) / hard;
  const double hM = hard > 0.0 ? hard / (lam_h + 2.0 * mu_h) : 0.0;
  const double hG = hard > 0.0 ? hard / mu_h : 0.0;
  const double gapM = F[9] / Mk[9], gapG = F[9] / mk[9];
  const double layM = (F[3] / Mk[3] + F[7] / Mk[7]) + gapM;
  const double layG = (F[3] / mk[3] + F[7] / mk[7]) + gapG;
  const double fch = lay + hard + fgap;
  const double occ_g = occ + fgap;
  const double M_ch = fch / (layM + hM), G_ch = fch / (layG + hG);
  const double lam_soft = F[5] * lk[5] + F[8] * lk[8], mu_soft = F[5] * mk[5] + F[8] * mk[8];
  const double lam_c = (fch * (M_ch - 2.0 * G_ch) + lam_soft) / occ_g;
  const double mu_c = (fch * G_ch + mu_soft) / occ_g;
  const double softM = F[5] / Mk[5] + F[8] / Mk[8], softG = F[5] / mk[5] + F[8] / mk[8];
  const double M_s = occ_g / (softM + hM + gapM), G_s = occ_g / (softG + hG + gapG);
  a.mix[0 * ncell + c] = lamV;
  a.mix[1 * ncell + c] = muV;
  a.mix[2 * ncell + c] = lam_c;
  a.mix[3 * ncell + c] = mu_c;
  a.mix[4 * ncell + c] = M_s - 2.0 * G_s;
  a.mix[5 * ncell + c] = G_s;
  a.mix[6 * ncell + c] = occ;
  a.cart[c] = lay > 0.0;
}

// Joint cells, the cell's (E, nu), activity and bone part.
__global__ __launch_bounds__(256) void k_lattice_finish(LatticeArgs a, double* __restrict__ E, double* __restrict__ nu,
                                                        uint8_t* __restrict__ active, uint8_t* __restrict__ ser,
                                                        int* __restrict__ part) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long ncell = (long long)nx * ny * nz;
  if (c >= ncell) return;
  const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((long long)ny * nz));
  unsigned long long m = 0ull;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int ii = i + di, jj = j + dj, kk = k + dk;
        if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
        m |= a.pres[((long long)ii * ny + jj) * nz + kk];
      }
  const bool crt = a.cart[c] != 0;
  const bool sr = crt || __popcll(m) >= 2;
  const double occ = a.mix[6 * ncell + c];
  const double lamX = crt ? a.mix[2 * ncell + c] : (sr ? a.mix[4 * ncell + c] : a.mix[0 * ncell + c]);
  const double muX = crt ? a.mix[3 * ncell + c] : (sr ? a.mix[5 * ncell + c] : a.mix[1 * ncell + c]);
  double e = muX * (3.0 * lamX + 2.0 * muX) / (lamX + muX);
  double n = lamX / (2.0 * (lamX + muX));
  const bool act = sr ? (occ > 0.15) : (occ >= 0.5);
  if (!act) {
    e = 1.0;
    n = 0.3;
  } else {
    n = n < 0.0 ? 0.0 : (n > 0.4995 ? 0.4995 : n);    // NaN stays NaN (clip semantics)
  }
  // nan -> 1.0 / 0.3, +-inf -> +-largest finite (numpy nan_to_num)
  if (isnan(e)) e = 1.0;
  else if (isinf(e)) e = e > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
  if (isnan(n)) n = 0.3;
  else if (isinf(n)) n = n > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
  E[c] = e;
  nu[c] = n;
  active[c] = act;
  ser[c] = sr;
  // part: the first bone group with the largest share, when >= 0.5 and no joint layer
  const float* Sc = a.S + c * a.ngrp;
  float best = Sc[0];
  int arg = 0;
  for (int g = 1; g < a.ngrp; ++g)
    if (Sc[g] > best) { best = Sc[g]; arg = g; }
  part[c] = (best >= 0.5f && !crt) ? arg : 0;
}

// ---- split connectivity -----------------------------------------------------------------------
// cnt[node][bone] += 1 for every corner of every active bone cell (part > 0)
__global__ __launch_bounds__(256) void k_split_count(int nx, int ny, int nz, const uint8_t* __restrict__ active,
                                                     const int* __restrict__ part, int nb, int* __restrict__ cnt) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncell = (long long)nx * ny * nz;
  if (c >= ncell || !active[c]) return;
  const int p = part[c];
  if (p <= 0) return;
  const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((long long)ny * nz));
  const long long nxn = nx + 1, nyn = ny + 1;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const long long n = (i + (l & 1)) + nxn * ((j + ((l >> 1) & 1)) + nyn * (k + (l >> 2)));
    atomicAdd(&cnt[n * nb + p], 1);
  }
}

// per node: number of extra copies (bones - 1 when >= 2 bones), counted split nodes
__global__ __launch_bounds__(256) void k_split_nodes(long long nn, int nb, const int* __restrict__ cnt,
                                                     int* __restrict__ ncopy, unsigned long long* __restrict__ nsplit) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int b = 0;
  for (int q = 1; q < nb; ++q) b += cnt[n * nb + q] > 0;
  ncopy[n] = b >= 2 ? b - 1 : 0;
  if (b >= 2) atomicAdd(nsplit, 1ull);
}

// Exclusive scan of int counts into 64-bit offsets: per-block scan (1024 items), scan of the
// block totals (one block), add.
constexpr int SCAN_ITEMS = 1024;
__global__ __launch_bounds__(256) void k_scan_blocks(long long n, const int* __restrict__ in, long long* __restrict__ out,
                                                     long long* __restrict__ btot) {
  __shared__ long long sh[256];
  const long long base = (long long)blockIdx.x * SCAN_ITEMS;
  long long v[4], s = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long long idx = base + threadIdx.x * 4 + q;
    v[q] = idx < n ? in[idx] : 0;
    s += v[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  long long run = sh[threadIdx.x] - s;   // exclusive prefix of this thread's 4 items
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long long idx = base + threadIdx.x * 4 + q;
    if (idx < n) out[idx] = run;
    run += v[q];
  }
  if (threadIdx.x == 255) btot[blockIdx.x] = sh[255];
}
__global__ __launch_bounds__(256) void k_scan_totals(long long nblk, long long* __restrict__ btot,
                                                     long long* __restrict__ total) {
  // one block: thread t owns the contiguous chunk [t*per, (t+1)*per) of the block totals
  __shared__ long long sh[256];
  const long long per = (nblk + 255) / 256;
  const long long b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
  long long s = 0;
  for (long long b = b0; b < b1; ++b) s += btot[b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  long long run = sh[threadIdx.x] - s;
  for (long long b = b0; b < b1; ++b) {
    const long long t = btot[b];
    btot[b] = run;
    run += t;
  }
  if (threadIdx.x == 255) *total = sh[255];
}
__global__ __launch_bounds__(256) void k_scan_add(long long n, long long* __restrict__ out,
                                                  const long long* __restrict__ btot) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += btot[idx / SCAN_ITEMS];
}

// copy_orig[base[n] + r] = n for the r-th extra bone of node n
__global__ __launch_bounds__(256) void k_split_orig(long long nn, int nb, const int* __restrict__ cnt,
                                                    const int* __restrict__ ncopy, const long long* __restrict__ base,
                                                    long long* __restrict__ copy_orig) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn || ncopy[n] == 0) return;
  for (int r = 0; r < ncopy[n]; ++r) copy_orig[base[n] + r] = n;
}

// cn[c][l]: grid corner ids, replaced at split nodes by the copy of the bone the cell uses
__global__ __launch_bounds__(256) void k_split_cells(int nx, int ny, int nz, const uint8_t* __restrict__ active,
                                                     const int* __restrict__ part, const float* __restrict__ share,
                                                     int nb, const int* __restrict__ cnt, const int* __restrict__ ncopy,
                                                     const long long* __restrict__ base, long long* __restrict__ cn) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncell = (long long)nx * ny * nz;
  if (c >= ncell) return;
  const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((long long)ny * nz));
  const long long nxn = nx + 1, nyn = ny + 1, nnode = nxn * nyn * (nz + 1);
  const bool act = active[c] != 0;
  const int p = part[c];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const long long n = (i + (l & 1)) + nxn * ((j + ((l >> 1) & 1)) + nyn * (k + (l >> 2)));
    long long id = n;
    if (act && ncopy[n] > 0) {
      const int* cc = cnt + n * nb;
      int b = p;
      if (b == 0) {
        float best_s = -1.0f;
        int bs = 0, bm = 0, best_m = -1;
        for (int q = 1; q < nb; ++q) {
          if (cc[q] <= 0) continue;
          const float sh = share[c * nb + q];
          if (sh > best_s) { best_s = sh; bs = q; }
          if (cc[q] > best_m) { best_m = cc[q]; bm = q; }
        }
        b = best_s > 0.0f ? bs : bm;
      }
      // rank of b among the node's bones; the lowest keeps the grid id
      int rank = -1;
      for (int q = 1; q <= b && q < nb; ++q) rank += cc[q] > 0;
      if (rank > 0) id = nnode + base[n] + (rank - 1);
    }
    cn[c * 8 + l] = id;
  }
}

}  // namespace cm

namespace cm {

// -------------------------------------------------------------------------------------------
// Engine preparation around the solve (rb/rbfe.py:146-178): the node sets of the bones, the bone
// nearest to every ligament end, and the node patch of every end.

// mask[node] |= 1 << p for the 8 nodes of every active cell of part p (1..63); cnt[p] = active
// cells of part p (block histogram in LDS, one global atomic per block and part).
__global__ __launch_bounds__(256) void k_part_mask(long long ncell, const long long* __restrict__ cn,
                                                   const int* __restrict__ part, const uint8_t* __restrict__ active,
                                                   unsigned long long* __restrict__ mask, int* __restrict__ cnt) {
  __shared__ int h[64];
  if (threadIdx.x < 64) h[threadIdx.x] = 0;
  __syncthreads();
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c < ncell) {
    const int p = part[c];
    if (active[c] && p > 0 && p < 64) {
      const unsigned long long bit = 1ULL << p;
#pragma unroll
      for (int l = 0; l < 8; ++l) atomicOr(mask + cn[8 * c + l], bit);
      atomicAdd(&h[p], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(cnt + threadIdx.x, h[threadIdx.x]);
}

// Node coordinates: node n sits where its original orig[n] (identity when orig is null) sits on
// the grid, (xe[i], ye[j], ze[k]) with orig = i + (nx+1)(j + (ny+1)k).
struct NodeGeo {
  int nx, ny;
  const double* __restrict__ xe;
  const double* __restrict__ ye;
  const double* __restrict__ ze;
  const long long* __restrict__ orig;
};
__device__ __forceinline__ double node_dist2(const NodeGeo& g, long long n, double px, double py, double pz) {
  const long long o = g.orig ? g.orig[n] : n;
  const long long nxn = g.nx + 1, nyn = g.ny + 1;
  const double dx = g.xe[o % nxn] - px, dy = g.ye[(o / nxn) % nyn] - py, dz = g.ze[o / (nxn * nyn)] - pz;
  // (dx^2 + dy^2) + dz^2, each operation rounded once: the order numpy sums a 3-vector in
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// d2min[e * nbone + b] = min over the nodes of bone b of |x - p_e|^2 (inf for a bone without
// nodes).  One block per (point, bone).
__global__ __launch_bounds__(256) void k_nearest_bone(int npt, const double* __restrict__ pts, int nbone,
                                                      const long long* __restrict__ bptr,
                                                      const long long* __restrict__ bnodes, NodeGeo g,
                                                      double* __restrict__ d2min) {
  __shared__ double red[4];
  const int e = blockIdx.x / nbone, b = blockIdx.x % nbone;
  const double px = pts[3 * e], py = pts[3 * e + 1], pz = pts[3 * e + 2];
  double m = HUGE_VAL;
  for (long long q = bptr[b] + threadIdx.x; q < bptr[b + 1]; q += 256) m = fmin(m, node_dist2(g, bnodes[q], px, py, pz));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) d2min[(long long)e * nbone + b] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}

// Sorted (ascending) 8 smallest values seen so far: insert v.
__device__ __forceinline__ void top8_insert(double (&t)[8], double v) {
  if (!(v < t[7])) return;
#pragma unroll
  for (int q = 7; q > 0; --q) t[q] = (v < t[q - 1]) ? t[q - 1] : ((v < t[q]) ? v : t[q]);
  t[0] = (v < t[0]) ? v : t[0];
}

// The patch of each spring end e (rb/replica3.py:68-86): the candidate nodes are bone b = bone[e]'s
// nodes that are used (used == null: all); with d the distances from the end point, radius =
// max(4, the min(7, count - 1)-th smallest d) and flag = d <= radius.  One block per end;
// flags[foff[e] + q] for the bone's q-th node (0 for unused nodes), radius[e], count[e] = used
// candidates.
__global__ __launch_bounds__(256) void k_spring_patches(int nend, const double* __restrict__ pts,
                                                        const int* __restrict__ bone,
                                                        const long long* __restrict__ bptr,
                                                        const long long* __restrict__ bnodes,
                                                        const uint8_t* __restrict__ used, NodeGeo g,
                                                        const long long* __restrict__ foff,
                                                        uint8_t* __restrict__ flags, double* __restrict__ radius,
                                                        int* __restrict__ count) {
  __shared__ double pool[256 * 8];
  __shared__ int cnt_s[4];
  __shared__ double r_s;
  const int e = blockIdx.x;
  const int b = bone[e];
  const double px = pts[3 * e], py = pts[3 * e + 1], pz = pts[3 * e + 2];
  const long long q0 = bptr[b], q1 = bptr[b + 1];
  double t[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) t[k] = HUGE_VAL;
  int n = 0;
  for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
    const long long nd = bnodes[q];
    if (used && !used[nd]) continue;
    top8_insert(t, __dsqrt_rn(node_dist2(g, nd, px, py, pz)));
    ++n;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) pool[threadIdx.x * 8 + k] = t[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if ((threadIdx.x & 63) == 0) cnt_s[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = cnt_s[0] + cnt_s[1] + cnt_s[2] + cnt_s[3];
    double s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = HUGE_VAL;
    for (int q = 0; q < 256 * 8; ++q) top8_insert(s, pool[q]);
    const int kth = tot < 8 ? tot - 1 : 7;
    double dk = s[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) dk = (k == kth) ? s[k] : dk;
    r_s = tot > 0 ? fmax(4.0, dk) : -1.0;
    radius[e] = r_s;
    count[e] = tot;
  }
  __syncthreads();
  const double r = r_s;
  for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
    const long long nd = bnodes[q];
    const bool ok = !(used && !used[nd]) && __dsqrt_rn(node_dist2(g, nd, px, py, pz)) <= r;
    flags[foff[e] + (q - q0)] = ok ? 1 : 0;
  }
}

// Fractions of a cell's sub-samples that carry a body of each of nset body sets (at most
// BF_MAX_SETS): mask[b] has bit s set when body b belongs to set s; out[s][c] = count / ns3, the
// FP64 mean of the membership booleans (exact counts, one division).  One thread per cell; the
// 16-bit labels of a cell are contiguous (row of L).
constexpr int BF_MAX_SETS = 16;
__global__ __launch_bounds__(256) void k_body_fractions(long long ncell, int ns3, const int16_t* __restrict__ L,
                                                        int nlab, const uint32_t* __restrict__ mask, int nset,
                                                        double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int cnt[BF_MAX_SETS];
#pragma unroll
  for (int s = 0; s < BF_MAX_SETS; ++s) cnt[s] = 0;
  const int16_t* row = L + c * ns3;
  for (int q = 0; q < ns3; ++q) {
    const int b = row[q];
    const uint32_t m = (b >= 0 && b < nlab) ? mask[b] : 0u;
#pragma unroll
    for (int s = 0; s < BF_MAX_SETS; ++s) cnt[s] += (int)((m >> s) & 1u);
  }
  const double n = (double)ns3;
  for (int s = 0; s < nset; ++s) out[(long long)s * ncell + c] = (double)cnt[s] / n;
}

}  // namespace cm
