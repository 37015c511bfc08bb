// Voxel FE system on one GPU: geometric-multigrid hierarchy (matrix-free fine level, exact
// Galerkin moment cells on coarse levels, dense FP64 coarsest inverse), V-cycle, PCG drivers and
// the extern "C" boundary declared in include/comech_fe.h.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <tuple>
#include <unordered_map>

#include "../../include/comech_fe.h"
#include "kernels_fe.hpp"
#include "kernels_post.hpp"
#include "kernels_prep.hpp"
#include "kernels_setup.hpp"
#include "kernels_vec.hpp"

namespace cm {

static thread_local std::string g_last_error;

// cm_solve_info.flag of a solve stopped by the early-exit rule of cm_sys_set_fallback
constexpr int FLAG_FALLBACK = -2;

// ---------------------------------------------------------------------------------------------
// Library-held scratch of the allocation-free entry point (cm_sample_nodes_noalloc): the two
// stage-1 tables of the node resampling, CM_SAMPLE_NODES_MAX_COLUMNS ints each (32 MiB).  Every
// device the library runs on holds its own copy until the process ends; calls on one device share
// it.  Host code never uses the array's name as an address: the per-device address comes from
// hipGetSymbolAddress with that device current, once, at the entry point's first call there.
static __device__ __align__(256) int scratch[2 * CM_SAMPLE_NODES_MAX_COLUMNS];
static std::mutex scratch_mu;
static std::vector<int*> scratch_dev;

static int* scratch_on_current_device();

// ---------------------------------------------------------------------------------------------
// Per-level, per-precision device data
template <typename T>
struct LevelT {
  DevBuf<T> hx, hy, hz;
  DevBuf<T> tabL, tabM;     // level 0, label mode (256 entries)
  DevBuf<T> cellL, cellM;   // level 0, value mode (x-fastest, irregular cells zeroed)
  DevBuf<T> irrL, irrM;     // level 0 irregular cells
  DevBuf<T> mom;            // levels >= 2: [54][ncell]
  DevBuf<T> cw, ph[3];      // level 1: child weights [16][ncell], per-axis child factor tables
  DevBuf<T> cl1, cm1, q1[3];  // level 1 split: lam V, mu V of one-material cells, lam second moments
  DevBuf<T> hf;             // level 1 split: forces of the several-material cells [cell][corner][3]
  DevBuf<T> Dinv;           // [nn][6] symmetric 3x3 block inverses
  DevBuf<T> tc;             // sparse-pass accumulator, 3 per touched node (kept zero between uses)
  DevBuf<T> bval;           // coarse levels: block values (3x3 row-major per block entry)
  DevBuf<T> ex;             // coarse levels: extra-stiffness accumulator, 3 per node (kept zero)
  DevBuf<T> pw[3];          // prolongation weights, this level as the fine side
  DevBuf<T> rw[3];          // restriction weights of the next level's coarse nodes
  DevBuf<T> xval;           // extra stiffness values
  DevBuf<T> r1w, r1k;       // rank-one terms
  DevBuf<T> x, b, r, t, d;  // work vectors
  DevBuf<T> ginv;           // level 0: dense inverses of the split-copy groups
  double lmax = 0.0, lmin = 0.0;
  bool ready = false;
};

struct Level {
  int nx = 0, ny = 0, nz = 0;
  long long nng = 0, nn = 0, ncell = 0;
  std::vector<double> e[3];
  std::vector<uint8_t> used_h;
  DevBuf<uint8_t> used;
  // transfer to the next level
  std::vector<int> pia_h[3], rs_h[3];
  std::vector<double> pw_h[3], rw_h[3];
  DevBuf<int> pia[3], rs[3];
  // nesting of this level's cells in the next level's cells (for moments)
  std::vector<int> ncs_h[3], ncn_h[3];     // stored on the coarse level: per coarse cell
  std::vector<double> na_h[3], nb_h[3];    // stored on the coarse level: per fine cell
  DevBuf<int> ncs[3], ncn[3];
  DevBuf<double> na[3], nb[3];
  // extra stiffness (CSR over unique rows, full DOF index of this level)
  std::vector<int> xrow_h, xptr_h, xcol_h;
  std::vector<double> xval_h;
  DevBuf<int> xrow, xptr, xcol;
  int xw = 1;                 // lanes per extra row
  // rank-one extra stiffness sum_t k_t w_t w_t^T (springs; exact coarsening w -> P^T w):
  // t* = all terms of this level (host, coarsening); r1* = the terms applied as rank-one (support
  // wider than one 3x3x3 node neighbourhood), the rest is folded into the CSR
  std::vector<int> tptr_h, tdof_h;
  std::vector<double> tw_h, tk_h;
  // generic CSR part (off-diagonal input entries and their Galerkin products), host
  std::vector<std::tuple<long long, double, double>> gen_h;
  std::vector<int> r1ptr_h, r1dof_h;
  std::vector<double> r1w_h, r1k_h;
  DevBuf<int> r1ptr, r1dof;
  DevBuf<double> r1w64, r1k64;
  // z extent of a structured block per precision and epilogue class (0: products and residuals,
  // 1: fused smoothing steps); 0 = heuristic
  int kz[2][2] = {{0, 0}, {0, 0}};
  // nodes the sparse pass writes (irregular cells, extra rows, rank-one terms, all split copies),
  // sorted; the sparse pass addresses them through compact indices
  int ntouch = 0;
  std::vector<int> tnode_h;
  DevBuf<int> tnode, irr_ci, xrc, r1c;
  double iso_fac = 0.0;       // level 0: h / 16 when all cells are the same cube, else 0
  // coarse levels: the extra-stiffness rows regrouped as 3x3 blocks (row node, column node)
  DevBuf<int> brow, bcol;
  std::vector<double> bval_h;
  // level 1: lam V, mu V of the level-0 children of every cell, child factor tables per axis
  DevBuf<double> cw64;
  // level 1 split: lam V, mu V of the cells whose children share one material (0 elsewhere), the
  // cells with several materials (x-fastest ids), per-axis lam second moments per coarse index
  DevBuf<double> cl64, cm64;
  DevBuf<int> het;
  int nhet = 0;
  // per node the row of its corner slots (8 * cell + corner, in slot order) or -1
  DevBuf<int> hmap, hptr, hslot;
  std::vector<double> q_h[3];
  std::vector<double> ph_h[3];
  long long nblk = 0, nrn = 0;  // block-row entries and rank-one list entries (totals)
  DevBuf<uint8_t> bflag[2];     // per precision: stiffness flags of the structured-march tiles
  DevBuf<int> border[2][2];     // per precision and epilogue class: block order of the march
                                // (~tile = no stiffness)
  int bflag_kz[2][2] = {{0, 0}, {0, 0}};   // the z extent each order was built for (0 = none)
  int blk_max = 0, rn_max = 0;  // longest block row and longest rank-one list of one node
  // level 0 only
  bool label_mode = false;
  DevBuf<uint8_t> lab;                      // apply labels (x-fastest, irregular -> 255)
  std::vector<double> tabL_h, tabM_h;       // 256
  DevBuf<double> matL, matM;                // real per-cell materials, x-fastest (setup)
  DevBuf<double> applyL, applyM;            // per-cell apply materials (value mode / diag setup)
  int ncopy = 0, nirr = 0;
  std::vector<int> copy_orig_h;
  DevBuf<int> copy_orig, irr_nodes, irr_ijk;
  // split-copy groups of the smoother (a grid node and its copies coupled by extra stiffness)
  int ngroup = 0;
  DevBuf<int> gptr, gslot, goff;
  DevBuf<uint8_t> gmem;
  DevBuf<double> goffd64, gwork64;
  std::vector<int> irr_nodes_h;
  std::vector<double> irrL_h, irrM_h;
  // coarse-level moments (FP64 master)
  DevBuf<double> mom64;
  LevelT<double> d64;
  LevelT<float> d32;
  template <typename T>
  LevelT<T>& dat();
};
template <>
LevelT<double>& Level::dat<double>() { return d64; }
template <>
LevelT<float>& Level::dat<float>() { return d32; }

struct Options {
  int smoother = 0;   // 0 chebyshev (fe/gmg.py:61-76), 1 damped block Jacobi
  int degree = 3;
  double omega = 0.6;
  int npre = 2, npost = 2;
};

template <typename T>
static void upload_cast(DevBuf<T>& dst, const std::vector<double>& src, hipStream_t s) {
  std::vector<T> tmp(src.begin(), src.end());
  dst.upload(tmp, s);
}

static inline dim3 grid1(long long n, int bs = 256) { return dim3((unsigned)div_up(n, bs)); }

// One recorded launch: the record's bytes, its block count and the launcher of the same record
// type over several systems at once (k_multi).
using MultiFn = void (*)(const void* args, const int* start, int nsys, int nblocks, hipStream_t st);
struct LaunchRec {
  MultiFn multi;
  size_t size;
  std::vector<char> bytes;
  long long nb;
};
template <class L>
static void multi_launch(const void* args, const int* start, int nsys, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_multi<L>, dim3((unsigned)nblocks), dim3(256), 0, st, static_cast<const L*>(args), start, nsys);
}

struct System {
  hipStream_t stream = nullptr;
  // launches of the operator, V-cycle and PCG steps go through go(): launched on the stream, or,
  // while rec is set, appended to it (batched solves replay the records of several systems as
  // single launches)
  std::vector<LaunchRec>* rec = nullptr;
  template <class L>
  void go(const L& a, long long nb) {
    static_assert(std::is_trivially_copyable<L>::value, "launch records are copied as bytes");
    if (rec) {
      LaunchRec r{&multi_launch<L>, sizeof(L), std::vector<char>(sizeof(L)), nb};
      std::memcpy(r.bytes.data(), &a, sizeof(L));
      rec->push_back(std::move(r));
      return;
    }
    if (nb > 0) hipLaunchKernelGGL(k_one<L>, dim3((unsigned)nb), dim3(256), 0, stream, a);
  }
  int sri = 1;
  std::vector<std::unique_ptr<Level>> L;
  Options opt;
  bool built = false;
  // PCG state
  DevBuf<double> S, partial;
  DevBuf<double> r64, z64, p64, Ap64;
  DevBuf<float> r32, z32, p32, Ap32, d32;
  // rigid-mode correction
  int zcol = 0, zchunks = 0;
  DevBuf<int> zchk, zcchk, zrows_c, zrow, zptr, zcols;   // zchk: chunk bounds in the CSC; zcchk: chunks per column
  DevBuf<double> zvals_c, zvals, Einv, zw, zv, zpart;
  // coarsest dense inverse
  DevBuf<double> Ainv;        // inverse of the coarsest operator over its used DOFs
  DevBuf<int> dmap;           // compact DOF -> full DOF of the coarsest level
  int nc_dense = 0;
  double t_setup_ms = 0.0;
  // build stage times (ms): level-0 prep, hierarchy structure, uploads, moments, diagonals,
  // coarsest inverse, total
  double t_stage[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // + structure split: used masks, extra coarsening
  // staged level-0 input (converted in prepare_level0)
  int st_nmat = -1;
  std::vector<double> st_tabL, st_tabM, st_lam, st_mu;
  std::vector<uint8_t> st_lab, st_active;
  std::vector<long long> st_copy_orig, st_irr_cell, st_irr_nodes;
  std::vector<long long> st_xr, st_xc;
  std::vector<double> st_xv;
  std::vector<uint8_t> st_used_extra;
  std::vector<long long> st_r1ptr, st_r1dof;
  std::vector<double> st_r1w, st_r1k;

  // Level-0 data in the layouts the kernels use.
  void prepare_level0() {
    Level& v = lv(0);
    const int nx = v.nx, ny = v.ny, nz = v.nz;
    const long long ncell = v.ncell;
    CM_CHECK(st_nmat >= 0 || !st_lam.empty(), "materials not set (cm_sys_set_labels or cm_sys_set_cellmat)");
    v.label_mode = st_nmat >= 0;
    if (v.label_mode) {
      v.tabL_h.assign(256, 0.0);
      v.tabM_h.assign(256, 0.0);
      for (int m = 0; m < st_nmat; ++m) { v.tabL_h[m] = st_tabL[m]; v.tabM_h[m] = st_tabM[m]; }
      for (long long c = 0; c < ncell; ++c)
        CM_CHECK(st_lab[c] == LAB_NONE || (int)st_lab[c] < st_nmat, "label outside the material table");
    }
    auto xf = [&](long long c) {
      const long long k = c % nz, j = (c / nz) % ny, i = c / ((long long)ny * nz);
      return i + (long long)nx * (j + (long long)ny * k);
    };
    auto active_ref = [&](long long c) { return v.label_mode ? st_lab[c] != LAB_NONE : st_active[c] != 0; };
    // split copies and irregular cells (host lists)
    v.ncopy = (int)st_copy_orig.size();
    v.nn = v.nng + v.ncopy;
    v.copy_orig_h.resize(v.ncopy);
    for (int q = 0; q < v.ncopy; ++q) {
      CM_CHECK(st_copy_orig[q] >= 0 && st_copy_orig[q] < v.nng, "copy origin outside the grid");
      v.copy_orig_h[q] = (int)st_copy_orig[q];
    }
    v.nirr = (int)st_irr_cell.size();
    std::vector<int> ijk(3 * (size_t)v.nirr), nodes(8 * (size_t)v.nirr), irr_x(v.nirr);
    v.irrL_h.resize(v.nirr); v.irrM_h.resize(v.nirr);
    for (int e = 0; e < v.nirr; ++e) {
      const long long c = st_irr_cell[e];
      CM_CHECK(c >= 0 && c < ncell, "irregular cell id out of range");
      CM_CHECK(active_ref(c), "irregular cell is not active");
      irr_x[e] = (int)xf(c);
      ijk[3 * e] = (int)(c / ((long long)ny * nz)); ijk[3 * e + 1] = (int)((c / nz) % ny); ijk[3 * e + 2] = (int)(c % nz);
      v.irrL_h[e] = v.label_mode ? st_tabL[st_lab[c]] : st_lam[c];
      v.irrM_h[e] = v.label_mode ? st_tabM[st_lab[c]] : st_mu[c];
      for (int l = 0; l < 8; ++l) {
        const long long nd = st_irr_nodes[8 * (size_t)e + l];
        CM_CHECK(nd >= 0 && nd < v.nn, "irregular node id out of range");
        nodes[8 * (size_t)e + l] = (int)nd;
      }
    }
    // cell arrays on the GPU: x-fastest real labels/materials, apply arrays without the
    // irregular cells, used nodes (corners of active regular cells + irregular cells' nodes)
    {
      DevBuf<uint8_t> lab_ref, act_ref, actx, irr;
      DevBuf<double> tl, tm, lam_ref, mu_ref;
      v.matL.alloc(ncell); v.matM.alloc(ncell);
      actx.alloc(ncell);
      if (v.label_mode) {
        lab_ref.upload(st_lab, stream);
        tl.upload(v.tabL_h, stream); tm.upload(v.tabM_h, stream);
        v.lab.alloc(ncell);
      } else {
        lam_ref.upload(st_lam, stream); mu_ref.upload(st_mu, stream); act_ref.upload(st_active, stream);
      }
      hipLaunchKernelGGL(k_prep_cells, grid1(ncell), dim3(256), 0, stream, nx, ny, nz, lab_ref.p, tl.p, tm.p,
                         lam_ref.p, mu_ref.p, act_ref.p, v.lab.p, v.matL.p, v.matM.p, actx.p);
      CM_LAUNCH_CHECK();
      v.applyL.alloc(ncell); v.applyM.alloc(ncell);
      CM_HIP(hipMemcpyAsync(v.applyL.p, v.matL.p, ncell * sizeof(double), hipMemcpyDeviceToDevice, stream));
      CM_HIP(hipMemcpyAsync(v.applyM.p, v.matM.p, ncell * sizeof(double), hipMemcpyDeviceToDevice, stream));
      irr.alloc(ncell);
      hipLaunchKernelGGL(k_fill<uint8_t>, grid1(ncell), dim3(256), 0, stream, irr.p, ncell, uint8_t(0));
      if (v.nirr) {
        DevBuf<int> ix;
        ix.upload(irr_x, stream);
        hipLaunchKernelGGL(k_prep_irregular, grid1(v.nirr), dim3(256), 0, stream, v.nirr, ix.p,
                           v.label_mode ? v.lab.p : nullptr, v.applyL.p, v.applyM.p, irr.p);
        CM_LAUNCH_CHECK();
        CM_HIP(hipStreamSynchronize(stream));
      }
      DevBuf<uint8_t> used(v.nn);
      if (v.nn > v.nng)
        hipLaunchKernelGGL(k_fill<uint8_t>, grid1(v.nn - v.nng), dim3(256), 0, stream, used.p + v.nng, v.nn - v.nng,
                           uint8_t(0));
      hipLaunchKernelGGL(k_used_grid, grid1(v.nng), dim3(256), 0, stream, nx, ny, nz, actx.p, irr.p, used.p);
      v.irr_nodes.upload(nodes, stream);
      if (!nodes.empty())
        hipLaunchKernelGGL(k_mark_nodes, grid1((long long)nodes.size()), dim3(256), 0, stream, (long long)nodes.size(),
                           v.irr_nodes.p, used.p);
      CM_LAUNCH_CHECK();
      v.used_h.resize(v.nn);
      CM_HIP(hipMemcpyAsync(v.used_h.data(), used.p, v.nn, hipMemcpyDeviceToHost, stream));
      CM_HIP(hipStreamSynchronize(stream));
    }
    if (!st_used_extra.empty()) {
      CM_CHECK((long long)st_used_extra.size() == v.nn, "used mask length must equal grid nodes + copies");
      for (long long q = 0; q < v.nn; ++q) v.used_h[q] |= st_used_extra[q];
    }
    v.irr_ijk.upload(ijk, stream);
    v.irr_nodes_h.swap(nodes);
    v.copy_orig.upload(v.copy_orig_h, stream);
    // extra stiffness COO -> unique-row CSR (duplicates summed)
    const long long nd3 = 3 * v.nn;
    std::vector<std::pair<long long, double>> ent;
    ent.reserve(st_xr.size());
    for (size_t q = 0; q < st_xr.size(); ++q) {
      CM_CHECK(st_xr[q] >= 0 && st_xr[q] < nd3 && st_xc[q] >= 0 && st_xc[q] < nd3, "extra entry out of range");
      ent.emplace_back(st_xr[q] * nd3 + st_xc[q], st_xv[q]);
    }
    std::sort(ent.begin(), ent.end(), [](auto& a, auto& b) { return a.first < b.first; });
    std::vector<std::pair<long long, double>> merged;
    for (auto& e : ent) {
      if (!merged.empty() && merged.back().first == e.first) merged.back().second += e.second;
      else merged.push_back(e);
    }
    // diagonal entries are rank-one terms k e_r e_r^T (cheap exact coarsening); off-diagonal
    // entries form the generic part
    v.tptr_h.assign(1, 0); v.tdof_h.clear(); v.tw_h.clear(); v.tk_h.clear(); v.gen_h.clear();
    for (auto& e : merged) {
      const long long R = e.first / nd3, C = e.first % nd3;
      if (R == C) {
        if (e.second != 0.0) {
          v.tdof_h.push_back((int)R); v.tw_h.push_back(1.0); v.tk_h.push_back(e.second);
          v.tptr_h.push_back((int)v.tdof_h.size());
        }
      } else {
        v.gen_h.emplace_back(e.first, e.second, std::fabs(e.second));
      }
    }
    const size_t nt_diag = v.tk_h.size();
    // user rank-one terms k w w^T (DOFs sorted and merged per term): coarsened as terms and, on
    // level 0, applied in rank-one form k (w . u) w -- for a stiff spring between two nodes
    // that move together this forms u_a - u_b before scaling by k, where the expanded rows
    // k u_a - k u_b would cancel (the residual of a fused joint is only accurate this way)
    v.r1ptr_h.assign(1, 0); v.r1dof_h.clear(); v.r1w_h.clear(); v.r1k_h.clear();
    const long long nt = st_r1ptr.empty() ? 0 : (long long)st_r1ptr.size() - 1;
    std::vector<std::pair<int, double>> tw;
    for (long long t = 0; t < nt; ++t) {
      tw.clear();
      for (long long q = st_r1ptr[t]; q < st_r1ptr[t + 1]; ++q) {
        CM_CHECK(st_r1dof[q] >= 0 && st_r1dof[q] < nd3, "rank-one entry out of range");
        tw.emplace_back((int)st_r1dof[q], st_r1w[q]);
      }
      std::sort(tw.begin(), tw.end(), [](auto& a, auto& b) { return a.first < b.first; });
      const size_t before = v.r1dof_h.size();
      for (size_t q = 0; q < tw.size();) {
        const int d = tw[q].first;
        double w = 0.0;
        for (; q < tw.size() && tw[q].first == d; ++q) w += tw[q].second;
        if (w != 0.0) { v.r1dof_h.push_back(d); v.r1w_h.push_back(w); }
      }
      if (v.r1dof_h.size() == before || st_r1k[t] == 0.0) {
        v.r1dof_h.resize(before); v.r1w_h.resize(before);
        continue;
      }
      v.r1k_h.push_back(st_r1k[t]);
      v.r1ptr_h.push_back((int)v.r1dof_h.size());
      for (size_t q = before; q < v.r1dof_h.size(); ++q) { v.tdof_h.push_back(v.r1dof_h[q]); v.tw_h.push_back(v.r1w_h[q]); }
      v.tk_h.push_back(st_r1k[t]);
      v.tptr_h.push_back((int)v.tdof_h.size());
    }
    // level-0 CSR: generic entries + the diagonal springs
    std::vector<std::tuple<long long, double, double>> ent0(v.gen_h);
    for (size_t t = 0; t < nt_diag; ++t) {
      const int a = v.tptr_h[t];
      ent0.emplace_back((long long)v.tdof_h[a] * nd3 + v.tdof_h[a], v.tk_h[t], std::fabs(v.tk_h[t]));
    }
    set_extra_sorted(v, ent0, nd3);
  }

  Level& lv(int l) { return *L[l]; }
  int nlev() const { return (int)L.size(); }

  // ------------------------------------------------------------------ construction helpers
  // keep = this axis is not coarsened (every fine node is a coarse node; identity transfer)
  static void axis_transfer(const std::vector<double>& ef, Level& fine, Level& coarse, int ax, bool keep) {
    const int nfn = (int)ef.size();          // fine nodes
    std::vector<int> cidx;
    for (int i = 0; i < nfn; i += (keep ? 1 : 2)) cidx.push_back(i);
    if (cidx.back() != nfn - 1) cidx.push_back(nfn - 1);
    const int ncn = (int)cidx.size();
    std::vector<double> ec(ncn);
    for (int I = 0; I < ncn; ++I) ec[I] = ef[cidx[I]];
    coarse.e[ax] = ec;
    // prolongation (fe/gmg.py:7-26)
    std::vector<int> ia(nfn);
    std::vector<double> w(nfn, 0.0);
    for (int i = 0, a = 0; i < nfn; ++i) {
      while (a + 1 < ncn && cidx[a + 1] <= i) ++a;
      ia[i] = a;
      if (cidx[a] != i) w[i] = (ef[i] - ef[cidx[a]]) / (ef[cidx[a + 1]] - ef[cidx[a]]);
    }
    // restriction: coarse I gathers fine cidx[I]-1 .. cidx[I]+1
    std::vector<int> rs(ncn);
    std::vector<double> rw(3 * ncn, 0.0);
    for (int I = 0; I < ncn; ++I) {
      const int c = cidx[I];
      rs[I] = c - 1;
      rw[3 * I + 1] = 1.0;
      if (I > 0 && c - 1 > cidx[I - 1]) rw[3 * I] = w[c - 1];
      if (I + 1 < ncn && c + 1 < cidx[I + 1]) rw[3 * I + 2] = 1.0 - w[c + 1];
    }
    fine.pia_h[ax] = ia; fine.pw_h[ax] = w; fine.rs_h[ax] = rs; fine.rw_h[ax] = rw;
    // nesting of fine cells in coarse cells
    const int ncc = ncn - 1, nfc = nfn - 1;
    std::vector<int> cs(ncc), cn(ncc);
    std::vector<double> na(nfc), nb(nfc);
    for (int C = 0; C < ncc; ++C) {
      cs[C] = cidx[C];
      cn[C] = cidx[C + 1] - cidx[C];
      const double X0 = ec[C], X1 = ec[C + 1], H = X1 - X0;
      for (int i = cidx[C]; i < cidx[C + 1]; ++i) {
        na[i] = (ef[i + 1] - ef[i]) / H;
        nb[i] = (ef[i] + ef[i + 1] - X0 - X1) / H;
      }
    }
    coarse.ncs_h[ax] = cs; coarse.ncn_h[ax] = cn; coarse.na_h[ax] = na; coarse.nb_h[ax] = nb;
  }

  // Kept coarse nodes: coarse nodes with a nonzero prolongation weight from a used fine node
  // (fe/gmg.py:94-96, keep = columns of P[act] with nonzero sum).
  static void coarse_used(Level& f, Level& c) {
    const int fx = f.nx + 1, fy = f.ny + 1, fz = f.nz + 1;
    const int cx = c.nx + 1, cy = c.ny + 1, cz = c.nz + 1;
    c.used_h.assign((size_t)cx * cy * cz, 0);
    // fold copies onto their originals: a copy is interpolated like its original
    std::vector<uint8_t> fu(f.used_h.begin(), f.used_h.begin() + f.nng);
    for (long long q = 0; q < (long long)f.nn - f.nng; ++q)
      if (f.used_h[f.nng + q]) fu[f.copy_orig_h[q]] = 1;
    auto sup = [](const std::vector<int>& ia, const std::vector<double>& w, int i, int* out) {
      int n = 0;
      out[n++] = ia[i];
      if (w[i] != 0.0) out[n++] = ia[i] + 1;
      return n;
    };
    for (int k = 0; k < fz; ++k)
      for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i) {
          if (!fu[(size_t)i + (size_t)fx * (j + (size_t)fy * k)]) continue;
          int si[2], sj[2], sk[2];
          const int ni = sup(f.pia_h[0], f.pw_h[0], i, si);
          const int nj = sup(f.pia_h[1], f.pw_h[1], j, sj);
          const int nk = sup(f.pia_h[2], f.pw_h[2], k, sk);
          for (int c3 = 0; c3 < nk; ++c3)
            for (int b = 0; b < nj; ++b)
              for (int a = 0; a < ni; ++a) c.used_h[(size_t)si[a] + (size_t)cx * (sj[b] + (size_t)cy * sk[c3])] = 1;
        }
  }

  // Coarse support of a fine node (copies interpolate like their originals): <= 8 coarse nodes.
  static int node_support(const Level& f, const Level& c, int fnode, long long* nodes, double* w) {
    const int fx = f.nx + 1, fy = f.ny + 1;
    const int cx = c.nx + 1, cy = c.ny + 1;
    int g = fnode;
    if (g >= f.nng) g = f.copy_orig_h[g - f.nng];
    const int i = g % fx, j = (g / fx) % fy, k = g / (fx * fy);
    int n = 0;
    for (int dz = 0; dz < 2; ++dz) {
      const double wz = dz ? f.pw_h[2][k] : 1.0 - f.pw_h[2][k];
      if (wz == 0.0) continue;
      for (int dy = 0; dy < 2; ++dy) {
        const double wy = dy ? f.pw_h[1][j] : 1.0 - f.pw_h[1][j];
        if (wy == 0.0) continue;
        for (int dx = 0; dx < 2; ++dx) {
          const double wx = dx ? f.pw_h[0][i] : 1.0 - f.pw_h[0][i];
          if (wx == 0.0) continue;
          nodes[n] = (f.pia_h[0][i] + dx) + (long long)cx * ((f.pia_h[1][j] + dy) + (long long)cy * (f.pia_h[2][k] + dz));
          w[n] = wx * wy * wz;
          ++n;
        }
      }
    }
    return n;
  }

  // Galerkin coarsening of the extra stiffness.  Generic part: P^T S P by sorted accumulation (an
  // entry whose contributions cancel to 1e-12 of their magnitude is dropped).  Rank-one terms:
  // w -> P^T w (entries cancelling to 1e-12 dropped; node-pair springs between a copy and its
  // original vanish exactly).  For applying, terms whose support fits one 3x3x3 node
  // neighbourhood are summed into 27-point node blocks and emitted as CSR rows; wider terms
  // (ligaments) stay rank-one.
  static void coarse_extra(Level& f, Level& c) {
    const int cx = c.nx + 1, cy = c.ny + 1;
    const long long ncd = 3LL * (long long)cx * cy * (c.nz + 1);
    c.gen_h.clear();
    for (auto& e : f.gen_h) {
      const long long key = std::get<0>(e);
      const long long row = key / (3 * f.nn), col = key % (3 * f.nn);
      long long rn[8], cnod[8];
      double rw[8], cw[8];
      const int nr = node_support(f, c, (int)(row / 3), rn, rw);
      const int ncs = node_support(f, c, (int)(col / 3), cnod, cw);
      for (int a2 = 0; a2 < nr; ++a2)
        for (int b2 = 0; b2 < ncs; ++b2) {
          const double val = rw[a2] * cw[b2] * std::get<1>(e);
          c.gen_h.emplace_back((3 * rn[a2] + row % 3) * ncd + 3 * cnod[b2] + col % 3, val, std::fabs(val));
        }
    }
    if (!c.gen_h.empty()) {
      // merge duplicates now so that the next level starts from unique entries
      std::vector<std::tuple<long long, double, double>> m;
      std::sort(c.gen_h.begin(), c.gen_h.end(), [](auto& x, auto& y) { return std::get<0>(x) < std::get<0>(y); });
      for (size_t q = 0; q < c.gen_h.size();) {
        const long long key = std::get<0>(c.gen_h[q]);
        double sv = 0.0, sa = 0.0;
        while (q < c.gen_h.size() && std::get<0>(c.gen_h[q]) == key) { sv += std::get<1>(c.gen_h[q]); sa += std::get<2>(c.gen_h[q]); ++q; }
        if (std::fabs(sv) > 1e-12 * sa) m.emplace_back(key, sv, std::fabs(sv));
      }
      c.gen_h.swap(m);
    }
    // coarsen the rank-one terms
    c.tptr_h.assign(1, 0); c.tdof_h.clear(); c.tw_h.clear(); c.tk_h.clear();
    std::vector<std::tuple<int, double, double>> tw;
    for (size_t t = 0; t + 1 < f.tptr_h.size(); ++t) {
      tw.clear();
      for (int q = f.tptr_h[t]; q < f.tptr_h[t + 1]; ++q) {
        const int d = f.tdof_h[q];
        long long cn[8];
        double cw[8];
        const int ns = node_support(f, c, d / 3, cn, cw);
        for (int a2 = 0; a2 < ns; ++a2) {
          const double val = cw[a2] * f.tw_h[q];
          tw.emplace_back((int)(3 * cn[a2] + d % 3), val, std::fabs(val));
        }
      }
      std::sort(tw.begin(), tw.end(), [](auto& x, auto& y) { return std::get<0>(x) < std::get<0>(y); });
      const size_t before = c.tdof_h.size();
      for (size_t q = 0; q < tw.size();) {
        const int d = std::get<0>(tw[q]);
        double sv = 0.0, sa = 0.0;
        while (q < tw.size() && std::get<0>(tw[q]) == d) { sv += std::get<1>(tw[q]); sa += std::get<2>(tw[q]); ++q; }
        if (std::fabs(sv) > 1e-12 * sa) { c.tdof_h.push_back(d); c.tw_h.push_back(sv); }
      }
      if (c.tdof_h.size() > before) {
        c.tk_h.push_back(f.tk_h[t]);
        c.tptr_h.push_back((int)c.tdof_h.size());
      }
    }
    // apply structures: local terms -> 27-point node blocks -> CSR; wide terms -> rank-one
    c.r1ptr_h.assign(1, 0); c.r1dof_h.clear(); c.r1w_h.clear(); c.r1k_h.clear();
    std::unordered_map<long long, int> slot;
    std::vector<double> blk;          // per slot: [3 comps][27 offsets][3 comps]
    std::vector<long long> slot_node;
    slot.reserve(2 * c.tdof_h.size());
    slot_node.reserve(c.tdof_h.size());
    auto ijk = [&](long long nd, int& i, int& j, int& k) {
      i = (int)(nd % cx); j = (int)((nd / cx) % cy); k = (int)(nd / ((long long)cx * cy));
    };
    for (size_t t = 0; t + 1 < c.tptr_h.size(); ++t) {
      const int q0 = c.tptr_h[t], q1 = c.tptr_h[t + 1];
      int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
      for (int q = q0; q < q1; ++q) {
        int p3[3];
        ijk(c.tdof_h[q] / 3, p3[0], p3[1], p3[2]);
        for (int a2 = 0; a2 < 3; ++a2) { lo[a2] = std::min(lo[a2], p3[a2]); hi[a2] = std::max(hi[a2], p3[a2]); }
      }
      const bool local = (hi[0] - lo[0] <= 1) && (hi[1] - lo[1] <= 1) && (hi[2] - lo[2] <= 1);
      if (!local) {
        for (int q = q0; q < q1; ++q) { c.r1dof_h.push_back(c.tdof_h[q]); c.r1w_h.push_back(c.tw_h[q]); }
        c.r1k_h.push_back(c.tk_h[t]);
        c.r1ptr_h.push_back((int)c.r1dof_h.size());
        continue;
      }
      for (int qa = q0; qa < q1; ++qa) {
        const long long na = c.tdof_h[qa] / 3;
        auto it = slot.find(na);
        int sl;
        if (it == slot.end()) {
          sl = (int)slot_node.size();
          slot.emplace(na, sl);
          slot_node.push_back(na);
          blk.resize(blk.size() + 243, 0.0);
        } else {
          sl = it->second;
        }
        int ia, ja, ka;
        ijk(na, ia, ja, ka);
        for (int qb = q0; qb < q1; ++qb) {
          int ib, jb, kb;
          ijk(c.tdof_h[qb] / 3, ib, jb, kb);
          const int off = (ib - ia + 1) + 3 * (jb - ja + 1) + 9 * (kb - ka + 1);
          blk[(size_t)sl * 243 + (c.tdof_h[qa] % 3) * 81 + off * 3 + c.tdof_h[qb] % 3] +=
              c.tk_h[t] * c.tw_h[qa] * c.tw_h[qb];
        }
      }
    }
    std::vector<std::tuple<long long, double, double>> ent(c.gen_h);
    std::vector<int> order(slot_node.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return slot_node[x] < slot_node[y]; });
    if (ent.empty()) {
      // rows straight from the node blocks, columns in increasing order (offsets z, y, x)
      c.xrow_h.clear(); c.xptr_h.clear(); c.xcol_h.clear(); c.xval_h.clear();
      for (int sl : order) {
        const long long na = slot_node[sl];
        int ia, ja, ka;
        ijk(na, ia, ja, ka);
        for (int ca = 0; ca < 3; ++ca) {
          const size_t start = c.xcol_h.size();
          for (int off = 0; off < 27; ++off) {
            const int ib = ia + off % 3 - 1, jb = ja + (off / 3) % 3 - 1, kb = ka + off / 9 - 1;
            for (int cb = 0; cb < 3; ++cb) {
              const double v2 = blk[(size_t)sl * 243 + ca * 81 + off * 3 + cb];
              if (v2 == 0.0) continue;
              c.xcol_h.push_back((int)(3 * ((long long)ib + (long long)cx * (jb + (long long)cy * kb)) + cb));
              c.xval_h.push_back(v2);
            }
          }
          if (c.xcol_h.size() > start) {
            c.xrow_h.push_back((int)(3 * na + ca));
            c.xptr_h.push_back((int)start);
          }
        }
      }
      c.xptr_h.push_back((int)c.xcol_h.size());
      const size_t nr = c.xrow_h.size();
      const double avg = nr ? (double)c.xcol_h.size() / nr : 1.0;
      int w = 1;
      while (w < 64 && w < avg) w <<= 1;
      c.xw = w;
    } else {
      for (int sl : order) {
        const long long na = slot_node[sl];
        int ia, ja, ka;
        ijk(na, ia, ja, ka);
        for (int ca = 0; ca < 3; ++ca)
          for (int off = 0; off < 27; ++off)
            for (int cb = 0; cb < 3; ++cb) {
              const double v2 = blk[(size_t)sl * 243 + ca * 81 + off * 3 + cb];
              if (v2 == 0.0) continue;
              const long long nb2 = (ia + off % 3 - 1) + (long long)cx * ((ja + (off / 3) % 3 - 1) + (long long)cy * (ka + off / 9 - 1));
              ent.emplace_back((3 * na + ca) * ncd + 3 * nb2 + cb, v2, std::fabs(v2));
            }
      }
      set_extra_sorted(c, ent, ncd);
    }
  }

  static void set_extra_sorted(Level& lv, std::vector<std::tuple<long long, double, double>>& ent, long long ncd) {
    lv.xrow_h.clear(); lv.xptr_h.clear(); lv.xcol_h.clear(); lv.xval_h.clear();
    std::sort(ent.begin(), ent.end(), [](auto& x, auto& y) { return std::get<0>(x) < std::get<0>(y); });
    long long last = -1;
    for (size_t q = 0; q < ent.size();) {
      const long long key = std::get<0>(ent[q]);
      double sv = 0.0, sa = 0.0;
      while (q < ent.size() && std::get<0>(ent[q]) == key) { sv += std::get<1>(ent[q]); sa += std::get<2>(ent[q]); ++q; }
      if (!(std::fabs(sv) > 1e-12 * sa)) continue;     // cancelled (or zero)
      const long long R = key / ncd, C = key % ncd;
      if (R != last) {
        lv.xrow_h.push_back((int)R);
        lv.xptr_h.push_back((int)lv.xcol_h.size());
        last = R;
      }
      lv.xcol_h.push_back((int)C);
      lv.xval_h.push_back(sv);
    }
    lv.xptr_h.push_back((int)lv.xcol_h.size());
    const size_t nr = lv.xrow_h.size();
    const double avg = nr ? (double)lv.xcol_h.size() / nr : 1.0;
    int w = 1;
    while (w < 64 && w < avg) w <<= 1;
    lv.xw = w;
  }

  static long long compact_dofs_of(const Level& v) {
    long long n = 0;
    for (auto u : v.used_h) n += u;
    return 3 * n;
  }
  long long compact_dofs(int l) { return compact_dofs_of(lv(l)); }
  int last_axes = 3;   // axes the step to the coarsest level coarsens (7 = all, the reference rule)
  bool copy_blocks = true;   // level-0 smoother blocks over split-copy groups (false: 3x3 per node)

  // ------------------------------------------------------------------ build
  // h / 16 when every cell of the level is the same cube (edge lengths equal to 1e-13 relative),
  // else 0.  The structured march then folds the gradient factors into the material weights.
  static double iso_factor(const Level& v) {
    const double h = v.e[0][1] - v.e[0][0];
    if (!(h > 0.0)) return 0.0;
    for (int ax = 0; ax < 3; ++ax)
      for (size_t i = 0; i + 1 < v.e[ax].size(); ++i)
        if (std::fabs((v.e[ax][i + 1] - v.e[ax][i]) - h) > 1e-13 * h) return 0.0;
    return h / 16.0;
  }

  // Compact node list of the sparse pass: every node an irregular cell, an extra row or a
  // rank-one term writes, plus every split copy (copies have no structured part).  Uploads the
  // node list and the compact indices of the sparse pass's outputs.
  static int compact_of(const std::vector<int>& sorted, long long node) {
    const auto it = std::lower_bound(sorted.begin(), sorted.end(), (int)node);
    return (int)(it - sorted.begin());
  }
  // Coarse levels: the extra-stiffness CSR (rows sorted and unique) regrouped as 3x3 blocks in
  // row-node order; k_coarse_extra applies them (with the rank-one terms) before the march, so
  // coarse levels need no sparse pass.  The per-node counts are kept for level_stats.
  void build_coarse_rows(Level& v) {
    const long long nn = v.nn;
    v.ntouch = 0;
    if (!v.xrow_h.empty()) {
      std::vector<int> bptr(nn + 1, 0), bcol, brow;
      v.bval_h.clear();
      std::map<int, std::array<double, 9>> blocks;
      const size_t nrow = v.xrow_h.size();
      for (size_t r = 0; r < nrow;) {
        const int node = v.xrow_h[r] / 3;
        blocks.clear();
        for (; r < nrow && v.xrow_h[r] / 3 == node; ++r) {
          const int comp = v.xrow_h[r] % 3;
          for (int q = v.xptr_h[r]; q < v.xptr_h[r + 1]; ++q) blocks[v.xcol_h[q] / 3][3 * comp + v.xcol_h[q] % 3] += v.xval_h[q];
        }
        bptr[node + 1] = (int)blocks.size();
        for (auto& kv : blocks) {
          brow.push_back(node);
          bcol.push_back(kv.first);
          v.bval_h.insert(v.bval_h.end(), kv.second.begin(), kv.second.end());
        }
      }
      for (long long n = 0; n < nn; ++n) v.blk_max = std::max(v.blk_max, bptr[n + 1]);
      for (long long n = 0; n < nn; ++n) bptr[n + 1] += bptr[n];
      v.nblk = bptr[nn];
      v.brow.upload(brow, stream);
      v.bcol.upload(bcol, stream);
    }
    if (!v.r1k_h.empty()) {
      // per node: the number of rank-one terms whose support includes it (entries of a term are
      // sorted by DOF, so a node's entries are adjacent)
      std::vector<int> cnt(nn, 0);
      v.nrn = 0;
      for (size_t t = 0; t + 1 < v.r1ptr_h.size(); ++t) {
        int last = -1;
        for (int q = v.r1ptr_h[t]; q < v.r1ptr_h[t + 1]; ++q) {
          const int node = v.r1dof_h[q] / 3;
          if (node != last) { ++cnt[node]; ++v.nrn; last = node; }
        }
      }
      for (long long n = 0; n < nn; ++n) v.rn_max = std::max(v.rn_max, cnt[n]);
    }
  }

  void build_touch(Level& v) {
    std::vector<int> nodes;
    nodes.reserve(v.irr_nodes_h.size() + v.xrow_h.size() + v.r1dof_h.size() + (size_t)(v.nn - v.nng));
    for (int nd : v.irr_nodes_h) nodes.push_back(nd);
    for (int r : v.xrow_h) nodes.push_back(r / 3);
    for (int q : v.r1dof_h) nodes.push_back(q / 3);
    for (long long n = v.nng; n < v.nn; ++n) nodes.push_back((int)n);
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    v.tnode_h.swap(nodes);
    v.ntouch = (int)v.tnode_h.size();
    std::vector<int> ci(v.irr_nodes_h.size()), xr(v.xrow_h.size()), rc(v.r1dof_h.size());
    for (size_t q = 0; q < ci.size(); ++q) ci[q] = compact_of(v.tnode_h, v.irr_nodes_h[q]);
    for (size_t q = 0; q < xr.size(); ++q) xr[q] = 3 * compact_of(v.tnode_h, v.xrow_h[q] / 3) + v.xrow_h[q] % 3;
    for (size_t q = 0; q < rc.size(); ++q) rc[q] = 3 * compact_of(v.tnode_h, v.r1dof_h[q] / 3) + v.r1dof_h[q] % 3;
    v.tnode.upload(v.tnode_h, stream);
    v.irr_ci.upload(ci, stream);
    v.xrc.upload(xr, stream);
    v.r1c.upload(rc, stream);
  }

  // Split-copy groups of level 0: a grid node and its split copies, all in the touched list, with
  // extra stiffness between them (interface springs).  Keeps the members (as touched slots) and the
  // off-diagonal member blocks of A, dense per group; k_group_inverse adds the diagonal blocks.
  void build_groups(Level& v) {
    v.ngroup = 0;
    if (!copy_blocks || v.nn == v.nng || v.ntouch == 0) return;
    auto touched = [&](long long n) { return std::binary_search(v.tnode_h.begin(), v.tnode_h.end(), (int)n); };
    std::map<int, std::vector<int>> by_orig;
    for (long long c = v.nng; c < v.nn; ++c) by_orig[v.copy_orig_h[c - v.nng]].push_back((int)c);
    std::vector<std::vector<int>> mem;
    std::vector<int> gid(v.nn, -1), gix(v.nn, -1);
    for (auto& kv : by_orig) {
      std::vector<int> m;
      if (touched(kv.first)) m.push_back(kv.first);
      for (int c : kv.second)
        if (touched(c)) m.push_back(c);
      if (m.size() < 2 || m.size() > (size_t)GROUP_MAX) continue;
      for (size_t j = 0; j < m.size(); ++j) { gid[m[j]] = (int)mem.size(); gix[m[j]] = (int)j; }
      mem.push_back(std::move(m));
    }
    if (mem.empty()) return;
    std::vector<std::vector<double>> off(mem.size());
    for (size_t g = 0; g < mem.size(); ++g) off[g].assign(9 * mem[g].size() * mem[g].size(), 0.0);
    auto add = [&](int da, int db, double val) {
      const int a = da / 3, b = db / 3;
      if (a == b || gid[a] < 0 || gid[a] != gid[b]) return;
      const int N = 3 * (int)mem[gid[a]].size();
      off[gid[a]][(size_t)(3 * gix[a] + da % 3) * N + 3 * gix[b] + db % 3] += val;
    };
    for (size_t r = 0; r < v.xrow_h.size(); ++r)
      for (int q = v.xptr_h[r]; q < v.xptr_h[r + 1]; ++q) add(v.xrow_h[r], v.xcol_h[q], v.xval_h[q]);
    for (size_t k = 0; k + 1 < v.r1ptr_h.size(); ++k)
      for (int p = v.r1ptr_h[k]; p < v.r1ptr_h[k + 1]; ++p)
        for (int q = v.r1ptr_h[k]; q < v.r1ptr_h[k + 1]; ++q)
          add(v.r1dof_h[p], v.r1dof_h[q], v.r1k_h[k] * v.r1w_h[p] * v.r1w_h[q]);
    std::vector<int> gptr{0}, gslot, goff{0};
    std::vector<double> offd;
    std::vector<uint8_t> gmem(v.ntouch, 0);
    for (size_t g = 0; g < mem.size(); ++g) {
      bool coupled = false;
      for (double x : off[g]) coupled |= (x != 0.0);
      if (!coupled) continue;
      for (int n : mem[g]) {
        const int s = compact_of(v.tnode_h, n);
        gslot.push_back(s);
        gmem[s] = 1;
      }
      gptr.push_back((int)gslot.size());
      offd.insert(offd.end(), off[g].begin(), off[g].end());
      goff.push_back((int)offd.size());
    }
    v.ngroup = (int)gptr.size() - 1;
    if (!v.ngroup) return;
    v.gptr.upload(gptr, stream);
    v.gslot.upload(gslot, stream);
    v.goff.upload(goff, stream);
    v.gmem.upload(gmem, stream);
    v.goffd64.upload(offd, stream);
    v.gwork64.alloc(offd.size());
  }

  template <typename T>
  CopyGroups<T> groups_of(int l) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    CopyGroups<T> ga{};
    if (l != 0 || !v.ngroup || !d.ginv.n) return ga;
    ga.ng = v.ngroup;
    ga.gptr = v.gptr.p; ga.gslot = v.gslot.p; ga.tnode = v.tnode.p; ga.gmem = v.gmem.p;
    ga.goff = v.goff.p; ga.ginv = d.ginv.p;
    return ga;
  }

  void build(int min_dofs, int max_levels, int fixed_levels, const Options& o) {
    CM_CHECK(!L.empty(), "system has no fine level");
    opt = o;
    hipEvent_t e0, e1;
    CM_HIP(hipEventCreate(&e0)); CM_HIP(hipEventCreate(&e1));
    CM_HIP(hipEventRecord(e0, stream));
    CM_CHECK(!built, "system already built");
    using clk = std::chrono::steady_clock;
    auto tick = clk::now();
    auto lap = [&](int k) {
      CM_HIP(hipStreamSynchronize(stream));
      const auto now = clk::now();
      t_stage[k] = std::chrono::duration<double, std::milli>(now - tick).count();
      tick = now;
    };
    const auto t_begin = clk::now();
    prepare_level0();
    lap(0);
    Level& f0 = lv(0);
    f0.used.upload(f0.used_h, stream);
    f0.iso_fac = iso_factor(f0);
    // level structure (fe/gmg.py:86-101, rb/gpusolve.py:44-46)
    const int nl_max = fixed_levels > 0 ? fixed_levels : max_levels;
    while (true) {
      const int l = nlev() - 1;
      const bool last = fixed_levels > 0 ? (l == fixed_levels - 1) : (compact_dofs(l) <= min_dofs || l == nl_max - 1);
      if (last) break;
      Level& f = lv(l);
      auto coarsen = [&](int axes) {
        auto c = std::make_unique<Level>();
        for (int ax = 0; ax < 3; ++ax) axis_transfer(f.e[ax], f, *c, ax, !((axes >> ax) & 1));
        c->nx = (int)c->e[0].size() - 1; c->ny = (int)c->e[1].size() - 1; c->nz = (int)c->e[2].size() - 1;
        c->nng = c->nn = (long long)(c->nx + 1) * (c->ny + 1) * (c->nz + 1);
        c->ncell = (long long)c->nx * c->ny * c->nz;
        auto ta = clk::now();
        coarse_used(f, *c);
        t_stage[7] += std::chrono::duration<double, std::milli>(clk::now() - ta).count();
        return c;
      };
      auto c = coarsen(7);
      // The level just built is the coarsest (dense direct solve): coarsen it along last_axes only
      // when that level still has <= min_dofs DOFs.  Keeping the vertical axis (the direction of
      // the ground springs and of the load) keeps the bones stacked along it apart on the coarsest
      // grid.  (Rebuilding restores the fine level's transfer tables for the variant kept.)
      if (fixed_levels <= 0 && last_axes != 7 && compact_dofs_of(*c) <= min_dofs) {
        auto cs = coarsen(last_axes);
        if (compact_dofs_of(*cs) <= min_dofs) c = std::move(cs);
        else c = coarsen(7);
      }
      // the springs are coarsened once, through the transfer tables of the variant kept
      const auto tb = clk::now();
      coarse_extra(f, *c);
      t_stage[8] += std::chrono::duration<double, std::milli>(clk::now() - tb).count();
      L.push_back(std::move(c));
    }
    lap(1);
    // device transfer tables and per-level constant data
    for (int l = 0; l < nlev(); ++l) {
      Level& v = lv(l);
      if (l > 0) v.used.upload(v.used_h, stream);
      if (l + 1 < nlev()) {
        for (int ax = 0; ax < 3; ++ax) {
          v.pia[ax].upload(v.pia_h[ax], stream);
          v.rs[ax].upload(v.rs_h[ax], stream);
        }
      }
      if (l > 0) {
        for (int ax = 0; ax < 3; ++ax) {
          v.ncs[ax].upload(v.ncs_h[ax], stream);
          v.ncn[ax].upload(v.ncn_h[ax], stream);
          v.na[ax].upload(v.na_h[ax], stream);
          v.nb[ax].upload(v.nb_h[ax], stream);
        }
      }
      v.xrow.upload(v.xrow_h, stream);
      v.xptr.upload(v.xptr_h, stream);
      v.xcol.upload(v.xcol_h, stream);
      v.r1ptr.upload(v.r1ptr_h, stream);
      v.r1dof.upload(v.r1dof_h, stream);
      v.r1w64.upload(v.r1w_h, stream);
      v.r1k64.upload(v.r1k_h, stream);
      if (l == 0) {
        build_touch(v);
        build_groups(v);
      } else {
        build_coarse_rows(v);
      }
    }
    prepare_prec<double>(0, false);
    lap(2);
    // moments of coarse cells (FP64)
    for (int l = 1; l < nlev(); ++l) {
      Level& c = lv(l);
      Level& f = lv(l - 1);
      c.mom64.alloc(54 * (size_t)c.ncell);
      AxisNest ax{c.ncs[0].p, c.ncn[0].p, c.na[0].p, c.nb[0].p};
      AxisNest ay{c.ncs[1].p, c.ncn[1].p, c.na[1].p, c.nb[1].p};
      AxisNest az{c.ncs[2].p, c.ncn[2].p, c.na[2].p, c.nb[2].p};
      if (l == 1) {
        hipLaunchKernelGGL(k_moments_from_fine, grid1(c.ncell), dim3(256), 0, stream, c.nx, c.ny, c.nz, f.nx, f.ny,
                           ax, ay, az, f.d64.hx.p, f.d64.hy.p, f.d64.hz.p, f.matL.p, f.matM.p, sri, c.mom64.p);
        c.cw64.alloc(16 * (size_t)c.ncell);
        hipLaunchKernelGGL(k_child_weights, grid1(c.ncell), dim3(256), 0, stream, c.nx, c.ny, c.nz, f.nx, f.ny, ax,
                           ay, az, f.d64.hx.p, f.d64.hy.p, f.d64.hz.p, f.matL.p, f.matM.p, c.cw64.p);
        // child factor tables: per coarse index, per child d: b, a^2/3 + b^2, b^2 (xi_p = a xi_c + b)
        const int ncoarse[3] = {c.nx, c.ny, c.nz};
        for (int a2 = 0; a2 < 3; ++a2) {
          c.ph_h[a2].assign(6 * (size_t)ncoarse[a2], 0.0);
          for (int C = 0; C < ncoarse[a2]; ++C)
            for (int d = 0; d < c.ncn_h[a2][C]; ++d) {
              const int i = c.ncs_h[a2][C] + d;
              const double A = c.na_h[a2][i], B = c.nb_h[a2][i];
              c.ph_h[a2][6 * C + 3 * d] = B;
              c.ph_h[a2][6 * C + 3 * d + 1] = A * A / 3.0 + B * B;
              c.ph_h[a2][6 * C + 3 * d + 2] = B * B;
            }
        }
        // split of level 1: cells of one material for the structured march, the others listed for
        // the heterogeneous-cell pass; per-axis lam second moments of one-material cells
        c.cl64.alloc(c.ncell);
        c.cm64.alloc(c.ncell);
        DevBuf<uint8_t> hetf(c.ncell);
        hipLaunchKernelGGL(k_level1_split, grid1(c.ncell), dim3(256), 0, stream, c.nx, c.ny, c.nz, f.nx, f.ny, ax, ay,
                           az, f.matL.p, f.matM.p, c.cw64.p, c.cl64.p, c.cm64.p, hetf.p);
        CM_LAUNCH_CHECK();
        std::vector<uint8_t> hf(c.ncell);
        CM_HIP(hipMemcpyAsync(hf.data(), hetf.p, c.ncell, hipMemcpyDeviceToHost, stream));
        CM_HIP(hipStreamSynchronize(stream));
        std::vector<int> hl;
        for (long long q = 0; q < c.ncell; ++q)
          if (hf[q]) hl.push_back((int)q);
        c.nhet = (int)hl.size();
        c.het.upload(hl, stream);
        {
          const long long cxn = c.nx + 1, lay1 = cxn * (c.ny + 1), ncl1 = (long long)c.nx * c.ny;
          std::vector<std::pair<long long, int>> ns;   // (node, slot)
          ns.reserve(8 * hl.size());
          for (int t = 0; t < (int)hl.size(); ++t) {
            const long long q = hl[t];
            const long long i = q % c.nx, j = (q / c.nx) % c.ny, k = q / ncl1;
            for (int l = 0; l < 8; ++l)
              ns.emplace_back((i + (l & 1)) + cxn * (j + ((l >> 1) & 1)) + lay1 * (k + (l >> 2)), 8 * t + l);
          }
          std::sort(ns.begin(), ns.end());
          std::vector<int> hn, hp{0}, hs;
          for (size_t q = 0; q < ns.size(); ++q) {
            if (q == 0 || ns[q].first != ns[q - 1].first) {
              if (q) hp.push_back((int)hs.size());
              hn.push_back((int)ns[q].first);
            }
            hs.push_back(ns[q].second);
          }
          if (!ns.empty()) hp.push_back((int)hs.size());
          std::vector<int> hm((size_t)c.nn, -1);
          for (size_t q = 0; q < hn.size(); ++q) hm[hn[q]] = (int)q;
          c.hmap.upload(hm, stream);
          c.hptr.upload(hp, stream);
          c.hslot.upload(hs, stream);
        }
        for (int a2 = 0; a2 < 3; ++a2) {
          c.q_h[a2].assign((size_t)ncoarse[a2], 1.0 / 3.0);
          if (sri)
            for (int C = 0; C < ncoarse[a2]; ++C) c.q_h[a2][C] = c.ncn_h[a2][C] == 2 ? 0.25 : 0.0;
        }
      } else {
        hipLaunchKernelGGL(k_moments_from_moments, grid1(c.ncell), dim3(256), 0, stream, c.nx, c.ny, c.nz, f.nx, f.ny,
                           f.nz, ax, ay, az, f.mom64.p, c.mom64.p);
      }
      CM_LAUNCH_CHECK();
      prepare_prec<double>(l, false);   // casts the moments just computed
    }
    lap(3);
    // diagonal blocks (FP64) -> Dinv per precision is derived lazily
    for (int l = 0; l + 1 < nlev(); ++l) compute_diag(l);
    lap(4);
    build_coarse_dense();
    lap(5);
    t_stage[6] = std::chrono::duration<double, std::milli>(clk::now() - t_begin).count();
    CM_HIP(hipEventRecord(e1, stream));
    CM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CM_HIP(hipEventElapsedTime(&ms, e0, e1));
    t_setup_ms = ms;
    CM_HIP(hipEventDestroy(e0)); CM_HIP(hipEventDestroy(e1));
    S.alloc(64);
    partial.alloc(RED_BLOCKS);
    built = true;
  }

  std::vector<DevBuf<double>> D64;   // per level FP64 diagonal blocks

  void compute_diag(int l) {
    Level& v = lv(l);
    if ((int)D64.size() < nlev()) D64.resize(nlev());
    DevBuf<double>& D = D64[l];
    D.alloc(9 * (size_t)v.nn);
    hipLaunchKernelGGL(k_fill<double>, grid1(9 * v.nn), dim3(256), 0, stream, D.p, 9 * v.nn, 0.0);
    if (l == 0) {
      if (sri)
        hipLaunchKernelGGL((k_diag_structured<false, true>), grid1(v.nng), dim3(256), 0, stream, v.nx, v.ny, v.nz,
                           v.d64.hx.p, v.d64.hy.p, v.d64.hz.p, v.applyL.p, v.applyM.p, (const double*)nullptr, D.p);
      else
        hipLaunchKernelGGL((k_diag_structured<false, false>), grid1(v.nng), dim3(256), 0, stream, v.nx, v.ny, v.nz,
                           v.d64.hx.p, v.d64.hy.p, v.d64.hz.p, v.applyL.p, v.applyM.p, (const double*)nullptr, D.p);
      CM_LAUNCH_CHECK();
      if (v.nirr) {
        if (sri)
          hipLaunchKernelGGL((k_diag_irregular<true>), grid1(v.nirr), dim3(256), 0, stream, v.nirr, v.irr_nodes.p,
                             v.irr_ijk.p, v.d64.irrL.p, v.d64.irrM.p, v.d64.hx.p, v.d64.hy.p, v.d64.hz.p, D.p);
        else
          hipLaunchKernelGGL((k_diag_irregular<false>), grid1(v.nirr), dim3(256), 0, stream, v.nirr, v.irr_nodes.p,
                             v.irr_ijk.p, v.d64.irrL.p, v.d64.irrM.p, v.d64.hx.p, v.d64.hy.p, v.d64.hz.p, D.p);
        CM_LAUNCH_CHECK();
      }
    } else {
      hipLaunchKernelGGL((k_diag_structured<true, false>), grid1(v.nng), dim3(256), 0, stream, v.nx, v.ny, v.nz,
                         v.d64.hx.p, v.d64.hy.p, v.d64.hz.p, (const double*)nullptr, (const double*)nullptr,
                         v.mom64.p, D.p);
      CM_LAUNCH_CHECK();
    }
    const int nrow = (int)v.xrow_h.size();
    if (nrow) {
      DevBuf<double> xv;
      xv.upload(v.xval_h, stream);
      hipLaunchKernelGGL(k_diag_extra, grid1(nrow), dim3(256), 0, stream, nrow, v.xrow.p, v.xptr.p, v.xcol.p, xv.p,
                         D.p);
      CM_LAUNCH_CHECK();
      CM_HIP(hipStreamSynchronize(stream));
    }
    const int nterm = (int)v.r1k_h.size();
    if (nterm) {
      hipLaunchKernelGGL(k_diag_rank1, grid1(64LL * nterm), dim3(256), 0, stream, nterm, v.r1ptr.p, v.r1dof.p, v.r1w64.p,
                         v.r1k64.p, D.p);
      CM_LAUNCH_CHECK();
    }
  }

  void build_coarse_dense() {
    const int l = nlev() - 1;
    Level& c = lv(l);
    const long long n = 3 * c.nn;
    CM_CHECK(n <= 30000, "coarsest level too large for the dense direct solve; use more levels");
    // assembled over all DOFs of the level, then restricted to the used ones (the others have
    // empty rows and columns) and inverted there
    DevBuf<double> Afull((size_t)n * n);
    hipLaunchKernelGGL(k_fill<double>, grid1(n * n), dim3(256), 0, stream, Afull.p, n * n, 0.0);
    CM_LAUNCH_CHECK();
    if (l == 0) {
      // single-level hierarchy: the fine operator is assembled column by column
      DevBuf<double> e(n), y(n);
      for (long long j = 0; j < n; ++j) {
        hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, stream, e.p, n, 0.0);
        const double one = 1.0;
        CM_HIP(hipMemcpyAsync(e.p + j, &one, sizeof(double), hipMemcpyHostToDevice, stream));
        apply<double>(0, e.p, y.p, MODE_APPLY, nullptr);
        CM_HIP(hipMemcpyAsync(Afull.p + j * n, y.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        CM_HIP(hipStreamSynchronize(stream));
      }
    } else {
      hipLaunchKernelGGL(k_dense_assemble, dim3((unsigned)c.ncell), dim3(64), 0, stream, c.nx, c.ny, c.nz,
                         c.d64.hx.p, c.d64.hy.p, c.d64.hz.p, c.mom64.p, (int)n, Afull.p);
      CM_LAUNCH_CHECK();
      const int nrow = (int)c.xrow_h.size();
      if (nrow) {
        DevBuf<double> xv;
        xv.upload(c.xval_h, stream);
        hipLaunchKernelGGL(k_dense_add_extra, grid1(nrow), dim3(256), 0, stream, nrow, c.xrow.p, c.xptr.p, c.xcol.p,
                           xv.p, (int)n, Afull.p);
        CM_LAUNCH_CHECK();
        CM_HIP(hipStreamSynchronize(stream));
      }
      const int nterm = (int)c.r1k_h.size();
      if (nterm) {
        hipLaunchKernelGGL(k_dense_add_rank1, dim3((unsigned)nterm), dim3(256), 0, stream, c.r1ptr.p, c.r1dof.p,
                           c.r1w64.p, c.r1k64.p, (int)n, Afull.p);
        CM_LAUNCH_CHECK();
      }
    }
    std::vector<int> map;
    for (long long nd = 0; nd < c.nn; ++nd)
      if (c.used_h[nd])
        for (int q = 0; q < 3; ++q) map.push_back((int)(3 * nd + q));
    const long long m = (long long)map.size();
    CM_CHECK(m > 0, "coarsest level has no used DOF");
    dmap.upload(map, stream);
    nc_dense = (int)m;
    Ainv.alloc((size_t)m * m);
    hipLaunchKernelGGL(k_dense_gather, grid1(m * m), dim3(256), 0, stream, (int)m, dmap.p, (int)n, Afull.p, Ainv.p);
    hipLaunchKernelGGL(k_dense_fix_empty, dim3((unsigned)m), dim3(256), 0, stream, (int)m, Ainv.p);
    CM_LAUNCH_CHECK();
    CM_HIP(hipStreamSynchronize(stream));
    Afull.release();
    invert_dense((int)m);
  }

  // In-place blocked Gauss-Jordan inverse of the SPD matrix Ainv (m x m).
  void invert_dense(int m) {
    const long long n = m;
    DevBuf<double> Dk(GJB * GJB), Ct((size_t)GJB * n), Wt((size_t)GJB * n);
    const unsigned nt = (unsigned)div_up(n, 64);
    for (int k0 = 0; k0 < n; k0 += GJB) {
      const int nb = (int)std::min<long long>(GJB, n - k0);
      hipLaunchKernelGGL(k_gjb_diag, dim3(1), dim3(256), 0, stream, (int)n, k0, nb, Ainv.p, Dk.p);
      hipLaunchKernelGGL(k_gjs_panel, dim3(nt), dim3(256), 0, stream, (int)n, k0, nb, Ainv.p, Dk.p, Ct.p, Wt.p);
      hipLaunchKernelGGL(k_gjs_update, dim3(nt, nt), dim3(256), 0, stream, (int)n, k0, Ct.p, Wt.p, Ainv.p);
    }
    hipLaunchKernelGGL(k_gjs_mirror, dim3(nt, nt), dim3(256), 0, stream, (int)n, Ainv.p);
    CM_LAUNCH_CHECK();
    CM_HIP(hipStreamSynchronize(stream));
  }

  // Level data in precision T (geometry, materials, moments, Dinv, weights, work vectors).
  template <typename T>
  void prepare_prec(int l, bool with_dinv) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    if (!d.ready) {
      std::vector<double> h[3];
      for (int ax = 0; ax < 3; ++ax) {
        h[ax].resize(v.e[ax].size() - 1);
        for (size_t i = 0; i + 1 < v.e[ax].size(); ++i) h[ax][i] = v.e[ax][i + 1] - v.e[ax][i];
      }
      upload_cast(d.hx, h[0], stream); upload_cast(d.hy, h[1], stream); upload_cast(d.hz, h[2], stream);
      if (l == 0) {
        if (v.label_mode) {
          upload_cast(d.tabL, v.tabL_h, stream);
          upload_cast(d.tabM, v.tabM_h, stream);
        } else {
          d.cellL.alloc(v.ncell); d.cellM.alloc(v.ncell);
          hipLaunchKernelGGL((k_cast<T, double>), grid1(v.ncell), dim3(256), 0, stream, v.ncell, d.cellL.p, v.applyL.p);
          hipLaunchKernelGGL((k_cast<T, double>), grid1(v.ncell), dim3(256), 0, stream, v.ncell, d.cellM.p, v.applyM.p);
          CM_LAUNCH_CHECK();
        }
        if (v.nirr) { upload_cast(d.irrL, v.irrL_h, stream); upload_cast(d.irrM, v.irrM_h, stream); }
      } else if (l == 1) {
        d.cw.alloc(v.cw64.n);
        hipLaunchKernelGGL((k_cast<T, double>), grid1((long long)v.cw64.n), dim3(256), 0, stream,
                           (long long)v.cw64.n, d.cw.p, v.cw64.p);
        d.cl1.alloc(v.ncell);
        d.cm1.alloc(v.ncell);
        hipLaunchKernelGGL((k_cast<T, double>), grid1(v.ncell), dim3(256), 0, stream, v.ncell, d.cl1.p, v.cl64.p);
        hipLaunchKernelGGL((k_cast<T, double>), grid1(v.ncell), dim3(256), 0, stream, v.ncell, d.cm1.p, v.cm64.p);
        CM_LAUNCH_CHECK();
        for (int ax = 0; ax < 3; ++ax) {
          upload_cast(d.ph[ax], v.ph_h[ax], stream);
          upload_cast(d.q1[ax], v.q_h[ax], stream);
        }
        d.hf.alloc(24 * (size_t)v.nhet);
      } else if (v.mom64.n) {
        d.mom.alloc(v.mom64.n);
        hipLaunchKernelGGL((k_cast<T, double>), grid1((long long)v.mom64.n), dim3(256), 0, stream,
                           (long long)v.mom64.n, d.mom.p, v.mom64.p);
        CM_LAUNCH_CHECK();
      }
      if (l + 1 < nlev()) {
        for (int ax = 0; ax < 3; ++ax) { upload_cast(d.pw[ax], v.pw_h[ax], stream); upload_cast(d.rw[ax], v.rw_h[ax], stream); }
      }
      if (!v.xval_h.empty()) upload_cast(d.xval, v.xval_h, stream);
      if (!v.r1k_h.empty()) { upload_cast(d.r1w, v.r1w_h, stream); upload_cast(d.r1k, v.r1k_h, stream); }
      if (!v.bval_h.empty()) upload_cast(d.bval, v.bval_h, stream);
      // level 1 always has the accumulator: its heterogeneous cells add into it
      if (l > 0 && (l == 1 || v.nblk || !v.r1k_h.empty())) {
        d.ex.alloc(3 * (size_t)v.nn);
        hipLaunchKernelGGL(k_fill<T>, grid1(3LL * v.nn), dim3(256), 0, stream, d.ex.p, 3LL * v.nn, T(0));
        CM_LAUNCH_CHECK();
      }
      const size_t nv = 3 * (size_t)v.nn;
      d.x.alloc(nv); d.b.alloc(nv); d.r.alloc(nv); d.t.alloc(nv); d.d.alloc(nv);
      if (l == nlev() - 1) {   // the coarse solve writes only the used DOFs of x
        hipLaunchKernelGGL(k_fill<T>, grid1((long long)nv), dim3(256), 0, stream, d.x.p, (long long)nv, T(0));
        CM_LAUNCH_CHECK();
      }
      if (v.ntouch) {
        d.tc.alloc(3 * (size_t)v.ntouch);
        hipLaunchKernelGGL(k_fill<T>, grid1(3LL * v.ntouch), dim3(256), 0, stream, d.tc.p, 3LL * v.ntouch, T(0));
        CM_LAUNCH_CHECK();
      }
      d.ready = true;
    }
    if (with_dinv && !d.Dinv.n && l + 1 < nlev()) {
      d.Dinv.alloc(6 * (size_t)v.nn);
      hipLaunchKernelGGL(k_block_inverse<T>, grid1(v.nn), dim3(256), 0, stream, v.nn, D64[l].p, d.Dinv.p);
      CM_LAUNCH_CHECK();
      if (l == 0 && v.ngroup) {
        d.ginv.alloc(v.gwork64.n);
        hipLaunchKernelGGL(k_group_inverse<T>, grid1(v.ngroup), dim3(256), 0, stream, v.ngroup, v.gptr.p, v.gslot.p,
                           v.tnode.p, v.goff.p, v.goffd64.p, D64[l].p, v.gwork64.p, d.Dinv.p, d.ginv.p);
        CM_LAUNCH_CHECK();
      }
      if (opt.smoother == 0) power_iteration<T>(l);
    }
    refresh_flags<T>(l);
  }

  // every level ready for the V-cycle in precision T (data, block-Jacobi inverses, Chebyshev bounds)
  template <typename T>
  void ensure_prec() {
    for (int l = 0; l < nlev(); ++l) prepare_prec<T>(l, true);
  }
  // every level's operator data in precision T (no smoother set-up)
  template <typename T>
  void ensure_data() {
    for (int l = 0; l < nlev(); ++l) prepare_prec<T>(l, false);
  }

  // ------------------------------------------------------------------ operator
  template <typename T>
  GridArgs<T> grid_args(int l) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    return GridArgs<T>{v.nx, v.ny, v.nz, d.hx.p, d.hy.p, d.hz.p, T(v.iso_fac)};
  }

  // y = A u (MODE_APPLY) or y = b - A u (MODE_RESID); b may alias y.
  template <typename T>
  void apply(int l, const T* u, T* y, int mode, const T* b) {
    Epi<T> ep{b, y, nullptr, nullptr, nullptr, T(0), T(0), T(0)};
    if (mode == MODE_APPLY) apply_ep<T, EPI_APPLY>(l, u, ep);
    else apply_ep<T, EPI_RESID>(l, u, ep);
  }

  // One operator application on s with a fused epilogue: sparse pass -> structured march ->
  // fix-up of the touched nodes (which also clears the sparse accumulator).
  template <typename T, int EM>
  void apply_ep(int l, const T* s, const Epi<T>& ep) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    if (l > 0 && d.ex.p) launch_extra<T>(l, s);
    launch_march<T, EM>(l, s, ep, v.ntouch ? sparse_args<T>(v, d) : SparseArgs<T>{});
    if (v.ntouch) {
      const CopyGroups<T> ga = EM == EPI_CHEB ? groups_of<T>(l) : CopyGroups<T>{};
      CM_CHECK(!ga.ng || ep.y, "split-copy groups need the residual output");
      go(FixupL<T, EM>{(int)v.ntouch, v.tnode.p, v.nng, d.tc.p, s, ep, ga}, div_up(v.ntouch + ga.ng, 256));
    }
    CM_LAUNCH_CHECK();
  }

  // extra stiffness of a coarse level into its accumulator ex (consumed by the next march)
  template <typename T>
  void launch_extra(int l, const T* s) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    CoarseExtra<T> a{};
    a.nterm = (int)v.r1k_h.size();
    a.nblk = (int)v.nblk;
    a.rw = term_lanes(v);
    a.nb_term = (int)div_up((long long)a.rw * a.nterm, 256);
    a.r1ptr = v.r1ptr.p; a.r1dof = v.r1dof.p; a.r1w = d.r1w.p; a.r1k = d.r1k.p;
    a.brow = v.brow.p; a.bcol = v.bcol.p; a.bval = d.bval.p;
    const int nbx = a.nb_term + div_up(a.nblk, 256);
    // level 1: its several-material cells ride on this launch (the march gathers their forces)
    HetArgs<T> h{};
    if (l == 1)
      h = HetArgs<T>{v.nhet, sri, v.het.p, v.nx, v.ny, v.nz, d.hx.p, d.hy.p, d.hz.p, d.cw.p, d.ph[0].p, d.ph[1].p,
                     d.ph[2].p, d.hf.p};
    // recorded even when empty: batched systems keep one launch order
    go(CoarseExtraL<T>{a, s, d.ex.p, nbx, h}, (long long)nbx + div_up(h.n, 256));
  }

  // lanes per rank-one term in the sparse passes: the power of two >= the mean support (<= 64)
  static int term_lanes(const Level& v) {
    const size_t nt = v.r1ptr_h.empty() ? 0 : v.r1ptr_h.size() - 1;
    const double avg = nt ? (double)v.r1dof_h.size() / nt : 1.0;
    int w = 1;
    while (w < 64 && w < avg) w <<= 1;
    return w;
  }

  // Ap = A p with the dot p . Ap into S[slot] (FP64 PCG): the march and the fix-up leave per-block
  // partials, summed in fixed order (no separate pass over p and Ap).
  void apply_dot(int l, const double* p, double* Ap, int slot) {
    Level& v = lv(l);
    const int nbm = (int)march_shape<double>(l, pick_kz<double>(v, 0)).nb;
    const int nbf = v.ntouch ? div_up(v.ntouch, 256) : 0;
    CM_CHECK(partial.n >= (size_t)(nbm + nbf), "partial buffer too small for the fused dot");
    Epi<double> ep{nullptr, Ap, nullptr, nullptr, nullptr, 0.0, 0.0, 0.0, partial.p};
    LevelT<double>& d = v.dat<double>();
    if (l > 0 && d.ex.p) launch_extra<double>(l, p);
    launch_march<double, EPI_APPLY>(l, p, ep, v.ntouch ? sparse_args<double>(v, d) : SparseArgs<double>{});
    if (v.ntouch) {
      Epi<double> ef = ep;
      ef.part = partial.p + nbm;
      go(FixupL<double, EPI_APPLY>{(int)v.ntouch, v.tnode.p, v.nng, d.tc.p, p, ef, CopyGroups<double>{}},
         div_up(v.ntouch, 256));
    }
    go(FinalizeL{nbm + nbf, partial.p, S.p, slot}, 1);
    CM_LAUNCH_CHECK();
  }
  // partials needed by apply_dot on level l (FP64)
  size_t dot_partials(int l) {
    Level& v = lv(l);
    return (size_t)march_shape<double>(l, pick_kz<double>(v, 0)).nb + (v.ntouch ? div_up(v.ntouch, 256) : 0);
  }

  // the level-0 sparse pass, run as extra blocks of the march launch
  template <typename T>
  SparseArgs<T> sparse_args(Level& v, LevelT<T>& d) {
    SparseArgs<T> a{};
    a.nirr = v.nirr;
    a.nrow = (int)v.xrow_h.size();
    a.nterm = (int)v.r1k_h.size();
    a.xw = v.xw;
    a.nb_irr = (int)div_up(a.nirr, 256);
    a.nb_row = (int)div_up((long long)a.nrow * v.xw, 256);
    a.rw = term_lanes(v);
    const int nb_r1 = (int)div_up((long long)a.rw * a.nterm, 256);
    a.nb = a.nb_irr + a.nb_row + nb_r1;
    a.irr_nodes = v.irr_nodes.p; a.irr_ci = v.irr_ci.p; a.irr_ijk = v.irr_ijk.p; a.irrL = d.irrL.p; a.irrM = d.irrM.p;
    a.xrc = v.xrc.p; a.xptr = v.xptr.p; a.xcol = v.xcol.p; a.xval = d.xval.p;
    a.r1ptr = v.r1ptr.p; a.r1dof = v.r1dof.p; a.r1c = v.r1c.p; a.r1w = d.r1w.p; a.r1k = d.r1k.p;
    a.tc = d.tc.p;
    return a;
  }

  // Calls f with the cell-material policy of level l's structured march.
  template <typename T, class F>
  void with_mat(int l, F&& f) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    if (l == 1) {
      if (sri)
        f(SplitMat<T, true>{d.cl1.p, d.cm1.p, d.q1[0].p, d.q1[1].p, d.q1[2].p, d.ex.p, v.hmap.p, v.hptr.p, v.hslot.p,
                            d.hf.p});
      else
        f(SplitMat<T, false>{d.cl1.p, d.cm1.p, d.q1[0].p, d.q1[1].p, d.q1[2].p, d.ex.p, v.hmap.p, v.hptr.p,
                             v.hslot.p, d.hf.p});
      return;
    }
    if (l > 1) {
      f(MomentMat<T>{d.mom.p, d.ex.p});
      return;
    }
    const bool iso = v.iso_fac > 0.0;
#define CM_FINE(MAT, SRI_, ISO_, ...)      \
  if (sri == SRI_ && iso == ISO_) {        \
    f(MAT<T, SRI_, ISO_>{__VA_ARGS__});    \
    return;                                \
  }
    if (v.label_mode) {
      CM_FINE(LabelMat, true, true, v.lab.p, d.tabL.p, d.tabM.p)
      CM_FINE(LabelMat, true, false, v.lab.p, d.tabL.p, d.tabM.p)
      CM_FINE(LabelMat, false, true, v.lab.p, d.tabL.p, d.tabM.p)
      CM_FINE(LabelMat, false, false, v.lab.p, d.tabL.p, d.tabM.p)
    } else {
      CM_FINE(ValueMat, true, true, d.cellL.p, d.cellM.p)
      CM_FINE(ValueMat, true, false, d.cellL.p, d.cellM.p)
      CM_FINE(ValueMat, false, true, d.cellL.p, d.cellM.p)
      CM_FINE(ValueMat, false, false, d.cellL.p, d.cellM.p)
    }
#undef CM_FINE
  }

  struct MarchShape {
    int kz, ntx, nty, ntz;
    long long nb;
  };
  template <typename T>
  MarchShape march_shape(int l, int kz) {
    Level& v = lv(l);
    MarchShape m{kz, (int)div_up(v.nx + 1, TOUT), (int)div_up(v.ny + 1, TOUT), (int)div_up(v.nz + 1, kz), 0};
    m.nb = (long long)m.ntx * m.nty * m.ntz;
    CM_CHECK(m.nb < (1LL << 31), "grid too large for one launch");
    return m;
  }

  // Empty-block flags of level l for the z extent the march uses now (rebuilt when it changed).
  // Runs outside graph capture: entry points refresh before they launch or capture.
  template <typename T>
  void refresh_flags(int l) {
    refresh_flags_ec<T>(l, 0);
    refresh_flags_ec<T>(l, 1);
  }
  template <typename T>
  void refresh_flags_ec(int l, int ec) {
    Level& v = lv(l);
    if (!v.dat<T>().ready || l == nlev() - 1) return;
    const int slot = std::is_same<T, float>::value ? 1 : 0;
    const int kz = pick_kz<T>(v, ec);
    if (v.bflag_kz[slot][ec] == kz) return;
    const MarchShape ms = march_shape<T>(l, kz);
    if (v.bflag[slot].n < (size_t)ms.nb) v.bflag[slot].alloc((size_t)march_shape<T>(l, 1).nb);
    const GridArgs<T> g = grid_args<T>(l);
    with_mat<T>(l, [&](auto m) {
      using M = decltype(m);
      hipLaunchKernelGGL((k_block_flags<T, M>), dim3((unsigned)ms.nb), dim3(BLOCK), 0, stream, g, m, kz, ms.ntx,
                         ms.nty, v.bflag[slot].p);
    });
    CM_LAUNCH_CHECK();
    const int nb = (int)ms.nb;
    std::vector<uint8_t> fl(nb);
    CM_HIP(hipMemcpyAsync(fl.data(), v.bflag[slot].p, nb, hipMemcpyDeviceToHost, stream));
    CM_HIP(hipStreamSynchronize(stream));
    // blocks 0 .. nf-1 take the tiles with stiffness, the rest the others; block b runs on XCD
    // b % 8, and every XCD gets one contiguous run of each list (tile order: x, then y, then z)
    std::vector<int> full, none;
    for (int id = 0; id < nb; ++id) (fl[id] ? full : none).push_back(id);
    const int nf = (int)full.size();
    std::vector<int> ord(nb);
    auto deal = [&](const std::vector<int>& list, int b0, bool neg) {
      const int n = (int)list.size();
      int cnt[8], start[8], first[8];
      for (int x = 0; x < 8; ++x) {
        first[x] = b0 + ((x - b0 % 8) + 8) % 8;                   // first block >= b0 on XCD x
        cnt[x] = first[x] < b0 + n ? (b0 + n - first[x] + 7) / 8 : 0;
      }
      start[0] = 0;
      for (int x = 1; x < 8; ++x) start[x] = start[x - 1] + cnt[x - 1];
      for (int b = b0; b < b0 + n; ++b) {
        const int x = b % 8, pos = (b - first[x]) / 8;
        const int id = list[start[x] + pos];
        ord[b] = neg ? ~id : id;
      }
    };
    deal(full, 0, false);
    deal(none, nf, true);
    v.border[slot][ec].upload(ord, stream);
    v.bflag_kz[slot][ec] = kz;
  }
  template <typename T>
  void refresh_all_flags() {
    for (int l = 0; l < nlev(); ++l) refresh_flags<T>(l);
  }

  // structured march of level l with epilogue EM
  template <typename T, int EM>
  void launch_march(int l, const T* s, const Epi<T>& ep, const SparseArgs<T>& sp = SparseArgs<T>{}) {
    Level& v = lv(l);
    const GridArgs<T> g = grid_args<T>(l);
    const int slot = std::is_same<T, float>::value ? 1 : 0;
    const int ec = EM == EPI_CHEB ? 1 : 0;
    const MarchShape ms = march_shape<T>(l, pick_kz<T>(v, ec));
    const TileMap tm{ms.ntx, ms.nty, ms.ntz, (int)(ms.nb / 8), (int)(ms.nb % 8)};
    const int* fl = (v.bflag_kz[slot][ec] == ms.kz && use_flags) ? v.border[slot][ec].p : nullptr;
    with_mat<T>(l, [&](auto m) {
      using M = decltype(m);
      CM_CHECK(sp.nb == 0 || M::kind <= 1, "the sparse pass rides only on level-0 marches");
      go(MarchL<T, M, EM>{g, m, s, ep, ms.kz, tm, fl, sp}, ms.nb + sp.nb);
    });
  }

  // structured part only, plain product (timing, tuning)
  template <typename T>
  void apply_structured(int l, const T* u, T* y) {
    Epi<T> ep{nullptr, y, nullptr, nullptr, nullptr, T(0), T(0), T(0)};
    launch_march<T, EPI_APPLY>(l, u, ep);
    CM_LAUNCH_CHECK();
  }
  // The structured march of one middle Chebyshev step alone (timing): r = b - A s, d = c1 s +
  // c2 Dinv r, x += s + d on the level's own work vectors (b, x cleared by the caller once)
  template <typename T>
  void smooth_step_structured(int l, const T* s) {
    LevelT<T>& d = lv(l).dat<T>();
    Epi<T> ep{d.b.p, d.r.p, d.Dinv.p, d.d.p, d.x.p, T(0.5), T(0.25), T(1)};
    launch_march<T, EPI_CHEB>(l, s, ep);
    CM_LAUNCH_CHECK();
  }

  int kz_override = 0;
  long long march_blocks = 600;   // the heuristic's minimum number of march blocks per launch
  bool use_flags = true;      // empty blocks of the structured march take the epilogue-only path
  template <typename T>
  int pick_kz(const Level& v, int ec) const {
    if (kz_override > 0) return kz_override;
    const int slot = std::is_same<T, float>::value ? 1 : 0;
    const int tuned = v.kz[slot][ec] > 0 ? v.kz[slot][ec] : v.kz[slot][0];
    if (tuned > 0) return tuned;
    const long long xy = (long long)div_up(v.nx + 1, TOUT) * div_up(v.ny + 1, TOUT);
    // the longest z chain that still gives >= march_blocks blocks (default 600, about 2.3 blocks
    // per CU); measured on MI355X this is the fastest or within 5 % of it on every level of the
    // hindfoot raster solved alone
    for (int kz : {32, 16, 8, 4, 2}) {
      if (xy * div_up(v.nz + 1, kz) >= march_blocks) return kz;
    }
    return 1;
  }

  // Tile autotuner: time the structured kernel of every level for each z extent and keep the
  // fastest (HIP events, on this system's stream).  Results are kept process-wide per level shape
  // and kernel variant, so later systems of the same shape skip the timing.
  struct KzKey {
    int nx, ny, nz, variant;
    bool operator==(const KzKey& o) const { return nx == o.nx && ny == o.ny && nz == o.nz && variant == o.variant; }
  };
  struct KzHash {
    size_t operator()(const KzKey& k) const {
      return ((size_t)k.nx * 73856093u) ^ ((size_t)k.ny * 19349663u) ^ ((size_t)k.nz * 83492791u) ^ (size_t)k.variant;
    }
  };
  static std::mutex& kz_mu() { static std::mutex m; return m; }
  static std::unordered_map<KzKey, int, KzHash>& kz_cache() {
    static auto* c = new std::unordered_map<KzKey, int, KzHash>;
    return *c;
  }
  template <typename T>
  int kz_variant(int l) const {
    const Level& v = *L[l];
    const int prec = std::is_same<T, float>::value ? 1 : 0;
    const int kind = l > 0 ? 3 : (v.label_mode ? 1 : 2);
    return prec + 2 * (kind + 4 * ((l == 0 && sri) + 2 * (l == 0 && v.iso_fac > 0.0)));
  }
  // Class 0 (products, residuals) is timed on the plain product, class 1 on the fused smoothing
  // step where the level has its block-Jacobi inverse in precision T (else it takes class 0's).
  template <typename T>
  void autotune_kz(int nrep, bool retime = false) {
    hipEvent_t e0, e1;
    CM_HIP(hipEventCreate(&e0)); CM_HIP(hipEventCreate(&e1));
    const int slot = std::is_same<T, float>::value ? 1 : 0;
    for (int l = 0; l < nlev() - 1; ++l) {
      Level& v = lv(l);
      LevelT<T>& d = v.dat<T>();
      for (int ec = 0; ec < 2; ++ec) {
        if (ec == 1 && !d.Dinv.n) {
          v.kz[slot][1] = v.kz[slot][0];
          refresh_flags_ec<T>(l, 1);
          continue;
        }
        const KzKey key{v.nx, v.ny, v.nz, 2 * kz_variant<T>(l) + ec};
        if (!retime) {
          std::lock_guard<std::mutex> g(kz_mu());
          auto it = kz_cache().find(key);
          if (it != kz_cache().end()) {
            v.kz[slot][ec] = it->second;
            refresh_flags_ec<T>(l, ec);
            continue;
          }
        }
        auto run = [&]() {
          if (ec == 0) apply_structured<T>(l, d.x.p, d.t.p);
          else smooth_step_structured<T>(l, d.t.p);
        };
        // two passes over the extents, the faster of each extent's two timings kept (a single
        // pass picked within-noise losers on large grids)
        constexpr int KZS[6] = {1, 2, 4, 8, 16, 32};
        float tbest[6];
        for (float& t : tbest) t = 1e30f;
        for (int pass = 0; pass < 2; ++pass)
          for (int q = 0; q < 6; ++q) {
            const int kz = KZS[q];
            if (kz > v.nz + 1 && kz != 1) continue;
            v.kz[slot][ec] = kz;
            refresh_flags_ec<T>(l, ec);
            run();
            CM_HIP(hipEventRecord(e0, stream));
            for (int r = 0; r < nrep; ++r) run();
            CM_HIP(hipEventRecord(e1, stream));
            CM_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CM_HIP(hipEventElapsedTime(&ms, e0, e1));
            tbest[q] = std::min(tbest[q], ms);
          }
        float best = 1e30f;
        int best_kz = 1;
        for (int q = 0; q < 6; ++q)
          if (tbest[q] < best) { best = tbest[q]; best_kz = KZS[q]; }
        v.kz[slot][ec] = best_kz;
        refresh_flags_ec<T>(l, ec);
        std::lock_guard<std::mutex> g(kz_mu());
        kz_cache()[key] = best_kz;
      }
    }
    CM_HIP(hipEventDestroy(e0)); CM_HIP(hipEventDestroy(e1));
  }

  // ------------------------------------------------------------------ smoothers, V-cycle
  // lambda_max of M A by 12 power steps from a fixed pseudo-normal start (fe/gmg.py:51-59):
  // lam = |y| / |x| of the last step, then lmax = 1.1 lam, lmin = lmax / 30.  The iterate is
  // normalised on the device; the host reads the result once.
  // x0: start vector in the level's layout (null: the built-in pseudo-normal start)
  template <typename T>
  double power_iteration(int l, const T* x0 = nullptr) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    const long long n = 3 * v.nn;
    if (x0)
      CM_HIP(hipMemcpyAsync(d.x.p, x0, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
    else
      hipLaunchKernelGGL(k_randn<T>, grid1(v.nn), dim3(256), 0, stream, v.nn, v.used.p, 0ULL, d.x.p);
    DevBuf<double> s(4);
    for (int it = 0; it < 12; ++it) {
      apply<T>(l, d.x.p, d.t.p, MODE_APPLY, nullptr);
      const CopyGroups<T> ga = groups_of<T>(l);
      hipLaunchKernelGGL(k_block_apply<T>, grid1(v.nn + ga.ng), dim3(256), 0, stream, v.nn, d.Dinv.p, d.t.p, d.d.p,
                         ga);
      hipLaunchKernelGGL((k_dot<T, T>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, d.d.p, d.d.p, partial.p);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_THREADS), 0, stream, RED_BLOCKS, partial.p, s.p, 0);
      if (it == 11) {
        hipLaunchKernelGGL((k_dot<T, T>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, d.x.p, d.x.p, partial.p);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_THREADS), 0, stream, RED_BLOCKS, partial.p, s.p, 1);
      } else {
        // x = y / |y|
        CM_HIP(hipMemcpyAsync(d.x.p, d.d.p, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
        hipLaunchKernelGGL(k_scale<T>, grid1(n), dim3(256), 0, stream, n, d.x.p, s.p, 0);
      }
      CM_LAUNCH_CHECK();
    }
    double hs[2];
    CM_HIP(hipMemcpyAsync(hs, s.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    CM_HIP(hipStreamSynchronize(stream));
    const double lam = std::sqrt(hs[0]) / std::sqrt(hs[1]);
    d.lmax = 1.1 * lam;
    d.lmin = d.lmax / 30.0;
    return lam;
  }

  // coefficient c of the first smoothing step from zero, x = c Dinv b
  template <typename T>
  double first_step_coef(int l) {
    LevelT<T>& d = lv(l).dat<T>();
    return opt.smoother == 1 ? opt.omega : 1.0 / (0.5 * (d.lmax + d.lmin));
  }
  // x_started: zero_start with that first step already taken (by the restriction)
  template <typename T>
  void smooth(int l, const T* b, T* x, bool zero_start, int nsweep, bool x_started = false) {
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    const long long nn = v.nn;
    // split-copy groups need every step's residual (their fix-up completes it)
    const CopyGroups<T> ga = groups_of<T>(l);
    T* rkeep = ga.ng ? d.r.p : nullptr;
    if (opt.smoother == 1) {
      // damped block Jacobi: x += omega Dinv (b - A x); the step lands in d, then x += d
      int s0 = 0;
      if (zero_start) {
        if (!x_started) go(Cheb0L<T>{nn, x, b, d.Dinv.p, T(opt.omega), ga}, div_up(nn + ga.ng, 256));
        s0 = 1;
      }
      for (int s = s0; s < nsweep; ++s) {
        const Epi<T> ep{b, rkeep, d.Dinv.p, d.d.p, nullptr, T(0), T(opt.omega), T(0)};
        apply_ep<T, EPI_CHEB>(l, x, ep);
        go(AddL<T, T>{3 * nn, x, d.d.p}, div_up(3 * nn, 256));
      }
      CM_LAUNCH_CHECK();
      return;
    }
    // Chebyshev of degree opt.degree (fe/gmg.py:61-76):
    //   r = b - A x;  d = M r / theta;  x += d;
    //   repeat: r -= A d;  rho' = 1 / (2 sigma - rho);  d = rho' rho d + 2 rho' / delta M r;  x += d
    // Every "r -= A d; d = ...; x += ..." is one fused operator application with d as its input.
    // Directions alternate between two buffers; x is updated in place whenever it is not the
    // operator's input (a direction that could not be added yet is added by the next step: xa).
    const double theta = 0.5 * (d.lmax + d.lmin), delta = 0.5 * (d.lmax - d.lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    T* r = d.r.p;
    T* dirs[2] = {d.d.p, d.t.p};
    const T* s;            // input of the next application
    const T* rin = b;
    bool pending;          // s is a direction not yet added to x
    int nd = 0;            // next direction buffer
    if (zero_start) {
      if (!x_started) go(Cheb0L<T>{nn, x, b, d.Dinv.p, T(1.0 / theta), ga}, div_up(nn + ga.ng, 256));
      s = x;
      pending = false;
    } else {
      const Epi<T> ep{b, r, d.Dinv.p, dirs[nd], nullptr, T(0), T(1.0 / theta), T(0)};
      apply_ep<T, EPI_CHEB>(l, x, ep);
      s = dirs[nd];
      nd ^= 1;
      rin = r;
      pending = true;
    }
    for (int k = 1; k < opt.degree; ++k) {
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      const bool last = (k == opt.degree - 1);
      const bool upd = (s != x);
      T* dout = (!last || !upd) ? dirs[nd] : nullptr;
      const Epi<T> ep{rin, last ? rkeep : r, d.Dinv.p, dout, upd ? x : nullptr, T(rho_new * rho),
                      T(2.0 * rho_new / delta), T(pending ? 1 : 0)};
      apply_ep<T, EPI_CHEB>(l, s, ep);
      pending = !upd;
      if (dout) {
        s = dout;
        nd ^= 1;
      }
      rin = r;
      rho = rho_new;
    }
    if (pending && s != x) go(AddL<T, T>{3 * nn, x, s}, div_up(3 * nn, 256));
    CM_LAUNCH_CHECK();
  }

  template <typename T>
  void vcycle(int l, const T* b, T* x, bool x_started = false) {
    if (l == nlev() - 1) {
      if (l == 0) {   // one level: x is the caller's vector; the solve writes only the used DOFs
        const long long nv = 3 * lv(0).nn;
        go(FillL<T>{x, nv, T(0)}, div_up(nv, 256));
      }
      go(GemvL<T>{nc_dense, Ainv.p, dmap.p, b, x}, nc_dense);
      CM_LAUNCH_CHECK();
      return;
    }
    Level& v = lv(l);
    LevelT<T>& d = v.dat<T>();
    Level& c = lv(l + 1);
    LevelT<T>& dc = c.dat<T>();
    smooth<T>(l, b, x, true, opt.npre, x_started);
    const int ncopy = (int)(v.nn - v.nng);
    {
      // residual for the restriction; split copies fold theirs into their originals
      Epi<T> ep{b, d.r.p, nullptr, nullptr, nullptr, T(0), T(0), T(0)};
      ep.fold = ncopy ? v.copy_orig.p : nullptr;
      apply_ep<T, EPI_RESID>(l, x, ep);
    }
    AxisR<T> rx{v.rs[0].p, d.rw[0].p}, ry{v.rs[1].p, d.rw[1].p}, rz{v.rs[2].p, d.rw[2].p};
    AxisP<T> px{v.pia[0].p, d.pw[0].p}, py{v.pia[1].p, d.pw[1].p}, pz{v.pia[2].p, d.pw[2].p};
    // the restriction also takes the coarse level's first smoothing step from zero (one launch
    // less); the coarsest level is solved directly instead
    const bool start_c = l + 2 < nlev();
    go(RestrictL<T>{v.nx + 1, v.ny + 1, v.nz + 1, c.nx + 1, c.ny + 1, c.nz + 1, rx, ry, rz, d.r.p, dc.b.p,
                    start_c ? dc.Dinv.p : (const T*)nullptr, start_c ? dc.x.p : (T*)nullptr,
                    T(start_c ? first_step_coef<T>(l + 1) : 0.0)},
       div_up(c.nn, 256));
    CM_LAUNCH_CHECK();
    vcycle<T>(l + 1, dc.b.p, dc.x.p, start_c);
    const int nbg = div_up(v.nng, 256);
    go(ProlongL<T>{v.nx + 1, v.ny + 1, v.nz + 1, c.nx + 1, c.ny + 1, px, py, pz, dc.x.p, v.used.p, x, nbg, ncopy,
                   v.copy_orig.p},
       nbg + div_up(ncopy, 256));
    CM_LAUNCH_CHECK();
    smooth<T>(l, b, x, false, opt.npost);
  }

  // ------------------------------------------------------------------ V-cycle graphs
  // The top-level V-cycle (all levels, ~100 launches) is captured once per precision into a HIP
  // graph and replayed while its right-hand side and result buffers stay the same.  Capture runs
  // on the system's own stream (non-blocking, default priority), entered through OwnStream.
  bool graphs = true;
  hipStream_t own = nullptr;
  struct VGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    const void* b = nullptr;
    void* x = nullptr;
  };
  VGraph vg[2];
  void drop_graph(VGraph& G) {
    if (G.e) (void)hipGraphExecDestroy(G.e);
    if (G.g) (void)hipGraphDestroy(G.g);
    G = VGraph{};
  }
  void drop_graphs() {
    for (auto& G : vg) drop_graph(G);
  }
  template <typename T>
  void vcycle_top(const T* b, T* x) {
    if (!graphs || own == nullptr || stream != own) {
      vcycle<T>(0, b, x);
      return;
    }
    VGraph& G = vg[std::is_same<T, float>::value ? 1 : 0];
    if (!G.e || G.b != (const void*)b || G.x != (void*)x) {
      drop_graph(G);
      CM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        vcycle<T>(0, b, x);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      CM_HIP(hipStreamEndCapture(stream, &G.g));
      CM_HIP(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
      G.b = b;
      G.x = x;
    }
    CM_HIP(hipGraphLaunch(G.e, stream));
  }
  ~System() {
    drop_graphs();
    if (own) (void)hipStreamDestroy(own);
  }

  // ------------------------------------------------------------------ PCG
  template <typename T>
  void finalize_to(int slot) { go(FinalizeL{RED_BLOCKS, partial.p, S.p, slot}, 1); }
  template <typename Ta, typename Tb>
  void dot(long long n, const Ta* a, const Tb* b, int slot) {
    go(DotL<Ta, Tb>{n, a, b, partial.p}, RED_BLOCKS);
    go(FinalizeL{RED_BLOCKS, partial.p, S.p, slot}, 1);
  }
  double read_slot(int slot) {
    double v;
    CM_HIP(hipMemcpyAsync(&v, S.p + slot, sizeof(double), hipMemcpyDeviceToHost, stream));
    CM_HIP(hipStreamSynchronize(stream));
    return v;
  }

  // Recurrence threshold after a failed confirmation that saw the drift (true minus recurrence
  // residual): at most tol minus that drift and half the previous threshold, at least tol / 16.
  // Halving keeps a margin when the drift grows from one restart to the next (stiff joints beside
  // soft tissue, whose attainable true residual lies near tol).
  static double next_target(double target, double thr, double drift) {
    return std::max(thr / 16.0, std::min(thr - drift, 0.5 * target));
  }

  // rigid-mode correction z += Z Einv Z^T r (FP64)
  template <typename T>
  void rigid_correct(const double* r, T* z) {
    if (!zcol) return;
    hipLaunchKernelGGL(k_zt_r, dim3(zchunks), dim3(RED_THREADS), 0, stream, zchk.p, zrows_c.p, zvals_c.p, r, zpart.p);
    hipLaunchKernelGGL(k_rigid_small, dim3(1), dim3(RED_THREADS), 0, stream, zcol, zcchk.p, zpart.p, Einv.p, zw.p, zv.p);
    const int nr = (int)(zrow.n);
    hipLaunchKernelGGL(k_z_v<T>, grid1(nr), dim3(256), 0, stream, nr, zrow.p, zptr.p, zcols.p, zvals.p, zv.p, z);
    CM_LAUNCH_CHECK();
  }

  // z = prec(r), FP64 in/out; vcycle in precision VP
  void precondition64(int vprec, const double* r, double* z) {
    Level& v0 = lv(0);
    const long long n = 3 * v0.nn;
    if (vprec == 0) {
      vcycle_top<double>(r, z);
      rigid_correct<double>(r, z);
    } else {
      hipLaunchKernelGGL((k_cast<float, double>), grid1(n), dim3(256), 0, stream, n, r32.p, r);
      vcycle_top<float>(r32.p, z32.p);
      hipLaunchKernelGGL((k_cast<double, float>), grid1(n), dim3(256), 0, stream, n, z, z32.p);
      rigid_correct<double>(r, z);
    }
  }

  // everything a mode-1 solve needs before its first launch: work vectors, FP32 V-cycle data,
  // FP64 operator data, the partials of the fused dot, the march's block order (the last one reads
  // flags back to the host, so it runs where a wait on the stream is harmless)
  void prepare_solve() {
    alloc_pcg64();
    alloc_pcg32();
    ensure_prec<float>();
    ensure_data<double>();
    if (partial.n < dot_partials(0)) partial.alloc(std::max((size_t)RED_BLOCKS, dot_partials(0)));
    refresh_all_flags<double>();
    refresh_all_flags<float>();
  }

  void alloc_pcg64() {
    const size_t n = 3 * (size_t)lv(0).nn;
    if (r64.n != n) { r64.alloc(n); z64.alloc(n); p64.alloc(n); Ap64.alloc(n); }
  }
  void alloc_pcg32() {
    const size_t n = 3 * (size_t)lv(0).nn;
    if (r32.n != n) { r32.alloc(n); z32.alloc(n); p32.alloc(n); Ap32.alloc(n); d32.alloc(n); }
  }

  // Slots of S
  enum { S_RR = 0, S_BB = 1, S_PAP = 2, S_RZP = 3, S_RZ0 = 4, S_RZ1 = 5, S_IR_RR = 8 };
  // early-exit rule for a solve started from a predicted x0 (modes fp64 / ref): after fb_iters
  // iterations the FP64 true residual must be below 1 / fb_drop of the initial one, else the
  // solve stops with flag FLAG_FALLBACK (the caller then solves plainly); fb_iters = 0: off
  int fb_iters = 0;
  double fb_drop = 10.0;

  // the answer to b = 0: x = 0, converged without an iteration
  void zero_result(double* x, long long n, cm_solve_info* info) {
    hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, stream, x, n, 0.0);
    CM_LAUNCH_CHECK();
    info->iters = 0;
    info->flag = 0;
    info->outer = 0;
    info->relres = 0.0;
  }

  // FP64 PCG (rb/gpusolve.py:229-254), V-cycle in FP64 (vprec 0) or FP32 (vprec 1)
  void pcg64(int vprec, const double* b, double* x, double tol, int maxiter, int check_every, cm_solve_info* info) {
    Level& v0 = lv(0);
    const long long n = 3 * v0.nn;
    alloc_pcg64();
    if (vprec) {   // FP32 V-cycle: FP64 is needed for the operator only
      alloc_pcg32();
      ensure_prec<float>();
      ensure_data<double>();
    } else {
      ensure_prec<double>();
    }
    if (partial.n < dot_partials(0)) partial.alloc(std::max((size_t)RED_BLOCKS, dot_partials(0)));
    apply<double>(0, x, r64.p, MODE_RESID, b);
    dot(n, b, b, S_BB);
    const double bn = std::sqrt(read_slot(S_BB));
    if (bn == 0.0) {   // b = 0: x = 0 whatever x0 is (scipy's cg returns b itself)
      zero_result(x, n, info);
      return;
    }
    const double thr = tol * bn;
    precondition64(vprec, r64.p, z64.p);
    CM_HIP(hipMemcpyAsync(p64.p, z64.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
    int cur = S_RZ0, nxt = S_RZ1;
    dot(n, r64.p, z64.p, cur);
    dot(n, r64.p, r64.p, S_RR);
    double rn = std::sqrt(read_slot(S_RR));
    const double r_initial = rn;   // r64 = b - A x0 exactly here
    int it = 0, restarts = 0, stalls = 0, flag = rn <= thr ? 0 : 1;
    bool flexible = false;
    double last_true = 1e300;
    double target = thr;   // threshold of the recurrence; lowered by the drift a confirmation saw
    while (flag && it < maxiter) {
      apply_dot(0, p64.p, Ap64.p, S_PAP);
      // FP32 V-cycle without rigid-mode correction: the residual's FP32 copy comes out of the
      // update kernel and z stays FP32 (the dots and the direction update read it directly)
      const bool z32_path = vprec == 1 && zcol == 0;
      hipLaunchKernelGGL((k_pcg_xr<double, double, float>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, x,
                         r64.p, p64.p, Ap64.p, S.p, cur, S_PAP, partial.p, z32_path ? r32.p : (float*)nullptr);
      finalize_to<double>(S_RR);
      ++it;
      if (fb_iters > 0 && it == fb_iters) {
        // Ap64 is free until the next product: the true residual b - A x goes there
        apply<double>(0, x, Ap64.p, MODE_RESID, b);
        dot(n, Ap64.p, Ap64.p, S_IR_RR);
        const double rt = std::sqrt(read_slot(S_IR_RR));
        if (rt > thr && rt * fb_drop > r_initial) {
          rn = rt;
          flag = cm::FLAG_FALLBACK;
          break;
        }
      }
      if (it < check_every || it % check_every == 0) {
        rn = std::sqrt(read_slot(S_RR));
        if (rn <= target) {
          // the recurrence says converged: confirm with the FP64 true residual b - A x; on a
          // stiff system (springs 1e6 beside soft tissue) the two drift apart, and then the
          // recurrence restarts from the true residual (residual replacement) and aims lower by
          // the drift it saw, so that the next confirmation has that margin
          const double rec = rn;
          apply<double>(0, x, r64.p, MODE_RESID, b);
          dot(n, r64.p, r64.p, S_RR);
          rn = std::sqrt(read_slot(S_RR));
          if (rn <= thr) { flag = 0; break; }
          target = next_target(target, thr, rn - rec);
          // three restarts in a row that do not halve the true residual: the attainable accuracy
          // of this system (the residual of its solution rounded to FP64) is above tol; stop (not
          // converged) instead of iterating to maxiter.  A system whose floor lies just below tol
          // gets a third sample of it before giving up.
          stalls = (rn > 0.5 * last_true) ? stalls + 1 : 0;
          last_true = rn;
          if (stalls >= 3) break;
          ++restarts;
          flexible = true;
          precondition64(vprec, r64.p, z64.p);
          CM_HIP(hipMemcpyAsync(p64.p, z64.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
          dot(n, r64.p, z64.p, cur);
          continue;
        }
      }
      // after a restart the preconditioner's inexactness (FP32 V-cycle, atomics) has shown up:
      // flexible CG from then on, beta = r.(z - z_old) / (r_old.z_old)
      if (z32_path) {
        if (flexible) dot(n, r64.p, z32.p, S_RZP);   // z32 still holds z_old here
        vcycle_top<float>(r32.p, z32.p);
        dot(n, r64.p, z32.p, nxt);
        hipLaunchKernelGGL((k_pcg_p<double, float>), grid1(n), dim3(256), 0, stream, n, p64.p, z32.p, S.p, nxt, cur,
                           flexible ? (int)S_RZP : -1);
        CM_LAUNCH_CHECK();
        std::swap(cur, nxt);
        continue;
      }
      if (flexible) dot(n, r64.p, z64.p, S_RZP);     // z64 still holds z_old here
      precondition64(vprec, r64.p, z64.p);
      dot(n, r64.p, z64.p, nxt);
      hipLaunchKernelGGL(k_pcg_p<double>, grid1(n), dim3(256), 0, stream, n, p64.p, z64.p, S.p, nxt, cur,
                         flexible ? (int)S_RZP : -1);
      CM_LAUNCH_CHECK();
      std::swap(cur, nxt);
    }
    if (flag > 0) {
      apply<double>(0, x, r64.p, MODE_RESID, b);
      dot(n, r64.p, r64.p, S_RR);
      rn = std::sqrt(read_slot(S_RR));
      flag = rn <= thr ? 0 : (it > 0 ? it : maxiter);
    }
    info->iters = it;
    info->flag = flag;
    info->outer = 1 + restarts;
    info->relres = rn / (bn > 0 ? bn : 1.0);
  }

  // z32 = prec(r32) in FP32 (+ the FP64 rigid-mode correction when present)
  void precondition32(const float* r, float* z) {
    const long long n = 3 * lv(0).nn;
    vcycle_top<float>(r, z);
    if (zcol) {
      hipLaunchKernelGGL((k_cast<double, float>), grid1(n), dim3(256), 0, stream, n, z64.p, z);
      hipLaunchKernelGGL((k_cast<double, float>), grid1(n), dim3(256), 0, stream, n, Ap64.p, r);
      rigid_correct<double>(Ap64.p, z64.p);
      hipLaunchKernelGGL((k_cast<float, double>), grid1(n), dim3(256), 0, stream, n, z, z64.p);
    }
  }

  // Mixed precision as iterative refinement: FP64 outer residual r = b - A64 x; inner FP32 PCG
  // (operator, V-cycle and Krylov vectors in FP32, alpha/beta reductions accumulated in FP64,
  // coarse solve FP64) on A d = r/||r|| until its recurrence residual falls below inner_tol (or
  // the outer target); x += ||r|| d in FP64.  Convergence is decided by the FP64 true residual.
  // With an FP32 operator each outer step contracts the true residual by at most about
  // eps32 * cond(A), so inner tolerances much below that only cost iterations.
  void pcg_mixed(const double* b, double* x, double tol, int maxiter, int check_every, double inner_tol,
                 cm_solve_info* info) {
    Level& v0 = lv(0);
    const long long n = 3 * v0.nn;
    alloc_pcg64();
    alloc_pcg32();
    ensure_prec<float>();
    ensure_data<double>();
    dot(n, b, b, S_BB);
    const double bn = std::sqrt(read_slot(S_BB));
    if (bn == 0.0) {
      zero_result(x, n, info);
      return;
    }
    const double thr = tol * bn;
    int it = 0, outer = 0, flag = 1, stalls = 0;
    double rn = 0.0, last_true = 1e300;
    while (true) {
      apply<double>(0, x, Ap64.p, MODE_APPLY, nullptr);
      hipLaunchKernelGGL(k_resid_cast<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, r64.p, b, Ap64.p,
                         partial.p, (float*)nullptr, 1.0);
      finalize_to<double>(S_IR_RR);
      rn = std::sqrt(read_slot(S_IR_RR));
      if (rn <= thr) { flag = 0; break; }
      if (it >= maxiter) { flag = maxiter; break; }
      // three refinement steps in a row that do not halve the FP64 true residual: the FP32 inner
      // solve no longer contracts it (eps32 * cond(A) near 1 on stiff joints, or the FP64 floor of
      // the system lies above tol); stop, not converged, instead of iterating to maxiter
      stalls = (rn > 0.5 * last_true) ? stalls + 1 : 0;
      last_true = rn;
      if (stalls >= 3) { flag = it > 0 ? it : maxiter; break; }
      ++outer;
      const double sc = 1.0 / rn;
      hipLaunchKernelGGL(k_resid_cast<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, r64.p, b, Ap64.p,
                         partial.p, r32.p, sc);
      hipLaunchKernelGGL(k_fill<float>, grid1(n), dim3(256), 0, stream, d32.p, n, 0.f);
      precondition32(r32.p, z32.p);
      CM_HIP(hipMemcpyAsync(p32.p, z32.p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
      int cur = S_RZ0, nxt = S_RZ1;
      dot(n, r32.p, z32.p, cur);
      int inner = 0;
      const double ithr = std::max(inner_tol, thr / rn);
      while (it < maxiter) {
        apply<float>(0, p32.p, Ap32.p, MODE_APPLY, nullptr);
        dot(n, p32.p, Ap32.p, S_PAP);
        hipLaunchKernelGGL((k_pcg_xr<float, float, float>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, n, d32.p,
                           r32.p, p32.p, Ap32.p, S.p, cur, S_PAP, partial.p, (float*)nullptr);
        finalize_to<double>(S_RR);
        ++it; ++inner;
        if (inner < check_every || inner % check_every == 0) {
          if (std::sqrt(read_slot(S_RR)) <= ithr) break;
        }
        precondition32(r32.p, z32.p);
        dot(n, r32.p, z32.p, nxt);
        hipLaunchKernelGGL(k_pcg_p<float>, grid1(n), dim3(256), 0, stream, n, p32.p, z32.p, S.p, nxt, cur);
        CM_LAUNCH_CHECK();
        std::swap(cur, nxt);
      }
      hipLaunchKernelGGL(k_axpy_scaled<float>, grid1(n), dim3(256), 0, stream, n, x, d32.p, rn);
      CM_LAUNCH_CHECK();
    }
    info->iters = it;
    info->flag = flag;
    info->outer = outer;
    info->relres = rn / (bn > 0 ? bn : 1.0);
  }
};

// ---------------------------------------------------------------------------------------------
// Level-0 description from the boundary
static void set_fine_common(System& s, int nx, int ny, int nz, const double* xe, const double* ye, const double* ze) {
  auto lv = std::make_unique<Level>();
  lv->nx = nx; lv->ny = ny; lv->nz = nz;
  lv->e[0].assign(xe, xe + nx + 1);
  lv->e[1].assign(ye, ye + ny + 1);
  lv->e[2].assign(ze, ze + nz + 1);
  lv->nng = lv->nn = (long long)(nx + 1) * (ny + 1) * (nz + 1);
  lv->ncell = (long long)nx * ny * nz;
  s.L.clear();
  s.L.push_back(std::move(lv));
}

}  // namespace cm


using cm::System;

#define CM_API_BEGIN try {
#define CM_API_END                                        \
  return 0;                                               \
  }                                                       \
  catch (const std::exception& e) {                       \
    cm::g_last_error = e.what();                          \
    return -1;                                            \
  }                                                       \
  catch (...) {                                           \
    cm::g_last_error = "comech: unknown error";           \
    return -1;                                            \
  }

namespace cm {
static int* scratch_on_current_device() {
  int dev = 0;
  CM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(scratch_mu);
  if ((int)scratch_dev.size() <= dev) scratch_dev.resize(dev + 1, nullptr);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    CM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CM_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_dev[dev] = static_cast<int*>(p);
  }
  return scratch_dev[dev];
}
}  // namespace cm

struct cm_sys {
  System s;
};

namespace {
// Runs an entry point on the system's own stream, ordered after the work already queued on the
// caller's stream and before whatever the caller queues next (events, no host wait).
struct OwnStream {
  System& s;
  hipStream_t user;
  hipEvent_t ev = nullptr;
  explicit OwnStream(System& sys) : s(sys), user(sys.stream) {
    if (!s.own) CM_HIP(hipStreamCreateWithFlags(&s.own, hipStreamNonBlocking));
    CM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    CM_HIP(hipEventRecord(ev, user));
    CM_HIP(hipStreamWaitEvent(s.own, ev, 0));
    s.stream = s.own;
  }
  ~OwnStream() {
    (void)hipEventRecord(ev, s.own);
    (void)hipStreamWaitEvent(user, ev, 0);
    (void)hipEventDestroy(ev);
    s.stream = user;
  }
};
template <typename T>
const T* cp(const void* p) { return static_cast<const T*>(p); }
template <typename T>
T* mp(void* p) { return static_cast<T*>(p); }
void need_built(cm_sys* h) {
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(h->s.built, "system not built (call cm_sys_build)");
}
void need_level(cm_sys* h, int level) {
  need_built(h);
  CM_CHECK(level >= 0 && level < h->s.nlev(), "level out of range");
}
}  // namespace

// ---------------------------------------------------------------------------------------------
// Lock-step batched PCG over several systems on one stream (mode "ref": FP64 Krylov vectors,
// alpha/beta reductions and operator, FP32 V-cycle; the same iteration as System::pcg64).  Every
// launch of an iteration -- the FP64 product with its dot, the vector updates, every kernel of
// the V-cycle on every level -- covers all systems of the batch: their launch records
// (System::rec) are laid out side by side and run as one grid (k_multi).  Systems join and leave
// between iterations (continuous batching).  The recurrence residual is checked on the host every
// check_every iterations of the batch, and a system whose recurrence has converged is confirmed,
// restarted or stopped by its FP64 true residual exactly as pcg64 does.
struct cm_batch {
  hipStream_t st = nullptr;
  int check_every = 5;
  long long iter = 0;                 // iterations of the batch
  struct Slot {
    cm_sys* h = nullptr;
    hipStream_t user = nullptr;       // the system's stream, restored when it leaves
    const double* b = nullptr;
    double* x = nullptr;
    double tol = 0, thr = 0, bn = 0, rn = 0, target = 0, last_true = 1e300;
    int it = 0, maxiter = 0, restarts = 0, stalls = 0;
    bool flexible = false, live = false, done = false;
    std::vector<cm::LaunchRec> seq;   // one iteration's launches
    cm_solve_info info{};
    std::chrono::steady_clock::time_point t0;
  };
  std::vector<Slot> slots;
  std::vector<cm::MultiFn> sig;       // record types of one iteration (the batch's first system)
  // device layout of the live slots' records
  bool plan_ok = false;
  std::vector<int> live;
  std::vector<char> h_args;
  std::vector<int> h_start;
  cm::DevBuf<char> d_args;
  cm::DevBuf<int> d_start;
  std::vector<size_t> aoff, soff;
  std::vector<long long> tot;
  cm::DevBuf<double*> d_sptr;
  cm::DevBuf<double> d_rr;
  std::vector<double> h_rr;

  ~cm_batch() {
    if (st) (void)hipStreamSynchronize(st);
  }

  static long long n3(System& s) { return 3 * s.lv(0).nn; }

  // One iteration, preconditioner first: z = V(r) in FP32, r.z (and r.z_old, the flexible form),
  // p = z + beta p, rz_old = rz; then Ap = A p with p.Ap, x += alpha p, r -= alpha Ap, r32 = r and
  // ||r||^2.  A system (re)starts with p = 0 and rz_old = +inf (beta = 0: p = z), so joining or
  // restarting needs no V-cycle of its own and no host wait.
  void record(Slot& sl) {
    System& s = sl.h->s;
    const long long n = n3(s);
    sl.seq.clear();
    s.rec = &sl.seq;
    try {
      s.dot(n, s.r64.p, s.z32.p, System::S_RZP);
      s.vcycle<float>(0, s.r32.p, s.z32.p);
      s.dot(n, s.r64.p, s.z32.p, System::S_RZ1);
      s.go(cm::PcgPL<double, float>{n, s.p64.p, s.z32.p, s.S.p, System::S_RZ1, System::S_RZ0,
                                    sl.flexible ? (int)System::S_RZP : -1},
           cm::div_up(n, 256));
      s.go(cm::SlotCopyL{s.S.p, System::S_RZ0, System::S_RZ1}, 1);
      s.apply_dot(0, s.p64.p, s.Ap64.p, System::S_PAP);
      s.go(cm::PcgXrL<double, double, float>{n, sl.x, s.r64.p, s.p64.p, s.Ap64.p, s.S.p, System::S_RZ0,
                                             System::S_PAP, s.partial.p, s.r32.p},
           cm::RED_BLOCKS);
      s.finalize_to<double>(System::S_RR);
    } catch (...) {
      s.rec = nullptr;
      throw;
    }
    s.rec = nullptr;
    plan_ok = false;
  }

  bool same_sig(const Slot& sl) const {
    if (sig.size() != sl.seq.size()) return false;
    for (size_t j = 0; j < sig.size(); ++j)
      if (sig[j] != sl.seq[j].multi) return false;
    return true;
  }

  void plan() {
    live.clear();
    for (int k = 0; k < (int)slots.size(); ++k)
      if (slots[k].live) live.push_back(k);
    plan_ok = true;
    if (live.empty()) return;
    const int K = (int)live.size();
    const size_t nrec = sig.size();
    aoff.assign(nrec, 0);
    soff.assign(nrec, 0);
    tot.assign(nrec, 0);
    size_t asz = 0;
    for (size_t j = 0; j < nrec; ++j) {
      aoff[j] = asz;
      asz += ((slots[live[0]].seq[j].size * K + 255) / 256) * 256;
    }
    h_args.assign(asz, 0);
    h_start.assign(nrec * (K + 1), 0);
    for (size_t j = 0; j < nrec; ++j) {
      soff[j] = j * (K + 1);
      long long acc = 0;
      for (int k = 0; k < K; ++k) {
        const cm::LaunchRec& r = slots[live[k]].seq[j];
        std::memcpy(h_args.data() + aoff[j] + k * r.size, r.bytes.data(), r.size);
        h_start[soff[j] + k] = (int)acc;
        acc += std::max(0LL, r.nb);
      }
      CM_CHECK(acc < (1LL << 31), "batched launch too large");
      h_start[soff[j] + K] = (int)acc;
      tot[j] = acc;
    }
    if (d_args.n < asz) d_args.alloc(asz);
    if (d_start.n < h_start.size()) d_start.alloc(h_start.size());
    CM_HIP(hipMemcpyAsync(d_args.p, h_args.data(), asz, hipMemcpyHostToDevice, st));
    CM_HIP(hipMemcpyAsync(d_start.p, h_start.data(), h_start.size() * sizeof(int), hipMemcpyHostToDevice, st));
    std::vector<double*> sp(K);
    for (int k = 0; k < K; ++k) sp[k] = slots[live[k]].h->s.S.p;
    if (d_sptr.n < (size_t)K) d_sptr.alloc(std::max(K, 64));
    if (d_rr.n < (size_t)(2 * K)) d_rr.alloc(std::max(2 * K, 128));
    CM_HIP(hipMemcpyAsync(d_sptr.p, sp.data(), K * sizeof(double*), hipMemcpyHostToDevice, st));
    CM_HIP(hipStreamSynchronize(st));   // the host copies above must outlive the transfers
  }

  void replay() {
    const int K = (int)live.size();
    for (size_t j = 0; j < sig.size(); ++j)
      if (tot[j] > 0) sig[j](d_args.p + aoff[j], d_start.p + soff[j], K, (int)tot[j], st);
    CM_LAUNCH_CHECK();
  }

  void finish(Slot& sl, int flag) {
    System& s = sl.h->s;
    sl.info.iters = sl.it;
    sl.info.flag = flag;
    sl.info.outer = 1 + sl.restarts;
    sl.info.relres = sl.rn / (sl.bn > 0 ? sl.bn : 1.0);
    sl.info.levels = s.nlev();
    sl.info.t_setup_ms = s.t_setup_ms;
    sl.info.t_solve_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - sl.t0).count();
    // hand the system back to its own stream after the batch's work on it
    hipEvent_t ev;
    CM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    CM_HIP(hipEventRecord(ev, st));
    CM_HIP(hipStreamWaitEvent(sl.user, ev, 0));
    CM_HIP(hipEventDestroy(ev));
    s.stream = sl.user;
    sl.live = false;
    sl.done = true;
    sl.seq.clear();
    plan_ok = false;
  }

  // r32 = r64, p = 0, rz_old = +inf: the next iteration starts the recurrence from r64
  void restart_from_r(Slot& sl) {
    System& s = sl.h->s;
    const long long n = n3(s);
    s.go(cm::CastL<float, double>{n, s.r32.p, s.r64.p}, cm::div_up(n, 256));
    s.go(cm::FillL<double>{s.p64.p, n, 0.0}, cm::div_up(n, 256));
    s.go(cm::SlotSetL{s.S.p, System::S_RZ0, INFINITY}, 1);
  }

  // true residual of the slot's x (FP64), into rn
  double true_res(Slot& sl) {
    System& s = sl.h->s;
    s.apply<double>(0, sl.x, s.r64.p, cm::MODE_RESID, sl.b);
    s.dot(n3(s), s.r64.p, s.r64.p, System::S_RR);
    return std::sqrt(s.read_slot(System::S_RR));
  }

  int add(cm_sys* h, const double* b, double* x, double tol, int maxiter) {
    System& s = h->s;
    CM_CHECK(s.built, "system not built (call cm_sys_build)");
    if (s.zcol != 0) return -1;   // a rigid-mode correction: solved on its own
    for (auto& o : slots)
      CM_CHECK(!(o.live && o.h == h), "a system can be in the batch once at a time");
    int k = 0;   // a free slot: not live and its result taken
    while (k < (int)slots.size() && (slots[k].live || slots[k].h != nullptr)) ++k;
    if (k == (int)slots.size()) slots.emplace_back();
    Slot& sl = slots[k];
    sl = Slot{};
    sl.h = h;
    sl.b = b;
    sl.x = x;
    sl.maxiter = maxiter;
    sl.t0 = std::chrono::steady_clock::now();
    // the system's queued work (its build) before the batch's
    sl.user = s.stream;
    hipEvent_t ev;
    CM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    CM_HIP(hipEventRecord(ev, sl.user));
    CM_HIP(hipStreamWaitEvent(st, ev, 0));
    CM_HIP(hipEventDestroy(ev));
    s.stream = st;
    try {
      s.prepare_solve();   // once per system (cm_sys_prepare_solve does it on the builder's stream)
      const long long n = n3(s);
      // r = b - A x; ||b|| is read at the batch's next check (tol and thr then)
      s.apply<double>(0, x, s.r64.p, cm::MODE_RESID, b);
      s.dot(n, b, b, System::S_BB);
      sl.tol = tol;
      sl.thr = -1.0;
      restart_from_r(sl);
      if (maxiter <= 0) {
        sl.bn = std::sqrt(s.read_slot(System::S_BB));
        sl.rn = true_res(sl);
        finish(sl, maxiter);
        return k;
      }
      record(sl);
      if (sig.empty() || std::none_of(slots.begin(), slots.end(), [](const Slot& o) { return o.live; })) {
        sig.clear();
        for (auto& r : sl.seq) sig.push_back(r.multi);
      } else if (!same_sig(sl)) {
        sl.seq.clear();
        s.stream = sl.user;
        sl = Slot{};
        sl.done = false;
        sl.h = nullptr;
        return -1;   // another launch structure (level count, springs): solve it on its own
      }
      sl.live = true;
      plan_ok = false;
    } catch (...) {
      s.stream = sl.user;
      s.rec = nullptr;
      sl = Slot{};
      throw;
    }
    return k;
  }

  // host decisions of pcg64 at a check, for one live slot
  void check(Slot& sl, double rr, double bb) {
    if (sl.thr < 0) {   // the first check of this solve: ||b|| and the threshold
      sl.bn = std::sqrt(bb);
      sl.thr = sl.tol * (sl.bn > 0 ? sl.bn : 1.0);
      sl.target = sl.thr;
    }
    sl.rn = std::sqrt(rr);
    if (sl.rn <= sl.target) {
      const double rec = sl.rn;
      sl.rn = true_res(sl);
      if (sl.rn <= sl.thr) {
        finish(sl, 0);
        return;
      }
      sl.target = System::next_target(sl.target, sl.thr, sl.rn - rec);
      sl.stalls = (sl.rn > 0.5 * sl.last_true) ? sl.stalls + 1 : 0;
      sl.last_true = sl.rn;
      if (sl.stalls >= 3) {
        finish(sl, sl.it > 0 ? sl.it : sl.maxiter);
        return;
      }
      ++sl.restarts;
      restart_from_r(sl);   // r64 holds the true residual
      if (!sl.flexible) {
        sl.flexible = true;
        record(sl);
      }
      return;
    }
    if (sl.it >= sl.maxiter) {
      sl.rn = true_res(sl);
      finish(sl, sl.rn <= sl.thr ? 0 : sl.it);
    }
  }

  // iterations until a slot has finished or max_iters ran; returns the live slots left
  int step(int max_iters) {
    int ran = 0;
    while (ran < max_iters) {
      if (!plan_ok) plan();
      if (live.empty()) break;
      replay();
      for (int k : live) ++slots[k].it;
      ++iter;
      ++ran;
      if (iter % check_every) continue;
      const int K = (int)live.size();
      hipLaunchKernelGGL(cm::k_gather_slot, dim3(1), dim3(256), 0, st, K, d_sptr.p, (int)System::S_RR, d_rr.p);
      hipLaunchKernelGGL(cm::k_gather_slot, dim3(1), dim3(256), 0, st, K, d_sptr.p, (int)System::S_BB, d_rr.p + K);
      h_rr.resize(2 * K);
      CM_HIP(hipMemcpyAsync(h_rr.data(), d_rr.p, 2 * K * sizeof(double), hipMemcpyDeviceToHost, st));
      CM_HIP(hipStreamSynchronize(st));
      const std::vector<int> now = live;
      bool any_done = false;
      for (int q = 0; q < K; ++q) {
        Slot& sl = slots[now[q]];
        check(sl, h_rr[q], h_rr[K + q]);
        any_done |= !sl.live;
      }
      if (any_done) break;
    }
    int n = 0;
    for (auto& sl : slots) n += sl.live;
    return n;
  }

  // a finished slot's result (slot -1: none finished)
  void poll(int* slot, cm_solve_info* info) {
    *slot = -1;
    for (int k = 0; k < (int)slots.size(); ++k) {
      Slot& sl = slots[k];
      if (sl.done && sl.h) {
        *slot = k;
        *info = sl.info;
        sl.h = nullptr;
        return;
      }
    }
  }
};

extern "C" {

const char* cm_last_error(void) { return cm::g_last_error.c_str(); }

int cm_version(void) { return CM_ABI_VERSION; }

int cm_device_count(int* n) {
  CM_API_BEGIN
  CM_HIP(hipGetDeviceCount(n));
  CM_API_END
}

int cm_sys_create(cm_sys** out, void* stream, int nx, int ny, int nz, const double* xe, const double* ye,
                  const double* ze, int sri) {
  CM_API_BEGIN
  CM_CHECK(out != nullptr, "out is null");
  CM_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid must have at least one cell per axis");
  for (int q = 0; q < nx; ++q) CM_CHECK(xe[q + 1] > xe[q], "xe must increase");
  for (int q = 0; q < ny; ++q) CM_CHECK(ye[q + 1] > ye[q], "ye must increase");
  for (int q = 0; q < nz; ++q) CM_CHECK(ze[q + 1] > ze[q], "ze must increase");
  CM_CHECK((long long)(nx + 1) * (ny + 1) * (nz + 1) * 3 < (1LL << 31), "grid too large for 32-bit DOF indices");
  auto* h = new cm_sys();
  h->s.stream = (hipStream_t)stream;
  h->s.sri = sri ? 1 : 0;
  cm::set_fine_common(h->s, nx, ny, nz, xe, ye, ze);
  *out = h;
  CM_API_END
}

int cm_sys_destroy(cm_sys* h) {
  CM_API_BEGIN
  if (h) {
    // nothing of this system is in flight after these waits, so its buffers may be reused
    const bool idle = hipStreamSynchronize(h->s.stream) == hipSuccess &&
                      (h->s.own == nullptr || hipStreamSynchronize(h->s.own) == hipSuccess);
    cm::release_to_cache() = idle;
    delete h;
    cm::release_to_cache() = false;
  }
  CM_API_END
}

int cm_sys_set_labels(cm_sys* h, int nmat, const double* lam_tab, const double* mu_tab, const uint8_t* lab_ref) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  CM_CHECK(nmat >= 1 && nmat <= 255, "label mode holds 1..255 materials");
  System& s = h->s;
  const long long nc = s.lv(0).ncell;
  s.st_nmat = nmat;
  s.st_tabL.assign(lam_tab, lam_tab + nmat);
  s.st_tabM.assign(mu_tab, mu_tab + nmat);
  s.st_lab.assign(lab_ref, lab_ref + nc);
  s.st_lam.clear(); s.st_mu.clear(); s.st_active.clear();
  CM_API_END
}

int cm_sys_set_cellmat(cm_sys* h, const double* lam_ref, const double* mu_ref, const uint8_t* active_ref) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  System& s = h->s;
  const long long nc = s.lv(0).ncell;
  s.st_nmat = -1;
  s.st_lam.assign(lam_ref, lam_ref + nc);
  s.st_mu.assign(mu_ref, mu_ref + nc);
  s.st_active.assign(active_ref, active_ref + nc);
  CM_API_END
}

int cm_sys_set_split(cm_sys* h, int64_t ncopy, const int64_t* copy_orig, int64_t nirr, const int64_t* irr_cell_ref,
                     const int64_t* irr_nodes) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  CM_CHECK(ncopy >= 0 && nirr >= 0, "negative counts");
  System& s = h->s;
  s.st_copy_orig.assign(copy_orig, copy_orig + ncopy);
  s.st_irr_cell.assign(irr_cell_ref, irr_cell_ref + nirr);
  s.st_irr_nodes.assign(irr_nodes, irr_nodes + 8 * nirr);
  CM_API_END
}

int cm_sys_set_extra(cm_sys* h, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  CM_CHECK(nnz >= 0, "negative nnz");
  System& s = h->s;
  s.st_xr.assign(rows, rows + nnz);
  s.st_xc.assign(cols, cols + nnz);
  s.st_xv.assign(vals, vals + nnz);
  CM_API_END
}

int cm_sys_set_rank1(cm_sys* h, int64_t nterm, const int64_t* ptr, const int64_t* dofs, const double* w,
                     const double* k) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  CM_CHECK(nterm >= 0, "negative term count");
  System& s = h->s;
  s.st_r1ptr.assign(ptr, ptr + nterm + 1);
  CM_CHECK(s.st_r1ptr[0] == 0, "rank-one ptr must start at 0");
  for (int64_t t = 0; t < nterm; ++t) CM_CHECK(s.st_r1ptr[t + 1] >= s.st_r1ptr[t], "rank-one ptr must not decrease");
  const int64_t nnz = s.st_r1ptr[nterm];
  s.st_r1dof.assign(dofs, dofs + nnz);
  s.st_r1w.assign(w, w + nnz);
  s.st_r1k.assign(k, k + nterm);
  CM_API_END
}

int cm_sys_set_used(cm_sys* h, int64_t nn, const uint8_t* used) {
  CM_API_BEGIN
  CM_CHECK(h && !h->s.built, "system missing or already built");
  h->s.st_used_extra.assign(used, used + nn);
  CM_API_END
}

int cm_sys_set_rigid(cm_sys* h, int ncol, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                     const double* Einv) {
  CM_API_BEGIN
  need_built(h);
  System& s = h->s;
  const long long nd3 = 3 * s.lv(0).nn;
  s.zcol = 0;
  if (ncol <= 0 || nnz <= 0) return 0;
  // CSC for Z^T r, unique-row CSR for Z v
  std::vector<int> cptr(ncol + 1, 0), crow(nnz);
  std::vector<double> cval(nnz);
  for (long long q = 0; q < nnz; ++q) {
    CM_CHECK(cols[q] >= 0 && cols[q] < ncol && rows[q] >= 0 && rows[q] < nd3, "rigid-mode entry out of range");
    cptr[cols[q] + 1]++;
  }
  for (int j = 0; j < ncol; ++j) cptr[j + 1] += cptr[j];
  std::vector<int> fill(cptr.begin(), cptr.end() - 1);
  for (long long q = 0; q < nnz; ++q) {
    const int pos = fill[cols[q]]++;
    crow[pos] = (int)rows[q];
    cval[pos] = vals[q];
  }
  std::vector<long long> ord(nnz);
  std::iota(ord.begin(), ord.end(), 0);
  auto before = [&](long long a, long long b) { return rows[a] < rows[b] || (rows[a] == rows[b] && cols[a] < cols[b]); };
  bool sorted = true;   // callers that pass (row, column) order skip the sort
  for (long long q = 1; q < nnz && sorted; ++q) sorted = !before(q, q - 1);
  if (!sorted) std::sort(ord.begin(), ord.end(), before);
  std::vector<int> rrow, rptr, rcol;
  std::vector<double> rval;
  long long last = -1;
  for (long long q : ord) {
    if (rows[q] != last) { rrow.push_back((int)rows[q]); rptr.push_back((int)rcol.size()); last = rows[q]; }
    rcol.push_back((int)cols[q]);
    rval.push_back(vals[q]);
  }
  rptr.push_back((int)rcol.size());
  // chunks of at most ZT_CHUNK entries inside each column
  std::vector<int> chk{0}, cchk{0};
  for (int j = 0; j < ncol; ++j) {
    for (int q = cptr[j]; q < cptr[j + 1];) {
      q = std::min(q + cm::ZT_CHUNK, cptr[j + 1]);
      chk.push_back(q);
    }
    cchk.push_back((int)chk.size() - 1);
  }
  s.zchk.upload(chk, s.stream); s.zcchk.upload(cchk, s.stream);
  s.zchunks = (int)chk.size() - 1;
  s.zpart.alloc(std::max(1, s.zchunks));
  s.zrows_c.upload(crow, s.stream); s.zvals_c.upload(cval, s.stream);
  s.zrow.upload(rrow, s.stream); s.zptr.upload(rptr, s.stream); s.zcols.upload(rcol, s.stream);
  s.zvals.upload(rval, s.stream);
  std::vector<double> E(Einv, Einv + (size_t)ncol * ncol);
  s.Einv.upload(E, s.stream);
  s.zw.alloc(ncol); s.zv.alloc(ncol);
  s.zcol = ncol;
  CM_API_END
}

int cm_sys_build(cm_sys* h, int min_dofs, int max_levels, int fixed_levels, int smoother, int degree, double omega,
                 int npre, int npost) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(smoother == 0 || smoother == 1, "smoother must be 0 (Chebyshev) or 1 (Jacobi)");
  CM_CHECK(max_levels >= 1 && degree >= 1 && npre >= 1 && npost >= 1, "bad multigrid options");
  cm::Options o;
  o.smoother = smoother; o.degree = degree; o.omega = omega; o.npre = npre; o.npost = npost;
  h->s.build(min_dofs, max_levels, fixed_levels, o);
  CM_API_END
}

int cm_sys_levels(cm_sys* h, int* nlev, int64_t* desc, int cap) {
  CM_API_BEGIN
  need_built(h);
  System& s = h->s;
  *nlev = s.nlev();
  for (int l = 0; l < s.nlev() && l < cap; ++l) {
    cm::Level& v = s.lv(l);
    desc[6 * l] = v.nx; desc[6 * l + 1] = v.ny; desc[6 * l + 2] = v.nz;
    desc[6 * l + 3] = v.nn; desc[6 * l + 4] = s.compact_dofs(l); desc[6 * l + 5] = (int64_t)(v.xrow_h.size() + v.r1k_h.size());
  }
  CM_API_END
}

int cm_sys_level_stats(cm_sys* h, int level, int64_t* st, int cap) {
  CM_API_BEGIN
  need_level(h, level);
  cm::Level& v = h->s.lv(level);
  const System& s = h->s;
  const int64_t vals[15] = {v.ncell, v.nirr, (int64_t)v.xrow_h.size(), (int64_t)v.xcol_h.size(),
                            (int64_t)(v.r1ptr_h.empty() ? 0 : v.r1ptr_h.size() - 1), (int64_t)v.r1dof_h.size(),
                            v.nblk, v.blk_max, v.nrn, v.rn_max,
                            s.pick_kz<double>(v, 0), s.pick_kz<double>(v, 1), s.pick_kz<float>(v, 0),
                            s.pick_kz<float>(v, 1), v.nhet};
  for (int q = 0; q < 15 && q < cap; ++q) st[q] = vals[q];
  CM_API_END
}

int cm_sys_apply(cm_sys* h, int level, int prec, const void* u, void* y) {
  CM_API_BEGIN
  need_level(h, level);
  System& s = h->s;
  if (prec == 0) {
    s.prepare_prec<double>(level, false);
    s.apply<double>(level, cp<double>(u), mp<double>(y), cm::MODE_APPLY, nullptr);
  } else {
    s.prepare_prec<float>(level, false);
    s.apply<float>(level, cp<float>(u), mp<float>(y), cm::MODE_APPLY, nullptr);
  }
  CM_API_END
}

int cm_sys_vcycle(cm_sys* h, int prec, const void* b, void* x) {
  CM_API_BEGIN
  need_built(h);
  System& s = h->s;
  if (prec == 0) s.ensure_prec<double>();
  else s.ensure_prec<float>();
  OwnStream os(s);
  if (prec == 0) s.vcycle_top<double>(cp<double>(b), mp<double>(x));
  else s.vcycle_top<float>(cp<float>(b), mp<float>(x));
  CM_API_END
}

int cm_sys_set_fallback(cm_sys* h, int iters, double drop) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(iters >= 0 && drop >= 1.0, "fallback: iters >= 0 and drop >= 1");
  h->s.fb_iters = iters;
  h->s.fb_drop = drop;
  CM_API_END
}

int cm_sys_level_data(cm_sys* h, int level, int prec, int what, void* out) {
  CM_API_BEGIN
  need_level(h, level);
  CM_CHECK(out != nullptr, "out is null");
  System& s = h->s;
  if (what == 2) {
    const cm::Level& v = s.lv(level);
    std::memcpy(out, v.used_h.data(), v.used_h.size());
    return 0;
  }
  CM_CHECK(level + 1 < s.nlev(), "the coarsest level has no smoother");
  if (prec == 0) s.prepare_prec<double>(level, true);
  else s.prepare_prec<float>(level, true);
  cm::Level& v = s.lv(level);
  if (what == 0) {
    double* o = static_cast<double*>(out);
    o[0] = prec == 0 ? v.d64.lmax : v.d32.lmax;
    o[1] = prec == 0 ? v.d64.lmin : v.d32.lmin;
  } else if (what == 1) {
    const size_t bytes = 6 * (size_t)v.nn * (prec == 0 ? sizeof(double) : sizeof(float));
    const void* src = prec == 0 ? (const void*)v.d64.Dinv.p : (const void*)v.d32.Dinv.p;
    CM_HIP(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToDevice, s.stream));
    CM_HIP(hipStreamSynchronize(s.stream));
  } else {
    CM_CHECK(false, "what must be 0 (Chebyshev bounds), 1 (block-Jacobi inverses) or 2 (used nodes)");
  }
  CM_API_END
}

int cm_sys_power(cm_sys* h, int level, int prec, const void* x0, double* lam) {
  CM_API_BEGIN
  need_level(h, level);
  CM_CHECK(x0 != nullptr && lam != nullptr, "x0 and lam must not be null");
  System& s = h->s;
  CM_CHECK(level + 1 < s.nlev(), "the coarsest level has no smoother");
  if (prec == 0) {
    s.prepare_prec<double>(level, true);
    *lam = s.power_iteration<double>(level, cp<double>(x0));
  } else {
    s.prepare_prec<float>(level, true);
    *lam = s.power_iteration<float>(level, cp<float>(x0));
  }
  s.drop_graphs();   // the captured V-cycle holds the old Chebyshev coefficients
  CM_API_END
}

int cm_sys_smooth(cm_sys* h, int level, int prec, const void* b, void* x, int zero_start) {
  CM_API_BEGIN
  need_level(h, level);
  System& s = h->s;
  CM_CHECK(level + 1 < s.nlev(), "the coarsest level has no smoother");
  if (prec == 0) {
    s.prepare_prec<double>(level, true);
    s.smooth<double>(level, cp<double>(b), mp<double>(x), zero_start != 0, s.opt.npre);
  } else {
    s.prepare_prec<float>(level, true);
    s.smooth<float>(level, cp<float>(b), mp<float>(x), zero_start != 0, s.opt.npre);
  }
  CM_API_END
}

int cm_sys_solve(cm_sys* h, int mode, const double* b, double* x, double tol, int maxiter, int check_every,
                 double inner_tol, cm_solve_info* info) {
  CM_API_BEGIN
  need_built(h);
  CM_CHECK(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
  CM_CHECK(info != nullptr, "info is null");
  System& s = h->s;
  if (check_every < 1) check_every = 1;
  OwnStream os(s);
  hipEvent_t e0, e1;
  CM_HIP(hipEventCreate(&e0)); CM_HIP(hipEventCreate(&e1));
  CM_HIP(hipEventRecord(e0, s.stream));
  std::memset(info, 0, sizeof(*info));
  if (mode == 2) s.pcg_mixed(b, x, tol, maxiter, check_every, inner_tol, info);
  else s.pcg64(mode, b, x, tol, maxiter, check_every, info);
  CM_HIP(hipEventRecord(e1, s.stream));
  CM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CM_HIP(hipEventElapsedTime(&ms, e0, e1));
  CM_HIP(hipEventDestroy(e0)); CM_HIP(hipEventDestroy(e1));
  info->t_solve_ms = ms;
  info->t_setup_ms = s.t_setup_ms;
  info->levels = s.nlev();
  CM_API_END
}

int cm_cell_stress(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                   int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* lam_ref,
                   const double* mu_ref, const double* U, double* sig, double* vm, double* eps) {
  CM_API_BEGIN
  CM_CHECK(nx > 0 && ny > 0 && nz > 0, "bad grid");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(cm::k_cell_stress, cm::grid1(n), dim3(256), 0, (hipStream_t)stream, (long long)n,
                     (const long long*)cells_ref, (const long long*)nodes8, nx, ny, nz, hx, hy, hz, lam_ref, mu_ref, U,
                     sig, vm, eps);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_cell_stress_ortho(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                         int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* coeffs_ref,
                         const double* U, double* sig, double* vm, double* eps) {
  CM_API_BEGIN
  CM_CHECK(nx > 0 && ny > 0 && nz > 0, "bad grid");
  CM_CHECK(coeffs_ref != nullptr, "coeffs must not be null");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(cm::k_cell_stress, cm::grid1(n), dim3(256), 0, (hipStream_t)stream, (long long)n,
                     (const long long*)cells_ref, (const long long*)nodes8, nx, ny, nz, hx, hy, hz, (const double*)nullptr,
                     (const double*)nullptr, U, sig, vm, eps, coeffs_ref);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_element_forces(void* stream, int nx, int ny, int nz, const double* hx, const double* hy, const double* hz,
                      int sri, int64_t n, const int64_t* cells_ref, const int64_t* nodes8, const double* lam_ref,
                      const double* mu_ref, const double* U, double* out) {
  CM_API_BEGIN
  CM_CHECK(nx > 0 && ny > 0 && nz > 0, "bad grid");
  if (n <= 0) return 0;
  if (sri)
    hipLaunchKernelGGL(cm::k_element_forces<true>, cm::grid1(n), dim3(256), 0, (hipStream_t)stream, (long long)n,
                       (const long long*)cells_ref, (const long long*)nodes8, nx, ny, nz, hx, hy, hz, lam_ref, mu_ref,
                       U, out);
  else
    hipLaunchKernelGGL(cm::k_element_forces<false>, cm::grid1(n), dim3(256), 0, (hipStream_t)stream, (long long)n,
                       (const long long*)cells_ref, (const long long*)nodes8, nx, ny, nz, hx, hy, hz, lam_ref, mu_ref,
                       U, out);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_sys_time_apply(cm_sys* h, int level, int prec, const void* u, void* y, int nrep, int structured_only,
                      double* avg_ms) {
  CM_API_BEGIN
  need_level(h, level);
  CM_CHECK(nrep >= 1, "nrep must be positive");
  CM_CHECK(structured_only >= 0 && structured_only <= 2, "structured_only must be 0, 1 or 2");
  System& s = h->s;
  const bool step = structured_only == 2;
  CM_CHECK(!step || level + 1 < s.nlev(), "the coarsest level has no smoother");
  if (prec == 0) s.prepare_prec<double>(level, step);
  else s.prepare_prec<float>(level, step);
  if (step) {
    cm::Level& v = s.lv(level);
    const long long nv = 3 * v.nn;
    if (prec == 0) {
      hipLaunchKernelGGL(cm::k_fill<double>, cm::grid1(nv), dim3(256), 0, s.stream, v.d64.b.p, nv, 0.0);
      hipLaunchKernelGGL(cm::k_fill<double>, cm::grid1(nv), dim3(256), 0, s.stream, v.d64.x.p, nv, 0.0);
    } else {
      hipLaunchKernelGGL(cm::k_fill<float>, cm::grid1(nv), dim3(256), 0, s.stream, v.d32.b.p, nv, 0.f);
      hipLaunchKernelGGL(cm::k_fill<float>, cm::grid1(nv), dim3(256), 0, s.stream, v.d32.x.p, nv, 0.f);
    }
    CM_LAUNCH_CHECK();
  }
  hipEvent_t e0, e1;
  CM_HIP(hipEventCreate(&e0)); CM_HIP(hipEventCreate(&e1));
  auto once = [&]() {
    if (prec == 0) {
      if (step) s.smooth_step_structured<double>(level, cp<double>(u));
      else if (structured_only) s.apply_structured<double>(level, cp<double>(u), mp<double>(y));
      else s.apply<double>(level, cp<double>(u), mp<double>(y), cm::MODE_APPLY, nullptr);
    } else {
      if (step) s.smooth_step_structured<float>(level, cp<float>(u));
      else if (structured_only) s.apply_structured<float>(level, cp<float>(u), mp<float>(y));
      else s.apply<float>(level, cp<float>(u), mp<float>(y), cm::MODE_APPLY, nullptr);
    }
  };
  once();
  CM_HIP(hipEventRecord(e0, s.stream));
  for (int r = 0; r < nrep; ++r) once();
  CM_HIP(hipEventRecord(e1, s.stream));
  CM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CM_HIP(hipEventElapsedTime(&ms, e0, e1));
  CM_HIP(hipEventDestroy(e0)); CM_HIP(hipEventDestroy(e1));
  *avg_ms = ms / nrep;
  CM_API_END
}

int cm_sys_autotune(cm_sys* h, int prec, int nrep) {
  CM_API_BEGIN
  need_built(h);
  CM_CHECK(nrep != 0, "nrep must not be 0");
  System& s = h->s;
  const bool retime = nrep < 0;
  if (prec == 0) {
    s.ensure_data<double>();
    s.autotune_kz<double>(retime ? -nrep : nrep, retime);
  } else {
    s.ensure_prec<float>();   // an FP32 V-cycle always has its smoother data
    s.autotune_kz<float>(retime ? -nrep : nrep, retime);
  }
  s.drop_graphs();
  CM_API_END
}

int cm_sys_timings(cm_sys* h, double* out, int cap) {
  CM_API_BEGIN
  need_built(h);
  for (int k = 0; k < 9 && k < cap; ++k) out[k] = h->s.t_stage[k];
  CM_API_END
}

int cm_sys_set_kz(cm_sys* h, int kz) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(kz >= 0 && kz <= 4096, "kz out of range");
  h->s.kz_override = kz;
  h->s.drop_graphs();
  if (h->s.built) {
    h->s.refresh_all_flags<double>();
    h->s.refresh_all_flags<float>();
  }
  CM_API_END
}

int cm_sys_set_march_blocks(cm_sys* h, int min_blocks) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(min_blocks >= 1, "min_blocks must be positive");
  System& s = h->s;
  s.march_blocks = min_blocks;
  s.drop_graphs();
  if (s.built) {
    s.refresh_all_flags<double>();
    s.refresh_all_flags<float>();
  }
  CM_API_END
}

int cm_sys_set_tuned_kz(cm_sys* h, int level, int prec, int smooth, int kz) {
  CM_API_BEGIN
  need_level(h, level);
  CM_CHECK(prec == 0 || prec == 1, "prec must be 0 (FP64) or 1 (FP32)");
  CM_CHECK(kz >= 0 && kz <= 4096, "kz out of range");
  System& s = h->s;
  s.lv(level).kz[prec][smooth ? 1 : 0] = kz;
  s.drop_graphs();
  if (prec == 0) s.refresh_flags<double>(level);
  else s.refresh_flags<float>(level);
  CM_API_END
}

int cm_sys_set_last_coarsening(cm_sys* h, int axes) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(!h->s.built, "set the coarsening before cm_sys_build");
  CM_CHECK(axes >= 1 && axes <= 7, "axes is a bit mask of x (1), y (2), z (4)");
  h->s.last_axes = axes;
  CM_API_END
}

int cm_sys_set_copy_blocks(cm_sys* h, int on) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_CHECK(!h->s.built, "set the smoother blocks before cm_sys_build");
  h->s.copy_blocks = on != 0;
  CM_API_END
}

int cm_sys_set_block_skip(cm_sys* h, int on) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  h->s.use_flags = on != 0;
  h->s.drop_graphs();
  CM_API_END
}

int cm_sys_set_graphs(cm_sys* h, int on) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  h->s.graphs = on != 0;
  h->s.drop_graphs();
  CM_API_END
}

int cm_sys_sync(cm_sys* h) {
  CM_API_BEGIN
  CM_CHECK(h != nullptr, "null system");
  CM_HIP(hipStreamSynchronize(h->s.stream));
  CM_API_END
}

int cm_lattice_materials(void* stream, int nx, int ny, int nz, int ns3, const int16_t* L, int nlab,
                         const int8_t* lut_cls, const int8_t* lut_grp, int ngrp, const double* lam_k,
                         const double* mu_k, double* E, double* nu, uint8_t* active, double* F, uint8_t* ser, float* S,
                         int32_t* part) {
  CM_API_BEGIN
  CM_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "empty lattice");
  CM_CHECK(ns3 >= 1 && ns3 <= cm::LAT_MAX_NS3, "1..64 sub-samples per cell");
  CM_CHECK(nlab >= 1, "empty label table");
  CM_CHECK(ngrp >= 1 && ngrp <= 64, "at most 63 bone groups");
  hipStream_t st = (hipStream_t)stream;
  const long long ncell = (long long)nx * ny * nz;
  cm::DevBuf<double> mix(7 * (size_t)ncell);
  cm::DevBuf<unsigned long long> pres((size_t)ncell);
  cm::DevBuf<uint8_t> cart((size_t)ncell);
  cm::LatticeArgs a{nx, ny, nz, ns3, L, lut_cls, lut_grp, nlab, ngrp, lam_k, mu_k, F, S, mix.p, pres.p, cart.p};
  hipLaunchKernelGGL(cm::k_lattice_cells, cm::grid1(ncell), dim3(256), 0, st, a);
  hipLaunchKernelGGL(cm::k_lattice_finish, cm::grid1(ncell), dim3(256), 0, st, a, E, nu, active, ser, part);
  CM_LAUNCH_CHECK();
  CM_HIP(hipStreamSynchronize(st));
  CM_API_END
}

int cm_body_fractions(void* stream, int64_t ncell, int ns3, const int16_t* L, int nlab, const uint32_t* mask, int nset,
                      double* out) {
  CM_API_BEGIN
  CM_CHECK(ncell >= 0, "negative size");
  CM_CHECK(ns3 >= 1 && ns3 <= cm::LAT_MAX_NS3, "1..64 sub-samples per cell");
  CM_CHECK(nlab >= 1, "empty body table");
  CM_CHECK(nset >= 1 && nset <= cm::BF_MAX_SETS, "1..16 body sets");
  hipStream_t st = (hipStream_t)stream;
  if (ncell)
    hipLaunchKernelGGL(cm::k_body_fractions, cm::grid1(ncell), dim3(256), 0, st, (long long)ncell, ns3, L, nlab, mask,
                       nset, out);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_part_nodes(void* stream, int64_t ncell, const int64_t* cn, const int32_t* part, const uint8_t* active,
                  int64_t nnode, uint64_t* node_mask, int32_t* part_cells) {
  CM_API_BEGIN
  CM_CHECK(ncell >= 0 && nnode >= 0, "negative size");
  hipStream_t st = (hipStream_t)stream;
  if (nnode)
    hipLaunchKernelGGL(cm::k_fill<unsigned long long>, cm::grid1(nnode), dim3(256), 0, st,
                       (unsigned long long*)node_mask, (long long)nnode, 0ull);
  hipLaunchKernelGGL(cm::k_fill<int>, dim3(1), dim3(256), 0, st, (int*)part_cells, 64LL, 0);
  if (ncell)
    hipLaunchKernelGGL(cm::k_part_mask, cm::grid1(ncell), dim3(256), 0, st, (long long)ncell, (const long long*)cn,
                       (const int*)part, active, (unsigned long long*)node_mask, (int*)part_cells);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_nearest_bone(void* stream, int npt, const double* pts, int nbone, const int64_t* bptr, const int64_t* bnodes,
                    int nx, int ny, const double* xe, const double* ye, const double* ze, const int64_t* orig,
                    double* d2min) {
  CM_API_BEGIN
  CM_CHECK(npt >= 0 && nbone >= 0 && (long long)npt * nbone < (1LL << 31), "point x bone count out of range");
  if (npt == 0 || nbone == 0) return 0;
  const cm::NodeGeo g{nx, ny, xe, ye, ze, (const long long*)orig};
  hipLaunchKernelGGL(cm::k_nearest_bone, dim3((unsigned)(npt * nbone)), dim3(256), 0, (hipStream_t)stream, npt, pts,
                     nbone, (const long long*)bptr, (const long long*)bnodes, g, d2min);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_spring_patches(void* stream, int nend, const double* pts, const int32_t* bone, const int64_t* bptr,
                      const int64_t* bnodes, const uint8_t* used, int nx, int ny, const double* xe, const double* ye,
                      const double* ze, const int64_t* orig, const int64_t* foff, uint8_t* flags, double* radius,
                      int32_t* count) {
  CM_API_BEGIN
  CM_CHECK(nend >= 0, "negative end count");
  if (nend == 0) return 0;
  const cm::NodeGeo g{nx, ny, xe, ye, ze, (const long long*)orig};
  hipLaunchKernelGGL(cm::k_spring_patches, dim3((unsigned)nend), dim3(256), 0, (hipStream_t)stream, nend, pts,
                     (const int*)bone, (const long long*)bptr, (const long long*)bnodes, used, g,
                     (const long long*)foff, flags, radius, (int*)count);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_split_connectivity(void* stream, int nx, int ny, int nz, const uint8_t* active, const int32_t* part,
                          const float* share, int nb, int64_t* cn, int64_t* copy_orig, int64_t cap, int64_t* ncopy,
                          int64_t* nsplit) {
  CM_API_BEGIN
  CM_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "empty grid");
  CM_CHECK(nb >= 2 && nb <= 4096, "bone count out of range");
  CM_CHECK(ncopy != nullptr && nsplit != nullptr, "null output");
  hipStream_t st = (hipStream_t)stream;
  const long long ncell = (long long)nx * ny * nz;
  const long long nn = (long long)(nx + 1) * (ny + 1) * (nz + 1);
  CM_CHECK(nn * nb < (1LL << 40), "grid x bones too large");
  cm::DevBuf<int> cnt((size_t)(nn * nb)), nc((size_t)nn);
  cm::DevBuf<long long> base((size_t)nn), btot((size_t)cm::div_up(nn, cm::SCAN_ITEMS)), tot(1);
  cm::DevBuf<unsigned long long> ns(1);
  hipLaunchKernelGGL(cm::k_fill<int>, cm::grid1(nn * nb), dim3(256), 0, st, cnt.p, nn * nb, 0);
  hipLaunchKernelGGL(cm::k_fill<unsigned long long>, dim3(1), dim3(256), 0, st, ns.p, 1LL, 0ull);
  hipLaunchKernelGGL(cm::k_split_count, cm::grid1(ncell), dim3(256), 0, st, nx, ny, nz, active, (const int*)part, nb,
                     cnt.p);
  hipLaunchKernelGGL(cm::k_split_nodes, cm::grid1(nn), dim3(256), 0, st, nn, nb, cnt.p, nc.p, ns.p);
  const long long nblk = cm::div_up(nn, cm::SCAN_ITEMS);
  hipLaunchKernelGGL(cm::k_scan_blocks, dim3((unsigned)nblk), dim3(256), 0, st, nn, nc.p, base.p, btot.p);
  hipLaunchKernelGGL(cm::k_scan_totals, dim3(1), dim3(256), 0, st, nblk, btot.p, tot.p);
  hipLaunchKernelGGL(cm::k_scan_add, cm::grid1(nn), dim3(256), 0, st, nn, base.p, btot.p);
  CM_LAUNCH_CHECK();
  long long total = 0;
  unsigned long long nsp = 0;
  CM_HIP(hipMemcpyAsync(&total, tot.p, sizeof(total), hipMemcpyDeviceToHost, st));
  CM_HIP(hipMemcpyAsync(&nsp, ns.p, sizeof(nsp), hipMemcpyDeviceToHost, st));
  CM_HIP(hipStreamSynchronize(st));
  *ncopy = total;
  *nsplit = (int64_t)nsp;
  CM_CHECK(total <= cap, "copy_orig capacity too small: need " + std::to_string(total));
  if (total > 0)
    hipLaunchKernelGGL(cm::k_split_orig, cm::grid1(nn), dim3(256), 0, st, nn, nb, cnt.p, nc.p, base.p,
                       (long long*)copy_orig);
  hipLaunchKernelGGL(cm::k_split_cells, cm::grid1(ncell), dim3(256), 0, st, nx, ny, nz, active, (const int*)part,
                     share, nb, cnt.p, nc.p, base.p, (long long*)cn);
  CM_LAUNCH_CHECK();
  CM_HIP(hipStreamSynchronize(st));
  CM_API_END
}

int cm_bank_pack(void* stream, int64_t ncell, int ncls, const float* vm, const float* u_nodes, const int64_t* cell_nodes,
                 const int32_t* lab, const float* frac, void* vm16, void* uc16, int8_t* lab8, uint8_t* frac8) {
  CM_API_BEGIN
  CM_CHECK(ncell >= 0 && ncls >= 0, "negative sizes");
  if (ncell == 0) return 0;
  hipLaunchKernelGGL(cm::k_bank_pack, cm::grid1(ncell), dim3(256), 0, (hipStream_t)stream, (long long)ncell, ncls, vm,
                     u_nodes, (const long long*)cell_nodes, (const int*)lab, frac, (__half*)vm16, (__half*)uc16, lab8,
                     frac8);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_sample_fe(void* stream, int mx, int my, int mz, int sub, int nx, int ny, int nz, const int* si, const double* st,
                 const uint8_t* sin_, const int* sj, const double* su, const uint8_t* sjn, const int* sk,
                 const double* sv, const uint8_t* skn, const uint8_t* active, const int64_t* cn, const float* VM,
                 const float* SIG, const double* Urel, float* vm_out, float* u_out, double* cover, float* sig_out) {
  CM_API_BEGIN
  CM_CHECK(mx > 0 && my > 0 && mz > 0 && sub > 0, "bad analysis grid");
  const long long ncm = (long long)mx * my * mz;
  hipLaunchKernelGGL(cm::k_sample_fe, cm::grid1(ncm), dim3(256), 0, (hipStream_t)stream, mx, my, mz, sub, nx, ny, nz,
                     si, st, sin_, sj, su, sjn, sk, sv, skn, active, (const long long*)cn, VM, SIG, Urel, vm_out,
                     u_out, cover, sig_out);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_sample_nodes(void* stream, int mx, int my, int mz, const double* mxe, const double* mye, const double* mze,
                    const int* ci, const int* cj, const int* ck, int nq, const int* kq, const int* qz, int nx, int ny,
                    int nz, const double* xe, const double* ye, const double* ze, const uint8_t* active,
                    const int64_t* cn, const double* Urel, int* dz2, int* kbest, float* out) {
  CM_API_BEGIN
  CM_CHECK(mx > 0 && my > 0 && mz > 0 && nx > 0 && ny > 0 && nz > 0, "bad grids");
  CM_CHECK(nq >= 1 && nq <= mz + 1, "bad query layers");
  const hipStream_t st = (hipStream_t)stream;
  const long long ncol = (long long)nx * ny;
  hipLaunchKernelGGL(cm::k_column_nearest, cm::grid1(ncol * nq), dim3(256), 0, st, nx, ny, nz, nq, kq, active, dz2,
                     kbest);
  const long long mnode = (long long)(mx + 1) * (my + 1) * (mz + 1);
  hipLaunchKernelGGL(cm::k_sample_nodes, cm::grid1(mnode), dim3(256), 0, st, mx + 1, my + 1, mz + 1, mxe, mye, mze, ci,
                     cj, ck, qz, nx, ny, nz, xe, ye, ze, active, dz2, kbest, (const long long*)cn, Urel, out);
  CM_LAUNCH_CHECK();
  CM_API_END
}

int cm_sample_nodes_noalloc(void* stream, int mx, int my, int mz, const double* mxe, const double* mye,
                            const double* mze, const int* ci, const int* cj, const int* ck, int nq, const int* kq,
                            const int* qz, int nx, int ny, int nz, const double* xe, const double* ye, const double* ze,
                            const uint8_t* active, const int64_t* cn, const double* Urel, float* out) {
  CM_API_BEGIN
  CM_CHECK(mx > 0 && my > 0 && mz > 0 && nx > 0 && ny > 0 && nz > 0, "bad grids");
  CM_CHECK(nq >= 1 && nq <= mz + 1, "bad query layers");
  const long long ncol = (long long)nx * ny;
  CM_CHECK(ncol * nq <= (long long)CM_SAMPLE_NODES_MAX_COLUMNS,
           "cm_sample_nodes_noalloc: nq * nx * ny = " + std::to_string(ncol * nq) + " exceeds the scratch limit of " +
               std::to_string((long long)CM_SAMPLE_NODES_MAX_COLUMNS) + " (use cm_sample_nodes with work arrays)");
  // both stage-1 tables live in this device's copy of the scratch array; k_column_nearest writes
  // every entry the second kernel reads, so nothing needs clearing between calls
  int* dz2 = cm::scratch_on_current_device();
  int* kbest = dz2 + CM_SAMPLE_NODES_MAX_COLUMNS;
  const hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(cm::k_column_nearest, cm::grid1(ncol * nq), dim3(256), 0, st, nx, ny, nz, nq, kq, active, dz2,
                     kbest);
  const long long mnode = (long long)(mx + 1) * (my + 1) * (mz + 1);
  hipLaunchKernelGGL(cm::k_sample_nodes, cm::grid1(mnode), dim3(256), 0, st, mx + 1, my + 1, mz + 1, mxe, mye, mze, ci,
                     cj, ck, qz, nx, ny, nz, xe, ye, ze, active, dz2, kbest, (const long long*)cn, Urel, out);
  CM_LAUNCH_CHECK();
  CM_API_END
}


int cm_sys_prepare_solve(cm_sys* h) {
  CM_API_BEGIN
  need_built(h);
  h->s.prepare_solve();
  CM_HIP(hipStreamSynchronize(h->s.stream));
  CM_API_END
}

int cm_batch_create(cm_batch** out, void* stream, int check_every) {
  CM_API_BEGIN
  CM_CHECK(out != nullptr, "out is null");
  auto* B = new cm_batch();
  B->st = (hipStream_t)stream;
  B->check_every = check_every > 0 ? check_every : 5;
  *out = B;
  CM_API_END
}

int cm_batch_destroy(cm_batch* B) {
  CM_API_BEGIN
  if (B) {
    for (auto& sl : B->slots)
      if (sl.live && sl.h) sl.h->s.stream = sl.user;
    delete B;
  }
  CM_API_END
}

int cm_batch_add(cm_batch* B, cm_sys* h, const double* b, double* x, double tol, int maxiter, int* slot) {
  CM_API_BEGIN
  CM_CHECK(B != nullptr && h != nullptr && slot != nullptr, "null argument");
  need_built(h);
  *slot = B->add(h, b, x, tol, maxiter);
  CM_API_END
}

int cm_batch_step(cm_batch* B, int max_iters, int* nlive) {
  CM_API_BEGIN
  CM_CHECK(B != nullptr && nlive != nullptr, "null argument");
  *nlive = B->step(max_iters > 0 ? max_iters : 1);
  CM_API_END
}

int cm_batch_poll(cm_batch* B, int* slot, cm_solve_info* info) {
  CM_API_BEGIN
  CM_CHECK(B != nullptr && slot != nullptr && info != nullptr, "null argument");
  B->poll(slot, info);
  CM_API_END
}

}  // extern "C"
