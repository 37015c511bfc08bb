// Element operator of the 8-node trilinear hexahedron, written in the Walsh-Hadamard
// (hierarchical) basis of the Q1 space.
//
// Reference formulation (what this must reproduce bit-for-bit up to FP rounding):
//   fe/solver.py:15-90   K_e = lam * KL + mu * KM, 2x2x2 Gauss (exact for Q1 on boxes)
//   rb/splitfe.py:58-69  SRI: the lam term integrated at the cell centre only (rank one)
//   fe/gmg.py:79-102     Galerkin coarse operators P^T A P with trilinear P
//
// Local node l = a + 2b + 4c, corner signs s = (2a-1, 2b-1, 2c-1); element DOF 3*l + comp
// (fe/solver.py:11-12,145).  With modes alpha in {0,1}^3 (index m = a1 + 2 a2 + 4 a3) the
// unnormalised transform  C[m] = sum_l s_l^alpha u_l  gives u(xi) = sum_m (C[m]/8) xi^alpha.
// Physical gradient G_ij = (2/h_j) d u_i / d xi_j.  The strain energy
//   E = 1/2 int  lam (div u)^2 + 2 mu sum_i G_ii^2 + mu sum_{i<j} (G_ij + G_ji)^2  dV
// equals the reference's 1/2 u^T (lam KL + mu KM) u (D_lambda, D_mu of fe/solver.py:39-40).
// Every term is a polynomial p_t(xi) = sum_beta a_t[beta] xi^beta with beta in {0,1}^3, so
//   E = 1/2 sum_t kappa_t sum_{beta,beta'} a_t[beta] Mom[beta+beta'] a_t[beta']
// with material moments Mom[gamma] = int m(x) xi^gamma dV, gamma in {0,1,2}^3 (27 values).
// For a constant material on the cell the moment matrix is diagonal (monomials of degree <= 1
// are orthogonal under exact integration), which makes the fine-level operator cheap.  Coarse
// levels carry the moments of the fine cells they cover: that is exact Galerkin coarsening for
// nested Q1 spaces with trilinear interpolation.
#pragma once

#ifndef __HIPCC__
#define CM_HD
#else
#define CM_HD __host__ __device__
#endif

#define CM_INLINE CM_HD inline __attribute__((always_inline))

namespace cm {

// Forward transform on 8 values (in place): C[m] = sum_l s_l^alpha(m) v_l.
template <typename T>
CM_INLINE void wht_fwd(T* v) {
#pragma unroll
  for (int s = 1; s < 8; s <<= 1) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      if (l & s) continue;
      T lo = v[l], hi = v[l + s];
      v[l] = hi + lo;        // mode bit 0 along this axis
      v[l + s] = hi - lo;    // mode bit 1: s = +1 at the high corner, -1 at the low corner
    }
  }
}

// Transposed transform on 8 values (in place): f_l = sum_m s_l^alpha(m) F[m].
template <typename T>
CM_INLINE void wht_bwd(T* v) {
#pragma unroll
  for (int s = 1; s < 8; s <<= 1) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      if (l & s) continue;
      T f0 = v[l], f1 = v[l + s];
      v[l] = f0 - f1;        // low corner
      v[l + s] = f0 + f1;    // high corner
    }
  }
}

// Transforms of one cell face (4 nodes, corner a + 2b), used by the z-marching kernels: the 3-D
// transform of a cell = xy transforms of its bottom (lo) and top (hi) faces, then
// C[m] = hi[m] + lo[m], C[m + 4] = hi[m] - lo[m]; the forces split back as lo = F[m] - F[m+4],
// hi = F[m] + F[m+4], then an xy back-transform per face.
//   face_fwd: v[corner][comp] -> f[4 comp + (mx + 2 my)]
template <typename T>
CM_INLINE void face_fwd(const T (&v)[4][3], T* __restrict__ f) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T s0 = v[1][c] + v[0][c], s1 = v[1][c] - v[0][c];
    const T s2 = v[3][c] + v[2][c], s3 = v[3][c] - v[2][c];
    f[4 * c + 0] = s2 + s0;
    f[4 * c + 1] = s3 + s1;
    f[4 * c + 2] = s2 - s0;
    f[4 * c + 3] = s3 - s1;
  }
}
//   face_bwd: F[4 comp + m] -> cr[corner][comp]
template <typename T>
CM_INLINE void face_bwd(const T* __restrict__ F, T (&cr)[4][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T t0 = F[4 * c] - F[4 * c + 1], t1 = F[4 * c] + F[4 * c + 1];
    const T t2 = F[4 * c + 2] - F[4 * c + 3], t3 = F[4 * c + 2] + F[4 * c + 3];
    cr[0][c] = t0 - t2;
    cr[2][c] = t0 + t2;
    cr[1][c] = t1 - t3;
    cr[3][c] = t1 + t3;
  }
}
//   z stage: faces -> 3-D transform c3[comp][m]; forces F3 -> faces (lo, hi)
template <typename T>
CM_INLINE void faces_to_cell(const T* __restrict__ lo, const T* __restrict__ hi, T (&c3)[3][8]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      c3[c][m] = hi[4 * c + m] + lo[4 * c + m];
      c3[c][m + 4] = hi[4 * c + m] - lo[4 * c + m];
    }
}
template <typename T>
CM_INLINE void cell_to_faces(const T (&F3)[3][8], T* __restrict__ lo, T* __restrict__ hi) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      lo[4 * c + m] = F3[c][m] - F3[c][m + 4];
      hi[4 * c + m] = F3[c][m] + F3[c][m + 4];
    }
}

// Per-cell constants of the constant-material operator.
//   g_j = (2/h_j)/8 converts transform coefficients into physical gradients.
//   wl = lam * V, wm = mu * V.
template <typename T>
struct CellW {
  T gx, gy, gz, wl, wm;
};

// Transform-space core of the constant-material operator: (cx, cy, cz) = forward transforms of
// the three displacement components, (Fx, Fy, Fz) = the transform-space forces (the element
// forces are wht_bwd of them).  ISO = true is the cubic cell (h_x = h_y = h_z = h): every term
// carries exactly two gradient factors, so the caller passes wl = lam V g^2, wm = mu V g^2
// (V g^2 = h / 16) and the gradient factors are compile-time ones.
// The mu terms (exact integration of a constant mu over the cell); sets F.
template <typename T, bool ISO>
CM_INLINE void elem_mu_part(const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
                            T* __restrict__ Fx, T* __restrict__ Fy, T* __restrict__ Fz, const CellW<T>& w) {
  const T gx = ISO ? T(1) : w.gx, gy = ISO ? T(1) : w.gy, gz = ISO ? T(1) : w.gz;
  const T third = T(1) / T(3), ninth = T(1) / T(9);
  // monomial weights V * 3^-|beta| folded with the material factor
  const T m0 = w.wm, m1 = w.wm * third, m2 = w.wm * ninth;
#pragma unroll
  for (int m = 0; m < 8; ++m) { Fx[m] = T(0); Fy[m] = T(0); Fz[m] = T(0); }

  // ---- mu: 2 mu G_ii^2.  Mode alpha with alpha_i = 1 maps to beta = alpha - e_i.
  {
    const T ax = T(2) * gx * gx, ay = T(2) * gy * gy, az = T(2) * gz * gz;
    // x component, modes 1(100) 3(110) 5(101) 7(111): |beta| = 0,1,1,2
    Fx[1] += ax * m0 * cx[1]; Fx[3] += ax * m1 * cx[3]; Fx[5] += ax * m1 * cx[5]; Fx[7] += ax * m2 * cx[7];
    // y component, modes 2(010) 3(110) 6(011) 7(111)
    Fy[2] += ay * m0 * cy[2]; Fy[3] += ay * m1 * cy[3]; Fy[6] += ay * m1 * cy[6]; Fy[7] += ay * m2 * cy[7];
    // z component, modes 4(001) 5(101) 6(011) 7(111)
    Fz[4] += az * m0 * cz[4]; Fz[5] += az * m1 * cz[5]; Fz[6] += az * m1 * cz[6]; Fz[7] += az * m2 * cz[7];
  }
  // ---- mu: (G_ij + G_ji)^2 for the three shear pairs.
  // pair (x,y): beta with beta_x = beta_y = 0 couples cx[beta+e_y] and cy[beta+e_x]:
  //   beta=000: cx[2], cy[1] (w m0);  beta=001: cx[6], cy[5] (w m1)
  // beta_x=1, beta_y=0 (only G_xy): cx[3] (beta 100, m1), cx[7] (beta 101, m2)
  // beta_y=1, beta_x=0 (only G_yx): cy[3] (beta 010, m1), cy[7] (beta 011, m2)
  {
    T g;
    g = gy * cx[2] + gx * cy[1]; g *= m0; Fx[2] += gy * g; Fy[1] += gx * g;
    g = gy * cx[6] + gx * cy[5]; g *= m1; Fx[6] += gy * g; Fy[5] += gx * g;
    const T yy = gy * gy, xx = gx * gx;
    Fx[3] += yy * m1 * cx[3]; Fx[7] += yy * m2 * cx[7];
    Fy[3] += xx * m1 * cy[3]; Fy[7] += xx * m2 * cy[7];
  }
  // pair (y,z): beta_y = beta_z = 0 couples cy[beta+e_z] and cz[beta+e_y]:
  //   beta=000: cy[4], cz[2] (m0);  beta=100: cy[5], cz[3] (m1)
  // beta_y=1 (only G_yz): cy[6] (beta 010, m1), cy[7] (beta 110, m2)
  // beta_z=1 (only G_zy): cz[6] (beta 001, m1), cz[7] (beta 101, m2)
  {
    T g;
    g = gz * cy[4] + gy * cz[2]; g *= m0; Fy[4] += gz * g; Fz[2] += gy * g;
    g = gz * cy[5] + gy * cz[3]; g *= m1; Fy[5] += gz * g; Fz[3] += gy * g;
    const T zz = gz * gz, yy = gy * gy;
    Fy[6] += zz * m1 * cy[6]; Fy[7] += zz * m2 * cy[7];
    Fz[6] += yy * m1 * cz[6]; Fz[7] += yy * m2 * cz[7];
  }
  // pair (x,z): beta_x = beta_z = 0 couples cx[beta+e_z] and cz[beta+e_x]:
  //   beta=000: cx[4], cz[1] (m0);  beta=010: cx[6], cz[3] (m1)
  // beta_x=1 (only G_xz): cx[5] (beta 100, m1), cx[7] (beta 110, m2)
  // beta_z=1 (only G_zx): cz[5] (beta 001, m1), cz[7] (beta 011, m2)
  {
    T g;
    g = gz * cx[4] + gx * cz[1]; g *= m0; Fx[4] += gz * g; Fz[1] += gx * g;
    g = gz * cx[6] + gx * cz[3]; g *= m1; Fx[6] += gz * g; Fz[3] += gx * g;
    const T zz = gz * gz, xx = gx * gx;
    Fx[5] += zz * m1 * cx[5]; Fx[7] += zz * m2 * cx[7];
    Fz[5] += xx * m1 * cz[5]; Fz[7] += xx * m2 * cz[7];
  }
}

template <typename T, bool SRI, bool ISO>
CM_INLINE void elem_mid(const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
                        T* __restrict__ Fx, T* __restrict__ Fy, T* __restrict__ Fz, const CellW<T>& w) {
  // ISO: cubic cell, the common factor g^2 is folded into wl and wm by the caller
  const T gx = ISO ? T(1) : w.gx, gy = ISO ? T(1) : w.gy, gz = ISO ? T(1) : w.gz;
  const T third = T(1) / T(3), ninth = T(1) / T(9);
  elem_mu_part<T, ISO>(cx, cy, cz, Fx, Fy, Fz, w);
  // ---- lam: (div u)^2.  div monomial beta collects G_ii from components with beta_i = 0.
  {
    const T l0 = w.wl;
    T g = gx * cx[1] + gy * cy[2] + gz * cz[4];            // beta = 000
    g *= l0; Fx[1] += gx * g; Fy[2] += gy * g; Fz[4] += gz * g;
    if (!SRI) {
      const T l1 = w.wl * third, l2 = w.wl * ninth;
      g = gy * cy[3] + gz * cz[5]; g *= l1; Fy[3] += gy * g; Fz[5] += gz * g;   // beta 100
      g = gx * cx[3] + gz * cz[6]; g *= l1; Fx[3] += gx * g; Fz[6] += gz * g;   // beta 010
      g = gx * cx[5] + gy * cy[6]; g *= l1; Fx[5] += gx * g; Fy[6] += gy * g;   // beta 001
      Fz[7] += gz * gz * l2 * cz[7];                                             // beta 110
      Fy[7] += gy * gy * l2 * cy[7];                                             // beta 101
      Fx[7] += gx * gx * l2 * cx[7];                                             // beta 011
    }
  }
}

// Constant-material operator whose lam term is integrated with per-axis second moments q
// (int xi^2 dV / V along that axis): q = 1/3 is exact integration, 0 the centre point (SRI), 1/4
// the two centre points of a cell split into equal halves -- a level-1 cell whose level-0 children
// share one material integrates its lam term with the children's SRI points, and its Galerkin
// operator is this one (the mu term is exact on every child, so exact on the cell).
template <typename T>
CM_INLINE void elem_mid_q(const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
                          T* __restrict__ Fx, T* __restrict__ Fy, T* __restrict__ Fz, const CellW<T>& w, T qx,
                          T qy, T qz) {
  const T gx = w.gx, gy = w.gy, gz = w.gz;
  elem_mu_part<T, false>(cx, cy, cz, Fx, Fy, Fz, w);
  const T lx = w.wl * qx, ly = w.wl * qy, lz = w.wl * qz;
  T g = gx * cx[1] + gy * cy[2] + gz * cz[4];            // beta = 000
  g *= w.wl; Fx[1] += gx * g; Fy[2] += gy * g; Fz[4] += gz * g;
  g = gy * cy[3] + gz * cz[5]; g *= lx; Fy[3] += gy * g; Fz[5] += gz * g;   // beta 100
  g = gx * cx[3] + gz * cz[6]; g *= ly; Fx[3] += gx * g; Fz[6] += gz * g;   // beta 010
  g = gx * cx[5] + gy * cy[6]; g *= lz; Fx[5] += gx * g; Fy[6] += gy * g;   // beta 001
  Fz[7] += gz * gz * (lx * qy) * cz[7];                                      // beta 110
  Fy[7] += gy * gy * (lx * qz) * cy[7];                                      // beta 101
  Fx[7] += gx * gx * (ly * qz) * cx[7];                                      // beta 011
}

// The cubic-cell, SRI operator of one z step in face form: lo, hi = xy transforms of the cell's
// bottom and top faces (f[4 comp + q], q = mx + 2 my), car = top-face forces of the cell below;
// out: flo = car + the cell's bottom-face forces, fhi = its top-face forces.  The same operator as
// faces_to_cell + elem_mid<T, true, true> + cell_to_faces + the carry add, with the z butterflies
// folded into the element terms.  With A = hi + lo, D = hi - lo (the z modes) and m0 = wm,
// m1 = wm / 3, the transform-space forces of elem_mid are
//   x: A1 = 2 m0 A1 + g7, D1 = m0 D1, A2 = g1, D2 = g2 + g6, A3 = m0 A3, D3 = 4/9 m0 D3, D0 = g5
//   y: A1 = g1, D1 = g2 + g4, A2 = 2 m0 A2 + g7, D2 = m0 D2, A3 = m0 A3, D3 = 4/9 m0 D3, D0 = g3
//   z: A1 = g5, A2 = g3, A3 = g4 + g6, D0 = 2 m0 D0 + g7, D1 = m0 D1, D2 = m0 D2, D3 = 4/9 m0 D3
// (the shear terms of a mode add to its normal term: 2 m1 + m1 = m0, 2 m2 + m2 + m2 = 4/9 m0),
// the A0 forces are zero, and the faces take lo = FA - FD, hi = FA + FD.
template <typename T>
CM_INLINE void cell_faces_iso_sri(const T* __restrict__ lo, const T* __restrict__ hi, const T* __restrict__ car,
                                  T* __restrict__ flo, T* __restrict__ fhi, T wl, T wm) {
  const T m0 = wm, m1 = wm * (T(1) / T(3)), k2 = T(2) * wm, k49 = wm * (T(4) / T(9));
  const T kp = m0 + k49, km = m0 - k49;
  // z modes the coupled terms need
  const T xD0 = hi[0] - lo[0], xA1 = hi[1] + lo[1], xD1 = hi[1] - lo[1], xA2 = hi[2] + lo[2], xD2 = hi[2] - lo[2];
  const T yD0 = hi[4] - lo[4], yA1 = hi[5] + lo[5], yD1 = hi[5] - lo[5], yA2 = hi[6] + lo[6], yD2 = hi[6] - lo[6];
  const T zD0 = hi[8] - lo[8], zA1 = hi[9] + lo[9], zD1 = hi[9] - lo[9], zA2 = hi[10] + lo[10], zD2 = hi[10] - lo[10];
  const T zA3 = hi[11] + lo[11], zD3 = hi[11] - lo[11];
  const T g1 = m0 * (xA2 + yA1), g2 = m1 * (xD2 + yD1), g3 = m0 * (yD0 + zA2);
  const T g4 = m1 * (yD1 + zA3), g5 = m0 * (xD0 + zA1), g6 = m1 * (xD2 + zA3);
  const T g7 = wl * (xA1 + yA2 + zD0);
  // x component
  flo[0] = car[0] - g5;
  fhi[0] = g5;
  {
    const T FA = k2 * xA1 + g7;
    flo[1] = (car[1] + FA) - m0 * xD1;
    fhi[1] = FA + m0 * xD1;
  }
  {
    const T t = g2 + g6;
    flo[2] = (car[2] + g1) - t;
    fhi[2] = g1 + t;
  }
  flo[3] = car[3] + km * hi[3] + kp * lo[3];
  fhi[3] = kp * hi[3] + km * lo[3];
  // y component
  flo[4] = car[4] - g3;
  fhi[4] = g3;
  {
    const T t = g2 + g4;
    flo[5] = (car[5] + g1) - t;
    fhi[5] = g1 + t;
  }
  {
    const T FA = k2 * yA2 + g7;
    flo[6] = (car[6] + FA) - m0 * yD2;
    fhi[6] = FA + m0 * yD2;
  }
  flo[7] = car[7] + km * hi[7] + kp * lo[7];
  fhi[7] = kp * hi[7] + km * lo[7];
  // z component
  {
    const T FD = k2 * zD0 + g7;
    flo[8] = car[8] - FD;
    fhi[8] = FD;
  }
  flo[9] = (car[9] + g5) - m0 * zD1;
  fhi[9] = g5 + m0 * zD1;
  flo[10] = (car[10] + g3) - m0 * zD2;
  fhi[10] = g3 + m0 * zD2;
  {
    const T t = g4 + g6;
    flo[11] = (car[11] + t) - k49 * zD3;
    fhi[11] = t + k49 * zD3;
  }
}

// f = (lam KL + mu KM) u for one cell with constant material.  u, f: 24 values, DOF 3*l+comp.
// SRI = true integrates the lam term at the cell centre only (rb/splitfe.py:58-69).
template <typename T, bool SRI>
CM_INLINE void elem_apply_const(const T* __restrict__ u, T* __restrict__ f, const CellW<T>& w) {
  T cx[8], cy[8], cz[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) { cx[l] = u[3 * l]; cy[l] = u[3 * l + 1]; cz[l] = u[3 * l + 2]; }
  wht_fwd(cx); wht_fwd(cy); wht_fwd(cz);
  T Fx[8], Fy[8], Fz[8];
  elem_mid<T, SRI, false>(cx, cy, cz, Fx, Fy, Fz, w);
  wht_bwd(Fx); wht_bwd(Fy); wht_bwd(Fz);
#pragma unroll
  for (int l = 0; l < 8; ++l) { f[3 * l] = Fx[l]; f[3 * l + 1] = Fy[l]; f[3 * l + 2] = Fz[l]; }
}

// ---------------------------------------------------------------------------------------------
// General (moment) operator used on coarse levels.
//   momL, momM: 27 moments each, index gamma = g1 + 3 g2 + 9 g3, physical-volume integrals.
//   gx, gy, gz: (2/h_j)/8 of the coarse cell.
// Term monomials are encoded by beta index b = b1 + 2 b2 + 4 b3; the moment of beta+beta' is
// at MIDX[b][b'] = (b1+b1') + 3 (b2+b2') + 9 (b3+b3').
CM_INLINE int mom_index(int b, int bp) {
  return ((b & 1) + (bp & 1)) + 3 * (((b >> 1) & 1) + ((bp >> 1) & 1)) + 9 * (((b >> 2) & 1) + ((bp >> 2) & 1));
}

// q[k] = sum_k' M[beta_k + beta_k'] a[k'] for a term with N monomials listed in B.
template <typename T, int N>
CM_INLINE void mom_mul(const T* __restrict__ mom, const int (&B)[N], const T* a, T* q) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    T s = T(0);
#pragma unroll
    for (int kp = 0; kp < N; ++kp) s += mom[mom_index(B[k], B[kp])] * a[kp];
    q[k] = s;
  }
}

// Transform-space core of the moment operator: c = forward transforms of the three displacement
// components, F = transform-space forces (element forces = wht_bwd of them).  In two parts, the
// lam part (sets F) and the mu part (adds to F), so that a caller can form each part's 27 moments
// just before it (the two sets are not needed at once).
template <typename T>
CM_INLINE void elem_mid_mom_lam(const T (&c)[3][8], T (&F)[3][8], const T* __restrict__ momL, T gx, T gy, T gz) {
  const T g[3] = {gx, gy, gz};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int m = 0; m < 8; ++m) F[i][m] = T(0);

  // ---- lam (div u)^2: beta over {0,1}^3 \ {111}; a[beta] = sum_{i: beta_i=0} g_i c_i[beta + e_i]
  {
    constexpr int B[7] = {0, 1, 2, 4, 3, 5, 6};
    T a[7], q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      T s = T(0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (!((B[k] >> i) & 1)) s += g[i] * c[i][B[k] | (1 << i)];
      a[k] = s;
    }
    mom_mul<T, 7>(momL, B, a, q);
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (!((B[k] >> i) & 1)) F[i][B[k] | (1 << i)] += g[i] * q[k];
  }
}
template <typename T>
CM_INLINE void elem_mid_mom_mu(const T (&c)[3][8], T (&F)[3][8], const T* __restrict__ momM, T gx, T gy, T gz) {
  const T g[3] = {gx, gy, gz};
  // ---- mu 2 G_ii^2: beta with beta_i = 0
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int B[4];
    {
      int n = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b)
        if (!((b >> i) & 1)) B[n++] = b;
    }
    T a[4], q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = g[i] * c[i][B[k] | (1 << i)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T s = T(0);
#pragma unroll
      for (int kp = 0; kp < 4; ++kp) s += momM[mom_index(B[k], B[kp])] * a[kp];
      q[k] = s;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) F[i][B[k] | (1 << i)] += T(2) * g[i] * q[k];
  }
  // ---- mu (G_ij + G_ji)^2: beta with beta_i = 0 or beta_j = 0 (6 monomials);
  //      a[beta] = [beta_j=0] g_j c_i[beta+e_j] + [beta_i=0] g_i c_j[beta+e_i]
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int i = (p == 2) ? 0 : p;        // pairs (0,1) (1,2) (0,2)
    const int j = (p == 0) ? 1 : 2;
    int B[6];
    {
      int n = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b)
        if (!(((b >> i) & 1) && ((b >> j) & 1))) B[n++] = b;
    }
    T a[6], q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      T s = T(0);
      if (!((B[k] >> j) & 1)) s += g[j] * c[i][B[k] | (1 << j)];
      if (!((B[k] >> i) & 1)) s += g[i] * c[j][B[k] | (1 << i)];
      a[k] = s;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      T s = T(0);
#pragma unroll
      for (int kp = 0; kp < 6; ++kp) s += momM[mom_index(B[k], B[kp])] * a[kp];
      q[k] = s;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (!((B[k] >> j) & 1)) F[i][B[k] | (1 << j)] += g[j] * q[k];
      if (!((B[k] >> i) & 1)) F[j][B[k] | (1 << i)] += g[i] * q[k];
    }
  }
}

template <typename T>
CM_INLINE void elem_mid_mom(const T (&c)[3][8], T (&F)[3][8], const T* __restrict__ momL, const T* __restrict__ momM,
                            T gx, T gy, T gz) {
  elem_mid_mom_lam<T>(c, F, momL, gx, gy, gz);
  elem_mid_mom_mu<T>(c, F, momM, gx, gy, gz);
}

template <typename T>
CM_INLINE void elem_apply_mom(const T* __restrict__ u, T* __restrict__ f, const T* __restrict__ momL,
                              const T* __restrict__ momM, T gx, T gy, T gz) {
  T c[3][8];
#pragma unroll
  for (int l = 0; l < 8; ++l) { c[0][l] = u[3 * l]; c[1][l] = u[3 * l + 1]; c[2][l] = u[3 * l + 2]; }
  wht_fwd(c[0]); wht_fwd(c[1]); wht_fwd(c[2]);
  T F[3][8];
  elem_mid_mom<T>(c, F, momL, momM, gx, gy, gz);
  wht_bwd(F[0]); wht_bwd(F[1]); wht_bwd(F[2]);
#pragma unroll
  for (int l = 0; l < 8; ++l) { f[3 * l] = F[0][l]; f[3 * l + 1] = F[1][l]; f[3 * l + 2] = F[2][l]; }
}

// Moments of a constant material m over a cell of volume V in its own coordinates:
// int xi^gamma dV = V prod_k mu1(gamma_k), mu1 = (1, 0, 1/3).  SRI lam: point at the centre.
template <typename T>
CM_INLINE void mom_const(T m, T V, bool centre_point, T* mom) {
#pragma unroll
  for (int gi = 0; gi < 27; ++gi) {
    const int g1 = gi % 3, g2 = (gi / 3) % 3, g3 = gi / 9;
    T v = m * V;
    if (centre_point) {
      v = (gi == 0) ? v : T(0);
    } else {
      v *= (g1 == 1 ? T(0) : (g1 == 2 ? T(1) / T(3) : T(1)));
      v *= (g2 == 1 ? T(0) : (g2 == 2 ? T(1) / T(3) : T(1)));
      v *= (g3 == 1 ? T(0) : (g3 == 2 ? T(1) / T(3) : T(1)));
    }
    mom[gi] = v;
  }
}

// Re-express moments of a child cell in its parent's coordinates: xi_p = a_k xi_c + b_k per axis.
// (a xi + b)^0 = 1, ^1 = a xi + b, ^2 = a^2 xi^2 + 2ab xi + b^2.  Accumulates into out.
template <typename T>
CM_INLINE void mom_shift_add(const T* __restrict__ in, const T a[3], const T b[3], T* __restrict__ out) {
  T t[27], s[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) t[i] = in[i];
  // axis 0 (stride 1), axis 1 (stride 3), axis 2 (stride 9)
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int st = (ax == 0) ? 1 : (ax == 1 ? 3 : 9);
    const T A = a[ax], Bv = b[ax];
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const int gk = (i / st) % 3;
      if (gk != 0) continue;
      const T m0 = t[i], m1 = t[i + st], m2 = t[i + 2 * st];
      s[i] = m0;
      s[i + st] = A * m1 + Bv * m0;
      s[i + 2 * st] = A * A * m2 + T(2) * A * Bv * m1 + Bv * Bv * m0;
    }
#pragma unroll
    for (int i = 0; i < 27; ++i) t[i] = s[i];
  }
#pragma unroll
  for (int i = 0; i < 27; ++i) out[i] += t[i];
}

// Centre strain (engineering order xx,yy,zz,xy,yz,xz), stress and von Mises of one cell
// (fe/solver.py:257-272, rb/splitfe.py:161-175).  At the centre only the linear modes survive:
// G_ij = g_j C_i[e_j].
template <typename T>
CM_INLINE void cell_centre_stress(const T* __restrict__ u, T gx, T gy, T gz, T lam, T mu, T* eps, T* sig, T* vm) {
  T cx[8], cy[8], cz[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) { cx[l] = u[3 * l]; cy[l] = u[3 * l + 1]; cz[l] = u[3 * l + 2]; }
  wht_fwd(cx); wht_fwd(cy); wht_fwd(cz);
  const T exx = gx * cx[1], eyy = gy * cy[2], ezz = gz * cz[4];
  const T gxy = gy * cx[2] + gx * cy[1];
  const T gyz = gz * cy[4] + gy * cz[2];
  const T gxz = gz * cx[4] + gx * cz[1];
  eps[0] = exx; eps[1] = eyy; eps[2] = ezz; eps[3] = gxy; eps[4] = gyz; eps[5] = gxz;
  const T tr = exx + eyy + ezz;
  const T sx = lam * tr + T(2) * mu * exx, sy = lam * tr + T(2) * mu * eyy, sz = lam * tr + T(2) * mu * ezz;
  const T txy = mu * gxy, tyz = mu * gyz, txz = mu * gxz;
  sig[0] = sx; sig[1] = sy; sig[2] = sz; sig[3] = txy; sig[4] = tyz; sig[5] = txz;
  const T d1 = sx - sy, d2 = sy - sz, d3 = sz - sx;
  const T v2 = T(0.5) * (d1 * d1 + d2 * d2 + d3 * d3) + T(3) * (txy * txy + tyz * tyz + txz * txz);
  *vm = sqrt(v2);
}

}  // namespace cm
