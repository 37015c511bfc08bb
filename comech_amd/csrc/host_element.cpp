// CPU build of the element operator used by the GPU kernels (element.hpp), so the kernel math is
// testable on machines without a GPU.  Not used by the solver.
#include <cmath>

#include "element.hpp"

using namespace cm;

extern "C" {

// Dense 24x24 element matrix (row-major) of the constant-material operator, column by column.
void cm_host_elem_const(const double* dims, double lam, double mu, int sri, double* K) {
  const double V = dims[0] * dims[1] * dims[2];
  CellW<double> w{0.25 / dims[0], 0.25 / dims[1], 0.25 / dims[2], lam * V, mu * V};
  for (int c = 0; c < 24; ++c) {
    double u[24] = {0}, f[24];
    u[c] = 1.0;
    if (sri) elem_apply_const<double, true>(u, f, w);
    else elem_apply_const<double, false>(u, f, w);
    for (int r = 0; r < 24; ++r) K[r * 24 + c] = f[r];
  }
}

// Same in single precision (returns the FP32 result widened).
void cm_host_elem_const_f32(const double* dims, double lam, double mu, int sri, const double* u64, double* f64) {
  const float V = float(dims[0] * dims[1] * dims[2]);
  CellW<float> w{0.25f / float(dims[0]), 0.25f / float(dims[1]), 0.25f / float(dims[2]), float(lam) * V, float(mu) * V};
  float u[24], f[24];
  for (int q = 0; q < 24; ++q) u[q] = float(u64[q]);
  if (sri) elem_apply_const<float, true>(u, f, w);
  else elem_apply_const<float, false>(u, f, w);
  for (int q = 0; q < 24; ++q) f64[q] = f[q];
}

// Dense 24x24 matrix of the constant-material operator with per-axis lam second moments q
// (elem_mid_q), cell dims h, lam and mu per unit volume.
void cm_host_elem_q(const double* h, double lam, double mu, const double* q, double* K) {
  const double V = h[0] * h[1] * h[2];
  CellW<double> w{0.25 / h[0], 0.25 / h[1], 0.25 / h[2], lam * V, mu * V};
  for (int c = 0; c < 24; ++c) {
    double cx[8], cy[8], cz[8];
    for (int l = 0; l < 8; ++l) { cx[l] = (3 * l == c); cy[l] = (3 * l + 1 == c); cz[l] = (3 * l + 2 == c); }
    wht_fwd(cx); wht_fwd(cy); wht_fwd(cz);
    double Fx[8], Fy[8], Fz[8];
    elem_mid_q<double>(cx, cy, cz, Fx, Fy, Fz, w, q[0], q[1], q[2]);
    wht_bwd(Fx); wht_bwd(Fy); wht_bwd(Fz);
    for (int l = 0; l < 8; ++l) { K[(3 * l) * 24 + c] = Fx[l]; K[(3 * l + 1) * 24 + c] = Fy[l]; K[(3 * l + 2) * 24 + c] = Fz[l]; }
  }
}

void cm_host_elem_mom(const double* momL, const double* momM, const double* h, double* K) {
  for (int c = 0; c < 24; ++c) {
    double u[24] = {0}, f[24];
    u[c] = 1.0;
    elem_apply_mom<double>(u, f, momL, momM, 0.25 / h[0], 0.25 / h[1], 0.25 / h[2]);
    for (int r = 0; r < 24; ++r) K[r * 24 + c] = f[r];
  }
}

void cm_host_mom_const(double m, double V, int centre_point, double* mom) { mom_const<double>(m, V, centre_point != 0, mom); }

void cm_host_mom_shift_add(const double* in, const double* a, const double* b, double* out) {
  mom_shift_add<double>(in, a, b, out);
}

void cm_host_cell_stress(const double* u, const double* dims, double lam, double mu, double* eps, double* sig,
                         double* vm) {
  cell_centre_stress<double>(u, 0.25 / dims[0], 0.25 / dims[1], 0.25 / dims[2], lam, mu, eps, sig, vm);
}

// CPU emulation of the structured march on one cell column (1 x 1 x nz cells; node (a, b, k) at
// index 4k + a + 2b, 3 values each): the same face transforms, z stage, transform-space cell
// operator, force carry and back-transform as the GPU kernel.  iso != 0 uses the cubic-cell form
// (requires hxy[0] = hxy[1] = hz[k] for all k).
void cm_host_march_column(int nz, const double* hxy, const double* hz, const double* lam, const double* mu, int sri,
                          int iso, const double* u, double* y) {
  const double gx = 0.25 / hxy[0], gy = 0.25 / hxy[1];
  double ulo[12], uhi[12], car[12] = {0};
  auto face = [&](int kk, double* f) {
    double v[4][3];
    for (int ab = 0; ab < 4; ++ab)
      for (int c = 0; c < 3; ++c) v[ab][c] = u[(4 * kk + ab) * 3 + c];
    face_fwd(v, f);
  };
  face(0, ulo);
  for (int k = 0; k <= nz; ++k) {
    double Flo[12] = {0}, Fhi[12] = {0};
    if (k < nz) {
      face(k + 1, uhi);
      double c3[3][8], F3[3][8];
      faces_to_cell(ulo, uhi, c3);
      if (iso && sri) {   // the kernel's face form (carry included)
        const double f = hxy[0] / 16.0;
        double Fc[12];
        cell_faces_iso_sri<double>(ulo, uhi, car, Fc, Fhi, lam[k] * f, mu[k] * f);
        for (int q = 0; q < 12; ++q) Flo[q] = Fc[q] - car[q];
      } else if (iso) {
        const double f = hxy[0] / 16.0;
        CellW<double> w{1.0, 1.0, 1.0, lam[k] * f, mu[k] * f};
        elem_mid<double, false, true>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w);
      } else {
        const double V = hxy[0] * hxy[1] * hz[k];
        CellW<double> w{gx, gy, 0.25 / hz[k], lam[k] * V, mu[k] * V};
        if (sri) elem_mid<double, true, false>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w);
        else elem_mid<double, false, false>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w);
      }
      if (!(iso && sri)) cell_to_faces(F3, Flo, Fhi);
    }
    for (int q = 0; q < 12; ++q) {
      Flo[q] += car[q];
      car[q] = Fhi[q];
    }
    double cr[4][3];
    face_bwd(Flo, cr);
    for (int ab = 0; ab < 4; ++ab)
      for (int c = 0; c < 3; ++c) y[(4 * k + ab) * 3 + c] = cr[ab][c];
    for (int q = 0; q < 12; ++q) ulo[q] = uhi[q];
  }
}

// One z step of the cubic-cell SRI operator on face transforms (lo, hi, car: 12 values each):
// fused != 0 runs the kernel's face form (cell_faces_iso_sri), 0 the composition faces_to_cell +
// elem_mid + cell_to_faces + carry add it replaces.
void cm_host_cell_faces(int fused, const double* lo, const double* hi, const double* car, double wl, double wm,
                        double* flo, double* fhi) {
  if (fused) {
    cell_faces_iso_sri<double>(lo, hi, car, flo, fhi, wl, wm);
    return;
  }
  double c3[3][8], F3[3][8];
  faces_to_cell(lo, hi, c3);
  CellW<double> w{1.0, 1.0, 1.0, wl, wm};
  elem_mid<double, true, true>(c3[0], c3[1], c3[2], F3[0], F3[1], F3[2], w);
  cell_to_faces(F3, flo, fhi);
  for (int q = 0; q < 12; ++q) flo[q] += car[q];
}

}  // extern "C"
