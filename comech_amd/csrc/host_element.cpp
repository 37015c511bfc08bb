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

}  // extern "C"
