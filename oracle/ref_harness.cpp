// C-ABI harness over the reference's own header-only math and BLAS wrappers,
// compiled from the reference sources in place (Makefile target "ref").
// Test infrastructure: used only to pin the oracle's restatement bitwise.
#include "MathUtilities.h"              // Core/MathUtilities.h
#include "PCGSolver/BLASWrapper.h"      // Core/PCGSolver/BLASWrapper.h

extern "C" {
double ref_quad_kernel(double x) { return mathutils::quad_kernel(x); }
double ref_N_kernel2(double a, double b, double c) {
  return mathutils::N_kernel<2>(Vector3s(a, b, c));
}
double ref_fraction_inside2(double l, double r) {
  return mathutils::fraction_inside(l, r);
}
double ref_fraction_inside4(double bl, double br, double tl, double tr) {
  return mathutils::fraction_inside(bl, br, tl, tr);
}
double ref_inverse_D_coeff(double h, int order) {
  return mathutils::inverse_D_coeff(h, order);
}
double ref_lerp(double a, double b, double f) { return mathutils::lerp(a, b, f); }
double ref_clamp(double a, double lo, double hi) {
  return mathutils::clamp(a, lo, hi);
}
double ref_blas_dot(const double* x, const double* y, long n) {
  std::vector<double> a(x, x + n), b(y, y + n);
  return robertbridson::BLAS::dot(a, b);
}
double ref_rotated_drag(const double* a, const double* d, int k) {
  const Vector3s av(a[0], a[1], a[2]), dv(d[0], d[1], d[2]);
  return k == 0 ? mathutils::get_rotated_drag_x(av, dv)
                : (k == 1 ? mathutils::get_rotated_drag_y(av, dv) : mathutils::get_rotated_drag_z(av, dv));
}
double ref_rotated_pore_coeff(const double* a, double sf, int k) {
  const Vector3s av(a[0], a[1], a[2]);
  return k == 0 ? mathutils::get_rotated_pore_coeff_x(av, sf)
                : (k == 1 ? mathutils::get_rotated_pore_coeff_y(av, sf)
                          : mathutils::get_rotated_pore_coeff_z(av, sf));
}
double ref_blas_abs_max(const double* x, long n) {
  std::vector<double> a(x, x + n);
  return robertbridson::BLAS::abs_max(a);
}
}
