// Host build of the device element math (vgpu_elem.hpp) for CPU tests:
// returns the dense 8x8 element matrix A_e (or A_e^T) by applying the
// operator to unit vectors, exactly as the kernels do per element.
#define VGPU_HD inline
#include "../../vanadis_solver-directional_residual_amd/csrc/vgpu_elem.hpp"

extern "C" void elem_matrix(const double h[3], double lam, const double v[3],
                            double tau, double cH, double mu, double ra,
                            int robin, int transpose, double *A /* 64 row-major */) {
  vgpu::ElemConst<double> k;
  const double ih[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
  vgpu::elem_setup<double>(h[0], h[1], h[2], ih[0], ih[1], ih[2], lam, v[0], v[1],
                           v[2], tau, cH, mu, k);
  for (int a = 0; a < 8; ++a) {
    double x[8], xh[8], yh[8], y[8];
    for (int c = 0; c < 8; ++c) x[c] = (c == a) ? 1.0 : 0.0;
    vgpu::haar_fwd<double>(x, xh);
    if (transpose) {
      vgpu::haar_apply<double, true>(k, xh, yh);
      vgpu::haar_inv<double>(yh, y);
      if (robin) vgpu::robin_add<double, true>(k, h[0], tau, v[0], v[1], v[2], ra, x, y);
    } else {
      vgpu::haar_apply<double, false>(k, xh, yh);
      vgpu::haar_inv<double>(yh, y);
      if (robin) vgpu::robin_add<double, false>(k, h[0], tau, v[0], v[1], v[2], ra, x, y);
    }
    for (int i = 0; i < 8; ++i) A[8 * i + a] = y[i];  // column a
  }
}

// The same matrix through the factored path of the marching kernel: 2-D face
// transforms of the bottom and top corners, z stage, Haar-space operator
// (transpose as W B^T), Robin term in Haar space, transposed z stage and the
// transposed face butterflies.
extern "C" void elem_matrix_faces(const double h[3], double lam, const double v[3],
                                  double tau, double cH, double mu, double ra,
                                  int robin, int transpose, double *A) {
  vgpu::ElemConst<double> k;
  const double ih[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
  vgpu::elem_setup<double>(h[0], h[1], h[2], ih[0], ih[1], ih[2], lam, v[0], v[1],
                           v[2], tau, cH, mu, k);
  for (int a = 0; a < 8; ++a) {
    double x[8], fb[4], ft[4], xh[8], yh[8], ob[4], ot[4];
    for (int c = 0; c < 8; ++c) x[c] = (c == a) ? 1.0 : 0.0;
    vgpu::face_fwd<double>(x, fb);
    vgpu::face_fwd<double>(x + 4, ft);
    vgpu::zfwd<double>(fb, ft, xh);
    if (transpose) {
      vgpu::haar_apply_tr<double>(k, xh, yh);
      if (robin) vgpu::robin_haar<double, true>(k, ih[0], tau, v[0], v[1], v[2], ra, xh, yh);
    } else {
      vgpu::haar_apply<double, false>(k, xh, yh);
      if (robin) vgpu::robin_haar<double, false>(k, ih[0], tau, v[0], v[1], v[2], ra, xh, yh);
    }
    vgpu::zinv<double>(yh, ob, ot);
    double y[8];
    vgpu::face_inv<double>(ob, y);
    vgpu::face_inv<double>(ot, y + 4);
    for (int i = 0; i < 8; ++i) A[8 * i + a] = y[i];
  }
}
