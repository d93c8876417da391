/*
 * CPU oracle: plain-C restatement of the reference's element operator,
 * assembly, masked EbE products, BiCG, mass integral and diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): reference values for the tests and the CPU baseline, never part
 * of the product path.  Built with -ffp-contract=off so every expression
 * rounds like the reference built by gfortran/flang on baseline x86-64.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* par_3d reference-element sign tables, VANADIS.FOR:332-357 */
static const double KSI[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const double ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const double ZETA[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
/* Gauss abscissa literal, VANADIS.FOR:360-362 and 484 */
static const double GP = 0.577350269189626;
/* face -> local node table i_pow, VANADIS.FOR:365-393 (1-based nodes) */
static const int I_POW[6][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}, {1, 4, 8, 5},
                                {2, 3, 7, 6}, {4, 3, 7, 8}, {1, 2, 6, 5}};

/* siatka, VANADIS.FOR:226-244: x-fastest tensor grid */
void orc_siatka(const double *dx, const double *dy, const double *dz, int nax,
                int nay, int naz, double *xx, double *yy, double *zz) {
  int i = 0;
  for (int jz = 0; jz < naz; jz++)
    for (int jy = 0; jy < nay; jy++)
      for (int jx = 0; jx < nax; jx++) {
        xx[i] = dx[jx];
        yy[i] = dy[jy];
        zz[i] = dz[jz];
        i++;
      }
}

/* w_el, VANADIS.FOR:246-276 (the same running counter k / layer ik) */
void orc_w_el(int nax, int nay, int naz, int *iss) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  int k = 1, ik = 0;
  const int ik2 = nay * nax;
  for (int j = 1; j <= E; j++) {
    int *c = iss + (j - 1);
    c[0 * E] = k + ik2 * ik;
    c[1 * E] = k + 1 + ik2 * ik;
    c[2 * E] = k + 1 + nax + ik2 * ik;
    c[3 * E] = k + nax + ik2 * ik;
    c[4 * E] = k + nax * nay + ik2 * ik;
    c[5 * E] = k + 1 + nax * nay + ik2 * ik;
    c[6 * E] = k + 1 + nax * nay + nax + ik2 * ik;
    c[7 * E] = k + nax * nay + nax + ik2 * ik;
    k = k + 1;
    if (j % (nax - 1) == 0) k = k + 1;
    if (j % ((nax - 1) * (nay - 1)) == 0) {
      k = 1;
      ik = ik + 1;
    }
  }
}

/* kod_el_dr, VANADIS.FOR:278-320 */
void orc_kod_el_dr(int nax, int nay, int naz, int *kod) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int L = (nax - 1) * (nay - 1);
#define KOD(i, j) kod[((j)-1) * E + ((i)-1)]
  memset(kod, 0, sizeof(int) * 6 * (size_t)E);
  for (int i = 1; i <= L; i++) KOD(i, 1) = 1;
  for (int i = E - L + 1; i <= E; i++) KOD(i, 2) = 2;
  for (int i = nax - 1; i <= E; i += nax - 1) KOD(i, 4) = 4;
  for (int i = 1; i <= E; i += nax - 1) KOD(i, 3) = 3;
  for (int k = 0; k <= naz - 2; k++)
    for (int i = k * L + 1; i <= k * L + nax - 1; i++) KOD(i, 6) = 6;
  for (int k = 1; k <= naz - 1; k++)
    for (int i = k * L - nax + 1 + 1; i <= k * L; i++) KOD(i, 5) = 5;
#undef KOD
}

/* mark_face_nodes: V1 VANADIS.FOR:2063-2117; V3 omp_timing.for:1756-1810
 * (faces 5 and 6 carry each other's node sets in V3). */
static void mark_face_nodes(int facemap, int iwb, const int nodes[8],
                            int *mask) {
  static const int F1[6][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}, {1, 4, 8, 5},
                               {2, 3, 7, 6}, {4, 3, 7, 8}, {1, 2, 6, 5}};
  int f = iwb - 1;
  if (facemap == ORC_FACEMAP_V3 && (iwb == 5 || iwb == 6)) f = (iwb == 5) ? 5 : 4;
  for (int q = 0; q < 4; q++) mask[nodes[F1[f][q] - 1] - 1] = 1;
}

/* build_bc_nodes, VANADIS.FOR:2026-2061: Dirichlet on faces 1,2,4,5,6 */
void orc_build_bc_nodes(int facemap, const int *iss, const int *kod, int nax,
                        int nay, int naz, int *mask) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  for (int i = 0; i < N; i++) mask[i] = 0;
  for (int e = 0; e < E; e++) {
    int nodes[8];
    for (int j = 0; j < 8; j++) nodes[j] = iss[j * E + e];
    for (int iwb = 1; iwb <= 6; iwb++)
      if (iwb != 3 && kod[(iwb - 1) * E + e] == iwb)
        mark_face_nodes(facemap, iwb, nodes, mask);
  }
}

/* shape_hex8DR, VANADIS.FOR:1698-1774.  dn[3*i + k] = dN_i/d(xi_k). */
void orc_shape_hex8(double xi, double eta, double zeta, double n[8],
                    double dn[24]) {
  const double xm = 1.0 - xi, xp = 1.0 + xi;
  const double ym = 1.0 - eta, yp = 1.0 + eta;
  const double zm = 1.0 - zeta, zp = 1.0 + zeta;
  n[0] = 0.125 * xm * ym * zm;
  n[1] = 0.125 * xp * ym * zm;
  n[2] = 0.125 * xp * yp * zm;
  n[3] = 0.125 * xm * yp * zm;
  n[4] = 0.125 * xm * ym * zp;
  n[5] = 0.125 * xp * ym * zp;
  n[6] = 0.125 * xp * yp * zp;
  n[7] = 0.125 * xm * yp * zp;
  dn[3 * 0 + 0] = -0.125 * ym * zm;
  dn[3 * 1 + 0] = 0.125 * ym * zm;
  dn[3 * 2 + 0] = 0.125 * yp * zm;
  dn[3 * 3 + 0] = -0.125 * yp * zm;
  dn[3 * 4 + 0] = -0.125 * ym * zp;
  dn[3 * 5 + 0] = 0.125 * ym * zp;
  dn[3 * 6 + 0] = 0.125 * yp * zp;
  dn[3 * 7 + 0] = -0.125 * yp * zp;
  dn[3 * 0 + 1] = -0.125 * xm * zm;
  dn[3 * 1 + 1] = -0.125 * xp * zm;
  dn[3 * 2 + 1] = 0.125 * xp * zm;
  dn[3 * 3 + 1] = 0.125 * xm * zm;
  dn[3 * 4 + 1] = -0.125 * xm * zp;
  dn[3 * 5 + 1] = -0.125 * xp * zp;
  dn[3 * 6 + 1] = 0.125 * xp * zp;
  dn[3 * 7 + 1] = 0.125 * xm * zp;
  dn[3 * 0 + 2] = -0.125 * xm * ym;
  dn[3 * 1 + 2] = -0.125 * xp * ym;
  dn[3 * 2 + 2] = -0.125 * xp * yp;
  dn[3 * 3 + 2] = -0.125 * xm * yp;
  dn[3 * 4 + 2] = 0.125 * xm * ym;
  dn[3 * 5 + 2] = 0.125 * xp * ym;
  dn[3 * 6 + 2] = 0.125 * xp * yp;
  dn[3 * 7 + 2] = 0.125 * xm * yp;
}

/* quad4_3d, VANADIS.FOR:1499-1561: surface Jacobian |x_r x x_s| */
void orc_quad4_3d(const double x[4], const double y[4], const double z[4],
                  double r, double s, double *detjs) {
  double dndr[4], dnds[4];
  dndr[0] = -0.25 * (1.0 - s);
  dndr[1] = 0.25 * (1.0 - s);
  dndr[2] = 0.25 * (1.0 + s);
  dndr[3] = -0.25 * (1.0 + s);
  dnds[0] = -0.25 * (1.0 - r);
  dnds[1] = -0.25 * (1.0 + r);
  dnds[2] = 0.25 * (1.0 + r);
  dnds[3] = 0.25 * (1.0 - r);
  double xr = 0, xs = 0, yr = 0, ys = 0, zr = 0, zs = 0;
  for (int i = 0; i < 4; i++) {
    xr = xr + dndr[i] * x[i];
    xs = xs + dnds[i] * x[i];
    yr = yr + dndr[i] * y[i];
    ys = ys + dnds[i] * y[i];
    zr = zr + dndr[i] * z[i];
    zs = zs + dnds[i] * z[i];
  }
  const double cx = yr * zs - zr * ys;
  const double cy = zr * xs - xr * zs;
  const double cz = xr * ys - yr * xs;
  *detjs = sqrt(cx * cx + cy * cy + cz * cz);
}

/* gauss_point_hex8DR: V1 VANADIS.FOR:1868-2023, V2 gauss_point_hex8DR.for:
 * 1-174, V3 omp_timing.for:1567-1716.  grad W = grad N for every variant. */
void orc_gauss_point(int variant, double xi, double eta, double zeta,
                     const double xn[8], const double yn[8], const double zn[8],
                     const double lam[3], const double v[3], double *detj_out,
                     double n[8], double w[8], double dnx[8], double dny[8],
                     double dnz[8], double *tau_out) {
  double dn[24];
  orc_shape_hex8(xi, eta, zeta, n, dn);
  double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
         j32 = 0, j33 = 0;
  for (int i = 0; i < 8; i++) {
    j11 = j11 + xn[i] * dn[3 * i + 0];
    j12 = j12 + xn[i] * dn[3 * i + 1];
    j13 = j13 + xn[i] * dn[3 * i + 2];
    j21 = j21 + yn[i] * dn[3 * i + 0];
    j22 = j22 + yn[i] * dn[3 * i + 1];
    j23 = j23 + yn[i] * dn[3 * i + 2];
    j31 = j31 + zn[i] * dn[3 * i + 0];
    j32 = j32 + zn[i] * dn[3 * i + 1];
    j33 = j33 + zn[i] * dn[3 * i + 2];
  }
  const double detj = j11 * (j22 * j33 - j23 * j32) -
                      j12 * (j21 * j33 - j23 * j31) +
                      j13 * (j21 * j32 - j22 * j31);
  const double detji = 1.0 / detj;
  const double i11 = detji * (j22 * j33 - j23 * j32);
  const double i12 = detji * (j13 * j32 - j12 * j33);
  const double i13 = detji * (j12 * j23 - j13 * j22);
  const double i21 = detji * (j23 * j31 - j21 * j33);
  const double i22 = detji * (j11 * j33 - j13 * j31);
  const double i23 = detji * (j13 * j21 - j11 * j23);
  const double i31 = detji * (j21 * j32 - j22 * j31);
  const double i32 = detji * (j12 * j31 - j11 * j32);
  const double i33 = detji * (j11 * j22 - j12 * j21);
  for (int i = 0; i < 8; i++) {
    const double a = dn[3 * i + 0], b = dn[3 * i + 1], c = dn[3 * i + 2];
    dnx[i] = i11 * a + i21 * b + i31 * c;
    dny[i] = i12 * a + i22 * b + i32 * c;
    dnz[i] = i13 * a + i23 * b + i33 * c;
  }
  const double eps = 1.0e-14;
  const double vmag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double h_stream, lambda_s;
  if (variant == ORC_V3) {
    /* axis lengths from nodes 2-1, 4-1, 5-1 (omp_timing.for:1652-1666) */
    const double dlx = fabs(xn[1] - xn[0]);
    const double dly = fabs(yn[3] - yn[0]);
    const double dlz = fabs(zn[4] - zn[0]);
    if (vmag > eps)
      h_stream =
          2.0 / (fabs(v[0]) / dlx + fabs(v[1]) / dly + fabs(v[2]) / dlz + eps);
    else
      h_stream = 0.0;
    lambda_s = lam[0];
  } else {
    double ex = 0, ey = 0, ez = 0;
    if (vmag > eps) {
      ex = v[0] / vmag;
      ey = v[1] / vmag;
      ez = v[2] / vmag;
    }
    const double gx = i11 * ex + i12 * ey + i13 * ez;
    const double gy = i21 * ex + i22 * ey + i23 * ez;
    const double gz = i31 * ex + i32 * ey + i33 * ez;
    const double gnorm = sqrt(gx * gx + gy * gy + gz * gz);
    h_stream = (gnorm > eps) ? 2.0 / gnorm : 0.0;
    if (variant == ORC_V1)
      lambda_s = lam[0] * ex * ex + lam[1] * ey * ey + lam[2] * ez * ez;
    else
      lambda_s = lam[0];
  }
  double Pe = 0.0;
  if (vmag > eps && lambda_s > eps) Pe = vmag * h_stream / (2.0 * lambda_s);
  double tau = 0.0;
  if (Pe > 1.0e-12 && vmag > eps) {
    if (variant == ORC_V1) {
      double beta;
      if (Pe < 1.0e-3)
        beta = Pe / 3.0 - (Pe * Pe * Pe) / 45.0 +
               2.0 * ((Pe * Pe) * (Pe * Pe) * Pe) / 945.0;
      else if (Pe > 50.0)
        beta = 1.0 - 1.0 / Pe;
      else
        beta = 1.0 / tanh(Pe) - 1.0 / Pe;
      tau = 0.5 * h_stream / vmag * beta;
    } else {
      const double epos = exp(Pe), eneg = exp(-Pe);
      const double coshPe = 0.5 * (epos + eneg);
      const double sinhPe = 0.5 * (epos - eneg);
      const double cothPe = coshPe / sinhPe;
      tau = 0.5 * h_stream / vmag * (cothPe - 1.0 / Pe);
    }
  }
  for (int i = 0; i < 8; i++) {
    const double vdotg = v[0] * dnx[i] + v[1] * dny[i] + v[2] * dnz[i];
    w[i] = n[i] + tau * vdotg;
  }
  *detj_out = detj;
  if (tau_out) *tau_out = tau;
}

/* el3d, VANADIS.FOR:456-636 (V3 omp_timing.for:431-611 differs only in the
 * Robin source term, which carries no Gauss weights; qs = 0 in both).
 * H = hma (diffusion + DR advection + P*M_w + Robin), C = cma_n (M_w),
 * F = fmv.  Row-major outputs. */
void orc_el3d(int variant, const double xe[8], const double ye[8],
              const double ze[8], const double lam[3], double P,
              const double v[3], double q_in, int robin, double alpha,
              double qs, double H[64], double C[64], double F[8]) {
  double hma[64], cma[64], hmav[64], cma_n[64], fmv[8];
  double n1[8], w[8], dnx[8], dny[8], dnz[8], djac;
  memset(hma, 0, sizeof hma);
  memset(cma, 0, sizeof cma);
  memset(hmav, 0, sizeof hmav);
  memset(cma_n, 0, sizeof cma_n);
  memset(fmv, 0, sizeof fmv);
  double vol = 0.0;
  for (int g = 0; g < 8; g++) {
    orc_gauss_point(variant, GP * KSI[g], GP * ETA[g], GP * ZETA[g], xe, ye, ze,
                    lam, v, &djac, n1, w, dnx, dny, dnz, NULL);
    vol = vol + djac;
  }
  const double qv = q_in / vol;
  for (int g = 0; g < 8; g++) {
    orc_gauss_point(variant, GP * KSI[g], GP * ETA[g], GP * ZETA[g], xe, ye, ze,
                    lam, v, &djac, n1, w, dnx, dny, dnz, NULL);
    const double z1 = djac * lam[0], z2 = djac * lam[1], z3 = djac * lam[2];
    for (int i = 0; i < 8; i++) fmv[i] = fmv[i] + qv * w[i] * djac;
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) {
        const double hx = dnx[i] * dnx[j], hy = dny[i] * dny[j],
                     hz = dnz[i] * dnz[j];
        const double cg = w[i] * n1[j];
        const double hvx = w[i] * dnx[j], hvy = w[i] * dny[j],
                     hvz = w[i] * dnz[j];
        const int k = 8 * i + j;
        hma[k] = hma[k] + hx * z1 + hy * z2 + hz * z3;
        hmav[k] = hmav[k] + (hvx * v[0] + hvy * v[1] + hvz * v[2]) * djac;
        cma[k] = cma[k] + cg * djac * P;
        cma_n[k] = cma_n[k] + cg * djac;
      }
  }
  if (robin) {
    /* Robin on face 3 (xi = -1), VANADIS.FOR:579-627 */
    double zx1[4], zx2[4], zx3[4];
    for (int q = 0; q < 4; q++) {
      zx1[q] = xe[I_POW[2][q] - 1];
      zx2[q] = ye[I_POW[2][q] - 1];
      zx3[q] = ze[I_POW[2][q] - 1];
    }
    const double gp2[2] = {-GP, GP};
    const double gw = 1.0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const double ze1 = -1.0, ze2 = gp2[a], ze3 = gp2[b];
        double detjs;
        orc_quad4_3d(zx1, zx2, zx3, ze2, ze3, &detjs);
        orc_gauss_point(variant, ze1, ze2, ze3, xe, ye, ze, lam, v, &djac, n1,
                        w, dnx, dny, dnz, NULL);
        const double z_nic = detjs * gw * gw * alpha;
        for (int i = 0; i < 8; i++) {
          if (variant == ORC_V3)
            fmv[i] = fmv[i] - qs * w[i] * detjs;
          else
            fmv[i] = fmv[i] - qs * w[i] * detjs * gw * gw;
          for (int j = 0; j < 8; j++) {
            const double cg = w[i] * n1[j];
            cma[8 * i + j] = cma[8 * i + j] + cg * z_nic;
          }
        }
      }
  }
  for (int k = 0; k < 64; k++) H[k] = hma[k] + cma[k] + hmav[k];
  memcpy(C, cma_n, sizeof cma_n);
  memcpy(F, fmv, sizeof fmv);
}

/* przest3d element loop, VANADIS.FOR:713-771 (theta scheme, serial order).
 * scheme ORC_STEADY: A = H, b = F.  ORC_EULER: A = H + C/dt,
 * b = F + (C/dt) T_old.  Returns 0. */
int orc_assemble(int variant, int scheme, double dt, int nax, int nay, int naz,
                 const double *xx, const double *yy, const double *zz,
                 const int *iss, const int *kod, const double *lam, double P,
                 const double *vel, const double *q_in, double alpha, double qs,
                 const double *t_old, double *bb, double *fmv3, double *diag) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  for (int i = 0; i < N; i++) {
    diag[i] = 0.0;
    fmv3[i] = 0.0;
  }
  for (int e = 0; e < E; e++) {
    int nd[8];
    double xe[8], ye[8], ze[8];
    for (int j = 0; j < 8; j++) {
      nd[j] = iss[j * E + e];
      xe[j] = xx[nd[j] - 1];
      ye[j] = yy[nd[j] - 1];
      ze[j] = zz[nd[j] - 1];
    }
    const double l3[3] = {lam[e], lam[e], lam[e]};
    const double v[3] = {vel[e], vel[E + e], vel[2 * (size_t)E + e]};
    const int robin = (kod[2 * E + e] == 3);
    double H[64], Cn[64], F[8], A[64], f2[8];
    orc_el3d(variant, xe, ye, ze, l3, P, v, q_in ? q_in[e] : 0.0, robin, alpha,
             qs, H, Cn, F);
    double wt[8];
    for (int j = 0; j < 8; j++) wt[j] = t_old ? t_old[nd[j] - 1] : 0.0;
    if (scheme == ORC_THETA) {
      double ht[64], ft[8];
      for (int k = 0; k < 64; k++) ht[k] = H[k] - 3.0 / dt * Cn[k];
      for (int i = 0; i < 8; i++) {
        ft[i] = 0.0;
        for (int k = 0; k < 8; k++) ft[i] = ft[i] + ht[8 * i + k] * wt[k];
      }
      for (int i = 0; i < 8; i++) f2[i] = 3.0 * F[i] - ft[i];
      for (int k = 0; k < 64; k++) A[k] = 2.0 * H[k] + 3.0 / dt * Cn[k];
    } else if (scheme == ORC_STEADY) {
      for (int i = 0; i < 8; i++) f2[i] = F[i];
      for (int k = 0; k < 64; k++) A[k] = H[k];
    } else {
      double ft[8];
      for (int i = 0; i < 8; i++) {
        ft[i] = 0.0;
        for (int k = 0; k < 8; k++)
          ft[i] = ft[i] + (1.0 / dt * Cn[8 * i + k]) * wt[k];
      }
      for (int i = 0; i < 8; i++) f2[i] = F[i] + ft[i];
      for (int k = 0; k < 64; k++) A[k] = H[k] + 1.0 / dt * Cn[k];
    }
    memcpy(bb + 64 * (size_t)e, A, sizeof A);
    for (int j = 0; j < 8; j++) fmv3[nd[j] - 1] = fmv3[nd[j] - 1] + f2[j];
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++)
        if (nd[i] == nd[j]) diag[nd[i] - 1] = diag[nd[i] - 1] + A[8 * i + j];
  }
  return 0;
}

/* matvec_0, VANADIS.FOR:2149-2193: y = (I-D) A (I-D) x, zeroing xin in place */
void orc_matvec_0(double *xin, double *xout, const int *iss, const double *bb,
                  int nax, int nay, int naz, const int *mask) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  for (int j = 0; j < N; j++) xout[j] = 0.0;
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xin[j] = 0.0;
  size_t ik = 0;
  for (int e = 0; e < E; e++) {
    int nd[8];
    for (int j = 0; j < 8; j++) nd[j] = iss[j * E + e] - 1;
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) {
        xout[nd[i]] = xout[nd[i]] + bb[ik] * xin[nd[j]];
        ik++;
      }
  }
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xout[j] = 0.0;
}

/* matvec_0 with the elements in 8 parity colours, colour (i&1) | (j&1)<<1 |
 * (k&1)<<2 of the element's lowest node (i, j, k): elements of one colour
 * share no node, so a colour's elements run in parallel (OpenMP, `threads`
 * threads) and each node receives its contributions colour by colour -- the
 * CPU counterpart of the reference's coloured scatter (coloring.cu:50-106,
 * 289-360) over the stored element matrices of VANADIS.FOR:2149-2193.  Same
 * result as orc_matvec_0 up to the order of a node's additions, and the same
 * bits for every thread count.  Returns 0, or -1 if out of memory. */
int orc_matvec_0_colored(double *xin, double *xout, const int *iss, const double *bb,
                         int nax, int nay, int naz, const int *mask, int threads) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  int *list = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
  if (!list) return -1;
  int start[9] = {0};
  for (int e = 0; e < E; e++) {
    const int n0 = iss[e] - 1;
    const int c = (n0 % nax & 1) | ((n0 / nax) % nay & 1) << 1 | (n0 / (nax * nay) & 1) << 2;
    start[c + 1]++;
  }
  for (int c = 0; c < 8; c++) start[c + 1] += start[c];
  {
    int fill[8];
    for (int c = 0; c < 8; c++) fill[c] = start[c];
    for (int e = 0; e < E; e++) {
      const int n0 = iss[e] - 1;
      const int c = (n0 % nax & 1) | ((n0 / nax) % nay & 1) << 1 | (n0 / (nax * nay) & 1) << 2;
      list[fill[c]++] = e;
    }
  }
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int j = 0; j < N; j++) {
    xout[j] = 0.0;
    if (mask[j] == 1) xin[j] = 0.0;
  }
  for (int c = 0; c < 8; c++) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int q = start[c]; q < start[c + 1]; q++) {
      const int e = list[q];
      int nd[8];
      for (int j = 0; j < 8; j++) nd[j] = iss[j * E + e] - 1;
      const double *b3 = bb + 64 * (size_t)e; /* b3(i,j) = b3[8i + j] */
      for (int i = 0; i < 8; i++) {
        double s = xout[nd[i]];
        for (int j = 0; j < 8; j++) s = s + b3[8 * i + j] * xin[nd[j]];
        xout[nd[i]] = s;
      }
    }
  }
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xout[j] = 0.0;
  free(list);
  return 0;
}

/* mttvec_0, VANADIS.FOR:2195-2245: y = (I-D) A^T (I-D) x */
void orc_mttvec_0(double *xin, double *xout, const int *iss, const double *bb,
                  int nax, int nay, int naz, const int *mask) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  for (int j = 0; j < N; j++) xout[j] = 0.0;
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xin[j] = 0.0;
  for (int e = 0; e < E; e++) {
    int nd[8];
    for (int j = 0; j < 8; j++) nd[j] = iss[j * E + e] - 1;
    const double *b3 = bb + 64 * (size_t)e; /* b3(i,j) = b3[8i + j] */
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++)
        xout[nd[i]] = xout[nd[i]] + b3[8 * j + i] * xin[nd[j]];
  }
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xout[j] = 0.0;
}

/* dnrm2, VANADIS.FOR:1078-1087 (no scaling) */
double orc_dnrm2(int n, const double *x) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + x[i] * x[i];
  return sqrt(s);
}

/* dbcg_simple, VANADIS.FOR:977-1055: Jacobi-preconditioned BiCG.
 * Returns 0 when converged, 1 when itmax was reached. */
int orc_dbcg_simple(int n, const double *b, double *x, double tol, int itmax,
                    int *iter, double *err, const double *bb, const int *iss,
                    const double *diag, int nax, int nay, int naz,
                    const int *mask) {
  double *p = calloc((size_t)n, sizeof(double));
  double *pp = calloc((size_t)n, sizeof(double));
  double *r = calloc((size_t)n, sizeof(double));
  double *rr = calloc((size_t)n, sizeof(double));
  double *z = calloc((size_t)n, sizeof(double));
  double *zz = calloc((size_t)n, sizeof(double));
  int rc = 1;
  double bkden = 0.0;
  orc_matvec_0(x, r, iss, bb, nax, nay, naz, mask);
  for (int i = 0; i < n; i++) {
    r[i] = b[i] - r[i];
    rr[i] = r[i];
  }
  for (int i = 0; i < n; i++) z[i] = r[i] / diag[i];
  for (int i = 0; i < n; i++) zz[i] = rr[i] / diag[i];
  const double bnrm = orc_dnrm2(n, b);
  *err = orc_dnrm2(n, r) / bnrm;
  if (*err <= tol) {
    rc = 0;
    goto done;
  }
  for (int k = 1; k <= itmax; k++) {
    *iter = k;
    double bknum = 0.0;
    for (int i = 0; i < n; i++) bknum = bknum + z[i] * rr[i];
    if (k == 1) {
      for (int i = 0; i < n; i++) {
        p[i] = z[i];
        pp[i] = zz[i];
      }
    } else {
      const double bk = bknum / bkden;
      for (int i = 0; i < n; i++) {
        p[i] = z[i] + bk * p[i];
        pp[i] = zz[i] + bk * pp[i];
      }
    }
    bkden = bknum;
    orc_matvec_0(p, z, iss, bb, nax, nay, naz, mask);
    double akden = 0.0;
    for (int i = 0; i < n; i++) akden = akden + z[i] * pp[i];
    const double ak = bknum / akden;
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + ak * p[i];
      r[i] = r[i] - ak * z[i];
    }
    orc_mttvec_0(pp, zz, iss, bb, nax, nay, naz, mask);
    for (int i = 0; i < n; i++) rr[i] = rr[i] - ak * zz[i];
    for (int i = 0; i < n; i++) z[i] = r[i] / diag[i];
    for (int i = 0; i < n; i++) zz[i] = rr[i] / diag[i];
    *err = orc_dnrm2(n, r) / bnrm;
    if (*err <= tol) {
      rc = 0;
      break;
    }
  }
done:
  free(p);
  free(pp);
  free(r);
  free(rr);
  free(z);
  free(zz);
  return rc;
}

/* integrate_mass_hex8, VANADIS.FOR:2493-2561 (+ jacobian_hex8 2564-2598) */
double orc_integrate_mass_hex8(int nelem, const int *iss, const double *xx,
                               const double *yy, const double *zz,
                               const double *c) {
  const double xg[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
  double mass = 0.0;
  for (int e = 0; e < nelem; e++) {
    double xc[8], yc[8], zc[8], cc[8];
    for (int a = 0; a < 8; a++) {
      const int nd = iss[a * nelem + e] - 1;
      xc[a] = xx[nd];
      yc[a] = yy[nd];
      zc[a] = zz[nd];
      cc[a] = c[nd];
    }
    for (int ig = 0; ig < 2; ig++)
      for (int jg = 0; jg < 2; jg++)
        for (int kg = 0; kg < 2; kg++) {
          const double wgt = 1.0 * 1.0 * 1.0;
          double n[8], dn[24];
          orc_shape_hex8(xg[ig], xg[jg], xg[kg], n, dn);
          double J[9] = {0};
          for (int a = 0; a < 8; a++) {
            J[0] = J[0] + dn[3 * a + 0] * xc[a];
            J[1] = J[1] + dn[3 * a + 1] * xc[a];
            J[2] = J[2] + dn[3 * a + 2] * xc[a];
            J[3] = J[3] + dn[3 * a + 0] * yc[a];
            J[4] = J[4] + dn[3 * a + 1] * yc[a];
            J[5] = J[5] + dn[3 * a + 2] * yc[a];
            J[6] = J[6] + dn[3 * a + 0] * zc[a];
            J[7] = J[7] + dn[3 * a + 1] * zc[a];
            J[8] = J[8] + dn[3 * a + 2] * zc[a];
          }
          const double detJ = J[0] * (J[4] * J[8] - J[5] * J[7]) -
                              J[1] * (J[3] * J[8] - J[5] * J[6]) +
                              J[2] * (J[3] * J[7] - J[4] * J[6]);
          double cv = 0.0;
          for (int a = 0; a < 8; a++) cv = cv + n[a] * cc[a];
          mass = mass + cv * detJ * wgt;
        }
  }
  return mass;
}

/* diagonal_dominance_factor, VANADIS.FOR:2434-2489 */
double orc_diagonal_dominance_factor(int nelem, int nnode, const int *iss,
                                     const double *bb) {
  double *dg = calloc((size_t)nnode, sizeof(double));
  double *sr = calloc((size_t)nnode, sizeof(double));
  for (int e = 0; e < nelem; e++)
    for (int i = 0; i < 8; i++) {
      const int gi = iss[i * nelem + e] - 1;
      for (int j = 0; j < 8; j++) {
        const int gj = iss[j * nelem + e] - 1;
        const double aij = bb[64 * (size_t)e + 8 * i + j];
        if (gi == gj)
          dg[gi] = dg[gi] + fabs(aij);
        else
          sr[gi] = sr[gi] + fabs(aij);
      }
    }
  double M = 0.0;
  for (int i = 0; i < nnode; i++)
    if (dg[i] > 0.0) {
      const double ti = sr[i] / dg[i];
      if (ti > M) M = ti;
    }
  free(dg);
  free(sr);
  return M;
}

/* is_diagonally_dominant, VANADIS.FOR:2248-2303 */
int orc_is_diagonally_dominant(int nelem, int nnode, const int *iss,
                               const double *bb) {
  double *dg = calloc((size_t)nnode, sizeof(double));
  double *sr = calloc((size_t)nnode, sizeof(double));
  for (int e = 0; e < nelem; e++)
    for (int i = 0; i < 8; i++) {
      const int gi = iss[i * nelem + e] - 1;
      for (int j = 0; j < 8; j++) {
        const int gj = iss[j * nelem + e] - 1;
        const double aij = bb[64 * (size_t)e + 8 * i + j];
        if (gi == gj)
          dg[gi] = dg[gi] + fabs(aij);
        else
          sr[gi] = sr[gi] + fabs(aij);
      }
    }
  int dd = 1;
  for (int i = 0; i < nnode; i++)
    if (dg[i] <= sr[i]) {
      dd = 0;
      break;
    }
  free(dg);
  free(sr);
  return dd;
}

/* generate_hex8_colors, cudatest_coloring_gtx1060_optimized.cu:50-74 */
void orc_generate_hex8_colors(int nax, int nay, int naz, int *color) {
  for (int k = 0; k < naz - 1; k++)
    for (int j = 0; j < nay - 1; j++)
      for (int i = 0; i < nax - 1; i++) {
        const int e = i + (nax - 1) * j + (nax - 1) * (nay - 1) * k;
        color[e] = (i & 1) + 2 * (j & 1) + 4 * (k & 1);
      }
}

/* build_color_element_lists, cudatest_coloring_gtx1060_optimized.cu:82-106 */
void orc_build_color_element_lists(int ile, const int *color, int *elems,
                                   int offsets[9], int counts[8]) {
  int cursor[8];
  for (int c = 0; c < 8; c++) counts[c] = 0;
  for (int e = 0; e < ile; e++) counts[color[e]]++;
  offsets[0] = 0;
  for (int c = 0; c < 8; c++) offsets[c + 1] = offsets[c] + counts[c];
  for (int c = 0; c < 8; c++) cursor[c] = offsets[c];
  for (int e = 0; e < ile; e++) elems[cursor[color[e]]++] = e;
}
