// Element math of the matrix-free DR-HEX8 operator, shared by the kernels
// and by a host build used in tests (no HIP headers here).
//
// Box (axis-aligned) HEX8 elements of a tensor grid have a constant
// Jacobian, so the 2x2x2 Gauss rule of el3d (VANADIS.FOR:527-633) integrates
// every term exactly and each element matrix is a combination of Kronecker
// products of 2x2 one-dimensional matrices (mass M, stiffness K, convection
// int phi_a phi_b' and its transpose).  In the per-axis basis
// (s, d) = (u0 + u1, u1 - u0) M and K are diagonal and the convection
// matrices map d -> s and s -> d, so A_e x_e is a 3-stage butterfly, a sparse
// map with ~44 entries and the inverse butterfly.
#pragma once

#ifndef VGPU_HD
#define VGPU_HD __host__ __device__ __forceinline__
#endif

namespace vgpu {

// ---------------------------------------------------------------------------
// Element math.  Local node order (VANADIS.FOR:255-262, par_3d 332-357):
// 0:(0,0,0) 1:(1,0,0) 2:(1,1,0) 3:(0,1,0) 4:(0,0,1) 5:(1,0,1) 6:(1,1,1)
// 7:(0,1,1).  Haar index b = bx + 2*by + 4*bz with 0 = s, 1 = d.
// ---------------------------------------------------------------------------
template <typename T>
struct ElemConst {
  T m[3][2];  // 1-D mass eigenvalues / 2: h/4 (s mode), h/12 (d mode)
  T W[8];     // mass weight of each Haar mode: m_x(bx) m_y(by) m_z(bz)
  T D[8];     // diagonal multiplier: mu + sum over d-mode axes of bra_k
  T g[3];     // d -> s coupling per axis on u = W xh: 6 cH v_k / h_k
  T c[3];     // s -> d coupling per axis: 2 mu tau v_k / h_k
  T t[3];     // symmetric DR cross terms, by free axis: 12 cH tau v_k v_l / (h_k h_l)
};

// Factors of an element column that stay fixed while a kernel marches up z:
// the halved x and y mass eigenvalues of the s mode and the reciprocal sizes.
template <typename T>
struct ColConst {
  T mx0, my0;  // hx / 4, hy / 4
  T ihx, ihy;  // 1 / hx, 1 / hy
};

template <typename T>
VGPU_HD void col_setup(T hx, T hy, T ihx, T ihy, ColConst<T> &q) {
  q.mx0 = hx * T(0.25);
  q.my0 = hy * T(0.25);
  q.ihx = ihx;
  q.ihy = ihy;
}

// Per-element constants of one layer from the column factors, the z size and
// the element's fields.  bra_k = 12 cH (lam + tau v_k^2) / h_k^2 is the K/M
// eigenvalue ratio of the d mode times the diffusion plus DR streamline
// diffusion coefficient.
template <typename T>
VGPU_HD void elem_layer_setup(const ColConst<T> &q, T hz, T ihz, T lam, T vx,
                              T vy, T vz, T tau, T cH, T mu,
                              ElemConst<T> &k) {
  const T third = T(1.0 / 3.0);
  k.m[0][0] = q.mx0;
  k.m[0][1] = q.mx0 * third;
  k.m[1][0] = q.my0;
  k.m[1][1] = q.my0 * third;
  k.m[2][0] = hz * T(0.25);
  k.m[2][1] = hz * T(1.0 / 12.0);
  const T tvx = tau * vx, tvy = tau * vy, tvz = tau * vz;
  const T c12 = T(12) * cH, c6 = T(6) * cH, m2 = T(2) * mu;
  const T brx = (c12 * (q.ihx * q.ihx)) * (lam + tvx * vx);
  const T bry = (c12 * (q.ihy * q.ihy)) * (lam + tvy * vy);
  const T brz = (c12 * (ihz * ihz)) * (lam + tvz * vz);
  k.g[0] = (c6 * q.ihx) * vx;
  k.g[1] = (c6 * q.ihy) * vy;
  k.g[2] = (c6 * ihz) * vz;
  k.c[0] = (m2 * q.ihx) * tvx;
  k.c[1] = (m2 * q.ihy) * tvy;
  k.c[2] = (m2 * ihz) * tvz;
  const T c12z = c12 * ihz;
  k.t[0] = (c12z * q.ihy) * tvy * vz;          // y-z, free axis x
  k.t[1] = (c12z * q.ihx) * tvx * vz;          // x-z, free axis y
  k.t[2] = (c12 * (q.ihx * q.ihy)) * tvx * vy;  // x-y, free axis z
#pragma unroll
  for (int b = 0; b < 8; ++b)
    k.W[b] = (k.m[0][b & 1] * k.m[1][(b >> 1) & 1]) * k.m[2][b >> 2];
  k.D[0] = mu;
  k.D[1] = mu + brx;
  k.D[2] = mu + bry;
  k.D[3] = k.D[1] + bry;
#pragma unroll
  for (int b = 0; b < 4; ++b) k.D[b + 4] = k.D[b] + brz;
}

// All per-element constants from the sizes h, reciprocals ih and the fields.
template <typename T>
VGPU_HD void elem_setup(T hx, T hy, T hz, T ihx, T ihy, T ihz, T lam, T vx, T vy,
                        T vz, T tau, T cH, T mu, ElemConst<T> &k) {
  ColConst<T> q;
  col_setup<T>(hx, hy, ihx, ihy, q);
  elem_layer_setup<T>(q, hz, ihz, lam, vx, vy, vz, tau, cH, mu, k);
}

// forward per-axis butterflies: nodal x[8] -> Haar h[8]
template <typename T>
VGPU_HD void haar_fwd(const T x[8], T h[8]) {
  // x-axis: pairs (0,1) (3,2) (4,5) (7,6) -> index by (by,bz)
  const T a0s = x[0] + x[1], a0d = x[1] - x[0];  // by=0,bz=0
  const T a1s = x[3] + x[2], a1d = x[2] - x[3];  // by=1,bz=0
  const T a2s = x[4] + x[5], a2d = x[5] - x[4];  // by=0,bz=1
  const T a3s = x[7] + x[6], a3d = x[6] - x[7];  // by=1,bz=1
  // y-axis
  const T b0 = a0s + a1s, b1 = a0d + a1d, b2 = a1s - a0s, b3 = a1d - a0d;  // bz=0
  const T b4 = a2s + a3s, b5 = a2d + a3d, b6 = a3s - a2s, b7 = a3d - a2d;  // bz=1
  // z-axis
  h[0] = b0 + b4;
  h[1] = b1 + b5;
  h[2] = b2 + b6;
  h[3] = b3 + b7;
  h[4] = b4 - b0;
  h[5] = b5 - b1;
  h[6] = b6 - b2;
  h[7] = b7 - b3;
}

// inverse butterflies (scale folded into the operator): Haar h -> nodal y
template <typename T>
VGPU_HD void haar_inv(const T h[8], T y[8]) {
  // z-axis
  const T c0 = h[0] - h[4], c4 = h[0] + h[4];
  const T c1 = h[1] - h[5], c5 = h[1] + h[5];
  const T c2 = h[2] - h[6], c6 = h[2] + h[6];
  const T c3 = h[3] - h[7], c7 = h[3] + h[7];
  // y-axis (index bx + 2 by within a z-slab)
  const T d0 = c0 - c2, d2 = c0 + c2, d1 = c1 - c3, d3 = c1 + c3;  // az=0
  const T d4 = c4 - c6, d6 = c4 + c6, d5 = c5 - c7, d7 = c5 + c7;  // az=1
  // x-axis: (s,d) -> (u0,u1)
  y[0] = d0 - d1;
  y[1] = d0 + d1;
  y[3] = d2 - d3;
  y[2] = d2 + d3;
  y[4] = d4 - d5;
  y[5] = d4 + d5;
  y[7] = d6 - d7;
  y[6] = d6 + d7;
}

// The butterflies factor into a 2-D (x, y) part per horizontal face and a
// z stage per element, so a kernel marching up z transforms each node plane
// once: the top face of one layer is the bottom face of the next, and the
// two faces' contributions to a plane are added before one inverse.
// face_fwd: corners in local order (0,0) (1,0) (1,1) (0,1) -> F[bx + 2 by].
template <typename T>
VGPU_HD void face_fwd(const T n[4], T F[4]) {
  const T as = n[0] + n[1], ad = n[1] - n[0];  // y = 0 edge
  const T bs = n[3] + n[2], bd = n[2] - n[3];  // y = 1 edge
  F[0] = as + bs;
  F[1] = ad + bd;
  F[2] = bs - as;
  F[3] = bd - ad;
}

// transposed 2-D butterfly: face Haar values -> contributions to the corners
template <typename T>
VGPU_HD void face_inv(const T S[4], T y[4]) {
  const T d0 = S[0] - S[2], d2 = S[0] + S[2], d1 = S[1] - S[3], d3 = S[1] + S[3];
  y[0] = d0 - d1;
  y[1] = d0 + d1;
  y[3] = d2 - d3;
  y[2] = d2 + d3;
}

// z stage: faces -> element Haar xh[b], b = bx + 2 by + 4 bz
template <typename T>
VGPU_HD void zfwd(const T fb[4], const T ft[4], T xh[8]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    xh[i] = fb[i] + ft[i];
    xh[i + 4] = ft[i] - fb[i];
  }
}

// transposed z stage: element Haar output -> bottom and top face values
template <typename T>
VGPU_HD void zinv(const T yh[8], T ob[4], T ot[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ob[i] = yh[i] - yh[i + 4];
    ot[i] = yh[i] + yh[i + 4];
  }
}

// Transposed element operator in Haar space.  With A_e = H^T B W H (B the
// Haar-space map haar_apply uses on u = W xh), A_e^T = H^T W B^T H: apply
// B^T to xh, then the mass weights.
template <typename T>
VGPU_HD void haar_apply_tr(const ElemConst<T> &k, const T xh[8], T yh[8]) {
  T v[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) v[b] = k.D[b] * xh[b];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int bit = 1 << ax;
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
#pragma unroll
    for (int pq = 0; pq < 4; ++pq) {
      const int b0 = ((pq & 1) << o1) | ((pq >> 1) << o2);
      const int b1 = b0 | bit;
      v[b0] += k.c[ax] * xh[b1];
      v[b1] += k.g[ax] * xh[b0];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int kx = (j == 0) ? 1 : 0, lx = (j == 2) ? 1 : 2;
    const int bk = 1 << kx, bl = 1 << lx, bj = 1 << j;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const int base = f ? bj : 0;
      const int i1 = base | bl, i2 = base | bk;
      const T u1 = xh[i1], u2 = xh[i2];
      v[i2] += k.t[j] * u1;
      v[i1] += k.t[j] * u2;
    }
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) yh[b] = k.W[b] * v[b];
}

// The Robin term of robin_add in Haar space, added to the element's Haar
// output yh (before zinv).  With i = by + 2 bz, the x-min face's 2-D Haar is
// w0[i] = (xh[2i] - xh[2i+1]) / 2; nodal face contributions (f0, f1) on the
// x = 0 and x = 1 faces are yh[2i] += (f0 + f1) / 2, yh[2i+1] += (f1 - f0) / 2.
template <typename T, bool TR>
VGPU_HD void robin_haar(const ElemConst<T> &k, T ihx, T tau, T vx, T vy, T vz,
                        T ra, const T xh[8], T yh[8]) {
  const T rx = tau * vx * ihx, hr = T(0.5) * ra;
  const T cy = tau * vy * T(0.5), cz = tau * vz * T(0.5);
  T w0[4], mm[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    w0[i] = T(0.5) * (xh[2 * i] - xh[2 * i + 1]);
    mm[i] = k.m[1][i & 1] * k.m[2][i >> 1];
  }
  if (!TR) {
    // f0 = (1 - rx) M w0 + C w0, f1 = rx M w0
    T cc[4];
    cc[0] = T(0);
    cc[1] = cy * k.m[2][0] * w0[0];
    cc[2] = cz * k.m[1][0] * w0[0];
    cc[3] = cy * k.m[2][1] * w0[2] + cz * k.m[1][1] * w0[1];
    const T r2 = T(2) * rx - T(1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const T t1 = mm[i] * w0[i];
      yh[2 * i] += hr * (t1 + cc[i]);
      yh[2 * i + 1] += hr * (r2 * t1 - cc[i]);
    }
  } else {
    // f0 = M w0 + rx M (w1 - w0) + C^T w0, f1 = 0; w1 - w0 = xh[2i+1]
    T o0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) o0[i] = mm[i] * w0[i] + rx * (mm[i] * xh[2 * i + 1]);
    o0[0] += cy * k.m[2][0] * w0[1];
    o0[2] += cy * k.m[2][1] * w0[3];
    o0[0] += cz * k.m[1][0] * w0[2];
    o0[1] += cz * k.m[1][1] * w0[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      yh[2 * i] += hr * o0[i];
      yh[2 * i + 1] -= hr * o0[i];
    }
  }
}

// Haar-space operator on u = W xh; TR = transpose (the diagonal and the DR
// cross terms are symmetric on u; the per-axis couplings are not).
template <typename T, bool TR>
VGPU_HD void haar_apply(const ElemConst<T> &k, const T xh[8],
                                           T yh[8]) {
  T u[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    u[b] = k.W[b] * xh[b];
    yh[b] = k.D[b] * u[b];
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int bit = 1 << ax;
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    // on u the s and d modes carry mass weights h/4 and h/12, so the
    // transpose exchanges the couplings with a factor 3 (and 1/3)
    const T g = TR ? T(3) * k.c[ax] : k.g[ax];
    const T c = TR ? k.g[ax] * T(1.0 / 3.0) : k.c[ax];
#pragma unroll
    for (int pq = 0; pq < 4; ++pq) {
      const int b0 = ((pq & 1) << o1) | ((pq >> 1) << o2);  // s on axis ax
      const int b1 = b0 | bit;                             // d on axis ax
      yh[b0] += g * u[b1];
      yh[b1] += c * u[b0];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int kx = (j == 0) ? 1 : 0, lx = (j == 2) ? 1 : 2;
    const int bk = 1 << kx, bl = 1 << lx, bj = 1 << j;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const int base = f ? bj : 0;
      const int i1 = base | bl, i2 = base | bk;  // (s,d) and (d,s) on (k,l)
      const T u1 = u[i1], u2 = u[i2];
      yh[i2] += k.t[j] * u1;
      yh[i1] += k.t[j] * u2;
    }
  }
}

// Robin term on the x-min face (face 3, VANADIS.FOR:579-627):
// R = ra [E0 (x) M (x) M + tau(vx E0' (x) M (x) M + vy E0 (x) C^T (x) M
//                                + vz E0 (x) M (x) C^T)], added to nodal y.
template <typename T, bool TR>
VGPU_HD void robin_add(const ElemConst<T> &k, T hx, T tau,
                                          T vx, T vy, T vz, T ra, const T x[8],
                                          T y[8]) {
  // 2-D Haar of a face slice w[ay][az] given as (n00, n10, n01, n11)
  auto h2 = [](T n00, T n10, T n01, T n11, T o[4]) {
    const T ps = n00 + n10, pd = n10 - n00, qs = n01 + n11, qd = n11 - n01;
    o[0] = ps + qs;  // [s][s]  index by + 2 bz
    o[1] = pd + qd;  // [d][s]
    o[2] = qs - ps;  // [s][d]
    o[3] = qd - pd;  // [d][d]
  };
  auto ih2 = [](const T o[4], T &n00, T &n10, T &n01, T &n11) {
    const T v00 = o[0] - o[2], v01 = o[0] + o[2];  // by = s: az = 0, 1
    const T v10 = o[1] - o[3], v11 = o[1] + o[3];  // by = d
    n00 = v00 - v10;
    n10 = v00 + v10;
    n01 = v01 - v11;
    n11 = v01 + v11;
  };
  const T rx = tau * vx / hx;
  T w0[4], o0[4];
  h2(x[0], x[3], x[4], x[7], w0);
  T mm[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) mm[i] = k.m[1][i & 1] * k.m[2][i >> 1];
  const T cy = tau * vy * T(0.5), cz = tau * vz * T(0.5);
  if (!TR) {
    T o1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const T t1 = mm[i] * w0[i];
      o0[i] = t1 - rx * t1;
      o1[i] = rx * t1;
    }
    // C^T in y: s -> d ; C^T in z: s -> d
    o0[1] += cy * k.m[2][0] * w0[0];
    o0[3] += cy * k.m[2][1] * w0[2];
    o0[2] += cz * k.m[1][0] * w0[0];
    o0[3] += cz * k.m[1][1] * w0[1];
    T a, b, c, d;
    ih2(o0, a, b, c, d);
    y[0] += ra * a;
    y[3] += ra * b;
    y[4] += ra * c;
    y[7] += ra * d;
    ih2(o1, a, b, c, d);
    y[1] += ra * a;
    y[2] += ra * b;
    y[5] += ra * c;
    y[6] += ra * d;
  } else {
    T w1[4];
    h2(x[1], x[2], x[5], x[6], w1);
#pragma unroll
    for (int i = 0; i < 4; ++i) o0[i] = mm[i] * w0[i] + rx * (mm[i] * (w1[i] - w0[i]));
    // C in y: d -> s ; C in z: d -> s
    o0[0] += cy * k.m[2][0] * w0[1];
    o0[2] += cy * k.m[2][1] * w0[3];
    o0[0] += cz * k.m[1][0] * w0[2];
    o0[1] += cz * k.m[1][1] * w0[3];
    T a, b, c, d;
    ih2(o0, a, b, c, d);
    y[0] += ra * a;
    y[3] += ra * b;
    y[4] += ra * c;
    y[7] += ra * d;
  }
}

}  // namespace vgpu
