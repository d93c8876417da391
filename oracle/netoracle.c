/* CPU oracle of the learned-guess network (test infrastructure only: tests/
 * call it through oracle/pyoracle.py to check the GPU forward and backward
 * kernels; nothing in the product links it).
 *
 * The network: 5 layers of 3x3x3 convolution on the node grid (16 channels,
 * zero "same" padding), leaky ReLU (slope 0.1) after layers 0..3; tap
 * t = dx + 3 dy + 9 dz
 * reads node (x + dx - 1, y + dy - 1, z + dz - 1); weights w[l][t][co][ci],
 * biases b[l][co].  Inputs: channels 0..5 = x1, x1 - x2, src, vx, vy, vz
 * times in_scale[c]; channels 6..15 zero.  Loss over free nodes (mask == 0):
 *   L = sum (out0 - (target - x1 - extrap (x1 - x2)) / out_scale)^2 / count.
 *
 * quant = 1 mirrors the storage points of the GPU kernels: inputs, weights,
 * activations and layer-output gradients are rounded to BF16 (round to
 * nearest even), every sum is taken in FP64 (the GPU accumulates in FP32),
 * the pre-activation is rounded to FP32 before the bias add.  quant = 0 is
 * FP64 throughout (for finite-difference checks of the gradients). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NC 16
#define NL 5
#define NTAP 27
#define WL (NTAP * NC * NC)
#define SLOPE 0.1

static float bf16r(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static double qb(double v, int quant) { return quant ? (double)bf16r((float)v) : v; }

/* one forward layer: y[n][co] = act(b + sum_t sum_ci w[t][co][ci] x[n+off][ci]) */
static void conv_fwd(int nx, int ny, int nz, const double *x, const double *w,
                     const double *b, int relu, int quant, double *y) {
  for (int z = 0; z < nz; ++z)
    for (int yy = 0; yy < ny; ++yy)
      for (int xx = 0; xx < nx; ++xx) {
        const long n = xx + (long)nx * (yy + (long)ny * z);
        for (int co = 0; co < NC; ++co) {
          double s = 0.0;
          for (int t = 0; t < NTAP; ++t) {
            const int ax = xx + t % 3 - 1, ay = yy + (t / 3) % 3 - 1, az = z + t / 9 - 1;
            if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) continue;
            const double *xi = x + (ax + (long)nx * (ay + (long)ny * az)) * NC;
            const double *wt = w + (t * NC + co) * NC;
            for (int ci = 0; ci < NC; ++ci) s += wt[ci] * xi[ci];
          }
          double v = quant ? (double)((float)s + (float)b[co]) : s + b[co];
          if (relu && !(v > 0)) v = quant ? (double)((float)SLOPE * (float)v) : SLOPE * v;
          y[n * NC + co] = v;
        }
      }
}

int orc_net_grads(int nx, int ny, int nz, const double *x1, const double *x2,
                  const double *src, const double *vx, const double *vy,
                  const double *vz, const double *target, const int *mask,
                  const float *in_scale, double out_scale, double extrap, const float *wf,
                  const float *bf, int quant, double *loss, double *dw, double *db,
                  double *out0) {
  const long N = (long)nx * ny * nz;
  double *w = malloc(sizeof(double) * NL * WL), *b = malloc(sizeof(double) * NL * NC);
  double *act[NL + 1];
  for (int l = 0; l <= NL; ++l) act[l] = calloc((size_t)N * NC, sizeof(double));
  double *dz = calloc((size_t)N * NC, sizeof(double)), *dp = calloc((size_t)N * NC, sizeof(double));
  if (!w || !b || !dz || !dp) return 1;
  for (int l = 0; l <= NL; ++l)
    if (!act[l]) return 1;
  for (long i = 0; i < NL * WL; ++i) w[i] = quant ? (double)bf16r(wf[i]) : (double)wf[i];
  for (long i = 0; i < NL * NC; ++i) b[i] = bf[i];
  /* inputs */
  const double *ch[6] = {x1, x2, src, vx, vy, vz};
  for (long n = 0; n < N; ++n)
    for (int c = 0; c < 6; ++c) {
      const double raw = c == 1 ? x1[n] - x2[n] : ch[c][n];
      const double v = raw * (double)in_scale[c];
      act[0][n * NC + c] = quant ? (double)bf16r((float)v) : v;
    }
  /* forward; act[l+1] = output of layer l (stored rounded when quant) */
  for (int l = 0; l < NL; ++l) {
    conv_fwd(nx, ny, nz, act[l], w + l * WL, b + l * NC, l < NL - 1, quant, act[l + 1]);
    if (l < NL - 1 && quant)
      for (long i = 0; i < N * NC; ++i) act[l + 1][i] = qb(act[l + 1][i], 1);
  }
  /* loss and its gradient w.r.t. output channel 0 */
  long count = 0;
  for (long n = 0; n < N; ++n) count += mask[n] == 0;
  if (count < 1) count = 1;
  const double inv_count = quant ? (double)(float)(1.0 / (double)count) : 1.0 / (double)count;
  double L = 0.0;
  memset(dz, 0, sizeof(double) * N * NC);
  for (long n = 0; n < N; ++n) {
    const double o = act[NL][n * NC];
    if (out0) out0[n] = o;
    double d = 0.0;
    if (!mask[n]) {
      const double t = (target[n] - x1[n] - extrap * (x1[n] - x2[n])) * (1.0 / out_scale);
      d = quant ? (double)((float)o - (float)t) : o - t;
    }
    L += d * d;
    dz[n * NC] = qb(2.0 * d * inv_count, quant);
  }
  *loss = L / (double)count;
  /* backward */
  for (int l = NL - 1; l >= 0; --l) {
    const double *xin = act[l];
    double *dwl = dw + l * WL, *dbl = db + l * NC;
    memset(dwl, 0, sizeof(double) * WL);
    memset(dbl, 0, sizeof(double) * NC);
    for (int z = 0; z < nz; ++z)
      for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
          const long n = xx + (long)nx * (yy + (long)ny * z);
          for (int co = 0; co < NC; ++co) dbl[co] += dz[n * NC + co];
          for (int t = 0; t < NTAP; ++t) {
            const int ax = xx + t % 3 - 1, ay = yy + (t / 3) % 3 - 1, az = z + t / 9 - 1;
            if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) continue;
            const double *xi = xin + (ax + (long)nx * (ay + (long)ny * az)) * NC;
            for (int co = 0; co < NC; ++co) {
              const double g = dz[n * NC + co];
              if (g == 0.0) continue;
              double *row = dwl + (t * NC + co) * NC;
              for (int ci = 0; ci < NC; ++ci) row[ci] += g * xi[ci];
            }
          }
        }
    if (l == 0) break;
    /* dX of layer l -> gradient of layer l-1's output, through its leaky ReLU */
    memset(dp, 0, sizeof(double) * N * NC);
    const double *wl = w + l * WL;
    for (int z = 0; z < nz; ++z)
      for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
          const long n = xx + (long)nx * (yy + (long)ny * z);
          for (int t = 0; t < NTAP; ++t) {
            const int ax = xx - (t % 3 - 1), ay = yy - ((t / 3) % 3 - 1), az = z - (t / 9 - 1);
            if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) continue;
            const double *g = dz + (ax + (long)nx * (ay + (long)ny * az)) * NC;
            for (int co = 0; co < NC; ++co) {
              if (g[co] == 0.0) continue;
              const double *wr = wl + (t * NC + co) * NC;
              for (int ci = 0; ci < NC; ++ci) dp[n * NC + ci] += wr[ci] * g[co];
            }
          }
        }
    for (long i = 0; i < N * NC; ++i) {
      const double v = act[l][i] > 0.0 ? dp[i]
                       : (quant ? (double)((float)SLOPE * (float)dp[i]) : SLOPE * dp[i]);
      dz[i] = quant ? (double)bf16r((float)v) : v;
    }
  }
  for (int l = 0; l <= NL; ++l) free(act[l]);
  free(dz);
  free(dp);
  free(w);
  free(b);
  return 0;
}
