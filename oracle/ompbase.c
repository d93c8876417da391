/*
 * OpenMP CPU path of the oracle: the reference's Jacobi-BiCG (dbcg_simple,
 * VANADIS.FOR:977-1055) over stored element matrices with the element
 * products matvec_0 / mttvec_0 (VANADIS.FOR:2149-2245) run colour by colour
 * (the 8 parity colours of cudatest_coloring_gtx1060_optimized.cu:50-106:
 * elements of one colour share no node, so a colour's elements run in
 * parallel), the vector loops and dot products split over threads in fixed
 * static ranges whose partial sums are added in thread order (the same bits
 * for a fixed thread count), and the element assembly of przest3d
 * (VANADIS.FOR:713-771) with the element matrices computed in parallel and
 * scattered into the right-hand side and diagonal in the serial element
 * order (bitwise orc_assemble).
 *
 * TEST INFRASTRUCTURE ONLY: the OpenMP CPU baseline of the benchmark and a
 * faster reference for large systems; never part of the product path.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Element lists of the 8 parity colours, colour (i&1) | (j&1)<<1 | (k&1)<<2
 * of the element's lowest node (i, j, k), in increasing element order. */
static int color_lists(const int *iss, int nax, int nay, int naz, int **list,
                       int start[9]) {
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  *list = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
  if (!*list) return -1;
  int cnt[9] = {0};
  for (int e = 0; e < E; e++) {
    const int n0 = iss[e] - 1;
    const int c = (n0 % nax & 1) | ((n0 / nax) % nay & 1) << 1 | (n0 / (nax * nay) & 1) << 2;
    cnt[c + 1]++;
  }
  start[0] = 0;
  for (int c = 0; c < 8; c++) start[c + 1] = start[c] + cnt[c + 1];
  int fill[8];
  for (int c = 0; c < 8; c++) fill[c] = start[c];
  for (int e = 0; e < E; e++) {
    const int n0 = iss[e] - 1;
    const int c = (n0 % nax & 1) | ((n0 / nax) % nay & 1) << 1 | (n0 / (nax * nay) & 1) << 2;
    (*list)[fill[c]++] = e;
  }
  return 0;
}

/* y = (I-D) A (I-D) x (tr = 0) or (I-D) A^T (I-D) x (tr = 1); xin zeroed in
 * place at the mask like matvec_0 (VANADIS.FOR:2167-2169). */
static void cprod(int tr, double *xin, double *xout, const int *iss, const double *bb,
                  int E, int N, const int *mask, const int *list, const int start[9],
                  int threads) {
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
        if (tr)
          for (int j = 0; j < 8; j++) s = s + b3[8 * j + i] * xin[nd[j]];
        else
          for (int j = 0; j < 8; j++) s = s + b3[8 * i + j] * xin[nd[j]];
        xout[nd[i]] = s;
      }
    }
  }
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int j = 0; j < N; j++)
    if (mask[j] == 1) xout[j] = 0.0;
}

int orc_mttvec_0_colored(double *xin, double *xout, const int *iss, const double *bb,
                         int nax, int nay, int naz, const int *mask, int threads) {
  int *list, start[9];
  if (color_lists(iss, nax, nay, naz, &list, start)) return -1;
  cprod(1, xin, xout, iss, bb, (nax - 1) * (nay - 1) * (naz - 1), nax * nay * naz, mask,
        list, start, threads);
  free(list);
  return 0;
}

/* sum of a[i] * b[i]: thread t adds its static range in order, the thread
 * partials are added in thread order */
static double pdot(int n, const double *a, const double *b, int threads, double *part) {
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const long lo = (long)n * t / nt, hi = (long)n * (t + 1) / nt;
    double s = 0.0;
    for (long i = lo; i < hi; i++) s = s + a[i] * b[i];
    part[t] = s;
    if (t == 0)
      for (int u = nt; u < threads; u++) part[u] = 0.0;
  }
  double s = 0.0;
  for (int t = 0; t < threads; t++) s = s + part[t];
  return s;
}

/* dbcg_simple (VANADIS.FOR:977-1055) with the OpenMP products and loops
 * above: the same recurrence, stopping rule and return convention as
 * orc_dbcg_simple (0 converged, 1 itmax; -1 out of memory). */
int orc_dbcg_simple_omp(int n, const double *b, double *x, double tol, int itmax,
                        int *iter, double *err, const double *bb, const int *iss,
                        const double *diag, int nax, int nay, int naz, const int *mask,
                        int threads) {
  if (threads < 1) threads = 1;
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  int *list, start[9];
  if (color_lists(iss, nax, nay, naz, &list, start)) return -1;
  double *v = calloc((size_t)n * 6, sizeof(double));
  double *part = calloc((size_t)threads, sizeof(double));
  if (!v || !part) {
    free(v);
    free(part);
    free(list);
    return -1;
  }
  double *p = v, *pp = v + n, *r = v + 2 * (size_t)n, *rr = v + 3 * (size_t)n,
         *z = v + 4 * (size_t)n, *zz = v + 5 * (size_t)n;
  int rc = 1;
  double bkden = 0.0;
  cprod(0, x, r, iss, bb, E, n, mask, list, start, threads);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int i = 0; i < n; i++) {
    r[i] = b[i] - r[i];
    rr[i] = r[i];
    z[i] = r[i] / diag[i];
    zz[i] = rr[i] / diag[i];
  }
  const double bnrm = sqrt(pdot(n, b, b, threads, part));
  *err = sqrt(pdot(n, r, r, threads, part)) / bnrm;
  if (*err <= tol) {
    rc = 0;
    goto done;
  }
  for (int k = 1; k <= itmax; k++) {
    *iter = k;
    const double bknum = pdot(n, z, rr, threads, part);
    const double bk = k == 1 ? 0.0 : bknum / bkden;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++) {
      if (k == 1) {
        p[i] = z[i];
        pp[i] = zz[i];
      } else {
        p[i] = z[i] + bk * p[i];
        pp[i] = zz[i] + bk * pp[i];
      }
    }
    bkden = bknum;
    cprod(0, p, z, iss, bb, E, n, mask, list, start, threads);
    const double akden = pdot(n, z, pp, threads, part);
    const double ak = bknum / akden;
    cprod(1, pp, zz, iss, bb, E, n, mask, list, start, threads);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + ak * p[i];
      r[i] = r[i] - ak * z[i];
      rr[i] = rr[i] - ak * zz[i];
      z[i] = r[i] / diag[i];
      zz[i] = rr[i] / diag[i];
    }
    *err = sqrt(pdot(n, r, r, threads, part)) / bnrm;
    if (*err <= tol) {
      rc = 0;
      break;
    }
  }
done:
  free(v);
  free(part);
  free(list);
  return rc;
}

/* przest3d's element loop (VANADIS.FOR:713-771) as orc_assemble, with the
 * element systems computed in parallel and the node sums (right-hand side,
 * diagonal) done afterwards in the serial element order: bitwise
 * orc_assemble for any thread count.  Returns 0, or -1 if out of memory. */
int orc_assemble_omp(int variant, int scheme, double dt, int nax, int nay, int naz,
                     const double *xx, const double *yy, const double *zz,
                     const int *iss, const int *kod, const double *lam, double P,
                     const double *vel, const double *q_in, double alpha, double qs,
                     const double *t_old, double *bb, double *fmv3, double *diag,
                     int threads) {
  if (threads < 1) threads = 1;
  const int E = (nax - 1) * (nay - 1) * (naz - 1);
  const int N = nax * nay * naz;
  double *f2all = malloc(sizeof(double) * 8 * (size_t)(E > 0 ? E : 1));
  if (!f2all) return -1;
#pragma omp parallel for num_threads(threads) schedule(static)
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
    double H[64], Cn[64], F[8], A[64];
    double *f2 = f2all + 8 * (size_t)e;
    orc_el3d(variant, xe, ye, ze, l3, P, v, q_in ? q_in[e] : 0.0, robin, alpha, qs, H, Cn,
             F);
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
        for (int k = 0; k < 8; k++) ft[i] = ft[i] + (1.0 / dt * Cn[8 * i + k]) * wt[k];
      }
      for (int i = 0; i < 8; i++) f2[i] = F[i] + ft[i];
      for (int k = 0; k < 64; k++) A[k] = H[k] + 1.0 / dt * Cn[k];
    }
    memcpy(bb + 64 * (size_t)e, A, sizeof A);
  }
  for (int i = 0; i < N; i++) {
    diag[i] = 0.0;
    fmv3[i] = 0.0;
  }
  for (int e = 0; e < E; e++) {
    int nd[8];
    for (int j = 0; j < 8; j++) nd[j] = iss[j * E + e];
    const double *A = bb + 64 * (size_t)e;
    const double *f2 = f2all + 8 * (size_t)e;
    for (int j = 0; j < 8; j++) fmv3[nd[j] - 1] = fmv3[nd[j] - 1] + f2[j];
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++)
        if (nd[i] == nd[j]) diag[nd[i] - 1] = diag[nd[i] - 1] + A[8 * i + j];
  }
  free(f2all);
  return 0;
}
