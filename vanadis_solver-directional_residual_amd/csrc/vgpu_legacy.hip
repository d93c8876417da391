// Legacy drop-in: kernel_dbcg_simple_ with the reference's argument list and
// semantics (src_v2026_2_1/cudatest.cu:301-602, colour variant
// cudatest_coloring_gtx1060_optimized.cu:434-803): Jacobi-PBiCG on
// (I-D) A (I-D) with A given as stored 8x8 element matrices BB.
//
// MI355X organisation: the dual product (A p and A^T pp) reads each element
// matrix once instead of twice; the colour variant uses the reference's
// colouring and compact per-colour element lists, bit for bit
// (generate_hex8_colors / build_color_element_lists, lines 50-106); the BiCG
// is the graph-captured engine of vgpu_bicg.hpp; device buffers persist
// between calls of the same size instead of 19 allocations per call.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "vgpu_bicg.hpp"
#include "vgpu_internal.hpp"

namespace vgpu {
namespace {

// Element products on stored BB (row-major per element: BB[64e + 8i + j]).
// y1 += A_e x1 (rows), y2 += A_e^T x2 (columns); masked in and out.
template <bool ATOM, bool DUAL>
__global__ __launch_bounds__(256) void k_bb(int64_t E, const int *__restrict__ iss,
                                            const double *__restrict__ bb,
                                            const uint8_t *__restrict__ mask,
                                            const int *__restrict__ elems,
                                            int64_t count, const double *x1,
                                            double *y1, const double *x2,
                                            double *y2, const int *done) {
  if (done && *done) return;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = elems ? elems[t] : t;
    int nd[8];
    bool mk[8];
    double xa[8], xb[8], ya[8], yb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      nd[j] = iss[j * E + e] - 1;
      mk[j] = mask[nd[j]] != 0;
      xa[j] = mk[j] ? 0.0 : x1[nd[j]];
      xb[j] = (DUAL && !mk[j]) ? x2[nd[j]] : 0.0;
      yb[j] = 0.0;
    }
    const double2 *row = reinterpret_cast<const double2 *>(bb + 64 * e);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double a[8];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const double2 v = row[4 * i + h];
        a[2 * h] = v.x;
        a[2 * h + 1] = v.y;
      }
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j) s += a[j] * xa[j];
      ya[i] = s;
      if (DUAL) {
#pragma unroll
        for (int j = 0; j < 8; ++j) yb[j] += a[j] * xb[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (mk[i]) continue;
      if (ATOM) {
        atomicAdd(&y1[nd[i]], ya[i]);
        if (DUAL) atomicAdd(&y2[nd[i]], yb[i]);
      } else {
        y1[nd[i]] += ya[i];
        if (DUAL) y2[nd[i]] += yb[i];
      }
    }
  }
}

__global__ void k_zero(const int *done, double *a, double *b, int64_t n) {
  if (done && *done) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i] = 0.0;
    if (b) b[i] = 0.0;
  }
}

__global__ void k_mask_from_int(int64_t n, const int *bc, uint8_t *m) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m[i] = bc[i] == 1 ? 1 : 0;
}

// Colouring and compact per-colour element lists of the colour variant
// (generate_hex8_colors / build_color_element_lists, coloring.cu:50-106):
// colour (i&1) + 2(j&1) + 4(k&1) of element (i, j, k), numbered
// e = i + (nax-1) j + (nax-1)(nay-1) k; the lists are a stable counting sort
// by colour.  color and elems hold E entries, coff 9, ccnt 8.
void colour_lists(int nax, int nay, int naz, int *color, int *elems, int coff[9],
                  int ccnt[8]) {
  const int64_t E = (int64_t)(nax - 1) * (nay - 1) * (naz - 1);
  for (int k = 0; k < naz - 1; ++k)
    for (int j = 0; j < nay - 1; ++j)
      for (int i = 0; i < nax - 1; ++i) {
        const int64_t e = i + (int64_t)(nax - 1) * j + (int64_t)(nax - 1) * (nay - 1) * k;
        color[e] = (i & 1) + 2 * (j & 1) + 4 * (k & 1);
      }
  int cursor[8];
  for (int c = 0; c < 8; ++c) ccnt[c] = 0;
  for (int64_t e = 0; e < E; ++e) ccnt[color[e]]++;
  coff[0] = 0;
  for (int c = 0; c < 8; ++c) coff[c + 1] = coff[c] + ccnt[c];
  for (int c = 0; c < 8; ++c) cursor[c] = coff[c];
  for (int64_t e = 0; e < E; ++e) elems[cursor[color[e]]++] = (int)e;
}

struct Legacy {
  int N = 0;
  int64_t E = 0;
  int nax = 0, nay = 0, naz = 0;
  hipStream_t s = nullptr;
  double *x = nullptr, *b = nullptr, *bb = nullptr, *diag = nullptr;
  int *iss = nullptr, *bc = nullptr, *celems = nullptr;
  uint8_t *mask = nullptr;
  int coff[9] = {0}, ccnt[8] = {0};
  Bicg<double> bicg;
  int scatter = -1;  // -1 = not chosen yet (env / default)
  // wall times of the last call: uploads (incl. the buffer set-up of a new
  // size), the solve, the solution download
  double t_upload = 0, t_solve = 0, t_download = 0;

  void release() {
    bicg.release();
    for (double **p : {&x, &b, &bb, &diag})
      if (*p) {
        hipFree(*p);
        *p = nullptr;
      }
    for (int **p : {&iss, &bc, &celems})
      if (*p) {
        hipFree(*p);
        *p = nullptr;
      }
    if (mask) hipFree(mask);
    mask = nullptr;
    if (s) hipStreamDestroy(s);
    s = nullptr;
    N = 0;
    E = 0;
  }
  int ensure(int N_, int nax_, int nay_, int naz_) {
    const int64_t E_ = (int64_t)(nax_ - 1) * (nay_ - 1) * (naz_ - 1);
    if (N_ == N && E_ == E && nax_ == nax && nay_ == nay && naz_ == naz && s)
      return VGPU_OK;
    const int keep_scatter = scatter;
    release();
    scatter = keep_scatter;
    N = N_;
    E = E_;
    nax = nax_;
    nay = nay_;
    naz = naz_;
    VGPU_HIP(hipStreamCreate(&s));
    VGPU_HIP(hipMalloc(&x, N * sizeof(double)));
    VGPU_HIP(hipMalloc(&b, N * sizeof(double)));
    VGPU_HIP(hipMalloc(&diag, N * sizeof(double)));
    VGPU_HIP(hipMalloc(&bb, 64 * E * sizeof(double)));
    VGPU_HIP(hipMalloc(&iss, 8 * E * sizeof(int)));
    VGPU_HIP(hipMalloc(&bc, N * sizeof(int)));
    VGPU_HIP(hipMalloc(&mask, N));
    VGPU_HIP(hipMalloc(&celems, E * sizeof(int)));
    std::vector<int> color(E), elems(E);
    colour_lists(nax, nay, naz, color.data(), elems.data(), coff, ccnt);
    VGPU_HIP(hipMemcpy(celems, elems.data(), E * sizeof(int), hipMemcpyHostToDevice));
    return bicg.alloc(N, 512);
  }
  int launch_dual(const double *p, double *z, const double *pp, double *zz,
                  const int *done, hipStream_t st) {
    const bool dual = pp != nullptr;
    if (scatter == VGPU_SCATTER_ATOMIC) {
      const unsigned g = (unsigned)std::min<int64_t>(4096, (E + 255) / 256);
      if (dual)
        hipLaunchKernelGGL((k_bb<true, true>), dim3(g), dim3(256), 0, st, E, iss,
                           bb, mask, nullptr, E, p, z, pp, zz, done);
      else
        hipLaunchKernelGGL((k_bb<true, false>), dim3(g), dim3(256), 0, st, E,
                           iss, bb, mask, nullptr, E, p, z, pp, zz, done);
    } else {
      for (int c = 0; c < 8; ++c) {
        if (ccnt[c] == 0) continue;
        const unsigned g = (unsigned)std::min<int64_t>(4096, (ccnt[c] + 255) / 256);
        if (dual)
          hipLaunchKernelGGL((k_bb<false, true>), dim3(g), dim3(256), 0, st, E,
                             iss, bb, mask, celems + coff[c], (int64_t)ccnt[c], p,
                             z, pp, zz, done);
        else
          hipLaunchKernelGGL((k_bb<false, false>), dim3(g), dim3(256), 0, st, E,
                             iss, bb, mask, celems + coff[c], (int64_t)ccnt[c], p,
                             z, pp, zz, done);
      }
    }
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
};

// Process-lifetime cache; never destroyed, so no HIP call runs during static
// destruction after the runtime may already be gone.
Legacy &legacy() {
  static Legacy *L = new Legacy();
  return *L;
}

int legacy_scatter_default() {
  const char *e = std::getenv("VGPU_LEGACY_SCATTER");
  if (e && std::strcmp(e, "atomic") == 0) return VGPU_SCATTER_ATOMIC;
  return VGPU_SCATTER_COLORED;
}

int legacy_solve(int N, const double *B, double *X, double tol, int itmax,
                 int *iter, double *err, const double *BB, const int *ISS,
                 const double *DIAG, int nax, int nay, int naz, const int *bc) {
  if (N != nax * nay * naz || nax < 2 || nay < 2 || naz < 2 || itmax < 0) {
    set_error("kernel_dbcg_simple_: inconsistent sizes");
    return VGPU_ERR_ARG;
  }
  Legacy &L = legacy();
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  int rc = L.ensure(N, nax, nay, naz);
  if (rc) return rc;
  if (L.scatter < 0) L.scatter = legacy_scatter_default();
  hipStream_t s = L.s;
  VGPU_HIP(hipMemcpyAsync(L.x, X, N * sizeof(double), hipMemcpyHostToDevice, s));
  VGPU_HIP(hipMemcpyAsync(L.b, B, N * sizeof(double), hipMemcpyHostToDevice, s));
  VGPU_HIP(hipMemcpyAsync(L.bb, BB, 64 * L.E * sizeof(double), hipMemcpyHostToDevice, s));
  VGPU_HIP(hipMemcpyAsync(L.iss, ISS, 8 * L.E * sizeof(int), hipMemcpyHostToDevice, s));
  VGPU_HIP(hipMemcpyAsync(L.diag, DIAG, N * sizeof(double), hipMemcpyHostToDevice, s));
  VGPU_HIP(hipMemcpyAsync(L.bc, bc, N * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mask_from_int, dim3(256), dim3(256), 0, s, (int64_t)N,
                     L.bc, L.mask);
  VGPU_HIP(hipGetLastError());
  VGPU_HIP(hipStreamSynchronize(s));
  const auto t1 = clk::now();
  DualOp<double> dual = [&L](const double *p, double *z, const double *pp,
                             double *zz, double *pa, int *na, const int *done,
                             hipStream_t st) {
    int r = L.launch_dual(p, z, pp, zz, done, st);
    if (r) return r;
    *na = L.bicg.nb;
    hipLaunchKernelGGL((k_dot<double>), dim3(L.bicg.nb), dim3(VB), 0, st,
                       (int64_t)L.N,
                       reinterpret_cast<const SolveStatus *>(done), z, pp, pa);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  };
  SingleOp<double> single = [&L](const double *xx, double *y, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, st, nullptr, y, nullptr,
                       (int64_t)L.N);
    return L.launch_dual(xx, y, nullptr, nullptr, nullptr, st);
  };
  int status = 0;
  rc = L.bicg.solve((uint64_t)(100 + L.scatter), dual, single, L.mask, L.b, L.x,
                    L.diag, tol, itmax, true, s, iter, err, &status);
  if (rc) return rc;
  const auto t2 = clk::now();
  VGPU_HIP(hipMemcpyAsync(X, L.x, N * sizeof(double), hipMemcpyDeviceToHost, s));
  VGPU_HIP(hipStreamSynchronize(s));
  const auto t3 = clk::now();
  L.t_upload = std::chrono::duration<double>(t1 - t0).count();
  L.t_solve = std::chrono::duration<double>(t2 - t1).count();
  L.t_download = std::chrono::duration<double>(t3 - t2).count();
  return VGPU_OK;
}

}  // namespace
}  // namespace vgpu

extern "C" {

void kernel_dbcg_simple_(int *N, double *B, double *X, double *TOL, int *ITMAX,
                         int *ITER, double *ERR, double *BB, int *ISS,
                         double *DIAG, int *NAX, int *NAY, int *NAZ,
                         int *bc_dirichlet) {
  if (!N || !B || !X || !TOL || !ITMAX || !ITER || !ERR || !BB || !ISS ||
      !DIAG || !NAX || !NAY || !NAZ || !bc_dirichlet) {
    if (ERR) *ERR = std::numeric_limits<double>::infinity();
    return;
  }
  int iter = 0;
  double err = 0.0;
  const int rc = vgpu::legacy_solve(*N, B, X, *TOL, *ITMAX, &iter, &err, BB,
                                    ISS, DIAG, *NAX, *NAY, *NAZ, bc_dirichlet);
  if (rc) {
    std::fprintf(stderr, "kernel_dbcg_simple_: %s\n", vgpu_last_error());
    *ERR = std::numeric_limits<double>::infinity();
    return;
  }
  *ITER = iter;
  *ERR = err;
}

int vgpu_legacy_set_scatter(int scatter) {
  if (scatter != VGPU_SCATTER_ATOMIC && scatter != VGPU_SCATTER_COLORED) {
    vgpu::set_error("legacy path supports colored and atomic scatter only");
    return VGPU_ERR_ARG;
  }
  vgpu::legacy().scatter = scatter;
  vgpu::legacy().bicg.drop_graph();
  return VGPU_OK;
}

int vgpu_legacy_colors(int nax, int nay, int naz, int *color, int *elems, int *offsets,
                       int *counts) {
  if (nax < 2 || nay < 2 || naz < 2 || !color || !elems || !offsets || !counts ||
      (int64_t)(nax - 1) * (nay - 1) * (naz - 1) > INT32_MAX) {
    vgpu::set_error("vgpu_legacy_colors: bad arguments");
    return VGPU_ERR_ARG;
  }
  vgpu::colour_lists(nax, nay, naz, color, elems, offsets, counts);
  return VGPU_OK;
}

int vgpu_legacy_get_color_lists(int64_t capacity, int *elems, int *offsets, int *counts) {
  vgpu::Legacy &L = vgpu::legacy();
  if (!L.celems || L.E == 0) {
    vgpu::set_error("vgpu_legacy_get_color_lists: no legacy call has run");
    return VGPU_ERR_STATE;
  }
  if (!elems || !offsets || !counts || capacity < L.E) {
    vgpu::set_error("vgpu_legacy_get_color_lists: bad arguments");
    return VGPU_ERR_ARG;
  }
  VGPU_HIP(hipMemcpy(elems, L.celems, L.E * sizeof(int), hipMemcpyDeviceToHost));
  for (int c = 0; c < 9; ++c) offsets[c] = L.coff[c];
  for (int c = 0; c < 8; ++c) counts[c] = L.ccnt[c];
  return VGPU_OK;
}

int vgpu_legacy_last_times(double *upload_s, double *solve_s, double *download_s) {
  const vgpu::Legacy &L = vgpu::legacy();
  if (upload_s) *upload_s = L.t_upload;
  if (solve_s) *solve_s = L.t_solve;
  if (download_s) *download_s = L.t_download;
  return VGPU_OK;
}

int vgpu_legacy_release(void) {
  vgpu::legacy().release();
  return VGPU_OK;
}

}  // extern "C"
