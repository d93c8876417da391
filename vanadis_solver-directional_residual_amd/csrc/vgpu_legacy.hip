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
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "vgpu_bicg.hpp"
#include "vgpu_internal.hpp"

namespace vgpu {
namespace {

// Element products on stored BB (row-major per element: BB[64e + 8i + j]).
// y1 += A_e x1 (rows), y2 += A_e^T x2 (columns); masked in and out.  Eight
// lanes per element: lane j of an element's group handles local node j.  It
// loads the node's inputs, row j of A_e (four 16-byte loads: the group reads
// the element's 512 bytes in order) and column j (eight 8-byte loads: the
// group reads one 64-byte row per load, from cache), takes the other nodes'
// inputs from its neighbours in the group, and forms
//   y1_j = sum_i A[j][i] x1_i,   y2_j = sum_i A[i][j] x2_i   (i = 0..7 in order),
// the same sums as the reference's per-element loops (MATVEC_C_cu /
// MTTVEC_C_cu, cudatest.cu:184-227), so the coloured results do not depend on
// the launch shape.
template <bool ATOM, bool DUAL>
__global__ __launch_bounds__(256) void k_bb(int64_t E, const int *__restrict__ iss,
                                            const double *__restrict__ bb,
                                            const uint8_t *__restrict__ mask,
                                            const int *__restrict__ elems,
                                            int64_t count, const double *x1,
                                            double *y1, const double *x2,
                                            double *y2, const int *done) {
  if (done && *done) return;
  const int j = threadIdx.x & 7, gbase = threadIdx.x & 63 & ~7;
  const int64_t step = (int64_t)gridDim.x * (blockDim.x >> 3);
  // the eight lanes of a group share t, so they run the loop together
  for (int64_t t = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 3; t < count;
       t += step) {
    const int64_t e = elems ? elems[t] : t;
    const int nd = iss[j * E + e] - 1;
    const bool mk = mask[nd] != 0;
    const double xa = mk ? 0.0 : x1[nd];
    const double xb = (DUAL && !mk) ? x2[nd] : 0.0;
    const double *A = bb + 64 * e;
    double r[8], c[8];
    const double2 *row = reinterpret_cast<const double2 *>(A + 8 * j);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const double2 v = row[h];
      r[2 * h] = v.x;
      r[2 * h + 1] = v.y;
    }
    if (DUAL)
#pragma unroll
      for (int i = 0; i < 8; ++i) c[i] = A[8 * i + j];
    double ya = 0.0, yb = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      ya += r[i] * __shfl(xa, gbase + i, 64);
      if (DUAL) yb += c[i] * __shfl(xb, gbase + i, 64);
    }
    if (!mk) {
      if (ATOM) {
        atomicAdd(&y1[nd], ya);
        if (DUAL) atomicAdd(&y2[nd], yb);
      } else {
        y1[nd] += ya;
        if (DUAL) y2[nd] += yb;
      }
    }
  }
}

// workgroups of a k_bb launch over count elements (eight lanes each)
inline unsigned bb_grid(int64_t count) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (8 * count + 255) / 256));
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
  // (a failed allocation releases everything, so that the next call, the
  // host's next time step, allocates again instead of finding null buffers)
  int ensure(int N_, int nax_, int nay_, int naz_) {
    const int64_t E_ = (int64_t)(nax_ - 1) * (nay_ - 1) * (naz_ - 1);
    if (N_ == N && E_ == E && nax_ == nax && nay_ == nay && naz_ == naz && s)
      return VGPU_OK;
    const int rc = ensure_parts(N_, nax_, nay_, naz_, E_);
    if (rc) {
      const int keep_scatter = scatter;
      release();
      scatter = keep_scatter;
    }
    return rc;
  }
  int ensure_parts(int N_, int nax_, int nay_, int naz_, int64_t E_) {
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
      const unsigned g = bb_grid(E);
      if (dual)
        hipLaunchKernelGGL((k_bb<true, true>), dim3(g), dim3(256), 0, st, E, iss,
                           bb, mask, nullptr, E, p, z, pp, zz, done);
      else
        hipLaunchKernelGGL((k_bb<true, false>), dim3(g), dim3(256), 0, st, E,
                           iss, bb, mask, nullptr, E, p, z, pp, zz, done);
    } else {
      for (int c = 0; c < 8; ++c) {
        if (ccnt[c] == 0) continue;
        const unsigned g = bb_grid(ccnt[c]);
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
  // b = 0 ends BiCG in a breakdown (|r| / |b| = 0/0 or |r|/0; the reference's
  // loop turns x into NaN there): x = 0 is the exact solution, returned as
  // converged after 0 iterations with err 0
  if (status >= 3 && std::all_of(B, B + N, [](double v) { return v == 0.0; })) {
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, s, nullptr, L.x, nullptr, (int64_t)N);
    VGPU_HIP(hipGetLastError());
    *iter = 0;
    *err = 0.0;
  }
  const auto t2 = clk::now();
  VGPU_HIP(hipMemcpyAsync(X, L.x, N * sizeof(double), hipMemcpyDeviceToHost, s));
  VGPU_HIP(hipStreamSynchronize(s));
  const auto t3 = clk::now();
  L.t_upload = std::chrono::duration<double>(t1 - t0).count();
  L.t_solve = std::chrono::duration<double>(t2 - t1).count();
  L.t_download = std::chrono::duration<double>(t3 - t2).count();
  return VGPU_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free form of the drop-in (kernel_dbcg_simple_noalloc): the
// caller's device buffers and, for everything else the solve needs, one
// file-scope device array of fixed size.  Layout of the array (each part
// 256-byte aligned): r, rr, p, pp, z, zz [kNaMaxNodes doubles each], the
// Dirichlet flags [kNaMaxNodes bytes], the colour lists [kNaMaxNodes ints],
// the reduction partials (pr 2 x 256, pb 2 x 256, pbn 256, pa 256 doubles)
// and the solve status.
// ---------------------------------------------------------------------------
constexpr int64_t kNaMaxNodes = int64_t(1) << 18;
constexpr int64_t kNaAlign = 256;
constexpr int64_t na_round(int64_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr int64_t kNaVecBytes = 6 * kNaMaxNodes * (int64_t)sizeof(double);
constexpr int64_t kNaMaskOff = kNaVecBytes;
constexpr int64_t kNaElemOff = kNaMaskOff + na_round(kNaMaxNodes);
constexpr int64_t kNaPartOff = kNaElemOff + na_round(kNaMaxNodes * (int64_t)sizeof(int));
constexpr int64_t kNaPartCount = 2 * 256 + 2 * 256 + 256 + 256;
constexpr int64_t kNaStatOff = kNaPartOff + na_round(kNaPartCount * (int64_t)sizeof(double));
constexpr int64_t kNaScratchBytes = kNaStatOff + na_round(sizeof(SolveStatus));

// Element lists of the colour variant, computed on the device in closed form:
// element e = i + ex j + ex ey k has colour c = (i&1) + 2(j&1) + 4(k&1)
// (generate_hex8_colors, coloring.cu:50-74); the stable counting sort of
// build_color_element_lists (coloring.cu:82-106) puts it at
// offset_c + (k>>1) ni nj + (j>>1) ni + (i>>1), with ni (nj) the number of
// i in [0, ex) (j in [0, ey)) of its parity.  off[c] as colour_counts.
__global__ void k_colour_lists(int ex, int ey, int ez, int o0, int o1, int o2, int o3, int o4,
                               int o5, int o6, int o7, int *elems) {
  const int off[8] = {o0, o1, o2, o3, o4, o5, o6, o7};
  const int64_t E = (int64_t)ex * ey * ez;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % ex), j = (int)((e / ex) % ey), k = (int)(e / ((int64_t)ex * ey));
    const int pi = i & 1, pj = j & 1, pk = k & 1;
    const int ni = (ex - pi + 1) >> 1, nj = (ey - pj + 1) >> 1;
    const int64_t rank = (int64_t)(k >> 1) * ni * nj + (int64_t)(j >> 1) * ni + (i >> 1);
    elems[off[pi + 2 * pj + 4 * pk] + rank] = (int)e;
  }
}

// Colour counts and offsets of an ex x ey x ez element grid (the same as
// colour_lists computes by counting).
void colour_counts(int ex, int ey, int ez, int coff[9], int ccnt[8]) {
  coff[0] = 0;
  for (int c = 0; c < 8; ++c) {
    const int pi = c & 1, pj = (c >> 1) & 1, pk = (c >> 2) & 1;
    ccnt[c] = ((ex - pi + 1) >> 1) * ((ey - pj + 1) >> 1) * ((ez - pk + 1) >> 1);
    coff[c + 1] = coff[c] + ccnt[c];
  }
}

}  // namespace

// The scratch array of kernel_dbcg_simple_noalloc (one copy per device,
// resident for the life of the process).
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace {

std::mutex g_na_mu;
void *g_na_ptr[64] = {};  // device address of `scratch` per device

// The device address of `scratch` on the current device, looked up once per
// device (host code must never pass the array's name to a kernel: on the host
// it is a stand-in, not a device address).
int noalloc_scratch(void **out) {
  int dev = 0;
  VGPU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("kernel_dbcg_simple_noalloc: device index above 63");
    return VGPU_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    VGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = p;
  }
  *out = g_na_ptr[dev];
  return VGPU_OK;
}

int legacy_solve_noalloc(int N, const double *b, double *x, double tol, int itmax, int *iter,
                         double *err, const double *bb, const int *iss, const double *diag,
                         int nax, int nay, int naz, const int *bc, int scatter,
                         hipStream_t s) {
  if (N != nax * nay * naz || nax < 2 || nay < 2 || naz < 2 || itmax < 0 || !(tol >= 0) ||
      !b || !x || !iter || !err || !bb || !iss || !diag || !bc) {
    set_error("kernel_dbcg_simple_noalloc: bad arguments");
    return VGPU_ERR_ARG;
  }
  if ((int64_t)N > kNaMaxNodes) {
    set_error("kernel_dbcg_simple_noalloc: " + std::to_string(N) + " nodes exceed the " +
              std::to_string(kNaMaxNodes) + " of its fixed scratch array");
    return VGPU_ERR_ARG;
  }
  if (scatter != VGPU_SCATTER_COLORED && scatter != VGPU_SCATTER_ATOMIC) {
    set_error("kernel_dbcg_simple_noalloc: scatter must be colored or atomic");
    return VGPU_ERR_ARG;
  }
  void *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  unsigned char *sc = static_cast<unsigned char *>(base);
  const int64_t n = N, E = (int64_t)(nax - 1) * (nay - 1) * (naz - 1);
  double *v = reinterpret_cast<double *>(sc);
  double *r = v, *rr = v + kNaMaxNodes, *p = v + 2 * kNaMaxNodes, *pp = v + 3 * kNaMaxNodes,
         *z = v + 4 * kNaMaxNodes, *zz = v + 5 * kNaMaxNodes;
  uint8_t *mask = sc + kNaMaskOff;
  int *elems = reinterpret_cast<int *>(sc + kNaElemOff);
  double *part = reinterpret_cast<double *>(sc + kNaPartOff);
  double *pr = part, *pb = part + 512, *pbn = part + 1024, *pa = part + 1280;
  SolveStatus *st = reinterpret_cast<SolveStatus *>(sc + kNaStatOff);
  // the legacy path's launch shapes (bitwise the same iterates)
  const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + VB - 1) / VB));
  int coff[9], ccnt[8];
  colour_counts(nax - 1, nay - 1, naz - 1, coff, ccnt);
  hipLaunchKernelGGL(k_mask_from_int, dim3(256), dim3(256), 0, s, n, bc, mask);
  if (scatter == VGPU_SCATTER_COLORED)
    hipLaunchKernelGGL(k_colour_lists, dim3((unsigned)std::min<int64_t>(4096, (E + 255) / 256)),
                       dim3(256), 0, s, nax - 1, nay - 1, naz - 1, coff[0], coff[1], coff[2],
                       coff[3], coff[4], coff[5], coff[6], coff[7], elems);
  VGPU_HIP(hipGetLastError());
  // y1 (+)= A x1, y2 (+)= A^T x2 (y2 == nullptr: forward only) on stored BB
  auto dual = [&](const double *x1, double *y1, const double *x2, double *y2,
                  const int *done) {
    const bool two = x2 != nullptr;
    if (scatter == VGPU_SCATTER_ATOMIC) {
      const unsigned g = bb_grid(E);
      if (two)
        hipLaunchKernelGGL((k_bb<true, true>), dim3(g), dim3(256), 0, s, E, iss, bb, mask,
                           nullptr, E, x1, y1, x2, y2, done);
      else
        hipLaunchKernelGGL((k_bb<true, false>), dim3(g), dim3(256), 0, s, E, iss, bb, mask,
                           nullptr, E, x1, y1, x2, y2, done);
    } else {
      for (int c = 0; c < 8; ++c) {
        if (ccnt[c] == 0) continue;
        const unsigned g = bb_grid(ccnt[c]);
        if (two)
          hipLaunchKernelGGL((k_bb<false, true>), dim3(g), dim3(256), 0, s, E, iss, bb, mask,
                             elems + coff[c], (int64_t)ccnt[c], x1, y1, x2, y2, done);
        else
          hipLaunchKernelGGL((k_bb<false, false>), dim3(g), dim3(256), 0, s, E, iss, bb, mask,
                             elems + coff[c], (int64_t)ccnt[c], x1, y1, x2, y2, done);
      }
    }
  };
  // r = A x0, then the initial residuals, partials and (zero) status
  hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, s, nullptr, r, nullptr, n);
  dual(x, r, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL((k_init<double>), dim3(nb), dim3(VB), 0, s, n, mask, b, x, r, rr, z, zz,
                     diag, pr, pb, pbn, st);
  VGPU_HIP(hipGetLastError());
  // iterations in pairs (partials by parity), the status read after each
  // pair; kernels of iterations past the end return at once
  const int max_pairs = itmax / 2 + 2;
  SolveStatus hs{};
  for (int pair = 0; pair < max_pairs; ++pair) {
    for (int q = 1; q >= 0; --q) {
      hipLaunchKernelGGL((k_dir<double>), dim3(nb), dim3(VB), 0, s, n, nb, q, pr, pb, pbn, st,
                         tol, itmax, 1e300, p, pp, z, zz, 1);
      dual(p, z, pp, zz, &st->done);
      hipLaunchKernelGGL((k_dot<double>), dim3(nb), dim3(VB), 0, s, n, st, z, pp, pa);
      hipLaunchKernelGGL((k_update<double>), dim3(nb), dim3(VB), 0, s, n, nb, nb, q, pa, pr,
                         pb, st, x, p, r, rr, z, zz, diag);
    }
    VGPU_HIP(hipGetLastError());
    VGPU_HIP(hipMemcpyAsync(&hs, st, sizeof hs, hipMemcpyDeviceToHost, s));
    VGPU_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
  }
  *iter = hs.iter;
  *err = hs.err;
  if (hs.done >= 3 && hs.bnrm2 == 0.0) {  // b = 0: x = 0 exactly (as the allocating form)
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, s, nullptr, x, nullptr, n);
    VGPU_HIP(hipGetLastError());
    VGPU_HIP(hipStreamSynchronize(s));
    *iter = 0;
    *err = 0.0;
    return VGPU_OK;
  }
  if (hs.done == 1) return VGPU_OK;
  if (hs.done == 3) {
    set_error("kernel_dbcg_simple_noalloc: BiCG breakdown");
    return VGPU_BREAKDOWN;
  }
  set_error("kernel_dbcg_simple_noalloc: itmax reached");
  return VGPU_NOT_CONVERGED;
}

}  // namespace
}  // namespace vgpu

extern "C" {

int kernel_dbcg_simple_noalloc(int n, const double *b_dev, double *x_dev, double tol,
                               int itmax, int *iter, double *err, const double *bb_dev,
                               const int *iss_dev, const double *diag_dev, int nax, int nay,
                               int naz, const int *bc_dev, int scatter, void *stream) {
  return vgpu::legacy_solve_noalloc(n, b_dev, x_dev, tol, itmax, iter, err, bb_dev, iss_dev,
                                    diag_dev, nax, nay, naz, bc_dev, scatter,
                                    static_cast<hipStream_t>(stream));
}

int vgpu_noalloc_max_nodes(int64_t *max_nodes) {
  if (!max_nodes) return VGPU_ERR_ARG;
  *max_nodes = vgpu::kNaMaxNodes;
  return VGPU_OK;
}

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
