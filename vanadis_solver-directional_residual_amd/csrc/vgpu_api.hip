// C ABI of libvanadis_gpu.so: context lifecycle, coefficients, operator
// apply, RHS/diagonal assembly, FP64 and mixed-precision solves, timing.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vgpu_bbicg.hpp"
#include "vgpu_internal.hpp"

namespace vgpu {

static thread_local std::string g_err = "";
void set_error(const std::string &msg) { g_err = msg; }

namespace {

template <typename TA, typename TB>
__global__ void k_convert(int64_t n, const TA *a, TB *b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = TB(a[i]);
}

// out = (mask ? 0 : a*u + c*v)
__global__ void k_lincomb(int64_t n, const uint8_t *mask, double a,
                          const double *u, double c, const double *v,
                          double *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = mask[i] ? 0.0 : a * u[i] + (v ? c * v[i] : 0.0);
}

// out = a u + b v + c w on free nodes, 0 on Dirichlet nodes
__global__ void k_lincomb3(int64_t n, const uint8_t *mask, double a, const double *u, double b,
                           const double *v, double c, const double *w, double *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = mask[i] ? 0.0 : a * u[i] + b * v[i] + c * w[i];
}

// r = b - r; partials of r.r
__global__ __launch_bounds__(VB) void k_resid(int64_t n, const double *b,
                                              double *r, double *part) {
  __shared__ double red[VB / 64];
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * VB) {
    const double ri = b[i] - r[i];
    r[i] = ri;
    s += ri * ri;
  }
  block_partial(s, part, red);
}

// f = float(r * scale)
__global__ void k_scale_to_f32(int64_t n, const double *r, double scale,
                               float *f, float *x0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    f[i] = float(r[i] * scale);
    x0[i] = 0.0f;
  }
}

// x += s * double(d)
__global__ void k_refine(int64_t n, double *x, double s, const float *d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += s * (double)d[i];
}

// Gaussian source exp(-|c - x0|^2 / (2 sigma^2)) at element centroids times
// the element volume (the q_in of a unit box), and at the nodes.
__global__ void k_gauss_elem(Grid3 g, const double *xc, const double *yc,
                             const double *zc, double cx, double cy, double cz,
                             double inv2s2, double *q) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.ex), j = (int)((e / g.ex) % g.ey);
    const int k = (int)(e / ((int64_t)g.ex * g.ey));
    const double x = 0.5 * (xc[i] + xc[i + 1]), y = 0.5 * (yc[j] + yc[j + 1]),
                 z = 0.5 * (zc[k] + zc[k + 1]);
    const double vol = (xc[i + 1] - xc[i]) * (yc[j + 1] - yc[j]) * (zc[k + 1] - zc[k]);
    const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
    q[e] = exp(-d2 * inv2s2) * vol;
  }
}

__global__ void k_gauss_node(Grid3 g, const double *xc, const double *yc,
                             const double *zc, double cx, double cy, double cz,
                             double inv2s2, double *s) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.N;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.nax), j = (int)((n / g.nax) % g.nay);
    const int k = (int)(n / ((int64_t)g.nax * g.nay));
    const double d2 = (xc[i] - cx) * (xc[i] - cx) + (yc[j] - cy) * (yc[j] - cy) +
                      (zc[k] - cz) * (zc[k] - cz);
    s[n] = exp(-d2 * inv2s2);
  }
}

// x += omega (b - A x) / diag on free nodes (Ax holds A x)
__global__ void k_jacobi(int64_t n, const uint8_t *mask, const double *b, const double *ax,
                         const double *diag, double omega, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = mask[i] ? 0.0 : x[i] + omega * (b[i] - ax[i]) / diag[i];
}

__global__ void k_fill(int64_t n, double v, double *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v;
}

__global__ void k_copy(int64_t n, const double *a, double *b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Copy of n2 16-byte words, four per thread per step (the HBM-sized
// calibration copy of vgpu_time_kernel, what = 7).
__global__ __launch_bounds__(256) void k_copy_wide(int64_t n2, const double2 *__restrict__ a,
                                                   double2 *__restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * 1024;
  for (int64_t i0 = blockIdx.x * (int64_t)1024 + threadIdx.x; i0 < n2; i0 += stride) {
    double2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t i = i0 + 256 * k;
      v[k] = i < n2 ? a[i] : double2{0.0, 0.0};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t i = i0 + 256 * k;
      if (i < n2) b[i] = v[k];
    }
  }
}

inline dim3 grid_for(int64_t n, int bs = 256) {
  int64_t g = (n + bs - 1) / bs;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

}  // namespace
}  // namespace vgpu

using namespace vgpu;

#include "vgpu_ctx.hpp"

namespace {

int scheme_coeffs(int scheme, double dt, double *cH, double *cC) {
  switch (scheme) {
    case VGPU_SCHEME_THETA:  // A = 2H + (3/dt) C   (VANADIS.FOR:743-748)
      if (!(dt > 0)) break;
      *cH = 2.0;
      *cC = 3.0 / dt;
      return VGPU_OK;
    case VGPU_SCHEME_STEADY:
      *cH = 1.0;
      *cC = 0.0;
      return VGPU_OK;
    case VGPU_SCHEME_EULER:
      if (!(dt > 0)) break;
      *cH = 1.0;
      *cC = 1.0 / dt;
      return VGPU_OK;
    default:
      break;
  }
  set_error("invalid scheme/dt");
  return VGPU_ERR_ARG;
}

template <typename T>
void free_dev(T *&p) {
  if (p) hipFree(p);
  p = nullptr;
}

int ensure_diag(vgpu_ctx *c) {
  if (c->diag_valid) return VGPU_OK;
  OpData<double> op = c->op64(c->cH, c->cC);
  int rc = launch_diag(op, c->diag, c->stream);
  if (rc) return rc;
  hipLaunchKernelGGL((k_convert<double, float>), grid_for(c->g.N), dim3(256), 0,
                     c->stream, c->g.N, c->diag, c->fdiag);
  VGPU_HIP(hipGetLastError());
  c->diag_valid = true;
  return VGPU_OK;
}

double host_sum(const double *dev, int n, hipStream_t s, int *rc) {
  std::vector<double> h(n);
  if (hipMemcpyAsync(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("reduction readback failed");
    *rc = VGPU_ERR_HIP;
    return 0;
  }
  double t = 0;
  for (int i = 0; i < n; ++i) t += h[i];
  return t;
}

// dot-partial capacity of the brick-mode solver (k_dot partials)
int max_owner_tiles(const Grid3 &) { return 256; }

int nparts_for(int64_t n) {
  return (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + VB - 1) / VB));
}

// The context's FP32 right-hand side and iterate, allocated on first use;
// one that failed is tried again by the next call (no half-set pair is used).
int ensure_f32_vectors(vgpu_ctx *c) {
  const int64_t N = c->g.N;
  if (!c->fb && hipMalloc(&c->fb, N * sizeof(float)) != hipSuccess) c->fb = nullptr;
  if (c->fb && !c->fx && hipMalloc(&c->fx, N * sizeof(float)) != hipSuccess) c->fx = nullptr;
  if (c->fb && c->fx) return VGPU_OK;
  (void)hipGetLastError();
  set_error("FP32 vector allocation failed");
  return VGPU_ERR_NOMEM;
}

// A zero right-hand side: x = 0 solves A x = b exactly from any start, while
// the solvers' relative residual |r| / |b| is 0/0 or |r|/0 there and ends
// them in a breakdown.  Called for a solve that ended in one: with |b| = 0
// (b and x of n entries on the context's device and stream) x is zeroed and
// *zero set, so that the caller reports the solve converged after 0
// iterations with err 0.
int zero_rhs_solution(vgpu_ctx *c, const double *b, double *x, int64_t n, bool *zero) {
  const int np = nparts_for(n);
  hipLaunchKernelGGL((k_dot<double>), dim3(np), dim3(VB), 0, c->stream, n, nullptr, b, b,
                     c->red);
  VGPU_HIP(hipGetLastError());
  int rc = VGPU_OK;
  const double bb = host_sum(c->red, np, c->stream, &rc);
  if (rc) return rc;
  *zero = bb == 0.0;
  if (*zero) {
    hipLaunchKernelGGL(k_fill, grid_for(n), dim3(256), 0, c->stream, n, 0.0, x);
    VGPU_HIP(hipGetLastError());
    VGPU_HIP(hipStreamSynchronize(c->stream));
  }
  return VGPU_OK;
}

// FP64 dual-operator callback for the BiCG on the context's operator
DualOp<double> make_dual64(vgpu_ctx *c, int scatter) {
  return [c, scatter](const double *p, double *z, const double *pp, double *zz,
                      double *pa, int *na, const int *done, hipStream_t s) {
    OpData<double> op = c->op64(c->cH, c->cC);
    op.done = done;
    int rc = launch_apply<double, double>(op, scatter, c->zchunk, p, z, pp, zz,
                                          nullptr, s);
    if (rc) return rc;
    const int nb = nparts_for(c->g.N);
    *na = nb;
    SolveStatus *st = reinterpret_cast<SolveStatus *>(
        const_cast<int *>(done));  // done is the first member
    hipLaunchKernelGGL((k_dot<double>), dim3(nb), dim3(VB), 0, s, c->g.N, st, z,
                       pp, pa);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  };
}

DualOp<float> make_dual32(vgpu_ctx *c, int scatter) {
  return [c, scatter](const float *p, float *z, const float *pp, float *zz,
                      double *pa, int *na, const int *done, hipStream_t s) {
    OpData<float> op = c->op32();
    op.done = done;
    int rc = launch_apply<float, float>(op, scatter, c->zchunk, p, z, pp, zz,
                                        nullptr, s);
    if (rc) return rc;
    const int nb = nparts_for(c->g.N);
    *na = nb;
    SolveStatus *st = reinterpret_cast<SolveStatus *>(const_cast<int *>(done));
    hipLaunchKernelGGL((k_dot<float>), dim3(nb), dim3(VB), 0, s, c->g.N, st, z,
                       pp, pa);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  };
}

template <typename TV>
OpData<TV> op_for(const vgpu_ctx *c);
template <>
OpData<double> op_for<double>(const vgpu_ctx *c) {
  return c->op64(c->cH, c->cC);
}
template <>
OpData<float> op_for<float>(const vgpu_ctx *c) {
  return c->op32();
}

// One system on the batched engine (owner scatter).  status pointer of the
// dual callback doubles as OpData::done (SolveStatus for owner launches).
template <typename TV>
int solve_owner(vgpu_ctx *c, BBicg<TV> &eng, const TV *b, TV *x, const TV *diag,
                double tol, int itmax, int *iters, double *err, int *status,
                double divtol = 1e300) {
  int rc = eng.alloc(make_segments({c->g.N}, {owner_units(c->g, c->tile)},
                                   {owner_groups(c->g, c->tile, c->zchunk)}));
  if (rc) return rc;
  const Segment sg = eng.hseg[0];
  BDirOp<TV> dual = [c, sg](const TV *z, const TV *zz, const DirData<TV> &dd, TV *w,
                            TV *ww, double *pa, const ActiveSet &, hipStream_t s) {
    Member<TV> m0{};
    m0.op = op_for<TV>(c);
    m0.seg = sg;
    return launch_owner_dir<TV, TV>(m0, nullptr, ActiveSet{}, false, c->tile, c->zchunk,
                                    m0.op.ra != 0.0, dd, z, zz, w, ww, pa, s);
  };
  BSingleOp<TV> single = [c](const TV *xx, TV *y, const ActiveSet &, hipStream_t s) {
    OpData<TV> op = op_for<TV>(c);
    return launch_apply<TV, TV>(op, VGPU_SCATTER_OWNER, c->zchunk, xx, y, nullptr,
                                nullptr, nullptr, s);
  };
  const uint64_t key = (c->version << 8) | (uint64_t)(c->zchunk % 64) |
                       (sizeof(TV) == 4 ? (1ull << 7) : 0);
  return eng.solve(key, dual, single, c->mask, b, x, diag, tol, itmax, c->stream,
                   iters, err, status, divtol);
}

// Z chunk of the BiCGSTAB launches: the context's, or with none set (0, the
// balanced launch, whose per-plane dot partials every workgroup of both
// launches would re-sum) the smallest chunk that keeps every (chunk, tile)
// group of a launch resident at once (two 256-thread workgroups per CU).
int stab_zchunk(const vgpu_ctx *c) {
  if (c->zchunk > 0) return c->zchunk;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess || cus <= 0)
    cus = 256;
  const int64_t tiles = owner_groups(c->g, c->tile, c->g.naz);  // one chunk per tile
  const int64_t chunks = std::max<int64_t>(1, 2 * (int64_t)cus / std::max<int64_t>(1, tiles));
  return (int)((c->g.naz + chunks - 1) / chunks);
}

// One system by BiCGSTAB on the owner launches (whatever the context's
// scatter mode: the fused launches exist for the owner kernel only).
template <typename TV>
int solve_stab(vgpu_ctx *c, Stab<TV> &eng, const TV *b, TV *x, const TV *diag, double tol,
               int itmax, int *iters, double *err, int *status, double divtol = 1e300) {
  if (!owner_fits(c->g)) {
    set_error("BiCGSTAB: system too large for the owner launches (8 N and 40 E must "
              "stay below 2^31 bytes)");
    return VGPU_ERR_ARG;
  }
  const int zc = stab_zchunk(c);
  const int64_t na = owner_groups(c->g, c->tile, zc);
  int rc = eng.alloc(c->g.N, (int)na);
  if (rc) return rc;
  StabSingleOp<TV> single = [c](const TV *xx, TV *y, hipStream_t s) {
    OpData<TV> op = op_for<TV>(c);
    return launch_apply<TV, TV>(op, VGPU_SCATTER_OWNER, c->zchunk, xx, y, nullptr, nullptr,
                                nullptr, s);
  };
  StabOwnerOp<TV> owner = [c, zc](int mode, const StabData<TV> &sd, const TV *r, TV *y,
                                  hipStream_t s) {
    OpData<TV> op = op_for<TV>(c);
    return launch_owner_stab<TV, TV>(op, mode, zc, sd, r, y, s);
  };
  const uint64_t key = (c->version << 8) | (uint64_t)(zc % 64) |
                       (sizeof(TV) == 4 ? (1ull << 7) : 0);
  return eng.solve(key, single, owner, c->mask, b, x, diag, tol, itmax, divtol, c->stream,
                   iters, err, status);
}

int solve_fp64_dev(vgpu_ctx *c, const double *b, double *x, double rtol,
                   int itmax, int *iters, double *err, int *status) {
  if (c->method == VGPU_METHOD_BICGSTAB)
    return solve_stab<double>(c, c->stab64, b, x, c->diag, rtol, itmax, iters, err, status);
  const int scatter = c->scatter;
  if (scatter == VGPU_SCATTER_OWNER)
    return solve_owner<double>(c, c->bb64, b, x, c->diag, rtol, itmax, iters, err,
                               status);
  if (c->bicg64.n != c->g.N) {
    int rc = c->bicg64.alloc(c->g.N, max_owner_tiles(c->g) + 256);
    if (rc) return rc;
  }
  DualOp<double> dual = make_dual64(c, scatter);
  SingleOp<double> single = [c, scatter](const double *xx, double *y,
                                         hipStream_t s) {
    OpData<double> op = c->op64(c->cH, c->cC);
    return launch_apply<double, double>(op, scatter, c->zchunk, xx, y, nullptr,
                                        nullptr, nullptr, s);
  };
  const uint64_t key = (c->version << 8) | (uint64_t)(scatter * 16 + c->zchunk % 16);
  return c->bicg64.solve(key, dual, single, c->mask, b, x, c->diag, rtol, itmax,
                         scatter != VGPU_SCATTER_OWNER, c->stream, iters, err,
                         status);
}

// Pure FP32 BiCG (FP32 vectors and operator, FP64 scalars): b, x FP64 in/out
int solve_fp32_dev(vgpu_ctx *c, const double *b, double *x, double rtol,
                   int itmax, int *iters, double *err, int *status) {
  const int64_t N = c->g.N;
  const int scatter = c->scatter;
  if (c->bicg32.n != N) {
    int rc = c->bicg32.alloc(N, max_owner_tiles(c->g) + 256);
    if (rc) return rc;
  }
  if (int r = ensure_f32_vectors(c)) return r;
  hipLaunchKernelGGL((k_convert<double, float>), grid_for(N), dim3(256), 0,
                     c->stream, N, b, c->fb);
  hipLaunchKernelGGL((k_convert<double, float>), grid_for(N), dim3(256), 0,
                     c->stream, N, x, c->fx);
  VGPU_HIP(hipGetLastError());
  int rc = VGPU_OK;
  if (c->method == VGPU_METHOD_BICGSTAB) {
    rc = solve_stab<float>(c, c->stab32, c->fb, c->fx, c->fdiag, rtol, itmax, iters, err,
                           status);
  } else if (scatter == VGPU_SCATTER_OWNER) {
    rc = solve_owner<float>(c, c->bb32, c->fb, c->fx, c->fdiag, rtol, itmax, iters,
                            err, status);
  } else {
    DualOp<float> dual = make_dual32(c, scatter);
    SingleOp<float> single = [c, scatter](const float *xx, float *y,
                                          hipStream_t s) {
      OpData<float> op = c->op32();
      return launch_apply<float, float>(op, scatter, c->zchunk, xx, y, nullptr,
                                        nullptr, nullptr, s);
    };
    const uint64_t key = (c->version << 8) | (uint64_t)(scatter * 16 + c->zchunk % 16) | (1ull << 7);
    rc = c->bicg32.solve(key, dual, single, c->mask, c->fb, c->fx, c->fdiag, rtol,
                         itmax, true, c->stream, iters, err, status);
  }
  if (rc) return rc;
  hipLaunchKernelGGL((k_convert<float, double>), grid_for(N), dim3(256), 0,
                     c->stream, N, c->fx, x);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

// Inner iterations allowed per refinement step of the mixed solve.
constexpr int kInnerCap = 1000;

// FP64 pass of a mixed solve whose FP32 inner solves stopped reducing the
// true residual: BiCGSTAB on the owner launches (the method of the inner
// solves; plain BiCG can stagnate on the systems where the FP32 inner solves
// give up), the context's method for systems too large for them.
int mixed_fp64_pass(vgpu_ctx *c, const double *b, double *x, double rtol, int itmax,
                    int *iters, double *err, int *status) {
  if (owner_fits(c->g))
    return solve_stab<double>(c, c->stab64, b, x, c->diag, rtol, itmax, iters, err, status);
  return solve_fp64_dev(c, b, x, rtol, itmax, iters, err, status);
}

// Mixed precision: FP64 true residual + refinement, FP32 inner BiCGSTAB on
// the owner launches (FP32 BiCG loses its recurrences across the large
// residual peaks of advection-dominated systems; BiCGSTAB's residual stays
// within a few orders of |b|), FP32 BiCG for systems too large for them
int solve_mixed_dev(vgpu_ctx *c, const double *b, double *x, double rtol,
                    int itmax, double inner_rtol, int *iters, double *err,
                    int *status) {
  const int64_t N = c->g.N;
  const int scatter = c->scatter;
  if (c->bicg32.n != N) {
    int rc = c->bicg32.alloc(N, max_owner_tiles(c->g) + 256);
    if (rc) return rc;
  }
  if (int r = ensure_f32_vectors(c)) return r;
  const int np = nparts_for(N);
  int rc = VGPU_OK;
  // ||b||
  hipLaunchKernelGGL((k_dot<double>), dim3(np), dim3(VB), 0, c->stream, N,
                     nullptr, b, b, c->red);
  VGPU_HIP(hipGetLastError());
  const double bn = std::sqrt(host_sum(c->red, np, c->stream, &rc));
  if (rc) return rc;
  if (bn == 0.0) {  // b = 0: x = 0 exactly, from any start
    hipLaunchKernelGGL(k_fill, grid_for(N), dim3(256), 0, c->stream, N, 0.0, x);
    VGPU_HIP(hipGetLastError());
    VGPU_HIP(hipStreamSynchronize(c->stream));
    *iters = 0;
    *err = 0.0;
    *status = 1;
    return VGPU_OK;
  }
  DualOp<float> dual = make_dual32(c, scatter);
  SingleOp<float> single = [c, scatter](const float *xx, float *y,
                                        hipStream_t s) {
    OpData<float> op = c->op32();
    return launch_apply<float, float>(op, scatter, c->zchunk, xx, y, nullptr,
                                      nullptr, nullptr, s);
  };
  const uint64_t key = (c->version << 8) | (uint64_t)(scatter * 16 + c->zchunk % 16) | (1ull << 7);
  int total = 0, st_inner = 0;
  double rel = 0, prev_rel = 0;
  *status = 2;
  // x masked like the FP64 path
  hipLaunchKernelGGL(k_lincomb, grid_for(N), dim3(256), 0, c->stream, N, c->mask,
                     1.0, x, 0.0, nullptr, x);
  // true relative residual |b - A x| / |b| of the current iterate (r in vt)
  auto true_rel = [&](double *rn_out) -> int {
    OpData<double> op = c->op64(c->cH, c->cC);
    int r = launch_apply<double, double>(op, VGPU_SCATTER_OWNER, c->zchunk, x, c->vt,
                                         nullptr, nullptr, nullptr, c->stream);
    if (r) return r;
    hipLaunchKernelGGL(k_resid, dim3(np), dim3(VB), 0, c->stream, N, b, c->vt,
                       c->red);
    VGPU_HIP(hipGetLastError());
    *rn_out = std::sqrt(host_sum(c->red, np, c->stream, &r));
    return r;
  };
  bool fp64 = false;
  for (int outer = 0; outer < 64; ++outer) {
    double rn = 0;
    rc = true_rel(&rn);
    if (rc) return rc;
    rel = rn / bn;
    if (rel <= rtol) {
      *status = 1;
      break;
    }
    if (!(rel == rel)) {
      *status = 3;
      break;
    }
    if (total >= itmax) {
      *status = 2;
      break;
    }
    // the FP32 inner solve stopped reducing the true residual: continue with
    // FP64 passes from the current iterate (the fallback of the mixed mode);
    // each pass restarts from the true residual checked above
    if (!fp64 && outer > 0 && (rel > 0.5 * prev_rel || st_inner >= 3)) fp64 = true;
    prev_rel = rel;
    if (fp64) {
      int it64 = 0, st64 = 0;
      double e64 = 0;
      rc = mixed_fp64_pass(c, b, x, rtol, itmax - total, &it64, &e64, &st64);
      if (rc) return rc;
      total += it64;
      if (st64 >= 3) {  // breakdown or divergence of the FP64 pass
        *status = st64;
        double rn2 = 0;
        rc = true_rel(&rn2);
        if (rc) return rc;
        rel = rn2 / bn;
        break;
      }
      continue;
    }
    hipLaunchKernelGGL(k_scale_to_f32, grid_for(N), dim3(256), 0, c->stream, N,
                       c->vt, 1.0 / rn, c->fb, c->fx);
    VGPU_HIP(hipGetLastError());
    int it = 0;
    double e = 0;
    // inner FP32 solve on the scaled residual: capped, and stopped if its
    // recursive residual grows 1e4-fold (BiCGSTAB) or 100-fold (BiCG); a
    // stopped inner solve is not applied
    if (owner_fits(c->g))
      rc = solve_stab<float>(c, c->stab32, c->fb, c->fx, c->fdiag, inner_rtol, kInnerCap,
                             &it, &e, &st_inner, 1e4);
    else
      rc = c->bicg32.solve(key, dual, single, c->mask, c->fb, c->fx, c->fdiag,
                           inner_rtol, kInnerCap, true, c->stream, &it, &e,
                           &st_inner, 1e2);
    if (rc) return rc;
    total += it;
    if (st_inner >= 3) continue;  // next pass falls back to FP64
    hipLaunchKernelGGL(k_refine, grid_for(N), dim3(256), 0, c->stream, N, x, rn,
                       c->fx);
    VGPU_HIP(hipGetLastError());
  }
  if (*status == 2) {  // pass limit or itmax: report the final iterate
    double rn = 0;
    rc = true_rel(&rn);
    if (rc) return rc;
    rel = rn / bn;
  }
  VGPU_HIP(hipStreamSynchronize(c->stream));
  *iters = total;
  *err = rel;
  return VGPU_OK;
}

}  // namespace

extern "C" {

const char *vgpu_last_error(void) { return g_err.c_str(); }
const char *vgpu_version(void) { return "vanadis-mi355x 0.1 (gfx950)"; }

int vgpu_device_count(int *count) {
  if (!count) return VGPU_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return VGPU_OK;
}

int vgpu_create(int nax, int nay, int naz, const double *xx, const double *yy,
                const double *zz, const int *iss, const int *bc,
                vgpu_ctx **out) {
  if (!out || !xx || !yy || !zz || !bc || nax < 2 || nay < 2 || naz < 2) {
    set_error("vgpu_create: bad arguments");
    return VGPU_ERR_ARG;
  }
  *out = nullptr;
  const int64_t N = (int64_t)nax * nay * naz;
  const int64_t E = (int64_t)(nax - 1) * (nay - 1) * (naz - 1);
  if (E > INT32_MAX / 8) {
    set_error("vgpu_create: mesh too large for 32-bit connectivity");
    return VGPU_ERR_ARG;
  }
  // axis coordinates and tensor-grid check (siatka, VANADIS.FOR:226-244)
  std::vector<double> ax(nax), ay(nay), az(naz);
  for (int i = 0; i < nax; ++i) ax[i] = xx[i];
  for (int j = 0; j < nay; ++j) ay[j] = yy[(int64_t)nax * j];
  for (int k = 0; k < naz; ++k) az[k] = zz[(int64_t)nax * nay * k];
  for (int64_t n = 0; n < N; ++n) {
    const int i = (int)(n % nax), j = (int)((n / nax) % nay),
              k = (int)(n / ((int64_t)nax * nay));
    if (xx[n] != ax[i] || yy[n] != ay[j] || zz[n] != az[k]) {
      set_error("vgpu_create: node coordinates are not a tensor grid");
      return VGPU_ERR_MESH;
    }
  }
  for (int i = 0; i + 1 < nax; ++i)
    if (!(ax[i + 1] > ax[i])) {
      set_error("vgpu_create: x coordinates not increasing");
      return VGPU_ERR_MESH;
    }
  for (int j = 0; j + 1 < nay; ++j)
    if (!(ay[j + 1] > ay[j])) {
      set_error("vgpu_create: y coordinates not increasing");
      return VGPU_ERR_MESH;
    }
  for (int k = 0; k + 1 < naz; ++k)
    if (!(az[k + 1] > az[k])) {
      set_error("vgpu_create: z coordinates not increasing");
      return VGPU_ERR_MESH;
    }
  // connectivity must be the structured w_el numbering (VANADIS.FOR:246-276)
  if (iss) {
    const int ex = nax - 1, ey = nay - 1;
    for (int64_t e = 0; e < E; ++e) {
      const int i = (int)(e % ex), j = (int)((e / ex) % ey),
                k = (int)(e / ((int64_t)ex * ey));
      const int64_t n0 = 1 + i + (int64_t)nax * j + (int64_t)nax * nay * k;
      const int64_t off[8] = {0,
                              1,
                              1 + nax,
                              nax,
                              (int64_t)nax * nay,
                              1 + (int64_t)nax * nay,
                              1 + (int64_t)nax * nay + nax,
                              (int64_t)nax * nay + nax};
      for (int c = 0; c < 8; ++c)
        if (iss[c * E + e] != n0 + off[c]) {
          set_error("vgpu_create: ISS is not the structured HEX8 connectivity");
          return VGPU_ERR_MESH;
        }
    }
  }
  std::vector<uint8_t> m(N);
  for (int64_t n = 0; n < N; ++n) {
    if (bc[n] != 0 && bc[n] != 1) {
      set_error("vgpu_create: bc_dirichlet must be 0/1");
      return VGPU_ERR_ARG;
    }
    m[n] = (uint8_t)bc[n];
  }
  vgpu_ctx *c = new vgpu_ctx();
  auto fail = [&](int rc) {
    vgpu_destroy(c);
    return rc;
  };
#define CK(call)                                                         \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));      \
      return fail(VGPU_ERR_HIP);                                         \
    }                                                                    \
  } while (0)
  CK(hipGetDevice(&c->device));
  CK(hipStreamCreate(&c->stream));
  c->g = Grid3{nax, nay, naz, nax - 1, nay - 1, naz - 1, N, E};
  // the owner kernel addresses each vector and the element records with
  // 32-bit offsets; larger systems use the coloured scatter
  if (!owner_fits(c->g)) c->scatter = VGPU_SCATTER_COLORED;
  CK(hipMalloc(&c->xc, nax * sizeof(double)));
  CK(hipMalloc(&c->yc, nay * sizeof(double)));
  CK(hipMalloc(&c->zc, naz * sizeof(double)));
  CK(hipMalloc(&c->mask, N));
  CK(hipMemcpy(c->xc, ax.data(), nax * sizeof(double), hipMemcpyHostToDevice));
  CK(hipMemcpy(c->yc, ay.data(), nay * sizeof(double), hipMemcpyHostToDevice));
  CK(hipMemcpy(c->zc, az.data(), naz * sizeof(double), hipMemcpyHostToDevice));
  CK(hipMemcpy(c->mask, m.data(), N, hipMemcpyHostToDevice));
  {
    // element sizes and reciprocals per axis (FP64 and FP32 copies)
    const std::vector<double> *axes[3] = {&ax, &ay, &az};
    for (int d = 0; d < 3; ++d) {
      const std::vector<double> &a = *axes[d];
      const size_t ne = a.size() - 1;
      std::vector<double> h(ne), ih(ne);
      std::vector<float> fh(ne), fih(ne);
      for (size_t i = 0; i < ne; ++i) {
        h[i] = a[i + 1] - a[i];
        ih[i] = 1.0 / h[i];
        fh[i] = (float)h[i];
        fih[i] = (float)ih[i];
      }
      CK(hipMalloc(&c->ah[d], ne * sizeof(double)));
      CK(hipMalloc(&c->aih[d], ne * sizeof(double)));
      CK(hipMalloc(&c->fah[d], ne * sizeof(float)));
      CK(hipMalloc(&c->faih[d], ne * sizeof(float)));
      CK(hipMemcpy(c->ah[d], h.data(), ne * sizeof(double), hipMemcpyHostToDevice));
      CK(hipMemcpy(c->aih[d], ih.data(), ne * sizeof(double), hipMemcpyHostToDevice));
      CK(hipMemcpy(c->fah[d], fh.data(), ne * sizeof(float), hipMemcpyHostToDevice));
      CK(hipMemcpy(c->faih[d], fih.data(), ne * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  CK(hipMalloc(&c->rec, 5 * E * sizeof(double)));
  CK(hipMalloc(&c->frec, 5 * E * sizeof(float)));
  CK(hipMalloc(&c->diag, N * sizeof(double)));
  CK(hipMalloc(&c->fdiag, N * sizeof(float)));
  for (double **p : {&c->vb, &c->vx_, &c->vt, &c->vt2})
    CK(hipMalloc(p, N * sizeof(double)));
  CK(hipMalloc(&c->red, 4096 * sizeof(double)));
#undef CK
  *out = c;
  return VGPU_OK;
}

int vgpu_destroy(vgpu_ctx *c) {
  if (!c) return VGPU_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  c->bicg64.release();
  c->bicg32.release();
  c->bb64.release();
  c->bb32.release();
  c->stab64.release();
  c->stab32.release();
  free_dev(c->xc);
  free_dev(c->yc);
  free_dev(c->zc);
  free_dev(c->mask);
  for (int d = 0; d < 3; ++d) {
    free_dev(c->ah[d]);
    free_dev(c->aih[d]);
    free_dev(c->fah[d]);
    free_dev(c->faih[d]);
  }
  for (double **p : {&c->rec, &c->diag, &c->vb, &c->vx_, &c->vt, &c->vt2, &c->red,
                     &c->tpart, &c->tdir})
    free_dev(*p);
  for (float **p : {&c->frec, &c->fdiag,
                    &c->fb, &c->fx})
    free_dev(*p);
  for (void *p : c->user_allocs) hipFree(p);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return VGPU_OK;
}

int vgpu_get_sizes(const vgpu_ctx *c, int64_t *n_nodes, int64_t *n_elems) {
  if (!c || !n_nodes || !n_elems) return VGPU_ERR_ARG;
  *n_nodes = c->g.N;
  *n_elems = c->g.E;
  return VGPU_OK;
}

int vgpu_set_coeffs(vgpu_ctx *c, const double *lam, const double *vel,
                    double P, double alpha, int dr_variant, int scheme,
                    double dt) {
  if (!c || !lam || !vel) {
    set_error("vgpu_set_coeffs: null argument");
    return VGPU_ERR_ARG;
  }
  if (dr_variant < VGPU_DR_V1 || dr_variant > VGPU_DR_V3) {
    set_error("vgpu_set_coeffs: dr_variant must be 1, 2 or 3");
    return VGPU_ERR_ARG;
  }
  double cH, cC;
  int rc = scheme_coeffs(scheme, dt, &cH, &cC);
  if (rc) return rc;
  const int64_t E = c->g.E;
  VGPU_HIP(hipSetDevice(c->device));
  {
    std::vector<double> r(5 * (size_t)E);
    for (int64_t e = 0; e < E; ++e) {
      r[5 * e + 0] = vel[e];
      r[5 * e + 1] = vel[E + e];
      r[5 * e + 2] = vel[2 * E + e];
      r[5 * e + 3] = lam[e];
      r[5 * e + 4] = 0.0;
    }
    VGPU_HIP(hipMemcpy(c->rec, r.data(), 5 * E * sizeof(double), hipMemcpyHostToDevice));
  }
  rc = launch_tau(c->g, c->xc, c->yc, c->zc, c->rec, dr_variant, c->stream);
  if (rc) return rc;
  hipLaunchKernelGGL((k_convert<double, float>), grid_for(5 * E), dim3(256), 0,
                     c->stream, 5 * E, c->rec, c->frec);
  VGPU_HIP(hipGetLastError());
  VGPU_HIP(hipStreamSynchronize(c->stream));
  c->P = P;
  c->alpha = alpha;
  c->variant = dr_variant;
  c->scheme = scheme;
  c->dt = dt;
  c->cH = cH;
  c->cC = cC;
  c->have_coeffs = true;
  c->diag_valid = false;
  c->version++;
  return VGPU_OK;
}

int vgpu_set_scheme(vgpu_ctx *c, int scheme, double dt) {
  if (!c) return VGPU_ERR_ARG;
  double cH, cC;
  int rc = scheme_coeffs(scheme, dt, &cH, &cC);
  if (rc) return rc;
  c->scheme = scheme;
  c->dt = dt;
  c->cH = cH;
  c->cC = cC;
  c->diag_valid = false;
  c->version++;
  return VGPU_OK;
}

int vgpu_set_options(vgpu_ctx *c, int scatter, int zchunk, int tile) {
  if (!c || scatter < 0 || scatter > 2 || zchunk < -1 || zchunk > 256 || tile > 2) {
    set_error("vgpu_set_options: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (scatter == VGPU_SCATTER_OWNER && !owner_fits(c->g)) {
    set_error("vgpu_set_options: system too large for the owner scatter (8 N and 40 E "
              "must stay below 2^31 bytes); use the coloured or atomic scatter");
    return VGPU_ERR_ARG;
  }
  c->scatter = scatter;
  if (zchunk >= 0) c->zchunk = zchunk;
  if (tile >= 0) c->tile = tile;
  c->version++;
  return VGPU_OK;
}

int vgpu_set_method(vgpu_ctx *c, int method) {
  if (!c || (method != VGPU_METHOD_BICG && method != VGPU_METHOD_BICGSTAB)) {
    set_error("vgpu_set_method: method must be VGPU_METHOD_BICG or VGPU_METHOD_BICGSTAB");
    return VGPU_ERR_ARG;
  }
  if (method == VGPU_METHOD_BICGSTAB && !owner_fits(c->g)) {
    set_error("vgpu_set_method: BiCGSTAB runs on the owner launches; system too large");
    return VGPU_ERR_ARG;
  }
  c->method = method;
  return VGPU_OK;
}

int vgpu_apply_dev(vgpu_ctx *c, int scatter, int transpose, const double *x,
                   double *y) {
  if (!c || !x || !y || scatter < 0 || scatter > 2) return VGPU_ERR_ARG;
  if (!c->have_coeffs) {
    set_error("vgpu_apply: coefficients not set");
    return VGPU_ERR_STATE;
  }
  OpData<double> op = c->op64(c->cH, c->cC);
  if (!transpose)
    return launch_apply<double, double>(op, scatter, c->zchunk, x, y, nullptr,
                                        nullptr, nullptr, c->stream);
  // transpose alone: dual apply with the forward result in scratch
  return launch_apply<double, double>(op, scatter, c->zchunk, x, c->vt2, x, y,
                                      nullptr, c->stream);
}

int vgpu_apply(vgpu_ctx *c, int scatter, int precision, int transpose,
               const double *x, double *y) {
  if (!c || !x || !y || scatter < 0 || scatter > 2) {
    set_error("vgpu_apply: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (!c->have_coeffs) {
    set_error("vgpu_apply: coefficients not set");
    return VGPU_ERR_STATE;
  }
  const int64_t N = c->g.N;
  VGPU_HIP(hipSetDevice(c->device));
  VGPU_HIP(hipMemcpyAsync(c->vx_, x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int rc;
  if (precision == VGPU_FP64) {
    rc = vgpu_apply_dev(c, scatter, transpose, c->vx_, c->vb);
  } else if (precision == VGPU_FP32) {
    OpData<float> op = c->op32();
    if (!transpose)
      rc = launch_apply<float, double>(op, scatter, c->zchunk, c->vx_, c->vb,
                                       nullptr, nullptr, nullptr, c->stream);
    else
      rc = launch_apply<float, double>(op, scatter, c->zchunk, c->vx_, c->vt2,
                                       c->vx_, c->vb, nullptr, c->stream);
  } else {
    set_error("vgpu_apply: precision must be FP64 or FP32");
    return VGPU_ERR_ARG;
  }
  if (rc) return rc;
  VGPU_HIP(hipMemcpyAsync(y, c->vb, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  VGPU_HIP(hipStreamSynchronize(c->stream));
  return VGPU_OK;
}

int vgpu_diag(vgpu_ctx *c, double *diag) {
  if (!c || !diag) return VGPU_ERR_ARG;
  if (!c->have_coeffs) {
    set_error("vgpu_diag: coefficients not set");
    return VGPU_ERR_STATE;
  }
  int rc = ensure_diag(c);
  if (rc) return rc;
  VGPU_HIP(hipMemcpyAsync(diag, c->diag, c->g.N * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
  VGPU_HIP(hipStreamSynchronize(c->stream));
  return VGPU_OK;
}

int vgpu_assemble_rhs_dev(vgpu_ctx *c, const double *q_dev, const double *t_old_dev,
                          double *b_dev) {
  if (!c || !q_dev || !b_dev) return VGPU_ERR_ARG;
  if (!c->have_coeffs) {
    set_error("vgpu_assemble_rhs: coefficients not set");
    return VGPU_ERR_STATE;
  }
  VGPU_HIP(hipSetDevice(c->device));
  const int64_t N = c->g.N;
  OpData<double> op = c->op64(c->cH, c->cC);
  int rc = launch_source(op, q_dev, c->vt, c->stream);  // F
  if (rc) return rc;
  double fa = 1.0, ya = 0.0;
  const bool use_t = t_old_dev != nullptr && c->scheme != VGPU_SCHEME_STEADY;
  if (use_t) {
    OpData<double> o2;
    if (c->scheme == VGPU_SCHEME_THETA) {
      // b = 3F - (H - (3/dt) C) t_old   (VANADIS.FOR:725-741)
      o2 = c->op64(1.0, -3.0 / c->dt);
      fa = 3.0;
      ya = -1.0;
    } else {
      // b = F + (C/dt) t_old
      o2 = c->op64(0.0, 1.0 / c->dt);
      ya = 1.0;
    }
    rc = launch_apply<double, double>(o2, VGPU_SCATTER_OWNER, c->zchunk, t_old_dev,
                                      c->vt2, nullptr, nullptr, nullptr, c->stream);
    if (rc) return rc;
  } else if (c->scheme == VGPU_SCHEME_THETA) {
    fa = 3.0;
  }
  hipLaunchKernelGGL(k_lincomb, grid_for(N), dim3(256), 0, c->stream, N, c->mask, fa,
                     c->vt, ya, use_t ? c->vt2 : nullptr, b_dev);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

int vgpu_assemble_rhs(vgpu_ctx *c, const double *q_in, const double *t_old,
                      double *b) {
  if (!c || !q_in || !b) return VGPU_ERR_ARG;
  if (!c->have_coeffs) {
    set_error("vgpu_assemble_rhs: coefficients not set");
    return VGPU_ERR_STATE;
  }
  VGPU_HIP(hipSetDevice(c->device));
  const int64_t N = c->g.N, E = c->g.E;
  double *q_dev = nullptr;
  VGPU_HIP(hipMalloc(&q_dev, E * sizeof(double)));
  int rc = VGPU_OK;
  do {
    const bool use_t = t_old != nullptr && c->scheme != VGPU_SCHEME_STEADY;
    if (hipMemcpyAsync(q_dev, q_in, E * sizeof(double), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        (use_t && hipMemcpyAsync(c->vx_, t_old, N * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
      set_error("vgpu_assemble_rhs: upload failed");
      rc = VGPU_ERR_HIP;
      break;
    }
    rc = vgpu_assemble_rhs_dev(c, q_dev, use_t ? c->vx_ : nullptr, c->vb);
    if (rc) break;
    if (hipMemcpyAsync(b, c->vb, N * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      set_error("vgpu_assemble_rhs: HIP failure");
      rc = VGPU_ERR_HIP;
    }
  } while (0);
  hipFree(q_dev);
  return rc;
}

int vgpu_copy_dev(vgpu_ctx *c, void *dst, const void *src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(c->device));
  VGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return VGPU_OK;
}

int vgpu_axpby_dev(vgpu_ctx *c, double a, const double *x, double b, const double *y,
                   double *out) {
  if (!c || !x || !out || (b != 0.0 && !y)) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(c->device));
  const int64_t N = c->g.N;
  hipLaunchKernelGGL(k_lincomb, grid_for(N), dim3(256), 0, c->stream, N, c->mask, a, x,
                     b, b != 0.0 ? y : nullptr, out);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

int vgpu_axpbypcz_dev(vgpu_ctx *c, double a, const double *x, double b, const double *y,
                      double cz, const double *z, double *out) {
  if (!c || !x || !y || !z || !out) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(c->device));
  const int64_t N = c->g.N;
  hipLaunchKernelGGL(k_lincomb3, grid_for(N), dim3(256), 0, c->stream, N, c->mask, a, x, b, y,
                     cz, z, out);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

int vgpu_residual_norm_dev(vgpu_ctx *c, const double *b, const double *x, double *rel) {
  if (!c || !b || !x || !rel) return VGPU_ERR_ARG;
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  VGPU_HIP(hipSetDevice(c->device));
  const int64_t N = c->g.N;
  OpData<double> op = c->op64(c->cH, c->cC);
  int rc = launch_apply<double, double>(op, VGPU_SCATTER_OWNER, c->zchunk, x, c->vt,
                                        nullptr, nullptr, nullptr, c->stream);
  if (rc) return rc;
  const int np = nparts_for(N);
  hipLaunchKernelGGL(k_resid, dim3(np), dim3(VB), 0, c->stream, N, b, c->vt, c->red);
  hipLaunchKernelGGL((k_dot<double>), dim3(np), dim3(VB), 0, c->stream, N, nullptr, b, b,
                     c->red + np);
  VGPU_HIP(hipGetLastError());
  const double rn = host_sum(c->red, np, c->stream, &rc);
  if (rc) return rc;
  const double bn = host_sum(c->red + np, np, c->stream, &rc);
  if (rc) return rc;
  *rel = bn > 0 ? std::sqrt(rn / bn) : std::sqrt(rn);
  return VGPU_OK;
}

int vgpu_jacobi_dev(vgpu_ctx *c, const double *b, double *x, int sweeps, double omega) {
  if (!c || !b || !x || sweeps < 0) return VGPU_ERR_ARG;
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  VGPU_HIP(hipSetDevice(c->device));
  int rc = ensure_diag(c);
  if (rc) return rc;
  const int64_t N = c->g.N;
  OpData<double> op = c->op64(c->cH, c->cC);
  if (owner_fits(c->g)) {
    // the update fused into the owner launch, alternating between x and vt
    // (other tiles read a sweep's input at their halo while it writes)
    for (int k = 0; k < sweeps; ++k) {
      const double *in = (k & 1) ? c->vt : x;
      double *out = (k & 1) ? x : c->vt;
      rc = launch_owner_jacobi<double, double>(op, c->zchunk, omega, in, b, c->diag, out,
                                               c->stream);
      if (rc) return rc;
    }
    if (sweeps & 1) {
      hipLaunchKernelGGL(k_copy, grid_for(N), dim3(256), 0, c->stream, N, c->vt, x);
      VGPU_HIP(hipGetLastError());
    }
    return VGPU_OK;
  }
  for (int k = 0; k < sweeps; ++k) {
    rc = launch_apply<double, double>(op, c->scatter, c->zchunk, x, c->vt, nullptr, nullptr,
                                      nullptr, c->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_jacobi, grid_for(N), dim3(256), 0, c->stream, N, c->mask, b, c->vt,
                       c->diag, omega, x);
    VGPU_HIP(hipGetLastError());
  }
  return VGPU_OK;
}

int vgpu_gaussian_dev(vgpu_ctx *c, double cx, double cy, double cz, double sigma,
                      double *q_elem_dev, double *src_node_dev) {
  if (!c || !(sigma > 0)) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(c->device));
  const double inv2s2 = 0.5 / (sigma * sigma);
  if (q_elem_dev)
    hipLaunchKernelGGL(k_gauss_elem, grid_for(c->g.E), dim3(256), 0, c->stream, c->g,
                       c->xc, c->yc, c->zc, cx, cy, cz, inv2s2, q_elem_dev);
  if (src_node_dev)
    hipLaunchKernelGGL(k_gauss_node, grid_for(c->g.N), dim3(256), 0, c->stream, c->g,
                       c->xc, c->yc, c->zc, cx, cy, cz, inv2s2, src_node_dev);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

int vgpu_solve_dev(vgpu_ctx *c, int precision, const double *b, double *x,
                   double rtol, int itmax, double inner_rtol, int *iters,
                   double *err, double *seconds) {
  if (!c || !b || !x || !iters || !err || itmax < 0 || !(rtol > 0)) {
    set_error("vgpu_solve: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (!c->have_coeffs) {
    set_error("vgpu_solve: coefficients not set");
    return VGPU_ERR_STATE;
  }
  VGPU_HIP(hipSetDevice(c->device));
  int rc = ensure_diag(c);
  if (rc) return rc;
  // the timer brackets the solve only when asked for (a sync before it);
  // without `seconds` the solve queues behind the caller's work
  if (seconds) VGPU_HIP(hipStreamSynchronize(c->stream));
  const auto t0 = std::chrono::steady_clock::now();
  int status = 0;
  if (precision == VGPU_FP64)
    rc = solve_fp64_dev(c, b, x, rtol, itmax, iters, err, &status);
  else if (precision == VGPU_FP32)
    rc = solve_fp32_dev(c, b, x, rtol, itmax, iters, err, &status);
  else if (precision == VGPU_MIXED)
    rc = solve_mixed_dev(c, b, x, rtol, itmax,
                         inner_rtol > 0 ? inner_rtol : 1e-4, iters, err, &status);
  else {
    set_error("vgpu_solve: precision must be FP64, FP32 or MIXED");
    return VGPU_ERR_ARG;
  }
  if (rc) return rc;
  VGPU_HIP(hipStreamSynchronize(c->stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  if (status >= 3) {  // a breakdown on b = 0 is the exact solution x = 0
    bool zero = false;
    rc = zero_rhs_solution(c, b, x, c->g.N, &zero);
    if (rc) return rc;
    if (zero) {
      *iters = 0;
      *err = 0.0;
      status = 1;
    }
  }
  if (status == 1) return VGPU_OK;
  if (status == 3) {
    set_error("BiCG breakdown (zero or non-finite scalar)");
    return VGPU_BREAKDOWN;
  }
  set_error("BiCG: itmax reached before rtol");
  return VGPU_NOT_CONVERGED;
}

int vgpu_solve(vgpu_ctx *c, int precision, const double *b, double *x,
               double rtol, int itmax, double inner_rtol, int *iters,
               double *err) {
  if (!c || !b || !x) return VGPU_ERR_ARG;
  const int64_t N = c->g.N;
  VGPU_HIP(hipSetDevice(c->device));
  VGPU_HIP(hipMemcpyAsync(c->vb, b, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  VGPU_HIP(hipMemcpyAsync(c->vx_, x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const int rc = vgpu_solve_dev(c, precision, c->vb, c->vx_, rtol, itmax,
                                inner_rtol, iters, err, nullptr);
  if (rc == VGPU_OK || rc == VGPU_NOT_CONVERGED || rc == VGPU_BREAKDOWN) {
    VGPU_HIP(hipMemcpyAsync(x, c->vx_, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VGPU_HIP(hipStreamSynchronize(c->stream));
  }
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched solves: several contexts (own mesh, coefficients, right-hand side)
// on one device, solved by one engine: one persistent owner launch per
// iteration over all systems, one hipGraph for the batch.
// ---------------------------------------------------------------------------
struct vgpu_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<vgpu_ctx *> m;
  std::vector<uint64_t> versions;
  std::vector<Segment> seg;
  int tile = 0, zchunk = 0;
  bool robin = false;
  int64_t ntot = 0, utot = 0;
  Member<double> *md64 = nullptr;
  Member<float> *md32 = nullptr;
  uint8_t *mask = nullptr;
  double *diag = nullptr, *b = nullptr, *x = nullptr;
  float *fdiag = nullptr, *fb = nullptr, *fx = nullptr;
  BBicg<double> e64;
  BBicg<float> e32;
  BStab<float> es32;  // inner solves of the mixed-precision batch
  double *vr = nullptr;  // mixed precision: true residuals (batch layout)
  double *bred = nullptr;  // mixed precision: per-member reduction partials
  int bred_n = 0;
  uint64_t gen = 0;
  int fallback = VGPU_METHOD_BICG;  // VGPU_METHOD_BICGSTAB: finish unconverged members
  int fallback_itmax = 0;  // iterations allowed to each member's fallback solve

  ~vgpu_batch() {
    e64.release();
    e32.release();
    es32.release();
    void *ptrs[] = {md64, md32, mask, diag, b, x, fdiag, fb, fx, vr, bred};
    for (void *q : ptrs)
      if (q) hipFree(q);
    if (stream) hipStreamDestroy(stream);
  }

  // rebuild descriptors and concatenated mask / diagonals after any member
  // changed its coefficients, scheme or options
  int refresh() {
    bool stale = !md64;
    for (size_t i = 0; i < m.size(); ++i)
      stale = stale || versions[i] != m[i]->version || !m[i]->diag_valid;
    if (!stale) return VGPU_OK;
    const int nm = (int)m.size();
    std::vector<Member<double>> h64(nm);
    std::vector<Member<float>> h32(nm);
    robin = false;
    for (int i = 0; i < nm; ++i) {
      vgpu_ctx *c = m[i];
      if (!c->have_coeffs) {
        set_error("vgpu_batch: member " + std::to_string(i) + " has no coefficients");
        return VGPU_ERR_STATE;
      }
      int rc = ensure_diag(c);
      if (rc) return rc;
      VGPU_HIP(hipStreamSynchronize(c->stream));
      h64[i].op = c->op64(c->cH, c->cC);
      h32[i].op = c->op32();
      h64[i].op.tile = h32[i].op.tile = tile;
      h64[i].seg = h32[i].seg = seg[i];
      robin = robin || h64[i].op.ra != 0.0;
      const int64_t o = seg[i].off, n = seg[i].n;
      VGPU_HIP(hipMemcpyAsync(mask + o, c->mask, n, hipMemcpyDeviceToDevice, stream));
      VGPU_HIP(hipMemcpyAsync(diag + o, c->diag, n * sizeof(double),
                              hipMemcpyDeviceToDevice, stream));
      VGPU_HIP(hipMemcpyAsync(fdiag + o, c->fdiag, n * sizeof(float),
                              hipMemcpyDeviceToDevice, stream));
      versions[i] = c->version;
    }
    VGPU_HIP(hipMemcpyAsync(md64, h64.data(), nm * sizeof(Member<double>),
                            hipMemcpyHostToDevice, stream));
    VGPU_HIP(hipMemcpyAsync(md32, h32.data(), nm * sizeof(Member<float>),
                            hipMemcpyHostToDevice, stream));
    VGPU_HIP(hipStreamSynchronize(stream));
    ++gen;
    return VGPU_OK;
  }

  template <typename TV>
  int solve_tv(BBicg<TV> &eng, const Member<TV> *md, const TV *bb, TV *xx,
               const TV *dg, double rtol, int itmax, int *iters, double *errs,
               int *status) {
    int rc = eng.alloc(seg);
    if (rc) return rc;
    BDirOp<TV> dual = [this, md](const TV *z, const TV *zz, const DirData<TV> &dd, TV *w,
                                 TV *ww, double *pa, const ActiveSet &as, hipStream_t s) {
      return launch_owner_dir<TV, TV>(Member<TV>{}, md, as, true, tile, zchunk, robin, dd,
                                      z, zz, w, ww, pa, s);
    };
    BSingleOp<TV> single = [this, md](const TV *xi, TV *y, const ActiveSet &as,
                                      hipStream_t s) {
      return launch_owner_batch<TV, TV>(md, as, tile, zchunk, robin, nullptr, xi, y,
                                        nullptr, nullptr, nullptr, s);
    };
    const uint64_t key = (gen << 8) | (uint64_t)(zchunk % 64) |
                         (sizeof(TV) == 4 ? (1ull << 7) : 0);
    return eng.solve(key, dual, single, mask, bb, xx, dg, rtol, itmax, stream, iters,
                     errs, status);
  }
};

namespace {

// Mixed precision over a batch: for every member the steps of
// solve_mixed_dev (FP64 true residual r = b - A x and refinement
// x += |r| dx, inner FP32 BiCGSTAB on r/|r| to inner_rtol, capped at
// kInnerCap iterations and stopped on a 1e4-fold residual growth; a member
// whose refinement fails to halve its true residual continues with FP64
// BiCGSTAB passes, mixed_fp64_pass), with the FP64 operator applied to all
// members in one owner launch and the inner solves of all members that need
// one run together by the batched BiCGSTAB (two owner launches per
// iteration over all members, captured in one hipGraph).  The vector
// kernels run per member with the grids of the single-context solve, so with
// the same owner tile shape and a common z chunk > 0 a member's result is
// bitwise that of vgpu_solve(VGPU_MIXED) on its context.
int batch_solve_mixed(vgpu_batch *bt, const double *b, double *x, double rtol, int itmax,
                      double inner_rtol, int *iters, double *errs, int *status) {
  const int nm = (int)bt->m.size();
  hipStream_t s = bt->stream;
  int rc = bt->es32.alloc(bt->seg);
  if (rc) return rc;
  if (!bt->vr) VGPU_HIP(hipMalloc(&bt->vr, bt->ntot * sizeof(double)));
  std::vector<int> np(nm), pofs(nm + 1, 0);
  for (int m = 0; m < nm; ++m) {
    np[m] = nparts_for(bt->seg[m].n);
    pofs[m + 1] = pofs[m] + np[m];
  }
  if (bt->bred_n < pofs[nm]) {
    if (bt->bred) hipFree(bt->bred);
    bt->bred = nullptr;
    bt->bred_n = 0;
    VGPU_HIP(hipMalloc(&bt->bred, pofs[nm] * sizeof(double)));
    bt->bred_n = pofs[nm];
  }
  std::vector<double> hpart(pofs[nm]);
  // per-member sums of the partials, each in order (as host_sum)
  auto member_sums = [&](std::vector<double> &out) -> int {
    VGPU_HIP(hipMemcpyAsync(hpart.data(), bt->bred, pofs[nm] * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    VGPU_HIP(hipStreamSynchronize(s));
    for (int m = 0; m < nm; ++m) {
      double t = 0;
      for (int i = 0; i < np[m]; ++i) t += hpart[pofs[m] + i];
      out[m] = t;
    }
    return VGPU_OK;
  };
  const ActiveSet &all = bt->es32.all.view;
  std::vector<double> bn(nm), rn(nm), rel(nm, 0.0), prev_rel(nm, 0.0), sums(nm);
  std::vector<int> total(nm, 0), st_inner(nm, 0), fp64(nm, 0), running(nm, 1), run(nm);
  for (int m = 0; m < nm; ++m) {
    const Segment &sg = bt->seg[m];
    status[m] = 2;
    hipLaunchKernelGGL((k_dot<double>), dim3(np[m]), dim3(VB), 0, s, sg.n, nullptr,
                       b + sg.off, b + sg.off, bt->bred + pofs[m]);
    hipLaunchKernelGGL(k_lincomb, grid_for(sg.n), dim3(256), 0, s, sg.n, bt->mask + sg.off,
                       1.0, x + sg.off, 0.0, nullptr, x + sg.off);
  }
  VGPU_HIP(hipGetLastError());
  rc = member_sums(sums);
  if (rc) return rc;
  for (int m = 0; m < nm; ++m) {
    bn[m] = std::sqrt(sums[m]);
    if (bn[m] == 0.0) {  // b = 0: x = 0 exactly, from any start
      const Segment &sg = bt->seg[m];
      hipLaunchKernelGGL(k_fill, grid_for(sg.n), dim3(256), 0, s, (int64_t)sg.n, 0.0,
                         x + sg.off);
      status[m] = 1;
      running[m] = 0;
    }
  }
  VGPU_HIP(hipGetLastError());
  // true residuals r = b - A x of every member (in vr) and their norms
  auto true_rel = [&]() -> int {
    int r = launch_owner_batch<double, double>(bt->md64, all, bt->tile, bt->zchunk, bt->robin,
                                               nullptr, x, bt->vr, nullptr, nullptr, nullptr,
                                               s);
    if (r) return r;
    for (int m = 0; m < nm; ++m) {
      const Segment &sg = bt->seg[m];
      hipLaunchKernelGGL(k_resid, dim3(np[m]), dim3(VB), 0, s, sg.n, b + sg.off,
                         bt->vr + sg.off, bt->bred + pofs[m]);
    }
    VGPU_HIP(hipGetLastError());
    r = member_sums(sums);
    if (r) return r;
    for (int m = 0; m < nm; ++m) {
      rn[m] = std::sqrt(sums[m]);
      rel[m] = bn[m] > 0.0 ? rn[m] / bn[m] : 0.0;  // b = 0: x = 0, set above
    }
    return VGPU_OK;
  };
  BStabSingleOp<float> single = [bt, &all](const float *xi, float *y, hipStream_t st) {
    return launch_owner_batch<float, float>(bt->md32, all, bt->tile, bt->zchunk, bt->robin,
                                            nullptr, xi, y, nullptr, nullptr, nullptr, st);
  };
  BStabOwnerOp<float> owner = [bt, &all](int mode, const StabData<float> &sd, const float *r,
                                         float *y, hipStream_t st) {
    return launch_owner_stab_batch<float, float>(bt->md32, all, bt->tile, bt->zchunk,
                                                 bt->robin, mode, sd, r, y, st);
  };
  const uint64_t key = (bt->gen << 8) | (uint64_t)(bt->zchunk % 64) | (1ull << 7);
  std::vector<int> it_in(nm), stin(nm);
  std::vector<double> e_in(nm);
  for (int outer = 0; outer < 64; ++outer) {
    bool any = false;
    for (int m = 0; m < nm; ++m) any = any || running[m];
    if (!any) break;
    rc = true_rel();
    if (rc) return rc;
    for (int m = 0; m < nm; ++m) {
      if (!running[m]) continue;
      if (rel[m] <= rtol) {
        status[m] = 1;
        running[m] = 0;
      } else if (!(rel[m] == rel[m])) {
        status[m] = 3;
        running[m] = 0;
      } else if (total[m] >= itmax) {
        status[m] = 2;
        running[m] = 0;
      } else {
        if (!fp64[m] && outer > 0 && (rel[m] > 0.5 * prev_rel[m] || st_inner[m] >= 3))
          fp64[m] = 1;
        prev_rel[m] = rel[m];
      }
    }
    // members in FP64 mode: one FP64 pass from the current iterate (their b
    // and x are the batch vectors' ranges)
    bool need_rel = false;
    for (int m = 0; m < nm; ++m) {
      if (!running[m] || !fp64[m]) continue;
      vgpu_ctx *c = bt->m[m];
      const Segment &sg = bt->seg[m];
      VGPU_HIP(hipStreamSynchronize(s));
      VGPU_HIP(hipSetDevice(c->device));
      int it64 = 0, st64 = 0;
      double e64 = 0;
      rc = mixed_fp64_pass(c, b + sg.off, x + sg.off, rtol, itmax - total[m], &it64, &e64,
                           &st64);
      if (rc) return rc;
      VGPU_HIP(hipStreamSynchronize(c->stream));
      VGPU_HIP(hipSetDevice(bt->device));
      total[m] += it64;
      if (st64 >= 3) {  // breakdown or divergence of the FP64 pass
        status[m] = st64;
        running[m] = 0;
        need_rel = true;
      }
    }
    if (need_rel) {  // the final true residual of those members
      std::vector<double> keep = rel;
      rc = true_rel();
      if (rc) return rc;
      for (int m = 0; m < nm; ++m)
        if (running[m] || status[m] < 3) rel[m] = keep[m];
    }
    // inner FP32 solves of the other running members, together
    bool any_inner = false;
    for (int m = 0; m < nm; ++m) {
      run[m] = running[m] && !fp64[m];
      any_inner = any_inner || run[m];
    }
    if (!any_inner) continue;
    for (int m = 0; m < nm; ++m) {
      if (!run[m]) continue;
      const Segment &sg = bt->seg[m];
      hipLaunchKernelGGL(k_scale_to_f32, grid_for(sg.n), dim3(256), 0, s, sg.n,
                         bt->vr + sg.off, 1.0 / rn[m], bt->fb + sg.off, bt->fx + sg.off);
    }
    VGPU_HIP(hipGetLastError());
    rc = bt->es32.solve(key, single, owner, bt->mask, bt->fb, bt->fx, bt->fdiag, inner_rtol,
                        kInnerCap, 1e4, run.data(), s, it_in.data(), e_in.data(),
                        stin.data());
    if (rc) return rc;
    for (int m = 0; m < nm; ++m) {
      if (!run[m]) continue;
      total[m] += it_in[m];
      st_inner[m] = stin[m];
      if (stin[m] >= 3) continue;  // next pass falls back to FP64
      const Segment &sg = bt->seg[m];
      hipLaunchKernelGGL(k_refine, grid_for(sg.n), dim3(256), 0, s, sg.n, x + sg.off, rn[m],
                         bt->fx + sg.off);
    }
    VGPU_HIP(hipGetLastError());
  }
  // pass limit or itmax: report the final iterate's true residual
  bool any2 = false;
  for (int m = 0; m < nm; ++m) any2 = any2 || status[m] == 2;
  if (any2) {
    std::vector<double> keep = rel;
    rc = true_rel();
    if (rc) return rc;
    for (int m = 0; m < nm; ++m)
      if (status[m] != 2) rel[m] = keep[m];
  }
  VGPU_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < nm; ++m) {
    iters[m] = total[m];
    errs[m] = rel[m];
  }
  return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_batch_create(int nm, vgpu_ctx *const *ctxs, vgpu_batch **out) {
  if (!out || nm < 1 || !ctxs) {
    set_error("vgpu_batch_create: bad arguments");
    return VGPU_ERR_ARG;
  }
  *out = nullptr;
  for (int i = 0; i < nm; ++i) {
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) {
      set_error("vgpu_batch_create: null context or contexts on different devices");
      return VGPU_ERR_ARG;
    }
    if (!owner_fits(ctxs[i]->g)) {
      set_error("vgpu_batch_create: system too large for the owner scatter");
      return VGPU_ERR_ARG;
    }
  }
  auto *bt = new (std::nothrow) vgpu_batch();
  if (!bt) return VGPU_ERR_NOMEM;
  bt->device = ctxs[0]->device;
  bt->m.assign(ctxs, ctxs + nm);
  bt->versions.assign(nm, 0);
  bt->tile = ctxs[0]->tile;
  bt->zchunk = ctxs[0]->zchunk;
  std::vector<int64_t> n(nm), nu(nm), ng(nm);
  for (int i = 0; i < nm; ++i) {
    n[i] = ctxs[i]->g.N;
    nu[i] = owner_units(ctxs[i]->g, bt->tile);
    ng[i] = owner_groups(ctxs[i]->g, bt->tile, bt->zchunk);
  }
  bt->seg = make_segments(n, nu, ng);
  bt->ntot = bt->seg.back().off + seg_pad(bt->seg.back().n);
  bt->utot = bt->seg.back().u0 + bt->seg.back().nu;
  auto fail = [&](const char *what, hipError_t e) {
    set_error(std::string("vgpu_batch_create: ") + what + ": " + hipGetErrorString(e));
    delete bt;
    return e == hipErrorOutOfMemory ? VGPU_ERR_NOMEM : VGPU_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipSetDevice(bt->device)) != hipSuccess) return fail("device", e);
  if ((e = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail("stream", e);
  const size_t N = (size_t)bt->ntot;
  if ((e = hipMalloc(&bt->md64, nm * sizeof(Member<double>))) != hipSuccess ||
      (e = hipMalloc(&bt->md32, nm * sizeof(Member<float>))) != hipSuccess ||
      (e = hipMalloc(&bt->mask, N)) != hipSuccess ||
      (e = hipMalloc(&bt->diag, N * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&bt->b, N * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&bt->x, N * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&bt->fdiag, N * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&bt->fb, N * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&bt->fx, N * sizeof(float))) != hipSuccess)
    return fail("allocation", e);
  // padding between systems: masked, unit diagonal, zero vectors
  if ((e = hipMemset(bt->mask, 1, N)) != hipSuccess ||
      (e = hipMemset(bt->b, 0, N * sizeof(double))) != hipSuccess ||
      (e = hipMemset(bt->x, 0, N * sizeof(double))) != hipSuccess)
    return fail("memset", e);
  hipLaunchKernelGGL(k_fill, grid_for(N), dim3(256), 0, bt->stream, (int64_t)N, 1.0,
                     bt->diag);
  hipLaunchKernelGGL((k_convert<double, float>), grid_for(N), dim3(256), 0, bt->stream,
                     (int64_t)N, bt->diag, bt->fdiag);
  if ((e = hipStreamSynchronize(bt->stream)) != hipSuccess) return fail("init", e);
  *out = bt;
  return VGPU_OK;
}

int vgpu_batch_destroy(vgpu_batch *bt) {
  if (!bt) return VGPU_ERR_ARG;
  hipSetDevice(bt->device);
  delete bt;
  return VGPU_OK;
}

int vgpu_batch_layout(vgpu_batch *bt, int64_t *offsets, int64_t *total) {
  if (!bt || !offsets || !total) return VGPU_ERR_ARG;
  for (size_t i = 0; i < bt->seg.size(); ++i) offsets[i] = bt->seg[i].off;
  *total = bt->ntot;
  return VGPU_OK;
}

int vgpu_batch_solve_dev(vgpu_batch *bt, int precision, const double *b, double *x,
                         double rtol, int itmax, int *iters, double *errs,
                         int *status, double *seconds) {
  if (!bt || !b || !x || !iters || !errs || !status || itmax < 0 || !(rtol > 0)) {
    set_error("vgpu_batch_solve: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (precision != VGPU_FP64 && precision != VGPU_FP32 && precision != VGPU_MIXED) {
    set_error("vgpu_batch_solve: precision must be FP64, FP32 or MIXED");
    return VGPU_ERR_ARG;
  }
  VGPU_HIP(hipSetDevice(bt->device));
  int rc = bt->refresh();
  if (rc) return rc;
  const int64_t N = bt->ntot;
  VGPU_HIP(hipStreamSynchronize(bt->stream));
  const auto t0 = std::chrono::steady_clock::now();
  if (precision == VGPU_FP64) {
    rc = bt->solve_tv<double>(bt->e64, bt->md64, b, x, bt->diag, rtol, itmax, iters,
                              errs, status);
  } else if (precision == VGPU_MIXED) {
    rc = batch_solve_mixed(bt, b, x, rtol, itmax, 1e-4, iters, errs, status);
  } else {
    hipLaunchKernelGGL((k_convert<double, float>), grid_for(N), dim3(256), 0,
                       bt->stream, N, b, bt->fb);
    hipLaunchKernelGGL((k_convert<double, float>), grid_for(N), dim3(256), 0,
                       bt->stream, N, x, bt->fx);
    VGPU_HIP(hipGetLastError());
    rc = bt->solve_tv<float>(bt->e32, bt->md32, bt->fb, bt->fx, bt->fdiag, rtol,
                             itmax, iters, errs, status);
    if (rc) return rc;
    hipLaunchKernelGGL((k_convert<float, double>), grid_for(N), dim3(256), 0,
                       bt->stream, N, bt->fx, x);
    VGPU_HIP(hipGetLastError());
  }
  if (rc) return rc;
  VGPU_HIP(hipStreamSynchronize(bt->stream));
  // members that ended in a breakdown on b = 0: x = 0 is their exact
  // solution, converged after 0 iterations
  std::vector<char> zero(bt->m.size(), 0);
  for (size_t i = 0; i < bt->m.size(); ++i) {
    if (status[i] < 3) continue;
    vgpu_ctx *c = bt->m[i];
    const Segment &sg = bt->seg[i];
    bool z = false;
    VGPU_HIP(hipSetDevice(c->device));
    rc = zero_rhs_solution(c, b + sg.off, x + sg.off, sg.n, &z);
    if (rc) return rc;
    if (z) {
      zero[i] = 1;
      iters[i] = 0;
      errs[i] = 0.0;
      status[i] = 1;
    }
  }
  VGPU_HIP(hipSetDevice(bt->device));
  // FP64 members the batched BiCG left at itmax or in breakdown, or stopped
  // on its recursive residual while the true one, |b - A x| / |b|, is above
  // rtol (the recurrence drifts from the truth on members whose residual
  // peaks high): BiCGSTAB on the member's own context from its current
  // iterate, or from zero if that iterate is worse than zero (true residual
  // above 1 or not finite), to the same rtol
  if (precision == VGPU_FP64 && bt->fallback == VGPU_METHOD_BICGSTAB) {
    for (size_t i = 0; i < bt->m.size(); ++i) {
      if (zero[i] || (status[i] != 1 && status[i] != 2 && status[i] != 3)) continue;
      vgpu_ctx *c = bt->m[i];
      const Segment &sg = bt->seg[i];
      int it = 0, st = 0;
      double e = 0, tr = 0;
      VGPU_HIP(hipSetDevice(c->device));
      rc = vgpu_residual_norm_dev(c, b + sg.off, x + sg.off, &tr);
      if (rc) return rc;
      if (status[i] == 1 && tr <= rtol) continue;
      if (!(tr <= 1.0)) {
        hipLaunchKernelGGL(k_fill, grid_for(sg.n), dim3(256), 0, c->stream, (int64_t)sg.n, 0.0,
                           x + sg.off);
        VGPU_HIP(hipGetLastError());
      }
      // BiCGSTAB stops on its recursive residual too: a pass that ends with
      // the true residual above rtol is followed by another from the current
      // iterate (its recurrence restarts from the true residual), at most
      // kFallbackPasses in all; a member still above rtol reports status 2
      constexpr int kFallbackPasses = 3;
      for (int pass = 0; pass < kFallbackPasses; ++pass) {
        rc = solve_stab<double>(c, c->stab64, b + sg.off, x + sg.off, c->diag, rtol,
                                bt->fallback_itmax, &it, &e, &st);
        if (rc) return rc;
        VGPU_HIP(hipStreamSynchronize(c->stream));
        iters[i] += it;
        errs[i] = e;
        status[i] = st;
        if (st != 1) break;
        rc = vgpu_residual_norm_dev(c, b + sg.off, x + sg.off, &tr);
        if (rc) return rc;
        if (tr <= rtol) break;
        if (pass == kFallbackPasses - 1) status[i] = 2;
      }
    }
    VGPU_HIP(hipSetDevice(bt->device));
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  for (size_t i = 0; i < bt->m.size(); ++i) {
    if (status[i] == 3) {
      set_error("batched BiCG: breakdown in system " + std::to_string(i));
      return VGPU_BREAKDOWN;
    }
  }
  for (size_t i = 0; i < bt->m.size(); ++i)
    if (status[i] != 1) {
      set_error("batched BiCG: system " + std::to_string(i) + " did not reach rtol");
      return VGPU_NOT_CONVERGED;
    }
  return VGPU_OK;
}

int vgpu_batch_set_fallback(vgpu_batch *bt, int method, int itmax) {
  if (!bt || (method != VGPU_METHOD_BICG && method != VGPU_METHOD_BICGSTAB) || itmax < 0) {
    set_error("vgpu_batch_set_fallback: method must be VGPU_METHOD_BICG (off) or "
              "VGPU_METHOD_BICGSTAB, itmax >= 0");
    return VGPU_ERR_ARG;
  }
  if (method == VGPU_METHOD_BICGSTAB) {
    for (vgpu_ctx *c : bt->m)
      if (!owner_fits(c->g)) {
        set_error("vgpu_batch_set_fallback: a member is too large for the BiCGSTAB owner "
                  "launches (8 N and 40 E must stay below 2^31 bytes)");
        return VGPU_ERR_ARG;
      }
  }
  bt->fallback = method;
  bt->fallback_itmax = itmax;
  return VGPU_OK;
}

int vgpu_batch_solve(vgpu_batch *bt, int precision, const double *const *b,
                     double *const *x, double rtol, int itmax, int *iters,
                     double *errs, int *status) {
  if (!bt || !b || !x) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(bt->device));
  for (size_t i = 0; i < bt->m.size(); ++i) {
    if (!b[i] || !x[i]) return VGPU_ERR_ARG;
    const Segment &sg = bt->seg[i];
    VGPU_HIP(hipMemcpyAsync(bt->b + sg.off, b[i], sg.n * sizeof(double),
                            hipMemcpyHostToDevice, bt->stream));
    VGPU_HIP(hipMemcpyAsync(bt->x + sg.off, x[i], sg.n * sizeof(double),
                            hipMemcpyHostToDevice, bt->stream));
  }
  const int rc = vgpu_batch_solve_dev(bt, precision, bt->b, bt->x, rtol, itmax, iters,
                                      errs, status, nullptr);
  if (rc == VGPU_OK || rc == VGPU_NOT_CONVERGED || rc == VGPU_BREAKDOWN) {
    for (size_t i = 0; i < bt->m.size(); ++i) {
      const Segment &sg = bt->seg[i];
      VGPU_HIP(hipMemcpyAsync(x[i], bt->x + sg.off, sg.n * sizeof(double),
                              hipMemcpyDeviceToHost, bt->stream));
    }
    VGPU_HIP(hipStreamSynchronize(bt->stream));
  }
  return rc;
}

int vgpu_true_residual(vgpu_ctx *c, const double *b, const double *x,
                       double *rel) {
  if (!c || !b || !x || !rel) return VGPU_ERR_ARG;
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  const int64_t N = c->g.N;
  VGPU_HIP(hipMemcpyAsync(c->vb, b, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  VGPU_HIP(hipMemcpyAsync(c->vx_, x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  OpData<double> op = c->op64(c->cH, c->cC);
  int rc = launch_apply<double, double>(op, VGPU_SCATTER_OWNER, c->zchunk,
                                        c->vx_, c->vt, nullptr, nullptr, nullptr,
                                        c->stream);
  if (rc) return rc;
  const int np = nparts_for(N);
  hipLaunchKernelGGL(k_resid, dim3(np), dim3(VB), 0, c->stream, N, c->vb, c->vt,
                     c->red);
  hipLaunchKernelGGL((k_dot<double>), dim3(np), dim3(VB), 0, c->stream, N,
                     nullptr, c->vb, c->vb, c->red + np);
  VGPU_HIP(hipGetLastError());
  const double rn = host_sum(c->red, np, c->stream, &rc);
  if (rc) return rc;
  const double bn = host_sum(c->red + np, np, c->stream, &rc);
  if (rc) return rc;
  *rel = std::sqrt(rn / bn);
  return VGPU_OK;
}

}  // extern "C"

namespace {
// Fortran Ew.d edit descriptor with w = 12, d = 5: [-]0.ddddd E+ee, right
// justified; exponents beyond 99 drop the E ([-]0.ddddd+eee); zero prints as
// 0.00000E+00; NaN and infinities right-justified, spelled as gfortran
// spells them in a 12-wide field (the reference's build compiles VANADIS.FOR
// with gfortran, makefile:296): "NaN", "Infinity", "-Infinity" (flang writes
// "Inf" / "-Inf" instead).
void fortran_e12_5(double v, char out[13]) {
  char t[40];
  if (std::isnan(v)) {
    std::snprintf(out, 13, "%12s", "NaN");
    return;
  }
  if (std::isinf(v)) {
    std::snprintf(out, 13, "%12s", v < 0 ? "-Infinity" : "Infinity");
    return;
  }
  const bool neg = std::signbit(v);
  int e10 = 0;
  char dig[6] = "00000";
  if (v != 0.0) {
    std::snprintf(t, sizeof t, "%.4e", std::fabs(v));  // d.dddde+XX
    dig[0] = t[0];
    for (int i = 0; i < 4; ++i) dig[i + 1] = t[2 + i];
    e10 = std::atoi(t + 7) + 1;
  }
  if (e10 >= -99 && e10 <= 99)
    std::snprintf(t, sizeof t, "%s0.%sE%c%02d", neg ? "-" : "", dig, e10 < 0 ? '-' : '+',
                  e10 < 0 ? -e10 : e10);
  else
    std::snprintf(t, sizeof t, "%s0.%s%c%03d", neg ? "-" : "", dig, e10 < 0 ? '-' : '+',
                  e10 < 0 ? -e10 : e10);
  std::snprintf(out, 13, "%12s", t);
}
}  // namespace

extern "C" {

int vgpu_write_tfile(const char *path, int64_t n, const double *xx, const double *yy,
                     const double *zz, const double *c) {
  if (!path || n < 0 || (n > 0 && (!xx || !yy || !zz || !c))) return VGPU_ERR_ARG;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("vgpu_write_tfile: cannot open ") + path);
    return VGPU_ERR_ARG;
  }
  char a[13], b[13], d[13], e[13];
  for (int64_t i = 0; i < n; ++i) {
    fortran_e12_5(xx[i], a);
    fortran_e12_5(yy[i], b);
    fortran_e12_5(zz[i], d);
    fortran_e12_5(c[i] * 1e6, e);
    std::fprintf(f, "%s  %s  %s  %s\n", a, b, d, e);
  }
  if (std::fclose(f) != 0) {
    set_error("vgpu_write_tfile: write failed");
    return VGPU_ERR_ARG;
  }
  return VGPU_OK;
}

int vgpu_ddf(vgpu_ctx *c, double *ddf, int *dominant) {
  if (!c || !ddf) return VGPU_ERR_ARG;
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  VGPU_HIP(hipSetDevice(c->device));
  constexpr int NB = 1024;
  double *pm = nullptr;
  int *pb = nullptr;
  VGPU_HIP(hipMalloc(&pm, NB * sizeof(double)));
  if (hipMalloc(&pb, NB * sizeof(int)) != hipSuccess) {
    hipFree(pm);
    set_error("vgpu_ddf: allocation failed");
    return VGPU_ERR_NOMEM;
  }
  OpData<double> op = c->op64(c->cH, c->cC);
  int rc = launch_ddf(op, pm, pb, NB, c->stream);
  std::vector<double> hm(NB);
  std::vector<int> hb(NB);
  if (!rc && (hipMemcpyAsync(hm.data(), pm, NB * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream) != hipSuccess ||
              hipMemcpyAsync(hb.data(), pb, NB * sizeof(int), hipMemcpyDeviceToHost,
                             c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess)) {
    set_error("vgpu_ddf: readback failed");
    rc = VGPU_ERR_HIP;
  }
  hipFree(pm);
  hipFree(pb);
  if (rc) return rc;
  double m = 0.0;
  int64_t bad = 0;
  for (int i = 0; i < NB; ++i) {
    m = std::max(m, hm[i]);
    bad += hb[i];
  }
  *ddf = m;
  if (dominant) *dominant = bad == 0;
  return VGPU_OK;
}

int vgpu_mass(vgpu_ctx *c, const double *cvec, double *mass) {
  if (!c || !cvec || !mass) return VGPU_ERR_ARG;
  const int64_t N = c->g.N;
  VGPU_HIP(hipMemcpyAsync(c->vx_, cvec, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const int np = 512;
  int rc = launch_mass(c->g, c->xc, c->yc, c->zc, c->vx_, c->red, np, c->stream);
  if (rc) return rc;
  *mass = host_sum(c->red, np, c->stream, &rc);
  return rc;
}

int vgpu_dev_alloc(vgpu_ctx *c, int64_t bytes, void **ptr) {
  if (!c || !ptr || bytes <= 0) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(c->device));
  VGPU_HIP(hipMalloc(ptr, bytes));
  c->user_allocs.push_back(*ptr);
  return VGPU_OK;
}

int vgpu_dev_free(vgpu_ctx *c, void *ptr) {
  if (!c || !ptr) return VGPU_ERR_ARG;
  for (size_t i = 0; i < c->user_allocs.size(); ++i)
    if (c->user_allocs[i] == ptr) {
      c->user_allocs.erase(c->user_allocs.begin() + i);
      VGPU_HIP(hipFree(ptr));
      return VGPU_OK;
    }
  set_error("vgpu_dev_free: pointer not allocated by this context");
  return VGPU_ERR_ARG;
}

int vgpu_memcpy_h2d(vgpu_ctx *c, void *dst, const void *src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return VGPU_ERR_ARG;
  VGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  VGPU_HIP(hipStreamSynchronize(c->stream));
  return VGPU_OK;
}

int vgpu_memcpy_d2h(vgpu_ctx *c, void *dst, const void *src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return VGPU_ERR_ARG;
  VGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  VGPU_HIP(hipStreamSynchronize(c->stream));
  return VGPU_OK;
}

int vgpu_autotune(vgpu_ctx *c, int precision, int reps, int *scatter,
                  int *tile, int *zchunk, double *best_ms) {
  if (!c || reps < 1 ||
      (precision != VGPU_FP64 && precision != VGPU_FP32 && precision != VGPU_MIXED)) {
    set_error("vgpu_autotune: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  // a mixed solve's work is its FP32 inner launches: the direction and S
  // launches of its inner BiCGSTAB (the FP32 BiCG iteration launch where the
  // system exceeds the owner launches' limits)
  const int tp = precision == VGPU_MIXED ? VGPU_FP32 : precision;
  const bool stab_pair = precision == VGPU_MIXED && owner_fits(c->g);
  // balanced launch, and z-chunk launches with chunks of naz / k planes (the
  // chunkings whose groups fill k rounds of workgroups evenly) and a few
  // fixed sizes
  std::vector<int> zcs = {0, 4, 8, 16};
  for (int k = 1; k <= 8; ++k) zcs.push_back((c->g.naz + k - 1) / k);
  std::sort(zcs.begin(), zcs.end());
  zcs.erase(std::unique(zcs.begin(), zcs.end()), zcs.end());
  const int save_tile = c->tile, save_zc = c->zchunk, save_sc = c->scatter;
  double best = 1e300;
  int bt = -1, bz = 0;
  for (int t = 0; t < 3; ++t)
    for (int zc : zcs) {
      if (zc > 256) continue;
      c->tile = t;
      c->zchunk = zc;
      double ms = 0;
      int rc = VGPU_OK;
      if (stab_pair) {
        double ms2 = 0;
        rc = vgpu_time_kernel(c, 5, VGPU_SCATTER_OWNER, tp, reps, &ms);
        if (!rc) rc = vgpu_time_kernel(c, 6, VGPU_SCATTER_OWNER, tp, reps, &ms2);
        ms += ms2;
      } else {
        rc = vgpu_time_kernel(c, 3, VGPU_SCATTER_OWNER, tp, reps, &ms);
      }
      if (rc == VGPU_ERR_ARG) continue;  // more partials than the timing buffer
      if (rc) {
        c->tile = save_tile;
        c->zchunk = save_zc;
        c->scatter = save_sc;
        return rc;
      }
      if (ms < best) {
        best = ms;
        bt = t;
        bz = zc;
      }
    }
  if (bt < 0) {
    c->tile = save_tile;
    c->zchunk = save_zc;
    set_error("vgpu_autotune: no candidate could be timed");
    return VGPU_ERR_ARG;
  }
  c->scatter = VGPU_SCATTER_OWNER;
  c->tile = bt;
  c->zchunk = bz;
  c->version++;
  if (scatter) *scatter = VGPU_SCATTER_OWNER;
  if (tile) *tile = bt;
  if (zchunk) *zchunk = bz;
  if (best_ms) *best_ms = best;
  return VGPU_OK;
}

int vgpu_autotune_bricks(vgpu_ctx *c, int scatter, int precision, int reps, int *tile,
                         int *zchunk, double *best_ms) {
  if (!c || reps < 1 || (precision != VGPU_FP64 && precision != VGPU_FP32) ||
      (scatter != VGPU_SCATTER_COLORED && scatter != VGPU_SCATTER_ATOMIC)) {
    set_error("vgpu_autotune_bricks: bad arguments");
    return VGPU_ERR_ARG;
  }
  if (!c->have_coeffs) return VGPU_ERR_STATE;
  // brick depths: the library's rule (0) and fixed layer counts
  const int depths[] = {0, 1, 2, 3, 4, 6, 8, 12, 16};
  const int save_tile = c->tile, save_zc = c->zchunk, save_sc = c->scatter;
  double best = 1e300;
  int bt = -1, bz = 0;
  for (int t = 0; t < 3; ++t)
    for (int zc : depths) {
      if (zc > c->g.ez) continue;
      c->tile = t;
      c->zchunk = zc;
      double ms = 0;
      const int rc = vgpu_time_kernel(c, 1, scatter, precision, reps, &ms);
      if (rc) {
        c->tile = save_tile;
        c->zchunk = save_zc;
        c->scatter = save_sc;
        return rc;
      }
      if (ms < best) {
        best = ms;
        bt = t;
        bz = zc;
      }
    }
  c->scatter = scatter;
  c->tile = bt;
  c->zchunk = bz;
  c->version++;
  if (tile) *tile = bt;
  if (zchunk) *zchunk = bz;
  if (best_ms) *best_ms = best;
  return VGPU_OK;
}

int vgpu_time_kernel(vgpu_ctx *c, int what, int scatter, int precision,
                     int reps, double *avg_ms) {
  if (!c || !avg_ms || reps < 1 || what < 0 || what > 7) return VGPU_ERR_ARG;
  if (what != 2 && what != 4 && what != 7 && !c->have_coeffs) return VGPU_ERR_STATE;
  if (what == 3 && scatter != VGPU_SCATTER_OWNER) return VGPU_ERR_ARG;
  if (what >= 5 && (scatter != VGPU_SCATTER_OWNER ||
                    (precision != VGPU_FP64 && precision != VGPU_FP32) || !owner_fits(c->g)))
    return VGPU_ERR_ARG;
  const int64_t N = c->g.N;
  VGPU_HIP(hipSetDevice(c->device));
  // deterministic non-trivial inputs
  hipLaunchKernelGGL(k_lincomb, grid_for(N), dim3(256), 0, c->stream, N, c->mask,
                     0.0, c->vb, 0.0, nullptr, c->vt);
  if (precision == VGPU_FP32) {
    if (int r = ensure_f32_vectors(c)) return r;
  }
  float *f1 = nullptr, *f2 = nullptr, *f3 = nullptr, *f4 = nullptr;
  if (precision == VGPU_FP32) {
    f1 = c->fb;
    f2 = c->fx;
    f3 = reinterpret_cast<float *>(c->vt);
    f4 = reinterpret_cast<float *>(c->vt2);
  }
  // fused iteration launch on one block of nine vectors, as the BiCG engine
  // lays them out (z, zz, previous directions, new directions, operator
  // outputs, x: the launch takes the block-addressed kernel the solver uses):
  // z, zz, po, ppo copied from the deterministic vectors vx_, vb, vt, vt2; a
  // status that never stops (tol 0, alpha 0) and unit partials, so beta = 1
  // and every system runs
  const Segment sg = make_segments({N}, {owner_units(c->g, c->tile)})[0];
  const int64_t tail = 2048;
  if (what == 3) {
    if (!c->tdir) {
      VGPU_HIP(hipMalloc(&c->tdir, (9 * N + tail) * sizeof(double)));
      VGPU_HIP(hipMemset(c->tdir, 0, 9 * N * sizeof(double)));
    }
    if (precision == VGPU_FP32) {
      float *v = reinterpret_cast<float *>(c->tdir);
      const float *src[] = {f1, f2, f3, f4};
      for (int i = 0; i < 4; ++i)
        VGPU_HIP(hipMemcpyAsync(v + i * N, src[i], N * sizeof(float),
                                hipMemcpyDeviceToDevice, c->stream));
    } else {
      const double *src[] = {c->vx_, c->vb, c->vt, c->vt2};
      for (int i = 0; i < 4; ++i)
        VGPU_HIP(hipMemcpyAsync(c->tdir + i * N, src[i], N * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
    }
    double *parts = c->tdir + 9 * N;
    hipLaunchKernelGGL(k_fill, grid_for(tail), dim3(256), 0, c->stream, tail, 1.0, parts);
    SolveStatus s0{};
    s0.iter = 1;
    VGPU_HIP(hipMemcpyAsync(parts, &s0, sizeof(SolveStatus), hipMemcpyHostToDevice,
                            c->stream));
    VGPU_HIP(hipStreamSynchronize(c->stream));
  }
  auto launch_dir = [&]() -> int {
    double *parts = c->tdir + 9 * N;
    const int ns = (int)((sizeof(SolveStatus) + 7) / 8);
    if (precision == VGPU_FP32) {
      float *v = reinterpret_cast<float *>(c->tdir);
      DirData<float> dd;
      dd.po = v + 2 * N;
      dd.ppo = v + 3 * N;
      dd.pn = v + 4 * N;
      dd.ppn = v + 5 * N;
      dd.x = v + 8 * N;
      dd.pr = dd.pb = dd.pbn = parts + ns;
      dd.stv = reinterpret_cast<SolveStatus *>(parts);
      dd.tol = 0.0;
      dd.divtol = 1e300;
      dd.itmax = 1 << 30;
      dd.NB = sg.nb;
      dd.q = 0;
      dd.blk = v;
      dd.blk_bytes = 9 * N * (int64_t)sizeof(float);
      Member<float> m0{};
      m0.op = c->op32();
      m0.seg = sg;
      return launch_owner_dir<float, float>(m0, nullptr, ActiveSet{}, false, c->tile,
                                            c->zchunk, m0.op.ra != 0.0, dd, v, v + N,
                                            v + 6 * N, v + 7 * N, c->tpart, c->stream);
    }
    double *v = c->tdir;
    DirData<double> dd;
    dd.po = v + 2 * N;
    dd.ppo = v + 3 * N;
    dd.pn = v + 4 * N;
    dd.ppn = v + 5 * N;
    dd.x = v + 8 * N;
    dd.pr = dd.pb = dd.pbn = parts + ns;
    dd.stv = reinterpret_cast<SolveStatus *>(parts);
    dd.tol = 0.0;
    dd.divtol = 1e300;
    dd.itmax = 1 << 30;
    dd.NB = sg.nb;
    dd.q = 0;
    dd.blk = v;
    dd.blk_bytes = 9 * N * (int64_t)sizeof(double);
    Member<double> m0{};
    m0.op = c->op64(c->cH, c->cC);
    m0.seg = sg;
    return launch_owner_dir<double, double>(m0, nullptr, ActiveSet{}, false, c->tile,
                                            c->zchunk, m0.op.ra != 0.0, dd, v, v + N,
                                            v + 6 * N, v + 7 * N, c->tpart, c->stream);
  };
  // what 5 / 6: the BiCGSTAB direction / S launch (FP64, or FP32 as the
  // inner iterations of a mixed solve run it) on the BiCGSTAB engine's
  // buffers, inputs from the deterministic vectors above, fixed reduction
  // partials and a status that never stops (tol 0), so every workgroup runs
  // its march
  auto stab_setup = [&](auto &e, const auto *dg) -> int {
    using TV = std::remove_const_t<std::remove_pointer_t<std::decay_t<decltype(dg)>>>;
    int rc = e.alloc(N, (int)owner_groups(c->g, c->tile, stab_zchunk(c)));
    if (rc) return rc;
    // partials 1 (r~.s: 2, so that rho = r~.s - omega r~.t is not 0)
    double *fills[] = {e.prt, e.prr, e.pbn, e.pga, e.prn, e.pts, e.ptt, e.prs, e.prq};
    const int64_t counts[] = {e.NB, e.NB, e.NB, e.cap, e.cap, e.cap, e.cap, e.cap, e.cap};
    for (int i = 0; i < 9; ++i)
      hipLaunchKernelGGL(k_fill, grid_for(counts[i]), dim3(256), 0, c->stream, counts[i],
                         fills[i] == e.prs ? 2.0 : 1.0, fills[i]);
    // iteration 1 of a solve that never stops (tol 0): both launches march
    SolveStatus s0{};
    s0.iter = s0.pad = 1;
    s0.rho = s0.alpha = s0.omega = s0.rho_next = 1.0;
    VGPU_HIP(hipMemcpyAsync(e.st, &s0, sizeof(SolveStatus), hipMemcpyHostToDevice,
                            c->stream));
    // inputs from the deterministic vectors, in the engine's own buffers (the
    // launches then address its block as a solve does): R = vx_, x = vt2
    const double *src[] = {c->vx_, c->vx_, c->vt2};
    TV *dst[] = {e.R[0], e.R[1], e.xw};
    for (int i = 0; i < 3; ++i)
      hipLaunchKernelGGL((k_convert<double, TV>), grid_for(N), dim3(256), 0, c->stream, N,
                         src[i], dst[i]);
    VGPU_HIP(hipGetLastError());
    VGPU_HIP(hipStreamSynchronize(c->stream));
    return VGPU_OK;
  };
  if (what >= 5) {
    int rc = ensure_diag(c);
    if (rc) return rc;
    rc = precision == VGPU_FP32 ? stab_setup(c->stab32, c->fdiag)
                                : stab_setup(c->stab64, c->diag);
    if (rc) return rc;
  }
  auto launch_stab = [&]() -> int {
    const int zc = stab_zchunk(c);
    if (precision == VGPU_FP32) {
      Stab<float> &e = c->stab32;
      OpData<float> op = c->op32();
      if (what == 5)
        return launch_owner_stab<float, float>(
            op, kStabDir, zc, e.dir_data(0, c->fdiag, e.xw, 0.0, 1 << 30, 1e300), e.R[1],
            e.V[0], c->stream);
      return launch_owner_stab<float, float>(
          op, kStabS, zc, e.s_data(0, c->fdiag, 0.0, 1 << 30, 1e300), e.R[0], e.T, c->stream);
    }
    Stab<double> &e = c->stab64;
    OpData<double> op = c->op64(c->cH, c->cC);
    if (what == 5)
      return launch_owner_stab<double, double>(
          op, kStabDir, zc, e.dir_data(0, c->diag, e.xw, 0.0, 1 << 30, 1e300), e.R[1],
          e.V[0], c->stream);
    return launch_owner_stab<double, double>(
        op, kStabS, zc, e.s_data(0, c->diag, 0.0, 1 << 30, 1e300), e.R[0], e.T, c->stream);
  };
  // what 4: a copy of 8 N doubles, the size class of a BiCG iteration's
  // working set (buffers held for this call only)
  // what 7: a copy of 64 N doubles (0.55 GB each way at 128x128x64, well
  // above the 256 MB last-level cache): the practical HBM roof
  // the copy buffer, graph and events of this call, released on every return
  struct Held {
    double *big = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Held() {
      if (gexec) hipGraphExecDestroy(gexec);
      if (graph) hipGraphDestroy(graph);
      if (e0) hipEventDestroy(e0);
      if (e1) hipEventDestroy(e1);
      if (big) hipFree(big);
    }
  } held;
  double *&big = held.big;
  if (what == 4) VGPU_HIP(hipMalloc(&big, 16 * N * sizeof(double)));
  if (what == 7) VGPU_HIP(hipMalloc(&big, 128 * N * sizeof(double)));
  auto launch = [&]() -> int {
    if (what == 3) return launch_dir();
    if (what == 7) {
      const int64_t n2 = 32 * N;  // 64 N doubles
      const int64_t g = std::min<int64_t>(32768, (n2 + 1023) / 1024);
      hipLaunchKernelGGL(k_copy_wide, dim3((unsigned)g), dim3(256), 0, c->stream, n2,
                         reinterpret_cast<const double2 *>(big),
                         reinterpret_cast<double2 *>(big + 64 * N));
      VGPU_HIP(hipGetLastError());
      return VGPU_OK;
    }
    if (what >= 5) return launch_stab();
    if (what == 4) {
      hipLaunchKernelGGL(k_copy, grid_for(8 * N, 256), dim3(256), 0, c->stream, 8 * N, big,
                         big + 8 * N);
      VGPU_HIP(hipGetLastError());
      return VGPU_OK;
    }
    if (what == 2) {
      hipLaunchKernelGGL(k_copy, grid_for(N, 256), dim3(256), 0, c->stream, N,
                         c->vx_, c->vb);
      VGPU_HIP(hipGetLastError());
      return VGPU_OK;
    }
    if (precision == VGPU_FP32) {
      OpData<float> op = c->op32();
      return launch_apply<float, float>(op, scatter, c->zchunk, f1, f2,
                                        what == 1 ? f3 : nullptr,
                                        what == 1 ? f4 : nullptr, c->tpart,
                                        c->stream);
    }
    OpData<double> op = c->op64(c->cH, c->cC);
    return launch_apply<double, double>(op, scatter, c->zchunk, c->vx_, c->vb,
                                        what == 1 ? c->vt : nullptr,
                                        what == 1 ? c->vt2 : nullptr,
                                        what == 1 ? c->tpart : nullptr, c->stream);
  };
  // dot partials: one per owner unit (brick modes: none)
  const int64_t nparts = std::max<int64_t>(owner_units(c->g, c->tile), 4096);
  if (c->tpart_n < nparts) {
    free_dev(c->tpart);
    c->tpart_n = 0;
    VGPU_HIP(hipMalloc(&c->tpart, nparts * sizeof(double)));
    c->tpart_n = nparts;
  }
  for (int i = 0; i < 3; ++i) {
    int rc = launch();
    if (rc) return rc;
  }
  // the timed launches are captured in one graph, so the events bracket GPU
  // work only (host enqueue gaps between launches do not count)
  VGPU_HIP(hipStreamSynchronize(c->stream));
  VGPU_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int lrc = VGPU_OK;
  for (int i = 0; i < reps && lrc == VGPU_OK; ++i) lrc = launch();
  const hipError_t ce = hipStreamEndCapture(c->stream, &held.graph);
  if (lrc) return lrc;
  VGPU_HIP(ce);
  VGPU_HIP(hipGraphInstantiate(&held.gexec, held.graph, nullptr, nullptr, 0));
  VGPU_HIP(hipEventCreate(&held.e0));
  VGPU_HIP(hipEventCreate(&held.e1));
  VGPU_HIP(hipGraphLaunch(held.gexec, c->stream));  // warm-up replay
  VGPU_HIP(hipEventRecord(held.e0, c->stream));
  VGPU_HIP(hipGraphLaunch(held.gexec, c->stream));
  VGPU_HIP(hipEventRecord(held.e1, c->stream));
  VGPU_HIP(hipEventSynchronize(held.e1));
  float ms = 0;
  VGPU_HIP(hipEventElapsedTime(&ms, held.e0, held.e1));
  *avg_ms = ms / reps;
  return VGPU_OK;
}

}  // extern "C"
