// Jacobi-preconditioned BiCG on the GPU (dbcg_simple, VANADIS.FOR:977-1055;
// kernel_dbcg_simple_, cudatest.cu:301-566), organised for gfx950:
//
//  * one iteration = 3 kernels (+ the operator's own): K1 direction update
//    (p = z + beta p, pp = zz + beta pp, convergence test of the previous
//    iteration), K2 the dual operator z = A p, zz = A^T pp (+ partials of
//    z.pp), K3 the fused update x += a p, r -= a z, rr -= a zz, z = r/diag,
//    zz = rr/diag with partials of r.r and z.rr;
//  * reductions are deterministic two-level sums: every producer block writes
//    one partial, and every block of the consuming kernel re-sums the partial
//    array in the same fixed order, so no atomics, no spin locks and no
//    cross-XCD fences are needed and all blocks agree bit-for-bit;
//  * partials are double-buffered by iteration parity, so a pair of
//    iterations is a fixed kernel sequence that is captured once in a
//    hipGraph and replayed; a device status word (done / iteration / err)
//    turns the remaining kernels of a replay into no-ops after convergence.
// Vectors are TV (double, or float for the inner solve of mixed precision);
// every scalar and partial is double.
#pragma once
#include <algorithm>
#include <functional>

#include "vgpu_internal.hpp"

namespace vgpu {

constexpr int VB = 1024;  // threads per vector block

// Block-wide sum of parts[0..n) in a fixed order; result valid in all threads.
__device__ __forceinline__ double block_sum(const double *__restrict__ parts,
                                            int n, double *red) {
  double s = 0.0;
  // eight loads in flight per step; the adds keep the sequential order
  const int bd = blockDim.x;
  int i = threadIdx.x;
  for (; i + 7 * bd < n; i += 8 * bd) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = parts[i + k * bd];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; i < n; i += bd) s += parts[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) t += red[w];
  __syncthreads();
  return t;
}

// Block-wide sum of a per-thread value, written by thread 0 to out[blockIdx].
__device__ __forceinline__ void block_partial(double v, double *out,
                                              double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

// r = b - r (r holds A x0), rr = r, z = r/d, zz = rr/d, x zeroed on the mask;
// partials of r.r, z.rr and b.b (cudatest.cu:464-491, VANADIS.FOR:991-1000);
// the solve's status starts at zero (written here, on the device, so a solve
// queues no host-to-device copy).
template <typename TV>
__global__ __launch_bounds__(VB) void k_init(int64_t n, const uint8_t *mask,
                                             const TV *b, TV *x, TV *r, TV *rr,
                                             TV *z, TV *zz, const TV *diag,
                                             double *pr, double *pb,
                                             double *pbn, SolveStatus *st) {
  __shared__ double red[VB / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *st = SolveStatus{};
  double sr = 0, sb = 0, sn = 0;
  for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * VB) {
    if (mask[i]) x[i] = TV(0);
    const double bi = b[i];
    const TV ri = TV(bi - (double)r[i]);
    r[i] = ri;
    rr[i] = ri;
    const TV zi = TV((double)ri / (double)diag[i]);
    z[i] = zi;
    zz[i] = zi;
    sr += (double)ri * (double)ri;
    sb += (double)zi * (double)ri;
    sn += bi * bi;
  }
  block_partial(sr, pr, red);
  block_partial(sb, pb, red);
  block_partial(sn, pbn, red);
}

// Elements per thread staged in registers before the block reductions, so
// the first chunk of vector traffic overlaps the partial-sum latency.
constexpr int PF = 4;

// K1: convergence test of iteration k-1, then p = z + beta p, pp = zz + beta pp.
// q = k & 1.  zero_out: also clear z, zz (brick scatter modes accumulate).
template <typename TV>
__global__ __launch_bounds__(VB) void k_dir(int64_t n, int nb, int q,
                                            const double *pr, const double *pb,
                                            const double *pbn, SolveStatus *st,
                                            double tol, int itmax,
                                            double divtol, TV *p, TV *pp,
                                            TV *z, TV *zz, int zero_out) {
  __shared__ double red[VB / 64];
  if (st->done) return;
  const int k = st->iter + 1;
  const int64_t stride = (int64_t)gridDim.x * VB;
  const int64_t i0 = blockIdx.x * (int64_t)VB + threadIdx.x;
  TV zr[PF], zzr[PF], pr0[PF], ppr0[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int64_t i = i0 + j * stride;
    const bool ok = i < n;
    zr[j] = ok ? z[i] : TV(0);
    zzr[j] = ok ? zz[i] : TV(0);
    pr0[j] = ok ? p[i] : TV(0);
    ppr0[j] = ok ? pp[i] : TV(0);
  }
  const double rn2 = block_sum(pr + (q ^ 1) * nb, nb, red);
  const double bn2 = block_sum(pbn, nb, red);
  const double err = sqrt(rn2 / bn2);
  int done = 0;
  if (err <= tol) done = 1;
  else if (k > itmax) done = 2;
  else if (!(err == err)) done = 3;
  else if (err > divtol) done = 4;
  double beta = 0.0;
  if (!done && k > 1) {
    const double bknum = block_sum(pb + (q ^ 1) * nb, nb, red);
    const double bkden = block_sum(pb + q * nb, nb, red);
    beta = bknum / bkden;
    if (!(beta == beta) || bkden == 0.0) done = 3;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->err = err;
    st->bnrm2 = bn2;
    if (done) st->done = done;
  }
  if (done) return;
  const TV bt = TV(beta);
  // iteration 1 copies (cudatest.cu:504-505): no 0 * stale-value products
  const bool first = k == 1;
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < n) {
      p[i] = first ? zr[j] : zr[j] + bt * pr0[j];
      pp[i] = first ? zzr[j] : zzr[j] + bt * ppr0[j];
      if (zero_out) {
        z[i] = TV(0);
        zz[i] = TV(0);
      }
    }
  }
  for (int64_t i = i0 + PF * stride; i < n; i += stride) {
    const TV zi = z[i], zzi = zz[i];
    p[i] = first ? zi : zi + bt * p[i];
    pp[i] = first ? zzi : zzi + bt * pp[i];
    if (zero_out) {
      z[i] = TV(0);
      zz[i] = TV(0);
    }
  }
}

// partials of a.b (akden = z.pp for the brick scatter modes)
template <typename TV>
__global__ __launch_bounds__(VB) void k_dot(int64_t n, const SolveStatus *st,
                                            const TV *a, const TV *b,
                                            double *part) {
  __shared__ double red[VB / 64];
  if (st && st->done) return;
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * VB)
    s += (double)a[i] * (double)b[i];
  block_partial(s, part, red);
}

// K3: alpha = bknum/akden; x += a p; r -= a z; rr -= a zz; z = r/d; zz = rr/d;
// partials of r.r and z.rr into parity q.
template <typename TV>
__device__ __forceinline__ void update_one(int64_t i, TV a, TV xi, TV pi, TV ri,
                                           TV rri, TV zi, TV zzi, TV di, TV *x,
                                           TV *r, TV *rr, TV *z, TV *zz,
                                           double &sr, double &sb) {
  x[i] = xi + a * pi;
  const TV rn = ri - a * zi;
  const TV rrn = rri - a * zzi;
  r[i] = rn;
  rr[i] = rrn;
  const TV zn = rn / di, zzn = rrn / di;
  z[i] = zn;
  zz[i] = zzn;
  sr += (double)rn * (double)rn;
  sb += (double)zn * (double)rrn;
}

template <typename TV>
__global__ __launch_bounds__(VB) void k_update(
    int64_t n, int nb, int na, int q, const double *pa, double *pr, double *pb,
    SolveStatus *st, TV *x, const TV *p, TV *r, TV *rr, TV *z, TV *zz,
    const TV *diag) {
  __shared__ double red[VB / 64];
  if (st->done) return;
  const int64_t stride = (int64_t)gridDim.x * VB;
  const int64_t i0 = blockIdx.x * (int64_t)VB + threadIdx.x;
  TV xv[PF], pv[PF], rv[PF], rrv[PF], zv[PF], zzv[PF], dv[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int64_t i = i0 + j * stride;
    const bool ok = i < n;
    xv[j] = ok ? x[i] : TV(0);
    pv[j] = ok ? p[i] : TV(0);
    rv[j] = ok ? r[i] : TV(0);
    rrv[j] = ok ? rr[i] : TV(0);
    zv[j] = ok ? z[i] : TV(0);
    zzv[j] = ok ? zz[i] : TV(0);
    dv[j] = ok ? diag[i] : TV(1);
  }
  const double akden = block_sum(pa, na, red);
  const double bknum = block_sum(pb + (q ^ 1) * nb, nb, red);
  const double ak = bknum / akden;
  if (!(ak == ak) || akden == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 3;
    return;
  }
  const TV a = TV(ak);
  double sr = 0, sb = 0;
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < n)
      update_one<TV>(i, a, xv[j], pv[j], rv[j], rrv[j], zv[j], zzv[j], dv[j], x, r,
                     rr, z, zz, sr, sb);
  }
  for (int64_t i = i0 + PF * stride; i < n; i += stride)
    update_one<TV>(i, a, x[i], p[i], r[i], rr[i], z[i], zz[i], diag[i], x, r, rr,
                   z, zz, sr, sb);
  block_partial(sr, pr + q * nb, red);
  block_partial(sb, pb + q * nb, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->iter = st->iter + 1;
}

// Dual operator callback: launches z = A p, zz = A^T pp on stream s, writing
// the akden partials (z.pp) to pa and returning their count in *na.  Kernels
// must return at once when *done != 0.
template <typename TV>
using DualOp = std::function<int(const TV *p, TV *z, const TV *pp, TV *zz,
                                 double *pa, int *na, const int *done,
                                 hipStream_t s)>;
// Single operator: y = A x (no done flag).
template <typename TV>
using SingleOp = std::function<int(const TV *x, TV *y, hipStream_t s)>;

// Solver workspace for vectors of type TV.
template <typename TV>
struct Bicg {
  int64_t n = 0;
  int nb = 0;        // vector-kernel blocks (= partial count)
  int na_max = 0;    // capacity of the akden partial array
  TV *r = nullptr, *rr = nullptr, *p = nullptr, *pp = nullptr, *z = nullptr,
     *zz = nullptr;
  double *pr = nullptr, *pb = nullptr, *pbn = nullptr, *pa = nullptr;
  SolveStatus *st = nullptr;       // device
  SolveStatus *st_host = nullptr;  // pinned
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int pairs_per_graph = 4;

  // (a failed allocation releases everything, n = 0: the next call starts
  // again)
  int alloc(int64_t n_, int na_cap) {
    const int rc = alloc_parts(n_, na_cap);
    if (rc) release();
    return rc;
  }
  int alloc_parts(int64_t n_, int na_cap) {
    release();
    n = n_;
    nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + VB - 1) / VB));
    na_max = std::max(na_cap, nb);
    VGPU_HIP(hipMalloc(&r, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&rr, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&p, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&pp, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&z, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&zz, n * sizeof(TV)));
    VGPU_HIP(hipMemset(p, 0, n * sizeof(TV)));
    VGPU_HIP(hipMemset(pp, 0, n * sizeof(TV)));
    VGPU_HIP(hipMalloc(&pr, 2 * nb * sizeof(double)));
    VGPU_HIP(hipMalloc(&pb, 2 * nb * sizeof(double)));
    VGPU_HIP(hipMalloc(&pbn, nb * sizeof(double)));
    VGPU_HIP(hipMalloc(&pa, na_max * sizeof(double)));
    VGPU_HIP(hipMemset(pb, 0, 2 * nb * sizeof(double)));
    VGPU_HIP(hipMalloc(&st, sizeof(SolveStatus)));
    VGPU_HIP(hipHostMalloc(&st_host, sizeof(SolveStatus), hipHostMallocDefault));
    return VGPU_OK;
  }
  void drop_graph() {
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    key_valid = false;
  }
  void release() {
    drop_graph();
    TV **v[] = {&r, &rr, &p, &pp, &z, &zz};
    for (auto pv : v)
      if (*pv) {
        hipFree(*pv);
        *pv = nullptr;
      }
    double **d[] = {&pr, &pb, &pbn, &pa};
    for (auto pd : d)
      if (*pd) {
        hipFree(*pd);
        *pd = nullptr;
      }
    if (st) hipFree(st);
    st = nullptr;
    if (st_host) hipHostFree(st_host);
    st_host = nullptr;
    n = 0;
  }
  ~Bicg() { release(); }

  // Solve A x = b with x, b, diag on the device.  op_key identifies the
  // operator (kernels + coefficient buffers); the captured graph is reused
  // while (op_key, buffers, tol, itmax) are unchanged.  zero_out: the dual
  // operator accumulates into zeroed outputs (brick scatter modes).
  int solve(uint64_t op_key, const DualOp<TV> &dual, const SingleOp<TV> &single,
            const uint8_t *mask, const TV *b, TV *x, const TV *diag,
            double tol, int itmax, bool zero_out, hipStream_t s, int *iters,
            double *err, int *status, double divtol = 1e300) {
    int rc = single(x, r, s);  // r = A x0 (x0 masked on load)
    if (rc) return rc;
    hipLaunchKernelGGL((k_init<TV>), dim3(nb), dim3(VB), 0, s, n, mask, b, x,
                       r, rr, z, zz, diag, pr, pb, pbn, st);
    VGPU_HIP(hipGetLastError());
    const bool same = key_valid && key_op == op_key && key_b == b &&
                      key_x == x && key_diag == diag && key_tol == tol &&
                      key_itmax == itmax && key_zero == zero_out &&
                      key_div == divtol;
    if (!same) {
      drop_graph();
      VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int crc = VGPU_OK;
      for (int pi = 0; pi < pairs_per_graph && crc == VGPU_OK; ++pi)
        for (int qq = 1; qq >= 0 && crc == VGPU_OK; --qq)
          crc = enqueue_iteration(dual, diag, x, tol, itmax, divtol, zero_out, qq, s);
      hipError_t ce = hipStreamEndCapture(s, &graph);
      if (crc) {
        drop_graph();
        return crc;
      }
      VGPU_HIP(ce);
      VGPU_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      key_valid = true;
      key_op = op_key;
      key_b = b;
      key_x = x;
      key_diag = diag;
      key_tol = tol;
      key_itmax = itmax;
      key_zero = zero_out;
      key_div = divtol;
    }
    // replay until the device reports done, at most two replays in flight
    const int max_replays = itmax / (2 * pairs_per_graph) + 2;
    hipEvent_t ev[2];
    VGPU_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    VGPU_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int result = VGPU_OK, launched = 0, completed = 0;
    for (;;) {
      while (launched - completed < 2 && launched < max_replays) {
        if (hipGraphLaunch(gexec, s) != hipSuccess ||
            hipMemcpyAsync(st_host, st, sizeof(SolveStatus),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(ev[launched & 1], s) != hipSuccess) {
          set_error("BiCG graph replay failed");
          result = VGPU_ERR_HIP;
          break;
        }
        ++launched;
      }
      if (result || completed == launched) break;
      if (hipEventSynchronize(ev[completed & 1]) != hipSuccess) {
        set_error("BiCG event wait failed");
        result = VGPU_ERR_HIP;
        break;
      }
      ++completed;
      if (st_host->done) break;
    }
    if (hipStreamSynchronize(s) != hipSuccess && !result) {
      set_error("BiCG stream sync failed");
      result = VGPU_ERR_HIP;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (result) return result;
    *iters = st_host->iter;
    *err = st_host->err;
    *status = st_host->done;
    return VGPU_OK;
  }

 private:
  bool key_valid = false;
  uint64_t key_op = 0;
  const void *key_b = nullptr, *key_x = nullptr, *key_diag = nullptr;
  double key_tol = 0;
  int key_itmax = 0;
  bool key_zero = false;
  double key_div = 0;

  int enqueue_iteration(const DualOp<TV> &dual, const TV *diag, TV *x,
                        double tol, int itmax, double divtol, bool zero_out,
                        int q, hipStream_t s) {
    hipLaunchKernelGGL((k_dir<TV>), dim3(nb), dim3(VB), 0, s, n, nb, q, pr, pb,
                       pbn, st, tol, itmax, divtol, p, pp, z, zz, (int)zero_out);
    int na = 0;
    int rc = dual(p, z, pp, zz, pa, &na, &st->done, s);
    if (rc) return rc;
    if (na > na_max) {
      set_error("akden partial buffer too small");
      return VGPU_ERR_STATE;
    }
    hipLaunchKernelGGL((k_update<TV>), dim3(nb), dim3(VB), 0, s, n, nb, na, q,
                       pa, pr, pb, st, x, p, r, rr, z, zz, diag);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
};

}  // namespace vgpu
