// Batched Jacobi-preconditioned BiCG for the owner (matrix-free) operator:
// nm independent systems solved together, each with the iteration of
// dbcg_simple (VANADIS.FOR:977-1055) and the stopping rule of
// kernel_dbcg_simple_ (recursive ||r||/||b|| <= tol, cudatest.cu:491-553).
//
// Layout.  The systems' vectors are concatenated (system m at nodes
// [off_m, off_m + n_m)).  System m owns the vector workgroups
// [vb0_m, vb0_m + nb_m), nb_m = min(256, ceil(n_m / 1024)), each striding over
// the system's own range exactly as a solve of that system alone would, and
// its reduction partials sit at those workgroup indices; the operator writes
// its z.pp partials per owner unit.  Every scalar of system m is therefore
// summed in the same order whatever else is in the batch: a member's iterates
// are bitwise those of its solve alone (same tile shape).
//
// One iteration = K1 (direction), K2 (one persistent owner launch over all
// systems), K3 (update); pairs of iterations are captured in one hipGraph and
// replayed until every system reports done (converged, itmax, breakdown or
// divergence).  Kernels skip the systems that are done.
#pragma once
#include <algorithm>
#include <functional>
#include <vector>

#include "vgpu_bicg.hpp"

namespace vgpu {

// system of a vector workgroup (segments sorted by vb0)
__device__ __forceinline__ int seg_of_block(const Segment *seg, int nm, int b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].vb0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// r = b - r (r holds A x0), rr = r, z = r/d, zz = rr/d, x zeroed on the mask;
// partials of r.r, z.rr and b.b (cudatest.cu:464-491, VANADIS.FOR:991-1000).
template <typename TV>
__global__ __launch_bounds__(VB) void kb_init(const Segment *seg, int nm,
                                              const uint8_t *mask, const TV *b,
                                              TV *x, TV *r, TV *rr, TV *z,
                                              TV *zz, const TV *diag,
                                              double *pr, double *pb,
                                              double *pbn) {
  __shared__ double red[VB / 64];
  const Segment sg = seg[seg_of_block(seg, nm, blockIdx.x)];
  const int j = blockIdx.x - sg.vb0;
  double sr = 0, sb = 0, sn = 0;
  for (int64_t li = j * (int64_t)VB + threadIdx.x; li < sg.n;
       li += (int64_t)sg.nb * VB) {
    const int64_t i = sg.off + li;
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

// K1: convergence test of iteration k-1, then p = z + beta p, pp = zz + beta pp
// (iteration 1 copies, cudatest.cu:504-505).  q = k & 1; NB = all workgroups.
template <typename TV>
__global__ __launch_bounds__(VB) void kb_dir(const Segment *seg, int nm, int NB,
                                             int q, const double *pr,
                                             const double *pb, const double *pbn,
                                             SolveStatus *stv, double tol,
                                             int itmax, double divtol, TV *p,
                                             TV *pp, const TV *z,
                                             const TV *zz) {
  __shared__ double red[VB / 64];
  const int m = seg_of_block(seg, nm, blockIdx.x);
  SolveStatus *st = stv + m;
  if (st->done) return;
  const Segment sg = seg[m];
  const int j = blockIdx.x - sg.vb0;
  const int k = st->iter + 1;
  const int64_t stride = (int64_t)sg.nb * VB;
  const int64_t i0 = j * (int64_t)VB + threadIdx.x;
  TV zr[PF], zzr[PF], pr0[PF], ppr0[PF];
#pragma unroll
  for (int t = 0; t < PF; ++t) {
    const int64_t li = i0 + t * stride;
    const bool ok = li < sg.n;
    const int64_t i = sg.off + (ok ? li : 0);
    zr[t] = ok ? z[i] : TV(0);
    zzr[t] = ok ? zz[i] : TV(0);
    pr0[t] = ok ? p[i] : TV(0);
    ppr0[t] = ok ? pp[i] : TV(0);
  }
  const double rn2 = block_sum(pr + (q ^ 1) * NB + sg.vb0, sg.nb, red);
  const double bn2 = block_sum(pbn + sg.vb0, sg.nb, red);
  const double err = sqrt(rn2 / bn2);
  int done = 0;
  if (err <= tol) done = 1;
  else if (k > itmax) done = 2;
  else if (!(err == err)) done = 3;
  else if (err > divtol) done = 4;
  double beta = 0.0;
  if (!done && k > 1) {
    const double bknum = block_sum(pb + (q ^ 1) * NB + sg.vb0, sg.nb, red);
    const double bkden = block_sum(pb + q * NB + sg.vb0, sg.nb, red);
    beta = bknum / bkden;
    if (!(beta == beta) || bkden == 0.0) done = 3;
  }
  if (j == 0 && threadIdx.x == 0) {
    st->err = err;
    st->bnrm2 = bn2;
    if (done) st->done = done;
  }
  if (done) return;
  const TV bt = TV(beta);
  const bool first = k == 1;
#pragma unroll
  for (int t = 0; t < PF; ++t) {
    const int64_t li = i0 + t * stride;
    if (li < sg.n) {
      const int64_t i = sg.off + li;
      p[i] = first ? zr[t] : zr[t] + bt * pr0[t];
      pp[i] = first ? zzr[t] : zzr[t] + bt * ppr0[t];
    }
  }
  for (int64_t li = i0 + PF * stride; li < sg.n; li += stride) {
    const int64_t i = sg.off + li;
    const TV zi = z[i], zzi = zz[i];
    p[i] = first ? zi : zi + bt * p[i];
    pp[i] = first ? zzi : zzi + bt * pp[i];
  }
}

// K3: alpha = bknum/akden (akden = sum of the system's owner-unit partials in
// unit order); x += a p; r -= a z; rr -= a zz; z = r/d; zz = rr/d; partials of
// r.r and z.rr into parity q.
template <typename TV>
__global__ __launch_bounds__(VB) void kb_update(const Segment *seg, int nm,
                                                int NB, int q, const double *pa,
                                                double *pr, double *pb,
                                                SolveStatus *stv, TV *x,
                                                const TV *p, TV *r, TV *rr,
                                                TV *z, TV *zz, const TV *diag) {
  __shared__ double red[VB / 64];
  const int m = seg_of_block(seg, nm, blockIdx.x);
  SolveStatus *st = stv + m;
  if (st->done) return;
  const Segment sg = seg[m];
  const int j = blockIdx.x - sg.vb0;
  const int64_t stride = (int64_t)sg.nb * VB;
  const int64_t i0 = j * (int64_t)VB + threadIdx.x;
  TV xv[PF], pv[PF], rv[PF], rrv[PF], zv[PF], zzv[PF], dv[PF];
#pragma unroll
  for (int t = 0; t < PF; ++t) {
    const int64_t li = i0 + t * stride;
    const bool ok = li < sg.n;
    const int64_t i = sg.off + (ok ? li : 0);
    xv[t] = ok ? x[i] : TV(0);
    pv[t] = ok ? p[i] : TV(0);
    rv[t] = ok ? r[i] : TV(0);
    rrv[t] = ok ? rr[i] : TV(0);
    zv[t] = ok ? z[i] : TV(0);
    zzv[t] = ok ? zz[i] : TV(0);
    dv[t] = ok ? diag[i] : TV(1);
  }
  const double akden = block_sum(pa + sg.u0, (int)sg.nu, red);
  const double bknum = block_sum(pb + (q ^ 1) * NB + sg.vb0, sg.nb, red);
  const double ak = bknum / akden;
  if (!(ak == ak) || akden == 0.0) {
    if (j == 0 && threadIdx.x == 0) st->done = 3;
    return;
  }
  const TV a = TV(ak);
  double sr = 0, sb = 0;
#pragma unroll
  for (int t = 0; t < PF; ++t) {
    const int64_t li = i0 + t * stride;
    if (li < sg.n)
      update_one<TV>(sg.off + li, a, xv[t], pv[t], rv[t], rrv[t], zv[t], zzv[t],
                     dv[t], x, r, rr, z, zz, sr, sb);
  }
  for (int64_t li = i0 + PF * stride; li < sg.n; li += stride) {
    const int64_t i = sg.off + li;
    update_one<TV>(i, a, x[i], p[i], r[i], rr[i], z[i], zz[i], diag[i], x, r, rr,
                   z, zz, sr, sb);
  }
  block_partial(sr, pr + q * NB, red);
  block_partial(sb, pb + q * NB, red);
  if (j == 0 && threadIdx.x == 0) st->iter = st->iter + 1;
}

// Operator callbacks of the batched engine: the dual operator writes the
// per-unit z.pp partials to pa and skips systems whose status is done; the
// single operator (y = A x for every system) ignores the status.
template <typename TV>
using BDualOp = std::function<int(const TV *p, TV *z, const TV *pp, TV *zz,
                                  double *pa, const SolveStatus *st,
                                  hipStream_t s)>;
template <typename TV>
using BSingleOp = std::function<int(const TV *x, TV *y, hipStream_t s)>;

// Nodes of a system rounded up so that every system starts on a 256-node
// boundary of the batch vectors.
inline int64_t seg_pad(int64_t n) { return (n + 255) / 256 * 256; }

// Segments of nm systems with n[m] nodes and nu[m] owner units.
inline std::vector<Segment> make_segments(const std::vector<int64_t> &n,
                                          const std::vector<int64_t> &nu) {
  std::vector<Segment> s(n.size());
  int64_t off = 0, u0 = 0;
  int vb0 = 0;
  for (size_t m = 0; m < n.size(); ++m) {
    s[m].off = off;
    s[m].n = n[m];
    s[m].u0 = u0;
    s[m].nu = nu[m];
    s[m].vb0 = vb0;
    s[m].nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n[m] + VB - 1) / VB));
    off += seg_pad(n[m]);
    u0 += nu[m];
    vb0 += s[m].nb;
  }
  return s;
}

template <typename TV>
struct BBicg {
  int nm = 0, NB = 0;
  int64_t ntot = 0, utot = 0;
  std::vector<Segment> hseg;
  Segment *seg = nullptr;
  TV *r = nullptr, *rr = nullptr, *p = nullptr, *pp = nullptr, *z = nullptr,
     *zz = nullptr;
  double *pr = nullptr, *pb = nullptr, *pbn = nullptr, *pa = nullptr;
  SolveStatus *st = nullptr;       // device [nm]
  SolveStatus *st_host = nullptr;  // pinned [nm]
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int pairs_per_graph = 4;

  bool matches(const std::vector<Segment> &s) const {
    if ((int)s.size() != nm) return false;
    for (int m = 0; m < nm; ++m)
      if (s[m].off != hseg[m].off || s[m].n != hseg[m].n || s[m].u0 != hseg[m].u0 ||
          s[m].nu != hseg[m].nu || s[m].vb0 != hseg[m].vb0 || s[m].nb != hseg[m].nb)
        return false;
    return true;
  }

  int alloc(const std::vector<Segment> &s) {
    if (seg && matches(s)) return VGPU_OK;
    release();
    hseg = s;
    nm = (int)s.size();
    NB = s.back().vb0 + s.back().nb;
    ntot = s.back().off + seg_pad(s.back().n);
    utot = s.back().u0 + s.back().nu;
    VGPU_HIP(hipMalloc(&seg, nm * sizeof(Segment)));
    VGPU_HIP(hipMemcpy(seg, s.data(), nm * sizeof(Segment), hipMemcpyHostToDevice));
    TV **v[] = {&r, &rr, &p, &pp, &z, &zz};
    for (auto pv : v) {
      VGPU_HIP(hipMalloc(pv, ntot * sizeof(TV)));
      VGPU_HIP(hipMemset(*pv, 0, ntot * sizeof(TV)));
    }
    VGPU_HIP(hipMalloc(&pr, 2 * NB * sizeof(double)));
    VGPU_HIP(hipMalloc(&pb, 2 * NB * sizeof(double)));
    VGPU_HIP(hipMalloc(&pbn, NB * sizeof(double)));
    VGPU_HIP(hipMalloc(&pa, std::max<int64_t>(1, utot) * sizeof(double)));
    VGPU_HIP(hipMemset(pb, 0, 2 * NB * sizeof(double)));
    VGPU_HIP(hipMalloc(&st, nm * sizeof(SolveStatus)));
    VGPU_HIP(hipHostMalloc(&st_host, nm * sizeof(SolveStatus), hipHostMallocDefault));
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
    if (seg) hipFree(seg);
    seg = nullptr;
    if (st) hipFree(st);
    st = nullptr;
    if (st_host) hipHostFree(st_host);
    st_host = nullptr;
    nm = 0;
    hseg.clear();
  }
  ~BBicg() { release(); }

  // Solve every system with x, b, diag, mask on the device (batch layout).
  // op_key identifies the operators; the graph is reused while the key,
  // buffers, tol, itmax and divtol are unchanged.  Per-system results go to
  // iters/err/status [nm].
  int solve(uint64_t op_key, const BDualOp<TV> &dual, const BSingleOp<TV> &single,
            const uint8_t *mask, const TV *b, TV *x, const TV *diag, double tol,
            int itmax, hipStream_t s, int *iters, double *err, int *status,
            double divtol = 1e300) {
    int rc = single(x, r, s);  // r = A x0 (x0 masked on load)
    if (rc) return rc;
    hipLaunchKernelGGL((kb_init<TV>), dim3(NB), dim3(VB), 0, s, seg, nm, mask, b,
                       x, r, rr, z, zz, diag, pr, pb, pbn);
    VGPU_HIP(hipGetLastError());
    std::vector<SolveStatus> s0(nm, SolveStatus{});
    VGPU_HIP(hipMemcpyAsync(st, s0.data(), nm * sizeof(SolveStatus),
                            hipMemcpyHostToDevice, s));
    const bool same = key_valid && key_op == op_key && key_b == b && key_x == x &&
                      key_diag == diag && key_tol == tol && key_itmax == itmax &&
                      key_div == divtol;
    if (!same) {
      drop_graph();
      VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int crc = VGPU_OK;
      for (int pi = 0; pi < pairs_per_graph && crc == VGPU_OK; ++pi)
        for (int q = 1; q >= 0 && crc == VGPU_OK; --q)
          crc = enqueue_iteration(dual, diag, x, tol, itmax, divtol, q, s);
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
      key_div = divtol;
    }
    // replay until every system is done, at most two replays in flight
    const int max_replays = itmax / (2 * pairs_per_graph) + 2;
    hipEvent_t ev[2];
    VGPU_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    VGPU_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int result = VGPU_OK, launched = 0, completed = 0;
    for (;;) {
      while (launched - completed < 2 && launched < max_replays) {
        if (hipGraphLaunch(gexec, s) != hipSuccess ||
            hipMemcpyAsync(st_host, st, nm * sizeof(SolveStatus),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(ev[launched & 1], s) != hipSuccess) {
          set_error("batched BiCG graph replay failed");
          result = VGPU_ERR_HIP;
          break;
        }
        ++launched;
      }
      if (result || completed == launched) break;
      if (hipEventSynchronize(ev[completed & 1]) != hipSuccess) {
        set_error("batched BiCG event wait failed");
        result = VGPU_ERR_HIP;
        break;
      }
      ++completed;
      bool all = true;
      for (int m = 0; m < nm; ++m) all = all && st_host[m].done != 0;
      if (all) break;
    }
    if (hipStreamSynchronize(s) != hipSuccess && !result) {
      set_error("batched BiCG stream sync failed");
      result = VGPU_ERR_HIP;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (result) return result;
    for (int m = 0; m < nm; ++m) {
      iters[m] = st_host[m].iter;
      err[m] = st_host[m].err;
      status[m] = st_host[m].done;
    }
    return VGPU_OK;
  }

 private:
  bool key_valid = false;
  uint64_t key_op = 0;
  const void *key_b = nullptr, *key_x = nullptr, *key_diag = nullptr;
  double key_tol = 0;
  int key_itmax = 0;
  double key_div = 0;

  int enqueue_iteration(const BDualOp<TV> &dual, const TV *diag, TV *x, double tol,
                        int itmax, double divtol, int q, hipStream_t s) {
    hipLaunchKernelGGL((kb_dir<TV>), dim3(NB), dim3(VB), 0, s, seg, nm, NB, q, pr,
                       pb, pbn, st, tol, itmax, divtol, p, pp, z, zz);
    int rc = dual(p, z, pp, zz, pa, st, s);
    if (rc) return rc;
    hipLaunchKernelGGL((kb_update<TV>), dim3(NB), dim3(VB), 0, s, seg, nm, NB, q,
                       pa, pr, pb, st, x, p, r, rr, z, zz, diag);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
};

}  // namespace vgpu
