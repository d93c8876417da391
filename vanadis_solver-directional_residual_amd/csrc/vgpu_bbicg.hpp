// Batched Jacobi-preconditioned BiCG for the owner (matrix-free) operator:
// nm independent systems solved together, each with the iteration of
// dbcg_simple (VANADIS.FOR:977-1055) and the stopping rule of
// kernel_dbcg_simple_ (recursive ||r||/||b|| <= tol, cudatest.cu:491-553).
//
// Layout.  The systems' vectors are concatenated (system m at nodes
// [off_m, off_m + n_m)).  System m owns the vector workgroups
// [vb0_m, vb0_m + nb_m), nb_m = min(512, ceil(n_m / 2048)), each striding over
// the system's own range exactly as a solve of that system alone would, and
// its reduction partials sit at those workgroup indices; the operator writes
// its z.pp partials per owner unit.  Every scalar of system m is therefore
// summed in the same order whatever else is in the batch: a member's iterates
// are bitwise those of its solve alone (same tile shape).
//
// One iteration = K1 (one owner launch over all systems: convergence test,
// beta and the new directions fused into the dual operator) and K2 (update);
// pairs of iterations are captured in one hipGraph and replayed until every
// system reports done (converged, itmax, breakdown or divergence).  Kernels
// skip the systems that are done.
#pragma once
#include <algorithm>
#include <functional>
#include <vector>

#include "vgpu_bicg.hpp"

namespace vgpu {

// Active system of vector workgroup b and the workgroup's index within it.
struct BlockOf {
  int m, j;
};
__device__ __forceinline__ BlockOf block_of(const int *__restrict__ act,
                                            const int *__restrict__ actvb, int nact, int b) {
  int lo = 0, hi = nact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (actvb[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return {act[lo], b - actvb[lo]};
}

// Block-wide sum of a per-thread value, written by thread 0 to out[slot].
__device__ __forceinline__ void block_partial_at(double v, double *out, int slot,
                                                 double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) t += red[w];
    out[slot] = t;
  }
}

#define VGPU_ACT_ARGS \
  const int *__restrict__ act, const int *__restrict__ actvb, int nact

// Vector workgroups of the batched engine: BVB threads, at most BNB per
// system (about four nodes per thread; a large system gets two workgroups of
// 8 waves per CU).  Every owner launch of an iteration re-sums a system's nb
// partials of r.r, b.b and z.rr in each of its workgroups, all reading the
// same lines at the launch's start: fewer partials shorten that read, while
// workgroups of 16 waves made the update launch itself slower (their block
// reductions and barriers); 512 x 512 is the measured best of 256 x 1024,
// 512 x 512 and 1024 x 256 per iteration.
constexpr int BVB = 512;
constexpr int BNB = 512;

// r = b - A x0 (ax holds A x0), rr = r; stores only the preconditioned
// z = r/d, zz = rr/d (the engine carries z, zz: r = d z) and the engine's
// iterate xw = x0 zeroed on the mask; partials of r.r, z.rr and b.b
// (cudatest.cu:464-491, VANADIS.FOR:991-1000); each system's initial (zero)
// status, written on the device.
template <typename TV>
__global__ __launch_bounds__(BVB) void kb_init(SolveStatus *__restrict__ stv,
                                              const Segment *__restrict__ seg, VGPU_ACT_ARGS,
                                              const uint8_t *mask, const TV *b,
                                              const TV *x, TV *xw, const TV *ax, TV *z,
                                              TV *zz, const TV *diag,
                                              double *pr, double *pb,
                                              double *pbn) {
  __shared__ double red[BVB / 64];
  const BlockOf bo = block_of(act, actvb, nact, blockIdx.x);
  const Segment sg = seg[bo.m];
  const int j = bo.j;
  if (j == 0 && threadIdx.x == 0) stv[bo.m] = SolveStatus{};
  double sr = 0, sb = 0, sn = 0;
  for (int64_t li = j * (int64_t)BVB + threadIdx.x; li < sg.n;
       li += (int64_t)sg.nb * BVB) {
    const int64_t i = sg.off + li;
    xw[i] = mask[i] ? TV(0) : x[i];
    const double bi = b[i];
    const TV ri = TV(bi - (double)ax[i]);
    const TV zi = TV((double)ri / (double)diag[i]);
    z[i] = zi;
    zz[i] = zi;
    sr += (double)ri * (double)ri;
    sb += (double)zi * (double)ri;
    sn += bi * bi;
  }
  block_partial_at(sr, pr, sg.vb0 + j, red);
  block_partial_at(sb, pb, sg.vb0 + j, red);
  block_partial_at(sn, pbn, sg.vb0 + j, red);
}

// y = x over n values
template <typename TV>
__global__ __launch_bounds__(BVB) void kb_copy(int64_t n, const TV *__restrict__ x,
                                              TV *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * BVB;
  for (int64_t i = blockIdx.x * (int64_t)BVB + threadIdx.x; i < n; i += stride) y[i] = x[i];
}

// Block-wide sums of two per-thread values, each in the order block_sum
// uses (shuffles within waves, waves in order); results in every thread.
__device__ __forceinline__ void block_sum_pair(double u, double v, double (*red2)[BVB / 64],
                                               double &su, double &sv) {
  for (int off = 32; off > 0; off >>= 1) {
    u += __shfl_down(u, off, 64);
    v += __shfl_down(v, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red2[0][threadIdx.x >> 6] = u;
    red2[1][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  su = sv = 0.0;
  for (int w = 0; w < BVB / 64; ++w) {
    su += red2[0][w];
    sv += red2[1][w];
  }
  __syncthreads();
}

// One node of K2: the residual update r -= a w, rr -= a ww carried on
// z = r/d, zz = rr/d (w = A p, ww = A^T pp); partials of r.r and z.rr with
// r = d z, rr = d zz.  (x += a p is applied by the next iteration launch.)
template <typename TV>
__device__ __forceinline__ void update_z(int64_t i, TV a, TV zi, TV zzi, TV wi, TV wwi,
                                         TV di, TV *z, TV *zz, double &sr, double &sb) {
  const TV zn = zi - a * (wi / di), zzn = zzi - a * (wwi / di);
  z[i] = zn;
  zz[i] = zzn;
  const double rn = (double)zn * (double)di, rrn = (double)zzn * (double)di;
  sr += rn * rn;
  sb += (double)zn * rrn;
}

// K2: alpha = bknum/akden (akden = sum of the system's owner partials in
// order); z -= a w/d; zz -= a ww/d; partials of r.r and z.rr into parity q;
// alpha to the status for the next iteration launch's x += alpha p.
template <typename TV>
__global__ __launch_bounds__(BVB) void kb_update(const Segment *__restrict__ seg,
                                                 VGPU_ACT_ARGS, int NB, int q,
                                                 const double *__restrict__ pa,
                                                 double *__restrict__ pr,
                                                 double *__restrict__ pb,
                                                 SolveStatus *__restrict__ stv,
                                                 const TV *__restrict__ w,
                                                 const TV *__restrict__ ww,
                                                 TV *__restrict__ z, TV *__restrict__ zz,
                                                 const TV *__restrict__ diag) {
  __shared__ double red[BVB / 64];
  __shared__ double red2[2][BVB / 64];
  __shared__ int sst[2];
  const BlockOf bo = block_of(act, actvb, nact, blockIdx.x);
  SolveStatus *st = stv + bo.m;
  const Segment sg = seg[bo.m];
  const int j = bo.j;
  // the status, read once per block (block 0 of the system may write it
  // below); a result the iteration launch found becomes `done` here
  if (threadIdx.x == 0) {
    sst[0] = st->done;
    sst[1] = st->pend;
  }
  __syncthreads();
  if (sst[0] || sst[1]) {
    if (!sst[0] && j == 0 && threadIdx.x == 0) st->done = sst[1];
    return;
  }
  const int64_t stride = (int64_t)sg.nb * BVB;
  const int tid = threadIdx.x;
  const int64_t i0 = j * (int64_t)BVB + tid;
  // the owner launch's w.pp partials (one per (chunk, tile) group of a
  // z-chunk launch, else one per unit) and the z.rr partials: their first
  // loads are issued before the node loads, so alpha is known as soon as
  // they return (loads complete in order) and the first nodes' stores need
  // not wait for the last nodes' loads
  const int na = (int)(sg.ng > 0 ? sg.ng : sg.nu);
  const double *pas = pa + sg.u0, *pbs = pb + (q ^ 1) * NB + sg.vb0;
  constexpr int KA = 2, KB = 1;
  double av[KA], bv[KB];
  // unconditional loads at clamped indices (na, nb >= 1), then selects
#pragma unroll
  for (int k = 0; k < KA; ++k) av[k] = pas[min(tid + k * BVB, na - 1)];
#pragma unroll
  for (int k = 0; k < KB; ++k) bv[k] = pbs[min(tid + k * BVB, sg.nb - 1)];
#pragma unroll
  for (int k = 0; k < KA; ++k) av[k] = tid + k * BVB < na ? av[k] : 0.0;
#pragma unroll
  for (int k = 0; k < KB; ++k) bv[k] = tid + k * BVB < sg.nb ? bv[k] : 0.0;
  TV zv[PF], zzv[PF], wv[PF], wwv[PF], dv[PF];
#pragma unroll
  for (int t = 0; t < PF; ++t) {
    const int64_t li = i0 + t * stride;
    const bool ok = li < sg.n;
    const int64_t i = sg.off + (ok ? li : 0);
    zv[t] = ok ? z[i] : TV(0);
    zzv[t] = ok ? zz[i] : TV(0);
    wv[t] = ok ? w[i] : TV(0);
    wwv[t] = ok ? ww[i] : TV(0);
    dv[t] = ok ? diag[i] : TV(1);
  }
  // per thread in index order (as block_sum), then over the block
  double sa = 0.0, sbk = 0.0;
#pragma unroll
  for (int k = 0; k < KA; ++k) sa += av[k];
  for (int i = tid + KA * BVB; i < na; i += BVB) sa += pas[i];
#pragma unroll
  for (int k = 0; k < KB; ++k) sbk += bv[k];
  for (int i = tid + KB * BVB; i < sg.nb; i += BVB) sbk += pbs[i];
  double akden, bknum;
  block_sum_pair(sa, sbk, red2, akden, bknum);
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
      update_z<TV>(sg.off + li, a, zv[t], zzv[t], wv[t], wwv[t], dv[t], z, zz, sr, sb);
  }
  for (int64_t li = i0 + PF * stride; li < sg.n; li += stride) {
    const int64_t i = sg.off + li;
    update_z<TV>(i, a, z[i], zz[i], w[i], ww[i], diag[i], z, zz, sr, sb);
  }
  block_partial_at(sr, pr + q * NB, sg.vb0 + j, red);
  block_partial_at(sb, pb + q * NB, sg.vb0 + j, red);
  if (j == 0 && threadIdx.x == 0) {
    st->alpha = ak;
    st->iter = st->iter + 1;
  }
}

// Operator callbacks of the batched engine.  The direction-and-dual operator
// (K1 of an iteration) tests convergence, forms p = z + beta po and
// pp = zz + beta ppo into dd.pn / dd.ppn and applies w = A p, ww = A^T pp
// with the per-unit w.pp partials in pa, skipping systems that are done; the
// single operator (y = A x for every system) ignores the status.
template <typename TV>
using BDirOp = std::function<int(const TV *z, const TV *zz, const DirData<TV> &dd,
                                 TV *w, TV *ww, double *pa, const ActiveSet &as,
                                 hipStream_t s)>;
template <typename TV>
using BSingleOp = std::function<int(const TV *x, TV *y, const ActiveSet &as,
                                    hipStream_t s)>;

// Nodes of a system rounded up so that every system starts on a 256-node
// boundary of the batch vectors.
inline int64_t seg_pad(int64_t n) { return (n + 255) / 256 * 256; }

// Segments of nm systems with n[m] nodes, nu[m] owner units and ng[m] owner
// groups (empty: none).
inline std::vector<Segment> make_segments(const std::vector<int64_t> &n,
                                          const std::vector<int64_t> &nu,
                                          const std::vector<int64_t> &ng = {}) {
  std::vector<Segment> s(n.size());
  int64_t off = 0, u0 = 0;
  int vb0 = 0;
  for (size_t m = 0; m < n.size(); ++m) {
    s[m].off = off;
    s[m].n = n[m];
    s[m].u0 = u0;
    s[m].nu = nu[m];
    s[m].vb0 = vb0;
    s[m].nb = (int)std::min<int64_t>(BNB, std::max<int64_t>(1, (n[m] + 4 * BVB - 1) / (4 * BVB)));
    s[m].ng = ng.empty() ? 0 : ng[m];
    off += seg_pad(n[m]);
    u0 += nu[m];
    vb0 += s[m].nb;
  }
  return s;
}

// Device copy of an active set (member list and prefix sums).
struct ActiveBuf {
  int *act = nullptr, *actvb = nullptr;
  int64_t *actu = nullptr, *actg = nullptr;
  ActiveSet view;
  std::vector<int> members;

  int alloc(int nm) {
    VGPU_HIP(hipMalloc(&act, nm * sizeof(int)));
    VGPU_HIP(hipMalloc(&actvb, (nm + 1) * sizeof(int)));
    VGPU_HIP(hipMalloc(&actu, (nm + 1) * sizeof(int64_t)));
    VGPU_HIP(hipMalloc(&actg, (nm + 1) * sizeof(int64_t)));
    return VGPU_OK;
  }
  // members (increasing) of segments s; synchronous upload
  int set(const std::vector<Segment> &s, const std::vector<int> &mem) {
    members = mem;
    std::vector<int> vb(mem.size() + 1, 0);
    std::vector<int64_t> u(mem.size() + 1, 0), g(mem.size() + 1, 0);
    for (size_t i = 0; i < mem.size(); ++i) {
      vb[i + 1] = vb[i] + s[mem[i]].nb;
      u[i + 1] = u[i] + s[mem[i]].nu;
      g[i + 1] = g[i] + s[mem[i]].ng;
    }
    if (!mem.empty()) {
      VGPU_HIP(hipMemcpy(act, mem.data(), mem.size() * sizeof(int), hipMemcpyHostToDevice));
      VGPU_HIP(hipMemcpy(actvb, vb.data(), vb.size() * sizeof(int), hipMemcpyHostToDevice));
      VGPU_HIP(hipMemcpy(actu, u.data(), u.size() * sizeof(int64_t), hipMemcpyHostToDevice));
      VGPU_HIP(hipMemcpy(actg, g.data(), g.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    view.act = act;
    view.actvb = actvb;
    view.actu = actu;
    view.actg = actg;
    view.nact = (int)mem.size();
    view.nvb = vb.back();
    view.nu = u.back();
    view.ng = g.back();
    return VGPU_OK;
  }
  void release() {
    if (act) hipFree(act);
    if (actvb) hipFree(actvb);
    if (actu) hipFree(actu);
    if (actg) hipFree(actg);
    act = actvb = nullptr;
    actu = actg = nullptr;
    view = ActiveSet{};
    members.clear();
  }
};

template <typename TV>
struct BBicg {
  int nm = 0, NB = 0;
  int64_t ntot = 0, utot = 0;
  std::vector<Segment> hseg;
  Segment *seg = nullptr;
  ActiveBuf full, cur;  // every system / the systems the graph launches
  // directions by iteration parity (an iteration reads those of the previous
  // one while neighbouring tiles still load them), operator outputs w, ww,
  // preconditioned residuals z, zz (the residuals are d z, d zz); ax holds
  // A x0 at the start
  TV *ax = nullptr, *p[2] = {nullptr, nullptr}, *pp[2] = {nullptr, nullptr},
     *z = nullptr, *zz = nullptr, *w = nullptr, *ww = nullptr;
  TV *xw = nullptr;  // the iterates (copied to the caller's x at the end)
  VecBlock vblk;     // one allocation for the ten vectors above
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
          s[m].nu != hseg[m].nu || s[m].vb0 != hseg[m].vb0 || s[m].nb != hseg[m].nb ||
          s[m].ng != hseg[m].ng)
        return false;
    return true;
  }

  // A failed allocation releases everything, so that the next call
  // allocates again instead of finding a half-built engine.
  int alloc(const std::vector<Segment> &s) {
    if (seg && matches(s)) return VGPU_OK;
    const int rc = alloc_parts(s);
    if (rc) release();
    return rc;
  }
  int alloc_parts(const std::vector<Segment> &s) {
    release();
    hseg = s;
    nm = (int)s.size();
    NB = s.back().vb0 + s.back().nb;
    ntot = s.back().off + seg_pad(s.back().n);
    utot = s.back().u0 + s.back().nu;
    VGPU_HIP(hipMalloc(&seg, nm * sizeof(Segment)));
    VGPU_HIP(hipMemcpy(seg, s.data(), nm * sizeof(Segment), hipMemcpyHostToDevice));
    int rc = full.alloc(nm);
    if (!rc) rc = cur.alloc(nm);
    if (rc) return rc;
    std::vector<int> all(nm);
    for (int m = 0; m < nm; ++m) all[m] = m;
    rc = full.set(hseg, all);
    if (!rc) rc = cur.set(hseg, all);
    if (rc) return rc;
    TV **v[] = {&ax, &p[0], &p[1], &pp[0], &pp[1], &z, &zz, &w, &ww, &xw};
    rc = vblk.alloc<TV>(v, 10, ntot);
    if (rc) {
      set_error("batched BiCG: vector allocation failed");
      return rc;
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
    TV **v[] = {&ax, &p[0], &p[1], &pp[0], &pp[1], &z, &zz, &w, &ww, &xw};
    for (auto pv : v) *pv = nullptr;
    vblk.release();
    double **d[] = {&pr, &pb, &pbn, &pa};
    for (auto pd : d)
      if (*pd) {
        hipFree(*pd);
        *pd = nullptr;
      }
    if (seg) hipFree(seg);
    seg = nullptr;
    full.release();
    cur.release();
    if (st) hipFree(st);
    st = nullptr;
    if (st_host) hipHostFree(st_host);
    st_host = nullptr;
    nm = 0;
    hseg.clear();
  }
  ~BBicg() { release(); }

  // Solve every system with x, b, diag, mask on the device (batch layout).
  // op_key identifies the operators; the iterations run on the engine's own
  // buffers (x0 copied in by the init kernel, the result copied out at the
  // end), so the graph is reused while the key, diag, tol, itmax, divtol and
  // the active set are unchanged, whatever b and x.  When a
  // quarter of the launched systems have finished, the launches are narrowed
  // to the running ones (graph re-captured between replays).  Per-system
  // results go to iters/err/status [nm].
  int solve(uint64_t op_key, const BDirOp<TV> &dual, const BSingleOp<TV> &single,
            const uint8_t *mask, const TV *b, TV *x, const TV *diag, double tol,
            int itmax, hipStream_t s, int *iters, double *err, int *status,
            double divtol = 1e300) {
    int rc = single(x, ax, full.view, s);  // A x0 (x0 masked on load)
    if (rc) return rc;
    hipLaunchKernelGGL((kb_init<TV>), dim3(full.view.nvb), dim3(BVB), 0, s, st, seg,
                       full.view.act, full.view.actvb, full.view.nact, mask, b, x, xw, ax,
                       z, zz, diag, pr, pb, pbn);
    VGPU_HIP(hipGetLastError());
    if ((int)cur.members.size() != nm) {  // start with every system launched
      VGPU_HIP(hipStreamSynchronize(s));
      drop_graph();
      std::vector<int> all(nm);
      for (int m = 0; m < nm; ++m) all[m] = m;
      rc = cur.set(hseg, all);
      if (rc) return rc;
    }
    const bool same = key_valid && key_op == op_key && key_diag == diag && key_tol == tol &&
                      key_itmax == itmax && key_div == divtol;
    if (!same) {
      rc = capture(op_key, dual, diag, tol, itmax, divtol, s);
      if (rc) return rc;
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
      int running = 0;
      for (int m = 0; m < nm; ++m) running += st_host[m].done == 0;
      if (running == 0) break;
      if (4 * running <= 3 * (int)cur.members.size() && launched < max_replays) {
        // drain, then launch only the running systems
        if (hipStreamSynchronize(s) != hipSuccess) {
          set_error("batched BiCG stream sync failed");
          result = VGPU_ERR_HIP;
          break;
        }
        completed = launched;
        std::vector<int> run;
        for (int m = 0; m < nm; ++m)
          if (st_host[m].done == 0) run.push_back(m);
        if (run.empty()) break;
        result = cur.set(hseg, run);
        if (!result)
          result = capture(op_key, dual, diag, tol, itmax, divtol, s);
        if (result) break;
      }
    }
    if (!result) {  // the iterates to the caller (stream order: after the replays)
      const int64_t nx = hseg.back().off + hseg.back().n;
      hipLaunchKernelGGL((kb_copy<TV>), dim3(BNB), dim3(BVB), 0, s, nx, xw, x);
      if (hipGetLastError() != hipSuccess) {
        set_error("batched BiCG result copy failed");
        result = VGPU_ERR_HIP;
      }
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
  const void *key_diag = nullptr;
  double key_tol = 0;
  int key_itmax = 0;
  double key_div = 0;

  int capture(uint64_t op_key, const BDirOp<TV> &dual, const TV *diag, double tol,
              int itmax, double divtol, hipStream_t s) {
    drop_graph();
    VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = VGPU_OK;
    for (int pi = 0; pi < pairs_per_graph && crc == VGPU_OK; ++pi)
      for (int q = 1; q >= 0 && crc == VGPU_OK; --q)
        crc = enqueue_iteration(dual, diag, xw, tol, itmax, divtol, q, s);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (crc) {
      drop_graph();
      return crc;
    }
    VGPU_HIP(ce);
    VGPU_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    key_valid = true;
    key_op = op_key;
    key_diag = diag;
    key_tol = tol;
    key_itmax = itmax;
    key_div = divtol;
    return VGPU_OK;
  }

  int enqueue_iteration(const BDirOp<TV> &dual, const TV *diag, TV *x, double tol,
                        int itmax, double divtol, int q, hipStream_t s) {
    const ActiveSet &a = cur.view;
    DirData<TV> dd;
    dd.po = p[q ^ 1];
    dd.ppo = pp[q ^ 1];
    dd.pn = p[q];
    dd.ppn = pp[q];
    dd.x = x;
    dd.pr = pr;
    dd.pb = pb;
    dd.pbn = pbn;
    dd.stv = st;
    dd.tol = tol;
    dd.divtol = divtol;
    dd.itmax = itmax;
    dd.NB = NB;
    dd.q = q;
    dd.blk = vblk.base;
    dd.blk_bytes = vblk.bytes;
    int rc = dual(z, zz, dd, w, ww, pa, a, s);
    if (rc) return rc;
    hipLaunchKernelGGL((kb_update<TV>), dim3(a.nvb), dim3(BVB), 0, s, seg, a.act,
                       a.actvb, a.nact, NB, q, pa, pr, pb, st, w, ww, z, zz, diag);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
};

}  // namespace vgpu
