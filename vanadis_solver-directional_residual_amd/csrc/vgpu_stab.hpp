// Jacobi-preconditioned BiCGSTAB (van der Vorst 1992, right preconditioning
// by the diagonal D) for one system on the owner operator, with FP64
// reductions and scalars whatever the vector type TV.  The reference solves
// with BiCG (dbcg_simple, VANADIS.FOR:977-1055); BiCGSTAB is the inner
// method of the mixed-precision solve and an optional method of the FP64 and
// FP32 solves, with the reference's stopping rule (recursive
// ||r||/||b|| <= tol, cudatest.cu:491-553).
//
// The recurrences run on the scaled vectors R = D^-1 r, P = D^-1 p,
// V = D^-1 A P, S = D^-1 s, T = D^-1 t and the shadow RT = D r0 (r0 the
// initial residual, so r0 . r = RT . R), iteration j:
//   rho_j = r~.s - omega r~.t (= RT . R_j);  beta = (rho_j / rho_{j-1}) (alpha / omega)
//   S = R - alpha V;  R_j = S - omega T;  x += alpha P + omega S
//   P_j = R_j + beta (P - omega V);  V_j = D^-1 A P_j;  alpha_j = rho_j / (RT . V_j)
//   S_j = R_j - alpha_j V_j;  T_j = D^-1 A S_j;  omega_j = (t . s) / (t . t)
// with t = d T, s = d S.  A step that solves the system exactly leaves S = 0
// and t . t = 0: omega and beta are then taken as 0 (R_j = S = 0, x gains
// alpha P), and the convergence test ends the solve (Barrett et al.,
// Templates, 1994, sec. 2.3.8: stop when s is small, with x += alpha p).
// One iteration = the kStabDir owner launch (R_j,
// P_j and x formed as the plane nodes are loaded, V_j, partials of RT . V_j
// and |r_j|^2) and the kStabS owner launch (the test of |r_j|, alpha_j, S_j
// formed on load, T_j, partials of t.s, t.t, r~.s, r~.t); pairs of
// iterations are captured in one hipGraph and replayed until the status
// reports done.  R, P and V alternate between two buffers by iteration
// parity: a launch reads the previous ones at its tiles' halo while it
// stores the new ones.
#pragma once
#include <algorithm>
#include <functional>

#include "vgpu_bbicg.hpp"
#include "vgpu_bicg.hpp"

namespace vgpu {

// threads of the vector workgroups and at most this many of them (= partials,
// re-summed by the owner launches: few and large, as in the BiCG engine)
constexpr int SVB = 1024;
constexpr int SNB = 256;

inline int stab_blocks(int64_t n) {
  return (int)std::min<int64_t>(SNB, std::max<int64_t>(1, (n + 4 * SVB - 1) / (4 * SVB)));
}

// Block sums of two per-thread values (shuffles within waves, waves in
// order), written by thread 0 to a[slot], b[slot].
__device__ __forceinline__ void stab_partial2(double u, double v, double *a, double *b,
                                              int slot) {
  __shared__ double red[2][SVB / 64];
  for (int off = 32; off > 0; off >>= 1) {
    u += __shfl_down(u, off, 64);
    v += __shfl_down(v, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = u;
    red[1][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double su = 0.0, sv = 0.0;
    for (int w = 0; w < SVB / 64; ++w) {
      su += red[0][w];
      sv += red[1][w];
    }
    a[slot] = su;
    b[slot] = sv;
  }
}

// r = b - A x0 (ax = A x0), R = r / d, RT = d r (shadow r0 = r); the
// engine's iterate xw = x0 zeroed on the mask; partials of RT . R
// (= r0 . r0), |r|^2 and |b|^2; the solve's initial status (written on the
// device: no host copy on the stream).
template <typename TV>
__global__ __launch_bounds__(SVB) void ks_init(SolveStatus *__restrict__ st, int64_t n,
                                              const uint8_t *__restrict__ mask,
                                              const TV *__restrict__ b,
                                              const TV *__restrict__ x, TV *__restrict__ xw,
                                              const TV *__restrict__ ax,
                                              const TV *__restrict__ diag, TV *__restrict__ R,
                                              TV *__restrict__ RT, double *__restrict__ prt,
                                              double *__restrict__ prr,
                                              double *__restrict__ pbn) {
  __shared__ double red[SVB / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SolveStatus s0{};
    s0.rho = s0.alpha = s0.omega = 1.0;
    *st = s0;
  }
  double srt = 0, sr = 0, sb = 0;
  const int64_t stride = (int64_t)gridDim.x * SVB;
  for (int64_t i = blockIdx.x * (int64_t)SVB + threadIdx.x; i < n; i += stride) {
    xw[i] = mask[i] ? TV(0) : x[i];
    const TV bi = b[i], di = diag[i];
    const TV ri = bi - ax[i];
    const TV Ri = ri / di, RTi = ri * di;
    R[i] = Ri;
    RT[i] = RTi;
    srt += (double)RTi * (double)Ri;
    sr += (double)ri * (double)ri;
    sb += (double)bi * (double)bi;
  }
  stab_partial2(srt, sr, prt, prr, blockIdx.x);
  for (int off = 32; off > 0; off >>= 1) sb += __shfl_down(sb, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sb;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < SVB / 64; ++w) t += red[w];
    pbn[blockIdx.x] = t;
  }
}

// y = x over n values
template <typename TV>
__global__ __launch_bounds__(SVB) void ks_copy(int64_t n, const TV *__restrict__ x,
                                              TV *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * SVB;
  for (int64_t i = blockIdx.x * (int64_t)SVB + threadIdx.x; i < n; i += stride) y[i] = x[i];
}

// Operator callbacks: y = A x of the system (ignores the status), and one
// BiCGSTAB owner launch (mode kStabDir or kStabS).
template <typename TV>
using StabSingleOp = std::function<int(const TV *x, TV *y, hipStream_t s)>;
template <typename TV>
using StabOwnerOp = std::function<int(int mode, const StabData<TV> &sd, const TV *r, TV *y,
                                      hipStream_t s)>;

template <typename TV>
struct Stab {
  int64_t n = 0;
  int NB = 0, na = 0;
  TV *ax = nullptr, *RT = nullptr, *T = nullptr;
  TV *xw = nullptr;  // the iterate (copied to the caller's x at the end)
  TV *R[2] = {nullptr, nullptr}, *P[2] = {nullptr, nullptr}, *V[2] = {nullptr, nullptr};
  VecBlock vblk;  // one allocation for the ten vectors above
  double *prt = nullptr, *prr = nullptr, *pbn = nullptr;  // initial partials [NB]
  // direction launch partials (RT.V, |r|^2), S launch partials (t.s, t.t,
  // r~.s, r~.t) [cap]
  double *pga = nullptr, *prn = nullptr;
  double *pts = nullptr, *ptt = nullptr, *prs = nullptr, *prq = nullptr;
  SolveStatus *st = nullptr, *st_host = nullptr;
  // graphs of 1, 2 and 4 iteration pairs, captured on demand: a solve replays
  // the size that the previous solve's iteration count suggests, so short
  // (time-step) solves launch few no-op iterations past their end and long
  // ones few graph launches
  static constexpr int kSizes = 3;
  hipGraphExec_t gexec[kSizes] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[kSizes] = {nullptr, nullptr, nullptr};
  int last_iters = -1;

  int cap = 0;  // capacity of the owner partial arrays

  // n nodes, na owner dot partials per launch (the launches' count: units or
  // (chunk, tile) groups)
  // (a failed allocation releases everything: the next call starts again)
  int alloc(int64_t nn, int nna) {
    nna = std::max(1, nna);
    if (RT && nn == n && nna <= cap) {
      if (nna != na) drop_graph();
      na = nna;
      return VGPU_OK;
    }
    const int rc = alloc_parts(nn, nna);
    if (rc) release();
    return rc;
  }
  int alloc_parts(int64_t nn, int nna) {
    release();
    n = nn;
    NB = stab_blocks(n);
    na = cap = nna;
    TV **v[] = {&ax, &RT, &T, &xw, &R[0], &R[1], &P[0], &P[1], &V[0], &V[1]};
    if (vblk.alloc<TV>(v, 10, n) != VGPU_OK) {
      set_error("BiCGSTAB: vector allocation failed");
      return VGPU_ERR_NOMEM;
    }
    double **pn[] = {&prt, &prr, &pbn};
    for (auto pd : pn) VGPU_HIP(hipMalloc(pd, NB * sizeof(double)));
    double **pa[] = {&pga, &prn, &pts, &ptt, &prs, &prq};
    for (auto pd : pa) {
      VGPU_HIP(hipMalloc(pd, cap * sizeof(double)));
      VGPU_HIP(hipMemset(*pd, 0, cap * sizeof(double)));
    }
    VGPU_HIP(hipMalloc(&st, sizeof(SolveStatus)));
    VGPU_HIP(hipHostMalloc(&st_host, sizeof(SolveStatus), hipHostMallocDefault));
    return VGPU_OK;
  }
  void drop_graph() {
    for (int i = 0; i < kSizes; ++i) {
      if (gexec[i]) hipGraphExecDestroy(gexec[i]);
      if (graph[i]) hipGraphDestroy(graph[i]);
      gexec[i] = nullptr;
      graph[i] = nullptr;
    }
    key_valid = false;
  }
  static int pairs_of(int size) { return 1 << size; }
  void release() {
    drop_graph();
    TV **v[] = {&ax, &RT, &T, &xw, &R[0], &R[1], &P[0], &P[1], &V[0], &V[1]};
    for (auto pv : v) *pv = nullptr;
    vblk.release();
    double **d[] = {&prt, &prr, &pbn, &pga, &prn, &pts, &ptt, &prs, &prq};
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
    NB = na = cap = 0;
  }
  ~Stab() { release(); }

  // Solve A x = b (device vectors; x in/out, zeroed on the mask).  The
  // iterations run on the engine's own buffers (x0 copied in by the init
  // kernel, the result copied out at the end), so the graph is reused while
  // op_key, diag, tol, itmax and divtol are unchanged, whatever b and x.
  int solve(uint64_t op_key, const StabSingleOp<TV> &single, const StabOwnerOp<TV> &owner,
            const uint8_t *mask, const TV *b, TV *x, const TV *diag, double tol, int itmax,
            double divtol, hipStream_t s, int *iters, double *err, int *status) {
    int rc = single(x, ax, s);  // A x0 (x0 masked on load)
    if (rc) return rc;
    // R_0 goes to the buffer the first captured iteration (parity 1) reads
    hipLaunchKernelGGL((ks_init<TV>), dim3(NB), dim3(SVB), 0, s, st, n, mask, b, x, xw, ax,
                       diag, R[0], RT, prt, prr, pbn);
    VGPU_HIP(hipGetLastError());
    const bool same = key_valid && key_op == op_key && key_diag == diag && key_tol == tol &&
                      key_itmax == itmax && key_div == divtol;
    if (!same) drop_graph();
    const int size = last_iters < 0 ? 2 : last_iters <= 6 ? 0 : last_iters <= 24 ? 1 : 2;
    if (!gexec[size]) {
      rc = capture(op_key, owner, diag, tol, itmax, divtol, size, s);
      if (rc) return rc;
    }
    hipGraphExec_t gx = gexec[size];
    // replay until done, at most two replays in flight
    const int max_replays = itmax / (2 * pairs_of(size)) + 2;
    hipEvent_t ev[2];
    VGPU_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    VGPU_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int result = VGPU_OK, launched = 0, completed = 0;
    for (;;) {
      while (launched - completed < 2 && launched < max_replays) {
        if (hipGraphLaunch(gx, s) != hipSuccess ||
            hipMemcpyAsync(st_host, st, sizeof(SolveStatus), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipEventRecord(ev[launched & 1], s) != hipSuccess) {
          set_error("BiCGSTAB graph replay failed");
          result = VGPU_ERR_HIP;
          break;
        }
        ++launched;
      }
      if (result || completed == launched) break;
      if (hipEventSynchronize(ev[completed & 1]) != hipSuccess) {
        set_error("BiCGSTAB event wait failed");
        result = VGPU_ERR_HIP;
        break;
      }
      ++completed;
      if (st_host->done) break;
    }
    if (!result) {  // the iterate to the caller (stream order: after the replays)
      hipLaunchKernelGGL((ks_copy<TV>), dim3(NB), dim3(SVB), 0, s, n, xw, x);
      if (hipGetLastError() != hipSuccess) {
        set_error("BiCGSTAB result copy failed");
        result = VGPU_ERR_HIP;
      }
    }
    if (hipStreamSynchronize(s) != hipSuccess && !result) {
      set_error("BiCGSTAB stream sync failed");
      result = VGPU_ERR_HIP;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (result) return result;
    *iters = st_host->iter;
    *err = st_host->err;
    *status = st_host->done;
    last_iters = st_host->iter;
    return VGPU_OK;
  }

 private:
  bool key_valid = false;
  uint64_t key_op = 0;
  const void *key_diag = nullptr;
  double key_tol = 0, key_div = 0;
  int key_itmax = 0;

  int capture(uint64_t op_key, const StabOwnerOp<TV> &owner, const TV *diag, double tol,
              int itmax, double divtol, int size, hipStream_t s) {
    VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = VGPU_OK;
    for (int pi = 0; pi < pairs_of(size) && crc == VGPU_OK; ++pi)
      for (int q = 1; q >= 0 && crc == VGPU_OK; --q)
        crc = enqueue_iteration(owner, diag, xw, tol, itmax, divtol, q, size < 2, s);
    hipError_t ce = hipStreamEndCapture(s, &graph[size]);
    if (crc) {
      drop_graph();
      return crc;
    }
    VGPU_HIP(ce);
    VGPU_HIP(hipGraphInstantiate(&gexec[size], graph[size], nullptr, nullptr, 0));
    key_valid = true;
    key_op = op_key;
    key_diag = diag;
    key_tol = tol;
    key_itmax = itmax;
    key_div = divtol;
    return VGPU_OK;
  }

  // early: the launches of short solves' graphs return at once after the end
  int enqueue_iteration(const StabOwnerOp<TV> &owner, const TV *diag, TV *x, double tol,
                        int itmax, double divtol, int q, bool early, hipStream_t s) {
    StabData<TV> a = dir_data(q, diag, x, tol, itmax, divtol);
    StabData<TV> b = s_data(q, diag, tol, itmax, divtol);
    a.early = b.early = early ? 1 : 0;
    int rc = owner(kStabDir, a, R[q ^ 1], V[q], s);
    if (rc) return rc;
    return owner(kStabS, b, R[q], T, s);
  }

 public:
  // Launch data of the direction launch (input R[q^1], output V[q]) and of
  // the S launch (input R[q], output T) of an iteration of parity q.
  StabData<TV> dir_data(int q, const TV *diag, TV *x, double tol, int itmax,
                        double divtol) const {
    StabData<TV> sa;
    sa.po = P[q ^ 1];
    sa.vo = V[q ^ 1];
    sa.to = T;
    sa.pn = P[q];
    sa.rn = R[q];
    sa.x = x;
    sa.d = diag;
    sa.rt = RT;
    double *const ain[4] = {pts, ptt, prs, prq};
    for (int i = 0; i < 4; ++i) {
      sa.pin[i] = ain[i];
      sa.nin[i] = na;
    }
    sa.prt0 = prt;
    sa.nb0 = NB;
    sa.pout[0] = pga;
    sa.pout[1] = prn;
    sa.st = st;
    sa.blk = vblk.base;
    sa.blk_bytes = vblk.bytes;
    sa.tol = tol;
    sa.divtol = divtol;
    sa.itmax = itmax;
    return sa;
  }
  StabData<TV> s_data(int q, const TV *diag, double tol, int itmax, double divtol) const {
    StabData<TV> sb;
    sb.v = V[q];
    sb.d = diag;
    sb.rt = RT;
    sb.pin[0] = pga;
    sb.pin[1] = prn;
    sb.pin[2] = pbn;
    sb.nin[0] = sb.nin[1] = na;
    sb.nin[2] = NB;
    double *const bout[4] = {pts, ptt, prs, prq};
    for (int i = 0; i < 4; ++i) sb.pout[i] = bout[i];
    sb.st = st;
    sb.blk = vblk.base;
    sb.blk_bytes = vblk.bytes;
    sb.tol = tol;
    sb.divtol = divtol;
    sb.itmax = itmax;
    return sb;
  }
};

// ---------------------------------------------------------------------------
// Batched BiCGSTAB: nm systems in the batch layout of vgpu_bbicg.hpp (system
// m at nodes [off_m, off_m + n_m), vector workgroups [vb0_m, vb0_m + nb_m)
// with nb_m = min(256, ceil(n_m / 4096)), the count a solve of the system
// alone takes, owner partials from u0_m).  Every iteration is one kStabDir
// and one kStabS owner launch over all systems; each system reads and writes
// only its own partial ranges and status, so its iterates are bitwise those
// of Stab on the system alone with the same tile shape and z chunk.  A
// solve takes a run flag per system: systems not run are marked done by the
// init kernel (iterations 0) and their marches return at once.
// ---------------------------------------------------------------------------

// ks_init over the members of a batch (per-member workgroups as above).
template <typename TV>
__global__ __launch_bounds__(SVB) void ksb_init(SolveStatus *__restrict__ stv,
                                               const Segment *__restrict__ seg, VGPU_ACT_ARGS,
                                               const int *__restrict__ run,
                                               const uint8_t *__restrict__ mask,
                                               const TV *__restrict__ b,
                                               const TV *__restrict__ x, TV *__restrict__ xw,
                                               const TV *__restrict__ ax,
                                               const TV *__restrict__ diag,
                                               TV *__restrict__ R, TV *__restrict__ RT,
                                               double *__restrict__ prt,
                                               double *__restrict__ prr,
                                               double *__restrict__ pbn) {
  __shared__ double red[SVB / 64];
  const BlockOf bo = block_of(act, actvb, nact, blockIdx.x);
  const Segment sg = seg[bo.m];
  const int j = bo.j;
  if (j == 0 && threadIdx.x == 0) {
    SolveStatus s0{};
    s0.rho = s0.alpha = s0.omega = 1.0;
    if (!run[bo.m]) s0.done = 1;
    stv[bo.m] = s0;
  }
  double srt = 0, sr = 0, sb = 0;
  for (int64_t li = j * (int64_t)SVB + threadIdx.x; li < sg.n; li += (int64_t)sg.nb * SVB) {
    const int64_t i = sg.off + li;
    xw[i] = mask[i] ? TV(0) : x[i];
    const TV bi = b[i], di = diag[i];
    const TV ri = bi - ax[i];
    const TV Ri = ri / di, RTi = ri * di;
    R[i] = Ri;
    RT[i] = RTi;
    srt += (double)RTi * (double)Ri;
    sr += (double)ri * (double)ri;
    sb += (double)bi * (double)bi;
  }
  stab_partial2(srt, sr, prt, prr, sg.vb0 + j);
  for (int off = 32; off > 0; off >>= 1) sb += __shfl_down(sb, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sb;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < SVB / 64; ++w) t += red[w];
    pbn[sg.vb0 + j] = t;
  }
}

// Batched operator callbacks: y = A x over every member (ignores statuses),
// and one BiCGSTAB owner launch over every member.
template <typename TV>
using BStabSingleOp = std::function<int(const TV *x, TV *y, hipStream_t s)>;
template <typename TV>
using BStabOwnerOp = std::function<int(int mode, const StabData<TV> &sd, const TV *r, TV *y,
                                       hipStream_t s)>;

template <typename TV>
struct BStab {
  int nm = 0;
  int64_t ntot = 0;
  int NB = 0;   // vector workgroups of all members
  int cap = 0;  // owner partials of all members (units; groups never exceed them)
  std::vector<Segment> hseg;
  Segment *seg = nullptr;
  ActiveBuf all;  // every member, for the vector launches
  TV *ax = nullptr, *RT = nullptr, *T = nullptr, *xw = nullptr;
  TV *R[2] = {nullptr, nullptr}, *P[2] = {nullptr, nullptr}, *V[2] = {nullptr, nullptr};
  VecBlock vblk;  // one allocation for the ten vectors above
  double *prt = nullptr, *prr = nullptr, *pbn = nullptr;
  double *pga = nullptr, *prn = nullptr, *pts = nullptr, *ptt = nullptr, *prs = nullptr,
         *prq = nullptr;
  SolveStatus *st = nullptr, *st_host = nullptr;
  int *run_dev = nullptr, *run_host = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  static constexpr int kPairs = 4;

  int alloc(const std::vector<Segment> &sg) {
    if (seg && sg.size() == hseg.size()) {
      bool same = true;
      for (size_t i = 0; i < sg.size() && same; ++i)
        same = sg[i].off == hseg[i].off && sg[i].n == hseg[i].n && sg[i].u0 == hseg[i].u0 &&
               sg[i].nu == hseg[i].nu && sg[i].vb0 == hseg[i].vb0 && sg[i].nb == hseg[i].nb &&
               sg[i].ng == hseg[i].ng;
      if (same) return VGPU_OK;
    }
    // a failed allocation releases everything: the next call starts again
    const int rc = alloc_parts(sg);
    if (rc) release();
    return rc;
  }
  int alloc_parts(const std::vector<Segment> &sg) {
    release();
    hseg = sg;
    nm = (int)sg.size();
    ntot = sg.back().off + seg_pad(sg.back().n);
    NB = sg.back().vb0 + sg.back().nb;
    cap = (int)std::max<int64_t>(1, sg.back().u0 + sg.back().nu);
    VGPU_HIP(hipMalloc(&seg, nm * sizeof(Segment)));
    VGPU_HIP(hipMemcpy(seg, sg.data(), nm * sizeof(Segment), hipMemcpyHostToDevice));
    int rc = all.alloc(nm);
    if (rc) return rc;
    std::vector<int> members(nm);
    for (int i = 0; i < nm; ++i) members[i] = i;
    rc = all.set(sg, members);
    if (rc) return rc;
    TV **v[] = {&ax, &RT, &T, &xw, &R[0], &R[1], &P[0], &P[1], &V[0], &V[1]};
    if (vblk.alloc<TV>(v, 10, ntot) != VGPU_OK) {
      set_error("batched BiCGSTAB: vector allocation failed");
      return VGPU_ERR_NOMEM;
    }
    double **pn[] = {&prt, &prr, &pbn};
    for (auto pd : pn) VGPU_HIP(hipMalloc(pd, NB * sizeof(double)));
    double **pa[] = {&pga, &prn, &pts, &ptt, &prs, &prq};
    for (auto pd : pa) {
      VGPU_HIP(hipMalloc(pd, cap * sizeof(double)));
      VGPU_HIP(hipMemset(*pd, 0, cap * sizeof(double)));
    }
    VGPU_HIP(hipMalloc(&st, nm * sizeof(SolveStatus)));
    VGPU_HIP(hipHostMalloc(&st_host, nm * sizeof(SolveStatus), hipHostMallocDefault));
    VGPU_HIP(hipMalloc(&run_dev, nm * sizeof(int)));
    VGPU_HIP(hipHostMalloc(&run_host, nm * sizeof(int), hipHostMallocDefault));
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
    TV **v[] = {&ax, &RT, &T, &xw, &R[0], &R[1], &P[0], &P[1], &V[0], &V[1]};
    for (auto pv : v) *pv = nullptr;
    vblk.release();
    double **d[] = {&prt, &prr, &pbn, &pga, &prn, &pts, &ptt, &prs, &prq};
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
    if (run_dev) hipFree(run_dev);
    run_dev = nullptr;
    if (run_host) hipHostFree(run_host);
    run_host = nullptr;
    all.release();
    hseg.clear();
    nm = 0;
  }
  ~BStab() { release(); }

  // Solve A_m x_m = b_m for the members with run[m] != 0 (device vectors in
  // the batch layout; x in/out, zeroed on the mask); the other members' x
  // are left as they are.  iters / errs / status per member (status 1
  // converged, 2 itmax, 3 breakdown, 4 diverged; members not run: 1 with 0
  // iterations).
  int solve(uint64_t op_key, const BStabSingleOp<TV> &single, const BStabOwnerOp<TV> &owner,
            const uint8_t *mask, const TV *b, TV *x, const TV *diag, double tol, int itmax,
            double divtol, const int *run, hipStream_t s, int *iters, double *errs,
            int *status) {
    for (int m = 0; m < nm; ++m) run_host[m] = run[m] ? 1 : 0;
    VGPU_HIP(hipMemcpyAsync(run_dev, run_host, nm * sizeof(int), hipMemcpyHostToDevice, s));
    int rc = single(x, ax, s);  // A x0 (x0 masked on load)
    if (rc) return rc;
    const ActiveSet &as = all.view;
    hipLaunchKernelGGL((ksb_init<TV>), dim3(NB), dim3(SVB), 0, s, st, seg, as.act, as.actvb,
                       as.nact, run_dev, mask, b, x, xw, ax, diag, R[0], RT, prt, prr, pbn);
    VGPU_HIP(hipGetLastError());
    const bool same = key_valid && key_op == op_key && key_diag == diag && key_tol == tol &&
                      key_itmax == itmax && key_div == divtol;
    if (!same) {
      drop_graph();
      VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int crc = VGPU_OK;
      for (int pi = 0; pi < kPairs && crc == VGPU_OK; ++pi)
        for (int q = 1; q >= 0 && crc == VGPU_OK; --q) {
          crc = owner(kStabDir, dir_data(q, diag, tol, itmax, divtol), R[q ^ 1], V[q], s);
          if (crc == VGPU_OK)
            crc = owner(kStabS, s_data(q, diag, tol, itmax, divtol), R[q], T, s);
        }
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
    }
    // replay until every member is done, at most two replays in flight
    const int max_replays = itmax / (2 * kPairs) + 2;
    hipEvent_t ev[2];
    VGPU_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    VGPU_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int result = VGPU_OK, launched = 0, completed = 0;
    for (;;) {
      while (launched - completed < 2 && launched < max_replays) {
        if (hipGraphLaunch(gexec, s) != hipSuccess ||
            hipMemcpyAsync(st_host, st, nm * sizeof(SolveStatus), hipMemcpyDeviceToHost,
                           s) != hipSuccess ||
            hipEventRecord(ev[launched & 1], s) != hipSuccess) {
          set_error("batched BiCGSTAB graph replay failed");
          result = VGPU_ERR_HIP;
          break;
        }
        ++launched;
      }
      if (result || completed == launched) break;
      if (hipEventSynchronize(ev[completed & 1]) != hipSuccess) {
        set_error("batched BiCGSTAB event wait failed");
        result = VGPU_ERR_HIP;
        break;
      }
      ++completed;
      bool all_done = true;
      for (int m = 0; m < nm && all_done; ++m) all_done = st_host[m].done != 0;
      if (all_done) break;
    }
    if (!result) {  // the iterates to the caller (members not run: x unchanged
                    // up to the mask, which their x already has)
      hipLaunchKernelGGL((ks_copy<TV>), dim3(NB), dim3(SVB), 0, s, ntot, xw, x);
      if (hipGetLastError() != hipSuccess) {
        set_error("batched BiCGSTAB result copy failed");
        result = VGPU_ERR_HIP;
      }
    }
    if (hipStreamSynchronize(s) != hipSuccess && !result) {
      set_error("batched BiCGSTAB stream sync failed");
      result = VGPU_ERR_HIP;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (result) return result;
    for (int m = 0; m < nm; ++m) {
      iters[m] = st_host[m].iter;
      errs[m] = st_host[m].err;
      status[m] = st_host[m].done;
    }
    return VGPU_OK;
  }

 private:
  bool key_valid = false;
  uint64_t key_op = 0;
  const void *key_diag = nullptr;
  double key_tol = 0, key_div = 0;
  int key_itmax = 0;

  StabData<TV> dir_data(int q, const TV *diag, double tol, int itmax, double divtol) const {
    StabData<TV> sa;
    sa.po = P[q ^ 1];
    sa.vo = V[q ^ 1];
    sa.to = T;
    sa.pn = P[q];
    sa.rn = R[q];
    sa.x = xw;
    sa.d = diag;
    sa.rt = RT;
    double *const ain[4] = {pts, ptt, prs, prq};
    for (int i = 0; i < 4; ++i) sa.pin[i] = ain[i];
    sa.prt0 = prt;
    sa.pout[0] = pga;
    sa.pout[1] = prn;
    sa.st = st;
    sa.blk = vblk.base;
    sa.blk_bytes = vblk.bytes;
    sa.tol = tol;
    sa.divtol = divtol;
    sa.itmax = itmax;
    return sa;
  }
  StabData<TV> s_data(int q, const TV *diag, double tol, int itmax, double divtol) const {
    StabData<TV> sb;
    sb.v = V[q];
    sb.d = diag;
    sb.rt = RT;
    sb.pin[0] = pga;
    sb.pin[1] = prn;
    sb.pin[2] = pbn;
    double *const bout[4] = {pts, ptt, prs, prq};
    for (int i = 0; i < 4; ++i) sb.pout[i] = bout[i];
    sb.st = st;
    sb.blk = vblk.base;
    sb.blk_bytes = vblk.bytes;
    sb.tol = tol;
    sb.divtol = divtol;
    sb.itmax = itmax;
    return sb;
  }
};

}  // namespace vgpu
