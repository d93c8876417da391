// Internal declarations of libvanadis_gpu.so (gfx950 / CDNA4): operator
// data, tile geometry, barriers and launchers.  The element math lives in
// vgpu_elem.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/vanadis_gpu.h"

namespace vgpu {

void set_error(const std::string &msg);

#define VGPU_HIP(call)                                                       \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::vgpu::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + \
                        " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
      return VGPU_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

// ---------------------------------------------------------------------------
// Operator description shared by every kernel.
// ---------------------------------------------------------------------------
struct Grid3 {
  int nax, nay, naz;  // nodes per axis
  int ex, ey, ez;     // elements per axis
  int64_t N, E;
};

template <typename T>
struct OpData {
  Grid3 g;
  const double *xc, *yc, *zc;  // axis coordinates (device, FP64)
  const T *hx, *hy, *hz;       // element sizes along each axis [ex], [ey], [ez]
  const T *ihx, *ihy, *ihz;    // their reciprocals
  const uint8_t *mask;         // 1 = Dirichlet node
  const T *rec;  // per-element records [5E]: vx, vy, vz, lambda, tau
  double cH;     // weight of H (diffusion + advection + DR + Robin)
  double mu;     // weight of C = M_w: cH*P + cC
  double ra;     // Robin weight cH*alpha on x-min elements (0 = none)
  const int *done;  // optional device flag: nonzero -> kernels return at once
  int tile;         // owner tile shape: 0 = 32x8, 1 = 16x16, 2 = 64x4
};

// Workgroup barrier that waits only for LDS (and scalar) traffic: global
// loads issued before it stay in flight, which is what lets a z-marching
// kernel prefetch the next layer while it computes the current one.  The
// compiler still inserts the vmcnt wait before any use of a prefetched value.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

}  // namespace vgpu

#include "vgpu_elem.hpp"

namespace vgpu {

// Workgroups are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8),
// each XCD with its own L2.  Map block b to a logical tile so that every XCD
// works on a contiguous range of tiles: neighbouring tiles, which share halo
// node planes and element columns, then meet in the same L2.  Bijective on
// [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, slot = b >> 3;
  return x * q + min(x, r) + slot;
}

// Status of one BiCG system on the device; its kernels return at once when
// done != 0.
struct SolveStatus {
  int done;   // 0 running, 1 converged, 2 itmax, 3 breakdown, 4 diverged
  int iter;   // completed iterations
  int pend;   // batched engine: done code found by an iteration launch, made
              // `done` by the update launch that follows (so no workgroup of
              // the iteration launch sees it change)
  int pad;    // BiCGSTAB: index of the iteration whose direction launch ran
  double err; // ||r||/||b|| of the last test
  double bnrm2;
  double alpha;  // alpha of the last update (batched engine: x += alpha p
                 // is applied by the next iteration launch)
  double rho;       // BiCGSTAB: rho of the last S launch's iteration (r~ . r)
  double omega;     // BiCGSTAB: omega of the last completed iteration
  double rho_next;  // BiCGSTAB: rho of the iteration whose direction launch ran
};

// One system of a batch: its node range in the concatenated batch vectors,
// its vector workgroups and its owner work units (tile-planes).  A single
// solve is a batch of one.
struct Segment {
  int64_t off, n;  // nodes [off, off + n) of the batch vectors
  int64_t u0, nu;  // owner units [u0, u0 + nu) (dot partial slots)
  int vb0, nb;     // vector workgroups [vb0, vb0 + nb) (= partial slots)
  int64_t ng;      // owner groups of the system (z-chunk launches), else 0
};

// Operator data of one batch member for the owner kernel (device array).
template <typename T>
struct Member {
  OpData<T> op;
  Segment seg;
};

// The systems of a batch that launches cover: member indices act[0..nact)
// (increasing) with the prefix sums of their vector workgroups actvb[0..nact]
// and owner units actu[0..nact] (device arrays).  Partial slots stay at each
// member's own vb0 / u0, so dropping finished systems from the launches does
// not move anything the others read.
struct ActiveSet {
  const int *act = nullptr;
  const int *actvb = nullptr;
  const int64_t *actu = nullptr;
  const int64_t *actg = nullptr;  // prefix sums of the members' owner groups
  int nact = 0;
  int nvb = 0;     // actvb[nact]
  int64_t nu = 0;  // actu[nact]
  int64_t ng = 0;  // actg[nact]
};

// Sum of v over the 64 lanes of a wave, in a fixed order, valid in lane 63:
// DPP adds across lane pairs (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1),
// half rows (row_half_mirror) and rows (row_mirror), then row r += lane 15 of
// row r-1 (row_bcast:15 on rows 1, 3) and rows 2, 3 += lane 31
// (row_bcast:31).  VALU moves only: no LDS round trips on the critical path.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = __double_as_longlong(v);
  // lanes of rows a broadcast step leaves out keep an unspecified value (the
  // destination is not initialised); lane 63 never reads them
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ double wave_sum_lane63(double v) {
  v += dpp_f64<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xf>(v);  // row_mirror
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 -> rows 2, 3
  return v;
}

// ---------------------------------------------------------------------------
// Host-side launchers (vgpu_operator.hip)
// ---------------------------------------------------------------------------
// Node planes per owner work unit: a unit is this many consecutive planes of
// one tile, the granularity of the workgroup split and of the dot partials.
constexpr int kUnitPlanes = 1;
// Owner work units of one system for a tile shape.
int64_t owner_units(const Grid3 &g, int tile);
// Owner groups of one system for a tile shape and z chunk (> 0): the
// (z chunk, tile) pairs a z-chunk launch gives one workgroup each; 0 if
// zchunk <= 0.
int64_t owner_groups(const Grid3 &g, int tile, int zchunk);
// Whether the owner kernel can take a system: it addresses each vector
// (8 N bytes in FP64) and the element records (40 E bytes) with 32-bit offsets.
bool owner_fits(const Grid3 &g);
// BiCG direction update fused into the dual owner launch of an iteration
// (batched engine, owner scatter): per system, the convergence test of the
// previous iteration and beta from the reduction partials, then
// p = z + beta po and pp = zz + beta ppo formed as the nodes are loaded
// (iteration 1: p = z, pp = zz); the new directions of the nodes a tile owns
// go to pn / ppn, the operator results to the launch's outputs, and the
// previous update's x += alpha po (alpha from the status) to x — also when
// the test ends the solve, so x is complete when the system is done.
template <typename TIO>
struct DirData {
  const TIO *po = nullptr, *ppo = nullptr;  // previous directions
  TIO *pn = nullptr, *ppn = nullptr;        // new directions
  TIO *x = nullptr;  // solution: x += alpha po of the previous update (owned nodes)
  const double *pr = nullptr, *pb = nullptr, *pbn = nullptr;  // r.r, z.rr, b.b partials
  SolveStatus *stv = nullptr;
  double tol = 0, divtol = 0;
  int itmax = 0, NB = 0, q = 0;
  // the engine's vector block (VecBlock): when every vector of the launch
  // lies in it, the launch addresses them through one buffer resource
  const void *blk = nullptr;
  int64_t blk_bytes = 0;
};
// y1 = A p, y2 = A^T pp with the fused direction update (inputs z, zz);
// batch: the active members of md (else the single member m0); dot partials
// of y1 . pp per owner unit as launch_owner_batch.
template <typename T, typename TIO>
int launch_owner_dir(const Member<T> &m0, const Member<T> *md, const ActiveSet &as,
                     bool batch, int tile, int zchunk, bool robin, const DirData<TIO> &dd,
                     const TIO *z, const TIO *zz, TIO *y1, TIO *y2, double *dot,
                     hipStream_t s);
// Owner launches of one Jacobi-BiCGSTAB iteration j (one system, owner
// scatter), written for the scaled vectors R = D^-1 r, P = D^-1 p,
// V = D^-1 A P, S = D^-1 s, T = D^-1 t (D the diagonal, right
// preconditioning) and the shadow RT = D r0 (so that r0 . r = RT . R):
//  * kStabDir: omega_{j-1} = t.s / t.t and rho_j = r~.s - omega r~.t from the
//    previous S launch's partials, beta_j = (rho_j / rho_{j-1})
//    (alpha_{j-1} / omega_{j-1}); the plane loads (R, V, P of iteration j-1
//    and T) form S = R - alpha V, R_j = S - omega T and
//    P_j = R_j + beta (P - omega V) (j = 0: R_j = R, P_j = R); the nodes the
//    tile owns store R_j, P_j and x += alpha P + omega S; output
//    V_j = (A P_j)/d; partials of RT . V_j and |r_j|^2 = sum (d R_j)^2;
//  * kStabS: the convergence test of r_j (|r_j|/|b| from those partials),
//    alpha_j = rho_j / (RT . V_j); the plane loads form S = R_j - alpha V_j;
//    output T = (A S)/d; partials of t.s, t.t, r~.s and r~.t (t = d T,
//    s = d S).
// Two launches per iteration, no vector pass in between.
enum StabMode { kStabDir = 2, kStabS = 3 };
// (4, 5: the same with the early status test; 6: a damped Jacobi sweep;
// 7: the operator over the z-slabs of one colour of the colored scatter)
constexpr int kJacobi = 6;
constexpr int kSlab = 7;
// added to a BiCG / BiCGSTAB mode (1..5): the same launch with every vector
// but the diagonal addressed through the engine's block (DirData::blk)
constexpr int kBlk = 10;
template <typename TIO>
struct StabData {
  const TIO *po = nullptr, *vo = nullptr;  // kStabDir: P, V of iteration j-1
  const TIO *to = nullptr;                 // kStabDir: T of iteration j-1
  TIO *pn = nullptr, *rn = nullptr;        // kStabDir: P_j, R_j (owned nodes)
  TIO *x = nullptr;                        // kStabDir: solution (owned nodes)
  const TIO *v = nullptr;                  // kStabS: V_j
  const TIO *d = nullptr;                  // diagonal (owned nodes)
  const TIO *rt = nullptr;                 // shadow RT (owned nodes)
  // partial arrays the launch's prologue sums (in a fixed order) and their
  // counts: kStabDir t.s, t.t, r~.s, r~.t; kStabS RT.V, |r|^2, |b|^2
  const double *pin[4] = {nullptr, nullptr, nullptr, nullptr};
  int nin[4] = {0, 0, 0, 0};
  const double *prt0 = nullptr;  // kStabDir, iteration 0: r0 . r0 partials
  int nb0 = 0;
  // partials the launch writes, one per unit (zchunk 0) or (chunk, tile)
  // group (zchunk > 0): kStabDir RT.V, |r|^2; kStabS t.s, t.t, r~.s, r~.t
  double *pout[4] = {nullptr, nullptr, nullptr, nullptr};
  SolveStatus *st = nullptr;
  double tol = 0, divtol = 0;
  int itmax = 0;
  // 1: every workgroup reads the status first and returns at once when the
  // solve has ended (cheap launches past the end of a short solve, one load
  // latency more for the others; a kernel variant of its own)
  int early = 0;
  double jomega = 0;  // Jacobi sweep: damping (d, rt hold the diagonal and b)
  // the engine's vector block, as in DirData (d is addressed on its own)
  const void *blk = nullptr;
  int64_t blk_bytes = 0;
};

// The vectors of a solver engine carved from one allocation (each padded to
// a multiple of 128 values), so that the engine's owner launches can address
// all of them through one buffer resource and a scalar offset each: fewer
// scalar registers than a resource per vector (4 each).  The owner launches
// need the block below 2^31 bytes; a larger block still works, through a
// resource per vector.
struct VecBlock {
  void *base = nullptr;
  int64_t bytes = 0;
  // k vectors of n values of type TV into *v[0..k), zeroed
  template <typename TV>
  int alloc(TV **const *v, int k, int64_t n) {
    release();
    const int64_t np = (std::max<int64_t>(1, n) + 127) / 128 * 128;
    bytes = k * np * (int64_t)sizeof(TV);
    if (hipMalloc(&base, bytes) != hipSuccess) {
      base = nullptr;
      bytes = 0;
      return VGPU_ERR_NOMEM;
    }
    if (hipMemset(base, 0, bytes) != hipSuccess) return VGPU_ERR_HIP;
    for (int i = 0; i < k; ++i) *v[i] = static_cast<TV *>(base) + i * np;
    return VGPU_OK;
  }
  void release() {
    if (base) hipFree(base);
    base = nullptr;
    bytes = 0;
  }
};
// The same two launches over the members of a batch (md, as as
// launch_owner_batch): sd.st points at the members' statuses (indexed by
// member), sd.pin / sd.pout / sd.prt0 at the batch-wide partial arrays, of
// which member m reads and writes its own ranges (owner partials from
// seg.u0, vector-workgroup partials from seg.vb0); sd.nin / sd.nb0 are not
// used.  A member whose status is done at the launch skips its marches.
template <typename T, typename TIO>
int launch_owner_stab_batch(const Member<T> *md, const ActiveSet &as, int tile, int zchunk,
                            bool robin, int mode, const StabData<TIO> &sd, const TIO *r,
                            TIO *y1, hipStream_t s);
// One damped Jacobi sweep on the owner kernel: xn = x + omega (b - A x) / d
// (0 on the mask), the update fused into the operator's output stage; xn must
// not alias x (other tiles read x at their halo).
template <typename T, typename TIO>
int launch_owner_jacobi(const OpData<T> &op, int zchunk, double omega, const TIO *x,
                        const TIO *b, const TIO *d, TIO *xn, hipStream_t s);
// One BiCGSTAB owner launch (mode kStabDir or kStabS) on R (x1) into y1.
template <typename T, typename TIO>
int launch_owner_stab(const OpData<T> &op, int mode, int zchunk, const StabData<TIO> &sd,
                      const TIO *r, TIO *y1, hipStream_t s);
// Owner operator over the active systems of a batch (units split evenly over
// as.nu): for every active member m whose status is not done, y1 = A_m x1 (and
// y2 = A_m^T x2) on the member's node range of the batch vectors.  dot (dual
// only, optional): partials of y1 . x2 at dot[u0 + i], one per unit of a
// balanced launch (zchunk 0; they depend on the tile shape only) or one per
// (chunk, tile) group of a z-chunk launch (zchunk > 0).
template <typename T, typename TIO>
int launch_owner_batch(const Member<T> *md, const ActiveSet &as, int tile,
                       int zchunk, bool robin, const SolveStatus *st,
                       const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                       double *dot, hipStream_t s);
// Apply y1 = A x1 (and, if x2, y2 = A^T x2) with masking; TIO is the vector
// type (double or float), T the element arithmetic type.  dot_part (optional,
// owner mode, dual only): partials of y1 . x2 as launch_owner_batch (at most
// owner_units(g, op.tile) of them).  op.done, if set, points at a SolveStatus
// (owner mode) or its done flag (brick modes).  zchunk: 0 balanced launch,
// > 0 z-chunk launch / brick z chunk.
template <typename T, typename TIO>
int launch_apply(const OpData<T> &op, int scatter, int zchunk,
                 const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                 double *dot_part, hipStream_t s);

// Fills the tau field of the per-element records from the other fields.
int launch_tau(const Grid3 &g, const double *xc, const double *yc,
               const double *zc, double *rec, int variant, hipStream_t s);
int launch_diag(const OpData<double> &op, double *diag, hipStream_t s);
// Per-block largest off-diagonal / diagonal ratio of the element-matrix rows
// and the count of rows that are not diagonally dominant (nblocks x 256).
int launch_ddf(const OpData<double> &op, double *pmax, int *pbad, int nblocks,
               hipStream_t s);
int launch_source(const OpData<double> &op, const double *q_in, double *f,
                  hipStream_t s);
int launch_mass(const Grid3 &g, const double *xc, const double *yc,
                const double *zc, const double *c, double *partials,
                int nparts, hipStream_t s);

}  // namespace vgpu
