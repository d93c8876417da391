// Internal declarations of libvanadis_gpu.so (gfx950 / CDNA4): operator
// data, tile geometry, barriers and launchers.  The element math lives in
// vgpu_elem.hpp.
#pragma once
#include <hip/hip_runtime.h>

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
  double err; // ||r||/||b|| of the last test
  double bnrm2;
  double pad;
};

// One system of a batch: its node range in the concatenated batch vectors,
// its vector workgroups and its owner work units (tile-planes).  A single
// solve is a batch of one.
struct Segment {
  int64_t off, n;  // nodes [off, off + n) of the batch vectors
  int64_t u0, nu;  // owner units [u0, u0 + nu) (dot partial slots)
  int vb0, nb;     // vector workgroups [vb0, vb0 + nb) (= partial slots)
};

// Operator data of one batch member for the owner kernel (device array).
template <typename T>
struct Member {
  OpData<T> op;
  Segment seg;
};

// ---------------------------------------------------------------------------
// Host-side launchers (vgpu_operator.hip)
// ---------------------------------------------------------------------------
// Owner work units (tile-planes) of one system for a tile shape.
int64_t owner_units(const Grid3 &g, int tile);
// Owner operator over a batch of nm systems with utot units in total: for
// every member m whose status is not done, y1 = A_m x1 (and y2 = A_m^T x2) on
// the member's node range of the batch vectors.  dot (dual only, optional):
// per-unit partials of y1 . x2 at dot[u0 + unit]; they depend on the tile
// shape only, not on the grid.  zchunk: as owner_num_blocks.
template <typename T, typename TIO>
int launch_owner_batch(const Member<T> *md, int nm, int64_t utot, int tile,
                       int zchunk, bool robin, const SolveStatus *st,
                       const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                       double *dot, hipStream_t s);
// Apply y1 = A x1 (and, if x2, y2 = A^T x2) with masking; TIO is the vector
// type (double or float), T the element arithmetic type.  dot_part (optional,
// owner mode, dual only): owner_units(g, op.tile) partials of y1 . x2, one per
// owner unit.  op.done, if set, points at a SolveStatus (owner mode) or its
// done flag (brick modes).  zchunk: as owner_num_blocks / brick z chunk.
template <typename T, typename TIO>
int launch_apply(const OpData<T> &op, int scatter, int zchunk,
                 const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                 double *dot_part, hipStream_t s);

// Fills the tau field of the per-element records from the other fields.
int launch_tau(const Grid3 &g, const double *xc, const double *yc,
               const double *zc, double *rec, int variant, hipStream_t s);
int launch_diag(const OpData<double> &op, double *diag, hipStream_t s);
int launch_source(const OpData<double> &op, const double *q_in, double *f,
                  hipStream_t s);
int launch_mass(const Grid3 &g, const double *xc, const double *yc,
                const double *zc, const double *c, double *partials,
                int nparts, hipStream_t s);

}  // namespace vgpu
