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

// ---------------------------------------------------------------------------
// Host-side launchers (vgpu_operator.hip)
// ---------------------------------------------------------------------------
// Apply y1 = A x1 (and, if x2, y2 = A^T x2) with masking; TIO is the vector
// type (double or float), T the element arithmetic type.  dot_part (optional,
// owner mode only): per-workgroup partials of y1 . x2.  Returns # partials.
template <typename T, typename TIO>
int launch_apply(const OpData<T> &op, int scatter, int zchunk,
                 const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                 double *dot_part, hipStream_t s);
// number of dot partials the owner launch writes for this geometry
// Workgroups of an owner launch (= dot partials it writes): zchunk > 0 gives
// about zchunk node planes per workgroup, zchunk = 0 one workgroup per
// resident slot of the device.  elem_bytes / io_bytes: sizeof(T) / sizeof(TIO).
int owner_num_blocks(const Grid3 &g, int zchunk, int tile, int elem_bytes,
                     int io_bytes, bool dual, bool robin);
// Upper bound of owner_num_blocks over every option (partial buffer size).
int owner_max_blocks(const Grid3 &g);

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
