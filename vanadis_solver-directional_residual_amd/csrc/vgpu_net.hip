// Learned initial guess for implicit time stepping: a 5-layer 3x3x3
// convolution network on the node grid (16 channels, "same" zero padding,
// leaky ReLU (slope 0.1) between layers), BF16 operands on the MFMA units with FP32
// accumulation, trained on the device with hand-written backward kernels and
// Adam.  No counterpart in the reference (its host restarts BiCG from the
// previous step's solution, VANADIS.FOR:809-810 with TNW2 in/out); the guess
// only changes x0, BiCG still runs to the same rtol.
//
// Inputs (channels 0..5, 6..15 zero): x_{n-1}, the increment x_{n-1} - x_{n-2}
// (differenced in FP64: separately rounded states would lose it), the nodal
// source, vx, vy, vz, each times a scale; output channel 0 of layer 5, times
// out_scale, is the predicted correction of a base guess:
// guess = x_{n-1} + extrap (x_{n-1} - x_{n-2}) + out_scale * net.
//
// Tensors are node-major, 16 BF16 channels per node ([N][16], 32 bytes).
// The layer kernels march tiles of x rows up z through an LDS ring of node
// planes, one v_mfma_f32_16x16x32_bf16 per pair of taps; the weight-gradient
// kernel reduces over nodes with v_mfma_f32_16x16x16_bf16, whose operand
// layout (lane l, i = 0..3) is A[l%16][4(l/16)+i], B[4(l/16)+i][l%16],
// D[4(l/16)+i][l%16].
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "vgpu_ctx.hpp"

namespace vgpu {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NC = 16;            // channels
constexpr int NL = 5;             // layers
constexpr int NTAP = 27;          // 3x3x3 taps, t = dx + 3 dy + 9 dz, offset d-1
constexpr int WL = NTAP * NC * NC;  // weights per layer
constexpr int NBLK_W = 512;       // workgroups of the weight-gradient kernel (= partials)
constexpr float SLOPE = 0.1f;     // leaky ReLU slope for negative inputs

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u)
    return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ v4s load4(const uint16_t *p) {
  return *reinterpret_cast<const v4s *>(p);
}

struct RowSeg {
  int x0, y, z;
};
__device__ __forceinline__ RowSeg row_seg(const Grid3 &g, int64_t s) {
  const int nsx = (g.nax + 15) >> 4;
  const int64_t r = s / nsx;
  return {(int)(s % nsx) * 16, (int)(r % g.nay), (int)(r / g.nay)};
}

__device__ __forceinline__ bool inside(const Grid3 &g, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < g.nax && y < g.nay && z < g.naz;
}

// Layer kernels.  A workgroup (4 waves) owns a tile of CT_TY x rows of 16
// nodes, (x0 .. x0+15) x (y0 .. y0+CT_TY-1), over a chunk of zc node planes,
// and marches it up z.  The input planes z-1, z, z+1 around the output plane
// z, each with one halo node or row on every side of the tile, sit in LDS in
// a ring of four plane slots.  Each step the workgroup writes plane z+1,
// which it loaded two steps earlier, into the slot of plane z-3 and issues
// the global loads of plane z+3 into the registers just freed, so a load has
// two steps of MFMAs to arrive; one barrier per step orders the writes
// against the reads of the step before.  A wave computes two rows of the
// tile for all 16 output channels: per pair of taps one
// v_mfma_f32_16x16x32_bf16 with A = the two taps' weights (16 output x 2 x 16
// input channels, in registers for the whole march) and B = the 16 nodes'
// input channels at the two taps' offsets, 16 bytes per lane from LDS.  The
// 27 taps make 13 pairs and tap 26 alone (zero weights beside it); in 12 of
// the pairs both taps have the same dy, so row r at dy and row r + 1 at
// dy - 1 read the same operand and a wave's two rows need 20 operand reads
// instead of 28.  An
// output row's MFMAs run in the same order whatever the chunking, so the
// results do not depend on it.
//   FWD_HIDDEN: Y[n][co] = leaky(b[co] + sum_t sum_ci W[t][co][ci] X[n+off_t][ci])
//   FWD_LAST:   out[n] = b[0] + sum_t sum_ci W[t][0][ci] X[n+off_t][ci] (FP32)
//   BWD_DATA:   dZprev[n][ci] = leaky'(Yprev[n][ci]) sum_t sum_co WT[t][ci][co]
//               dZ[n-off_t][co]
// (off_t = (t%3-1, t/3%3-1, t/9-1)).  MFMA operand layout (16x16x32, lane l,
// i = 0..7): A[l%16][8(l/16)+i], B[8(l/16)+i][l%16], D[4(l/16)+i][l%16]; the
// reduction index k = 8(l/16)+i stands for input channel 8((l/16)%2)+i of
// tap ct_lo(j) (l < 32) or ct_hi(j) (l >= 32) in A and B alike.  In LDS a node takes its 32 bytes of
// channels, unpadded: the 16-byte reads of 16 consecutive nodes by each
// quarter wave run free of bank conflicts on gfx950 (its LDS conflict counter
// reads zero; a 48-byte node stride doubled the LDS busy cycles).  Global
// accesses are buffer operations on one node plane each (planes up to
// CT_OOB bytes).  Tiles are numbered x-fastest within a chunk
// and workgroups mapped XCD-aware, so the workgroups of one L2 march
// neighbouring tiles over the same planes.
enum ConvMode { FWD_HIDDEN = 0, FWD_LAST = 1, BWD_DATA = 2 };

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int CT_RW = 2;                     // x rows per wave
constexpr int CT_TY = 4 * CT_RW;             // x rows per tile
constexpr int CT_ROWS = CT_TY + 2;           // staged rows: y0 - 1 .. y0 + CT_TY
constexpr int CT_NS = 18;                    // staged nodes per row: x0 - 1 .. x0 + 16
constexpr int CT_NB = 2 * NC;                // LDS bytes per node
constexpr int CT_RB = CT_NS * CT_NB;         // LDS bytes per staged row
constexpr int CT_SB = CT_ROWS * CT_RB;       // LDS bytes per plane slot
constexpr int CT_NPC = CT_ROWS * CT_NS * 2;  // 16-byte pieces per staged plane
constexpr int CT_NP = (NTAP + 1) / 2;        // tap pairs
constexpr int CT_NQ = (CT_NPC + 255) / 256;  // pieces per thread and plane

__device__ __forceinline__ v8s load8(const uint16_t *p) {
  return *reinterpret_cast<const v8s *>(p);
}
// operand of tap t (forward X[n + off_t], backward dZ[n - off_t]) relative
// to the lower corner (-1, -1, -1) of the output node's 3x3x3 neighbourhood
template <int MODE>
__device__ __forceinline__ constexpr int ct_dz(int t) {
  return MODE == BWD_DATA ? 2 - t / 9 : t / 9;
}
// tap pairs of one MFMA, in MFMA order (lo taps on lanes 0..31, hi taps on
// 32..63; t = dx + 3 dy + 9 dz): j < 9 the z neighbours (dx, dy, 0) and
// (dx, dy, 1) with dx = j / 3, dy = j % 3; j = 9..11 the x neighbours (0, dy, 2)
// and (1, dy, 2) with dy = j - 9 -- in both, the two taps share dy, so row r at
// dy and row r + 1 at dy - 1 read the same operand; j = 12 (2, 0, 2) and
// (2, 1, 2); j = 13 (2, 2, 2) alone (-1: zero weights)
__host__ __device__ constexpr int ct_lo(int j) {
  return j < 9 ? j / 3 + 3 * (j % 3) : j < 12 ? 3 * (j - 9) + 18 : j == 12 ? 20 : 26;
}
__host__ __device__ constexpr int ct_hi(int j) {
  return j < 9 ? ct_lo(j) + 9 : j < 12 ? ct_lo(j) + 1 : j == 12 ? 23 : -1;
}
constexpr int CT_NG = 4;  // groups of three pairs that share their operands
template <int MODE>
__device__ __forceinline__ constexpr int ct_dy(int t) {
  return MODE == BWD_DATA ? 2 - (t / 3) % 3 : (t / 3) % 3;
}
template <int MODE>
__device__ __forceinline__ constexpr int ct_dx(int t) {
  return MODE == BWD_DATA ? 2 - t % 3 : t % 3;
}

typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
constexpr int CT_OOB = 0x40000000;  // buffer offset past every plane: loads 0, stores dropped
// buffer resource over node plane p of a [N][16] BF16 tensor (or an FP32
// node vector when bytes_per_node is 4); no records outside the grid, so
// every access to such a plane reads zeros or is dropped
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, const Grid3 &g,
                                                            int p, int bytes_per_node) {
  const int pb = g.nax * g.nay * bytes_per_node;
  const bool in = p >= 0 && p < g.naz;
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char *>(static_cast<const char *>(base)) + (in ? (int64_t)p * pb : 0), 0,
      in ? pb : 0, 0x00020000);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_conv_t(Grid3 g, int zc, const uint16_t *__restrict__ X,
                                                const uint16_t *__restrict__ W,
                                                const float *__restrict__ bias,
                                                const uint16_t *__restrict__ Yprev,
                                                uint16_t *__restrict__ Y,
                                                float *__restrict__ out) {
  __shared__ __align__(16) unsigned char lds[4 * CT_SB];
  const int tid = threadIdx.x, lane = tid & 63, wl = tid >> 6;
  const int ntx = (g.nax + 15) >> 4, nty = (g.nay + CT_TY - 1) / CT_TY;
  const int tpc = ntx * nty;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);  // the grid is exactly tpc x chunks
  const int chunk = tile / tpc, tr = tile % tpc;
  const int x0 = (tr % ntx) * 16, y0 = (tr / ntx) * CT_TY;
  const int z0 = chunk * zc, z1 = min(z0 + zc, g.naz);
  const int col = lane & 15, cg = lane >> 4, hi = cg >> 1;
  const v8s zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  // A: weights of tap ct_hi(j) (lanes 32..63) or ct_lo(j), input channels
  // 8 (cg % 2) .. + 7
  v8s a[CT_NP];
#pragma unroll
  for (int j = 0; j < CT_NP; ++j) {
    const int t = hi ? ct_hi(j) : ct_lo(j);
    a[j] = t >= 0 ? load8(W + (t * NC + col) * NC + 8 * (cg & 1)) : zero8;
  }
  float bs[4] = {0.f, 0.f, 0.f, 0.f};
  if (MODE == FWD_HIDDEN)
#pragma unroll
    for (int i = 0; i < 4; ++i) bs[i] = bias[4 * cg + i];
  else if (MODE == FWD_LAST)
    bs[0] = bias[0];
  // The two 16-byte pieces this thread stages per plane: piece c is channel
  // half c % 2 of staged node (c % 36) / 2 in staged row c / 36.  Global
  // accesses are buffer operations whose range check supplies the zero
  // padding (pieces outside the grid get an offset past the plane, planes
  // outside the grid or past the chunk no records) and drops the stores of
  // nodes outside the grid, so every access is issued and the waits before
  // the LDS writes count exactly the accesses issued since.
  int soff[CT_NQ], goff[CT_NQ];
  bool pin[CT_NQ];
#pragma unroll
  for (int q = 0; q < CT_NQ; ++q) {
    const int c = tid + 256 * q;
    const int r = c / (2 * CT_NS), s = (c % (2 * CT_NS)) >> 1, hh = c & 1;
    const int xx = x0 - 1 + s, yy = y0 - 1 + r;
    pin[q] = c < CT_NPC;
    goff[q] = pin[q] && xx >= 0 && xx < g.nax && yy >= 0 && yy < g.nay
                  ? (xx + g.nax * yy) * (2 * NC) + 16 * hh
                  : CT_OOB;
    soff[q] = r * CT_RB + s * CT_NB + 16 * hh;
  }
  // plane p's pieces (zeros outside the grid; planes past z1 are never
  // written, so they are not read either)
  auto ld = [&](v4u *R, int p) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(X, g, p <= z1 ? p : -1, 2 * NC);
#pragma unroll
    for (int q = 0; q < CT_NQ; ++q) R[q] = __builtin_amdgcn_raw_buffer_load_b128(rs, goff[q], 0, 0);
    __builtin_amdgcn_sched_barrier(0);  // issued here, ahead of the step's MFMAs
  };
  auto st = [&](const v4u *R, int p) {
#pragma unroll
    for (int q = 0; q < CT_NQ; ++q)
      if (pin[q]) *reinterpret_cast<v4u *>(lds + ((p - z0 + 1) & 3) * CT_SB + soff[q]) = R[q];
  };
  // the wave's rows y0 + wr, y0 + wr + 1: the lane's node offsets in a plane
  // (past the plane for nodes outside the grid) and its B operand position
  const int wr = CT_RW * wl, x = x0 + col;
  int noff[CT_RW];
#pragma unroll
  for (int rr = 0; rr < CT_RW; ++rr) {
    const int y = y0 + wr + rr;
    noff[rr] = x < g.nax && y < g.nay ? x + g.nax * y : -1;
  }
  int colx[3];  // byte offset of staged node col + d, channel half cg % 2
#pragma unroll
  for (int d = 0; d < 3; ++d) colx[d] = (col + d) * CT_NB + 16 * (cg & 1);
  const int lbase = wr * CT_RB;
  // BWD_DATA: Yprev of the wave's two rows, planes two steps apart
  v2u yp[2][CT_RW];
  auto ld_yp = [&](v2u *dst, int z) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(Yprev, g, z < z1 ? z : -1, 2 * NC);
#pragma unroll
    for (int rr = 0; rr < CT_RW; ++rr)
      dst[rr] = __builtin_amdgcn_raw_buffer_load_b64(
          rs, noff[rr] >= 0 ? noff[rr] * (2 * NC) + 8 * cg : CT_OOB, 0, 0);
  };
  auto plane = [&](int z, const v2u *ypz) {
    int sb[3];  // slots of planes z - 1, z, z + 1
#pragma unroll
    for (int d = 0; d < 3; ++d) sb[d] = ((z - z0 + d) & 3) * CT_SB;
    v4f acc[CT_RW];
#pragma unroll
    for (int rr = 0; rr < CT_RW; ++rr) acc[rr] = v4f{0.f, 0.f, 0.f, 0.f};
    auto mfma = [&](int j, int rr, const v8s &b) {
      acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a[j]),
                                                        __builtin_bit_cast(v8bf, b), acc[rr],
                                                        0, 0, 0);
    };
    // pairs 3g .. 3g + 2 (g < CT_NG) differ only in dy: the B operands of
    // the staged rows wr .. wr + 3 at each lane's plane and column serve all
    // three pairs for both rows
#pragma unroll
    for (int gx = 0; gx < CT_NG; ++gx) {
      const int t = hi ? ct_hi(3 * gx) : ct_lo(3 * gx);
      const int o = sb[ct_dz<MODE>(t)] + colx[ct_dx<MODE>(t)];
      v8s b[CT_RW + 2];
#pragma unroll
      for (int k = 0; k < CT_RW + 2; ++k)
        b[k] = *reinterpret_cast<const v8s *>(lds + lbase + o + k * CT_RB);
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int rr = 0; rr < CT_RW; ++rr)
          mfma(3 * gx + dy, rr, b[rr + ct_dy<MODE>(ct_lo(3 * gx + dy))]);
    }
    // pairs 12, 13: each lane at its own tap's row and column
#pragma unroll
    for (int j = 3 * CT_NG; j < CT_NP; ++j) {
      const int tl = ct_lo(j), th = ct_hi(j) >= 0 ? ct_hi(j) : tl;
      const int o = hi ? sb[ct_dz<MODE>(th)] + ct_dy<MODE>(th) * CT_RB + colx[ct_dx<MODE>(th)]
                       : sb[ct_dz<MODE>(tl)] + ct_dy<MODE>(tl) * CT_RB + colx[ct_dx<MODE>(tl)];
#pragma unroll
      for (int rr = 0; rr < CT_RW; ++rr)
        mfma(j, rr, *reinterpret_cast<const v8s *>(lds + lbase + o + rr * CT_RB));
    }
    if (MODE == FWD_LAST) {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(out, g, z, 4);
#pragma unroll
      for (int rr = 0; rr < CT_RW; ++rr)
        __builtin_amdgcn_raw_buffer_store_b32(
            __float_as_uint(acc[rr][0] + bs[0]), rs,
            noff[rr] >= 0 && cg == 0 ? noff[rr] * 4 : CT_OOB, 0, 0);
    } else {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(Y, g, z, 2 * NC);
#pragma unroll
      for (int rr = 0; rr < CT_RW; ++rr) {
        v4f v;
        if (MODE == FWD_HIDDEN) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float u = acc[rr][i] + bs[i];
            v[i] = fmaxf(u, SLOPE * u);  // leaky ReLU (slope < 1)
          }
        } else {
          const v4s yv = __builtin_bit_cast(v4s, ypz[rr]);
#pragma unroll
          for (int i = 0; i < 4; ++i)
            v[i] = bf2f((uint16_t)yv[i]) > 0.f ? acc[rr][i] : SLOPE * acc[rr][i];
        }
        // round to nearest even, two values per conversion
        v2u o;
        o[0] = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{v[0], v[1]}, v2bf));
        o[1] = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{v[2], v[3]}, v2bf));
        __builtin_amdgcn_raw_buffer_store_b64(o, rs,
                                              noff[rr] >= 0 ? noff[rr] * (2 * NC) + 8 * cg
                                                            : CT_OOB,
                                              0, 0);
      }
    }
  };
  // register sets: set s holds the planes p with (p - z0 + 1) % 2 == s
  v4u R0[CT_NQ], R1[CT_NQ];
  ld(R0, z0 - 1);
  ld(R1, z0);
  st(R0, z0 - 1);
  st(R1, z0);
  ld(R0, z0 + 1);
  ld(R1, z0 + 2);
  if (MODE == BWD_DATA) {
    ld_yp(yp[0], z0);
    ld_yp(yp[1], z0 + 1);
  }
  // Step z: plane z + 1 into its slot, barrier, loads of plane z + 3 (and
  // Yprev of plane z + 2), output plane z.  A trip runs two steps, starting
  // after the first step's barrier, so that the register sets are named
  // statically and every path into a wait before an LDS write has the same
  // accesses in flight: the waits count only the accesses issued since the
  // set was loaded.  An odd last step follows the loop.
  st(R0, z0 + 1);
  __syncthreads();
  int z = z0;
  for (; z + 1 < z1; z += 2) {
    ld(R0, z + 3);
    plane(z, yp[0]);
    if (MODE == BWD_DATA) ld_yp(yp[0], z + 2);
    st(R1, z + 2);
    __syncthreads();
    ld(R1, z + 4);
    plane(z + 1, yp[1]);
    if (MODE == BWD_DATA) ld_yp(yp[1], z + 3);
    st(R0, z + 3);  // past the chunk (no step follows): zeros into a slot no step reads
    __syncthreads();
  }
  if (z < z1) plane(z, yp[0]);
}

// z chunk of the layer kernels.  A workgroup's time grows with the planes it
// marches plus the two halo planes it loads first, and the workgroups run in
// rounds of the device's resident slots, so the chunk count minimises
// ceil(workgroups / slots) x (planes per chunk + 2).
inline int conv_zchunk(const Grid3 &g) {
  // resident workgroups of the layer kernel on the current device (cached per device)
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0, slots = 0;
  if (hipGetDevice(&dev) != hipSuccess) return g.naz;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) slots = it->second;
  }
  if (slots == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&k_conv_t<FWD_HIDDEN>), 256, 0) !=
            hipSuccess)
      return g.naz;
    slots = std::max(1, per_cu) * std::max(1, cus);
    std::lock_guard<std::mutex> lk(mu);
    cache[dev] = slots;
  }
  const int64_t tpc = (int64_t)((g.nax + 15) / 16) * ((g.nay + CT_TY - 1) / CT_TY);
  int best = g.naz;
  int64_t best_cost = -1;
  for (int nch = 1; nch <= g.naz; ++nch) {
    const int zc = (g.naz + nch - 1) / nch;
    const int64_t wgs = tpc * ((g.naz + zc - 1) / zc);
    const int64_t cost = ((wgs + slots - 1) / slots) * (zc + 2);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = zc;
    }
  }
  return best;
}

// tile of the weight-gradient kernel (below)
constexpr int WG_TY = 8;                      // x rows per tile (two per wave)
constexpr int WG_ROWS = WG_TY + 2;            // staged X rows: y0 - 1 .. y0 + 8
constexpr int WG_CB = 16 * 2;                 // LDS bytes per (row, channel): 16 nodes
constexpr int WG_CPY = NC * WG_CB;            // bytes per shifted copy of a staged row
constexpr int WG_RB = 3 * WG_CPY;             // bytes per staged X row (three copies)
constexpr int WG_SB = WG_ROWS * WG_RB;        // bytes per X plane slot
constexpr int WG_DB = WG_TY * WG_CPY;         // bytes per staged dZ plane
constexpr int WG_XP = WG_ROWS * CT_NS * 2;    // 16-byte X pieces per plane
constexpr int WG_XQ = (WG_XP + 255) / 256;    // X pieces per thread and plane
static_assert(WG_TY * 16 * 2 == 256, "one dZ piece per thread and plane");
// z chunk of the weight-gradient march: the (tile, chunk) items run in
// rounds of the NBLK_W workgroups, so the chunk count minimises
// ceil(items / NBLK_W) x (planes per chunk + 2)
inline int wgrad_zchunk(const Grid3 &g) {
  const int64_t tpc = (int64_t)((g.nax + 15) / 16) * ((g.nay + WG_TY - 1) / WG_TY);
  int best = g.naz;
  int64_t best_cost = -1;
  for (int nch = 1; nch <= g.naz; ++nch) {
    const int zc = (g.naz + nch - 1) / nch;
    const int64_t items = tpc * ((g.naz + zc - 1) / zc);
    const int64_t cost = ((items + NBLK_W - 1) / NBLK_W) * (zc + 2);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = zc;
    }
  }
  return best;
}
inline dim3 conv_grid(const Grid3 &g, int zc) {
  const int64_t tpc = (int64_t)((g.nax + 15) / 16) * ((g.nay + CT_TY - 1) / CT_TY);
  return dim3((unsigned)(tpc * ((g.naz + zc - 1) / zc)));
}

// Weight and bias gradients of a layer:
//   dW[t][co][ci] = sum_n dZ[n][co] X[n+off_t][ci],  db[co] = sum_n dZ[n][co]
// as v_mfma_f32_16x16x32_bf16 over 32 nodes (K) at a time, two x-rows of 16:
// A[co][k] = dZ[node k][co], B[k][ci] = X[node k + off_t][ci].  A lane needs
// one channel of 8 consecutive nodes, while a node's channels are contiguous
// in memory, so the planes go to LDS transposed: per staged row and channel
// the 16 nodes as 32 bytes, the X rows in three copies shifted by one node
// (dx = -1, 0, +1), so that every operand is one aligned 16-byte read.  A
// workgroup (4 waves, two rows each) marches a tile of 8 x-rows of 16 nodes
// up a chunk of zc planes like the layer kernels: X planes z-1, z, z+1 in an
// LDS ring of four slots, written one step after their loads were issued
// (two steps of work for a load to arrive), the dZ plane of the step in one
// of two slots, one barrier per step.  A persistent grid of NBLK_W
// workgroups takes the (tile, chunk) items blockIdx, blockIdx + NBLK_W, ...;
// its four waves add their accumulators in wave order through LDS and write
// one partial per workgroup: part[blk][t][co][ci], partb[blk][lane].
__global__ __launch_bounds__(256) void k_conv_wg(Grid3 g, int zc,
                                                 const uint16_t *__restrict__ X,
                                                 const uint16_t *__restrict__ dZ,
                                                 float *__restrict__ part,
                                                 float *__restrict__ partb) {
  __shared__ __align__(16) unsigned char lds[4 * WG_SB + 2 * WG_DB];
  static_assert(sizeof(lds) >= (NTAP * 4 * 64 + 64) * sizeof(float), "LDS sum area");
  const int tid = threadIdx.x, lane = tid & 63, wl = tid >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int ntx = (g.nax + 15) >> 4, nty = (g.nay + WG_TY - 1) / WG_TY;
  const int tpc = ntx * nty;
  const int nitem = tpc * ((g.naz + zc - 1) / zc);
  v4f acc[NTAP];
#pragma unroll
  for (int t = 0; t < NTAP; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  // the wave's operand positions: rows 2 wl + grp / 2, channel col, nodes
  // 8 (grp % 2) .. + 7
  const int orow = 2 * wl + (grp >> 1);
  const int obyte = col * WG_CB + 16 * (grp & 1);
  for (int it = xcd_tile(blockIdx.x, gridDim.x); it < nitem; it += gridDim.x) {
    const int chunk = it / tpc, tr = it % tpc;
    const int x0 = (tr % ntx) * 16, y0 = (tr / ntx) * WG_TY;
    const int z0 = chunk * zc, z1 = min(z0 + zc, g.naz);
    // X pieces: piece c is channel half c % 2 of staged node (c % 36) / 2
    // (node x0 - 1 + s) in staged row c / 36; dZ piece: channel half tid % 2
    // of node x0 + (tid % 32) / 2 in row y0 + tid / 32.  Offsets past the
    // plane for nodes outside the grid (loads return zeros).
    int goff[WG_XQ], xs[WG_XQ], xr[WG_XQ], xh[WG_XQ];
    bool pin[WG_XQ];
#pragma unroll
    for (int q = 0; q < WG_XQ; ++q) {
      const int c = tid + 256 * q;
      xr[q] = c / (2 * CT_NS);
      xs[q] = (c % (2 * CT_NS)) >> 1;
      xh[q] = c & 1;
      const int xx = x0 - 1 + xs[q], yy = y0 - 1 + xr[q];
      pin[q] = c < WG_XP;
      goff[q] = pin[q] && xx >= 0 && xx < g.nax && yy >= 0 && yy < g.nay
                    ? (xx + g.nax * yy) * (2 * NC) + 16 * xh[q]
                    : CT_OOB;
    }
    const int dr = tid >> 5, dsn = (tid & 31) >> 1, dh = tid & 1;
    const int doff = x0 + dsn < g.nax && y0 + dr < g.nay
                         ? ((x0 + dsn) + g.nax * (y0 + dr)) * (2 * NC) + 16 * dh
                         : CT_OOB;
    auto ldx = [&](v4u *R, int p) {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(X, g, p <= z1 ? p : -1, 2 * NC);
#pragma unroll
      for (int q = 0; q < WG_XQ; ++q)
        R[q] = __builtin_amdgcn_raw_buffer_load_b128(rs, goff[q], 0, 0);
    };
    auto ldd = [&](v4u &D, int p) {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(dZ, g, p < z1 ? p : -1, 2 * NC);
      D = __builtin_amdgcn_raw_buffer_load_b128(rs, doff, 0, 0);
      __builtin_amdgcn_sched_barrier(0);  // issued here, ahead of the step's MFMAs
    };
    // transposed writes: channel 8 h + c of staged node s to each copy d
    // that holds it (copy d, position j = s - d, 0 <= j < 16)
    auto stx = [&](const v4u *R, int p) {
      unsigned char *base = lds + ((p - z0 + 1) & 3) * WG_SB;
#pragma unroll
      for (int q = 0; q < WG_XQ; ++q) {
        if (!pin[q]) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int j = xs[q] - d;
          if (j < 0 || j > 15) continue;
          uint16_t *dst = reinterpret_cast<uint16_t *>(base + xr[q] * WG_RB + d * WG_CPY +
                                                       8 * xh[q] * WG_CB) + j;
#pragma unroll
          for (int c = 0; c < 8; ++c) dst[c * (WG_CB / 2)] = (uint16_t)(R[q][c >> 1] >> (16 * (c & 1)));
        }
      }
    };
    auto std_ = [&](const v4u &D, int p) {
      uint16_t *dst = reinterpret_cast<uint16_t *>(lds + 4 * WG_SB + ((p - z0) & 1) * WG_DB +
                                                   dr * WG_CPY + 8 * dh * WG_CB) + dsn;
#pragma unroll
      for (int c = 0; c < 8; ++c) dst[c * (WG_CB / 2)] = (uint16_t)(D[c >> 1] >> (16 * (c & 1)));
    };
    auto step = [&](int z) {
      const v8s a = *reinterpret_cast<const v8s *>(lds + 4 * WG_SB + ((z - z0) & 1) * WG_DB +
                                                   orow * WG_CPY + obyte);
#pragma unroll
      for (int i = 0; i < 8; ++i) bsum += bf2f((uint16_t)a[i]);
      const v8bf av = __builtin_bit_cast(v8bf, a);
      int sb[3];  // slots of planes z - 1, z, z + 1
#pragma unroll
      for (int d = 0; d < 3; ++d) sb[d] = ((z - z0 + d) & 3) * WG_SB;
#pragma unroll
      for (int t = 0; t < NTAP; ++t) {
        const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
        const v8s b = *reinterpret_cast<const v8s *>(lds + sb[dz] + (orow + dy) * WG_RB +
                                                     dx * WG_CPY + obyte);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(v8bf, b),
                                                         acc[t], 0, 0, 0);
      }
    };
    v4u R0[WG_XQ], R1[WG_XQ], D0, D1;
    ldx(R0, z0 - 1);
    ldx(R1, z0);
    ldd(D0, z0);
    ldd(D1, z0 + 1);
    __syncthreads();  // the previous item's reads are done
    stx(R0, z0 - 1);
    stx(R1, z0);
    ldx(R0, z0 + 1);
    ldx(R1, z0 + 2);
    stx(R0, z0 + 1);
    std_(D0, z0);
    __syncthreads();
    // step z: loads of X plane z + 3 and dZ plane z + 2, the MFMAs of plane
    // z, then X plane z + 2 and dZ plane z + 1 into their slots and the
    // barrier; two steps per trip name the register sets statically
    int z = z0;
    for (; z + 1 < z1; z += 2) {
      ldx(R0, z + 3);
      ldd(D0, z + 2);
      step(z);
      stx(R1, z + 2);
      std_(D1, z + 1);
      __syncthreads();
      ldx(R1, z + 4);
      ldd(D1, z + 3);
      step(z + 1);
      stx(R0, z + 3);  // past the chunk (no step follows): zeros into a slot no step reads
      std_(D0, z + 2);
      __syncthreads();
    }
    if (z < z1) step(z);
  }
  // the workgroup's sum in wave order through LDS
  float *red = reinterpret_cast<float *>(lds);  // [NTAP][4][64] + [64]
  __syncthreads();
  for (int v = 0; v < 4; ++v) {
    if (wl == v) {
#pragma unroll
      for (int t = 0; t < NTAP; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float &r = red[(t * 4 + i) * 64 + lane];
          r = v == 0 ? acc[t][i] : r + acc[t][i];
        }
      float &rb = red[NTAP * 4 * 64 + lane];
      rb = v == 0 ? bsum : rb + bsum;
    }
    __syncthreads();
  }
  float *pw = part + (int64_t)blockIdx.x * WL;
  for (int e = threadIdx.x; e < NTAP * 4 * 64; e += blockDim.x) {
    const int t = e / 256, i = (e / 64) % 4, l = e % 64;
    pw[(t * NC + 4 * (l >> 4) + i) * NC + (l & 15)] = red[e];
  }
  if (threadIdx.x < 64) partb[(int64_t)blockIdx.x * 64 + threadIdx.x] = red[NTAP * 4 * 64 + threadIdx.x];
}

// dW = sum of the workgroup partials in two fixed-order stages: stage 1
// sums slice j of WS_SPLIT consecutive partials of each element (eight
// interleaved running sums) into part2[j], the bias lanes likewise; stage 2
// adds the slices in order (db[co] also over the four lane groups).
constexpr int WS_SPLIT = 16;
constexpr int WS_P2 = WL + 64;  // elements of one slice sum (weights, bias lanes)
__global__ void k_wsum1(const float *__restrict__ part, const float *__restrict__ partb,
                        float *__restrict__ part2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  constexpr int per = NBLK_W / WS_SPLIT;
  static_assert(NBLK_W % (WS_SPLIT * 8) == 0, "slices of whole groups of 8");
  if (i >= WS_P2) return;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int w = j * per; w < (j + 1) * per; w += 8)
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[k] += i < WL ? part[(int64_t)(w + k) * WL + i] : partb[(int64_t)(w + k) * 64 + i - WL];
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += acc[k];
  part2[(int64_t)j * WS_P2 + i] = s;
}

__global__ void k_wsum2(const float *__restrict__ part2, float *__restrict__ dW,
                        float *__restrict__ db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < WL) {
    float s = 0.f;
    for (int j = 0; j < WS_SPLIT; ++j) s += part2[(int64_t)j * WS_P2 + i];
    dW[i] = s;
  } else if (i < WL + NC) {
    const int co = i - WL;
    float s = 0.f;
    for (int j = 0; j < WS_SPLIT; ++j)
      for (int c = 0; c < 4; ++c) s += part2[(int64_t)j * WS_P2 + WL + c * 16 + co];
    db[co] = s;
  }
}

// 16-channel input tensor from FP64 node vectors (channels 6..15 zero); the
// step increment x1 - x2 is formed in FP64 before the BF16 rounding
__global__ void k_pack(int64_t n, const double *x1, const double *x2, const double *src,
                       const double *vx, const double *vy, const double *vz,
                       float s0, float s1, float s2, float s3, float s4, float s5,
                       uint16_t *X0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    v4s o0, o1;
    o0[0] = (short)f2bf((float)(x1[i] * (double)s0));
    o0[1] = (short)f2bf((float)((x1[i] - x2[i]) * (double)s1));
    o0[2] = (short)f2bf((float)(src[i] * (double)s2));
    o0[3] = (short)f2bf((float)(vx[i] * (double)s3));
    o1[0] = (short)f2bf((float)(vy[i] * (double)s4));
    o1[1] = (short)f2bf((float)(vz[i] * (double)s5));
    o1[2] = o1[3] = 0;
    v4s z = {0, 0, 0, 0};
    v4s *p = reinterpret_cast<v4s *>(X0 + i * NC);
    p[0] = o0;
    p[1] = o1;
    p[2] = z;
    p[3] = z;
  }
}

// guess = x1 + ex (x1 - x2) + os * out
__global__ void k_guess(int64_t n, const double *x1, const double *x2, const float *out,
                        double ex, double os, double *guess) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    guess[i] = x1[i] + ex * (x1[i] - x2[i]) + os * (double)out[i];
}

// Loss = sum over free nodes of (out - t)^2 / count with
// t = (target - x1 - ex (x1 - x2)) / os; gout = dLoss/dout = 2 (out - t) /
// count; block partials.
__global__ __launch_bounds__(256) void k_loss(int64_t n, const float *out, const double *x1,
                                              const double *x2, const double *target,
                                              const uint8_t *mask, double ex, double inv_os,
                                              float inv_count, float *gout,
                                              double *lpart) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float d = 0.f;
    if (!mask[i]) d = out[i] - (float)((target[i] - x1[i] - ex * (x1[i] - x2[i])) * inv_os);
    s += (double)d * (double)d;
    gout[i] = 2.f * d * inv_count;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Residual loss, step 1: base = x1 + ex (x1 - x2), g = base + os out,
// e = target - g and e0 = target - base on free nodes (0 on Dirichlet nodes).
__global__ void k_res_prep(int64_t n, const double *x1, const double *x2, const float *out,
                           const double *target, const uint8_t *mask, double ex, double os,
                           double *e, double *e0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double base = x1[i] + ex * (x1[i] - x2[i]);
    const bool m = mask[i] != 0;
    e[i] = m ? 0.0 : target[i] - (base + os * (double)out[i]);
    e0[i] = m ? 0.0 : target[i] - base;
  }
}

// block partials of a.a and b.b
__global__ __launch_bounds__(256) void k_sq2(int64_t n, const double *a, const double *b,
                                             double *pa, double *pb) {
  __shared__ double ra[4], rb[4];
  double sa = 0.0, sb = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    sa += a[i] * a[i];
    sb += b[i] * b[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    sa += __shfl_down(sa, o, 64);
    sb += __shfl_down(sb, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    ra[threadIdx.x >> 6] = sa;
    rb[threadIdx.x >> 6] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = (ra[0] + ra[1]) + (ra[2] + ra[3]);
    pb[blockIdx.x] = (rb[0] + rb[1]) + (rb[2] + rb[3]);
  }
}

// Residual loss, step 2: dL/dout = coef * u (u = A^T A e) on free nodes
__global__ void k_res_grad(int64_t n, const double *u, const uint8_t *mask, double coef,
                           float *gout) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    gout[i] = mask[i] ? 0.f : (float)(coef * u[i]);
}

// One axis of the [1 4 1] / 6 filter (zero padding, so the filter is
// symmetric and its own transpose): damps the grid-scale (checkerboard)
// modes, the slowest ones of a mass-dominated operator, by 1/3 per axis.
__global__ void k_smooth_axis(Grid3 g, int axis, const float *in, float *out) {
  const int64_t n = g.N;
  const int64_t st = axis == 0 ? 1 : (axis == 1 ? (int64_t)g.nax : (int64_t)g.nax * g.nay);
  const int len = axis == 0 ? g.nax : (axis == 1 ? g.nay : g.naz);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)((i / st) % len);
    float v = 4.f * in[i];
    if (c > 0) v += in[i - st];
    if (c < len - 1) v += in[i + st];
    out[i] = v * (1.f / 6.f);
  }
}

__global__ void k_copy_f(int64_t n, const float *a, float *b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// dZ4[n][0] = bf16(g[n]) (other channels stay zero)
__global__ void k_pack_grad(int64_t n, const float *g, uint16_t *dZ4) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dZ4[i * NC] = f2bf(g[i]);
}

// Adam (bias-corrected) on n parameters
__global__ void k_adam(int n, float *p, const float *gr, float *m, float *v, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of the weights: W [l][t][co][ci] and per-tap transposes
__global__ void k_cast_w(const float *W, uint16_t *Wb, uint16_t *WbT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NL * WL) return;
  const int l = i / WL, r = i % WL, t = r / (NC * NC), co = (r / NC) % NC, ci = r % NC;
  const uint16_t h = f2bf(W[i]);
  Wb[i] = h;
  WbT[l * WL + (t * NC + ci) * NC + co] = h;
}

inline dim3 grid1(int64_t n, int bs = 256) {
  int64_t gsz = (n + bs - 1) / bs;
  if (gsz > 4096) gsz = 4096;
  if (gsz < 1) gsz = 1;
  return dim3((unsigned)gsz);
}

}  // namespace
}  // namespace vgpu

struct vgpu_net {
  vgpu_ctx *c = nullptr;
  int device = 0;  // the context's device (destroy does not touch the context)
  int64_t N = 0, nseg = 0, count = 0;
  uint16_t *X0 = nullptr, *Y[NL - 1] = {}, *dZ4 = nullptr, *dZa = nullptr, *dZb = nullptr;
  float *out = nullptr;
  float *W = nullptr, *B = nullptr, *mW = nullptr, *vW = nullptr, *mB = nullptr,
        *vB = nullptr, *dW = nullptr, *dB = nullptr, *part = nullptr, *partb = nullptr,
        *part2 = nullptr;
  uint16_t *Wb = nullptr, *WbT = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr, *lpart = nullptr;
  float in_scale[6] = {1, 1, 1, 1, 1, 1};
  double out_scale = 1.0, extrap = 0.0;
  int adam_t = 0;
  int loss_kind = 0;  // 0: increment L2, 1: operator residual
  int smooth = 0;     // passes of the [1 4 1]^3 / 216 filter on the output
  int zchunk = 0;     // node planes per wave of the layer kernels (0: automatic)
  int layer_zc() const { return zchunk > 0 ? std::min(zchunk, c->g.naz) : conv_zchunk(c->g); }
  double *e = nullptr, *e0 = nullptr, *r = nullptr, *r0 = nullptr;
  float *fo = nullptr, *ft = nullptr;  // filtered output / gradient, scratch

  ~vgpu_net() {
    void *ptrs[] = {X0, Y[0], Y[1], Y[2], Y[3], dZ4, dZa, dZb, out, W,  B, mW, vW, mB,
                    vB, dW, dB, part, partb, part2, Wb, WbT, vx, vy, vz, lpart, e, e0, r, r0,
                    fo, ft};
    for (void *p : ptrs)
      if (p) hipFree(p);
  }

  int forward(const double *x1, const double *x2, const double *src) {
    hipStream_t s = c->stream;
    const Grid3 &g = c->g;
    hipLaunchKernelGGL(k_pack, grid1(N), dim3(256), 0, s, N, x1, x2, src, vx, vy, vz,
                       in_scale[0], in_scale[1], in_scale[2], in_scale[3], in_scale[4],
                       in_scale[5], X0);
    const int zc = layer_zc();
    const dim3 grid = conv_grid(g, zc), blk(256);
    const uint16_t *in = X0;
    for (int l = 0; l < NL - 1; ++l) {
      hipLaunchKernelGGL((k_conv_t<FWD_HIDDEN>), grid, blk, 0, s, g, zc, in, Wb + l * WL,
                         B + l * NC, nullptr, Y[l], nullptr);
      in = Y[l];
    }
    hipLaunchKernelGGL((k_conv_t<FWD_LAST>), grid, blk, 0, s, g, zc, in, Wb + (NL - 1) * WL,
                       B + (NL - 1) * NC, nullptr, nullptr, out);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }

  // v <- S^passes v (S = the three axis filters); uses ft as scratch
  int filter(float *v) {
    hipStream_t s = c->stream;
    for (int p = 0; p < smooth; ++p)
      for (int ax = 0; ax < 3; ++ax) {
        hipLaunchKernelGGL(k_smooth_axis, grid1(N), dim3(256), 0, s, c->g, ax, v, ft);
        hipLaunchKernelGGL(k_copy_f, grid1(N), dim3(256), 0, s, N, ft, v);
      }
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
  // the output the guess uses: out, or S^passes out in fo
  const float *final_out() {
    if (!smooth) return out;
    hipLaunchKernelGGL(k_copy_f, grid1(N), dim3(256), 0, c->stream, N, out, fo);
    filter(fo);
    return fo;
  }

  int weight_grad(int l, const uint16_t *Xin, const uint16_t *dZl) {
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_conv_wg, dim3(NBLK_W), dim3(256), 0, s, c->g, wgrad_zchunk(c->g), Xin,
                       dZl, part, partb);
    hipLaunchKernelGGL(k_wsum1, dim3((WS_P2 + 255) / 256, WS_SPLIT), dim3(256), 0, s, part,
                       partb, part2);
    hipLaunchKernelGGL(k_wsum2, dim3((WL + NC + 255) / 256), dim3(256), 0, s, part2,
                       dW + l * WL, dB + l * NC);
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  }
};

extern "C" {

int vgpu_net_create(vgpu_ctx *c, vgpu_net **out) {
  if (!c || !out) return VGPU_ERR_ARG;
  *out = nullptr;
  // the layer kernels address one node plane per buffer resource
  if ((int64_t)c->g.nax * c->g.nay * NC * (int64_t)sizeof(uint16_t) >= CT_OOB) {
    set_error("vgpu_net_create: node planes above 2^30 bytes of 16 BF16 channels");
    return VGPU_ERR_ARG;
  }
  VGPU_HIP(hipSetDevice(c->device));
  auto *n = new (std::nothrow) vgpu_net();
  if (!n) return VGPU_ERR_NOMEM;
  n->c = c;
  n->device = c->device;
  n->N = c->g.N;
  n->nseg = (int64_t)((c->g.nax + 15) / 16) * c->g.nay * c->g.naz;
  const size_t T = (size_t)n->N * NC * sizeof(uint16_t);
  hipError_t e = hipSuccess;
  auto A = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
    if (e == hipSuccess) e = hipMemset(*p, 0, bytes);
  };
  A((void **)&n->X0, T);
  for (int l = 0; l < NL - 1; ++l) A((void **)&n->Y[l], T);
  A((void **)&n->dZ4, T);
  A((void **)&n->dZa, T);
  A((void **)&n->dZb, T);
  A((void **)&n->out, n->N * sizeof(float));
  A((void **)&n->fo, n->N * sizeof(float));
  A((void **)&n->ft, n->N * sizeof(float));
  for (float **p : {&n->W, &n->mW, &n->vW, &n->dW}) A((void **)p, NL * WL * sizeof(float));
  for (float **p : {&n->B, &n->mB, &n->vB, &n->dB}) A((void **)p, NL * NC * sizeof(float));
  A((void **)&n->part, (size_t)NBLK_W * WL * sizeof(float));
  A((void **)&n->partb, (size_t)NBLK_W * 64 * sizeof(float));
  A((void **)&n->part2, (size_t)WS_SPLIT * WS_P2 * sizeof(float));
  A((void **)&n->Wb, NL * WL * sizeof(uint16_t));
  A((void **)&n->WbT, NL * WL * sizeof(uint16_t));
  for (double **p : {&n->vx, &n->vy, &n->vz, &n->e, &n->e0, &n->r, &n->r0})
    A((void **)p, n->N * sizeof(double));
  A((void **)&n->lpart, 4096 * sizeof(double));
  if (e != hipSuccess) {
    set_error(std::string("vgpu_net_create: ") + hipGetErrorString(e));
    delete n;
    return e == hipErrorOutOfMemory ? VGPU_ERR_NOMEM : VGPU_ERR_HIP;
  }
  // free (non-Dirichlet) nodes: the loss normaliser
  std::vector<uint8_t> m(n->N);
  if (hipMemcpy(m.data(), c->mask, n->N, hipMemcpyDeviceToHost) != hipSuccess) {
    delete n;
    return VGPU_ERR_HIP;
  }
  int64_t cnt = 0;
  for (uint8_t v : m) cnt += v == 0;
  n->count = std::max<int64_t>(1, cnt);
  *out = n;
  return VGPU_OK;
}

int vgpu_net_destroy(vgpu_net *n) {
  if (!n) return VGPU_ERR_ARG;
  // the context may already be destroyed (its stream with it): wait for the
  // device instead of the context's stream
  hipSetDevice(n->device);
  hipDeviceSynchronize();
  delete n;
  return VGPU_OK;
}

int vgpu_net_set_params(vgpu_net *n, const float *w, const float *b) {
  if (!n || !w || !b) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  VGPU_HIP(hipMemcpy(n->W, w, NL * WL * sizeof(float), hipMemcpyHostToDevice));
  VGPU_HIP(hipMemcpy(n->B, b, NL * NC * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_cast_w, dim3((NL * WL + 255) / 256), dim3(256), 0, n->c->stream,
                     n->W, n->Wb, n->WbT);
  VGPU_HIP(hipGetLastError());
  VGPU_HIP(hipStreamSynchronize(n->c->stream));
  return VGPU_OK;
}

int vgpu_net_get_params(vgpu_net *n, float *w, float *b) {
  if (!n || !w || !b) return VGPU_ERR_ARG;
  VGPU_HIP(hipStreamSynchronize(n->c->stream));
  VGPU_HIP(hipMemcpy(w, n->W, NL * WL * sizeof(float), hipMemcpyDeviceToHost));
  VGPU_HIP(hipMemcpy(b, n->B, NL * NC * sizeof(float), hipMemcpyDeviceToHost));
  return VGPU_OK;
}

int vgpu_net_get_grads(vgpu_net *n, float *dw, float *db) {
  if (!n || !dw || !db) return VGPU_ERR_ARG;
  VGPU_HIP(hipStreamSynchronize(n->c->stream));
  VGPU_HIP(hipMemcpy(dw, n->dW, NL * WL * sizeof(float), hipMemcpyDeviceToHost));
  VGPU_HIP(hipMemcpy(db, n->dB, NL * NC * sizeof(float), hipMemcpyDeviceToHost));
  return VGPU_OK;
}

int vgpu_net_set_static(vgpu_net *n, const double *vx, const double *vy, const double *vz,
                        const float *in_scale, double out_scale, double extrap) {
  if (!n || !vx || !vy || !vz || !in_scale || !(out_scale > 0)) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  VGPU_HIP(hipMemcpy(n->vx, vx, n->N * sizeof(double), hipMemcpyHostToDevice));
  VGPU_HIP(hipMemcpy(n->vy, vy, n->N * sizeof(double), hipMemcpyHostToDevice));
  VGPU_HIP(hipMemcpy(n->vz, vz, n->N * sizeof(double), hipMemcpyHostToDevice));
  std::memcpy(n->in_scale, in_scale, sizeof n->in_scale);
  n->out_scale = out_scale;
  n->extrap = extrap;
  return VGPU_OK;
}

int vgpu_net_predict_dev(vgpu_net *n, const double *x1, const double *x2,
                         const double *src, double *guess) {
  if (!n || !x1 || !x2 || !src || !guess) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  int rc = n->forward(x1, x2, src);
  if (rc) return rc;
  const float *o = n->final_out();
  hipLaunchKernelGGL(k_guess, grid1(n->N), dim3(256), 0, n->c->stream, n->N, x1, x2, o,
                     n->extrap, n->out_scale, guess);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

int vgpu_net_train_step_dev(vgpu_net *n, const double *x1, const double *x2,
                            const double *src, const double *target, float lr,
                            double *loss) {
  if (!n || !x1 || !x2 || !src || !target) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  hipStream_t s = n->c->stream;
  int rc = n->forward(x1, x2, src);
  if (rc) return rc;
  const int nlb = 1024;
  double lval = 0.0;
  const float *o = n->final_out();  // fo when filtered
  float *gout = n->fo;              // dLoss/d(final output), then through S^T = S
  if (n->loss_kind == 0) {
    hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, s, n->N, o, x1, x2, target,
                       n->c->mask, n->extrap, 1.0 / n->out_scale,
                       (float)(1.0 / (double)n->count), gout, n->lpart);
    VGPU_HIP(hipGetLastError());
    std::vector<double> h(nlb);
    VGPU_HIP(hipMemcpyAsync(h.data(), n->lpart, nlb * sizeof(double), hipMemcpyDeviceToHost, s));
    VGPU_HIP(hipStreamSynchronize(s));
    for (double v : h) lval += v;
    lval /= (double)n->count;
  } else {
    // L = |A e|^2 / |A e0|^2, e = x_n - guess, e0 = x_n - base guess;
    // dL/dout = -2 os A^T A e / |A e0|^2 (A = the masked FP64 operator)
    hipLaunchKernelGGL(k_res_prep, grid1(n->N), dim3(256), 0, s, n->N, x1, x2, o,
                       target, n->c->mask, n->extrap, n->out_scale, n->e, n->e0);
    VGPU_HIP(hipGetLastError());
    rc = vgpu_apply_dev(n->c, VGPU_SCATTER_OWNER, 0, n->e, n->r);
    if (!rc) rc = vgpu_apply_dev(n->c, VGPU_SCATTER_OWNER, 0, n->e0, n->r0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sq2, dim3(nlb), dim3(256), 0, s, n->N, n->r, n->r0, n->lpart,
                       n->lpart + nlb);
    VGPU_HIP(hipGetLastError());
    std::vector<double> h(2 * nlb);
    VGPU_HIP(hipMemcpyAsync(h.data(), n->lpart, 2 * nlb * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    VGPU_HIP(hipStreamSynchronize(s));
    double ra = 0.0, rb = 0.0;
    for (int i = 0; i < nlb; ++i) {
      ra += h[i];
      rb += h[nlb + i];
    }
    const double D = rb > 0.0 ? rb : 1.0;
    lval = ra / D;
    rc = vgpu_apply_dev(n->c, VGPU_SCATTER_OWNER, 1, n->r, n->e0);  // u = A^T r
    if (rc) return rc;
    hipLaunchKernelGGL(k_res_grad, grid1(n->N), dim3(256), 0, s, n->N, n->e0, n->c->mask,
                       -2.0 * n->out_scale / D, gout);
    VGPU_HIP(hipGetLastError());
  }
  rc = n->filter(gout);  // the filter is symmetric: S^T g = S g
  if (rc) return rc;
  hipLaunchKernelGGL(k_pack_grad, grid1(n->N), dim3(256), 0, s, n->N, gout, n->dZ4);
  VGPU_HIP(hipGetLastError());
  // backward: layer 4 (no activation), then 3..0 through the ReLUs
  const int zc = n->layer_zc();
  const dim3 grid = conv_grid(n->c->g, zc), blk(256);
  const uint16_t *dz = n->dZ4;
  uint16_t *bufs[2] = {n->dZa, n->dZb};
  for (int l = NL - 1; l >= 0; --l) {
    const uint16_t *xin = l == 0 ? n->X0 : n->Y[l - 1];
    rc = n->weight_grad(l, xin, dz);
    if (rc) return rc;
    if (l > 0) {
      uint16_t *dprev = bufs[l & 1];
      hipLaunchKernelGGL((k_conv_t<BWD_DATA>), grid, blk, 0, s, n->c->g, zc, dz,
                         n->WbT + l * WL, nullptr, n->Y[l - 1], dprev, nullptr);
      dz = dprev;
    }
  }
  VGPU_HIP(hipGetLastError());
  if (loss) *loss = lval;
  if (lr > 0.f) {
    n->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->adam_t),
                c2 = 1.f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((NL * WL + 255) / 256), dim3(256), 0, s, NL * WL, n->W,
                       n->dW, n->mW, n->vW, lr, b1, b2, 1e-8f, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, s, NL * NC, n->B, n->dB, n->mB, n->vB,
                       lr, b1, b2, 1e-8f, c1, c2);
    hipLaunchKernelGGL(k_cast_w, dim3((NL * WL + 255) / 256), dim3(256), 0, s, n->W, n->Wb,
                       n->WbT);
    VGPU_HIP(hipGetLastError());
  }
  return VGPU_OK;
}

int vgpu_net_set_smoothing(vgpu_net *n, int passes) {
  if (!n || passes < 0 || passes > 8) return VGPU_ERR_ARG;
  n->smooth = passes;
  return VGPU_OK;
}

int vgpu_net_set_zchunk(vgpu_net *n, int zchunk) {
  if (!n || zchunk < 0) return VGPU_ERR_ARG;
  n->zchunk = zchunk;
  return VGPU_OK;
}

int vgpu_net_set_loss(vgpu_net *n, int kind) {
  if (!n || kind < 0 || kind > 1) return VGPU_ERR_ARG;
  n->loss_kind = kind;
  return VGPU_OK;
}

int vgpu_net_reset_optimizer(vgpu_net *n) {
  if (!n) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  for (float *p : {n->mW, n->vW}) VGPU_HIP(hipMemset(p, 0, NL * WL * sizeof(float)));
  for (float *p : {n->mB, n->vB}) VGPU_HIP(hipMemset(p, 0, NL * NC * sizeof(float)));
  n->adam_t = 0;
  return VGPU_OK;
}

int vgpu_net_time_layers(vgpu_net *n, int what, int reps, double *avg_ms) {
  if (!n || !avg_ms || reps < 1 || what < 0 || what > 3) return VGPU_ERR_ARG;
  VGPU_HIP(hipSetDevice(n->c->device));
  hipStream_t s = n->c->stream;
  const Grid3 &g = n->c->g;
  const int zc = n->layer_zc();
  const dim3 grid = conv_grid(g, zc), blk(256);
  // layer 1 (16 -> 16 channels) on the network's own buffers
  auto launch = [&]() -> int {
    switch (what) {
      case 0:
        hipLaunchKernelGGL((k_conv_t<FWD_HIDDEN>), grid, blk, 0, s, g, zc, n->Y[0],
                           n->Wb + WL, n->B + NC, nullptr, n->Y[1], nullptr);
        break;
      case 1:
        hipLaunchKernelGGL((k_conv_t<BWD_DATA>), grid, blk, 0, s, g, zc, n->dZa,
                           n->WbT + WL, nullptr, n->Y[0], n->dZb, nullptr);
        break;
      case 2:
        return n->weight_grad(1, n->Y[0], n->dZa);
      default:  // the whole prediction network on zero states
        return n->forward(n->e, n->e0, n->r);
    }
    VGPU_HIP(hipGetLastError());
    return VGPU_OK;
  };
  for (int i = 0; i < 2; ++i) {
    int rc = launch();
    if (rc) return rc;
  }
  // the timed launches are captured in one graph, so the events bracket GPU
  // work only
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  VGPU_HIP(hipStreamSynchronize(s));
  VGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int lrc = VGPU_OK;
  for (int i = 0; i < reps && lrc == VGPU_OK; ++i) lrc = launch();
  const hipError_t ce = hipStreamEndCapture(s, &graph);
  if (lrc) {
    if (graph) hipGraphDestroy(graph);
    return lrc;
  }
  VGPU_HIP(ce);
  VGPU_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  hipEvent_t e0, e1;
  VGPU_HIP(hipEventCreate(&e0));
  VGPU_HIP(hipEventCreate(&e1));
  VGPU_HIP(hipGraphLaunch(gexec, s));  // warm-up replay
  VGPU_HIP(hipEventRecord(e0, s));
  VGPU_HIP(hipGraphLaunch(gexec, s));
  VGPU_HIP(hipEventRecord(e1, s));
  VGPU_HIP(hipEventSynchronize(e1));
  float ms = 0;
  VGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipGraphExecDestroy(gexec);
  hipGraphDestroy(graph);
  *avg_ms = ms / reps;
  return VGPU_OK;
}

}  // extern "C"
