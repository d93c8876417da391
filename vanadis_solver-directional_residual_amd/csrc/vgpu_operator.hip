// Matrix-free DR-HEX8 operator kernels for gfx950.
//
// Every kernel marches a 32 x 8 element-column tile up the z axis, one element
// layer per step: the node plane above the layer is staged in LDS (x-fastest,
// coalesced), each of the 256 threads applies one element (Haar-space closed
// form, vgpu_internal.hpp), the 8 nodal results go to an LDS staging array and
// each node sums its (up to) 4 in-plane contributions in a fixed order.
//
//  * OWNER   (deterministic, one launch): tiles overlap by one element column,
//            each tile owns 31 x 7 node columns and writes them once with plain
//            stores -- no atomics, no pre-zeroing, no read-modify-write.
//  * COLORED (deterministic): disjoint 32 x 8 x zc tiles in 8 parity colours,
//            one launch per colour, tile-face nodes read-modify-written.
//  * ATOMIC  : disjoint tiles in one launch, tile-face nodes added with
//            hardware atomics (global_atomic_add_f64 / _f32).
// Reference semantics: y = (I-D) A (I-D) x as matvec_0 / mttvec_0
// (VANADIS.FOR:2149-2245) with A the element systems of przest3d (723-748).
#include "vgpu_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <map>
#include <mutex>

namespace vgpu {

namespace {

template <typename T>
struct Coef {
  T lam, vx, vy, vz, tau, hz, ihz;
};

// per-element fields (zero outside the grid) and the layer's z size
template <typename T>
__device__ __forceinline__ Coef<T> load_coef(const OpData<T> &op, bool valid,
                                             int64_t e, int kz) {
  Coef<T> c;
  // kz is uniform over the workgroup: scalar loads of the layer size
  const int kc = min(max(kz, 0), op.g.ez - 1);
  c.hz = op.hz[kc];
  c.ihz = op.ihz[kc];
  if (valid) {
    const T *r = op.rec + 5 * e;
    c.vx = r[0];
    c.vy = r[1];
    c.vz = r[2];
    c.lam = r[3];
    c.tau = r[4];
  } else {
    c.lam = c.vx = c.vy = c.vz = c.tau = T(0);
  }
  return c;
}

// Sum of the four in-plane contributions to tile node (lx, ly); `top` selects
// the upper face (local nodes 4..7) of the layer.  Elements outside the tile
// contribute nothing.  Fixed order -> deterministic.
template <typename T, int TX, int TY>
__device__ __forceinline__ double node_sum(const T *stage, int lx, int ly,
                                           bool top) {
  const int o = top ? 4 : 0;
  double s = 0.0;
  if (lx >= 1 && ly >= 1) s += (double)stage[(o + 2) * (TX * TY) + (ly - 1) * TX + lx - 1];
  if (lx < TX && ly >= 1) s += (double)stage[(o + 3) * (TX * TY) + (ly - 1) * TX + lx];
  if (lx >= 1 && ly < TY) s += (double)stage[(o + 1) * (TX * TY) + ly * TX + lx - 1];
  if (lx < TX && ly < TY) s += (double)stage[(o + 0) * (TX * TY) + ly * TX + lx];
  return s;
}

// Plane traffic of one thread: tile-plane nodes q0 = tid and q1 = tid + NT
// (q1 < (TX+1)(TY+1) only): raw values of each input and the Dirichlet flag
// (1 also outside the grid).  The loads are only issued here; the mask is
// applied when the plane is written to LDS one layer later, so no wait on
// the loads is needed before then.
template <typename TIO, bool DUAL>
struct PlaneRegs {
  TIO a0, a1, b0, b1;
  uint8_t m0, m1;
};

template <typename TIO, bool DUAL>
__device__ __forceinline__ void node_raw(const Grid3 &g, const uint8_t *mask,
                                         const TIO *x1, const TIO *x2, int gx,
                                         int gy, int gz, TIO &a, TIO &b,
                                         uint8_t &m) {
  a = b = TIO(0);
  m = 1;
  if (gx >= 0 && gy >= 0 && gz >= 0 && gx < g.nax && gy < g.nay && gz < g.naz) {
    const int64_t n = gx + (int64_t)g.nax * (gy + (int64_t)g.nay * gz);
    m = mask[n];
    a = x1[n];
    if (DUAL) b = x2[n];
  }
}

template <typename T, typename TIO, bool DUAL>
__device__ __forceinline__ void plane_load(const OpData<T> &op, const TIO *x1,
                                           const TIO *x2, int gx0, int gy0,
                                           int gz, int lx0, int ly0, int lx1,
                                           int ly1, bool has1,
                                           PlaneRegs<TIO, DUAL> &r) {
  node_raw<TIO, DUAL>(op.g, op.mask, x1, x2, gx0 + lx0, gy0 + ly0, gz, r.a0, r.b0, r.m0);
  r.a1 = r.b1 = TIO(0);
  r.m1 = 1;
  if (has1)
    node_raw<TIO, DUAL>(op.g, op.mask, x1, x2, gx0 + lx1, gy0 + ly1, gz, r.a1, r.b1, r.m1);
}

template <typename T, typename TIO, bool DUAL>
__device__ __forceinline__ void plane_store(const PlaneRegs<TIO, DUAL> &r, T *p1,
                                            T *p2, uint8_t *pm, int q0, int q1,
                                            bool has1) {
  p1[q0] = r.m0 ? T(0) : T(r.a0);
  pm[q0] = r.m0;
  if (DUAL) p2[q0] = r.m0 ? T(0) : T(r.b0);
  if (has1) {
    p1[q1] = r.m1 ? T(0) : T(r.a1);
    pm[q1] = r.m1;
    if (DUAL) p2[q1] = r.m1 ? T(0) : T(r.b1);
  }
}

// ---------------------------------------------------------------------------
// Element apply in registers: nodal inputs x1 (and x2) in local node order,
// outputs y1 = A_e x1 (and y2 = A_e^T x2).  Invalid elements give zeros.
// ---------------------------------------------------------------------------
template <typename T, bool DUAL, bool ROBIN>
__device__ __forceinline__ void elem_eval(const OpData<T> &op, bool valid,
                                          bool robin, const ColConst<T> &q,
                                          const Coef<T> &cf, const T x1[8],
                                          const T x2[8], T y1[8], T y2[8]) {
  // an element outside the grid has zero fields (load_coef); with a zero mass
  // weight and no Robin term every constant vanishes and so do y1, y2
  ElemConst<T> k;
  elem_layer_setup<T>(q, cf.hz, cf.ihz, cf.lam, cf.vx, cf.vy, cf.vz, cf.tau,
                      T(op.cH), valid ? T(op.mu) : T(0), k);
  const bool rb = ROBIN && robin && valid;
  T xh[8], yh[8];
  haar_fwd<T>(x1, xh);
  haar_apply<T, false>(k, xh, yh);
  haar_inv<T>(yh, y1);
  if (rb)
    robin_add<T, false>(k, q.mx0 * T(4), cf.tau, cf.vx, cf.vy, cf.vz, T(op.ra), x1, y1);
  if (DUAL) {
    haar_fwd<T>(x2, xh);
    haar_apply<T, true>(k, xh, yh);
    haar_inv<T>(yh, y2);
    if (rb)
      robin_add<T, true>(k, q.mx0 * T(4), cf.tau, cf.vx, cf.vy, cf.vz, T(op.ra), x2, y2);
  }
}

// ---------------------------------------------------------------------------
// Raw buffer access for the owner kernel.  A load whose offset lies outside
// [0, bytes) returns zero and such a store is dropped, so a thread whose node,
// element or owned column lies outside the grid passes kOob as its vector
// offset instead of branching.  The per-layer part of an address is a scalar
// offset (uniform over the workgroup, no vector address arithmetic per layer).
// Every resource must stay below 2 GiB (checked on the host, owner_fits).
// ---------------------------------------------------------------------------
using Res = __amdgpu_buffer_rsrc_t;
constexpr int kOob = int(0x80000000u);

// a value the compiler may not prove uniform (a strength-reduced induction
// variable, say) but that is: keep it in a scalar register
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ Res make_res(const void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}

template <typename V>
__device__ __forceinline__ V bload(Res r, int v, int s);
template <>
__device__ __forceinline__ double bload<double>(Res r, int v, int s) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
}
template <>
__device__ __forceinline__ float bload<float>(Res r, int v, int s) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
}
template <>
__device__ __forceinline__ uint8_t bload<uint8_t>(Res r, int v, int s) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, v, s, 0);
}

__device__ __forceinline__ void bstore(double x, Res r, int v, int s) {
  __builtin_amdgcn_raw_buffer_store_b64(
      __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r, v, s, 0);
}
__device__ __forceinline__ void bstore(float x, Res r, int v, int s) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, s, 0);
}

// A vector of an owner launch: a resource of its own (offset 0), or the
// engine's block resource and the vector's byte offset in it (added to the
// scalar offset of every access)
struct VRef {
  Res r;
  int o;
};
template <typename V>
__device__ __forceinline__ V bload(const VRef &v, int vo, int so) {
  return bload<V>(v.r, vo, so + v.o);
}
__device__ __forceinline__ void bstore(double x, const VRef &v, int vo, int so) {
  bstore(x, v.r, vo, so + v.o);
}
__device__ __forceinline__ void bstore(float x, const VRef &v, int vo, int so) {
  bstore(x, v.r, vo, so + v.o);
}

// element record (vx, vy, vz, lambda, tau) at byte offset v + s
template <typename T>
__device__ __forceinline__ void load_rec(Res r, int v, int s, Coef<T> &c) {
  c.vx = bload<T>(r, v, s);
  c.vy = bload<T>(r, v + int(sizeof(T)), s);
  c.vz = bload<T>(r, v + 2 * int(sizeof(T)), s);
  c.lam = bload<T>(r, v + 3 * int(sizeof(T)), s);
  c.tau = bload<T>(r, v + 4 * int(sizeof(T)), s);
}

// ---------------------------------------------------------------------------
// OWNER kernel (persistent, balanced, batched).  The node columns of each
// system are cut into overlapping TX x TY element-column tiles that own
// (TX-1) x (TY-1) node columns each; a work unit is kUnitPlanes consecutive
// node planes of one tile, units are numbered system by system, tile-major
// (z fastest), and block b takes the contiguous range [b U / G, (b+1) U / G)
// of all U units, split into one march per (system, tile) it touches.
// Producing planes k0..k1-1 of a tile computes element layers k0-1..k1-1.
//
// Per layer kl (one barrier): the thread of element column (tx, ty) reads the
// 4 corners of the layer's top face from LDS and transforms them to 2-D Haar
// space once; the bottom face's transform is the previous layer's top face,
// kept in registers.  After the z stage and the Haar-space operator, the
// layer's bottom-face output is added to the previous layer's top-face output
// (both in Haar space) and one transposed face butterfly gives the 4 column
// contributions to plane kl, which are staged in LDS; the owner of a node adds
// the 4 columns around it one layer later.  Plane kl+1 is read only in layer
// kl and a stage only in the layer after it is written, so two LDS buffers of
// each suffice, and the march is unrolled by two so that their indices are
// constants.  The next plane (values, Dirichlet flags) and the next layer's
// element fields are loaded one layer ahead; the Dirichlet select is applied
// when the plane is written to LDS.  The dual mode also reduces y1 . x2 over
// each unit's owned nodes (per thread over the unit's planes in order, then
// over the block in a fixed order): a balanced launch stores one partial per
// unit (they do not depend on the grid), a z-chunk launch one per (chunk,
// tile) group (per thread over the group's planes, then over the block).
// ---------------------------------------------------------------------------
// With DIR (dual only) the launch also does the BiCG direction update of the
// iteration: x1, x2 are z, zz; the plane loads also read the previous
// directions, form p = z + beta po, pp = zz + beta ppo (masked in LDS as the
// operator's inputs), and the nodes of the tile's own columns and planes are
// stored to pn, ppn.
//
// With kStabDir / kStabS (single operator) the launch is one of the two
// operator launches of a BiCGSTAB iteration (see StabData): the plane loads
// form R, P (and x at owned nodes) or S, and the owned node's output is
// scaled by 1/d and reduced with values loaded at that node.
template <typename TIO, bool DIR>
struct PlanePair {
  TIO a0, a1, b0, b1;
  TIO c0, c1, d0, d1;  // DIR: previous directions of the two nodes; BiCGSTAB:
                       // c = V, d = previous P (kStabDir)
  TIO e0, e1;          // DIR, kStabDir: x of the nodes the march owns (else 0)
  TIO t0, t1;          // kStabDir: T of the previous S launch
  TIO f, g;            // BiCGSTAB: d and RT at the thread's owned node
  uint8_t m0, m1;
};

template <typename T, typename TIO, bool DUAL, bool ROBIN, bool BATCH, int DM, int TX, int TY>
__global__ __launch_bounds__(TX * TY)
__attribute__((amdgpu_waves_per_eu(
    sizeof(T) == 8 ? ((DUAL || (DM % kBlk >= 2 && DM % kBlk != kSlab)) ? 2 : 3)
                   : ((DUAL || (DM % kBlk >= 2 && DM % kBlk != kSlab)) ? 3 : 4)))) void k_owner(
    Member<T> m0, const Member<T> *__restrict__ md, const int *__restrict__ act,
    const int64_t *__restrict__ actu, const int64_t *__restrict__ actg, int nact,
    int64_t U, int zc, const SolveStatus *__restrict__ st, const TIO *__restrict__ x1,
    TIO *__restrict__ y1, const TIO *__restrict__ x2, TIO *__restrict__ y2,
    double *__restrict__ dot, DirData<TIO> dd, StabData<TIO> sd, int slab_color) {
  // BLK (DM + kBlk): the vectors of a BiCG or BiCGSTAB launch, but the
  // diagonal, lie in the engine's block: one resource for all of them
  constexpr bool BLK = DM >= kBlk;
  constexpr int DMM = DM % kBlk;
  constexpr bool DIR = DMM == 1;  // BiCG iteration launch
  // BiCGSTAB launches; DM + 2: the same with the early status test
  constexpr bool SA = DMM == kStabDir || DMM == kStabDir + 2;
  constexpr bool SB = DMM == kStabS || DMM == kStabS + 2, STAB = SA || SB;
  constexpr bool EARLY = DMM == kStabDir + 2 || DMM == kStabS + 2;
  // damped Jacobi sweep: output x + omega (b - A x) / d at the owned nodes
  constexpr bool JAC = DMM == kJacobi;
  // colored scatter (z-chunk launch only): the march of a z-slab of element
  // layers [k0, k1 - 1) with the layers just below and above it taken as
  // zero (no recomputed halo layer), so its bottom and top planes get the
  // slab's partial sums: the colour-0 launch (slab_color 0, the even slabs)
  // stores them, the colour-1 launch adds its own to them (the grid's top
  // plane, which has no slab above, is stored)
  constexpr bool SLAB = DMM == kSlab;
  static_assert(!BLK || DIR || STAB, "block addressing is for the solver launches");
  static_assert(DUAL || !DIR, "the direction update needs the dual operator");
  static_assert(!(DUAL && (STAB || JAC)), "the BiCGSTAB and Jacobi launches are single");
  constexpr int NP = (TX + 1) * (TY + 1);
  constexpr int NT = TX * TY;
  constexpr int NW = NT / 64;
  constexpr int SI = sizeof(TIO), SR = 5 * sizeof(T);
  __shared__ T pl1[2][NP];
  __shared__ T pl2[DUAL ? 2 : 1][DUAL ? NP : 1];
  __shared__ uint8_t pm[2][NP];
  __shared__ T st1[2][4 * NT];
  __shared__ T st2[DUAL ? 2 : 1][DUAL ? 4 * NT : 1];
  // dot sums per launch: the dual apply's w.pp; kStabDir RT.V and |r|^2;
  // kStabS t.s, t.t, r~.s and r~.t
  constexpr int ND = DUAL ? 1 : SA ? 2 : SB ? 4 : 1;
  __shared__ double red[ND][2][NW];
  __shared__ double rsum[4][NW];
  __shared__ T plR[SA ? 2 : 1][SA ? NP : 1];  // kStabDir: the new R of the plane
  __shared__ int sflag[3];
  __shared__ double salpha;
  __shared__ double sstab[2];  // BiCGSTAB: scalars of the status

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int64_t G = gridDim.x, lb = xcd_tile(blockIdx.x, gridDim.x);
  const bool want_dot = (DUAL && dot != nullptr) || STAB;

  const int q0 = tid, q1 = tid + NT;
  const int lx0 = q0 % (TX + 1), ly0 = q0 / (TX + 1);
  const int lx1 = q1 % (TX + 1), ly1 = q1 / (TX + 1);
  const bool has1 = q1 < NP;
  const bool own = tx < (TX - 1) && ty < (TY - 1);  // owns node (tx+1, ty+1)
  const int oq = (ty + 1) * (TX + 1) + tx + 1;
  const int c00 = ty * (TX + 1) + tx;  // tile node of local corner 0

  // BATCH: active members act[i], their units at [actu[i], actu[i+1]) of
  // the split, their groups at [actg[i], actg[i+1]) and their partials at
  // md[act[i]].seg.u0; else the member m0.  Binary search: the last i with
  // a[i] <= v.
  auto find = [&](const int64_t *a, int64_t v) {
    int lo = 0, hi = nact - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
  };

  // Block-wide sums of four partial arrays a[j][0..n[j]) in one pass, each in a
  // fixed order (per thread strided, DPP within waves, waves in order); the
  // results in every thread.
  auto blk_sum4 = [&](const double *const a[4], const int n[4], double out[4]) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int nmax = max(max(n[0], n[1]), max(n[2], n[3]));
    for (int i = tid; i < nmax; i += NT) {
      double x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = i < n[j] ? a[j][i] : 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] += x[j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = wave_sum_lane63(v[j]);
      if ((tid & 63) == 63) rsum[j][tid >> 6] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double t = 0.0;
      for (int w = 0; w < NW; ++w) t += rsum[j][w];
      out[j] = t;
    }
    __syncthreads();
  };

  // One march of member m: node planes k0..k1-1 of tile `tile` (element
  // layers k0-1..k1-1).  Every thread of the block takes the same path.
  // gidx >= 0 (z-chunk launch): the dot partial of the whole march goes to
  // slot gidx of the member's partials, the sum of its planes in order;
  // otherwise one partial per (tile, plane) unit.
  auto march = [&](int m, int tile, int k0, int k1, int64_t gidx) {
    if (!DIR && !STAB && st && st[m].done) return;
    if (EARLY && (BATCH ? sd.st[m].done : sd.st->done)) return;
    const OpData<T> op = BATCH ? md[m].op : m0.op;
    const int64_t off = BATCH ? md[m].seg.off : m0.seg.off;
    const int64_t ub = BATCH ? md[m].seg.u0 : m0.seg.u0;
    const int naz = op.g.naz;
    const int ngz = (naz + kUnitPlanes - 1) / kUnitPlanes;  // units per tile
    const int ntx = (op.g.nax + TX - 2) / (TX - 1);
    const int nax = op.g.nax, nay = op.g.nay, ex = op.g.ex, ey = op.g.ey, ez = op.g.ez;
    const int plane = nax * nay, eplane = ex * ey;
    // BLK: one resource over the engine's block, a scalar byte offset per
    // vector (four scalar registers in all instead of four per vector)
    const char *const vbase = BLK ? (const char *)(DIR ? dd.blk : sd.blk) : nullptr;
    const Res rvb = make_res(vbase, BLK ? (DIR ? dd.blk_bytes : sd.blk_bytes) : 0);
    auto vref = [&](const TIO *p) {
      VRef v;
      if constexpr (BLK) {
        v.r = rvb;
        v.o = (int)((const char *)p - vbase);
      } else {
        v.r = make_res(p, op.g.N * SI);
        v.o = 0;
      }
      return v;
    };
    const VRef rx1 = vref(x1 + off);
    const VRef rx2 = vref(DUAL ? x2 + off : x1 + off);
    const VRef ry1 = vref(y1 + off);
    const VRef ry2 = vref(DUAL ? y2 + off : y1 + off);
    const Res rmk = make_res(op.mask, op.g.N);
    const Res rrc = make_res(op.rec, (int64_t)eplane * ez * SR);
    const int64_t dslot = ub + (gidx >= 0 ? gidx : (int64_t)tile * ngz);
    double *dm[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
      dm[i] = STAB ? sd.pout[i] + dslot : want_dot ? dot + dslot : nullptr;

    const int gx0 = (tile % ntx) * (TX - 1) - 1;  // tile node/element origin
    const int gy0 = (tile / ntx) * (TY - 1) - 1;
    const int exg = gx0 + tx, eyg = gy0 + ty;
    const bool colok = exg >= 0 && exg < ex && eyg >= 0 && eyg < ey;
    ColConst<T> cq;
    col_setup<T>(colok ? op.hx[exg] : T(1), colok ? op.hy[eyg] : T(1),
                 colok ? op.ihx[exg] : T(1), colok ? op.ihy[eyg] : T(1), cq);
    // outside the grid the fields load as zero; with a zero mass weight every
    // element constant and so the element's result vanish
    const T mu_c = colok ? T(op.mu) : T(0);
    const bool robin = ROBIN && op.ra != 0.0 && colok && exg == 0;
    const int ngx = gx0 + 1 + tx, ngy = gy0 + 1 + ty;
    const bool owng = own && ngx < nax && ngy < nay;
    const int vo = owng ? (ngx + nax * ngy) * SI : kOob;  // owned node column
    const int vr = colok ? (exg + ex * eyg) * SR : kOob;  // element column
    auto colofs = [&](int ax, int ay) {
      return (ax >= 0 && ay >= 0 && ax < nax && ay < nay) ? ax + nax * ay : -1;
    };
    const int pc0 = colofs(gx0 + lx0, gy0 + ly0);
    const int pc1 = has1 ? colofs(gx0 + lx1, gy0 + ly1) : -1;
    const int vp0 = pc0 >= 0 ? pc0 * SI : kOob, vp1 = pc1 >= 0 ? pc1 * SI : kOob;
    const int vm0 = pc0 >= 0 ? pc0 : kOob, vm1 = pc1 >= 0 ? pc1 : kOob;
    // DIR: previous and new directions; a plane node of the tile's own node
    // columns stores its new direction (kOob elsewhere)
    // BiCGSTAB: c = V (previous V for kStabDir), d = previous P, t = T of the
    // previous S launch; new P, R and x of owned nodes to rpn, rrn, rxs; d
    // and RT read at the owned node column vo
    const TIO *cptr = DIR ? dd.po : SA ? sd.vo : SB ? sd.v : x1;
    const TIO *dptr = DIR ? dd.ppo : SA ? sd.po : x1;
    const VRef rpo = vref(cptr + off);
    const VRef rppo = vref(dptr + off);
    const VRef rpn = vref(DIR ? dd.pn + off : SA ? sd.pn + off : y1 + off);
    const VRef rrn = vref(SA ? sd.rn + off : y1 + off);
    const VRef rto = vref(SA ? sd.to + off : x1 + off);
    // owned-node values: d and RT (BiCGSTAB), d and b (Jacobi sweep); the
    // diagonal is the context's, outside the engine's block
    const Res rdg = make_res((STAB || JAC) ? sd.d + off : x1 + off, op.g.N * SI);
    const VRef rrt = vref((STAB || JAC) ? sd.rt + off : x1 + off);
    const VRef rppn = vref(DIR ? dd.ppn + off : y1 + off);
    auto owned_col = [&](int lx, int ly) {
      return lx >= 1 && lx <= TX - 1 && ly >= 1 && ly <= TY - 1;
    };
    const int vn0 = (pc0 >= 0 && owned_col(lx0, ly0)) ? vp0 : kOob;
    const int vn1 = (pc1 >= 0 && owned_col(lx1, ly1)) ? vp1 : kOob;
    const VRef rxs = vref(DIR ? dd.x + off : SA ? sd.x + off : y1 + off);
    TIO ap = TIO(0);  // DIR: alpha of the previous update (x += ap po)
    // DIR: beta and the first-iteration flag, known after the status test
    // below; dir() is only evaluated after it.  BiCGSTAB: beta, omega
    // (kStabDir) or alpha (kStabS) and the first-iteration flag.
    TIO bt = TIO(0), om = TIO(0), al = TIO(0);
    bool first = false;
    auto dir = [&](TIO zv, TIO pv) { return first ? zv : zv + bt * pv; };
    // kStabDir, one node from its loads R, V, P, T of the previous iteration:
    // S = R - alpha V, the new R = S - omega T and P = R + beta (P - omega V)
    // (iteration 0: R and P = R as loaded)
    auto stab_rp = [&](TIO a, TIO c, TIO d, TIO t, TIO &sv, TIO &rv, TIO &pv) {
      sv = a - al * c;
      rv = first ? a : sv - om * t;
      pv = first ? a : rv + bt * (d - om * c);
    };

    // this thread's two plane nodes of plane gz (raw values and flags; value
    // 0 outside the grid: the lane offsets of a plane above the top get the
    // out-of-range bit, at a clamped scalar offset, so every resource stays
    // one and the same for the whole march)
    auto plane_fetch = [&](int gz, PlanePair<TIO, DIR> &r) {
      const int ob = uni(gz < (SLAB ? k1 : naz) ? 0 : kOob);
      const int gc = min(gz, naz - 1);
      const int sm = uni(gc * plane), sv = uni(gc * plane * SI);
      const int p0 = vp0 | ob;
      r.m0 = bload<uint8_t>(rmk, vm0 | ob, sm);
      r.a0 = bload<TIO>(rx1, p0, sv);
      r.b0 = DUAL ? bload<TIO>(rx2, p0, sv) : TIO(0);
      r.c0 = (DIR || STAB) ? bload<TIO>(rpo, p0, sv) : TIO(0);
      r.d0 = (DIR || SA) ? bload<TIO>(rppo, p0, sv) : TIO(0);
      r.t0 = SA ? bload<TIO>(rto, p0, sv) : TIO(0);
      r.f = (STAB || JAC) ? bload<TIO>(rdg, vo | ob, sv) : TIO(0);
      r.g = (STAB || JAC) ? bload<TIO>(rrt, vo | ob, sv) : TIO(0);
      // x of owned nodes of the march's own planes only
      const int oe = uni(gz < k1 ? 0 : kOob);
      r.e0 = (DIR || SA) ? bload<TIO>(rxs, vn0 | oe, sv) : TIO(0);
      r.a1 = r.b1 = r.c1 = r.d1 = r.e1 = r.t1 = TIO(0);
      r.m1 = 1;
      if (has1) {
        const int p1 = vp1 | ob;
        r.m1 = bload<uint8_t>(rmk, vm1 | ob, sm);
        r.a1 = bload<TIO>(rx1, p1, sv);
        if (DUAL) r.b1 = bload<TIO>(rx2, p1, sv);
        if (DIR || STAB) r.c1 = bload<TIO>(rpo, p1, sv);
        if (DIR || SA) {
          r.d1 = bload<TIO>(rppo, p1, sv);
          r.e1 = bload<TIO>(rxs, vn1 | oe, sv);
        }
        if (SA) r.t1 = bload<TIO>(rto, p1, sv);
      }
    };
    // plane gz into LDS; DIR: the directions are formed first and those of
    // the march's own planes stored
    auto plane_put = [&](const PlanePair<TIO, DIR> &r, int gz, T *p1, T *p2, uint8_t *pk,
                         T *prr) {
      TIO a0 = r.a0, b0 = r.b0, a1 = r.a1, b1 = r.b1;
      if (DIR) {
        a0 = dir(r.a0, r.c0);
        b0 = dir(r.b0, r.d0);
        a1 = dir(r.a1, r.c1);
        b1 = dir(r.b1, r.d1);
        if (gz >= k0 && gz < k1) {
          const int sv = uni(gz * plane * SI);
          bstore(a0, rpn, vn0, sv);
          bstore(b0, rppn, vn0, sv);
          if (!first) bstore(r.e0 + ap * r.c0, rxs, vn0, sv);
          if (has1) {
            bstore(a1, rpn, vn1, sv);
            bstore(b1, rppn, vn1, sv);
            if (!first) bstore(r.e1 + ap * r.c1, rxs, vn1, sv);
          }
        }
      }
      if (SA) {
        TIO s0, r0, s1, r1;
        stab_rp(r.a0, r.c0, r.d0, r.t0, s0, r0, a0);
        stab_rp(r.a1, r.c1, r.d1, r.t1, s1, r1, a1);
        if (gz >= k0 && gz < k1) {
          const int sv = uni(gz * plane * SI);
          bstore(a0, rpn, vn0, sv);
          bstore(r0, rrn, vn0, sv);
          if (!first) bstore(r.e0 + al * r.d0 + om * s0, rxs, vn0, sv);
          if (has1) {
            bstore(a1, rpn, vn1, sv);
            bstore(r1, rrn, vn1, sv);
            if (!first) bstore(r.e1 + al * r.d1 + om * s1, rxs, vn1, sv);
          }
        }
        prr[q0] = r.m0 ? T(0) : T(r0);
        if (has1) prr[q1] = r.m1 ? T(0) : T(r1);
      }
      if (SB) {
        a0 = r.a0 - al * r.c0;
        a1 = r.a1 - al * r.c1;
      }
      p1[q0] = r.m0 ? T(0) : T(a0);
      if (DUAL) p2[q0] = r.m0 ? T(0) : T(b0);
      pk[q0] = r.m0;
      if (has1) {
        p1[q1] = r.m1 ? T(0) : T(a1);
        if (DUAL) p2[q1] = r.m1 ? T(0) : T(b1);
        pk[q1] = r.m1;
      }
    };
    // fields of element layer kz (zero outside the grid) and its z size
    auto coef_fetch = [&](int kz, Coef<T> &c) {
      const int kc = min(max(kz, 0), ez - 1);
      c.hz = op.hz[kc];
      c.ihz = op.ihz[kc];
      const bool live = kz == kc && (!SLAB || (kz >= k0 && kz < k1 - 1));
      load_rec<T>(rrc, vr | uni(live ? 0 : kOob), uni(kc * eplane * SR), c);
    };

    // dot terms (ND sums): each thread sums its owned nodes of the current
    // unit (its planes in order); at the unit's last plane the block sums go
    // to red[.][unit & 1] and thread 0 stores them after the next barrier
    double dacc[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) dacc[i] = 0.0;
    int pend = -1;  // unit whose block sum waits for the next barrier
    // a z-chunk launch stores one partial per group: the per-thread sum over
    // the march is reduced once at its end
    auto dot_emit = [&](const double (&dv)[ND], int k) {
#pragma unroll
      for (int i = 0; i < ND; ++i) dacc[i] += dv[i];
      if (gidx < 0 && (k % kUnitPlanes == kUnitPlanes - 1 || k == naz - 1)) {
#pragma unroll
        for (int i = 0; i < ND; ++i) {
          const double v = wave_sum_lane63(dacc[i]);
          if ((tid & 63) == 63) red[i][(k / kUnitPlanes) & 1][tid >> 6] = v;
          dacc[i] = 0.0;
        }
        pend = k / kUnitPlanes;
      }
    };
    // every thread reads the wave sums (LDS broadcast, no branch around the
    // reads, so they issue with the layer's other LDS reads); thread 0 stores
    auto dot_flush = [&]() {
      if (gidx >= 0) return;
      const int pb = pend < 0 ? 0 : pend;
      double t[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        t[i] = 0.0;
        for (int w = 0; w < NW; ++w) t[i] += red[i][pb & 1][w];
      }
      if (pend >= 0 && tid == 0) {
#pragma unroll
        for (int i = 0; i < ND; ++i) dm[i][pend] = t[i];
      }
      pend = -1;
    };

    // prologue: the bottom face of layer k0-1 (plane k0-1) straight from
    // memory, plane k0 and the fields of layer k0-1 ahead; the loads are
    // issued before the status test so that they overlap its reductions
    TIO ca[4], cb[4], cc[4], cd[4], ct[4];
    uint8_t cm[4];
    {
      const int gz = max(k0 - 1, 0);
      const int sm = uni(gz * plane), sv = uni(gz * plane * SI);
      const int oz = uni(!SLAB && k0 >= 1 ? 0 : kOob);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = colofs(exg + (c == 1 || c == 2), eyg + (c >= 2));
        const int vc = (col >= 0 ? col : kOob) | oz, vcs = (col >= 0 ? col * SI : kOob) | oz;
        cm[c] = bload<uint8_t>(rmk, vc, sm);
        ca[c] = bload<TIO>(rx1, vcs, sv);
        cb[c] = DUAL ? bload<TIO>(rx2, vcs, sv) : TIO(0);
        cc[c] = (DIR || STAB) ? bload<TIO>(rpo, vcs, sv) : TIO(0);
        cd[c] = (DIR || SA) ? bload<TIO>(rppo, vcs, sv) : TIO(0);
        ct[c] = SA ? bload<TIO>(rto, vcs, sv) : TIO(0);
      }
    }
    PlanePair<TIO, DIR> pr;
    plane_fetch(k0, pr);
    Coef<T> cn, cf;
    coef_fetch(k0 - 1, cn);

    if (DIR) {
      // the status of the previous iteration, read once per block and
      // broadcast (one block of this launch writes its err and pend; done,
      // iter and alpha change only in the update launches)
      if (tid == 0) {
        sflag[0] = dd.stv[m].done;
        sflag[1] = dd.stv[m].iter;
        salpha = dd.stv[m].alpha;
      }
      const Segment sg = BATCH ? md[m].seg : m0.seg;
      const int q = dd.q;
      const double *const parts[4] = {dd.pr + (q ^ 1) * dd.NB + sg.vb0, dd.pbn + sg.vb0,
                                      dd.pb + (q ^ 1) * dd.NB + sg.vb0,
                                      dd.pb + q * dd.NB + sg.vb0};
      double sums[4];
      const int ns[4] = {sg.nb, sg.nb, sg.nb, sg.nb};
      blk_sum4(parts, ns, sums);  // its barriers also publish sflag
      const int pdone = sflag[0], k = sflag[1] + 1;
      ap = TIO(salpha);
      __syncthreads();  // sflag is rewritten by the block's next march
      if (pdone) return;
      const double rn2 = sums[0], bn2 = sums[1];
      const double err = sqrt(rn2 / bn2);
      int done = 0;
      if (err <= dd.tol) done = 1;
      else if (k > dd.itmax) done = 2;
      else if (!(err == err)) done = 3;
      else if (err > dd.divtol) done = 4;
      double beta = 0.0;
      if (!done && k > 1) {
        const double bknum = sums[2], bkden = sums[3];
        beta = bknum / bkden;
        if (!(beta == beta) || bkden == 0.0) done = 3;
      }
      if (tile == 0 && k0 == 0 && tid == 0) {  // one block per system
        dd.stv[m].err = err;
        dd.stv[m].bnrm2 = bn2;
        if (done) dd.stv[m].pend = done;
      }
      first = k == 1;
      if (done) {
        // the solve ends here: complete x with the last update's alpha po
        // on the nodes this march owns, then stop
        if (!first) {
          for (int gz = k0; gz < k1; ++gz) {
            const int sv = uni(gz * plane * SI);
            const TIO x0 = bload<TIO>(rxs, vn0, sv), p0 = bload<TIO>(rpo, vn0, sv);
            bstore(x0 + ap * p0, rxs, vn0, sv);
            if (has1) {
              const TIO x1v = bload<TIO>(rxs, vn1, sv), p1v = bload<TIO>(rpo, vn1, sv);
              bstore(x1v + ap * p1v, rxs, vn1, sv);
            }
          }
        }
        return;
      }
      bt = TIO(beta);
    }
    if (STAB) {
      // the status and the partials of the previous launch, read once per
      // block and broadcast.  Each launch writes only fields the other one
      // reads (kStabDir: rho_next, omega, pad, pend; kStabS: rho, alpha,
      // iter, err, done), so no block sees a field change that it reads.
      // BATCH: member m's status, and its ranges of the partial arrays: the
      // owner partials at its units / groups (from seg.u0), the vector
      // workgroup partials (|b|^2, r0 . r0) at its workgroups (from seg.vb0)
      SolveStatus *stm = BATCH ? sd.st + m : sd.st;
      if (tid == 0) {
        sflag[0] = stm->done;
        sflag[1] = SA ? stm->iter : stm->pad;  // index j of this iteration
        sflag[2] = stm->pend;
        sstab[0] = SA ? stm->rho : stm->rho_next;
        sstab[1] = stm->alpha;
      }
      const bool writer = tile == 0 && k0 == 0 && tid == 0;  // one block
      double sums[4];
      if (BATCH) {
        const Segment sg = md[m].seg;
        const int no = (int)(zc > 0 ? sg.ng : sg.nu);
        const double *pin[4];
        int nin[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const bool vec = SB && i == 2;  // |b|^2: vector-workgroup partials
          pin[i] = sd.pin[i] ? sd.pin[i] + (vec ? (int64_t)sg.vb0 : sg.u0) : nullptr;
          nin[i] = sd.pin[i] ? (vec ? sg.nb : no) : 0;
        }
        blk_sum4(pin, nin, sums);  // its barriers also publish sflag
      } else {
        blk_sum4(sd.pin, sd.nin, sums);
      }
      const int pdone = sflag[0], j = sflag[1], ppend = sflag[2];
      const double rho_a = sstab[0], al_a = sstab[1];
      __syncthreads();  // sflag is rewritten by the block's next march
      if (pdone) return;
      if (SA) {
        // sums: t.s, t.t, r~.s, r~.t of the previous S launch
        first = j == 0;
        double rho = 0.0, omega = 0.0, beta = 0.0;
        if (first) {
          // rho_0 = r0 . r0 from the initial partials
          double s0[4];
          const double *p0b = BATCH ? sd.prt0 + md[m].seg.vb0 : sd.prt0;
          const double *const p0[4] = {p0b, p0b, p0b, p0b};
          const int n0[4] = {BATCH ? md[m].seg.nb : sd.nb0, 0, 0, 0};
          blk_sum4(p0, n0, s0);
          rho = s0[0];
        } else {
          // t . t = 0: the previous step's S is zero, it solved the system
          // exactly (t = D^-1 A S vanishes only with S for a regular A):
          // omega = beta = 0, so R = S, x gains alpha P only, and the S
          // launch's convergence test ends the solve; were A singular on S,
          // that launch reports the breakdown (r~ . V = 0)
          const bool sz = sums[1] == 0.0;
          omega = sz ? 0.0 : sums[0] / sums[1];
          rho = sums[2] - omega * sums[3];
          beta = sz ? 0.0 : (rho / rho_a) * (al_a / omega);
          if (!sz && (!(beta == beta) || !(omega == omega) || omega == 0.0 || rho == 0.0)) {
            if (writer) stm->pend = 3;
            return;
          }
        }
        if (writer) {
          stm->rho_next = rho;
          stm->omega = omega;
          stm->pad = j;
        }
        bt = TIO(beta);
        om = TIO(omega);
        al = TIO(al_a);
      } else {
        // sums: RT.V and |r_j|^2 of the kStabDir launch, |b|^2
        if (ppend) {  // the kStabDir launch found a breakdown
          if (writer) stm->done = ppend;
          return;
        }
        const double err = sqrt(sums[1] / sums[2]);
        const double gam = sums[0];
        const double alpha = rho_a / gam;
        int done = 0;
        if (err <= sd.tol) done = 1;
        else if (j >= sd.itmax) done = 2;
        else if (!(err == err)) done = 3;
        else if (err > sd.divtol) done = 4;
        else if (!(alpha == alpha) || gam == 0.0) done = 3;
        if (writer) {
          stm->err = err;
          stm->bnrm2 = sums[2];
          stm->iter = done ? j : j + 1;
          if (done) {
            stm->done = done;
          } else {
            stm->rho = rho_a;
            stm->alpha = alpha;
          }
        }
        if (done) return;
        al = TIO(alpha);
      }
    }
    T fb1[4], fb2[4];
    {
      T n1[4], n2[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        TIO a = DIR ? dir(ca[c], cc[c]) : ca[c];
        if (SA) {
          TIO sv_, rv_;
          stab_rp(ca[c], cc[c], cd[c], ct[c], sv_, rv_, a);
        }
        if (SB) a = ca[c] - al * cc[c];
        const TIO b = DIR ? dir(cb[c], cd[c]) : cb[c];
        n1[c] = cm[c] ? T(0) : T(a);
        n2[c] = cm[c] ? T(0) : T(b);
      }
      face_fwd<T>(n1, fb1);
      if (DUAL) face_fwd<T>(n2, fb2);
    }

    T ot1[4], ot2[4];  // Haar-space top-face outputs of the previous layer
#pragma unroll
    for (int c = 0; c < 4; ++c) ot1[c] = ot2[c] = T(0);
    uint8_t omk[2] = {1, 1};    // flag of the owned node of the plane read 2 layers ago
    double xok[2] = {0.0, 0.0};  // its x2 value (for the dot; kStabS: its S;
                                 // kStabDir: its new R)
    TIO fok[2] = {TIO(0), TIO(0)}, gok[2] = {TIO(0), TIO(0)};  // BiCGSTAB: its d, RT
    int kl = k0 - 1;

    // output of plane kl-1 (summed from the stage written in the previous
    // layer, buffer Q) with the flag and x2 value read two layers ago
    auto emit_plane = [&](auto Qc, auto Pc) {
      constexpr int Q = decltype(Qc)::value, P = decltype(Pc)::value;
      double v1 = 0.0, v2 = 0.0;
      if (own) {
        const T *s = st1[Q];
        v1 = (double)s[2 * NT + tid] + (double)s[3 * NT + tid + 1] +
             (double)s[1 * NT + tid + TX] + (double)s[tid + TX + 1];
        if (DUAL) {
          const T *s2 = st2[DUAL ? Q : 0];
          v2 = (double)s2[2 * NT + tid] + (double)s2[3 * NT + tid + 1] +
               (double)s2[1 * NT + tid + TX] + (double)s2[tid + TX + 1];
        }
      }
      const int so = uni((kl - 1) * plane * SI);
      // BiCGSTAB: V = (A P)/d, T = (A S)/d (d is 0 only where the node is
      // not stored)
      const double dg = (double)fok[P];
      TIO o1 = omk[P] ? TIO(0) : STAB ? TIO(dg != 0.0 ? v1 / dg : 0.0) : TIO(v1);
      if (JAC)  // the input x of the node is its masked LDS value
        o1 = omk[P] ? TIO(0)
                    : TIO(xok[P] + sd.jomega * (((double)gok[P] - v1) / (dg != 0.0 ? dg : 1.0)));
      TIO o2 = omk[P] ? TIO(0) : TIO(v2);
      if (SLAB && slab_color == 1 && !omk[P] &&
          (kl - 1 == k0 || (kl - 1 == k1 - 1 && kl - 1 < naz - 1))) {
        // colour 1: a slab face plane adds to the colour-0 slab's sums
        o1 = TIO((double)bload<TIO>(ry1, vo, so) + v1);
        if (DUAL) o2 = TIO((double)bload<TIO>(ry2, vo, so) + v2);
      }
      bstore(o1, ry1, vo, so);
      double dv[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i) dv[i] = 0.0;
      if (DUAL) {
        bstore(o2, ry2, vo, so);
        dv[0] = (double)o1 * xok[P];
      }
      const double rtv = (double)gok[P];
      if (SA) {
        dv[0] = (double)o1 * rtv;                 // RT . V
        const double rn = dg * xok[P];
        dv[ND > 1 ? 1 : 0] = rn * rn;             // |r|^2, r = d R
      }
      if (SB) {
        const double tv = dg * (double)o1, sv = dg * xok[P];  // t, s at the node
        dv[0] = tv * sv;                          // t . s
        dv[ND > 1 ? 1 : 0] = tv * tv;             // t . t
        dv[ND > 2 ? 2 : 0] = rtv * xok[P];        // r~ . s = RT . S
        dv[ND > 3 ? 3 : 0] = rtv * (double)o1;    // r~ . t = RT . T
      }
      if (want_dot) dot_emit(dv, kl - 1);
    };

    // one element layer kl with plane kl+1 in buffer P
    auto layer = [&](auto Pc) {
      constexpr int P = decltype(Pc)::value;
      plane_put(pr, kl + 1, pl1[P], DUAL ? pl2[DUAL ? P : 0] : nullptr, pm[P],
                plR[SA ? P : 0]);
      const TIO fon = pr.f, gon = pr.g;  // owned node of plane kl+1
      cf = cn;
      plane_fetch(kl + 2, pr);  // prefetch plane kl+2 and layer kl+1
      coef_fetch(kl + 1, cn);
      lds_barrier();
      if (want_dot) dot_flush();
      // top face of layer kl
      T xt[4], ft1[4], ft2[4];
      xt[0] = pl1[P][c00];
      xt[1] = pl1[P][c00 + 1];
      xt[2] = pl1[P][c00 + TX + 2];
      xt[3] = pl1[P][c00 + TX + 1];
      face_fwd<T>(xt, ft1);
      if (DUAL) {
        const T *q = pl2[DUAL ? P : 0];
        xt[0] = q[c00];
        xt[1] = q[c00 + 1];
        xt[2] = q[c00 + TX + 2];
        xt[3] = q[c00 + TX + 1];
        face_fwd<T>(xt, ft2);
      }
      const uint8_t omn = pm[P][oq];
      const double xon = (DUAL && own) ? (double)pl2[DUAL ? P : 0][oq]
                         : ((SB || JAC) && own) ? (double)pl1[P][oq]
                         : (SA && own) ? (double)plR[SA ? P : 0][oq] : 0.0;
      // owner sums and output of plane kl-1
      if (kl - 1 >= k0) emit_plane(std::integral_constant<int, P ^ 1>(), Pc);
      omk[P] = omn;
      xok[P] = xon;
      fok[P] = fon;
      gok[P] = gon;
      // element layer kl; a layer outside the grid has zero fields and a
      // zero mass weight, so its result is zero
      const bool lv = kl >= 0 && kl < ez && (!SLAB || (kl >= k0 && kl < k1 - 1));
      T ob1[4], on1[4], ob2[4], on2[4];
      {
        ElemConst<T> k;
        elem_layer_setup<T>(cq, cf.hz, cf.ihz, cf.lam, cf.vx, cf.vy, cf.vz, cf.tau,
                            T(op.cH), lv ? mu_c : T(0), k);
        T xh[8], yh[8];
        zfwd<T>(fb1, ft1, xh);
        haar_apply<T, false>(k, xh, yh);
        if (ROBIN && robin && lv)
          robin_haar<T, false>(k, cq.ihx, cf.tau, cf.vx, cf.vy, cf.vz, T(op.ra), xh, yh);
        zinv<T>(yh, ob1, on1);
        if (DUAL) {
          zfwd<T>(fb2, ft2, xh);
          haar_apply_tr<T>(k, xh, yh);
          if (ROBIN && robin && lv)
            robin_haar<T, true>(k, cq.ihx, cf.tau, cf.vx, cf.vy, cf.vz, T(op.ra), xh, yh);
          zinv<T>(yh, ob2, on2);
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) ob2[c] = on2[c] = T(0);
        }
      }
      // column contributions to plane kl
      if (kl >= k0) {
        T s[4], y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) s[c] = ot1[c] + ob1[c];
        face_inv<T>(s, y);
#pragma unroll
        for (int c = 0; c < 4; ++c) st1[P][c * NT + tid] = y[c];
        if (DUAL) {
#pragma unroll
          for (int c = 0; c < 4; ++c) s[c] = ot2[c] + ob2[c];
          face_inv<T>(s, y);
#pragma unroll
          for (int c = 0; c < 4; ++c) st2[DUAL ? P : 0][c * NT + tid] = y[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        fb1[c] = ft1[c];
        ot1[c] = on1[c];
        if (DUAL) {
          fb2[c] = ft2[c];
          ot2[c] = on2[c];
        }
      }
    };

    // march: layers k0-1 .. k1-1 (at least two), then plane k1-1
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    for (;;) {
      layer(I0());
      if (++kl >= k1) {  // plane k1-1 was staged in buffer 0
        lds_barrier();
        if (want_dot) dot_flush();
        emit_plane(I0(), I1());
        break;
      }
      layer(I1());
      if (++kl >= k1) {
        lds_barrier();
        if (want_dot) dot_flush();
        emit_plane(I1(), I0());
        break;
      }
    }
    lds_barrier();  // buffers are refilled by the next march
    if (want_dot) {
      if (gidx >= 0) {
#pragma unroll
        for (int i = 0; i < ND; ++i) {
          const double v = wave_sum_lane63(dacc[i]);
          if ((tid & 63) == 63) red[i][0][tid >> 6] = v;
        }
        lds_barrier();
        if (tid == 0) {
#pragma unroll
          for (int i = 0; i < ND; ++i) {
            double t = 0.0;
            for (int w = 0; w < NW; ++w) t += red[i][0][w];
            dm[i][0] = t;
          }
        }
      } else {
        dot_flush();
      }
    }
  };

  if (zc > 0) {
    // z-chunk launch: workgroup lb takes group lb, the (chunk, tile) groups
    // numbered chunk-major per system, so the workgroups of one XCD march
    // neighbouring tiles over the same planes at the same time and share
    // their halo in L2
    const int ia = BATCH ? find(actg, lb) : 0;
    const int m = BATCH ? act[ia] : 0;
    const Grid3 g = BATCH ? md[m].op.g : m0.op.g;
    const int nt = ((g.nax + TX - 2) / (TX - 1)) * ((g.nay + TY - 2) / (TY - 1));
    const int64_t lg = lb - (BATCH ? actg[ia] : (int64_t)0);
    const int chunk = (int)(lg / nt), tile = (int)(lg - (int64_t)chunk * nt);
    if (SLAB) {  // slab 2 chunk + colour: element layers [e0, e1), planes e0..e1
      const int e0 = (2 * chunk + slab_color) * zc;
      if (e0 < g.ez) march(m, tile, e0, min(g.ez, e0 + zc) + 1, lg);
      return;
    }
    const int k0 = chunk * zc;
    if (k0 < g.naz) march(m, tile, k0, min(g.naz, k0 + zc), lg);
    return;
  }
  // balanced launch: block lb takes the units [lb U / G, (lb+1) U / G),
  // split into one march per (system, tile) it touches
  int64_t u = U * lb / G;
  const int64_t uend = U * (lb + 1) / G;
  int ia = BATCH ? find(actu, u) : 0;
  while (u < uend) {
    if (BATCH)
      while (u >= actu[ia + 1]) ++ia;
    const int m = BATCH ? act[ia] : 0;
    const int naz = BATCH ? md[m].op.g.naz : m0.op.g.naz;
    const int ngz = (naz + kUnitPlanes - 1) / kUnitPlanes;
    const int64_t lu = u - (BATCH ? actu[ia] : (int64_t)0);
    const int tile = (int)(lu / ngz);
    const int g0 = (int)(lu - (int64_t)tile * ngz);
    const int g1 = (int)min<int64_t>((int64_t)ngz, g0 + (uend - u));
    u += g1 - g0;
    march(m, tile, g0 * kUnitPlanes, min(naz, g1 * kUnitPlanes), -1);
  }
}

// Element apply of the brick kernels: corners gathered from the two LDS
// planes, the 8 nodal results (8 transposed results when DUAL) written to the
// staging arrays, corner-major.
template <typename T, bool DUAL, int TX, int TY>
__device__ __forceinline__ void elem_stage(const OpData<T> &op, bool valid,
                                           bool robin, const ColConst<T> &q,
                                           const Coef<T> &cf, const T *pb,
                                           const T *pt, const T *pb2,
                                           const T *pt2, int tx, int ty,
                                           T *stage, T *stage2, int tid) {
  const int c00 = ty * (TX + 1) + tx;
  const int off[4] = {0, 1, TX + 2, TX + 1};
  T xa[8], xc[8], ya[8], yc[8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    xa[c] = pb[c00 + off[c]];
    xa[c + 4] = pt[c00 + off[c]];
    xc[c] = DUAL ? pb2[c00 + off[c]] : T(0);
    xc[c + 4] = DUAL ? pt2[c00 + off[c]] : T(0);
  }
  elem_eval<T, DUAL, true>(op, valid, robin, q, cf, xa, xc, ya, yc);
#pragma unroll
  for (int c = 0; c < 8; ++c) stage[c * (TX * TY) + tid] = ya[c];
  if (DUAL) {
#pragma unroll
    for (int c = 0; c < 8; ++c) stage2[c * (TX * TY) + tid] = yc[c];
  }
}

// ---------------------------------------------------------------------------
// ATOMIC kernel: disjoint bricks of TX x TY element columns x zc layers (one
// launch, XCD-aware brick order).  Nodes strictly inside the brick (x and y)
// of planes strictly inside it are stored; all other brick nodes are added
// with hardware atomics.  The output must be zero on entry.
// ---------------------------------------------------------------------------
template <typename T, typename TIO, bool DUAL, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_brick(OpData<T> op, int zc,
                                                const TIO *__restrict__ x1,
                                                TIO *__restrict__ y1,
                                                const TIO *__restrict__ x2,
                                                TIO *__restrict__ y2) {
  __shared__ T pl1[2][((TX + 1) * (TY + 1))];
  __shared__ T pl2[DUAL ? 2 : 1][DUAL ? ((TX + 1) * (TY + 1)) : 1];
  __shared__ uint8_t pm[2][((TX + 1) * (TY + 1))];
  __shared__ T st1[8 * (TX * TY)];
  __shared__ T st2[DUAL ? 8 * (TX * TY) : 1];

  if (op.done && *op.done) return;
  const int ntbx = (op.g.ex + TX - 1) / TX, ntby = (op.g.ey + TY - 1) / TY;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = t % ntbx, by = (t / ntbx) % ntby, bz = t / (ntbx * ntby);
  const int kz0 = bz * zc;
  if (bx * TX >= op.g.ex || by * TY >= op.g.ey || kz0 >= op.g.ez) return;
  const int kz1 = min(kz0 + zc, op.g.ez);  // layers [kz0, kz1)

  const int tid = threadIdx.x;
  const int tx = tid & (TX - 1), ty = tid / TX;
  const int gx0 = bx * TX, gy0 = by * TY;
  const int exg = gx0 + tx, eyg = gy0 + ty;
  const bool colok = exg < op.g.ex && eyg < op.g.ey;
  ColConst<T> cq;
  col_setup<T>(colok ? op.hx[exg] : T(1), colok ? op.hy[eyg] : T(1),
               colok ? op.ihx[exg] : T(1), colok ? op.ihy[eyg] : T(1), cq);
  const bool robin = colok && exg == 0 && op.ra != 0.0;

  const int q0 = tid, q1 = tid + (TX * TY);
  const int lx0 = q0 % (TX + 1), ly0 = q0 / (TX + 1);
  const int lx1 = q1 % (TX + 1), ly1 = q1 / (TX + 1);
  const bool has1 = q1 < ((TX + 1) * (TY + 1));

  auto eidx = [&](int k) {
    return exg + (int64_t)op.g.ex * (eyg + (int64_t)op.g.ey * k);
  };
  // tile node (lx,ly) of plane kz: interior nodes are plain stores
  auto emit = [&](TIO *y, int lx, int ly, int kz, double v, bool masked,
                  bool complete_z) {
    const int gx = gx0 + lx, gy = gy0 + ly;
    if (masked || gx >= op.g.nax || gy >= op.g.nay || kz >= op.g.naz) return;
    const int64_t n = gx + (int64_t)op.g.nax * (gy + (int64_t)op.g.nay * kz);
    const bool inner = complete_z && lx >= 1 && lx < TX && ly >= 1 && ly < TY;
    if (inner)
      y[n] = TIO(v);
    else
      atomicAdd(&y[n], TIO(v));
  };

  PlaneRegs<TIO, DUAL> pr;
  plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kz0, lx0, ly0, lx1, ly1, has1, pr);
  plane_store<T, TIO, DUAL>(pr, pl1[0], DUAL ? pl2[0] : nullptr, pm[0], q0, q1, has1);
  plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kz0 + 1, lx0, ly0, lx1, ly1, has1, pr);
  Coef<T> cn = load_coef<T>(op, colok, eidx(kz0), kz0);
  double a1[2] = {0.0, 0.0}, a2[2] = {0.0, 0.0};
  for (int kl = kz0; kl < kz1; ++kl) {
    const int cb = (kl - kz0) & 1, ct = cb ^ 1;
    plane_store<T, TIO, DUAL>(pr, pl1[ct], DUAL ? pl2[ct] : nullptr, pm[ct], q0, q1, has1);
    const Coef<T> cf = cn;
    if (kl + 1 < kz1) {
      plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kl + 2, lx0, ly0, lx1, ly1,
                               has1, pr);
      cn = load_coef<T>(op, colok, eidx(kl + 1), kl + 1);
    }
    lds_barrier();
    const uint8_t mk0 = pm[cb][q0], mk1 = has1 ? pm[cb][q1] : uint8_t(1);
    elem_stage<T, DUAL, TX, TY>(op, colok, robin, cq, cf, pl1[cb], pl1[ct],
                       DUAL ? pl2[cb] : nullptr, DUAL ? pl2[ct] : nullptr, tx,
                       ty, st1, DUAL ? st2 : nullptr, tid);
    lds_barrier();
    const bool zin = kl > kz0;  // plane kl has both layers inside the chunk
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h && !has1) break;
      const int lx = h ? lx1 : lx0, ly = h ? ly1 : ly0;
      const bool mk = h ? mk1 : mk0;
      emit(y1, lx, ly, kl, a1[h] + node_sum<T, TX, TY>(st1, lx, ly, false), mk, zin);
      a1[h] = node_sum<T, TX, TY>(st1, lx, ly, true);
      if (DUAL) {
        emit(y2, lx, ly, kl, a2[h] + node_sum<T, TX, TY>(st2, lx, ly, false), mk, zin);
        a2[h] = node_sum<T, TX, TY>(st2, lx, ly, true);
      }
    }
  }
  // top plane of the chunk (shared with the next chunk); its flags are in the
  // buffer written last, which no thread refills
  const int cl = (kz1 - kz0) & 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h && !has1) break;
    const int lx = h ? lx1 : lx0, ly = h ? ly1 : ly0;
    const bool mk = pm[cl][h ? q1 : q0] != 0;
    emit(y1, lx, ly, kz1, a1[h], mk, false);
    if (DUAL) emit(y2, lx, ly, kz1, a2[h], mk, false);
  }
}

// ---------------------------------------------------------------------------
// COLORED kernel: two colours of z-slabs of zc element layers (colour = slab
// index parity), x-y tiles of TX x TY element columns overlapping by one
// column, every node column owned by one tile (the owner kernel's x-y
// split: a tile's owned nodes see all four in-plane elements).  A workgroup
// marches one (tile, slab): the owned nodes of the planes inside the slab get
// complete sums and plain stores; the slab's bottom and top planes get its
// partial sums, which colour 0 stores and colour 1 adds to what colour 0
// stored (its bottom plane always, its top plane unless it is the grid's
// top).  Slabs of one colour share no node, the colour-0 launch runs before
// the colour-1 launch, so every node's sum has a fixed order: deterministic,
// no atomics, no zeroing pass.
// ---------------------------------------------------------------------------
template <typename T, typename TIO, bool DUAL, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_slab(OpData<T> op, int zc, int color,
                                               const TIO *__restrict__ x1,
                                               TIO *__restrict__ y1,
                                               const TIO *__restrict__ x2,
                                               TIO *__restrict__ y2) {
  constexpr int NP = (TX + 1) * (TY + 1), NT = TX * TY;
  __shared__ T pl1[2][NP];
  __shared__ T pl2[DUAL ? 2 : 1][DUAL ? NP : 1];
  __shared__ uint8_t pm[2][NP];
  __shared__ T st1[8 * NT];
  __shared__ T st2[DUAL ? 8 * NT : 1];

  if (op.done && *op.done) return;
  const int ntx = (op.g.nax + TX - 2) / (TX - 1), nty = (op.g.nay + TY - 2) / (TY - 1);
  const int ntiles = ntx * nty;
  // XCD-aware: each XCD takes a contiguous range of tiles of one slab row
  const int lin = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = lin % ntiles, bz = 2 * (lin / ntiles) + color;
  const int kz0 = bz * zc;
  if (kz0 >= op.g.ez) return;
  const int kz1 = min(kz0 + zc, op.g.ez);  // element layers [kz0, kz1)

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int gx0 = (tile % ntx) * (TX - 1) - 1, gy0 = (tile / ntx) * (TY - 1) - 1;
  const int exg = gx0 + tx, eyg = gy0 + ty;
  const bool colok = exg >= 0 && eyg >= 0 && exg < op.g.ex && eyg < op.g.ey;
  ColConst<T> cq;
  col_setup<T>(colok ? op.hx[exg] : T(1), colok ? op.hy[eyg] : T(1),
               colok ? op.ihx[exg] : T(1), colok ? op.ihy[eyg] : T(1), cq);
  const bool robin = colok && exg == 0 && op.ra != 0.0;

  const int q0 = tid, q1 = tid + NT;
  const int lx0 = q0 % (TX + 1), ly0 = q0 / (TX + 1);
  const int lx1 = q1 % (TX + 1), ly1 = q1 / (TX + 1);
  const bool has1 = q1 < NP;

  auto eidx = [&](int k) {
    return exg + (int64_t)op.g.ex * (eyg + (int64_t)op.g.ey * k);
  };
  // owned node (lx, ly) of plane kz: store v (add to the stored value when
  // add); Dirichlet nodes get 0
  auto emit = [&](TIO *y, int lx, int ly, int kz, double v, bool masked, bool add) {
    const int gx = gx0 + lx, gy = gy0 + ly;
    if (lx < 1 || ly < 1 || lx > TX - 1 || ly > TY - 1) return;
    if (gx >= op.g.nax || gy >= op.g.nay || kz >= op.g.naz) return;
    const int64_t n = gx + (int64_t)op.g.nax * (gy + (int64_t)op.g.nay * kz);
    if (masked)
      y[n] = TIO(0);
    else if (add)
      y[n] = TIO((double)y[n] + v);
    else
      y[n] = TIO(v);
  };

  PlaneRegs<TIO, DUAL> pr;
  plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kz0, lx0, ly0, lx1, ly1, has1, pr);
  plane_store<T, TIO, DUAL>(pr, pl1[0], DUAL ? pl2[0] : nullptr, pm[0], q0, q1, has1);
  plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kz0 + 1, lx0, ly0, lx1, ly1, has1, pr);
  Coef<T> cn = load_coef<T>(op, colok, eidx(kz0), kz0);
  double a1[2] = {0.0, 0.0}, a2[2] = {0.0, 0.0};
  for (int kl = kz0; kl < kz1; ++kl) {
    const int cb = (kl - kz0) & 1, ct = cb ^ 1;
    plane_store<T, TIO, DUAL>(pr, pl1[ct], DUAL ? pl2[ct] : nullptr, pm[ct], q0, q1, has1);
    const Coef<T> cf = cn;
    if (kl + 1 < kz1) {
      plane_load<T, TIO, DUAL>(op, x1, x2, gx0, gy0, kl + 2, lx0, ly0, lx1, ly1, has1, pr);
      cn = load_coef<T>(op, colok, eidx(kl + 1), kl + 1);
    }
    lds_barrier();
    const uint8_t mk0 = pm[cb][q0], mk1 = has1 ? pm[cb][q1] : uint8_t(1);
    elem_stage<T, DUAL, TX, TY>(op, colok, robin, cq, cf, pl1[cb], pl1[ct],
                                DUAL ? pl2[cb] : nullptr, DUAL ? pl2[ct] : nullptr, tx, ty,
                                st1, DUAL ? st2 : nullptr, tid);
    lds_barrier();
    // plane kl: complete inside the slab, partial at its bottom
    const bool add = kl == kz0 && color == 1;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h && !has1) break;
      const int lx = h ? lx1 : lx0, ly = h ? ly1 : ly0;
      const bool mk = h ? mk1 : mk0;
      emit(y1, lx, ly, kl, a1[h] + node_sum<T, TX, TY>(st1, lx, ly, false), mk, add);
      a1[h] = node_sum<T, TX, TY>(st1, lx, ly, true);
      if (DUAL) {
        emit(y2, lx, ly, kl, a2[h] + node_sum<T, TX, TY>(st2, lx, ly, false), mk, add);
        a2[h] = node_sum<T, TX, TY>(st2, lx, ly, true);
      }
    }
  }
  // top plane of the slab (its flags are in the buffer written last)
  const int cl = (kz1 - kz0) & 1;
  const bool add = color == 1 && kz1 < op.g.ez;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h && !has1) break;
    const int lx = h ? lx1 : lx0, ly = h ? ly1 : ly0;
    const bool mk = pm[cl][h ? q1 : q0] != 0;
    emit(y1, lx, ly, kz1, a1[h], mk, add);
    if (DUAL) emit(y2, lx, ly, kz1, a2[h], mk, add);
  }
}

template <typename TIO>
__global__ void k_zero2(const int *done, TIO *a, TIO *b, int64_t n) {
  if (done && *done) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i] = TIO(0);
    if (b) b[i] = TIO(0);
  }
}

}  // namespace

// Owner tile shapes (element columns per workgroup): 0 = 32x8, 1 = 16x16,
// 2 = 64x4; each owns (TX-1) x (TY-1) node columns.
static void owner_shape(int tile, int *tx, int *ty) {
  switch (tile) {
    case 1: *tx = 16; *ty = 16; break;
    case 2: *tx = 64; *ty = 4; break;
    default: *tx = 32; *ty = 8; break;
  }
}

int64_t owner_units(const Grid3 &g, int tile) {
  int tx, ty;
  owner_shape(tile, &tx, &ty);
  return (int64_t)((g.nax + tx - 2) / (tx - 1)) * ((g.nay + ty - 2) / (ty - 1)) *
         ((g.naz + kUnitPlanes - 1) / kUnitPlanes);
}

int64_t owner_groups(const Grid3 &g, int tile, int zchunk) {
  if (zchunk <= 0) return 0;
  int tx, ty;
  owner_shape(tile, &tx, &ty);
  return (int64_t)((g.nax + tx - 2) / (tx - 1)) * ((g.nay + ty - 2) / (ty - 1)) *
         ((g.naz + zchunk - 1) / zchunk);
}

bool owner_fits(const Grid3 &g) {
  const int64_t lim = (int64_t)1 << 31;
  return 8 * g.N < lim && 40 * g.E < lim;
}

namespace {
template <typename T, typename TIO, bool DUAL, bool ROBIN, bool BATCH, int DM, int TX, int TY>
const void *owner_kernel() {
  return reinterpret_cast<const void *>(&k_owner<T, TIO, DUAL, ROBIN, BATCH, DM, TX, TY>);
}

// Resident workgroups of a kernel on the whole device (cached per kernel).
int resident_blocks(const void *fn, int threads) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({dev, fn});
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int r = std::max(1, per_cu) * std::max(1, cus);
  cache[{dev, fn}] = r;
  return r;
}

// Compute units of the current device (cached per device).
int device_cus() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  cache[dev] = cus;
  return cus;
}

// Element layers per brick of the ATOMIC launch when the caller leaves the
// choice to the library: the deepest brick (at most 8 layers) whose launch
// still gives every compute unit two workgroups, else 1 (at 128x128x64 with
// 32 x 8 bricks: 8 layers, 512 bricks).
int brick_depth(const Grid3 &g, int bx, int by) {
  const int64_t cols = (int64_t)((g.ex + bx - 1) / bx) * ((g.ey + by - 1) / by);
  const int64_t want = 2 * (int64_t)device_cus();
  for (int zc = 8; zc > 1; --zc)
    if (cols * ((g.ez + zc - 1) / zc) >= want) return zc;
  return 1;
}

// VGPU_COLORED_LARGE_PATH=1 sends every COLORED apply to k_slab, the path of
// systems too large for the owner kernel, so that it can be checked on
// small meshes (read at every launch).
bool colored_large_path() {
  const char *e = std::getenv("VGPU_COLORED_LARGE_PATH");
  return e && e[0] == '1';
}

// Element layers per slab of the COLORED launches when the caller leaves the
// choice to the library: the deepest slab (at most 8 layers) whose colour
// launches (each with half of the slabs) still give every compute unit two
// workgroups, else 1.
int slab_depth(const Grid3 &g, int tx, int ty) {
  const int64_t tiles = (int64_t)((g.nax + tx - 2) / (tx - 1)) * ((g.nay + ty - 2) / (ty - 1));
  const int64_t want = 2 * (int64_t)device_cus();
  for (int zc = 8; zc > 1; --zc) {
    const int64_t nsl = (g.ez + zc - 1) / zc;
    if (tiles * (nsl / 2) >= want) return zc;
  }
  return 1;
}

// Grid of an owner launch: zchunk > 0 one workgroup per (z chunk, tile)
// group (gtot of them), zchunk = 0 one workgroup per resident slot over utot
// units.
int owner_grid(const void *fn, int threads, int64_t utot, int64_t gtot, int zchunk) {
  if (zchunk > 0) return (int)std::max<int64_t>(1, gtot);
  int r = resident_blocks(fn, threads);
  if (r <= 0) r = 1024;
  return (int)std::max<int64_t>(1, std::min<int64_t>(r, utot));
}

// Everything one owner launch needs besides the template choices.
template <typename T, typename TIO>
struct OwnerLaunch {
  Member<T> m0;
  const Member<T> *md;
  ActiveSet as;
  int64_t utot, gtot;
  int zchunk;
  const SolveStatus *st;
  const TIO *x1;
  TIO *y1;
  const TIO *x2;
  TIO *y2;
  double *dot;
  DirData<TIO> dd;
  StabData<TIO> sd;
  int slab_color = 0;
};

template <typename T, typename TIO, bool DUAL, bool ROBIN, bool BATCH, int DM, int TX, int TY>
void launch_owner_k(const OwnerLaunch<T, TIO> &L, hipStream_t s) {
  const void *fn = owner_kernel<T, TIO, DUAL, ROBIN, BATCH, DM, TX, TY>();
  const int nb = owner_grid(fn, TX * TY, L.utot, L.gtot, L.zchunk);
  hipLaunchKernelGGL((k_owner<T, TIO, DUAL, ROBIN, BATCH, DM, TX, TY>), dim3(nb),
                     dim3(TX * TY), 0, s, L.m0, L.md, L.as.act, L.as.actu, L.as.actg,
                     L.as.nact, L.utot, L.zchunk > 0 ? L.zchunk : 0, L.st, L.x1, L.y1,
                     L.x2, L.y2, L.dot, L.dd, L.sd, L.slab_color);
}

// DM: 0 plain operator, 1 BiCG iteration launch (dual), kStabDir / kStabS
// BiCGSTAB launches (single)
template <typename T, typename TIO, bool BATCH, int DM, int TX, int TY>
void launch_owner_shape(const OwnerLaunch<T, TIO> &L, bool robin, hipStream_t s) {
  constexpr int DMM = DM % kBlk;
  if (L.x2) {
    if constexpr (DMM <= 1 || DMM == kSlab) {
      if (robin)
        launch_owner_k<T, TIO, true, true, BATCH, DM, TX, TY>(L, s);
      else
        launch_owner_k<T, TIO, true, false, BATCH, DM, TX, TY>(L, s);
    }
  } else if constexpr (DMM != 1) {
    if (robin)
      launch_owner_k<T, TIO, false, true, BATCH, DM, TX, TY>(L, s);
    else
      launch_owner_k<T, TIO, false, false, BATCH, DM, TX, TY>(L, s);
  }
}

// utot units (balanced launch) or gtot groups (z-chunk launch, zchunk > 0)
template <typename T, typename TIO, bool BATCH, int DM = 0>
int launch_owner_any(const OwnerLaunch<T, TIO> &L, int tile, bool robin, hipStream_t s) {
  if (L.zchunk > 0 ? L.gtot <= 0 : L.utot <= 0) return VGPU_OK;
  switch (tile) {
    case 1:
      launch_owner_shape<T, TIO, BATCH, DM, 16, 16>(L, robin, s);
      break;
    case 2:
      launch_owner_shape<T, TIO, BATCH, DM, 64, 4>(L, robin, s);
      break;
    default:
      launch_owner_shape<T, TIO, BATCH, DM, 32, 8>(L, robin, s);
      break;
  }
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}
}  // namespace

template <typename T, typename TIO>
int launch_owner_batch(const Member<T> *md, const ActiveSet &as, int tile,
                       int zchunk, bool robin, const SolveStatus *st,
                       const TIO *x1, TIO *y1, const TIO *x2, TIO *y2,
                       double *dot, hipStream_t s) {
  OwnerLaunch<T, TIO> L{};
  L.md = md;
  L.as = as;
  L.utot = as.nu;
  L.gtot = as.ng;
  L.zchunk = zchunk;
  L.st = st;
  L.x1 = x1;
  L.y1 = y1;
  L.x2 = x2;
  L.y2 = y2;
  L.dot = dot;
  return launch_owner_any<T, TIO, true>(L, tile, robin, s);
}

template int launch_owner_batch<double, double>(const Member<double> *, const ActiveSet &,
                                                int, int, bool, const SolveStatus *,
                                                const double *, double *,
                                                const double *, double *, double *,
                                                hipStream_t);
template int launch_owner_batch<float, float>(const Member<float> *, const ActiveSet &,
                                              int, int, bool, const SolveStatus *,
                                              const float *, float *, const float *,
                                              float *, double *, hipStream_t);

// Whether every vector p[0..k) of a solver launch lies in the engine's block
// [blk, blk + bytes), with len bytes from its start (0: the start only, for
// the batch vectors, which the engines carve whole from their block); the
// launch then takes the block-addressed kernel (its resource needs the block
// below 2^31 bytes).  VGPU_NO_BLOCK_LAUNCH=1 sends every solver launch to the
// form with a resource per vector, the path of engines whose block is too
// large, so that it can be checked on small meshes (read at every launch).
template <typename TIO>
bool in_block(const void *blk, int64_t bytes, const TIO *const *p, int k, int64_t len) {
  const char *e = std::getenv("VGPU_NO_BLOCK_LAUNCH");
  if (e && e[0] == '1') return false;
  if (!blk || bytes <= 0 || bytes >= ((int64_t)1 << 31)) return false;
  const char *b = static_cast<const char *>(blk);
  for (int i = 0; i < k; ++i) {
    const char *q = reinterpret_cast<const char *>(p[i]);
    if (!q || q < b || q + std::max<int64_t>(len, 1) > b + bytes) return false;
  }
  return true;
}

template <typename T, typename TIO>
int launch_owner_dir(const Member<T> &m0, const Member<T> *md, const ActiveSet &as,
                     bool batch, int tile, int zchunk, bool robin, const DirData<TIO> &dd,
                     const TIO *z, const TIO *zz, TIO *y1, TIO *y2, double *dot,
                     hipStream_t s) {
  const TIO *const vs[] = {z, zz, y1, y2, dd.po, dd.ppo, dd.pn, dd.ppn, dd.x};
  const bool blk = in_block(dd.blk, dd.blk_bytes, vs, 9,
                            batch ? 0 : m0.op.g.N * (int64_t)sizeof(TIO));
  OwnerLaunch<T, TIO> L{};
  L.zchunk = zchunk;
  L.x1 = z;
  L.y1 = y1;
  L.x2 = zz;
  L.y2 = y2;
  L.dot = dot;
  L.dd = dd;
  if (batch) {
    L.md = md;
    L.as = as;
    L.utot = as.nu;
    L.gtot = as.ng;
    return blk ? launch_owner_any<T, TIO, true, 1 + kBlk>(L, tile, robin, s)
               : launch_owner_any<T, TIO, true, 1>(L, tile, robin, s);
  }
  if (!owner_fits(m0.op.g)) {
    set_error("owner scatter: system too large for 32-bit offsets");
    return VGPU_ERR_ARG;
  }
  L.m0 = m0;
  L.utot = m0.seg.nu;
  L.gtot = owner_groups(m0.op.g, tile, zchunk);
  return blk ? launch_owner_any<T, TIO, false, 1 + kBlk>(L, tile, robin, s)
             : launch_owner_any<T, TIO, false, 1>(L, tile, robin, s);
}

template int launch_owner_dir<double, double>(const Member<double> &, const Member<double> *,
                                              const ActiveSet &, bool, int, int, bool,
                                              const DirData<double> &, const double *,
                                              const double *, double *, double *, double *,
                                              hipStream_t);
template int launch_owner_dir<float, float>(const Member<float> &, const Member<float> *,
                                            const ActiveSet &, bool, int, int, bool,
                                            const DirData<float> &, const float *,
                                            const float *, float *, float *, double *,
                                            hipStream_t);

template <typename T, typename TIO>
int launch_owner_stab(const OpData<T> &op, int mode, int zchunk, const StabData<TIO> &sd,
                      const TIO *r, TIO *y1, hipStream_t s) {
  const Grid3 &g = op.g;
  if (!owner_fits(g)) {
    set_error("owner scatter: system too large for 32-bit offsets");
    return VGPU_ERR_ARG;
  }
  if (mode != kStabDir && mode != kStabS) {
    set_error("BiCGSTAB launch: unknown mode");
    return VGPU_ERR_ARG;
  }
  OwnerLaunch<T, TIO> L{};
  L.m0.op = op;
  L.m0.seg.n = g.N;
  L.m0.seg.nu = owner_units(g, op.tile);
  L.utot = L.m0.seg.nu;
  L.gtot = owner_groups(g, op.tile, zchunk);
  L.zchunk = zchunk;
  L.x1 = r;
  L.y1 = y1;
  L.sd = sd;
  const bool rb = op.ra != 0.0;
  const int t = op.tile;
  const TIO *const va[] = {r, y1, sd.po, sd.vo, sd.to, sd.pn, sd.rn, sd.x, sd.rt};
  const TIO *const vb[] = {r, y1, sd.v, sd.rt};
  const int64_t len = g.N * (int64_t)sizeof(TIO);
  if (mode == kStabDir && in_block(sd.blk, sd.blk_bytes, va, 9, len))
    return sd.early ? launch_owner_any<T, TIO, false, kStabDir + 2 + kBlk>(L, t, rb, s)
                    : launch_owner_any<T, TIO, false, kStabDir + kBlk>(L, t, rb, s);
  if (mode == kStabS && in_block(sd.blk, sd.blk_bytes, vb, 4, len))
    return sd.early ? launch_owner_any<T, TIO, false, kStabS + 2 + kBlk>(L, t, rb, s)
                    : launch_owner_any<T, TIO, false, kStabS + kBlk>(L, t, rb, s);
  if (sd.early) {
    if (mode == kStabDir) return launch_owner_any<T, TIO, false, kStabDir + 2>(L, t, rb, s);
    return launch_owner_any<T, TIO, false, kStabS + 2>(L, t, rb, s);
  }
  if (mode == kStabDir) return launch_owner_any<T, TIO, false, kStabDir>(L, t, rb, s);
  return launch_owner_any<T, TIO, false, kStabS>(L, t, rb, s);
}

template <typename T, typename TIO>
int launch_owner_stab_batch(const Member<T> *md, const ActiveSet &as, int tile, int zchunk,
                            bool robin, int mode, const StabData<TIO> &sd, const TIO *r,
                            TIO *y1, hipStream_t s) {
  if (mode != kStabDir && mode != kStabS) {
    set_error("BiCGSTAB launch: unknown mode");
    return VGPU_ERR_ARG;
  }
  OwnerLaunch<T, TIO> L{};
  L.md = md;
  L.as = as;
  L.utot = as.nu;
  L.gtot = as.ng;
  L.zchunk = zchunk;
  L.x1 = r;
  L.y1 = y1;
  L.sd = sd;
  const TIO *const va[] = {r, y1, sd.po, sd.vo, sd.to, sd.pn, sd.rn, sd.x, sd.rt};
  const TIO *const vb[] = {r, y1, sd.v, sd.rt};
  if (mode == kStabDir && in_block(sd.blk, sd.blk_bytes, va, 9, 0))
    return launch_owner_any<T, TIO, true, kStabDir + kBlk>(L, tile, robin, s);
  if (mode == kStabS && in_block(sd.blk, sd.blk_bytes, vb, 4, 0))
    return launch_owner_any<T, TIO, true, kStabS + kBlk>(L, tile, robin, s);
  if (mode == kStabDir) return launch_owner_any<T, TIO, true, kStabDir>(L, tile, robin, s);
  return launch_owner_any<T, TIO, true, kStabS>(L, tile, robin, s);
}

template int launch_owner_stab_batch<float, float>(const Member<float> *, const ActiveSet &,
                                                   int, int, bool, int,
                                                   const StabData<float> &, const float *,
                                                   float *, hipStream_t);

template <typename T, typename TIO>
int launch_owner_jacobi(const OpData<T> &op, int zchunk, double omega, const TIO *x, const TIO *b,
                        const TIO *d, TIO *xn, hipStream_t s) {
  if (!owner_fits(op.g)) {
    set_error("owner scatter: system too large for 32-bit offsets");
    return VGPU_ERR_ARG;
  }
  OwnerLaunch<T, TIO> L{};
  L.m0.op = op;
  L.m0.seg.n = op.g.N;
  L.m0.seg.nu = owner_units(op.g, op.tile);
  L.utot = L.m0.seg.nu;
  L.gtot = owner_groups(op.g, op.tile, zchunk);
  L.zchunk = zchunk;
  L.x1 = x;
  L.y1 = xn;
  L.sd.d = d;
  L.sd.rt = b;
  L.sd.jomega = omega;
  return launch_owner_any<T, TIO, false, kJacobi>(L, op.tile, op.ra != 0.0, s);
}
template int launch_owner_jacobi<double, double>(const OpData<double> &, int, double,
                                                 const double *, const double *,
                                                 const double *, double *, hipStream_t);

template int launch_owner_stab<double, double>(const OpData<double> &, int, int,
                                               const StabData<double> &, const double *,
                                               double *, hipStream_t);
template int launch_owner_stab<float, float>(const OpData<float> &, int, int,
                                             const StabData<float> &, const float *, float *,
                                             hipStream_t);

// The ATOMIC brick launch of one apply for bricks of BX x BY element columns
// and zc element layers (zc <= 0: brick_depth's rule); the outputs must be
// zero on entry.
template <typename T, typename TIO, int BX, int BY>
void launch_bricks(const OpData<T> &op, int zc, const TIO *x1, TIO *y1, const TIO *x2,
                   TIO *y2, hipStream_t s) {
  const Grid3 &g = op.g;
  const bool dual = x2 != nullptr;
  if (zc <= 0) zc = brick_depth(g, BX, BY);
  const int ntx = (g.ex + BX - 1) / BX, nty = (g.ey + BY - 1) / BY;
  const int ntz = (g.ez + zc - 1) / zc;
  {
    dim3 grid(ntx * nty * ntz);
    if (dual)
      hipLaunchKernelGGL((k_brick<T, TIO, true, BX, BY>), grid, dim3(BX * BY), 0, s, op, zc,
                         x1, y1, x2, y2);
    else
      hipLaunchKernelGGL((k_brick<T, TIO, false, BX, BY>), grid, dim3(BX * BY), 0, s, op, zc,
                         x1, y1, x2, y2);
  }
}

// The two colour launches of the COLORED scatter on the owner kernel
// (kSlab): tiles of the context's owner shape, slabs of zc element layers
// (zc <= 0: slab_depth's rule), one workgroup per (slab, tile) of a colour.
template <typename T, typename TIO>
int launch_owner_slabs(const OpData<T> &op, int zc, const TIO *x1, TIO *y1, const TIO *x2,
                       TIO *y2, hipStream_t s) {
  int tx, ty;
  owner_shape(op.tile, &tx, &ty);
  const Grid3 &g = op.g;
  if (zc <= 0) zc = slab_depth(g, tx, ty);
  const int64_t tiles =
      (int64_t)((g.nax + tx - 2) / (tx - 1)) * ((g.nay + ty - 2) / (ty - 1));
  const int nsl = (g.ez + zc - 1) / zc;
  for (int c = 0; c < 2; ++c) {
    const int nslc = (nsl - c + 1) / 2;  // slabs of parity c
    if (nslc <= 0) continue;
    OwnerLaunch<T, TIO> L{};
    L.m0.op = op;
    L.m0.seg.n = g.N;
    L.m0.seg.nu = owner_units(g, op.tile);
    L.utot = L.m0.seg.nu;
    L.gtot = tiles * nslc;
    L.zchunk = zc;
    // the solver's done flag is the first member of its status
    L.st = reinterpret_cast<const SolveStatus *>(op.done);
    L.x1 = x1;
    L.y1 = y1;
    L.x2 = x2;
    L.y2 = y2;
    L.slab_color = c;
    const int rc = launch_owner_any<T, TIO, false, kSlab>(L, op.tile, op.ra != 0.0, s);
    if (rc) return rc;
  }
  return VGPU_OK;
}

// The two colour launches of the COLORED scatter with k_slab (systems too
// large for the owner kernel's 32-bit offsets) for tiles of BX x BY element
// columns and slabs of zc element layers (zc <= 0: slab_depth's rule).
template <typename T, typename TIO, int BX, int BY>
void launch_slabs(const OpData<T> &op, int zc, const TIO *x1, TIO *y1, const TIO *x2,
                  TIO *y2, hipStream_t s) {
  const Grid3 &g = op.g;
  const bool dual = x2 != nullptr;
  if (zc <= 0) zc = slab_depth(g, BX, BY);
  const int64_t tiles =
      (int64_t)((g.nax + BX - 2) / (BX - 1)) * ((g.nay + BY - 2) / (BY - 1));
  const int nsl = (g.ez + zc - 1) / zc;
  for (int c = 0; c < 2; ++c) {
    const int nslc = (nsl - c + 1) / 2;  // slabs of parity c
    if (nslc <= 0) continue;
    const dim3 grid((unsigned)(tiles * nslc));
    if (dual)
      hipLaunchKernelGGL((k_slab<T, TIO, true, BX, BY>), grid, dim3(BX * BY), 0, s, op, zc,
                         c, x1, y1, x2, y2);
    else
      hipLaunchKernelGGL((k_slab<T, TIO, false, BX, BY>), grid, dim3(BX * BY), 0, s, op, zc,
                         c, x1, y1, x2, y2);
  }
}

template <typename T, typename TIO>
int launch_apply(const OpData<T> &op, int scatter, int zc, const TIO *x1,
                 TIO *y1, const TIO *x2, TIO *y2, double *dot_part,
                 hipStream_t s) {
  const bool dual = x2 != nullptr;
  const Grid3 &g = op.g;
  if (scatter == VGPU_SCATTER_OWNER) {
    if (!owner_fits(g)) {
      set_error("owner scatter: system too large for 32-bit offsets (8 N and 40 E "
                "must stay below 2^31 bytes)");
      return VGPU_ERR_ARG;
    }
    Member<T> m0{};
    m0.op = op;
    m0.seg.off = 0;
    m0.seg.n = g.N;
    m0.seg.u0 = 0;
    m0.seg.nu = owner_units(g, op.tile);
    const SolveStatus *st = reinterpret_cast<const SolveStatus *>(op.done);
    OwnerLaunch<T, TIO> L{};
    L.m0 = m0;
    L.utot = m0.seg.nu;
    L.gtot = owner_groups(g, op.tile, zc);
    L.zchunk = zc;
    L.st = st;
    L.x1 = x1;
    L.y1 = y1;
    L.x2 = x2;
    L.y2 = y2;
    L.dot = dot_part;
    return launch_owner_any<T, TIO, false>(L, op.tile, op.ra != 0.0, s);
  }
  // brick modes need zeroed outputs; the brick shape follows op.tile
  // (0 = 32 x 8, 1 = 16 x 16, 2 = 64 x 4 element columns)
  if (scatter != VGPU_SCATTER_ATOMIC && scatter != VGPU_SCATTER_COLORED) {
    set_error("unknown scatter mode " + std::to_string(scatter));
    return VGPU_ERR_ARG;
  }
  if (scatter == VGPU_SCATTER_COLORED && owner_fits(g) && !colored_large_path()) {
    return launch_owner_slabs<T, TIO>(op, zc, x1, y1, x2, y2, s);
  } else if (scatter == VGPU_SCATTER_COLORED) {
    // systems whose vectors exceed the owner kernel's 32-bit offsets (or
    // any system with VGPU_COLORED_LARGE_PATH=1)
    switch (op.tile) {
      case 1:
        launch_slabs<T, TIO, 16, 16>(op, zc, x1, y1, x2, y2, s);
        break;
      case 2:
        launch_slabs<T, TIO, 64, 4>(op, zc, x1, y1, x2, y2, s);
        break;
      default:
        launch_slabs<T, TIO, 32, 8>(op, zc, x1, y1, x2, y2, s);
        break;
    }
  } else {
    hipLaunchKernelGGL((k_zero2<TIO>), dim3(1024), dim3(256), 0, s, op.done, y1,
                       dual ? y2 : nullptr, g.N);
    switch (op.tile) {
      case 1:
        launch_bricks<T, TIO, 16, 16>(op, zc, x1, y1, x2, y2, s);
        break;
      case 2:
        launch_bricks<T, TIO, 64, 4>(op, zc, x1, y1, x2, y2, s);
        break;
      default:
        launch_bricks<T, TIO, 32, 8>(op, zc, x1, y1, x2, y2, s);
        break;
    }
  }
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

template int launch_apply<double, double>(const OpData<double> &, int, int,
                                          const double *, double *,
                                          const double *, double *, double *,
                                          hipStream_t);
template int launch_apply<float, float>(const OpData<float> &, int, int,
                                        const float *, float *, const float *,
                                        float *, double *, hipStream_t);
template int launch_apply<float, double>(const OpData<float> &, int, int,
                                         const double *, double *,
                                         const double *, double *, double *,
                                         hipStream_t);

// ---------------------------------------------------------------------------
// Per-element tau of the DR test function (closed form of
// gauss_point_hex8DR for a box: J = diag(h/2)).
//   V1 (VANADIS.FOR:1952-2006): h_s = 2/|J^-1 e|, lambda_s = lam e.e,
//       beta = series / 1-1/Pe / coth-1/Pe.
//   V2 (gauss_point_hex8DR.for:106-157): h_s as V1, lambda_s = lam,
//       coth from exp.
//   V3 (omp_timing.for:1650-1688): h_s = 2/(sum |v_k|/h_k + 1e-14),
//       lambda_s = lam, coth from exp.
// ---------------------------------------------------------------------------
__global__ void k_tau(Grid3 g, const double *xc, const double *yc,
                      const double *zc, double *rec, int variant) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(e % g.ex);
    const int iy = (int)((e / g.ex) % g.ey);
    const int iz = (int)(e / ((int64_t)g.ex * g.ey));
    const double hx = xc[ix + 1] - xc[ix], hy = yc[iy + 1] - yc[iy],
                 hz = zc[iz + 1] - zc[iz];
    double *r = rec + 5 * e;
    const double v1 = r[0], v2 = r[1], v3 = r[2], l = r[3];
    const double eps = 1.0e-14;
    const double vmag = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
    double hs, ls;
    if (variant == VGPU_DR_V3) {
      hs = vmag > eps ? 2.0 / (fabs(v1) / hx + fabs(v2) / hy + fabs(v3) / hz + eps) : 0.0;
      ls = l;
    } else {
      double ex = 0, ey = 0, ez = 0;
      if (vmag > eps) {
        ex = v1 / vmag;
        ey = v2 / vmag;
        ez = v3 / vmag;
      }
      const double gx = 2.0 / hx * ex, gy = 2.0 / hy * ey, gz = 2.0 / hz * ez;
      const double gn = sqrt(gx * gx + gy * gy + gz * gz);
      hs = gn > eps ? 2.0 / gn : 0.0;
      ls = (variant == VGPU_DR_V1) ? l * ex * ex + l * ey * ey + l * ez * ez : l;
    }
    double Pe = 0.0;
    if (vmag > eps && ls > eps) Pe = vmag * hs / (2.0 * ls);
    double t = 0.0;
    if (Pe > 1.0e-12 && vmag > eps) {
      double beta;
      if (variant == VGPU_DR_V1) {
        if (Pe < 1.0e-3)
          beta = Pe / 3.0 - (Pe * Pe * Pe) / 45.0 + 2.0 * ((Pe * Pe) * (Pe * Pe) * Pe) / 945.0;
        else if (Pe > 50.0)
          beta = 1.0 - 1.0 / Pe;
        else
          beta = 1.0 / tanh(Pe) - 1.0 / Pe;
      } else {
        const double ep = exp(Pe), en = exp(-Pe);
        beta = (0.5 * (ep + en)) / (0.5 * (ep - en)) - 1.0 / Pe;
      }
      t = 0.5 * hs / vmag * beta;
    }
    r[4] = t;
  }
}

int launch_tau(const Grid3 &g, const double *xc, const double *yc,
               const double *zc, double *rec, int variant, hipStream_t s) {
  hipLaunchKernelGGL(k_tau, dim3(2048), dim3(256), 0, s, g, xc, yc, zc, rec,
                     variant);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

// ---------------------------------------------------------------------------
// Jacobi diagonal: node-centric, deterministic.  diag(n) = sum over the (up
// to 8) elements around n of A_e(a,a), a = local index of n; 1 on the mask
// (VANADIS.FOR:761-769, 799-805).  A_e(a,a) via the element apply of e_a.
// ---------------------------------------------------------------------------
__device__ double elem_entry(const OpData<double> &op, int ix, int iy, int iz,
                             int a) {
  const int64_t e = ix + (int64_t)op.g.ex * (iy + (int64_t)op.g.ey * iz);
  ElemConst<double> k;
  const double *r = op.rec + 5 * e;
  elem_setup<double>(op.hx[ix], op.hy[iy], op.hz[iz], op.ihx[ix], op.ihy[iy],
                     op.ihz[iz], r[3], r[0], r[1], r[2], r[4], op.cH, op.mu, k);
  double x[8], xh[8], yh[8], y[8];
  for (int c = 0; c < 8; ++c) x[c] = (c == a) ? 1.0 : 0.0;
  haar_fwd<double>(x, xh);
  haar_apply<double, false>(k, xh, yh);
  haar_inv<double>(yh, y);
  if (ix == 0 && op.ra != 0.0)
    robin_add<double, false>(k, op.hx[ix], r[4], r[0], r[1], r[2], op.ra, x, y);
  return y[a];
}

__global__ void k_diag(OpData<double> op, double *diag) {
  // local index of a node at corner offset (cx,cy,cz) of its element
  const int LOC[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // index cx + 2cy + 4cz -> local
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < op.g.N;
       n += (int64_t)gridDim.x * blockDim.x) {
    if (op.mask[n]) {
      diag[n] = 1.0;
      continue;
    }
    const int gx = (int)(n % op.g.nax);
    const int gy = (int)((n / op.g.nax) % op.g.nay);
    const int gz = (int)(n / ((int64_t)op.g.nax * op.g.nay));
    double s = 0.0;
    for (int c = 0; c < 8; ++c) {
      const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
      // element whose corner (cx,cy,cz) is this node
      const int ix = gx - cx, iy = gy - cy, iz = gz - cz;
      if (ix < 0 || iy < 0 || iz < 0 || ix >= op.g.ex || iy >= op.g.ey || iz >= op.g.ez)
        continue;
      s += elem_entry(op, ix, iy, iz, LOC[c]);
    }
    diag[n] = s;
  }
}

int launch_diag(const OpData<double> &op, double *diag, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3(2048), dim3(256), 0, s, op, diag);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

// Row a of an element matrix: y = A_e^T e_a (closed form, with the Robin term
// on x-min elements).
__device__ void elem_row(const OpData<double> &op, int ix, int iy, int iz, int a,
                         double y[8]) {
  const int64_t e = ix + (int64_t)op.g.ex * (iy + (int64_t)op.g.ey * iz);
  ElemConst<double> k;
  const double *r = op.rec + 5 * e;
  elem_setup<double>(op.hx[ix], op.hy[iy], op.hz[iz], op.ihx[ix], op.ihy[iy],
                     op.ihz[iz], r[3], r[0], r[1], r[2], r[4], op.cH, op.mu, k);
  double x[8], xh[8], yh[8];
  for (int c = 0; c < 8; ++c) x[c] = (c == a) ? 1.0 : 0.0;
  haar_fwd<double>(x, xh);
  haar_apply<double, true>(k, xh, yh);
  haar_inv<double>(yh, y);
  if (ix == 0 && op.ra != 0.0)
    robin_add<double, true>(k, op.hx[ix], r[4], r[0], r[1], r[2], op.ra, x, y);
}

// Diagonal dominance of the element matrices as diagonal_dominance_factor and
// is_diagonally_dominant compute it (VANADIS.FOR:2434-2489, 2248-2303): per
// node i, dg = sum over its elements of |A_e(i,i)|, sr = sum of |A_e(i,j)|,
// j != i; per block the largest sr/dg (rows with dg > 0) and the number of
// rows with dg <= sr.
__global__ __launch_bounds__(256) void k_ddf(OpData<double> op, double *pmax, int *pbad) {
  const int LOC[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // index cx + 2cy + 4cz -> local
  __shared__ double rm[4];
  __shared__ int rb[4];
  double m = 0.0;
  int bad = 0;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < op.g.N;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int gx = (int)(n % op.g.nax);
    const int gy = (int)((n / op.g.nax) % op.g.nay);
    const int gz = (int)(n / ((int64_t)op.g.nax * op.g.nay));
    double dg = 0.0, sr = 0.0;
    for (int c = 0; c < 8; ++c) {
      const int ix = gx - (c & 1), iy = gy - ((c >> 1) & 1), iz = gz - (c >> 2);
      if (ix < 0 || iy < 0 || iz < 0 || ix >= op.g.ex || iy >= op.g.ey || iz >= op.g.ez)
        continue;
      const int a = LOC[c];
      double y[8];
      elem_row(op, ix, iy, iz, a, y);
      for (int b = 0; b < 8; ++b) {
        if (b == a) dg += fabs(y[b]);
        else sr += fabs(y[b]);
      }
    }
    if (dg > 0.0) m = fmax(m, sr / dg);
    bad += dg <= sr;
  }
  for (int o = 32; o > 0; o >>= 1) {
    m = fmax(m, __shfl_down(m, o, 64));
    bad += __shfl_down(bad, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    rm[threadIdx.x >> 6] = m;
    rb[threadIdx.x >> 6] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = fmax(fmax(rm[0], rm[1]), fmax(rm[2], rm[3]));
    pbad[blockIdx.x] = rb[0] + rb[1] + rb[2] + rb[3];
  }
}

int launch_ddf(const OpData<double> &op, double *pmax, int *pbad, int nblocks,
               hipStream_t s) {
  hipLaunchKernelGGL(k_ddf, dim3(nblocks), dim3(256), 0, s, op, pmax, pbad);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

// Element source F_a = q_in [1/8 + tau sum_k v_k s_k(a) / (4 h_k)]
// (el3d fmv, VANADIS.FOR:525, 551-553: qv = q_in/vol, int w_a), node-centric.
__global__ void k_source(OpData<double> op, const double *q_in, double *f) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < op.g.N;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int gx = (int)(n % op.g.nax);
    const int gy = (int)((n / op.g.nax) % op.g.nay);
    const int gz = (int)(n / ((int64_t)op.g.nax * op.g.nay));
    double s = 0.0;
    for (int c = 0; c < 8; ++c) {
      const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
      const int ix = gx - cx, iy = gy - cy, iz = gz - cz;
      if (ix < 0 || iy < 0 || iz < 0 || ix >= op.g.ex || iy >= op.g.ey || iz >= op.g.ez)
        continue;
      const int64_t e = ix + (int64_t)op.g.ex * (iy + (int64_t)op.g.ey * iz);
      const double q = q_in[e];
      if (q == 0.0) continue;
      const double hx = op.xc[ix + 1] - op.xc[ix], hy = op.yc[iy + 1] - op.yc[iy],
                   hz = op.zc[iz + 1] - op.zc[iz];
      const double sx = cx ? 1.0 : -1.0, sy = cy ? 1.0 : -1.0, sz = cz ? 1.0 : -1.0;
      const double *r = op.rec + 5 * e;
      s += q * (0.125 + r[4] * (r[0] * sx / (4.0 * hx) + r[1] * sy / (4.0 * hy) +
                                r[2] * sz / (4.0 * hz)));
    }
    f[n] = s;
  }
}

int launch_source(const OpData<double> &op, const double *q_in, double *f,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_source, dim3(2048), dim3(256), 0, s, op, q_in, f);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

// Mass integral: for trilinear c on a box the 2x2x2 rule is exact, so
// int_e c = vol_e * mean of the 8 nodal values (integrate_mass_hex8,
// VANADIS.FOR:2493-2561).  Per-block partials, summed on the host in order.
__global__ void k_mass(Grid3 g, const double *xc, const double *yc,
                       const double *zc, const double *c, double *partials) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(e % g.ex);
    const int iy = (int)((e / g.ex) % g.ey);
    const int iz = (int)(e / ((int64_t)g.ex * g.ey));
    const int64_t n0 = ix + (int64_t)g.nax * (iy + (int64_t)g.nay * iz);
    const int64_t px = 1, py = g.nax, pz = (int64_t)g.nax * g.nay;
    const double sum = c[n0] + c[n0 + px] + c[n0 + px + py] + c[n0 + py] + c[n0 + pz] +
                       c[n0 + pz + px] + c[n0 + pz + px + py] + c[n0 + pz + py];
    const double vol = (xc[ix + 1] - xc[ix]) * (yc[iy + 1] - yc[iy]) * (zc[iz + 1] - zc[iz]);
    s += 0.125 * sum * vol;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

int launch_mass(const Grid3 &g, const double *xc, const double *yc,
                const double *zc, const double *c, double *partials, int nparts,
                hipStream_t s) {
  hipLaunchKernelGGL(k_mass, dim3(nparts), dim3(256), 0, s, g, xc, yc, zc, c,
                     partials);
  VGPU_HIP(hipGetLastError());
  return VGPU_OK;
}

}  // namespace vgpu
