// The context object behind the vgpu_* C ABI (shared by the translation
// units that implement it).
#pragma once
#include <vector>

#include "vgpu_bbicg.hpp"
#include "vgpu_stab.hpp"

using namespace vgpu;

struct vgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Grid3 g{};
  double *xc = nullptr, *yc = nullptr, *zc = nullptr;
  double *ah[3] = {nullptr, nullptr, nullptr}, *aih[3] = {nullptr, nullptr, nullptr};
  float *fah[3] = {nullptr, nullptr, nullptr}, *faih[3] = {nullptr, nullptr, nullptr};
  uint8_t *mask = nullptr;
  double *rec = nullptr;  // per-element records vx, vy, vz, lambda, tau
  float *frec = nullptr;
  bool have_coeffs = false;
  double P = 0, alpha = 0, dt = 1;
  int variant = VGPU_DR_V1, scheme = VGPU_SCHEME_STEADY;
  double cH = 1, cC = 0;
  int scatter = VGPU_SCATTER_OWNER, zchunk = 0, tile = 0;
  double *diag = nullptr;
  float *fdiag = nullptr;
  bool diag_valid = false;
  double *vb = nullptr, *vx_ = nullptr, *vt = nullptr, *vt2 = nullptr,
         *red = nullptr;
  float *fb = nullptr, *fx = nullptr;
  double *tpart = nullptr;  // dot partials of vgpu_time_kernel
  int64_t tpart_n = 0;
  // vgpu_time_kernel's fused-iteration launch: four vectors of N doubles
  // (directions and operator outputs), then status and reduction partials
  double *tdir = nullptr;
  Bicg<double> bicg64;  // brick scatter modes
  Bicg<float> bicg32;
  BBicg<double> bb64;   // owner scatter: batched engine, one system
  BBicg<float> bb32;
  Stab<double> stab64;  // BiCGSTAB (owner launches)
  Stab<float> stab32;
  int method = VGPU_METHOD_BICG;
  uint64_t version = 1;
  std::vector<void *> user_allocs;

  OpData<double> op64(double cH_, double cC_) const {
    OpData<double> o{};
    o.g = g;
    o.xc = xc;
    o.yc = yc;
    o.zc = zc;
    o.mask = mask;
    o.hx = ah[0];
    o.hy = ah[1];
    o.hz = ah[2];
    o.ihx = aih[0];
    o.ihy = aih[1];
    o.ihz = aih[2];
    o.rec = rec;
    o.cH = cH_;
    o.mu = cH_ * P + cC_;
    o.ra = cH_ * alpha;
    o.done = nullptr;
    o.tile = tile;
    return o;
  }
  OpData<float> op32() const {
    OpData<float> o{};
    o.g = g;
    o.xc = xc;
    o.yc = yc;
    o.zc = zc;
    o.mask = mask;
    o.hx = fah[0];
    o.hy = fah[1];
    o.hz = fah[2];
    o.ihx = faih[0];
    o.ihy = faih[1];
    o.ihz = faih[2];
    o.rec = frec;
    o.cH = cH;
    o.mu = cH * P + cC;
    o.ra = cH * alpha;
    o.done = nullptr;
    o.tile = tile;
    return o;
  }
};
