// lbm_impl.h -- kernel-side parameter block and slot addressing shared by
// lbm_step.hip, lbm_tile.hip and lbm_aux.hip.
#pragma once

#include "lbm_kernels.h"

namespace lbm {

template <typename S>
struct Params {
  using R = typename Store<S>::R;
  Trt<R> trt;
  R f1, f2;       // body force terms, added to f_E,f_NE,... and subtracted
                  // from f_W,f_NW,... (3 w c_x g of src/lbm.c:919)
  R scale, inv;   // storage scale and its inverse (exact powers of two)
  uint64_t plane;
  int nx, ny, nz, segs, nrows;
  int addForce;
};

template <typename R>
inline Trt<R> toTrt(const double *d, const float *f);
template <>
inline Trt<double> toTrt<double>(const double *d, const float *)
{
  return Trt<double>{ d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8] };
}
template <>
inline Trt<float> toTrt<float>(const double *, const float *f)
{
  return Trt<float>{ f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8] };
}

template <typename S>
inline Params<S> makeParams(const StepLaunch &L)
{
  using R = typename Store<S>::R;
  Params<S> p;
  p.trt      = toTrt<R>(L.trt, L.trtf);
  p.f1       = std::is_same<R, double>::value ? (R)L.f1 : (R)L.f1f;
  p.f2       = std::is_same<R, double>::value ? (R)L.f2 : (R)L.f2f;
  p.scale    = (R)L.scale;
  p.inv      = (R)L.invScale;
  p.plane    = L.g.plane;
  p.nx       = L.g.nx;
  p.ny       = L.g.ny;
  p.nz       = L.g.nz;
  p.segs     = L.g.segs;
  p.nrows    = L.g.nx * L.g.ny;
  p.addForce = L.addForce;
  return p;
}

// element index of (cell, d)
template <bool AOS>
__device__ __forceinline__ uint64_t slot(uint64_t cell, int d, uint64_t plane)
{
  return AOS ? cell * Q + (uint64_t)d : (uint64_t)d * plane + cell;
}

// body force on the 10 directions with c_x != 0
template <typename R>
__device__ __forceinline__ void addForce(R (&f)[Q], R f1, R f2)
{
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    if (cx(d) == 0)
      continue;
    const R fd = wclass(d) == 1 ? f1 : f2;
    if (cx(d) > 0)
      f[d] += fd;
    else
      f[d] += -fd;
  }
}

} // namespace lbm
