// lbm_impl.h -- kernel-side parameter block and slot addressing shared by
// lbm_step.hip and lbm_aux.hip.
#pragma once

#include "lbm_kernels.h"

namespace lbm {

template <typename S>
struct Params {
  using R = typename Store<S>::R;
  Trt<R> trt;
  R f1, f2;       // body force terms, added to f_E,f_NE,... and subtracted
                  // from f_W,f_NW,... (3 w c_x g of src/lbm.c:919)
  R cf1, cf2;     // the same terms as the force pass forms them (the double
                  // values converted to R), added at the stores of a closing step
  R scale, inv;   // storage scale and its inverse (exact powers of two)
  uint64_t plane;
  int nx, ny, nz, segs, nrows;
  int addForce;
  int dither;   // FP16/BF16: stochastic rounding instead of to nearest
  const uint32_t *seqBase; // step number of the rounding seeds:
  uint32_t seq;            // *seqBase + seq
};

template <typename R>
inline Trt<R> toTrt(const double *d, const float *f);
template <>
inline Trt<double> toTrt<double>(const double *d, const float *)
{
  return Trt<double>{ d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], 1.5 };
}
template <>
inline Trt<float> toTrt<float>(const double *, const float *f)
{
  return Trt<float>{ f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], 1.5f };
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
  p.cf1      = (R)L.f1;
  p.cf2      = (R)L.f2;
  p.dither   = L.dither;
  p.seqBase  = L.seqBase;
  p.seq      = L.seq;
  if (Store<S>::scaled) {
    // the step kernels collide on the stored (scaled) values: the terms
    // quadratic in j and the force take the scale in (exact: a power of two)
    p.trt.w1nh *= p.inv;
    p.trt.w2nh *= p.inv;
    p.trt.d15 *= p.inv;
    p.f1 *= p.scale;
    p.f2 *= p.scale;
    p.cf1 *= p.scale;
    p.cf2 *= p.scale;
  }
  return p;
}

// seed input of a step: its step number, spread over the 32 bits
template <typename S>
__device__ __forceinline__ uint32_t stepSeq(const Params<S> &p)
{
  return (*p.seqBase + p.seq) * 0x9e3779b9u;
}

// element index of (cell, d)
template <bool AOS>
__device__ __forceinline__ uint64_t slot(uint64_t cell, int d, uint64_t plane)
{
  return AOS ? cell * Q + (uint64_t)d : (uint64_t)d * plane + cell;
}

// A pointer the compiler must take as it is: the slot offset d * plane
// stays in the (wavefront-uniform, scalar) base instead of being folded
// into every lane's address as a 64-bit vector multiply-add, so a lane's
// access costs one 64-bit add.  Not volatile: equal bases are computed once.
template <typename T>
__device__ __forceinline__ T *scalarBase(T *p)
{
  asm("" : "+s"(p));
  return p;
}
// element `cell` of SoA slot d
template <typename T>
__device__ __forceinline__ T *slotPtr(T *a, int d, uint64_t plane, uint64_t cell)
{
  return scalarBase(a + (uint64_t)d * plane) + cell;
}
// element rowCell + z of SoA slot d, rowCell wavefront-uniform: the base
// goes to SGPRs as a global-address-space pointer and the lane adds a
// 32-bit byte offset, so the access is one global load / store with a
// scalar base and no 64-bit address arithmetic per lane (z < nZ, so the
// offset fits for any lattice size)
template <typename T>
__device__ __forceinline__ T *rowSlot(T *a, int d, uint64_t plane, uint64_t rowCell, int z)
{
  typedef __attribute__((address_space(1))) char gchar;
  gchar *base = (gchar *)(a + (uint64_t)d * plane + rowCell);
  asm("" : "+s"(base));
  return (T *)(base + (uint32_t)z * (uint32_t)sizeof(T));
}

// the step's stores to the wavefront's own row: slot d (pull) or inv d (AA
// even) of cell rowCell + z
template <bool NTS, bool INV, typename S, typename R>
__device__ __forceinline__ void storeRow(S *o, uint64_t plane, uint64_t rowCell, int z,
    const R (&f)[Q], Dither dz)
{
#pragma unroll
  for (int d = 0; d < Q; ++d)
    putS<NTS>(rowSlot(o, INV ? inv(d) : d, plane, rowCell, z), Store<S>::rawStepD(f[d], dz, d));
}

// Closing steps.  The value a push or AA odd step streams to x + c_d is
// the logical f_d there after the step, and the bounce-back value it
// stores at x is the logical f_inv d of x.  The last step of a run can
// therefore add its own body force at the store, +t_d to a streamed value
// and -t_d to a bounce-back one, as the force pass after the step would
// (src/lbm.c:902-922).  The force pass loads the stored value and adds the
// term with one rounding: for storage without stochastic rounding (FP64,
// FP32, FP32 deviation) that is the same arithmetic, bit for bit; FP16 and
// BF16 storage round f*_d + t_d once, stochastically, where the force pass
// rounds f*_d and then the sum.
// t_d for a direction with c_x != 0
template <typename R>
__device__ __forceinline__ R closeTerm(int d, R cf1, R cf2)
{
  const R fd = wclass(d) == 1 ? cf1 : cf2;
  return cx(d) > 0 ? fd : -fd;
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

// One wavefront owns 64 consecutive z of one (x, y) row.  The row, the
// segment and the 3x3 neighbour rows are wavefront-uniform (scalar); each
// lane holds its z and the wrapped z-1 / z+1.  init() returns false for
// lanes past the lattice end and, unless keepSolid, for solid cells; with
// keepSolid a solid cell's lane stays with fluid == false.
struct RowCtx {
  const MaskWin *mask;
  int segs, ny, nz;
  int lane, row, seg, z;
  int xs[3], ys[3], zs[3];
  uint64_t cell, rowCell; // rowCell = row * nZ, wavefront-uniform
  bool fluid;

  // keepSolid: solid lanes stay (they store zeros into their own slots) when
  // their store line, lineLanes lanes wide, also holds a fluid cell.
  // tiled: rows are taken in tiles of 2 x 2 (x, y), tiles y fastest, so
  // that the wavefronts of consecutive rows of the grid cover a 2 x 2 x nZ
  // block; otherwise row after row.  Wavefront w of the grid owns segment
  // w % segs of the (w / segs)-th row in that order.
  __device__ __forceinline__ bool init(const MaskWin *__restrict__ m, int nx, int ny_, int nz_,
      int segs_, int remap, bool keepSolid = false, int lineLanes = 64, int tiled = 0)
  {
    mask = m;
    segs = segs_;
    ny   = ny_;
    nz   = nz_;
    const unsigned blk = remap ? xcdLogicalBlock(blockIdx.x, gridDim.x) : blockIdx.x;
    lane = (int)(threadIdx.x & 63u);
    const int wv = __builtin_amdgcn_readfirstlane((int)(blk * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    const int r  = wv / segs;
    seg          = wv - r * segs;
    if (tiled) {
      const int tY = (ny + 1) / 2, t = r >> 2, q = r & 3;
      const int xr = (t / tY) * 2 + (q >> 1), yr = (t % tY) * 2 + (q & 1);
      if (xr >= nx || yr >= ny)
        return false;
      row = xr * ny + yr;
    } else {
      row = r;
      if (row >= nx * ny)
        return false;
    }
    z   = seg * 64 + lane;
    if (z >= nz)
      return false;
    const int x = row / ny, y = row - x * ny;
    const uint64_t solid = mask[(uint64_t)row * segs + seg].lo;
    fluid                = !bitAt(solid, lane);
    if (!fluid) {
      if (!keepSolid)
        return false;
      const uint64_t line = lineLanes >= 64 ? ~0ull
                                            : (((1ull << lineLanes) - 1ull) << (lane & ~(lineLanes - 1)));
      if ((~solid & line) == 0ull)
        return false;
    }
    xs[0] = wrapDec(x, nx);
    xs[1] = x;
    xs[2] = wrapInc(x, nx);
    ys[0] = wrapDec(y, ny);
    ys[1] = y;
    ys[2] = wrapInc(y, ny);
    zs[0] = wrapDec(z, nz);
    zs[1] = z;
    zs[2] = wrapInc(z, nz);
    rowCell = (uint64_t)row * nz;
    cell    = rowCell + z;
    return true;
  }
  // row index of the neighbour row (x+ox, y+oy)
  __device__ __forceinline__ uint64_t rowOf(int ox, int oy) const
  {
    return (uint64_t)xs[ox + 1] * ny + ys[oy + 1];
  }
  __device__ __forceinline__ uint64_t cellOf(int ox, int oy, int oz) const
  {
    return rowOf(ox, oy) * nz + zs[oz + 1];
  }
  __device__ __forceinline__ bool solidOf(int ox, int oy, int oz) const
  {
    const MaskWin w  = mask[rowOf(ox, oy) * segs + seg];
    const RowBits rb = rowBits(w);
    return bitAt(oz < 0 ? rb.m : (oz > 0 ? rb.p : rb.c), lane);
  }
};

// launch geometry of the row kernels
inline unsigned rowBlocks(const Geom &g, int wavesPerBlock)
{
  const uint64_t wv = (uint64_t)g.nx * g.ny * g.segs;
  return (unsigned)((wv + wavesPerBlock - 1) / wavesPerBlock);
}
// tiles of 2 x 2 rows (4 wavefronts each) of a tiled launch, times segs
inline unsigned tileBlocks(const Geom &g, int segs)
{
  return (unsigned)((uint64_t)((g.nx + 1) / 2) * (uint64_t)((g.ny + 1) / 2) * (uint64_t)segs);
}

} // namespace lbm
