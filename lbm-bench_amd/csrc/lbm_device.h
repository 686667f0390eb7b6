// lbm_device.h -- device-side building blocks of the D3Q19 TRT step for
// gfx950: velocity set, TRT collision, storage conversion, solid-mask
// windows and the per-variant slot addressing.
//
// Data layout in HBM (one lattice, no halo):
//   cell   = (x*nY + y)*nZ + z           z is the unit-stride axis
//   SoA    element (d, cell) at d*plane + cell, plane >= nX*nY*nZ (padded)
//   AoS    element (d, cell) at cell*19 + d
//   mask   one MaskWin per (row = x*nY + y, seg = z/64): 66 solid bits
//          w[k] = solid(x, y, (64*seg - 1 + k) mod nZ), k = 0..65,
//          stored as lo = w[1..64], hi bit0 = w[0], hi bit1 = w[65];
//          hi bit 63 flags a solid cell among the windows of the 3x3
//          rows around (x, y) (bounce-back needed by this wavefront),
//          hi bit 2 + 3 (ox + 1) + (oy + 1) one in the window of row
//          (x + ox, y + oy).
//          Lane l of the wavefront that owns z = 64*seg + l reads its
//          z-1, z, z+1 bits as w[l], w[l+1], w[l+2].
// Periodicity is index wrap-around on every axis; a non-periodic axis has
// solid boundary layers (as LBM-Bench's geometryCreate builds them,
// src/lbm.c:104-129), which no fluid cell ever looks past.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace lbm {

constexpr int Q = 19;
// direction ids of LBM-Bench (src/lbm.h:20-38)
enum Dir { N = 0, S, E, W, NE, SE, NW, SW, T, TN, TE, TW, TS, B, BS, BN, BW, BE, C };

__host__ __device__ constexpr int cx(int d)
{
  constexpr int v[Q] = { 0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0 };
  return v[d];
}
__host__ __device__ constexpr int cy(int d)
{
  constexpr int v[Q] = { 1, -1, 0, 0, 1, -1, 1, -1, 0, 1, 0, 0, -1, 0, -1, 1, 0, 0, 0 };
  return v[d];
}
__host__ __device__ constexpr int cz(int d)
{
  constexpr int v[Q] = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, 0 };
  return v[d];
}
__host__ __device__ constexpr int inv(int d)
{
  constexpr int v[Q] = { S, N, W, E, SW, NW, SE, NE, B, BS, BW, BE, BN, T, TN, TS, TE, TW, C };
  return v[d];
}
// weight class: 0 centre (1/3), 1 axis (1/18), 2 diagonal (1/36)
__host__ __device__ constexpr int wclass(int d)
{
  return d == C ? 0 : ((cx(d) != 0) + (cy(d) != 0) + (cz(d) != 0) == 1 ? 1 : 2);
}

// TRT constants, formed on the host in the order of
// src/kernel_pull_soa.c:18-28 with Lambda as a parameter.  d15 is the 3/2
// of D = rho - 3/2 |j|^2.  A collision on values scaled by s (the storage
// scale, a power of two) uses w1nh / s, w2nh / s and d15 / s: every term is
// then s times the unscaled one, rounding included.
template <typename R>
struct Trt {
  R we, wo, w0, w1, w2, w1x3, w2x3, w1nh, w2nh, d15;
};

// TRT collision in place (src/kernel_pull_soa.c:118-198): f holds the 19
// pre-collision values and receives the post-collision ones, both indexed
// by direction.  Moments are summed in the reference's order.
template <typename R>
__device__ __forceinline__ void collide(R (&f)[Q], const Trt<R> &p)
{
  const R ux = f[E] + f[NE] + f[SE] + f[TE] + f[BE] - f[W] - f[NW] - f[SW] - f[TW] - f[BW];
  const R uy = f[N] + f[NE] + f[NW] + f[TN] + f[BN] - f[S] - f[SE] - f[SW] - f[TS] - f[BS];
  const R uz = f[T] + f[TE] + f[TW] + f[TN] + f[TS] - f[B] - f[BE] - f[BW] - f[BN] - f[BS];
  const R rho = f[C] + f[N] + f[E] + f[S] + f[W] + f[NE] + f[SE] + f[SW] + f[NW] + f[T] +
                f[TN] + f[TE] + f[TS] + f[TW] + f[B] + f[BN] + f[BE] + f[BS] + f[BW];
  const R D = rho - (ux * ux + uy * uy + uz * uz) * p.d15;

  f[C] = f[C] - p.we * (f[C] - p.w0 * D);

  const R w1D = p.w1 * D, w2D = p.w2 * D;
  auto pair = [&](int i, int j, R ui, R wX3, R wNH, R wD) {
    const R fi = f[i], fj = f[j];
    const R even = p.we * (R(0.5) * (fi + fj) - ui * ui * wNH - wD);
    const R odd  = p.wo * (R(0.5) * (fi - fj) - ui * wX3);
    f[i] = fi - even - odd;
    f[j] = fj - even + odd;
  };
  pair(N, S, uy, p.w1x3, p.w1nh, w1D);
  pair(E, W, ux, p.w1x3, p.w1nh, w1D);
  pair(T, B, uz, p.w1x3, p.w1nh, w1D);
  pair(NW, SE, -ux + uy, p.w2x3, p.w2nh, w2D);
  pair(NE, SW, ux + uy, p.w2x3, p.w2nh, w2D);
  pair(TW, BE, -ux + uz, p.w2x3, p.w2nh, w2D);
  pair(TE, BW, ux + uz, p.w2x3, p.w2nh, w2D);
  pair(TS, BN, -uy + uz, p.w2x3, p.w2nh, w2D);
  pair(TN, BS, uy + uz, p.w2x3, p.w2nh, w2D);
}

// ---- storage forms ------------------------------------------------------
// Native storage keeps f itself; the low-precision forms keep
// (f - w*RhoRef) * scale with scale a power of two, so the conversion to
// the FP32 compute type is one exact multiply.  The collision is applied
// to the deviation directly: the incompressible equilibrium is linear in
// rho, so the w*RhoRef part drops out of every term.
//
// store() rounds to nearest even.  raw() and rawStep() convert without the
// scale: the step kernels collide on scaled values (constants in Params).
// storeStep() scales and rounds like rawStep(): for FP16 and BF16 it rounds with gfx950's stochastic-rounding
// converts (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32).  These add bits 16..31
// of a 32-bit seed below the last kept bit and truncate, so a seed that is
// uniform over those bits rounds up with probability (dropped part /
// quantum): each rounding is unbiased, and a slowly varying field (or a
// small body force added to it) does not collect the same rounding error
// step after step.  The seed is (value bits & mask) ^ salt.  FP16 drops 13
// and BF16 16 bits, so bits 16..31 of the value are kept bits only and the
// seed never depends on the part being rounded.  The salt hashes the cell
// index with a per-step sequence number and differs per direction.  With
// dithering off, mask is 0 and salt 0x80000000: half a quantum is added
// and the result truncated, which is round to nearest, ties away from 0.

// 32-bit integer hash (lowbias32: Wellons 2018)
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

struct Dither {
  uint32_t mask, salt;
};
// per cell: on != 0 dithers, seq varies from step to step.  The salt
// passes through an empty asm so that it is formed once per cell:
// otherwise the compiler re-derives the hash's last shift-xor inside every
// direction's seed.
__device__ __forceinline__ Dither ditherOf(uint64_t cell, uint32_t seq, int on)
{
  uint32_t salt = on ? mix32((uint32_t)cell ^ (uint32_t)(cell >> 32) ^ seq) : 0x80000000u;
  asm("" : "+v"(salt));
  return Dither{ on ? 0xffffffffu : 0u, salt };
}
// ditherOf(cell, seq, 1) and ditherOf(cell, seq, 0) with the mask visible
// to the compiler, for a kernel that tests the flag once and keeps a store
// loop for each case
__device__ __forceinline__ Dither ditherOn(uint64_t cell, uint32_t seq)
{
  uint32_t h = mix32((uint32_t)cell ^ (uint32_t)(cell >> 32) ^ seq);
  asm("" : "+v"(h));
  return Dither{ 0xffffffffu, h };
}
__device__ __forceinline__ Dither ditherNone() { return Dither{ 0u, 0x80000000u }; }
// per direction: the cell's salt xor'ed with a direction constant
__device__ __forceinline__ Dither ditherDir(Dither z, int d)
{
  return Dither{ z.mask, z.salt ^ (((uint32_t)d + 1u) * 0x9e3779b9u & z.mask) };
}
// the rounding seed of direction d for a value with bits `bits`:
// (bits & mask) ^ ditherDir(z, d).salt, formed with gfx950's three-input
// bitwise instruction (one instruction when the mask is known to be all
// ones, as after ditherOn, two otherwise, instead of three)
__device__ __forceinline__ uint32_t seedOf(uint32_t bits, Dither z, int d)
{
  const uint32_t cd = ((uint32_t)d + 1u) * 0x9e3779b9u;
  if (__builtin_constant_p(z.mask) && z.mask == 0u)
    return z.salt; // rounding to nearest: the seed is the salt alone
  if (__builtin_constant_p(z.mask) && z.mask == 0xffffffffu)
    return __builtin_amdgcn_bitop3_b32(bits, z.salt, cd, 0x96); // a ^ b ^ c
  return __builtin_amdgcn_bitop3_b32(bits, z.mask, z.salt ^ (cd & z.mask), 0x6a); // (a & b) ^ c
}

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf162v __attribute__((ext_vector_type(2)));

template <typename S>
struct Store;
template <>
struct Store<double> {
  using R = double;
  static constexpr bool scaled = false;
  __device__ static R load(double v, R) { return v; }
  __device__ static double store(R v, R) { return v; }
  __device__ static double storeStep(R v, R, Dither) { return v; }
  __device__ static R raw(double v) { return v; }
  __device__ static double rawStep(R v, Dither) { return v; }
  __device__ static double rawStepD(R v, Dither, int) { return v; }
};
template <>
struct Store<float> {
  using R = float;
  static constexpr bool scaled = false;
  __device__ static R load(float v, R) { return v; }
  __device__ static float store(R v, R) { return v; }
  __device__ static float storeStep(R v, R, Dither) { return v; }
  __device__ static R raw(float v) { return v; }
  __device__ static float rawStep(R v, Dither) { return v; }
  __device__ static float rawStepD(R v, Dither, int) { return v; }
};
// FP32 shifted-deviation storage: same bytes as float, scaled
struct f32dev {
  float v;
};
template <>
struct Store<f32dev> {
  using R = float;
  static constexpr bool scaled = true;
  __device__ static R load(f32dev v, R inv) { return v.v * inv; }
  __device__ static f32dev store(R v, R sc) { return f32dev{ v * sc }; }
  __device__ static f32dev storeStep(R v, R sc, Dither) { return f32dev{ v * sc }; }
  __device__ static R raw(f32dev v) { return v.v; }
  __device__ static f32dev rawStep(R v, Dither) { return f32dev{ v }; }
  __device__ static f32dev rawStepD(R v, Dither, int) { return f32dev{ v }; }
};
template <>
struct Store<__half> {
  using R = float;
  static constexpr bool scaled = true;
  __device__ static R load(__half v, R inv) { return __half2float(v) * inv; }
  __device__ static __half store(R v, R sc) { return __float2half_rn(v * sc); }
  __device__ static __half storeStep(R v, R sc, Dither z) { return rawStep(v * sc, z); }
  __device__ static R raw(__half v) { return __half2float(v); }
  __device__ static __half rawStep(R x, Dither z)
  {
    const uint32_t seed = (__float_as_uint(x) & z.mask) ^ z.salt;
    // the tied operand is the value's own register (its upper half is
    // not used), so no register needs clearing first
    const half2v r      = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(half2v, x), x, seed, false);
    return __half(__half_raw{ r[0] });
  }
  // rawStep(x, ditherDir(z, d)), the seed formed by seedOf
  __device__ static __half rawStepD(R x, Dither z, int d)
  {
    const uint32_t seed = seedOf(__float_as_uint(x), z, d);
    const half2v r      = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(half2v, x), x, seed, false);
    return __half(__half_raw{ r[0] });
  }
};
template <>
struct Store<__hip_bfloat16> {
  using R = float;
  static constexpr bool scaled = true;
  __device__ static R load(__hip_bfloat16 v, R inv) { return __bfloat162float(v) * inv; }
  __device__ static __hip_bfloat16 store(R v, R sc) { return __float2bfloat16(v * sc); }
  __device__ static __hip_bfloat16 storeStep(R v, R sc, Dither z) { return rawStep(v * sc, z); }
  __device__ static R raw(__hip_bfloat16 v) { return __bfloat162float(v); }
  __device__ static __hip_bfloat16 rawStep(R x, Dither z)
  {
    const uint32_t seed = (__float_as_uint(x) & z.mask) ^ z.salt;
    const bf162v r      = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf162v, x), x, seed, false);
    return __hip_bfloat16(__hip_bfloat16_raw{ __builtin_bit_cast(unsigned short, r[0]) });
  }
  // rawStep(x, ditherDir(z, d)), the seed formed by seedOf
  __device__ static __hip_bfloat16 rawStepD(R x, Dither z, int d)
  {
    const uint32_t seed = seedOf(__float_as_uint(x), z, d);
    const bf162v r      = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf162v, x), x, seed, false);
    return __hip_bfloat16(__hip_bfloat16_raw{ __builtin_bit_cast(unsigned short, r[0]) });
  }
};

// ---- solid mask -----------------------------------------------------------
// lo: bit l = cell z0 + l of the row segment; hi: bit 0 = z0 - 1, bit 1 =
// z0 + 64 (both wrapped), bit 63 = some cell of the 3x3 rows around the
// segment (z halo included) is solid, bits 2 .. 10 = the same per row of
// the 3x3 (bit 2 + 3 (ox + 1) + (oy + 1) for row (x + ox, y + oy))
struct alignas(16) MaskWin {
  uint64_t lo, hi;
};
__device__ __forceinline__ bool nearSolid(MaskWin w) { return (w.hi >> 63) != 0; }

// window shifted so that bit l is the z-1 / z / z+1 neighbour of lane l
struct RowBits {
  uint64_t m, c, p; // z-1, z, z+1
};
__device__ __forceinline__ RowBits rowBits(MaskWin w)
{
  RowBits r;
  r.c = w.lo;
  r.m = (w.lo << 1) | (w.hi & 1ull);
  r.p = (w.lo >> 1) | ((w.hi >> 1) << 63);
  return r;
}
__device__ __forceinline__ bool bitAt(uint64_t word, int lane)
{
  return (word >> lane) & 1ull;
}

// per-lane lookup of an arbitrary cell (used by the generic passes: init,
// force, reductions, the adjoint objective)
__device__ __forceinline__ bool solidAt(const MaskWin *__restrict__ mask, int segs, int nY,
    int x, int y, int z)
{
  const MaskWin w = mask[((size_t)x * nY + y) * segs + (z >> 6)];
  return (w.lo >> (z & 63)) & 1ull;
}

// stores and loads, optionally non-temporal (streaming: the data is not
// reused by this kernel, so it should not displace lines that are)
template <bool NT, typename S>
__device__ __forceinline__ void putS(S *p, S v)
{
  if constexpr (NT) {
    if constexpr (sizeof(S) == 8)
      __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v), (unsigned long long *)p);
    else if constexpr (sizeof(S) == 4)
      __builtin_nontemporal_store(__builtin_bit_cast(unsigned int, v), (unsigned int *)p);
    else
      __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, v), (unsigned short *)p);
  } else {
    *p = v;
  }
}
template <bool NT, typename S>
__device__ __forceinline__ S getS(const S *p)
{
  if constexpr (NT) {
    if constexpr (sizeof(S) == 8)
      return __builtin_bit_cast(S, __builtin_nontemporal_load((const unsigned long long *)p));
    else if constexpr (sizeof(S) == 4)
      return __builtin_bit_cast(S, __builtin_nontemporal_load((const unsigned int *)p));
    else
      return __builtin_bit_cast(S, __builtin_nontemporal_load((const unsigned short *)p));
  } else {
    return *p;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This maps the hardware order onto a logical order in which
// consecutive logical workgroups sit on the same XCD: a bijection of
// [0, n) for any n.
__device__ __forceinline__ unsigned xcdLogicalBlock(unsigned b, unsigned n)
{
  const unsigned xcd = b % 8u, pos = b / 8u, q = n / 8u, r = n % 8u;
  return xcd < r ? xcd * (q + 1u) + pos : r * (q + 1u) + (xcd - r) * q + pos;
}

__device__ __forceinline__ int wrapInc(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int wrapDec(int v, int n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ int wrapAdd(int v, int c, int n)
{
  return c > 0 ? wrapInc(v, n) : (c < 0 ? wrapDec(v, n) : v);
}

} // namespace lbm
