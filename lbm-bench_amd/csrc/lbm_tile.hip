// lbm_tile.hip -- the blocked variants (LBM-Bench src/kernel_blk_pull_soa.c,
// src/kernel_blk_push_soa.c): the same pull / push step as lbm_step.hip,
// traversed in tiles of TZ x TY x TX cells, one workgroup of 256 lanes per
// tile.  The reference blocks in 8^3 cubes for CPU caches
// (src/lbm.c:629-650); here a tile is 64 z x 2 y x 2 x, so each wavefront
// covers 64 consecutive z of one row (lane -> z = t & 63, y = (t >> 6) & 1,
// x = t >> 7) and its loads and stores are whole 512-byte runs.  An 8^3
// tile gave each wavefront eight 64-byte runs and ran at half the speed.
// Each cell is still independent of the others, so the results equal the
// unblocked step.  Solid bits are looked up per lane.
#include "lbm_impl.h"

namespace lbm {

constexpr int TZ = 64, TY = 2, TX = 2; // tile edge lengths (TZ * TY * TX = 256 lanes)

template <int V, typename S>
__global__ __launch_bounds__(256) void kStepTile(S *a, S *b, const MaskWin *__restrict__ mask,
    Params<S> p, int tilesY, int tilesZ)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const int t   = (int)threadIdx.x;
  const int bid = (int)blockIdx.x;
  const int tz  = bid % tilesZ;
  const int ty  = (bid / tilesZ) % tilesY;
  const int tx  = bid / (tilesZ * tilesY);
  const int z   = tz * TZ + (t & (TZ - 1));
  const int y   = ty * TY + ((t / TZ) & (TY - 1));
  const int x   = tx * TX + t / (TZ * TY);
  if (x >= p.nx || y >= p.ny || z >= p.nz)
    return;
  if (solidAt(mask, p.segs, p.ny, x, y, z))
    return;
  const uint64_t cell = ((uint64_t)x * p.ny + y) * p.nz + z;

  auto nbr = [&](int ox, int oy, int oz, bool &solid) -> uint64_t {
    const int xn = wrapAdd(x, ox, p.nx), yn = wrapAdd(y, oy, p.ny), zn = wrapAdd(z, oz, p.nz);
    solid = solidAt(mask, p.segs, p.ny, xn, yn, zn);
    return ((uint64_t)xn * p.ny + yn) * p.nz + zn;
  };

  R f[Q];
  if constexpr (V == V_PULL) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      uint64_t i = slot<false>(cell, d, p.plane);
      if (d != C) {
        bool bb;
        const uint64_t n = nbr(-cx(d), -cy(d), -cz(d), bb);
        i = bb ? slot<false>(cell, inv(d), p.plane) : slot<false>(n, d, p.plane);
      }
      f[d] = St::raw(a[i]);
    }
  } else {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = St::raw(a[slot<false>(cell, d, p.plane)]);
  }
  if (p.addForce)
    addForce(f, p.f1, p.f2);
  collide(f, p.trt);
  const Dither dz = ditherOf(cell, stepSeq(p), p.dither);
  if constexpr (V == V_PULL) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const uint64_t i = slot<false>(cell, d, p.plane);
      b[i] = St::rawStep(f[d], ditherDir(dz, d));
    }
  } else {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      uint64_t i = slot<false>(cell, d, p.plane);
      if (d != C) {
        bool bb;
        const uint64_t n = nbr(cx(d), cy(d), cz(d), bb);
        i = bb ? slot<false>(cell, inv(d), p.plane) : slot<false>(n, d, p.plane);
      }
      b[i] = St::rawStep(f[d], ditherDir(dz, d));
    }
  }
}

template <int V, typename S>
static void launchT(const StepLaunch &L, hipStream_t s)
{
  Params<S> p    = makeParams<S>(L);
  const int tX = (p.nx + TX - 1) / TX, tY = (p.ny + TY - 1) / TY, tZ = (p.nz + TZ - 1) / TZ;
  const unsigned blocks = (unsigned)((uint64_t)tX * tY * tZ);
  hipLaunchKernelGGL((kStepTile<V, S>), dim3(blocks), dim3(TZ * TY * TX), 0, s, (S *)L.src,
      (S *)L.dst, L.g.mask, p, tY, tZ);
}

template <typename S>
static void launchS(const StepLaunch &L, hipStream_t s)
{
  if (L.variant == V_PULL)
    launchT<V_PULL, S>(L, s);
  else
    launchT<V_PUSH, S>(L, s);
}

void launchStepTile(const StepLaunch &L, hipStream_t s)
{
  switch (L.storage) {
  case ST_F64:
    return launchS<double>(L, s);
  case ST_F32:
    return launchS<float>(L, s);
  case ST_F32DEV:
    return launchS<f32dev>(L, s);
  case ST_F16:
    return launchS<__half>(L, s);
  default:
    return launchS<__hip_bfloat16>(L, s);
  }
}

} // namespace lbm
