// lbm_step.hip -- the stream-and-collide step, one thread per cell along
// the unit-stride z axis.
//
// Mapping: a wavefront owns 64 consecutive z of one (x, y) row
// (wave w -> row w / segs, segment w % segs); 4 wavefronts per workgroup.
// Row, segment and the nine neighbouring rows are wavefront-uniform, so the
// nine mask windows of the 3x3 row neighbourhood come in as scalar loads
// and each lane picks its z-1 / z / z+1 solid bits by shifting them.
//
// Variants (LBM-Bench src/kernel_*.c), all with half-way bounce-back folded
// into the addressing through the mask instead of the reference's
// post-pass copy lists (src/lbm.c:465-568):
//   pull    f_d = solid(x-c_d) ? src[x, inv d] : src[x-c_d, d]; dst[x, d] = f*_d
//   push    f_d = src[x, d]; solid(x+c_d) ? dst[x, inv d] : dst[x+c_d, d] = f*_d
//   AA even f_d = a[x, d];   a[x, inv d] = f*_d            (in place)
//   AA odd  f_d = solid(x-c_d) ? a[x, d] : a[x-c_d, inv d]
//           solid(x+c_d) ? a[x, inv d] : a[x+c_d, d] = f*_d   (in place)
// In the AA odd step the slot (y, k) is read and written only by cell
// y - c_k, or, when that cell is solid, by nobody else than y itself
// through the bounce-back branch, so cells never race.  Solid cells are
// skipped entirely.
#include "lbm_impl.h"

namespace lbm {

// minimum wavefronts per SIMD the register allocation must allow
#ifndef LBM_MIN_WAVES
#define LBM_MIN_WAVES 1
#endif

template <int V, bool AOS, int NT, typename S>
__global__ __launch_bounds__(512, LBM_MIN_WAVES) void kStepRow(S *a, S *b, const MaskWin *__restrict__ mask,
    Params<S> p, int remap)
{
  using R = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;

  RowCtx c;
  if (!c.init(mask, p.nx, p.ny, p.nz, p.segs, remap))
    return;
  const uint64_t cell = c.cell;
  auto cellOf  = [&](int ox, int oy, int oz) { return c.cellOf(ox, oy, oz); };
  auto solidOf = [&](int ox, int oy, int oz) { return c.solidOf(ox, oy, oz); };

  R f[Q];
  if constexpr (V == V_PULL || V == V_AA_ODD) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      uint64_t i;
      if (d == C) {
        i = slot<AOS>(cell, C, p.plane);
      } else {
        const bool bb = solidOf(-cx(d), -cy(d), -cz(d));
        if (V == V_PULL)
          i = bb ? slot<AOS>(cell, inv(d), p.plane) : slot<AOS>(cellOf(-cx(d), -cy(d), -cz(d)), d, p.plane);
        else
          i = bb ? slot<AOS>(cell, d, p.plane) : slot<AOS>(cellOf(-cx(d), -cy(d), -cz(d)), inv(d), p.plane);
      }
      f[d] = St::load(getS<NTL>(&a[i]), p.inv);
    }
  } else {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = St::load(getS<NTL>(&a[slot<AOS>(cell, d, p.plane)]), p.inv);
  }

  if (p.addForce)
    addForce(f, p.f1, p.f2);

  collide(f, p.trt);
  const Dither dz = ditherOf(cell, stepSeq(p), p.dither);

  if constexpr (V == V_PULL) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const uint64_t i = slot<AOS>(cell, d, p.plane);
      putS<NTS>(&b[i], St::storeStep(f[d], p.scale, ditherDir(dz, d)));
    }
  } else if constexpr (V == V_AA_EVEN) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const uint64_t i = slot<AOS>(cell, inv(d), p.plane);
      putS<NTS>(&a[i], St::storeStep(f[d], p.scale, ditherDir(dz, d)));
    }
  } else {
    S *o = V == V_PUSH ? b : a;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      uint64_t i;
      if (d == C) {
        i = slot<AOS>(cell, C, p.plane);
      } else {
        const bool bb = solidOf(cx(d), cy(d), cz(d));
        i = bb ? slot<AOS>(cell, inv(d), p.plane) : slot<AOS>(cellOf(cx(d), cy(d), cz(d)), d, p.plane);
      }
      putS<NTS>(&o[i], St::storeStep(f[d], p.scale, ditherDir(dz, d)));
    }
  }
}

template <int V, bool AOS, int NT, typename S>
static void launchT(const StepLaunch &L, hipStream_t s)
{
  Params<S> p       = makeParams<S>(L);
  const int wpb     = L.wavesPerBlock == 8 ? 8 : 4;
  const unsigned blocks = rowBlocks(L.g, wpb);
  hipLaunchKernelGGL((kStepRow<V, AOS, NT, S>), dim3(blocks), dim3(64 * wpb), 0, s, (S *)L.src,
      (S *)L.dst, L.g.mask, p, L.remap);
}

template <int V, bool AOS, typename S>
static void launchNT(const StepLaunch &L, hipStream_t s)
{
  switch (L.nt & 3) {
  case 0:
    return launchT<V, AOS, 0, S>(L, s);
  case 1:
    return launchT<V, AOS, 1, S>(L, s);
  case 2:
    return launchT<V, AOS, 2, S>(L, s);
  default:
    return launchT<V, AOS, 3, S>(L, s);
  }
}

template <typename S>
static void launchS(const StepLaunch &L, hipStream_t s)
{
  // AoS is the layout of the push/pull variants only (src/kernel_*_aos.c)
  if (L.g.aos)
    return L.variant == V_PULL ? launchNT<V_PULL, true, S>(L, s) : launchNT<V_PUSH, true, S>(L, s);
  switch (L.variant) {
  case V_PULL:
    return launchNT<V_PULL, false, S>(L, s);
  case V_PUSH:
    return launchNT<V_PUSH, false, S>(L, s);
  case V_AA_EVEN:
    return launchNT<V_AA_EVEN, false, S>(L, s);
  default:
    return launchNT<V_AA_ODD, false, S>(L, s);
  }
}

void launchStepRow(const StepLaunch &L, hipStream_t s)
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
