// lbm_step.hip -- the stream-and-collide step of the seven variants
// (LBM-Bench src/kernel_{pull,push,aa}_soa.c, src/kernel_{pull,push}_aos.c,
// src/kernel_blk_{pull,push}_soa.c).
//
// Five kernels share the slot rules below:
//   kStepRow   SoA, one wavefront per 64 consecutive z of one (x, y) row
//              (wave w -> segment w % segs of the (w / segs)-th row), 4 per
//              workgroup; pull, push, AA even and AA odd.
//   kRowSweep  SoA push and AA odd with nZ % 64 == 0: one wavefront walks a
//              whole row so that its z-shifted stores (and, for 8-byte
//              elements, loads) are line-aligned through DPP lane shifts.
//   kStepAoS   AoS push and pull: the side that touches the wavefront's own
//              64 records goes through LDS.
//   kPushCollideAoS, kPushGatherAoS  push-aos in two passes (the default):
//              collide into SoA planes of post-collision values, then take
//              each target slot from there, so every access is whole lines.
// The blocked variants run kStepRow / kRowSweep with the rows taken in
// tiles of 2 x 2 (x, y) instead of row after row (`tiled`); every cell is
// independent, so the results equal the unblocked step.
// Row, segment and the nine neighbouring rows are wavefront-uniform, so the
// mask windows of the 3x3 row neighbourhood come in as scalar loads and
// each lane picks its z-1 / z / z+1 solid bits by shifting them.
//
// Half-way bounce-back is folded into the addressing through the mask
// instead of the reference's post-pass copy lists (src/lbm.c:465-568):
//   pull    f_d = solid(x-c_d) ? src[x, inv d] : src[x-c_d, d]; dst[x, d] = f*_d
//   push    f_d = src[x, d]; solid(x+c_d) ? dst[x, inv d] : dst[x+c_d, d] = f*_d
//   AA even f_d = a[x, d];   a[x, inv d] = f*_d            (in place)
//   AA odd  f_d = solid(x-c_d) ? a[x, d] : a[x-c_d, inv d]
//           solid(x+c_d) ? a[x, inv d] : a[x+c_d, d] = f*_d   (in place)
// In the AA odd step the slot (y, k) is read and written only by cell
// y - c_k, or, when that cell is solid, by nobody else than y itself
// through the bounce-back branch, so cells never race.  No step reads the
// slots of a solid cell; kernels that store to their own cell write zeros
// there so that their store lines are whole.
#include "lbm_impl.h"

namespace lbm {

// minimum wavefronts per SIMD the register allocation must allow
#ifndef LBM_MIN_WAVES
#define LBM_MIN_WAVES 1
#endif

template <int V, int NT, typename S, bool CL = false>
__global__ __launch_bounds__(512, LBM_MIN_WAVES) void kStepRow(S *a, S *b, const MaskWin *__restrict__ mask,
    Params<S> p, int remap, int tiled)
{
  static_assert(!CL || V == V_PUSH || V == V_AA_ODD, "closing: push / AA odd");
  using R = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;

  // AA even and pull store only to their own cell, so they keep the lanes
  // of solid cells whose 128-byte store line holds a fluid cell: these
  // store the collision of zeros into their own slots (which no step
  // reads), and a store line shared by fluid and solid cells is written
  // whole instead of leaving holes for HBM to fill by read-modify-write (8 %
  // of the FP64 lines of a 512^3 sphere packing, 22 % of the FP16 ones;
  // measured +2-3 % for FP64 and FP32 AA).  Lines without fluid cells are
  // not written.
  constexpr bool keepSolid = V == V_AA_EVEN || V == V_PULL;
  RowCtx c;
  if (!c.init(mask, p.nx, p.ny, p.nz, p.segs, remap, keepSolid, (int)(128 / sizeof(S)), tiled))
    return;
  const uint64_t cell = c.cell;
  auto cellOf  = [&](int ox, int oy, int oz) { return c.cellOf(ox, oy, oz); };
  auto solidOf = [&](int ox, int oy, int oz) { return c.solidOf(ox, oy, oz); };

  R f[Q];
  if (keepSolid && !c.fluid) {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = R(0);
  } else if constexpr (V == V_PULL || V == V_AA_ODD) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const S *ptr;
      if (d == C) {
        ptr = rowSlot(a, C, p.plane, c.rowCell, c.z);
      } else {
        const bool bb = solidOf(-cx(d), -cy(d), -cz(d));
        if (V == V_PULL)
          ptr = bb ? slotPtr(a, inv(d), p.plane, cell) : slotPtr(a, d, p.plane, cellOf(-cx(d), -cy(d), -cz(d)));
        else
          ptr = bb ? slotPtr(a, d, p.plane, cell) : slotPtr(a, inv(d), p.plane, cellOf(-cx(d), -cy(d), -cz(d)));
      }
      f[d] = St::raw(getS<NTL>(ptr));
    }
  } else {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = St::raw(getS<NTL>(rowSlot(a, d, p.plane, c.rowCell, c.z)));
  }

  if (p.addForce)
    addForce(f, p.f1, p.f2);

  collide(f, p.trt);

  if constexpr (V == V_PULL || V == V_AA_EVEN) {
    // own-row stores; one uniform test of the rounding flag, so that each
    // copy of the store loop has its seeds in their plain form (two xors
    // per direction)
    constexpr bool SR = std::is_same<S, __half>::value || std::is_same<S, __hip_bfloat16>::value;
    S *o              = V == V_PULL ? b : a;
    const uint64_t rc = c.rowCell;
    const int zc      = c.z;
    if (SR && p.dither)
      storeRow<NTS, V == V_AA_EVEN>(o, p.plane, rc, zc, f, ditherOn(cell, stepSeq(p)));
    else
      storeRow<NTS, V == V_AA_EVEN>(o, p.plane, rc, zc, f, ditherNone());
  } else {
    const Dither dz = ditherOf(cell, stepSeq(p), p.dither);
    S *o            = V == V_PUSH ? b : a;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      S *ptr;
      R fv = f[d];
      if (d == C) {
        ptr = rowSlot(o, C, p.plane, c.rowCell, c.z);
      } else {
        const bool bb = solidOf(cx(d), cy(d), cz(d));
        ptr = bb ? slotPtr(o, inv(d), p.plane, cell) : slotPtr(o, d, p.plane, cellOf(cx(d), cy(d), cz(d)));
        if constexpr (CL) {
          if (cx(d) != 0) {
            const R t = closeTerm(d, p.cf1, p.cf2);
            fv += bb ? -t : t;
          }
        }
      }
      putS<NTS>(ptr, St::rawStepD(fv, dz, d));
    }
  }
}

// ---- AoS: the contiguous side through LDS ----------------------------------------
// In the AoS layout (src/kernel_{push,pull}_aos.c) a cell's 19 values are one
// record, so lane l addressing f_d of its own cell touches a different line
// for every lane: 64 lines per wavefront instruction, on both sides of the
// step.  One side is contiguous for a wavefront's 64 cells: push loads its
// own records, pull stores its own records (64 x 19 elements in a row).
// That side goes through LDS: lane i moves elements i, i + 64, ... of the
// block, so every global access covers 64 consecutive elements, and each
// lane reads / writes its record in LDS.  The other side (push's scatter,
// pull's gather, 1 to 3 slots of each of nine neighbour rows) stays per
// lane and is always cached: a line of records collects its 8-byte pieces
// from nine neighbour rows, and L2 has to hold it until they are all there
// (measured at 256^3: non-temporal scatter 1,196 MLUPS against 4,191
// cached for push, non-temporal gather 2,429 against 5,329 for pull).  The
// non-temporal bits apply to the contiguous side.  All lanes take part in
// the copies: solid cells collide zeros (pull writes zeros into their
// records, which no step reads), lanes past nZ move nothing.
// band > 0 takes the rows in bands of `band` y-rows with x swept inside
// each band, so that the nine rows sharing a record line run close
// together and L2 still holds the line (the launcher also limits the
// workgroups per CU through unused LDS for the same reason).
template <int V, int NT, typename S>
__global__ __launch_bounds__(256) void kStepAoS(const S *a, S *b, const MaskWin *__restrict__ mask,
    Params<S> p, int remap, int band)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  static_assert(V == V_PUSH || V == V_PULL, "AoS: push and pull only");
  __shared__ S stage[4][64 * Q];

  const unsigned blk = remap ? xcdLogicalBlock(blockIdx.x, gridDim.x) : blockIdx.x;
  const int w        = (int)(threadIdx.x >> 6);
  const int wv       = __builtin_amdgcn_readfirstlane((int)(blk * 4u + (unsigned)w));
  const int lane     = (int)(threadIdx.x & 63u);
  const int r        = wv / p.segs; // row in launch order
  if (r >= p.nrows)
    return; // the whole wavefront
  const int seg    = wv - r * p.segs;
  const int z0     = seg * 64;
  const int nvalid = min(64, p.nz - z0);
  const int z      = z0 + lane;
  int x, y;
  if (band > 0) {
    // band k holds y-rows [k band, k band + h) of every x, x slowest
    const int k = r / (band * p.nx), h = min(band, p.ny - k * band);
    const int i = r - k * band * p.nx;
    x           = i / h;
    y           = k * band + (i - x * h);
  } else {
    x = r / p.ny;
    y = r - x * p.ny;
  }
  const int row = x * p.ny + y;
  const MaskWin wc  = mask[(uint64_t)row * p.segs + seg];
  const bool fluid  = lane < nvalid && !bitAt(wc.lo, lane);
  const uint64_t c0 = (uint64_t)row * p.nz + (uint64_t)z0; // first cell of the block
  const int nel     = nvalid * Q;                            // elements of the block's records
  S *st             = stage[w];

  R f[Q];
  if constexpr (V == V_PUSH) {
    const S *src = a + c0 * Q;
    for (int e = lane; e < nel; e += 64)
      st[e] = getS<NTL>(&src[e]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = fluid ? St::raw(st[lane * Q + d]) : R(0);
  } else {
    const int xs[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
    const int ys[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
    const int zw    = min(z, p.nz - 1); // lanes past nZ read a valid cell, unused
    const int zs[3] = { wrapDec(zw, p.nz), zw, wrapInc(zw, p.nz) };
    auto rowOf      = [&](int ox, int oy) {
      return (uint64_t)xs[ox + 1] * (uint64_t)p.ny + (uint64_t)ys[oy + 1];
    };
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      f[d] = R(0);
      if (fluid) {
        uint64_t i = c0 * Q + (uint64_t)lane * Q + (uint64_t)d;
        if (d != C) {
          const int ox = -cx(d), oy = -cy(d), oz = -cz(d);
          const RowBits rb = rowBits(mask[rowOf(ox, oy) * p.segs + seg]);
          const bool bb    = bitAt(oz < 0 ? rb.m : (oz > 0 ? rb.p : rb.c), lane);
          i = bb ? c0 * Q + (uint64_t)lane * Q + (uint64_t)inv(d)
                 : (rowOf(ox, oy) * p.nz + (uint64_t)zs[oz + 1]) * Q + (uint64_t)d;
        }
        f[d] = St::raw(getS<false>(&a[i]));
      }
    }
  }
  if (p.addForce)
    addForce(f, p.f1, p.f2);
  collide(f, p.trt);
  const Dither dz = ditherOf(c0 + (uint64_t)lane, stepSeq(p), p.dither);

  if constexpr (V == V_PULL) {
    if (lane < nvalid) {
#pragma unroll
      for (int d = 0; d < Q; ++d)
        st[lane * Q + d] = St::rawStepD(f[d], dz, d);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    S *dst = b + c0 * Q;
    for (int e = lane; e < nel; e += 64)
      putS<NTS>(&dst[e], st[e]);
  } else if (fluid) {
    const int xs[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
    const int ys[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
    const int zs[3] = { wrapDec(z, p.nz), z, wrapInc(z, p.nz) };
    auto rowOf      = [&](int ox, int oy) {
      return (uint64_t)xs[ox + 1] * (uint64_t)p.ny + (uint64_t)ys[oy + 1];
    };
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      uint64_t i = (c0 + (uint64_t)lane) * Q + (uint64_t)d;
      if (d != C) {
        const RowBits rb = rowBits(mask[rowOf(cx(d), cy(d)) * p.segs + seg]);
        const bool bb    = bitAt(cz(d) < 0 ? rb.m : (cz(d) > 0 ? rb.p : rb.c), lane);
        i = bb ? (c0 + (uint64_t)lane) * Q + (uint64_t)inv(d)
               : (rowOf(cx(d), cy(d)) * p.nz + (uint64_t)zs[cz(d) + 1]) * Q + (uint64_t)d;
      }
      putS<false>(&b[i], St::rawStepD(f[d], dz, d));
    }
  }
}

// ---- push-aos in two passes: whole lines on both sides ---------------------------
// The single-pass push scatter leaves each record line of the target rows
// partly written by nine rows' workgroups; lines that leave L2 before all
// nine have written reach HBM as partial writes, which it completes by
// read-modify-write (at 256^3, 4.1x the algorithmic write bytes even in
// band order).  The two passes move twice the bytes, all in whole lines:
//   pass 1 (kPushCollideAoS): a block's records come in through LDS as in
//     the push step, collide, and f*_d goes to plane d of a SoA lattice
//     `post` at the cell itself;
//   pass 2 (kPushGatherAoS): target slot d of fluid cell y is what the
//     push scatter stores there, post[inv d][y] when y - c_d is solid (the
//     bounce-back of y's own f*_{inv d}) and post[d][y - c_d] otherwise;
//     the block's records leave through LDS as whole lines.
// The values are the stored (rounded) ones of pass 1, moved unchanged, so
// both forms give the same lattice bit for bit.
template <int NT, typename S>
__global__ __launch_bounds__(256) void kPushCollideAoS(const S *a, S *post, const MaskWin *__restrict__ mask,
    Params<S> p)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  __shared__ S stage[4][64 * Q];

  const int w    = (int)(threadIdx.x >> 6);
  const int wv   = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (unsigned)w));
  const int lane = (int)(threadIdx.x & 63u);
  const int row  = wv / p.segs;
  if (row >= p.nrows)
    return; // the whole wavefront
  const int seg     = wv - row * p.segs;
  const int z0      = seg * 64;
  const int nvalid  = min(64, p.nz - z0);
  const MaskWin wc  = mask[(uint64_t)row * p.segs + seg];
  const bool fluid  = lane < nvalid && !bitAt(wc.lo, lane);
  const uint64_t c0 = (uint64_t)row * p.nz + (uint64_t)z0;
  const int nel     = nvalid * Q;
  S *st             = stage[w];

  const S *src = a + c0 * Q;
  for (int e = lane; e < nel; e += 64)
    st[e] = getS<NTL>(&src[e]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  R f[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d)
    f[d] = fluid ? St::raw(st[lane * Q + d]) : R(0);
  if (p.addForce)
    addForce(f, p.f1, p.f2);
  collide(f, p.trt);
  const Dither dz = ditherOf(c0 + (uint64_t)lane, stepSeq(p), p.dither);
  if (lane < nvalid) {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      putS<NTS>(&post[(uint64_t)d * p.plane + c0 + (uint64_t)lane], St::rawStepD(f[d], dz, d));
  }
}

template <int NT, typename S>
__global__ __launch_bounds__(256) void kPushGatherAoS(const S *post, S *b, const MaskWin *__restrict__ mask,
    Params<S> p)
{
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  __shared__ S stage[4][64 * Q];

  const int w    = (int)(threadIdx.x >> 6);
  const int wv   = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (unsigned)w));
  const int lane = (int)(threadIdx.x & 63u);
  const int row  = wv / p.segs;
  if (row >= p.nrows)
    return; // the whole wavefront
  const int seg     = wv - row * p.segs;
  const int z0      = seg * 64;
  const int nvalid  = min(64, p.nz - z0);
  const int z       = z0 + lane;
  const int x       = row / p.ny, y = row - x * p.ny;
  const MaskWin wc  = mask[(uint64_t)row * p.segs + seg];
  const bool fluid  = lane < nvalid && !bitAt(wc.lo, lane);
  const uint64_t c0 = (uint64_t)row * p.nz + (uint64_t)z0;
  const uint64_t c  = c0 + (uint64_t)lane;
  const int nel     = nvalid * Q;
  S *st             = stage[w];

  const int xs[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
  const int ys[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
  const int zw    = min(z, p.nz - 1); // lanes past nZ read a valid cell, unused
  const int zs[3] = { wrapDec(zw, p.nz), zw, wrapInc(zw, p.nz) };
  auto rowOf      = [&](int ox, int oy) {
    return (uint64_t)xs[ox + 1] * (uint64_t)p.ny + (uint64_t)ys[oy + 1];
  };
  S v[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    v[d] = S{};
    if (fluid) {
      uint64_t i = (uint64_t)C * p.plane + c;
      if (d != C) {
        const int ox = -cx(d), oy = -cy(d), oz = -cz(d);
        const RowBits rb = rowBits(mask[rowOf(ox, oy) * p.segs + seg]);
        const bool bb    = bitAt(oz < 0 ? rb.m : (oz > 0 ? rb.p : rb.c), lane);
        i = bb ? (uint64_t)inv(d) * p.plane + c
               : (uint64_t)d * p.plane + rowOf(ox, oy) * p.nz + (uint64_t)zs[oz + 1];
      }
      v[d] = getS<NTL>(&post[i]);
    }
  }
  if (lane < nvalid) {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      st[lane * Q + d] = v[d];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  S *dst = b + c0 * Q;
  for (int e = lane; e < nel; e += 64)
    putS<NTS>(&dst[e], st[e]);
}

template <int V, int NT, typename S>
static void launchAoS(const StepLaunch &L, hipStream_t s)
{
  Params<S> p           = makeParams<S>(L);
  const unsigned blocks = rowBlocks(L.g, 4);
  if constexpr (V == V_PUSH) {
    if (L.post) {
      Params<S> q = p;
      q.plane     = (uint64_t)L.g.nx * L.g.ny * L.g.nz; // planes of `post`
      hipLaunchKernelGGL((kPushCollideAoS<NT, S>), dim3(blocks), dim3(256), 0, s, (const S *)L.src,
          (S *)L.post, L.g.mask, q);
      hipLaunchKernelGGL((kPushGatherAoS<NT, S>), dim3(blocks), dim3(256), 0, s, (const S *)L.post,
          (S *)L.dst, L.g.mask, q);
      return;
    }
  }
  // unused dynamic LDS that leaves room for L.wgPerCu workgroups of the
  // 160 KiB of a CU (one workgroup takes at most 64 KiB)
  constexpr int kStage = 4 * 64 * Q * (int)sizeof(S);
  int pad              = 0;
  if (L.wgPerCu > 0)
    pad = std::min(65536, 163840 / L.wgPerCu - 256) - kStage;
  pad = std::max(pad, 0);
  hipLaunchKernelGGL((kStepAoS<V, NT, S>), dim3(blocks), dim3(256), (unsigned)pad, s, (const S *)L.src,
      (S *)L.dst, L.g.mask, p, L.remap, L.band);
}

// ---- row sweep: aligned scatter of the z-shifted directions ---------------------
// The push and AA odd steps write f*_d to x + c_d.  For the ten directions
// with c_z != 0 that target is one element off the wavefront's 64-element
// block, so a plain store leaves two partial 128-byte lines per block and
// direction, which HBM completes by read-modify-write.  Here one wavefront
// owns a whole (x, y) row (nZ a multiple of 64, planes below 2^32
// elements) and walks its blocks in z
// order.  Lane l of the store for block k writes element 64k + l of the
// target row, taking the value of lane l - c_z through a one-lane DPP shift:
//   c_z = +1: lane 0 takes lane 63 of block k - 1 (carried as a scalar);
//             block 0's element 0 waits for block nZ/64 - 1 and is stored
//             alone after the loop (one partial line per row).
//   c_z = -1: lane 63 takes lane 0 of block k + 1, so block k is stored one
//             iteration later; the last block takes lane 0 of block 0,
//             saved as a scalar.
// A streamed store is masked where the producing cell is solid (its fluid
// target fills that slot by bounce-back).  A fluid producer with a solid
// target stores into the solid cell's slot (no step reads it) and also to
// its own slot inv d, the bounce-back, so store lines are written whole
// except next to solid producers.  Every slot keeps a single writer (the
// cell that reads it in the AA odd step, or the push source), so the
// delayed stores do not race.
template <typename S>
__device__ __forceinline__ S laneShift(S v, S fill, bool up)
{
  // up: lane l <- lane l - 1 (wave_shr:1), lane 0 <- fill
  // down: lane l <- lane l + 1 (wave_shl:1), lane 63 <- fill
  if constexpr (sizeof(S) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v), f = __builtin_bit_cast(uint64_t, fill);
    const int lo = up ? __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)u, 0x138, 0xf, 0xf, false)
                      : __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)u, 0x130, 0xf, 0xf, false);
    const int hi = up ? __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(u >> 32), 0x138, 0xf, 0xf, false)
                      : __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(u >> 32), 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(S, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
  } else if constexpr (sizeof(S) == 4) {
    const int u = __builtin_bit_cast(int, v), f = __builtin_bit_cast(int, fill);
    const int r = up ? __builtin_amdgcn_update_dpp(f, u, 0x138, 0xf, 0xf, false)
                     : __builtin_amdgcn_update_dpp(f, u, 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(S, r);
  } else {
    const int u = (int)__builtin_bit_cast(unsigned short, v);
    const int f = (int)__builtin_bit_cast(unsigned short, fill);
    const int r = up ? __builtin_amdgcn_update_dpp(f, u, 0x138, 0xf, 0xf, false)
                     : __builtin_amdgcn_update_dpp(f, u, 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(S, (unsigned short)r);
  }
}

// value the compiler must treat as unknown at this point (no hoisting or CSE)
template <typename T>
__device__ __forceinline__ T opaqueS(T v)
{
  asm volatile("" : "+s"(v));
  return v;
}

// value of lane `lane` as a wavefront-uniform value
template <typename S>
__device__ __forceinline__ S laneValue(S v, int lane)
{
  if constexpr (sizeof(S) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(S, ((uint64_t)hi << 32) | lo);
  } else if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  } else {
    const int r = __builtin_amdgcn_readlane((int)__builtin_bit_cast(unsigned short, v), lane);
    return __builtin_bit_cast(S, (unsigned short)r);
  }
}

// DI: stochastic rounding on (FP16/BF16 with the rounding flag set), as a
// compile-time switch, so that each direction's seed is one three-way xor
template <int V, int NT, typename S, bool CL = false, bool DI = false>
__global__ __launch_bounds__(256, LBM_MIN_WAVES) void kRowSweep(S *a, S *b,
    const MaskWin *__restrict__ mask, Params<S> p, int remap, int tiled)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  static_assert(V == V_PUSH || V == V_AA_ODD, "row sweep: push and AA odd only");

  const unsigned blk = remap ? xcdLogicalBlock(blockIdx.x, gridDim.x) : blockIdx.x;
  // tiled: the workgroup's 4 wavefronts are a tile of 2 x 2 rows (y faster)
  int row;
  if (tiled) {
    const int w       = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const unsigned tY = (unsigned)(p.ny + 1) / 2u;
    const int xr = (int)(blk / tY) * 2 + (w >> 1), yr = (int)(blk % tY) * 2 + (w & 1);
    if (xr >= p.nx || yr >= p.ny)
      return;
    row = __builtin_amdgcn_readfirstlane(xr * p.ny + yr);
  } else {
    row = __builtin_amdgcn_readfirstlane((int)(blk * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (row >= p.nrows)
      return;
  }
  const int lane = (int)(threadIdx.x & 63u);
  const int nz = p.nz, K = p.segs;
  const int x = row / p.ny, y = row - x * p.ny;
  const int xs[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
  const int ys[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
  auto rowOf = [&](int ox, int oy) -> uint64_t {
    return (uint64_t)xs[ox + 1] * (uint64_t)p.ny + (uint64_t)ys[oy + 1];
  };
  // Slot and row offsets are formed at each use from strides the compiler
  // cannot see through: kept live across the block loop, the ~40 distinct
  // 64-bit offsets exceed the SGPRs and spill.
  // (the launcher guarantees plane < 2^32, so cell indices are 32-bit)
  auto sl = [&](uint64_t c, int d) -> uint64_t {
    return c + (uint64_t)(uint32_t)d * (uint64_t)opaqueS((uint32_t)p.plane);
  };
  auto rowStart = [&](int ox, int oy) -> uint64_t {
    const uint32_t r = (uint32_t)xs[ox + 1] * (uint32_t)p.ny + (uint32_t)ys[oy + 1];
    return (uint64_t)(r * opaqueS((uint32_t)nz));
  };
  auto winOf = [&](int ox, int oy, int k) -> MaskWin { return mask[rowOf(ox, oy) * K + k]; };
  auto bitsOf = [&](int ox, int oy, int oz, int k) -> uint64_t {
    const RowBits rb = rowBits(winOf(ox, oy, k));
    return oz < 0 ? rb.m : (oz > 0 ? rb.p : rb.c);
  };
  // lanes of block k whose streamed store (to row (ox, oy), z) is skipped:
  // those whose producer (z - oz) is solid; its target fills that slot by
  // bounce-back.  A solid target's slot is written (no step reads it), so
  // lines shared by fluid and solid cells are written whole.
  auto skipOf = [&](int ox, int oy, int oz, int k) -> uint64_t {
    (void)ox;
    (void)oy;
    return bitsOf(0, 0, -oz, k);
  };
  S *o = V == V_PUSH ? b : a;

  S carry[Q], pend[Q], first[Q], lcarry[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    carry[d]  = S{};
    pend[d]   = S{};
    first[d]  = S{};
    lcarry[d] = S{};
  }
  bool prevClean = true; // block k - 1 had no solid cell near it
  const uint32_t seqv = stepSeq(p); // read once, not once per block

  for (int k = 0; k < K; ++k) {
    const int z  = k * 64 + lane;
    const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
    const uint64_t cell = rowStart(0, 0) + (uint64_t)z;
    const MaskWin wc    = winOf(0, 0, k);
    auto cellN = [&](int ox, int oy, int oz) -> uint64_t {
      return rowStart(ox, oy) + (uint64_t)(oz < 0 ? zm : (oz > 0 ? zp : z));
    };

    // One block.  walls = false: no cell of the 3x3 rows around the block
    // (z halo included) is solid, so there is no bounce-back and no store
    // to skip.
    auto block = [&](auto walls) {
      constexpr bool W = decltype(walls)::value;
      const bool fluid = W ? !__builtin_amdgcn_inverse_ballot_w64(wc.lo) : true;
      R f[Q];
      if constexpr (V == V_AA_ODD && sizeof(S) == 8) {
        // gather from x - c_d.  The z-shifted directions load the block of
        // the neighbouring row aligned and shift it by one lane.  From
        // z - 1, lane 0 takes lane 63 of the previous block (element nZ - 1
        // for block 0).  From z + 1, lane 63 takes element 64k + 64 (0 for
        // the last block), one uniform load.  Each of these slots is read
        // and later written only by a cell of this row that has not stored
        // yet.  All lanes load (a fluid lane may take its value from the lane
        // of a solid cell); solid cells' results are never stored.  Every
        // load of the block, bounce-back loads included, is issued before
        // the first lane shift: a shift right after its own load would make
        // the wavefront wait for all loads issued so far, once per shifted
        // direction.  Used for 8-byte elements only: measured at 512^3, the
        // lane shifts cost 2-byte elements 16 %.
        S al[Q], edge[Q], own[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) {
          if (cz(d) == 0) {
            al[d] = getS<NTL>(&a[sl(cellN(-cx(d), -cy(d), 0), inv(d))]);
          } else {
            const uint64_t rb = rowStart(-cx(d), -cy(d));
            al[d]             = getS<NTL>(&a[sl(rb + (uint64_t)z, inv(d))]);
            if (cz(d) > 0)
              edge[d] = k == 0 ? a[sl(rb + (uint64_t)(nz - 1), inv(d))] : lcarry[d];
            else
              edge[d] = a[sl(rb + (uint64_t)(k == K - 1 ? 0 : k * 64 + 64), inv(d))];
          }
          own[d] = S{};
          if (W && d != C && __builtin_amdgcn_inverse_ballot_w64(bitsOf(-cx(d), -cy(d), -cz(d), k)))
            own[d] = getS<NTL>(&a[sl(cell, d)]); // bounce-back: own slot d
        }
#pragma unroll
        for (int d = 0; d < Q; ++d) {
          S v = al[d];
          if (cz(d) != 0) {
            v = laneShift(al[d], edge[d], cz(d) > 0);
            if (cz(d) > 0)
              lcarry[d] = laneValue(al[d], 63);
          }
          if (W && d != C && __builtin_amdgcn_inverse_ballot_w64(bitsOf(-cx(d), -cy(d), -cz(d), k)))
            v = own[d];
          f[d] = St::raw(v);
        }
      } else if constexpr (V == V_AA_ODD) {
        if (fluid) {
#pragma unroll
          for (int d = 0; d < Q; ++d) {
            uint64_t i = sl(cell, d);
            if (d != C) {
              i = sl(cellN(-cx(d), -cy(d), -cz(d)), inv(d));
              if (W && __builtin_amdgcn_inverse_ballot_w64(bitsOf(-cx(d), -cy(d), -cz(d), k)))
                i = sl(cell, d);
            }
            f[d] = St::raw(getS<NTL>(&a[i]));
          }
        } else {
#pragma unroll
          for (int d = 0; d < Q; ++d)
            f[d] = R(0);
        }
      } else if (fluid) {
#pragma unroll
        for (int d = 0; d < Q; ++d)
          f[d] = St::raw(getS<NTL>(&a[sl(cell, d)]));
      } else {
#pragma unroll
        for (int d = 0; d < Q; ++d)
          f[d] = R(0);
      }
      if (p.addForce)
        addForce(f, p.f1, p.f2);
      collide(f, p.trt);
      const Dither dz = DI ? ditherOn(cell, seqv) : ditherNone();

#pragma unroll
      for (int d = 0; d < Q; ++d) {
        // streamed value v; bounce-back value vb (they differ in a closing step)
        S v, vb;
        if constexpr (CL) {
          if (cx(d) != 0) {
            const R t = closeTerm(d, p.cf1, p.cf2);
            v         = St::rawStepD(f[d] + t, dz, d);
            vb        = St::rawStepD(f[d] - t, dz, d);
          } else {
            v = vb = St::rawStepD(f[d], dz, d);
          }
        } else {
          v = vb = St::rawStepD(f[d], dz, d);
        }
        if (d == C) {
          if (fluid)
            putS<NTS>(&o[sl(cell, C)], v);
          continue;
        }
        if (cz(d) == 0) {
          if (fluid)
            putS<NTS>(&o[sl(cellN(cx(d), cy(d), 0), d)], v);
          if (W && fluid && __builtin_amdgcn_inverse_ballot_w64(bitsOf(cx(d), cy(d), 0, k)))
            putS<NTS>(&o[sl(cell, inv(d))], vb); // bounce-back: own slot inv d
          continue;
        }
        if (W && fluid && __builtin_amdgcn_inverse_ballot_w64(bitsOf(cx(d), cy(d), cz(d), k)))
          putS<NTS>(&o[sl(cell, inv(d))], vb); // bounce-back
        const uint64_t tgt = rowStart(cx(d), cy(d));
        if (cz(d) > 0) {
          const S sh = laneShift(v, carry[d], true);
          carry[d]   = laneValue(v, 63);
          uint64_t skip = W ? skipOf(cx(d), cy(d), 1, k) : 0ull;
          if (k == 0)
            skip |= 1ull; // element 0 comes from the last block
          if (!__builtin_amdgcn_inverse_ballot_w64(skip))
            putS<NTS>(&o[sl(tgt + (uint64_t)z, d)], sh);
        } else {
          if (k == 0) {
            first[d] = laneValue(v, 0);
          } else {
            // block k - 1, completed by lane 0 of this block
            const S done = laneShift(pend[d], laneValue(v, 0), false);
            const uint64_t skip = prevClean ? 0ull : skipOf(cx(d), cy(d), -1, k - 1);
            if (!__builtin_amdgcn_inverse_ballot_w64(skip))
              putS<NTS>(&o[sl(tgt + (uint64_t)(z - 64), d)], done);
          }
          pend[d] = v;
        }
      }
    };
    const bool clean = !nearSolid(wc);
    if (clean)
      block(std::false_type{});
    else
      block(std::true_type{});
    prevClean = clean;
  }

  // the wrap-around: element 0 (c_z = +1) and the last block (c_z = -1)
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    if (cz(d) == 0)
      continue;
    const uint64_t tgt = rowStart(cx(d), cy(d));
    if (cz(d) > 0) {
      const uint64_t skip = skipOf(cx(d), cy(d), 1, 0) | ~1ull;
      if (!__builtin_amdgcn_inverse_ballot_w64(skip))
        putS<NTS>(&o[sl(tgt, d)], carry[d]);
    } else {
      const S done        = laneShift(pend[d], first[d], false);
      const uint64_t skip = prevClean ? 0ull : skipOf(cx(d), cy(d), -1, K - 1);
      if (!__builtin_amdgcn_inverse_ballot_w64(skip))
        putS<NTS>(&o[sl(tgt + (uint64_t)((K - 1) * 64 + lane), d)], done);
    }
  }
}

// ---- row group: one workgroup per row, one wavefront per 64-block ------------
// The row sweep's line-aligned z-shifted loads and stores without its serial
// walk.  The nZ/64 wavefronts of a workgroup take the blocks of one (x, y)
// row together; the one element per block and direction that a lane shift
// takes from the neighbouring block (lane 63 of block k - 1 for c_z = +1,
// lane 0 of block k + 1 for c_z = -1, wrapping in z) comes through LDS.
// A row's blocks then reach HBM together, as in kStepRow, instead of one
// block after the other.  AA odd (in place): every load of the workgroup
// completes before its first barrier and every store follows its second,
// so the slots that a lane shift moves between wavefronts are read before
// they are written (each slot is read and written by the one cell that
// owns it).  Results equal the row sweep's bit for bit.
template <int V, int NT, typename S, int WNT = (NT | ((NT & 2) << 1)), bool CL = false>
__global__ __launch_bounds__(512) void kRowGroup(S *a, S *b, const MaskWin *__restrict__ mask,
    Params<S> p, int tiled, int remap)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  // blocks with a solid cell nearby: the non-temporal bits of their
  // stores (1), aligned gathers (2) and bounce-back loads (4).  A line that
  // a bounce-back load reads is also read whole by the aligned gather of
  // the row at x - c_d (and, for T and B, by this wavefront's own gather),
  // so the FP64 AA odd step keeps these blocks' loads cached: measured at
  // 512^3 spheres, 8.25 -> 7.89 ms (loads of clean blocks cached as well:
  // 8.78; stores of these blocks cached, loads not: 9.33)
  constexpr bool WNTS = WNT & 1, WNTL = WNT & 2, WNTB = WNT & 4;
  static_assert(V == V_PUSH || V == V_AA_ODD, "row group: push and AA odd only");
  constexpr int E0 = T; // the ten directions with c_z != 0 are T .. BE
  __shared__ S xg[8][10], xs[8][10];

  int row;
  const unsigned blk = remap ? xcdLogicalBlock(blockIdx.x, gridDim.x) : blockIdx.x;
  if (tiled) {
    // the rows of a 2 x 2 tile are consecutive workgroups (y faster)
    const unsigned t = blk >> 2, i = blk & 3u;
    const unsigned tY = (unsigned)(p.ny + 1) / 2u;
    const int xr = (int)(t / tY) * 2 + (int)(i >> 1), yr = (int)(t % tY) * 2 + (int)(i & 1u);
    if (xr >= p.nx || yr >= p.ny)
      return; // the whole workgroup: no barrier is left waiting
    row = xr * p.ny + yr;
  } else {
    row = (int)blk;
  }
  const int w    = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63u);
  const int nz = p.nz, K = p.segs, k = w;
  const int wl = w == 0 ? K - 1 : w - 1, wr = w == K - 1 ? 0 : w + 1;
  const int x = row / p.ny, y = row - x * p.ny;
  const int xs_[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
  const int ys_[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
  auto rowOf = [&](int ox, int oy) -> uint64_t {
    return (uint64_t)xs_[ox + 1] * (uint64_t)p.ny + (uint64_t)ys_[oy + 1];
  };
  // slot offsets from a stride the compiler cannot see through (kept live,
  // the distinct 64-bit offsets would exceed the SGPRs)
  auto sl = [&](uint64_t c, int d) -> uint64_t {
    return c + (uint64_t)(uint32_t)d * (uint64_t)opaqueS((uint32_t)p.plane);
  };
  auto rowStart = [&](int ox, int oy) -> uint64_t {
    const uint32_t r = (uint32_t)xs_[ox + 1] * (uint32_t)p.ny + (uint32_t)ys_[oy + 1];
    return (uint64_t)(r * opaqueS((uint32_t)nz));
  };
  auto bitsOf = [&](int ox, int oy, int oz) -> uint64_t {
    const RowBits rb = rowBits(mask[rowOf(ox, oy) * K + k]);
    return oz < 0 ? rb.m : (oz > 0 ? rb.p : rb.c);
  };
  S *o = V == V_PUSH ? b : a;

  const int z         = k * 64 + lane;
  const uint64_t cell = rowStart(0, 0) + (uint64_t)z;
  const MaskWin wc    = mask[rowOf(0, 0) * K + k];
  // clean: no cell of the 3x3 rows around the block (z halo included) is
  // solid, so there is no bounce-back and no store to skip
  const bool clean = !nearSolid(wc);
  const bool fluid = clean || !__builtin_amdgcn_inverse_ballot_w64(wc.lo);

  // loads: AA odd gathers from x - c_d, aligned (element z of the
  // neighbour row, shifted below); push loads its own cell
  S al[Q], own[Q];
  auto gather = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      own[d] = S{};
      constexpr bool L_ = W ? WNTL : NTL, B_ = W ? WNTB : NTL;
      if constexpr (V == V_AA_ODD) {
        al[d] = getS<L_>(&a[sl(rowStart(-cx(d), -cy(d)) + (uint64_t)z, inv(d))]);
        if (W && d != C && __builtin_amdgcn_inverse_ballot_w64(bitsOf(-cx(d), -cy(d), -cz(d))))
          own[d] = getS<B_>(&a[sl(cell, d)]); // bounce-back: own slot d
      } else {
        al[d] = fluid ? getS<L_>(&a[sl(cell, d)]) : S{};
      }
    }
  };
  if (clean)
    gather(std::false_type{});
  else
    gather(std::true_type{});
  if constexpr (V == V_AA_ODD) {
#pragma unroll
    for (int d = E0; d < E0 + 10; ++d) {
      if ((cz(d) > 0 && lane == 63) || (cz(d) < 0 && lane == 0))
        xg[w][d - E0] = al[d];
    }
  }
  __syncthreads();

  R f[Q];
  auto take = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      S v = al[d];
      if constexpr (V == V_AA_ODD) {
        if (cz(d) != 0)
          v = laneShift(al[d], cz(d) > 0 ? xg[wl][d - E0] : xg[wr][d - E0], cz(d) > 0);
        if (W && d != C && __builtin_amdgcn_inverse_ballot_w64(bitsOf(-cx(d), -cy(d), -cz(d))))
          v = own[d];
      }
      f[d] = St::raw(v);
    }
  };
  if (clean)
    take(std::false_type{});
  else
    take(std::true_type{});
  if (p.addForce)
    addForce(f, p.f1, p.f2);
  collide(f, p.trt);
  // the collided values as they are: the closing compilation uses them
  // twice (streamed and bounce-back values), which would otherwise let the
  // compiler fuse the collision's multiply-adds differently from the plain
  // compilation (last-bit differences between a run closed by this kernel
  // and one continued past the same step)
#pragma unroll
  for (int d = 0; d < Q; ++d)
    asm("" : "+v"(f[d]));
  const Dither dz = ditherOf(cell, stepSeq(p), p.dither);
  S v[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    R fv = f[d];
    if constexpr (CL) {
      if (cx(d) != 0)
        fv += closeTerm(d, p.cf1, p.cf2);
    }
    v[d] = St::rawStepD(fv, dz, d);
    if ((cz(d) > 0 && lane == 63) || (cz(d) < 0 && lane == 0))
      xs[w][d - E0] = v[d];
  }
  // bounce-back value of direction d (it differs from v[d] in a closing step)
  auto bbValue = [&](int d) -> S {
    if constexpr (CL) {
      if (cx(d) != 0)
        return St::rawStepD(f[d] - closeTerm(d, p.cf1, p.cf2), dz, d);
    }
    return v[d];
  };
  __syncthreads();

  // stores to x + c_d; the z-shifted ones line-aligned: lane l of block k
  // writes element 64k + l with the value of cell 64k + l - c_z, skipped
  // where that producer is solid (its fluid target fills the slot by
  // bounce-back)
  auto scatter = [&](auto walls) {
    constexpr bool W  = decltype(walls)::value;
    constexpr bool S_ = W ? WNTS : NTS;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      if (d == C) {
        if (fluid)
          putS<S_>(&o[sl(cell, C)], v[C]);
        continue;
      }
      if (W && fluid && __builtin_amdgcn_inverse_ballot_w64(bitsOf(cx(d), cy(d), cz(d))))
        putS<S_>(&o[sl(cell, inv(d))], bbValue(d)); // bounce-back: own slot inv d
      const uint64_t tgt = rowStart(cx(d), cy(d)) + (uint64_t)z;
      if (cz(d) == 0) {
        if (fluid)
          putS<S_>(&o[sl(tgt, d)], v[d]);
        continue;
      }
      const S sh = laneShift(v[d], cz(d) > 0 ? xs[wl][d - E0] : xs[wr][d - E0], cz(d) > 0);
      const uint64_t skip = W ? bitsOf(0, 0, -cz(d)) : 0ull;
      if (!__builtin_amdgcn_inverse_ballot_w64(skip))
        putS<S_>(&o[sl(tgt, d)], sh);
    }
  };
  if (clean)
    scatter(std::false_type{});
  else
    scatter(std::true_type{});
}

// ---- vector row group: the AA odd step of 4- and 2-byte elements ----------------
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// kRowGroup's shape for elements narrower than 8 bytes.  Each lane holds
// CPL = WB / sizeof(S) consecutive cells of the row as one word of WB
// bytes (8, or 4 for 2-byte elements; cell z0 + j in bits
// [j * 8 sizeof(S), ...) of the word), so a wavefront covers 64 * CPL cells
// and every load and store instruction moves 64 * WB line-aligned bytes; a
// 2-byte element per lane moves 128 bytes, and the wavefronts of that shape
// (row sweep, per-block kernel) do more address and shift work per cell.
// The z shift by one cell is a funnel shift of the word with the
// neighbouring lane's word (one DPP lane shift of 32 bits), the
// neighbouring wavefront's end word coming through LDS as in kRowGroup.  Near a solid (a wavefront
// whose 3x3 row neighbourhood holds one) each lane reads the solid bits of
// its cells z0 - 1 .. z0 + CPL from the mask windows and stores per cell
// where only some of its cells may be written: a streamed store is skipped
// for a solid producer (its fluid target fills the slot by bounce-back) and
// a bounce-back store is made for a solid target, as in the other kernels.
// The loads complete before the first barrier (their values reach LDS or
// the collision before the second) and every store follows the second, so
// the slots that move between wavefronts are read before they are written.
// DI: stochastic rounding on, as a compile-time switch (as in kRowSweep)
template <int NT, typename S, int WB = 8, bool CL = false, bool DI = false>
__global__ __launch_bounds__(512) void kRowGroupV(S *a, const MaskWin *__restrict__ mask,
    Params<S> p, int remap)
{
  static_assert(sizeof(S) == 2 || sizeof(S) == 4, "vector row group: 4- and 2-byte elements");
  using R  = typename Store<S>::R;
  using St = Store<S>;
  using E  = std::conditional_t<sizeof(S) == 2, unsigned short, unsigned int>;
  using Wd = std::conditional_t<WB == 8, uint64_t, uint32_t>; // the lane's word
  static_assert(WB == 8 || (WB == 4 && sizeof(S) == 2), "words of 8 bytes, or 4 for 2-byte elements");
  constexpr int CPL   = WB / (int)sizeof(S);
  constexpr int BITS  = 8 * (int)sizeof(S);
  constexpr int SH    = 32 - BITS; // funnel shift of a one-cell move toward higher z
  constexpr uint32_t FULL = (1u << CPL) - 1u;
  constexpr bool NTS = NT & 1, NTL = NT & 2;
  constexpr int E0 = T; // the ten directions with c_z != 0 are T .. BE
  __shared__ uint32_t xg[8][10], xs[8][10];

  const int row  = (int)(remap ? xcdLogicalBlock(blockIdx.x, gridDim.x) : blockIdx.x);
  const int w    = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int nw   = (int)(blockDim.x >> 6);
  const int lane = (int)(threadIdx.x & 63u);
  const int nz = p.nz, K = p.segs;
  const int wl = w == 0 ? nw - 1 : w - 1, wr = w == nw - 1 ? 0 : w + 1;
  const int x = row / p.ny, y = row - x * p.ny;
  const int xs_[3] = { wrapDec(x, p.nx), x, wrapInc(x, p.nx) };
  const int ys_[3] = { wrapDec(y, p.ny), y, wrapInc(y, p.ny) };
  auto rowOf = [&](int ox, int oy) -> uint64_t {
    return (uint64_t)xs_[ox + 1] * (uint64_t)p.ny + (uint64_t)ys_[oy + 1];
  };
  // slot offsets from strides the compiler cannot see through (as in kRowGroup)
  auto sl = [&](uint64_t c, int d) -> uint64_t {
    return c + (uint64_t)(uint32_t)d * (uint64_t)opaqueS((uint32_t)p.plane);
  };
  auto rowStart = [&](int ox, int oy) -> uint64_t {
    const uint32_t r = (uint32_t)xs_[ox + 1] * (uint32_t)p.ny + (uint32_t)ys_[oy + 1];
    return (uint64_t)(r * opaqueS((uint32_t)nz));
  };
  const int z0 = (w * 64 + lane) * CPL;
  // the lane's word of slot d in row (ox, oy): cells z0 .. z0 + CPL - 1,
  // as a scalar base (the row's start in the plane) plus the lane's 32-bit
  // byte offset, so an access costs no vector address arithmetic and all
  // of them share one offset register
  const uint32_t zoff = (uint32_t)z0 * (uint32_t)sizeof(S);
  typedef __attribute__((address_space(1))) char gchar;
  auto rowBase = [&](int ox, int oy, int d) -> gchar * {
    gchar *base = (gchar *)(a + sl(rowStart(ox, oy), d));
    asm("" : "+s"(base));
    return base;
  };
  auto wordPtr = [&](int ox, int oy, int d) -> Wd * { return (Wd *)(rowBase(ox, oy, d) + zoff); };
  auto ld = [&](const Wd *ptr, auto nt) -> Wd {
    if constexpr (decltype(nt)::value)
      return __builtin_nontemporal_load(ptr);
    else
      return *ptr;
  };
  auto st = [&](Wd *ptr, Wd v) {
    if constexpr (NTS)
      __builtin_nontemporal_store(v, ptr);
    else
      *ptr = v;
  };
  // Near a solid the stores go through a buffer resource on the target
  // row's slot plane (nZ elements): a lane that stores nothing takes an
  // offset beyond the resource's end, and the hardware drops such a store,
  // so the lane's cell mask m costs a select rather than a branch.  A lane
  // whose word holds two cells and has one of them to store does so in a
  // second store, which every lane makes (out of range where it has no
  // such cell): without a branch on the wavefront's masks it ran faster
  // than with one.  Words of four cells store their cells one by one, only
  // where the wavefront has a lane with some but not all of its cells.
  constexpr uint32_t OOB = 0x80000000u; // beyond any row's bytes
  constexpr int AUX      = NTS ? 2 : 0; // non-temporal
  constexpr int WORD3    = 0x00020000;  // untyped access, 32-bit data format (gfx9 family)
  auto stNear = [&](int ox, int oy, int d, Wd v, uint32_t m) {
    if constexpr (CPL > 2) {
      if (!__builtin_amdgcn_ballot_w64(m != 0u))
        return;
    }
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        (void *)rowBase(ox, oy, d), (short)0, nz * (int)sizeof(S), WORD3);
    const uint32_t off = m == FULL ? zoff : OOB;
    if constexpr (WB == 8)
      __builtin_amdgcn_raw_buffer_store_b64(
          u32x2{(uint32_t)v, (uint32_t)(v >> 32)}, r, (int)off, 0, AUX);
    else
      __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, (int)off, 0, AUX);
    auto stCell = [&](Wd e, uint32_t o) {
      if constexpr (sizeof(S) == 2)
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)e, r, (int)o, 0, AUX);
      else
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)e, r, (int)o, 0, AUX);
    };
    if constexpr (CPL == 2) {
      const bool hi = m == 2u;
      stCell(hi ? v >> BITS : v, m == 1u || hi ? zoff + (hi ? (uint32_t)sizeof(S) : 0u) : OOB);
    } else if (__builtin_amdgcn_ballot_w64(m != 0u && m != FULL)) {
#pragma unroll
      for (int j = 0; j < CPL; ++j)
        stCell(v >> (BITS * j),
            m != FULL && ((m >> j) & 1u) ? zoff + (uint32_t)(j * (int)sizeof(S)) : OOB);
    }
  };
  // one-cell moves of a word: toward higher z (cell z takes z - 1; `prev`
  // is the top 32 bits of the word below) and toward lower z (cell z takes
  // z + 1; `next` is the bottom 32 bits of the word above)
  auto funnel = [](uint32_t hi, uint32_t lo, int s) -> uint32_t {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
  };
  auto moveUp = [&](Wd v, uint32_t prev) -> Wd {
    if constexpr (WB == 8) {
      const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
      return ((uint64_t)funnel(hi, lo, SH) << 32) | funnel(lo, prev, SH);
    } else {
      return funnel(v, prev, SH);
    }
  };
  auto moveDown = [&](Wd v, uint32_t next) -> Wd {
    if constexpr (WB == 8) {
      const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
      return ((uint64_t)funnel(next, hi, BITS) << 32) | funnel(hi, lo, BITS);
    } else {
      return funnel(next, v, BITS);
    }
  };
  auto top32 = [](Wd v) -> uint32_t { return (uint32_t)(v >> (8 * WB - 32)); };
  // top 32 bits of lane l - 1's word (lane 0: fill) / bottom 32 bits of
  // lane l + 1's word (lane 63: fill)
  auto topOfBelow = [&](Wd v, uint32_t fill) -> uint32_t {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)top32(v), 0x138, 0xf, 0xf, false);
  };
  auto bottomOfAbove = [](Wd v, uint32_t fill) -> uint32_t {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)(uint32_t)v, 0x130, 0xf, 0xf, false);
  };

  // The wavefront's CPL mask windows of its own row (scalar).  clean: no
  // solid cell in the 3x3 rows around its cells (z halo included).  Near a
  // solid, the per-row flags of the windows say which 64-cell groups of
  // lanes may need a bounce-back value from a row: those lanes load their
  // own slot (one line per group), in the same round trip as the streamed
  // loads, and the per-cell solid bits come from per-lane window loads
  // issued with them.
  constexpr int LPW = 64 / CPL; // lanes per mask window
  MaskWin ow[CPL];
  bool clean = true;
#pragma unroll
  for (int i = 0; i < CPL; ++i) {
    ow[i] = mask[rowOf(0, 0) * K + w * CPL + i];
    clean = clean && !nearSolid(ow[i]);
  }
  // lanes whose window flags a solid in row (ox, oy)
  auto groupLanes = [&](int ox, int oy) -> uint64_t {
    const int r = 2 + 3 * (ox + 1) + (oy + 1);
    uint64_t m  = 0;
#pragma unroll
    for (int i = 0; i < CPL; ++i)
      m |= ((ow[i].hi >> r) & 1ull) ? (~0ull >> (64 - LPW)) << (LPW * i) : 0ull;
    return m;
  };
  // solid bits of cells z0 - 1 .. z0 + CPL of the nine rows (ox, oy): bit
  // i is cell z0 - 1 + i (wrapped in z through the window's halo bits)
  const int seg = z0 >> 6, b = z0 & 63;
  MaskWin mw[9];
  uint32_t fld9[9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
    fld9[i] = 0u;
  auto fields = [&]() {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const uint32_t own   = (uint32_t)(mw[i].lo >> b) & FULL;
      const uint32_t below = b == 0 ? (uint32_t)(mw[i].hi & 1ull) : (uint32_t)(mw[i].lo >> (b - 1)) & 1u;
      const uint32_t above =
          b + CPL == 64 ? (uint32_t)(mw[i].hi >> 1) & 1u : (uint32_t)(mw[i].lo >> (b + CPL)) & 1u;
      fld9[i] = below | (own << 1) | (above << (CPL + 1));
    }
  };
  auto field = [&](int ox, int oy) -> uint32_t { return fld9[(ox + 1) * 3 + oy + 1]; };
  // the CPL bits of cells z0 + j + oz, j = 0 .. CPL - 1
  auto cellsAt = [&](uint32_t fld, int oz) -> uint32_t { return (fld >> (1 + oz)) & FULL; };

  // gather from x - c_d: slot inv d of the neighbour row, moved by -c_z;
  // bounce-back cells take their own slot d
  Wd v[Q], own[Q];
  auto gather = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
    if constexpr (W) {
#pragma unroll
      for (int i = 0; i < 9; ++i)
        mw[i] = mask[rowOf(i / 3 - 1, i % 3 - 1) * K + seg];
    }
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      v[d]   = ld(wordPtr(-cx(d), -cy(d), inv(d)), std::integral_constant<bool, NTL>{});
      own[d] = 0;
      if (W && d != C && __builtin_amdgcn_inverse_ballot_w64(groupLanes(-cx(d), -cy(d))))
        own[d] = ld(wordPtr(0, 0, d), std::integral_constant<bool, NTL>{});
    }
    // the window loads went out first: waiting for them leaves the
    // gather in flight, and the windows' registers free before the barrier
    if constexpr (W)
      fields();
  };
  if (clean)
    gather(std::false_type{});
  else
    gather(std::true_type{});
#pragma unroll
  for (int d = E0; d < E0 + 10; ++d) {
    if (cz(d) > 0 && lane == 63)
      xg[w][d - E0] = top32(v[d]);
    if (cz(d) < 0 && lane == 0)
      xg[w][d - E0] = (uint32_t)v[d];
  }
  __syncthreads();
  auto take = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      if (cz(d) > 0)
        v[d] = moveUp(v[d], topOfBelow(v[d], xg[wl][d - E0]));
      else if (cz(d) < 0)
        v[d] = moveDown(v[d], bottomOfAbove(v[d], xg[wr][d - E0]));
      if (W && d != C) {
        const uint32_t m = cellsAt(field(-cx(d), -cy(d)), -cz(d));
        Wd sel           = 0;
#pragma unroll
        for (int j = 0; j < CPL; ++j)
          sel |= ((m >> j) & 1u) ? (Wd)(((1ull << BITS) - 1ull) << (BITS * j)) : (Wd)0;
        v[d] = (own[d] & sel) | (v[d] & ~sel);
      }
    }
  };
  if (clean)
    take(std::false_type{});
  else
    take(std::true_type{});

  // collide the lane's cells one after the other; out: streamed values,
  // written over the gathered word cell by cell (a cell's gathered value
  // is not used after its own collision, and a second array of words
  // would cost the registers of a wavefront per SIMD); bb: bounce-back
  // values (they differ in a closing step only)
  Wd(&out)[Q] = v;
  Wd bb[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d)
    bb[d] = 0;
  const uint32_t seqv = stepSeq(p);
  const uint64_t rc   = rowStart(0, 0);
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    R f[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d)
      f[d] = St::raw(__builtin_bit_cast(S, (E)(v[d] >> (BITS * j))));
    if (p.addForce)
      addForce(f, p.f1, p.f2);
    collide(f, p.trt);
#pragma unroll
    for (int d = 0; d < Q; ++d)
      asm("" : "+v"(f[d]));
    const Dither dz = DI ? ditherOn(rc + (uint64_t)(z0 + j), seqv) : ditherOf(rc + (uint64_t)(z0 + j), seqv, 0);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      R fv = f[d], fb = f[d];
      if constexpr (CL) {
        if (cx(d) != 0) {
          fv += closeTerm(d, p.cf1, p.cf2);
          fb -= closeTerm(d, p.cf1, p.cf2);
        }
      }
      const E ev = __builtin_bit_cast(E, St::rawStepD(fv, dz, d));
      out[d] = (out[d] & ~((Wd)(E)~E(0) << (BITS * j))) | ((Wd)ev << (BITS * j));
      if constexpr (CL) {
        const E eb = __builtin_bit_cast(E, St::rawStepD(fb, dz, d));
        bb[d] |= (Wd)eb << (BITS * j);
      }
    }
  }
  if constexpr (!CL) {
#pragma unroll
    for (int d = 0; d < Q; ++d)
      bb[d] = out[d];
  }
#pragma unroll
  for (int d = E0; d < E0 + 10; ++d) {
    if (cz(d) > 0 && lane == 63)
      xs[w][d - E0] = top32(out[d]);
    if (cz(d) < 0 && lane == 0)
      xs[w][d - E0] = (uint32_t)out[d];
  }
  __syncthreads();

  // stores to x + c_d, the z-shifted ones moved by +c_z: the lane's word of
  // the target row takes the values of cells z0 + j - c_z
  auto scatter = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
    const uint32_t fl = W ? FULL & ~cellsAt(field(0, 0), 0) : FULL; // fluid cells of the lane
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      if (d == C) {
        if (W)
          stNear(0, 0, C, out[C], fl);
        else
          st(wordPtr(0, 0, C), out[C]);
        continue;
      }
      // bounce-back: fluid cells whose target x + c_d is solid, own slot inv d
      if (W)
        stNear(0, 0, inv(d), bb[d], fl & cellsAt(field(cx(d), cy(d)), cz(d)));
      Wd sv = out[d];
      if (cz(d) > 0)
        sv = moveUp(sv, topOfBelow(sv, xs[wl][d - E0]));
      else if (cz(d) < 0)
        sv = moveDown(sv, bottomOfAbove(sv, xs[wr][d - E0]));
      // producers of the word's cells: z0 + j - c_z, skipped where solid
      if (W)
        stNear(cx(d), cy(d), d, sv, FULL & ~cellsAt(field(0, 0), -cz(d)));
      else
        st(wordPtr(cx(d), cy(d), d), sv);
    }
  };
  if (clean)
    scatter(std::false_type{});
  else
    scatter(std::true_type{});
}

template <int V, int NT, typename S>
static void launchRowGroup(const StepLaunch &L, hipStream_t s)
{
  Params<S> p           = makeParams<S>(L);
  const unsigned blocks = L.tiled ? tileBlocks(L.g, 1) * 4u : (unsigned)((uint64_t)L.g.nx * L.g.ny);
#define LAUNCH_RG(W)                                                                         \
  hipLaunchKernelGGL((kRowGroup<V, NT, S, W>), dim3(blocks), dim3(64 * L.g.segs), 0, s,     \
      (S *)L.src, (S *)L.dst, L.g.mask, p, L.tiled, L.remap)
  if (L.close) {
    // the closing step: one compilation per variant and storage, with the
    // default non-temporal bits (they do not change results)
    constexpr int NTC = sizeof(S) == 8 ? 3 : 1;
    constexpr int WC  = V == V_AA_ODD && sizeof(S) == 8 ? 1 : (NTC | ((NTC & 2) << 1));
    hipLaunchKernelGGL((kRowGroup<V, NTC, S, WC, true>), dim3(blocks), dim3(64 * L.g.segs), 0,
        s, (S *)L.src, (S *)L.dst, L.g.mask, p, L.tiled, L.remap);
    return;
  }
  if constexpr (V == V_AA_ODD && sizeof(S) == 8 && NT == 3) {
    switch (L.wallNt) {
    case 0: LAUNCH_RG(0); return;
    case 1: LAUNCH_RG(1); return;
    case 2: LAUNCH_RG(2); return;
    case 3: LAUNCH_RG(3); return;
    case 4: LAUNCH_RG(4); return;
    case 5: LAUNCH_RG(5); return;
    case 6: LAUNCH_RG(6); return;
    case 7: LAUNCH_RG(7); return;
    default: LAUNCH_RG(1); return;
    }
  }
  LAUNCH_RG((NT | ((NT & 2) << 1)));
#undef LAUNCH_RG
}

template <int NT, typename S, int WB>
static void launchRowGroupV(const StepLaunch &L, hipStream_t s)
{
  Params<S> p           = makeParams<S>(L);
  const unsigned blocks = (unsigned)((uint64_t)L.g.nx * L.g.ny);
  const unsigned lanes  = (unsigned)(L.g.nz / (WB / (int)sizeof(S))); // 64 per wavefront
  constexpr bool SR = std::is_same<S, __half>::value || std::is_same<S, __hip_bfloat16>::value;
#define LAUNCH_RGV(NT_, CL_, DI_)                                                               \
  hipLaunchKernelGGL((kRowGroupV<NT_, S, WB, CL_, DI_>), dim3(blocks), dim3(lanes), 0, s,        \
      (S *)L.src, L.g.mask, p, L.remap)
  if (SR && p.dither) {
    if (L.close)
      LAUNCH_RGV(1, true, SR);
    else
      LAUNCH_RGV(NT, false, SR);
  } else {
    if (L.close)
      LAUNCH_RGV(1, true, false);
    else
      LAUNCH_RGV(NT, false, false);
  }
#undef LAUNCH_RGV
}

template <int V, int NT, typename S>
static void launchRowSweep(const StepLaunch &L, hipStream_t s)
{
  Params<S> p           = makeParams<S>(L);
  const unsigned blocks = L.tiled ? tileBlocks(L.g, 1) : (unsigned)(((uint64_t)L.g.nx * L.g.ny + 3) / 4);
  constexpr bool SR = std::is_same<S, __half>::value || std::is_same<S, __hip_bfloat16>::value;
  constexpr int NTC  = sizeof(S) == 8 ? 3 : 1;
#define LAUNCH_RS(NT_, CL_, DI_)                                                                \
  hipLaunchKernelGGL((kRowSweep<V, NT_, S, CL_, DI_>), dim3(blocks), dim3(256), 0, s, (S *)L.src, \
      (S *)L.dst, L.g.mask, p, L.remap, L.tiled)
  if (SR && p.dither) {
    if (L.close)
      LAUNCH_RS(NTC, true, SR);
    else
      LAUNCH_RS(NT, false, SR);
  } else {
    if (L.close)
      LAUNCH_RS(NTC, true, false);
    else
      LAUNCH_RS(NT, false, false);
  }
#undef LAUNCH_RS
}

template <int V, bool AOS, int NT, typename S>
static void launchT(const StepLaunch &L, hipStream_t s)
{
  if constexpr (AOS) {
    launchAoS<V, NT, S>(L, s);
  } else {
    if constexpr (V == V_PUSH || V == V_AA_ODD) {
      if (L.rowSweep && L.g.nz % 64 == 0 && L.g.plane < (1ull << 32)) {
        if constexpr (V == V_AA_ODD && sizeof(S) < 8) {
          // whole words per lane (8 bytes; 4 bytes for 2-byte elements with
          // LBM_ROW_SWEEP=4), at most 8 wavefronts per row
          constexpr int C8 = 8 / (int)sizeof(S), C4 = 4 / (int)sizeof(S);
          if (L.rowSweep == 3 && !L.tiled && L.g.nz % (64 * C8) == 0 && L.g.nz <= 512 * C8 &&
              L.g.plane % C8 == 0)
            return launchRowGroupV<NT, S, 8>(L, s);
          if constexpr (sizeof(S) == 2) {
            if (L.rowSweep == 4 && !L.tiled && L.g.nz % (64 * C4) == 0 && L.g.nz <= 512 * C4 &&
                L.g.plane % C4 == 0)
              return launchRowGroupV<NT, S, 4>(L, s);
          }
        }
        if (L.rowSweep == 3 && L.g.segs <= 8)
          return launchRowGroup<V, NT, S>(L, s);
        return launchRowSweep<V, NT, S>(L, s);
      }
    }
    Params<S> p       = makeParams<S>(L);
    const int wpb     = L.tiled ? 4 : (L.wavesPerBlock == 8 ? 8 : 4);
    const unsigned blocks = L.tiled ? tileBlocks(L.g, L.g.segs) : rowBlocks(L.g, wpb);
    if constexpr (V == V_PUSH || V == V_AA_ODD) {
      if (L.close) {
        constexpr int NTC = sizeof(S) == 8 ? 3 : 1;
        hipLaunchKernelGGL((kStepRow<V, NTC, S, true>), dim3(blocks), dim3(64 * wpb), 0, s,
            (S *)L.src, (S *)L.dst, L.g.mask, p, L.remap, L.tiled);
        return;
      }
    }
    hipLaunchKernelGGL((kStepRow<V, NT, S>), dim3(blocks), dim3(64 * wpb), 0, s, (S *)L.src,
        (S *)L.dst, L.g.mask, p, L.remap, L.tiled);
  }
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
