// lbm_aux.hip -- passes around the step: body force at the end of a run,
// seeded initialisation, FP64 moment reductions, field comparison and a
// plain copy kernel.  One thread per cell (cell = (x*nY + y)*nZ + z); they
// run once per run or per check, not per step, so the logical-slot rule and
// layout are runtime arguments here.
#include <algorithm>

#include "lbm_impl.h"

namespace lbm {

__device__ __forceinline__ double weightOf(int d)
{
  return d == C ? 1.0 / 3.0 : (wclass(d) == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
}

// slot holding logical f_d of cell (x, y, z) under rule (src/lbm.c:193-294
// restated for the masked, halo-free layout)
__device__ __forceinline__ uint64_t logicalSlot(const Geom &g, int rule, int x, int y, int z,
    uint64_t cell, int d)
{
  auto sl = [&](uint64_t c, int k) -> uint64_t {
    return g.aos ? c * Q + (uint64_t)k : (uint64_t)k * g.plane + c;
  };
  if (rule == SR_LOCAL || d == C)
    return sl(cell, d);
  const int xn = wrapAdd(x, -cx(d), g.nx), yn = wrapAdd(y, -cy(d), g.ny),
            zn = wrapAdd(z, -cz(d), g.nz);
  const bool bb    = solidAt(g.mask, g.segs, g.ny, xn, yn, zn);
  const uint64_t n = ((uint64_t)xn * g.ny + yn) * g.nz + zn;
  if (rule == SR_PULL)
    return bb ? sl(cell, inv(d)) : sl(n, d);
  return bb ? sl(cell, d) : sl(n, inv(d));
}

struct CellPos {
  int x, y, z;
  uint64_t cell;
  bool valid;
};
// the thread's cell among [first, end) (logical order, x slowest); valid
// for in-range fluid cells
__device__ __forceinline__ CellPos cellPosIn(const Geom &g, uint64_t first, uint64_t end)
{
  CellPos c;
  c.cell  = first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  c.valid = c.cell < end;
  if (!c.valid)
    return c;
  c.z              = (int)(c.cell % g.nz);
  const uint64_t r = c.cell / g.nz;
  c.y              = (int)(r % g.ny);
  c.x              = (int)(r / g.ny);
  c.valid          = !solidAt(g.mask, g.segs, g.ny, c.x, c.y, c.z);
  return c;
}
__device__ __forceinline__ CellPos cellPos(const Geom &g)
{
  return cellPosIn(g, 0, (uint64_t)g.nx * g.ny * g.nz);
}

// ---- body force pass -------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void kForce(S *a, Geom g, int rule, float scale, float inv,
    double f1, double f2, int dither, uint32_t seq)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPos(g);
  if (!c.valid)
    return;
  const Dither dz = ditherOf(c.cell, seq * 0x9e3779b9u, dither);
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    if (cx(d) == 0)
      continue;
    const R fd       = (R)(wclass(d) == 1 ? f1 : f2);
    const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
    R v              = St::load(a[i], (R)inv);
    v += cx(d) > 0 ? fd : -fd;
    a[i] = St::storeStep(v, (R)scale, ditherDir(dz, d));
  }
}

__global__ void kBumpSeq(uint32_t *counter, uint32_t n)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    counter[0] += n;
}

void launchBumpSeq(uint32_t *counter, uint32_t n, hipStream_t s)
{
  hipLaunchKernelGGL(kBumpSeq, dim3(1), dim3(64), 0, s, counter, n);
}

void launchForce(const AuxLaunch &A, double f1, double f2, hipStream_t s)
{
  const uint64_t n      = (uint64_t)A.g.nx * A.g.ny * A.g.nz;
  const unsigned blocks = (unsigned)((n + 255) / 256);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kForce<T>, dim3(blocks), dim3(256), 0, s, (T *)A.pdf, A.g, A.rule, A.scale, \
      A.invScale, f1, f2, A.dither, A.seq)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
}

// ---- Zou-He pressure inlet / outlet -------------------------------------------
// kernelComputeBoundaryConditions (src/lbm.c:676-785) on the logical PDFs.
// One thread per (y, z) of the x faces, y and z in 1 .. n-2 as in the
// reference's loops; face[y*nZ + z] bit 0 marks (0, y, z) an inlet cell,
// bit 1 marks (nX-1, y, z) an outlet cell.  The thread treats the inlet
// cell and then the outlet cell, the reference's order, so a lattice one
// cell long sees the same sequence.  Distinct threads touch distinct
// logical (cell, direction) pairs, and the logical-slot map is one to one
// on fluid cells, so they never share an element.
// pending != 0: the state lacks the body force of the step just run (the
// next step's load or the closing force pass adds 3 w_d c_dx g to every
// fluid cell), so the values read get it added, as the completed state
// holds them, and the five values written get it subtracted.
// The arithmetic keeps the reference's grouping; contraction is off.
template <typename S, typename Cm>
__global__ __launch_bounds__(256) void kZouHe(S *a, Geom g, int rule, const uint8_t *__restrict__ face,
    Cm rhoIn, Cm rhoOut, Cm f1, Cm f2, int pending, double rhoRef, float scale, float inv,
    int dither, const uint32_t *seqBase, uint32_t seq)
{
#pragma clang fp contract(off)
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)g.ny * g.nz)
    return;
  const int y = (int)(t / (uint64_t)g.nz), z = (int)(t % (uint64_t)g.nz);
  if (y < 1 || y > g.ny - 2 || z < 1 || z > g.nz - 2)
    return;
  const unsigned bits = face[t];
  if (bits == 0)
    return;
  const uint32_t seqv = (*seqBase + seq) * 0x9e3779b9u ^ 0x2545f491u;
  auto force = [&](int d) -> Cm {
    return cx(d) == 0 ? Cm(0) : (cx(d) > 0 ? (wclass(d) == 1 ? f1 : f2) : -(wclass(d) == 1 ? f1 : f2));
  };
  const Cm third = Cm(1.0) / Cm(3.0), fourth = Cm(1.0) / Cm(4.0), sixth = Cm(1.0) / Cm(6.0);
  for (int side = 0; side < 2; ++side) {
    if (!(bits & (1u << side)))
      continue;
    const int x         = side == 0 ? 0 : g.nx - 1;
    const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
    const Dither dz     = ditherOf(cell, seqv, dither);
    auto get = [&](int d) -> Cm {
      const uint64_t i = logicalSlot(g, rule, x, y, z, cell, d);
      Cm v = St::scaled ? (Cm)(weightOf(d) * rhoRef) + (Cm)St::load(a[i], (R)inv)
                        : (Cm)St::load(a[i], (R)inv);
      if (pending)
        v += force(d);
      return v;
    };
    auto put = [&](int d, Cm v) {
      if (pending)
        v -= force(d);
      const uint64_t i = logicalSlot(g, rule, x, y, z, cell, d);
      a[i] = St::scaled ? St::storeStep((R)((double)v - weightOf(d) * rhoRef), (R)scale, ditherDir(dz, d))
                        : St::storeStep((R)v, (R)scale, ditherDir(dz, d));
    };
    const Cm fC = get(C), fT = get(T), fB = get(B), fS = get(lbm::S), fN = get(N);
    const Cm fTS = get(TS), fBS = get(BS), fTN = get(TN), fBN = get(BN);
    if (side == 0) {
      const Cm fSW = get(SW), fTW = get(TW), fW = get(W), fBW = get(BW), fNW = get(NW);
      const Cm dens = rhoIn;
      const Cm ux   = Cm(1.0) - (fC + (fT + fB + fS + fN) + (fTS + fBS + fTN + fBN) +
                                  Cm(2.0) * (fSW + fTW + fW + fBW + fNW)) * (Cm(1.0) / rhoIn);
      const Cm indep = sixth * dens * ux;
      put(E, fW + third * dens * ux);
      put(NE, fSW - fourth * (fN - fS) + indep);
      put(SE, fNW + fourth * (fN - fS) + indep);
      put(TE, fBW - fourth * (fT - fB) + indep);
      put(BE, fTW + fourth * (fT - fB) + indep);
    } else {
      const Cm fNE = get(NE), fBE = get(BE), fE = get(E), fTE = get(TE), fSE = get(SE);
      const Cm dens = rhoOut;
      const Cm ux   = Cm(-1.0) + (fC + (fT + fB + fS + fN) + (fTS + fBS + fTN + fBN) +
                                   Cm(2.0) * (fNE + fBE + fE + fTE + fSE)) * (Cm(1.0) / rhoOut);
      const Cm indep = sixth * dens * ux;
      put(W, fE - third * dens * ux);
      put(SW, fNE + fourth * (fN - fS) - indep);
      put(NW, fSE - fourth * (fN - fS) - indep);
      put(BW, fTE + fourth * (fT - fB) - indep);
      put(TW, fBE - fourth * (fT - fB) - indep);
    }
  }
}

void launchZouHe(const AuxLaunch &A, const ZouHeLaunch &Z, hipStream_t s)
{
  const uint64_t n      = (uint64_t)A.g.ny * A.g.nz;
  const unsigned blocks = (unsigned)((n + 255) / 256);
#define LAUNCH(T, Cm)                                                                      \
  hipLaunchKernelGGL((kZouHe<T, Cm>), dim3(blocks), dim3(256), 0, s, (T *)A.pdf, A.g, A.rule,   \
      Z.face, (Cm)Z.rhoIn, (Cm)Z.rhoOut, (Cm)Z.f1, (Cm)Z.f2, Z.pending, A.rhoRef, A.scale,    \
      A.invScale, A.dither, Z.seqBase, Z.seq)
  switch (A.storage) {
  case ST_F64: LAUNCH(double, double); break;
  case ST_F32: LAUNCH(float, float); break;
  case ST_F32DEV: LAUNCH(f32dev, double); break;
  case ST_F16: LAUNCH(__half, double); break;
  default: LAUNCH(__hip_bfloat16, double); break;
  }
#undef LAUNCH
}

// ---- geometry (src/lbm.c:52-187, plus spheres) -------------------------------
// One thread per cell in LatticeDescType order (x fastest): obstacle (0) on
// the boundary layer of a non-periodic axis, outside the pipe's radius, or
// inside a block; fluid (1) elsewhere.  The pipe's test is the reference's
// expression in double, contraction off, so the cells match a host loop.
__global__ __launch_bounds__(256) void kGeometryCells(int *__restrict__ lat, int nx, int ny,
    int nz, int px, int py, int pz, int pipe, int blocks)
{
#pragma clang fp contract(off)
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint64_t)nx * ny * nz)
    return;
  const int x = (int)(t % (uint64_t)nx);
  const int y = (int)((t / (uint64_t)nx) % (uint64_t)ny);
  const int z = (int)(t / ((uint64_t)nx * ny));
  bool obst = (!px && (x == 0 || x == nx - 1)) || (!py && (y == 0 || y == ny - 1)) ||
              (!pz && (z == 0 || z == nz - 1));
  if (pipe) {
    const double cz = nz / 2.0 - 0.5, cy = ny / 2.0 - 0.5;
    const double dmin = (double)(ny < nz ? ny : nz);
    const double r2   = (dmin / 2 - 1) * (dmin / 2 - 1);
    if ((z - cz) * (z - cz) + (y - cy) * (y - cy) >= r2)
      obst = true;
  } else if (blocks > 0) {
    const int nb = blocks;
    if ((z % (2 * nb)) >= nb && (y % (2 * nb)) >= nb && (x % (2 * nb)) >= nb)
      obst = true;
  }
  lat[t] = obst ? 0 : 1;
}

// One workgroup per sphere: its threads walk the sphere's bounding box
// (lo..hi per axis).  A cell whose centre lies within the radius
// (dx^2 + dy^2 + dz^2 <= r^2 in double, contraction off) becomes an
// obstacle; a box cell beyond the lattice wraps through a periodic axis and
// is dropped on a non-periodic one.  Overlapping spheres store the same
// value into a shared cell.
__global__ __launch_bounds__(256) void kCarveSpheres(int *__restrict__ lat, int nx, int ny,
    int nz, int px, int py, int pz, const double *__restrict__ sph, const int *__restrict__ box)
{
#pragma clang fp contract(off)
  const int k = blockIdx.x;
  const double c0 = sph[4 * k], c1 = sph[4 * k + 1], c2 = sph[4 * k + 2], r = sph[4 * k + 3];
  const int lo0 = box[6 * k], lo1 = box[6 * k + 1], lo2 = box[6 * k + 2];
  const int64_t ex = (int64_t)box[6 * k + 3] - lo0 + 1, ey = (int64_t)box[6 * k + 4] - lo1 + 1,
                ez = (int64_t)box[6 * k + 5] - lo2 + 1;
  const int64_t total = ex * ey * ez;
  const int dims[3] = { nx, ny, nz }, per[3] = { px, py, pz };
  auto wrapAxis = [&](int v, int a, int &out) -> bool {
    if (v >= 0 && v < dims[a]) {
      out = v;
      return true;
    }
    if (!per[a])
      return false;
    out = ((v % dims[a]) + dims[a]) % dims[a];
    return true;
  };
  for (int64_t i = threadIdx.x; i < total; i += blockDim.x) {
    const int x = lo0 + (int)(i % ex), y = lo1 + (int)((i / ex) % ey), z = lo2 + (int)(i / (ex * ey));
    const double dx = x - c0, dy = y - c1, dz = z - c2;
    if (dx * dx + dy * dy + dz * dz > r * r)
      continue;
    int wx, wy, wz;
    if (!wrapAxis(z, 2, wz) || !wrapAxis(y, 1, wy) || !wrapAxis(x, 0, wx))
      continue;
    lat[((uint64_t)wz * ny + wy) * nx + wx] = 0;
  }
}

// number of non-obstacle flags, one 64-bit atomic per workgroup (exact)
__global__ __launch_bounds__(256) void kCountFluid(const int *__restrict__ lat, uint64_t n,
    unsigned long long *count)
{
  __shared__ unsigned int part[4];
  unsigned int c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
    c += lat[i] != 0;
  for (int o = 32; o > 0; o >>= 1)
    c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0)
    part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(count, (unsigned long long)part[0] + part[1] + part[2] + part[3]);
}

void launchGeometry(int *lat, const int dims[3], const int per[3], int pipe, int blocks,
    int nSpheres, const double *sph, const int *box, unsigned long long *count, hipStream_t s)
{
  const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(kGeometryCells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lat,
      dims[0], dims[1], dims[2], per[0], per[1], per[2], pipe, blocks);
  if (nSpheres > 0)
    hipLaunchKernelGGL(kCarveSpheres, dim3((unsigned)nSpheres), dim3(256), 0, s, lat, dims[0],
        dims[1], dims[2], per[0], per[1], per[2], sph, box);
  const uint64_t groups = std::min<uint64_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(kCountFluid, dim3((unsigned)std::max<uint64_t>(groups, 1)), dim3(256), 0, s,
      lat, n, count);
}

// ---- x velocity of one x-plane (kernelVerify's profile) -----------------------
// ux[y*nZ + z] = j_x of the logical PDFs of (x, y, z), summed in the order
// of src/lbm.c:981-983 in the compute type Cm (PdfType for native storage:
// what GetNode would hand the host loop); 0 on obstacle cells
template <typename S, typename Cm>
__global__ __launch_bounds__(256) void kPlaneUx(const S *a, Geom g, int rule, int x,
    double rhoRef, float inv, double *ux)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint64_t)g.ny * g.nz)
    return;
  const int y = (int)(t / (uint64_t)g.nz), z = (int)(t % (uint64_t)g.nz);
  if (solidAt(g.mask, g.segs, g.ny, x, y, z)) {
    ux[t] = 0.0;
    return;
  }
  const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
  auto f = [&](int d) -> Cm {
    const uint64_t i = logicalSlot(g, rule, x, y, z, cell, d);
    return St::scaled ? (Cm)(weightOf(d) * rhoRef + (double)St::load(a[i], (R)inv))
                      : (Cm)St::load(a[i], (R)inv);
  };
  const Cm v = f(E) + f(NE) + f(SE) + f(TE) + f(BE) - f(W) - f(NW) - f(SW) - f(TW) - f(BW);
  ux[t] = (double)v;
}

void launchPlaneUx(const AuxLaunch &A, int x, double *ux, hipStream_t s)
{
  const uint64_t n      = (uint64_t)A.g.ny * A.g.nz;
  const unsigned blocks = (unsigned)((n + 255) / 256);
#define LAUNCH(T, Cm)                                                                      \
  hipLaunchKernelGGL((kPlaneUx<T, Cm>), dim3(blocks), dim3(256), 0, s, (const T *)A.pdf, A.g,    \
      A.rule, x, A.rhoRef, A.invScale, ux)
  switch (A.storage) {
  case ST_F64: LAUNCH(double, double); break;
  case ST_F32: LAUNCH(float, float); break;
  case ST_F32DEV: LAUNCH(f32dev, double); break;
  case ST_F16: LAUNCH(__half, double); break;
  default: LAUNCH(__hip_bfloat16, double); break;
  }
#undef LAUNCH
}

// ---- seeded initialisation -------------------------------------------------
// splitmix64 finalizer (Steele, Lea & Flood 2014) chained over (seed, a, b)
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b)
{
  return mix64(mix64(mix64(seed) ^ a) ^ b);
}

template <typename S>
__global__ __launch_bounds__(256) void kInit(S *a, Geom g, int rule, int mode, uint64_t seed,
    double rho, double amp, double rhoRef, float scale)
{
#pragma clang fp contract(off)
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPos(g);
  if (!c.valid)
    return;
  const uint64_t lcell = c.cell; // logical cell index: same ordering
  for (int d = 0; d < Q; ++d) {
    double e = 0.0;
    if (mode == 2) {
      const double u = (double)(hash3(seed, lcell, (uint64_t)d) >> 11) * (1.0 / 9007199254740992.0);
      e              = (2.0 * u - 1.0) * amp;
    }
    double v;
    if (St::scaled)
      v = mode == 0 ? -weightOf(d) * rhoRef : weightOf(d) * ((rho - rhoRef) + e);
    else
      v = mode == 0 ? 0.0 : weightOf(d) * (rho + e);
    const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
    a[i]             = St::store((R)v, (R)scale);
  }
}

void launchInit(const AuxLaunch &A, int mode, uint64_t seed, double rho, double amp,
    hipStream_t s)
{
  const uint64_t n      = (uint64_t)A.g.nx * A.g.ny * A.g.nz;
  const unsigned blocks = (unsigned)((n + 255) / 256);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kInit<T>, dim3(blocks), dim3(256), 0, s, (T *)A.pdf, A.g, A.rule, mode, seed, \
      rho, amp, A.rhoRef, A.scale)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
}

// ---- logical gather / scatter ------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void kGather(const S *a, Geom g, int rule, double rhoRef,
    float inv, double *out, uint64_t first, uint64_t end)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPosIn(g, first, end);
  if (c.cell >= end)
    return;
  for (int d = 0; d < Q; ++d) {
    double v = 0.0;
    if (c.valid) {
      const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
      v = St::scaled ? weightOf(d) * rhoRef + (double)St::load(a[i], (R)inv)
                     : (double)St::load(a[i], (R)inv);
    }
    out[(c.cell - first) * Q + d] = v;
  }
}

template <typename S>
__global__ __launch_bounds__(256) void kScatter(S *a, Geom g, int rule, double rhoRef,
    float scale, const double *in, uint64_t first, uint64_t end)
{
#pragma clang fp contract(off)
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPosIn(g, first, end);
  if (!c.valid)
    return;
  for (int d = 0; d < Q; ++d) {
    const double v   = in[(c.cell - first) * Q + d];
    const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
    a[i] = St::scaled ? St::store((R)(v - weightOf(d) * rhoRef), (R)scale) : St::store((R)v, (R)scale);
  }
}

static void cellRange(const Geom &g, uint64_t &first, uint64_t &count)
{
  const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
  if (first > n)
    first = n;
  if (count > n - first)
    count = n - first;
}

void launchGather(const AuxLaunch &A, double *out, hipStream_t s, uint64_t first, uint64_t count)
{
  cellRange(A.g, first, count);
  if (count == 0)
    return;
  const unsigned blocks = (unsigned)((count + 255) / 256);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kGather<T>, dim3(blocks), dim3(256), 0, s, (const T *)A.pdf, A.g, A.rule,    \
      A.rhoRef, A.invScale, out, first, first + count)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
}

void launchScatter(const AuxLaunch &A, const double *in, hipStream_t s, uint64_t first,
    uint64_t count)
{
  cellRange(A.g, first, count);
  if (count == 0)
    return;
  const unsigned blocks = (unsigned)((count + 255) / 256);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kScatter<T>, dim3(blocks), dim3(256), 0, s, (T *)A.pdf, A.g, A.rule,         \
      A.rhoRef, A.scale, in, first, first + count)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
}

// ---- per-cell density and momentum -----------------------------------------
// out[(cell - first)*4 + {0,1,2,3}] = rho, jx, jy, jz of the logical PDFs
// (the incompressible model's velocity is j); 0 on obstacle cells
template <typename S>
__global__ __launch_bounds__(256) void kCellMoments(const S *a, Geom g, int rule, double rhoRef,
    float inv, double *out, uint64_t first, uint64_t end)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPosIn(g, first, end);
  if (c.cell >= end)
    return;
  double m[4] = { 0.0, 0.0, 0.0, 0.0 };
  if (c.valid) {
    for (int d = 0; d < Q; ++d) {
      const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
      const double v   = (double)St::load(a[i], (R)inv);
      m[0] += v;
      m[1] += cx(d) * v;
      m[2] += cy(d) * v;
      m[3] += cz(d) * v;
    }
    m[0] += St::scaled ? rhoRef : 0.0; // sum_d w_d rhoRef
  }
  for (int k = 0; k < 4; ++k)
    out[(c.cell - first) * 4 + k] = m[k];
}

void launchCellMoments(const AuxLaunch &A, double *out, hipStream_t s, uint64_t first,
    uint64_t count)
{
  cellRange(A.g, first, count);
  if (count == 0)
    return;
  const unsigned blocks = (unsigned)((count + 255) / 256);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kCellMoments<T>, dim3(blocks), dim3(256), 0, s, (const T *)A.pdf, A.g,      \
      A.rule, A.rhoRef, A.invScale, out, first, first + count)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
}

// ---- FP64 reductions -------------------------------------------------------
// wavefront sum by xor-shuffles (DPP / ds_swizzle lane exchange), then
// the four wavefronts of the workgroup through LDS, in a fixed order
__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double waveMax(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// reduce k values per thread across a 256-lane workgroup; lane 0 of
// wavefront 0 ends with the totals.  isMax selects max instead of sum
// for value index kmax (-1: none).
template <int K>
__device__ __forceinline__ void blockReduce(double (&v)[K], int kmax)
{
  __shared__ double part[4][K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k)
    v[k] = k == kmax ? waveMax(v[k]) : waveSum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k)
      part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = part[0][k];
      for (int j = 1; j < 4; ++j)
        t = k == kmax ? fmax(t, part[j][k]) : t + part[j][k];
      v[k] = t;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void kMoments(const S *a, Geom g, int rule, float inv,
    double rhoRef, double *partials, double *field)
{
  using R  = typename Store<S>::R;
  using St = Store<S>;
  const CellPos c = cellPos(g);
  double m[4]     = { 0.0, 0.0, 0.0, 0.0 };
  const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
  if (c.valid) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const uint64_t i = logicalSlot(g, rule, c.x, c.y, c.z, c.cell, d);
      const double v   = (double)St::load(a[i], (R)inv);
      m[0] += v;
      m[1] += cx(d) * v;
      m[2] += cy(d) * v;
      m[3] += cz(d) * v;
    }
  }
  if (field != nullptr && c.cell < n) {
    // the field holds the density itself: the deviation forms add back
    // sum_d w_d RhoRef = RhoRef (j needs nothing: sum_d c_d w_d = 0)
    const double nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 4; ++k)
      field[k * n + c.cell] = c.valid ? m[k] + (k == 0 && St::scaled ? rhoRef : 0.0) : nan;
  }
  blockReduce<4>(m, -1);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      partials[4 * (uint64_t)blockIdx.x + k] = m[k];
}

int momentsBlocks(const Geom &g)
{
  const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
  return (int)((n + 255) / 256);
}

int launchMoments(const AuxLaunch &A, double *partials, double *field, hipStream_t s)
{
  const int blocks = momentsBlocks(A.g);
#define LAUNCH(T)                                                                          \
  hipLaunchKernelGGL(kMoments<T>, dim3(blocks), dim3(256), 0, s, (const T *)A.pdf, A.g, A.rule,   \
      A.invScale, A.rhoRef, partials, field)
  switch (A.storage) {
  case ST_F64: LAUNCH(double); break;
  case ST_F32: LAUNCH(float); break;
  case ST_F32DEV: LAUNCH(f32dev); break;
  case ST_F16: LAUNCH(__half); break;
  default: LAUNCH(__hip_bfloat16); break;
  }
#undef LAUNCH
  return blocks;
}

// final pass: one workgroup sums the partials in a fixed order
template <int K>
__global__ __launch_bounds__(256) void kReduceK(const double *partials, int n, double *out,
    int kmax)
{
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    v[k] = k == kmax ? 0.0 : 0.0;
  for (int i = threadIdx.x; i < n; i += 256)
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = k == kmax ? fmax(v[k], partials[(uint64_t)K * i + k]) : v[k] + partials[(uint64_t)K * i + k];
  blockReduce<K>(v, kmax);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k)
      out[k] = v[k];
}

void launchReduce4(const double *partials, int n, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(kReduceK<4>, dim3(1), dim3(256), 0, s, partials, n, out, -1);
}

// {sum |jA - jB|^2, sum |jB|^2, max |jA - jB|, sum (rhoA - rhoB)^2} over
// cells where both fields hold numbers
__global__ __launch_bounds__(256) void kCompare(const double *fa, const double *fb, uint64_t n,
    double *partials)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[4]      = { 0.0, 0.0, 0.0, 0.0 };
  if (i < n && !__builtin_isnan(fa[i]) && !__builtin_isnan(fb[i])) {
    const double dx = fa[n + i] - fb[n + i], dy = fa[2 * n + i] - fb[2 * n + i],
                 dz = fa[3 * n + i] - fb[3 * n + i];
    const double d2 = dx * dx + dy * dy + dz * dz;
    v[0]            = d2;
    v[1] = fb[n + i] * fb[n + i] + fb[2 * n + i] * fb[2 * n + i] + fb[3 * n + i] * fb[3 * n + i];
    v[2] = sqrt(d2);
    v[3] = (fa[i] - fb[i]) * (fa[i] - fb[i]);
  }
  blockReduce<4>(v, 2);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      partials[4 * (uint64_t)blockIdx.x + k] = v[k];
}

int launchCompareFields(const double *fa, const double *fb, uint64_t n, double *partials,
    hipStream_t s)
{
  const int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(kCompare, dim3(blocks), dim3(256), 0, s, fa, fb, n, partials);
  return blocks;
}

void launchReduceCompare(const double *partials, int n, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(kReduceK<4>, dim3(1), dim3(256), 0, s, partials, n, out, 2);
}

// ---- comparison of two lattices without field buffers -------------------------
// The moments of kMoments, per cell of two lattices of the same dimensions
// (any storage forms), and the four terms of kCompare, in one pass: each
// workgroup sums its chunk of `chunk` consecutive cells into one slot of
// partials.  A cell counts when it is fluid in both lattices.
__device__ __forceinline__ double rawLoad(const void *pdf, int storage, uint64_t i, float inv)
{
  switch (storage) {
  case ST_F64: return ((const double *)pdf)[i];
  case ST_F32: return (double)((const float *)pdf)[i];
  case ST_F32DEV: return (double)(((const float *)pdf)[i] * inv);
  case ST_F16: return (double)(__half2float(((const __half *)pdf)[i]) * inv);
  default: return (double)(__bfloat162float(((const __hip_bfloat16 *)pdf)[i]) * inv);
  }
}

struct CompareSide {
  const void *pdf;
  Geom g;
  int storage, rule;
  double rhoRef; // added to rho for the deviation forms
  float inv;
};

__device__ __forceinline__ void cellMoments(const CompareSide &L, int x, int y, int z,
    uint64_t cell, double (&m)[4])
{
  m[0] = m[1] = m[2] = m[3] = 0.0;
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const double v = rawLoad(L.pdf, L.storage, logicalSlot(L.g, L.rule, x, y, z, cell, d), L.inv);
    m[0] += v;
    m[1] += cx(d) * v;
    m[2] += cy(d) * v;
    m[3] += cz(d) * v;
  }
  m[0] += L.storage >= ST_F32DEV ? L.rhoRef : 0.0;
}

__global__ __launch_bounds__(256) void kCompareLattices(CompareSide A, CompareSide B, uint64_t n,
    uint64_t chunk, double *partials)
{
  double v[4] = { 0.0, 0.0, 0.0, 0.0 };
  const uint64_t first = (uint64_t)blockIdx.x * chunk;
  const uint64_t end   = first + chunk < n ? first + chunk : n;
  for (uint64_t c = first + threadIdx.x; c < end; c += 256u) {
    const int z      = (int)(c % A.g.nz);
    const uint64_t r = c / A.g.nz;
    const int y = (int)(r % A.g.ny), x = (int)(r / A.g.ny);
    if (solidAt(A.g.mask, A.g.segs, A.g.ny, x, y, z) || solidAt(B.g.mask, B.g.segs, B.g.ny, x, y, z))
      continue;
    double a[4], b[4];
    cellMoments(A, x, y, z, c, a);
    cellMoments(B, x, y, z, c, b);
    const double dx = a[1] - b[1], dy = a[2] - b[2], dz = a[3] - b[3];
    const double d2 = dx * dx + dy * dy + dz * dz;
    v[0] += d2;
    v[1] += b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
    v[2] = fmax(v[2], sqrt(d2));
    v[3] += (a[0] - b[0]) * (a[0] - b[0]);
  }
  blockReduce<4>(v, 2);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      partials[4 * (uint64_t)blockIdx.x + k] = v[k];
}

void launchCompareLattices(const AuxLaunch &A, const AuxLaunch &B, uint64_t chunk,
    double *partials, hipStream_t s)
{
  const uint64_t n = (uint64_t)A.g.nx * A.g.ny * A.g.nz;
  const CompareSide a{ A.pdf, A.g, A.storage, A.rule, A.rhoRef, A.invScale };
  const CompareSide b{ B.pdf, B.g, B.storage, B.rule, B.rhoRef, B.invScale };
  hipLaunchKernelGGL(kCompareLattices, dim3((unsigned)((n + chunk - 1) / chunk)), dim3(256), 0, s,
      a, b, n, chunk, partials);
}

// ---- copy ------------------------------------------------------------------
// 4 x 16 B per thread, non-temporal: the streaming pattern of the step
// kernels without the arithmetic
typedef double v2d __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void kCopy(const v2d *__restrict__ src, v2d *__restrict__ dst,
    uint64_t n2)
{
  const uint64_t base = (uint64_t)blockIdx.x * 1024u + threadIdx.x;
  v2d v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + 256u * k;
    if (i < n2)
      v[k] = __builtin_nontemporal_load(&src[i]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + 256u * k;
    if (i < n2)
      __builtin_nontemporal_store(v[k], &dst[i]);
  }
}

void launchCopy(const double *src, double *dst, uint64_t n, hipStream_t s)
{
  const uint64_t n2     = n / 2;
  const unsigned blocks = (unsigned)((n2 + 1023) / 1024);
  hipLaunchKernelGGL(kCopy, dim3(blocks), dim3(256), 0, s, (const v2d *)src, (v2d *)dst, n2);
}

// ---- copy kernels of other shapes (the achievable-bandwidth sweep) -------------
// V 16-byte vectors per lane, 256 lanes per workgroup, workgroup b copying
// V * 256 consecutive vectors; NT: non-temporal loads and stores;
// persistent: a grid of `groups` workgroups strides over the array.  All
// loads of a lane are issued before its stores.
template <int V, bool NT>
__global__ __launch_bounds__(256) void kCopyShape(const v2d *__restrict__ src, v2d *__restrict__ dst,
    uint64_t n2, int persistent)
{
  const uint64_t tile   = (uint64_t)V * 256u;
  const uint64_t stride = persistent ? (uint64_t)gridDim.x * tile : 0;
  for (uint64_t base = (uint64_t)blockIdx.x * tile + threadIdx.x; base < n2;
       base += persistent ? stride : n2) {
    v2d v[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const uint64_t i = base + 256u * k;
      if (i < n2)
        v[k] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const uint64_t i = base + 256u * k;
      if (i < n2) {
        if (NT)
          __builtin_nontemporal_store(v[k], &dst[i]);
        else
          dst[i] = v[k];
      }
    }
  }
}

bool launchCopyShape(const double *src, double *dst, uint64_t n, int vectors, int nt,
    int persistentGroups, hipStream_t s)
{
  const uint64_t n2     = n / 2;
  const uint64_t tile   = (uint64_t)vectors * 256u;
  const unsigned blocks = persistentGroups > 0 ? (unsigned)persistentGroups
                                               : (unsigned)((n2 + tile - 1) / tile);
  const int persistent  = persistentGroups > 0;
  const v2d *a = (const v2d *)src;
  v2d *b       = (v2d *)dst;
#define LAUNCH(V, NTV)                                                                     \
  hipLaunchKernelGGL((kCopyShape<V, NTV>), dim3(blocks), dim3(256), 0, s, a, b, n2, persistent)
  switch (vectors * 2 + (nt ? 1 : 0)) {
  case 2: LAUNCH(1, false); return true;
  case 3: LAUNCH(1, true); return true;
  case 4: LAUNCH(2, false); return true;
  case 5: LAUNCH(2, true); return true;
  case 8: LAUNCH(4, false); return true;
  case 9: LAUNCH(4, true); return true;
  case 16: LAUNCH(8, false); return true;
  case 17: LAUNCH(8, true); return true;
  default: return false;
  }
#undef LAUNCH
}

// ---- porous solid-fraction field (differentiation form) ---------------------
// The smoothing passes compute each output as the host formula does, tap by
// tap from -rad to rad without fused multiply-adds, so the smoothed field
// is the one a serial host loop gives, bit for bit.
__global__ __launch_bounds__(256) void kPorousNoise(double *s, uint64_t n, uint64_t seed)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i < n)
    s[i] = (double)(hash3(seed, i, 0) >> 11) * (1.0 / 9007199254740992.0);
}

// one axis of the separable periodic Gaussian, field [x][y][z]
template <int AXIS>
__global__ __launch_bounds__(256) void kSmoothAxis(const double *__restrict__ in,
    double *__restrict__ out, int nx, int ny, int nz, const double *__restrict__ kern, int rad)
{
#pragma clang fp contract(off)
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= (uint64_t)nx * ny * nz)
    return;
  const int z = (int)(i % (uint64_t)nz);
  const int y = (int)((i / (uint64_t)nz) % (uint64_t)ny);
  const int x = (int)(i / ((uint64_t)nz * (uint64_t)ny));
  const int n      = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
  const int p      = AXIS == 0 ? x : (AXIS == 1 ? y : z);
  const int64_t st = AXIS == 0 ? (int64_t)ny * nz : (AXIS == 1 ? (int64_t)nz : 1);
  const int64_t base = (int64_t)i - (int64_t)p * st;
  double acc = 0.0;
  for (int k = -rad; k <= rad; ++k) {
    int q = (p + k) % n;
    q     = q < 0 ? q + n : q;
    acc += kern[k + rad] * in[base + (int64_t)q * st];
  }
  out[i] = acc;
}

__global__ __launch_bounds__(256) void kLogistic(double *s, uint64_t n, double mean, double sd,
    double a, double b)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= n)
    return;
  const double z = sd > 0.0 ? (s[i] - mean) / sd : 0.0;
  s[i]           = 1.0 / (1.0 + exp(-(a * z + b)));
}

void launchPorousNoise(double *s, uint64_t n, uint64_t seed, hipStream_t st)
{
  hipLaunchKernelGGL(kPorousNoise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, n, seed);
}

void launchSmoothAxis(int axis, const double *in, double *out, const int dims[3],
    const double *kern, int rad, hipStream_t st)
{
  const uint64_t n      = (uint64_t)dims[0] * dims[1] * dims[2];
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (axis == 0)
    hipLaunchKernelGGL(kSmoothAxis<0>, dim3(blocks), dim3(256), 0, st, in, out, dims[0], dims[1],
        dims[2], kern, rad);
  else if (axis == 1)
    hipLaunchKernelGGL(kSmoothAxis<1>, dim3(blocks), dim3(256), 0, st, in, out, dims[0], dims[1],
        dims[2], kern, rad);
  else
    hipLaunchKernelGGL(kSmoothAxis<2>, dim3(blocks), dim3(256), 0, st, in, out, dims[0], dims[1],
        dims[2], kern, rad);
}

void launchLogistic(double *s, uint64_t n, double mean, double sd, double a, double b,
    hipStream_t st)
{
  hipLaunchKernelGGL(kLogistic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, n, mean, sd,
      a, b);
}

// ---- access-pattern ceiling --------------------------------------------------
// The memory traffic of one SoA step without its arithmetic: lane `cell`
// loads element `cell` of P planes (non-temporal) and stores each value
// plus one, to plane P-1-q of the same array (in place, the AA step's
// traffic) or to plane q of a second array (push and pull).
template <typename T, int P>
__global__ __launch_bounds__(512) void kPattern(const T *src, T *dst, uint64_t plane, uint64_t cells,
    int inPlace)
{
  const uint64_t c = (uint64_t)blockIdx.x * 512u + threadIdx.x;
  if (c >= cells)
    return;
  T v[P];
#pragma unroll
  for (int q = 0; q < P; ++q)
    v[q] = __builtin_nontemporal_load(&src[c + (uint64_t)q * plane]);
#pragma unroll
  for (int q = 0; q < P; ++q)
    __builtin_nontemporal_store((T)(v[q] + 1), &dst[c + (uint64_t)(inPlace ? P - 1 - q : q) * plane]);
}

template <typename T>
static bool launchPatternT(int planes, const void *src, void *dst, uint64_t cells, bool inPlace,
    hipStream_t s)
{
  const unsigned blocks = (unsigned)((cells + 511) / 512);
  const T *a            = (const T *)src;
  T *b                  = (T *)dst;
  switch (planes) {
  case 1: hipLaunchKernelGGL((kPattern<T, 1>), dim3(blocks), dim3(512), 0, s, a, b, cells, cells, (int)inPlace); return true;
  case 2: hipLaunchKernelGGL((kPattern<T, 2>), dim3(blocks), dim3(512), 0, s, a, b, cells, cells, (int)inPlace); return true;
  case 4: hipLaunchKernelGGL((kPattern<T, 4>), dim3(blocks), dim3(512), 0, s, a, b, cells, cells, (int)inPlace); return true;
  case 8: hipLaunchKernelGGL((kPattern<T, 8>), dim3(blocks), dim3(512), 0, s, a, b, cells, cells, (int)inPlace); return true;
  case 19: hipLaunchKernelGGL((kPattern<T, 19>), dim3(blocks), dim3(512), 0, s, a, b, cells, cells, (int)inPlace); return true;
  default: return false;
  }
}

bool launchPattern(int planes, int elementBytes, const void *src, void *dst, uint64_t cells,
    bool inPlace, hipStream_t s)
{
  switch (elementBytes) {
  case 2: return launchPatternT<uint16_t>(planes, src, dst, cells, inPlace, s);
  case 4: return launchPatternT<uint32_t>(planes, src, dst, cells, inPlace, s);
  case 8: return launchPatternT<uint64_t>(planes, src, dst, cells, inPlace, s);
  default: return false;
  }
}

} // namespace lbm

namespace lbm {

// ---- solid-mask windows ------------------------------------------------------
// one thread per (seg, y, x), x fastest so that the flag reads of a
// wavefront are contiguous
__global__ __launch_bounds__(256) void kBuildMask(const int *__restrict__ lat,
    MaskWin *__restrict__ mask, int nx, int ny, int nz, int segs)
{
  const uint64_t t     = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)nx * ny * segs;
  if (t >= total)
    return;
  const int x   = (int)(t % nx);
  const int y   = (int)((t / nx) % ny);
  const int seg = (int)(t / ((uint64_t)nx * ny));
  uint64_t lo = 0, hi = 0;
  for (int k = 0; k < 66; ++k) {
    int64_t zz = ((int64_t)seg * 64 - 1 + k) % nz;
    if (zz < 0)
      zz += nz;
    const uint64_t bit = lat[((uint64_t)zz * ny + y) * nx + x] == 0 ? 1ull : 0ull;
    if (k == 0)
      hi |= bit;
    else if (k <= 64)
      lo |= bit << (k - 1);
    else
      hi |= bit << 1;
  }
  mask[((uint64_t)x * ny + y) * segs + seg] = MaskWin{ lo, hi };
}

// bit 63 of hi: a cell of the 3x3 rows around the window (z halo included)
// is solid, so the wavefront of this window needs bounce-back; bit
// 2 + 3 (ox + 1) + (oy + 1): the same for the one row (x + ox, y + oy).
// Reads the windows from src and writes them with the flags to dst.
__global__ __launch_bounds__(256) void kMaskNeighbourhood(const MaskWin *__restrict__ src,
    MaskWin *__restrict__ dst, int nx, int ny, int segs)
{
  const uint64_t t     = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)nx * ny * segs;
  if (t >= total)
    return;
  const int x   = (int)(t % nx);
  const int y   = (int)((t / nx) % ny);
  const int seg = (int)(t / ((uint64_t)nx * ny));
  uint64_t rows = 0;
  for (int ox = -1; ox <= 1; ++ox)
    for (int oy = -1; oy <= 1; ++oy) {
      const int xn = wrapAdd(x, ox, nx), yn = wrapAdd(y, oy, ny);
      const MaskWin w = src[((uint64_t)xn * ny + yn) * segs + seg];
      if ((w.lo | (w.hi & 3ull)) != 0)
        rows |= 1ull << (2 + 3 * (ox + 1) + (oy + 1));
    }
  MaskWin w = src[((uint64_t)x * ny + y) * segs + seg];
  w.hi |= rows | ((rows != 0 ? 1ull : 0ull) << 63);
  dst[((uint64_t)x * ny + y) * segs + seg] = w;
}

void launchBuildMask(const int *lattice, MaskWin *mask, int nx, int ny, int nz, int segs,
    hipStream_t s)
{
  const uint64_t total  = (uint64_t)nx * ny * segs;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  MaskWin *tmp          = nullptr;
  if (hipMalloc(&tmp, total * sizeof(MaskWin)) != hipSuccess) {
    fprintf(stderr, "ERROR: hipMalloc of %zu bytes for the mask windows failed\n",
        (size_t)(total * sizeof(MaskWin)));
    exit(EXIT_FAILURE);
  }
  hipLaunchKernelGGL(kBuildMask, dim3(blocks), dim3(256), 0, s, lattice, tmp, nx, ny, nz, segs);
  hipLaunchKernelGGL(kMaskNeighbourhood, dim3(blocks), dim3(256), 0, s, tmp, mask, nx, ny, segs);
  if (hipStreamSynchronize(s) != hipSuccess || hipFree(tmp) != hipSuccess) {
    fprintf(stderr, "ERROR: building the mask windows failed\n");
    exit(EXIT_FAILURE);
  }
}

} // namespace lbm
