// lbm_runtime.hip -- host side of the C-ABI: the kernel registry entries of
// LBM-Bench's lbm.h (kernelInit<Variant>, kernel<Variant>, kernelDeinit,
// GetNode/SetNode, kernelAddBodyForce, kernelDensity) on top of device
// arrays, plus the entry points of lbm_hip.h.
//
// Device state per KernelData (struct LbmDevice):
//   one (AA) or two (push/pull) PDF arrays, hipMalloc'd, padded SoA planes
//   or AoS; the solid-mask windows; a non-blocking stream of default
//   priority; two hipEvents; hipGraphs of G consecutive steps; a device
//   step counter that seeds the FP16/BF16 stochastic rounding; a host
//   mirror of the array holding the state, allocated on the first
//   GetNode/SetNode.
//
// Step loop of kernel<Variant> (the time loop of src/kernel_*.c):
//   the first step of a run collides the state as it is; every later step
//   adds the body force of the step before it to the values it loads
//   (the reference adds it after bounce-back, src/kernel_pull_soa.c:206-209,
//   and the value it stores is the value the next step reads); one force
//   pass after the last step completes the state.  AA runs steps in
//   even/odd pairs, rounding an odd count up (src/kernel_aa_soa.c:36).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <vector>

#include "lbm.h"
#include "lbm_geometry.h"
#include "lbm_hip.h"
#include "lbm_kernels.h"

using namespace lbm;

#define HIP_OK(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "ERROR: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_),      \
          __FILE__, __LINE__);                                                             \
      exit(EXIT_FAILURE);                                                                  \
    }                                                                                      \
  } while (0)

static void fail(const char *msg)
{
  fprintf(stderr, "ERROR: %s\n", msg);
  exit(EXIT_FAILURE);
}

#ifdef PRECISION_SP
static const int kNativeStorage = ST_F32;
#else
static const int kNativeStorage = ST_F64;
#endif

static const double kWeight[Q] = { 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36,
  1.0 / 36, 1.0 / 36, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 18, 1.0 / 36,
  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 3 };

struct LbmDevice {
  int prop, layout, tiled, storage;
  size_t esize;
  int nx, ny, nz, segs;
  uint64_t ncells, plane, nelem;
  uint64_t nfluid;
  int narrays;
  void *d[2];
  void *post; // push-aos in two passes: post-collision planes, Q x ncells
  MaskWin *dmask;
  std::vector<MaskWin> hmask;
  int active;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  unsigned char *mirror;
  bool mirrorValid, mirrorDirty;
  hipGraphExec_t gexec[2];
  bool gvalid[2];
  int gsteps;
  double gkey[9];
  double rhoRef;
  float scale, invScale;
  double *dpart, *dout;
  int npart;
  int remap, nt, wavesPerBlock, dither, rowSweep, band, wgPerCu, wallNt, closeFused;
  int placementProbe, placementCandidates;
  double placementGbs;
  // FP16/BF16 rounding seeds.  Every step is seeded with its absolute
  // number, steps, however the steps are grouped into graph replays and
  // eager launches, so the rounding does not depend on LBM_GRAPH_STEPS.
  // A graph node adds its index to the device counter dseq[0], and the
  // graph's last node advances it; before a replay the host brings it to
  // `steps` (dseqHost: its value as the host last set it).  Eager steps
  // add their number to dseq[1], which stays 0.  Force passes number
  // themselves from seq, offset by 2^31 so that the two ranges stay apart.
  uint32_t *dseq;
  uint32_t seq, steps, dseqHost;
  // Zou-He pressure boundaries (kernelComputeBoundaryConditions): one byte
  // per (y, z) of the x faces, bit 0 inlet at x = 0, bit 1 outlet at
  // x = nX-1, uploaded from the lattice's flags (dface, allocated on first
  // use); zhOn: applied after every step of the current run (zh: its
  // densities and force terms)
  uint8_t *dface;
  int zhOn;
  ZouHeLaunch zh;
  // per-launch timing (lbmRunStepsTimed): an event after every step launch,
  // the first one before it; tkind: 1 for an AA odd step, else 0
  bool timing;
  std::vector<hipEvent_t> tev;
  std::vector<int> tkind;
};

// ---- device scratch of the no-allocation entry point ----------------------------
// lbmCompareMomentumNoAlloc keeps its partial sums here instead of in
// allocated memory: one slot of 4 doubles {sum |dj|^2, sum |jB|^2,
// max |dj|, sum drho^2} per chunk of kNoAllocChunk cells, then the 4
// totals.  4 * 32768 + 4 doubles = 1 MiB + 32 B, so lattices of up to
// 32768 * 8192 = 2^28 cells.  Each GPU the library runs on holds its own
// copy until the process ends.  In host code the name `scratch` is a
// host-side stand-in, not a device address: the address comes from
// hipGetSymbolAddress, once per device with that device current
// (ensureScratch, called from Init), and only that address is used.
constexpr uint64_t kNoAllocChunk  = 8192;
constexpr uint64_t kNoAllocGroups = 32768;
constexpr int kMaxDevices         = 64;
static __device__ __align__(256) double scratch[4 * kNoAllocGroups + 4];
static double *gScratch[kMaxDevices]; // device address of scratch, per device

// the current device's scratch address (looked up on first use on a device);
// null after a message when there is none
static double *ensureScratch()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "ERROR: no usable HIP device for the scratch buffer\n");
    return nullptr;
  }
  if (!gScratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      fprintf(stderr, "ERROR: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return nullptr;
    }
    gScratch[dev] = (double *)p;
  }
  return gScratch[dev];
}

// ---- small helpers ---------------------------------------------------------
static int envInt(const char *name, int dflt)
{
  const char *v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

static bool solidHost(const LbmDevice *D, int x, int y, int z)
{
  const MaskWin &w = D->hmask[((size_t)x * D->ny + y) * D->segs + (z >> 6)];
  return (w.lo >> (z & 63)) & 1ull;
}

static int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

static int currentRule(const LbmDevice *D, const KernelDataType *kd)
{
  if (D->prop == PROP_PULL)
    return SR_PULL;
  if (D->prop == PROP_AA && kd->Iteration >= 0 && kd->Iteration % 2 == 0)
    return SR_AA_EVEN;
  return SR_LOCAL;
}

// element index of logical f_d(x, y, z) (src/lbm.c:193-294, masked form)
static uint64_t hostSlot(const LbmDevice *D, int rule, int x, int y, int z, int d)
{
  auto sl = [&](uint64_t c, int k) -> uint64_t {
    return D->layout == LAYOUT_AOS ? c * Q + (uint64_t)k : (uint64_t)k * D->plane + c;
  };
  const uint64_t cell = ((uint64_t)x * D->ny + y) * D->nz + z;
  if (rule == SR_LOCAL || d == C)
    return sl(cell, d);
  const int xn = wrap(x - cx(d), D->nx), yn = wrap(y - cy(d), D->ny), zn = wrap(z - cz(d), D->nz);
  const bool bb    = solidHost(D, xn, yn, zn);
  const uint64_t n = ((uint64_t)xn * D->ny + yn) * D->nz + zn;
  if (rule == SR_PULL)
    return bb ? sl(cell, inv(d)) : sl(n, d);
  return bb ? sl(cell, d) : sl(n, inv(d));
}

// storage <-> double on the host
static float halfToFloat(uint16_t h)
{
  _Float16 v;
  memcpy(&v, &h, 2);
  return (float)v;
}
static uint16_t floatToHalf(float f)
{
  _Float16 v = (_Float16)f;
  uint16_t h;
  memcpy(&h, &v, 2);
  return h;
}
static float bf16ToFloat(uint16_t b)
{
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint16_t floatToBf16(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u)
    return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static double loadHost(const LbmDevice *D, uint64_t i, int d)
{
  const unsigned char *m = D->mirror;
  switch (D->storage) {
  case ST_F64:
    return ((const double *)m)[i];
  case ST_F32:
    return ((const float *)m)[i];
  case ST_F32DEV:
    return kWeight[d] * D->rhoRef + (double)(((const float *)m)[i] * D->invScale);
  case ST_F16:
    return kWeight[d] * D->rhoRef + (double)(halfToFloat(((const uint16_t *)m)[i]) * D->invScale);
  default:
    return kWeight[d] * D->rhoRef + (double)(bf16ToFloat(((const uint16_t *)m)[i]) * D->invScale);
  }
}

static void storeHost(LbmDevice *D, uint64_t i, int d, double v)
{
  unsigned char *m = D->mirror;
  switch (D->storage) {
  case ST_F64:
    ((double *)m)[i] = v;
    return;
  case ST_F32:
    ((float *)m)[i] = (float)v;
    return;
  default:
    break;
  }
  const float dev = (float)(v - kWeight[d] * D->rhoRef) * D->scale;
  if (D->storage == ST_F32DEV)
    ((float *)m)[i] = dev;
  else if (D->storage == ST_F16)
    ((uint16_t *)m)[i] = floatToHalf(dev);
  else
    ((uint16_t *)m)[i] = floatToBf16(dev);
}

// ---- host mirror -----------------------------------------------------------
static void pointMirror(KernelDataType *kd)
{
  LbmDevice *D = kd->Dev;
  PdfType *m   = (D->mirror && D->storage == kNativeStorage) ? (PdfType *)D->mirror : NULL;
  kd->Pdfs[0] = kd->Pdfs[1] = NULL;
  kd->Pdfs[D->active]       = m;
  kd->PdfsActive            = m;
}

static void ensureMirror(KernelDataType *kd)
{
  LbmDevice *D = kd->Dev;
  if (!D->mirror) {
    D->mirror = (unsigned char *)malloc(D->nelem * D->esize);
    if (!D->mirror)
      fail("insufficient host memory for the node mirror");
    D->mirrorValid = false;
  }
  if (!D->mirrorValid) {
    HIP_OK(hipStreamSynchronize(D->stream));
    HIP_OK(hipMemcpy(D->mirror, D->d[D->active], D->nelem * D->esize, hipMemcpyDeviceToHost));
    D->mirrorValid = true;
    D->mirrorDirty = false;
  }
  pointMirror(kd);
}

static void flushMirror(LbmDevice *D)
{
  if (D->mirror && D->mirrorDirty) {
    HIP_OK(hipStreamSynchronize(D->stream));
    HIP_OK(hipMemcpy(D->d[D->active], D->mirror, D->nelem * D->esize, hipMemcpyHostToDevice));
    D->mirrorDirty = false;
  }
}

static void getNodeImpl(KernelDataType *kd, int x, int y, int z, PdfType *pdfs)
{
  ensureMirror(kd);
  const int rule = currentRule(kd->Dev, kd);
  for (int d = 0; d < Q; ++d)
    pdfs[d] = (PdfType)loadHost(kd->Dev, hostSlot(kd->Dev, rule, x, y, z, d), d);
}

static void setNodeImpl(KernelDataType *kd, int x, int y, int z, PdfType *pdfs)
{
  ensureMirror(kd);
  const int rule = currentRule(kd->Dev, kd);
  for (int d = 0; d < Q; ++d)
    storeHost(kd->Dev, hostSlot(kd->Dev, rule, x, y, z, d), d, (double)pdfs[d]);
  kd->Dev->mirrorDirty = true;
}

static void bcGetImpl(KernelDataType *kd, int x, int y, int z, int dir, PdfType *pdf)
{
  ensureMirror(kd);
  const int rule = currentRule(kd->Dev, kd);
  *pdf           = (PdfType)loadHost(kd->Dev, hostSlot(kd->Dev, rule, x, y, z, dir), dir);
}

static void bcSetImpl(KernelDataType *kd, int x, int y, int z, int dir, PdfType pdf)
{
  ensureMirror(kd);
  const int rule = currentRule(kd->Dev, kd);
  storeHost(kd->Dev, hostSlot(kd->Dev, rule, x, y, z, dir), dir, (double)pdf);
  kd->Dev->mirrorDirty = true;
}

// ---- launches --------------------------------------------------------------
static Geom geomOf(const LbmDevice *D)
{
  Geom g;
  g.mask  = D->dmask;
  g.plane = D->plane;
  g.nx    = D->nx;
  g.ny    = D->ny;
  g.nz    = D->nz;
  g.segs  = D->segs;
  g.aos   = D->layout == LAYOUT_AOS;
  return g;
}

template <typename R>
static void trtConstants(R omega, R magic, R *c)
{
  // src/kernel_pull_soa.c:18-28, with magic as a parameter
  const R w1 = R(1.0) / R(18.0), w2 = R(1.0) / R(36.0);
  c[0] = omega;
  c[1] = R(1.0) / (R(0.5) + magic / (R(1.0) / omega - R(0.5)));
  c[2] = R(1.0) / R(3.0);
  c[3] = w1;
  c[4] = w2;
  c[5] = w1 * R(3.0);
  c[6] = w2 * R(3.0);
  c[7] = w1 * R(9.0) / R(2.0);
  c[8] = w2 * R(9.0) / R(2.0);
}

static StepLaunch baseLaunch(const LbmDevice *D, const CaseDataType *cd)
{
  StepLaunch L;
  memset(&L, 0, sizeof(L));
  L.g       = geomOf(D);
  L.tiled   = D->tiled;
  L.storage = D->storage;
  trtConstants<double>((double)cd->Omega, cd->MagicParam, L.trt);
  trtConstants<float>((float)cd->Omega, (float)cd->MagicParam, L.trtf);
  // 3 * w * c_x * g, evaluated as in src/lbm.c:919
  L.f1       = 3.0 * (1.0 / 18.0) * (double)cd->XForce;
  L.f2       = 3.0 * (1.0 / 36.0) * (double)cd->XForce;
  L.f1f      = 3.0f * (1.0f / 18.0f) * (float)cd->XForce;
  L.f2f      = 3.0f * (1.0f / 36.0f) * (float)cd->XForce;
  L.scale    = D->scale;
  L.invScale = D->invScale;
  L.remap         = D->remap;
  L.band          = D->band;
  L.wgPerCu       = D->wgPerCu;
  L.post          = D->post;
  L.nt            = D->nt;
  L.wavesPerBlock = D->wavesPerBlock;
  L.rowSweep      = D->rowSweep;
  L.dither        = D->dither;
  L.wallNt        = D->wallNt;
  return L;
}

static uint32_t forceSeq(LbmDevice *D) { return 0x80000000u + D->seq++; }

// a graph replay of G steps, with the device step counter brought to the
// number of steps run so far first
static void replaySteps(LbmDevice *D, hipGraphExec_t g, int G)
{
  if (D->dseqHost != D->steps) {
    launchBumpSeq(D->dseq, D->steps - D->dseqHost, D->stream);
    HIP_OK(hipGetLastError());
    D->dseqHost = D->steps;
  }
  HIP_OK(hipGraphLaunch(g, D->stream));
  D->steps += (uint32_t)G;
  D->dseqHost += (uint32_t)G;
}

// one step; for two-lattice variants src is D->d[srcIdx].  node >= 0 is the
// step's index inside a graph being captured, -1 an eager launch.
static void launchStep(LbmDevice *D, StepLaunch L, int variant, int srcIdx, int addForce,
    hipStream_t s, int node = -1, int close = 0)
{
  L.variant  = variant;
  L.addForce = addForce;
  L.close    = close;
  // XCD-aware block order: on for the AoS layout, whose scattered side
  // collects each record line from nine neighbour rows, so rows that share
  // lines should share an L2 (measured at 256^3: pull-aos +66 %, push-aos
  // +16 %).  Off for SoA: -7 % pull-soa, -2 % push-soa, -1-3 % row sweep,
  // -3 % FP32 AA even; blocked variants neutral.
  if (D->remap < 0)
    L.remap = D->layout == LAYOUT_AOS ? 1 : 0;
  // AoS row order and residency: the scattered side (pull's gather, push's
  // scatter) reads or writes each record line from nine neighbour rows, and
  // L2 serves those nine only when they run close together.  Rows are taken
  // in bands of y-rows swept along x (the nine users of a line then lie
  // within about two bands' rows), and at most three workgroups stay on a
  // CU, so less of the lattice is in flight per XCD than its 4 MB L2 holds.
  // Measured, channel, FP64: pull-aos 5,338 -> 9,762 MLUPS at 256^3 (16
  // rows of 256 per band) and 8,159 at 512^3 (8 rows); push-aos 4,176 ->
  // 5,634 at 256^3 (8 rows), 4,846 at 512^3 (4 rows), in one pass.  In
  // y-fastest order pull-aos read 8.6x and push-aos wrote 5.2x the
  // algorithmic bytes; in bands pull-aos reads 1.14x.
  if (D->band < 0)
    L.band = D->layout != LAYOUT_AOS ? 0
           : std::max(1, std::min(L.g.ny, (variant == V_PULL ? 4096 : 2048) / L.g.nz));
  if (D->wgPerCu < 0)
    L.wgPerCu = D->layout == LAYOUT_AOS ? 3 : 0;
  // measured at 512^3 (aa-soa): non-temporal loads and stores for 8-byte
  // elements and for the AA even step of 4- and 2-byte elements (-7 % and
  // -11 % time against cached loads); the other steps of 4- and 2-byte
  // elements read lines that neighbouring wavefronts share, and keep their
  // loads cached (non-temporal there costs 23 % and 27 %)
  if (D->nt < 0)
    L.nt = D->esize == 8 || variant == V_AA_EVEN ? 3 : 1;
  if (node >= 0) {
    L.seqBase = D->dseq;
    L.seq     = (uint32_t)node;
  } else {
    L.seqBase = D->dseq + 1;
    L.seq     = D->steps++;
  }
  if (D->prop == PROP_AA) {
    L.src = L.dst = D->d[0];
  } else {
    L.src = D->d[srcIdx];
    L.dst = D->d[srcIdx ^ 1];
  }
  launchStepRow(L, s); // L.tiled orders the blocked variants' wavefronts in tiles
  HIP_OK(hipGetLastError());
  if (D->zhOn) {
    // the pressure boundaries on the state this step wrote, read through
    // the slot rule that holds after it; the step's own force is pending
    AuxLaunch A;
    A.g        = L.g;
    A.storage  = D->storage;
    A.rule     = variant == V_PULL ? SR_PULL : (variant == V_AA_EVEN ? SR_AA_EVEN : SR_LOCAL);
    A.pdf      = L.dst;
    A.rhoRef   = D->rhoRef;
    A.scale    = D->scale;
    A.invScale = D->invScale;
    A.dither   = D->dither;
    A.seq      = 0;
    ZouHeLaunch Z = D->zh;
    Z.seqBase     = L.seqBase;
    Z.seq         = L.seq;
    launchZouHe(A, Z, s);
    HIP_OK(hipGetLastError());
  }
  if (D->timing) {
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, s));
    D->tev.push_back(e);
    D->tkind.push_back(variant == V_AA_ODD ? 1 : 0);
  }
}

static int stepVariant(const LbmDevice *D) { return D->prop == PROP_PULL ? V_PULL : V_PUSH; }

static void destroyGraphs(LbmDevice *D)
{
  for (int i = 0; i < 2; ++i)
    if (D->gvalid[i]) {
      HIP_OK(hipGraphExecDestroy(D->gexec[i]));
      D->gvalid[i] = false;
    }
}

// hipGraphs of G steps that all add the force of the step before them:
// AA -> one graph (even, odd, ...); push/pull -> one per source array.
static void prepareGraphs(LbmDevice *D, const CaseDataType *cd, const StepLaunch &base)
{
  const double key[9] = { (double)cd->Omega, cd->MagicParam, (double)cd->XForce,
    (double)cd->BodyForce, (double)D->gsteps, (double)D->storage, (double)D->zhOn,
    D->zhOn ? D->zh.rhoIn : 0.0, D->zhOn ? D->zh.rhoOut : 0.0 };
  bool same = D->gvalid[0];
  for (int i = 0; i < 9 && same; ++i)
    same = key[i] == D->gkey[i];
  if (same || D->gsteps < 2)
    return;
  destroyGraphs(D);
  memcpy(D->gkey, key, sizeof(key));
  const int ngraphs = D->prop == PROP_AA ? 1 : 2;
  for (int gi = 0; gi < ngraphs; ++gi) {
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(D->stream, hipStreamCaptureModeThreadLocal));
    int src = gi;
    for (int s = 0; s < D->gsteps; ++s) {
      if (D->prop == PROP_AA)
        launchStep(D, base, (s & 1) ? V_AA_ODD : V_AA_EVEN, 0, cd->BodyForce, D->stream, s);
      else {
        launchStep(D, base, stepVariant(D), src, cd->BodyForce, D->stream, s);
        src ^= 1;
      }
    }
    launchBumpSeq(D->dseq, (uint32_t)D->gsteps, D->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamEndCapture(D->stream, &graph));
    HIP_OK(hipGraphInstantiate(&D->gexec[gi], graph, nullptr, nullptr, 0));
    HIP_OK(hipGraphDestroy(graph));
    D->gvalid[gi] = true;
  }
}

static AuxLaunch auxOf(const LbmDevice *D, int rule)
{
  AuxLaunch A;
  A.g        = geomOf(D);
  A.storage  = D->storage;
  A.rule     = rule;
  A.pdf      = D->d[D->active];
  A.rhoRef   = D->rhoRef;
  A.scale    = D->scale;
  A.invScale = D->invScale;
  A.dither   = D->dither;
  A.seq      = 0;
  return A;
}

// Uploads the inlet / outlet flags of the x faces (y, z in 1 .. n-2, as
// src/lbm.c:693-694 visits them) from the lattice's flags; returns whether
// any cell is flagged.  Reads 2 nY nZ flags, never the whole lattice.
static bool uploadFaces(LbmDevice *D, const LatticeDescType *ld)
{
  const size_t n = (size_t)D->ny * D->nz;
  std::vector<uint8_t> face(n, 0);
  int dims[3]  = { D->nx, D->ny, D->nz };
  bool any     = false;
  for (int y = 1; y < D->ny - 1; ++y)
    for (int z = 1; z < D->nz - 1; ++z) {
      uint8_t b = 0;
      if (ld->Lattice[latticeIndex(dims, 0, y, z)] == LAT_CELL_INLET)
        b |= 1;
      if (ld->Lattice[latticeIndex(dims, D->nx - 1, y, z)] == LAT_CELL_OUTLET)
        b |= 2;
      face[(size_t)y * D->nz + z] = b;
      any |= b != 0;
    }
  if (!any)
    return false;
  if (!D->dface)
    HIP_OK(hipMalloc(&D->dface, n));
  HIP_OK(hipMemcpyAsync(D->dface, face.data(), n, hipMemcpyHostToDevice, D->stream));
  HIP_OK(hipStreamSynchronize(D->stream));
  return true;
}

static ZouHeLaunch zouHeOf(const LbmDevice *D, const CaseDataType *cd, const StepLaunch &base,
    int pending)
{
  ZouHeLaunch Z;
  Z.face    = D->dface;
  Z.rhoIn   = (double)cd->RhoIn;
  Z.rhoOut  = (double)cd->RhoOut;
  Z.f1      = D->storage == ST_F32 ? (double)base.f1f : base.f1;
  Z.f2      = D->storage == ST_F32 ? (double)base.f2f : base.f2;
  Z.pending = pending;
  Z.seqBase = D->dseq + 1;
  Z.seq     = 0;
  return Z;
}

// timed != null: every step is launched eagerly (no graph), with a hipEvent
// after each launch; timed[0] / timed[1] receive the summed durations of
// the step launches (AA: even steps) and of the AA odd steps, counts[0..1]
// their numbers
static void runKernel(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cdIn,
    double *timed = nullptr, int *counts = nullptr)
{
  const CaseDataType eff = caseDataEffective(cdIn);
  const CaseDataType *cd = &eff;
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  int n = cd->MaxIterations;
  if (n < 0)
    n = 0;
  if (D->prop == PROP_AA)
    n = (n + 1) / 2 * 2;
  const int bf          = cd->BodyForce ? 1 : 0;
  const StepLaunch base = baseLaunch(D, cd);
  // pressure boundaries after every step: the flags are read now, so a
  // relabelling between runs takes effect at the next run
  D->zhOn = cd->PressureBoundaries && uploadFaces(D, ld) ? 1 : 0;
  if (D->zhOn)
    D->zh = zouHeOf(D, cd, base, bf);
  prepareGraphs(D, cd, base);
  const int G = D->gvalid[0] && !timed ? D->gsteps : 0;

  // the reference's LIKWID region around the kernel (main.c:122-131), as a
  // ROCm marker range: rocprofv3 --marker-trace shows it
  roctxRangePush("PROPKERNEL");
  HIP_OK(hipEventRecord(D->ev0, D->stream));
  if (timed) {
    D->timing = true;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, D->stream));
    D->tev.assign(1, e);
    D->tkind.assign(1, -1);
  }
  // The last step adds its own force at the store where it can (push and
  // AA odd on SoA storage, no pressure boundaries after it): the closing
  // force pass then drops out.  That step is launched directly, after the
  // graph replays.
  const bool fuse = bf && !D->zhOn && D->layout == LAYOUT_SOA &&
                    (D->prop == PROP_AA || D->prop == PROP_PUSH) && D->closeFused;
  if (n > 0) {
    int left = n;
    if (D->prop == PROP_AA) {
      launchStep(D, base, V_AA_EVEN, 0, 0, D->stream);
      launchStep(D, base, V_AA_ODD, 0, bf, D->stream, -1, fuse && left == 2);
      left -= 2;
      while (G > 0 && left >= G + (fuse ? 2 : 0)) {
        replaySteps(D, D->gexec[0], G);
        left -= G;
      }
      while (left > 0) {
        launchStep(D, base, V_AA_EVEN, 0, bf, D->stream);
        launchStep(D, base, V_AA_ODD, 0, bf, D->stream, -1, fuse && left == 2);
        left -= 2;
      }
    } else {
      launchStep(D, base, stepVariant(D), D->active, 0, D->stream, -1, fuse && left == 1);
      D->active ^= 1;
      left -= 1;
      while (G > 0 && left >= G + (fuse ? 1 : 0)) {
        replaySteps(D, D->gexec[D->active], G);
        left -= G; // G is even: the state ends in the same array
      }
      while (left > 0) {
        launchStep(D, base, stepVariant(D), D->active, bf, D->stream, -1, fuse && left == 1);
        D->active ^= 1;
        left -= 1;
      }
    }
    if (bf && !fuse) {
      const int rule = D->prop == PROP_PULL ? SR_PULL : SR_LOCAL;
      AuxLaunch A    = auxOf(D, rule);
      A.seq          = forceSeq(D);
      launchForce(A, base.f1, base.f2, D->stream);
      HIP_OK(hipGetLastError());
    }
  }
  HIP_OK(hipEventRecord(D->ev1, D->stream));
  HIP_OK(hipEventSynchronize(D->ev1));
  roctxRangePop();
  D->zhOn = 0;
  if (timed) {
    timed[0] = timed[1] = 0.0;
    counts[0] = counts[1] = 0;
    for (size_t i = 1; i < D->tev.size(); ++i) {
      float t = 0.0f;
      HIP_OK(hipEventElapsedTime(&t, D->tev[i - 1], D->tev[i]));
      timed[D->tkind[i]] += (double)t * 1e-3;
      counts[D->tkind[i]] += 1;
    }
    for (hipEvent_t e : D->tev)
      HIP_OK(hipEventDestroy(e));
    D->tev.clear();
    D->tkind.clear();
    D->timing = false;
  }
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, D->ev0, D->ev1));
  kd->Duration = (double)ms * 1e-3;
  if (D->prop == PROP_AA && n > 0)
    kd->Iteration = n - 1;
  D->mirrorValid = false;
  pointMirror(kd);
}

// ---- init / deinit ------------------------------------------------------------
static int storageKind(const CaseDataType *cd)
{
  switch (cd->Storage) {
  case LBM_STORE_NATIVE:
    return kNativeStorage;
  case LBM_STORE_FP32:
    return ST_F32DEV;
  case LBM_STORE_FP16:
    return ST_F16;
  case LBM_STORE_BF16:
    return ST_BF16;
  default:
    fail("unknown storage form");
  }
  return kNativeStorage;
}

static size_t storageBytes(int st)
{
  return st == ST_F64 ? 8 : (st == ST_F32 || st == ST_F32DEV ? 4 : 2);
}

// The kernels wrap every axis, so the boundary layers of a non-periodic
// axis must be obstacles.  Only the faces are read (O(n^2) flags).
static void checkGeometry(const LatticeDescType *ld)
{
  const int n[3]   = { ld->Dims[0], ld->Dims[1], ld->Dims[2] };
  const int per[3] = { ld->PeriodicX, ld->PeriodicY, ld->PeriodicZ };
  int dims[3]      = { n[0], n[1], n[2] };
  for (int a = 0; a < 3; ++a) {
    if (per[a])
      continue;
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    for (int side = 0; side < 2; ++side)
      for (int j = 0; j < n[b]; ++j)
        for (int k = 0; k < n[c]; ++k) {
          int p[3];
          p[a] = side ? n[a] - 1 : 0;
          p[b] = j;
          p[c] = k;
          if (ld->Lattice[latticeIndex(dims, p[0], p[1], p[2])] != LAT_CELL_OBSTACLE)
            fail("fluid cell on a non-periodic boundary face; the boundary layers of "
                 "non-periodic axes must be obstacles");
        }
  }
}

// GB/s of the step's stream pattern (19 planes loaded and stored back in
// place, one element per lane; the pattern kernel of lbmPatternBandwidth)
// over a candidate PDF array, whose contents it overwrites
static double probePlacement(LbmDevice *D, void *pdf)
{
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  const int eb = (int)D->esize;
  launchPattern(Q, eb, pdf, pdf, D->plane, true, D->stream);
  HIP_OK(hipEventRecord(e0, D->stream));
  for (int r = 0; r < 2; ++r)
    launchPattern(Q, eb, pdf, pdf, D->plane, true, D->stream);
  HIP_OK(hipEventRecord(e1, D->stream));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  return 2.0 * 2.0 * Q * (double)D->plane * eb / ((double)ms * 1e-3) / 1e9;
}

// A PDF array.  How fast HBM serves the step's 19 streams depends on where
// an allocation lands: over fresh 512^3 FP64 lattices the even + odd pair
// took 15.3-15.4 ms for most and 14.0 ms for about one in four, and the
// same pattern over the array itself (probePlacement) tells them apart
// (5.5 against 5.9-6.0 TB/s).  For SoA arrays of at least 4 GiB, up to
// LBM_PLACEMENT_PROBE candidates (default 16) are allocated side by side
// while all of them together fit in half the device memory free before
// the first (seven of a 512^3 FP64 array on a 288 GB device, sixteen of
// an FP16 one), each probed (three launches of the pattern, about 20 ms
// at 512^3 FP64), and the fastest kept.  More candidates raise the odds
// of a fast placement; in 20-step bench runs paired on one box, 16
// against 6 gave FP64 18,985-19,343 against 18,359-19,287 MLUPS and BF16
// 67,978-68,629 against 64,646-68,364, but the low runs with 6 were each
// box's first bench, so the difference is not shown beyond the spread.  (Stopping at the first candidate 4 % above the slowest kept a
// 5.88 TB/s one whose steps ran between the two groups, so every
// candidate is probed.)  The others are freed before the lattice is
// initialised.
static void *allocPlaced(LbmDevice *D, size_t bytes)
{
  size_t free0 = 0, total0 = 0;
  HIP_OK(hipMemGetInfo(&free0, &total0));
  void *best = nullptr;
  HIP_OK(hipMalloc(&best, bytes));
  if (D->layout != LAYOUT_SOA || D->placementProbe < 2 || bytes < (4ull << 30) ||
      D->plane * (uint64_t)Q * D->esize != bytes)
    return best;
  double bestRate = probePlacement(D, best);
  std::vector<void *> others;
  for (int k = 1; k < D->placementProbe; ++k) {
    size_t freeB = 0, totalB = 0;
    HIP_OK(hipMemGetInfo(&freeB, &totalB));
    if ((uint64_t)(k + 1) * bytes > free0 / 2 || bytes > freeB / 2)
      break;
    void *c = nullptr;
    if (hipMalloc(&c, bytes) != hipSuccess) {
      (void)hipGetLastError(); // no room for another candidate: keep the best so far
      break;
    }
    const double r = probePlacement(D, c);
    if (r > bestRate) {
      others.push_back(best);
      best     = c;
      bestRate = r;
    } else {
      others.push_back(c);
    }
  }
  for (void *p : others)
    HIP_OK(hipFree(p));
  D->placementGbs = bestRate;
  D->placementCandidates = 1 + (int)others.size();
  return best;
}

static void initInternal(LatticeDescType *ld, KernelDataType **kdOut, CaseDataType *cdIn,
    int prop, int layout, int tiled, const char *name,
    void (*kernel)(LatticeDescType *, KernelDataType *, CaseDataType *))
{
  const CaseDataType eff = caseDataEffective(cdIn);
  const CaseDataType *cd = &eff;
  checkGeometry(ld);
  KernelDataType *kd = (KernelDataType *)calloc(1, sizeof(KernelDataType));
  LbmDevice *D       = new LbmDevice();
  if (!kd)
    fail("insufficient host memory");
  *kdOut   = kd;
  kd->Dev  = D;
  kd->Name = name;
  for (int a = 0; a < 3; ++a) {
    kd->Dims[a]       = ld->Dims[a];
    kd->GlobalDims[a] = ld->Dims[a];
    kd->Offsets[a]    = 0;
    kd->Blk[a]        = tiled ? (a == 2 ? 64 : 2) : ld->Dims[a]; // 2 x 2 rows, 64-cell segments
  }
  kd->Iteration = -1;
  kd->Kernel    = kernel;
  kd->GetNode   = getNodeImpl;
  kd->SetNode   = setNodeImpl;
  kd->BoundaryConditionsGetPdf = bcGetImpl;
  kd->BoundaryConditionsSetPdf = bcSetImpl;

  D->prop    = prop;
  D->layout  = layout;
  D->tiled   = tiled;
  D->storage = storageKind(cd);
  D->esize   = storageBytes(D->storage);
  D->nx      = ld->Dims[0];
  D->ny      = ld->Dims[1];
  D->nz      = ld->Dims[2];
  D->segs    = (D->nz + 63) / 64;
  D->ncells  = (uint64_t)D->nx * D->ny * D->nz;
  D->nfluid  = (uint64_t)ld->nFluid;
  // SoA planes padded to a multiple of 64 elements plus a stagger, so the 19
  // streams of a wavefront need not map to the same HBM channels (the planes
  // are 1 GiB apart at 512^3 FP64).  Measured over 0 ... 1 MiB with the
  // current kernels: none is best for 8-byte elements (+3 %), 64 KiB for
  // 2-byte elements (none loses 4 %), either for 4-byte elements.
  const size_t ebytes = storageBytes(storageKind(cd));
  const uint64_t pad  = (uint64_t)envInt("LBM_PLANE_PAD", ebytes == 8 ? 0 : (int)(65536 / ebytes));
  D->plane           = layout == LAYOUT_SOA ? ((D->ncells + 63) / 64) * 64 + pad : D->ncells;
  D->nelem           = layout == LAYOUT_SOA ? D->plane * Q : D->ncells * Q;
  D->narrays         = prop == PROP_AA ? 1 : 2;
  D->rhoRef          = (double)cd->RhoRef;
  double sc          = cd->StorageScale;
  if (!(sc > 0.0))
    sc = D->storage == ST_F16 ? 4096.0 : 1.0;
  if (D->storage == kNativeStorage)
    sc = 1.0;
  int ex;
  if (frexp(sc, &ex) != 0.5)
    fail("StorageScale must be a power of two");
  D->scale    = (float)sc;
  D->invScale = (float)(1.0 / sc);
  if (D->storage == kNativeStorage)
    D->rhoRef = 0.0;
  D->gsteps  = envInt("LBM_GRAPH_STEPS", 16) & ~1;
  // XCD-aware block order (LBM_XCD_REMAP 0 / 1; unset: per launch, see
  // launchStep)
  D->remap   = envInt("LBM_XCD_REMAP", -1);
  // row order of the AoS steps: bands of LBM_ROW_BAND y-rows, x swept
  // inside each band (0: x-major, y fastest; unset: per launch, see
  // launchStep)
  D->band    = envInt("LBM_ROW_BAND", -1);
  // resident workgroups per CU of the AoS steps (LBM_AOS_WG_PER_CU 1-4, 0:
  // no limit; unset: per launch, see launchStep)
  D->wgPerCu = envInt("LBM_AOS_WG_PER_CU", -1);
  // non-temporal hints (LBM_NT: bit 0 stores, bit 1 loads; unset: per
  // launch, see launchStep)
  D->nt      = envInt("LBM_NT", -1);
  D->wavesPerBlock = envInt("LBM_WAVES_PER_BLOCK", 4);
  D->dither        = envInt("LBM_DITHER", 1);
  // kRowGroup (AA odd, FP64): non-temporal bits of the blocks near a
  // solid (LBM_WALL_NT 0-7; unset: cached loads, non-temporal stores)
  D->wallNt        = envInt("LBM_WALL_NT", -1);
  // the last step of a run adds its own body force at the store
  // (LBM_CLOSE_FUSED 0: a separate force pass after it, as before)
  D->closeFused    = envInt("LBM_CLOSE_FUSED", 1);
  D->placementProbe      = envInt("LBM_PLACEMENT_PROBE", 16);
  D->placementCandidates = 1;
  D->placementGbs        = 0.0;
  // push / AA odd with nZ % 64 == 0: 1 the row sweep (one wavefront walks
  // a row), 3 the row group (one workgroup per row, nZ <= 512; else the
  // sweep; the AA odd step of 4- and 2-byte elements with nZ a multiple of
  // 64 x the cells of an 8-byte word runs the row group of 8-byte words),
  // 4 for 2-byte AA odd the row group of 4-byte words, 0 the per-block
  // kernel.  Measured at 512^3 spheres (aa-soa odd step, one process):
  // 8-byte elements row group 8.40 ms, sweep 8.80, per-block 9.30; 2-byte
  // sweep 2.32, row group 3.75, in 8-byte words 2.50, in 4-byte words 2.31;
  // 4-byte per-block 3.98, row group 4.51, sweep 4.69, in 8-byte words
  // 4.3.
  D->rowSweep      = envInt("LBM_ROW_SWEEP", D->esize == 8 ? 3 : (D->esize == 4 ? 0 : 1));
  // The AA odd step of 2-byte storage, LBM_ROW_SWEEP unset: the row group of
  // 4-byte words (kRowGroupV, two cells per lane) where its shape fits.
  // Measured at 512^3 (FP16 odd step, words against the row sweep, by the
  // share of 128-cell wavefronts near a solid): plain channel (0.8 %) 1.83
  // against 2.08 ms; 2000 spheres (46 %) 2.16 against 2.30; 6000 spheres
  // (83 %) 2.43-2.45 against 2.53-2.55; 15000 spheres (99 %) 2.82-2.85
  // against 2.91-2.92.
  if (envInt("LBM_ROW_SWEEP", -1) < 0 && D->esize == 2 && D->prop == PROP_AA &&
      D->layout == LAYOUT_SOA && D->nz % 128 == 0 && D->nz <= 1024 && D->plane % 2 == 0 &&
      D->plane < (1ull << 32)) {
    D->rowSweep = 4;
    printf("# AA odd step: row group of 4-byte words\n");
  }
  D->seq           = 0;
  D->steps         = 0;
  D->dseqHost      = 0;
  D->dface         = nullptr;
  D->zhOn          = 0;
  D->timing        = false;
  kd->LoopBalance = 2.0 * Q * (double)D->esize;

  HIP_OK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&D->ev0));
  HIP_OK(hipEventCreate(&D->ev1));

  const size_t bytes = D->nelem * D->esize;
  printf("# allocating data for %llu LB nodes on the GPU (%llu bytes = %.1f MiB for %s)\n",
      (unsigned long long)D->ncells, (unsigned long long)(bytes * D->narrays),
      (double)(bytes * D->narrays) / 1024.0 / 1024.0,
      D->narrays == 1 ? "single lattice" : "both lattices");
  // Refuse a lattice that cannot fit before anything is allocated or
  // queued: the PDF arrays, the mask windows (twice while they are built)
  // and the flag upload (as src/allocate.c:17-31 refuses on the host).
  {
    size_t freeB = 0, totalB = 0;
    HIP_OK(hipMemGetInfo(&freeB, &totalB));
    const size_t nwin   = (size_t)D->nx * D->ny * D->segs;
    const size_t needed = bytes * (size_t)D->narrays + 2 * nwin * sizeof(MaskWin) +
                          D->ncells * sizeof(int);
    if (needed > freeB) {
      fprintf(stderr,
          "ERROR: insufficient device memory: the lattice needs %.2f GiB, %.2f GiB of %.2f GiB "
          "are free\n",
          (double)needed / 1073741824.0, (double)freeB / 1073741824.0,
          (double)totalB / 1073741824.0);
      exit(EXIT_FAILURE);
    }
  }
  for (int i = 0; i < D->narrays; ++i) {
    D->d[i] = allocPlaced(D, bytes);
    HIP_OK(hipMemsetAsync(D->d[i], 0, bytes, D->stream));
  }
  // push-aos runs as two passes through SoA planes of post-collision
  // values (LBM_AOS_PUSH_PASSES 1: the single-pass scatter, no planes).
  // Without room for the planes it falls back to the single pass.
  D->post = nullptr;
  if (D->layout == LAYOUT_AOS && D->prop == PROP_PUSH && envInt("LBM_AOS_PUSH_PASSES", 2) == 2) {
    const size_t pbytes = (size_t)Q * D->ncells * D->esize;
    if (hipMalloc(&D->post, pbytes) == hipSuccess) {
      printf("# push-aos in two passes: %.1f MiB of post-collision planes\n",
          (double)pbytes / 1024.0 / 1024.0);
    } else {
      (void)hipGetLastError();
      D->post = nullptr;
      printf("# push-aos in one pass: no room for %.1f MiB of post-collision planes\n",
          (double)pbytes / 1024.0 / 1024.0);
    }
  }

  // mask windows, built on the device from the flag array
  const size_t nwin = (size_t)D->nx * D->ny * D->segs;
  HIP_OK(hipMalloc(&D->dmask, nwin * sizeof(MaskWin)));
  int *dlat = nullptr;
  HIP_OK(hipMalloc(&dlat, D->ncells * sizeof(int)));
  HIP_OK(hipMemcpyAsync(dlat, ld->Lattice, D->ncells * sizeof(int), hipMemcpyHostToDevice,
      D->stream));
  launchBuildMask(dlat, D->dmask, D->nx, D->ny, D->nz, D->segs, D->stream);
  HIP_OK(hipGetLastError());
  D->hmask.resize(nwin);
  HIP_OK(hipMemcpyAsync(D->hmask.data(), D->dmask, nwin * sizeof(MaskWin),
      hipMemcpyDeviceToHost, D->stream));
  HIP_OK(hipStreamSynchronize(D->stream));
  HIP_OK(hipFree(dlat));

  if (!ensureScratch())
    fail("the device scratch buffer has no address on this device");
  D->npart = momentsBlocks(geomOf(D));
  HIP_OK(hipMalloc(&D->dpart, sizeof(double) * 4 * (size_t)D->npart));
  HIP_OK(hipMalloc(&D->dout, sizeof(double) * 4));
  HIP_OK(hipMalloc(&D->dseq, 256));
  HIP_OK(hipMemset(D->dseq, 0, 256));
  pointMirror(kd);
}

extern "C" {

void kernelDeinit(LatticeDescType *ld, KernelDataType **kdp)
{
  if (kdp == NULL || *kdp == NULL)
    return;
  KernelDataType *kd = *kdp;
  LbmDevice *D       = kd->Dev;
  if (D) {
    HIP_OK(hipStreamSynchronize(D->stream));
    destroyGraphs(D);
    for (int i = 0; i < D->narrays; ++i)
      HIP_OK(hipFree(D->d[i]));
    if (D->post)
      HIP_OK(hipFree(D->post));
    HIP_OK(hipFree(D->dmask));
    HIP_OK(hipFree(D->dpart));
    HIP_OK(hipFree(D->dout));
    HIP_OK(hipFree(D->dseq));
    if (D->dface)
      HIP_OK(hipFree(D->dface));
    HIP_OK(hipEventDestroy(D->ev0));
    HIP_OK(hipEventDestroy(D->ev1));
    HIP_OK(hipStreamDestroy(D->stream));
    free(D->mirror);
    delete D;
  }
  free(kd);
  *kdp = NULL;
}

void kernelPushSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelPushAoS(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelPullSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelPullAoS(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelBlkPushSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelBlkPullSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }
void kernelAaSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd) { runKernel(ld, kd, cd); }

void kernelInitPushSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PUSH, LAYOUT_SOA, 0, "push-soa", kernelPushSoA);
}
void kernelInitPushAoS(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PUSH, LAYOUT_AOS, 0, "push-aos", kernelPushAoS);
}
void kernelInitPullSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PULL, LAYOUT_SOA, 0, "pull-soa", kernelPullSoA);
}
void kernelInitPullAoS(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PULL, LAYOUT_AOS, 0, "pull-aos", kernelPullAoS);
}
void kernelInitBlkPushSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PUSH, LAYOUT_SOA, 1, "blk-push-soa", kernelBlkPushSoA);
  printf("# blocking x: %3d y: %3d z: %3d\n", (*kd)->Blk[0], (*kd)->Blk[1], (*kd)->Blk[2]);
}
void kernelInitBlkPullSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_PULL, LAYOUT_SOA, 1, "blk-pull-soa", kernelBlkPullSoA);
  printf("# blocking x: %3d y: %3d z: %3d\n", (*kd)->Blk[0], (*kd)->Blk[1], (*kd)->Blk[2]);
}
void kernelInitAaSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)
{
  initInternal(ld, kd, cd, PROP_AA, LAYOUT_SOA, 0, "aa-soa", kernelAaSoA);
}

void kernelAddBodyForce(KernelDataType *kd, LatticeDescType *ld, CaseDataType *cdIn)
{
  const CaseDataType eff = caseDataEffective(cdIn);
  const CaseDataType *cd = &eff;
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const StepLaunch base = baseLaunch(D, cd);
  AuxLaunch A = auxOf(D, currentRule(D, kd));
  A.seq       = forceSeq(D);
  launchForce(A, base.f1, base.f2, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  D->mirrorValid = false;
}

void kernelComputeBoundaryConditions(KernelDataType *kd, LatticeDescType *ld, CaseDataType *cdIn)
{
  const CaseDataType eff = caseDataEffective(cdIn);
  const CaseDataType *cd = &eff;
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  if (!uploadFaces(D, ld))
    return; // no inlet or outlet cell: nothing to do, as in the reference's loop
  const StepLaunch base = baseLaunch(D, cd);
  // between runs the state is complete (the closing force pass has run)
  launchZouHe(auxOf(D, currentRule(D, kd)), zouHeOf(D, cd, base, 0), D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  D->mirrorValid = false;
}

void lbmPlaneVelocity(LatticeDescType *ld, KernelDataType *kd, int x, double *ux)
{
  LbmDevice *D = kd->Dev;
  if (x < 0 || x >= D->nx) {
    fprintf(stderr, "ERROR: lbmPlaneVelocity: plane x = %d outside 0 .. %d\n", x, D->nx - 1);
    exit(EXIT_FAILURE);
  }
  flushMirror(D);
  const size_t bytes = sizeof(double) * (size_t)D->ny * D->nz;
  double *buf        = nullptr;
  HIP_OK(hipMalloc(&buf, bytes));
  launchPlaneUx(auxOf(D, currentRule(D, kd)), x, buf, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(ux, buf, bytes, hipMemcpyDeviceToHost, D->stream));
  HIP_OK(hipStreamSynchronize(D->stream));
  HIP_OK(hipFree(buf));
}

long long lbmBuildFlags(const int dims[3], const int periodic[3], int type, int blockSize,
    int nSpheres, const double *spheres, const int *bounds, LatticeType *lattice)
{
  const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
  hipStream_t s;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  int *dlat                 = nullptr;
  double *dsph              = nullptr;
  int *dbox                 = nullptr;
  unsigned long long *dcnt  = nullptr;
  HIP_OK(hipMalloc(&dlat, n * sizeof(int)));
  HIP_OK(hipMalloc(&dcnt, sizeof(unsigned long long)));
  HIP_OK(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), s));
  if (nSpheres > 0) {
    HIP_OK(hipMalloc(&dsph, (size_t)nSpheres * 4 * sizeof(double)));
    HIP_OK(hipMalloc(&dbox, (size_t)nSpheres * 6 * sizeof(int)));
    HIP_OK(hipMemcpyAsync(dsph, spheres, (size_t)nSpheres * 4 * sizeof(double),
        hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(dbox, bounds, (size_t)nSpheres * 6 * sizeof(int), hipMemcpyHostToDevice, s));
  }
  launchGeometry(dlat, dims, periodic, type == LBM_GEO_PIPE, type == LBM_GEO_BLOCKS ? blockSize : 0,
      type == LBM_GEO_SPHERES ? nSpheres : 0, dsph, dbox, dcnt, s);
  HIP_OK(hipGetLastError());
  unsigned long long nfluid = 0;
  HIP_OK(hipMemcpyAsync(&nfluid, dcnt, sizeof(nfluid), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(lattice, dlat, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(hipFree(dlat));
  HIP_OK(hipFree(dcnt));
  if (dsph)
    HIP_OK(hipFree(dsph));
  if (dbox)
    HIP_OK(hipFree(dbox));
  HIP_OK(hipStreamDestroy(s));
  return (long long)nfluid;
}

void lbmMassMomentum(LatticeDescType *ld, KernelDataType *kd, double out[4])
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const int nb = launchMoments(auxOf(D, currentRule(D, kd)), D->dpart, nullptr, D->stream);
  HIP_OK(hipGetLastError());
  launchReduce4(D->dpart, nb, D->dout, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, D->dout, 4 * sizeof(double), hipMemcpyDeviceToHost, D->stream));
  HIP_OK(hipStreamSynchronize(D->stream));
  // the deviation forms add back sum_i w_i RhoRef = RhoRef per fluid cell
  out[0] += (double)D->nfluid * D->rhoRef;
}

PdfType kernelDensity(KernelDataType *kd, LatticeDescType *ld)
{
  double m[4];
  lbmMassMomentum(ld, kd, m);
  return (PdfType)(m[0] / (double)ld->nFluid);
}

void lbmCompareMomentum(LatticeDescType *ld, KernelDataType *a, KernelDataType *b, double out[4])
{
  LbmDevice *A = a->Dev, *B = b->Dev;
  if (A->ncells != B->ncells)
    fail("lbmCompareMomentum: lattices differ in size");
  flushMirror(A);
  flushMirror(B);
  HIP_OK(hipStreamSynchronize(B->stream));
  double *fa = nullptr, *fb = nullptr;
  HIP_OK(hipMalloc(&fa, sizeof(double) * 4 * A->ncells));
  HIP_OK(hipMalloc(&fb, sizeof(double) * 4 * A->ncells));
  launchMoments(auxOf(A, currentRule(A, a)), A->dpart, fa, A->stream);
  HIP_OK(hipGetLastError());
  launchMoments(auxOf(B, currentRule(B, b)), A->dpart, fb, A->stream);
  HIP_OK(hipGetLastError());
  // A->dpart holds momentsBlocks() partials, the same block count as compare
  const int nb = launchCompareFields(fa, fb, A->ncells, A->dpart, A->stream);
  HIP_OK(hipGetLastError());
  launchReduceCompare(A->dpart, nb, A->dout, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, A->dout, 4 * sizeof(double), hipMemcpyDeviceToHost, A->stream));
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipFree(fa));
  HIP_OK(hipFree(fb));
}

int lbmCompareMomentumNoAlloc(LatticeDescType *ld, KernelDataType *a, KernelDataType *b,
    double out[4])
{
  LbmDevice *A = a->Dev, *B = b->Dev;
  if (A->nx != B->nx || A->ny != B->ny || A->nz != B->nz) {
    fprintf(stderr, "ERROR: lbmCompareMomentumNoAlloc: lattices of %d x %d x %d and %d x %d x %d "
                    "cells\n", A->nx, A->ny, A->nz, B->nx, B->ny, B->nz);
    return -1;
  }
  if (A->ncells > kNoAllocChunk * kNoAllocGroups) {
    fprintf(stderr, "ERROR: lbmCompareMomentumNoAlloc: %llu cells exceed the %llu its scratch "
                    "buffer covers\n", (unsigned long long)A->ncells,
        (unsigned long long)(kNoAllocChunk * kNoAllocGroups));
    return -1;
  }
  double *sp = ensureScratch();
  if (!sp)
    return -1;
  flushMirror(A);
  flushMirror(B);
  HIP_OK(hipStreamSynchronize(B->stream));
  const uint64_t groups = (A->ncells + kNoAllocChunk - 1) / kNoAllocChunk;
  launchCompareLattices(auxOf(A, currentRule(A, a)), auxOf(B, currentRule(B, b)), kNoAllocChunk,
      sp, A->stream);
  HIP_OK(hipGetLastError());
  launchReduceCompare(sp, (int)groups, sp + 4 * kNoAllocGroups, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, sp + 4 * kNoAllocGroups, 4 * sizeof(double), hipMemcpyDeviceToHost,
      A->stream));
  HIP_OK(hipStreamSynchronize(A->stream));
  return 0;
}

void lbmGetNodes(LatticeDescType *ld, KernelDataType *kd, double *out)
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const size_t bytes = sizeof(double) * Q * D->ncells;
  double *buf        = nullptr;
  HIP_OK(hipMalloc(&buf, bytes));
  launchGather(auxOf(D, currentRule(D, kd)), buf, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  HIP_OK(hipMemcpy(out, buf, bytes, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(buf));
}

void lbmCellMoments(LatticeDescType *ld, KernelDataType *kd, double *out)
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const size_t bytes = sizeof(double) * 4 * D->ncells;
  double *buf        = nullptr;
  HIP_OK(hipMalloc(&buf, bytes));
  launchCellMoments(auxOf(D, currentRule(D, kd)), buf, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  HIP_OK(hipMemcpy(out, buf, bytes, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(buf));
}

} // extern "C"

// ---- state files (format in lbm_hip.h) --------------------------------------
namespace {

const char kStateMagic[8]      = { 'L', 'B', 'M', 'S', 'T', 'A', 'T', '1' };
const uint32_t kStateVersion   = 1;
const uint32_t kStateHeader    = 128;
const uint64_t kFnvBasis       = 0xcbf29ce484222325ull;
const uint64_t kFnvPrime       = 0x100000001b3ull;
const size_t kSlabTargetBytes  = 256u << 20; // PDF bytes moved per x-slab

struct StateHeader {
  char magic[8];
  uint32_t version, headerBytes;
  int32_t dims[3], periodic[3], q, iteration;
  double omega, xforce, magic_;
  uint64_t flagBytes, pdfBytes, checksum;
  char kernel[32];
};
static_assert(sizeof(StateHeader) == 128, "state header layout");

uint64_t fnvWords(uint64_t h, const void *data, size_t bytes)
{
  const uint64_t *w = (const uint64_t *)data;
  for (size_t i = 0; i < bytes / 8; ++i)
    h = (h ^ w[i]) * kFnvPrime;
  return h;
}

int stateFail(FILE *f, const char *path, const char *what)
{
  fprintf(stderr, "ERROR: state file %s: %s\n", path, what);
  if (f)
    fclose(f);
  return -1;
}

// x-slab length so that one slab holds about kSlabTargetBytes of PDFs
int slabPlanes(const LbmDevice *D)
{
  const size_t plane = (size_t)D->ny * D->nz * Q * sizeof(double);
  const size_t k     = plane ? kSlabTargetBytes / plane : 1;
  return (int)std::max<size_t>(1, std::min<size_t>(k, (size_t)D->nx));
}

std::vector<uint8_t> flagSection(const LatticeDescType *ld, uint64_t bytes)
{
  std::vector<uint8_t> v(bytes, 0);
  const size_t n = (size_t)ld->Dims[0] * ld->Dims[1] * ld->Dims[2];
  for (size_t i = 0; i < n; ++i)
    v[i] = (uint8_t)(int8_t)ld->Lattice[i];
  return v;
}

} // namespace

extern "C" {

int lbmSaveState(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, const char *path)
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const size_t ncells = (size_t)D->nx * D->ny * D->nz;
  StateHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kStateMagic, 8);
  h.version     = kStateVersion;
  h.headerBytes = kStateHeader;
  h.dims[0] = D->nx, h.dims[1] = D->ny, h.dims[2] = D->nz;
  h.periodic[0] = ld->PeriodicX, h.periodic[1] = ld->PeriodicY, h.periodic[2] = ld->PeriodicZ;
  h.q         = Q;
  h.iteration = kd->Iteration;
  if (cd) {
    const CaseDataType eff = caseDataEffective(cd);
    h.omega = (double)eff.Omega, h.xforce = (double)eff.XForce, h.magic_ = eff.MagicParam;
  }
  h.flagBytes = (ncells + 7) / 8 * 8;
  h.pdfBytes  = (uint64_t)ncells * Q * sizeof(double);
  if (kd->Name)
    strncpy(h.kernel, kd->Name, sizeof(h.kernel) - 1);

  FILE *f = fopen(path, "wb");
  if (!f)
    return stateFail(nullptr, path, "cannot open for writing");
  if (fwrite(&h, sizeof(h), 1, f) != 1)
    return stateFail(f, path, "write failed");
  std::vector<uint8_t> flags = flagSection(ld, h.flagBytes);
  uint64_t sum = fnvWords(kFnvBasis, flags.data(), flags.size());
  if (fwrite(flags.data(), 1, flags.size(), f) != flags.size())
    return stateFail(f, path, "write failed");

  const int planes       = slabPlanes(D);
  const size_t slabCells = (size_t)planes * D->ny * D->nz;
  double *dbuf           = nullptr;
  HIP_OK(hipMalloc(&dbuf, slabCells * Q * sizeof(double)));
  std::vector<double> hbuf(slabCells * Q);
  const AuxLaunch A = auxOf(D, currentRule(D, kd));
  int rc            = 0;
  for (size_t first = 0; first < ncells && rc == 0; first += slabCells) {
    const size_t count = std::min(slabCells, ncells - first);
    launchGather(A, dbuf, D->stream, first, count);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(hbuf.data(), dbuf, count * Q * sizeof(double), hipMemcpyDeviceToHost,
        D->stream));
    HIP_OK(hipStreamSynchronize(D->stream));
    sum = fnvWords(sum, hbuf.data(), count * Q * sizeof(double));
    if (fwrite(hbuf.data(), sizeof(double), count * Q, f) != count * Q)
      rc = -1;
  }
  HIP_OK(hipFree(dbuf));
  if (rc)
    return stateFail(f, path, "write failed");
  h.checksum = sum;
  if (fseek(f, 0, SEEK_SET) != 0 || fwrite(&h, sizeof(h), 1, f) != 1)
    return stateFail(f, path, "write failed");
  if (fclose(f) != 0)
    return stateFail(nullptr, path, "close failed");
  return 0;
}

int lbmLoadState(LatticeDescType *ld, KernelDataType *kd, const char *path)
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const size_t ncells = (size_t)D->nx * D->ny * D->nz;
  FILE *f             = fopen(path, "rb");
  if (!f)
    return stateFail(nullptr, path, "cannot open for reading");
  StateHeader h;
  if (fread(&h, sizeof(h), 1, f) != 1)
    return stateFail(f, path, "truncated header");
  if (memcmp(h.magic, kStateMagic, 8) != 0)
    return stateFail(f, path, "not an LBM state file");
  if (h.version != kStateVersion || h.headerBytes != kStateHeader || h.q != Q)
    return stateFail(f, path, "unsupported format version");
  if (h.dims[0] != D->nx || h.dims[1] != D->ny || h.dims[2] != D->nz)
    return stateFail(f, path, "lattice dimensions differ");
  if (h.flagBytes != (ncells + 7) / 8 * 8 || h.pdfBytes != (uint64_t)ncells * Q * sizeof(double))
    return stateFail(f, path, "section sizes do not match the dimensions");
  std::vector<uint8_t> flags(h.flagBytes);
  if (fread(flags.data(), 1, flags.size(), f) != flags.size())
    return stateFail(f, path, "truncated flag section");
  for (size_t i = 0; i < ncells; ++i)
    if ((flags[i] == LAT_CELL_OBSTACLE) != (ld->Lattice[i] == LAT_CELL_OBSTACLE))
      return stateFail(f, path, "obstacle cells differ from the lattice's");
  uint64_t sum = fnvWords(kFnvBasis, flags.data(), flags.size());

  // Pass 1 reads the PDF section and checks its length and checksum
  // without touching the lattice; only a file that passes is loaded, in
  // pass 2, so a truncated or corrupt file leaves the state as it was.
  const long pdfStart    = ftell(f);
  const int planes       = slabPlanes(D);
  const size_t slabCells = (size_t)planes * D->ny * D->nz;
  std::vector<double> hbuf(slabCells * Q);
  for (size_t first = 0; first < ncells; first += slabCells) {
    const size_t count = std::min(slabCells, ncells - first);
    if (fread(hbuf.data(), sizeof(double), count * Q, f) != count * Q)
      return stateFail(f, path, "truncated PDF section (the lattice is unchanged)");
    sum = fnvWords(sum, hbuf.data(), count * Q * sizeof(double));
  }
  if (fgetc(f) != EOF)
    return stateFail(f, path, "trailing bytes after the PDF section (the lattice is unchanged)");
  if (sum != h.checksum)
    return stateFail(f, path, "checksum mismatch (the lattice is unchanged)");
  if (pdfStart < 0 || fseek(f, pdfStart, SEEK_SET) != 0)
    return stateFail(f, path, "cannot seek back to the PDF section (the lattice is unchanged)");

  double *dbuf = nullptr;
  HIP_OK(hipMalloc(&dbuf, slabCells * Q * sizeof(double)));
  const AuxLaunch A = auxOf(D, currentRule(D, kd));
  uint64_t again    = fnvWords(kFnvBasis, flags.data(), flags.size());
  int rc            = 0;
  for (size_t first = 0; first < ncells; first += slabCells) {
    const size_t count = std::min(slabCells, ncells - first);
    if (fread(hbuf.data(), sizeof(double), count * Q, f) != count * Q) {
      rc = -1;
      break;
    }
    again = fnvWords(again, hbuf.data(), count * Q * sizeof(double));
    HIP_OK(hipMemcpyAsync(dbuf, hbuf.data(), count * Q * sizeof(double), hipMemcpyHostToDevice,
        D->stream));
    launchScatter(A, dbuf, D->stream, first, count);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(D->stream));
  }
  HIP_OK(hipFree(dbuf));
  D->mirrorValid = false;
  fclose(f);
  // the file changed between the two passes: the loaded PDFs are not the
  // checked ones
  if (rc || again != h.checksum) {
    fprintf(stderr, "ERROR: state file %s: changed while it was loaded (the lattice's PDFs are "
                    "undefined)\n", path);
    return -1;
  }
  return 0;
}

void lbmSetNodes(LatticeDescType *ld, KernelDataType *kd, const double *in)
{
  LbmDevice *D = kd->Dev;
  flushMirror(D);
  const size_t bytes = sizeof(double) * Q * D->ncells;
  double *buf        = nullptr;
  HIP_OK(hipMalloc(&buf, bytes));
  HIP_OK(hipMemcpy(buf, in, bytes, hipMemcpyHostToDevice));
  launchScatter(auxOf(D, currentRule(D, kd)), buf, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  HIP_OK(hipFree(buf));
  D->mirrorValid = false;
}

void lbmInitPdfs(LatticeDescType *ld, KernelDataType *kd, int mode, uint64_t seed, double rho,
    double amp)
{
  LbmDevice *D = kd->Dev;
  if (mode < LBM_INIT_ZERO || mode > LBM_INIT_RANDOM)
    fail("lbmInitPdfs: unknown mode");
  launchInit(auxOf(D, currentRule(D, kd)), mode, seed, rho, amp, D->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(D->stream));
  D->mirrorValid = false;
  D->mirrorDirty = false;
}

double lbmRunStepsTimed(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, int steps,
    double kernelSeconds[2], int launches[2])
{
  const int keep    = cd->MaxIterations;
  cd->MaxIterations = steps;
  runKernel(ld, kd, cd, kernelSeconds, launches);
  cd->MaxIterations = keep;
  return kd->Duration;
}

double lbmRunSteps(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, int steps)
{
  const int keep    = cd->MaxIterations;
  cd->MaxIterations = steps;
  kd->Kernel(ld, kd, cd);
  cd->MaxIterations = keep;
  return kd->Duration;
}

void lbmPlacement(KernelDataType *kd, int *candidates, double *gbs)
{
  const LbmDevice *D = kd->Dev;
  *candidates        = D->placementCandidates;
  *gbs               = D->placementGbs;
}

size_t lbmDeviceBytes(KernelDataType *kd)
{
  const LbmDevice *D = kd->Dev;
  const size_t post  = D->post ? (size_t)Q * D->ncells * D->esize : 0;
  return (size_t)D->narrays * D->nelem * D->esize + post;
}

double lbmCopyBandwidth(size_t bytes, int reps)
{
  double *a = nullptr, *b = nullptr;
  const uint64_t n = bytes / sizeof(double);
  hipStream_t s;
  hipEvent_t e0, e1;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipMalloc(&a, n * sizeof(double)));
  HIP_OK(hipMalloc(&b, n * sizeof(double)));
  HIP_OK(hipMemsetAsync(a, 0, n * sizeof(double), s));
  launchCopy(a, b, n, s);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    launchCopy(a, b, n, s);
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  HIP_OK(hipFree(a));
  HIP_OK(hipFree(b));
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  HIP_OK(hipStreamDestroy(s));
  return 2.0 * (double)(n / 2 * 2) * sizeof(double) * reps / ((double)ms * 1e-3) / 1e9;
}

void lbmPorousField(const int dims[3], uint64_t seed, double sigma, double a, double b,
    double *s)
{
  // noise and smoothing on the device; the mean and the variance are summed
  // on the host in cell order (so the field does not depend on a device
  // reduction's order); the logistic on the device
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) {
    fprintf(stderr, "ERROR: lbmPorousField: dimensions %d x %d x %d (each must be >= 1)\n",
        dims[0], dims[1], dims[2]);
    return;
  }
  const uint64_t n   = (uint64_t)dims[0] * dims[1] * dims[2];
  const size_t bytes = n * sizeof(double);
  double *u = nullptr, *t = nullptr, *kd = nullptr;
  hipStream_t st;
  HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIP_OK(hipMalloc(&u, bytes));
  HIP_OK(hipMalloc(&t, bytes));
  launchPorousNoise(u, n, seed, st);
  HIP_OK(hipGetLastError());
  double *field = u;
  if (sigma > 0.0) {
    const int rad = (int)ceil(3.0 * sigma);
    std::vector<double> kern((size_t)(2 * rad + 1));
    double norm = 0.0;
    for (int k = -rad; k <= rad; ++k) {
      kern[(size_t)(k + rad)] = exp(-0.5 * (double)k * k / (sigma * sigma));
      norm += kern[(size_t)(k + rad)];
    }
    for (double &w : kern)
      w /= norm;
    HIP_OK(hipMalloc(&kd, kern.size() * sizeof(double)));
    HIP_OK(hipMemcpyAsync(kd, kern.data(), kern.size() * sizeof(double), hipMemcpyHostToDevice, st));
    launchSmoothAxis(0, u, t, dims, kd, rad, st);
    launchSmoothAxis(1, t, u, dims, kd, rad, st);
    launchSmoothAxis(2, u, t, dims, kd, rad, st);
    HIP_OK(hipGetLastError());
    field = t;
  }
  HIP_OK(hipMemcpyAsync(s, field, bytes, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  double mean = 0.0, var = 0.0;
  for (uint64_t i = 0; i < n; ++i)
    mean += s[i];
  mean /= (double)n;
  for (uint64_t i = 0; i < n; ++i)
    var += (s[i] - mean) * (s[i] - mean);
  const double sd = sqrt(var / (double)n);
  launchLogistic(field, n, mean, sd, a, b, st);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(s, field, bytes, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (kd)
    HIP_OK(hipFree(kd));
  HIP_OK(hipFree(u));
  HIP_OK(hipFree(t));
  HIP_OK(hipStreamDestroy(st));
}

double lbmCopyBandwidthShape(size_t bytes, int reps, int vectors, int nonTemporal,
    int persistentGroups)
{
  const bool ok = (vectors == 1 || vectors == 2 || vectors == 4 || vectors == 8) && reps >= 1 &&
                  persistentGroups >= 0 && bytes >= 16;
  if (!ok) {
    fprintf(stderr, "lbmCopyBandwidthShape: vectors must be 1, 2, 4 or 8, reps positive, "
                    "persistentGroups >= 0, bytes >= 16 (got %d, %d, %d, %zu)\n",
        vectors, reps, persistentGroups, bytes);
    return -1.0;
  }
  double *a = nullptr, *b = nullptr;
  const uint64_t n = bytes / sizeof(double);
  hipStream_t s;
  hipEvent_t e0, e1;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipMalloc(&a, n * sizeof(double)));
  HIP_OK(hipMalloc(&b, n * sizeof(double)));
  HIP_OK(hipMemsetAsync(a, 0, n * sizeof(double), s));
  launchCopyShape(a, b, n, vectors, nonTemporal, persistentGroups, s);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    launchCopyShape(a, b, n, vectors, nonTemporal, persistentGroups, s);
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  HIP_OK(hipFree(a));
  HIP_OK(hipFree(b));
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  HIP_OK(hipStreamDestroy(s));
  return 2.0 * (double)(n / 2 * 2) * sizeof(double) * reps / ((double)ms * 1e-3) / 1e9;
}

double lbmPatternBandwidth(int planes, int elementBytes, int inPlace, size_t cells, int reps)
{
  const bool okPlanes = planes == 1 || planes == 2 || planes == 4 || planes == 8 || planes == 19;
  const bool okBytes  = elementBytes == 2 || elementBytes == 4 || elementBytes == 8;
  if (!okPlanes || !okBytes || cells == 0 || reps < 1) {
    fprintf(stderr, "lbmPatternBandwidth: planes must be 1, 2, 4, 8 or 19, elementBytes 2, 4 "
                    "or 8, cells and reps positive (got %d, %d, %zu, %d)\n",
        planes, elementBytes, cells, reps);
    return -1.0;
  }
  const size_t bytes = (size_t)planes * cells * (size_t)elementBytes;
  void *a = nullptr, *b = nullptr;
  hipStream_t s;
  hipEvent_t e0, e1;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipMalloc(&a, bytes));
  HIP_OK(hipMemsetAsync(a, 0, bytes, s));
  if (inPlace) {
    b = a;
  } else {
    HIP_OK(hipMalloc(&b, bytes));
    HIP_OK(hipMemsetAsync(b, 0, bytes, s));
  }
  launchPattern(planes, elementBytes, a, b, cells, inPlace != 0, s);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    launchPattern(planes, elementBytes, a, b, cells, inPlace != 0, s);
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  HIP_OK(hipFree(a));
  if (!inPlace)
    HIP_OK(hipFree(b));
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  HIP_OK(hipStreamDestroy(s));
  return 2.0 * (double)bytes * reps / ((double)ms * 1e-3) / 1e9;
}

int lbmDeviceSynchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 1; }

int lbmSelectDevice(int index)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
    n = 0;
  if (index < 0 || index >= n) {
    fprintf(stderr, "ERROR: no HIP device %d (%d present)\n", index, n);
    return -1;
  }
  return hipSetDevice(index) == hipSuccess ? 0 : -1;
}

int lbmDeviceInfo(char *name, int len, int *computeUnits, size_t *globalMem)
{
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return 1;
  if (name && len > 0) {
    snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (computeUnits)
    *computeUnits = prop.multiProcessorCount;
  if (globalMem)
    *globalMem = prop.totalGlobalMem;
  return 0;
}

} // extern "C"
