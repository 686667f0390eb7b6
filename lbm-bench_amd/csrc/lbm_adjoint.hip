// lbm_adjoint.hip -- the differentiation form: a Brinkman-penalised TRT
// forward step and its hand-written discrete adjoint, with multi-level
// checkpointing, behind the C-ABI of lbm_hip.h.  No counterpart exists in
// LBM-Bench; the forward step at s = 0 is the reference's push step.
//
// Model, per fluid cell y with solid fraction s(y) in [0, 1]:
//   rho = sum f, j = sum c f, j_s = (1 - s) j                (Brinkman:
//   the equilibrium velocity is scaled down, the drag is s j per step)
//   g = TRT(f) with j_s in place of j in D = rho - 1.5 |j_s|^2 and in
//       u_i = c_i . j_s  (src/kernel_pull_soa.c:118-198 otherwise)
//   f'_d(x) = (solid(x - c_d) ? g_{inv d}(x) : g_d(x - c_d)) + 3 w_d c_dx g
// Objective: J = sum over fluid cells of the plane x = X of 1/2 rho |j|^2
// (j is the velocity of the incompressible model, as in the reference).
//
// Adjoint of one step, lambda = dJ/df at t+1 -> at t, and dJ/ds:
//   mu_k(y)   = solid(y + c_k) ? lambda_{inv k}(y) : lambda_k(y + c_k)
//   M0 = w+ sum_k w_k mu_k,  V = sum_k w_k mu_k (9 w+ u_k + 3 w-) c_k
//   lambda_d  = a mu_d + b mu_{inv d} + M0 + (1 - s) c_d . (V - 3 M0 j_s)
//               a = 1 - (w+ + w-)/2,  b = (w- - w+)/2
//   dJ/ds(y) += 3 M0 (j_s . j) - j . V
// (w+ = omega, w- the odd rate; the transposed Jacobian of the TRT update
// in closed form: its even part depends on rho and j_s, its odd part on j_s.)
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "lbm.h"
#include "lbm_hip.h"
#include "lbm_impl.h"

namespace lbm {

template <typename R>
struct AdjParams {
  Trt<R> trt;
  R f1, f2;
  R scale, inv;  // forward storage scale
  double rhoRef; // forward deviation offset (0 for FP64 storage)
  uint64_t plane;
  int nx, ny, nz, segs, outletX, remap, dither;
  uint32_t seq; // forward step index, for the rounding seeds
};

// TRT collision with the Brinkman-scaled momentum (keep = 1 - s); at
// keep == 1 every product keep*u is exact and this is collide() bit for bit
template <typename R>
__device__ __forceinline__ void collideBrinkman(R (&f)[Q], const Trt<R> &p, R keep)
{
  const R ux = f[E] + f[NE] + f[SE] + f[TE] + f[BE] - f[W] - f[NW] - f[SW] - f[TW] - f[BW];
  const R uy = f[N] + f[NE] + f[NW] + f[TN] + f[BN] - f[S] - f[SE] - f[SW] - f[TS] - f[BS];
  const R uz = f[T] + f[TE] + f[TW] + f[TN] + f[TS] - f[B] - f[BE] - f[BW] - f[BN] - f[BS];
  const R rho = f[C] + f[N] + f[E] + f[S] + f[W] + f[NE] + f[SE] + f[SW] + f[NW] + f[T] +
                f[TN] + f[TE] + f[TS] + f[TW] + f[B] + f[BN] + f[BE] + f[BS] + f[BW];
  const R sx = keep * ux, sy = keep * uy, sz = keep * uz;
  const R D  = rho - (sx * sx + sy * sy + sz * sz) * R(3.0) / R(2.0);
  f[C]       = f[C] - p.we * (f[C] - p.w0 * D);
  const R w1D = p.w1 * D, w2D = p.w2 * D;
  auto pair = [&](int i, int j, R ui, R wX3, R wNH, R wD) {
    const R fi = f[i], fj = f[j];
    const R even = p.we * (R(0.5) * (fi + fj) - ui * ui * wNH - wD);
    const R odd  = p.wo * (R(0.5) * (fi - fj) - ui * wX3);
    f[i] = fi - even - odd;
    f[j] = fj - even + odd;
  };
  pair(N, S, sy, p.w1x3, p.w1nh, w1D);
  pair(E, W, sx, p.w1x3, p.w1nh, w1D);
  pair(T, B, sz, p.w1x3, p.w1nh, w1D);
  pair(NW, SE, -sx + sy, p.w2x3, p.w2nh, w2D);
  pair(NE, SW, sx + sy, p.w2x3, p.w2nh, w2D);
  pair(TW, BE, -sx + sz, p.w2x3, p.w2nh, w2D);
  pair(TE, BW, sx + sz, p.w2x3, p.w2nh, w2D);
  pair(TS, BN, -sy + sz, p.w2x3, p.w2nh, w2D);
  pair(TN, BS, sy + sz, p.w2x3, p.w2nh, w2D);
}

template <typename R>
__device__ __forceinline__ R forceOf(int d, R f1, R f2)
{
  if (cx(d) == 0)
    return R(0);
  const R v = wclass(d) == 1 ? f1 : f2;
  return cx(d) > 0 ? v : -v;
}

template <typename R>
__device__ __forceinline__ R wOf(int d)
{
  return d == C ? R(1.0) / R(3.0) : (wclass(d) == 1 ? R(1.0) / R(18.0) : R(1.0) / R(36.0));
}

// ---- forward: push step with the Brinkman collision, force on store --------
// The z-shifted stores of a wavefront straddle two lines, each completed by
// a neighbouring wavefront.  Stores of 8- and 4-byte elements go through
// L2 (cached), where the two halves of such a line meet, and 2-byte ones
// are non-temporal.  Measured forward step (porous channel, cached against
// non-temporal stores): 128^3 FP64 0.101 against 0.127 ms, FP32 0.054
// against 0.090; 256^3 FP64 0.96-1.06 against 1.11-1.13, FP32 0.49-0.50
// against 0.51; FP16 level at 128^3 and 2.5-3 % slower cached at 256^3
// and 384^3 (0.357-0.362 against 0.349, 1.064 against 1.038).
template <typename SF>
__global__ __launch_bounds__(256) void kBrinkmanStep(const SF *src, SF *dst,
    const typename Store<SF>::R *__restrict__ sfield, const MaskWin *__restrict__ mask,
    AdjParams<typename Store<SF>::R> p)
{
  using R  = typename Store<SF>::R;
  using St = Store<SF>;
  RowCtx c;
  if (!c.init(mask, p.nx, p.ny, p.nz, p.segs, p.remap))
    return;
  R f[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d)
    f[d] = St::load(getS<true>(&src[(uint64_t)d * p.plane + c.cell]), p.inv);
  collideBrinkman(f, p.trt, R(1) - sfield[c.cell]);
  const Dither dz = ditherOf(c.cell, p.seq, p.dither);
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    uint64_t i;
    R v;
    if (d == C) {
      i = (uint64_t)C * p.plane + c.cell;
      v = f[C];
    } else if (c.solidOf(cx(d), cy(d), cz(d))) {
      i = (uint64_t)inv(d) * p.plane + c.cell;
      v = f[d] + forceOf<R>(inv(d), p.f1, p.f2);
    } else {
      i = (uint64_t)d * p.plane + c.cellOf(cx(d), cy(d), cz(d));
      v = f[d] + forceOf<R>(d, p.f1, p.f2);
    }
    putS<(sizeof(SF) < 4)>(&dst[i], St::storeStep(v, p.scale, ditherDir(dz, d)));
  }
}

// ---- adjoint step ---------------------------------------------------------------
template <typename SF, typename SA>
__global__ __launch_bounds__(256) void kAdjointStep(const SF *fstate, const SA *lamIn, SA *lamOut,
    const typename Store<SF>::R *__restrict__ sfield, double *grad,
    const MaskWin *__restrict__ mask, AdjParams<typename Store<SF>::R> p)
{
  using R  = typename Store<SF>::R;
  using St = Store<SF>;
  RowCtx c;
  if (!c.init(mask, p.nx, p.ny, p.nz, p.segs, p.remap))
    return;
  R f[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d)
    f[d] = St::load(getS<true>(&fstate[(uint64_t)d * p.plane + c.cell]), p.inv);
  R mu[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    uint64_t i;
    if (k == C)
      i = (uint64_t)C * p.plane + c.cell;
    else if (c.solidOf(cx(k), cy(k), cz(k)))
      i = (uint64_t)inv(k) * p.plane + c.cell;
    else
      i = (uint64_t)k * p.plane + c.cellOf(cx(k), cy(k), cz(k));
    mu[k] = (R)getS<true>(&lamIn[i]);
  }
  const R ux = f[E] + f[NE] + f[SE] + f[TE] + f[BE] - f[W] - f[NW] - f[SW] - f[TW] - f[BW];
  const R uy = f[N] + f[NE] + f[NW] + f[TN] + f[BN] - f[S] - f[SE] - f[SW] - f[TS] - f[BS];
  const R uz = f[T] + f[TE] + f[TW] + f[TN] + f[TS] - f[B] - f[BE] - f[BW] - f[BN] - f[BS];
  const R keep = R(1) - sfield[c.cell];
  const R sx = keep * ux, sy = keep * uy, sz = keep * uz;
  const Trt<R> &t = p.trt;
  R M0 = 0, Vx = 0, Vy = 0, Vz = 0;
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    const R tk = wOf<R>(k) * mu[k];
    M0 += tk;
    if (k == C)
      continue;
    const R uk = R(cx(k)) * sx + R(cy(k)) * sy + R(cz(k)) * sz;
    const R q  = tk * (R(9.0) * t.we * uk + R(3.0) * t.wo);
    Vx += R(cx(k)) * q;
    Vy += R(cy(k)) * q;
    Vz += R(cz(k)) * q;
  }
  M0 *= t.we;
  const R a  = R(1) - R(0.5) * (t.we + t.wo);
  const R b  = R(0.5) * (t.wo - t.we);
  const R Wx = keep * (Vx - R(3.0) * M0 * sx), Wy = keep * (Vy - R(3.0) * M0 * sy),
          Wz = keep * (Vz - R(3.0) * M0 * sz);
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const R lam = a * mu[d] + b * mu[inv(d)] + M0 + R(cx(d)) * Wx + R(cy(d)) * Wy + R(cz(d)) * Wz;
    putS<true>(&lamOut[(uint64_t)d * p.plane + c.cell], (SA)lam);
  }
  const double dg = 3.0 * (double)M0 * ((double)sx * ux + (double)sy * uy + (double)sz * uz) -
                    ((double)ux * Vx + (double)uy * Vy + (double)uz * Vz);
  grad[c.cell] += dg;
}

// ---- objective at the final state: J partials and lambda^T -------------------------
template <typename SF, typename SA>
__global__ __launch_bounds__(256) void kObjective(const SF *fstate, SA *lam, Geom g,
    AdjParams<typename Store<SF>::R> p, double *partials)
{
  using St = Store<SF>;
  const uint64_t n    = (uint64_t)g.nx * g.ny * g.nz;
  const uint64_t cell = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double Jc           = 0.0;
  if (cell < n) {
    const int z      = (int)(cell % g.nz);
    const uint64_t r = cell / g.nz;
    const int y      = (int)(r % g.ny);
    const int x      = (int)(r / g.ny);
    const bool on    = x == p.outletX && !solidAt(g.mask, g.segs, g.ny, x, y, z);
    double rho = p.rhoRef, jx = 0.0, jy = 0.0, jz = 0.0;
    if (on) {
      for (int d = 0; d < Q; ++d) {
        const double v = (double)St::load(fstate[(uint64_t)d * p.plane + cell], p.inv);
        rho += v;
        jx += cx(d) * v;
        jy += cy(d) * v;
        jz += cz(d) * v;
      }
      const double j2 = jx * jx + jy * jy + jz * jz;
      Jc              = 0.5 * rho * j2;
      for (int d = 0; d < Q; ++d)
        lam[(uint64_t)d * p.plane + cell] = (SA)(0.5 * j2 + rho * (cx(d) * jx + cy(d) * jy + cz(d) * jz));
    } else {
      for (int d = 0; d < Q; ++d)
        lam[(uint64_t)d * p.plane + cell] = (SA)0.0;
    }
  }
  // block sum in a fixed order
  __shared__ double part[256];
  part[threadIdx.x] = Jc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partials[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(256) void kSum(const double *partials, int n, double *out)
{
  __shared__ double part[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256)
    v += partials[i];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *out = part[0];
}

template <typename R>
__global__ __launch_bounds__(256) void kFieldToR(const double *in, R *out, uint64_t n)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = (R)in[i];
}

} // namespace lbm

// ==================================================================================
// host side
// ==================================================================================
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

static void adjFail(const char *msg)
{
  fprintf(stderr, "ERROR: %s\n", msg);
  exit(EXIT_FAILURE);
}

struct LbmAdjoint {
  Geom g;
  int sf, sa;          // forward / adjoint storage kinds
  size_t esf, esa, esr; // element sizes (forward, adjoint, field)
  uint64_t ncells;
  double rhoRef;
  float scale, inv;
  int outletX, remap, dither;
  CaseDataType cd;
  void *sfield;        // R per cell
  double *grad;
  void *lam[2];
  void *state0;        // initial state
  void *work[2];       // forward ping-pong
  std::vector<void *> pool; // checkpoint / segment states
  size_t peak;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  double *dpart, *dout;
  int npart;
  double tForward, tRecompute, tAdjoint;
  int finalIn;         // work index holding the last forward state
};

static bool adjDouble(const LbmAdjoint *A) { return A->sf == ST_F64; }

template <typename R>
static AdjParams<R> adjParams(const LbmAdjoint *A)
{
  AdjParams<R> p;
  double td[9];
  float tf[9];
  const double omega = (double)A->cd.Omega, magic = A->cd.MagicParam;
  {
    const double w1 = 1.0 / 18.0, w2 = 1.0 / 36.0;
    td[0] = omega;
    td[1] = 1.0 / (0.5 + magic / (1.0 / omega - 0.5));
    td[2] = 1.0 / 3.0;
    td[3] = w1;
    td[4] = w2;
    td[5] = w1 * 3.0;
    td[6] = w2 * 3.0;
    td[7] = w1 * 9.0 / 2.0;
    td[8] = w2 * 9.0 / 2.0;
    const float o = (float)omega, m = (float)magic, v1 = 1.0f / 18.0f, v2 = 1.0f / 36.0f;
    tf[0] = o;
    tf[1] = 1.0f / (0.5f + m / (1.0f / o - 0.5f));
    tf[2] = 1.0f / 3.0f;
    tf[3] = v1;
    tf[4] = v2;
    tf[5] = v1 * 3.0f;
    tf[6] = v2 * 3.0f;
    tf[7] = v1 * 9.0f / 2.0f;
    tf[8] = v2 * 9.0f / 2.0f;
  }
  p.trt    = toTrt<R>(td, tf);
  const double g = (double)A->cd.XForce;
  p.f1     = std::is_same<R, double>::value ? (R)(3.0 * (1.0 / 18.0) * g) : (R)(3.0f * (1.0f / 18.0f) * (float)g);
  p.f2     = std::is_same<R, double>::value ? (R)(3.0 * (1.0 / 36.0) * g) : (R)(3.0f * (1.0f / 36.0f) * (float)g);
  p.scale  = (R)A->scale;
  p.inv    = (R)A->inv;
  p.rhoRef = A->rhoRef;
  p.plane  = A->g.plane;
  p.nx     = A->g.nx;
  p.ny     = A->g.ny;
  p.nz     = A->g.nz;
  p.segs   = A->g.segs;
  p.outletX = A->outletX;
  p.remap  = A->remap;
  p.dither = A->dither;
  p.seq    = 0;
  return p;
}

static size_t stateBytes(const LbmAdjoint *A) { return A->g.plane * Q * A->esf; }

static void *adjAlloc(LbmAdjoint *A, size_t bytes)
{
  void *p = nullptr;
  HIP_OK(hipMalloc(&p, bytes));
  A->peak += bytes;
  return p;
}

// forward step t -> t + 1.  The FP16 rounding seeds depend on t, so a state
// recomputed from a checkpoint equals the one of the first sweep bit for bit.
static void adjForwardStep(LbmAdjoint *A, const void *src, void *dst, int t)
{
  const unsigned blocks = rowBlocks(A->g, 4);
  const uint32_t seq    = (uint32_t)t * 0x9e3779b9u;
  switch (A->sf) {
  case ST_F64: {
    AdjParams<double> p = adjParams<double>(A);
    p.seq               = seq;
    hipLaunchKernelGGL(kBrinkmanStep<double>, dim3(blocks), dim3(256), 0, A->stream,
        (const double *)src, (double *)dst, (const double *)A->sfield, A->g.mask, p);
    break;
  }
  case ST_F32DEV: {
    AdjParams<float> p = adjParams<float>(A);
    p.seq              = seq;
    hipLaunchKernelGGL(kBrinkmanStep<f32dev>, dim3(blocks), dim3(256), 0, A->stream,
        (const f32dev *)src, (f32dev *)dst, (const float *)A->sfield, A->g.mask, p);
    break;
  }
  default: {
    AdjParams<float> p = adjParams<float>(A);
    p.seq              = seq;
    hipLaunchKernelGGL(kBrinkmanStep<__half>, dim3(blocks), dim3(256), 0, A->stream,
        (const __half *)src, (__half *)dst, (const float *)A->sfield, A->g.mask, p);
    break;
  }
  }
  HIP_OK(hipGetLastError());
}

static void adjBackwardStep(LbmAdjoint *A, const void *fstate, const void *lamIn, void *lamOut)
{
  const unsigned blocks = rowBlocks(A->g, 4);
  switch (A->sf) {
  case ST_F64:
    hipLaunchKernelGGL((kAdjointStep<double, double>), dim3(blocks), dim3(256), 0, A->stream,
        (const double *)fstate, (const double *)lamIn, (double *)lamOut, (const double *)A->sfield,
        A->grad, A->g.mask, adjParams<double>(A));
    break;
  case ST_F32DEV:
    hipLaunchKernelGGL((kAdjointStep<f32dev, float>), dim3(blocks), dim3(256), 0, A->stream,
        (const f32dev *)fstate, (const float *)lamIn, (float *)lamOut, (const float *)A->sfield,
        A->grad, A->g.mask, adjParams<float>(A));
    break;
  default:
    hipLaunchKernelGGL((kAdjointStep<__half, float>), dim3(blocks), dim3(256), 0, A->stream,
        (const __half *)fstate, (const float *)lamIn, (float *)lamOut, (const float *)A->sfield,
        A->grad, A->g.mask, adjParams<float>(A));
    break;
  }
  HIP_OK(hipGetLastError());
}

// J of a state (FP64 sum) and, when lam != nullptr, lambda^T = dJ/df into lam
static double adjObjective(LbmAdjoint *A, const void *fstate, void *lam)
{
  void *lamOut = lam ? lam : A->lam[1];
  const unsigned blocks = (unsigned)((A->ncells + 255) / 256);
  switch (A->sf) {
  case ST_F64:
    hipLaunchKernelGGL((kObjective<double, double>), dim3(blocks), dim3(256), 0, A->stream,
        (const double *)fstate, (double *)lamOut, A->g, adjParams<double>(A), A->dpart);
    break;
  case ST_F32DEV:
    hipLaunchKernelGGL((kObjective<f32dev, float>), dim3(blocks), dim3(256), 0, A->stream,
        (const f32dev *)fstate, (float *)lamOut, A->g, adjParams<float>(A), A->dpart);
    break;
  default:
    hipLaunchKernelGGL((kObjective<__half, float>), dim3(blocks), dim3(256), 0, A->stream,
        (const __half *)fstate, (float *)lamOut, A->g, adjParams<float>(A), A->dpart);
    break;
  }
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(kSum, dim3(1), dim3(256), 0, A->stream, A->dpart, (int)blocks, A->dout);
  HIP_OK(hipGetLastError());
  double J = 0.0;
  HIP_OK(hipMemcpyAsync(&J, A->dout, sizeof(double), hipMemcpyDeviceToHost, A->stream));
  HIP_OK(hipStreamSynchronize(A->stream));
  return J;
}

static AuxLaunch adjAux(const LbmAdjoint *A, void *pdf)
{
  AuxLaunch X;
  X.g        = A->g;
  X.storage  = A->sf;
  X.rule     = SR_LOCAL;
  X.pdf      = pdf;
  X.rhoRef   = A->rhoRef;
  X.scale    = A->scale;
  X.invScale = A->inv;
  return X;
}

static float eventSeconds(LbmAdjoint *A)
{
  float ms = 0.0f;
  HIP_OK(hipEventSynchronize(A->ev1));
  HIP_OK(hipEventElapsedTime(&ms, A->ev0, A->ev1));
  return ms * 1e-3f;
}

// run n forward steps from src = f^{t0}; the last state lands in dst (src
// is kept).  Outputs alternate between dst and a scratch buffer so that the
// last one is dst: step i writes dst when n - 1 - i is even.
static void adjForwardRun(LbmAdjoint *A, const void *src, void *dst, int n, int t0)
{
  if (n <= 0) {
    HIP_OK(hipMemcpyAsync(dst, src, stateBytes(A), hipMemcpyDeviceToDevice, A->stream));
    return;
  }
  void *scratch = A->work[0];
  if (scratch == dst || scratch == src)
    scratch = A->work[1];
  if (scratch == dst || scratch == src)
    adjFail("adjForwardRun: no scratch buffer");
  const void *cur = src;
  for (int i = 0; i < n; ++i) {
    void *out = ((n - 1 - i) % 2 == 0) ? dst : scratch;
    adjForwardStep(A, cur, out, t0 + i);
    cur = out;
  }
}

// lay checkpoints every K steps over [t0, t1) starting from `start`:
// ck[0] = start, ck[i] = f^{t0 + iK} in pool[base + i - 1]
static std::vector<void *> adjCheckpoints(LbmAdjoint *A, int K, int t0, int t1, void *start,
    size_t base)
{
  const int m = (t1 - t0 + K - 1) / K;
  std::vector<void *> ck(m);
  ck[0] = start;
  for (int i = 1; i < m; ++i) {
    ck[i] = A->pool[base + i - 1];
    adjForwardRun(A, ck[i - 1], ck[i], K, t0 + (i - 1) * K);
  }
  return ck;
}

static void adjBackwardRange(LbmAdjoint *A, const int *intervals, int nlevels, int level, int t0,
    int t1, void *start, size_t base, int *lamIdx);

// the segments of level `level`, last first, each one level deeper
static void adjSegments(LbmAdjoint *A, const int *intervals, int nlevels, int level, int t0,
    int t1, const std::vector<void *> &ck, size_t next, int *lamIdx)
{
  const int K = intervals[level];
  for (int i = (int)ck.size() - 1; i >= 0; --i) {
    const int a = t0 + i * K, b = MIN(t1, t0 + (i + 1) * K);
    adjBackwardRange(A, intervals, nlevels, level + 1, a, b, ck[i], next, lamIdx);
  }
}

// Adjoint over [t0, t1): lambda^{t1} is in A->lam[*lamIdx] on entry and
// lambda^{t0} on exit; start holds f^{t0}.  Level `level` of intervals
// splits the range into checkpointed segments; past the last level every
// state of the range is stored.  The pool holds the buffers from `base`.
static void adjBackwardRange(LbmAdjoint *A, const int *intervals, int nlevels, int level, int t0,
    int t1, void *start, size_t base, int *lamIdx)
{
  const int n = t1 - t0;
  if (level >= nlevels) {
    std::vector<void *> st(n);
    st[0] = start;
    HIP_OK(hipEventRecord(A->ev0, A->stream));
    for (int i = 1; i < n; ++i) {
      st[i] = A->pool[base + i - 1];
      adjForwardStep(A, st[i - 1], st[i], t0 + i - 1);
    }
    HIP_OK(hipEventRecord(A->ev1, A->stream));
    A->tRecompute += eventSeconds(A);
    HIP_OK(hipEventRecord(A->ev0, A->stream));
    for (int i = n - 1; i >= 0; --i) {
      adjBackwardStep(A, st[i], A->lam[*lamIdx], A->lam[*lamIdx ^ 1]);
      *lamIdx ^= 1;
    }
    HIP_OK(hipEventRecord(A->ev1, A->stream));
    A->tAdjoint += eventSeconds(A);
    return;
  }
  HIP_OK(hipEventRecord(A->ev0, A->stream));
  const std::vector<void *> ck = adjCheckpoints(A, intervals[level], t0, t1, start, base);
  HIP_OK(hipEventRecord(A->ev1, A->stream));
  A->tRecompute += eventSeconds(A);
  adjSegments(A, intervals, nlevels, level, t0, t1, ck, base + ck.size() - 1, lamIdx);
}

// buffers the recursion needs at most: level l holds ceil(n_l/K_l) - 1
// checkpoints besides its start, the last level n_last - 1 states
static size_t poolNeeded(const int *intervals, int nlevels, int T)
{
  size_t need = 0;
  int n       = T;
  for (int l = 0; l < nlevels; ++l) {
    const int K = intervals[l];
    need += (size_t)((n + K - 1) / K - 1);
    n = K < n ? K : n;
  }
  return need + (size_t)(n - 1);
}

extern "C" {

LbmAdjoint *lbmAdjointCreate(LatticeDescType *ld, CaseDataType *cd, int outletX)
{
  const int nx = ld->Dims[0], ny = ld->Dims[1], nz = ld->Dims[2];
  const int per[3] = { ld->PeriodicX, ld->PeriodicY, ld->PeriodicZ };
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const bool face = (!per[0] && (x == 0 || x == nx - 1)) ||
                          (!per[1] && (y == 0 || y == ny - 1)) ||
                          (!per[2] && (z == 0 || z == nz - 1));
        if (face && ld->Lattice[((size_t)z * ny + y) * nx + x] != LAT_CELL_OBSTACLE)
          adjFail("fluid cell on a non-periodic boundary face");
      }
  if (outletX < 0 || outletX >= nx)
    adjFail("lbmAdjointCreate: outlet plane outside the lattice");
  LbmAdjoint *A = new LbmAdjoint();
  A->cd         = caseDataEffective(cd);
  cd            = &A->cd;
  A->outletX    = outletX;
  const char *rm = getenv("LBM_XCD_REMAP");
  A->remap      = rm && *rm ? atoi(rm) : 1;
  const char *di = getenv("LBM_DITHER");
  A->dither     = di && *di ? atoi(di) : 1;
  switch (cd->Storage) {
  case LBM_STORE_NATIVE:
    A->sf = ST_F64;
    A->sa = ST_F64;
    break;
  case LBM_STORE_FP32:
    A->sf = ST_F32DEV;
    A->sa = ST_F32;
    break;
  case LBM_STORE_FP16:
    A->sf = ST_F16;
    A->sa = ST_F32;
    break;
  default:
    adjFail("lbmAdjointCreate: storage must be native (FP64), FP32 or FP16");
  }
  A->esf    = A->sf == ST_F64 ? 8 : (A->sf == ST_F16 ? 2 : 4);
  A->esa    = A->sa == ST_F64 ? 8 : 4;
  A->esr    = adjDouble(A) ? 8 : 4;
  A->rhoRef = adjDouble(A) ? 0.0 : cd->RhoRef;
  double sc = cd->StorageScale;
  if (!(sc > 0.0))
    sc = A->sf == ST_F16 ? 4096.0 : 1.0;
  if (adjDouble(A))
    sc = 1.0;
  int ex;
  if (frexp(sc, &ex) != 0.5)
    adjFail("StorageScale must be a power of two");
  A->scale   = (float)sc;
  A->inv     = (float)(1.0 / sc);
  A->ncells  = (uint64_t)nx * ny * nz;
  A->g.nx    = nx;
  A->g.ny    = ny;
  A->g.nz    = nz;
  A->g.segs  = (nz + 63) / 64;
  A->g.aos   = 0;
  A->g.plane = ((A->ncells + 63) / 64) * 64 + 65536 / A->esf;
  HIP_OK(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&A->ev0));
  HIP_OK(hipEventCreate(&A->ev1));
  const size_t nwin = (size_t)nx * ny * A->g.segs;
  MaskWin *mask     = (MaskWin *)adjAlloc(A, nwin * sizeof(MaskWin));
  int *dlat         = nullptr;
  HIP_OK(hipMalloc(&dlat, A->ncells * sizeof(int)));
  HIP_OK(hipMemcpy(dlat, ld->Lattice, A->ncells * sizeof(int), hipMemcpyHostToDevice));
  launchBuildMask(dlat, mask, nx, ny, nz, A->g.segs, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipFree(dlat));
  A->g.mask  = mask;
  A->sfield  = adjAlloc(A, A->ncells * A->esr);
  HIP_OK(hipMemsetAsync(A->sfield, 0, A->ncells * A->esr, A->stream));
  A->grad    = (double *)adjAlloc(A, A->ncells * sizeof(double));
  const size_t sb = stateBytes(A), lb = A->g.plane * Q * A->esa;
  A->state0  = adjAlloc(A, sb);
  A->work[0] = adjAlloc(A, sb);
  A->work[1] = adjAlloc(A, sb);
  A->lam[0]  = adjAlloc(A, lb);
  A->lam[1]  = adjAlloc(A, lb);
  for (void *p : { A->state0, A->work[0], A->work[1] })
    HIP_OK(hipMemsetAsync(p, 0, sb, A->stream));
  A->npart = (int)((A->ncells + 255) / 256);
  A->dpart = (double *)adjAlloc(A, sizeof(double) * (size_t)A->npart);
  A->dout  = (double *)adjAlloc(A, sizeof(double) * 4);
  A->finalIn = -1;
  HIP_OK(hipStreamSynchronize(A->stream));
  return A;
}

void lbmAdjointDestroy(LbmAdjoint *A)
{
  if (!A)
    return;
  HIP_OK(hipStreamSynchronize(A->stream));
  for (void *p : A->pool)
    HIP_OK(hipFree(p));
  for (void *p : { (void *)A->g.mask, A->sfield, (void *)A->grad, A->state0, A->work[0],
           A->work[1], A->lam[0], A->lam[1], (void *)A->dpart, (void *)A->dout })
    HIP_OK(hipFree(p));
  HIP_OK(hipEventDestroy(A->ev0));
  HIP_OK(hipEventDestroy(A->ev1));
  HIP_OK(hipStreamDestroy(A->stream));
  delete A;
}

void lbmAdjointSetField(LbmAdjoint *A, const double *s)
{
  double *tmp = nullptr;
  HIP_OK(hipMalloc(&tmp, A->ncells * sizeof(double)));
  HIP_OK(hipMemcpy(tmp, s, A->ncells * sizeof(double), hipMemcpyHostToDevice));
  const unsigned blocks = (unsigned)((A->ncells + 255) / 256);
  if (adjDouble(A))
    hipLaunchKernelGGL(kFieldToR<double>, dim3(blocks), dim3(256), 0, A->stream, tmp,
        (double *)A->sfield, A->ncells);
  else
    hipLaunchKernelGGL(kFieldToR<float>, dim3(blocks), dim3(256), 0, A->stream, tmp,
        (float *)A->sfield, A->ncells);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipFree(tmp));
}

void lbmAdjointSetState(LbmAdjoint *A, const double *f)
{
  double *tmp = nullptr;
  const size_t bytes = sizeof(double) * Q * A->ncells;
  HIP_OK(hipMalloc(&tmp, bytes));
  HIP_OK(hipMemcpy(tmp, f, bytes, hipMemcpyHostToDevice));
  launchScatter(adjAux(A, A->state0), tmp, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipFree(tmp));
}

void lbmAdjointInitState(LbmAdjoint *A, int mode, uint64_t seed, double rho, double amp)
{
  launchInit(adjAux(A, A->state0), mode, seed, rho, amp, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(A->stream));
}

double lbmAdjointForward(LbmAdjoint *A, int steps, double *seconds)
{
  HIP_OK(hipEventRecord(A->ev0, A->stream));
  adjForwardRun(A, A->state0, A->work[0], steps, 0);
  HIP_OK(hipEventRecord(A->ev1, A->stream));
  const double t = eventSeconds(A);
  if (seconds)
    *seconds = t;
  A->finalIn = 0;
  return adjObjective(A, A->work[0], nullptr);
}

void lbmAdjointGetState(LbmAdjoint *A, double *f)
{
  const void *src = A->finalIn >= 0 ? A->work[A->finalIn] : A->state0;
  double *tmp     = nullptr;
  const size_t bytes = sizeof(double) * Q * A->ncells;
  HIP_OK(hipMalloc(&tmp, bytes));
  launchGather(adjAux(A, (void *)src), tmp, A->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipMemcpy(f, tmp, bytes, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(tmp));
}

double lbmAdjointGradient(LbmAdjoint *A, int steps, const int *intervals, int nlevels,
    double *grad, double stats[4])
{
  if (steps < 1)
    adjFail("lbmAdjointGradient: steps must be >= 1");
  for (int l = 0; l < nlevels; ++l)
    if (intervals[l] < 1)
      adjFail("lbmAdjointGradient: checkpoint intervals must be >= 1");
  const size_t need = nlevels == 0 ? (size_t)(steps - 1) : poolNeeded(intervals, nlevels, steps);
  while (A->pool.size() < need)
    A->pool.push_back(adjAlloc(A, stateBytes(A)));
  A->tForward = A->tRecompute = A->tAdjoint = 0.0;
  HIP_OK(hipMemsetAsync(A->grad, 0, A->ncells * sizeof(double), A->stream));
  int lamIdx = 0;
  double J   = 0.0;
  if (nlevels == 0) {
    // every state kept: f^0 .. f^{T-1} in state0 and the pool, f^T in work[0]
    HIP_OK(hipEventRecord(A->ev0, A->stream));
    std::vector<void *> st(steps);
    st[0] = A->state0;
    for (int i = 1; i < steps; ++i) {
      st[i] = A->pool[i - 1];
      adjForwardStep(A, st[i - 1], st[i], i - 1);
    }
    adjForwardStep(A, st[steps - 1], A->work[0], steps - 1);
    HIP_OK(hipEventRecord(A->ev1, A->stream));
    A->tForward = eventSeconds(A);
    J           = adjObjective(A, A->work[0], A->lam[0]);
    HIP_OK(hipEventRecord(A->ev0, A->stream));
    for (int i = steps - 1; i >= 0; --i) {
      adjBackwardStep(A, st[i], A->lam[lamIdx], A->lam[lamIdx ^ 1]);
      lamIdx ^= 1;
    }
    HIP_OK(hipEventRecord(A->ev1, A->stream));
    A->tAdjoint = eventSeconds(A);
  } else {
    // the forward sweep lays the first level of checkpoints and ends at f^T
    HIP_OK(hipEventRecord(A->ev0, A->stream));
    const std::vector<void *> ck = adjCheckpoints(A, intervals[0], 0, steps, A->state0, 0);
    const int last = (int)(ck.size() - 1) * intervals[0];
    adjForwardRun(A, ck.back(), A->work[0], steps - last, last);
    HIP_OK(hipEventRecord(A->ev1, A->stream));
    A->tForward = eventSeconds(A);
    J           = adjObjective(A, A->work[0], A->lam[0]);
    adjSegments(A, intervals, nlevels, 0, 0, steps, ck, ck.size() - 1, &lamIdx);
  }
  A->finalIn = 0;
  HIP_OK(hipStreamSynchronize(A->stream));
  HIP_OK(hipMemcpy(grad, A->grad, A->ncells * sizeof(double), hipMemcpyDeviceToHost));
  if (stats) {
    stats[0] = A->tForward;
    stats[1] = A->tRecompute;
    stats[2] = A->tAdjoint;
    stats[3] = (double)A->peak;
  }
  return J;
}

void lbmAdjointStepTimes(LbmAdjoint *A, int reps, double out[2])
{
  HIP_OK(hipEventRecord(A->ev0, A->stream));
  for (int r = 0; r < reps; ++r)
    adjForwardStep(A, A->work[r & 1], A->work[(r + 1) & 1], r);
  HIP_OK(hipEventRecord(A->ev1, A->stream));
  out[0] = eventSeconds(A) / reps;
  HIP_OK(hipEventRecord(A->ev0, A->stream));
  for (int r = 0; r < reps; ++r)
    adjBackwardStep(A, A->work[0], A->lam[r & 1], A->lam[(r + 1) & 1]);
  HIP_OK(hipEventRecord(A->ev1, A->stream));
  out[1] = eventSeconds(A) / reps;
}

size_t lbmAdjointDeviceBytes(LbmAdjoint *A) { return A->peak; }

} // extern "C"
