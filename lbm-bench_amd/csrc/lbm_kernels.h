// lbm_kernels.h -- launch interface between the host runtime
// (lbm_runtime.hip) and the kernel translation units (lbm_step.hip,
// lbm_aux.hip).  Plain structs, no templates, so each kernel
// file instantiates its own templates and the runtime stays small.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_device.h"

namespace lbm {

// one stream-and-collide step
enum Variant { V_PULL = 0, V_PUSH = 1, V_AA_EVEN = 2, V_AA_ODD = 3 };

// element type of the PDF arrays
enum StorageKind { ST_F64 = 0, ST_F32 = 1, ST_F32DEV = 2, ST_F16 = 3, ST_BF16 = 4 };

// where logical f_d(x) is stored between steps (LBM-Bench's GetNode,
// src/lbm.c:193-294): locally (push, AA after an odd step, every variant
// before its first step), at x - c_d (pull), or AA after an even step.
enum SlotRule { SR_LOCAL = 0, SR_PULL = 1, SR_AA_EVEN = 2 };

struct Geom {
  const MaskWin *mask;
  uint64_t plane; // SoA elements per direction (>= nx*ny*nz)
  int nx, ny, nz, segs;
  int aos;
};

struct StepLaunch {
  Geom g;
  int variant, tiled, storage;
  void *src, *dst; // AA: src == dst
  double trt[9];   // we, wo, w0, w1, w2, w1x3, w2x3, w1nh, w2nh
  float trtf[9];   // the same in FP32, formed in FP32
  double f1, f2;   // body force terms 3*w1*g and 3*w2*g
  float f1f, f2f;
  float scale, invScale;
  int addForce;
  // the run's last step (push or AA odd, storage without stochastic
  // rounding): add this step's own body force at the store, so that no
  // closing force pass is needed
  int close;
  // launch shape: XCD-aware block order, non-temporal stores (bit 0) /
  // loads (bit 1), wavefronts per workgroup (4 or 8)
  int remap, nt, wavesPerBlock;
  int band;    // AoS row order: bands of `band` y-rows with x swept inside each (0: y fastest)
  int wgPerCu; // AoS: resident workgroups per CU, limited through the LDS allocation (0: none)
  void *post;  // push-aos in two passes: Q planes of nX nY nZ post-collision values (else null)
  int rowSweep; // push / AA odd SoA with nZ % 64 == 0, aligned stores: 1 one wavefront per
                // row (kRowSweep), 3 one workgroup per row (kRowGroup, nZ <= 512)
  int dither;    // FP16/BF16 storage: stochastic rounding in the step
  int wallNt;    // kRowGroup AA odd, 8-byte, nt 3: non-temporal bits of the blocks near a
                 // solid, stores (1) / aligned gathers (2) / bounce-back loads (4); -1: 1
  // rounding seeds: step number = *seqBase + seq (seqBase in device memory,
  // advanced by every replay of a step graph; see launchBumpSeq)
  const uint32_t *seqBase;
  uint32_t seq;
};

void launchStepRow(const StepLaunch &L, hipStream_t s);  // lbm_step.hip

// auxiliary passes over every fluid cell (lbm_aux.hip)
struct AuxLaunch {
  Geom g;
  int storage, rule;
  void *pdf;
  double rhoRef;       // shifted-deviation reference density
  float scale, invScale;
  int dither;          // kForce: stochastic rounding of FP16/BF16, seeded by seq
  uint32_t seq;
};
// f_d += 3 w_d c_dx g at the slot holding logical f_d (src/lbm.c:902-922)
void launchForce(const AuxLaunch &A, double f1, double f2, hipStream_t s);
// *counter += n, one lane: the last node of a step graph, so that each
// replay seeds its rounding with new step numbers
void launchBumpSeq(uint32_t *counter, uint32_t n, hipStream_t s);
// Zou-He pressure inlet (x = 0) / outlet (x = nX-1) on the logical PDFs
// (src/lbm.c:676-785); face[y*nZ + z]: bit 0 inlet at (0, y, z), bit 1
// outlet at (nX-1, y, z).  pending: the state lacks the force 3 w c_x g
// (f1 = 3 w1 g, f2 = 3 w2 g) of the step just run.  FP16/BF16 stores round
// stochastically, seeded by *seqBase + seq.
struct ZouHeLaunch {
  const uint8_t *face;
  double rhoIn, rhoOut, f1, f2;
  int pending;
  const uint32_t *seqBase;
  uint32_t seq;
};
void launchZouHe(const AuxLaunch &A, const ZouHeLaunch &Z, hipStream_t s);
// cell flags of geometryCreate into lat (x fastest; 0 obstacle, 1 fluid):
// boundary layers of non-periodic axes, pipe (pipe != 0), blocks of edge
// `blocks` (> 0), nSpheres spheres (sph: centre and radius, box: lo / hi
// cell per axis); *count (zeroed by the caller) += fluid cells
void launchGeometry(int *lat, const int dims[3], const int per[3], int pipe, int blocks,
    int nSpheres, const double *sph, const int *box, unsigned long long *count, hipStream_t s);
// ux[y*nZ + z]: x momentum of the logical PDFs of plane x (0 on obstacles)
void launchPlaneUx(const AuxLaunch &A, int x, double *ux, hipStream_t s);
// fill logical PDFs: mode 0 zero, 1 w_d*rho, 2 w_d*(rho + (2u-1)*amp)
void launchInit(const AuxLaunch &A, int mode, uint64_t seed, double rho, double amp,
    hipStream_t s);
// per-block FP64 partial sums {rho, jx, jy, jz} of the deviation moments
// into partials[4*blocks]; optionally the per-cell moments into field
// (4 planes of nx*ny*nz doubles, NaN on solid cells); returns block count
int launchMoments(const AuxLaunch &A, double *partials, double *field, hipStream_t s);
int momentsBlocks(const Geom &g);
// sum partials[4*n] into out[4] in a fixed order (deterministic)
void launchReduce4(const double *partials, int n, double *out, hipStream_t s);
// compare two moment fields: partials of {sum |dj|^2, sum |jB|^2,
// max |dj|, sum drho^2}; returns block count
int launchCompareFields(const double *fa, const double *fb, uint64_t ncells, double *partials,
    hipStream_t s);
void launchReduceCompare(const double *partials, int n, double *out, hipStream_t s);
// the comparison above straight from two lattices of the same dimensions
// (no field buffers): partials[4 * k ...] of the k-th chunk of `chunk`
// cells, ceil(cells / chunk) chunks
void launchCompareLattices(const AuxLaunch &A, const AuxLaunch &B, uint64_t chunk,
    double *partials, hipStream_t s);
// plain device copy, for the measured bandwidth line and counter scaling
void launchCopy(const double *src, double *dst, uint64_t n, hipStream_t s);
// copies of other shapes: `vectors` 16-byte vectors per lane (1, 2, 4, 8),
// non-temporal or cached, one workgroup per tile or a persistent grid of
// persistentGroups workgroups (0: one per tile); false for other shapes
bool launchCopyShape(const double *src, double *dst, uint64_t n, int vectors, int nt,
    int persistentGroups, hipStream_t s);
// porous solid-fraction field: uniform noise from the splitmix64 chain over
// (seed, cell), one periodic Gaussian pass per axis (field [x][y][z], taps
// kern[0 .. 2 rad]), and s = 1 / (1 + exp(-(a (s - mean) / sd + b)))
void launchPorousNoise(double *s, uint64_t n, uint64_t seed, hipStream_t st);
void launchSmoothAxis(int axis, const double *in, double *out, const int dims[3],
    const double *kern, int rad, hipStream_t st);
void launchLogistic(double *s, uint64_t n, double mean, double sd, double a, double b,
    hipStream_t st);
// the step's memory traffic without arithmetic over `planes` planes of
// `cells` elements of `elementBytes` bytes (planes 1, 2, 4, 8 or 19;
// elementBytes 2, 4 or 8; false for any other); src == dst when in place
bool launchPattern(int planes, int elementBytes, const void *src, void *dst, uint64_t cells,
    bool inPlace, hipStream_t s);
// logical PDFs of every cell to / from a float64 [cell][19] buffer
// (obstacle cells: written as 0 / skipped), converting the storage form
// logical [cell][19] float64 of the cells [first, first + count) (x slowest;
// obstacle cells read as 0 and are not written)
void launchGather(const AuxLaunch &A, double *out, hipStream_t s, uint64_t first = 0,
    uint64_t count = ~0ull);
void launchScatter(const AuxLaunch &A, const double *in, hipStream_t s, uint64_t first = 0,
    uint64_t count = ~0ull);
// per-cell rho, jx, jy, jz (float64 [cell][4]) of the cells [first, first + count)
void launchCellMoments(const AuxLaunch &A, double *out, hipStream_t s, uint64_t first = 0,
    uint64_t count = ~0ull);
// solid-mask windows from LatticeDescType.Lattice flags (x fastest):
// window (row, seg) bit k = (flag at z = (64*seg - 1 + k) mod nz) == obstacle
void launchBuildMask(const int *lattice, MaskWin *mask, int nx, int ny, int nz, int segs,
    hipStream_t s);

} // namespace lbm
