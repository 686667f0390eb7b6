/* lbm_hip.h -- entry points this implementation adds to the lbm.h
 * boundary.  None of them has a counterpart in LBM-Bench; each says what
 * reference behaviour it stands in for.  Plain C types only.
 */
#ifndef LBM_AMD_LBM_HIP_H
#define LBM_AMD_LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bulk logical PDF access.  Element ((x*nY + y)*nZ + z)*19 + d of the
 * float64 array is f_d of cell (x, y, z), exactly what GetNode/SetNode
 * (src/lbm.c:193-294) return/take for that cell; obstacle cells read as 0
 * and are ignored on write.  Stands in for a GetNode/SetNode loop over all
 * cells (src/lbm.c:928-947 style). */
void lbmGetNodes(LatticeDescType *ld, KernelDataType *kd, double *out);
void lbmSetNodes(LatticeDescType *ld, KernelDataType *kd, const double *in);

/* Initial state written on the device, without a host copy.
 *   LBM_INIT_ZERO    f = 0 (the reference's benchmark start, src/lbm.c:452-463)
 *   LBM_INIT_REST    f_i = w_i * rho
 *   LBM_INIT_RANDOM  f_i = w_i * (rho + eps_i), eps_i = (2u - 1) * amp with u
 *                    from the splitmix64 chain over (seed, cell, i),
 *                    cell = (x*nY + y)*nZ + z */
enum { LBM_INIT_ZERO = 0, LBM_INIT_REST = 1, LBM_INIT_RANDOM = 2 };
void lbmInitPdfs(LatticeDescType *ld, KernelDataType *kd, int mode, uint64_t seed,
    double rho, double amp);

/* FP64 sums over fluid cells on the device: out = {sum rho, sum jx, sum jy,
 * sum jz}, j = sum c_i f_i.  Generalises kernelDensity (src/lbm.c:928-947),
 * which only returns the mean density. */
void lbmMassMomentum(LatticeDescType *ld, KernelDataType *kd, double out[4]);

/* Momentum fields of two states on the same lattice, compared on the
 * device: out = {sum |jA - jB|^2, sum |jB|^2, max |jA - jB|,
 * sum (rhoA - rhoB)^2}.  Used to state the error of a storage form
 * against the FP64 run. */
void lbmCompareMomentum(LatticeDescType *ld, KernelDataType *a, KernelDataType *b,
    double out[4]);

/* lbmCompareMomentum without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same four
 * outputs, computed in one pass over both lattices (same dimensions, any
 * storage forms, both on the current device) with the partial sums in a
 * static device buffer of the library (1 MiB + 32 B per GPU, held until
 * the process ends): no hipMalloc or hipFree.  Lattices of more than 2^28
 * cells are refused.  The buffer is shared by every call on one GPU, so
 * calls that overlap on one GPU race.  Returns 0, or -1 after a message
 * on stderr (dimensions differ, lattice too large, no device). */
int lbmCompareMomentumNoAlloc(LatticeDescType *ld, KernelDataType *a, KernelDataType *b,
    double out[4]);

/* Per-cell density and momentum: out[((x*nY + y)*nZ + z)*4 + k] holds rho,
 * jx, jy, jz of the logical PDFs (j is the velocity of the incompressible
 * model); obstacle cells hold 0.  The velocity field the reference can
 * only form through a GetNode loop (src/lbm.c:964-1023 does it for one
 * plane in kernelVerify). */
void lbmCellMoments(LatticeDescType *ld, KernelDataType *kd, double *out);

/* x momentum of every cell of the plane x: ux[y*nZ + z] = j_x of the
 * logical PDFs, f_E + f_NE + f_SE + f_TE + f_BE - f_W - f_NW - f_SW - f_TW
 * - f_BW summed in that order in PdfType (the profile src/lbm.c:976-988
 * builds through GetNode), 0 on obstacle cells.  Reduced on the device:
 * only the plane's nY nZ values reach the host.  kernelVerify uses it. */
void lbmPlaneVelocity(LatticeDescType *ld, KernelDataType *kd, int x, double *ux);

/* On-disk state: the logical PDFs and the cell flags in one file, for
 * resume and external validation.  The reference has no state file.
 * Layout, little-endian:
 *   bytes   0-7    magic "LBMSTAT1"
 *           8-11   uint32 format version (1)
 *          12-15   uint32 header size (128)
 *          16-27   int32 Dims[3]
 *          28-39   int32 Periodic[3]
 *          40-43   int32 19 (velocities)
 *          44-47   int32 kd->Iteration when saved
 *          48-71   float64 Omega, XForce, MagicParam of cd (informational)
 *          72-79   uint64 F: bytes of the flag section
 *          80-87   uint64 P: bytes of the PDF section
 *          88-95   uint64 checksum
 *          96-127  kernel name, NUL padded
 *   then   F bytes: int8 LatticeDescType flag per cell in latticeIndex
 *                   order (x fastest), zero-padded to a multiple of 8
 *   then   P bytes: float64 logical PDFs [x][y][z][19], as lbmGetNodes
 * The checksum is FNV-1a over the little-endian 64-bit words of the two
 * sections (offset basis 0xcbf29ce484222325, prime 0x100000001b3).
 * lbmSaveState streams the lattice to `path` in x-slabs; lbmLoadState
 * checks magic, version, dimensions and that the obstacle cells are the
 * same, then loads the PDFs, so a run resumed from a file continues as the
 * uninterrupted run would.  Both return 0, or -1 after a message on stderr
 * (I/O error, mismatch, truncated or trailing data, bad checksum).
 * lbmLoadState reads the PDF section twice: the first pass checks its
 * length and checksum, and only a file that passes is loaded, so a failed
 * load leaves the lattice's PDFs as they were (unless the file changes
 * between the passes, which is reported). */
int lbmSaveState(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, const char *path);
int lbmLoadState(LatticeDescType *ld, KernelDataType *kd, const char *path);

/* Runs `steps` steps (as Kernel with MaxIterations = steps) and returns the
 * hipEvent time of the launches in seconds. */
double lbmRunSteps(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, int steps);

/* lbmRunSteps with every step launched on its own (no graph replay) and a
 * hipEvent recorded on the library's stream after each launch: returns the
 * run's time as lbmRunSteps does, and kernelSeconds[0] / launches[0] the
 * summed durations and number of the step launches (aa-soa: its even
 * steps), kernelSeconds[1] / launches[1] those of aa-soa's odd steps (0
 * for the other variants).  A pressure-boundary pass after a step counts
 * with that step.  For a per-kernel roofline. */
double lbmRunStepsTimed(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, int steps,
    double kernelSeconds[2], int launches[2]);

/* bytes of device memory held by the PDF arrays of kd (with push-aos's
 * post-collision planes when it runs in two passes) */
size_t lbmDeviceBytes(KernelDataType *kd);

/* How the PDF arrays were placed: *candidates allocations were probed with
 * the step's 19-stream pattern and the fastest kept, at *gbs GB/s (1 and
 * 0.0 when no probe ran: arrays below 4 GiB, AoS, LBM_PLACEMENT_PROBE < 2).
 * No counterpart in the reference. */
void lbmPlacement(KernelDataType *kd, int *candidates, double *gbs);

/* Plain device-to-device copy of `bytes` bytes, repeated `reps` times on one
 * stream; returns GB/s (read + write bytes / second) from hipEvents.  The
 * achievable-bandwidth line beside the roofline. */
double lbmCopyBandwidth(size_t bytes, int reps);

/* lbmCopyBandwidth with the copy's shape chosen: `vectors` 16-byte vectors
 * per lane (1, 2, 4 or 8; lbmCopyBandwidth uses 4), nonTemporal != 0 for
 * non-temporal loads and stores (else cached), persistentGroups > 0 for a
 * persistent grid of that many workgroups striding over the array (0: one
 * workgroup per tile).  Returns GB/s, or -1 after a message on stderr for
 * another shape.  The sweep behind the achievable line. */
double lbmCopyBandwidthShape(size_t bytes, int reps, int vectors, int nonTemporal,
    int persistentGroups);

/* The memory traffic of one SoA step without its arithmetic, as GB/s (read
 * + write bytes / second, hipEvents over `reps` launches): every lane loads
 * element `cell` of `planes` planes of `cells` elements and stores it back,
 * to the reversed planes of the same array (inPlace != 0, the AA step's
 * traffic) or to a second array (push and pull).  planes 1, 2, 4, 8 or 19;
 * elementBytes 2, 4 or 8; returns -1 after a message on stderr otherwise.
 * The ceiling the step kernels are measured against, beside the copy. */
double lbmPatternBandwidth(int planes, int elementBytes, int inPlace, size_t cells, int reps);

/* hipDeviceSynchronize; 0 on success */
int lbmDeviceSynchronize(void);

/* name and compute-unit count of the current device; 0 on success */
int lbmDeviceInfo(char *name, int len, int *computeUnits, size_t *globalMem);

/* makes HIP device `index` current for the calls that follow (Init,
 * measurement entry points); 0 on success, -1 after a message on stderr
 * when there is no such device */
int lbmSelectDevice(int index);

/* ---------------------------------------------------------------------
 * Differentiation form (no counterpart in LBM-Bench).
 *
 * A Brinkman-penalised TRT forward model on the same lattice, mask, force
 * and periodicity rules as the kernels above, and its discrete adjoint.
 * The collision uses j_s = (1 - s) j in place of j (s = solid fraction per
 * cell, 0 gives the reference's TRT exactly); the objective is
 *   J = sum over the fluid cells of the plane x = outletX of 1/2 rho |j|^2
 * at the final step.  Storage (CaseDataType.Storage): native -> FP64
 * forward and adjoint; FP32 -> FP32 deviation forward, FP32 adjoint;
 * FP16 -> FP16 deviation forward, FP32 adjoint.  Gradients accumulate in
 * FP64.  Arrays: fields are float64 [x][y][z]; states float64 logical
 * [x][y][z][19] as lbmGetNodes.
 * ------------------------------------------------------------------- */
typedef struct LbmAdjoint LbmAdjoint;

LbmAdjoint *lbmAdjointCreate(LatticeDescType *ld, CaseDataType *cd, int outletX);
void lbmAdjointDestroy(LbmAdjoint *adj);
void lbmAdjointSetField(LbmAdjoint *adj, const double *s);
void lbmAdjointSetState(LbmAdjoint *adj, const double *f);
void lbmAdjointInitState(LbmAdjoint *adj, int mode, uint64_t seed, double rho, double amp);
/* steps forward from the initial state; returns J of the final state */
double lbmAdjointForward(LbmAdjoint *adj, int steps, double *seconds);
/* the state after the last forward run (or the initial state) */
void lbmAdjointGetState(LbmAdjoint *adj, double *f);
/* dJ/ds of a `steps`-step run into grad; returns J.  intervals[0..nlevels)
 * are the checkpoint spacings per level (e.g. {32} for sqrt(T) at
 * T = 1024; nlevels = 0 keeps every state).  stats = {forward s,
 * recompute s, adjoint s, device bytes} */
double lbmAdjointGradient(LbmAdjoint *adj, int steps, const int *intervals, int nlevels,
    double *grad, double stats[4]);
/* mean time of one forward step and one adjoint step over `reps` launches */
void lbmAdjointStepTimes(LbmAdjoint *adj, int reps, double out[2]);
size_t lbmAdjointDeviceBytes(LbmAdjoint *adj);

/* Porous solid-fraction field: u ~ U[0,1) per cell from the splitmix64
 * chain over (seed, cell), Gaussian-smoothed with standard deviation sigma
 * (periodic, radius 3 sigma), standardised to zero mean and unit variance
 * (z), then s = 1 / (1 + exp(-(a z + b))).  Output float64 [x][y][z] in
 * host memory.  Noise, smoothing and the logistic run on the current
 * device (temporary buffers of 2 n doubles); mean and variance are summed
 * on the host in cell order.  Dimensions below 1 print an error and leave
 * s untouched, before any device call. */
void lbmPorousField(const int dims[3], uint64_t seed, double sigma, double a, double b,
    double *s);

#ifdef __cplusplus
}
#endif

#endif
