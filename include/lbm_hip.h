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

/* Runs `steps` steps (as Kernel with MaxIterations = steps) and returns the
 * hipEvent time of the launches in seconds. */
double lbmRunSteps(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, int steps);

/* bytes of device memory held by the PDF arrays of kd */
size_t lbmDeviceBytes(KernelDataType *kd);

/* Plain device-to-device copy of `bytes` bytes, repeated `reps` times on one
 * stream; returns GB/s (read + write bytes / second) from hipEvents.  The
 * achievable-bandwidth line beside the roofline. */
double lbmCopyBandwidth(size_t bytes, int reps);

/* hipDeviceSynchronize; 0 on success */
int lbmDeviceSynchronize(void);

/* name and compute-unit count of the current device; 0 on success */
int lbmDeviceInfo(char *name, int len, int *computeUnits, size_t *globalMem);

#ifdef __cplusplus
}
#endif

#endif
