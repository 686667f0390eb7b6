/* lbm.h -- lattice, case and kernel interface of the MI355X LBM-Bench path.
 *
 * This header is the drop-in boundary.  It keeps the names, argument
 * meaning and ownership rules of LBM-Bench's src/lbm.h so that a driver
 * written against the reference (src/main.c) compiles against it
 * unchanged; every kernel registered in GKernels[] runs on the GPU.
 *
 * Reference interface each declaration replaces (file:line in LBM-Bench):
 *   PdfType / F()                   src/lbm.h:10-16
 *   D3Q19_* direction ids           src/lbm.h:18-38   (same numbering)
 *   D3Q19X/Y/Z/Inv                  src/lbm.h:40-43, src/lbm.c:20-29
 *   LAT_CELL_*                      src/lbm.h:47-52
 *   LatticeDescType                 src/lbm.h:54-66
 *   CaseDataType                    src/lbm.h:68-75   (fields appended)
 *   KernelDataType                  src/lbm.h:77-97   (fields appended)
 *   KernelFunctionsType / GKernels  src/lbm.h:99-106, src/lbm.c:33-43
 *   latticeIndex                    src/lbm.h:108-111
 *   geometryCreate                  src/lbm.h:130, src/lbm.c:52-187
 *   kernelDeinit                    src/lbm.h:132, src/lbm.c:660-670
 *   kernelSetInitialDensity         src/lbm.h:133, src/lbm.c:791-843
 *   kernelSetInitialVelocity        src/lbm.h:134, src/lbm.c:845-896
 *   kernelAddBodyForce              src/lbm.h:137, src/lbm.c:902-922
 *   kernelDensity                   src/lbm.h:138, src/lbm.c:928-947
 *   kernelStatistics                src/lbm.h:139-140, src/lbm.c:953-958
 *   kernelVerify                    src/lbm.h:141-142, src/lbm.c:964-1023
 *   PROP_* / LAYOUT_*               src/lbm.h:144-145
 *   kernelInit<Variant>             src/lbm.h:147-153, src/lbm.c:601-658
 *   kernel<Variant>                 src/lbm.h:155-161, src/kernel_*.c
 *
 * Differences a caller can observe (all deliberate):
 *   - PDFs live in device memory (hipMalloc).  Pdfs[]/PdfsActive point at a
 *     host mirror that GetNode/SetNode use; it is allocated on first use and
 *     synchronised with the device lazily (download after a run, upload
 *     before the next run).
 *   - There is no halo: GlobalDims == Dims and Offsets == 0.  Periodicity is
 *     index wrap-around and half-way bounce-back is folded into the
 *     gather/scatter through a bit-packed solid mask, so BounceBackPdfsSrc/
 *     Dst stay NULL and nBounceBackPdfs is 0.
 *   - The TRT magic parameter is CaseDataType.MagicParam (the reference
 *     hard-codes 1/12, src/kernel_pull_soa.c:20) and the body force is
 *     switched at run time by CaseDataType.BodyForce (the reference compiles
 *     it in with -DVERIFICATION, src/kernel_pull_soa.c:206-209).
 *   - LoopBalance is the real traffic per lattice update: 2*19*sizeof(store)
 *     for every variant, AA included (the reference reports half of that
 *     for aa-soa, src/lbm.c:657).
 *   - Index arithmetic is 64-bit, so lattices beyond 480^3 work.
 *   - Errors print a message on stderr and exit(EXIT_FAILURE), as the
 *     reference's allocate() does (src/allocate.c:17-31).
 */
#ifndef LBM_AMD_LBM_H
#define LBM_AMD_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef PRECISION_SP
typedef float PdfType;
#else
typedef double PdfType;
#endif

#define F(number) (PdfType)(number)

#define N_D3Q19 19

#define D3Q19_N 0
#define D3Q19_S 1
#define D3Q19_E 2
#define D3Q19_W 3
#define D3Q19_NE 4
#define D3Q19_SE 5
#define D3Q19_NW 6
#define D3Q19_SW 7
#define D3Q19_T 8
#define D3Q19_TN 9
#define D3Q19_TE 10
#define D3Q19_TW 11
#define D3Q19_TS 12
#define D3Q19_B 13
#define D3Q19_BS 14
#define D3Q19_BN 15
#define D3Q19_BW 16
#define D3Q19_BE 17
#define D3Q19_C 18

extern int D3Q19X[N_D3Q19];
extern int D3Q19Y[N_D3Q19];
extern int D3Q19Z[N_D3Q19];
extern int D3Q19Inv[N_D3Q19];

typedef int LatticeType;

enum {
  LAT_CELL_OBSTACLE = 0,
  LAT_CELL_FLUID    = 1,
  LAT_CELL_INLET    = 2,
  LAT_CELL_OUTLET   = 4
};

/* Storage forms of the PDF arrays (CaseDataType.Storage).
 * LBM_STORE_NATIVE keeps f_i in PdfType, exactly as the reference does.
 * The other forms keep the shifted deviation (f_i - w_i*RhoRef) times
 * StorageScale in HBM and collide in FP32. */
enum {
  LBM_STORE_NATIVE = 0,
  LBM_STORE_FP32   = 1,
  LBM_STORE_FP16   = 2,
  LBM_STORE_BF16   = 3
};

typedef struct LatticeDesc {
  int Dims[3];
  LatticeType *Lattice; /* flags, x fastest: latticeIndex() */
  int nCells;
  int nFluid;
  int nObst;
  int nInlet;
  int nOutlet;
  int PeriodicX;
  int PeriodicY;
  int PeriodicZ;
  const char *Name;
} LatticeDescType;

#define LBM_CASE_EXTENSIONS 0x4C424D31 /* "LBM1" */

typedef struct CaseData {
  PdfType Omega;
  PdfType RhoIn;
  PdfType RhoOut;
  PdfType XForce;
  int MaxIterations;
  int StatisticsModulus;
  /* --- appended by this implementation --- */
  int Extensions;      /* LBM_CASE_EXTENSIONS when the fields below are set
                          (caseDataDefaults); otherwise the library uses the
                          reference's behaviour: Lambda 1/12, native storage,
                          body force from the LBM_BODY_FORCE environment
                          variable (the reference's -DVERIFICATION switch) */
  double MagicParam;   /* TRT Lambda; the reference's value is 1/12       */
  int BodyForce;       /* 1: add 3 w_i c_ix XForce after every step        */
  int Storage;         /* LBM_STORE_*                                       */
  double StorageScale; /* power of two; 0 selects the default for Storage  */
  double RhoRef;       /* reference density of the shifted deviation       */
  int PressureBoundaries; /* 1: after every step, the Zou-He pressure inlet
                             (x = 0, RhoIn) and outlet (x = nX-1, RhoOut) of
                             kernelComputeBoundaryConditions on the cells
                             flagged LAT_CELL_INLET / LAT_CELL_OUTLET; the
                             reference defines it (src/lbm.c:676-785) but
                             its time loop never calls it */
} CaseDataType;

struct LbmDevice; /* opaque device-side state (device arrays, mask, stream) */

typedef struct KernelData {
  PdfType *Pdfs[2];
  PdfType *PdfsActive;
  int Dims[3];
  int GlobalDims[3];
  int Offsets[3];
  int *BounceBackPdfsSrc;
  int *BounceBackPdfsDst;
  int nBounceBackPdfs;
  int Blk[3];
  int Iteration;
  double LoopBalance;
  double Duration;
  void (*GetNode)(struct KernelData *kd, int x, int y, int z, PdfType *pdfs);
  void (*SetNode)(struct KernelData *kd, int x, int y, int z, PdfType *pdfs);
  void (*BoundaryConditionsGetPdf)(
      struct KernelData *kd, int x, int y, int z, int dir, PdfType *pdf);
  void (*BoundaryConditionsSetPdf)(
      struct KernelData *kd, int x, int y, int z, int dir, PdfType pdf);
  void (*Kernel)(LatticeDescType *ld, struct KernelData *kd, CaseDataType *cd);
  /* --- appended by this implementation --- */
  struct LbmDevice *Dev;
  const char *Name;
} KernelDataType;

typedef struct KernelFunctions {
  const char *Name;
  void (*Init)(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
  void (*Deinit)(LatticeDescType *ld, KernelDataType **kd);
} KernelFunctionsType;

extern KernelFunctionsType GKernels[];
extern const int G_N_KERNELS;

static inline int latticeIndex(int dims[3], int x, int y, int z)
{
  return z * dims[0] * dims[1] + y * dims[0] + x;
}

#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif

/* Sets CaseDataType's appended fields to the reference's behaviour:
 * Lambda = 1/12, no body force, native storage, RhoRef = 1, no pressure
 * boundaries in the time loop, and marks them
 * valid.  A driver that never calls it (LBM-Bench's own main.c) gets the
 * same behaviour, with the body force taken from LBM_BODY_FORCE. */
void caseDataDefaults(CaseDataType *cd);

/* The case as the kernels use it: cd itself when its appended fields are
 * marked valid, else cd's reference fields plus the defaults above. */
CaseDataType caseDataEffective(const CaseDataType *cd);

/* Geometry types: box, channel, pipe, blocks[-N], fluid (src/lbm.c:49-156)
 * and spheres[-N] (N seeded random spheres inside a channel, radius
 * U[4,12]; seed from geometrySetSeed()). */
void geometryCreate(const char *type, int dims[3], int periodic[3], LatticeDescType *ld);
void geometrySetSeed(uint64_t seed);
void geometryDestroy(LatticeDescType *ld);

void kernelDeinit(LatticeDescType *ld, KernelDataType **kd);
void kernelSetInitialDensity(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelSetInitialVelocity(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
/* Zou-He pressure inlet / outlet (src/lbm.c:676-785), applied once to the
 * current state on the device: the cells of x = 0 flagged LAT_CELL_INLET
 * (y, z in 1 .. n-2) get density RhoIn, those of x = nX-1 flagged
 * LAT_CELL_OUTLET density RhoOut.  CaseDataType.PressureBoundaries applies
 * the same after every step inside Kernel. */
void kernelComputeBoundaryConditions(
    KernelDataType *kd, LatticeDescType *ld, CaseDataType *cd);
void kernelAddBodyForce(KernelDataType *kd, LatticeDescType *ld, CaseDataType *cd);
PdfType kernelDensity(KernelDataType *kd, LatticeDescType *ld);
void kernelStatistics(
    KernelDataType *kd, LatticeDescType *ld, CaseDataType *cd, int iter);
void kernelVerify(
    LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd, PdfType *errorNorm);

enum { PROP_PUSH = 0, PROP_PULL = 1, PROP_AA = 2 };
enum { LAYOUT_SOA = 0, LAYOUT_AOS = 1 };

void kernelInitPushSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitPushAoS(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitPullSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitPullAoS(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitBlkPushSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitBlkPullSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);
void kernelInitAaSoA(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd);

void kernelPushSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelPushAoS(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelPullSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelPullAoS(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelBlkPushSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelBlkPullSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);
void kernelAaSoA(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd);

#ifdef __cplusplus
}
#endif

#endif
