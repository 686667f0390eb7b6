/* lbm_host.c -- host-side parts of the lbm.h boundary: the D3Q19 tables
 * and kernel registry, geometryCreate's type parsing and sphere
 * parameters, LBM-Bench's initial conditions, statistics and verification.
 * The cell flags and the verification profile are computed on the device
 * (lbm_geometry.h, lbmPlaneVelocity); the -V initialisers write through
 * the kernel's SetNode as the reference's lbm.c does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "lbm.h"
#include "lbm_geometry.h"
#include "lbm_hip.h"

/* velocity set and inverse directions (src/lbm.c:20-29) */
int D3Q19X[N_D3Q19]   = { 0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0 };
int D3Q19Y[N_D3Q19]   = { 1, -1, 0, 0, 1, -1, 1, -1, 0, 1, 0, 0, -1, 0, -1, 1, 0, 0, 0 };
int D3Q19Z[N_D3Q19]   = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, 0 };
int D3Q19Inv[N_D3Q19] = { D3Q19_S, D3Q19_N, D3Q19_W, D3Q19_E, D3Q19_SW, D3Q19_NW, D3Q19_SE,
  D3Q19_NE, D3Q19_B, D3Q19_BS, D3Q19_BW, D3Q19_BE, D3Q19_BN, D3Q19_T, D3Q19_TN, D3Q19_TS,
  D3Q19_TE, D3Q19_TW, D3Q19_C };

/* registry: same names and order as src/lbm.c:33-41; every entry is a GPU
 * kernel */
KernelFunctionsType GKernels[] = {
  { "pull-soa", kernelInitPullSoA, kernelDeinit },
  { "pull-aos", kernelInitPullAoS, kernelDeinit },
  { "push-soa", kernelInitPushSoA, kernelDeinit },
  { "push-aos", kernelInitPushAoS, kernelDeinit },
  { "blk-push-soa", kernelInitBlkPushSoA, kernelDeinit },
  { "blk-pull-soa", kernelInitBlkPullSoA, kernelDeinit },
  { "aa-soa", kernelInitAaSoA, kernelDeinit },
};
const int G_N_KERNELS = sizeof(GKernels) / sizeof(GKernels[0]);

void caseDataDefaults(CaseDataType *cd)
{
  cd->Extensions   = LBM_CASE_EXTENSIONS;
  cd->MagicParam   = 1.0 / 12.0;
  cd->BodyForce    = 0;
  cd->Storage      = LBM_STORE_NATIVE;
  cd->StorageScale = 0.0;
  cd->RhoRef       = 1.0;
  cd->PressureBoundaries = 0;
}

CaseDataType caseDataEffective(const CaseDataType *cd)
{
  CaseDataType e = *cd;
  if (cd->Extensions != LBM_CASE_EXTENSIONS) {
    caseDataDefaults(&e);
    const char *bf = getenv("LBM_BODY_FORCE");
    e.BodyForce    = bf != NULL && atoi(bf) != 0;
  }
  return e;
}

/* ------------------------------------------------------------------------ */
/* geometry (src/lbm.c:49-187) and the seeded sphere packing                 */
/* ------------------------------------------------------------------------ */
static uint64_t GGeometrySeed = 0;

void geometrySetSeed(uint64_t seed) { GGeometrySeed = seed; }

/* splitmix64 finalizer (Steele, Lea & Flood 2014) chained over (seed, a, b);
 * the same chain seeds the device initialiser */
static uint64_t mix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static double hashU01(uint64_t seed, uint64_t a, uint64_t b)
{
  return (double)(mix64(mix64(mix64(seed) ^ a) ^ b) >> 11) * (1.0 / 9007199254740992.0);
}

static const char *GGeoTypeStr[] = { "box", "channel", "pipe", "blocks", "fluid", "spheres" };

#define SPHERE_SALT 0x5350484552455321ull

/* centre and radius of sphere k (4 doubles) and the lowest and highest cell
 * of its bounding box per axis (6 ints) */
static void sphereParams(int k, const int dims[3], uint64_t seed, double *sph, int *box)
{
  double r;
  for (int a = 0; a < 3; ++a)
    sph[a] = hashU01(seed ^ SPHERE_SALT, (uint64_t)k, (uint64_t)a) * dims[a];
  r      = 4.0 + 8.0 * hashU01(seed ^ SPHERE_SALT, (uint64_t)k, 3);
  sph[3] = r;
  for (int a = 0; a < 3; ++a) {
    box[a]     = (int)floor(sph[a] - r);
    box[3 + a] = (int)ceil(sph[a] + r);
  }
}

void geometryCreate(const char *geometryType, int dims[3], int periodic[3], LatticeDescType *ld)
{
  int type = -1, blockSize = 8, nSpheres = 0;
  if (strncasecmp("channel", geometryType, 7) == 0)
    type = LBM_GEO_CHANNEL;
  else if (strncasecmp("box", geometryType, 3) == 0)
    type = LBM_GEO_BOX;
  else if (strncasecmp("pipe", geometryType, 4) == 0)
    type = LBM_GEO_PIPE;
  else if (strncasecmp("blocks", geometryType, 6) == 0) {
    type = LBM_GEO_BLOCKS;
    if (strlen(geometryType) > 7) {
      blockSize = atoi(&geometryType[7]);
      if (blockSize <= 0)
        blockSize = 8;
    }
  } else if (strncasecmp("fluid", geometryType, 5) == 0)
    type = LBM_GEO_FLUID;
  else if (strncasecmp("spheres", geometryType, 7) == 0) {
    type     = LBM_GEO_SPHERES;
    nSpheres = strlen(geometryType) > 8 ? atoi(&geometryType[8]) : 0;
    if (nSpheres < 0)
      nSpheres = 0;
  } else {
    printf("ERROR: unknown geometry type '%s'.\n", geometryType);
    exit(1);
  }

  const size_t ncells = (size_t)dims[0] * dims[1] * dims[2];
  ld->Dims[0]         = dims[0];
  ld->Dims[1]         = dims[1];
  ld->Dims[2]         = dims[2];
  ld->nCells          = (int)ncells;
  ld->Name            = GGeoTypeStr[type];
  LatticeType *lat    = (LatticeType *)malloc(sizeof(LatticeType) * ncells);
  double *sph         = (double *)malloc(sizeof(double) * 4 * (size_t)(nSpheres > 0 ? nSpheres : 1));
  int *box            = (int *)malloc(sizeof(int) * 6 * (size_t)(nSpheres > 0 ? nSpheres : 1));
  if (lat == NULL || sph == NULL || box == NULL) {
    fprintf(stderr, "Error: Insufficient memory to fulfill the request\n");
    exit(EXIT_FAILURE);
  }
  ld->Lattice = lat;

  /* periodicity forced by the type (src/lbm.c:91-102) */
  if (type == LBM_GEO_CHANNEL || type == LBM_GEO_BLOCKS || type == LBM_GEO_PIPE ||
      type == LBM_GEO_SPHERES)
    periodic[0] = 1;
  else if (type == LBM_GEO_FLUID)
    periodic[0] = periodic[1] = periodic[2] = 1;

  for (int k = 0; k < nSpheres; ++k)
    sphereParams(k, dims, GGeometrySeed, sph + 4 * k, box + 6 * k);
  /* the cells themselves are built on the device and copied into lat */
  const long long nFluid = lbmBuildFlags(dims, periodic, type, blockSize, nSpheres, sph, box, lat);
  free(sph);
  free(box);

  ld->PeriodicX = periodic[0];
  ld->PeriodicY = periodic[1];
  ld->PeriodicZ = periodic[2];
  ld->nFluid    = (int)nFluid;
  ld->nObst     = (int)((long long)ncells - nFluid);
  ld->nInlet    = 0;
  ld->nOutlet   = 0;
}

void geometryDestroy(LatticeDescType *ld)
{
  free(ld->Lattice);
  ld->Lattice = NULL;
}

/* ------------------------------------------------------------------------ */
/* initial conditions (src/lbm.c:791-896): the incompressible equilibrium     */
/* scaled by omega, written through SetNode.                                  */
/* ------------------------------------------------------------------------ */
static void equilibriumScaled(PdfType dens, PdfType ux, PdfType uy, PdfType uz, PdfType omega,
    PdfType *pdfs)
{
  const PdfType w0 = F(1.0) / F(3.0), w1 = F(1.0) / F(18.0), w2 = F(1.0) / F(36.0);
  const PdfType ow0 = F(3.0) * w0 * omega, ow1 = F(3.0) * w1 * omega, ow2 = F(3.0) * w2 * omega;
  const PdfType indep = F(1.0) / F(3.0) * dens - F(0.5) * (ux * ux + uy * uy + uz * uz);
  /* direction d gets ow_d * (indep + u_d + 1.5 u_d^2), u_d = c_d . u */
  pdfs[D3Q19_C] = ow0 * indep;
  for (int d = 0; d < N_D3Q19 - 1; ++d) {
    const PdfType ud = (PdfType)D3Q19X[d] * ux + (PdfType)D3Q19Y[d] * uy + (PdfType)D3Q19Z[d] * uz;
    const int axis   = (D3Q19X[d] != 0) + (D3Q19Y[d] != 0) + (D3Q19Z[d] != 0) == 1;
    pdfs[d]          = (axis ? ow1 : ow2) * (indep + ud + F(1.5) * (ud * ud));
  }
}

void kernelSetInitialDensity(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd)
{
  const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
  PdfType pdfs[N_D3Q19];
  for (int z = 0; z < nZ; ++z)
    for (int y = 0; y < nY; ++y)
      for (int x = 0; x < nX; ++x) {
        if (ld->Lattice[latticeIndex(ld->Dims, x, y, z)] == LAT_CELL_OBSTACLE)
          continue;
        const PdfType dens = cd->RhoIn + (cd->RhoOut - cd->RhoIn) * x / (nX - F(1.0));
        equilibriumScaled(dens, F(0.0), F(0.0), F(0.0), cd->Omega, pdfs);
        kd->SetNode(kd, x, y, z, pdfs);
      }
}

void kernelSetInitialVelocity(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cd)
{
  const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
  PdfType pdfs[N_D3Q19];
  for (int z = 0; z < nZ; ++z)
    for (int y = 0; y < nY; ++y)
      for (int x = 0; x < nX; ++x) {
        if (ld->Lattice[latticeIndex(ld->Dims, x, y, z)] != LAT_CELL_FLUID)
          continue;
        equilibriumScaled(F(1.0), F(0.0), F(0.0), F(0.0), cd->Omega, pdfs);
        kd->SetNode(kd, x, y, z, pdfs);
      }
}

/* ------------------------------------------------------------------------ */
/* statistics and verification (src/lbm.c:953-1023)                           */
/* ------------------------------------------------------------------------ */
void kernelStatistics(KernelDataType *kd, LatticeDescType *ld, CaseDataType *cd, int iter)
{
  if (iter % cd->StatisticsModulus == 0 || iter == cd->MaxIterations)
    printf("# iter: %4d   avg density: %e\n", iter, (double)kernelDensity(kd, ld));
}

void kernelVerify(LatticeDescType *ld, KernelDataType *kd, CaseDataType *cdIn, PdfType *errorNorm)
{
  const CaseDataType eff = caseDataEffective(cdIn);
  const CaseDataType *cd = &eff;
  const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
  const PdfType visc = (F(1.0) / cd->Omega - F(0.5)) / F(3.0);
  PdfType *prof      = (PdfType *)calloc((size_t)nZ * nY, sizeof(PdfType));
  double *ux         = (double *)malloc(sizeof(double) * (size_t)nZ * nY);
  if (prof == NULL || ux == NULL) {
    fprintf(stderr, "Error: Insufficient memory to fulfill the request\n");
    exit(EXIT_FAILURE);
  }
  const int xPos = nX / 2;
  /* the plane's x momentum, summed on the device as src/lbm.c:981-983 */
  lbmPlaneVelocity(ld, kd, xPos, ux);
  for (int z = 0; z < nZ; ++z)
    for (int y = 0; y < nY; ++y) {
      if (ld->Lattice[latticeIndex(ld->Dims, xPos, y, z)] == LAT_CELL_OBSTACLE)
        continue;
      PdfType v = (PdfType)ux[(size_t)y * nZ + z];
      if (cd->BodyForce)
        v += F(0.5) * cd->XForce; /* half-step force correction */
      prof[y * nZ + z] = v;
    }
  free(ux);
  const PdfType centre = nY / F(2.0);
  const PdfType r2max  = (nY / F(2.0) - F(1.0)) * (nY / F(2.0) - F(1.0));
  const int even       = nY % 2 == 0;
  const int yc         = (nY - even - 1) / 2;
  PdfType dev          = F(0.0);
  for (int z = 0; z < nZ; ++z) {
    const PdfType r2 = (z - centre + F(0.5)) * (z - centre + F(0.5));
    const PdfType an = r2 >= r2max ? F(0.0) : cd->XForce * (r2max - r2) / (F(2.0) * visc);
    const PdfType avg = even ? (prof[yc * nZ + z] + prof[(yc + 1) * nZ + z]) / F(2.0) : prof[yc * nZ + z];
    if (an != F(0.0)) {
      const PdfType rel = (PdfType)fabs((double)(an - avg)) / an;
      dev += rel * rel;
    }
  }
  *errorNorm = (PdfType)sqrt((double)dev);
  printf("# Kernel validation: L2 error norm of relative error: %e\n", (double)*errorNorm);
  free(prof);
}
