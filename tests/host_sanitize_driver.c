/* host_sanitize_driver.c -- runs the host C of the library
 * (lbm-bench_amd/csrc/lbm_host.c: geometryCreate's parsing and sphere
 * parameters, the -V initialisers, statistics and verify; the cell flags,
 * the verification profile and the Zou-He boundaries are computed on the
 * device, and host stand-ins below take the place of those entry points) under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_host_sanitizers.py).
 *
 * lbm_host.c reaches the PDFs only through KernelDataType's accessors, so
 * a mock KernelDataType over a host array stands in for the device
 * lattice here.  The registry entries that create device lattices are
 * stubs that abort: nothing here may reach them.  Host code only.
 *
 * Exit status 0 when every call returns and the checks hold; any
 * out-of-bounds access, overflow or misaligned load aborts the program
 * through the sanitizers.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm.h"
#include "lbm_hip.h"

#define Q 19

/* ---- stubs for the device side of the registry (never called) ---------- */
#define STUB_INIT(name)                                                                          \
  void name(LatticeDescType *ld, KernelDataType **kd, CaseDataType *cd)                          \
  {                                                                                              \
    (void)ld;                                                                                    \
    (void)kd;                                                                                    \
    (void)cd;                                                                                    \
    fprintf(stderr, #name " reached from host code\n");                                          \
    abort();                                                                                     \
  }
STUB_INIT(kernelInitPushSoA)
STUB_INIT(kernelInitPushAoS)
STUB_INIT(kernelInitPullSoA)
STUB_INIT(kernelInitPullAoS)
STUB_INIT(kernelInitBlkPushSoA)
STUB_INIT(kernelInitBlkPullSoA)
STUB_INIT(kernelInitAaSoA)

void kernelDeinit(LatticeDescType *ld, KernelDataType **kd)
{
  (void)ld;
  (void)kd;
  fprintf(stderr, "kernelDeinit reached from host code\n");
  abort();
}

/* ---- stand-ins for the device entry points lbm_host.c calls -------------- */
/* geometryCreate's flags: fluid inside, obstacles on the boundary layers of
 * the non-periodic axes (the device builder adds pipe, blocks and spheres;
 * here only their parameters are checked) */
long long lbmBuildFlags(const int dims[3], const int periodic[3], int type, int blockSize,
    int nSpheres, const double *spheres, const int *bounds, LatticeType *lattice)
{
  (void)type;
  (void)blockSize;
  long long fluid = 0;
  for (int k = 0; k < nSpheres; ++k)
    for (int a = 0; a < 3; ++a)
      if (bounds[6 * k + a] > bounds[6 * k + 3 + a] || !(spheres[4 * k + 3] >= 4.0)) {
        fprintf(stderr, "sphere %d: bad parameters\n", k);
        abort();
      }
  for (int z = 0; z < dims[2]; ++z)
    for (int y = 0; y < dims[1]; ++y)
      for (int x = 0; x < dims[0]; ++x) {
        const int face = (!periodic[0] && (x == 0 || x == dims[0] - 1)) ||
                         (!periodic[1] && (y == 0 || y == dims[1] - 1)) ||
                         (!periodic[2] && (z == 0 || z == dims[2] - 1));
        lattice[latticeIndex((int *)dims, x, y, z)] = face ? LAT_CELL_OBSTACLE : LAT_CELL_FLUID;
        fluid += !face;
      }
  return fluid;
}

static void getNode(KernelDataType *kd, int x, int y, int z, PdfType *p);

/* the plane's x momentum through the mock accessors */
void lbmPlaneVelocity(LatticeDescType *ld, KernelDataType *kd, int x, double *ux)
{
  PdfType p[Q];
  for (int y = 0; y < ld->Dims[1]; ++y)
    for (int z = 0; z < ld->Dims[2]; ++z) {
      double v = 0.0;
      if (ld->Lattice[latticeIndex(ld->Dims, x, y, z)] != LAT_CELL_OBSTACLE) {
        getNode(kd, x, y, z, p);
        v = p[D3Q19_E] + p[D3Q19_NE] + p[D3Q19_SE] + p[D3Q19_TE] + p[D3Q19_BE] - p[D3Q19_W] -
            p[D3Q19_NW] - p[D3Q19_SW] - p[D3Q19_TW] - p[D3Q19_BW];
      }
      ux[(size_t)y * ld->Dims[2] + z] = v;
    }
}

/* ---- the mock lattice: logical PDFs [cell][19], cell = latticeIndex ------ */
static LatticeDescType *gLd;

static PdfType *slot(KernelDataType *kd, int x, int y, int z, int d)
{
  if (x < 0 || y < 0 || z < 0 || x >= kd->Dims[0] || y >= kd->Dims[1] || z >= kd->Dims[2] ||
      d < 0 || d >= Q) {
    fprintf(stderr, "accessor out of range: %d %d %d %d\n", x, y, z, d);
    abort();
  }
  return &kd->PdfsActive[(size_t)latticeIndex(kd->Dims, x, y, z) * Q + (size_t)d];
}
static void getNode(KernelDataType *kd, int x, int y, int z, PdfType *p)
{
  memcpy(p, slot(kd, x, y, z, 0), Q * sizeof(PdfType));
}
static void setNode(KernelDataType *kd, int x, int y, int z, PdfType *p)
{
  memcpy(slot(kd, x, y, z, 0), p, Q * sizeof(PdfType));
}
static void bcGet(KernelDataType *kd, int x, int y, int z, int d, PdfType *p) { *p = *slot(kd, x, y, z, d); }
static void bcSet(KernelDataType *kd, int x, int y, int z, int d, PdfType p) { *slot(kd, x, y, z, d) = p; }

PdfType kernelDensity(KernelDataType *kd, LatticeDescType *ld)
{
  double sum = 0.0;
  PdfType p[Q];
  for (int z = 0; z < ld->Dims[2]; ++z)
    for (int y = 0; y < ld->Dims[1]; ++y)
      for (int x = 0; x < ld->Dims[0]; ++x) {
        if (ld->Lattice[latticeIndex(ld->Dims, x, y, z)] == LAT_CELL_OBSTACLE)
          continue;
        getNode(kd, x, y, z, p);
        for (int d = 0; d < Q; ++d)
          sum += (double)p[d];
      }
  return (PdfType)(sum / (ld->nFluid > 0 ? ld->nFluid : 1));
}

static int finiteAll(const PdfType *v, size_t n)
{
  for (size_t i = 0; i < n; ++i)
    if (!isfinite((double)v[i]))
      return 0;
  return 1;
}

static int runCase(const char *geometry, int nx, int ny, int nz, int px, int py, int pz)
{
  int dims[3] = { nx, ny, nz }, per[3] = { px, py, pz };
  LatticeDescType ld;
  memset(&ld, 0, sizeof(ld));
  geometrySetSeed(5);
  geometryCreate(geometry, dims, per, &ld);
  gLd = &ld;
  if (ld.nCells != nx * ny * nz || ld.nFluid + ld.nObst != ld.nCells) {
    fprintf(stderr, "%s: cell counts %d %d %d\n", geometry, ld.nCells, ld.nFluid, ld.nObst);
    return 1;
  }
  for (int i = 0; i < ld.nCells; ++i)
    if (ld.Lattice[i] != LAT_CELL_OBSTACLE && ld.Lattice[i] != LAT_CELL_FLUID) {
      fprintf(stderr, "%s: flag %d\n", geometry, (int)ld.Lattice[i]);
      return 1;
    }

  KernelDataType kd;
  memset(&kd, 0, sizeof(kd));
  memcpy(kd.Dims, ld.Dims, sizeof(kd.Dims));
  memcpy(kd.GlobalDims, ld.Dims, sizeof(kd.GlobalDims));
  kd.PdfsActive = calloc((size_t)ld.nCells * Q, sizeof(PdfType));
  kd.Pdfs[0]    = kd.PdfsActive;
  if (!kd.PdfsActive)
    return 1;
  kd.GetNode                  = getNode;
  kd.SetNode                  = setNode;
  kd.BoundaryConditionsGetPdf = bcGet;
  kd.BoundaryConditionsSetPdf = bcSet;

  CaseDataType cd;
  memset(&cd, 0, sizeof(cd));
  cd.Omega             = (PdfType)1.25;
  cd.RhoIn             = (PdfType)1.01;
  cd.RhoOut            = (PdfType)0.99;
  cd.XForce            = (PdfType)1e-5;
  cd.MaxIterations     = 10;
  cd.StatisticsModulus = 5;
  caseDataDefaults(&cd);
  cd.BodyForce = 1;

  kernelSetInitialDensity(&ld, &kd, &cd);
  kernelSetInitialVelocity(&ld, &kd, &cd);
  kernelStatistics(&kd, &ld, &cd, 10);
  PdfType norm = 0;
  kernelVerify(&ld, &kd, &cd, &norm);
  if (!finiteAll(kd.PdfsActive, (size_t)ld.nCells * Q) || !isfinite((double)norm)) {
    fprintf(stderr, "%s: non-finite PDFs or norm\n", geometry);
    return 1;
  }
  printf("%-10s %2dx%2dx%2d fluid %5d: ok\n", geometry, nx, ny, nz, ld.nFluid);
  free(kd.PdfsActive);
  geometryDestroy(&ld);
  return 0;
}

int main(void)
{
  CaseDataType cd;
  memset(&cd, 0, sizeof(cd));
  cd.Omega = (PdfType)1.25;
  caseDataDefaults(&cd);
  const CaseDataType eff = caseDataEffective(&cd);
  if (eff.Extensions != LBM_CASE_EXTENSIONS)
    return 1;

  struct {
    const char *geometry;
    int nx, ny, nz, px, py, pz;
  } cases[] = {
    { "channel", 9, 8, 10, 0, 0, 1 },
    { "channel", 16, 11, 7, 1, 0, 1 },
    { "box", 7, 7, 7, 0, 0, 0 },
    { "pipe", 9, 11, 11, 0, 0, 0 },
    { "blocks-3", 12, 10, 9, 0, 0, 1 },
    { "blocks", 17, 17, 17, 0, 0, 1 },
    { "fluid", 5, 6, 4, 1, 1, 1 },
    { "fluid", 1, 2, 5, 1, 1, 1 },
    { "spheres-4", 20, 18, 22, 0, 0, 1 },
    { "spheres-1", 3, 3, 3, 1, 1, 1 },
  };
  for (size_t i = 0; i < sizeof(cases) / sizeof(cases[0]); ++i)
    if (runCase(cases[i].geometry, cases[i].nx, cases[i].ny, cases[i].nz, cases[i].px, cases[i].py,
            cases[i].pz) != 0)
      return 1;
  return 0;
}
