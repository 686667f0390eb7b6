/* ref_harness.c -- flat entry points around LBM-Bench's own, unmodified
 * sources (src/lbm.c, src/kernel_*.c, src/allocate.c, src/timing.c), so the
 * oracle restatement can be pinned against the compiled reference.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/
 * from the reference tree, with -DVERIFICATION (the body force is compiled
 * in, src/kernel_pull_soa.c:206-209).  Nothing in the product links this.
 *
 * Logical PDF arrays are float64 [x][y][z][19] (z fastest inside a cell
 * block), read and written through the variant's own GetNode/SetNode
 * (src/lbm.c:193-294), so every layout is compared on the same footing.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "lbm.h"

static KernelFunctionsType *findKernel(const char *name)
{
  for (int j = 0; j < G_N_KERNELS; ++j)
    if (!strcasecmp(name, GKernels[j].Name))
      return &GKernels[j];
  return NULL;
}

/* Geometry as the reference builds it (src/lbm.c:52-187).  lattice_out
 * receives nX*nY*nZ flags in latticeIndex order (x fastest);
 * periodic_io is updated with the forced periodicity; counts_out gets
 * {nFluid, nObst}. */
int ref_geometry(const char *type, const int dims_in[3], int periodic_io[3],
    int *lattice_out, int counts_out[2])
{
  int dims[3] = { dims_in[0], dims_in[1], dims_in[2] };
  LatticeDescType ld;
  geometryCreate(type, dims, periodic_io, &ld);
  memcpy(lattice_out, ld.Lattice, sizeof(int) * (size_t)ld.nCells);
  counts_out[0] = ld.nFluid;
  counts_out[1] = ld.nObst;
  free(ld.Lattice);
  return 0;
}

/* One run of a reference kernel.
 *   init      logical PDFs to SetNode on every non-obstacle cell before the
 *             run (Iteration == -1, as main.c:112-117), or NULL for the
 *             reference's all-zero start (src/lbm.c:452-463).
 *   ref_init  1: use kernelSetInitialDensity/Velocity (main.c:113-116).
 *   out       logical PDFs after the run (GetNode on non-obstacle cells;
 *             obstacle cells are written as 0).
 *   stats     {kernelDensity, kernelVerify error norm, Duration}.
 * Returns 0, or -1 for an unknown kernel name. */
static int runCase(const char *kernel, const char *geometry, const int dims_in[3],
    const int periodic_in[3], const double *init, int ref_init, int iterations,
    double omega, double xforce, int zouHe, double rhoIn, double rhoOut, double *out,
    double stats[3])
{
  int dims[3]     = { dims_in[0], dims_in[1], dims_in[2] };
  int periodic[3] = { periodic_in[0], periodic_in[1], periodic_in[2] };

  KernelFunctionsType *kf = findKernel(kernel);
  if (kf == NULL)
    return -1;

  CaseDataType cd;
  cd.MaxIterations     = iterations;
  cd.RhoIn             = (PdfType)rhoIn;
  cd.RhoOut            = (PdfType)rhoOut;
  cd.Omega             = (PdfType)omega;
  cd.XForce            = (PdfType)xforce;
  cd.StatisticsModulus = 100;

  LatticeDescType ld;
  geometryCreate(geometry, dims, periodic, &ld);

  KernelDataType *kd = NULL;
  kf->Init(&ld, &kd, &cd);

  int nX = dims[0], nY = dims[1], nZ = dims[2];
  PdfType pdfs[N_D3Q19];

  if (ref_init) {
    kernelSetInitialDensity(&ld, kd, &cd);
    kernelSetInitialVelocity(&ld, kd, &cd);
  } else if (init != NULL) {
    for (int x = 0; x < nX; ++x)
      for (int y = 0; y < nY; ++y)
        for (int z = 0; z < nZ; ++z) {
          if (ld.Lattice[latticeIndex(ld.Dims, x, y, z)] == LAT_CELL_OBSTACLE)
            continue;
          const double *src = init + (((size_t)x * nY + y) * nZ + z) * N_D3Q19;
          for (int d = 0; d < N_D3Q19; ++d)
            pdfs[d] = (PdfType)src[d];
          kd->SetNode(kd, x, y, z, pdfs);
        }
  }

  kd->Kernel(&ld, kd, &cd);

  if (zouHe) {
    /* the non-obstacle cells of the planes x = 0 and x = nX-1 become inlet
     * and outlet cells, then the reference's boundary pass runs once */
    for (int z = 1; z < dims[2] - 1; ++z)
      for (int y = 1; y < dims[1] - 1; ++y) {
        int *in  = &ld.Lattice[latticeIndex(ld.Dims, 0, y, z)];
        int *out_ = &ld.Lattice[latticeIndex(ld.Dims, dims[0] - 1, y, z)];
        if (*in != LAT_CELL_OBSTACLE)
          *in = LAT_CELL_INLET;
        if (*out_ != LAT_CELL_OBSTACLE)
          *out_ = LAT_CELL_OUTLET;
      }
    kernelComputeBoundaryConditions(kd, &ld, &cd);
  }

  for (int x = 0; x < nX; ++x)
    for (int y = 0; y < nY; ++y)
      for (int z = 0; z < nZ; ++z) {
        double *dst = out + (((size_t)x * nY + y) * nZ + z) * N_D3Q19;
        if (ld.Lattice[latticeIndex(ld.Dims, x, y, z)] == LAT_CELL_OBSTACLE) {
          for (int d = 0; d < N_D3Q19; ++d)
            dst[d] = 0.0;
          continue;
        }
        kd->GetNode(kd, x, y, z, pdfs);
        for (int d = 0; d < N_D3Q19; ++d)
          dst[d] = (double)pdfs[d];
      }

  PdfType norm = -1.0;
  stats[0]     = (double)kernelDensity(kd, &ld);
  kernelVerify(&ld, kd, &cd, &norm);
  stats[1] = (double)norm;
  stats[2] = kd->Duration;

  kf->Deinit(&ld, &kd);
  free(ld.Lattice);
  return 0;
}

int ref_run(const char *kernel, const char *geometry, const int dims_in[3],
    const int periodic_in[3], const double *init, int ref_init, int iterations,
    double omega, double xforce, double *out, double stats[3])
{
  return runCase(kernel, geometry, dims_in, periodic_in, init, ref_init, iterations, omega,
      xforce, 0, 1.0, 1.0, out, stats);
}

/* As ref_run without ref_init, then the non-obstacle cells of x = 0 and
 * x = nX-1 (y, z in 1 .. n-2) are marked inlet and outlet and the
 * reference's kernelComputeBoundaryConditions (src/lbm.c:676-785) runs once
 * with rhoIn / rhoOut before the logical PDFs are read back. */
int ref_zou_he(const char *kernel, const char *geometry, const int dims_in[3],
    const int periodic_in[3], const double *init, int iterations, double omega, double xforce,
    double rhoIn, double rhoOut, double *out)
{
  double stats[3];
  return runCase(kernel, geometry, dims_in, periodic_in, init, 0, iterations, omega, xforce, 1,
      rhoIn, rhoOut, out, stats);
}

/* The reference's functions composed into a time loop with pressure
 * boundaries: the non-obstacle cells of x = 0 and x = nX-1 (y, z in
 * 1 .. n-2) are marked inlet and outlet, init is set through SetNode, and
 * then `iterations` times: the variant's Kernel for one step (its
 * bounce-back copies and, under -DVERIFICATION, its body force) followed by
 * kernelComputeBoundaryConditions (src/lbm.c:676-785).  The two-lattice
 * kernels start every call from Pdfs[0] (src/kernel_pull_soa.c:35-36), so
 * the harness swaps Pdfs[0] and Pdfs[1] after each one-step call.  aa-soa
 * runs its steps in pairs (src/kernel_aa_soa.c:36) and has no one-step
 * call: it returns -2.  Returns 0, or -1 for an unknown kernel. */
int ref_run_bc(const char *kernel, const char *geometry, const int dims_in[3],
    const int periodic_in[3], const double *init, int iterations, double omega, double xforce,
    double rhoIn, double rhoOut, double *out)
{
  int dims[3]     = { dims_in[0], dims_in[1], dims_in[2] };
  int periodic[3] = { periodic_in[0], periodic_in[1], periodic_in[2] };
  KernelFunctionsType *kf = findKernel(kernel);
  if (kf == NULL)
    return -1;
  if (!strcasecmp(kernel, "aa-soa"))
    return -2;
  CaseDataType cd;
  cd.MaxIterations     = 1;
  cd.RhoIn             = (PdfType)rhoIn;
  cd.RhoOut            = (PdfType)rhoOut;
  cd.Omega             = (PdfType)omega;
  cd.XForce            = (PdfType)xforce;
  cd.StatisticsModulus = 100;
  LatticeDescType ld;
  geometryCreate(geometry, dims, periodic, &ld);
  for (int z = 1; z < dims[2] - 1; ++z)
    for (int y = 1; y < dims[1] - 1; ++y) {
      int *in   = &ld.Lattice[latticeIndex(ld.Dims, 0, y, z)];
      int *out_ = &ld.Lattice[latticeIndex(ld.Dims, dims[0] - 1, y, z)];
      if (*in != LAT_CELL_OBSTACLE)
        *in = LAT_CELL_INLET;
      if (*out_ != LAT_CELL_OBSTACLE)
        *out_ = LAT_CELL_OUTLET;
    }
  KernelDataType *kd = NULL;
  kf->Init(&ld, &kd, &cd);
  const int nX = dims[0], nY = dims[1], nZ = dims[2];
  PdfType pdfs[N_D3Q19];
  for (int x = 0; x < nX; ++x)
    for (int y = 0; y < nY; ++y)
      for (int z = 0; z < nZ; ++z) {
        if (ld.Lattice[latticeIndex(ld.Dims, x, y, z)] == LAT_CELL_OBSTACLE)
          continue;
        const double *src = init + (((size_t)x * nY + y) * nZ + z) * N_D3Q19;
        for (int d = 0; d < N_D3Q19; ++d)
          pdfs[d] = (PdfType)src[d];
        kd->SetNode(kd, x, y, z, pdfs);
      }
  for (int it = 0; it < iterations; ++it) {
    kd->Kernel(&ld, kd, &cd);
    kernelComputeBoundaryConditions(kd, &ld, &cd);
    PdfType *t  = kd->Pdfs[0];
    kd->Pdfs[0] = kd->Pdfs[1];
    kd->Pdfs[1] = t;
  }
  for (int x = 0; x < nX; ++x)
    for (int y = 0; y < nY; ++y)
      for (int z = 0; z < nZ; ++z) {
        double *dst = out + (((size_t)x * nY + y) * nZ + z) * N_D3Q19;
        if (ld.Lattice[latticeIndex(ld.Dims, x, y, z)] == LAT_CELL_OBSTACLE) {
          for (int d = 0; d < N_D3Q19; ++d)
            dst[d] = 0.0;
          continue;
        }
        kd->GetNode(kd, x, y, z, pdfs);
        for (int d = 0; d < N_D3Q19; ++d)
          dst[d] = (double)pdfs[d];
      }
  kf->Deinit(&ld, &kd);
  free(ld.Lattice);
  return 0;
}

/* The reference's -V run (cli.c:84-99 + main.c:112-142): box 16^3,
 * periodic x,y, omega 1, g 1e-5, 1000 iterations, kernelSetInitial*.
 * Returns the L2 error norm printed as "VERIFICATION: deviation". */
double ref_verification_norm(const char *kernel)
{
  int dims[3]     = { 16, 16, 16 };
  int periodic[3] = { 1, 1, 0 };
  size_t n        = (size_t)16 * 16 * 16 * N_D3Q19;
  double *out     = (double *)malloc(n * sizeof(double));
  double stats[3];
  int rc = ref_run(kernel, "box", dims, periodic, NULL, 1, 1000, 1.0, 0.00001, out, stats);
  free(out);
  return rc == 0 ? stats[1] : -1.0;
}
