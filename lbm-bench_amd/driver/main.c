/* Portions derived from LBM-Bench (src/main.c), which is
 *   Copyright (C) NHR@FAU, University Erlangen-Nuremberg.
 *   Copyright (c) 2025 HPC Dwarfs.
 * and distributed under the MIT License:
 *
 * Permission is hereby granted, free of charge, to any person obtaining a copy
 * of this software and associated documentation files (the "Software"), to deal
 * in the Software without restriction, including without limitation the rights
 * to use, copy, modify, merge, publish, distribute, sublicense, and/or sell
 * copies of the Software, and to permit persons to whom the Software is
 * furnished to do so, subject to the following conditions:
 *
 * The above copyright notice and this permission notice shall be included in all
 * copies or substantial portions of the Software.
 *
 * THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
 * IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
 * FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE
 * AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
 * LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM,
 * OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE
 * SOFTWARE.
 */
/* main.c -- the benchmark driver: the flow and output lines of LBM-Bench's
 * src/main.c:34-189, with every kernel running on the GPU through the
 * lbm.h boundary (liblbmhip).  Extra lines: device, HBM roofline, mass. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "cli.h"
#include "lbm.h"
#include "lbm_hip.h"

#define HBM_PEAK_GBS 8000.0 /* MI355X HBM3E, published peak */

int main(int argc, char *argv[])
{
  int dims[3]              = { 20, 20, 20 };
  const char *geometryType = "channel";
  int verify               = 0;
  char *kernelToUse        = "push-soa";
  int periodic[3]          = { 0, 0, 0 };
  RunOptions ro            = { LBM_INIT_ZERO, 0, 1e-3, -1 };

  CaseDataType cd;
  cd.MaxIterations     = 10;
  cd.RhoIn             = F(1.0);
  cd.RhoOut            = F(1.0);
  cd.Omega             = F(1.0);
  cd.XForce            = F(0.00001);
  cd.StatisticsModulus = 100;
  caseDataDefaults(&cd);

  printf("LBM Benchmark Kernels, compiled %s %s, type: %s\n", __DATE__, __TIME__, "gpu (HIP)");

  parseArguments(argc, argv, dims, &geometryType, &kernelToUse, periodic, &verify, &cd, &ro);

  /* the kernel name is checked before anything runs, and the device is
   * selected before geometryCreate, which builds the flags on it */
  KernelFunctionsType *kf = NULL;
  for (int j = 0; j < G_N_KERNELS; ++j)
    if (!strcasecmp(kernelToUse, GKernels[j].Name)) {
      kf = &GKernels[j];
      break;
    }
  if (kf == NULL) {
    printf("ERROR: kernel \"%s\" not found.\n", kernelToUse);
    return 1;
  }
  if (ro.device >= 0 && lbmSelectDevice(ro.device) != 0)
    exit(EXIT_FAILURE);

  geometrySetSeed(ro.seed);
  LatticeDescType ld;
  geometryCreate(geometryType, dims, periodic, &ld);
  char dev[256] = "unknown";
  int cus       = 0;
  lbmDeviceInfo(dev, sizeof(dev), &cus, NULL);

  printf("#\n");
  printf("# - iterations:        %d\n", cd.MaxIterations);
  printf("# - geometry:\n");
  printf("#   type:              %s\n", ld.Name);
  printf("#   dimensions:        %d x %d x %d\n", ld.Dims[0], ld.Dims[1], ld.Dims[2]);
  printf("#   nodes total:       %10d\n", ld.nObst + ld.nFluid);
  printf("#   nodes fluid:       %10d\n", ld.nFluid);
  printf("#   nodes obstacles:   %10d\n", ld.nObst);
  printf("#   periodicity:       x: %d y: %d z: %d\n", ld.PeriodicX, ld.PeriodicY, ld.PeriodicZ);
  printf("# - flow:\n");
  printf("#   omega:             %f\n", (double)cd.Omega);
  printf("#   magic parameter:   %f\n", cd.MagicParam);
  printf("#   body force:        %s (%e)\n", cd.BodyForce ? "on" : "off", (double)cd.XForce);
  printf("#   rho in:            %e\n", (double)cd.RhoIn);
  printf("#   rho out:           %e\n", (double)cd.RhoOut);
  printf("# - device:            %s, %d CUs\n", dev, cus);

  printf("#\n# - kernel:            %s\n#\n", kf->Name);

  KernelDataType *kd = NULL;
  kf->Init(&ld, &kd, &cd);

  if (verify) {
    kernelSetInitialDensity(&ld, kd, &cd);
    kernelSetInitialVelocity(&ld, kd, &cd);
  } else if (ro.initMode != LBM_INIT_ZERO) {
    lbmInitPdfs(&ld, kd, ro.initMode, ro.seed, 1.0, ro.amplitude);
  }

  double m0[4] = { 0, 0, 0, 0 }, m1[4];
  lbmMassMomentum(&ld, kd, m0);

  printf("# starting kernel...\n");
  kd->Kernel(&ld, kd, &cd);

  kernelStatistics(kd, &ld, &cd, cd.MaxIterations);
  lbmMassMomentum(&ld, kd, m1);

  PdfType errorNorm = -1.0;
  if (verify)
    kernelVerify(&ld, kd, &cd, &errorNorm);

  const double duration    = kd->Duration;
  const double loopBalance = kd->LoopBalance;
  const double dataVolGByte =
      loopBalance * ld.nFluid * cd.MaxIterations / 1024.0 / 1024.0 / 1024.0;
  const double bandwidthGBytePerS = dataVolGByte / duration;
  const double perf      = (double)ld.nFluid * (double)cd.MaxIterations / duration / 1.0e6;
  const double hbmGBs    = perf * 1e6 * loopBalance / 1e9;

  kf->Deinit(&ld, &kd);

  printf("#\n");
  printf("# Evaluation Stats\n");
  printf("#   runtime:           \t%.3f s\n", duration);
  printf("#   iterations:        \t%d\n", cd.MaxIterations);
  printf("#   fluid cells:       \t%d\n", ld.nFluid);
  printf("# Derived metrics\n");
  printf("#   MEM data vol.:     \t%.2f GByte\n", dataVolGByte);
  printf("#   MEM bandwidth:     \t%.2f GByte/s\n", bandwidthGBytePerS);
  printf("#   performance:       \t%.3f MFLUP/s\n", perf);
  printf("#   HBM roofline:      \t%.1f GB/s of %.0f GB/s (%.1f %%), %.0f B/LUP\n", hbmGBs,
      HBM_PEAK_GBS, 100.0 * hbmGBs / HBM_PEAK_GBS, loopBalance);
  printf("#   mass drift:        \t%e (relative)\n", m0[0] != 0.0 ? (m1[0] - m0[0]) / m0[0] : 0.0);

  printf("P:   %f MFLUP/s  d: %f s  iter: %d  fnodes: %f x1e6  geo: %s  kernel: %s\n", perf,
      duration, cd.MaxIterations, ld.nFluid / 1e6, geometryType, kernelToUse);

  int exitCode = 0;
  /* an unstable run (e.g. omega far outside (0, 2)) ends with a non-finite
   * mass; the reference would print NaN and succeed */
  if (!isfinite(m1[0])) {
    fprintf(stderr, "ERROR: the mass is not finite after %d steps: the run is unstable "
                    "(omega %g)\n", cd.MaxIterations, (double)cd.Omega);
    exitCode = EXIT_FAILURE;
  }
  if (verify) {
    printf("# VERIFICATION: deviation from analytical solution: %e\n", (double)errorNorm);
    if (errorNorm > 0.1) {
      printf("# VERIFICATION FAILED.\n");
      exitCode = 1;
    } else {
      printf("# VERIFICATION SUCCEEDED.\n");
    }
  }
  geometryDestroy(&ld);
  return exitCode;
}
