/* Portions derived from LBM-Bench (src/cli.c), which is
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
/* cli.c -- getopt parsing for the GPU driver; the reference's flags keep
 * their meaning and error behaviour (src/cli.c:13-123). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>

#include "cli.h"
#include "lbm_hip.h"

static int parseDimensions(const char *param, int *nX, int *nY, int *nZ)
{
  const char *p = param;
  int *out[3]   = { nX, nY, nZ };
  static const char *axis[3] = { "X", "Y", "Z" };
  for (int a = 0; a < 3; ++a) {
    if (a > 0) {
      p = strchr(p, 'x');
      if (p == NULL) {
        printf("ERROR: %s dimension is missing.\n", axis[a]);
        return 0;
      }
      ++p;
    }
    *out[a] = atoi(p);
    if (*out[a] <= 0) {
      printf("ERROR: %s dimension must be > 0.\n", axis[a]);
      return 0;
    }
  }
  return 1;
}

void parseArguments(int argc, char **argv, int dims[3], const char **geometryType,
    char **kernelToUse, int periodic[3], int *verify, CaseDataType *cd, RunOptions *ro)
{
  int opt;
  while ((opt = getopt(argc, argv, BASE_ARGS)) != -1) {
    switch (opt) {
    case 'd':
      if (!parseDimensions(optarg, &dims[0], &dims[1], &dims[2]))
        exit(EXIT_FAILURE);
      break;
    case 'g':
      *geometryType = optarg;
      break;
    case 'i':
      cd->MaxIterations = (int)strtol(optarg, NULL, 0);
      if (cd->MaxIterations <= 0) {
        printf("ERROR: iterations must be > 0.\n");
        exit(EXIT_FAILURE);
      }
      break;
    case 'o':
      cd->Omega = F(strtod(optarg, NULL));
      break;
    case 'f':
      cd->XForce = F(strtod(optarg, NULL));
      break;
    case 'k':
      *kernelToUse = optarg;
      break;
    case 'l':
      printf("Available kernels:\n");
      for (int j = 0; j < G_N_KERNELS; ++j)
        printf("   %s\n", GKernels[j].Name);
      exit(EXIT_SUCCESS);
    case 'V':
      *verify           = 1;
      cd->Omega         = F(1.0);
      cd->RhoIn         = F(1.0);
      cd->RhoOut        = F(1.0);
      cd->BodyForce     = 1;
      *geometryType     = "box";
      dims[0]           = 16;
      dims[1]           = 16;
      dims[2]           = 16;
      cd->XForce        = F(0.00001);
      cd->MaxIterations = 1000;
      periodic[0]       = 1;
      periodic[1]       = 1;
      periodic[2]       = 0;
      printf("#\n# VERIFICATION: verifying flow profile of channel flow.\n#\n");
      break;
    case 'x':
      periodic[0] = 1;
      break;
    case 'y':
      periodic[1] = 1;
      break;
    case 'z':
      periodic[2] = 1;
      break;
    case 'm':
      cd->MagicParam = strtod(optarg, NULL);
      break;
    case 'B':
      cd->BodyForce = 1;
      break;
    case 'I':
      if (!strcasecmp(optarg, "zero"))
        ro->initMode = LBM_INIT_ZERO;
      else if (!strcasecmp(optarg, "rest"))
        ro->initMode = LBM_INIT_REST;
      else if (!strcasecmp(optarg, "random"))
        ro->initMode = LBM_INIT_RANDOM;
      else {
        printf("ERROR: unknown initial state '%s'.\n", optarg);
        exit(EXIT_FAILURE);
      }
      break;
    case 's':
      ro->seed = strtoull(optarg, NULL, 0);
      break;
    case 'a':
      ro->amplitude = strtod(optarg, NULL);
      break;
    case 'p':
      if (!strcasecmp(optarg, "native") || !strcasecmp(optarg, "fp64"))
        cd->Storage = LBM_STORE_NATIVE;
      else if (!strcasecmp(optarg, "fp32"))
        cd->Storage = LBM_STORE_FP32;
      else if (!strcasecmp(optarg, "fp16"))
        cd->Storage = LBM_STORE_FP16;
      else if (!strcasecmp(optarg, "bf16"))
        cd->Storage = LBM_STORE_BF16;
      else {
        printf("ERROR: unknown storage '%s'.\n", optarg);
        exit(EXIT_FAILURE);
      }
      break;
    case 'S':
      cd->StorageScale = strtod(optarg, NULL);
      break;
    case 'D': {
      /* a HIP device index: a non-negative decimal integer, nothing else */
      char *end = NULL;
      const long v = strtol(optarg, &end, 10);
      if (end == optarg || *end != '\0' || v < 0 || v > 1024) {
        printf("ERROR: device index must be a non-negative integer, got '%s'.\n", optarg);
        exit(EXIT_FAILURE);
      }
      ro->device = (int)v;
      break;
    }
    case 'h':
      printf(HELPTEXT);
      exit(EXIT_SUCCESS);
    default:
      printf("ERROR: unknown option.\n");
      printf(HELPTEXT);
      exit(EXIT_FAILURE);
    }
  }
}
