/* Portions derived from LBM-Bench (src/cli.h), which is
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
/* cli.h -- command line of the GPU benchmark driver.
 *
 * Same flags and meaning as LBM-Bench's src/cli.h:10-23 (d g i o f k l V t
 * x y z h; -t is accepted by getopt and then rejected, as there), plus:
 *   -m V   TRT magic parameter Lambda (reference: fixed 1/12)
 *   -B     add the body force every step (reference: -DVERIFICATION builds)
 *   -I T   initial state: zero (reference benchmark start), rest, random
 *   -s N   seed of the random initial state and of spheres-N geometries
 *   -a V   amplitude of the random initial state (default 1e-3)
 *   -p T   PDF storage: native, fp32, fp16, bf16 (shifted deviation)
 *   -S V   storage scale (power of two; default per form)
 */
#ifndef LBM_AMD_CLI_H
#define LBM_AMD_CLI_H

#include <stdint.h>

#include "lbm.h"

#define BASE_ARGS "d:g:i:o:f:k:lVt:xyzhm:BI:s:a:p:S:D:"

#define HELPTEXT                                                                         \
  "Usage: lbmbench [options]\n"                                                          \
  "  -d XxYxZ             Geometry dimensions\n"                                         \
  "  -g TYPE              box|channel|pipe|blocks-N|fluid|spheres-N\n"                   \
  "  -i N                 Number of iterations\n"                                        \
  "  -o V                 Relaxation parameter (omega)\n"                                \
  "  -f V                 X-direction body force\n"                                      \
  "  -k NAME              Kernel to use\n"                                               \
  "  -l                   List available kernels\n"                                      \
  "  -V                   Run verification\n"                                            \
  "  -x/-y/-z             Enable periodic BC in x/y/z\n"                                 \
  "  -m V                 TRT magic parameter (default 1/12)\n"                          \
  "  -B                   Apply the body force every step\n"                             \
  "  -I TYPE              Initial state: zero|rest|random (default zero)\n"              \
  "  -s N                 Seed for random initial state and spheres\n"                   \
  "  -a V                 Amplitude of the random initial state\n"                       \
  "  -p TYPE              PDF storage: native|fp32|fp16|bf16\n"                          \
  "  -S V                 Storage scale (power of two)\n"                                \
  "  -D N                 GPU to run on (default: HIP's default device)\n"                \
  "  -h                   Print this help\n"

typedef struct {
  int initMode; /* LBM_INIT_* */
  uint64_t seed;
  double amplitude;
  int device; /* -D: HIP device index; -1: HIP's default device */
} RunOptions;

void parseArguments(int argc, char **argv, int dims[3], const char **geometryType,
    char **kernelToUse, int periodic[3], int *verify, CaseDataType *cd, RunOptions *ro);

#endif
