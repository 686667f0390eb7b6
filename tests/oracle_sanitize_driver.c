/* oracle_sanitize_driver.c -- runs every function of the CPU oracle
 * (oracle/lbm_oracle.c) on small lattices, for a build with AddressSanitizer
 * and UndefinedBehaviorSanitizer (tests/test_oracle_sanitizers.py).  Test
 * infrastructure: it checks the checker, not the GPU path.
 *
 * Exit status 0 when every call returns and the results are finite; any
 * out-of-bounds access, overflow or misaligned load aborts the program
 * through the sanitizers.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

#define Q 19

static int finite_all(const double *v, size_t n)
{
  for (size_t i = 0; i < n; ++i)
    if (!isfinite(v[i]))
      return 0;
  return 1;
}

static int run_case(const char *geometry, const int dims[3], const int per_in[3])
{
  static const char *kernels[] = { "pull-soa", "push-soa", "pull-aos", "push-aos",
    "blk-pull-soa", "blk-push-soa", "aa-soa" };
  const size_t ncells = (size_t)dims[0] * dims[1] * dims[2];
  int per[3]          = { per_in[0], per_in[1], per_in[2] };
  int counts[2]       = { 0, 0 };
  int *lat            = malloc(ncells * sizeof(int));
  double *init        = malloc(ncells * Q * sizeof(double));
  double *out         = malloc(ncells * Q * sizeof(double));
  if (!lat || !init || !out)
    return 1;
  if (oracle_geometry(geometry, dims, per, 7, lat, counts) != 0) {
    fprintf(stderr, "geometry %s rejected\n", geometry);
    return 1;
  }
  oracle_init_random(dims, lat, 3, 1e-3, 1.0, init);
  for (size_t k = 0; k < sizeof(kernels) / sizeof(kernels[0]); ++k) {
    for (int steps = 1; steps <= 3; ++steps) {
      int done = 0;
      if (oracle_run(kernels[k], dims, lat, per, init, steps, 1.25, 3.0 / 16.0, 1e-5, 1, out,
              &done) != 0) {
        fprintf(stderr, "%s rejected\n", kernels[k]);
        return 1;
      }
      if (!finite_all(out, ncells * Q)) {
        fprintf(stderr, "%s %s: non-finite PDFs\n", geometry, kernels[k]);
        return 1;
      }
    }
  }
  double m[4];
  oracle_moments(dims, lat, out, m);
  const double rho = oracle_density(dims, lat, out);
  const double err = oracle_verify(dims, lat, out, 1.25, 1e-5, 1);
  oracle_zou_he(dims, lat, out, 1.001, 0.999);
  if (!isfinite(m[0]) || !isfinite(rho) || !finite_all(out, ncells * Q)) {
    fprintf(stderr, "%s: non-finite moments\n", geometry);
    return 1;
  }
  (void)err; /* the verify norm of a noisy 3-step run carries no meaning */
  printf("%-10s %2dx%2dx%2d fluid %d: ok\n", geometry, dims[0], dims[1], dims[2], counts[0]);
  free(lat);
  free(init);
  free(out);
  return 0;
}

static int run_adjoint(void)
{
  const int dims[3]   = { 8, 6, 5 };
  const size_t ncells = (size_t)dims[0] * dims[1] * dims[2];
  int per[3]          = { 1, 0, 1 };
  int counts[2];
  int *lat      = malloc(ncells * sizeof(int));
  double *s     = malloc(ncells * sizeof(double));
  double *init  = malloc(ncells * Q * sizeof(double));
  double *state = malloc(ncells * Q * sizeof(double));
  double *grad  = malloc(ncells * sizeof(double));
  if (!lat || !s || !init || !state || !grad)
    return 1;
  oracle_geometry("channel", dims, per, 0, lat, counts);
  oracle_porous_field(dims, 11, 1.5, 4.0, -2.0, s);
  oracle_init_random(dims, lat, 5, 1e-3, 1.0, init);
  const double J = oracle_brinkman(dims, lat, s, init, 6, 1.25, 3.0 / 16.0, 1e-5, dims[0] - 1,
      state, grad);
  if (!isfinite(J) || !finite_all(grad, ncells) || !finite_all(state, ncells * Q)) {
    fprintf(stderr, "adjoint: non-finite results\n");
    return 1;
  }
  printf("adjoint    J %.6e: ok\n", J);
  free(lat);
  free(s);
  free(init);
  free(state);
  free(grad);
  return 0;
}

int main(void)
{
  struct {
    const char *geometry;
    int dims[3], per[3];
  } cases[] = {
    { "channel", { 6, 7, 5 }, { 0, 0, 1 } },
    { "fluid", { 4, 5, 6 }, { 0, 0, 0 } },
    { "box", { 5, 5, 5 }, { 0, 0, 0 } },
    { "blocks-2", { 8, 6, 7 }, { 0, 0, 1 } },
    { "spheres-3", { 9, 8, 10 }, { 0, 0, 1 } },
    { "pipe", { 6, 7, 7 }, { 0, 0, 0 } },
  };
  for (size_t i = 0; i < sizeof(cases) / sizeof(cases[0]); ++i)
    if (run_case(cases[i].geometry, cases[i].dims, cases[i].per) != 0)
      return 1;
  return run_adjoint();
}
