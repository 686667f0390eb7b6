/* lbm_oracle.h -- CPU restatement of LBM-Bench's D3Q19 TRT path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the GPU path is compared
 * with.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (lbm-bench_amd/) never links it.
 *
 * Logical PDF arrays are float64 [x][y][z][19]: element
 * ((x*nY + y)*nZ + z)*19 + d holds f_d of cell (x,y,z).
 * Lattice flag arrays are int [z][y][x] (x fastest), LAT_CELL_* values,
 * exactly like LBM-Bench's LatticeDescType.Lattice (src/lbm.h:108-111).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

/* geometry (src/lbm.c:52-187, plus "spheres-N"); returns 0 or -1 */
int oracle_geometry(const char *type, const int dims[3], int periodic_io[3],
    uint64_t seed, int *lattice_out, int counts_out[2]);

/* one run of a variant; see lbm_oracle.c for the argument contract */
int oracle_run(const char *kernel, const int dims[3], const int *lattice,
    const int periodic[3], const double *init, int iterations, double omega,
    double magic, double xforce, int bodyforce, double *out, int *steps_done);

/* oracle_run with kernelComputeBoundaryConditions (src/lbm.c:676-785)
 * applied after every step, after its bounce-back and body force */
int oracle_run_bc(const char *kernel, const int dims[3], const int *lattice,
    const int periodic[3], const double *init, int iterations, double omega,
    double magic, double xforce, int bodyforce, double rhoIn, double rhoOut,
    double *out, int *steps_done);

/* f_i = w_i (1 + eps), eps ~ U(-amp, amp) keyed on (seed, cell, i) */
void oracle_init_random(const int dims[3], const int *lattice, uint64_t seed,
    double amp, double rho, double *out);

/* mean density over non-obstacle cells (src/lbm.c:928-947) */
double oracle_density(const int dims[3], const int *lattice, const double *pdfs);

/* src/lbm.c:964-1023 (wall-normal z, averaged over the central y rows) */
double oracle_verify(const int dims[3], const int *lattice, const double *pdfs,
    double omega, double xforce, int force_correction);

/* sums over non-obstacle cells: out = {rho, jx, jy, jz} */
void oracle_moments(const int dims[3], const int *lattice, const double *pdfs,
    double out[4]);

/* differentiation form: Brinkman TRT forward, J = sum over x = outletX of
 * 1/2 rho |j|^2, and dJ/ds by the discrete adjoint (grad may be NULL) */
double oracle_brinkman(const int dims[3], const int *lattice, const double *s,
    const double *init, int T, double omega, double magic, double xforce, int outletX,
    double *final_state, double *grad);

/* porous solid-fraction field (smoothed, standardised, logistic) */
/* Zou-He pressure inlet/outlet on logical PDFs, in place (src/lbm.c:676-785) */
void oracle_zou_he(const int dims[3], const int *lattice, double *pdfs, double rhoIn,
    double rhoOut);

void oracle_porous_field(const int dims[3], uint64_t seed, double sigma, double a, double b,
    double *s);

/* the hash used for every seeded input */
uint64_t oracle_hash(uint64_t seed, uint64_t a, uint64_t b);

#endif
