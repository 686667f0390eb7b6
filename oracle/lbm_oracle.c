/* lbm_oracle.c -- CPU restatement of LBM-Bench's D3Q19 TRT
 * stream-and-collide path with half-way bounce-back.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke
 * check and bench.py's cpu_baseline leg load it as liboracle.so to check
 * (never to replace) the GPU path.  The product never links it.
 *
 * It follows the reference's own mechanism, not the GPU path's: a lattice
 * with one halo layer, explicit bounce-back/periodic copy lists and the
 * per-variant node accessors.  The GPU path folds all of that into a
 * masked gather/scatter, so agreement between the two is a real check.
 *
 * Differences from the reference, each a parameter the reference fixes:
 *   - Lambda (magic) is an argument; the reference hard-codes 1/12
 *     (src/kernel_pull_soa.c:20).  magic = 1/12 reproduces it.
 *   - The body force is an argument; the reference compiles it in with
 *     -DVERIFICATION (src/kernel_pull_soa.c:206-209).
 *   - Indices are 64-bit (the reference's int overflows beyond 480^3,
 *     src/lbm.h:113-121).
 *
 * Parity status: the collision, streaming, bounce-back, force, AA parity,
 * density and verify functions are pinned against the compiled reference
 * (oracle/_ref, tests/test_oracle_vs_reference.py) and against golden
 * fixtures generated from it (tests/golden/).  The Lambda dependence
 * (magic != 1/12) is pinned by closed forms: Lambda = (1/omega - 1/2)^2
 * reduces TRT to BGK, and Lambda = 3/16 makes steady Poiseuille flow exact
 * (tests/test_oracle_closed_forms.py).  The "spheres-N" geometry and the
 * seeded initialiser have no reference counterpart; they are specified
 * here and pinned by brute-force restatements in the tests.  The Zou-He
 * boundaries are pinned against the reference's own
 * kernelComputeBoundaryConditions (oracle/_ref) after runs of every
 * variant.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

/* direction ids and velocity set (src/lbm.h:20-38, src/lbm.c:20-29) */
enum { N = 0, S, E, W, NE, SE, NW, SW, T, TN, TE, TW, TS, B, BS, BN, BW, BE, C, Q };
static const int CX[Q]  = { 0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0 };
static const int CY[Q]  = { 1, -1, 0, 0, 1, -1, 1, -1, 0, 1, 0, 0, -1, 0, -1, 1, 0, 0, 0 };
static const int CZ[Q]  = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, 0 };
static const int INV[Q] = { S, N, W, E, SW, NW, SE, NE, B, BS, BW, BE, BN, T, TN, TS, TE, TW, C };

enum { OBST = 0, FLUID = 1 };
enum { P_PUSH = 0, P_PULL = 1, P_AA = 2 };
enum { L_SOA = 0, L_AOS = 1 };

/* --------------------------------------------------------------------- */
/* seeded inputs: splitmix64 (Steele, Lea & Flood 2014) finalizer chained */
/* over (seed, a, b); u01 takes the top 53 bits.                          */
/* --------------------------------------------------------------------- */
static uint64_t mix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t oracle_hash(uint64_t seed, uint64_t a, uint64_t b)
{
  return mix64(mix64(mix64(seed) ^ a) ^ b);
}

static double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

/* --------------------------------------------------------------------- */
/* geometry: src/lbm.c:52-187                                             */
/* --------------------------------------------------------------------- */
static inline int64_t lat_index(const int n[3], int x, int y, int z)
{
  return ((int64_t)z * n[1] + y) * n[0] + x;
}

#define SPHERE_SALT 0x5350484552455321ull

int oracle_geometry(const char *type, const int dims[3], int per[3], uint64_t seed,
    int *lat, int counts[2])
{
  enum { G_BOX, G_CHANNEL, G_PIPE, G_BLOCKS, G_FLUID, G_SPHERES } g;
  int blockSize = 8, nSpheres = 0;
  if (strncasecmp("channel", type, 7) == 0)
    g = G_CHANNEL;
  else if (strncasecmp("box", type, 3) == 0)
    g = G_BOX;
  else if (strncasecmp("pipe", type, 4) == 0)
    g = G_PIPE;
  else if (strncasecmp("blocks", type, 6) == 0) {
    g = G_BLOCKS;
    if (strlen(type) > 7) {
      blockSize = atoi(type + 7);
      if (blockSize <= 0)
        blockSize = 8;
    }
  } else if (strncasecmp("fluid", type, 5) == 0)
    g = G_FLUID;
  else if (strncasecmp("spheres", type, 7) == 0) {
    g        = G_SPHERES;
    nSpheres = strlen(type) > 8 ? atoi(type + 8) : 0;
  } else
    return -1;

  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int64_t ncells = (int64_t)nx * ny * nz;
  for (int64_t i = 0; i < ncells; ++i)
    lat[i] = FLUID;

  /* forced periodicity (src/lbm.c:96-102) */
  if (g == G_CHANNEL || g == G_BLOCKS || g == G_PIPE || g == G_SPHERES)
    per[0] = 1;
  else if (g == G_FLUID)
    per[0] = per[1] = per[2] = 1;

  /* boundary layers of non-periodic axes are obstacles (src/lbm.c:104-129) */
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        int onFace = (!per[0] && (x == 0 || x == nx - 1)) ||
                     (!per[1] && (y == 0 || y == ny - 1)) ||
                     (!per[2] && (z == 0 || z == nz - 1));
        if (onFace)
          lat[lat_index(dims, x, y, z)] = OBST;
      }

  if (g == G_PIPE) { /* src/lbm.c:131-142 */
    double cz = nz / 2.0 - 0.5, cy = ny / 2.0 - 0.5;
    double dmin = ny < nz ? ny : nz;
    double r2   = (dmin / 2 - 1) * (dmin / 2 - 1);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        if ((z - cz) * (z - cz) + (y - cy) * (y - cy) >= r2)
          for (int x = 0; x < nx; ++x)
            lat[lat_index(dims, x, y, z)] = OBST;
  } else if (g == G_BLOCKS) { /* src/lbm.c:143-156 */
    int nb = blockSize;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
          if ((z % (2 * nb)) >= nb && (y % (2 * nb)) >= nb && (x % (2 * nb)) >= nb)
            lat[lat_index(dims, x, y, z)] = OBST;
  } else if (g == G_SPHERES) {
    /* sphere k: centre U[0,n) per axis, radius U[4,12]; a cell is solid when
     * its integer position lies within r of the centre, the bounding box
     * wrapping around periodic axes and clipped on the others. */
    const int ncoord[3] = { nx, ny, nz };
    for (int k = 0; k < nSpheres; ++k) {
      double c[3], r;
      for (int a = 0; a < 3; ++a)
        c[a] = u01(oracle_hash(seed ^ SPHERE_SALT, (uint64_t)k, (uint64_t)a)) * ncoord[a];
      r = 4.0 + 8.0 * u01(oracle_hash(seed ^ SPHERE_SALT, (uint64_t)k, 3));
      int lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        lo[a] = (int)floor(c[a] - r);
        hi[a] = (int)ceil(c[a] + r);
      }
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) {
            double dx = x - c[0], dy = y - c[1], dz = z - c[2];
            if (dx * dx + dy * dy + dz * dz > r * r)
              continue;
            int p[3] = { x, y, z }, ok = 1;
            for (int a = 0; a < 3; ++a) {
              if (p[a] < 0 || p[a] >= ncoord[a]) {
                if (!per[a]) {
                  ok = 0;
                  break;
                }
                p[a] = ((p[a] % ncoord[a]) + ncoord[a]) % ncoord[a];
              }
            }
            if (ok)
              lat[lat_index(dims, p[0], p[1], p[2])] = OBST;
          }
    }
  }

  int64_t nf = 0, no = 0;
  for (int64_t i = 0; i < ncells; ++i) {
    if (lat[i] == OBST)
      ++no;
    else
      ++nf;
  }
  counts[0] = (int)nf;
  counts[1] = (int)no;
  return 0;
}

/* --------------------------------------------------------------------- */
/* TRT collision, src/kernel_pull_soa.c:18-28 (constants) and :118-198    */
/* (update).  f holds the 19 pre-collision values, o receives the 19      */
/* post-collision values, both indexed by direction.                      */
/* --------------------------------------------------------------------- */
typedef struct {
  double we, wo, w0, w1, w2, w1x3, w2x3, w1nh, w2nh;
} trt_t;

static void trt_setup(trt_t *p, double omega, double magic)
{
  p->we   = omega;
  p->wo   = 1.0 / (0.5 + magic / (1.0 / omega - 0.5));
  p->w0   = 1.0 / 3.0;
  p->w1   = 1.0 / 18.0;
  p->w2   = 1.0 / 36.0;
  p->w1x3 = p->w1 * 3.0;
  p->w1nh = p->w1 * 9.0 / 2.0;
  p->w2x3 = p->w2 * 3.0;
  p->w2nh = p->w2 * 9.0 / 2.0;
}

/* summation orders of the moments as the reference writes them */
static const int RHO_ORDER[Q] = { C, N, E, S, W, NE, SE, SW, NW, T, TN, TE, TS, TW, B, BN, BE, BS, BW };
static const int JX_PLUS[5] = { E, NE, SE, TE, BE }, JX_MINUS[5] = { W, NW, SW, TW, BW };
static const int JY_PLUS[5] = { N, NE, NW, TN, BN }, JY_MINUS[5] = { S, SE, SW, TS, BS };
static const int JZ_PLUS[5] = { T, TE, TW, TN, TS }, JZ_MINUS[5] = { B, BE, BW, BN, BS };

/* the nine (i, inv i) pairs in the reference's order; u_i = c_i . j */
static const int PAIR[9][2] = { { N, S }, { E, W }, { T, B }, { NW, SE }, { NE, SW },
  { TW, BE }, { TE, BW }, { TS, BN }, { TN, BS } };

static inline double sum5(const double *f, const int *plus, const int *minus)
{
  double s = f[plus[0]];
#pragma GCC unroll 5
  for (int k = 1; k < 5; ++k)
    s += f[plus[k]];
#pragma GCC unroll 5
  for (int k = 0; k < 5; ++k)
    s -= f[minus[k]];
  return s;
}

static inline void trt_collide(const double *f, double *o, const trt_t *p)
{
  double ux = sum5(f, JX_PLUS, JX_MINUS);
  double uy = sum5(f, JY_PLUS, JY_MINUS);
  double uz = sum5(f, JZ_PLUS, JZ_MINUS);
  double rho = f[RHO_ORDER[0]];
#pragma GCC unroll 19
  for (int k = 1; k < Q; ++k)
    rho += f[RHO_ORDER[k]];
  double D = rho - (ux * ux + uy * uy + uz * uz) * 3.0 / 2.0;

  o[C] = f[C] - p->we * (f[C] - p->w0 * D);
#pragma GCC unroll 9
  for (int k = 0; k < 9; ++k) {
    int i = PAIR[k][0], j = PAIR[k][1];
    double ui = CX[i] * ux + CY[i] * uy + CZ[i] * uz; /* exact: coefficients are 0, +-1 */
    /* the reference forms -ux+uy as (-ux)+uy; the sum of two terms is the
     * same either way, and one- or two-term sums are exact re-statements */
    int first = k < 3;
    double wX3 = first ? p->w1x3 : p->w2x3, wNH = first ? p->w1nh : p->w2nh;
    double wIndep = (first ? p->w1 : p->w2) * D;
    double even   = p->we * (0.5 * (f[i] + f[j]) - ui * ui * wNH - wIndep);
    double odd    = p->wo * (0.5 * (f[i] - f[j]) - ui * wX3);
    o[i]          = f[i] - even - odd;
    o[j]          = f[j] - even + odd;
  }
}

/* --------------------------------------------------------------------- */
/* lattice state with a halo, src/lbm.c:400-599                          */
/* --------------------------------------------------------------------- */
typedef struct {
  int prop, layout, blocked;
  int n[3];
  int64_t g[3], gcells;
  double *pdf[2];
  int64_t *bbSrc, *bbDst, nbb;
  const int *lat;
  int per[3];
  int iteration;
  int64_t nb[Q]; /* halo'd cell index of x + c_d minus that of x */
} ost_t;

/* element of slot d of halo'd cell `cell` (src/lbm.h:113-121) */
static inline int64_t eidx(const ost_t *s, int64_t cell, int d)
{
  if (s->layout == L_SOA)
    return d * s->gcells + cell;
  return cell * Q + d;
}

static inline int64_t hcell(const ost_t *s, int64_t x, int64_t y, int64_t z)
{
  return (x * s->g[1] + y) * s->g[2] + z;
}

static inline int64_t pidx(const ost_t *s, int64_t x, int64_t y, int64_t z, int d)
{
  return eidx(s, hcell(s, x, y, z), d);
}

static inline int obstacle(const ost_t *s, int x, int y, int z)
{
  return s->lat[lat_index(s->n, x, y, z)] == OBST;
}

/* Bounce-back / periodic copy lists, the rule of src/lbm.c:465-568:
 * for every fluid cell and direction, n = x - c (pull) or x + c (push, AA).
 * Leaving through a periodic face: wrap to p; p obstacle -> bounce-back
 * entry, else a periodic copy.  In range and obstacle -> bounce-back entry.
 * Leaving through a non-periodic face -> nothing. */
static void build_lists(ost_t *s)
{
  const int lx = s->n[0], ly = s->n[1], lz = s->n[2];
  for (int pass = 0; pass < 2; ++pass) {
    int64_t cnt = 0;
    for (int x = 0; x < lx; ++x)
      for (int y = 0; y < ly; ++y)
        for (int z = 0; z < lz; ++z) {
          if (obstacle(s, x, y, z))
            continue;
          for (int d = 0; d < Q; ++d) {
            int sg = s->prop == P_PULL ? -1 : 1;
            int nx = x + sg * CX[d], ny = y + sg * CY[d], nz = z + sg * CZ[d];
            int outX = nx < 0 || nx >= lx, outY = ny < 0 || ny >= ly, outZ = nz < 0 || nz >= lz;
            int64_t src, dst;
            if ((outX && s->per[0]) || (outY && s->per[1]) || (outZ && s->per[2])) {
              int px = nx < 0 ? lx - 1 : (nx >= lx ? 0 : nx);
              int py = ny < 0 ? ly - 1 : (ny >= ly ? 0 : ny);
              int pz = nz < 0 ? lz - 1 : (nz >= lz ? 0 : nz);
              if (obstacle(s, px, py, pz)) {
                if (s->prop == P_PULL) {
                  src = pidx(s, x + 1, y + 1, z + 1, INV[d]);
                  dst = pidx(s, nx + 1, ny + 1, nz + 1, d);
                } else {
                  src = pidx(s, nx + 1, ny + 1, nz + 1, d);
                  dst = pidx(s, x + 1, y + 1, z + 1, INV[d]);
                }
              } else {
                if (s->prop == P_PULL) {
                  src = pidx(s, px + 1, py + 1, pz + 1, d);
                  dst = pidx(s, nx + 1, ny + 1, nz + 1, d);
                } else {
                  src = pidx(s, nx + 1, ny + 1, nz + 1, d);
                  dst = pidx(s, px + 1, py + 1, pz + 1, d);
                }
              }
            } else if (outX || outY || outZ) {
              continue;
            } else if (obstacle(s, nx, ny, nz)) {
              if (s->prop == P_PULL) {
                src = pidx(s, x + 1, y + 1, z + 1, INV[d]);
                dst = pidx(s, nx + 1, ny + 1, nz + 1, d);
              } else {
                src = pidx(s, nx + 1, ny + 1, nz + 1, d);
                dst = pidx(s, x + 1, y + 1, z + 1, INV[d]);
              }
            } else {
              continue;
            }
            if (pass == 1) {
              s->bbSrc[cnt] = src;
              s->bbDst[cnt] = dst;
            }
            ++cnt;
          }
        }
    if (pass == 0) {
      s->nbb   = cnt;
      s->bbSrc = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cnt + 1));
      s->bbDst = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cnt + 1));
    }
  }
}

/* node accessors: which stored slot holds logical f_d of cell (x,y,z)
 * (src/lbm.c:193-294).  AA: even Iteration -> slot inv d of x - c_d,
 * otherwise (odd, or -1 before the first step) the local slot d. */
static inline int64_t node_slot_at(const ost_t *s, int64_t cell, int d)
{
  switch (s->prop) {
  case P_PUSH:
    return eidx(s, cell, d);
  case P_PULL:
    return eidx(s, cell - s->nb[d], d);
  default:
    if (s->iteration % 2 == 0)
      return eidx(s, cell - s->nb[d], INV[d]);
    return eidx(s, cell, d);
  }
}

static int64_t node_slot(const ost_t *s, int x, int y, int z, int d)
{
  return node_slot_at(s, hcell(s, x + 1, y + 1, z + 1), d);
}

/* body force after streaming, src/lbm.c:902-922: f_d += 3 w_d c_dx g on
 * the logical PDFs of every non-obstacle cell (the slots node_slot_at
 * names, taken as fixed offsets from the cell's first element) */
static void add_force(ost_t *s, double *a, double xforce)
{
  const double w[Q] = { 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
    1.0 / 36, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 18, 1.0 / 36, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 3 };
  double fd[Q];
  int64_t off[Q];
  const int64_t base0 = eidx(s, 0, 0);
  for (int d = 0; d < Q; ++d) {
    fd[d]  = 3.0 * w[d] * CX[d] * xforce;
    off[d] = node_slot_at(s, 0, d) - base0;
  }
  const int aos = s->layout == L_AOS;
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < s->n[0]; ++x)
    for (int y = 0; y < s->n[1]; ++y)
      for (int z = 0; z < s->n[2]; ++z) {
        if (obstacle(s, x, y, z))
          continue;
        const int64_t cell = hcell(s, x + 1, y + 1, z + 1);
        double *b          = a + (aos ? cell * Q : cell);
        for (int d = 0; d < Q; ++d)
          b[off[d]] += fd[d];
      }
}

static void copy_lists(const ost_t *s, double *a, int forward)
{
  if (forward)
    for (int64_t i = 0; i < s->nbb; ++i)
      a[s->bbDst[i]] = a[s->bbSrc[i]];
  else
    for (int64_t i = 0; i < s->nbb; ++i)
      a[s->bbSrc[i]] = a[s->bbDst[i]];
}

/* one two-lattice step: push (local read, write to x + c) or pull (read
 * from x - c, local write) over every interior cell, obstacles included
 * (src/kernel_push_soa.c:94-199, src/kernel_pull_soa.c:94-199); the
 * blocked variants visit the same cells in 8^3 tiles
 * (src/kernel_blk_pull_soa.c:38-47). */
static inline void cell_twolattice(const ost_t *s, const double *src, double *dst, int x,
    int y, int z, const trt_t *p)
{
  double f[Q], o[Q];
  const int64_t cell = hcell(s, x, y, z);
  if (s->prop == P_PULL) {
#pragma GCC unroll 19
    for (int d = 0; d < Q; ++d)
      f[d] = src[eidx(s, cell - s->nb[d], d)];
    trt_collide(f, o, p);
#pragma GCC unroll 19
    for (int d = 0; d < Q; ++d)
      dst[eidx(s, cell, d)] = o[d];
  } else {
#pragma GCC unroll 19
    for (int d = 0; d < Q; ++d)
      f[d] = src[eidx(s, cell, d)];
    trt_collide(f, o, p);
#pragma GCC unroll 19
    for (int d = 0; d < Q; ++d)
      dst[eidx(s, cell + s->nb[d], d)] = o[d];
  }
}

static void step_twolattice(ost_t *s, const double *src, double *dst, const trt_t *p)
{
  const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  if (!s->blocked) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int x = 1; x <= nx; ++x)
      for (int y = 1; y <= ny; ++y)
        for (int z = 1; z <= nz; ++z)
          cell_twolattice(s, src, dst, x, y, z, p);
  } else {
    const int bk = 8;
#pragma omp parallel for collapse(2) schedule(static)
    for (int bx = 1; bx <= nx; bx += bk)
      for (int by = 1; by <= ny; by += bk)
        for (int bz = 1; bz <= nz; bz += bk) {
          int ex = bx + bk <= nx + 1 ? bx + bk : nx + 1;
          int ey = by + bk <= ny + 1 ? by + bk : ny + 1;
          int ez = bz + bk <= nz + 1 ? bz + bk : nz + 1;
          for (int x = bx; x < ex; ++x)
            for (int y = by; y < ey; ++y)
              for (int z = bz; z < ez; ++z)
                cell_twolattice(s, src, dst, x, y, z, p);
        }
  }
}

/* AA pattern on one array (src/kernel_aa_soa.c:36-286).
 * Even: fluid cells only, read slot d locally, write post-collision f_d
 * to the local slot inv d.  Odd: every interior cell, read f_d from slot
 * inv d of x - c_d, write post-collision f_d to slot d of x + c_d.
 * Each cell owns the slots it reads and writes, so cells are independent. */
static void step_aa(ost_t *s, double *a, int odd, const trt_t *p)
{
  const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 1; x <= nx; ++x)
    for (int y = 1; y <= ny; ++y)
      for (int z = 1; z <= nz; ++z) {
        double f[Q], o[Q];
        const int64_t cell = hcell(s, x, y, z);
        if (!odd) {
          if (obstacle(s, x - 1, y - 1, z - 1))
            continue;
          for (int d = 0; d < Q; ++d)
            f[d] = a[eidx(s, cell, d)];
          trt_collide(f, o, p);
          for (int d = 0; d < Q; ++d)
            a[eidx(s, cell, INV[d])] = o[d];
        } else {
          for (int d = 0; d < Q; ++d)
            f[d] = a[eidx(s, cell - s->nb[d], INV[d])];
          trt_collide(f, o, p);
          for (int d = 0; d < Q; ++d)
            a[eidx(s, cell + s->nb[d], d)] = o[d];
        }
      }
}

static void zou_he_cell(double *f, int inlet, double rho);

/* kernelComputeBoundaryConditions (src/lbm.c:676-785) on the running
 * state, through the node accessors: the BoundaryConditionsGetPdf /
 * SetPdf of each variant (src/lbm.c:296-396) read and write the slot that
 * GetNode / SetNode use for that direction.  Only the five entering
 * populations are written. */
static void zou_he_state(const ost_t *s, double *a, double rhoIn, double rhoOut)
{
  static const int inSet[5] = { E, NE, SE, TE, BE }, outSet[5] = { W, SW, NW, BW, TW };
  const int nX = s->n[0], nY = s->n[1], nZ = s->n[2];
  double f[Q];
  for (int z = 1; z < nZ - 1; ++z)
    for (int y = 1; y < nY - 1; ++y)
      for (int side = 0; side < 2; ++side) {
        const int x = side ? nX - 1 : 0;
        if (s->lat[lat_index(s->n, x, y, z)] != (side ? 4 : 2)) /* OUTLET : INLET */
          continue;
        const int64_t cell = hcell(s, x + 1, y + 1, z + 1);
        for (int d = 0; d < Q; ++d)
          f[d] = a[node_slot_at(s, cell, d)];
        zou_he_cell(f, !side, side ? rhoOut : rhoIn);
        const int *set = side ? outSet : inSet;
        for (int k = 0; k < 5; ++k)
          a[node_slot_at(s, cell, set[k])] = f[set[k]];
      }
}

static int parse_kernel(const char *name, int *prop, int *layout, int *blocked)
{
  static const struct {
    const char *name;
    int prop, layout, blocked;
  } tab[] = {
    { "pull-soa", P_PULL, L_SOA, 0 },
    { "pull-aos", P_PULL, L_AOS, 0 },
    { "push-soa", P_PUSH, L_SOA, 0 },
    { "push-aos", P_PUSH, L_AOS, 0 },
    { "blk-push-soa", P_PUSH, L_SOA, 1 },
    { "blk-pull-soa", P_PULL, L_SOA, 1 },
    { "aa-soa", P_AA, L_SOA, 0 },
  };
  for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i)
    if (!strcasecmp(name, tab[i].name)) {
      *prop    = tab[i].prop;
      *layout  = tab[i].layout;
      *blocked = tab[i].blocked;
      return 0;
    }
  return -1;
}

/* One run.
 *   kernel      one of the seven reference names (src/lbm.c:33-41)
 *   lattice     flags from oracle_geometry
 *   periodic    periodicity after geometry forcing
 *   init        logical PDFs set on every non-obstacle cell before the run
 *               (as SetNode with Iteration -1), or NULL for all zero
 *   iterations  requested steps; aa-soa runs them in pairs, so an odd
 *               count runs one more (src/kernel_aa_soa.c:36)
 *   bodyforce   1: add 3 w_i c_ix xforce after each step's bounce-back
 *   out         logical PDFs after the run (obstacle cells: 0)
 *   steps_done  steps actually run
 * Returns 0, or -1 for an unknown kernel. */
static int run_case(const char *kernel, const int dims[3], const int *lattice,
    const int periodic[3], const double *init, int iterations, double omega, double magic,
    double xforce, int bodyforce, int zouHe, double rhoIn, double rhoOut, double *out,
    int *steps_done)
{
  ost_t s;
  memset(&s, 0, sizeof(s));
  if (parse_kernel(kernel, &s.prop, &s.layout, &s.blocked))
    return -1;
  for (int a = 0; a < 3; ++a) {
    s.n[a]   = dims[a];
    s.g[a]   = dims[a] + 2;
    s.per[a] = periodic[a];
  }
  s.gcells    = s.g[0] * s.g[1] * s.g[2];
  for (int d = 0; d < Q; ++d)
    s.nb[d] = (CX[d] * s.g[1] + CY[d]) * s.g[2] + CZ[d];
  s.lat       = lattice;
  s.iteration = -1;
  int narrays = s.prop == P_AA ? 1 : 2;
  for (int i = 0; i < narrays; ++i)
    s.pdf[i] = (double *)calloc((size_t)(s.gcells * Q), sizeof(double));
  build_lists(&s);

  const int nx = dims[0], ny = dims[1], nz = dims[2];
  if (init) {
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          if (obstacle(&s, x, y, z))
            continue;
          const double *v = init + (((int64_t)x * ny + y) * nz + z) * Q;
          for (int d = 0; d < Q; ++d)
            s.pdf[0][node_slot(&s, x, y, z, d)] = v[d];
        }
  }

  trt_t p;
  trt_setup(&p, omega, magic);
  double *final = s.pdf[0];
  int done      = 0;
  if (s.prop == P_AA) {
    for (int iter = 0; iter < iterations; iter += 2) {
      step_aa(&s, s.pdf[0], 0, &p);
      copy_lists(&s, s.pdf[0], 0);
      s.iteration = iter;
      if (bodyforce)
        add_force(&s, s.pdf[0], xforce);
      if (zouHe)
        zou_he_state(&s, s.pdf[0], rhoIn, rhoOut);
      step_aa(&s, s.pdf[0], 1, &p);
      copy_lists(&s, s.pdf[0], 1);
      s.iteration = iter + 1;
      if (bodyforce)
        add_force(&s, s.pdf[0], xforce);
      if (zouHe)
        zou_he_state(&s, s.pdf[0], rhoIn, rhoOut);
      done += 2;
    }
  } else {
    double *src = s.pdf[0], *dst = s.pdf[1];
    for (int iter = 0; iter < iterations; ++iter) {
      step_twolattice(&s, src, dst, &p);
      copy_lists(&s, dst, 1);
      if (bodyforce)
        add_force(&s, dst, xforce);
      if (zouHe)
        zou_he_state(&s, dst, rhoIn, rhoOut);
      double *t = src;
      src       = dst;
      dst       = t;
      ++done;
    }
    final = src;
  }

  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        double *v = out + (((int64_t)x * ny + y) * nz + z) * Q;
        for (int d = 0; d < Q; ++d)
          v[d] = obstacle(&s, x, y, z) ? 0.0 : final[node_slot(&s, x, y, z, d)];
      }
  if (steps_done)
    *steps_done = done;
  free(s.pdf[0]);
  free(s.pdf[1]);
  free(s.bbSrc);
  free(s.bbDst);
  return 0;
}

int oracle_run(const char *kernel, const int dims[3], const int *lattice,
    const int periodic[3], const double *init, int iterations, double omega, double magic,
    double xforce, int bodyforce, double *out, int *steps_done)
{
  return run_case(kernel, dims, lattice, periodic, init, iterations, omega, magic, xforce,
      bodyforce, 0, 0.0, 0.0, out, steps_done);
}

/* oracle_run with the Zou-He inlet / outlet applied after every step (after
 * its bounce-back copies and its body force), on the cells the lattice
 * flags LAT_CELL_INLET (x = 0) and LAT_CELL_OUTLET (x = nX-1) */
int oracle_run_bc(const char *kernel, const int dims[3], const int *lattice,
    const int periodic[3], const double *init, int iterations, double omega, double magic,
    double xforce, int bodyforce, double rhoIn, double rhoOut, double *out, int *steps_done)
{
  return run_case(kernel, dims, lattice, periodic, init, iterations, omega, magic, xforce,
      bodyforce, 1, rhoIn, rhoOut, out, steps_done);
}

/* --------------------------------------------------------------------- */
/* seeded initial state                                                   */
/* --------------------------------------------------------------------- */
static const double WEIGHT[Q] = { 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36,
  1.0 / 36, 1.0 / 36, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 18, 1.0 / 36,
  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 3 };

/* f_d = w_d * (rho + eps), eps = (2u - 1) * amp with u from
 * oracle_hash(seed, cell, d), cell = (x*nY + y)*nZ + z.  The operations are
 * written so that no two of them may be contracted into one FMA. */
void oracle_init_random(const int dims[3], const int *lattice, uint64_t seed, double amp,
    double rho, double *out)
{
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        int64_t cell = ((int64_t)x * ny + y) * nz + z;
        double *v    = out + cell * Q;
        int fluid    = lattice[lat_index(dims, x, y, z)] != OBST;
        for (int d = 0; d < Q; ++d) {
          if (!fluid) {
            v[d] = 0.0;
            continue;
          }
          double u   = u01(oracle_hash(seed, (uint64_t)cell, (uint64_t)d));
          volatile double t = 2.0 * u - 1.0;
          volatile double e = t * amp;
          volatile double r = rho + e;
          v[d]              = WEIGHT[d] * r;
        }
      }
}

/* --------------------------------------------------------------------- */
/* statistics and verification, src/lbm.c:928-1023                       */
/* --------------------------------------------------------------------- */
double oracle_density(const int dims[3], const int *lat, const double *pdfs)
{
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  double dens = 0.0;
  int64_t nf  = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        if (lat[lat_index(dims, x, y, z)] == OBST)
          continue;
        const double *v = pdfs + (((int64_t)x * ny + y) * nz + z) * Q;
        double loc      = 0.0;
        for (int d = 0; d < Q; ++d)
          loc += v[d];
        dens += loc;
        ++nf;
      }
  return dens / (double)nf;
}

void oracle_moments(const int dims[3], const int *lat, const double *pdfs, double out[4])
{
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  double m[4] = { 0, 0, 0, 0 };
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        if (lat[lat_index(dims, x, y, z)] == OBST)
          continue;
        const double *v = pdfs + (((int64_t)x * ny + y) * nz + z) * Q;
        for (int d = 0; d < Q; ++d) {
          m[0] += v[d];
          m[1] += CX[d] * v[d];
          m[2] += CY[d] * v[d];
          m[3] += CZ[d] * v[d];
        }
      }
  for (int k = 0; k < 4; ++k)
    out[k] = m[k];
}

/* Poiseuille check of src/lbm.c:964-1023: u_x at x = nX/2 (plus g/2 when
 * the force correction is on, as under VERIFICATION), analytic parabola
 * over z with centre nY/2 and radius nY/2 - 1, averaged over the one or two
 * central y rows; returns the L2 norm of the relative error. */
double oracle_verify(const int dims[3], const int *lat, const double *pdfs, double omega,
    double xforce, int force_correction)
{
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  double visc   = (1.0 / omega - 0.5) / 3.0;
  double *prof  = (double *)calloc((size_t)ny * nz, sizeof(double));
  int xpos      = nx / 2;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      if (lat[lat_index(dims, xpos, y, z)] == OBST)
        continue;
      const double *v = pdfs + (((int64_t)xpos * ny + y) * nz + z) * Q;
      double ux = v[E] + v[NE] + v[SE] + v[TE] + v[BE] - v[W] - v[NW] - v[SW] - v[TW] - v[BW];
      if (force_correction)
        ux += 0.5 * xforce;
      prof[y * nz + z] = ux;
    }
  double centre = ny / 2.0;
  double r2max  = (ny / 2.0 - 1.0) * (ny / 2.0 - 1.0);
  int even      = ny % 2 == 0;
  int yc        = (ny - even - 1) / 2;
  double dev    = 0.0;
  for (int z = 0; z < nz; ++z) {
    double r2 = (z - centre + 0.5) * (z - centre + 0.5);
    double an = r2 >= r2max ? 0.0 : xforce * (r2max - r2) / (2.0 * visc);
    double avg = even ? (prof[yc * nz + z] + prof[(yc + 1) * nz + z]) / 2.0 : prof[yc * nz + z];
    if (an != 0.0) {
      double rel = fabs(an - avg) / an;
      dev += rel * rel;
    }
  }
  free(prof);
  return sqrt(dev);
}

/* --------------------------------------------------------------------- */
/* differentiation form: Brinkman-penalised TRT and its adjoint           */
/*                                                                        */
/* Not in the reference.  The forward model is the push step above on a   */
/* halo-free lattice (periodic wrap, half-way bounce-back off obstacles,  */
/* force 3 w_d c_dx g on the streamed value) with j_s = (1 - s) j in the   */
/* equilibrium.  The adjoint transposes the streaming permutation and     */
/* obtains the per-cell Jacobian of the collision by forward-mode         */
/* automatic differentiation (dual numbers), so it shares no hand-derived */
/* formula with the GPU kernel it checks.                                 */
/* --------------------------------------------------------------------- */
typedef struct {
  double v, d;
} dual;

static inline dual dd(double v) { dual r = { v, 0.0 }; return r; }
static inline dual dadd(dual a, dual b) { dual r = { a.v + b.v, a.d + b.d }; return r; }
static inline dual dsub(dual a, dual b) { dual r = { a.v - b.v, a.d - b.d }; return r; }
static inline dual dmul(dual a, dual b) { dual r = { a.v * b.v, a.d * b.v + a.v * b.d }; return r; }
static inline dual dscale(dual a, double c) { dual r = { a.v * c, a.d * c }; return r; }
static inline dual dneg(dual a) { dual r = { -a.v, -a.d }; return r; }

static dual dsum5(const dual *f, const int *plus, const int *minus)
{
  dual s = f[plus[0]];
  for (int k = 1; k < 5; ++k)
    s = dadd(s, f[plus[k]]);
  for (int k = 0; k < 5; ++k)
    s = dsub(s, f[minus[k]]);
  return s;
}

/* the TRT update of trt_collide with j_s = keep * j, in dual numbers */
static void brinkman_collide(const dual *f, dual *o, const trt_t *p, dual keep)
{
  const dual ux = dsum5(f, JX_PLUS, JX_MINUS), uy = dsum5(f, JY_PLUS, JY_MINUS),
             uz = dsum5(f, JZ_PLUS, JZ_MINUS);
  dual rho = f[RHO_ORDER[0]];
  for (int k = 1; k < Q; ++k)
    rho = dadd(rho, f[RHO_ORDER[k]]);
  const dual sx = dmul(keep, ux), sy = dmul(keep, uy), sz = dmul(keep, uz);
  const dual q  = dadd(dadd(dmul(sx, sx), dmul(sy, sy)), dmul(sz, sz));
  const dual D  = dsub(rho, dscale(dscale(q, 3.0), 0.5));
  o[C] = dsub(f[C], dscale(dsub(f[C], dscale(D, p->w0)), p->we));
  const dual s3[3] = { sx, sy, sz };
  for (int k = 0; k < 9; ++k) {
    const int i = PAIR[k][0], j = PAIR[k][1];
    const int ci[3] = { CX[i], CY[i], CZ[i] };
    dual ui = dd(0.0);
    int first = 1;
    for (int a = 0; a < 3; ++a) {
      if (ci[a] == 0)
        continue;
      const dual t = ci[a] > 0 ? s3[a] : dneg(s3[a]);
      ui           = first ? t : dadd(ui, t);
      first        = 0;
    }
    const int axis  = k < 3;
    const double wX3 = axis ? p->w1x3 : p->w2x3, wNH = axis ? p->w1nh : p->w2nh;
    const dual wD   = dscale(D, axis ? p->w1 : p->w2);
    const dual even = dscale(dsub(dsub(dscale(dadd(f[i], f[j]), 0.5), dscale(dmul(ui, ui), wNH)), wD), p->we);
    const dual odd  = dscale(dsub(dscale(dsub(f[i], f[j]), 0.5), dscale(ui, wX3)), p->wo);
    o[i] = dsub(dsub(f[i], even), odd);
    o[j] = dadd(dsub(f[j], even), odd);
  }
}

typedef struct {
  int n[3];
  const int *lat;
  const double *s;
  trt_t p;
  double fd[Q];
} bk_t;

static inline int bk_wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }
static inline int64_t bk_cell(const bk_t *b, int x, int y, int z)
{
  return ((int64_t)x * b->n[1] + y) * b->n[2] + z;
}
static inline int bk_fluid(const bk_t *b, int x, int y, int z)
{
  return b->lat[lat_index(b->n, x, y, z)] != OBST;
}

static void bk_step(const bk_t *b, const double *in, double *out)
{
  const int nx = b->n[0], ny = b->n[1], nz = b->n[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        if (!bk_fluid(b, x, y, z))
          continue;
        const int64_t c = bk_cell(b, x, y, z);
        dual f[Q], g[Q];
        for (int d = 0; d < Q; ++d)
          f[d] = dd(in[c * Q + d]);
        brinkman_collide(f, g, &b->p, dd(1.0 - b->s[c]));
        for (int d = 0; d < Q; ++d) {
          const int tx = bk_wrap(x + CX[d], nx), ty = bk_wrap(y + CY[d], ny), tz = bk_wrap(z + CZ[d], nz);
          if (bk_fluid(b, tx, ty, tz))
            out[bk_cell(b, tx, ty, tz) * Q + d] = g[d].v + b->fd[d];
          else
            out[c * Q + INV[d]] = g[d].v + b->fd[INV[d]];
        }
      }
}

static double bk_objective(const bk_t *b, const double *f, int outletX, double *lam)
{
  const int nx = b->n[0], ny = b->n[1], nz = b->n[2];
  double J = 0.0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t c = bk_cell(b, x, y, z);
        double rho = 0.0, j[3] = { 0, 0, 0 };
        const int on = x == outletX && bk_fluid(b, x, y, z);
        if (on)
          for (int d = 0; d < Q; ++d) {
            rho += f[c * Q + d];
            j[0] += CX[d] * f[c * Q + d];
            j[1] += CY[d] * f[c * Q + d];
            j[2] += CZ[d] * f[c * Q + d];
          }
        const double j2 = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        if (on)
          J += 0.5 * rho * j2;
        if (lam)
          for (int d = 0; d < Q; ++d)
            lam[c * Q + d] = on ? 0.5 * j2 + rho * (CX[d] * j[0] + CY[d] * j[1] + CZ[d] * j[2]) : 0.0;
      }
  return J;
}

/* T steps of the Brinkman model from init (logical [x][y][z][19]) with
 * solid fraction s ([x][y][z]); returns J at x = outletX.  final_state and
 * grad (dJ/ds, [x][y][z]) may be NULL. */
double oracle_brinkman(const int dims[3], const int *lattice, const double *s,
    const double *init, int T, double omega, double magic, double xforce, int outletX,
    double *final_state, double *grad)
{
  bk_t b;
  for (int a = 0; a < 3; ++a)
    b.n[a] = dims[a];
  b.lat = lattice;
  b.s   = s;
  trt_setup(&b.p, omega, magic);
  for (int d = 0; d < Q; ++d)
    b.fd[d] = 3.0 * WEIGHT[d] * CX[d] * xforce;
  const int64_t nc = (int64_t)dims[0] * dims[1] * dims[2];
  const size_t sz  = (size_t)(nc * Q);
  const int keep_all = grad != NULL;
  double *states = (double *)calloc(sz * (size_t)(keep_all ? T + 1 : 2), sizeof(double));
  memcpy(states, init, sz * sizeof(double));
  for (int t = 0; t < T; ++t) {
    double *in  = states + sz * (size_t)(keep_all ? t : (t & 1));
    double *out = states + sz * (size_t)(keep_all ? t + 1 : ((t + 1) & 1));
    memcpy(out, in, sz * sizeof(double)); /* obstacle slots keep their values */
    bk_step(&b, in, out);
  }
  const double *fT = states + sz * (size_t)(keep_all ? T : (T & 1));
  if (final_state)
    memcpy(final_state, fT, sz * sizeof(double));
  double *lam = grad ? (double *)calloc(sz, sizeof(double)) : NULL;
  const double J = bk_objective(&b, fT, outletX, lam);
  if (grad) {
    double *lam2 = (double *)calloc(sz, sizeof(double));
    for (int64_t i = 0; i < nc; ++i)
      grad[i] = 0.0;
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    for (int t = T - 1; t >= 0; --t) {
      const double *ft = states + sz * (size_t)t;
#pragma omp parallel for collapse(2) schedule(static)
      for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
          for (int z = 0; z < nz; ++z) {
            if (!bk_fluid(&b, x, y, z))
              continue;
            const int64_t c = bk_cell(&b, x, y, z);
            double mu[Q];
            /* transpose of the streaming permutation */
            for (int k = 0; k < Q; ++k) {
              const int tx = bk_wrap(x + CX[k], nx), ty = bk_wrap(y + CY[k], ny), tz = bk_wrap(z + CZ[k], nz);
              mu[k] = bk_fluid(&b, tx, ty, tz) ? lam[bk_cell(&b, tx, ty, tz) * Q + k] : lam[c * Q + INV[k]];
            }
            dual f[Q], g[Q];
            for (int d = 0; d < Q; ++d)
              f[d] = dd(ft[c * Q + d]);
            /* column d of the collision Jacobian: seed f_d */
            for (int d = 0; d < Q; ++d) {
              f[d].d = 1.0;
              brinkman_collide(f, g, &b.p, dd(1.0 - s[c]));
              f[d].d = 0.0;
              double acc = 0.0;
              for (int k = 0; k < Q; ++k)
                acc += mu[k] * g[k].d;
              lam2[c * Q + d] = acc;
            }
            /* d g / d s: seed keep = 1 - s with derivative -1 */
            dual keep = { 1.0 - s[c], -1.0 };
            brinkman_collide(f, g, &b.p, keep);
            double acc = 0.0;
            for (int k = 0; k < Q; ++k)
              acc += mu[k] * g[k].d;
            grad[c] += acc;
          }
      double *tmp = lam;
      lam         = lam2;
      lam2        = tmp;
    }
    free(lam2);
    free(lam);
  }
  free(states);
  return J;
}

/* the porous field of the library, restated: u from the hash chain over
 * (seed, cell, 0), periodic Gaussian smoothing (radius ceil(3 sigma)),
 * standardised, then the logistic a z + b */
void oracle_porous_field(const int dims[3], uint64_t seed, double sigma, double a, double b,
    double *s)
{
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int64_t n = (int64_t)nx * ny * nz;
  double *u = (double *)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i)
    u[i] = u01(oracle_hash(seed, (uint64_t)i, 0));
  const int rad = sigma > 0.0 ? (int)ceil(3.0 * sigma) : 0;
  double *w = (double *)malloc(sizeof(double) * (size_t)(2 * rad + 1));
  double wsum = 0.0;
  for (int k = -rad; k <= rad; ++k) {
    w[k + rad] = sigma > 0.0 ? exp(-0.5 * (double)k * k / (sigma * sigma)) : 1.0;
    wsum += w[k + rad];
  }
  for (int k = 0; k <= 2 * rad; ++k)
    w[k] /= wsum;
  double *t = (double *)malloc(sizeof(double) * (size_t)n);
  const int nn[3] = { nx, ny, nz };
  for (int axis = 0; axis < 3; ++axis) {
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          double acc = 0.0;
          for (int k = -rad; k <= rad; ++k) {
            int p[3] = { x, y, z };
            p[axis]  = ((p[axis] + k) % nn[axis] + nn[axis]) % nn[axis];
            acc += w[k + rad] * u[((int64_t)p[0] * ny + p[1]) * nz + p[2]];
          }
          t[((int64_t)x * ny + y) * nz + z] = acc;
        }
    memcpy(u, t, sizeof(double) * (size_t)n);
  }
  double mean = 0.0, var = 0.0;
  for (int64_t i = 0; i < n; ++i)
    mean += u[i];
  mean /= (double)n;
  for (int64_t i = 0; i < n; ++i)
    var += (u[i] - mean) * (u[i] - mean);
  const double sd = sqrt(var / (double)n);
  for (int64_t i = 0; i < n; ++i)
    s[i] = 1.0 / (1.0 + exp(-(a * (sd > 0.0 ? (u[i] - mean) / sd : 0.0) + b)));
  free(u);
  free(t);
  free(w);
}

/* ------------------------------------------------------------------------ */
/* Zou-He pressure boundaries (src/lbm.c:676-785)                           */
/* ------------------------------------------------------------------------ */
/* On the logical PDFs: at every inlet cell of the plane x = 0 (flag
 * LAT_CELL_INLET) the density is set to rhoIn by the five populations
 * entering along +x; at every outlet cell of x = nX-1 (LAT_CELL_OUTLET) to
 * rhoOut by the five entering along -x.  y and z run over 1 .. n-2, as in
 * the reference, and the arithmetic keeps its grouping (rho^-1 as a
 * multiplier, the 1/4 and 1/6 factors, the sum order). */
/* one cell's update on its 19 logical PDFs f (inlet: the five populations
 * entering along +x; outlet: along -x) */
static void zou_he_cell(double *f, int inlet, double rho)
{
  const double third = 1.0 / 3.0, fourth = 1.0 / 4.0, sixth = 1.0 / 6.0;
  const double inv = 1.0 / rho;
  if (inlet) {
    const double ux  = 1.0 - (f[C] + (f[T] + f[B] + f[S] + f[N]) +
                                 (f[TS] + f[BS] + f[TN] + f[BN]) +
                                 2.0 * (f[SW] + f[TW] + f[W] + f[BW] + f[NW])) *
                                inv;
    const double ind = sixth * rho * ux;
    f[E]  = f[W] + third * rho * ux;
    f[NE] = f[SW] - fourth * (f[N] - f[S]) + ind;
    f[SE] = f[NW] + fourth * (f[N] - f[S]) + ind;
    f[TE] = f[BW] - fourth * (f[T] - f[B]) + ind;
    f[BE] = f[TW] + fourth * (f[T] - f[B]) + ind;
  } else {
    const double ux  = -1.0 + (f[C] + (f[T] + f[B] + f[S] + f[N]) +
                                  (f[TS] + f[BS] + f[TN] + f[BN]) +
                                  2.0 * (f[NE] + f[BE] + f[E] + f[TE] + f[SE])) *
                                 inv;
    const double ind = sixth * rho * ux;
    f[W]  = f[E] - third * rho * ux;
    f[SW] = f[NE] + fourth * (f[N] - f[S]) - ind;
    f[NW] = f[SE] - fourth * (f[N] - f[S]) - ind;
    f[BW] = f[TE] + fourth * (f[T] - f[B]) - ind;
    f[TW] = f[BE] - fourth * (f[T] - f[B]) - ind;
  }
}

void oracle_zou_he(const int dims[3], const int *lat, double *pdfs, double rhoIn, double rhoOut)
{
  const int nX = dims[0], nY = dims[1], nZ = dims[2];
  for (int z = 1; z < nZ - 1; ++z)
    for (int y = 1; y < nY - 1; ++y) {
      if (lat[((size_t)z * nY + y) * nX + 0] == 2) /* LAT_CELL_INLET */
        zou_he_cell(pdfs + (((size_t)0 * nY + y) * nZ + z) * Q, 1, rhoIn);
      if (lat[((size_t)z * nY + y) * nX + (nX - 1)] == 4) /* LAT_CELL_OUTLET */
        zou_he_cell(pdfs + (((size_t)(nX - 1) * nY + y) * nZ + z) * Q, 0, rhoOut);
    }
}
