/* lbm_geometry.h -- the device side of geometryCreate and kernelVerify,
 * called by the host C of lbm_host.c (implemented in lbm_runtime.hip). */
#ifndef LBM_AMD_LBM_GEOMETRY_H
#define LBM_AMD_LBM_GEOMETRY_H

#include "lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* geometry types of geometryCreate (src/lbm.c:49-156, plus spheres) */
enum { LBM_GEO_BOX, LBM_GEO_CHANNEL, LBM_GEO_PIPE, LBM_GEO_BLOCKS, LBM_GEO_FLUID,
       LBM_GEO_SPHERES };

/* Builds the cell flags of geometryCreate on the current HIP device and
 * copies them into lattice (dims[0]*dims[1]*dims[2] flags, x fastest,
 * LatticeDescType.Lattice order): obstacle on the boundary layers of the
 * non-periodic axes, inside the pipe's wall (type pipe), in the blocks of
 * edge blockSize (type blocks) and inside the spheres (4 doubles per
 * sphere: centre x, y, z and radius; 6 ints: lowest and highest cell of
 * its bounding box per axis, wrapped through the periodic axes), fluid
 * elsewhere.  Returns the number of fluid cells.  Device errors print a
 * message and exit(EXIT_FAILURE). */
long long lbmBuildFlags(const int dims[3], const int periodic[3], int type, int blockSize,
    int nSpheres, const double *spheres, const int *bounds, LatticeType *lattice);

#ifdef __cplusplus
}
#endif

#endif
