// Drop-in ICM API (reference src/include/icm.h:9-21), without the
// reference's CUDA include: all three entry points run the gfx950 kernels of
// this library through ginf_icm_host.  decodeIcmSync and gpuDecodeIcm are the
// synchronous update (src/icm.cc:5-37, src/gpu/icm.cu:21-83);
// decodeIcmAsync is the in-place row-major update (src/icm.cc:39-62),
// reproduced exactly by an anti-diagonal wavefront.  Errors print a message
// to stderr and exit(1).
#ifndef GRIDINF_INCLUDE_ICM_H
#define GRIDINF_INCLUDE_ICM_H

#include "core.h"

// Kept for source compatibility (src/include/icm.h:7); not used by the kernels.
#define GINF_ICM_BLOCK_SIZE 16

namespace ginf {
template <typename T>
void decodeIcmSync(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result);

template <typename T>
void gpuDecodeIcm(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *results);

template <typename T>
void decodeIcmAsync(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result);
}  // namespace ginf

#endif
