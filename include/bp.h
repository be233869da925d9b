// Drop-in belief propagation API (reference src/include/bp.h:13-26), without
// the reference's CUDA include: all four entry points run the gfx950 HIP
// kernels of this library through the C ABI in ginf_bp.h.  decodeBp and
// decodeHbp give the reference CPU path's labels (src/bp.cc); the gpu*
// names are the same computation.  Errors print a message to stderr and
// exit(1), the reference's GPU error policy (src/include/gpu.h:42-47).
#ifndef GRIDINF_INCLUDE_BP_H
#define GRIDINF_INCLUDE_BP_H

#include "core.h"

// Kept for source compatibility with the reference (src/include/bp.h:8-9);
// the MI355X kernels do not use a 2-D block shape.
#define GINF_BP_BLOCK_WIDTH 32
#define GINF_BP_BLOCK_HEIGHT 4

namespace ginf {
// Loopy BP: numIters red-black half-sweeps.
template <typename T>
void decodeBp(Grid<T> *grid, int numIters, Matrix<int> *result);

template <typename T>
void gpuDecodeBp(Grid<T> *grid, int numIters, Matrix<int> *result);

// Hierarchical BP: numLevels pyramid levels, numItersPerLevel half-sweeps each.
template <typename T>
void decodeHbp(Grid<T> *grid, int numLevels, int numItersPerLevel, Matrix<int> *result);

template <typename T>
void gpuDecodeHbp(Grid<T> *grid, int numLevels, int numItersPerLevel, Matrix<int> *result);

// No reference counterpart.  The four entry points above keep the device
// state of the last problem they ran (its plan, costs and labels: about 8 GB
// for 1920x1080 with 128 labels) so that a repeated call on the same problem
// neither reallocates nor re-zeroes it; the reference frees its state at the
// end of every call (src/bp.cc:176-179).  This frees it now; the next call
// allocates afresh.
void releaseBpDeviceState();
}  // namespace ginf

#endif
