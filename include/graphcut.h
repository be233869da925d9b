// Drop-in alpha-beta swap API (reference src/include/graphcut.h:13-26),
// without the reference's CUDA include: both entry points run the gfx950
// push-relabel of this library through ginf_alpha_beta_host.  As in the
// reference only Grid<int> is supported.  Each label pair's graph is built
// from a fresh state (the reference CPU path carries state of nodes outside
// the pair over from the previous pair).  Errors print a message to stderr
// and exit(1).
#ifndef GRIDINF_INCLUDE_GRAPHCUT_H
#define GRIDINF_INCLUDE_GRAPHCUT_H

#include "core.h"

// Kept for source compatibility (src/include/graphcut.h:7-11); the kernels do
// not use them.
#define GINF_GRAPHCUT_NUM_ITERS_PER_GLOBAL_RELABEL 10
#define GINF_GRAPHCUT_BLOCK_WIDTH 32
#define GINF_GRAPHCUT_BLOCK_HEIGHT 8
#define GINF_GRAPHCUT_NUM_THREADS_PER_BLOCK (GINF_GRAPHCUT_BLOCK_WIDTH * GINF_GRAPHCUT_BLOCK_HEIGHT)

namespace ginf {
enum TerminalDirections { GINF_DIR_SOURCE = GINF_NUM_DIR, GINF_DIR_SINK };

template <typename T>
void decodeAlphaBeta(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result);

template <typename T>
void gpuDecodeAlphaBeta(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result);
}  // namespace ginf

#endif
