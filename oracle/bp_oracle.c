/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C restatement of the CPU min-sum belief propagation of
 * GridInfCuda (src/bp.cc, with the energy of src/grid.cc:102-125 and the
 * Matrix layout of src/matrix.cc).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker or as the
 * timed CPU baseline; the HIP library never calls it.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   oracle_pyramid      brute-force child sums, ceil(w/2) dims (numpy)
 *   oracle_upsample     closed-form parent copy (numpy)
 *   send_msg O(L) path  brute-force min_j(h_j + V(i,j)) on random integer
 *                       vectors; Potts identity; mean/min normalisation
 *                       invariants (both through oracle_hbp on 3x3 grids)
 *   send_msg O(L^2)     brute force on 3x3 grids (EXPLICIT / TRUNC_QUAD)
 *   oracle_half_sweep   colour parity known answers (iters=1: nodes with
 *                       x+y even keep label 0), an independent dict-based
 *                       Python message-passing restatement on tiny grids,
 *                       and exact Viterbi MAP on chain grids (H=3)
 *   oracle_labels       lowest-index argmin; 3x3 grid = decodeMaxPrior
 *   oracle_hbp          HBP on H=3 equals plain BP (coarse levels have no
 *                       interior); level handling vs the numpy restatement
 *   oracle_energy*      brute-force FP64/numpy sums
 * The reference itself cannot be compiled here (src/include/bp.h:5 pulls in
 * src/include/gpu.h, which includes <cuda.h>/<cuda_runtime.h>; the image has
 * no CUDA toolkit), so no oracle/_ref build exists.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORACLE_SM_EXPLICIT = 0, ORACLE_SM_POTTS = 1, ORACLE_SM_TRUNC_LINEAR = 2, ORACLE_SM_TRUNC_QUAD = 3 };
enum { ORACLE_NORM_MEAN = 0, ORACLE_NORM_MIN = 1 };

/* Direction order S, E, N, W and their offsets (src/include/utils.h:26-37). */
static const int DIR_X[4] = {0, 1, 0, -1};
static const int DIR_Y[4] = {1, 0, -1, 0};
#define OPP_DIR(d) ((d) >= 2 ? (d) - 2 : (d) + 2)
/* The reference MIN keeps the second operand on ties (src/include/utils.h:16). */
#define OMIN(a, b) (((a) < (b)) ? (a) : (b))

#define OT float
#define OS f32
#include "bp_oracle_t.h"
#undef OT
#undef OS

#define OT int
#define OS i32
#include "bp_oracle_t.h"
#undef OT
#undef OS

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
