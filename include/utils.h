// Drop-in constants and helpers of the grid inference API
// (same names and values as the reference's src/include/utils.h:4-37).
#ifndef GRIDINF_INCLUDE_UTILS_H
#define GRIDINF_INCLUDE_UTILS_H

#define GINF_UNDEFINED -1

// Four neighbours per node.
#define GINF_NUM_DIR 4

// S <-> N, E <-> W.
#define GINF_OPP_DIR(x) ((x) >= 2 ? (x) - 2 : (x) + 2)

#define GINF_IS_VALID_NODE(x, y, w, h) ((x) >= 0 && (y) >= 0 && (x) < (w) && (y) < (h))

// MIN keeps the second operand on ties, as the reference's macro does.
#define GINF_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define GINF_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define GINF_ABS(x) ((x) < 0 ? (-(x)) : (x))
#define GINF_SQ(x) ((x) * (x))

// The reference capped labels at 16 on its GPU path (a 16-entry local
// array; its CPU path has no cap); the MI355X kernels support up to 1024
// labels for every model (bit-exact against the reference CPU path up to
// 256 labels for FP32 with mean normalisation, at every count for int and
// for integer-valued FP32 costs with min normalisation).  The macro keeps
// its old value for source compatibility.
#define GINF_MAX_NUM_LABELS 16
#define GINF_MAX_NUM_LABELS_LINEAR 1024
#define GINF_MAX_NUM_LABELS_GENERAL 1024

namespace ginf {
enum Directions { GINF_DIR_S, GINF_DIR_E, GINF_DIR_N, GINF_DIR_W };

// (dirX[d], dirY[d]) is the offset of the neighbour in direction d.
const int dirX[] = {0, 1, 0, -1};
const int dirY[] = {1, 0, -1, 0};
}  // namespace ginf

#endif
