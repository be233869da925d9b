// Drop-in simple estimators (reference src/include/estimators.h:8-34).
// decodeMaxPrior runs on the GPU through ginf_max_prior_host (every node's
// lowest-index label of minimal data cost); decodeRandom draws labels from
// the C library's rand(), as the reference does, so it stays a host routine.
#ifndef GRIDINF_INCLUDE_ESTIMATORS_H
#define GRIDINF_INCLUDE_ESTIMATORS_H

#include "core.h"

#include <cstdlib>

namespace ginf {
template <typename T>
void decodeMaxPrior(Grid<T> *grid, Matrix<int> *result);

template <typename T>
void decodeRandom(Grid<T> *grid, Matrix<int> *result) {
    for (int y = 0; y < grid->getHeight(); y++)
        for (int x = 0; x < grid->getWidth(); x++)
            result->get(x, y) = int(rand() * 1.0 / RAND_MAX * grid->getNumLabels());
}
}  // namespace ginf

#endif
