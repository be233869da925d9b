// Drop-in dense N-d array (up to 4 dims, first index fastest), same interface
// as the reference's Matrix<T> (src/include/matrix.h:9-51).  Instantiated for
// int and float.  Memory is zero-initialised on construction.
#ifndef GRIDINF_INCLUDE_MATRIX_H
#define GRIDINF_INCLUDE_MATRIX_H

#include "utils.h"

#define GINF_MAT_MAX_NUM_DIMS 4

namespace ginf {
template <typename T>
class Matrix {
private:
    int numDims;
    int dimSize[GINF_MAT_MAX_NUM_DIMS];
    long long dimIdx[GINF_MAT_MAX_NUM_DIMS];  // running products of the sizes

public:
    // Flat element storage (exposed, as in the reference, for bulk copies).
    T *data;

    // Matrix(numDims, size0, size1, ...) with numDims in 1..4.
    Matrix(int numDims, ...);
    ~Matrix();

    int getNumDims();
    int getSize(int dim);
    int getTotalSize();

    T at(int x, int y, int z = 0, int w = 0);
    T &get(int x, int y, int z = 0, int w = 0);

    void copyFrom(Matrix<T> *m);
    void clear();
};
}  // namespace ginf

#endif
