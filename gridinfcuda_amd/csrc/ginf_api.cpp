// C++ drop-in API (include/matrix.h, grid.h, bp.h, icm.h, estimators.h,
// graphcut.h) over the C ABI.
// Matrix and Grid keep the reference host data model (src/matrix.cc,
// src/grid.cc); the four BP templates run the HIP path via ginf_hbp_host.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bp.h"
#include "estimators.h"
#include "ginf_bp.h"
#include "graphcut.h"
#include "icm.h"

namespace ginf {

// ------------------------------------------------------------------ Matrix

template <typename T>
Matrix<T>::Matrix(int nd, ...) {
    numDims = nd;
    va_list ap;
    va_start(ap, nd);
    long long run = 1;
    for (int i = 0; i < nd && i < GINF_MAT_MAX_NUM_DIMS; ++i) {
        dimSize[i] = va_arg(ap, int);
        run *= dimSize[i];
        dimIdx[i] = run;
    }
    va_end(ap);
    data = (T *)calloc((size_t)getTotalSize() > 0 ? (size_t)run : 1, sizeof(T));
    if (!data) {
        fprintf(stderr, "ginf::Matrix: allocation of %lld elements failed\n", run);
        exit(1);
    }
}

template <typename T>
Matrix<T>::~Matrix() {
    free(data);
    data = nullptr;
}

template <typename T>
int Matrix<T>::getNumDims() { return numDims; }

template <typename T>
int Matrix<T>::getSize(int dim) { return dimSize[dim]; }

template <typename T>
int Matrix<T>::getTotalSize() { return (int)dimIdx[numDims - 1]; }

template <typename T>
T Matrix<T>::at(int x, int y, int z, int w) {
    return data[x + dimIdx[0] * y + dimIdx[1] * z + dimIdx[2] * w];
}

template <typename T>
T &Matrix<T>::get(int x, int y, int z, int w) {
    return data[x + dimIdx[0] * y + dimIdx[1] * z + dimIdx[2] * w];
}

template <typename T>
void Matrix<T>::copyFrom(Matrix<T> *m) { memcpy(data, m->data, sizeof(T) * (size_t)dimIdx[numDims - 1]); }

template <typename T>
void Matrix<T>::clear() { memset(data, 0, sizeof(T) * (size_t)dimIdx[numDims - 1]); }

template class Matrix<int>;
template class Matrix<float>;

// ------------------------------------------------------------------ Grid

template <typename T>
Grid<T>::Grid(int w, int h, int n) {
    dtCosts = new Matrix<T>(3, w, h, n);
    smCosts = new Matrix<T>(2, n, n);
    smModel = GINF_SM_POTTS;  // an all-zero table is Potts with lambda 0
    width = w;
    height = h;
    numLabels = n;
}

template <typename T>
Grid<T>::~Grid() {
    delete dtCosts;
    delete smCosts;
    dtCosts = smCosts = nullptr;
}

template <typename T> int Grid<T>::getWidth() { return width; }
template <typename T> int Grid<T>::getHeight() { return height; }
template <typename T> int Grid<T>::getNumLabels() { return numLabels; }
template <typename T> int Grid<T>::getNumNodes() { return width * height; }
template <typename T> SmoothnessFunctions Grid<T>::getSmModel() { return smModel; }
template <typename T> T Grid<T>::getDataCost(int x, int y, int f) { return dtCosts->at(x, y, f); }
template <typename T> void Grid<T>::setDataCost(int x, int y, int f, T c) { dtCosts->get(x, y, f) = c; }
template <typename T> T Grid<T>::getSmoothnessCost(int a, int b) { return smCosts->at(a, b); }

template <typename T>
void Grid<T>::setSmoothnessCost(int a, int b, T c) {
    smModel = GINF_SM_EXPLICIT;
    smCosts->get(a, b) = c;
}

template <typename T>
void Grid<T>::useSmoothnessPotts(T lambda) {
    smModel = GINF_SM_POTTS;
    for (int a = 0; a < numLabels; ++a)
        for (int b = 0; b < numLabels; ++b) smCosts->get(a, b) = lambda * (T)(a != b);
}

template <typename T>
void Grid<T>::useSmoothnessTruncLinear(T scale, T d) {
    smModel = GINF_SM_TRUNC_LINEAR;
    for (int a = 0; a < numLabels; ++a)
        for (int b = 0; b < numLabels; ++b) smCosts->get(a, b) = GINF_MIN(scale * GINF_ABS(a - b), d);
}

template <typename T>
void Grid<T>::useSmoothnessTruncQuad(T scale, T d) {
    smModel = GINF_SM_TRUNC_QUAD;
    for (int a = 0; a < numLabels; ++a)
        for (int b = 0; b < numLabels; ++b) smCosts->get(a, b) = GINF_MIN(scale * GINF_SQ(a - b), d);
}

template <typename T>
T Grid<T>::getLabelingCost(Matrix<int> *f) {
    T total = 0;
    for (int y = 1; y < height - 1; ++y)
        for (int x = 1; x < width - 1; ++x) {
            const int l = f->at(x, y);
            total += getDataCost(x, y, l);
            total += getSmoothnessCost(l, f->at(x + dirX[GINF_DIR_N], y + dirY[GINF_DIR_N]));
            total += getSmoothnessCost(l, f->at(x + dirX[GINF_DIR_E], y + dirY[GINF_DIR_E]));
        }
    return total;
}

template <typename T>
T Grid<T>::getLabelingCost(Matrix<int> *f, int x, int y, int label) {
    T total = getDataCost(x, y, label);
    for (int d = 0; d < GINF_NUM_DIR; ++d) total += getSmoothnessCost(label, f->at(x + dirX[d], y + dirY[d]));
    return total;
}

template <typename T> Matrix<T> *Grid<T>::getDataCosts() { return dtCosts; }
template <typename T> Matrix<T> *Grid<T>::getSmCosts() { return smCosts; }

template class Grid<int>;
template class Grid<float>;

// ------------------------------------------------------------------ BP

namespace {

template <typename T> int dtype_of();
template <> int dtype_of<float>() { return GINF_F32; }
template <> int dtype_of<int>() { return GINF_I32; }

template <typename T>
void run_hbp(const char *who, Grid<T> *grid, int levels, int iters, Matrix<int> *result) {
    ginf_bp_desc d;
    memset(&d, 0, sizeof d);
    d.width = grid->getWidth();
    d.height = grid->getHeight();
    d.num_labels = grid->getNumLabels();
    d.dtype = dtype_of<T>();
    d.sm_model = (int)grid->getSmModel();
    d.num_levels = levels;
    d.iters_per_level = iters;
    d.norm = GINF_NORM_MEAN;
    const int rc = ginf_hbp_host(&d, grid->getDataCosts()->data, grid->getSmCosts()->data, result->data);
    if (rc != GINF_OK) {
        fprintf(stderr, "%s failed (code %d): %s\n", who, rc, ginf_last_error());
        exit(1);
    }
}

}  // namespace

template <typename T>
void decodeBp(Grid<T> *grid, int numIters, Matrix<int> *result) { run_hbp("ginf::decodeBp", grid, 1, numIters, result); }

template <typename T>
void gpuDecodeBp(Grid<T> *grid, int numIters, Matrix<int> *result) {
    run_hbp("ginf::gpuDecodeBp", grid, 1, numIters, result);
}

template <typename T>
void decodeHbp(Grid<T> *grid, int numLevels, int numItersPerLevel, Matrix<int> *result) {
    run_hbp("ginf::decodeHbp", grid, numLevels, numItersPerLevel, result);
}

template <typename T>
void gpuDecodeHbp(Grid<T> *grid, int numLevels, int numItersPerLevel, Matrix<int> *result) {
    run_hbp("ginf::gpuDecodeHbp", grid, numLevels, numItersPerLevel, result);
}

void releaseBpDeviceState() {
    const int rc = ginf_hbp_host_release();
    if (rc != GINF_OK) {
        fprintf(stderr, "ginf::releaseBpDeviceState failed (code %d): %s\n", rc, ginf_last_error());
        exit(1);
    }
}

template void decodeBp<float>(Grid<float> *, int, Matrix<int> *);
template void decodeBp<int>(Grid<int> *, int, Matrix<int> *);
template void gpuDecodeBp<float>(Grid<float> *, int, Matrix<int> *);
template void gpuDecodeBp<int>(Grid<int> *, int, Matrix<int> *);
template void decodeHbp<float>(Grid<float> *, int, int, Matrix<int> *);
template void decodeHbp<int>(Grid<int> *, int, int, Matrix<int> *);
template void gpuDecodeHbp<float>(Grid<float> *, int, int, Matrix<int> *);
template void gpuDecodeHbp<int>(Grid<int> *, int, int, Matrix<int> *);

// ------------------------------------------------------------------ ICM, estimators

namespace {

template <typename T>
void run_icm(const char *who, Grid<T> *grid, Matrix<int> *initial, int iters, int mode, Matrix<int> *result) {
    const int rc = ginf_icm_host(grid->getWidth(), grid->getHeight(), grid->getNumLabels(), dtype_of<T>(),
                                 grid->getDataCosts()->data, grid->getSmCosts()->data, initial->data, iters, mode,
                                 result->data);
    if (rc != GINF_OK) {
        fprintf(stderr, "%s failed (code %d): %s\n", who, rc, ginf_last_error());
        exit(1);
    }
}

}  // namespace

template <typename T>
void decodeIcmSync(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result) {
    run_icm("ginf::decodeIcmSync", grid, initial, numIters, GINF_ICM_SYNC, result);
}

template <typename T>
void gpuDecodeIcm(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result) {
    run_icm("ginf::gpuDecodeIcm", grid, initial, numIters, GINF_ICM_SYNC, result);
}

template <typename T>
void decodeIcmAsync(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result) {
    run_icm("ginf::decodeIcmAsync", grid, initial, numIters, GINF_ICM_ASYNC, result);
}

template <typename T>
void decodeMaxPrior(Grid<T> *grid, Matrix<int> *result) {
    const int rc = ginf_max_prior_host(grid->getWidth(), grid->getHeight(), grid->getNumLabels(), dtype_of<T>(),
                                       grid->getDataCosts()->data, result->data);
    if (rc != GINF_OK) {
        fprintf(stderr, "ginf::decodeMaxPrior failed (code %d): %s\n", rc, ginf_last_error());
        exit(1);
    }
}

template void decodeIcmSync<float>(Grid<float> *, Matrix<int> *, int, Matrix<int> *);
template void decodeIcmSync<int>(Grid<int> *, Matrix<int> *, int, Matrix<int> *);
template void gpuDecodeIcm<float>(Grid<float> *, Matrix<int> *, int, Matrix<int> *);
template void gpuDecodeIcm<int>(Grid<int> *, Matrix<int> *, int, Matrix<int> *);
template void decodeIcmAsync<float>(Grid<float> *, Matrix<int> *, int, Matrix<int> *);
template void decodeIcmAsync<int>(Grid<int> *, Matrix<int> *, int, Matrix<int> *);
template void decodeMaxPrior<float>(Grid<float> *, Matrix<int> *);
template void decodeMaxPrior<int>(Grid<int> *, Matrix<int> *);

// ------------------------------------------------------------------ alpha-beta swap

template <typename T>
void decodeAlphaBeta(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result) {
    const int rc = ginf_alpha_beta_host(grid->getWidth(), grid->getHeight(), grid->getNumLabels(), dtype_of<T>(),
                                        grid->getDataCosts()->data, grid->getSmCosts()->data, initial->data, numIters,
                                        result->data);
    if (rc != GINF_OK) {
        fprintf(stderr, "ginf::decodeAlphaBeta failed (code %d): %s\n", rc, ginf_last_error());
        exit(1);
    }
}

template <typename T>
void gpuDecodeAlphaBeta(Grid<T> *grid, Matrix<int> *initial, int numIters, Matrix<int> *result) {
    decodeAlphaBeta(grid, initial, numIters, result);
}

// Grid<int> only, as the reference instantiates (src/graphcut.cc:281, src/gpu/graphcut.cu:250).
template void decodeAlphaBeta<int>(Grid<int> *, Matrix<int> *, int, Matrix<int> *);
template void gpuDecodeAlphaBeta<int>(Grid<int> *, Matrix<int> *, int, Matrix<int> *);

}  // namespace ginf
