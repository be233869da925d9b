// Drop-in grid MRF model, same interface as the reference's Grid<T>
// (src/include/grid.h:9-81): W x H nodes, L labels, data costs D(x,y,l) and
// an L x L smoothness table V(a,b) with a model tag.  Instantiated for int
// and float.
#ifndef GRIDINF_INCLUDE_GRID_H
#define GRIDINF_INCLUDE_GRID_H

#include "matrix.h"
#include "utils.h"

namespace ginf {
enum SmoothnessFunctions { GINF_SM_EXPLICIT, GINF_SM_POTTS, GINF_SM_TRUNC_LINEAR, GINF_SM_TRUNC_QUAD };

template <typename T>
class Grid {
private:
    Matrix<T> *dtCosts;  // (W, H, L)
    Matrix<T> *smCosts;  // (L, L): V(a, b) at a + L*b
    SmoothnessFunctions smModel;
    int width, height, numLabels;

public:
    Grid(int w, int h, int n);
    ~Grid();

    int getWidth();
    int getHeight();
    int getNumLabels();
    int getNumNodes();
    SmoothnessFunctions getSmModel();

    T getDataCost(int x, int y, int fp);
    void setDataCost(int x, int y, int fp, T cost);

    T getSmoothnessCost(int fp, int fq);
    // Sets one entry and switches the model to GINF_SM_EXPLICIT.
    void setSmoothnessCost(int fp, int fq, T cost);

    // V(a,b) = lambda * (a != b)
    void useSmoothnessPotts(T lambda);
    // V(a,b) = min(scale * |a - b|, d)
    void useSmoothnessTruncLinear(T scale, T d);
    // V(a,b) = min(scale * (a - b)^2, d)
    void useSmoothnessTruncQuad(T scale, T d);

    // Energy of a labelling: interior nodes' D(f) + V(f, f_N) + V(f, f_E),
    // accumulated in T.
    T getLabelingCost(Matrix<int> *f);
    // D(x,y,label) + sum over the 4 neighbours of V(label, f_neighbour).
    T getLabelingCost(Matrix<int> *f, int x, int y, int label);

    Matrix<T> *getDataCosts();
    Matrix<T> *getSmCosts();
};
}  // namespace ginf

#endif
