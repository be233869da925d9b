// Drop-in check of the C++ API: the reference's stereo example flow
// (examples/stereo/stereo.cc:25-43) on a seeded synthetic pair instead of
// the PNGs: Grid<int>, 16 labels, costs min(|L - R(x-k)|, 20) for x >= 15,
// TruncLinear(10, 20), gpuDecodeBp 30 iterations.  Prints the labels, one row
// per line, for the test to compare with the oracle.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bp.h"
#include "core.h"
#include "ginf_bp.h"

int main(int argc, char **argv) {
    const int W = argc > 1 ? atoi(argv[1]) : 96, H = argc > 2 ? atoi(argv[2]) : 64, L = 16;
    const int iters = argc > 3 ? atoi(argv[3]) : 30;
    const int levels = argc > 4 ? atoi(argv[4]) : 1;
    std::vector<uint8_t> left(W * H), right(W * H);
    if (ginf_gen_stereo(1, W, H, L, 8, 15, 20, left.data(), right.data(), nullptr, nullptr, nullptr) != 0) return 2;
    ginf::Grid<int> grid(W, H, L);
    for (int y = 0; y < grid.getHeight(); y++)
        for (int x = grid.getNumLabels() - 1; x < grid.getWidth(); x++)
            for (int k = 0; k < L; k++) {
                int c = abs((int)left[y * W + x] - (int)right[y * W + x - k]);
                grid.setDataCost(x, y, k, c < 20 ? c : 20);
            }
    grid.useSmoothnessTruncLinear(10, 20);
    ginf::Matrix<int> result(2, grid.getWidth(), grid.getHeight());
    if (levels > 1) ginf::gpuDecodeHbp(&grid, levels, iters, &result);
    else ginf::gpuDecodeBp(&grid, iters, &result);
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) printf("%d ", result.get(x, y));
        printf("\n");
    }
    printf("energy %d\n", grid.getLabelingCost(&result));
    return 0;
}
