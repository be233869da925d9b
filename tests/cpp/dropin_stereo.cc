// Drop-in check of the C++ API: the reference's stereo example flow
// (examples/stereo/stereo.cc:25-43) -- Grid<int>, 16 labels, costs
// min(|L - R(x-k)|, 20) for x >= 15, TruncLinear(10, 20), gpuDecodeBp -- on a
// seeded synthetic pair or on a binary PGM (P5) pair, optionally followed by
// ICM refinement (include/icm.h).  Prints the labels, one row per line, and
// the energies, for the tests to compare with the oracle.
//
//   dropin_stereo [W H iters levels [refine_iters [sync|async|ab]]]   synthetic pair (seed 1)
//   dropin_stereo --pgm left.pgm right.pgm iters levels [refine_iters [sync|async|ab]]
// Refinement: ICM (gpuDecodeIcm / decodeIcmAsync) or alpha-beta swap
// (decodeAlphaBeta, include/graphcut.h) started from the BP labels.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bp.h"
#include "core.h"
#include "estimators.h"
#include "ginf_bp.h"
#include "graphcut.h"
#include "icm.h"

namespace {

bool read_pgm(const char *path, int &w, int &h, std::vector<uint8_t> &px) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char magic[3] = {0};
    int maxv = 0;
    bool ok = fscanf(f, "%2s %d %d %d", magic, &w, &h, &maxv) == 4 && strcmp(magic, "P5") == 0 && maxv == 255;
    if (ok) {
        fgetc(f);  // the single whitespace after the header
        px.resize((size_t)w * h);
        ok = fread(px.data(), 1, px.size(), f) == px.size();
    }
    fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char **argv) {
    const int L = 16;
    int W = 96, H = 64, iters = 30, levels = 1, icm_iters = 0;
    const char *refine = "sync";
    std::vector<uint8_t> left, right;
    if (argc > 1 && strcmp(argv[1], "--pgm") == 0) {
        int w2 = 0, h2 = 0;
        if (argc < 6 || !read_pgm(argv[2], W, H, left) || !read_pgm(argv[3], w2, h2, right) || w2 != W || h2 != H) {
            fprintf(stderr, "Error in loading stereo image pair\n");
            return 1;
        }
        iters = atoi(argv[4]);
        levels = atoi(argv[5]);
        if (argc > 6) icm_iters = atoi(argv[6]);
        if (argc > 7) refine = argv[7];
    } else {
        if (argc > 1) W = atoi(argv[1]);
        if (argc > 2) H = atoi(argv[2]);
        if (argc > 3) iters = atoi(argv[3]);
        if (argc > 4) levels = atoi(argv[4]);
        if (argc > 5) icm_iters = atoi(argv[5]);
        if (argc > 6) refine = argv[6];
        left.resize((size_t)W * H);
        right.resize((size_t)W * H);
        if (ginf_gen_stereo(1, W, H, L, 8, 15, 20, left.data(), right.data(), nullptr, nullptr, nullptr) != 0) return 2;
    }
    ginf::Grid<int> grid(W, H, L);
    for (int y = 0; y < grid.getHeight(); y++)
        for (int x = grid.getNumLabels() - 1; x < grid.getWidth(); x++)
            for (int k = 0; k < L; k++) {
                int c = abs((int)left[(size_t)y * W + x] - (int)right[(size_t)y * W + x - k]);
                grid.setDataCost(x, y, k, c < 20 ? c : 20);
            }
    grid.useSmoothnessTruncLinear(10, 20);
    ginf::Matrix<int> result(2, grid.getWidth(), grid.getHeight());
    if (levels > 1) ginf::gpuDecodeHbp(&grid, levels, iters, &result);
    else ginf::gpuDecodeBp(&grid, iters, &result);
    const int e_bp = grid.getLabelingCost(&result);
    int e_icm = e_bp;
    if (icm_iters > 0) {
        ginf::Matrix<int> refined(2, grid.getWidth(), grid.getHeight());
        if (strcmp(refine, "async") == 0) ginf::decodeIcmAsync(&grid, &result, icm_iters, &refined);
        else if (strcmp(refine, "ab") == 0) ginf::decodeAlphaBeta(&grid, &result, icm_iters, &refined);
        else ginf::gpuDecodeIcm(&grid, &result, icm_iters, &refined);
        result.copyFrom(&refined);
        e_icm = grid.getLabelingCost(&result);
    }
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) printf("%d ", result.get(x, y));
        printf("\n");
    }
    ginf::Matrix<int> prior(2, grid.getWidth(), grid.getHeight());
    ginf::decodeMaxPrior(&grid, &prior);
    printf("energy %d %d %d\n", e_bp, e_icm, grid.getLabelingCost(&prior));
    return 0;
}
