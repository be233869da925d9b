// Drop-in check of the C++ BP API beyond the stereo example: Grid<int> and
// Grid<float>, every smoothness setter (useSmoothnessPotts,
// useSmoothnessTruncLinear, useSmoothnessTruncQuad, and setSmoothnessCost,
// which switches the grid to an explicit table and the O(L^2) kernel), the
// four decode templates, and releaseBpDeviceState between two calls.
//
//   dropin_models <int|float> <potts|tl|tq|explicit> W H L levels iters costs.bin [table.bin] [entry]
//
// costs.bin holds L*H*W values of the grid's type, layout [l][y][x]; table.bin
// (explicit model only) L*L values, V(a,b) at a + L*b.  The model parameters
// of potts/tl/tq are fixed: Potts(7), TruncLinear(3, 12), TruncQuad(2, 30).
// entry: hbp (decodeHbp, default), gpuhbp, bp (decodeBp, levels ignored) or
// gpubp.  The decode runs twice, with releaseBpDeviceState() in between, and
// the program fails unless both runs give the same labels.  Prints the
// labels, one row per line, then "energy <getLabelingCost>".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bp.h"
#include "core.h"

namespace {

template <typename T>
bool read_values(const char *path, size_t n, std::vector<T> &out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    out.resize(n);
    const bool ok = fread(out.data(), sizeof(T), n, f) == n;
    fclose(f);
    return ok;
}

template <typename T>
void print_energy(T e);
template <>
void print_energy<int>(int e) { printf("energy %d\n", e); }
template <>
void print_energy<float>(float e) { printf("energy %.9g\n", (double)e); }

template <typename T>
int run(int argc, char **argv) {
    const char *model = argv[2];
    const int W = atoi(argv[3]), H = atoi(argv[4]), L = atoi(argv[5]);
    const int levels = atoi(argv[6]), iters = atoi(argv[7]);
    const char *entry = argc > 10 ? argv[10] : "hbp";
    std::vector<T> costs, table;
    if (!read_values(argv[8], (size_t)L * H * W, costs)) {
        fprintf(stderr, "cannot read %s\n", argv[8]);
        return 1;
    }
    ginf::Grid<T> grid(W, H, L);
    for (int l = 0; l < L; ++l)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) grid.setDataCost(x, y, l, costs[((size_t)l * H + y) * W + x]);
    if (strcmp(model, "potts") == 0) {
        grid.useSmoothnessPotts((T)7);
    } else if (strcmp(model, "tl") == 0) {
        grid.useSmoothnessTruncLinear((T)3, (T)12);
    } else if (strcmp(model, "tq") == 0) {
        grid.useSmoothnessTruncQuad((T)2, (T)30);
    } else if (strcmp(model, "explicit") == 0) {
        if (argc < 10 || !read_values(argv[9], (size_t)L * L, table)) {
            fprintf(stderr, "explicit model needs a table file\n");
            return 1;
        }
        for (int b = 0; b < L; ++b)
            for (int a = 0; a < L; ++a) grid.setSmoothnessCost(a, b, table[(size_t)a + (size_t)L * b]);
        if (grid.getSmModel() != ginf::GINF_SM_EXPLICIT) return 3;
    } else {
        fprintf(stderr, "unknown model %s\n", model);
        return 1;
    }
    ginf::Matrix<int> first(2, W, H), second(2, W, H);
    for (int pass = 0; pass < 2; ++pass) {
        ginf::Matrix<int> *r = pass == 0 ? &first : &second;
        if (strcmp(entry, "bp") == 0) ginf::decodeBp(&grid, iters, r);
        else if (strcmp(entry, "gpubp") == 0) ginf::gpuDecodeBp(&grid, iters, r);
        else if (strcmp(entry, "gpuhbp") == 0) ginf::gpuDecodeHbp(&grid, levels, iters, r);
        else ginf::decodeHbp(&grid, levels, iters, r);
        if (pass == 0) ginf::releaseBpDeviceState();
    }
    if (memcmp(first.data, second.data, sizeof(int) * (size_t)W * H) != 0) {
        fprintf(stderr, "labels differ after releaseBpDeviceState\n");
        return 4;
    }
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) printf("%d ", first.get(x, y));
        printf("\n");
    }
    print_energy<T>(grid.getLabelingCost(&first));
    ginf::releaseBpDeviceState();
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: dropin_models <int|float> <potts|tl|tq|explicit> W H L levels iters costs.bin "
                        "[table.bin] [hbp|gpuhbp|bp|gpubp]\n");
        return 1;
    }
    if (strcmp(argv[1], "int") == 0) return run<int>(argc, argv);
    if (strcmp(argv[1], "float") == 0) return run<float>(argc, argv);
    fprintf(stderr, "unknown type %s\n", argv[1]);
    return 1;
}
