// Argument errors of the C++ drop-in header that need no device: rows of
// unequal length refused by the constructor, an empty mesh refused by
// solve() (fvm2d_solve rejects the null state before any HIP call).
// Prints "OK" when every case raises the expected exception.
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "fvm2d_rgrp.hpp"

int main() {
    std::vector<std::vector<vec5d>> ragged(8, std::vector<vec5d>(8, vec5d(1.0, 1.0, 0.0, 0.0, 0.0)));
    ragged[5].pop_back();
    bool ok = false;
    try {
        RGRP_2D_CUDA s(ragged);
    } catch (const std::invalid_argument &) {
        ok = true;
    }
    if (!ok) {
        std::printf("ragged rows accepted\n");
        return 1;
    }
    ok = false;
    try {
        RGRP_2D_CUDA s(std::vector<std::vector<vec5d>>{});
        s.setTimeLayout(0.1, 0.5);
        s.solve();
    } catch (const std::runtime_error &) {
        ok = true;
    }
    if (!ok) {
        std::printf("empty mesh accepted\n");
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
