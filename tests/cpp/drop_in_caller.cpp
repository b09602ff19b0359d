// test2D-style caller (test2D.cu:46-80, 135-152) of the drop-in header.
// Usage: drop_in_caller N t_end [mode [record_dir|- gap [txt_dir|- [alpha]]]]
//   -> prints "steps ms rho_sum"; with txt_dir, writeFinalResultTo_txt(txt_dir);
//   alpha defaults to 1.9 (test2D.cu's main uses 2.0 and record_dir ./pngs, gap 18)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "fvm2d_rgrp.hpp"

static std::vector<std::vector<vec5d>> initialize_4shocks(int N) {
    vec5d U1 = {1.5, 1.5, 0.0, 0.0, 0};
    vec5d U2 = {0.5323, 0.3, 1.206, 0.0, 0};
    vec5d U3 = {0.138, 0.029, 1.206, 1.206, 0};
    vec5d U4 = {0.5323, 0.3, 0.0, 1.206, 0};
    std::vector<std::vector<vec5d>> ret;
    for (int i = 0; i < N; i++) {
        std::vector<vec5d> row(N);
        for (int j = 0; j < N; j++) row[j] = (i < N / 5) ? (j < 4 * N / 5 ? U2 : U1) : (j < 4 * N / 5 ? U3 : U4);
        ret.push_back(row);
    }
    return ret;
}

int main(int argc, char **argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 64;
    const double tend = argc > 2 ? std::atof(argv[2]) : 0.05;
    RGRP_2D_CUDA solver(initialize_4shocks(N));
    solver.setSpatialLayout(1.0, 1.0);
    solver.setTimeLayout(tend, 0.5);
    solver.setGhostCellStrategy(ghostCellStrategy::outflow);
    solver.setAlpha(argc > 7 ? std::atof(argv[7]) : 1.9);
    solver.setGamma(1.4);
    if (argc > 3) solver.setArithmeticMode(std::atoi(argv[3]));
    if (argc > 5 && std::strcmp(argv[4], "-") != 0) solver.setRecordModel(true, std::atoi(argv[5]), argv[4]);
    auto U = solver.solve();
    double s = 0.0;
    for (auto &row : U)
        for (auto &c : row) s += c(0);
    std::printf("%d %.3f %.17g\n", solver.getTimeStagesCount(), solver.getComputationTime_ms(), s);
    if (argc > 6 && std::strcmp(argv[6], "-") != 0) solver.writeFinalResultTo_txt(argv[6]);
    solver.clear();
    return 0;
}
