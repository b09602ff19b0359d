// writeFinalResultTo_txt of the C++ drop-in header on a given state, with no
// device: a subclass sets the result the way solve() leaves it.
// Usage: txt_writer <in.bin> <ny> <nx> <t_end> <dir>; in.bin holds ny*nx*5
// doubles ([ny][nx][5]).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "fvm2d_rgrp.hpp"

struct TxtOnly : RGRP_2D_CUDA {
    TxtOnly(const std::vector<double> &v, int ny, int nx, double t_end) {
        U_.assign(ny, std::vector<vec5d>(nx));
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                for (int k = 0; k < 5; k++) U_[j][i](k) = v[((size_t)j * nx + i) * 5 + k];
        setTimeLayout(t_end);
    }
};

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    const int ny = std::atoi(argv[2]), nx = std::atoi(argv[3]);
    std::vector<double> v((size_t)ny * nx * 5);
    std::ifstream f(argv[1], std::ios::binary);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    if (!f) return 3;
    TxtOnly(v, ny, nx, std::atof(argv[4])).writeFinalResultTo_txt(argv[5]);
    std::printf("OK\n");
    return 0;
}
