// Host driver of the oracle for the address / undefined-behaviour sanitizer
// build (tests/test_sanitizers.py): every entry point of oracle/fvm2d_oracle.h
// on small meshes, sizes that are not multiples of 16, every ghost preset.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../oracle/fvm2d_oracle.h"

static std::vector<double> state(int nx, int ny, unsigned seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> pos(0.2, 2.0), vel(-1.0, 1.0);
    std::vector<double> U((size_t)nx * ny * 5, 0.0);
    for (size_t k = 0; k < (size_t)nx * ny; k++) {
        U[5 * k] = pos(g);
        U[5 * k + 1] = pos(g);
        U[5 * k + 2] = vel(g);
        U[5 * k + 3] = vel(g);
    }
    return U;
}

static int check_finite(const std::vector<double> &a, const char *what) {
    for (double x : a)
        if (!std::isfinite(x)) {
            std::printf("non-finite value in %s\n", what);
            return 1;
        }
    return 0;
}

int main() {
    int bad = 0;
    const int sizes[][2] = {{6, 6}, {37, 23}, {48, 40}, {17, 90}};
    for (auto &sz : sizes) {
        for (int strategy = 0; strategy <= 4; strategy++) {
            std::vector<double> U = state(sz[0], sz[1], 7u + strategy);
            orc_params p = {1.0, 1.0, 1.0, 0.5, 1.9, 1.4, strategy, 4, 0, 0};
            int steps = 0;
            double t = 0.0, ms = 0.0;
            if (orc_solve(U.data(), sz[0], sz[1], &p, 2, &steps, &t, &ms) != 0 || steps != 4) bad++;
            bad += check_finite(U, "orc_solve");
            std::vector<double> V = state(sz[0], sz[1], 11u);
            std::vector<double> frames((size_t)3 * sz[0] * sz[1] * 5);
            int nf = 0, idx[3] = {0, 0, 0};
            if (orc_solve_record(V.data(), sz[0], sz[1], &p, 1, 2, frames.data(), 3, &nf, idx, &steps, &t) != 0)
                bad++;
            if (nf != 2 || idx[0] != 0 || idx[1] != 2) bad++;
            orc_fill(V.data(), sz[0], sz[1], strategy, 1.0, 1.0, 0.01);
            orc_sweep_x(V.data(), sz[0], sz[1], 1.0, 1e-3, 1.9, 1.4, 1);
            std::vector<double> R((size_t)sz[0] * sz[1] * 5);
            orc_rotate(V.data(), R.data(), sz[0], sz[1]);
            bad += check_finite(V, "fill/sweep");
        }
    }
    std::mt19937_64 g(3);
    std::uniform_real_distribution<double> pos(0.2, 3.0), nrm(-2.0, 2.0);
    for (int k = 0; k < 2000; k++) {
        double fl[4] = {pos(g), pos(g), nrm(g) * 3, nrm(g)}, fr[4] = {pos(g), pos(g), nrm(g) * 3, nrm(g)};
        double sl[4] = {nrm(g), nrm(g), nrm(g), nrm(g)}, sr[4] = {nrm(g), nrm(g), nrm(g), nrm(g)};
        double W[5], Wt[5], F[4], c[4], q[4];
        int br = -1;
        orc_grp(fl, fr, sl, sr, 1.4, W, Wt, &br);
        orc_flux(W, Wt, 1e-3, F);
        orc_cons(fl, 1.4, c);
        orc_prim(c, 1.4, q);
        orc_slope(fl, fr, sl, 0.01, 1.9, q);
        if (br < 0 || br > 3 || !std::isfinite(orc_speed(fr, 1.4))) bad++;
    }
    {
        const int nx = 20, ny = 18;
        std::vector<double> U = state(nx, ny, 5u), C((size_t)4 * nx * ny), CN(C.size()), G(C.size()), G0(C.size());
        for (size_t k = 0; k < (size_t)nx * ny; k++) {
            const double r = U[5 * k], p = U[5 * k + 1], u = U[5 * k + 2], v = U[5 * k + 3];
            C[k] = r;
            C[nx * ny + k] = r * u;
            C[2 * nx * ny + k] = r * v;
            C[3 * nx * ny + k] = p / 0.4 + 0.5 * r * (u * u + v * v);
            for (int c = 0; c < 4; c++) G[c * nx * ny + k] = std::sin(1.0 + k + c);
        }
        orc_params p = {1.0, 1.0, 0.0, 0.5, 1.9, 1.4, 0, -1, 0, 0};
        orc_adj_forward(C.data(), nx, ny, &p, 3, CN.data());
        orc_adj_backward(C.data(), nx, ny, &p, 3, G.data(), G0.data());
        bad += check_finite(CN, "adj_forward") + check_finite(G0, "adj_backward");
    }
    std::printf(bad ? "FAILED %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}
