// Host driver of the engine's launch-layout logic and parameter checks
// (2d_fvm_cuda_amd/csrc/fvm2d_engine.h) for the address / undefined-behaviour
// sanitizer build (tests/test_sanitizers.py): for many mesh sizes, strip
// heights, block shapes and device sizes, every window's interior rows are
// covered exactly once by the task list.
#include <cstdio>
#include <vector>

#include "../../2d_fvm_cuda_amd/csrc/fvm2d_engine.h"

using fvm2d::Engine;

int main() {
    int bad = 0, cases = 0;
    const int nxs[] = {6, 7, 64, 65, 127, 300, 2000};
    const int nys[] = {6, 9, 40, 211, 1000, 2000};
    const int Rs[] = {1, 2, 3, 7, 16, 34, 64, 200};
    for (int nx : nxs)
        for (int ny : nys)
            for (int R : Rs)
                for (int wpb = 1; wpb <= 4; wpb++)
                    for (int pers = 0; pers <= 1; pers++)
                        for (int cus : {1, 8, 256})
                            for (int bpc : {1, 2}) {
                                const int nw = (nx - 4 + fvm2d::kOwn - 1) / fvm2d::kOwn;
                                const long long slots = (long long)bpc * wpb * cus;
                                std::vector<int4> tasks;
                                Engine::mesh_strips(0, nw, ny, R, pers == 0, slots, 4LL * cus, wpb, tasks);
                                if (pers == 0) Engine::mirror_second_round(tasks, 4LL * cus);
                                std::vector<int> cover((size_t)nw * ny, 0);
                                for (const int4 &t : tasks) {
                                    if (t.y < 0 || t.y >= nw || t.z < 2 || t.w > ny - 2 || t.z >= t.w) {
                                        bad++;
                                        continue;
                                    }
                                    for (int j = t.z; j < t.w; j++) cover[(size_t)t.y * ny + j]++;
                                }
                                for (int w = 0; w < nw; w++)
                                    for (int j = 2; j < ny - 2; j++) bad += cover[(size_t)w * ny + j] != 1;
                                cases++;
                            }
    fvm2d_params p;
    p.width = p.height = 1.0;
    p.t_end = 0.1;
    p.cfl = 0.5;
    p.alpha = 1.9;
    p.gamma = 1.4;
    p.strategy = 0;
    p.max_steps = -1;
    p.mode = 0;
    p.reserved = 0;
    bad += fvm2d::check_params(64, 64, &p) != FVM2D_OK;
    bad += fvm2d::check_params(5, 64, &p) != FVM2D_ERR_ARG;
    bad += fvm2d::check_params(1 << 16, 1 << 16, &p) != FVM2D_ERR_ARG;
    p.gamma = 1.0;
    bad += fvm2d::check_params(64, 64, &p) != FVM2D_ERR_ARG;
    bad += fvm2d::check_params(64, 64, nullptr) != FVM2D_ERR_ARG;
    std::printf(bad ? "FAILED %d of %d\n" : "OK %d layouts\n", bad ? bad : cases, cases);
    return bad ? 1 : 0;
}
