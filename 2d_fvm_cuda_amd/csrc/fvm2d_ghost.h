// Ghost-cell presets of the reference (src/ghostCellPresets.cuh:47-241)
// expressed as a per-cell source map: after the preset's two-phase fill
// (all ghost-row writes, then all ghost-column writes, which is what the
// reference's warps execute), every ghost cell holds either a constant
// state or a copy of one interior cell, possibly with v negated.
#pragma once

#include <hip/hip_runtime.h>

#include "fvm2d_device.h"

namespace fvm2d {

struct GhostSrc {
    int is_const;
    int sj, si;
    int negv;
    double r, p, u, v;
};

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Value of ghost cell (gj, gi) after the preset's two-phase fill, as a copy
// of an interior cell (optionally with v negated) or a constant state.
__device__ GhostSrc ghost_source(int strategy, int gj, int gi, int nx, int ny, double t) {
    GhostSrc g{0, clampi(gj, 2, ny - 3), clampi(gi, 2, nx - 3), 0, 0.0, 0.0, 0.0, 0.0};
    auto set_const = [&](double r, double p, double u, double v) {
        g.is_const = 1;
        g.r = r;
        g.p = p;
        g.u = u;
        g.v = v;
    };
    if (strategy == kOutflow) return g;  // ghostCellPresets.cuh:106-145
    if (strategy == kJet) {              // ghostCellPresets.cuh:148-193
        if (gi <= 1) {
            if (gj >= 2 * ny / 5 + 1 && gj <= 3 * ny / 5) set_const(5.0, 0.4127, 800.0, 0.0);
            else set_const(0.5, 0.4127, 0.0, 0.0);
        }
        return g;
    }
    if (strategy == kJet800) {  // ghostCellPresets.cuh:196-241
        if (gi <= 1) {
            if (gj >= 9 * ny / 10) set_const(1.4, 1.0, 800.0, 0.0);
            else set_const(0.14, 1.0, 0.0, 0.0);
        } else if (gj == ny - 1) {
            g.sj = ny - 4;
        }
        return g;
    }
    // double Mach (also selected by "reflective"), ghostCellPresets.cuh:48-103
    const double post_u = 4.125 * sqrt(3.0);
    if (gi <= 1) {
        set_const(8.0, 116.5, post_u, -4.125);
    } else if (gi >= nx - 2) {
        set_const(1.4, 1.0, 0.0, 0.0);
    } else if (gj <= 1) {
        double temp = double(ny) / sqrt(3.0);
        double dist = double(ny) * 20 * t / sqrt(3.0);
        if (gi <= nx / 24 + int(temp) + int(dist)) set_const(8.0, 116.5, post_u, -4.125);
        else set_const(1.4, 1.0, 0.0, 0.0);
    } else if (gj == ny - 2) {
        g.sj = ny - 3;
        g.negv = gi > nx / 24;
    } else {  // gj == ny - 1
        if (gi <= nx / 24) {
            g.sj = ny - 3;
        } else {
            g.sj = ny - 4;
            g.negv = 1;
        }
    }
    return g;
}

}  // namespace fvm2d
