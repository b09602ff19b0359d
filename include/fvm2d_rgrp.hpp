// Drop-in C++ front end with the reference's class interface, forwarding to
// the C ABI of libfvm2d.so (include/fvm2d.h).  A caller written against the
// reference's src/wrappers.cuh (e.g. test2D.cu's main) compiles with a host
// compiler and links -lfvm2d; no CUDA/HIP toolchain is needed on its side.
//
//   vec5d              <- src/ghostCellPresets.cuh:4-45 (same 40-byte layout,
//                         +, - and *scalar on slots 0..3, slot 4 -> 0)
//   ghostCellStrategy  <- src/wrappers.cuh:17-24 (same enumerators, same order)
//   RGRP_2D_CUDA       <- src/wrappers.cuh:26-65 (same methods and defaults)
//
// Record mode (setRecordModel, src/wrappers.cuh:147-152,253-261,323-345)
// writes a frame of the density after the first step and every gapPerFrame
// steps, named by the step index, as <index>.png: the min-max normalised
// 8-bit image coloured with viridis, as the reference writes it through
// OpenCV (include/fvm2d_frame.hpp writes the PNG without OpenCV).  Like the
// reference's, a frame holds the state before that step's ghost fill.
#pragma once

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <deque>
#include <filesystem>
#include <fstream>
#include <future>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "fvm2d.h"
#include "fvm2d_frame.hpp"

struct vec5d {
    vec5d() : data{0.0, 0.0, 0.0, 0.0, 0.0} {}
    vec5d(double x0, double x1, double x2, double x3, double x4) : data{x0, x1, x2, x3, x4} {}
    double &operator[](size_t i) { return data[i]; }
    const double &operator[](size_t i) const { return data[i]; }
    double &operator()(size_t i) { return data[i]; }
    const double &operator()(size_t i) const { return data[i]; }
    vec5d operator+(const vec5d &v) const {
        vec5d r;
        for (int i = 0; i < 4; i++) r.data[i] = data[i] + v.data[i];
        return r;
    }
    vec5d operator-(const vec5d &v) const {
        vec5d r;
        for (int i = 0; i < 4; i++) r.data[i] = data[i] - v.data[i];
        return r;
    }
    vec5d operator*(double a) const {
        vec5d r;
        for (int i = 0; i < 4; i++) r.data[i] = data[i] * a;
        return r;
    }
    double data[5];
};
static_assert(sizeof(vec5d) == 5 * sizeof(double), "vec5d must be five packed doubles");

enum class ghostCellStrategy { outflow, reflective, jet, jet800, doubleMach };

class RGRP_2D_CUDA {
  public:
    // Rows of unequal length (undefined in the reference) are refused with
    // std::invalid_argument; an empty U0 gives a 0 x 0 mesh, which solve()
    // refuses.
    explicit RGRP_2D_CUDA(const std::vector<std::vector<vec5d>> &U0) {
        Ny_ = (int)U0.size();
        Nx_ = U0.empty() ? 0 : (int)U0[0].size();
        for (auto &row : U0) {
            if (row.size() != (size_t)Nx_)
                throw std::invalid_argument("RGRP_2D_CUDA: the rows of U0 differ in length");
            for (auto &u : row) U0_.push_back(u);
        }
    }
    RGRP_2D_CUDA() {}

    void setRPInitialData(const vec5d &U1, const vec5d &U2, const vec5d &U3, const vec5d &U4, int Nx,
                          int Ny) {
        U0_.assign((size_t)Nx * Ny, vec5d());
        Nx_ = Nx;
        Ny_ = Ny;
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const bool west = i < Nx / 2, north = j < Ny / 2;
                U0_[(size_t)j * Nx + i] = west ? (north ? U2 : U3) : (north ? U1 : U4);
            }
    }
    void setSpatialLayout(double width, double height) {
        width_ = width;
        height_ = height;
    }
    void setTimeLayout(double endTime, double CFL) {
        endTime_ = endTime;
        CFL_ = CFL;
    }
    void setTimeLayout(double endTime) { endTime_ = endTime; }
    void setCFL(double CFL) { CFL_ = CFL; }
    void setAlpha(double alpha) { alpha_ = alpha; }
    void setGamma(double gamma) { gamma_ = gamma; }
    void setGhostCellStrategy(ghostCellStrategy s) { strategy_ = s; }
    void setRecordModel(bool on_off, int gapPerFrame = 10, const std::string &recordDir = ".") {
        record_ = on_off;
        gap_ = gapPerFrame;
        recordDir_ = recordDir;
    }
    // Extension: FVM2D_MODE_FAST (default) or FVM2D_MODE_EXACT.
    void setArithmeticMode(int mode) { mode_ = mode; }

    std::vector<std::vector<vec5d>> solve() {
        fvm2d_params p;
        fvm2d_params_default(&p);
        p.width = width_;
        p.height = height_;
        p.t_end = endTime_;
        p.cfl = CFL_;
        p.alpha = alpha_;
        p.gamma = gamma_;
        p.strategy = (int)strategy_;
        p.mode = mode_;
        int steps = 0;
        double ms = 0.0;
        if (record_) {
            solve_recording(p, &steps, &ms);
        } else {
            const int rc = fvm2d_solve(reinterpret_cast<double *>(U0_.data()), Nx_, Ny_, &p, &steps, &ms);
            if (rc != FVM2D_OK) throw std::runtime_error(std::string("fvm2d_solve: ") + fvm2d_last_error());
        }
        steps_ += steps;
        ms_ = ms;
        U_ = vectorized();
        return U_;
    }
    double getComputationTime_ms() { return ms_; }
    int getTimeStagesCount() { return steps_; }
    // src/wrappers.cuh:370-397: rho_/p_/u_/v_<std::to_string(t_end)>s.txt in
    // dir, one line per row, each value as the default ostream writes it
    // (printf "%.6g", which std::to_chars in general format with precision
    // 6 reproduces) and a space.  The four files are formatted on four
    // threads.  Like the reference's, a file that cannot be opened is
    // skipped silently.
    void writeFinalResultTo_txt(const std::string &dir) {
        const char *names[4] = {"rho", "p", "u", "v"};
        std::future<void> files[4];
        for (int k = 0; k < 4; k++)
            files[k] = std::async(std::launch::async, [this, k, path = dir + "/" + names[k] + "_" +
                                                                    std::to_string(endTime_) + "s.txt"] {
                std::string text;
                text.reserve(U_.size() * (U_.empty() ? 0 : U_[0].size()) * 10);
                char num[32];
                for (const auto &row : U_) {
                    for (const auto &e : row) {
                        const auto r = std::to_chars(num, num + sizeof(num), e(k), std::chars_format::general, 6);
                        text.append(num, r.ptr);
                        text.push_back(' ');
                    }
                    text.push_back('\n');
                }
                std::ofstream f(path, std::ios::binary);
                f.write(text.data(), (std::streamsize)text.size());
            });
        for (auto &f : files) f.get();
    }
    void clear() {
        U0_.clear();
        U_.clear();
    }

  protected:
    static void check(int rc, const char *what) {
        if (rc != FVM2D_OK) throw std::runtime_error(std::string(what) + ": " + fvm2d_last_error());
    }

    // The time loop on the device in chunks of gap_ steps; after the first
    // step and every gap_ further steps the density of the state as the
    // reference copies it out (after the branch average, before the ghost
    // fill) is written as <recordDir_>/<step index>.png.  Its 8-bit image
    // (the min-max normalisation of src/wrappers.cuh:340) is made on the
    // device (fvm2d_download_frame_gray); the colouring, compression and
    // file write run on worker threads (at most one per hardware thread,
    // up to 16, in flight) while the device steps on.  Like the
    // reference's, the computation time includes writing the frames: it
    // ends when the last one is written.
    void solve_recording(const fvm2d_params &p, int *steps_out, double *ms_out) {
        fvm2d_solver *sv = nullptr;
        check(fvm2d_create(Nx_, Ny_, &p, &sv), "fvm2d_create");
        int steps = 0;
        const size_t max_pending = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::deque<std::future<void>> pending;
        try {
            check(fvm2d_upload(sv, reinterpret_cast<const double *>(U0_.data())), "fvm2d_upload");
            std::filesystem::create_directories(recordDir_);
            const auto t0 = std::chrono::steady_clock::now();
            check(fvm2d_begin(sv), "fvm2d_begin");
            const int gap = std::max(1, gap_);
            int due = 1, done = 0;
            for (;;) {
                check(fvm2d_advance(sv, due - done), "fvm2d_advance");
                double t = 0.0, dt = 0.0;
                check(fvm2d_sync(sv, &steps, &t, &dt), "fvm2d_sync");
                if (steps < due) break;
                done = steps;
                std::vector<unsigned char> gray((size_t)Nx_ * Ny_);
                check(fvm2d_download_frame_gray(sv, gray.data()), "fvm2d_download_frame_gray");
                if (pending.size() >= max_pending) {
                    pending.front().get();
                    pending.pop_front();
                }
                pending.push_back(std::async(std::launch::async,
                                             [path = recordDir_ + "/" + std::to_string(due - 1) + ".png",
                                              gray = std::move(gray), w = Nx_, h = Ny_] {
                                                 fvm2d::write_frame_png(path, w, h, gray.data());
                                             }));
                due += gap;
            }
            while (!pending.empty()) {
                pending.front().get();
                pending.pop_front();
            }
            *ms_out = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            check(fvm2d_download(sv, reinterpret_cast<double *>(U0_.data())), "fvm2d_download");
        } catch (...) {
            pending.clear();  // each future's destructor waits for its write
            fvm2d_destroy(sv);
            throw;
        }
        fvm2d_destroy(sv);
        *steps_out = steps;
    }

    std::vector<std::vector<vec5d>> vectorized() const {
        std::vector<std::vector<vec5d>> r(Ny_);
        for (int j = 0; j < Ny_; ++j)
            r[j].assign(U0_.begin() + (size_t)j * Nx_, U0_.begin() + (size_t)(j + 1) * Nx_);
        return r;
    }
    std::vector<vec5d> U0_;
    std::vector<std::vector<vec5d>> U_;
    double width_ = 1.0, height_ = 1.0, endTime_ = 0.0, CFL_ = 0.5;
    double alpha_ = 1.9, gamma_ = 1.4;
    int Nx_ = 100, Ny_ = 100;
    ghostCellStrategy strategy_ = ghostCellStrategy::outflow;
    bool record_ = false;
    int gap_ = 1;
    std::string recordDir_ = ".";
    double ms_ = 0.0;
    int steps_ = 0;
    int mode_ = FVM2D_MODE_FAST;
};
