// Measurement probes of libfvm2d_probes.so (include/fvm2d_probes.h): HBM
// counter calibration, FP64 VALU throughput and the accuracy of the device
// reciprocals the FAST arithmetic uses.  Diagnostics only; the solver
// library libfvm2d.so neither contains nor calls them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/fvm2d_probes.h"
#include "fvm2d_cell.h"

using namespace fvm2d;

namespace {

thread_local std::string g_probe_error;

int fail(int code, const std::string &msg) {
    g_probe_error = msg;
    return code;
}

int device_ok() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(FVM2D_PROBE_ERR_NODEVICE, "no HIP device");
    return 0;
}

}  // namespace

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FVM2D_PROBE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Plain device-to-device copy of a known size, used to calibrate the HBM
// byte counters (FETCH_SIZE / WRITE_SIZE) against the step kernel's.
__global__ void __launch_bounds__(256) hbm_copy_kernel(const double2 *__restrict__ src,
                                                       double2 *__restrict__ dst, long long n2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += stride) dst[k] = src[k];
}

extern "C" {

const char *fvm2d_probe_last_error(void) { return g_probe_error.c_str(); }

int fvm2d_hbm_copy_probe(long long n_doubles, int reps, double *avg_ms_out) {
    if (n_doubles < 2 || (n_doubles & 1) || reps < 1) return fail(FVM2D_PROBE_ERR_ARG, "need an even n_doubles >= 2 and reps >= 1");
    if (int rc0 = device_ok()) return rc0;
    double *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, n_doubles * sizeof(double)));
    if (hipMalloc((void **)&b, n_doubles * sizeof(double)) != hipSuccess) {
        hipFree(a);
        return fail(FVM2D_PROBE_ERR_HIP, "hipMalloc failed");
    }
    hipEvent_t e0, e1;
    int rc = 0;
    if (hipMemset(a, 0, n_doubles * sizeof(double)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(FVM2D_PROBE_ERR_HIP, "probe setup failed");
    } else {
        const long long n2 = n_doubles / 2;
        const int grid = (int)std::min<long long>((n2 + 255) / 256, 256LL * 8 * 4);
        hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; r++) hbm_copy_kernel<<<grid, 256, 0, nullptr>>>((const double2 *)a, (double2 *)b, n2);
        hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(FVM2D_PROBE_ERR_HIP, "copy probe failed");
        } else {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            if (avg_ms_out) *avg_ms_out = ms / reps;
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    }
    hipFree(a);
    hipFree(b);
    return rc;
}

}  // extern "C"

// VALU throughput probe: 16 independent chains per lane of one instruction
// kind, 2048 blocks of 256 threads (8 waves per CU), `iters` iterations.
template <int OP>
__global__ void __launch_bounds__(256) valu_probe_kernel(double *out, double b, double c, int iters) {
    double a[16];
#pragma unroll
    for (int k = 0; k < 16; k++) a[k] = threadIdx.x * 1e-3 + k;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (OP == 0) a[k] = fma(a[k], b, c);
            if (OP == 1) a[k] = a[k] * b;
            if (OP == 2) a[k] = a[k] + c;
            if (OP == 3) a[k] = __builtin_amdgcn_rcp(a[k]);
            if (OP == 4) a[k] = __hiloint2double(__double2hiint(a[k]) ^ 0x10, __double2loint(a[k]) + 3);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; k++) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

extern "C" {

int fvm2d_valu_probe(int kind, double *per_simd_per_s) {
    if (kind < 0 || kind > 4) return fail(FVM2D_PROBE_ERR_ARG, "probe kind must be 0..4");
    if (int rc0 = device_ok()) return rc0;
    const int grid = 2048, block = 256, iters = 4096;
    double *out = nullptr;
    HIPCHK(hipMalloc((void **)&out, (size_t)grid * block * sizeof(double)));
    hipEvent_t e0, e1;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        hipFree(out);
        return fail(FVM2D_PROBE_ERR_HIP, "probe setup failed");
    }
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    float ms = 0.f;
    for (int rep = 0; rep < 2 && !rc; rep++) {
        hipEventRecord(e0, nullptr);
        switch (kind) {
            case 0: valu_probe_kernel<0><<<grid, block>>>(out, 1.0000001, 1e-9, iters); break;
            case 1: valu_probe_kernel<1><<<grid, block>>>(out, 1.0000001, 1e-9, iters); break;
            case 2: valu_probe_kernel<2><<<grid, block>>>(out, 1.0000001, 1e-9, iters); break;
            case 3: valu_probe_kernel<3><<<grid, block>>>(out, 1.0000001, 1e-9, iters); break;
            default: valu_probe_kernel<4><<<grid, block>>>(out, 1.0000001, 1e-9, iters); break;
        }
        hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FVM2D_PROBE_ERR_HIP, "probe failed");
        hipEventElapsedTime(&ms, e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(out);
    if (!rc && per_simd_per_s) {
        // wave-level instructions of the probed kind (kind 4: two 32-bit ops per chain step)
        const double waves = (double)grid * block / 64.0;
        const double ops = waves * iters * 16.0 * (kind == 4 ? 2.0 : 1.0);
        *per_simd_per_s = ops / (ms * 1e-3) / (4.0 * cus);
    }
    return rc;
}

}  // extern "C"

// Accuracy of the fast path's reciprocal and reciprocal square root: raw
// v_rcp_f64 / v_rsq_f64 and the refined rcp() / rsq() of fvm2d_cell.h.
__global__ void recip_probe_kernel(const double *x, double *out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i];
    out[4 * i + 0] = __builtin_amdgcn_rcp(a);
    out[4 * i + 1] = rcp(a);
    out[4 * i + 2] = __builtin_amdgcn_rsq(a);
    out[4 * i + 3] = rsq(a);
}

extern "C" {

int fvm2d_recip_probe(double *max_rel_err4) {
    if (!max_rel_err4) return fail(FVM2D_PROBE_ERR_ARG, "null output");
    if (int rc0 = device_ok()) return rc0;
    const int n = 1 << 18;
    std::vector<double> x(n), o(4 * (size_t)n);
    unsigned long long st = 0x9e3779b97f4a7c15ull;
    for (int i = 0; i < n; i++) {  // 2^u, u uniform in [-8, 8) (splitmix64)
        st += 0x9e3779b97f4a7c15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[i] = std::exp2(-8.0 + 16.0 * (double)(z >> 11) * 0x1.0p-53);
    }
    double *dx = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc((void **)&dx, n * sizeof(double)));
    if (hipMalloc((void **)&dout, 4 * (size_t)n * sizeof(double)) != hipSuccess) {
        hipFree(dx);
        return fail(FVM2D_PROBE_ERR_HIP, "hipMalloc failed");
    }
    int rc = 0;
    if (hipMemcpy(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(FVM2D_PROBE_ERR_HIP, "copy failed");
    } else {
        recip_probe_kernel<<<(n + 255) / 256, 256>>>(dx, dout, n);
        if (hipMemcpy(o.data(), dout, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(FVM2D_PROBE_ERR_HIP, "recip probe failed");
    }
    hipFree(dx);
    hipFree(dout);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) max_rel_err4[k] = 0.0;
    for (int i = 0; i < n; i++) {
        const long double r = 1.0L / (long double)x[i], q = 1.0L / sqrtl((long double)x[i]);
        const long double ref[4] = {r, r, q, q};
        for (int k = 0; k < 4; k++)
            max_rel_err4[k] = std::max(max_rel_err4[k], (double)fabsl((o[4 * (size_t)i + k] - ref[k]) / ref[k]));
    }
    return 0;
}

}  // extern "C"
