// The FDK program's per-voxel angle window (prepareAngles, F/misc.cpp:87-180) and the panel-offset class
// that selects it (F/projectionDataMVkV.cpp:277-281).  The reference computes the [angles][R][R] mask in
// its main program (F/main.cpp:57) and never applies it; it is kept here as a set-up product.
//
// The window of a voxel depends only on its position, so the host computes the R x R window bounds once,
// with the same FP64 arctangent as the reference, and the device only compares every angle against them
// (one byte written per voxel and angle, coalesced along x).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {

__global__ __launch_bounds__(256) void k_angle_mask(const double2* __restrict__ win, const double* __restrict__ ang,
                                                    int n, int RR, uint8_t* __restrict__ mask) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= RR) return;
    const double pi = 3.141592653589793238462643383279;
    const double2 w = win[v];  // (start, end)
    for (int k = 0; k < n; ++k) {
        const double a = ang[k];
        uint8_t m;
        if (w.x > w.y) m = ((a > w.x && a < pi) || (a > -pi && a < w.y)) ? 1 : 0;
        else m = (a > w.x && a < w.y) ? 1 : 0;
        mask[(size_t)k * RR + v] = m;
    }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* __restrict__ p, size_t n, uint8_t v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_fdk_size_class(int n, const double* xshift, char* out) {
    if (n < 1 || !xshift || !out) return fail(VCBCT_EINVAL, "fdk_size_class: bad arguments");
    double xa = 0.0;
    for (int i = 0; i < n; ++i) xa += xshift[i];
    xa = xa / n;
    out[0] = xa < 6 ? 'S' : (xa > 15 ? 'L' : 'M');
    out[1] = 0;
    return 0;
}

extern "C" int vcbct_fdk_angle_mask(int n, const double* angles, int R, double scale, const char* size_class,
                                    uint8_t* mask, void* stream) {
    if (n < 1 || !angles || R < 1 || R > 16384 || !size_class || !mask)
        return fail(VCBCT_EINVAL, "fdk_angle_mask: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const size_t RR = (size_t)R * R;
    if (strcmp(size_class, "S") == 0) {  // a small panel offset keeps every angle
        const unsigned nb = (unsigned)std::min<size_t>((n * RR + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_u8, dim3(nb), dim3(256), 0, st, mask, n * RR, (uint8_t)1);
        VCBCT_HIP(hipGetLastError());
        return 0;
    }
    // window (th1 + 170 deg, th1 + 10 deg) through +-180 deg, th1 the voxel's angle from +y, clockwise
    const double pi = 3.141592653589793238462643383279;
    std::vector<double> win(2 * RR);
    for (int j = 0; j < R; ++j) {
        const double y = -(-R / 2 + j + 0.5) * scale;
        for (int i = 0; i < R; ++i) {
            const double x = (-R / 2 + i + 0.5) * scale;
            double th1 = 0.0;  // every voxel centre has x != 0 and y != 0
            if (y > 0) th1 = -atan(x / y);
            else if (x < 0 && y < 0) th1 = (-atan(x / y) + pi);
            else if (x > 0 && y < 0) th1 = -(atan(x / y) + pi);
            double ts = th1 + pi - 10 * pi / 180;
            double te = th1 + 10 * pi / 180;
            if (ts < -pi) ts = ts + 2 * pi;
            else if (ts > pi) ts = ts - 2 * pi;
            if (te > pi) te = te - 2 * pi;
            else if (te < -pi) te = te + 2 * pi;
            win[2 * ((size_t)j * R + i)] = ts;
            win[2 * ((size_t)j * R + i) + 1] = te;
        }
    }
    void *d_win = nullptr, *d_ang = nullptr;
    if (dev_malloc(&d_win, win.size() * sizeof(double)) != hipSuccess ||
        dev_malloc(&d_ang, (size_t)n * sizeof(double)) != hipSuccess) {
        (void)hipFree(d_win);
        return fail(VCBCT_ENOMEM, "fdk_angle_mask: device allocation failed");
    }
    int rc = 0;
    hipError_t e = hipMemcpy(d_win, win.data(), win.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ang, angles, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_angle_mask, dim3((unsigned)((RR + 255) / 256)), dim3(256), 0, st, (const double2*)d_win,
                           (const double*)d_ang, n, (int)RR, mask);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) rc = fail(VCBCT_EHIP, "fdk_angle_mask: %s", hipGetErrorString(e));
    (void)hipFree(d_win);
    (void)hipFree(d_ang);
    return rc;
}
