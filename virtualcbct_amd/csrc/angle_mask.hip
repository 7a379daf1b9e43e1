// The FDK program's per-voxel angle window (prepareAngles, F/misc.cpp:87-180) and the panel-offset class
// that selects it (F/projectionDataMVkV.cpp:277-281).  The reference computes the [angles][R][R] mask in
// its main program (F/main.cpp:57) and never applies it; it is kept here as a set-up product.
// Also the offset-detector (half-fan) redundancy weighting that such a panel offset needs and the reference
// does not apply (below, k_offset_weight).
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


// Offset-detector (half-fan) weighting.  A 360-degree scan with the panel shifted sideways measures the rays
// near the central ray twice (once from each side) and the rays of the long side once; plain FDK then
// counts the centre of the field twice as often as its rim.  Each projection column gets a weight w(t), t the
// column's distance from the central ray (pixels, positive towards the long side), with
//   w = 0 for t <= -tau,  w = sin^2(pi/4 (1 + t/tau)) for |t| < tau,  w = 1 for t >= tau,
// so that w(t) + w(-t) = 1 for every pair of conjugate rays (G. Wang, Med. Phys. 29 (2002) 1634).  tau is
// the overlap half-width, the distance from the central ray to the short edge (W/2 - |shift|), or a
// narrower taper.  The central ray falls on column c = W/2 + shift - 0.5 (the column rule of the
// backprojector, F/reconstructImage.cpp:83; shift = xshift / (0.05 * 800 / W) pixels).  The weight applies
// to the line integral q = -log p that the filter forms, times 2 so that both conjugates together keep the
// full-scan normalisation of a centred detector: p' = p^(2w), i.e. -log p' = 2 w q.
// The ramp-filtered weighted projection does not vanish beyond the short edge (its negative tails reach
// the voxels that only the long side sees), so the panel is also padded on the short side with P =
// round(2 max|shift|) columns of p = 1 (no attenuation), making it symmetric about the central ray; the
// filter and the backprojector then run on the W + P columns with the same pixel size.
__host__ __device__ inline double offset_exponent(const OffsetPar& q, int col) {
    const double t = q.side * ((double)col - q.c);
    if (t <= -q.tau) return 0.0;
    if (t >= q.tau) return 2.0;
    const double s = sin(0.78539816339744830962 * (1.0 + t / q.tau));
    return 2.0 * (s * s);
}

void offset_exponent_table(const OffsetPar* par, int n, int W, float* out) {
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < W; ++c) out[(size_t)k * W + c] = (float)offset_exponent(par[k], c);
}

constexpr int OW_ROWS = 16;  // rows of one projection per block

// out [n][H][Wo]: column j of the padded panel holds real column j - off (off = P when the padding goes
// before the real columns, i.e. the short side is towards column 0), or 1 in the padding.
__global__ __launch_bounds__(256) void k_offset_weight(const float* __restrict__ p, float* __restrict__ out, int H,
                                                       int W, int Wo, int off, const OffsetPar* __restrict__ par) {
    extern __shared__ double e_col[];  // exponent 2w per real column of this projection
    const int k = blockIdx.y;
    const OffsetPar q = par[k];
    for (int c = threadIdx.x; c < W; c += 256) e_col[c] = offset_exponent(q, c);
    __syncthreads();
    const int j0 = blockIdx.x * OW_ROWS;
    const int nr = min(OW_ROWS, H - j0);
    const float* src = p + ((size_t)k * H + j0) * W;
    float* dst = out + ((size_t)k * H + j0) * Wo;
    // the block's nr x Wo outputs in order, 256 apart; a thread's (row, column) advances by 256 outputs (no
    // division per output)
    int row = threadIdx.x / Wo, oc = threadIdx.x - row * Wo;
    for (int i = threadIdx.x; i < nr * Wo; i += 256) {
        const int col = oc - off;  // real column (outside [0, W): padding)
        float v = 1.0f;
        if (col >= 0 && col < W) {
            const double e = e_col[col];
            const double x = (double)src[(size_t)row * W + col];
            // p^e across the overlap as 2^(e log2 p) in FP64: ~1e-15 relative error, far below the FP32 output's
            v = (float)(e == 2.0 ? x * x : (e == 0.0 ? 1.0 : exp2(e * log2(x))));
        }
        dst[i] = v;
        oc += 256;
        while (oc >= Wo) {
            oc -= Wo;
            ++row;
        }
    }
}

// Padding of a half-fan panel: P = round(2 max |shift|) pixels, and the side it goes on (+1 after the real
// columns, -1 before, 0 none).  Returns false when the shifts change sign or leave no overlap.
static bool half_fan_pad(int W, int n, const double* xshift, int& P, int& side) {
    double mx = 0.0;
    int sg = 0;
    for (int k = 0; k < n; ++k) {
        const double sh = xshift[k] / (0.05 * 800 / W);
        if (!(W / 2.0 - fabs(sh) > 0.0)) return false;
        const int s = sh > 0.0 ? 1 : (sh < 0.0 ? -1 : 0);
        if (s != 0 && sg != 0 && s != sg) return false;
        if (s != 0) sg = s;
        mx = fmax(mx, fabs(sh));
    }
    P = (int)lround(2.0 * mx);
    side = P > 0 ? sg : 0;
    return true;
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

namespace vcbct {

int half_fan_params(int W, int n, const double* xshift, const double* yshift, double taper_px, int& Wo, int& off,
                    OffsetPar* par, double* xs_out, double* ys_out) {
    int P, side;
    if (!half_fan_pad(W, n, xshift, P, side))
        return fail(VCBCT_EINVAL, "half-fan weighting: the shifts change sign or one leaves no overlap");
    Wo = W + P;
    if (Wo > 4096) return fail(VCBCT_EINVAL, "half-fan weighting: padded width %d above 4096", Wo);
    off = side < 0 ? P : 0;
    const double pix = 0.05 * 800 / W, pix_o = 0.05 * 800 / Wo;
    for (int k = 0; k < n; ++k) {
        const double sh = xshift[k] / pix;       // panel shift in pixels
        const double delta = W / 2.0 - fabs(sh);  // central ray to the short edge
        par[k].c = W / 2.0 + sh - 0.5;
        par[k].tau = taper_px > 0.0 ? fmin(taper_px, delta) : delta;
        par[k].side = sh > 0.0 ? -1.0 : 1.0;
        par[k].pad = 0.0;
        // the same central-ray column on the padded panel (Wo/2 + shift' - 0.5 = c + off), and the row
        // centre the filter takes from the panel width (-Wo/2 - yshift' = -W/2 - yshift) kept
        xs_out[k] = (sh + off - P / 2.0) * pix_o;
        ys_out[k] = (yshift[k] / pix - P / 2.0) * pix_o;
    }
    return 0;
}

int launch_offset_weight(const float* in, float* out, int H, int W, int Wo, int off, int n, const OffsetPar* d_par,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_offset_weight, dim3((unsigned)((H + OW_ROWS - 1) / OW_ROWS), (unsigned)n), dim3(256),
                       (size_t)W * sizeof(double), st, in, out, H, W, Wo, off, d_par);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

}  // namespace vcbct

extern "C" int vcbct_fdk_half_fan_width(int W, int n, const double* xshift) {
    if (W < 1 || n < 1 || !xshift) return fail(VCBCT_EINVAL, "fdk_half_fan_width: bad arguments");
    int P, side;
    if (!half_fan_pad(W, n, xshift, P, side))
        return fail(VCBCT_EINVAL, "fdk_half_fan_width: the shifts change sign or one leaves no overlap");
    return W + P;
}

extern "C" int vcbct_fdk_offset_weight(int H, int W, int n, const float* dev_proj, const double* xshift,
                                       const double* yshift, double taper_px, int out_width, float* dev_out,
                                       double* xshift_out, double* yshift_out, void* stream) {
    if (H < 1 || W < 1 || W > 4096 || n < 1 || n > 65535 || !dev_proj || !dev_out || !xshift || !yshift ||
        !xshift_out || !yshift_out)
        return fail(VCBCT_EINVAL, "fdk_offset_weight: bad arguments");
    std::vector<OffsetPar> par(n);
    int Wo, off;
    int rc = half_fan_params(W, n, xshift, yshift, taper_px, Wo, off, par.data(), xshift_out, yshift_out);
    if (rc) return rc;
    if (out_width != Wo)
        return fail(VCBCT_EINVAL, "fdk_offset_weight: out_width %d, the padded panel has %d columns", out_width, Wo);
    if ((const void*)dev_proj == (const void*)dev_out && Wo != W)
        return fail(VCBCT_EINVAL, "fdk_offset_weight: in place only when nothing is padded");
    hipStream_t st = (hipStream_t)stream;
    void* d_par = nullptr;
    if (dev_malloc(&d_par, (size_t)n * sizeof(OffsetPar)) != hipSuccess)
        return fail(VCBCT_ENOMEM, "fdk_offset_weight: device allocation failed");
    hipError_t e = hipMemcpy(d_par, par.data(), (size_t)n * sizeof(OffsetPar), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = launch_offset_weight(dev_proj, dev_out, H, W, Wo, off, n, (const OffsetPar*)d_par, st);
        if (!rc) e = hipStreamSynchronize(st);
    }
    (void)hipFree(d_par);
    if (rc) return rc;
    if (e != hipSuccess) return fail(VCBCT_EHIP, "fdk_offset_weight: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int vcbct_fdk_offset_weight_host(int H, int W, int n, const float* host_proj, const double* xshift,
                                            const double* yshift, double taper_px, int out_width, float* host_out,
                                            double* xshift_out, double* yshift_out) {
    if (H < 1 || W < 1 || n < 1 || out_width < W || !host_proj || !host_out)
        return fail(VCBCT_EINVAL, "fdk_offset_weight_host: bad arguments");
    float *dp = nullptr, *dq = nullptr;
    const size_t in_b = (size_t)n * H * W * sizeof(float), out_b = (size_t)n * H * out_width * sizeof(float);
    if (dev_malloc((void**)&dp, in_b) != hipSuccess || dev_malloc((void**)&dq, out_b) != hipSuccess) {
        (void)hipFree(dp);
        return fail(VCBCT_ENOMEM, "fdk_offset_weight_host: device allocation failed");
    }
    int rc = 0;
    if (hipMemcpy(dp, host_proj, in_b, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(VCBCT_EHIP, "fdk_offset_weight_host: upload failed");
    if (!rc)
        rc = vcbct_fdk_offset_weight(H, W, n, dp, xshift, yshift, taper_px, out_width, dq, xshift_out, yshift_out,
                                     nullptr);
    if (!rc && hipMemcpy(host_out, dq, out_b, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(VCBCT_EHIP, "fdk_offset_weight_host: download failed");
    (void)hipFree(dp);
    (void)hipFree(dq);
    return rc;
}
