// Detector model: the stage between the forward projector and FDK that the reference leaves to other
// programs (its projection file names, F/txtfile, and the scatter U-Net, Unet_scatter.py:349-423, sit on
// either side of it).  lambda = I0 (primary + scatter) expected quanta per pixel, the scatter estimate (e.g.
// the U-Net's prediction) optional; a Poisson count k drawn from a counter-based generator keyed by
// (seed, flat pixel index), so that a pixel's draw does not depend on the launch shape, the grid or the
// order of execution; output k (raw) or max(k, 1) / I0 (normalised intensity for the filter's -log p,
// zero counts floored at one quantum).
//
// Generator: Philox4x32-10 (J. Salmon et al., SC 2011).  Block b of pixel i's stream is Philox of counter
// (i_lo, i_hi, b, 0) under key (seed_lo, seed_hi); a block gives two 53-bit uniforms (words 0-1, 2-3).
// Sampler: lambda < 10, inversion by sequential search on the first uniform (k capped at 256); lambda >= 10,
// PTRS transformed rejection (W. Hoermann, Insurance: Math. Econ. 12 (1993) 39-45), one block per trial, at
// most 64 trials (acceptance is above 0.9 per trial; the cap only bounds the loop: floor(lambda + 0.5)).
// The rejection test compares log(V alpha^-1 / (a/us^2 + b)) with log of the pmf, k log(lambda) - lambda -
// log k!, with log k! exact below k = 10 (a table of factorials) and Stirling's series to the 1/x^5 term
// above (x = k + 1; the next term is below 2e-12 of log k!).  FP64 throughout; exp and log are the device's
// FP64 functions, so a draw can differ from a host restatement only where a comparison falls within an ulp
// or two of its threshold.
//
// Bound: instruction issue (Philox's 20 integer multiplies per block, the sampler's FP64 square root,
// divisions and logarithms),
// 8-12 B of HBM traffic per pixel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "vcbct_internal.h"

namespace vcbct {

struct Philox4 {
    uint32_t w[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{{c0, c1, c2, c3}};
}

__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

// log k! (k >= 0, integral): the factorial table below 10, Stirling's series above.
__device__ __forceinline__ double log_factorial(double k) {
    if (k < 10.0) {
        const double f[10] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0};
        return log(f[(int)k]);
    }
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 +
           r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0)));
}

__device__ double poisson_draw(double lam, uint64_t i, uint32_t k0, uint32_t k1) {
    if (!(lam > 0.0)) return 0.0;
    const uint32_t i0 = (uint32_t)i, i1 = (uint32_t)(i >> 32);
    if (lam < 10.0) {
        const Philox4 r = philox4x32_10(i0, i1, 0u, 0u, k0, k1);
        const double u = u53(r.w[0], r.w[1]);
        double p = exp(-lam), F = p;
        int k = 0;
        while (u > F && k < 256) {
            ++k;
            p = p * lam / (double)k;
            F = F + p;
        }
        return (double)k;
    }
    const double slam = sqrt(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t t = 0; t < 64; ++t) {
        const Philox4 r = philox4x32_10(i0, i1, t, 0u, k0, k1);
        const double U = u53(r.w[0], r.w[1]) - 0.5;
        const double V = u53(r.w[2], r.w[3]);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;  // the fast acceptance (most trials)
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        // the full test; its constants are formed here, off the fast path (same values)
        const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
        if (log(V * invalpha / (a / (us * us) + b)) <= -lam + k * log(lam) - log_factorial(k)) return k;
    }
    return floor(lam + 0.5);
}

__global__ __launch_bounds__(256) void k_detector_counts(const float* __restrict__ primary,
                                                         const float* __restrict__ scatter, size_t n, double I0,
                                                         uint32_t k0, uint32_t k1, int raw, float* out) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double lam = I0 * ((double)primary[i] + (scatter ? (double)scatter[i] : 0.0));
        const double k = poisson_draw(lam, i, k0, k1);
        out[i] = raw ? (float)k : (float)((k < 1.0 ? 1.0 : k) / I0);
    }
}

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_detector_counts(size_t n, const float* dev_primary, const float* dev_scatter, double I0,
                                     uint64_t seed, int raw, float* dev_out, void* stream) {
    if (n == 0) return 0;
    if (!dev_primary || !dev_out) return fail(VCBCT_EINVAL, "detector_counts: NULL argument");
    if (!(I0 > 0.0) || !std::isfinite(I0)) return fail(VCBCT_EINVAL, "detector_counts: I0 must be positive and finite");
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_detector_counts, dim3(nb), dim3(256), 0, (hipStream_t)stream, dev_primary, dev_scatter, n,
                       I0, (uint32_t)seed, (uint32_t)(seed >> 32), raw ? 1 : 0, dev_out);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

extern "C" int vcbct_detector_counts_host(size_t n, const float* host_primary, const float* host_scatter, double I0,
                                          uint64_t seed, int raw, float* host_out) {
    if (n == 0) return 0;
    if (!host_primary || !host_out) return fail(VCBCT_EINVAL, "detector_counts_host: NULL argument");
    float *dp = nullptr, *ds = nullptr;
    const size_t bytes = n * sizeof(float);
    if (dev_malloc((void**)&dp, bytes) != hipSuccess ||
        (host_scatter && dev_malloc((void**)&ds, bytes) != hipSuccess)) {
        (void)hipFree(dp);
        return fail(VCBCT_ENOMEM, "detector_counts_host: device allocation failed");
    }
    int rc = 0;
    if (hipMemcpy(dp, host_primary, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        (ds && hipMemcpy(ds, host_scatter, bytes, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(VCBCT_EHIP, "detector_counts_host: upload failed");
    if (!rc) rc = vcbct_detector_counts(n, dp, ds, I0, seed, raw, dp, nullptr);
    if (!rc && hipMemcpy(host_out, dp, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(VCBCT_EHIP, "detector_counts_host: download failed");
    (void)hipFree(dp);
    (void)hipFree(ds);
    return rc;
}
