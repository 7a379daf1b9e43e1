// Device-side Adam step for the decomposition loop (no reference counterpart): the maps, the gradient and the two
// moment buffers are updated in place by one elementwise kernel, with the step count kept on the device, so a
// whole iteration (loss and gradient, then this update) can be captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "vcbct_internal.h"

namespace vcbct {

// t <- t + 1 (one thread); the update kernel then reads the new t for its bias corrections.
__global__ void k_adam_tick(int* __restrict__ t) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *t += 1;
}

// m <- b1 m + (1 - b1) g;  v <- b2 v + (1 - b2) g^2;  x <- x - lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps),
// then x <- max(x, lo) when lo is finite (the decomposition keeps densities non-negative).  The moments and the
// update in FP32 (torch.optim.Adam's order of operations), the bias corrections in FP64 once per block.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ x, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, size_t n,
                                              const int* __restrict__ tp, float lr, float b1, float b2, float eps,
                                              float lo) {
    __shared__ float corr[2];
    if (threadIdx.x == 0) {
        const int t = *tp;
        corr[0] = (float)(1.0 - pow((double)b1, (double)t));
        corr[1] = (float)sqrt(1.0 - pow((double)b2, (double)t));
    }
    __syncthreads();
    const float bc1 = corr[0], bc2s = corr[1];
    const float step = lr / bc1;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
        const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
        m[i] = mi;
        v[i] = vi;
        const float denom = sqrtf(vi) / bc2s + eps;
        float xi = x[i] - step * (mi / denom);
        if (lo > -INFINITY) xi = fmaxf(xi, lo);
        x[i] = xi;
    }
}

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_adam_step(size_t n, float* dev_x, const float* dev_grad, float* dev_m, float* dev_v,
                               int* dev_step, double lr, double beta1, double beta2, double eps, double lower,
                               void* stream) {
    if (n == 0) return 0;
    if (!dev_x || !dev_grad || !dev_m || !dev_v || !dev_step) return fail(VCBCT_EINVAL, "adam_step: NULL argument");
    if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return fail(VCBCT_EINVAL, "adam_step: bad hyper-parameters");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_adam_tick, dim3(1), dim3(64), 0, st, dev_step);
    VCBCT_HIP(hipGetLastError());
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_adam, dim3(nb), dim3(256), 0, st, dev_x, dev_grad, dev_m, dev_v, n, dev_step, (float)lr,
                       (float)beta1, (float)beta2, (float)eps, std::isfinite(lower) ? (float)lower : -INFINITY);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
