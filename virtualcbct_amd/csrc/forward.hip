// Forward projector kernels: volume re-layout, energy-collapsed polychromatic projection, line
// integrals, log projection, and the adjoint (transposed traversal) — gfx950.
//
// Replaces reprojectionKernel and the host energy loop of reprojectionOnGPU
// (P/reprojectionOnGPU.cu:56-690, 1004-1056): the reference recomputes the whole traversal once per
// energy bin and copies every projection to the host after each bin; here one traversal per ray yields
// the material line integrals L_m, and the spectral sum runs in the same kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include <climits>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>
#include <cmath>

#include "traverse.h"
#include "vcbct_internal.h"

namespace vcbct {

// Occupancy of the ray kernels.  Their step loops wait on dependent LDS reads, scalar broadcasts and
// gathers, and more resident waves hide that better than spilling a few registers costs: on the
// 512x512x128 / 360-angle workload, k_forward takes 47.1 ms at its natural 75 VGPRs (6 waves per SIMD),
// 45.0 ms held to 7 waves and 43.1 ms held to 8 (64 VGPRs, 56 bytes per lane spilled to scratch).  The
// persistent batch kernel is fastest held to 7.  With the later step loop (76 VGPRs unbounded) the intensity
// form takes 22.4 ms at 6 waves and 21.1 ms at both 7 and 8, while the log projection of the 256x256x64 /
// 90-angle workload takes 1.58 ms at 7 (72 VGPRs, 28 bytes spilled) against 1.74 ms at 8: the projection
// forms other than the intensity are held to 7.
#define VCBCT_WAVES8 __attribute__((amdgpu_waves_per_eu(8, 8)))
#define VCBCT_WAVES7 __attribute__((amdgpu_waves_per_eu(7, 8)))
#define VCBCT_WAVES_FWD(MODE) __attribute__((amdgpu_waves_per_eu((MODE) == VCBCT_MODE_INTENSITY ? 8 : 7, 8)))

// ------------------------------------------------------------------------------------------------
// Volume re-layout.  Reference u16 [z][mat][y][x] -> device [y][x][z][8] (16-byte voxel, all materials;
// the 64 rows of a wave read one contiguous run of a z column).  Tiled through LDS: a block reads a
// 32(x) x 32(z) tile of every material plane coalesced along x and writes 32 z-runs of 16-byte voxels.
constexpr int TT = 32;

__global__ __launch_bounds__(256) void k_transpose_u16(const uint16_t* __restrict__ src, uint4* __restrict__ dst,
                                                       int N, int Nz, int nmat) {
    __shared__ uint16_t tile[VCBCT_MAX_MAT][TT][TT + 2];
    const int x0 = blockIdx.x * TT, z0 = blockIdx.y * TT, y = blockIdx.z;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8
    const size_t n2 = (size_t)N * N;
    for (int m = 0; m < nmat; ++m)
        for (int zz = ty; zz < TT; zz += 8) {
            int z = z0 + zz, x = x0 + tx;
            uint16_t v = 0;
            if (z < Nz && x < N) v = src[((size_t)z * nmat + m) * n2 + (size_t)y * N + x];
            tile[m][zz][tx] = v;
        }
    __syncthreads();
    // write: thread -> (x = x0 + i / TT, z = z0 + i % TT), z fastest
    for (int i = threadIdx.x; i < TT * TT; i += 256) {
        int xx = i / TT, zz = i % TT;
        int x = x0 + xx, z = z0 + zz;
        if (x >= N || z >= Nz) continue;
        uint16_t w[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) w[m] = (m < nmat) ? tile[m][zz][xx] : (uint16_t)0;
        uint4 o;
        o.x = (uint32_t)w[0] | ((uint32_t)w[1] << 16);
        o.y = (uint32_t)w[2] | ((uint32_t)w[3] << 16);
        o.z = (uint32_t)w[4] | ((uint32_t)w[5] << 16);
        o.w = (uint32_t)w[6] | ((uint32_t)w[7] << 16);
        dst[((size_t)y * N + x) * Nz + z] = o;
    }
}

// FP32 maps [mat][z][y][x] -> [y][x][z][8] float (two float4 per voxel).
__global__ __launch_bounds__(256) void k_transpose_f32(const float* __restrict__ src, float4* __restrict__ dst, int N,
                                                       int Nz, int nmat) {
    __shared__ float tile[VCBCT_MAX_MAT][TT][TT + 1];
    const int x0 = blockIdx.x * TT, z0 = blockIdx.y * TT, y = blockIdx.z;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const size_t n2 = (size_t)N * N, n3 = n2 * Nz;
    for (int m = 0; m < nmat; ++m)
        for (int zz = ty; zz < TT; zz += 8) {
            int z = z0 + zz, x = x0 + tx;
            float v = 0.f;
            if (z < Nz && x < N) v = src[(size_t)m * n3 + (size_t)z * n2 + (size_t)y * N + x];
            tile[m][zz][tx] = v;
        }
    __syncthreads();
    for (int i = threadIdx.x; i < TT * TT; i += 256) {
        int xx = i / TT, zz = i % TT;
        int x = x0 + xx, z = z0 + zz;
        if (x >= N || z >= Nz) continue;
        float w[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) w[m] = (m < nmat) ? tile[m][zz][xx] : 0.f;
        size_t o = (((size_t)y * N + x) * Nz + z) * 2;
        dst[o] = make_float4(w[0], w[1], w[2], w[3]);
        dst[o + 1] = make_float4(w[4], w[5], w[6], w[7]);
    }
}

// [y][x][z][8] float -> [mat][z][y][x]
// Gradient maps from the scatter layout [mat][y][x][z] back to the caller's [mat][z][y][x] (LDS-tiled
// transpose of each (y, material) plane: z-runs in, x-runs out).
__global__ __launch_bounds__(256) void k_untranspose_f32(const float* __restrict__ src, float* __restrict__ dst, int N,
                                                         int Nz, int nmat) {
    __shared__ float tile[TT][TT + 1];
    const int x0 = blockIdx.x * TT, z0 = blockIdx.y * TT;
    const int y = blockIdx.z % N, m = blockIdx.z / N;
    const size_t n2 = (size_t)N * N, n3 = n2 * Nz;
    const float* s = src + (size_t)m * n3 + (size_t)y * N * Nz;
    for (int i = threadIdx.x; i < TT * TT; i += 256) {
        int xx = i / TT, zz = i % TT;
        int x = x0 + xx, z = z0 + zz;
        tile[zz][xx] = (x < N && z < Nz) ? s[(size_t)x * Nz + z] : 0.f;
    }
    __syncthreads();
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    for (int zz = ty; zz < TT; zz += 8) {
        int z = z0 + zz, x = x0 + tx;
        if (z < Nz && x < N) dst[(size_t)m * n3 + (size_t)z * n2 + (size_t)y * N + x] = tile[zz][tx];
    }
}

// The adjoint's channel planes [nch][y][x][z] -> the maps [nmat][z][y][x], grad_m = sum_k b[m][k] G_k: a block
// stages a 32 (x) x 32 (z) tile of every channel of one y in LDS (z-runs in), then writes each map's tile
// (x-runs out).  Without a spectral basis the mix is the identity (nch = nmat).
__global__ __launch_bounds__(256) void k_untranspose_mix(const float* __restrict__ src, float* __restrict__ dst, int N,
                                                         int Nz, int nmat, ChannelMix mix) {
    __shared__ float tile[VCBCT_MAX_MAT][TT][TT + 1];
    const int x0 = blockIdx.x * TT, z0 = blockIdx.y * TT, y = blockIdx.z;
    const size_t n2 = (size_t)N * N, n3 = n2 * Nz;
    const int nch = mix.nch;
    for (int k = 0; k < nch; ++k) {
        const float* s = src + (size_t)k * n3 + (size_t)y * N * Nz;
        for (int i = threadIdx.x; i < TT * TT; i += 256) {
            const int xx = i / TT, zz = i % TT;
            const int x = x0 + xx, z = z0 + zz;
            tile[k][zz][xx] = (x < N && z < Nz) ? s[(size_t)x * Nz + z] : 0.f;
        }
    }
    __syncthreads();
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    for (int zz = ty; zz < TT; zz += 8) {
        const int z = z0 + zz, x = x0 + tx;
        if (z >= Nz || x >= N) continue;
        float g[VCBCT_MAX_MAT];
#pragma unroll
        for (int k = 0; k < VCBCT_MAX_MAT; ++k) g[k] = k < nch ? tile[k][zz][tx] : 0.f;
        for (int m = 0; m < nmat; ++m) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < VCBCT_MAX_MAT; ++k) v = fmaf(mix.b[m][k], g[k], v);
            dst[(size_t)m * n3 + (size_t)z * n2 + (size_t)y * N + x] = v;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Voxel access

// Six material line integrals of one ray, sum_seg w_m * rr, accumulated in FP32 (a ray has a few hundred
// segments; the relative rounding of the sums stays near 1e-6).  Voxel layout [y][x][z][8]: u16 maps are
// one 16-byte record per voxel, FP32 maps two.  The gathers of a plane step are all issued before the
// first of them is consumed, so a wave keeps up to three in flight.
template <class VoxT>
struct AccumL;

// u16 maps: the first gather of a step is software-pipelined.  A step adds the previous step's, which has had a
// whole step of arithmetic to arrive, then issues its own (lanes that charge nothing read voxel 0, so the load
// is unconditional); finish() adds the last.
// (Pipelining the second slot too costs registers down to 5 waves per SIMD and was measured slower.)
template <>
struct AccumL<uint4> {
    const uint4* __restrict__ vox;
    int Nz;
    float L[VCBCT_MAX_MAT];
    uint3 p0 = make_uint3(0, 0, 0);  // the previous step's first gather
    float r0 = 0.f;                  // and its length (0: nothing charged)
    __device__ __forceinline__ void step(const Seg& s0, const Seg& s1, const Seg& s2) {
#pragma clang fp contract(fast)
        add(p0, r0);  // the previous step's first gather, issued a step ago
        uint3 q1, q2;
        if (s1.ok) q1 = ld(s1.cell);
        if (s2.ok) q2 = ld(s2.cell);
        // issued into the register it is added from in the next step, after that register's add: no copy at
        // the step's end, whose wait would hold the wave until this gather arrives (measured 4 % faster)
        p0 = ld(s0.ok ? s0.cell : 0);
        r0 = s0.ok ? (float)s0.rr : 0.f;
        if (s1.ok) add(q1, (float)s1.rr);
        if (s2.ok) add(q2, (float)s2.rr);
    }
    __device__ __forceinline__ void finish() {
        add(p0, r0);
        r0 = 0.f;
    }
    __device__ __forceinline__ uint3 ld(int cell) const {  // (cells are nonnegative)
        const uint4 q = ldg(vox, (size_t)(unsigned)cell);
        return make_uint3(q.x, q.y, q.z);
    }
    __device__ __forceinline__ void add(const uint3& q, float rr) {
#pragma clang fp contract(fast)
        L[0] += (float)(q.x & 0xffffu) * rr;
        L[1] += (float)(q.x >> 16) * rr;
        L[2] += (float)(q.y & 0xffffu) * rr;
        L[3] += (float)(q.y >> 16) * rr;
        L[4] += (float)(q.z & 0xffffu) * rr;
        L[5] += (float)(q.z >> 16) * rr;
    }
};

template <>
struct AccumL<float4> {
    const float4* __restrict__ vox;
    int Nz;
    float L[VCBCT_MAX_MAT];
    struct Q {
        float4 a;
        float2 b;
    };
    __device__ __forceinline__ void step(const Seg& s0, const Seg& s1, const Seg& s2) {
        Q q0, q1, q2;  // read only where the matching segment is charged
        if (s0.ok) q0 = ld(s0);
        if (s1.ok) q1 = ld(s1);
        if (s2.ok) q2 = ld(s2);
        if (s0.ok) add(q0, (float)s0.rr);
        if (s1.ok) add(q1, (float)s1.rr);
        if (s2.ok) add(q2, (float)s2.rr);
    }
    __device__ __forceinline__ void finish() {}
    __device__ __forceinline__ Q ld(const Seg& s) const {
        const size_t i = 2 * (size_t)(unsigned)s.cell;
        const float4 b = ldg(vox, i + 1);
        return Q{ldg(vox, i), make_float2(b.x, b.y)};
    }
    __device__ __forceinline__ void add(const Q& q, float rr) {
#pragma clang fp contract(fast)
        L[0] += q.a.x * rr;
        L[1] += q.a.y * rr;
        L[2] += q.a.z * rr;
        L[3] += q.a.w * rr;
        L[4] += q.b.x * rr;
        L[5] += q.b.y * rr;
    }
};

// Channel volumes: the spectral channels c_k = sum_m B_mk rho_m of the maps in FP32, two (float2, 8 bytes) or
// four (Ch4, 16 bytes) to a voxel record; L[k] accumulates the channel line integrals (L[k] = 0 beyond the
// channel count).  The first gather of a step is software-pipelined as for the u16 maps.  With the identity
// basis (channels = maps) the channels of a u16 volume are its values exactly, and every product and sum is the
// u16 path's, in its order, bit for bit.
struct alignas(16) Ch4 {
    float x, y, z, w;
};

template <>
struct AccumL<float2> {
    const float2* __restrict__ vox;
    int Nz;
    float L[VCBCT_MAX_MAT];
    float2 p0 = make_float2(0.f, 0.f);  // the previous step's first gather
    float r0 = 0.f;                     // and its length (0: nothing charged)
    __device__ __forceinline__ void step(const Seg& s0, const Seg& s1, const Seg& s2) {
        add(p0, r0);
        float2 q1, q2;
        if (s1.ok) q1 = ldg(vox, (size_t)(unsigned)s1.cell);
        if (s2.ok) q2 = ldg(vox, (size_t)(unsigned)s2.cell);
        p0 = ldg(vox, (size_t)(unsigned)(s0.ok ? s0.cell : 0));
        r0 = s0.ok ? (float)s0.rr : 0.f;
        if (s1.ok) add(q1, (float)s1.rr);
        if (s2.ok) add(q2, (float)s2.rr);
    }
    __device__ __forceinline__ void finish() {
        add(p0, r0);
        r0 = 0.f;
    }
    __device__ __forceinline__ void add(const float2& q, float rr) {
#pragma clang fp contract(fast)
        L[0] += q.x * rr;
        L[1] += q.y * rr;
    }
};

template <>
struct AccumL<Ch4> {
    const Ch4* __restrict__ vox;
    int Nz;
    float L[VCBCT_MAX_MAT];
    Ch4 p0 = {0.f, 0.f, 0.f, 0.f};
    float r0 = 0.f;
    __device__ __forceinline__ void step(const Seg& s0, const Seg& s1, const Seg& s2) {
        add(p0, r0);
        Ch4 q1, q2;
        if (s1.ok) q1 = ldg(vox, (size_t)(unsigned)s1.cell);
        if (s2.ok) q2 = ldg(vox, (size_t)(unsigned)s2.cell);
        p0 = ldg(vox, (size_t)(unsigned)(s0.ok ? s0.cell : 0));
        r0 = s0.ok ? (float)s0.rr : 0.f;
        if (s1.ok) add(q1, (float)s1.rr);
        if (s2.ok) add(q2, (float)s2.rr);
    }
    __device__ __forceinline__ void finish() {
        add(p0, r0);
        r0 = 0.f;
    }
    __device__ __forceinline__ void add(const Ch4& q, float rr) {
#pragma clang fp contract(fast)
        L[0] += q.x * rr;
        L[1] += q.y * rr;
        L[2] += q.z * rr;
        L[3] += q.w * rr;
    }
};

// The channel volume of a material volume, one thread per voxel: c_k = sum_m b[m][k] rho_m in FP32, m
// ascending (with the identity mix c_k = rho_k exactly).  SrcT uint4: one u16 record per voxel; float4: two
// FP32 records per voxel.  DstT float2 or Ch4 (channels past mix.nch are zero).
template <class SrcT, class DstT>
__global__ __launch_bounds__(256) void k_channels(const SrcT* __restrict__ src, DstT* __restrict__ dst, size_t nvox,
                                                  ChannelMix mix, int nmat) {
    constexpr int C = sizeof(DstT) / sizeof(float);
    for (size_t v = (size_t)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (size_t)gridDim.x * 256) {
        float rho[VCBCT_MAX_MAT];
        if constexpr (std::is_same<SrcT, uint4>::value) {
            const uint4 q = ldg(src, v);
            rho[0] = (float)(q.x & 0xffffu);
            rho[1] = (float)(q.x >> 16);
            rho[2] = (float)(q.y & 0xffffu);
            rho[3] = (float)(q.y >> 16);
            rho[4] = (float)(q.z & 0xffffu);
            rho[5] = (float)(q.z >> 16);
        } else {
            const float4 a = ldg(src, 2 * v), b = ldg(src, 2 * v + 1);
            rho[0] = a.x;
            rho[1] = a.y;
            rho[2] = a.z;
            rho[3] = a.w;
            rho[4] = b.x;
            rho[5] = b.y;
        }
        float c[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
            float acc = 0.0f;
            if (k < mix.nch) {
#pragma unroll
                for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                    if (m < nmat) acc = fmaf(mix.b[m][k], rho[m], acc);
            }
            c[k] = acc;
        }
        DstT r;
        __builtin_memcpy(&r, c, sizeof r);
        dst[v] = r;  // (a kernel argument: the compiler knows it addresses global memory)
    }
}

// ------------------------------------------------------------------------------------------------
// Spectral sums of opaque rays.  The fast paths form each energy term with an FP32 exponential, which
// underflows once mu.L / xf exceeds ~87 (103 with denormals) in a bin; a ray that is that opaque in every bin
// would give I = 0, p = -log(0) = +inf and a coefficient 0/0.  Such rays are recognised by their FP32 sum
// (below 1e-30 of the spectrum's weight: every term then sits within a few decades of underflow) and their
// terms are re-formed in FP64 from the FP64 tables, relative to the smallest exponent x_min (over the bins
// with weight):  sum_e pdf_e e^{-x_e} = e^{-x_min} sum_e pdf_e e^{-(x_e - x_min)}.  Rare, so the FP64
// exponentials cost nothing measurable.

// All six material loops below run over the compile-time VCBCT_MAX_MAT with a guard on nmat, so the per-ray
// arrays keep constant indices and stay in registers.

// x_e = sum_m mu_me L_m / xf in FP64.
__device__ __forceinline__ double spectral_x(const float (&L)[VCBCT_MAX_MAT], int nmat, const double* __restrict__ mu,
                                             int NE, int e, double inv_xf) {
    double x = 0.0;
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m)
        if (m < nmat) x += ldg(mu, m * NE + e) * (double)L[m];
    return x * inv_xf;
}

// x_min over the bins with pdf_e > 0 (0 if there are none).
__device__ __forceinline__ double spectral_xmin(const float (&L)[VCBCT_MAX_MAT], int nmat,
                                                const double* __restrict__ mu, const double* __restrict__ pdf, int NE,
                                                double inv_xf) {
    double xmin = 0.0;
    bool first = true;
    for (int e = 0; e < NE; ++e) {
        if (!(ldg(pdf, e) > 0.0)) continue;
        const double x = spectral_x(L, nmat, mu, NE, e, inv_xf);
        if (first || x < xmin) xmin = x;
        first = false;
    }
    return xmin;
}

// sum_e pdf_e e^{-(x_e - x_min)}, and with MUB also sum_e pdf_e e^{-(x_e - x_min)} ch_ke into mub (FP64; ch the
// adjoint's channel table, nch channels).
template <bool MUB>
__device__ __forceinline__ double spectral_shifted(const float (&L)[VCBCT_MAX_MAT], int nmat,
                                                   const double* __restrict__ mu, const double* __restrict__ pdf,
                                                   int NE, double inv_xf, double xmin, double (&mub)[VCBCT_MAX_MAT],
                                                   const double* __restrict__ ch = nullptr, int nch = 0) {
    double se = 0.0;
    for (int e = 0; e < NE; ++e) {
        const double we = ldg(pdf, e) * exp(-(spectral_x(L, nmat, mu, NE, e, inv_xf) - xmin));
        se += we;
        if (MUB) {
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nch) mub[m] += we * ldg(ch, m * NE + e);
        }
    }
    return se;
}

// The intensity of an opaque ray with the reference's own terms, float(pdf_e * exp(-transp_e / xf) * leng)
// summed in FP32 in e order (P/reprojectionOnGPU.cu:432, 1038-1044): most are zero or denormal in FP32.
__device__ __forceinline__ float intensity_exact(const float (&L)[VCBCT_MAX_MAT], int nmat,
                                                 const double* __restrict__ mu, const double* __restrict__ pdf, int NE,
                                                 double inv_xf, double leng) {
    float s = 0.0f;
    for (int e = 0; e < NE; ++e) s += (float)(ldg(pdf, e) * exp(-spectral_x(L, nmat, mu, NE, e, inv_xf)) * leng);
    return s;
}

// -log(sum_e pdf_e e^{-x_e} / sw) of an opaque ray.
__device__ __forceinline__ double logproj_opaque(const float (&L)[VCBCT_MAX_MAT], int nmat,
                                                 const double* __restrict__ mu, const double* __restrict__ pdf, int NE,
                                                 double inv_xf, double sw) {
    const double xmin = spectral_xmin(L, nmat, mu, pdf, NE, inv_xf);
    double unused[VCBCT_MAX_MAT];
    return xmin - log(spectral_shifted<false>(L, nmat, mu, pdf, NE, inv_xf, xmin, unused) / sw);
}

// The transmission-weighted spectral mean of the channel table of an opaque ray (mub[k], already divided by the
// sum) and its log projection.
__device__ __forceinline__ double mubar_opaque(const float (&L)[VCBCT_MAX_MAT], int nmat, const double* __restrict__ mu,
                                               const double* __restrict__ pdf, int NE, double inv_xf, double sw,
                                               double (&mub)[VCBCT_MAX_MAT], const double* __restrict__ ch, int nch) {
    const double xmin = spectral_xmin(L, nmat, mu, pdf, NE, inv_xf);
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) mub[m] = 0.0;
    const double se = spectral_shifted<true>(L, nmat, mu, pdf, NE, inv_xf, xmin, mub, ch, nch);
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) mub[m] /= se;
    return xmin - log(se / sw);
}

// Counts the charged segments of a ray (the algorithmic work unit of the roofline model).
struct CountSeg {
    int n = 0;
    __device__ __forceinline__ void step(const Seg& a, const Seg& b, const Seg& c) { n += (int)a.ok + (int)b.ok + (int)c.ok; }
};

// Work decomposition of the ray kernels (TileShape).  One wave = 64 detector rows of one column at one
// angle; row group g covers rows [64 g, 64 g + 64).  A block is rgb row groups by `cols` adjacent columns of
// one angle, wave w taking row group w / cols and column w % cols of the block's tile.  Blocks are dealt
// round-robin to the 8 XCDs (block b runs on XCD b % 8).  With split 0 the column groups of every (row block,
// angle) are cut into 8 contiguous runs, one per XCD: all XCDs work on the same angle and each XCD's L2 sees
// adjacent columns.  With split 1 the angles are cut into 8 runs instead: each XCD walks its own run of
// angles over all columns, so the voxels that neighbouring angles share stay in that XCD's L2.  order 0 walks
// the angles inside each row block (one z band of the volume at a time), order 1 the row blocks inside each
// angle.
struct Tile {
    int rg, a, col;
};
__device__ __forceinline__ Tile tile_map(const TileMap& m, int nA, int b) {
    const int xcd = b & 7, local = b >> 3, w = threadIdx.x >> 6;
    int cg, a, rb;
    if (m.split == 0) {
        const int q = local % m.gpx, ra = local / m.gpx;
        cg = xcd * m.gpx + q;
        if (m.order == 0) a = ra % nA, rb = ra / nA;
        else rb = ra % m.nrb, a = ra / m.nrb;
    } else {
        cg = local % m.ncg;
        const int ra = local / m.ncg;
        int al;
        if (m.order == 0) al = ra % m.gpx, rb = ra / m.gpx;
        else rb = ra % m.nrb, al = ra / m.nrb;
        a = xcd * m.gpx + al;
    }
    Tile t;
    t.a = a;
    t.rg = rb * m.rgb + w / m.cols;
    t.col = cg * m.cols + w % m.cols;
    return t;
}

__device__ __forceinline__ Tile tile_map(const TileMap& m, int nA) { return tile_map(m, nA, blockIdx.x); }

// Host side: the map of shape s over nA angles and its block count (-1 when the grid is too large).  Every
// (row group, angle, column) of the launch falls to exactly one wave; waves past the last column, angle or
// row group exit at once.
long tile_map_host(const ProjGeom& g, int nA, const TileShape& s, TileMap& m) {
    m.rgb = s.rgb;
    m.cols = s.cols;
    m.order = s.order;
    m.split = s.split;
    const int nrg = (g.H + 63) / 64;
    m.nrb = (nrg + s.rgb - 1) / s.rgb;
    m.ncg = (g.W + s.cols - 1) / s.cols;
    m.gpx = s.split == 0 ? (m.ncg + 7) / 8 : (nA + 7) / 8;
    const long nb = s.split == 0 ? 8L * m.gpx * nA * m.nrb : 8L * m.gpx * m.ncg * m.nrb;
    return nb > INT_MAX ? -1 : nb;
}

bool tile_shape_valid(const TileShape& s) {
    const int w = s.rgb * s.cols;
    return s.rgb >= 1 && s.cols >= 1 && (w == 2 || w == 4 || w == 8) && (s.order == 0 || s.order == 1) &&
           (s.split == 0 || s.split == 1);
}

// One wave: 64 detector rows [64 rg, 64 rg + 64) of column jj at angle a.  Wave-uniform arguments.
template <int MODE, class VoxT>
__device__ __forceinline__ void forward_wave(const ProjGeom& g, const double* __restrict__ tab,
                                             const VoxT* __restrict__ vox, const double* __restrict__ mu,
                                             const double* __restrict__ pdf, const float* __restrict__ muf,
                                             const float* __restrict__ pdff, int NE, double inv_xf, int nl, int a,
                                             int rg, int jj, float* __restrict__ out, float* __restrict__ aux,
                                             WaveLds* steps) {
    const int lane = threadIdx.x & 63;
    const int j = rg * 64 + lane;
    const bool active = j < g.H;

    const double c = ldg(tab, a), sn = ldg(tab, g.nAng + a), xc = ldg(tab, 2 * g.nAng + a),
                 yc = ldg(tab, 3 * g.nAng + a);
    const size_t o = (size_t)a * g.H * g.W + (size_t)(g.H - 1 - j) * g.W + jj;
    if (MODE == VCBCT_MODE_SEGCOUNT) {
        CountSeg cnt;
        traverse_wave(g, c, sn, xc, yc, j, jj, active, cnt, steps);
        if (active) stg(aux, o, (float)cnt.n);
        return;
    }
    AccumL<VoxT> acc;
    acc.vox = vox;
    acc.Nz = g.Nz;
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) acc.L[m] = 0.0f;
    const double leng = traverse_wave(g, c, sn, xc, yc, j, jj, active, acc, steps);
    acc.finish();
    if (!active) return;

    if (MODE == VCBCT_MODE_LINEINT) {
        for (int m = 0; m < g.nmat; ++m) stg(aux, o * g.nmat + m, (float)acc.L[m]);
        return;
    }
    if (MODE == VCBCT_MODE_LOGPROJ && aux) {  // the nl line integrals, kept for the backward pass ([term][ray])
        const size_t nrays = (size_t)g.nAng * g.H * g.W;
#pragma unroll
        for (int m = 0; m < VCBCT_MAX_MAT; ++m)
            if (m < nl) stg(aux, (size_t)m * nrays + o, acc.L[m]);
    }
    bool zero = true;
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) zero = zero && (acc.L[m] == 0.0);
    if (MODE == VCBCT_MODE_INTENSITY) {
        // sum_e pdf_e * exp(-transp_e / xf) * leng, accumulated in FP32 in e order (P/reprojectionOnGPU.cu:432,
        // 1038-1044).  A ray that charges nothing sums the reference's own terms, float(pdf_e * 1.0 * leng), bit
        // for bit.  Otherwise the terms are FP32: the tables are FP32 as given, the line integrals FP32 sums,
        // and each term carries ~1e-6 relative error against the reference's FP64 term rounded to float, far
        // below the FP32 accumulation of the line integrals.  exp(-x) = 2^(-x log2 e).
        float s = 0.0f;
        if (zero) {
            for (int e = 0; e < NE; ++e) s += (float)(ldg(pdf, e) * 1.0 * leng);
        } else {
            const float lf = (float)leng;
            const float k2 = (float)(-inv_xf * 1.4426950408889634);
            for (int e = 0; e < NE; ++e) {
                float tr = 0.0f;
#pragma unroll
                for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                    if (m < nl) tr = fmaf(ldg(muf, m * NE + e), acc.L[m], tr);
                s += ldg(pdff, e) * exp2f(tr * k2) * lf;
            }
            // opaque ray: the FP32 exponentials underflow; the reference's own FP64 terms instead
            if (s < 1e-30f) s = intensity_exact(acc.L, nl, mu, pdf, NE, inv_xf, leng);
        }
        stg(out, o, s);
    } else {  // VCBCT_MODE_LOGPROJ: p = -log(sum_e pdf_e e^{-mu.L/xf} / sum_e pdf_e); FP32 terms, FP64 sums
        double sw = 0.0, se = 0.0;
        const float k2 = (float)(-inv_xf * 1.4426950408889634);
        for (int e = 0; e < NE; ++e) {
            float tr = 0.0f;
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nl) tr = fmaf(ldg(muf, m * NE + e), acc.L[m], tr);
            const double pe = ldg(pdf, e);
            sw += pe;
            se += pe * (double)exp2f(tr * k2);
        }
        // opaque ray (the FP32 exponentials underflow): FP64 terms relative to the smallest exponent
        const double p = se < 1e-30 * sw ? logproj_opaque(acc.L, nl, mu, pdf, NE, inv_xf, sw) : -log(se / sw);
        stg(out, o, (float)p);
    }
}

// WPB waves per block (TileShape: rgb * cols); the block's waves share nothing but the LDS allocation.
template <int MODE, class VoxT, int WPB>
__global__ __launch_bounds__(64 * WPB) VCBCT_WAVES_FWD(MODE) void k_forward(ProjGeom g, const double* __restrict__ tab,
                                                                 const VoxT* __restrict__ vox,
                                                                 const double* __restrict__ mu,
                                                                 const double* __restrict__ pdf,
                                                                 const float* __restrict__ muf,
                                                                 const float* __restrict__ pdff, int NE,
                                                                 double inv_xf, int nl, int a0, int nA, TileMap tm,
                                                                 float* __restrict__ out, float* __restrict__ aux) {
    __shared__ WaveLds steps[WPB];
    const Tile t = tile_map(tm, nA);
    if (t.col >= g.W || t.a >= nA || t.rg * 64 >= g.H) return;  // wave out of range: uniform exit
    forward_wave<MODE, VoxT>(g, tab, vox, mu, pdf, muf, pdff, NE, inv_xf, nl, a0 + t.a, t.rg, t.col, out, aux,
                             &steps[threadIdx.x >> 6]);
}

// ------------------------------------------------------------------------------------------------
// Batched forward: many problems (volumes of different slice counts, angle counts and spectra) in one
// persistent launch.  Work item = one wave = (problem, row group, angle, column).  Each XCD has its own
// queue (work counter) holding, per problem, either a contiguous eighth of the columns of every (row group,
// angle) (split 0) or a contiguous eighth of the angles with all their columns (split 1); consecutive items
// are neighbouring columns of one angle, so the waves of an XCD share voxels in its L2.  A wave that finds
// its XCD's queue empty steals from the other queues, so every queue drains and every wave exits.
__device__ __forceinline__ int find_problem(const BatchProblem* __restrict__ pr, int np, int xcd, int item) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {  // last problem whose start <= item
        int mid = (lo + hi + 1) >> 1;
        if (pr[mid].xcd_start[xcd] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class VoxT>
__global__ __launch_bounds__(256) VCBCT_WAVES7 void k_forward_batch(const BatchProblem* __restrict__ pr, int np,
                                                       const int* __restrict__ xcd_total, int* __restrict__ counters) {
    __shared__ WaveLds steps[4];
    const int home = blockIdx.x & 7;
    const int lane = threadIdx.x & 63;
    for (int qi = 0; qi < 8; ++qi) {
        const int q = (home + qi) & 7;
        const int total = xcd_total[q];
        for (;;) {
            int item = 0;
            if (lane == 0) item = atomicAdd(&counters[q], 1);
            // broadcast from lane 0 into a scalar register: the problem and its parameters are then
            // wave-uniform (scalar loads), as in the plain kernel
            item = __builtin_amdgcn_readfirstlane(item);
            if (item >= total) break;
            const int p = find_problem(pr, np, q, item);
            const BatchProblem& P = pr[p];
            const int local = item - P.xcd_start[q];
            const int cl = local % P.pcols;
            const int ra = local / P.pcols;  // rg * pang + angle within the queue's range
            const int al = ra % P.pang, rg = ra / P.pang;
            const int a = P.split ? q * P.pang + al : al;
            const int jj = P.split ? cl : q * P.pcols + cl;
            if (jj >= P.g.W || a >= P.g.nAng) continue;
            forward_wave<VCBCT_MODE_INTENSITY, VoxT>(P.g, P.tab, (const VoxT*)P.vox, P.mu, P.pdf, P.muf, P.pdff, P.NE,
                                                     P.inv_xf, P.g.nmat, a, rg, jj, P.out, nullptr,
                                                     &steps[threadIdx.x >> 6]);
        }
    }
}

// Flat batched grid: the blocks of every problem of one voxel type laid end to end, each problem decoding its
// own blocks with its tile map (each problem's block count is a multiple of 8, so block b still runs on XCD
// b % 8 and the map's XCD assignment carries over).
template <class VoxT>
__global__ __launch_bounds__(256) VCBCT_WAVES8 void k_forward_flat(const BatchProblem* __restrict__ pr, int np) {
    __shared__ WaveLds steps[4];
    const int b = blockIdx.x;
    int lo = 0, hi = np - 1;
    while (lo < hi) {  // last problem whose first block <= b
        int mid = (lo + hi + 1) >> 1;
        if (pr[mid].block_start <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchProblem& P = pr[lo];
    const Tile t = tile_map(P.tm, P.g.nAng, b - P.block_start);
    if (t.col >= P.g.W || t.a >= P.g.nAng || t.rg * 64 >= P.g.H) return;  // wave out of range: uniform exit
    forward_wave<VCBCT_MODE_INTENSITY, VoxT>(P.g, P.tab, (const VoxT*)P.vox, P.mu, P.pdf, P.muf, P.pdff, P.NE,
                                             P.inv_xf, P.g.nmat, t.a, t.rg, t.col, P.out, nullptr,
                                             &steps[threadIdx.x >> 6]);
}

int launch_forward_flat(const BatchProblem* d_problems, int np, int total_blocks, int vox_is_f32, hipStream_t st) {
    if (np <= 0 || total_blocks <= 0) return 0;
    if (vox_is_f32)
        hipLaunchKernelGGL(k_forward_flat<float4>, dim3(total_blocks), dim3(256), 0, st, d_problems, np);
    else
        hipLaunchKernelGGL(k_forward_flat<uint4>, dim3(total_blocks), dim3(256), 0, st, d_problems, np);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

int launch_forward_batch(const BatchProblem* d_problems, int np, const int* d_xcd_total, int* d_counters, int nblocks,
                         int vox_is_f32, hipStream_t st) {
    if (np <= 0) return 0;
    int rc = launch_zero(d_counters, 8 * sizeof(int), st);
    if (rc) return rc;
    if (vox_is_f32)
        hipLaunchKernelGGL(k_forward_batch<float4>, dim3(nblocks), dim3(256), 0, st, d_problems, np, d_xcd_total,
                           d_counters);
    else
        hipLaunchKernelGGL(k_forward_batch<uint4>, dim3(nblocks), dim3(256), 0, st, d_problems, np, d_xcd_total,
                           d_counters);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------------------------------------
// Adjoint, in two kernels.
//
// k_adjoint_coef: recompute L_m for the ray (traversal over the FP32 maps), then the per-ray, per-material
// coefficient
//   c_m = w_ray * mubar_m / xf,  mubar_m = sum_e pdf_e e_e mu_me / sum_e pdf_e e_e,
// with w_ray = 2 (p - target) (loss mode, the loss summed in FP64) or the upstream gradient dL/dp (upstream
// mode); c is stored [mat][ray].
// k_adjoint_scatter: walk the ray again and add c_m * rr into the gradient for each charged segment (FP32
// atomics into the planar [mat][y][x][z] layout, transposed to the caller's [mat][z][y][x] afterwards).
// (Splitting the scatter into z-slab passes, to keep each pass's gradient lines in the last-level cache, was
// measured slower: the atomics are executed beyond L2 either way, and the passes repeat traversal work.)
// Adds c_m * rr of each charged segment to the gradient.  A float atomic instruction whose lanes name the same
// address more than once is processed lane by lane (measured on MI355X: two lanes per address cost 6x, four
// per address as much as 64 scattered lanes), and neighbouring detector rows often charge the same voxel
// (rows closer than slices).  So the lanes of a slot that name the same voxel are first combined: cells
// form runs over consecutive lanes (the lanes of a wave share the xy cell of a slot and their z is monotone
// in the row); the run head sums its run (a segmented sum in log steps) and alone issues the atomics.
// NCH: the scattered channels (compile time, so the run merge and the atomics touch only those).
template <int NCH>
struct ScatterG {
    float* __restrict__ gvox;
    float c[NCH];
    size_t plane;  // N * N * Nz: one channel plane of the [channel][y][x][z] scatter layout
    __device__ __forceinline__ void seg(const Seg& s) {
        const int lane = threadIdx.x & 63;
        const bool ok = s.ok;
        if (!__any(ok)) return;
        const int key = ok ? s.cell : -1 - lane;  // lanes that charge nothing never join a run
        const int prev = __shfl_up(key, 1, 64);
        const bool start = lane == 0 || prev != key;
        // lanes outside the current exec mask (a caller's divergent branch) end every run
        const unsigned long long starts = __ballot(start) | ~__ballot(true);
        // [lane, next start): the rest of the lane's run
        const unsigned long long after = lane == 63 ? 0ull : (starts >> (lane + 1));
        const int len = after ? __ffsll((long long)after) : 64 - lane;
        const float rf = (float)s.rr;
        float v[NCH];
#pragma unroll
        for (int m = 0; m < NCH; ++m) v[m] = ok ? c[m] * rf : 0.f;
        // Segmented sum in log steps: after the step of offset d every lane holds the sum of its run's values
        // over [lane, min(lane + 2d, run end)), so the heads hold their runs' sums once 2d reaches the longest
        // run.  cont (wave-uniform) marks the lanes that continue a run; after the step of offset d its bit i
        // is set where lanes i .. i + 2d - 1 all continue one, i.e. where a run is still longer than 2d.
        unsigned long long cont = ~starts;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            if (!cont) break;
#pragma unroll
            for (int m = 0; m < NCH; ++m) {
                const float w = __shfl_down(v[m], d, 64);
                if (d < len) v[m] += w;
            }
            cont &= cont >> d;
        }
        if (!(ok && start)) return;
        // planar layout: the heads (neighbouring z of one (x, y)) hit a contiguous run of one channel plane
        float* p = gvox + (size_t)(unsigned)s.cell;
#pragma unroll
        for (int m = 0; m < NCH; ++m)
            if (v[m] != 0.f) atomicAdd(p + m * plane, v[m]);
    }
    __device__ __forceinline__ void step(const Seg& a, const Seg& b, const Seg& c2) {
        seg(a);
        seg(b);
        seg(c2);
    }
};

// VoxT float4: the FP32 maps (mu / muf the material tables, nl = nmat); float2 / Ch4: their channel volume
// (mu / muf the channel tables, nl = nch).
template <int MODE, class VoxT>
__global__ __launch_bounds__(256) VCBCT_WAVES8 void k_adjoint_coef(ProjGeom g, const double* __restrict__ tab,
                                                      const VoxT* __restrict__ vox, const double* __restrict__ mu,
                                                      const double* __restrict__ pdf, const float* __restrict__ muf,
                                                      const double* __restrict__ ch, int nch, int NE, double inv_xf,
                                                      int nl, TileMap tm, const float* __restrict__ tg,
                                                      float* __restrict__ coef, double* __restrict__ loss) {
    __shared__ WaveLds step_tabs[4];
    WaveLds* steps = &step_tabs[threadIdx.x >> 6];
    const Tile t = tile_map(tm, g.nAng);
    const int lane = threadIdx.x & 63;
    const int j = t.rg * 64 + lane;
    const int jj = t.col;
    const int a = t.a;
    if (jj >= g.W || a >= g.nAng || t.rg * 64 >= g.H) return;  // wave out of range: uniform exit
    const bool active = j < g.H;
    const double c = ldg(tab, a), sn = ldg(tab, g.nAng + a), xc = ldg(tab, 2 * g.nAng + a),
                 yc = ldg(tab, 3 * g.nAng + a);
    AccumL<VoxT> acc;
    acc.vox = vox;
    acc.Nz = g.Nz;
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) acc.L[m] = 0.0f;
    traverse_wave(g, c, sn, xc, yc, j, jj, active, acc, steps);
    acc.finish();  // (the pipelined accumulators add their last gather here)

    double lossc = 0.0;
    if (active) {
        // the exponent and its exponential in FP32 (relative error ~1e-6 per term), the sums in FP64
        double sw = 0.0, se = 0.0, mub[VCBCT_MAX_MAT] = {0, 0, 0, 0, 0, 0};
        const float k2 = (float)(-inv_xf * 1.4426950408889634);
        for (int e = 0; e < NE; ++e) {
            float tr = 0.0f;
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nl) tr = fmaf(ldg(muf, m * NE + e), acc.L[m], tr);
            const double pe = ldg(pdf, e);
            const double we = pe * (double)exp2f(tr * k2);
            sw += pe;
            se += we;
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nch) mub[m] += we * ldg(ch, m * NE + e);
        }
        const size_t nrays = (size_t)g.nAng * g.H * g.W;
        const size_t o = (size_t)a * g.H * g.W + (size_t)(g.H - 1 - j) * g.W + jj;
        double p;
        if (se < 1e-30 * sw) {  // opaque ray: FP64 terms relative to the smallest exponent
            p = mubar_opaque(acc.L, nl, mu, pdf, NE, inv_xf, sw, mub, ch, nch);
        } else {
            p = -log(se / sw);
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m) mub[m] /= se;
        }
        double w;
        if (MODE == VCBCT_ADJ_LOSS) {
            double res = p - (double)ldg(tg, o);
            lossc = res * res;
            w = 2.0 * res;
        } else {
            w = (double)ldg(tg, o);
        }
#pragma unroll
        for (int m = 0; m < VCBCT_MAX_MAT; ++m)
            if (m < nch) stg(coef, (size_t)m * nrays + o, (float)(w * mub[m] * inv_xf));
    }
    if (MODE == VCBCT_ADJ_LOSS && loss) {
        // wave reduction of the loss, one atomic per wave
        double v = lossc;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) atomicAdd(loss, v);
    }
}

// The multi-spectrum form of k_adjoint_coef (loss mode): one traversal per ray gives the line integrals, and
// for every spectrum s the log projection p_s, its residual against the target t_s ([s][ray]) and its
// coefficients; the loss sums (p_s - t_s)^2 over spectra and rays, and the stored coefficient of a channel is
// the sum over the spectra, c_k = sum_s 2 (p_s - t_s) chbar_sk / xf, so that one scatter serves all spectra.
// NCH: the channel count (the arrays of the spectral loop stay in registers).  VoxT float4: the FP32 maps (the
// exponents over the material tables); float2 / Ch4: the joint basis's channel volume (over the channel tables).
template <int NCH, class VoxT>
__global__ __launch_bounds__(256) VCBCT_WAVES7 void k_adjoint_coef_multi(ProjGeom g, const double* __restrict__ tab,
                                                                         const VoxT* __restrict__ vox, SpecSet ss,
                                                                         double inv_xf, TileMap tm,
                                                                         const float* __restrict__ tg,
                                                                         float* __restrict__ coef,
                                                                         double* __restrict__ loss) {
    __shared__ WaveLds step_tabs[4];
    WaveLds* steps = &step_tabs[threadIdx.x >> 6];
    const Tile t = tile_map(tm, g.nAng);
    const int lane = threadIdx.x & 63;
    const int j = t.rg * 64 + lane;
    const int jj = t.col;
    const int a = t.a;
    if (jj >= g.W || a >= g.nAng || t.rg * 64 >= g.H) return;  // wave out of range: uniform exit
    const bool active = j < g.H;
    const double c = ldg(tab, a), sn = ldg(tab, g.nAng + a), xc = ldg(tab, 2 * g.nAng + a),
                 yc = ldg(tab, 3 * g.nAng + a);
    AccumL<VoxT> acc;
    acc.vox = vox;
    acc.Nz = g.Nz;
#pragma unroll
    for (int m = 0; m < VCBCT_MAX_MAT; ++m) acc.L[m] = 0.0f;
    traverse_wave(g, c, sn, xc, yc, j, jj, active, acc, steps);
    acc.finish();  // (the pipelined accumulators add their last gather here)

    double lossc = 0.0;
    if (active) {
        const size_t nrays = (size_t)g.nAng * g.H * g.W;
        const size_t o = (size_t)a * g.H * g.W + (size_t)(g.H - 1 - j) * g.W + jj;
        const float k2 = (float)(-inv_xf * 1.4426950408889634);
        double cacc[NCH];
#pragma unroll
        for (int m = 0; m < NCH; ++m) cacc[m] = 0.0;
        for (int s = 0; s < ss.nspec; ++s) {
            const int NE = ss.NE[s];
            constexpr bool CHV = !std::is_same<VoxT, float4>::value;  // channel volume
            const int nl = CHV ? NCH : g.nmat;
            const double* __restrict__ mu = CHV ? ss.ch[s] : ss.mu[s];
            const double* __restrict__ pdf = ss.pdf[s];
            const float* __restrict__ muf = CHV ? ss.chf[s] : ss.muf[s];
            const double* __restrict__ ch = ss.ch[s];
            double sw = 0.0, se = 0.0, mub[VCBCT_MAX_MAT] = {0, 0, 0, 0, 0, 0};
            for (int e = 0; e < NE; ++e) {
                float tr = 0.0f;
#pragma unroll
                for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                    if (m < nl) tr = fmaf(ldg(muf, m * NE + e), acc.L[m], tr);
                const double pe = ldg(pdf, e);
                const double we = pe * (double)exp2f(tr * k2);
                sw += pe;
                se += we;
#pragma unroll
                for (int m = 0; m < NCH; ++m) mub[m] += we * ldg(ch, m * NE + e);
            }
            double p;
            if (se < 1e-30 * sw) {  // opaque ray: FP64 terms relative to the smallest exponent
                p = mubar_opaque(acc.L, nl, mu, pdf, NE, inv_xf, sw, mub, ch, NCH);
            } else {
                p = -log(se / sw);
#pragma unroll
                for (int m = 0; m < NCH; ++m) mub[m] /= se;
            }
            const double res = p - (double)ldg(tg, (size_t)s * nrays + o);
            lossc += res * res;
#pragma unroll
            for (int m = 0; m < NCH; ++m) cacc[m] += 2.0 * res * mub[m];
        }
#pragma unroll
        for (int m = 0; m < NCH; ++m) stg(coef, (size_t)m * nrays + o, (float)(cacc[m] * inv_xf));
    }
    if (loss) {
        double v = lossc;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) atomicAdd(loss, v);
    }
}

// The coefficients of the upstream-gradient mode from line integrals saved by the forward log projection
// (the same traversal, so the same FP32 sums): one thread per ray, the epilogue of k_adjoint_coef.
__global__ __launch_bounds__(256) void k_adjoint_coef_saved(ProjGeom g, const float* __restrict__ Ls,
                                                            const double* __restrict__ mu,
                                                            const double* __restrict__ pdf,
                                                            const float* __restrict__ muf,
                                                            const double* __restrict__ ch, int nch, int NE,
                                                            double inv_xf, int nl, const float* __restrict__ gp,
                                                            float* __restrict__ coef) {
    const size_t nrays = (size_t)g.nAng * g.H * g.W;
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < nrays; o += (size_t)gridDim.x * 256) {
        float L[VCBCT_MAX_MAT];
#pragma unroll
        for (int m = 0; m < VCBCT_MAX_MAT; ++m) L[m] = m < nl ? ldg(Ls, (size_t)m * nrays + o) : 0.0f;
        double sw = 0.0, se = 0.0, mub[VCBCT_MAX_MAT] = {0, 0, 0, 0, 0, 0};
        const float k2 = (float)(-inv_xf * 1.4426950408889634);
        for (int e = 0; e < NE; ++e) {
            float tr = 0.0f;
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nl) tr = fmaf(ldg(muf, m * NE + e), L[m], tr);
            const double pe = ldg(pdf, e);
            const double we = pe * (double)exp2f(tr * k2);
            sw += pe;
            se += we;
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                if (m < nch) mub[m] += we * ldg(ch, m * NE + e);
        }
        if (se < 1e-30 * sw) {  // opaque ray: FP64 terms relative to the smallest exponent
            (void)mubar_opaque(L, nl, mu, pdf, NE, inv_xf, sw, mub, ch, nch);
        } else {
#pragma unroll
            for (int m = 0; m < VCBCT_MAX_MAT; ++m) mub[m] /= se;
        }
        const double w = (double)ldg(gp, o);
#pragma unroll
        for (int m = 0; m < VCBCT_MAX_MAT; ++m)
            if (m < nch) stg(coef, (size_t)m * nrays + o, (float)(w * mub[m] * inv_xf));
    }
}

template <int NCH>
__global__ __launch_bounds__(256) VCBCT_WAVES8 void k_adjoint_scatter(ProjGeom g, const double* __restrict__ tab,
                                                                     TileMap tm, const float* __restrict__ coef,
                                                                     float* __restrict__ gvox) {
    __shared__ WaveLds step_tabs[4];
    WaveLds* steps = &step_tabs[threadIdx.x >> 6];
    const Tile t = tile_map(tm, g.nAng);
    const int lane = threadIdx.x & 63;
    const int j = t.rg * 64 + lane;
    const int jj = t.col;
    const int a = t.a;
    if (jj >= g.W || a >= g.nAng || t.rg * 64 >= g.H) return;  // wave out of range: uniform exit
    const bool active = j < g.H;
    ScatterG<NCH> sc;  // the scattered channels: the maps, or their spectral basis
    sc.gvox = gvox;
    sc.plane = (size_t)g.N * g.N * g.Nz;
    bool any = false;
    const size_t nrays = (size_t)g.nAng * g.H * g.W;
    const size_t o = (size_t)a * g.H * g.W + (size_t)(g.H - 1 - (active ? j : 0)) * g.W + jj;
#pragma unroll
    for (int m = 0; m < NCH; ++m) {
        sc.c[m] = active ? ldg(coef, (size_t)m * nrays + o) : 0.f;
        any = any || sc.c[m] != 0.f;
    }
    if (!__any(any)) return;  // nothing to scatter for this wave (e.g. zero upstream gradient)
    const double c = ldg(tab, a), sn = ldg(tab, g.nAng + a), xc = ldg(tab, 2 * g.nAng + a),
                 yc = ldg(tab, 3 * g.nAng + a);
    traverse_wave(g, c, sn, xc, yc, j, jj, active, sc, steps);
}

// ------------------------------------------------------------------------------------------------
// Launchers

__global__ void k_zero(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

int launch_zero(void* p, size_t bytes, hipStream_t st) {
    size_t n = bytes / 4;
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, st, (uint32_t*)p, n);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

int launch_transpose_u16(const uint16_t* src, uint16_t* dst, const ProjGeom& g, hipStream_t st) {
    dim3 grid((g.N + TT - 1) / TT, (g.Nz + TT - 1) / TT, g.N);
    hipLaunchKernelGGL(k_transpose_u16, grid, dim3(256), 0, st, src, (uint4*)dst, g.N, g.Nz, g.nmat);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
int launch_transpose_f32(const float* src, float* dst, const ProjGeom& g, hipStream_t st) {
    dim3 grid((g.N + TT - 1) / TT, (g.Nz + TT - 1) / TT, g.N);
    hipLaunchKernelGGL(k_transpose_f32, grid, dim3(256), 0, st, src, (float4*)dst, g.N, g.Nz, g.nmat);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
int launch_untranspose_mix(const float* src, float* dst, const ProjGeom& g, const ChannelMix& mix, hipStream_t st) {
    dim3 grid((g.N + TT - 1) / TT, (g.Nz + TT - 1) / TT, g.N);
    hipLaunchKernelGGL(k_untranspose_mix, grid, dim3(256), 0, st, src, dst, g.N, g.Nz, g.nmat, mix);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
int launch_untranspose_f32(const float* src, float* dst, const ProjGeom& g, hipStream_t st) {
    dim3 grid((g.N + TT - 1) / TT, (g.Nz + TT - 1) / TT, g.N * g.nmat);
    hipLaunchKernelGGL(k_untranspose_f32, grid, dim3(256), 0, st, src, dst, g.N, g.Nz, g.nmat);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// One forward launch of shape s (valid, and WPB 4 for the line-integral and segment-count modes).
template <int M, class T>
static void launch_fwd_t(int wpb, dim3 grid, hipStream_t st, const ProjGeom& g, const double* tables, const void* vox,
                         const Spectrum& sp, double inv_xf, int a0, int nA, const TileMap& tm, float* out,
                         float* aux) {
#define VCBCT_LAUNCH_FWD(W)                                                                                  \
    hipLaunchKernelGGL((k_forward<M, T, W>), grid, dim3(64 * W), 0, st, g, tables, (const T*)vox, sp.mu, sp.pdf, \
                       sp.muf, sp.pdff, sp.NE, inv_xf, sp.nl, a0, nA, tm, out, aux)
    if constexpr (M == VCBCT_MODE_INTENSITY || M == VCBCT_MODE_LOGPROJ) {
        if (wpb == 2) VCBCT_LAUNCH_FWD(2);
        else if (wpb == 8) VCBCT_LAUNCH_FWD(8);
        else VCBCT_LAUNCH_FWD(4);
    } else {
        VCBCT_LAUNCH_FWD(4);
    }
#undef VCBCT_LAUNCH_FWD
}

template <int M>
static void launch_fwd_m(int kind, int wpb, dim3 grid, hipStream_t st, const ProjGeom& g, const double* tables,
                         const void* vox, const Spectrum& sp, double inv_xf, int a0, int nA, const TileMap& tm,
                         float* out, float* aux) {
    if constexpr (M == VCBCT_MODE_INTENSITY || M == VCBCT_MODE_LOGPROJ) {
        if (kind == VOX_CH2) {
            launch_fwd_t<M, float2>(wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
            return;
        }
        if (kind == VOX_CH4) {
            launch_fwd_t<M, Ch4>(wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
            return;
        }
    }
    if (kind == VOX_F32) {
        if constexpr (M != VCBCT_MODE_SEGCOUNT)
            launch_fwd_t<M, float4>(wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    } else {
        launch_fwd_t<M, uint4>(wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    }
}

static int launch_fwd_shape(const ProjGeom& g, const double* tables, const void* vox, int vox_kind,
                            const Spectrum& sp, double inv_xf, int a0, int nA, float* out, int mode, float* aux,
                            hipStream_t st, const TileShape& s) {
    TileMap tm;
    const long nb = tile_map_host(g, nA, s, tm);
    if (nb < 0) return fail(VCBCT_EINVAL, "forward: launch too large");
    const bool chv = vox_kind == VOX_CH2 || vox_kind == VOX_CH4;
    if (chv && !(mode == VCBCT_MODE_INTENSITY || mode == VCBCT_MODE_LOGPROJ))
        return fail(VCBCT_EINVAL, "forward: channel volumes serve the intensity and log projections only");
    if (mode == VCBCT_MODE_SEGCOUNT && vox_kind != VOX_U16) return fail(VCBCT_EINVAL, "forward: bad volume kind");
    const int wpb = s.rgb * s.cols;
    const dim3 grid((unsigned)nb);
    if (mode == VCBCT_MODE_INTENSITY)
        launch_fwd_m<VCBCT_MODE_INTENSITY>(vox_kind, wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    else if (mode == VCBCT_MODE_LINEINT)
        launch_fwd_m<VCBCT_MODE_LINEINT>(vox_kind, wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    else if (mode == VCBCT_MODE_SEGCOUNT)
        launch_fwd_m<VCBCT_MODE_SEGCOUNT>(vox_kind, wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    else
        launch_fwd_m<VCBCT_MODE_LOGPROJ>(vox_kind, wpb, grid, st, g, tables, vox, sp, inv_xf, a0, nA, tm, out, aux);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

int launch_channels(const void* vox, int src_is_f32, float* dst, const ProjGeom& g, const ChannelMix& mix,
                    hipStream_t st) {
    const size_t nv = (size_t)g.N * g.N * g.Nz;
    if (mix.nch < 1 || mix.nch > 4) return fail(VCBCT_EINVAL, "channels: %d channels (1 to 4)", mix.nch);
    const unsigned nb = (unsigned)std::min<size_t>((nv + 255) / 256, 1 << 20);
    if (nb == 0) return 0;
    if (mix.nch <= 2) {
        if (src_is_f32)
            hipLaunchKernelGGL((k_channels<float4, float2>), dim3(nb), dim3(256), 0, st, (const float4*)vox,
                               (float2*)dst, nv, mix, g.nmat);
        else
            hipLaunchKernelGGL((k_channels<uint4, float2>), dim3(nb), dim3(256), 0, st, (const uint4*)vox,
                               (float2*)dst, nv, mix, g.nmat);
    } else {
        if (src_is_f32)
            hipLaunchKernelGGL((k_channels<float4, Ch4>), dim3(nb), dim3(256), 0, st, (const float4*)vox, (Ch4*)dst,
                               nv, mix, g.nmat);
        else
            hipLaunchKernelGGL((k_channels<uint4, Ch4>), dim3(nb), dim3(256), 0, st, (const uint4*)vox, (Ch4*)dst, nv,
                               mix, g.nmat);
    }
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------------------------------------
// Tile-shape autotuner.  Per key (kernel, geometry, launch parameters, device) the shape with the shortest
// launch, timed on the caller's stream and buffers by coordinate descent: the (order, split) pairs at the
// default block shape, then the other block shapes at the best pair, then the pairs again at the best block
// shape (a candidate must beat the best so far by 1 %).  Every candidate computes the same values (a shape is
// a schedule only); a kernel that accumulates is timed without its accumulator (the loss) or has it cleared
// afterwards (the gradient).  Nothing is timed under stream capture (the default shape is recorded, not cached)
// or with VCBCT_NO_AUTOTUNE=1.  It replaces hand-timed launch shapes (P/reprojectionOnGPU.cu:900-983 fixes them
// per kernel).
static std::mutex g_ftune_mu;
static std::unordered_map<std::string, TileShape> g_ftune;
static std::string g_ftune_log;

std::string forward_tune_log() {
    std::lock_guard<std::mutex> lk(g_ftune_mu);
    return g_ftune_log;
}

// Block shapes (row groups x columns) of 2, 4 and 8 waves (the forward kernel), and of 4 waves (the adjoint
// kernels, built for 256 lanes).
static const int kBlocksAll[8][2] = {{1, 4}, {1, 2}, {2, 1}, {2, 2}, {4, 1}, {1, 8}, {2, 4}, {4, 2}};
static const int kBlocks4[3][2] = {{1, 4}, {2, 2}, {4, 1}};

// 1: `key` is tuned (its shape in s); 0: tune it now; -1: do not tune (s = the default shape).
static int tuned_shape(const std::string& key, hipStream_t st, TileShape& s) {
    {
        std::lock_guard<std::mutex> lk(g_ftune_mu);
        auto it = g_ftune.find(key);
        if (it != g_ftune.end()) {
            s = it->second;
            return 1;
        }
    }
    s = kDefaultTile;
    const char* off = getenv("VCBCT_NO_AUTOTUNE");
    if (off && *off == '1') return -1;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return -1;
    return 0;
}

template <class F>
static int tune_tiles(const std::string& key, hipStream_t st, const int (*blocks)[2], int nblocks, F&& launch,
                      TileShape& best) {
    hipEvent_t e0, e1;
    VCBCT_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(VCBCT_EHIP, "autotune: event creation failed");
    }
    // warm-up, timed: launches longer than 10 ms are timed once per candidate, shorter ones twice
    float warm_ms = 0.0f;
    (void)hipEventRecord(e0, st);
    int rc = launch(kDefaultTile);
    (void)hipEventRecord(e1, st);
    if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&warm_ms, e0, e1) != hipSuccess))
        rc = fail(VCBCT_EHIP, "autotune: timing failed");
    const int reps = warm_ms > 10.0f ? 1 : 2;
    std::string rep = key + ":";
    std::vector<TileShape> done;
    float best_ms = 1e30f;
    best = kDefaultTile;
    auto time_shape = [&](const TileShape& s) {
        if (rc) return;
        for (const TileShape& d : done)
            if (d.rgb == s.rgb && d.cols == s.cols && d.order == s.order && d.split == s.split) return;
        done.push_back(s);
        float ms = 0.0f;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && !rc; ++r) rc = launch(s);
        (void)hipEventRecord(e1, st);
        if (rc) return;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(VCBCT_EHIP, "autotune: timing failed");
            return;
        }
        char b[64];
        ms /= reps;
        snprintf(b, sizeof b, " %dx%d/o%d/s%d=%.3fms", s.rgb, s.cols, s.order, s.split, ms);
        rep += b;
        // a candidate replaces the best only when it is faster by more than 1 % (timing noise does not flip
        // the choice between equal shapes from run to run)
        if (ms < best_ms * 0.99f) {
            best_ms = ms;
            best = s;
        }
    };
    for (int o = 0; o < 2; ++o)
        for (int sp = 0; sp < 2; ++sp) time_shape(TileShape{kDefaultTile.rgb, kDefaultTile.cols, o, sp});
    const TileShape p = best;
    for (int i = 0; i < nblocks; ++i) time_shape(TileShape{blocks[i][0], blocks[i][1], p.order, p.split});
    const TileShape q = best;
    for (int o = 0; o < 2; ++o)
        for (int sp = 0; sp < 2; ++sp) time_shape(TileShape{q.rgb, q.cols, o, sp});
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    char b[64];
    snprintf(b, sizeof b, " -> %dx%d/o%d/s%d\n", best.rgb, best.cols, best.order, best.split);
    std::lock_guard<std::mutex> lk(g_ftune_mu);
    g_ftune[key] = best;
    g_ftune_log += rep + b;
    return 0;
}

static std::string tune_key(const char* kernel, const ProjGeom& g, int nA, const char* extra) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    char key[224];
    snprintf(key, sizeof key, "%s dev%d N%d Nz%d H%d W%d nA%d nmat%d %s", kernel, dev, g.N, g.Nz, g.H, g.W, nA,
             g.nmat, extra);
    return key;
}

static const char* vox_name(int kind) {
    return kind == VOX_F32 ? "f32" : kind == VOX_CH2 ? "ch2" : kind == VOX_CH4 ? "ch4" : "u16";
}

int launch_forward(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const Spectrum& sp,
                   double xfactor, int a0, int a1, float* out, int mode, float* aux, hipStream_t st,
                   TileShape* tile) {
    const int nA = a1 - a0;
    if (nA <= 0) return 0;
    const double inv_xf = 1.0 / xfactor;
    const bool tunable = mode == VCBCT_MODE_INTENSITY || mode == VCBCT_MODE_LOGPROJ;
    TileShape s = kDefaultTile;
    if (tunable && tile && tile->rgb > 0) {
        if (!tile_shape_valid(*tile)) return fail(VCBCT_EINVAL, "forward: invalid tile shape");
        s = *tile;
    } else if (tunable) {
        char extra[32];
        snprintf(extra, sizeof extra, "mode%d %s", mode, vox_name(vox_kind));
        const std::string key = tune_key("fwd", g, nA, extra);
        if (tuned_shape(key, st, s) == 0) {
            auto launch = [&](const TileShape& t) {
                return launch_fwd_shape(g, tables, vox, vox_kind, sp, inv_xf, a0, nA, out, mode, aux, st, t);
            };
            int rc = tune_tiles(key, st, kBlocksAll, 8, launch, s);
            if (rc) return rc;
        }
    }
    if (tile) *tile = s;
    return launch_fwd_shape(g, tables, vox, vox_kind, sp, inv_xf, a0, nA, out, mode, aux, st, s);
}

// The adjoint's kernels, one launch of shape s each (4-wave shapes).
static int launch_coef(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const Spectrum& sp,
                       double inv_xf, const float* tg, int mode, float* coef, double* loss, hipStream_t st,
                       const TileShape& s) {
    TileMap tm;
    const long nb = tile_map_host(g, g.nAng, s, tm);
    if (nb < 0) return fail(VCBCT_EINVAL, "adjoint: launch too large");
#define VCBCT_LAUNCH_COEF(M, T)                                                                               \
    hipLaunchKernelGGL((k_adjoint_coef<M, T>), dim3((unsigned)nb), dim3(256), 0, st, g, tables, (const T*)vox,  \
                       sp.mu, sp.pdf, sp.muf, sp.ch, sp.nch, sp.NE, inv_xf, sp.nl, tm, tg, coef, loss)
    if (mode == VCBCT_ADJ_LOSS) {
        if (vox_kind == VOX_CH2) VCBCT_LAUNCH_COEF(VCBCT_ADJ_LOSS, float2);
        else if (vox_kind == VOX_CH4) VCBCT_LAUNCH_COEF(VCBCT_ADJ_LOSS, Ch4);
        else VCBCT_LAUNCH_COEF(VCBCT_ADJ_LOSS, float4);
    } else {
        if (vox_kind == VOX_CH2) VCBCT_LAUNCH_COEF(VCBCT_ADJ_UPSTREAM, float2);
        else if (vox_kind == VOX_CH4) VCBCT_LAUNCH_COEF(VCBCT_ADJ_UPSTREAM, Ch4);
        else VCBCT_LAUNCH_COEF(VCBCT_ADJ_UPSTREAM, float4);
    }
#undef VCBCT_LAUNCH_COEF
    VCBCT_HIP(hipGetLastError());
    return 0;
}

static int launch_coef_multi(const ProjGeom& g, const double* tables, const void* vox, int vox_kind,
                             const SpecSet& ss, double inv_xf, const float* targets, float* coef, double* loss,
                             hipStream_t st, const TileShape& s) {
    TileMap tm;
    const long nb = tile_map_host(g, g.nAng, s, tm);
    if (nb < 0) return fail(VCBCT_EINVAL, "adjoint: launch too large");
#define VCBCT_LAUNCH_MULTI(C, T)                                                                                \
    hipLaunchKernelGGL((k_adjoint_coef_multi<C, T>), dim3((unsigned)nb), dim3(256), 0, st, g, tables,          \
                       (const T*)vox, ss, inv_xf, tm, targets, coef, loss)
    if (vox_kind == VOX_CH2) {
        if (ss.nch == 1) VCBCT_LAUNCH_MULTI(1, float2);
        else if (ss.nch == 2) VCBCT_LAUNCH_MULTI(2, float2);
        else return fail(VCBCT_EINVAL, "adjoint: %d channels in a two-channel volume", ss.nch);
    } else if (vox_kind == VOX_CH4) {
        if (ss.nch == 3) VCBCT_LAUNCH_MULTI(3, Ch4);
        else if (ss.nch == 4) VCBCT_LAUNCH_MULTI(4, Ch4);
        else return fail(VCBCT_EINVAL, "adjoint: %d channels in a four-channel volume", ss.nch);
    } else {
        switch (ss.nch) {
            case 1: VCBCT_LAUNCH_MULTI(1, float4); break;
            case 2: VCBCT_LAUNCH_MULTI(2, float4); break;
            case 3: VCBCT_LAUNCH_MULTI(3, float4); break;
            case 4: VCBCT_LAUNCH_MULTI(4, float4); break;
            case 5: VCBCT_LAUNCH_MULTI(5, float4); break;
            case 6: VCBCT_LAUNCH_MULTI(6, float4); break;
            default: return fail(VCBCT_EINVAL, "adjoint: %d channels", ss.nch);
        }
    }
#undef VCBCT_LAUNCH_MULTI
    VCBCT_HIP(hipGetLastError());
    return 0;
}

static int launch_scatter(const ProjGeom& g, const double* tables, const float* coef, int nch, float* grad_vox,
                          hipStream_t st, const TileShape& s) {
    TileMap tm;
    const long nb = tile_map_host(g, g.nAng, s, tm);
    if (nb < 0) return fail(VCBCT_EINVAL, "adjoint: launch too large");
#define VCBCT_LAUNCH_SCATTER(C)                                                                                 \
    hipLaunchKernelGGL(k_adjoint_scatter<C>, dim3((unsigned)nb), dim3(256), 0, st, g, tables, tm, coef, grad_vox)
    switch (nch) {
        case 1: VCBCT_LAUNCH_SCATTER(1); break;
        case 2: VCBCT_LAUNCH_SCATTER(2); break;
        case 3: VCBCT_LAUNCH_SCATTER(3); break;
        case 4: VCBCT_LAUNCH_SCATTER(4); break;
        case 5: VCBCT_LAUNCH_SCATTER(5); break;
        case 6: VCBCT_LAUNCH_SCATTER(6); break;
        default: return fail(VCBCT_EINVAL, "adjoint: %d channels", nch);
    }
#undef VCBCT_LAUNCH_SCATTER
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// The shape of an adjoint kernel: the caller's (forced.rgb > 0, 4 waves), else the tuned one for `key`, tuned
// now with `launch` if there is none.
template <class F>
static int adjoint_shape(const TileShape& forced, const std::string& key, hipStream_t st, F&& launch, TileShape& s,
                         bool& tuned_now) {
    tuned_now = false;
    if (forced.rgb > 0) {
        if (!tile_shape_valid(forced) || forced.rgb * forced.cols != 4)
            return fail(VCBCT_EINVAL, "adjoint: tile shape must have 4 waves");
        s = forced;
        return 0;
    }
    if (tuned_shape(key, st, s) != 0) return 0;
    tuned_now = true;
    return tune_tiles(key, st, kBlocks4, 3, launch, s);
}

// The scatter's shape, tuned on the caller's coefficients: the timing launches add into grad_vox, which is
// cleared again afterwards (nch planes of N * N * Nz).
static int scatter_shape(const ProjGeom& g, const double* tables, const float* coef, int nch, float* grad_vox,
                         hipStream_t st, const TileShape& forced, TileShape& s) {
    char extra[32];
    snprintf(extra, sizeof extra, "nch%d", nch);
    bool tuned_now = false;
    auto launch = [&](const TileShape& t) { return launch_scatter(g, tables, coef, nch, grad_vox, st, t); };
    int rc = adjoint_shape(forced, tune_key("adj-scatter", g, g.nAng, extra), st, launch, s, tuned_now);
    if (rc || !tuned_now) return rc;
    return launch_zero(grad_vox, (size_t)nch * g.N * g.N * g.Nz * sizeof(float), st);
}

int launch_adjoint(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const Spectrum& sp,
                   double xfactor, const float* target_or_gp, int mode, float* coef, float* grad_vox, double* loss,
                   hipStream_t st, const TileShape& forced) {
    const double inv_xf = 1.0 / xfactor;
    char extra[64];
    snprintf(extra, sizeof extra, "mode%d nch%d NE%d %s", mode, sp.nch, sp.NE, vox_name(vox_kind));
    TileShape sc, ss;
    bool tuned_now = false;
    // tuning launches pass no loss accumulator
    auto launch = [&](const TileShape& t) {
        return launch_coef(g, tables, vox, vox_kind, sp, inv_xf, target_or_gp, mode, coef, nullptr, st, t);
    };
    int rc = adjoint_shape(forced, tune_key("adj-coef", g, g.nAng, extra), st, launch, sc, tuned_now);
    if (!rc) rc = launch_coef(g, tables, vox, vox_kind, sp, inv_xf, target_or_gp, mode, coef, loss, st, sc);
    if (!rc) rc = scatter_shape(g, tables, coef, sp.nch, grad_vox, st, forced, ss);
    if (!rc) rc = launch_scatter(g, tables, coef, sp.nch, grad_vox, st, ss);
    return rc;
}

int launch_adjoint_multi(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const SpecSet& ss,
                         double xfactor, const float* targets, float* coef, float* grad_vox, double* loss,
                         hipStream_t st, const TileShape& forced) {
    const double inv_xf = 1.0 / xfactor;
    char extra[64];
    int ne = 0;
    for (int s = 0; s < ss.nspec; ++s) ne += ss.NE[s];
    snprintf(extra, sizeof extra, "nspec%d nch%d NE%d %s", ss.nspec, ss.nch, ne, vox_name(vox_kind));
    TileShape sc, sg;
    bool tuned_now = false;
    auto launch = [&](const TileShape& t) {
        return launch_coef_multi(g, tables, vox, vox_kind, ss, inv_xf, targets, coef, nullptr, st, t);
    };
    int rc = adjoint_shape(forced, tune_key("adj-multi", g, g.nAng, extra), st, launch, sc, tuned_now);
    if (!rc) rc = launch_coef_multi(g, tables, vox, vox_kind, ss, inv_xf, targets, coef, loss, st, sc);
    if (!rc) rc = scatter_shape(g, tables, coef, ss.nch, grad_vox, st, forced, sg);
    if (!rc) rc = launch_scatter(g, tables, coef, ss.nch, grad_vox, st, sg);
    return rc;
}

int launch_adjoint_saved(const ProjGeom& g, const double* tables, const float* Ls, const Spectrum& sp,
                         double xfactor, const float* gp, float* coef, float* grad_vox, hipStream_t st,
                         const TileShape& forced) {
    const size_t nrays = (size_t)g.nAng * g.H * g.W;
    const unsigned nbc = (unsigned)std::min<size_t>((nrays + 255) / 256, 65536);
    hipLaunchKernelGGL(k_adjoint_coef_saved, dim3(nbc), dim3(256), 0, st, g, Ls, sp.mu, sp.pdf, sp.muf, sp.ch, sp.nch,
                       sp.NE, 1.0 / xfactor, sp.nl, gp, coef);
    VCBCT_HIP(hipGetLastError());
    TileShape sg;
    int rc = scatter_shape(g, tables, coef, sp.nch, grad_vox, st, forced, sg);
    if (!rc) rc = launch_scatter(g, tables, coef, sp.nch, grad_vox, st, sg);
    return rc;
}

}  // namespace vcbct
