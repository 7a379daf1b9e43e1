// FDK on gfx950: ramp filter as a batched FFT convolution and voxel-driven cone-beam backprojection.
//
// Replaces filterSinogram (F/filterSinogram.cpp:67-131, a direct O(W^2) spatial convolution with
// -log() inside the inner loop) and reconstructImage (F/reconstructImage.cpp:10-184, OpenMP over z).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <climits>
#include <complex>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {

static const double kPI = 3.141592653589793238462643383279;

// ------------------------------------------------------------------------------------------------
// Filter kernel h(d), F/filterSinogram.cpp:10-65 (off-centre taps) and :78-80 (centre coefficient).
// FP64, the reference's operation order.  ALP is the window parameter: 0.5 is the reference's (ALPHA,
// F/physParams.h:39-41); 1 drops the window terms, leaving the band-limited ramp (Ram-Lak), an opt-in of the
// FDK plan (vcbct_fdk_plan_set_window).
constexpr double kRefAlpha = 0.50;
static double ramp_offcentre(int md, double pitch_iso, int W, double ALP = kRefAlpha) {
    const double CUT = 1.60, GAM = 1.00;
    double theConst = 1.0 / (8.0 * kPI * kPI * pitch_iso * 800 / W);
    double gamma_x_cutoffFreq = GAM * CUT;
    double PI_x_cutoffFreq = kPI * CUT;
    double one_minus_alpha = 1.0 - ALP;
    double omega = kPI / GAM;
    double cos_omega = cos(omega), sin_omega = sin(omega);
    double MN = (double)md;
    double theta = PI_x_cutoffFreq * MN;
    double delta = gamma_x_cutoffFreq * MN;
    double eta = 1.0 / (1.0 - 1.0 / delta);
    double xi = 1.0 / (1.0 + 1.0 / delta);
    double ct = cos(theta), st = sin(theta);
    return theConst / MN / MN *
           (2.0 * ALP * theta * st + 2.0 * ALP * (ct - 1.0) + xi * one_minus_alpha * theta * (st * cos_omega + ct * sin_omega) +
            eta * one_minus_alpha * theta * (st * cos_omega - ct * sin_omega) +
            xi * xi * one_minus_alpha * (ct * cos_omega - st * sin_omega - 1.0) +
            eta * eta * one_minus_alpha * (ct * cos_omega + st * sin_omega - 1.0));
}
static double ramp_centre(int W, double pitch_iso, double ALP = kRefAlpha) {
    const double CUT = 1.60, GAM = 1.00;
    double omega = kPI / GAM;
    return CUT * CUT / 4.0 / (pitch_iso * 800 / W) *
           (ALP * 0.5 + (1.0 - ALP) * (1.0 / omega * sin(omega) + 1.0 / omega / omega * (cos(omega) - 1.0)));
}

// Host FP64 radix-2 FFT (used once per (W, pitch) to build the filter's frequency response).
static void fft64(std::vector<std::complex<double>>& a) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        double ang = -2.0 * kPI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                std::complex<double> w(cos(ang * k), sin(ang * k));
                std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

struct FilterPlan {
    int W, L, logL, r5;  // L = 2^logL, or (r5) 5 x 2^logL: the shorter one holding the 2W - 1 taps
    float2* d_H;    // [L] frequency response of the circularly wrapped h, scaled by 1/L
    float2* d_tw;   // twiddles per stage, contiguous in k (see fft_lds); L entries
};

static std::mutex g_plan_mu;
struct FilterKey {
    int dev, W;
    double pitch, alpha;
    bool operator==(const FilterKey& o) const {
        return dev == o.dev && W == o.W && pitch == o.pitch && alpha == o.alpha;
    }
};
struct FilterKeyHash {
    size_t operator()(const FilterKey& k) const {
        unsigned long long pb, ab;
        memcpy(&pb, &k.pitch, 8);
        memcpy(&ab, &k.alpha, 8);
        return (size_t)(((unsigned long long)k.dev << 56) ^ ((unsigned long long)k.W << 40) ^
                        (pb * 0x9E3779B97F4A7C15ull) ^ (ab * 0xC2B2AE3D27D4EB4Full));
    }
};
static std::unordered_map<FilterKey, FilterPlan, FilterKeyHash> g_plans;  // per device, width, pitch and window

// The transform length for a W-column panel: the circular convolution of length L equals the linear one on the
// W outputs when L >= 2W - 1; L is the shorter of the smallest power of two (at least 64) and the smallest
// 5 x 2^m (at least 80) that hold them (630 columns: 1280, not 2048).  Returns L; logL is log2 of its
// power-of-two factor, r5 whether it has the factor 5.
static int filter_length(int W, int* logL, int* r5) {
    int L = 64, lg = 6;
    while (L < 2 * W - 1) { L <<= 1; lg++; }
    int L5 = 80, m = 4;
    while (L5 < 2 * W - 1) { L5 <<= 1; m++; }
    *r5 = L5 < L;
    *logL = *r5 ? m : lg;
    return *r5 ? L5 : L;
}

static int get_plan(int W, double pitch_iso, FilterPlan** out, double alpha = kRefAlpha) {
    int dev = 0;
    VCBCT_HIP(hipGetDevice(&dev));
    const FilterKey key{dev, W, pitch_iso, alpha};
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        *out = &it->second;
        return 0;
    }
    FilterPlan P;
    P.W = W;
    P.L = filter_length(W, &P.logL, &P.r5);
    if (P.L > 4096) return fail(VCBCT_EINVAL, "filter: projection width %d too large (max 2048)", W);
    std::vector<std::complex<double>> h(P.L, 0.0);
    h[0] = ramp_centre(W, pitch_iso, alpha);
    for (int d = 1; d <= W - 1; ++d) {
        h[d] = ramp_offcentre(d, pitch_iso, W, alpha);           // out[m] gets h(d) q[m-d]
        h[P.L - d] = ramp_offcentre(-d, pitch_iso, W, alpha);
    }
    if (!P.r5) {
        fft64(h);
    } else {  // a direct FP64 DFT over the 2W - 1 nonzero taps (once per plan)
        std::vector<std::complex<double>> e(P.L), Hd(P.L, 0.0);
        for (int m = 0; m < P.L; ++m) e[m] = std::complex<double>(cos(-2.0 * kPI * m / P.L), sin(-2.0 * kPI * m / P.L));
        for (int k = 0; k < P.L; ++k) {
            std::complex<double> acc = h[0];
            for (int d = 1; d <= W - 1; ++d)
                acc += h[d] * e[(size_t)d * k % P.L] + h[P.L - d] * e[(size_t)(P.L - d) * k % P.L];
            Hd[k] = acc;
        }
        h.swap(Hd);
    }
    std::vector<float2> Hf(P.L), tw(P.L);
    for (int k = 0; k < P.L; ++k) Hf[k] = make_float2((float)(h[k].real() / P.L), (float)(h[k].imag() / P.L));
    for (int k = 0; k < P.L; ++k) tw[k] = make_float2((float)cos(-2.0 * kPI * k / P.L), (float)sin(-2.0 * kPI * k / P.L));
    // per-stage layout: radix-4 stage s (Ns = 4^s) holds w^(k r), w = exp(-2 pi i / (4 Ns)), at
    // [base_s + (r - 1) Ns + k], r = 1..3, k < Ns; a radix-2 stage (odd logL) holds w^k, w = exp(-2 pi i /
    // (2 Ns)), after them; with r5 a closing radix-5 stage (Ns = L / 5) holds w^(k r), w = exp(-2 pi i / L),
    // r = 1..4, last.  For a power of two these are tw[] entries (w^(k r) = tw[(k r) << (logL - 2 - 2s)]).
    std::vector<float2> tc(P.L, make_float2(0.f, 0.f));
    auto wr = [](double num, double den) {
        return make_float2((float)cos(-2.0 * kPI * num / den), (float)sin(-2.0 * kPI * num / den));
    };
    int base = 0, Ns = 1;
    for (int st = 0; st < P.logL / 2; ++st) {
        const int shift = P.logL - 2 - 2 * st;
        for (int r = 1; r <= 3; ++r)
            for (int k = 0; k < Ns; ++k)
                tc[base + (r - 1) * Ns + k] = P.r5 ? wr((double)k * r, 4.0 * Ns) : tw[(k * r) << shift];
        base += 3 * Ns;
        Ns <<= 2;
    }
    if (P.logL & 1) {
        for (int k = 0; k < Ns; ++k) tc[base + k] = P.r5 ? wr((double)k, 2.0 * Ns) : tw[k];
        base += Ns;
        Ns <<= 1;
    }
    if (P.r5)
        for (int r = 1; r <= 4; ++r)
            for (int k = 0; k < Ns; ++k) tc[base + (r - 1) * Ns + k] = tw[k * r];
    VCBCT_HIP(dev_malloc((void**)&P.d_H, P.L * sizeof(float2)));
    VCBCT_HIP(dev_malloc((void**)&P.d_tw, P.L * sizeof(float2)));
    VCBCT_HIP(hipMemcpy(P.d_H, Hf.data(), P.L * sizeof(float2), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(P.d_tw, tc.data(), P.L * sizeof(float2), hipMemcpyHostToDevice));
    g_plans[key] = P;
    *out = &g_plans[key];
    return 0;
}

// Two detector rows form one complex signal z = q1 + i q2 (the kernel is real, so the two convolutions
// separate into the real and imaginary parts).  Each is zero-padded to L (2^logL, or 5 x 2^logL when that is
// shorter), transformed with a Stockham FFT in LDS (radix-4 stages, one radix-2 stage when logL is odd, a
// closing radix-5 stage for 5 x 2^logL; the 256 threads share each stage's butterflies), multiplied by the
// response H, transformed back, and cosine-weighted at the output
// pixel.  A block filters FSIG row pairs in turn.  The twiddle and response tables are read from global
// memory (they stay in the caches): staging them in LDS as well doubled the block's LDS, and the occupancy it
// cost was more than the loads it saved (L = 2048: 64 KB per block, 2 blocks per CU).
constexpr int FT = 256;
constexpr int FSIG = 8;
constexpr int FPF = 8;  // pixels per thread per row held in registers: W <= FPF * FT

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// x -> DFT(x) (INV: the unscaled inverse) through the ping-pong buffers x, y; returns the buffer holding it.
// L = 2^logL, or (R5) 5 x 2^logL.  A stage of radix R whose earlier stages' radices multiply to Ns (its
// sub-transform size) has L / R butterflies: butterfly j reads x[j + r L / R], scales input r by
// w^(k r), w = exp(-2 pi i / (R Ns)), k = j mod Ns, and writes y[(j / Ns) R Ns + k + r Ns].  T holds each
// stage's twiddles contiguous in k (w^(k r) at [base + (r - 1) Ns + k], base advancing by (R - 1) Ns per
// stage), so the lanes of a wave read consecutive or equal addresses (a single exp(-2 pi i m / L) table read
// at (k r) << shift had lanes 16 to 64 float2 apart).
// Pruning, for the filter's signals (W <= L / 2 real samples, since L >= 2W - 1, and W <= L / 2 outputs
// read): the first stage (Ns = 1) has unit twiddles, and in the forward transform its inputs at L / 2 and
// beyond are zero padding, so it neither loads nor multiplies them; the inverse's last stage forms only the
// outputs below L / 2.  The values are the unpruned ones (x + 0 = x and x (1 + 0i) = x up to the sign of a
// zero), and the buffer returned by the inverse holds [0, L / 2) only.
template <bool INV, bool R5>
__device__ __forceinline__ float2* fft_lds(float2* x, float2* y, const float2* T, int L, int logL) {
    int Ns = 1, base = 0;
    const int q = L >> 2;
#pragma unroll
    for (int s = 0; s < logL / 2; ++s) {
        const float2* Ts = T + base;
        const bool first = s == 0;
        const bool last_half = INV && !R5 && !(logL & 1) && s == logL / 2 - 1;  // outputs >= L / 2 unused
        for (int j = threadIdx.x; j < q; j += FT) {
            const int k = j & (Ns - 1);
            float2 a0, a1, a2, d;
            if (first && !INV) {  // v2 = v3 = 0, unit twiddles
                const float2 v0 = x[j], v1 = x[j + q];
                a0 = v0;
                a1 = v0;
                a2 = v1;
                d = v1;
            } else {
                float2 v0 = x[j], v1 = x[j + q], v2 = x[j + 2 * q], v3 = x[j + 3 * q];
                if (!first) {
                    float2 w1 = Ts[k], w2 = Ts[Ns + k], w3 = Ts[2 * Ns + k];
                    if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                    v1 = cmul(v1, w1);
                    v2 = cmul(v2, w2);
                    v3 = cmul(v3, w3);
                }
                a0 = make_float2(v0.x + v2.x, v0.y + v2.y);
                a1 = make_float2(v0.x - v2.x, v0.y - v2.y);
                a2 = make_float2(v1.x + v3.x, v1.y + v3.y);
                d = make_float2(v1.x - v3.x, v1.y - v3.y);
            }
            const float2 a3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);  // (+-i) (v1 - v3)
            const int o = ((j - k) << 2) + k;
            y[o] = make_float2(a0.x + a2.x, a0.y + a2.y);
            y[o + Ns] = make_float2(a1.x + a3.x, a1.y + a3.y);
            if (!last_half) {
                y[o + 2 * Ns] = make_float2(a0.x - a2.x, a0.y - a2.y);
                y[o + 3 * Ns] = make_float2(a1.x - a3.x, a1.y - a3.y);
            }
        }
        __syncthreads();
        float2* t = x; x = y; y = t;
        base += 3 * Ns;
        Ns <<= 2;
    }
    if (logL & 1) {  // last stage radix 2: Ns = L / 2, w^k at [base + k]
        const int half = L >> 1;
        for (int j = threadIdx.x; j < half; j += FT) {
            const int k = j & (Ns - 1);
            const float2 a = x[j];
            float2 w = T[base + k];
            if (INV) w.y = -w.y;
            const float2 bw = cmul(x[j + half], w);
            const int o = ((j - k) << 1) + k;
            y[o] = make_float2(a.x + bw.x, a.y + bw.y);
            if (!(INV && !R5)) y[o + Ns] = make_float2(a.x - bw.x, a.y - bw.y);  // the inverse's last: >= L / 2
        }
        __syncthreads();
        float2* t = x; x = y; y = t;
        base += Ns;
        Ns <<= 1;
    }
    if (R5) {  // closing radix 5: Ns = L / 5, so k = j and the outputs go to j + r Ns
        // DFT5 with c_m = cos(2 pi m / 5), s_m = sin(2 pi m / 5): out0 = v0 + t1 + t2,
        // out1,4 = v0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4), out2,3 = v0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4)
        // (t1 = v1 + v4, t2 = v2 + v3, t3 = v1 - v4, t4 = v2 - v3; the inverse swaps the signs of i)
        const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;
        const float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;
        const float2* Ts = T + base;
        for (int j = threadIdx.x; j < Ns; j += FT) {
            float2 v[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) v[r] = x[j + r * Ns];
#pragma unroll
            for (int r = 1; r < 5; ++r) {
                float2 w = Ts[(r - 1) * Ns + j];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
            const float2 t1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y);
            const float2 t2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
            const float2 t3 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
            const float2 t4 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
            const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
            const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
            const float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
            const float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
            // -i b = (b.y, -b.x); +i b = (-b.y, b.x)
            const float2 ib1 = INV ? make_float2(-b1.y, b1.x) : make_float2(b1.y, -b1.x);
            const float2 ib2 = INV ? make_float2(-b2.y, b2.x) : make_float2(b2.y, -b2.x);
            y[j] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
            y[j + Ns] = make_float2(a1.x + ib1.x, a1.y + ib1.y);
            y[j + 2 * Ns] = make_float2(a2.x + ib2.x, a2.y + ib2.y);
            if (!INV) {  // the inverse's last stage: 3 Ns and 4 Ns are at or beyond L / 2
                y[j + 3 * Ns] = make_float2(a2.x - ib2.x, a2.y - ib2.y);
                y[j + 4 * Ns] = make_float2(a1.x - ib1.x, a1.y - ib1.y);
            }
        }
        __syncthreads();
        float2* t = x; x = y; y = t;
    }
    return x;
}

// ecol (half-fan plans): the offset weighting fused into the load.  The caller's panel has Wsrc columns,
// placed at column `off` of the W-column padded panel; a real column c of projection th contributes
// ecol[th][c] (-log p) (= -log p^(2w), the weighted pixel's line integral; exponent 0 gives 0 whatever p), a
// padding column 0 (-log 1).  HF false: no weighting, Wsrc = W and off = 0.
// LOGLC: logL as a compile-time constant for the common lengths (1024; 1280 = 5 x 256), so that the stage
// loops unroll and every stride and twiddle offset is an immediate; 0: L and logL as given at run time.
template <bool HF, bool R5, int LOGLC>
__global__ __launch_bounds__(FT) void k_filter(const float* __restrict__ proj, float* __restrict__ out, int H, int W,
                                               long nrows, int L_rt, int logL_rt, const float2* __restrict__ Hf,
                                               const float2* __restrict__ tw, const float* __restrict__ xc_tab,
                                               const float* __restrict__ yc_tab, float pitchW,
                                               const float* __restrict__ ecol, int Wsrc, int off) {
    const int L = LOGLC ? ((R5 ? 5 : 1) << LOGLC) : L_rt;
    const int logL = LOGLC ? LOGLC : logL_rt;
    // LDS: the [A][B] ping-pong buffers
    extern __shared__ float2 lds[];
    float2* A = lds;
    float2* B = lds + L;
    const float2* T = tw;
    const float2* HL = Hf;
    // the next row pair's pixels are loaded into registers while the current pair is transformed (the
    // block barriers wait for LDS traffic only, so these loads stay in flight across them)
    float pa[FPF], pb[FPF];
    const int ws = HF ? Wsrc : W, of = HF ? off : 0;
    auto prefetch = [&](long r1) {
        const float* p1 = proj + r1 * ws;
        const float* p2 = proj + (r1 + 1) * ws;
        const bool has2 = r1 + 1 < nrows;
#pragma unroll
        for (int i = 0; i < FPF; ++i) {
            const int c = threadIdx.x + i * FT - of;  // the caller's column (outside [0, ws): padding)
            pa[i] = 1.f;
            pb[i] = 1.f;
            if (c >= 0 && c < ws) {
                pa[i] = p1[c];
                if (has2) pb[i] = p2[c];
            }
        }
    };
    const long first = 2 * (long)blockIdx.x * FSIG;
    if (first < nrows) prefetch(first);
    for (int sig = 0; sig < FSIG; ++sig) {
        const long r1 = first + 2 * sig, r2 = r1 + 1;
        if (r1 >= nrows) break;  // uniform over the block
        const bool has2 = r2 < nrows;
        const float* e1 = HF ? ecol + (r1 / H) * Wsrc : nullptr;
        const float* e2 = HF ? ecol + (r2 / H) * Wsrc : nullptr;
#pragma unroll
        for (int i = 0; i < FPF; ++i) {
            const int k = threadIdx.x + i * FT;
            if (k >= W) continue;
            float q1 = -logf(pa[i]), q2 = has2 ? -logf(pb[i]) : 0.f;
            if (HF) {
                const int c = k - off;
                if (c >= 0 && c < Wsrc) {
                    const float x1 = e1[c], x2 = has2 ? e2[c] : 0.f;
                    q1 = x1 == 0.f ? 0.f : x1 * q1;
                    q2 = x2 == 0.f ? 0.f : x2 * q2;
                } else {
                    q1 = q2 = 0.f;
                }
            }
            A[k] = make_float2(q1, q2);
        }
        for (int k = W + threadIdx.x; k < (L >> 1); k += FT) A[k] = make_float2(0.f, 0.f);  // read: [0, L / 2)
        __syncthreads();
        if (sig + 1 < FSIG && r1 + 2 < nrows) prefetch(r1 + 2);
        float2* X = fft_lds<false, R5>(A, B, T, L, logL);
        float2* Y = (X == A) ? B : A;
        for (int k = threadIdx.x; k < L; k += FT) X[k] = cmul(X[k], HL[k]);
        __syncthreads();
        X = fft_lds<true, R5>(X, Y, T, L, logL);
        // cosine weight at the output pixel (F/filterSinogram.cpp:92-99, applied at :103 and :108)
        for (int s = 0; s < 2; ++s) {
            const long r = r1 + s;
            if (s == 1 && !has2) break;
            const int th = (int)(r / H), n = (int)(r % H);
            const float pos2 = fabsf(yc_tab[th] + (float)n) * pitchW;
            for (int m = threadIdx.x; m < W; m += FT) {
                float pos1 = fabsf(xc_tab[th] + (float)m) * pitchW;
                float cm = 100.0f / sqrtf(10000.0f + pos1 * pos1 + pos2 * pos2);
                float v = (s == 0) ? X[m].x : X[m].y;
                out[r * W + m] = v * cm;
            }
        }
        __syncthreads();  // X is read above; the next pair overwrites A and B
    }
}

// ------------------------------------------------------------------------------------------------
// The FP32 tap interpolation both backprojectors share (so that they agree bit for bit): a + w (b - a) as one
// FMA.  (FP32 and continuous in the row coordinate; the FP64 horizontal part keeps the reference's order.)
__device__ __forceinline__ float bp_lerp(float a, float b, float w) { return fmaf(w, b - a, a); }

// One projection as a buffer resource (GFX9 raw buffer: byte-addressed, stride 0, loads past `bytes` read 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bp_rsrc(const float* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ float bp_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// ------------------------------------------------------------------------------------------------
// Backprojection.  Thread = one (x, y) voxel column, ZB consecutive slices (ZB picked by the autotuner
// among 2, 4, 8); block = 16 x 16 columns.
// Per angle the geometry (U, its interpolation weight and base column, 1/L2) is computed once per column
// and reused for the ZB slices: U in FP64 in the reference's order, the vertical coordinate in FP32.  The
// interpolation follows F/reconstructImage.cpp:76-175: the horizontal weight is frac(U) but the base column
// is floor(W/2 + U + xshift_px - 0.5), and the three edge cases at the last column / row sample one tap pair
// or one tap.  (A precomputed per-(column, angle) table was measured slower: each z-block re-reads it.)
// A slice's taps are raw buffer loads of one projection: one 32-bit byte offset per slice (row * 4W + 4 column,
// a 24-bit multiply-add; the launcher checks H < 2^24, 4W < 2^24 and 4HW < 2^31), the next column as the
// instruction's immediate offset and the next row as its scalar offset.  The resource's range check covers
// the sum of the vector, scalar and immediate offsets (on gfx950 a load whose scalar offset alone carries it
// past the record count reads 0), so the unused next-row taps of the last row read 0 on every projection.
template <int ZB>
__global__ __launch_bounds__(256) void k_backproject(float* __restrict__ recon, const float* __restrict__ sino, int R,
                                                     int Z, double zstart, double step, double scale, int H, int W,
                                                     int n, const double* __restrict__ atab, double xfactor,
                                                     double pitchW) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z0 = blockIdx.z * ZB;
    if (x >= R || y >= R) return;
    const double rad = (double)(R - 1) / 2.0;
    if (((double)y - rad) * ((double)y - rad) + ((double)x - rad) * ((double)x - rad) > rad * rad) return;
    const double centerPos = (double)R / 2.0 - 0.5;
    const double xx = (x - centerPos) * scale;
    const double yy = -(y - centerPos) * scale;
    float zz[ZB];
    float acc[ZB];
#pragma unroll
    for (int i = 0; i < ZB; ++i) {
        zz[i] = (float)(zstart + (double)(z0 + i) * step);
        acc[i] = 0.f;
    }
    const double hW = (double)(W / 2);
    const float hH = (float)(H / 2);
    const float ipW = (float)(1.0 / pitchW);
    const double rpW = 1.0 / pitchW;  // correctly rounded: the divisions by pitchW below are exact products
    const int rowb = W * 4;           // byte offset of the next row
    const float hm1 = (float)(H - 1);  // the last row (rows are exact in FP32: H < 2^24)
    for (int th = 0; th < n; ++th) {
        const double c = atab[th], s = atab[n + th], dbeta = atab[2 * n + th], xs = atab[3 * n + th];
        // horizontal: FP64 in the reference's operation order (its base column floor(W/2 + U + xs - 0.5)
        // and weight frac(U) jump at different points, so U must round as the reference's does)
        double X = xx * c + yy * s;
        double Y = -xx * s + yy * c;
        double L2 = 1.0 - div_known(Y, 100.0, 0.01);  // = 1.0 - Y / 100.0 bit for bit
        // = X / L2 / pitchW bit for bit (|X| below ~1e3 cm and either 0 or above ~1e-17, L2 near 1)
        double U = div_known(div_unscaled(X, L2), pitchW, rpW);
        double fU = floor(U);
        double w = U - fU;
        int UU = (int)floor(hW + U + xs - 0.5);
        const bool ucol = (UU >= 0 && UU < W - 1), uedge = (UU == W - 1);
        const bool inpanel = ucol || uedge;
        // vertical: an ordinary bilinear tap, continuous in the row coordinate, and the angle weight: FP32
        const float rL2 = __builtin_amdgcn_rcpf((float)L2);
        const float iL = rL2 * ipW;
        const float wgt = rL2 * rL2 * (float)(dbeta * xfactor);
        const __amdgpu_buffer_rsrc_t P = bp_rsrc(sino + (size_t)th * H * W, H * W * 4);
        // the last column: horizontal weight 0, so the lerps keep its one tap (a + 0 (b - a) = a for finite taps;
        // the unused neighbour is the next row's first column, or 0 past the projection's end)
        const float wf = ucol ? (float)w : 0.f;
        const unsigned UU4 = inpanel ? (unsigned)UU * 4u : 0u;
        // Branch-free: the four taps are loaded unconditionally (row clamped into the panel; a tap past the
        // projection's end reads 0) and weights and selects keep the reference's cases, so for finite filtered
        // projections every kept value is the one the branching form computes, and no angle waits for the
        // previous angle's loads.
#pragma unroll
        for (int i = 0; i < ZB; ++i) {
            const float ZETA = fmaf(zz[i], iL, hH);
            const float fz = floorf(ZETA);
            const float wz = ZETA - fz;
            const float fzc = __builtin_amdgcn_fmed3f(fz, 0.f, hm1);  // the row clamped into the panel
            const int vo = (int)(__umul24((unsigned)(int)fzc, (unsigned)rowb) + UU4);
            const float a0 = bp_ld(P, vo, 0), a1 = bp_ld(P, vo + 4, 0);
            const float b0 = bp_ld(P, vo, rowb), b1 = bp_ld(P, vo + 4, rowb);
            const float h0 = bp_lerp(a0, a1, wf), h1 = bp_lerp(b0, b1, wf);
            const float sum = fz < hm1 ? bp_lerp(h0, h1, wz) : h0;  // the last row: its one row of taps
            // a tap set the reference skips (a row outside the panel: fz clamped) adds 0 x wgt (the weight is
            // always finite): acc unchanged
            acc[i] = fmaf((inpanel && fzc == fz) ? sum : 0.f, wgt, acc[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < ZB; ++i)
        if (z0 + i < Z) recon[((size_t)(z0 + i) * R + y) * R + x] += acc[i];
}


// The same backprojection with the projection data staged in LDS.  Per angle, the rows and columns that the
// block's 16 x 16 columns and ZB slices can sample (its footprint on the panel) are loaded into an LDS tile by
// the whole block, coalesced along the rows, and the taps are read from the tile.  The footprint's columns
// come from a per-(block, angle) table built with the plan: U is a linear-fractional function of (x, y), so
// over the block's square its extremes lie at the four corner voxels; the base column floor(W/2 + U + xs -
// 0.5) is taken there in the same FP64 expression, widened by one column before and two after.  Its rows
// come from the extremes of 1 / (L2 pitch) at the corners (also in the table) and the block's slice range,
// widened likewise.  A tap outside the tile (a footprint larger than the tile, or a rounding beyond the
// margins) is read from global memory at the address the unstaged kernel reads, so the result is identical
// to k_backproject's bit for bit.  Two tiles alternate, so one barrier per angle suffices: a tile is
// rewritten two angles after it was read, and every thread has passed the barrier in between.
struct BpFoot {
    int u0, u1;      // columns [u0, u1] of the footprint (clamped to the panel; u0 > u1: none)
    float ilo, ihi;  // extremes of (1 / L2) / pitch over the block's corner voxels
};
constexpr int BP_TILE = 2048;  // floats per staged tile (8 KB; two per block)

template <int ZB>
__global__ __launch_bounds__(256) void k_backproject_lds(float* __restrict__ recon, const float* __restrict__ sino,
                                                         int R, int Z, double zstart, double step, double scale,
                                                         int H, int W, int n, const double* __restrict__ atab,
                                                         double xfactor, double pitchW,
                                                         const BpFoot* __restrict__ foot) {
    __shared__ float tiles[2][BP_TILE];
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z0 = blockIdx.z * ZB;
    const double rad = (double)(R - 1) / 2.0;
    // every thread stages and meets the barriers; only the voxels inside the mask accumulate
    const bool active = x < R && y < R &&
                        ((double)y - rad) * ((double)y - rad) + ((double)x - rad) * ((double)x - rad) <= rad * rad;
    const double centerPos = (double)R / 2.0 - 0.5;
    const double xx = (x - centerPos) * scale;
    const double yy = -(y - centerPos) * scale;
    float zz[ZB];
    float acc[ZB];
#pragma unroll
    for (int i = 0; i < ZB; ++i) {
        zz[i] = (float)(zstart + (double)(z0 + i) * step);
        acc[i] = 0.f;
    }
    const double hW = (double)(W / 2);
    const float hH = (float)(H / 2);
    const float ipW = (float)(1.0 / pitchW);
    const double rpW = 1.0 / pitchW;  // correctly rounded: the divisions by pitchW below are exact products
    const BpFoot* fb = foot + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * n;
    const float zlo = zz[0], zhi = zz[ZB - 1];
    for (int th = 0; th < n; ++th) {
        const BpFoot f = fb[th];
        const float e0 = zlo * f.ilo, e1 = zlo * f.ihi, e2 = zhi * f.ilo, e3 = zhi * f.ihi;
        const int r0 = max(0, (int)floorf(hH + fminf(fminf(e0, e1), fminf(e2, e3))) - 1);
        const int r1 = min(H - 1, (int)floorf(hH + fmaxf(fmaxf(e0, e1), fmaxf(e2, e3))) + 2);
        const int tw = f.u1 - f.u0 + 1, tr = r1 - r0 + 1;
        const bool use_tile = tw > 0 && tr > 0 && tw * tr <= BP_TILE;  // uniform over the block
        float* tile = tiles[th & 1];
        const float* P = sino + (size_t)th * H * W;
        if (use_tile) {
            for (int i = threadIdx.x; i < tw * tr; i += 256) {
                const int rr = i / tw, cc = i - rr * tw;
                tile[i] = P[(size_t)(r0 + rr) * W + f.u0 + cc];
            }
        }
        __syncthreads();
        if (!active) continue;
        const double c = atab[th], s = atab[n + th], dbeta = atab[2 * n + th], xs = atab[3 * n + th];
        double X = xx * c + yy * s;
        double Y = -xx * s + yy * c;
        double L2 = 1.0 - div_known(Y, 100.0, 0.01);  // = 1.0 - Y / 100.0 bit for bit
        // = X / L2 / pitchW bit for bit (|X| below ~1e3 cm and either 0 or above ~1e-17, L2 near 1)
        double U = div_known(div_unscaled(X, L2), pitchW, rpW);
        double fU = floor(U);
        double w = U - fU;
        int UU = (int)floor(hW + U + xs - 0.5);
        const bool ucol = (UU >= 0 && UU < W - 1), uedge = (UU == W - 1);
        if (!ucol && !uedge) continue;
        const float rL2 = __builtin_amdgcn_rcpf((float)L2);
        const float iL = rL2 * ipW;
        const float wgt = rL2 * rL2 * (float)(dbeta * xfactor);
        const float wf = (float)w;
        auto tap = [&](int r, int col) -> float {
            const int a = r - r0, b = col - f.u0;
            return (use_tile && (unsigned)a < (unsigned)tr && (unsigned)b < (unsigned)tw) ? tile[a * tw + b]
                                                                                      : P[(size_t)r * W + col];
        };
#pragma unroll
        for (int i = 0; i < ZB; ++i) {
            const float ZETA = fmaf(zz[i], iL, hH);
            const float fz = floorf(ZETA);
            const int NZ = (int)fz;
            const float wz = ZETA - fz;
            float sum;
            if (NZ >= 0 && NZ < H - 1) {
                if (ucol) {
                    float a0 = tap(NZ, UU), a1 = tap(NZ, UU + 1), b0 = tap(NZ + 1, UU), b1 = tap(NZ + 1, UU + 1);
                    sum = bp_lerp(bp_lerp(a0, a1, wf), bp_lerp(b0, b1, wf), wz);
                } else {
                    sum = bp_lerp(tap(NZ, UU), tap(NZ + 1, UU), wz);
                }
            } else if (NZ == H - 1) {
                sum = ucol ? bp_lerp(tap(NZ, UU), tap(NZ, UU + 1), wf) : tap(NZ, UU);
            } else {
                continue;
            }
            acc[i] = fmaf(sum, wgt, acc[i]);
        }
    }
    if (!active) return;
#pragma unroll
    for (int i = 0; i < ZB; ++i)
        if (z0 + i < Z) recon[((size_t)(z0 + i) * R + y) * R + x] += acc[i];
}

// Footprint table of the staged backprojector: per (16 x 16 column block, angle), in block-row order.
static std::vector<BpFoot> bp_foot_table(int R, int W, int n, double scale, double pitch_iso, const double* tab) {
    const int nb = (R + 15) / 16;
    const double pitchW = pitch_iso * 800 / W;
    const double hW = (double)(W / 2);
    const float ipW = (float)(1.0 / pitchW);
    const double centerPos = (double)R / 2.0 - 0.5;
    std::vector<BpFoot> f((size_t)nb * nb * n);
    for (int by = 0; by < nb; ++by)
        for (int bx = 0; bx < nb; ++bx)
            for (int th = 0; th < n; ++th) {
                const double c = tab[th], s = tab[n + th], xs = tab[3 * n + th];
                int umin = INT_MAX, umax = INT_MIN;
                float ilo = 1e30f, ihi = -1e30f;
                for (int cy = 0; cy < 2; ++cy)
                    for (int cx = 0; cx < 2; ++cx) {
                        const int x = cx ? std::min(bx * 16 + 15, R - 1) : bx * 16;
                        const int y = cy ? std::min(by * 16 + 15, R - 1) : by * 16;
                        const double xx = (x - centerPos) * scale;
                        const double yy = -(y - centerPos) * scale;
                        double X = xx * c + yy * s;
                        double Y = -xx * s + yy * c;
                        double L2 = 1.0 - Y / 100.0;
                        double U = X / L2 / pitchW;
                        const int UU = (int)floor(hW + U + xs - 0.5);
                        umin = std::min(umin, UU);
                        umax = std::max(umax, UU);
                        const float iL = (1.0f / (float)L2) * ipW;
                        ilo = std::min(ilo, iL);
                        ihi = std::max(ihi, iL);
                    }
                BpFoot& e = f[((size_t)by * nb + bx) * n + th];
                e.u0 = std::max(0, umin - 1);
                e.u1 = std::min(W - 1, umax + 2);
                e.ilo = ilo;
                e.ihi = ihi;
            }
    return f;
}

// ------------------------------------------------------------------------------------------------
// Plans: per-geometry device tables built once, so that execution allocates nothing and synchronises
// nothing (it can be captured in a hipGraph).

static std::vector<double> bp_table(int n, int W, const double* angles, const double* xshift) {
    // per-angle FP64 table: cos, sin, dbeta (F/reconstructImage.cpp:44-71), xshift in detector pixels
    std::vector<double> tab(4 * (size_t)n);
    for (int th = 0; th < n; ++th) {
        double dbeta, thdb0, thdb1, thdb2;
        if (th == 0) {
            thdb1 = angles[th];
            thdb2 = angles[th + 1];
            dbeta = fabs(thdb2 - thdb1) / 2;
        } else if (th == n - 1) {
            thdb0 = angles[th - 1];
            thdb1 = angles[th];
            dbeta = fabs(thdb1 - thdb0) / 2;
        } else {
            thdb0 = angles[th - 1];
            thdb2 = angles[th + 1];
            thdb1 = angles[th + 1] - 2 * kPI;
            if (thdb2 < 0) thdb1 = angles[th + 1] + 2 * kPI;
            dbeta = fabs(thdb2 - thdb0) / 2;
            if (dbeta > kPI / 2) dbeta = fabs(thdb1 - thdb0) / 2;
        }
        tab[th] = cos(angles[th]);
        tab[n + th] = sin(angles[th]);
        tab[2 * n + th] = dbeta;
        tab[3 * n + th] = xshift[th] / (0.05 * 800 / W);
    }
    return tab;
}

static std::vector<float> filter_table(int n, int W, const double* xshift, const double* yshift) {
    // per-projection centre offsets, F/filterSinogram.cpp:86-87 (both with the projection WIDTH)
    std::vector<float> tab(2 * (size_t)n);
    for (int th = 0; th < n; ++th) {
        tab[th] = (float)(-W / 2.0 - xshift[th] / (0.05 * 800 / W) - 1);
        tab[n + th] = (float)(-W / 2.0 - yshift[th] / (0.05 * 800 / W) - 1);
    }
    return tab;
}

// ecol non-null: the half-fan weighting fused into the filter (k_filter), the caller's panel Wsrc columns wide at
// column `off` of the W-column padded panel.
static int launch_filter(const FilterPlan* P, int H, int W, int n, const float* dp, float* dq, const float* d_ctab,
                         double pitch_iso, hipStream_t st, const float* ecol = nullptr, int Wsrc = 0, int off = 0) {
    const long nrows = (long)n * H;
    const long npairs = (nrows + 1) / 2;
    const long nblk = (npairs + FSIG - 1) / FSIG;
    if (nblk > INT_MAX) return fail(VCBCT_EINVAL, "filter_sinogram: too many rows");
    const float pitchW = (float)(pitch_iso * 800 / W);
    const size_t sig_bytes = 2 * (size_t)P->L * sizeof(float2);  // the two signal buffers
#define VCBCT_FILT(HFV, R5V, LG, EC, WS, OF)                                                                          \
    hipLaunchKernelGGL((k_filter<HFV, R5V, LG>), dim3((unsigned)nblk), dim3(FT), sig_bytes, st, dp, dq, H, W, nrows,  \
                       P->L, P->logL, P->d_H, P->d_tw, d_ctab, d_ctab + n, pitchW, EC, WS, OF)
    // the lengths with a compile-time instantiation: 1024 (400 columns) and 1280 (630 columns)
    const bool l1024 = !P->r5 && P->logL == 10, l1280 = P->r5 && P->logL == 8;
    if (ecol) {
        if (l1280) VCBCT_FILT(true, true, 8, ecol, Wsrc, off);
        else if (l1024) VCBCT_FILT(true, false, 10, ecol, Wsrc, off);
        else if (P->r5) VCBCT_FILT(true, true, 0, ecol, Wsrc, off);
        else VCBCT_FILT(true, false, 0, ecol, Wsrc, off);
    } else {
        if (l1280) VCBCT_FILT(false, true, 8, nullptr, W, 0);
        else if (l1024) VCBCT_FILT(false, false, 10, nullptr, W, 0);
        else if (P->r5) VCBCT_FILT(false, true, 0, nullptr, W, 0);
        else VCBCT_FILT(false, false, 0, nullptr, W, 0);
    }
#undef VCBCT_FILT
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// zb: slices per thread; d_foot non-null: the LDS-staged form with that footprint table.
static int launch_bp(int zb, int R, int Z, int zstart, double step, double scale, float* drec, int H, int W, int n,
                     const float* ds, const double* d_atab, int xfactor, double pitch_iso, hipStream_t st,
                     const BpFoot* d_foot = nullptr) {
    dim3 grid((R + 15) / 16, (R + 15) / 16, (Z + zb - 1) / zb);
    const double pitchW = pitch_iso * 800 / W;
    // the taps' 24-bit row multiply and 32-bit byte offsets within one projection
    if (H >= (1 << 24) || W >= (1 << 22) || (long long)H * W * 4 >= (1ll << 31))
        return fail(VCBCT_EINVAL, "reconstruct_image: projection too large (H W 4 bytes must stay below 2^31)");
#define VCBCT_BP(ZBV)                                                                                                  \
    hipLaunchKernelGGL(k_backproject<ZBV>, grid, dim3(256), 0, st, drec, ds, R, Z, (double)zstart, step, scale, H, W, n, \
                       d_atab, (double)xfactor, pitchW)
#define VCBCT_BPL(ZBV)                                                                                                 \
    hipLaunchKernelGGL(k_backproject_lds<ZBV>, grid, dim3(256), 0, st, drec, ds, R, Z, (double)zstart, step, scale, H, \
                       W, n, d_atab, (double)xfactor, pitchW, d_foot)
    if (d_foot) {
        if (zb == 2) VCBCT_BPL(2);
        else if (zb == 8) VCBCT_BPL(8);
        else VCBCT_BPL(4);
    } else {
        if (zb == 2) VCBCT_BP(2);
        else if (zb == 8) VCBCT_BP(8);
        else VCBCT_BP(4);
    }
#undef VCBCT_BP
#undef VCBCT_BPL
    VCBCT_HIP(hipGetLastError());
    return 0;
}

// Autotuner cache: slices per thread of the backprojector, per (R, Z, n, H, W).
static std::mutex g_tune_mu;
static std::unordered_map<std::string, int> g_bp_zb;
static std::string g_tune_log;

}  // namespace vcbct

struct vcbct_fdk_plan {
    int device, H, W, n, R, Z, zstart, xfactor, zb;
    double step, scale, pitch;
    vcbct::FilterPlan* fp;
    float* d_ctab;
    double* d_atab;
    // half-fan plans: W is the padded panel; the caller's projections have hf_W columns at column hf_off of
    // it, weighted inside the filter's load with the exponents d_ecol [n][hf_W]
    int hf_W, hf_off;
    float* d_ecol;
    // backprojector form: LDS-staged (with its footprint table) or not, picked with zb by the autotuner
    int bp_lds;
    vcbct::BpFoot* d_foot;
};

namespace vcbct {

static int tune_bp(vcbct_fdk_plan* p) {
    // VCBCT_BP_LDS=0 / 1 restricts the candidates to the unstaged / LDS-staged form
    const char* lds_env = getenv("VCBCT_BP_LDS");
    const int lds_only = (lds_env && (*lds_env == '0' || *lds_env == '1')) ? *lds_env - '0' : -1;
    char key[160];
    snprintf(key, sizeof key, "bp R%d Z%d n%d H%d W%d lds%d", p->R, p->Z, p->n, p->H, p->W, lds_only);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_bp_zb.find(key);
        if (it != g_bp_zb.end()) {
            p->zb = it->second % 100;
            p->bp_lds = it->second / 100;
            return 0;
        }
    }
    const char* off = getenv("VCBCT_NO_AUTOTUNE");
    if (off && *off == '1') {
        p->zb = 4;
        p->bp_lds = lds_only == 1 ? 1 : 0;
        return 0;
    }
    float *ds = nullptr, *dr = nullptr;
    const size_t sb = (size_t)p->n * p->H * p->W * sizeof(float), rb = (size_t)p->Z * p->R * p->R * sizeof(float);
    if (dev_malloc((void**)&ds, sb) != hipSuccess || dev_malloc((void**)&dr, rb) != hipSuccess) {
        (void)hipFree(ds);
        return fail(VCBCT_ENOMEM, "autotune buffers");
    }
    int rc = launch_zero(ds, sb, nullptr);
    if (!rc) rc = launch_zero(dr, rb, nullptr);
    hipEvent_t e0, e1;
    VCBCT_HIP(hipEventCreate(&e0));
    VCBCT_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_code = 4;
    std::string rep = std::string(key) + ":";
    for (int lds = 0; lds < 2; ++lds) {
        if (lds_only >= 0 && lds != lds_only) continue;
        const BpFoot* foot = lds ? p->d_foot : nullptr;
        for (int zb : {2, 4, 8}) {
            if (rc) break;
            rc = launch_bp(zb, p->R, p->Z, p->zstart, p->step, p->scale, dr, p->H, p->W, p->n, ds, p->d_atab,
                           p->xfactor, p->pitch, nullptr, foot);  // warm
            float ms = 0;
            (void)hipEventRecord(e0, nullptr);
            for (int r = 0; r < 2 && !rc; ++r)
                rc = launch_bp(zb, p->R, p->Z, p->zstart, p->step, p->scale, dr, p->H, p->W, p->n, ds, p->d_atab,
                               p->xfactor, p->pitch, nullptr, foot);
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
            char b[64];
            snprintf(b, sizeof b, " %szb%d=%.3fms", lds ? "lds-" : "", zb, ms / 2);
            rep += b;
            if (ms < best) {
                best = ms;
                best_code = zb + 100 * lds;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(ds);
    (void)hipFree(dr);
    if (rc) return rc;
    p->zb = best_code % 100;
    p->bp_lds = best_code / 100;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_bp_zb[key] = best_code;
    g_tune_log += rep + " -> " + (p->bp_lds ? "lds-" : "") + "zb" + std::to_string(p->zb) + "\n";
    return 0;
}

const char* tune_log() { return g_tune_log.c_str(); }

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_fdk_plan_create(vcbct_fdk_plan** out, int H, int W, int n, int R, int Z, int zstart, double step,
                                     double scale, int xfactor, const double* angles, const double* xshift,
                                     const double* yshift, double pitch_iso, int device) {
    if (!out) return fail(VCBCT_EINVAL, "fdk_plan_create: out is NULL");
    *out = nullptr;
    if (H < 2 || W < 2 || n < 2 || R < 1 || Z < 1 || !angles || !xshift || !yshift)
        return fail(VCBCT_EINVAL, "fdk_plan_create: bad arguments");
    VCBCT_HIP(hipSetDevice(device));
    vcbct_fdk_plan* p = (vcbct_fdk_plan*)calloc(1, sizeof(vcbct_fdk_plan));
    if (!p) return fail(VCBCT_ENOMEM, "host allocation failed");
    p->device = device; p->H = H; p->W = W; p->n = n; p->R = R; p->Z = Z; p->zstart = zstart; p->xfactor = xfactor;
    p->step = step; p->scale = scale; p->pitch = pitch_iso; p->zb = 4;
    int rc = get_plan(W, pitch_iso, &p->fp);
    std::vector<float> ct = filter_table(n, W, xshift, yshift);
    std::vector<double> at = bp_table(n, W, angles, xshift);
    if (!rc && (dev_malloc((void**)&p->d_ctab, ct.size() * sizeof(float)) != hipSuccess ||
                dev_malloc((void**)&p->d_atab, at.size() * sizeof(double)) != hipSuccess))
        rc = fail(VCBCT_ENOMEM, "fdk plan tables");
    if (!rc && (hipMemcpy(p->d_ctab, ct.data(), ct.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(p->d_atab, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(VCBCT_EHIP, "fdk plan upload");
    if (!rc) {  // footprint table of the LDS-staged backprojector
        std::vector<BpFoot> ft = bp_foot_table(R, W, n, scale, pitch_iso, at.data());
        if (dev_malloc((void**)&p->d_foot, ft.size() * sizeof(BpFoot)) != hipSuccess)
            rc = fail(VCBCT_ENOMEM, "fdk plan footprint table");
        else if (hipMemcpy(p->d_foot, ft.data(), ft.size() * sizeof(BpFoot), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(VCBCT_EHIP, "fdk plan upload");
    }
    if (!rc) rc = tune_bp(p);
    if (rc) {
        (void)hipFree(p->d_ctab);
        (void)hipFree(p->d_atab);
        (void)hipFree(p->d_foot);
        free(p);
        return rc;
    }
    *out = p;
    return 0;
}

extern "C" int vcbct_fdk_plan_create_half_fan(vcbct_fdk_plan** out, int H, int W, int n, int R, int Z, int zstart,
                                              double step, double scale, int xfactor, const double* angles,
                                              const double* xshift, const double* yshift, double pitch_iso,
                                              double taper_px, int device) {
    if (!out) return fail(VCBCT_EINVAL, "fdk_plan_create_half_fan: out is NULL");
    *out = nullptr;
    if (H < 2 || W < 2 || W > 4096 || n < 2 || n > 65535 || !angles || !xshift || !yshift)
        return fail(VCBCT_EINVAL, "fdk_plan_create_half_fan: bad arguments");
    std::vector<OffsetPar> par(n);
    std::vector<double> xs(n), ys(n);
    int Wo, off;
    int rc = half_fan_params(W, n, xshift, yshift, taper_px, Wo, off, par.data(), xs.data(), ys.data());
    if (rc) return rc;
    // the plan proper runs on the padded panel: same pixel size, shifts that keep the central ray
    vcbct_fdk_plan* p = nullptr;
    rc = vcbct_fdk_plan_create(&p, H, Wo, n, R, Z, zstart, step, scale, xfactor, angles, xs.data(), ys.data(),
                               pitch_iso * Wo / W, device);
    if (rc) return rc;
    p->hf_W = W;
    p->hf_off = off;
    std::vector<float> ecol((size_t)n * W);
    offset_exponent_table(par.data(), n, W, ecol.data());
    if (dev_malloc((void**)&p->d_ecol, ecol.size() * sizeof(float)) != hipSuccess)
        rc = fail(VCBCT_ENOMEM, "fdk_plan_create_half_fan: device allocation of the weighting table failed");
    if (!rc && hipMemcpy(p->d_ecol, ecol.data(), ecol.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(VCBCT_EHIP, "fdk_plan_create_half_fan: upload failed");
    if (rc) {
        vcbct_fdk_plan_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

extern "C" int vcbct_fdk_plan_width(const vcbct_fdk_plan* p) { return p ? p->W : -1; }

extern "C" int vcbct_fdk_plan_set_window(vcbct_fdk_plan* p, double alpha) {
    if (!p) return fail(VCBCT_EINVAL, "fdk_plan_set_window: plan is NULL");
    if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(VCBCT_EINVAL, "fdk_plan_set_window: alpha %g not in [0, 1]", alpha);
    VCBCT_HIP(hipSetDevice(p->device));
    FilterPlan* fp = nullptr;
    int rc = get_plan(p->W, p->pitch, &fp, alpha);
    if (rc) return rc;
    p->fp = fp;
    return 0;
}

extern "C" void vcbct_fdk_plan_destroy(vcbct_fdk_plan* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipFree(p->d_ecol);
    (void)hipFree(p->d_foot);
    (void)hipFree(p->d_ctab);
    (void)hipFree(p->d_atab);
    free(p);
}

extern "C" int vcbct_fdk_plan_filter(vcbct_fdk_plan* p, const float* dev_proj, float* dev_filtered, void* stream) {
    if (!p || !dev_proj || !dev_filtered) return fail(VCBCT_EINVAL, "fdk_plan_filter: NULL argument");
    if (p->d_ecol)  // half-fan: weighted and padded inside the filter's load
        return launch_filter(p->fp, p->H, p->W, p->n, dev_proj, dev_filtered, p->d_ctab, p->pitch,
                             (hipStream_t)stream, p->d_ecol, p->hf_W, p->hf_off);
    return launch_filter(p->fp, p->H, p->W, p->n, dev_proj, dev_filtered, p->d_ctab, p->pitch, (hipStream_t)stream);
}

extern "C" int vcbct_fdk_plan_backproject(vcbct_fdk_plan* p, const float* dev_filtered, float* dev_recon, void* stream) {
    if (!p || !dev_filtered || !dev_recon) return fail(VCBCT_EINVAL, "fdk_plan_backproject: NULL argument");
    return launch_bp(p->zb, p->R, p->Z, p->zstart, p->step, p->scale, dev_recon, p->H, p->W, p->n, dev_filtered,
                     p->d_atab, p->xfactor, p->pitch, (hipStream_t)stream, p->bp_lds ? p->d_foot : nullptr);
}

extern "C" int vcbct_fdk_plan_bp_staged(const vcbct_fdk_plan* p) { return p ? p->bp_lds : -1; }

extern "C" int vcbct_fdk_plan_execute(vcbct_fdk_plan* p, const float* dev_proj, float* dev_filtered, float* dev_recon,
                                      void* stream) {
    int rc = vcbct_fdk_plan_filter(p, dev_proj, dev_filtered, stream);
    if (rc) return rc;
    return vcbct_fdk_plan_backproject(p, dev_filtered, dev_recon, stream);
}

extern "C" int vcbct_fdk_plan_slices_per_thread(const vcbct_fdk_plan* p) { return p ? p->zb : -1; }


extern "C" int vcbct_prepare_const_map(int W, double pitch_iso, double* out) {
    if (W < 1 || !out) return fail(VCBCT_EINVAL, "prepare_const_map: bad arguments");
    for (int m = 0; m < W; ++m)
        for (int n = 0; n < W; ++n) out[(size_t)m * W + n] = (m == n) ? 0.0 : ramp_offcentre(m - n, pitch_iso, W);
    return 0;
}

extern "C" int vcbct_filter_fft_length(int W) {
    if (W < 1 || W > 2048) return fail(VCBCT_EINVAL, "filter_fft_length: projection width %d outside 1..2048", W);
    int lg = 0, r5 = 0;
    return filter_length(W, &lg, &r5);
}

extern "C" int vcbct_filter_sinogram(int H, int W, int n, const float* dev_proj, float* dev_out, const double* xshift,
                                     const double* yshift, double pitch_iso, void* stream) {
    if (H < 1 || W < 2 || n < 1 || !dev_proj || !dev_out || !xshift || !yshift)
        return fail(VCBCT_EINVAL, "filter_sinogram: bad arguments");
    FilterPlan* P;
    int rc = get_plan(W, pitch_iso, &P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<float> tab = filter_table(n, W, xshift, yshift);
    float* d_tab;
    VCBCT_HIP(dev_malloc((void**)&d_tab, tab.size() * sizeof(float)));
    VCBCT_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, st));
    rc = launch_filter(P, H, W, n, dev_proj, dev_out, d_tab, pitch_iso, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_tab);
    return rc;
}

extern "C" int vcbct_reconstruct_image(int R, int Z, int zstart, double step, double scale, float* dev_recon, int H,
                                       int W, int n, const float* dev_sino, int xfactor, const double* angles,
                                       const double* xshift, double pitch_iso, void* stream) {
    if (R < 1 || Z < 1 || H < 2 || W < 2 || n < 2 || !dev_recon || !dev_sino || !angles || !xshift)
        return fail(VCBCT_EINVAL, "reconstruct_image: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> tab = bp_table(n, W, angles, xshift);
    double* d_tab;
    VCBCT_HIP(dev_malloc((void**)&d_tab, tab.size() * sizeof(double)));
    VCBCT_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    int zb = 4;
    {
        char key[128];
        snprintf(key, sizeof key, "bp R%d Z%d n%d H%d W%d", R, Z, n, H, W);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_bp_zb.find(key);
        if (it != g_bp_zb.end()) zb = it->second;
    }
    int rc = launch_bp(zb, R, Z, zstart, step, scale, dev_recon, H, W, n, dev_sino, d_tab, xfactor, pitch_iso, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_tab);
    return rc;
}

// ------------------------------------------------------------------------------------------------
// Allocation-free backprojection (no reference counterpart): vcbct_reconstruct_image with its per-call
// angle table (cos, sin, dbeta, xshift in pixels; 4 doubles per projection) built by a kernel into a
// file-scope device array instead of a hipMalloc'd buffer, for callers that cannot allocate.  The array
// lives in every GPU the library runs on until the process ends; its device address is looked up once
// per device (host code never uses the symbol itself: the name stands for a host-side shadow).

namespace vcbct {
constexpr int NOALLOC_MAX_PROJ = 16384;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[4 * NOALLOC_MAX_PROJ];  // 512 KiB per device
static std::mutex g_scratch_mu;
static double* g_scratch_dev[NOALLOC_MAX_DEVICES];  // device address of `scratch` per device, 0 until looked up

// The current device's copy of `scratch` (hipGetSymbolAddress on its first use on that device).
static int scratch_on_current_device(double** out) {
    int dev = -1;
    VCBCT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(VCBCT_EINVAL, "noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VCBCT_EHIP, "noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

// The per-angle table of bp_table, one thread per projection, from device copies of the angles and shifts
// (same expressions; FP64 division and the neighbour rule of F/reconstructImage.cpp:44-71).
__global__ __launch_bounds__(256) void k_bp_table(const double* __restrict__ angles, const double* __restrict__ xshift,
                                                  int n, int W, double* __restrict__ tab) {
    const int th = blockIdx.x * 256 + threadIdx.x;
    if (th >= n) return;
    const double kp = 3.141592653589793238462643383279;
    double dbeta;
    if (th == 0) {
        dbeta = fabs(angles[1] - angles[0]) / 2;
    } else if (th == n - 1) {
        dbeta = fabs(angles[th] - angles[th - 1]) / 2;
    } else {
        const double thdb0 = angles[th - 1], thdb2 = angles[th + 1];
        double thdb1 = angles[th + 1] - 2 * kp;
        if (thdb2 < 0) thdb1 = angles[th + 1] + 2 * kp;
        dbeta = fabs(thdb2 - thdb0) / 2;
        if (dbeta > kp / 2) dbeta = fabs(thdb1 - thdb0) / 2;
    }
    tab[th] = cos(angles[th]);
    tab[n + th] = sin(angles[th]);
    tab[2 * n + th] = dbeta;
    tab[3 * n + th] = xshift[th] / (0.05 * 800 / W);
}
}  // namespace vcbct

extern "C" int vcbct_reconstruct_image_noalloc(int R, int Z, int zstart, double step, double scale, float* dev_recon,
                                               int H, int W, int n, const float* dev_sino, int xfactor,
                                               const double* dev_angles, const double* dev_xshift, double pitch_iso,
                                               void* stream) {
    if (R < 1 || Z < 1 || H < 2 || W < 2 || n < 2 || !dev_recon || !dev_sino || !dev_angles || !dev_xshift)
        return fail(VCBCT_EINVAL, "reconstruct_image_noalloc: bad arguments");
    if (n > NOALLOC_MAX_PROJ)
        return fail(VCBCT_EINVAL, "reconstruct_image_noalloc: %d projections exceed the maximum of %d", n,
                    NOALLOC_MAX_PROJ);
    double* tab = nullptr;
    int rc = scratch_on_current_device(&tab);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_bp_table, dim3((n + 255) / 256), dim3(256), 0, st, dev_angles, dev_xshift, n, W, tab);
    VCBCT_HIP(hipGetLastError());
    // four slices per thread (the kernels give the same sums for every slice count)
    return launch_bp(4, R, Z, zstart, step, scale, dev_recon, H, W, n, dev_sino, tab, xfactor, pitch_iso, st);
}

extern "C" int vcbct_filter_sinogram_host(int H, int W, int n, const float* hp, float* ho, const double* xs,
                                          const double* ys, double pitch_iso) {
    if (!hp || !ho) return fail(VCBCT_EINVAL, "filter_sinogram_host: NULL argument");
    const size_t bytes = (size_t)n * H * W * sizeof(float);
    float *dp = nullptr, *dq = nullptr;
    if (dev_malloc((void**)&dp, bytes) != hipSuccess || dev_malloc((void**)&dq, bytes) != hipSuccess) {
        (void)hipFree(dp);
        return fail(VCBCT_ENOMEM, "device allocation of %zu bytes failed", 2 * bytes);
    }
    int rc = 0;
    if (hipMemcpy(dp, hp, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(VCBCT_EHIP, "upload failed");
    if (!rc) rc = vcbct_filter_sinogram(H, W, n, dp, dq, xs, ys, pitch_iso, nullptr);
    if (!rc && hipMemcpy(ho, dq, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(VCBCT_EHIP, "download failed");
    (void)hipFree(dp);
    (void)hipFree(dq);
    return rc;
}

extern "C" int vcbct_reconstruct_image_host(int R, int Z, int zstart, double step, double scale, double* hrec, int H,
                                            int W, int n, const float* hs, int xfactor, const double* angles,
                                            const double* xs, double pitch_iso) {
    if (!hrec || !hs) return fail(VCBCT_EINVAL, "reconstruct_image_host: NULL argument");
    const size_t sb = (size_t)n * H * W * sizeof(float), vn = (size_t)Z * R * R;
    float *ds = nullptr, *dr = nullptr;
    if (dev_malloc((void**)&ds, sb) != hipSuccess || dev_malloc((void**)&dr, vn * sizeof(float)) != hipSuccess) {
        (void)hipFree(ds);
        return fail(VCBCT_ENOMEM, "device allocation failed");
    }
    int rc = 0;
    if (hipMemcpy(ds, hs, sb, hipMemcpyHostToDevice) != hipSuccess) rc = fail(VCBCT_EHIP, "upload failed");
    if (!rc) rc = launch_zero(dr, vn * sizeof(float), nullptr);
    if (!rc) rc = vcbct_reconstruct_image(R, Z, zstart, step, scale, dr, H, W, n, ds, xfactor, angles, xs, pitch_iso, nullptr);
    std::vector<float> tmp;
    if (!rc) {
        tmp.resize(vn);
        if (hipMemcpy(tmp.data(), dr, vn * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(VCBCT_EHIP, "download failed");
    }
    if (!rc)
        for (size_t i = 0; i < vn; ++i) hrec[i] += tmp[i];
    (void)hipFree(ds);
    (void)hipFree(dr);
    return rc;
}
