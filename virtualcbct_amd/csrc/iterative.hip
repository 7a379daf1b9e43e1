// The iterative-reconstruction pieces of the projector tree: the voxel-driven chord backprojection
// lengthproj (P/lengthproj.cpp:9-131) and one MLEM + total-variation update IR_ImageCBCT
// (P/IR_ImageCBCT.cpp:9-94).  Both are per-voxel FP64 computations; a thread owns one voxel (x fastest,
// so neighbouring lanes read neighbouring detector taps and image voxels) and, for lengthproj, walks the
// angles in order, accumulating in FP32 as the reference does, so the results are the reference's bit
// for bit.  The per-angle source and detector constants are computed once on the host, in the
// reference's operation order, and read as wave-uniform scalars.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {

// per-angle constants of lengthproj (host-computed)
struct LpAngle {
    double sx, sy, sz, ox, oy, c, s, xshift, yshift;
    double ko;  // (oy - sy) / (ox - sx), the detector line's slope: per angle, formed once on the host
};

// (int)v as the reference's x86 build converts it; false where that conversion yields INT_MIN (NaN or out
// of range), an index every later bounds test rejects.
__device__ __host__ inline bool trunc_index(double v, int* out) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return false;
    *out = (int)v;
    return true;
}

__global__ __launch_bounds__(256) void k_lengthproj(const LpAngle* __restrict__ tab, int a0, int a1,
                                                    const uint16_t* __restrict__ proj, float* __restrict__ npf, int N,
                                                    int Nz, double s, double t, int W, int H, double pix,
                                                    double z_reg, double x_reg) {
    const int x = blockIdx.x * 64 + threadIdx.x % 64;
    const int y = blockIdx.y * 4 + threadIdx.x / 64;
    const int z = blockIdx.z;
    if (x >= N || y >= N) return;
    const double rz = Nz * t * 0.5 - (z + 0.5) * t;
    const double ry = N * s * 0.5 - (y + 0.5) * s;
    const double rx = -N * s * 0.5 + (x + 0.5) * s;
    const int h2 = H / 2, w2 = W / 2;
    const double wh = (double)((W - H) / 2);
    const double rpix = 1.0 / pix;
    float acc = 0.0f;
    for (int a = a0; a < a1; ++a) {
        const LpAngle A = tab[a];
        double rr;
        // Quotients that share a divisor (or divide by the constant pitch) are formed from one correctly
        // rounded reciprocal with div_known: the IEEE quotients bit for bit, one division instead of two.
        // (A zero divisor gives NaN there where the division gives an infinity; it occurs only where the
        // reference's result is NaN or rejected anyway: the angle then contributes nothing on both builds.)
        if (fabs(rx - A.sx) > fabs(ry - A.sy)) {
            const double x0 = rx - 0.5 * s, x1 = rx + 0.5 * s;
            const double dx = rx - A.sx, rdx = 1.0 / dx;
            const double f0 = div_known(x0 - A.sx, dx, rdx), f1 = div_known(x1 - A.sx, dx, rdx);
            const double y0 = f0 * (ry - A.sy) + A.sy, z0 = f0 * (rz - A.sz) + A.sz;
            const double y1 = f1 * (ry - A.sy) + A.sy, z1 = f1 * (rz - A.sz) + A.sz;
            rr = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1));
        } else {
            const double y0 = ry - 0.5 * s, y1 = ry + 0.5 * s;
            const double dy = ry - A.sy, rdy = 1.0 / dy;
            const double f0 = div_known(y0 - A.sy, dy, rdy), f1 = div_known(y1 - A.sy, dy, rdy);
            const double x0 = f0 * (rx - A.sx) + A.sx, z0 = f0 * (rz - A.sz) + A.sz;
            const double x1 = f1 * (rx - A.sx) + A.sx, z1 = f1 * (rz - A.sz) + A.sz;
            rr = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1));
        }
        // meet of the source -> voxel line with the detector line through o
        const double dys = A.sy - ry, rdys = 1.0 / dys;
        const double kr = div_known(A.sx - rx, dys, rdys), ko = A.ko;
        const double py = (A.ox - rx + kr * ry + ko * A.oy) / (kr + ko);
        const double px = (kr * (py - ry) + rx);
        const double pz = (div_known(A.sz - rz, dys, rdys) * (py - ry) + rz);
        const double vv = -(pz - z_reg - A.yshift);
        const double uu = (px - x_reg) * A.c + (py - x_reg) * A.s + A.xshift - wh * pix;
        const double vfac = vv < 0 ? -0.5 : 0.5, ufac = uu < 0 ? -0.5 : 0.5;
        const double vq = div_known(vv, pix, rpix), uq = div_known(uu, pix, rpix);
        int v1, u1;
        if (!trunc_index(vq + vfac, &v1) || !trunc_index(uq + ufac, &u1)) continue;
        const double dv1 = vq - v1, du1 = uq - u1;
        const int v2 = fabs(vv - v1) < 0.000000001 ? v1 : v1 + ((vv - v1) > 0 ? 1 : -1);
        const int u2 = fabs(uu - u1) < 0.000000001 ? u1 : u1 + ((uu - u1) > 0 ? 1 : -1);
        if (v1 + h2 < 0 || v1 + h2 >= H || v2 + h2 < 0 || v2 + h2 >= H || u1 + w2 < 0 || u1 + w2 >= W || u2 + w2 < 0 ||
            u2 + w2 >= W)
            continue;
        const uint16_t* P = proj + (size_t)a * W * H;
        const double p11 = P[(v1 + h2) * W + u1 + w2], p12 = P[(v2 + h2) * W + u1 + w2];
        const double p21 = P[(v1 + h2) * W + u2 + w2], p22 = P[(v2 + h2) * W + u2 + w2];
        // the v fraction blends the u taps and the u fraction the v taps, as the reference writes it
        const double q1 = ((1 - dv1) * p11 + dv1 * p21);
        const double q2 = ((1 - dv1) * p12 + dv1 * p22);
        const double q = ((1 - du1) * q1 + du1 * q2);
        acc += (float)((65535.0 - q) * rr);
    }
    npf[((size_t)z * N + y) * N + x] = acc;
}

// The reference sums the squared differences in int; its x86 build wraps them modulo 2^32 past INT_MAX (a
// negative sum and a NaN square root follow).  The wrap is explicit here: left to signed arithmetic, the
// compiler may fold the sum into exact FP64 additions and never wrap.
__device__ inline int sq3_wrapped(int d1, int d2, int d3) {
    const uint32_t u = (uint32_t)d1 * (uint32_t)d1 + (uint32_t)d2 * (uint32_t)d2 + (uint32_t)d3 * (uint32_t)d3;
    return (int)u;
}

__device__ inline double tv_term(int a, int ax, int ay, int az, int ax1, int ay1, int az1, int ayx1, int azx1,
                                 int axy1, int azy1, int axz1, int ayz1) {
    const double d1 = sqrt(double(sq3_wrapped(a - ax, a - ay, a - az))) + 0.00000001;
    const double d2 = sqrt(double(sq3_wrapped(a - ax1, ax1 - ayx1, ax1 - azx1))) + 0.00000001;
    const double d3 = sqrt(double(sq3_wrapped(ay1 - axy1, a - ay1, ay1 - azy1))) + 0.00000001;
    const double d4 = sqrt(double(sq3_wrapped(az1 - axz1, az1 - ayz1, a - az1))) + 0.00000001;
    return ((a - ax) + (a - ay) + (a - az)) / d1 + (a - ax1) / d2 + (a - ay1) / d3 + (a - az1) / d4;
}

__global__ __launch_bounds__(256) void k_ir_update(const int16_t* __restrict__ b, const float* __restrict__ npf,
                                                   const float* __restrict__ npf_re, float* __restrict__ out, int N,
                                                   int Nz, double scale, double wbefore, double weight_tv,
                                                   double weight_prior) {
    const int x = blockIdx.x * 64 + threadIdx.x % 64;
    const int y = blockIdx.y * 4 + threadIdx.x / 64;
    const int z = blockIdx.z;
    if (x >= N || y >= N) return;
    const long long NN = (long long)N * N;
    const long long k = z * NN + (long long)y * N + x;
    double tv = 0.0, prior = 0.0;
    if (z < Nz - 1 && y < N - 1 && x < N - 1 && x > 0 && y > 0 && z > 0) {
        auto at = [&](int dz, int dy, int dx) { return (int)b[k + dz * NN + (long long)dy * N + dx]; };
        tv = tv_term(at(0, 0, 0), at(0, 0, 1), at(0, 1, 0), at(1, 0, 0), at(0, 0, -1), at(0, -1, 0), at(-1, 0, 0),
                     at(0, 1, -1), at(1, 0, -1), at(0, -1, 1), at(1, -1, 0), at(-1, 0, 1), at(-1, 1, 0));
    }
    const double lambda = weight_tv * tv + weight_prior * prior + 0.000001;
    float v = (float)(b[k] * (wbefore + (1 - wbefore) * npf_re[k] / (npf[k] + lambda)));
    if (v < 0.0) v = 0.0f;
    else if (v > 65535 / 2) v = 5000.0f;
    const double xx = -N * scale * 0.5 + (x + 0.5) * scale;
    const double yy = N * scale * 0.5 - (y + 0.5) * scale;
    const double rmax = N * 0.5 * 1.0;
    if (sqrt(xx * xx + yy * yy) >= rmax) v = 0.0f;
    out[k] = v;
}

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_projector_lengthproj(vcbct_projector* p, int ite, const uint16_t* proj, int a0, int a1, float* npf,
                                          void* stream) {
    if (!p || !proj || !npf) return fail(VCBCT_EINVAL, "lengthproj: NULL argument");
    const ProjGeom& g = p->g;
    if (a1 < 0) a1 = g.nAng;
    if (a0 < 0 || a1 > g.nAng || a0 > a1) return fail(VCBCT_EINVAL, "lengthproj: angle range [%d, %d) of %d", a0, a1, g.nAng);
    VCBCT_HIP(hipSetDevice(p->device));
    if (!p->d_lp_tab) {
        // both forms of the table: registration as given (ite == 0) and zeroed (ite != 0)
        std::vector<double> host((size_t)g.nAng * 4);
        VCBCT_HIP(hipMemcpy(host.data(), p->d_tables, host.size() * sizeof(double), hipMemcpyDeviceToHost));
        const double* cs = host.data();
        const double* sn = cs + g.nAng;
        const double* xc = sn + g.nAng;
        const double* yc = xc + g.nAng;
        std::vector<LpAngle> tab((size_t)2 * g.nAng);
        const double pix = g.pitch * (g.pixnum / g.H);
        for (int form = 0; form < 2; ++form) {
            const double xr = form ? 0.0 : g.x_reg, yr = form ? 0.0 : g.y_reg, zr = form ? 0.0 : g.z_reg;
            for (int a = 0; a < g.nAng; ++a) {
                LpAngle& A = tab[(size_t)form * g.nAng + a];
                A.sx = -g.sad * sn[a] + xr;
                A.sy = g.sad * cs[a] + yr;
                A.sz = zr;
                A.ox = A.sx + g.pitch / g.pitch_iso * sn[a] * g.sad;
                A.oy = A.sy - g.pitch / g.pitch_iso * cs[a] * g.sad;
                A.c = cs[a];
                A.s = sn[a];
                A.xshift = -(xc[a] + g.H / 2.0) * pix;
                A.yshift = -(yc[a] + g.H / 2.0) * pix;
                A.ko = (A.oy - A.sy) / (A.ox - A.sx);
            }
        }
        void* d = nullptr;
        if (dev_malloc(&d, tab.size() * sizeof(LpAngle)) != hipSuccess)
            return fail(VCBCT_ENOMEM, "lengthproj: device allocation of the angle table failed");
        hipError_t e = hipMemcpy(d, tab.data(), tab.size() * sizeof(LpAngle), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return fail(VCBCT_EHIP, "lengthproj: table upload failed: %s", hipGetErrorString(e));
        }
        p->d_lp_tab = d;
    }
    const LpAngle* tab = (const LpAngle*)p->d_lp_tab + (ite ? g.nAng : 0);
    const double pix = g.pitch * (g.pixnum / g.H);
    dim3 grid((g.N + 63) / 64, (g.N + 3) / 4, g.Nz);
    hipLaunchKernelGGL(k_lengthproj, grid, dim3(256), 0, (hipStream_t)stream, tab, a0, a1, proj, npf, g.N, g.Nz, g.s,
                       g.t, g.W, g.H, pix, ite ? 0.0 : g.z_reg, ite ? 0.0 : g.x_reg);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

extern "C" int vcbct_ir_update(int N, int Nz, double scale, const int16_t* before, const float* npf, const float* npf_re,
                               double wbefore, double weight_tv, double weight_prior, float* out, void* stream) {
    if (!before || !npf || !npf_re || !out) return fail(VCBCT_EINVAL, "ir_update: NULL argument");
    if (N < 1 || Nz < 1 || N > 65535 * 4 || Nz > 65535) return fail(VCBCT_EINVAL, "ir_update: size %d x %d", N, Nz);
    dim3 grid((N + 63) / 64, (N + 3) / 4, Nz);
    hipLaunchKernelGGL(k_ir_update, grid, dim3(256), 0, (hipStream_t)stream, before, npf, npf_re, out, N, Nz, scale,
                       wbefore, weight_tv, weight_prior);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
