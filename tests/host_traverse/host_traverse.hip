// CPU build of the GPU ray traversal (virtualcbct_amd/csrc/traverse.h) for tests: each 64-row wave is
// emulated on the host (the wave-wide step range is the min/max over its lanes, as the kernel's shuffle
// reduction computes it), so the lockstep walk, the step culling and the quirk handling can be checked
// against the oracle without a GPU.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "traverse.h"

namespace {
struct HostAcc {
    const uint16_t* vol;  // reference layout [z][mat][y][x]
    int N, Nz, nmat;
    double L[VCBCT_MAX_MAT];
    int nseg;
    void step(const vcbct::Seg& a, const vcbct::Seg& b, const vcbct::Seg& c) {
        if (a.ok) seg(a.cell / Nz, a.cell % Nz, a.rr);
        if (b.ok) seg(b.cell / Nz, b.cell % Nz, b.rr);
        if (c.ok) seg(c.cell / Nz, c.cell % Nz, c.rr);
    }
    void seg(int xy, int zmc, double rr) {
        const size_t n2 = (size_t)N * N;
        const int ymc = xy / N, xmc = xy % N;
        for (int m = 0; m < nmat; ++m) L[m] += (double)vol[((size_t)zmc * nmat + m) * n2 + (size_t)ymc * N + xmc] * rr;
        ++nseg;
    }
};
}  // namespace

// 1: waves whose lanes share the box-entry choice walk with the 64-step table (as on the GPU); 0: per lane.
static int g_use_table = 1;
extern "C" void ht_set_table(int on) { g_use_table = on; }

extern "C" int ht_line_integrals(int N, double s, int Nz, double t, int H, int W, int nAng, double x_reg, double y_reg,
                                 double z_reg, double pitch, double pitch_iso, double sad, double pixnum, int nmat,
                                 const double* tab, const uint16_t* vol, double* L, int32_t* nseg) {
    vcbct::ProjGeom g;
    g.N = N; g.Nz = Nz; g.H = H; g.W = W; g.nAng = nAng; g.nmat = nmat;
    g.s = s; g.t = t; g.x_reg = x_reg; g.y_reg = y_reg; g.z_reg = z_reg;
    g.pitch = pitch; g.pitch_iso = pitch_iso; g.sad = sad; g.pixnum = pixnum;
    vcbct::derive_geom(g);
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int a = 0; a < nAng; ++a)
        for (int jj = 0; jj < W; ++jj)
            for (int rg = 0; rg < (H + 63) / 64; ++rg) {
                vcbct::RayCtx r[64];
                bool act[64];
                int kmin = INT_MAX, kmax = INT_MIN;
                for (int l = 0; l < 64; ++l) {
                    int j = rg * 64 + l;
                    act[l] = j < H;
                    vcbct::ray_setup(g, tab[a], tab[nAng + a], tab[2 * nAng + a], tab[3 * nAng + a], act[l] ? j : 0, jj, r[l]);
                    int lo, hi;
                    vcbct::z_step_range(g, r[l], act[l], lo, hi);
                    kmin = std::min(kmin, lo);
                    kmax = std::max(kmax, hi);
                }
                bool same = true;
                for (int l = 1; l < 64; ++l) same = same && r[l].swap == r[0].swap;
                vcbct::WaveLds tab;
                for (int l = 0; l < 64; ++l) {
                    if (!act[l]) continue;
                    int j = rg * 64 + l;
                    HostAcc acc{vol, N, Nz, nmat, {0, 0, 0, 0, 0, 0}, 0};
                    vcbct::traverse_lane(g, r[l], act[l], kmin, kmax, acc, (same && g_use_table) ? &tab : nullptr);
                    size_t o = (size_t)a * H * W + (size_t)(H - 1 - j) * W + jj;
                    for (int m = 0; m < nmat; ++m) L[o * nmat + m] = acc.L[m];
                    if (nseg) nseg[o] = acc.nseg;
                }
            }
    return 0;
}

namespace {
struct HostRec {
    double* out;
    int cap, n, N, Nz;
    void step(const vcbct::Seg& a, const vcbct::Seg& b, const vcbct::Seg& c) {
        if (a.ok) seg(a.cell / Nz, a.cell % Nz, a.rr);
        if (b.ok) seg(b.cell / Nz, b.cell % Nz, b.rr);
        if (c.ok) seg(c.cell / Nz, c.cell % Nz, c.rr);
    }
    void seg(int xy, int zmc, double rr) {
        const int ymc = xy / N, xmc = xy % N;
        if (n < cap) {
            out[5 * n + 0] = xmc; out[5 * n + 1] = ymc; out[5 * n + 2] = zmc; out[5 * n + 3] = rr; out[5 * n + 4] = 1;
        }
        ++n;
    }
};
}  // namespace

// Charged segments of one ray, with the wave-wide step range of its 64-row wave.
extern "C" int ht_trace_ray(int N, double s, int Nz, double t, int H, int W, int nAng, double x_reg, double y_reg,
                            double z_reg, double pitch, double pitch_iso, double sad, double pixnum, int nmat,
                            const double* tab, int a, int j, int jj, double* out, int cap) {
    vcbct::ProjGeom g;
    g.N = N; g.Nz = Nz; g.H = H; g.W = W; g.nAng = nAng; g.nmat = nmat;
    g.s = s; g.t = t; g.x_reg = x_reg; g.y_reg = y_reg; g.z_reg = z_reg;
    g.pitch = pitch; g.pitch_iso = pitch_iso; g.sad = sad; g.pixnum = pixnum;
    vcbct::derive_geom(g);
    const int rg = j / 64;
    int kmin = INT_MAX, kmax = INT_MIN;
    vcbct::RayCtx me;
    for (int l = 0; l < 64; ++l) {
        int jl = rg * 64 + l;
        bool act = jl < H;
        vcbct::RayCtx r;
        vcbct::ray_setup(g, tab[a], tab[nAng + a], tab[2 * nAng + a], tab[3 * nAng + a], act ? jl : 0, jj, r);
        int lo, hi;
        vcbct::z_step_range(g, r, act, lo, hi);
        kmin = std::min(kmin, lo);
        kmax = std::max(kmax, hi);
        if (jl == j) me = r;
    }
    HostRec rec{out, cap, 0, N, Nz};
    vcbct::traverse_lane(g, me, true, kmin, kmax, rec);
    return rec.n;
}

// Quotients a / b over n pseudo-random a (mantissas and exponents drawn from a 64-bit LCG, |a| in
// [2^-lo, 2^hi)): how many differ from the IEEE division when formed by div_known.
extern "C" long ht_div_known_mismatches(double b, long n, unsigned long long seed, int lo, int hi) {
    const double rb = 1.0 / b;
    unsigned long long x = seed * 6364136223846793005ull + 1442695040888963407ull;
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const double m = 1.0 + (double)(x >> 11) * (1.0 / 9007199254740992.0);  // [1, 2)
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const int e = (int)((x >> 33) % (unsigned)(lo + hi)) - lo;
        const double a = ldexp((x & 1) ? -m : m, e);
        if (vcbct::div_known(a, b, rb) != a / b) ++bad;
    }
    return bad;
}

// The same with a pseudo-random divisor per quotient (its reciprocal formed by one IEEE division).
extern "C" long ht_div_known_random_mismatches(long n, unsigned long long seed, int lo, int hi) {
    unsigned long long x = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto draw = [&]() {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const double m = 1.0 + (double)(x >> 11) * (1.0 / 9007199254740992.0);
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const int e = (int)((x >> 33) % (unsigned)(lo + hi)) - lo;
        return ldexp((x & 1) ? -m : m, e);
    };
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        const double a = draw(), b = draw();
        if (vcbct::div_known(a, b, 1.0 / b) != a / b) ++bad;
    }
    return bad;
}
