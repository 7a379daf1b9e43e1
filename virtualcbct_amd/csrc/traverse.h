// Ray traversal of the material-based polychromatic forward projector, shared by the forward,
// line-integral, log-projection and adjoint kernels.
//
// Semantics follow reprojectionKernel (P/reprojectionOnGPU.cu:56-690): ray set-up (:128-150),
// box entry/exit (:154-192), major axis (:198), lumped first segment to the whole-cm plane int(xi)
// (:205-209), per-plane minor/z crossing detection by truncation (:212-247, z=0 never split), crossing
// order by distance from the plane (:249-404), midpoint voxel by truncation (:256-259) and the
// inscribed-circle mask (:261-267).
//
// Work shape: one wavefront = 64 consecutive detector rows of one column at one angle.  Such rays share
// their xy path (source and detector point differ only in z), so the plane sequence, the minor-axis
// crossings, the loop bounds and the major-axis choice are wave-uniform; only z crossings differ per
// lane.  The wave walks the planes in lockstep and each lane charges 1-3 segments per plane step.
// Steps where no lane of the wave can hit the volume's z band are skipped (their segments are not
// charged by the reference either: the midpoint slice is out of range).
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cmath>

#include "vcbct_internal.h"

namespace vcbct {

struct RayCtx {
    // literal ray set-up
    double sx, sy, sz, px, py, pz, leng;
    bool hit;  // at least two box hits
    // u = major axis coordinate, v = minor
    bool xmaj;
    double su, sv, pu, pv;
    double ui, vi, zi, uf, vf, zf;
    double pha;       // +-1, direction along u
    double uin;       // plane 0 position (whole cm)
    double dv, dz;    // dv/du, dz/du
    double duv, duz;  // du/dv, du/dz
    double K;         // segment length per |du|
    double dirv, dirz;
    bool vmove;       // pv != sv
    bool swap;        // entry point is the second box hit (the only lane-dependent choice of the xy path)
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// P/reprojectionOnGPU.cu:128-198 (FP64, the reference's operation order; no contraction in this file's
// users: they are compiled with -ffp-contract=off).
__host__ __device__ __forceinline__ void ray_setup(const ProjGeom& g, double c, double sn, double xc, double yc, int j, int jj,
                                          RayCtx& r) {
    r.sx = -g.sad * sn + g.x_reg;
    r.sy = g.sad * c + g.y_reg;
    r.sz = g.z_reg;
    r.pz = -(yc + j) * g.pitch_iso * (g.pixnum / g.H) + g.z_reg * g.pitch_iso / g.pitch;
    double pr = g.pitch_iso * (xc + jj) * (g.pixnum / g.H);
    r.px = pr * c + g.x_reg * g.pitch_iso / g.pitch;
    r.py = pr * sn + g.y_reg * g.pitch_iso / g.pitch;
    double leng = g.sad * g.sad / (g.sad * g.sad + (pr * pr + r.pz * r.pz));
    r.leng = leng * sqrt(leng);

    const int nxm = (int)(g.N * 0.5), nym = nxm;
    const double sx = r.sx, sy = r.sy, sz = r.sz, px = r.px, py = r.py, pz = r.pz;
    // candidate box hits in the reference's order: y = +nym s, y = -nym s, x = +nxm s, x = -nxm s; the
    // reference keeps them in order (at most four) and uses the first two
    const double lim = double(nym) * g.s, limx = double(nxm) * g.s;
    double cx[4], cy[4], cz[4];
    bool cv[4];
    {
        double y0 = lim;
        double x0 = (y0 - sy) / (py - sy) * (px - sx) + sx;
        cx[0] = x0; cy[0] = y0; cz[0] = (y0 - sy) / (py - sy) * (pz - sz) + sz; cv[0] = fabs(x0) < limx;
    }
    {
        double y0 = -lim;
        double x0 = (y0 - sy) / (py - sy) * (px - sx) + sx;
        cx[1] = x0; cy[1] = y0; cz[1] = (y0 - sy) / (py - sy) * (pz - sz) + sz; cv[1] = fabs(x0) < limx;
    }
    {
        double x0 = limx;
        double y0 = (x0 - sx) / (px - sx) * (py - sy) + sy;
        cx[2] = x0; cy[2] = y0; cz[2] = (x0 - sx) / (px - sx) * (pz - sz) + sz; cv[2] = fabs(y0) < lim;
    }
    {
        double x0 = -limx;
        double y0 = (x0 - sx) / (px - sx) * (py - sy) + sy;
        cx[3] = x0; cy[3] = y0; cz[3] = (x0 - sx) / (px - sx) * (pz - sz) + sz; cv[3] = fabs(y0) < lim;
    }
    const int m = (int)cv[0] + (int)cv[1] + (int)cv[2] + (int)cv[3];
    // first and second accepted hits, chosen with selects (no indexed arrays: they would live in scratch)
    const int f0 = cv[0] ? 0 : cv[1] ? 1 : cv[2] ? 2 : 3;
    const int f1 = (f0 < 1 && cv[1]) ? 1 : (f0 < 2 && cv[2]) ? 2 : 3;
    double xx[2], yy[2], zz[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int f = q ? f1 : f0;
        xx[q] = f == 0 ? cx[0] : f == 1 ? cx[1] : f == 2 ? cx[2] : cx[3];
        yy[q] = f == 0 ? cy[0] : f == 1 ? cy[1] : f == 2 ? cy[2] : cy[3];
        zz[q] = f == 0 ? cz[0] : f == 1 ? cz[1] : f == 2 ? cz[2] : cz[3];
    }
    r.hit = m >= 2;
    if (!r.hit) { xx[0] = xx[1] = yy[0] = yy[1] = zz[0] = zz[1] = 0.0; }
    double xi, yi, zi, xf, yf, zf;
    r.swap = !((xx[0] - sx) * (xx[0] - sx) + (yy[0] - sy) * (yy[0] - sy) + (zz[0] - sz) * (zz[0] - sz) <
               (xx[1] - sx) * (xx[1] - sx) + (yy[1] - sy) * (yy[1] - sy) + (zz[1] - sz) * (zz[1] - sz));
    if (!r.swap) {
        xi = xx[0]; yi = yy[0]; zi = zz[0]; xf = xx[1]; yf = yy[1]; zf = zz[1];
    } else {
        xi = xx[1]; yi = yy[1]; zi = zz[1]; xf = xx[0]; yf = yy[0]; zf = zz[0];
    }
    r.xmaj = fabs(xi - xf) > fabs(yi - yf);
    r.su = r.xmaj ? sx : sy;
    r.sv = r.xmaj ? sy : sx;
    r.pu = r.xmaj ? px : py;
    r.pv = r.xmaj ? py : px;
    r.ui = r.xmaj ? xi : yi;
    r.uf = r.xmaj ? xf : yf;
    r.vi = r.xmaj ? yi : xi;
    r.vf = r.xmaj ? yf : xf;
    r.zi = zi;
    r.zf = zf;
    r.pha = (r.uf - r.ui) / fabs(r.ui - r.uf);
    int uin;
    if ((r.ui < 0 && (r.uf - r.ui) < 0) || (r.ui > 0 && (r.uf - r.ui) > 0)) uin = int(r.ui) + (r.uf - r.ui) / fabs(r.uf - r.ui);
    else uin = int(r.ui);
    r.uin = (double)uin;
    r.dv = (r.pv - r.sv) / (r.pu - r.su);
    r.dz = (pz - sz) / (r.pu - r.su);
    r.duv = (r.pu - r.su) / (r.pv - r.sv);
    r.duz = (r.pu - r.su) / (pz - sz);
    r.K = sqrt(1.0 + r.dv * r.dv + r.dz * r.dz);
    r.dirv = (r.vf - r.vi) / fabs(r.vf - r.vi);
    r.dirz = (zf - zi) / fabs(zf - zi);
    r.vmove = (r.pv != r.sv) && (r.vf != r.vi);
    if (!r.hit) r.pha = 1.0;
}

// Range of plane steps k (k >= 1) whose segments can have a midpoint slice inside [0, Nz-1].  Slice i holds
// the midpoints with (int)(-z / t + nzm) == i, i.e. z in ((nzm - i - 1) t, (nzm - i) t], and slice 0 also
// the truncated (nzm t, (nzm + 1) t].
__host__ __device__ __forceinline__ void z_step_range(const ProjGeom& g, const RayCtx& r, bool active, int& klo, int& khi) {
    klo = INT_MAX;
    khi = INT_MIN;
    if (!active || !r.hit) return;
    const int nzm = (int)(g.Nz * 0.5);
    const double zlo = (double)(nzm - g.Nz) * g.t, zhi = (double)(nzm + 1) * g.t;
    const double z0 = (r.uin - r.su) * r.dz + r.sz;   // z at plane 0
    const double dk = r.pha * g.s * r.dz;             // z change per plane step
    if (fabs(dk) < 1e-300) {
        if (z0 > zlo - g.t && z0 < zhi + g.t) { klo = 0; khi = INT_MAX - 8; }
        return;
    }
    double k1 = (zlo - z0) / dk, k2 = (zhi - z0) / dk;
    double a = fmin(k1, k2), b = fmax(k1, k2);
    // steps whose end planes both lie beyond the band on one side have midpoints outside it;
    // k = 0 stands for the lumped entry segment, which spans plane indices [ke, 0]
    a = floor(a) - 1.0;
    b = ceil(b) + 2.0;
    const double ke = (r.ui - r.uin) * r.pha / g.s - 1.0;
    if (b < ke) return;
    if (a > 1.0e8) return;
    klo = (int)fmax(a, 0.0);
    khi = (int)fmin(fmax(b, 0.0), 1.0e8);
}

// ((a + b) * 0.5) / s + n truncated, unfused, in the reference's operation order.
__host__ __device__ __forceinline__ int exact_index(double a, double b, double s, int n) {
#pragma clang fp contract(off)
    double m = (a + b) * 0.5;
    double q = m / s;
    return (int)(q + n);
}

// Plane k of the major axis, xin + pha * ik * s in the reference's order (P/reprojectionOnGPU.cu:214), unfused:
// its value feeds the tie-deciding index above.
__host__ __device__ __forceinline__ double plane_u(const RayCtx& r, int k, double s) {
#pragma clang fp contract(off)
    double t = r.pha * (double)k;
    t = t * s;
    return r.uin + t;
}

// Per-ray constants of segment charging.  Voxel indices follow the reference's midpoint truncation,
// (int)(coord / pitch + n): along the major axis u -> (int)(u * smaj + nmaj), along the minor axis
// v -> (int)(v * smin + nmaj), in z -> (int)(z * ninv_t + nzm) (x index = x / s + nxm, y index = -y / s + nym,
// z index = -z / t + nzm; nxm == nym).  A voxel's cell in the [y][x][z] layout is
// (y * N + x) * Nz + z = imin * mstride + imaj * majmul + z.
struct ChargeC {
    double smaj, smin, nmaj, ninv_t, nzm, rmax2;
    int mstride, majmul;
};

// imin * mstride with a full-rate 24-bit multiply when both factors fit in 24 bits.
template <bool MUL24>
__host__ __device__ __forceinline__ int minor_term(int imin, int mstride) {
#ifdef __HIP_DEVICE_COMPILE__
    if (MUL24) return (int)__umul24((unsigned)imin, (unsigned)mstride);
#endif
    return (int)((unsigned)imin * (unsigned)mstride);  // wraps harmlessly for uncharged segments
}

__host__ __device__ __forceinline__ ChargeC charge_consts(const ProjGeom& g, const RayCtx& r) {
    ChargeC c;
    c.smaj = r.xmaj ? g.inv_s : -g.inv_s;
    c.smin = r.xmaj ? -g.inv_s : g.inv_s;
    c.nmaj = g.nmaj;
    c.ninv_t = g.ninv_t;
    c.nzm = g.nzm;
    c.rmax2 = g.rmax2;
    c.mstride = r.xmaj ? g.N * g.Nz : g.Nz;
    c.majmul = r.xmaj ? g.Nz : g.N * g.Nz;
    return c;
}

// Major-axis voxel index of the segment [ua, ub] from its midpoint.  The index can sit exactly on a voxel
// boundary for whole families of segments: the lumped entry segment whenever the ray enters through a
// major-axis face ((12.8 + 12) / 2 = 248 s), and every plane-to-plane segment when the plane grid is
// offset from the voxel grid (s = 0.8 cm: planes at whole cm, boundaries at multiples of 0.8).  Near such a
// tie the index is recomputed with the reference's own expression, ((a + b) * 0.5) / s + n, which decides
// it (P/reprojectionOnGPU.cu:256-257); the fast value is within a few ulps of it otherwise.
__host__ __device__ __forceinline__ int major_index(const ProjGeom& g, const RayCtx& r, const ChargeC& c, double ua,
                                                    double ub) {
#pragma clang fp contract(fast)
    const double um = (ua + ub) * 0.5;
    const double fu = um * c.smaj + c.nmaj;
    int iu = (int)fu;
    if (fabs(fu - rint(fu)) < 1e-9) {
        const int n = (int)c.nmaj;
        iu = r.xmaj ? exact_index(ua, ub, g.s, n) : exact_index(-ua, -ub, g.s, n);
    }
    return iu;
}

// One charged segment: voxel cell ((y * N + x) * Nz + z), length; ok = charged by this lane.
struct Seg {
    int cell;
    double rr;
    bool ok;
};

// The segment [ua, ub] (P/reprojectionOnGPU.cu:256-269): midpoint voxel by truncation, in-range test,
// inscribed-circle mask at the midpoint, length |ub - ua| * sqrt(1 + dv^2 + dz^2).  `majterm`/`majok` carry
// the major-axis part of the cell when the caller knows it for the whole plane step (STEPMAJ); otherwise it
// is taken from this segment's midpoint.
template <bool STEPMAJ, bool MUL24>
__host__ __device__ __forceinline__ Seg make_seg(const ProjGeom& g, const RayCtx& r, const ChargeC& c, double ua,
                                                 double ub, bool on, int majterm, bool majok) {
    // The per-segment arithmetic may fuse multiply-adds: it already differs from the reference's operation
    // order (slopes instead of divisions), and the set-up, whose roundings are restated, stays unfused.
#pragma clang fp contract(fast)
    Seg sg;
    sg.ok = false;
    if (!on) return sg;  // the other fields are read only where ok is set
    const double um = (ua + ub) * 0.5;
    const double d = um - r.su;
    const double vm = d * r.dv + r.sv;
    const double zm = d * r.dz + r.sz;
    const int imin = (int)(vm * c.smin + c.nmaj);
    const int zmc = (int)(zm * c.ninv_t + c.nzm);
    if (!STEPMAJ) {
        const int imaj = major_index(g, r, c, ua, ub);
        majok = (unsigned)imaj < (unsigned)g.N;
        majterm = (int)((unsigned)imaj * (unsigned)c.majmul);
    }
    sg.ok = majok && (unsigned)imin < (unsigned)g.N && (unsigned)zmc < (unsigned)g.Nz && um * um + vm * vm < c.rmax2;
    sg.cell = (int)((unsigned)minor_term<MUL24>(imin, c.mstride) + (unsigned)majterm + (unsigned)zmc);
    sg.rr = fabs(ub - ua) * r.K;
    return sg;
}

// Hands the (up to three) segments of one plane step to the visitor.  Visitors implement
// step(s0, s1, s2) and see each segment with ok set exactly once, in ray order.
// One lane's walk over plane steps [kmin, kmax] (the wave-wide union of the lanes' z ranges).
// ALIGNED: the whole-cm plane grid lies on voxel boundaries (1/s is an integer), so every segment of a plane
// step has the major-axis index of the step's midpoint, computed once per step.
template <bool ALIGNED, bool MUL24, class V>
__host__ __device__ __forceinline__ void traverse_steps(const ProjGeom& g, const RayCtx& r, bool active, int kmin,
                                                        int kmax, V& vis) {
#pragma clang fp contract(fast)
    const double lim_u = g.lim_u;  // nxm == nym
    const double inv_s = g.inv_s, inv_t = g.inv_t;
    const bool on = active && r.hit;
    const ChargeC c = charge_consts(g, r);

    // plane 0 and the lumped entry segment (wave-uniform box test)
    double u0 = r.uin;
    double v0 = (u0 - r.su) * r.dv + r.sv;
    bool inbox0 = (u0 <= lim_u && u0 >= -lim_u && v0 <= lim_u && v0 >= -lim_u);
    if (!inbox0) return;
    int kstart = kmin;
    double uprev;
    if (kstart <= 1) {  // kmin 0 or 1: the lumped entry segment is needed (and cheap)
        const Seg none{0, 0.0, false};
        vis.step(make_seg<false, MUL24>(g, r, c, r.ui, u0, on, 0, false), none, none);
        kstart = 1;
        uprev = u0;
    } else {
        uprev = plane_u(r, kstart - 1, g.s);
    }
    double vprev = (uprev - r.su) * r.dv + r.sv;
    double zprev = (uprev - r.su) * r.dz + r.sz;
    int vmaprev = (int)(vprev * inv_s);
    int zmaprev = (int)(zprev * inv_t);
    const double ds = r.dirv * g.s, dt = r.dirz * g.t;

    for (int k = kstart; k <= kmax; ++k) {
        // the reference continues while the previous plane lies in the box (inclusive)
        if (!(uprev <= lim_u && uprev >= -lim_u && vprev <= lim_u && vprev >= -lim_u)) break;
        const double uk = plane_u(r, k, g.s);
        const double duk = uk - r.su;
        const double vk = duk * r.dv + r.sv;
        const double zk = duk * r.dz + r.sz;
        const int vma = (int)(vk * inv_s);
        const int zma = (int)(zk * inv_t);
        // sign change of v counts as a crossing (truncation maps (-s, s) to one index)
        const bool hv = r.vmove && ((vma != vmaprev) || !(vk * vprev > 0.0));
        const bool hz = zma != zmaprev;
        double uv = uk, uz = uk;
        if (hv) {
            const double vb = (vma * g.s) - ((vk <= 0 && r.dirv < 0) || (vk >= 0 && r.dirv > 0) ? 0.0 : ds);
            uv = (vb - r.sv) * r.duv + r.su;
        }
        if (hz) {
            const double zb = (zma * g.t) - ((zk < 0 && r.dirz < 0) || (zk > 0 && r.dirz > 0) ? 0.0 : dt);
            uz = (zb - r.sz) * r.duz + r.su;
        }
        // crossings ordered by distance from the end plane (the farther one comes first)
        const double chv = fabs(uk - uv), chz = fabs(uk - uz);
        double e1 = uk, e2 = uk;
        if (hv && hz) {
            if (chz > chv) { e1 = uz; e2 = uv; }
            else if (chz < chv) { e1 = uv; e2 = uz; }
        } else if (hz && chz != 0) {
            e1 = uz;
        } else if (hv && chv != 0) {
            e1 = uv;
        }
        const double b0 = uprev;
        if (ALIGNED) {
            const int imaj = (int)((b0 + uk) * 0.5 * c.smaj + c.nmaj);
            if ((unsigned)imaj < (unsigned)g.N) {
                const int mt = imaj * c.majmul;
                vis.step(make_seg<true, MUL24>(g, r, c, b0, e1, on, mt, true),
                         make_seg<true, MUL24>(g, r, c, e1, e2, on && (e2 != e1), mt, true),
                         make_seg<true, MUL24>(g, r, c, e2, uk, on && (uk != e2), mt, true));
            }
        } else {
            vis.step(make_seg<false, MUL24>(g, r, c, b0, e1, on, 0, false),
                     make_seg<false, MUL24>(g, r, c, e1, e2, on && (e2 != e1), 0, false),
                     make_seg<false, MUL24>(g, r, c, e2, uk, on && (uk != e2), 0, false));
        }
        uprev = uk;
        vprev = vk;
        vmaprev = vma;
        zmaprev = zma;
    }
}

// The wave-uniform part of plane step k (the xy path is shared by the lanes of a wave), computed for 64 steps
// at a time, one step per lane, and read back by every lane in step order.  A step's segments are the one
// or two pieces of its xy chord split at the minor-axis crossing (the same on every lane), one of which a
// lane's own z crossing may split again.  Everything about a piece that depends only on its xy chord is
// decided here, once per step: the plane positions, the minor crossing, the piece's voxel column (its minor
// index from the piece's midpoint, and the step's major index), whether that column is in range, and how the
// inscribed-circle mask treats the piece.  The piece of a z split lies inside its xy piece, so it has the
// same column (a piece's minor cell is that of any of its points; only a piece shorter than ~1e-12 cm could
// round its midpoint across a minor boundary, and such a piece charges nothing measurable), and the same
// circle decision unless the circle's boundary crosses the xy piece: those rare steps carry STEP_CIRC and
// every lane tests its own midpoints there.  Positions are kept relative to the source along the major axis
// (d = u - su), the form in which a midpoint's z is formed: zm = d_mid * dz + sz.
struct StepRec {
    double dk;     // plane k, uk - su
    double dv;     // minor crossing, uv - su (== dk when there is none)
    int cb0, cb1;  // voxel-column base (minor term + major term) of the first and second xy piece
    int flags;     // STEP_* below
    int majterm;   // imaj * majmul (for a segment that spans the minor crossing)
};
enum {
    STEP_SKIP = 1,    // no segment of the step can be charged (column out of range, or wholly outside the circle)
    STEP_SPLITV = 2,  // the minor crossing splits the chord into two xy pieces
    STEP_OK0 = 4,     // first xy piece chargeable (column in range, not wholly outside the circle)
    STEP_OK1 = 8,     // second xy piece chargeable
    STEP_CIRC = 16,   // the circle's boundary crosses the chord: test each midpoint
};

// A lane's held z-split piece (see traverse_table): [a, b] relative to the source, its column base and its
// xy piece's STEP_OK / STEP_CIRC bits.
struct HeldPiece {
    double a, b;
    int cb, flags;
};

// Per-wave scratch of the table walk (LDS on the GPU): the step table and one held piece per lane.
struct WaveLds {
    StepRec steps[64];
    HeldPiece held[64];
};

// Squared distance bounds of the xy chord from (u0, v0) to (u1, v1) to the volume's axis: the largest is at an
// end (r^2 is convex along a line), the smallest is at least that of the chord's bounding box.
__host__ __device__ __forceinline__ void chord_r2(double u0, double v0, double u1, double v1, double& lo, double& hi) {
    const double au = (u0 > 0.0) == (u1 > 0.0) ? fmin(fabs(u0), fabs(u1)) : 0.0;
    const double av = (v0 > 0.0) == (v1 > 0.0) ? fmin(fabs(v0), fabs(v1)) : 0.0;
    lo = au * au + av * av;
    hi = fmax(u0 * u0 + v0 * v0, u1 * u1 + v1 * v1);
}

__host__ __device__ __forceinline__ bool make_step(const ProjGeom& g, const RayCtx& r, const ChargeC& c, int k,
                                                   StepRec& t) {
#pragma clang fp contract(fast)
    const double up = plane_u(r, k - 1, g.s);
    const double uk = plane_u(r, k, g.s);
    const double vp = (up - r.su) * r.dv + r.sv;
    const double duk = uk - r.su;
    const double vk = duk * r.dv + r.sv;
    const bool inprev = up <= g.lim_u && up >= -g.lim_u && vp <= g.lim_u && vp >= -g.lim_u;
    const int vma = (int)(vk * g.inv_s);
    const int vmap = (int)(vp * g.inv_s);
    const bool hv = r.vmove && ((vma != vmap) || !(vk * vp > 0.0));
    double uv = uk;
    if (hv) {
        const double vb = (vma * g.s) - ((vk <= 0 && r.dirv < 0) || (vk >= 0 && r.dirv > 0) ? 0.0 : r.dirv * g.s);
        uv = (vb - r.sv) * r.duv + r.su;
    }
    // a crossing at the end plane splits nothing (the reference's crossing distance is 0 there)
    const bool splitv = hv && fabs(uk - uv) != 0.0;
    const int imaj = (int)((up + uk) * 0.5 * c.smaj + c.nmaj);
    const bool majok = (unsigned)imaj < (unsigned)g.N;
    const int majterm = imaj * c.majmul;
    // xy pieces [up, a1] and [uv, uk]; their columns from their midpoints, as a segment's (make_seg)
    const double a1 = splitv ? uv : uk;
    const double vv = (uv - r.su) * r.dv + r.sv;  // v at the minor crossing
    const double va1 = splitv ? vv : vk;
    const double m0 = (up + a1) * 0.5 - r.su;
    const double m1 = (uv + uk) * 0.5 - r.su;
    const int imin0 = (int)((m0 * r.dv + r.sv) * c.smin + c.nmaj);
    const int imin1 = (int)((m1 * r.dv + r.sv) * c.smin + c.nmaj);
    // circle mask over each piece: wholly inside (margin 1e-9 cm^2, while a midpoint's rounding is ~1e-13 cm^2),
    // wholly outside, or crossed (then tested per midpoint)
    double lo0, hi0, lo1, hi1;
    chord_r2(up, vp, a1, va1, lo0, hi0);
    chord_r2(uv, vv, uk, vk, lo1, hi1);
    const bool out0 = lo0 > c.rmax2 + 1e-9, out1 = !splitv || lo1 > c.rmax2 + 1e-9;
    const bool circ = (!out0 && hi0 >= c.rmax2 - 1e-9) || (!out1 && hi1 >= c.rmax2 - 1e-9);
    const bool ok0 = majok && (unsigned)imin0 < (unsigned)g.N && !out0;
    const bool ok1 = splitv && majok && (unsigned)imin1 < (unsigned)g.N && !out1;
    t.dk = duk;
    t.dv = uv - r.su;
    t.cb0 = (int)((unsigned)imin0 * (unsigned)c.mstride + (unsigned)majterm);
    t.cb1 = (int)((unsigned)imin1 * (unsigned)c.mstride + (unsigned)majterm);
    t.majterm = majterm;
    // a step none of whose pieces can be charged only moves the z state on; a tie of the minor and z crossing
    // distances makes the whole chord one segment, so such a step is kept whenever its column is in range
    t.flags = (!majok || (out0 && out1) ? STEP_SKIP : 0) | (splitv ? STEP_SPLITV : 0) | (ok0 ? STEP_OK0 : 0) |
              (ok1 ? STEP_OK1 : 0) | (circ ? STEP_CIRC : 0);
    return inprev;
}

// The first lane's value of x (in scalar registers).
__device__ __forceinline__ double first_lane(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// True when the condition holds on any lane of the wave (the host build walks one lane at a time).
__host__ __device__ __forceinline__ bool any_lane(bool b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __any(b);
#else
    return b;
#endif
}

// Per-lane constants of the table walk: the z slope and its inverse, the segment length per unit of d, and
// the z offset of the midpoint rule.
struct LaneZ {
    // the slice index of the piece [a, b] is (int)(zm * ninv_t + nzm) with zm = (a + b) dz / 2 + sz; folded
    // into one FMA of (a + b): zA = dz / 2 * ninv_t, zB = sz * ninv_t + nzm (per-lane constants)
    double zA, zB;
    double K;     // segment length per |du|
};

// One piece [a, b] (relative to the source) of column base cb: its slice from the midpoint and its length.
__host__ __device__ __forceinline__ Seg make_piece(const ProjGeom& g, const RayCtx& r, const ChargeC& c,
                                                   const LaneZ& lz, double a, double b, int cb, bool on) {
#pragma clang fp contract(fast)
    Seg sg;
    const int zmc = (int)((a + b) * lz.zA + lz.zB);
    sg.ok = on && (unsigned)zmc < (unsigned)g.Nz;
    sg.cell = (int)((unsigned)cb + (unsigned)zmc);
    sg.rr = fabs(b - a) * lz.K;
    return sg;
}

// The column base of a segment that spans the minor crossing (a tie of the crossing distances), from its own
// midpoint; -1 when the column is out of range.
__host__ __device__ __forceinline__ int span_base(const ProjGeom& g, const RayCtx& r, const ChargeC& c, double a,
                                                  double b, int majterm) {
#pragma clang fp contract(fast)
    const double dm = (a + b) * 0.5;
    const int imin = (int)((dm * r.dv + r.sv) * c.smin + c.nmaj);
    return (unsigned)imin < (unsigned)g.N ? (int)((unsigned)imin * (unsigned)c.mstride + (unsigned)majterm) : -1;
}

// The inscribed-circle mask at the midpoint of [a, b] (relative to the source), as make_seg tests it.
__host__ __device__ __forceinline__ bool in_circle(const RayCtx& r, const ChargeC& c, double a, double b) {
#pragma clang fp contract(fast)
    const double dm = (a + b) * 0.5;
    const double um = dm + r.su;
    const double vm = dm * r.dv + r.sv;
    return um * um + vm * vm < c.rmax2;
}

// The aligned-grid walk with the step records of 64 steps at a time in a per-wave table (`lds`; LDS on the
// GPU).  Only the z part stays per lane: the z crossing of each step, which piece it splits, and each
// piece's slice and length.  Requires the lanes to share their xy path bit for bit (caller checks).
// The z crossings are sparse per lane (a few per cent of the lanes in a step) but, over 64 rows, present in
// almost every step: so the second piece of a z split is not handed over in the step (which would run a
// mostly idle extra slot for the whole wave) but held by the lane (in LDS), and the wave hands over all held
// pieces together when a lane that holds one meets its next z crossing, and at the end.  The segments are
// the same; only the order of their sums changes.  The step body has one hand-over to the visitor, with both
// xy pieces (the second charges nothing on steps without a minor crossing), and no branch that depends on a
// single lane's data: the rare cases (a piece near the circle's boundary, a tie of the crossing distances)
// are wave-uniform branches around their extra arithmetic.
template <class V>
__host__ __device__ __forceinline__ void traverse_table(const ProjGeom& g, const RayCtx& r_in, bool active, int kmin,
                                                        int kmax, V& vis, WaveLds* lds) {
#pragma clang fp contract(fast)
#ifdef __HIP_DEVICE_COMPILE__
    // Every lane took the same box entry here (the caller's condition), so the entry and exit along both axes,
    // the plane grid (uin, pha) and the minor direction are the same on every lane, bit for bit: read from
    // the first lane they sit in scalar registers, not in vector registers live across the walk.
    RayCtx r = r_in;
    r.ui = first_lane(r_in.ui);
    r.uf = first_lane(r_in.uf);
    r.vi = first_lane(r_in.vi);
    r.vf = first_lane(r_in.vf);
    r.uin = first_lane(r_in.uin);
    r.pha = first_lane(r_in.pha);
    r.dirv = first_lane(r_in.dirv);
#else
    const RayCtx& r = r_in;
#endif
    const double lim_u = g.lim_u;
    const bool on = active && r.hit;
    const ChargeC c = charge_consts(g, r);
    LaneZ lz;
    lz.zA = 0.5 * r.dz * c.ninv_t;
    lz.zB = r.sz * c.ninv_t + c.nzm;
    lz.K = r.K;

    double u0 = r.uin;
    double v0 = (u0 - r.su) * r.dv + r.sv;
    bool inbox0 = (u0 <= lim_u && u0 >= -lim_u && v0 <= lim_u && v0 >= -lim_u);
    if (!inbox0) return;
    int k = kmin;
    double b0;  // the previous plane, relative to the source
    if (k <= 1) {
        const Seg none{0, 0.0, false};
        vis.step(make_seg<false, false>(g, r, c, r.ui, u0, on, 0, false), none, none);
        k = 1;
        b0 = u0 - r.su;
    } else {
        b0 = plane_u(r, k - 1, g.s) - r.su;
    }
    const double inv_t = g.inv_t;
    int zmaprev = (int)((b0 * r.dz + r.sz) * inv_t);
    const double dt = r.dirz * g.t;
    const Seg none{0, 0.0, false};
    StepRec* tab = lds->steps;
#ifdef __HIP_DEVICE_COMPILE__
    HeldPiece* hp = &lds->held[threadIdx.x & 63];  // this lane's held piece (kept in LDS, not registers)
#else
    HeldPiece* hp = &lds->held[0];
#endif
    bool held = false;

    // hands the held pieces over (a wave-wide slot; lanes without one charge nothing); the midpoint test of
    // the circle is exact for every piece, so it is applied to all of them
    auto flush = [&]() {
        const HeldPiece q = *hp;
        const bool qon = held && (q.flags & STEP_OK0) && ((q.flags & STEP_CIRC) == 0 || in_circle(r, c, q.a, q.b));
        vis.step(make_piece(g, r, c, lz, q.a, q.b, q.cb, qon), none, none);
        held = false;
    };

    for (;;) {
        // steps k .. k+63: record and validity (k' <= kmax and the reference's loop still running)
        int nsteps;
#ifdef __HIP_DEVICE_COMPILE__
        {
            const int lane = threadIdx.x & 63;
            StepRec t;
            const bool ok = make_step(g, r, c, k + lane, t) && k + lane <= kmax;
            tab[lane] = t;
            const unsigned long long bad = __ballot(!ok);
            nsteps = bad ? __ffsll((long long)bad) - 1 : 64;
            __builtin_amdgcn_wave_barrier();
        }
#else
        nsteps = 64;
        for (int l = 0; l < 64; ++l) {
            const bool ok = make_step(g, r, c, k + l, tab[l]) && k + l <= kmax;
            if (!ok && nsteps == 64) nsteps = l;
        }
#endif
        for (int i = 0; i < nsteps; ++i) {
            const double dk = tab[i].dk;
#ifdef __HIP_DEVICE_COMPILE__
            // the record is the same for every lane: its integer fields in scalar registers make the
            // branches on them scalar
            const int flags = __builtin_amdgcn_readfirstlane(tab[i].flags);
#else
            const int flags = tab[i].flags;
#endif
            const double zk = dk * r.dz + r.sz;  // (one FMA)
            const int zma = (int)(zk * inv_t);
            if (flags & STEP_SKIP) {  // nothing charged in this step: only the z state moves on
                b0 = dk;
                zmaprev = zma;
                continue;
            }
#ifdef __HIP_DEVICE_COMPILE__
            int cb0 = __builtin_amdgcn_readfirstlane(tab[i].cb0);
            const int cb1 = __builtin_amdgcn_readfirstlane(tab[i].cb1);
#else
            int cb0 = tab[i].cb0;
            const int cb1 = tab[i].cb1;
#endif
            const double dv = tab[i].dv;
            const bool hz = zma != zmaprev;
            if (any_lane(hz && held)) flush();  // a lane that holds a piece meets a z crossing: hand over first
            // the z crossing, relative to the source (formed on every lane, used where hz)
            const double zb = (zma * g.t) - ((zk < 0 && r.dirz < 0) || (zk > 0 && r.dirz > 0) ? 0.0 : dt);
            const double dzc = hz ? (zb - r.sz) * r.duz : dk;
            // Crossings ordered by distance from the end plane (the farther one comes first; equal distances
            // drop both, as no branch of the reference matches); a crossing at the end plane splits nothing.
            const bool splitv = (flags & STEP_SPLITV) != 0;  // wave-uniform
            const double chz = fabs(dk - dzc), chv = fabs(dk - dv);
            const bool splitz = hz && chz != 0.0;
            const bool tie = splitv && splitz && chz == chv;
            const bool z0 = splitz && !tie && (!splitv || chz > chv);  // z crossing inside the first piece
            const bool z1 = splitz && !tie && !z0;                     // inside the second
            // first piece [b0, a1]: a1 = dv, as the table's dv equals dk when the minor crossing splits nothing
            const double a1 = dv;
            if (on && (z0 || z1)) {  // the piece beyond the z crossing is held
                *hp = HeldPiece{dzc, z0 ? a1 : dk, z0 ? cb0 : cb1, z0 ? (flags & (STEP_OK0 | STEP_CIRC))
                                                                      : ((flags & STEP_OK1) ? STEP_OK0 : 0) |
                                                                            (flags & STEP_CIRC)};
                held = true;
            }
            double e0 = z0 ? dzc : a1;
            const double e1 = z1 ? dzc : dk;
            bool on0 = on && (flags & STEP_OK0);
            const bool on1 = on && !tie && (flags & STEP_OK1);
            bool circ = (flags & STEP_CIRC) != 0;
            if (any_lane(on && tie)) {  // lanes with a tie charge the whole chord as one segment, from its midpoint
                const int cbt = span_base(g, r, c, b0, dk, tab[i].majterm);
                if (tie) {
                    e0 = dk;
                    cb0 = cbt;
                    on0 = on && cbt >= 0;
                }
                circ = true;
            }
            bool on1c = on1;
            if (circ) {  // a piece near the circle's boundary: the midpoint tests
                on0 = on0 && in_circle(r, c, b0, e0);
                if (splitv) on1c = on1c && in_circle(r, c, dv, e1);
            }
            // the second xy piece exists only where the minor crossing splits the step (wave-uniform): its slice,
            // cell and length are formed only then
            const Seg q0 = make_piece(g, r, c, lz, b0, e0, cb0, on0);
            Seg q1 = none;
            if (splitv) q1 = make_piece(g, r, c, lz, dv, e1, cb1, on1c);
            vis.step(q0, q1, none);
            b0 = dk;
            zmaprev = zma;
        }
#ifdef __HIP_DEVICE_COMPILE__
        __builtin_amdgcn_wave_barrier();  // all reads of this table before the next fill
#endif
        if (nsteps < 64) break;
        k += 64;
    }
    if (any_lane(held)) flush();
}

// Walks the compat traversal of one ray (lanes of a wave in lockstep over plane steps).
// `active`: lane holds a real detector pixel.  The visitor accumulates.
template <class V>
__host__ __device__ __forceinline__ void traverse_lane(const ProjGeom& g, const RayCtx& r, bool active, int kmin,
                                                       int kmax, V& vis, WaveLds* lds = nullptr) {
    if (kmin > kmax) return;  // no lane of this wave can reach the volume's z band
    if (g.aligned && lds) {  // the caller vouches that the lanes share their xy path
        traverse_table(g, r, active, kmin, kmax, vis, lds);
    } else if (g.aligned) {
        if (g.mul24) traverse_steps<true, true>(g, r, active, kmin, kmax, vis);
        else traverse_steps<true, false>(g, r, active, kmin, kmax, vis);
    } else {
        if (g.mul24) traverse_steps<false, true>(g, r, active, kmin, kmax, vis);
        else traverse_steps<false, false>(g, r, active, kmin, kmax, vis);
    }
}

}  // namespace vcbct

namespace vcbct {
// Walks the rays of one wave: lane = detector row j (inactive lanes, j >= H, walk row 0 and charge
// nothing), all lanes one detector column jj at the angle whose trig row is (c, sn, xc, yc).  The plane
// steps walked are the union of the lanes' z ranges (wave-wide min/max).  Returns the lane's cos^3 factor.

// The fields of a ray's xy path that do not depend on which box hit is the entry are the same, bit for
// bit, on every lane of the wave (the lanes differ only in z): the major axis (|xi - xf| > |yi - yf| is
// symmetric in the two hits), the source position in (u, v), the slopes dv/du and du/dv, and the
// hit/minor-move flags.  Reading them from the first lane puts them in scalar registers.
__device__ __forceinline__ void share_xy_path(RayCtx& r) {
    r.xmaj = __builtin_amdgcn_readfirstlane((int)r.xmaj) != 0;
    r.hit = __builtin_amdgcn_readfirstlane((int)r.hit) != 0;
    r.vmove = __builtin_amdgcn_readfirstlane((int)r.vmove) != 0;
    r.su = first_lane(r.su);
    r.sv = first_lane(r.sv);
    r.dv = first_lane(r.dv);
    r.duv = first_lane(r.duv);
}

// `lds`: this wave's step table and held pieces in LDS.  Waves whose lanes all chose the same box entry
// (every wave but those of rays through a box corner) share their whole xy path and walk with the step
// table; the others walk per lane.
template <class V>
__device__ __forceinline__ double traverse_wave(const ProjGeom& g, double c, double sn, double xc, double yc, int j,
                                                int jj, bool active, V& vis, WaveLds* lds) {
    RayCtx r;
    ray_setup(g, c, sn, xc, yc, active ? j : 0, jj, r);
    share_xy_path(r);
    int klo, khi;
    z_step_range(g, r, active, klo, khi);
    const int kmin = wave_min(klo);
    const int kmax = wave_max(khi);
    const bool same = __all((int)r.swap == __builtin_amdgcn_readfirstlane((int)r.swap));
    traverse_lane(g, r, active, kmin, kmax, vis, same ? lds : nullptr);
    return r.leng;
}
}  // namespace vcbct
