// Internal declarations shared by the vcbct translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>

#include <cmath>

#include "vcbct.h"

namespace vcbct {

// Thread-local last-error message; returns `code` so call sites can `return fail(...)`.
int fail(int code, const char* fmt, ...);

#define VCBCT_HIP(call)                                                                          \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return ::vcbct::fail(VCBCT_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,        \
                                 hipGetErrorString(_e));                                         \
    } while (0)

// Device copy of the projector geometry: everything the traversal needs, by value in a kernel argument.
struct ProjGeom {
    int N, Nz, H, W, nAng, nmat;
    double s, t;            // voxel pitch x/y and z [cm]
    double x_reg, y_reg, z_reg;
    double pitch_iso, pitch, sad, pixnum;
    // derived on the host by derive_geom() (kernel arguments, so they sit in scalar registers)
    double inv_s, inv_t;    // 1/s, 1/t
    double ninv_t;          // -1/t
    double nmaj, nzm;       // (int)(N/2), (int)(Nz/2) as doubles: index offsets of the midpoint rule
    double lim_u;           // box half-width (int)(N/2) * s
    double rmax2;           // squared radius of the inscribed-circle mask, (N/2 * s)^2
    int aligned;            // 1/s integral: the whole-cm plane grid lies on voxel boundaries
    int mul24;              // N * Nz < 2^24: voxel strides multiply in 24 bits
};

// Fills the derived fields of a ProjGeom whose primary fields are set.
inline void derive_geom(ProjGeom& g) {
    g.inv_s = 1.0 / g.s;
    g.inv_t = 1.0 / g.t;
    g.ninv_t = -1.0 / g.t;
    g.nmaj = (double)(int)(g.N * 0.5);
    g.nzm = (double)(int)(g.Nz * 0.5);
    g.lim_u = double((int)(g.N * 0.5)) * g.s;
    const double rmax = g.N * 0.5 * g.s;
    g.rmax2 = rmax * rmax;
    const double q = 1.0 / g.s;
    g.aligned = (q >= 1.0 && std::fabs(q - std::rint(q)) < 1e-9) ? 1 : 0;
    g.mul24 = ((long long)g.N * g.Nz < (1LL << 24)) ? 1 : 0;
}

#ifdef __HIPCC__
// Loads and stores through pointers known to address global memory.  Pointers read from device-resident
// descriptors (the batched launch) are otherwise generic, and generic accesses compile to FLAT
// instructions; through a global pointer the compiler emits global (or, for uniform read-only data,
// scalar) memory instructions.
template <class T>
__device__ __forceinline__ T ldg(const T* p, size_t i) {
    typedef __attribute__((address_space(1))) const T GT;
    T r;
    __builtin_memcpy(&r, (GT*)p + i, sizeof(T));  // (HIP vector types have no copy from a qualified object)
    return r;
}
template <class T>
__device__ __forceinline__ void stg(T* p, size_t i, T v) {
    typedef __attribute__((address_space(1))) T GT;
    ((GT*)p)[i] = v;
}
#endif

// Per-call spectrum view (device pointers to FP64 copies of the reference's float tables).
struct Spectrum {
    const double* mu;   // [nmat][NE]
    const double* pdf;  // [NE]
    const float* muf;   // the same tables as given (FP32): the intensity projection's spectral sum
    const float* pdff;
    int NE;
    // The adjoint's channels: the gradient with respect to the nmat maps is formed as sum_k B_mk G_k from nch
    // scattered channels G_k, whose per-ray coefficients use the channel table ch [nch][NE] in place of mu.
    // ch = mu, nch = nmat, B = identity when the table has no lower-rank form (see spectral_basis).
    const double* ch;   // [nch][NE]
    int nch;
    // The forward's spectral sum over the line integrals of the traversed volume: nl = nmat with the tables
    // mu / muf above for material volumes; for channel volumes (VOX_CH2 / VOX_CH4) the caller passes the channel
    // tables as mu (FP64) and muf (FP32) and nl = nch.  chf: the channel table in FP32 (as muf for mu).
    int nl;
    const float* chf;
};
// Voxel records of the traversed volume ([y][x][z][record]): the u16 elemental maps (8 x u16), the FP32 maps
// (8 x FP32), or the spectral channels c_k = sum_m B_mk rho_m of either in FP32, two (float2) or four (float4)
// to a record.
enum VoxKind { VOX_U16 = 0, VOX_F32 = 1, VOX_CH2 = 2, VOX_CH4 = 3 };

// Several spectra over the same rays (dual-energy decomposition): each spectrum's tables, and the channels of
// the joint basis of all of them (the basis of the concatenated attenuation tables).
#define VCBCT_MAX_SPECTRA 4
struct SpecSet {
    const double* mu[VCBCT_MAX_SPECTRA];   // [nmat][NE_s]
    const double* pdf[VCBCT_MAX_SPECTRA];  // [NE_s]
    const float* muf[VCBCT_MAX_SPECTRA];   // [nmat][NE_s] as given
    const double* ch[VCBCT_MAX_SPECTRA];   // [nch][NE_s], the channel table of spectrum s
    const float* chf[VCBCT_MAX_SPECTRA];   // the same in FP32
    int NE[VCBCT_MAX_SPECTRA];
    int nspec, nch;
};

// The mix from the adjoint's channels to the maps: grad_m = sum_k b[m][k] G_k (identity without a basis).
struct ChannelMix {
    float b[VCBCT_MAX_MAT][VCBCT_MAX_MAT];
    int nch;
};

// Spectral basis of an attenuation table mu [nmat][NE] (FP64): the smallest r and orthonormal columns
// B [nmat][r] with max |mu - B (B^T mu)| <= tol * max |mu|, from the eigenvectors of mu mu^T (Jacobi).
// Writes B (row-major [nmat][VCBCT_MAX_MAT], first r columns used) and returns r (nmat when no smaller r does).
int spectral_basis(const double* mu, int nmat, int NE, double tol, double B[VCBCT_MAX_MAT][VCBCT_MAX_MAT]);

// Tile shape of the ray kernels: a block is rgb row groups (64 detector rows each) by `cols` adjacent
// columns of one angle, one wave per (row group, column), rgb * cols in {2, 4, 8}; blocks walk angles inside
// row blocks (order 0) or row blocks inside angles (order 1); the 8 XCDs divide the columns (split 0) or the
// angles (split 1) among them.  rgb == 0: not set (the autotuner picks one per geometry).
struct TileShape {
    int rgb, cols, order, split;
};
// The shape the ray launches use when nothing else is known: 1 x 4, angles inner, columns split.
constexpr TileShape kDefaultTile = {1, 4, 0, 0};
bool tile_shape_valid(const TileShape& s);
// The block -> (row group, angle, column) map of a shape over nA angles (the kernels decode it), and its block
// count, a multiple of 8 (-1 when it exceeds the 32-bit grid).
struct TileMap {
    int rgb, cols, order, split;
    int nrb;  // row blocks, ceil(ceil(H / 64) / rgb)
    int ncg;  // column groups, ceil(W / cols)
    int gpx;  // per XCD: column groups (split 0) or angles (split 1)
};
long tile_map_host(const ProjGeom& g, int nA, const TileShape& s, TileMap& m);

struct vcbct_projector_impl;

// One problem of a batched forward launch (device-resident descriptor).
struct BatchProblem {
    ProjGeom g;
    const double* tab;
    const void* vox;
    const double* mu;
    const double* pdf;
    const float* muf;
    const float* pdff;
    float* out;
    double inv_xf;
    double xfactor;     // as given (the sequence form launches the plain forward with it)
    vcbct_projector* proj;  // host pointer (the sequence form launches through the projector)
    int NE, vox_is_f32;
    // persistent form: the XCD queues hold (row group, angle, column) items; with split 0 queue q holds columns
    // [q pcols, (q + 1) pcols) of every angle (pang = nAng), with split 1 angles [q pang, (q + 1) pang) of
    // every column (pcols = W)
    int split, pcols, pang;
    int xcd_start[8];   // first work item of this problem in each XCD queue
    TileMap tm;         // flat form: the problem's tile map (4-wave shapes)
    int block_start;    // flat form: first block of this problem in its voxel type's grid (a multiple of 8)
};
int launch_forward_batch(const BatchProblem* d_problems, int np, const int* d_xcd_total, int* d_counters, int nblocks,
                         int vox_is_f32, hipStream_t st);
int launch_forward_flat(const BatchProblem* d_problems, int np, int total_blocks, int vox_is_f32, hipStream_t st);

// Kernels launchers (forward.hip)
int launch_transpose_u16(const uint16_t* src, uint16_t* dst, const ProjGeom& g, hipStream_t st);
int launch_transpose_f32(const float* src, float* dst, const ProjGeom& g, hipStream_t st);
int launch_untranspose_f32(const float* src, float* dst, const ProjGeom& g, hipStream_t st);
// The adjoint's channel planes [nch][y][x][z] -> the caller's maps [nmat][z][y][x] through the mix.
int launch_untranspose_mix(const float* src, float* dst, const ProjGeom& g, const ChannelMix& mix, hipStream_t st);
// Forward launch over angles [a0, a1).  tile: in, the caller's fixed shape (rgb > 0) or none (rgb == 0: the
// autotuner's choice for this geometry, timed on the first launch outside stream capture); out, the shape used.
int launch_forward(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const Spectrum& sp,
                   double xfactor, int a0, int a1, float* out, int mode, float* aux, hipStream_t st,
                   TileShape* tile = nullptr);
// The channel volume of a material volume (u16 or FP32 maps in the [y][x][z][8] layout): per voxel
// c_k = sum_m mix.b[m][k] rho_m for k < mix.nch, as float2 (nch <= 2) or float4 records (nch <= 4), zero-padded.
int launch_channels(const void* vox, int src_is_f32, float* dst, const ProjGeom& g, const ChannelMix& mix,
                    hipStream_t st);
// The forward launch of a projector (intensity or log projection, aux: the saved line integrals) from its u16
// volume or its FP32 maps (use_f32), over their channel volume when its spectrum allows it (built on demand;
// rebuild_channels: regardless of the cached one), with the projector's fixed or the autotuned tile shape.
int projector_forward(vcbct_projector* p, int mode, bool use_f32, float* out, float* aux, double xfactor, int a0,
                      int a1, hipStream_t st, bool rebuild_channels);
// The forward autotuner's record (one line per tuned geometry).
std::string forward_tune_log();
// The adjoint launches take the caller's 4-wave tile shape for their traversal kernels (forced.rgb > 0), or
// autotune them per geometry (forced.rgb == 0) as launch_forward does.  vox: the FP32 maps (VOX_F32, the
// spectral sum over sp.mu / sp.muf) or their channel volume (VOX_CH2 / VOX_CH4, over the channel tables).
int launch_adjoint(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const Spectrum& sp,
                   double xfactor,
                   const float* target_or_gp, int mode, float* coef, float* grad_vox, double* loss, hipStream_t st,
                   const TileShape& forced);
// Ls: the sp.nl line integrals per ray logproj_save kept (materials, or channels when sp holds the channel
// tables in mu / muf).
int launch_adjoint_saved(const ProjGeom& g, const double* tables, const float* Ls, const Spectrum& sp,
                         double xfactor, const float* gp, float* coef, float* grad_vox, hipStream_t st,
                         const TileShape& forced);
// loss = sum_s sum_rays (p_s - t_s)^2 over the spectra of `ss` and its gradient: one traversal per ray for all
// spectra (their line integrals are the same), the coefficients summed over the spectra, one scatter.
int launch_adjoint_multi(const ProjGeom& g, const double* tables, const void* vox, int vox_kind, const SpecSet& ss,
                         double xfactor, const float* targets, float* coef, float* grad_vox, double* loss,
                         hipStream_t st, const TileShape& forced);

// a / b for a divisor b known in advance, with rb = 1 / b correctly rounded: the product a * rb is within an
// ulp of a / b, and one correction with the exact remainder a - q b (an FMA) rounds it correctly (Markstein's
// theorem, round-to-nearest, no overflow or underflow), so the result equals the IEEE quotient bit for bit in
// three operations instead of a full division sequence.
__host__ __device__ inline double div_known(double a, double b, double rb) {
    const double q = a * rb;
    const double r = fma(-q, b, a);
    return fma(r, rb, q);
}

// n / d correctly rounded, for d and n / d normal and far from overflow and n normal or zero: the FP64
// division sequence the compiler emits (hardware reciprocal, two Newton steps, quotient, one correction with
// the remainder) without its scaling and special-case steps (v_div_scale, v_div_fixup, and v_div_fmas's
// scaling), which leave such operands unchanged, so the quotient is the IEEE one bit for bit (for n = 0 up to
// the sign of the zero).  Three instructions fewer than n / d.
__device__ inline double div_unscaled(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    const double q = n * r;
    const double res = fma(-d, q, n);
    return fma(res, r, q);
}

// Zero-fill kernel (bytes must be a multiple of 4).
int launch_zero(void* p, size_t bytes, hipStream_t st);

// Half-fan (offset-panel) weighting and short-side padding (angle_mask.hip).  Per projection: the central
// ray's column on the unpadded panel, the transition half-width, the side of the long half.
struct OffsetPar {
    double c;     // column of the central ray (unpadded panel)
    double tau;   // transition half-width (pixels)
    double side;  // +1: the long side is towards higher columns; -1: towards lower
    double pad;
};
// The padded width Wo, the offset of the real columns in it, the per-projection parameters (n entries) and
// the shifts of the padded panel (xs_out / ys_out, n entries each).  Returns 0 or an error code.
int half_fan_params(int W, int n, const double* xshift, const double* yshift, double taper_px, int& Wo, int& off,
                    OffsetPar* par, double* xs_out, double* ys_out);
// in [n][H][W] -> out [n][H][Wo]; d_par: n parameters in device memory.  Allocates nothing.
int launch_offset_weight(const float* in, float* out, int H, int W, int Wo, int off, int n, const OffsetPar* d_par,
                         hipStream_t st);
// The weighting's exponent 2w per (projection, real column), FP32, [n][W] (host).
void offset_exponent_table(const OffsetPar* par, int n, int W, float* out);

// Fault injection for allocation failure (VCBCT_FAULT_ALLOC=<n>: the n-th device allocation fails).
hipError_t dev_malloc(void** p, size_t bytes);

}  // namespace vcbct

// Forward/logproj kernel modes
#define VCBCT_MODE_INTENSITY 0
#define VCBCT_MODE_LINEINT 1
#define VCBCT_MODE_LOGPROJ 2
#define VCBCT_MODE_SEGCOUNT 3
// Adjoint modes
#define VCBCT_ADJ_LOSS 0
#define VCBCT_ADJ_UPSTREAM 1

struct vcbct_projector {
    vcbct::ProjGeom g;
    int device;
    double* d_tables;      // [4][nAng]: cos, sin, Xc, Yc
    uint16_t* d_vox_u16;   // [N][N][Nz][8] (y, x, z, mat)
    float* d_vox_f32;      // [N][N][Nz][8]
    float* d_grad_vox;     // adjoint scatter target, planar [mat][y][x][z]
    float* d_coef;         // adjoint per-ray coefficients [mat][angle][row][col]
    void* d_lp_tab;        // lengthproj per-angle constants (two forms, see iterative.hip)
    double* d_mu;          // [nmat][NE]
    double* d_pdf;         // [NE]
    float* d_muf;          // [nmat][NE] as given
    float* d_pdff;         // [NE] as given
    double* d_ch;          // the adjoint's channel table [nch][NE] (the basis form of mu; unused when nch == nmat)
    float* d_chf;          // the same in FP32 (when nch == nmat: a copy of the table as given)
    vcbct::ChannelMix mix; // channels -> maps
    int fwd_ch;            // the forward and the coefficient kernels traverse the channel volume (nch <= 4)
    int NE;
    int have_u16, have_f32;
    // spectra of the multi-spectrum loss (vcbct_projector_set_spectra), with their joint channel basis
    struct {
        double *d_mu, *d_pdf, *d_ch;
        float *d_muf, *d_chf;
        int NE;
    } ms[VCBCT_MAX_SPECTRA];
    int nspec_multi;
    vcbct::ChannelMix mix_multi;
    int fwd_ch_multi;      // the multi-spectrum coefficient kernel traverses the joint basis's channel volume
    // The channel volume (float2 or float4 records), built on demand from the volume and the basis it was
    // asked for; chv_key names what it holds (0: nothing valid; 1: the u16 volume in the spectrum's basis; 2: the
    // FP32 maps in the spectrum's basis; 3: the FP32 maps in the spectra's joint basis).  Setting a volume or a
    // spectrum clears it.
    float* d_chv;
    size_t chv_bytes;
    int chv_key;
    vcbct::TileShape tile;       // fixed by vcbct_projector_set_tile (rgb == 0: autotuned)
    vcbct::TileShape tile_used;  // the shape of the last intensity / log-projection launch
    vcbct::TileShape adj_tile;   // the adjoint kernels' shape, fixed by vcbct_projector_set_adjoint_tile (4 waves)
};
