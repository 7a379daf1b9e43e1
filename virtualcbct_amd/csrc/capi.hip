// C ABI of the projector: context lifetime, uploads, and the forward / log-projection / adjoint calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {

static thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static std::atomic<long> g_alloc_count{0};

hipError_t dev_malloc(void** p, size_t bytes) {
    // VCBCT_FAULT_ALLOC=n makes the n-th device allocation of the process fail (host-side fault
    // injection for the error paths; no kernel is involved).
    static long fault_at = [] {
        const char* e = getenv("VCBCT_FAULT_ALLOC");
        return e ? atol(e) : 0L;
    }();
    long k = ++g_alloc_count;
    if (fault_at > 0 && k == fault_at) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

int spectral_basis(const double* mu, int nmat, int NE, double tol, double B[VCBCT_MAX_MAT][VCBCT_MAX_MAT]) {
    const int n = nmat;
    // Gram matrix A = mu mu^T and its eigenvectors by cyclic Jacobi rotations (n <= 6)
    double A[VCBCT_MAX_MAT][VCBCT_MAX_MAT] = {}, V[VCBCT_MAX_MAT][VCBCT_MAX_MAT] = {};
    double amax = 0.0;
    for (int i = 0; i < n; ++i) {
        V[i][i] = 1.0;
        for (int j = 0; j < n; ++j) {
            double a = 0.0;
            for (int e = 0; e < NE; ++e) a += mu[(size_t)i * NE + e] * mu[(size_t)j * NE + e];
            A[i][j] = a;
        }
        for (int e = 0; e < NE; ++e) amax = std::max(amax, std::fabs(mu[(size_t)i * NE + e]));
    }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
        }
        if (off <= 1e-32 * diag) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (A[p][q] == 0.0) continue;
                const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J, then J^T A
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - sn * vkq;
                    V[k][q] = sn * vkp + c * vkq;
                }
            }
    }
    int order[VCBCT_MAX_MAT];
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order, order + n, [&](int a, int b) { return A[a][a] > A[b][b]; });
    for (int r = 1; r <= n; ++r) {
        double err = 0.0;
        for (int e = 0; e < NE; ++e) {
            double cfs[VCBCT_MAX_MAT];
            for (int k = 0; k < r; ++k) {
                double s = 0.0;
                for (int m = 0; m < n; ++m) s += V[m][order[k]] * mu[(size_t)m * NE + e];
                cfs[k] = s;
            }
            for (int m = 0; m < n; ++m) {
                double s = 0.0;
                for (int k = 0; k < r; ++k) s += V[m][order[k]] * cfs[k];
                err = std::max(err, std::fabs(s - mu[(size_t)m * NE + e]));
            }
        }
        if (err <= tol * amax || r == n) {
            for (int m = 0; m < VCBCT_MAX_MAT; ++m)
                for (int k = 0; k < VCBCT_MAX_MAT; ++k) B[m][k] = (m < n && k < r) ? V[m][order[k]] : 0.0;
            return r;
        }
    }
    return n;
}

static int check_geometry(const vcbct_geometry* g) {
    if (!g) return fail(VCBCT_EINVAL, "geometry is NULL");
    if (g->recon_size < 2 || g->recon_size > 8192) return fail(VCBCT_EINVAL, "recon_size %d out of range [2, 8192]", g->recon_size);
    if (g->recon_slices < 1 || g->recon_slices > 8192) return fail(VCBCT_EINVAL, "recon_slices %d out of range", g->recon_slices);
    if (g->proj_height < 1 || g->proj_height > 16384 || g->proj_width < 1 || g->proj_width > 16384)
        return fail(VCBCT_EINVAL, "projection size %dx%d out of range", g->proj_width, g->proj_height);
    if (g->n_angles < 1 || g->n_angles > 100000) return fail(VCBCT_EINVAL, "n_angles %d out of range", g->n_angles);
    if (g->num_material < 1 || g->num_material > VCBCT_MAX_MAT)
        return fail(VCBCT_EINVAL, "num_material %d out of range [1, %d]", g->num_material, VCBCT_MAX_MAT);
    if (!(g->recon_scale > 0) || !(g->recon_step > 0) || !(g->sad_cm > 0) || !(g->det_pitch_cm > 0) ||
        !(g->det_pitch_iso_cm > 0) || !(g->det_pixel_num > 0))
        return fail(VCBCT_EINVAL, "geometry pitches and distances must be positive");
    // voxel indices are 32-bit in the ray kernels
    double vox = (double)g->recon_size * g->recon_size * g->recon_slices;
    if (vox >= 2147483648.0) return fail(VCBCT_EINVAL, "volume of %.0f voxels exceeds 2^31 - 1", vox);
    return 0;
}

}  // namespace vcbct

using namespace vcbct;

static void free_multi(vcbct_projector* p);

extern "C" const char* vcbct_last_error(void) { return g_err; }

extern "C" const char* vcbct_version(void) { return "vcbct 0.1 (gfx950)"; }

extern "C" int vcbct_projector_create(vcbct_projector** out, const vcbct_geometry* g, const double* cos_t,
                                      const double* sin_t, const double* xc, const double* yc, int device) {
    if (!out) return fail(VCBCT_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = check_geometry(g);
    if (rc) return rc;
    if (!cos_t || !sin_t || !xc || !yc) return fail(VCBCT_EINVAL, "angle tables are NULL");
    int ndev = 0;
    VCBCT_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(VCBCT_EINVAL, "device %d not present (%d devices)", device, ndev);
    VCBCT_HIP(hipSetDevice(device));
    vcbct_projector* p = (vcbct_projector*)calloc(1, sizeof(vcbct_projector));
    if (!p) return fail(VCBCT_ENOMEM, "host allocation failed");
    p->device = device;
    ProjGeom& G = p->g;
    G.N = g->recon_size;
    G.Nz = g->recon_slices;
    G.H = g->proj_height;
    G.W = g->proj_width;
    G.nAng = g->n_angles;
    G.nmat = g->num_material;
    G.s = g->recon_scale;
    G.t = g->recon_step;
    G.x_reg = g->x_reg;
    G.y_reg = g->y_reg;
    G.z_reg = g->z_reg;
    G.pitch_iso = g->det_pitch_iso_cm;
    G.pitch = g->det_pitch_cm;
    G.sad = g->sad_cm;
    G.pixnum = g->det_pixel_num;
    derive_geom(G);
    std::vector<double> tab(4 * (size_t)G.nAng);
    memcpy(tab.data(), cos_t, G.nAng * sizeof(double));
    memcpy(tab.data() + G.nAng, sin_t, G.nAng * sizeof(double));
    memcpy(tab.data() + 2 * G.nAng, xc, G.nAng * sizeof(double));
    memcpy(tab.data() + 3 * G.nAng, yc, G.nAng * sizeof(double));
    if (dev_malloc((void**)&p->d_tables, tab.size() * sizeof(double)) != hipSuccess) {
        free(p);
        return fail(VCBCT_ENOMEM, "device allocation of angle tables failed");
    }
    if (hipMemcpy(p->d_tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->d_tables);
        free(p);
        return fail(VCBCT_EHIP, "copy of angle tables failed");
    }
    *out = p;
    return 0;
}

extern "C" void vcbct_projector_destroy(vcbct_projector* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipFree(p->d_tables);
    (void)hipFree(p->d_vox_u16);
    (void)hipFree(p->d_vox_f32);
    (void)hipFree(p->d_grad_vox);
    (void)hipFree(p->d_coef);
    (void)hipFree(p->d_lp_tab);
    (void)hipFree(p->d_mu);
    (void)hipFree(p->d_pdf);
    (void)hipFree(p->d_muf);
    (void)hipFree(p->d_pdff);
    (void)hipFree(p->d_ch);
    (void)hipFree(p->d_chf);
    (void)hipFree(p->d_chv);
    free_multi(p);
    free(p);
}

static size_t vox_count(const ProjGeom& g) { return (size_t)g.N * g.N * g.Nz; }

extern "C" int vcbct_projector_set_volume_u16(vcbct_projector* p, const uint16_t* vol, int is_device, void* stream) {
    if (!p || !vol) return fail(VCBCT_EINVAL, "set_volume_u16: NULL argument");
    VCBCT_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = vox_count(p->g);
    if (!p->d_vox_u16 && dev_malloc((void**)&p->d_vox_u16, nv * 16) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the volume (%zu bytes) failed", nv * 16);
    const uint16_t* src = vol;
    uint16_t* stage = nullptr;
    const size_t raw = nv * p->g.nmat * sizeof(uint16_t);
    if (!is_device) {
        if (dev_malloc((void**)&stage, raw) != hipSuccess)
            return fail(VCBCT_ENOMEM, "device allocation of the upload stage (%zu bytes) failed", raw);
        hipError_t e = hipMemcpyAsync(stage, vol, raw, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            (void)hipFree(stage);
            return fail(VCBCT_EHIP, "volume upload failed: %s", hipGetErrorString(e));
        }
        src = stage;
    }
    int rc = launch_transpose_u16(src, p->d_vox_u16, p->g, st);
    if (stage) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(stage);
    }
    if (rc) return rc;
    p->have_u16 = 1;
    p->chv_key = 0;
    return 0;
}

extern "C" int vcbct_projector_set_volume_f32(vcbct_projector* p, const float* dev_maps, void* stream) {
    if (!p || !dev_maps) return fail(VCBCT_EINVAL, "set_volume_f32: NULL argument");
    VCBCT_HIP(hipSetDevice(p->device));
    const size_t nv = vox_count(p->g);
    if (!p->d_vox_f32 && dev_malloc((void**)&p->d_vox_f32, nv * 32) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the FP32 volume (%zu bytes) failed", nv * 32);
    int rc = launch_transpose_f32(dev_maps, p->d_vox_f32, p->g, (hipStream_t)stream);
    if (rc) return rc;
    p->have_f32 = 1;
    p->chv_key = 0;
    return 0;
}

// Whether the forward may sum over the channels of a basis B (r columns) of the table mu [nm][NE]: every
// energy column reproduced to 1e-6 of its largest entry, |mu_me - (B B^T mu)_me| <= 1e-6 max_m' |mu_m'e|, so
// that each bin's exponent sum_m mu_me L_m (L >= 0) is off by at most ~1e-6 of its own size (nonnegative
// tables: the largest entry bounds the sum from below once it is weighted by the line integrals that carry it;
// in practice the bins that dominate a projection hold similar entries).  r == nm: the identity, exact.
static bool basis_columns_ok(const double* mu, int nm, int NE, int r, const double B[VCBCT_MAX_MAT][VCBCT_MAX_MAT]) {
    if (r >= nm) return true;
    for (int e = 0; e < NE; ++e) {
        double cmax = 0.0, rmax = 0.0;
        double proj[VCBCT_MAX_MAT];
        for (int k = 0; k < r; ++k) {
            proj[k] = 0.0;
            for (int m = 0; m < nm; ++m) proj[k] += B[m][k] * mu[(size_t)m * NE + e];
        }
        for (int m = 0; m < nm; ++m) {
            double v = 0.0;
            for (int k = 0; k < r; ++k) v += B[m][k] * proj[k];
            cmax = std::max(cmax, std::fabs(mu[(size_t)m * NE + e]));
            rmax = std::max(rmax, std::fabs(mu[(size_t)m * NE + e] - v));
        }
        if (rmax > 1e-6 * cmax) return false;
    }
    return true;
}

extern "C" int vcbct_projector_set_spectrum(vcbct_projector* p, const float* atten, const float* pdf, int NE) {
    if (!p || !atten || !pdf) return fail(VCBCT_EINVAL, "set_spectrum: NULL argument");
    if (NE < 1 || NE > VCBCT_MAX_NE) return fail(VCBCT_EINVAL, "NE %d out of range [1, %d]", NE, VCBCT_MAX_NE);
    VCBCT_HIP(hipSetDevice(p->device));
    const int nm = p->g.nmat;
    std::vector<double> mu((size_t)nm * NE), pd(NE);
    for (size_t i = 0; i < mu.size(); ++i) mu[i] = atten[i];
    for (int e = 0; e < NE; ++e) pd[e] = pdf[e];
    if (NE != p->NE) {
        (void)hipFree(p->d_mu);
        (void)hipFree(p->d_pdf);
        (void)hipFree(p->d_muf);
        (void)hipFree(p->d_pdff);
        p->d_mu = nullptr;
        p->d_pdf = nullptr;
        p->d_muf = nullptr;
        p->d_pdff = nullptr;
        p->NE = 0;
        if (dev_malloc((void**)&p->d_mu, mu.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&p->d_pdf, pd.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&p->d_muf, mu.size() * sizeof(float)) != hipSuccess ||
            dev_malloc((void**)&p->d_pdff, pd.size() * sizeof(float)) != hipSuccess)
            return fail(VCBCT_ENOMEM, "device allocation of the spectrum failed");
    }
    VCBCT_HIP(hipMemcpy(p->d_mu, mu.data(), mu.size() * sizeof(double), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(p->d_pdf, pd.data(), pd.size() * sizeof(double), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(p->d_muf, atten, mu.size() * sizeof(float), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(p->d_pdff, pdf, pd.size() * sizeof(float), hipMemcpyHostToDevice));
    // The adjoint's channels.  When the table has a lower-rank form (mu = B (B^T mu) to 1e-6 of its largest
    // entry: attenuation in this energy range is close to a photoelectric and a Compton basis function, and
    // exactly so for tables built from those two terms) the gradient is scattered in r < nmat channels with the
    // coefficient table B^T mu and mixed back to the maps per voxel; otherwise the channels are the maps.
    double B[VCBCT_MAX_MAT][VCBCT_MAX_MAT];
    const int r = spectral_basis(mu.data(), nm, NE, 1e-6, B);
    p->mix.nch = r < nm ? r : nm;
    for (int m = 0; m < VCBCT_MAX_MAT; ++m)
        for (int k = 0; k < VCBCT_MAX_MAT; ++k)
            p->mix.b[m][k] = r < nm ? (float)B[m][k] : (m == k && m < nm ? 1.0f : 0.0f);
    (void)hipFree(p->d_ch);
    (void)hipFree(p->d_chf);
    p->d_ch = nullptr;
    p->d_chf = nullptr;
    p->chv_key = 0;
    p->fwd_ch = 0;
    if (r < nm) {
        std::vector<double> ch((size_t)r * NE);
        std::vector<float> chf((size_t)r * NE);
        for (int k = 0; k < r; ++k)
            for (int e = 0; e < NE; ++e) {
                double s = 0.0;
                for (int m = 0; m < nm; ++m) s += B[m][k] * mu[(size_t)m * NE + e];
                ch[(size_t)k * NE + e] = s;
                chf[(size_t)k * NE + e] = (float)s;
            }
        if (dev_malloc((void**)&p->d_ch, ch.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&p->d_chf, chf.size() * sizeof(float)) != hipSuccess)
            return fail(VCBCT_ENOMEM, "device allocation of the channel tables failed");
        VCBCT_HIP(hipMemcpy(p->d_ch, ch.data(), ch.size() * sizeof(double), hipMemcpyHostToDevice));
        VCBCT_HIP(hipMemcpy(p->d_chf, chf.data(), chf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // The forward and the coefficient kernels traverse the channel volume (2 or 4 FP32 channels per voxel in
    // place of 6 u16 or 6 FP32 maps) when there are at most 4 channels and the basis keeps every bin's
    // exponent to ~1e-6 of itself; without a basis the channels are the maps themselves (nmat <= 4).
    p->fwd_ch = (p->mix.nch <= 4 && basis_columns_ok(mu.data(), nm, NE, r, B)) ? 1 : 0;
    p->NE = NE;
    return 0;
}

static void free_multi(vcbct_projector* p) {
    for (int s = 0; s < VCBCT_MAX_SPECTRA; ++s) {
        (void)hipFree(p->ms[s].d_mu);
        (void)hipFree(p->ms[s].d_pdf);
        (void)hipFree(p->ms[s].d_muf);
        (void)hipFree(p->ms[s].d_ch);
        (void)hipFree(p->ms[s].d_chf);
        p->ms[s].d_mu = p->ms[s].d_pdf = p->ms[s].d_ch = nullptr;
        p->ms[s].d_muf = p->ms[s].d_chf = nullptr;
        p->ms[s].NE = 0;
    }
    p->nspec_multi = 0;
    p->fwd_ch_multi = 0;
    if (p->chv_key == 3) p->chv_key = 0;
}

extern "C" int vcbct_projector_set_spectra(vcbct_projector* p, int nspec, const float* atten, const float* pdf,
                                           const int* NE) {
    if (!p || !atten || !pdf || !NE) return fail(VCBCT_EINVAL, "set_spectra: NULL argument");
    if (nspec < 1 || nspec > VCBCT_MAX_SPECTRA)
        return fail(VCBCT_EINVAL, "set_spectra: %d spectra, 1 to %d supported", nspec, VCBCT_MAX_SPECTRA);
    int tot = 0;
    for (int s = 0; s < nspec; ++s) {
        if (NE[s] < 1 || NE[s] > VCBCT_MAX_NE) return fail(VCBCT_EINVAL, "set_spectra: NE[%d] = %d out of range", s, NE[s]);
        tot += NE[s];
    }
    VCBCT_HIP(hipSetDevice(p->device));
    free_multi(p);
    const int nm = p->g.nmat;
    // the joint table [nmat][sum NE] (spectrum s's bins at columns off_s ..), its basis, and per spectrum the
    // FP64 copies and the channel table B^T mu_s
    std::vector<double> joint((size_t)nm * tot);
    std::vector<int> off(nspec + 1, 0);
    for (int s = 0; s < nspec; ++s) off[s + 1] = off[s] + NE[s];
    for (int s = 0; s < nspec; ++s)
        for (int m = 0; m < nm; ++m)
            for (int e = 0; e < NE[s]; ++e)
                joint[(size_t)m * tot + off[s] + e] = atten[(size_t)nm * off[s] + (size_t)m * NE[s] + e];
    double B[VCBCT_MAX_MAT][VCBCT_MAX_MAT];
    const int r = spectral_basis(joint.data(), nm, tot, 1e-6, B);
    const int nch = r < nm ? r : nm;
    p->mix_multi.nch = nch;
    for (int m = 0; m < VCBCT_MAX_MAT; ++m)
        for (int k = 0; k < VCBCT_MAX_MAT; ++k)
            p->mix_multi.b[m][k] = r < nm ? (float)B[m][k] : (m == k && m < nm ? 1.0f : 0.0f);
    for (int s = 0; s < nspec; ++s) {
        const int ne = NE[s];
        const float* at = atten + (size_t)nm * off[s];
        const float* pd = pdf + off[s];
        std::vector<double> mu((size_t)nm * ne), pdd(ne), ch((size_t)nch * ne);
        std::vector<float> chf((size_t)nch * ne);
        for (size_t i = 0; i < mu.size(); ++i) mu[i] = at[i];
        for (int e = 0; e < ne; ++e) pdd[e] = pd[e];
        for (int k = 0; k < nch; ++k)
            for (int e = 0; e < ne; ++e) {
                double v = 0.0;
                if (r < nm)
                    for (int m = 0; m < nm; ++m) v += B[m][k] * mu[(size_t)m * ne + e];
                else
                    v = mu[(size_t)k * ne + e];
                ch[(size_t)k * ne + e] = v;
                // without a basis the FP32 channel table is the table as given
                chf[(size_t)k * ne + e] = r < nm ? (float)v : at[(size_t)k * ne + e];
            }
        auto& t = p->ms[s];
        if (dev_malloc((void**)&t.d_mu, mu.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&t.d_pdf, pdd.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&t.d_muf, mu.size() * sizeof(float)) != hipSuccess ||
            dev_malloc((void**)&t.d_ch, ch.size() * sizeof(double)) != hipSuccess ||
            dev_malloc((void**)&t.d_chf, chf.size() * sizeof(float)) != hipSuccess) {
            free_multi(p);
            return fail(VCBCT_ENOMEM, "device allocation of the spectra failed");
        }
        VCBCT_HIP(hipMemcpy(t.d_mu, mu.data(), mu.size() * sizeof(double), hipMemcpyHostToDevice));
        VCBCT_HIP(hipMemcpy(t.d_pdf, pdd.data(), pdd.size() * sizeof(double), hipMemcpyHostToDevice));
        VCBCT_HIP(hipMemcpy(t.d_muf, at, mu.size() * sizeof(float), hipMemcpyHostToDevice));
        VCBCT_HIP(hipMemcpy(t.d_ch, ch.data(), ch.size() * sizeof(double), hipMemcpyHostToDevice));
        VCBCT_HIP(hipMemcpy(t.d_chf, chf.data(), chf.size() * sizeof(float), hipMemcpyHostToDevice));
        t.NE = ne;
    }
    p->nspec_multi = nspec;
    p->fwd_ch_multi = (nch <= 4 && basis_columns_ok(joint.data(), nm, tot, r, B)) ? 1 : 0;
    return 0;
}

extern "C" int vcbct_projector_multi_channels(const vcbct_projector* p) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (p->nspec_multi < 1) return fail(VCBCT_ESTATE, "no spectra set (vcbct_projector_set_spectra)");
    return p->mix_multi.nch;
}

extern "C" int vcbct_projector_forward_channels(const vcbct_projector* p) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (p->NE < 1) return fail(VCBCT_ESTATE, "no spectrum set (vcbct_projector_set_spectrum)");
    return p->fwd_ch ? p->mix.nch : 0;
}

extern "C" int vcbct_projector_adjoint_channels(const vcbct_projector* p) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (p->NE < 1) return fail(VCBCT_ESTATE, "no spectrum set (vcbct_projector_set_spectrum)");
    return p->mix.nch;
}

static int need(const vcbct_projector* p, bool u16, bool f32, bool spec) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (u16 && !p->have_u16) return fail(VCBCT_ESTATE, "no uint16 volume set (vcbct_projector_set_volume_u16)");
    if (f32 && !p->have_f32) return fail(VCBCT_ESTATE, "no FP32 volume set (vcbct_projector_set_volume_f32)");
    if (spec && p->NE < 1) return fail(VCBCT_ESTATE, "no spectrum set (vcbct_projector_set_spectrum)");
    return 0;
}

static Spectrum spec_of(const vcbct_projector* p) {
    Spectrum s;
    s.mu = p->d_mu;
    s.pdf = p->d_pdf;
    s.muf = p->d_muf;
    s.pdff = p->d_pdff;
    s.NE = p->NE;
    s.ch = p->d_ch ? p->d_ch : p->d_mu;
    s.chf = p->d_chf ? p->d_chf : p->d_muf;
    s.nch = p->mix.nch;
    s.nl = p->g.nmat;
    return s;
}

// The spectrum as the kernels over the channel volume take it: the exponents over the channel tables.
static Spectrum spec_channels(const vcbct_projector* p) {
    Spectrum s = spec_of(p);
    s.mu = s.ch;
    s.muf = s.chf;
    s.nl = s.nch;
    return s;
}

// The channel volume for key (1: the u16 volume, 2: the FP32 maps, both in the spectrum's basis; 3: the FP32 maps
// in the spectra's joint basis), built now unless it already holds that (force: rebuild regardless, for callers
// that replay captured work after the volume may have changed).  Sets *kind to VOX_CH2 / VOX_CH4.
static int ensure_channels(vcbct_projector* p, int key, hipStream_t st, bool force, int* kind) {
    const ChannelMix& mix = key == 3 ? p->mix_multi : p->mix;
    const int c = mix.nch <= 2 ? 2 : 4;
    *kind = c == 2 ? VOX_CH2 : VOX_CH4;
    if (p->chv_key == key && !force) return 0;
    const size_t bytes = vox_count(p->g) * c * sizeof(float);
    if (p->chv_bytes < bytes) {
        (void)hipFree(p->d_chv);
        p->d_chv = nullptr;
        p->chv_bytes = 0;
        p->chv_key = 0;
        if (dev_malloc((void**)&p->d_chv, bytes) != hipSuccess)
            return fail(VCBCT_ENOMEM, "device allocation of the channel volume (%zu bytes) failed", bytes);
        p->chv_bytes = bytes;
    }
    const bool f32 = key != 1;
    int rc = launch_channels(f32 ? (const void*)p->d_vox_f32 : (const void*)p->d_vox_u16, f32 ? 1 : 0, p->d_chv, p->g,
                             mix, st);
    if (rc) return rc;
    p->chv_key = key;
    return 0;
}

namespace vcbct {
// The forward launch of a projector (intensity or log projection) over its channel volume when its spectrum
// allows it, else over its maps; the tile shape fixed on the projector or autotuned.
int projector_forward(vcbct_projector* p, int mode, bool use_f32, float* out, float* aux, double xfactor, int a0,
                      int a1, hipStream_t st, bool rebuild_channels) {
    TileShape ts = p->tile;
    int rc;
    if (p->fwd_ch) {
        int kind = 0;
        rc = ensure_channels(p, use_f32 ? 2 : 1, st, rebuild_channels, &kind);
        if (rc) return rc;
        rc = launch_forward(p->g, p->d_tables, p->d_chv, kind, spec_channels(p), xfactor, a0, a1, out, mode, aux, st,
                            &ts);
    } else {
        rc = launch_forward(p->g, p->d_tables, use_f32 ? (const void*)p->d_vox_f32 : (const void*)p->d_vox_u16,
                            use_f32 ? VOX_F32 : VOX_U16, spec_of(p), xfactor, a0, a1, out, mode, aux, st, &ts);
    }
    if (!rc) p->tile_used = ts;
    return rc;
}
}  // namespace vcbct

extern "C" int vcbct_projector_forward(vcbct_projector* p, float* dev_out, double xfactor, int a0, int a1, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    const bool use_f32 = !p->have_u16 && p->have_f32;
    int rc = need(p, !use_f32, use_f32, true);
    if (rc) return rc;
    if (!dev_out) return fail(VCBCT_EINVAL, "forward: output is NULL");
    if (!(xfactor > 0) || !std::isfinite(xfactor)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    if (a1 < 0) a1 = p->g.nAng;
    if (a0 < 0 || a0 > a1 || a1 > p->g.nAng) return fail(VCBCT_EINVAL, "angle range [%d, %d) invalid", a0, a1);
    VCBCT_HIP(hipSetDevice(p->device));
    return projector_forward(p, VCBCT_MODE_INTENSITY, use_f32, dev_out, nullptr, xfactor, a0, a1, (hipStream_t)stream,
                             false);
}

extern "C" int vcbct_projector_line_integrals(vcbct_projector* p, float* dev_L, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    const bool use_f32 = !p->have_u16 && p->have_f32;
    int rc = need(p, !use_f32, use_f32, false);
    if (rc) return rc;
    if (!dev_L) return fail(VCBCT_EINVAL, "line_integrals: output is NULL");
    VCBCT_HIP(hipSetDevice(p->device));
    Spectrum sp{nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, p->g.nmat, nullptr};
    return launch_forward(p->g, p->d_tables, use_f32 ? (const void*)p->d_vox_f32 : (const void*)p->d_vox_u16,
                          use_f32 ? VOX_F32 : VOX_U16, sp, 1.0, 0, p->g.nAng, nullptr, VCBCT_MODE_LINEINT, dev_L,
                          (hipStream_t)stream);
}

extern "C" int vcbct_reprojection_angles(vcbct_projector* p, const uint16_t* host_vol, const float* atten,
                                         const float* pdf, int NE, double xfactor, int n_out_angles,
                                         float* host_out) {
    if (!p || !host_vol || !host_out) return fail(VCBCT_EINVAL, "reprojection: NULL argument");
    // the reference's grid covers angles [0, endProjNumber) of the tables initializeGPU uploaded
    // (P/reprojectionOnGPU.cu:976); more than those tables hold has no defined result
    if (n_out_angles < 1 || n_out_angles > p->g.nAng)
        return fail(VCBCT_EINVAL, "reprojection: %d angles requested, the projector was created for %d",
                    n_out_angles, p->g.nAng);
    int rc = vcbct_projector_set_volume_u16(p, host_vol, 0, nullptr);
    if (rc) return rc;
    rc = vcbct_projector_set_spectrum(p, atten, pdf, NE);
    if (rc) return rc;
    const size_t n = (size_t)n_out_angles * p->g.H * p->g.W;
    float* d_out = nullptr;
    if (dev_malloc((void**)&d_out, n * sizeof(float)) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the projections (%zu bytes) failed", n * sizeof(float));
    // angle a lands at a * H * W: the first n_out_angles projections fill d_out exactly
    rc = vcbct_projector_forward(p, d_out, xfactor, 0, n_out_angles, nullptr);
    if (rc == 0) {
        hipError_t e = hipMemcpy(host_out, d_out, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(VCBCT_EHIP, "projection download failed: %s", hipGetErrorString(e));
    }
    (void)hipFree(d_out);
    return rc;
}

extern "C" int vcbct_reprojection(vcbct_projector* p, const uint16_t* host_vol, const float* atten, const float* pdf,
                                  int NE, double xfactor, float* host_out) {
    if (!p) return fail(VCBCT_EINVAL, "reprojection: NULL argument");
    return vcbct_reprojection_angles(p, host_vol, atten, pdf, NE, xfactor, p->g.nAng, host_out);
}

extern "C" int vcbct_projector_logproj(vcbct_projector* p, float* dev_p, double xfactor, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    const bool use_f32 = p->have_f32;
    int rc = need(p, !use_f32, use_f32, true);
    if (rc) return rc;
    if (!dev_p) return fail(VCBCT_EINVAL, "logproj: output is NULL");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    VCBCT_HIP(hipSetDevice(p->device));
    return projector_forward(p, VCBCT_MODE_LOGPROJ, use_f32, dev_p, nullptr, xfactor, 0, p->g.nAng,
                             (hipStream_t)stream, false);
}

static int run_adjoint(vcbct_projector* p, const float* tg, float* dev_grad, double* dev_loss, double xfactor, int mode,
                       void* stream) {
    int rc = need(p, false, true, true);
    if (rc) return rc;
    if (!tg || !dev_grad) return fail(VCBCT_EINVAL, "adjoint: NULL argument");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    VCBCT_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = vox_count(p->g);
    if (!p->d_grad_vox && dev_malloc((void**)&p->d_grad_vox, nv * 32) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the gradient volume failed");
    const size_t nrays = (size_t)p->g.nAng * p->g.H * p->g.W;
    if (!p->d_coef && dev_malloc((void**)&p->d_coef, nrays * p->g.nmat * sizeof(float)) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the adjoint coefficients failed");
    // scatter target [mat][y][x][z] (FP32 planes); zero with a fill kernel, not hipMemsetAsync, so the call
    // stays safe under stream capture
    rc = launch_zero(p->d_grad_vox, nv * 4 * p->mix.nch, st);
    if (rc) return rc;
    if (dev_loss && (rc = launch_zero(dev_loss, sizeof(double), st))) return rc;
    if (p->fwd_ch) {  // the coefficient kernel over the channel volume
        int kind = 0;
        rc = ensure_channels(p, 2, st, false, &kind);
        if (!rc)
            rc = launch_adjoint(p->g, p->d_tables, p->d_chv, kind, spec_channels(p), xfactor, tg, mode, p->d_coef,
                                p->d_grad_vox, dev_loss, st, p->adj_tile);
    } else {
        rc = launch_adjoint(p->g, p->d_tables, p->d_vox_f32, VOX_F32, spec_of(p), xfactor, tg, mode, p->d_coef,
                            p->d_grad_vox, dev_loss, st, p->adj_tile);
    }
    if (rc) return rc;
    return launch_untranspose_mix(p->d_grad_vox, dev_grad, p->g, p->mix, st);
}

extern "C" int vcbct_projector_adjoint(vcbct_projector* p, const float* dev_target, float* dev_grad, double* dev_loss,
                                       double xfactor, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    return run_adjoint(p, dev_target, dev_grad, dev_loss, xfactor, VCBCT_ADJ_LOSS, stream);
}

extern "C" int vcbct_projector_logproj_backward(vcbct_projector* p, const float* dev_gp, float* dev_grad, double xfactor,
                                                void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    return run_adjoint(p, dev_gp, dev_grad, nullptr, xfactor, VCBCT_ADJ_UPSTREAM, stream);
}

extern "C" int vcbct_projector_multi_loss_grad(vcbct_projector* p, const float* dev_targets, float* dev_grad,
                                               double* dev_loss, double xfactor, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    int rc = need(p, false, true, false);
    if (rc) return rc;
    if (p->nspec_multi < 1) return fail(VCBCT_ESTATE, "no spectra set (vcbct_projector_set_spectra)");
    if (!dev_targets || !dev_grad) return fail(VCBCT_EINVAL, "multi_loss_grad: NULL argument");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    VCBCT_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = vox_count(p->g);
    if (!p->d_grad_vox && dev_malloc((void**)&p->d_grad_vox, nv * 32) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the gradient volume failed");
    const size_t nrays = (size_t)p->g.nAng * p->g.H * p->g.W;
    if (!p->d_coef && dev_malloc((void**)&p->d_coef, nrays * p->g.nmat * sizeof(float)) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the adjoint coefficients failed");
    SpecSet ss{};
    for (int s = 0; s < p->nspec_multi; ++s) {
        ss.mu[s] = p->ms[s].d_mu;
        ss.pdf[s] = p->ms[s].d_pdf;
        ss.muf[s] = p->ms[s].d_muf;
        ss.ch[s] = p->ms[s].d_ch;
        ss.chf[s] = p->ms[s].d_chf;
        ss.NE[s] = p->ms[s].NE;
    }
    ss.nspec = p->nspec_multi;
    ss.nch = p->mix_multi.nch;
    rc = launch_zero(p->d_grad_vox, nv * 4 * ss.nch, st);
    if (rc) return rc;
    if (dev_loss && (rc = launch_zero(dev_loss, sizeof(double), st))) return rc;
    if (p->fwd_ch_multi) {  // the coefficient kernel over the joint basis's channel volume
        int kind = 0;
        rc = ensure_channels(p, 3, st, false, &kind);
        if (!rc)
            rc = launch_adjoint_multi(p->g, p->d_tables, p->d_chv, kind, ss, xfactor, dev_targets, p->d_coef,
                                      p->d_grad_vox, dev_loss, st, p->adj_tile);
    } else {
        rc = launch_adjoint_multi(p->g, p->d_tables, p->d_vox_f32, VOX_F32, ss, xfactor, dev_targets, p->d_coef,
                                  p->d_grad_vox, dev_loss, st, p->adj_tile);
    }
    if (rc) return rc;
    return launch_untranspose_mix(p->d_grad_vox, dev_grad, p->g, p->mix_multi, st);
}

extern "C" int vcbct_projector_logproj_save(vcbct_projector* p, float* dev_p, float* dev_L, double xfactor,
                                            void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    const bool use_f32 = p->have_f32;
    int rc = need(p, !use_f32, use_f32, true);
    if (rc) return rc;
    if (!dev_p || !dev_L) return fail(VCBCT_EINVAL, "logproj_save: NULL argument");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    VCBCT_HIP(hipSetDevice(p->device));
    return projector_forward(p, VCBCT_MODE_LOGPROJ, use_f32, dev_p, dev_L, xfactor, 0, p->g.nAng, (hipStream_t)stream,
                             false);
}

extern "C" int vcbct_projector_set_tile(vcbct_projector* p, int row_groups, int cols, int order, int split) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (row_groups == 0) {
        p->tile = TileShape{0, 0, 0, 0};
        return 0;
    }
    const TileShape s{row_groups, cols, order, split};
    if (!tile_shape_valid(s))
        return fail(VCBCT_EINVAL, "tile shape %dx%d (order %d, split %d): row groups x columns must be 2, 4 or 8, "
                    "order and split 0 or 1", row_groups, cols, order, split);
    p->tile = s;
    return 0;
}

extern "C" int vcbct_projector_set_adjoint_tile(vcbct_projector* p, int row_groups, int cols, int order, int split) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (row_groups == 0) {
        p->adj_tile = TileShape{0, 0, 0, 0};
        return 0;
    }
    const TileShape s{row_groups, cols, order, split};
    if (!tile_shape_valid(s) || row_groups * cols != 4)
        return fail(VCBCT_EINVAL, "adjoint tile shape %dx%d (order %d, split %d): row groups x columns must be 4, "
                    "order and split 0 or 1", row_groups, cols, order, split);
    p->adj_tile = s;
    return 0;
}

extern "C" int vcbct_projector_get_tile(const vcbct_projector* p, int* shape4) {
    if (!p || !shape4) return fail(VCBCT_EINVAL, "get_tile: NULL argument");
    shape4[0] = p->tile_used.rgb;
    shape4[1] = p->tile_used.cols;
    shape4[2] = p->tile_used.order;
    shape4[3] = p->tile_used.split;
    return 0;
}

extern "C" int vcbct_projector_logproj_backward_saved(vcbct_projector* p, const float* dev_L, const float* dev_gp,
                                                      float* dev_grad, double xfactor, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    int rc = need(p, false, false, true);  // the volume itself is not read: the line integrals are given
    if (rc) return rc;
    if (!dev_L || !dev_gp || !dev_grad) return fail(VCBCT_EINVAL, "logproj_backward_saved: NULL argument");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    VCBCT_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = vox_count(p->g);
    if (!p->d_grad_vox && dev_malloc((void**)&p->d_grad_vox, nv * 32) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the gradient volume failed");
    const size_t nrays = (size_t)p->g.nAng * p->g.H * p->g.W;
    if (!p->d_coef && dev_malloc((void**)&p->d_coef, nrays * p->g.nmat * sizeof(float)) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the adjoint coefficients failed");
    rc = launch_zero(p->d_grad_vox, nv * 4 * p->mix.nch, st);
    if (rc) return rc;
    // dev_L holds what logproj_save kept: the channel line integrals when the forward runs over channels
    rc = launch_adjoint_saved(p->g, p->d_tables, dev_L, p->fwd_ch ? spec_channels(p) : spec_of(p), xfactor, dev_gp,
                              p->d_coef, p->d_grad_vox, st, p->adj_tile);
    if (rc) return rc;
    return launch_untranspose_mix(p->d_grad_vox, dev_grad, p->g, p->mix, st);
}

extern "C" int vcbct_projector_segment_counts(vcbct_projector* p, float* dev_counts, void* stream) {
    if (!p) return fail(VCBCT_EINVAL, "projector is NULL");
    if (!dev_counts) return fail(VCBCT_EINVAL, "segment_counts: output is NULL");
    VCBCT_HIP(hipSetDevice(p->device));
    Spectrum sp{nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, p->g.nmat, nullptr};
    return launch_forward(p->g, p->d_tables, nullptr, VOX_U16, sp, 1.0, 0, p->g.nAng, nullptr, VCBCT_MODE_SEGCOUNT,
                          dev_counts, (hipStream_t)stream);
}
