// Organ-label phantom -> six elemental maps (H, C, N, O, P, Ca), the step before the projector.
//
// Replaces ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp (R/ below): the media and
// organ tables are read and (optionally) perturbed on the host exactly as R/ :104-201 does, then reduced
// to a 256-entry table indexed by the 8-bit organ label, whose entries are already in the projector's
// 16-byte voxel format.  The per-voxel work (R/ :212-226, an O(voxels x organs) scan on the host) becomes
// one LDS lookup per voxel on the device, written either in the reference's per-slice file layout or
// straight into the projector's [y][x][z][8] volume (crop / pad placement fused in).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {

// ------------------------------------------------------------------------------------------------
// Host: tables.

// glibc's rand() sequence (the additive-feedback generator behind random(), 128-byte state) on a
// private state, so the caller's rand() stream is left alone.
struct GlibcRand {
    random_data d;
    int32_t state[32];
    explicit GlibcRand(unsigned seed) {
        memset(&d, 0, sizeof(d));
        memset(state, 0, sizeof(state));
        initstate_r(seed, (char*)state, sizeof(state), &d);
    }
    // Uniform0 (R/ :14-17): (rand() + 1) / (RAND_MAX + 2), never 0 or 1
    double uniform() {
        int32_t v = 0;
        random_r(&d, &v);
        return ((double)v + 1.0) / ((double)RAND_MAX + 2.0);
    }
    // rand_gauss0 (R/ :19-24): Box-Muller with the reference's truncated pi.  The first draw feeds the
    // log term, the second the sine (the order the reference's compiled expression evaluates them in).
    double gauss(double mu, double sigma) {
        const double pi = 3.14159265;
        const double u1 = uniform();
        const double r = sqrt(-2.0 * log(u1));
        const double u2 = uniform();
        const double z = r * sin(2.0 * pi * u2);
        return mu + sigma * z;
    }
};

struct Medium {
    int id;
    float w[13];  // H C N O Na Mg P S Cl K Ca Fe I
};
enum { W_H = 0, W_C = 1, W_N = 2, W_O = 3, W_P = 6, W_CA = 10 };

// one weight redrawn from N(w, 0.1 w), clamped at 0 (R/ :121-124)
static float perturb(GlibcRand& rng, float w) {
    const double sigma = (double)w * 0.1;
    float v = (float)rng.gauss((double)w, sigma);
    if (v < 0.0) v = 0.0f;
    return v;
}

static int read_media(const char* path, std::vector<Medium>& out) {
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "media table %s not found", path);
    for (;;) {
        Medium m;
        float* w = m.w;
        int n = fscanf(fp, "%d,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f", &m.id, w, w + 1, w + 2, w + 3, w + 4, w + 5,
                       w + 6, w + 7, w + 8, w + 9, w + 10, w + 11, w + 12);
        if (n == EOF) break;
        if (n != 14) {
            fclose(fp);
            return fail(VCBCT_EIO, "media table %s: row %zu is malformed", path, out.size() + 1);
        }
        out.push_back(m);
    }
    fclose(fp);
    if (out.empty()) return fail(VCBCT_EIO, "media table %s is empty", path);
    return 0;
}

struct Organ {
    int label, medium;
    float dens;
};

static int read_organs(const char* path, std::vector<Organ>& out) {
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "organ table %s not found", path);
    for (;;) {
        Organ o;
        int n = fscanf(fp, "%d,%d,%f", &o.label, &o.medium, &o.dens);
        if (n == EOF) break;
        if (n != 3) {
            fclose(fp);
            return fail(VCBCT_EIO, "organ table %s: row %zu is malformed", path, out.size() + 1);
        }
        out.push_back(o);
    }
    fclose(fp);
    if (out.empty()) return fail(VCBCT_EIO, "organ table %s is empty", path);
    return 0;
}

}  // namespace vcbct

using namespace vcbct;

extern "C" int vcbct_material_table(const char* media_csv, const char* organs_csv, int randomize, unsigned int seed,
                                    uint16_t* lut, float* media_w, int media_cap) {
    if (!media_csv || !organs_csv || !lut || (media_w && media_cap < 0))
        return fail(VCBCT_EINVAL, "material_table: bad arguments");
    std::vector<Medium> media;
    std::vector<Organ> organs;
    int rc = read_media(media_csv, media);
    if (rc) return rc;
    rc = read_organs(organs_csv, organs);
    if (rc) return rc;
    const int nm = (int)media.size();
    if (media_w && media_cap < nm) return fail(VCBCT_EINVAL, "material_table: %d media rows, capacity %d", nm, media_cap);

    // per-medium weights H, C, N, O, P, Ca (R/ :117-165), rows in file order
    std::vector<float> W((size_t)nm * 6);
    GlibcRand rng(seed);
    for (int i = 0; i < nm; ++i) {
        float H = media[i].w[W_H], Cw = media[i].w[W_C], Nw = media[i].w[W_N];
        float P = media[i].w[W_P], Ca = media[i].w[W_CA];
        if (randomize && media[i].id != 53) {
            H = perturb(rng, H);
            Cw = perturb(rng, Cw);
            Nw = perturb(rng, Nw);
            if (P > 0.09) P = perturb(rng, P);
            if (Ca > 0.09) Ca = perturb(rng, Ca);
        }
        // oxygen takes the remainder of the six (the other elements' share included); FP32 as the reference
        float O = 100.0f - H - Cw - Nw - P - Ca;
        if (O < 0.0) {
            const double s = (double)(H + Cw + Nw + P + Ca) * 0.01;
            O = 0.0f;
            H = (float)(H / s);
            Cw = (float)(Cw / s);
            Nw = (float)(Nw / s);
            P = (float)(P / s);
            Ca = (float)(Ca / s);
        }
        const float v[6] = {H, Cw, Nw, O, P, Ca};
        for (int m = 0; m < 6; ++m) W[(size_t)i * 6 + m] = v[m];
    }

    // label -> six values; the last organ row carrying a label wins (R/ :214-225)
    std::vector<uint16_t> tab(256 * 8, 0);
    for (const Organ& o : organs) {
        if (o.label < 0 || o.label > 255) continue;  // an 8-bit label never equals it
        if (o.medium < 1 || o.medium > nm)
            return fail(VCBCT_EIO, "organ %d names medium %d; the media table has %d rows", o.label, o.medium, nm);
        for (int m = 0; m < 6; ++m) {
            const float v = W[(size_t)(o.medium - 1) * 6 + m] * o.dens * 10;
            if (!(v > -1.0f && v < 32768.0f))
                return fail(VCBCT_EIO, "organ %d: map value %g does not fit a 16-bit voxel", o.label, (double)v);
            tab[(size_t)o.label * 8 + m] = (uint16_t)(short)v;
        }
    }
    memcpy(lut, tab.data(), tab.size() * sizeof(uint16_t));
    if (media_w) memcpy(media_w, W.data(), W.size() * sizeof(float));
    return nm;
}

namespace vcbct {

// ------------------------------------------------------------------------------------------------
// Device: lookups.

// labels [nz][n2] -> maps [nz][6][n2]; one voxel per thread per iteration, the table in LDS.  Grid (bx, nz):
// a few thousand blocks in all, each striding over its slice, so that the table is loaded once per
// several thousand voxels.
__global__ __launch_bounds__(256) void k_labels_to_maps(const uint8_t* __restrict__ labels,
                                                        const uint4* __restrict__ lut, uint16_t* __restrict__ maps,
                                                        int n2) {
    __shared__ uint4 tab[256];
    tab[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const size_t z = blockIdx.y;
    const uint8_t* src = labels + z * n2;
    uint16_t* dst = maps + z * 6 * (size_t)n2;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n2; i += gridDim.x * 256) {
        const uint4 e = tab[src[i]];
        dst[i] = (uint16_t)e.x;
        dst[(size_t)n2 + i] = (uint16_t)(e.x >> 16);
        dst[2 * (size_t)n2 + i] = (uint16_t)e.y;
        dst[3 * (size_t)n2 + i] = (uint16_t)(e.y >> 16);
        dst[4 * (size_t)n2 + i] = (uint16_t)e.z;
        dst[5 * (size_t)n2 + i] = (uint16_t)(e.z >> 16);
    }
}

// labels [nz][nrow][ncol] -> projector voxels [N][N][Nz] (y, x, z; 16 bytes each).  A block covers one y
// row and LX = 32 x: the labels of up to LZC slices are read along x (one slice row at a time) into LDS,
// then the block writes its output, which for fixed y and consecutive x is one contiguous range of
// memory, linearly (whole cache lines, 4 KB per block-wide store).  Grid: (ceil(N/LX), N).
constexpr int LT = 32;
constexpr int LX = 32;
constexpr int LZC = 1024;

// LDS row length of one x for z chunks of up to zmax slices: a multiple of 4 bytes whose word count is
// odd, so the 32 x of one z fall in distinct banks.
static inline int label_pitch(int zmax) {
    int p = (zmax + 3) / 4 * 4;
    if ((p / 4) % 2 == 0) p += 4;
    return p;
}

__global__ __launch_bounds__(256) void k_labels_to_vox(const uint8_t* __restrict__ labels, int ncol, int nrow, int nz,
                                                       int col0, int row0, int slice0, const uint4* __restrict__ lut,
                                                       uint4 keep, uint4* __restrict__ vox, int N, int Nz, int pitch) {
    __shared__ uint4 tab[256];
    extern __shared__ uint8_t tile[];  // [LX][pitch]
    {
        uint4 e = lut[threadIdx.x];
        e.x &= keep.x;
        e.y &= keep.y;
        e.z &= keep.z;
        e.w &= keep.w;
        tab[threadIdx.x] = e;
    }
    const int x0 = blockIdx.x * LX, y = blockIdx.y;
    const int nx = min(LX, N - x0);
    const int row = y + row0;
    const bool row_in = row >= 0 && row < nrow;
    const size_t n2 = (size_t)nrow * ncol;
    uint4* out = vox + ((size_t)y * N + x0) * Nz;
    for (int zc = 0; zc < Nz; zc += LZC) {
        const int nzc = min(LZC, Nz - zc);
        for (int j = threadIdx.x; j < nzc * LX; j += 256) {
            const int zz = j / LX, xx = j % LX;
            const int sl = zc + zz + slice0, col = x0 + xx + col0;
            uint8_t v = 0;
            if (row_in && col >= 0 && col < ncol && sl >= 0 && sl < nz) v = labels[(size_t)sl * n2 + (size_t)row * ncol + col];
            tile[xx * pitch + zz] = v;
        }
        __syncthreads();
        // j -> (x, z) with z fastest; the FP32 quotient is exact for j < 32 * nzc and nzc <= 1024
        const float inv = 1.0f / (float)nzc;
        for (int j = threadIdx.x; j < nx * nzc; j += 256) {
            const int xx = (int)(((float)j + 0.5f) * inv);
            const int zz = j - xx * nzc;
            out[(size_t)xx * Nz + zc + zz] = tab[tile[xx * pitch + zz]];
        }
        __syncthreads();
    }
}

// projector voxels [N][N][Nz] -> [Nz][nmat][N][N] u16: the inverse of the upload re-layout.  A block
// stages 32 x-columns of 32 z-voxels of one y row in LDS, then writes 32-wide x-runs of each material.
__global__ __launch_bounds__(256) void k_vox_to_u16(const uint4* __restrict__ vox, uint16_t* __restrict__ dst, int N,
                                                    int Nz, int nmat) {
    __shared__ uint4 tile[LT][LT + 1];
    const int x0 = blockIdx.x * LT, z0 = blockIdx.y * LT, y = blockIdx.z;
    for (int i = threadIdx.x; i < LT * LT; i += 256) {
        const int xx = i / LT, zz = i % LT;
        const int x = x0 + xx, z = z0 + zz;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (x < N && z < Nz) v = vox[((size_t)y * N + x) * Nz + z];
        tile[xx][zz] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x % LT, ty = threadIdx.x / LT;
    const int x = x0 + tx;
    if (x >= N) return;
    for (int zz = ty; zz < LT; zz += 256 / LT) {
        const int z = z0 + zz;
        if (z >= Nz) break;
        const uint4 v = tile[tx][zz];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        for (int m = 0; m < nmat; ++m)
            dst[(((size_t)z * nmat + m) * N + y) * N + x] = (uint16_t)(w[m / 2] >> (16 * (m % 2)));
    }
}

}  // namespace vcbct

extern "C" int vcbct_labels_to_maps(const uint8_t* labels, int ncol, int nrow, int nz, const uint16_t* lut,
                                    uint16_t* maps, void* stream) {
    if (!labels || !lut || !maps) return fail(VCBCT_EINVAL, "labels_to_maps: NULL argument");
    if (ncol < 1 || nrow < 1 || nz < 1 || nz > 65535 || (double)ncol * nrow >= 2147483648.0 - 65536.0 * 256)
        return fail(VCBCT_EINVAL, "labels_to_maps: label grid %d x %d x %d out of range", ncol, nrow, nz);
    const int n2 = ncol * nrow;
    const unsigned bx = (unsigned)std::min((n2 + 255) / 256, std::max(1, (4096 + nz - 1) / nz));
    hipLaunchKernelGGL(k_labels_to_maps, dim3(bx, nz), dim3(256), 0, (hipStream_t)stream, labels, (const uint4*)lut,
                       maps, n2);
    VCBCT_HIP(hipGetLastError());
    return 0;
}

extern "C" int vcbct_projector_set_volume_labels(vcbct_projector* p, const uint8_t* labels, int ncol, int nrow, int nz,
                                                 int col0, int row0, int slice0, const uint16_t* lut, void* stream) {
    if (!p || !labels || !lut) return fail(VCBCT_EINVAL, "set_volume_labels: NULL argument");
    if (ncol < 1 || nrow < 1 || nz < 1 || (double)ncol * nrow * nz >= 9.2e18)
        return fail(VCBCT_EINVAL, "set_volume_labels: label grid %d x %d x %d out of range", ncol, nrow, nz);
    VCBCT_HIP(hipSetDevice(p->device));
    const ProjGeom& g = p->g;
    const size_t nv = (size_t)g.N * g.N * g.Nz;
    if (!p->d_vox_u16 && dev_malloc((void**)&p->d_vox_u16, nv * 16) != hipSuccess)
        return fail(VCBCT_ENOMEM, "device allocation of the volume (%zu bytes) failed", nv * 16);
    // materials beyond num_material are zero in the device volume, as the u16 upload leaves them
    uint32_t k[4] = {0, 0, 0, 0};
    for (int m = 0; m < g.nmat; ++m) k[m / 2] |= 0xffffu << (16 * (m % 2));
    const uint4 keep = make_uint4(k[0], k[1], k[2], k[3]);
    dim3 grid((g.N + LX - 1) / LX, g.N);
    const int pitch = label_pitch(std::min(g.Nz, LZC));
    hipLaunchKernelGGL(k_labels_to_vox, grid, dim3(256), (size_t)LX * pitch, (hipStream_t)stream, labels, ncol, nrow,
                       nz, col0, row0, slice0, (const uint4*)lut, keep, (uint4*)p->d_vox_u16, g.N, g.Nz, pitch);
    VCBCT_HIP(hipGetLastError());
    p->have_u16 = 1;
    p->chv_key = 0;
    return 0;
}

extern "C" int vcbct_projector_get_volume_u16(vcbct_projector* p, uint16_t* vol, void* stream) {
    if (!p || !vol) return fail(VCBCT_EINVAL, "get_volume_u16: NULL argument");
    if (!p->have_u16) return fail(VCBCT_ESTATE, "get_volume_u16: no u16 volume has been set");
    VCBCT_HIP(hipSetDevice(p->device));
    const ProjGeom& g = p->g;
    dim3 grid((g.N + LT - 1) / LT, (g.Nz + LT - 1) / LT, g.N);
    hipLaunchKernelGGL(k_vox_to_u16, grid, dim3(256), 0, (hipStream_t)stream, (const uint4*)p->d_vox_u16, vol, g.N,
                       g.Nz, g.nmat);
    VCBCT_HIP(hipGetLastError());
    return 0;
}
