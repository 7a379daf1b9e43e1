// vcbct_proj <info.txt>: drop-in for the reference's proj.exe (P/main.cpp:15-175), GPU build.
// VCBCT_NOISE_I0 / VCBCT_NOISE_SEED (not in the reference): Poisson quantum noise on the output (below).
// Reads the same txt, spectrum CSV, xraydb tables (attenuation_xraydb/ relative to the working directory),
// flexmap and elemental slices, and writes <outputname>.raw (FP32 [angle][row][col]).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vcbct.h"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, vcbct_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "\n[usage]$ vcbct_proj <info filename>\n\n");
        return 2;
    }
    vcbct_proj_config c;
    if (vcbct_read_proj_config(argv[1], &c)) return die("loadData");
    if (c.mat_num < 1 || c.mat_num > VCBCT_MAX_MAT) {
        fprintf(stderr, "mat_num %d unsupported\n", c.mat_num);
        return 1;
    }
    // angles (P/projectionData.cpp:152-171)
    std::vector<double> ang(c.total_proj), xs(c.total_proj), ys(c.total_proj);
    int nrow = vcbct_read_flexmap(c.flexmap, ang.data(), xs.data(), ys.data(), c.total_proj);
    if (nrow < 0) return die("angle data file");
    if (nrow < c.total_proj) fprintf(stderr, "WARNING: flexmap has %d rows for %d projections\n", nrow, c.total_proj);
    // slice window and stack (P/projectionData.cpp:173-218)
    int lo, hi;
    if (vcbct_proj_slice_window(&c, &lo, &hi)) return die("slice window");
    const int usedZ = hi - lo;
    std::vector<uint16_t> vol((size_t)usedZ * c.mat_num * c.recon_size * c.recon_size);
    int missing = 0;
    if (vcbct_load_material_slices(c.mat_prefix, lo, hi, c.recon_size, c.mat_num, 1, vol.data(), &missing))
        return die("material slices");
    if (missing) fprintf(stderr, "WARNING: %d slice files missing or short; previous slice reused\n", missing);
    const int zic = c.zic - lo;
    // spectrum and mu table (P/main.cpp:56-64, nthin = 1)
    std::vector<float> atten(1024 * c.mat_num), pdf(1024);
    int NE = vcbct_attenuation_table(c.spectrum, "attenuation_xraydb", c.mat_num, 1, atten.data(), pdf.data(), 1024);
    if (NE < 0) return die("attenuation_coefficient");
    // tables and registration (P/main.cpp:88-103)
    std::vector<double> cs(c.total_proj), sn(c.total_proj), xc(c.total_proj), yc(c.total_proj);
    if (vcbct_trig_table(c.proj_width, c.proj_height, c.total_proj, ang.data(), xs.data(), ys.data(), c.det_pitch_cm,
                         c.det_pixel_num, cs.data(), sn.data(), xc.data(), yc.data()))
        return die("prepareTrigonometricTable");
    const double z_reg = (usedZ / 2 - zic) * c.z_scale;
    fprintf(stderr, "matImage num = %d, center = %d, z_reg = %f\n", usedZ, zic, z_reg);
    vcbct_geometry g;
    g.recon_size = c.recon_size;
    g.recon_scale = c.x_scale;
    g.recon_slices = usedZ;
    g.recon_step = c.z_scale;
    g.proj_height = c.proj_height;
    g.proj_width = c.proj_width;
    g.n_angles = c.total_proj;
    g.x_reg = c.x_reg;
    g.y_reg = c.y_reg;
    g.z_reg = z_reg;
    g.det_pitch_cm = c.det_pitch_cm;
    g.det_pitch_iso_cm = c.det_pitch_iso_cm;
    g.sad_cm = c.sad_cm;
    g.det_pixel_num = c.det_pixel_num;
    g.num_material = c.mat_num;
    const char* dv = getenv("VCBCT_DEVICE");
    vcbct_projector* p = nullptr;
    if (vcbct_projector_create(&p, &g, cs.data(), sn.data(), xc.data(), yc.data(), dv ? atoi(dv) : 0))
        return die("initializeGPU");
    std::vector<float> out((size_t)c.total_proj * c.proj_width * c.proj_height);
    if (vcbct_reprojection(p, vol.data(), atten.data(), pdf.data(), NE, c.xfactor, out.data())) {
        vcbct_projector_destroy(p);
        return die("reprojectionOnGPU");
    }
    vcbct_projector_destroy(p);
    // VCBCT_NOISE_I0=<quanta> (not in the reference): Poisson counts of I0 x the projection per pixel, written
    // as max(k, 1) / I0 (vcbct_detector_counts; VCBCT_NOISE_SEED=<integer> picks the generator's key)
    const char* ni = getenv("VCBCT_NOISE_I0");
    if (ni && atof(ni) > 0.0) {
        const char* nsd = getenv("VCBCT_NOISE_SEED");
        const unsigned long long seed = nsd ? strtoull(nsd, nullptr, 10) : 0ull;
        if (vcbct_detector_counts_host(out.size(), out.data(), nullptr, atof(ni), seed, 0, out.data()))
            return die("detector model");
        fprintf(stderr, "detector model: Poisson counts at I0 = %g, seed %llu\n", atof(ni), seed);
    }
    char fn[1024];
    snprintf(fn, sizeof fn, "%s.raw", c.outputname);
    FILE* fp = fopen(fn, "wb");
    if (!fp || fwrite(out.data(), sizeof(float), out.size(), fp) != out.size()) {
        fprintf(stderr, "cannot write %s\n", fn);
        if (fp) fclose(fp);
        return 1;
    }
    fclose(fp);
    fprintf(stderr, "wrote %s (%d x %d x %d float32)\n", fn, c.total_proj, c.proj_height, c.proj_width);
    return 0;
}
