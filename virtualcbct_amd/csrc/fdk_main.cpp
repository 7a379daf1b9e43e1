// vcbct_fdk <info.txt>: drop-in for the reference's cbct2 (F/main.cpp:13-247).  Reads <dir>.<ext>
// (FP32 [th][row][col] projections), filters, writes fltsngrm.raw, backprojects and writes the int16
// volume ((short) truncation of the FP64 accumulation, F/main.cpp:179-190).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vcbct.h"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, vcbct_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "\n[usage]$ vcbct_fdk <info filename>\n\n");
        return 2;
    }
    vcbct_fdk_config c;
    if (vcbct_read_fdk_config(argv[1], &c)) return die("loadData");
    const int n = c.total_proj, H = c.proj_height, W = c.proj_width, R = c.recon_size, Z = c.recon_slices;
    std::vector<double> ang(n), xs(n), ys(n);
    int used = vcbct_read_fdk_angles(c.flexmap, c.tgt_class, n, ang.data(), xs.data(), ys.data());
    if (used < 0) return die("angle data file");
    if (used != n) fprintf(stderr, "WARNING: total projection number mismatch (%d != %d)\n", used, n);
    char fn[1024];
    snprintf(fn, sizeof fn, "%s.%s", c.proj_path, c.ext);
    FILE* fp = fopen(fn, "rb");
    if (!fp) {
        fprintf(stderr, "ERROR: projection file %s not readable\n", fn);
        return 1;
    }
    std::vector<float> proj((size_t)n * H * W), flt((size_t)n * H * W);
    size_t got = fread(proj.data(), sizeof(float), proj.size(), fp);
    fclose(fp);
    if (got != proj.size()) {
        fprintf(stderr, "ERROR: projection file %s is short (%zu of %zu floats)\n", fn, got, proj.size());
        return 1;
    }
    if (vcbct_filter_sinogram_host(H, W, n, proj.data(), flt.data(), xs.data(), ys.data(), c.det_pitch_iso_cm))
        return die("filterSinogram");
    if ((fp = fopen("fltsngrm.raw", "wb")) == nullptr) {
        fprintf(stderr, "ERROR: fltsngrm file NOT writable\n");
        return 1;
    }
    fwrite(flt.data(), sizeof(float), flt.size(), fp);
    fclose(fp);
    std::vector<double> rec((size_t)Z * R * R, 0.0);
    if (vcbct_reconstruct_image_host(R, Z, c.recon_offset_cm, c.recon_step, c.recon_scale, rec.data(), H, W, n,
                                     flt.data(), c.xfactor, ang.data(), xs.data(), c.det_pitch_iso_cm))
        return die("reconstructImage");
    std::vector<short> s(rec.size());
    for (size_t i = 0; i < rec.size(); ++i) s[i] = (short)rec[i];
    if ((fp = fopen(c.output, "wb")) == nullptr) {
        fprintf(stderr, "ERROR: output file NOT writable\n");
        return 1;
    }
    fwrite(s.data(), sizeof(short), s.size(), fp);
    fclose(fp);
    fprintf(stderr, "wrote %s (%d x %d x %d int16)\n", c.output, Z, R, R);
    return 0;
}
