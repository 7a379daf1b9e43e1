// vcbct_fdk <info.txt>: drop-in for the reference's cbct2 (F/main.cpp:13-247).  Reads <dir>.<ext>
// (FP32 [th][row][col] projections), filters, writes fltsngrm.raw, backprojects and writes the int16
// volume ((short) truncation of the FP64 accumulation, F/main.cpp:179-190).
// VCBCT_HALF_FAN=1 (not in the reference) weights an offset-panel scan and pads its short side before the
// filter (vcbct_fdk_offset_weight; VCBCT_HALF_FAN_TAPER=<pixels> narrows the transition); fltsngrm.raw then
// holds the padded panel.
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
    std::vector<float> proj((size_t)n * H * W), flt;
    size_t got = fread(proj.data(), sizeof(float), proj.size(), fp);
    fclose(fp);
    if (got != proj.size()) {
        fprintf(stderr, "ERROR: projection file %s is short (%zu of %zu floats)\n", fn, got, proj.size());
        return 1;
    }
    int Wf = W;  // columns of the panel that is filtered and backprojected
    double pitch = c.det_pitch_iso_cm;
    const char* hf = getenv("VCBCT_HALF_FAN");
    if (hf && hf[0] == '1') {
        const char* tp = getenv("VCBCT_HALF_FAN_TAPER");
        Wf = vcbct_fdk_half_fan_width(W, n, xs.data());
        if (Wf < 0) return die("half-fan weighting");
        std::vector<float> padded((size_t)n * H * Wf);
        std::vector<double> xs2(n), ys2(n);
        if (vcbct_fdk_offset_weight_host(H, W, n, proj.data(), xs.data(), ys.data(), tp ? atof(tp) : 0.0, Wf,
                                         padded.data(), xs2.data(), ys2.data()))
            return die("half-fan weighting");
        proj.swap(padded);
        xs.swap(xs2);
        ys.swap(ys2);
        pitch = c.det_pitch_iso_cm * Wf / W;  // the same pixel size on the wider panel
        fprintf(stderr, "half-fan weighting: panel padded from %d to %d columns\n", W, Wf);
    }
    flt.resize((size_t)n * H * Wf);
    if (vcbct_filter_sinogram_host(H, Wf, n, proj.data(), flt.data(), xs.data(), ys.data(), pitch))
        return die("filterSinogram");
    if ((fp = fopen("fltsngrm.raw", "wb")) == nullptr) {
        fprintf(stderr, "ERROR: fltsngrm file NOT writable\n");
        return 1;
    }
    fwrite(flt.data(), sizeof(float), flt.size(), fp);
    fclose(fp);
    std::vector<double> rec((size_t)Z * R * R, 0.0);
    if (vcbct_reconstruct_image_host(R, Z, c.recon_offset_cm, c.recon_step, c.recon_scale, rec.data(), H, Wf, n,
                                     flt.data(), c.xfactor, ang.data(), xs.data(), pitch))
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
