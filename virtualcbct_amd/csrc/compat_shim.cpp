// The reference's C++-linkage entry points, re-exported over the C ABI so that the reference's own host
// programs link unchanged against libvcbct_compat.so:
//   initializeGPU / reprojectionOnGPU / terminateGPU   (P/reprojectionOnGPU.h:2-19)
//   prepareConstMap / filterSinogram                   (F/filterSinogram.h:2-6)
//   reconstructImage                                   (F/reconstructImage.h:2-5)
// Behaviour follows the reference (same output file, same array ownership), except that failures print
// the library's message and exit with status 1 instead of 0.  The device is VCBCT_DEVICE (default 0),
// not the reference's hard-coded device 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vcbct.h"

namespace {
vcbct_projector* g_proj = nullptr;
int g_width = 0, g_height = 0;  // projection size fixed by initializeGPU

[[noreturn]] void die(const char* where) {
    fprintf(stderr, "%s: %s\n", where, vcbct_last_error());
    exit(1);
}
}  // namespace

void initializeGPU(int reconSize, double reconScale, int reconSlices, double reconStep, int projImgHeight,
                   int projImgWidth, int totalProjNumber, double* cosTable, double* sinTable, double* XcenterShift,
                   double* YcenterShift, double x_reg, double y_reg, double z_reg, double DETECTOR_PITCH_CM,
                   double DETECTOR_PITCH_CM_AT_ISO_CENT, double DIST_BTWN_SRC_AND_ISOCENT_CM,
                   double DETECTOR_PIXEL_NUM, int num_material) {
    vcbct_geometry g;
    g.recon_size = reconSize;
    g.recon_scale = reconScale;
    g.recon_slices = reconSlices;
    g.recon_step = reconStep;
    g.proj_height = projImgHeight;
    g.proj_width = projImgWidth;
    g.n_angles = totalProjNumber;
    g.x_reg = x_reg;
    g.y_reg = y_reg;
    g.z_reg = z_reg;
    g.det_pitch_cm = DETECTOR_PITCH_CM;
    g.det_pitch_iso_cm = DETECTOR_PITCH_CM_AT_ISO_CENT;
    g.sad_cm = DIST_BTWN_SRC_AND_ISOCENT_CM;
    g.det_pixel_num = DETECTOR_PIXEL_NUM;
    g.num_material = num_material;
    const char* dv = getenv("VCBCT_DEVICE");
    if (g_proj) vcbct_projector_destroy(g_proj);
    if (vcbct_projector_create(&g_proj, &g, cosTable, sinTable, XcenterShift, YcenterShift, dv ? atoi(dv) : 0))
        die("initializeGPU");
    g_width = projImgWidth;
    g_height = projImgHeight;
}

void terminateGPU() {
    vcbct_projector_destroy(g_proj);
    g_proj = nullptr;
}

void reprojectionOnGPU(int ite, int reconSize, int reconSlices, unsigned short* reconImageshort,
                       float* reprojection_float, float* npf_double, float* projVolume, double reconScale,
                       double reconStep, int projImgWidth, int projImgHeight, int startProjNumber, int endProjNumber,
                       double* anglesRad, double* cosTable, double* sinTable, double* XcenterShift,
                       double* YcenterShift, double xfactor, double* xalpha, double DETECTOR_PITCH_CM,
                       double DETECTOR_PITCH_CM_AT_ISO_CENT, double DIST_BTWN_SRC_AND_ISOCENT_CM,
                       double DETECTOR_PIXEL_NUM, double x_reg, double y_reg, double z_reg, float* Attenuation,
                       float* pdf_pe, int num_mat, int NE, char* outputname) {
    (void)ite; (void)reprojection_float; (void)npf_double; (void)projVolume; (void)reconScale; (void)reconStep;
    (void)startProjNumber; (void)anglesRad; (void)cosTable; (void)sinTable; (void)XcenterShift; (void)YcenterShift;
    (void)xalpha; (void)DETECTOR_PITCH_CM; (void)DETECTOR_PITCH_CM_AT_ISO_CENT; (void)DIST_BTWN_SRC_AND_ISOCENT_CM;
    (void)DETECTOR_PIXEL_NUM; (void)x_reg; (void)y_reg; (void)z_reg; (void)reconSize; (void)reconSlices; (void)num_mat;
    if (!g_proj) {
        fprintf(stderr, "reprojectionOnGPU: initializeGPU was not called\n");
        exit(1);
    }
    // The reference computes angles [0, endProjNumber) (grid z, P/reprojectionOnGPU.cu:976; startProjNumber is
    // not used) and writes them all.  The buffer holds exactly those angles, and the library computes exactly
    // those; endProjNumber beyond initializeGPU's totalProjNumber (or below 1) is an error.  The image size
    // is the one initializeGPU fixed; a different one here would not match the projector's buffers.
    if (endProjNumber < 1) {
        fprintf(stderr, "reprojectionOnGPU: endProjNumber %d must be positive\n", endProjNumber);
        exit(1);
    }
    const size_t n = (size_t)endProjNumber * g_width * g_height;
    if (projImgWidth != g_width || projImgHeight != g_height) {
        fprintf(stderr, "reprojectionOnGPU: projection size %dx%d differs from initializeGPU's %dx%d\n",
                projImgWidth, projImgHeight, g_width, g_height);
        exit(1);
    }
    std::vector<float> out(n);
    if (vcbct_reprojection_angles(g_proj, reconImageshort, Attenuation, pdf_pe, NE, xfactor, endProjNumber,
                                  out.data()))
        die("reprojectionOnGPU");
    char filename[1024];
    snprintf(filename, sizeof filename, "%s.raw", outputname);
    FILE* fp = fopen(filename, "wb");
    if (!fp) {
        fprintf(stderr, "reprojectionOnGPU: cannot write %s\n", filename);
        exit(1);
    }
    size_t w = fwrite(out.data(), sizeof(float), n, fp);
    fclose(fp);
    if (w != n) {
        fprintf(stderr, "reprojectionOnGPU: short write to %s\n", filename);
        exit(1);
    }
}

// new2DArray layout (P/mallocMD.cpp:46-68): one calloc'd block plus a row-pointer table, so that the
// caller's delete2DArray frees it.
double** prepareConstMap(int projWidth, double DETECTOR_PITCH_CM_AT_ISO_CENT) {
    double* block = (double*)calloc((size_t)projWidth * projWidth, sizeof(double));
    double** rows = (double**)malloc(projWidth * sizeof(double*));
    if (!block || !rows) {
        free(block);
        free(rows);
        fprintf(stderr, "ERROR: not enough memory for prepareConstMap\n");
        return nullptr;
    }
    for (int j = 0; j < projWidth; ++j) rows[j] = block + (size_t)j * projWidth;
    if (vcbct_prepare_const_map(projWidth, DETECTOR_PITCH_CM_AT_ISO_CENT, block)) die("prepareConstMap");
    return rows;
}

void filterSinogram(int projHeight, int projWidth, int sngrmHeight, float*** projVolume, float*** fltdVoxData,
                    double** constMap, double* xshift, double* yshift, double DETECTOR_PITCH_CM_AT_ISO_CENT) {
    (void)constMap;  // the kernel is rebuilt on the device from the same formula
    if (fltdVoxData == nullptr) {
        fprintf(stderr, "ERROR: fltdVoxData is NULL\n");
        return;
    }
    if (vcbct_filter_sinogram_host(projHeight, projWidth, sngrmHeight, &projVolume[0][0][0], &fltdVoxData[0][0][0],
                                   xshift, yshift, DETECTOR_PITCH_CM_AT_ISO_CENT))
        die("filterSinogram");
}

void reconstructImage(int reconSize, int reconSlices, int zstart, double reconStep, double reconScale,
                      double*** reconVoxData, char** maskData, int projHeight, int sngrmWidth, int sngrmHeight,
                      float*** sngrmVoxData, int xfactor, double* anglesRad, double* cosTable, double* sinTable,
                      double* xshift, double DETECTOR_PITCH_CM_AT_ISO_CENT) {
    (void)maskData; (void)cosTable; (void)sinTable;  // the (R-1)/2 circle mask and the trig are recomputed
    if (reconVoxData == nullptr) {
        fprintf(stderr, "ERROR: reconVoxData is NULL\n");
        return;
    }
    if (vcbct_reconstruct_image_host(reconSize, reconSlices, zstart, reconStep, reconScale, &reconVoxData[0][0][0],
                                     projHeight, sngrmWidth, sngrmHeight, &sngrmVoxData[0][0][0], xfactor, anglesRad,
                                     xshift, DETECTOR_PITCH_CM_AT_ISO_CENT))
        die("reconstructImage");
}
