// Calls the shim's initializeGPU / reprojectionOnGPU with the reference's argument lists
// (P/reprojectionOnGPU.h:2-19), as P/main.cpp:111-125 does, with the angle count and end angle from the command
// line: drive_shim <totalProjNumber> <endProjNumber> <outputname> [width height of the reprojection call].
#include <cstdlib>
#include <vector>

void initializeGPU(int reconSize, double reconScale, int reconSlices, double reconStep, int projImgHeight,
                   int projImgWidth, int totalProjNumber, double* cosTable, double* sinTable, double* XcenterShift,
                   double* YcenterShift, double x_reg, double y_reg, double z_reg, double DETECTOR_PITCH_CM,
                   double DETECTOR_PITCH_CM_AT_ISO_CENT, double DIST_BTWN_SRC_AND_ISOCENT_CM,
                   double DETECTOR_PIXEL_NUM, int num_material);
void terminateGPU();
void reprojectionOnGPU(int ite, int reconSize, int reconSlices, unsigned short* reconImageshort,
                       float* reprojection_float, float* npf_double, float* projVolume, double reconScale,
                       double reconStep, int projImgWidth, int projImgHeight, int startProjNumber, int endProjNumber,
                       double* anglesRad, double* cosTable, double* sinTable, double* XcenterShift,
                       double* YcenterShift, double xfactor, double* xalpha, double DETECTOR_PITCH_CM,
                       double DETECTOR_PITCH_CM_AT_ISO_CENT, double DIST_BTWN_SRC_AND_ISOCENT_CM,
                       double DETECTOR_PIXEL_NUM, double x_reg, double y_reg, double z_reg, float* Attenuation,
                       float* pdf_pe, int num_mat, int NE, char* outputname);

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const int total = atoi(argv[1]), end = atoi(argv[2]);
    const int H = 3, W = 5, N = 4, Z = 2;
    const int wc = argc > 5 ? atoi(argv[4]) : W, hc = argc > 5 ? atoi(argv[5]) : H;
    std::vector<double> tab(total > 0 ? total : 1, 0.0);
    std::vector<unsigned short> vol((size_t)N * N * Z * 6, 0);
    std::vector<float> atten(6, 0.2f), pdf(1, 1.0f);
    initializeGPU(N, 0.5, Z, 0.5, H, W, total, tab.data(), tab.data(), tab.data(), tab.data(), 0, 0, 0, 0.05,
                  0.0325, 100, 800, 6);
    reprojectionOnGPU(0, N, Z, vol.data(), nullptr, nullptr, nullptr, 0.5, 0.5, wc, hc, 0, end, tab.data(),
                      tab.data(), tab.data(), tab.data(), tab.data(), 1000.0, nullptr, 0.05, 0.0325, 100, 800, 0, 0, 0,
                      atten.data(), pdf.data(), 6, 1, argv[3]);
    terminateGPU();
    return 0;
}
