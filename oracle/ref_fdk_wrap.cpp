/*
 * extern "C" entry points over the reference's own FDK CPU code, compiled from the reference sources
 * where they lie (recipe: oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used to pin the oracle and as the
 * timed CPU baseline.  The functions called are the reference's, unmodified:
 *   prepareConstMap / filterSinogram   F/filterSinogram.cpp:10-65, 67-131
 *   reconstructImage                   F/reconstructImage.cpp:10-184
 *   prepareMaskData, prepareTrigonometricTable  F/misc.cpp:9-37, 181-200
 *   prepareAngles                      F/misc.cpp:87-180
 *   loadData                           F/projectionDataMVkV.cpp:51-321
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <chrono>

#include "physParams.h"
#include "mallocMD.h"
#include "misc.h"
#include "projectionDataMVkV.h"
#include "filterSinogram.h"
#include "reconstructImage.h"

extern "C" {

void ref_const_map(int W, double pitch, double* out) {
    double** cm = prepareConstMap(W, pitch);
    memcpy(out, cm[0], (size_t)W * W * sizeof(double));
    delete2DArray((void**)cm, W, W, UNIT_FLOAT64);
}

/* filterSinogram on flat [n_ang][H][W] buffers; returns the wall seconds of the filterSinogram call. */
double ref_filter_sinogram(int H, int W, int n_ang, const float* proj, float* out, const double* xs, const double* ys,
                           double pitch) {
    float*** P = (float***)new3DArray(W, H, n_ang, UNIT_FLOAT32);
    float*** O = (float***)new3DArray(W, H, n_ang, UNIT_FLOAT32);
    memcpy(get1DArrayOf3DArray((void***)P, UNIT_FLOAT32), proj, (size_t)W * H * n_ang * sizeof(float));
    clearSinogramImage(H, W, n_ang, O);
    double** cm = prepareConstMap(W, pitch);
    auto t0 = std::chrono::steady_clock::now();
    filterSinogram(H, W, n_ang, P, O, cm, (double*)xs, (double*)ys, pitch);
    auto t1 = std::chrono::steady_clock::now();
    memcpy(out, get1DArrayOf3DArray((void***)O, UNIT_FLOAT32), (size_t)W * H * n_ang * sizeof(float));
    delete2DArray((void**)cm, W, W, UNIT_FLOAT64);
    delete3DArray((void***)P, W, H, n_ang, UNIT_FLOAT32);
    delete3DArray((void***)O, W, H, n_ang, UNIT_FLOAT32);
    return std::chrono::duration<double>(t1 - t0).count();
}

/* reconstructImage into a zeroed double volume [Z][R][R]; returns wall seconds of the call. */
double ref_reconstruct_image(int R, int Z, int zstart, double step, double scale, double* recon, int H, int W,
                             int n_ang, const float* sino, int xfactor, const double* ang, const double* xs,
                             double pitch) {
    double*** V = (double***)new3DArray(R, R, Z, UNIT_FLOAT64);
    short*** S = (short***)new3DArray(R, R, Z, UNIT_SINT16);
    float*** F = (float***)new3DArray(W, H, n_ang, UNIT_FLOAT32);
    memcpy(get1DArrayOf3DArray((void***)F, UNIT_FLOAT32), sino, (size_t)W * H * n_ang * sizeof(float));
    clearReconstructedImage(R, Z, n_ang, V, S);
    char** mask = prepareMaskData(R);
    double *c, *s;
    prepareTrigonometricTable(n_ang, (double*)ang, &c, &s);
    auto t0 = std::chrono::steady_clock::now();
    reconstructImage(R, Z, zstart, step, scale, V, mask, H, W, n_ang, F, xfactor, (double*)ang, c, s, (double*)xs,
                     pitch);
    auto t1 = std::chrono::steady_clock::now();
    memcpy(recon, get1DArrayOf3DArray((void***)V, UNIT_FLOAT64), (size_t)R * R * Z * sizeof(double));
    delete1DArray(c); delete1DArray(s);
    delete2DArray((void**)mask, R, R, UNIT_UINT8);
    delete3DArray((void***)V, R, R, Z, UNIT_FLOAT64);
    delete3DArray((void***)S, R, R, Z, UNIT_SINT16);
    delete3DArray((void***)F, W, H, n_ang, UNIT_FLOAT32);
    return std::chrono::duration<double>(t1 - t0).count();
}

void ref_mask(int R, char* out) {
    char** m = prepareMaskData(R);
    memcpy(out, m[0], (size_t)R * R);
    delete2DArray((void**)m, R, R, UNIT_UINT8);
}

/* loadData for the FDK txt, run in workdir.  ints: total, W, H, R, reconOffset, Z, xfactor, used;
 * dbls: step, scale, pitch_iso; strs: 5 x 256 (proj path, ext, flexmap, output, panel-offset class). */
int ref_fdk_load(const char* txt, const char* workdir, int* ints, double* dbls, char* strs, double* ang, double* xs,
                 double* ys, int cap) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(workdir) != 0) return -2;
    PROJECTION_DATA* pd = newProjectionData();
    int ok = loadData((char*)txt, pd);
    int rc = chdir(cwd);
    (void)rc;
    if (ok == 0) { free(pd); return -3; }
    ints[0] = pd->totalProjNumber; ints[1] = pd->projImgWidth; ints[2] = pd->projImgHeight;
    ints[3] = pd->reconSize; ints[4] = pd->reconOffset; ints[5] = pd->reconSlices; ints[6] = pd->xfactor;
    ints[7] = pd->usedProjNumber;
    dbls[0] = pd->reconStep; dbls[1] = pd->reconScale; dbls[2] = pd->DETECTOR_PITCH_CM_AT_ISO_CENT;
    strncpy(strs, pd->projImgFileDirname, 255); strncpy(strs + 256, pd->projImgFileExt, 255);
    strncpy(strs + 512, pd->angleDataFilename, 255); strncpy(strs + 768, pd->reconDataFilename, 255);
    strncpy(strs + 1024, pd->SMLClassString, 255);
    int n = pd->usedProjNumber < cap ? pd->usedProjNumber : cap;
    for (int i = 0; i < n; i++) { ang[i] = pd->anglesRad[i]; xs[i] = pd->xshift[i]; ys[i] = pd->yshift[i]; }
    /* projVolume is never allocated by the vCBCT loader; free the rest by hand */
    free(pd->anglesRad); free(pd->xshift); free(pd->yshift); free(pd);
    return n;
}

/* prepareAngles: mask [n][R][R] (1 = angle inside the voxel's window). */
void ref_prepare_angles(int n, const double* ang, int R, double scale, const char* cls, char* mask) {
    prepareAngles(n, (double*)ang, R, mask, scale, cls);
}

}  // extern "C"
