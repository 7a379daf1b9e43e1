/*
 * extern "C" entry points over the reference's own projector CPU code, compiled from the reference
 * sources where they lie (recipe: oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used to pin the oracle
 * and as the timed CPU baseline.  This file holds only call wrappers written for this project; the
 * functions it calls are the reference's, unmodified:
 *   reprojection             P/reprojection.cpp:9-710
 *   attenuation_coefficient  P/prior_weight_production.cpp:200-523
 *   prepareTrigonometricTable P/misc.cpp:9-37
 *   loadData                 P/projectionData.cpp:63-237
 *   lengthproj               P/lengthproj.cpp:9-131
 *   IR_ImageCBCT             P/IR_ImageCBCT.cpp:19-94
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <chrono>
#include <vector>

#include "physParams.h"
#include "mallocMD.h"
#include "misc.h"
#include "projectionData.h"
#include "reprojection.h"
#include "virtual_projection.h"
#include "lengthproj.h"
#include "IR_ImageCBCT.h"

extern "C" {

/* Runs attenuation_coefficient with the working directory set to xraydb_parent (the reference opens
 * attenuation_xraydb/*.csv relative to the working directory).  Returns NE. */
int ref_attenuation_coefficient(const char* spectrum, const char* xraydb_parent, int nmat, int nthin,
                                const char* mat_type, float* atten, float* pdf) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(xraydb_parent) != 0) return -2;
    int NE = -1;
    attenuation_coefficient((char*)spectrum, nmat, atten, pdf, &NE, nthin, (char*)mat_type);
    if (chdir(cwd) != 0) return -3;
    return NE;
}

void ref_trig_table(int W, int H, int n, const double* ang, const double* xs, const double* ys, double pitch,
                    double pixnum, double* c, double* s, double* xc, double* yc) {
    double *c0, *s0, *x0, *y0;
    prepareTrigonometricTable(W, H, n, (double*)ang, (double*)xs, (double*)ys, pitch, pixnum, &c0, &s0, &x0, &y0);
    memcpy(c, c0, n * sizeof(double));
    memcpy(s, s0, n * sizeof(double));
    memcpy(xc, x0, n * sizeof(double));
    memcpy(yc, y0, n * sizeof(double));
    delete1DArray(c0); delete1DArray(s0); delete1DArray(x0); delete1DArray(y0);
}

/* One call of the CPU twin: angle a (the tables are offset so that the reference's angle-0 loop runs
 * angle a), energy bin ke.  out receives [H][W] un-flipped rows (the CPU twin's layout). */
void ref_reprojection_one(int N, int Nz, const unsigned short* vol, double s, double t, int W, int H, int a,
                          const double* ang, const double* c, const double* sn, const double* xc, const double* yc,
                          double xfactor, double pitch, double pitch_iso, double sad, double pixnum, double xr,
                          double yr, double zr, const float* atten, const float* pdf, int nmat, int NE, int ke,
                          float* out) {
    std::vector<float> pv((size_t)W * H, 0.0f);
    double xalpha[1] = {0};
    reprojection(0, N, Nz, (unsigned short*)vol, out, NULL, pv.data(), s, t, W, H, 0, 1, (double*)ang + a,
                 (double*)c + a, (double*)sn + a, (double*)xc + a, (double*)yc + a, xfactor, xalpha, pitch,
                 pitch_iso, sad, pixnum, xr, yr, zr, (float*)atten, (float*)pdf, nmat, NE, ke);
}

/* CPU baseline: n_pairs calls of the CPU twin over (angle, bin) pairs spread over nthreads host threads
 * (the twin itself is single-threaded per call).  Returns wall seconds. */
double ref_reprojection_timed(int N, int Nz, const unsigned short* vol, double s, double t, int W, int H,
                              const double* ang, const double* c, const double* sn, const double* xc,
                              const double* yc, double xfactor, double pitch, double pitch_iso, double sad,
                              double pixnum, double xr, double yr, double zr, const float* atten, const float* pdf,
                              int nmat, int NE, int n_pairs, const int* pair_ang, const int* pair_ke,
                              int nthreads, float* out /* n_pairs*H*W */) {
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int p = 0; p < n_pairs; p++) {
        ref_reprojection_one(N, Nz, vol, s, t, W, H, pair_ang[p], ang, c, sn, xc, yc, xfactor, pitch, pitch_iso,
                             sad, pixnum, xr, yr, zr, atten, pdf, nmat, NE, pair_ke[p], out + (size_t)p * W * H);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

/* loadData for the projector txt, run in the directory holding it.  Returns the fields a host program
 * needs; vol (optional, capacity vol_cap u16) receives matVolume [z][mat][y][x]. */
int ref_proj_load(const char* txt, const char* workdir, int* ints /* 9 */, double* dbls /* 7 */,
                  char* strs /* 4 x 256 */, double* ang, double* xs, double* ys, int cap, unsigned short* vol,
                  long vol_cap) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(workdir) != 0) return -2;
    PROJECTION_DATA* pd = newProjectionData();
    int ok = loadData((char*)txt, pd);
    int rc = chdir(cwd);
    (void)rc;
    if (ok == 0) { deleteProjectionData(pd); return -3; }
    ints[0] = pd->totalProjNumber; ints[1] = pd->projImgWidth; ints[2] = pd->projImgHeight;
    ints[3] = pd->zICposition; ints[4] = pd->reconSize; ints[5] = pd->reconSlices; ints[6] = pd->mat_num;
    ints[7] = pd->usedZNumber; ints[8] = 0;
    dbls[0] = pd->x_scale; dbls[1] = pd->y_scale; dbls[2] = pd->z_scale; dbls[3] = pd->xfactor;
    dbls[4] = pd->x_reg; dbls[5] = pd->y_reg; dbls[6] = pd->z_reg;
    strncpy(strs, pd->SpectrumString, 255); strncpy(strs + 256, pd->angleDataFilename, 255);
    strncpy(strs + 512, pd->matImgFileDirname, 255); strncpy(strs + 768, pd->outputname, 255);
    int n = pd->totalProjNumber < cap ? pd->totalProjNumber : cap;
    for (int i = 0; i < n; i++) { ang[i] = pd->anglesRad[i]; xs[i] = pd->xshift[i]; ys[i] = pd->yshift[i]; }
    long nv = (long)pd->reconSize * pd->reconSize * pd->mat_num * pd->usedZNumber;
    if (vol && nv <= vol_cap) memcpy(vol, pd->matVolume, nv * sizeof(unsigned short));
    deleteProjectionData(pd);
    return n;
}

/* lengthproj over angles [a0, a1) of proj [nAng][H][W] u16; npf [Nz][N][N] (overwritten). */
void ref_lengthproj(int ite, int N, int Nz, const unsigned short* proj, float* npf, double s, double t, int W, int H,
                    int a0, int a1, const double* ang, const double* c, const double* sn, const double* xc,
                    const double* yc, double pitch, double pitch_iso, double sad, double pixnum, double xr, double yr,
                    double zr) {
    double xalpha[1] = {0};
    lengthproj(ite, N, Nz, NULL, NULL, npf, (unsigned short*)proj, s, t, W, H, a0, a1, (double*)ang, (double*)c,
               (double*)sn, (double*)xc, (double*)yc, 1000.0, xalpha, pitch, pitch_iso, sad, pixnum, xr, yr, zr);
}

/* One IR_ImageCBCT update: before [Nz][N][N] short -> out [Nz][N][N] float. */
void ref_ir_update(int N, int Nz, double s, const short* before, float* out, const float* npf, const float* npf_re,
                   double wbefore, double wtv, double wprior) {
    IR_ImageCBCT(0, N, Nz, s, (short*)before, out, (float*)npf, (float*)npf_re, wbefore, wtv, wprior, 1);
}

}  // extern "C"
