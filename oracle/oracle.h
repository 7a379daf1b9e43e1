/*
 * FP64 CPU oracle for the virtual-CBCT chain (forward projector + FDK).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this
 * code; only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it, and only as the checker.
 *
 * Every function restates one reference routine (paths relative to the reference
 * repository root; P/ = Release_CBCT_Projection_CM, F/ = Release_recon_from_forward_pro_float_CM):
 *
 *   orc_trig_table_proj      P/misc.cpp:9-37            (prepareTrigonometricTable, projector)
 *   orc_trig_table_fdk       F/misc.cpp:181-200         (prepareTrigonometricTable, FDK)
 *   orc_attenuation_table    P/prior_weight_production.cpp:200-523 (attenuation_coefficient)
 *   orc_forward_compat       P/reprojectionOnGPU.cu:56-690 (reprojectionKernel) + :1004-1044 (host
 *                            energy loop and FP32 accumulation)
 *   orc_line_integrals       the same traversal, six material line integrals per ray (energy collapsed)
 *   orc_forward_collapsed    one traversal per ray, then the spectral sum (new form of the same result)
 *   orc_logproj / orc_adjoint  polychromatic -log(I/I0) and its exact gradient (no reference counterpart)
 *   orc_const_map            F/filterSinogram.cpp:10-65 (prepareConstMap)
 *   orc_filter_sinogram      F/filterSinogram.cpp:67-131 (filterSinogram; FP32 accumulator like the
 *                            reference, or FP64 when acc64 != 0)
 *   orc_mask                 F/misc.cpp:9-37            (prepareMaskData)
 *   orc_reconstruct_image    F/reconstructImage.cpp:10-184 (reconstructImage)
 *   orc_fdk_angles           F/projectionDataMVkV.cpp:140-281 (flexmap read with CW/CCW reversal)
 *   orc_rand_sequence        glibc rand() after srand(seed), the generator R/ :14-24 draws from
 *   orc_material_weights     R/ :104-166 (media table with the weight draws; R/ = ICRPphantoms_2mmslice/
 *                            To_material_distribution_for512_random.cpp)
 *   orc_organ_table          R/ :168-201
 *   orc_labels_to_maps       R/ :212-226 (voxel loop) in the per-slice layout of R/ :257-270
 */
#ifndef VCBCT_ORACLE_H
#define VCBCT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The argument block of initializeGPU (P/reprojectionOnGPU.h:2-8). */
typedef struct orc_geom {
    int N;              /* reconSize (square, x and y) */
    double s;           /* reconScale, voxel pitch in x/y [cm] */
    int Nz;             /* reconSlices (usedZNumber) */
    double t;           /* reconStep, voxel pitch in z [cm] */
    int H;              /* projImgHeight */
    int W;              /* projImgWidth */
    int n_ang;          /* totalProjNumber */
    double x_reg, y_reg, z_reg;
    double det_pitch;       /* DETECTOR_PITCH_CM */
    double det_pitch_iso;   /* DETECTOR_PITCH_CM_AT_ISO_CENT */
    double sad;             /* DIST_BTWN_SRC_AND_ISOCENT_CM */
    double det_pixel_num;   /* DETECTOR_PIXEL_NUM */
    int nmat;
} orc_geom;

void orc_trig_table_proj(int W, int H, int n, const double* angles_rad, const double* xshift,
                         const double* yshift, double det_pitch, double det_pixel_num,
                         double* cos_t, double* sin_t, double* xc, double* yc);
void orc_trig_table_fdk(int n, const double* angles_rad, double* cos_t, double* sin_t);

/* Returns NE (>0) or a negative error code.  atten is [nmat][NE] (cm^2/g), pdf is [NE]. */
int orc_attenuation_table(const char* spectrum_csv, const char* xraydb_dir, int nmat, int nthin,
                          float* atten, float* pdf, int cap);

/* vol_u16 is the reference layout [z][mat][y][x].  out is [ang][H][W] with the reference's row flip.
 * Angles [a0, a1) are computed; out points at angle 0's image. */
int orc_forward_compat(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc,
                       const double* yc, const uint16_t* vol_u16, const float* atten, const float* pdf,
                       int NE, double xfactor, int a0, int a1, float* out);

/* Six (nmat) line integrals sum_seg vol*rr*mask per ray: L is [ang][H][W][nmat] in output-row order.
 * nseg (optional) receives the number of charged (in-range) segments per ray.  leng (optional)
 * receives the cos^3 factor.  vol is u16 (vol_f32 == NULL) or FP32 (vol_u16 == NULL), same layout. */
int orc_line_integrals(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc,
                       const double* yc, const uint16_t* vol_u16, const float* vol_f32, int a0, int a1,
                       double* L, int32_t* nseg, double* leng);

int orc_forward_collapsed(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc,
                          const double* yc, const uint16_t* vol_u16, const float* vol_f32,
                          const float* atten, const float* pdf, int NE, double xfactor, int a0, int a1,
                          float* out);

/* p = -log( sum_e pdf_e exp(-sum_m mu_me L_m / xf) / sum_e pdf_e ) per ray.  FP32 maps [z][mat][y][x]. */
int orc_logproj(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc,
                const double* yc, const float* vol_f32, const float* atten, const float* pdf, int NE,
                double xfactor, double* p);
/* Gradient of sum_rays (p - target)^2 w.r.t. the FP32 maps, layout [z][mat][y][x], FP64. */
int orc_adjoint(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc,
                const double* yc, const float* vol_f32, const float* atten, const float* pdf, int NE,
                double xfactor, const double* target, double* grad, double* loss);

int orc_trace_ray(const orc_geom* g, const double* cos_t, const double* sin_t, const double* xc, const double* yc,
                  int a, int j, int jj, double* out, int cap);
void orc_const_map(int W, double pitch_iso, double* cmap /* [W][W] */);
/* the same with the window parameter alpha (0.5: the reference's; 1: the band-limited ramp) */
void orc_const_map_w(int W, double pitch_iso, double alpha, double* cmap);
void orc_filter_sinogram(int H, int W, int n_ang, const float* proj, float* out, const double* cmap,
                         const double* xshift, const double* yshift, double pitch_iso);
void orc_filter_sinogram64_w(int H, int W, int n_ang, const float* proj, double* out, const double* cmap,
                             const double* xshift, const double* yshift, double pitch_iso, double alpha);
void orc_filter_sinogram64(int H, int W, int n_ang, const float* proj, double* out, const double* cmap,
                           const double* xshift, const double* yshift, double pitch_iso);
void orc_mask(int R, char* mask);
void orc_reconstruct_image(int R, int Z, int zstart, double step, double scale, double* recon,
                           const char* mask, int H, int W, int n_ang, const float* sino, int xfactor,
                           const double* angles, const double* cos_t, const double* sin_t,
                           const double* xshift, double pitch_iso);
/* Reads a flexmap with the FDK loader's class filter and CW/CCW reversal.  Returns the number of
 * used projections written into angles_rad/xshift/yshift (capacity cap), or a negative error. */
int orc_fdk_angles(const char* flexmap, const char* tgt_class, int total, double* angles_rad,
                   double* xshift, double* yshift, int cap);

/* Organ-label phantom path.  The first n values of glibc's rand() after srand(seed). */
int orc_rand_sequence(unsigned seed, int n, int* out);
/* Returns the number of media rows (<= cap) or a negative error; w [cap][6] = H C N O P Ca weight %. */
int orc_material_weights(const char* media_csv, int randomize, unsigned seed, float* w, int cap);
/* Returns the number of organ rows (<= cap) or a negative error. */
int orc_organ_table(const char* organs_csv, int* organ, int* medium, float* dens, int cap);
/* labels u8 [nz][nrow][ncol] -> maps int16 [nz][6][nrow][ncol]. */
int orc_labels_to_maps(const uint8_t* labels, int ncol, int nrow, int nz, const float* w, int nmedia,
                       const int* organ, const int* medium, const float* dens, int norg, int16_t* maps);

#ifdef __cplusplus
}
#endif
#endif
