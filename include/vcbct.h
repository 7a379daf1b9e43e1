/*
 * vcbct — MI355X-native virtual-CBCT chain: material-based polychromatic forward projector (MBFPA)
 * and FDK reconstruction, behind a plain C ABI (extern "C", plain pointers and sizes).
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference
 * repository; P/ = Release_CBCT_Projection_CM, F/ = Release_recon_from_forward_pro_float_CM).
 *
 * Conventions
 *  - Every function returns 0 on success and a negative VCBCT_E* code on failure; the message of the
 *    last failure on the calling thread is vcbct_last_error().  Nothing calls exit().
 *  - "dev_" pointers are device (HBM) pointers; "host_" pointers are host memory.  `stream` is a
 *    hipStream_t passed as void* (NULL = the null stream).  Device entry points are asynchronous on
 *    that stream; host entry points synchronise before returning.
 *  - Array layouts are the reference's: volumes [z][mat][y][x] uint16 (g/cc x 1000, material order
 *    H, C, N, O, P, Ca), projections [angle][row][col] float32 with the reference's row flip,
 *    reconstructions [z][y][x].
 */
#ifndef VCBCT_H
#define VCBCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCBCT_OK 0
#define VCBCT_EINVAL -1   /* bad argument or unsupported size */
#define VCBCT_EHIP -2     /* HIP runtime error */
#define VCBCT_ENOMEM -3   /* device or host allocation failed */
#define VCBCT_EIO -4      /* file missing, unreadable or malformed */
#define VCBCT_ESTATE -5   /* call out of order (e.g. forward before a volume was set) */

#define VCBCT_MAX_MAT 6
#define VCBCT_MAX_NE 1024

const char* vcbct_last_error(void);
/* Library version string and the offload target it was built for. */
const char* vcbct_version(void);

/* ------------------------------------------------------------------------------------------------
 * Forward projector.
 *
 * The argument block of initializeGPU (P/reprojectionOnGPU.h:2-8 and P/reprojectionOnGPU.cu:720-757),
 * field for field.
 */
typedef struct vcbct_geometry {
    int recon_size;          /* reconSize N (volume is N x N in x, y) */
    double recon_scale;      /* reconScale, x/y voxel pitch [cm] */
    int recon_slices;        /* reconSlices (the z_required_range slices actually loaded) */
    double recon_step;       /* reconStep, z voxel pitch [cm] */
    int proj_height;         /* projImgHeight */
    int proj_width;          /* projImgWidth */
    int n_angles;            /* totalProjNumber */
    double x_reg, y_reg, z_reg;
    double det_pitch_cm;     /* DETECTOR_PITCH_CM */
    double det_pitch_iso_cm; /* DETECTOR_PITCH_CM_AT_ISO_CENT */
    double sad_cm;           /* DIST_BTWN_SRC_AND_ISOCENT_CM */
    double det_pixel_num;    /* DETECTOR_PIXEL_NUM */
    int num_material;        /* 1..6 */
} vcbct_geometry;

typedef struct vcbct_projector vcbct_projector;

/* Replaces initializeGPU (P/reprojectionOnGPU.cu:720-825): geometry and the per-angle tables
 * (cos, sin, X/Y centre shift from prepareTrigonometricTable, P/misc.cpp:9-37) move to the device.
 * Unlike the reference there is no 360-angle limit and no hard-coded device: `device` selects it. */
int vcbct_projector_create(vcbct_projector** out, const vcbct_geometry* g, const double* cos_table,
                           const double* sin_table, const double* xcenter_shift, const double* ycenter_shift,
                           int device);
/* Replaces terminateGPU (P/reprojectionOnGPU.cu:827-834). */
void vcbct_projector_destroy(vcbct_projector* p);

/* Volume upload, the first step of reprojectionOnGPU (P/reprojectionOnGPU.cu:867).  The volume is the
 * reference's uint16 [z][mat][y][x]; it is re-laid out on the device to [y][x][z][8 x u16] (one 16-byte
 * voxel holds all materials).  host_vol may be a host pointer (is_device = 0) or a device pointer. */
int vcbct_projector_set_volume_u16(vcbct_projector* p, const uint16_t* vol, int is_device, void* stream);
/* FP32 elemental maps, device pointer, layout [mat][z][y][x] (the differentiable path's maps). */
int vcbct_projector_set_volume_f32(vcbct_projector* p, const float* dev_maps, void* stream);

/* Spectrum upload (P/reprojectionOnGPU.cu:873-882): atten [mat][NE] (cm^2/g, as written by
 * attenuation_coefficient), pdf [NE]. */
int vcbct_projector_set_spectrum(vcbct_projector* p, const float* host_atten, const float* host_pdf, int NE);

/* The compute of reprojectionOnGPU (P/reprojectionOnGPU.cu:1004-1056), energy-collapsed: one traversal
 * per ray yields the material line integrals, then the in-kernel spectral sum
 *   I = sum_e float(pdf_e * exp(-sum_m mu_me L_m / xfactor) * leng)   (FP32 accumulation, e ascending)
 * is written to dev_out [n_angles][H][W] (row H-1-j for kernel row j, as the reference).
 * Angles [a0, a1) are computed (a1 = -1: all). */
int vcbct_projector_forward(vcbct_projector* p, float* dev_out, double xfactor, int a0, int a1, void* stream);

/* Six line integrals per ray, dev_L [n_angles][H][W][num_material] FP32 (for checks and tools). */
int vcbct_projector_line_integrals(vcbct_projector* p, float* dev_L, void* stream);

/* Number of charged segments (in range, inside the inscribed circle) per ray, dev_counts [n_angles][H][W]
 * FP32: the work unit of the roofline model (12 bytes of voxel data per segment). */
int vcbct_projector_segment_counts(vcbct_projector* p, float* dev_counts, void* stream);

/* Tile shape of the intensity and log-projection launches (vcbct_projector_forward, _logproj, _logproj_save):
 * a workgroup is row_groups groups of 64 detector rows by cols adjacent columns of one angle, one wave each
 * (row_groups * cols = 2, 4 or 8); order 0 walks the angles inside each row block, 1 the row blocks inside
 * each angle; split 0 divides the columns among the 8 XCDs, 1 the angles.  row_groups = 0 (the default)
 * hands the choice to the autotuner, which times the candidates on the first launch of each geometry outside
 * stream capture (VCBCT_NO_AUTOTUNE=1: 1 x 4, order 0, split 0) and records them in vcbct_autotune_log.  The
 * shape changes only the schedule: every ray's value is the same under all of them.  Replaces the reference's
 * hand-set launch shapes (P/reprojectionOnGPU.cu:900-983). */
int vcbct_projector_set_tile(vcbct_projector* p, int row_groups, int cols, int order, int split);
/* The same for the traversal kernels of the adjoint (vcbct_projector_adjoint, _adjoint_upstream,
 * _logproj_backward_saved, _multi_loss_grad: the coefficient kernel and the scatter), whose workgroups are
 * 4 waves (row_groups * cols = 4).  row_groups = 0 (the default): autotuned per geometry, the scatter on the
 * caller's own coefficients with the gradient cleared again after the timing launches.  The scatter's
 * floating-point additions are unordered under every shape. */
int vcbct_projector_set_adjoint_tile(vcbct_projector* p, int row_groups, int cols, int order, int split);
/* The shape the last intensity / log-projection launch used: shape4 = {row_groups, cols, order, split}
 * (zeros before the first). */
int vcbct_projector_get_tile(const vcbct_projector* p, int* shape4);

/* Host drop-in for reprojectionOnGPU (P/reprojectionOnGPU.cu:838-1100, declared at
 * P/reprojectionOnGPU.h:12-19): uploads host_vol/atten/pdf, runs the forward projection and copies the
 * result to host_out [n_angles][H][W]. */
int vcbct_reprojection(vcbct_projector* p, const uint16_t* host_vol, const float* host_atten,
                       const float* host_pdf, int NE, double xfactor, float* host_out);
/* The same for the first n_out_angles angles only, host_out [n_out_angles][H][W]: the reference's launch
 * covers angles [0, endProjNumber) (P/reprojectionOnGPU.cu:976) and writes exactly those.  n_out_angles
 * must lie in [1, n_angles] (VCBCT_EINVAL otherwise). */
int vcbct_reprojection_angles(vcbct_projector* p, const uint16_t* host_vol, const float* host_atten,
                              const float* host_pdf, int NE, double xfactor, int n_out_angles, float* host_out);

/* ------------------------------------------------------------------------------------------------
 * Organ-label phantom -> six elemental maps, the step before the projector
 * (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp; R/ below).
 *
 * Replaces the table part of R/ :104-201 and the per-voxel lookup rule of :212-226.  media_csv rows are
 * "ID,H,C,N,O,Na,Mg,P,S,Cl,K,Ca,Fe,I" (weight %, R/AM_media.csv), organs_csv rows "organ,media,density"
 * (g/cc, R/AM_organs.csv).  host_lut [256][8] receives, for every 8-bit organ label, the six map values
 * H, C, N, O, P, Ca = (short)(weight% x density x 10) in FP32 as the reference computes them (the
 * organ's media is row media-1; the last organ row with that label wins); labels with no organ row map
 * to 0; entries 6 and 7 are 0.  With randomize != 0 the H, C, N weights, and P and Ca where above 0.09 %,
 * of every medium except ID 53 are redrawn from N(w, 0.1 w) and clamped at 0 (:117-150), by glibc's
 * rand() sequence after srand(seed) (seed 1 is the reference's unseeded rand(); a private state, the
 * process's rand() is untouched).  O then takes 100 - H - C - N - P - Ca, renormalised when negative
 * (:151-162).  The reference's density draw (:178-198) is not made: its result is never used.
 * host_media_w (optional, [media_cap][6]) receives the H, C, N, O, P, Ca weights used.
 * Returns the number of media rows, or a negative error. */
int vcbct_material_table(const char* media_csv, const char* organs_csv, int randomize, unsigned int seed,
                         uint16_t* host_lut, float* host_media_w, int media_cap);

/* Replaces the voxel loop of R/ :212-226 and the slice writer's layout (:242-274): labels u8
 * [nz][nrow][ncol] -> maps u16 [nz][6][nrow][ncol] (slice iz of dev_maps is the reference's per-slice
 * material file MD_<ncol>_<nrow>_<iz>.raw).  dev_lut [256][8] as vcbct_material_table writes it, 16-byte
 * aligned (it is read one 16-byte entry at a time). */
int vcbct_labels_to_maps(const uint8_t* dev_labels, int ncol, int nrow, int nz, const uint16_t* dev_lut,
                         uint16_t* dev_maps, void* stream);

/* The same lookup written straight into the projector's device volume (no reference counterpart: the
 * reference leaves the step from the converter's ncol x nrow slices to the projector's N x N slices to
 * the user).  Projector voxel (x, y, z) takes the label at column x + col0, row y + row0, slice
 * z + slice0 of dev_labels [nz][nrow][ncol]; outside that grid it takes label 0.  Materials beyond the
 * projector's num_material are dropped.  Equivalent to vcbct_labels_to_maps + cropping or padding to
 * N x N + vcbct_projector_set_volume_u16, without the intermediate maps. */
int vcbct_projector_set_volume_labels(vcbct_projector* p, const uint8_t* dev_labels, int ncol, int nrow, int nz,
                                      int col0, int row0, int slice0, const uint16_t* dev_lut, void* stream);
/* The projector's u16 device volume back in the reference layout [z][mat][y][x] (no reference
 * counterpart): the N x N slices a label phantom produced, as the projector's slice loader would read them
 * (P/projectionData.cpp:190-215), for writing them out or checking them. */
int vcbct_projector_get_volume_u16(vcbct_projector* p, uint16_t* dev_vol, void* stream);

/* ------------------------------------------------------------------------------------------------
 * Iterative-reconstruction pieces of the projector tree (P/main.cpp includes them but no longer calls them).
 *
 * Replaces lengthproj (P/lengthproj.cpp:9-131): dev_npf [Nz][N][N] FP32 (overwritten) receives, per voxel,
 * the sum over angles [a0, a1) (a1 = -1: all) of (float)((65535 - q) * chord), where chord is the voxel's
 * length along the major axis of the source -> voxel-centre line and q the bilinear sample of dev_proj
 * [n_angles][H][W] u16 where that line meets the detector (the reference's tap order and weights, its
 * x_reg in place of y_reg in the u coordinate, and its registration reset for ite != 0).  Angles whose
 * detector geometry divides by zero (sin = 0) contribute nothing, as on the reference's x86 build. */
int vcbct_projector_lengthproj(vcbct_projector* p, int ite, const uint16_t* dev_proj, int a0, int a1, float* dev_npf,
                               void* stream);
/* Replaces IR_ImageCBCT (P/IR_ImageCBCT.cpp:19-94), one update of a [Nz][N][N] int16 image:
 * out = before (wbefore + (1 - wbefore) npf_re / (npf + lambda)), lambda = weight_tv TV'(before) + 1e-6
 * (the 13-neighbour TV gradient of :9-16, interior voxels only; the prior term is 0 as in the reference),
 * clamped (< 0 -> 0, > 32767 -> 5000) and zeroed where the in-plane radius in cm reaches N/2. */
int vcbct_ir_update(int recon_size, int recon_slices, double recon_scale, const int16_t* dev_before,
                    const float* dev_npf, const float* dev_npf_re, double wbefore, double weight_tv,
                    double weight_prior, float* dev_out, void* stream);

/* ------------------------------------------------------------------------------------------------
 * Differentiable form (no reference counterpart; the reference's scatter U-Net, Unet_scatter.py, is the
 * intended consumer).  Log projection p = -log( sum_e pdf_e exp(-sum_m mu_me L_m/xf) / sum_e pdf_e ).
 */
int vcbct_projector_logproj(vcbct_projector* p, float* dev_p, double xfactor, void* stream);
/* Adjoint: dev_grad [mat][z][y][x] FP32 (overwritten) receives d/drho of
 *   loss = sum_rays w_ray (p - target)^2  with w_ray = 1,
 * by recomputing the per-ray line integrals (none stored) and scattering 2(p-t)*mubar_m/xf * rr along
 * the same traversal (FP32 atomics).  dev_target [n_angles][H][W].  dev_loss (optional, 1 double)
 * receives the loss.  Uses the volume set with vcbct_projector_set_volume_f32. */
int vcbct_projector_adjoint(vcbct_projector* p, const float* dev_target, float* dev_grad, double* dev_loss,
                            double xfactor, void* stream);
/* Backward of logproj for an arbitrary upstream gradient dev_gp = dL/dp [n_angles][H][W]:
 * dev_grad [mat][z][y][x] (overwritten) = sum_rays gp * mubar_m/xf * rr. */
int vcbct_projector_logproj_backward(vcbct_projector* p, const float* dev_gp, float* dev_grad, double xfactor,
                                     void* stream);
/* The log projection that also keeps its line integrals per ray (dev_L: FP32 [nmat][n_angles*H*W], the rays
 * in the projection layout; the six maps' integrals, or in the first vcbct_projector_forward_channels rows the
 * spectral channels' when the forward traverses channels), and the backward pass from them: the coefficients
 * are formed per ray from dev_L (the same FP32 sums the backward's own traversal would form, so the same
 * values) and only the scatter traverses the rays again.  No reference counterpart; the autograd path uses this
 * pair (at most 24 B per ray kept instead of a second traversal; the per-energy terms are never stored).  The
 * backward does not read the volume; the spectrum must be the one the log projection used. */
int vcbct_projector_logproj_save(vcbct_projector* p, float* dev_p, float* dev_L, double xfactor, void* stream);
/* Several spectra over the same rays, for a multi-spectrum (dual-energy) loss (no reference counterpart):
 * nspec (1..4) attenuation tables and pdfs concatenated, host_atten [spec][mat][NE[spec]] and host_pdf
 * [spec][NE[spec]], as vcbct_projector_set_spectrum takes one.  Their joint spectral basis (of the tables side
 * by side) gives the channels the gradient is scattered in.  Independent of the single spectrum of
 * vcbct_projector_set_spectrum. */
int vcbct_projector_set_spectra(vcbct_projector* p, int nspec, const float* host_atten, const float* host_pdf,
                                const int* NE);
/* Channels of the multi-spectrum loss's scatter (num_material, or the joint basis's rank). */
int vcbct_projector_multi_channels(const vcbct_projector* p);
/* loss = sum_s sum_rays (p_s - t_s)^2, p_s the log projection of the FP32 maps (vcbct_projector_set_volume_f32)
 * under spectrum s, and dev_grad [mat][z][y][x] (overwritten) its gradient: ONE traversal per ray for all the
 * spectra (their line integrals are the same), the per-ray coefficients summed over the spectra, ONE scatter.
 * dev_targets [spec][n_angles][H][W] FP32; dev_loss (optional) one double.  Equal to the sum over the spectra of
 * vcbct_projector_adjoint with each spectrum, at half the traversals.  Allocates on its first call only. */
int vcbct_projector_multi_loss_grad(vcbct_projector* p, const float* dev_targets, float* dev_grad, double* dev_loss,
                                    double xfactor, void* stream);
/* One Adam step on n FP32 values (no reference counterpart; the decomposition's optimiser on the device):
 * dev_step (one int on the device) is incremented first, then m <- b1 m + (1 - b1) g, v <- b2 v + (1 - b2) g^2,
 * x <- x - lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps), and x <- max(x, lower) unless lower is -inf.
 * Nothing allocated, nothing synchronised: with vcbct_projector_multi_loss_grad an iteration can be captured
 * in a hipGraph and replayed (the step count lives on the device). */
int vcbct_adam_step(size_t n, float* dev_x, const float* dev_grad, float* dev_m, float* dev_v, int* dev_step,
                    double lr, double beta1, double beta2, double eps, double lower, void* stream);

/* Number of spectral channels the forward traverses (no reference counterpart), 0 when it traverses the maps.
 * With at most 4 channels (vcbct_projector_adjoint_channels) and a basis that reproduces every energy column of
 * the table to 1e-6 of its largest entry, the intensity and log projections and the adjoint's coefficient
 * kernels gather the channel volume c_k = sum_m B_mk rho_m (2 or 4 FP32 per voxel, built on the device from the
 * volume when first needed after a volume or spectrum is set) in place of the six maps, and sum the spectrum over
 * the channel table B^T mu; without a basis (num_material <= 4) the channels are the maps (from a u16 volume
 * every sum is then the u16 path's, bit for bit).
 * vcbct_projector_logproj_save then keeps the channels' line integrals (this many rows of dev_L).  Requires a
 * spectrum (VCBCT_ESTATE otherwise). */
int vcbct_projector_forward_channels(const vcbct_projector* p);
/* Number of channels the adjoint scatters (no reference counterpart): num_material, or r < num_material when
 * the spectrum's attenuation table has a rank-r form mu = B (B^T mu) to 1e-6 of its largest entry (B orthonormal,
 * from the table's eigenvectors).  Attenuation between ~10 and ~150 keV is close to a photoelectric and a
 * Compton basis function, and tables built from just those two terms have r = 2: the gradient is then scattered
 * in r channels (per-ray coefficients with the table B^T mu) and mixed back to the maps per voxel, which cuts
 * the scatter's atomics by num_material / r.  Requires a spectrum (VCBCT_ESTATE otherwise). */
int vcbct_projector_adjoint_channels(const vcbct_projector* p);
int vcbct_projector_logproj_backward_saved(vcbct_projector* p, const float* dev_L, const float* dev_gp,
                                           float* dev_grad, double xfactor, void* stream);

/* ------------------------------------------------------------------------------------------------
 * FDK.
 *
 * Replaces prepareConstMap (F/filterSinogram.cpp:10-65): the W x W filter matrix h(m-n), FP64,
 * host output [W][W]. */
int vcbct_prepare_const_map(int proj_width, double det_pitch_iso_cm, double* host_const_map);

/* Replaces filterSinogram (F/filterSinogram.cpp:67-131): q = -log(p), convolution with the reference's
 * truncated kernel (centre tap coeff, off-centre h(m-k), |m-k| <= W-1) done as a batched FFT
 * convolution, then the cosine weight Cm at the output pixel.  Device [n][H][W] in and out.
 * xshift/yshift are host arrays of n_proj entries. */
/* No counterpart in the reference (its filter is a direct convolution): the FFT length that
 * vcbct_filter_sinogram and the FDK plans use for a panel of proj_width columns, the shorter of the smallest
 * power of two >= 64 and the smallest 5 x 2^m >= 80 holding the 2 proj_width - 1 samples of the linear
 * convolution.  Host only, no device call.  VCBCT_EINVAL for proj_width outside 1..2048. */
int vcbct_filter_fft_length(int proj_width);

int vcbct_filter_sinogram(int proj_height, int proj_width, int n_proj, const float* dev_proj, float* dev_filtered,
                          const double* host_xshift, const double* host_yshift, double det_pitch_iso_cm,
                          void* stream);

/* Replaces reconstructImage (F/reconstructImage.cpp:10-184): voxel-driven cone-beam backprojection
 * with the reference's interpolation rule, angle weights and edge cases; accumulates (+=) into
 * dev_recon [Z][R][R] FP32 (inside the (R-1)/2 circle mask).  Angles/xshift are host arrays.  One
 * projection (proj_height x proj_width FP32) must stay below 2^31 bytes, as must 4 proj_width below 2^24
 * (VCBCT_EINVAL otherwise; the same holds for every backprojection entry point and FDK plan). */
int vcbct_reconstruct_image(int recon_size, int recon_slices, int zstart_cm, double recon_step,
                            double recon_scale, float* dev_recon, int proj_height, int proj_width, int n_proj,
                            const float* dev_filtered, int xfactor, const double* host_angles_rad,
                            const double* host_xshift, double det_pitch_iso_cm, void* stream);

/* vcbct_reconstruct_image without any allocation (no reference counterpart; for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The angles and x shifts are DEVICE arrays of n_proj doubles here; the per-angle table the backprojector
 * reads (4 doubles per projection) is built by a kernel into a file-scope device array of the library
 * (512 KiB per GPU, held until the process ends), so n_proj is limited to 16384 (VCBCT_EINVAL beyond).
 * Same arithmetic and result as vcbct_reconstruct_image (up to an ulp of the device cos/sin in the table).
 * Asynchronous on `stream`.  Every call on one GPU uses that GPU's single table: calls that overlap on one
 * GPU (other streams or threads) race and must be serialised by the caller. */
int vcbct_reconstruct_image_noalloc(int recon_size, int recon_slices, int zstart_cm, double recon_step,
                                    double recon_scale, float* dev_recon, int proj_height, int proj_width, int n_proj,
                                    const float* dev_filtered, int xfactor, const double* dev_angles_rad,
                                    const double* dev_xshift, double det_pitch_iso_cm, void* stream);

/* Host-array forms of the two FDK calls for drop-in callers (F/main.cpp:133-137, 170-175): inputs are the
 * contiguous new3DArray blocks; reconstruct_image_host accumulates (+=) into the caller's FP64 volume. */
int vcbct_filter_sinogram_host(int proj_height, int proj_width, int n_proj, const float* host_proj,
                               float* host_filtered, const double* host_xshift, const double* host_yshift,
                               double det_pitch_iso_cm);
int vcbct_reconstruct_image_host(int recon_size, int recon_slices, int zstart_cm, double recon_step,
                                 double recon_scale, double* host_recon, int proj_height, int proj_width, int n_proj,
                                 const float* host_filtered, int xfactor, const double* host_angles_rad,
                                 const double* host_xshift, double det_pitch_iso_cm);

/* FDK plan (no reference counterpart: the reference recomputes its tables per call): the per-projection
 * tables and the filter's frequency response are built once; execution allocates nothing and does not
 * synchronise, so it can be captured in a hipGraph.  Creating a plan also picks the backprojector's
 * slices-per-thread by timing 2, 4 and 8 (autotuner; VCBCT_NO_AUTOTUNE=1 keeps 4). */
typedef struct vcbct_fdk_plan vcbct_fdk_plan;
int vcbct_fdk_plan_create(vcbct_fdk_plan** out, int proj_height, int proj_width, int n_proj, int recon_size,
                          int recon_slices, int zstart_cm, double recon_step, double recon_scale, int xfactor,
                          const double* host_angles_rad, const double* host_xshift, const double* host_yshift,
                          double det_pitch_iso_cm, int device);
void vcbct_fdk_plan_destroy(vcbct_fdk_plan* p);
int vcbct_fdk_plan_filter(vcbct_fdk_plan* p, const float* dev_proj, float* dev_filtered, void* stream);
int vcbct_fdk_plan_backproject(vcbct_fdk_plan* p, const float* dev_filtered, float* dev_recon, void* stream);
int vcbct_fdk_plan_execute(vcbct_fdk_plan* p, const float* dev_proj, float* dev_filtered, float* dev_recon,
                           void* stream);
int vcbct_fdk_plan_slices_per_thread(const vcbct_fdk_plan* p);
/* 1 when the plan's backprojector stages the projection footprint of each block in LDS (picked with the
 * slices per thread by the autotuner; VCBCT_BP_LDS=0 / 1 restricts it), 0 otherwise.  Both forms give the
 * same result bit for bit. */
int vcbct_fdk_plan_bp_staged(const vcbct_fdk_plan* p);
/* Half-fan FDK plan (no reference counterpart; the weighting of vcbct_fdk_offset_weight below, inside the
 * plan): the plan runs on the padded panel (its width: vcbct_fdk_plan_width), and its filter reads the caller's
 * [n][H][W] projections, weighting and padding them as it loads them (2w (-log p) per real column, from a table
 * built at creation; 0 in the padding), so dev_filtered is [n][H][vcbct_fdk_plan_width(p)] and equals the
 * separate weighting + filter to FP32 rounding.  Execution allocates nothing, writes no plan state and does not
 * synchronise (capturable).  Errors as vcbct_fdk_offset_weight. */
int vcbct_fdk_plan_create_half_fan(vcbct_fdk_plan** out, int proj_height, int proj_width, int n_proj,
                                   int recon_size, int recon_slices, int zstart_cm, double recon_step,
                                   double recon_scale, int xfactor, const double* host_angles_rad,
                                   const double* host_xshift, const double* host_yshift, double det_pitch_iso_cm,
                                   double taper_px, int device);
/* Columns of the panel the plan filters and backprojects (proj_width, or the padded width of a half-fan plan). */
int vcbct_fdk_plan_width(const vcbct_fdk_plan* p);
/* The filter's window parameter (no reference counterpart: the reference fixes ALPHA = 0.5 in
 * F/physParams.h:39-41, used by prepareConstMap, F/filterSinogram.cpp:10-65, and its centre coefficient,
 * :78-80).  alpha in [0, 1]; 0.5 (the default) is the reference's kernel, 1 the band-limited ramp without a
 * window (Ram-Lak, cutoff 1.6 x Nyquist as the reference's).  Rebuilds the plan's frequency response (set-up
 * call: allocates and synchronises once per new (width, pitch, alpha)). */
int vcbct_fdk_plan_set_window(vcbct_fdk_plan* p, double alpha);

/* Panel-offset class of an FDK angle list (F/projectionDataMVkV.cpp:277-281): "S" when the mean panel x
 * shift of the used projections is below 6, "L" above 15, else "M".  out: at least 2 chars. */
int vcbct_fdk_size_class(int n_proj, const double* host_xshift, char* out);
/* prepareAngles (F/misc.cpp:87-180): dev_mask [n_proj][R][R] u8 = 1 where angle k lies in voxel (i, j)'s
 * window (th1 + 170 deg, th1 + 10 deg) through +-180 deg, th1 the voxel centre's angle from +y measured
 * clockwise; all 1 for size class "S".  The reference builds this mask in its main program and never
 * applies it.  Synchronises the stream (set-up call). */
int vcbct_fdk_angle_mask(int n_proj, const double* host_angles_rad, int recon_size, double recon_scale,
                         const char* size_class, uint8_t* dev_mask, void* stream);
/* Offset-detector (half-fan) weighting, the step that the panel-offset class and angle mask above are computed
 * for (F/main.cpp:57) and that the reference never takes: no reference counterpart.  Each column of
 * projection k gets w(t) = sin^2(pi/4 (1 + t/tau)) on |t| < tau, 0 below, 1 above, t its distance in pixels
 * from the central ray (column W/2 + xshift[k]/(0.05*800/W) - 0.5, the backprojector's column rule,
 * F/reconstructImage.cpp:83) counted positive towards the long side, tau = min(taper_px, W/2 - |shift|)
 * (taper_px <= 0: the whole overlap), so that conjugate rays sum to one (G. Wang, Med. Phys. 29 (2002)
 * 1634).  The weight acts on the filter's line integral q = -log p, doubled to keep a full 360-degree scan's
 * normalisation (p^(2w)).  The panel is padded on the short side with P = round(2 max|shift|) columns of
 * p = 1, so that the filtered projection's tails reach the voxels only the long side sees: dev_out is
 * [n][H][W + P] (out_width = vcbct_fdk_half_fan_width()), and host_xshift_out / host_yshift_out (n each) put
 * the central ray and the filter's row centre where they were.  Filter and backproject dev_out with these
 * shifts, width W + P and det_pitch_iso_cm * (W + P) / W (the same pixel size).  dev_out may alias dev_proj
 * only when P = 0.  Errors: shifts of both signs, a shift that leaves no overlap, W + P > 4096,
 * n_proj > 65535.  Synchronises the stream. */
int vcbct_fdk_half_fan_width(int proj_width, int n_proj, const double* host_xshift);
int vcbct_fdk_offset_weight(int proj_height, int proj_width, int n_proj, const float* dev_proj,
                            const double* host_xshift, const double* host_yshift, double taper_px, int out_width,
                            float* dev_out, double* host_xshift_out, double* host_yshift_out, void* stream);
/* Host-array form of vcbct_fdk_offset_weight for the host programs (host_out: [n][H][out_width]). */
int vcbct_fdk_offset_weight_host(int proj_height, int proj_width, int n_proj, const float* host_proj,
                                 const double* host_xshift, const double* host_yshift, double taper_px, int out_width,
                                 float* host_out, double* host_xshift_out, double* host_yshift_out);

/* ------------------------------------------------------------------------------------------------
 * Detector model (no reference counterpart: the stage between the projector and FDK that the reference
 * leaves to other programs — its projection file names, F/txtfile, and the scatter U-Net,
 * Unet_scatter.py:349-423, on either side).  For each of the n pixels (any flat layout, e.g. [n][H][W]):
 * lambda = I0 * (primary[i] + scatter[i]) quanta (dev_scatter may be NULL), k ~ Poisson(lambda) from
 * Philox4x32-10 keyed by (seed, i) — inversion below lambda = 10, Hoermann's PTRS rejection above — and
 * dev_out[i] = k (raw != 0) or max(k, 1) / I0 (normalised intensity for vcbct_filter_sinogram).  The draw
 * of pixel i depends only on (seed, i, lambda).  dev_out may alias dev_primary.  Errors: I0 not positive. */
int vcbct_detector_counts(size_t n, const float* dev_primary, const float* dev_scatter, double I0, uint64_t seed,
                          int raw, float* dev_out, void* stream);
/* Host-array form of vcbct_detector_counts (host_scatter may be NULL; host_out may alias host_primary). */
int vcbct_detector_counts_host(size_t n, const float* host_primary, const float* host_scatter, double I0,
                               uint64_t seed, int raw, float* host_out);

/* ------------------------------------------------------------------------------------------------
 * Batched launch (no reference counterpart): many forward problems — volumes of different slice counts,
 * angle counts and spectra, each set up as a vcbct_projector — in ONE persistent kernel launch with
 * per-XCD work queues, optionally followed by per-problem FDK plans; the whole sequence can be captured
 * in a hipGraph and replayed.  The persistent grid size is autotuned on the first run. */
typedef struct vcbct_batch vcbct_batch;
int vcbct_batch_create(vcbct_batch** out, int device);
void vcbct_batch_destroy(vcbct_batch* b);
int vcbct_batch_add_forward(vcbct_batch* b, vcbct_projector* p, float* dev_out, double xfactor);
/* recon_bytes: size of dev_recon, which each run zeroes before backprojecting into it */
int vcbct_batch_add_fdk(vcbct_batch* b, vcbct_fdk_plan* plan, const float* dev_proj, float* dev_filtered,
                        float* dev_recon, long long recon_bytes);
int vcbct_batch_run(vcbct_batch* b, void* stream);
int vcbct_batch_capture_graph(vcbct_batch* b);
int vcbct_batch_launch_graph(vcbct_batch* b, void* stream);
/* Grid the autotuner chose for the batch's forward problems: the persistent grid size in blocks, 0 for the
 * flat grid (every problem's blocks laid end to end, no work queue), -2 for the sequence form (one plain
 * forward launch per problem, each with the tile shape its projector had fixed when it was added, else its
 * geometry's autotuned one; replayed together from the captured graph); -1 for NULL. */
int vcbct_batch_blocks(const vcbct_batch* b);
/* Fixes the forward grid instead of autotuning it: 0 = flat grid, n > 0 = persistent grid of n blocks
 * (rounded up to a multiple of 8), -2 = sequence form, -1 = autotune (default).  No reference counterpart
 * (tuning control). */
int vcbct_batch_set_grid(vcbct_batch* b, int blocks);
/* Fixes the tile shape instead of autotuning it: the flat grid's workgroup of row_groups x cols = 4 waves
 * and its order, and for either grid the split (0: columns, 1: angles divided among the 8 XCDs), as
 * vcbct_projector_set_tile describes them; row_groups = 0: autotune (default; the tuner searches grid and
 * split together, then the flat grid's block shape and order).  No reference counterpart (tuning control). */
int vcbct_batch_set_tile(vcbct_batch* b, int row_groups, int cols, int order, int split);
/* The shape of the flat or persistent grid in use after the first run: shape4 = {row_groups, cols, order,
 * split} (the sequence form's launches each use their projector's fixed or their geometry's autotuned shape
 * instead). */
int vcbct_batch_get_tile(const vcbct_batch* b, int* shape4);
/* Text log of the autotuner's decisions (candidates and timings); returns its length. */
int vcbct_autotune_log(char* buf, int capacity);

/* ------------------------------------------------------------------------------------------------
 * Host-side formats kept byte-compatible with the reference.
 *
 * attenuation_coefficient (P/prior_weight_production.cpp:200-523): spectrum CSV + the three xraydb
 * CSVs under xraydb_dir -> atten [mat][NE], pdf [NE].  Returns NE (> 0) or a negative error. */
int vcbct_attenuation_table(const char* spectrum_csv, const char* xraydb_dir, int num_material, int nthin,
                            float* host_atten, float* host_pdf, int capacity);

/* Projector input txt (P/projectionData.cpp:63-102, 15 fields after a 6-line header; format in
 * P/txtfile/IR_waterphantom_36pro.txt).  Geometry constants come from physParams.h (P/physParams.h:47-51). */
typedef struct vcbct_proj_config {
    char spectrum[256];
    int total_proj, proj_width, proj_height;
    char flexmap[256];
    char mat_prefix[256];
    int zic;              /* IC z position (slice number) */
    char outputname[256];
    double x_scale, y_scale, z_scale;
    int recon_size, recon_slices, mat_num;
    char mat_type[16];
    double xfactor;
    double x_reg, y_reg, z_reg;
    double det_pitch_cm, det_pitch_iso_cm, sad_cm, det_pixel_num, sdd_iso_cm;
} vcbct_proj_config;
int vcbct_read_proj_config(const char* txt_path, vcbct_proj_config* out);
/* Flexmap rows "angle_deg xshift yshift class" -> radians (P/projectionData.cpp:152-171).  Returns the
 * number of rows read (at most capacity). */
int vcbct_read_flexmap(const char* path, double* angles_rad, double* xshift, double* yshift, int capacity);
/* Slice window inside the cone (P/projectionData.cpp:173-185): [linemin, linemax). */
int vcbct_proj_slice_window(const vcbct_proj_config* c, int* linemin, int* linemax);
/* Stacks <prefix>_%03d.raw slices [linemin, linemax) into vol [z][mat][y][x] (P/projectionData.cpp:190-215).
 * missing_policy 0: a missing or short file is an error (VCBCT_EIO).  1: the reference's behaviour — the
 * previous slice's pixels are kept for it (P/projectionData.cpp:201, 246-259) and *n_missing counts it. */
int vcbct_load_material_slices(const char* prefix, int linemin, int linemax, int recon_size, int num_material,
                               int missing_policy, uint16_t* vol, int* n_missing);

/* FDK input txt (F/projectionDataMVkV.cpp:51-97, 18 fields; reconOffset and xfactor are integers). */
typedef struct vcbct_fdk_config {
    char kvmv[16];
    int total_proj, proj_width, proj_height, header_bytes;
    char proj_path[256];
    int num_digits;
    char ext[16];
    char flexmap[256];
    int recon_size;
    char tgt_class[16];
    int recon_offset_cm;
    double recon_step;
    int recon_slices;
    double recon_scale;
    int xfactor;
    char output[256];
    double det_pitch_iso_cm;
} vcbct_fdk_config;
int vcbct_read_fdk_config(const char* txt_path, vcbct_fdk_config* out);
/* FDK angle list (F/projectionDataMVkV.cpp:140-274): class filter, reversal when angle[row 1] - angle[row 30]
 * < 0.  Returns the number of projections used (= total, the txt count) or a negative error. */
int vcbct_read_fdk_angles(const char* path, const char* tgt_class, int total, double* angles_rad, double* xshift,
                          double* yshift);

/* prepareTrigonometricTable (P/misc.cpp:9-37). */
int vcbct_trig_table(int proj_width, int proj_height, int n, const double* angles_rad, const double* xshift,
                     const double* yshift, double det_pitch_cm, double det_pixel_num, double* cos_t,
                     double* sin_t, double* xcenter_shift, double* ycenter_shift);

#ifdef __cplusplus
}
#endif
#endif
