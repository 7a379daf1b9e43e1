/*
 * FP64 CPU oracle for the virtual-CBCT chain.  TEST INFRASTRUCTURE ONLY: the product library never
 * links or calls this file; tests, the smoke check and the benchmark's CPU-baseline leg load it as the
 * checker.  Each function cites the reference routine it restates (paths relative to the reference
 * repository; P/ = Release_CBCT_Projection_CM, F/ = Release_recon_from_forward_pro_float_CM).
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math) so that FP64 expressions written in
 * the reference's operation order round exactly as the reference's CPU code does.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPI = 3.141592653589793238462643383279;

/* One charged segment of the compat traversal. */
struct Seg {
    int xmc, ymc, zmc;
    double rr;
    double ph; /* 1, or 0 when the midpoint radius is >= rmax */
};

struct Ray {
    double sx, sy, sz, px, py, pz, leng;
};

/* Ray set-up: P/reprojectionOnGPU.cu:128-150.  j = detector row index of the kernel (before the
 * output flip), jj = column, a = angle. */
Ray make_ray(const orc_geom& g, const double* cos_t, const double* sin_t, const double* xc,
             const double* yc, int a, int j, int jj) {
    Ray r;
    r.sx = -g.sad * sin_t[a] + g.x_reg;
    r.sy = g.sad * cos_t[a] + g.y_reg;
    r.sz = g.z_reg;
    r.pz = -(yc[a] + j) * g.det_pitch_iso * (g.det_pixel_num / g.H) + g.z_reg * g.det_pitch_iso / g.det_pitch;
    double pr = g.det_pitch_iso * (xc[a] + jj) * (g.det_pixel_num / g.H);
    r.px = pr * cos_t[a] + g.x_reg * g.det_pitch_iso / g.det_pitch;
    r.py = pr * sin_t[a] + g.y_reg * g.det_pitch_iso / g.det_pitch;
    double leng = g.sad * g.sad / (g.sad * g.sad + (pr * pr + r.pz * r.pz));
    r.leng = leng * sqrt(leng);
    return r;
}

/* Literal restatement of the segment enumeration of reprojectionKernel
 * (P/reprojectionOnGPU.cu:154-429 for |dx|>|dy|, :441-666 otherwise).  The five charge sites of each
 * branch are folded into charge(); the operation order of every FP64 expression is the reference's.
 * Behaviour the reference leaves undefined is defined here: a ray with fewer than two box hits has no
 * segments, and a NaN midpoint is not charged. */
struct Tracer {
    const orc_geom& g;
    int nxm, nym, nzm;
    double rmax;
    std::vector<Seg>* out;

    explicit Tracer(const orc_geom& geom) : g(geom) {
        nxm = (int)(g.N * 0.5);
        nym = (int)(g.N * 0.5);
        nzm = (int)(g.Nz * 0.5);
        rmax = g.N * 0.5 * g.s;
        out = nullptr;
    }

    /* P/reprojectionOnGPU.cu:255-269 (one of the 14 copies) */
    void charge(double xb0, double yb0, double zb0, double xb1, double yb1, double zb1) {
        double rr = sqrt((xb0 - xb1) * (xb0 - xb1) + (yb0 - yb1) * (yb0 - yb1) + (zb0 - zb1) * (zb0 - zb1));
        double fx = (xb0 + xb1) * 0.5 / g.s + nxm;
        double fy = (-yb0 - yb1) * 0.5 / g.s + nym;
        double fz = (-zb0 - zb1) * 0.5 / g.t + nzm;
        if (!(fabs(fx) < 1e9 && fabs(fy) < 1e9 && fabs(fz) < 1e9)) return; /* NaN/inf guard */
        int xmc = (int)fx, ymc = (int)fy, zmc = (int)fz;
        if (xmc >= 0 && ymc >= 0 && zmc >= 0 && xmc <= g.N - 1 && ymc <= g.N - 1 && zmc <= g.Nz - 1) {
            double r = sqrt((xb0 + xb1) * (xb0 + xb1) + (-yb0 - yb1) * (-yb0 - yb1)) * 0.5;
            Seg sg;
            sg.xmc = xmc;
            sg.ymc = ymc;
            sg.zmc = zmc;
            sg.rr = rr;
            sg.ph = (r >= rmax) ? 0.0 : 1.0;
            out->push_back(sg);
        }
    }

    void trace(const Ray& R, std::vector<Seg>& segs) {
        out = &segs;
        segs.clear();
        const double sx = R.sx, sy = R.sy, sz = R.sz, px = R.px, py = R.py, pz = R.pz;
        double xx[4], yy[4], zz[4];
        int m = 0;
        yy[m] = double(nym) * g.s;
        if (fabs((yy[m] - sy) / (py - sy) * (px - sx) + sx) < double(nxm) * g.s) {
            xx[m] = (yy[m] - sy) / (py - sy) * (px - sx) + sx;
            zz[m] = (yy[m] - sy) / (py - sy) * (pz - sz) + sz;
            m++;
        }
        yy[m] = -double(nym) * g.s;
        if (fabs((yy[m] - sy) / (py - sy) * (px - sx) + sx) < double(nxm) * g.s) {
            xx[m] = (yy[m] - sy) / (py - sy) * (px - sx) + sx;
            zz[m] = (yy[m] - sy) / (py - sy) * (pz - sz) + sz;
            m++;
        }
        if (m < 4) {
            xx[m] = double(nxm) * g.s;
            if (fabs((xx[m] - sx) / (px - sx) * (py - sy) + sy) < double(nym) * g.s) {
                yy[m] = (xx[m] - sx) / (px - sx) * (py - sy) + sy;
                zz[m] = (xx[m] - sx) / (px - sx) * (pz - sz) + sz;
                m++;
            }
        }
        if (m < 4) {
            xx[m] = -double(nxm) * g.s;
            if (fabs((xx[m] - sx) / (px - sx) * (py - sy) + sy) < double(nym) * g.s) {
                yy[m] = (xx[m] - sx) / (px - sx) * (py - sy) + sy;
                zz[m] = (xx[m] - sx) / (px - sx) * (pz - sz) + sz;
                m++;
            }
        }
        if (m < 2) return; /* undefined in the reference (uninitialised hits); defined as a miss */
        double xi, yi, zi, xf, yf, zf;
        if ((xx[0] - sx) * (xx[0] - sx) + (yy[0] - sy) * (yy[0] - sy) + (zz[0] - sz) * (zz[0] - sz) <
            (xx[1] - sx) * (xx[1] - sx) + (yy[1] - sy) * (yy[1] - sy) + (zz[1] - sz) * (zz[1] - sz)) {
            xi = xx[0]; yi = yy[0]; zi = zz[0]; xf = xx[1]; yf = yy[1]; zf = zz[1];
        } else {
            xi = xx[1]; yi = yy[1]; zi = zz[1]; xf = xx[0]; yf = yy[0]; zf = zz[0];
        }
        double xb0 = xi, yb0 = yi, zb0 = zi;
        double xii, yii, zii, xiii = 0, yiii = 0, ziii = 0, xiiii = 0, yiiii = 0, ziiii = 0, yj, xj, zj;
        int xma = 0, yma = 0, zma = 0, xmb, ymb, zmb, icheck;
        const double lim_x = double(nxm) * g.s, lim_y = double(nym) * g.s;
        const long max_steps = 4L * (g.N + 8) + 64; /* the reference's loop is bounded by the box */

        if (fabs(xi - xf) > fabs(yi - yf)) {
            double pha = (xf - xi) / fabs(xi - xf);
            int xin, ik = 0;
            double chy, chz;
            if ((xi < 0 && (xf - xi) < 0) || (xi > 0 && (xf - xi) > 0)) xin = int(xi) + (xf - xi) / fabs(xf - xi);
            else xin = int(xi);
            xii = xin;
            yii = (xii - sx) / (px - sx) * (py - sy) + sy;
            zii = (xii - sx) / (px - sx) * (pz - sz) + sz;
            while (xii <= lim_x && xii >= -lim_x && yii <= lim_y && yii >= -lim_y && ik < max_steps) {
                xii = xin + pha * ik * g.s;
                ymb = yma;
                yj = yii;
                yii = (xii - sx) / (px - sx) * (py - sy) + sy;
                yma = (int)(yii / g.s);
                zmb = zma;
                zj = zii;
                zii = (xii - sx) / (px - sx) * (pz - sz) + sz;
                zma = (int)(zii / g.t);
                icheck = 0;
                chy = 0;
                chz = 0;
                if ((ik != 0 && yma != ymb) || (ik != 0 && (yii / fabs(yii) != yj / fabs(yj)))) {
                    if ((yii <= 0 && (yf - yi) < 0) || (yii >= 0 && (yf - yi) > 0)) yiii = (double)(yma * g.s);
                    else yiii = (double)(yma * g.s) - (yf - yi) / fabs(yf - yi) * g.s;
                    xiii = (yiii - sy) / (py - sy) * (px - sx) + sx;
                    ziii = (yiii - sy) / (py - sy) * (pz - sz) + sz;
                    chy = (xii - xiii) * (xii - xiii) + (yii - yiii) * (yii - yiii) + (zii - ziii) * (zii - ziii);
                    icheck++;
                }
                if (ik != 0 && zma != zmb) {
                    if ((zii < 0 && (zf - zi) < 0) || (zii > 0 && (zf - zi) > 0)) ziiii = (double)(zma * g.t);
                    else ziiii = (double)(zma * g.t) - (zf - zi) / fabs(zf - zi) * g.t;
                    xiiii = (ziiii - sz) / (pz - sz) * (px - sx) + sx;
                    yiiii = (ziiii - sz) / (pz - sz) * (py - sy) + sy;
                    chz = (xii - xiiii) * (xii - xiiii) + (yii - yiiii) * (yii - yiiii) + (zii - ziiii) * (zii - ziiii);
                    icheck++;
                }
                step_tail(icheck, chy, chz, xb0, yb0, zb0, xiii, yiii, ziii, xiiii, yiiii, ziiii, xii, yii, zii);
                ik++;
            }
        } else {
            double pha = (yf - yi) / fabs(yi - yf);
            int yin, ik = 0;
            double chy, chz;
            if ((yi < 0 && (yf - yi) < 0) || (yi > 0 && (yf - yi) > 0)) yin = int(yi) + (yf - yi) / fabs(yf - yi);
            else yin = int(yi);
            yii = yin;
            xii = (yii - sy) / (py - sy) * (px - sx) + sx;
            zii = (yii - sy) / (py - sy) * (pz - sz) + sz;
            while (xii <= lim_x && xii >= -lim_x && yii <= lim_y && yii >= -lim_y && ik < max_steps) {
                yii = yin + pha * ik * g.s;
                xmb = xma;
                xj = xii;
                xii = (yii - sy) / (py - sy) * (px - sx) + sx;
                xma = (int)(xii / g.s);
                zmb = zma;
                zj = zii;
                zii = (yii - sy) / (py - sy) * (pz - sz) + sz;
                zma = (int)(zii / g.t);
                icheck = 0;
                chy = 0;
                chz = 0;
                if ((ik != 0 && xma != xmb) || (ik != 0 && (xii / fabs(xii) != xj / fabs(xj)))) {
                    if ((xii <= 0 && (xf - xi) < 0) || (xii >= 0 && (xf - xi) > 0)) xiii = (double)(xma * g.s);
                    else xiii = (double)(xma * g.s) - (xf - xi) / fabs(xf - xi) * g.s;
                    yiii = (xiii - sx) / (px - sx) * (py - sy) + sy;
                    ziii = (xiii - sx) / (px - sx) * (pz - sz) + sz;
                    chy = (xii - xiii) * (xii - xiii) + (yii - yiii) * (yii - yiii) + (zii - ziii) * (zii - ziii);
                    icheck++;
                }
                if (ik != 0 && zma != zmb) {
                    if ((zii < 0 && (zf - zi) < 0) || (zii > 0 && (zf - zi) > 0)) ziiii = (double)(zma * g.t);
                    else ziiii = (double)(zma * g.t) - (zf - zi) / fabs(zf - zi) * g.t;
                    xiiii = (ziiii - sz) / (pz - sz) * (px - sx) + sx;
                    yiiii = (ziiii - sz) / (pz - sz) * (py - sy) + sy;
                    chz = (xii - xiiii) * (xii - xiiii) + (yii - yiiii) * (yii - yiiii) + (zii - ziiii) * (zii - ziiii);
                    icheck++;
                }
                step_tail(icheck, chy, chz, xb0, yb0, zb0, xiii, yiii, ziii, xiiii, yiiii, ziiii, xii, yii, zii);
                ik++;
            }
        }
    }

    /* The segment emission of one step (P/reprojectionOnGPU.cu:249-427, identical in both branches). */
    void step_tail(int icheck, double chy, double chz, double& xb0, double& yb0, double& zb0, double xiii,
                   double yiii, double ziii, double xiiii, double yiiii, double ziiii, double xii, double yii,
                   double zii) {
        if (icheck == 2 && chz > chy) {
            charge(xb0, yb0, zb0, xiiii, yiiii, ziiii);
            xb0 = xiiii; yb0 = yiiii; zb0 = ziiii;
            charge(xb0, yb0, zb0, xiii, yiii, ziii);
            xb0 = xiii; yb0 = yiii; zb0 = ziii;
        } else if (icheck == 2 && chz < chy) {
            charge(xb0, yb0, zb0, xiii, yiii, ziii);
            xb0 = xiii; yb0 = yiii; zb0 = ziii;
            charge(xb0, yb0, zb0, xiiii, yiiii, ziiii);
            xb0 = xiiii; yb0 = yiiii; zb0 = ziiii;
        } else if (icheck == 1 && chz != 0) {
            charge(xb0, yb0, zb0, xiiii, yiiii, ziiii);
            xb0 = xiiii; yb0 = yiiii; zb0 = ziiii;
        } else if (icheck == 1 && chy != 0) {
            charge(xb0, yb0, zb0, xiii, yiii, ziii);
            xb0 = xiii; yb0 = yiii; zb0 = ziii;
        }
        charge(xb0, yb0, zb0, xii, yii, zii);
        xb0 = xii; yb0 = yii; zb0 = zii;
    }
};

inline size_t out_index(const orc_geom& g, int a, int j, int jj) {
    /* P/reprojectionOnGPU.cu:106: row H-1-j */
    return (size_t)a * g.H * g.W + (size_t)(g.H - 1 - j) * g.W + jj;
}

inline double vox(const uint16_t* vu, const float* vf, size_t idx) {
    return vu ? (double)vu[idx] : (double)vf[idx];
}

}  // namespace

extern "C" {

void orc_trig_table_proj(int W, int H, int n, const double* angles_rad, const double* xshift,
                         const double* yshift, double det_pitch, double det_pixel_num, double* cos_t,
                         double* sin_t, double* xc, double* yc) {
    (void)W;
    /* P/misc.cpp:25-34: both shifts use the projection HEIGHT. */
    for (int i = 0; i < n; i++) {
        cos_t[i] = cos(angles_rad[i]);
        sin_t[i] = sin(angles_rad[i]);
        xc[i] = -H / 2.0 - xshift[i] * 10.0 / (det_pitch * (det_pixel_num / H));
        yc[i] = -H / 2.0 - yshift[i] * 10.0 / (det_pitch * (det_pixel_num / H));
    }
}

void orc_trig_table_fdk(int n, const double* angles_rad, double* cos_t, double* sin_t) {
    for (int i = 0; i < n; i++) {
        cos_t[i] = cos(angles_rad[i]);
        sin_t[i] = sin(angles_rad[i]);
    }
}

int orc_attenuation_table(const char* spectrum_csv, const char* xraydb_dir, int nmat, int nthin, float* atten,
                          float* pdf, int cap) {
    /* P/prior_weight_production.cpp:200-270: spectrum read, thinning, normalisation. */
    FILE* fp = fopen(spectrum_csv, "r");
    if (!fp) return -1;
    char row[1024];
    float d0 = 0, d1 = 0, d2 = 0, sum = 0;
    int nne = 0;
    std::vector<double> energy(1024);
    std::vector<float> pdf_raw(1024);
    if (fgets(row, sizeof(row), fp) == NULL) { fclose(fp); return -2; }
    while (fgets(row, sizeof(row), fp) != NULL) {
        if (nne >= 1024) { fclose(fp); return -3; }
        /* the reference does not check the conversion count: an unparsable row repeats the previous
         * row's values (P/prior_weight_production.cpp:220-228) */
        sscanf(row, "%f,%f,%f", &d0, &d1, &d2);
        energy[nne] = (d1 + d0) * 0.5;
        pdf_raw[nne] = d2;
        sum += d2;
        nne++;
    }
    fclose(fp);
    if (nne == 0 || nthin < 1) return -4;
    double psum = 0.0, esum = 0.0, sum_thin = 0.0;
    int nne_thin = 0;
    float pdf_thin[1024], pe_thin[1024];
    for (int i = 0; i < nne; i++) {
        psum += pdf_raw[i];
        esum += energy[i];
        if (i - nthin + 1 == (i - nthin + 1) / nthin * nthin) {
            int k = (i - nthin + 1) / nthin;
            if (k < 0) continue;
            pdf_thin[k] = psum;
            pe_thin[k] = esum / nthin;
            sum_thin += psum;
            psum = 0.0;
            esum = 0.0;
            nne_thin = k + 1;
        }
    }
    nne = nne_thin;
    sum = sum_thin;
    if (nne > cap) return -5;
    for (int i = 0; i < nne; i++) {
        pdf[i] = pdf_thin[i];
        energy[i] = pe_thin[i];
    }
    for (int i = 0; i < nne; i++) pdf[i] = pdf[i] / sum;

    /* P/prior_weight_production.cpp:318-507: xraydb rows, 5 eV grid, match |dE| < 30e-9 MeV,
     * mu/rho = photo + compton + rayleigh.  A bin with no match keeps the previous bin's values. */
    const int ndb = 29001;
    if (nmat < 1 || nmat > 6) return -6;
    std::vector<double> edb(ndb), ph(ndb * 6), co(ndb * 6), ra(ndb * 6);
    const char* names[3] = {"photo_5keVto150keV.csv", "comp_5keVto150keV.csv", "rayl_5keVto150keV.csv"};
    std::vector<double>* dst[3] = {&ph, &co, &ra};
    for (int f = 0; f < 3; f++) {
        char path[2048];
        snprintf(path, sizeof(path), "%s/%s", xraydb_dir, names[f]);
        FILE* fd = fopen(path, "rb");
        if (!fd) return -7;
        char lb[512];
        if (!fgets(lb, sizeof(lb), fd)) { fclose(fd); return -8; }
        for (int e = 0; e < ndb; e++) {
            double v[6], en;
            if (!fgets(lb, sizeof(lb), fd)) { fclose(fd); return -9; }
            sscanf(lb, "%lf,%lf,%lf,%lf,%lf,%lf,%lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &en);
            if (f == 0) edb[e] = en / 1000 / 1000;
            for (int k = 0; k < 6; k++) (*dst[f])[k * ndb + e] = v[k];
        }
        fclose(fd);
    }
    double photo = 0, compton = 0, rayleigh = 0;
    for (int iz = 0; iz < nmat; iz++) {
        for (int i = 0; i < nne; i++) {
            for (int e = 0; e < ndb; e++) {
                double aaa = edb[e] * 1000 * 1000 * 1000 - energy[i] * 1000 * 1000 * 1000;
                if (-30 < aaa && aaa < 30) {
                    photo = ph[iz * ndb + e];
                    compton = co[iz * ndb + e];
                    rayleigh = ra[iz * ndb + e];
                }
            }
            atten[iz * nne + i] = (photo + compton + rayleigh + 0.0);
        }
    }
    return nne;
}

int orc_forward_compat(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc,
                       const double* yc, const uint16_t* vol, const float* atten, const float* pdf, int NE,
                       double xfactor, int a0, int a1, float* out) {
    const orc_geom& g = *gp;
    const size_t n2 = (size_t)g.N * g.N;
#pragma omp parallel
    {
        Tracer tr(g);
        std::vector<Seg> segs;
        segs.reserve(8192);
#pragma omp for collapse(2) schedule(dynamic, 4)
        for (int a = a0; a < a1; a++)
            for (int j = 0; j < g.H; j++)
                for (int jj = 0; jj < g.W; jj++) {
                    Ray R = make_ray(g, cos_t, sin_t, xc, yc, a, j, jj);
                    tr.trace(R, segs);
                    float acc = 0.0f; /* P/reprojectionOnGPU.cu:1038-1044 (FP32, ke order) */
                    for (int ke = 0; ke < NE; ke++) {
                        double transp = 0.0;
                        for (const Seg& sg : segs)
                            for (int mat = 0; mat < g.nmat; mat++) {
                                size_t vnum = sg.xmc + (size_t)sg.ymc * g.N + ((size_t)sg.zmc * g.nmat + mat) * n2;
                                /* u16 * float is an FP32 product in the reference (:268) */
                                float vm = (float)vol[vnum] * atten[mat * NE + ke];
                                transp += vm * sg.rr * sg.ph;
                            }
                        double intens = pdf[ke] * exp(-transp / xfactor) * R.leng;
                        acc += (float)intens;
                    }
                    out[out_index(g, a, j, jj)] = acc;
                }
    }
    return 0;
}

int orc_line_integrals(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc,
                       const double* yc, const uint16_t* vu, const float* vf, int a0, int a1, double* L,
                       int32_t* nseg, double* leng) {
    const orc_geom& g = *gp;
    const size_t n2 = (size_t)g.N * g.N;
#pragma omp parallel
    {
        Tracer tr(g);
        std::vector<Seg> segs;
        segs.reserve(8192);
#pragma omp for collapse(2) schedule(dynamic, 4)
        for (int a = a0; a < a1; a++)
            for (int j = 0; j < g.H; j++)
                for (int jj = 0; jj < g.W; jj++) {
                    Ray R = make_ray(g, cos_t, sin_t, xc, yc, a, j, jj);
                    tr.trace(R, segs);
                    size_t o = out_index(g, a, j, jj);
                    double Lm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (const Seg& sg : segs)
                        for (int mat = 0; mat < g.nmat; mat++) {
                            size_t vnum = sg.xmc + (size_t)sg.ymc * g.N + ((size_t)sg.zmc * g.nmat + mat) * n2;
                            Lm[mat] += vox(vu, vf, vnum) * sg.rr * sg.ph;
                        }
                    for (int mat = 0; mat < g.nmat; mat++) L[o * g.nmat + mat] = Lm[mat];
                    if (nseg) nseg[o] = (int32_t)segs.size();
                    if (leng) leng[o] = R.leng;
                }
    }
    return 0;
}

int orc_forward_collapsed(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc,
                          const double* yc, const uint16_t* vu, const float* vf, const float* atten,
                          const float* pdf, int NE, double xfactor, int a0, int a1, float* out) {
    const orc_geom& g = *gp;
    size_t nray = (size_t)(a1 - a0) * g.H * g.W;
    std::vector<double> L(nray * g.nmat), lg(nray);
    /* L and leng are indexed from angle a0 */
    std::vector<double> cs(cos_t + a0, cos_t + a1), sn(sin_t + a0, sin_t + a1), xs(xc + a0, xc + a1),
        ys(yc + a0, yc + a1);
    orc_line_integrals(gp, cs.data(), sn.data(), xs.data(), ys.data(), vu, vf, 0, a1 - a0, L.data(), nullptr,
                       lg.data());
    float* o = out + (size_t)a0 * g.H * g.W;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nray; i++) {
        float acc = 0.0f;
        for (int ke = 0; ke < NE; ke++) {
            double transp = 0.0;
            for (int mat = 0; mat < g.nmat; mat++) transp += (double)atten[mat * NE + ke] * L[i * g.nmat + mat];
            double intens = pdf[ke] * exp(-transp / xfactor) * lg[i];
            acc += (float)intens;
        }
        o[i] = acc;
    }
    return 0;
}

int orc_logproj(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc, const double* yc,
                const float* vf, const float* atten, const float* pdf, int NE, double xfactor, double* p) {
    const orc_geom& g = *gp;
    size_t nray = (size_t)g.n_ang * g.H * g.W;
    std::vector<double> L(nray * g.nmat);
    orc_line_integrals(gp, cos_t, sin_t, xc, yc, nullptr, vf, 0, g.n_ang, L.data(), nullptr, nullptr);
    double wsum = 0.0;
    for (int e = 0; e < NE; e++) wsum += pdf[e];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nray; i++) {
        double s = 0.0;
        for (int e = 0; e < NE; e++) {
            double tr = 0.0;
            for (int m = 0; m < g.nmat; m++) tr += (double)atten[m * NE + e] * L[i * g.nmat + m];
            s += pdf[e] * exp(-tr / xfactor);
        }
        p[i] = -log(s / wsum);
    }
    return 0;
}

int orc_adjoint(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc, const double* yc,
                const float* vf, const float* atten, const float* pdf, int NE, double xfactor, const double* target,
                double* grad, double* loss) {
    const orc_geom& g = *gp;
    const size_t n2 = (size_t)g.N * g.N;
    const size_t nvox = n2 * g.Nz * g.nmat;
    memset(grad, 0, nvox * sizeof(double));
    double wsum = 0.0;
    for (int e = 0; e < NE; e++) wsum += pdf[e];
    double total = 0.0;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> part(nthr);
#pragma omp parallel reduction(+ : total)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::vector<double>& gl = part[tid];
        gl.assign(nvox, 0.0);
        Tracer tr(g);
        std::vector<Seg> segs;
        segs.reserve(8192);
#pragma omp for collapse(2) schedule(dynamic, 4)
        for (int a = 0; a < g.n_ang; a++)
            for (int j = 0; j < g.H; j++)
                for (int jj = 0; jj < g.W; jj++) {
                    Ray R = make_ray(g, cos_t, sin_t, xc, yc, a, j, jj);
                    tr.trace(R, segs);
                    double Lm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (const Seg& sg : segs)
                        for (int m = 0; m < g.nmat; m++) {
                            size_t vnum = sg.xmc + (size_t)sg.ymc * g.N + ((size_t)sg.zmc * g.nmat + m) * n2;
                            Lm[m] += (double)vf[vnum] * sg.rr * sg.ph;
                        }
                    double s = 0.0, mub[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (int e = 0; e < NE; e++) {
                        double trn = 0.0;
                        for (int m = 0; m < g.nmat; m++) trn += (double)atten[m * NE + e] * Lm[m];
                        double we = pdf[e] * exp(-trn / xfactor);
                        s += we;
                        for (int m = 0; m < g.nmat; m++) mub[m] += we * atten[m * NE + e];
                    }
                    double p = -log(s / wsum);
                    double r = p - target[out_index(g, a, j, jj)];
                    total += r * r;
                    /* dp/dL_m = mubar_m / xf with mubar the exp-weighted spectral mean of mu */
                    double coef[8];
                    for (int m = 0; m < g.nmat; m++) coef[m] = 2.0 * r * (mub[m] / s) / xfactor;
                    for (const Seg& sg : segs)
                        for (int m = 0; m < g.nmat; m++) {
                            size_t vnum = sg.xmc + (size_t)sg.ymc * g.N + ((size_t)sg.zmc * g.nmat + m) * n2;
                            gl[vnum] += coef[m] * sg.rr * sg.ph;
                        }
                }
    }
    for (int t = 0; t < nthr; t++)
        if (!part[t].empty())
            for (size_t i = 0; i < nvox; i++) grad[i] += part[t][i];
    if (loss) *loss = total;
    return 0;
}

/* Debug aid: the charged segments of one ray (xmc, ymc, zmc, rr, ph per row of `out`, capacity cap). */
int orc_trace_ray(const orc_geom* gp, const double* cos_t, const double* sin_t, const double* xc, const double* yc,
                  int a, int j, int jj, double* out, int cap) {
    const orc_geom& g = *gp;
    Tracer tr(g);
    std::vector<Seg> segs;
    Ray R = make_ray(g, cos_t, sin_t, xc, yc, a, j, jj);
    tr.trace(R, segs);
    int n = 0;
    for (const Seg& sg : segs) {
        if (n >= cap) break;
        out[5 * n + 0] = sg.xmc; out[5 * n + 1] = sg.ymc; out[5 * n + 2] = sg.zmc;
        out[5 * n + 3] = sg.rr; out[5 * n + 4] = sg.ph;
        ++n;
    }
    return (int)segs.size();
}

void orc_const_map(int W, double pitch_iso, double* cmap) { orc_const_map_w(W, pitch_iso, 0.5, cmap); }

void orc_const_map_w(int W, double pitch_iso, double ALP, double* cmap) {
    /* F/filterSinogram.cpp:10-65 with MAX_PIC_NUM 800, CUTOFF_FREQ 1.6, GAMMA 1, and the window parameter ALPHA
       (0.5 in F/physParams.h:39-41; 1 is the build's opt-in band-limited ramp) */
    const double CUT = 1.60, GAM = 1.00;
    double theConst = 1.0 / (8.0 * kPI * kPI * pitch_iso * 800 / W);
    double gamma_x_cutoffFreq = GAM * CUT;
    double PI_x_cutoffFreq = kPI * CUT;
    double one_minus_alpha = 1.0 - ALP;
    double omega = kPI / GAM;
    double cos_omega = cos(omega);
    double sin_omega = sin(omega);
    for (int m = 0; m < W; m++)
        for (int n = 0; n < W; n++) {
            if (m == n) {
                cmap[(size_t)m * W + n] = 0.0;
            } else {
                double MN = (double)(m - n);
                double theta = PI_x_cutoffFreq * MN;
                double delta = gamma_x_cutoffFreq * MN;
                double eta = 1.0 / (1.0 - 1.0 / delta);
                double xi = 1.0 / (1.0 + 1.0 / delta);
                double cos_theta = cos(theta);
                double sin_theta = sin(theta);
                cmap[(size_t)m * W + n] =
                    theConst / MN / MN *
                    (2.0 * ALP * theta * sin_theta + 2.0 * ALP * (cos_theta - 1.0) +
                     xi * one_minus_alpha * theta * (sin_theta * cos_omega + cos_theta * sin_omega) +
                     eta * one_minus_alpha * theta * (sin_theta * cos_omega - cos_theta * sin_omega) +
                     xi * xi * one_minus_alpha * (cos_theta * cos_omega - sin_theta * sin_omega - 1.0) +
                     eta * eta * one_minus_alpha * (cos_theta * cos_omega + sin_theta * sin_omega - 1.0));
            }
        }
}

static double filter_coeff(int W, double pitch_iso, double ALP = 0.50) {
    /* F/filterSinogram.cpp:78-80 */
    const double CUT = 1.60, GAM = 1.00;
    double omega = kPI / GAM;
    return CUT * CUT / 4.0 / (pitch_iso * 800 / W) *
           (ALP * 0.5 + (1.0 - ALP) * (1.0 / omega * sin(omega) + 1.0 / omega / omega * (cos(omega) - 1.0)));
}

static inline double filter_cm(int W, int n, int m, double xs, double ys, double pitch_iso) {
    /* F/filterSinogram.cpp:86-99: both centres use the projection WIDTH */
    double XcenterShift = -W / 2.0 - xs / (0.05 * 800 / W) - 1;
    double YcenterShift = -W / 2.0 - ys / (0.05 * 800 / W) - 1;
    double pos2 = fabs((YcenterShift + n)) * (pitch_iso * 800 / W);
    double pos1 = fabs((XcenterShift + m)) * (pitch_iso * 800 / W);
    return 100.0 / sqrt(100.0 * 100.0 + pos1 * pos1 + pos2 * pos2);
}

void orc_filter_sinogram(int H, int W, int n_ang, const float* proj, float* out, const double* cmap,
                         const double* xshift, const double* yshift, double pitch_iso) {
    /* F/filterSinogram.cpp:84-118: FP64 terms accumulated into an FP32 lvalue, k ascending */
    double coeff = filter_coeff(W, pitch_iso);
#pragma omp parallel for
    for (int th = 0; th < n_ang; th++)
        for (int n = 0; n < H; n++) {
            const float* q = proj + ((size_t)th * H + n) * W;
            float* o = out + ((size_t)th * H + n) * W;
            for (int m = 0; m < W; m++) {
                double Cm = filter_cm(W, n, m, xshift[th], yshift[th], pitch_iso);
                o[m] = -log(q[m]) * Cm * coeff;
                for (int k = 0; k < W; k++) o[m] += cmap[(size_t)m * W + k] * (-log(q[k])) * Cm;
            }
        }
}

void orc_filter_sinogram64(int H, int W, int n_ang, const float* proj, double* out, const double* cmap,
                           const double* xshift, const double* yshift, double pitch_iso) {
    orc_filter_sinogram64_w(H, W, n_ang, proj, out, cmap, xshift, yshift, pitch_iso, 0.5);
}

void orc_filter_sinogram64_w(int H, int W, int n_ang, const float* proj, double* out, const double* cmap,
                             const double* xshift, const double* yshift, double pitch_iso, double alpha) {
    double coeff = filter_coeff(W, pitch_iso, alpha);
#pragma omp parallel for
    for (int th = 0; th < n_ang; th++)
        for (int n = 0; n < H; n++) {
            const float* q = proj + ((size_t)th * H + n) * W;
            double* o = out + ((size_t)th * H + n) * W;
            for (int m = 0; m < W; m++) {
                double Cm = filter_cm(W, n, m, xshift[th], yshift[th], pitch_iso);
                double acc = -log(q[m]) * coeff;
                for (int k = 0; k < W; k++) acc += cmap[(size_t)m * W + k] * (-log(q[k]));
                o[m] = acc * Cm;
            }
        }
}

void orc_mask(int R, char* mask) {
    /* F/misc.cpp:20-35 */
    double rad = (double)(R - 1) / 2.0, sq = rad * rad;
    for (int j = 0; j < R; j++)
        for (int i = 0; i < R; i++)
            mask[(size_t)j * R + i] =
                (((double)j - rad) * ((double)j - rad) + ((double)i - rad) * ((double)i - rad) <= sq) ? 1 : 0;
}

void orc_reconstruct_image(int R, int Z, int zstart, double step, double scale, double* recon, const char* mask,
                           int H, int W, int n_ang, const float* sino, int xfactor, const double* angles,
                           const double* cos_t, const double* sin_t, const double* xshift, double pitch_iso) {
    /* F/reconstructImage.cpp:10-184, accumulating into the caller's buffer (+=) */
    double centerPos = (double)R / 2.0 - 0.5;
#pragma omp parallel for
    for (int z = 0; z < Z; z++) {
        double zz = zstart + z * step;
        for (int y = 0; y < R; y++)
            for (int x = 0; x < R; x++) {
                if (mask[(size_t)y * R + x] == 0) continue;
                double xx = (x - centerPos) * scale;
                double yy = -(y - centerPos) * scale;
                double& acc = recon[((size_t)z * R + y) * R + x];
                for (int th = 0; th < n_ang; th++) {
                    double dbeta, thdb0, thdb1, thdb2;
                    if (th == 0) {
                        thdb1 = angles[th];
                        thdb2 = angles[th + 1];
                        dbeta = fabs(thdb2 - thdb1) / 2;
                    } else if (th == n_ang - 1) {
                        thdb0 = angles[th - 1];
                        thdb1 = angles[th];
                        dbeta = fabs(thdb1 - thdb0) / 2;
                    } else {
                        thdb0 = angles[th - 1];
                        thdb2 = angles[th + 1];
                        thdb1 = angles[th + 1] - 2 * kPI;
                        if (thdb2 < 0) thdb1 = angles[th + 1] + 2 * kPI;
                        dbeta = fabs(thdb2 - thdb0) / 2;
                        if (dbeta > kPI / 2) dbeta = fabs(thdb1 - thdb0) / 2;
                    }
                    double c = cos_t[th], s = sin_t[th];
                    double X = xx * c + yy * s;
                    double Y = -xx * s + yy * c;
                    double L2 = 1.0 - Y / 100.0;
                    double U = X / L2 / (pitch_iso * 800 / W);
                    double w = fabs(fabs(U) - fabs(floor(U)));
                    double UUU = (double)(W / 2 + U + xshift[th] / (0.05 * 800 / W) - 0.5);
                    int UU = (int)floor(UUU);
                    double ZETA = (double)(H / 2) + zz / L2 / (pitch_iso * 800 / W);
                    int NZ = (int)floor(ZETA);
                    double wz = fabs(fabs(ZETA) - fabs(floor(ZETA)));
                    const float* P = sino + (size_t)th * H * W;
                    double sum;
                    bool hit = false;
                    if ((UU >= 0 && UU < W - 1) && (NZ >= 0 && NZ < H - 1)) {
                        sum = P[(size_t)NZ * W + UU] * (1.0 - w) * (1.0 - wz) + P[(size_t)NZ * W + UU + 1] * w * (1.0 - wz) +
                              P[(size_t)(NZ + 1) * W + UU] * (1.0 - w) * wz + P[(size_t)(NZ + 1) * W + UU + 1] * w * wz;
                        hit = true;
                    } else if ((UU == W - 1) && (NZ >= 0 && NZ < H - 1)) {
                        sum = P[(size_t)NZ * W + UU] * (1.0 - wz) + P[(size_t)(NZ + 1) * W + UU] * wz;
                        hit = true;
                    } else if ((UU >= 0 && UU < W - 1) && (NZ == H - 1)) {
                        sum = P[(size_t)NZ * W + UU] * (1.0 - w) + P[(size_t)NZ * W + UU + 1] * w;
                        hit = true;
                    } else if ((UU == W - 1) && (NZ == H - 1)) {
                        sum = P[(size_t)NZ * W + UU];
                        hit = true;
                    }
                    if (hit) acc += (1.0 / L2 / L2 * sum * dbeta * xfactor);
                }
            }
    }
}

int orc_fdk_angles(const char* flexmap, const char* tgt_class, int total, double* angles_rad, double* xshift,
                   double* yshift, int cap) {
    /* F/projectionDataMVkV.cpp:155-274 */
    FILE* fp = fopen(flexmap, "r");
    if (!fp) return -1;
    std::vector<double> pa, px, py;
    std::vector<int> ph;
    double angle, x, y, cw0 = 0, cw = 0;
    char cls[64];
    int i = 0;
    while (fscanf(fp, "%lf %lf %lf %63s", &angle, &x, &y, cls) == 4) {
        i++;
        if (i == 1) cw0 = angle;
        if (i == 30) cw = (cw0 - angle);
        ph.push_back((!strcmp(tgt_class, "*") || !strcmp(cls, tgt_class)) ? 1 : 0);
        pa.push_back(angle);
        px.push_back(x);
        py.push_back(y);
    }
    fclose(fp);
    int nl = (int)pa.size();
    /* a flexmap shorter than 30 rows leaves cwccw uninitialised in the reference; here it is 0 (CCW) */
    bool rev = cw < 0;
    int used = total;
    if (used > cap) return -2;
    int cnt = 0;
    for (int k = 0; k < nl; k++) {
        int src = rev ? (nl - 1 - k) : k;
        if (ph[src] == 1) {
            if (cnt >= used) break; /* the reference writes past its arrays here */
            angles_rad[cnt] = pa[src] * kPI / 180.0;
            xshift[cnt] = px[src];
            yshift[cnt] = py[src];
            cnt++;
        }
    }
    return cnt;
}

}  // extern "C"

/* ------------------------------------------------------------------------------------------------
 * Organ-label phantom -> elemental maps (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp,
 * R/ below).
 *
 * rand() of glibc, restated from its published description (the additive feedback generator behind
 * random() with the default 128-byte state): h[0] = seed (0 is taken as 1), h[i] = 16807 h[i-1] mod
 * (2^31 - 1) for 0 < i < 31, h[i] = h[i-31] for 31 <= i < 34, h[i] = h[i-31] + h[i-3] mod 2^32 from
 * i = 34 on; the k-th rand() after srand(seed) is h[344 + k] >> 1. */
struct OrcRand {
    std::vector<uint32_t> h;
    explicit OrcRand(unsigned seed) : h(344) {
        h[0] = seed ? seed : 1u;
        for (int i = 1; i < 31; i++) {
            long long v = (16807LL * (int32_t)h[i - 1]) % 2147483647LL;
            if (v < 0) v += 2147483647LL;
            h[i] = (uint32_t)v;
        }
        for (int i = 31; i < 34; i++) h[i] = h[i - 31];
        for (int i = 34; i < 344; i++) h[i] = h[i - 31] + h[i - 3];
    }
    int next() {
        size_t n = h.size();
        h.push_back(h[n - 31] + h[n - 3]);
        return (int)(h.back() >> 1);
    }
    /* Uniform0, R/ :14-17 */
    double uniform() { return ((double)next() + 1.0) / ((double)2147483647 + 2.0); }
    /* rand_gauss0, R/ :19-24; the log factor's draw is taken before the sine's */
    double gauss(double mu, double sigma) {
        double PI = 3.14159265;
        double a = sqrt(-2.0 * log(uniform()));
        double z = a * sin(2.0 * PI * uniform());
        return mu + sigma * z;
    }
};

extern "C" {

int orc_rand_sequence(unsigned seed, int n, int* out) {
    OrcRand r(seed);
    for (int i = 0; i < n; i++) out[i] = r.next();
    return n;
}

int orc_material_weights(const char* media_csv, int randomize, unsigned seed, float* wout, int cap) {
    /* R/ :104-166: rows in file order; H, C, N always and P, Ca above 0.09 % redrawn from N(w, 0.1 w)
     * except medium 53; O = 100 minus the five, renormalised when negative.  wout [cap][6]: H C N O P Ca */
    FILE* fp = fopen(media_csv, "r");
    if (!fp) return -1;
    float Hw, Cw, Nw, Ow, Naw, Maw, Pw, Sw, Clw, Kw, Caw, Few, Iw;
    int ID, i = 0;
    OrcRand rng(seed);
    double sigma, mu;
    while (fscanf(fp, "%d,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f", &ID, &Hw, &Cw, &Nw, &Ow, &Naw, &Maw, &Pw, &Sw,
                  &Clw, &Kw, &Caw, &Few, &Iw) == 14) {
        if (i >= cap) { fclose(fp); return -2; }
        if (randomize && ID != 53) {
            sigma = double(Hw * 0.1); mu = double(Hw);
            Hw = (float)rng.gauss(mu, sigma); if (Hw < 0.0) Hw = 0.0;
            sigma = double(Cw * 0.1); mu = double(Cw);
            Cw = (float)rng.gauss(mu, sigma); if (Cw < 0.0) Cw = 0.0;
            sigma = double(Nw * 0.1); mu = double(Nw);
            Nw = (float)rng.gauss(mu, sigma); if (Nw < 0.0) Nw = 0.0;
            if (Pw > 0.09) {
                sigma = double(Pw * 0.1); mu = double(Pw);
                Pw = (float)rng.gauss(mu, sigma); if (Pw < 0.0) Pw = 0.0;
            }
            if (Caw > 0.09) {
                sigma = double(Caw * 0.1); mu = double(Caw);
                Caw = (float)rng.gauss(mu, sigma); if (Caw < 0.0) Caw = 0.0;
            }
        }
        Ow = 100 - Hw - Cw - Nw - Pw - Caw;
        if (Ow < 0.0) {
            double s = (Hw + Cw + Nw + Pw + Caw) * 0.01;
            Ow = 0.0;
            Hw = Hw / s; Cw = Cw / s; Nw = Nw / s; Pw = Pw / s; Caw = Caw / s;
        }
        float* w = wout + (size_t)i * 6;
        w[0] = Hw; w[1] = Cw; w[2] = Nw; w[3] = Ow; w[4] = Pw; w[5] = Caw;
        i++;
    }
    fclose(fp);
    return i;
}

int orc_organ_table(const char* organs_csv, int* organ, int* medium, float* dens, int cap) {
    /* R/ :168-201 (the density draw there is never used and is not made) */
    FILE* fp = fopen(organs_csv, "r");
    if (!fp) return -1;
    int i = 0, a, b;
    float d;
    while (fscanf(fp, "%d,%d,%f", &a, &b, &d) == 3) {
        if (i >= cap) { fclose(fp); return -2; }
        organ[i] = a; medium[i] = b; dens[i] = d;
        i++;
    }
    fclose(fp);
    return i;
}

int orc_labels_to_maps(const uint8_t* labels, int ncol, int nrow, int nz, const float* w, int nmedia,
                       const int* organ, const int* medium, const float* dens, int norg, int16_t* maps) {
    /* R/ :212-226 per voxel (every organ row tested, the last match wins; a label without a row leaves the
     * voxel at the zero of a fresh allocation), laid out as the per-slice files of R/ :257-270:
     * maps [nz][6][nrow][ncol] */
    for (int j = 0; j < norg; j++)
        if (medium[j] < 1 || medium[j] > nmedia) return -3;
    const size_t n2 = (size_t)ncol * nrow;
#pragma omp parallel for schedule(static)
    for (long long zz = 0; zz < (long long)nz; zz++) {
        for (size_t n = 0; n < n2; n++) {
            int16_t v[6] = {0, 0, 0, 0, 0, 0};
            const int lab = labels[(size_t)zz * n2 + n];
            for (int j = 0; j < norg; j++)
                if (organ[j] == lab) {
                    const float* m = w + (size_t)(medium[j] - 1) * 6;
                    for (int k = 0; k < 6; k++) v[k] = (short)(m[k] * dens[j] * 10);
                }
            for (int k = 0; k < 6; k++) maps[((size_t)zz * 6 + k) * n2 + n] = v[k];
        }
    }
    return 0;
}

}  // extern "C"

/* ------------------------------------------------------------------------------------------------
 * Iterative-reconstruction pieces of the projector tree (not called by its main program).
 *
 * The (int) casts of the reference run on x86, where a NaN or out-of-range double converts to INT_MIN;
 * every later test on such an index fails, so here an invalid conversion simply drops the term. */
static inline bool x86_trunc(double v, int* out) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return false;
    *out = (int)v;
    return true;
}

extern "C" {

int orc_lengthproj(const orc_geom* gp, int ite, const double* cos_t, const double* sin_t, const double* xc,
                   const double* yc, const uint16_t* proj, int a0, int a1, float* npf) {
    /* P/lengthproj.cpp:9-131: for every voxel centre r and angle, the chord of the voxel along the major
     * axis of the source-voxel line times (65535 - the projection sampled where that line meets the detector),
     * accumulated in FP32 over the angles in order. */
    const orc_geom& g = *gp;
    double x_reg = g.x_reg, y_reg = g.y_reg, z_reg = g.z_reg;
    if (ite != 0) { x_reg = 0.0; y_reg = 0.0; z_reg = 0.0; }
    const int N = g.N, Nz = g.Nz, W = g.W, H = g.H;
    const double s = g.s, t = g.t;
    for (size_t i = 0; i < (size_t)N * N * Nz; i++) npf[i] = 0.0f;
#pragma omp parallel for schedule(static)
    for (int jz = 0; jz < Nz; jz++) {
        const double rz = Nz * t * 0.5 - (jz + 0.5) * t;
        for (int jy = 0; jy < N; jy++) {
            const double ry = N * s * 0.5 - (jy + 0.5) * s;
            for (int jx = 0; jx < N; jx++) {
                const double rx = -N * s * 0.5 + (jx + 0.5) * s;
                float acc = 0.0f;
                for (int a = a0; a < a1; a++) {
                    const double sx = -g.sad * sin_t[a] + x_reg;
                    const double sy = g.sad * cos_t[a] + y_reg;
                    const double sz = z_reg;
                    double rr;
                    if (fabs(rx - sx) > fabs(ry - sy)) {
                        const double x0 = rx - 0.5 * s, x1 = rx + 0.5 * s;
                        const double y0 = (x0 - sx) / (rx - sx) * (ry - sy) + sy;
                        const double z0 = (x0 - sx) / (rx - sx) * (rz - sz) + sz;
                        const double y1 = (x1 - sx) / (rx - sx) * (ry - sy) + sy;
                        const double z1 = (x1 - sx) / (rx - sx) * (rz - sz) + sz;
                        rr = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1));
                    } else {
                        const double y0 = ry - 0.5 * s, y1 = ry + 0.5 * s;
                        const double x0 = (y0 - sy) / (ry - sy) * (rx - sx) + sx;
                        const double z0 = (y0 - sy) / (ry - sy) * (rz - sz) + sz;
                        const double x1 = (y1 - sy) / (ry - sy) * (rx - sx) + sx;
                        const double z1 = (y1 - sy) / (ry - sy) * (rz - sz) + sz;
                        rr = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1));
                    }
                    /* detector line through o, perpendicular to the central ray; p = its meet with s->r */
                    const double ox = sx + g.det_pitch / g.det_pitch_iso * sin_t[a] * g.sad;
                    const double oy = sy - g.det_pitch / g.det_pitch_iso * cos_t[a] * g.sad;
                    const double kr = (sx - rx) / (sy - ry), ko = (oy - sy) / (ox - sx);
                    const double py = (ox - rx + kr * ry + ko * oy) / (kr + ko);
                    const double px = (kr * (py - ry) + rx);
                    const double pz = ((sz - rz) / (sy - ry) * (py - ry) + rz);
                    const double pix = g.det_pitch * (g.det_pixel_num / H);
                    const double xshift = -(xc[a] + H / 2.0) * pix;
                    const double yshift = -(yc[a] + H / 2.0) * pix;
                    const double vv = -(pz - z_reg - yshift);
                    /* the reference subtracts x_reg from py too */
                    const double uu = (px - x_reg) * cos_t[a] + (py - x_reg) * sin_t[a] + xshift - (W - H) / 2 * pix;
                    const double vfac = vv < 0 ? -0.5 : 0.5, ufac = uu < 0 ? -0.5 : 0.5;
                    int v1, u1;
                    if (!x86_trunc(vv / pix + vfac, &v1) || !x86_trunc(uu / pix + ufac, &u1)) continue;
                    const double dv1 = vv / pix - v1, du1 = uu / pix - u1;
                    /* the second tap steps toward the sign of (coordinate in cm - index) */
                    const int v2 = fabs(vv - v1) < 0.000000001 ? v1 : v1 + ((vv - v1) > 0 ? 1 : -1);
                    const int u2 = fabs(uu - u1) < 0.000000001 ? u1 : u1 + ((uu - u1) > 0 ? 1 : -1);
                    const int h2 = H / 2, w2 = W / 2;
                    if (v1 + h2 < 0 || v1 + h2 >= H || v2 + h2 < 0 || v2 + h2 >= H || u1 + w2 < 0 || u1 + w2 >= W ||
                        u2 + w2 < 0 || u2 + w2 >= W)
                        continue;
                    const uint16_t* P = proj + (size_t)a * W * H;
                    const double p11 = P[(v1 + h2) * W + u1 + w2], p12 = P[(v2 + h2) * W + u1 + w2];
                    const double p21 = P[(v1 + h2) * W + u2 + w2], p22 = P[(v2 + h2) * W + u2 + w2];
                    /* as written: the v fraction blends the u taps and the u fraction the v taps */
                    const double q1 = ((1 - dv1) * p11 + dv1 * p21);
                    const double q2 = ((1 - dv1) * p12 + dv1 * p22);
                    const double q = ((1 - du1) * q1 + du1 * q2);
                    acc += (float)((65535.0 - q) * rr);
                }
                npf[(size_t)jz * N * N + (size_t)jy * N + jx] = acc;
            }
        }
    }
    return 0;
}

/* The TV term's sums of squared differences are int arithmetic in the reference; on its x86 build they wrap
 * modulo 2^32 when they pass INT_MAX (a negative sum, then a NaN square root).  The wrap is made explicit. */
static inline int sq3_wrapped(int d1, int d2, int d3) {
    const uint32_t u = (uint32_t)d1 * (uint32_t)d1 + (uint32_t)d2 * (uint32_t)d2 + (uint32_t)d3 * (uint32_t)d3;
    return (int)u;
}

static double tv_term(int a, int ax, int ay, int az, int ax1, int ay1, int az1, int ayx1, int azx1, int axy1, int azy1,
                      int axz1, int ayz1) {
    /* P/IR_ImageCBCT.cpp:9-16 */
    const double d1 = sqrt(double(sq3_wrapped(a - ax, a - ay, a - az))) + 0.00000001;
    const double d2 = sqrt(double(sq3_wrapped(a - ax1, ax1 - ayx1, ax1 - azx1))) + 0.00000001;
    const double d3 = sqrt(double(sq3_wrapped(ay1 - axy1, a - ay1, ay1 - azy1))) + 0.00000001;
    const double d4 = sqrt(double(sq3_wrapped(az1 - axz1, az1 - ayz1, a - az1))) + 0.00000001;
    return ((a - ax) + (a - ay) + (a - az)) / d1 + (a - ax1) / d2 + (a - ay1) / d3 + (a - az1) / d4;
}

int orc_ir_update(int N, int Nz, double scale, const int16_t* before, const float* npf, const float* npf_re,
                  double wbefore, double weight_tv, double weight_prior, float* out) {
    /* P/IR_ImageCBCT.cpp:19-94: x' = x (w + (1 - w) npf_re / (npf + lambda)), lambda = weight_tv TV'(x) + 1e-6
     * (the prior term is 0), then clamped (< 0 -> 0, > 32767 -> 5000) and zeroed outside r >= N/2 (in cm). */
    const double rmax = N * 0.5 * 1.0;
    const size_t NN = (size_t)N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const size_t k = (size_t)z * NN + (size_t)y * N + x;
                double tv = 0.0, prior = 0.0;
                if (z < Nz - 1 && y < N - 1 && x < N - 1 && x > 0 && y > 0 && z > 0) {
                    auto at = [&](int dz, int dy, int dx) {
                        return (int)before[(long long)k + (long long)dz * (long long)NN + (long long)dy * N + dx];
                    };
                    tv = tv_term(at(0, 0, 0), at(0, 0, 1), at(0, 1, 0), at(1, 0, 0), at(0, 0, -1), at(0, -1, 0),
                                 at(-1, 0, 0), at(0, 1, -1), at(1, 0, -1), at(0, -1, 1), at(1, -1, 0), at(-1, 0, 1),
                                 at(-1, 1, 0));
                }
                const double lambda = weight_tv * tv + weight_prior * prior + 0.000001;
                float v = (float)(before[k] * (wbefore + (1 - wbefore) * npf_re[k] / (npf[k] + lambda)));
                if (v < 0.0) v = 0.0f;
                else if (v > 65535 / 2) v = 5000.0f;
                const double xx = -N * scale * 0.5 + (x + 0.5) * scale;
                const double yy = N * scale * 0.5 - (y + 0.5) * scale;
                if (sqrt(xx * xx + yy * yy) >= rmax) v = 0.0f;
                out[k] = v;
            }
    return 0;
}

}  // extern "C"

extern "C" {

int orc_size_class(int n, const double* xshift, char* out) {
    /* F/projectionDataMVkV.cpp:277-281: the mean panel x shift of the used projections */
    double xa = 0.0;
    for (int i = 0; i < n; i++) xa += xshift[i];
    xa = xa / n;
    strcpy(out, xa < 6 ? "S" : (xa > 15 ? "L" : "M"));
    return 0;
}

int orc_angle_mask(int n, const double* angles, int R, double scale, const char* cls, char* mask) {
    /* F/misc.cpp:87-180: for M and L panels, voxel (i, j) keeps angle k when k lies in the window
     * (th1 + 170 deg, th1 + 10 deg) taken through +-180 deg, th1 the voxel's polar angle from +y (clockwise
     * positive); an S panel keeps every angle. */
    const double pi = 3.141592653589793238462643383279;
    const size_t RR = (size_t)R * R;
    if (strcmp(cls, "S") == 0) {
        memset(mask, 1, (size_t)n * RR);
        return 0;
    }
    for (int j = 0; j < R; j++) {
        const double y = -(-R / 2 + j + 0.5) * scale;
        for (int i = 0; i < R; i++) {
            const double x = (-R / 2 + i + 0.5) * scale;
            double th1 = 0.0;
            if (y > 0) th1 = -atan(x / y);
            else if (x < 0 && y < 0) th1 = (-atan(x / y) + pi);
            else if (x > 0 && y < 0) th1 = -(atan(x / y) + pi);
            double ts = th1 + pi - 10 * pi / 180;
            double te = th1 + 10 * pi / 180;
            if (ts < -pi) ts = ts + 2 * pi;
            else if (ts > pi) ts = ts - 2 * pi;
            if (te > pi) te = te - 2 * pi;
            else if (te < -pi) te = te + 2 * pi;
            for (int k = 0; k < n; k++) {
                const double a = angles[k];
                char v;
                if (ts > te) v = ((a > ts && a < pi) || (a > -pi && a < te)) ? 1 : 0;
                else v = (a > ts && a < te) ? 1 : 0;
                mask[(size_t)k * RR + (size_t)j * R + i] = v;
            }
        }
    }
    return 0;
}


/* Offset-detector (half-fan) weighting; no reference counterpart (the reference computes the angle mask
 * above for it and never applies either).  Column col of a projection with panel shift xshift (cm at the
 * detector, shift = xshift / (0.05 * 800 / W) pixels) sees the central ray at c = W/2 + shift - 0.5
 * (F/reconstructImage.cpp:83); t = side * (col - c), side = -1 when shift > 0 (long side towards column 0),
 * else +1; tau = min(taper, W/2 - |shift|) (taper <= 0: W/2 - |shift|).  Returns 2 w(t):
 * 0 for t <= -tau, 2 for t >= tau, 2 sin^2(pi/4 (1 + t/tau)) between.  -1 when the shift leaves no overlap. */
double orc_offset_exponent(int W, double xshift, double taper, int col) {
    const double sh = xshift / (0.05 * 800 / W);
    const double delta = W / 2.0 - fabs(sh);
    if (!(delta > 0.0)) return -1.0;
    const double c = W / 2.0 + sh - 0.5;
    const double tau = taper > 0.0 ? fmin(taper, delta) : delta;
    const double t = (sh > 0.0 ? -1.0 : 1.0) * ((double)col - c);
    if (t <= -tau) return 0.0;
    if (t >= tau) return 2.0;
    const double s = sin(0.78539816339744830962 * (1.0 + t / tau));
    return 2.0 * (s * s);
}

/* Padding of the half-fan panel: P = round(2 max |shift|) columns on the short side (after the real columns
 * when the shifts are positive, before when negative).  Returns W + P, or -1 when the shifts change sign or
 * one leaves no overlap. */
int orc_half_fan_width(int W, int n, const double* xshift) {
    double mx = 0.0;
    int sg = 0;
    for (int k = 0; k < n; k++) {
        const double sh = xshift[k] / (0.05 * 800 / W);
        if (!(W / 2.0 - fabs(sh) > 0.0)) return -1;
        const int s = sh > 0.0 ? 1 : (sh < 0.0 ? -1 : 0);
        if (s != 0 && sg != 0 && s != sg) return -1;
        if (s != 0) sg = s;
        if (fabs(sh) > mx) mx = fabs(sh);
    }
    return W + (int)lround(2.0 * mx);
}

/* p' = p^(2w) per real column (the weight on q = -log p, doubled for the 360-degree normalisation), written
 * into the padded panel out [n][H][Wo] (padding columns 1); xs_out / ys_out: the shifts that keep the central
 * ray's column (Wo/2 + shift' - 0.5 = c + off) and the filter's row centre (-Wo/2 - yshift' = -W/2 - yshift,
 * in pixels) on the padded panel. */
int orc_offset_weight(int n, int H, int W, const float* p, const double* xshift, const double* yshift, double taper,
                      int Wo, float* out, double* xs_out, double* ys_out) {
    if (orc_half_fan_width(W, n, xshift) != Wo) return -1;
    const int P = Wo - W;
    int neg = 0;
    for (int k = 0; k < n; k++)
        if (xshift[k] < 0.0) neg = 1;
    const int off = neg ? P : 0;
    const double pix = 0.05 * 800 / W, pix_o = 0.05 * 800 / Wo;
    for (int k = 0; k < n; k++) {
        xs_out[k] = (xshift[k] / pix + off - P / 2.0) * pix_o;
        ys_out[k] = (yshift[k] / pix - P / 2.0) * pix_o;
        for (int col = 0; col < Wo; col++) {
            const int rc = col - off;
            const double e = (rc >= 0 && rc < W) ? orc_offset_exponent(W, xshift[k], taper, rc) : -2.0;
            for (int j = 0; j < H; j++) {
                float v = 1.0f;
                if (e > -2.0) {
                    const double x = (double)p[((size_t)k * H + j) * W + rc];
                    v = (float)(e == 2.0 ? x * x : (e == 0.0 ? 1.0 : pow(x, e)));
                }
                out[((size_t)k * H + j) * Wo + col] = v;
            }
        }
    }
    return 0;
}


/* ---- Detector model: scatter composition and quantum noise (no reference counterpart) ----------------
 * The reference leaves the stage between the forward projector and FDK to other programs (projection file
 * names in F/txtfile, the scatter U-Net Unet_scatter.py:349-423).  Restated here as the library defines
 * it: lambda = I0 (primary + scatter) expected quanta per pixel; a Poisson count k drawn from a
 * counter-based generator keyed by (seed, flat pixel index), so every pixel's draw is independent of the
 * launch shape; output k (raw) or max(k, 1) / I0 (normalised intensity, zero counts floored at one quantum).
 *
 * Philox4x32-10 (J. Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ten rounds of
 * the two 32 x 32 -> 64 multiplies by 0xD2511F53 and 0xCD9E8D57, key bumped by 0x9E3779B9 / 0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* 53-bit uniform in [0, 1) from two words */
static double orc_u53(uint32_t lo, uint32_t hi) {
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

/* log k! for integral k >= 0: exact factorials below 10, Stirling's series to the 1/x^5 term above (x = k + 1;
 * the omitted terms are below 2e-12 of log k!). */
double orc_log_factorial(double k) {
    if (k < 10.0) {
        static const double f[10] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0};
        return log(f[(int)k]);
    }
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 + r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0)));
}

/* Poisson(lambda) for the pixel with flat index i.  Block b = 0, 1, ... of the pixel's stream is Philox of
 * counter (i_lo, i_hi, b, 0) under key (seed_lo, seed_hi); each block gives two uniforms (words 0-1, 2-3).
 * lambda < 10: inversion by sequential search on the first uniform (capped at k = 256).
 * lambda >= 10: PTRS transformed rejection (W. Hoermann, Insurance: Math. Econ. 12 (1993) 39-45), one block
 * per trial (U from words 0-1, V from words 2-3), at most 64 trials (then floor(lambda + 0.5)); the full test
 * as log(V / alpha / (a/us^2 + b)) <= k log(lambda) - lambda - log k!. */
long long orc_poisson(double lam, uint64_t i, uint64_t seed) {
    if (!(lam > 0.0)) return 0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u}, r[4];
    if (lam < 10.0) {
        orc_philox4x32_10(ctr, key, r);
        const double u = orc_u53(r[0], r[1]);
        double p = exp(-lam), F = p;
        long long k = 0;
        while (u > F && k < 256) {
            k++;
            p = p * lam / (double)k;
            F = F + p;
        }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t t = 0; t < 64; t++) {
        ctr[2] = t;
        orc_philox4x32_10(ctr, key, r);
        const double U = orc_u53(r[0], r[1]) - 0.5;
        const double V = orc_u53(r[2], r[3]);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (long long)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V * invalpha / (a / (us * us) + b)) <= -lam + k * loglam - orc_log_factorial(k))
            return (long long)k;
    }
    return (long long)floor(lam + 0.5);
}

/* out[i] = k (raw != 0) or max(k, 1) / I0, k ~ Poisson(I0 (primary[i] + scatter[i])) (scatter may be NULL). */
int orc_detector_counts(size_t n, const float* primary, const float* scatter, double I0, uint64_t seed, int raw,
                        float* out) {
    for (size_t i = 0; i < n; i++) {
        const double lam = I0 * ((double)primary[i] + (scatter ? (double)scatter[i] : 0.0));
        const long long k = orc_poisson(lam, i, seed);
        out[i] = raw ? (float)k : (float)((double)(k < 1 ? 1 : k) / I0);
    }
    return 0;
}

}  // extern "C"
