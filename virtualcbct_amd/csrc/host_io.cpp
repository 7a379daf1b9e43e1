// Host-side formats of the virtual-CBCT chain, byte-compatible with the reference: the two input txt
// files, flexmaps, elemental slice stacks, the spectrum/xraydb CSVs and the projector's angle tables.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vcbct.h"

namespace vcbct {
int fail(int code, const char* fmt, ...);
}
using vcbct::fail;

namespace {

const double kPI = 3.141592653589793238462643383279;

// The txt readers of both programs: a block is "two skipped lines, one value line"
// (P/projectionData.cpp:19-59, F/projectionDataMVkV.cpp:19-47).
struct TxtReader {
    FILE* fp;
    char buf[512];
    bool ok = true;
    explicit TxtReader(FILE* f) : fp(f) {}
    void skip(int n) {
        for (int i = 0; i < n; ++i)
            if (!fgets(buf, sizeof buf, fp)) ok = false;
    }
    const char* value() {
        skip(2);
        if (!fgets(buf, sizeof buf, fp)) {
            ok = false;
            buf[0] = 0;
        }
        return buf;
    }
    void str(char* dst, size_t cap) {
        const char* v = value();
        char tmp[512] = "";
        if (sscanf(v, "%511s", tmp) != 1) ok = false;
        snprintf(dst, cap, "%s", tmp);
    }
    void i32(int* d) {
        if (sscanf(value(), "%d", d) != 1) ok = false;
    }
    void hex(int* d) {
        if (sscanf(value(), "%X", d) != 1) { /* "hore" in the vCBCT txt: the field is unused */ }
    }
    void f64(double* d) {
        if (sscanf(value(), "%lf", d) != 1) ok = false;
    }
    void f64x3(double* a, double* b, double* c) {
        if (sscanf(value(), "%lf %lf %lf", a, b, c) != 3) ok = false;
    }
};

}  // namespace

extern "C" int vcbct_read_proj_config(const char* path, vcbct_proj_config* c) {
    if (!path || !c) return fail(VCBCT_EINVAL, "read_proj_config: NULL argument");
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "info file %s not found", path);
    memset(c, 0, sizeof(*c));
    TxtReader r(fp);
    r.skip(6);
    r.str(c->spectrum, sizeof c->spectrum);
    r.i32(&c->total_proj);
    r.i32(&c->proj_width);
    r.i32(&c->proj_height);
    r.str(c->flexmap, sizeof c->flexmap);
    r.str(c->mat_prefix, sizeof c->mat_prefix);
    r.i32(&c->zic);
    r.str(c->outputname, sizeof c->outputname);
    r.f64x3(&c->x_scale, &c->y_scale, &c->z_scale);
    r.i32(&c->recon_size);
    r.i32(&c->recon_slices);
    r.i32(&c->mat_num);
    r.str(c->mat_type, sizeof c->mat_type);
    r.f64(&c->xfactor);
    r.f64x3(&c->x_reg, &c->y_reg, &c->z_reg);
    fclose(fp);
    if (!r.ok) return fail(VCBCT_EIO, "info file %s is malformed", path);
    // P/physParams.h:47-51 (kV, CM geometry), copied in P/projectionData.cpp:95-99
    c->det_pitch_cm = 0.05;
    c->det_pitch_iso_cm = 0.032509752;
    c->sad_cm = 100.0;
    c->det_pixel_num = 800;
    c->sdd_iso_cm = 53.8;
    return 0;
}

extern "C" int vcbct_read_flexmap(const char* path, double* ang, double* xs, double* ys, int cap) {
    if (!path || !ang || !xs || !ys || cap < 0) return fail(VCBCT_EINVAL, "read_flexmap: bad arguments");
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "angle data file %s not found", path);
    double a, x, y;
    char cls[64];
    int n = 0;
    while (fscanf(fp, "%lf %lf %lf %63s", &a, &x, &y, cls) == 4) {
        if (n >= cap) {  // the reference writes past its totalProjNumber-sized arrays here
            fclose(fp);
            return fail(VCBCT_EIO, "flexmap %s has more than %d rows", path, cap);
        }
        ang[n] = a * kPI / 180.0;
        xs[n] = x;
        ys[n] = y;
        ++n;
    }
    fclose(fp);
    return n;
}

extern "C" int vcbct_proj_slice_window(const vcbct_proj_config* c, int* linemin, int* linemax) {
    if (!c || !linemin || !linemax) return fail(VCBCT_EINVAL, "proj_slice_window: NULL argument");
    // P/projectionData.cpp:173-185 (reconSize/2 is an integer division)
    double z_length = (c->sad_cm + c->recon_size / 2 * c->y_scale) / (c->sdd_iso_cm + c->sad_cm) * c->det_pitch_cm *
                      c->det_pixel_num;
    int iz_length = (int)(z_length / c->z_scale);
    int lo = c->zic - (iz_length / 2 + 1);
    int hi = c->zic + (iz_length / 2 + 1);
    if (lo < 0) lo = 0;
    if (hi >= c->recon_slices) hi = c->recon_slices;
    if (hi <= lo) return fail(VCBCT_EINVAL, "empty slice window [%d, %d)", lo, hi);
    *linemin = lo;
    *linemax = hi;
    return 0;
}

extern "C" int vcbct_load_material_slices(const char* prefix, int lo, int hi, int N, int nmat, int policy,
                                          uint16_t* vol, int* n_missing) {
    if (!prefix || !vol || N < 1 || nmat < 1 || hi <= lo) return fail(VCBCT_EINVAL, "load_material_slices: bad arguments");
    const size_t per = (size_t)N * N * nmat;
    std::vector<uint16_t> pixels(per, 0);  // reused across slices, as the reference's buffer
    int missing = 0;
    for (int i = lo; i < hi; ++i) {
        char name[1024];
        snprintf(name, sizeof name, "%s_%03d.raw", prefix, i);
        FILE* fp = fopen(name, "rb");
        size_t got = 0;
        if (fp) {
            got = fread(pixels.data(), 1, per * sizeof(uint16_t), fp);
            fclose(fp);
        }
        if (!fp || got != per * sizeof(uint16_t)) {
            if (policy == 0)
                return fail(VCBCT_EIO, "material slice %s %s", name, fp ? "is truncated" : "not found");
            ++missing;
        }
        memcpy(vol + (size_t)(i - lo) * per, pixels.data(), per * sizeof(uint16_t));
    }
    if (n_missing) *n_missing = missing;
    return 0;
}

extern "C" int vcbct_read_fdk_config(const char* path, vcbct_fdk_config* c) {
    if (!path || !c) return fail(VCBCT_EINVAL, "read_fdk_config: NULL argument");
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "info file %s not found", path);
    memset(c, 0, sizeof(*c));
    TxtReader r(fp);
    r.skip(6);
    r.str(c->kvmv, sizeof c->kvmv);
    r.i32(&c->total_proj);
    r.i32(&c->proj_width);
    r.i32(&c->proj_height);
    r.i32(&c->header_bytes);
    r.str(c->proj_path, sizeof c->proj_path);
    r.i32(&c->num_digits);
    int first = 0;
    if (!strcmp(c->kvmv, "kV")) r.hex(&first);
    else if (!strcmp(c->kvmv, "MV")) r.i32(&first);
    // F/projectionDataMVkV.cpp:73-82: with neither label the first-number block is not consumed
    r.str(c->ext, sizeof c->ext);
    r.str(c->flexmap, sizeof c->flexmap);
    r.i32(&c->recon_size);
    r.str(c->tgt_class, sizeof c->tgt_class);
    r.i32(&c->recon_offset_cm);
    r.f64(&c->recon_step);
    r.i32(&c->recon_slices);
    r.f64(&c->recon_scale);
    r.i32(&c->xfactor);
    r.str(c->output, sizeof c->output);
    fclose(fp);
    if (!r.ok) return fail(VCBCT_EIO, "info file %s is malformed", path);
    c->det_pitch_iso_cm = 0.032509752;  // F/physParams.h:24-25 (kV and MV are equal)
    // F/projectionDataMVkV.cpp:284-302: slice-count correction (in detector rows, as the reference)
    if (c->recon_step > 0 && c->recon_offset_cm + c->recon_step * (c->recon_slices - 1) >= c->proj_height)
        c->recon_slices = (int)((c->proj_height - c->recon_offset_cm) / c->recon_step);
    else if (c->recon_step < 0 && c->recon_offset_cm + c->recon_step * (c->recon_slices - 1) < 0)
        c->recon_slices = (int)((c->recon_offset_cm + 1) / (-c->recon_step));
    return 0;
}

extern "C" int vcbct_read_fdk_angles(const char* path, const char* tgt, int total, double* ang, double* xs, double* ys) {
    if (!path || !tgt || !ang || !xs || !ys || total < 1) return fail(VCBCT_EINVAL, "read_fdk_angles: bad arguments");
    FILE* fp = fopen(path, "r");
    if (!fp) return fail(VCBCT_EIO, "angle data file %s not found", path);
    std::vector<double> pa, px, py;
    std::vector<char> use;
    double a, x, y, cw0 = 0, cw = 0;
    char cls[64];
    int i = 0;
    while (fscanf(fp, "%lf %lf %lf %63s", &a, &x, &y, cls) == 4) {
        ++i;
        if (i == 1) cw0 = a;
        if (i == 30) cw = cw0 - a;
        use.push_back(!strcmp(tgt, "*") || !strcmp(cls, tgt));
        pa.push_back(a);
        px.push_back(x);
        py.push_back(y);
    }
    fclose(fp);
    // Reverse when angle(row 1) - angle(row 30) < 0 (F/projectionDataMVkV.cpp:184-211).  A flexmap of fewer
    // than 30 rows leaves that difference uninitialised in the reference; it is 0 (no reversal) here.
    const bool rev = cw < 0;
    const int nl = (int)pa.size();
    int cnt = 0;
    for (int k = 0; k < nl; ++k) {
        int s = rev ? nl - 1 - k : k;
        if (!use[s]) continue;
        if (cnt >= total) return fail(VCBCT_EIO, "flexmap %s selects more than %d projections", path, total);
        ang[cnt] = pa[s] * kPI / 180.0;
        xs[cnt] = px[s];
        ys[cnt] = py[s];
        ++cnt;
    }
    return cnt;
}

extern "C" int vcbct_trig_table(int W, int H, int n, const double* ang, const double* xs, const double* ys,
                                double pitch, double pixnum, double* c, double* s, double* xc, double* yc) {
    (void)W;
    if (n < 1 || H < 1 || !ang || !xs || !ys || !c || !s || !xc || !yc) return fail(VCBCT_EINVAL, "trig_table: bad arguments");
    // P/misc.cpp:25-34: both centre shifts use the projection HEIGHT
    for (int i = 0; i < n; ++i) {
        c[i] = cos(ang[i]);
        s[i] = sin(ang[i]);
        xc[i] = -H / 2.0 - xs[i] * 10.0 / (pitch * (pixnum / H));
        yc[i] = -H / 2.0 - ys[i] * 10.0 / (pitch * (pixnum / H));
    }
    return 0;
}

// attenuation_coefficient (P/prior_weight_production.cpp:200-523).  The spectrum keeps the reference's
// float arithmetic (bin centre (lo+hi)*0.5 in float, thinning sums in double stored to float, float
// normalisation); the xraydb match |E_db - E| < 30e-9 MeV is found by binary search on the sorted
// database energies instead of a full scan (the last matching row wins, as in the scan).
extern "C" int vcbct_attenuation_table(const char* spectrum, const char* dbdir, int nmat, int nthin, float* atten,
                                       float* pdf, int cap) {
    if (!spectrum || !dbdir || !atten || !pdf || nmat < 1 || nmat > 6 || nthin < 1)
        return fail(VCBCT_EINVAL, "attenuation_table: bad arguments");
    FILE* fp = fopen(spectrum, "r");
    if (!fp) return fail(VCBCT_EIO, "no photon energy file %s", spectrum);
    char row[1024];
    if (!fgets(row, sizeof row, fp)) {
        fclose(fp);
        return fail(VCBCT_EIO, "photon energy file %s is empty", spectrum);
    }
    std::vector<double> energy;
    std::vector<float> raw;
    float d0 = 0, d1 = 0, d2 = 0;
    while (fgets(row, sizeof row, fp)) {
        // unparsable rows repeat the previous row's values, as in the reference (no count check)
        sscanf(row, "%f,%f,%f", &d0, &d1, &d2);
        energy.push_back((d1 + d0) * 0.5);
        raw.push_back(d2);
        if (energy.size() > 1024) {
            fclose(fp);
            return fail(VCBCT_EIO, "photon energy file %s has more than 1024 rows", spectrum);
        }
    }
    fclose(fp);
    const int nraw = (int)energy.size();
    std::vector<float> pthin, ethin;
    double ps = 0.0, es = 0.0, sum_thin = 0.0;
    for (int i = 0; i < nraw; ++i) {
        ps += raw[i];
        es += energy[i];
        if (i - nthin + 1 == (i - nthin + 1) / nthin * nthin) {
            pthin.push_back((float)ps);
            ethin.push_back((float)(es / nthin));
            sum_thin += ps;
            ps = 0.0;
            es = 0.0;
        }
    }
    const int ne = (int)pthin.size();
    if (ne < 1) return fail(VCBCT_EIO, "photon energy file %s has no bins", spectrum);
    if (ne > cap) return fail(VCBCT_EINVAL, "%d energy bins exceed capacity %d", ne, cap);
    const float sum = (float)sum_thin;
    for (int i = 0; i < ne; ++i) pdf[i] = pthin[i] / sum;

    const int ndb = 29001;
    const char* names[3] = {"photo_5keVto150keV.csv", "comp_5keVto150keV.csv", "rayl_5keVto150keV.csv"};
    std::vector<double> edb(ndb), tab[3];
    for (int f = 0; f < 3; ++f) {
        std::string p = std::string(dbdir) + "/" + names[f];
        FILE* fd = fopen(p.c_str(), "rb");
        if (!fd) return fail(VCBCT_EIO, "xraydb file %s not found", p.c_str());
        tab[f].assign((size_t)6 * ndb, 0.0);
        char lb[512];
        if (!fgets(lb, sizeof lb, fd)) {
            fclose(fd);
            return fail(VCBCT_EIO, "xraydb file %s is empty", p.c_str());
        }
        for (int e = 0; e < ndb; ++e) {
            double v[6], en;
            if (!fgets(lb, sizeof lb, fd) ||
                sscanf(lb, "%lf,%lf,%lf,%lf,%lf,%lf,%lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &en) != 7) {
                fclose(fd);
                return fail(VCBCT_EIO, "xraydb file %s: row %d missing or malformed", p.c_str(), e + 2);
            }
            if (f == 0) edb[e] = en / 1000 / 1000;
            for (int k = 0; k < 6; ++k) tab[f][(size_t)k * ndb + e] = v[k];
        }
        fclose(fd);
    }
    std::vector<int> order(ndb);
    for (int e = 0; e < ndb; ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return edb[a] < edb[b]; });
    std::vector<double> esorted(ndb);
    for (int e = 0; e < ndb; ++e) esorted[e] = edb[order[e]];
    // match row per bin (shared by all materials); -1: none, the previous bin's values are kept
    std::vector<int> match(ne, -1);
    for (int i = 0; i < ne; ++i) {
        const double E = (double)ethin[i];
        auto it = std::lower_bound(esorted.begin(), esorted.end(), E - 1e-7);
        int best = -1;
        for (auto k = it - esorted.begin(); k < ndb && esorted[k] <= E + 1e-7; ++k) {
            int e = order[k];
            double aaa = edb[e] * 1000 * 1000 * 1000 - E * 1000 * 1000 * 1000;
            if (-30 < aaa && aaa < 30) best = std::max(best, e);
        }
        match[i] = best;
    }
    double photo = 0, compton = 0, rayleigh = 0;  // kept across bins and materials, as the reference's
    for (int iz = 0; iz < nmat; ++iz) {
        for (int i = 0; i < ne; ++i) {
            int e = match[i];
            if (e >= 0) {
                photo = tab[0][(size_t)iz * ndb + e];
                compton = tab[1][(size_t)iz * ndb + e];
                rayleigh = tab[2][(size_t)iz * ndb + e];
            }
            atten[iz * ne + i] = (float)(photo + compton + rayleigh + 0.0);
        }
    }
    return ne;
}
