// Host-only stand-in for the parts of libvcbct.so that the compat shim (virtualcbct_amd/csrc/compat_shim.cpp)
// calls, for the CPU test of the shim's own logic (buffer sizing, argument checks, the output file).  It keeps
// the C ABI's contract for vcbct_reprojection_angles (VCBCT_EINVAL outside [1, n_angles]) and fills angle a,
// row r, column c of the output with a * 1e6 + r * 1e3 + c, so the test can check what was written where.
// Test infrastructure only: no kernel, no GPU.
#include <cstdio>
#include <cstdlib>

#include "vcbct.h"

struct vcbct_projector {
    int n_angles, H, W;
};

static char g_msg[256] = "";

extern "C" const char* vcbct_last_error(void) { return g_msg; }

extern "C" int vcbct_projector_create(vcbct_projector** out, const vcbct_geometry* g, const double*, const double*,
                                      const double*, const double*, int) {
    *out = new vcbct_projector{g->n_angles, g->proj_height, g->proj_width};
    return 0;
}

extern "C" void vcbct_projector_destroy(vcbct_projector* p) { delete p; }

extern "C" int vcbct_reprojection_angles(vcbct_projector* p, const uint16_t*, const float*, const float*, int, double,
                                         int n_out_angles, float* host_out) {
    if (n_out_angles < 1 || n_out_angles > p->n_angles) {
        snprintf(g_msg, sizeof g_msg, "reprojection: %d angles requested, the projector was created for %d",
                 n_out_angles, p->n_angles);
        return VCBCT_EINVAL;
    }
    for (int a = 0; a < n_out_angles; ++a)
        for (int r = 0; r < p->H; ++r)
            for (int c = 0; c < p->W; ++c) host_out[((size_t)a * p->H + r) * p->W + c] = a * 1e6f + r * 1e3f + c;
    return 0;
}

extern "C" int vcbct_prepare_const_map(int, double, double*) { return 0; }
extern "C" int vcbct_filter_sinogram_host(int, int, int, const float*, float*, const double*, const double*, double) {
    return 0;
}
extern "C" int vcbct_reconstruct_image_host(int, int, int, double, double, double*, int, int, int, const float*, int,
                                            const double*, const double*, double) {
    return 0;
}
