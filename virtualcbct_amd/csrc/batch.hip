// Batched launch: many forward problems (phantoms of uneven slice counts, angle counts and spectra) in one
// launch per voxel type, optionally followed by per-problem FDK, captured in a hipGraph.  The grid (flat, or
// persistent with per-XCD work queues and 2/4/6 blocks per CU) is chosen by timing the candidates on the
// first run (autotuner, cached per batch signature) unless the caller fixes it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "vcbct_internal.h"

namespace vcbct {
const char* tune_log();
static std::mutex g_btune_mu;
static std::unordered_map<std::string, int> g_batch_blocks;
static std::string g_btune_log;
}  // namespace vcbct

using namespace vcbct;

struct vcbct_batch {
    int device;
    std::vector<BatchProblem> probs;
    int totals[16];   // persistent queues: items per XCD queue, u16 problems [0, 8), FP32 maps [8, 16)
    BatchProblem* d_probs;
    int* d_total;
    int* d_counters;
    int cap;
    int nblocks;      // persistent grid size, or 0 for the flat grid
    int forced;       // grid fixed by the caller (-1: autotune)
    int flat_blocks;  // running total of flat-grid blocks (range check while problems are added)
    long items_total;  // running total of persistent work items (same check)
    int n_u16;        // after finalize: problems [0, n_u16) have u16 volumes, the rest FP32 maps
    int flat_u16, flat_f32;  // flat-grid blocks of each voxel type
    bool dirty;
    struct Fdk {
        vcbct_fdk_plan* plan;
        const float* proj;
        float* filt;
        float* rec;
        size_t rec_bytes;
    };
    std::vector<Fdk> fdk;
    hipStream_t cap_stream;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool have_graph;
    std::string signature;
};

extern "C" int vcbct_fdk_plan_execute(vcbct_fdk_plan* p, const float* dev_proj, float* dev_filtered, float* dev_recon,
                                      void* stream);

extern "C" int vcbct_batch_create(vcbct_batch** out, int device) {
    if (!out) return fail(VCBCT_EINVAL, "batch_create: out is NULL");
    *out = nullptr;
    VCBCT_HIP(hipSetDevice(device));
    vcbct_batch* b = new (std::nothrow) vcbct_batch();
    if (!b) return fail(VCBCT_ENOMEM, "host allocation failed");
    b->device = device;
    memset(b->totals, 0, sizeof b->totals);
    b->d_probs = nullptr;
    b->d_total = nullptr;
    b->d_counters = nullptr;
    b->cap = 0;
    b->nblocks = 0;
    b->forced = -1;
    b->flat_blocks = 0;
    b->items_total = 0;
    b->n_u16 = 0;
    b->flat_u16 = b->flat_f32 = 0;
    b->dirty = true;
    b->have_graph = false;
    if (hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(VCBCT_EHIP, "stream creation failed");
    }
    *out = b;
    return 0;
}

extern "C" void vcbct_batch_destroy(vcbct_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->have_graph) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
    }
    (void)hipStreamDestroy(b->cap_stream);
    (void)hipFree(b->d_probs);
    (void)hipFree(b->d_total);
    (void)hipFree(b->d_counters);
    delete b;
}

extern "C" int vcbct_batch_add_forward(vcbct_batch* b, vcbct_projector* p, float* dev_out, double xfactor) {
    if (!b || !p || !dev_out) return fail(VCBCT_EINVAL, "batch_add_forward: NULL argument");
    if (p->device != b->device) return fail(VCBCT_EINVAL, "projector lives on another device");
    if (!(p->have_u16 || p->have_f32)) return fail(VCBCT_ESTATE, "projector has no volume");
    if (p->NE < 1) return fail(VCBCT_ESTATE, "projector has no spectrum");
    if (!(xfactor > 0)) return fail(VCBCT_EINVAL, "xfactor must be positive");
    if (b->have_graph) return fail(VCBCT_ESTATE, "batch already captured; create a new batch");
    BatchProblem P;
    memset(&P, 0, sizeof P);
    P.g = p->g;
    P.tab = p->d_tables;
    P.vox_is_f32 = p->have_u16 ? 0 : 1;
    P.vox = P.vox_is_f32 ? (const void*)p->d_vox_f32 : (const void*)p->d_vox_u16;
    P.mu = p->d_mu;
    P.pdf = p->d_pdf;
    P.muf = p->d_muf;
    P.pdff = p->d_pdff;
    P.NE = p->NE;
    P.out = dev_out;
    P.inv_xf = 1.0 / xfactor;
    P.cols_per_xcd = (p->g.W + 7) / 8;
    P.quads_per_xcd = ((p->g.W + 3) / 4 + 7) / 8;
    // block and work-item offsets are assigned at finalize (problems grouped by voxel type); here only the
    // totals are checked against the 32-bit grid and queue ranges
    const long blocks = (long)((p->g.H + 63) / 64) * p->g.nAng * P.quads_per_xcd * 8;
    const long items = (long)((p->g.H + 63) / 64) * p->g.nAng * P.cols_per_xcd;
    if ((long)b->flat_blocks + blocks > 0x7fffffffL || b->items_total + items > 0x7fffffffL - 1024)
        return fail(VCBCT_EINVAL, "batch too large");
    b->flat_blocks += (int)blocks;
    b->items_total += items;
    b->probs.push_back(P);
    char sig[128];
    snprintf(sig, sizeof sig, "[N%d Z%d A%d H%d W%d E%d]", p->g.N, p->g.Nz, p->g.nAng, p->g.H, p->g.W, p->NE);
    b->signature += sig;
    b->dirty = true;
    return 0;
}

extern "C" int vcbct_batch_add_fdk(vcbct_batch* b, vcbct_fdk_plan* plan, const float* dev_proj, float* dev_filtered,
                                   float* dev_recon, long long recon_bytes) {
    if (!b || !plan || !dev_proj || !dev_filtered || !dev_recon || recon_bytes <= 0 || (recon_bytes & 3))
        return fail(VCBCT_EINVAL, "batch_add_fdk: bad arguments");
    if (b->have_graph) return fail(VCBCT_ESTATE, "batch already captured; create a new batch");
    b->fdk.push_back({plan, dev_proj, dev_filtered, dev_recon, (size_t)recon_bytes});
    return 0;
}

static int batch_forward(vcbct_batch* b, int nblocks, hipStream_t st) {
    if (b->probs.empty()) return 0;
    if (nblocks == 0) {  // one flat launch per voxel type present
        const int np = (int)b->probs.size();
        int rc = launch_forward_flat(b->d_probs, b->n_u16, b->flat_u16, 0, st);
        if (!rc) rc = launch_forward_flat(b->d_probs + b->n_u16, np - b->n_u16, b->flat_f32, 1, st);
        return rc;
    }
    const int np = (int)b->probs.size();
    int rc = launch_forward_batch(b->d_probs, b->n_u16, b->d_total, b->d_counters, nblocks, 0, st);
    if (!rc)
        rc = launch_forward_batch(b->d_probs + b->n_u16, np - b->n_u16, b->d_total + 8, b->d_counters + 8, nblocks,
                                  1, st);
    return rc;
}

static int batch_launch(vcbct_batch* b, int nblocks, hipStream_t st) {
    int rc = batch_forward(b, nblocks, st);
    for (size_t i = 0; !rc && i < b->fdk.size(); ++i) {
        rc = launch_zero(b->fdk[i].rec, b->fdk[i].rec_bytes, st);
        if (!rc) rc = vcbct_fdk_plan_execute(b->fdk[i].plan, b->fdk[i].proj, b->fdk[i].filt, b->fdk[i].rec, st);
    }
    return rc;
}

static int batch_finalize(vcbct_batch* b) {
    if (!b->dirty) return 0;
    const size_t np = b->probs.size();
    // u16 problems first, then FP32 maps (stable), each type with its own flat grid; the persistent queues
    // run over all problems in that order
    std::stable_sort(b->probs.begin(), b->probs.end(),
                     [](const BatchProblem& x, const BatchProblem& y) { return x.vox_is_f32 < y.vox_is_f32; });
    b->n_u16 = 0;
    b->flat_u16 = b->flat_f32 = 0;
    memset(b->totals, 0, sizeof b->totals);
    for (BatchProblem& P : b->probs) {
        const int nrg = (P.g.H + 63) / 64;
        int& flat = P.vox_is_f32 ? b->flat_f32 : b->flat_u16;
        P.block_start = flat;
        flat += nrg * P.g.nAng * P.quads_per_xcd * 8;
        if (!P.vox_is_f32) ++b->n_u16;
        const int items = nrg * P.g.nAng * P.cols_per_xcd;
        int* tot = b->totals + (P.vox_is_f32 ? 8 : 0);
        for (int q = 0; q < 8; ++q) {
            P.xcd_start[q] = tot[q];
            tot[q] += items;
        }
    }
    if (np > (size_t)b->cap) {
        (void)hipFree(b->d_probs);
        b->d_probs = nullptr;
        if (dev_malloc((void**)&b->d_probs, np * sizeof(BatchProblem)) != hipSuccess)
            return fail(VCBCT_ENOMEM, "batch descriptors");
        b->cap = (int)np;
    }
    if (!b->d_total && (dev_malloc((void**)&b->d_total, 16 * sizeof(int)) != hipSuccess ||
                        dev_malloc((void**)&b->d_counters, 16 * sizeof(int)) != hipSuccess))
        return fail(VCBCT_ENOMEM, "batch counters");
    if (np) VCBCT_HIP(hipMemcpy(b->d_probs, b->probs.data(), np * sizeof(BatchProblem), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(b->d_total, b->totals, 16 * sizeof(int), hipMemcpyHostToDevice));
    // candidates: the flat grid (0) and persistent grids of 2, 4, 6 blocks per CU (one queue per XCD)
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    cus = (cus + 7) / 8 * 8;
    int chosen = b->forced;
    if (chosen < 0) {
        std::lock_guard<std::mutex> lk(g_btune_mu);
        auto it = g_batch_blocks.find(b->signature);
        if (it != g_batch_blocks.end()) chosen = it->second;
    }
    const char* off = getenv("VCBCT_NO_AUTOTUNE");
    if (chosen < 0 && np && !(off && *off == '1')) {
        hipEvent_t e0, e1;
        VCBCT_HIP(hipEventCreate(&e0));
        VCBCT_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        std::string rep = "batch " + b->signature + ":";
        for (int per : {0, 2, 4, 6}) {
            int nb = cus * per;
            int rc = batch_forward(b, nb, nullptr);
            if (rc) return rc;
            (void)hipEventRecord(e0, nullptr);
            rc = batch_forward(b, nb, nullptr);
            if (rc) return rc;
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            char s[64];
            if (per == 0) snprintf(s, sizeof s, " flat=%.3fms", ms);
            else snprintf(s, sizeof s, " persistent %dx%d=%.3fms", per, cus, ms);
            rep += s;
            if (ms < best) { best = ms; chosen = nb; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> lk(g_btune_mu);
        g_batch_blocks[b->signature] = chosen;
        g_btune_log += rep + " -> " + (chosen ? "persistent " + std::to_string(chosen) + " blocks" : std::string("flat")) +
                       "\n";
    }
    b->nblocks = chosen >= 0 ? chosen : 0;
    b->dirty = false;
    return 0;
}

extern "C" int vcbct_batch_run(vcbct_batch* b, void* stream) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    VCBCT_HIP(hipSetDevice(b->device));
    int rc = batch_finalize(b);
    if (rc) return rc;
    return batch_launch(b, b->nblocks, (hipStream_t)stream);
}

extern "C" int vcbct_batch_capture_graph(vcbct_batch* b) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    VCBCT_HIP(hipSetDevice(b->device));
    int rc = batch_finalize(b);  // uploads and autotuning happen outside the capture
    if (rc) return rc;
    VCBCT_HIP(hipDeviceSynchronize());
    if (b->have_graph) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->have_graph = false;
    }
    VCBCT_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    rc = batch_launch(b, b->nblocks, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(VCBCT_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(VCBCT_EHIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    b->graph = g;
    b->have_graph = true;
    return 0;
}

extern "C" int vcbct_batch_launch_graph(vcbct_batch* b, void* stream) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    if (!b->have_graph) return fail(VCBCT_ESTATE, "no captured graph (vcbct_batch_capture_graph)");
    VCBCT_HIP(hipSetDevice(b->device));
    VCBCT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return 0;
}

extern "C" int vcbct_batch_set_grid(vcbct_batch* b, int blocks) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    if (blocks < -1 || blocks > (1 << 20)) return fail(VCBCT_EINVAL, "batch grid %d out of range", blocks);
    if (b->have_graph) return fail(VCBCT_ESTATE, "batch already captured; create a new batch");
    b->forced = blocks > 0 ? (blocks + 7) / 8 * 8 : blocks;
    b->dirty = true;
    return 0;
}

// Persistent grid size chosen for the batch, 0 when the flat grid was chosen, -1 for NULL.
extern "C" int vcbct_batch_blocks(const vcbct_batch* b) { return b ? b->nblocks : -1; }

extern "C" int vcbct_autotune_log(char* buf, int cap) {
    if (!buf || cap < 1) return fail(VCBCT_EINVAL, "autotune_log: bad buffer");
    std::string s;
    {
        std::lock_guard<std::mutex> lk(g_btune_mu);
        s = g_btune_log;
    }
    s += tune_log();
    s += forward_tune_log();
    snprintf(buf, (size_t)cap, "%s", s.c_str());
    return (int)s.size();
}
