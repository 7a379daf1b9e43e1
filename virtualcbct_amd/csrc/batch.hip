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
struct BatchChoice {
    int nblocks;  // 0: flat grid
    TileShape shape;
};
static std::unordered_map<std::string, BatchChoice> g_batch_choice;
static std::string g_btune_log;
}  // namespace vcbct

using namespace vcbct;

// The grid value of the sequence form: one plain forward launch per problem (each with the tile shape the
// forward autotuner picked for its geometry), in submission order; in a captured graph they replay together.
static constexpr int kSequence = -2;

// The 4-wave block shapes (row groups x columns) of the flat grid's kernel.
static const int kBlockShapes4[3][2] = {{1, 4}, {2, 2}, {4, 1}};

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
    TileShape tile_forced;  // shape fixed by the caller (rgb == 0: autotune)
    TileShape tile;   // the shape in use (after finalize)
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
    b->tile_forced = TileShape{0, 0, 0, 0};
    b->tile = kDefaultTile;
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
    P.xfactor = xfactor;
    P.proj = p;
    // block and work-item offsets are assigned at finalize (problems grouped by voxel type, laid out for the
    // chosen shape); here only the totals are checked against the 32-bit grid and queue ranges, for the largest
    // layout any 4-wave shape gives
    long blocks = 0;
    for (const int* bs : {kBlockShapes4[0], kBlockShapes4[1], kBlockShapes4[2]})
        for (int sp = 0; sp < 2; ++sp) {
            TileMap tm;
            const long nb = tile_map_host(p->g, p->g.nAng, TileShape{bs[0], bs[1], 0, sp}, tm);
            blocks = nb < 0 ? 0x7fffffffL : std::max(blocks, nb);
        }
    const long nrg = (p->g.H + 63) / 64;
    const long items = nrg * std::max((long)p->g.nAng * ((p->g.W + 7) / 8), (long)((p->g.nAng + 7) / 8) * p->g.W);
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
    if (nblocks == kSequence) {  // one plain launch per problem, each with its projector's or geometry's shape
        for (const BatchProblem& P : b->probs) {
            // (the channel volume is rebuilt each time: a replayed graph then projects the volume as it is now)
            const int rc = projector_forward(P.proj, VCBCT_MODE_INTENSITY, P.vox_is_f32 != 0, P.out, nullptr,
                                             P.xfactor, 0, P.g.nAng, st, true);
            if (rc) return rc;
        }
        return 0;
    }
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

// Lays the problems out for shape s (flat grid: each problem's tile map and first block; persistent queues:
// the split's item ranges) and uploads the descriptors and queue totals.
static int batch_layout(vcbct_batch* b, const TileShape& s) {
    b->n_u16 = 0;
    b->flat_u16 = b->flat_f32 = 0;
    memset(b->totals, 0, sizeof b->totals);
    for (BatchProblem& P : b->probs) {
        const int nrg = (P.g.H + 63) / 64;
        int& flat = P.vox_is_f32 ? b->flat_f32 : b->flat_u16;
        const long nb = tile_map_host(P.g, P.g.nAng, s, P.tm);
        if (nb < 0 || (long)flat + nb > 0x7fffffffL) return fail(VCBCT_EINVAL, "batch too large");
        P.block_start = flat;
        flat += (int)nb;
        if (!P.vox_is_f32) ++b->n_u16;
        P.split = s.split;
        P.pcols = s.split ? P.g.W : (P.g.W + 7) / 8;
        P.pang = s.split ? (P.g.nAng + 7) / 8 : P.g.nAng;
        const int items = nrg * P.pang * P.pcols;
        int* tot = b->totals + (P.vox_is_f32 ? 8 : 0);
        for (int q = 0; q < 8; ++q) {
            P.xcd_start[q] = tot[q];
            tot[q] += items;
        }
    }
    const size_t np = b->probs.size();
    if (np) VCBCT_HIP(hipMemcpy(b->d_probs, b->probs.data(), np * sizeof(BatchProblem), hipMemcpyHostToDevice));
    VCBCT_HIP(hipMemcpy(b->d_total, b->totals, 16 * sizeof(int), hipMemcpyHostToDevice));
    b->tile = s;
    return 0;
}

static bool same_shape(const TileShape& x, const TileShape& y) {
    return x.rgb == y.rgb && x.cols == y.cols && x.order == y.order && x.split == y.split;
}

static int batch_finalize(vcbct_batch* b) {
    if (!b->dirty) return 0;
    const size_t np = b->probs.size();
    // u16 problems first, then FP32 maps (stable), each type with its own flat grid; the persistent queues
    // run over all problems in that order
    std::stable_sort(b->probs.begin(), b->probs.end(),
                     [](const BatchProblem& x, const BatchProblem& y) { return x.vox_is_f32 < y.vox_is_f32; });
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
    // Candidates: the grid (flat, or persistent with 2, 4, 6 blocks per CU, one queue per XCD) by the split
    // (columns or angles divided among the XCDs); then, when the flat grid wins, its other 4-wave block
    // shapes and the other order at the best split.  What the caller fixed is not searched.
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    cus = (cus + 7) / 8 * 8;
    const bool grid_fixed = b->forced != -1, tile_fixed = b->tile_forced.rgb > 0;
    char fx[96];
    snprintf(fx, sizeof fx, " grid%d tile%d,%d,%d,%d", b->forced, b->tile_forced.rgb, b->tile_forced.cols,
             b->tile_forced.order, b->tile_forced.split);
    const std::string key = b->signature + fx;
    BatchChoice choice{grid_fixed ? b->forced : 0, tile_fixed ? b->tile_forced : kDefaultTile};
    bool found = false;
    {
        std::lock_guard<std::mutex> lk(g_btune_mu);
        auto it = g_batch_choice.find(key);
        if (it != g_batch_choice.end()) choice = it->second, found = true;
    }
    const char* off = getenv("VCBCT_NO_AUTOTUNE");
    // (a fixed sequence form has no shape of its own to search)
    const bool fixed = (grid_fixed && tile_fixed) || b->forced == kSequence;
    if (!found && np && !fixed && !(off && *off == '1')) {
        hipEvent_t e0, e1;
        VCBCT_HIP(hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return fail(VCBCT_EHIP, "batch autotune: event creation failed");
        }
        float best = 1e30f;
        int rc = 0;
        std::string rep = "batch " + key + ":";
        auto time_one = [&](int nb, const TileShape& s) {
            if (rc) return;
            if (!same_shape(b->tile, s) || b->dirty) {
                rc = batch_layout(b, s);
                b->dirty = false;
                if (rc) return;
            }
            rc = batch_forward(b, nb, nullptr);  // warm
            if (rc) return;
            (void)hipEventRecord(e0, nullptr);
            rc = batch_forward(b, nb, nullptr);
            if (rc) return;
            (void)hipEventRecord(e1, nullptr);
            float ms = 0;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(VCBCT_EHIP, "batch autotune: timing failed");
                return;
            }
            char t[96];
            if (nb == kSequence) snprintf(t, sizeof t, " sequence=%.3fms", ms);
            else if (nb == 0) snprintf(t, sizeof t, " flat/%dx%d/o%d/s%d=%.3fms", s.rgb, s.cols, s.order, s.split, ms);
            else snprintf(t, sizeof t, " persistent %d/s%d=%.3fms", nb, s.split, ms);
            rep += t;
            if (ms < best) {
                best = ms;
                choice = BatchChoice{nb, s};
            }
        };
        std::vector<int> grids;
        if (grid_fixed) grids.push_back(b->forced);
        else grids = {0, 2 * cus, 4 * cus, 6 * cus};
        if (!grid_fixed) time_one(kSequence, kDefaultTile);  // (its shapes are the forward tuner's, per problem)
        for (int sp = 0; sp < 2; ++sp) {
            if (tile_fixed && sp != b->tile_forced.split) continue;
            const TileShape s = tile_fixed ? b->tile_forced : TileShape{1, 4, 0, sp};
            for (int nb : grids) time_one(nb, s);
        }
        if (!tile_fixed && choice.nblocks == 0) {
            const int sp = choice.shape.split;
            for (const int* bs : {kBlockShapes4[1], kBlockShapes4[2]}) time_one(0, TileShape{bs[0], bs[1], 0, sp});
            const TileShape c = choice.shape;
            time_one(0, TileShape{c.rgb, c.cols, 1, sp});
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_btune_mu);
        g_batch_choice[key] = choice;
        char t[96];
        if (choice.nblocks == kSequence)
            snprintf(t, sizeof t, " -> sequence\n");
        else if (choice.nblocks == 0)
            snprintf(t, sizeof t, " -> flat %dx%d/o%d/s%d\n", choice.shape.rgb, choice.shape.cols, choice.shape.order,
                     choice.shape.split);
        else
            snprintf(t, sizeof t, " -> persistent %d blocks, s%d\n", choice.nblocks, choice.shape.split);
        g_btune_log += rep + t;
    }
    int rc = batch_layout(b, choice.shape);
    if (rc) return rc;
    b->nblocks = choice.nblocks;
    b->dirty = false;
    // the sequence form's per-problem shapes are tuned by their first launch; make that happen here, outside
    // any capture, rather than in a captured graph (which would record the default shapes)
    if (b->nblocks == kSequence && np && !(off && *off == '1')) rc = batch_forward(b, kSequence, nullptr);
    return rc;
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

extern "C" int vcbct_batch_set_tile(vcbct_batch* b, int row_groups, int cols, int order, int split) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    if (b->have_graph) return fail(VCBCT_ESTATE, "batch already captured; create a new batch");
    if (row_groups == 0) {
        b->tile_forced = TileShape{0, 0, 0, 0};
    } else {
        const TileShape s{row_groups, cols, order, split};
        if (!tile_shape_valid(s) || row_groups * cols != 4)
            return fail(VCBCT_EINVAL, "batch tile shape %dx%d (order %d, split %d): row groups x columns must be "
                        "4, order and split 0 or 1", row_groups, cols, order, split);
        b->tile_forced = s;
    }
    b->dirty = true;
    return 0;
}

extern "C" int vcbct_batch_get_tile(const vcbct_batch* b, int* shape4) {
    if (!b || !shape4) return fail(VCBCT_EINVAL, "batch_get_tile: NULL argument");
    shape4[0] = b->tile.rgb;
    shape4[1] = b->tile.cols;
    shape4[2] = b->tile.order;
    shape4[3] = b->tile.split;
    return 0;
}

extern "C" int vcbct_batch_set_grid(vcbct_batch* b, int blocks) {
    if (!b) return fail(VCBCT_EINVAL, "batch is NULL");
    if (blocks < kSequence || blocks > (1 << 20)) return fail(VCBCT_EINVAL, "batch grid %d out of range", blocks);
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
