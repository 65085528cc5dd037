/*
 * batch.hip -- conversion of many volumes of uneven sizes: one persistent
 * launch per stage over all volumes, the whole sequence in one HIP graph.
 *
 * Plan creation
 *  1. records the u16 pipeline of every volume (the same code as a single
 *     get_apr, with the context in recording mode): each stage launch becomes
 *     (body, arguments, tiles) in the volume's list; buffers come from the
 *     plan's pool in keep mode, so every volume owns all of its scratch;
 *  2. merges the lists stage by stage: the next merged launch takes the body
 *     that heads the most lists and the volumes whose head it is, so every
 *     volume's order is kept and volumes with fewer levels simply sit out the
 *     extra pulling-scheme launches;
 *  3. autotunes each merged launch: the whole sequence runs once per
 *     candidate (tile-shape alternative of the stages that have them -- the
 *     x/z smoothing's checkpoint spacing, the running means' look-ahead, the
 *     pulling scheme's z run, the row fill's lanes per work item -- times
 *     the grid: one block per tile, or 1, 2, 4, 8, 16 blocks per CU walking
 *     the tiles persistently), each launch timed with events, and the
 *     fastest shape and grid of each launch kept;
 *  4. captures the tuned sequence into one graph.
 * aprh_batch_run replays the graph and reads each volume's particle count,
 * B-spline offset and overflow flags from device slots.
 */
#include <algorithm>
#include <map>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {
void get_apr_u16_dispatch(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params& par,
                          aprh_apr* a);
}

using namespace aprh;

namespace {

/* the merged launch of one tile-shape alternative: arguments and tile prefix of every volume */
struct GroupAlt {
    const TileAlt* proto = nullptr;           /* body, block size, launcher, name, shape */
    int64_t tiles = 0;
    Buf d_args, d_prefix;
};

struct Group {
    std::vector<const TileRec*> recs;         /* one per volume, in launch order  */
    std::vector<GroupAlt> alts;               /* the stage's tile-shape alternatives */
    int alt = 0;                              /* the one launched                */
    unsigned grid = 0;
    float ms = 0.0f;
    const GroupAlt& cur() const { return alts[(size_t)alt]; }
};

constexpr int N_CAND = 6;   /* one block per tile; 1, 2, 4, 8, 16 blocks per CU */

}  // namespace

struct aprh_batch {
    aprh_context* ctx = nullptr;
    int n = 0;
    std::vector<int64_t> dims;                 /* y, x, z per volume        */
    aprh_params par{};
    Pool* pool = nullptr;                      /* outputs + all scratch (keep mode) */
    std::vector<Buf*> inputs;                  /* u16 images                */
    std::vector<aprh_apr*> aprs;               /* results (owned)           */
    Buf slots;                                 /* u64 totals | f32 offsets | int errs[3] */
    std::vector<TileRecorder> recs;
    std::vector<Group> groups;
    int cus = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t voxels = 0;
    ~aprh_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        groups.clear();
        for (auto* a : aprs) delete a;
        for (auto* b : inputs) delete b;
        slots.release();
        delete pool;
    }
};

namespace {

uint64_t* slot_totals(aprh_batch* b) { return b->slots.as<uint64_t>(); }
float* slot_offsets(aprh_batch* b) { return (float*)(b->slots.as<uint64_t>() + b->n); }
int* slot_errs(aprh_batch* b) { return (int*)(slot_offsets(b) + b->n); }

/* step 1: every volume's stage launches, recorded with the plan's pool in keep mode */
void record(aprh_batch* b) {
    aprh_context* ctx = b->ctx;
    Pool* saved = ctx->pool;
    const int saved_timing = ctx->timing, saved_keep = ctx->keep_intermediates;
    auto restore = [&] {
        ctx->pool = saved;
        ctx->async_mode = 0;
        ctx->timing = saved_timing;
        ctx->keep_intermediates = saved_keep;
        ctx->recorder = nullptr;
        ctx->slot_total = nullptr;
        ctx->slot_offset = nullptr;
        ctx->slot_err = nullptr;
    };
    ctx->pool = b->pool;
    ctx->async_mode = 1;
    ctx->timing = 0;
    ctx->keep_intermediates = 0;
    b->recs.assign(b->n, TileRecorder());
    try {
        for (int i = 0; i < b->n; ++i) {
            ctx->recorder = &b->recs[i];
            ctx->slot_total = slot_totals(b) + i;
            ctx->slot_offset = slot_offsets(b) + i;
            ctx->slot_err = slot_errs(b) + 3 * i;
            get_apr_u16_dispatch(ctx, b->inputs[i]->as<uint16_t>(), b->dims[3 * i], b->dims[3 * i + 1], b->dims[3 * i + 2], b->par,
                                 b->aprs[i]);
        }
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

/* step 2: merge the per-volume lists into launches */
void merge(aprh_batch* b) {
    std::vector<size_t> head(b->n, 0);
    for (;;) {
        std::map<const std::type_info*, int> count;
        const std::type_info* best = nullptr;
        int best_n = 0;
        for (int v = 0; v < b->n; ++v) {
            if (head[v] >= b->recs[v].recs.size()) continue;
            const std::type_info* t = b->recs[v].recs[head[v]].body;
            const int c = ++count[t];
            if (c > best_n) {
                best_n = c;
                best = t;
            }
        }
        if (!best) break;
        Group g;
        for (int v = 0; v < b->n; ++v) {
            if (head[v] >= b->recs[v].recs.size()) continue;
            const TileRec& r = b->recs[v].recs[head[v]];
            if (r.body != best) continue;
            g.recs.push_back(&r);
            ++head[v];
        }
        /* the alternatives every volume of the launch has (the recording passes them in a fixed
           order per stage, so the lists agree) */
        size_t na = g.recs.front()->alts.size();
        for (const TileRec* r : g.recs) na = std::min(na, r->alts.size());
        for (size_t k = 0; k < na; ++k) {
            bool same = true;
            for (const TileRec* r : g.recs) same = same && r->alts[k].body == g.recs.front()->alts[k].body;
            if (!same) {
                na = k;
                break;
            }
        }
        g.alts.resize(std::max<size_t>(na, 1));
        g.alt = std::min<int>(g.recs.front()->prefer, (int)g.alts.size() - 1);
        b->groups.push_back(std::move(g));
    }
    /* arguments and tile prefixes of each launch and alternative, uploaded once */
    for (auto& g : b->groups) {
        for (size_t k = 0; k < g.alts.size(); ++k) {
            GroupAlt& ga = g.alts[k];
            ga.proto = &g.recs.front()->alts[k];
            const size_t as = ga.proto->args.size();
            std::vector<uint8_t> args(as * g.recs.size());
            std::vector<int64_t> prefix(g.recs.size() + 1, 0);
            for (size_t v = 0; v < g.recs.size(); ++v) {
                std::memcpy(args.data() + v * as, g.recs[v]->alts[k].args.data(), as);
                prefix[v + 1] = prefix[v] + g.recs[v]->alts[k].tiles;
            }
            ga.tiles = prefix.back();
            ga.d_args.reset(b->pool, args.size());
            ga.d_prefix.reset(b->pool, prefix.size() * 8);
            HIP_CHECK(hipMemcpy(ga.d_args.ptr(), args.data(), args.size(), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(ga.d_prefix.ptr(), prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice));
        }
    }
}

unsigned candidate_grid(const aprh_batch* b, const GroupAlt& ga, int c) {
    const int64_t cap = (int64_t)(0xFFFFFFFFu / (unsigned)ga.proto->threads);
    const int64_t want = (c == 0) ? ga.tiles : (int64_t)b->cus * (1 << (c - 1));
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>({want, ga.tiles, cap}));
}

void launch_all(aprh_batch* b, const std::vector<hipEvent_t>* ev) {
    hipStream_t st = b->ctx->stream;
    for (size_t k = 0; k < b->groups.size(); ++k) {
        const Group& g = b->groups[k];
        const GroupAlt& ga = g.cur();
        if (ev) HIP_CHECK(hipEventRecord((*ev)[k], st));
        ga.proto->multi(st, ga.d_args.ptr(), ga.d_prefix.as<int64_t>(), (int)g.recs.size(), g.grid);
    }
    if (ev) HIP_CHECK(hipEventRecord(ev->back(), st));
}

/* step 3: per-launch tile shape and grid from timed runs of the whole sequence (each run
   recomputes every stage from the inputs with alternatives that give the same results, so the
   runs leave valid results behind) */
void autotune(aprh_batch* b) {
    hipStream_t st = b->ctx->stream;
    std::vector<hipEvent_t> ev(b->groups.size() + 1);
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    size_t max_alts = 1;
    for (const auto& g : b->groups) max_alts = std::max(max_alts, g.alts.size());
    /* t[launch][alternative][grid candidate] */
    std::vector<std::vector<std::vector<float>>> t(b->groups.size());
    for (size_t k = 0; k < b->groups.size(); ++k)
        t[k].assign(b->groups[k].alts.size(), std::vector<float>(N_CAND, 1e30f));
    try {
        for (size_t a = 0; a < max_alts; ++a)
            for (int c = -1; c < N_CAND; ++c) {   /* c = -1: warm-up */
                for (auto& g : b->groups) {
                    g.alt = (int)std::min(a, g.alts.size() - 1);
                    g.grid = candidate_grid(b, g.cur(), std::max(c, 0));
                }
                launch_all(b, &ev);
                HIP_CHECK(hipStreamSynchronize(st));
                if (c < 0) continue;
                for (size_t k = 0; k < b->groups.size(); ++k)
                    if (a < b->groups[k].alts.size()) HIP_CHECK(hipEventElapsedTime(&t[k][a][c], ev[k], ev[k + 1]));
            }
    } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (size_t k = 0; k < b->groups.size(); ++k) {
        Group& g = b->groups[k];
        int ba = 0, bc = 0;
        for (size_t a = 0; a < g.alts.size(); ++a)
            for (int c = 0; c < N_CAND; ++c)
                if (t[k][a][c] < t[k][ba][bc]) {
                    ba = (int)a;
                    bc = c;
                }
        g.alt = ba;
        g.grid = candidate_grid(b, g.cur(), bc);
        g.ms = t[k][ba][bc];
    }
}

void read_slots(aprh_batch* b) {
    aprh_context* ctx = b->ctx;
    std::vector<uint64_t> tot(b->n);
    std::vector<float> off(b->n);
    std::vector<int> err(3 * (size_t)b->n);
    HIP_CHECK(hipMemcpyAsync(tot.data(), slot_totals(b), 8 * (size_t)b->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(off.data(), slot_offsets(b), 4 * (size_t)b->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(err.data(), slot_errs(b), 12 * (size_t)b->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < b->n; ++i) {
        b->aprs[i]->total = tot[i];
        b->aprs[i]->bspline_offset = off[i];
        b->aprs[i]->tree_built = false;
        if (err[3 * i] || err[3 * i + 1] || err[3 * i + 2])
            throw Error(APRH_ERR_OVERFLOW, "integer under-/overflow in the B-spline smoothing of batch volume " + std::to_string(i));
    }
}

}  // namespace

extern "C" {

int aprh_batch_create(aprh_context* ctx, int32_t n, const int64_t* dims3, const aprh_params* par, aprh_batch** out) {
    APRH_TRY
    check_ctx(ctx);
    if (n < 1 || !dims3 || !par || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad batch arguments");
    aprh_batch* b = new aprh_batch();
    try {
        b->ctx = ctx;
        b->n = n;
        b->dims.assign(dims3, dims3 + 3 * (size_t)n);
        b->par = *par;
        b->pool = new Pool(ctx->device);
        b->pool->set_keep(true);
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cus > 0) b->cus = cus;
        b->slots.reset(b->pool, (size_t)n * (8 + 4 + 12));
        for (int i = 0; i < n; ++i) {
            const int64_t y = dims3[3 * i], x = dims3[3 * i + 1], z = dims3[3 * i + 2];
            if (y < 1 || x < 1 || z < 1) throw Error(APRH_ERR_BAD_DIMS, "dimensions must be positive");
            if (y > 65535 || x > 65535 || z > 65535) throw Error(APRH_ERR_UNSUPPORTED, "dimensions above 65535");
            aprh_apr* a = new aprh_apr();
            a->ctx = ctx;
            a->borrowed = true;
            geninfo_init(*a, y, x, z);
            if (a->l_max <= 2) {
                delete a;
                throw Error(APRH_ERR_UNSUPPORTED, "batched conversion needs l_max >= 3 (a side above 4 pixels)");
            }
            const size_t nv = (size_t)(y * x * z);
            b->voxels += nv;
            b->inputs.push_back(new Buf(b->pool, nv * 2));
            a->xz_end.reset(b->pool, a->n_xz * 8);
            a->y_vec.reset(b->pool, nv * 2);   /* P <= N */
            b->aprs.push_back(a);
        }
        launch_zero(ctx, b->slots.ptr(), b->slots.bytes());
        for (auto* in : b->inputs) launch_zero(ctx, in->ptr(), in->bytes());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        record(b);
        merge(b);
        autotune(b);
        HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        try {
            launch_all(b, nullptr);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(ctx->stream, &b->graph));
        HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        delete b;
        throw;
    }
    *out = b;
    APRH_CATCH
}

int aprh_batch_input(aprh_batch* b, int32_t i, uint16_t** dev_ptr) {
    APRH_TRY
    if (!b || !dev_ptr || i < 0 || i >= b->n) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad batch index");
    *dev_ptr = b->inputs[(size_t)i]->as<uint16_t>();
    APRH_CATCH
}

int aprh_batch_run(aprh_context* ctx, aprh_batch* b) {
    APRH_TRY
    check_ctx(ctx);
    if (!b || b->ctx != ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "batch belongs to another context");
    HIP_CHECK(hipGraphLaunch(b->exec, ctx->stream));
    read_slots(b);
    APRH_CATCH
}

int aprh_batch_get(aprh_batch* b, int32_t i, aprh_apr** apr) {
    APRH_TRY
    if (!b || !apr || i < 0 || i >= b->n) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad batch index");
    *apr = b->aprs[(size_t)i];
    APRH_CATCH
}

int aprh_batch_stage_count(const aprh_batch* b) { return b ? (int)b->groups.size() : 0; }

int aprh_batch_stage_info(const aprh_batch* b, int32_t i, aprh_batch_stage* info) {
    APRH_TRY
    if (!b || !info || i < 0 || i >= (int)b->groups.size()) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad stage index");
    const Group& g = b->groups[(size_t)i];
    const GroupAlt& ga = g.cur();
    info->name = ga.proto->name;
    info->volumes = (int32_t)g.recs.size();
    info->tiles = ga.tiles;
    info->grid = g.grid;
    info->threads = (uint32_t)ga.proto->threads;
    info->tuned_ms = g.ms;
    info->shape = ga.proto->shape;
    info->n_shapes = (int32_t)g.alts.size();
    APRH_CATCH
}

int aprh_batch_destroy(aprh_batch* b) {
    APRH_TRY
    if (!b) return APRH_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    APRH_CATCH
}

}  // extern "C"
