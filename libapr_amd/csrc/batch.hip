/*
 * batch.hip -- conversion of many volumes of uneven sizes as one HIP graph.
 *
 * A plan owns, for every volume, the input image buffer and the output
 * structure (xz_end_vec, and y_vec sized for the bound P <= N), plus a private
 * caching pool for all pipeline scratch.  The whole sequence of conversions is
 * run once eagerly (which fills the pool with every scratch block the graph
 * will address), then captured into a single hipGraph: the pipeline runs in its
 * asynchronous mode (no host synchronisation, no memset nodes, per-volume
 * particle counts, B-spline offsets and overflow counters left in device
 * slots).  aprh_batch_run replays the graph and reads the slots back.
 *
 * Volumes run one after the other inside the graph (they share the context's
 * flag words and, through the pool, scratch blocks), so the graph is a chain.
 */
#include "aprh_internal.hpp"

namespace aprh {
void get_apr_u16_dispatch(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params& par,
                          aprh_apr* a);
}

using namespace aprh;

struct aprh_batch {
    aprh_context* ctx = nullptr;
    int n = 0;
    std::vector<int64_t> dims;                 /* y, x, z per volume        */
    aprh_params par{};
    Pool* pool = nullptr;                      /* scratch + outputs         */
    std::vector<Buf*> inputs;                  /* u16 images                */
    std::vector<aprh_apr*> aprs;               /* results (owned)           */
    Buf slots;                                 /* u64 totals | f32 offsets | int errs[3] */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t voxels = 0;
    ~aprh_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
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

/* every conversion of the plan, in order, on the context's stream */
void record(aprh_batch* b) {
    aprh_context* ctx = b->ctx;
    for (int i = 0; i < b->n; ++i) {
        ctx->slot_total = slot_totals(b) + i;
        ctx->slot_offset = slot_offsets(b) + i;
        ctx->slot_err = slot_errs(b) + 3 * i;
        get_apr_u16_dispatch(ctx, b->inputs[i]->as<uint16_t>(), b->dims[3 * i], b->dims[3 * i + 1], b->dims[3 * i + 2], b->par,
                             b->aprs[i]);
    }
}

/* run `f` with the context switched to the plan's pool and asynchronous mode */
template <typename F>
void in_plan(aprh_batch* b, F f) {
    aprh_context* ctx = b->ctx;
    Pool* saved = ctx->pool;
    const int saved_timing = ctx->timing, saved_keep = ctx->keep_intermediates;
    ctx->pool = b->pool;
    ctx->async_mode = 1;
    ctx->timing = 0;
    ctx->keep_intermediates = 0;
    try {
        f();
    } catch (...) {
        ctx->pool = saved;
        ctx->async_mode = 0;
        ctx->timing = saved_timing;
        ctx->keep_intermediates = saved_keep;
        ctx->slot_total = nullptr;
        ctx->slot_offset = nullptr;
        ctx->slot_err = nullptr;
        throw;
    }
    ctx->pool = saved;
    ctx->async_mode = 0;
    ctx->timing = saved_timing;
    ctx->keep_intermediates = saved_keep;
    ctx->slot_total = nullptr;
    ctx->slot_offset = nullptr;
    ctx->slot_err = nullptr;
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
        b->slots.reset(b->pool, (size_t)n * (8 + 4 + 12));
        for (int i = 0; i < n; ++i) {
            const int64_t y = dims3[3 * i], x = dims3[3 * i + 1], z = dims3[3 * i + 2];
            if (y < 1 || x < 1 || z < 1) throw Error(APRH_ERR_BAD_DIMS, "dimensions must be positive");
            if (y > 65535 || x > 65535 || z > 65535) throw Error(APRH_ERR_UNSUPPORTED, "dimensions above 65535");
            const size_t nv = (size_t)(y * x * z);
            b->voxels += nv;
            b->inputs.push_back(new Buf(b->pool, nv * 2));
            aprh_apr* a = new aprh_apr();
            a->ctx = ctx;
            a->borrowed = true;
            geninfo_init(*a, y, x, z);
            if (a->l_max <= 2) {
                delete a;
                throw Error(APRH_ERR_UNSUPPORTED, "batched conversion needs l_max >= 3 (a side above 4 pixels)");
            }
            a->xz_end.reset(b->pool, a->n_xz * 8);
            a->y_vec.reset(b->pool, nv * 2);   /* P <= N */
            b->aprs.push_back(a);
        }
        launch_zero(ctx, b->slots.ptr(), b->slots.bytes());
        for (auto* in : b->inputs) launch_zero(ctx, in->ptr(), in->bytes());
        /* eager pass: fills the pool with every scratch block the capture will reuse */
        in_plan(b, [&] { record(b); });
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        /* capture */
        in_plan(b, [&] {
            HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            try {
                record(b);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(ctx->stream, &b->graph));
        });
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

int aprh_batch_destroy(aprh_batch* b) {
    APRH_TRY
    if (!b) return APRH_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    APRH_CATCH
}

}  // extern "C"
