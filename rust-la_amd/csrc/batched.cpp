// Batched uneven solvers: plans, graph capture and the tile autotuner.
//
// A plan sorts the problems by order (largest first, so the persistent
// workgroups finish together), uploads the order and sizes once, owns the
// device queue counters, and records the solver launch in a hipGraph the
// first time it runs with a given set of buffers; later calls with the same
// buffers replay the graph.  The launch is a single persistent kernel; the
// queue is reset by the kernel itself, so the graph has no memset node.
#include "la_common.hpp"
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace lahip {

// kernels (batched_chol.hip / batched_lu.hip)
int launch_chol_batched(la_ctx *ctx, hipStream_t s, int grid, int count, const int32_t *order,
                        const int32_t *nvec, const int64_t *offA, const int64_t *offB,
                        const double *A, double *L, const double *B, double *X, int32_t *status,
                        uint32_t *queue, const int cfg[4]);
int chol_batched_grid(la_ctx *ctx);
int fill_spd_batch(la_ctx *ctx, int count, const int32_t *nvec, const int64_t *offA, double *A);
int lu_ir_batched_grid(la_ctx *ctx, int prec);
int launch_lu_ir_batched(la_ctx *ctx, hipStream_t s, int prec, int grid, int count,
                         const int32_t *order, const int32_t *nvec, const int32_t *rhsvec,
                         const int64_t *offA, const int64_t *offB, const int64_t *offW,
                         const int64_t *offN, const double *A, const double *B, double *X,
                         void *Wbuf, int32_t *permbuf, double tol, int max_iter, int32_t *status,
                         int32_t *iters, double *berr, uint32_t *queue);

}  // namespace lahip

using namespace lahip;

static constexpr int kCholDefault = 16 | (2 << 8);

struct la_batch_plan {
    la_ctx *ctx = nullptr;
    int kind = 0;
    int64_t count = 0;
    int prec = 0;
    int32_t *d_order = nullptr, *d_n = nullptr, *d_nrhs = nullptr;
    uint32_t *d_queue = nullptr;
    // LU-IR plans: working copies W_i (offsets d_offW, elements of the
    // working precision) and pivot orders (offsets d_offN)
    int64_t *d_offW = nullptr, *d_offN = nullptr;
    void *d_W = nullptr;
    int32_t *d_perm = nullptr;
    // Cholesky tile configuration per size bucket: panel width | row tiles << 8
    int nb[4] = {kCholDefault, kCholDefault, kCholDefault, kCholDefault};
    int grid = 0;
    std::string desc;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *key[12] = {nullptr};
};

static const int kCholCandidates[] = {8 | (1 << 8), 8 | (2 << 8), 16 | (1 << 8), 16 | (2 << 8)};

static void describe(la_batch_plan *p) {
    char buf[256];
    if (p->kind == 1) {
        static const char *pn[] = {"fp64", "fp32", "bf16", "fp16"};
        snprintf(buf, sizeof buf,
                 "kind=lu-ir count=%lld grid=%d working precision=%s panel width=%d%s",
                 (long long)p->count, p->grid, pn[p->prec], p->prec == LA_PREC_FP64 ? 16 : 32,
                 p->ctx->lu_tuned ? " (grid autotuned)" : " (grid at the occupancy limit)");
        p->desc = buf;
        return;
    }
    snprintf(buf, sizeof buf,
             "kind=cholesky count=%lld grid=%d (panel width, row tiles) by bucket [<=32, <=64, "
             "<=128, <=256] = (%d,%d) (%d,%d) (%d,%d) (%d,%d)%s",
             (long long)p->count, p->grid, p->nb[0] & 255, p->nb[0] >> 8, p->nb[1] & 255,
             p->nb[1] >> 8, p->nb[2] & 255, p->nb[2] >> 8, p->nb[3] & 255, p->nb[3] >> 8,
             p->ctx->chol_tuned ? " (autotuned)" : " (defaults)");
    p->desc = buf;
}

static int create_lu_ir_plan(la_ctx *ctx, int64_t count, const int32_t *h_n,
                             const int32_t *h_nrhs, int prec, la_batch_plan **out)
{
    *out = nullptr;
    if (!h_nrhs) return set_error(ctx, LA_ERR_ARG, "batched LU-IR: nrhs array required");
    if (prec != LA_PREC_FP64 && prec != LA_PREC_FP32 && prec != LA_PREC_FP16)
        return set_error(ctx, LA_ERR_ARG, "batched LU-IR: precision must be FP64, FP32 or FP16");
    for (int64_t i = 0; i < count; ++i) {
        if (h_n[i] <= 0 || h_n[i] > 1024)
            return set_error(ctx, LA_ERR_ARG, "batched LU-IR: order must be in [1, 1024]");
        if (h_nrhs[i] <= 0 || h_nrhs[i] > 8)
            return set_error(ctx, LA_ERR_ARG, "batched LU-IR: nrhs must be in [1, 8]");
    }
    la_batch_plan *p = new la_batch_plan();
    p->ctx = ctx;
    p->kind = 1;
    p->count = count;
    p->prec = prec;
    std::vector<int32_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return h_n[a] > h_n[b]; });
    std::vector<int64_t> offW(count), offN(count);
    int64_t w = 0, q = 0;
    for (int64_t i = 0; i < count; ++i) {
        offW[i] = w;
        offN[i] = q;
        w += (int64_t)h_n[i] * h_n[i];
        q += h_n[i];
    }
    const size_t esz = prec == LA_PREC_FP64 ? 8 : 4;
    auto fail = [&](hipError_t e, const char *what) {
        la_batch_plan_destroy(p);
        return hip_fail(ctx, e, what);
    };
    hipError_t e;
    if ((e = hipMalloc(&p->d_order, count * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_n, count * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_nrhs, count * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_offW, count * 8)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_offN, count * 8)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_W, (size_t)w * esz)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_perm, (size_t)q * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_queue, 256)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMemcpy(p->d_order, order.data(), count * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_n, h_n, count * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_nrhs, h_nrhs, count * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_offW, offW.data(), count * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_offN, offN.data(), count * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "plan upload");
    if ((e = hipMemset(p->d_queue, 0, 256)) != hipSuccess) return fail(e, "plan queue");
    if ((e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "plan stream");
    p->grid = lu_ir_batched_grid(ctx, prec);
    if (ctx->lu_tuned && ctx->lu_wg_per_cu[prec] > 0)
        p->grid = std::min(p->grid, std::max(1, ctx->num_cus) * ctx->lu_wg_per_cu[prec]);
    describe(p);
    *out = p;
    return LA_OK;
}

extern "C" int la_batch_plan_create(la_ctx *ctx, int kind, int64_t count, const int32_t *h_n,
                                    const int32_t *h_nrhs, int prec, la_batch_plan **out)
{
    if (!ctx || !out || !h_n || count <= 0 || count > INT32_MAX) return LA_ERR_ARG;
    if (kind == 1) return create_lu_ir_plan(ctx, count, h_n, h_nrhs, prec, out);
    if (kind != 0) return set_error(ctx, LA_ERR_ARG, "batch plan: kind must be 0 or 1");
    *out = nullptr;
    for (int64_t i = 0; i < count; ++i)
        if (h_n[i] <= 0 || h_n[i] > 256)
            return set_error(ctx, LA_ERR_ARG, "batched Cholesky: order must be in [1, 256]");
    if (h_nrhs)
        for (int64_t i = 0; i < count; ++i)
            if (h_nrhs[i] != 1) return set_error(ctx, LA_ERR_ARG, "batched Cholesky: nrhs must be 1");
    la_batch_plan *p = new la_batch_plan();
    p->ctx = ctx;
    p->kind = kind;
    p->count = count;
    p->prec = prec;
    std::vector<int32_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return h_n[a] > h_n[b]; });
    auto fail = [&](hipError_t e, const char *w) {
        la_batch_plan_destroy(p);
        return hip_fail(ctx, e, w);
    };
    hipError_t e;
    if ((e = hipMalloc(&p->d_order, count * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_n, count * 4)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMalloc(&p->d_queue, 256)) != hipSuccess) return fail(e, "plan alloc");
    if ((e = hipMemcpy(p->d_order, order.data(), count * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "plan upload");
    if ((e = hipMemcpy(p->d_n, h_n, count * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "plan upload");
    if ((e = hipMemset(p->d_queue, 0, 256)) != hipSuccess) return fail(e, "plan queue");
    if ((e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "plan stream");
    if (ctx->chol_tuned) memcpy(p->nb, ctx->chol_nb, sizeof p->nb);
    p->grid = chol_batched_grid(ctx);
    describe(p);
    *out = p;
    return LA_OK;
}

extern "C" void la_batch_plan_destroy(la_batch_plan *p)
{
    if (!p) return;
    // a replay may still be in flight on the context stream
    if (p->ctx && p->ctx->stream) (void)hipStreamSynchronize(p->ctx->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_order) (void)hipFree(p->d_order);
    if (p->d_n) (void)hipFree(p->d_n);
    if (p->d_nrhs) (void)hipFree(p->d_nrhs);
    if (p->d_queue) (void)hipFree(p->d_queue);
    if (p->d_offW) (void)hipFree(p->d_offW);
    if (p->d_offN) (void)hipFree(p->d_offN);
    if (p->d_W) (void)hipFree(p->d_W);
    if (p->d_perm) (void)hipFree(p->d_perm);
    delete p;
}

extern "C" const char *la_batch_plan_describe(la_batch_plan *p)
{
    return p ? p->desc.c_str() : "null plan";
}

extern "C" int la_dpotrf_potrs_batched_dev(la_ctx *ctx, la_batch_plan *p, const int64_t *d_offA,
                                           const int64_t *d_offB, const double *A, double *L,
                                           const double *B, double *X, int32_t *d_status)
{
    if (!ctx || !p || !d_offA || !d_offB || !A || !L || !B || !X || !d_status) return LA_ERR_ARG;
    if (p->kind != 0 || p->ctx != ctx) return LA_ERR_ARG;
    const void *key[12] = {d_offA, d_offB, A, L, B, X, d_status, nullptr};
    const bool same = p->exec && memcmp(key, p->key, sizeof key) == 0;
    ProfScope ps(ctx, LA_K_BATCH, 0.0, 0.0);
    if (!same) {
        if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        LA_HIP_CHECK(ctx, hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        int st = launch_chol_batched(ctx, p->cap_stream, p->grid, (int)p->count, p->d_order,
                                     p->d_n, d_offA, d_offB, A, L, B, X, d_status, p->d_queue,
                                     p->nb);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        p->graph = g;
        LA_HIP_CHECK(ctx, hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
        memcpy(p->key, key, sizeof key);
    }
    LA_HIP_CHECK(ctx, hipGraphLaunch(p->exec, ctx->stream));
    return LA_OK;
}

extern "C" int la_dgesv_ir_batched_dev(la_ctx *ctx, la_batch_plan *p, const int64_t *d_offA,
                                       const int64_t *d_offB, const double *A, const double *B,
                                       double *X, double tol, int max_iter, int32_t *d_status,
                                       int32_t *d_iters, double *d_berr)
{
    if (!ctx || !p || !d_offA || !d_offB || !A || !B || !X || !d_status || !d_iters || !d_berr)
        return LA_ERR_ARG;
    if (p->kind != 1 || p->ctx != ctx || max_iter < 0 || !(tol >= 0.0)) return LA_ERR_ARG;
    union { double d; const void *v; } tkey;
    tkey.v = nullptr;
    tkey.d = tol;
    const void *key[12] = {d_offA, d_offB, A, B, X, d_status, d_iters, d_berr, tkey.v,
                           (const void *)(intptr_t)max_iter, nullptr, nullptr};
    const bool same = p->exec && memcmp(key, p->key, sizeof key) == 0;
    ProfScope ps(ctx, LA_K_BATCH, 0.0, 0.0);
    if (!same) {
        if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        LA_HIP_CHECK(ctx, hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        int st = launch_lu_ir_batched(ctx, p->cap_stream, p->prec, p->grid, (int)p->count,
                                      p->d_order, p->d_n, p->d_nrhs, d_offA, d_offB, p->d_offW,
                                      p->d_offN, A, B, X, p->d_W, p->d_perm, tol, max_iter,
                                      d_status, d_iters, d_berr, p->d_queue);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        p->graph = g;
        LA_HIP_CHECK(ctx, hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
        memcpy(p->key, key, sizeof key);
    }
    LA_HIP_CHECK(ctx, hipGraphLaunch(p->exec, ctx->stream));
    return LA_OK;
}

// ---------------------------------------------------------------- autotune
// Kind 1 (batched LU + refinement): for each working precision, time the
// persistent kernel at 1, 2 and the occupancy limit of workgroups per CU on
// a synthetic batch of 4096 systems (orders log-uniform in [nmin, nmax],
// kappa = 1e3, 1-8 right-hand sides) and keep the fastest.  The batch size
// matters: few large problems favour one workgroup per CU (the largest
// problem's latency sets the time), many favour more (throughput).
static int autotune_lu(la_ctx *ctx, int nmin, int nmax, const char *cache_path)
{
    const int count = 4096;
    std::vector<int32_t> hn(count), hr(count);
    std::vector<int64_t> ha(count), hb(count);
    uint64_t s = 0x13198A2E03707344ull;
    int64_t oa = 0, ob = 0;
    const double lr = std::log((double)nmax / nmin);
    for (int i = 0; i < count; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double u = (double)(s >> 11) * 0x1.0p-53;
        hn[i] = std::max(nmin, std::min(nmax, (int)std::lround(nmin * std::exp(u * lr))));
        hr[i] = 1 + (int)((s >> 3) % 8);
        ha[i] = oa; hb[i] = ob;
        oa += (int64_t)hn[i] * hn[i];
        ob += (int64_t)hn[i] * hr[i];
    }
    double *dA = nullptr, *dB = nullptr, *dX = nullptr, *dXs = nullptr, *dE = nullptr;
    int64_t *dOA = nullptr, *dOB = nullptr;
    int32_t *dN = nullptr, *dR = nullptr, *dS = nullptr, *dI = nullptr;
    hipError_t he = hipMalloc(&dA, oa * 8);
    if (he == hipSuccess) he = hipMalloc(&dB, ob * 8);
    if (he == hipSuccess) he = hipMalloc(&dX, ob * 8);
    if (he == hipSuccess) he = hipMalloc(&dXs, ob * 8);
    if (he == hipSuccess) he = hipMalloc(&dE, count * 8);
    if (he == hipSuccess) he = hipMalloc(&dOA, count * 8);
    if (he == hipSuccess) he = hipMalloc(&dOB, count * 8);
    if (he == hipSuccess) he = hipMalloc(&dN, count * 4);
    if (he == hipSuccess) he = hipMalloc(&dR, count * 4);
    if (he == hipSuccess) he = hipMalloc(&dS, count * 4);
    if (he == hipSuccess) he = hipMalloc(&dI, count * 4);
    if (he == hipSuccess) he = hipMemcpy(dOA, ha.data(), count * 8, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dOB, hb.data(), count * 8, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dN, hn.data(), count * 4, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dR, hr.data(), count * 4, hipMemcpyHostToDevice);
    int st = he == hipSuccess ? LA_OK : hip_fail(ctx, he, "autotune alloc");
    if (st == LA_OK)
        st = la_gen_svd_batch_dev(ctx, count, dN, dR, dOA, dOB, 1e3, 16, 3, 0, dA, dB, dXs);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == LA_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = set_error(ctx, LA_ERR_HIP, "autotune events");
    int best[4] = {0, 0, 0, 0};
    float best_ms[4] = {0, 0, 0, 0};
    const int precs[3] = {LA_PREC_FP64, LA_PREC_FP32, LA_PREC_FP16};
    for (int pi = 0; pi < 3 && st == LA_OK; ++pi) {
        const int prec = precs[pi];
        la_batch_plan *p = nullptr;
        if ((st = la_batch_plan_create(ctx, 1, count, hn.data(), hr.data(), prec, &p))) break;
        const int cus = std::max(1, ctx->num_cus);
        const int maxg = lu_ir_batched_grid(ctx, prec) / cus;
        float bestms = 1e30f;
        for (int g : {1, 2, maxg}) {
            if (g < 1 || g > maxg || st) continue;
            float ms = 0.f;
            for (int rep = 0; rep < 3 && st == LA_OK; ++rep) {
                (void)hipEventRecord(e0, ctx->stream);
                st = launch_lu_ir_batched(ctx, ctx->stream, prec, g * cus, count, p->d_order,
                                          p->d_n, p->d_nrhs, dOA, dOB, p->d_offW, p->d_offN, dA,
                                          dB, dX, p->d_W, p->d_perm, 1e-14, 30, dS, dI, dE,
                                          p->d_queue);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float t = 0.f;
                (void)hipEventElapsedTime(&t, e0, e1);
                if (rep > 0) ms += t;
            }
            if (st == LA_OK && ms < bestms) { bestms = ms; best[prec] = g; best_ms[prec] = ms / 2; }
        }
        la_batch_plan_destroy(p);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dX); (void)hipFree(dXs); (void)hipFree(dE);
    (void)hipFree(dOA); (void)hipFree(dOB); (void)hipFree(dN); (void)hipFree(dR); (void)hipFree(dS);
    (void)hipFree(dI);
    if (st) return st;
    memcpy(ctx->lu_wg_per_cu, best, sizeof best);
    ctx->lu_tuned = true;
    if (cache_path) {
        FILE *f = fopen(cache_path, "w");
        if (!f) return set_error(ctx, LA_ERR_ARG, "autotune: cannot write cache");
        fprintf(f,
                "{\"kind\": \"lu-ir\", \"orders\": [%d, %d], "
                "\"workgroups_per_cu\": [%d, %d, %d], \"ms_per_4096\": [%.3f, %.3f, %.3f]}\n",
                nmin, nmax, best[LA_PREC_FP64], best[LA_PREC_FP32], best[LA_PREC_FP16],
                best_ms[LA_PREC_FP64], best_ms[LA_PREC_FP32], best_ms[LA_PREC_FP16]);
        fclose(f);
    }
    return LA_OK;
}

// Kind 0 (batched Cholesky): for each size bucket, time every candidate
// panel width on a synthetic batch of that bucket's sizes and keep the
// fastest.
extern "C" int la_autotune(la_ctx *ctx, int kind, int nmin, int nmax, const char *cache_path)
{
    if (ctx && kind == 1 && nmin >= 1 && nmax <= 1024 && nmin <= nmax)
        return autotune_lu(ctx, nmin, nmax, cache_path);
    if (!ctx || kind != 0 || nmin < 1 || nmax > 256 || nmin > nmax) return LA_ERR_ARG;
    const int lo[4] = {1, 33, 65, 129}, hi[4] = {32, 64, 128, 256};
    int best[4];
    for (int b = 0; b < 4; ++b) best[b] = ctx->chol_tuned ? ctx->chol_nb[b] : kCholDefault;
    hipEvent_t e0, e1;
    LA_HIP_CHECK(ctx, hipEventCreate(&e0));
    LA_HIP_CHECK(ctx, hipEventCreate(&e1));
    int st = LA_OK;
    for (int b = 0; b < 4 && st == LA_OK; ++b) {
        const int a = std::max(lo[b], nmin), z = std::min(hi[b], nmax);
        if (a > z) continue;
        const int count = 2048;
        std::vector<int32_t> hn(count);
        std::vector<int64_t> ha(count), hb(count);
        uint64_t s = 0x243F6A8885A308D3ull + b;
        int64_t oa = 0, ob = 0;
        for (int i = 0; i < count; ++i) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            hn[i] = a + (int)(s % (uint64_t)(z - a + 1));
            ha[i] = oa; hb[i] = ob;
            oa += (int64_t)hn[i] * hn[i];
            ob += hn[i];
        }
        la_batch_plan *p = nullptr;
        if ((st = la_batch_plan_create(ctx, 0, count, hn.data(), nullptr, 0, &p))) break;
        double *dA = nullptr, *dL = nullptr, *dB = nullptr, *dX = nullptr;
        int64_t *dOA = nullptr, *dOB = nullptr;
        int32_t *dS = nullptr;
        hipError_t he = hipSuccess;
        he = hipMalloc(&dA, oa * 8); if (he == hipSuccess) he = hipMalloc(&dL, oa * 8);
        if (he == hipSuccess) he = hipMalloc(&dB, ob * 8);
        if (he == hipSuccess) he = hipMalloc(&dX, ob * 8);
        if (he == hipSuccess) he = hipMalloc(&dOA, count * 8);
        if (he == hipSuccess) he = hipMalloc(&dOB, count * 8);
        if (he == hipSuccess) he = hipMalloc(&dS, count * 4);
        if (he == hipSuccess) he = hipMemcpy(dOA, ha.data(), count * 8, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(dOB, hb.data(), count * 8, hipMemcpyHostToDevice);
        if (he != hipSuccess) st = hip_fail(ctx, he, "autotune alloc");
        if (st == LA_OK) st = fill_spd_batch(ctx, count, p->d_n, dOA, dA);
        if (st == LA_OK) st = la_fill_uniform_dev(ctx, 1, 1, ob, dB);
        float bestms = 1e30f;
        for (int cand : kCholCandidates) {
            if (st) break;
            for (int q = 0; q < 4; ++q) p->nb[q] = cand;
            float ms = 0.f;
            for (int rep = 0; rep < 3 && st == LA_OK; ++rep) {
                (void)hipEventRecord(e0, ctx->stream);
                st = launch_chol_batched(ctx, ctx->stream, p->grid, count, p->d_order, p->d_n, dOA,
                                         dOB, dA, dL, dB, dX, dS, p->d_queue, p->nb);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float t = 0.f;
                (void)hipEventElapsedTime(&t, e0, e1);
                if (rep > 0) ms += t;
            }
            if (st == LA_OK && ms < bestms) { bestms = ms; best[b] = cand; }
        }
        (void)hipFree(dA); (void)hipFree(dL); (void)hipFree(dB); (void)hipFree(dX);
        (void)hipFree(dOA); (void)hipFree(dOB); (void)hipFree(dS);
        la_batch_plan_destroy(p);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    memcpy(ctx->chol_nb, best, sizeof best);
    ctx->chol_tuned = true;
    if (cache_path) {
        FILE *f = fopen(cache_path, "w");
        if (!f) return set_error(ctx, LA_ERR_ARG, "autotune: cannot write cache");
        fprintf(f,
                "{\"kind\": \"cholesky\", \"buckets\": [32, 64, 128, 256], "
                "\"panel_width\": [%d, %d, %d, %d], \"row_tiles\": [%d, %d, %d, %d]}\n",
                best[0] & 255, best[1] & 255, best[2] & 255, best[3] & 255, best[0] >> 8,
                best[1] >> 8, best[2] >> 8, best[3] >> 8);
        fclose(f);
    }
    return LA_OK;
}

extern "C" int la_autotune_load(la_ctx *ctx, const char *cache_path)
{
    if (!ctx || !cache_path) return LA_ERR_ARG;
    FILE *f = fopen(cache_path, "r");
    if (!f) return set_error(ctx, LA_ERR_ARG, "autotune: cannot read cache");
    char buf[512] = {0};
    size_t nread = fread(buf, 1, sizeof buf - 1, f);
    fclose(f);
    buf[nread] = 0;
    if (strstr(buf, "\"lu-ir\"")) {
        const char *kw = strstr(buf, "\"workgroups_per_cu\"");
        int g[3];
        if (!kw || !strchr(kw, '[') ||
            sscanf(strchr(kw, '['), "[%d, %d, %d]", &g[0], &g[1], &g[2]) != 3)
            return set_error(ctx, LA_ERR_ARG, "autotune: malformed cache");
        for (int i = 0; i < 3; ++i)
            if (g[i] < 0 || g[i] > 64) return set_error(ctx, LA_ERR_ARG, "autotune: bad grid in cache");
        ctx->lu_wg_per_cu[LA_PREC_FP64] = g[0];
        ctx->lu_wg_per_cu[LA_PREC_FP32] = g[1];
        ctx->lu_wg_per_cu[LA_PREC_FP16] = g[2];
        ctx->lu_tuned = true;
        return LA_OK;
    }
    const char *k = strstr(buf, "\"panel_width\"");
    int v[4], t[4] = {2, 2, 2, 2};
    if (!k || !strchr(k, '[') ||
        sscanf(strchr(k, '['), "[%d, %d, %d, %d]", &v[0], &v[1], &v[2], &v[3]) != 4)
        return set_error(ctx, LA_ERR_ARG, "autotune: malformed cache");
    const char *kt = strstr(buf, "\"row_tiles\"");     // optional, default 2
    if (kt && (!strchr(kt, '[') ||
               sscanf(strchr(kt, '['), "[%d, %d, %d, %d]", &t[0], &t[1], &t[2], &t[3]) != 4))
        return set_error(ctx, LA_ERR_ARG, "autotune: malformed cache");
    for (int i = 0; i < 4; ++i) {
        if (v[i] != 8 && v[i] != 16)
            return set_error(ctx, LA_ERR_ARG, "autotune: bad panel width in cache");
        if (t[i] != 1 && t[i] != 2)
            return set_error(ctx, LA_ERR_ARG, "autotune: bad row-tile count in cache");
    }
    for (int i = 0; i < 4; ++i) ctx->chol_nb[i] = v[i] | (t[i] << 8);
    ctx->chol_tuned = true;
    return LA_OK;
}
