// C ABI: context, workspaces, and the drop-in entry points that stand under
// rust-la's Matrix<f64> multiply, LUDecomposition and CholeskyDecomposition.
#include "la_common.hpp"
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <cstdlib>

namespace lahip {

int set_error(la_ctx *ctx, int status, const char *what) {
    if (ctx) ctx->last_error = what ? what : "";
    return status;
}

int hip_fail(la_ctx *ctx, hipError_t e, const char *what) {
    if (ctx) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
        ctx->last_error = buf;
    }
    return (e == hipErrorOutOfMemory) ? LA_ERR_NOMEM : LA_ERR_HIP;
}

int ws_get(la_ctx *ctx, int id, size_t bytes, void **out) {
    Workspace &w = ctx->ws[id];
    if (bytes == 0) bytes = 256;
    if (w.size < bytes) {
        size_t total = bytes;
        for (int i = 0; i < WS_COUNT; ++i)
            if (i != id) total += ctx->ws[i].size;
        if (ctx->ws_limit > 0 && (int64_t)total > ctx->ws_limit)
            return set_error(ctx, LA_ERR_NOMEM, "workspace limit exceeded");
        if (w.ptr) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return hip_fail(ctx, e, "ws_get sync");
            (void)hipFree(w.ptr);
            w.ptr = nullptr;
            w.size = 0;
        }
        size_t sz = (bytes + 4095) & ~(size_t)4095;
        hipError_t e = hipMalloc(&w.ptr, sz);
        if (e != hipSuccess) {
            w.ptr = nullptr;
            return hip_fail(ctx, e, "hipMalloc(workspace)");
        }
        w.size = sz;
    }
    *out = w.ptr;
    return LA_OK;
}

static hipEvent_t ev_take(la_ctx *ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

ProfScope::ProfScope(la_ctx *c, int kclass, double flop, double bytes) : ctx(c), on(false) {
    if (!ctx || !ctx->prof) return;
    rec.kclass = kclass;
    rec.flop = flop;
    rec.bytes = bytes;
    rec.start = ev_take(ctx);
    rec.stop = ev_take(ctx);
    if (!rec.start || !rec.stop) return;
    on = hipEventRecord(rec.start, ctx->stream) == hipSuccess;
}

ProfScope::~ProfScope() {
    if (!on) return;
    if (hipEventRecord(rec.stop, ctx->stream) == hipSuccess) ctx->prof_pending.push_back(rec);
}

static int prof_resolve(la_ctx *ctx) {
    if (ctx->prof_pending.empty()) return LA_OK;
    // kernels may sit on several streams (the LU look-ahead): wait for all
    for (auto &r : ctx->prof_pending) LA_HIP_CHECK(ctx, hipEventSynchronize(r.stop));
    for (auto &r : ctx->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) {
            ProfStats &s = ctx->prof_stats[r.kclass];
            s.calls += 1;
            s.ms += ms;
            s.flop += r.flop;
            s.bytes += r.bytes;
        }
        ctx->ev_pool.push_back(r.start);
        ctx->ev_pool.push_back(r.stop);
    }
    ctx->prof_pending.clear();
    return LA_OK;
}

}  // namespace lahip

using namespace lahip;

extern "C" int la_ctx_set_profiling(la_ctx *ctx, int enable) {
    if (!ctx) return LA_ERR_ARG;
    int st = prof_resolve(ctx);
    ctx->prof = enable != 0;
    return st;
}

extern "C" int la_ctx_reset_kernel_stats(la_ctx *ctx) {
    if (!ctx) return LA_ERR_ARG;
    int st = prof_resolve(ctx);
    for (auto &s : ctx->prof_stats) s = ProfStats();
    return st;
}

extern "C" int la_ctx_kernel_stats(la_ctx *ctx, int kclass, int64_t *calls, double *total_ms,
                                   double *flop, double *bytes) {
    if (!ctx || kclass < 0 || kclass >= LA_K_COUNT) return LA_ERR_ARG;
    int st = prof_resolve(ctx);
    const ProfStats &s = ctx->prof_stats[kclass];
    if (calls) *calls = s.calls;
    if (total_ms) *total_ms = s.ms;
    if (flop) *flop = s.flop;
    if (bytes) *bytes = s.bytes;
    return st;
}

// ------------------------------------------------------------------ context
extern "C" int la_ctx_create(int device, la_ctx **out) {
    if (!out) return LA_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LA_ERR_HIP;
    if (device < 0 || device >= ndev) return LA_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return LA_ERR_HIP;
    la_ctx *ctx = new la_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return LA_ERR_HIP; }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->hbm_bytes = (int64_t)prop.totalGlobalMem;
    ctx->clock_khz = prop.clockRate;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return LA_ERR_HIP;
    }
    ctx->own_stream = true;
    if (hipMalloc(&ctx->d_scratch, 4096) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return LA_ERR_NOMEM;
    }
    if (hipHostMalloc((void **)&ctx->h_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        hipFree(ctx->d_scratch);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return LA_ERR_NOMEM;
    }
    if (hipMemsetAsync(ctx->d_scratch, 0, 4096, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess ||
        noalloc_scratch_init(device) != LA_OK) {
        la_ctx_destroy(ctx);
        return LA_ERR_HIP;
    }
    *out = ctx;
    return LA_OK;
}

extern "C" void la_ctx_destroy(la_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->prof_pending) {
        hipEventDestroy(r.start);
        hipEventDestroy(r.stop);
    }
    for (auto e : ctx->ev_pool) hipEventDestroy(e);
    for (auto e : ctx->la_ev)
        if (e) hipEventDestroy(e);
    for (auto &w : ctx->ws)
        if (w.ptr) hipFree(w.ptr);
    if (ctx->d_scratch) hipFree(ctx->d_scratch);
    if (ctx->h_pinned) hipHostFree(ctx->h_pinned);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int la_ctx_synchronize(la_ctx *ctx) {
    if (!ctx) return LA_ERR_ARG;
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" void *la_ctx_stream(la_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int la_ctx_set_stream(la_ctx *ctx, void *s) {
    if (!ctx) return LA_ERR_ARG;
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    ctx->stream = (hipStream_t)s;
    ctx->own_stream = false;
    return LA_OK;
}

extern "C" const char *la_status_string(int s) {
    switch (s) {
    case LA_OK: return "ok";
    case LA_ERR_SHAPE: return "shape mismatch (reference: assert! panic)";
    case LA_SINGULAR: return "singular: exact zero on U's diagonal (reference: None)";
    case LA_NOT_SQUARE: return "not square (reference: None)";
    case LA_NOT_SYMMETRIC: return "not symmetric (reference: None)";
    case LA_NOT_SPD: return "not positive definite (reference: None)";
    case LA_ERR_HIP: return "HIP runtime error";
    case LA_ERR_NOMEM: return "out of device memory";
    case LA_NOT_CONVERGED: return "iterative refinement did not converge";
    case LA_ERR_ARG: return "invalid argument";
    case LA_ERR_TIMEOUT: return "device wait exceeded its spin limit";
    case LA_ERR_IO: return "file missing or item not an f64 (reference: unwrap panic)";
    default: return "unknown status";
    }
}

extern "C" const char *la_ctx_last_error(la_ctx *ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" int la_device_info(la_ctx *ctx, int *num_cus, int64_t *hbm_bytes, int *clock_khz) {
    if (!ctx) return LA_ERR_ARG;
    if (num_cus) *num_cus = ctx->num_cus;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    if (clock_khz) *clock_khz = ctx->clock_khz;
    return LA_OK;
}

extern "C" int la_ctx_set_workspace_limit(la_ctx *ctx, int64_t bytes) {
    if (!ctx || bytes < 0) return LA_ERR_ARG;
    ctx->ws_limit = bytes;
    return LA_OK;
}

// ------------------------------------------------------------ helpers
namespace {

struct HostBufs {
    char *base = nullptr;
    size_t off = 0;
    template <typename T> T *take(int64_t count) {
        T *p = reinterpret_cast<T *>(base + off);
        off += (((size_t)std::max<int64_t>(count, 1) * sizeof(T)) + 255) & ~(size_t)255;
        return p;
    }
};

size_t rounded(int64_t count, size_t elem) {
    return (((size_t)std::max<int64_t>(count, 1) * elem) + 255) & ~(size_t)255;
}

}  // namespace

namespace lahip {
int read_scratch(la_ctx *ctx, const int32_t *d, int32_t *h, int n) {
    LA_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_pinned, d, n * sizeof(int32_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(h, ctx->h_pinned, n * sizeof(int32_t));
    return LA_OK;
}

int check_sticky(la_ctx *ctx) {
    int32_t e = 0;
    int st = read_scratch(ctx, ctx->d_scratch + SCR_ERR, &e, 1);
    if (st) return st;
    if (e != 0) {
        // clear it so the next call starts clean
        fill_i32(ctx, ctx->d_scratch + SCR_ERR, 1, 0);
        return set_error(ctx, e, "device wait timed out");
    }
    return LA_OK;
}

}  // namespace lahip


// ------------------------------------------------------------------ GEMM
extern "C" int la_dgemm_dev(la_ctx *ctx, int64_t m, int64_t n, int64_t k, double alpha,
                            const double *A, int64_t lda, const double *B, int64_t ldb,
                            double beta, double *C, int64_t ldc) {
    if (!ctx || !A || !B || !C) return LA_ERR_ARG;
    if (m < 0 || n < 0 || k < 0 || lda < std::max<int64_t>(k, 1) ||
        ldb < std::max<int64_t>(n, 1) || ldc < std::max<int64_t>(n, 1))
        return LA_ERR_ARG;
    return gemm_rm<double>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, false, false);
}

extern "C" int la_dgemm_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                           const double *A, const double *B, double *C) {
    if (!ctx) return LA_ERR_ARG;
    if (k != k2) return set_error(ctx, LA_ERR_SHAPE, "Mul: self.cols() != m.rows()");
    if (m <= 0 || k <= 0 || p <= 0) return set_error(ctx, LA_ERR_SHAPE, "Mul: empty matrix");
    if (!A || !B || !C) return LA_ERR_ARG;
    void *base = nullptr;
    size_t total = rounded(m * k, 8) + rounded(k * p, 8) + rounded(m * p, 8);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dA = hb.take<double>(m * k), *dB = hb.take<double>(k * p), *dC = hb.take<double>(m * p);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, m * k * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dB, B, k * p * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = gemm_rm<double>(ctx, m, p, k, 1.0, dA, k, dB, p, 0.0, dC, p, false, false)))
        return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(C, dC, m * p * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" int la_dmmul_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                           int64_t dst_rows, int64_t dst_cols, const double *A, const double *B,
                           double *dst) {
    if (!ctx) return LA_ERR_ARG;
    if (k != k2) return set_error(ctx, LA_ERR_SHAPE, "mmul: self.cols() != m.rows()");
    if (dst_rows != m) return set_error(ctx, LA_ERR_SHAPE, "mmul: dst.rows() != self.rows()");
    if (dst_cols != p) return set_error(ctx, LA_ERR_SHAPE, "mmul: dst.cols() != m.cols()");
    return la_dgemm_rm(ctx, m, k, k2, p, A, B, dst);
}

// -------------------------------------------------------------------- LU
extern "C" int la_dgetrf_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda,
                             uint64_t *d_piv, int32_t *d_info) {
    if (!ctx || !A || !d_piv || !d_info) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "getrf: empty matrix");
    if (lda < n) return LA_ERR_ARG;
    const int64_t kmax = std::min(m, n);
    void *p = nullptr;
    int st = ws_get(ctx, WS_PIV, (size_t)kmax * sizeof(int64_t), &p);
    if (st) return st;
    int64_t *ipiv = (int64_t *)p;
    if ((st = getrf_core<double>(ctx, m, n, A, lda, ipiv, d_piv, d_info + 1))) return st;
    return pivsign(ctx, ipiv, kmax, d_info);
}

extern "C" int la_dgetrf_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, double *LU,
                            uint64_t *piv, int *pospivsign) {
    if (!ctx) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "LUDecomposition: empty matrix");
    if (!A || !LU || !piv || !pospivsign) return LA_ERR_ARG;
    void *base = nullptr;
    size_t total = rounded(m * n, 8) + rounded(m, 8) + rounded(4, 4);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dA = hb.take<double>(m * n);
    uint64_t *dpiv = hb.take<uint64_t>(m);
    int32_t *dinfo = hb.take<int32_t>(4);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = la_dgetrf_dev(ctx, m, n, dA, n, dpiv, dinfo))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(LU, dA, m * n * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(piv, dpiv, m * 8, hipMemcpyDeviceToHost, ctx->stream));
    int32_t info[2];
    if ((st = read_scratch(ctx, dinfo, info, 2))) return st;
    if (info[1] != LA_OK) return set_error(ctx, info[1], "getrf: device wait failed");
    *pospivsign = info[0];
    return check_sticky(ctx);
}

// Shared solve body on device buffers: X(m x nx) from LU (m x n), perm.
static int getrs_device(la_ctx *ctx, int64_t m, int64_t n, const double *dLU, int64_t ldlu,
                        const uint64_t *dpiv, int64_t nx, const double *dB, int64_t ldb,
                        double *dX, int64_t ldx) {
    int st;
    if ((st = gather_rows(ctx, m, nx, dpiv, dB, ldb, dX, ldx))) return st;
    if (nx >= TRSM_MIN_RHS && n >= TRSM_MIN_N) {     // many columns: MFMA TRSM
        if ((st = trsm_blocked(ctx, 'L', n, dLU, ldlu, nx, dX, ldx))) return st;
        return trsm_blocked(ctx, 'U', n, dLU, ldlu, nx, dX, ldx);
    }
    if ((st = trsv_blocked<double>(ctx, 'L', false, true, n, dLU, ldlu, nx, dX, ldx))) return st;
    return trsv_blocked<double>(ctx, 'U', false, false, n, dLU, ldlu, nx, dX, ldx);
}

extern "C" int la_dgetrs_dev(la_ctx *ctx, int64_t m, int64_t n, const double *LU, int64_t ldlu,
                             const uint64_t *d_piv, int64_t nx, const double *B, int64_t ldb,
                             double *X, int64_t ldx, int32_t *d_status) {
    if (!ctx || !LU || !d_piv || !B || !X) return LA_ERR_ARG;
    if (nx <= 0 || m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "solve: empty");
    if (B == X) return set_error(ctx, LA_ERR_ARG, "getrs_dev: X must not alias B");
    int st;
    if (m < n) {
        // lu.rs:174-182: a zero among the m in-bounds diagonal entries is
        // None; otherwise the scan indexes out of bounds (a panic)
        int32_t *probe = d_status ? d_status : ctx->d_scratch + SCR_STATUS;
        if ((st = check_diag_zero(ctx, m, LU, ldlu, probe))) return st;
        int32_t h = 0;
        if ((st = read_scratch(ctx, probe, &h, 1))) return st;
        if (h == LA_SINGULAR) return LA_SINGULAR;
        return set_error(ctx, LA_ERR_SHAPE, "is_non_singular: index out of bounds (rows < cols)");
    }
    if (d_status && (st = check_diag_zero(ctx, n, LU, ldlu, d_status))) return st;
    return getrs_device(ctx, m, n, LU, ldlu, d_piv, nx, B, ldb, X, ldx);
}

extern "C" int la_dgetrs_rm(la_ctx *ctx, int64_t m, int64_t n, const double *LU,
                            const uint64_t *piv, int64_t b_rows, int64_t nx, const double *B,
                            double *X) {
    if (!ctx) return LA_ERR_ARG;
    if (b_rows != m) return set_error(ctx, LA_ERR_SHAPE, "solve: b.rows() != lu.rows()");
    if (nx <= 0 || m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "solve: empty matrix");
    if (!LU || !piv || !B || !X) return LA_ERR_ARG;
    // lu.rs:174-182 scans j = 0..n-1 and reads lu[j*n + j]: an exact zero
    // returns None before an out-of-bounds index (j*n + j >= m*n, first at
    // j = m when m < n) panics.
    for (int64_t j = 0; j < n; ++j) {
        if (j >= m) return set_error(ctx, LA_ERR_SHAPE, "is_non_singular: index out of bounds (rows < cols)");
        if (LU[j * n + j] == 0.0) return LA_SINGULAR;
    }
    for (int64_t i = 0; i < m; ++i)
        if (piv[i] >= (uint64_t)m) return set_error(ctx, LA_ERR_ARG, "solve: piv out of range");
    void *base = nullptr;
    size_t total = rounded(m * n, 8) + rounded(m, 8) + 2 * rounded(m * nx, 8);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dLU = hb.take<double>(m * n);
    uint64_t *dpiv = hb.take<uint64_t>(m);
    double *dB = hb.take<double>(m * nx), *dX = hb.take<double>(m * nx);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dLU, LU, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dpiv, piv, m * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dB, B, m * nx * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = getrs_device(ctx, m, n, dLU, n, dpiv, nx, dB, nx, dX, nx))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(X, dX, m * nx * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

extern "C" int la_dgesv_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, int64_t b_rows,
                           int64_t nx, const double *B, double *X) {
    if (!ctx) return LA_ERR_ARG;
    if (m <= 0 || n <= 0 || nx <= 0) return set_error(ctx, LA_ERR_SHAPE, "solve: empty matrix");
    if (b_rows != m) return set_error(ctx, LA_ERR_SHAPE, "solve: b.rows() != a.rows()");
    if (!A || !B || !X) return LA_ERR_ARG;
    void *base = nullptr;
    size_t total = rounded(m * n, 8) + rounded(m, 8) + 2 * rounded(m * nx, 8) + rounded(4, 4);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dA = hb.take<double>(m * n);
    uint64_t *dpiv = hb.take<uint64_t>(m);
    double *dB = hb.take<double>(m * nx), *dX = hb.take<double>(m * nx);
    int32_t *dinfo = hb.take<int32_t>(4);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dB, B, m * nx * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = la_dgetrf_dev(ctx, m, n, dA, n, dpiv, dinfo))) return st;
    // the first min(m, n) diagonal entries are in bounds (lu.rs:174-182);
    // with m < n a zero among them is None, otherwise index m panics
    if ((st = check_diag_zero(ctx, std::min(m, n), dA, n, dinfo + 2))) return st;
    int32_t info[3];
    if ((st = read_scratch(ctx, dinfo, info, 3))) return st;
    if (info[1] != LA_OK) return set_error(ctx, info[1], "getrf: device wait failed");
    if (info[2] == LA_SINGULAR) return LA_SINGULAR;
    if (m < n) return set_error(ctx, LA_ERR_SHAPE, "is_non_singular: index out of bounds (rows < cols)");
    if ((st = getrs_device(ctx, m, n, dA, n, dpiv, nx, dB, nx, dX, nx))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(X, dX, m * nx * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

// --------------------------------------------------------------- inverse
namespace {
// X = P I: X[i][j] = 1 where j == piv[i]  (lu.rs:246-254 with B = id)
__global__ void k_perm_identity(int64_t n, const uint64_t *piv, double *X)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e / n, j = e - i * n;
    X[e] = (piv[i] == (uint64_t)j) ? 1.0 : 0.0;
}
}  // namespace

extern "C" int la_dinverse_dev(la_ctx *ctx, int64_t n, double *A, int64_t lda, uint64_t *d_piv,
                               double *X, int64_t ldx, int32_t *d_info)
{
    if (!ctx || !A || !d_piv || !X || !d_info || n <= 0 || lda < n) return LA_ERR_ARG;
    if (ldx != n) return set_error(ctx, LA_ERR_ARG, "inverse_dev: ldx must equal n");
    int st;
    if ((st = la_dgetrf_dev(ctx, n, n, A, lda, d_piv, d_info))) return st;
    if ((st = check_diag_zero(ctx, n, A, lda, d_info + 2))) return st;
    // X = P I, then the two triangular solves with n right-hand sides
    hipLaunchKernelGGL(k_perm_identity, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0,
                       ctx->stream, n, (const uint64_t *)d_piv, X);
    LA_HIP_CHECK(ctx, hipGetLastError());
    if (n >= TRSM_MIN_N) {
        if ((st = trsm_blocked(ctx, 'L', n, A, lda, n, X, ldx))) return st;
        return trsm_blocked(ctx, 'U', n, A, lda, n, X, ldx);
    }
    if ((st = trsv_blocked<double>(ctx, 'L', false, true, n, A, lda, n, X, ldx))) return st;
    return trsv_blocked<double>(ctx, 'U', false, false, n, A, lda, n, X, ldx);
}

extern "C" int la_dinverse_rm(la_ctx *ctx, int64_t rows, int64_t cols, const double *A,
                              double *Ainv)
{
    if (!ctx) return LA_ERR_ARG;
    if (rows != cols) return set_error(ctx, LA_ERR_SHAPE, "inverse: not square");  // mod.rs:1035
    if (rows <= 0) return set_error(ctx, LA_ERR_SHAPE, "inverse: empty matrix");
    if (!A || !Ainv) return LA_ERR_ARG;
    const int64_t n = rows;
    void *base = nullptr;
    size_t total = rounded(n * n, 8) * 2 + rounded(n, 8) + rounded(4, 4);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dA = hb.take<double>(n * n), *dX = hb.take<double>(n * n);
    uint64_t *dpiv = hb.take<uint64_t>(n);
    int32_t *dinfo = hb.take<int32_t>(4);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, n * n * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = la_dinverse_dev(ctx, n, dA, n, dpiv, dX, n, dinfo))) return st;
    int32_t info[3];
    if ((st = read_scratch(ctx, dinfo, info, 3))) return st;
    if (info[1] != LA_OK) return set_error(ctx, info[1], "getrf: device wait failed");
    if (info[2] == LA_SINGULAR) return LA_SINGULAR;                 // lu.rs:241-243: None
    LA_HIP_CHECK(ctx, hipMemcpyAsync(Ainv, dX, n * n * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

// -------------------------------------------------------------- Cholesky
extern "C" int la_dpotrf_dev(la_ctx *ctx, int64_t n, double *A, int64_t lda, int32_t *d_status) {
    if (!ctx || !A || !d_status) return LA_ERR_ARG;
    if (n <= 0 || lda < n) return LA_ERR_ARG;
    return potrf_blocked(ctx, n, A, lda, d_status);
}

extern "C" int la_dpotrs_dev(la_ctx *ctx, int64_t n, const double *L, int64_t ldl, int64_t nx,
                             const double *B, int64_t ldb, double *X, int64_t ldx) {
    if (!ctx || !L || !B || !X) return LA_ERR_ARG;
    if (n <= 0 || nx <= 0) return LA_ERR_ARG;
    int st;
    if (X != B && (st = copy2d(ctx, n, nx, B, ldb, X, ldx))) return st;
    if ((st = trsv_blocked<double>(ctx, 'L', false, false, n, L, ldl, nx, X, ldx))) return st;
    return trsv_blocked<double>(ctx, 'L', true, false, n, L, ldl, nx, X, ldx);
}

extern "C" int la_dpotrf_rm(la_ctx *ctx, int64_t rows, int64_t cols, const double *A,
                            double *L) {
    if (!ctx) return LA_ERR_ARG;
    if (rows != cols) return LA_NOT_SQUARE;              // cholesky.rs:57-59
    if (rows <= 0) return set_error(ctx, LA_ERR_SHAPE, "Cholesky: empty matrix");
    if (!A || !L) return LA_ERR_ARG;
    const int64_t n = rows;
    void *base = nullptr;
    size_t total = rounded(n * n, 8) + rounded(4, 4);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dA = hb.take<double>(n * n);
    int32_t *dst = hb.take<int32_t>(4);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, n * n * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = potrf_blocked(ctx, n, dA, n, dst))) return st;
    int32_t status = 0;
    if ((st = read_scratch(ctx, dst, &status, 1))) return st;
    if (status != LA_OK) return status;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(L, dA, n * n * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" int la_dpotrs_rm(la_ctx *ctx, int64_t n, const double *L, int64_t b_rows, int64_t nx,
                            const double *B, double *X) {
    if (!ctx) return LA_ERR_ARG;
    if (b_rows != n) return set_error(ctx, LA_ERR_SHAPE, "Cholesky solve: l.rows() != b.rows()");
    if (n <= 0 || nx <= 0) return set_error(ctx, LA_ERR_SHAPE, "Cholesky solve: empty matrix");
    if (!L || !B || !X) return LA_ERR_ARG;
    void *base = nullptr;
    size_t total = rounded(n * n, 8) + rounded(n * nx, 8);
    int st = ws_get(ctx, WS_HOST, total, &base);
    if (st) return st;
    HostBufs hb{(char *)base, 0};
    double *dL = hb.take<double>(n * n);
    double *dX = hb.take<double>(n * nx);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dL, L, n * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dX, B, n * nx * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = la_dpotrs_dev(ctx, n, dL, n, nx, dX, nx, dX, nx))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(X, dX, n * nx * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}
