// Mixed-precision LU with FP64 iterative refinement (no reference
// counterpart; the FP64 solve it accelerates is LUDecomposition::new +
// solve, src/decomp/lu.rs:104-278).
//
//   A32 = fl32(A)                       conversion kernel (HBM: 8 B in, 4 B out)
//   P A32 = L U                         getrf_core<float>: FP32 panel, FP32 MFMA
//   x = U^-1 L^-1 P b                   FP32 factors, FP64 arithmetic
//   loop: r = b - A x  (FP64, A read from HBM once per step)
//         eta = ||r||inf / (||A||inf ||x||inf + ||b||inf)
//         stop if eta <= tol;  d = U^-1 L^-1 P r;  x += d
// FP64 is used only for the residual, the update and the check.
#include "la_common.hpp"
#include <cmath>
#include <cstring>

namespace lahip {

__global__ void k_convert_f32(int64_t rows, int64_t cols, const double *A, int64_t lda,
                              float *B, int64_t ldb)
{
    int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const int64_t total = rows * cols;
    if (e >= total) return;
    if ((cols & 3) == 0 && (lda & 1) == 0) {
        const int64_t r = e / cols, c = e - r * cols;
        const double2 *src = reinterpret_cast<const double2 *>(A + r * lda + c);
        double2 a = src[0], b = src[1];
        float4 o = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
        *reinterpret_cast<float4 *>(B + r * ldb + c) = o;
    } else {
        for (int q = 0; q < 4 && e + q < total; ++q) {
            const int64_t r = (e + q) / cols, c = (e + q) - r * cols;
            B[r * ldb + c] = (float)A[r * lda + c];
        }
    }
}

// Column maxima of |A| (non-negative doubles compared as bit patterns):
// grid (column blocks of 256, row chunks of 256).
__global__ __launch_bounds__(256) void k_colmax(int64_t n, const double *A, int64_t lda,
                                                unsigned long long *colmax)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int64_t i0 = (int64_t)blockIdx.y * 256, i1 = min(n, i0 + 256);
    double m = 0.0;
    for (int64_t i = i0; i < i1; ++i) m = fmax(m, fabs(A[i * lda + j]));
    atomicMax(colmax + j, (unsigned long long)__double_as_longlong(m));
}

// s_j = 2^-e_j with max_i |A_ij| = f * 2^e_j, f in [1, 2): powers of two,
// so scaling is exact; A32 = fl32(A S).
__global__ void k_colscale(int64_t n, const unsigned long long *colmax, double *s)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double m = __longlong_as_double((long long)colmax[j]);
    s[j] = (m > 0.0 && isfinite(m)) ? ldexp(1.0, -ilogb(m)) : 1.0;
}

__global__ void k_convert_f32_scaled(int64_t n, const double *A, int64_t lda, const double *s,
                                     float *B, int64_t ldb)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t r = e / n, c = e - r * n;
    B[r * ldb + c] = (float)(A[r * lda + c] * s[c]);
}

// x_j *= s_j for the solution rows (x = S y)
__global__ void k_scale_rows_by(int64_t n, int nrhs, const double *s, double *X, int64_t ldx)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t i = e / nrhs, r = e - i * nrhs;
    X[i * ldx + r] *= s[i];
}

// Non-negative doubles order like their bit patterns, so a running maximum
// can use the 64-bit integer atomicMax.
__device__ inline void atomic_max_nonneg(double *addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(addr),
              (unsigned long long)__double_as_longlong(v));
}

// R = B - A X for nrhs <= 16 right-hand sides, one workgroup per row.
// norms[0] = max |R|, norms[1] = max |X|, norms[2] = max |B|,
// norms[3] = max_i sum_j |A_ij| (only when want_anorm).
__global__ __launch_bounds__(256)
void k_residual(int64_t n, int nrhs, const double *__restrict__ A, int64_t lda,
                const double *__restrict__ X, int64_t ldx, const double *__restrict__ B,
                int64_t ldb, double *R, int64_t ldr, double *norms, int want_anorm)
{
    constexpr int NRM = 16;
    __shared__ double red[4][NRM + 1];
    const int64_t i = blockIdx.x;
    const double *arow = A + i * lda;
    double acc[NRM];
#pragma unroll
    for (int r = 0; r < NRM; ++r) acc[r] = 0.0;
    double asum = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
        const double a = arow[j];
        asum += fabs(a);
        const double *xr = X + j * ldx;
#pragma unroll
        for (int r = 0; r < NRM; ++r)
            if (r < nrhs) acc[r] += a * xr[r];
    }
    // reduce
#pragma unroll
    for (int r = 0; r < NRM; ++r) {
        double v = (r < nrhs) ? acc[r] : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        acc[r] = v;
    }
    for (int off = 32; off > 0; off >>= 1) asum += __shfl_xor(asum, off);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NRM; ++r) red[wv][r] = acc[r];
        red[wv][NRM] = asum;
    }
    __syncthreads();
    if (threadIdx.x < nrhs) {
        const int r = threadIdx.x;
        const double ax = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
        const double b = B[i * ldb + r];
        const double res = b - ax;
        R[i * ldr + r] = res;
        atomic_max_nonneg(norms + 0, fabs(res));
        atomic_max_nonneg(norms + 1, fabs(X[i * ldx + r]));
        atomic_max_nonneg(norms + 2, fabs(b));
    }
    if (threadIdx.x == 0 && want_anorm) {
        const double s = (red[0][NRM] + red[1][NRM]) + (red[2][NRM] + red[3][NRM]);
        atomic_max_nonneg(norms + 3, s);
    }
}

__global__ void k_axpy_rows(int64_t n, int nrhs, const double *D, int64_t ldd, double *X,
                            int64_t ldx)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    int64_t i = e / nrhs, r = e - i * nrhs;
    X[i * ldx + r] += D[i * ldd + r];
}

__global__ void k_zero_f64(double *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

int residual(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A, int64_t lda,
             const double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
             int64_t ldr, double *d_norms, bool want_anorm)
{
    if (nrhs > 16) return set_error(ctx, LA_ERR_ARG, "residual: nrhs > 16");
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, ctx->stream, d_norms, want_anorm ? 4 : 3);
    ProfScope ps(ctx, LA_K_RESID, 2.0 * n * n * nrhs, 8.0 * ((double)n * n + 3.0 * n * nrhs));
    hipLaunchKernelGGL(k_residual, dim3((unsigned)n), dim3(256), 0, ctx->stream, n, (int)nrhs, A,
                       lda, X, ldx, B, ldb, R, ldr, d_norms, want_anorm ? 1 : 0);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// Low-precision factor + solve helpers ------------------------------------
struct IrWork {
    float *A32 = nullptr;
    int64_t ld32 = 0;
    double *colscale = nullptr;          // per-column 2^k scale (BF16/FP16 modes)
    unsigned long long *colmax = nullptr;
    int prec = LA_PREC_FP32;
    uint64_t *perm = nullptr;
    int64_t *ipiv = nullptr;
    double *R = nullptr, *D = nullptr, *norms = nullptr;
    int32_t *status = nullptr;
};

static int ir_alloc(la_ctx *ctx, int64_t n, int64_t nrhs, IrWork &w)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t ld32 = (n + 63) & ~(int64_t)63;   // pad rows to 256 B
    size_t b32 = al((size_t)n * ld32 * sizeof(float));
    size_t bperm = al((size_t)n * 8), bipiv = al((size_t)n * 8);
    size_t brd = al((size_t)n * nrhs * 8);
    size_t bn = al(8 * 8), bst = al(4 * 4), bsc = al((size_t)n * 8);
    void *base = nullptr;
    int st = ws_get(ctx, WS_LP, b32 + bperm + bipiv + 2 * brd + bn + bst + 2 * bsc, &base);
    if (st) return st;
    char *p = (char *)base;
    w.A32 = (float *)p; p += b32;
    w.ld32 = ld32;
    w.perm = (uint64_t *)p; p += bperm;
    w.ipiv = (int64_t *)p; p += bipiv;
    w.R = (double *)p; p += brd;
    w.D = (double *)p; p += brd;
    w.norms = (double *)p; p += bn;
    w.status = (int32_t *)p; p += bst;
    w.colscale = (double *)p; p += bsc;
    w.colmax = (unsigned long long *)p;
    return LA_OK;
}

// d = U^-1 L^-1 P r with the FP32 factor (FP64 arithmetic)
static int lp_solve(la_ctx *ctx, int64_t n, int64_t nrhs, const IrWork &w, const double *rhs,
                    int64_t ldr, double *out, int64_t ldo)
{
    int st;
    if ((st = gather_rows(ctx, n, nrhs, w.perm, rhs, ldr, out, ldo))) return st;
    if ((st = trsv_blocked<float>(ctx, 'L', false, true, n, w.A32, w.ld32, nrhs, out, ldo)))
        return st;
    if ((st = trsv_blocked<float>(ctx, 'U', false, false, n, w.A32, w.ld32, nrhs, out, ldo)))
        return st;
    if (w.prec == LA_PREC_BF16 || w.prec == LA_PREC_FP16) {   // x = S y
        hipLaunchKernelGGL(k_scale_rows_by, dim3((unsigned)ceil_div(n * nrhs, 256)), dim3(256), 0,
                           ctx->stream, n, (int)nrhs, w.colscale, out, ldo);
        LA_HIP_CHECK(ctx, hipGetLastError());
    }
    return LA_OK;
}

}  // namespace lahip

using namespace lahip;

namespace {
struct EvTimer {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    EvTimer() { for (auto &x : e) (void)hipEventCreate(&x); }
    ~EvTimer() { for (auto &x : e) if (x) (void)hipEventDestroy(x); }
    float ms(int a, int b) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e[a], e[b]);
        return t;
    }
};
}  // namespace

extern "C" int la_dgesv_ir_dev(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A,
                               int64_t lda, const double *B, int64_t ldb, double *X,
                               int64_t ldx, int prec, double tol, int max_iter,
                               la_ir_report *report)
{
    if (!ctx || !A || !B || !X || !report) return LA_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || nrhs > 16 || lda < n || ldb < nrhs || ldx < nrhs)
        return LA_ERR_ARG;
    if (prec != LA_PREC_FP32 && prec != LA_PREC_BF16 && prec != LA_PREC_FP16)
        return set_error(ctx, LA_ERR_ARG, "dgesv_ir: precision not supported");
    std::memset(report, 0, sizeof(*report));
    report->prec = prec;
    IrWork w;
    int st = ir_alloc(ctx, n, nrhs, w);
    if (st) return st;
    EvTimer ev;
    w.prec = prec;
    LA_HIP_CHECK(ctx, hipEventRecord(ev.e[0], ctx->stream));
    if (prec == LA_PREC_FP32) {
        hipLaunchKernelGGL(k_convert_f32, dim3((unsigned)ceil_div(n * n, 1024)), dim3(256), 0,
                           ctx->stream, n, n, A, lda, w.A32, w.ld32);
    } else {
        // per-column power-of-two scaling keeps the rounded operands in range
        if ((st = fill_u64(ctx, (uint64_t *)w.colmax, n, 0))) return st;
        hipLaunchKernelGGL(k_colmax, dim3((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, 256)),
                           dim3(256), 0, ctx->stream, n, A, lda, w.colmax);
        hipLaunchKernelGGL(k_colscale, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                           ctx->stream, n, w.colmax, w.colscale);
        hipLaunchKernelGGL(k_convert_f32_scaled, dim3((unsigned)ceil_div(n * n, 256)), dim3(256),
                           0, ctx->stream, n, A, lda, w.colscale, w.A32, w.ld32);
    }
    LA_HIP_CHECK(ctx, hipGetLastError());
    if ((st = getrf_core<float>(ctx, n, n, w.A32, w.ld32, w.ipiv, w.perm, w.status, prec)))
        return st;
    LA_HIP_CHECK(ctx, hipEventRecord(ev.e[1], ctx->stream));
    // singular low-precision factor: report and stop
    // (exact zero on U's diagonal makes every correction infinite)
    if ((st = lp_solve(ctx, n, nrhs, w, B, ldb, X, ldx))) return st;
    double h[4] = {0, 0, 0, 0};
    int it = 0;
    double eta = INFINITY;
    for (;; ++it) {
        if ((st = residual(ctx, n, nrhs, A, lda, X, ldx, B, ldb, w.R, nrhs, w.norms, it == 0)))
            return st;
        LA_HIP_CHECK(ctx, hipMemcpyAsync(h, w.norms, 4 * sizeof(double), hipMemcpyDeviceToHost,
                                         ctx->stream));
        LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (it == 0) report->anorm_inf = h[3];
        const double denom = report->anorm_inf * h[1] + h[2];
        eta = denom > 0 ? h[0] / denom : (h[0] == 0 ? 0.0 : INFINITY);
        if (it == 0) report->initial_backward_error = eta;
        if (!(eta > tol) || it >= max_iter || !std::isfinite(eta)) break;
        if ((st = lp_solve(ctx, n, nrhs, w, w.R, nrhs, w.D, nrhs))) return st;
        hipLaunchKernelGGL(k_axpy_rows, dim3((unsigned)ceil_div(n * nrhs, 256)), dim3(256), 0,
                           ctx->stream, n, (int)nrhs, w.D, nrhs, X, ldx);
        LA_HIP_CHECK(ctx, hipGetLastError());
    }
    LA_HIP_CHECK(ctx, hipEventRecord(ev.e[2], ctx->stream));
    LA_HIP_CHECK(ctx, hipEventSynchronize(ev.e[2]));
    int32_t fst = 0, derr = 0;
    LA_HIP_CHECK(ctx, hipMemcpy(&fst, w.status, 4, hipMemcpyDeviceToHost));
    LA_HIP_CHECK(ctx, hipMemcpy(&derr, ctx->d_scratch + SCR_ERR, 4, hipMemcpyDeviceToHost));
    if (derr != 0) {
        fill_i32(ctx, ctx->d_scratch + SCR_ERR, 1, 0);
        if (fst == LA_OK) fst = derr;
    }
    report->iterations = it;
    report->backward_error = eta;
    report->factor_ms = ev.ms(0, 1);
    report->refine_ms = ev.ms(1, 2);
    report->total_ms = ev.ms(0, 2);
    if (fst != LA_OK) report->status = fst;
    else if (!(eta <= tol)) report->status = std::isfinite(eta) ? LA_NOT_CONVERGED : LA_SINGULAR;
    else report->status = LA_OK;
    return report->status == LA_ERR_TIMEOUT ? LA_ERR_TIMEOUT : report->status;
}

extern "C" int la_dgesv_ir_rm(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A,
                              const double *B, double *X, int prec, double tol, int max_iter,
                              la_ir_report *report)
{
    if (!ctx || !A || !B || !X || !report) return LA_ERR_ARG;
    if (n <= 0 || nrhs <= 0) return set_error(ctx, LA_ERR_SHAPE, "dgesv_ir: empty");
    void *base = nullptr;
    size_t ba = ((size_t)n * n * 8 + 255) & ~(size_t)255, bb = ((size_t)n * nrhs * 8 + 255) & ~(size_t)255;
    int st = ws_get(ctx, WS_HOST, ba + 2 * bb, &base);
    if (st) return st;
    double *dA = (double *)base, *dB = (double *)((char *)base + ba),
           *dX = (double *)((char *)base + ba + bb);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, n * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dB, B, n * nrhs * 8, hipMemcpyHostToDevice, ctx->stream));
    int rs = la_dgesv_ir_dev(ctx, n, nrhs, dA, n, dB, nrhs, dX, nrhs, prec, tol, max_iter, report);
    if (rs != LA_OK && rs != LA_NOT_CONVERGED) return rs;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(X, dX, n * nrhs * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return rs;
}

extern "C" int la_backward_error_dev(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A,
                                     int64_t lda, const double *X, int64_t ldx,
                                     const double *B, int64_t ldb, double *eta)
{
    if (!ctx || !A || !X || !B || !eta) return LA_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || nrhs > 16) return LA_ERR_ARG;
    void *base = nullptr;
    int st = ws_get(ctx, WS_IR, (size_t)n * nrhs * 8 + 256, &base);
    if (st) return st;
    double *R = (double *)base;
    double *norms = (double *)((char *)base + (((size_t)n * nrhs * 8 + 127) & ~(size_t)127));
    if ((st = residual(ctx, n, nrhs, A, lda, X, ldx, B, ldb, R, nrhs, norms, true))) return st;
    double h[4];
    LA_HIP_CHECK(ctx, hipMemcpyAsync(h, norms, 32, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double denom = h[3] * h[1] + h[2];
    *eta = denom > 0 ? h[0] / denom : (h[0] == 0 ? 0.0 : INFINITY);
    return LA_OK;
}
