// Blocked right-looking Cholesky of a single SPD matrix (row-major, lower).
//
// Semantics of CholeskyDecomposition::new (src/decomp/cholesky.rs:56-110):
//   - any exactly asymmetric pair A[k][j] != A[j][k] (NaN included) -> None
//   - a diagonal pivot d <= 0 -> None (a NaN pivot is NOT rejected: d <= 0 is
//     false for NaN, so the factor carries the NaN on, as the reference does)
//   - the result is the full n x n L with the strict upper triangle zeroed.
// Three-level blocking.  Per 64-column panel: the diagonal block is factored
// by one wavefront and the panel below it is solved against L11^T one row
// per lane; the rest of its 256-column middle block gets the panel's
// rank-64 update.  Once a middle block is done, the rest of its 1024-column
// outer block gets one rank-256 update, and once an outer block is done the
// trailing lower triangle gets one rank-1024 update.  Every update is the
// FP64 MFMA GEMM with B = panel^T and the tiles above the diagonal skipped;
// wider ranks amortise reading and writing the trailing triangle.
#include "la_common.hpp"

namespace lahip {

constexpr int PB = 64;          // panel: factored and solved by latency-bound kernels
constexpr int PB_MID = 256;     // middle block: rank of the updates inside an outer block
constexpr int PB_OUT = 1024;    // outer block: rank of the trailing update

// status: bit 1 = asymmetric pair seen.  Compares 64 x 64 tile (bi, bj),
// bi > bj or bi == bj, against its mirror through LDS.
__global__ __launch_bounds__(256)
void k_symcheck(int n, const double *A, int64_t lda, int32_t *flag)
{
    __shared__ double t1[64][65];
    __shared__ double t2[64][65];
    // decode lower-triangular tile index
    const int64_t idx = blockIdx.x;
    int bi = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((int64_t)bi * (bi + 1) / 2 > idx) --bi;
    while ((int64_t)(bi + 1) * (bi + 2) / 2 <= idx) ++bi;
    const int bj = (int)(idx - (int64_t)bi * (bi + 1) / 2);
    const int r0 = bi * 64, c0 = bj * 64;
    // both tiles' 32 loads per thread from clamped (in-bounds) positions,
    // issued together; positions outside the matrix are never compared
    double v1[16], v2[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + 256 * k, r = e >> 6, c = e & 63;
        v1[k] = A[(int64_t)min(r0 + r, n - 1) * lda + min(c0 + c, n - 1)];
        v2[k] = A[(int64_t)min(c0 + r, n - 1) * lda + min(r0 + c, n - 1)];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + 256 * k, r = e >> 6, c = e & 63;
        t1[r][c] = v1[k];
        t2[r][c] = v2[k];
    }
    __syncthreads();
    int bad = 0;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        int r = e >> 6, c = e & 63;
        if (r0 + r < n && c0 + c < n && (r0 + r) > (c0 + c))
            if (t1[r][c] != t2[c][r]) bad = 1;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr((int *)flag, 1);
}

// v of lane `lane` (a wave-uniform index) to every lane of the wavefront
__device__ inline double potf2_bcast(double v, int lane)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, lane);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The kb x kb diagonal block (kb <= 64) by one wavefront, lane r owning row
// r, eight columns at a time.  Element (r, c) is
//   l[r][c] = (a[r][c] - sum_{k<c} l[r][k] * l[c][k]) / sqrt(d_c)
// with the sum taken as fma(-l[r][k], l[c][k], .) in k order and d_c the
// diagonal element after its sum: the same operations in the same order as
// a right-looking sweep.  For columns c0..c0+7 the terms k < c0 (finished
// columns, LDS) go to eight independent register accumulators; the terms
// inside the block follow column by column, l[c][k] taken from lane c by
// readlane.  Eight independent chains keep the FP64 pipeline busy where one
// dot product per column would wait out the latency of every fma.  A pivot
// d <= 0 sets bit 2 of the flag (NaN is not rejected and propagates, as in
// the reference).
__global__ __launch_bounds__(64)
void k_potf2(double *A, int64_t lda, int kb, int32_t *flag)
{
    __shared__ double a[64][65];
    const int r = threadIdx.x;
    const int rc = min(r, kb - 1);
    {
        double v[64];
#pragma unroll
        for (int i = 0; i < 64; ++i) v[i] = A[(int64_t)min(i, kb - 1) * lda + rc];
#pragma unroll
        for (int i = 0; i < 64; ++i) a[i][r] = (i < kb && r <= i) ? v[i] : 0.0;
    }
    __syncthreads();
    bool bad = false;
    for (int c0 = 0; c0 < kb; c0 += 8) {
        double s[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] = a[r][c0 + j];
        for (int k = 0; k < c0; ++k) {
            const double lrk = a[r][k];
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] = fma(-lrk, a[c0 + j][k], s[j]);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int c = c0 + t;
            if (c < kb) {
                const double d = potf2_bcast(s[t], c);
                if (d <= 0.0) bad = true;
                const double dt = sqrt(d);
                s[t] = (r == c) ? dt : s[t] / dt;
#pragma unroll
                for (int j = t + 1; j < 8; ++j) s[j] = fma(-s[t], potf2_bcast(s[t], c0 + j), s[j]);
            }
        }
        // entries right of the diagonal (lanes r < c) are scratch: never
        // read as l[.][.] and not stored
#pragma unroll
        for (int j = 0; j < 8; ++j) a[r][c0 + j] = s[j];
        __syncthreads();
    }
    if (r < kb) {
        for (int i = 0; i < kb; ++i) A[(int64_t)i * lda + r] = (r <= i) ? a[i][r] : 0.0;
    }
    if (r == 0 && bad) atomicOr((int *)flag, 2);
}

// Panel solve: rows of B (nrows x kb, row-major) <- B * L11^-T with L11 the
// kb x kb lower factor at Lm.  Rows are independent: a one-wavefront
// workgroup stages 64 rows and L11 (transposed, Lt[k][c] = L[c][k]) in LDS
// and lane r solves its own row,
//   x_c = (b_c - sum_{k<c} x_k * L[c][k]) / L[c][c]
// with the sum taken as fma(-x_k, L[c][k], .) in k order (per element the
// operations of the right-looking sweep).  Eight columns at a time: the
// terms of finished columns go to eight independent register accumulators,
// then the block is solved column by column.  No barrier inside the solve.
__global__ __launch_bounds__(64)
void k_trsm_right_lt(const double *Lm, int64_t ldl, int kb, double *B, int64_t ldb, int nrows)
{
    __shared__ double Lt[64][65];
    __shared__ double Bs[64][65];
    const int r = threadIdx.x;
    const int row0 = blockIdx.x * 64;
    const int rc = min(r, kb - 1);
    {
        double v[64];
#pragma unroll
        for (int i = 0; i < 64; ++i) v[i] = Lm[(int64_t)min(i, kb - 1) * ldl + rc];
#pragma unroll
        for (int i = 0; i < 64; ++i) Lt[r][i] = (i < kb && r <= i) ? v[i] : 0.0;
    }
    {
        double v[64];
#pragma unroll
        for (int i = 0; i < 64; ++i) v[i] = B[(int64_t)min(row0 + i, nrows - 1) * ldb + rc];
#pragma unroll
        for (int i = 0; i < 64; ++i) Bs[i][r] = (row0 + i < nrows && r < kb) ? v[i] : 0.0;
    }
    __syncthreads();
    for (int c0 = 0; c0 < kb; c0 += 8) {
        double s[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] = Bs[r][c0 + j];
        for (int k = 0; k < c0; ++k) {
            const double xk = Bs[r][k];
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] = fma(-xk, Lt[k][c0 + j], s[j]);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int c = c0 + t;
            if (c < kb) {
                s[t] = s[t] / Lt[c][c];
#pragma unroll
                for (int j = t + 1; j < 8; ++j) s[j] = fma(-s[t], Lt[c][c0 + j], s[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[r][c0 + j] = s[j];
    }
    __syncthreads();
    if (r < kb) {
        const int nr = min(64, nrows - row0);
        for (int i = 0; i < nr; ++i) B[(int64_t)(row0 + i) * ldb + r] = Bs[i][r];
    }
}

__global__ void k_zero_upper(int64_t n, double *A, int64_t lda)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    int64_t r = e / n, c = e - r * n;
    if (c > r) A[r * lda + c] = 0.0;
}

__global__ void k_flag_to_status(const int32_t *flag, int32_t *status)
{
    int f = *flag;
    *status = (f & 1) ? LA_NOT_SYMMETRIC : ((f & 2) ? LA_NOT_SPD : LA_OK);
}

int potrf_blocked(la_ctx *ctx, int64_t n, double *A, int64_t lda, int32_t *d_status)
{
    if (n <= 0) return LA_OK;
    if (n > INT32_MAX / 2) return set_error(ctx, LA_ERR_ARG, "potrf: n too large");
    int32_t *flag = ctx->d_scratch + SCR_CHOL;
    int fst = fill_i32(ctx, flag, 1, 0);
    if (fst) return fst;
    const int64_t nbt = ceil_div(n, 64);
    hipLaunchKernelGGL(k_symcheck, dim3((unsigned)(nbt * (nbt + 1) / 2)), dim3(256), 0,
                       ctx->stream, (int)n, (const double *)A, lda, flag);
    LA_HIP_CHECK(ctx, hipGetLastError());
    // lower-triangular update of the block whose top-left corner (r0, r0)
    // lies on the diagonal: rows r0..n-1, columns r0..c1-1, minus the
    // product of the factored columns k0..k1-1 of those rows with the
    // transpose of their first c1 - r0 rows
    auto syrk = [&](int64_t r0, int64_t c1, int64_t k0, int64_t k1) -> int {
        const int64_t rows = n - r0, cols = c1 - r0;
        if (rows <= 0 || cols <= 0) return LA_OK;
        const double *P = A + r0 * lda + k0;
        return gemm_rm<double>(ctx, rows, cols, k1 - k0, -1.0, P, lda, P, lda, 1.0,
                               A + r0 * lda + r0, lda, true, true);
    };
    for (int64_t K0 = 0; K0 < n; K0 += PB_OUT) {
        const int64_t K1 = std::min<int64_t>(n, K0 + PB_OUT);
        for (int64_t M0 = K0; M0 < K1; M0 += PB_MID) {
            const int64_t M1 = std::min<int64_t>(K1, M0 + PB_MID);
            for (int64_t k = M0; k < M1; k += PB) {
                const int kb = (int)std::min<int64_t>(PB, M1 - k);
                double *Akk = A + k * lda + k;
                hipLaunchKernelGGL(k_potf2, dim3(1), dim3(64), 0, ctx->stream, Akk, lda, kb,
                                   flag);
                const int64_t below = n - k - kb;
                if (below <= 0) continue;
                hipLaunchKernelGGL(k_trsm_right_lt, dim3((unsigned)ceil_div(below, 64)), dim3(64),
                                   0, ctx->stream, (const double *)Akk, lda, kb,
                                   A + (k + kb) * lda + k, lda, (int)below);
                LA_HIP_CHECK(ctx, hipGetLastError());
                int st = syrk(k + kb, M1, k, k + kb);          // rest of the middle block
                if (st) return st;
            }
            int st = syrk(M1, K1, M0, M1);                      // rest of the outer block
            if (st) return st;
        }
        int st = syrk(K1, n, K0, K1);                           // the trailing triangle
        if (st) return st;
    }
    hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0,
                       ctx->stream, n, A, lda);
    hipLaunchKernelGGL(k_flag_to_status, dim3(1), dim3(1), 0, ctx->stream, flag, d_status);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

}  // namespace lahip
