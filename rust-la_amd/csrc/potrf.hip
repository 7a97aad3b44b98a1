// Blocked right-looking Cholesky of a single SPD matrix (row-major, lower).
//
// Semantics of CholeskyDecomposition::new (src/decomp/cholesky.rs:56-110):
//   - any exactly asymmetric pair A[k][j] != A[j][k] (NaN included) -> None
//   - a diagonal pivot d <= 0 -> None (a NaN pivot is NOT rejected: d <= 0 is
//     false for NaN, so the factor carries the NaN on, as the reference does)
//   - the result is the full n x n L with the strict upper triangle zeroed.
// Two-level blocking.  Per 64-column panel: the diagonal block is factored
// in LDS by one workgroup and the panel below it is solved against L11^T row
// by row; the rest of its 256-column outer block gets the panel's rank-64
// update, and once the outer block is done the trailing lower triangle gets
// one rank-256 update on FP64 MFMA (the GEMM kernel with B = panel^T, tiles
// above the diagonal skipped).  A rank-64 update of the whole trailing
// triangle per panel would read and write it four times as often.
#include "la_common.hpp"

namespace lahip {

constexpr int PB = 64;
constexpr int PB_OUT = 256;     // outer block: K of the trailing update

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
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        int r = e >> 6, c = e & 63;
        t1[r][c] = (r0 + r < n && c0 + c < n) ? A[(int64_t)(r0 + r) * lda + c0 + c] : 0.0;
        t2[r][c] = (c0 + r < n && r0 + c < n) ? A[(int64_t)(c0 + r) * lda + r0 + c] : 0.0;
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

// Factor the kb x kb diagonal block at A (kb <= 64) in LDS.
// status bit 2: non-positive pivot (d <= 0).
__device__ inline double potf2_bcast(double v, int lane)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, lane);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The kb x kb diagonal block (kb <= 64), right-looking, by one wavefront:
// lane r keeps row r in registers (fully unrolled, so every index is
// static); per column t the pivot and each a[c][t] reach all lanes by
// readlane.  Per element the same updates in the same order as a
// workgroup-wide LDS version: a[r][t] /= sqrt(d_t), then
// a[r][c] -= a[r][t] * a[c][t] for t < c <= r.  A pivot d <= 0 sets bit 2 of
// the flag (NaN is not rejected and propagates, as in the reference).
__global__ __launch_bounds__(64)
void k_potf2(double *A, int64_t lda, int kb, int32_t *flag)
{
    const int r = threadIdx.x;
    double a[64];
    const int rc = min(r, kb - 1);
#pragma unroll
    for (int c = 0; c < 64; ++c) {
        const double v = A[(int64_t)rc * lda + min(c, kb - 1)];
        a[c] = (r < kb && c <= r) ? v : 0.0;
    }
    bool bad = false;
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        if (t < kb) {
            const double d = potf2_bcast(a[t], t);
            if (d <= 0.0) bad = true;
            const double dt = sqrt(d);
            if (r == t) a[t] = dt;
            if (r > t) a[t] = a[t] / dt;
#pragma unroll
            for (int c = t + 1; c < 64; ++c) {
                if (c < kb) {
                    const double act = potf2_bcast(a[t], c);
                    if (c <= r) a[c] -= a[t] * act;
                }
            }
        }
    }
    if (r < kb) {
#pragma unroll
        for (int c = 0; c < 64; ++c)
            if (c < kb) A[(int64_t)r * lda + c] = (c <= r) ? a[c] : 0.0;
    }
    if (r == 0 && bad) atomicOr((int *)flag, 2);
}

// Panel solve: rows of B (nrows x kb, row-major) <- B * L11^-T with L11 the
// kb x kb lower factor at Lm.  A workgroup stages 64 rows of B in LDS and
// sweeps the columns: x_t = b_t / L[t][t], then b_c -= x_t * L[c][t] for c > t.
__global__ __launch_bounds__(256)
void k_trsm_right_lt(const double *Lm, int64_t ldl, int kb, double *B, int64_t ldb, int nrows)
{
    __shared__ double Ls[64][65];
    __shared__ double Bs[64][65];
    const int tid = threadIdx.x;
    const int row0 = blockIdx.x * 64;
    // unconditional loads from clamped positions (issued together), then
    // the out-of-range elements are zeroed
    double lv[16], bv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, r = e >> 6, c = e & 63;
        lv[k] = Lm[(int64_t)min(r, kb - 1) * ldl + min(c, kb - 1)];
        bv[k] = B[(int64_t)min(row0 + r, nrows - 1) * ldb + min(c, kb - 1)];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, r = e >> 6, c = e & 63;
        Ls[r][c] = (r < kb && c <= r) ? lv[k] : 0.0;
        Bs[r][c] = (row0 + r < nrows && c < kb) ? bv[k] : 0.0;
    }
    __syncthreads();
    const int r = tid & 63, cq = tid >> 6;
    for (int t = 0; t < kb; ++t) {
        if (cq == 0) Bs[r][t] = Bs[r][t] / Ls[t][t];
        __syncthreads();
        const double xt = Bs[r][t];
        for (int c = t + 1 + ((cq - (t + 1)) & 3); c < kb; c += 4) Bs[r][c] -= xt * Ls[c][t];
        __syncthreads();
    }
    for (int e = tid; e < 64 * 64; e += 256) {
        int rr = e >> 6, c = e & 63;
        if (row0 + rr < nrows && c < kb) B[(int64_t)(row0 + rr) * ldb + c] = Bs[rr][c];
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
    for (int64_t K0 = 0; K0 < n; K0 += PB_OUT) {
        const int64_t K1 = std::min<int64_t>(n, K0 + PB_OUT);
        for (int64_t k = K0; k < K1; k += PB) {
            const int kb = (int)std::min<int64_t>(PB, K1 - k);
            double *Akk = A + k * lda + k;
            hipLaunchKernelGGL(k_potf2, dim3(1), dim3(64), 0, ctx->stream, Akk, lda, kb, flag);
            const int64_t below = n - k - kb;
            if (below <= 0) continue;
            double *A21 = A + (k + kb) * lda + k;
            hipLaunchKernelGGL(k_trsm_right_lt, dim3((unsigned)ceil_div(below, 64)), dim3(256), 0,
                               ctx->stream, (const double *)Akk, lda, kb, A21, lda, (int)below);
            LA_HIP_CHECK(ctx, hipGetLastError());
            // the rest of the outer block: columns [k + kb, K1), rows below
            // them (the tiles above the block's diagonal skipped)
            const int64_t cols = K1 - (k + kb);
            if (cols > 0) {
                int st = gemm_rm<double>(ctx, below, cols, kb, -1.0, A21, lda, A21, lda, 1.0,
                                         A + (k + kb) * lda + k + kb, lda, true, true);
                if (st) return st;
            }
        }
        // the trailing lower triangle, rank K1 - K0
        const int64_t rest = n - K1;
        if (rest > 0) {
            const double *L21 = A + K1 * lda + K0;
            int st = gemm_rm<double>(ctx, rest, rest, K1 - K0, -1.0, L21, lda, L21, lda, 1.0,
                                     A + K1 * lda + K1, lda, true, true);
            if (st) return st;
        }
    }
    hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0,
                       ctx->stream, n, A, lda);
    hipLaunchKernelGGL(k_flag_to_status, dim3(1), dim3(1), 0, ctx->stream, flag, d_status);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

}  // namespace lahip
