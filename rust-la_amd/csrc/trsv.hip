// Blocked triangular solves for the LU and Cholesky solve paths.
//
//   MODE_LU_FWD  : L y = b, L unit lower from a packed LU     (lu.rs:257-263)
//   MODE_LU_BWD  : U x = y, U upper from a packed LU          (lu.rs:266-275)
//   MODE_CH_FWD  : L y = b, L non-unit lower (Cholesky)       (cholesky.rs:124-131)
//   MODE_CH_BWD  : L^T x = y with the same L                  (cholesky.rs:134-141)
//
// One launch per triangle.  Workgroups take 64-row blocks in dependency
// order from a ticket counter (so the lowest unfinished block is always
// resident and the launch cannot deadlock, whatever the dispatch order).
// A block first folds every finished block's contribution into private
// partial sums - the bulk of the traffic, overlapped with the blocks still
// in flight - then solves its own 64 x 64 diagonal triangle with one lane
// per row and publishes its slice of x through a per-block flag.
// Arithmetic is FP64; the factor may be FP64 or FP32 (the refinement path).
#include "la_common.hpp"

namespace lahip {

enum { MODE_LU_FWD = 0, MODE_LU_BWD = 1, MODE_CH_FWD = 2, MODE_CH_BWD = 3 };
constexpr int TB = 64;      // rows per block
constexpr int NR = 16;      // right-hand sides per launch

template <typename TF, int MODE>
__global__ __launch_bounds__(256)
void k_trsv(int n, const TF *__restrict__ Tm, int64_t ldt, int nrhs, double *X, int64_t ldx,
            int32_t *flags, int32_t *ticket, int32_t *err, uint64_t spin_ticks)
{
    // No contraction: within a block the sweeps below apply the reference's
    // update order (lu.rs:257-275, cholesky.rs:124-131) with a rounded
    // product and a rounded subtraction per term.
#pragma clang fp contract(off)
    constexpr bool FWD = (MODE == MODE_LU_FWD || MODE == MODE_CH_FWD);
    constexpr bool TRANS = (MODE == MODE_CH_BWD);
    constexpr bool UNIT = (MODE == MODE_LU_FWD);
    __shared__ double xs[TB][NR];
    __shared__ double red[4][TB][NR + 1];
    __shared__ double Ts[TB][TB + 1];
    __shared__ int s_blk;

    const int tid = threadIdx.x;
    const int nblocks = (n + TB - 1) / TB;
    if (tid == 0) s_blk = atomicAdd(ticket, 1);
    __syncthreads();
    const int t = s_blk;
    const int b = FWD ? t : nblocks - 1 - t;
    const int r0 = b * TB;
    const int rb = min(TB, n - r0);

    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;

    // thread mapping for the off-diagonal products
    //   non-transposed: row i = tid/4, 16-column quarter q = tid%4
    //   transposed    : row i = tid%64 (coalesced along L's row), quarter q = tid/64
    const int i = TRANS ? (tid & 63) : (tid >> 2);
    const int q = TRANS ? (tid >> 6) : (tid & 3);

    for (int step = 0; step < t; ++step) {
        const int c = FWD ? step : nblocks - 1 - step;
        const int c0 = c * TB;
        const int cb = min(TB, n - c0);
        if (tid == 0) {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(flags + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                if (wall_clock64() - t0 > spin_ticks) {
                    atomicCAS((int *)err, 0, (int)LA_ERR_TIMEOUT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        for (int e = tid; e < cb * nrhs; e += blockDim.x) {
            int col = e / nrhs, r = e - col * nrhs;
            xs[col][r] = ld_dev(X + (int64_t)(c0 + col) * ldx + r);
        }
        __syncthreads();
        if (i < rb) {
#pragma unroll 4
            for (int cc = 0; cc < 16; ++cc) {
                const int col = q * 16 + cc;
                if (col < cb) {
                    double tv;
                    if (!TRANS) tv = (double)Tm[(int64_t)(r0 + i) * ldt + c0 + col];
                    else        tv = (double)Tm[(int64_t)(c0 + col) * ldt + r0 + i];
#pragma unroll
                    for (int r = 0; r < NR; ++r)
                        if (r < nrhs) acc[r] += tv * xs[col][r];
                }
            }
        }
        __syncthreads();
    }

    // fold partial sums, form the right-hand side of the diagonal solve
#pragma unroll
    for (int r = 0; r < NR; ++r) red[q][i][r] = acc[r];
    for (int e = tid; e < TB * TB; e += blockDim.x) {
        const int a = e / TB, bcol = e - a * TB;
        Ts[a][bcol] = (a < rb && bcol < rb) ? (double)Tm[(int64_t)(r0 + a) * ldt + r0 + bcol] : 0.0;
    }
    __syncthreads();

    // diagonal triangle: wave w solves right-hand sides w, w+4, ...; lane = row
    const int lane = tid & 63, wv = tid >> 6;
    for (int r = wv; r < nrhs; r += 4) {
        double y = 0.0;
        if (lane < rb)
            y = X[(int64_t)(r0 + lane) * ldx + r] -
                ((red[0][lane][r] + red[1][lane][r]) + (red[2][lane][r] + red[3][lane][r]));
        if (FWD) {
            for (int s = 0; s < rb; ++s) {
                if (!UNIT && lane == s) y = y / Ts[s][s];
                const double ys = __shfl(y, s);
                if (lane > s && lane < rb) y -= Ts[lane][s] * ys;
            }
        } else {
            for (int s = rb - 1; s >= 0; --s) {
                if (lane == s) y = y / Ts[s][s];
                const double ys = __shfl(y, s);
                if (lane < s) y -= (TRANS ? Ts[s][lane] : Ts[lane][s]) * ys;
            }
        }
        if (lane < rb) st_dev(X + (int64_t)(r0 + lane) * ldx + r, y);
    }
    stores_done();
    __syncthreads();
    if (tid == 0) st_dev(flags + b, 1);
}

__global__ void k_zero_i32(int32_t *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// X[i][r] = B[piv[i]][r]   (lu.rs:246-254)
__global__ void k_gather_rows(int64_t m, int64_t nx, const uint64_t *piv, const double *B,
                              int64_t ldb, double *X, int64_t ldx)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * nx) return;
    int64_t i = e / nx, r = e - i * nx;
    X[i * ldx + r] = B[(int64_t)piv[i] * ldb + r];
}

template <typename TF, int MODE>
static int launch_trsv_mode(la_ctx *ctx, int64_t n, const TF *Tm, int64_t ldt, int64_t nrhs,
                            double *X, int64_t ldx)
{
    const int nblocks = (int)ceil_div(n, TB);
    // per-block flags followed by the ticket counter
    void *p = nullptr;
    int st = ws_get(ctx, WS_SYNC, ((size_t)nblocks + 1) * sizeof(int32_t), &p);
    if (st) return st;
    int32_t *flags = (int32_t *)p;
    int32_t *ticket = flags + nblocks;
    int32_t *d_err = ctx->d_scratch;   // word 0: sticky device error
    const uint64_t ticks = spin_limit_ticks(ctx);
    // algorithmic: one triangle of the factor + right-hand sides in and out
    ProfScope ps(ctx, LA_K_TRSV, (double)n * n * nrhs,
                 0.5 * n * n * sizeof(TF) + 2.0 * n * nrhs * 8.0);
    for (int64_t r0 = 0; r0 < nrhs; r0 += NR) {
        const int nr = (int)std::min<int64_t>(NR, nrhs - r0);
        hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)ceil_div(nblocks + 1, 256)), dim3(256), 0,
                           ctx->stream, flags, nblocks + 1);
        hipLaunchKernelGGL((k_trsv<TF, MODE>), dim3(nblocks), dim3(256), 0, ctx->stream,
                           (int)n, Tm, ldt, nr, X + r0, ldx, flags, ticket, d_err, ticks);
        LA_HIP_CHECK(ctx, hipGetLastError());
    }
    return LA_OK;
}

template <typename TF>
int trsv_blocked(la_ctx *ctx, char uplo, bool trans, bool unit, int64_t n, const TF *Tm,
                 int64_t ldt, int64_t nrhs, double *X, int64_t ldx)
{
    if (n <= 0 || nrhs <= 0) return LA_OK;
    if (n > INT32_MAX) return set_error(ctx, LA_ERR_ARG, "trsv: n too large");
    if (uplo == 'L' && !trans && unit)
        return launch_trsv_mode<TF, MODE_LU_FWD>(ctx, n, Tm, ldt, nrhs, X, ldx);
    if (uplo == 'U' && !trans && !unit)
        return launch_trsv_mode<TF, MODE_LU_BWD>(ctx, n, Tm, ldt, nrhs, X, ldx);
    if (uplo == 'L' && !trans && !unit)
        return launch_trsv_mode<TF, MODE_CH_FWD>(ctx, n, Tm, ldt, nrhs, X, ldx);
    if (uplo == 'L' && trans && !unit)
        return launch_trsv_mode<TF, MODE_CH_BWD>(ctx, n, Tm, ldt, nrhs, X, ldx);
    return set_error(ctx, LA_ERR_ARG, "trsv: unsupported triangle");
}

int gather_rows(la_ctx *ctx, int64_t m, int64_t nx, const uint64_t *d_piv, const double *B,
                int64_t ldb, double *X, int64_t ldx)
{
    if (m <= 0 || nx <= 0) return LA_OK;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)ceil_div(m * nx, 256)), dim3(256), 0,
                       ctx->stream, m, nx, d_piv, B, ldb, X, ldx);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template int trsv_blocked<double>(la_ctx *, char, bool, bool, int64_t, const double *, int64_t,
                                  int64_t, double *, int64_t);
template int trsv_blocked<float>(la_ctx *, char, bool, bool, int64_t, const float *, int64_t,
                                 int64_t, double *, int64_t);

}  // namespace lahip
