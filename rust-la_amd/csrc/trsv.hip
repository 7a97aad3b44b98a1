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

// Inverse of the rb x rb diagonal block held in Ts (LDS, FP64), written to
// Xi (LDS).  Lower modes sweep rows downwards, upper modes upwards; for the
// transposed Cholesky mode the upper triangle is L^T (U[i][t] = L[t][i]).
// Four threads share each column's dot product.
template <int MODE>
__device__ void invert_diag_block(const double (*Ts)[TB + 1], int rb, double (*Xi)[TB + 1],
                                  double (*part)[TB])
{
    constexpr bool LOWER = (MODE == MODE_LU_FWD || MODE == MODE_CH_FWD);
    constexpr bool TRANS = (MODE == MODE_CH_BWD);
    constexpr bool UNIT = (MODE == MODE_LU_FWD);
    const int tid = threadIdx.x, c = tid & 63, q = tid >> 6;
    for (int e = tid; e < TB * TB; e += blockDim.x) Xi[e / TB][e % TB] = 0.0;
    __syncthreads();
    for (int step = 0; step < rb; ++step) {
        const int i = LOWER ? step : rb - 1 - step;
        double s = 0.0;
        if (LOWER) {
            if (c < i)
                for (int t = c + q; t < i; t += 4) s += Ts[i][t] * Xi[t][c];
        } else {
            if (c > i && c < rb)
                for (int t = i + 1 + q; t <= c; t += 4)
                    s += (TRANS ? Ts[t][i] : Ts[i][t]) * Xi[t][c];
        }
        part[q][c] = s;
        __syncthreads();
        if (q == 0 && c < rb) {
            const double d = UNIT ? 1.0 : Ts[i][i];
            const double tot = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
            if (c == i) Xi[i][c] = 1.0 / d;
            else if (LOWER ? (c < i) : (c > i)) Xi[i][c] = -tot / d;
        }
        __syncthreads();
    }
}

// K factor tiles of blocks first .. first+count-1 (dependency order) for
// this thread's row ir and sixteen columns q*16 .. q*16+15, into registers.
// Every load is unconditional so that they issue back to back (a load under
// a per-element branch waits for its data before the branch joins): the
// column is clamped into the factor instead, and tiles past count repeat
// the last one.  A clamped column meets a padding row of x, which is
// published as zero.
template <bool FWD, bool TRANS, int K, typename TF>
__device__ inline void load_tiles(TF (&tv)[K][16], const TF *__restrict__ Tm, int64_t ldt, int n,
                                  int nblocks, int r0, int ir, int q, int first, int count)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int step = first + min(k, count - 1);
        const int c = FWD ? step : nblocks - 1 - step;
        const int c0 = c * TB, cb = min(TB, n - c0);
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int col = min(q * 16 + cc, cb - 1);
            tv[k][cc] = !TRANS ? Tm[(int64_t)(r0 + ir) * ldt + c0 + col]
                               : Tm[(int64_t)(c0 + col) * ldt + r0 + ir];
        }
    }
}

// Gather the tagged x words of blocks first .. first+count-1 into xs32:
// sixteen independent loads in flight per thread, then validate; only
// words that are not yet current are polled again.
template <bool FWD>
__device__ inline void gather_words(uint32_t *xs32, const uint64_t *xw, int words_per_block,
                                    int nblocks, int first, int count, uint32_t epoch,
                                    int32_t *err, uint64_t spin_ticks)
{
    const int tid = threadIdx.x;
    const int nwords = count * words_per_block;
    for (int e0 = tid; e0 < nwords; e0 += 16 * 256) {
        uint64_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = min(e0 + k * 256, nwords - 1);
            const int kb = e / words_per_block, wi = e - kb * words_per_block;
            const int step = first + kb;
            const int c = FWD ? step : nblocks - 1 - step;
            v[k] = ld_dev(xw + (size_t)c * words_per_block + wi);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = e0 + k * 256;
            if (e >= nwords) continue;
            if ((uint32_t)(v[k] >> 32) != epoch) {
                const int kb = e / words_per_block, wi = e - kb * words_per_block;
                const int step = first + kb;
                const int c = FWD ? step : nblocks - 1 - step;
                const uint64_t *src = xw + (size_t)c * words_per_block + wi;
                const uint64_t t0 = wall_clock64();
                for (uint32_t spin = 0;; ++spin) {
                    __builtin_amdgcn_s_sleep(1);
                    v[k] = ld_dev(src);
                    if ((uint32_t)(v[k] >> 32) == epoch) break;
                    // a wait that failed anywhere ends every other wait quickly
                    if ((spin & 63) == 63 && ld_dev(err) != 0) break;
                    if (wall_clock64() - t0 > spin_ticks) {
                        atomicCAS((int *)err, 0, (int)LA_ERR_TIMEOUT);
                        break;
                    }
                }
            }
            xs32[e] = (uint32_t)v[k];
        }
    }
}

// acc[r] += sum over the first count tiles and this thread's sixteen
// columns of tile * x, tiles and columns ascending.
template <int NRB, int K, typename TF>
__device__ inline void fold_tiles(double (&acc)[NRB], const TF (&tv)[K][16], const double *xs,
                                  int nrhs, int q, int count)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k < count) {
            const double *xk = xs + (size_t)k * TB * nrhs;
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) {
                const int col = q * 16 + cc;
#pragma unroll
                for (int r = 0; r < NRB; ++r)
                    if (r < nrhs)
                        acc[r] = __builtin_fma((double)tv[k][cc], xk[col * nrhs + r], acc[r]);
            }
        }
    }
}

template <typename TF, int MODE, int NRB>
__global__ __launch_bounds__(256)
void k_trsv(int n, const TF *__restrict__ Tm, int64_t ldt, int nrhs_arg, double *X, int64_t ldx,
            uint64_t *xw, int32_t *ticket, uint32_t epoch, int32_t *err, uint64_t spin_ticks)
{
    // Single-block solves (n <= 64) keep the reference's update order with a
    // rounded product and a rounded subtraction per term (lu.rs:257-275,
    // cholesky.rs:124-131); larger solves multiply by inverted diagonal blocks.
#pragma clang fp contract(off)
    constexpr bool FWD = (MODE == MODE_LU_FWD || MODE == MODE_CH_FWD);
    constexpr bool TRANS = (MODE == MODE_CH_BWD);
    constexpr bool UNIT = (MODE == MODE_LU_FWD);
    // scratch: the diagonal block (during inversion), then the x slices of
    // finished blocks, then the partial-sum reduction
    __shared__ __align__(16) double scratch[TB * (TB + 1)];
    __shared__ double Xi[TB][TB + 1];
    __shared__ double part[4][TB];
    __shared__ int s_blk;
    double (*Ts)[TB + 1] = reinterpret_cast<double (*)[TB + 1]>(scratch);
    uint32_t *xs32 = reinterpret_cast<uint32_t *>(scratch);
    const double *xs = scratch;                              // [chunk][64][nrhs]
    double (*red)[TB][NRB] = reinterpret_cast<double (*)[TB][NRB]>(scratch);  // [4][64][NRB]

    // NRB = 1 compiles the single right-hand side without per-column
    // guards; NRB = NR serves 2 .. NR columns
    const int nrhs = (NRB == 1) ? 1 : nrhs_arg;
    const int tid = threadIdx.x;
    const int nblocks = (n + TB - 1) / TB;
    if (tid == 0) s_blk = atomicAdd(ticket, 1);
    __syncthreads();
    const int t = s_blk;
    const int b = FWD ? t : nblocks - 1 - t;
    const int r0 = b * TB;
    const int rb = min(TB, n - r0);
    const bool single = (nblocks == 1);

    // invert the diagonal block first: this overlaps the wait for the
    // blocks this one depends on
    if (!single) {
        // sixteen unconditional (clamped) loads per thread, then the stores
        TF dv[TB * TB / 256];
#pragma unroll
        for (int k = 0; k < TB * TB / 256; ++k) {
            const int e = tid + 256 * k, a = e / TB, c = e - a * TB;
            dv[k] = Tm[(int64_t)(r0 + min(a, rb - 1)) * ldt + r0 + min(c, rb - 1)];
        }
#pragma unroll
        for (int k = 0; k < TB * TB / 256; ++k) {
            const int e = tid + 256 * k, a = e / TB, c = e - a * TB;
            Ts[a][c] = (a < rb && c < rb) ? (double)dv[k] : 0.0;
        }
        __syncthreads();
        invert_diag_block<MODE>(Ts, rb, Xi, part);
        __syncthreads();
    }

    double acc[NRB];
#pragma unroll
    for (int r = 0; r < NRB; ++r) acc[r] = 0.0;
    const int i = TRANS ? (tid & 63) : (tid >> 2);
    const int q = TRANS ? (tid >> 6) : (tid & 3);
    // rows past rb load clamped rows and are never stored
    const int ir = min(i, rb - 1);
    double bpre = 0.0;
    const int ba = tid / max(nrhs, 1), br = tid - ba * max(nrhs, 1);
    if (tid < TB * nrhs && ba < rb) bpre = X[(int64_t)(r0 + ba) * ldx + br];
    // The last two dependencies (blocks t-2, t-1) form the final chunk, on
    // the critical path: their tiles do not depend on x and are loaded now,
    // before any wait.
    const int nlast = min(t, 2);
    TF tpre[2][16];
    load_tiles<FWD, TRANS>(tpre, Tm, ldt, n, nblocks, r0, ir, q, max(t - nlast, 0), max(nlast, 1));
    // The other dependencies go in chunks whose factor tiles are loaded
    // together before its x words are polled; small chunks keep a
    // workgroup from sitting on a large batch of finished blocks while its
    // immediate predecessor completes.
    constexpr int CHT = sizeof(TF) == 8 ? 4 : 8;
    const int CH = max(1, min(CHT, 64 / max(nrhs, 1)));
    const int words_per_block = TB * nrhs * 2;

    int done = 0;
    while (done < t - nlast) {
        const int chunk = min(CH, t - nlast - done);
        TF tv[CHT][16];
        load_tiles<FWD, TRANS>(tv, Tm, ldt, n, nblocks, r0, ir, q, done, chunk);
        gather_words<FWD>(xs32, xw, words_per_block, nblocks, done, chunk, epoch, err, spin_ticks);
        __syncthreads();
        fold_tiles<NRB>(acc, tv, xs, nrhs, q, chunk);
        __syncthreads();
        done += chunk;
    }
    if (nlast > 0) {
        gather_words<FWD>(xs32, xw, words_per_block, nblocks, done, nlast, epoch, err, spin_ticks);
        __syncthreads();
        fold_tiles<NRB>(acc, tpre, xs, nrhs, q, nlast);
        __syncthreads();
    }

    // reduce partial sums over the four column groups
#pragma unroll
    for (int r = 0; r < NRB; ++r) red[q][i][r] = acc[r];
    __syncthreads();

    const int lane = tid & 63, wv = tid >> 6;
    if (single) {
        // substitution in the reference's order (no contraction)
        for (int e = tid; e < TB * TB; e += blockDim.x) {
            const int a = e / TB, c = e - a * TB;
            Xi[a][c] = (a < rb && c < rb) ? (double)Tm[(int64_t)(r0 + a) * ldt + r0 + c] : 0.0;
        }
        __syncthreads();
        for (int r = wv; r < nrhs; r += 4) {
            double y = (lane < rb) ? X[(int64_t)(r0 + lane) * ldx + r] : 0.0;
            if (FWD) {
                for (int s = 0; s < rb; ++s) {
                    if (!UNIT && lane == s) y = y / Xi[s][s];
                    const double ys = __shfl(y, s);
                    if (lane > s && lane < rb) y -= Xi[lane][s] * ys;
                }
            } else {
                for (int s = rb - 1; s >= 0; --s) {
                    if (lane == s) y = y / Xi[s][s];
                    const double ys = __shfl(y, s);
                    if (lane < s) y -= (TRANS ? Xi[s][lane] : Xi[lane][s]) * ys;
                }
            }
            if (lane < rb) X[(int64_t)(r0 + lane) * ldx + r] = y;
        }
        return;
    }
    // y = b - partial sums (into part-like scratch: reuse red[0])
    __shared__ double ys[TB][NRB];
    for (int e = tid; e < TB * nrhs; e += blockDim.x) {
        const int a = e / nrhs, r = e - a * nrhs;
        double y = 0.0;
        if (a < rb)
            y = (e == tid ? bpre : X[(int64_t)(r0 + a) * ldx + r]) -
                ((red[0][a][r] + red[1][a][r]) + (red[2][a][r] + red[3][a][r]));
        ys[a][r] = y;
    }
    __syncthreads();
    // x = inv(T_bb) y, published as tagged words and stored to X
    const uint64_t tag = (uint64_t)epoch << 32;
    // every one of the 64 rows is published (padding rows as zero): readers
    // fetch whole blocks
    for (int e = tid; e < TB * nrhs; e += blockDim.x) {
        const int a = e / nrhs, r = e - a * nrhs;
        double s = 0.0;
        if (a < rb) {
            // fixed trip count (Xi and ys are zero past rb, each such term
            // leaves s unchanged) so the LDS reads are issued ahead of the
            // FMA chain; the order of the sum is c ascending as before
#pragma unroll 16
            for (int c = 0; c < TB; ++c) s = __builtin_fma(Xi[a][c], ys[c][r], s);
            X[(int64_t)(r0 + a) * ldx + r] = s;
        }
        const uint64_t bits = (uint64_t)__double_as_longlong(s);
        uint64_t *dst = xw + (size_t)b * words_per_block + ((size_t)a * nrhs + r) * 2;
        st_dev(dst + 0, tag | (bits & 0xffffffffull));
        st_dev(dst + 1, tag | (bits >> 32));
    }
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
    // ticket word, then the tagged x words (2 per value): blocks x 64 x NR x 2
    const size_t words = (size_t)nblocks * TB * NR * 2;
    const size_t bytes = 256 + words * sizeof(uint64_t);
    void *p = nullptr;
    const size_t before = ctx->ws[WS_SYNC].size;
    int st = ws_get(ctx, WS_SYNC, bytes, &p);
    if (st) return st;
    int32_t *ticket = (int32_t *)p;
    uint64_t *xw = (uint64_t *)((char *)p + 256);
    if (ctx->ws[WS_SYNC].size != before) {
        // fresh buffer: tag 0 never matches a live epoch (epochs start at 1)
        if ((st = fill_u64(ctx, (uint64_t *)p, (int64_t)(ctx->ws[WS_SYNC].size / 8), 0)))
            return st;
        ctx->trsv_epoch = 0;
    }
    int32_t *d_err = ctx->d_scratch + SCR_ERR;
    const uint64_t ticks = spin_limit_ticks(ctx);
    // algorithmic: one triangle of the factor + right-hand sides in and out
    ProfScope ps(ctx, LA_K_TRSV, (double)n * n * nrhs,
                 0.5 * n * n * sizeof(TF) + 2.0 * n * nrhs * 8.0);
    for (int64_t rr = 0; rr < nrhs; rr += NR) {
        const int nr = (int)std::min<int64_t>(NR, nrhs - rr);
        uint32_t epoch = ++ctx->trsv_epoch;
        if (epoch == 0) epoch = ++ctx->trsv_epoch;
        if ((st = fill_i32(ctx, ticket, 1, 0))) return st;
        if (nr == 1)
            hipLaunchKernelGGL((k_trsv<TF, MODE, 1>), dim3(nblocks), dim3(256), 0, ctx->stream,
                               (int)n, Tm, ldt, nr, X + rr, ldx, xw, ticket, epoch, d_err, ticks);
        else
            hipLaunchKernelGGL((k_trsv<TF, MODE, NR>), dim3(nblocks), dim3(256), 0, ctx->stream,
                               (int)n, Tm, ldt, nr, X + rr, ldx, xw, ticket, epoch, d_err, ticks);
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

// ------------------------------------------- many right-hand sides (TRSM)
// For nrhs > NR the substitution becomes a blocked TRSM on the FP64 MFMA
// GEMM: the 64 x 64 diagonal blocks are inverted once (one workgroup each,
// in parallel), then, in dependency order, per super-block of eight blocks
//   per block b:  Y = inv(T_bb) X_b;  X_(rest of super-block) -= T_.,b Y;  X_b = Y
//   then          X_(rest of the triangle) -= T_.,super-block X_super-block
// (right-looking: the update of the remaining rows is one GEMM of K = 512,
// which runs the MFMA kernel far closer to its peak than eight of K = 64).
template <int MODE, typename TF>
__global__ __launch_bounds__(256)
void k_diag_inverses(int n, const TF *__restrict__ Tm, int64_t ldt, double *Dinv)
{
    __shared__ double Ts[TB][TB + 1];
    __shared__ double Xi[TB][TB + 1];
    __shared__ double part[4][TB];
    const int b = blockIdx.x, r0 = b * TB, rb = min(TB, n - r0), tid = threadIdx.x;
    TF dv[TB * TB / 256];
#pragma unroll
    for (int k = 0; k < TB * TB / 256; ++k) {
        const int e = tid + 256 * k, a = e / TB, c = e - a * TB;
        dv[k] = Tm[(int64_t)(r0 + min(a, rb - 1)) * ldt + r0 + min(c, rb - 1)];
    }
#pragma unroll
    for (int k = 0; k < TB * TB / 256; ++k) {
        const int e = tid + 256 * k, a = e / TB, c = e - a * TB;
        Ts[a][c] = (a < rb && c < rb) ? (double)dv[k] : 0.0;
    }
    __syncthreads();
    invert_diag_block<MODE>(Ts, rb, Xi, part);
    __syncthreads();
    double *out = Dinv + (size_t)b * TB * TB;
    for (int e = tid; e < TB * TB; e += blockDim.x) out[e] = Xi[e / TB][e % TB];
}

__global__ void k_copy_rows(int64_t rows, int64_t cols, const double *src, int64_t lds,
                            double *dst, int64_t ldd)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int64_t r = e / cols, c = e - r * cols;
    dst[r * ldd + c] = src[r * lds + c];
}

// X (n x nrhs, ld ldx) <- T^-1 X for the LU triangles ('L' unit lower,
// 'U' upper) of a packed FP64 factor.
int trsm_blocked(la_ctx *ctx, char uplo, int64_t n, const double *Tm, int64_t ldt, int64_t nrhs,
                 double *X, int64_t ldx)
{
    if (n <= 0 || nrhs <= 0) return LA_OK;
    const int64_t nblocks = ceil_div(n, TB);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t dbytes = al((size_t)nblocks * TB * TB * 8), ybytes = al((size_t)TB * nrhs * 8);
    void *base = nullptr;
    int st = ws_get(ctx, WS_LP, dbytes + ybytes, &base);
    if (st) return st;
    double *Dinv = (double *)base, *Y = (double *)((char *)base + dbytes);
    const bool lower = uplo == 'L';
    if (lower)
        hipLaunchKernelGGL((k_diag_inverses<MODE_LU_FWD, double>), dim3((unsigned)nblocks),
                           dim3(256), 0, ctx->stream, (int)n, Tm, ldt, Dinv);
    else
        hipLaunchKernelGGL((k_diag_inverses<MODE_LU_BWD, double>), dim3((unsigned)nblocks),
                           dim3(256), 0, ctx->stream, (int)n, Tm, ldt, Dinv);
    LA_HIP_CHECK(ctx, hipGetLastError());
    constexpr int64_t SBB = 8;                 // blocks per super-block
    const int64_t nsuper = ceil_div(nblocks, SBB);
    for (int64_t ss = 0; ss < nsuper; ++ss) {
        const int64_t sb = lower ? ss : nsuper - 1 - ss;
        const int64_t b_lo = sb * SBB, b_hi = std::min(nblocks, b_lo + SBB);
        const int64_t R0 = b_lo * TB, R1 = std::min<int64_t>(n, b_hi * TB);
        for (int64_t s = 0; s < b_hi - b_lo; ++s) {
            const int64_t b = lower ? b_lo + s : b_hi - 1 - s;
            const int64_t r0 = b * TB, rb = std::min<int64_t>(TB, n - r0);
            double *Xb = X + r0 * ldx;
            if ((st = gemm_rm<double>(ctx, rb, nrhs, rb, 1.0, Dinv + (size_t)b * TB * TB, TB, Xb,
                                      ldx, 0.0, Y, nrhs, false, false)))
                return st;
            // the rows of this super-block still to be solved
            if (lower && r0 + rb < R1) {
                if ((st = gemm_rm<double>(ctx, R1 - r0 - rb, nrhs, rb, -1.0,
                                          Tm + (r0 + rb) * ldt + r0, ldt, Y, nrhs, 1.0,
                                          X + (r0 + rb) * ldx, ldx, false, false)))
                    return st;
            } else if (!lower && r0 > R0) {
                if ((st = gemm_rm<double>(ctx, r0 - R0, nrhs, rb, -1.0, Tm + R0 * ldt + r0, ldt, Y,
                                          nrhs, 1.0, X + R0 * ldx, ldx, false, false)))
                    return st;
            }
            hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)ceil_div(rb * nrhs, 256)), dim3(256), 0,
                               ctx->stream, rb, nrhs, (const double *)Y, nrhs, Xb, ldx);
            LA_HIP_CHECK(ctx, hipGetLastError());
        }
        // every row outside the super-block, with the super-block's rows of X
        if (lower && R1 < n) {
            if ((st = gemm_rm<double>(ctx, n - R1, nrhs, R1 - R0, -1.0, Tm + R1 * ldt + R0, ldt,
                                      X + R0 * ldx, ldx, 1.0, X + R1 * ldx, ldx, false, false)))
                return st;
        } else if (!lower && R0 > 0) {
            if ((st = gemm_rm<double>(ctx, R0, nrhs, R1 - R0, -1.0, Tm + R0, ldt, X + R0 * ldx, ldx,
                                      1.0, X, ldx, false, false)))
                return st;
        }
    }
    return LA_OK;
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
