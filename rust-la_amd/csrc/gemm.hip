// MFMA GEMM for row-major operands on gfx950.
//
// C[M x N] = alpha * A[M x K] * op(B) + beta * C, every matrix row-major with
// its own leading dimension.  op(B) = B (K x N) or, with TRANS_B, the
// transpose of an N x K row-major matrix.  This one kernel serves the
// drop-in `Mul`/`mmul` (src/matrix/mod.rs:957-980, mmatrix.rs:82-98) and the
// trailing updates of the blocked LU and Cholesky factorisations.
//
// Tiling: a 256-thread workgroup owns a 128 x 128 tile of C; its four
// wavefronts each own 64 x 64 (4 x 4 MFMA 16x16 blocks, 16 accumulators).
// K advances in slabs of 16 staged through double-buffered LDS, k-major
// (A is transposed on the way in), so every MFMA operand fetch is a
// 16-lane contiguous LDS read.  FP64 uses v_mfma_f64_16x16x4f64, FP32 uses
// v_mfma_f32_16x16x4f32; both take A and B as
//   A: lane l holds A[l%16][l/16]        B: lane l holds B[l/16][l%16]
// and return D with lane l, element r holding (measured on gfx950)
//   FP64: D[l/16 + 4*r][l%16]            FP32: D[4*(l/16) + r][l%16].
#include "la_common.hpp"

namespace lahip {

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4_t acc_t;
    // row of accumulator element r held by lane-group lhi = lane/16
    __device__ static inline int drow(int lhi, int r) { return lhi + 4 * r; }
    __device__ static inline acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4_t acc_t;
    __device__ static inline int drow(int lhi, int r) { return 4 * lhi + r; }
    __device__ static inline acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

constexpr int GBM = 128, GBN = 128, GBK = 16;

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs, so remap ids so that each XCD receives a contiguous range of
// logical tiles, then walk tiles in groups of 8 tile-rows for L2 reuse.
__device__ inline void tile_coords(int bid, int ntiles, int tiles_m, int tiles_n,
                                   int &tm, int &tn) {
    int logical = bid;
    if ((ntiles & 7) == 0) {
        int per = ntiles >> 3;
        logical = (bid & 7) * per + (bid >> 3);
    }
    const int G = 8;
    int group = logical / (G * tiles_n);
    int first_m = group * G;
    int gsize = min(tiles_m - first_m, G);
    int in_group = logical - group * G * tiles_n;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

// Tiles on or below the diagonal (tn <= tm, tn < tiles_n) of a
// lower-triangular update: sum over tm of min(tm + 1, tiles_n).
__host__ __device__ inline int lower_tile_count(int tiles_m, int tiles_n) {
    return tiles_m <= tiles_n ? tiles_m * (tiles_m + 1) / 2
                              : tiles_n * (tiles_n + 1) / 2 + (tiles_m - tiles_n) * tiles_n;
}

// Logical index -> tile of a lower-triangular update, in the order of
// tile_coords restricted to the lower tiles: groups of 8 tile-rows, inside
// a group by column, then by row.  A group holds every row of the columns
// left of its diagonal and a shrinking run of rows for the columns it
// crosses.  The grid enumerates only these tiles, so with the XCD range
// remap every XCD receives an equal, contiguous share of real work.
__device__ inline void lower_tile_coords(int logical, int tiles_m, int tiles_n, int &tm,
                                         int &tn) {
    const int G = 8;
    for (int first_m = 0; first_m < tiles_m; first_m += G) {
        const int gsize = min(G, tiles_m - first_m);
        const int full_cols = min(first_m, tiles_n);
        const int tri = max(0, min(gsize, tiles_n - first_m));
        const int cnt = full_cols * gsize + tri * gsize - tri * (tri - 1) / 2;
        if (logical < cnt) {
            if (logical < full_cols * gsize) {
                tn = logical / gsize;
                tm = first_m + logical % gsize;
                return;
            }
            int o = logical - full_cols * gsize;
            for (int j = 0; j < tri; ++j) {
                const int c = gsize - j;          // rows first_m + j .. first_m + gsize - 1
                if (o < c) {
                    tn = first_m + j;
                    tm = first_m + j + o;
                    return;
                }
                o -= c;
            }
        }
        logical -= cnt;
    }
    tm = tiles_m;
    tn = 0;
}

template <typename T, bool TRANS_B>
__global__ __launch_bounds__(256, 2)
void k_gemm_rm(int M, int N, int K, T alpha, const T *__restrict__ A, int64_t lda,
               const T *__restrict__ B, int64_t ldb, T beta, T *__restrict__ C,
               int64_t ldc, int lower_only, int tiles_m, int tiles_n)
{
    constexpr int PAD = (sizeof(T) == 8) ? 2 : 4;
    constexpr int LDA_S = GBM + PAD, LDB_S = GBN + PAD;
    __shared__ T As[2][GBK][LDA_S];
    __shared__ T Bs[2][GBK][LDB_S];

    int tm, tn;
    if (lower_only) {
        // the grid is the lower tiles rounded up to a multiple of 8
        const int per = gridDim.x >> 3;
        const int logical = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
        if (logical >= lower_tile_count(tiles_m, tiles_n)) return;
        lower_tile_coords(logical, tiles_m, tiles_n, tm, tn);
    } else {
        tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    }
    const int m0 = tm * GBM, n0 = tn * GBN;
    if (tm >= tiles_m || (lower_only && n0 > m0 + GBM - 1)) return;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l16 = lane & 15, lhi = lane >> 4;

    // global -> register staging assignments
    // A tile (BM x BK): thread -> row ar, k offset ak (8 consecutive k)
    const int ar = tid >> 1, ak = (tid & 1) * 8;
    // B tile: non-trans (BK x BN): row bk_r, 8 consecutive columns bc
    //         trans (BN x BK in memory): row (=n) bt_n, 8 consecutive k bt_k
    const int bk_r = tid >> 4, bc = (tid & 15) * 8;
    const int bt_n = tid >> 1, bt_k = (tid & 1) * 8;

    const bool full_mn = (m0 + GBM <= M) && (n0 + GBN <= N);
    const int nk = (K + GBK - 1) / GBK;

    T ra[8], rb[8];
    auto load_tile = [&](int kt) {
        const int k0 = kt * GBK;
        const bool full = full_mn && (k0 + GBK <= K);
        // edge tiles load from clamped (in-bounds) positions unconditionally,
        // so the eight loads issue together, and zero what lies outside
        {
            const int gr = m0 + ar;
            if (full) {
                const T *src = A + (int64_t)gr * lda + k0 + ak;
#pragma unroll
                for (int e = 0; e < 8; ++e) ra[e] = src[e];
            } else {
                const T *src = A + (int64_t)min(gr, M - 1) * lda;
#pragma unroll
                for (int e = 0; e < 8; ++e) ra[e] = src[min(k0 + ak + e, K - 1)];
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (!(gr < M && k0 + ak + e < K)) ra[e] = T(0);
            }
        }
        if (!TRANS_B) {
            const int gk = k0 + bk_r;
            if (full) {
                const T *src = B + (int64_t)gk * ldb + n0 + bc;
#pragma unroll
                for (int e = 0; e < 8; ++e) rb[e] = src[e];
            } else {
                const T *src = B + (int64_t)min(gk, K - 1) * ldb;
#pragma unroll
                for (int e = 0; e < 8; ++e) rb[e] = src[min(n0 + bc + e, N - 1)];
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (!(gk < K && n0 + bc + e < N)) rb[e] = T(0);
            }
        } else {
            const int gn = n0 + bt_n;
            if (full) {
                const T *src = B + (int64_t)gn * ldb + k0 + bt_k;
#pragma unroll
                for (int e = 0; e < 8; ++e) rb[e] = src[e];
            } else {
                const T *src = B + (int64_t)min(gn, N - 1) * ldb;
#pragma unroll
                for (int e = 0; e < 8; ++e) rb[e] = src[min(k0 + bt_k + e, K - 1)];
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (!(gn < N && k0 + bt_k + e < K)) rb[e] = T(0);
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 8; ++e) As[buf][ak + e][ar] = ra[e];
        if (!TRANS_B) {
#pragma unroll
            for (int e = 0; e < 8; ++e) Bs[buf][bk_r][bc + e] = rb[e];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) Bs[buf][bt_k + e][bt_n] = rb[e];
        }
    };

    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

    if (nk > 0) {
        load_tile(0);
        store_tile(0);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile(kt + 1);
#pragma unroll
        for (int ks = 0; ks < GBK / 4; ++ks) {
            T a[4], b[4];
            const int kr = ks * 4 + lhi;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][kr][wm * 64 + i * 16 + l16];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[buf][kr][wn * 64 + j * 16 + l16];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }

    // epilogue.  Interior tiles (with lower_only: those entirely below the
    // diagonal, last column <= first row) read C (beta != 0) sixteen
    // elements at a time with no per-element condition, so those loads issue
    // together; edge tiles and the tiles the diagonal crosses keep
    // per-element guards.  Both forms evaluate alpha * acc + beta * c in the
    // same order.
    if (full_mn && (!lower_only || n0 + GBN - 1 <= m0)) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            T cv[4][4];
            T *cbase = C + (int64_t)(m0 + wm * 64 + i * 16) * ldc + n0 + wn * 64 + l16;
            if (beta != T(0)) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        cv[r][j] = cbase[(int64_t)Mfma<T>::drow(lhi, r) * ldc + j * 16];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                T *crow = cbase + (int64_t)Mfma<T>::drow(lhi, r) * ldc;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * cv[r][j];
                    crow[j * 16] = v;
                }
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + wm * 64 + i * 16 + Mfma<T>::drow(lhi, r);
            if (row >= M) continue;
            T *crow = C + (int64_t)row * ldc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = n0 + wn * 64 + j * 16 + l16;
                if (col >= N) continue;
                if (lower_only && col > row) continue;
                T v = alpha * acc[i][j][r];
                if (beta != T(0)) v += beta * crow[col];
                crow[col] = v;
            }
        }
    }
}

template <typename T>
int gemm_rm(la_ctx *ctx, int64_t M, int64_t N, int64_t K, T alpha,
            const T *A, int64_t lda, const T *B, int64_t ldb, T beta,
            T *C, int64_t ldc, bool trans_b, bool lower_only)
{
    if (M <= 0 || N <= 0) return LA_OK;
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX)
        return set_error(ctx, LA_ERR_ARG, "gemm: extent exceeds int32");
    const int tiles_m = (int)ceil_div(M, GBM), tiles_n = (int)ceil_div(N, GBN);
    const int64_t ntiles = (int64_t)tiles_m * tiles_n;
    if (ntiles > INT32_MAX / 2) return set_error(ctx, LA_ERR_ARG, "gemm: too many tiles");
    const int64_t nlaunch = lower_only ? (int64_t)(lower_tile_count(tiles_m, tiles_n) + 7) / 8 * 8
                                       : ntiles;
    dim3 grid((unsigned)nlaunch), block(256);
    const double flop = (lower_only ? 1.0 : 2.0) * (double)M * (double)N * (double)K;
    const double bytes = (double)sizeof(T) * ((double)M * K + (double)K * N +
                                              (beta != T(0) ? 2.0 : 1.0) * (double)M * N);
    ProfScope ps(ctx, sizeof(T) == 8 ? LA_K_GEMM64 : LA_K_GEMM32, flop, bytes);
    if (trans_b)
        hipLaunchKernelGGL((k_gemm_rm<T, true>), grid, block, 0, ctx->stream, (int)M, (int)N,
                           (int)K, alpha, A, lda, B, ldb, beta, C, ldc, lower_only ? 1 : 0,
                           tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((k_gemm_rm<T, false>), grid, block, 0, ctx->stream, (int)M, (int)N,
                           (int)K, alpha, A, lda, B, ldb, beta, C, ldc, lower_only ? 1 : 0,
                           tiles_m, tiles_n);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template int gemm_rm<double>(la_ctx *, int64_t, int64_t, int64_t, double, const double *,
                             int64_t, const double *, int64_t, double, double *, int64_t,
                             bool, bool);
template int gemm_rm<float>(la_ctx *, int64_t, int64_t, int64_t, float, const float *,
                            int64_t, const float *, int64_t, float, float *, int64_t, bool,
                            bool);

}  // namespace lahip
