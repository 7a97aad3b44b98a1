// Low-precision-operand MFMA GEMM for the BF16 / FP16 factorisation modes.
//
// C[M x N] (FP32, row-major) = alpha * A[M x K] * B[K x N] + beta * C with A
// and B held in FP32 in memory and rounded (to nearest even) to BF16 or
// FP16 while they are staged into LDS; products accumulate in FP32 on
// v_mfma_f32_16x16x32_{bf16,f16}.  This is the trailing update of the
// low-precision LU: the panel, the multipliers' division and the TRSM stay
// FP32, so the factor's error is that of the rounded update operands.
//
// Tiling: 128 x 128 C tile per 256-thread workgroup, 64 x 64 per wave
// (4 x 4 MFMA blocks), K in slabs of 32 (one MFMA k-step).  Both operands
// are stored k-contiguous in LDS ([row][k] for A, [col][k] for B) so that a
// lane's eight consecutive k values are one 16-byte LDS read.
// Lane layout (gfx950, 16x16x32): A lane l holds A[l%16][8*(l/16)+e],
// B lane l holds B[8*(l/16)+e][l%16], e = 0..7; D lane l element r holds
// D[4*(l/16)+r][l%16].  A and B use the same (l/16, e) -> k assignment, so
// the k order inside the slab does not matter.
#include "la_common.hpp"

namespace lahip {

template <typename E> struct LpOp;
template <> struct LpOp<__bf16> {
    typedef bf8_t vec_t;
    __device__ static inline f4_t op(vec_t a, vec_t b, f4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct LpOp<_Float16> {
    typedef h8_t vec_t;
    __device__ static inline f4_t op(vec_t a, vec_t b, f4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

constexpr int LBM = 128, LBN = 128, LBK = 32, LPAD = 8;

__device__ inline void lp_tile_coords(int bid, int ntiles, int tiles_m, int tiles_n, int &tm,
                                      int &tn) {
    int logical = bid;
    if ((ntiles & 7) == 0) logical = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    const int G = 8;
    int group = logical / (G * tiles_n);
    int first_m = group * G;
    int gsize = min(tiles_m - first_m, G);
    int in_group = logical - group * G * tiles_n;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

template <typename E>
__global__ __launch_bounds__(256, 2)
void k_gemm_lp(int M, int N, int K, float alpha, const float *__restrict__ A, int64_t lda,
               const float *__restrict__ B, int64_t ldb, float beta, float *__restrict__ C,
               int64_t ldc, int tiles_m, int tiles_n)
{
    typedef typename LpOp<E>::vec_t vec_t;
    __shared__ __align__(16) E As[2][LBM][LBK + LPAD];
    __shared__ __align__(16) E Bs[2][LBN][LBK + LPAD];

    int tm, tn;
    lp_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * LBM, n0 = tn * LBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1, l16 = lane & 15, lhi = lane >> 4;

    // staging: A -> thread (row ar, k half ak: 16 consecutive k)
    const int ar = tid >> 1, ak = (tid & 1) * 16;
    // B -> thread (k pair kp: rows 2kp, 2kp+1; 8 consecutive columns bc)
    const int kp = tid >> 4, bc = (tid & 15) * 8;
    const bool full_mn = (m0 + LBM <= M) && (n0 + LBN <= N);
    const int nk = (K + LBK - 1) / LBK;

    float ra[16], rb0[8], rb1[8];
    auto load_tile = [&](int kt) {
        const int k0 = kt * LBK;
        const bool full = full_mn && (k0 + LBK <= K);
        const int gr = m0 + ar;
        const float *sa = A + (int64_t)gr * lda + k0 + ak;
        const int gk = k0 + 2 * kp;
        const float *sb0 = B + (int64_t)gk * ldb + n0 + bc;
        const float *sb1 = sb0 + ldb;
        if (full) {
#pragma unroll
            for (int e = 0; e < 16; ++e) ra[e] = sa[e];
#pragma unroll
            for (int e = 0; e < 8; ++e) { rb0[e] = sb0[e]; rb1[e] = sb1[e]; }
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) ra[e] = (gr < M && k0 + ak + e < K) ? sa[e] : 0.f;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool cn = n0 + bc + e < N;
                rb0[e] = (gk < K && cn) ? sb0[e] : 0.f;
                rb1[e] = (gk + 1 < K && cn) ? sb1[e] : 0.f;
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 16; ++e) As[buf][ar][ak + e] = (E)ra[e];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            Bs[buf][bc + e][2 * kp] = (E)rb0[e];
            Bs[buf][bc + e][2 * kp + 1] = (E)rb1[e];
        }
    };

    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.f;

    if (nk > 0) {
        load_tile(0);
        store_tile(0);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile(kt + 1);
        vec_t a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *reinterpret_cast<const vec_t *>(&As[buf][wm * 64 + i * 16 + l16][8 * lhi]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            b[j] = *reinterpret_cast<const vec_t *>(&Bs[buf][wn * 64 + j * 16 + l16][8 * lhi]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = LpOp<E>::op(a[i], b[j], acc[i][j]);
        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + wm * 64 + i * 16 + 4 * lhi + r;
            if (row >= M) continue;
            float *crow = C + (int64_t)row * ldc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = n0 + wn * 64 + j * 16 + l16;
                if (col >= N) continue;
                float v = alpha * acc[i][j][r];
                if (beta != 0.f) v += beta * crow[col];
                crow[col] = v;
            }
        }
}

// prec: LA_PREC_BF16 or LA_PREC_FP16
int gemm_lp(la_ctx *ctx, int prec, int64_t M, int64_t N, int64_t K, float alpha, const float *A,
            int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc)
{
    if (M <= 0 || N <= 0) return LA_OK;
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX)
        return set_error(ctx, LA_ERR_ARG, "gemm_lp: extent exceeds int32");
    const int tiles_m = (int)ceil_div(M, LBM), tiles_n = (int)ceil_div(N, LBN);
    const int64_t ntiles = (int64_t)tiles_m * tiles_n;
    ProfScope ps(ctx, LA_K_GEMM16, 2.0 * M * N * K,
                 4.0 * ((double)M * K + (double)K * N + (beta != 0.f ? 2.0 : 1.0) * M * N));
    if (prec == LA_PREC_BF16)
        hipLaunchKernelGGL(k_gemm_lp<__bf16>, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream,
                           (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m,
                           tiles_n);
    else if (prec == LA_PREC_FP16)
        hipLaunchKernelGGL(k_gemm_lp<_Float16>, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream,
                           (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m,
                           tiles_n);
    else
        return set_error(ctx, LA_ERR_ARG, "gemm_lp: precision must be BF16 or FP16");
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

}  // namespace lahip
