// The f32 surface of Matrix<f32> (src/matrix/simd.rs:6-219) on gfx950:
// simd_mul on the FP32 MFMA GEMM of the refinement solver's factorisation,
// the elementwise operations (scale, add, sub, elem_mul, elem_div, neg) as
// one memory-bound kernel, and simd_dot as a reduction.
//
// Two defects of the reference are not reproduced:
//  * its elementwise loops leave the result uninitialised for matrices of
//    fewer than four elements (the scalar tail runs only when
//    len & !3 != 0, simd.rs:19-24 and alike); here every element is formed;
//  * simd_dot's tail adds a[i] instead of a[i] * b[i] over
//    [len & 3, len) (simd.rs:179-184); here the dot product is the sum of
//    every a[i] * b[i], accumulated in FP64 and rounded once to FP32.
#include "la_common.hpp"

namespace lahip {

// out[i] = op(a[i], b[i]) with op from la_elementwise_op
__global__ void k_elementwise32(int op, int64_t count, const float *__restrict__ a,
                                const float *__restrict__ b, float s, float *__restrict__ out)
{
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= count) return;
    float x[4], y[4] = {0.f, 0.f, 0.f, 0.f}, r[4];
    const int nv = (int)min((int64_t)4, count - i0);
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = a[min(i0 + e, count - 1)];
    if (b) {
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] = b[min(i0 + e, count - 1)];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        switch (op) {
        case LA_EW_SCALE: r[e] = s * x[e]; break;
        case LA_EW_ADD: r[e] = x[e] + y[e]; break;
        case LA_EW_SUB: r[e] = x[e] - y[e]; break;
        case LA_EW_MUL: r[e] = x[e] * y[e]; break;
        case LA_EW_DIV: r[e] = x[e] / y[e]; break;
        default: r[e] = -x[e]; break;
        }
    }
    for (int e = 0; e < nv; ++e) out[i0 + e] = r[e];
}

// partial[blk] = sum over the block's elements of (double)a[i] * b[i]
__global__ __launch_bounds__(256)
void k_dot32_partial(int64_t count, const float *__restrict__ a, const float *__restrict__ b,
                     double *partial)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * 256)
        acc = fma((double)a[i], (double)b[i], acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_dot32_final(int nblk, const double *partial, float *out)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[b];
    *out = (float)s;
}

constexpr int DOT_BLOCKS = 256;

}  // namespace lahip

using namespace lahip;

extern "C" int la_sgemm_dev(la_ctx *ctx, int64_t m, int64_t n, int64_t k, float alpha,
                            const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
                            float *C, int64_t ldc)
{
    if (!ctx || !A || !B || !C) return LA_ERR_ARG;
    if (m < 0 || n < 0 || k < 0 || lda < std::max<int64_t>(k, 1) ||
        ldb < std::max<int64_t>(n, 1) || ldc < std::max<int64_t>(n, 1))
        return LA_ERR_ARG;
    return gemm_rm<float>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, false, false);
}

extern "C" int la_sgemm_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                           const float *A, const float *B, float *C)
{
    if (!ctx) return LA_ERR_ARG;
    if (k != k2) return set_error(ctx, LA_ERR_SHAPE, "simd_mul: self.cols() != m2.rows()");
    if (m <= 0 || k <= 0 || p <= 0) return set_error(ctx, LA_ERR_SHAPE, "simd_mul: empty matrix");
    if (!A || !B || !C) return LA_ERR_ARG;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t ba = al((size_t)m * k * 4), bb = al((size_t)k * p * 4), bc = al((size_t)m * p * 4);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST, ba + bb + bc, &base);
    if (st) return st;
    float *dA = (float *)base, *dB = (float *)((char *)base + ba),
          *dC = (float *)((char *)base + ba + bb);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, (size_t)m * k * 4, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dB, B, (size_t)k * p * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((st = gemm_rm<float>(ctx, m, p, k, 1.f, dA, k, dB, p, 0.f, dC, p, false, false))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(C, dC, (size_t)m * p * 4, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" int la_selementwise_dev(la_ctx *ctx, int op, int64_t count, const float *a,
                                   const float *b, float scalar, float *out)
{
    if (!ctx || !a || !out || count < 0 || op < LA_EW_SCALE || op > LA_EW_NEG) return LA_ERR_ARG;
    const bool binary = op == LA_EW_ADD || op == LA_EW_SUB || op == LA_EW_MUL || op == LA_EW_DIV;
    if (binary && !b) return LA_ERR_ARG;
    if (count == 0) return LA_OK;
    hipLaunchKernelGGL(k_elementwise32, dim3((unsigned)ceil_div(count, 1024)), dim3(256), 0,
                       ctx->stream, op, count, a, binary ? b : nullptr, scalar, out);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

extern "C" int la_selementwise_rm(la_ctx *ctx, int op, int64_t rows, int64_t cols,
                                  int64_t b_rows, int64_t b_cols, const float *a, const float *b,
                                  float scalar, float *out)
{
    if (!ctx) return LA_ERR_ARG;
    const bool binary = op == LA_EW_ADD || op == LA_EW_SUB || op == LA_EW_MUL || op == LA_EW_DIV;
    if (op < LA_EW_SCALE || op > LA_EW_NEG || !a || !out || (binary && !b)) return LA_ERR_ARG;
    if (rows <= 0 || cols <= 0) return set_error(ctx, LA_ERR_SHAPE, "simd: empty matrix");
    // simd.rs:32-33 and alike: the operands' shapes must match
    if (binary && (b_rows != rows || b_cols != cols))
        return set_error(ctx, LA_ERR_SHAPE, "simd: operand shapes differ");
    const int64_t count = rows * cols;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al((size_t)count * 4);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST, 3 * bytes, &base);
    if (st) return st;
    float *da = (float *)base, *db = (float *)((char *)base + bytes),
          *dout = (float *)((char *)base + 2 * bytes);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(da, a, (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
    if (binary)
        LA_HIP_CHECK(ctx, hipMemcpyAsync(db, b, (size_t)count * 4, hipMemcpyHostToDevice,
                                         ctx->stream));
    if ((st = la_selementwise_dev(ctx, op, count, da, binary ? db : nullptr, scalar, dout)))
        return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(out, dout, (size_t)count * 4, hipMemcpyDeviceToHost,
                                     ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" int la_sdot_dev(la_ctx *ctx, int64_t count, const float *a, const float *b,
                           float *d_out)
{
    if (!ctx || !a || !b || !d_out || count < 0) return LA_ERR_ARG;
    void *base = nullptr;
    int st = ws_get(ctx, WS_IR, DOT_BLOCKS * sizeof(double), &base);
    if (st) return st;
    hipLaunchKernelGGL(k_dot32_partial, dim3(DOT_BLOCKS), dim3(256), 0, ctx->stream, count, a, b,
                       (double *)base);
    hipLaunchKernelGGL(k_dot32_final, dim3(1), dim3(64), 0, ctx->stream, DOT_BLOCKS,
                       (const double *)base, d_out);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

extern "C" int la_sdot_rm(la_ctx *ctx, int64_t rows, int64_t cols, int64_t b_rows,
                          int64_t b_cols, const float *a, const float *b, float *out)
{
    if (!ctx || !a || !b || !out) return LA_ERR_ARG;
    // simd.rs:167-168: equal row counts, both single columns
    if (rows != b_rows || cols != 1 || b_cols != 1)
        return set_error(ctx, LA_ERR_SHAPE, "simd_dot: operands must be equal-length columns");
    if (rows <= 0) return set_error(ctx, LA_ERR_SHAPE, "simd_dot: empty matrix");
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al((size_t)rows * 4);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST, 2 * bytes + 256, &base);
    if (st) return st;
    float *da = (float *)base, *db = (float *)((char *)base + bytes),
          *dr = (float *)((char *)base + 2 * bytes);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(da, a, (size_t)rows * 4, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(db, b, (size_t)rows * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((st = la_sdot_dev(ctx, rows, da, db, dr))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(out, dr, 4, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}
