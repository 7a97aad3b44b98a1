// Seeded test-matrix generator: A = U diag(sigma) V^T with U and V products
// of Householder reflectors, so kappa_2(A) = kappa exactly and the exact
// solution of A x = b is available through the reflectors.  (No reference
// counterpart: the reference's only random source, Matrix::random at
// src/matrix/mod.rs:842-851, is unseeded.)
//
//   U = H_1 ... H_k,  V = G_1 ... G_k,  H_i = I - 2 u_i u_i^T / (u_i^T u_i)
//   u_i = normal(seed, u_stream + i), v_i = normal(seed, v_stream + i)
//   sigma_j = kappa^(-j/(n-1))
//   A = H_1 (H_2 ( ... (H_k diag(sigma)) ... )) G_k ... G_1
//   x* = G_1 ... G_k diag(1/sigma) H_k ... H_1 b
#include "la_common.hpp"
#include <cmath>

namespace lahip {

__global__ void k_gen_diag(int64_t n, double kappa, double *A, int64_t lda) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    int64_t r = e / n, c = e - r * n;
    double v = 0.0;
    if (r == c) v = (n > 1) ? pow(kappa, -(double)r / (double)(n - 1)) : 1.0;
    A[r * lda + c] = v;
}

__global__ __launch_bounds__(256) void k_gen_sqnorm(int64_t n, const double *V, double *out) {
    __shared__ double s[4];
    const double *v = V + (int64_t)blockIdx.x * n;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += v[i] * v[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// part[y][j] = sum_{i in row chunk y} v[i] A[i][j], grid (col blocks, row chunks);
// w = A^T v is then summed over the chunks in order (no atomics: the
// generated matrix is the same bit for bit in every run)
__global__ __launch_bounds__(256) void k_gen_gemv_t(int64_t n, const double *A, int64_t lda,
                                                    const double *v, double *part) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int64_t i0 = (int64_t)blockIdx.y * 256, i1 = min(n, i0 + 256);
    double acc = 0.0;
    for (int64_t i = i0; i < i1; ++i) acc += v[i] * A[i * lda + j];
    part[(int64_t)blockIdx.y * n + j] = acc;
}

__global__ void k_gen_sum_parts(int64_t n, int nparts, const double *part, double *w) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (int y = 0; y < nparts; ++y) s += part[(int64_t)y * n + j];
    w[j] = s;
}

// w[i] = sum_j A[i][j] v[j]   (w = A v), one workgroup per row
__global__ __launch_bounds__(256) void k_gen_gemv(int64_t n, const double *A, int64_t lda,
                                                  const double *v, double *w) {
    __shared__ double s[4];
    const int64_t i = blockIdx.x;
    double acc = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) acc += A[i * lda + j] * v[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) w[i] = (s[0] + s[1]) + (s[2] + s[3]);
}

// A[i][j] -= c * x[i] * y[j],  c = 2 / sq[0]
__global__ void k_gen_rank1(int64_t n, double *A, int64_t lda, const double *x,
                            const double *y, const double *sq) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    int64_t r = e / n, c = e - r * n;
    const double cf = 2.0 / sq[0];
    A[r * lda + c] -= (cf * x[r]) * y[c];
}

// y <- (I - 2 v v^T / sq) y for nrhs columns (n x nrhs row-major), one WG.
__global__ __launch_bounds__(1024) void k_gen_reflect_vec(int64_t n, int nrhs, const double *v,
                                                          const double *sq, double *y) {
    __shared__ double part[16][16];
    __shared__ double dots[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int r = 0; r < nrhs; ++r) {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += 1024) acc += v[i] * y[i * nrhs + r];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) part[r][wv] = acc;
    }
    __syncthreads();
    if (threadIdx.x < nrhs) {
        double s = 0.0;
        for (int q = 0; q < 16; ++q) s += part[threadIdx.x][q];
        dots[threadIdx.x] = s * (2.0 / sq[0]);
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < n * nrhs; e += 1024) {
        const int64_t i = e / nrhs;
        const int r = (int)(e - i * nrhs);
        y[e] -= v[i] * dots[r];
    }
}

__global__ void k_gen_scale_rows(int64_t n, int nrhs, double kappa, double *y) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    int64_t i = e / nrhs;
    const double s = (n > 1) ? pow(kappa, -(double)i / (double)(n - 1)) : 1.0;
    y[e] = y[e] / s;
}

// Batch of SPD matrices A_i = G_i G_i^T / n_i + I with G_i = normal(seed,
// stream_base + i) (n_i x n_i, row-major): one workgroup per problem draws
// G_i into the scratch slot (same layout as A), then forms the product tile
// by tile through LDS and writes both triangles of A_i.
__global__ __launch_bounds__(256) void k_gen_spd_batch(int count, const int32_t *nvec,
                                                       const int64_t *offA, uint64_t seed,
                                                       uint64_t stream_base, double *A, double *G)
{
    const int prob = blockIdx.x;
    if (prob >= count) return;
    const int n = nvec[prob];
    double *Gp = G + offA[prob];
    const uint64_t key = mix64(seed + (stream_base + prob + 1) * GOLDEN);
    const double two_pi = 6.283185307179586476925286766559;
    const int64_t cnt = (int64_t)n * n;
    for (int64_t p = threadIdx.x; 2 * p < cnt; p += blockDim.x) {
        double u1 = (double)((mix64(key + (uint64_t)(2 * p + 1) * GOLDEN) >> 11) + 1) * 0x1.0p-53;
        double u2 = (double)(mix64(key + (uint64_t)(2 * p + 2) * GOLDEN) >> 11) * 0x1.0p-53;
        double r = sqrt(-2.0 * log(u1));
        Gp[2 * p] = r * cos(two_pi * u2);
        if (2 * p + 1 < cnt) Gp[2 * p + 1] = r * sin(two_pi * u2);
    }
    __syncthreads();
    __shared__ double gi[32][33], gj[32][33];
    const double inv_n = 1.0 / (double)n;
    for (int i0 = 0; i0 < n; i0 += 32)
        for (int j0 = 0; j0 <= i0; j0 += 32) {
            double acc[4] = {0, 0, 0, 0};
            const int tr = threadIdx.x >> 5, tc = threadIdx.x & 31;   // 8 x 32 threads
            for (int k0 = 0; k0 < n; k0 += 32) {
                for (int e = threadIdx.x; e < 32 * 32; e += 256) {
                    const int r = e >> 5, c = e & 31;
                    gi[r][c] = (i0 + r < n && k0 + c < n) ? Gp[(int64_t)(i0 + r) * n + k0 + c] : 0.0;
                    gj[r][c] = (j0 + r < n && k0 + c < n) ? Gp[(int64_t)(j0 + r) * n + k0 + c] : 0.0;
                }
                __syncthreads();
                for (int q = 0; q < 4; ++q)
                    for (int k = 0; k < 32; ++k) acc[q] += gi[tr + 8 * q][k] * gj[tc][k];
                __syncthreads();
            }
            for (int q = 0; q < 4; ++q) {
                const int r = i0 + tr + 8 * q, c = j0 + tc;
                if (r < n && c < n && c <= r) {
                    const double v = acc[q] * inv_n + (r == c ? 1.0 : 0.0);
                    A[offA[prob] + (int64_t)r * n + c] = v;
                    A[offA[prob] + (int64_t)c * n + r] = v;
                }
            }
        }
}

}  // namespace lahip

using namespace lahip;

extern "C" int la_gen_spd_batch_dev(la_ctx *ctx, int64_t count, const int32_t *d_n,
                                    const int64_t *d_offA, uint64_t seed, uint64_t stream_base,
                                    double *A, double *G_scratch)
{
    if (!ctx || !d_n || !d_offA || !A || !G_scratch || count <= 0 || count > INT32_MAX)
        return LA_ERR_ARG;
    hipLaunchKernelGGL(k_gen_spd_batch, dim3((unsigned)count), dim3(256), 0, ctx->stream,
                       (int)count, d_n, d_offA, seed, stream_base, A, G_scratch);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

extern "C" int la_gen_svd_matrix_dev(la_ctx *ctx, int64_t n, double kappa, int nrefl,
                                     uint64_t seed, uint64_t u_stream, uint64_t v_stream,
                                     double *A, int64_t lda, int64_t nrhs, uint64_t b_stream,
                                     double *b, double *xstar)
{
    if (!ctx || !A || n <= 0 || lda < n || nrefl < 0 || !(kappa >= 1.0)) return LA_ERR_ARG;
    if (nrhs < 0 || nrhs > 16 || (nrhs > 0 && (!b || !xstar))) return LA_ERR_ARG;
    const size_t vbytes = (size_t)n * 8;
    void *base = nullptr;
    const int nparts = (int)ceil_div(n, 256);
    int st = ws_get(ctx, WS_GEN,
                    (size_t)(2 * nrefl + 1 + nparts) * vbytes + 2 * nrefl * 8 + 1024, &base);
    if (st) return st;
    double *U = (double *)base;
    double *V = U + (size_t)nrefl * n;
    double *w = V + (size_t)nrefl * n;
    double *part = w + n;  // nparts x n partial sums of A^T v
    double *sq = part + (size_t)nparts * n;  // [0, nrefl): |u_i|^2, [nrefl, 2 nrefl): |v_i|^2
    for (int i = 0; i < nrefl; ++i) {
        if ((st = la_fill_normal_dev(ctx, seed, u_stream + i, n, U + (size_t)i * n))) return st;
        if ((st = la_fill_normal_dev(ctx, seed, v_stream + i, n, V + (size_t)i * n))) return st;
    }
    if (nrefl > 0) {
        hipLaunchKernelGGL(k_gen_sqnorm, dim3(nrefl), dim3(256), 0, ctx->stream, n, U, sq);
        hipLaunchKernelGGL(k_gen_sqnorm, dim3(nrefl), dim3(256), 0, ctx->stream, n, V, sq + nrefl);
    }
    const unsigned g2 = (unsigned)ceil_div(n * n, 256);
    hipLaunchKernelGGL(k_gen_diag, dim3(g2), dim3(256), 0, ctx->stream, n, kappa, A, lda);
    LA_HIP_CHECK(ctx, hipGetLastError());
    const dim3 gt((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, 256));
    for (int i = nrefl - 1; i >= 0; --i) {       // A <- H_i A
        hipLaunchKernelGGL(k_gen_gemv_t, gt, dim3(256), 0, ctx->stream, n, A, lda,
                           U + (size_t)i * n, part);
        hipLaunchKernelGGL(k_gen_sum_parts, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                           ctx->stream, n, nparts, part, w);
        hipLaunchKernelGGL(k_gen_rank1, dim3(g2), dim3(256), 0, ctx->stream, n, A, lda,
                           U + (size_t)i * n, w, sq + i);
    }
    for (int i = nrefl - 1; i >= 0; --i) {       // A <- A G_i
        hipLaunchKernelGGL(k_gen_gemv, dim3((unsigned)n), dim3(256), 0, ctx->stream, n, A, lda,
                           V + (size_t)i * n, w);
        hipLaunchKernelGGL(k_gen_rank1, dim3(g2), dim3(256), 0, ctx->stream, n, A, lda, w,
                           V + (size_t)i * n, sq + nrefl + i);
    }
    LA_HIP_CHECK(ctx, hipGetLastError());
    if (nrhs > 0) {
        if ((st = la_fill_normal_dev(ctx, seed, b_stream, n * nrhs, b))) return st;
        LA_HIP_CHECK(ctx, hipMemcpyAsync(xstar, b, n * nrhs * 8, hipMemcpyDeviceToDevice,
                                         ctx->stream));
        for (int i = 0; i < nrefl; ++i)          // H_k ... H_1 b: H_1 first
            hipLaunchKernelGGL(k_gen_reflect_vec, dim3(1), dim3(1024), 0, ctx->stream, n,
                               (int)nrhs, U + (size_t)i * n, sq + i, xstar);
        hipLaunchKernelGGL(k_gen_scale_rows, dim3((unsigned)ceil_div(n * nrhs, 256)), dim3(256),
                           0, ctx->stream, n, (int)nrhs, kappa, xstar);
        for (int i = nrefl - 1; i >= 0; --i)     // G_1 ... G_k y: G_k first
            hipLaunchKernelGGL(k_gen_reflect_vec, dim3(1), dim3(1024), 0, ctx->stream, n,
                               (int)nrhs, V + (size_t)i * n, sq + nrefl + i, xstar);
        LA_HIP_CHECK(ctx, hipGetLastError());
    }
    return LA_OK;
}
