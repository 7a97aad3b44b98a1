// Small utility kernels shared by the factorisations, the seeded input
// generators and the backward-error measurement.
#include "la_common.hpp"
#include <math.h>

namespace lahip {

__global__ void k_util_fill_i32(int32_t *p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_util_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_util_iota_u64(uint64_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint64_t)i;
}
__global__ void k_util_copy2d(int64_t rows, int64_t cols, const double *src, int64_t lds,
                              double *dst, int64_t ldd) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    int64_t r = e / cols, c = e - r * cols;
    dst[r * ldd + c] = src[r * lds + c];
}
__global__ void k_util_diag_zero(int64_t n, const double *LU, int64_t ld, int32_t *status) {
    __shared__ int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    int z = 0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) z |= (LU[j * ld + j] == 0.0);
    if (z) atomicOr(&s, 1);
    __syncthreads();
    if (threadIdx.x == 0) *status = s ? LA_SINGULAR : LA_OK;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)ceil_div(n, b); }

size_t allow_dynamic_lds(const void *kernel) {
    hipFuncAttributes attr;
    size_t stat = 0;
    if (hipFuncGetAttributes(&attr, kernel) == hipSuccess) stat = attr.sharedSizeBytes;
    const size_t total = 160 * 1024;
    size_t dyn = stat < total ? total - stat : 0;
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) !=
        hipSuccess)
        dyn = 64 * 1024 - stat;
    (void)hipGetLastError();   // never leave a sticky error behind
    return dyn;
}

uint64_t spin_limit_ticks(la_ctx *ctx) {
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) !=
            hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    // 5 s: far above any legitimate wait, short enough to fail fast.
    return (uint64_t)rate_khz * 1000ull * 5ull;
}

int fill_i32(la_ctx *ctx, int32_t *p, int64_t n, int32_t v) {
    if (n <= 0) return LA_OK;
    hipLaunchKernelGGL(k_util_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
int fill_u64(la_ctx *ctx, uint64_t *p, int64_t n, uint64_t v) {
    if (n <= 0) return LA_OK;
    hipLaunchKernelGGL(k_util_fill_u64, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
int iota_u64(la_ctx *ctx, uint64_t *p, int64_t n) {
    if (n <= 0) return LA_OK;
    hipLaunchKernelGGL(k_util_iota_u64, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, p, n);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
int copy2d(la_ctx *ctx, int64_t rows, int64_t cols, const double *src, int64_t lds,
           double *dst, int64_t ldd) {
    if (rows <= 0 || cols <= 0) return LA_OK;
    hipLaunchKernelGGL(k_util_copy2d, dim3(nblk(rows * cols, 256)), dim3(256), 0, ctx->stream,
                       rows, cols, src, lds, dst, ldd);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
int check_diag_zero(la_ctx *ctx, int64_t n, const double *LU, int64_t ld, int32_t *status) {
    hipLaunchKernelGGL(k_util_diag_zero, dim3(1), dim3(256), 0, ctx->stream, n, LU, ld, status);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// ------------------------------------------------------------ generators
// SplitMix64 finaliser over a counter; identical to the CPU generator the
// tests use, so uniform draws are bit-identical on both sides.

__global__ void k_fill_uniform(uint64_t key, int64_t count, double *out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double u = (double)(mix64(key + (uint64_t)(i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
    out[i] = 2.0 * u - 1.0;
}
__global__ void k_fill_normal(uint64_t key, int64_t count, double *out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * p >= count) return;
    const double two_pi = 6.283185307179586476925286766559;
    double u1 = (double)((mix64(key + (uint64_t)(2 * p + 1) * GOLDEN) >> 11) + 1) * 0x1.0p-53;
    double u2 = (double)(mix64(key + (uint64_t)(2 * p + 2) * GOLDEN) >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    double t = two_pi * u2;
    out[2 * p] = r * cos(t);
    if (2 * p + 1 < count) out[2 * p + 1] = r * sin(t);
}

}  // namespace lahip

using namespace lahip;

static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed + (stream + 1) * GOLDEN);
}

extern "C" int la_fill_uniform_dev(la_ctx *ctx, uint64_t seed, uint64_t stream, int64_t count,
                                   double *out) {
    if (!ctx || (!out && count > 0) || count < 0) return LA_ERR_ARG;
    if (count == 0) return LA_OK;
    hipLaunchKernelGGL(k_fill_uniform, dim3(nblk(count, 256)), dim3(256), 0, ctx->stream,
                       stream_key(seed, stream), count, out);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
extern "C" int la_fill_normal_dev(la_ctx *ctx, uint64_t seed, uint64_t stream, int64_t count,
                                  double *out) {
    if (!ctx || (!out && count > 0) || count < 0) return LA_ERR_ARG;
    if (count == 0) return LA_OK;
    hipLaunchKernelGGL(k_fill_normal, dim3(nblk(ceil_div(count, 2), 256)), dim3(256), 0,
                       ctx->stream, stream_key(seed, stream), count, out);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
