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
//
// With LA_IR_RBT the FP32 factor is of a randomised butterfly transform of
// A instead, factored without pivoting (Parker's random butterflies, depth 2,
// as in Baboulin, Dongarra, Herrmann and Tomov, "Accelerating linear system
// solutions using randomization techniques", ACM TOMS 39(2), 2013):
//   Ap = [A 0; 0 I] of order np = 4 ceil(n/4);   Ar = U^T Ap V
//   Ar = L U (no pivoting, static pivots);  correction d = V U^-1 L^-1 U^T r
// The refinement itself (residual, update, check) is unchanged and runs on
// the original A; a refinement that does not converge falls back to the
// partial-pivoting factor.
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

// R = B - A X for nrhs <= 16 right-hand sides.  Each wavefront takes
// RES_RPW rows in turn, a workgroup 4 * RES_RPW rows; every lane keeps eight
// independent loads of A (and of X) in flight.  NRB = 1 is the single
// right-hand side without per-column guards.  norms[0] = max |R|,
// norms[1] = max |X|, norms[2] = max |B|, norms[3] = max_i sum_j |A_ij|
// (only when want_anorm): reduced over the workgroup first, so each
// workgroup issues one atomic per norm.  Maxima are taken over the bit
// patterns of the non-negative values, so a NaN (above +inf) propagates.
constexpr int RES_RPW = 4;
template <int NRB>
__global__ __launch_bounds__(256)
void k_residual(int64_t n, int nrhs_arg, const double *__restrict__ A, int64_t lda,
                const double *__restrict__ X, int64_t ldx, const double *__restrict__ B,
                int64_t ldb, double *R, int64_t ldr, double *norms, int want_anorm)
{
    __shared__ unsigned long long red[4][4];
    const int nrhs = (NRB == 1) ? 1 : nrhs_arg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    typedef unsigned long long u64;
    auto bits = [](double v) { return (u64)__double_as_longlong(fabs(v)); };
    u64 mr = 0, mx = 0, mb = 0, ma = 0;
    for (int rr = 0; rr < RES_RPW; ++rr) {
        const int64_t i = ((int64_t)blockIdx.x * 4 + wv) * RES_RPW + rr;
        if (i >= n) break;
        const double *arow = A + i * lda;
        double acc[NRB];
#pragma unroll
        for (int r = 0; r < NRB; ++r) acc[r] = 0.0;
        double asum = 0.0;
        int64_t j = lane;
        for (; j + 64 * 7 < n; j += 64 * 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = arow[j + 64 * u];
            if (NRB == 1) {
                double xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) xv[u] = X[(j + 64 * u) * ldx];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    asum += fabs(a[u]);
                    acc[0] = fma(a[u], xv[u], acc[0]);
                }
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    asum += fabs(a[u]);
                    const double *xr = X + (j + 64 * u) * ldx;
#pragma unroll
                    for (int r = 0; r < NRB; ++r)
                        if (r < nrhs) acc[r] = fma(a[u], xr[r], acc[r]);
                }
            }
        }
        for (; j < n; j += 64) {
            const double a = arow[j];
            asum += fabs(a);
            const double *xr = X + j * ldx;
#pragma unroll
            for (int r = 0; r < NRB; ++r)
                if (r < nrhs) acc[r] = fma(a, xr[r], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < NRB; ++r) {
            double v = acc[r];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            acc[r] = v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) asum += __shfl_xor(asum, off);
        ma = max(ma, bits(asum));
        if (lane < nrhs) {
            double ax = acc[0];
#pragma unroll
            for (int r = 1; r < NRB; ++r) ax = (r == lane) ? acc[r] : ax;
            const double b = B[i * ldb + lane];
            const double res = b - ax;
            R[i * ldr + lane] = res;
            mr = max(mr, bits(res));
            mx = max(mx, bits(X[i * ldx + lane]));
            mb = max(mb, bits(b));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mr = max(mr, (u64)__shfl_xor((long long)mr, off));
        mx = max(mx, (u64)__shfl_xor((long long)mx, off));
        mb = max(mb, (u64)__shfl_xor((long long)mb, off));
    }
    if (lane == 0) {
        red[wv][0] = mr;
        red[wv][1] = mx;
        red[wv][2] = mb;
        red[wv][3] = ma;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        const u64 v = max(max(red[0][q], red[1][q]), max(red[2][q], red[3][q]));
        if (q < 3 || want_anorm) atomicMax(reinterpret_cast<u64 *>(norms + q), v);
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

// ---------------------------------------------- randomised butterflies
// A butterfly of order m is (1/sqrt 2) [R S; R -S], R and S diagonal with
// entries exp(u/10), u uniform in [-1/2, 1/2).  The depth-2 transform of
// order np is W = diag(B1, B2) B0 (B0 of order np, B1 and B2 of order np/2);
// its parameters for one side are g0 (B0: R in [0, h), S in [h, np)) and g1
// (B1: R in [0, q), S in [q, h); B2: R in [h, h+q), S in [h+q, np)), with
// h = np/2 and q = np/4.  The four indices i, i+q, h+i, h+q+i (i < q) are
// closed under both levels.  g holds U's g0, g1 then V's g0, g1.
__global__ void k_rbt_params(int64_t np, uint64_t seed, double *g)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * np) return;
    const uint64_t z = mix64(seed + GOLDEN * (uint64_t)(e + 1));
    const double u = (double)(z >> 11) * 0x1.0p-53;
    g[e] = exp((u - 0.5) / 10.0);
}

// v <- W^T v on the index group (i, i+q, h+i, h+q+i)
__device__ inline void rbt_t4(double v[4], const double *g0, const double *g1, int64_t i,
                              int64_t q, int64_t h)
{
    const double z0 = g1[i] * (v[0] + v[1]);
    const double z1 = g1[q + i] * (v[0] - v[1]);
    const double z2 = g1[h + i] * (v[2] + v[3]);
    const double z3 = g1[h + q + i] * (v[2] - v[3]);
    v[0] = 0.5 * g0[i] * (z0 + z2);
    v[1] = 0.5 * g0[q + i] * (z1 + z3);
    v[2] = 0.5 * g0[h + i] * (z0 - z2);
    v[3] = 0.5 * g0[h + q + i] * (z1 - z3);
}

// v <- W v on the index group (i, i+q, h+i, h+q+i)
__device__ inline void rbt_f4(double v[4], const double *g0, const double *g1, int64_t i,
                              int64_t q, int64_t h)
{
    const double u0 = g0[i] * v[0] + g0[h + i] * v[2];
    const double u2 = g0[i] * v[0] - g0[h + i] * v[2];
    const double u1 = g0[q + i] * v[1] + g0[h + q + i] * v[3];
    const double u3 = g0[q + i] * v[1] - g0[h + q + i] * v[3];
    v[0] = 0.5 * (g1[i] * u0 + g1[q + i] * u1);
    v[1] = 0.5 * (g1[i] * u0 - g1[q + i] * u1);
    v[2] = 0.5 * (g1[h + i] * u2 + g1[h + q + i] * u3);
    v[3] = 0.5 * (g1[h + i] * u2 - g1[h + q + i] * u3);
}

// W32 = fl32(U^T Ap V), Ap = [A 0; 0 I]; *amax_bits = max |W32| (as bits).
// Thread (row group i = blockIdx.y, column group j): a 4 x 4 tile of
// indices (i, i+q, h+i, h+q+i) x (j, j+q, h+j, h+q+j); consecutive threads
// take consecutive columns, so every load and store is coalesced.
__global__ __launch_bounds__(256)
void k_rbt_matrix(int64_t n, int64_t np, const double *__restrict__ A, int64_t lda,
                  const double *__restrict__ g, float *W, int64_t ldw, unsigned *amax_bits)
{
    __shared__ float wmax[4];
    const int64_t q = np / 4, h = np / 2;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = blockIdx.y;
    float amax = 0.f;
    if (j < q) {
        const int64_t ri[4] = {i, i + q, h + i, h + q + i};
        const int64_t cj[4] = {j, j + q, h + j, h + q + j};
        double v[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                v[a][b] = A[min(ri[a], n - 1) * lda + min(cj[b], n - 1)];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (!(ri[a] < n && cj[b] < n)) v[a][b] = (ri[a] == cj[b]) ? 1.0 : 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) rbt_t4(v[a], g + 2 * np, g + 3 * np, j, q, h);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double t[4] = {v[0][b], v[1][b], v[2][b], v[3][b]};
            rbt_t4(t, g, g + np, i, q, h);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const float f = (float)t[a];
                W[ri[a] * ldw + cj[b]] = f;
                amax = fmaxf(amax, fabsf(f));
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) amax = fmaxf(amax, __shfl_xor(amax, m));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = amax;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(amax_bits, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]),
                                                   fmaxf(wmax[2], wmax[3]))));
}

// static-pivot threshold: 2^-20 max |W32| (at least the smallest normal)
__global__ void k_rbt_tau(const unsigned *amax_bits, float *tau)
{
    if (threadIdx.x == 0) *tau = fmaxf(__uint_as_float(*amax_bits) * 0x1.0p-20f, 1.17549435e-38f);
}

// Y (np x nrhs, ld nrhs) = U^T [R; 0]
__global__ void k_rbt_left_vec(int64_t n, int64_t np, int nrhs, const double *__restrict__ g,
                               const double *__restrict__ R, int64_t ldr, double *Y)
{
    const int64_t q = np / 4, h = np / 2;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= q * nrhs) return;
    const int64_t i = e / nrhs, r = e - i * nrhs;
    const int64_t ri[4] = {i, i + q, h + i, h + q + i};
    double v[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) v[a] = R[min(ri[a], n - 1) * ldr + r];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        if (ri[a] >= n) v[a] = 0.0;
    rbt_t4(v, g, g + np, i, q, h);
#pragma unroll
    for (int a = 0; a < 4; ++a) Y[ri[a] * nrhs + r] = v[a];
}

// X[0:n] = (V Y)[0:n]
__global__ void k_rbt_right_vec(int64_t n, int64_t np, int nrhs, const double *__restrict__ g,
                                const double *__restrict__ Y, double *X, int64_t ldx)
{
    const int64_t q = np / 4, h = np / 2;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= q * nrhs) return;
    const int64_t i = e / nrhs, r = e - i * nrhs;
    const int64_t ri[4] = {i, i + q, h + i, h + q + i};
    double v[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) v[a] = Y[ri[a] * nrhs + r];
    rbt_f4(v, g + 2 * np, g + 3 * np, i, q, h);
#pragma unroll
    for (int a = 0; a < 4; ++a)
        if (ri[a] < n) X[ri[a] * ldx + r] = v[a];
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
    const unsigned grid = (unsigned)ceil_div(n, (int64_t)4 * RES_RPW);
    if (nrhs == 1)
        hipLaunchKernelGGL(k_residual<1>, dim3(grid), dim3(256), 0, ctx->stream, n, (int)nrhs, A,
                           lda, X, ldx, B, ldb, R, ldr, d_norms, want_anorm ? 1 : 0);
    else
        hipLaunchKernelGGL(k_residual<16>, dim3(grid), dim3(256), 0, ctx->stream, n, (int)nrhs,
                           A, lda, X, ldx, B, ldb, R, ldr, d_norms, want_anorm ? 1 : 0);
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
    // randomised-butterfly form
    bool rbt = false;
    int64_t np = 0;                      // padded order (multiple of 4)
    double *g = nullptr;                 // butterfly parameters (4 np)
    double *Y = nullptr;                 // transformed right-hand sides (np x nrhs)
    unsigned *amax = nullptr;
    float *tau = nullptr;
    int32_t *npert = nullptr;
};

static int ir_alloc(la_ctx *ctx, int64_t n, int64_t nrhs, IrWork &w)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t np = (n + 3) & ~(int64_t)3;       // butterfly order
    const int64_t ld32 = (np + 63) & ~(int64_t)63;  // pad rows to 256 B
    size_t b32 = al((size_t)np * ld32 * sizeof(float));
    size_t bperm = al((size_t)n * 8), bipiv = al((size_t)n * 8);
    size_t brd = al((size_t)n * nrhs * 8);
    size_t bn = al(8 * 8), bst = al(4 * 4), bsc = al((size_t)n * 8);
    size_t bg = al((size_t)4 * np * 8), by = al((size_t)np * nrhs * 8), bsm = al(64);
    void *base = nullptr;
    int st = ws_get(ctx, WS_LP,
                    b32 + bperm + bipiv + 2 * brd + bn + bst + 2 * bsc + bg + by + bsm, &base);
    if (st) return st;
    char *p = (char *)base;
    w.np = np;
    w.g = (double *)p; p += bg;
    w.Y = (double *)p; p += by;
    w.amax = (unsigned *)p;
    w.tau = (float *)(p + 16);
    w.npert = (int32_t *)(p + 32);
    p += bsm;
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
    if (w.rbt) {                     // d = V U^-1 L^-1 U^T r
        const int64_t q = w.np / 4;
        hipLaunchKernelGGL(k_rbt_left_vec, dim3((unsigned)ceil_div(q * nrhs, 256)), dim3(256), 0,
                           ctx->stream, n, w.np, (int)nrhs, w.g, rhs, ldr, w.Y);
        LA_HIP_CHECK(ctx, hipGetLastError());
        if ((st = trsv_blocked<float>(ctx, 'L', false, true, w.np, w.A32, w.ld32, nrhs, w.Y,
                                      nrhs)))
            return st;
        if ((st = trsv_blocked<float>(ctx, 'U', false, false, w.np, w.A32, w.ld32, nrhs, w.Y,
                                      nrhs)))
            return st;
        hipLaunchKernelGGL(k_rbt_right_vec, dim3((unsigned)ceil_div(q * nrhs, 256)), dim3(256), 0,
                           ctx->stream, n, w.np, (int)nrhs, w.g, w.Y, out, ldo);
        LA_HIP_CHECK(ctx, hipGetLastError());
        return LA_OK;
    }
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

// One factor-and-refine attempt (partial pivoting, or the butterfly form
// when w.rbt); fills report and returns its status.
constexpr int RBT_MAX_STEPS = 15;
static int ir_attempt(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A, int64_t lda,
                      const double *B, int64_t ldb, double *X, int64_t ldx, int prec,
                      double tol, int max_iter, IrWork &w, la_ir_report *report)
{
    int st;
    EvTimer ev;
    LA_HIP_CHECK(ctx, hipEventRecord(ev.e[0], ctx->stream));
    if (w.rbt) {
        const int64_t q = w.np / 4;
        hipLaunchKernelGGL(k_rbt_params, dim3((unsigned)ceil_div(4 * w.np, 256)), dim3(256), 0,
                           ctx->stream, w.np, (uint64_t)0x5EED0F0B0773ull, w.g);
        if ((st = fill_i32(ctx, (int32_t *)w.amax, 1, 0))) return st;
        hipLaunchKernelGGL(k_rbt_matrix, dim3((unsigned)ceil_div(q, 256), (unsigned)q), dim3(256),
                           0, ctx->stream, n, w.np, A, lda, w.g, w.A32, w.ld32, w.amax);
        hipLaunchKernelGGL(k_rbt_tau, dim3(1), dim3(64), 0, ctx->stream, w.amax, w.tau);
        LA_HIP_CHECK(ctx, hipGetLastError());
        if ((st = fill_i32(ctx, w.status, 1, 0))) return st;
        if ((st = getrf_nopiv_core<float>(ctx, w.np, w.np, w.A32, w.ld32, w.tau, w.npert, prec)))
            return st;
    } else {
        if (prec == LA_PREC_FP32) {
            hipLaunchKernelGGL(k_convert_f32, dim3((unsigned)ceil_div(n * n, 1024)), dim3(256), 0,
                               ctx->stream, n, n, A, lda, w.A32, w.ld32);
        } else {
            // per-column power-of-two scaling keeps the rounded operands in range
            if ((st = fill_u64(ctx, (uint64_t *)w.colmax, n, 0))) return st;
            hipLaunchKernelGGL(k_colmax,
                               dim3((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, 256)),
                               dim3(256), 0, ctx->stream, n, A, lda, w.colmax);
            hipLaunchKernelGGL(k_colscale, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                               ctx->stream, n, w.colmax, w.colscale);
            hipLaunchKernelGGL(k_convert_f32_scaled, dim3((unsigned)ceil_div(n * n, 256)),
                               dim3(256), 0, ctx->stream, n, A, lda, w.colscale, w.A32, w.ld32);
        }
        LA_HIP_CHECK(ctx, hipGetLastError());
        if ((st = getrf_core<float>(ctx, n, n, w.A32, w.ld32, w.ipiv, w.perm, w.status, prec)))
            return st;
    }
    LA_HIP_CHECK(ctx, hipEventRecord(ev.e[1], ctx->stream));
    if ((st = lp_solve(ctx, n, nrhs, w, B, ldb, X, ldx))) return st;
    double h[4] = {0, 0, 0, 0};
    int it = 0;
    double eta = INFINITY, eta_prev = INFINITY;
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
        // the unpivoted factor is abandoned for the pivoted one as soon as
        // refinement stagnates (less than a halving per step) or runs long
        if (w.rbt && ((it >= 2 && eta > 0.5 * eta_prev) || it >= RBT_MAX_STEPS)) break;
        eta_prev = eta;
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
    report->factor_ms += ev.ms(0, 1);
    report->refine_ms += ev.ms(1, 2);
    report->total_ms += ev.ms(0, 2);
    if (fst != LA_OK) report->status = fst;
    else if (!(eta <= tol)) report->status = std::isfinite(eta) ? LA_NOT_CONVERGED : LA_SINGULAR;
    else report->status = LA_OK;
    return report->status;
}

extern "C" int la_dgesv_ir_dev(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A,
                               int64_t lda, const double *B, int64_t ldb, double *X,
                               int64_t ldx, int prec, double tol, int max_iter,
                               la_ir_report *report)
{
    if (!ctx || !A || !B || !X || !report) return LA_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || nrhs > 16 || lda < n || ldb < nrhs || ldx < nrhs)
        return LA_ERR_ARG;
    const bool rbt = (prec & LA_IR_RBT) != 0;
    prec &= ~LA_IR_RBT;
    if (prec != LA_PREC_FP32 && prec != LA_PREC_BF16 && prec != LA_PREC_FP16)
        return set_error(ctx, LA_ERR_ARG, "dgesv_ir: precision not supported");
    if (rbt && (prec != LA_PREC_FP32 || ((n + 3) / 4) > 65535))
        return set_error(ctx, LA_ERR_ARG,
                         "dgesv_ir: the butterfly form needs FP32 and n <= 262140");
    std::memset(report, 0, sizeof(*report));
    report->prec = prec;
    IrWork w;
    int st = ir_alloc(ctx, n, nrhs, w);
    if (st) return st;
    w.prec = prec;
    w.rbt = rbt;
    report->method = rbt ? LA_IR_METHOD_RBT : LA_IR_METHOD_PIVOTED;
    st = ir_attempt(ctx, n, nrhs, A, lda, B, ldb, X, ldx, prec, tol, max_iter, w, report);
    if (rbt && (st == LA_NOT_CONVERGED || st == LA_SINGULAR)) {
        // the unpivoted factor did not serve: partial pivoting, same budget
        int32_t np = 0;
        LA_HIP_CHECK(ctx, hipMemcpy(&np, w.npert, 4, hipMemcpyDeviceToHost));
        report->perturbed_pivots = np;
        w.rbt = false;
        report->method = LA_IR_METHOD_RBT_THEN_PIVOTED;
        st = ir_attempt(ctx, n, nrhs, A, lda, B, ldb, X, ldx, prec, tol, max_iter, w, report);
    } else if (rbt) {
        int32_t np = 0;
        LA_HIP_CHECK(ctx, hipMemcpy(&np, w.npert, 4, hipMemcpyDeviceToHost));
        report->perturbed_pivots = np;
    }
    return st;
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
