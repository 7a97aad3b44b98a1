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
//
// With LA_IR_GMRES each correction solves A d = r by GMRES on the left-
// preconditioned system M^-1 A d = M^-1 r instead of d = M^-1 r, M the
// low-precision factor (GMRES-IR, Carson and Higham, SIAM J. Sci. Comput.
// 39(6) 2017 and 40(2) 2018).  The Krylov basis, the products with A, the
// orthogonalisation and the update are FP64; only M^-1 applies the
// low-precision factors.  Refinement then converges for kappa * u_f well
// above 1 (BF16 factors at kappa = 1e6), where plain refinement diverges.
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

// A32 = fl32(A) as k_convert_f32, and *amax_bits = max |A32| (as bits)
__global__ __launch_bounds__(256)
void k_convert_f32_amax(int64_t n, const double *__restrict__ A, int64_t lda, float *B,
                        int64_t ldb, unsigned *amax_bits)
{
    // grid-stride over groups of four elements, so that the grid (and the
    // number of atomics on the one maximum word) stays small
    __shared__ float wmax[4];
    float m = 0.f;
    const int64_t total = n * n;
    for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; e < total;
         e += (int64_t)gridDim.x * blockDim.x * 4) {
        const int64_t r = e / n, c = e - r * n;
        if ((n & 3) == 0 && (lda & 1) == 0) {
            const double2 *src = reinterpret_cast<const double2 *>(A + r * lda + c);
            const double2 a = src[0], b = src[1];
            const float4 o = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
            *reinterpret_cast<float4 *>(B + r * ldb + c) = o;
            m = fmaxf(m, fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), fabsf(o.w))));
        } else {
            for (int q = 0; q < 4 && e + q < total; ++q) {
                const int64_t rr = (e + q) / n, cc = (e + q) - rr * n;
                const float f = (float)A[rr * lda + cc];
                B[rr * ldb + cc] = f;
                m = fmaxf(m, fabsf(f));
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(amax_bits, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
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
// right-hand side without per-column guards.  Per column c < nrhs:
// norms[NRM_R + c] = max_i |R_ic|, norms[NRM_X + c] = max_i |X_ic|,
// norms[NRM_B + c] = max_i |B_ic|; norms[NRM_A] = max_i sum_j |A_ij| (only
// when want_anorm).  Lane c of a wave carries column c, so the maxima are
// reduced over the workgroup's waves first and each workgroup issues one
// atomic per column and norm.  Maxima are taken over the bit patterns of
// the non-negative values, so a NaN (above +inf) propagates.
constexpr int RES_RPW = 4;
// norm word layout of k_residual (NRM_WORDS doubles)
enum { NRM_R = 0, NRM_X = 16, NRM_B = 32, NRM_A = 48, NRM_WORDS = 64 };
template <int NRB>
__global__ __launch_bounds__(256)
void k_residual(int64_t n, int nrhs_arg, const double *__restrict__ A, int64_t lda,
                const double *__restrict__ X, int64_t ldx, const double *__restrict__ B,
                int64_t ldb, double *R, int64_t ldr, double *norms, int want_anorm)
{
    __shared__ unsigned long long red[4][3][16];
    __shared__ unsigned long long reda[4];
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
    if (lane < 16) {
        red[wv][0][lane] = mr;
        red[wv][1][lane] = mx;
        red[wv][2][lane] = mb;
    }
    if (lane == 0) reda[wv] = ma;
    __syncthreads();
    if (threadIdx.x < 48) {
        const int q = threadIdx.x >> 4, c = threadIdx.x & 15;
        if (c < nrhs) {
            const u64 v = max(max(red[0][q][c], red[1][q][c]), max(red[2][q][c], red[3][q][c]));
            atomicMax(reinterpret_cast<u64 *>(norms + NRM_R + 16 * q + c), v);
        }
    } else if (threadIdx.x == 48 && want_anorm) {
        const u64 v = max(max(reda[0], reda[1]), max(reda[2], reda[3]));
        atomicMax(reinterpret_cast<u64 *>(norms + NRM_A), v);
    }
}

// eta = max over columns c of |R_c|max / (||A||inf |X_c|max + |B_c|max),
// from the norm words k_residual leaves (a NaN or an Inf propagates).
static double eta_from_norms(const double *h, int64_t nrhs, double anorm)
{
    double eta = 0.0;
    for (int64_t c = 0; c < nrhs; ++c) {
        const double r = h[NRM_R + c], denom = anorm * h[NRM_X + c] + h[NRM_B + c];
        const double e = denom > 0 ? r / denom : (r == 0 ? 0.0 : INFINITY);
        if (e != e) return e;            // a NaN column makes eta NaN
        eta = std::max(eta, e);
    }
    return eta;
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
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, ctx->stream, d_norms,
                       want_anorm ? (int)NRM_WORDS : (int)NRM_A);
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
    // tournament-pivoted (CALU) form: perm / A32 as the pivoted form
    bool calu = false;
    // randomised-butterfly form
    bool rbt = false;
    int64_t np = 0;                      // padded order (multiple of 4)
    double *g = nullptr;                 // butterfly parameters (4 np)
    double *Y = nullptr;                 // transformed right-hand sides (np x nrhs)
    unsigned *amax = nullptr;
    float *tau = nullptr;
    int32_t *npert = nullptr;
    // GMRES-IR: Krylov basis (GM_MAXV vectors of n), work vectors, dot
    // partials and the projection coefficients
    bool gmres = false;
    double *V = nullptr, *gw = nullptr, *gz = nullptr, *zero = nullptr;
    double *partial = nullptr, *hcoef = nullptr;
};

constexpr int GM_RESTART = 200;                // Krylov vectors per correction
constexpr int GM_MAXV = GM_RESTART + 1;         // <= 256: one k_dot_reduce thread each
static_assert(GM_MAXV <= 256, "k_dot_reduce has 256 threads");
constexpr int GM_DOT_BLOCKS = 128;             // row chunks of the dot kernel
constexpr double GM_TOL = 1e-6;                // relative reduction per correction

static int gmres_alloc(la_ctx *ctx, int64_t n, IrWork &w)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bv = al((size_t)n * 8);
    const size_t total = (GM_MAXV + 3) * bv + al((size_t)GM_DOT_BLOCKS * GM_MAXV * 8) +
                         al((size_t)GM_MAXV * 8);
    void *base = nullptr;
    int st = ws_get(ctx, WS_IR, total, &base);
    if (st) return st;
    char *p = (char *)base;
    w.V = (double *)p; p += GM_MAXV * bv;
    w.gw = (double *)p; p += bv;
    w.gz = (double *)p; p += bv;
    w.zero = (double *)p; p += bv;
    w.partial = (double *)p; p += al((size_t)GM_DOT_BLOCKS * GM_MAXV * 8);
    w.hcoef = (double *)p;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream,
                       w.zero, (int)n);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

static int ir_alloc(la_ctx *ctx, int64_t n, int64_t nrhs, IrWork &w)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t np = (n + 3) & ~(int64_t)3;       // butterfly order
    const int64_t ld32 = (np + 63) & ~(int64_t)63;  // pad rows to 256 B
    size_t b32 = al((size_t)np * ld32 * sizeof(float));
    size_t bperm = al((size_t)n * 8), bipiv = al((size_t)n * 8);
    size_t brd = al((size_t)n * nrhs * 8);
    size_t bn = al(NRM_WORDS * 8), bst = al(4 * 4), bsc = al((size_t)n * 8);
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

// ------------------------------------------------------------- GMRES-IR
// partial[k * GM_DOT_BLOCKS + blk] = sum over the block's rows of V_k[i] w[i]
// (grid: GM_DOT_BLOCKS x nv); reduced in a fixed order by k_dot_reduce, so
// the coefficients are the same in every run.
__global__ __launch_bounds__(256)
void k_dot_partial(int64_t n, const double *__restrict__ V, int64_t ldv,
                   const double *__restrict__ w, double *partial)
{
    __shared__ double red[4];
    const int k = blockIdx.y;
    const int64_t chunk = (n + GM_DOT_BLOCKS - 1) / GM_DOT_BLOCKS;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    const double *v = V + (int64_t)k * ldv;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc = fma(v[i], w[i], acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[k * GM_DOT_BLOCKS + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_dot_reduce(int nv, const double *partial, double *h)
{
    const int k = threadIdx.x;
    if (k >= nv) return;
    double s = 0.0;
    for (int b = 0; b < GM_DOT_BLOCKS; ++b) s += partial[k * GM_DOT_BLOCKS + b];
    h[k] = s;
}

// w -= sum_k h[k] V_k  (one row per thread)
__global__ void k_sub_basis(int64_t n, int nv, const double *__restrict__ V, int64_t ldv,
                            const double *__restrict__ h, double *w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = w[i];
    for (int k = 0; k < nv; ++k) acc = fma(-h[k], V[(int64_t)k * ldv + i], acc);
    w[i] = acc;
}

// out = s * in (strided out, for a column of the correction block)
__global__ void k_scale_vec(int64_t n, double s, const double *in, int64_t ldi, double *out,
                            int64_t ldo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i * ldo] = s * in[i * ldi];
}

// out[i * ldo] = sum_k y[k] V_k[i]
__global__ void k_combine_basis(int64_t n, int nv, const double *__restrict__ V, int64_t ldv,
                                const double *__restrict__ y, double *out, int64_t ldo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < nv; ++k) acc = fma(y[k], V[(int64_t)k * ldv + i], acc);
    out[i * ldo] = acc;
}

// h[0:nv] = V[0:nv]^T x, copied to host (synchronises the stream)
static int basis_dots(la_ctx *ctx, int64_t n, int nv, const IrWork &w, const double *x,
                      double *h_host)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(GM_DOT_BLOCKS, nv), dim3(256), 0, ctx->stream, n,
                       w.V, n, x, w.partial);
    hipLaunchKernelGGL(k_dot_reduce, dim3(1), dim3(256), 0, ctx->stream, nv, w.partial, w.hcoef);
    LA_HIP_CHECK(ctx, hipGetLastError());
    LA_HIP_CHECK(ctx, hipMemcpyAsync(h_host, w.hcoef, nv * sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

// ||x||_2 of a device vector (synchronises)
static int vec_norm2(la_ctx *ctx, int64_t n, const IrWork &w, const double *x, double *nrm)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(GM_DOT_BLOCKS, 1), dim3(256), 0, ctx->stream, n, x, n,
                       x, w.partial);
    hipLaunchKernelGGL(k_dot_reduce, dim3(1), dim3(64), 0, ctx->stream, 1, w.partial, w.hcoef);
    LA_HIP_CHECK(ctx, hipGetLastError());
    double v = 0.0;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(&v, w.hcoef, sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *nrm = std::sqrt(v);
    return LA_OK;
}

// One GMRES-IR correction for every column c of R (n x nrhs, ld nrhs):
// D[:, c] approximately solves A d = R[:, c]; GMRES (restarted never: at
// most GM_RESTART iterations) on M^-1 A d = M^-1 r, classical Gram-Schmidt
// applied twice, Givens rotations on the host.  *inner accumulates the
// iterations taken.
static int gmres_correction(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A, int64_t lda,
                            const IrWork &w, const double *R, double *D, int *inner)
{
    int st;
    const unsigned gb = (unsigned)ceil_div(n, 256);
    double *V = w.V;
    for (int64_t c = 0; c < nrhs; ++c) {
        // v_0 = z / beta, z = M^-1 r
        if ((st = lp_solve(ctx, n, 1, w, R + c, nrhs, w.gz, 1))) return st;
        double beta = 0.0;
        if ((st = vec_norm2(ctx, n, w, w.gz, &beta))) return st;
        if (!(beta > 0.0) || !std::isfinite(beta)) {
            hipLaunchKernelGGL(k_scale_vec, dim3(gb), dim3(256), 0, ctx->stream, n, 0.0, w.gz, 1,
                               D + c, nrhs);
            LA_HIP_CHECK(ctx, hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL(k_scale_vec, dim3(gb), dim3(256), 0, ctx->stream, n, 1.0 / beta, w.gz,
                           1, V, 1);
        LA_HIP_CHECK(ctx, hipGetLastError());
        // H is (GM_MAXV x GM_RESTART), column-major by iteration
        std::vector<double> H((size_t)GM_MAXV * GM_RESTART, 0.0), cs(GM_RESTART), sn(GM_RESTART),
            g(GM_MAXV, 0.0), h1(GM_MAXV), h2(GM_MAXV);
        g[0] = beta;
        int j = 0;
        for (; j < GM_RESTART; ++j) {
            // w = M^-1 A v_j:  R' = 0 - A v_j, then -M^-1 R'
            if ((st = residual(ctx, n, 1, A, lda, V + (int64_t)j * n, 1, w.zero, 1, w.gw, 1,
                               w.norms, false)))
                return st;
            if ((st = lp_solve(ctx, n, 1, w, w.gw, 1, w.gz, 1))) return st;
            hipLaunchKernelGGL(k_scale_vec, dim3(gb), dim3(256), 0, ctx->stream, n, -1.0, w.gz,
                               1, w.gz, 1);
            LA_HIP_CHECK(ctx, hipGetLastError());
            // classical Gram-Schmidt, twice
            const int nv = j + 1;
            if ((st = basis_dots(ctx, n, nv, w, w.gz, h1.data()))) return st;
            hipLaunchKernelGGL(k_sub_basis, dim3(gb), dim3(256), 0, ctx->stream, n, nv, V, n,
                               w.hcoef, w.gz);
            if ((st = basis_dots(ctx, n, nv, w, w.gz, h2.data()))) return st;
            hipLaunchKernelGGL(k_sub_basis, dim3(gb), dim3(256), 0, ctx->stream, n, nv, V, n,
                               w.hcoef, w.gz);
            LA_HIP_CHECK(ctx, hipGetLastError());
            double hn = 0.0;
            if ((st = vec_norm2(ctx, n, w, w.gz, &hn))) return st;
            double *hc = &H[(size_t)j * GM_MAXV];
            for (int k = 0; k < nv; ++k) hc[k] = h1[k] + h2[k];
            hc[nv] = hn;
            if (hn > 0.0 && std::isfinite(hn)) {
                hipLaunchKernelGGL(k_scale_vec, dim3(gb), dim3(256), 0, ctx->stream, n, 1.0 / hn,
                                   w.gz, 1, V + (int64_t)nv * n, 1);
                LA_HIP_CHECK(ctx, hipGetLastError());
            }
            // apply the previous rotations, then form the new one
            for (int k = 0; k < j; ++k) {
                const double t = cs[k] * hc[k] + sn[k] * hc[k + 1];
                hc[k + 1] = -sn[k] * hc[k] + cs[k] * hc[k + 1];
                hc[k] = t;
            }
            const double rr = std::hypot(hc[j], hc[j + 1]);
            cs[j] = rr > 0.0 ? hc[j] / rr : 1.0;
            sn[j] = rr > 0.0 ? hc[j + 1] / rr : 0.0;
            hc[j] = rr;
            hc[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++*inner;
            if (!(std::fabs(g[j + 1]) > GM_TOL * beta) || !(hn > 0.0) || !std::isfinite(hn)) {
                ++j;
                break;
            }
        }
        // y = H(0:j, 0:j)^-1 g(0:j), d = V y
        std::vector<double> y(j, 0.0);
        for (int k = j - 1; k >= 0; --k) {
            double t = g[k];
            for (int q = k + 1; q < j; ++q) t -= H[(size_t)q * GM_MAXV + k] * y[q];
            const double d = H[(size_t)k * GM_MAXV + k];
            y[k] = d != 0.0 ? t / d : 0.0;
        }
        LA_HIP_CHECK(ctx, hipMemcpy(w.hcoef, y.data(), j * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_combine_basis, dim3(gb), dim3(256), 0, ctx->stream, n, j, V, n,
                           w.hcoef, D + c, nrhs);
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
    } else if (w.calu) {
        if ((st = fill_i32(ctx, (int32_t *)w.amax, 1, 0))) return st;
        hipLaunchKernelGGL(k_convert_f32_amax,
                           dim3((unsigned)std::min<int64_t>(ceil_div(n * n, 1024), 4096)),
                           dim3(256), 0, ctx->stream, n, A, lda, w.A32, w.ld32, w.amax);
        hipLaunchKernelGGL(k_rbt_tau, dim3(1), dim3(64), 0, ctx->stream, w.amax, w.tau);
        LA_HIP_CHECK(ctx, hipGetLastError());
        if ((st = fill_i32(ctx, w.status, 1, 0))) return st;
        if ((st = iota_u64(ctx, w.perm, n))) return st;
        if ((st = getrf_nopiv_core<float>(ctx, n, n, w.A32, w.ld32, w.tau, w.npert, prec, w.perm)))
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
    double h[NRM_WORDS] = {};
    int it = 0;
    double eta = INFINITY, eta_prev = INFINITY;
    for (;; ++it) {
        if ((st = residual(ctx, n, nrhs, A, lda, X, ldx, B, ldb, w.R, nrhs, w.norms, it == 0)))
            return st;
        LA_HIP_CHECK(ctx, hipMemcpyAsync(h, w.norms, NRM_WORDS * sizeof(double),
                                         hipMemcpyDeviceToHost, ctx->stream));
        LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (it == 0) report->anorm_inf = h[NRM_A];
        eta = eta_from_norms(h, nrhs, report->anorm_inf);
        if (it == 0) report->initial_backward_error = eta;
        if (!(eta > tol) || it >= max_iter || !std::isfinite(eta)) break;
        // the unpivoted factor is abandoned for the pivoted one as soon as
        // refinement stagnates (less than a halving per step) or runs long
        if ((w.rbt || w.calu) && ((it >= 2 && eta > 0.5 * eta_prev) || it >= RBT_MAX_STEPS))
            break;
        eta_prev = eta;
        if (w.gmres) {
            if ((st = gmres_correction(ctx, n, nrhs, A, lda, w, w.R, w.D, &report->inner_iterations)))
                return st;
        } else if ((st = lp_solve(ctx, n, nrhs, w, w.R, nrhs, w.D, nrhs))) {
            return st;
        }
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
    const bool gm = (prec & LA_IR_GMRES) != 0;
    const bool calu = (prec & LA_IR_TOURNAMENT) != 0;
    prec &= ~(LA_IR_RBT | LA_IR_GMRES | LA_IR_TOURNAMENT);
    if (prec != LA_PREC_FP32 && prec != LA_PREC_BF16 && prec != LA_PREC_FP16)
        return set_error(ctx, LA_ERR_ARG, "dgesv_ir: precision not supported");
    if (rbt && (prec != LA_PREC_FP32 || ((n + 3) / 4) > 65535))
        return set_error(ctx, LA_ERR_ARG,
                         "dgesv_ir: the butterfly form needs FP32 and n <= 262140");
    if (calu && (rbt || prec != LA_PREC_FP32))
        return set_error(ctx, LA_ERR_ARG,
                         "dgesv_ir: tournament pivoting needs FP32 and excludes the butterfly form");
    std::memset(report, 0, sizeof(*report));
    report->prec = prec;
    IrWork w;
    int st = ir_alloc(ctx, n, nrhs, w);
    if (st) return st;
    w.prec = prec;
    w.rbt = rbt;
    w.calu = calu;
    w.gmres = gm;
    if (gm && (st = gmres_alloc(ctx, n, w))) return st;
    report->method = rbt ? LA_IR_METHOD_RBT : calu ? LA_IR_METHOD_TOURNAMENT : LA_IR_METHOD_PIVOTED;
    st = ir_attempt(ctx, n, nrhs, A, lda, B, ldb, X, ldx, prec, tol, max_iter, w, report);
    if ((rbt || calu) && (st == LA_NOT_CONVERGED || st == LA_SINGULAR)) {
        // the factor did not serve: partial pivoting, same budget
        int32_t np = 0;
        LA_HIP_CHECK(ctx, hipMemcpy(&np, w.npert, 4, hipMemcpyDeviceToHost));
        report->perturbed_pivots = np;
        w.rbt = false;
        w.calu = false;
        report->method = rbt ? LA_IR_METHOD_RBT_THEN_PIVOTED : LA_IR_METHOD_TOURNAMENT_THEN_PIVOTED;
        st = ir_attempt(ctx, n, nrhs, A, lda, B, ldb, X, ldx, prec, tol, max_iter, w, report);
    } else if (rbt || calu) {
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
    const size_t rbytes = ((size_t)n * nrhs * 8 + 255) & ~(size_t)255;
    int st = ws_get(ctx, WS_IR, rbytes + NRM_WORDS * 8, &base);
    if (st) return st;
    double *R = (double *)base;
    double *norms = (double *)((char *)base + rbytes);
    if ((st = residual(ctx, n, nrhs, A, lda, X, ldx, B, ldb, R, nrhs, norms, true))) return st;
    double h[NRM_WORDS];
    LA_HIP_CHECK(ctx, hipMemcpyAsync(h, norms, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *eta = eta_from_norms(h, nrhs, h[NRM_A]);
    return LA_OK;
}
