// Householder QR (QRDecomposition, src/decomp/qr.rs:26-237) and the
// pseudo-inverse (Matrix::pinverse, src/matrix/mod.rs:1049-1057).
//
// The factor is the reference's: for each column c < min(m, n),
//   a_c = -sign(x_cc) |x|  (+|x| when x_cc <= 0),   x = A[c.., c],
//   u = x - a_c e_c stored in place (A[c][c] = x_cc - a_c),  rdiag[c] = a_c,
//   every later column y -> y + <y,u>/(a_c u_c) u  (the reflection H_c),
// skipped when a_c = 0.  It is computed blocked, 32 columns per panel:
//   * inside the panel, column by column in one cooperative kernel
//     (k_qr_panel) with one exchange between workgroups per column: the
//     norm of x and the products <x, y> for the panel's other columns are
//     summed in one pass, and <u, y> = <x, y> - a_c y_c then needs only the
//     diagonal row, which travels in the same exchange; every sum is taken
//     in a fixed order, so the result is the same in every run and in every
//     workgroup.  Panels of more than 65536 rows use two small launches per
//     column instead (k_qr_dots, k_qr_apply);
//   * the panel's reflections as one block, H_c H_c+1 ... = I - V T V^T
//     (T upper triangular, T_jj = tau_j = -1/(a_j u_j), forward column-wise
//     accumulation of Schreiber and Van Loan), applied to the trailing
//     columns as A2 -= V (T^T (V^T A2)): V^T A2 (32 rows, long K) split over
//     row chunks and summed in chunk order, the other two on MFMA.
// get_q applies the blocks to the partial identity in reverse order
// (Q = (I - V T V^T) Q), the least-squares solve applies them to B in
// forward order and finishes with R x = (Q^T B)[0..n).
#include "la_common.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

namespace lahip {

constexpr int QR_NB = 32;       // panel width
constexpr int QR_PARTS = 256;   // max workgroups of the column kernels
constexpr int QRP_MAXP = 128;   // max workgroups of the cooperative panel
constexpr int QRP_SW = 4 * QR_NB;   // slot words: 32 partial sums, then (workgroup 0) the diagonal row
constexpr int QRP_LS = QR_NB + 1;   // LDS row stride of the panel rows (doubles)

__device__ inline void put_tagged(uint64_t *w, uint32_t tag, double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v), t = (uint64_t)tag << 32;
    st_dev(w + 0, t | (b & 0xffffffffull));
    st_dev(w + 1, t | (b >> 32));
}

// Waits (bounded) until the 64-bit word at p carries `tag` in its high half.
// Returns false when the wait is abandoned: the sticky status is set (by
// this or another workgroup) or the spin limit passes.
__device__ inline bool wait_tagged(const uint64_t *p, uint32_t tag, uint64_t &v,
                                   int32_t *status, uint64_t spin_ticks)
{
    if ((uint32_t)(v >> 32) == tag) return true;
    const uint64_t t0 = wall_clock64();
    for (uint32_t spin = 0;; ++spin) {
        v = ld_dev(p);
        if ((uint32_t)(v >> 32) == tag) return true;
        if ((spin & 31) == 31 && __hip_atomic_load(status, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
        if (wall_clock64() - t0 > spin_ticks) {
            atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The panel of columns [k0, k0 + nb) over rows [k0, m) in one cooperative
// launch of P workgroups; workgroup w keeps rows [k0 + w R, k0 + (w+1) R),
// R = 256 RPT, in LDS.  Per column c (global gc = k0 + c, whose diagonal row
// is row c of workgroup 0):
//   1. partials over the workgroup's rows r >= gc: q_j = sum_r x_r A[r][j]
//      for j in [c, nb) (j = c is |x|^2), thread (g, j) over row group g,
//      then the eight groups in order;
//   2. publish q (and, workgroup 0, the diagonal row A[gc][c..nb)) as tagged
//      words, buffers alternating by column parity;
//   3. every workgroup reads all P partials (thread (g, j) the workgroups
//      g, g + 8, ...) and sums them in one fixed order;
//   4. a = -sign(x_c)|x|, u_c = x_c - a, f_j = (q_j - a A[gc][j]) / (a u_c),
//      and its rows r >= gc get A[r][j] += f_j u_r (u_gc = u_c), A[gc][c] = u_c.
// With want_vg the kernel also writes its rows of V (the Householder vectors,
// as k_qr_make_v builds them) and its partial Gram matrix V_w^T V_w (on MFMA,
// a fixed order) as chunk w of the partial products that k_qr_vtc_reduce
// adds up, so the block's V^T V needs no pass over V.
// A workgroup can publish column c + 2 into a buffer only after every
// workgroup published column c + 1, which each did after reading column c,
// so the alternating buffers are never overwritten while read.
template <int RPT, int KMAX>
__global__ __launch_bounds__(256)
void k_qr_panel(int m, double *__restrict__ A, int64_t lda, int k0, int nb,
                double *__restrict__ rdiag, uint64_t *__restrict__ slots,
                int32_t *__restrict__ status, uint64_t spin_ticks, double *__restrict__ V,
                double *__restrict__ gpart, int want_vg)
{
    constexpr int R = 256 * RPT;
    constexpr int GR = 32 * RPT;                // rows per group of 32 threads
    // KMAX >= ceil(P / 8): workgroups read per thread
    extern __shared__ __align__(16) double qx[];    // R x QRP_LS
    __shared__ double red[8][QR_NB], red2[8][QR_NB];
    __shared__ double tot[QR_NB], drow[QR_NB], fac[QR_NB], s_av[QR_NB];
    __shared__ int s_abort;
    const int t = threadIdx.x, w = blockIdx.x, P = gridDim.x;
    const int r0 = k0 + w * R;
    // sixteen loads per thread (clamped, in-bounds positions) issued before
    // any is stored: one memory round trip per sixteen elements
    for (int i0 = t; i0 < R * QR_NB; i0 += 16 * 256) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int i = i0 + 256 * k, lr = i / QR_NB, j = i % QR_NB;
            v[k] = A[(int64_t)min(r0 + lr, m - 1) * lda + k0 + min(j, nb - 1)];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int i = i0 + 256 * k, lr = i / QR_NB, j = i % QR_NB;
            qx[lr * QRP_LS + j] = (r0 + lr < m && j < nb) ? v[k] : 0.0;
        }
    }
    if (t == 0) s_abort = 0;
    if (t < QR_NB) s_av[t] = 0.0;
    __syncthreads();
    const int g = t >> 5, jt = t & 31;
    for (int c = 0; c < nb; ++c) {
        const int gc = k0 + c;
        const uint32_t tag = (uint32_t)gc;
        const int buf = c & 1;
        // 1. this workgroup's partial sums: four interleaved accumulators
        //    (rows q mod 4), added as (a0 + a1) + (a2 + a3), so the
        //    dependent FP64 chain is GR / 4 fmas long instead of GR
        {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q = 0; q < GR; q += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int lr = g * GR + q + u;
                    const double x = r0 + lr >= gc ? qx[lr * QRP_LS + c] : 0.0;
                    acc[u] = fma(x, qx[lr * QRP_LS + jt], acc[u]);
                }
            }
            red[g][jt] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
        __syncthreads();
        // 2. publish
        if (t < QR_NB && t >= c && t < nb) {
            double sq = 0.0;
            for (int q = 0; q < 8; ++q) sq += red[q][t];
            uint64_t *slot = slots + ((size_t)buf * QRP_MAXP + w) * QRP_SW;
            put_tagged(slot + 2 * t, tag, sq);
            if (w == 0) put_tagged(slot + 2 * QR_NB + 2 * t, tag, qx[c * QRP_LS + t]);
        }
        // 3. gather: loads clamped to words that will carry this tag, so every
        // wait ends; out-of-range ones are not summed
        {
            const int jc = min(max(jt, c), nb - 1);
            const bool jon = jt >= c && jt < nb;
            uint64_t v[2 * KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                const int ww = min(g + 8 * k, P - 1);
                const uint64_t *src = slots + ((size_t)buf * QRP_MAXP + ww) * QRP_SW + 2 * jc;
                v[2 * k] = ld_dev(src);
                v[2 * k + 1] = ld_dev(src + 1);
            }
            uint64_t dv[2] = {0, 0};
            const uint64_t *dsrc = slots + (size_t)buf * QRP_MAXP * QRP_SW + 2 * QR_NB + 2 * jc;
            if (g == 0) { dv[0] = ld_dev(dsrc); dv[1] = ld_dev(dsrc + 1); }
            bool ok = true;
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                const int ww = min(g + 8 * k, P - 1);
                const uint64_t *src = slots + ((size_t)buf * QRP_MAXP + ww) * QRP_SW + 2 * jc;
                ok = ok && wait_tagged(src, tag, v[2 * k], status, spin_ticks) &&
                     wait_tagged(src + 1, tag, v[2 * k + 1], status, spin_ticks);
                const double d = __longlong_as_double(
                    (long long)((v[2 * k] & 0xffffffffull) | (v[2 * k + 1] << 32)));
                if (g + 8 * k < P) sum += d;
            }
            if (g == 0) {
                ok = ok && wait_tagged(dsrc, tag, dv[0], status, spin_ticks) &&
                     wait_tagged(dsrc + 1, tag, dv[1], status, spin_ticks);
                drow[jt] = jon ? __longlong_as_double((long long)((dv[0] & 0xffffffffull) |
                                                                  (dv[1] << 32)))
                               : 0.0;
            }
            red2[g][jt] = jon ? sum : 0.0;
            if (!ok) s_abort = 1;
        }
        __syncthreads();
        if (s_abort) break;
        if (t < QR_NB) {
            double s2 = 0.0;
            for (int q = 0; q < 8; ++q) s2 += red2[q][t];
            tot[t] = s2;
        }
        __syncthreads();
        // 4. the reflection
        const double xc = drow[c];
        const double av = xc > 0.0 ? -sqrt(tot[c]) : sqrt(tot[c]);
        const double uc = xc - av;
        if (t < QR_NB)
            fac[t] = (av != 0.0 && t > c && t < nb) ? (tot[t] - av * drow[t]) / (av * uc) : 0.0;
        __syncthreads();
        if (av != 0.0) {
            for (int q = 0; q < RPT; ++q) {
                const int lr = t + 256 * q, gr = r0 + lr;
                if (gr < gc || gr >= m) continue;
                double *row = qx + lr * QRP_LS;
                const double u = gr == gc ? uc : row[c];
                for (int j = c + 1; j < nb; ++j) row[j] = fma(fac[j], u, row[j]);
                if (gr == gc) row[c] = uc;
            }
        }
        if (w == 0 && t == 0) rdiag[gc] = av;
        if (t == 0) s_av[c] = av;
        __syncthreads();
    }
    for (int i = t; i < R * QR_NB; i += 256) {
        const int lr = i / QR_NB, j = i % QR_NB;
        if (r0 + lr < m && j < nb) A[(int64_t)(r0 + lr) * lda + k0 + j] = qx[lr * QRP_LS + j];
    }
    if (!want_vg || s_abort) return;
    // V in place: column j zero above row k0 + j and where a_j = 0 (rows past
    // m and columns past nb are zero already)
    __syncthreads();
    for (int i = t; i < R * QR_NB; i += 256) {
        const int lr = i / QR_NB, j = i % QR_NB;
        if (!(r0 + lr >= k0 + j && s_av[j] != 0.0)) qx[lr * QRP_LS + j] = 0.0;
    }
    __syncthreads();
    for (int i = t; i < R * QR_NB; i += 256) {
        const int lr = i / QR_NB, j = i % QR_NB;
        if (r0 + lr < m && j < nb) V[(int64_t)(r0 + lr - k0) * nb + j] = qx[lr * QRP_LS + j];
    }
    // V_w^T V_w on v_mfma_f64_16x16x4: wave wv owns the 16 x 16 tile
    // (16 ti, 16 tj) and runs K over the workgroup's rows four at a time;
    // lane l supplies A[l%16][l/16] = V[k][16 ti + l%16] and
    // B[l/16][l%16] = V[k][16 tj + l%16], k = row l/16 of the step, and
    // receives D[l/16 + 4 r][l%16]
    {
        const int lane = t & 63, wv = t >> 6, ti = wv >> 1, tj = wv & 1;
        d4_t acc = {0.0, 0.0, 0.0, 0.0};
        for (int k4 = 0; k4 < R; k4 += 4) {
            const double* rowp = qx + (k4 + (lane >> 4)) * QRP_LS;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rowp[16 * ti + (lane & 15)],
                                                       rowp[16 * tj + (lane & 15)], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = 16 * ti + (lane >> 4) + 4 * r, gj = 16 * tj + (lane & 15);
            if (gi < nb && gj < nb) gpart[((int64_t)w * QR_NB + gi) * nb + gj] = acc[r];
        }
    }
}

// Row r of the panel's current column is handled by thread (r - r0) of the
// grid (rows beyond the grid's size wrap by the grid size).

// Partial sums of x^2 over rows [r0, m) of column c.
__global__ __launch_bounds__(256)
void k_qr_norm_partial(int m, const double *__restrict__ A, int64_t lda, int c,
                       double *__restrict__ part)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int stride = gridDim.x * 256;
    double s = 0.0;
    for (int r = c + blockIdx.x * 256 + t; r < m; r += stride) {
        const double x = A[(int64_t)r * lda + c];
        s = fma(x, x, s);
    }
    red[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = red[0];
}

// Sum of the P partials in a fixed tree order; every workgroup gets the same
// value.  Needs a 256-entry LDS buffer; all threads receive the sum.
__device__ inline double tree_sum(const double *__restrict__ part, int P, double *red)
{
    const int t = threadIdx.x;
    red[t] = t < P ? part[t] : 0.0;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// a_c, u_c from the column's norm (every workgroup, same order), then the
// partial dot products <u, A[., col]> for the panel columns col in (c, cend).
// Workgroup 0 stores (a_c, u_c) in scal[2c..2c+1].
__global__ __launch_bounds__(256)
void k_qr_dots(int m, const double *__restrict__ A, int64_t lda, int c, int cend,
               const double *__restrict__ npart, int P, double *__restrict__ scal,
               double *__restrict__ dpart)
{
    __shared__ double red[256];
    __shared__ double dsum[8][QR_NB];
    const int t = threadIdx.x;
    const double nrm2 = tree_sum(npart, P, red);
    const double x0 = A[(int64_t)c * lda + c];
    const double av = x0 > 0.0 ? -sqrt(nrm2) : sqrt(nrm2);
    const double u0 = x0 - av;
    if (blockIdx.x == 0 && t == 0) { scal[2 * c] = av; scal[2 * c + 1] = u0; }
    const int ncol = cend - c - 1;
    if (av == 0.0 || ncol <= 0) return;
    double acc[QR_NB - 1];
#pragma unroll
    for (int j = 0; j < QR_NB - 1; ++j) acc[j] = 0.0;
    const int stride = gridDim.x * 256;
    for (int r = c + blockIdx.x * 256 + t; r < m; r += stride) {
        const double *row = A + (int64_t)r * lda;
        const double u = (r == c) ? u0 : row[c];
#pragma unroll
        for (int j = 0; j < QR_NB - 1; ++j)
            if (j < ncol) acc[j] = fma(u, row[c + 1 + j], acc[j]);
    }
    // workgroup sum per column: eight groups of 32 lanes in a fixed order
    // (j is static in the unrolled loop, ncol is the same in every thread)
    const int g = t >> 5, l = t & 31;
#pragma unroll
    for (int j = 0; j < QR_NB - 1; ++j) {
        if (j < ncol) {
            red[t] = acc[j];
            __syncthreads();
            if (l == 0) {
                double s = 0.0;
                for (int q = 0; q < 32; ++q) s += red[32 * g + q];
                dsum[g][j] = s;
            }
            __syncthreads();
        }
    }
    if (t < ncol) {
        double s = 0.0;
        for (int q = 0; q < 8; ++q) s += dsum[q][t];
        dpart[(int64_t)blockIdx.x * QR_NB + t] = s;
    }
}

// y += <y,u>/(a_c u_c) u for the panel columns in (c, cend), then the partial
// norm of column c + 1 over rows > c (if c + 1 < cend), and A[c][c] = u_c,
// rdiag[c] = a_c.  The dot products are summed in the same fixed order by
// every workgroup.
__global__ __launch_bounds__(256)
void k_qr_apply(int m, double *__restrict__ A, int64_t lda, int c, int cend,
                const double *__restrict__ dpart, int P, const double *__restrict__ scal,
                double *__restrict__ rdiag, double *__restrict__ npart)
{
    __shared__ double fac[QR_NB];
    __shared__ double red[256];
    const int t = threadIdx.x;
    const double av = scal[2 * c], u0 = scal[2 * c + 1];
    const int ncol = cend - c - 1;
    if (t < QR_NB) {
        double f = 0.0;
        if (av != 0.0 && t < ncol) {
            double s = 0.0;
            for (int b = 0; b < P; ++b) s += dpart[(int64_t)b * QR_NB + t];
            f = s / (av * u0);
        }
        fac[t] = f;
    }
    __syncthreads();
    const int stride = gridDim.x * 256;
    double nrm = 0.0;
    for (int r = c + blockIdx.x * 256 + t; r < m; r += stride) {
        double *row = A + (int64_t)r * lda;
        const double u = (r == c) ? u0 : row[c];
        if (av != 0.0) {
            for (int j = 0; j < ncol; ++j) {
                const double y = fma(fac[j], u, row[c + 1 + j]);
                row[c + 1 + j] = y;
                if (j == 0 && r > c) nrm = fma(y, y, nrm);
            }
        } else if (ncol > 0 && r > c) {
            const double y = row[c + 1];
            nrm = fma(y, y, nrm);
        }
    }
    if (blockIdx.x == 0 && t == 0) {
        rdiag[c] = av;
        if (av != 0.0) A[(int64_t)c * lda + c] = u0;
    }
    if (ncol > 0) {
        red[t] = nrm;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) red[t] += red[t + h];
            __syncthreads();
        }
        if (t == 0) npart[blockIdx.x] = red[0];
    }
}

// W = V^T C (nb x cols) for a tall V (rows x nb, ld nb) and C (rows x cols,
// ld ldc): a product with a short output and a long K, split over row chunks
// (blockIdx.y) whose partial products are summed by k_qr_vtc_reduce in chunk
// order.  Thread t owns column (t & 127) of a 128-column tile and 16 of the
// 32 rows of W; V is staged in LDS 64 rows at a time, zero past nb.
constexpr int VT_COLS = 128, VT_SUB = 64, VT_MAXCHUNK = 16;
__global__ __launch_bounds__(256)
void k_qr_vtc_partial(int rows, int nb, int cols, int rch, const double *__restrict__ V,
                      const double *__restrict__ C, int64_t ldc, double *__restrict__ P)
{
    __shared__ double vs[VT_SUB][QR_NB];
    const int t = threadIdx.x;
    const int col = blockIdx.x * VT_COLS + (t & 127), jh = (t >> 7) * 16;
    const int r0 = blockIdx.y * rch, r1 = min(rows, r0 + rch);
    const bool okc = col < cols;
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    for (int rs = r0; rs < r1; rs += VT_SUB) {
        const int rn = min(VT_SUB, r1 - rs);
        for (int e = t; e < VT_SUB * QR_NB; e += 256) {
            const int rr = e / QR_NB, j = e - rr * QR_NB;
            vs[rr][j] = (rr < rn && j < nb) ? V[(int64_t)(rs + rr) * nb + j] : 0.0;
        }
        __syncthreads();
        const double *cp = C + (int64_t)rs * ldc + (okc ? col : 0);
        // eight rows of C in flight per step (clamped loads, uniform guard)
        for (int rr0 = 0; rr0 < rn; rr0 += 8) {
            double cv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) cv[u] = cp[(int64_t)min(rr0 + u, rn - 1) * ldc];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (rr0 + u < rn) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) acc[j] = fma(vs[rr0 + u][jh + j], cv[u], acc[j]);
                }
            }
        }
        __syncthreads();
    }
    if (okc)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (jh + j < nb) P[((int64_t)blockIdx.y * QR_NB + jh + j) * cols + col] = acc[j];
}

__global__ __launch_bounds__(256)
void k_qr_vtc_reduce(int nchunks, int nb, int cols, const double *__restrict__ P,
                     double *__restrict__ W)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)nb * cols) return;
    const int j = (int)(e / cols), col = (int)(e - (int64_t)j * cols);
    double s = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) s += P[((int64_t)ch * QR_NB + j) * cols + col];
    W[(int64_t)j * cols + col] = s;
}

// V (rows x nb, ld nb) of the panel's Householder vectors: column j is zero
// above row j, u_j at row j (already stored in A), A below; columns whose
// a_j = 0 are zero.
__global__ __launch_bounds__(256)
void k_qr_make_v(int rows, int nb, const double *__restrict__ A, int64_t lda,
                 const double *__restrict__ rd, double *__restrict__ V)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)rows * nb) return;
    const int r = (int)(e / nb), j = (int)(e - (int64_t)r * nb);
    double v = 0.0;
    if (r >= j && rd[j] != 0.0) v = A[(int64_t)r * lda + j];
    V[(int64_t)r * nb + j] = v;
}

// T (nb x nb, upper) of I - V T V^T = H_0 ... H_nb-1 from the Gram matrix
// G = V^T V: T_jj = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j]
// (tau_j = -1/(a_j u_j), 0 for a skipped column).  One workgroup; also
// writes Tt = T^T.
__global__ __launch_bounds__(64)
void k_qr_larft(int nb, const double *__restrict__ G, const double *__restrict__ rd,
                const double *__restrict__ diagu, int64_t ldu, double *__restrict__ T,
                double *__restrict__ Tt)
{
    __shared__ double sT[QR_NB][QR_NB + 1], sG[QR_NB][QR_NB + 1];
    __shared__ double stau[QR_NB];
    const int t = threadIdx.x;
    // G, rdiag and the diagonal of u staged in LDS first: the recurrence
    // below then issues no global load
    for (int e = t; e < QR_NB * QR_NB; e += 64) {
        const int i = e / QR_NB, j = e % QR_NB;
        sT[i][j] = 0.0;
        const double g = G[min(i, nb - 1) * nb + min(j, nb - 1)];
        sG[i][j] = (i < nb && j < nb) ? g : 0.0;
    }
    if (t < QR_NB) {
        const int jc = min(t, nb - 1);
        const double a = rd[jc], u = diagu[(int64_t)jc * ldu + jc];
        stau[t] = (t < nb && a != 0.0) ? -1.0 / (a * u) : 0.0;
    }
    __syncthreads();
    // thread t < 32 keeps row t of T in registers; per column j it forms
    // w_t = sum_{q = t}^{j-1} T[t][q] G[q][j] (q ascending, as the textbook
    // recurrence) from LDS broadcasts of G's column, with no barrier
    double tr[QR_NB];
#pragma unroll
    for (int q = 0; q < QR_NB; ++q) tr[q] = 0.0;
#pragma unroll
    for (int j = 0; j < QR_NB; ++j) {
        if (j < nb) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < j; ++q) {
                const double g = sG[q][j];
                if (q >= t) s += tr[q] * g;
            }
            if (t < j) tr[j] = -stau[j] * s;
            else if (t == j) tr[j] = stau[j];
        }
    }
    if (t < QR_NB)
#pragma unroll
        for (int q = 0; q < QR_NB; ++q) sT[t][q] = tr[q];
    __syncthreads();
    for (int e = t; e < nb * nb; e += 64) {
        const int i = e / nb, j = e % nb;
        T[e] = sT[i][j];
        Tt[e] = sT[j][i];
    }
}

// Q = partial identity (ones at (i, i) for i < dc), m x m, ld ldq.
__global__ __launch_bounds__(256)
void k_qr_partial_identity(int m, int dc, double *__restrict__ Q, int64_t ldq)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const int i = (int)(e / m), j = (int)(e - (int64_t)i * m);
    Q[(int64_t)i * ldq + j] = (i == j && i < dc) ? 1.0 : 0.0;
}

// R (rows x cols, ld ldr) from the packed factor: strict upper part of qr,
// rdiag on the diagonal, zero below; with Rt (cols x rows) its transpose.
__global__ __launch_bounds__(256)
void k_qr_make_r(int rows, int cols, const double *__restrict__ QR, int64_t ldq,
                 const double *__restrict__ rd, double *__restrict__ R, int64_t ldr,
                 double *__restrict__ Rt)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int i = (int)(e / cols), j = (int)(e - (int64_t)i * cols);
    const double v = i < j ? QR[(int64_t)i * ldq + j] : (i == j ? rd[i] : 0.0);
    R[(int64_t)i * ldr + j] = v;
    if (Rt) Rt[(int64_t)j * rows + i] = v;
}

// B (rows x cols, ld ldb) -> X (ld ldx)
__global__ __launch_bounds__(256)
void k_qr_copy(int rows, int cols, const double *__restrict__ B, int64_t ldb,
               double *__restrict__ X, int64_t ldx)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int i = (int)(e / cols), j = (int)(e - (int64_t)i * cols);
    X[(int64_t)i * ldx + j] = B[(int64_t)i * ldb + j];
}

namespace {

// Workspace of one QR call.
struct QrWs {
    double *V, *G, *T, *Tt, *W, *W2, *npart, *dpart, *scal, *P;
    uint64_t *slots;
};

int qr_ws(la_ctx *ctx, int64_t m, int64_t n, int64_t wcols, QrWs &w)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t dc = std::min(m, n);
    const size_t bv = al((size_t)m * QR_NB * 8), bg = al((size_t)QR_NB * QR_NB * 8),
                 bw = al((size_t)QR_NB * std::max<int64_t>(wcols, 1) * 8),
                 bp = al((size_t)QR_PARTS * 8), bd = al((size_t)QR_PARTS * QR_NB * 8),
                 bs = al((size_t)2 * std::max<int64_t>(dc, 1) * 8),
                 bpk = al(std::max((size_t)VT_MAXCHUNK * QR_NB * std::max<int64_t>(std::max<int64_t>(wcols, 1), QR_NB) * 8,
                                   (size_t)QRP_MAXP * QR_NB * QR_NB * 8));
    const size_t bsl = al((size_t)2 * QRP_MAXP * QRP_SW * 8);
    void *base = nullptr;
    int st = ws_get(ctx, WS_QR, bv + 3 * bg + 2 * bw + bp + bd + bs + bpk + bsl, &base);
    if (st) return st;
    char *p = (char *)base;
    w.V = (double *)p; p += bv;
    w.G = (double *)p; p += bg;
    w.T = (double *)p; p += bg;
    w.Tt = (double *)p; p += bg;
    w.W = (double *)p; p += bw;
    w.W2 = (double *)p; p += bw;
    w.npart = (double *)p; p += bp;
    w.dpart = (double *)p; p += bd;
    w.scal = (double *)p; p += bs;
    w.P = (double *)p; p += bpk;
    w.slots = (uint64_t *)p;
    return LA_OK;
}

int grid_rows(int64_t rows) { return (int)std::min<int64_t>(QR_PARTS, std::max<int64_t>(1, ceil_div(rows, 256))); }

// W (nb x cols, ld cols) = V^T C, split over at most VT_MAXCHUNK row chunks
int qr_vtc(la_ctx *ctx, int64_t rows, int nb, int64_t cols, const double *C, int64_t ldc,
           double *W, QrWs &w)
{
    const int64_t rch = std::max<int64_t>(VT_SUB, ceil_div(ceil_div(rows, VT_MAXCHUNK), VT_SUB) * VT_SUB);
    const int nch = (int)ceil_div(rows, rch);
    hipLaunchKernelGGL(k_qr_vtc_partial, dim3((unsigned)ceil_div(cols, VT_COLS), (unsigned)nch),
                       dim3(256), 0, ctx->stream, (int)rows, nb, (int)cols, (int)rch,
                       (const double *)w.V, C, ldc, w.P);
    hipLaunchKernelGGL(k_qr_vtc_reduce, dim3((unsigned)ceil_div(nb * cols, 256)), dim3(256), 0,
                       ctx->stream, nch, nb, (int)cols, (const double *)w.P, W);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// T of a panel from its G = V^T V (in w.G); diagu: the panel's diagonal block.
int qr_larft(la_ctx *ctx, int nb, const double *rd, const double *diagu, int64_t ldu, QrWs &w)
{
    hipLaunchKernelGGL(k_qr_larft, dim3(1), dim3(64), 0, ctx->stream, nb, (const double *)w.G,
                       rd, diagu, ldu, w.T, w.Tt);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// V and T of the panel at column k0 (width nb) of the packed factor.
int qr_block(la_ctx *ctx, int64_t m, const double *QR, int64_t ldq, const double *rdiag,
             int64_t k0, int nb, QrWs &w)
{
    const int64_t rows = m - k0;
    hipLaunchKernelGGL(k_qr_make_v, dim3((unsigned)ceil_div(rows * nb, 256)), dim3(256), 0,
                       ctx->stream, (int)rows, nb, QR + k0 * ldq + k0, ldq, rdiag + k0, w.V);
    LA_HIP_CHECK(ctx, hipGetLastError());
    int st = qr_vtc(ctx, rows, nb, nb, w.V, nb, w.G, w);
    if (st) return st;
    return qr_larft(ctx, nb, rdiag + k0, QR + k0 * ldq + k0, ldq, w);
}

// C (rows x cols, ld ldc, the rows k0.. of the operand) <- (I - V op(T) V^T) C
// with op(T) = T^T (trans) or T.
int qr_apply_block(la_ctx *ctx, int64_t rows, int nb, int64_t cols, double *C, int64_t ldc,
                   bool trans, QrWs &w)
{
    if (cols <= 0) return LA_OK;
    int st = qr_vtc(ctx, rows, nb, cols, C, ldc, w.W, w);
    if (st) return st;
    st = gemm_rm<double>(ctx, nb, cols, nb, 1.0, trans ? w.Tt : w.T, nb, w.W, cols, 0.0, w.W2,
                         cols, false, false);
    if (st) return st;
    return gemm_rm<double>(ctx, rows, cols, nb, -1.0, w.V, nb, w.W2, cols, 1.0, C, ldc, false,
                           false);
}

}  // namespace

// In-place factor of A (m x n, ld lda); rdiag (device, min(m, n)).
int geqrf_device(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda, double *rdiag)
{
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2) return set_error(ctx, LA_ERR_ARG, "qr: extent too large");
    const int64_t dc = std::min(m, n);
    QrWs w;
    int st = qr_ws(ctx, m, n, n, w);
    if (st) return st;
    // cooperative panels (LAHIP_QR_COLUMN_KERNELS=1 selects the per-column
    // launches everywhere, for testing); the exchange slots start with a tag
    // no column uses
    const char *ck = getenv("LAHIP_QR_COLUMN_KERNELS");
    const bool coop = !(ck && atoi(ck) != 0);
    const uint64_t ticks = spin_limit_ticks(ctx);
    if (coop && (st = fill_u64(ctx, w.slots, (int64_t)2 * QRP_MAXP * QRP_SW,
                               0xffffffff00000000ull)))
        return st;
    for (int64_t k0 = 0; k0 < dc; k0 += QR_NB) {
        const int nb = (int)std::min<int64_t>(QR_NB, dc - k0);
        const int64_t cend = k0 + nb;
        const int64_t rows = m - k0;
        const int rpt = rows <= (int64_t)256 * QRP_MAXP ? 1 : rows <= (int64_t)512 * QRP_MAXP ? 2 : 0;
        if (coop && rpt > 0) {
            const int Pq = (int)ceil_div(rows, (int64_t)256 * rpt);
            const int kneed = (Pq + 7) / 8;
            const int ki = kneed <= 1 ? 0 : kneed <= 2 ? 1 : kneed <= 4 ? 2 : kneed <= 8 ? 3 : 4;
            static const void *const fns[2][5] = {
                {(const void *)k_qr_panel<1, 1>, (const void *)k_qr_panel<1, 2>,
                 (const void *)k_qr_panel<1, 4>, (const void *)k_qr_panel<1, 8>,
                 (const void *)k_qr_panel<1, 16>},
                {(const void *)k_qr_panel<2, 1>, (const void *)k_qr_panel<2, 2>,
                 (const void *)k_qr_panel<2, 4>, (const void *)k_qr_panel<2, 8>,
                 (const void *)k_qr_panel<2, 16>}};
            const void *fn = fns[rpt - 1][ki];
            const size_t shmem = (size_t)256 * rpt * QRP_LS * sizeof(double);
            static size_t lds_max[2][5] = {};
            size_t &lm = lds_max[rpt - 1][ki];
            if (!lm) lm = allow_dynamic_lds(fn);
            if (shmem > lm) return set_error(ctx, LA_ERR_ARG, "qr panel: LDS");
            int mi = (int)m, k0i = (int)k0;
            int32_t *stp = ctx->d_scratch + SCR_ERR;
            uint64_t tk = ticks;
            int want = cend < n ? 1 : 0;
            void *args[] = {&mi, &A, &lda, &k0i, (void *)&nb, &rdiag, &w.slots, &stp, &tk,
                            &w.V, &w.P, &want};
            // an ordinary launch: at most 128 workgroups of one per CU are
            // co-resident on the full device, every wait is bounded, and the
            // cooperative-launch API would tie the device to this process
            // (another process's cooperative launch then fails)
            LA_HIP_CHECK(ctx, hipLaunchKernel(fn, dim3(Pq), dim3(256), args, shmem, ctx->stream));
            if (want) {
                // the block's V^T V from the panel's per-workgroup partials, then T
                hipLaunchKernelGGL(k_qr_vtc_reduce, dim3((unsigned)ceil_div(nb * nb, 256)),
                                   dim3(256), 0, ctx->stream, Pq, nb, nb, (const double *)w.P,
                                   w.G);
                LA_HIP_CHECK(ctx, hipGetLastError());
                if ((st = qr_larft(ctx, nb, rdiag + k0, A + k0 * lda + k0, lda, w))) return st;
                if ((st = qr_apply_block(ctx, m - k0, nb, n - cend, A + k0 * lda + cend, lda,
                                         true, w)))
                    return st;
            }
            continue;
        } else {
        // column by column inside the panel; the first column's norm first
        int P = grid_rows(m - k0);
        hipLaunchKernelGGL(k_qr_norm_partial, dim3(P), dim3(256), 0, ctx->stream, (int)m,
                           (const double *)A, lda, (int)k0, w.npart);
        LA_HIP_CHECK(ctx, hipGetLastError());
        for (int64_t c = k0; c < cend; ++c) {
            // the grid of both launches covers rows [c, m); its size also
            // fixes the number of partials the next step reduces
            const int Pc = grid_rows(m - c);
            hipLaunchKernelGGL(k_qr_dots, dim3(Pc), dim3(256), 0, ctx->stream, (int)m,
                               (const double *)A, lda, (int)c, (int)cend, (const double *)w.npart,
                               P, w.scal, w.dpart);
            hipLaunchKernelGGL(k_qr_apply, dim3(Pc), dim3(256), 0, ctx->stream, (int)m, A, lda,
                               (int)c, (int)cend, (const double *)w.dpart, Pc,
                               (const double *)w.scal, rdiag, w.npart);
            LA_HIP_CHECK(ctx, hipGetLastError());
            P = Pc;
        }
        }
        // the block's reflections on the trailing columns
        if (cend < n) {
            if ((st = qr_block(ctx, m, A, lda, rdiag, k0, nb, w))) return st;
            if ((st = qr_apply_block(ctx, m - k0, nb, n - cend, A + k0 * lda + cend, lda, true, w)))
                return st;
        }
    }
    return LA_OK;
}

// Q (m x m, ld ldq) of the packed factor QR (m x n, ld ldqr).
int qr_q_device(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                const double *rdiag, double *Q, int64_t ldq)
{
    const int64_t dc = std::min(m, n);
    QrWs w;
    int st = qr_ws(ctx, m, n, m, w);
    if (st) return st;
    hipLaunchKernelGGL(k_qr_partial_identity, dim3((unsigned)ceil_div(m * m, 256)), dim3(256), 0,
                       ctx->stream, (int)m, (int)dc, Q, ldq);
    LA_HIP_CHECK(ctx, hipGetLastError());
    const int64_t last = ((dc - 1) / QR_NB) * QR_NB;
    for (int64_t k0 = last; k0 >= 0; k0 -= QR_NB) {
        const int nb = (int)std::min<int64_t>(QR_NB, dc - k0);
        if ((st = qr_block(ctx, m, QR, ldqr, rdiag, k0, nb, w))) return st;
        if ((st = qr_apply_block(ctx, m - k0, nb, m, Q + k0 * ldq, ldq, false, w))) return st;
    }
    return LA_OK;
}

// X (n x nx) least-squares solution; Y (m x nx, ld ldy) holds B on entry and
// Q^T B on exit.  Full rank and m >= n checked by the caller.
int qr_solve_device(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                    const double *rdiag, int64_t nx, double *Y, int64_t ldy, double *R,
                    double *X, int64_t ldx)
{
    QrWs w;
    int st = qr_ws(ctx, m, n, nx, w);
    if (st) return st;
    for (int64_t k0 = 0; k0 < n; k0 += QR_NB) {
        const int nb = (int)std::min<int64_t>(QR_NB, n - k0);
        if ((st = qr_block(ctx, m, QR, ldqr, rdiag, k0, nb, w))) return st;
        if ((st = qr_apply_block(ctx, m - k0, nb, nx, Y + k0 * ldy, ldy, true, w))) return st;
    }
    // R x = y[0..n): R with rdiag on its diagonal, solved as an upper factor
    hipLaunchKernelGGL(k_qr_make_r, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0,
                       ctx->stream, (int)n, (int)n, QR, ldqr, rdiag, R, n, (double *)nullptr);
    hipLaunchKernelGGL(k_qr_copy, dim3((unsigned)ceil_div(n * nx, 256)), dim3(256), 0,
                       ctx->stream, (int)n, (int)nx, (const double *)Y, ldy, X, ldx);
    LA_HIP_CHECK(ctx, hipGetLastError());
    if (nx >= TRSM_MIN_RHS && n >= TRSM_MIN_N) return trsm_blocked(ctx, 'U', n, R, n, nx, X, ldx);
    return trsv_blocked<double>(ctx, 'U', false, false, n, R, n, nx, X, ldx);
}

}  // namespace lahip

using namespace lahip;

namespace {
struct Bufs {
    char *base;
    size_t off = 0;
    template <typename T> T *take(int64_t count) {
        T *p = reinterpret_cast<T *>(base + off);
        off += (((size_t)std::max<int64_t>(count, 1) * sizeof(T)) + 255) & ~(size_t)255;
        return p;
    }
};
size_t rbytes(int64_t count, size_t elem)
{
    return (((size_t)std::max<int64_t>(count, 1) * elem) + 255) & ~(size_t)255;
}
// rdiag -> full-rank flag as QRDecomposition::is_full_rank (qr.rs:108-115)
// sees it: LA_ERR_SHAPE when n > m (it reads rdiag past its end, a panic),
// LA_SINGULAR when some rdiag[j] == 0, else LA_OK.
int rank_status(la_ctx *ctx, int64_t m, int64_t n, const double *d_rdiag)
{
    if (n > m) return set_error(ctx, LA_ERR_SHAPE, "qr is_full_rank: index out of bounds (cols > rows)");
    std::vector<double> h((size_t)n);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), d_rdiag, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t j = 0; j < n; ++j)
        if (h[(size_t)j] == 0.0) return LA_SINGULAR;
    return LA_OK;
}
}  // namespace

extern "C" int la_dgeqrf_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda,
                             double *d_rdiag)
{
    if (!ctx || !A || !d_rdiag) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr: empty matrix");
    if (lda < n) return LA_ERR_ARG;
    return geqrf_device(ctx, m, n, A, lda, d_rdiag);
}

extern "C" int la_dgeqrf_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, double *QR,
                            double *rdiag)
{
    if (!ctx) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr: empty matrix");  // mod.rs:209
    if (!A || !QR || !rdiag) return LA_ERR_ARG;
    const int64_t dc = std::min(m, n);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST, rbytes(m * n, 8) + rbytes(dc, 8), &base);
    if (st) return st;
    Bufs b{(char *)base};
    double *dA = b.take<double>(m * n), *dr = b.take<double>(dc);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = geqrf_device(ctx, m, n, dA, n, dr))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(QR, dA, m * n * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(rdiag, dr, dc * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

extern "C" int la_dqr_q_dev(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                            const double *d_rdiag, double *Q, int64_t ldq)
{
    if (!ctx || !QR || !d_rdiag || !Q) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr: empty matrix");
    if (ldqr < n || ldq < m) return LA_ERR_ARG;
    return qr_q_device(ctx, m, n, QR, ldqr, d_rdiag, Q, ldq);
}

extern "C" int la_dqr_q_rm(la_ctx *ctx, int64_t m, int64_t n, const double *QR,
                           const double *rdiag, double *Q)
{
    if (!ctx) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr: empty matrix");
    if (!QR || !rdiag || !Q) return LA_ERR_ARG;
    const int64_t dc = std::min(m, n);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST, rbytes(m * n, 8) + rbytes(dc, 8) + rbytes(m * m, 8), &base);
    if (st) return st;
    Bufs b{(char *)base};
    double *dqr = b.take<double>(m * n), *dr = b.take<double>(dc), *dq = b.take<double>(m * m);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dqr, QR, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dr, rdiag, dc * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = qr_q_device(ctx, m, n, dqr, n, dr, dq, m))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(Q, dq, m * m * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

extern "C" int la_dqrs_dev(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                           const double *d_rdiag, int64_t nx, const double *B, int64_t ldb,
                           double *X, int64_t ldx)
{
    if (!ctx || !QR || !d_rdiag || !B || !X) return LA_ERR_ARG;
    if (m <= 0 || n <= 0 || nx <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr solve: empty");
    if (ldqr < n || ldb < nx || ldx < nx) return LA_ERR_ARG;
    int st = rank_status(ctx, m, n, d_rdiag);
    if (st) return st;
    void *base = nullptr;
    if ((st = ws_get(ctx, WS_HOST, rbytes(m * nx, 8) + rbytes(n * n, 8), &base))) return st;
    Bufs b{(char *)base};
    double *Y = b.take<double>(m * nx), *R = b.take<double>(n * n);
    hipLaunchKernelGGL(k_qr_copy, dim3((unsigned)ceil_div(m * nx, 256)), dim3(256), 0, ctx->stream,
                       (int)m, (int)nx, B, ldb, Y, nx);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return qr_solve_device(ctx, m, n, QR, ldqr, d_rdiag, nx, Y, nx, R, X, ldx);
}

extern "C" int la_dqrs_rm(la_ctx *ctx, int64_t m, int64_t n, const double *QR,
                          const double *rdiag, int64_t b_rows, int64_t nx, const double *B,
                          double *X)
{
    if (!ctx) return LA_ERR_ARG;
    if (b_rows != m) return set_error(ctx, LA_ERR_SHAPE, "qr solve: b.rows() != rows");  // qr.rs:200
    if (m <= 0 || n <= 0 || nx <= 0) return set_error(ctx, LA_ERR_SHAPE, "qr solve: empty");
    if (!QR || !rdiag || !B || !X) return LA_ERR_ARG;
    if (n > m) return set_error(ctx, LA_ERR_SHAPE, "qr is_full_rank: index out of bounds (cols > rows)");
    for (int64_t j = 0; j < n; ++j)
        if (rdiag[j] == 0.0) return LA_SINGULAR;                   // qr.rs:201-203: None
    const int64_t dc = n;
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST,
                    rbytes(m * n, 8) + rbytes(dc, 8) + rbytes(m * nx, 8) + rbytes(n * n, 8) +
                        rbytes(n * nx, 8),
                    &base);
    if (st) return st;
    Bufs b{(char *)base};
    double *dqr = b.take<double>(m * n), *dr = b.take<double>(dc), *Y = b.take<double>(m * nx);
    double *R = b.take<double>(n * n), *dX = b.take<double>(n * nx);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dqr, QR, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dr, rdiag, dc * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(Y, B, m * nx * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((st = qr_solve_device(ctx, m, n, dqr, n, dr, nx, Y, nx, R, dX, nx))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(X, dX, n * nx * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}

extern "C" int la_dpinverse_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, double *P)
{
    if (!ctx) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "pinverse: empty matrix");
    if (!A || !P) return LA_ERR_ARG;
    const int64_t dc = std::min(m, n);
    void *base = nullptr;
    int st = ws_get(ctx, WS_HOST,
                    2 * rbytes(m * n, 8) + rbytes(dc, 8) + 2 * rbytes(m * n, 8) +
                        2 * rbytes(n * n, 8) + rbytes(n, 8) + rbytes(4, 4) + rbytes(n * m, 8),
                    &base);
    if (st) return st;
    Bufs b{(char *)base};
    double *dA = b.take<double>(m * n), *dqr = b.take<double>(m * n), *dr = b.take<double>(dc);
    double *R = b.take<double>(m * n), *Rt = b.take<double>(m * n);
    double *G = b.take<double>(n * n), *Gi = b.take<double>(n * n);
    uint64_t *piv = b.take<uint64_t>(n);
    int32_t *info = b.take<int32_t>(4);
    double *dP = b.take<double>(n * m);
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dA, A, m * n * 8, hipMemcpyHostToDevice, ctx->stream));
    LA_HIP_CHECK(ctx, hipMemcpyAsync(dqr, dA, m * n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if ((st = geqrf_device(ctx, m, n, dqr, n, dr))) return st;
    // R = get_r() (m x n), G = R^T R (n x n)
    hipLaunchKernelGGL(k_qr_make_r, dim3((unsigned)ceil_div(m * n, 256)), dim3(256), 0,
                       ctx->stream, (int)m, (int)n, (const double *)dqr, n, (const double *)dr, R,
                       n, Rt);
    LA_HIP_CHECK(ctx, hipGetLastError());
    if ((st = gemm_rm<double>(ctx, n, n, m, 1.0, Rt, m, R, n, 0.0, G, n, false, false))) return st;
    // (R^T R)^-1 by LU and two triangular solves; .unwrap() of None panics
    if ((st = la_dinverse_dev(ctx, n, G, n, piv, Gi, n, info))) return st;
    int32_t hinfo[3];
    LA_HIP_CHECK(ctx, hipMemcpyAsync(hinfo, info, 12, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (hinfo[1] != LA_OK) return set_error(ctx, hinfo[1], "getrf: device wait failed");
    if (hinfo[2] == LA_SINGULAR)
        return set_error(ctx, LA_SINGULAR, "pinverse: R^T R is singular (inverse().unwrap() on None)");
    // P = (R^T R)^-1 A^T
    if ((st = gemm_rm<double>(ctx, n, m, n, 1.0, Gi, n, dA, n, 0.0, dP, m, true, false))) return st;
    LA_HIP_CHECK(ctx, hipMemcpyAsync(P, dP, n * m * 8, hipMemcpyDeviceToHost, ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_sticky(ctx);
}
