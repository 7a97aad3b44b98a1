// Batched uneven Cholesky factor + solve, FP64, one persistent launch.
//
// Each problem i is CholeskyDecomposition::new(A_i) followed by .solve(b_i)
// (src/decomp/cholesky.rs:56-144): exact-symmetry check, d <= 0 rejection
// (NaN is not rejected), full L with the strict upper triangle zeroed,
// forward then backward substitution.  Problems (order <= 256) are pulled
// from an atomic queue in descending size order by a grid of persistent
// workgroups; the last workgroup to finish resets the queue, so a captured
// graph can be replayed without any host-side reset.
//
// Per problem, one 256-thread workgroup, left-looking in panels of NB
// columns (NB = 8 or 16, chosen per size bucket by the autotuner):
//   1. P = A[J0:n, J0:J0+NB] - L[J0:n, 0:J0] L[J0:J0+NB, 0:J0]^T on
//      v_mfma_f64_16x16x4 straight from the L written so far, A read in
//      the accumulator layout, every strictly-lower element compared with
//      its mirror; P goes to LDS.  Wave 0 updates only the first row tile
//      (the diagonal block's rows) and goes straight on to step 2 while
//      waves 1-3 update the other row tiles, so the serial diagonal block
//      is hidden behind the rest of the panel update;
//   2. wave 0 factors the NB x NB diagonal block: lane r holds row r, the
//      column being eliminated is always register 0 (rows rotate left one
//      step per column) and column values are broadcast with readlane, so
//      the block needs no barrier; then it solves this block of y;
//   3. every row below solves against the diagonal block in registers
//      (row per thread), the panel and the zeros right of the diagonal block
//      go to L, the rows below update their part of b from registers;
//   backward substitution block by block, last block first: one wave solves
//   the transposed diagonal block with readlane broadcasts, then the block's
//   rows of L are pushed into the part of y above it in one coalesced sweep.
// L is both output and left-looking operand; it is written and re-read by
// the same workgroup with a workgroup barrier in between.
#include "la_common.hpp"
#include <cmath>
#include <type_traits>

namespace lahip {

constexpr int CB_MAXN = 256;
constexpr int CB_NT = 256;
constexpr int CB_NBMAX = 16;
constexpr int CB_NBUCKET = 4;   // size buckets: [1,32], (32,64], (64,128], (128,256]

__host__ __device__ inline int chol_bucket(int n) {
    return n <= 32 ? 0 : (n <= 64 ? 1 : (n <= 128 ? 2 : 3));
}

struct CbLds {
    double P[CB_MAXN][CB_NBMAX + 1];   // current panel, rows J0..n-1
    double y[CB_MAXN];                 // b -> y -> x
    double dinv[CB_NBMAX];             // reciprocals of the diagonal block's pivots
    int bad, prob;
};

__device__ inline double cb_bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// NB: panel width; RTW: row tiles whose MFMA K loops run together (more
// loads in flight per round trip against more registers)
template <int NB, int CB_RTW>
__device__ __forceinline__ void chol_solve_one(int n, const double *__restrict__ Ap, double *Lp,
                               const double *__restrict__ bp, double *xp, int32_t *status,
                               CbLds &S)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lhi = lane >> 4;
    if (tid == 0) S.bad = 0;
    for (int i = tid; i < n; i += CB_NT) S.y[i] = bp[i];
    __syncthreads();

    for (int J0 = 0; J0 < n; J0 += NB) {
        const int jb = min(NB, n - J0), m = n - J0;
        // ---- 1. panel update on MFMA, symmetry check, P -> LDS, and (wave 0)
        //         the diagonal block right after its own row tile
        // Wave 0 updates row tile 0 (the diagonal block's rows) alone and goes
        // on to factor the diagonal block while waves 1-3 update row tiles
        // w, w+3, w+6, ... (m <= 256), running the K loops of RTW of them
        // together (4 + 4*RTW independent loads per lane per 16-wide chunk).
        const int rt_n = (m + 15) >> 4;
        int asym = 0;
        const bool okB = l16 < jb;
        const double *pb = Lp + (int64_t)(J0 + min(l16, jb - 1)) * n;
        auto update_tiles = [&](auto rtw_c, int base, int step) {
            constexpr int RTW = decltype(rtw_c)::value;
            for (int pass = 0;; ++pass) {
                if (base + step * RTW * pass >= rt_n) break;
                d4_t acc[RTW];
#pragma unroll
                for (int i = 0; i < RTW; ++i) acc[i] = (d4_t){0.0, 0.0, 0.0, 0.0};
                if (J0 > 0) {
                    const double *pa[RTW];
                    bool okA[RTW];
#pragma unroll
                    for (int i = 0; i < RTW; ++i) {
                        const int rowA = J0 + 16 * (base + step * (RTW * pass + i)) + l16;
                        okA[i] = rowA < n;
                        pa[i] = Lp + (int64_t)min(rowA, n - 1) * n;
                    }
                    // k = k0 + 4*(lane/16) + s for MFMA step s, the same for A
                    // and B; J0 is a multiple of 8, so a K chunk of 16 may end
                    // inside the current panel (or past J0): those columns are
                    // masked out.  KU chunks are loaded together before their
                    // MFMAs: every load is unconditional from a clamped column
                    // (< J0) of a clamped row, and out-of-range values are
                    // zeroed afterwards.
                    constexpr int KU = 4 / RTW;
                    for (int k0 = 0; k0 < J0; k0 += 16 * KU) {
                        double a[KU][RTW][4], b[KU][4];
#pragma unroll
                        for (int u = 0; u < KU; ++u) {
                            // J0 is a multiple of 4: kc..kc+3 all below J0 or none
                            const int kc = min(k0 + 16 * u + 4 * lhi, J0 - 4);
#pragma unroll
                            for (int s = 0; s < 4; ++s) b[u][s] = pb[kc + s];
#pragma unroll
                            for (int i = 0; i < RTW; ++i)
#pragma unroll
                                for (int s = 0; s < 4; ++s) a[u][i][s] = pa[i][kc + s];
                        }
#pragma unroll
                        for (int u = 0; u < KU; ++u) {
                            const bool kin = k0 + 16 * u + 4 * lhi < J0;
#pragma unroll
                            for (int s = 0; s < 4; ++s)
                                if (!(okB && kin)) b[u][s] = 0.0;
#pragma unroll
                            for (int i = 0; i < RTW; ++i)
#pragma unroll
                                for (int s = 0; s < 4; ++s)
                                    if (!(okA[i] && kin)) a[u][i][s] = 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < KU; ++u)
#pragma unroll
                            for (int i = 0; i < RTW; ++i)
#pragma unroll
                                for (int s = 0; s < 4; ++s)
                                    acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                        a[u][i][s], b[u][s], acc[i], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int i = 0; i < RTW; ++i) {
                    const int rt = base + step * (RTW * pass + i);
                    if (rt < rt_n) {
                        // the element and its mirror, loaded unconditionally
                        // from clamped (in-bounds) positions so that the eight
                        // loads issue together; out-of-range lanes discard them
                        double v[4], w[4];
                        const int gcc = J0 + min(l16, jb - 1);
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int grc = J0 + min(16 * rt + lhi + 4 * r, m - 1);
                            v[r] = Ap[(int64_t)grc * n + gcc];
                            w[r] = Ap[(int64_t)gcc * n + grc];
                        }
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int pr = 16 * rt + lhi + 4 * r;      // f64 MFMA row layout
                            if (pr < m && l16 < jb) {
                                const int gr = J0 + pr, gc = J0 + l16;
                                if (gr > gc && !(v[r] == w[r])) asym = 1;   // NaN too
                                S.P[pr][l16] = v[r] - acc[i][r];
                            }
                        }
                    }
                }
            }
        };
        if (wave == 0) {
            update_tiles(std::integral_constant<int, 1>{}, 0, 1 << 20);   // tile 0 only
            // the diagonal block's rows were written by this wave's own lanes:
            // keep the compiler from moving the reads above the writes (LDS
            // executes one wave's operations in order)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // ---- 2. diagonal block, rows in registers (rotated)
            double a[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) a[c] = (lane < jb && c < jb) ? S.P[lane][c] : 0.0;
#pragma unroll 1
            for (int t = 0; t < NB; ++t) {
                if (t < jb) {
                    const double dd = cb_bcast(a[0], t);      // A[t][t] after the updates
                    if (lane == 0 && dd <= 0.0) atomicOr(&S.bad, 2);
                    const double d = sqrt(dd);
                    const double rd = 1.0 / d;
                    if (lane == 0) S.dinv[t] = rd;
                    const double l = a[0] * rd;
                    if (lane == t) a[0] = d;
                    else if (lane > t) a[0] = l;
                    // columns t+1..: lane r >= t+j subtracts L[r][t] L[t+j][t];
                    // readlane indices stay below 32, masked lanes add zero
#pragma unroll
                    for (int j = 1; j < NB; ++j) {
                        const double ltj = cb_bcast(l, t + j);
                        const double lm = (lane >= t + j && t + j < jb) ? l : 0.0;
                        a[j] = fma(-lm, ltj, a[j]);
                    }
                }
                const double a0 = a[0];
#pragma unroll
                for (int c = 0; c < NB - 1; ++c) a[c] = a[c + 1];
                a[NB - 1] = a0;
            }
            if (lane < jb) {
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c <= lane) S.P[lane][c] = a[c];
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // this block of y: forward substitution with the diagonal block
            double yv = lane < jb ? S.y[J0 + lane] : 0.0;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < jb) {
                    if (lane == c) yv = yv * S.dinv[c];
                    const double yc = cb_bcast(yv, c);
                    if (lane > c && lane < jb) yv -= S.P[lane][c] * yc;
                }
            }
            if (lane < jb) S.y[J0 + lane] = yv;
        } else {
            update_tiles(std::integral_constant<int, CB_RTW>{}, wave, 3);
        }
        if (asym) atomicOr(&S.bad, 1);
        __syncthreads();
        // ---- 3. rows below: solve against the diagonal block; write L
        const int r = jb + tid;
        double x[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) x[c] = 0.0;
        if (r < m) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < jb) {
                    double s = S.P[r][c];
#pragma unroll
                    for (int k = 0; k < c; ++k) s -= x[k] * S.P[c][k];
                    x[c] = s * S.dinv[c];
                }
            }
            double *lr = Lp + (int64_t)(J0 + r) * n + J0;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < jb) lr[c] = x[c];
        }
        // diagonal block rows, with the zeros of the upper triangle
        for (int e = tid; e < jb * m; e += CB_NT) {
            const int rr = e / m, c = e - rr * m;
            Lp[(int64_t)(J0 + rr) * n + J0 + c] = (c < jb && c <= rr) ? S.P[rr][c] : 0.0;
        }
        __syncthreads();
        if (r < m) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < jb) s = fma(x[c], S.y[J0 + c], s);
            S.y[J0 + r] -= s;
        }
        __syncthreads();
    }

    // ---- backward substitution L^T x = y, last block first, right-looking:
    // once block B of x is known, its rows of L are pushed into y[0, J0) at
    // once.  The loads of the push (rows J0..J0+jb of L, one column per
    // thread) and those of the diagonal block (wave 0) do not depend on x_B
    // and are issued together at the start of the block, so a block costs
    // one memory round trip.
    const int nblk = (n + NB - 1) / NB;
    for (int B = nblk - 1; B >= 0; --B) {
        const int J0 = B * NB, jb = min(NB, n - J0);
        double lp[NB];
        const int c = tid;                 // J0 <= CB_MAXN - NB < CB_NT
        const bool push = c < J0;
#pragma unroll
        for (int rr = 0; rr < NB; ++rr)
            lp[rr] = Lp[(int64_t)(J0 + min(rr, jb - 1)) * n + min(c, J0 > 0 ? J0 - 1 : 0)];
        if (wave == 0) {
            double colv[NB];          // lane c: column c of the diagonal block
#pragma unroll
            for (int rr = 0; rr < NB; ++rr)
                colv[rr] = Lp[(int64_t)(J0 + min(rr, jb - 1)) * n + J0 + min(lane, jb - 1)];
            const double dgi = 1.0 / Lp[(int64_t)(J0 + min(lane, jb - 1)) * n + J0 + min(lane, jb - 1)];
            double xv = lane < jb ? S.y[J0 + lane] : 0.0;
#pragma unroll
            for (int rr = NB - 1; rr >= 0; --rr) {
                if (rr < jb) {
                    if (lane == rr) xv = xv * dgi;
                    const double xr = cb_bcast(xv, rr);
                    if (lane < rr) xv -= colv[rr] * xr;
                }
            }
            if (lane < jb) S.y[J0 + lane] = xv;
        }
        __syncthreads();
        if (push) {
            double s = 0.0;
#pragma unroll
            for (int rr = 0; rr < NB; ++rr)
                if (rr < jb) s = fma(lp[rr], S.y[J0 + rr], s);
            S.y[c] -= s;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += CB_NT) xp[i] = S.y[i];
    if (tid == 0) {
        const int b = S.bad;
        *status = (b & 1) ? LA_NOT_SYMMETRIC : ((b & 2) ? LA_NOT_SPD : LA_OK);
    }
    __syncthreads();
}

__global__ __launch_bounds__(CB_NT, 4)
void k_chol_batched(int count, const int32_t *__restrict__ order, const int32_t *__restrict__ nvec,
                    const int64_t *__restrict__ offA, const int64_t *__restrict__ offB,
                    const double *A, double *L, const double *B, double *X, int32_t *status,
                    uint32_t *queue, int4 cfg_by_bucket)
{
    __shared__ CbLds S;
    for (;;) {
        if (threadIdx.x == 0) S.prob = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int pi = S.prob;
        __syncthreads();
        if (pi >= count) break;
        const int prob = order[pi];
        const int n = nvec[prob];
        if (n <= 0 || n > CB_MAXN) {
            if (threadIdx.x == 0) status[prob] = LA_ERR_SHAPE;
            continue;
        }
        // tile configuration of this size bucket: panel width | row tiles << 8
        const int bk = chol_bucket(n);
        const int cfg = bk == 0 ? cfg_by_bucket.x : bk == 1 ? cfg_by_bucket.y
                      : bk == 2 ? cfg_by_bucket.z : cfg_by_bucket.w;
        const double *Ap = A + offA[prob];
        double *Lp = L + offA[prob];
        const double *bp = B + offB[prob];
        double *xp = X + offB[prob];
        switch (cfg) {
        case 8 | (1 << 8): chol_solve_one<8, 1>(n, Ap, Lp, bp, xp, status + prob, S); break;
        case 8 | (2 << 8): chol_solve_one<8, 2>(n, Ap, Lp, bp, xp, status + prob, S); break;
        case 16 | (1 << 8): chol_solve_one<16, 1>(n, Ap, Lp, bp, xp, status + prob, S); break;
        default: chol_solve_one<16, 2>(n, Ap, Lp, bp, xp, status + prob, S); break;
        }
    }
    // the last workgroup out resets the queue for the next launch or replay
    if (threadIdx.x == 0) {
        const unsigned d = atomicAdd(&queue[1], 1u);
        if (d == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// Diagonally dominant symmetric test matrices for the autotuner:
// A[i][j] = 1/(1+|i-j|) + n * delta_ij.
__global__ void k_fill_spd_batch(int count, const int32_t *nvec, const int64_t *offA, double *A)
{
    const int prob = blockIdx.x;
    if (prob >= count) return;
    const int n = nvec[prob];
    double *Ap = A + offA[prob];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e / n, c = e - r * n;
        Ap[e] = 1.0 / (1.0 + (double)abs(r - c)) + (r == c ? (double)n : 0.0);
    }
}

int fill_spd_batch(la_ctx *ctx, int count, const int32_t *nvec, const int64_t *offA, double *A)
{
    hipLaunchKernelGGL(k_fill_spd_batch, dim3(count), dim3(256), 0, ctx->stream, count, nvec,
                       offA, A);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

int chol_batched_grid(la_ctx *ctx)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_chol_batched, CB_NT,
                                                     0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    (void)hipGetLastError();
    return std::max(1, ctx->num_cus) * per_cu;
}

int launch_chol_batched(la_ctx *ctx, hipStream_t s, int grid, int count, const int32_t *order,
                        const int32_t *nvec, const int64_t *offA, const int64_t *offB,
                        const double *A, double *L, const double *B, double *X, int32_t *status,
                        uint32_t *queue, const int cfg[4])
{
    const int4 nbv = make_int4(cfg[0], cfg[1], cfg[2], cfg[3]);
    hipLaunchKernelGGL(k_chol_batched, dim3(grid), dim3(CB_NT), 0, s, count, order, nvec, offA,
                       offB, A, L, B, X, status, queue, nbv);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

}  // namespace lahip
