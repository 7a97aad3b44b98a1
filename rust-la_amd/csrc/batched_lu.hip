// Batched uneven LU with partial pivoting and FP64 iterative refinement in
// one persistent launch.
//
// Problem i: A_i (n_i x n_i, FP64, row-major, n_i <= 1024) and B_i
// (n_i x nrhs_i, nrhs_i <= 8).  One 256-thread workgroup owns a problem from
// start to finish, so no problem ever waits on another workgroup:
//   1. W_i = A_i in the working precision, ||A_i||_inf and max|a| on the way;
//   2. blocked right-looking LU of W_i with partial pivoting: the NB-column
//      panel lives in registers (4 rows per thread, one barrier per column,
//      candidates exchanged through LDS), its row interchanges are applied to
//      the other columns as one permutation of at most 2*NB rows, then
//      U12 = L11^-1 A12 and the trailing update A22 -= L21 U12 on MFMA;
//   3. refinement: x = 0, then repeat { r = b - A x in FP64 from A_i;
//      eta = max_c ||r_c|| / (||A|| ||x_c|| + ||b_c||); stop when eta <= tol;
//      solve L U d = P r in working precision (vector in LDS); x += d }.
// Working precisions (template MODE):
//   FP64  W double, NB 16, v_mfma_f64_16x16x4; one solve, eta reported
//   FP32  W float,  NB 32, v_mfma_f32_16x16x4f32
//   FP16  W float,  NB 32, trailing-update operands rounded to FP16: L21 as
//         is (|l| <= 1), U12 after a per-matrix power-of-two scale that puts
//         max|a| in [64, 128) (undone exactly in the epilogue);
//         v_mfma_f32_16x16x32_f16 with FP32 accumulation.
// Each problem is LUDecomposition::new + solve (src/decomp/lu.rs:104-278)
// in structure; pivots follow the reference's rule (largest |v|, lowest row
// on ties) in the working precision.
#include "la_common.hpp"
#include <cmath>

namespace lahip {

constexpr int BL_MAXN = 1024;
constexpr int BL_NT = 256;
constexpr int BL_SLOTS = BL_MAXN / BL_NT;   // panel rows per thread
constexpr int BL_CHUNK = BL_NT;             // columns per TRSM + update chunk
constexpr int BL_MVC = 64;                  // columns per row-move chunk
constexpr int BL_MAXRHS = 8;
constexpr int BL_Y_BYTES = 32768;           // right-hand-side block in LDS
constexpr int BL_STG_BYTES = 36864;         // U12 chunk / row-move staging
constexpr int BL_BIG = 1 << 30;             // position of an absent row

template <int MODE> struct BlCfg;
template <> struct BlCfg<LA_PREC_FP64> {
    typedef double T; typedef double E;
    static constexpr int NB = 16, PAD = 1;
};
template <> struct BlCfg<LA_PREC_FP32> {
    typedef float T; typedef float E;
    static constexpr int NB = 32, PAD = 4;
};
template <> struct BlCfg<LA_PREC_FP16> {
    typedef float T; typedef _Float16 E;
    static constexpr int NB = 32, PAD = 8;
};

// pivot key: |v| bits + 1, NaN -> 0 (never preferred over a number)
__device__ inline uint64_t bl_key(float v) {
    const uint32_t b = __float_as_uint(v) & 0x7fffffffu;
    return b > 0x7f800000u ? 0ull : (uint64_t)b + 1ull;
}
__device__ inline uint64_t bl_key(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v) & 0x7fffffffffffffffull;
    return b > 0x7ff0000000000000ull ? 0ull : b + 1ull;
}
__device__ inline uint64_t bl_shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
// (key, q) lexicographic with q = ~position: larger |v| first, lower row on ties
__device__ inline bool bl_better(uint64_t k1, uint32_t q1, uint64_t k0, uint32_t q0) {
    return k1 > k0 || (k1 == k0 && q1 > q0);
}
__device__ inline float bl_bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline double bl_bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double bl_wave_max(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
    return v;
}
__device__ inline double bl_wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

template <int MODE> struct BlMfma;
template <> struct BlMfma<LA_PREC_FP64> {
    __device__ static inline int drow(int lhi, int r) { return lhi + 4 * r; }
};
template <> struct BlMfma<LA_PREC_FP32> {
    __device__ static inline int drow(int lhi, int r) { return 4 * lhi + r; }
};
template <> struct BlMfma<LA_PREC_FP16> {
    __device__ static inline int drow(int lhi, int r) { return 4 * lhi + r; }
};

// LDS of one workgroup (about 77 KB: two workgroups fit a CU)
template <int MODE> struct BlLds {
    typedef typename BlCfg<MODE>::T T;
    static constexpr int NB = BlCfg<MODE>::NB;
    __align__(16) unsigned char y[BL_Y_BYTES];       // RHS block, T[n][g]
    __align__(16) unsigned char stg[BL_STG_BYTES];   // E[BL_CHUNK][NB+PAD] or T[2NB][BL_MVC]
    T l11[NB][NB + 1];
    T cand_row[2][4][NB];
    uint64_t cand_key[2][4];
    uint32_t cand_q[2][4];
    int mv_src[2 * NB], mv_dst[2 * NB], mv_tmp[2 * NB];
    int nmv, prob, singular;
    double red[4 * 2 * BL_MAXRHS];
};

// ---------------------------------------------------------------- convert
template <int MODE>
__device__ void bl_convert(int n, const double *__restrict__ Ap,
                           typename BlCfg<MODE>::T *__restrict__ Wp, int *__restrict__ perm,
                           BlLds<MODE> &L, double &anorm, double &amax)
{
    typedef typename BlCfg<MODE>::T T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double rs_max = 0.0, vmax = 0.0;
    for (int r = wave; r < n; r += 4) {
        double s = 0.0;
        const double *ar = Ap + (int64_t)r * n;
        T *wr = Wp + (int64_t)r * n;
#pragma unroll 4
        for (int c = lane; c < n; c += 64) {
            const double v = ar[c];
            wr[c] = (T)v;
            const double av = fabs(v);
            s += av;
            vmax = fmax(vmax, av);
        }
        rs_max = fmax(rs_max, bl_wave_sum(s));
    }
    for (int i = tid; i < n; i += BL_NT) perm[i] = i;
    vmax = bl_wave_max(vmax);
    if (lane == 0) { L.red[wave] = rs_max; L.red[4 + wave] = vmax; }
    __syncthreads();
    anorm = fmax(fmax(L.red[0], L.red[1]), fmax(L.red[2], L.red[3]));
    amax = fmax(fmax(L.red[4], L.red[5]), fmax(L.red[6], L.red[7]));
    __syncthreads();
}

// ------------------------------------------------------------- factorise
template <int MODE>
__device__ void bl_getrf(int n, typename BlCfg<MODE>::T *__restrict__ W, int *__restrict__ perm,
                         float uscale, BlLds<MODE> &L)
{
    typedef typename BlCfg<MODE>::T T;
    typedef typename BlCfg<MODE>::E E;
    constexpr int NB = BlCfg<MODE>::NB, PAD = BlCfg<MODE>::PAD;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lhi = lane >> 4;
    const float uinv = 1.0f / uscale;
    if (tid == 0) L.singular = 0;

    for (int k0 = 0; k0 < n; k0 += NB) {
        const int jb = min(NB, n - k0), m = n - k0;
        T *Wp = W + (int64_t)k0 * n + k0;
        // ---- panel: rows k0..n-1, columns k0..k0+jb-1, in registers.
        // The column being eliminated is always a[.][0]: after each step the
        // row arrays rotate left by one (the finished multiplier moves to the
        // end), so the column loop is not unrolled and every register index
        // stays static; NB rotations restore the original column order.
        T a[BL_SLOTS][NB];
        int pos[BL_SLOTS];
#pragma unroll
        for (int s = 0; s < BL_SLOTS; ++s) {
            const int r = tid + BL_NT * s;
            pos[s] = r < m ? r : BL_BIG;
            // unconditional loads from clamped positions, then zeroed
            const T *wr = Wp + (int64_t)min(r, m - 1) * n;
#pragma unroll
            for (int c = 0; c < NB; ++c) a[s][c] = wr[min(c, jb - 1)];
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (!(r < m && c < jb)) a[s][c] = (T)0;
        }
        if (tid == 0) L.nmv = 0;
#pragma unroll 1
        for (int j = 0; j < NB; ++j) {
            if (j < jb) {
                const int par = j & 1;
                uint64_t bk = 0;
                uint32_t bq = 0;
#pragma unroll
                for (int s = 0; s < BL_SLOTS; ++s)
                    if (pos[s] >= j && pos[s] < m) {
                        const uint64_t kk = bl_key(a[s][0]);
                        const uint32_t qq = ~(uint32_t)pos[s];
                        if (bl_better(kk, qq, bk, bq)) { bk = kk; bq = qq; }
                    }
#pragma unroll
                for (int msk = 32; msk >= 1; msk >>= 1) {
                    const uint64_t ok = bl_shfl_xor_u64(bk, msk);
                    const uint32_t oq = (uint32_t)__shfl_xor((int)bq, msk);
                    if (bl_better(ok, oq, bk, bq)) { bk = ok; bq = oq; }
                }
                const int wpos = (int)~bq;
#pragma unroll
                for (int s = 0; s < BL_SLOTS; ++s)
                    if (pos[s] == wpos && bq != 0) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) L.cand_row[par][wave][c] = a[s][c];
                    }
                if (lane == 0) { L.cand_key[par][wave] = bk; L.cand_q[par][wave] = bq; }
                __syncthreads();
                uint64_t gk = L.cand_key[par][0];
                uint32_t gq = L.cand_q[par][0];
                int gw = 0;
#pragma unroll
                for (int w = 1; w < 4; ++w) {
                    const uint64_t k2 = L.cand_key[par][w];
                    const uint32_t q2 = L.cand_q[par][w];
                    if (bl_better(k2, q2, gk, gq)) { gk = k2; gq = q2; gw = w; }
                }
                const int ppos = (int)~gq;      // current position of the pivot row
                // pivot row: active columns only (the tail holds its multipliers)
                T prow[NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) prow[c] = c < NB - j ? L.cand_row[par][gw][c] : (T)0;
                const T piv = prow[0];
#pragma unroll
                for (int s = 0; s < BL_SLOTS; ++s) {
                    if (pos[s] == ppos) pos[s] = j;
                    else if (pos[s] == j) pos[s] = ppos;
                }
                if (!(piv != (T)0)) {           // exact zero (or NaN) pivot: no scaling
                    if (tid == 0) L.singular = 1;
                } else {
                    const T rinv = (T)1 / piv;
#pragma unroll
                    for (int s = 0; s < BL_SLOTS; ++s)
                        if (pos[s] > j && pos[s] < m) {
                            const T l = (MODE == LA_PREC_FP64) ? a[s][0] / piv : a[s][0] * rinv;
#pragma unroll
                            for (int c = 1; c < NB; ++c) a[s][c] -= l * prow[c];
                            a[s][0] = l;
                        }
                }
            }
#pragma unroll
            for (int s = 0; s < BL_SLOTS; ++s) {
                const T t0 = a[s][0];
#pragma unroll
                for (int c = 0; c < NB - 1; ++c) a[s][c] = a[s][c + 1];
                a[s][NB - 1] = t0;
            }
        }
        // ---- panel write-back at the final positions, L11, move list
#pragma unroll
        for (int s = 0; s < BL_SLOTS; ++s) {
            const int r = tid + BL_NT * s;
            if (r < m) {
                const int q = pos[s];
                T *dst = Wp + (int64_t)q * n;
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < jb) dst[c] = a[s][c];
                if (q < jb) {
#pragma unroll
                    for (int c = 0; c < NB; ++c) L.l11[q][c] = c < q ? a[s][c] : (T)0;
                }
                if (q != r) {
                    const int idx = atomicAdd(&L.nmv, 1);
                    L.mv_src[idx] = r;
                    L.mv_dst[idx] = q;
                }
            }
        }
        __syncthreads();
        const int nmv = L.nmv;
        // ---- the same row permutation on every column outside the panel
        if (nmv > 0) {
            T (*mstg)[BL_MVC] = reinterpret_cast<T (*)[BL_MVC]>(L.stg);
            const int outside = n - jb;
            for (int x0 = 0; x0 < outside; x0 += BL_MVC) {
                const int cw = min(BL_MVC, outside - x0);
                for (int e = tid; e < nmv * cw; e += BL_NT) {
                    const int i = e / cw, cc = e - i * cw;
                    const int x = x0 + cc, col = x < k0 ? x : x + jb;
                    mstg[i][cc] = W[(int64_t)(k0 + L.mv_src[i]) * n + col];
                }
                __syncthreads();
                for (int e = tid; e < nmv * cw; e += BL_NT) {
                    const int i = e / cw, cc = e - i * cw;
                    const int x = x0 + cc, col = x < k0 ? x : x + jb;
                    W[(int64_t)(k0 + L.mv_dst[i]) * n + col] = mstg[i][cc];
                }
                __syncthreads();
            }
            if (tid < nmv) L.mv_tmp[tid] = perm[k0 + L.mv_src[tid]];
            __syncthreads();
            if (tid < nmv) perm[k0 + L.mv_dst[tid]] = L.mv_tmp[tid];
            __syncthreads();
        }
        // ---- U12 = L11^-1 A12 and A22 -= L21 U12, in column chunks
        const int c_begin = k0 + jb;
        const int R = n - c_begin;
        E (*ustg)[NB + PAD] = reinterpret_cast<E (*)[NB + PAD]>(L.stg);
        for (int cc0 = c_begin; cc0 < n; cc0 += BL_CHUNK) {
            const int cw = min(BL_CHUNK, n - cc0);
            {
                const int col = cc0 + tid;
                T x[NB];
                const int colc = min(col, n - 1);
#pragma unroll
                for (int i = 0; i < NB; ++i) x[i] = W[(int64_t)(k0 + min(i, jb - 1)) * n + colc];
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (!(tid < cw && i < jb)) x[i] = (T)0;
#pragma unroll
                for (int i = 1; i < NB; ++i) {
                    T acc = x[i];
#pragma unroll
                    for (int k = 0; k < i; ++k) acc -= L.l11[i][k] * x[k];
                    x[i] = acc;
                }
                if (tid < cw) {
#pragma unroll
                    for (int i = 0; i < NB; ++i)
                        if (i < jb) W[(int64_t)(k0 + i) * n + col] = x[i];
                }
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    ustg[tid][i] = (MODE == LA_PREC_FP16) ? (E)((float)x[i] * uscale) : (E)x[i];
            }
            __syncthreads();
            // each wave: one 16-row tile at a time against all column tiles of
            // the chunk, G column tiles per step with their C values loaded
            // together so that 4*G reads per lane are in flight
            const int rt_n = (R + 15) >> 4, ct_n = (cw + 15) >> 4;
            constexpr int G = sizeof(T) == 8 ? 4 : 8;
            for (int rt = wave; rt < rt_n; rt += 4) {
                const int row = c_begin + 16 * rt + l16;
                const bool rok = row < n;
                // unconditional loads from clamped positions (issued
                // together); out-of-range elements are replaced by zero
                const T *lrow = W + (int64_t)min(row, n - 1) * n + k0;
                h8_t ah;
                T af[NB / 4];
                if constexpr (MODE == LA_PREC_FP16) {
                    T lv[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) lv[e] = lrow[min(8 * lhi + e, jb - 1)];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = 8 * lhi + e;
                        ah[e] = (_Float16)((rok && k < jb) ? lv[e] : 0.0f);
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < NB / 4; ++s) af[s] = lrow[min(4 * s + lhi, jb - 1)];
#pragma unroll
                    for (int s = 0; s < NB / 4; ++s) {
                        const int k = 4 * s + lhi;
                        if (!(rok && k < jb)) af[s] = (T)0;
                    }
                }
                for (int ct0 = 0; ct0 < ct_n; ct0 += G) {
                    T cv[G][4];
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const int col = cc0 + 16 * (ct0 + g) + l16;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int crow = c_begin + 16 * rt + BlMfma<MODE>::drow(lhi, r);
                            cv[g][r] = (crow < n && col < cc0 + cw) ? W[(int64_t)crow * n + col]
                                                                    : (T)0;
                        }
                    }
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const int cl = 16 * (ct0 + g) + l16;     // < BL_CHUNK: ct < 16
                        T d[4];
                        if constexpr (MODE == LA_PREC_FP16) {
                            const h8_t bfv = *reinterpret_cast<const h8_t *>(&ustg[cl][8 * lhi]);
                            f4_t acc = {0.f, 0.f, 0.f, 0.f};
                            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bfv, acc, 0, 0, 0);
#pragma unroll
                            for (int r = 0; r < 4; ++r) d[r] = acc[r] * uinv;
                        } else if constexpr (MODE == LA_PREC_FP64) {
                            d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                            for (int s = 0; s < NB / 4; ++s)
                                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], ustg[cl][4 * s + lhi],
                                                                           acc, 0, 0, 0);
#pragma unroll
                            for (int r = 0; r < 4; ++r) d[r] = acc[r];
                        } else {
                            f4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                            for (int s = 0; s < NB / 4; ++s)
                                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[s], ustg[cl][4 * s + lhi],
                                                                           acc, 0, 0, 0);
#pragma unroll
                            for (int r = 0; r < 4; ++r) d[r] = acc[r];
                        }
                        const int col = cc0 + cl;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int crow = c_begin + 16 * rt + BlMfma<MODE>::drow(lhi, r);
                            if (crow < n && col < cc0 + cw)
                                W[(int64_t)crow * n + col] = cv[g][r] - d[r];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------ triangular solves in LDS
// Y (n x g, row-major in LDS) <- L^-1 Y (unit lower), then U^-1 Y, with the
// packed factor in W.  64-row blocks: the off-diagonal part is a 16-lanes-
// per-row GEMV, the diagonal block is solved by one wave per right-hand side
// with the block row in registers and readlane broadcasts.
template <typename T>
__device__ void bl_gemv_sub(int n, const T *__restrict__ W, T *Y, int g, int b0, int bs, int j0,
                            int j1)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = lane >> 4, l16 = lane & 15;
    for (int i0 = wave * 4; i0 < bs; i0 += 16) {
        const int il = i0 + sub;
        const int i = b0 + il;
        T acc[BL_MAXRHS];
#pragma unroll
        for (int c = 0; c < BL_MAXRHS; ++c) acc[c] = (T)0;
        if (il < bs) {
            const T *wr = W + (int64_t)i * n;
#pragma unroll 8
            for (int j = j0 + l16; j < j1; j += 16) {
                const T w = wr[j];
#pragma unroll
                for (int c = 0; c < BL_MAXRHS; ++c)
                    if (c < g) acc[c] += w * Y[j * g + c];
            }
        }
#pragma unroll
        for (int c = 0; c < BL_MAXRHS; ++c) {
            if (c < g) {
#pragma unroll
                for (int msk = 8; msk >= 1; msk >>= 1) acc[c] += __shfl_xor(acc[c], msk);
            }
        }
        if (il < bs && l16 == 0) {
#pragma unroll
            for (int c = 0; c < BL_MAXRHS; ++c)
                if (c < g) Y[i * g + c] -= acc[c];
        }
    }
}

template <typename T>
__device__ void bl_trsv_lu(int n, const T *__restrict__ W, T *Y, int g)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // forward: unit lower
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int bs = min(64, n - b0);
        if (b0 > 0) {
            bl_gemv_sub<T>(n, W, Y, g, b0, bs, 0, b0);
            __syncthreads();
        }
        for (int c = wave; c < g; c += 4) {
            const int i = lane;
            T lrow[64];
            // unconditional loads from clamped positions, then zeroed
            const T *wr = W + (int64_t)(b0 + min(i, bs - 1)) * n + b0;
#pragma unroll
            for (int k = 0; k < 64; ++k) lrow[k] = wr[min(k, bs - 1)];
#pragma unroll
            for (int k = 0; k < 64; ++k)
                if (!(i < bs && k < i)) lrow[k] = (T)0;
            T y = i < bs ? Y[(b0 + i) * g + c] : (T)0;
#pragma unroll
            for (int k = 0; k < 63; ++k) {
                if (k + 1 < bs) {
                    const T yk = bl_bcast(y, k);
                    if (i > k) y -= lrow[k] * yk;
                }
            }
            if (i < bs) Y[(b0 + i) * g + c] = y;
        }
        __syncthreads();
    }
    // backward: upper with diagonal
    const int nblk = (n + 63) / 64;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int b0 = bi * 64, bs = min(64, n - b0);
        if (b0 + bs < n) {
            bl_gemv_sub<T>(n, W, Y, g, b0, bs, b0 + bs, n);
            __syncthreads();
        }
        for (int c = wave; c < g; c += 4) {
            const int i = lane;
            T urow[64];
            const T *wr = W + (int64_t)(b0 + min(i, bs - 1)) * n + b0;
#pragma unroll
            for (int k = 0; k < 64; ++k) urow[k] = wr[min(k, bs - 1)];
#pragma unroll
            for (int k = 0; k < 64; ++k)
                if (!(i < bs && k >= i && k < bs)) urow[k] = (T)1;
            T y = i < bs ? Y[(b0 + i) * g + c] : (T)0;
#pragma unroll
            for (int k = 63; k >= 0; --k) {
                if (k < bs) {
                    if (i == k) y = y / urow[k];
                    const T yk = bl_bcast(y, k);
                    if (i < k) y -= urow[k] * yk;
                }
            }
            if (i < bs) Y[(b0 + i) * g + c] = y;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- refinement
// r = b - A x (FP64) for the rows in pivot order; Y[i][c] = r[perm[i]][c]
// in working precision; rn[c] = max |r_c| (all threads get it).
template <int MODE>
__device__ void bl_residual(int n, int nrhs, int c0, int g, bool first,
                            const double *__restrict__ Ap, const double *__restrict__ Bp,
                            const double *Xp, const int *__restrict__ perm, BlLds<MODE> &L,
                            double *rn)
{
    typedef typename BlCfg<MODE>::T T;
    T *Y = reinterpret_cast<T *>(L.y);
    double *xs = reinterpret_cast<double *>(L.stg);      // x block, [j][g], in LDS
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = lane >> 4, l16 = lane & 15;
    if (!first) {
        for (int e = tid; e < n * g; e += BL_NT) {
            const int j = e / g, c = e - j * g;
            xs[e] = Xp[(int64_t)j * nrhs + c0 + c];
        }
        __syncthreads();
    }
    double rmax[BL_MAXRHS];
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c) rmax[c] = 0.0;
    for (int i0 = wave * 4; i0 < n; i0 += 16) {
        const int i = i0 + sub;
        double acc[BL_MAXRHS];
#pragma unroll
        for (int c = 0; c < BL_MAXRHS; ++c) acc[c] = 0.0;
        const int row = perm[min(i, n - 1)];       // unconditional; unused past n
        if (i < n && !first) {
            const double *ar = Ap + (int64_t)row * n;
            // 8 independent row loads in flight per lane
#pragma unroll 8
            for (int j = l16; j < n; j += 16) {
                const double a = ar[j];
#pragma unroll
                for (int c = 0; c < BL_MAXRHS; ++c)
                    if (c < g) acc[c] = fma(a, xs[j * g + c], acc[c]);
            }
        }
        if (!first) {
#pragma unroll
            for (int c = 0; c < BL_MAXRHS; ++c)
                if (c < g) {
#pragma unroll
                    for (int msk = 8; msk >= 1; msk >>= 1) acc[c] += __shfl_xor(acc[c], msk);
                }
        }
        // the right-hand side row, loaded unconditionally (clamped column)
        const double *br = Bp + (int64_t)row * nrhs + c0;
        double bv[BL_MAXRHS];
#pragma unroll
        for (int c = 0; c < BL_MAXRHS; ++c) bv[c] = br[min(c, g - 1)];
        if (i < n && l16 == 0) {
#pragma unroll
            for (int c = 0; c < BL_MAXRHS; ++c)
                if (c < g) {
                    const double r = bv[c] - acc[c];
                    Y[i * g + c] = (T)r;
                    rmax[c] = fmax(rmax[c], fabs(r));
                }
        }
    }
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c) {
        const double v = bl_wave_max(rmax[c]);
        if (lane == 0) L.red[wave * BL_MAXRHS + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c)
        rn[c] = fmax(fmax(L.red[c], L.red[BL_MAXRHS + c]),
                     fmax(L.red[2 * BL_MAXRHS + c], L.red[3 * BL_MAXRHS + c]));
    __syncthreads();
}

// x += Y (columns c0..c0+g-1); xn[c] = max |x_c| afterwards
template <int MODE>
__device__ void bl_update_x(int n, int nrhs, int c0, int g, double *Xp, BlLds<MODE> &L,
                            bool first, double *xn)
{
    typedef typename BlCfg<MODE>::T T;
    const T *Y = reinterpret_cast<const T *>(L.y);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double xm[BL_MAXRHS];
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c) xm[c] = 0.0;
    for (int e = tid; e < n * g; e += BL_NT) {
        const int i = e / g, c = e - i * g;
        double *xp = Xp + (int64_t)i * nrhs + c0 + c;
        const double v = (first ? 0.0 : *xp) + (double)Y[e];
        *xp = v;
#pragma unroll
        for (int q = 0; q < BL_MAXRHS; ++q)
            if (q == c) xm[q] = fmax(xm[q], fabs(v));
    }
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c) {
        const double v = bl_wave_max(xm[c]);
        if (lane == 0) L.red[4 * BL_MAXRHS + wave * BL_MAXRHS + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < BL_MAXRHS; ++c)
        xn[c] = fmax(fmax(L.red[4 * BL_MAXRHS + c], L.red[5 * BL_MAXRHS + c]),
                     fmax(L.red[6 * BL_MAXRHS + c], L.red[7 * BL_MAXRHS + c]));
    __syncthreads();
}

template <int MODE>
__device__ void bl_refine(int n, int nrhs, const double *__restrict__ Ap,
                          const double *__restrict__ Bp, double *Xp,
                          const typename BlCfg<MODE>::T *__restrict__ W,
                          const int *__restrict__ perm, double anorm, double tol, int max_iter,
                          BlLds<MODE> &L, int &status, int &iters, double &eta_out)
{
    typedef typename BlCfg<MODE>::T T;
    T *Y = reinterpret_cast<T *>(L.y);
    const int cap = BL_Y_BYTES / (int)sizeof(T);
    const int xcap = BL_STG_BYTES / (int)sizeof(double);
    const int gmax = max(1, min(min(nrhs, BL_MAXRHS), min(cap, xcap) / n));
    status = LA_OK;
    iters = 0;
    eta_out = 0.0;
    for (int c0 = 0; c0 < nrhs; c0 += gmax) {
        const int g = min(gmax, nrhs - c0);
        double bn[BL_MAXRHS], xn[BL_MAXRHS], rn[BL_MAXRHS];
#pragma unroll
        for (int c = 0; c < BL_MAXRHS; ++c) { bn[c] = 0.0; xn[c] = 0.0; }
        double eta = 1.0, eta_prev = 1e300;
        int stalls = 0, it = 0;
        for (;; ++it) {
            bl_residual<MODE>(n, nrhs, c0, g, it == 0, Ap, Bp, Xp, perm, L, rn);
            if (it == 0) {
#pragma unroll
                for (int c = 0; c < BL_MAXRHS; ++c) bn[c] = rn[c];
            }
            eta = 0.0;
#pragma unroll
            for (int c = 0; c < BL_MAXRHS; ++c)
                if (c < g) {
                    const double den = anorm * xn[c] + bn[c];
                    const double e = den > 0.0 ? rn[c] / den : (rn[c] > 0.0 ? 1e300 : 0.0);
                    eta = (e != e) ? 1e300 : fmax(eta, e);
                }
            if (it > 0) {
                if (eta <= tol) break;
                if (MODE == LA_PREC_FP64) break;
                if (it > max_iter) { status = LA_NOT_CONVERGED; break; }
                if (eta > 0.5 * eta_prev) {
                    if (++stalls >= 3) { status = LA_NOT_CONVERGED; break; }
                } else {
                    stalls = 0;
                }
            }
            eta_prev = eta;
            bl_trsv_lu<T>(n, W, Y, g);
            bl_update_x<MODE>(n, nrhs, c0, g, Xp, L, it == 0, xn);
        }
        iters = max(iters, it - 1);
        eta_out = fmax(eta_out, eta);
    }
}

// ---------------------------------------------------------------- kernel
template <int MODE>
__global__ __launch_bounds__(BL_NT, 2)
void k_lu_ir_batched(int count, const int32_t *__restrict__ order, const int32_t *__restrict__ nvec,
                     const int32_t *__restrict__ rhsvec, const int64_t *__restrict__ offA,
                     const int64_t *__restrict__ offB, const int64_t *__restrict__ offW,
                     const int64_t *__restrict__ offN, const double *A, const double *B, double *X,
                     typename BlCfg<MODE>::T *Wbuf, int32_t *permbuf, double tol, int max_iter,
                     int32_t *status, int32_t *iters, double *berr, uint32_t *queue)
{
    __shared__ BlLds<MODE> L;
    for (;;) {
        if (threadIdx.x == 0) L.prob = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int pi = L.prob;
        __syncthreads();
        if (pi >= count) break;
        const int prob = order[pi];
        const int n = nvec[prob], nrhs = rhsvec[prob];
        if (n <= 0 || n > BL_MAXN || nrhs <= 0 || nrhs > BL_MAXRHS) {
            if (threadIdx.x == 0) {
                status[prob] = LA_ERR_SHAPE;
                iters[prob] = 0;
                berr[prob] = 0.0;
            }
            continue;
        }
        const double *Ap = A + offA[prob];
        const double *Bp = B + offB[prob];
        double *Xp = X + offB[prob];
        typename BlCfg<MODE>::T *Wp = Wbuf + offW[prob];
        int *perm = permbuf + offN[prob];
        double anorm, amax;
        bl_convert<MODE>(n, Ap, Wp, perm, L, anorm, amax);
        float uscale = 1.0f;
        if (MODE == LA_PREC_FP16 && amax > 0.0 && amax < 1e300)
            uscale = ldexpf(1.0f, 6 - ilogb(amax));
        bl_getrf<MODE>(n, Wp, perm, uscale, L);
        __syncthreads();
        const int sing = L.singular;
        int st = LA_SINGULAR, it = 0;
        double eta = 0.0;
        if (!sing) bl_refine<MODE>(n, nrhs, Ap, Bp, Xp, Wp, perm, anorm, tol, max_iter, L, st, it, eta);
        if (threadIdx.x == 0) {
            status[prob] = st;
            iters[prob] = it;
            berr[prob] = eta;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned d = atomicAdd(&queue[1], 1u);
        if (d == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

static const void *lu_ir_kernel(int prec) {
    switch (prec) {
    case LA_PREC_FP64: return (const void *)k_lu_ir_batched<LA_PREC_FP64>;
    case LA_PREC_FP32: return (const void *)k_lu_ir_batched<LA_PREC_FP32>;
    case LA_PREC_FP16: return (const void *)k_lu_ir_batched<LA_PREC_FP16>;
    default: return nullptr;
    }
}

int lu_ir_batched_grid(la_ctx *ctx, int prec)
{
    const void *k = lu_ir_kernel(prec);
    int per_cu = 0;
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, BL_NT, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    (void)hipGetLastError();
    return std::max(1, ctx->num_cus) * per_cu;
}

int launch_lu_ir_batched(la_ctx *ctx, hipStream_t s, int prec, int grid, int count,
                         const int32_t *order, const int32_t *nvec, const int32_t *rhsvec,
                         const int64_t *offA, const int64_t *offB, const int64_t *offW,
                         const int64_t *offN, const double *A, const double *B, double *X,
                         void *Wbuf, int32_t *permbuf, double tol, int max_iter, int32_t *status,
                         int32_t *iters, double *berr, uint32_t *queue)
{
    switch (prec) {
    case LA_PREC_FP64:
        hipLaunchKernelGGL(k_lu_ir_batched<LA_PREC_FP64>, dim3(grid), dim3(BL_NT), 0, s, count,
                           order, nvec, rhsvec, offA, offB, offW, offN, A, B, X, (double *)Wbuf,
                           permbuf, tol, max_iter, status, iters, berr, queue);
        break;
    case LA_PREC_FP32:
        hipLaunchKernelGGL(k_lu_ir_batched<LA_PREC_FP32>, dim3(grid), dim3(BL_NT), 0, s, count,
                           order, nvec, rhsvec, offA, offB, offW, offN, A, B, X, (float *)Wbuf,
                           permbuf, tol, max_iter, status, iters, berr, queue);
        break;
    case LA_PREC_FP16:
        hipLaunchKernelGGL(k_lu_ir_batched<LA_PREC_FP16>, dim3(grid), dim3(BL_NT), 0, s, count,
                           order, nvec, rhsvec, offA, offB, offW, offN, A, B, X, (float *)Wbuf,
                           permbuf, tol, max_iter, status, iters, berr, queue);
        break;
    default:
        return set_error(ctx, LA_ERR_ARG, "batched LU-IR: precision must be FP64, FP32 or FP16");
    }
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// ------------------------------------------------------------- generator
// Batch of A_i = U_i diag(sigma) V_i^T (Householder products, as
// la_gen_svd_matrix_dev) with B_i standard normal (n_i x nrhs_i) and the
// exact solutions X*_i = V_i diag(1/sigma) U_i^T B_i.  Problem p draws its
// reflector vectors from streams stream_base + 256 p + i (U) and
// stream_base + 256 p + 128 + i (V), i < nrefl <= 128, and B_i from stream
// stream_base + 256 p + 255 (row-major order).  One workgroup per problem.
__device__ inline double bl_normal_at(uint64_t key, int64_t e) {
    const int64_t p = e >> 1;
    const double two_pi = 6.283185307179586476925286766559;
    const double u1 = (double)((mix64(key + (uint64_t)(2 * p + 1) * GOLDEN) >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(mix64(key + (uint64_t)(2 * p + 2) * GOLDEN) >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    const double t = two_pi * u2;
    return (e & 1) ? r * sin(t) : r * cos(t);
}

__device__ double bl_load_reflector(uint64_t seed, uint64_t stream, int n, double *v, double *red)
{
    const uint64_t key = mix64(seed + (stream + 1) * GOLDEN);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double s = 0.0;
    for (int i = tid; i < n; i += BL_NT) {
        const double x = bl_normal_at(key, i);
        v[i] = x;
        s += x * x;
    }
    s = bl_wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    const double vtv = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return vtv;
}

__global__ __launch_bounds__(BL_NT)
void k_gen_svd_batch(int count, const int32_t *nvec, const int32_t *rhsvec, const int64_t *offA,
                     const int64_t *offB, double kappa, int nrefl, uint64_t seed,
                     uint64_t stream_base, double *A, double *B, double *Xs)
{
    __shared__ double v[BL_MAXN], w[BL_MAXN], red[8];
    const int prob = blockIdx.x;
    if (prob >= count) return;
    const int n = nvec[prob], nrhs = rhsvec[prob];
    if (n <= 0 || n > BL_MAXN || nrhs <= 0 || nrhs > BL_MAXRHS) return;
    double *Ap = A + offA[prob];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t base = stream_base + 256ull * (uint64_t)prob;
    auto sigma = [&](int i) { return n > 1 ? pow(kappa, -(double)i / (double)(n - 1)) : 1.0; };
    for (int64_t e = tid; e < (int64_t)n * n; e += BL_NT) {
        const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
        Ap[e] = r == c ? sigma(r) : 0.0;
    }
    __syncthreads();
    for (int i = nrefl - 1; i >= 0; --i) {           // A <- H_i A
        const double vtv = bl_load_reflector(seed, base + i, n, v, red);
        const double f = 2.0 / vtv;
        for (int c = tid; c < n; c += BL_NT) {
            double s = 0.0;
            for (int r = 0; r < n; ++r) s += v[r] * Ap[(int64_t)r * n + c];
            w[c] = s;
        }
        __syncthreads();
        for (int64_t e = tid; e < (int64_t)n * n; e += BL_NT) {
            const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
            Ap[e] -= f * v[r] * w[c];
        }
        __syncthreads();
    }
    for (int i = nrefl - 1; i >= 0; --i) {           // A <- A H'_i
        const double vtv = bl_load_reflector(seed, base + 128 + i, n, v, red);
        const double f = 2.0 / vtv;
        for (int r = wave; r < n; r += 4) {
            double s = 0.0;
            for (int c = lane; c < n; c += 64) s += Ap[(int64_t)r * n + c] * v[c];
            s = bl_wave_sum(s);
            if (lane == 0) w[r] = s;
        }
        __syncthreads();
        for (int64_t e = tid; e < (int64_t)n * n; e += BL_NT) {
            const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
            Ap[e] -= f * w[r] * v[c];
        }
        __syncthreads();
    }
    // right-hand sides and exact solutions
    double *Bp = B + offB[prob], *Xp = Xs + offB[prob];
    const uint64_t bkey = mix64(seed + (base + 255 + 1) * GOLDEN);
    for (int e = tid; e < n * nrhs; e += BL_NT) {
        const double x = bl_normal_at(bkey, e);
        Bp[e] = x;
        Xp[e] = x;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = 0; q < nrefl; ++q) {
            // pass 0: U^T y = H_nrefl..H_1 y (H_1 first); pass 1: V y (H'_nrefl first)
            const int i = pass == 0 ? q : nrefl - 1 - q;
            const double vtv = bl_load_reflector(seed, base + (pass ? 128 : 0) + i, n, v, red);
            const double f = 2.0 / vtv;
            for (int c = wave; c < nrhs; c += 4) {
                double s = 0.0;
                for (int r = lane; r < n; r += 64) s += v[r] * Xp[r * nrhs + c];
                s = bl_wave_sum(s);
                if (lane == 0) w[c] = s;
            }
            __syncthreads();
            for (int e = tid; e < n * nrhs; e += BL_NT) {
                const int r = e / nrhs, c = e - r * nrhs;
                Xp[e] -= f * v[r] * w[c];
            }
            __syncthreads();
        }
        if (pass == 0) {
            for (int e = tid; e < n * nrhs; e += BL_NT) Xp[e] /= sigma(e / nrhs);
            __syncthreads();
        }
    }
}

}  // namespace lahip

using namespace lahip;

extern "C" int la_gen_svd_batch_dev(la_ctx *ctx, int64_t count, const int32_t *d_n,
                                    const int32_t *d_nrhs, const int64_t *d_offA,
                                    const int64_t *d_offB, double kappa, int nrefl, uint64_t seed,
                                    uint64_t stream_base, double *A, double *B, double *Xstar)
{
    if (!ctx || !d_n || !d_nrhs || !d_offA || !d_offB || !A || !B || !Xstar || count <= 0 ||
        count > INT32_MAX || nrefl < 0 || nrefl > 128 || !(kappa >= 1.0))
        return LA_ERR_ARG;
    hipLaunchKernelGGL(k_gen_svd_batch, dim3((unsigned)count), dim3(BL_NT), 0, ctx->stream,
                       (int)count, d_n, d_nrhs, d_offA, d_offB, kappa, nrefl, seed, stream_base, A,
                       B, Xstar);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}
