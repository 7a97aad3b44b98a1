// Blocked right-looking LU with partial pivoting (row-major, in place).
//
// Pivot semantics follow LUDecomposition::new (src/decomp/lu.rs:131-160):
// the pivot of column j is the FIRST row at or below j with the strictly
// largest |a_ij| (ties go to the lowest row index); a NaN never displaces the
// current candidate, so a NaN on the diagonal stays the pivot; rows are
// swapped whole; the multipliers are a_ij / a_jj (a true division) and are
// left unscaled when the pivot is exactly zero.
//
// Structure (two-level blocking):
//   for each outer block of NB_OUT columns
//     for each inner panel of NB_IN columns inside it
//       panel kernel     : cooperative, P workgroups, rows held in LDS,
//                          one cross-workgroup exchange per column
//       row-swap kernel  : net permutation of the panel applied to every
//                          column outside the panel and to the perm vector
//       trsm + gemm      : update the rest of the outer block
//     trsm + gemm        : U12 = L11^-1 A12, A22 -= L21 U12 (MFMA, K = NB_OUT)
#include "la_common.hpp"

namespace lahip {

constexpr int PANEL_MAXP = 256;     // max cooperative workgroups in a panel
constexpr int NB_IN = 64;           // inner panel width
constexpr int NB_OUT = 256;         // outer block width (trailing GEMM K)

// ---------------------------------------------------------------- keys
// Ordered key for |v|: NaN -> 0 (never wins), otherwise bits(|v|)+1, so that
// unsigned comparison of keys is comparison of |v| and 0 beats NaN.
__device__ inline uint64_t abs_key(double v) {
    if (v != v) return 0;
    return (uint64_t)(__double_as_longlong(v) & 0x7fffffffffffffffLL) + 1;
}
__device__ inline uint64_t abs_key(float v) {
    if (v != v) return 0;
    return (uint64_t)(__float_as_uint(v) & 0x7fffffffu) + 1;
}
constexpr uint64_t KEY_MAX = 0x7fffffffffffffffULL;
constexpr uint64_t TAG_BIT = 0x8000000000000000ULL;
constexpr int IDX_NONE = 0x7fffffff;

__device__ inline bool better(uint64_t k1, int i1, uint64_t k0, int i0) {
    return k1 > k0 || (k1 == k0 && i1 < i0);
}

__device__ inline void wave_argmax(uint64_t &key, int &idx) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)key, off);
        uint32_t hi = __shfl_xor((uint32_t)(key >> 32), off);
        int oi = __shfl_xor(idx, off);
        uint64_t ok = ((uint64_t)hi << 32) | lo;
        if (better(ok, oi, key, idx)) { key = ok; idx = oi; }
    }
}

// Workgroup argmax; every thread returns the winner.  Uses two barriers.
__device__ inline void wg_argmax(uint64_t &key, int &idx, uint64_t *s_key, int *s_idx) {
    wave_argmax(key, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_key[wave] = key; s_idx[wave] = idx; }
    __syncthreads();
    key = s_key[0]; idx = s_idx[0];
    const int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; ++w)
        if (better(s_key[w], s_idx[w], key, idx)) { key = s_key[w]; idx = s_idx[w]; }
    __syncthreads();
}

__device__ inline int64_t shfl_i64(int64_t v, int src_lane) {
    uint32_t lo = __shfl((uint32_t)v, src_lane);
    uint32_t hi = __shfl((uint32_t)((uint64_t)v >> 32), src_lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ inline double shfl_t(double v, int src_lane) { return __shfl(v, src_lane); }
__device__ inline float shfl_t(float v, int src_lane) { return __shfl(v, src_lane); }

__device__ inline uint64_t ld_relaxed(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_status(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Net row movement of a panel, computed by one wave entirely in registers:
// lane l holds entry l (panel row kglob+l) and extra entry l (a row below
// the panel that received a swap).  The swaps are replayed on source-row
// labels; swapmap receives count, then (dst, src) pairs of global rows whose
// content moved.
__device__ void panel_net_moves(int NB, int kglob, const int *s_piv, int64_t *swapmap)
{
    const int lane = threadIdx.x & 63;
    int64_t pa = (lane < NB) ? (int64_t)kglob + lane : -1, sa = pa;
    int64_t pb = -1, sb = -1;
    int cnt = 0;
    for (int s = 0; s < NB; ++s) {
        const int b = s_piv[s];
        if (b == s) continue;
        const int64_t v_s = shfl_i64(sa, s);
        if (b < NB) {
            const int64_t v_b = shfl_i64(sa, b);
            if (lane == s) sa = v_b;
            if (lane == b) sa = v_s;
        } else {
            const int64_t gb = (int64_t)kglob + b;
            const uint64_t m = __ballot(lane < cnt && pb == gb);
            int L;
            if (m) {
                L = __ffsll((unsigned long long)m) - 1;
            } else {
                L = cnt;
                if (lane == L) { pb = gb; sb = gb; }
                ++cnt;
            }
            const int64_t v_b = shfl_i64(sb, L);
            if (lane == s) sa = v_b;
            if (lane == L) sb = v_s;
        }
    }
    const bool ma = (lane < NB) && pa != sa;
    const bool mb = (lane < cnt) && pb != sb;
    const uint64_t bal_a = __ballot(ma), bal_b = __ballot(mb);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int na = __popcll(bal_a);
    if (ma) {
        const int o = __popcll(bal_a & lt);
        swapmap[1 + 2 * o] = pa;
        swapmap[2 + 2 * o] = sa;
    }
    if (mb) {
        const int o = na + __popcll(bal_b & lt);
        swapmap[1 + 2 * o] = pb;
        swapmap[2 + 2 * o] = sb;
    }
    if (lane == 0) swapmap[0] = na + __popcll(bal_b);
}

// ------------------------------------------------------------- panel kernel
// Factors the M x NB panel whose top-left element is A (row-major, lda),
// global column/row offset kglob.  Workgroup w owns panel rows
// [w*R, min(w*R+R, M)) in LDS for the whole kernel.
//
// Per column j, every workgroup publishes its best candidate (key, row) in
// slot[buf][w] (two 64-bit words, each carrying the step tag so that a torn
// read is detected and retried) together with the candidate's full panel
// row; the owner of row j also publishes row j.  After all P slots show the
// current step, each workgroup reduces them to the same winner, swaps rows j
// and p in its own rows, scales the multipliers and applies the rank-1
// update.  Slots and row buffers are double-buffered by step parity: a
// workgroup can only reach step s+2 after every workgroup published step
// s+1, i.e. after every workgroup finished reading step s.
//
// swapmap receives the net row movement of the panel: count, then pairs
// (dst, src) of global rows, where dst's final content is src's content
// before the panel.
template <typename T>
__global__ __launch_bounds__(256)
void k_getrf_panel(T *A, int64_t lda, int M, int NB, int R, int kglob,
                   int64_t *ipiv, uint64_t *slots, T *rowbuf, T *diagbuf,
                   int64_t *swapmap, int32_t *status, uint64_t spin_ticks)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int LDB = NB + 1;
    T *blk = reinterpret_cast<T *>(smem_raw);
    T *urow = blk + (size_t)R * LDB;
    T *drow = urow + NB;
    __shared__ uint64_t s_key[4];
    __shared__ int s_idx[4];
    __shared__ int s_p;
    __shared__ int s_abort;
    __shared__ int s_piv[NB_IN];

    const int tid = threadIdx.x;
    const int w = blockIdx.x;
    const int P = gridDim.x;
    const int rowbase = w * R;
    const int nrows = min(R, M - rowbase);

    // load owned rows
    for (int e = tid; e < nrows * NB; e += blockDim.x) {
        int r = e / NB, c = e - r * NB;
        blk[r * LDB + c] = A[(int64_t)(rowbase + r) * lda + c];
    }
    if (tid == 0) s_abort = 0;
    __syncthreads();

    for (int j = 0; j < NB; ++j) {
        const uint32_t gs = (uint32_t)(kglob + j);
        const int buf = gs & 1;
        const uint64_t tag = ((gs >> 1) & 1) ? TAG_BIT : 0;

        // 1. local candidate
        uint64_t key = 0;
        int idx = IDX_NONE;
        for (int r = tid; r < nrows; r += blockDim.x) {
            const int gr = rowbase + r;
            if (gr < j) continue;
            T v = blk[r * LDB + j];
            uint64_t k = abs_key(v);
            if (gr == j && v != v) k = KEY_MAX;
            if (better(k, gr, key, idx)) { key = k; idx = gr; }
        }
        wg_argmax(key, idx, s_key, s_idx);

        // 2. publish candidate row (and row j if owned)
        T *myrow = rowbuf + ((size_t)buf * PANEL_MAXP + w) * NB;
        if (idx != IDX_NONE) {
            const int rl = idx - rowbase;
            for (int c = tid; c < NB; c += blockDim.x) st_dev(myrow + c, blk[rl * LDB + c]);
        }
        if (j >= rowbase && j < rowbase + nrows) {
            const int rl = j - rowbase;
            T *dr = diagbuf + (size_t)buf * NB;
            for (int c = tid; c < NB; c += blockDim.x) st_dev(dr + c, blk[rl * LDB + c]);
        }
        stores_done();
        __syncthreads();
        if (tid == 0) {
            uint64_t *slot = slots + ((size_t)buf * PANEL_MAXP + w) * 2;
            st_relaxed(slot + 0, tag | key);
            st_relaxed(slot + 1, ((uint64_t)gs << 32) | (uint32_t)idx);
        }

        // 3. wait for all P slots of this step and reduce them
        uint64_t ck = 0;
        int ci = IDX_NONE;
        if (tid < P) {
            const uint64_t *slot = slots + ((size_t)buf * PANEL_MAXP + tid) * 2;
            uint64_t t0 = wall_clock64();
            while (true) {
                uint64_t w1 = ld_relaxed(slot + 1);
                uint64_t w0 = ld_relaxed(slot + 0);
                if ((uint32_t)(w1 >> 32) == gs && (w0 & TAG_BIT) == tag) {
                    ck = w0 & ~TAG_BIT;
                    ci = (int)(uint32_t)w1;
                    break;
                }
                if (ld_status(status) != 0) { s_abort = 1; break; }
                if (wall_clock64() - t0 > spin_ticks) {
                    atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
                    s_abort = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (s_abort) return;
        wg_argmax(ck, ci, s_key, s_idx);
        // winning panel row (relative); a well-formed exchange always yields
        // j <= ci < M, anything else is reported and replaced by j
        int p = ci;
        if (p < j || p >= M) {
            if (tid == 0) atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
            p = j;
        }
        const int wp = p / R;

        // 4. fetch the pivot row (and old row j for the owner of p)
        const T *wrow = rowbuf + ((size_t)buf * PANEL_MAXP + wp) * NB;
        for (int c = tid; c < NB; c += blockDim.x) urow[c] = ld_dev(wrow + c);
        if (p != j && p >= rowbase && p < rowbase + nrows) {
            const T *dr = diagbuf + (size_t)buf * NB;
            for (int c = tid; c < NB; c += blockDim.x) drow[c] = ld_dev(dr + c);
        }
        if (tid == 0) {
            s_piv[j] = p;
            if (w == 0) ipiv[kglob + j] = (int64_t)kglob + p;
        }
        __syncthreads();
        if (p != j) {
            if (j >= rowbase && j < rowbase + nrows) {
                const int rl = j - rowbase;
                for (int c = tid; c < NB; c += blockDim.x) blk[rl * LDB + c] = urow[c];
            }
            if (p >= rowbase && p < rowbase + nrows) {
                const int rl = p - rowbase;
                for (int c = tid; c < NB; c += blockDim.x) blk[rl * LDB + c] = drow[c];
            }
        }
        __syncthreads();

        // 5. multipliers (true division, skipped for a zero pivot) + rank-1
        const T piv = urow[j];
        const int r_first = max(0, j + 1 - rowbase);
        if (piv != T(0)) {
            for (int r = r_first + tid; r < nrows; r += blockDim.x)
                blk[r * LDB + j] = blk[r * LDB + j] / piv;
        }
        __syncthreads();
        const int ncols = NB - j - 1;
        if (ncols > 0 && r_first < nrows) {
            const int nr = nrows - r_first;
            for (int e = tid; e < nr * ncols; e += blockDim.x) {
                const int r = r_first + e / ncols;
                const int c = j + 1 + e % ncols;
                blk[r * LDB + c] -= blk[r * LDB + j] * urow[c];
            }
        }
        __syncthreads();
    }

    // write back owned rows
    for (int e = tid; e < nrows * NB; e += blockDim.x) {
        int r = e / NB, c = e - r * NB;
        A[(int64_t)(rowbase + r) * lda + c] = blk[r * LDB + c];
    }

    __syncthreads();
    if (w == 0 && tid < 64) panel_net_moves(NB, kglob, s_piv, swapmap);
}


// ------------------------------------------------------ single-WG panel
// Used when the whole M x NB panel fits one workgroup's LDS.  Column j is
// formed LEFT-looking in exactly the reference's operation order
// (lu.rs:122-129): for every row i, s = sum_{k < min(i,j)} l[i][k]*u[k][j]
// accumulated for k ascending from zero, each product rounded, then
// lu[i][j] = a[i][j] - s; no fused multiply-adds.  Rows above the diagonal
// form a sequential chain (u[k][j] feeds row i > k), run by one wave with a
// shuffle broadcast per k.  When the panel is the whole matrix (kglob == 0,
// n <= NB) the factor is therefore bit-identical to the reference.
template <typename T>
__global__ __launch_bounds__(256)
void k_getrf_panel_local(T *A, int64_t lda, int M, int NB, int kglob, int64_t *ipiv,
                         int64_t *swapmap)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int LDB = NB + 1;
    T *blk = reinterpret_cast<T *>(smem_raw);
    __shared__ uint64_t s_key[4];
    __shared__ int s_idx[4];
    __shared__ int s_piv[NB_IN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (int e = tid; e < M * NB; e += blockDim.x) {
        int r = e / NB, c = e - r * NB;
        blk[r * LDB + c] = A[(int64_t)r * lda + c];
    }
    __syncthreads();

    for (int j = 0; j < NB; ++j) {
        // rows above the diagonal: forward chain, lane i = row i < j
        if (wave == 0 && j > 0) {
            T s = T(0);
            const T a = (lane < j) ? blk[lane * LDB + j] : T(0);
            for (int k = 0; k < j; ++k) {
                const T fin = a - s;
                const T uk = shfl_t(fin, k);
                if (lane == k) blk[k * LDB + j] = uk;
                if (lane > k && lane < j) {
                    const T prod = blk[lane * LDB + k] * uk;
                    s = s + prod;
                }
            }
        }
        __syncthreads();
        // rows at and below the diagonal: independent dot products over k < j
        for (int i = j + tid; i < M; i += blockDim.x) {
            T s = T(0);
            for (int k = 0; k < j; ++k) {
                const T prod = blk[i * LDB + k] * blk[k * LDB + j];
                s = s + prod;
            }
            blk[i * LDB + j] = blk[i * LDB + j] - s;
        }
        __syncthreads();
        // pivot: first strictly largest |.|, NaN never wins, NaN diagonal stays
        uint64_t key = 0;
        int idx = IDX_NONE;
        for (int i = j + tid; i < M; i += blockDim.x) {
            const T v = blk[i * LDB + j];
            uint64_t k = abs_key(v);
            if (i == j && v != v) k = KEY_MAX;
            if (better(k, i, key, idx)) { key = k; idx = i; }
        }
        wg_argmax(key, idx, s_key, s_idx);
        const int p = (idx >= j && idx < M) ? idx : j;
        if (p != j) {
            for (int c = tid; c < NB; c += blockDim.x) {
                const T t = blk[p * LDB + c];
                blk[p * LDB + c] = blk[j * LDB + c];
                blk[j * LDB + c] = t;
            }
        }
        if (tid == 0) {
            s_piv[j] = p;
            ipiv[kglob + j] = (int64_t)kglob + p;
        }
        __syncthreads();
        const T piv = blk[j * LDB + j];
        if (piv != T(0))
            for (int i = j + 1 + tid; i < M; i += blockDim.x)
                blk[i * LDB + j] = blk[i * LDB + j] / piv;
        __syncthreads();
    }
    for (int e = tid; e < M * NB; e += blockDim.x) {
        int r = e / NB, c = e - r * NB;
        A[(int64_t)r * lda + c] = blk[r * LDB + c];
    }
    __syncthreads();
    if (tid < 64) panel_net_moves(NB, kglob, s_piv, swapmap);
}

// ------------------------------------------------------------ row swaps
// Applies the panel's net row movement to columns [0, c_lo) and [c_hi, n)
// (the panel's own columns are already final) and to the permutation vector.
// Each workgroup handles 64 columns: it first gathers every source row of
// its columns into LDS, then scatters them to their destinations, so cycles
// in the permutation need no ordering.
template <typename T>
__global__ __launch_bounds__(256)
void k_apply_rowmoves(T *A, int64_t lda, int n, int c_lo, int c_hi,
                      const int64_t *swapmap, uint64_t *perm)
{
    __shared__ T tile[2 * NB_IN][64];
    __shared__ int64_t s_dst[2 * NB_IN], s_src[2 * NB_IN];
    __shared__ int s_cnt;
    const int tid = threadIdx.x;
    if (tid == 0) s_cnt = (int)swapmap[0];
    __syncthreads();
    const int cnt = s_cnt;
    if (cnt == 0) return;
    for (int e = tid; e < cnt; e += blockDim.x) {
        s_dst[e] = swapmap[1 + 2 * e];
        s_src[e] = swapmap[2 + 2 * e];
    }
    __syncthreads();

    if (blockIdx.x == gridDim.x - 1 && perm) {
        // permutation vector: perm_new[dst] = perm_old[src]
        uint64_t v = 0;
        if (tid < cnt) v = perm[s_src[tid]];
        __syncthreads();
        if (tid < cnt) perm[s_dst[tid]] = v;
        if (cnt > (int)blockDim.x) {
            // cnt <= 2*NB_IN <= blockDim.x always; kept for clarity
        }
        return;
    }
    // map this block's 64 columns, skipping the panel columns [c_lo, c_hi)
    const int left = c_lo;
    int c0 = blockIdx.x * 64;
    int col;
    const int cl = tid & 63;
    {
        int v = c0 + cl;
        col = (v < left) ? v : v + (c_hi - c_lo);
    }
    // blocks are laid out over (left + right) columns contiguously
    const bool valid = (c0 + cl) < (left + (n - c_hi));
    for (int e = tid >> 6; e < cnt; e += blockDim.x >> 6)
        tile[e][cl] = valid ? A[s_src[e] * lda + col] : T(0);
    __syncthreads();
    if (valid)
        for (int e = tid >> 6; e < cnt; e += blockDim.x >> 6)
            A[s_dst[e] * lda + col] = tile[e][cl];
}

// ------------------------------------------------------------ unit TRSM
// B (nb x ncols, in place, row-major) <- L^-1 B with L the nb x nb unit lower
// triangle at Lm (ld ldl), nb <= 64.  A workgroup stages 64 columns of B and
// the triangle in LDS and sweeps the columns of L: after step t, row t of
// the block is final and rows below it have had row t's contribution
// removed.
template <typename T>
__global__ __launch_bounds__(256)
void k_trsm_unit_lower(const T *Lm, int64_t ldl, int nb, T *B, int64_t ldb, int ncols)
{
    __shared__ T Ls[64][65];
    __shared__ T Bs[64][65];
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * 64;
    const int cl = tid & 63, rq = tid >> 6;
    for (int e = tid; e < 64 * 64; e += 256) {
        int i = e >> 6, t = e & 63;
        Ls[i][t] = (i < nb && t < i) ? Lm[(int64_t)i * ldl + t] : T(0);
    }
    for (int i = rq; i < 64; i += 4)
        Bs[i][cl] = (i < nb && c0 + cl < ncols) ? B[(int64_t)i * ldb + c0 + cl] : T(0);
    __syncthreads();
    for (int t = 0; t < nb - 1; ++t) {
        const T xt = Bs[t][cl];
        for (int i = t + 1 + ((rq - (t + 1)) & 3); i < nb; i += 4) Bs[i][cl] -= Ls[i][t] * xt;
        __syncthreads();
    }
    if (c0 + cl < ncols)
        for (int i = rq; i < nb; i += 4) B[(int64_t)i * ldb + c0 + cl] = Bs[i][cl];
}

// sign = parity of the number of real swaps in ipiv (lu.rs:140-152 flips the
// sign once per p != j).
__global__ __launch_bounds__(256)
void k_pivsign(const int64_t *ipiv, int64_t kmax, int32_t *info)
{
    __shared__ int s[256];
    int c = 0;
    for (int64_t i = threadIdx.x; i < kmax; i += blockDim.x) c += (ipiv[i] != i);
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) info[0] = (s[0] & 1) ? 0 : 1;
}

// ------------------------------------------------------------------ driver
struct PanelWS {
    uint64_t *slots;
    void *rowbuf;
    void *diagbuf;
    int64_t *swapmap;
};


constexpr size_t PANEL_LDS_MAX = 150 * 1024;

template <typename T>
static int launch_panel(la_ctx *ctx, T *Apanel, int64_t lda, int M, int NB, int kglob,
                        int64_t *d_ipiv, const PanelWS &pw, int32_t *d_status, uint64_t ticks)
{
    static bool attr_local[2] = {false, false};
    const int ai0 = sizeof(T) == 8 ? 0 : 1;
    const size_t local_bytes = (size_t)M * (NB + 1) * sizeof(T);
    if (local_bytes <= 136 * 1024) {
        if (!attr_local[ai0]) {
            hipFuncSetAttribute((const void *)k_getrf_panel_local<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)PANEL_LDS_MAX);
            attr_local[ai0] = true;
        }
        hipLaunchKernelGGL(k_getrf_panel_local<T>, dim3(1), dim3(256), local_bytes, ctx->stream,
                           Apanel, lda, M, NB, kglob, d_ipiv, pw.swapmap);
        LA_HIP_CHECK(ctx, hipGetLastError());
        return LA_OK;
    }
    int maxp = std::min(PANEL_MAXP, ctx->num_cus > 0 ? ctx->num_cus : 256);
    int P = (int)std::min<int64_t>(maxp, ceil_div(M, 32));
    if (P < 1) P = 1;
    int R = (int)ceil_div(M, P);
    P = (int)ceil_div(M, R);
    size_t shmem = ((size_t)R * (NB + 1) + 2 * NB) * sizeof(T);
    if (shmem > 150 * 1024)
        return set_error(ctx, LA_ERR_ARG, "getrf panel: rows per workgroup exceed LDS");
    static bool attr_set[2] = {false, false};
    const int ai = sizeof(T) == 8 ? 0 : 1;
    if (!attr_set[ai]) {
        hipFuncSetAttribute((const void *)k_getrf_panel<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        attr_set[ai] = true;
    }
    T *rowbuf = reinterpret_cast<T *>(pw.rowbuf);
    T *diagbuf = reinterpret_cast<T *>(pw.diagbuf);
    int64_t *swapmap = pw.swapmap;
    uint64_t *slots = pw.slots;
    void *args[] = {&Apanel, &lda, &M, &NB, &R, &kglob, &d_ipiv, &slots, &rowbuf,
                    &diagbuf, &swapmap, &d_status, &ticks};
    LA_HIP_CHECK(ctx, hipLaunchCooperativeKernel((const void *)k_getrf_panel<T>, dim3(P),
                                                 dim3(256), args, (unsigned)shmem,
                                                 ctx->stream));
    return LA_OK;
}

template <typename T>
static int launch_rowmoves(la_ctx *ctx, T *A, int64_t lda, int64_t n, int64_t c_lo,
                           int64_t c_hi, const int64_t *swapmap, uint64_t *perm)
{
    int64_t ncols = c_lo + (n - c_hi);
    int nblk = (int)ceil_div(ncols, 64);
    // one extra block permutes the perm vector
    hipLaunchKernelGGL(k_apply_rowmoves<T>, dim3(nblk + 1), dim3(256), 0, ctx->stream, A,
                       lda, (int)n, (int)c_lo, (int)c_hi, swapmap, perm);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template <typename T>
static int launch_trsm(la_ctx *ctx, const T *L, int64_t ldl, int nb, T *B, int64_t ldb,
                       int64_t ncols)
{
    if (ncols <= 0) return LA_OK;
    hipLaunchKernelGGL(k_trsm_unit_lower<T>, dim3((unsigned)ceil_div(ncols, 64)), dim3(256), 0,
                       ctx->stream, L, ldl, nb, B, ldb, (int)ncols);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

// In-place LU.  d_ipiv receives LAPACK-style swaps (global rows), d_perm the
// permutation vector (may be null), d_status LA_OK / LA_ERR_TIMEOUT.
template <typename T>
int getrf_core(la_ctx *ctx, int64_t m, int64_t n, T *A, int64_t lda, int64_t *d_ipiv,
               uint64_t *d_perm, int32_t *d_status)
{
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2)
        return set_error(ctx, LA_ERR_ARG, "getrf: extent too large");
    const int64_t kmax = m < n ? m : n;
    // panel workspace
    size_t slot_bytes = (size_t)2 * PANEL_MAXP * 2 * sizeof(uint64_t);
    size_t row_bytes = (size_t)2 * PANEL_MAXP * NB_IN * sizeof(T);
    size_t diag_bytes = (size_t)2 * NB_IN * sizeof(T);
    size_t map_bytes = (size_t)(1 + 4 * NB_IN) * sizeof(int64_t);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t total = al(slot_bytes) + al(row_bytes) + al(diag_bytes) + al(map_bytes);
    void *base = nullptr;
    int st = ws_get(ctx, WS_PANEL, total, &base);
    if (st) return st;
    PanelWS pw;
    char *b = (char *)base;
    pw.slots = (uint64_t *)b; b += al(slot_bytes);
    pw.rowbuf = b; b += al(row_bytes);
    pw.diagbuf = b; b += al(diag_bytes);
    pw.swapmap = (int64_t *)b;

    // reset slots: step tag 0xffffffff never occurs
    const int64_t nslot_words = 2 * PANEL_MAXP * 2;
    if ((st = fill_u64(ctx, pw.slots, nslot_words, 0xffffffff00000000ull))) return st;
    if ((st = fill_i32(ctx, d_status, 1, 0))) return st;
    if (d_perm && (st = iota_u64(ctx, d_perm, m))) return st;
    const uint64_t ticks = spin_limit_ticks(ctx);

    for (int64_t K0 = 0; K0 < kmax; K0 += NB_OUT) {
        const int64_t nbo = std::min<int64_t>(NB_OUT, kmax - K0);
        for (int64_t k = K0; k < K0 + nbo; k += NB_IN) {
            const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - k);
            T *Ap = A + k * lda + k;
            if ((st = launch_panel<T>(ctx, Ap, lda, (int)(m - k), nbi, (int)k, d_ipiv, pw,
                                      d_status, ticks)))
                return st;
            if ((st = launch_rowmoves<T>(ctx, A, lda, n, k, k + nbi, pw.swapmap, d_perm)))
                return st;
            const int64_t rest = K0 + nbo - (k + nbi);
            if (rest > 0) {
                if ((st = launch_trsm<T>(ctx, Ap, lda, nbi, A + k * lda + k + nbi, lda, rest)))
                    return st;
                if (m - k - nbi > 0)
                    if ((st = gemm_rm<T>(ctx, m - k - nbi, rest, nbi, T(-1),
                                         A + (k + nbi) * lda + k, lda, A + k * lda + k + nbi,
                                         lda, T(1), A + (k + nbi) * lda + k + nbi, lda, false,
                                         false)))
                        return st;
            }
        }
        const int64_t ncol = n - (K0 + nbo);
        if (ncol > 0) {
            T *B12 = A + K0 * lda + K0 + nbo;
            for (int64_t kb = K0; kb < K0 + nbo; kb += NB_IN) {
                const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - kb);
                if ((st = launch_trsm<T>(ctx, A + kb * lda + kb, lda, nbi,
                                         A + kb * lda + K0 + nbo, lda, ncol)))
                    return st;
                const int64_t below = K0 + nbo - (kb + nbi);
                if (below > 0)
                    if ((st = gemm_rm<T>(ctx, below, ncol, nbi, T(-1), A + (kb + nbi) * lda + kb,
                                         lda, A + kb * lda + K0 + nbo, lda, T(1),
                                         A + (kb + nbi) * lda + K0 + nbo, lda, false, false)))
                        return st;
            }
            if (K0 + nbo < m)
                if ((st = gemm_rm<T>(ctx, m - K0 - nbo, ncol, nbo, T(-1),
                                     A + (K0 + nbo) * lda + K0, lda, B12, lda, T(1),
                                     A + (K0 + nbo) * lda + K0 + nbo, lda, false, false)))
                    return st;
        }
    }
    return LA_OK;
}

int pivsign(la_ctx *ctx, const int64_t *d_ipiv, int64_t kmax, int32_t *d_info)
{
    hipLaunchKernelGGL(k_pivsign, dim3(1), dim3(256), 0, ctx->stream, d_ipiv, kmax, d_info);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template int getrf_core<double>(la_ctx *, int64_t, int64_t, double *, int64_t, int64_t *,
                                uint64_t *, int32_t *);
template int getrf_core<float>(la_ctx *, int64_t, int64_t, float *, int64_t, int64_t *,
                               uint64_t *, int32_t *);

}  // namespace lahip
