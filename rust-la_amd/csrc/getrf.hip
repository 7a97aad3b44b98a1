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
//     u12 + gemm         : U12 = L11^-1 A12 (one kernel per block),
//                          A22 -= L21 U12 (MFMA, K = NB_OUT)
// Once the remaining rows fit the CUs reserved for panels (the panel
// stream's), outer blocks overlap: block j+1's panels run on those CUs while
// the other CUs apply block j (look-ahead; see getrf_core).
#include "la_common.hpp"
#include <mutex>

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

// Maximum of an unsigned value over the wavefront, by DPP row shifts and
// row broadcasts (no LDS round trip): lane 63 ends with the maximum, which
// every lane receives.
__device__ inline uint32_t wave_max_u32(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Wavefront argmax under better(): largest 64-bit key, lowest index among
// equal keys; every lane returns the winner.  Three DPP maxima (high key
// word, low key word among the lanes at the high maximum, complemented
// index among the lanes at the full maximum) instead of a shuffle tree.
__device__ inline void wave_argmax(uint64_t &key, int &idx) {
    const uint32_t hi = (uint32_t)(key >> 32), lo = (uint32_t)key;
    const uint32_t mhi = wave_max_u32(hi);
    const bool at_hi = hi == mhi;
    const uint32_t mlo = wave_max_u32(at_hi ? lo : 0u);
    const bool at_max = at_hi && lo == mlo;
    const uint32_t nidx = wave_max_u32(at_max ? ~(uint32_t)idx : 0u);
    key = ((uint64_t)mhi << 32) | mlo;
    idx = (int)~nidx;
}

// Value of row r of a wave's rows (lane r % 64, register r / 64) broadcast
// to every lane; r is the same in every lane.
__device__ inline float bcast_lane(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ inline double bcast_lane(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename T, int RPL>
__device__ inline T lane_value(const T (&v)[RPL], int r) {
    const int q = r >> 6;
    T x = v[0];
#pragma unroll
    for (int i = 1; i < RPL; ++i)
        if (q == i) x = v[i];
    return bcast_lane(x, r & 63);
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


// X = L^-1 for the unit lower nb x nb triangle L (LDS, row stride ldl; only
// the strict lower part is read), by ONE wavefront with no barrier: lane c
// forms column c, X[i][c] = -sum_{t=c}^{i-1} L[i][t] X[t][c], X[c][c] = 1.
// The sum is kept as four interleaved partial sums over t = c+q, c+q+4, ...
// (q = 0..3) combined as (p0 + p1) + (p2 + p3): terms with t < c are exact
// zeros, so the accumulators indexed by t mod 4 are the same partial sums
// rotated by c mod 4.  Every L[i][t] is an LDS broadcast.  out is
// NB_IN x NB_IN, zero outside the nb x nb triangle.
template <typename T>
__device__ void invert_unit_lower_wave(const T *L, int ldl, int nb, T *out)
{
    const int c = threadIdx.x & 63, cm = c & 3;
    T x[NB_IN];
#pragma unroll
    for (int i = 0; i < NB_IN; ++i) {
        T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int t = 0; t < i; ++t) acc[t & 3] += L[i * ldl + t] * x[t];
        // part q sums t = c+q (mod 4): accumulator (q + c) mod 4
        const T p0 = cm == 0 ? acc[0] : cm == 1 ? acc[1] : cm == 2 ? acc[2] : acc[3];
        const T p1 = cm == 0 ? acc[1] : cm == 1 ? acc[2] : cm == 2 ? acc[3] : acc[0];
        const T p2 = cm == 0 ? acc[2] : cm == 1 ? acc[3] : cm == 2 ? acc[0] : acc[1];
        const T p3 = cm == 0 ? acc[3] : cm == 1 ? acc[0] : cm == 2 ? acc[1] : acc[2];
        x[i] = (i == c) ? T(1) : (i > c ? -((p0 + p1) + (p2 + p3)) : T(0));
    }
#pragma unroll
    for (int i = 0; i < NB_IN; ++i) out[i * NB_IN + c] = (i < nb && c < nb) ? x[i] : T(0);
}

// ------------------------------------------------------------- panel kernel
// Factors the M x NB panel whose top-left element is A (row-major, lda),
// global column/row offset kglob, with P cooperative workgroups; workgroup
// w owns panel rows [w*R, min(w*R+R, M)) in LDS for the whole kernel and
// its first wavefront owns the per-column critical path (R <= 64*RPL).
//
// Exchange: per column j every workgroup publishes its best candidate
// (ordered |value| key, row index, sign) and that row's NB values, and the
// owner of row j publishes row j.  Every published 64-bit word carries a
// tag of the global step number gs, so a reader validates each word on its
// own and never needs the writer to order its stores: no store-completion
// wait and no separate flag.  Value words carry gs itself in their upper
// 32 bits; the key word has room for only ONE tag bit, bit 1 of gs.
// Doubles travel as two words (low and high halves), floats as one.
// Buffers are double-buffered by step parity (bit 0 of gs): a workgroup
// reaches its step-(s+2) publication only after seeing every workgroup's
// step-(s+1) slot, which each publishes after it finished reading step s.
//
// Why one key bit suffices: when a reader polls slot (buffer gs & 1,
// workgroup q) for step gs, that slot holds either step gs or the step it
// last validated there, gs - 2, whose bit 1 differs.  This needs (a) every
// reader to poll the same slots at every step, (b) every panel of a
// factorisation to write the last two steps of each slot a later panel
// reads: panels are NB_IN = 64 columns wide except the last, and when the
// number of workgroups P grows from one panel to the next, launch_panel
// resets the slots first, and (c) a reset (tag 0xffffffff in every value
// word and bit 63 set in every key word, at the start of a factorisation
// and before a panel whose P grew) to be followed by a panel starting at a
// step kglob = 0 mod 4, whose first two steps (bit 1 clear) reject it.
//
// Per column, once the winner p is known: rows j and p are swapped, then the
// first wavefront forms the multipliers (true division, skipped for a zero
// pivot), updates column j+1, picks the next local candidate, brings that
// row (and row j+1 if owned) fully up to date and publishes step j+1 before
// the whole workgroup applies the rest of the rank-1 update, so the bulk of
// the arithmetic overlaps the exchange latency.
template <typename T> struct PanelWords;
template <> struct PanelWords<double> { static constexpr int W = 2; };
template <> struct PanelWords<float> { static constexpr int W = 1; };
constexpr int CAND_WORDS = 2 + 2 * NB_IN;        // slot words + row words

__device__ inline void put_value(uint64_t *w, uint64_t tag32, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    st_dev(w + 0, tag32 | (b & 0xffffffffull));
    st_dev(w + 1, tag32 | (b >> 32));
}
__device__ inline void put_value(uint64_t *w, uint64_t tag32, float v) {
    st_dev(w, tag32 | (uint64_t)__float_as_uint(v));
}

#ifdef LAHIP_PANEL_TIMING
// Phase clock of workgroup 0, thread 0 (wall-clock ticks summed over
// columns, kept in registers and added once at the end of the kernel):
// [0] header poll, [1] winner-row fetch, [2] wait for the bulk update,
// [3] swap, [4] multipliers and column j+1, [5] next local candidate,
// [6] candidate-row update, [7] publication
__device__ unsigned long long g_panel_phase[8];
#define PHASE_MARK(k)                                                          \
    do {                                                                       \
        if (w == 0 && tid == 0) {                                              \
            const uint64_t now_ = wall_clock64();                              \
            ph_acc[k] += now_ - ph_t;                                          \
            ph_t = now_;                                                       \
        }                                                                      \
    } while (0)
#else
#define PHASE_MARK(k) do {} while (0)
#endif

// The rows x NB block at src (row stride lds) into LDS at dst (row stride
// ldb), and back: sixteen loads per thread are issued before any is
// stored, so the block costs one memory round trip per sixteen elements a
// thread, not one per element.
template <typename T>
__device__ inline void stage_block(T *dst, int ldb, const T *src, int64_t lds, int rows, int NB)
{
    const int total = rows * NB, nt = (int)blockDim.x;
    for (int e0 = (int)threadIdx.x; e0 < total; e0 += 16 * nt) {
        T v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = min(e0 + k * nt, total - 1), r = e / NB, c = e - r * NB;
            v[k] = src[(int64_t)r * lds + c];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = e0 + k * nt, r = e / NB, c = e - r * NB;
            if (e < total) dst[r * ldb + c] = v[k];
        }
    }
}

template <typename T>
__device__ inline void unstage_block(T *dst, int64_t ldd, const T *src, int ldb, int rows, int NB)
{
    const int total = rows * NB, nt = (int)blockDim.x;
    for (int e0 = (int)threadIdx.x; e0 < total; e0 += 16 * nt) {
        T v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = min(e0 + k * nt, total - 1), r = e / NB, c = e - r * NB;
            v[k] = src[r * ldb + c];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = e0 + k * nt, r = e / NB, c = e - r * NB;
            if (e < total) dst[(int64_t)r * ldd + c] = v[k];
        }
    }
}

template <typename T, int RPL>
__global__ __launch_bounds__(256)
void k_getrf_panel(T *A, int64_t lda, int M, int NB, int R, int kglob, int64_t *ipiv,
                   uint64_t *cand, uint64_t *diagw, int64_t *swapmap, int32_t *status,
                   uint64_t spin_ticks, T *linv)
{
    constexpr int W = PanelWords<T>::W;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int LDB = NB + 1;
    T *blk = reinterpret_cast<T *>(smem_raw);
    // winner row (raw halves), double-buffered by step parity: the bulk
    // update of step j-1 reads one buffer while step j's row lands in the other
    __shared__ __align__(16) uint32_t s_urow[2][W * NB_IN];
    __shared__ __align__(16) uint32_t s_drow[W * NB_IN];   // old row j (raw halves)
    __shared__ uint64_t s_key[4];
    __shared__ int s_idx[4];
    __shared__ int s_abort;
    __shared__ int s_piv[NB_IN];
    __shared__ int s_skip[2];
    const T *drow = reinterpret_cast<const T *>(s_drow);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int w = blockIdx.x;
    const int P = gridDim.x;
    const int rowbase = w * R;
    const int nrows = min(R, M - rowbase);

    stage_block(blk, LDB, A + (int64_t)rowbase * lda, lda, nrows, NB);
    if (tid == 0) s_abort = 0;
    __syncthreads();

    // publish the candidate of step s (wave 0 only): rows >= s of column s
    auto publish = [&](int s, uint64_t key, int idx, bool neg) {
        const uint32_t gs = (uint32_t)(kglob + s);
        const int buf = gs & 1;
        const uint64_t tag32 = (uint64_t)gs << 32;
        uint64_t *slot = cand + ((size_t)buf * PANEL_MAXP + w) * CAND_WORDS;
        if (idx != IDX_NONE) {
            const int rl = idx - rowbase;
            for (int c = lane; c < NB; c += 64) put_value(slot + 2 + W * c, tag32, blk[rl * LDB + c]);
        }
        if (s >= rowbase && s < rowbase + nrows) {
            const int rl = s - rowbase;
            uint64_t *dw = diagw + (size_t)buf * W * NB_IN;
            for (int c = lane; c < NB; c += 64) put_value(dw + W * c, tag32, blk[rl * LDB + c]);
        }
        if (lane == 0) {
            const uint64_t tagb = ((gs >> 1) & 1) ? TAG_BIT : 0;
            st_dev(slot + 0, tagb | key);
            st_dev(slot + 1, tag32 | (neg ? 0x80000000ull : 0ull) | (uint32_t)idx);
        }
    };
    // publish the candidate of step s from registers: lane c holds column c
    // of the candidate row (vc) and of row s (vd, used if this workgroup
    // owns row s)
    auto publish_row = [&](int s, uint64_t key, int idx, bool neg, T vc, T vd) {
        const uint32_t gs = (uint32_t)(kglob + s);
        const int buf = gs & 1;
        const uint64_t tag32 = (uint64_t)gs << 32;
        uint64_t *slot = cand + ((size_t)buf * PANEL_MAXP + w) * CAND_WORDS;
        if (lane < NB) {
            if (idx != IDX_NONE) put_value(slot + 2 + W * lane, tag32, vc);
            if (s >= rowbase && s < rowbase + nrows)
                put_value(diagw + (size_t)buf * W * NB_IN + W * lane, tag32, vd);
        }
        if (lane == 0) {
            const uint64_t tagb = ((gs >> 1) & 1) ? TAG_BIT : 0;
            st_dev(slot + 0, tagb | key);
            st_dev(slot + 1, tag32 | (neg ? 0x80000000ull : 0ull) | (uint32_t)idx);
        }
    };
    // local candidate of column s over rows >= s (wave 0, RPL rows per lane)
    auto local_cand = [&](int s, uint64_t &key, int &idx, bool &neg) {
        key = 0; idx = IDX_NONE;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int r = lane + 64 * q;
            const int gr = rowbase + r;
            if (r < nrows && gr >= s) {
                const T v = blk[r * LDB + s];
                uint64_t k = abs_key(v);
                if (gr == s && v != v) k = KEY_MAX;
                if (better(k, gr, key, idx)) { key = k; idx = gr; }
            }
        }
        wave_argmax(key, idx);
        neg = false;
        if (idx != IDX_NONE) neg = blk[(idx - rowbase) * LDB + s] < T(0);
    };

    if (wave == 0) {
        uint64_t key; int idx; bool neg;
        local_cand(0, key, idx, neg);
        publish(0, key, idx, neg);
    }
    int pend_j = -1;          // step whose bulk update is still pending
#ifdef LAHIP_PANEL_TIMING
    uint64_t ph_t = wall_clock64();
    uint64_t ph_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#endif
    for (int j = 0; j < NB; ++j) {
        const uint32_t gs = (uint32_t)(kglob + j);
        const int buf = gs & 1;
        const uint64_t tagb = ((gs >> 1) & 1) ? TAG_BIT : 0;
        const T *urow = reinterpret_cast<const T *>(s_urow[j & 1]);

        if (wave == 0) {
            // 1. gather the P candidates of step j (lane q polls workgroup q)
            uint64_t ck = 0;
            int ci = IDX_NONE;
            int abort = 0;
            // lane q polls workgroups q, q+64, ... (P <= 256): every slot's
            // words are loaded before any is checked, and only slots not yet
            // current are polled again, so a pass costs one round trip
            {
                constexpr int SPL = PANEL_MAXP / 64;     // slots per lane
                uint64_t w0[SPL], w1[SPL];
                uint32_t pend = 0;
#pragma unroll
                for (int i = 0; i < SPL; ++i)
                    if (lane + 64 * i < P) pend |= 1u << i;
                const uint64_t t0 = wall_clock64();
                for (uint32_t spin = 0; pend; ++spin) {
#pragma unroll
                    for (int i = 0; i < SPL; ++i)
                        if ((pend >> i) & 1) {
                            const uint64_t *slot =
                                cand + ((size_t)buf * PANEL_MAXP + lane + 64 * i) * CAND_WORDS;
                            w1[i] = ld_dev(slot + 1);
                            w0[i] = ld_dev(slot + 0);
                        }
#pragma unroll
                    for (int i = 0; i < SPL; ++i)
                        if (((pend >> i) & 1) && (uint32_t)(w1[i] >> 32) == gs &&
                            (w0[i] & TAG_BIT) == tagb) {
                            pend &= ~(1u << i);
                            const uint64_t k = w0[i] & ~TAG_BIT;
                            const int ix = (int)((uint32_t)w1[i] & 0x7fffffffu);
                            if (better(k, ix, ck, ci)) { ck = k; ci = ix; }
                        }
                    if (!pend) break;
                    if ((spin & 31) == 31 && ld_status(status) != 0) { abort = 1; break; }
                    if (wall_clock64() - t0 > spin_ticks) {
                        atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
                        abort = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            wave_argmax(ck, ci);
            abort = __any(abort);
            PHASE_MARK(0);
            int p = ci;
            if (!abort && (p < j || p >= M)) {
                if (lane == 0) atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
                p = j;
            }
            // 2. fetch the winner row (and old row j if this workgroup owns p)
            if (!abort) {
                const int wp = p / R;
                const bool need_d = (p != j) && (p >= rowbase && p < rowbase + nrows);
                const uint64_t *wsrc = cand + ((size_t)buf * PANEL_MAXP + wp) * CAND_WORDS + 2;
                const uint64_t *dsrc = diagw + (size_t)buf * W * NB_IN;
                // all of a lane's words (up to 4) are loaded before any is
                // checked, so they cost one round trip, not one each
                const int nw = W * NB * (need_d ? 2 : 1);
                uint64_t v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int wi = lane + 64 * i;
                    if (wi < nw) {
                        const bool isd = wi >= W * NB;
                        v[i] = ld_dev((isd ? dsrc : wsrc) + (isd ? wi - W * NB : wi));
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int wi = lane + 64 * i;
                    if (wi >= nw) continue;
                    const bool isd = wi >= W * NB;
                    const int k = isd ? wi - W * NB : wi;
                    if ((uint32_t)(v[i] >> 32) != gs) {
                        const uint64_t *src = (isd ? dsrc : wsrc) + k;
                        const uint64_t t0 = wall_clock64();
                        for (uint32_t spin = 0;; ++spin) {
                            v[i] = ld_dev(src);
                            if ((uint32_t)(v[i] >> 32) == gs) break;
                            if ((spin & 31) == 31 && ld_status(status) != 0) break;
                            if (wall_clock64() - t0 > spin_ticks) {
                                atomicCAS((int *)status, 0, (int)LA_ERR_TIMEOUT);
                                break;
                            }
                        }
                    }
                    (isd ? s_drow : s_urow[j & 1])[k] = (uint32_t)v[i];
                }
            }
            if (lane == 0) {
                s_abort = abort;
                s_piv[j] = p;
                if (w == 0) ipiv[kglob + j] = (int64_t)kglob + p;
            }
            PHASE_MARK(1);
            // 3. swap rows j and p inside the panel, here rather than after
            //    the barrier: waves 1-3 never write those two rows in the
            //    meantime (they are the rows the critical wave of step j - 1
            //    brought up to date itself), so the swap costs no barrier
            if (!abort && p != j) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const bool own_j = (j >= rowbase && j < rowbase + nrows);
                const bool own_p = (p >= rowbase && p < rowbase + nrows);
                if (lane < NB) {
                    if (own_j) blk[(j - rowbase) * LDB + lane] = urow[lane];
                    if (own_p) blk[(p - rowbase) * LDB + lane] = drow[lane];
                }
            }
            PHASE_MARK(3);
        } else if (pend_j >= 0) {
            // 1'. meanwhile waves 1-3 finish step pend_j's rank-1 update:
            //     columns pend_j+2 .. NB-1 of every row below pend_j except
            //     the rows the critical wave already brought up to date
            const int jj = pend_j;
            const T *uprev = reinterpret_cast<const T *>(s_urow[jj & 1]);
            const int sk0 = s_skip[0], sk1 = s_skip[1];
            const int r_first = max(0, jj + 1 - rowbase);
            // each wave takes a contiguous third of the rows; loads are
            // issued eight rows at a time ahead of the stores so the LDS
            // latency is paid once per batch, not once per row
            const int nr = nrows - r_first;
            const int per = (nr + 2) / 3;
            const int rb0 = r_first + (wave - 1) * per;
            const int rb1 = min(nrows, rb0 + per);
            for (int cc = jj + 2 + lane; cc < NB; cc += 64) {
                const T u = uprev[cc];
                int r = rb0;
                for (; r + 8 <= rb1; r += 8) {
                    T l[8], a[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        l[k] = blk[(r + k) * LDB + jj];
                        a[k] = blk[(r + k) * LDB + cc];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (r + k != sk0 && r + k != sk1) blk[(r + k) * LDB + cc] = a[k] - l[k] * u;
                }
                for (; r < rb1; ++r)
                    if (r != sk0 && r != sk1) blk[r * LDB + cc] -= blk[r * LDB + jj] * u;
            }
        }
        __syncthreads();
        PHASE_MARK(2);
        if (s_abort) return;

        // 4. critical path (wave 0): multipliers, column j+1, next candidate.
        //    This panel never runs for matrices small enough to be factored
        //    in one piece, so the multipliers use one reciprocal per column.
        //    The new column j+1 stays in registers for the candidate search,
        //    the candidate's sign and the multipliers of the two rows brought
        //    up to date come from their lanes by readlane, and the updated
        //    rows are published from registers: two LDS round trips in all.
        const T piv = urow[j];
        const int j1 = j + 1;
        if (wave == 0) {
            const T rpiv = (piv != T(0)) ? T(1) / piv : T(1);
            const T u1 = (j1 < NB) ? urow[j1] : T(0);
            const T uc = urow[min(lane, NB - 1)];       // column lane of the winner row
            T lv[RPL], av[RPL];
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int r = min(lane + 64 * q, nrows - 1);
                lv[q] = blk[r * LDB + j];
                av[q] = (j1 < NB) ? blk[r * LDB + j1] : T(0);
            }
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int r = lane + 64 * q;
                if (r < nrows && rowbase + r > j) {
                    const T l = lv[q] * rpiv;
                    lv[q] = l;
                    blk[r * LDB + j] = l;
                    if (j1 < NB) {
                        av[q] = av[q] - l * u1;
                        blk[r * LDB + j1] = av[q];
                    }
                }
            }
            PHASE_MARK(4);
            if (j1 < NB) {
                // local candidate of column j1 (rows >= j1) from the registers
                uint64_t key = 0;
                int idx = IDX_NONE;
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int r = lane + 64 * q, gr = rowbase + r;
                    if (r < nrows && gr >= j1) {
                        uint64_t k = abs_key(av[q]);
                        if (gr == j1 && av[q] != av[q]) k = KEY_MAX;
                        if (better(k, gr, key, idx)) { key = k; idx = gr; }
                    }
                }
                wave_argmax(key, idx);
                const int rc = (idx != IDX_NONE) ? idx - rowbase : -1;
                const int rd = (j1 >= rowbase && j1 < rowbase + nrows && j1 != idx) ? j1 - rowbase : -1;
                bool neg = false;
                T lrc = T(0), lrd = T(0);
                if (rc >= 0) {
                    neg = lane_value(av, rc) < T(0);
                    lrc = lane_value(lv, rc);
                }
                if (rd >= 0) lrd = lane_value(lv, rd);
                PHASE_MARK(5);
                // bring the candidate row and row j1 fully up to date (their
                // columns j and j1 were written above by their own lanes)
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                T vc = T(0), vd = T(0);
                if (lane < NB) {
                    if (rc >= 0) vc = blk[rc * LDB + lane];
                    if (rd >= 0) vd = blk[rd * LDB + lane];
                    if (lane > j1) {
                        if (rc >= 0) { vc -= lrc * uc; blk[rc * LDB + lane] = vc; }
                        if (rd >= 0) { vd -= lrd * uc; blk[rd * LDB + lane] = vd; }
                    }
                }
                PHASE_MARK(6);
                publish_row(j1, key, idx, neg, vc, rd >= 0 ? vd : vc);
                if (lane == 0) { s_skip[0] = rc; s_skip[1] = rd; }
            }
            PHASE_MARK(7);
        }
        __syncthreads();
        pend_j = (j1 + 1 < NB) ? j : -1;
    }

#ifdef LAHIP_PANEL_TIMING
    if (w == 0 && tid == 0)
        for (int k = 0; k < 8; ++k) atomicAdd(&g_panel_phase[k], (unsigned long long)ph_acc[k]);
#endif
    unstage_block(A + (int64_t)rowbase * lda, lda, blk, LDB, nrows, NB);
    __syncthreads();
    if (w == 0 && tid < 64) panel_net_moves(NB, kglob, s_piv, swapmap);
    // L11 (rows 0..NB-1 of the panel, all owned by workgroup 0 since R >= NB)
    if (w == 0 && wave == 1) invert_unit_lower_wave<T>(blk, LDB, NB, linv);
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
                         int64_t *swapmap, T *linv)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int LDB = NB + 1;
    T *blk = reinterpret_cast<T *>(smem_raw);
    __shared__ uint64_t s_key[4];
    __shared__ int s_idx[4];
    __shared__ int s_piv[NB_IN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    stage_block(blk, LDB, A, lda, M, NB);
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
    unstage_block(A, lda, blk, LDB, M, NB);
    __syncthreads();
    if (tid < 64) panel_net_moves(NB, kglob, s_piv, swapmap);
    if (wave == 1) invert_unit_lower_wave<T>(blk, LDB, NB, linv);
}

// ------------------------------------------------------------ row swaps
// Applies the net row movement of nmaps consecutive panels (swap maps
// map_stride words apart, in panel order) to columns [0, c_lo) and
// [c_hi, n) (the panels' own columns are already final) and to the
// permutation vector.  Each workgroup handles 64 columns: per map it first
// gathers every source row of its columns into LDS, then scatters them to
// their destinations, so cycles in a permutation need no ordering; the maps
// follow one another behind a barrier.
template <typename T>
__global__ __launch_bounds__(256)
void k_apply_rowmoves(T *A, int64_t lda, int n, int c_lo, int c_hi,
                      const int64_t *swapmap, int nmaps, int64_t map_stride, uint64_t *perm)
{
    __shared__ T tile[2 * NB_IN][64];
    __shared__ int64_t s_dst[2 * NB_IN], s_src[2 * NB_IN];
    __shared__ int s_cnt;
    const int tid = threadIdx.x;
    const bool perm_block = blockIdx.x == gridDim.x - 1 && perm;
    // map this block's 64 columns, skipping the panel columns [c_lo, c_hi);
    // blocks are laid out over (left + right) columns contiguously
    const int cl = tid & 63;
    const int v = blockIdx.x * 64 + cl;
    const int col = v < c_lo ? v : v + (c_hi - c_lo);
    const bool valid = v < c_lo + (n - c_hi);
    for (int mi = 0; mi < nmaps; ++mi) {
        const int64_t *map = swapmap + (size_t)mi * map_stride;
        __syncthreads();                // the previous map's scatter is done
        if (tid == 0) s_cnt = (int)map[0];
        __syncthreads();
        const int cnt = s_cnt;          // <= 2 * NB_IN <= blockDim.x
        if (cnt == 0) continue;
        for (int e = tid; e < cnt; e += blockDim.x) {
            s_dst[e] = map[1 + 2 * e];
            s_src[e] = map[2 + 2 * e];
        }
        __syncthreads();
        if (perm_block) {
            // permutation vector: perm_new[dst] = perm_old[src]
            uint64_t pv = 0;
            if (tid < cnt) pv = perm[s_src[tid]];
            __syncthreads();
            if (tid < cnt) perm[s_dst[tid]] = pv;
            continue;
        }
        // eight rows per wave in flight: loads from clamped positions (a
        // lane past the last column reads column 0, never stored), issued
        // before any is stored
        for (int e0 = tid >> 6; e0 < cnt; e0 += 8 * (blockDim.x >> 6)) {
            T rv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = min(e0 + k * (int)(blockDim.x >> 6), cnt - 1);
                rv[k] = A[s_src[e] * lda + (valid ? col : 0)];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = e0 + k * (int)(blockDim.x >> 6);
                if (e < cnt) tile[e][cl] = rv[k];
            }
        }
        __syncthreads();
        if (valid)
            for (int e = tid >> 6; e < cnt; e += blockDim.x >> 6)
                A[s_dst[e] * lda + col] = tile[e][cl];
    }
}

// ------------------------------------------------------------ unit TRSM
// B (nb x ncols, in place, row-major) <- L^-1 B, given Linv = L^-1 (NB_IN x
// NB_IN, unit diagonal stored, zero above it, produced by the panel kernel).
// Linv and 64 columns of B staged in LDS; thread (column cl, row group rq)
// keeps 16 independent accumulators (rows rq + 4k), so the inner loop is a
// stream of LDS broadcasts and FMAs with no dependent chain; each row sums
// B[i] + Linv[i][t] B[t] for t ascending below the diagonal.
template <typename T>
__global__ __launch_bounds__(256)
void k_trsm_unit_lower(const T *Linv, int nb, T *B, int64_t ldb, int ncols)
{
    // rows of Li are 16-byte aligned so four consecutive t are one vector read
    constexpr int LIS = 64 + 16 / (int)sizeof(T);
    constexpr int VW = 16 / (int)sizeof(T);           // elements per 16-byte read
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    __shared__ __align__(16) T Li[64][LIS];
    __shared__ T Bs[64][65];
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * 64;
    const int cl = tid & 63, rq = tid >> 6;
    // staged with a zero diagonal: row i then sums B[i] + Linv[i][t] B[t]
    // over t ascending, the t >= i terms adding exact zeros.  All 32 loads
    // per thread are unconditional (clamped positions) and issue together;
    // out-of-range elements are replaced by zero.
    T lv[16], bv[16];
    const int cc = min(c0 + cl, ncols - 1);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, i = e >> 6, t = e & 63;
        lv[k] = Linv[min(i, nb - 1) * NB_IN + min(t, nb - 1)];
        bv[k] = B[(int64_t)min(rq + 4 * k, nb - 1) * ldb + cc];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, i = e >> 6, t = e & 63;
        Li[i][t] = (i < nb && t < nb && t != i) ? lv[k] : T(0);
        const int ib = rq + 4 * k;
        Bs[ib][cl] = (ib < nb && c0 + cl < ncols) ? bv[k] : T(0);
    }
    __syncthreads();
    // row i = rq + 4k: s = B[i] + sum_{t<i} Linv[i][t] B[t], t ascending
    T acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = Bs[rq + 4 * k][cl];
#pragma unroll 2
    for (int t0 = 0; t0 < 64; t0 += VW) {
        T b[VW];
#pragma unroll
        for (int u = 0; u < VW; ++u) b[u] = Bs[t0 + u][cl];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const vec_t lv = *reinterpret_cast<const vec_t *>(&Li[rq + 4 * k][t0]);
#pragma unroll
            for (int u = 0; u < VW; ++u) acc[k] += lv[u] * b[u];
        }
    }
    if (c0 + cl >= ncols) return;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = rq + 4 * k;
        if (i < nb) B[(int64_t)i * ldb + c0 + cl] = acc[k];
    }
}

// ------------------------------------------------------ block U12 = L11^-1 A12
// U12 for an outer block of nbo <= NB_OUT rows in one pass.  Each workgroup
// keeps CW columns of A12 (nbo x CW) in LDS and runs the block forward
// substitution: for each 64-row panel b, X_b <- Linv_b X_b (the panel
// kernel's inverse, unit diagonal implied; same operation order as
// k_trsm_unit_lower), then X_r -= L_rb X_b for every panel r below it, with
// L_rb read from the factored block (Lblk = A + K0 * lda + K0).  Replaces
// per block four TRSM and three GEMM launches.  CW = 8 (FP32) or 16 (FP64).
template <typename T, int CW>
__global__ __launch_bounds__(256)
void k_u12_block(const T *__restrict__ Lblk, int64_t lda, const T *__restrict__ linv, int nbo,
                 T *B, int64_t ldb, int ncols)
{
    constexpr int RQ = 256 / CW;            // row groups
    constexpr int KR = 64 / RQ;             // rows per thread per 64-row panel
    constexpr int XS = CW + 1, LS = 64 + 1;
    extern __shared__ __align__(16) unsigned char u12_smem[];
    T *X = reinterpret_cast<T *>(u12_smem);           // NB_OUT x XS
    T *Ls = X + NB_OUT * XS;                          // 2 x 64 x LS
    const int tid = threadIdx.x, cl = tid % CW, rq = tid / CW;
    const int c0 = blockIdx.x * CW;
    const int col = min(c0 + cl, ncols - 1);
    const int nb = (nbo + 63) / 64;
    // stages in order: for each panel b, its inverse (diagonal stage), then
    // L_rb for r = b+1 .. nb-1; stage s + 1's 64 x 64 operand is loaded into
    // registers while stage s computes, so the loads overlap the FMAs
    const int nstage = nb + nb * (nb - 1) / 2;
    auto stage_br = [&](int st, int &b, int &r) {
        b = 0;
        while (st >= nb - b) { st -= nb - b; ++b; }
        r = b + st;                         // r == b: diagonal stage
    };
    T pre[16];
    auto load_stage = [&](int st) {
        int b, r;
        stage_br(st, b, r);
        const int nbb = min(64, nbo - 64 * b), nbr = min(64, nbo - 64 * r);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k, i = e >> 6, t = e & 63;
            pre[k] = r == b ? linv[(size_t)b * NB_IN * NB_IN + min(i, nbb - 1) * NB_IN + min(t, nbb - 1)]
                            : Lblk[(int64_t)(64 * r + min(i, nbr - 1)) * lda + 64 * b + min(t, nbb - 1)];
        }
    };
    auto store_stage = [&](int st, T *dst) {
        int b, r;
        stage_br(st, b, r);
        const int nbb = min(64, nbo - 64 * b), nbr = min(64, nbo - 64 * r);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k, i = e >> 6, t = e & 63;
            const bool in = r == b ? (i < nbb && t < nbb && t != i) : (i < nbr && t < nbb);
            dst[i * LS + t] = in ? pre[k] : T(0);
        }
    };
    load_stage(0);
    if constexpr (sizeof(T) == 8) {
        // FP64: this thread's NB_OUT / RQ rows of the block, all loads in
        // flight before the first store (measured: 1% off the FP64 LU; the
        // same change slows the FP32 instance, which keeps the loop)
        constexpr int NX = NB_OUT / RQ;
        T xv[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) xv[k] = B[(int64_t)min(rq + RQ * k, nbo - 1) * ldb + col];
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int i = rq + RQ * k;
            X[i * XS + cl] = i < nbo ? xv[k] : T(0);
        }
    } else {
        for (int i = rq; i < NB_OUT; i += RQ) {
            const T v = B[(int64_t)min(i, nbo - 1) * ldb + col];
            X[i * XS + cl] = i < nbo ? v : T(0);
        }
    }
    store_stage(0, Ls);
    for (int st = 0; st < nstage; ++st) {
        T *cur = Ls + (st & 1) * 64 * LS;
        __syncthreads();                    // stage operand and X rows in place
        if (st + 1 < nstage) load_stage(st + 1);
        int b, r;
        stage_br(st, b, r);
        T acc[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) acc[k] = r == b ? X[(64 * b + rq + RQ * k) * XS + cl] : T(0);
#pragma unroll 4
        for (int t = 0; t < 64; ++t) {
            const T xb = X[(64 * b + t) * XS + cl];
#pragma unroll
            for (int k = 0; k < KR; ++k) acc[k] += cur[(rq + RQ * k) * LS + t] * xb;
        }
        if (r == b) {
            __syncthreads();                // every read of X_b is done
#pragma unroll
            for (int k = 0; k < KR; ++k) X[(64 * b + rq + RQ * k) * XS + cl] = acc[k];
        } else {
#pragma unroll
            for (int k = 0; k < KR; ++k) X[(64 * r + rq + RQ * k) * XS + cl] -= acc[k];
        }
        if (st + 1 < nstage) store_stage(st + 1, Ls + ((st + 1) & 1) * 64 * LS);
    }
    __syncthreads();
    if (c0 + cl < ncols)
        for (int i = rq; i < nbo; i += RQ) B[(int64_t)i * ldb + c0 + cl] = X[i * XS + cl];
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
    int last_p = 1 << 30;       // workgroups of the previous cooperative panel
};


constexpr size_t PANEL_LDS_MAX = 150 * 1024;

template <typename T>
static const void *panel_kernel(int rpl)
{
    return rpl == 1 ? (const void *)k_getrf_panel<T, 1>
         : rpl == 2 ? (const void *)k_getrf_panel<T, 2>
         : rpl == 3 ? (const void *)k_getrf_panel<T, 3>
                    : (const void *)k_getrf_panel<T, 4>;
}

// Rows per lane for a panel of M rows confined to `cus` CUs (an ordinary
// launch on a CU-masked stream: every workgroup must be resident at once, so
// the grid may not exceed cus x the kernel's occupancy at the largest LDS
// block), or 0 when no row count fits.  Aims at `wpc` workgroups per CU
// (more, shorter workgroups shorten the per-column row work; more than a
// few per CU contend for the CU), then more rows, then fewer.
template <typename T>
static int masked_panel_rpl(int M, int cus)
{
    const int rpl_max = sizeof(T) == 8 ? 3 : 4;
    static int occ[2][5] = {{-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}};
    static std::mutex mu;
    const int ai = sizeof(T) == 8 ? 0 : 1;
    auto fits = [&](int rpl) {
        const int R = 64 * rpl;
        const int64_t P = ceil_div((int64_t)M, (int64_t)R);
        int o;
        {
            std::lock_guard<std::mutex> g(mu);
            if (occ[ai][rpl] < 0) {
                const void *fn = panel_kernel<T>(rpl);
                const size_t shmem = (size_t)R * (NB_IN + 1) * sizeof(T);
                int v = 0;
                if (shmem > allow_dynamic_lds(fn) ||
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, fn, 256, shmem) != hipSuccess)
                    v = 0;
                (void)hipGetLastError();
                occ[ai][rpl] = v;
            }
            o = occ[ai][rpl];
        }
        return o > 0 && P <= std::min<int64_t>(PANEL_MAXP, (int64_t)cus * o);
    };
    const int wpc = 2;
    const int r0 = (int)std::min<int64_t>(
        rpl_max, std::max<int64_t>(1, ceil_div((int64_t)M, (int64_t)64 * cus * wpc)));
    for (int r = r0; r <= rpl_max; ++r)
        if (fits(r)) return r;
    for (int r = r0 - 1; r >= 1; --r)
        if (fits(r)) return r;
    return 0;
}

// mask_cus > 0: the panel runs on a stream whose CU mask holds mask_cus CUs
// (masked_panel_rpl must have accepted M); 0: on the whole device, at most
// one workgroup per CU.
template <typename T>
static int launch_panel(la_ctx *ctx, T *Apanel, int64_t lda, int M, int NB, int kglob,
                        int64_t *d_ipiv, PanelWS &pw, int64_t *swapmap, int32_t *d_status,
                        uint64_t ticks, T *linv, int mask_cus = 0)
{
    // algorithmic traffic: the panel is read and written once
    ProfScope ps(ctx, LA_K_PANEL, (double)M * NB * NB, 2.0 * M * NB * sizeof(T));
    static size_t local_max[2] = {0, 0};
    const int ai0 = sizeof(T) == 8 ? 0 : 1;
    if (!local_max[ai0]) local_max[ai0] = allow_dynamic_lds((const void *)k_getrf_panel_local<T>);
    const size_t local_bytes = (size_t)M * (NB + 1) * sizeof(T);
    if (local_bytes <= std::min<size_t>(96 * 1024, local_max[ai0])) {
        hipLaunchKernelGGL(k_getrf_panel_local<T>, dim3(1), dim3(256), local_bytes, ctx->stream,
                           Apanel, lda, M, NB, kglob, d_ipiv, swapmap, linv);
        LA_HIP_CHECK(ctx, hipGetLastError());
        return LA_OK;
    }
    // More, shorter workgroups shorten the per-column critical path (the
    // rank-1 update and the candidate search scale with the rows each group
    // holds), so a panel prefers 64 rows per workgroup (FP32) or 128 (FP64)
    // and grows the rows per group only when more than maxp groups would be
    // needed; the LDS holds at most 4 (FP32) or 3 (FP64) lane-rows.
    const int maxp = std::min(PANEL_MAXP, ctx->num_cus > 0 ? ctx->num_cus : 256);
    const int rpl_max = sizeof(T) == 8 ? 3 : 4;
    const int rpl_pref = sizeof(T) == 8 ? 2 : 1;
    const int rpl_fit = (int)std::max<int64_t>(1, ceil_div(M, (int64_t)64 * maxp));
    const int rpl = mask_cus > 0 ? masked_panel_rpl<T>(M, mask_cus)
                                 : std::min(rpl_max, std::max(rpl_pref, std::min(rpl_fit, rpl_max)));
    if (rpl < 1) return set_error(ctx, LA_ERR_ARG, "getrf: panel does not fit the masked CUs");
    int R = 64 * rpl;
    const int P = (int)ceil_div(M, R);
    if (mask_cus == 0 && P > maxp)
        return set_error(ctx, LA_ERR_ARG,
                         "getrf: more rows than the cooperative panel holds (m > 64 * 3 * "
                         "min(256, CUs) for FP64, 64 * 4 * min(256, CUs) for FP32)");
    // the key word's single tag bit relies on every slot read by this panel
    // having been written by the previous one (see the exchange comment of
    // k_getrf_panel); when P grows, the slots are reset instead, which the
    // first two steps reject as long as kglob is a multiple of 4
    if (P > pw.last_p) {
        if (kglob % 4 != 0) return set_error(ctx, LA_ERR_ARG, "getrf panel: misaligned reset");
        int st = fill_u64(ctx, pw.slots, (int64_t)2 * PANEL_MAXP * CAND_WORDS,
                          0xffffffff00000000ull);
        if (st) return st;
    }
    pw.last_p = P;
    size_t shmem = (size_t)R * (NB + 1) * sizeof(T);
    static size_t coop_max[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    const int ai = sizeof(T) == 8 ? 0 : 1;
    const void *fn = panel_kernel<T>(rpl);
    if (!coop_max[ai][rpl - 1]) coop_max[ai][rpl - 1] = allow_dynamic_lds(fn);
    if (shmem > coop_max[ai][rpl - 1])
        return set_error(ctx, LA_ERR_ARG, "getrf panel: rows per workgroup exceed LDS");
    uint64_t *candp = pw.slots;
    uint64_t *diagp = reinterpret_cast<uint64_t *>(pw.diagbuf);
    void *args[] = {&Apanel, &lda, &M, &NB, &R, &kglob, &d_ipiv, &candp, &diagp,
                    &swapmap, &d_status, &ticks, &linv};
    // An ordinary launch in both cases.  The grid is co-resident by
    // construction (P <= min(256, CUs) at one workgroup per CU or, masked,
    // P <= the masked CUs x occupancy; dispatch is balanced over XCDs and
    // shader engines), and every wait in the kernel is bounded.  The
    // cooperative-launch API is not used: on this ROCm one process at a time
    // can hold its device-wide queue, and another process's cooperative
    // launch then fails ("operation cannot be performed in the present
    // state"), so two processes factoring on one GPU would break each other.
    LA_HIP_CHECK(ctx, hipLaunchKernel(fn, dim3(P), dim3(256), args, shmem, ctx->stream));
    return LA_OK;
}

template <typename T>
static int launch_rowmoves(la_ctx *ctx, T *A, int64_t lda, int64_t n, int64_t c_lo,
                           int64_t c_hi, const int64_t *swapmap, uint64_t *perm,
                           int nmaps = 1, int64_t map_stride = 0)
{
    int64_t ncols = c_lo + (n - c_hi);
    if (ncols <= 0 && !perm) return LA_OK;
    int nblk = (int)ceil_div(ncols, 64);
    // up to 2*NB_IN rows move per map: read + write of each moved row segment
    ProfScope ps(ctx, LA_K_ROWMOVE, 0.0, 2.0 * 2 * NB_IN * ncols * sizeof(T) * nmaps);
    // one extra block permutes the perm vector
    hipLaunchKernelGGL(k_apply_rowmoves<T>, dim3(nblk + 1), dim3(256), 0, ctx->stream, A,
                       lda, (int)n, (int)c_lo, (int)c_hi, swapmap, nmaps, map_stride, perm);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template <typename T>
static int launch_trsm(la_ctx *ctx, const T *Linv, int nb, T *B, int64_t ldb, int64_t ncols)
{
    if (ncols <= 0) return LA_OK;
    ProfScope ps(ctx, LA_K_TRSM, (double)nb * nb * ncols,
                 (double)sizeof(T) * (2.0 * nb * ncols + (double)nb * nb));
    hipLaunchKernelGGL(k_trsm_unit_lower<T>, dim3((unsigned)ceil_div(ncols, 64)), dim3(256), 0,
                       ctx->stream, Linv, nb, B, ldb, (int)ncols);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template <typename T, int CW>
static int launch_u12_cw(la_ctx *ctx, const T *Lblk, int64_t lda, const T *linv, int nbo,
                         T *B, int64_t ldb, int64_t ncols)
{
    const void *fn = (const void *)k_u12_block<T, CW>;
    const size_t shmem = ((size_t)NB_OUT * (CW + 1) + 2 * 64 * 65) * sizeof(T);
    static size_t lm = 0;
    if (!lm) lm = allow_dynamic_lds(fn);
    if (shmem > lm) return set_error(ctx, LA_ERR_ARG, "getrf: U12 block exceeds LDS");
    ProfScope ps(ctx, LA_K_TRSM, (double)nbo * nbo * ncols,
                 (double)sizeof(T) * (2.0 * nbo * ncols + (double)nbo * nbo));
    hipLaunchKernelGGL((k_u12_block<T, CW>), dim3((unsigned)ceil_div(ncols, (int64_t)CW)),
                       dim3(256), shmem, ctx->stream, Lblk, lda, linv, nbo, B, ldb, (int)ncols);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return LA_OK;
}

template <typename T>
static int launch_u12_block(la_ctx *ctx, const T *Lblk, int64_t lda, const T *linv, int nbo,
                            T *B, int64_t ldb, int64_t ncols)
{
    // columns per workgroup, measured at n = 16384 with look-ahead (FP32-IR
    // 92.0 / 92.4 / 93.2 ms, FP64 LU 119.0 / 114.9 / 116.8 ms for 8 / 16 / 32):
    // narrow workgroups shorten the U12 of the next block's 256 columns,
    // which the next panel waits for
    if constexpr (sizeof(T) == 4)
        return launch_u12_cw<T, 8>(ctx, Lblk, lda, linv, nbo, B, ldb, ncols);
    else
        return launch_u12_cw<T, 16>(ctx, Lblk, lda, linv, nbo, B, ldb, ncols);
}

// Trailing updates: native MFMA for T, or FP32 data with BF16/FP16-rounded
// operands on the low-precision MFMA (upd_prec = LA_PREC_BF16 / LA_PREC_FP16).
template <typename T>
static int trail_gemm(la_ctx *ctx, int upd_prec, int64_t M, int64_t N, int64_t K, const T *A,
                      int64_t lda, const T *B, int64_t ldb, T *C, int64_t ldc)
{
    if constexpr (sizeof(T) == 4) {
        if (upd_prec == LA_PREC_BF16 || upd_prec == LA_PREC_FP16)
            return gemm_lp(ctx, upd_prec, M, N, K, -1.f, A, lda, B, ldb, 1.f, C, ldc);
    }
    return gemm_rm<T>(ctx, M, N, K, T(-1), A, lda, B, ldb, T(1), C, ldc, false, false);
}

// Panel workspace of getrf_core: exchange slots, row buffer, diagonal-row
// words, swap map and the ring of L11^-1 blocks, each 256-byte aligned.
constexpr size_t ws_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t panel_ws_bytes(size_t elem)
{
    return ws_align((size_t)2 * PANEL_MAXP * CAND_WORDS * sizeof(uint64_t)) + ws_align(256) +
           ws_align((size_t)2 * 2 * NB_IN * sizeof(uint64_t)) +
           ws_align((size_t)(1 + 4 * NB_IN) * sizeof(int64_t)) +
           ws_align((size_t)(NB_OUT / NB_IN) * NB_IN * NB_IN * elem);
}

// ------------------------------------------------ look-ahead streams
// Two process-wide streams per device with complementary CU masks: the
// panel stream owns `cus_per_xcd` CUs of every XCD, the update stream the
// rest.  On MI355X mask bit i selects a CU of XCD i % 8 and, within it, of
// shader engine (i / 8) % 4, so bits [0, 8 * cus_per_xcd) spread the panel
// evenly over XCDs and, when cus_per_xcd is a multiple of 4, over the four
// shader engines of each XCD.  That evenness is required, not cosmetic: an
// XCD deals workgroups to its shader engines in turn, so with an uneven
// split (say 3, 3, 2, 2 CUs) the smaller engines overflow and part of the
// grid waits for CUs the rest of it never frees.  The runtime also refuses
// a mask that leaves an XCD with no CU.
// Cooperative launches ignore CU masks, which is why the masked panel is an
// ordinary launch sized by masked_panel_rpl.  The streams live until the
// process ends (the runtime allows only a few masked hardware queues).
struct LookAhead {
    hipStream_t panel = nullptr, update = nullptr;
    int panel_cus = 0;
    bool tried = false;
};

static const LookAhead *lookahead_streams(la_ctx *ctx)
{
    static std::mutex mu;
    static LookAhead la[16];
    const char *off = getenv("LAHIP_LOOKAHEAD");
    if (off && atoi(off) == 0) return nullptr;
    if (ctx->device < 0 || ctx->device >= 16 || ctx->num_cus < 16 || ctx->num_cus % 8 != 0)
        return nullptr;
    std::lock_guard<std::mutex> g(mu);
    LookAhead &l = la[ctx->device];
    if (!l.tried) {
        l.tried = true;
        int per_xcd = 12;
        if (const char *e = getenv("LAHIP_PANEL_CUS_PER_XCD")) per_xcd = atoi(e);
        per_xcd -= per_xcd % 4;
        const int xcd_cus = ctx->num_cus / 8;
        if (per_xcd < 4 || per_xcd >= xcd_cus || xcd_cus % 4 != 0) return nullptr;
        const int words = (ctx->num_cus + 31) / 32;
        std::vector<uint32_t> mp(words, 0), mu_(words, 0);
        for (int i = 0; i < ctx->num_cus; ++i)
            (i < 8 * per_xcd ? mp : mu_)[i / 32] |= 1u << (i % 32);
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return nullptr;
        if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess) return nullptr;
        hipStream_t sp = nullptr, su = nullptr;
        bool ok = hipExtStreamCreateWithCUMask(&sp, (uint32_t)words, mp.data()) == hipSuccess &&
                  hipExtStreamCreateWithCUMask(&su, (uint32_t)words, mu_.data()) == hipSuccess;
        (void)hipGetLastError();
        if (cur != ctx->device) hipSetDevice(cur);
        if (!ok) return nullptr;
        l.panel = sp;
        l.update = su;
        l.panel_cus = 8 * per_xcd;
    }
    return l.panel ? &l : nullptr;
}

// Restores ctx->stream when the look-ahead schedule returns.
struct StreamSwap {
    la_ctx *ctx;
    hipStream_t saved;
    explicit StreamSwap(la_ctx *c) : ctx(c), saved(c->stream) {}
    void use(hipStream_t s) { ctx->stream = s; }
    ~StreamSwap() { ctx->stream = saved; }
};

// In-place LU.  d_ipiv receives LAPACK-style swaps (global rows), d_perm the
// permutation vector (may be null), d_status LA_OK / LA_ERR_TIMEOUT.
// ext_ws: caller-provided panel workspace of panel_ws_bytes(sizeof(T))
// bytes, or null for the context's grow-only workspace.
//
// Outer blocks run one of two schedules.  The plain schedule factors a block
// and then updates everything to its right, all on ctx->stream.  Once the
// remaining rows fit the panel stream's CUs, blocks run with look-ahead: the
// panel stream factors block j+1 (panels, their row moves inside the block,
// the in-block updates) while the update stream applies block j to the
// columns right of block j+1.  Per block j the update stream does, in order:
// block j's row moves, U12 = L11^-1 A12 and update on block j+1's columns
// (after which block j+1 may start), then block j's row moves on the columns
// left of it and right of block j+1 (and the permutation vector), U12 and
// the update there; while the updates are large, the panel stream takes the
// rightmost columns of that update once block j+1 is factored (the share).
// Buffers that block j's update still reads while block j+1 is factored
// (swap maps, L11^-1 blocks) alternate between two copies by block parity.
// Every location is written by one stream between two event waits, and row
// moves on a column keep their block order, so the factor is identical to
// the plain schedule's.
template <typename T>
int getrf_core(la_ctx *ctx, int64_t m, int64_t n, T *A, int64_t lda, int64_t *d_ipiv,
               uint64_t *d_perm, int32_t *d_status, int upd_prec, void *ext_ws)
{
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2)
        return set_error(ctx, LA_ERR_ARG, "getrf: extent too large");
    const int64_t kmax = m < n ? m : n;
    // outer block from which the look-ahead schedule runs (kmax: never)
    const LookAhead *la = ext_ws ? nullptr : lookahead_streams(ctx);
    int64_t K0la = kmax;
    if (la)
        for (int64_t K0 = 0; K0 + NB_OUT < kmax && K0 + NB_OUT < n; K0 += NB_OUT)
            if (masked_panel_rpl<T>((int)(m - K0), la->panel_cus) > 0) { K0la = K0; break; }
    // panel workspace
    size_t slot_bytes = (size_t)2 * PANEL_MAXP * CAND_WORDS * sizeof(uint64_t);
    size_t row_bytes = 256;
    size_t diag_bytes = (size_t)2 * 2 * NB_IN * sizeof(uint64_t);
    size_t map_bytes = (size_t)(1 + 4 * NB_IN) * sizeof(int64_t);
    constexpr int RING = 2 * (NB_OUT / NB_IN);           // swap maps of two blocks
    auto al = [](size_t x) { return ws_align(x); };
    void *base = ext_ws;
    int st;
    // look-ahead: swap maps and L11^-1 blocks of two outer blocks, since the
    // update stream still applies block j (row moves, U12) while the panel
    // stream factors block j+1
    const size_t linv_bytes = al((size_t)(NB_OUT / NB_IN) * NB_IN * NB_IN * sizeof(T));
    const size_t ring_bytes = K0la < kmax ? RING * al(map_bytes) + linv_bytes : 0;
    if (!base && (st = ws_get(ctx, WS_PANEL, panel_ws_bytes(sizeof(T)) + ring_bytes, &base)))
        return st;
    PanelWS pw;
    char *b = (char *)base;
    pw.slots = (uint64_t *)b; b += al(slot_bytes);
    pw.rowbuf = b; b += al(row_bytes);
    pw.diagbuf = b; b += al(diag_bytes);
    pw.swapmap = (int64_t *)b; b += al(map_bytes);
    T *linv_ring = (T *)b;
    b += linv_bytes;
    int64_t *swap_ring = (int64_t *)b;
    const size_t ring_stride = al(map_bytes) / sizeof(int64_t);
    T *linv_ring2 = (T *)(b + RING * al(map_bytes));      // look-ahead only

    // step tag 0xffffffff never occurs: no stale word can validate
    const int64_t nslot_words = 2 * PANEL_MAXP * CAND_WORDS;
    if ((st = fill_u64(ctx, pw.slots, nslot_words, 0xffffffff00000000ull))) return st;
    if ((st = fill_u64(ctx, (uint64_t *)pw.diagbuf, 2 * 2 * NB_IN, 0xffffffff00000000ull)))
        return st;
    if ((st = fill_i32(ctx, d_status, 1, 0))) return st;
    if (d_perm && (st = iota_u64(ctx, d_perm, m))) return st;
    const uint64_t ticks = spin_limit_ticks(ctx);

    // Factors the nbo columns of block K0 (panels, their row moves on the
    // block's other columns, in-block updates).  outside_moves: also apply
    // each panel's moves to the columns outside the block and to perm.
    auto factor_block = [&](int64_t K0, int64_t nbo, int64_t *maps, size_t map_stride,
                            T *linvs, int mask_cus, bool outside_moves) -> int {
        for (int64_t k = K0; k < K0 + nbo; k += NB_IN) {
            const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - k);
            T *Ap = A + k * lda + k;
            T *linv = linvs + (size_t)((k - K0) / NB_IN) * NB_IN * NB_IN;
            int64_t *smap = maps + (size_t)((k - K0) / NB_IN) * map_stride;
            int s;
            if ((s = launch_panel<T>(ctx, Ap, lda, (int)(m - k), nbi, (int)k, d_ipiv, pw, smap,
                                     d_status, ticks, linv, mask_cus)))
                return s;
            if (outside_moves) {
                if ((s = launch_rowmoves<T>(ctx, A, lda, n, k, k + nbi, smap, d_perm))) return s;
            } else if (nbo > nbi) {
                if ((s = launch_rowmoves<T>(ctx, A + K0, lda, nbo, k - K0, k - K0 + nbi, smap,
                                            nullptr)))
                    return s;
            }
            const int64_t rest = K0 + nbo - (k + nbi);
            if (rest > 0) {
                if ((s = launch_trsm<T>(ctx, linv, nbi, A + k * lda + k + nbi, lda, rest)))
                    return s;
                if (m - k - nbi > 0)
                    if ((s = trail_gemm<T>(ctx, upd_prec, m - k - nbi, rest, nbi,
                                           A + (k + nbi) * lda + k, lda, A + k * lda + k + nbi,
                                           lda, A + (k + nbi) * lda + k + nbi, lda)))
                        return s;
            }
        }
        return LA_OK;
    };
    // U12 = L11^-1 A12 for block K0 on columns [c0, c0 + ncol) (the block's rows).
    auto block_u12 = [&](int64_t K0, int64_t nbo, int64_t c0, int64_t ncol,
                         const T *linvs) -> int {
        if (ncol <= 0) return LA_OK;
        if (upd_prec != LA_PREC_BF16 && upd_prec != LA_PREC_FP16)
            return launch_u12_block<T>(ctx, A + K0 * lda + K0, lda, linvs, (int)nbo,
                                       A + K0 * lda + c0, lda, ncol);
        for (int64_t kb = K0; kb < K0 + nbo; kb += NB_IN) {
            const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - kb);
            const T *linv_kb = linvs + (size_t)((kb - K0) / NB_IN) * NB_IN * NB_IN;
            int s;
            if ((s = launch_trsm<T>(ctx, linv_kb, nbi, A + kb * lda + c0, lda, ncol))) return s;
            const int64_t below = K0 + nbo - (kb + nbi);
            if (below > 0)
                if ((s = trail_gemm<T>(ctx, upd_prec, below, ncol, nbi, A + (kb + nbi) * lda + kb,
                                       lda, A + kb * lda + c0, lda, A + (kb + nbi) * lda + c0,
                                       lda)))
                    return s;
        }
        return LA_OK;
    };
    // A22 -= L21 U12 for block K0 on columns [c0, c0 + ncol).
    auto block_gemm = [&](int64_t K0, int64_t nbo, int64_t c0, int64_t ncol) -> int {
        if (ncol <= 0 || K0 + nbo >= m) return LA_OK;
        return trail_gemm<T>(ctx, upd_prec, m - K0 - nbo, ncol, nbo, A + (K0 + nbo) * lda + K0,
                             lda, A + K0 * lda + c0, lda, A + (K0 + nbo) * lda + c0, lda);
    };

    // plain schedule
    for (int64_t K0 = 0; K0 < K0la; K0 += NB_OUT) {
        const int64_t nbo = std::min<int64_t>(NB_OUT, kmax - K0);
        if ((st = factor_block(K0, nbo, pw.swapmap, 0, linv_ring, 0, true))) return st;
        if (n - (K0 + nbo) > 0) {
            if ((st = block_u12(K0, nbo, K0 + nbo, n - (K0 + nbo), linv_ring))) return st;
            if ((st = block_gemm(K0, nbo, K0 + nbo, n - (K0 + nbo)))) return st;
        }
    }
    if (K0la >= kmax) return LA_OK;

    // look-ahead schedule
    if (!ctx->la_ev[0])
        for (auto &e : ctx->la_ev)
            LA_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipEvent_t ev_start = ctx->la_ev[0], ev_pan = ctx->la_ev[1], ev_la = ctx->la_ev[2],
               ev_end = ctx->la_ev[3], ev_u12 = ctx->la_ev[4], ev_share = ctx->la_ev[5];
    const hipStream_t home = ctx->stream;
    StreamSwap sw(ctx);
    LA_HIP_CHECK(ctx, hipEventRecord(ev_start, home));
    LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->panel, ev_start, 0));
    LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->update, ev_start, 0));
    // Share of the update handed to the panel stream.  While the updates are
    // large the panel stream finishes block j+1 long before the update
    // stream finishes block j's update, so it takes the rightmost columns of
    // that update once block j+1 is factored.  The share is sized from a
    // rate model (GEMM FLOP/s per CU at K = 256 and panel time per column,
    // both measured on MI355X at n = 16384) so that both streams finish
    // together (scaling the modelled share by 0.5 or 1.5 was slower); it never includes the columns of
    // blocks j+1 and j+2, which the update stream needs first.  Fencing: the
    // panel stream's share waits for block j's U12 there (ev_u12); the update
    // stream waits for the share (ev_share) before block j+1's row moves,
    // which rewrite rows of block j's L21 that the share reads.
    const double gemm_tf_per_cu = sizeof(T) == 8 ? 0.20e12 : 0.354e12;
    const double panel_s_per_col = sizeof(T) == 8 ? 4.9e-6 : 4.3e-6;
    const double r_upd = gemm_tf_per_cu * (ctx->num_cus - la->panel_cus);
    const double r_pan = gemm_tf_per_cu * la->panel_cus;
    struct Share { int64_t K0 = 0, nbo = 0, c0 = 0, ncol = 0; } share;   // pending (block j-1)
    bool share_pending = false;
    int parity = 0;
    for (int64_t K0 = K0la; K0 < kmax; K0 += NB_OUT, parity ^= 1) {
        const int64_t nbo = std::min<int64_t>(NB_OUT, kmax - K0);
        int64_t *maps = swap_ring + (size_t)parity * (NB_OUT / NB_IN) * ring_stride;
        T *linvs = parity ? linv_ring2 : linv_ring;
        const int npan = (int)ceil_div(nbo, (int64_t)NB_IN);
        sw.use(la->panel);
        if (K0 > K0la) LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->panel, ev_la, 0));
        if ((st = factor_block(K0, nbo, maps, ring_stride, linvs, la->panel_cus, false)))
            return st;
        LA_HIP_CHECK(ctx, hipEventRecord(ev_pan, la->panel));
        const bool shared_prev = share_pending;
        if (share_pending) {
            LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->panel, ev_u12, 0));
            if ((st = block_gemm(share.K0, share.nbo, share.c0, share.ncol))) return st;
            LA_HIP_CHECK(ctx, hipEventRecord(ev_share, la->panel));
            share_pending = false;
        }

        // update stream: first what block j+1 needs (block j's row moves,
        // U12 and update on block j+1's columns [c1, c1 + nla)), then block
        // j's moves on the columns left of it and right of block j+1 (one
        // launch, with the permutation vector), U12 and the update there
        sw.use(la->update);
        LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->update, ev_pan, 0));
        const int64_t c1 = K0 + nbo, ncol = n - c1;
        const int64_t nla = std::min<int64_t>(NB_OUT, std::max<int64_t>(ncol, 0));
        if (nla > 0) {
            if ((st = launch_rowmoves<T>(ctx, A + c1, lda, nla, 0, 0, maps, nullptr, npan,
                                         (int64_t)ring_stride)))
                return st;
            if ((st = block_u12(K0, nbo, c1, nla, linvs))) return st;
            if ((st = block_gemm(K0, nbo, c1, nla))) return st;
        }
        LA_HIP_CHECK(ctx, hipEventRecord(ev_la, la->update));
        if (shared_prev) LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->update, ev_share, 0));
        if ((st = launch_rowmoves<T>(ctx, A, lda, n, K0, c1 + nla, maps, d_perm, npan,
                                     (int64_t)ring_stride)))
            return st;
        if ((st = block_u12(K0, nbo, c1 + nla, ncol - nla, linvs))) return st;
        const int64_t nbulk = ncol - nla, rows = m - c1;
        int64_t nshare = 0;
        if (nbulk > 2 * NB_OUT && rows > 0 && K0 + 2 * NB_OUT < kmax) {
            const double w = 2.0 * rows * nbulk * nbo;
            const double t_next = panel_s_per_col * std::min<int64_t>(NB_OUT, kmax - c1);
            const double ws = (w / r_upd - t_next) / (1.0 / r_pan + 1.0 / r_upd);
            if (ws > 0) nshare = (int64_t)(ws / (2.0 * rows * nbo)) / 128 * 128;
            nshare = std::min<int64_t>(nshare, nbulk - NB_OUT);
            if (nshare < 256) nshare = 0;
        }
        if (nshare > 0) {
            LA_HIP_CHECK(ctx, hipEventRecord(ev_u12, la->update));
            share = Share{K0, nbo, n - nshare, nshare};
            share_pending = true;
        }
        if ((st = block_gemm(K0, nbo, c1 + nla, nbulk - nshare))) return st;
    }
    if (share_pending) {            // not reached: the last two blocks never share
        sw.use(la->panel);
        LA_HIP_CHECK(ctx, hipStreamWaitEvent(la->panel, ev_u12, 0));
        if ((st = block_gemm(share.K0, share.nbo, share.c0, share.ncol))) return st;
    }
    LA_HIP_CHECK(ctx, hipEventRecord(ev_pan, la->panel));
    LA_HIP_CHECK(ctx, hipEventRecord(ev_end, la->update));
    LA_HIP_CHECK(ctx, hipStreamWaitEvent(home, ev_pan, 0));
    LA_HIP_CHECK(ctx, hipStreamWaitEvent(home, ev_end, 0));
    return LA_OK;
}

// ------------------------------------------------ LU without pivoting
// For the randomised (butterfly-transformed) factor of the refinement
// solver, whose pivots need not follow partial pivoting: the panel becomes
// a one-wavefront factorisation of the diagonal block plus an independent
// product per 64-row block, with no exchange between workgroups.

// Diagonal block (nb x nb at A, nb <= NB_IN) factored in place without
// pivoting, in FP64, by one workgroup (padding rows and columns act as an
// identity block).  A pivot smaller than tau in magnitude is replaced by
// +-tau (static pivoting; the refinement corrects the perturbation) and
// counted in *npert.  Also writes linv = L^-1 (unit lower) and uinv = U^-1,
// NB_IN x NB_IN, zero outside the nb x nb triangles, formed in FP64 from the
// FP64 factors.
//
// Factorisation: lane r of wave g holds the sixteen elements [r][16g ..
// 16g+15] in registers.  Per pair of steps j, j+1 (one barrier; LDS
// buffers double-buffered by pair parity; unrolled over the sixteen columns
// of a register block so that column j's register index is static) the
// wave holding columns j, j+1 publishes them, lanes j and j+1 of every wave
// publish their row segments, and every
// element takes a[k] -= l_r * u_c with l_r = 0 for rows r <= j: one FMA
// and no select.  Elements left of the diagonal of rows below it become
// don't-cares; the multipliers l_r and the pivots are logged in LDS and
// rebuild L and U's diagonal at the end.  Inverses: 2 x 2 blocked, the
// four 32 x 32 diagonal-block inverses one per wave, then two small
// products for the off-diagonal blocks.
__global__ __launch_bounds__(256)
void k_diag_nopiv(float *A, int64_t lda, int nb, const float *tau_p, float *linv, float *uinv,
                  int32_t *npert)
{
    typedef double d2_t __attribute__((ext_vector_type(2)));
    constexpr int LS = NB_IN + 2;                        // 16-byte aligned rows
    __shared__ __align__(16) double LT[NB_IN][LS];       // LT[t][i] = L[i][t] (i > t)
    __shared__ __align__(16) double UT[NB_IN][LS];       // UT[t][i] = U[i][t] (i <= t)
    __shared__ __align__(16) double prow[2][2][NB_IN];   // rows j, j+1, published by lanes j, j+1
    __shared__ double pcol[2][2][NB_IN];                 // columns j, j+1, published by wave gj
    __shared__ double piv[NB_IN], rpiv[NB_IN];
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const double tau = fmax((double)*tau_p, 1e-30);
    double a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        a[k] = (double)A[(int64_t)min(r, nb - 1) * lda + min(16 * g + k, nb - 1)];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int c = 16 * g + k;
        if (!(r < nb && c < nb)) a[k] = (r == c) ? 1.0 : 0.0;
    }
    for (int gj = 0; gj < NB_IN / 16; ++gj) {
#pragma unroll
        for (int kj = 0; kj < 16; kj += 2) {
            // steps j and j+1 share one barrier: rows and columns j, j+1 are
            // published as they stand, and every thread forms step j's
            // update of row and column j+1 itself (the same operations, in
            // the same order, as two single steps)
            const int j = 16 * gj + kj, buf = (kj >> 1) & 1;
            if (g == gj) {
                pcol[buf][0][r] = a[kj];
                pcol[buf][1][r] = a[kj + 1];
            }
            if (r == j || r == j + 1) {
#pragma unroll
                for (int k = 0; k < 16; k += 2)
                    *reinterpret_cast<d2_t *>(&prow[buf][r - j][16 * g + k]) = (d2_t){a[k], a[k + 1]};
            }
            __syncthreads();
            // every LDS read of the pair is issued before the reciprocals
            double p0 = pcol[buf][0][j];
            const double q0 = pcol[buf][0][j + 1];       // a(j+1, j)
            const double p1raw = pcol[buf][1][j + 1];    // a(j+1, j+1)
            const double c0 = pcol[buf][0][r], c1 = pcol[buf][1][r];
            const double s01 = prow[buf][0][j + 1];      // a(j, j+1)
            d2_t u0[8], u1[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                u0[v] = *reinterpret_cast<const d2_t *>(&prow[buf][0][16 * g + 2 * v]);
                u1[v] = *reinterpret_cast<const d2_t *>(&prow[buf][1][16 * g + 2 * v]);
            }
            if (!(fabs(p0) >= tau)) p0 = (p0 < 0.0) ? -tau : tau;
            // reciprocal: hardware estimate refined by two Newton steps
            double rp0 = __builtin_amdgcn_rcp(p0);
            rp0 = fma(rp0, fma(-p0, rp0, 1.0), rp0);
            rp0 = fma(rp0, fma(-p0, rp0, 1.0), rp0);
            const double l0n = q0 * rp0;                  // row j+1's multiplier for column j
            double p1 = fma(-l0n, s01, p1raw);
            if (!(fabs(p1) >= tau)) p1 = (p1 < 0.0) ? -tau : tau;
            double rp1 = __builtin_amdgcn_rcp(p1);
            rp1 = fma(rp1, fma(-p1, rp1, 1.0), rp1);
            rp1 = fma(rp1, fma(-p1, rp1, 1.0), rp1);
            const double l0 = (r > j) ? c0 * rp0 : 0.0;
            const double l1 = (r > j + 1) ? fma(-l0, s01, c1) * rp1 : 0.0;
            if (g == 0) {
                LT[j][r] = l0;
                LT[j + 1][r] = l1;
                if (r == 0) {
                    piv[j] = p0;
                    rpiv[j] = rp0;
                    piv[j + 1] = p1;
                    rpiv[j + 1] = rp1;
                }
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double u0k = u0[k >> 1][k & 1];
                const double u1k = fma(-l0n, u0k, u1[k >> 1][k & 1]);   // row j+1 after step j
                a[k] = fma(-l1, u1k, fma(-l0, u0k, a[k]));
            }
        }
    }
    __syncthreads();
    int pert = 0;
    // U: rows r <= c as updated; diagonal: the (possibly raised) pivots;
    // L: the logged multipliers (LT already holds them)
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int c = 16 * g + k;
        const double v = (r < c) ? a[k] : (r == c ? piv[c] : LT[c][r]);
        if (r == c && v != a[k]) ++pert;
        if (r < nb && c < nb) A[(int64_t)r * lda + c] = (float)v;
        UT[c][r] = (r <= c) ? v : 0.0;
    }
    if (pert) atomicAdd(npert, pert);
    __syncthreads();
    // ---- inverses, 2 x 2 blocked with 32 x 32 blocks:
    //   L^-1 = [X11 0; X21 X22],  X21 = -X22 L21 X11
    //   U^-1 = [Y11 Y12; 0 Y22],  Y12 = -Y11 U12 Y22
    // the four diagonal-block inverses one per wave (lane c < 32 forms
    // column c by right-looking substitution), then the two products by the
    // whole workgroup
    constexpr int H = NB_IN / 2;
    __shared__ double Xd[4][H][H + 1];                    // X11, X22, Y11, Y22
    __shared__ double Tp[4][H][H + 1];                    // T1, T2, X21, Y12
    {
        const int c = r;
        const int o = (g & 1) * H;                        // block offset
        if (c < H) {
            double sv[H];
#pragma unroll
            for (int i = 0; i < H; ++i) sv[i] = (i == c) ? 1.0 : 0.0;
            if (g < 2) {                                  // unit lower: t ascending
#pragma unroll
                for (int t = 0; t < H - 1; ++t) {
#pragma unroll
                    for (int i0 = ((t + 1) / 2) * 2; i0 < H; i0 += 2) {
                        const d2_t l = *reinterpret_cast<const d2_t *>(&LT[o + t][o + i0]);
                        if (i0 > t) sv[i0] = fma(-l[0], sv[t], sv[i0]);
                        sv[i0 + 1] = fma(-l[1], sv[t], sv[i0 + 1]);
                    }
                }
            } else {                                      // upper: t descending
#pragma unroll
                for (int t = H - 1; t >= 0; --t) {
                    sv[t] *= rpiv[o + t];
#pragma unroll
                    for (int i0 = 0; i0 < t; i0 += 2) {
                        const d2_t u = *reinterpret_cast<const d2_t *>(&UT[o + t][o + i0]);
                        sv[i0] = fma(-u[0], sv[t], sv[i0]);
                        if (i0 + 1 < t) sv[i0 + 1] = fma(-u[1], sv[t], sv[i0 + 1]);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < H; ++i) Xd[g][i][c] = sv[i];
        }
    }
    __syncthreads();
    const int pr = tid >> 3, pc = (tid & 7) * 4;          // 32 rows x 4 columns per thread
    // T1 = L21 X11 (L21[r][t] = L[H+r][t]),  T2 = U12 Y22 (U12[r][t] = U[r][H+t])
    {
        double t1[4] = {0.0, 0.0, 0.0, 0.0}, t2[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int t = 0; t < H; ++t) {
            const double lv = LT[t][H + pr], uv = UT[H + t][pr];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                t1[q] = fma(lv, Xd[0][t][pc + q], t1[q]);
                t2[q] = fma(uv, Xd[3][t][pc + q], t2[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Tp[0][pr][pc + q] = t1[q];
            Tp[1][pr][pc + q] = t2[q];
        }
    }
    __syncthreads();
    // X21 = -X22 T1,  Y12 = -Y11 T2
    {
        double x21[4] = {0.0, 0.0, 0.0, 0.0}, y12[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int t = 0; t < H; ++t) {
            const double xv = Xd[1][pr][t], yv = Xd[2][pr][t];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                x21[q] = fma(xv, Tp[0][t][pc + q], x21[q]);
                y12[q] = fma(yv, Tp[1][t][pc + q], y12[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Tp[2][pr][pc + q] = -x21[q];
            Tp[3][pr][pc + q] = -y12[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, i = e >> 6, c = e & 63;
        const int bi = i >= H, bc = c >= H, ii = i - bi * H, cc2 = c - bc * H;
        const bool in = i < nb && c < nb;
        // L^-1 blocks: (0,0) X11, (1,0) X21, (1,1) X22, (0,1) zero
        const double lval = !bi ? (!bc ? Xd[0][ii][cc2] : 0.0)
                                : (!bc ? Tp[2][ii][cc2] : Xd[1][ii][cc2]);
        // U^-1 blocks: (0,0) Y11, (0,1) Y12, (1,1) Y22, (1,0) zero
        const double uval = !bi ? (!bc ? Xd[2][ii][cc2] : Tp[3][ii][cc2])
                                : (!bc ? 0.0 : Xd[3][ii][cc2]);
        linv[i * NB_IN + c] = in ? (float)lval : 0.f;
        uinv[i * NB_IN + c] = in ? (float)uval : 0.f;
    }
}

// L21 = A21 U11^-1 in place for rows [0, rows) of the nb-column block at A:
// one workgroup per 64 rows; the block and U^-1 staged in LDS, thread (row
// r, column group g) forms sixteen columns.
template <typename T>
__global__ __launch_bounds__(256)
void k_l21_nopiv(T *A, int64_t lda, int rows, int nb, const T *__restrict__ uinv)
{
    // U^-1 rows 16-byte aligned: a thread's sixteen columns are four vector reads
    constexpr int VW = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    __shared__ __align__(16) T Ui[NB_IN][NB_IN + VW];
    __shared__ T As[NB_IN][NB_IN + 1];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * 64;
    T av[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, i = e >> 6, t = e & 63;
        av[k] = A[(int64_t)min(r0 + i, rows - 1) * lda + min(t, nb - 1)];
        Ui[i][t] = uinv[i * NB_IN + t];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, i = e >> 6, t = e & 63;
        As[i][t] = (t < nb) ? av[k] : T(0);
    }
    __syncthreads();
    const int r = tid >> 2, g = tid & 3;
    T acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = T(0);
#pragma unroll 4
    for (int t = 0; t < NB_IN; ++t) {
        const T a = As[r][t];
        vec_t u[16 / VW];
#pragma unroll
        for (int v = 0; v < 16 / VW; ++v)
            u[v] = *reinterpret_cast<const vec_t *>(&Ui[t][g * 16 + v * VW]);
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] = fma(a, u[k / VW][k % VW], acc[k]);
    }
    if (r0 + r >= rows) return;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (g * 16 + k < nb) A[(int64_t)(r0 + r) * lda + g * 16 + k] = acc[k];
}

// ------------------------------------------ tournament pivoting (CALU)
// For the refinement solver's FP32 factor the pivots of a 64-column panel
// are chosen by a tournament instead of a search per column (communication-
// avoiding LU, Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32(4),
// 2011): every node of a reduction tree runs Gaussian elimination with
// partial pivoting on at most TS_ROWS candidate rows of the panel (their
// current values) and passes on the rows it chose as pivots, in order; the
// root's choice becomes the panel's pivot rows.  Those rows are moved to
// the top of the panel and the panel is then factored without pivoting
// (k_diag_nopiv, k_l21_nopiv), so the only synchronisation per panel is one
// launch per tree level instead of one cross-workgroup exchange per column.
//
// Node kernel: thread t holds candidate t's row (up to 64 values) in
// registers, rotated so that the current column is always element 0 (the
// step's code is the same for every column: no dynamic register index).
// Per step: wave argmax of (|v[0]|, lowest row), each wave's winning lane
// writes its row to LDS, one barrier, every thread picks the overall winner
// and eliminates with it.  LDS slots alternate by step parity, so one
// barrier per step suffices.
constexpr int TS_ROWS = 256;

__global__ __launch_bounds__(256)
void k_tsel(const float *__restrict__ P, int64_t ld, int M, int nb,
            const int *__restrict__ cand_in, int ncand, int *__restrict__ cand_out)
{
    __shared__ __align__(16) float prow[2][4][NB_IN];
    __shared__ uint32_t skey[2][4];
    __shared__ int sidx[2][4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int c0 = blockIdx.x * TS_ROWS;
    const int cnt = min(TS_ROWS, ncand - c0);
    int row = -1;
    if (t < cnt) row = cand_in ? cand_in[c0 + t] : c0 + t;
    bool active = row >= 0 && row < M;
    float v[NB_IN];
    {
        const float *src = P + (int64_t)(active ? row : 0) * ld;
#pragma unroll
        for (int c = 0; c < NB_IN; ++c) v[c] = src[min(c, nb - 1)];
#pragma unroll
        for (int c = 0; c < NB_IN; ++c)
            if (c >= nb) v[c] = 0.f;
    }
    int *out = cand_out + blockIdx.x * NB_IN;
    // Columns go in blocks of 16: inside a block the step index s is static,
    // so the current column is register v[s]; after each block the row is
    // shifted down by 16 registers (slot c then holds column 16 jb + c).
    // All 64 steps run (a panel narrower than 64 has zero columns past nb,
    // whose picks are discarded at the end); the step code has no branch on
    // the column index.
#pragma unroll 1
    for (int jb = 0; jb < NB_IN / 16; ++jb) {
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int j = 16 * jb + s, buf = j & 1;
            // key = bits(|v|) + 2 (a NaN value 1, the lowest), so every
            // active row has a key >= 1 and the node always chooses
            // min(nb, rows) pivots; 0 marks "no row left".  Ties go to the
            // lowest lane, then the lowest wave: a fixed rule (any rule
            // serves a tournament).
            const uint32_t ab = __float_as_uint(v[s]) & 0x7fffffffu;
            const uint32_t key = active ? (ab > 0x7f800000u ? 1u : ab + 2u) : 0u;
            const uint32_t wkey = wave_max_u32(key);
            const uint64_t hit = __ballot(key == wkey);
            const int wlane = __ffsll((unsigned long long)hit) - 1;
            const int widx = __builtin_amdgcn_readlane(row, wlane);
            if (lane == wlane && wkey != 0) {
#pragma unroll
                for (int c = 4 * (s / 4); c < NB_IN; c += 4)
                    *reinterpret_cast<float4 *>(&prow[buf][wv][c]) =
                        make_float4(v[c], v[c + 1], v[c + 2], v[c + 3]);
            }
            if (lane == 0) {
                skey[buf][wv] = wkey;
                sidx[buf][wv] = widx;
            }
            __syncthreads();
            uint32_t bk = skey[buf][0];
            int bi = sidx[buf][0], bw = 0;
#pragma unroll
            for (int q = 1; q < 4; ++q)
                if (skey[buf][q] > bk) {
                    bk = skey[buf][q];
                    bi = sidx[buf][q];
                    bw = q;
                }
            // bk == 0: no candidate left (fewer rows than columns), the
            // step records -1 and changes nothing
            if (t == 0 && j < nb) out[j] = (bk != 0) ? bi : -1;
            if (active && row == bi) active = false;
            const float4 *pr4 = reinterpret_cast<const float4 *>(prow[buf][bw]);
            float pr[NB_IN];
#pragma unroll
            for (int c4 = s / 4; c4 < NB_IN / 4; ++c4) {
                const float4 q = pr4[c4];
                pr[4 * c4] = q.x;
                pr[4 * c4 + 1] = q.y;
                pr[4 * c4 + 2] = q.z;
                pr[4 * c4 + 3] = q.w;
            }
            const float p0 = pr[s];
            // the multipliers only steer the choice of later pivots: a
            // reciprocal (within 1 ulp) serves
            const float rp = (p0 != 0.f) ? __builtin_amdgcn_rcpf(p0) : 0.f;
            const float l = (active && bk != 0) ? v[s] * rp : 0.f;
            // eliminate every column right of s (columns past nb are zero)
#pragma unroll
            for (int c = s + 1; c < NB_IN; ++c) v[c] = fmaf(-l, pr[c], v[c]);
        }
#pragma unroll
        for (int c = 0; c < NB_IN; ++c) v[c] = (c + 16 < NB_IN) ? v[c + 16] : 0.f;
    }
    for (int r = nb + t; r < NB_IN; r += blockDim.x) out[r] = -1;
}

// Net row movement that brings the winners w[0..nb) (panel-relative rows)
// to rows 0..nb-1 of the panel: row i receives w[i]; every non-winner row
// among 0..nb-1 moves to a row vacated by a winner at or below nb, both
// lists in ascending order.  One wavefront, ranks by ballots.  swapmap:
// count, then (dst, src) global rows.
__global__ __launch_bounds__(64)
void k_tsel_moves(int nb, int kglob, const int *__restrict__ w, int64_t *swapmap)
{
    __shared__ int s_w[NB_IN];
    const int t = threadIdx.x;
    const uint64_t below = (t == 0) ? 0ull : (~0ull >> (64 - t));    // lanes < t
    const int wt = t < nb ? w[t] : -1;
    s_w[t] = wt;
    __syncthreads();
    // winners already among the top nb rows
    bool top_is_winner = false;
    for (int i = 0; i < nb; ++i) top_is_winner |= (s_w[i] == t);
    const bool displaced = t < nb && !top_is_winner;           // row t leaves
    const bool vacates = t < nb && wt >= nb;                    // winner wt leaves a hole
    const bool moves = t < nb && wt != t;
    const uint64_t bm = __ballot(moves), bd = __ballot(displaced);
    const int nmoves = __popcll(bm);
    if (moves) {
        const int o = __popcll(bm & below);
        swapmap[1 + 2 * o] = (int64_t)kglob + t;
        swapmap[2 + 2 * o] = (int64_t)kglob + wt;
    }
    // the vacated rows in ascending order: rank of wt among them
    int vrank = 0;
    if (vacates)
        for (int i = 0; i < nb; ++i) vrank += (s_w[i] >= nb && s_w[i] < wt);
    __shared__ int s_vac[NB_IN];
    if (vacates) s_vac[vrank] = wt;
    __syncthreads();
    if (displaced) {
        const int a = __popcll(bd & below);                    // rank among displaced rows
        swapmap[1 + 2 * (nmoves + a)] = (int64_t)kglob + s_vac[a];
        swapmap[2 + 2 * (nmoves + a)] = (int64_t)kglob + t;
    }
    if (t == 0) swapmap[0] = nmoves + __popcll(bd);
}

// Chooses the panel's pivot rows by the tournament and applies the row
// movement to every column of A (ld lda, n columns) and to perm.
static int tournament_pivots(la_ctx *ctx, float *A, int64_t lda, int64_t n, int64_t k, int M,
                             int nb, int *cand0, int *cand1, int64_t *swapmap, uint64_t *perm)
{
    const float *P = A + k * lda + k;
    int ncand = M;
    const int *in = nullptr;
    int *out = cand0;
    for (;;) {
        const int nodes = (int)ceil_div(ncand, TS_ROWS);
        hipLaunchKernelGGL(k_tsel, dim3(nodes), dim3(256), 0, ctx->stream, P, lda, M, nb, in, ncand,
                           out);
        LA_HIP_CHECK(ctx, hipGetLastError());
        if (nodes == 1) break;
        ncand = nodes * NB_IN;
        in = out;
        out = (out == cand0) ? cand1 : cand0;
    }
    hipLaunchKernelGGL(k_tsel_moves, dim3(1), dim3(64), 0, ctx->stream, nb, (int)k,
                       (const int *)out, swapmap);
    LA_HIP_CHECK(ctx, hipGetLastError());
    return launch_rowmoves<float>(ctx, A, lda, n, 0, 0, swapmap, perm);
}

// In-place LU of the m x n matrix A without pivoting (right-looking, the
// same outer/inner blocking and trailing updates as getrf_core).  *d_tau is
// the static-pivot threshold; *d_npert counts perturbed pivots.  With d_perm
// non-null the pivots of every panel are chosen by the tournament first
// (rows moved, d_perm permuted alike; d_perm must hold the identity on
// entry), and the panel is then factored without pivoting.
template <typename T>
int getrf_nopiv_core(la_ctx *ctx, int64_t m, int64_t n, T *A, int64_t lda, const float *d_tau,
                     int32_t *d_npert, int upd_prec, uint64_t *d_perm)
{
    static_assert(sizeof(T) == 4, "the unpivoted factor is FP32");
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2)
        return set_error(ctx, LA_ERR_ARG, "getrf_nopiv: extent too large");
    const int64_t kmax = m < n ? m : n;
    constexpr int NBO = 2 * NB_OUT;               // outer block width (trailing GEMM K)
    constexpr int NLINV = NBO / NB_IN;
    const size_t inv_bytes = (size_t)(NLINV + 1) * NB_IN * NB_IN * sizeof(T);
    const size_t cand_bytes = ws_align((size_t)ceil_div(m, TS_ROWS) * NB_IN * sizeof(int));
    const size_t map_bytes = ws_align((size_t)(1 + 4 * NB_IN) * sizeof(int64_t));
    void *base = nullptr;
    int st = ws_get(ctx, WS_PANEL, ws_align(inv_bytes) + 2 * cand_bytes + map_bytes, &base);
    if (st) return st;
    T *linv_ring = (T *)base;
    T *uinv = linv_ring + (size_t)NLINV * NB_IN * NB_IN;
    int *cand0 = (int *)((char *)base + ws_align(inv_bytes));
    int *cand1 = (int *)((char *)cand0 + cand_bytes);
    int64_t *swapmap = (int64_t *)((char *)cand1 + cand_bytes);
    if ((st = fill_i32(ctx, d_npert, 1, 0))) return st;

    for (int64_t K0 = 0; K0 < kmax; K0 += NBO) {
        const int64_t nbo = std::min<int64_t>(NBO, kmax - K0);
        for (int64_t k = K0; k < K0 + nbo; k += NB_IN) {
            const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - k);
            T *Ap = A + k * lda + k;
            T *linv = linv_ring + (size_t)((k - K0) / NB_IN) * NB_IN * NB_IN;
            if (d_perm) {
                ProfScope ps(ctx, LA_K_PIVSEL, (double)(m - k) * nbi * nbi,
                             (double)(m - k) * nbi * sizeof(T));
                if ((st = tournament_pivots(ctx, A, lda, n, k, (int)(m - k), nbi, cand0, cand1,
                                            swapmap, d_perm)))
                    return st;
            }
            {
                ProfScope ps(ctx, LA_K_PANEL, (double)(m - k) * nbi * nbi,
                             2.0 * (m - k) * nbi * sizeof(T));
                hipLaunchKernelGGL(k_diag_nopiv, dim3(1), dim3(256), 0, ctx->stream, Ap, lda,
                                   nbi, d_tau, linv, uinv, d_npert);
                LA_HIP_CHECK(ctx, hipGetLastError());
                const int64_t below = m - k - nbi;
                if (below > 0) {
                    hipLaunchKernelGGL(k_l21_nopiv<T>, dim3((unsigned)ceil_div(below, 64)),
                                       dim3(256), 0, ctx->stream, Ap + (int64_t)nbi * lda, lda,
                                       (int)below, nbi, (const T *)uinv);
                    LA_HIP_CHECK(ctx, hipGetLastError());
                }
            }
            const int64_t rest = K0 + nbo - (k + nbi);
            if (rest > 0) {
                if ((st = launch_trsm<T>(ctx, linv, nbi, A + k * lda + k + nbi, lda, rest)))
                    return st;
                if (m - k - nbi > 0)
                    if ((st = trail_gemm<T>(ctx, upd_prec, m - k - nbi, rest, nbi,
                                            A + (k + nbi) * lda + k, lda, A + k * lda + k + nbi,
                                            lda, A + (k + nbi) * lda + k + nbi, lda)))
                        return st;
            }
        }
        const int64_t ncol = n - (K0 + nbo);
        if (ncol > 0) {
            T *B12 = A + K0 * lda + K0 + nbo;
            for (int64_t kb = K0; kb < K0 + nbo; kb += NB_IN) {
                const int nbi = (int)std::min<int64_t>(NB_IN, K0 + nbo - kb);
                const T *linv_kb = linv_ring + (size_t)((kb - K0) / NB_IN) * NB_IN * NB_IN;
                if ((st = launch_trsm<T>(ctx, linv_kb, nbi, A + kb * lda + K0 + nbo, lda, ncol)))
                    return st;
                const int64_t below = K0 + nbo - (kb + nbi);
                if (below > 0)
                    if ((st = trail_gemm<T>(ctx, upd_prec, below, ncol, nbi,
                                            A + (kb + nbi) * lda + kb, lda,
                                            A + kb * lda + K0 + nbo, lda,
                                            A + (kb + nbi) * lda + K0 + nbo, lda)))
                        return st;
            }
            if (K0 + nbo < m)
                if ((st = trail_gemm<T>(ctx, upd_prec, m - K0 - nbo, ncol, nbo,
                                        A + (K0 + nbo) * lda + K0, lda, B12, lda,
                                        A + (K0 + nbo) * lda + K0 + nbo, lda)))
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
                                uint64_t *, int32_t *, int, void *);
template int getrf_core<float>(la_ctx *, int64_t, int64_t, float *, int64_t, int64_t *,
                               uint64_t *, int32_t *, int, void *);

// ------------------------------------------------ allocation-free entry
// la_dgetrf_noalloc_dev takes its device workspace (the panel exchange
// slots, swap map, L11^-1 ring and the LAPACK-style swap list) from this
// fixed array instead of the context's grow-only workspace, so it calls no
// allocator.  Every device that loads the library holds its own copy until
// the process ends.  Its device address is looked up once per device (in
// la_ctx_create, with that device current) and kept in g_scratch_ptr; host
// code uses only that address, never the array's name.
constexpr int64_t NOALLOC_MAX_K = 65536;
constexpr size_t NOALLOC_PANEL_BYTES = panel_ws_bytes(sizeof(double));
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_PANEL_BYTES + NOALLOC_MAX_K * sizeof(int64_t);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return LA_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return LA_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LA_ERR_HIP;
    g_scratch_ptr[device] = p;
    return LA_OK;
}
template int getrf_nopiv_core<float>(la_ctx *, int64_t, int64_t, float *, int64_t, const float *,
                                     int32_t *, int, uint64_t *);

}  // namespace lahip

using namespace lahip;

extern "C" int la_dgetrf_noalloc_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda,
                                     uint64_t *d_piv, int32_t *d_info)
{
    if (!ctx || !A || !d_piv || !d_info) return LA_ERR_ARG;
    if (m <= 0 || n <= 0) return set_error(ctx, LA_ERR_SHAPE, "getrf: empty matrix");
    if (lda < n) return LA_ERR_ARG;
    const int64_t kmax = std::min(m, n);
    if (kmax > NOALLOC_MAX_K)
        return set_error(ctx, LA_ERR_ARG, "getrf_noalloc: min(m, n) exceeds 65536, the size "
                                          "of the fixed scratch");
    void *base = (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES)
                     ? g_scratch_ptr[ctx->device] : nullptr;
    if (!base) return set_error(ctx, LA_ERR_HIP, "getrf_noalloc: scratch address not set up");
    int64_t *ipiv = (int64_t *)((char *)base + NOALLOC_PANEL_BYTES);
    // the timing harness creates events; keep it out of this call
    const bool prof = ctx->prof;
    ctx->prof = false;
    int st = getrf_core<double>(ctx, m, n, A, lda, ipiv, d_piv, d_info + 1, LA_PREC_FP64, base);
    if (st == LA_OK) st = pivsign(ctx, ipiv, kmax, d_info);
    ctx->prof = prof;
    return st;
}

#ifdef LAHIP_PANEL_TIMING
// Probe-only export of the panel phase clock (and reset).
extern "C" int la_debug_panel_phase(unsigned long long *out8, int reset)
{
    if (hipMemcpyFromSymbol(out8, HIP_SYMBOL(lahip::g_panel_phase), 8 * sizeof(unsigned long long)) !=
        hipSuccess)
        return LA_ERR_HIP;
    if (reset) {
        unsigned long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (hipMemcpyToSymbol(HIP_SYMBOL(lahip::g_panel_phase), z, sizeof z) != hipSuccess)
            return LA_ERR_HIP;
    }
    return LA_OK;
}
#endif
