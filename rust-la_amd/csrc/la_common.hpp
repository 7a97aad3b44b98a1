// Shared definitions for the gfx950 kernels and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <algorithm>
#include <type_traits>
#include "../../include/la_hip.h"

namespace lahip {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// ---- cross-workgroup exchange without cache-wide fences -------------------
// Payload written for other workgroups of the same kernel goes out with
// device-coherent (sc1) stores, then each writer waits for its stores to
// complete (s_waitcnt vmcnt(0)) before the workgroup barrier that precedes
// the flag store.  Readers poll the flag with sc1 loads and fetch the
// payload with sc1 loads issued after the flag was seen.  This orders the
// exchange without the agent-scope acquire/release fences, which on gfx950
// invalidate / write back the whole L2 (buffer_inv sc1 / buffer_wbl2 sc1).
template <typename T>
__device__ inline void st_dev(T *p, typename std::remove_reference<T>::type v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T ld_dev(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every store issued by this thread has completed
__device__ inline void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// SplitMix64 finaliser over a counter: the seeded generators' bit source.
__device__ __host__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ULL;

// Context: one device, one stream, a grow-only workspace, small device scratch
// for status words.
struct Workspace {
    void *ptr = nullptr;
    size_t size = 0;
};

}  // namespace lahip

namespace lahip {
struct ProfRecord {
    hipEvent_t start, stop;
    int kclass;
    double flop, bytes;
};
struct ProfStats {
    int64_t calls = 0;
    double ms = 0, flop = 0, bytes = 0;
};
}  // namespace lahip

struct la_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 0;
    int64_t hbm_bytes = 0;
    int clock_khz = 0;
    int64_t ws_limit = 0;
    lahip::Workspace ws[9];     // independent grow-only workspaces (WS_* ids)
    int32_t *d_scratch = nullptr;   // 4 KiB of device status/counter words
    int32_t *h_pinned = nullptr;    // 4 KiB pinned host mirror
    std::string last_error;
    // kernel timing harness
    uint32_t trsv_epoch = 0;        // tag of the triangular-solve exchange
    // batched Cholesky tile configuration per size bucket: panel width | row tiles << 8
    int chol_nb[4] = {16 | (2 << 8), 16 | (2 << 8), 16 | (2 << 8), 16 | (2 << 8)};
    bool chol_tuned = false;
    // batched LU + refinement: workgroups per CU by working precision
    // (LA_PREC_FP64 / FP32 / FP16 index), 0 = occupancy limit
    int lu_wg_per_cu[4] = {0, 0, 0, 0};
    bool lu_tuned = false;
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    hipEvent_t la_ev[6] = {};       // LU look-ahead stream joins
    std::vector<lahip::ProfRecord> prof_pending;
    lahip::ProfStats prof_stats[LA_K_COUNT];
};

namespace lahip {

// Workspace slots.  A slot may be re-allocated (grown) by ws_get, which
// invalidates earlier pointers from the SAME slot, so each slot has one user
// per API call.
enum { WS_HOST = 0, WS_PANEL = 1, WS_SYNC = 2, WS_PIV = 3, WS_LP = 4, WS_IR = 5,
       WS_BATCH = 6, WS_GEN = 7, WS_QR = 8, WS_COUNT = 9 };

// Device scratch words (ctx->d_scratch, 1024 int32):
//   [0]  sticky device error (LA_ERR_TIMEOUT from bounded waits)
//   [8]  Cholesky flag word      [16] singular flag      [24..31] status words
enum { SCR_ERR = 0, SCR_CHOL = 8, SCR_SING = 16, SCR_STATUS = 24, SCR_RED = 64 };

int set_error(la_ctx *ctx, int status, const char *what);
int hip_fail(la_ctx *ctx, hipError_t e, const char *what);
// Copies n device status words to the host (synchronous).
int read_scratch(la_ctx *ctx, const int32_t *d, int32_t *h, int n);
// LA_OK, or the sticky device error of a bounded wait (then cleared).
int check_sticky(la_ctx *ctx);
// Returns a device buffer of at least `bytes` from workspace slot `id`.
int ws_get(la_ctx *ctx, int id, size_t bytes, void **out);

#define LA_HIP_CHECK(ctx, expr)                                         \
    do {                                                                \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess) return ::lahip::hip_fail((ctx), _e, #expr); \
    } while (0)

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Brackets the launches made during its lifetime with events when the
// context's timing harness is on (see la_ctx_set_profiling).
struct ProfScope {
    la_ctx *ctx;
    ProfRecord rec;
    bool on;
    ProfScope(la_ctx *c, int kclass, double flop, double bytes);
    ~ProfScope();
};

// ---- helpers implemented in util.hip --------------------------------------
// Allow a kernel the dynamic LDS left over by its static LDS (160 KiB per CU
// on gfx950); returns that limit in bytes and leaves no sticky HIP error.
size_t allow_dynamic_lds(const void *kernel);
uint64_t spin_limit_ticks(la_ctx *ctx);
int fill_i32(la_ctx *ctx, int32_t *p, int64_t n, int32_t v);
int fill_u64(la_ctx *ctx, uint64_t *p, int64_t n, uint64_t v);
int iota_u64(la_ctx *ctx, uint64_t *p, int64_t n);
int copy2d(la_ctx *ctx, int64_t rows, int64_t cols, const double *src, int64_t lds,
           double *dst, int64_t ldd);
// status := LA_SINGULAR if any of the first n diagonal entries is exactly 0
int check_diag_zero(la_ctx *ctx, int64_t n, const double *LU, int64_t ld, int32_t *status);

// ---- kernel launchers implemented in the .hip translation units ----------
// GEMM: C = alpha*A*op(B) + beta*C, row-major.  trans_b: B is given as an
// N x K row-major matrix (B^T).  lower_only: skip tiles strictly above the
// diagonal (symmetric rank-k updates).
template <typename T>
int gemm_rm(la_ctx *ctx, int64_t M, int64_t N, int64_t K, T alpha,
            const T *A, int64_t lda, const T *B, int64_t ldb, T beta,
            T *C, int64_t ldc, bool trans_b, bool lower_only);

// Blocked right-looking LU with partial pivoting in precision T, in place.
// d_ipiv (device, min(m,n)) receives LAPACK-style swaps (row k exchanged
// with row d_ipiv[k], global 0-based); d_perm (device, m, may be null) the
// permutation vector; d_status LA_OK or LA_ERR_TIMEOUT.
// upd_prec: LA_PREC_FP32/FP64 = native MFMA trailing update; LA_PREC_BF16 /
// LA_PREC_FP16 (T = float only) = operands rounded to that format.
template <typename T>
int getrf_core(la_ctx *ctx, int64_t m, int64_t n, T *A, int64_t lda, int64_t *d_ipiv,
               uint64_t *d_perm, int32_t *d_status, int upd_prec = LA_PREC_FP32,
               void *ext_ws = nullptr);
// Looks up this device's copy of the fixed scratch of la_dgetrf_noalloc_dev
// (once per device; the device must be current).
int noalloc_scratch_init(int device);
// LU without pivoting (static pivots below *d_tau, counted in *d_npert)
// With d_perm (identity on entry) the pivots of every 64-column panel are
// chosen by a tournament (CALU) and applied to A and d_perm first.
template <typename T>
int getrf_nopiv_core(la_ctx *ctx, int64_t m, int64_t n, T *A, int64_t lda, const float *d_tau,
                     int32_t *d_npert, int upd_prec = LA_PREC_FP32, uint64_t *d_perm = nullptr);
int gemm_lp(la_ctx *ctx, int prec, int64_t M, int64_t N, int64_t K, float alpha, const float *A,
            int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc);
// d_info[0] = pospivsign (1 = even number of swaps)
int pivsign(la_ctx *ctx, const int64_t *d_ipiv, int64_t kmax, int32_t *d_info);

// Blocked triangular solves, FP64 arithmetic, factor TF:
//   ('L', trans=false, unit=true)  L y = b  (packed LU)
//   ('U', false, false)            U x = y  (packed LU)
//   ('L', false, false)            L y = b  (Cholesky)
//   ('L', true,  false)            L^T x = y (Cholesky)
template <typename TF>
int trsv_blocked(la_ctx *ctx, char uplo, bool trans, bool unit, int64_t n,
                 const TF *Tm, int64_t ldt, int64_t nrhs, double *X, int64_t ldx);
int gather_rows(la_ctx *ctx, int64_t m, int64_t nx, const uint64_t *d_piv, const double *B,
                int64_t ldb, double *X, int64_t ldx);
// X <- T^-1 X for many right-hand sides: blocked TRSM on the FP64 MFMA GEMM
// with inverted 64 x 64 diagonal blocks ('L' unit lower / 'U' upper of a
// packed LU)
int trsm_blocked(la_ctx *ctx, char uplo, int64_t n, const double *Tm, int64_t ldt, int64_t nrhs,
                 double *X, int64_t ldx);
// right-hand-side counts from which the LU solves use trsm_blocked
constexpr int64_t TRSM_MIN_RHS = 17, TRSM_MIN_N = 65;

int potrf_blocked(la_ctx *ctx, int64_t n, double *A, int64_t lda, int32_t *d_status);

}  // namespace lahip
