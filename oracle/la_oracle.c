/*
 * la_oracle.c -- CPU restatement of the rust-la (`la` 0.2.0) dense hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library is
 * compared against.  Only tests/, the smoke() entry and bench.py's
 * cpu_baseline leg may load it; the shipped library never links or calls it.
 *
 * Every routine restates the reference loop-for-loop, in the same order, in
 * IEEE double with no contraction (build with -O2 -ffp-contract=off, no
 * fast-math).  Rust never fuses a*b+c and never reassociates, so with those
 * flags the results are bit-identical to the reference by language semantics.
 *
 *   orc_mul            src/matrix/mod.rs:957-980     (impl Mul for &Matrix)
 *   orc_mmul           src/matrix/mmatrix.rs:82-98   (Matrix::mmul)
 *   orc_lu_new         src/decomp/lu.rs:104-168      (LUDecomposition::new)
 *   orc_lu_is_non_singular lu.rs:174-182
 *   orc_lu_det         src/decomp/lu.rs:224-232
 *   orc_lu_solve       src/decomp/lu.rs:237-278
 *   orc_chol_new       src/decomp/cholesky.rs:56-110
 *   orc_chol_solve     src/decomp/cholesky.rs:116-144
 *   orc_smul           src/matrix/simd.rs:189-219    (Matrix<f32>::simd_mul)
 *   orc_sdot           src/matrix/simd.rs:166-187    (simd_dot, tail defect
 *                      corrected: the tail adds a[i]*b[i], see below)
 *   orc_qr_new         src/decomp/qr.rs:26-106       (QRDecomposition::new)
 *   orc_qr_is_full_rank qr.rs:108-115
 *   orc_qr_q           src/decomp/qr.rs:150-194      (get_q)
 *   orc_qr_solve_literal qr.rs:196-237 as written; orc_qr_solve the
 *                      least-squares solve it means (see below)
 *   orc_pinverse       src/matrix/mod.rs:1049-1057
 *
 * Status codes (mirroring the reference's failure signals):
 *   0  OK
 *   1  shape error     -> the reference panics (assert!)
 *   2  singular        -> LUDecomposition::solve returns None
 *   3  not square      -> CholeskyDecomposition::new returns None
 *   4  not symmetric   -> CholeskyDecomposition::new returns None
 *   5  not SPD         -> CholeskyDecomposition::new returns None
 *
 * Parity pins for each routine are listed in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <math.h>
#include <stdlib.h>

#define ORC_OK 0
#define ORC_SHAPE 1
#define ORC_SINGULAR 2
#define ORC_NOT_SQUARE 3
#define ORC_NOT_SYMMETRIC 4
#define ORC_NOT_SPD 5

/* ---- C = A * B  (mod.rs:957-980).  res starts at zero and accumulates
 * res = res + a*b for idx ascending; the product is rounded before the add. */
int orc_mul(int64_t m, int64_t k, int64_t k2, int64_t p,
            const double *a, const double *b, double *c)
{
    if (k != k2) return ORC_SHAPE;                  /* mod.rs:961 assert! */
    if (m <= 0 || k <= 0 || p <= 0) return ORC_SHAPE; /* Matrix::new asserts rows,cols > 0 (mod.rs:209) */
    for (int64_t row = 0; row < m; ++row) {
        for (int64_t col = 0; col < p; ++col) {
            double res = 0.0;
            for (int64_t idx = 0; idx < k; ++idx) {
                double prod = a[row * k + idx] * b[idx * p + col];
                res = res + prod;
            }
            c[row * p + col] = res;
        }
    }
    return ORC_OK;
}

/* ---- mmul (mmatrix.rs:82-98): same loop, shape of dst asserted. */
int orc_mmul(int64_t m, int64_t k, int64_t k2, int64_t p,
             int64_t dst_rows, int64_t dst_cols,
             const double *a, const double *b, double *dst)
{
    if (k != k2) return ORC_SHAPE;          /* mmatrix.rs:83 */
    if (dst_rows != m) return ORC_SHAPE;    /* mmatrix.rs:84 */
    if (dst_cols != p) return ORC_SHAPE;    /* mmatrix.rs:85 */
    return orc_mul(m, k, k2, p, a, b, dst);
}

/* ---- LUDecomposition::new (lu.rs:104-168), JAMA left-looking Doolittle.
 * lu (m*n, row-major) receives the packed factors, piv (m) the permutation
 * vector (row i of PA is row piv[i] of A), *pospivsign the sign flag.
 * If min_gap is non-NULL it receives the smallest relative gap between the
 * largest and second-largest |candidate| over all pivot steps with at least
 * two candidates and a nonzero maximum (1.0 when no such step exists); this
 * is the tie-freeness certificate used by the GPU pivot-parity tests. */
int orc_lu_new(int64_t m, int64_t n, const double *a, double *lu,
               uint64_t *piv, int *pospivsign, double *min_gap)
{
    if (m <= 0 || n <= 0) return ORC_SHAPE;
    memcpy(lu, a, (size_t)(m * n) * sizeof(double));     /* lu.rs:105 */
    for (int64_t i = 0; i < m; ++i) piv[i] = (uint64_t)i; /* lu.rs:108-111 */
    int sign = 1;
    double gap = 1.0;
    for (int64_t j = 0; j < n; ++j) {
        /* lu.rs:122-129: lu[i][j] -= sum_{k<min(i,j)} lu[i][k]*lu[k][j] */
        for (int64_t i = 0; i < m; ++i) {
            double s = 0.0;
            int64_t kmax = i < j ? i : j;
            for (int64_t kk = 0; kk < kmax; ++kk) {
                double prod = lu[i * n + kk] * lu[kk * n + j];
                s = s + prod;
            }
            lu[i * n + j] = lu[i * n + j] - s;
        }
        /* lu.rs:132-137: first row with strictly largest |.| at or below j */
        int64_t p = j;
        for (int64_t i = j + 1; i < m; ++i) {
            if (fabs(lu[i * n + j]) > fabs(lu[p * n + j])) p = i;
        }
        if (min_gap && j + 1 < m) {
            double best = -1.0, second = -1.0;
            for (int64_t i = j; i < m; ++i) {
                double v = fabs(lu[i * n + j]);
                if (v != v) continue;
                if (v > best) { second = best; best = v; }
                else if (v > second) second = v;
            }
            if (best > 0.0 && second >= 0.0) {
                double g = (best - second) / best;
                if (g < gap) gap = g;
            }
        }
        /* lu.rs:140-152: swap whole rows, swap piv entries, flip sign */
        if (p != j) {
            for (int64_t kk = 0; kk < n; ++kk) {
                double t = lu[p * n + kk];
                lu[p * n + kk] = lu[j * n + kk];
                lu[j * n + kk] = t;
            }
            uint64_t t = piv[p];
            piv[p] = piv[j];
            piv[j] = t;
            sign = !sign;
        }
        /* lu.rs:156-160: divide (not reciprocal-multiply) unless pivot == 0 */
        if (j < m && lu[j * n + j] != 0.0) {
            for (int64_t i = j + 1; i < m; ++i)
                lu[i * n + j] = lu[i * n + j] / lu[j * n + j];
        }
    }
    *pospivsign = sign;
    if (min_gap) *min_gap = gap;
    return ORC_OK;
}

/* ---- is_non_singular (lu.rs:174-182): no exact zero on the diagonal for
 * j < cols.  The reference indexes lu[j*n+j] for every j < n, which runs
 * past the end of the data (a panic) when m < n; that case returns -1. */
/* The same factorisation with column j of lu copied to a contiguous buffer
 * before its dot products (JAMA's LUcolj): every product and sum is the
 * same operation on the same operands in the same order as orc_lu_new, so
 * the results are bit-identical; only the memory walk differs (the
 * reference's strided column walk makes n = 8192 take hours).  Used to
 * write large fixtures; orc_lu_new stays the statement of the reference
 * and its timing is the CPU baseline. */
int orc_lu_new_colcopy(int64_t m, int64_t n, const double *a, double *lu,
                       uint64_t *piv, int *pospivsign, double *min_gap)
{
    if (m <= 0 || n <= 0) return ORC_SHAPE;
    double *col = (double *)malloc((size_t)m * sizeof(double));
    if (!col) return ORC_SHAPE;
    memcpy(lu, a, (size_t)(m * n) * sizeof(double));
    for (int64_t i = 0; i < m; ++i) piv[i] = (uint64_t)i;
    int sign = 1;
    double gap = 1.0;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) col[i] = lu[i * n + j];
        for (int64_t i = 0; i < m; ++i) {
            double s = 0.0;
            int64_t kmax = i < j ? i : j;
            const double *row = lu + i * n;
            for (int64_t kk = 0; kk < kmax; ++kk) {
                double prod = row[kk] * col[kk];
                s = s + prod;
            }
            col[i] = col[i] - s;
            lu[i * n + j] = col[i];
        }
        int64_t p = j;
        for (int64_t i = j + 1; i < m; ++i) {
            if (fabs(lu[i * n + j]) > fabs(lu[p * n + j])) p = i;
        }
        if (min_gap && j + 1 < m) {
            double best = -1.0, second = -1.0;
            for (int64_t i = j; i < m; ++i) {
                double v = fabs(lu[i * n + j]);
                if (v != v) continue;
                if (v > best) { second = best; best = v; }
                else if (v > second) second = v;
            }
            if (best > 0.0 && second >= 0.0) {
                double g = (best - second) / best;
                if (g < gap) gap = g;
            }
        }
        if (p != j) {
            for (int64_t kk = 0; kk < n; ++kk) {
                double t = lu[p * n + kk];
                lu[p * n + kk] = lu[j * n + kk];
                lu[j * n + kk] = t;
            }
            uint64_t t = piv[p];
            piv[p] = piv[j];
            piv[j] = t;
            sign = !sign;
        }
        if (j < m && lu[j * n + j] != 0.0) {
            for (int64_t i = j + 1; i < m; ++i)
                lu[i * n + j] = lu[i * n + j] / lu[j * n + j];
        }
    }
    free(col);
    *pospivsign = sign;
    if (min_gap) *min_gap = gap;
    return ORC_OK;
}

int orc_lu_is_non_singular(int64_t m, int64_t n, const double *lu)
{
    for (int64_t j = 0; j < n; ++j) {
        if (j * n + j >= m * n) return -1;
        if (lu[j * n + j] == 0.0) return 0;
    }
    return 1;
}

/* ---- det (lu.rs:224-232): sign * prod_j lu[j][j], product in j order. */
int orc_lu_det(int64_t m, int64_t n, const double *lu, int pospivsign, double *det)
{
    if (m != n) return ORC_SHAPE;                  /* lu.rs:225 */
    double d = pospivsign ? 1.0 : -1.0;
    for (int64_t j = 0; j < n; ++j) d = d * lu[j * n + j];
    *det = d;
    return ORC_OK;
}

/* ---- LUDecomposition::solve (lu.rs:237-278).  x is m*nx. */
int orc_lu_solve(int64_t m, int64_t n, const double *lu, const uint64_t *piv,
                 int64_t b_rows, int64_t nx, const double *b, double *x)
{
    if (b_rows != m) return ORC_SHAPE;             /* lu.rs:240 */
    int ns = orc_lu_is_non_singular(m, n, lu);     /* lu.rs:241-243 */
    if (ns < 0) return ORC_SHAPE;
    if (ns == 0) return ORC_SINGULAR;
    int64_t dest = 0;                              /* lu.rs:246-254 */
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < nx; ++j)
            x[dest++] = b[(int64_t)piv[i] * nx + j];
    for (int64_t k = 0; k < n; ++k)                /* lu.rs:257-263 */
        for (int64_t i = k + 1; i < n; ++i)
            for (int64_t j = 0; j < nx; ++j) {
                double prod = x[k * nx + j] * lu[i * n + k];
                x[i * nx + j] = x[i * nx + j] - prod;
            }
    for (int64_t k = n - 1; k >= 0; --k) {         /* lu.rs:266-275 */
        for (int64_t j = 0; j < nx; ++j)
            x[k * nx + j] = x[k * nx + j] / lu[k * n + k];
        for (int64_t i = 0; i < k; ++i)
            for (int64_t j = 0; j < nx; ++j) {
                double prod = x[k * nx + j] * lu[i * n + k];
                x[i * nx + j] = x[i * nx + j] - prod;
            }
    }
    return ORC_OK;
}

/* ---- CholeskyDecomposition::new (cholesky.rs:56-110), row by row.
 * l receives the full n*n factor with the strict upper triangle zeroed.
 * *fail_row (if non-NULL) receives the row j at which None was returned. */
int orc_chol_new(int64_t rows, int64_t cols, const double *a, double *l, int64_t *fail_row)
{
    if (rows != cols) return ORC_NOT_SQUARE;       /* cholesky.rs:57-59 */
    int64_t n = rows;
    for (int64_t j = 0; j < n; ++j) {
        double d = 0.0;
        for (int64_t k = 0; k < j; ++k) {
            double s = 0.0;                        /* cholesky.rs:76-79 */
            for (int64_t i = 0; i < k; ++i) {
                double prod = l[k * n + i] * l[j * n + i];
                s = s + prod;
            }
            s = (a[j * n + k] - s) / l[k * n + k]; /* cholesky.rs:82-83 */
            l[j * n + k] = s;
            double sq = s * s;                     /* cholesky.rs:86 */
            d = d + sq;
            if (a[k * n + j] != a[j * n + k]) {    /* cholesky.rs:89-91 */
                if (fail_row) *fail_row = j;
                return ORC_NOT_SYMMETRIC;
            }
        }
        d = a[j * n + j] - d;                      /* cholesky.rs:96 */
        if (d <= 0.0) {                            /* cholesky.rs:97-100 */
            if (fail_row) *fail_row = j;
            return ORC_NOT_SPD;
        }
        l[j * n + j] = sqrt(d);                    /* cholesky.rs:101 */
        for (int64_t k = j + 1; k < n; ++k) l[j * n + k] = 0.0; /* :104-106 */
    }
    return ORC_OK;
}

/* ---- CholeskyDecomposition::solve (cholesky.rs:116-144). x is n*nx. */
int orc_chol_solve(int64_t n, const double *l, int64_t b_rows, int64_t nx,
                   const double *b, double *x)
{
    if (b_rows != n) return ORC_SHAPE;             /* cholesky.rs:118 */
    memcpy(x, b, (size_t)(n * nx) * sizeof(double));
    for (int64_t k = 0; k < n; ++k)                /* cholesky.rs:124-131 */
        for (int64_t j = 0; j < nx; ++j) {
            for (int64_t i = 0; i < k; ++i) {
                double prod = x[i * nx + j] * l[k * n + i];
                x[k * nx + j] = x[k * nx + j] - prod;
            }
            x[k * nx + j] = x[k * nx + j] / l[k * n + k];
        }
    for (int64_t k = n - 1; k >= 0; --k)           /* cholesky.rs:134-141 */
        for (int64_t j = 0; j < nx; ++j) {
            for (int64_t i = k + 1; i < n; ++i) {
                double prod = x[i * nx + j] * l[i * n + k];
                x[k * nx + j] = x[k * nx + j] - prod;
            }
            x[k * nx + j] = x[k * nx + j] / l[k * n + k];
        }
    return ORC_OK;
}

/* ---- Seeded generators (no reference counterpart: the reference's only
 * random source, Matrix::random at mod.rs:842-851, is unseeded).
 * SplitMix64 finaliser over a counter; stream s of seed q has key
 * K = mix(q + (s+1)*G); element i is mix(K + (i+1)*G), G = 0x9E3779B97F4A7C15. */
static inline uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t key = orc_mix64(seed + (stream + 1) * ORC_GOLDEN);
    return orc_mix64(key + (i + 1) * ORC_GOLDEN);
}

/* uniform in [-1,1): 2*u - 1 with u = (bits >> 11) * 2^-53 (exact). */
void orc_fill_uniform(uint64_t seed, uint64_t stream, int64_t count, double *out)
{
    for (int64_t i = 0; i < count; ++i) {
        double u = (double)(orc_rng_u64(seed, stream, (uint64_t)i) >> 11) * 0x1.0p-53;
        out[i] = 2.0 * u - 1.0;
    }
}

/* standard normal by Box-Muller over counter pairs (2p, 2p+1):
 * u1 = ((bits(2p)>>11) + 1) * 2^-53 in (0,1], u2 = (bits(2p+1)>>11) * 2^-53,
 * r = sqrt(-2 ln u1); out[2p] = r cos(2 pi u2), out[2p+1] = r sin(2 pi u2). */
void orc_fill_normal(uint64_t seed, uint64_t stream, int64_t count, double *out)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t p = 0; 2 * p < count; ++p) {
        double u1 = (double)((orc_rng_u64(seed, stream, (uint64_t)(2 * p)) >> 11) + 1) * 0x1.0p-53;
        double u2 = (double)(orc_rng_u64(seed, stream, (uint64_t)(2 * p + 1)) >> 11) * 0x1.0p-53;
        double r = sqrt(-2.0 * log(u1));
        double t = two_pi * u2;
        out[2 * p] = r * cos(t);
        if (2 * p + 1 < count) out[2 * p + 1] = r * sin(t);
    }
}

/* ---------------------------------------------------------------- f32
 * Matrix<f32>::simd_mul (src/matrix/simd.rs:189-219): for each row, each
 * block of four output columns accumulates res = res + splat(a[row][i]) *
 * b[i][4c..4c+4] over i ascending, each lane a separately rounded FP32
 * product then sum (f32x4 mul and add, no fusion); the remaining columns
 * (p mod 4) use the scalar loop res += a * b, i ascending. */
int orc_smul(int64_t m, int64_t k, int64_t k2, int64_t p, const float *a, const float *b,
             float *c)
{
    if (k != k2) return 1;                       /* simd.rs:190 */
    for (int64_t row = 0; row < m; ++row) {
        for (int64_t cb = 0; cb < p / 4; ++cb) {
            float res[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int64_t i = 0; i < k; ++i) {
                const float av = a[row * k + i];
                for (int l = 0; l < 4; ++l) {
                    const float prod = av * b[i * p + 4 * cb + l];
                    res[l] = res[l] + prod;
                }
            }
            for (int l = 0; l < 4; ++l) c[row * p + 4 * cb + l] = res[l];
        }
        for (int64_t col = p & ~(int64_t)3; col < p; ++col) {
            float res = 0.0f;
            for (int64_t i = 0; i < k; ++i) {
                const float prod = a[row * k + i] * b[i * p + col];
                res += prod;
            }
            c[row * p + col] = res;
        }
    }
    return 0;
}

/* simd_dot (src/matrix/simd.rs:166-187): four lanes accumulate a*b over the
 * 4-element groups, then sum = ((l0 + l1) + l2) + l3.  The reference's tail
 * loop runs over [len - (len & !3), len) and adds a[i] alone; that is a
 * defect, so this restatement adds a[i] * b[i] for the len mod 4 trailing
 * elements instead (the mathematically intended dot product). */
float orc_sdot(int64_t len, const float *a, const float *b)
{
    float lanes[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t i = 0; i < len / 4; ++i)
        for (int l = 0; l < 4; ++l) {
            const float prod = a[4 * i + l] * b[4 * i + l];
            lanes[l] = lanes[l] + prod;
        }
    float sum = ((lanes[0] + lanes[1]) + lanes[2]) + lanes[3];
    for (int64_t i = len & ~(int64_t)3; i < len; ++i) {
        const float prod = a[i] * b[i];
        sum += prod;
    }
    return sum;
}

/* ---- QRDecomposition::new (qr.rs:26-106): unblocked Householder, column by
 * column.  qr (m*n) receives the packed result: R strictly above the
 * diagonal, the Householder vector u_k = x - a e_k on and below the diagonal
 * of column k (zero columns are left as they were); rdiag (min(m,n))
 * receives a_k = -sign(x_kk) |x| (+|x| when x_kk <= 0). */
int orc_qr_new(int64_t m, int64_t n, const double *a, double *qr, double *rdiag)
{
    if (m <= 0 || n <= 0) return ORC_SHAPE;        /* Matrix::new asserts */
    memcpy(qr, a, (size_t)(m * n) * sizeof(double));
    const int64_t diag_count = m < n ? m : n;
    for (int64_t minor = 0; minor < diag_count; ++minor) {
        double x_norm_sqr = 0.0;                   /* qr.rs:49-53 */
        for (int64_t i = minor; i < m; ++i) {
            double c = qr[i * n + minor];
            x_norm_sqr = x_norm_sqr + c * c;
        }
        double av = qr[minor * n + minor] > 0.0 ? -sqrt(x_norm_sqr) : sqrt(x_norm_sqr);
        rdiag[minor] = av;                         /* qr.rs:58-59 */
        if (av != 0.0) {
            qr[minor * n + minor] = qr[minor * n + minor] - av;   /* qr.rs:76 */
            for (int64_t column = minor + 1; column < n; ++column) {   /* qr.rs:92-104 */
                double x_dot_u = 0.0;
                for (int64_t row = minor; row < m; ++row) {
                    double p = qr[row * n + minor] * qr[row * n + column];
                    x_dot_u = x_dot_u + p;
                }
                double den = av * qr[minor * n + minor];
                double factor = x_dot_u / den;
                for (int64_t row = minor; row < m; ++row) {
                    double p = factor * qr[row * n + minor];
                    qr[row * n + column] = qr[row * n + column] + p;
                }
            }
        }
    }
    return ORC_OK;
}

/* ---- QRDecomposition::is_full_rank (qr.rs:108-115): 1 / 0; it indexes
 * rdiag[j] for every column j, which is out of bounds (a panic) when n > m. */
int orc_qr_is_full_rank(int64_t m, int64_t n, const double *rdiag)
{
    for (int64_t j = 0; j < n; ++j) {
        if (j >= (m < n ? m : n)) return -1;
        if (rdiag[j] == 0.0) return 0;
    }
    return 1;
}

/* ---- QRDecomposition::get_q (qr.rs:150-194): q (m*m) = H_0 H_1 ... applied
 * to the identity restricted to the first min(m,n) diagonal entries. */
int orc_qr_q(int64_t m, int64_t n, const double *qr, const double *rdiag, double *q)
{
    memset(q, 0, (size_t)(m * m) * sizeof(double));
    const int64_t diag_count = m < n ? m : n;
    for (int64_t minor = 0; minor < diag_count; ++minor) q[minor * m + minor] = 1.0;
    for (int64_t minor = diag_count - 1; minor >= 0; --minor) {
        if (qr[minor * n + minor] != 0.0) {
            for (int64_t column = minor; column < m; ++column) {
                double x_dot_u = 0.0;
                for (int64_t row = minor; row < m; ++row) {
                    double p = qr[row * n + minor] * q[row * m + column];
                    x_dot_u = x_dot_u + p;
                }
                double den = rdiag[minor] * qr[minor * n + minor];
                double factor = x_dot_u / den;
                for (int64_t row = minor; row < m; ++row) {
                    double p = factor * qr[row * n + minor];
                    q[row * m + column] = q[row * m + column] + p;
                }
            }
        }
    }
    return ORC_OK;
}

/* Shared checks of QRDecomposition::solve (qr.rs:196-237): b.rows() == m
 * (assert :193), is_full_rank (:194-196, out-of-bounds panic when n > m). */
static int orc_qr_solve_checks(int64_t m, int64_t n, const double *rdiag, int64_t b_rows)
{
    if (b_rows != m) return ORC_SHAPE;
    int fr = orc_qr_is_full_rank(m, n, rdiag);
    if (fr < 0) return ORC_SHAPE;
    if (fr == 0) return ORC_SINGULAR;
    return ORC_OK;
}

/* ---- QRDecomposition::solve exactly as written (qr.rs:196-237), for m == n
 * (for m > n the reference's Matrix::new(n, nx, data of m*nx) asserts).  Its
 * "Y = Q^T B" step applies x - (<u,x>/u_kk) u, JAMA's formula for vectors
 * normalised to |u|^2 = 2 u_kk, to the vectors of qr.rs:76, whose norm is
 * |u|^2 = -2 a u_kk: it is not the reflection unless a = -1.  Kept to show
 * that defect; orc_qr_solve is the least-squares solve the routine means. */
int orc_qr_solve_literal(int64_t m, int64_t n, const double *qr, const double *rdiag,
                         int64_t b_rows, int64_t nx, const double *b, double *x)
{
    int st = orc_qr_solve_checks(m, n, rdiag, b_rows);
    if (st) return st;
    if (m != n) return ORC_SHAPE;
    memcpy(x, b, (size_t)(m * nx) * sizeof(double));
    for (int64_t k = 0; k < n; ++k)
        for (int64_t j = 0; j < nx; ++j) {
            double s = 0.0;
            for (int64_t i = k; i < m; ++i) {
                double p = qr[i * n + k] * x[i * nx + j];
                s = s + p;
            }
            s = -s / qr[k * n + k];
            for (int64_t i = k; i < m; ++i) {
                double p = s * qr[i * n + k];
                x[i * nx + j] = x[i * nx + j] + p;
            }
        }
    for (int64_t k = n - 1; k >= 0; --k) {
        for (int64_t j = 0; j < nx; ++j) x[k * nx + j] = x[k * nx + j] / rdiag[k];
        for (int64_t i = 0; i < k; ++i)
            for (int64_t j = 0; j < nx; ++j) {
                double p = x[k * nx + j] * qr[i * n + k];
                x[i * nx + j] = x[i * nx + j] - p;
            }
    }
    return ORC_OK;
}

/* ---- The least-squares solve of QRDecomposition::solve with the
 * reflections of qr.rs:76-104 (Hx = x + <x,u>/(a u_kk) u, as in new() and
 * get_q), then R x = (Q^T b)[0..n) with R's diagonal rdiag.  m >= n; x is
 * n*nx. */
int orc_qr_solve(int64_t m, int64_t n, const double *qr, const double *rdiag,
                 int64_t b_rows, int64_t nx, const double *b, double *x)
{
    int st = orc_qr_solve_checks(m, n, rdiag, b_rows);
    if (st) return st;
    double *y = (double *)malloc((size_t)(m * nx) * sizeof(double));
    if (!y) return ORC_SHAPE;
    memcpy(y, b, (size_t)(m * nx) * sizeof(double));
    for (int64_t k = 0; k < n; ++k) {
        if (qr[k * n + k] == 0.0) continue;
        for (int64_t j = 0; j < nx; ++j) {
            double s = 0.0;
            for (int64_t i = k; i < m; ++i) {
                double p = qr[i * n + k] * y[i * nx + j];
                s = s + p;
            }
            double den = rdiag[k] * qr[k * n + k];
            double f = s / den;
            for (int64_t i = k; i < m; ++i) {
                double p = f * qr[i * n + k];
                y[i * nx + j] = y[i * nx + j] + p;
            }
        }
    }
    for (int64_t k = n - 1; k >= 0; --k) {
        for (int64_t j = 0; j < nx; ++j) y[k * nx + j] = y[k * nx + j] / rdiag[k];
        for (int64_t i = 0; i < k; ++i)
            for (int64_t j = 0; j < nx; ++j) {
                double p = y[k * nx + j] * qr[i * n + k];
                y[i * nx + j] = y[i * nx + j] - p;
            }
    }
    memcpy(x, y, (size_t)(n * nx) * sizeof(double));
    free(y);
    return ORC_OK;
}

/* ---- Matrix::pinverse (mod.rs:1049-1057): (R^T R)^-1 A^T with R = get_r()
 * (qr.rs:134-148: m*n, rdiag on the diagonal, zero below), the products by
 * impl Mul and the inverse by LU solve with the identity (mod.rs:1034-1037;
 * .unwrap() on None is a panic: ORC_SINGULAR here).  p is n*m. */
int orc_pinverse(int64_t m, int64_t n, const double *a, double *p)
{
    int st = ORC_OK;
    double *qr = (double *)malloc((size_t)(m * n) * sizeof(double));
    double *rd = (double *)malloc((size_t)((m < n ? m : n)) * sizeof(double));
    double *r = (double *)malloc((size_t)(m * n) * sizeof(double));
    double *rt = (double *)malloc((size_t)(m * n) * sizeof(double));
    double *g = (double *)malloc((size_t)(n * n) * sizeof(double));
    double *lu = (double *)malloc((size_t)(n * n) * sizeof(double));
    uint64_t *piv = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    double *id = (double *)calloc((size_t)(n * n), sizeof(double));
    double *gi = (double *)malloc((size_t)(n * n) * sizeof(double));
    double *at = (double *)malloc((size_t)(m * n) * sizeof(double));
    if (!qr || !rd || !r || !rt || !g || !lu || !piv || !id || !gi || !at) { st = ORC_SHAPE; goto done; }
    if ((st = orc_qr_new(m, n, a, qr, rd))) goto done;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j)
            r[i * n + j] = i < j ? qr[i * n + j] : (i == j ? rd[i] : 0.0);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) rt[j * m + i] = r[i * n + j];
    if ((st = orc_mul(n, m, m, n, rt, r, g))) goto done;
    {
        int sign;
        double gap;
        if ((st = orc_lu_new(n, n, g, lu, piv, &sign, &gap))) goto done;
    }
    for (int64_t i = 0; i < n; ++i) id[i * n + i] = 1.0;
    if ((st = orc_lu_solve(n, n, lu, piv, n, n, id, gi))) goto done;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) at[j * m + i] = a[i * n + j];
    st = orc_mul(n, n, n, m, gi, at, p);
done:
    free(qr); free(rd); free(r); free(rt); free(g); free(lu); free(piv); free(id); free(gi); free(at);
    return st;
}
