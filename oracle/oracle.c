/* oracle.c — CPU restatement of the reference batched-LU path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product library never links,
 * loads or calls this file; tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are its only users.
 *
 * Each routine restates one reference algorithm, cited file:line
 * (paths relative to the rocSOLVER source tree):
 *
 *  getf2   library/src/specialized/roclapack_getf2_specialized_kernels.hpp:72-126
 *          (small kernel: pivot = first max |a| with a strict '<' scan, so ties
 *          go to the lowest row and NaN never displaces the candidate; zero
 *          pivot -> info = first k+1, no swap; reciprocal-then-multiply scaling;
 *          rank-1 update a -= l*u, contracted to one fma as hipcc does)
 *          library/src/lapack/roclapack_getf2.hpp:41-99 (generic path: row swap
 *          across all n columns, ipiv = pivot+1, info rule)
 *  laswp   library/src/auxiliary/rocauxiliary_laswp.hpp:41-62
 *  getrs   library/src/lapack/roclapack_getrs.hpp:117-150 (order of laswp and
 *          the two triangular solves for N and T) and
 *          library/src/specialized/roclapack_trsm_specialized_kernels.hpp:166-566
 *          (column-oriented substitution; non-unit diagonal by true division)
 *  gesv    library/src/lapack/roclapack_gesv.hpp:145-175 (B unchanged where
 *          info != 0)
 *  getri   library/src/lapack/roclapack_getri.hpp:265-378 (inv(U) in place by
 *          trtri, upper non-unit, whose singularity check leaves info = first
 *          zero diagonal; then inv(A) from inv(A)*L = inv(U) column by column
 *          from the right; then the column interchanges of ipiv in reverse,
 *          getri_pivot :56-70), in LAPACK dgetri's unblocked form (dtrti2 +
 *          dgemv sweep)
 *  potf2   library/src/lapack/roclapack_potf2.hpp:10-37,160-250 (left-looking:
 *          t = a_jj - dot; t <= 0 -> info = first j+1, a_jj = t, scale 0, and
 *          the factorisation goes on; else a_jj = sqrt(t), then the rest of
 *          column j (lower) / row j (upper) minus a gemv, times 1/a_jj).  The
 *          dot and gemv sums (rocBLAS reductions of unspecified order there)
 *          are taken here term by term in increasing k, one fma each
 *  potrs   library/src/lapack/roclapack_potrs.hpp:121-148 (L then L^T, or U^T
 *          then U, non-unit triangular solves)
 *  posv    library/src/lapack/roclapack_posv.hpp:140-170 (potrf, potrs, B
 *          unchanged where info != 0)
 *  dsgesv  no rocSOLVER counterpart; LAPACK 3.x dsgesv.f semantics (range
 *          check -> ITER=-2, FP32 LU zero pivot -> ITER=-3, ITERMAX steps,
 *          converged when ||r||_inf <= ||x||_inf ||A||_inf sqrt(n) eps,
 *          otherwise ITER=-(ITERMAX+1) and an FP64 getf2/getrs)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------ getf2 */
#define DEF_GETF2(T, P, ABS, FMA)                                                       \
    int orc_##P##getf2(int m, int n, T* A, int lda, int* ipiv, double* gap)            \
    {                                                                                   \
        int info = 0;                                                                   \
        int kmax = m < n ? m : n;                                                       \
        for(int k = 0; k < kmax; ++k)                                                   \
        {                                                                               \
            int p = k;                                                                  \
            T pv = A[IDX(k, k, lda)];                                                   \
            for(int i = k + 1; i < m; ++i)                                              \
            {                                                                           \
                T t = A[IDX(i, k, lda)];                                                \
                if(ABS(pv) < ABS(t))                                                    \
                {                                                                       \
                    pv = t;                                                             \
                    p = i;                                                              \
                }                                                                       \
            }                                                                           \
            if(gap)                                                                     \
            {                                                                           \
                double best = ABS(pv), second = -1.0;                                  \
                for(int i = k; i < m; ++i)                                              \
                    if(i != p && (double)ABS(A[IDX(i, k, lda)]) > second)               \
                        second = ABS(A[IDX(i, k, lda)]);                                \
                gap[k] = best == 0.0 ? 0.0 : (second < 0.0 ? 1.0 : (best - second) / best); \
            }                                                                           \
            ipiv[k] = p + 1;                                                            \
            if(pv != (T)0)                                                              \
            {                                                                           \
                if(p != k)                                                              \
                    for(int j = 0; j < n; ++j)                                          \
                    {                                                                   \
                        T t = A[IDX(k, j, lda)];                                        \
                        A[IDX(k, j, lda)] = A[IDX(p, j, lda)];                          \
                        A[IDX(p, j, lda)] = t;                                          \
                    }                                                                   \
                T r = (T)1 / pv;                                                        \
                for(int i = k + 1; i < m; ++i)                                          \
                    A[IDX(i, k, lda)] *= r;                                             \
            }                                                                           \
            else if(info == 0)                                                          \
                info = k + 1;                                                           \
            for(int j = k + 1; j < n; ++j)                                              \
            {                                                                           \
                T u = A[IDX(k, j, lda)];                                                \
                for(int i = k + 1; i < m; ++i)                                          \
                    A[IDX(i, j, lda)] = FMA(-A[IDX(i, k, lda)], u, A[IDX(i, j, lda)]);  \
            }                                                                           \
        }                                                                               \
        return info;                                                                    \
    }

DEF_GETF2(double, d, fabs, fma)
DEF_GETF2(float, s, fabsf, fmaf)

/* ------------------------------------------------------------- getf2_npvt
 * No-pivoting LU, restating the reference small kernel
 * library/src/specialized/roclapack_getf2_specialized_kernels.hpp:138-215
 * (getf2_npvt_small_kernel, the path of m <= 64 batched npvt): the pivot is
 * the diagonal entry; its reciprocal scales the column below it, and a zero
 * pivot leaves the scale at 0 (:185-188), so the whole sub-column is
 * multiplied by zero; info = first zero pivot; rank-1 update
 * a -= l*u as one fma. */
#define DEF_GETF2_NPVT(T, P, FMA)                                                       \
    int orc_##P##getf2_npvt(int m, int n, T* A, int lda)                               \
    {                                                                                   \
        int info = 0;                                                                   \
        int kmax = m < n ? m : n;                                                       \
        for(int k = 0; k < kmax; ++k)                                                   \
        {                                                                               \
            T pv = A[IDX(k, k, lda)];                                                   \
            T r = (T)0;                                                                 \
            if(pv != (T)0)                                                              \
                r = (T)1 / pv;                                                          \
            else if(info == 0)                                                          \
                info = k + 1;                                                           \
            for(int i = k + 1; i < m; ++i)                                              \
                A[IDX(i, k, lda)] *= r;                                                 \
            for(int j = k + 1; j < n; ++j)                                              \
            {                                                                           \
                T u = A[IDX(k, j, lda)];                                                \
                for(int i = k + 1; i < m; ++i)                                          \
                    A[IDX(i, j, lda)] = FMA(-A[IDX(i, k, lda)], u, A[IDX(i, j, lda)]);  \
            }                                                                           \
        }                                                                               \
        return info;                                                                    \
    }

DEF_GETF2_NPVT(double, d, fma)
DEF_GETF2_NPVT(float, s, fmaf)

/* ------------------------------------------------------------------ laswp */
#define DEF_LASWP(T, P)                                                                 \
    void orc_##P##laswp(int n, T* A, int lda, int k1, int k2, const int* ipiv, int incx) \
    {                                                                                   \
        int start, end, inc;                                                            \
        if(incx < 0)                                                                    \
        {                                                                               \
            start = k2;                                                                 \
            end = k1 - 1;                                                               \
            inc = -1;                                                                   \
            incx = -incx;                                                               \
        }                                                                               \
        else                                                                            \
        {                                                                               \
            start = k1;                                                                 \
            end = k2 + 1;                                                               \
            inc = 1;                                                                    \
        }                                                                               \
        for(int c = 0; c < n; ++c)                                                      \
            for(int i = start; i != end; i += inc)                                      \
            {                                                                           \
                int ex = ipiv[k1 + (i - k1) * incx - 1];                                \
                if(ex != i)                                                             \
                {                                                                       \
                    T t = A[IDX(i - 1, c, lda)];                                        \
                    A[IDX(i - 1, c, lda)] = A[IDX(ex - 1, c, lda)];                     \
                    A[IDX(ex - 1, c, lda)] = t;                                         \
                }                                                                       \
            }                                                                           \
    }

DEF_LASWP(double, d)
DEF_LASWP(float, s)

/* ------------------------------------------------------------------ getrs */
#define DEF_GETRS(T, P, FMA)                                                            \
    void orc_##P##getrs(char trans, int n, int nrhs, const T* A, int lda, const int* ipiv, \
                        T* B, int ldb)                                                  \
    {                                                                                   \
        if(n == 0 || nrhs == 0)                                                         \
            return;                                                                     \
        int notrans = (trans == 'N' || trans == 'n');                                   \
        if(notrans)                                                                     \
            orc_##P##laswp(nrhs, B, ldb, 1, n, ipiv, 1);                                \
        for(int c = 0; c < nrhs; ++c)                                                   \
        {                                                                               \
            T* b = B + IDX(0, c, ldb);                                                  \
            if(notrans)                                                                 \
            {                                                                           \
                /* L y = b, unit lower, column oriented */                              \
                for(int k = 0; k < n - 1; ++k)                                          \
                    for(int x = k + 1; x < n; ++x)                                      \
                        b[x] = FMA(-A[IDX(x, k, lda)], b[k], b[x]);                     \
                /* U x = y, non-unit upper */                                           \
                for(int k = n - 1; k >= 0; --k)                                         \
                {                                                                       \
                    b[k] = b[k] / A[IDX(k, k, lda)];                                    \
                    for(int x = 0; x < k; ++x)                                          \
                        b[x] = FMA(-A[IDX(x, k, lda)], b[k], b[x]);                     \
                }                                                                       \
            }                                                                           \
            else                                                                        \
            {                                                                           \
                /* U^T y = b, non-unit lower */                                         \
                for(int k = 0; k < n; ++k)                                              \
                {                                                                       \
                    b[k] = b[k] / A[IDX(k, k, lda)];                                    \
                    for(int x = k + 1; x < n; ++x)                                      \
                        b[x] = FMA(-A[IDX(k, x, lda)], b[k], b[x]);                     \
                }                                                                       \
                /* L^T x = y, unit upper */                                             \
                for(int k = n - 1; k > 0; --k)                                          \
                    for(int x = 0; x < k; ++x)                                          \
                        b[x] = FMA(-A[IDX(k, x, lda)], b[k], b[x]);                     \
            }                                                                           \
        }                                                                               \
        if(!notrans)                                                                    \
            orc_##P##laswp(nrhs, B, ldb, 1, n, ipiv, -1);                               \
    }

DEF_GETRS(double, d, fma)
DEF_GETRS(float, s, fmaf)

/* ------------------------------------------------------------------- gesv */
#define DEF_GESV(T, P)                                                                  \
    int orc_##P##gesv(int n, int nrhs, T* A, int lda, int* ipiv, T* B, int ldb)        \
    {                                                                                   \
        int info = orc_##P##getf2(n, n, A, lda, ipiv, NULL);                            \
        if(info == 0)                                                                   \
            orc_##P##getrs('N', n, nrhs, A, lda, ipiv, B, ldb);                         \
        return info;                                                                    \
    }

DEF_GESV(double, d)
DEF_GESV(float, s)

/* ------------------------------------------------------------------ getri */
#define DEF_GETRI(T, P, FMA)                                                            \
    int orc_##P##getri(int n, T* A, int lda, const int* ipiv)                         \
    {                                                                                   \
        for(int j = 0; j < n; ++j)                                                      \
            if(A[IDX(j, j, lda)] == (T)0)                                               \
                return j + 1;                                                           \
        /* inv(U): column j of the inverse from the inverted leading block */           \
        for(int j = 0; j < n; ++j)                                                      \
        {                                                                               \
            A[IDX(j, j, lda)] = (T)1 / A[IDX(j, j, lda)];                               \
            T ajj = -A[IDX(j, j, lda)];                                                 \
            for(int jj = 0; jj < j; ++jj)                                               \
            {                                                                           \
                T t = A[IDX(jj, j, lda)];                                               \
                for(int i = 0; i < jj; ++i)                                             \
                    A[IDX(i, j, lda)] = FMA(t, A[IDX(i, jj, lda)], A[IDX(i, j, lda)]);  \
                A[IDX(jj, j, lda)] = t * A[IDX(jj, jj, lda)];                           \
            }                                                                           \
            for(int i = 0; i < j; ++i)                                                  \
                A[IDX(i, j, lda)] *= ajj;                                               \
        }                                                                               \
        /* inv(A) * L = inv(U), columns from the right */                               \
        T* work = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                      \
        for(int j = n - 1; j >= 0; --j)                                                 \
        {                                                                               \
            for(int i = j + 1; i < n; ++i)                                              \
            {                                                                           \
                work[i] = A[IDX(i, j, lda)];                                            \
                A[IDX(i, j, lda)] = (T)0;                                               \
            }                                                                           \
            for(int jj = j + 1; jj < n; ++jj)                                           \
                for(int i = 0; i < n; ++i)                                              \
                    A[IDX(i, j, lda)] = FMA(-A[IDX(i, jj, lda)], work[jj], A[IDX(i, j, lda)]); \
        }                                                                               \
        free(work);                                                                     \
        if(ipiv)                                                                        \
            for(int j = n - 2; j >= 0; --j)                                             \
            {                                                                           \
                int jp = ipiv[j] - 1;                                                   \
                if(jp != j)                                                             \
                    for(int i = 0; i < n; ++i)                                          \
                    {                                                                   \
                        T t = A[IDX(i, j, lda)];                                        \
                        A[IDX(i, j, lda)] = A[IDX(i, jp, lda)];                         \
                        A[IDX(i, jp, lda)] = t;                                         \
                    }                                                                   \
            }                                                                           \
        return 0;                                                                       \
    }

DEF_GETRI(double, d, fma)
DEF_GETRI(float, s, fmaf)

/* ------------------------------------------------------------ potf2 etc. */
/* element (i, j) of the lower factor L: A(i, j) (lower) or A(j, i) (upper,
 * where the factor is U = L^T) */
#define LEL(i, j) (*(up ? &A[IDX(j, i, lda)] : &A[IDX(i, j, lda)]))
#define DEF_POTF2(T, P, SQRT, FMA)                                                      \
    int orc_##P##potf2(char uplo, int n, T* A, int lda)                                \
    {                                                                                   \
        int up = (uplo == 'U' || uplo == 'u');                                          \
        int info = 0;                                                                   \
        for(int j = 0; j < n; ++j)                                                      \
        {                                                                               \
            /* t = a_jj - sum_k l_jk^2, subtracted term by term in k order */            \
            T t = LEL(j, j);                                                            \
            for(int k = 0; k < j; ++k)                                                  \
                t = FMA(-LEL(j, k), LEL(j, k), t);                                      \
            T r;                                                                        \
            if(t <= (T)0)                                                               \
            {                                                                           \
                if(info == 0)                                                           \
                    info = j + 1;                                                       \
                LEL(j, j) = t;                                                          \
                r = (T)0;                                                               \
            }                                                                           \
            else                                                                        \
            {                                                                           \
                LEL(j, j) = SQRT(t);                                                    \
                r = (T)1 / LEL(j, j);                                                   \
            }                                                                           \
            for(int i = j + 1; i < n; ++i)                                              \
            {                                                                           \
                T v = LEL(i, j);                                                        \
                for(int k = 0; k < j; ++k)                                              \
                    v = FMA(-LEL(i, k), LEL(j, k), v);                                  \
                LEL(i, j) = v * r;                                                      \
            }                                                                           \
        }                                                                               \
        return info;                                                                    \
    }                                                                                   \
    void orc_##P##potrs(char uplo, int n, int nrhs, const T* A, int lda, T* B, int ldb) \
    {                                                                                   \
        int up = (uplo == 'U' || uplo == 'u');                                          \
        for(int c = 0; c < nrhs; ++c)                                                   \
        {                                                                               \
            T* b = B + IDX(0, c, ldb);                                                  \
            /* L y = b (column oriented) */                                             \
            for(int k = 0; k < n; ++k)                                                  \
            {                                                                           \
                b[k] = b[k] / LEL(k, k);                                                \
                for(int i = k + 1; i < n; ++i)                                          \
                    b[i] = FMA(-LEL(i, k), b[k], b[i]);                                 \
            }                                                                           \
            /* L^T x = y */                                                             \
            for(int k = n - 1; k >= 0; --k)                                             \
            {                                                                           \
                b[k] = b[k] / LEL(k, k);                                                \
                for(int i = 0; i < k; ++i)                                              \
                    b[i] = FMA(-LEL(k, i), b[k], b[i]);                                 \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
    int orc_##P##posv(char uplo, int n, int nrhs, T* A, int lda, T* B, int ldb)        \
    {                                                                                   \
        int info = orc_##P##potf2(uplo, n, A, lda);                                     \
        if(info == 0)                                                                   \
            orc_##P##potrs(uplo, n, nrhs, A, lda, B, ldb);                              \
        return info;                                                                    \
    }

DEF_POTF2(double, d, sqrt, fma)
DEF_POTF2(float, s, sqrtf, fmaf)
#undef LEL

/* ------------------------------------------------------------- residuals */
void orc_dresidual(char trans, int n, int nrhs, const double* A, int lda, const double* X,
                   int ldx, const double* B, int ldb, double* R, int ldr)
{
    int notrans = (trans == 'N' || trans == 'n');
    for(int c = 0; c < nrhs; ++c)
        for(int i = 0; i < n; ++i)
        {
            long double s = B[IDX(i, c, ldb)];
            for(int j = 0; j < n; ++j)
            {
                double a = notrans ? A[IDX(i, j, lda)] : A[IDX(j, i, lda)];
                s -= (long double)a * (long double)X[IDX(j, c, ldx)];
            }
            R[IDX(i, c, ldr)] = (double)s;
        }
}

double orc_dnorm_fro(int m, int n, const double* A, int lda)
{
    long double s = 0;
    for(int j = 0; j < n; ++j)
        for(int i = 0; i < m; ++i)
            s += (long double)A[IDX(i, j, lda)] * A[IDX(i, j, lda)];
    return (double)sqrtl(s);
}

double orc_dnorm_inf(int m, int n, const double* A, int lda)
{
    long double best = 0;
    for(int i = 0; i < m; ++i)
    {
        long double s = 0;
        for(int j = 0; j < n; ++j)
            s += fabsl((long double)A[IDX(i, j, lda)]);
        if(s > best || s != s)
            best = s;
    }
    return (double)best;
}

double orc_dlu_residual(int m, int n, const double* A0, int lda0, const double* LU, int lda,
                        const int* ipiv)
{
    /* apply the row interchanges to a copy of A0, then subtract L*U */
    int kmax = m < n ? m : n;
    double* PA = (double*)malloc(sizeof(double) * (size_t)m * (size_t)(n > 0 ? n : 1));
    for(int j = 0; j < n; ++j)
        for(int i = 0; i < m; ++i)
            PA[IDX(i, j, m)] = A0[IDX(i, j, lda0)];
    if(kmax > 0)
        orc_dlaswp(n, PA, m, 1, kmax, ipiv, 1);
    long double s = 0;
    for(int j = 0; j < n; ++j)
        for(int i = 0; i < m; ++i)
        {
            long double acc = 0;
            int kk = i < j ? i : j;
            if(kk >= kmax)
                kk = kmax - 1;
            for(int k = 0; k <= kk && k < kmax; ++k)
            {
                long double l = (i == k) ? 1.0L : (i > k ? LU[IDX(i, k, lda)] : 0.0L);
                long double u = (k <= j) ? LU[IDX(k, j, lda)] : 0.0L;
                acc += l * u;
            }
            long double d = (long double)PA[IDX(i, j, m)] - acc;
            s += d * d;
        }
    free(PA);
    return (double)sqrtl(s);
}

/* ---------------------------------------------------------------- dsgesv */
int orc_dsgesv(int n, int nrhs, const double* A, int lda, const double* B, int ldb,
               double* X, int ldx, int itermax, int* iter)
{
    /* LAPACK dsgesv.f: EPS = DLAMCH('Epsilon') = 2^-53, CTE = ||A||_inf eps sqrt(n) */
    const double eps = 0x1p-53;
    *iter = 0;
    if(n == 0 || nrhs == 0)
        return 0;
    double anrm = 0;
    for(int i = 0; i < n; ++i)
    {
        double s = 0;
        for(int j = 0; j < n; ++j)
            s += fabs(A[IDX(i, j, lda)]);
        if(s > anrm)
            anrm = s;
    }
    const double cte = anrm * eps * sqrt((double)n);
    float* SA = (float*)malloc(sizeof(float) * (size_t)n * n);
    float* SR = (float*)malloc(sizeof(float) * (size_t)n * nrhs);
    double* R = (double*)malloc(sizeof(double) * (size_t)n * nrhs);
    int* ipiv = (int*)malloc(sizeof(int) * (size_t)n);
    int info = 0, converged = 0;

    /* DLAG2S: an entry outside the FP32 range sends the solve to FP64 (ITER=-2) */
    for(int j = 0; j < n; ++j)
        for(int i = 0; i < n; ++i)
        {
            double a = A[IDX(i, j, lda)];
            if(a > 0x1.fffffep127 || a < -0x1.fffffep127)
            {
                *iter = -2;
                goto fallback;
            }
            SA[IDX(i, j, n)] = (float)a;
        }
    for(int j = 0; j < nrhs; ++j)
        for(int i = 0; i < n; ++i)
        {
            double b = B[IDX(i, j, ldb)];
            if(b > 0x1.fffffep127 || b < -0x1.fffffep127)
            {
                *iter = -2;
                goto fallback;
            }
            SR[IDX(i, j, n)] = (float)b;
        }

    if(orc_sgetf2(n, n, SA, n, ipiv, NULL) != 0)
    {
        *iter = -3;
        goto fallback;
    }
    orc_sgetrs('N', n, nrhs, SA, n, ipiv, SR, n);
    for(int j = 0; j < nrhs; ++j)
        for(int i = 0; i < n; ++i)
            X[IDX(i, j, ldx)] = (double)SR[IDX(i, j, n)];

    for(int it = 0; it <= itermax; ++it)
    {
        /* R = B - A X in FP64 (plain double accumulation, as DGEMM) */
        for(int c = 0; c < nrhs; ++c)
            for(int i = 0; i < n; ++i)
            {
                double s = B[IDX(i, c, ldb)];
                for(int j = 0; j < n; ++j)
                    s = fma(-A[IDX(i, j, lda)], X[IDX(j, c, ldx)], s);
                R[IDX(i, c, n)] = s;
            }
        converged = 1;
        for(int c = 0; c < nrhs; ++c)
        {
            double xn = 0, rn = 0;
            for(int i = 0; i < n; ++i)
            {
                if(fabs(X[IDX(i, c, ldx)]) > xn)
                    xn = fabs(X[IDX(i, c, ldx)]);
                if(fabs(R[IDX(i, c, n)]) > rn)
                    rn = fabs(R[IDX(i, c, n)]);
            }
            if(rn > xn * cte) /* dsgesv.f: IF( RNRM.GT.XNRM*CTE ) not converged */
                converged = 0;
        }
        if(converged)
        {
            *iter = it;
            goto done;
        }
        if(it == itermax)
            break;
        for(int j = 0; j < nrhs; ++j)
            for(int i = 0; i < n; ++i)
                SR[IDX(i, j, n)] = (float)R[IDX(i, j, n)];
        orc_sgetrs('N', n, nrhs, SA, n, ipiv, SR, n);
        for(int j = 0; j < nrhs; ++j)
            for(int i = 0; i < n; ++i)
                X[IDX(i, j, ldx)] += (double)SR[IDX(i, j, n)];
    }
    *iter = -(itermax + 1);

fallback:
{
    double* DA = (double*)malloc(sizeof(double) * (size_t)n * n);
    for(int j = 0; j < n; ++j)
        for(int i = 0; i < n; ++i)
            DA[IDX(i, j, n)] = A[IDX(i, j, lda)];
    for(int j = 0; j < nrhs; ++j)
        for(int i = 0; i < n; ++i)
            X[IDX(i, j, ldx)] = B[IDX(i, j, ldb)];
    info = orc_dgesv(n, nrhs, DA, n, ipiv, X, ldx);
    free(DA);
}
done:
    free(SA);
    free(SR);
    free(R);
    free(ipiv);
    return info;
}
