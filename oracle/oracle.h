/* oracle.h — CPU restatement of the reference LU path (getf2/getrf, laswp,
 * getrs, gesv, getri, the Cholesky potf2/potrs/posv, and the
 * mixed-precision refinement of LAPACK dsgesv).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Storage is column-major, A[i + j*lda]; pivots are 1-based (LAPACK and
 * rocsolver convention).  Every routine works on ONE matrix; batching is the
 * caller's loop.
 */
#ifndef ROCSOLVER_AMD_ORACLE_H
#define ROCSOLVER_AMD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Unblocked right-looking LU with partial pivoting.  Returns info (first zero
 * pivot, 1-based, or 0).  If gap != NULL, gap[k] receives the relative
 * separation (|a_p| - |a_second|) / |a_p| of the two largest candidates of
 * column k (1.0 when there is a single candidate, 0.0 for a zero pivot). */
int orc_dgetf2(int m, int n, double* A, int lda, int* ipiv, double* gap);
int orc_sgetf2(int m, int n, float* A, int lda, int* ipiv, double* gap);

/* LU without pivoting (the reference npvt small kernel): a zero diagonal
 * pivot sets info (first, 1-based) and multiplies the column below it by
 * zero.  Returns info. */
int orc_dgetf2_npvt(int m, int n, double* A, int lda);
int orc_sgetf2_npvt(int m, int n, float* A, int lda);

/* Row interchanges, reference laswp semantics (k1, k2 1-based, incx = +-1). */
void orc_dlaswp(int n, double* A, int lda, int k1, int k2, const int* ipiv, int incx);
void orc_slaswp(int n, float* A, int lda, int k1, int k2, const int* ipiv, int incx);

/* Solve op(A) X = B with the LU from getf2.  trans: 'N' or 'T' ('C' == 'T'
 * for real data). */
void orc_dgetrs(char trans, int n, int nrhs, const double* A, int lda, const int* ipiv,
                double* B, int ldb);
void orc_sgetrs(char trans, int n, int nrhs, const float* A, int lda, const int* ipiv,
                float* B, int ldb);

/* getf2 + getrs; B is left unchanged when info != 0.  Returns info. */
int orc_dgesv(int n, int nrhs, double* A, int lda, int* ipiv, double* B, int ldb);
int orc_sgesv(int n, int nrhs, float* A, int lda, int* ipiv, float* B, int ldb);

/* Inverse from the getf2 output (LU, ipiv), in place (LAPACK dgetri, unblocked).
 * ipiv NULL: no column interchanges (the no-pivoting variants).  Returns info
 * (first zero diagonal of U, 1-based; A untouched then) or 0. */
int orc_dgetri(int n, double* A, int lda, const int* ipiv);
int orc_sgetri(int n, float* A, int lda, const int* ipiv);

/* Cholesky factorisation, uplo 'L' (A = L L^T) or 'U' (A = U^T U), of the
 * referenced triangle in place.  A non-positive pivot t sets info (first,
 * 1-based), stores t on the diagonal, scales its column (row) by 0 and the
 * factorisation goes on, as the reference does.  Returns info. */
int orc_dpotf2(char uplo, int n, double* A, int lda);
int orc_spotf2(char uplo, int n, float* A, int lda);
/* Solve A X = B with the factor of potf2. */
void orc_dpotrs(char uplo, int n, int nrhs, const double* A, int lda, double* B, int ldb);
void orc_spotrs(char uplo, int n, int nrhs, const float* A, int lda, float* B, int ldb);
/* potf2 + potrs; B unchanged when info != 0.  Returns info. */
int orc_dposv(char uplo, int n, int nrhs, double* A, int lda, double* B, int ldb);
int orc_sposv(char uplo, int n, int nrhs, float* A, int lda, float* B, int ldb);

/* Mixed-precision solve (LAPACK dsgesv semantics): FP32 LU of A, FP32
 * correction solves, FP64 residual and update, stop when every column
 * satisfies ||r||_inf <= ||x||_inf * ||A||_inf * sqrt(n) * eps (eps = 2^-53),
 * at most itermax refinement steps, otherwise a full FP64 getf2/getrs.
 * A is not modified.  X (ldx) receives the solution.  Returns info of the
 * path that produced X; *iter follows dsgesv: >=0 refinement steps used,
 * -3 FP32 LU hit a zero pivot, -(itermax+1) no convergence. */
int orc_dsgesv(int n, int nrhs, const double* A, int lda, const double* B, int ldb,
               double* X, int ldx, int itermax, int* iter);

/* Residual R = B - op(A) X accumulated in long double, returned in double. */
void orc_dresidual(char trans, int n, int nrhs, const double* A, int lda, const double* X,
                   int ldx, const double* B, int ldb, double* R, int ldr);

/* ||A||_F and ||A||_inf (long double accumulation). */
double orc_dnorm_fro(int m, int n, const double* A, int lda);
double orc_dnorm_inf(int m, int n, const double* A, int lda);

/* ||P*A - L*U||_F for the getf2 output (LU, ipiv) of the original A. */
double orc_dlu_residual(int m, int n, const double* A0, int lda0, const double* LU, int lda,
                        const int* ipiv);

#ifdef __cplusplus
}
#endif
#endif
