/*
 * oracle_blas.h -- dense kernels for the CPU oracle.
 *
 * TEST INFRASTRUCTURE: the oracle is the checker and the CPU baseline; it is
 * never linked into the product library.
 *
 * The reference calls CBLAS (cblas_dgemm/dtrsm/dsyrk, kernels/core_z*.c).  At
 * load time the oracle looks for a CBLAS in the process (libmkl_rt.so, then an
 * OpenBLAS); when none is found it uses the portable loops below, which compute
 * the same products in a fixed order.  Which one ran is reported by
 * orc_blas_name().
 */
#ifndef ORACLE_BLAS_H
#define ORACLE_BLAS_H

#include <stdint.h>

enum { ORC_NOTRANS = 111, ORC_TRANS = 112 };

/* C = alpha * A * op(B) + beta * C, column major (op = B or B^T) */
typedef void ( *orc_dgemm_fn )( int transb, int64_t m, int64_t n, int64_t k, double alpha,
                                const double *A, int64_t lda, const double *B, int64_t ldb,
                                double beta, double *C, int64_t ldc );
typedef void ( *orc_sgemm_fn )( int transb, int64_t m, int64_t n, int64_t k, float alpha,
                                const float *A, int64_t lda, const float *B, int64_t ldb,
                                float beta, float *C, int64_t ldc );
/* B = B * L^{-T}, L lower n x n non unit, B m x n */
typedef void ( *orc_dtrsm_fn )( int64_t m, int64_t n, const double *L, int64_t ldl, double *B, int64_t ldb );
typedef void ( *orc_strsm_fn )( int64_t m, int64_t n, const float *L, int64_t ldl, float *B, int64_t ldb );

void         orc_blas_init( void );
const char  *orc_blas_name( void );
extern orc_dgemm_fn orc_dgemm;
extern orc_sgemm_fn orc_sgemm;
extern orc_dtrsm_fn orc_dtrsm;
extern orc_strsm_fn orc_strsm;

#endif
