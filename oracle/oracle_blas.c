/*
 * oracle_blas.c -- CBLAS lookup and portable fallbacks for the CPU oracle.
 * TEST INFRASTRUCTURE (checker and CPU baseline only).
 */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include "oracle_blas.h"

typedef void ( *cblas_dgemm_t )( int, int, int, int, int, int, double, const double *, int,
                                 const double *, int, double, double *, int );
typedef void ( *cblas_sgemm_t )( int, int, int, int, int, int, float, const float *, int,
                                 const float *, int, float, float *, int );
typedef void ( *cblas_dtrsm_t )( int, int, int, int, int, int, int, double, const double *, int,
                                 double *, int );
typedef void ( *cblas_strsm_t )( int, int, int, int, int, int, int, float, const float *, int,
                                 float *, int );

static cblas_dgemm_t p_dgemm;
static cblas_sgemm_t p_sgemm;
static cblas_dtrsm_t p_dtrsm;
static cblas_strsm_t p_strsm;
static const char   *blas_name = "portable loops";

#define CB_COL 102
#define CB_NT  111
#define CB_T   112
#define CB_R   142
#define CB_LO  122
#define CB_NU  131

/* ---- portable loops ---- */
#define DEF_GEMM( NAME, T )                                                                        \
    static void NAME( int transb, int64_t m, int64_t n, int64_t k, T alpha, const T *A,           \
                      int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc )            \
    {                                                                                              \
        for ( int64_t j = 0; j < n; j++ ) {                                                        \
            T *c = C + j * ldc;                                                                    \
            if ( beta != (T)1 ) {                                                                  \
                for ( int64_t i = 0; i < m; i++ ) c[i] = ( beta == (T)0 ) ? (T)0 : beta * c[i];    \
            }                                                                                      \
            for ( int64_t l = 0; l < k; l++ ) {                                                    \
                T b = ( transb == ORC_TRANS ) ? B[j + l * ldb] : B[l + j * ldb];                   \
                T s = alpha * b;                                                                   \
                const T *a = A + l * lda;                                                          \
                for ( int64_t i = 0; i < m; i++ ) c[i] += a[i] * s;                                \
            }                                                                                      \
        }                                                                                          \
    }

#define DEF_TRSM( NAME, T )                                                                        \
    static void NAME( int64_t m, int64_t n, const T *L, int64_t ldl, T *B, int64_t ldb )           \
    {                                                                                              \
        /* X L^T = B: column j of X = (B_j - sum_{l<j} X_l L_jl) / L_jj */                         \
        for ( int64_t j = 0; j < n; j++ ) {                                                        \
            T *bj = B + j * ldb;                                                                   \
            for ( int64_t l = 0; l < j; l++ ) {                                                    \
                T        f  = L[j + l * ldl];                                                      \
                const T *bl = B + l * ldb;                                                         \
                for ( int64_t i = 0; i < m; i++ ) bj[i] -= bl[i] * f;                              \
            }                                                                                      \
            T d = L[j + j * ldl];                                                                  \
            for ( int64_t i = 0; i < m; i++ ) bj[i] /= d;                                          \
        }                                                                                          \
    }

DEF_GEMM( loop_dgemm, double )
DEF_GEMM( loop_sgemm, float )
DEF_TRSM( loop_dtrsm, double )
DEF_TRSM( loop_strsm, float )

/* ---- CBLAS wrappers ---- */
static void
cb_dgemm( int transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
          const double *B, int64_t ldb, double beta, double *C, int64_t ldc )
{
    p_dgemm( CB_COL, CB_NT, transb == ORC_TRANS ? CB_T : CB_NT, (int)m, (int)n, (int)k, alpha, A,
             (int)lda, B, (int)ldb, beta, C, (int)ldc );
}
static void
cb_sgemm( int transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
          const float *B, int64_t ldb, float beta, float *C, int64_t ldc )
{
    p_sgemm( CB_COL, CB_NT, transb == ORC_TRANS ? CB_T : CB_NT, (int)m, (int)n, (int)k, alpha, A,
             (int)lda, B, (int)ldb, beta, C, (int)ldc );
}
static void
cb_dtrsm( int64_t m, int64_t n, const double *L, int64_t ldl, double *B, int64_t ldb )
{
    p_dtrsm( CB_COL, CB_R, CB_LO, CB_T, CB_NU, (int)m, (int)n, 1.0, L, (int)ldl, B, (int)ldb );
}
static void
cb_strsm( int64_t m, int64_t n, const float *L, int64_t ldl, float *B, int64_t ldb )
{
    p_strsm( CB_COL, CB_R, CB_LO, CB_T, CB_NU, (int)m, (int)n, 1.0f, L, (int)ldl, B, (int)ldb );
}

orc_dgemm_fn orc_dgemm = loop_dgemm;
orc_sgemm_fn orc_sgemm = loop_sgemm;
orc_dtrsm_fn orc_dtrsm = loop_dtrsm;
orc_strsm_fn orc_strsm = loop_strsm;

void
orc_blas_init( void )
{
    static int done = 0;
    const char *cands[] = { "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libopenblas.so.0", NULL };
    if ( done ) {
        return;
    }
    done = 1;
    if ( getenv( "ORACLE_NO_BLAS" ) ) {
        return;
    }
    for ( int i = 0; cands[i]; i++ ) {
        void *h = dlopen( cands[i], RTLD_NOW | RTLD_LOCAL );
        if ( h == NULL ) {
            continue;
        }
        p_dgemm = (cblas_dgemm_t)dlsym( h, "cblas_dgemm" );
        p_sgemm = (cblas_sgemm_t)dlsym( h, "cblas_sgemm" );
        p_dtrsm = (cblas_dtrsm_t)dlsym( h, "cblas_dtrsm" );
        p_strsm = (cblas_strsm_t)dlsym( h, "cblas_strsm" );
        if ( p_dgemm && p_sgemm && p_dtrsm && p_strsm ) {
            /* one BLAS thread per scheduler thread, as the reference runs it */
            void ( *setn )( int ) = (void ( * )( int ))dlsym( h, "MKL_Set_Num_Threads" );
            if ( setn == NULL ) setn = (void ( * )( int ))dlsym( h, "openblas_set_num_threads" );
            if ( setn ) setn( 1 );
            orc_dgemm = cb_dgemm;
            orc_sgemm = cb_sgemm;
            orc_dtrsm = cb_dtrsm;
            orc_strsm = cb_strsm;
            blas_name = cands[i];
            return;
        }
        dlclose( h );
    }
}

const char *
orc_blas_name( void )
{
    return blas_name;
}
