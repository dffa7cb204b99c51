/*
 * oracle.c -- CPU restatement of the PaStiX sopalin path (fill, static-pivot
 * Cholesky, TRSM, 2D updates, static scheduler, supernodal solve, SpMV,
 * GMRES) on a solver matrix given as plain arrays.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library: it is the checker for the device engine
 * and the CPU baseline timed beside it, never part of the product.
 *
 * Each routine follows a reference file:
 *   orc_fill           kernels/cpucblk_zinit.c:392-509 (cpucblk_zfillin_fr)
 *   potf2 / potrf      kernels/core_zpotrfsp.c:67-189 (64-column blocks)
 *   trsm               kernels/core_ztrsmsp.c:148-183 (2D: one dtrsm per blok)
 *   update             kernels/core_zgemmsp.c:401-472 (core_zgemmsp_2d2d)
 *   cblk task          kernels/core_zpotrfsp.c:333-367, dependency counters of
 *                      kernels/cpucblk_zmpi_coeftab.c:541-622
 *   static scheduler   sopalin/sequential_zpotrf.c:102-147 (per-thread task
 *                      lists in a simulated order, spin on ctrbcnt)
 *   solve              kernels/solve_ztrsmsp.c:74-625
 *   spmv               bcsc/bcsc_zspmv.c:57-76
 *   gmres              refinement/z_refine_gmres.c:48-371
 *   LDL^T              kernels/core_zsytrfsp.c:40-204 (sytf2 / 64-blocked
 *                      sytrf), :445-537 (TRSM with D L^T, scalo workspace,
 *                      gemmsp updates); kernels/core_zscalo.c:72-138
 *   LU                 kernels/core_zgetrfsp.c:64-184 (getf2 / 64-blocked getrf
 *                      with static pivot), :216-272 (geadd, getrf, getmo),
 *                      :354-397 (two TRSMs, updates of both sides)
 *
 * Structure arrays: cblk[6*(cblknbr+1)] = {fcolnum, lcolnum, fblok, stride,
 * coefoff, brownum}, blok[5*bloknbr] = {frownum, lrownum, lcblknm, fcblknm,
 * coefind}, browtab[bloknbr-cblknbr].
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle_blas.h"

#define CB( k, f ) ( cblk[6 * ( k ) + ( f )] )
#define BK( b, f ) ( blok[5 * ( b ) + ( f )] )
enum { C_FCOL, C_LCOL, C_FBLOK, C_STRIDE, C_COEFOFF, C_BROWNUM };
enum { B_FROW, B_LROW, B_LCBLK, B_FCBLK, B_COEFIND };

static double
wtime( void )
{
    struct timespec ts;
    clock_gettime( CLOCK_MONOTONIC, &ts );
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

const char *
orc_blas( void )
{
    orc_blas_init();
    return orc_blas_name();
}

/* ------------------------------------------------------------------------ */
/* Typed kernels, instantiated for double and float.                        */

#define ORC_TYPED( T, SFX, GEMM, TRSM, SQRT, FABS )                                                 \
    static void fill_##SFX( int64_t n, const int64_t *cblk, const int64_t *blok,                    \
                            const int64_t *colptr, const int64_t *rowtab, const double *val,        \
                            const int64_t *col2cblk, T *L )                                         \
    {                                                                                               \
        for ( int64_t j = 0; j < n; j++ ) {                                                         \
            int64_t k = col2cblk[j], ncol = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                  \
            int64_t b = CB( k, C_FBLOK ), lb = CB( k + 1, C_FBLOK );                                \
            for ( int64_t e = colptr[j]; e < colptr[j + 1]; e++ ) {                                 \
                int64_t i = rowtab[e];                                                              \
                if ( i < j ) continue;                                                              \
                while ( b < lb && ( BK( b, B_LROW ) < i || BK( b, B_FROW ) > i ) ) b++;             \
                if ( b >= lb ) continue;                                                            \
                int64_t rows = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                               \
                L[CB( k, C_COEFOFF ) + BK( b, B_COEFIND ) + ( i - BK( b, B_FROW ) ) +               \
                  ( j - CB( k, C_FCOL ) ) * rows] = (T)val[e];                                      \
            }                                                                                       \
            (void)ncol;                                                                             \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* unblocked Cholesky with static pivoting (core_zpotf2sp) */                                   \
    static void potf2_##SFX( int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )              \
    {                                                                                               \
        for ( int64_t k = 0; k < n; k++ ) {                                                         \
            T *akk = A + k * ( lda + 1 );                                                           \
            if ( FABS( *akk ) < ( __typeof__( FABS( *akk ) ) )tau ) {                               \
                *akk = (T)tau;                                                                      \
                ( *npiv )++;                                                                        \
            }                                                                                       \
            *akk = SQRT( *akk );                                                                    \
            T a  = (T)1 / *akk;                                                                     \
            for ( int64_t i = k + 1; i < n; i++ ) A[i + k * lda] *= a;                              \
            for ( int64_t j = k + 1; j < n; j++ ) {                                                 \
                T f = A[j + k * lda];                                                               \
                for ( int64_t i = j; i < n; i++ ) A[i + j * lda] -= A[i + k * lda] * f;             \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* blocked (64) Cholesky of the diagonal tile (core_zpotrfsp) */                                \
    static void potrf_##SFX( int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )              \
    {                                                                                               \
        for ( int64_t k = 0; k < n; k += 64 ) {                                                     \
            int64_t bs = ( n - k < 64 ) ? n - k : 64;                                               \
            T      *a  = A + k * ( lda + 1 );                                                       \
            potf2_##SFX( bs, a, lda, npiv, tau );                                                   \
            if ( k + bs < n ) {                                                                     \
                int64_t m  = n - k - bs;                                                            \
                T      *a1 = a + bs;                                                                \
                T      *a2 = a1 + bs * lda;                                                         \
                TRSM( m, bs, a, lda, a1, lda );                                                     \
                /* lower part of a2 -= a1 a1^T, by 64-column slabs */                               \
                for ( int64_t j = 0; j < m; j += 64 ) {                                             \
                    int64_t w = ( m - j < 64 ) ? m - j : 64;                                        \
                    GEMM( ORC_TRANS, m - j, w, bs, (T)-1, a1 + j, lda, a1 + j, lda, (T)1,           \
                          a2 + j + j * lda, lda );                                                  \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    static void trsm_cblk_##SFX( const int64_t *cblk, const int64_t *blok, int64_t k, T *L )        \
    {                                                                                               \
        int64_t n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                          \
        T      *A = L + CB( k, C_COEFOFF );                                                         \
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {                   \
            int64_t m = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                      \
            TRSM( m, n, A, n, A + BK( b, B_COEFIND ), m );                                          \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* C[f(i)] -= A_i A_b^T for every blok i >= b (core_zgemmsp_2d2d) */                            \
    static void update_##SFX( const int64_t *cblk, const int64_t *blok, int64_t k, int64_t b,       \
                              T *L, pthread_mutex_t *locks )                                        \
    {                                                                                               \
        int64_t K  = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                         \
        int64_t F  = BK( b, B_FCBLK );                                                              \
        int64_t N  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                         \
        int64_t fb = CB( F, C_FBLOK );                                                              \
        const T *Bm = L + CB( k, C_COEFOFF ) + BK( b, B_COEFIND );                                  \
        for ( int64_t i = b; i < CB( k + 1, C_FBLOK ); i++ ) {                                      \
            while ( !( BK( i, B_FROW ) >= BK( fb, B_FROW ) && BK( i, B_LROW ) <= BK( fb, B_LROW ) ) ) \
                fb++;                                                                               \
            int64_t M   = BK( i, B_LROW ) - BK( i, B_FROW ) + 1;                                    \
            int64_t ldc = BK( fb, B_LROW ) - BK( fb, B_FROW ) + 1;                                  \
            T *C = L + CB( F, C_COEFOFF ) + BK( fb, B_COEFIND ) + ( BK( i, B_FROW ) - BK( fb, B_FROW ) ) + \
                   ( BK( b, B_FROW ) - CB( F, C_FCOL ) ) * ldc;                                     \
            if ( locks ) pthread_mutex_lock( locks + F );                                           \
            GEMM( ORC_TRANS, M, N, K, (T)-1, L + CB( k, C_COEFOFF ) + BK( i, B_COEFIND ), M, Bm, N, \
                  (T)1, C, ldc );                                                                   \
            if ( locks ) pthread_mutex_unlock( locks + F );                                         \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* forward and backward sweeps (solve_cblk_ztrsmsp_forward/backward) */                         \
    static void solve_##SFX( int64_t nc, const int64_t *cblk, const int64_t *blok,                  \
                             const int64_t *browtab, const T *L, int64_t nrhs, T *X, int64_t ldx )  \
    {                                                                                               \
        for ( int64_t k = 0; k < nc; k++ ) {                                                        \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = 0; j < n; j++ ) {                                                 \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = 0; l < j; l++ ) s -= A[j + l * n] * xq[l];                    \
                    xq[j] = s / A[j + j * n];                                                       \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {               \
                int64_t m = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                  \
                GEMM( ORC_NOTRANS, m, nrhs, n, (T)-1, A + BK( b, B_COEFIND ), m, x, ldx, (T)1,      \
                      X + BK( b, B_FROW ), ldx );                                                   \
            }                                                                                       \
        }                                                                                           \
        for ( int64_t k = nc - 1; k >= 0; k-- ) {                                                   \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = n - 1; j >= 0; j-- ) {                                            \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = j + 1; l < n; l++ ) s -= A[l + j * n] * xq[l];                \
                    xq[j] = s / A[j + j * n];                                                       \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t e = CB( k + 1, C_BROWNUM ) - 1; e >= CB( k, C_BROWNUM ); e-- ) {          \
                int64_t  b  = browtab[e];                                                           \
                int64_t  s  = BK( b, B_LCBLK );                                                     \
                int64_t  m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                \
                int64_t  ns = CB( s, C_LCOL ) - CB( s, C_FCOL ) + 1;                                \
                const T *Ab = L + CB( s, C_COEFOFF ) + BK( b, B_COEFIND );                          \
                /* X_s -= A_b^T X_k(rows of b) */                                                   \
                for ( int64_t q = 0; q < nrhs; q++ ) {                                              \
                    const T *xk = X + BK( b, B_FROW ) + q * ldx;                                    \
                    T       *xs = X + CB( s, C_FCOL ) + q * ldx;                                    \
                    for ( int64_t c = 0; c < ns; c++ ) {                                            \
                        T acc = 0;                                                                  \
                        for ( int64_t r = 0; r < m; r++ ) acc += Ab[r + c * m] * xk[r];             \
                        xs[c] -= acc;                                                               \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }

static double dgemm_tr_wrap_dummy;
static void dgemm_w( int tb, int64_t m, int64_t n, int64_t k, double al, const double *A, int64_t lda,
                     const double *B, int64_t ldb, double be, double *C, int64_t ldc )
{
    orc_dgemm( tb, m, n, k, al, A, lda, B, ldb, be, C, ldc );
}
static void sgemm_w( int tb, int64_t m, int64_t n, int64_t k, float al, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float be, float *C, int64_t ldc )
{
    orc_sgemm( tb, m, n, k, al, A, lda, B, ldb, be, C, ldc );
}
static void dtrsm_w( int64_t m, int64_t n, const double *L, int64_t ldl, double *B, int64_t ldb )
{
    orc_dtrsm( m, n, L, ldl, B, ldb );
}
static void strsm_w( int64_t m, int64_t n, const float *L, int64_t ldl, float *B, int64_t ldb )
{
    orc_strsm( m, n, L, ldl, B, ldb );
}

ORC_TYPED( double, d, dgemm_w, dtrsm_w, sqrt, fabs )
ORC_TYPED( float, s, sgemm_w, strsm_w, sqrtf, fabsf )

/* ------------------------------------------------------------------------ */
/* Static scheduler.                                                        */

typedef struct sched_s {
    int             nthreads;
    int64_t         nc;
    const int64_t  *cblk, *blok;
    void           *L;
    int             flt;
    double          tau;
    int64_t       **lists;
    int64_t        *nlist;
    _Atomic int64_t *ctrbcnt;
    pthread_mutex_t *locks;
    _Atomic int64_t  npiv;
} sched_t;

typedef struct { sched_t *s; int rank; } targ_t;

static void
cblk_task( sched_t *s, int64_t k )
{
    const int64_t *cblk = s->cblk, *blok = s->blok;
    int64_t        n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1, piv = 0;
    /* wait for the incoming contributions */
    while ( atomic_load_explicit( s->ctrbcnt + k, memory_order_acquire ) > 0 ) {
        sched_yield();
    }
    if ( s->flt == 3 ) {
        double *L = (double *)s->L;
        potrf_d( n, L + CB( k, C_COEFOFF ), n, &piv, s->tau );
        trsm_cblk_d( cblk, blok, k, L );
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {
            update_d( cblk, blok, k, b, L, s->locks );
            atomic_fetch_sub_explicit( s->ctrbcnt + BK( b, B_FCBLK ), 1, memory_order_release );
        }
    }
    else {
        float *L = (float *)s->L;
        potrf_s( n, L + CB( k, C_COEFOFF ), n, &piv, s->tau );
        trsm_cblk_s( cblk, blok, k, L );
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {
            update_s( cblk, blok, k, b, L, s->locks );
            atomic_fetch_sub_explicit( s->ctrbcnt + BK( b, B_FCBLK ), 1, memory_order_release );
        }
    }
    if ( piv ) {
        atomic_fetch_add( &s->npiv, piv );
    }
}

static void *
thread_static( void *varg )
{
    targ_t  *a = (targ_t *)varg;
    sched_t *s = a->s;
    for ( int64_t i = 0; i < s->nlist[a->rank]; i++ ) {
        cblk_task( s, s->lists[a->rank][i] );
    }
    return NULL;
}

static double
cblk_cost( const int64_t *cblk, const int64_t *blok, int64_t k )
{
    double n = (double)( CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1 );
    double m = (double)( CB( k, C_STRIDE ) ) - n;
    double c = n * n * n / 3. + m * n * n;
    for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {
        double nb = (double)( BK( b, B_LROW ) - BK( b, B_FROW ) + 1 );
        c += 2. * nb * m * n;
        m -= nb;
    }
    return c + 1e4; /* plus a fixed task overhead */
}

/* Greedy list schedule in simulated time: a cblk becomes ready when its last
 * contributor finishes; ready cblks go, in ready-time order, to the thread
 * that frees first.  Every thread's list is in simulated start order, which
 * keeps the runtime spin-waits deadlock free. */
static void
build_static_lists( sched_t *s )
{
    const int64_t *cblk = s->cblk, *blok = s->blok;
    int64_t        nc = s->nc;
    int            P  = s->nthreads;
    double        *ready  = calloc( nc, sizeof( double ) );
    double        *finish = calloc( nc, sizeof( double ) );
    int64_t       *left   = calloc( nc, sizeof( int64_t ) );
    double        *tfree  = calloc( P, sizeof( double ) );
    int64_t       *heap   = malloc( ( nc > 0 ? nc : 1 ) * sizeof( int64_t ) );
    int64_t        hn = 0;

    for ( int64_t b = 0; b < CB( nc, C_FBLOK ); b++ ) {
        if ( BK( b, B_FCBLK ) != BK( b, B_LCBLK ) ) left[BK( b, B_FCBLK )]++;
    }
    s->lists = malloc( P * sizeof( int64_t * ) );
    s->nlist = calloc( P, sizeof( int64_t ) );
    for ( int p = 0; p < P; p++ ) s->lists[p] = malloc( ( nc > 0 ? nc : 1 ) * sizeof( int64_t ) );

#define HLESS( a, b ) ( ready[a] < ready[b] || ( ready[a] == ready[b] && ( a ) < ( b ) ) )
#define HPUSH( x )                                                                                  \
    do {                                                                                            \
        int64_t i_ = hn++;                                                                          \
        heap[i_]   = ( x );                                                                         \
        while ( i_ > 0 && HLESS( heap[i_], heap[( i_ - 1 ) / 2] ) ) {                               \
            int64_t t_ = heap[i_]; heap[i_] = heap[( i_ - 1 ) / 2]; heap[( i_ - 1 ) / 2] = t_;      \
            i_ = ( i_ - 1 ) / 2;                                                                    \
        }                                                                                           \
    } while ( 0 )

    for ( int64_t k = 0; k < nc; k++ ) {
        if ( left[k] == 0 ) HPUSH( k );
    }
    while ( hn > 0 ) {
        int64_t k = heap[0], i = 0;
        heap[0] = heap[--hn];
        for ( ;; ) {
            int64_t l = 2 * i + 1, r = l + 1, m = i;
            if ( l < hn && HLESS( heap[l], heap[m] ) ) m = l;
            if ( r < hn && HLESS( heap[r], heap[m] ) ) m = r;
            if ( m == i ) break;
            int64_t t = heap[i]; heap[i] = heap[m]; heap[m] = t;
            i = m;
        }
        int p = 0;
        for ( int q = 1; q < P; q++ ) if ( tfree[q] < tfree[p] ) p = q;
        double st = ( ready[k] > tfree[p] ) ? ready[k] : tfree[p];
        finish[k] = st + cblk_cost( cblk, blok, k );
        tfree[p]  = finish[k];
        s->lists[p][s->nlist[p]++] = k;
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {
            int64_t F = BK( b, B_FCBLK );
            if ( finish[k] > ready[F] ) ready[F] = finish[k];
            if ( --left[F] == 0 ) HPUSH( F );
        }
    }
#undef HPUSH
#undef HLESS
    free( ready ); free( finish ); free( left ); free( tfree ); free( heap );
}

/*
 * Fill + numerical Cholesky factorization.  flt: 3 = double, 2 = float.
 * L (coefnbr entries of the factor type) is overwritten.  Returns 0.
 */
int
orc_factorize( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
               int64_t coefnbr, const int64_t *colptr, const int64_t *rowtab, const double *val,
               double tau, void *L, int nthreads, int64_t *npiv, double *seconds )
{
    sched_t  s;
    int64_t *col2cblk = malloc( ( n > 0 ? n : 1 ) * sizeof( int64_t ) );
    double   t0;

    orc_blas_init();
    for ( int64_t k = 0; k < nc; k++ ) {
        for ( int64_t j = CB( k, C_FCOL ); j <= CB( k, C_LCOL ); j++ ) col2cblk[j] = k;
    }
    memset( &s, 0, sizeof( s ) );
    s.nthreads = nthreads > 0 ? nthreads : 1;
    s.nc = nc;
    s.cblk = cblk;
    s.blok = blok;
    s.L = L;
    s.flt = flt;
    s.tau = tau;
    s.ctrbcnt = malloc( ( nc > 0 ? nc : 1 ) * sizeof( _Atomic int64_t ) );
    s.locks   = malloc( ( nc > 0 ? nc : 1 ) * sizeof( pthread_mutex_t ) );
    for ( int64_t k = 0; k < nc; k++ ) {
        atomic_init( s.ctrbcnt + k, 0 );
        pthread_mutex_init( s.locks + k, NULL );
    }
    for ( int64_t b = 0; b < CB( nc, C_FBLOK ); b++ ) {
        if ( BK( b, B_FCBLK ) != BK( b, B_LCBLK ) ) atomic_fetch_add( s.ctrbcnt + BK( b, B_FCBLK ), 1 );
    }
    atomic_init( &s.npiv, 0 );
    build_static_lists( &s );

    t0 = wtime();
    if ( flt == 3 ) {
        memset( L, 0, coefnbr * sizeof( double ) );
        fill_d( n, cblk, blok, colptr, rowtab, val, col2cblk, (double *)L );
    }
    else {
        memset( L, 0, coefnbr * sizeof( float ) );
        fill_s( n, cblk, blok, colptr, rowtab, val, col2cblk, (float *)L );
    }
    {
        pthread_t *th = malloc( s.nthreads * sizeof( pthread_t ) );
        targ_t    *ta = malloc( s.nthreads * sizeof( targ_t ) );
        for ( int p = 0; p < s.nthreads; p++ ) {
            ta[p].s    = &s;
            ta[p].rank = p;
            if ( p > 0 ) pthread_create( th + p, NULL, thread_static, ta + p );
        }
        thread_static( ta );
        for ( int p = 1; p < s.nthreads; p++ ) pthread_join( th[p], NULL );
        free( th );
        free( ta );
    }
    if ( seconds ) *seconds = wtime() - t0;
    if ( npiv ) *npiv = atomic_load( &s.npiv );

    for ( int p = 0; p < s.nthreads; p++ ) free( s.lists[p] );
    free( s.lists );
    free( s.nlist );
    for ( int64_t k = 0; k < nc; k++ ) pthread_mutex_destroy( s.locks + k );
    free( s.locks );
    free( (void *)s.ctrbcnt );
    free( col2cblk );
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Complex arithmetic for the symmetric (non Hermitian) LDL^T and the LU:    */
/* plain loops, C = alpha A op(B) + beta C with op = B or B^T (no conjugate) */

#define ORC_GEMM_LOOP( T, NAME )                                                                     \
    static void NAME( int tb, int64_t m, int64_t n, int64_t k, T al, const T *A, int64_t lda,       \
                      const T *B, int64_t ldb, T be, T *C, int64_t ldc )                             \
    {                                                                                               \
        for ( int64_t j = 0; j < n; j++ ) {                                                         \
            for ( int64_t i = 0; i < m; i++ ) {                                                     \
                T acc = 0;                                                                          \
                for ( int64_t l = 0; l < k; l++ ) {                                                 \
                    T b = ( tb == ORC_TRANS ) ? B[j + l * ldb] : B[l + j * ldb];                    \
                    acc += A[i + l * lda] * b;                                                      \
                }                                                                                   \
                C[i + j * ldc] = al * acc + be * C[i + j * ldc];                                    \
            }                                                                                       \
        }                                                                                           \
    }
ORC_GEMM_LOOP( double complex, zgemm_w )
ORC_GEMM_LOOP( float complex, cgemm_w )

/* lower triangle of the matrix (complex input) into the L tiles */
#define ORC_FILLC( T, SFX )                                                                          \
    static void fill_##SFX( int64_t n, const int64_t *cblk, const int64_t *blok,                    \
                            const int64_t *colptr, const int64_t *rowtab, const double complex *val, \
                            const int64_t *col2cblk, T *L )                                         \
    {                                                                                               \
        for ( int64_t j = 0; j < n; j++ ) {                                                         \
            int64_t k = col2cblk[j];                                                                \
            int64_t b = CB( k, C_FBLOK ), lb = CB( k + 1, C_FBLOK );                                \
            for ( int64_t e = colptr[j]; e < colptr[j + 1]; e++ ) {                                 \
                int64_t i = rowtab[e];                                                              \
                if ( i < j ) continue;                                                              \
                while ( b < lb && ( BK( b, B_LROW ) < i || BK( b, B_FROW ) > i ) ) b++;             \
                if ( b >= lb ) continue;                                                            \
                int64_t rows = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                               \
                L[CB( k, C_COEFOFF ) + BK( b, B_COEFIND ) + ( i - BK( b, B_FROW ) ) +               \
                  ( j - CB( k, C_FCOL ) ) * rows] = (T)val[e];                                      \
            }                                                                                       \
        }                                                                                           \
    }
ORC_FILLC( double complex, z )
ORC_FILLC( float complex, c )

/* ------------------------------------------------------------------------ */
/* LU with static pivoting (kernels/core_zgetrfsp.c), sequential in cblk     */
/* order (every source of cblk k precedes k).  The U side holds U^T in the   */
/* layout of L (ucoeftab).                                                   */

#define ORC_LU_TYPED( T, SFX, GEMM, FABS, VT, NEG )                                                \
    /* U tiles: A[j][i] at (i, j) for i > j (the strict upper part, transposed) */                  \
    static void fillu_##SFX( int64_t n, const int64_t *cblk, const int64_t *blok,                   \
                             const int64_t *tcolptr, const int64_t *trowtab, const VT *tval,        \
                             const int64_t *col2cblk, T *U )                                        \
    {                                                                                               \
        for ( int64_t j = 0; j < n; j++ ) {                                                         \
            int64_t k = col2cblk[j];                                                                \
            int64_t b = CB( k, C_FBLOK ), lb = CB( k + 1, C_FBLOK );                                \
            for ( int64_t e = tcolptr[j]; e < tcolptr[j + 1]; e++ ) {                               \
                int64_t i = trowtab[e];                                                             \
                if ( i <= j ) continue;                                                             \
                while ( b < lb && ( BK( b, B_LROW ) < i || BK( b, B_FROW ) > i ) ) b++;             \
                if ( b >= lb ) continue;                                                            \
                int64_t rows = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                               \
                U[CB( k, C_COEFOFF ) + BK( b, B_COEFIND ) + ( i - BK( b, B_FROW ) ) +               \
                  ( j - CB( k, C_FCOL ) ) * rows] = (T)tval[e];                                     \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* unblocked LU of an m x n panel, static pivot +-tau (core_zgetf2sp) */                        \
    static void getf2_##SFX( int64_t m, int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )   \
    {                                                                                               \
        int64_t mn = ( m < n ) ? m : n;                                                             \
        for ( int64_t k = 0; k < mn; k++ ) {                                                        \
            T *akk = A + k * ( lda + 1 );                                                           \
            if ( FABS( *akk ) < ( __typeof__( FABS( *akk ) ) )tau ) {                               \
                *akk = NEG( *akk ) ? (T)-tau : (T)tau;                                              \
                ( *npiv )++;                                                                        \
            }                                                                                       \
            T a = (T)1 / *akk;                                                                      \
            for ( int64_t i = k + 1; i < m; i++ ) A[i + k * lda] *= a;                              \
            for ( int64_t j = k + 1; j < n; j++ ) {                                                 \
                T f = A[k + j * lda];                                                               \
                for ( int64_t i = k + 1; i < m; i++ ) A[i + j * lda] -= A[i + k * lda] * f;         \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* blocked (64) LU of the diagonal tile (core_zgetrfsp) */                                      \
    static void getrf_##SFX( int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )              \
    {                                                                                               \
        for ( int64_t k = 0; k < n; k += 64 ) {                                                     \
            int64_t bs = ( n - k < 64 ) ? n - k : 64, tm = n - k, ms = tm - bs;                     \
            T      *akk = A + k * ( lda + 1 );                                                      \
            getf2_##SFX( tm, bs, akk, lda, npiv, tau );                                             \
            if ( ms > 0 ) {                                                                         \
                T *ukj = akk + bs * lda, *lik = akk + bs, *aij = ukj + bs;                          \
                /* U_k,k+1 = L_kk^{-1} A_k,k+1 (left, lower, unit) */                               \
                for ( int64_t c = 0; c < ms; c++ )                                                  \
                    for ( int64_t r = 1; r < bs; r++ ) {                                            \
                        T acc = ukj[r + c * lda];                                                   \
                        for ( int64_t l = 0; l < r; l++ ) acc -= akk[r + l * lda] * ukj[l + c * lda]; \
                        ukj[r + c * lda] = acc;                                                     \
                    }                                                                               \
                GEMM( ORC_NOTRANS, ms, ms, bs, (T)-1, lik, lda, ukj, lda, (T)1, aij, lda );         \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* B <- B U^{-1}, U the upper part of the n x n tile D (unit: no division) */                   \
    static void trsm_ru_##SFX( int64_t m, int64_t n, const T *D, int unit, T *B )                   \
    {                                                                                               \
        for ( int64_t j = 0; j < n; j++ ) {                                                         \
            for ( int64_t l = 0; l < j; l++ ) {                                                     \
                T u = D[l + j * n];                                                                 \
                if ( u != (T)0 )                                                                    \
                    for ( int64_t i = 0; i < m; i++ ) B[i + j * m] -= B[i + l * m] * u;             \
            }                                                                                       \
            if ( !unit ) {                                                                          \
                T a = (T)1 / D[j + j * n];                                                          \
                for ( int64_t i = 0; i < m; i++ ) B[i + j * m] *= a;                                \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* C[f(i)] -= A_i B_b^T for the bloks i >= b + shift (core_zgemmsp_2d2d) */                     \
    static void update2_##SFX( const int64_t *cblk, const int64_t *blok, int64_t k, int64_t b,      \
                               int shift, const T *A, const T *B, T *C )                            \
    {                                                                                               \
        int64_t  K  = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                        \
        int64_t  F  = BK( b, B_FCBLK );                                                             \
        int64_t  N  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                        \
        int64_t  fb = CB( F, C_FBLOK );                                                             \
        const T *Bm = B + CB( k, C_COEFOFF ) + BK( b, B_COEFIND );                                  \
        for ( int64_t i = b + shift; i < CB( k + 1, C_FBLOK ); i++ ) {                              \
            while ( !( BK( i, B_FROW ) >= BK( fb, B_FROW ) && BK( i, B_LROW ) <= BK( fb, B_LROW ) ) ) \
                fb++;                                                                               \
            int64_t M   = BK( i, B_LROW ) - BK( i, B_FROW ) + 1;                                    \
            int64_t ldc = BK( fb, B_LROW ) - BK( fb, B_FROW ) + 1;                                  \
            T *Cp = C + CB( F, C_COEFOFF ) + BK( fb, B_COEFIND ) + ( BK( i, B_FROW ) - BK( fb, B_FROW ) ) + \
                    ( BK( b, B_FROW ) - CB( F, C_FCOL ) ) * ldc;                                    \
            GEMM( ORC_TRANS, M, N, K, (T)-1, A + CB( k, C_COEFOFF ) + BK( i, B_COEFIND ), M, Bm, N, \
                  (T)1, Cp, ldc );                                                                  \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    static void lu_cblk_##SFX( const int64_t *cblk, const int64_t *blok, int64_t k, T *L, T *U,     \
                               int64_t *npiv, double tau )                                          \
    {                                                                                               \
        int64_t n  = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                         \
        T      *Ld = L + CB( k, C_COEFOFF ), *Ud = U + CB( k, C_COEFOFF );                          \
        /* geadd: L_kk += U_kk^T, then GETRF, then getmo: U_kk <- L_kk^T */                         \
        for ( int64_t j = 0; j < n; j++ )                                                           \
            for ( int64_t i = 0; i < n; i++ ) Ld[i + j * n] += Ud[j + i * n];                       \
        getrf_##SFX( n, Ld, n, npiv, tau );                                                         \
        for ( int64_t j = 0; j < n; j++ )                                                           \
            for ( int64_t i = 0; i < n; i++ ) Ud[i + j * n] = Ld[j + i * n];                        \
        /* L_off <- L_off U_kk^{-1};  U^T_off <- U^T_off L_kk^{-T} (the upper part */               \
        /* of the U tile is L_kk^T, unit)                                          */               \
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {                   \
            int64_t m = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                      \
            trsm_ru_##SFX( m, n, Ld, 0, L + CB( k, C_COEFOFF ) + BK( b, B_COEFIND ) );              \
            trsm_ru_##SFX( m, n, Ud, 1, U + CB( k, C_COEFOFF ) + BK( b, B_COEFIND ) );              \
        }                                                                                           \
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {                   \
            update2_##SFX( cblk, blok, k, b, 0, L, U, L );                                          \
            if ( b + 1 < CB( k + 1, C_FBLOK ) ) update2_##SFX( cblk, blok, k, b, 1, U, L, U );      \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* L y = b (unit L), then U x = y with U read from the U^T tiles */                             \
    static void lusolve_##SFX( int64_t nc, const int64_t *cblk, const int64_t *blok,                \
                               const int64_t *browtab, const T *L, const T *U, int64_t nrhs, T *X,  \
                               int64_t ldx )                                                        \
    {                                                                                               \
        for ( int64_t k = 0; k < nc; k++ ) {                                                        \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = 0; j < n; j++ ) {                                                 \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = 0; l < j; l++ ) s -= A[j + l * n] * xq[l];                    \
                    xq[j] = s;                                                                      \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {               \
                int64_t m = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                  \
                GEMM( ORC_NOTRANS, m, nrhs, n, (T)-1, A + BK( b, B_COEFIND ), m, x, ldx, (T)1,      \
                      X + BK( b, B_FROW ), ldx );                                                   \
            }                                                                                       \
        }                                                                                           \
        for ( int64_t k = nc - 1; k >= 0; k-- ) {                                                   \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = U + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = n - 1; j >= 0; j-- ) {                                            \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = j + 1; l < n; l++ ) s -= A[l + j * n] * xq[l];                \
                    xq[j] = s / A[j + j * n];                                                       \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t e = CB( k + 1, C_BROWNUM ) - 1; e >= CB( k, C_BROWNUM ); e-- ) {          \
                int64_t  b  = browtab[e];                                                           \
                int64_t  s  = BK( b, B_LCBLK );                                                     \
                int64_t  m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                \
                int64_t  ns = CB( s, C_LCOL ) - CB( s, C_FCOL ) + 1;                                \
                const T *Ab = U + CB( s, C_COEFOFF ) + BK( b, B_COEFIND );                          \
                for ( int64_t q = 0; q < nrhs; q++ ) {                                              \
                    const T *xk = X + BK( b, B_FROW ) + q * ldx;                                    \
                    T       *xs = X + CB( s, C_FCOL ) + q * ldx;                                    \
                    for ( int64_t c = 0; c < ns; c++ ) {                                            \
                        T acc = 0;                                                                  \
                        for ( int64_t r = 0; r < m; r++ ) acc += Ab[r + c * m] * xk[r];             \
                        xs[c] -= acc;                                                               \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }

#define ORC_NEG_R( x ) ( ( x ) < 0 )
#define ORC_NEG_C( x ) ( creal( x ) < 0 )
ORC_LU_TYPED( double, d, dgemm_w, fabs, double, ORC_NEG_R )
ORC_LU_TYPED( float, s, sgemm_w, fabsf, double, ORC_NEG_R )
ORC_LU_TYPED( double complex, z, zgemm_w, cabs, double complex, ORC_NEG_C )
ORC_LU_TYPED( float complex, c, cgemm_w, cabsf, double complex, ORC_NEG_C )

/*
 * Fill + LU factorization with static pivoting.  L and U (coefnbr entries
 * each, factor type flt) are overwritten; tcolptr/trowtab/tval is the CSC of
 * A^T (for a symmetric matrix, the CSC itself).
 */
static size_t
orc_es( int flt )
{
    return ( flt == 3 ) ? 8 : ( flt == 2 ) ? 4 : ( flt == 5 ) ? 16 : 8;
}

int
orc_factorize_lu( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
                  int64_t coefnbr, const int64_t *colptr, const int64_t *rowtab, const void *val,
                  const int64_t *tcolptr, const int64_t *trowtab, const void *tval, double tau,
                  void *L, void *U, int64_t *npiv )
{
    int64_t *col2cblk = malloc( ( n > 0 ? n : 1 ) * sizeof( int64_t ) );
    int64_t  piv = 0;
    orc_blas_init();
    for ( int64_t k = 0; k < nc; k++ ) {
        for ( int64_t j = CB( k, C_FCOL ); j <= CB( k, C_LCOL ); j++ ) col2cblk[j] = k;
    }
    memset( L, 0, coefnbr * orc_es( flt ) );
    memset( U, 0, coefnbr * orc_es( flt ) );
#define ORC_LU_RUN( T, SFX, VT )                                                                     \
    fill_##SFX( n, cblk, blok, colptr, rowtab, (const VT *)val, col2cblk, (T *)L );                 \
    fillu_##SFX( n, cblk, blok, tcolptr, trowtab, (const VT *)tval, col2cblk, (T *)U );             \
    for ( int64_t k = 0; k < nc; k++ ) lu_cblk_##SFX( cblk, blok, k, (T *)L, (T *)U, &piv, tau );
    switch ( flt ) {
    case 3: ORC_LU_RUN( double, d, double ) break;
    case 2: ORC_LU_RUN( float, s, double ) break;
    case 5: ORC_LU_RUN( double complex, z, double complex ) break;
    default: ORC_LU_RUN( float complex, c, double complex ) break;
    }
#undef ORC_LU_RUN
    if ( npiv ) *npiv = piv;
    free( col2cblk );
    return 0;
}

/* x <- A^{-1} x with an LU factor (double x, factor in flt). */
int
orc_solve_lu( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
              const int64_t *browtab, const void *L, const void *U, int64_t nrhs, void *x, int64_t ldx )
{
    orc_blas_init();
    if ( flt == 3 ) {
        lusolve_d( nc, cblk, blok, browtab, (const double *)L, (const double *)U, nrhs, (double *)x, ldx );
    }
    else if ( flt == 5 ) {
        lusolve_z( nc, cblk, blok, browtab, (const double complex *)L, (const double complex *)U, nrhs,
                   (double complex *)x, ldx );
    }
    else if ( flt == 2 ) {
        float  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float ) );
        double *xd = (double *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float)xd[i + q * ldx];
        lusolve_s( nc, cblk, blok, browtab, (const float *)L, (const float *)U, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double)xs[i + q * n];
        free( xs );
    }
    else {
        float complex  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float complex ) );
        double complex *xd = (double complex *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float complex)xd[i + q * ldx];
        lusolve_c( nc, cblk, blok, browtab, (const float complex *)L, (const float complex *)U, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double complex)xs[i + q * n];
        free( xs );
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LDL^T without pivoting, static pivot +-tau (kernels/core_zsytrfsp.c),    */
/* sequential in cblk order.  The diagonal tile holds L (strictly below), D  */
/* and D L^T (above); the off-diagonal tiles hold L.                         */

#define ORC_LDLT_TYPED( T, SFX, GEMM, FABS, NEG )                                                   \
    /* unblocked (core_zsytf2sp): copy the column to the row, scale, update */                      \
    static void sytf2_##SFX( int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )              \
    {                                                                                               \
        for ( int64_t k = 0; k < n; k++ ) {                                                         \
            T *akk = A + k * ( lda + 1 );                                                           \
            if ( FABS( *akk ) < ( __typeof__( FABS( *akk ) ) )tau ) {                               \
                *akk = NEG( *akk ) ? (T)-tau : (T)tau;                                              \
                ( *npiv )++;                                                                        \
            }                                                                                       \
            T a = (T)1 / *akk;                                                                      \
            for ( int64_t i = k + 1; i < n; i++ ) A[k + i * lda] = A[i + k * lda];                  \
            for ( int64_t i = k + 1; i < n; i++ ) A[i + k * lda] *= a;                              \
            for ( int64_t j = k + 1; j < n; j++ )                                                   \
                for ( int64_t i = j; i < n; i++ ) A[i + j * lda] -= *akk * A[i + k * lda] * A[j + k * lda]; \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* blocked (64) LDL^T of the diagonal tile (core_zsytrfsp) */                                   \
    static void sytrf_##SFX( int64_t n, T *A, int64_t lda, int64_t *npiv, double tau )              \
    {                                                                                               \
        for ( int64_t k = 0; k < n; k += 64 ) {                                                     \
            int64_t bs = ( n - k < 64 ) ? n - k : 64, ms = n - k - bs;                              \
            T      *akk = A + k * ( lda + 1 ), *amk = akk + bs, *akm = akk + bs * lda;              \
            sytf2_##SFX( bs, akk, lda, npiv, tau );                                                 \
            if ( ms > 0 ) {                                                                         \
                /* A(k+1:n,k) <- A(k+1:n,k) L(k,k)^{-T} (right, lower, trans, unit) */              \
                for ( int64_t j = 0; j < bs; j++ )                                                  \
                    for ( int64_t l = 0; l < j; l++ ) {                                             \
                        T f = akk[j + l * lda];                                                     \
                        for ( int64_t i = 0; i < ms; i++ ) amk[i + j * lda] -= amk[i + l * lda] * f; \
                    }                                                                               \
                /* D L^T into the upper part, then L = . / D */                                     \
                for ( int64_t j = 0; j < bs; j++ ) {                                                \
                    T a = (T)1 / akk[j + j * lda];                                                  \
                    for ( int64_t i = 0; i < ms; i++ ) {                                            \
                        akm[j + i * lda] = amk[i + j * lda];                                        \
                        amk[i + j * lda] *= a;                                                      \
                    }                                                                               \
                }                                                                                   \
                GEMM( ORC_NOTRANS, ms, ms, bs, (T)-1, amk, lda, akm, lda, (T)1, akm + bs, lda );    \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    static void ldlt_cblk_##SFX( const int64_t *cblk, const int64_t *blok, int64_t k, T *L, T *W,   \
                                 int64_t *npiv, double tau )                                        \
    {                                                                                               \
        int64_t n  = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                         \
        T      *Ld = L + CB( k, C_COEFOFF );                                                        \
        sytrf_##SFX( n, Ld, n, npiv, tau );                                                         \
        /* L_off <- L_off (D L^T)^{-1} (upper part of the tile), W_off = L_off D (scalo) */         \
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {                   \
            int64_t m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                     \
            T      *Lb = L + CB( k, C_COEFOFF ) + BK( b, B_COEFIND );                               \
            T      *Wb = W + CB( k, C_COEFOFF ) + BK( b, B_COEFIND );                               \
            trsm_ru_##SFX( m, n, Ld, 0, Lb );                                                       \
            for ( int64_t j = 0; j < n; j++ )                                                       \
                for ( int64_t i = 0; i < m; i++ ) Wb[i + j * m] = Lb[i + j * m] * Ld[j + j * n];    \
        }                                                                                           \
        for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {                   \
            update2_##SFX( cblk, blok, k, b, 0, L, W, L );                                          \
        }                                                                                           \
    }                                                                                               \
                                                                                                    \
    /* L y = b (unit), y <- D^{-1} y, L^T x = y (unit) */                                           \
    static void ldltsolve_##SFX( int64_t nc, const int64_t *cblk, const int64_t *blok,              \
                                 const int64_t *browtab, const T *L, int64_t nrhs, T *X, int64_t ldx ) \
    {                                                                                               \
        for ( int64_t k = 0; k < nc; k++ ) {                                                        \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = 0; j < n; j++ ) {                                                 \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = 0; l < j; l++ ) s -= A[j + l * n] * xq[l];                    \
                    xq[j] = s;                                                                      \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {               \
                int64_t m = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                  \
                GEMM( ORC_NOTRANS, m, nrhs, n, (T)-1, A + BK( b, B_COEFIND ), m, x, ldx, (T)1,      \
                      X + BK( b, B_FROW ), ldx );                                                   \
            }                                                                                       \
        }                                                                                           \
        for ( int64_t k = 0; k < nc; k++ ) {                                                        \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            for ( int64_t q = 0; q < nrhs; q++ )                                                    \
                for ( int64_t j = 0; j < n; j++ ) X[CB( k, C_FCOL ) + j + q * ldx] /= A[j + j * n]; \
        }                                                                                           \
        for ( int64_t k = nc - 1; k >= 0; k-- ) {                                                   \
            int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                     \
            const T *A = L + CB( k, C_COEFOFF );                                                    \
            T       *x = X + CB( k, C_FCOL );                                                       \
            for ( int64_t q = 0; q < nrhs; q++ ) {                                                  \
                T *xq = x + q * ldx;                                                                \
                for ( int64_t j = n - 1; j >= 0; j-- ) {                                            \
                    T s = xq[j];                                                                    \
                    for ( int64_t l = j + 1; l < n; l++ ) s -= A[l + j * n] * xq[l];                \
                    xq[j] = s;                                                                      \
                }                                                                                   \
            }                                                                                       \
            for ( int64_t e = CB( k + 1, C_BROWNUM ) - 1; e >= CB( k, C_BROWNUM ); e-- ) {          \
                int64_t  b  = browtab[e];                                                           \
                int64_t  s  = BK( b, B_LCBLK );                                                     \
                int64_t  m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                                \
                int64_t  ns = CB( s, C_LCOL ) - CB( s, C_FCOL ) + 1;                                \
                const T *Ab = L + CB( s, C_COEFOFF ) + BK( b, B_COEFIND );                          \
                for ( int64_t q = 0; q < nrhs; q++ ) {                                              \
                    const T *xk = X + BK( b, B_FROW ) + q * ldx;                                    \
                    T       *xs = X + CB( s, C_FCOL ) + q * ldx;                                    \
                    for ( int64_t c = 0; c < ns; c++ ) {                                            \
                        T acc = 0;                                                                  \
                        for ( int64_t r = 0; r < m; r++ ) acc += Ab[r + c * m] * xk[r];             \
                        xs[c] -= acc;                                                               \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }

ORC_LDLT_TYPED( double, d, dgemm_w, fabs, ORC_NEG_R )
ORC_LDLT_TYPED( float, s, sgemm_w, fabsf, ORC_NEG_R )
ORC_LDLT_TYPED( double complex, z, zgemm_w, cabs, ORC_NEG_C )
ORC_LDLT_TYPED( float complex, c, cgemm_w, cabsf, ORC_NEG_C )

/* Fill + LDL^T factorization with static pivoting (flt 3 double, 2 float). */
int
orc_factorize_ldlt( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
                    int64_t coefnbr, const int64_t *colptr, const int64_t *rowtab, const void *val,
                    double tau, void *L, int64_t *npiv )
{
    int64_t *col2cblk = malloc( ( n > 0 ? n : 1 ) * sizeof( int64_t ) );
    int64_t  piv = 0;
    void    *W = calloc( coefnbr > 0 ? coefnbr : 1, orc_es( flt ) );
    orc_blas_init();
    for ( int64_t k = 0; k < nc; k++ ) {
        for ( int64_t j = CB( k, C_FCOL ); j <= CB( k, C_LCOL ); j++ ) col2cblk[j] = k;
    }
    memset( L, 0, coefnbr * orc_es( flt ) );
#define ORC_LDLT_RUN( T, SFX, VT )                                                                   \
    fill_##SFX( n, cblk, blok, colptr, rowtab, (const VT *)val, col2cblk, (T *)L );                 \
    for ( int64_t k = 0; k < nc; k++ ) ldlt_cblk_##SFX( cblk, blok, k, (T *)L, (T *)W, &piv, tau );
    switch ( flt ) {
    case 3: ORC_LDLT_RUN( double, d, double ) break;
    case 2: ORC_LDLT_RUN( float, s, double ) break;
    case 5: ORC_LDLT_RUN( double complex, z, double complex ) break;
    default: ORC_LDLT_RUN( float complex, c, double complex ) break;
    }
#undef ORC_LDLT_RUN
    if ( npiv ) *npiv = piv;
    free( W );
    free( col2cblk );
    return 0;
}

/* x <- A^{-1} x with an LDL^T factor (double x, factor in flt). */
int
orc_solve_ldlt( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
                const int64_t *browtab, const void *L, int64_t nrhs, void *x, int64_t ldx )
{
    orc_blas_init();
    if ( flt == 3 ) {
        ldltsolve_d( nc, cblk, blok, browtab, (const double *)L, nrhs, (double *)x, ldx );
    }
    else if ( flt == 5 ) {
        ldltsolve_z( nc, cblk, blok, browtab, (const double complex *)L, nrhs, (double complex *)x, ldx );
    }
    else if ( flt == 2 ) {
        float  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float ) );
        double *xd = (double *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float)xd[i + q * ldx];
        ldltsolve_s( nc, cblk, blok, browtab, (const float *)L, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double)xs[i + q * n];
        free( xs );
    }
    else {
        float complex  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float complex ) );
        double complex *xd = (double complex *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float complex)xd[i + q * ldx];
        ldltsolve_c( nc, cblk, blok, browtab, (const float complex *)L, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double complex)xs[i + q * n];
        free( xs );
    }
    return 0;
}

/* x <- A^{-1} x with the factor (double x, factor in flt). */
int
orc_solve( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
           const int64_t *browtab, const void *L, int64_t nrhs, double *x, int64_t ldx )
{
    orc_blas_init();
    if ( flt == 3 ) {
        solve_d( nc, cblk, blok, browtab, (const double *)L, nrhs, x, ldx );
    }
    else {
        float *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float ) );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float)x[i + q * ldx];
        solve_s( nc, cblk, blok, browtab, (const float *)L, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) x[i + q * ldx] = (double)xs[i + q * n];
        free( xs );
    }
    return 0;
}

/* y = alpha A x + beta y on the full-pattern CSC of a symmetric matrix
 * (column j is row j), or on the transposed CSC of a general one. */
int
orc_spmv( int64_t n, const int64_t *colptr, const int64_t *rowtab, const double *val, double alpha,
          const double *x, double beta, double *y )
{
    for ( int64_t j = 0; j < n; j++ ) {
        double s = 0.;
        for ( int64_t e = colptr[j]; e < colptr[j + 1]; e++ ) s += val[e] * x[rowtab[e]];
        y[j] = alpha * s + ( beta == 0. ? 0. : beta * y[j] );
    }
    return 0;
}

/* the same for complex double values (complex symmetric or Hermitian full
 * pattern, or the transposed CSC of a general matrix) */
int
orc_spmv_z( int64_t n, const int64_t *colptr, const int64_t *rowtab, const double complex *val, double alpha,
            const double complex *x, double beta, double complex *y )
{
    for ( int64_t j = 0; j < n; j++ ) {
        double complex s = 0.;
        for ( int64_t e = colptr[j]; e < colptr[j + 1]; e++ ) s += val[e] * x[rowtab[e]];
        y[j] = alpha * s + ( beta == 0. ? 0. : beta * y[j] );
    }
    return 0;
}

static double
nrm2( int64_t n, const double *x )
{
    double s = 0.;
    for ( int64_t i = 0; i < n; i++ ) s += x[i] * x[i];
    return sqrt( s );
}

/*
 * GMRES(im) right-preconditioned by the factor, one right-hand side, as
 * z_gmres_smp: modified Gram-Schmidt, Givens rotations, stop on
 * ||r||_2/||b||_2 <= eps.  Returns the number of iterations.
 */
int64_t
orc_gmres( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
           const int64_t *browtab, const void *L, const int64_t *colptr, const int64_t *rowtab,
           const double *val, const double *b, double *x, int64_t im, int64_t itermax, double eps )
{
    int64_t im1 = im + 1, iters = 0, i;
    double *V = malloc( n * im1 * sizeof( double ) ), *W = malloc( n * im * sizeof( double ) );
    double *H = calloc( im * im1, sizeof( double ) ), *cs = calloc( im, sizeof( double ) );
    double *sn = calloc( im, sizeof( double ) ), *g = calloc( im1, sizeof( double ) );
    double  normb = nrm2( n, b ), resid;
    if ( normb == 0. ) normb = 1.;
    for ( ;; ) {
        memcpy( V, b, n * sizeof( double ) );
        orc_spmv( n, colptr, rowtab, val, -1., x, 1., V );
        resid = nrm2( n, V );
        if ( resid / normb <= eps ) break;
        for ( int64_t k = 0; k < n; k++ ) V[k] /= resid;
        memset( g, 0, im1 * sizeof( double ) );
        g[0] = resid;
        for ( i = 0; i < im && iters < itermax; ) {
            double *vi = V + i * n, *vi1 = V + ( i + 1 ) * n, *wi = W + i * n, *Hi = H + i * im1, t;
            memcpy( wi, vi, n * sizeof( double ) );
            orc_solve( flt, n, nc, cblk, blok, browtab, L, 1, wi, n );
            orc_spmv( n, colptr, rowtab, val, 1., wi, 0., vi1 );
            for ( int64_t j = 0; j <= i; j++ ) {
                double h = 0.;
                for ( int64_t k = 0; k < n; k++ ) h += vi1[k] * V[j * n + k];
                Hi[j] = h;
                for ( int64_t k = 0; k < n; k++ ) vi1[k] -= h * V[j * n + k];
            }
            Hi[i + 1] = nrm2( n, vi1 );
            if ( Hi[i + 1] > 1e-50 ) {
                for ( int64_t k = 0; k < n; k++ ) vi1[k] /= Hi[i + 1];
            }
            for ( int64_t j = 0; j < i; j++ ) {
                t         = Hi[j];
                Hi[j]     = cs[j] * t + sn[j] * Hi[j + 1];
                Hi[j + 1] = cs[j] * Hi[j + 1] - sn[j] * t;
            }
            t = sqrt( Hi[i] * Hi[i] + Hi[i + 1] * Hi[i + 1] );
            if ( fabs( t ) <= eps ) t = eps;
            cs[i]    = Hi[i] / t;
            sn[i]    = Hi[i + 1] / t;
            g[i + 1] = -sn[i] * g[i];
            g[i]     = cs[i] * g[i];
            Hi[i]    = cs[i] * Hi[i] + sn[i] * Hi[i + 1];
            i++;
            iters++;
            if ( fabs( g[i] ) / normb <= eps ) break;
        }
        for ( int64_t k = i - 1; k >= 0; k-- ) {
            double s = g[k];
            for ( int64_t l = k + 1; l < i; l++ ) s -= H[l * im1 + k] * g[l];
            g[k] = s / H[k * im1 + k];
        }
        for ( int64_t j = 0; j < i; j++ )
            for ( int64_t k = 0; k < n; k++ ) x[k] += W[j * n + k] * g[j];
        if ( iters >= itermax ) break;
    }
    free( V ); free( W ); free( H ); free( cs ); free( sn ); free( g );
    (void)dgemm_tr_wrap_dummy;
    return iters;
}

/* ------------------------------------------------------------------------ */
/* One step of a solve on one side of a factor (the sweeps of               */
/* solve_cblk_ztrsmsp_forward / _backward, kernels/solve_ztrsmsp.c:300-625, */
/* and solve_cblk_zdiag :629-679), as pastix_subtask_trsm and                */
/* pastix_subtask_diag apply them.  A is a coeftab whose diagonal tiles are  */
/* read by their lower triangle (Cholesky / LDL^T L, LU L, LU U^T); unit:    */
/* ones on the diagonal instead of the stored entries.                       */
/*   step 0: forward  A y = b (lower)                                        */
/*   step 1: backward A^T x = y                                              */
/*   step 2: x <- x / diag(A_kk)                                             */

#define ORC_STEP_TYPED( T, SFX )                                                                    \
    static void solvestep_##SFX( int64_t nc, const int64_t *cblk, const int64_t *blok,              \
                                 const int64_t *browtab, const T *L, int step, int unit,            \
                                 int64_t nrhs, T *X, int64_t ldx )                                  \
    {                                                                                               \
        if ( step == 0 ) {                                                                          \
            for ( int64_t k = 0; k < nc; k++ ) {                                                    \
                int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                 \
                const T *A = L + CB( k, C_COEFOFF );                                                \
                for ( int64_t q = 0; q < nrhs; q++ ) {                                              \
                    T *xq = X + CB( k, C_FCOL ) + q * ldx;                                          \
                    for ( int64_t j = 0; j < n; j++ ) {                                             \
                        T s = xq[j];                                                                \
                        for ( int64_t l = 0; l < j; l++ ) s -= A[j + l * n] * xq[l];                \
                        xq[j] = unit ? s : s / A[j + j * n];                                        \
                    }                                                                               \
                    /* X[rows of b] -= A_b x_k for the off-diagonal bloks */                        \
                    for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {       \
                        int64_t  m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                        \
                        const T *Ab = A + BK( b, B_COEFIND );                                       \
                        T       *xr = X + BK( b, B_FROW ) + q * ldx;                                \
                        for ( int64_t c = 0; c < n; c++ )                                           \
                            for ( int64_t r = 0; r < m; r++ ) xr[r] -= Ab[r + c * m] * xq[c];       \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
        else if ( step == 1 ) {                                                                     \
            for ( int64_t k = nc - 1; k >= 0; k-- ) {                                               \
                int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                 \
                const T *A = L + CB( k, C_COEFOFF );                                                \
                for ( int64_t q = 0; q < nrhs; q++ ) {                                              \
                    T *xq = X + CB( k, C_FCOL ) + q * ldx;                                          \
                    /* x_k -= A_b^T X[rows of b] for the cblk's own off-diagonal bloks */           \
                    for ( int64_t b = CB( k, C_FBLOK ) + 1; b < CB( k + 1, C_FBLOK ); b++ ) {       \
                        int64_t  m  = BK( b, B_LROW ) - BK( b, B_FROW ) + 1;                        \
                        const T *Ab = A + BK( b, B_COEFIND );                                       \
                        const T *xr = X + BK( b, B_FROW ) + q * ldx;                                \
                        for ( int64_t c = 0; c < n; c++ ) {                                         \
                            T acc = 0;                                                              \
                            for ( int64_t r = 0; r < m; r++ ) acc += Ab[r + c * m] * xr[r];         \
                            xq[c] -= acc;                                                           \
                        }                                                                           \
                    }                                                                               \
                    for ( int64_t j = n - 1; j >= 0; j-- ) {                                        \
                        T s = xq[j];                                                                \
                        for ( int64_t l = j + 1; l < n; l++ ) s -= A[l + j * n] * xq[l];            \
                        xq[j] = unit ? s : s / A[j + j * n];                                        \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
        else {                                                                                      \
            for ( int64_t k = 0; k < nc; k++ ) {                                                    \
                int64_t  n = CB( k, C_LCOL ) - CB( k, C_FCOL ) + 1;                                 \
                const T *A = L + CB( k, C_COEFOFF );                                                \
                for ( int64_t q = 0; q < nrhs; q++ )                                                \
                    for ( int64_t j = 0; j < n; j++ ) X[CB( k, C_FCOL ) + j + q * ldx] /= A[j + j * n]; \
            }                                                                                       \
        }                                                                                           \
    }

ORC_STEP_TYPED( double, d )
ORC_STEP_TYPED( float, s )
ORC_STEP_TYPED( double complex, z )
ORC_STEP_TYPED( float complex, c )

/* x (double / complex double) <- one solve step with the tiles A of a factor
 * in flt (a single factor works on x rounded to single and back). */
int
orc_solve_step( int flt, int64_t n, int64_t nc, const int64_t *cblk, const int64_t *blok,
                const int64_t *browtab, const void *A, int step, int unit, int64_t nrhs, void *x, int64_t ldx )
{
    if ( flt == 3 ) {
        solvestep_d( nc, cblk, blok, browtab, (const double *)A, step, unit, nrhs, (double *)x, ldx );
    }
    else if ( flt == 5 ) {
        solvestep_z( nc, cblk, blok, browtab, (const double complex *)A, step, unit, nrhs, (double complex *)x, ldx );
    }
    else if ( flt == 2 ) {
        float  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float ) );
        double *xd = (double *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float)xd[i + q * ldx];
        solvestep_s( nc, cblk, blok, browtab, (const float *)A, step, unit, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double)xs[i + q * n];
        free( xs );
    }
    else {
        float complex  *xs = malloc( ( n * nrhs > 0 ? n * nrhs : 1 ) * sizeof( float complex ) );
        double complex *xd = (double complex *)x;
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xs[i + q * n] = (float complex)xd[i + q * ldx];
        solvestep_c( nc, cblk, blok, browtab, (const float complex *)A, step, unit, nrhs, xs, n );
        for ( int64_t q = 0; q < nrhs; q++ )
            for ( int64_t i = 0; i < n; i++ ) xd[i + q * ldx] = (double complex)xs[i + q * n];
        free( xs );
    }
    return 0;
}
