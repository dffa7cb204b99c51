/*
 * bcsc.c -- the user matrix permuted into the solver numbering.
 *
 * Equivalent of bcscInit (reference bcsc/bcsc.c:1887, layout bcsc/bcsc.h:
 * 119-133) for one process: column j of the result is column peritab[j] of A
 * with rows renumbered by permtab and sorted.  Symmetric and hermitian
 * matrices are stored full (both triangles), as the reference does for its
 * SpMV.  For general matrices the transpose is kept as a second CSC, the
 * reference's Uvalues, so that the device SpMV gathers rows without atomics.
 */
#include <complex.h>
#include <math.h>
#include "px_internal.h"

void
px_bcsc_exit( px_bcsc_t *b )
{
    free( b->colptr );
    free( b->rowtab );
    free( b->Lvalues );
    free( b->trowptr );
    free( b->tcoltab );
    free( b->Uvalues );
    memset( b, 0, sizeof( *b ) );
}

typedef struct { pastix_int_t r; double re, im; } entry_t;

static int
cmp_entry( const void *a, const void *b )
{
    pastix_int_t x = ( (const entry_t *)a )->r, y = ( (const entry_t *)b )->r;
    return ( x > y ) - ( x < y );
}

/* Build a CSC from (row, col, value) triplets, summing duplicates. */
static void
build_csc( pastix_int_t n, pastix_int_t ne, const pastix_int_t *ri, const pastix_int_t *ci,
           const double complex *val, int cplx, pastix_int_t **colptr_o, pastix_int_t **rowtab_o,
           void **values_o, pastix_int_t *nnz_o )
{
    pastix_int_t *cnt = px_calloc( n + 1, sizeof( pastix_int_t ) );
    pastix_int_t *ptr, *fill, *rows, nnz = 0;
    entry_t      *tmp;
    void         *vals;

    for ( pastix_int_t e = 0; e < ne; e++ ) {
        cnt[ci[e] + 1]++;
    }
    for ( pastix_int_t j = 0; j < n; j++ ) {
        cnt[j + 1] += cnt[j];
    }
    tmp  = px_malloc( ( ne > 0 ? ne : 1 ) * sizeof( entry_t ) );
    fill = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    memcpy( fill, cnt, ( n + 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t e = 0; e < ne; e++ ) {
        entry_t *t = tmp + fill[ci[e]]++;
        t->r  = ri[e];
        t->re = creal( val[e] );
        t->im = cimag( val[e] );
    }
    ptr  = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    rows = px_malloc( ( ne > 0 ? ne : 1 ) * sizeof( pastix_int_t ) );
    vals = px_malloc( ( ne > 0 ? ne : 1 ) * ( cplx ? 16 : 8 ) );
    ptr[0] = 0;
    for ( pastix_int_t j = 0; j < n; j++ ) {
        pastix_int_t b = cnt[j], e = cnt[j + 1];
        qsort( tmp + b, e - b, sizeof( entry_t ), cmp_entry );
        for ( pastix_int_t k = b; k < e; k++ ) {
            if ( nnz > ptr[j] && rows[nnz - 1] == tmp[k].r ) {
                if ( cplx ) {
                    ( (double complex *)vals )[nnz - 1] += tmp[k].re + I * tmp[k].im;
                }
                else {
                    ( (double *)vals )[nnz - 1] += tmp[k].re;
                }
                continue;
            }
            rows[nnz] = tmp[k].r;
            if ( cplx ) {
                ( (double complex *)vals )[nnz] = tmp[k].re + I * tmp[k].im;
            }
            else {
                ( (double *)vals )[nnz] = tmp[k].re;
            }
            nnz++;
        }
        ptr[j + 1] = nnz;
    }
    free( cnt );
    free( fill );
    free( tmp );
    *colptr_o = ptr;
    *rowtab_o = rows;
    *values_o = vals;
    *nnz_o    = nnz;
}

int
px_bcsc_init( px_bcsc_t *b, const spmatrix_t *spm, const pastix_order_t *ord )
{
    pastix_int_t          n = spm->n, base = spm->baseval, ne = 0, cap;
    int                   cplx = ( spm->flttype == SpmComplex64 );
    pastix_int_t         *ri, *ci;
    double complex       *vv;
    const double         *dv = (const double *)spm->values;
    const double complex *zv = (const double complex *)spm->values;

    if ( spm->flttype != SpmDouble && spm->flttype != SpmComplex64 ) {
        px_print_error( "bcsc: only double and complex double input matrices are supported" );
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( b, 0, sizeof( *b ) );
    b->n       = n;
    b->flttype = spm->flttype;
    b->mtxtype = spm->mtxtype;

    cap = 2 * spm->nnz;
    ri  = px_malloc( ( cap > 0 ? cap : 1 ) * sizeof( pastix_int_t ) );
    ci  = px_malloc( ( cap > 0 ? cap : 1 ) * sizeof( pastix_int_t ) );
    vv  = px_malloc( ( cap > 0 ? cap : 1 ) * sizeof( double complex ) );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        for ( pastix_int_t k = spm->colptr[j] - base; k < spm->colptr[j + 1] - base; k++ ) {
            pastix_int_t   i = spm->rowptr[k] - base;
            double complex v = cplx ? zv[k] : (double complex)dv[k];
            pastix_int_t   pi = ord->permtab[i], pj = ord->permtab[j];
            if ( spm->fmttype == SpmCSR ) {
                pastix_int_t t = pi;
                pi = pj;
                pj = t;
            }
            ri[ne] = pi;
            ci[ne] = pj;
            vv[ne] = v;
            ne++;
            if ( spm->mtxtype != SpmGeneral && i != j ) {
                ri[ne] = pj;
                ci[ne] = pi;
                vv[ne] = ( spm->mtxtype == SpmHermitian ) ? conj( v ) : v;
                ne++;
            }
        }
    }
    build_csc( n, ne, ri, ci, vv, cplx, &b->colptr, &b->rowtab, &b->Lvalues, &b->nnz );
    /* the transpose, for the row gathers of the SpMV and the U side of LU:
     * a general matrix, and a Hermitian one (A^T = conj(A) != A) */
    if ( spm->mtxtype == SpmGeneral || ( cplx && spm->mtxtype == SpmHermitian ) ) {
        pastix_int_t tnnz;
        build_csc( n, ne, ci, ri, vv, cplx, &b->trowptr, &b->tcoltab, &b->Uvalues, &tnnz );
    }
    free( ri );
    free( ci );
    free( vv );

    /* ||A||_F (scaled sum of squares) and ||A||_inf */
    {
        double  scale = 0., ssq = 1.;
        double *rs = px_calloc( n, sizeof( double ) );
        for ( pastix_int_t k = 0; k < b->nnz; k++ ) {
            double a = cplx ? cabs( ( (double complex *)b->Lvalues )[k] ) : fabs( ( (double *)b->Lvalues )[k] );
            rs[b->rowtab[k]] += a;
            if ( a > 0. ) {
                if ( scale < a ) {
                    ssq   = 1. + ssq * ( scale / a ) * ( scale / a );
                    scale = a;
                }
                else {
                    ssq += ( a / scale ) * ( a / scale );
                }
            }
        }
        b->Anorm_f   = scale * sqrt( ssq );
        b->Anorm_inf = 0.;
        for ( pastix_int_t i = 0; i < n; i++ ) {
            b->Anorm_inf = fmax( b->Anorm_inf, rs[i] );
        }
        free( rs );
    }
    return PASTIX_SUCCESS;
}
