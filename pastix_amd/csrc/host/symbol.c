/*
 * symbol.c -- block symbolic factorization, intra-supernode row reordering,
 * supernode splitting and flop accounting.
 *
 * The block structure produced here must be identical to the reference's
 * (integer work), so each step restates the reference's rule:
 *  - px_symbol_fax_direct: the block elimination of symbol/symbol_fax.c
 *    (included by symbol/symbol_fax_direct.c:66).  The row set of a cblk is
 *    its columns' pattern below the cblk plus, for every child cblk (whose
 *    first off-diagonal blok faces it), the child's rows below the cblk; bloks
 *    are the maximal runs of consecutive rows inside one facing cblk.
 *  - px_symbol_reordering: the pseudo-TSP of symbol/symbol_reordering.c:
 *    124-304 (hamming distances, cheapest insertion with the cut tie-break)
 *    and the split-level search of :383-418.
 *  - px_symbol_split: splitSmart (blend/splitsymbol.c:78-432) and the symbol
 *    rebuild of extraCblkMerge (blend/extracblk.c).
 *  - px_symbol_flops: pastixSymbolGetFlops (symbol/symbol_cost.c:52-88,
 *    :287-297, :421-443) with the formulas of common/flops.h, summed with the
 *    same pairwise recursion so that THFLOPS matches to the last bit.
 */
#include <math.h>
#include <pthread.h>
#include "px_internal.h"

/* ------------------------------------------------------------------------ */
/* Graph preparation: symmetrized pattern without the diagonal. */

static int
cmp_int( const void *a, const void *b )
{
    pastix_int_t x = *(const pastix_int_t *)a, y = *(const pastix_int_t *)b;
    return ( x > y ) - ( x < y );
}

int
px_graph_prepare( const spmatrix_t *spm, px_graph_t *g )
{
    pastix_int_t  n = spm->n, base = spm->baseval;
    pastix_int_t *cnt = px_calloc( n + 1, sizeof( pastix_int_t ) );
    pastix_int_t *tmpc, *tmpr, *fill, pos;

    /* count both (i,j) and (j,i) for every off-diagonal entry */
    for ( pastix_int_t j = 0; j < n; j++ ) {
        for ( pastix_int_t k = spm->colptr[j] - base; k < spm->colptr[j + 1] - base; k++ ) {
            pastix_int_t i = spm->rowptr[k] - base;
            if ( i != j ) {
                cnt[i]++;
                cnt[j]++;
            }
        }
    }
    tmpc = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    tmpc[0] = 0;
    for ( pastix_int_t j = 0; j < n; j++ ) {
        tmpc[j + 1] = tmpc[j] + cnt[j];
    }
    tmpr = px_malloc( ( tmpc[n] > 0 ? tmpc[n] : 1 ) * sizeof( pastix_int_t ) );
    fill = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    memcpy( fill, tmpc, ( n + 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        for ( pastix_int_t k = spm->colptr[j] - base; k < spm->colptr[j + 1] - base; k++ ) {
            pastix_int_t i = spm->rowptr[k] - base;
            if ( i != j ) {
                tmpr[fill[j]++] = i;
                tmpr[fill[i]++] = j;
            }
        }
    }
    /* sort and remove duplicates in place */
    g->n      = n;
    g->colptr = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pos = 0;
    g->colptr[0] = 0;
    for ( pastix_int_t j = 0; j < n; j++ ) {
        pastix_int_t b = tmpc[j], e = tmpc[j + 1];
        qsort( tmpr + b, e - b, sizeof( pastix_int_t ), cmp_int );
        for ( pastix_int_t k = b; k < e; k++ ) {
            if ( k == b || tmpr[k] != tmpr[k - 1] ) {
                tmpr[pos++] = tmpr[k];
            }
        }
        g->colptr[j + 1] = pos;
    }
    g->rows = px_malloc( ( pos > 0 ? pos : 1 ) * sizeof( pastix_int_t ) );
    memcpy( g->rows, tmpr, pos * sizeof( pastix_int_t ) );
    free( tmpc );
    free( tmpr );
    free( fill );
    free( cnt );
    return PASTIX_SUCCESS;
}

void
px_graph_exit( px_graph_t *g )
{
    free( g->colptr );
    free( g->rows );
    g->colptr = g->rows = NULL;
}

/* ------------------------------------------------------------------------ */

void
px_symbol_exit( px_symbol_t *s )
{
    free( s->cblktab );
    free( s->bloktab );
    free( s->browtab );
    memset( s, 0, sizeof( *s ) );
}

void
px_symbol_build_rowtab( px_symbol_t *s )
{
    pastix_int_t  nc = s->cblknbr;
    pastix_int_t  ne = s->bloknbr - nc;
    pastix_int_t *cnt = px_calloc( nc + 1, sizeof( pastix_int_t ) );

    for ( pastix_int_t c = 0; c < nc; c++ ) {
        for ( pastix_int_t b = s->cblktab[c].bloknum + 1; b < s->cblktab[c + 1].bloknum; b++ ) {
            cnt[s->bloktab[b].fcblknm]++;
        }
    }
    s->browmax = 0;
    s->cblktab[0].brownum = 0;
    for ( pastix_int_t c = 0; c < nc; c++ ) {
        s->browmax = px_imax( s->browmax, cnt[c] );
        s->cblktab[c + 1].brownum = s->cblktab[c].brownum + cnt[c];
        cnt[c] = s->cblktab[c].brownum;
    }
    free( s->browtab );
    s->browtab = px_malloc( ( ne > 0 ? ne : 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t c = 0; c < nc; c++ ) {
        for ( pastix_int_t b = s->cblktab[c].bloknum + 1; b < s->cblktab[c + 1].bloknum; b++ ) {
            s->browtab[cnt[s->bloktab[b].fcblknm]++] = b;
        }
    }
    free( cnt );
}

/*
 * Block symbolic factorization.  Cblks are processed in increasing order;
 * the rows below a cblk are collected with a stamp array, sorted, then cut
 * into bloks.  The rows a child passes to its father are kept as bloks and
 * expanded when the father is processed.
 */
int
px_symbol_fax_direct( px_symbol_t *s, const px_graph_t *g, const pastix_order_t *ord )
{
    pastix_int_t  n = g->n, nc = ord->cblknbr;
    pastix_int_t *col2cblk = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t *stamp    = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t *rows     = px_malloc( ( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *father   = px_malloc( nc * sizeof( pastix_int_t ) );
    pastix_int_t *chead    = px_malloc( nc * sizeof( pastix_int_t ) ); /* first child */
    pastix_int_t *cnext    = px_malloc( nc * sizeof( pastix_int_t ) ); /* next sibling */
    pastix_int_t  blokmax  = 2 * nc + 16, bloknbr = 0;
    px_symblok_t *bt       = px_malloc( blokmax * sizeof( px_symblok_t ) );

    memset( s, 0, sizeof( *s ) );
    s->cblknbr   = nc;
    s->nodenbr   = n;
    s->schurfcol = n;
    s->dof       = 1;
    s->cblktab   = px_calloc( nc + 1, sizeof( px_symcblk_t ) );

    for ( pastix_int_t c = 0; c < nc; c++ ) {
        for ( pastix_int_t j = ord->rangtab[c]; j < ord->rangtab[c + 1]; j++ ) {
            col2cblk[j] = c;
        }
        chead[c] = -1;
        cnext[c] = -1;
    }
    for ( pastix_int_t i = 0; i < n; i++ ) {
        stamp[i] = -1;
    }

    for ( pastix_int_t c = 0; c < nc; c++ ) {
        pastix_int_t fcol = ord->rangtab[c], lcol = ord->rangtab[c + 1] - 1;
        pastix_int_t nr = 0;

        /* the bloks of cblk c start here: this also closes the blok range of
         * cblk c - 1, which may be one of the children read below */
        s->cblktab[c].bloknum = bloknbr;

        /* pattern of A below the cblk */
        for ( pastix_int_t j = fcol; j <= lcol; j++ ) {
            pastix_int_t v = ord->peritab[j];
            for ( pastix_int_t k = g->colptr[v]; k < g->colptr[v + 1]; k++ ) {
                pastix_int_t r = ord->permtab[g->rows[k]];
                if ( r > lcol && stamp[r] != c ) {
                    stamp[r]   = c;
                    rows[nr++] = r;
                }
            }
        }
        /* contributions of the children */
        for ( pastix_int_t ch = chead[c]; ch != -1; ch = cnext[ch] ) {
            for ( pastix_int_t b = s->cblktab[ch].bloknum + 1; b < s->cblktab[ch + 1].bloknum; b++ ) {
                for ( pastix_int_t r = px_imax( bt[b].frownum, lcol + 1 ); r <= bt[b].lrownum; r++ ) {
                    if ( stamp[r] != c ) {
                        stamp[r]   = c;
                        rows[nr++] = r;
                    }
                }
            }
        }
        qsort( rows, nr, sizeof( pastix_int_t ), cmp_int );

        /* bloks: diagonal, then maximal runs inside one facing cblk */
        if ( bloknbr + nr + 2 > blokmax ) {
            while ( bloknbr + nr + 2 > blokmax ) {
                blokmax = blokmax + blokmax / 4 + 16;
            }
            bt = realloc( bt, blokmax * sizeof( px_symblok_t ) );
            if ( bt == NULL ) {
                px_print_error( "symbolic factorization: out of memory" );
                abort();
            }
        }
        s->cblktab[c].fcolnum = fcol;
        s->cblktab[c].lcolnum = lcol;
        s->cblktab[c].bloknum = bloknbr;
        bt[bloknbr].frownum = fcol;
        bt[bloknbr].lrownum = lcol;
        bt[bloknbr].lcblknm = c;
        bt[bloknbr].fcblknm = c;
        bloknbr++;
        for ( pastix_int_t k = 0; k < nr; ) {
            pastix_int_t f  = col2cblk[rows[k]];
            pastix_int_t fr = rows[k];
            pastix_int_t lim = ord->rangtab[f + 1];
            while ( k + 1 < nr && rows[k + 1] == rows[k] + 1 && rows[k + 1] < lim ) {
                k++;
            }
            bt[bloknbr].frownum = fr;
            bt[bloknbr].lrownum = rows[k];
            bt[bloknbr].lcblknm = c;
            bt[bloknbr].fcblknm = f;
            bloknbr++;
            k++;
        }
        /* the father is the cblk facing the first off-diagonal blok */
        if ( nr > 0 ) {
            pastix_int_t f = bt[s->cblktab[c].bloknum + 1].fcblknm;
            father[c] = f;
            cnext[c]  = chead[f];
            chead[f]  = c;
        }
        else {
            father[c] = -1;
        }
    }
    s->cblktab[nc].fcolnum = n;
    s->cblktab[nc].lcolnum = n;
    s->cblktab[nc].bloknum = bloknbr;
    s->bloknbr = bloknbr;
    s->bloktab = realloc( bt, ( bloknbr > 0 ? bloknbr : 1 ) * sizeof( px_symblok_t ) );

    free( col2cblk );
    free( stamp );
    free( rows );
    free( father );
    free( chead );
    free( cnext );
    px_symbol_build_rowtab( s );
    return PASTIX_SUCCESS;
}

int
px_symbol_check( const px_symbol_t *s )
{
    for ( pastix_int_t c = 0; c < s->cblknbr; c++ ) {
        const px_symcblk_t *cb = s->cblktab + c;
        pastix_int_t        b0 = cb->bloknum, b1 = cb[1].bloknum;
        if ( cb->fcolnum > cb->lcolnum || b1 <= b0 ) {
            return PASTIX_ERR_BADPARAMETER;
        }
        if ( s->bloktab[b0].frownum != cb->fcolnum || s->bloktab[b0].lrownum != cb->lcolnum ||
             s->bloktab[b0].fcblknm != c ) {
            return PASTIX_ERR_BADPARAMETER;
        }
        if ( c > 0 && cb->fcolnum != cb[-1].lcolnum + 1 ) {
            return PASTIX_ERR_BADPARAMETER;
        }
        for ( pastix_int_t b = b0 + 1; b < b1; b++ ) {
            const px_symblok_t *bk = s->bloktab + b;
            const px_symcblk_t *fc = s->cblktab + bk->fcblknm;
            if ( bk->lcblknm != c || bk->fcblknm <= c || bk->frownum > bk->lrownum ||
                 bk->frownum <= s->bloktab[b - 1].lrownum ||
                 bk->frownum < fc->fcolnum || bk->lrownum > fc->lcolnum ) {
                return PASTIX_ERR_BADPARAMETER;
            }
        }
    }
    return PASTIX_SUCCESS;
}

pastix_int_t
px_symbol_nnz( const px_symbol_t *s )
{
    pastix_int_t nnz = 0;
    for ( pastix_int_t c = 0; c < s->cblknbr; c++ ) {
        pastix_int_t ncol = s->cblktab[c].lcolnum - s->cblktab[c].fcolnum + 1;
        pastix_int_t nrow = 0;
        for ( pastix_int_t b = s->cblktab[c].bloknum + 1; b < s->cblktab[c + 1].bloknum; b++ ) {
            nrow += s->bloktab[b].lrownum - s->bloktab[b].frownum + 1;
        }
        nnz += ( ncol * ( ncol + 1 ) ) / 2 + ncol * nrow;
    }
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Flop formulas (common/flops.h): FMULS/FADDS pairs, complex = 6 mul + 2 add. */

static double fm_potrf( double n ) { return n * ( ( ( 1. / 6. ) * n + 0.5 ) * n + ( 1. / 3. ) ); }
static double fa_potrf( double n ) { return n * ( ( ( 1. / 6. ) * n ) * n - ( 1. / 6. ) ); }
static double fm_sytrf( double n ) { return n * ( ( ( 1. / 6. ) * n + 0.5 ) * n + ( 10. / 3. ) ); }
static double fa_sytrf( double n ) { return n * ( ( ( 1. / 6. ) * n ) * n - ( 1. / 6. ) ); }
static double fm_getrf( double m, double n )
{
    return ( m < n ) ? ( 0.5 * m * ( m * ( n - ( 1. / 3. ) * m - 1. ) + n ) + ( 2. / 3. ) * m )
                     : ( 0.5 * n * ( n * ( m - ( 1. / 3. ) * n - 1. ) + m ) + ( 2. / 3. ) * n );
}
static double fa_getrf( double m, double n )
{
    return ( m < n ) ? ( 0.5 * m * ( m * ( n - ( 1. / 3. ) * m ) - n ) + ( 1. / 6. ) * m )
                     : ( 0.5 * n * ( n * ( m - ( 1. / 3. ) * n ) - m ) + ( 1. / 6. ) * n );
}
/* TRSM right-side: FMULS_TRMM_2(N, M); the reference counts adds as muls */
static double fm_trsm_r( double m, double n ) { return 0.5 * m * n * ( n + 1. ); }
static double fm_syrk( double k, double n ) { return 0.5 * k * n * ( n + 1. ); }
static double fm_gemm( double m, double n, double k ) { return m * n * k; }

typedef struct flops_fn_s {
    int cplx;
    int facto;
} flops_fn_t;

static double
fl_combine( int cplx, double fm, double fa )
{
    return cplx ? ( 6. * fm + 2. * fa ) : ( fm + fa );
}

static double
fl_diag( const flops_fn_t *f, double N )
{
    switch ( f->facto ) {
    case PastixFactSYTRF:
    case PastixFactHETRF:
        return fl_combine( f->cplx, fm_sytrf( N ), fa_sytrf( N ) );
    case PastixFactGETRF:
        return fl_combine( f->cplx, fm_getrf( N, N ), fa_getrf( N, N ) );
    default:
        return fl_combine( f->cplx, fm_potrf( N ), fa_potrf( N ) );
    }
}

static double
fl_trsm( const flops_fn_t *f, double M, double N )
{
    double t = fl_combine( f->cplx, fm_trsm_r( M, N ), fm_trsm_r( M, N ) );
    switch ( f->facto ) {
    case PastixFactSYTRF:
    case PastixFactHETRF:
        return t + ( f->cplx ? 6. : 1. ) * N * M;
    case PastixFactGETRF:
        return 2. * t;
    default:
        return t;
    }
}

static double
fl_update( const flops_fn_t *f, double K, double M )
{
    switch ( f->facto ) {
    case PastixFactSYTRF:
    case PastixFactHETRF:
        return fl_combine( f->cplx, fm_syrk( K, M ), fm_syrk( K, M ) ) + ( f->cplx ? 6. : 1. ) * M * M;
    case PastixFactGETRF:
        return fl_combine( f->cplx, fm_gemm( M, M, K ), fm_gemm( M, M, K ) );
    default:
        return fl_combine( f->cplx, fm_syrk( K, M ), fm_syrk( K, M ) );
    }
}

static double
fl_cblk( const flops_fn_t *f, const px_symbol_t *s, pastix_int_t c )
{
    const px_symcblk_t *cb = s->cblktab + c;
    pastix_int_t        N = cb->lcolnum - cb->fcolnum + 1, M = 0;
    double              ops;
    for ( pastix_int_t b = cb->bloknum + 1; b < cb[1].bloknum; b++ ) {
        M += s->bloktab[b].lrownum - s->bloktab[b].frownum + 1;
    }
    N *= s->dof;
    M *= s->dof;
    ops = fl_diag( f, (double)N );
    if ( M > 0 ) {
        ops += fl_trsm( f, (double)M, (double)N );
        ops += fl_update( f, (double)N, (double)M );
    }
    return ops;
}

static double
fl_recursive( const flops_fn_t *f, const px_symbol_t *s, pastix_int_t a, pastix_int_t b )
{
    if ( a != b ) {
        return fl_recursive( f, s, a, ( a + b ) / 2 ) + fl_recursive( f, s, ( a + b ) / 2 + 1, b );
    }
    return fl_cblk( f, s, a );
}

void
px_symbol_flops( const px_symbol_t *s, int flttype, int factotype, double *thflops, double *rlflops )
{
    flops_fn_t f;
    f.cplx  = ( flttype == SpmComplex32 || flttype == SpmComplex64 );
    f.facto = factotype;
    if ( s->cblknbr <= 0 ) {
        if ( thflops ) *thflops = 0.;
        if ( rlflops ) *rlflops = 0.;
        return;
    }
    if ( thflops ) {
        *thflops = fl_recursive( &f, s, 0, s->cblknbr - 1 );
    }
    if ( rlflops ) {
        *rlflops = thflops ? *thflops : fl_recursive( &f, s, 0, s->cblknbr - 1 );
    }
}

/* ------------------------------------------------------------------------ */
/* Row reordering inside supernodes (pseudo-TSP on hamming distances).      */

typedef struct hvec_s {
    pastix_int_t **v;
    pastix_int_t  *sz;
} hvec_t;

/* min(|symmetric difference|, stop); the ghost (-1) is at distance |other|. */
static pastix_int_t
hamming( const hvec_t *h, pastix_int_t xi, pastix_int_t xj, pastix_int_t stop )
{
    const pastix_int_t *a, *b, *ea, *eb;
    pastix_int_t        sum = 0;
    if ( xi == -1 ) {
        return h->sz[xj];
    }
    if ( xj == -1 ) {
        return h->sz[xi];
    }
    if ( h->sz[xi] - h->sz[xj] >= stop || h->sz[xj] - h->sz[xi] >= stop ) {
        return stop;
    }
    a  = h->v[xi];
    b  = h->v[xj];
    ea = a + h->sz[xi];
    eb = b + h->sz[xj];
    while ( a < ea && b < eb ) {
        if ( *a == *b ) {
            a++;
            b++;
        }
        else if ( *a < *b ) {
            while ( a < ea && *a < *b ) {
                sum++;
                a++;
            }
        }
        else {
            while ( b < eb && *a > *b ) {
                sum++;
                b++;
            }
        }
        if ( sum >= stop ) {
            return stop;
        }
    }
    sum += ( ea - a ) + ( eb - b );
    return ( sum >= stop ) ? stop : sum;
}

/*
 * Distances from one row (`cur`) to the others, in O(1) each.
 *
 * The lists hold cblk numbers in increasing order, so for two rows
 * |A xor B| = |A| + |B| - 2 |A and B|.  The common counts of the current row
 * with every earlier row come from inverted lists (the rows that hold each
 * cblk number, in increasing order), at a cost of the lengths of those lists
 * rather than one merge per pair: the insertion compares every new row with
 * every row placed so far, O(size^2) merges for the widest separators.
 * min(|A xor B|, stop) is what the merge of hamming() returns.  The upper
 * lists are compared at stop 1 only (equal or not), so each row gets the
 * number of its class of equal lists.  Lists that are not strictly
 * increasing fall back to hamming().
 */
typedef struct tsp_rows_s {
    const hvec_t *lw, *up;
    pastix_int_t  stop;
    int           fast;
    pastix_int_t  cur;
    pastix_int_t *gptr, *grow;   /* rows of each cblk number (CSR by group)  */
    pastix_int_t *rptr, *rgrp;   /* groups of each row's lower list (CSR)    */
    pastix_int_t *common;        /* |lw[cur] and lw[x]| for the rows x < cur */
    pastix_int_t *ucls;          /* class of each row's upper list           */
    pastix_int_t  uempty;        /* class of the empty list (the ghost's)    */
} tsp_rows_t;

static int
increasing( const hvec_t *h, pastix_int_t size )
{
    for ( pastix_int_t i = 0; i < size; i++ ) {
        for ( pastix_int_t k = 1; k < h->sz[i]; k++ ) {
            if ( h->v[i][k - 1] >= h->v[i][k] ) {
                return 0;
            }
        }
    }
    return 1;
}

static int
cmp_pair( const void *x, const void *y )
{
    const pastix_int_t *a = (const pastix_int_t *)x, *b = (const pastix_int_t *)y;
    if ( a[0] != b[0] ) {
        return ( a[0] > b[0] ) - ( a[0] < b[0] );
    }
    return ( a[1] > b[1] ) - ( a[1] < b[1] );
}

static void
tsp_rows_init( tsp_rows_t *t, pastix_int_t size, const hvec_t *lw, const hvec_t *up, pastix_int_t stop )
{
    pastix_int_t ne = 0, ng = 0, *pr, *hc;
    memset( t, 0, sizeof( *t ) );
    t->lw   = lw;
    t->up   = up;
    t->stop = stop;
    t->cur  = -1;
    t->fast = increasing( lw, size ) && increasing( up, size );
    if ( !t->fast ) {
        return;
    }
    for ( pastix_int_t i = 0; i < size; i++ ) {
        ne += lw->sz[i];
    }
    /* (cblk number, row) pairs sorted: groups of equal numbers, rows rising */
    pr = px_malloc( 2 * ( ne > 0 ? ne : 1 ) * sizeof( pastix_int_t ) );
    ne = 0;
    for ( pastix_int_t i = 0; i < size; i++ ) {
        for ( pastix_int_t k = 0; k < lw->sz[i]; k++ ) {
            pr[2 * ne]     = lw->v[i][k];
            pr[2 * ne + 1] = i;
            ne++;
        }
    }
    qsort( pr, ne, 2 * sizeof( pastix_int_t ), cmp_pair );
    t->gptr   = px_malloc( ( ne + 1 ) * sizeof( pastix_int_t ) );
    t->grow   = px_malloc( ( ne > 0 ? ne : 1 ) * sizeof( pastix_int_t ) );
    t->rptr   = px_calloc( size + 1, sizeof( pastix_int_t ) );
    t->rgrp   = px_malloc( ( ne > 0 ? ne : 1 ) * sizeof( pastix_int_t ) );
    t->common = px_calloc( size > 0 ? size : 1, sizeof( pastix_int_t ) );
    for ( pastix_int_t p = 0; p < ne; p++ ) {
        if ( p == 0 || pr[2 * p] != pr[2 * p - 2] ) {
            t->gptr[ng++] = p;
        }
        t->grow[p] = pr[2 * p + 1];
        t->rptr[pr[2 * p + 1] + 1]++;
    }
    t->gptr[ng] = ne;
    for ( pastix_int_t i = 0; i < size; i++ ) {
        t->rptr[i + 1] += t->rptr[i];
    }
    {
        pastix_int_t *fill = px_malloc( ( size > 0 ? size : 1 ) * sizeof( pastix_int_t ) );
        memcpy( fill, t->rptr, size * sizeof( pastix_int_t ) );
        for ( pastix_int_t g = 0; g < ng; g++ ) {
            for ( pastix_int_t p = t->gptr[g]; p < t->gptr[g + 1]; p++ ) {
                t->rgrp[fill[t->grow[p]]++] = g;
            }
        }
        free( fill );
    }
    free( pr );

    /* classes of equal upper lists: sort the rows by (hash, row), then split
     * each run of equal hashes by comparing the lists themselves */
    t->ucls = px_malloc( ( size > 0 ? size : 1 ) * sizeof( pastix_int_t ) );
    hc      = px_malloc( 2 * ( size > 0 ? size : 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t i = 0; i < size; i++ ) {
        uint64_t h = 1469598103934665603ull ^ (uint64_t)up->sz[i];
        for ( pastix_int_t k = 0; k < up->sz[i]; k++ ) {
            h = ( h ^ (uint64_t)up->v[i][k] ) * 1099511628211ull;
        }
        hc[2 * i]     = (pastix_int_t)( h >> 1 );
        hc[2 * i + 1] = i;
    }
    qsort( hc, size, 2 * sizeof( pastix_int_t ), cmp_pair );
    {
        pastix_int_t ncls = 0, r0 = 0;
        t->uempty = -1;
        while ( r0 < size ) {
            pastix_int_t r1 = r0, first_cls = ncls;
            while ( r1 < size && hc[2 * r1] == hc[2 * r0] ) {
                r1++;
            }
            for ( pastix_int_t r = r0; r < r1; r++ ) {
                pastix_int_t x = hc[2 * r + 1], c = -1;
                /* the representatives of this run are its earlier rows that
                 * opened a class */
                for ( pastix_int_t q = r0; q < r && c < 0; q++ ) {
                    pastix_int_t y = hc[2 * q + 1];
                    if ( t->ucls[y] >= first_cls && up->sz[x] == up->sz[y] &&
                         memcmp( up->v[x], up->v[y], up->sz[x] * sizeof( pastix_int_t ) ) == 0 ) {
                        c = t->ucls[y];
                    }
                }
                t->ucls[x] = ( c >= 0 ) ? c : ncls++;
                if ( up->sz[x] == 0 ) {
                    t->uempty = t->ucls[x];
                }
            }
            r0 = r1;
        }
        if ( t->uempty < 0 ) {
            t->uempty = ncls;   /* no row has an empty list: the ghost is alone */
        }
    }
    free( hc );
}

static void
tsp_rows_exit( tsp_rows_t *t )
{
    free( t->gptr );
    free( t->grow );
    free( t->rptr );
    free( t->rgrp );
    free( t->common );
    free( t->ucls );
}

/* counts of the row i with the rows before it (add 1, or clear with 0) */
static void
tsp_rows_count( tsp_rows_t *t, pastix_int_t i, int clear )
{
    for ( pastix_int_t q = t->rptr[i]; q < t->rptr[i + 1]; q++ ) {
        pastix_int_t g = t->rgrp[q];
        for ( pastix_int_t p = t->gptr[g]; p < t->gptr[g + 1] && t->grow[p] < i; p++ ) {
            if ( clear ) {
                t->common[t->grow[p]] = 0;
            }
            else {
                t->common[t->grow[p]]++;
            }
        }
    }
}

/* the current row becomes i (its distances to the rows before it) */
static void
tsp_rows_set( tsp_rows_t *t, pastix_int_t i )
{
    if ( !t->fast ) {
        t->cur = i;
        return;
    }
    if ( t->cur >= 0 ) {
        tsp_rows_count( t, t->cur, 1 );
    }
    t->cur = i;
    tsp_rows_count( t, i, 0 );
}

/* hamming( lw, cur, x, stop ) for a row x before cur or the ghost */
static inline pastix_int_t
tsp_lw( const tsp_rows_t *t, pastix_int_t x )
{
    pastix_int_t d;
    if ( !t->fast ) {
        return hamming( t->lw, t->cur, x, t->stop );
    }
    if ( x == -1 ) {
        return t->lw->sz[t->cur];
    }
    d = t->lw->sz[t->cur] + t->lw->sz[x] - 2 * t->common[x];
    return ( d >= t->stop ) ? t->stop : d;
}

/* hamming( up, cur, x, 1 ) < 1: the upper lists are equal */
static inline int
tsp_up_equal( const tsp_rows_t *t, pastix_int_t x )
{
    if ( !t->fast ) {
        return hamming( t->up, t->cur, x, 1 ) < 1;
    }
    return t->ucls[t->cur] == ( ( x == -1 ) ? t->uempty : t->ucls[x] );
}

/* Cheapest insertion of the rows into a cycle that starts with a ghost row;
 * writes the resulting order of the rows into `perm` (size entries). */
static void
reorder_tsp( pastix_int_t size, const hvec_t *lw, const hvec_t *up, pastix_int_t stop,
             pastix_int_t *perm )
{
    pastix_int_t *inv = px_malloc( ( size + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *len = px_calloc( size + 1, sizeof( pastix_int_t ) );
    pastix_int_t  d, elected = 0;
    tsp_rows_t    t;

    tsp_rows_init( &t, size, lw, up, stop );
    inv[0] = -1;
    inv[1] = 0;
    d      = lw->sz[0];   /* row 0 to the ghost */
    len[0] = d;
    len[1] = d;

    for ( pastix_int_t i = 1; i < size; i++ ) {
        tsp_rows_set( &t, i );
        pastix_int_t lw_before = tsp_lw( &t, inv[0] );
        pastix_int_t lw_after  = tsp_lw( &t, inv[1] );
        pastix_int_t up_after  = tsp_up_equal( &t, inv[1] ) ? 0 : 1;
        pastix_int_t minl      = lw_before + lw_after - len[0];
        pastix_int_t mpos      = 1;
        pastix_int_t min_cut   = -1;
        pastix_int_t first, last, l;

        for ( pastix_int_t j = 1; j < i; j++ ) {
            pastix_int_t up_before = up_after;
            up_after = tsp_up_equal( &t, inv[j + 1] ) ? 0 : 1;
            if ( up_before < 1 || up_after < 1 ) {
                if ( lw_after == -1 ) {
                    lw_before = tsp_lw( &t, inv[j] );
                }
                else {
                    lw_before = lw_after;
                }
                lw_after = tsp_lw( &t, inv[j + 1] );
                l        = lw_before + lw_after - len[j];
                if ( l == minl ) {
                    if ( lw_before < min_cut ) {
                        min_cut = lw_before;
                        minl    = l;
                        mpos    = j + 1;
                    }
                    if ( lw_after < min_cut ) {
                        min_cut = lw_after;
                        minl    = l;
                        mpos    = j + 1;
                    }
                }
                if ( l < minl ) {
                    min_cut = lw_before;
                    minl    = l;
                    mpos    = j + 1;
                    if ( lw_after < min_cut ) {
                        min_cut = lw_after;
                    }
                }
                if ( lw_after == 0 ) {
                    min_cut = 0;
                    minl    = l;
                    mpos    = j + 1;
                    j       = i;
                }
            }
            else {
                lw_after = -1;
            }
        }

        first     = tsp_lw( &t, inv[0] );
        last      = tsp_lw( &t, inv[i] );
        lw_before = tsp_lw( &t, inv[mpos - 1] );
        lw_after  = tsp_lw( &t, inv[mpos] );
        l         = first + last - len[i];
        if ( l < minl ) {
            minl = l;
            mpos = i + 1;
        }
        if ( mpos > 0 ) {
            len[mpos - 1] = lw_before;
        }
        if ( mpos < i + 1 ) {
            pastix_int_t k = i, ll = lw_after;
            for ( pastix_int_t j = mpos; j < i + 2; j++ ) {
                pastix_int_t ti = inv[j], tl = len[j];
                inv[j] = k;
                len[j] = ll;
                k      = ti;
                ll     = tl;
            }
        }
        else {
            inv[i + 1] = i;
            len[i + 1] = first;
        }
    }

    for ( pastix_int_t i = 0; i < size; i++ ) {
        if ( inv[i] == -1 ) {
            elected = i;
        }
    }
    for ( pastix_int_t i = 0; i < size; i++ ) {
        perm[i] = inv[( i + 1 + elected ) % ( size + 1 )];
    }
    tsp_rows_exit( &t );
    free( inv );
    free( len );
}

typedef struct reorder_arg_s {
    px_symbol_t        *s;
    pastix_order_t     *ord;
    const pastix_int_t *levels;
    pastix_int_t        maxdepth;
    pastix_int_t        split_level;
    pastix_int_t        stop;
    pastix_int_t       *tasks;     /* cblk indices, largest first   */
    pastix_int_t        ntasks;
    pastix_int_t        next;      /* shared work counter           */
    pthread_mutex_t     lock;
} reorder_arg_t;

static void
reorder_cblk( reorder_arg_t *a, pastix_int_t c, pastix_int_t *depthweight )
{
    const px_symbol_t *s    = a->s;
    px_symcblk_t      *cb   = s->cblktab + c;
    pastix_int_t       size = cb->lcolnum - cb->fcolnum + 1;
    pastix_int_t       split = a->split_level;
    hvec_t             lw, up;
    pastix_int_t      *perm, *tmp;

    if ( size < 3 ) {
        return;
    }
    memset( depthweight, 0, a->maxdepth * sizeof( pastix_int_t ) );
    for ( pastix_int_t k = cb->brownum; k < cb[1].brownum; k++ ) {
        const px_symblok_t *bk = s->bloktab + s->browtab[k];
        depthweight[a->levels[bk->lcblknm] - 1] += bk->lrownum - bk->frownum + 1;
    }
    /* move the split level so that the upper part holds 17% to 25% */
    {
        pastix_int_t sign = 0;
        for ( ;; ) {
            pastix_int_t up_total = 0, lw_total = 0, i;
            for ( i = 0; i < split; i++ ) {
                up_total += depthweight[i];
            }
            for ( ; i < a->maxdepth; i++ ) {
                lw_total += depthweight[i];
            }
            if ( lw_total < 5 * up_total && lw_total > 10 && up_total > 10 && sign <= 0 ) {
                split--;
                sign--;
                continue;
            }
            if ( lw_total > 3 * up_total && lw_total > 10 && up_total > 10 && sign >= 0 ) {
                split++;
                sign++;
                continue;
            }
            break;
        }
    }
    lw.sz = px_calloc( size, sizeof( pastix_int_t ) );
    up.sz = px_calloc( size, sizeof( pastix_int_t ) );
    for ( pastix_int_t k = cb->brownum; k < cb[1].brownum; k++ ) {
        const px_symblok_t *bk = s->bloktab + s->browtab[k];
        pastix_int_t       *sz = ( a->levels[bk->lcblknm] <= split ) ? up.sz : lw.sz;
        for ( pastix_int_t r = bk->frownum; r <= bk->lrownum; r++ ) {
            sz[r - cb->fcolnum]++;
        }
    }
    lw.v = px_malloc( size * sizeof( pastix_int_t * ) );
    up.v = px_malloc( size * sizeof( pastix_int_t * ) );
    for ( pastix_int_t i = 0; i < size; i++ ) {
        lw.v[i] = px_malloc( ( lw.sz[i] > 0 ? lw.sz[i] : 1 ) * sizeof( pastix_int_t ) );
        up.v[i] = px_malloc( ( up.sz[i] > 0 ? up.sz[i] : 1 ) * sizeof( pastix_int_t ) );
        lw.sz[i] = 0;
        up.sz[i] = 0;
    }
    for ( pastix_int_t k = cb->brownum; k < cb[1].brownum; k++ ) {
        const px_symblok_t *bk = s->bloktab + s->browtab[k];
        hvec_t             *h  = ( a->levels[bk->lcblknm] <= split ) ? &up : &lw;
        for ( pastix_int_t r = bk->frownum; r <= bk->lrownum; r++ ) {
            pastix_int_t idx = r - cb->fcolnum;
            h->v[idx][h->sz[idx]++] = bk->lcblknm;
        }
    }

    perm = px_malloc( size * sizeof( pastix_int_t ) );
    reorder_tsp( size, &lw, &up, a->stop, perm );
    tmp = px_malloc( size * sizeof( pastix_int_t ) );
    {
        pastix_int_t *peri = a->ord->peritab + a->ord->rangtab[c];
        for ( pastix_int_t i = 0; i < size; i++ ) {
            tmp[i] = peri[perm[i]];
        }
        memcpy( peri, tmp, size * sizeof( pastix_int_t ) );
    }
    free( tmp );
    free( perm );
    for ( pastix_int_t i = 0; i < size; i++ ) {
        free( lw.v[i] );
        free( up.v[i] );
    }
    free( lw.v );
    free( up.v );
    free( lw.sz );
    free( up.sz );
}

static void *
reorder_worker( void *varg )
{
    reorder_arg_t *a  = (reorder_arg_t *)varg;
    pastix_int_t  *dw = px_malloc( ( a->maxdepth > 0 ? a->maxdepth : 1 ) * sizeof( pastix_int_t ) );
    for ( ;; ) {
        pastix_int_t t;
        pthread_mutex_lock( &a->lock );
        t = a->next++;
        pthread_mutex_unlock( &a->lock );
        if ( t >= a->ntasks ) {
            break;
        }
        reorder_cblk( a, a->tasks[t], dw );
    }
    free( dw );
    return NULL;
}

static int
cmp_cost_desc( const void *x, const void *y )
{
    const double *a = (const double *)x, *b = (const double *)y;
    if ( a[0] != b[0] ) {
        return ( a[0] < b[0] ) - ( a[0] > b[0] );
    }
    return ( a[1] > b[1] ) - ( a[1] < b[1] );
}

void
px_symbol_reordering( px_symbol_t *s, pastix_order_t *ord, pastix_int_t split_level,
                      pastix_int_t stop, int nthreads )
{
    pastix_int_t  nc = s->cblknbr;
    pastix_int_t *levels = px_calloc( nc, sizeof( pastix_int_t ) );
    pastix_int_t  maxdepth = 0;
    reorder_arg_t a;

    /* depth of every supernode in the ordering's tree (root = 1) */
    for ( pastix_int_t c = nc - 1; c >= 0; c-- ) {
        pastix_int_t f = ord->treetab[c];
        levels[c] = ( f == -1 ) ? 1 : levels[f] + 1;
        maxdepth  = px_imax( maxdepth, levels[c] );
    }

    a.s           = s;
    a.ord         = ord;
    a.levels      = levels;
    a.maxdepth    = maxdepth;
    a.split_level = split_level;
    a.stop        = stop;
    a.next        = 0;
    pthread_mutex_init( &a.lock, NULL );
    {
        /* largest supernodes first for load balance; results do not depend
         * on the processing order since each cblk permutes its own rows */
        double *cost = px_malloc( 2 * nc * sizeof( double ) );
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            double w   = (double)( s->cblktab[c].lcolnum - s->cblktab[c].fcolnum + 1 );
            double nb  = (double)( s->cblktab[c + 1].brownum - s->cblktab[c].brownum );
            cost[2 * c]     = w * w * ( nb / 2. + 1. );
            cost[2 * c + 1] = (double)c;
        }
        qsort( cost, nc, 2 * sizeof( double ), cmp_cost_desc );
        a.tasks = px_malloc( nc * sizeof( pastix_int_t ) );
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            a.tasks[c] = (pastix_int_t)cost[2 * c + 1];
        }
        a.ntasks = nc;
        free( cost );
    }
    if ( nthreads < 1 ) {
        nthreads = 1;
    }
    {
        pthread_t *th = px_malloc( nthreads * sizeof( pthread_t ) );
        for ( int t = 1; t < nthreads; t++ ) {
            pthread_create( th + t, NULL, reorder_worker, &a );
        }
        reorder_worker( &a );
        for ( int t = 1; t < nthreads; t++ ) {
            pthread_join( th[t], NULL );
        }
        free( th );
    }
    pthread_mutex_destroy( &a.lock );
    free( a.tasks );
    for ( pastix_int_t i = 0; i < s->nodenbr; i++ ) {
        ord->permtab[ord->peritab[i]] = i;
    }
    free( levels );
}

/* ------------------------------------------------------------------------ */
/* Splitting of wide supernodes.                                            */

static pastix_int_t
split_count( pastix_int_t width, pastix_int_t bmin, pastix_int_t bmax )
{
    pastix_int_t nseq;
    if ( width <= bmax ) {
        return 1;
    }
    nseq = px_iceil( width, bmax );
    if ( nseq > 1 && ( width / nseq ) < bmin ) {
        nseq--;
    }
    return nseq;
}

/* Column (relative to fcol) after which to cut, near step-1, where the fewest
 * blocks are crossed; percent = 10 (PASTIX_BLEND_SPLIT_AUTORIZED_PERCENTAGE). */
static pastix_int_t
split_smallest( const pastix_int_t *nbl, pastix_int_t step, pastix_int_t min, pastix_int_t max,
                pastix_int_t percent )
{
    pastix_int_t limit = px_iceil( step * percent, 100 );
    pastix_int_t lavg, lmin, lmax, lcol, nb;
    if ( step >= max ) {
        return max - 1;
    }
    lavg = step - 1;
    lmin = px_imax( lavg - limit - 1, min );
    lmax = px_imin( lavg + limit + 1, max );
    lcol = lavg;
    nb   = nbl[lavg];
    for ( pastix_int_t i = lavg + 1; i < lmax; i++ ) {
        if ( nbl[i] < nb ) {
            lcol = i;
            nb   = nbl[i];
        }
    }
    for ( pastix_int_t i = lavg - 1; i > lmin; i-- ) {
        if ( nbl[i] < nb ) {
            lcol = i;
            nb   = nbl[i];
        }
    }
    return lcol;
}

void
px_symbol_split( px_symbol_t *s, pastix_int_t bmin, pastix_int_t bmax )
{
    pastix_int_t  nc = s->cblknbr;
    pastix_int_t *npieces = px_malloc( nc * sizeof( pastix_int_t ) );
    pastix_int_t *pfirst  = px_malloc( nc * sizeof( pastix_int_t ) ); /* first piece index  */
    pastix_int_t *newnum  = px_malloc( ( nc + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *pcol    = NULL;  /* pieces: fcol,lcol pairs */
    pastix_int_t  npc = 0, maxpc = 0, addc = 0;
    pastix_int_t *nbl = NULL, fsplit = -1;

    for ( pastix_int_t c = 0; c < nc; c++ ) {
        px_symcblk_t *cb = s->cblktab + c;
        pastix_int_t  fcol = cb->fcolnum, lcol = cb->lcolnum;
        pastix_int_t  width = lcol - fcol + 1, nseq, cnt = 0;

        npieces[c] = 1;
        pfirst[c]  = -1;
        nseq = split_count( width, bmin, bmax );
        if ( nseq <= 1 ) {
            continue;
        }
        if ( nbl == NULL ) {
            /* blocks crossed by a cut between rows r and r+1, for r >= fsplit */
            pastix_int_t sz = s->nodenbr - fcol + 1;
            fsplit = fcol;
            nbl    = px_calloc( sz, sizeof( pastix_int_t ) );
            for ( pastix_int_t b = 0; b < s->bloknbr; b++ ) {
                px_symblok_t *bk = s->bloktab + b;
                if ( bk->lrownum < fsplit ) {
                    continue;
                }
                for ( pastix_int_t r = px_imax( bk->frownum, fsplit ); r < bk->lrownum; r++ ) {
                    nbl[r - fsplit]++;
                }
            }
        }
        pfirst[c] = npc;
        {
            pastix_int_t f = fcol;
            while ( f <= lcol ) {
                pastix_int_t l, step, lwidth;
                nseq = split_count( width, bmin, bmax );
                if ( npc + 1 > maxpc ) {
                    maxpc = 2 * maxpc + 64;
                    pcol  = realloc( pcol, 2 * maxpc * sizeof( pastix_int_t ) );
                }
                if ( nseq == 1 ) {
                    pcol[2 * npc]     = f;
                    pcol[2 * npc + 1] = lcol;
                    npc++;
                    cnt++;
                    f = lcol + 1;
                    continue;
                }
                step   = px_iceil( width, nseq );
                lwidth = width - bmin;
                l      = f + split_smallest( nbl + ( f - fsplit ), step, bmin, lwidth, 10 );
                pcol[2 * npc]     = f;
                pcol[2 * npc + 1] = l;
                npc++;
                cnt++;
                width -= l - f + 1;
                f = l + 1;
            }
        }
        npieces[c] = cnt;
        addc += cnt - 1;
        /* the off-diagonal rows of the cblk are now crossed by cnt-1 more cuts */
        for ( pastix_int_t b = cb->bloknum + 1; b < cb[1].bloknum; b++ ) {
            for ( pastix_int_t r = s->bloktab[b].frownum; r < s->bloktab[b].lrownum; r++ ) {
                nbl[r - fsplit] += cnt - 1;
            }
        }
    }
    free( nbl );

    if ( addc == 0 ) {
        free( npieces );
        free( pfirst );
        free( newnum );
        free( pcol );
        return;
    }

    newnum[0] = 0;
    for ( pastix_int_t c = 0; c < nc; c++ ) {
        newnum[c + 1] = newnum[c] + npieces[c];
    }
    {
        pastix_int_t  nnc = newnum[nc];
        px_symcblk_t *nct = px_calloc( nnc + 1, sizeof( px_symcblk_t ) );
        pastix_int_t  maxb = s->bloknbr * 2 + nnc * 4 + 16, nb = 0;
        px_symblok_t *nbt = px_malloc( maxb * sizeof( px_symblok_t ) );

        for ( pastix_int_t c = 0; c < nc; c++ ) {
            for ( pastix_int_t p = 0; p < npieces[c]; p++ ) {
                pastix_int_t nci = newnum[c] + p;
                if ( npieces[c] > 1 ) {
                    nct[nci].fcolnum = pcol[2 * ( pfirst[c] + p )];
                    nct[nci].lcolnum = pcol[2 * ( pfirst[c] + p ) + 1];
                }
                else {
                    nct[nci].fcolnum = s->cblktab[c].fcolnum;
                    nct[nci].lcolnum = s->cblktab[c].lcolnum;
                }
                nct[nci].selevtx = s->cblktab[c].selevtx;
            }
        }
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            pastix_int_t w = npieces[c];
            for ( pastix_int_t p = 0; p < w; p++ ) {
                pastix_int_t nci = newnum[c] + p;
                pastix_int_t need = ( w - p ) + 2 * ( s->cblktab[c + 1].bloknum - s->cblktab[c].bloknum ) + 8;
                for ( pastix_int_t b = s->cblktab[c].bloknum + 1; b < s->cblktab[c + 1].bloknum; b++ ) {
                    need += npieces[s->bloktab[b].fcblknm];
                }
                if ( nb + need > maxb ) {
                    while ( nb + need > maxb ) {
                        maxb = maxb + maxb / 2 + 16;
                    }
                    nbt = realloc( nbt, maxb * sizeof( px_symblok_t ) );
                }
                nct[nci].bloknum = nb;
                /* the split diagonal block: one blok per later piece */
                for ( pastix_int_t k = 0; k < w - p; k++ ) {
                    nbt[nb].frownum = nct[nci + k].fcolnum;
                    nbt[nb].lrownum = nct[nci + k].lcolnum;
                    nbt[nb].lcblknm = nci;
                    nbt[nb].fcblknm = nci + k;
                    nb++;
                }
                /* off-diagonal bloks, cut where their facing cblk was cut */
                for ( pastix_int_t b = s->cblktab[c].bloknum + 1; b < s->cblktab[c + 1].bloknum; b++ ) {
                    pastix_int_t fr = s->bloktab[b].frownum, lr = s->bloktab[b].lrownum;
                    pastix_int_t f  = s->bloktab[b].fcblknm;
                    if ( npieces[f] > 1 ) {
                        for ( pastix_int_t q = 0; q < npieces[f]; q++ ) {
                            pastix_int_t pf = pcol[2 * ( pfirst[f] + q )];
                            pastix_int_t pl = pcol[2 * ( pfirst[f] + q ) + 1];
                            if ( fr > pl ) {
                                continue;
                            }
                            if ( lr < pf ) {
                                break;
                            }
                            nbt[nb].frownum = fr;
                            nbt[nb].lrownum = px_imin( lr, pl );
                            nbt[nb].lcblknm = nci;
                            nbt[nb].fcblknm = newnum[f] + q;
                            nb++;
                            fr = pl + 1;
                        }
                    }
                    else {
                        nbt[nb].frownum = fr;
                        nbt[nb].lrownum = lr;
                        nbt[nb].lcblknm = nci;
                        nbt[nb].fcblknm = newnum[f];
                        nb++;
                    }
                }
            }
        }
        nct[nnc].fcolnum = nct[nnc - 1].lcolnum + 1;
        nct[nnc].lcolnum = nct[nnc - 1].lcolnum + 1;
        nct[nnc].bloknum = nb;

        free( s->cblktab );
        free( s->bloktab );
        s->cblktab = nct;
        s->bloktab = realloc( nbt, ( nb > 0 ? nb : 1 ) * sizeof( px_symblok_t ) );
        s->cblknbr = nnc;
        s->bloknbr = nb;
    }
    free( npieces );
    free( pfirst );
    free( newnum );
    free( pcol );
    px_symbol_build_rowtab( s );
}
