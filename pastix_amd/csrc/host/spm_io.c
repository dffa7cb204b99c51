/*
 * spm_io.c -- matrix input for the PaStiX C API: spmReadDriver and its file
 * formats, the entry the reference's examples and tests read matrices through
 * (example/simple.c:63-66, common/get_options.c:150-215; the test matrices of
 * test/CMakeLists.txt:144-153).
 *
 * The SPM library that provides spmReadDriver in the reference build is an
 * empty submodule in the reference tree, so these readers follow the
 * published file formats:
 *   - Harwell-Boeing (Duff, Grimes, Lewis, "Users' guide for the Harwell-
 *     Boeing sparse matrix collection", 1992): four header lines (five with
 *     right-hand sides), then column pointers, row indices and values in
 *     fixed-width Fortran fields.  Assembled real, complex and pattern
 *     matrices, unsymmetric, symmetric, Hermitian.  RSA files are
 *     Harwell-Boeing files (get_options.c:158-161 routes --rsa to the HB
 *     driver);
 *   - Matrix Market coordinate format (NIST, "The Matrix Market exchange
 *     formats", 1996): real, integer, complex or pattern values, general,
 *     symmetric, skew-symmetric or Hermitian;
 *   - IJV: "n nnz" then one "i j value" (1-based) per line;
 *   - the Laplacian generator string "[<type>:]<nx>[:<ny>[:<nz>[:<alpha>[:<beta>]]]]"
 *     (7-point stencil, diagonal alpha * 2 * dims, off-diagonals -beta).
 * Every matrix comes back as CSC with 0-based indices, duplicates summed,
 * rows sorted in each column.  Symmetric and Hermitian files keep the one
 * triangle they store.
 */
#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "px_internal.h"

typedef struct { spm_int_t i, j; double re, im; } px_ijv_t;

static int
cmp_ijv( const void *a, const void *b )
{
    const px_ijv_t *x = a, *y = b;
    if ( x->j != y->j ) return ( x->j < y->j ) ? -1 : 1;
    if ( x->i != y->i ) return ( x->i < y->i ) ? -1 : 1;
    return 0;
}

/* Sorted CSC (0-based) from coordinate triplets (0-based), duplicates summed. */
static int
px_ijv_to_csc( spmatrix_t *spm, spm_int_t n, spm_int_t nnz, px_ijv_t *t, int cplx, int pattern,
               spm_mtxtype_t mtx )
{
    spm_int_t k, ne = 0;
    qsort( t, (size_t)nnz, sizeof( px_ijv_t ), cmp_ijv );
    for ( k = 0; k < nnz; k++ ) {
        if ( ne > 0 && t[ne - 1].i == t[k].i && t[ne - 1].j == t[k].j ) {
            t[ne - 1].re += t[k].re;
            t[ne - 1].im += t[k].im;
        }
        else {
            t[ne++] = t[k];
        }
    }
    spmInit( spm );
    spm->mtxtype = mtx;
    spm->flttype = cplx ? SpmComplex64 : SpmDouble;
    spm->fmttype = SpmCSC;
    spm->baseval = 0;
    spm->n       = n;
    spm->nnz     = ne;
    spm->dof     = 1;
    spm->colptr  = px_calloc( (size_t)n + 1, sizeof( spm_int_t ) );
    spm->rowptr  = px_malloc( (size_t)( ne > 0 ? ne : 1 ) * sizeof( spm_int_t ) );
    spm->values  = px_malloc( (size_t)( ne > 0 ? ne : 1 ) * ( cplx ? 16 : 8 ) );
    for ( k = 0; k < ne; k++ ) {
        spm->colptr[t[k].j + 1]++;
        spm->rowptr[k] = t[k].i;
        if ( cplx ) {
            ( (double complex *)spm->values )[k] = t[k].re + I * t[k].im;
        }
        else {
            ( (double *)spm->values )[k] = pattern ? 1.0 : t[k].re;
        }
    }
    for ( spm_int_t j = 0; j < n; j++ ) {
        spm->colptr[j + 1] += spm->colptr[j];
    }
    spmUpdateComputedFields( spm );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Harwell-Boeing */

static size_t
px_min_size( size_t a, size_t b )
{
    return a < b ? a : b;
}

/* Fortran edit descriptor of a data format such as (13I6), (1P5E16.8),
 * (3E26.18) or (10I8): the fields per line and their width. */
static int
hb_format( const char *fmt, int *per_line, int *width )
{
    const char *p = fmt;
    while ( *p && *p != '(' ) p++;
    if ( *p == '(' ) p++;
    /* an optional scale factor such as 1P, then <count><letter><width> */
    for ( ;; ) {
        const char *q = p;
        int         cnt = 0;
        while ( isdigit( (unsigned char)*q ) ) cnt = 10 * cnt + ( *q++ - '0' );
        if ( *q == 'P' || *q == 'p' ) {
            p = q + 1;
            if ( *p == ',' ) p++;
            continue;
        }
        if ( strchr( "IiEeDdFfGg", *q ) && *q ) {
            int w = 0;
            q++;
            while ( isdigit( (unsigned char)*q ) ) w = 10 * w + ( *q++ - '0' );
            *per_line = cnt > 0 ? cnt : 1;
            *width    = w;
            return ( w > 0 ) ? 0 : -1;
        }
        return -1;
    }
}

/* Read `count` fixed-width fields (per_line per record, width wide). */
static int
hb_fields( FILE *f, spm_int_t count, int per_line, int width, int real, spm_int_t *iv, double *dv )
{
    char      line[4096];
    spm_int_t got = 0;
    while ( got < count ) {
        size_t len;
        if ( fgets( line, sizeof( line ), f ) == NULL ) {
            return -1;
        }
        len = strlen( line );
        while ( len > 0 && ( line[len - 1] == '\n' || line[len - 1] == '\r' ) ) line[--len] = '\0';
        for ( int c = 0; c < per_line && got < count; c++ ) {
            char   field[128];
            size_t off = (size_t)c * (size_t)width;
            if ( off >= len ) {
                break;   /* short last record */
            }
            size_t w = ( off + (size_t)width <= len ) ? (size_t)width : len - off;
            if ( w >= sizeof( field ) ) {
                return -1;
            }
            memcpy( field, line + off, w );
            field[w] = '\0';
            if ( real ) {
                /* Fortran D exponents; an exponent written without its letter
                 * ("1.0-3") needs one */
                char  buf[160];
                int   o = 0;
                for ( size_t s = 0; field[s] && o < 150; s++ ) {
                    char ch = field[s];
                    if ( ch == 'D' || ch == 'd' ) ch = 'E';
                    if ( ( ch == '-' || ch == '+' ) && o > 0 && ( isdigit( (unsigned char)buf[o - 1] ) || buf[o - 1] == '.' ) ) {
                        buf[o++] = 'E';
                    }
                    buf[o++] = ch;
                }
                buf[o] = '\0';
                char *end;
                dv[got] = strtod( buf, &end );
                if ( end == buf ) {
                    return -1;
                }
            }
            else {
                char *end;
                iv[got] = strtoll( field, &end, 10 );
                if ( end == field ) {
                    return -1;
                }
            }
            got++;
        }
    }
    return 0;
}

static int
px_read_hb( const char *filename, spmatrix_t *spm )
{
    FILE     *f = fopen( filename, "r" );
    char      l1[256], l2[256], l3[256], l4[256], l5[256];
    char      mxtype[4], ptrfmt[32], indfmt[32], valfmt[32];
    long long totcrd = 0, ptrcrd = 0, indcrd = 0, valcrd = 0, rhscrd = 0;
    long long nrow = 0, ncol = 0, nnz = 0, neltvl = 0;
    int       pl, pw, il, iw, vl, vw, cplx, pattern, rc = PASTIX_ERR_FILE;
    spm_int_t *colptr = NULL, *rowind = NULL;
    double   *val = NULL;
    px_ijv_t *t = NULL;

    if ( f == NULL ) {
        px_print_error( "spmReadDriver: cannot open %s", filename );
        return PASTIX_ERR_FILE;
    }
    if ( !fgets( l1, sizeof( l1 ), f ) || !fgets( l2, sizeof( l2 ), f ) || !fgets( l3, sizeof( l3 ), f ) ||
         !fgets( l4, sizeof( l4 ), f ) ) {
        goto out;
    }
    if ( sscanf( l2, "%lld %lld %lld %lld %lld", &totcrd, &ptrcrd, &indcrd, &valcrd, &rhscrd ) < 4 ) {
        goto out;
    }
    memcpy( mxtype, l3, 3 );
    mxtype[3] = '\0';
    if ( sscanf( l3 + 3, "%lld %lld %lld %lld", &nrow, &ncol, &nnz, &neltvl ) < 3 ) {
        goto out;
    }
    /* format fields: 16, 16, 20 (and 20) columns */
    memset( ptrfmt, 0, sizeof( ptrfmt ) );
    memset( indfmt, 0, sizeof( indfmt ) );
    memset( valfmt, 0, sizeof( valfmt ) );
    memcpy( ptrfmt, l4, px_min_size( strlen( l4 ), 16 ) );
    if ( strlen( l4 ) > 16 ) memcpy( indfmt, l4 + 16, px_min_size( strlen( l4 ) - 16, 16 ) );
    if ( strlen( l4 ) > 32 ) memcpy( valfmt, l4 + 32, px_min_size( strlen( l4 ) - 32, 20 ) );
    if ( rhscrd > 0 && !fgets( l5, sizeof( l5 ), f ) ) {
        goto out;
    }
    for ( int c = 0; c < 3; c++ ) mxtype[c] = (char)toupper( (unsigned char)mxtype[c] );
    if ( mxtype[2] != 'A' || nrow != ncol || nrow <= 0 ) {
        px_print_error( "spmReadDriver: %s: only assembled square Harwell-Boeing matrices are read (type %s)",
                        filename, mxtype );
        rc = PASTIX_ERR_BADPARAMETER;
        goto out;
    }
    pattern = ( mxtype[0] == 'P' );
    cplx    = ( mxtype[0] == 'C' );
    if ( hb_format( ptrfmt, &pl, &pw ) || hb_format( indfmt, &il, &iw ) ||
         ( !pattern && hb_format( valfmt, &vl, &vw ) ) ) {
        goto out;
    }
    colptr = px_malloc( (size_t)( ncol + 1 ) * sizeof( spm_int_t ) );
    rowind = px_malloc( (size_t)( nnz > 0 ? nnz : 1 ) * sizeof( spm_int_t ) );
    if ( hb_fields( f, ncol + 1, pl, pw, 0, colptr, NULL ) || hb_fields( f, nnz, il, iw, 0, rowind, NULL ) ) {
        goto out;
    }
    if ( !pattern ) {
        val = px_malloc( (size_t)( nnz > 0 ? nnz : 1 ) * ( cplx ? 2 : 1 ) * sizeof( double ) );
        if ( hb_fields( f, nnz * ( cplx ? 2 : 1 ), vl, vw, 1, NULL, val ) ) {
            goto out;
        }
    }
    t = px_malloc( (size_t)( nnz > 0 ? nnz : 1 ) * sizeof( px_ijv_t ) );
    for ( long long j = 0; j < ncol; j++ ) {
        for ( spm_int_t k = colptr[j] - 1; k < colptr[j + 1] - 1; k++ ) {
            if ( k < 0 || k >= nnz || rowind[k] < 1 || rowind[k] > nrow ) {
                goto out;
            }
            t[k].i  = rowind[k] - 1;
            t[k].j  = j;
            t[k].re = pattern ? 1.0 : ( cplx ? val[2 * k] : val[k] );
            t[k].im = cplx ? val[2 * k + 1] : 0.0;
        }
    }
    rc = px_ijv_to_csc( spm, (spm_int_t)nrow, (spm_int_t)nnz, t, cplx, pattern,
                        mxtype[1] == 'S' ? SpmSymmetric : mxtype[1] == 'H' ? SpmHermitian : SpmGeneral );
out:
    if ( rc == PASTIX_ERR_FILE ) {
        px_print_error( "spmReadDriver: %s is not a readable Harwell-Boeing file", filename );
    }
    free( colptr );
    free( rowind );
    free( val );
    free( t );
    fclose( f );
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Matrix Market coordinate */

static int
px_read_mm( const char *filename, spmatrix_t *spm )
{
    FILE     *f = fopen( filename, "r" );
    char      line[1024], obj[64] = "", fmt[64] = "", field[64] = "", sym[64] = "";
    long long m = 0, n = 0, nnz = 0, k = 0, extra = 0;
    int       cplx, pattern, skew, rc = PASTIX_ERR_FILE;
    px_ijv_t *t = NULL;
    spm_mtxtype_t mtx;

    if ( f == NULL ) {
        px_print_error( "spmReadDriver: cannot open %s", filename );
        return PASTIX_ERR_FILE;
    }
    if ( !fgets( line, sizeof( line ), f ) ||
         sscanf( line, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym ) != 4 ) {
        goto out;
    }
    for ( char *p = fmt; *p; p++ ) *p = (char)tolower( (unsigned char)*p );
    for ( char *p = field; *p; p++ ) *p = (char)tolower( (unsigned char)*p );
    for ( char *p = sym; *p; p++ ) *p = (char)tolower( (unsigned char)*p );
    if ( strcmp( fmt, "coordinate" ) != 0 ) {
        px_print_error( "spmReadDriver: %s: only the coordinate Matrix Market format is read", filename );
        rc = PASTIX_ERR_BADPARAMETER;
        goto out;
    }
    cplx    = !strcmp( field, "complex" );
    pattern = !strcmp( field, "pattern" );
    skew    = !strcmp( sym, "skew-symmetric" );
    mtx     = !strcmp( sym, "symmetric" ) ? SpmSymmetric : !strcmp( sym, "hermitian" ) ? SpmHermitian : SpmGeneral;
    do {
        if ( !fgets( line, sizeof( line ), f ) ) {
            goto out;
        }
    } while ( line[0] == '%' || line[strspn( line, " \t\r\n" )] == '\0' );
    if ( sscanf( line, "%lld %lld %lld", &m, &n, &nnz ) != 3 || m != n || m <= 0 ) {
        goto out;
    }
    /* a skew-symmetric file stores the strict lower triangle: both
     * triangles are expanded into a general matrix */
    if ( skew ) {
        extra = nnz;
    }
    t = px_malloc( (size_t)( nnz + extra > 0 ? nnz + extra : 1 ) * sizeof( px_ijv_t ) );
    for ( k = 0; k < nnz; ) {
        long long i, j;
        double    re = 1.0, im = 0.0;
        int       nf;
        if ( !fgets( line, sizeof( line ), f ) ) {
            goto out;
        }
        if ( line[0] == '%' || line[strspn( line, " \t\r\n" )] == '\0' ) {
            continue;
        }
        nf = sscanf( line, "%lld %lld %lf %lf", &i, &j, &re, &im );
        if ( nf < ( pattern ? 2 : cplx ? 4 : 3 ) || i < 1 || i > m || j < 1 || j > n ) {
            goto out;
        }
        t[k].i  = i - 1;
        t[k].j  = j - 1;
        t[k].re = re;
        t[k].im = cplx ? im : 0.0;
        k++;
    }
    if ( skew ) {
        for ( long long e = 0; e < nnz; e++ ) {
            t[nnz + e].i  = t[e].j;
            t[nnz + e].j  = t[e].i;
            t[nnz + e].re = -t[e].re;
            t[nnz + e].im = -t[e].im;
        }
        mtx = SpmGeneral;
    }
    rc = px_ijv_to_csc( spm, (spm_int_t)n, (spm_int_t)( nnz + extra ), t, cplx, pattern, mtx );
out:
    if ( rc == PASTIX_ERR_FILE ) {
        px_print_error( "spmReadDriver: %s is not a readable Matrix Market file", filename );
    }
    free( t );
    fclose( f );
    return rc;
}

/* ------------------------------------------------------------------------ */
/* IJV: "n nnz" then "i j value" (1-based) per line, general */

static int
px_read_ijv( const char *filename, spmatrix_t *spm )
{
    FILE     *f = fopen( filename, "r" );
    long long n = 0, nnz = 0;
    px_ijv_t *t;
    int       rc;
    if ( f == NULL ) {
        px_print_error( "spmReadDriver: cannot open %s", filename );
        return PASTIX_ERR_FILE;
    }
    if ( fscanf( f, "%lld %lld", &n, &nnz ) != 2 || n <= 0 || nnz < 0 ) {
        fclose( f );
        return PASTIX_ERR_FILE;
    }
    t = px_malloc( (size_t)( nnz > 0 ? nnz : 1 ) * sizeof( px_ijv_t ) );
    for ( long long k = 0; k < nnz; k++ ) {
        long long i, j;
        double    v;
        if ( fscanf( f, "%lld %lld %lf", &i, &j, &v ) != 3 || i < 1 || i > n || j < 1 || j > n ) {
            free( t );
            fclose( f );
            return PASTIX_ERR_FILE;
        }
        t[k].i  = i - 1;
        t[k].j  = j - 1;
        t[k].re = v;
        t[k].im = 0.0;
    }
    fclose( f );
    rc = px_ijv_to_csc( spm, (spm_int_t)n, (spm_int_t)nnz, t, 0, 0, SpmGeneral );
    free( t );
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Laplacian generator string */

static int
px_read_laplacian( const char *desc, spmatrix_t *spm )
{
    char           type = 'd';
    const char    *p    = desc;
    long long      dims[3] = { 1, 1, 1 };
    double         ab[2]   = { 1.0, 1.0 };
    int            nd = 0, na = 0;
    spm_coeftype_t flt;
    if ( isalpha( (unsigned char)p[0] ) && p[1] == ':' ) {
        type = (char)tolower( (unsigned char)p[0] );
        p += 2;
    }
    while ( *p ) {
        char *end;
        if ( nd < 3 ) {
            long long v = strtoll( p, &end, 10 );
            if ( end == p || v < 1 ) {
                return PASTIX_ERR_BADPARAMETER;
            }
            dims[nd++] = v;
        }
        else if ( na < 2 ) {
            double v = strtod( p, &end );
            if ( end == p ) {
                return PASTIX_ERR_BADPARAMETER;
            }
            ab[na++] = v;
        }
        else {
            return PASTIX_ERR_BADPARAMETER;
        }
        p = end;
        if ( *p == ':' ) p++;
        else if ( *p ) return PASTIX_ERR_BADPARAMETER;
    }
    if ( nd == 0 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    {
        int ndim = ( dims[2] > 1 ) ? 3 : ( dims[1] > 1 ) ? 2 : 1;
        flt      = ( type == 'z' || type == 'c' ) ? SpmComplex64 : SpmDouble;
        return spmGenLaplacian7( spm, flt, dims[0], dims[1], dims[2], ab[0] * 2.0 * ndim, ab[1] );
    }
}

int
spmReadDriver( spm_driver_t driver, const char *filename, spmatrix_t *spm )
{
    if ( filename == NULL || spm == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    switch ( driver ) {
    case SpmDriverRSA:
    case SpmDriverHB: return px_read_hb( filename, spm );
    case SpmDriverMM: return px_read_mm( filename, spm );
    case SpmDriverIJV: return px_read_ijv( filename, spm );
    case SpmDriverLaplacian: return px_read_laplacian( filename, spm );
    default:
        px_print_error( "spmReadDriver: driver %d is not available in this build", (int)driver );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
}
