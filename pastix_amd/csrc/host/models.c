/*
 * models.c -- per-kernel cost models (polynomials in the kernel sizes), their
 * model-file format and defaults, the kernel timing log of the device engine
 * written as model.csv, and a least-squares fit of the polynomials to a log.
 *
 * Reference: common/models.h:20-87 and common/models.c:41-593 (forms, file
 * reader, propagation rules, defaults); kernels/kernels_trace.c:222-249
 * (model.csv).  The reference fits its models outside the library; the fit
 * here is the engine's own.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "px_internal.h"
#include "pastix_amd_models.h"

/* ------------------------------------------------------------------------ */
/* Polynomial forms (common/models.h:38-83), clamped at zero. */

double
modelsGetCost1Param( const double *coefs, pastix_int_t N )
{
    const double n = (double)N;
    const double t = coefs[0] + n * ( coefs[1] + n * ( coefs[2] + n * coefs[3] ) );
    return ( t < 0. ) ? 0. : t;
}

double
modelsGetCost2Param( const double *coefs, pastix_int_t M, pastix_int_t N )
{
    const double m = (double)M, n = (double)N;
    const double t = coefs[0] + coefs[1] * n + coefs[2] * m + coefs[3] * n * n + coefs[4] * m * n +
                     coefs[5] * m * n * n;
    return ( t < 0. ) ? 0. : t;
}

double
modelsGetCost3Param( const double *coefs, pastix_int_t M, pastix_int_t N, pastix_int_t K )
{
    const double m = (double)M, n = (double)N, k = (double)K;
    const double t = coefs[0] + coefs[1] * k + coefs[2] * n + coefs[3] * m + coefs[4] * m * n +
                     coefs[5] * k * n + coefs[6] * m * k + coefs[7] * m * n * k;
    return ( t < 0. ) ? 0. : t;
}

/* ------------------------------------------------------------------------ */
/* Kernel names of the model files and their number of coefficients
 * (common/models.c:41-84). */
static const struct {
    const char *name;
    int         ktype;
    int         nbcoef;
} px_model_kernels[] = {
    { "getrf", PastixKernelGETRF, 4 },
    { "hetrf", PastixKernelHETRF, 4 },
    { "potrf", PastixKernelPOTRF, 4 },
    { "pxtrf", PastixKernelPXTRF, 4 },
    { "sytrf", PastixKernelSYTRF, 4 },
    { "trsmcblk1d", PastixKernelTRSMCblk1d, 6 },
    { "trsmcblk2d", PastixKernelTRSMCblk2d, 6 },
    { "trsmcblklr", PastixKernelTRSMCblkLR, 6 },
    { "trsmblok2d", PastixKernelTRSMBlok2d, 6 },
    { "trsmbloklr", PastixKernelTRSMBlokLR, 6 },
    { "gemmcblk1d1d", PastixKernelGEMMCblk1d1d, 8 },
    { "gemmcblk1d2d", PastixKernelGEMMCblk1d2d, 8 },
    { "gemmcblk2d2d", PastixKernelGEMMCblk2d2d, 8 },
    { "gemmcblkfrlr", PastixKernelGEMMCblkFRLR, 8 },
    { "gemmcblklrlr", PastixKernelGEMMCblkLRLR, 8 },
    { "gemmblok2d2d", PastixKernelGEMMBlok2d2d, 8 },
    { "gemmbloklrlr", PastixKernelGEMMBlokLRLR, 8 },
};
#define PX_NMODEL_KERNELS ( (int)( sizeof( px_model_kernels ) / sizeof( px_model_kernels[0] ) ) )

int
modelsGetKernelId( const char *kernelstr, int *nbcoef )
{
    int i;
    for ( i = 0; i < PX_NMODEL_KERNELS; i++ ) {
        if ( strcasecmp( px_model_kernels[i].name, kernelstr ) == 0 ) {
            *nbcoef = px_model_kernels[i].nbcoef;
            return px_model_kernels[i].ktype;
        }
    }
    *nbcoef = 0;
    return -1;
}

/* Name and coefficient count of a kernel type in the files; NULL if none. */
static const char *
px_model_kernel_name( int ktype, int *nbcoef )
{
    int i;
    for ( i = 0; i < PX_NMODEL_KERNELS; i++ ) {
        if ( px_model_kernels[i].ktype == ktype ) {
            *nbcoef = px_model_kernels[i].nbcoef;
            return px_model_kernels[i].name;
        }
    }
    *nbcoef = 0;
    return NULL;
}

/* ------------------------------------------------------------------------ */
/*
 * Propagation (common/models.c:100-201).  The kernel types form families:
 * the diagonal-block kernels GETRF..SYTRF (4 coefficients; GETRF costs twice
 * the others in its cubic and quadratic terms), the TRSMs TRSMCblk1d..
 * TRSMBlok2d and the GEMMs GEMMCblk1d1d..GEMMBlok2d2d (all 8 coefficients
 * copied).  Unset members of the family take the new kernel's coefficients;
 * then, in the other arithmetics, unset members take this arithmetic's
 * coefficients scaled by (a + 1) / (arithm + 1) (s : d : c : z = 1 : 2 : 3 : 4).
 * The scaling kernels have no family.
 */
static int
px_unset( const pastix_model_t *model, int a, int k )
{
    return model->coefficients[a][k][0] == PASTIX_MODEL_UNSET;
}

void
modelsPropagate( pastix_model_t *model, int arithm, pastix_ktype_t kernelid )
{
    const double *src = model->coefficients[arithm][kernelid];
    int first, last, ncoef, a, k, i;

    if ( kernelid <= PastixKernelSYTRF ) {
        first = PastixKernelGETRF;
        last  = PastixKernelSYTRF;
        ncoef = 4;
    }
    else if ( kernelid <= PastixKernelSCALOBlok ) {
        return;
    }
    else if ( kernelid < PastixKernelGEMMCblk1d1d ) {
        first = PastixKernelTRSMCblk1d;
        last  = PastixKernelTRSMBlok2d;
        ncoef = 8;
    }
    else {
        first = PastixKernelGEMMCblk1d1d;
        last  = PastixKernelGEMMBlok2d2d;
        ncoef = 8;
    }

    for ( k = first; k <= last; k++ ) {
        double *dst = model->coefficients[arithm][k];
        if ( k == (int)kernelid || !px_unset( model, arithm, k ) ) {
            continue;
        }
        for ( i = 0; i < ncoef; i++ ) {
            double w = 1.;
            if ( first == PastixKernelGETRF && i >= 2 ) {
                w = ( ( k == PastixKernelGETRF ) ? 2. : 1. ) / ( ( kernelid == PastixKernelGETRF ) ? 2. : 1. );
            }
            dst[i] = w * src[i];
        }
    }
    for ( a = 0; a < 4; a++ ) {
        const double ratio = (double)( a + 1 ) / (double)( arithm + 1 );
        if ( a == arithm ) {
            continue;
        }
        for ( k = first; k <= last; k++ ) {
            if ( !px_unset( model, a, k ) ) {
                continue;
            }
            for ( i = 0; i < ncoef; i++ ) {
                model->coefficients[a][k][i] = ratio * model->coefficients[arithm][k][i];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Model files (common/models.c:219-317). */

/* Next line of f into *buf (grown as needed); -1 at the end. */
static long
px_getline( char **buf, size_t *size, FILE *f )
{
    size_t len = 0;
    int    c;
    if ( *buf == NULL || *size < 2 ) {
        *size = 256;
        *buf  = px_malloc( *size );
    }
    while ( ( c = fgetc( f ) ) != EOF ) {
        if ( len + 2 > *size ) {
            *size *= 2;
            *buf = realloc( *buf, *size );
        }
        ( *buf )[len++] = (char)c;
        if ( c == '\n' ) {
            break;
        }
    }
    ( *buf )[len] = '\0';
    return ( len == 0 && c == EOF ) ? -1 : (long)len;
}

int
modelsRead( pastix_model_t *model, const char *modelfilename )
{
    FILE  *f = fopen( modelfilename, "r" );
    char  *line = NULL, kname[13];
    size_t size = 0;
    int    arith, nbcoef, ktype;

    if ( f == NULL ) {
        fprintf( stderr, "Can't open model file\n" );
        return -1;
    }
    /* the first line that is not a comment names the model */
    do {
        if ( px_getline( &line, &size, f ) < 0 ) {
            fprintf( stderr, "modelsRead: %s has no model name\n", modelfilename );
            free( line );
            fclose( f );
            return -1;
        }
    } while ( line[0] == '#' );
    line[strcspn( line, "\r\n" )] = '\0';
    free( model->name );
    model->name = strdup( line );

    while ( px_getline( &line, &size, f ) >= 0 ) {
        double      c[8];
        const char *p;
        int         i, got;
        if ( line[0] == '#' ) {
            continue;
        }
        if ( sscanf( line, "%d;%12[a-z0-9];", &arith, kname ) != 2 ) {
            fprintf( stderr, "modelRead: %s - Error reading line (%s)\n", model->name, line );
            continue;
        }
        if ( arith < 0 || arith > 3 ) {
            fprintf( stderr, "modelRead: %s - Incorrect arithmetic %d in line:\n\t%s\n", model->name, arith,
                     line );
            continue;
        }
        ktype = modelsGetKernelId( kname, &nbcoef );
        if ( ktype < 0 ) {
            fprintf( stderr, "modelRead: %s - Incorrect kernel type %s in line:\n\t%s\n", model->name, kname,
                     line );
            continue;
        }
        /* the coefficients follow "a;name;" */
        p   = line + 3 + strlen( kname );
        got = 0;
        for ( i = 0; i < nbcoef; i++ ) {
            int used = 0;
            if ( sscanf( p, i ? ";%le%n" : "%le%n", &c[i], &used ) != 1 ) {
                break;
            }
            p += used;
            got++;
        }
        if ( got != nbcoef ) {
            fprintf( stderr, "modelRead: %s - Pb reading the %d coefficients in line:\n\t%s\n", model->name,
                     nbcoef, line );
            continue;
        }
        memcpy( model->coefficients[arith][ktype], c, nbcoef * sizeof( double ) );
        modelsPropagate( model, arith, (pastix_ktype_t)ktype );
    }
    free( line );
    fclose( f );
    return 0;
}

int
pastix_amd_models_write( const pastix_model_t *model, const char *filename )
{
    FILE *f = fopen( filename, "w" );
    int   a, k, i, nbcoef;
    if ( f == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    fprintf( f, "# arith;kernel;coefficients (arith 0..3 = s, d, c, z)\n" );
    fprintf( f, "%s\n", model->name ? model->name : "unnamed" );
    for ( a = 0; a < 4; a++ ) {
        for ( k = 0; k < PastixKernelLvl1Nbr; k++ ) {
            const char *name = px_model_kernel_name( k, &nbcoef );
            if ( name == NULL || model->coefficients[a][k][0] == PASTIX_MODEL_UNSET ) {
                continue;
            }
            fprintf( f, "%d;%s", a, name );
            for ( i = 0; i < nbcoef; i++ ) {
                fprintf( f, ";%.17e", model->coefficients[a][k][i] );
            }
            fprintf( f, "\n" );
        }
    }
    fclose( f );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Defaults (common/models.c:337-474): real double coefficients, propagated. */

static void
px_model_set( pastix_model_t *model, int ktype, int n, const double *c )
{
    memcpy( model->coefficients[1][ktype], c, n * sizeof( double ) );
    modelsPropagate( model, 1, (pastix_ktype_t)ktype );
}

/* GEMM of a panel, shared by both defaults (common/models.c:379-390, 453-464) */
static const double px_gemm_cblk_default[8] = { 1.216278e-06,  0., -2.704179e-10, 1.148989e-07,
                                                2.724804e-10, 1.328900e-09, 0., 2.429169e-10 };

int
modelsInitDefaultCPU( pastix_model_t *model )
{
    static const double potrf[4] = { 4.071507e-07, -1.469893e-07, 1.707006e-08, 2.439599e-11 };
    static const double trsm[6]  = { 3.255168e-06, 3.976198e-08, 0., 0., 0., 2.626177e-10 };
    static const double gemmb[8] = { 0., 0., 0., 0., 0., 0., 0., 2. / 24.e9 };
    free( model->name );
    model->name = strdup( "AMD Opteron 6180 - Intel MKL" );
    px_model_set( model, PastixKernelPOTRF, 4, potrf );
    px_model_set( model, PastixKernelTRSMCblk2d, 6, trsm );
    px_model_set( model, PastixKernelGEMMCblk2d2d, 8, px_gemm_cblk_default );
    px_model_set( model, PastixKernelGEMMBlok2d2d, 8, gemmb );
    return 0;
}

int
modelsInitDefaultGPU( pastix_model_t *model )
{
    static const double trsmb[6] = { -3.16663635648446e-05, 2.63809317549331e-06, 5.86447245256688e-07,
                                     -1.57859559108480e-09, -4.74303242824929e-09, 5.36284121953867e-12 };
    static const double gemmb[8] = { 0., 0., 0., 0., 0., 0., 0., 2. / 1.2e12 };
    free( model->name );
    model->name = strdup( "Nvidia K40 GK1108L - CUDA 8.0" );
    px_model_set( model, PastixKernelTRSMBlok2d, 6, trsmb );
    px_model_set( model, PastixKernelGEMMCblk2d2d, 8, px_gemm_cblk_default );
    px_model_set( model, PastixKernelGEMMBlok2d2d, 8, gemmb );
    return 0;
}

pastix_model_t *
pastixModelsNew( void )
{
    pastix_model_t *model = px_calloc( 1, sizeof( pastix_model_t ) );
    int             a, k;
    for ( a = 0; a < 4; a++ ) {
        for ( k = 0; k < PastixKernelLvl1Nbr; k++ ) {
            model->coefficients[a][k][0] = PASTIX_MODEL_UNSET;
        }
    }
    return model;
}

void
pastixModelsFree( pastix_model_t *model )
{
    if ( model != NULL ) {
        free( model->name );
        free( model );
    }
}

/* One model of an instance: the file named by env, else the defaults. */
static pastix_model_t *
px_models_load_one( const char *env, int ( *defaults )( pastix_model_t * ) )
{
    pastix_model_t *m    = pastixModelsNew();
    const char     *file = getenv( env );
    int             rc   = ( file != NULL && file[0] ) ? modelsRead( m, file ) : defaults( m );
    if ( rc == -1 ) {
        pastixModelsFree( m );
        m = NULL;
    }
    return m;
}

void
pastixModelsLoad( pastix_data_t *pastix_data )
{
    pastix_data->cpu_models = px_models_load_one( "PASTIX_MODELS_CPU", modelsInitDefaultCPU );
    pastix_data->gpu_models = px_models_load_one( "PASTIX_MODELS_GPU", modelsInitDefaultGPU );
}

pastix_model_t *
pastix_amd_models_get( pastix_data_t *pd, int gpu )
{
    return pd ? ( gpu ? pd->gpu_models : pd->cpu_models ) : NULL;
}

double
pastix_amd_model_cost( const pastix_model_t *model, int arith, int ktype, pastix_int_t m, pastix_int_t n,
                       pastix_int_t k )
{
    const double *c;
    if ( model == NULL || arith < 0 || arith > 3 || ktype < 0 || ktype >= PastixKernelLvl1Nbr ) {
        return -1.;
    }
    c = model->coefficients[arith][ktype];
    if ( c[0] == PASTIX_MODEL_UNSET ) {
        return -1.;
    }
    if ( ktype <= PastixKernelSYTRF ) {
        return modelsGetCost1Param( c, m );
    }
    if ( ktype < PastixKernelGEMMCblk1d1d ) {
        return modelsGetCost2Param( c, m, n );
    }
    return modelsGetCost3Param( c, m, n, k );
}

/* ------------------------------------------------------------------------ */
/* Timing log of the device engine. */

int
pastix_amd_model_log( pastix_data_t *pd, pxd_model_entry_t *entries, pastix_int_t max, pastix_int_t *count )
{
    int64_t n = 0;
    int     rc;
    if ( pd == NULL || pd->dev == NULL ) {
        if ( count ) {
            *count = 0;
        }
        return ( pd == NULL ) ? PASTIX_ERR_BADPARAMETER : PASTIX_SUCCESS;
    }
    rc = pxd_model_log( pd->dev, entries, max, &n );
    if ( count ) {
        *count = n;
    }
    return rc;
}

int
pastix_amd_model_write_csv( pastix_data_t *pd, const char *filename )
{
    static const char *prec_names[4] = { "s - single real", "d - double real", "c - single complex",
                                         "z - double complex" };
    pastix_int_t       n = 0, i;
    pxd_model_entry_t *e;
    FILE              *f;
    int                prec;

    if ( pd == NULL || filename == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    pastix_amd_model_log( pd, NULL, 0, &n );
    e = px_malloc( ( n + 1 ) * sizeof( pxd_model_entry_t ) );
    pastix_amd_model_log( pd, e, n, &n );
    f = fopen( filename, "w" );
    if ( f == NULL ) {
        free( e );
        return PASTIX_ERR_BADPARAMETER;
    }
    /* the kernels ran in the precision of the factor */
    prec = ( pd->factored_flttype >= 2 && pd->factored_flttype <= 5 ) ? pd->factored_flttype - 2 : 1;
    fprintf( f, "# GPU Model data\n" );
    fprintf( f, "# Precision: %d - %s\n", prec, prec_names[prec] );
    fprintf( f, "Kernel;M;N;K;Time\n" );
    for ( i = 0; i < n; i++ ) {
        fprintf( f, "%d;%d;%d;%d;%e\n", e[i].ktype, e[i].m, e[i].n, e[i].k, e[i].time );
    }
    fclose( f );
    free( e );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/*
 * Fit: for each kernel type in the records, the coefficients c minimise
 * sum_i ((x_i . c - t_i) / t_i)^2 over the records of that type, where x_i
 * holds the monomials of the type's form.  The rows are folded one at a time
 * into a triangular factor by Givens rotations (a streaming QR), with the
 * columns scaled by their largest monomial so that the factor is well
 * conditioned.
 */
static int
px_model_monomials( int ktype, const pxd_model_entry_t *e, double *x )
{
    const double m = e->m, n = e->n, k = e->k;
    if ( ktype <= PastixKernelSYTRF ) {
        x[0] = 1.; x[1] = m; x[2] = m * m; x[3] = m * m * m;
        return 4;
    }
    if ( ktype < PastixKernelGEMMCblk1d1d ) {
        x[0] = 1.; x[1] = n; x[2] = m; x[3] = n * n; x[4] = m * n; x[5] = m * n * n;
        return 6;
    }
    x[0] = 1.; x[1] = k; x[2] = n; x[3] = m; x[4] = m * n; x[5] = k * n; x[6] = m * k; x[7] = m * n * k;
    return 8;
}

int
pastix_amd_models_fit( pastix_model_t *model, const char *name, int arith, const pxd_model_entry_t *entries,
                       pastix_int_t count )
{
    int kt;
    if ( model == NULL || arith < 0 || arith > 3 || ( count > 0 && entries == NULL ) ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( name ) {
        free( model->name );
        model->name = strdup( name );
    }
    for ( kt = 0; kt < PastixKernelLvl1Nbr; kt++ ) {
        double       R[8][9], scale[8], x[8], c[8];
        int          p = 0, i, j, nrows = 0;
        pastix_int_t r;
        memset( R, 0, sizeof( R ) );
        for ( i = 0; i < 8; i++ ) {
            scale[i] = 0.;
        }
        for ( r = 0; r < count; r++ ) {
            if ( entries[r].ktype != kt || !( entries[r].time > 0. ) ) {
                continue;
            }
            p = px_model_monomials( kt, &entries[r], x );
            for ( i = 0; i < p; i++ ) {
                scale[i] = fmax( scale[i], fabs( x[i] ) );
            }
            nrows++;
        }
        if ( nrows < p || p == 0 ) {
            continue;
        }
        for ( i = 0; i < p; i++ ) {
            scale[i] = ( scale[i] > 0. ) ? scale[i] : 1.;
        }
        for ( r = 0; r < count; r++ ) {
            double row[9];
            if ( entries[r].ktype != kt || !( entries[r].time > 0. ) ) {
                continue;
            }
            px_model_monomials( kt, &entries[r], x );
            for ( i = 0; i < p; i++ ) {
                row[i] = x[i] / scale[i] / entries[r].time;
            }
            row[p] = 1.;   /* t_i / t_i */
            for ( i = 0; i < p; i++ ) {
                double a = R[i][i], b = row[i], h, cs, sn;
                if ( b == 0. ) {
                    continue;
                }
                h  = hypot( a, b );
                cs = a / h;
                sn = b / h;
                for ( j = i; j <= p; j++ ) {
                    const double u = R[i][j], v = row[j];
                    R[i][j] = cs * u + sn * v;
                    row[j]  = -sn * u + cs * v;
                }
            }
        }
        /* back substitution; a column without information keeps 0 */
        for ( i = p - 1; i >= 0; i-- ) {
            double s = R[i][p];
            for ( j = i + 1; j < p; j++ ) {
                s -= R[i][j] * c[j];
            }
            c[i] = ( fabs( R[i][i] ) > 1e-12 * fabs( R[0][0] ) ) ? s / R[i][i] : 0.;
        }
        for ( i = 0; i < 8; i++ ) {
            model->coefficients[arith][kt][i] = ( i < p ) ? c[i] / scale[i] : 0.;
        }
        modelsPropagate( model, arith, (pastix_ktype_t)kt );
    }
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Learned model: standardised features and log times around the GPU MLP. */

#define PX_LM_F 10

struct pastix_amd_learned_s {
    pxd_mlp_t *mlp;
    double     mu[PX_LM_F], sd[PX_LM_F];   /* feature standardisation  */
    double     tmu, tsd;                    /* log-time standardisation */
};

static void
px_lm_raw_features( int arith, const pxd_model_entry_t *e, double *f )
{
    const double m = e->m, n = e->n, k = e->k;
    const int    upd = ( e->ktype >= PastixKernelGEMMCblk1d1d );
    double       flops;
    f[0] = ( e->ktype <= PastixKernelSYTRF ) ? 1. : 0.;
    f[1] = ( e->ktype > PastixKernelSCALOBlok && e->ktype < PastixKernelGEMMCblk1d1d ) ? 1. : 0.;
    /* updates: the kernel class that runs them (an unknown class counts as
     * the 64 x 64 tiles) */
    f[2] = ( upd && e->variant != 1 && e->variant != 2 ) ? 1. : 0.;
    f[3] = ( upd && e->variant == 1 ) ? 1. : 0.;
    f[4] = ( upd && e->variant == 2 ) ? 1. : 0.;
    if ( f[0] != 0. ) {
        flops = m * m * m / 3.;
    }
    else if ( f[1] != 0. ) {
        flops = m * n * n;
    }
    else {
        flops = 2. * m * n * k;
    }
    f[5] = log1p( m );
    f[6] = log1p( n );
    f[7] = log1p( k );
    f[8] = log1p( flops );
    f[9] = arith;
}

static float *
px_lm_features( const pastix_amd_learned_t *lm, int arith, const pxd_model_entry_t *e, pastix_int_t count )
{
    float       *X = px_malloc( ( count + 1 ) * PX_LM_F * sizeof( float ) );
    pastix_int_t r;
    int          i;
    for ( r = 0; r < count; r++ ) {
        double f[PX_LM_F];
        px_lm_raw_features( arith, &e[r], f );
        for ( i = 0; i < PX_LM_F; i++ ) {
            X[r * PX_LM_F + i] = (float)( ( f[i] - lm->mu[i] ) / lm->sd[i] );
        }
    }
    return X;
}

int
pastix_amd_learned_train( pastix_amd_learned_t **plm, int device, int arith, const pxd_model_entry_t *entries,
                          pastix_int_t count, int steps, double lr, uint64_t seed, double *loss )
{
    pastix_amd_learned_t *lm;
    pxd_model_entry_t    *e;
    float                *X, *t;
    pastix_int_t          r, n = 0;
    int                   i, q, rc;
    double                s2[PX_LM_F] = { 0 }, ts2 = 0.;

    if ( plm == NULL || entries == NULL || count <= 0 || steps < 0 || arith < 0 || arith > 3 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    /* the records with a measured time */
    e = px_malloc( count * sizeof( pxd_model_entry_t ) );
    for ( r = 0; r < count; r++ ) {
        if ( entries[r].time > 0. ) {
            e[n++] = entries[r];
        }
    }
    if ( n == 0 ) {
        free( e );
        return PASTIX_ERR_BADPARAMETER;
    }
    lm = px_calloc( 1, sizeof( pastix_amd_learned_t ) );
    for ( r = 0; r < n; r++ ) {
        double f[PX_LM_F], lt = log( e[r].time );
        px_lm_raw_features( arith, &e[r], f );
        for ( i = 0; i < PX_LM_F; i++ ) {
            lm->mu[i] += f[i];
            s2[i] += f[i] * f[i];
        }
        lm->tmu += lt;
        ts2 += lt * lt;
    }
    for ( i = 0; i < PX_LM_F; i++ ) {
        lm->mu[i] /= n;
        lm->sd[i] = sqrt( fmax( 0., s2[i] / n - lm->mu[i] * lm->mu[i] ) );
        lm->sd[i] = ( lm->sd[i] > 1e-12 ) ? lm->sd[i] : 1.;
    }
    lm->tmu /= n;
    lm->tsd = sqrt( fmax( 0., ts2 / n - lm->tmu * lm->tmu ) );
    lm->tsd = ( lm->tsd > 1e-12 ) ? lm->tsd : 1.;

    X = px_lm_features( lm, arith, e, n );
    t = px_malloc( n * sizeof( float ) );
    for ( r = 0; r < n; r++ ) {
        t[r] = (float)( ( log( e[r].time ) - lm->tmu ) / lm->tsd );
    }
    rc = pxd_mlp_create( &lm->mlp, device, seed );
    if ( rc == 0 ) {
        rc = pxd_mlp_set_data( lm->mlp, X, t, n );
    }
    for ( q = 0; q < 4 && rc == 0; q++ ) {
        const int s = ( steps * ( q + 1 ) ) / 4 - ( steps * q ) / 4;
        rc = pxd_mlp_train( lm->mlp, s, lr, ( q == 3 ) ? loss : NULL );
        lr /= 3.;
    }
    free( X );
    free( t );
    free( e );
    if ( rc != 0 ) {
        pastix_amd_learned_free( lm );
        return rc;
    }
    *plm = lm;
    return PASTIX_SUCCESS;
}

int
pastix_amd_learned_predict( pastix_amd_learned_t *lm, int arith, const pxd_model_entry_t *entries,
                            pastix_int_t count, double *time )
{
    float       *X, *y;
    pastix_int_t r;
    int          rc;
    if ( lm == NULL || count < 0 || ( count > 0 && ( entries == NULL || time == NULL ) ) ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( count == 0 ) {
        return PASTIX_SUCCESS;
    }
    X  = px_lm_features( lm, arith, entries, count );
    y  = px_malloc( count * sizeof( float ) );
    rc = pxd_mlp_predict( lm->mlp, X, count, y );
    for ( r = 0; r < count && rc == 0; r++ ) {
        time[r] = exp( lm->tmu + lm->tsd * (double)y[r] );
    }
    free( X );
    free( y );
    return rc;
}

/*
 * Kernel class of every blok update chosen by predicted cost: the learned
 * model when lm is given, else the polynomial models poly[0..2] of the three
 * classes (64 x 64 tiles, tall tiles, one wavefront); neither: the engine's
 * cut-offs again.  The models predict the time of an update alone on the
 * device; the updates of a level run side by side, so the cost compared is
 * the share of the device an update holds over that time: the predicted
 * time times the fraction of the CUs its workgroups fill (64 x 64 tiles:
 * one workgroup per tile; tall tiles: one per PX_TALL_ROWS(arith) x 64
 * tile; one wavefront: a quarter of a workgroup per 64 rows).  The updates and their sizes are those the engine
 * schedules (kernels/core_zgemmsp.c:401-472 sizes: M the stacked rows from
 * the facing blok on, less the blok itself on an LU factor's U side, N the
 * blok's rows, K the cblk's width); the one-wavefront class is offered only
 * where it fits (M <= 256, N <= 64, K <= 32).  *predicted, if given, gets the
 * predicted total update time of the choice.  Needs the structure on the
 * device (after pastix_subtask_bcsc2ctab); takes effect at the next
 * factorization.
 */
#define PX_DEVICE_CUS 256.   /* MI355X */
#define PX_TALL_ROWS( arith ) ( ( arith ) == 0 ? 256 : ( arith ) == 3 ? 64 : 128 )

int
pastix_amd_choose_update_kernels( pastix_data_t *pd, pastix_amd_learned_t *lm, pastix_model_t *const *poly,
                                  int arith, double *predicted )
{
    const px_solver_t *s;
    pxd_model_entry_t *rec;
    double            *t;
    int8_t            *cls;
    pastix_int_t       nrec = 0, r, nb;
    int                lu, rc = PASTIX_SUCCESS;

    if ( pd == NULL || pd->solvmatr == NULL || pd->dev == NULL || arith < 0 || arith > 3 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( lm == NULL && poly == NULL ) {
        if ( predicted ) {
            *predicted = 0.;
        }
        return pxd_set_update_classes( pd->dev, NULL, 0 );
    }
    s   = pd->solvmatr;
    nb  = s->bloknbr;
    lu  = ( pd->iparm[IPARM_FACTORIZATION] == PastixFactGETRF || pd->iparm[IPARM_FACTORIZATION] == PastixFactLU );
    rec = px_malloc( ( 6 * nb + 1 ) * sizeof( pxd_model_entry_t ) );
    /* three candidate records per (blok, side); m = 0 marks a missing one */
    for ( pastix_int_t k = 0; k < s->cblknbr; k++ ) {
        const px_cblk_t   *c = s->cblktab + k;
        const pastix_int_t K = c->lcolnum - c->fcolnum + 1;
        for ( pastix_int_t b = c->fblok; b < c[1].fblok; b++ ) {
            const px_blok_t   *bb = s->bloktab + b;
            const pastix_int_t N  = bb->lrownum - bb->frownum + 1;
            for ( int side = 0; side < 2; side++ ) {
                pxd_model_entry_t *e = rec + 3 * ( 2 * b + side );
                const pastix_int_t M = c->stride - bb->coefind / K - ( side ? N : 0 );
                const int          ok = ( b > c->fblok ) && ( side == 0 || lu ) && M > 0;
                for ( int v = 0; v < 3; v++ ) {
                    const int fits = ok && ( v < 2 || ( M <= 256 && N <= 64 && K <= 32 ) );
                    e[v] = (pxd_model_entry_t){ PastixKernelGEMMCblk2d2d, fits ? (int32_t)M : 0, (int32_t)N,
                                                (int32_t)K, 0., v, 0 };
                }
            }
        }
    }
    nrec = 6 * nb;
    t    = px_calloc( nrec + 1, sizeof( double ) );
    if ( lm != NULL ) {
        rc = pastix_amd_learned_predict( lm, arith, rec, nrec, t );
    }
    else {
        for ( r = 0; r < nrec; r++ ) {
            t[r] = pastix_amd_model_cost( poly[rec[r].variant], arith, PastixKernelGEMMCblk2d2d, rec[r].m, rec[r].n,
                                          rec[r].k );
        }
    }
    /* device share: time x the fraction of the CUs the workgroups fill */
    for ( r = 0; r < nrec; r++ ) {
        const double M = rec[r].m, N = rec[r].n;
        double       wg;
        if ( rec[r].variant == 0 ) {
            wg = ceil( M / 64. ) * ceil( N / 64. );
        }
        else if ( rec[r].variant == 1 ) {
            wg = ceil( M / PX_TALL_ROWS( arith ) ) * ceil( N / 64. );
        }
        else {
            wg = 0.25 * ceil( M / 64. );   /* one wavefront per 64 rows */
        }
        t[r] *= fmin( 1., wg / PX_DEVICE_CUS );
    }
    cls = px_malloc( 2 * nb + 1 );
    if ( predicted ) {
        *predicted = 0.;
    }
    for ( r = 0; r < 2 * nb && rc == PASTIX_SUCCESS; r++ ) {
        int best = -1;
        for ( int v = 0; v < 3; v++ ) {
            if ( rec[3 * r + v].m > 0 && ( best < 0 || t[3 * r + v] < t[3 * r + best] ) ) {
                best = v;
            }
        }
        cls[r] = (int8_t)best;
        if ( predicted && best >= 0 ) {
            *predicted += t[3 * r + best];
        }
    }
    if ( rc == PASTIX_SUCCESS ) {
        rc = pxd_set_update_classes( pd->dev, cls, 2 * nb );
    }
    free( cls );
    free( t );
    free( rec );
    return rc;
}

void
pastix_amd_learned_free( pastix_amd_learned_t *lm )
{
    if ( lm != NULL ) {
        pxd_mlp_destroy( lm->mlp );
        free( lm );
    }
}
