/*
 * spm.c -- the SPM routines the PaStiX path calls (norms, matvec, RHS
 * generation, backward-error check) and the grid problem generators.
 *
 * Only double and complex double values are produced or consumed: the user
 * matrix is the FP64 input of every configuration; single precision exists
 * only inside the factor.
 */
#include <complex.h>
#include <math.h>
#include "px_internal.h"

size_t
spm_size_of( spm_coeftype_t flttype )
{
    switch ( flttype ) {
    case SpmFloat:     return 4;
    case SpmDouble:    return 8;
    case SpmComplex32: return 8;
    case SpmComplex64: return 16;
    default:           return 0;
    }
}

void
spmInit( spmatrix_t *spm )
{
    memset( spm, 0, sizeof( *spm ) );
    spm->mtxtype  = SpmGeneral;
    spm->flttype  = SpmDouble;
    spm->fmttype  = SpmCSC;
    spm->dof      = 1;
    spm->layout   = PastixColMajor;
    spm->clustnbr = 1;
}

void
spmExit( spmatrix_t *spm )
{
    free( spm->colptr );
    free( spm->rowptr );
    free( spm->values );
    free( spm->dofs );
    free( spm->loc2glob );
    spm->colptr = spm->rowptr = NULL;
    spm->values = NULL;
    spm->dofs = spm->loc2glob = NULL;
}

void
spmUpdateComputedFields( spmatrix_t *spm )
{
    spm_int_t dof = spm->dof > 0 ? spm->dof : 1;
    spm->gN      = spm->n;
    spm->gnnz    = spm->nnz;
    spm->nexp    = spm->n * dof;
    spm->gNexp   = spm->nexp;
    spm->nnzexp  = spm->nnz * dof * dof;
    spm->gnnzexp = spm->nnzexp;
}

/* ------------------------------------------------------------------------ */
/* plrnt: 64-bit LCG with O(log n) jump, single-precision output in (-0.5,0.5]
 * (reference kernels/core_zplrnt.c:24-44,91-120). */
#define PX_RND_A 6364136223846793005ULL
#define PX_RND_C 1ULL

static unsigned long long
px_rnd_jump( unsigned long long n, unsigned long long seed )
{
    unsigned long long a_k = PX_RND_A, c_k = PX_RND_C, ran = seed;
    for ( ; n; n >>= 1 ) {
        if ( n & 1 ) {
            ran = a_k * ran + c_k;
        }
        c_k *= ( a_k + 1 );
        a_k *= a_k;
    }
    return ran;
}

void
spm_plrnt( spm_int_t n, unsigned long long seed, unsigned long long shift, double *out )
{
    unsigned long long ran = px_rnd_jump( shift, seed );
    for ( spm_int_t i = 0; i < n; i++ ) {
        float v = 0.5f - (float)ran * 5.4210108624275222e-20f;
        out[i]  = (double)v;
        ran     = PX_RND_A * ran + PX_RND_C;
    }
}

/* ------------------------------------------------------------------------ */
/* Expand the stored pattern into the full matrix (both triangles) and call
 * fn(row, col, value) for every entry. */
typedef void ( *px_entry_fn )( void *arg, spm_int_t i, spm_int_t j, double complex v );

static void
spm_foreach_full( const spmatrix_t *spm, px_entry_fn fn, void *arg )
{
    spm_int_t base = spm->baseval;
    int       cplx = ( spm->flttype == SpmComplex64 );
    const double         *dv = (const double *)spm->values;
    const double complex *zv = (const double complex *)spm->values;

    for ( spm_int_t j = 0; j < spm->n; j++ ) {
        for ( spm_int_t k = spm->colptr[j] - base; k < spm->colptr[j + 1] - base; k++ ) {
            spm_int_t      i = spm->rowptr[k] - base;
            double complex v = cplx ? zv[k] : (double complex)dv[k];
            if ( spm->fmttype == SpmCSR ) {
                fn( arg, j, i, v );
                if ( spm->mtxtype != SpmGeneral && i != j ) {
                    fn( arg, i, j, spm->mtxtype == SpmHermitian ? conj( v ) : v );
                }
            }
            else {
                fn( arg, i, j, v );
                if ( spm->mtxtype != SpmGeneral && i != j ) {
                    fn( arg, j, i, spm->mtxtype == SpmHermitian ? conj( v ) : v );
                }
            }
        }
    }
}

struct norm_arg {
    double *rows;
    double *cols;
    double  frob_scale, frob_ssq, maxv;
};

static void
norm_entry( void *varg, spm_int_t i, spm_int_t j, double complex v )
{
    struct norm_arg *a = (struct norm_arg *)varg;
    double           av = cabs( v );
    a->rows[i] += av;
    a->cols[j] += av;
    if ( av > a->maxv ) {
        a->maxv = av;
    }
    /* scaled sum of squares (LAPACK lassq) */
    if ( av > 0. ) {
        if ( a->frob_scale < av ) {
            a->frob_ssq   = 1. + a->frob_ssq * ( a->frob_scale / av ) * ( a->frob_scale / av );
            a->frob_scale = av;
        }
        else {
            a->frob_ssq += ( av / a->frob_scale ) * ( av / a->frob_scale );
        }
    }
}

double
spmNorm( spm_normtype_t ntype, const spmatrix_t *spm )
{
    struct norm_arg a;
    double          r = 0.;
    a.rows       = px_calloc( spm->n, sizeof( double ) );
    a.cols       = px_calloc( spm->n, sizeof( double ) );
    a.frob_scale = 0.;
    a.frob_ssq   = 1.;
    a.maxv       = 0.;
    spm_foreach_full( spm, norm_entry, &a );
    switch ( ntype ) {
    case SpmOneNorm:
        for ( spm_int_t j = 0; j < spm->n; j++ ) {
            r = fmax( r, a.cols[j] );
        }
        break;
    case SpmInfNorm:
        for ( spm_int_t i = 0; i < spm->n; i++ ) {
            r = fmax( r, a.rows[i] );
        }
        break;
    case SpmMaxNorm:
        r = a.maxv;
        break;
    case SpmFrobeniusNorm:
    default:
        r = a.frob_scale * sqrt( a.frob_ssq );
        break;
    }
    free( a.rows );
    free( a.cols );
    return r;
}

struct mv_arg {
    int                   cplx;
    int                   trans;
    spm_int_t             nrhs;
    const double         *dx;
    const double complex *zx;
    spm_int_t             ldx;
    double               *dy;
    double complex       *zy;
    spm_int_t             ldy;
    double                alpha;
};

static void
mv_entry( void *varg, spm_int_t i, spm_int_t j, double complex v )
{
    struct mv_arg *a = (struct mv_arg *)varg;
    if ( a->trans != PastixNoTrans ) {
        spm_int_t t = i;
        i = j;
        j = t;
        if ( a->trans == PastixConjTrans ) {
            v = conj( v );
        }
    }
    for ( spm_int_t c = 0; c < a->nrhs; c++ ) {
        if ( a->cplx ) {
            a->zy[i + c * a->ldy] += a->alpha * v * a->zx[j + c * a->ldx];
        }
        else {
            a->dy[i + c * a->ldy] += a->alpha * creal( v ) * a->dx[j + c * a->ldx];
        }
    }
}

int
spmMatMat( int trans, spm_int_t n, double alpha, const spmatrix_t *spm,
           const void *B, spm_int_t ldb, double beta, void *C, spm_int_t ldc )
{
    struct mv_arg a;
    spm_int_t     N = spm->n;
    a.cplx  = ( spm->flttype == SpmComplex64 );
    a.trans = trans;
    a.nrhs  = n;
    a.dx    = (const double *)B;
    a.zx    = (const double complex *)B;
    a.ldx   = ldb;
    a.dy    = (double *)C;
    a.zy    = (double complex *)C;
    a.ldy   = ldc;
    a.alpha = alpha;
    for ( spm_int_t c = 0; c < n; c++ ) {
        for ( spm_int_t i = 0; i < N; i++ ) {
            if ( a.cplx ) {
                a.zy[i + c * ldc] = ( beta == 0. ) ? 0. : beta * a.zy[i + c * ldc];
            }
            else {
                a.dy[i + c * ldc] = ( beta == 0. ) ? 0. : beta * a.dy[i + c * ldc];
            }
        }
    }
    spm_foreach_full( spm, mv_entry, &a );
    return PASTIX_SUCCESS;
}

int
spmMatVec( int trans, double alpha, const spmatrix_t *spm,
           const void *x, double beta, void *y )
{
    return spmMatMat( trans, 1, alpha, spm, x, spm->n, beta, y, spm->n );
}

int
spmGenRHSSeed( spm_rhstype_t type, spm_int_t nrhs, const spmatrix_t *spm,
               unsigned long long seed, void *x, spm_int_t ldx, void *b, spm_int_t ldb )
{
    spm_int_t n    = spm->n;
    int       cplx = ( spm->flttype == SpmComplex64 );
    void     *xx   = x;
    double   *tmp;

    if ( type == SpmRhsRndB ) {
        for ( spm_int_t c = 0; c < nrhs; c++ ) {
            tmp = px_malloc( n * ( cplx ? 2 : 1 ) * sizeof( double ) );
            spm_plrnt( n * ( cplx ? 2 : 1 ), seed + c, 0, tmp );
            for ( spm_int_t i = 0; i < n; i++ ) {
                if ( cplx ) {
                    ( (double complex *)b )[i + c * ldb] = 2. * tmp[2 * i] + 2. * I * tmp[2 * i + 1];
                }
                else {
                    ( (double *)b )[i + c * ldb] = 2. * tmp[i];
                }
            }
            free( tmp );
        }
        return PASTIX_SUCCESS;
    }

    if ( xx == NULL ) {
        xx  = px_malloc( n * nrhs * ( cplx ? 16 : 8 ) );
        ldx = n;
    }
    for ( spm_int_t c = 0; c < nrhs; c++ ) {
        if ( type == SpmRhsOne || type == SpmRhsI ) {
            for ( spm_int_t i = 0; i < n; i++ ) {
                double v = ( type == SpmRhsOne ) ? 1. : (double)( i + 1 );
                if ( cplx ) {
                    ( (double complex *)xx )[i + c * ldx] = v;
                }
                else {
                    ( (double *)xx )[i + c * ldx] = v;
                }
            }
        }
        else {
            /* x uniform in [-1, 1): twice the plrnt value, real and imaginary
             * parts drawn alternately as core_zplrnt does */
            tmp = px_malloc( n * ( cplx ? 2 : 1 ) * sizeof( double ) );
            spm_plrnt( n * ( cplx ? 2 : 1 ), seed + c, 0, tmp );
            for ( spm_int_t i = 0; i < n; i++ ) {
                if ( cplx ) {
                    ( (double complex *)xx )[i + c * ldx] = 2. * tmp[2 * i] + 2. * I * tmp[2 * i + 1];
                }
                else {
                    ( (double *)xx )[i + c * ldx] = 2. * tmp[i];
                }
            }
            free( tmp );
        }
    }
    spmMatMat( PastixNoTrans, nrhs, 1., spm, xx, ldx, 0., b, ldb );
    if ( xx != x ) {
        free( xx );
    }
    return PASTIX_SUCCESS;
}

int
spmGenRHS( spm_rhstype_t type, spm_int_t nrhs, const spmatrix_t *spm,
           void *x, spm_int_t ldx, void *b, spm_int_t ldb )
{
    return spmGenRHSSeed( type, nrhs, spm, 0, x, ldx, b, ldb );
}

/* max that keeps a NaN of either operand (fmax drops it) */
static inline double
max_nan( double a, double b )
{
    return ( a > b || a != a ) ? a : b;
}

double
spmBackwardError( spm_int_t nrhs, const spmatrix_t *spm, const void *b, spm_int_t ldb,
                  const void *x, spm_int_t ldx, double *berr_each )
{
    spm_int_t n    = spm->n;
    int       cplx = ( spm->flttype == SpmComplex64 );
    double    normA = spmNorm( SpmInfNorm, spm );
    double    worst = 0.;
    void     *r    = px_malloc( n * nrhs * ( cplx ? 16 : 8 ) );

    for ( spm_int_t c = 0; c < nrhs; c++ ) {
        for ( spm_int_t i = 0; i < n; i++ ) {
            if ( cplx ) {
                ( (double complex *)r )[i + c * n] = ( (const double complex *)b )[i + c * ldb];
            }
            else {
                ( (double *)r )[i + c * n] = ( (const double *)b )[i + c * ldb];
            }
        }
    }
    spmMatMat( PastixNoTrans, nrhs, -1., spm, x, ldx, 1., r, n );
    for ( spm_int_t c = 0; c < nrhs; c++ ) {
        double nr = 0., nx = 0., nb = 0.;
        for ( spm_int_t i = 0; i < n; i++ ) {
            if ( cplx ) {
                nr = max_nan( nr, cabs( ( (double complex *)r )[i + c * n] ) );
                nx = max_nan( nx, cabs( ( (const double complex *)x )[i + c * ldx] ) );
                nb = max_nan( nb, cabs( ( (const double complex *)b )[i + c * ldb] ) );
            }
            else {
                nr = max_nan( nr, fabs( ( (double *)r )[i + c * n] ) );
                nx = max_nan( nx, fabs( ( (const double *)x )[i + c * ldx] ) );
                nb = max_nan( nb, fabs( ( (const double *)b )[i + c * ldb] ) );
            }
        }
        double be = nr / ( normA * nx + nb );
        if ( berr_each ) {
            berr_each[c] = be;
        }
        worst = max_nan( worst, be );
    }
    free( r );
    return worst;
}

/* max over the columns of ||x - x0||_inf / ||x0||_inf (NaN kept) */
double
spmForwardError( spm_int_t nrhs, const spmatrix_t *spm, const void *x0, spm_int_t ldx0, const void *x,
                 spm_int_t ldx )
{
    spm_int_t n     = spm->n;
    int       cplx  = ( spm->flttype == SpmComplex64 );
    double    worst = 0.;
    for ( spm_int_t c = 0; c < nrhs; c++ ) {
        double nd = 0., n0 = 0.;
        for ( spm_int_t i = 0; i < n; i++ ) {
            if ( cplx ) {
                double complex a = ( (const double complex *)x0 )[i + c * ldx0];
                double complex v = ( (const double complex *)x )[i + c * ldx];
                nd = max_nan( nd, cabs( v - a ) );
                n0 = max_nan( n0, cabs( a ) );
            }
            else {
                double a = ( (const double *)x0 )[i + c * ldx0];
                double v = ( (const double *)x )[i + c * ldx];
                nd = max_nan( nd, fabs( v - a ) );
                n0 = max_nan( n0, fabs( a ) );
            }
        }
        worst = max_nan( worst, ( n0 > 0. ) ? nd / n0 : nd );
    }
    return worst;
}

/*
 * 0 when the backward error of every column is at most eps and, with x0
 * given, the forward error ||x - x0||_inf / ||x0||_inf is at most
 * PX_CHECK_FWD_FACTOR * eps (a forward error may exceed the backward error
 * by the condition number, hence the margin); 1 otherwise.  A NaN fails.
 * SPM's own check lives outside the reference tree, so its exact
 * tolerances are not reproduced here.
 */
#define PX_CHECK_FWD_FACTOR 1e4
int
spmCheckAxb( double eps, spm_int_t nrhs, const spmatrix_t *spm,
             const void *x0, spm_int_t ldx0, const void *b, spm_int_t ldb,
             const void *x, spm_int_t ldx )
{
    double berr;
    int    fail;
    if ( eps < 0. ) {
        eps = 1e-12;
    }
    berr = spmBackwardError( nrhs, spm, b, ldb, x, ldx, NULL );
    fail = !( berr <= eps );
    if ( x0 != NULL ) {
        double ferr = spmForwardError( nrhs, spm, x0, ldx0, x, ldx );
        fail        = fail || !( ferr <= PX_CHECK_FWD_FACTOR * eps );
    }
    return fail;
}

/* ------------------------------------------------------------------------ */
/* Grid generators.  Node (i,j,k) has number i + nx*j + nx*ny*k. */

static spm_int_t
gidx( spm_int_t i, spm_int_t j, spm_int_t k, spm_int_t nx, spm_int_t ny )
{
    return i + nx * j + nx * ny * k;
}

/* Lower-triangular CSC builder from per-column neighbour lists. */
int
spmGenLaplacian7( spmatrix_t *spm, spm_coeftype_t flttype, spm_int_t nx, spm_int_t ny,
                  spm_int_t nz, double diag, double offd )
{
    spm_int_t n = nx * ny * nz;
    spm_int_t nnz, pos = 0;
    int       cplx = ( flttype == SpmComplex64 );

    spmInit( spm );
    spm->mtxtype = SpmSymmetric;
    spm->flttype = cplx ? SpmComplex64 : SpmDouble;
    spm->n       = n;
    /* lower triangle: diagonal + the +x, +y, +z neighbours */
    nnz = n + ( nx - 1 ) * ny * nz + nx * ( ny - 1 ) * nz + nx * ny * ( nz - 1 );
    spm->nnz    = nnz;
    spm->colptr = px_malloc( ( n + 1 ) * sizeof( spm_int_t ) );
    spm->rowptr = px_malloc( nnz * sizeof( spm_int_t ) );
    spm->values = px_malloc( nnz * ( cplx ? 16 : 8 ) );
    for ( spm_int_t k = 0; k < nz; k++ ) {
        for ( spm_int_t j = 0; j < ny; j++ ) {
            for ( spm_int_t i = 0; i < nx; i++ ) {
                spm_int_t col = gidx( i, j, k, nx, ny );
                spm_int_t nb[4];
                double    vv[4];
                int       cnt = 0;
                spm->colptr[col] = pos;
                nb[cnt] = col; vv[cnt++] = diag;
                if ( i + 1 < nx ) { nb[cnt] = gidx( i + 1, j, k, nx, ny ); vv[cnt++] = -offd; }
                if ( j + 1 < ny ) { nb[cnt] = gidx( i, j + 1, k, nx, ny ); vv[cnt++] = -offd; }
                if ( k + 1 < nz ) { nb[cnt] = gidx( i, j, k + 1, nx, ny ); vv[cnt++] = -offd; }
                for ( int t = 0; t < cnt; t++ ) {
                    spm->rowptr[pos] = nb[t];
                    if ( cplx ) {
                        ( (double complex *)spm->values )[pos] = vv[t];
                    }
                    else {
                        ( (double *)spm->values )[pos] = vv[t];
                    }
                    pos++;
                }
            }
        }
    }
    spm->colptr[n] = pos;
    spmUpdateComputedFields( spm );
    return PASTIX_SUCCESS;
}

int
spmGenHelmholtz7( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz, double kr, double ki )
{
    spmGenLaplacian7( spm, SpmComplex64, nx, ny, nz, 6., 1. );
    /* subtract (kr + i ki) on the diagonal: the diagonal is the first entry of each column */
    for ( spm_int_t j = 0; j < spm->n; j++ ) {
        ( (double complex *)spm->values )[spm->colptr[j]] -= kr + I * ki;
    }
    return PASTIX_SUCCESS;
}

/*
 * 27-point convection-diffusion, general storage (full pattern, CSC):
 *   a(p,p) = 26 d,  a(p,q) = -d for the 26 neighbours q of p,
 *   plus central-difference convection on the 6 face neighbours:
 *   a(p, p +/- e_c) += +/- 0.5 * v_c(p), v(p) uniform in [-vmax, vmax]^3 per
 *   cell (plrnt stream seeded by `seed`, 3 draws per cell in cell order).
 */
int
spmGenConvDiff27( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz,
                  double diffusion, double vmax, unsigned long long seed )
{
    spm_int_t n = nx * ny * nz;
    spm_int_t pos = 0;
    double   *vel = px_malloc( 3 * n * sizeof( double ) );

    spm_plrnt( 3 * n, seed, 0, vel );
    for ( spm_int_t p = 0; p < 3 * n; p++ ) {
        vel[p] *= 2. * vmax;
    }
    spmInit( spm );
    spm->mtxtype = SpmGeneral;
    spm->flttype = SpmDouble;
    spm->n       = n;
    spm->nnz     = ( 3 * nx - 2 ) * ( 3 * ny - 2 ) * ( 3 * nz - 2 );
    spm->colptr  = px_malloc( ( n + 1 ) * sizeof( spm_int_t ) );
    spm->rowptr  = px_malloc( spm->nnz * sizeof( spm_int_t ) );
    spm->values  = px_malloc( spm->nnz * sizeof( double ) );
    /* Column q holds a(p,q) for the neighbours p of q: build by columns. */
    for ( spm_int_t k = 0; k < nz; k++ ) {
        for ( spm_int_t j = 0; j < ny; j++ ) {
            for ( spm_int_t i = 0; i < nx; i++ ) {
                spm_int_t q = gidx( i, j, k, nx, ny );
                spm->colptr[q] = pos;
                for ( int dk = -1; dk <= 1; dk++ ) {
                    for ( int dj = -1; dj <= 1; dj++ ) {
                        for ( int di = -1; di <= 1; di++ ) {
                            spm_int_t pi = i + di, pj = j + dj, pk = k + dk;
                            if ( pi < 0 || pi >= nx || pj < 0 || pj >= ny || pk < 0 || pk >= nz ) {
                                continue;
                            }
                            spm_int_t p = gidx( pi, pj, pk, nx, ny );
                            double    v;
                            if ( p == q ) {
                                v = 26. * diffusion;
                            }
                            else {
                                v = -diffusion;
                                /* q = p + s e_c  =>  a(p,q) += s * 0.5 * v_c(p) */
                                int nzc = ( di != 0 ) + ( dj != 0 ) + ( dk != 0 );
                                if ( nzc == 1 ) {
                                    int c = ( di != 0 ) ? 0 : ( ( dj != 0 ) ? 1 : 2 );
                                    int s = -( di + dj + dk ); /* q - p direction */
                                    v += 0.5 * s * vel[3 * p + c];
                                }
                            }
                            spm->rowptr[pos]             = p;
                            ( (double *)spm->values )[pos] = v;
                            pos++;
                        }
                    }
                }
            }
        }
    }
    spm->colptr[n] = pos;
    free( vel );
    spmUpdateComputedFields( spm );
    return PASTIX_SUCCESS;
}

/*
 * Vector Laplacian with dof unknowns per node, stored expanded (dof = 1 in
 * the CSC, node-blocked numbering: unknown = dof*node + d).  Cell coefficient
 * c_p = exp(sigma * g_p) with g_p a standard normal from Box-Muller over the
 * plrnt stream; the dof x dof block coupling p,q is -min(c_p,c_q) I and the
 * diagonal block is 6 c_p I (diagonally dominant, SPD).
 */
int
spmGenVLaplacian7( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz,
                   spm_int_t dof, double sigma, unsigned long long seed )
{
    spm_int_t nn = nx * ny * nz;
    spm_int_t n  = nn * dof;
    spm_int_t pos = 0;
    double   *u = px_malloc( 2 * nn * sizeof( double ) );
    double   *c = px_malloc( nn * sizeof( double ) );

    spm_plrnt( 2 * nn, seed, 0, u );
    for ( spm_int_t p = 0; p < nn; p++ ) {
        double u1 = u[2 * p] + 0.5, u2 = u[2 * p + 1] + 0.5; /* (0,1] */
        if ( u1 <= 0. ) {
            u1 = 1e-300;
        }
        double g = sqrt( -2. * log( u1 ) ) * cos( 2. * M_PI * u2 );
        c[p] = exp( sigma * g );
    }
    spmInit( spm );
    spm->mtxtype = SpmSymmetric;
    spm->flttype = SpmDouble;
    spm->n       = n;
    spm->nnz     = dof * ( nn + ( nx - 1 ) * ny * nz + nx * ( ny - 1 ) * nz + nx * ny * ( nz - 1 ) );
    spm->colptr  = px_malloc( ( n + 1 ) * sizeof( spm_int_t ) );
    spm->rowptr  = px_malloc( spm->nnz * sizeof( spm_int_t ) );
    spm->values  = px_malloc( spm->nnz * sizeof( double ) );
    for ( spm_int_t k = 0; k < nz; k++ ) {
        for ( spm_int_t j = 0; j < ny; j++ ) {
            for ( spm_int_t i = 0; i < nx; i++ ) {
                spm_int_t p = gidx( i, j, k, nx, ny );
                spm_int_t nb[3];
                int       cnt = 0;
                if ( i + 1 < nx ) nb[cnt++] = gidx( i + 1, j, k, nx, ny );
                if ( j + 1 < ny ) nb[cnt++] = gidx( i, j + 1, k, nx, ny );
                if ( k + 1 < nz ) nb[cnt++] = gidx( i, j, k + 1, nx, ny );
                for ( spm_int_t d = 0; d < dof; d++ ) {
                    spm_int_t col = p * dof + d;
                    spm->colptr[col] = pos;
                    spm->rowptr[pos] = col;
                    ( (double *)spm->values )[pos++] = 6. * c[p];
                    for ( int t = 0; t < cnt; t++ ) {
                        spm->rowptr[pos] = nb[t] * dof + d;
                        ( (double *)spm->values )[pos++] = -fmin( c[p], c[nb[t]] );
                    }
                }
            }
        }
    }
    spm->colptr[n] = pos;
    free( u );
    free( c );
    spmUpdateComputedFields( spm );
    return PASTIX_SUCCESS;
}
